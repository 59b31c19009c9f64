"""Python handle of an ``sgpe_plan``: a batch of same-shape instances on one GPU.

A plan owns, per slot, the phase tables derived from the instance's physics
(the reference precomputes these as 16 full-grid complex arrays per run,
spinor_gpe/pspinor/tensor_propagator.py:138-149). Operator landscapes are
handed over as float64 host arrays; :func:`split_separable` recognizes the
ones that are sums of an x-only and a y-only part (harmonic traps, uniform or
axis-aligned detuning, free or Raman-shifted dispersions) so that the engine
can rebuild them from two 1D tables instead of reading a full grid every
step. Anything else is passed densely.
"""

import ctypes
import os

import numpy as np
import torch

from spinor_gpe import _native as nat

MAGIC_GAMMA = 1 / (2 + 2 ** (1 / 3))


def split_separable(a, tol_ulps=8):
    """Return ``(ax, ay)`` with ``a[jy, jx] == ax[jx] + ay[jy]`` to rounding.

    The split is anchored at the array minimum, so both parts are
    non-negative when ``a`` is. Returns None if ``a`` is not additively
    separable within ``tol_ulps`` units of ``max|a|``.
    """
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError("expected a 2D array")
    iy, ix = np.unravel_index(np.argmin(a), a.shape)
    ax = np.ascontiguousarray(a[iy, :])
    ay = np.ascontiguousarray(a[:, ix] - a[iy, ix])
    scale = float(np.max(np.abs(a))) or 1.0
    err = np.max(np.abs((ay[:, None] + ax[None, :]) - a))
    if err <= tol_ulps * np.finfo(np.float64).eps * scale:
        return ax, ay
    return None


class InstanceSpec:
    """Physics of one instance, in the units and layout of a PSpinor.

    ``kin`` / ``pot``: two (ny, nx) float64 arrays each (the reference's
    ``kin_eng_spin`` / ``pot_eng_spin``); ``coupling``: (ny, nx) array or
    None for no coupling (``is_coupling``); ``theta_x``: the coupling phase
    ``2 kL x`` as a function of x, or None (rotated frame).
    """

    def __init__(self, kin, pot, dr, atom_num, g_sc, time, t_step,
                 coupling=None, theta_x=None):
        if time not in ('imag', 'real'):
            raise ValueError(f"time must be 'imag' or 'real', not {time!r}")
        self.kin = [np.asarray(k, dtype=np.float64) for k in kin]
        self.pot = [np.asarray(v, dtype=np.float64) for v in pot]
        self.ny, self.nx = self.kin[0].shape
        self.dr = np.asarray(dr, dtype=np.float64)
        self.atom_num = float(atom_num)
        self.g = (float(g_sc['uu']), float(g_sc['dd']), float(g_sc['ud']))
        self.time = time
        self.t_step = float(t_step)
        self.coupling = None if coupling is None else np.asarray(coupling, np.float64)
        self.theta_x = None if theta_x is None else np.asarray(theta_x, np.float64)

    @classmethod
    def from_spin(cls, spin, time, t_step):
        """Read the same PSpinor attributes TensorPropagator reads
        (tensor_propagator.py:111-129)."""
        coupling = np.asarray(spin.coupling) if spin.is_coupling else None
        theta = None
        if spin.is_coupling and not spin.rot_coupling:
            theta = 2 * spin.kL_recoil * np.asarray(spin.space['x_mesh'])[0, :]
        return cls(spin.kin_eng_spin, spin.pot_eng_spin, spin.space['dr'],
                   spin.atom_num, spin.g_sc, time, t_step, coupling, theta)

    def descriptor(self):
        """Build an ``sgpe_inst_desc``; returns (desc, keep-alive arrays)."""
        d = nat.InstDesc()
        keep = []

        def ptr(arr):
            arr = np.ascontiguousarray(arr, dtype=np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        d.ny, d.nx = self.ny, self.nx
        d.time_mode = nat.TIME_IMAG if self.time == 'imag' else nat.TIME_REAL
        d.t_step = self.t_step
        d.atom_num = self.atom_num
        d.g_uu, d.g_dd, d.g_ud = self.g
        d.dx, d.dy = float(self.dr[0]), float(self.dr[1])
        for s in range(2):
            for arr, fx, fy, fd in ((self.kin[s], 't_x', 't_y', 't_dense'),
                                    (self.pot[s], 'v_x', 'v_y', 'v_dense')):
                if arr.shape != (self.ny, self.nx):
                    raise ValueError("operator grids must match the state shape")
                parts = split_separable(arr)
                if parts is None:
                    getattr(d, fd)[s] = ptr(arr)
                else:
                    getattr(d, fx)[s] = ptr(parts[0])
                    getattr(d, fy)[s] = ptr(parts[1])
        if self.coupling is None:
            d.coupling = nat.COUPLING_OFF
        else:
            c = self.coupling
            if np.all(c == c.flat[0]):
                d.coupling = nat.COUPLING_UNIFORM
                d.omega = float(c.flat[0])
            else:
                d.coupling = nat.COUPLING_ARRAY
                d.omega_dense = ptr(c)
            if self.theta_x is not None:
                d.theta_x = ptr(self.theta_x)
        return d, keep


class Plan:
    """A batch of ``n_slots`` instances of one (ny, nx) grid and precision."""

    def __init__(self, device, precision, ny, nx, n_slots=1):
        self.device = nat.require_device(device)
        self.prec = nat.precision_of(precision)
        self.dtype = nat.TORCH_DTYPE[self.prec]
        self.ny, self.nx, self.n_slots = int(ny), int(nx), int(n_slots)
        self.ctx = nat.Context.get(self.device)
        handle = ctypes.c_void_p()
        nat.check(nat.lib().sgpe_plan_create(self.ctx.handle, self.prec, self.ny, self.nx,
                                             self.n_slots, ctypes.byref(handle)))
        self.handle = handle
        self.states = [None] * self.n_slots
        self.run_pops = [None] * self.n_slots
        self.n_launch = self.n_slots  # kernels run for the first n_launch slots (compact)
        if not os.environ.get('SGPE_NO_TUNE_CACHE'):
            from spinor_gpe.engine import autotune  # pylint: disable=import-outside-toplevel
            autotune.apply(self)

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h:
            try:
                nat.lib().sgpe_plan_destroy(h)
            except Exception:  # pylint: disable=broad-except
                pass
            self.handle = None

    def _call(self, name, *args):
        self.ctx.bind_stream()
        nat.check(getattr(nat.lib(), name)(self.handle, *args))

    def new_state(self, psik=None):
        """Allocate a [2, ny, nx] state tensor, optionally filled from psik."""
        st = torch.empty((2, self.ny, self.nx), dtype=self.dtype, device=self.device)
        if psik is not None:
            host = np.stack([np.asarray(p) for p in psik])
            st.copy_(torch.as_tensor(host, dtype=self.dtype))
        return st

    def set_slot(self, slot, spec, state):
        """Load physics ``spec`` into ``slot`` and bind ``state`` ([2,ny,nx])."""
        if state.shape != (2, self.ny, self.nx) or state.dtype != self.dtype:
            raise ValueError("state must be a [2, ny, nx] tensor of the plan dtype")
        if not state.is_contiguous() or state.device != self.device:
            raise ValueError("state must be contiguous on the plan's device")
        desc, keep = spec.descriptor()
        self._call('sgpe_plan_set_slot', slot, ctypes.byref(desc), ctypes.c_void_p(state.data_ptr()))
        del keep
        self.states[slot] = state

    def set_state(self, slot, state):
        """Rebind ``slot`` to a new k-space state (marks it normalized k-space)."""
        self._call('sgpe_plan_set_state', slot, ctypes.c_void_p(state.data_ptr()))
        self.states[slot] = state

    def set_run(self, slot, pops=None, conv_tol=0.0):
        """Start a run: ``pops`` is a float64 (n, 2) device tensor or None.

        The plan keeps a reference to ``pops`` until the slot's next run, so
        the device never writes populations into a freed buffer."""
        if pops is None:
            self._call('sgpe_plan_set_run', slot, None, 0, float(conv_tol))
        else:
            if pops.dtype != torch.float64 or pops.device != self.device or not pops.is_contiguous():
                raise ValueError("pops must be a contiguous float64 tensor on the plan device")
            self._call('sgpe_plan_set_run', slot, ctypes.c_void_p(pops.data_ptr()),
                       int(pops.shape[0]), float(conv_tol))
        self.run_pops[slot] = pops

    def set_active(self, slot, active):
        self._call('sgpe_plan_set_active', slot, int(bool(active)))

    def compact(self, order, n_launch):
        """New slot i is old slot ``order[i]`` (state, tables and run state
        move with it); kernels run for the first ``n_launch`` slots, and the
        slots after them must be inactive (see ``sgpe_plan_compact``)."""
        order = [int(o) for o in order]
        arr = (ctypes.c_int * len(order))(*order)
        self._call('sgpe_plan_compact', arr, int(n_launch))
        self.states = [self.states[o] for o in order]
        self.run_pops = [self.run_pops[o] for o in order]
        self.n_launch = int(n_launch)

    def tune(self, row_e=0, row_rows=0, col_e=0, col_w=0):
        self._call('sgpe_plan_tune', int(row_e), int(row_rows), int(col_e), int(col_w))

    def tuning(self):
        """(row E, rows per block, column E, columns per block); rows per
        block < 0: the split row layout with that many rows per block; E = 0:
        that pass runs on the runtime-length kernels."""
        out = (ctypes.c_int * 4)()
        nat.check(nat.lib().sgpe_plan_get_tuning(self.handle, out))
        return tuple(out)

    def set_streaming(self, row_stream, col_stream):
        """Non-temporal (1) or normal (0) state accesses of each pass."""
        self._call('sgpe_plan_set_streaming', int(bool(row_stream)), int(bool(col_stream)))

    def streaming(self):
        out = (ctypes.c_int * 2)()
        nat.check(nat.lib().sgpe_plan_get_streaming(self.handle, out))
        return tuple(out)

    def set_persistent(self, on, wait_ticks=0):
        """Run blocks of full steps as one persistent launch where available
        (see ``sgpe_plan_set_persistent``); ``wait_ticks`` bounds each wait
        of a workgroup, in 100 MHz ticks (0: the library default). When the
        plan's current shapes have no persistent kernel (the tuned shapes
        need not have one), the plan switches to the first compiled shape
        of its grid that has one and is resident."""
        self._call('sgpe_plan_set_persistent', int(bool(on)), int(wait_ticks))
        if on and self.persistent()[2] == 0 and self.ny == self.nx:
            from spinor_gpe.engine import autotune  # pylint: disable=import-outside-toplevel
            for shape in autotune.persistent_shapes(self.prec, self.nx):
                try:
                    self.tune(*shape)
                except nat.SgpeError:
                    continue
                if self.persistent()[2]:
                    break

    def persistent(self):
        """(requested, used for the next block, grid size or 0)."""
        out = (ctypes.c_int * 3)()
        nat.check(nat.lib().sgpe_plan_get_persistent(self.handle, out))
        return tuple(out)

    def check(self):
        """Synchronize; raise if a persistent launch stopped on its wait bound."""
        self._call('sgpe_plan_check')

    def profile_persistent(self, n):
        """Milliseconds of ``n`` full steps as one persistent launch."""
        out = ctypes.c_double()
        self.ctx.bind_stream()
        nat.check(nat.lib().sgpe_profile_persistent(self.handle, int(n), ctypes.byref(out)))
        return out.value

    def single_step(self, which):
        self._call('sgpe_single_step', int(which))

    def full_steps(self, n):
        self._call('sgpe_full_steps', int(n))

    def graph_steps(self, steps_per_graph, n_graphs):
        self._call('sgpe_graph_steps', int(steps_per_graph), int(n_graphs))

    def run(self, n, block=64):
        """``n`` full steps as captured graphs: ``n // block`` launches of a
        ``block``-step graph, then one graph of the remaining steps."""
        n = int(n)
        nb, rem = divmod(n, block)
        if nb:
            self.graph_steps(block, nb)
        if rem:
            self.graph_steps(rem, 1)

    def profile_steps(self, n):
        """Run ``n`` full steps eagerly with a hipEvent after every kernel.

        Returns (column-pass ms, row-pass ms, column launches, row launches).
        """
        out = (ctypes.c_double * 4)()
        self.ctx.bind_stream()
        nat.check(nat.lib().sgpe_profile_full_steps(self.handle, int(n), out))
        return tuple(out)

    def sync(self):
        self._call('sgpe_plan_sync')

    def status(self):
        """Host copy of the run state: (active, converged, steps) int array
        of shape (n_slots, 3) and the last energies, float64 (n_slots, 8)."""
        st = np.zeros((self.n_slots, 3), dtype=np.int32)
        en = np.zeros((self.n_slots, 8), dtype=np.float64)
        self._call('sgpe_plan_status', st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                   en.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return st, en

    def energy(self, out=None):
        """GP energy terms per slot, float64 tensor (n_slots, 8):
        E_tot, E_kin, E_pot, E_int, E_cpl, mu, N, relative change per step
        since the previous evaluation (-1 for the first). Leaves the states
        in normalized k-space."""
        if out is None:
            out = torch.empty((self.n_slots, 8), dtype=torch.float64, device=self.device)
        self._call('sgpe_energy', ctypes.c_void_p(out.data_ptr()))
        return out
