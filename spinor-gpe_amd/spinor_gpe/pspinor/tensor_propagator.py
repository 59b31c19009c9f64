"""Drop-in ``TensorPropagator`` backed by the HIP split-step engine.

Same constructor, attributes and methods as the reference
(``spinor_gpe/pspinor/tensor_propagator.py:11-324``). The state lives on the
GPU in one contiguous ``[2, Ny, Nx]`` tensor; ``full_step``/``single_step``/
``prop_loop`` enqueue libsgpe kernels (two sweeps per sub-step, no host
synchronization inside the loop; loops run as captured hipGraph blocks).
The only device accepted is a HIP GPU (``device='cuda'``).

Keyword-only additions: ``precision`` ('c128' default, or 'c64': complex64
state and arithmetic, float64 reductions), ``block_steps`` (full steps per
captured graph), ``progress`` (tqdm bar, default on as in the reference) and
``conv_tol`` (> 0: stop once the GP energy changes by at most this relative
amount per full step, checked every ``block_steps`` steps; the result then
holds the steps actually run) and ``persistent`` (run each block of steps as
one launch of the persistent step kernel, on a shape of the grid that has
one compiled; slower than the default pass launches on MI355X).
"""

import warnings

import numpy as np
import torch

from spinor_gpe import _native as nat
from spinor_gpe.engine.plan import MAGIC_GAMMA, InstanceSpec, Plan
from spinor_gpe.pspinor import prop_result
from spinor_gpe.pspinor import tensor_tools as ttools
from spinor_gpe.pspinor.plotting_tools import next_available_path

ENERGY_KEYS = ('E_tot', 'E_kin', 'E_pot', 'E_int', 'E_cpl', 'mu', 'N', 'rel_change')

try:
    from tqdm import tqdm as _tqdm
except ImportError:  # pragma: no cover
    _tqdm = None


class _DeviceDict(dict):
    """The reference's ``space`` dict of device tensors (tensor_propagator.py:
    119-120), holding host arrays and making each device tensor on first
    access."""

    def __init__(self, host, dev):
        super().__init__(host)
        self._dev = dev
        self._made = set()

    def __getitem__(self, key):
        val = super().__getitem__(key)
        if key not in self._made:
            val = ttools.remember_host_copy(torch.tensor(val, device=self._dev), val)
            super().__setitem__(key, val)
            self._made.add(key)
        return val

    def get(self, key, default=None):
        return self[key] if key in self else default

    def values(self):
        return [self[k] for k in self]

    def items(self):
        return [(k, self[k]) for k in self]


class TensorPropagator:
    """GPU propagator of the pseudospin-1/2 GPE (tensor_propagator.py:11-61)."""

    # pylint: disable=too-many-instance-attributes,too-many-arguments
    def __init__(self, spin, t_step, n_steps, device='cpu', time='imag',
                 is_sampling=False, n_samples=1, *, precision='c128',
                 block_steps=64, progress=True, conv_tol=0.0, persistent=False):
        """Set up a propagation (tensor_propagator.py:67-149).

        ``device='cpu'`` (the reference's default) runs the engine on the
        current HIP GPU, since there is no CPU propagation path, and presents
        the state and operator grids as host tensors, as the reference's CPU
        run would; any other device must be a HIP GPU.
        """
        self._view_dev = None
        if torch.device(device).type == 'cpu':
            warnings.warn("TensorPropagator: the split-step engine runs only on HIP GPUs; "
                          "device='cpu' runs it on the current GPU and returns host tensors",
                          RuntimeWarning, stacklevel=2)
            self._view_dev = torch.device('cpu')
            device = 'cuda'
        self._dev = nat.require_device(device)
        if self._view_dev is None:
            self._view_dev = self._dev
        self.n_steps = n_steps
        self.device = device
        self.paths = spin.paths
        if time == 'imag':
            self.t_step = -1.0j * t_step
        elif time == 'real':
            self.t_step = t_step
        else:
            raise ValueError(f"time must be 'imag' or 'real', not {time!r}")
        self.time = time
        self.dt_out = self.t_step * MAGIC_GAMMA
        self.dt_in = self.t_step * (1 - 2 * MAGIC_GAMMA)

        self.rand_seed = spin.rand_seed
        if self.rand_seed is not None:
            torch.manual_seed(self.rand_seed)
        self.is_sampling = is_sampling

        self.atom_num = spin.atom_num
        self.is_coupling = spin.is_coupling
        self.g_sc = spin.g_sc
        # The operator grids stay on the host: the engine builds its own
        # phase tables from them, and only the end-of-run energy reads them.
        # The reference's device copies (tensor_propagator.py:114-129) are
        # made on first access of the attributes below.
        self._kin_host = [np.asarray(k, dtype=np.float64) for k in spin.kin_eng_spin]
        self._pot_host = [np.asarray(v, dtype=np.float64) for v in spin.pot_eng_spin]
        keys_space = ['dr', 'dk', 'x_mesh', 'y_mesh', 'dv_r', 'dv_k']
        self._space_host = {k: np.asarray(spin.space[k]) for k in keys_space}
        self._coupling_host = np.asarray(spin.coupling, dtype=np.float64)
        self.space = _DeviceDict(self._space_host, self._view_dev)
        self._device_copies = {}
        # pylint: disable=invalid-name
        self.kL_recoil = spin.kL_recoil
        self._rot_coupling = spin.rot_coupling
        if self.is_sampling:
            assert self.n_steps % n_samples == 0, (
                f"The number of samples requested {n_samples} does not evenly "
                f"divide the total number of steps {self.n_steps}.")
        self.sample_rate = self.n_steps / n_samples
        # The reference stores 16 precomputed operator grids here; the engine
        # builds its own phase tables, so these name the sub-step only.
        self.eng_out = {'sub': nat.SUB_OUTER, 't_step': self.dt_out}
        self.eng_in = {'sub': nat.SUB_INNER, 't_step': self.dt_in}

        self.precision = precision
        self.block_steps = int(block_steps)
        self.progress = progress
        self.conv_tol = float(conv_tol)
        spec = InstanceSpec.from_spin(spin, time, t_step)
        self._plan = Plan(self._dev, precision, spec.ny, spec.nx, 1)
        self._state = self._plan.new_state(spin.psik)
        self._plan.set_slot(0, spec, self._state)
        if persistent:
            self._plan.set_persistent(True)
        self._pending = False  # True while the state awaits its closing half-step

    # ----------------------------------------- operator grids (on demand)
    def _device_copy(self, key, make):
        if key not in self._device_copies:
            self._device_copies[key] = make()
        return self._device_copies[key]

    @property
    def kin_eng_spin(self):
        """Kinetic energy grids as device tensors (tensor_propagator.py:114)."""
        return self._device_copy('kin', lambda: ttools.to_tensor(self._kin_host, dev=self._view_dev))

    @property
    def pot_eng_spin(self):
        """Potential energy grids as device tensors (tensor_propagator.py:115)."""
        return self._device_copy('pot', lambda: ttools.to_tensor(self._pot_host, dev=self._view_dev))

    @property
    def coupling(self):
        """Coupling grid as a device tensor (tensor_propagator.py:121)."""
        return self._device_copy('coupling', lambda: ttools.to_tensor(self._coupling_host, dev=self._view_dev))

    @property
    def expon(self):
        """Coupling phase: 0 in the rotated frame, else 2 kL x
        (tensor_propagator.py:126-129)."""
        if self._rot_coupling:
            return torch.tensor(0.0)
        return self._device_copy('expon', lambda: 2 * self.kL_recoil * self.space['x_mesh'])

    # ------------------------------------------------------------- state
    @property
    def psik(self):
        """Current normalized k-space state as ``[psik_up, psik_down]``: views
        of the device state, or host copies when the propagator was made with
        ``device='cpu'``."""
        self._settle()
        if self._view_dev.type == 'cpu':
            return [self._state[0].cpu(), self._state[1].cpu()]
        return [self._state[0], self._state[1]]

    @psik.setter
    def psik(self, value):
        vals = [torch.as_tensor(v) for v in value]
        self._state.copy_(torch.stack(vals).to(device=self._dev, dtype=self._state.dtype))
        self._plan.set_state(0, self._state)
        self._pending = False

    def _settle(self):
        if self._pending:
            self._plan.sync()
            self._pending = False

    # --------------------------------------------------------- stepping
    def full_step(self):
        """One split step: outer, inner, outer sub-steps (tensor_propagator.py:214-222)."""
        self._plan.full_steps(1)
        self._pending = True

    def single_step(self, t_step, eng):
        """One sub-step with the operators of ``eng`` (tensor_propagator.py:224-271).

        ``eng`` is ``self.eng_out`` or ``self.eng_in``; ``t_step`` must be its
        sub-step duration.
        """
        if isinstance(eng, dict) and 'sub' in eng:
            which = eng['sub']
        elif t_step == self.dt_out:
            which = nat.SUB_OUTER
        elif t_step == self.dt_in:
            which = nat.SUB_INNER
        else:
            raise ValueError("single_step: t_step must be dt_out or dt_in")
        self._plan.single_step(which)
        self._pending = True

    def _advance(self, n, bar):
        """Run up to ``n`` full steps; returns the number run (fewer only if
        the convergence test stopped the run)."""
        done = 0
        chunk = self.block_steps if self.conv_tol > 0 else self.block_steps * 16
        while done < n:
            k = min(chunk, n - done)
            self._plan.run(k, block=self.block_steps)
            self._pending = True
            done += k
            if bar is not None:
                bar.update(k)
            if self.conv_tol > 0:
                self._plan.energy()
                self._pending = False
                state, _ = self._plan.status()
                if state[0, 1]:
                    self._converged = True
                    break
        return done

    def prop_loop(self, n_steps):
        """Run ``n_steps`` full steps and return a :class:`PropResult`.

        Records the populations after every step and, with sampling, the
        normalized k-space state before every ``sample_rate``-th step
        (tensor_propagator.py:151-212).
        """
        self._settle()
        pop_times = np.linspace(0, self.n_steps * np.abs(self.t_step), n_steps)
        pops = {'times': pop_times, 'vals': np.empty((n_steps, 2))}
        pops_dev = torch.zeros((max(n_steps, 1), 2), dtype=torch.float64, device=self._dev)
        self._plan.set_run(0, pops_dev, conv_tol=self.conv_tol)
        self._converged = False

        sample_at = []
        if self.is_sampling:
            n_samples = int(n_steps / self.sample_rate)
            ny, nx = self._state.shape[1:]
            sampled_psik = np.empty((n_samples, 2, ny, nx), dtype=np.complex128)
            sampled_times = np.linspace(0, self.n_steps * np.abs(self.t_step), n_samples)
            sample_at = [i for i in range(n_steps) if i % self.sample_rate == 0]

        bar = _tqdm(total=n_steps) if (self.progress and _tqdm is not None) else None
        done = 0
        for i in sample_at:
            if self._converged:
                break
            done += self._advance(i - done, bar)
            if self._converged:
                break
            idx = int(i / self.sample_rate)
            sampled_psik[idx] = np.array(ttools.to_numpy(self.psik))
        if not self._converged:
            done += self._advance(n_steps - done, bar)
        if bar is not None:
            bar.close()
        self._settle()
        if done < n_steps:  # stopped by the convergence test
            pops['times'] = pop_times[:done]
            pops['vals'] = np.empty((done, 2))
        pops['vals'][:] = pops_dev[:done].cpu().numpy()
        self._plan.set_active(0, True)
        egp = self._plan.energy()[0].cpu().numpy()
        # Detach the population buffer before it goes out of scope: the run
        # state would otherwise keep its address, and later full_step() /
        # single_step() calls would write populations into freed memory.
        self._plan.set_run(0, None, conv_tol=0.0)
        del pops_dev
        self._plan.check()

        self._settle()
        on_device = [self._state[0], self._state[1]]
        psik = ttools.to_numpy(on_device)
        psi = ttools.to_numpy(ttools.ifft_2d(on_device, self._space_host['dr']))
        energy = self.eng_expect(psik, psi=psi)

        if self.is_sampling:
            test_name = self.paths['trial'] + 'psik_sampled'
            file_name = next_available_path(test_name, self.paths['folder'], '.npz')
            np.savez(file_name, psiks=sampled_psik, times=sampled_times)
        else:
            file_name = None

        result = prop_result.PropResult(psi, psik, energy, pops, file_name)
        result.steps_run = done
        result.converged = self._converged
        result.energy_gp = dict(zip(ENERGY_KEYS, (float(v) for v in egp)))
        result.chem_pot = result.energy_gp['mu']
        return result

    # ------------------------------------------------------------ energy
    def eng_expect(self, psik, psi=None):
        """End-of-run energy list ``[E_tot, E_kin, E_pot, E_int]`` on the host
        (tensor_propagator.py:273-324); see :func:`eng_expect_arrays`.
        ``psi`` may pass the already transformed real-space state."""
        assert len(psik) == 2, "Requires two spinor components to calculate the energy."
        return eng_expect_arrays(psik, self._space_host['dr'], self._pot_host, self._coupling_host,
                                 self.g_sc, self.kL_recoil, self.is_coupling, psi=psi)


def eng_expect_arrays(psik, delta_r, pot_eng_spin, coupling, g_sc, kl_recoil, is_coupling, psi=None):
    """The reference's end-of-run energy of a k-space spinor, host NumPy.

    Same functional and conventions as tensor_propagator.py:293-324: the
    real-space state by the NumPy ``ifft_2d``; finite-difference gradients
    (``np.gradient(f, dr[0], dr[1])``: dr[0] is the spacing of axis 0) of
    the square-root densities and of the unwrapped phases, masked to 0 where
    the density is below 1e-6 of its maximum (tensor_tools.py:528-539); the
    recoil term uses the axis-0 phase derivative; interaction terms have no
    1/2; the coupling term omits the coupling phase; no area element.
    Returns ``[E_tot, sum kin, sum pot, sum int]``.
    """
    if not isinstance(psik[0], np.ndarray):
        psik = ttools.to_numpy(psik)
    delta_r = np.asarray(ttools.to_numpy(delta_r) if not isinstance(delta_r, np.ndarray) else delta_r)
    if psi is None:
        psi = ttools.ifft_2d(psik, delta_r)
    dens = ttools.density(psi)
    root = [np.sqrt(d) for d in dens]
    ph = ttools.phase(psi, uwrap=True, dens=dens)
    ph_ax0 = [g[0] for g in ttools.grad(ph, delta_r)]
    ph_gsq = ttools.grad_sq(ph, delta_r)
    kin = (sum(ttools.grad_sq(root, delta_r))
           + sum([d * p for d, p in zip(dens, ph_gsq)])
           + sum([d * p for d, p in zip(dens, ph_ax0)]) * 2 * kl_recoil * is_coupling) / 2
    pot = sum([d * np.asarray(v) for d, v in zip(dens, pot_eng_spin)])
    int_e = (g_sc['uu'] * dens[0] ** 2 + g_sc['dd'] * dens[1] ** 2
             + g_sc['ud'] * dens[0] * dens[1])
    cpl = (np.conj(psi[0]) * psi[1] + np.conj(psi[1]) * psi[0]) * np.asarray(coupling) / 2
    total = np.real((kin + pot + int_e + cpl).sum())
    return [total, np.real(kin).sum(), np.real(pot).sum(), np.real(int_e).sum()]
