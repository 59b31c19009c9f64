"""Pseudospin-1/2 GPE facade: parameters, grids, initial state, entry points.

``PSpinor`` keeps the reference's constructor, attributes and entry points
(``spinor_gpe/pspinor/pspinor.py:21-962``) so user scripts run unchanged.
Everything here is one-off host set-up in float64 NumPy, as in the reference;
the propagation it hands off to (:class:`TensorPropagator`) runs on the GPU
through the HIP engine.

Differences from the reference, all deliberate:

* a relative ``path`` is created under ``$SPINOR_GPE_DATA`` (default:
  ``./data``) instead of under the reference's source tree;
* the plotting helpers import matplotlib on first use;
* ``imaginary``/``real`` accept keyword-only engine options (``precision``,
  ``block_steps``) after the reference's own parameters.
"""

import os
import shutil

import numpy as np
from scipy.ndimage import fourier_shift

from spinor_gpe import constants as const
from spinor_gpe.pspinor import plotting_tools as ptools
from spinor_gpe.pspinor import tensor_tools as ttools


def _data_root():
    return os.environ.get('SPINOR_GPE_DATA', os.path.join(os.getcwd(), 'data'))


class PSpinor:
    """Set-up object for a pseudospin-1/2 condensate (pspinor.py:21-110).

    Lengths are in units of the x harmonic-oscillator length ``a_x``,
    energies in ``hbar * omeg['x']`` and times in ``1 / omeg['x']``.
    """

    # pylint: disable=too-many-instance-attributes
    def __init__(self, path, omeg=None, g_sc=None, mesh_points=(256, 256),
                 r_sizes=(16, 16), atom_num=1e4, pop_frac=(0.5, 0.5),
                 **kwargs):
        """Build grids, energy landscapes and the Thomas-Fermi state.

        Follows ``pspinor.py:113-196``. ``g_sc`` is rescaled in place, as in
        the reference (``pspinor.py:310``): pass a fresh dict per object.
        """
        # pylint: disable=too-many-arguments
        phase_factor = kwargs.get('phase_factor', 1)
        overwrite = kwargs.get('overwrite', False)
        self.setup_data_path(path, overwrite)

        self.atom_num = atom_num
        self.space = {}

        assert sum(pop_frac) == 1.0, "Total population must equal 1."
        self.pop_frac = pop_frac

        if omeg is None:
            w0 = 2 * np.pi * 50
            self.omeg = {'x': w0, 'y': w0, 'z': 40 * w0}
        else:
            expected = {'x', 'y', 'z'}
            assert expected == omeg.keys(), (
                f"Keys for `omeg` must have the form: {expected}.")
            self.omeg = omeg

        if g_sc is None:
            self.g_sc = {'uu': 1.0, 'dd': 0.995, 'ud': 0.995}
        else:
            expected = {'uu', 'dd', 'ud'}
            assert expected == g_sc.keys(), (
                f"Keys for `g_sc` must have the form: {expected}.")
            self.g_sc = g_sc

        self.compute_tf_params()
        self.compute_spatial_grids(mesh_points, r_sizes)
        self.compute_energy_grids()
        self.compute_tf_psi(phase_factor)
        self.no_coupling_setup()

        self.rand_seed = None
        self.prop = None
        shape = tuple(np.flip(np.asarray(mesh_points)))
        self.coupling = np.zeros(shape)
        self.detuning = np.zeros(shape)
        self.rot_coupling = True

    # ------------------------------------------------------------------ I/O
    def setup_data_path(self, path, overwrite):
        """Create ``<path>/``, ``code/`` and ``trial_data/`` (pspinor.py:198-246).

        An existing directory raises ``FileExistsError`` unless ``overwrite``,
        in which case it is removed first.
        """
        base = path if os.path.isabs(path) else os.path.join(_data_root(), path)
        data_path = os.path.normpath(base) + os.sep
        if os.path.isdir(data_path):
            if not overwrite:
                raise FileExistsError(
                    f"The directory {data_path} already exists. To overwrite "
                    "this directory, supply the parameter `overwrite=True`.")
            shutil.rmtree(data_path)
        trial = os.path.normpath(data_path + 'trial_data') + os.sep
        code = os.path.normpath(data_path + 'code') + os.sep
        for p in (data_path, code, trial):
            os.makedirs(p, exist_ok=True)
        folder = os.path.basename(os.path.normpath(data_path))
        self.paths = {'data': data_path, 'trial': trial, 'code': code,
                      'folder': folder}

    # ----------------------------------------------------------- parameters
    def compute_tf_params(self, species='Rb87'):
        """Length/energy scales, TF chemical potential and g scaling.

        ``a_x = sqrt(hbar / (m omeg_x))``; ``a_sc = a / a_x``;
        ``mu_TF = sqrt(4 N a_sc (omeg_y/omeg_x) sqrt(omeg_z / (2 pi omeg_x)))``;
        ``g -> sqrt(8 pi omeg_z/omeg_x) a_sc g`` (pspinor.py:283-313).
        """
        ratio_y = self.omeg['y'] / self.omeg['x']
        ratio_z = self.omeg['z'] / self.omeg['x']
        self.a_x = np.sqrt(const.hbar / (const.Rb87['m'] * self.omeg['x']))
        self.a_sc = const.Rb87['a_sc'] / self.a_x if species == 'Rb87' else 1
        self.chem_pot = np.sqrt(4 * self.atom_num * self.a_sc * ratio_y
                                * np.sqrt(ratio_z / (2 * np.pi)))
        g_scale = np.sqrt(8 * ratio_z * np.pi) * self.a_sc
        self.g_sc.update({key: g_scale * val for key, val in self.g_sc.items()})
        self.rad_tf = np.sqrt(2 * self.chem_pot)
        self.time_scale = 1 / self.omeg['x']

    def compute_spatial_grids(self, mesh_points=(256, 256), r_sizes=(16, 16)):
        """Real- and k-space grids (pspinor.py:315-381).

        Arrays are ``(Ny, Nx)`` with x along axis 1. Both linear axes start
        at ``-size`` and exclude ``+size``; k-space arrays are thus in the
        fft-shifted order.
        """
        assert all(n % 2 == 0 for n in mesh_points), (
            f"Number of mesh points {mesh_points} should be powers of 2.")
        mesh_points = np.array(mesh_points)
        r_sizes = np.array(r_sizes)
        self.space['dr'] = 2 * r_sizes / mesh_points
        k_sizes = np.pi / self.space['dr']
        self.space['dk'] = np.pi / r_sizes
        self.space['x'] = self._compute_lin(r_sizes, mesh_points, axis=0)
        self.space['y'] = self._compute_lin(r_sizes, mesh_points, axis=1)
        self.space['kx'] = self._compute_lin(k_sizes, mesh_points, axis=0)
        self.space['ky'] = self._compute_lin(k_sizes, mesh_points, axis=1)
        x_mesh, y_mesh = np.meshgrid(self.space['x'], self.space['y'])
        kx_mesh, ky_mesh = np.meshgrid(self.space['kx'], self.space['ky'])
        self.space.update({'x_mesh': x_mesh, 'y_mesh': y_mesh,
                           'kx_mesh': kx_mesh, 'ky_mesh': ky_mesh})
        self.space['dv_r'] = np.prod(self.space['dr'])
        self.space['dv_k'] = np.prod(self.space['dk'])
        self.space['mesh_points'] = mesh_points
        self.space['r_sizes'] = r_sizes
        self.space['k_sizes'] = k_sizes

    @classmethod
    def _compute_lin(cls, sizes, points, axis=0):
        """``linspace(-size, size, n, endpoint=False)`` along one axis."""
        return np.linspace(-sizes[axis], sizes[axis], num=points[axis],
                           endpoint=False)

    @property
    def pot_eng(self):
        """Spin-independent potential energy grid [hbar omeg_x]."""
        return self._pot_eng

    @pot_eng.setter
    def pot_eng(self, array):
        # Assigning the potential resets both spin landscapes (and therefore
        # drops any detuning), as in pspinor.py:392-396.
        self._pot_eng = array
        self.pot_eng_spin = [self._pot_eng] * 2

    @property
    def kin_eng(self):
        """Kinetic energy grid on the shifted k mesh [hbar omeg_x]."""
        return self._kin_eng

    @kin_eng.setter
    def kin_eng(self, array):
        self._kin_eng = array
        self.kin_eng_spin = [self._kin_eng] * 2

    def compute_energy_grids(self):
        """Harmonic trap and free dispersion (pspinor.py:413-430)."""
        ratio_y = self.omeg['y'] / self.omeg['x']
        self.pot_eng = (self.space['x_mesh'] ** 2
                        + (ratio_y * self.space['y_mesh']) ** 2) / 2
        self.kin_eng = (self.space['kx_mesh'] ** 2
                        + self.space['ky_mesh'] ** 2) / 2

    def compute_tf_psi(self, phase_factor=1.0):
        """Thomas-Fermi initial state, its FFT and healing lengths.

        Follows pspinor.py:248-281 and writes ``trial_data/tf_wf-<folder>.npz``
        with keys ``psi`` and ``psik``.
        """
        assert abs(phase_factor) == 1.0, (
            "Relative phase factor must have unit magnitude.")
        g_bare = [self.g_sc['uu'], self.g_sc['dd']]
        profile = np.real(np.sqrt(self.chem_pot - self.pot_eng + 0.j))
        self.psi = [profile * np.sqrt(pop / abs(g))
                    for pop, g in zip(self.pop_frac, g_bare)]
        self.psi[1] = self.psi[1] * phase_factor
        self.psi, _ = ttools.norm(self.psi, self.space['dv_r'], self.atom_num)
        self.psik = ttools.fft_2d(self.psi, self.space['dr'])
        self.heal = [(8 * np.pi * np.max(np.abs(p) ** 2) * self.a_sc) ** (-1 / 2)
                     for p in self.psi]
        np.savez(self.paths['trial'] + 'tf_wf-' + self.paths['folder'],
                 psi=self.psi, psik=self.psik)

    def _calc_atoms(self, psi=None, space='r'):
        """Total atom number of ``psi`` in real ('r') or k ('k') space."""
        if space == 'r':
            psi = self.psi if psi is None else psi
            vol = self.space['dv_r']
        elif space == 'k':
            psi = self.psik if psi is None else psi
            vol = self.space['dv_k']
        else:
            raise ValueError(f"space must be 'r' or 'k', not {space!r}")
        return ttools.calc_atoms(psi, vol)

    # ------------------------------------------------------------- coupling
    def no_coupling_setup(self):
        """No Raman coupling (pspinor.py:462-467)."""
        self.is_coupling = False
        # pylint: disable=invalid-name
        self.kL_recoil = 1.0
        self.EL_recoil = 1.0

    def coupling_setup(self, wavel=790.1e-9, scale=1.0, kin_shift=False):
        """Raman recoil scales and spin-dependent dispersion (pspinor.py:469-501).

        ``kL = scale * sqrt(2) pi a_x / wavel``, ``E_L = kL^2 / 2``. With
        ``kin_shift`` the two dispersions are shifted by ``+/- kx kL``; each is
        then offset so that its minimum is zero.
        """
        self.is_coupling = True
        # pylint: disable=invalid-name
        self.kL_recoil = scale * (np.sqrt(2) * np.pi / wavel * self.a_x)
        self.EL_recoil = self.kL_recoil ** 2 / 2
        shift = self.space['kx_mesh'] * self.kL_recoil if kin_shift else 0
        spins = [self.kin_eng + shift, self.kin_eng - shift]
        self.kin_eng_spin = [k - np.min(k) for k in spins]

    def shift_momentum(self, psik=None, scale=1.0, frac=(0.5, 0.5)):
        """Split each component's momentum peak by ``+/- scale * kL``.

        Follows pspinor.py:503-545, including its use of the forward
        transform and of the sub-pixel ``scipy.ndimage.fourier_shift`` along
        x, and the alternation of ``frac`` between the two components.
        """
        assert self.is_coupling, (
            f"The `is_coupling` option is {self.is_coupling}. Initialize "
            "coupling with `coupling_setup()`.")
        if psik is None:
            psik = self.psik
        shift = scale * self.kL_recoil / self.space['dk'][0]
        spec = ttools.fft_2d(psik, self.space['dr'])
        out = [np.zeros_like(pk) for pk in psik]
        frac = np.asarray(frac)
        for i in range(len(psik)):
            right = fourier_shift(spec[i], shift=[0, shift], axis=1)
            left = fourier_shift(spec[i], shift=[0, -shift], axis=1)
            out[i] = frac[0] * right + frac[1] * left
            frac = np.flip(frac)
        self.psik = ttools.ifft_2d(out, self.space['dr'])
        self.psi = ttools.ifft_2d(self.psik, self.space['dr'])

    @property
    def coupling(self):
        """Raman coupling strength grid Omega [hbar omeg_x]."""
        return self._coupling

    @coupling.setter
    def coupling(self, array):
        self._coupling = array

    @property
    def detuning(self):
        """Raman detuning grid delta [hbar omeg_x]; sets V +/- delta/2."""
        return self._detuning

    @detuning.setter
    def detuning(self, array):
        self._detuning = array
        self.pot_eng_spin = [self.pot_eng + self._detuning / 2,
                             self.pot_eng - self._detuning / 2]

    def _axis_mesh(self, axis):
        if axis == 0:
            return self.space['x_mesh']
        if axis == 1:
            return self.space['y_mesh']
        raise ValueError(f"axis must be 0 (x) or 1 (y), not {axis!r}")

    def coupling_grad(self, slope, offset, axis=1):
        """Linear coupling gradient along x (axis=0) or y (axis=1)."""
        self.coupling = self._axis_mesh(axis) * slope + offset

    def coupling_uniform(self, value):
        """Uniform, non-negative coupling (pspinor.py:608-625)."""
        assert value >= 0, f"Cannot have a negative coupling value: {value}."
        self.coupling = np.ones_like(self.space['x_mesh']) * value

    def detuning_grad(self, slope, offset=0.0, axis=1):
        """Linear detuning gradient along x (axis=0) or y (axis=1)."""
        self.detuning = self._axis_mesh(axis) * slope + offset

    def detuning_uniform(self, value):
        """Uniform detuning (pspinor.py:662-678)."""
        self.detuning = np.ones_like(self.space['x_mesh']) * value

    def seed_vortices(self, positions, windings, *, components=(0, 1)):
        """Imprint vortices on the real-space state (pspinor.py:680-745).

        Each vortex multiplies component ``i`` by ``r/sqrt(r^2 + xi_i^2)``
        times ``exp(+/- i atan2(dy, dx))``, with ``xi_i`` the component's
        healing length. ``positions`` is ``(n, 2)`` (shared by both
        components) or ``(2, n, 2)``; ``windings`` is one value, one per
        position, or ``(2, n)``. Keyword-only addition: ``components`` lists
        the components to imprint (default both, as in the reference).
        """
        positions = np.array(positions)
        assert positions.shape[-1] == 2, (
            "Positions should be ordered pairs, e.g. (x, y).")
        shared = positions.ndim == 2
        if shared:
            n_pos = len(positions)
            positions = np.array([positions, positions])
        else:
            assert positions.ndim == 3, (
                "`positions` must be at most a three-dimensional array.")
        windings = np.array(windings)
        if windings.ndim == 0 or windings.shape == (1,):
            windings = np.full(positions.shape[:-1], windings.reshape(-1)[0])
        elif windings.ndim == 1:
            assert shared and len(windings) == n_pos, (
                "The number of supplied windings must match the number of "
                "supplied positions")
            windings = np.array([windings, windings])
        else:
            assert windings.shape == positions.shape[:-1], (
                "The number of supplied windings must match the number of "
                "supplied positions")
        for i in components:
            for j, (x0, y0) in enumerate(positions[i]):
                dx = self.space['x_mesh'] - x0
                dy = self.space['y_mesh'] - y0
                rho = np.sqrt(dx ** 2 + dy ** 2) / self.heal[i]
                core = rho / np.sqrt(rho ** 2 + 1)
                winding = np.exp(windings[i, j] * 1j * np.arctan2(dy, dx))
                self.psi[i] = self.psi[i] * core * winding
        self.psik = ttools.fft_2d(self.psi, delta_r=self.space['dr'])

    def seed_regular_vortices(self):
        """Not provided, as in the reference (pspinor.py:747-752)."""
        raise NotImplementedError()

    def seed_random_vortices(self, N):  # pylint: disable=invalid-name
        """Not provided, as in the reference (pspinor.py:754-756)."""
        raise NotImplementedError()

    # ------------------------------------------------------------- figures
    def _extent(self, key, scale):
        """``[-x, x, -y, y] / scale`` from the half-widths ``space[key]``."""
        half = np.asarray(self.space[key], dtype=float)
        return np.array([-half[0], half[0], -half[1], half[1]]) / scale

    def plot_rdens(self, psi=None, spin=None, cmap='viridis', scale=1.0):
        """Real-space density of both components (``spin=None``) or of one
        (0 up, 1 down), lengths in units of ``scale`` (pspinor.py:758-787)."""
        return ptools.plot_dens(self.psi if psi is None else psi, spin, cmap, scale,
                                extent=self._extent('r_sizes', scale))

    def plot_kdens(self, psik=None, spin=None, cmap='viridis', scale=1.0):
        """Momentum-space density, as :meth:`plot_rdens` (pspinor.py:789-818)."""
        return ptools.plot_dens(self.psik if psik is None else psik, spin, cmap, scale,
                                extent=self._extent('k_sizes', scale))

    def plot_rphase(self, psi=None, spin=None, cmap='twilight_shifted', scale=1.0):
        """Real-space phase, zero where the density is below 1e-6 of its
        maximum (pspinor.py:820-851)."""
        return ptools.plot_phase(self.psi if psi is None else psi, spin, cmap, scale,
                                 extent=self._extent('r_sizes', scale))

    # pylint: disable=too-many-arguments
    def plot_spins(self, rscale=1.0, kscale=1.0, cmap='viridis', save=True, ext='.pdf',
                   zoom=1.0):
        """Densities (real and k-space) and phases of both components of the
        current state (pspinor.py:853-887); saves
        ``spin_dens_phase<i>-<folder><ext>`` in the data folder when ``save``.
        Returns ``(fig, {'r', 'ph', 'k'})``."""
        extents = {'r': self._extent('r_sizes', rscale), 'k': self._extent('k_sizes', kscale)}
        return ptools.plot_spins(self.psi, self.psik, extents, self.paths, cmap=cmap,
                                 save=save, ext=ext, zoom=zoom)

    # ---------------------------------------------------------- propagation
    def _propagate(self, time, t_step, n_steps, device, is_sampling,
                   n_samples, **engine_opts):
        # Imported here: the propagator needs the native engine, the set-up
        # above does not.
        from spinor_gpe.pspinor import tensor_propagator as tprop
        prop = tprop.TensorPropagator(self, t_step, n_steps, device,
                                      time=time, is_sampling=is_sampling,
                                      n_samples=n_samples, **engine_opts)
        result = prop.prop_loop(prop.n_steps)
        result.paths = self.paths
        result.time_scale = self.time_scale
        result.space = self.space
        self.psik = result.psik
        self.psi = result.psi
        return result, prop

    # pylint: disable=too-many-arguments
    def imaginary(self, t_step, n_steps=1000, device='cpu',
                  is_sampling=False, n_samples=1, **engine_opts):
        """Imaginary-time propagation (pspinor.py:890-925).

        Returns ``(PropResult, TensorPropagator)`` and writes the final state
        back into ``self.psik`` / ``self.psi``.
        """
        return self._propagate('imag', t_step, n_steps, device, is_sampling,
                               n_samples, **engine_opts)

    def real(self, t_step, n_steps=1000, device='cpu', is_sampling=False,
             n_samples=1, **engine_opts):
        """Real-time propagation (pspinor.py:927-962)."""
        return self._propagate('real', t_step, n_steps, device, is_sampling,
                               n_samples, **engine_opts)
