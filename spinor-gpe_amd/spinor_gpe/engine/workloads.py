"""Seeded instance sets: Raman-coupled ground-state problems on mixed grids.

Every instance is a :class:`PSpinor` (pspinor.py:21-196) with the 50 Hz
isotropic trap (omega_z = 40 omega_x), a +/-16 oscillator-length box, a
Thomas-Fermi start split 0.5/0.5, and Raman coupling with recoil kL = 3 in the
rotated frame (``coupling_setup(kin_shift=True)``, pspinor.py:469-501). The
drawn parameters are the coupling Omega ~ U[0, 4] E_R, the detuning
delta ~ U[-1, 1] E_R, the atom number 10^U[3, 5] and g_ud / g_uu ~ U[0.9, 1.1]
(g_uu = g_dd = 1 before the reference's g scaling). All draws come from
``numpy.random.default_rng(seed)``.
"""

import dataclasses

import numpy as np

from spinor_gpe.pspinor import pspinor as spin

WAVEL = 790.1e-9
KL = 3.0
W50 = 2 * np.pi * 50
SWEEP_COUNTS = ((128, 32), (256, 48), (512, 16))


@dataclasses.dataclass
class Draw:
    """Parameters of one instance (energies in units of E_R = kL^2 / 2)."""
    grid: int
    omega_er: float
    delta_er: float
    atom_num: float
    g_ratio: float


def draw_params(rng, grid):
    """One seeded parameter draw for a ``grid`` x ``grid`` instance."""
    return Draw(grid, float(rng.uniform(0.0, 4.0)), float(rng.uniform(-1.0, 1.0)),
                float(10 ** rng.uniform(3.0, 5.0)), float(rng.uniform(0.9, 1.1)))


def raman_spinor(name, d, delta_shift_er=0.0):
    """The PSpinor of a draw (fresh data folder ``name``, overwritten)."""
    ps = spin.PSpinor(name, omeg={'x': W50, 'y': W50, 'z': 40 * W50},
                      g_sc={'uu': 1.0, 'dd': 1.0, 'ud': d.g_ratio},
                      mesh_points=(d.grid, d.grid), r_sizes=(16, 16),
                      atom_num=d.atom_num, pop_frac=(0.5, 0.5), overwrite=True)
    scale = KL / (np.sqrt(2) * np.pi / WAVEL * ps.a_x)
    ps.coupling_setup(wavel=WAVEL, scale=scale, kin_shift=True)
    e_r = ps.EL_recoil
    ps.coupling_uniform(d.omega_er * e_r)
    ps.detuning_uniform((d.delta_er + delta_shift_er) * e_r)
    ps.rot_coupling = True
    return ps


def sweep_draws(seed=2, counts=SWEEP_COUNTS):
    """The draws of a mixed-grid sweep: ``counts`` = ((grid, how many), ...)."""
    rng = np.random.default_rng(seed)
    return [draw_params(rng, g) for g, n in counts for _ in range(n)]


def c0_spinor(name='c0', seed=0):
    """The 256^2 imaginary-time ground-state problem: +/-16 oscillator
    lengths, isotropic 50 Hz trap, 1e4 atoms, g = (1, 1, 0.995), Thomas-Fermi
    start times seeded uniform phase noise of +/-0.1 rad (run with dt = 0.02
    for 4000 steps)."""
    ps = spin.PSpinor(name, omeg={'x': W50, 'y': W50, 'z': 40 * W50},
                      g_sc={'uu': 1.0, 'dd': 1.0, 'ud': 0.995}, mesh_points=(256, 256),
                      r_sizes=(16, 16), atom_num=1e4, pop_frac=(0.5, 0.5), overwrite=True)
    rng = np.random.default_rng(seed)
    ph = rng.uniform(-0.1, 0.1, (2, 256, 256))
    ps.psi = [p * np.exp(1j * f) for p, f in zip(ps.psi, ph)]
    c = ps.space['dv_r'] / (2 * np.pi)
    ps.psik = [np.fft.fftshift(np.fft.fftn(p) * c) for p in ps.psi]
    return ps


def c1_spinor(psi_ground, space_ground, name='c1', seed=1):
    """The 512^2 real-time Raman problem: +/-24 oscillator lengths, the same
    trap and g, kL = 3, Omega = 2 E_R, delta = 0.5 E_R, lab frame
    (theta = 2 kL x); the initial state is ``psi_ground`` (a 256^2 ground
    state on ``space_ground``) resampled bilinearly to the 512^2 grid,
    renormalized, times seeded Gaussian phase noise (sigma 0.05). Run with
    dt = 1e-3 for 20000 steps."""
    from scipy.interpolate import RegularGridInterpolator  # pylint: disable=import-outside-toplevel
    ps = spin.PSpinor(name, omeg={'x': W50, 'y': W50, 'z': 40 * W50},
                      g_sc={'uu': 1.0, 'dd': 1.0, 'ud': 0.995}, mesh_points=(512, 512),
                      r_sizes=(24, 24), atom_num=1e4, pop_frac=(0.5, 0.5), overwrite=True)
    scale = KL / (np.sqrt(2) * np.pi / WAVEL * ps.a_x)
    ps.coupling_setup(wavel=WAVEL, scale=scale, kin_shift=False)
    ps.coupling_uniform(2.0 * ps.EL_recoil)
    ps.detuning_uniform(0.5 * ps.EL_recoil)
    ps.rot_coupling = False
    pts = np.stack([ps.space['y_mesh'].ravel(), ps.space['x_mesh'].ravel()], axis=1)
    rng = np.random.default_rng(seed)
    psi = []
    for p in psi_ground:
        f = [RegularGridInterpolator((space_ground['y'], space_ground['x']), part,
                                     bounds_error=False, fill_value=0.0) for part in (p.real, p.imag)]
        v = (f[0](pts) + 1j * f[1](pts)).reshape(512, 512)
        psi.append(v * np.exp(1j * rng.normal(0.0, 0.05, (512, 512))))
    tot = sum(np.sum(np.abs(p) ** 2) for p in psi) * ps.space['dv_r']
    ps.psi = [p * np.sqrt(ps.atom_num / tot) for p in psi]
    c = ps.space['dv_r'] / (2 * np.pi)
    ps.psik = [np.fft.fftshift(np.fft.fftn(p) * c) for p in ps.psi]
    return ps


def c4_spinor(name='c4', seed=4, n_vortices=64, mesh=2048, r_size=64):
    """The large real-time problem: ``mesh``^2 over +/-``r_size`` oscillator
    lengths, 1e6 atoms, anisotropic trap 30 x 40 Hz (omega_z = 2 pi 1200 Hz),
    Raman kL = 3 in the lab frame with Omega = 1.5 E_R and delta = 0, and
    ``n_vortices`` phase-imprinted vortices (seed_vortices, pspinor.py:
    680-745) on component 0 only, at seeded positions uniform in a disc of
    0.7 Thomas-Fermi radii with seeded charges +/-1. Run with dt = 5e-4 for
    10000 steps."""
    w = 2 * np.pi
    ps = spin.PSpinor(name, omeg={'x': w * 30, 'y': w * 40, 'z': w * 1200},
                      g_sc={'uu': 1.0, 'dd': 1.0, 'ud': 0.995}, mesh_points=(mesh, mesh),
                      r_sizes=(r_size, r_size), atom_num=1e6, pop_frac=(0.5, 0.5), overwrite=True)
    scale = KL / (np.sqrt(2) * np.pi / WAVEL * ps.a_x)
    ps.coupling_setup(wavel=WAVEL, scale=scale, kin_shift=False)
    ps.coupling_uniform(1.5 * ps.EL_recoil)
    ps.detuning_uniform(0.0)
    ps.rot_coupling = False
    rng = np.random.default_rng(seed)
    rad = 0.7 * ps.rad_tf * np.sqrt(rng.uniform(0.0, 1.0, n_vortices))
    ang = rng.uniform(0.0, 2 * np.pi, n_vortices)
    pos = np.stack([rad * np.cos(ang), rad * np.sin(ang)], axis=1)
    charges = rng.choice([-1, 1], n_vortices)
    ps.seed_vortices(pos, charges, components=(0,))
    return ps
