"""Seeded test configurations built through the drop-in PSpinor API.

Shapes follow the benchmark configurations at sizes the CPU oracle finishes
in seconds: a Thomas-Fermi start with seeded phase noise, Raman coupling in
the lab frame (theta = 2 kL x) and in the rotated frame (shifted
dispersions), detuning, non-separable potentials and coupling gradients,
non-square grids.
"""

import numpy as np

from spinor_gpe.pspinor import pspinor as spin

W = 2 * np.pi * 50
WAVEL = 790.1e-9


def _noise(ps, seed, amp=0.1, gaussian=False):
    rng = np.random.default_rng(seed)
    shape = (2,) + ps.psi[0].shape
    ph = rng.normal(0, amp, shape) if gaussian else rng.uniform(-amp, amp, shape)
    ps.psi = [p * np.exp(1j * f) for p, f in zip(ps.psi, ph)]
    c = ps.space['dv_r'] / (2 * np.pi)
    ps.psik = [np.fft.fftshift(np.fft.fftn(p) * c) for p in ps.psi]


def base(name, mesh=(128, 128), r=(16, 16), atoms=1e4, g=(1.0, 1.0, 0.995), trap_y=1.0):
    return spin.PSpinor(name, omeg={'x': W, 'y': trap_y * W, 'z': 40 * W},
                        g_sc={'uu': g[0], 'dd': g[1], 'ud': g[2]}, mesh_points=mesh,
                        r_sizes=r, atom_num=atoms, pop_frac=(0.5, 0.5))


def kl3_scale(ps):
    """coupling_setup scale giving kL = 3 (inverse oscillator lengths)."""
    return 3.0 / (np.sqrt(2) * np.pi / WAVEL * ps.a_x)


def ground(name='gs', mesh=(128, 128), r=(16, 16), seed=0):
    """C0 recipe on a smaller grid: TF + uniform phase noise, no coupling."""
    ps = base(name, mesh, r)
    _noise(ps, seed)
    return ps, 'imag', 0.02


def raman_lab(name='lab', mesh=(128, 128), r=(24, 24), seed=1):
    """C1 recipe: kL = 3, Omega = 2 E_R, delta = 0.5 E_R, lab frame."""
    ps = base(name, mesh, r)
    ps.coupling_setup(wavel=WAVEL, scale=kl3_scale(ps), kin_shift=False)
    ps.coupling_uniform(2.0 * ps.EL_recoil)
    ps.detuning_uniform(0.5 * ps.EL_recoil)
    ps.rot_coupling = False
    _noise(ps, seed, amp=0.05, gaussian=True)
    return ps, 'real', 1e-3


def raman_rot(name='rot', mesh=(64, 64), r=(16, 16), seed=2):
    """Rotated frame: Raman-shifted dispersions, theta = 0, imaginary time."""
    ps = base(name, mesh, r, atoms=3e3)
    ps.coupling_setup(wavel=WAVEL, scale=kl3_scale(ps), kin_shift=True)
    ps.coupling_uniform(1.2 * ps.EL_recoil)
    ps.detuning_uniform(-0.3 * ps.EL_recoil)
    _noise(ps, seed)
    return ps, 'imag', 0.02


def dense_ops(name='dense', mesh=(128, 64), r=(16, 8), seed=3):
    """Non-separable potential, coupling gradient (dense Omega), detuning
    gradient, anisotropic trap, non-square grid, real time."""
    ps = base(name, mesh, r, atoms=2e3, trap_y=1.5)
    ps.coupling_setup(wavel=WAVEL, scale=0.3, kin_shift=False)
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    ps.pot_eng = ps.pot_eng + 3.0 * np.exp(-((x - 1.0) ** 2 + (y + 0.5) ** 2) / 2)
    ps.detuning_grad(0.2, 0.1, axis=1)
    ps.coupling_grad(0.3, 2.0, axis=0)
    ps.coupling = np.abs(ps.coupling)
    ps.rot_coupling = False
    _noise(ps, seed)
    return ps, 'real', 2e-3


def tall(name='tall', seed=4):
    """Non-square grid with Ny > Nx, imaginary time, coupling off."""
    ps = base(name, mesh=(64, 256), r=(8, 32), atoms=2e3, trap_y=0.25)
    _noise(ps, seed)
    return ps, 'imag', 0.01


def dense_kin(name='densekin', mesh=(64, 64), r=(8, 8), seed=5):
    """Non-separable kinetic energy (a kx*ky cross term) with Raman coupling
    in the rotated frame, imaginary time: the dense-kinetic kernels."""
    ps = base(name, mesh, r, atoms=1e3)
    ps.coupling_setup(wavel=WAVEL, scale=0.5, kin_shift=True)
    ps.coupling_uniform(1.5)
    kx, ky = ps.space['kx_mesh'], ps.space['ky_mesh']
    ps.kin_eng_spin = [k + 0.05 * kx * ky for k in ps.kin_eng_spin]
    ps.kin_eng_spin = [k - np.min(k) for k in ps.kin_eng_spin]
    _noise(ps, seed)
    return ps, 'imag', 5e-3


CASES = {'ground': ground, 'raman_lab': raman_lab, 'raman_rot': raman_rot,
         'dense_ops': dense_ops, 'dense_kin': dense_kin, 'tall': tall}
