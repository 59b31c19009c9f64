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
