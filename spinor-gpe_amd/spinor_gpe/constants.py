"""Physical constants (SI) for the pseudospin-1/2 GPE setup.

Mirrors the values the reference defines in ``spinor_gpe/constants.py:13-35``:
CODATA values come from :mod:`scipy.constants`; the Rb-87 mass is the rounded
``87 * 1.66e-27`` kg and the scattering length is 100.4 Bohr radii.
"""

import scipy.constants as _sc

h = _sc.h
hbar = _sc.hbar
c = _sc.c
eps0 = _sc.epsilon_0
a0 = _sc.physical_constants['Bohr radius'][0]
e = _sc.elementary_charge

#: Atomic data for Rubidium-87 (constants.py:22-35).
Rb87 = {}
Rb87['m'] = 87 * 1.66e-27
Rb87['D2'] = 780.1e-9
Rb87['a_sc'] = 100.4 * a0
Rb87['g'] = 4 * _sc.pi * hbar ** 2 * Rb87['a_sc'] / Rb87['m']
Rb87['k_r'] = 2 * _sc.pi / Rb87['D2']
Rb87['E_r'] = (h / (2 * Rb87['m'])) * (1 / Rb87['D2']) ** 2
Rb87['D1_alpha'] = h * 1.22306e-5
Rb87['gJ_S12'] = 2.002331
Rb87['gJ_P12'] = 0.666
Rb87['gJ_P32'] = 1.3362
Rb87['delta_FS'] = 6.834682e9
