"""Closed-form pins of the reference's end-of-run energy (eng_expect,
tensor_propagator.py:293-324), for the oracle's restatement and for the
engine's host implementation alike.

Each check builds a state on which the reference's functional has a value
fixed by arithmetic that does not repeat its formula:

* a uniform state (constant moduli and phases) on a grid with dx != dy:
  no gradients, so the kinetic term is 0; the interaction term is
  N_pix (g_uu n0^2 + g_dd n1^2 + g_ud n0 n1) -- no 1/2, no area element --
  and the coupling term N_pix Omega a0 a1 cos(chi) with chi the relative phase;
* a Gaussian pair (phi, +-phi) in the lab frame (theta = 2 kL x): the
  coupling term is +-Omega sum phi^2, because the reference leaves theta out;
* phase ramps exp(i q y) and exp(i q x) of uniform modulus: np.gradient with
  dr[0] as the axis-0 spacing returns q dy/dx along y (resp. q dx/dy along x)
  exactly at every pixel, so the kinetic term is (1/2) N_pix sum_s n_s
  (q'^2 + 2 kL q') for the y-ramp (the recoil term is on the axis-0
  derivative) and (1/2) N_pix sum_s n_s q'^2 for the x-ramp;
* a chirp exp(i (alpha y^2 + beta x^2)) that wraps several times along both
  axes without residues: any correct unwrap returns the phase up to a global
  2 pi multiple, and the finite differences of a square are known in closed
  form, so the kinetic term is fixed; leaving the unwrap out, or unwrapping
  along rows only, fails it.

The planted-mistake tests run each check on the oracle with one plausible
slip built in (a 1/2 on the interaction, the area element, the coupling phase
kept, the recoil on the x derivative) and require the check to fail.
"""

import numpy as np
import pytest

from oracle import energy as oe
from spinor_gpe.pspinor import pspinor as spin
from spinor_gpe.pspinor.tensor_propagator import eng_expect_arrays

W = 2 * np.pi * 50


def make(mesh=(48, 32), r=(6.0, 8.0), g=(0.7, 0.4, 0.9)):
    """A PSpinor with dx = 0.25 != dy = 0.5 (mesh (Nx, Ny), sizes (Rx, Ry))."""
    ps = spin.PSpinor('ee', omeg={'x': W, 'y': 1.3 * W, 'z': 40 * W},
                      g_sc={'uu': 1.0, 'dd': 1.0, 'ud': 1.0}, mesh_points=mesh,
                      r_sizes=r, atom_num=1e3, pop_frac=(0.5, 0.5), overwrite=True)
    ps.g_sc = {'uu': g[0], 'dd': g[1], 'ud': g[2]}  # unscaled, unequal strengths
    return ps


def to_k(ps, psi):
    c = ps.space['dv_r'] / (2 * np.pi)
    return [np.fft.fftshift(np.fft.fftn(np.asarray(p, dtype=np.complex128)) * c) for p in psi]


def run_oracle(ps, psik, coupling, kl, is_coupling, plant=None):
    theta = 2 * kl * ps.space['x_mesh']
    return oe.eng_expect(psik, ps.space['dr'], ps.pot_eng_spin, coupling, ps.g_sc, kl, is_coupling,
                         plant=plant, theta=theta)


def run_engine(ps, psik, coupling, kl, is_coupling):
    return eng_expect_arrays(psik, ps.space['dr'], ps.pot_eng_spin, coupling, ps.g_sc, kl, is_coupling)


IMPLS = {'oracle': run_oracle, 'engine': run_engine}


def coupling_part(e):
    """E_cpl = E_tot - kin - pot - int (eng_expect does not return it)."""
    return e[0] - e[1] - e[2] - e[3]


# ------------------------------------------------------------------ checks

def check_uniform(fn):
    ps = make()
    dx, dy = ps.space['dr']
    assert dx != dy
    ny, nx = ps.space['x_mesh'].shape
    npix = nx * ny
    a0, a1, chi = 1.7, 0.6, 0.9
    psi = [np.full((ny, nx), a0 + 0j), np.full((ny, nx), a1 * np.exp(1j * chi))]
    omega = 2.3
    coupling = np.full((ny, nx), omega)
    e = fn(ps, to_k(ps, psi), coupling, 1.1, True)
    g = ps.g_sc
    n0, n1 = a0 ** 2, a1 ** 2
    assert abs(e[1]) <= 1e-18 * npix, e[1]
    assert e[2] == pytest.approx(n0 * np.sum(ps.pot_eng_spin[0]) + n1 * np.sum(ps.pot_eng_spin[1]), rel=1e-12)
    assert e[3] == pytest.approx(npix * (g['uu'] * n0 ** 2 + g['dd'] * n1 ** 2 + g['ud'] * n0 * n1), rel=1e-12)
    assert coupling_part(e) == pytest.approx(npix * omega * a0 * a1 * np.cos(chi), rel=1e-10)
    total = e[2] + e[3] + npix * omega * a0 * a1 * np.cos(chi)
    assert e[0] == pytest.approx(total, rel=1e-12)


def check_coupling_pair(fn):
    ps = make(mesh=(64, 64), r=(8.0, 8.0))
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    phi = np.exp(-(x ** 2 + 1.5 * y ** 2) / 3.0)
    omega = 1.9
    coupling = np.full_like(x, omega)
    kl = 2.0  # lab frame: theta = 2 kL x varies over the cloud
    for sign in (1.0, -1.0):
        e = fn(ps, to_k(ps, [phi, sign * phi]), coupling, kl, True)
        assert coupling_part(e) == pytest.approx(sign * omega * np.sum(phi ** 2), rel=1e-9)


def check_ramp_y(fn):
    ps = make()
    dx, dy = ps.space['dr']
    y = ps.space['y_mesh']
    ny, nx = y.shape
    b0, b1, q, kl = 1.2, 0.8, 1.3, 1.7
    psi = [b0 * np.exp(1j * q * y), b1 * np.exp(1j * (q * y + 0.4))]
    e = fn(ps, to_k(ps, psi), np.zeros_like(y), kl, True)
    qp = q * dy / dx  # axis-0 derivative with dr[0] = dx as its spacing
    want = 0.5 * nx * ny * (b0 ** 2 + b1 ** 2) * (qp ** 2 + 2 * kl * qp)
    assert e[1] == pytest.approx(want, rel=1e-10)
    assert coupling_part(e) == pytest.approx(0.0, abs=1e-9 * abs(want))


def check_ramp_x(fn):
    ps = make()
    dx, dy = ps.space['dr']
    x = ps.space['x_mesh']
    ny, nx = x.shape
    b0, b1, q, kl = 0.9, 1.4, 1.1, 1.7
    psi = [b0 * np.exp(1j * q * x), b1 * np.exp(-1j * q * x)]
    e = fn(ps, to_k(ps, psi), np.zeros_like(x), kl, True)
    qp = q * dx / dy  # axis-1 derivative with dr[1] = dy as its spacing
    want = 0.5 * nx * ny * (b0 ** 2 + b1 ** 2) * qp ** 2
    assert e[1] == pytest.approx(want, rel=1e-10)


def fd_of_square(c, a, h):
    """np.gradient (edge_order 1, spacing h) of a*c**2 sampled on the uniform
    coordinates c, from the algebra of differences of squares: interior
    (a c[i+1]^2 - a c[i-1]^2) / 2h = 2 a c[i] d / h with d the coordinate
    step; first and last points a (c[0] + c[1]) d / h and
    a (c[-2] + c[-1]) d / h."""
    d = c[1] - c[0]
    out = 2 * a * c * d / h
    out[0] = a * (c[0] + c[1]) * d / h
    out[-1] = a * (c[-2] + c[-1]) * d / h
    return out


def check_chirp(fn):
    """Uniform moduli with the curved phase alpha y^2 + beta x^2, which wraps
    several times along both axes (ranges ~19 and ~18 rad) with neighbour
    steps below pi: the field has no residues, so every correct unwrapper
    returns it up to one global multiple of 2 pi, which no gradient sees.
    The kinetic term is then (1/2) sum_s n_s sum_pix (g0^2 + g1^2 + 2 kL g0)
    with g0, g1 the finite differences of the two quadratics (axis 0 with
    dr[0] = dx as its spacing)."""
    ps = make()
    dx, dy = ps.space['dr']
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    ny, nx = x.shape
    alpha, beta, kl, b0, b1 = 0.3, 0.5, 1.7, 1.1, 0.7
    phase = alpha * y ** 2 + beta * x ** 2
    assert np.ptp(alpha * y[:, 0] ** 2) > 4 * np.pi and np.ptp(beta * x[0] ** 2) > 4 * np.pi
    assert np.max(np.abs(np.diff(phase, axis=0))) < 0.9 * np.pi
    assert np.max(np.abs(np.diff(phase, axis=1))) < 0.9 * np.pi
    psi = [b0 * np.exp(1j * phase), b1 * np.exp(1j * (phase + 2.0))]
    e = fn(ps, to_k(ps, psi), np.zeros_like(x), kl, True)
    g0 = fd_of_square(y[:, 0], alpha, dx)[:, None] * np.ones((1, nx))
    g1 = fd_of_square(x[0], beta, dy)[None, :] * np.ones((ny, 1))
    want = 0.5 * (b0 ** 2 + b1 ** 2) * np.sum(g0 ** 2 + g1 ** 2 + 2 * kl * g0)
    assert e[1] == pytest.approx(want, rel=1e-9)


CHECKS = {'uniform': check_uniform, 'coupling_pair': check_coupling_pair,
          'ramp_y': check_ramp_y, 'ramp_x': check_ramp_x, 'chirp': check_chirp}


@pytest.mark.parametrize('impl', sorted(IMPLS))
@pytest.mark.parametrize('check', sorted(CHECKS))
def test_eng_expect_closed_form(impl, check):
    CHECKS[check](IMPLS[impl])


PLANT_CAUGHT_BY = {'half_int': ['uniform'], 'area': ['uniform'],
                   'cpl_theta': ['coupling_pair'], 'recoil_axis1': ['ramp_y', 'ramp_x', 'chirp']}


@pytest.mark.parametrize('plant', oe.PLANTS)
def test_planted_mistakes_fail(plant):
    def planted(ps, psik, coupling, kl, is_coupling):
        return run_oracle(ps, psik, coupling, kl, is_coupling, plant=plant)
    for name in PLANT_CAUGHT_BY[plant]:
        with pytest.raises(AssertionError):
            CHECKS[name](planted)


@pytest.mark.parametrize('unwrap', ['none', 'rows_only'])
def test_chirp_catches_a_missing_or_partial_unwrap(unwrap):
    """The chirp check fails for the oracle with the unwrap left out, or
    done along rows only (numpy's 1-D unwrap on axis 1), so it pins the
    two-dimensional unwrap the reference's functional relies on."""
    fn = None if unwrap == 'none' else (lambda a: np.unwrap(a, axis=1))

    def variant(ps, psik, coupling, kl, is_coupling):
        return oe.eng_expect(psik, ps.space['dr'], ps.pot_eng_spin, coupling, ps.g_sc, kl, is_coupling,
                             unwrap=fn, theta=2 * kl * ps.space['x_mesh'])
    with pytest.raises(AssertionError):
        check_chirp(variant)


def test_engine_and_oracle_agree_on_a_noisy_state():
    """Same numbers on a generic wrap-free state (Gaussian times smooth
    phase noise, coupling on)."""
    ps = make(mesh=(64, 64), r=(8.0, 8.0))
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    rng = np.random.default_rng(5)
    env = np.exp(-(x ** 2 + y ** 2) / 4.0)
    psi = [env * np.exp(0.3j * np.sin(x + rng.uniform())), 0.7 * env * np.exp(0.2j * np.cos(y))]
    coupling = 1.5 + 0.1 * x
    pk = to_k(ps, psi)
    want = run_oracle(ps, pk, coupling, 1.3, True)
    got = run_engine(ps, pk, coupling, 1.3, True)
    np.testing.assert_allclose(got, want, rtol=1e-12)
