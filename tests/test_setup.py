"""PSpinor set-up (host NumPy) against closed forms of pspinor.py:113-745."""

import numpy as np
import pytest

from spinor_gpe import constants as const
from spinor_gpe.pspinor import pspinor as spin

W = 2 * np.pi * 50


def c0(**kw):
    args = dict(omeg={'x': W, 'y': W, 'z': 40 * W}, g_sc={'uu': 1.0, 'dd': 1.0, 'ud': 0.995},
                mesh_points=(256, 256), r_sizes=(16, 16), atom_num=1e4, pop_frac=(0.5, 0.5))
    args.update(kw)
    return spin.PSpinor('c0', **args)


def test_scales_match_closed_forms():
    ps = c0()
    a_x = np.sqrt(const.hbar / (87 * 1.66e-27 * W))
    assert ps.a_x == pytest.approx(a_x, rel=1e-15)
    assert ps.a_x == pytest.approx(1.524577e-6, rel=1e-6)
    assert ps.a_sc == pytest.approx(3.48486e-3, rel=1e-5)
    g_scale = np.sqrt(8 * 40 * np.pi) * ps.a_sc
    assert g_scale == pytest.approx(0.1104931, rel=1e-6)
    assert ps.g_sc == pytest.approx({'uu': g_scale, 'dd': g_scale, 'ud': 0.995 * g_scale}, rel=1e-15)
    assert ps.chem_pot == pytest.approx(18.75395, rel=1e-6)
    assert ps.rad_tf == pytest.approx(6.1244, rel=1e-4)
    assert ps.time_scale == 1 / W


def test_g_sc_is_rescaled_in_place():
    g = {'uu': 1.0, 'dd': 1.0, 'ud': 0.5}
    ps = c0(g_sc=g)
    assert g is ps.g_sc and g['ud'] == pytest.approx(0.5 * g['uu'])
    assert g['uu'] != 1.0


def test_grids():
    ps = spin.PSpinor('g', mesh_points=(128, 64), r_sizes=(8, 4))
    sp = ps.space
    assert sp['dr'].tolist() == [0.125, 0.125]
    assert sp['dk'].tolist() == pytest.approx([np.pi / 8, np.pi / 4])
    assert sp['x'][0] == -8 and sp['x'][-1] == 8 - 0.125 and len(sp['x']) == 128
    assert sp['x_mesh'].shape == (64, 128)
    assert sp['kx'][64] == 0.0 and sp['kx'][0] == pytest.approx(-np.pi / 0.125)
    assert sp['dv_r'] == 0.125 ** 2 and sp['dv_k'] == pytest.approx(np.pi ** 2 / 32)
    np.testing.assert_array_equal(ps.kin_eng, (sp['kx_mesh'] ** 2 + sp['ky_mesh'] ** 2) / 2)


def test_odd_mesh_rejected():
    with pytest.raises(AssertionError):
        spin.PSpinor('odd', mesh_points=(63, 64))


def test_existing_path_requires_overwrite():
    spin.PSpinor('same', mesh_points=(64, 64))
    with pytest.raises(FileExistsError):
        spin.PSpinor('same', mesh_points=(64, 64))
    spin.PSpinor('same', mesh_points=(64, 64), overwrite=True)


def test_thomas_fermi_state():
    ps = c0(pop_frac=(0.7, 0.3))
    dv = ps.space['dv_r']
    pops = [np.sum(np.abs(p) ** 2) * dv for p in ps.psi]
    assert sum(pops) == pytest.approx(1e4, rel=1e-13)
    assert pops[0] / pops[1] == pytest.approx(0.7 / 0.3, rel=1e-12)
    # inverted paraboloid: zero outside the TF radius
    r2 = ps.space['x_mesh'] ** 2 + ps.space['y_mesh'] ** 2
    assert np.all(ps.psi[0][r2 > 2 * ps.chem_pot] == 0)
    # psik is the reference-convention transform of psi
    c = ps.space['dv_r'] / (2 * np.pi)
    np.testing.assert_allclose(ps.psik[0], np.fft.fftshift(np.fft.fftn(ps.psi[0]) * c), atol=1e-12)
    heal = (8 * np.pi * np.max(np.abs(ps.psi[0]) ** 2) * ps.a_sc) ** -0.5
    assert ps.heal[0] == pytest.approx(heal)
    saved = np.load(ps.paths['trial'] + 'tf_wf-c0.npz')
    np.testing.assert_array_equal(saved['psik'][1], ps.psik[1])


def test_coupling_setup_and_detuning():
    ps = c0(mesh_points=(64, 64), r_sizes=(8, 8))
    ps.coupling_setup(wavel=790.1e-9, scale=1.0, kin_shift=True)
    kl = np.sqrt(2) * np.pi / 790.1e-9 * ps.a_x
    assert ps.kL_recoil == pytest.approx(kl, rel=1e-15)
    assert ps.EL_recoil == pytest.approx(kl ** 2 / 2)
    for s, sign in ((0, 1), (1, -1)):
        raw = ps.kin_eng + sign * ps.space['kx_mesh'] * kl
        np.testing.assert_allclose(ps.kin_eng_spin[s], raw - raw.min(), rtol=0, atol=1e-12)
        assert ps.kin_eng_spin[s].min() == 0
    ps.detuning_uniform(2.0)
    np.testing.assert_array_equal(ps.pot_eng_spin[0], ps.pot_eng + 1.0)
    np.testing.assert_array_equal(ps.pot_eng_spin[1], ps.pot_eng - 1.0)
    ps.coupling_uniform(3.0)
    assert np.all(ps.coupling == 3.0)
    with pytest.raises(AssertionError):
        ps.coupling_uniform(-1.0)
    ps.coupling_grad(0.5, 1.0, axis=0)
    np.testing.assert_array_equal(ps.coupling, ps.space['x_mesh'] * 0.5 + 1.0)
    ps.detuning_grad(0.25, axis=1)
    np.testing.assert_array_equal(ps.pot_eng_spin[0], ps.pot_eng + ps.space['y_mesh'] * 0.125)
    # assigning the potential resets the spin landscapes (pspinor.py:392-396)
    ps.pot_eng = np.zeros_like(ps.pot_eng)
    assert np.all(ps.pot_eng_spin[0] == 0) and np.all(ps.pot_eng_spin[1] == 0)


def test_seed_vortices_winding():
    ps = c0(mesh_points=(128, 128), r_sizes=(8, 8))
    ps.seed_vortices([(1.0, 0.5)], [1])
    # circulation of the phase around the core is +2 pi on component 0
    x, y = ps.space['x'], ps.space['y']
    ix0, iy0 = np.argmin(abs(x - 1.0)), np.argmin(abs(y - 0.5))
    ring = [(iy0 - 3, ix0 - 3), (iy0 - 3, ix0 + 3), (iy0 + 3, ix0 + 3), (iy0 + 3, ix0 - 3)]
    path = []
    for (a, b), (c, d) in zip(ring, ring[1:] + ring[:1]):
        n = max(abs(c - a), abs(d - b))
        for t in range(n):
            path.append((a + (c - a) * t // n, b + (d - b) * t // n))
    ang = np.angle([ps.psi[0][p] for p in path])
    circ = np.sum(np.angle(np.exp(1j * np.diff(np.r_[ang, ang[0]]))))
    assert circ == pytest.approx(-2 * np.pi, abs=1e-6) or circ == pytest.approx(2 * np.pi, abs=1e-6)
    assert abs(ps.psi[0][iy0, ix0]) < 0.2 * np.max(abs(ps.psi[0]))


def test_shift_momentum_moves_peak():
    ps = c0(mesh_points=(128, 128), r_sizes=(16, 16))
    ps.coupling_setup(scale=0.4)
    shift = ps.kL_recoil / ps.space['dk'][0]
    norms = [np.sum(np.abs(p) ** 2) for p in ps.psik]

    def centroid(p):
        w = np.abs(p).sum(axis=0) ** 2
        return np.sum(np.arange(len(w)) * w) / np.sum(w)

    c_before = [centroid(p) for p in ps.psik]
    ps.shift_momentum(scale=1.0, frac=(1.0, 0.0))
    # frac alternates between components (pspinor.py:539-543): component 0
    # moves by +kL, component 1 by -kL; the Fourier shift is unitary.
    c_after = [centroid(p) for p in ps.psik]
    assert c_after[0] - c_before[0] == pytest.approx(shift, abs=0.1)
    assert c_after[1] - c_before[1] == pytest.approx(-shift, abs=0.1)
    assert [np.sum(np.abs(p) ** 2) for p in ps.psik] == pytest.approx(norms, rel=1e-13)
    np.testing.assert_allclose(ps.psi[0], np.fft.ifftn(np.fft.ifftshift(ps.psik[0]))
                               / (ps.space['dv_r'] / (2 * np.pi)), atol=1e-12)


@pytest.mark.parametrize('s', [3, 4])
def test_shift_momentum_integer_shift_closed_form(s):
    """For a whole-pixel shift s the Fourier shift of pspinor.py:534-545 is
    exact: with the transform pair of tensor_tools.py:201-258 around it, each
    component's k-array becomes (-1)^s (f_a roll(psik, +s) + f_b roll(psik, -s))
    along x, where (f_a, f_b) = frac for component 0 and its flip for
    component 1 (the sign comes from applying fftfreq phases to an
    fftshifted array). A random state, so that mirrored or unflipped shifts
    cannot pass."""
    ps = c0(mesh_points=(64, 128), r_sizes=(8, 16))
    ps.coupling_setup(scale=0.4)
    rng = np.random.default_rng(5)
    p = [rng.normal(size=(64, 128)) + 1j * rng.normal(size=(64, 128)) for _ in range(2)]
    frac = (0.3, 0.7)
    ps.shift_momentum(psik=[q.copy() for q in p], scale=s * ps.space['dk'][0] / ps.kL_recoil, frac=frac)
    sign = (-1) ** s
    want = [sign * (frac[0] * np.roll(p[0], s, axis=1) + frac[1] * np.roll(p[0], -s, axis=1)),
            sign * (frac[1] * np.roll(p[1], s, axis=1) + frac[0] * np.roll(p[1], -s, axis=1))]
    for got, w in zip(ps.psik, want):
        np.testing.assert_allclose(got, w, rtol=0, atol=1e-13 * np.max(np.abs(w)))


def test_shift_momentum_subpixel_composes_to_whole_pixel():
    """Sub-pixel shifts form a one-parameter group anchored at the exact
    whole-pixel case: with frac = (1, 0) component 0 is shifted by +a and
    component 1 by -a, and two half-pixel (or four quarter-pixel) shifts
    equal the whole-pixel result (-1) roll(psik, +-1), whatever the
    interpolation between pixels."""
    ps = c0(mesh_points=(64, 128), r_sizes=(8, 16))
    ps.coupling_setup(scale=0.4)
    rng = np.random.default_rng(11)
    p = [rng.normal(size=(64, 128)) + 1j * rng.normal(size=(64, 128)) for _ in range(2)]
    unit = ps.space['dk'][0] / ps.kL_recoil  # scale of a one-pixel shift
    want = [-np.roll(p[0], 1, axis=1), -np.roll(p[1], -1, axis=1)]
    for parts in (2, 4):
        q = [x.copy() for x in p]
        for _ in range(parts):
            ps.shift_momentum(psik=q, scale=unit / parts, frac=(1.0, 0.0))
            q = [x.copy() for x in ps.psik]
        for got, w in zip(q, want):
            np.testing.assert_allclose(got, w, rtol=0, atol=1e-12 * np.max(np.abs(w)))


def test_shift_momentum_subpixel_matches_published_filter():
    """scipy.ndimage.fourier_shift (the routine pspinor.py:534-545 calls;
    1.6.2 pinned by the reference, another version here) against its
    published definition, restated with NumPy: the input is taken as a
    Fourier transform along the axis and multiplied by
    exp(-2 pi i s fftfreq(N)). A 0.37-pixel shift, where a wrong frequency
    convention (non-negative frequencies k/N, say) would differ, around the
    transform pair of tensor_tools.py:201-258."""
    ps = c0(mesh_points=(64, 128), r_sizes=(8, 16))
    ps.coupling_setup(scale=0.4)
    rng = np.random.default_rng(13)
    p = [rng.normal(size=(64, 128)) + 1j * rng.normal(size=(64, 128)) for _ in range(2)]
    s = 0.37
    frac = (0.3, 0.7)
    ps.shift_momentum(psik=[q.copy() for q in p], scale=s * ps.space['dk'][0] / ps.kL_recoil, frac=frac)
    c = ps.space['dv_r'] / (2 * np.pi)
    ramp = np.exp(-2j * np.pi * s * np.fft.fftfreq(128))[None, :]
    fr = [frac, frac[::-1]]
    for i in range(2):
        spec = np.fft.fftshift(np.fft.fftn(p[i])) * c
        mixed = fr[i][0] * spec * ramp + fr[i][1] * spec * np.conj(ramp)
        want = np.fft.ifftn(np.fft.ifftshift(mixed)) / c
        np.testing.assert_allclose(ps.psik[i], want, rtol=0, atol=1e-12 * np.max(np.abs(want)))


@pytest.mark.parametrize('w', [1, -1])
def test_seed_vortices_profile_and_orientation(w):
    """The imprint of pspinor.py:731-743 on a grid point (x0, y0): k pixels
    along +x the state is scaled by rho/sqrt(rho^2 + 1), rho = k dx / xi, with
    no phase; k pixels along +y by the same modulus and the phase w*pi/2;
    along -x by the phase pi. This fixes the winding sign, the atan2(dy, dx)
    orientation and the healing-length scaling that a closed loop's
    circulation alone would not."""
    ps = c0(mesh_points=(128, 128), r_sizes=(8, 8))
    x, y = ps.space['x'], ps.space['y']
    ix0, iy0 = 70, 60
    before = [p.copy() for p in ps.psi]
    ps.seed_vortices([(x[ix0], y[iy0])], [w])
    dx, dy = ps.space['dr']
    for i in range(2):
        with np.errstate(divide='ignore', invalid='ignore'):  # zero outside the Thomas-Fermi radius
            ratio = ps.psi[i] / before[i]
        assert ratio[iy0, ix0] == 0
        for k in (1, 3, 8):
            for (jy, jx), dist, phase in (((iy0, ix0 + k), k * dx, 0.0), ((iy0 + k, ix0), k * dy, w * np.pi / 2),
                                          ((iy0, ix0 - k), k * dx, np.pi)):
                rho = dist / ps.heal[i]
                assert abs(ratio[jy, jx]) == pytest.approx(rho / np.sqrt(rho ** 2 + 1), rel=1e-12)
                assert np.angle(ratio[jy, jx] * np.exp(-1j * phase)) == pytest.approx(0.0, abs=1e-12)
    c = ps.space['dv_r'] / (2 * np.pi)
    np.testing.assert_allclose(ps.psik[1], np.fft.fftshift(np.fft.fftn(ps.psi[1])) * c, atol=1e-12)
