"""Pins of the oracle's propagation (tensor_propagator.py:67-271).

Closed forms: free-particle evolution, Rabi oscillations (real and imaginary
time), the harmonic ground state. Invariants: real-time unitarity of every
renormalized stage, time reversibility (which separates the reference's
reuse of the interaction operator from a recomputation, tensor_propagator.py:
262-267), second-order convergence in dt, the 2D virial identity. Each check
is made so that a plausible slip (a dropped factor, a sign, operator order,
recomputing U) fails it; the planted-mistake tests show that.
"""

import numpy as np
import pytest
import torch

from oracle import energy as oe
from oracle import reference_path as rp
from spinor_gpe.pspinor import pspinor as spin

W = 2 * np.pi * 50


def make(mesh=(64, 64), r=(8, 8), atoms=1e3, g=(1.0, 1.0, 0.995), pop=(0.5, 0.5), name='t'):
    return spin.PSpinor(name, omeg={'x': W, 'y': W, 'z': 40 * W},
                        g_sc={'uu': g[0], 'dd': g[1], 'ud': g[2]}, mesh_points=mesh,
                        r_sizes=r, atom_num=atoms, pop_frac=pop)


def set_psi(ps, psi):
    ps.psi = [np.asarray(p, dtype=np.complex128) for p in psi]
    c = ps.space['dv_r'] / (2 * np.pi)
    ps.psik = [np.fft.fftshift(np.fft.fftn(p) * c) for p in ps.psi]


def normalize_k(psik, ps):
    tot = sum(np.sum(np.abs(p) ** 2) for p in psik) * ps.space['dv_k']
    return [p * np.sqrt(ps.atom_num / tot) for p in psik]


def rel(a, b):
    a, b = np.stack(a), np.stack(b)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def packet(ps, k0=(1.5, -0.7), width=1.3, center=(0.5, -0.3)):
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    g = np.exp(-((x - center[0]) ** 2 + (y - center[1]) ** 2) / (2 * width ** 2))
    return g * np.exp(1j * (k0[0] * x + k0[1] * y))


# --------------------------------------------------------------- closed forms

@pytest.mark.parametrize('time', ['real', 'imag'])
def test_free_particle_exact(time):
    ps = make()
    ps.g_sc.update(uu=0.0, dd=0.0, ud=0.0)
    ps.pot_eng = np.zeros_like(ps.pot_eng)
    set_psi(ps, [packet(ps), 0.5 * packet(ps, k0=(-1.0, 0.2))])
    dt, n = 0.01, 25
    o = rp.OraclePropagator(ps, dt, time=time)
    o.run(n)
    t_tot = n * dt
    f = [np.exp(-1j * k * t_tot) if time == 'real' else np.exp(-k * t_tot) for k in ps.kin_eng_spin]
    want = normalize_k([fk * pk for fk, pk in zip(f, ps.psik)], ps)
    assert rel(o.psik_numpy(), want) < 1e-12


def test_rabi_real_time_exact():
    ps = make(pop=(1.0, 0.0))
    ps.g_sc.update(uu=0.0, dd=0.0, ud=0.0)
    ps.coupling_setup(kin_shift=False)
    omega = 7.0
    ps.coupling_uniform(omega)
    dt, n = 0.005, 60
    pops = rp.OraclePropagator(ps, dt, time='real').run(n)
    t = dt * np.arange(1, n + 1)
    want = ps.atom_num * np.stack([np.cos(omega * t / 2) ** 2, np.sin(omega * t / 2) ** 2], 1)
    assert np.max(np.abs(pops - want)) / ps.atom_num < 1e-12


def test_rabi_imaginary_time_exact():
    ps = make(pop=(1.0, 0.0))
    ps.g_sc.update(uu=0.0, dd=0.0, ud=0.0)
    ps.coupling_setup(kin_shift=False)
    omega = 3.0
    ps.coupling_uniform(omega)
    dt, n = 0.02, 40
    pops = rp.OraclePropagator(ps, dt, time='imag').run(n)
    tau = dt * np.arange(1, n + 1)
    want1 = np.sinh(omega * tau / 2) ** 2 / np.cosh(omega * tau)
    assert np.max(np.abs(pops[:, 1] / ps.atom_num - want1)) < 1e-12
    assert np.allclose(pops.sum(1), ps.atom_num, rtol=1e-14)


def test_harmonic_ground_state():
    ps = make(mesh=(64, 64), r=(8, 8))
    ps.g_sc.update(uu=0.0, dd=0.0, ud=0.0)
    o = rp.OraclePropagator(ps, 0.01, time='imag')
    o.run(1200)
    e = oe.gp_energy(o.psik_numpy(), ps.space['dr'], ps.kin_eng_spin, ps.pot_eng_spin, ps.g_sc)
    assert e['N'] == pytest.approx(ps.atom_num, rel=1e-13)
    assert e['mu'] == pytest.approx(1.0, abs=5e-5)
    psi = o.psi_numpy()[0]
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    gs = np.exp(-(x ** 2 + y ** 2) / 2)
    ov = abs(np.vdot(gs, psi)) ** 2 / (np.vdot(gs, gs).real * np.vdot(psi, psi).real)
    assert ov > 1 - 1e-6


# ------------------------------------------------------------- invariants

def coupled(name='c', frame_rot=False, g=(1.0, 1.0, 0.9), omega=4.0, delta=1.5):
    ps = make(g=g, name=name)
    ps.coupling_setup(scale=0.25, kin_shift=frame_rot)
    ps.rot_coupling = frame_rot
    ps.coupling_uniform(omega)
    ps.detuning_uniform(delta)
    set_psi(ps, [ps.psi[0] * np.exp(0.3j * ps.space['x_mesh']), ps.psi[1]])
    return ps


def test_real_time_renormalizations_are_identities(monkeypatch):
    ps = coupled()
    seen = []
    real_norm = rp.norm

    def spy(psi, vol, atoms):
        dens = rp.density(psi)
        seen.append((torch.sum(dens[0] + dens[1]) * vol / atoms).item())
        return real_norm(psi, vol, atoms)

    monkeypatch.setattr(rp, 'norm', spy)
    rp.OraclePropagator(ps, 0.002, time='real').run(10)
    assert len(seen) == 60
    assert max(abs(s - 1) for s in seen) < 1e-12


def reversal_error(ps, dt, n, prop=rp.OraclePropagator):
    fwd = prop(ps, dt, time='real')
    fwd.run(n)
    back = prop(ps, -dt, time='real')
    back.psik = fwd.psik
    back.run(n)
    return rel(back.psik_numpy(), normalize_k(ps.psik, ps))


class RecomputeU(rp.OraclePropagator):
    """Planted mistake: recompute U after C P C (the commented-out code at
    tensor_propagator.py:262-266)."""

    def single_step(self, t_step, eng):
        psik = [e * pk for e, pk in zip(eng['kin'], self.psik)]
        psi = rp.ifft_2d(psik, delta_r=self.space['dr'])
        psi, dens = rp.norm(psi, self.space['dv_r'], self.atom_num)

        def u_of(d):
            return rp.evolution_op(t_step / 2, [self.g_sc['uu'] * d[0] + self.g_sc['ud'] * d[1],
                                                self.g_sc['dd'] * d[1] + self.g_sc['ud'] * d[0]])

        psi = [op * p for op, p in zip(u_of(dens), psi)]
        psi = [sum([el * p for el, p in zip(row, psi)]) for row in eng['coupl']]
        psi = [e * p for e, p in zip(eng['pot'], psi)]
        psi = [sum([el * p for el, p in zip(row, psi)]) for row in eng['coupl']]
        psi = [op * p for op, p in zip(u_of(rp.density(psi)), psi)]
        psik = rp.fft_2d(psi, delta_r=self.space['dr'])
        psik = [e * pk for e, pk in zip(eng['kin'], psik)]
        self.psik, _ = rp.norm(psik, self.space['dv_k'], self.atom_num)


def test_reversibility_pins_interaction_reuse():
    # no coupling: each sub-step is exactly reversible
    ps = coupled(name='a')
    ps.is_coupling = False
    assert reversal_error(ps, 0.004, 15) < 1e-12
    # coupling without interactions: reversible
    ps = coupled(name='b', g=(1.0, 1.0, 0.9))
    ps.g_sc.update(uu=0.0, dd=0.0, ud=0.0)
    assert reversal_error(ps, 0.004, 15) < 1e-12
    # coupling and interactions: the reused U breaks reversibility at O(dt^2)
    ps = coupled(name='c')
    err_reuse = reversal_error(ps, 0.004, 15)
    err_recompute = reversal_error(ps, 0.004, 15, prop=RecomputeU)
    assert err_reuse > 1e-8
    assert err_recompute < 1e-12


def convergence_ratios(ps, prop=rp.OraclePropagator, t_end=0.08):
    def run(dt):
        o = prop(ps, dt, time='real')
        o.run(int(round(t_end / dt)))
        return o.psik_numpy()

    ref = run(t_end / 256)
    errs = [rel(run(t_end / n), ref) for n in (8, 16, 32)]
    return [errs[i] / errs[i + 1] for i in range(2)]


def test_convergence_order():
    """Second order in dt, except that reusing U after C P C (which changes the
    component densities) makes the coupled interacting scheme first order."""
    ps = coupled(name='o2', frame_rot=True)
    ps.g_sc.update(uu=0.0, dd=0.0, ud=0.0)
    assert all(3.5 < r < 4.5 for r in convergence_ratios(ps))
    ps = coupled(name='o1', frame_rot=True)
    assert all(1.8 < r < 2.4 for r in convergence_ratios(ps))
    assert all(3.5 < r < 4.5 for r in convergence_ratios(ps, prop=RecomputeU))


def test_virial_ground_state():
    ps = make(mesh=(64, 64), r=(8, 8), atoms=2e3, g=(1.0, 1.0, 0.5))
    o = rp.OraclePropagator(ps, 0.005, time='imag')
    o.run(1500)
    e = oe.gp_energy(o.psik_numpy(), ps.space['dr'], ps.kin_eng_spin, ps.pot_eng_spin, ps.g_sc)
    assert (e['E_kin'] + e['E_int']) == pytest.approx(e['E_pot'], rel=2e-4)
    assert e['E_int'] > 0.3 * e['E_pot']  # interactions matter in this state


# ----------------------------------------------------------- energy pins

def test_gp_energy_closed_forms():
    ps = make(mesh=(64, 64), r=(8, 8))
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    gs = np.exp(-(x ** 2 + y ** 2) / 2) + 0j
    set_psi(ps, [gs, gs])
    pk = normalize_k(ps.psik, ps)
    zero = {'uu': 0.0, 'dd': 0.0, 'ud': 0.0}
    e = oe.gp_energy(pk, ps.space['dr'], ps.kin_eng_spin, ps.pot_eng_spin, zero)
    assert e['E_kin'] == pytest.approx(ps.atom_num / 2, rel=1e-12)
    assert e['E_pot'] == pytest.approx(ps.atom_num / 2, rel=1e-12)
    assert e['mu'] == pytest.approx(1.0, rel=1e-12)
    # coupling energy sign: (phi, -phi) has E_cpl = -Omega N / 2
    set_psi(ps, [gs, -gs])
    pk = normalize_k(ps.psik, ps)
    omega = np.full_like(x, 2.5)
    e = oe.gp_energy(pk, ps.space['dr'], ps.kin_eng_spin, ps.pot_eng_spin, zero, coupling=omega)
    assert e['E_cpl'] == pytest.approx(-2.5 * ps.atom_num / 2, rel=1e-12)
    # interaction: uniform-density identity E_int = (1/2) g sum n^2 dv
    g = {'uu': 0.3, 'dd': 0.0, 'ud': 0.0}
    e = oe.gp_energy(pk, ps.space['dr'], ps.kin_eng_spin, ps.pot_eng_spin, g)
    n0 = np.abs(oe.ifft_2d_np(pk, ps.space['dr'])[0]) ** 2
    assert e['E_int'] == pytest.approx(0.5 * 0.3 * np.sum(n0 ** 2) * ps.space['dv_r'], rel=1e-13)


def test_eng_expect_gaussian():
    """Wrap-free state: finite-difference functional of tensor_propagator.py:308-324."""
    ps = make(mesh=(128, 128), r=(8, 8))
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    gs = np.exp(-(x ** 2 + y ** 2) / 2) + 0j
    set_psi(ps, [gs, gs])
    pk = normalize_k(ps.psik, ps)
    zero = {'uu': 0.0, 'dd': 0.0, 'ud': 0.0}
    e = oe.eng_expect(pk, ps.space['dr'], ps.pot_eng_spin, np.zeros_like(x), zero, 1.0, False)
    # Sum over pixels without the area element. Central differences with
    # spacing h scale each Fourier mode's k^2 by sin^2(kh)/(kh)^2, so for the
    # Gaussian the kinetic term is (N/2)(1 - h^2/2 + O(h^4)); the potential
    # term is N/2.
    h = ps.space['dr'][0]
    assert e[1] * ps.space['dv_r'] == pytest.approx(ps.atom_num / 2 * (1 - h ** 2 / 2), rel=2e-4)
    assert e[2] * ps.space['dv_r'] == pytest.approx(ps.atom_num / 2, rel=1e-10)
    assert e[0] == pytest.approx(e[1] + e[2], rel=1e-14)
    assert e[3] == 0.0


def test_unwrap_recovers_ramp():
    ny, nx = 24, 32
    yy, xx = np.mgrid[0:ny, 0:nx]
    true = 0.9 * xx - 0.4 * yy + 0.02 * xx * yy / 10
    wrapped = np.angle(np.exp(1j * true))
    got = oe.unwrap_phase_py(wrapped)
    d = got - true
    assert np.allclose(d, d.flat[0], atol=1e-12)
    assert abs(d.flat[0] / (2 * np.pi) - round(d.flat[0] / (2 * np.pi))) < 1e-12
    small = 0.3 * np.sin(xx / 5.0) * np.cos(yy / 7.0)
    np.testing.assert_array_equal(oe.unwrap_phase_py(small), small)


# ------------------------------------------------------- frames of the Raman coupling

WAVEL = 790.1e-9


def gauge_pair(kl_rot_sign, time, seed=21):
    """A lab-frame instance (theta = 2 kL x, tensor_propagator.py:126-129;
    unshifted dispersions) and a rotated-frame one (theta = 0, dispersions
    shifted by +/- kx kL', pspinor.py:493-501) with kL' = kl_rot_sign * kL,
    the rotated state being the lab state times exp(+i kL x) (component 0)
    and exp(-i kL x) (component 1). kL is a multiple of the grid's dk, so
    the phase factors are exact cyclic shifts in k; the grid is fine enough
    (k_max = 16 kL) that the state carries no weight where the shift wraps."""
    pss = []
    for i, (sign, frame) in enumerate(((1.0, 'lab'), (kl_rot_sign, 'rot'))):
        ps = make(mesh=(256, 256), r=(8, 8), atoms=2e3, name=f'gauge-{frame}-{kl_rot_sign}-{time}-{seed}')
        dk = ps.space['dk'][0]
        scale = 8 * dk / (np.sqrt(2) * np.pi / WAVEL * ps.a_x)
        ps.coupling_setup(wavel=WAVEL, scale=sign * scale, kin_shift=(frame == 'rot'))
        ps.coupling_uniform(3.0)
        ps.detuning_grad(0.4, 0.2, axis=0)  # breaks x -> -x symmetry
        ps.rot_coupling = frame == 'rot'
        pss.append(ps)
    lab, rot = pss
    kl = 8 * lab.space['dk'][0]
    assert lab.kL_recoil == pytest.approx(kl)
    # smooth (Gaussian) components: negligible weight near the k-grid edge,
    # where the cyclic shift by kL would wrap
    rng = np.random.default_rng(seed)
    psi = [packet(lab, k0=tuple(rng.uniform(-0.5, 0.5, 2)), width=1.2,
                  center=tuple(rng.uniform(-0.5, 0.5, 2))) for _ in range(2)]
    x = lab.space['x_mesh']
    set_psi(lab, psi)
    set_psi(rot, [psi[0] * np.exp(1j * kl * x), psi[1] * np.exp(-1j * kl * x)])
    return lab, rot, kl


@pytest.mark.parametrize('time', ['real', 'imag'])
def test_lab_and_rotated_frames_are_one_gauge(time):
    """The reference's two Raman frames are related by a gauge transform:
    the lab frame with recoil +kL equals the rotated frame with recoil -kL
    (the rotated frame's dispersion shift, T_0 = T + kx kL, pairs with the
    lab frame's exp(-i theta) on the (0, 1) coupling entry only with the
    opposite recoil sign). This pins both sign conventions together: a
    flipped theta or a flipped shift breaks it, and with the recoils of equal
    sign the frames disagree."""
    dt = 1e-3 if time == 'real' else 5e-3
    lab, rot, kl = gauge_pair(-1.0, time)
    x = lab.space['x_mesh']
    pl = rp.OraclePropagator(lab, dt, time=time)
    pr = rp.OraclePropagator(rot, dt, time=time)
    pl.run(40)
    pr.run(40)
    psi_l = oe.ifft_2d_np(pl.psik_numpy(), lab.space['dr'])
    psi_r = oe.ifft_2d_np(pr.psik_numpy(), rot.space['dr'])
    want = [psi_l[0] * np.exp(1j * kl * x), psi_l[1] * np.exp(-1j * kl * x)]
    assert rel(psi_r, want) < 1e-9
    # the same recoil sign in both frames is a different system
    if time == 'imag':
        return
    lab2, rot2, _ = gauge_pair(+1.0, time)
    pl2 = rp.OraclePropagator(lab2, dt, time=time)
    pr2 = rp.OraclePropagator(rot2, dt, time=time)
    pl2.run(10)
    pr2.run(10)
    psi_l2 = oe.ifft_2d_np(pl2.psik_numpy(), lab2.space['dr'])
    psi_r2 = oe.ifft_2d_np(pr2.psik_numpy(), rot2.space['dr'])
    want2 = [psi_l2[0] * np.exp(1j * kl * x), psi_l2[1] * np.exp(-1j * kl * x)]
    assert rel(psi_r2, want2) > 1e-3
