"""Energy evaluations for the oracle. TEST INFRASTRUCTURE ONLY.

``eng_expect``: NumPy restatement of the reference's end-of-run energy
(spinor_gpe/pspinor/tensor_propagator.py:273-324 with tensor_tools.py:
314-367, 487-539), including its conventions: np.gradient with dr[0] as the
axis-0 spacing, the recoil term on the axis-0 phase derivative, interaction
terms without 1/2, a coupling term without its phase, no area element.

``unwrap_phase_py``: an independent pure-Python implementation of the
reliability-sorted unwrapping algorithm (Herraez et al., Appl. Opt. 41, 7437
(2002)) that scikit-image 0.16's unwrap_phase implements; for small grids.

``gp_energy``: the GP energy functional and chemical potential the engine
defines (the reference has none): spectral kinetic energy
sum_s sum_k T_s |psik_s|^2 dv_k, potential sum_s V_s n_s dv_r, interaction
(1/2) sum (g_uu n0^2 + g_dd n1^2 + 2 g_ud n0 n1) dv_r, coupling
sum Omega Re(psi0* psi1 e^{-i theta}) dv_r; mu = (E_kin + E_pot + E_cpl +
2 E_int) / N.
"""

import numpy as np


def ifft_2d_np(psik, dr):
    """tensor_tools.py:248-252 (NumPy branch)."""
    c = dr[0] * dr[1] / (2 * np.pi)
    return [np.fft.ifftn(np.fft.ifftshift(pk)) / c for pk in psik]


def unwrap_phase_py(ph):
    """Reliability-sorted path-following unwrap of a 2D float array."""
    ph = np.asarray(ph, dtype=np.float64)
    ny, nx = ph.shape

    def wr(d):
        return d - 2 * np.pi * (d > np.pi) + 2 * np.pi * (d < -np.pi)

    rel = np.full((ny, nx), 1.0e7)
    c = ph[1:-1, 1:-1]
    h = wr(ph[1:-1, :-2] - c) - wr(c - ph[1:-1, 2:])
    v = wr(ph[:-2, 1:-1] - c) - wr(c - ph[2:, 1:-1])
    d1 = wr(ph[:-2, :-2] - c) - wr(c - ph[2:, 2:])
    d2 = wr(ph[:-2, 2:] - c) - wr(c - ph[2:, :-2])
    rel[1:-1, 1:-1] = np.sqrt(h * h + v * v + d1 * d1 + d2 * d2)
    flat = ph.ravel()
    r = rel.ravel()
    edges = []
    for i in range(ny):
        for j in range(nx):
            p = i * nx + j
            for q in ((p + 1) if j + 1 < nx else None, (p + nx) if i + 1 < ny else None):
                if q is None:
                    continue
                d = flat[p] - flat[q]
                inc = 1 if d > np.pi else (-1 if d < -np.pi else 0)
                edges.append((r[p] + r[q], p, q, inc))
    edges.sort(key=lambda e: e[0])  # stable: ties keep generation order
    members = {p: [p] for p in range(ny * nx)}
    group = list(range(ny * nx))
    turns = np.zeros(ny * nx, dtype=np.int64)
    for _, p1, p2, inc in edges:
        g1, g2 = group[p1], group[p2]
        if g1 == g2:
            continue
        if len(members[g1]) > len(members[g2]):
            dst, src, shift = g1, g2, turns[p1] + inc - turns[p2]
        else:
            dst, src, shift = g2, g1, turns[p2] - inc - turns[p1]
        for q in members[src]:
            group[q] = dst
            turns[q] += shift
        members[dst].extend(members.pop(src))
    return (flat + 2 * np.pi * turns).reshape(ny, nx)


def phase_masked(psi_comp, dens, unwrap):
    """tensor_tools.py:528-539: angle, optional unwrap, zeroed where dens < 1e-6 max."""
    ang = np.angle(psi_comp)
    if unwrap:
        ang = unwrap(ang)
    ang[dens < (dens.max() * 1e-6)] = 0
    return ang


PLANTS = ('half_int', 'area', 'cpl_theta', 'recoil_axis1')


def eng_expect(psik, dr, pot_eng_spin, coupling, g_sc, kl_recoil, is_coupling,
               unwrap=unwrap_phase_py, plant=None, theta=None):
    """tensor_propagator.py:293-324 on NumPy arrays.

    ``plant`` (tests only) introduces one named mistake, so that the tests
    can show their closed forms catch it: 'half_int' (a 1/2 on the
    interaction terms), 'area' (the total times dv_r), 'cpl_theta' (the
    coupling term with its phase ``theta``), 'recoil_axis1' (the recoil term
    on the axis-1 phase derivative).
    """
    if plant is not None and plant not in PLANTS:
        raise ValueError(plant)
    dr = np.asarray(dr)
    psi = ifft_2d_np(psik, dr)
    dens = [np.abs(p) ** 2 for p in psi]
    dens_sqrt = [np.sqrt(d) for d in dens]
    phase = [phase_masked(p, d, unwrap) for p, d in zip(psi, dens)]
    grads = [np.gradient(p, *dr) for p in phase]
    phase_gradx = [g[1] if plant == 'recoil_axis1' else g[0] for g in grads]
    phase_gradsq = [g[0] ** 2 + g[1] ** 2 for g in grads]
    root_gradsq = []
    for d in dens_sqrt:
        g0, g1 = np.gradient(d, *dr)
        root_gradsq.append(g0 ** 2 + g1 ** 2)
    kin = (sum(root_gradsq)
           + sum([d * p for d, p in zip(dens, phase_gradsq)])
           + sum([d * p for d, p in zip(dens, phase_gradx)]) * 2 * kl_recoil * is_coupling) / 2
    pot = sum([d * v for d, v in zip(dens, pot_eng_spin)])
    int_e = g_sc['uu'] * dens[0] ** 2 + g_sc['dd'] * dens[1] ** 2 + g_sc['ud'] * dens[0] * dens[1]
    if plant == 'half_int':
        int_e = 0.5 * int_e
    eth = np.exp(-1j * np.asarray(theta)) if plant == 'cpl_theta' else 1.0
    coupl_e = (np.conj(psi[0]) * psi[1] * eth + np.conj(psi[1]) * psi[0] * np.conj(eth)) * np.asarray(coupling) / 2
    total = np.real((kin + pot + int_e + coupl_e).sum())
    if plant == 'area':
        total = total * dr[0] * dr[1]
    return [total, np.real(kin).sum(), np.real(pot).sum(), np.real(int_e).sum()]


def gp_energy(psik, dr, kin_eng_spin, pot_eng_spin, g_sc, coupling=None, theta=None):
    """GP functional terms of a k-space spinor (float64).

    Returns dict with E_tot, E_kin, E_pot, E_int, E_cpl, N, mu. ``coupling``
    is the Omega grid (None: no coupling); ``theta`` the coupling phase grid
    (None: rotated frame, theta = 0).
    """
    dr = np.asarray(dr, dtype=np.float64)
    ny, nx = psik[0].shape
    dv_r = dr[0] * dr[1]
    dv_k = (np.pi / (nx * dr[0] / 2)) * (np.pi / (ny * dr[1] / 2))
    psi = ifft_2d_np(psik, dr)
    nk = [np.abs(p) ** 2 for p in psik]
    n = [np.abs(p) ** 2 for p in psi]
    e_kin = sum(np.sum(t * d) for t, d in zip(kin_eng_spin, nk)) * dv_k
    e_pot = sum(np.sum(v * d) for v, d in zip(pot_eng_spin, n)) * dv_r
    e_int = 0.5 * np.sum(g_sc['uu'] * n[0] ** 2 + g_sc['dd'] * n[1] ** 2
                         + 2 * g_sc['ud'] * n[0] * n[1]) * dv_r
    e_cpl = 0.0
    if coupling is not None:
        eth = np.exp(-1j * theta) if theta is not None else 1.0
        e_cpl = np.sum(coupling * np.real(np.conj(psi[0]) * psi[1] * eth)) * dv_r
    atoms = np.sum(n[0] + n[1]) * dv_r
    e_tot = e_kin + e_pot + e_int + e_cpl
    return {'E_tot': e_tot, 'E_kin': e_kin, 'E_pot': e_pot, 'E_int': e_int, 'E_cpl': e_cpl,
            'N': atoms, 'mu': (e_kin + e_pot + e_cpl + 2 * e_int) / atoms}
