"""Grids without compiled power-of-two passes (GPU): the runtime-length
mixed-radix kernels.

The reference only requires even mesh sizes (pspinor.py:331-332). Lengths
that are not powers of two from 64 to 4096 (or a power-of-two axis none of
whose compiled tiles divides the other axis) run on LDS-resident
mixed-radix Stockham kernels (radices 8, 4, 2, 3, 5, 7 and, for other prime
factors, a runtime radix). Checked here against the CPU oracle on the same
seeded inputs, to the tolerances of the power-of-two parity tests:

* fft_2d / ifft_2d vs NumPy, including prime factors 11 and 13 and the
  2-point axis;
* prop_loop in imaginary and real time (uniform and lab-frame coupling,
  dense potentials and coupling arrays, a dense kinetic operator), with
  the end-of-run energy from the device evaluation;
* plans mixing a compiled pass on one axis with a runtime-length pass on the
  other, a power-of-two axis forced onto the runtime-length kernels, a
  batched two-slot plan, a batched sweep (concurrent plans in one captured
  graph with on-device energy checks) and the shape report.
"""

import math

import numpy as np
import pytest
import torch

import cases
from oracle import energy as oe
from oracle.reference_path import OraclePropagator
from spinor_gpe import _native as nat
from spinor_gpe.engine.plan import InstanceSpec, Plan
from spinor_gpe.pspinor import tensor_tools as tt

pytestmark = pytest.mark.gpu

TOL = {'c128': (1e-11, 1e-11), 'c64': (2e-4, 1e-5)}


def rel(a, b):
    a, b = np.stack(a), np.stack(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# (ny, nx) arrays
FFT_SHAPES = [(96, 96), (96, 160), (160, 96), (32, 32), (2, 6), (22, 26), (90, 128), (128, 90),
              (1536, 1536), (2048, 96), (96, 2048), (4096, 90), (210, 198)]


@pytest.mark.parametrize('dtype', [torch.complex128, torch.complex64])
@pytest.mark.parametrize('shape', FFT_SHAPES)
def test_fft2_matches_numpy(gpu, shape, dtype):
    if dtype == torch.complex64 and shape == (4096, 90):
        pytest.skip("4096-point complex64 columns have no compiled 2-column tile")
    print(f"[grid-fft2] shape={shape} dtype={dtype}", flush=True)
    rng = np.random.default_rng(sum(shape))
    psi = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(2)]
    dr = np.array([0.3, 0.2])
    want = tt.fft_2d(psi, dr)
    dev = [torch.as_tensor(p, dtype=dtype, device=gpu) for p in psi]
    got = [g.cpu().numpy() for g in tt.fft_2d(dev, torch.tensor(dr, device=gpu))]
    tol = (4e-15 if dtype == torch.complex128 else 2e-6) * 2 * math.log2(shape[0] * shape[1])
    assert rel(got, want) < tol
    kdev = [torch.as_tensor(p, dtype=dtype, device=gpu) for p in want]
    back = [g.cpu().numpy() for g in tt.ifft_2d(kdev, dr)]
    assert rel(back, psi) < tol


def test_odd_and_oversized_grids_refused(gpu):
    for ny, nx in ((96, 95), (2050, 64), (64, 3000)):
        with pytest.raises(nat.SgpeError, match='even and at most 2048'):
            Plan(gpu, 'c128', ny, nx, 1)


# (case, mesh (nx, ny), r sizes, steps)
RUNS = [
    ('ground', (96, 96), (16, 16), 50),
    ('ground', (160, 96), (20, 12), 40),
    ('ground', (32, 32), (8, 8), 60),
    ('ground', (130, 66), (16, 8), 30),
    ('ground', (90, 128), (12, 16), 30),
    ('ground', (128, 96), (16, 12), 40),
    ('ground', (96, 128), (12, 16), 40),
    ('raman_lab', (96, 160), (16, 24), 60),
    ('raman_rot', (48, 80), (16, 16), 50),
    ('dense_ops', (120, 60), (16, 8), 50),
    ('dense_kin', (48, 48), (8, 8), 40),
]


def _build(case, mesh, r):
    return getattr(cases, case)(name=f'g-{case}-{mesh[0]}x{mesh[1]}', mesh=mesh, r=r)


@pytest.mark.parametrize('precision', ['c128', 'c64'])
@pytest.mark.parametrize('case,mesh,r,n', RUNS)
def test_prop_loop_matches_oracle(gpu, case, mesh, r, n, precision):
    ps, time, dt = _build(case, mesh, r)
    print(f"[grid-parity] case={case} mesh={mesh} precision={precision} steps={n}", flush=True)
    oracle = OraclePropagator(ps, dt, time=time)
    ref_pops = oracle.run(n)
    runner = ps.imaginary if time == 'imag' else ps.real
    res, prop = runner(dt, n, 'cuda', precision=precision, progress=False, block_steps=16)
    tol_psi, tol_pop = TOL[precision]
    err = rel(res.psik, oracle.psik_numpy())
    perr = float(np.max(np.abs(res.pops['vals'] - ref_pops)) / ps.atom_num)
    assert err < tol_psi, err
    assert perr < tol_pop, perr
    e_dev = res.energy_gp
    cpl = ps.coupling if ps.is_coupling else None
    theta = 2 * ps.kL_recoil * ps.space['x_mesh'] if (ps.is_coupling and not ps.rot_coupling) else None
    e_own = oe.gp_energy(res.psik, ps.space['dr'], ps.kin_eng_spin, ps.pot_eng_spin, ps.g_sc, cpl, theta)
    tol_e = 1e-12 if precision == 'c128' else 1e-5
    for key in ('E_tot', 'mu'):
        assert e_dev[key] == pytest.approx(e_own[key], rel=tol_e), key
    assert e_dev['N'] == pytest.approx(ps.atom_num, rel=1e-12 if precision == 'c128' else 1e-6)


def test_shape_report(gpu):
    """Runtime-length axes report E = 0; compiled axes their shapes."""
    # (ny, nx): (row pass on runtime-length kernels, column pass on them);
    # 128 x 90: no compiled 128-point column tile divides 90
    want = {(96, 96): (True, True), (96, 128): (False, True), (128, 96): (True, False),
            (128, 90): (True, True), (128, 128): (False, False)}
    for (ny, nx), (gen_x, gen_y) in want.items():
        plan = Plan(gpu, 'c128', ny, nx, 1)
        row_e, row_rows, col_e, col_w = plan.tuning()
        assert (row_e == 0) == gen_x and (col_e == 0) == gen_y, (ny, nx, plan.tuning())
        assert ny % row_rows == 0 and nx % col_w == 0


@pytest.mark.parametrize('prec', ['c128', 'c64'])
def test_batched_plan(gpu, prec):
    """Two instances of one 96 x 160 plan, stepped together and with one
    sub-step of one slot alone, vs the oracle."""
    (ps0, t0, dt0) = cases.raman_rot(name='gb0', mesh=(160, 96), r=(20, 12))
    (ps1, t1, dt1) = cases.raman_rot(name='gb1', mesh=(160, 96), r=(20, 12), seed=9)
    plan = Plan(gpu, prec, 96, 160, 2)
    states = [plan.new_state(ps0.psik), plan.new_state(ps1.psik)]
    plan.set_slot(0, InstanceSpec.from_spin(ps0, t0, dt0), states[0])
    plan.set_slot(1, InstanceSpec.from_spin(ps1, t1, dt1), states[1])
    plan.full_steps(3)
    plan.set_active(1, False)
    plan.single_step(nat.SUB_INNER)
    plan.set_active(1, True)
    plan.full_steps(2)
    plan.sync()
    o0 = OraclePropagator(ps0, dt0, time=t0)
    o0.run(3)
    o0.single_step(o0.dt_in, o0.eng_in)
    for _ in range(2):
        o0.full_step()
    o1 = OraclePropagator(ps1, dt1, time=t1)
    o1.run(5)
    tol = 1e-12 if prec == 'c128' else 2e-5
    for st, o in zip(states, (o0, o1)):
        err = rel(st.cpu().numpy(), np.stack(o.psik_numpy()))
        assert err <= tol, err


@pytest.mark.parametrize('precision', ['c128', 'c64'])
def test_sweep_on_runtime_length_grids(gpu, precision):
    """A batched sweep over grids on the runtime-length kernels (96^2, 160x96)
    and a compiled one (64^2), one captured graph with concurrent plans and
    on-device energy checks: 60 steps of every instance against the oracle
    (c128 1e-11, c64 2e-4), and the device GP energy against the oracle's
    functional on the engine's own state."""
    from spinor_gpe.engine.sweep import Sweep
    spins = [cases.raman_rot(name=f'gs{i}', mesh=mesh, r=r, seed=20 + i)[0]
             for i, (mesh, r) in enumerate([((96, 96), (16, 16)), ((160, 96), (20, 12)),
                                            ((64, 64), (16, 16)), ((96, 96), (12, 12))])]
    sw = Sweep(gpu, precision, conv_tol=0.0, check_every=20, max_steps=60)
    for ps in spins:
        sw.add_spin(ps, 0.02)
    res, _ = sw.run()
    tol = 1e-11 if precision == 'c128' else 2e-4
    for ps, r in zip(spins, res):
        assert r.steps == 60
        oracle = OraclePropagator(ps, 0.02, time='imag')
        oracle.run(60)
        err = rel(r.psik, oracle.psik_numpy())
        assert err < tol, (r.grid, err)
        cpl = ps.coupling if ps.is_coupling else None
        e_own = oe.gp_energy(r.psik, ps.space['dr'], ps.kin_eng_spin, ps.pot_eng_spin, ps.g_sc, cpl, None)
        assert r.energy['E_tot'] == pytest.approx(e_own['E_tot'], rel=1e-12 if precision == 'c128' else 1e-5)
