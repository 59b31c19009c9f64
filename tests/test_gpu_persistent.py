"""The persistent step kernel (one launch per block of full steps) against
the CPU oracle and the pass launches (GPU).

Every compiled persistent shape runs a single instance (state address as a
kernel argument) and a two-slot plan (slot table path) for three full steps,
eagerly and as captured graphs, each slot compared with its own oracle run:
complex128 to 1e-12, complex64 to 2e-5 relative L2. prop_loop with the
persistent path matches the oracle's states and populations; a slot disabled
between blocks is skipped by the work list; a wait bound of one tick makes
the launch stop, which the plan reports as an error, and rebinding the state
restores normal operation.
"""

import numpy as np
import pytest

from cases import ground, raman_lab
from oracle.reference_path import OraclePropagator
from spinor_gpe import _native as nat
from spinor_gpe.engine.plan import InstanceSpec, Plan

pytestmark = pytest.mark.gpu

# (grid, row E, rows per block, column E, columns per block): the compiled
# persistent shapes (sgpe_persist.h)
SHAPES = [(128, 4, 4, 8, 8), (128, 2, 2, 4, 4), (256, 2, 1, 8, 4), (256, 4, 2, 8, 4),
          (512, 4, 2, 8, 4), (512, 2, 1, 8, 4), (512, 4, 2, 4, 2), (512, 2, 1, 4, 2),
          (512, 4, 1, 8, 2), (512, 8, 2, 8, 2), (1024, 4, 1, 8, 2), (1024, 4, 1, 16, 4)]
_REF = {}


def reference(n, seed):
    if (n, seed) not in _REF:
        if seed == 1:
            ps, time, dt = raman_lab(name=f'pers{n}', mesh=(n, n), r=(24, 24))
        else:
            ps, _, _ = raman_lab(name=f'pers{n}s{seed}', mesh=(n, n), r=(24, 24), seed=seed)
            time, dt = 'imag', 0.01
        o = OraclePropagator(ps, dt, time=time)
        o.run(3)
        _REF[(n, seed)] = (ps, time, dt, np.stack(o.psik_numpy()))
    return _REF[(n, seed)]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize('graph', [False, True])
@pytest.mark.parametrize('slots', [1, 2])
@pytest.mark.parametrize('prec', ['c128', 'c64'])
@pytest.mark.parametrize('shape', SHAPES)
def test_persistent_matches_oracle(gpu, shape, prec, slots, graph):
    n = shape[0]
    print(f"[persistent] n={n} shape={shape[1:]} {prec} slots={slots} graph={graph}", flush=True)
    runs = [reference(n, 1)] + ([reference(n, 8)] if slots == 2 else [])
    plan = Plan(gpu, prec, n, n, slots)
    plan.tune(*shape[1:])
    plan.set_persistent(True)
    req, used, grid = plan.persistent()
    assert req == 1 and grid > 0
    states = []
    for slot, (ps, time, dt, _) in enumerate(runs):
        st = plan.new_state(ps.psik)
        plan.set_slot(slot, InstanceSpec.from_spin(ps, time, dt), st)
        states.append(st)
    assert plan.persistent()[1] == 1
    if graph:
        plan.graph_steps(3, 1)
    else:
        plan.full_steps(3)
    plan.sync()
    plan.check()
    tol = 1e-12 if prec == 'c128' else 2e-5
    for slot, (st, run) in enumerate(zip(states, runs)):
        err = rel(st.cpu().numpy(), run[3])
        assert err < tol, (slot, err)


@pytest.mark.parametrize('prec', ['c128', 'c64'])
def test_persistent_prop_loop_matches_oracle(gpu, prec):
    ps, time, dt = raman_lab(name=f'perloop{prec}', mesh=(512, 512), r=(24, 24))
    oracle = OraclePropagator(ps, dt, time=time)
    ref_pops = oracle.run(150)
    res, prop = ps.real(dt, 150, 'cuda', precision=prec, progress=False, block_steps=50, persistent=True)
    assert prop._plan.persistent()[1] == 1
    tol_psi, tol_pop = (1e-11, 1e-11) if prec == 'c128' else (2e-4, 1e-5)
    assert rel(np.stack(res.psik), np.stack(oracle.psik_numpy())) <= tol_psi
    assert np.max(np.abs(res.pops['vals'] - ref_pops)) / ps.atom_num <= tol_pop


def test_persistent_skips_inactive_slot(gpu):
    (ps0, t0, dt0), (ps1, _, _) = ground(name='pskip0', mesh=(256, 256)), ground(name='pskip1', mesh=(256, 256), seed=3)
    plan = Plan(gpu, 'c128', 256, 256, 2)
    plan.tune(4, 2, 8, 4)
    plan.set_persistent(True)
    states = [plan.new_state(ps0.psik), plan.new_state(ps1.psik)]
    plan.set_slot(0, InstanceSpec.from_spin(ps0, t0, dt0), states[0])
    plan.set_slot(1, InstanceSpec.from_spin(ps1, t0, dt0), states[1])
    plan.graph_steps(2, 1)
    plan.set_active(1, False)
    frozen = states[1].clone()
    plan.graph_steps(2, 2)
    plan.sync()
    plan.check()
    assert bool((states[1] == frozen).all())
    o0 = OraclePropagator(ps0, dt0, time=t0)
    o0.run(6)
    o1 = OraclePropagator(ps1, dt0, time=t0)
    o1.run(2)
    assert rel(states[0].cpu().numpy(), np.stack(o0.psik_numpy())) <= 1e-12
    assert rel(states[1].cpu().numpy(), np.stack(o1.psik_numpy())) <= 1e-12


def test_persistent_wait_bound_reports_and_recovers(gpu):
    ps, time, dt = raman_lab(name='pabort', mesh=(512, 512), r=(24, 24))
    plan = Plan(gpu, 'c128', 512, 512, 1)
    plan.tune(4, 2, 8, 4)
    st = plan.new_state(ps.psik)
    plan.set_slot(0, InstanceSpec.from_spin(ps, time, dt), st)
    # a bound of one tick: the first wait of the launch runs out, every
    # workgroup returns, and the plan reports the stop
    plan.set_persistent(True, wait_ticks=1)
    plan.full_steps(5)
    with pytest.raises(nat.SgpeError, match='persistent'):
        plan.check()
    # new state bound, normal bound: runs and matches the oracle
    plan.set_persistent(True)
    st2 = plan.new_state(ps.psik)
    plan.set_state(0, st2)
    plan.full_steps(3)
    plan.sync()
    plan.check()
    o = OraclePropagator(ps, dt, time=time)
    o.run(3)
    assert rel(st2.cpu().numpy(), np.stack(o.psik_numpy())) <= 1e-12
