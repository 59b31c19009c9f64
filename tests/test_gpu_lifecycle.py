"""Run-state and slot lifecycle of plans (GPU).

* After ``prop_loop`` the propagator no longer holds the address of its
  population buffer: a later ``full_step`` must not write into whatever
  tensor reuses that memory.
* Disabling a slot while a sub-step is pending brings it to normalized
  k-space first: a slot disabled between two full steps, while the other
  slot runs an inner sub-step, continues exactly as if it had run the two
  full steps alone; the other slot matches the oracle's full step, inner
  sub-step, full step.
* A slot whose tables grow keeps its previous physics when the new device
  allocation fails (NOMEM), and the plan still steps correctly.
* device='cpu', the reference's default, runs on the GPU and hands back host
  tensors.
* A compacted plan (running slots first, launched alone) steps on exactly as
  the uncompacted one.
"""

import numpy as np
import pytest
import torch

from cases import ground, raman_lab, raman_rot
from oracle.reference_path import OraclePropagator
from spinor_gpe import _native as nat
from spinor_gpe.engine.plan import InstanceSpec, Plan

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_prop_loop_detaches_population_buffer(gpu):
    ps, _, dt = ground(name='detach', mesh=(64, 64))
    n = 10
    _, prop = ps.imaginary(dt, n, 'cuda', progress=False)
    torch.cuda.synchronize()
    # tensors of the population buffer's size: one of them reuses its block
    canaries = [torch.full((n, 2), 12345.0, dtype=torch.float64, device=gpu) for _ in range(8)]
    torch.cuda.synchronize()
    prop.full_step()
    prop.single_step(prop.dt_out, prop.eng_out)
    _ = prop.psik
    torch.cuda.synchronize()
    for c in canaries:
        assert torch.all(c == 12345.0)


@pytest.mark.parametrize('prec', ['c128', 'c64'])
def test_disable_slot_mid_run(gpu, prec):
    (ps0, t0, dt0), (ps1, t1, dt1) = raman_rot(name='dis0'), raman_rot(name='dis1', seed=7)
    plan = Plan(gpu, prec, 64, 64, 2)
    states = [plan.new_state(ps0.psik), plan.new_state(ps1.psik)]
    plan.set_slot(0, InstanceSpec.from_spin(ps0, t0, dt0), states[0])
    plan.set_slot(1, InstanceSpec.from_spin(ps1, t1, dt1), states[1])
    plan.full_steps(1)
    plan.set_active(1, False)
    plan.single_step(nat.SUB_INNER)
    plan.set_active(1, True)
    plan.full_steps(1)
    plan.sync()
    o0 = OraclePropagator(ps0, dt0, time=t0)
    o0.full_step()
    o0.single_step(o0.dt_in, o0.eng_in)
    o0.full_step()
    o1 = OraclePropagator(ps1, dt1, time=t1)
    o1.run(2)
    tol = 1e-12 if prec == 'c128' else 2e-5
    for st, o in zip(states, (o0, o1)):
        err = rel(st.cpu().numpy(), np.stack(o.psik_numpy()))
        assert err <= tol, err


def test_failed_table_growth_keeps_slot(gpu):
    ps, time, dt = raman_lab(name='grow', mesh=(64, 64), r=(16, 16))
    spec = InstanceSpec.from_spin(ps, time, dt)
    plan = Plan(gpu, 'c128', 64, 64, 1)
    st = plan.new_state(ps.psik)
    plan.set_slot(0, spec, st)
    # a dense potential needs a larger table buffer
    dense = InstanceSpec.from_spin(ps, time, dt)
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    dense.pot = [v + 0.5 * np.exp(-((x - 1) ** 2 + y ** 2)) for v in dense.pot]
    nat.check(nat.lib().sgpe_debug_fail_allocs(1))
    try:
        with pytest.raises(MemoryError):
            plan.set_slot(0, dense, st)
    finally:
        nat.check(nat.lib().sgpe_debug_fail_allocs(0))
    plan.full_steps(3)
    plan.sync()
    o = OraclePropagator(ps, dt, time=time)
    o.run(3)
    assert rel(st.cpu().numpy(), np.stack(o.psik_numpy())) <= 1e-12


def test_cpu_device_argument_runs_on_the_gpu(gpu):
    """The reference's default device='cpu' (pspinor.py:890, 927): the engine
    runs on the GPU, warns, and presents host tensors."""
    ps, time, dt = ground(name='cpudev', mesh=(64, 64))
    oracle = OraclePropagator(ps, dt, time=time)
    ref_pops = oracle.run(20)
    with pytest.warns(RuntimeWarning, match='HIP'):
        res, prop = ps.imaginary(dt, 20, progress=False)
    assert all(p.device.type == 'cpu' for p in prop.psik)
    assert prop.space['dr'].device.type == 'cpu'
    assert rel(np.stack(res.psik), np.stack(oracle.psik_numpy())) <= 1e-12
    assert np.max(np.abs(res.pops['vals'] - ref_pops)) <= 1e-12 * ps.atom_num


def test_compact_continues_bit_identically(gpu):
    """Four slots, two of them stopped after 4 steps; the plan is compacted
    to the two running slots (reordered) and steps on. The running slots'
    states, steps and energies equal those of an uncompacted plan, bit for
    bit; the stopped ones keep theirs; bad permutations and an active slot
    left out of the launches are refused."""
    spins = [raman_rot(name=f'cmp{i}', seed=30 + i) for i in range(4)]

    def make():
        plan = Plan(gpu, 'c128', 64, 64, 4)
        st = [plan.new_state(ps.psik) for ps, _, _ in spins]
        for k, (ps, t, dt) in enumerate(spins):
            plan.set_slot(k, InstanceSpec.from_spin(ps, t, dt), st[k])
        plan.graph_steps(4, 1)
        plan.energy()
        plan.set_active(0, False)
        plan.set_active(2, False)
        return plan, st

    ref, ref_st = make()
    ref.graph_steps(4, 2)
    e_ref = ref.energy().cpu().numpy()
    ref.sync()
    plan, st = make()
    with pytest.raises(nat.SgpeError, match='permutation'):
        plan.compact([3, 1, 1, 0], 2)
    with pytest.raises(nat.SgpeError, match='still active'):
        plan.compact([3, 0, 1, 2], 2)
    plan.compact([3, 1, 0, 2], 2)
    with pytest.raises(nat.SgpeError, match='beyond the launched'):
        plan.set_active(2, True)
    plan.graph_steps(4, 2)
    e = plan.energy().cpu().numpy()
    plan.sync()
    status, _ = plan.status()
    ref_status, _ = ref.status()
    for new, old in enumerate([3, 1, 0, 2]):
        np.testing.assert_array_equal(st[old].cpu().numpy(), ref_st[old].cpu().numpy())
        assert tuple(status[new]) == tuple(ref_status[old])
        if new < 2:
            np.testing.assert_array_equal(e[new], e_ref[old])
