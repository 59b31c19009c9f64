"""Batched plans at a benchmark grid size against the CPU oracle (GPU).

Four 512^2 instances with different physics (the sweep's seeded Raman draws
in the rotated frame, imaginary time, and the lab-frame real-time Raman case)
share one plan, so the pass kernels take the slot-table path (state address
read per slot) with the shapes tuned for that batch size. Each slot after
three full steps must match the oracle run of its instance alone: complex128
to 1e-11 relative L2, complex64 to 2e-4.
"""

import numpy as np
import pytest

from cases import raman_lab
from oracle.reference_path import OraclePropagator
from spinor_gpe.engine import workloads as wl
from spinor_gpe.engine.plan import InstanceSpec, Plan

pytestmark = pytest.mark.gpu

STEPS = 3
N = 512
_REF = {}


def instances():
    if not _REF:
        out = []
        for i, d in enumerate(wl.sweep_draws(seed=4)[:3]):
            d.grid = N
            out.append((wl.raman_spinor(f'batch{i}', d), 'imag', 0.02))
        out.append(raman_lab(name='batch_lab', mesh=(N, N), r=(24, 24), seed=9))
        for ps, time, dt in out:
            o = OraclePropagator(ps, dt, time=time)
            o.run(STEPS)
            _REF.setdefault('runs', []).append((ps, time, dt, np.stack(o.psik_numpy())))
    return _REF['runs']


@pytest.mark.parametrize('prec', ['c128', 'c64'])
def test_batched_slots_match_oracle(gpu, prec):
    runs = instances()
    plan = Plan(gpu, prec, N, N, len(runs))
    print(f"[batched] {N}^2 x{len(runs)} {prec} tuning {plan.tuning()} streaming {plan.streaming()}", flush=True)
    states = []
    for slot, (ps, time, dt, _) in enumerate(runs):
        st = plan.new_state(ps.psik)
        plan.set_slot(slot, InstanceSpec.from_spin(ps, time, dt), st)
        states.append(st)
    plan.full_steps(STEPS)
    plan.sync()
    tol = 1e-11 if prec == 'c128' else 2e-4
    for slot, (st, (_, time, _, want)) in enumerate(zip(states, runs)):
        got = st.cpu().numpy()
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"[batched] slot {slot} ({time}) rel-L2 {err:.2e}", flush=True)
        assert err < tol, (slot, err)
