"""Continuous-batching server: results are the same alone and in the mix,
admission respects the HBM budget, and quench requests run their real-time
phase from the ground state."""

import numpy as np
import pytest
import torch

from spinor_gpe.engine import serve, workloads as wl

pytestmark = pytest.mark.gpu

SMALL = dict(block=50, max_steps=600, quench_steps=150, conv_tol=1e-7,
             slots={128: 8, 256: 4})


def _mix(n=7, seed=11):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        grid = 128 if i % 3 else 256
        kind = 'quench' if i % 4 == 1 else 'ground'
        reqs.append(serve.Request(i, wl.draw_params(rng, grid), kind, arrival=0.0))
    return reqs


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize('precision', ['c64', 'c128'])
def test_alone_and_in_mix_agree(gpu, precision):
    print(f"[serve] alone vs mix precision={precision}", flush=True)
    reqs = _mix()
    # staggered arrivals: some requests join while others are running
    for r in reqs[3:]:
        r.arrival = 0.05 * r.rid
    mixed = serve.Server(gpu, precision, keep_states=True, **SMALL).run(reqs)
    assert sorted(o.rid for o in mixed) == [r.rid for r in reqs]
    by_id = {o.rid: o for o in mixed}
    for rid in (1, 2, 6):  # a quench, a 128^2 ground state, a 256^2 ground state
        req = [r for r in reqs if r.rid == rid][0]
        alone = serve.Server(gpu, precision, keep_states=True, **SMALL).run(
            [serve.Request(rid, req.draw, req.kind, 0.0)])[0]
        got = by_id[rid]
        assert (got.steps_imag, got.steps_real) == (alone.steps_imag, alone.steps_real)
        assert _rel(got.psik, alone.psik) <= 1e-12, rid
        for key in ('E_tot', 'mu', 'N'):
            assert got.energy[key] == pytest.approx(alone.energy[key], rel=1e-12), (rid, key)


def test_quench_runs_real_time_after_ground_state(gpu):
    req = serve.Request(0, wl.draw_params(np.random.default_rng(5), 128), 'quench', 0.0)
    out = serve.Server(gpu, 'c128', **SMALL).run([req])[0]
    assert out.steps_real == SMALL['quench_steps']
    assert 0 < out.steps_imag <= SMALL['max_steps']
    # real time conserves the atom number exactly (renormalized each step)
    assert out.energy['N'] == pytest.approx(req.draw.atom_num, rel=1e-12)


def test_admission_waits_for_memory(gpu):
    reqs = _mix(n=4)
    for r in reqs:
        r.draw.grid = 128
        r.kind = 'ground'
    one = serve.state_bytes(128, 'c64')
    # room for one request's state and tables (no per-plan work buffer)
    srv = serve.Server(gpu, 'c64', hbm_budget=one, **SMALL)
    out = srv.run(reqs)
    assert len(out) == 4
    spans = sorted((o.admitted, o.finished) for o in out)
    for (a0, f0), (a1, _) in zip(spans, spans[1:]):
        assert a1 >= f0 - 1e-9  # at most one request held memory at a time
    assert srv.used == 0
    torch.cuda.synchronize()
