"""Host-side parts of the solve server (no GPU): the seeded request stream
and device refusal."""

import numpy as np
import pytest

from spinor_gpe.engine import serve


def test_request_stream_is_seeded_and_poisson():
    a = serve.request_stream(4000, rate=50.0, seed=3)
    b = serve.request_stream(4000, rate=50.0, seed=3)
    assert [(r.arrival, r.draw.grid, r.kind) for r in a] == [(r.arrival, r.draw.grid, r.kind) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert np.all(gaps > 0)
    assert np.mean(gaps) == pytest.approx(1 / 50.0, rel=0.05)       # exponential gaps, mean 1/rate
    assert np.std(gaps) == pytest.approx(1 / 50.0, rel=0.08)
    grids = np.array([r.draw.grid for r in a])
    for g, p in ((128, 0.3), (256, 0.4), (512, 0.2), (1024, 0.1)):
        assert np.mean(grids == g) == pytest.approx(p, abs=0.03)
    assert np.mean([r.kind == 'quench' for r in a]) == pytest.approx(0.3, abs=0.03)
    d = np.array([(r.draw.omega_er, r.draw.delta_er, np.log10(r.draw.atom_num), r.draw.g_ratio) for r in a])
    assert d[:, 0].min() >= 0 and d[:, 0].max() <= 4
    assert d[:, 1].min() >= -1 and d[:, 1].max() <= 1
    assert d[:, 2].min() >= 3 and d[:, 2].max() <= 5
    assert d[:, 3].min() >= 0.9 and d[:, 3].max() <= 1.1


def test_burst_stream_arrives_at_zero():
    assert all(r.arrival == 0.0 for r in serve.request_stream(10, rate=None))


def test_state_bytes():
    assert serve.state_bytes(512, 'c128') == 2 * 512 * 512 * 16 + 64 * 512 * 16
    assert serve.state_bytes(512, 'c64') == 2 * 512 * 512 * 8 + 64 * 512 * 8


def test_server_refuses_cpu():
    with pytest.raises(ValueError, match='HIP'):
        serve.Server('cpu')


class _FakePlan:
    """Stands in for a device plan: records what admission asks of it."""

    def __init__(self):
        self.loaded, self.inactive = [], []

    def new_state(self, psik):
        return np.stack([np.asarray(p) for p in psik])

    def set_slot(self, slot, spec, state):
        self.loaded.append(slot)

    def set_run(self, slot, pops=None, conv_tol=0.0):
        pass

    def set_active(self, slot, active):
        if not active:
            self.inactive.append(slot)


def _server(monkeypatch, free_bytes, slots):
    """A Server whose device query reports ``free_bytes`` free, with fake
    plans (admission and the memory budget are host logic)."""
    import torch
    monkeypatch.setattr(serve.nat, 'require_device', lambda d: torch.device('cpu'))
    monkeypatch.setattr(serve.torch.cuda, 'mem_get_info', lambda dev=None: (free_bytes, 2 * free_bytes))
    srv = serve.Server('cpu', 'c64', slots=slots)
    for g in slots:
        srv.plans[g] = _FakePlan()
        srv.free[g] = list(range(slots[g]))
    return srv


def _requests(n, grid=128):
    rng = np.random.default_rng(0)
    return [serve.Request(i, serve.wl.draw_params(rng, grid), 'ground', 0.0) for i in range(n)]


def test_admission_budget_from_free_memory(monkeypatch):
    """The default budget is 80 % of the free device memory; a request is
    admitted while a slot is free and its bytes fit, refused otherwise, and
    its bytes and slot return when it finishes."""
    need = serve.state_bytes(128, 'c64')
    srv = _server(monkeypatch, free_bytes=int(2.9 * need / 0.8), slots={128: 3})
    assert srv.budget == int(0.8 * int(2.9 * need / 0.8))
    r = _requests(4)
    j0, j1 = srv._admit(r[0], 0.0), srv._admit(r[1], 0.0)  # pylint: disable=protected-access
    assert j0 is not None and j1 is not None and srv.used == 2 * need
    assert srv._admit(r[2], 0.0) is None                    # a slot is free, the budget is not
    assert srv.free[128] == [2]
    srv._finish(j0, 100, True, np.zeros(8), 1.0)            # pylint: disable=protected-access
    assert srv.used == need and sorted(srv.free[128]) == [0, 2]
    assert srv.plans[128].inactive == [j0.slot]
    j2 = srv._admit(r[2], 1.0)                              # pylint: disable=protected-access
    assert j2 is not None and srv.used == 2 * need


def test_admission_waits_for_a_slot(monkeypatch):
    need = serve.state_bytes(128, 'c64')
    srv = _server(monkeypatch, free_bytes=100 * need, slots={128: 1})
    r = _requests(2)
    assert srv._admit(r[0], 0.0) is not None                # pylint: disable=protected-access
    assert srv._admit(r[1], 0.0) is None and srv.used == need  # pylint: disable=protected-access


def test_request_larger_than_the_budget_is_refused(monkeypatch):
    """A request that can never fit raises instead of waiting forever (no
    device work is issued)."""
    need = serve.state_bytes(128, 'c64')
    srv = _server(monkeypatch, free_bytes=int(0.5 * need), slots={128: 4})
    with pytest.raises(MemoryError, match='cannot be admitted'):
        srv.run(_requests(1))
    assert srv.used == 0 and srv.plans[128].loaded == []
