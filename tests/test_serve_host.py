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
