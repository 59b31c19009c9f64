"""Host-side pieces of the benchmark: the roofline arithmetic and the
consistency of the shipped MI355X tuning table with the measured-traffic
summary that ``bench.py`` reads (every pass kernel the shipped table selects
for the benchmark workloads has a rocprofv3 traffic entry)."""

import json
import os

import pytest

import bench
from spinor_gpe.engine import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(autouse=True)
def _packaged_table_only(monkeypatch, tmp_path):
    """Look shapes up in the shipped table, not in a user cache."""
    monkeypatch.setenv('SGPE_TUNE_FILE', str(tmp_path / 'none.json'))


def _instance(kind, prec, n, shape):
    real = 'double' if prec == 'c128' else 'float'
    e, p, stream = (shape[0], shape[1], shape[4]) if kind == 'row' else (shape[2], shape[3], shape[5])
    name = (f"row_split_kernel<{real}, {n}, {e}, {-p}>" if kind == 'row' and p < 0  # split row layout
            else f"{kind}_pass_kernel<{real}, {n}, {e}, {p}>")
    return name + (' streaming' if stream else '')


@pytest.mark.parametrize('prec', ['c128', 'c64'])
@pytest.mark.parametrize('n', [512, 2048])
def test_shipped_shapes_have_traffic_entries(prec, n):
    shape = autotune.lookup(prec, n, n, 1)
    assert shape is not None and len(shape) >= 6, shape
    for kind in ('row', 'col'):
        name = _instance(kind, prec, n, shape)
        traffic = bench.measured_traffic(name)
        assert traffic is not None, f"no measured traffic for {name}"
        alg = 2 * 2 * n * n * (16 if prec == 'c128' else 8)
        assert 1.0 <= traffic / alg < 1.1, (name, traffic / alg)


def test_shipped_table_names_compiled_shapes():
    """Every entry of the shipped MI355X table selects a row shape and a
    column shape the library compiles (negative counts: split rows, column
    tiles with global y tables), for its own grid lengths."""
    with open(autotune.PACKAGED) as f:
        table = json.load(f)
    assert table
    for k, entry in table.items():
        prec_code, grid, slots = k.split(':')
        ny, nx = (int(v) for v in grid.split('x'))
        prec = ('c128', 'c64')[int(prec_code)]
        shape = entry['shape']
        assert len(shape) == 6 and int(slots) >= 1, k
        assert tuple(shape[0:2]) in autotune.shapes(0, prec, nx), (k, shape)
        assert tuple(shape[2:4]) in autotune.shapes(1, prec, ny), (k, shape)
        assert shape[4] in (0, 1) and shape[5] in (0, 1), (k, shape)


def test_more_than_one_gpu_is_refused():
    """``bench.py --gpus N`` with N > 1 exits (status 2) with a message saying
    that the engine measures one GPU, before any GPU work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert 'measures one GPU' in r.stderr + r.stdout


def test_roofline_arithmetic():
    res = {'row_ms': 0.0125, 'col_ms': 0.010, 'tuning': (4, 2, 8, 4), 'streaming': (1, 1)}
    r = bench.roofline(res, 'c128', 512)
    assert r['kernel'] == 'row_pass'
    assert r['kernel_instance'] == 'row_pass_kernel<double, 512, 4, 2> streaming'
    assert r['algorithmic_bytes'] == 2 * 2 * 512 * 512 * 16
    assert r['achieved'] == pytest.approx(16777216 / 12.5e-6 / 1e9, abs=0.05)  # rounded to 0.1 GB/s
    assert r['frac'] == pytest.approx(r['achieved'] / 8000.0, abs=1e-4)
    res = {'row_ms': 0.005, 'col_ms': 0.007, 'tuning': (8, 1, 16, 4), 'streaming': (1, 0)}
    r = bench.roofline(res, 'c64', 2048)
    assert r['kernel'] == 'col_pass'
    assert r['kernel_instance'] == 'col_pass_kernel<float, 2048, 16, 4>'
    assert r['traffic'] == bench.measured_traffic('col_pass_kernel<float, 2048, 16, 4>')


def test_traffic_summary_is_consistent():
    with open(os.path.join(ROOT, bench.TRAFFIC_FILE), encoding='utf8') as f:
        data = json.load(f)
    for name, k in data['kernels'].items():
        prec = 'c64' if '<float' in name else 'c128'
        cal = [v for key, v in data['calibration'].items() if key.startswith(prec)]
        assert cal, name
        # traffic = fetch x fetch factor + write x write factor, in bytes
        ok = any(abs(k['traffic_bytes'] - 1024 * (c['fetch_factor'] * k['fetch_kib'] + c['write_factor']
                                                  * k['write_kib'])) <= 0.001 * k['traffic_bytes']
                 for c in cal)
        assert ok, name
        assert k['traffic_over_algorithmic'] == pytest.approx(k['traffic_bytes'] / k['algorithmic_bytes'],
                                                              abs=1e-3)


def test_c64_vs_c128_errors():
    import numpy as np  # pylint: disable=import-outside-toplevel
    rng = np.random.default_rng(0)
    b = rng.normal(size=(2, 8, 8)) + 1j * rng.normal(size=(2, 8, 8))
    r128 = {'state': b, 'energy': np.array([2.0]), 'steps_total': 10}
    same = bench.c64_vs_c128({'state': b * np.exp(0.3j), 'energy': np.array([2.0]), 'steps_total': 10},
                             r128, 0.5)
    assert same['state_fidelity'] == pytest.approx(1.0, abs=1e-14)  # a global phase is no error
    # a complex64 state is compared in float64 sums: its own fidelity is 1 to rounding of float64
    big = rng.normal(size=(2, 256, 256)) + 1j * rng.normal(size=(2, 256, 256))
    f32 = big.astype(np.complex64)
    own = bench.c64_vs_c128({'state': f32, 'energy': np.array([2.0]), 'steps_total': 1},
                            {'state': f32.astype(np.complex128), 'energy': np.array([2.0]), 'steps_total': 1}, 1.0)
    assert abs(own['state_fidelity'] - 1.0) < 1e-12
    assert same['energy_rel_err'] == 0.0 and same['pops_err_of_atoms'] == pytest.approx(0.0, abs=1e-15)
    swapped = np.stack([b[0], np.zeros_like(b[1])])
    other = {'state': np.stack([np.zeros_like(b[0]), b[1]]), 'energy': np.array([2.2]), 'steps_total': 10}
    res = bench.c64_vs_c128(other, {'state': swapped, 'energy': np.array([2.0]), 'steps_total': 10}, 1.0)
    assert res['state_fidelity'] == pytest.approx(0.0, abs=1e-15)  # orthogonal states
    assert res['energy_rel_err'] == pytest.approx(0.1)
    n0 = float(np.sum(np.abs(b[0]) ** 2))
    n1 = float(np.sum(np.abs(b[1]) ** 2))
    assert res['pops_err_of_atoms'] == pytest.approx(max(n0, n1) / n0)


def test_non_square_grids_combine_square_tunings():
    sq_rows = autotune.lookup('c128', 1024, 1024, 1)
    sq_cols = autotune.lookup('c128', 512, 512, 1)
    got = autotune.lookup('c128', 512, 1024, 1)  # ny = 512, nx = 1024
    assert got[:2] == sq_rows[:2] and got[2:4] == sq_cols[2:4]
    assert got[4] == sq_rows[4] and got[5] == sq_cols[5]
