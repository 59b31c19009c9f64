"""Host-side checks that need no GPU: the C ABI surface of libsgpe, the
separable-operator detection, descriptors, and the native phase unwrapper
(cross-checked against the oracle's independent implementation)."""

import os
import re
import subprocess

import numpy as np
import pytest

from oracle import energy as oe
from spinor_gpe import _native as nat
from spinor_gpe.engine.plan import InstanceSpec, split_separable
from spinor_gpe.pspinor import pspinor as spin

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'sgpe.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(sgpe_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = nat.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(['nm', '-D', '--defined-only', nat.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r'\bT (sgpe_[a-z0-9_]+)', out))
    assert set(names) <= exported
    # every binding in _native is declared in the header, and vice versa
    assert set(nat.EXPORTED) == set(names)


def test_abi_version_and_error_path():
    lib = nat.lib()
    assert lib.sgpe_abi_version() == nat.ABI_VERSION
    rc = lib.sgpe_plan_tune(None, 8, 1, 8, 4)
    assert rc == -1
    assert b'NULL' in lib.sgpe_last_error()
    with pytest.raises(nat.SgpeError):
        nat.check(rc)


def test_cpu_device_is_refused():
    with pytest.raises(ValueError, match='HIP'):
        nat.require_device('cpu')


def test_split_separable():
    ps = spin.PSpinor('sep', mesh_points=(128, 64), r_sizes=(16, 8))
    ps.coupling_setup(kin_shift=True)
    ps.detuning_uniform(0.7)
    for arr in list(ps.kin_eng_spin) + list(ps.pot_eng_spin):
        parts = split_separable(arr)
        assert parts is not None
        ax, ay = parts
        assert ax.shape == (128,) and ay.shape == (64,)
        assert np.all(ay >= 0) and np.min(ax) == np.min(arr)
        err = np.max(np.abs(ax[None, :] + ay[:, None] - arr))
        assert err <= 8 * np.finfo(float).eps * np.max(np.abs(arr))
    x, y = ps.space['x_mesh'], ps.space['y_mesh']
    assert split_separable(ps.pot_eng + np.exp(-(x ** 2 + y ** 2))) is None
    assert split_separable(x * y) is None


def test_descriptor_modes():
    ps = spin.PSpinor('desc', mesh_points=(64, 64), r_sizes=(8, 8))
    ps.coupling_setup(kin_shift=False)
    ps.coupling_uniform(2.0)
    ps.rot_coupling = False
    d, keep = InstanceSpec.from_spin(ps, 'real', 1e-3).descriptor()
    assert d.coupling == nat.COUPLING_UNIFORM and d.omega == 2.0
    assert bool(d.theta_x) and bool(d.t_x[0]) and not bool(d.t_dense[0])
    th = np.ctypeslib.as_array(d.theta_x, shape=(64,))
    np.testing.assert_allclose(th, 2 * ps.kL_recoil * ps.space['x'])
    ps.coupling_grad(0.1, 1.0, axis=0)
    ps.pot_eng = ps.pot_eng + np.exp(-ps.space['x_mesh'] ** 2 - ps.space['y_mesh'] ** 2)
    ps.rot_coupling = True
    d, keep = InstanceSpec.from_spin(ps, 'imag', 0.02).descriptor()
    assert d.coupling == nat.COUPLING_ARRAY and bool(d.omega_dense)
    assert not bool(d.theta_x)
    assert bool(d.v_dense[0]) and not bool(d.v_x[0])
    assert d.time_mode == nat.TIME_IMAG
    del keep


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_native_unwrap_matches_independent_implementation(seed):
    rng = np.random.default_rng(seed)
    ny, nx = 40, 56
    yy, xx = np.mgrid[0:ny, 0:nx]
    true = 1.3 * xx - 0.8 * yy + 4 * np.sin(xx / 9.0) + 0.15 * rng.standard_normal((ny, nx))
    # a vortex-like winding and a noisy corner make the ordering matter
    true += np.arctan2(yy - 20.3, xx - 30.7)
    true[:8, :8] = rng.uniform(-np.pi, np.pi, (8, 8))
    wrapped = np.angle(np.exp(1j * true))
    got = nat.unwrap_phase(wrapped)
    want = oe.unwrap_phase_py(wrapped)
    np.testing.assert_array_equal(got, want)
    k = (got - wrapped) / (2 * np.pi)
    assert np.allclose(k, np.round(k), atol=1e-9)


def test_native_unwrap_ramp_and_identity():
    yy, xx = np.mgrid[0:64, 0:64]
    ramp = 0.7 * xx + 0.2 * yy
    got = nat.unwrap_phase(np.angle(np.exp(1j * ramp)))
    d = got - ramp
    assert np.allclose(d, d[0, 0], atol=1e-12)
    small = 0.4 * np.cos(xx / 6.0) * np.sin(yy / 4.0)
    np.testing.assert_array_equal(nat.unwrap_phase(small), small)


def test_native_unwrap_rejects_bad_arguments():
    """Argument errors come back as SGPE_E_ARG without touching the arrays:
    null pointers, empty grids, and grids whose ny * nx overflows int."""
    import ctypes  # pylint: disable=import-outside-toplevel
    buf = np.zeros(4)
    ptr = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    f = nat.lib().sgpe_unwrap_phase_2d
    assert f(None, ptr, 2, 2) == -1
    assert f(ptr, ptr, 0, 2) == -1
    assert f(ptr, ptr, 50000, 50000) == -1
    assert f(ptr, buf[2:].ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, 2) == 0


def test_precision_codes():
    for v in ('c128', 'complex128', nat.C128, __import__('torch').complex128):
        assert nat.precision_of(v) == nat.C128
    for v in ('c64', 'complex64', nat.C64, __import__('torch').complex64):
        assert nat.precision_of(v) == nat.C64
    with pytest.raises(ValueError):
        nat.precision_of('f16')


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    from spinor_gpe.engine import autotune
    monkeypatch.setenv('SGPE_TUNE_FILE', str(tmp_path / 'tune.json'))
    assert autotune.batch_bucket(1) == 1 and autotune.batch_bucket(5) == 16
    assert autotune.key(nat.C64, 256, 512, 3) == autotune.key('c64', 256, 512, 4) == '1:256x512:4'
    assert autotune.lookup('c128', 64, 64, 1) in (None, tuple(autotune.lookup('c128', 64, 64, 1) or ()))
    autotune.store('c128', 64, 64, 1, (4, 4, 8, 8), {'row': {'4x4': 1.0}, 'col': {}})
    assert autotune.lookup(nat.C128, 64, 64, 1) == (4, 4, 8, 8)
    # every listed shape divides its length into power-of-two thread counts
    for pass_ in (0, 1):
        for prec in ('c128', 'c64'):
            for n in (64, 128, 256, 512, 1024, 2048, 4096):
                shp = autotune.shapes(pass_, prec, n)
                assert shp, (pass_, prec, n)
                for e, p in shp:
                    threads = p * n // e
                    assert n % e == 0 and threads % 64 == 0 and threads <= 1024
