"""Device tensor_tools primitives (HIP) vs their NumPy host forms (GPU).

fft_2d / ifft_2d on random inputs of every supported length (including
non-square and the 64 and 4096 extremes), the reference's own FFT checks on
ones (spinor_gpe/tests/fft_func_tests.py: round trip to 10*eps, atom number
to eps*N, Parseval with vol = 4 pi^2/N), norm and calc_pops.
"""

import itertools
import math

import numpy as np
import pytest
import torch

from spinor_gpe.pspinor import tensor_tools as tt

pytestmark = pytest.mark.gpu

SHAPES = [(64, 64), (128, 256), (256, 128), (512, 512), (1024, 1024), (64, 2048), (2048, 64),
          (4096, 64), (64, 4096), (2048, 2048)]
TOL = {torch.complex128: 4e-15, torch.complex64: 2e-6}


def rel(a, b):
    return np.linalg.norm(np.stack(a) - np.stack(b)) / np.linalg.norm(np.stack(b))


@pytest.mark.parametrize('dtype', [torch.complex128, torch.complex64])
@pytest.mark.parametrize('shape', SHAPES)
def test_fft2_matches_numpy(gpu, shape, dtype):
    print(f"[fft2] shape={shape} dtype={dtype}", flush=True)
    rng = np.random.default_rng(sum(shape))
    psi = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(2)]
    dr = np.array([0.3, 0.2])
    want = tt.fft_2d(psi, dr)
    dev = [torch.as_tensor(p, dtype=dtype, device=gpu) for p in psi]
    got = [g.cpu().numpy() for g in tt.fft_2d(dev, torch.tensor(dr, device=gpu))]
    logn = math.log2(shape[0] * shape[1])
    assert rel(got, want) < TOL[dtype] * logn
    back_want = tt.ifft_2d(want, dr)
    kdev = [torch.as_tensor(p, dtype=dtype, device=gpu) for p in want]
    back = [g.cpu().numpy() for g in tt.ifft_2d(kdev, dr)]
    assert rel(back, back_want) < TOL[dtype] * logn


def test_fft2_odd_count_in_place_and_overlap_checks(gpu):
    """Three arrays (one launch pair plus a single), in place and out of
    place, against numpy; outputs that overlap an array of their launch pair
    are refused before any launch."""
    from spinor_gpe import _native
    print("[fft2-pairs] shape=(128, 256) n=3", flush=True)
    rng = np.random.default_rng(7)
    shape = (128, 256)
    psi = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]
    dr = (0.3, 0.2)
    want = tt.fft_2d(psi, dr)
    dev = [torch.as_tensor(p, device=gpu) for p in psi]
    got = _native.fft2(dev, False, *dr)
    assert rel([g.cpu().numpy() for g in got], want) < 4e-15 * 15
    inplace = [d.clone() for d in dev]
    _native.fft2(inplace, False, *dr, out=inplace)
    assert rel([g.cpu().numpy() for g in inplace], want) < 4e-15 * 15
    big = torch.zeros((2 * shape[0] + 1,) + shape[1:], dtype=torch.complex128, device=gpu)
    shifted = big[1:shape[0] + 1]                     # partly overlaps big[:shape[0]]
    with pytest.raises(_native.SgpeError, match='launch pair'):
        _native.fft2(dev[:2], False, *dr, out=[big[:shape[0]], shifted])
    with pytest.raises(_native.SgpeError, match='launch pair'):
        _native.fft2(dev[:2], False, *dr, out=[dev[1], big[:shape[0]]])
    with pytest.raises(_native.SgpeError, match='its own input'):
        _native.fft2([big[:shape[0]]], False, *dr, out=[shifted])
    # Arrays of different pairs run in order and may overlap.
    a = [d.clone() for d in dev]
    _native.fft2(a, False, *dr, out=[a[0], a[1], a[0]])
    assert rel([a[0].cpu().numpy()], [want[2]]) < 4e-15 * 15


@pytest.mark.parametrize('shape', list(itertools.product([128, 256, 512, 1024], repeat=2)))
def test_reference_fft_checks_on_ones(gpu, shape):
    """fft_func_tests.py:27-56, 421-472 on device tensors."""
    eps = 10 * np.finfo(np.complex128).eps
    grid = [torch.ones(shape, dtype=torch.complex128, device=gpu)] * 2
    gk = tt.fft_2d(grid, (1, 1))
    back = tt.ifft_2d(gk, (1, 1))
    assert max(torch.max(abs(b - g)).item() for b, g in zip(back, grid)) <= eps
    atoms = [tt.calc_atoms(grid), tt.calc_atoms(back)]
    assert abs(atoms[0] - atoms[1]) <= eps * math.prod(shape)
    vol = 4 * np.pi ** 2 / math.prod(shape)
    assert abs(tt.calc_atoms(grid) - tt.calc_atoms(gk, vol)) <= eps * math.prod(shape)


@pytest.mark.parametrize('dtype', [torch.complex128, torch.complex64])
def test_norm_and_pops(gpu, dtype):
    rng = np.random.default_rng(5)
    psi = [rng.standard_normal((256, 128)) + 1j * rng.standard_normal((256, 128)) for _ in range(2)]
    want_psi, want_dens = tt.norm(psi, 0.05, 1234.0)
    dev = [torch.as_tensor(p, dtype=dtype, device=gpu) for p in psi]
    got_psi, got_dens = tt.norm(dev, 0.05, 1234.0)
    tol = 1e-14 if dtype == torch.complex128 else 1e-6
    assert rel([g.cpu().numpy() for g in got_psi], want_psi) < tol
    assert rel([g.cpu().numpy() for g in got_dens], want_dens) < tol
    pops = tt.calc_pops(got_psi, 0.05)
    assert sum(pops) == pytest.approx(1234.0, rel=tol * 10)
    assert pops == pytest.approx(tt.calc_pops(want_psi, 0.05), rel=tol * 10)
    # the inputs are not modified
    assert rel([d.cpu().numpy() for d in dev], psi) < tol


def test_cpu_tensors_rejected():
    with pytest.raises(TypeError, match='CPU tensors'):
        tt.fft_2d([torch.ones((64, 64), dtype=torch.complex128)] * 2)
