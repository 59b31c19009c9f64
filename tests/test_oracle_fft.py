"""Pins of the oracle's FFT conventions (tensor_tools.py:201-258).

The first group ports the assertions of the reference's own FFT tests
(spinor_gpe/tests/fft_func_tests.py:27-472: ones input, every shape in
{128..1024}^2, tolerance 10*eps, atom number to eps*N). Then brute-force
DFTs on tiny grids, the Gaussian known answer and Parseval.
"""

import itertools
import math

import numpy as np
import pytest
import torch

from oracle import reference_path as rp
from spinor_gpe.pspinor import tensor_tools as tt

LENGTHS = [2 ** s for s in range(7, 11)]
SHAPES = list(itertools.product(LENGTHS, LENGTHS))
EPS = 10 * np.finfo(np.complex128).eps


def _atoms_np(psi, vol=1.0):
    return sum(float(np.sum(np.abs(p) ** 2) * vol) for p in psi)


@pytest.mark.parametrize('shape', SHAPES)
def test_for_back_2d(shape):
    """for_back_torch_2d / for_back_np_2d (fft_func_tests.py:27-88)."""
    grid = [torch.ones(shape, dtype=torch.complex128)] * 2
    back = rp.ifft_2d(rp.fft_2d(grid, (1, 1)), (1, 1))
    assert max(torch.max(abs(b - g)).item() for b, g in zip(back, grid)) <= EPS
    atoms = [sum(rp.calc_pops(grid, 1.0)), sum(rp.calc_pops(back, 1.0))]
    assert abs(atoms[0] - atoms[1]) <= EPS * math.prod(shape)
    gnp = [np.ones(shape, dtype=np.complex128)] * 2
    bnp = tt.ifft_2d(tt.fft_2d(gnp, (1, 1)), (1, 1))
    assert max(np.max(abs(b - g)) for b, g in zip(bnp, gnp)) <= EPS


@pytest.mark.parametrize('shape', SHAPES[::3])
def test_1d_2d_consistency(shape):
    """compare1..4 (fft_func_tests.py:155-418): two 1D transforms == one 2D."""
    gnp = [np.ones(shape, dtype=np.complex128)] * 2
    half = tt.fft_1d(gnp, (1, 1), axis=0)
    full = tt.fft_1d(half, (1, 1), axis=1)
    back = tt.ifft_2d(full, (1, 1))
    assert max(np.max(abs(b - g)) for b, g in zip(back, gnp)) <= EPS
    k2 = tt.fft_2d(gnp, (1, 1))
    r1 = tt.ifft_1d(tt.ifft_1d(k2, (1, 1), axis=0), (1, 1), axis=1)
    assert max(np.max(abs(b - g)) for b, g in zip(r1, gnp)) <= EPS
    # compare3: 2D forward == 1D forward twice
    assert max(np.max(abs(a - b)) for a, b in zip(k2, full)) <= EPS * math.prod(shape)


@pytest.mark.parametrize('shape', SHAPES)
def test_parseval_ones(shape):
    """compare_norm_* (fft_func_tests.py:421-472): vol_k = 4 pi^2 / N."""
    grid = [torch.ones(shape, dtype=torch.complex128)] * 2
    gk = rp.fft_2d(grid, (1, 1))
    vol = 4 * np.pi ** 2 / math.prod(shape)
    assert abs(sum(rp.calc_pops(grid, 1.0)) - sum(rp.calc_pops(gk, vol))) <= EPS * math.prod(shape)


def _dft_brute(psi, dx, dy):
    """Explicit sum: psik[jy, jx] = c sum psi[y, x] e^{-2 pi i (my y/Ny + mx x/Nx)},
    m = j - N/2 (the fft-shifted frequency index)."""
    ny, nx = psi.shape
    my = np.arange(ny) - ny // 2
    mx = np.arange(nx) - nx // 2
    ey = np.exp(-2j * np.pi * np.outer(my, np.arange(ny)) / ny)
    ex = np.exp(-2j * np.pi * np.outer(np.arange(nx), mx) / nx)
    return ey @ psi @ ex * (dx * dy / (2 * np.pi))


@pytest.mark.parametrize('shape', [(4, 4), (8, 16), (16, 8)])
def test_brute_force_dft(shape):
    rng = np.random.default_rng(7)
    psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    want = _dft_brute(psi, 0.3, 0.7)
    got = rp.fft_2d([torch.as_tensor(psi)], torch.tensor([0.3, 0.7], dtype=torch.float64))[0].numpy()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())
    back = rp.ifft_2d([torch.as_tensor(want)], torch.tensor([0.3, 0.7], dtype=torch.float64))[0].numpy()
    np.testing.assert_allclose(back, psi, rtol=0, atol=1e-13)


def test_gaussian_known_answer():
    """e^{-r^2/2} on a grid starting at -R transforms to e^{-k^2/2} (-1)^{m_x + m_y}."""
    n, half = 128, 12.0
    x = np.linspace(-half, half, n, endpoint=False)
    xm, ym = np.meshgrid(x, x)
    psi = np.exp(-(xm ** 2 + ym ** 2) / 2)
    dr = 2 * half / n
    k = np.linspace(-np.pi / dr, np.pi / dr, n, endpoint=False)
    kx, ky = np.meshgrid(k, k)
    m = np.arange(n) - n // 2
    sign = np.outer((-1.0) ** m, (-1.0) ** m)
    want = np.exp(-(kx ** 2 + ky ** 2) / 2) * sign
    got = rp.fft_2d([torch.as_tensor(psi + 0j)], (dr, dr))[0].numpy()
    bound = np.exp(-min(half, np.pi / dr) ** 2 / 2) * 10 + 1e-14
    assert np.max(np.abs(got - want)) < bound


def test_parseval_random():
    rng = np.random.default_rng(3)
    psi = [torch.as_tensor(rng.standard_normal((32, 64)) + 1j * rng.standard_normal((32, 64)))
           for _ in range(2)]
    dr = (0.25, 0.5)
    dv_r = dr[0] * dr[1]
    dv_k = (np.pi / (64 * 0.25 / 2)) * (np.pi / (32 * 0.5 / 2))
    pk = rp.fft_2d(psi, dr)
    assert sum(rp.calc_pops(pk, dv_k)) == pytest.approx(sum(rp.calc_pops(psi, dv_r)), rel=1e-14)


def test_planted_shift_mistake_is_caught():
    """A variant without fftshift fails the brute-force check (the check has teeth)."""
    rng = np.random.default_rng(1)
    psi = rng.standard_normal((8, 8)) + 0j
    want = _dft_brute(psi, 1.0, 1.0)
    wrong = torch.fft.fftn(torch.as_tensor(psi)).numpy() / (2 * np.pi)
    assert np.max(np.abs(wrong - want)) > 1e-3
