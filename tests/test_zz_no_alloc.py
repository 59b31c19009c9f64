"""sgpe_norm_noalloc (the allocation-free form of sgpe_norm) against
sgpe_norm through the C ABI (GPU): same kernels, so the normalized states,
the densities and the norm factor must be bit-identical; two calls in a row,
the second on a smaller grid, in complex128 and complex64."""

import ctypes

import numpy as np
import pytest
import torch

from spinor_gpe import _native as nat

pytestmark = pytest.mark.gpu


def _call(name, ctx, prec, psi, vol, atoms):
    p0, p1 = (p.clone() for p in psi)
    ny, nx = p0.shape
    d0 = torch.empty((ny, nx), dtype=nat.REAL_DTYPE[prec], device=p0.device)
    d1 = torch.empty_like(d0)
    nf = ctypes.c_double()
    nat.check(getattr(nat.lib(), name)(ctx.handle, prec, ny, nx, ctypes.c_void_p(p0.data_ptr()),
                                       ctypes.c_void_p(p1.data_ptr()), float(vol), float(atoms),
                                       ctypes.c_void_p(d0.data_ptr()), ctypes.c_void_p(d1.data_ptr()),
                                       ctypes.byref(nf)))
    torch.cuda.synchronize()
    return [p0, p1, d0, d1], nf.value


@pytest.mark.parametrize('precision', ['c128', 'c64'])
def test_norm_noalloc_matches_norm(gpu, precision):
    prec = nat.precision_of(precision)
    ctx = nat.Context.get(gpu)
    rng = np.random.default_rng(17)
    for ny, nx in ((768, 1024), (96, 200)):  # the second call is smaller
        print(f"[no-alloc] {precision} grid {ny}x{nx}", flush=True)
        psi = [torch.as_tensor(rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx)),
                               dtype=nat.TORCH_DTYPE[prec], device=gpu) for _ in range(2)]
        want, nf_want = _call('sgpe_norm', ctx, prec, psi, 0.05, 1234.5)
        got, nf_got = _call('sgpe_norm_noalloc', ctx, prec, psi, 0.05, 1234.5)
        assert nf_got == nf_want
        for g, w in zip(got, want):
            assert torch.equal(g, w)
        atoms = sum(float((d.double()).sum()) for d in got[2:]) * 0.05
        assert atoms == pytest.approx(1234.5, rel=1e-5 if precision == 'c64' else 1e-12)
