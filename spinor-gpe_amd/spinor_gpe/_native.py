"""ctypes binding of libsgpe (the HIP engine, C ABI declared in include/sgpe.h).

The library is loaded after :mod:`torch`, so it resolves ``libamdhip64.so.7``
to the HIP runtime torch already loaded: device pointers and streams are then
shared with torch. There is no CPU implementation behind this module: if the
library is missing or no HIP device is present, every entry point raises.
"""

import ctypes
import os
import threading

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_lib', 'libsgpe.so')

ABI_VERSION = 1
C128, C64 = 0, 1
TIME_IMAG, TIME_REAL = 0, 1
COUPLING_OFF, COUPLING_UNIFORM, COUPLING_ARRAY = 0, 1, 2
SUB_OUTER, SUB_INNER = 0, 1
PRECISIONS = {'c128': C128, 'complex128': C128, torch.complex128: C128, C128: C128,
              'c64': C64, 'complex64': C64, torch.complex64: C64, C64: C64}
TORCH_DTYPE = {C128: torch.complex128, C64: torch.complex64}
REAL_DTYPE = {C128: torch.float64, C64: torch.float32}

_dp = ctypes.POINTER(ctypes.c_double)


class InstDesc(ctypes.Structure):
    """Mirror of ``sgpe_inst_desc``."""

    _fields_ = [
        ('ny', ctypes.c_int), ('nx', ctypes.c_int),
        ('time_mode', ctypes.c_int),
        ('t_step', ctypes.c_double),
        ('atom_num', ctypes.c_double),
        ('g_uu', ctypes.c_double), ('g_dd', ctypes.c_double), ('g_ud', ctypes.c_double),
        ('dx', ctypes.c_double), ('dy', ctypes.c_double),
        ('t_x', _dp * 2), ('t_y', _dp * 2), ('t_dense', _dp * 2),
        ('v_x', _dp * 2), ('v_y', _dp * 2), ('v_dense', _dp * 2),
        ('coupling', ctypes.c_int),
        ('omega', ctypes.c_double),
        ('omega_dense', _dp),
        ('theta_x', _dp),
    ]


class SgpeError(RuntimeError):
    """A libsgpe call returned an error code."""


_lib = None
_lib_lock = threading.Lock()


def _declare(lib):
    vp, ip, i, d = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_double
    sig = {
        'sgpe_abi_version': (i, []),
        'sgpe_last_error': (ctypes.c_char_p, []),
        'sgpe_debug_fail_allocs': (i, [i]),
        'sgpe_ctx_create': (i, [i, vp, ctypes.POINTER(vp)]),
        'sgpe_ctx_set_stream': (i, [vp, vp]),
        'sgpe_ctx_synchronize': (i, [vp]),
        'sgpe_ctx_destroy': (None, [vp]),
        'sgpe_fft2': (i, [vp, i, i, i, i, ctypes.POINTER(vp), ctypes.POINTER(vp), i, d, d]),
        'sgpe_norm': (i, [vp, i, i, i, vp, vp, d, d, vp, vp, _dp]),
        'sgpe_norm_noalloc': (i, [vp, i, i, i, vp, vp, d, d, vp, vp, _dp]),
        'sgpe_pops': (i, [vp, i, i, i, vp, vp, d, _dp]),
        'sgpe_plan_create': (i, [vp, i, i, i, i, ctypes.POINTER(vp)]),
        'sgpe_plan_destroy': (None, [vp]),
        'sgpe_plan_set_slot': (i, [vp, i, ctypes.POINTER(InstDesc), vp]),
        'sgpe_plan_set_state': (i, [vp, i, vp]),
        'sgpe_plan_set_run': (i, [vp, i, vp, i, d]),
        'sgpe_plan_set_active': (i, [vp, i, i]),
        'sgpe_plan_compact': (i, [vp, ip, i]),
        'sgpe_plan_status': (i, [vp, ip, _dp]),
        'sgpe_plan_tune': (i, [vp, i, i, i, i]),
        'sgpe_list_shapes': (i, [i, i, i, ip, i]),
        'sgpe_plan_get_tuning': (i, [vp, ip]),
        'sgpe_plan_set_streaming': (i, [vp, i, i]),
        'sgpe_plan_get_streaming': (i, [vp, ip]),
        'sgpe_plan_set_persistent': (i, [vp, i, ctypes.c_longlong]),
        'sgpe_plan_get_persistent': (i, [vp, ip]),
        'sgpe_plan_check': (i, [vp]),
        'sgpe_list_persistent': (i, [i, i, ip, i]),
        'sgpe_profile_persistent': (i, [vp, i, _dp]),
        'sgpe_single_step': (i, [vp, i]),
        'sgpe_full_steps': (i, [vp, i]),
        'sgpe_graph_steps': (i, [vp, i, i]),
        'sgpe_plan_sync': (i, [vp]),
        'sgpe_profile_full_steps': (i, [vp, i, _dp]),
        'sgpe_energy': (i, [vp, vp]),
        'sgpe_group_create': (i, [vp, i, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        'sgpe_group_destroy': (None, [vp]),
        'sgpe_group_run': (i, [vp, i, i, i]),
        'sgpe_unwrap_phase_2d': (i, [_dp, _dp, i, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return sig.keys()


EXPORTED = None


def lib():
    """Load (once) and return the ctypes handle of libsgpe."""
    global _lib, EXPORTED
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"libsgpe is not built ({LIB_PATH} missing); run "
                    "__graft_entry__.build() or spinor_gpe/build.py first.")
            handle = ctypes.CDLL(LIB_PATH)
            EXPORTED = tuple(_declare(handle))
            if handle.sgpe_abi_version() != ABI_VERSION:
                raise RuntimeError("libsgpe ABI version mismatch")
            _lib = handle
    return _lib


def check(rc):
    """Raise :class:`SgpeError` (or MemoryError) for a non-zero return code."""
    if rc != 0:
        msg = lib().sgpe_last_error().decode(errors='replace')
        if rc == -3:
            raise MemoryError(f"libsgpe: {msg}")
        raise SgpeError(f"libsgpe error {rc}: {msg}")


def require_device(device):
    """Return the torch device for ``device``; only HIP devices are accepted."""
    dev = torch.device(device)
    if dev.type != 'cuda':
        raise ValueError(
            f"device {device!r}: this engine runs only on HIP GPUs (use "
            "device='cuda'); there is no CPU propagation path.")
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is available to torch")
    if dev.index is None:
        dev = torch.device('cuda', torch.cuda.current_device())
    return dev


class Context:
    """One ``sgpe_ctx`` per device, bound to torch's current stream per call."""

    _by_device = {}

    def __init__(self, index):
        self.index = index
        handle = ctypes.c_void_p()
        with torch.cuda.device(index):
            check(lib().sgpe_ctx_create(index, None, ctypes.byref(handle)))
        self.handle = handle
        self._stream = None  # stream last bound (the context's own until the first bind)

    @classmethod
    def get(cls, device):
        if isinstance(device, torch.device) and device.type == 'cuda' and device.index is not None:
            ctx = cls._by_device.get(device.index)  # the per-call path: a device tensor's device
            if ctx is not None:
                ctx.bind_stream()
                return ctx
        dev = require_device(device)
        ctx = cls._by_device.get(dev.index)
        if ctx is None:
            ctx = cls._by_device[dev.index] = Context(dev.index)
        ctx.bind_stream()
        return ctx

    def bind_stream(self):
        stream = torch.cuda.current_stream(self.index).cuda_stream
        if stream != self._stream:
            check(lib().sgpe_ctx_set_stream(self.handle, ctypes.c_void_p(stream)))
            self._stream = stream

    def synchronize(self):
        check(lib().sgpe_ctx_synchronize(self.handle))


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def precision_of(dtype):
    try:
        return PRECISIONS[dtype]
    except KeyError:
        raise ValueError(f"unsupported precision/dtype {dtype!r}; use "
                         "complex128 ('c128') or complex64 ('c64')") from None


def fft2(arrays, inverse, dx, dy, out=None):
    """Device fft_2d / ifft_2d of a list of (ny, nx) complex tensors."""
    ref = arrays[0]
    prec = precision_of(ref.dtype)
    ny, nx = ref.shape
    ins = [a.contiguous() for a in arrays]
    for a in ins:
        if a.shape != ref.shape or a.dtype != ref.dtype or a.device != ref.device:
            raise ValueError("all arrays must share shape, dtype and device")
    outs = out if out is not None else [torch.empty_like(a) for a in ins]
    ctx = Context.get(ref.device)
    n = len(ins)
    in_arr = (ctypes.c_void_p * n)(*[a.data_ptr() for a in ins])
    out_arr = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
    check(lib().sgpe_fft2(ctx.handle, prec, ny, nx, n, in_arr, out_arr, int(bool(inverse)),
                          float(dx), float(dy)))
    return outs


def norm(psi, vol_elem, atom_num):
    """Device norm(): returns (normalized psi list, density list)."""
    p0, p1 = (p.clone() for p in psi)
    prec = precision_of(p0.dtype)
    ny, nx = p0.shape
    d0 = torch.empty((ny, nx), dtype=REAL_DTYPE[prec], device=p0.device)
    d1 = torch.empty_like(d0)
    ctx = Context.get(p0.device)
    nf = ctypes.c_double()
    check(lib().sgpe_norm(ctx.handle, prec, ny, nx, _ptr(p0), _ptr(p1), float(vol_elem),
                          float(atom_num), _ptr(d0), _ptr(d1), ctypes.byref(nf)))
    return [p0, p1], [d0, d1]


def pops(psi, vol_elem):
    """Device calc_pops(): per-component vol * sum |psi_s|^2 as floats."""
    p0, p1 = (p.contiguous() for p in psi)
    prec = precision_of(p0.dtype)
    ny, nx = p0.shape
    ctx = Context.get(p0.device)
    out = (ctypes.c_double * 2)()
    check(lib().sgpe_pops(ctx.handle, prec, ny, nx, _ptr(p0), _ptr(p1), float(vol_elem), out))
    return [out[0], out[1]]


def unwrap_phase(wrapped):
    """Reliability-ordered 2D phase unwrapping of a float64 (ny, nx) array."""
    w = np.ascontiguousarray(wrapped, dtype=np.float64)
    if w.ndim != 2:
        raise ValueError("unwrap_phase expects a 2D array")
    out = np.empty_like(w)
    check(lib().sgpe_unwrap_phase_2d(w.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                                     w.shape[0], w.shape[1]))
    return out
