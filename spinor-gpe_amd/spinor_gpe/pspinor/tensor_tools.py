"""Array primitives with the reference's NumPy / tensor dual dispatch.

Same functions and semantics as ``spinor_gpe/pspinor/tensor_tools.py``:

* NumPy arrays are handled on the host with NumPy, exactly as the reference
  does for set-up and end-of-run analysis;
* torch tensors on a HIP device go to the libsgpe kernels (FFT, norm,
  populations) through the C ABI;
* torch tensors on the CPU raise: this package has no CPU tensor path.

Lists of two arrays are spinor wavefunctions ``[psi_up, psi_down]`` of shape
``(Ny, Nx)``; k-space arrays are fft-shifted.
"""

import operator
import weakref
from functools import reduce

import numpy as np
import torch

# Host arrays that device tensors were copied from, by id(tensor): a weak
# reference to the tensor (checked by identity), its version counter at the
# copy and the values. Small parameters (the grid spacings ``dr``) are then
# read on the host without a device-to-host copy, which would synchronize the
# stream, for as long as the tensor has not been modified in place. (Tensors
# cannot be weak dictionary keys: key comparison would be elementwise.)
_HOST_COPIES = {}


def remember_host_copy(tensor, array):
    """Record that ``tensor`` holds the values of the host array ``array``."""
    key = id(tensor)
    _HOST_COPIES[key] = (weakref.ref(tensor), tensor._version,  # pylint: disable=protected-access
                         np.array(array, copy=True))
    weakref.finalize(tensor, _HOST_COPIES.pop, key, None)
    return tensor


def _host_copy(tensor):
    entry = _HOST_COPIES.get(id(tensor))
    if entry is not None and entry[0]() is tensor and entry[1] == tensor._version:  # pylint: disable=protected-access
        return entry[2]
    return None


def _native():
    from spinor_gpe import _native as nat  # pylint: disable=import-outside-toplevel
    return nat


def _is_tensor(x):
    return isinstance(x, torch.Tensor)


def _device_tensor(x, what):
    if not x.is_cuda:
        raise TypeError(f"{what}: CPU tensors are not supported; use NumPy arrays on "
                        "the host or tensors on device 'cuda'.")
    return x


def to_numpy(input_tens):
    """Tensor or list of tensors -> host NumPy (tensor_tools.py:14-38)."""
    if isinstance(input_tens, list):
        return [inp.cpu().numpy() for inp in input_tens]
    if _is_tensor(input_tens):
        return input_tens.cpu().numpy()
    raise TypeError(f"cannot convert {type(input_tens)} to NumPy")


def to_tensor(input_arr, dev='cpu', dtype=64):
    """Array or list of arrays -> tensors (tensor_tools.py:41-77).

    ``dtype``: 32 -> float32, 64 -> float64, 128 -> complex128 (the
    reference's codes), plus 'c64' -> complex64.
    """
    all_dtypes = {32: torch.float32, 64: torch.float64, 128: torch.complex128,
                  'c64': torch.complex64}
    if isinstance(input_arr, list):
        return [torch.as_tensor(inp, dtype=all_dtypes[dtype], device=dev) for inp in input_arr]
    if isinstance(input_arr, np.ndarray):
        return torch.as_tensor(input_arr, dtype=all_dtypes[dtype], device=dev)
    raise TypeError(f"cannot convert {type(input_arr)} to a tensor")


def to_cpu(input_tens):
    """Move a tensor (or list of tensors) to host memory."""
    if isinstance(input_tens, list):
        return [inp.cpu() for inp in input_tens]
    return input_tens.cpu()


def to_gpu(input_tens, dev='cuda'):
    """Move a tensor (or list of tensors) to ``dev``.

    The reference returns the un-moved tensor for a single tensor
    (tensor_tools.py:124-125); this version returns the moved one.
    """
    if isinstance(input_tens, list):
        assert isinstance(input_tens[0], torch.Tensor), (
            f"Cannot move a non-Tensor object of dtype `{type(input_tens[0])}` to GPU memory.")
        return [inp.to(dev) for inp in input_tens]
    return input_tens.to(dev)


def prod(factors):
    """Product of the elements of a 1D sequence (tensor_tools.py:594-600)."""
    return reduce(operator.mul, factors, 1)


def fft_1d(psi, delta_r=(1, 1), axis=0) -> list:
    """Forward 1D FFT along x (axis=0) or y (axis=1), NumPy only.

    Scaled by ``delta_r[axis]/sqrt(2 pi)`` and fft-shifted along that axis
    (tensor_tools.py:130-164).
    """
    true_ax = [1, 0]
    norm_c = delta_r[axis] / np.sqrt(2 * np.pi)
    if not isinstance(psi[0], np.ndarray):
        raise TypeError("fft_1d is provided for NumPy arrays only")
    out = [np.fft.fftn(p, axes=[true_ax[axis]]) * norm_c for p in psi]
    return [np.fft.fftshift(pk, axes=true_ax[axis]) for pk in out]


def ifft_1d(psik, delta_r=(1, 1), axis=0) -> list:
    """Inverse of :func:`fft_1d` (tensor_tools.py:167-198), NumPy only."""
    true_ax = [1, 0]
    norm_c = delta_r[axis] / np.sqrt(2 * np.pi)
    if not isinstance(psik[0], np.ndarray):
        raise TypeError("ifft_1d is provided for NumPy arrays only")
    out = [np.fft.ifftshift(pk, axes=true_ax[axis]) for pk in psik]
    return [np.fft.ifftn(p, axes=[true_ax[axis]]) / norm_c for p in out]


def _dr_pair(delta_r):
    if _is_tensor(delta_r):
        known = _host_copy(delta_r)
        delta_r = known if known is not None else delta_r.detach().cpu().numpy()
    return float(delta_r[0]), float(delta_r[1])


def fft_2d(psi, delta_r=(1, 1)) -> list:
    """``fftshift(fftn(psi)) * dx*dy/(2 pi)`` per component (tensor_tools.py:201-228)."""
    if isinstance(psi[0], np.ndarray):
        c = prod(delta_r) / (2 * np.pi)
        return [np.fft.fftshift(np.fft.fftn(p) * c) for p in psi]
    if _is_tensor(psi[0]):
        for p in psi:
            _device_tensor(p, 'fft_2d')
        dx, dy = _dr_pair(delta_r)
        return _native().fft2(list(psi), False, dx, dy)
    raise TypeError(f"unsupported wavefunction type {type(psi[0])}")


def ifft_2d(psik, delta_r=(1, 1)) -> list:
    """``ifftn(ifftshift(psik)) / (dx*dy/(2 pi))`` per component (tensor_tools.py:231-258)."""
    if isinstance(psik[0], np.ndarray):
        c = prod(delta_r) / (2 * np.pi)
        return [np.fft.ifftn(np.fft.ifftshift(pk)) / c for pk in psik]
    if _is_tensor(psik[0]):
        for p in psik:
            _device_tensor(p, 'ifft_2d')
        dx, dy = _dr_pair(delta_r)
        return _native().fft2(list(psik), True, dx, dy)
    raise TypeError(f"unsupported wavefunction type {type(psik[0])}")


def norm(psi, vol_elem, atom_num, pop_frac=None):
    """Normalize the spinor to ``atom_num`` atoms (tensor_tools.py:261-311).

    Returns ``(psi_norm, dens_norm)``. Population-fraction normalization is
    not implemented (as in the reference).
    """
    if pop_frac is not None:
        raise NotImplementedError("Normalizing to the expected population fractions is "
                                  "not implemented.")
    if isinstance(psi[0], np.ndarray):
        dens = density(psi)
        nf = np.sum(dens[0] + dens[1]) * vol_elem / atom_num
        return [p / np.sqrt(nf) for p in psi], [d / nf for d in dens]
    if _is_tensor(psi[0]):
        for p in psi:
            _device_tensor(p, 'norm')
        if _is_tensor(vol_elem):
            vol_elem = float(vol_elem.item())
        return _native().norm(psi, vol_elem, atom_num)
    raise TypeError(f"unsupported wavefunction type {type(psi[0])}")


def grad(psi, delta_r):
    """Finite-difference gradient of a component or list (tensor_tools.py:314-328)."""
    if isinstance(psi, list):
        return [grad_comp(p, delta_r) for p in psi]
    return grad_comp(psi, delta_r)


def grad_comp(psi_comp, delta_r):
    """``np.gradient(psi, *delta_r)`` (tensor_tools.py:331-350).

    As in the reference, ``delta_r[0]`` is used as the spacing of axis 0 (y)
    and ``delta_r[1]`` for axis 1 (x).
    """
    if isinstance(psi_comp, np.ndarray):
        return np.gradient(psi_comp, *np.array(delta_r))
    if _is_tensor(psi_comp):
        raise NotImplementedError("Spatial gradients for tensors are not implemented.")
    raise TypeError(f"`psi_comp` is of type {type(psi_comp)}.")


def grad_sq_comp(psi_comp, delta_r):
    """Squared gradient magnitude of one component."""
    grd_0, grd_1 = grad_comp(psi_comp, delta_r)
    return grd_0 ** 2 + grd_1 ** 2


def grad_sq(psi, delta_r):
    """Squared gradient magnitude of a component or list."""
    if isinstance(psi, list):
        return [grad_sq_comp(p, delta_r) for p in psi]
    return grad_sq_comp(psi, delta_r)


def conj(psi):
    """Complex conjugate of a component or list."""
    if isinstance(psi, list):
        return [conj_comp(p) for p in psi]
    return conj_comp(psi)


def conj_comp(psi_comp):
    """Complex conjugate of one component."""
    if isinstance(psi_comp, np.ndarray):
        return np.conj(psi_comp)
    if _is_tensor(psi_comp):
        return torch.conj(psi_comp)
    raise TypeError(f"`psi_comp` is of type {type(psi_comp)}.")


def density(psi):
    """``|psi|^2`` of a component or list (tensor_tools.py:392-411)."""
    if isinstance(psi, list):
        return [norm_sq(p) for p in psi]
    return norm_sq(psi)


def norm_sq(psi_comp):
    """``|psi_comp|^2`` (tensor_tools.py:414-441)."""
    if isinstance(psi_comp, np.ndarray):
        return np.abs(psi_comp) ** 2
    if _is_tensor(psi_comp):
        _device_tensor(psi_comp, 'norm_sq')
        return torch.abs(psi_comp) ** 2
    raise TypeError(f"`psi_comp` is of type {type(psi_comp)}.")


def calc_atoms(psi, vol_elem=1.0):
    """Total atom number ``sum(calc_pops)`` (tensor_tools.py:444-463)."""
    return sum(calc_pops(psi, vol_elem=vol_elem))


def calc_pops(psi, vol_elem=1.0):
    """Atom number per component, ``[sum |psi_s|^2 * vol]`` (tensor_tools.py:466-484)."""
    if _is_tensor(psi[0]):
        for p in psi:
            _device_tensor(p, 'calc_pops')
        if _is_tensor(vol_elem):
            vol_elem = float(vol_elem.item())
        return _native().pops(psi, vol_elem)
    return [float(d.sum() * vol_elem) for d in density(psi)]


def phase(psi, uwrap=False, dens=None):
    """Phase of a component or list (tensor_tools.py:487-511)."""
    if dens is not None:
        assert len(psi) == len(dens), "`psi` and `dens` should have the same length."
    elif isinstance(psi, list):
        dens = [None] * len(psi)
    if isinstance(psi, list):
        return [phase_comp(p, uwrap, d) for p, d in zip(psi, dens)]
    return phase_comp(psi, uwrap, dens)


def phase_comp(psi_comp, uwrap=False, dens=None):
    """Angle of one component, optionally unwrapped, zeroed where
    ``dens < 1e-6 max(dens)`` (tensor_tools.py:514-539).

    Unwrapping uses libsgpe's implementation of the reliability-sorted
    algorithm that scikit-image's ``unwrap_phase`` implements.
    """
    if isinstance(psi_comp, np.ndarray):
        ang = np.angle(psi_comp)
        if uwrap:
            ang = _native().unwrap_phase(ang)
    elif _is_tensor(psi_comp):
        ang = torch.angle(psi_comp)
        if uwrap:
            raise NotImplementedError("Unwrapping the complex phase is not implemented "
                                      "for tensors.")
    else:
        raise TypeError(f"`psi_comp` is of type {type(psi_comp)}.")
    if dens is not None:
        ang[dens < (dens.max() * 1e-6)] = 0
    return ang


def inner_prod():
    """Placeholder kept for API parity (empty in the reference, :542-543)."""


def evolution_op(t_step, energy):
    """``exp(-1j * energy * t_step)`` (tensor_tools.py:546-560).

    Kept for API parity; the propagator builds its phase tables natively.
    """
    if isinstance(energy, list):
        return [evolution_op(t_step, e) for e in energy]
    if isinstance(energy, np.ndarray):
        return np.exp(-1.0j * energy * t_step)
    return torch.exp(-1.0j * energy * t_step)


def coupling_op(t_step, coupling=None, expon=0.0):
    """2x2 coupling evolution operator (tensor_tools.py:563-591).

    ``[[cos a, -i sin a e^{-i expon}], [-i sin a e^{+i expon}, cos a]]`` with
    ``a = coupling * t_step / 2``. Kept for API parity.
    """
    lib = np if not _is_tensor(coupling) else torch
    if coupling is None:
        coupling = 0.0
    arg = coupling * t_step / 2
    cosine = lib.cos(arg)
    sine = -1.0j * lib.sin(arg)
    if _is_tensor(arg) and not _is_tensor(expon):
        expon = torch.tensor(expon)
    return [[cosine, sine * lib.exp(-1.0j * expon)],
            [sine * lib.exp(1.0j * expon), cosine]]


def expect_val(psi):
    """Not implemented in the reference either (tensor_tools.py:603-607)."""
    raise NotImplementedError("Function for computing the expectation value of an "
                              "arbitrary spatial operator is not implemented.")
