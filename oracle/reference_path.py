"""Op-for-op CPU restatement of the reference's propagation path (oracle).

TEST INFRASTRUCTURE: the checker for the HIP engine and the timed CPU
baseline; never imported by the product.

Follows ``spinor_gpe/pspinor/tensor_propagator.py:67-271`` and the
``spinor_gpe/pspinor/tensor_tools.py`` functions it calls, on torch CPU
tensors (complex128 / float64), with the same torch calls in the same order:
wavefunctions are lists of two ``(Ny, Nx)`` tensors, every FFT is
``torch.fft.fftn/ifftn`` with the fftshift pair and the ``dx*dy/(2 pi)``
scale, every normalization syncs its scalar with ``.item()``.
"""

import numpy as np
import torch


# ----------------------------------------------------------- tensor_tools

def prod(factors):
    """tensor_tools.py:594-600."""
    out = 1
    for f in factors:
        out = out * f
    return out


def fft_2d(psi, delta_r):
    """tensor_tools.py:218-228 (torch branch): fftshift(fftn(p) * c)."""
    c = prod(delta_r) / (2 * np.pi)
    out = [torch.fft.fftn(p) * c for p in psi]
    return [torch.fft.fftshift(pk) for pk in out]


def ifft_2d(psik, delta_r):
    """tensor_tools.py:248-258 (torch branch): ifftn(ifftshift(pk)) / c."""
    c = prod(delta_r) / (2 * np.pi)
    shifted = [torch.fft.ifftshift(pk) for pk in psik]
    return [torch.fft.ifftn(p) / c for p in shifted]


def density(psi):
    """tensor_tools.py:392-441: torch.abs(p)**2 per component."""
    return [torch.abs(p) ** 2 for p in psi]


def norm(psi, vol_elem, atom_num):
    """tensor_tools.py:289, 301-305: normalize to atom_num via .item()."""
    dens = density(psi)
    nf = torch.sum(dens[0] + dens[1]) * vol_elem / atom_num
    psi_n = [p / np.sqrt(nf.item()) for p in psi]
    dens_n = [d / nf.item() for d in dens]
    return psi_n, dens_n


def calc_pops(psi, vol_elem):
    """tensor_tools.py:466-484."""
    return [float(d.sum() * vol_elem) for d in density(psi)]


def evolution_op(t_step, energy):
    """tensor_tools.py:546-560."""
    return [torch.exp(-1.0j * e * t_step) for e in energy]


def coupling_op(t_step, coupling, expon):
    """tensor_tools.py:563-591."""
    arg = coupling * t_step / 2
    cosine = torch.cos(arg)
    sine = -1.0j * torch.sin(arg)
    return [[cosine, sine * torch.exp(-1.0j * expon)],
            [sine * torch.exp(1.0j * expon), cosine]]


# ------------------------------------------------------- TensorPropagator

class OraclePropagator:
    """CPU restatement of TensorPropagator (tensor_propagator.py:67-271).

    ``spin`` is any object with the PSpinor attributes the reference reads:
    ``atom_num, is_coupling, g_sc, kin_eng_spin, pot_eng_spin, psik, space,
    coupling, kL_recoil, rot_coupling``.
    """

    def __init__(self, spin, t_step, time='imag'):
        # tensor_propagator.py:96-103
        if time == 'imag':
            self.t_step = -1.0j * t_step
        elif time == 'real':
            self.t_step = t_step
        else:
            raise ValueError(time)
        magic_gamma = 1 / (2 + 2 ** (1 / 3))
        self.dt_out = self.t_step * magic_gamma
        self.dt_in = self.t_step * (1 - 2 * magic_gamma)
        # tensor_propagator.py:111-129
        self.atom_num = spin.atom_num
        self.is_coupling = spin.is_coupling
        self.g_sc = dict(spin.g_sc)
        self.kin_eng_spin = [torch.as_tensor(np.asarray(k), dtype=torch.float64) for k in spin.kin_eng_spin]
        self.pot_eng_spin = [torch.as_tensor(np.asarray(v), dtype=torch.float64) for v in spin.pot_eng_spin]
        self.psik = [torch.as_tensor(np.asarray(p), dtype=torch.complex128) for p in spin.psik]
        keys = ['dr', 'dk', 'x_mesh', 'y_mesh', 'dv_r', 'dv_k']
        self.space = {k: torch.tensor(spin.space[k]) for k in keys}
        self.coupling = torch.as_tensor(np.asarray(spin.coupling), dtype=torch.float64)
        self.kL_recoil = spin.kL_recoil
        if spin.rot_coupling:
            self.expon = torch.tensor(0.0)
        else:
            self.expon = 2 * self.kL_recoil * self.space['x_mesh']
        # tensor_propagator.py:138-149
        self.eng_out = {'kin': evolution_op(self.dt_out / 2, self.kin_eng_spin),
                        'pot': evolution_op(self.dt_out, self.pot_eng_spin),
                        'coupl': coupling_op(self.dt_out, self.coupling / 2, self.expon)}
        self.eng_in = {'kin': evolution_op(self.dt_in / 2, self.kin_eng_spin),
                       'pot': evolution_op(self.dt_in, self.pot_eng_spin),
                       'coupl': coupling_op(self.dt_in / 2, self.coupling, self.expon)}

    def full_step(self):
        """tensor_propagator.py:214-222."""
        self.single_step(self.dt_out, self.eng_out)
        self.single_step(self.dt_in, self.eng_in)
        self.single_step(self.dt_out, self.eng_out)

    def single_step(self, t_step, eng):
        """tensor_propagator.py:240-271."""
        psik = [e * pk for e, pk in zip(eng['kin'], self.psik)]
        psi = ifft_2d(psik, delta_r=self.space['dr'])
        psi, dens = norm(psi, self.space['dv_r'], self.atom_num)
        int_eng = [self.g_sc['uu'] * dens[0] + self.g_sc['ud'] * dens[1],
                   self.g_sc['dd'] * dens[1] + self.g_sc['ud'] * dens[0]]
        int_op = evolution_op(t_step / 2, int_eng)
        psi = [op * p for op, p in zip(int_op, psi)]
        if self.is_coupling:
            psi = [sum([elem * p for elem, p in zip(row, psi)]) for row in eng['coupl']]
        psi = [e * p for e, p in zip(eng['pot'], psi)]
        if self.is_coupling:
            psi = [sum([elem * p for elem, p in zip(row, psi)]) for row in eng['coupl']]
        psi = [op * p for op, p in zip(int_op, psi)]
        psik = fft_2d(psi, delta_r=self.space['dr'])
        psik = [e * pk for e, pk in zip(eng['kin'], psik)]
        self.psik, _ = norm(psik, self.space['dv_k'], self.atom_num)

    def run(self, n_steps):
        """The prop_loop body without sampling (tensor_propagator.py:185-194).

        Returns the (n_steps, 2) populations.
        """
        vals = np.empty((n_steps, 2))
        for i in range(n_steps):
            self.full_step()
            vals[i] = calc_pops(self.psik, self.space['dv_k'])
        return vals

    def psik_numpy(self):
        return [p.numpy().copy() for p in self.psik]

    def psi_numpy(self):
        """Real-space state via the NumPy ifft_2d (tensor_propagator.py:208-209)."""
        dr = self.space['dr'].numpy()
        c = dr[0] * dr[1] / (2 * np.pi)
        return [np.fft.ifftn(np.fft.ifftshift(p.numpy())) / c for p in self.psik]
