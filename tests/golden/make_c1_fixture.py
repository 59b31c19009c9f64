"""Write the oracle fixture of the 512^2 real-time Raman run (20000 steps).

Runs only the CPU oracle (``oracle/reference_path.py``), never the reference
or the engine: the 256^2 imaginary-time ground state (4000 steps, dt = 0.02,
complex128) is computed with the oracle, resampled to the 512^2 Raman
problem (``workloads.c1_spinor``: kL = 3, Omega = 2 E_R, delta = 0.5 E_R,
lab frame, seeded phase noise), and propagated by the oracle for 20000 real
time steps of dt = 1e-3. The fixture keeps, instead of the 8 MiB state:

* ``c0_sub``: the ground state's psik on the stride-8 sublattice (checks that
  the test's own oracle ground state is the one used here);
* at each checkpoint step ``ck_steps``: ``ck_sub`` (psik on the stride-16
  sublattice of both components) and ``ck_proj`` (``PROJ_K`` projections
  sum_i r_ki psik_i of the flattened [2, 512, 512] state on seeded
  Rademacher vectors r_k = +-1 from ``numpy.random.default_rng(PROJ_SEED)``:
  E|sum_i r_ki e_i|^2 = ||e||^2 for any error e, so they bound the full
  state's error, not just the sublattice's) and ``ck_norm`` (||psik||_2);
* ``pops``: the per-step populations at every 10th step (steps 10, 20, ...);
* ``gp``: the GP energy terms of the final state (E_tot, E_kin, E_pot,
  E_int, E_cpl, N, mu).

With ``--numpy-fft`` the oracle's transforms are NumPy's (pocketfft)
instead of torch's, everything else unchanged, and the fixture goes to
``c1_oracle_20000_npfft.npz``: two FP64 restatements of the same path that
differ only in the rounding of their FFTs. Their distance measures how far
rounding alone carries this run: the dynamics amplify per-step rounding
differences. Measured between the two fixtures (sublattice relative L2 /
projections relative to ||psik||): 2.3e-14 / 6.7e-14 after 100 steps,
1.8e-13 / 1.4e-12 after 1000, 1.7e-12 / 4.7e-12 after 5000, 9.6e-12 /
2.3e-11 after 10000, 4.4e-10 / 5.5e-10 after 20000; populations 7.2e-12 of
N_at; final GP energy terms 1e-12 to 1e-11 (relative).

Usage: ``python tests/golden/make_c1_fixture.py [--numpy-fft]`` (about 45
minutes (torch FFT) or 85 minutes (NumPy FFT) on 8 cores); writes
``tests/golden/c1_oracle_20000.npz`` (``..._npfft.npz``).
"""

import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for _p in (os.path.join(ROOT, 'spinor-gpe_amd'), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

CK_STEPS = (100, 1000, 5000, 10000, 20000)
PROJ_K = 8
PROJ_SEED = 20250101
SUB_C0 = 8
SUB_CK = 16
POPS_EVERY = 10
OUT = os.path.join(HERE, 'c1_oracle_20000.npz')


def rademacher(n, k=PROJ_K, seed=PROJ_SEED):
    """The k seeded +-1 projection vectors of length n (int8)."""
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 2, size=(k, n), dtype=np.int8) * 2 - 1).astype(np.int8)


def project(psik, r):
    """Projections sum_i r_ki psik_i of the flattened state (complex128)."""
    v = np.asarray(psik, dtype=np.complex128).ravel()
    return np.array([np.dot(rk.astype(np.float64), v.real) + 1j * np.dot(rk.astype(np.float64), v.imag)
                     for rk in r])


def c0_ground_oracle():
    """The 256^2 ground state by the oracle (workloads.c0_spinor, 4000 steps)."""
    from oracle.reference_path import OraclePropagator  # pylint: disable=import-outside-toplevel
    from spinor_gpe.engine import workloads as wl  # pylint: disable=import-outside-toplevel
    ps0 = wl.c0_spinor('c0-golden')
    o = OraclePropagator(ps0, 0.02, time='imag')
    o.run(4000)
    return ps0, o


def use_numpy_fft():
    """Swap the oracle's torch FFTs for NumPy's (same conventions)."""
    import torch  # pylint: disable=import-outside-toplevel
    from oracle import reference_path as rp  # pylint: disable=import-outside-toplevel

    def fft_2d(psi, delta_r):
        c = rp.prod(delta_r) / (2 * np.pi)
        return [torch.from_numpy(np.fft.fftshift(np.fft.fftn(p.numpy()) * c.item())) for p in psi]

    def ifft_2d(psik, delta_r):
        c = rp.prod(delta_r) / (2 * np.pi)
        return [torch.from_numpy(np.fft.ifftn(np.fft.ifftshift(pk.numpy())) / c.item()) for pk in psik]

    rp.fft_2d = fft_2d
    rp.ifft_2d = ifft_2d


def main():
    from oracle import energy as oe  # pylint: disable=import-outside-toplevel
    from oracle.reference_path import OraclePropagator  # pylint: disable=import-outside-toplevel
    from spinor_gpe.engine import workloads as wl  # pylint: disable=import-outside-toplevel
    out = OUT
    if '--numpy-fft' in sys.argv:
        use_numpy_fft()
        out = OUT.replace('.npz', '_npfft.npz')
    os.environ.setdefault('SPINOR_GPE_DATA', tempfile.mkdtemp(prefix='sgpe-golden-'))
    t0 = time.time()
    ps0, o0 = c0_ground_oracle()
    psik0 = np.stack(o0.psik_numpy())
    print(f"c0 ground state done ({time.time() - t0:.0f} s)", flush=True)
    ps = wl.c1_spinor(o0.psi_numpy(), ps0.space, name='c1-golden')
    init = np.stack(ps.psik)
    r = rademacher(init.size)
    o = OraclePropagator(ps, 1e-3, time='real')
    pops = []
    ck_sub, ck_proj, ck_norm = [], [], []
    done = 0
    for ck in CK_STEPS:
        vals = o.run(ck - done)
        pops.append(vals)
        done = ck
        st = np.stack(o.psik_numpy())
        ck_sub.append(st[:, ::SUB_CK, ::SUB_CK])
        ck_proj.append(project(st, r))
        ck_norm.append(np.linalg.norm(st))
        print(f"step {ck} done ({time.time() - t0:.0f} s)", flush=True)
    pops = np.concatenate(pops)[POPS_EVERY - 1::POPS_EVERY]
    final = np.stack(o.psik_numpy())
    theta = 2 * ps.kL_recoil * ps.space['x_mesh']
    g = oe.gp_energy(final, ps.space['dr'], ps.kin_eng_spin, ps.pot_eng_spin, ps.g_sc,
                     ps.coupling, theta)
    gp = np.array([g[k] for k in ('E_tot', 'E_kin', 'E_pot', 'E_int', 'E_cpl', 'N', 'mu')])
    np.savez_compressed(out, c0_sub=psik0[:, ::SUB_C0, ::SUB_C0], init_sub=init[:, ::SUB_CK, ::SUB_CK],
                        init_proj=project(init, r), ck_steps=np.array(CK_STEPS), ck_sub=np.stack(ck_sub),
                        ck_proj=np.stack(ck_proj), ck_norm=np.array(ck_norm), pops=pops,
                        pops_every=POPS_EVERY, gp=gp, proj_seed=PROJ_SEED, proj_k=PROJ_K,
                        sub_c0=SUB_C0, sub_ck=SUB_CK)
    print(f"wrote {out} ({time.time() - t0:.0f} s)", flush=True)


if __name__ == '__main__':
    main()
