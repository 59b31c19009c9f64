"""The reference's four example workflows through this package (GPU).

Each test follows one script of the reference's ``spinor_gpe/examples``
(``1_ground_state.py``, ``2_anisotropic_tof.py``, ``3_raman_rabi.py``,
``4_detuning_grad.py``): the same PSpinor arguments, set-up calls
(``coupling_setup``, ``shift_momentum``, ``coupling_uniform``,
``detuning_grad``, ``rot_coupling``, the ``pot_eng`` setter), time steps and
device string, then the analysis calls the script makes on the results
(``plot_spins``, ``plot_total``, ``plot_pops``, ``make_movie``). Run lengths
are shortened so that the CPU oracle, started from the same state before
every propagation, finishes in seconds; sampling keeps the scripts' interval
where the shortened run allows it. Every propagation is checked against the
oracle: k-space state to a relative L2 error of 1e-11 and the per-step
populations to 1e-11 of the atom number (complex128, the scripts' precision).
Example 1 asks for ``device='cpu'``, the reference's default: this engine
has no CPU propagation path, so that run goes to the GPU with a warning and
hands back host tensors, as the reference's CPU run would.
"""

import matplotlib
import numpy as np
import pytest

from oracle.reference_path import OraclePropagator
from spinor_gpe.pspinor import pspinor as spin

matplotlib.use('Agg')
import matplotlib.pyplot as plt  # noqa: E402  pylint: disable=wrong-import-position

pytestmark = pytest.mark.gpu

W = 2 * np.pi * 50


def rel(a, b):
    a, b = np.stack(a), np.stack(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def propagate(ps, kind, dt, n, device='cuda', **kw):
    """``ps.imaginary`` / ``ps.real`` checked against the oracle from the same start."""
    print(f"[examples] {ps.paths['folder']} {kind} dt={dt} steps={n} grid={ps.psi[0].shape}", flush=True)
    oracle = OraclePropagator(ps, dt, time=kind)
    ref_pops = oracle.run(n)
    runner = ps.imaginary if kind == 'imag' else ps.real
    res, prop = runner(dt, n, device, progress=False, **kw)
    err = rel(res.psik, oracle.psik_numpy())
    perr = float(np.max(np.abs(res.pops['vals'] - ref_pops)) / ps.atom_num)
    assert err < 1e-11, err
    assert perr < 1e-11, perr
    assert ps.psik is res.psik
    assert np.all(np.isfinite(res.eng_final))
    return res, prop


def analyze(ps, res, zoom, movie=True):
    """The scripts' analysis calls (movies written with pillow: no viewer)."""
    res.plot_spins(rscale=ps.rad_tf, kscale=ps.kL_recoil, zoom=zoom)
    res.plot_total(kscale=ps.kL_recoil, zoom=zoom)
    res.plot_pops()
    if movie:
        name = res.make_movie(rscale=ps.rad_tf, kscale=ps.kL_recoil, play=False, zoom=zoom,
                              norm_type='half', writer='pillow')
        assert name
    plt.close('all')


def test_example_1_ground_state(gpu):
    """1_ground_state.py: 64^2, 100 atoms, g_ud = 1.04, no Raman shift;
    100 imaginary steps of 1/50 with 50 samples (the script's full run)."""
    ps = spin.PSpinor('examples/Trial_011', overwrite=True, atom_num=1e2,
                      omeg={'x': W, 'y': W, 'z': 40.0 * W},
                      g_sc={'uu': 1, 'dd': 1, 'ud': 1.04}, pop_frac=(0.5, 0.5),
                      r_sizes=(8, 8), mesh_points=(64, 64))
    ps.coupling_setup(wavel=790.1e-9, kin_shift=False)
    ps.plot_spins(rscale=ps.rad_tf, kscale=ps.kL_recoil, zoom=4)
    ps.rand_seed = 99999
    with pytest.warns(RuntimeWarning, match='runs only on HIP GPUs'):
        res, prop = propagate(ps, 'imag', 1 / 50, 100, device='cpu', is_sampling=True, n_samples=50)
    assert all(p.device.type == 'cpu' for p in prop.psik)
    with np.load(res.sampled_path) as f:
        assert f['psiks'].shape == (50, 2, 64, 64)
    analyze(ps, res, 4)


def test_example_2_anisotropic_tof(gpu):
    """2_anisotropic_tof.py: 512^2 over +-32 in a 1:4 trap; imaginary time,
    then the trap removed through the ``pot_eng`` setter and a real-time
    expansion with samples (40 + 40 steps; samples every 4 steps)."""
    ps = spin.PSpinor('examples/Trial_002', overwrite=True, atom_num=1e4,
                      omeg={'x': W, 'y': 4 * W, 'z': 40.0 * W},
                      g_sc={'uu': 1, 'dd': 1, 'ud': 0.5}, phase_factor=1, pop_frac=(0.5, 0.5),
                      r_sizes=(32, 32), mesh_points=(512, 512))
    ps.coupling_setup(wavel=790.1e-9)
    ps.plot_spins(rscale=ps.rad_tf, kscale=ps.kL_recoil, zoom=4)
    ps.rand_seed = 99999
    res0, _ = propagate(ps, 'imag', 1 / 50, 40, is_sampling=False, n_samples=50)
    analyze(ps, res0, 4, movie=False)
    ps.pot_eng = np.zeros_like(ps.pot_eng)
    assert not np.any(ps.pot_eng_spin[0]) and not np.any(ps.pot_eng_spin[1])
    res1, _ = propagate(ps, 'real', 1 / 500, 40, is_sampling=True, n_samples=10)
    analyze(ps, res1, 2)


def test_example_3_raman_rabi(gpu):
    """3_raman_rabi.py: 256^2, all atoms in one component, Raman-shifted
    dispersions, momentum shift, lab-frame coupling; imaginary time, then
    Rabi flopping in real time at Omega = E_R (50 + 100 steps)."""
    ps = spin.PSpinor('examples/Trial_003', overwrite=True, atom_num=1e4,
                      omeg={'x': W, 'y': W, 'z': 40.0 * W},
                      g_sc={'uu': 1, 'dd': 1, 'ud': 0.0}, pop_frac=(1.0, 0.0),
                      r_sizes=(16, 16), mesh_points=(256, 256))
    ps.coupling_setup(wavel=790.1e-9, kin_shift=True)
    ps.shift_momentum(scale=1.0, frac=(0, 1.0))
    ps.rot_coupling = False
    ps.plot_spins(rscale=ps.rad_tf, kscale=ps.kL_recoil, zoom=2)
    ps.rand_seed = 99999
    res0, _ = propagate(ps, 'imag', 1 / 50, 50, is_sampling=True, n_samples=5)
    analyze(ps, res0, 2, movie=False)
    ps.coupling_uniform(1.0 * ps.EL_recoil)
    res1, _ = propagate(ps, 'real', 1 / 5000, 100, is_sampling=True, n_samples=10)
    # populations move from component 0 to component 1 under the coupling
    assert res1.pops['vals'][-1, 1] > res1.pops['vals'][0, 1]
    analyze(ps, res1, 1)


def test_example_4_detuning_grad(gpu):
    """4_detuning_grad.py: 256^2, Raman coupling 5 E_R with a detuning
    gradient of -12 in the rotated frame, 804 nm Raman beams (50 steps)."""
    ps = spin.PSpinor('examples/Trial_008', overwrite=True, atom_num=1e4,
                      omeg={'x': W, 'y': W, 'z': 40.0 * W},
                      g_sc={'uu': 1, 'dd': 0.995, 'ud': 0.995}, pop_frac=(0.5, 0.5),
                      r_sizes=(16, 16), mesh_points=(256, 256))
    ps.coupling_setup(wavel=804e-9, kin_shift=True)
    ps.shift_momentum(scale=0.6, frac=(0.5, 0.5))
    ps.coupling_uniform(5 * ps.EL_recoil)
    ps.detuning_grad(-12)
    ps.rot_coupling = True
    ps.plot_spins(rscale=ps.rad_tf, kscale=ps.kL_recoil, zoom=2)
    ps.rand_seed = 99999
    res, _ = propagate(ps, 'imag', 1 / 50, 50, is_sampling=True, n_samples=5)
    analyze(ps, res, 2)
