"""Result analysis and figures on the host: PropResult / PSpinor plotting,
the progress text, file naming and the sampled-state movie.

The figures' content is checked through what matplotlib holds: each image's
array is the density or masked phase of the state it was given, the extents
are the grid half-widths over the scale, and the saved file names follow the
``<stem><i>-<folder><ext>`` pattern with the first free index.
"""

import os
import warnings

import matplotlib
import numpy as np
import pytest

matplotlib.use('Agg')
from matplotlib import pyplot as plt  # noqa: E402  pylint: disable=wrong-import-position

from spinor_gpe.pspinor import plotting_tools as ptools  # noqa: E402
from spinor_gpe.pspinor import prop_result  # noqa: E402
from spinor_gpe.pspinor import tensor_tools as ttools  # noqa: E402

import cases  # noqa: E402

pytestmark = pytest.mark.filterwarnings('ignore:FigureCanvasAgg is non-interactive')


@pytest.fixture(autouse=True)
def _close_figures():
    yield
    plt.close('all')


def _spinor():
    ps, _, _ = cases.ground('plots', mesh=(64, 32), r=(8, 4), seed=7)
    # a vortex in component 1 so that its phase is not flat
    ps.seed_vortices([(1.0, 0.5)], [1], components=(1,))
    return ps


def _result(ps, n_pops=6):
    times = np.linspace(0.0, 1.0, n_pops)
    vals = np.stack([0.5 + 0.1 * times, 0.5 - 0.1 * times], axis=1) * ps.atom_num
    res = prop_result.PropResult(ps.psi, ps.psik, [1.0, 0.5, 0.3, 0.2],
                                 {'times': times, 'vals': vals})
    res.paths, res.time_scale, res.space = ps.paths, ps.time_scale, ps.space
    return res


def _masked_phase(psi):
    d = np.abs(psi) ** 2
    ph = np.angle(psi)
    ph[d < 1e-6 * d.max()] = 0
    return ph


def test_time_remaining_format():
    assert ptools.time_remaining(0, 3661, 1.0) == '[01:01:01]'
    assert ptools.time_remaining(10, 100, 2.0) == '[00:00:45]'
    assert ptools.time_remaining(0, 59.9, 1.0) == '[00:00:59]'  # rounded down
    assert ptools.time_remaining(0, 360000, 1.0) == '[100:00:00]'


def test_progress_message(capsys):
    ptools.progress_message(0, 10)
    assert capsys.readouterr().out == ''
    ptools.progress_message(1, 10)
    out = capsys.readouterr().out
    assert out.startswith('\r1/10, [') and out.endswith(' it/sec')


def test_next_available_path(tmp_path):
    stem = str(tmp_path / 'spin_dens_phase')
    assert ptools.next_available_path(stem, 'run', '.pdf') == stem + '1-run.pdf'
    open(stem + '1-run.pdf', 'w', encoding='utf8').close()
    open(stem + '2-run.pdf', 'w', encoding='utf8').close()
    assert ptools.next_available_path(stem, 'run', '.pdf') == stem + '3-run.pdf'


def test_result_plot_spins_content_and_file():
    ps = _spinor()
    res = _result(ps)
    fig, plots = res.plot_spins(rscale=2.0, kscale=4.0, save=True, ext='.png', show=False,
                                zoom=2.0)
    assert set(plots) == {'r', 'ph', 'k'} and all(len(v) == 2 for v in plots.values())
    for c in range(2):
        np.testing.assert_allclose(plots['r'][c].get_array(), np.abs(ps.psi[c]) ** 2)
        np.testing.assert_allclose(plots['ph'][c].get_array(), _masked_phase(ps.psi[c]))
        np.testing.assert_allclose(plots['k'][c].get_array(), np.abs(ps.psik[c]) ** 2)
    rs, ks = ps.space['r_sizes'], ps.space['k_sizes']
    np.testing.assert_allclose(plots['r'][0].get_extent(), [-rs[0] / 2, rs[0] / 2, -rs[1] / 2, rs[1] / 2])
    np.testing.assert_allclose(plots['k'][1].get_extent(), [-ks[0] / 4, ks[0] / 4, -ks[1] / 4, ks[1] / 4])
    k_ax = plots['k'][0].axes
    np.testing.assert_allclose(k_ax.get_xlim(), (-ks[0] / 8, ks[0] / 8))
    assert plots['ph'][0].get_clim() == (-np.pi, np.pi)
    assert len(fig.axes) == 12  # six images, six colour bars
    name = ps.paths['data'] + 'spin_dens_phase1-' + ps.paths['folder'] + '.png'
    assert os.path.getsize(name) > 0
    res.plot_spins(save=True, ext='.png', show=False)
    assert os.path.exists(name.replace('phase1-', 'phase2-'))


def test_result_plot_total():
    ps = _spinor()
    res = _result(ps)
    _, plots = res.plot_total(save=True, ext='.png', show=False)
    np.testing.assert_allclose(plots['r'].get_array(), sum(np.abs(p) ** 2 for p in ps.psi))
    np.testing.assert_allclose(plots['ph'].get_array(), _masked_phase(ps.psi[0] + ps.psi[1]))
    np.testing.assert_allclose(plots['k'].get_array(), sum(np.abs(p) ** 2 for p in ps.psik))
    assert os.path.exists(ps.paths['data'] + 'total_dens_phase1-' + ps.paths['folder'] + '.png')


def test_result_plot_pops():
    ps = _spinor()
    res = _result(ps)
    fig = res.plot_pops(scaled=True, save=True, ext='.png')
    ax0, ax1 = fig.axes
    lines = ax0.get_lines()
    np.testing.assert_allclose(lines[0].get_xdata(), res.pops['times'] * ps.time_scale)
    np.testing.assert_allclose(lines[1].get_ydata(), res.pops['vals'][:, 1])
    diff = ax1.get_lines()[0].get_ydata()
    np.testing.assert_allclose(diff, 0.2 * res.pops['times'] * ps.atom_num, rtol=1e-12, atol=1e-9)
    assert ax1.get_yscale() == 'log'
    assert os.path.exists(ps.paths['data'] + 'pop_evolution1-' + ps.paths['folder'] + '.png')
    fig = res.plot_pops(scaled=False, save=False)
    np.testing.assert_allclose(fig.axes[0].get_lines()[0].get_xdata(), res.pops['times'])


def test_not_provided_as_in_reference():
    ps = _spinor()
    res = _result(ps)
    with pytest.raises(NotImplementedError):
        res.plot_eng()
    with pytest.raises(NotImplementedError):
        res.analyze_vortex()
    with pytest.raises(NotImplementedError):
        ps.seed_regular_vortices()
    with pytest.raises(NotImplementedError):
        ps.seed_random_vortices(3)


def test_spinor_plots():
    ps = _spinor()
    fig, axs = ps.plot_rdens(spin=1, scale=2.0)
    assert len(axs) == 1
    np.testing.assert_allclose(axs[0].images[0].get_array(), np.abs(ps.psi[1]) ** 2)
    np.testing.assert_allclose(axs[0].images[0].get_extent(),
                               np.array([-8, 8, -4, 4]) / 2.0)
    _, axs = ps.plot_kdens()
    assert len(axs) == 2
    np.testing.assert_allclose(axs[0].images[0].get_array(), np.abs(ps.psik[0]) ** 2)
    _, axs = ps.plot_rphase(spin=1)
    np.testing.assert_allclose(axs[0].images[0].get_array(), _masked_phase(ps.psi[1]))
    with pytest.raises(AssertionError):
        ps.plot_rdens(spin=2)
    fig, _ = ps.plot_spins(save=False)
    assert len(fig.axes) == 12


def test_make_movie_frames(tmp_path):
    ps = _spinor()
    res = _result(ps)
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter('always')
        assert res.make_movie() is None
    assert any('No sampled' in str(w.message) for w in caught)

    # three samples: the state scaled by 1, 2, 3 (the last frame stays drawn)
    psiks = np.stack([np.stack(ps.psik) * s for s in (1.0, 2.0, 3.0)])
    res.sampled_path = str(tmp_path / 'psik_sampled1-plots.npz')
    np.savez(res.sampled_path, psiks=psiks, times=np.array([0.0, 0.5, 1.0]))
    real = []
    orig = res.plot_spins

    def spy(*args, **kwargs):
        out = orig(*args, **kwargs)
        real.append(out[1])
        return out
    res.plot_spins = spy
    name = res.make_movie(norm_type='half', writer='pillow')
    assert name == ps.paths['data'] + 'prop_movie1-' + ps.paths['folder'] + '.gif'
    assert os.path.getsize(name) > 0
    plots = real[0]
    last = ttools.ifft_2d(list(psiks[-1]), ps.space['dr'])
    dens = [np.abs(p) ** 2 for p in last]
    np.testing.assert_allclose(plots['r'][1].get_array(), dens[1], rtol=1e-12)
    np.testing.assert_allclose(plots['k'][0].get_array(), np.abs(psiks[-1][0]) ** 2, rtol=1e-12)
    top = (dens[0].max() + dens[1].max()) / 2
    np.testing.assert_allclose(plots['r'][0].get_clim(), (0, top), rtol=1e-12)
    # the real-space frame is the state itself (ifft of its own transform), times 3
    np.testing.assert_allclose(dens[0], 9 * np.abs(ps.psi[0]) ** 2, rtol=1e-9, atol=1e-12 * dens[0].max())
    with pytest.raises(ValueError):
        res.make_movie(writer='avconv')
