"""Result of a propagation (spinor_gpe/pspinor/prop_result.py).

Same constructor, fields and analysis/plotting methods as the reference
(prop_result.py:19-328). Engine-specific extras are added as attributes by
:class:`TensorPropagator`: ``energy_gp`` (the spectral GP energy terms),
``chem_pot``, ``steps_run`` and ``converged``. ``paths``, ``time_scale`` and
``space`` are filled in by ``PSpinor.imaginary/real``.
"""

import os
import subprocess
import sys
import warnings

import numpy as np

from spinor_gpe.pspinor import plotting_tools as ptools
from spinor_gpe.pspinor import tensor_tools as ttools

POP_LABELS = ('Pop. $| \\uparrow\\rangle$', 'Pop. $| \\downarrow\\rangle$')
# movie writers: matplotlib writer name -> file extension
MOVIE_EXT = {'ffmpeg': '.mp4', 'pillow': '.gif'}


class PropResult:
    """Final wavefunctions, energies and populations of a run."""

    def __init__(self, psi_final, psik_final, eng_final, pops, sampled_path=None):
        self.psi = psi_final
        self.psik = psik_final
        self.eng_final = eng_final
        self.pops = pops
        self.sampled_path = sampled_path
        self.dens = ttools.density(self.psi)
        self.densk = ttools.density(self.psik)
        self.phase = ttools.phase(self.psi, uwrap=False, dens=self.dens)
        self.paths = dict()
        self.time_scale = None
        self.space = dict()

    def calc_separation(self):
        """Spin separation ``1 - <n0 n1> / sqrt(<n0^2><n1^2>)`` (prop_result.py:84-88)."""
        return 1 - (np.sum(ttools.prod(self.dens))
                    / np.sqrt(np.sum(self.dens[0] ** 2) * np.sum(self.dens[1] ** 2)))

    # ------------------------------------------------------------- figures
    def _extents(self, rscale, kscale):
        """``{'r': [-x, x, -y, y] / rscale, 'k': [-kx, kx, -ky, ky] / kscale}``
        from the grid half-widths."""
        out = {}
        for key, scale in (('r', rscale), ('k', kscale)):
            half = np.asarray(self.space[key + '_sizes'], dtype=float)
            out[key] = np.array([-half[0], half[0], -half[1], half[1]]) / scale
        return out

    def plot_spins(self, rscale=1.0, kscale=1.0, cmap='viridis', save=True, ext='.pdf',
                   show=True, zoom=1.0):
        """Densities (real and k-space) and phases of both components
        (prop_result.py:90-127); lengths in units of ``rscale`` and
        ``kscale`` (1.0: the x oscillator length and its inverse). Saves
        ``spin_dens_phase<i>-<folder><ext>`` in the data folder when ``save``.
        Returns ``(fig, {'r', 'ph', 'k'})``."""
        return ptools.plot_spins(self.psi, self.psik, self._extents(rscale, kscale), self.paths,
                                 cmap=cmap, save=save, ext=ext, show=show, zoom=zoom)

    def plot_total(self, rscale=1.0, kscale=1.0, cmap='viridis', save=True, ext='.pdf',
                   show=True, zoom=1.0):
        """Total densities and the phase of ``psi_0 + psi_1``
        (prop_result.py:129-166); saves ``total_dens_phase<i>-<folder><ext>``."""
        return ptools.plot_total(self.psi, self.psik, self._extents(rscale, kscale), self.paths,
                                 cmap=cmap, save=save, ext=ext, show=show, zoom=zoom)

    def plot_eng(self):
        """Sampled energies: not provided, as in the reference (prop_result.py:168-170)."""
        raise NotImplementedError()

    def plot_pops(self, scaled=True, save=True, ext='.pdf'):
        """Populations against time (left) and ``|N_up - N_down|`` on a log
        scale (right) (prop_result.py:172-214). ``scaled``: time in seconds
        (times ``time_scale``) instead of ``1/omega_x``. Saves
        ``pop_evolution<i>-<folder><ext>`` when ``save``. Returns the figure."""
        plt = ptools._pyplot()  # pylint: disable=protected-access
        if scaled:
            xlabel, scale = 'Time [s]', self.time_scale
        else:
            xlabel, scale = 'Time [$1/\\omega_x$]', 1.0
        times = np.asarray(self.pops['times']) * scale
        vals = np.asarray(self.pops['vals'])
        fig = plt.figure(figsize=(12, 4))
        ax0 = fig.add_subplot(121)
        ax0.legend(ax0.plot(times, vals), POP_LABELS)
        ax0.set_ylabel('Population')
        ax1 = fig.add_subplot(122)
        ax1.plot(times, np.abs(np.diff(vals)))
        ax1.set_ylabel('Abs. Population Difference')
        ax1.set_yscale('log')
        ax1.set_ylim(2e-16, None)
        for ax in (ax0, ax1):
            ax.set_xlabel(xlabel)
            ax.grid(alpha=0.5)
        if save:
            fig.savefig(ptools.next_available_path(self.paths['data'] + 'pop_evolution',
                                                   self.paths['folder'], ext))
        plt.show()
        return fig

    def analyze_vortex(self):
        """Vorticity per component: not provided, as in the reference
        (prop_result.py:216-218)."""
        raise NotImplementedError()

    def make_movie(self, rscale=1.0, kscale=1.0, cmap='viridis', play=False, zoom=1.0,
                   norm_type='all', *, writer='ffmpeg', fps=5):
        """Movie of the sampled states' densities and phases
        (prop_result.py:220-302): one frame per sample of ``sampled_path``,
        laid out as :meth:`plot_spins`; each frame's density colour scales run
        from 0 to the sum of both components' maxima (``norm_type='all'``) or
        half of it (``'half'``). Saved as ``prop_movie<i>-<folder>.mp4`` in the
        data folder; ``play`` opens it in the system viewer. Warns and returns
        None when no samples were stored.

        ``writer`` (keyword-only) picks the matplotlib movie writer: the
        reference's ``'ffmpeg'`` (.mp4), or ``'pillow'`` (.gif) where ffmpeg
        is not installed. Returns the file name.
        """
        from matplotlib import animation  # pylint: disable=import-outside-toplevel
        plt = ptools._pyplot()  # pylint: disable=protected-access
        if not os.path.exists(str(self.sampled_path)):
            warnings.warn("Cannot generate propagation movie. No sampled "
                          "wavefuntion data exists.")
            return None
        divisor = {'all': 1.0, 'half': 2.0}[norm_type]
        if writer not in MOVIE_EXT:
            raise ValueError(f"writer must be one of {sorted(MOVIE_EXT)}, not {writer!r}")
        with np.load(self.sampled_path) as sampled:
            times = sampled['times']
            psiks = sampled['psiks']
        n_frames = len(times)
        fig, plots = self.plot_spins(rscale, kscale, cmap, save=False, show=False, zoom=zoom)

        def draw(frame):
            psik = list(psiks[frame])
            psi = ttools.ifft_2d(psik, self.space['dr'])
            dens = ttools.density(psi)
            densk = ttools.density(psik)
            phases = ttools.phase(psi, uwrap=False, dens=dens)
            for key, data in (('r', dens), ('ph', phases), ('k', densk)):
                for image, arr in zip(plots[key], data):
                    image.set_data(arr)
            for key, data in (('r', dens), ('k', densk)):
                top = sum(float(np.max(d)) for d in data) / divisor
                for image in plots[key]:
                    image.set_clim(0, top)
            ptools.progress_message(frame, n_frames)

        anim = animation.FuncAnimation(fig, draw, frames=n_frames, blit=False)
        file_name = ptools.next_available_path(self.paths['data'] + 'prop_movie',
                                               self.paths['folder'], MOVIE_EXT[writer])
        anim.save(file_name, writer=animation.writers[writer](fps=fps))
        plt.close(fig)
        if play:
            if sys.platform == 'win32':
                os.startfile(file_name)  # pylint: disable=no-member
            else:
                subprocess.call(['open' if sys.platform == 'darwin' else 'xdg-open', file_name])
        return file_name

    def rebin(self, arr, new_shape=(256, 256)):
        """Block-average both arrays of ``arr`` down to ``new_shape``
        (prop_result.py:304-328); returned unchanged if not smaller."""
        assert arr[0].shape == arr[1].shape
        if not tuple(new_shape) < tuple(arr[0].shape):
            return arr
        out = []
        for a in arr:
            shape = (new_shape[0], a.shape[0] // new_shape[0],
                     new_shape[1], a.shape[1] // new_shape[1])
            out.append(a.reshape(shape).mean(-1).mean(1))
        return out
