"""Figures, progress text and file naming for propagation results
(spinor_gpe/pspinor/plotting_tools.py).

Same functions, arguments and saved file names as the reference:

* ``next_available_path`` (plotting_tools.py:13-37): ``<file_name><i>-<trial_name><ext>``
  with the first unused ``i >= 1``;
* ``progress_message`` / ``time_remaining`` (:38-101): a one-line progress
  report ``'\\r<frame>/<total>, [HH:MM:SS], <rate> it/sec'``;
* ``plot_dens`` / ``plot_phase`` (:104-187): one image per spin component;
* ``plot_spins`` (:190-300): real-space densities, phases and k-space
  densities of both components on a 6 x 4 grid (file ``spin_dens_phase``);
* ``plot_total`` (:303-386): total density, phase of ``psi_0 + psi_1`` and
  total k-space density (file ``total_dens_phase``).

matplotlib is imported on first use, so the propagation path does not need
it. The figure functions also return what they draw (``plot_dens`` and
``plot_phase`` return ``(fig, axes)``; the reference returns nothing there).
"""

import os
import sys
import time

import numpy as np

from spinor_gpe.pspinor import tensor_tools as ttools

# time of the previous progress_message call
_PROGRESS = {'last': None}

PHASE_TICKS = np.linspace(-np.pi, np.pi, 5)
PHASE_LABELS = ['$-\\pi$', '', '$0$', '', '$\\pi$']


def _pyplot():
    from matplotlib import pyplot as plt  # pylint: disable=import-outside-toplevel
    return plt


def next_available_path(file_name, trial_name, ext=''):
    """Return the first ``file_name + str(i) + '-' + trial_name + ext`` that does not exist."""
    i = 1
    while True:
        path = f"{file_name}{i}-{trial_name}{ext}"
        if not os.path.exists(path):
            return path
        i += 1


def time_remaining(frame, n_total, its):
    """``'[HH:MM:SS]'`` left for ``n_total - frame`` iterations at ``its``
    iterations per second (whole seconds, rounded down; hours may exceed two
    digits)."""
    left = int(np.floor((n_total - frame) / its))
    hours, rest = divmod(left, 3600)
    minutes, seconds = divmod(rest, 60)
    return f'[{hours:02d}:{minutes:02d}:{seconds:02d}]'


def progress_message(frame, n_total):
    """Rewrite the progress line for ``frame`` of ``n_total``.

    Frame 0 only starts the clock; later frames print the count, the time
    left at the rate of the last iteration, and that rate. Returns the time
    of this call (``time.time()``).
    """
    now = time.time()
    last = _PROGRESS['last']
    _PROGRESS['last'] = now
    if frame != 0 and last is not None:
        rate = 1.0 / max(now - last, 1e-9)
        sys.stdout.write(f'\r{frame}/{n_total}, {time_remaining(frame, n_total, rate)}, '
                         f'{rate:.2f} it/sec')
        sys.stdout.flush()
    return now


def _components(psi, spin):
    if spin is None:
        return list(psi)
    assert spin in (0, 1), f"The `spin` parameter should be 0 or 1, not {spin}."
    return [psi[spin]]


def _image_row(images, cmap, extent):
    plt = _pyplot()
    fig, axs = plt.subplots(1, len(images), sharex=True, sharey=True, squeeze=False)
    axs = axs[0]
    for ax, img in zip(axs, images):
        ax.imshow(img, cmap=cmap, extent=extent)
    plt.show()
    return fig, axs


def plot_dens(psi, spin=None, cmap='viridis', scale=1., extent=None):
    """Density of both components (``spin=None``) or of component ``spin``
    (0 up, 1 down). ``scale`` is accepted for the reference's signature and,
    as there, not used; ``extent`` = ``[x_min, x_max, y_min, y_max]``."""
    del scale
    return _image_row(ttools.density(_components(psi, spin)), cmap, extent)


def plot_phase(psi, spin=None, cmap='twilight_shifted', scale=1, extent=None):
    """Phase of both components or of component ``spin``, zero where the
    density is below 1e-6 of its maximum."""
    del scale
    comps = _components(psi, spin)
    dens = ttools.density(comps)
    phases = ttools.phase(comps, uwrap=False, dens=dens)
    for ph, d in zip(phases, dens):
        ph[d < 1e-6 * np.max(d)] = 0
    return _image_row(phases, cmap, extent)


def _phase_colorbar(fig, image, ax):
    cbar = fig.colorbar(image, ax=ax)
    cbar.set_ticks(PHASE_TICKS)
    cbar.set_ticklabels(PHASE_LABELS)
    return cbar


def _label(axs, xlabel, ylabel):
    for ax in axs:
        ax.set_xlabel(xlabel)
        ax.set_ylabel(ylabel)


def _finish(fig, paths, stem, save, ext, show):
    plt = _pyplot()
    plt.tight_layout()
    if save:
        fig.savefig(next_available_path(paths['data'] + stem, paths['folder'], ext))
    if show:
        plt.show()


def plot_spins(psi, psik, extents, paths, cmap='viridis', save=True, ext='.pdf',
               show=True, zoom=1.0):
    """Real-space densities (top row), phases (middle) and k-space densities
    (bottom) of both components, with colour bars.

    ``extents``: ``{'r': [x_min, x_max, y_min, y_max], 'k': [...]}``;
    ``paths``: ``{'data', 'folder'}`` as set by ``PSpinor``; ``save`` writes
    ``<data>spin_dens_phase<i>-<folder><ext>``; ``zoom`` narrows the k-space
    axes to ``extents['k'] / zoom``. Returns ``(fig, {'r': [..], 'ph': [..],
    'k': [..]})``, one image per component in each list.
    """
    from matplotlib import gridspec  # pylint: disable=import-outside-toplevel
    plt = _pyplot()
    dens = ttools.density(psi)
    phases = ttools.phase(psi, uwrap=False, dens=dens)
    densk = ttools.density(psik)
    k_ext = np.asarray(extents['k'], dtype=float)

    fig = plt.figure(figsize=(5.5, 6.4))
    grid = gridspec.GridSpec(6, 4)
    rows = [[fig.add_subplot(grid[2 * r:2 * r + 2, 2 * c:2 * c + 2]) for c in range(2)]
            for r in range(3)]
    plots = {'r': [], 'ph': [], 'k': []}
    for c in range(2):
        ax_r, ax_ph, ax_k = rows[0][c], rows[1][c], rows[2][c]
        plots['r'].append(ax_r.imshow(dens[c], cmap=cmap, origin='lower', extent=extents['r'],
                                      vmin=0, aspect='equal'))
        fig.colorbar(plots['r'][c], ax=ax_r)
        plots['ph'].append(ax_ph.imshow(phases[c], cmap='twilight_shifted', origin='lower',
                                        extent=extents['r'], vmin=-np.pi, vmax=np.pi,
                                        aspect='equal'))
        _phase_colorbar(fig, plots['ph'][c], ax_ph)
        plots['k'].append(ax_k.imshow(densk[c], cmap=cmap, origin='lower', extent=k_ext,
                                      vmin=0, aspect='equal'))
        fig.colorbar(plots['k'][c], ax=ax_k)
        ax_k.set_xlim(k_ext[:2] / zoom)
        ax_k.set_ylim(k_ext[2:] / zoom)
    _label(rows[0] + rows[1], '$x$', '$y$')
    _label(rows[2], '$k_x$', '$k_y$')
    _finish(fig, paths, 'spin_dens_phase', save, ext, show)
    return fig, plots


def plot_total(psi, psik, extents, paths, cmap='viridis', save=True, ext='.pdf',
               show=True, zoom=1.0):
    """Total real-space density and the phase of ``psi_0 + psi_1`` (top),
    total k-space density (bottom, centred). Arguments as in
    :func:`plot_spins`; saves ``<data>total_dens_phase<i>-<folder><ext>``.
    Returns ``(fig, {'r': image, 'ph': image, 'k': image})``."""
    from matplotlib import gridspec  # pylint: disable=import-outside-toplevel
    plt = _pyplot()
    dens_r = sum(ttools.density(psi))
    phase_r = ttools.phase(psi[0] + psi[1], uwrap=False, dens=dens_r)
    dens_k = sum(ttools.density(psik))
    k_ext = np.asarray(extents['k'], dtype=float)

    fig = plt.figure()
    grid = gridspec.GridSpec(4, 4)
    ax_r = fig.add_subplot(grid[0:2, 0:2])
    ax_ph = fig.add_subplot(grid[0:2, 2:4])
    ax_k = fig.add_subplot(grid[2:4, 1:3])
    plots = {'r': ax_r.imshow(dens_r, cmap=cmap, origin='lower', extent=extents['r'], vmin=0),
             'ph': ax_ph.imshow(phase_r, cmap='twilight_shifted', origin='lower',
                                extent=extents['r'], vmin=-np.pi, vmax=np.pi),
             'k': ax_k.imshow(dens_k, cmap=cmap, origin='lower', extent=k_ext, vmin=0)}
    fig.colorbar(plots['r'], ax=ax_r)
    _phase_colorbar(fig, plots['ph'], ax_ph)
    fig.colorbar(plots['k'], ax=ax_k)
    _label([ax_r, ax_ph], '$x$', '$y$')
    _label([ax_k], '$k_x$', '$k_y$')
    ax_k.set_xlim(k_ext[:2] / zoom)
    ax_k.set_ylim(k_ext[2:] / zoom)
    _finish(fig, paths, 'total_dens_phase', save, ext, show)
    return fig, plots
