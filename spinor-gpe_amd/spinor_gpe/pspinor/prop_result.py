"""Result of a propagation (spinor_gpe/pspinor/prop_result.py:19-88).

Same constructor and fields as the reference. Engine-specific extras are
added as attributes by :class:`TensorPropagator`: ``energy_gp`` (the spectral
GP energy terms), ``chem_pot`` and ``steps_run``. The reference's plotting and
movie helpers are not provided.
"""

import numpy as np

from spinor_gpe.pspinor import tensor_tools as ttools


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
