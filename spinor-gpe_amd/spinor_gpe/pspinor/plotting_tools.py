"""File-naming helper shared with the reference's plotting module.

Only ``next_available_path`` (spinor_gpe/pspinor/plotting_tools.py:13-37) is
provided: sampled-state files keep the reference's naming,
``<file_name><i>-<trial_name><ext>`` with the first unused ``i >= 1``.
Figures and movies are outside this engine.
"""

import os


def next_available_path(file_name, trial_name, ext=''):
    """Return the first ``file_name + str(i) + '-' + trial_name + ext`` that does not exist."""
    i = 1
    while True:
        path = f"{file_name}{i}-{trial_name}{ext}"
        if not os.path.exists(path):
            return path
        i += 1
