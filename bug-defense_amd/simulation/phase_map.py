"""The four-phase tissue map and the char / latent-heat constants of the beetle runners.

``compute_phase_map`` is ``run_multispectral_matrix.py:168-192`` evaluated on the device
(``bd_phase_map``); ``H_LAT_WATER`` and ``CHAR_STATE`` are the constants the runners pass to the
engine (run_multispectral_matrix.py:54, 69-78).  The matrix runner itself is a caller of the
engine, not part of it: with the reference checkout on the package path (see
``simulation/__init__.py``) ``simulation.run_multispectral_matrix`` loads from there unchanged.
"""

import numpy as np

from . import _bdlib as B

T_AMBIENT = 20.0
T_BOIL = 100.0
T_ABL_CUTICLE = 350.0
H_LAT_WATER = 2.26
H_EFF = 15.0e-6

CHAR_STATE = {
    'A': 1.0e7,
    'Ea': 65000.0,
    'mu_a': 15.0,
    'mu_s': 20.0,
    'k': 0.35e-3,
    'rho_cp': 2.0e-3,
    't_remove': T_ABL_CUTICLE,
    'chi_remove': 0.95,
}


def compute_phase_map(T, chi, dry, ablated, H_acc, is_air_gap=None):
    """run_multispectral_matrix.py:168-192 evaluated on the device.

    0 intact, 1 char (dry and chi >= 0.10), 2 boiling (T >= 99.5 or H_acc > 0),
    3 void (ablated or air gap).  Inputs are host arrays; for an engine's own
    state use ``CoupledAblationEngine.phase_map`` (no host round trip).
    """
    T = np.asarray(T)
    ctx = B.scratch_context(T.shape, (1.0, 1.0, 1.0))
    ctx.upload('T', T)
    ctx.upload('CHI', chi)
    ctx.upload('H_ACC', H_acc)
    ctx.upload('DRY', dry)
    ctx.upload('ABLATED', ablated)
    return ctx.phase_map(is_air_gap)
