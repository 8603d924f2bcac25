"""MI355X (gfx950) implementation of the coupled laser-tissue epoch loop.

Same module layout and entry points as the reference's ``simulation`` package for the hot
path: ``voxel_mc_3d``, ``voxel_thermal_3d``, ``coupled_ablation``, ``plume``, the GPU-context
modules ``metal_ablation_solver``, ``voxel_thermal_metal`` and ``voxel_mc_metal``, and the
layer optics of ``tissue_stack``; compute runs in ``lib/libbugdef.so``.

Callers of the reference that live in modules this package does not restate (``beetle_sim``,
``run_multispectral_matrix``) keep working: when the environment variable
``BUGDEF_REFERENCE_SIMULATION`` names the reference checkout's ``simulation`` directory (several
directories may be joined with ``os.pathsep``), it is appended to this package's search path.
This package's modules come first, so ``beetle_sim``'s own ``from simulation.coupled_ablation
import CoupledAblationEngine`` resolves to the MI355X engine, while ``simulation.beetle_sim``
and ``simulation.run_multispectral_matrix`` themselves load from the reference unchanged.
"""

import os as _os

for _p in _os.environ.get("BUGDEF_REFERENCE_SIMULATION", "").split(_os.pathsep):
    if _p and _os.path.isdir(_p) and _os.path.abspath(_p) not in [_os.path.abspath(q) for q in __path__]:
        __path__.append(_p)
del _os
