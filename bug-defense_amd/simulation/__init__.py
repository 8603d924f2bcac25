"""MI355X (gfx950) implementation of the coupled laser-tissue epoch loop.

Same module layout and entry points as the reference's ``simulation``
package for the hot path: ``voxel_mc_3d``, ``voxel_thermal_3d``,
``coupled_ablation``, ``plume`` and the phase map of
``run_multispectral_matrix``; compute runs in ``lib/libbugdef.so``.
"""
