"""The GPU transport entry point of ``simulation/voxel_mc_metal.py``, on the MI355X.

``run_voxel_mc_metal`` (:388-586) keeps its name, keywords and output dictionary so callers of
the reference run unchanged; it is the gfx950 transport kernel of libbugdef.so, the same one
``run_voxel_mc3d`` runs, with the CPU semantics of the reference's MC (voxel_mc_3d.py:33-337).
``base_seed`` keys the Philox streams (random when None), ``max_steps`` is the per-photon event
budget, and ``lateral_escape_z`` holds the lateral loss per depth (the Metal path returned zeros).
``is_metal_available()`` keeps its name and reports whether this build's GPU engine (a HIP
device and the library) is usable.
"""

from __future__ import annotations

from . import _bdlib as B
from .voxel_mc_3d import run_voxel_mc3d


def is_metal_available() -> bool:
    """True when the GPU engine can run (voxel_mc_metal.py:23-40 contract)."""
    return B.is_gpu_available()


def run_voxel_mc_metal(power=1.0, pulse_duration=0.20, n_photons=100_000, Lx=2.0, Ly=2.0, Lz=1.0, Nx=40, Ny=40,
                       Nz=50, custom_3d_props=None, layers_config=None, theta_inc_deg=0.0, phi_inc_deg=0.0,
                       spot_dia=0.50, beam_center=None, n_ambient=1.0, r_max=0.0, base_seed=None, max_steps=50_000):
    """voxel_mc_metal.py:388-586 on the gfx950 transport kernel (Gaussian beam along +z)."""
    out = run_voxel_mc3d(power=power, pulse_duration=pulse_duration, n_photons=n_photons, spot_dia=spot_dia,
                         theta_inc_deg=theta_inc_deg, phi_inc_deg=phi_inc_deg, Lx=Lx, Ly=Ly, Lz=Lz, Nx=Nx, Ny=Ny,
                         Nz=Nz, r_max=r_max, layers_config=layers_config, custom_3d_props=custom_3d_props,
                         launch_axis='z', beam_center=beam_center, beam_profile='gaussian', engine='gpu',
                         seed=base_seed, n_ambient=n_ambient, max_events=max_steps)
    out['gpu_device_name'] = 'AMD Instinct MI355X (gfx950)'
    return out
