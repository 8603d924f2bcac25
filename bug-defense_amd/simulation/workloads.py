"""Synthetic inputs of the benchmark configurations (no datasets exist offline).

``layered_stack`` builds the coupled layered-stack run: a seeded planar stack
under ambient air, voxelised like the reference's beetle builders
(run_highres_knowledge_suite.py:126-185: initial ablated ambient air above the
dorsal surface, a subelytral air gap with ``air_gap_range``, dry cuticle layers
that char and ablate at 350 degC, wet layers that boil at 100 degC), and returns
the ``CoupledAblationEngine`` keyword arguments.
"""

from __future__ import annotations

import numpy as np

from .coupled_ablation import K_AIR, RHO_CP_AIR
from .phase_map import CHAR_STATE, H_LAT_WATER

# thicknesses [um] of the 7 planar layers; index 1 is the air gap
STACK_UM = (20.0, 300.0, 80.0, 150.0, 200.0, 400.0, 60.0)
AIR_LAYER = 1


def stack_engine_kwargs(shape, voxel_mm, n_air_top, layers, power, theta_deg, spot_dia, n_photons, dt_epoch,
                        dt_sub, plume_grid, seed, wavelength_nm=532.0):
    """Engine keywords for a planar stack under ambient air.

    layers: per entry of STACK_UM, None for the air gap or a tuple
    (mu_a [1/mm], mu_s [1/mm], g, n, k [W/(mm K)], rho_cp [J/(mm^3 K)], water fraction);
    layers with water fraction < 0.2 are dry (h_lat 0, char, removal at 350 degC),
    the others boil with h_lat = 2.26 w J/mm^3.  The deepest layer continues to the bottom.
    """
    Nx, Ny, Nz = shape
    mu_a = np.zeros(shape)
    mu_s = np.zeros(shape)
    g = np.zeros(shape)
    n = np.ones(shape)
    k = np.full(shape, K_AIR)
    rc = np.full(shape, RHO_CP_AIR)
    hl = np.full(shape, 1e12)
    dry = np.zeros(shape, bool)
    abl = np.zeros(shape, bool)
    abl[:, :, :n_air_top] = True
    z = n_air_top
    gap = None
    for i, um in enumerate(STACK_UM):
        nvox = max(1, int(round(um / (voxel_mm * 1000.0))))
        z1 = Nz if i == len(STACK_UM) - 1 else min(Nz, z + nvox)
        lay = layers[i]
        if lay is None:
            gap = (z, z1)
        else:
            ma, ms, gg, nn, kk, rcc, w = lay
            mu_a[:, :, z:z1], mu_s[:, :, z:z1], g[:, :, z:z1], n[:, :, z:z1] = ma, ms, gg, nn
            k[:, :, z:z1] = kk
            rc[:, :, z:z1] = rcc
            is_dry = w < 0.2
            hl[:, :, z:z1] = 0.0 if is_dry else H_LAT_WATER * w
            dry[:, :, z:z1] = is_dry
        z = z1
    optical = {'mu_a': mu_a, 'mu_s': mu_s, 'g': g, 'n': n}
    dry_state = {'mu_a': mu_a.copy(), 'mu_s': mu_s.copy(), 'g': g.copy(), 'n': n.copy(), 'k': k.copy(),
                 'rho_cp': rc.copy(), 't_remove': 350.0}
    L = (Nx * voxel_mm, Ny * voxel_mm, Nz * voxel_mm)
    return dict(
        optical_props=optical, k_grid=k, rho_cp_grid=rc, h_lat_grid=hl, Lx=L[0], Ly=L[1], Lz=L[2], power=power,
        spot_dia=spot_dia, t_abl=100.0, t_ambient=20.0, dt_epoch=dt_epoch, dt_sub=dt_sub, n_photons=n_photons,
        h_eff=15.0e-6, beam_center=(L[0] / 2.0, L[1] / 2.0), theta_inc_deg=theta_deg, phi_inc_deg=0.0,
        dry_state=dry_state, char_state=CHAR_STATE,
        plume_closure={'model': 'diffusion', 'D_xy': 1.0, 'D_z': 2.0, 'v_drift': 0.0, 'tau_linger_s': float('inf'),
                       'blast_expansion': False, 'blast_sigma_xy_mm': 0.080, 'blast_sigma_z_mm': 0.090,
                       'blast_z_bias_mm': 0.0},
        initial_dry_mask=dry, initial_ablated_mask=abl, air_gap_range=gap, wavelength_nm=wavelength_nm,
        plume_grid=plume_grid, seed=seed + 1)


def layered_stack(shape=(256, 256, 192), voxel_mm=0.015, n_air_top=21, seed=0, power=40.0, theta_deg=22.0,
                  spot_dia=0.5, n_photons=150_000, dt_epoch=1e-3, dt_sub=5e-5, plume_grid=(128, 128, 128)):
    """Engine keywords for the coupled layered-stack configuration.

    Per layer, from ``np.random.default_rng(seed)``: mu_a U[0.01, 20] /mm,
    mu_s U[1, 40] /mm, g U[0.7, 0.95], n U[1.0, 1.55], k U[0.1, 0.6] W/(m K),
    water fraction w U[0, 0.8]; rho_cp = w 4.184e-3 + (1 - w) 2.1e-3 J/(mm^3 K);
    the outermost layer's w is scaled by 0.1 (a cuticle); layers with w < 0.2
    are dry (h_lat 0, char, removal at 350 degC), the others boil with
    h_lat = 2.26 w J/mm^3.  The deepest layer continues to the bottom.
    """
    rng = np.random.default_rng(seed)
    layers = []
    for i in range(len(STACK_UM)):
        ma, ms, gg, nn, kk, w = (rng.uniform(0.01, 20.0), rng.uniform(1.0, 40.0), rng.uniform(0.7, 0.95),
                                 rng.uniform(1.0, 1.55), rng.uniform(0.1, 0.6) * 1e-3, rng.uniform(0.0, 0.8))
        if i == AIR_LAYER:
            layers.append(None)
            continue
        if i == 0:
            w *= 0.1  # the outermost layer is cuticle: dry, it chars and ablates at 350 degC
        layers.append((ma, ms, gg, nn, kk, w * 4.184e-3 + (1.0 - w) * 2.1e-3, w))
    return stack_engine_kwargs(shape, voxel_mm, n_air_top, layers, power, theta_deg, spot_dia, n_photons, dt_epoch,
                               dt_sub, plume_grid, seed)
