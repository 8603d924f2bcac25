"""3-D voxel bioheat stencil on the MI355X.

Drop-in for ``simulation/voxel_thermal_3d.py`` of the reference: the module
kernels ``_thermal3d_step_explicit`` (:30-112) and ``_apply_latent_heat_plateau``
(:115-143) keep their signatures and in-place semantics on numpy arrays, and
``VoxelThermalSolver3D`` (:146-309) keeps its constructor and ``solve``.  All
arithmetic runs in the fused FP32 sub-step kernel of libbugdef.so.
"""

from __future__ import annotations

import time

import numpy as np

from . import _bdlib as B

CFL_SAFETY_FACTOR = 0.40


def _physics(dt=1e-5, h_eff=0.0, t_ambient=20.0, t_abl=1e30, t_void=1e30, **kw) -> B.Physics:
    p = B.Physics()
    p.dt_sub = float(dt)
    p.h_eff = float(h_eff)
    p.t_ambient = float(t_ambient)
    p.t_abl = float(t_abl)
    p.t_void = float(t_void)
    p.air_gap_k0 = -1
    p.air_gap_k1 = -1
    p.alpha_max_stable = 1.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _thermal3d_step_explicit(T, T_new, Q, k_grid, rho_cp_grid, Nx, Ny, Nz, dx, dy, dz, dt, h_eff, t_ambient,
                             is_active):
    """One explicit step (voxel_thermal_3d.py:30-112): writes T_new in place."""
    ctx = B.scratch_context((Nx, Ny, Nz), (dx, dy, dz))
    ctx.set_physics(_physics(dt=dt, h_eff=h_eff, t_ambient=t_ambient))
    ctx.upload('T', T)
    ctx.upload('Q', Q)
    ctx.upload('K', k_grid)
    ctx.upload('RHO_CP', rho_cp_grid)
    ctx.substeps(1, 1 if is_active else 0, 0)
    T_new[...] = ctx.download('T', dtype=np.float64 if T_new.dtype == np.float64 else np.float32)


def _apply_latent_heat_plateau(T, H_acc, rho_cp_grid, ablated, t_abl, Nx, Ny, Nz, dry, t_void):
    """Enthalpy plateau (voxel_thermal_3d.py:115-143): updates T and H_acc in place."""
    ctx = B.scratch_context((Nx, Ny, Nz), (1.0, 1.0, 1.0))
    ctx.set_physics(_physics(t_abl=t_abl, t_void=t_void))
    ctx.upload('T', T)
    ctx.upload('RHO_CP', rho_cp_grid)
    ctx.upload('ABLATED', ablated)
    ctx.upload('DRY', dry)
    ctx.upload('H_ACC', H_acc)
    ctx.plateau()
    T[...] = ctx.download('T')
    H_acc[...] = ctx.download('H_ACC')


class VoxelThermalSolver3D:
    """voxel_thermal_3d.py:146-309 with the explicit steps on the device."""

    def __init__(self, grid_x, grid_y, grid_z, k_grid, rho_cp_grid, h_conv=15.0, emissivity=0.95,
                 t_ambient=20.0):
        self.grid_x, self.grid_y, self.grid_z = grid_x, grid_y, grid_z
        self.Nx, self.Ny, self.Nz = len(grid_x), len(grid_y), len(grid_z)
        self.dx = float(grid_x[1] - grid_x[0]) if self.Nx > 1 else 0.05
        self.dy = float(grid_y[1] - grid_y[0]) if self.Ny > 1 else 0.05
        self.dz = float(grid_z[1] - grid_z[0]) if self.Nz > 1 else 0.04
        # unit heuristics (voxel_thermal_3d.py:171-180)
        self.k_grid = k_grid * 1e-3 if np.max(k_grid) > 0.05 else np.array(k_grid, dtype=np.float64, copy=True)
        self.rho_cp_grid = (rho_cp_grid * 1e-9 if np.max(rho_cp_grid) > 100.0
                            else np.array(rho_cp_grid, dtype=np.float64, copy=True))
        sigma_sb = 5.670374e-8
        self.h_eff = (h_conv + 4.0 * sigma_sb * emissivity * (t_ambient + 273.15) ** 3) * 1e-6
        self.t_ambient = float(t_ambient)

    def cfl_conditioned_rho_cp(self, dt):
        """voxel_thermal_3d.py:191-203."""
        inv_d2 = 1.0 / self.dx ** 2 + 1.0 / self.dy ** 2 + 1.0 / self.dz ** 2
        amax = CFL_SAFETY_FACTOR / (dt * inv_d2)
        rc = self.rho_cp_grid.copy()
        unstable = self.k_grid / np.maximum(rc, 1e-12) > amax
        if np.any(unstable):
            rc[unstable] = self.k_grid[unstable] / amax
        return rc

    def solve(self, Q_3d, pulse_duration=0.20, total_time=1.00, dt=0.001, snapshot_times=None, engine='auto'):
        """voxel_thermal_3d.py:205-309: snapshots and centre/surface histories."""
        from .voxel_mc_3d import check_engine
        check_engine(engine)
        n_steps = int(round(total_time / dt))
        ctx = B.Context((self.Nx, self.Ny, self.Nz), (self.dx, self.dy, self.dz))
        ctx.set_physics(_physics(dt=dt, h_eff=self.h_eff, t_ambient=self.t_ambient))
        ctx.upload('K', self.k_grid)
        ctx.upload('RHO_CP', self.cfl_conditioned_rho_cp(dt))
        ctx.upload('Q', Q_3d)
        ctx.fill('T', self.t_ambient)
        if snapshot_times is None:
            snapshot_times = [0.0, pulse_duration, total_time]
        targets = sorted({round(float(t), 4) for t in snapshot_times})
        mx, my, mz = self.Nx // 2, self.Ny // 2, self.Nz // 2
        centre, surface, times, snaps = [], [], [], {}
        t0 = time.time()
        for step in range(n_steps + 1):
            cur = round(step * dt, 4)
            plane = ctx.slice('T', 1, my)  # x-z plane through the centre
            centre.append(float(plane[mx, mz]))
            surface.append(float(plane[mx, 0]))
            times.append(cur)
            if cur in targets:
                snaps[cur] = ctx.download('T')
            if step < n_steps:
                ctx.substeps(1, 1 if cur < pulse_duration else 0, 0)
        solve_time = time.time() - t0
        ctx.close()
        return {'snapshots': snaps, 'time_history': np.array(times), 'center_history': np.array(centre),
                'surface_history': np.array(surface), 'solve_time': round(solve_time, 3),
                'peak_surface_T': max(surface), 'peak_center_T': max(centre), 'engine': 'gpu'}
