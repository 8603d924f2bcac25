"""The GPU thermal stencil API of ``simulation/voxel_thermal_metal.py``, on the MI355X.

``run_thermal3d_substeps_metal`` (:179-285) and ``MetalThermalContext`` (:288-383) keep their
names, signatures and return conventions, so the reference's callers and tests run unchanged;
the sub-steps run in the fused gfx950 sub-step kernel of libbugdef.so with only the stencil
enabled, and the fields stay on the device between calls of a context.  The stencil is the CPU
kernel's (voxel_thermal_3d.py:30-112: series-resistance top/bottom faces), the oracle every
GPU path of this build follows.  ``is_metal_available()`` keeps its name and reports whether
this build's GPU engine (a HIP device and the library) is usable.
"""

from __future__ import annotations

import numpy as np

from . import _bdlib as B
from .voxel_thermal_3d import _physics


def is_metal_available() -> bool:
    """True when the GPU engine can run (voxel_thermal_metal.py:24-40 contract)."""
    return B.is_gpu_available()


def run_thermal3d_substeps_metal(T, Q, k_grid, rho_cp_grid, Nx, Ny, Nz, dx, dy, dz, dt, n_substeps, h_eff=0.0,
                                 t_ambient=20.0, is_active=1):
    """n_substeps explicit steps from T (voxel_thermal_metal.py:179-285); returns float64."""
    ctx = B.scratch_context((Nx, Ny, Nz), (dx, dy, dz))
    ctx.set_physics(_physics(dt=dt, h_eff=h_eff, t_ambient=t_ambient))
    ctx.upload('T', np.asarray(T, dtype=np.float32))
    ctx.upload('Q', np.asarray(Q, dtype=np.float32))
    ctx.upload('K', np.asarray(k_grid, dtype=np.float32))
    ctx.upload('RHO_CP', np.asarray(rho_cp_grid, dtype=np.float32))
    ctx.substeps(int(n_substeps), 1 if is_active else 0, 0)
    return ctx.download('T')


class MetalThermalContext:
    """voxel_thermal_metal.py:288-383: a persistent thermal context (k and rho_cp fixed)."""

    def __init__(self, Nx, Ny, Nz, dx, dy, dz, dt, k_grid, rho_cp_grid, h_eff=0.0, t_ambient=20.0):
        if not is_metal_available():
            raise RuntimeError("GPU engine is not available (no HIP device or libbugdef.so)")
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.dt = float(dt)
        self.h_eff = float(h_eff)
        self.t_ambient = float(t_ambient)
        self.n_voxels = self.Nx * self.Ny * self.Nz
        self._ctx = B.Context((self.Nx, self.Ny, self.Nz), (self.dx, self.dy, self.dz))
        self._ctx.set_physics(_physics(dt=self.dt, h_eff=self.h_eff, t_ambient=self.t_ambient))
        self._ctx.upload('K', np.asarray(k_grid, dtype=np.float32))
        self._ctx.upload('RHO_CP', np.asarray(rho_cp_grid, dtype=np.float32))
        self._ctx.fill('T', self.t_ambient)
        self._ctx.fill('Q', 0.0)

    def set_T(self, T):
        self._ctx.upload('T', np.asarray(T, dtype=np.float32))

    def set_Q(self, Q):
        self._ctx.upload('Q', np.asarray(Q, dtype=np.float32))

    def advance(self, n_substeps: int, is_active: int = 1):
        if n_substeps <= 0:
            return
        self._ctx.substeps(int(n_substeps), 1 if is_active else 0, 0)

    def get_T(self):
        return self._ctx.download('T')

    def get_T_view(self):
        """A float32 copy of T (the reference returns a view of unified memory; device memory
        here is not host-visible, so this is a fresh array)."""
        return self._ctx.download('T', dtype=np.float32)

    def close(self):
        self._ctx.close()
