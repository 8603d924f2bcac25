"""The fused ablation context of ``simulation/metal_ablation_solver.py``, on the MI355X.

The reference's GPU context (``MetalAblationContext``, :285-455) is what its tests and the
engine drive for the fused sub-step: stencil, enthalpy plateau, Arrhenius damage, char kinetics
and void-confined plume diffusion.  This module keeps the class name, constructor, methods and
array conventions, so callers written against it run unchanged; the work runs in the fused
gfx950 sub-step kernel of libbugdef.so (``bd_substeps``), with the state device-resident
between ``set_state`` and ``get_state``.

Semantics: the stencil, plateau, Arrhenius and char steps are the CPU kernels' (the reference's
oracle, voxel_thermal_3d.py:30-143, coupled_ablation.py:33-58), as everywhere in this build; the
plume takes the Metal solver's diffusion step ``min(dt, 0.45 / (2 d_rate))`` (:234-236), which is
what this context is for.  ``is_metal_available()`` keeps its name and reports whether this
build's GPU engine (a HIP device and the library) is usable.
"""

from __future__ import annotations

import numpy as np

from . import _bdlib as B


def is_metal_available() -> bool:
    """True when the GPU engine can run (metal_ablation_solver.py:26-41 contract)."""
    return B.is_gpu_available()


def _plume_params(plume_params, shape, dt):
    """metal_ablation_solver.py:328-342: D / D_plume fallback, D_xy, D_z, vz (default 15 mm/s),
    tau_linger (decay 1 for tau <= 0 or inf)."""
    pp = B.PlumeParams()
    if plume_params is None:
        pp.enabled = 0
        return pp
    d_fallback = float(plume_params.get('D', plume_params.get('D_plume', 0.5)))
    pp.enabled = 1
    pp.model = 0
    pp.nx, pp.ny, pp.nz = shape
    pp.D_xy = float(plume_params.get('D_xy', d_fallback))
    pp.D_z = float(plume_params.get('D_z', d_fallback))
    pp.v_drift = float(plume_params.get('vz', plume_params.get('v_drift', 15.0)))
    tau = float(plume_params.get('tau_linger', 0.0))
    pp.tau_linger = tau if tau > 0.0 else float('inf')
    pp.mu_max = pp.mu_a_max = 1.0
    pp.mu_s_max = 0.0
    pp.delta_mu = 0.0
    pp.blast_expansion = 0
    pp.dt_clamp = 1
    return pp


class MetalAblationContext:
    """metal_ablation_solver.py:285-455: persistent fused multi-phase context."""

    def __init__(self, Nx: int, Ny: int, Nz: int, dx: float, dy: float, dz: float, dt_sub: float, h_eff: float,
                 t_ambient: float, t_abl: float, t_void: float, arrhenius=None, char_params=None,
                 plume_params=None):
        if not is_metal_available():
            raise RuntimeError("GPU engine is not available (no HIP device or libbugdef.so)")
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.dt = float(dt_sub)
        self.h_eff = float(h_eff)
        self.t_ambient = float(t_ambient)
        self.t_abl = float(t_abl)
        self.t_void = float(t_void)
        self.has_arrhenius = 1 if arrhenius is not None else 0
        self.has_char = 1 if char_params is not None else 0
        self.has_plume = 1 if plume_params is not None else 0
        shape = (self.Nx, self.Ny, self.Nz)
        self.n_voxels = self.Nx * self.Ny * self.Nz
        self._ctx = B.Context(shape, (self.dx, self.dy, self.dz))
        p = B.Physics()
        p.dt_sub, p.h_eff, p.t_ambient, p.t_abl, p.t_void = self.dt, self.h_eff, self.t_ambient, self.t_abl, self.t_void
        if arrhenius is not None:
            p.has_arrhenius, p.arr_A, p.arr_Ea_over_R = 1, float(arrhenius[0]), float(arrhenius[1])
        if char_params is not None:
            p.has_char, p.char_A, p.char_Ea_over_R = 1, float(char_params[0]), float(char_params[1])
        p.alpha_max_stable = 1.0
        p.air_gap_k0 = p.air_gap_k1 = -1
        self._ctx.set_physics(p)
        self._ctx.set_plume(_plume_params(plume_params, shape, self.dt))
        # the fused step of the reference always applies the plateau; Q is used as given
        self._flags = B.SUB_PLATEAU
        if self.has_arrhenius:
            self._flags |= B.SUB_ARRHENIUS
        if self.has_char:
            self._flags |= B.SUB_CHAR
        if self.has_plume:
            self._flags |= B.SUB_PLUME

    def set_state(self, T, H_acc, Omega, chi, dry, ablated, Q, k_grid, rho_cp_grid, plume=None, plume_void=None):
        """Upload the full state (metal_ablation_solver.py:381-397); the plume void defaults to
        the ablated mask."""
        c = self._ctx
        for name, a in (('T', T), ('Q', Q), ('K', k_grid), ('RHO_CP', rho_cp_grid), ('CHI', chi)):
            c.upload(name, np.asarray(a, dtype=np.float32))
        c.upload('OMEGA', np.asarray(Omega, dtype=np.float32))
        c.upload('DRY', dry)
        c.upload('ABLATED', ablated)
        c.upload('H_ACC', np.asarray(H_acc, dtype=np.float32))
        if self.has_plume:
            if plume is not None:
                c.upload('PLUME_RAW', np.asarray(plume, dtype=np.float32))
            c.upload('VOID', plume_void if plume_void is not None else ablated)

    def advance_substeps(self, n_substeps: int, is_active: int = 1):
        """n fused sub-steps (metal_ablation_solver.py:399-437), one launch per sub-step (two per
        launch where the tuned kernel pairs them), no host round trip."""
        if n_substeps <= 0:
            return
        self._ctx.substeps(int(n_substeps), 1 if is_active else 0, self._flags)

    def get_state(self):
        """(T, H_acc, Omega, chi) as float64 (metal_ablation_solver.py:439-446)."""
        c = self._ctx
        return (c.download('T'), c.download('H_ACC'), c.download('OMEGA'), c.download('CHI'))

    def get_plume(self, out_arr=None):
        """The plume density field, float32 (metal_ablation_solver.py:448-455)."""
        if not self.has_plume:
            raise RuntimeError("get_plume: the context was created without plume_params")
        if out_arr is not None:
            out_arr[...] = self._ctx.download('PLUME_RAW', dtype=np.float32)
            return out_arr
        return self._ctx.download('PLUME_RAW', dtype=np.float32)

    def close(self):
        self._ctx.close()
