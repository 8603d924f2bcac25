"""Coupled voxel MC + bioheat + enthalpy ablation engine on the MI355X.

Drop-in for ``simulation/coupled_ablation.py`` of the reference:
``CoupledAblationEngine`` keeps the constructor keywords (:184-217), the
``step_epoch`` / ``compute_void_mask`` / ``centerline_depth_mm`` methods and
the state attributes its callers read (``T, Q, chi, dry, ablated, H_acc,
Omega, T_peak, optical, plume, last_mc, t, epoch``) and write (``power,
beam_center, optical[...] in place, dry_state[...], _wet, _ds_arrays,
dt_epoch, n_substeps``).  The module kernels ``_accumulate_arrhenius``,
``_accumulate_char``, ``_update_single_stage_ablation_numba`` and
``_update_phase_properties_numba`` keep their signatures.

All state lives on the device for the whole run; one ``step_epoch`` is one
``bd_epoch`` call (MC, fused sub-steps, phase update with removal, plume).
State attributes are read-only host copies fetched on first access after
each epoch (the reference returns fresh float64 arrays, so callers that only
read see the same behaviour).  ``optical`` is a writable host copy that is
uploaded back before the next epoch, because callers mutate it in place
(run_dual_permutation_suite.py:261-275, validation_benchmarks.py:119-122).

Extra keywords (not in the reference): ``seed`` (Philox key of the MC
streams; random when None), ``plume_grid`` (shape of a separate plume grid,
default the thermal grid), ``score_fluence`` (path-length tally for ``Phi``;
default True like the reference) and ``graph`` (replay each epoch from a
captured hipGraph when the configuration allows it).
"""

from __future__ import annotations

import time

import numpy as np

from . import _bdlib as B
from .plume import MultispectralPlumeField, PlumeAbsorptionField
from .voxel_mc_3d import check_engine, make_beam, spectro_dict

K_AIR = 0.026e-3       # W/(mm*K)
RHO_CP_AIR = 1.206e-6  # J/(mm^3*K)
R_GAS = 8.314          # J/(mol*K)


def mix_seed(seed: int, epoch: int) -> int:
    """SplitMix64 of (seed + (epoch + 1) * golden gamma): the per-epoch MC key."""
    z = (int(seed) + (int(epoch) + 1) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


# ------------------------------------------------------------------ module kernels
def _mask_ctx(shape):
    ctx = B.scratch_context(shape, (1.0, 1.0, 1.0))
    return ctx


def _accumulate_arrhenius(T, Omega, ablated, A, Ea_over_R, dt, Nx, Ny, Nz):
    """coupled_ablation.py:33-40 on the device (Omega updated in place)."""
    ctx = _mask_ctx((Nx, Ny, Nz))
    ctx.upload('T', T)
    ctx.upload('ABLATED', ablated)
    ctx.upload('OMEGA', Omega)
    ctx.arrhenius(A, Ea_over_R, dt)
    Omega[...] = ctx.download('OMEGA')


def _accumulate_char(T, chi, dry, ablated, A, Ea_over_R, dt, Nx, Ny, Nz):
    """coupled_ablation.py:43-58 on the device (chi updated in place)."""
    ctx = _mask_ctx((Nx, Ny, Nz))
    ctx.upload('T', T)
    ctx.upload('ABLATED', ablated)
    ctx.upload('DRY', dry)
    ctx.upload('CHI', chi)
    ctx.char(A, Ea_over_R, dt)
    chi[...] = ctx.download('CHI')


def _phase_physics(mode, has_cs=False, flags=(False, False, False, False), chi_rem=1.0, alpha=1.0,
                   t_abl=0.0) -> B.Physics:
    from .voxel_thermal_3d import _physics
    return _physics(t_abl=t_abl, phase_mode=mode, has_cs=int(has_cs), has_cs_mu_a=int(flags[0]),
                    has_cs_mu_s=int(flags[1]), has_cs_k=int(flags[2]), has_cs_rho_cp=int(flags[3]),
                    cs_chi_remove=float(chi_rem), alpha_max_stable=float(alpha), no_removal=1,
                    k_void=K_AIR, rho_cp_void=RHO_CP_AIR)


def _update_single_stage_ablation_numba(H_acc, h_lat, T, t_abl, ablated, to_remove):
    """coupled_ablation.py:61-74 on the device (to_remove written in place)."""
    ctx = _mask_ctx(T.shape)
    ctx.set_physics(_phase_physics(0, t_abl=t_abl))
    for name, a in (('T', T), ('H_LAT', h_lat), ('ABLATED', ablated), ('H_ACC', H_acc)):
        ctx.upload(name, a)
    ctx.fill('DRY', 0)
    ctx.phase_update(0.0)
    to_remove[...] = ctx.download('TO_REMOVE').astype(bool)


def _update_phase_properties_numba(H_acc, h_lat, T, chi, dry, ablated, mu_a, mu_s, g, n, k_grid, rho_cp_grid,
                                   wet_mu_a, wet_k, wet_rho_cp, ds_mu_a, ds_mu_s, ds_g, ds_n, ds_k, ds_rho_cp,
                                   ds_t_rem, has_cs, has_cs_mu_a, cs_mu_a, cs_mu_s, cs_k, cs_rho_cp, cs_t_rem,
                                   cs_chi_rem, has_cs_mu_s, has_cs_k, has_cs_rho_cp, alpha_max_stable, to_remove):
    """coupled_ablation.py:77-180 on the device (all outputs written in place)."""
    ctx = _mask_ctx(T.shape)
    ctx.set_physics(_phase_physics(1, has_cs, (has_cs_mu_a, has_cs_mu_s, has_cs_k, has_cs_rho_cp), cs_chi_rem,
                                   alpha_max_stable))
    ups = (('T', T), ('H_LAT', h_lat), ('CHI', chi), ('DRY', dry), ('ABLATED', ablated), ('H_ACC', H_acc),
           ('MU_A', mu_a), ('MU_S', mu_s), ('G', g), ('N', n), ('K', k_grid), ('RHO_CP', rho_cp_grid),
           ('WET_MU_A', wet_mu_a), ('WET_K', wet_k), ('WET_RHO_CP', wet_rho_cp),
           ('DS_MU_A', ds_mu_a), ('DS_MU_S', ds_mu_s), ('DS_G', ds_g), ('DS_N', ds_n), ('DS_K', ds_k),
           ('DS_RHO_CP', ds_rho_cp), ('DS_T_REMOVE', ds_t_rem))
    for name, a in ups:
        ctx.upload(name, a)
    if has_cs:
        for name, flag, a in (('CS_MU_A', has_cs_mu_a, cs_mu_a), ('CS_MU_S', has_cs_mu_s, cs_mu_s),
                              ('CS_K', has_cs_k, cs_k), ('CS_RHO_CP', has_cs_rho_cp, cs_rho_cp),
                              ('CS_T_REMOVE', True, cs_t_rem)):
            if flag:
                ctx.upload(name, a)
    ctx.phase_update(0.0)
    H_acc[...] = ctx.download('H_ACC')
    dry[...] = ctx.download('DRY').astype(bool)
    for name, a in (('MU_A', mu_a), ('MU_S', mu_s), ('G', g), ('N', n), ('K', k_grid), ('RHO_CP', rho_cp_grid)):
        a[...] = ctx.download(name)
    to_remove[...] = ctx.download('TO_REMOVE').astype(bool)


# ------------------------------------------------------------------ engine
class _OpticalDict(dict):
    """Host copy of the optical grids; the engine uploads it back before the next epoch."""


class CoupledAblationEngine:
    """coupled_ablation.py:183-667 with device-resident state (see module docstring)."""

    _STATE = {'T': 'T', 'H_acc': 'H_ACC', 'Omega': 'OMEGA', 'chi': 'CHI', 'T_peak': 'T_PEAK',
              'ablation_time': 'ABLATION_TIME', 'k_grid': 'K', 'rho_cp_grid': 'RHO_CP', 'h_lat': 'H_LAT'}
    _MASKS = {'dry': 'DRY', 'ablated': 'ABLATED'}

    def __init__(self, optical_props, k_grid, rho_cp_grid, h_lat_grid, Lx, Ly, Lz, power, spot_dia, t_abl,
                 t_ambient=20.0, dt_epoch=0.01, dt_sub=5e-5, n_photons=100000, h_eff=0.0, r_max=0.0,
                 beam_center=None, stability_factor=0.35, k_void=K_AIR, rho_cp_void=RHO_CP_AIR, arrhenius=None,
                 beam_profile='gaussian', dry_state=None, char_state=None, plume_closure=None, theta_inc_deg=0.0,
                 phi_inc_deg=0.0, initial_dry_mask=None, initial_ablated_mask=None, air_gap_range=None,
                 wavelength_nm=10600.0, mc_engine='auto', thermal_engine='auto', seed=None, plume_grid=None,
                 score_fluence=True, graph=False):
        check_engine(mc_engine)
        check_engine(thermal_engine)
        self.__dict__['_cache'] = {}
        optical = {key: np.array(optical_props[key], dtype=np.float64, copy=True) for key in ('mu_a', 'mu_s', 'g', 'n')}
        self.Nx, self.Ny, self.Nz = optical['mu_a'].shape
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.dx, self.dy, self.dz = self.Lx / self.Nx, self.Ly / self.Ny, self.Lz / self.Nz
        self.voxel_vol = self.dx * self.dy * self.dz
        self.power = float(power)
        self.spot_dia = float(spot_dia)
        self.t_abl = float(t_abl)
        self.t_ambient = float(t_ambient)
        self.dt_epoch = float(dt_epoch)
        self.dt_sub = float(dt_sub)
        self.n_substeps = max(1, int(round(self.dt_epoch / self.dt_sub)))
        self.n_photons = int(n_photons)
        self.h_eff = float(h_eff)
        self.r_max = float(r_max)
        self.beam_center = beam_center
        self.beam_profile = beam_profile
        self.theta_inc_deg = float(theta_inc_deg)
        self.phi_inc_deg = float(phi_inc_deg)
        self.mc_engine = mc_engine
        self.thermal_engine = str(thermal_engine).strip().lower()
        self.use_gpu_thermal = True
        self.seed = int(seed) if seed is not None else int(np.random.randint(1, 2 ** 62))
        self.score_fluence = bool(score_fluence)
        self.graph = bool(graph)
        shape = (self.Nx, self.Ny, self.Nz)
        self._ctx = B.Context(shape, (self.dx, self.dy, self.dz))
        ctx = self._ctx

        # one-time host preparation, exactly as coupled_ablation.py:257-311
        k_h = np.array(k_grid, dtype=np.float64, copy=True)
        rc_h = np.array(rho_cp_grid, dtype=np.float64, copy=True)
        inv_d2 = 1.0 / self.dx ** 2 + 1.0 / self.dy ** 2 + 1.0 / self.dz ** 2
        self.alpha_max_stable = stability_factor / (self.dt_sub * inv_d2)
        unstable = k_h / np.maximum(rc_h, 1e-12) > self.alpha_max_stable
        self.cfl_conditioned_mask = unstable.copy()
        if np.any(unstable):
            rc_h[unstable] = k_h[unstable] / self.alpha_max_stable
        self.k_void = float(k_void)
        self.rho_cp_void = max(float(rho_cp_void), self.k_void / self.alpha_max_stable)
        T_h = np.full(shape, self.t_ambient)
        abl_time = np.full(shape, -1.0)
        if initial_ablated_mask is not None:
            ablated = np.array(initial_ablated_mask, dtype=np.bool_, copy=True)
            abl_time[ablated] = 0.0
            optical['mu_a'][ablated] = 0.0
            optical['mu_s'][ablated] = 0.0
            optical['g'][ablated] = 0.0
            optical['n'][ablated] = 1.0
            k_h[ablated] = self.k_void
            rc_h[ablated] = self.rho_cp_void
            T_h[ablated] = self.t_ambient
        else:
            ablated = np.zeros(shape, dtype=np.bool_)
        self.arrhenius = None if arrhenius is None else (float(arrhenius[0]), float(arrhenius[1]) / R_GAS)
        self.dry_state = dry_state
        dry = (np.array(initial_dry_mask, dtype=np.bool_, copy=True) if initial_dry_mask is not None
               else np.zeros(shape, dtype=np.bool_))
        wet = None
        if dry_state is not None:
            if isinstance(dry_state.get('rho_cp'), np.ndarray):
                dry_state['rho_cp'] = np.maximum(dry_state['rho_cp'], k_h / self.alpha_max_stable)
            wet = (optical['mu_a'].copy(), k_h.copy(), rc_h.copy())
        self.t_void = float(np.max(dry_state['t_remove'])) if dry_state is not None else self.t_abl
        self.char_state = char_state
        if char_state is not None:
            self._arrhenius_char = (float(char_state['A']), float(char_state['Ea']) / R_GAS)
            if 't_remove' in char_state:
                self.t_void = float(np.max(char_state['t_remove']))
        else:
            self._arrhenius_char = None
        self.air_gap_range = air_gap_range
        self.wavelength_nm = float(wavelength_nm)

        # device state
        for name, arr in (('MU_A', optical['mu_a']), ('MU_S', optical['mu_s']), ('G', optical['g']),
                          ('N', optical['n']), ('K', k_h), ('RHO_CP', rc_h),
                          ('H_LAT', np.asarray(h_lat_grid, dtype=np.float64)), ('T', T_h), ('T_PEAK', T_h),
                          ('ABLATION_TIME', abl_time), ('DRY', dry), ('ABLATED', ablated)):
            ctx.upload(name, arr)
        for name in ('H_ACC', 'CHI', 'Q', 'OMEGA'):
            ctx.fill(name, 0.0)
        self._set_wet_device(wet)
        del optical, k_h, rc_h, T_h, abl_time

        # plume closure (coupled_ablation.py:318-367)
        self.plume_closure = plume_closure
        self.plume = None
        if isinstance(plume_closure, (PlumeAbsorptionField, MultispectralPlumeField)):
            raise ValueError("pass plume_closure as a dict: a field object owns its own device context")
        if isinstance(plume_closure, dict) and plume_closure.get('enabled', True):
            pc = plume_closure
            D_plume = float(pc.get('D_plume', pc.get('D', 2.0)))
            self.plume = MultispectralPlumeField(
                shape=shape, dx_mm=self.dx, dy_mm=self.dy, dz_mm=self.dz,
                wavelength_nm=float(pc.get('wavelength_nm', self.wavelength_nm)),
                tau_linger_s=float(pc.get('tau_linger_s', 0.00875)),
                sigma_xy_mm=float(pc.get('sigma_xy_mm', pc.get('sigma_mm', 0.060))),
                sigma_z_mm=float(pc.get('sigma_z_mm', pc.get('sigma_mm', 0.040))),
                z_drift_mm=float(pc.get('z_drift_mm', 0.030)),
                delta_mu_per_voxel=pc.get('delta_mu_per_voxel', pc.get('delta_mu')),
                mu_a_max=pc.get('mu_a_max', pc.get('mu_max')),
                model=str(pc.get('model', 'diffusion')).lower(), D_plume=D_plume,
                D_xy=float(pc['D_xy']) if 'D_xy' in pc else D_plume,
                D_z=float(pc['D_z']) if 'D_z' in pc else D_plume,
                v_drift=float(pc.get('v_drift', pc.get('vz', 80.0))),
                blast_expansion=bool(pc.get('blast_expansion', True)),
                blast_sigma_xy_mm=float(pc.get('blast_sigma_xy_mm', 0.080)),
                blast_sigma_z_mm=float(pc.get('blast_sigma_z_mm', 0.090)),
                blast_z_bias_mm=float(pc.get('blast_z_bias_mm', -0.050)),
                dt_clamp=bool(pc.get('dt_clamp', False)),
                _context=ctx, _grid_shape=tuple(plume_grid) if plume_grid is not None else shape)
        self.metal_ablation_ctx = None
        self._ds_arrays = None
        self._cs_arrays = None
        self._cs_chi_remove = 1.0
        self._init_ds_cs_arrays()
        self._push_physics()
        self.t = 0.0
        self.epoch = 0
        self.last_mc = None
        self.last_diag = None

    # ---------------------------------------------------------------- device plumbing
    def _set_wet_device(self, wet):
        self.__dict__['_wet_host'] = None
        if wet is not None:
            self._ctx.upload('WET_MU_A', wet[0])
            self._ctx.upload('WET_K', wet[1])
            self._ctx.upload('WET_RHO_CP', wet[2])
            self.__dict__['_has_wet'] = True
        else:
            self.__dict__['_has_wet'] = False

    @property
    def _wet(self):
        if not self.__dict__.get('_has_wet'):
            return None
        return (self._ctx.download('WET_MU_A'), self._ctx.download('WET_K'), self._ctx.download('WET_RHO_CP'))

    @_wet.setter
    def _wet(self, value):
        self._set_wet_device(value)

    def _push_physics(self):
        p = B.Physics()
        p.dt_sub = self.dt_sub
        p.h_eff = self.h_eff
        p.t_ambient = self.t_ambient
        p.t_abl = self.t_abl
        p.t_void = self.t_void
        p.has_arrhenius = 1 if self.arrhenius is not None else 0
        if self.arrhenius is not None:
            p.arr_A, p.arr_Ea_over_R = self.arrhenius
        p.has_char = 1 if self._arrhenius_char is not None else 0
        if self._arrhenius_char is not None:
            p.char_A, p.char_Ea_over_R = self._arrhenius_char
        p.phase_mode = 0 if self.dry_state is None else 1
        cs = self._cs_arrays
        p.has_cs = 1 if cs is not None else 0
        if cs is not None:
            p.has_cs_mu_a, p.has_cs_mu_s = int('mu_a' in cs), int('mu_s' in cs)
            p.has_cs_k, p.has_cs_rho_cp = int('k' in cs), int('rho_cp' in cs)
        p.cs_chi_remove = self._cs_chi_remove
        p.alpha_max_stable = self.alpha_max_stable
        p.k_void, p.rho_cp_void = self.k_void, self.rho_cp_void
        if self.air_gap_range is not None:
            p.air_gap_k0, p.air_gap_k1 = (int(v) for v in self.air_gap_range)
        else:
            p.air_gap_k0 = p.air_gap_k1 = -1
        p.no_removal = 0
        self._ctx.set_physics(p)

    def _init_ds_cs_arrays(self):
        """coupled_ablation.py:409-438: dry/char parameters (scalar or per voxel) to the device."""
        if self.dry_state is not None:
            self._ds_arrays = {}
            opt = None
            for k in ('mu_a', 'mu_s', 'g', 'n', 'k', 'rho_cp', 't_remove'):
                if k in self.dry_state:
                    v = self.dry_state[k]
                elif k == 't_remove':
                    v = self.t_abl
                elif k in ('mu_a', 'mu_s', 'g', 'n'):
                    opt = opt if opt is not None else self._download_optical()
                    v = opt[k]
                else:
                    v = 0.0
                a = np.array([v], dtype=np.float64) if np.ndim(v) == 0 else np.ascontiguousarray(
                    np.asarray(v).ravel(), dtype=np.float64)
                self._ds_arrays[k] = a
                self._ctx.upload('DS_' + k.upper(), a)
        else:
            self._ds_arrays = None
        if self.char_state is not None:
            cs = self.char_state
            self._cs_arrays = {}
            for k in ('mu_a', 'mu_s', 'k', 'rho_cp'):
                if k in cs:
                    v = cs[k]
                    a = np.array([v], dtype=np.float64) if np.ndim(v) == 0 else np.ascontiguousarray(
                        np.asarray(v).ravel(), dtype=np.float64)
                    self._cs_arrays[k] = a
                    self._ctx.upload('CS_' + k.upper(), a)
            t_rem = cs.get('t_remove', self.dry_state.get('t_remove', self.t_abl) if self.dry_state else self.t_abl)
            a = np.array([t_rem], dtype=np.float64) if np.ndim(t_rem) == 0 else np.ascontiguousarray(
                np.asarray(t_rem).ravel(), dtype=np.float64)
            self._cs_arrays['t_remove'] = a
            self._ctx.upload('CS_T_REMOVE', a)
            self._cs_chi_remove = float(cs.get('chi_remove', 1.0))
        else:
            self._cs_arrays = None
            self._cs_chi_remove = 1.0

    def _download_optical(self):
        return {k: self._ctx.download(k.upper()) for k in ('mu_a', 'mu_s', 'g', 'n')}

    # ---------------------------------------------------------------- host mirrors
    def __getattr__(self, name):
        # only reached for attributes not set on the instance: the device-state mirrors
        cache = self.__dict__.get('_cache')
        if cache is None:
            raise AttributeError(name)
        if name in CoupledAblationEngine._STATE or name in CoupledAblationEngine._MASKS:
            if name not in cache:
                if name in CoupledAblationEngine._MASKS:
                    a = self._ctx.download(CoupledAblationEngine._MASKS[name]).astype(np.bool_)
                else:
                    a = self._ctx.download(CoupledAblationEngine._STATE[name])
                a.flags.writeable = False
                cache[name] = a
            return cache[name]
        if name in ('Q', 'Q_opt'):
            if 'Q' not in cache:
                a = self._ctx.download('Q')
                a.flags.writeable = False
                cache['Q'] = a
            return cache['Q']
        if name == 'Phi':
            if 'Phi' not in cache:
                if self.power > 0.0 and self.score_fluence and self.epoch > 0:
                    a = self._ctx.download('L') / (self.n_photons * self.voxel_vol)
                else:
                    a = np.zeros((self.Nx, self.Ny, self.Nz))
                a.flags.writeable = False
                cache['Phi'] = a
            return cache['Phi']
        if name == 'optical':
            if 'optical' not in cache:
                cache['optical'] = _OpticalDict(self._download_optical())
            self.__dict__['_optical_dirty'] = True
            return cache['optical']
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in CoupledAblationEngine._STATE or name in CoupledAblationEngine._MASKS:
            field = CoupledAblationEngine._STATE.get(name) or CoupledAblationEngine._MASKS[name]
            self._ctx.upload(field, np.asarray(value))
            self.__dict__['_cache'].pop(name, None)
            return
        if name == 'optical':
            for k in ('mu_a', 'mu_s', 'g', 'n'):
                self._ctx.upload(k.upper(), np.asarray(value[k]))
            self.__dict__['_cache'].pop('optical', None)
            return
        object.__setattr__(self, name, value)

    def _invalidate(self):
        self.__dict__['_cache'].clear()

    def apply_damage_optics(self, musp_native, musp_coag, g):
        """The scan benchmark's DOP step (validation_benchmarks.py:116-121) on the device:
        every intact voxel's mu_s becomes ((1 - alpha) musp_native + alpha musp_coag) / (1 - g)
        with alpha = 1 - exp(-Omega).  Same result as the host loop body
        ``eng.optical['mu_s'][intact] = musp[intact] / (1 - g)``, without moving the grids."""
        self._flush_host_writes()
        self._ctx.damage_optics(musp_native, musp_coag, g)
        self.__dict__['_cache'].pop('optical', None)

    def _flush_host_writes(self):
        if self.__dict__.pop('_optical_dirty', False):
            opt = self.__dict__['_cache'].get('optical')
            if opt is not None:
                for k in ('mu_a', 'mu_s', 'g', 'n'):
                    self._ctx.upload(k.upper(), opt[k])

    # ---------------------------------------------------------------- epoch
    def _epoch_beam(self):
        """Flush host-side writes, push the physics, and form this epoch's beam."""
        self._flush_host_writes()
        if self._ds_arrays is None and self.dry_state is not None:
            self._init_ds_cs_arrays()
        self._push_physics()
        self.n_substeps = int(self.n_substeps)
        la_centre = [self.Lx / 2.0, self.Ly / 2.0, 0.0]
        if self.beam_center is not None:
            la_centre[0], la_centre[1] = float(self.beam_center[0]), float(self.beam_center[1])
        return make_beam(self.power, self.n_photons, self.spot_dia, self.theta_inc_deg, self.phi_inc_deg, la_centre,
                         self.r_max, 0, self.beam_profile, mix_seed(self.seed, self.epoch),
                         score_fluence=self.score_fluence, add_plume=self.plume is not None)

    def _finish_epoch(self, d, t_after, wall, mc=True) -> dict:
        self._invalidate()
        self.t = t_after
        self.epoch += 1
        self.last_diag = d
        if self.power > 0.0 and mc:
            self.last_mc = _LazyMC(self, d)
        return {
            'epoch': self.epoch,
            'time_s': self.t,
            'n_new': int(d.n_new),
            'absorbed_frac': float(d.absorbed_frac),
            'mc_energy_closure': float(d.mc_energy_closure),
            't_optical_s': d.t_optical_ms * 1e-3,
            't_thermal_s': d.t_thermal_ms * 1e-3,
            'plume_transmission': float(d.plume_transmission),
            'plume_tau': float(d.plume_tau),
            'plume_peak_mu_a': float(d.plume_peak_mu_a),
            'plume_mean_mu_a': float(d.plume_mean_mu_a),
            't_phase_s': d.t_phase_ms * 1e-3,
            't_wall_s': wall,
            'ledger_J': {'absorbed': d.absorbed_J, 'reflected': d.reflected_J, 'transmitted': d.transmitted_J,
                         'lateral': d.lateral_J, 'conducted': d.conducted_J, 'ablated': d.ablated_J},
        }

    def step_epoch(self) -> dict:
        """Advance one coupled epoch (coupled_ablation.py:440-617); returns its diagnostics."""
        beam = self._epoch_beam()
        t_after = self.t + self.dt_epoch
        t0 = time.perf_counter()
        if self.graph and self.graphable():
            # same MC key (mix_seed(seed, epoch)) and removal time (t + dt_epoch), formed on the device
            d = self._ctx.epoch_graph(beam, self.n_substeps, self.t, self.epoch, self.seed)
        else:
            d = self._ctx.epoch(beam, self.n_substeps, t_after)
        return self._finish_epoch(d, t_after, time.perf_counter() - t0)

    def graphable(self) -> bool:
        return (self.n_substeps % 2 == 0 and abs(self.n_substeps * self.dt_sub - self.dt_epoch)
                <= 1e-12 * self.dt_epoch and (self.plume is None or (self.plume.model == 'diffusion'
                                                                     and not self.plume.blast_expansion)))

    def step_epoch_launch(self):
        """Queue one graph-captured epoch without waiting (several engines can
        then have epochs in flight on their own streams); finish it with
        step_epoch_finish before any other call on this engine."""
        beam = self._epoch_beam()
        if not self.graphable():
            raise ValueError("step_epoch_launch needs a graph-capturable configuration (even sub-step count, "
                             "diffusion plume without blast expansion)")
        self._pending = (self.t + self.dt_epoch, time.perf_counter())
        self._ctx.epoch_graph_launch(beam, self.n_substeps, self.t, self.epoch, self.seed)

    @staticmethod
    def step_epochs_gang(engines):
        """One epoch of every engine, their photon transport in a single launch
        (bd_epoch_gang_launch); the rest of each epoch runs on the engine's own
        stream.  The engines need a graph-capturable configuration (even
        sub-step count, diffusion plume without blast expansion) and one
        sub-step count.  Returns each engine's diagnostics, as step_epoch."""
        if not engines:
            return []
        ns = {e.n_substeps for e in engines}
        if len(ns) != 1:
            raise ValueError("step_epochs_gang: the engines must share the sub-step count")
        beams = []
        for e in engines:
            beams.append(e._epoch_beam())
            if not e.graphable():
                raise ValueError("step_epochs_gang needs graph-capturable configurations")
        t0 = time.perf_counter()
        B.epoch_gang_launch([e._ctx for e in engines], beams, engines[0].n_substeps, [e.t for e in engines],
                            [e.epoch for e in engines], [e.seed for e in engines])
        out = []
        for e in engines:
            d = e._ctx.epoch_graph_finish()
            out.append(e._finish_epoch(d, e.t + e.dt_epoch, time.perf_counter() - t0))
        return out

    @staticmethod
    def step_epochs_batch(engines):
        """One epoch of every engine as a single graph launch (bd_epoch_batch_launch): each engine's
        captured epoch graph is a child node of one parent graph.  The engines need a
        graph-capturable configuration and one sub-step count.  Returns each engine's
        diagnostics, as step_epoch."""
        if not engines:
            return []
        ns = {e.n_substeps for e in engines}
        if len(ns) != 1:
            raise ValueError("step_epochs_batch: the engines must share the sub-step count")
        beams = []
        for e in engines:
            beams.append(e._epoch_beam())
            if not e.graphable():
                raise ValueError("step_epochs_batch needs graph-capturable configurations")
        t0 = time.perf_counter()
        B.epoch_batch_launch([e._ctx for e in engines], beams, engines[0].n_substeps, [e.t for e in engines],
                             [e.epoch for e in engines], [e.seed for e in engines])
        out = []
        for e in engines:
            d = e._ctx.epoch_graph_finish()
            out.append(e._finish_epoch(d, e.t + e.dt_epoch, time.perf_counter() - t0))
        return out

    def step_epoch_finish(self) -> dict:
        t_after, t0 = self._pending
        d = self._ctx.epoch_graph_finish()
        self._pending = None
        return self._finish_epoch(d, t_after, time.perf_counter() - t0)

    def step_epoch_with_q(self) -> dict:
        """One epoch whose absorbed-power field is already on the device (written
        by a surrogate into Q): sub-steps, phase update, removal and plume as in
        step_epoch, no transport.  ``mc_energy_closure`` is NaN and the
        reflected / transmitted / lateral ledger entries are zero."""
        beam = self._epoch_beam()
        t_after = self.t + self.dt_epoch
        t0 = time.perf_counter()
        d = self._ctx.epoch_q(beam, self.n_substeps, t_after)
        return self._finish_epoch(d, t_after, time.perf_counter() - t0, mc=False)

    def compute_void_mask(self) -> np.ndarray:
        """coupled_ablation.py:619-626."""
        void = np.array(self.ablated, copy=True)
        if self.air_gap_range is not None:
            k0, k1 = self.air_gap_range
            if k0 > 0 and np.any(void[:, :, k0 - 1]):
                void[:, :, k0:k1] = True
        return void

    def centerline_depth_mm(self, ix=None, iy=None, k_surface=0) -> float:
        """coupled_ablation.py:662-667."""
        ix = self.Nx // 2 if ix is None else ix
        iy = self.Ny // 2 if iy is None else iy
        idx = np.where(self.ablated[ix, iy, k_surface:])[0]
        return float((idx.max() + 1) * self.dz) if idx.size else 0.0

    def phase_map(self, is_air_gap=None) -> np.ndarray:
        """compute_phase_map (run_multispectral_matrix.py:168-192) evaluated on the device."""
        return self._ctx.phase_map(is_air_gap)

    def slice_xz(self, field: str, iy: int, stride: int = 1) -> np.ndarray:
        """x-z plane of a state field at y index iy (float32), without pulling the grid."""
        names = dict(CoupledAblationEngine._STATE, **CoupledAblationEngine._MASKS, Q='Q')
        return self._ctx.slice(names[field], 1, iy, stride)

    def close(self):
        if self.__dict__.get('_ctx') is not None:
            self._ctx.close()


class _LazyMC(dict):
    """last_mc dictionary of the reference (voxel_mc_3d.py:690-717) with the grids fetched on demand."""

    def __init__(self, eng, d):
        raw = {'R_spec': d.R_spec, 'R_diff': d.R_diff, 'Absorption': float(d.absorbed_frac),
               'Transmission': d.transmission, 'Lateral_loss': d.lateral_loss}
        spectro = {'R_spec': round(d.R_spec * 100.0, 2), 'R_diff': round(d.R_diff * 100.0, 2),
                   'R_tot': round((d.R_spec + d.R_diff) * 100.0, 2),
                   'Absorption': round(float(d.absorbed_frac) * 100.0, 2),
                   'Transmission': round(d.transmission * 100.0, 2), 'Lateral_loss': round(d.lateral_loss * 100.0, 2)}
        super().__init__(energy_closure=float(d.mc_energy_closure), engine='gpu', launch_axis='z',
                         elapsed_seconds=round(d.t_optical_ms * 1e-3, 3), spectro=spectro, spectro_raw=raw,
                         events=int(d.mc_events),
                         photons_per_sec=int(eng.n_photons / max(d.t_optical_ms * 1e-3, 1e-6)))
        self._eng = eng
        self._epoch = eng.epoch

    def __missing__(self, key):
        eng = self._eng
        if eng.epoch != self._epoch:
            raise KeyError(f"{key}: grids of an earlier epoch are no longer on the device")
        N = float(eng.n_photons)
        if key == 'A_3d':
            v = eng._ctx.download('A') / N
        elif key == 'Q_3d':
            v = np.array(eng.Q)
        elif key == 'Phi_3d':
            v = np.array(eng.Phi)
        else:
            raise KeyError(key)
        self[key] = v
        return v
