"""FP64 CPU oracle: restatement of the reference's CPU solver path.

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP product path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package never
does, and the product fails loudly instead of falling back to it.

The heavy kernels live in ``bd_oracle.c`` (compiled by ``oracle/Makefile``);
the per-epoch glue (property dictionaries, ``_remove``, void mask, plume field
bookkeeping) is restated here in numpy, line by line from the reference:

* ``simulation/voxel_mc_3d.py``        -> :func:`run_voxel_mc3d`
* ``simulation/voxel_thermal_3d.py``   -> :func:`thermal_step`, :func:`plateau`
* ``simulation/coupled_ablation.py``   -> :class:`OracleEngine`
* ``simulation/plume.py``              -> :class:`OraclePlume`
* ``simulation/run_multispectral_matrix.py:168-192`` -> :func:`compute_phase_map`

Parity pins for every function are listed in tests/test_oracle_*.py.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libbdoracle.so")
_lib = None

K_AIR = 0.026e-3
RHO_CP_AIR = 1.206e-6
R_GAS = 8.314
MC3D_ENERGY_CLOSURE_TOL = 0.01
MAX_EVENTS_PER_PHOTON = 10_000_000

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class _Param(ctypes.Structure):
    _fields_ = [("v", _dp), ("n", ctypes.c_int64)]


class _PhaseParams(ctypes.Structure):
    _fields_ = [
        ("ds_mu_a", _Param), ("ds_mu_s", _Param), ("ds_g", _Param), ("ds_n", _Param),
        ("ds_k", _Param), ("ds_rho_cp", _Param), ("ds_t_rem", _Param),
        ("has_cs", ctypes.c_int), ("has_cs_mu_a", ctypes.c_int), ("has_cs_mu_s", ctypes.c_int),
        ("has_cs_k", ctypes.c_int), ("has_cs_rho_cp", ctypes.c_int),
        ("cs_mu_a", _Param), ("cs_mu_s", _Param), ("cs_k", _Param), ("cs_rho_cp", _Param),
        ("cs_t_rem", _Param),
        ("cs_chi_rem", ctypes.c_double),
        ("alpha_max_stable", ctypes.c_double),
    ]


class _MCParams(ctypes.Structure):
    _fields_ = [
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
        ("sigma_axis", ctypes.c_double), ("beam_radius", ctypes.c_double),
        ("beam_profile", ctypes.c_int),
        ("theta_rad", ctypes.c_double), ("phi_rad", ctypes.c_double),
        ("x_center", ctypes.c_double), ("y_center", ctypes.c_double), ("z_center", ctypes.c_double),
        ("n_ambient", ctypes.c_double), ("r_max", ctypes.c_double),
        ("launch_axis", ctypes.c_int),
        ("n_photons", ctypes.c_int64), ("photon_offset", ctypes.c_int64),
        ("seed", ctypes.c_uint64), ("max_events", ctypes.c_int64),
        ("n_threads", ctypes.c_int),
    ]


def build():
    """Compile the C oracle (``make -C oracle``)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.bdo_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.bdo_thermal_step.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.bdo_plateau.argtypes = [_dp, _dp, _dp, _u8p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _u8p, ctypes.c_double]
        L.bdo_arrhenius.argtypes = [_dp, _dp, _u8p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int64]
        L.bdo_char.argtypes = [_dp, _dp, _u8p, _u8p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_int64]
        L.bdo_single_stage.argtypes = [_dp, _dp, _dp, ctypes.c_double, _u8p, _u8p, ctypes.c_int64]
        L.bdo_phase_update.argtypes = [_dp, _dp, _dp, _dp, _u8p, _u8p, _dp, _dp, _dp, _dp, _dp, _dp,
                                       _dp, _dp, _dp, ctypes.POINTER(_PhaseParams), _u8p, ctypes.c_int64]
        L.bdo_plume_diffuse.argtypes = [_fp, _fp, _u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.bdo_plume_convolve.argtypes = [_fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int,
                                         _fp, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.bdo_mc_trace.argtypes = [ctypes.POINTER(_MCParams), _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.bdo_mc_trace.restype = ctypes.c_int64
        L.bdo_mc_set_visit_grids.argtypes = [_dp, _dp]
        L.bdo_u01.argtypes = [ctypes.c_uint32]
        L.bdo_u01.restype = ctypes.c_double
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def _u8(a):
    assert a.dtype in (np.uint8, np.bool_) and a.flags.c_contiguous
    return a.ctypes.data_as(_u8p)


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (ctypes.c_uint32 * 4)()
    lib().bdo_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def mix_seed(seed: int, epoch: int) -> int:
    """SplitMix64 of (seed + epoch * golden gamma): the per-epoch MC key."""
    z = (int(seed) + (int(epoch) + 1) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


# --------------------------------------------------------------------------- kernels
def thermal_step(T, T_new, Q, k_grid, rho_cp_grid, Nx, Ny, Nz, dx, dy, dz, dt, h_eff, t_ambient, is_active):
    """voxel_thermal_3d.py:30-112."""
    lib().bdo_thermal_step(_d(T), _d(T_new), _d(Q), _d(k_grid), _d(rho_cp_grid), int(Nx), int(Ny), int(Nz),
                           float(dx), float(dy), float(dz), float(dt), float(h_eff), float(t_ambient),
                           int(bool(is_active)))


def plateau(T, H_acc, rho_cp_grid, ablated, t_abl, Nx, Ny, Nz, dry, t_void):
    """voxel_thermal_3d.py:115-143."""
    lib().bdo_plateau(_d(T), _d(H_acc), _d(rho_cp_grid), _u8(ablated), float(t_abl), int(Nx), int(Ny),
                      int(Nz), _u8(dry), float(t_void))


def arrhenius(T, Omega, ablated, A, Ea_over_R, dt):
    """coupled_ablation.py:33-40."""
    lib().bdo_arrhenius(_d(T), _d(Omega), _u8(ablated), float(A), float(Ea_over_R), float(dt), T.size)


def char(T, chi, dry, ablated, A, Ea_over_R, dt):
    """coupled_ablation.py:43-58."""
    lib().bdo_char(_d(T), _d(chi), _u8(dry), _u8(ablated), float(A), float(Ea_over_R), float(dt), T.size)


def single_stage(H_acc, h_lat, T, t_abl, ablated, to_remove):
    """coupled_ablation.py:61-74."""
    lib().bdo_single_stage(_d(H_acc), _d(h_lat), _d(T), float(t_abl), _u8(ablated), _u8(to_remove), T.size)


def _param(v, keep):
    a = np.ascontiguousarray(np.atleast_1d(np.asarray(v, dtype=np.float64)).ravel())
    keep.append(a)
    return _Param(_d(a), a.size)


def phase_update(H_acc, h_lat, T, chi, dry, ablated, mu_a, mu_s, g, n, k_grid, rho_cp_grid,
                 wet_mu_a, wet_k, wet_rho_cp, ds, cs, cs_chi_remove, alpha_max_stable, to_remove):
    """coupled_ablation.py:77-180.  ``ds``/``cs`` map names to scalars or per-voxel arrays."""
    keep = []
    p = _PhaseParams()
    for key in ("mu_a", "mu_s", "g", "n", "k", "rho_cp"):
        setattr(p, "ds_" + key, _param(ds[key], keep))
    p.ds_t_rem = _param(ds["t_remove"], keep)
    has_cs = cs is not None
    p.has_cs = int(has_cs)
    dummy = np.zeros(1)
    for key, attr in (("mu_a", "cs_mu_a"), ("mu_s", "cs_mu_s"), ("k", "cs_k"), ("rho_cp", "cs_rho_cp")):
        present = has_cs and key in cs
        setattr(p, "has_" + attr, int(present))
        setattr(p, attr, _param(cs[key] if present else dummy, keep))
    p.cs_t_rem = _param(cs["t_remove"] if has_cs else dummy, keep)
    p.cs_chi_rem = float(cs_chi_remove)
    p.alpha_max_stable = float(alpha_max_stable)
    lib().bdo_phase_update(_d(H_acc), _d(h_lat), _d(T), _d(chi), _u8(dry), _u8(ablated), _d(mu_a), _d(mu_s),
                           _d(g), _d(n), _d(k_grid), _d(rho_cp_grid), _d(wet_mu_a), _d(wet_k), _d(wet_rho_cp),
                           ctypes.byref(p), _u8(to_remove), T.size)


def plume_diffuse(C, C_new, void_mask, Nx, Ny, Nz, dx, dy, dz, dt, D_xy, D_z, vz, decay):
    """plume.py:18-80."""
    lib().bdo_plume_diffuse(_f(C), _f(C_new), _u8(void_mask), int(Nx), int(Ny), int(Nz), float(dx), float(dy),
                            float(dz), float(dt), float(D_xy), float(D_z), float(vz), float(decay))


def plume_convolve(raw, impulse, kx, ky, kz, delta_mu, x0, y0, z0):
    """plume.py:83-139."""
    impulse = np.ascontiguousarray(impulse, dtype=np.float32)
    kx = np.ascontiguousarray(kx, dtype=np.float32)
    ky = np.ascontiguousarray(ky, dtype=np.float32)
    kz = np.ascontiguousarray(kz, dtype=np.float32)
    Nx, Ny, Nz = raw.shape
    Lx, Ly, Lz = impulse.shape
    lib().bdo_plume_convolve(_f(raw), Nx, Ny, Nz, _f(impulse), Lx, Ly, Lz, _f(kx), kx.size, _f(ky), ky.size,
                             _f(kz), kz.size, float(delta_mu), int(x0), int(y0), int(z0))


# --------------------------------------------------------------------------- MC
_PROFILES = {'gaussian': 0, 'tophat': 1, 'kim_exp_disk': 2, 'kim_gauss_disk': 3}


def damage_optics(mu_s, Omega, ablated, musp_native, musp_coag, g):
    """validation_benchmarks.py:116-121 (the scan benchmark's DOP option): intact
    voxels' mu_s from their Arrhenius damage, in float64 on the given Omega."""
    mu_s = np.array(mu_s, dtype=np.float64, copy=True)
    alpha = 1.0 - np.exp(-np.asarray(Omega, dtype=np.float64))
    musp = (1.0 - alpha) * musp_native + alpha * musp_coag
    intact = ~np.asarray(ablated, dtype=bool)
    mu_s[intact] = musp[intact] / (1.0 - g)
    return mu_s


def layer_grids(Nx, Ny, Nz, Lx, Ly, Lz, layers_config, launch_axis='z'):
    """voxel_mc_3d.py:511-548: planar layer stack voxelised by cell centres."""
    dx, dy, dz = Lx / Nx, Ly / Ny, Lz / Nz
    mu_a = np.zeros((Nx, Ny, Nz))
    mu_s = np.zeros((Nx, Ny, Nz))
    g = np.zeros((Nx, Ny, Nz))
    n = np.ones((Nx, Ny, Nz))
    if layers_config is not None:
        if launch_axis.strip().lower() != 'x':
            grid = np.linspace(dz / 2.0, Lz - dz / 2.0, Nz)
            s = 0.0
            for lay in layers_config:
                m = (grid >= s) & (grid < lay['z_end'])
                if np.any(m):
                    mu_a[:, :, m] = lay['mu_a']
                    mu_s[:, :, m] = lay['mu_s']
                    g[:, :, m] = lay['g']
                    n[:, :, m] = lay['n']
                s = lay['z_end']
        else:
            grid = np.linspace(dx / 2.0, Lx - dx / 2.0, Nx)
            s = 0.0
            for lay in layers_config:
                m = (grid >= s) & (grid < lay['z_end'])
                if np.any(m):
                    mu_a[m] = lay['mu_a']
                    mu_s[m] = lay['mu_s']
                    g[m] = lay['g']
                    n[m] = lay['n']
                s = lay['z_end']
    return mu_a, mu_s, g, n


def run_voxel_mc3d(power=1.0, pulse_duration=0.20, n_photons=200000, spot_dia=0.50, theta_inc_deg=0.0,
                   phi_inc_deg=0.0, Lx=4.0, Ly=4.0, Lz=2.42, Nx=80, Ny=80, Nz=60, r_max=0.0,
                   layers_config=None, custom_3d_props=None, launch_axis='z', beam_center=None,
                   beam_profile='gaussian', seed=0, photon_offset=0, n_threads=None,
                   max_events=MAX_EVENTS_PER_PHOTON, visit_grids=None):
    """voxel_mc_3d.py:425-717 on the CPU (FP64), Philox stream ``seed``.

    visit_grids: optional pair of float64 (Nx, Ny, Nz) arrays to which the per-voxel counts of
    merged path-length and absorbed-weight tally additions are added (one per voxel visit, and
    one per visit with a collision): the additions a device transport that sums consecutive
    contributions to one voxel before issuing them would make."""
    dx, dy, dz = Lx / Nx, Ly / Ny, Lz / Nz
    la = 1 if launch_axis.strip().lower() == 'x' else 0
    if custom_3d_props is not None:
        mu_a, mu_s, g, n = (np.ascontiguousarray(custom_3d_props[k], dtype=np.float64)
                            for k in ('mu_a', 'mu_s', 'g', 'n'))
    else:
        mu_a, mu_s, g, n = layer_grids(Nx, Ny, Nz, Lx, Ly, Lz, layers_config, launch_axis)
    if la == 0:
        xc, yc, zc = Lx / 2.0, Ly / 2.0, 0.0
        if beam_center is not None:
            xc, yc = float(beam_center[0]), float(beam_center[1])
    else:
        xc, yc, zc = 0.0, Ly / 2.0, Lz / 2.0
        if beam_center is not None:
            yc, zc = float(beam_center[0]), float(beam_center[1])
    w0 = spot_dia / 2.0
    p = _MCParams()
    p.Nx, p.Ny, p.Nz = int(Nx), int(Ny), int(Nz)
    p.dx, p.dy, p.dz = dx, dy, dz
    p.Lx, p.Ly, p.Lz = float(Lx), float(Ly), float(Lz)
    p.sigma_axis = w0 / 2.0
    p.beam_radius = w0
    p.beam_profile = _PROFILES[beam_profile]
    p.theta_rad = math.radians(theta_inc_deg)
    p.phi_rad = math.radians(phi_inc_deg)
    p.x_center, p.y_center, p.z_center = xc, yc, zc
    p.n_ambient = 1.0
    p.r_max = float(r_max)
    p.launch_axis = la
    p.n_photons = int(n_photons)
    p.photon_offset = int(photon_offset)
    p.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    p.max_events = int(max_events)
    p.n_threads = int(n_threads if n_threads is not None else (os.cpu_count() or 1))
    A = np.zeros((Nx, Ny, Nz))
    L = np.zeros((Nx, Ny, Nz))
    tally = np.zeros(4)
    lat = np.zeros(Nz if la == 0 else Nx)
    if visit_grids is not None:
        vL, vA = visit_grids
        assert vL.shape == vA.shape == (Nx, Ny, Nz)
        lib().bdo_mc_set_visit_grids(_d(vL), _d(vA))
    t0 = time.perf_counter()
    try:
        events = lib().bdo_mc_trace(ctypes.byref(p), _d(mu_a), _d(mu_s), _d(g), _d(n), _d(A), _d(L), _d(tally),
                                    _d(lat))
    finally:
        if visit_grids is not None:
            lib().bdo_mc_set_visit_grids(None, None)
    elapsed = time.perf_counter() - t0
    N = float(n_photons)
    R_spec, R_diff, T_tot, Lat = (float(v) / N for v in tally)
    A_tot = float(A.sum()) / N
    closure = R_spec + R_diff + A_tot + T_tot + Lat
    vol = dx * dy * dz
    return {
        'grid_x': np.linspace(dx / 2, Lx - dx / 2, Nx), 'grid_y': np.linspace(dy / 2, Ly - dy / 2, Ny),
        'grid_z': np.linspace(dz / 2, Lz - dz / 2, Nz),
        'dx': dx, 'dy': dy, 'dz': dz, 'Nx': Nx, 'Ny': Ny, 'Nz': Nz,
        'A_3d': A / N, 'lateral_escape_z': lat / N, 'Q_3d': (A / N) * (power / vol), 'Phi_3d': (L / N) / vol,
        'spectro': {'R_spec': round(R_spec * 100, 2), 'R_diff': round(R_diff * 100, 2),
                    'R_tot': round((R_spec + R_diff) * 100, 2), 'Absorption': round(A_tot * 100, 2),
                    'Transmission': round(T_tot * 100, 2), 'Lateral_loss': round(Lat * 100, 2)},
        'energy_closure': float(closure),
        'spectro_raw': {'R_spec': R_spec, 'R_diff': R_diff, 'Absorption': A_tot, 'Transmission': T_tot,
                        'Lateral_loss': Lat},
        'launch_axis': 'x' if la else 'z', 'elapsed_seconds': elapsed,
        'photons_per_sec': int(n_photons / max(elapsed, 1e-4)), 'events': int(events), 'engine': 'oracle',
    }


# --------------------------------------------------------------------------- stand-alone thermal solver
class OracleThermalSolver3D:
    """voxel_thermal_3d.py:146-309 (VoxelThermalSolver3D) in FP64 on the CPU."""

    CFL_SAFETY_FACTOR = 0.40

    def __init__(self, grid_x, grid_y, grid_z, k_grid, rho_cp_grid, h_conv=15.0, emissivity=0.95, t_ambient=20.0):
        self.Nx, self.Ny, self.Nz = len(grid_x), len(grid_y), len(grid_z)
        self.dx = float(grid_x[1] - grid_x[0]) if self.Nx > 1 else 0.05
        self.dy = float(grid_y[1] - grid_y[0]) if self.Ny > 1 else 0.05
        self.dz = float(grid_z[1] - grid_z[0]) if self.Nz > 1 else 0.04
        k = np.asarray(k_grid, dtype=np.float64)
        rc = np.asarray(rho_cp_grid, dtype=np.float64)
        # :171-180: W/(m K) -> W/(mm K) when k > 0.05; J/(m^3 K) -> J/(mm^3 K) when rho_cp > 100
        self.k_grid = k * 1e-3 if np.max(k) > 0.05 else k.copy()
        self.rho_cp_grid = rc * 1e-9 if np.max(rc) > 100.0 else rc.copy()
        # :185-188: convection plus linearised radiation, W/(m^2 K) -> W/(mm^2 K)
        self.h_eff = (float(h_conv) + 4.0 * 5.670374e-8 * float(emissivity) * (t_ambient + 273.15) ** 3) * 1e-6
        self.t_ambient = float(t_ambient)

    def cfl_conditioned_rho_cp(self, dt):
        """:191-203."""
        inv_d2 = 1.0 / self.dx ** 2 + 1.0 / self.dy ** 2 + 1.0 / self.dz ** 2
        amax = self.CFL_SAFETY_FACTOR / (dt * inv_d2)
        rc = self.rho_cp_grid.copy()
        unstable = self.k_grid / np.maximum(rc, 1e-12) > amax
        if np.any(unstable):
            rc[unstable] = self.k_grid[unstable] / amax
        return rc

    def solve(self, Q_3d, pulse_duration=0.20, total_time=1.00, dt=0.001, snapshot_times=None):
        """:205-309: explicit steps, source on while t < pulse_duration, centre and surface
        histories at every step (step 0 included), snapshots at the rounded target times."""
        n_steps = int(round(total_time / dt))
        rc = self.cfl_conditioned_rho_cp(dt)
        Q = np.ascontiguousarray(Q_3d, dtype=np.float64)
        T = np.full((self.Nx, self.Ny, self.Nz), self.t_ambient)
        Tn = np.empty_like(T)
        if snapshot_times is None:
            snapshot_times = [0.0, pulse_duration, total_time]
        targets = sorted({round(float(t), 4) for t in snapshot_times})
        mx, my, mz = self.Nx // 2, self.Ny // 2, self.Nz // 2
        centre, surface, times, snaps = [], [], [], {}
        for step in range(n_steps + 1):
            cur = round(step * dt, 4)
            centre.append(float(T[mx, my, mz]))
            surface.append(float(T[mx, my, 0]))
            times.append(cur)
            if cur in targets:
                snaps[cur] = T.copy()
            if step < n_steps:
                thermal_step(T, Tn, Q, self.k_grid, rc, self.Nx, self.Ny, self.Nz, self.dx, self.dy, self.dz, dt,
                             self.h_eff, self.t_ambient, cur < pulse_duration)
                T, Tn = Tn, T
        return {'snapshots': snaps, 'time_history': np.array(times), 'center_history': np.array(centre),
                'surface_history': np.array(surface), 'peak_surface_T': max(surface), 'peak_center_T': max(centre),
                'rho_cp_eff': rc, 'T_final': T}


# --------------------------------------------------------------------------- plume
class OraclePlume:
    """plume.py:142-438 (MultispectralPlumeField) restated."""

    EXTINCTION_MAP = {
        10600: {'mu_a_max': 0.15, 'mu_s_max': 0.05, 'delta_mu': 0.075},
        1940: {'mu_a_max': 0.375, 'mu_s_max': 0.15, 'delta_mu': 0.18},
        1470: {'mu_a_max': 0.45, 'mu_s_max': 0.20, 'delta_mu': 0.22},
        1064: {'mu_a_max': 0.675, 'mu_s_max': 0.60, 'delta_mu': 0.33},
        532: {'mu_a_max': 1.275, 'mu_s_max': 0.80, 'delta_mu': 0.63},
        450: {'mu_a_max': 1.80, 'mu_s_max': 1.00, 'delta_mu': 0.90},
        320: {'mu_a_max': 2.40, 'mu_s_max': 1.20, 'delta_mu': 1.20},
    }

    def __init__(self, shape, dx_mm=0.030, dy_mm=0.030, dz_mm=0.020, wavelength_nm=None, tau_linger_s=0.00875,
                 sigma_xy_mm=0.060, sigma_z_mm=0.040, z_drift_mm=0.030, delta_mu_per_voxel=None, mu_a_max=None,
                 sigma_mm=None, model='separable', D_plume=2.0, D_xy=None, D_z=None, v_drift=80.0,
                 blast_expansion=True, blast_sigma_xy_mm=0.080, blast_sigma_z_mm=0.090, blast_z_bias_mm=-0.050):
        self.shape = tuple(shape)
        self.dx, self.dy, self.dz = float(dx_mm), float(dy_mm), float(dz_mm)
        self.tau_linger = float(tau_linger_s)
        self.wavelength = float(wavelength_nm) if wavelength_nm is not None else 10600.0
        self.model = str(model).lower()
        self.D_plume = float(D_plume)
        self.D_xy = float(D_xy) if D_xy is not None else self.D_plume
        self.D_z = float(D_z) if D_z is not None else self.D_plume
        self.v_drift = float(v_drift)
        self.blast_expansion = bool(blast_expansion)
        self.blast_sigma_xy = float(blast_sigma_xy_mm)
        self.blast_sigma_z = float(blast_sigma_z_mm)
        self.blast_z_bias = float(blast_z_bias_mm)
        if sigma_mm is not None:
            self.sigma_xy = self.sigma_z = float(sigma_mm)
            self.z_drift = float(z_drift_mm) if z_drift_mm != 0.030 else 0.0
        else:
            self.sigma_xy, self.sigma_z, self.z_drift = float(sigma_xy_mm), float(sigma_z_mm), float(z_drift_mm)
        cfg = self.EXTINCTION_MAP.get(int(round(self.wavelength)), {'mu_a_max': 0.30, 'mu_s_max': 0.15,
                                                                     'delta_mu': 0.15})
        self.mu_a_max = float(cfg.get('mu_a_max', 0.30))
        self.mu_s_max = float(cfg.get('mu_s_max', 0.25 * self.mu_a_max))
        self.mu_max = self.mu_a_max + self.mu_s_max
        if mu_a_max is not None:
            self.mu_a_max = float(mu_a_max)
            self.mu_max = self.mu_a_max + self.mu_s_max
        self.delta_mu = float(delta_mu_per_voxel) if delta_mu_per_voxel is not None else float(cfg.get('delta_mu', 0.15))
        self.raw_field = np.zeros(self.shape, dtype=np.float32)
        self.mu_ext_plume = np.zeros(self.shape, dtype=np.float32)
        self._tmp_field = np.zeros(self.shape, dtype=np.float32)
        # plume.py:238-252
        self.rx = max(1, int(np.ceil(3.0 * self.sigma_xy / self.dx)))
        self.ry = max(1, int(np.ceil(3.0 * self.sigma_xy / self.dy)))
        self.rz = max(1, int(np.ceil((3.0 * self.sigma_z + self.z_drift) / self.dz)))
        f32 = np.float32
        self.kx_1d = np.exp(-((np.arange(-self.rx, self.rx + 1, dtype=f32) * self.dx) ** 2)
                            / (2.0 * self.sigma_xy ** 2)).astype(f32)
        self.ky_1d = np.exp(-((np.arange(-self.ry, self.ry + 1, dtype=f32) * self.dy) ** 2)
                            / (2.0 * self.sigma_xy ** 2)).astype(f32)
        self.kz_1d = np.exp(-(((np.arange(-self.rz, self.rz + 1, dtype=f32) * self.dz) + self.z_drift) ** 2)
                            / (2.0 * self.sigma_z ** 2)).astype(f32)
        self.kernel = (self.kx_1d[:, None, None] * self.ky_1d[None, :, None] * self.kz_1d[None, None, :]).astype(f32)
        k_max = float(np.max(self.kernel))
        if k_max > 0.0:
            self.kernel /= k_max
            self.kz_1d /= k_max
        # plume.py:255-268
        self.rx_b = max(1, int(np.ceil(2.5 * self.blast_sigma_xy / self.dx)))
        self.ry_b = max(1, int(np.ceil(2.5 * self.blast_sigma_xy / self.dy)))
        self.rz_b = max(1, int(np.ceil((2.5 * self.blast_sigma_z + abs(self.blast_z_bias)) / self.dz)))
        self.kx_b = np.exp(-((np.arange(-self.rx_b, self.rx_b + 1, dtype=f32) * self.dx) ** 2)
                           / (2.0 * self.blast_sigma_xy ** 2)).astype(f32)
        self.ky_b = np.exp(-((np.arange(-self.ry_b, self.ry_b + 1, dtype=f32) * self.dy) ** 2)
                           / (2.0 * self.blast_sigma_xy ** 2)).astype(f32)
        self.kz_b = np.exp(-(((np.arange(-self.rz_b, self.rz_b + 1, dtype=f32) * self.dz) - self.blast_z_bias) ** 2)
                           / (2.0 * self.blast_sigma_z ** 2)).astype(f32)
        k_max_b = float(np.max(self.kx_b) * np.max(self.ky_b) * np.max(self.kz_b))
        if k_max_b > 0.0:
            s = float(0.35 / np.cbrt(k_max_b))
            self.kx_b *= s
            self.ky_b *= s
            self.kz_b *= s

    @property
    def mu_a_plume(self):
        return self.mu_ext_plume * (self.mu_a_max / max(1e-6, self.mu_max))

    @property
    def mu_s_plume(self):
        return self.mu_ext_plume * (self.mu_s_max / max(1e-6, self.mu_max))

    def _splat(self, mask, kx, ky, kz, rx, ry, rz):
        wx = np.where(np.any(mask, axis=(1, 2)))[0]
        wy = np.where(np.any(mask, axis=(0, 2)))[0]
        wz = np.where(np.any(mask, axis=(0, 1)))[0]
        x0, x1, y0, y1, z0, z1 = wx[0], wx[-1] + 1, wy[0], wy[-1] + 1, wz[0], wz[-1] + 1
        imp = mask[x0:x1, y0:y1, z0:z1].astype(np.float32)
        plume_convolve(self.raw_field, imp, kx, ky, kz, self.delta_mu, x0 - rx, y0 - ry, z0 - rz)

    def inject(self, newly_mask, void_mask=None):
        """plume.py:292-331."""
        if newly_mask is None or not np.any(newly_mask):
            return
        if self.blast_expansion:
            self._splat(newly_mask, self.kx_b, self.ky_b, self.kz_b, self.rx_b, self.ry_b, self.rz_b)
        else:
            self.raw_field[newly_mask] += np.float32(self.delta_mu)
        if void_mask is not None:
            self.raw_field[~void_mask] = 0.0

    def diffuse_substep(self, void_mask, dt_sub):
        """plume.py:333-350."""
        decay = float(np.exp(-dt_sub / self.tau_linger)) if (self.tau_linger > 0.0 and not np.isinf(self.tau_linger)) else 1.0
        plume_diffuse(self.raw_field, self._tmp_field, np.ascontiguousarray(void_mask, dtype=np.uint8),
                      *self.shape, self.dx, self.dy, self.dz, dt_sub, self.D_xy, self.D_z, self.v_drift, decay)
        self.raw_field[:] = self._tmp_field

    def set_wavelength(self, wavelength_nm):
        """plume.py:270-278."""
        self.wavelength = float(wavelength_nm)
        cfg = self.EXTINCTION_MAP.get(int(round(self.wavelength)), {'mu_a_max': 0.30, 'mu_s_max': 0.15,
                                                                      'delta_mu': 0.15})
        self.mu_a_max = float(cfg.get('mu_a_max', 0.30))
        self.mu_s_max = float(cfg.get('mu_s_max', 0.25 * self.mu_a_max))
        self.mu_max = self.mu_a_max + self.mu_s_max
        self.delta_mu = float(cfg.get('delta_mu', 0.15))

    def sync_extinction(self, void_mask=None):
        """plume.py:352-366."""
        if void_mask is not None:
            self.raw_field[~void_mask] = 0.0
        if self.mu_max is not None and self.mu_max > 0.0:
            np.tanh(self.raw_field / self.mu_max, out=self.mu_ext_plume)
            self.mu_ext_plume *= self.mu_max
        else:
            self.mu_ext_plume[:] = self.raw_field
        if void_mask is not None:
            self.mu_ext_plume[~void_mask] = 0.0
        return self.mu_ext_plume

    def update(self, dt_s, newly_ablated_mask, void_mask):
        """plume.py:368-432."""
        if not np.any(self.raw_field) and np.any(self.mu_ext_plume):
            self.raw_field[:] = self.mu_ext_plume
        if self.model == 'diffusion':
            self.inject(newly_ablated_mask)
            if dt_s > 0.0:
                n_sub = max(1, int(round(dt_s / 5e-5)))
                for _ in range(n_sub):
                    self.diffuse_substep(void_mask, dt_s / n_sub)
            return self.sync_extinction(void_mask)
        if self.tau_linger > 0.0 and dt_s > 0.0:
            self.raw_field *= float(np.exp(-dt_s / self.tau_linger))
        if np.any(newly_ablated_mask):
            self._splat(newly_ablated_mask, self.kx_1d, self.ky_1d, self.kz_1d, self.rx, self.ry, self.rz)
        self.raw_field *= void_mask
        if self.mu_max is not None and self.mu_max > 0.0:
            np.tanh(self.raw_field / self.mu_max, out=self.mu_ext_plume)
            self.mu_ext_plume *= self.mu_max
        else:
            self.mu_ext_plume[:] = self.raw_field
        return self.mu_ext_plume


# --------------------------------------------------------------------------- engine
class OracleEngine:
    """coupled_ablation.py:183-667 (CoupledAblationEngine) on the CPU in FP64.

    The MC call per epoch uses the Philox stream ``mix_seed(seed, epoch)`` so a
    device run with the same seed traces the same photon streams.
    """

    def __init__(self, optical_props, k_grid, rho_cp_grid, h_lat_grid, Lx, Ly, Lz, power, spot_dia, t_abl,
                 t_ambient=20.0, dt_epoch=0.01, dt_sub=5e-5, n_photons=100000, h_eff=0.0, r_max=0.0,
                 beam_center=None, stability_factor=0.35, k_void=K_AIR, rho_cp_void=RHO_CP_AIR, arrhenius=None,
                 beam_profile='gaussian', dry_state=None, char_state=None, plume_closure=None, theta_inc_deg=0.0,
                 phi_inc_deg=0.0, initial_dry_mask=None, initial_ablated_mask=None, air_gap_range=None,
                 wavelength_nm=10600.0, seed=0, mc_threads=None, plume_grid=None, **_ignored):
        self.optical = {k: np.array(optical_props[k], dtype=np.float64, copy=True) for k in ('mu_a', 'mu_s', 'g', 'n')}
        self.Nx, self.Ny, self.Nz = self.optical['mu_a'].shape
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
        self.seed = int(seed)
        self.mc_threads = mc_threads
        self.k_grid = np.array(k_grid, dtype=np.float64, copy=True)
        self.rho_cp_grid = np.array(rho_cp_grid, dtype=np.float64, copy=True)
        self.h_lat = np.array(h_lat_grid, dtype=np.float64, copy=True)
        inv_d2 = 1.0 / self.dx ** 2 + 1.0 / self.dy ** 2 + 1.0 / self.dz ** 2
        self.alpha_max_stable = stability_factor / (self.dt_sub * inv_d2)
        alpha = self.k_grid / np.maximum(self.rho_cp_grid, 1e-12)
        unstable = alpha > self.alpha_max_stable
        self.cfl_conditioned_mask = unstable.copy()
        if np.any(unstable):
            self.rho_cp_grid[unstable] = self.k_grid[unstable] / self.alpha_max_stable
        self.k_void = float(k_void)
        self.rho_cp_void = max(float(rho_cp_void), self.k_void / self.alpha_max_stable)
        shape = (self.Nx, self.Ny, self.Nz)
        self.T = np.full(shape, self.t_ambient)
        self._T_next = np.empty_like(self.T)
        self.H_acc = np.zeros(shape)
        self.ablation_time = np.full(shape, -1.0)
        if initial_ablated_mask is not None:
            self.ablated = np.array(initial_ablated_mask, dtype=np.bool_, copy=True)
            self.ablation_time[self.ablated] = 0.0
            self.optical['mu_a'][self.ablated] = 0.0
            self.optical['mu_s'][self.ablated] = 0.0
            self.optical['g'][self.ablated] = 0.0
            self.optical['n'][self.ablated] = 1.0
            self.k_grid[self.ablated] = self.k_void
            self.rho_cp_grid[self.ablated] = self.rho_cp_void
            self.T[self.ablated] = self.t_ambient
        else:
            self.ablated = np.zeros(shape, dtype=np.bool_)
        self.Q = np.zeros(shape)
        self.arrhenius = None if arrhenius is None else (float(arrhenius[0]), float(arrhenius[1]) / R_GAS)
        self.Omega = np.zeros(shape)
        self.T_peak = np.full(shape, self.t_ambient)
        self.dry_state = dry_state
        self.dry = (np.array(initial_dry_mask, dtype=np.bool_, copy=True) if initial_dry_mask is not None
                    else np.zeros(shape, dtype=np.bool_))
        self._wet = None
        if dry_state is not None:
            dry_state = dict(dry_state)
            self.dry_state = dry_state
            if isinstance(dry_state.get('rho_cp'), np.ndarray):
                dry_state['rho_cp'] = np.maximum(dry_state['rho_cp'], self.k_grid / self.alpha_max_stable)
            self._wet = (self.optical['mu_a'].copy(), self.k_grid.copy(), self.rho_cp_grid.copy())
        self.t_void = float(np.max(dry_state['t_remove'])) if dry_state is not None else self.t_abl
        self.char_state = char_state
        self.chi = np.zeros(shape)
        if char_state is not None:
            self._arrhenius_char = (float(char_state['A']), float(char_state['Ea']) / R_GAS)
            if 't_remove' in char_state:
                self.t_void = float(np.max(char_state['t_remove']))
        else:
            self._arrhenius_char = None
        self.air_gap_range = air_gap_range
        self.wavelength_nm = float(wavelength_nm)
        self.plume = None
        # plume grid: the thermal grid (reference layout) or a coarser grid over the same box,
        # coupled by cell-centre sampling (thermal -> plume index: (2i+1) P // (2N), and back)
        self.pshape = tuple(int(v) for v in plume_grid) if plume_grid is not None else shape
        self._same_plume = self.pshape == shape
        self._tp = [((2 * np.arange(N) + 1) * P) // (2 * N) for N, P in zip(shape, self.pshape)]
        self._pt = [((2 * np.arange(P) + 1) * N) // (2 * P) for N, P in zip(shape, self.pshape)]
        if isinstance(plume_closure, dict) and plume_closure.get('enabled', True):
            pc = plume_closure
            D_plume = float(pc.get('D_plume', pc.get('D', 2.0)))
            self.plume = OraclePlume(
                shape=self.pshape, dx_mm=self.Lx / self.pshape[0], dy_mm=self.Ly / self.pshape[1],
                dz_mm=self.Lz / self.pshape[2],
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
                blast_z_bias_mm=float(pc.get('blast_z_bias_mm', -0.050)))
            # the Metal solver's diffusion step (metal_ablation_solver.py:234-236), when asked for
            self.plume_dt_clamp = bool(pc.get('dt_clamp', False))
        self._to_remove = np.zeros(shape, dtype=np.bool_)
        self._init_ds_cs_arrays()
        self.t = 0.0
        self.epoch = 0
        self.last_mc = None

    def _init_ds_cs_arrays(self):
        """coupled_ablation.py:409-438."""
        if self.dry_state is not None:
            self._ds = {}
            for k in ('mu_a', 'mu_s', 'g', 'n', 'k', 'rho_cp', 't_remove'):
                if k in self.dry_state:
                    v = self.dry_state[k]
                elif k == 't_remove':
                    v = self.t_abl
                elif k in self.optical:
                    v = self.optical[k]
                else:
                    v = 0.0
                self._ds[k] = np.array(v, dtype=np.float64).ravel().copy()
        else:
            self._ds = None
        if self.char_state is not None:
            cs = self.char_state
            self._cs = {k: np.array(cs[k], dtype=np.float64).ravel().copy()
                        for k in ('mu_a', 'mu_s', 'k', 'rho_cp') if k in cs}
            t_rem = cs.get('t_remove', self.dry_state.get('t_remove', self.t_abl) if self.dry_state else self.t_abl)
            self._cs['t_remove'] = np.array(t_rem, dtype=np.float64).ravel().copy()
            self._cs_chi_remove = float(cs.get('chi_remove', 1.0))
        else:
            self._cs = None
            self._cs_chi_remove = 1.0

    def compute_void_mask(self):
        """coupled_ablation.py:619-626."""
        v = self.ablated.copy()
        if self.air_gap_range is not None:
            k0, k1 = self.air_gap_range
            if k0 > 0 and np.any(self.ablated[:, :, k0 - 1]):
                v[:, :, k0:k1] = True
        return v

    def _to_plume(self, a):
        """thermal-grid mask -> plume grid (value under each plume cell centre)"""
        return a if self._same_plume else a[np.ix_(*self._pt)]

    def _from_plume(self, a):
        """plume-grid field -> thermal grid (plume cell under each thermal cell centre)"""
        return a if self._same_plume else a[np.ix_(*self._tp)]

    def _inject(self, newly):
        if self._same_plume:
            self.plume.inject(newly, void_mask=self.compute_void_mask())
            return
        ii, jj, kk = np.nonzero(newly)
        amount = np.float32(self.plume.delta_mu * (np.prod(self.pshape) / newly.size))
        np.add.at(self.plume.raw_field, (self._tp[0][ii], self._tp[1][jj], self._tp[2][kk]), amount)
        self.plume.raw_field[~self._to_plume(self.compute_void_mask())] = 0.0

    def _remove(self, newly):
        """coupled_ablation.py:628-660."""
        n_new = int(np.sum(newly))
        if n_new > 0:
            self.ablated[newly] = True
            self.dry[newly] = False
            self.chi[newly] = 0.0
            self.ablation_time[newly] = self.t
            self.optical['mu_a'][newly] = 0.0
            self.optical['mu_s'][newly] = 0.0
            self.optical['g'][newly] = 0.0
            self.optical['n'][newly] = 1.0
            self.k_grid[newly] = self.k_void
            self.rho_cp_grid[newly] = self.rho_cp_void
            self.T[newly] = self.t_void
            self.H_acc[newly] = 0.0
            if self.air_gap_range is not None:
                k0, k1 = self.air_gap_range
                if k0 > 0:
                    cols = self.ablated[:, :, k0 - 1]
                    if np.any(cols):
                        for ix, iy in zip(*np.where(cols)):
                            if not self.ablated[ix, iy, k0]:
                                self.ablated[ix, iy, k0:k1] = True
                                self.ablation_time[ix, iy, k0:k1] = self.t
            if self.plume is not None:
                self._inject(newly)
        return n_new

    # -- one epoch, in the reference's order (coupled_ablation.py:440-617), split in three
    #    parts so a test can drive the second and third from given fields:
    #    optics (MC or a given Q), sub-steps, then phase update / removal / plume.
    def _optics(self):
        """coupled_ablation.py:442-486: MC of the current optics (+ plume), Q, Q_thermal."""
        t0 = time.perf_counter()
        if self.power > 0.0:
            if self.plume is not None and np.any(self.plume.mu_ext_plume > 0.0):
                opt = {k: v.copy() for k, v in self.optical.items()}
                opt['mu_a'] = self.optical['mu_a'] + self._from_plume(self.plume.mu_a_plume)
                opt['mu_s'] = self.optical['mu_s'] + self._from_plume(self.plume.mu_s_plume)
                opt['g'][(self._from_plume(self.plume.mu_ext_plume) > 0.001) & self.ablated] = 0.80
            else:
                opt = self.optical
            mc = run_voxel_mc3d(power=self.power, n_photons=self.n_photons, spot_dia=self.spot_dia, Lx=self.Lx,
                                Ly=self.Ly, Lz=self.Lz, Nx=self.Nx, Ny=self.Ny, Nz=self.Nz, r_max=self.r_max,
                                custom_3d_props=opt, launch_axis='z', beam_center=self.beam_center,
                                beam_profile=self.beam_profile, theta_inc_deg=self.theta_inc_deg,
                                phi_inc_deg=self.phi_inc_deg, seed=mix_seed(self.seed, self.epoch),
                                n_threads=self.mc_threads)
            self.last_mc = mc
            self.Q_opt = mc['Q_3d']
            self.Phi = mc['Phi_3d']
            self.Q = self.Q_opt.copy()
            Q_th = self.Q_opt.copy()
            if np.any(self.ablated):
                Q_th[self.ablated] = 0.0
            absorbed = float(np.sum(mc['A_3d']))
            closure = float(mc['energy_closure'])
        else:
            self.Q = np.zeros_like(self.T)
            self.Q_opt = self.Q
            self.Phi = np.zeros_like(self.T)
            Q_th = self.Q
            absorbed, closure = 0.0, 1.0
        return Q_th, absorbed, closure, time.perf_counter() - t0

    def _boundary_loss(self, T):
        """Energy [J] leaving through the top and bottom faces in one sub-step: the
        boundary term of voxel_thermal_3d.py:97,105 (half-cell conduction in series
        with h_eff), summed over the z = 0 and z = Nz-1 planes (both for Nz = 1)."""
        if self.h_eff <= 0.0:
            return 0.0
        loss = 0.0
        for kk in (0, self.Nz - 1):
            kc = self.k_grid[:, :, kk]
            rser = self.dz * (0.5 * self.dz / np.maximum(kc, 1e-30) + 1.0 / max(self.h_eff, 1e-30))
            loss += float(np.sum((T[:, :, kk] - self.t_ambient) / rser))
        return loss * self.voxel_vol * self.dt_sub

    def run_substeps(self, Q_th):
        """coupled_ablation.py:487-545: n_substeps fused steps (stencil, plateau, Arrhenius,
        char, void-confined plume diffusion), then T_peak.  Returns the conducted energy [J]."""
        has_diff = self.plume is not None and self.plume.model == 'diffusion'
        void = self.compute_void_mask() if has_diff else self.ablated
        ab8 = self.ablated.view(np.uint8)
        dr8 = self.dry.view(np.uint8)
        v8 = np.ascontiguousarray(self._to_plume(void), dtype=np.uint8)
        T, Tn = self.T, self._T_next
        if has_diff:
            P, Pn = self.plume.raw_field, self.plume._tmp_field
            tl = self.plume.tau_linger
            decay = float(np.exp(-self.dt_sub / tl)) if (tl > 0.0 and not np.isinf(tl)) else 1.0
            dt_p = self.dt_sub
            if getattr(self, 'plume_dt_clamp', False):
                pl = self.plume
                d_rate = pl.D_xy * (1.0 / pl.dx ** 2 + 1.0 / pl.dy ** 2) + pl.D_z / pl.dz ** 2
                if d_rate > 1e-6:
                    dt_p = min(dt_p, 0.45 / (2.0 * d_rate))
        N = self.Nx, self.Ny, self.Nz
        conducted = 0.0
        for _ in range(self.n_substeps):
            conducted += self._boundary_loss(T)
            thermal_step(T, Tn, Q_th, self.k_grid, self.rho_cp_grid, *N, self.dx, self.dy, self.dz, self.dt_sub,
                         self.h_eff, self.t_ambient, True)
            plateau(Tn, self.H_acc, self.rho_cp_grid, ab8, self.t_abl, *N, dr8, self.t_void)
            if self.arrhenius is not None:
                arrhenius(Tn, self.Omega, ab8, self.arrhenius[0], self.arrhenius[1], self.dt_sub)
            if self._arrhenius_char is not None:
                char(Tn, self.chi, dr8, ab8, self._arrhenius_char[0], self._arrhenius_char[1], self.dt_sub)
            if has_diff:
                plume_diffuse(P, Pn, v8, *self.pshape, self.plume.dx, self.plume.dy, self.plume.dz, dt_p,
                              self.plume.D_xy, self.plume.D_z, self.plume.v_drift, decay)
                P, Pn = Pn, P
            T, Tn = Tn, T
        self.T, self._T_next = T, Tn
        if has_diff:
            if P is not self.plume.raw_field:
                self.plume.raw_field[:] = P
        np.maximum(self.T_peak, self.T, out=self.T_peak)
        return conducted

    def finish_epoch(self):
        """coupled_ablation.py:547-617: time step, phase update (or the single-stage
        test), _remove with air-gap perforation, plume sync and centre-column diagnostics.
        Returns (n_new, ablated energy [J], plume tau, transmission, peak, mean)."""
        self.t += self.dt_epoch
        self.epoch += 1
        rem = self._to_remove
        ab8 = self.ablated.view(np.uint8)
        dr8 = self.dry.view(np.uint8)
        if self.dry_state is None:
            single_stage(self.H_acc, self.h_lat, self.T, self.t_abl, ab8, rem.view(np.uint8))
        else:
            if self._ds is None:
                self._init_ds_cs_arrays()
            phase_update(self.H_acc, self.h_lat, self.T, self.chi, dr8, ab8, self.optical['mu_a'],
                         self.optical['mu_s'], self.optical['g'], self.optical['n'], self.k_grid,
                         self.rho_cp_grid, self._wet[0], self._wet[1], self._wet[2], self._ds, self._cs,
                         self._cs_chi_remove, self.alpha_max_stable, rem.view(np.uint8))
        # energy the removed cells held above ambient, with their properties after the update
        ablated_J = float(np.sum(self.rho_cp_grid[rem] * (self.T[rem] - self.t_ambient) + self.H_acc[rem])) \
            * self.voxel_vol
        n_new = self._remove(rem)
        if self.plume is not None:
            vm = self._to_plume(self.compute_void_mask())
            if self.plume.model == 'diffusion':
                self.plume.sync_extinction(vm)
            else:
                self.plume.update(self.dt_epoch, rem, vm)
            mx, my = self._tp[0][self.Nx // 2], self._tp[1][self.Ny // 2]
            mua_p = self.plume.mu_a_plume
            tau = float(np.sum(mua_p[mx, my, :]) * self.plume.dz)
            trans = float(np.exp(-tau))
            peak = float(np.max(mua_p))
            mean = float(np.mean(self._from_plume(mua_p)[self.ablated])) if np.any(self.ablated) else 0.0
        else:
            tau, trans, peak, mean = 0.0, 1.0, 0.0, 0.0
        return n_new, ablated_J, tau, trans, peak, mean

    def _epoch(self, Q_th, absorbed, closure, t_opt):
        absorbed_J = float(np.sum(Q_th)) * self.voxel_vol * self.dt_epoch
        t1 = time.perf_counter()
        conducted_J = self.run_substeps(Q_th)
        t_th = time.perf_counter() - t1
        n_new, ablated_J, tau, trans, peak, mean = self.finish_epoch()
        self.ledger_J = {'absorbed': absorbed_J, 'conducted': conducted_J, 'ablated': ablated_J}
        return {'epoch': self.epoch, 'time_s': self.t, 'n_new': n_new, 'absorbed_frac': absorbed,
                'mc_energy_closure': closure, 't_optical_s': t_opt, 't_thermal_s': t_th,
                'plume_transmission': trans, 'plume_tau': tau, 'plume_peak_mu_a': peak, 'plume_mean_mu_a': mean,
                'ledger_J': dict(self.ledger_J)}

    def step_epoch(self):
        """coupled_ablation.py:440-617."""
        return self._epoch(*self._optics())

    def step_epoch_with_q(self, Q):
        """The epoch from the sub-steps on (coupled_ablation.py:476-617) with a given
        absorbed-power field Q [W/mm^3] in place of the transport: Q is zeroed in the
        void for the sub-steps, as after the MC.  ``absorbed_frac`` is the absorbed
        power over the beam power and the closure is NaN."""
        self.Q = np.array(Q, dtype=np.float64, copy=True)
        self.Q_opt = self.Q
        self.Phi = np.zeros_like(self.T)
        Q_th = self.Q.copy()
        if np.any(self.ablated):
            Q_th[self.ablated] = 0.0
        absorbed = float(np.sum(Q_th)) * self.voxel_vol / self.power if self.power > 0.0 else 0.0
        return self._epoch(Q_th, absorbed, float('nan'), 0.0)

    def state_energy_J(self):
        """Heat held above ambient by the intact cells: sum of V (rho_cp (T - T_amb) + H_acc)."""
        live = ~self.ablated
        return float(np.sum(self.rho_cp_grid[live] * (self.T[live] - self.t_ambient) + self.H_acc[live])) \
            * self.voxel_vol

    def load_state(self, **fields):
        """Overwrite state arrays in place (e.g. with a device engine's FP32 state as FP64),
        so one epoch of this engine can be compared from an identical start.  Keys: T,
        H_acc, chi, Omega, T_peak, ablation_time, k_grid, rho_cp_grid, h_lat, dry, ablated,
        mu_a, mu_s, g, n, wet (3-tuple), plume_raw, plume_ext."""
        for name, v in fields.items():
            if name in ('dry', 'ablated'):
                getattr(self, name)[...] = np.asarray(v, dtype=np.bool_)
            elif name in ('mu_a', 'mu_s', 'g', 'n'):
                self.optical[name][...] = np.asarray(v, dtype=np.float64)
            elif name == 'wet':
                self._wet = tuple(np.array(a, dtype=np.float64, copy=True) for a in v)
            elif name == 'plume_raw':
                self.plume.raw_field[...] = np.asarray(v, dtype=np.float32)
            elif name == 'plume_ext':
                self.plume.mu_ext_plume[...] = np.asarray(v, dtype=np.float32)
            elif name in ('T', 'H_acc', 'chi', 'Omega', 'T_peak', 'ablation_time', 'k_grid', 'rho_cp_grid',
                          'h_lat'):
                getattr(self, name)[...] = np.asarray(v, dtype=np.float64)
            else:
                raise KeyError(name)

    def centerline_depth_mm(self, ix=None, iy=None, k_surface=0):
        """coupled_ablation.py:662-667."""
        ix = self.Nx // 2 if ix is None else ix
        iy = self.Ny // 2 if iy is None else iy
        idx = np.where(self.ablated[ix, iy, k_surface:])[0]
        return float((idx.max() + 1) * self.dz) if idx.size else 0.0


def compute_phase_map(T, chi, dry, ablated, H_acc, is_air_gap=None):
    """run_multispectral_matrix.py:168-192."""
    phase = np.zeros(T.shape, dtype=np.uint8)
    if is_air_gap is not None:
        boiling = ((T >= 99.5) | (H_acc > 0.0)) & (~is_air_gap)
        charred = dry & (chi >= 0.10) & (~is_air_gap)
    else:
        boiling = (T >= 99.5) | (H_acc > 0.0)
        charred = dry & (chi >= 0.10)
    phase[boiling] = 2
    phase[charred] = 1
    phase[ablated] = 3
    if is_air_gap is not None:
        phase[is_air_gap] = 3
    return phase
