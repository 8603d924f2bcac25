"""3-D voxel Monte Carlo photon transport on the MI355X.

Drop-in for ``simulation/voxel_mc_3d.py`` of the reference: same entry point
(``run_voxel_mc3d``, signature :425-445), same output dictionary (:690-717),
same physics (:33-337), executed by the gfx950 transport kernel of
libbugdef.so.  There is one code path: ``engine='auto'`` and ``'gpu'`` run it,
``'cpu'`` raises (this build has no CPU engine; the FP64 CPU restatement lives
in the test oracle only).
"""

from __future__ import annotations

import math
import time
import warnings

import numpy as np

from . import _bdlib as B

MC3D_SCORING_MEMORY_BUDGET_BYTES = 512 * 1024 ** 2
MC3D_ENERGY_CLOSURE_TOL = 0.01
_MAX_EVENTS_PER_PHOTON = 10_000_000
BEAM_PROFILES = {'gaussian': 0, 'tophat': 1, 'kim_exp_disk': 2, 'kim_gauss_disk': 3}


def check_engine(engine: str) -> None:
    """voxel_mc_3d.py:556-576 engine keyword, restricted to the one device path."""
    e = str(engine).strip().lower()
    if e in ('auto', 'gpu'):
        if not B.is_gpu_available():
            raise RuntimeError("MI355X GPU engine requested but no HIP device is visible")
        return
    if e == 'cpu':
        raise ValueError("engine='cpu': this build has only the gfx950 engine (no CPU fallback)")
    raise ValueError(f"Unknown engine='{engine}'. Valid options: 'auto', 'gpu', 'cpu'.")


def layer_grids(Nx, Ny, Nz, Lx, Ly, Lz, layers_config, launch_axis_int):
    """voxel_mc_3d.py:517-548: planar stack voxelised by cell centres."""
    dx, dy, dz = Lx / Nx, Ly / Ny, Lz / Nz
    mu_a = np.zeros((Nx, Ny, Nz))
    mu_s = np.zeros((Nx, Ny, Nz))
    g = np.zeros((Nx, Ny, Nz))
    n = np.ones((Nx, Ny, Nz))
    if layers_config is not None:
        grid = (np.linspace(dz / 2.0, Lz - dz / 2.0, Nz) if launch_axis_int == 0
                else np.linspace(dx / 2.0, Lx - dx / 2.0, Nx))
        start = 0.0
        for lay in layers_config:
            m = (grid >= start) & (grid < lay['z_end'])
            if np.any(m):
                sl = (slice(None), slice(None), m) if launch_axis_int == 0 else (m,)
                mu_a[sl] = lay['mu_a']
                mu_s[sl] = lay['mu_s']
                g[sl] = lay['g']
                n[sl] = lay['n']
            start = lay['z_end']
    return mu_a, mu_s, g, n


def make_beam(power, n_photons, spot_dia, theta_inc_deg, phi_inc_deg, centre, r_max, launch_axis_int,
              beam_profile, seed, score_fluence=True, add_plume=False, photon_offset=0,
              max_events=_MAX_EVENTS_PER_PHOTON) -> B.Beam:
    if beam_profile not in BEAM_PROFILES:
        raise KeyError(beam_profile)
    b = B.Beam()
    b.power = float(power)
    b.spot_dia = float(spot_dia)
    b.theta_rad = math.radians(theta_inc_deg)
    b.phi_rad = math.radians(phi_inc_deg)
    b.x_center, b.y_center, b.z_center = (float(v) for v in centre)
    b.n_ambient = 1.0
    b.r_max = float(r_max)
    b.beam_profile = BEAM_PROFILES[beam_profile]
    b.launch_axis = int(launch_axis_int)
    b.n_photons = int(n_photons)
    b.photon_offset = int(photon_offset)
    b.max_events = int(max_events)
    b.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    b.score_fluence = 1 if score_fluence else 0
    b.add_plume = 1 if add_plume else 0
    return b


def beam_centre(Lx, Ly, Lz, launch_axis_int, beam_center):
    """voxel_mc_3d.py:599-610."""
    if launch_axis_int == 0:
        c = [Lx / 2.0, Ly / 2.0, 0.0]
        if beam_center is not None:
            c[0], c[1] = float(beam_center[0]), float(beam_center[1])
    else:
        c = [0.0, Ly / 2.0, Lz / 2.0]
        if beam_center is not None:
            c[1], c[2] = float(beam_center[0]), float(beam_center[1])
    return c


def spectro_dict(res: B.MCResult, n_photons: int, A_sum=None):
    """voxel_mc_3d.py:662-680,700-712."""
    N = float(n_photons)
    R_spec, R_diff = res.specular / N, res.diffuse / N
    A_tot = (res.absorbed if A_sum is None else A_sum) / N
    T_tot, Lat = res.transmitted / N, res.lateral / N
    closure = R_spec + R_diff + A_tot + T_tot + Lat
    if abs(closure - 1.0) > MC3D_ENERGY_CLOSURE_TOL:
        warnings.warn(f"Voxel MC energy closure {closure:.5f} deviates from 1 by more than "
                      f"{MC3D_ENERGY_CLOSURE_TOL:.3f} (R={R_spec + R_diff:.4f}, A={A_tot:.4f}, T={T_tot:.4f}, "
                      f"Lat={Lat:.4f})", RuntimeWarning)
    spectro = {'R_spec': round(R_spec * 100.0, 2), 'R_diff': round(R_diff * 100.0, 2),
               'R_tot': round((R_spec + R_diff) * 100.0, 2), 'Absorption': round(A_tot * 100.0, 2),
               'Transmission': round(T_tot * 100.0, 2), 'Lateral_loss': round(Lat * 100.0, 2)}
    raw = {'R_spec': R_spec, 'R_diff': R_diff, 'Absorption': A_tot, 'Transmission': T_tot, 'Lateral_loss': Lat}
    return spectro, raw, float(closure)


def run_voxel_mc3d(
    power: float = 1.0,
    pulse_duration: float = 0.20,
    n_photons: int = 200000,
    spot_dia: float = 0.50,
    theta_inc_deg: float = 0.0,
    phi_inc_deg: float = 0.0,
    Lx: float = 4.0,
    Ly: float = 4.0,
    Lz: float = 2.42,
    Nx: int = 80,
    Ny: int = 80,
    Nz: int = 60,
    r_max: float = 0.0,
    layers_config: list = None,
    custom_3d_props: dict = None,
    launch_axis: str = 'z',
    beam_center=None,
    beam_profile: str = 'gaussian',
    engine: str = 'auto',
    seed: int = None,
    score_fluence: bool = True,
    n_ambient: float = 1.0,
    max_events: int = _MAX_EVENTS_PER_PHOTON,
):
    """Voxel MC of voxel_mc_3d.py:425-717 on the gfx950 transport kernel.

    Extra keywords (not in the reference): ``seed`` keys the Philox streams
    (random when None, like the reference's unseeded generators) and
    ``score_fluence=False`` skips the path-length tally (``Phi_3d`` is then zero);
    ``n_ambient`` is the index outside the domain and ``max_events`` the per-photon event
    budget (voxel_mc_3d.py:30; the Metal entry point passes its ``max_steps``).
    """
    check_engine(engine)
    dx, dy, dz = Lx / Nx, Ly / Ny, Lz / Nz
    la = 1 if launch_axis.strip().lower() == 'x' else 0
    if custom_3d_props is not None:
        grids = [custom_3d_props[k] for k in ('mu_a', 'mu_s', 'g', 'n')]
    else:
        grids = layer_grids(Nx, Ny, Nz, Lx, Ly, Lz, layers_config, la)
    if seed is None:
        seed = int(np.random.randint(1, 2 ** 62))
    ctx = B.scratch_context((Nx, Ny, Nz), (dx, dy, dz))
    for name, arr in zip(('MU_A', 'MU_S', 'G', 'N'), grids):
        a = np.asarray(arr)
        if a.shape != (Nx, Ny, Nz):
            raise ValueError(f"custom_3d_props grid shape {a.shape} != {(Nx, Ny, Nz)}")
        ctx.upload(name, a)
    centre = beam_centre(Lx, Ly, Lz, la, beam_center)
    beam = make_beam(power, n_photons, spot_dia, theta_inc_deg, phi_inc_deg, centre, r_max, la, beam_profile, seed,
                     score_fluence=score_fluence, max_events=max_events)
    beam.n_ambient = float(n_ambient)
    t0 = time.perf_counter()
    res = ctx.mc_run(beam)
    elapsed = time.perf_counter() - t0
    N = float(n_photons)
    A = ctx.download('A') / N
    L = ctx.download('L') / N if score_fluence else np.zeros((Nx, Ny, Nz))
    lat = ctx.mc_lateral(Nz if la == 0 else Nx) / N
    spectro, raw, closure = spectro_dict(res, n_photons)
    vol = dx * dy * dz
    return {
        'grid_x': np.linspace(dx / 2.0, Lx - dx / 2.0, Nx),
        'grid_y': np.linspace(dy / 2.0, Ly - dy / 2.0, Ny),
        'grid_z': np.linspace(dz / 2.0, Lz - dz / 2.0, Nz),
        'dx': dx, 'dy': dy, 'dz': dz, 'Nx': Nx, 'Ny': Ny, 'Nz': Nz,
        'A_3d': A,
        'lateral_escape_z': lat,
        'Q_3d': A * (power / vol),
        'Phi_3d': L / vol,
        'spectro': spectro,
        'energy_closure': closure,
        'spectro_raw': raw,
        'launch_axis': 'x' if la else 'z',
        'elapsed_seconds': round(elapsed, 3),
        'photons_per_sec': int(n_photons / max(elapsed, 1e-4)),
        'kernel_ms': res.kernel_ms,
        'events': int(res.events),
        'engine': 'gpu',
    }
