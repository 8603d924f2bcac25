"""The reference's high-resolution beetle runs on this engine, reduced to the
campaign KPIs.

``build_beetle_model`` and ``run_single_highres`` follow
knowledge/beetle/scripts/run_highres_knowledge_suite.py:57-111 (grid, tilted
beam, pulse and plume constants), :126-185 (the 11 anatomical slabs under 21
slices of ambient air, optics from ``tissue_stack.eval_single_layer`` at the
wavelength) and :187-329 (the engine keywords and the per-epoch record).  The
record is kept by ``run_record.RunRecorder`` from device-side x-z slices
instead of six full host fields per epoch, and reduced by
``run_record.analyze_npz`` (compute_suite_metrics.py:28-118).

The KPIs read the state at the end of the 200 ms pulse, so a run records the
pulse and then only follows the plume transmission through the cooling epochs
(for ``min_trans``, the run's minimum centre-column transmission).

The reference's recorded outcomes of these runs (``kpi_cache.json``) come from
its Metal solver (the runner asks for ``mc_engine='gpu', thermal_engine='gpu'``
on macOS).  That solver takes the plume diffusion step
min(dt, 0.45 / (2 d_rate)) (metal_ablation_solver.py:234-236), where the CPU
scheme's step is unstable on this 15 um / 50 us grid (centre coefficient
-0.78).  The runs here take the same clamped step (``plume_closure['dt_clamp']``)
unless ``cpu_plume`` is set.

``run_dual`` and ``run_no_plume`` follow run_dual_permutation_suite.py:157-360
on the same beetle: 30 ms at the first wavelength and 40 W, a switch of the
intact voxels' optics to the second wavelength at 20 W until 200 ms (the
in-place ``optical[...]`` / ``dry_state[...]`` edits, ``_wet``,
``_ds_arrays = None`` and the plume's ``set_wavelength`` / ``sync_extinction``),
then cooling; or one wavelength at 40 W with the plume disabled.

    python -m simulation.beetle_suite --runs 532:40,1064:40 [--golden FILE] [--out DIR]
    python -m simulation.beetle_suite --runs dual:320:450,noplume:532 --golden FILE
    python -m simulation.beetle_suite --all --golden tests/golden/kpi_highres_reference.json
"""

from __future__ import annotations

import argparse
import json
import math
import os
import tempfile
import time

import numpy as np

from .coupled_ablation import CoupledAblationEngine
from .phase_map import CHAR_STATE, H_LAT_WATER
from .run_record import RunRecorder, analyze_npz
from .tissue_stack import BASE_LAYERS, eval_single_layer

# run_highres_knowledge_suite.py:57-111
DX = DY = DZ = 0.015
LX = LY = 3.0
NX, NY = 200, 200
NZ_AIR_TOP, NZ_BEETLE = 21, 163
NZ = NZ_AIR_TOP + NZ_BEETLE
LZ = NZ * DZ
L_AIR_TOP = NZ_AIR_TOP * DZ
X_DORSAL, X_VENTRAL = 1.000, 2.000
TAN_THETA = (X_VENTRAL - X_DORSAL) / (LZ - L_AIR_TOP)
THETA_INC_DEG = math.degrees(math.atan(TAN_THETA))
X_TOP = X_DORSAL - L_AIR_TOP * TAN_THETA
BEAM_CENTER = (X_TOP, LY / 2.0)
SPOT_DIA = 0.50
T_EXPOSURE, T_PULSE, DT_EPOCH, DT_SUB = 1.000, 0.200, 0.001, 1.0 / 20000.0
N_PHOTONS = 150_000
T_AMBIENT, T_BOIL, T_ABL_CUTICLE, H_EFF = 20.0, 100.0, 350.0, 15.0e-6
PLUME_CLOSURE = {'model': 'diffusion', 'D_xy': 1.0, 'D_z': 2.0, 'v_drift': 0.0, 'tau_linger_s': float('inf'),
                 'blast_expansion': False, 'blast_sigma_xy_mm': 0.080, 'blast_sigma_z_mm': 0.090,
                 'blast_z_bias_mm': 0.0}
WAVELENGTHS = (320, 450, 532, 1064, 1470, 1940, 10600)
POWERS = (20.0, 40.0, 80.0)
AIR_GAP = (NZ_AIR_TOP + 11, NZ_AIR_TOP + 31)
# run_dual_permutation_suite.py:77-96
T_PHASE1 = 0.030
DUAL_CHAR_STATE = {'A': 1.2e4, 'Ea': 5.8e4, 'mu_a': 15.0, 'mu_s': 1.0, 'g': 0.90, 't_remove': T_ABL_CUTICLE,
                   'chi_remove': 0.95}
DUAL_FIRST = (320, 1940, 10600)
DUAL_SECOND = (450, 532, 1064)

# (first slice, end slice below the dorsal surface, BASE_LAYERS index, k [W/(mm K)], rho c [J/(mm^3 K)],
# h_lat [J/mm^3], dry) -- run_highres_knowledge_suite.py:147-159
BEETLE_SLABS = (
    (0, 1, 0, 0.22e-3, 1.98e-3, 0.0, True),
    (1, 5, 1, 0.28e-3, 2.40e-3, 0.0, True),
    (5, 11, 2, 0.35e-3, 2.50e-3, 0.0, True),
    (11, 31, 3, 0.026e-3, 1.206e-6, 1e12, False),
    (31, 32, 4, 0.48e-3, 3.50e-3, H_LAT_WATER, False),
    (32, 52, 5, 0.32e-3, 2.80e-3, H_LAT_WATER, False),
    (52, 69, 6, 0.58e-3, 3.89e-3, H_LAT_WATER, False),
    (69, 137, 7, 0.50e-3, 3.65e-3, H_LAT_WATER, False),
    (137, 157, 8, 0.32e-3, 2.80e-3, H_LAT_WATER, False),
    (157, 160, 10, 0.35e-3, 2.50e-3, 0.0, True),
    (160, 163, 11, 0.28e-3, 2.40e-3, 0.0, True),
)


def build_beetle_model(wl_nm: float) -> dict:
    """run_highres_knowledge_suite.py:126-185: property grids of the beetle at wl_nm."""
    shape = (NX, NY, NZ)
    mu_a, mu_s, g = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    n = np.ones(shape)
    k_grid = np.full(shape, 0.026e-3)
    rho_cp = np.full(shape, 1.206e-6)
    h_lat = np.full(shape, 1e12)
    dry = np.zeros(shape, bool)
    ablated = np.zeros(shape, bool)
    ablated[:, :, :NZ_AIR_TOP] = True
    ev = [eval_single_layer(lay, float(wl_nm)) for lay in BASE_LAYERS]
    for k0, k1, li, kv, rc, hl, is_dry in BEETLE_SLABS:
        s = np.s_[:, :, NZ_AIR_TOP + k0:NZ_AIR_TOP + k1]
        mu_a[s], mu_s[s], g[s], n[s] = (float(ev[li][key]) for key in ('mu_a', 'mu_s', 'g', 'n'))
        k_grid[s], rho_cp[s], h_lat[s], dry[s] = kv, rc, hl, is_dry
    optical = {'mu_a': mu_a, 'mu_s': mu_s, 'g': g, 'n': n}
    dry_state = {'mu_a': mu_a.copy(), 'mu_s': mu_s.copy(), 'g': g.copy(), 'n': n.copy(), 'k': k_grid.copy(),
                 'rho_cp': rho_cp.copy(), 't_remove': T_ABL_CUTICLE}
    return {'optical': optical, 'k_grid': k_grid, 'rho_cp_grid': rho_cp, 'h_lat_grid': h_lat, 'dry_mask': dry,
            'initial_ablated': ablated, 'dry_state': dry_state}


def _engine(geo, wl_nm, power_w, char_state, plume, n_photons, seed):
    return CoupledAblationEngine(
        optical_props=geo['optical'], k_grid=geo['k_grid'], rho_cp_grid=geo['rho_cp_grid'],
        h_lat_grid=geo['h_lat_grid'], Lx=LX, Ly=LY, Lz=LZ, power=float(power_w), spot_dia=SPOT_DIA, t_abl=T_BOIL,
        t_ambient=T_AMBIENT, dt_epoch=DT_EPOCH, dt_sub=DT_SUB, n_photons=int(n_photons), h_eff=H_EFF,
        beam_center=BEAM_CENTER, beam_profile='gaussian', theta_inc_deg=THETA_INC_DEG, phi_inc_deg=0.0,
        dry_state=geo['dry_state'], char_state=char_state, plume_closure=plume,
        initial_dry_mask=geo['dry_mask'], initial_ablated_mask=geo['initial_ablated'], air_gap_range=AIR_GAP,
        wavelength_nm=float(wl_nm), mc_engine='gpu', thermal_engine='gpu', seed=seed)


def _run(eng, out_path, schedule, cooling, meta):
    """The epoch loop of the runners: schedule(ep) sets the laser before each epoch
    and returns its state label; the pulse is recorded, the transmission followed
    to the end.  Returns the KPIs with min_trans and the loop's wall time."""
    ix_beam = np.clip(np.round((X_TOP + TAN_THETA * (np.arange(NZ) * DZ)) / DX).astype(int), 0, NX - 1)
    regions = {'dorsal': (NZ_AIR_TOP, AIR_GAP[0]), 'core': (NZ_AIR_TOP + 52, NZ_AIR_TOP + 69),
               'ventral': (NZ_AIR_TOP + 157, NZ_AIR_TOP + 163)}
    rec = RunRecorder(eng, nz_air_top=NZ_AIR_TOP, air_gap=AIR_GAP, regions=regions, ix_beam=ix_beam)
    n_pulse = int(round(T_PULSE / DT_EPOCH))
    n_epochs = int(round(T_EXPOSURE / DT_EPOCH)) if cooling else n_pulse
    trans, states = [], []
    t0 = time.perf_counter()
    try:
        for ep in range(n_epochs):
            label = schedule(ep)
            d = eng.step_epoch()
            trans.append(float(d['plume_transmission']))
            if ep < n_pulse:
                rec.record(d)
                states.append(label)
        wall = time.perf_counter() - t0
        if meta.pop('laser_states', False):
            meta['laser_states'] = np.array(states)
        rec.save(out_path, **meta)
    finally:
        eng.close()
    k = analyze_npz(out_path, pulse_duration_s=T_PULSE)
    k['min_trans'] = 100.0 * min(trans)
    k['wall_s'] = wall
    return k


def run_single_highres(wl_nm: int, power_w: float, out_path: str, n_photons: int = N_PHOTONS, seed=None,
                       cooling=True, cpu_plume=False) -> dict:
    """run_highres_knowledge_suite.py:187-329 on this engine; returns the KPIs of
    the recorded pulse (analyze_npz) with min_trans over the whole exposure."""
    eng = _engine(build_beetle_model(wl_nm), wl_nm, power_w, CHAR_STATE, dict(PLUME_CLOSURE, dt_clamp=not cpu_plume),
                  n_photons, seed)

    def schedule(ep):
        if ep * DT_EPOCH >= T_PULSE:
            eng.power = 0.0
        return ""
    return _run(eng, out_path, schedule, cooling, {'wavelength_nm': float(wl_nm), 'power_w': float(power_w)})


def _dual_plume(enabled, cpu_plume):
    # run_dual_permutation_suite.py:189-197
    return {'enabled': enabled, 'model': 'diffusion', 'D_xy': 1.0, 'D_z': 2.0, 'v_drift': 0.0,
            'tau_linger_s': float('inf'), 'blast_expansion': False, 'dt_clamp': not cpu_plume}


def run_dual(wl1: int, wl2: int, out_path: str, p1: float = 40.0, p2: float = 20.0, n_photons: int = N_PHOTONS,
             seed=None, cooling=True, cpu_plume=False) -> dict:
    """run_dual_permutation_suite.py:157-360 (two wavelengths, plume on)."""
    geo2 = build_beetle_model(wl2)
    eng = _engine(build_beetle_model(wl1), wl1, p1, DUAL_CHAR_STATE, _dual_plume(True, cpu_plume), n_photons, seed)

    def schedule(ep):
        t = ep * DT_EPOCH
        if t < T_PHASE1:
            return f"{wl1} nm @ {int(p1)} W"
        if t < T_PULSE:
            if eng.wavelength_nm != float(wl2):   # :249-279, the switch of the intact voxels
                intact = ~np.asarray(eng.ablated)
                opt = eng.optical
                for key in ('mu_a', 'mu_s', 'g', 'n'):
                    opt[key][intact] = geo2['optical'][key][intact]
                if eng.dry_state is not None:
                    for key in ('mu_a', 'mu_s', 'g', 'n'):
                        eng.dry_state[key][intact] = geo2['dry_state'][key][intact]
                eng._wet = (eng.optical['mu_a'].copy(), np.array(eng.k_grid), np.array(eng.rho_cp_grid))
                eng._ds_arrays = None
                eng._init_ds_cs_arrays()
                if eng.plume is not None:
                    eng.plume.set_wavelength(wl2)
                    eng.plume.sync_extinction(eng.compute_void_mask())
                eng.wavelength_nm = float(wl2)
                eng.power = p2
            return f"{wl2} nm @ {int(p2)} W"
        eng.power = 0.0
        return "Laser OFF [Cooling]"
    return _run(eng, out_path, schedule, cooling, {'wavelength_nm': float(wl1), 'laser_states': True})


def run_no_plume(wl_nm: int, out_path: str, p1: float = 40.0, n_photons: int = N_PHOTONS, seed=None,
                 cooling=True) -> dict:
    """run_dual_permutation_suite.py:157-360 with --no-plume (one wavelength, plume off)."""
    eng = _engine(build_beetle_model(wl_nm), wl_nm, p1, DUAL_CHAR_STATE, _dual_plume(False, False), n_photons,
                  seed)

    def schedule(ep):
        if ep * DT_EPOCH < T_PULSE:
            return f"{wl_nm} nm @ {int(p1)} W [No Plume]"
        eng.power = 0.0
        return "Laser OFF [Cooling]"
    return _run(eng, out_path, schedule, cooling, {'wavelength_nm': float(wl_nm), 'power_w': float(p1),
                                                   'laser_states': True})


COMPARE = ("final_depth_um", "avg_depth_um", "avg_hole_dia_mm", "core_avg_T", "core_max_T", "ven_avg_T",
           "ven_max_T", "v_abl_mm3", "v_boil_mm3", "min_trans")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", default="532:40",
                    help="comma-separated runs: wl:power (single), dual:wl1:wl2, noplume:wl")
    ap.add_argument("--all", action="store_true",
                    help="the reference campaign: 21 single, 9 dual-wavelength and 7 no-plume runs")
    ap.add_argument("--golden", default=None, help="reference KPIs to compare with (JSON fixture)")
    ap.add_argument("--out", default=None, help="directory for the .npz records (default: a temporary one)")
    ap.add_argument("--photons", type=int, default=N_PHOTONS)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-plume", action="store_true",
                    help="the CPU scheme's plume step (unstable on this grid) instead of the Metal solver's")
    a = ap.parse_args(argv)
    if a.all:
        specs = ([f"{wl}:{int(p)}" for wl in WAVELENGTHS for p in POWERS] +
                 [f"dual:{w1}:{w2}" for w1 in DUAL_FIRST for w2 in DUAL_SECOND] +
                 [f"noplume:{wl}" for wl in WAVELENGTHS])
    else:
        specs = a.runs.split(",")
    ref = json.load(open(a.golden))["runs"] if a.golden else {}
    tmp = None if a.out else tempfile.TemporaryDirectory()
    out_dir = a.out or tmp.name
    os.makedirs(out_dir, exist_ok=True)
    try:
        for spec in specs:
            f = spec.split(":")
            if f[0] == "dual":
                key = f"dual_{f[1]}nm_{f[2]}nm"
                path = os.path.join(out_dir, f"matrix_{key}.npz")
                k = run_dual(int(f[1]), int(f[2]), path, n_photons=a.photons, seed=a.seed, cpu_plume=a.cpu_plume)
            elif f[0] == "noplume":
                key = f"{f[1]}nm_40w_no_plume"
                path = os.path.join(out_dir, f"matrix_{key}.npz")
                k = run_no_plume(int(f[1]), path, n_photons=a.photons, seed=a.seed)
            else:
                wl, p = int(f[0]), float(f[1])
                key = f"{wl}nm_{int(p)}w"
                path = os.path.join(out_dir, f"matrix_{key}.npz")
                k = run_single_highres(wl, p, path, n_photons=a.photons, seed=a.seed, cpu_plume=a.cpu_plume)
            if not a.out:
                os.remove(path)
            row = {"run": key, "wall_s": round(k["wall_s"], 2),
                   "kpis": {c: round(float(k[c]), 4) for c in COMPARE}}
            if key in ref:
                row["reference"] = {c: round(float(ref[key][c]), 4) for c in COMPARE}
            print(json.dumps(row), flush=True)
    finally:
        if tmp is not None:
            tmp.cleanup()


if __name__ == "__main__":
    main()
