"""The reference's scanned-beam liver benchmark on this engine.

``simulate_scan`` follows
knowledge/simulation_lit_comparison/code/validation_benchmarks.py:83-141: a
liver block (optics, conduction, Arrhenius and vaporisation constants of
:31-45, adiabatic boundaries, 22 C), a 2.8 mm Gaussian beam scanned along +y
from ``scan_margin`` to Ly - scan_margin with the beam centre moved every epoch,
optionally the damage-dependent scattering (DOP) of :116-121 before each epoch,
then ``t_off`` seconds of laser-off relaxation in 1 s epochs, and the x-z
cross-section at y = Ly / 2 reduced by ``cross_section_metrics`` (:54-80,
Shimojo's supplementary definitions).

The DOP step runs on the device (``CoupledAblationEngine.apply_damage_optics``)
unless ``host_dop`` is set, in which case the reference's own host loop body
(download Omega, update ``optical['mu_s']``, upload) runs instead; both give
the same mu_s.

The reference's recorded cross-sections of its six cases
(``raw_output/shimojo_cross_sections.json``, from
``validation_benchmarks.run_shimojo`` with 20 000 photons per 2 ms epoch and
DOP on: 50 W at 0.5, 0.75 and 1.0 mm/s on a 24 mm block with 30 s off, 10 W at
1.0, 1.25 and 1.5 mm/s on a 12 mm block with 10 s off) are the golden data.

    python -m simulation.scan_suite --cases 10:1.5,50:1.0 [--golden FILE]
    python -m simulation.scan_suite --all --golden tests/golden/shimojo_reference.json
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

LIVER = dict(mu_a=0.07, musp_native=0.51, musp_coag=1.85, g=0.9, n=1.44, k=5.43e-3 / 10.0,
             rho_cp=1.0e-3 * 3.542, A=7.39e37, Ea=2.577e5, Tv=100.0, Lv_vol=1.0e-3 * 2257.0, T0=22.0)
BEAM_DIA_MM = 2.8

# validation_benchmarks.py:160-175 and :145: the recorded cases
CASES = {(50.0, 0.5): dict(Lx=24.0, Ly=12.0, Lz=10.0, t_off=30.0),
         (50.0, 0.75): dict(Lx=24.0, Ly=12.0, Lz=10.0, t_off=30.0),
         (50.0, 1.0): dict(Lx=24.0, Ly=12.0, Lz=10.0, t_off=30.0),
         (10.0, 1.0): dict(Lx=12.0, Ly=12.0, Lz=10.0, t_off=10.0),
         (10.0, 1.25): dict(Lx=12.0, Ly=12.0, Lz=10.0, t_off=10.0),
         (10.0, 1.5): dict(Lx=12.0, Ly=12.0, Lz=10.0, t_off=10.0)}
RECORDED_DT_EPOCH = 0.002
RECORDED_PHOTONS = 20000


def case_key(power, speed) -> str:
    return f"{power:.0f}W_{speed}"


def _row_extent(mask_1d, d):
    idx = np.nonzero(mask_1d)[0]
    return float((idx.max() - idx.min() + 1) * d) if idx.size else 0.0


def cross_section_metrics(ablated_xz, coag_xz, dx, dz) -> dict:
    """validation_benchmarks.py:59-80: vaporisation depth and width, and the
    coagulation depth below the deepest crater column and the coagulation width
    one voxel above the crater floor (index [ix, iz])."""
    ablated_xz = np.asarray(ablated_xz, dtype=bool)
    coag_xz = np.asarray(coag_xz, dtype=bool)
    depth_of = lambda col: (np.flatnonzero(col).max() + 1) * dz if col.any() else 0.0  # noqa: E731
    depths = np.array([depth_of(ablated_xz[i]) for i in range(ablated_xz.shape[0])])
    out = {}
    if depths.max() > 0.0:
        out['vap_depth'] = float(depths.max())
        out['vap_width'] = max(_row_extent(ablated_xz[:, k], dx) for k in range(ablated_xz.shape[1]))
        deepest = np.flatnonzero(depths == depths.max())
        i_axis = int(deepest[len(deepest) // 2])
        k_floor = int(round(depths.max() / dz))
        below = coag_xz[i_axis, k_floor:]
        n_coag = below.size if below.all() else int(np.argmin(below))
        out['coag_depth'] = float(n_coag * dz)
        out['coag_width'] = _row_extent(coag_xz[:, k_floor - 1] | ablated_xz[:, k_floor - 1], dx)
    else:
        out['vap_depth'] = 0.0
        out['vap_width'] = 0.0
        out['coag_depth'] = float(max(depth_of(coag_xz[i]) for i in range(coag_xz.shape[0])))
        out['coag_width'] = max(_row_extent(coag_xz[:, k], dx) for k in range(coag_xz.shape[1]))
    return out


def build_engine(power, speed, n_photons=RECORDED_PHOTONS, voxel=0.1, Lx=24.0, Ly=12.0, Lz=10.0,
                 dt_sub=0.004, dt_epoch=None, seed=1, **engine_kw):
    """validation_benchmarks.py:92-111: the liver block and its engine."""
    from .coupled_ablation import CoupledAblationEngine
    Nx, Ny, Nz = int(round(Lx / voxel)), int(round(Ly / voxel)), int(round(Lz / voxel))
    shape = (Nx, Ny, Nz)
    optical = {'mu_a': np.full(shape, LIVER['mu_a']),
               'mu_s': np.full(shape, LIVER['musp_native'] / (1.0 - LIVER['g'])),
               'g': np.full(shape, LIVER['g']),
               'n': np.full(shape, LIVER['n'])}
    if dt_epoch is None:
        dt_epoch = min(0.1, voxel / speed)
    dt_sub = min(dt_sub, dt_epoch)
    return CoupledAblationEngine(
        optical_props=optical, k_grid=np.full(shape, LIVER['k']), rho_cp_grid=np.full(shape, LIVER['rho_cp']),
        h_lat_grid=np.full(shape, LIVER['Lv_vol']), Lx=Lx, Ly=Ly, Lz=Lz, power=power, spot_dia=BEAM_DIA_MM,
        t_abl=LIVER['Tv'], t_ambient=LIVER['T0'], dt_epoch=dt_epoch, dt_sub=dt_sub, n_photons=n_photons,
        h_eff=0.0, arrhenius=(LIVER['A'], LIVER['Ea']), seed=seed, **engine_kw), dt_epoch


def simulate_scan(power, speed, dop=True, n_photons=RECORDED_PHOTONS, voxel=0.1, Lx=24.0, Ly=12.0, Lz=10.0,
                  scan_margin=1.0, t_off=30.0, dt_sub=0.004, dt_epoch=None, seed=1, host_dop=False,
                  progress=None) -> dict:
    """validation_benchmarks.py:83-141 on the device engine.  Returns the metrics
    and the x-z masks ``xz_ablated`` / ``xz_coag`` at y = Ly / 2."""
    eng, dt_epoch = build_engine(power, speed, n_photons, voxel, Lx, Ly, Lz, dt_sub, dt_epoch, seed)
    g = LIVER['g']
    y0, y1 = scan_margin, Ly - scan_margin
    n_on = int(math.ceil((y1 - y0) / speed / dt_epoch))
    t_start = time.time()
    for e in range(n_on):
        eng.beam_center = (Lx / 2.0, y0 + speed * (e + 0.5) * dt_epoch)
        if dop:
            if host_dop:
                alpha = 1.0 - np.exp(-eng.Omega)
                musp = (1.0 - alpha) * LIVER['musp_native'] + alpha * LIVER['musp_coag']
                intact = ~eng.ablated
                eng.optical['mu_s'][intact] = musp[intact] / (1.0 - g)
            else:
                eng.apply_damage_optics(LIVER['musp_native'], LIVER['musp_coag'], g)
        eng.step_epoch()
        if progress is not None and (e + 1) % 1000 == 0:
            progress(f"  [{power:.0f} W, {speed} mm/s] epoch {e + 1}/{n_on}  ablated={int(eng.ablated.sum())}")
    eng.power = 0.0
    eng.dt_epoch = 1.0
    eng.n_substeps = int(round(1.0 / eng.dt_sub))
    for _ in range(int(round(t_off))):
        eng.step_epoch()
    iy = int(round((Ly / 2.0) / voxel - 0.5))
    abl = np.asarray(eng.ablated[:, iy, :], dtype=bool)
    coag = (np.asarray(eng.Omega[:, iy, :]) >= 1.0) & ~abl
    m = cross_section_metrics(abl, coag, eng.dx, eng.dz)
    m['runtime_s'] = time.time() - t_start
    m['epochs_on'] = n_on
    m['xz_ablated'] = abl
    m['xz_coag'] = coag
    return m


def run_case(power, speed, n_photons=RECORDED_PHOTONS, dt_epoch=RECORDED_DT_EPOCH, seed=1, host_dop=False,
             progress=None) -> dict:
    """One recorded case as ``validation_benchmarks.run_shimojo`` runs it (DOP on,
    dt_sub = dt_epoch = 2 ms)."""
    dom = CASES[(float(power), float(speed))]
    return simulate_scan(power, speed, dop=True, n_photons=n_photons, Lx=dom['Lx'], Ly=dom['Ly'], Lz=dom['Lz'],
                         t_off=dom['t_off'], dt_sub=dt_epoch, dt_epoch=dt_epoch, seed=seed, host_dop=host_dop,
                         progress=progress)


def unpack_mask(entry) -> np.ndarray:
    """A golden mask: np.packbits of the C-order bool array, as hex, with its shape."""
    bits = np.unpackbits(np.frombuffer(bytes.fromhex(entry['hex']), dtype=np.uint8))
    n = int(np.prod(entry['shape']))
    return bits[:n].astype(bool).reshape(entry['shape'])


def recorded_metrics(gold, voxel=0.1) -> dict:
    """All four metrics of a recorded case, from its recorded masks."""
    return cross_section_metrics(unpack_mask(gold['ablated']), unpack_mask(gold['coag']), voxel, voxel)


def compare(m, gold) -> dict:
    """Metric differences (mm) against the recorded case and the masks' intersection over union."""
    ref_m = recorded_metrics(gold)
    out = {k: m[k] - ref_m[k] for k in ('vap_depth', 'vap_width', 'coag_depth', 'coag_width')}
    for name, key in (('ablated', 'xz_ablated'), ('coag', 'xz_coag')):
        ref = unpack_mask(gold[name])
        mine = np.asarray(m[key], dtype=bool)
        union = int((ref | mine).sum())
        out[f'iou_{name}'] = 1.0 if union == 0 else int((ref & mine).sum()) / union
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--cases', default='10:1.5', help='comma list of power:speed')
    ap.add_argument('--all', action='store_true')
    ap.add_argument('--golden', default=None)
    ap.add_argument('--photons', type=int, default=RECORDED_PHOTONS)
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--host-dop', action='store_true')
    args = ap.parse_args(argv)
    cases = sorted(CASES) if args.all else [tuple(float(v) for v in c.split(':')) for c in args.cases.split(',')]
    gold = json.load(open(args.golden)) if args.golden else None
    say = lambda s: print(s, flush=True)  # noqa: E731
    for power, speed in cases:
        m = run_case(power, speed, n_photons=args.photons, seed=args.seed, host_dop=args.host_dop, progress=say)
        line = {'case': case_key(power, speed), **{k: round(m[k], 3) for k in
                ('vap_depth', 'vap_width', 'coag_depth', 'coag_width', 'runtime_s')}, 'epochs_on': m['epochs_on']}
        if gold is not None:
            line['vs_recorded'] = {k: round(v, 3) for k, v in compare(m, gold[case_key(power, speed)]).items()}
            line['recorded'] = {k: round(v, 3) for k, v in recorded_metrics(gold[case_key(power, speed)]).items()}
        say(json.dumps(line))
    return 0


if __name__ == '__main__':
    sys.exit(main())
