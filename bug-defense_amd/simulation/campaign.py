"""A wavelength x power campaign of coupled runs on one GPU.

24 runs: 8 wavelengths (320, 450, 532, 980, 1064, 1470, 1940, 10600 nm) x
powers (20, 40, 80 W) on grids cycled over 128x128x96, 192x192x144 and
256x256x192 voxels of 15 um.  Each run is the 7-layer stack under 21 voxels of
air (the air gap as layer 1), with per-layer optics and heat capacity evaluated
at the run's wavelength from a seeded chromophore mix (``tissue_stack``):
cuticle-like layers mix chitin, protein, water, melanin and ommochrome, soft
layers water, protein and lipid; reduced scattering a' U[20, 130] /cm at
500 nm, g500 U[0.85, 0.95].  The campaign is run the way the reference's
matrix runner sweeps wavelengths and powers (run_multispectral_matrix.py), but
with all runs resident at once.  Two ways to run a round (one epoch of every
run):
- concurrent (default): every epoch is a captured HIP graph per run, and each
  round launches all 24 graphs (one stream per run) before waiting for any of
  them, so small runs fill the gaps of large ones;
- gang: the photon transport of all 24 runs is one kernel launch
  (bd_epoch_gang_launch), whose blocks each trace photons of one run, so the
  long-photon tails of the strongly scattering runs fill with other runs'
  blocks; every other step of an epoch runs on the run's own stream.  The
  launch is a barrier between the rounds' transport and their sub-steps, which
  costs more than the fuller transport gains on this run set (measured:
  130.5 simulated ms/s against 144.3 concurrent, 71.4 sequential).

    python -m simulation.campaign --epochs 200 [--mode concurrent|gang|sequential]
"""

from __future__ import annotations

import argparse
import itertools
import json
import time

import numpy as np

from .tissue_stack import eval_single_layer
from .workloads import AIR_LAYER, STACK_UM, stack_engine_kwargs

WAVELENGTHS = (320.0, 450.0, 532.0, 980.0, 1064.0, 1470.0, 1940.0, 10600.0)
POWERS = (20.0, 40.0, 80.0)
GRIDS = ((128, 128, 96), (192, 192, 144), (256, 256, 192))
CUTICLE = ('chitin', 'protein', 'water', 'melanin', 'ommochrome')
SOFT = ('water', 'protein', 'lipid')
CUTICLE_ALPHA = (40.0, 33.0, 18.0, 7.5, 1.2)     # Dirichlet concentration, around the exocuticle mix
SOFT_ALPHA = (70.0, 20.0, 10.0)


def chromophore_layers(rng, wavelength_nm):
    """Per-layer property tuples (stack_engine_kwargs format) of one seeded mix."""
    layers = []
    for i in range(len(STACK_UM)):
        if i == AIR_LAYER:
            layers.append(None)
            continue
        cuticle = i in (0, 2)
        names, alpha = (CUTICLE, CUTICLE_ALPHA) if cuticle else (SOFT, SOFT_ALPHA)
        frac = rng.dirichlet(alpha)
        comp = {nme: float(f) for nme, f in zip(names, frac)}
        ld = {'name': f'layer{i}', 'thickness_um': STACK_UM[i], 'composition': comp,
              'a_prime': float(rng.uniform(20.0, 130.0)), 'f_rayleigh': 0.2, 'b_mie': 1.1,
              'g500': float(rng.uniform(0.85, 0.95))}
        ev = eval_single_layer(ld, wavelength_nm)
        w = comp.get('water', 0.0)
        k = (0.6 * w + 0.2 * (1.0 - w)) * 1e-3          # W/(mm K), water-weighted
        layers.append((ev['mu_a'], ev['mu_s'], ev['g'], ev['n'], k, ev['rho_cp'], w))
    return layers


def campaign_runs(seed=0, n_photons=150_000, voxel_mm=0.015, plume_grid=None, grids=GRIDS, wavelengths=WAVELENGTHS,
                  powers=POWERS):
    """Engine keyword sets of the runs, in (wavelength, power) order."""
    runs = []
    for i, (wl, P) in enumerate(itertools.product(wavelengths, powers)):
        rng = np.random.default_rng([seed, i])
        shape = grids[i % len(grids)]
        pg = plume_grid or tuple(max(32, s // 2) for s in shape)
        kw = stack_engine_kwargs(shape, voxel_mm, 21, chromophore_layers(rng, wl), P, 22.0, 0.5, n_photons, 1e-3,
                                 5e-5, pg, seed * 100 + i, wavelength_nm=wl)
        runs.append(kw)
    return runs


def run_campaign(runs, epochs, concurrent=True, progress=None, mode=None):
    """Run every engine for `epochs` epochs, mode "concurrent" (graph epochs on
    one stream per run), "batch" (every run's epoch graph a child of one parent
    graph, one launch per round), "gang" (one transport launch for all runs) or
    "sequential" (one run after another; the default mode follows `concurrent`).  Returns the wall time of the epoch loop,
    each run's final diagnostics and the voxel-epochs done."""
    from .coupled_ablation import CoupledAblationEngine
    mode = mode or ("concurrent" if concurrent else "sequential")
    if mode not in ("gang", "batch", "concurrent", "sequential"):
        raise ValueError(f"unknown campaign mode {mode!r}")
    engines = [CoupledAblationEngine(graph=True, **kw) for kw in runs]
    try:
        # first epoch of each run outside the timed loop: graph capture and tile autotuning
        for e in engines:
            e.step_epoch()
        for e in engines:
            e._ctx.sync()
        t0 = time.perf_counter()
        last = [None] * len(engines)
        if mode in ("gang", "batch"):
            step = CoupledAblationEngine.step_epochs_gang if mode == "gang" else CoupledAblationEngine.step_epochs_batch
            for ep in range(epochs):
                last = step(engines)
                if progress and (ep + 1) % progress == 0:
                    print(f"campaign: epoch {ep + 1}/{epochs}", flush=True)
        elif mode == "concurrent":
            for ep in range(epochs):
                for e in engines:
                    e.step_epoch_launch()
                for i, e in enumerate(engines):
                    last[i] = e.step_epoch_finish()
                if progress and (ep + 1) % progress == 0:
                    print(f"campaign: epoch {ep + 1}/{epochs}", flush=True)
        else:
            for i, e in enumerate(engines):
                for _ in range(epochs):
                    last[i] = e.step_epoch()
        wall = time.perf_counter() - t0
        summary = [{'wavelength_nm': kw['wavelength_nm'], 'power_W': kw['power'],
                    'grid': [int(round(kw['Lx'] / 0.015)), int(round(kw['Ly'] / 0.015)), int(round(kw['Lz'] / 0.015))],
                    'absorbed_frac': round(d['absorbed_frac'], 4), 'removed_total': int(e.ablated.sum())}
                   for kw, e, d in zip(runs, engines, last)]
        voxel_epochs = sum(int(e.Nx * e.Ny * e.Nz) for e in engines) * epochs
        return wall, summary, voxel_epochs
    finally:
        for e in engines:
            e.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--photons", type=int, default=150_000)
    ap.add_argument("--mode", default="concurrent", choices=("concurrent", "batch", "gang", "sequential"))
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    runs = campaign_runs(seed=a.seed, n_photons=a.photons)
    wall, summary, vox_ep = run_campaign(runs, a.epochs, progress=25, mode=a.mode)
    sim_ms = len(runs) * a.epochs * 1.0
    print(json.dumps({"runs": len(runs), "epochs": a.epochs, "mode": a.mode,
                      "wall_s": round(wall, 3), "simulated_ms_per_s": round(sim_ms / wall, 2),
                      "voxel_epochs_per_s": round(vox_ep / wall, 1), "summary": summary}), flush=True)


if __name__ == "__main__":
    main()
