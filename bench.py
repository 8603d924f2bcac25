"""Benchmark: simulated milliseconds per wall second of the coupled epoch loop,
with the surrogate in the loop against the plain MC loop.

Workload: the coupled layered-stack run at 256x256x192 voxels of 15 um (seeded
7-layer stack under 21 voxels of ambient air, 40 W at 22 deg, 0.5 mm spot,
1 ms epochs of 150k photons + 20 sub-steps of 50 us, boil 100 degC, ablate
350 degC, Arrhenius char, diffusion plume on a 128^3 grid), synthetic inputs,
state resident in HBM before the timed region.  One step = one epoch = 1 ms
simulated.  Two runs from the same initial state and seeds: the plain MC loop,
and the surrogate loop (network on 9 of 10 epochs, MC judging every 10th and
taking over while the judged error exceeds 5%).  ``value`` is the surrogate
loop's rate when every judgement in its timed window passed the accuracy gate,
and the plain loop's rate otherwise (``value_source`` says which); both rates,
the judged errors, the per-stage cost of a surrogate epoch, and the crater
depth, ablated volume and peak temperature of both runs are reported beside it.
``roofline_transport`` (the transport kernel k_mc) and ``roofline_stencil``
(the fused sub-step) are both against HBM with their algorithmic bytes;
``roofline`` repeats the one whose kernel takes the larger share of the measured
epoch.  The transport entry also carries its tally additions per second beside
the device's measured atomic-add ceiling.
Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--plain-only]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bug-defense_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per voxel of one fused sub-step with char on, Arrhenius off:
# read T, k, rho_cp, Q, write T (20 B) + H r/w (8) + chi r/w (8) + dry (1) + ablated (1)
SUBSTEP_BYTES_PER_VOXEL = 38
# algorithmic bytes per transport event (one voxel step): the optical properties gathered
# (mu_a, mu_s: 8 B) and the path-length tally addition (4 B)
MC_BYTES_PER_EVENT = 12
WEIGHTS = os.path.join(ROOT, "bug-defense_amd", "surrogate", "weights", "fno_layered_15um.npz")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def reference_mc_chunks(n_vox, n_photons, threads):
    """The reference CPU MC's chunk count (voxel_mc_3d.py:621-625): private FP64 tallies of
    16 bytes per voxel within a 512 MiB budget, at least 5000 photons per chunk."""
    return max(1, min(threads, (512 * 1024 ** 2) // (16 * n_vox), n_photons // 5000))


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_usable": affinity,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(kw, epochs, threads):
    """The FP64 oracle engine on the host cores, same workload, a bounded number of epochs;
    the MC runs `threads` private-tally chunks, the stencil every OpenMP thread."""
    from oracle import oracle as O
    okw = dict(kw)
    okw.pop('plume_grid', None)
    eng = O.OracleEngine(mc_threads=threads, plume_grid=kw.get('plume_grid'), **okw)
    eng.step_epoch()  # first-touch of the arrays outside the timed span
    t0 = time.perf_counter()
    for _ in range(epochs):
        eng.step_epoch()
    dt = time.perf_counter() - t0
    return epochs * eng.dt_epoch * 1e3 / dt, dt


def final_state(eng, n_air_top):
    """Crater depth below the initial surface [mm], ablated volume [mm^3], peak T [degC]."""
    import numpy as np
    abl = np.array(eng.ablated, copy=True)
    abl[:, :, :n_air_top] = False
    v = float(abl.sum()) * eng.dx * eng.dy * eng.dz
    zs = np.nonzero(abl.any(axis=(0, 1)))[0]
    depth = float((zs.max() + 1 - n_air_top) * eng.dz) if zs.size else 0.0
    return {"crater_depth_mm": round(depth, 4), "ablated_volume_mm3": round(v, 5),
            "peak_T_C": round(float(eng.T_peak.max()), 2)}


def run_plain(kw, steps, warmup, graph):
    import numpy as np
    from simulation.coupled_ablation import CoupledAblationEngine
    eng = CoupledAblationEngine(graph=graph, **kw)
    for _ in range(warmup):
        eng.step_epoch()
    log(f"bench: plain loop, {warmup} warmup epochs done")
    eng._ctx.sync()
    rec = {k: [] for k in ("sub", "mc", "recp", "opt", "th", "ph")}
    events, atomics, records, nnew = 0.0, 0.0, 0.0, 0
    t0 = time.perf_counter()
    for i in range(steps):
        d = eng.step_epoch()
        ld = eng.last_diag
        rec["sub"].append(ld.t_substeps_ms)
        rec["mc"].append(ld.t_mc_kernel_ms)
        rec["recp"].append(ld.t_mc_records_ms)
        rec["opt"].append(ld.t_optical_ms)
        rec["th"].append(ld.t_thermal_ms)
        rec["ph"].append(ld.t_phase_ms)
        events += ld.mc_events
        atomics += ld.mc_tally_atomics
        records += ld.mc_tally_records
        nnew += d['n_new']
        if (i + 1) % 25 == 0:
            log(f"bench: plain epoch {i + 1}/{steps}")
    eng._ctx.sync()
    wall = time.perf_counter() - t0
    out = {"wall": wall, "events": events, "atomics": atomics, "records": records, "nnew": nnew, "nvox": eng.Nx * eng.Ny * eng.Nz, "tuning": eng._ctx.tuning(),
           "n_sub": eng.n_substeps, "dt_epoch": eng.dt_epoch, "n_photons": eng.n_photons,
           **{k: np.asarray(v) for k, v in rec.items()}, "state": final_state(eng, 21)}
    eng.close()
    return out


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_substep_traffic_layered_256x256x192.md")


def measured_traffic(tuning):
    """HBM bytes per one-step sub-step launch on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated) for the tile shape
    this run picked, or None when the summary has no line for that tile."""
    want = f"{tuning['substep_tile_y']}x{tuning['substep_tile_x']}"
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("- tile ") and "traffic_bytes_per_launch" in line:
                    tile = line[len("- tile "):].split(":")[0].strip()
                    if tile == want:
                        return int(float(line.split("traffic_bytes_per_launch")[1]))
    except (OSError, ValueError):
        return None
    return None


def run_surrogate(kw, steps, warmup, weights, mc_lo, correct_block):
    import numpy as np
    from simulation.coupled_ablation import CoupledAblationEngine
    from surrogate.coupled import SurrogateCoupledEngine, excess
    eng = CoupledAblationEngine(**kw)
    sur = SurrogateCoupledEngine(eng, weights, judge_every=10, tol=0.05, mc_photons_lo=mc_lo,
                                 correct_block=correct_block)
    for _ in range(warmup):
        sur.step_epoch()
    log(f"bench: surrogate loop, {warmup} warmup epochs done")
    eng._ctx.sync()
    t0 = time.perf_counter()
    for i in range(steps):
        sur.step_epoch()
        if (i + 1) % 25 == 0:
            log(f"bench: surrogate epoch {i + 1}/{steps}")
    eng._ctx.sync()
    wall = time.perf_counter() - t0
    hist = sur.history
    state_end = final_state(eng, 21)   # before the extra profiling epochs below
    timed_judged = [h for h in hist[warmup:] if h["judge_error"] is not None]
    gate_met = bool(timed_judged) and all(not h["fallback"] for h in timed_judged)
    # cost of one surrogate epoch as it runs (the low-photon transport beside the forward pass),
    # then stage by stage (host waits between stages), and of one plain MC epoch, on the state the
    # timed loop left; forced predictions, whatever the gate says
    reps = 5
    eng._ctx.sync()
    t1 = time.perf_counter()
    for _ in range(reps):
        sur._predict(eng._epoch_beam(), into_q=True)
        eng.step_epoch_with_q()
    eng._ctx.sync()
    sur_epoch_ms = (time.perf_counter() - t1) * 1e3 / reps
    sur.profile = True
    sur.stage_ms = {}
    for _ in range(reps):
        beam = eng._epoch_beam()
        sur._predict(beam, into_q=True)
        t1 = time.perf_counter()
        eng.step_epoch_with_q()
        sur._mark("epoch_without_transport", t1)
    stage_ms = {k: round(v / reps, 4) for k, v in sur.stage_ms.items()}
    eng._ctx.sync()
    t1 = time.perf_counter()
    for _ in range(reps):
        eng.step_epoch()
    eng._ctx.sync()
    plain_epoch_ms = (time.perf_counter() - t1) * 1e3 / reps
    judged = [h["judge_error"] for h in hist if h["judge_error"] is not None]
    mm = [h["mc_vs_mc"] for h in hist if h["judge_error"] is not None]
    timed = hist[warmup:]
    out = {"wall": wall, "surrogate_epochs": sum(1 for h in timed if h["optics_source"] == "surrogate"),
           "mc_epochs": sum(1 for h in timed if h["optics_source"] == "mc"),
           "judged_block4": [round(e[1], 4) for e in judged], "judged_voxel": [round(e[0], 4) for e in judged],
           "mc_vs_mc_block4": [None if m is None else round(m, 4) for m in mm],
           "excess_block4": [None if m is None else round(excess(e[1], m), 4) for e, m in zip(judged, mm)],
           "judge_passes": sum(1 for h in hist if h["judge_error"] is not None and not h["fallback"]),
           "scale_after_judges": [round(h["scale"], 4) for h in hist if h["judge_error"] is not None],
           "gate_met": gate_met, "judgements_timed": len(timed_judged),
           "judge_passes_timed": sum(1 for h in timed_judged if not h["fallback"]),
           "stage_ms": stage_ms, "sur_epoch_ms": round(sur_epoch_ms, 4), "plain_epoch_ms": round(plain_epoch_ms, 4),
           "state": state_end}
    sur.close()
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-epochs", type=int, default=3)
    ap.add_argument("--graph", action="store_true", help="replay each plain epoch as a captured HIP graph")
    ap.add_argument("--plain-only", action="store_true", help="time the plain MC loop only")
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--mc-lo", type=int, default=37_500,
                    help="photons of the low-photon MC that sets the surrogate's 8^3-block totals (0: none)")
    ap.add_argument("--correct-block", type=int, default=8)
    ap.add_argument("--copy-kernel", action="store_true",
                    help="only run three 256 MiB device-to-device copies (calibrates the HBM byte counters)")
    ap.add_argument("--mc-probe", default=None, metavar="N1,N2,...",
                    help="only time the transport kernel on the workload's state after three epochs, at these "
                         "photon counts (three launches each, with its tally additions and record passes), then "
                         "the atomic-rate probe "
                         "over a range of address spans (for profiling the MC kernel)")
    args = ap.parse_args()
    if args.mc_probe:
        # launches of their own add path lengths inline unless told otherwise; the probe times the
        # kernel as the epochs run it, with the path lengths deferred to records
        os.environ.setdefault("BUGDEF_MC_DEFER_FLUENCE", "1")
        from simulation.coupled_ablation import CoupledAblationEngine
        from simulation.workloads import layered_stack
        eng = CoupledAblationEngine(**layered_stack())
        for _ in range(3):
            eng.step_epoch()
        beam = eng._epoch_beam()
        out = {}
        for n in (int(v) for v in args.mc_probe.split(",")):
            beam.n_photons = n
            ts, rp, atom, recs = [], [], 0.0, 0.0
            for rep in range(3):
                beam.seed = 77 + rep
                r = eng._ctx.mc_run(beam)
                ts.append(r.kernel_ms)
                rp.append(r.record_ms)
                atom, recs = r.tally_atomics, r.tally_records
            out[n] = {"kernel_ms_min": round(min(ts), 3), "kernel_ms_max": round(max(ts), 3),
                      "record_passes_ms_min": round(min(rp), 3),
                      "photons_per_s": round(n / (min(ts) * 1e-3), 1), "tally_atomics": int(atom),
                      "deferred_path_records": int(recs),
                      "g_tally_additions_per_s": round((atom + recs) / (min(ts) * 1e-3) / 1e9, 2)}
            log(f"bench: mc probe {n} photons done")
        nvox = eng.Nx * eng.Ny * eng.Nz
        eng.close()
        from simulation import _bdlib as B
        spans = {str(sp): round(B.probe_atomic_rate(nvox, sp) / 1e9, 2)
                 for sp in (nvox, 262144, 65536, 32768, 8192, 4096, 2048)}
        print(json.dumps({"mc_probe": out, "atomic_probe_g_per_s_by_span": spans}))
        return
    if args.copy_kernel:
        from simulation import _bdlib as B
        n = 256 * 1024 * 1024
        src, dst = B.DeviceArray((n,), dtype="u1"), B.DeviceArray((n,), dtype="u1")
        for _ in range(3):
            dst.copy_from(src)
        print(json.dumps({"copy_bytes_each_way": n, "copies": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; --gpus above 1 is not supported "
                                   "(a run would shard the domain in x slabs with a one-plane halo per "
                                   "sub-step and an all-reduce of the MC tallies)", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    from simulation import _bdlib as B
    from simulation.workloads import layered_stack

    if B.device_count() < 1:
        raise RuntimeError("bench: no HIP device visible")
    kw = layered_stack()
    p = run_plain(kw, args.steps, args.warmup, args.graph)
    plain_rate = args.steps * p["dt_epoch"] * 1e3 / p["wall"]
    # a launch of the pair kernel advances two sub-steps and moves each field once,
    # so its algorithmic bytes per voxel are those of one sub-step
    tun = p["tuning"]
    n_pair = 2 * (p["n_sub"] // 4) if tun["pairs"] else 0
    n_launch = n_pair + (p["n_sub"] - 2 * n_pair)
    sub_launch_ms = float(np.mean(p["sub"])) / n_launch
    achieved = SUBSTEP_BYTES_PER_VOXEL * p["nvox"] / (sub_launch_ms * 1e-3) / 1e9
    per_substep_ms = float(np.mean(p["sub"])) / p["n_sub"]
    mc_avg = float(np.mean(p["mc"]))
    # the transport's tally atomics against the device's measured atomic-add rate on
    # the same float2 grid: pairs hashed over the whole grid, and over 64k voxels
    probe = {f"uniform_{name}": round(B.probe_atomic_rate(p["nvox"], span) / 1e9, 2)
             for name, span in (("full_grid", p["nvox"]), ("65536_voxels", 65536))}
    # tally additions the transport kernel made: float atomics (absorbed weight, and path
    # lengths it could not defer) and deferred path-length records, per second of the kernel
    mc_add_rate = (p["atomics"] + p["records"]) / (float(np.sum(p["mc"])) * 1e-3) / 1e9
    log("bench: atomic-rate probe done")
    copy_gbs = B.probe_copy_rate(1 << 30) / 1e9
    log("bench: copy-rate probe done")
    s = None
    if not args.plain_only:
        if not os.path.exists(args.weights):
            raise RuntimeError(f"bench: surrogate weights not found at {args.weights}")
        s = run_surrogate(kw, args.steps, args.warmup, args.weights, args.mc_lo, args.correct_block)
    sur_rate = args.steps * p["dt_epoch"] * 1e3 / s["wall"] if s else None
    # the headline is the surrogate loop's rate only when its accuracy gate held for every
    # judgement of the timed window; otherwise the plain loop is what delivers the result
    use_sur = s is not None and s["gate_met"]
    rate = sur_rate if use_sur else plain_rate
    wall = s["wall"] if use_sur else p["wall"]
    if s is None:
        source = "plain MC loop (--plain-only)"
    elif use_sur:
        source = "surrogate loop: every judgement in the timed window passed the accuracy gate"
    else:
        source = (f"plain MC loop: the surrogate loop passed {s['judge_passes_timed']} of "
                  f"{s['judgements_timed']} judgements in its timed window, so its rate is not the headline")
    traffic = None if n_pair else measured_traffic(tun)
    epoch_ms = p["wall"] * 1e3 / args.steps
    result = {
        "metric": "simulated ms/s at 256x256x192: surrogate+MC vs plain MC at <=5% L2; photons/s",
        "value": round(rate, 3),
        "unit": "simulated ms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (transport, stencils); bf16 MFMA / f32 accumulate (surrogate); f64 ledgers",
        "data": "synthetic (seeded layered stack, no datasets); surrogate weights trained on device-MC samples",
        "config": {
            "workload": "coupled layered stack: 256x256x192 voxels of 15 um, 7 layers + 21 voxels ambient air, "
                        "40 W at 22 deg, 0.5 mm spot, 1 ms epochs of 150k photons + 20 x 50 us sub-steps, boil "
                        "100 C, ablate 350 C, Arrhenius char, diffusion plume on 128^3" +
                        ("" if args.plain_only else
                         "; surrogate on up to four 64^3 windows tiled along the beam" +
                         (f", its {args.correct_block}^3-block totals from a {args.mc_lo}-photon MC each epoch"
                          if args.mc_lo else "") +
                         "; judge every 10th epoch against the epoch's MC, fallback when the surrogate-vs-MC L2 "
                         "exceeds the MC-vs-MC L2 of two such runs by more than 5% in quadrature (MC-vs-MC " +
                         ("from the low-photon run, scaled by photon count)" if args.mc_lo else
                          "from a second independent run)")),
            "voxels": p["nvox"], "photons_per_epoch": p["n_photons"], "substeps_per_epoch": p["n_sub"],
            "mode": "plain MC" if args.plain_only else "surrogate+MC",
        },
        "value_source": source,
        "plain_mc": {"value": round(plain_rate, 3), "ms_per_step": round(epoch_ms, 4), **p["state"]},
        "photons_per_s": round(p["n_photons"] / (mc_avg * 1e-3), 1),
        "mc_events_per_s": round(p["events"] / (float(np.sum(p["mc"])) * 1e-3), 1),
        "breakdown_ms": {"optical": round(float(np.mean(p["opt"])), 4), "thermal": round(float(np.mean(p["th"])), 4),
                         "substeps": round(float(np.mean(p["sub"])), 4), "mc_kernel": round(mc_avg, 4),
                         "mc_record_passes": round(float(np.mean(p["recp"])), 4),
                         "phase_plume": round(float(np.mean(p["ph"])), 4)},
        "roofline_transport": {
            "kernel": "k_mc (photon transport)",
            "bound": "hbm",
            "achieved": round(MC_BYTES_PER_EVENT * p["events"] / (float(np.sum(p["mc"])) * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(MC_BYTES_PER_EVENT * p["events"] / (float(np.sum(p["mc"])) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_event": MC_BYTES_PER_EVENT,
            "events_per_launch": round(p["events"] / args.steps),
            "avg_launch_ms": round(mc_avg, 4),
            "share_of_epoch": round(mc_avg / epoch_ms, 4),
            "tally_additions_per_launch": round((p["atomics"] + p["records"]) / args.steps),
            "deferred_path_records_per_launch": round(p["records"] / args.steps),
            "tally_additions_G_per_s": round(mc_add_rate, 2),
            "atomic_ceiling_G_per_s": max(probe.values()),
            "atomic_probe": probe,
            "record_passes_ms": round(float(np.mean(p["recp"])), 4),
        },
        "roofline_stencil": {
            "kernel": "k_substep2 (two fused bioheat sub-steps per launch)" if n_pair else
                      "k_substep (fused bioheat sub-step)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_GBps": None if traffic is None else round(traffic / (sub_launch_ms * 1e-3) / 1e9, 1),
            "copy_GBps": round(copy_gbs, 1),
            "frac_of_copy": round(achieved / copy_gbs, 4),
            "bytes_per_voxel": SUBSTEP_BYTES_PER_VOXEL,
            "avg_launch_ms": round(sub_launch_ms, 5),
            "launches_per_epoch": n_launch,
            "ms_per_substep": round(per_substep_ms, 5),
            "share_of_epoch": round(float(np.mean(p["sub"])) / epoch_ms, 4),
            "tuning": tun,
        },
    }
    # the headline roofline is that of the kernel with the larger share of the measured epoch
    share_mc = result["roofline_transport"]["share_of_epoch"]
    share_sub = result["roofline_stencil"]["share_of_epoch"]
    result["roofline"] = dict(result["roofline_stencil"] if share_sub >= share_mc else result["roofline_transport"])
    result["roofline"]["kernel"] += " (the larger share of the epoch)"
    if s:
        jb = s["judged_block4"]
        sur_epoch_ms = s["sur_epoch_ms"]
        result["surrogate"] = {"value": round(sur_rate, 3), "speedup_vs_plain": round(sur_rate / plain_rate, 3),
                               "gate_met": s["gate_met"], "judgements_timed": s["judgements_timed"],
                               "judge_passes_timed": s["judge_passes_timed"],
                               "surrogate_epochs": s["surrogate_epochs"], "mc_epochs": s["mc_epochs"],
                               "judged_rel_l2_block4": jb, "judged_rel_l2_voxel": s["judged_voxel"],
                               "mc_vs_mc_block4": s["mc_vs_mc_block4"], "excess_block4": s["excess_block4"],
                               "judge_passes": s["judge_passes"], "judgements": len(jb),
                               "gate": "d_s^2 <= d_mm^2 + 0.05^2 (block-4 rel L2; d_mm: two independent "
                                       "150k-photon MC runs of the state, " +
                                       ("estimated from the low-photon run" if args.mc_lo else "measured") + ")",
                               "mc_anchored_scale": s["scale_after_judges"],
                               "max_judged_rel_l2_block4": max(jb) if jb else None,
                               "surrogate_epoch_stage_ms": s["stage_ms"],
                               "surrogate_epoch_ms_stages_in_series": round(sum(s["stage_ms"].values()), 4),
                               "surrogate_epoch_ms": round(sur_epoch_ms, 4),
                               "plain_epoch_ms_same_state": s["plain_epoch_ms"],
                               # every 10th epoch a plain MC epoch plus a prediction to judge, the other
                               # nine surrogate epochs: the rate if every judgement passed
                               "speedup_bound_if_every_judgement_passed": round(
                                   10.0 * s["plain_epoch_ms"] / (s["plain_epoch_ms"] + sur_epoch_ms -
                                                                 s["stage_ms"].get("epoch_without_transport", 0.0)
                                                                 + 9.0 * sur_epoch_ms), 3),
                               **s["state"]}
    if not args.no_cpu_baseline:
        info = cpu_info()
        threads = int(os.environ.get("OMP_NUM_THREADS", info["cpus_usable"] or 1))
        chunks = reference_mc_chunks(p["nvox"], p["n_photons"], threads)
        log(f"bench: cpu baseline ({args.cpu_epochs} epochs, MC on {threads} threads)")
        v, secs = cpu_baseline(kw, args.cpu_epochs, threads)
        log(f"bench: cpu baseline ({args.cpu_epochs} epochs, MC in the reference's {chunks} chunks)")
        v2, secs2 = cpu_baseline(kw, args.cpu_epochs, chunks)
        result["cpu_baseline"] = {"value": round(v, 4), "unit": "simulated ms/s", "cores": threads, "kind": "port",
                                  "sample": f"{args.cpu_epochs} epochs of the same workload on the FP64 oracle "
                                            f"engine ({secs:.1f} s), MC in {threads} private-tally chunks",
                                  "reference_chunking": {
                                      "value": round(v2, 4), "mc_chunks": chunks,
                                      "sample": f"{args.cpu_epochs} epochs ({secs2:.1f} s), MC in {chunks} "
                                                f"chunks as voxel_mc_3d.py:621-625 allows at this grid size, "
                                                f"stencil on {threads} threads"},
                                  **info}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
