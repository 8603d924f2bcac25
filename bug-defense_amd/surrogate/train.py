"""Train the surrogate on the GPU from MC-generated samples.

    python -m surrogate.train --samples 4096 --holdout 128 --epochs 40 --out weights.npz

Data: ``surrogate.data.generate`` (device MC, 1M photons per sample by
default).  Optimiser: AdamW on the device (bd_fno_adamw), cosine learning-rate
decay, batches of 8 tiles.  Reports the held-out relative L2 error of the
absorbed power (Q = mu_a Phi, from the network's log fluence) against the
samples' own MC, per sample and pooled.
"""

import argparse
import json
import math
import time

import numpy as np

from . import data
from .network import FourierSurrogate, init_params, T

PHI_EPS = 1e-4
LN1E4 = math.log(1e4)


def q_from(log_phi, feats_mu_a):
    """Absorbed power relative to E0 from a log-fluence field and log1p(mu_a) features."""
    return np.expm1(feats_mu_a) * np.maximum(np.exp(log_phi) - PHI_EPS, 0.0)


def block4(a):
    n = a.shape[-1] // 4
    return a.reshape(a.shape[:-3] + (n, 4, n, 4, n, 4)).sum(axis=(-5, -3, -1))


def rel_l2(p, t):
    s = float((t ** 2).sum())
    return math.sqrt(float(((p - t) ** 2).sum()) / s) if s > 0 else 0.0


def evaluate(net, ds, idx0, n, batch, other=None):
    """Relative L2 of Q, prediction vs the sample's MC, on samples [idx0, idx0 + n):
    median per voxel, median over 4^3 blocks, pooled per voxel.  With `other`
    (the same samples under independent MC streams) also the MC-vs-MC medians."""
    ev, eb, mv, mb, ib, num, den = [], [], [], [], [], 0.0, 0.0
    for c in range(idx0 // batch, (idx0 + n) // batch):
        X, t, _ = ds.chunk(c, batch)
        y = net.forward(X).download().astype(np.float64)
        tt = t.download().astype(np.float64)
        xh = X.download()
        mu = xh[..., 0].astype(np.float64)
        # the refining network's input run, as absorbed power (channel 5 inverted)
        qi = q_from((xh[..., 5].astype(np.float64) - 1.0) * LN1E4, mu) if net.mc_photons else None
        qp = q_from(y, mu)
        qt = q_from(tt, mu)
        qo = q_from(other.chunk(c - idx0 // batch, batch)[1].download().astype(np.float64), mu) if other else None
        for b in range(batch):
            ev.append(rel_l2(qp[b], qt[b]))
            eb.append(rel_l2(block4(qp[b]), block4(qt[b])))
            if qi is not None:
                ib.append(rel_l2(block4(qi[b]), block4(qt[b])))
            num += float(((qp[b] - qt[b]) ** 2).sum())
            den += float((qt[b] ** 2).sum())
            if other:
                mv.append(rel_l2(qo[b], qt[b]))
                mb.append(rel_l2(block4(qo[b]), block4(qt[b])))
    out = {"median_voxel": float(np.median(ev)), "median_block4": float(np.median(eb)),
           "pooled_voxel": math.sqrt(num / den)}
    if ib:
        out["input_run_median_block4"] = float(np.median(ib))
    if other:
        out.update(mc_vs_mc_median_voxel=float(np.median(mv)), mc_vs_mc_median_block4=float(np.median(mb)))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--holdout", type=int, default=128)
    ap.add_argument("--photons", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--wd", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from saved weights")
    ap.add_argument("--modes", default="6,6,8", help="kept modes kx,ky,kz")
    ap.add_argument("--eval-every", type=int, default=10)
    ap.add_argument("--coupled", type=int, default=0, help="windows taken from coupled runs (added to --samples)")
    ap.add_argument("--coupled-holdout", type=int, default=64)
    ap.add_argument("--np-len", type=int, default=80, help="transform length")
    ap.add_argument("--coupled-shape", default="192,192,192", help="grid of the coupled runs")
    ap.add_argument("--coupled-epochs", type=int, default=150, help="epochs per coupled run")
    ap.add_argument("--coupled-power", type=float, default=None, help="fixed power [W] (default: 20-80 W)")
    ap.add_argument("--coupled-theta", type=float, default=None, help="fixed angle [deg] (default: 0-30)")
    ap.add_argument("--coupled-spot", type=float, default=None, help="fixed spot [mm] (default: 0.35-0.65)")
    ap.add_argument("--mc-photons", type=int, default=0,
                    help="train the refining network: channel 5 is the fluence of a run of this many photons")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    nb = a.batch
    ntr = (a.samples // nb) * nb
    nho = (a.holdout // nb) * nb
    t0 = time.perf_counter()
    nco = (a.coupled // nb) * nb
    ncho = (a.coupled_holdout // nb) * nb if nco else 0
    mp = a.mc_photons
    ds = data.generate(ntr, n_photons=a.photons, seed=a.seed, progress=1024, mc_photons=mp) if ntr else None
    ho = data.generate(nho, n_photons=a.photons, seed=a.seed + 1000, mc_photons=mp)
    ho2 = data.generate(nho, n_photons=a.photons, seed=a.seed + 1000, mc_seed_offset=1, mc_photons=mp)
    ckw = dict(n_photons=a.photons, mc_photons=mp, shape=tuple(int(v) for v in a.coupled_shape.split(",")),
               epochs_per_run=a.coupled_epochs, power=a.coupled_power, theta_deg=a.coupled_theta,
               spot_dia=a.coupled_spot)
    cds = data.generate_coupled(nco, seed=a.seed + 7, progress=512, **ckw) if nco else None
    cho = data.generate_coupled(ncho, seed=a.seed + 2000, **ckw) if ncho else None
    t_data = time.perf_counter() - t0
    print(f"data: {ntr} train + {nho} held-out samples (+ independent MC of the held-out) in {t_data:.1f} s",
          flush=True)
    if a.init:
        net = FourierSurrogate.load(a.init, max_batch=nb, train=True)
    else:
        modes = tuple(int(v) for v in a.modes.split(","))
        net = FourierSurrogate(modes=modes, np_len=a.np_len, max_batch=nb, train=True)
        net.set_params(init_params(net.modes, seed=a.seed))
    net.mc_photons = mp
    rng = np.random.default_rng(a.seed + 1)
    chunks = [(ds, c) for c in range(ntr // nb)] + [(cds, c) for c in range(nco // nb)]
    nchunk = len(chunks)
    total = a.epochs * nchunk
    step = 0
    hist = []
    t1 = time.perf_counter()
    for ep in range(a.epochs):
        losses = []
        for ci in rng.permutation(nchunk):
            lr = a.lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            src, c = chunks[int(ci)]
            X, t, w = src.chunk(int(c), nb)
            losses.append(net.train_step(X, t, w, lr, weight_decay=a.wd))
            step += 1
        rec = {"epoch": ep + 1, "loss": float(np.mean(losses)), "s": round(time.perf_counter() - t1, 1)}
        if (ep + 1) % a.eval_every == 0 and ep + 1 < a.epochs:
            rec.update(evaluate(net, ho, 0, nho, nb))
            if cho:
                rec.update({"coupled_" + k: v for k, v in evaluate(net, cho, 0, ncho, nb).items()})
        hist.append(rec)
        print(json.dumps(rec), flush=True)
    t_train = time.perf_counter() - t1
    summary = {"train_samples": ntr, "heldout_samples": nho, "photons": a.photons, "mc_photons": mp,
               "epochs": a.epochs,
               "modes": list(net.modes), "np_len": net.np_len, "data_s": round(t_data, 1),
               "train_s": round(t_train, 1), "heldout": evaluate(net, ho, 0, nho, nb, other=ho2),
               "coupled_windows": nco, "coupled": {k: (list(v) if isinstance(v, tuple) else v)
                                                    for k, v in ckw.items() if k not in ("n_photons", "mc_photons")}, "coupled_heldout": evaluate(net, cho, 0, ncho, nb) if cho else None}
    net.save(a.out, history=json.dumps(hist), summary=json.dumps(summary))
    print(json.dumps(summary), flush=True)
    net.close()


if __name__ == "__main__":
    main()
