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


def q_from(log_phi, feats_mu_a):
    """Absorbed power relative to E0 from a log-fluence field and log1p(mu_a) features."""
    return np.expm1(feats_mu_a) * np.maximum(np.exp(log_phi) - PHI_EPS, 0.0)


def evaluate(net, ds, idx0, n, batch):
    """Relative L2 of Q (prediction vs the sample's MC) on samples [idx0, idx0 + n)."""
    errs, num, den = [], 0.0, 0.0
    for c in range(idx0 // batch, (idx0 + n) // batch):
        X, t, _ = ds.chunk(c, batch)
        y = net.forward(X).download()
        tt = t.download()
        mu = X.download()[..., 0]
        qp = q_from(y, mu)
        qt = q_from(tt, mu)
        for b in range(batch):
            d = float(((qp[b] - qt[b]) ** 2).sum())
            s = float((qt[b] ** 2).sum())
            errs.append(math.sqrt(d / s) if s > 0 else 0.0)
            num += d
            den += s
    return float(np.median(errs)), float(np.mean(errs)), math.sqrt(num / den)


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
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    nb = a.batch
    ntr = (a.samples // nb) * nb
    nho = (a.holdout // nb) * nb
    t0 = time.perf_counter()
    ds = data.generate(ntr + nho, n_photons=a.photons, seed=a.seed, progress=512)
    t_data = time.perf_counter() - t0
    print(f"data: {ntr} train + {nho} held-out samples in {t_data:.1f} s", flush=True)
    if a.init:
        net = FourierSurrogate.load(a.init, max_batch=nb, train=True)
    else:
        net = FourierSurrogate(max_batch=nb, train=True)
        net.set_params(init_params(net.modes, seed=a.seed))
    rng = np.random.default_rng(a.seed + 1)
    nchunk = ntr // nb
    total = a.epochs * nchunk
    step = 0
    hist = []
    t1 = time.perf_counter()
    for ep in range(a.epochs):
        losses = []
        for c in rng.permutation(nchunk):
            lr = a.lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            X, t, w = ds.chunk(int(c), nb)
            losses.append(net.train_step(X, t, w, lr, weight_decay=a.wd))
            step += 1
        rec = {"epoch": ep + 1, "loss": float(np.mean(losses)), "s": round(time.perf_counter() - t1, 1)}
        if (ep + 1) % 5 == 0 or ep + 1 == a.epochs:
            med, mean, pooled = evaluate(net, ds, ntr, nho, nb)
            rec.update(heldout_median=med, heldout_mean=mean, heldout_pooled=pooled)
        hist.append(rec)
        print(json.dumps(rec), flush=True)
    t_train = time.perf_counter() - t1
    med, mean, pooled = evaluate(net, ds, ntr, nho, nb)
    summary = {"train_samples": ntr, "heldout_samples": nho, "photons": a.photons, "epochs": a.epochs,
               "data_s": round(t_data, 1), "train_s": round(t_train, 1), "heldout_median_rel_l2": med,
               "heldout_mean_rel_l2": mean, "heldout_pooled_rel_l2": pooled}
    net.save(a.out, history=json.dumps(hist), summary=json.dumps(summary))
    print(json.dumps(summary), flush=True)
    net.close()


if __name__ == "__main__":
    main()
