"""The coupled epoch loop with the surrogate in it (judge / fallback).

``SurrogateCoupledEngine`` drives a ``CoupledAblationEngine``: on most epochs
the network supplies the absorbed power over the tiled beam windows (up to
four 64^3 windows stacked along the beam from below the ablation front to
above the tissue surface, overlapping by 8 voxels: features from the current
optics, one batched forward pass, Q composed over the windows and with the
field of the last MC epoch outside them) and the epoch runs without transport
(``step_epoch_with_q``); every ``judge_every``-th epoch the plain MC epoch runs
and judges the network, whose prediction for the same state is composed into a
separate field and compared with the MC field.  When the judged relative L2
error exceeds ``tol``, the loop falls back to MC on every epoch until a later
judge epoch finds the error back within ``tol``; the fallback epochs in between
are plain epochs, with no prediction.  While falling back, the interval to the
next judgement doubles after each failed one, up to 8 ``judge_every``
(``backoff``), so a network that keeps failing costs few judgements.

The judged error is the relative L2 over 4x4x4 block sums of the absorbed
power (non-ablated voxels), which averages down the MC's own shot noise; the
per-voxel figure is recorded beside it.

Noise-aware gate (``noise_aware``, on by default).  Two independent MC runs of
one state already differ, at the bench's 150k photons by 4-6 % (block-4) early
and 15 % once the crater is 2.5 mm deep, so no field can be within 5 % of one
MC run there, plain MC included.  Each judge epoch therefore also measures
the MC-vs-MC distance d_mm beside the surrogate-vs-MC distance d_s: from a
second MC of the same state (same photon count, disjoint photon ids), or,
when the low-photon run below exists, from that run's distance d_lo to the
epoch's MC, d_mm^2 = 2 d_lo^2 / (N / N_lo + 1) (independent photons: the
variance of each run is inversely proportional to its photon count).
The prediction passes when d_s^2 <= d_mm^2 + tol^2: it agrees with the epoch's
MC as well as a second MC run would, up to tol in quadrature.  Since
E d_s^2 = e^2 + s^2 and E d_mm^2 = 2 s^2 (e: the prediction's distance from
the expected field, s: one run's noise), this asks e^2 <= s^2 + tol^2, i.e.
the prediction may be off the expected field by what one MC run is, plus tol.

Coarse amplitudes (``mc_photons_lo`` > 0): each prediction also runs a
low-photon MC of the current state (photon ids disjoint from the epoch's own
stream), queued on the engine's stream beside the network's forward pass on
its own stream (the features do not use it), and the composed field takes its
totals over blocks of
``correct_block``^3 voxels from that run (bd_sur_correct); the network supplies
the shape inside each block, where the low-photon run is noisiest.  That run's
absorbed power also replaces the kept field outside the windows.

Refining networks (weights saved with ``mc_photons`` > 0): each prediction
first runs a low-photon MC of the current state (``mc_photons`` photons, photon
ids disjoint from the epoch's own stream), whose fluence is channel 5 of the
features; that run's absorbed power also replaces the kept field outside the
windows.  The network then refines the low-photon field into the high-photon
one the epoch would have had.

MC-anchored scale (``calibrate``, on by default): each judge epoch also
compares the total absorbed power of the MC field with that of the composed
prediction, and the following predictions are multiplied by the ratio (kept
within [0.5, 2]).  A judged error is always that of a prediction made before
its own epoch's MC, with the scale of the previous judge epoch.
"""

import time

from simulation._bdlib import DeviceArray

from .data import noisy_beam
from .network import CIN, T, FourierSurrogate

JUDGE_PHOTON_OFFSET = 1 << 41   # photon ids of the judge's second MC run


def gate_passes(d_s, d_mm, tol):
    """The judge's test: the surrogate-vs-MC distance d_s may exceed the
    MC-vs-MC distance d_mm by at most tol in quadrature (d_mm None: d_s <= tol)."""
    if d_mm is None:
        return d_s <= tol
    return d_s * d_s <= d_mm * d_mm + tol * tol


def mc_vs_mc_from_lo(d_lo, n_photons, n_lo):
    """The distance two independent n_photons runs would have, from the distance
    d_lo of an independent n_lo run to one of them: E d_lo^2 = (n/n_lo + 1) s^2 and
    E d_mm^2 = 2 s^2 with s^2 the variance of an n_photons run."""
    return d_lo * (2.0 / (n_photons / n_lo + 1.0)) ** 0.5


def excess(d_s, d_mm):
    """sqrt(max(d_s^2 - d_mm^2, 0)): the part of d_s that a second MC run does not show."""
    return (max(d_s * d_s - d_mm * d_mm, 0.0)) ** 0.5


class SurrogateCoupledEngine:
    def __init__(self, engine, weights, judge_every=10, tol=0.05, margin=12, metric="block4", max_windows=4,
                 calibrate=True, noise_aware=True, mc_photons_lo=0, correct_block=8, backoff=True):
        if engine.plume is not None and engine.plume.model != 'diffusion':
            raise ValueError("the surrogate loop supports the diffusion plume only")
        self.engine = engine
        self.max_windows = int(max_windows)
        self.net = weights if isinstance(weights, FourierSurrogate) else FourierSurrogate.load(
            weights, max_batch=self.max_windows)
        if self.net.max_batch < self.max_windows:
            raise ValueError("the network's max_batch is smaller than max_windows")
        self.judge_every = int(judge_every)
        self.backoff = bool(backoff)
        self.interval = self.judge_every
        self.next_judge = 0
        self.tol = float(tol)
        self.margin = int(margin)
        self.metric = metric
        self.x = DeviceArray((self.max_windows, T, T, T, CIN))
        self.y = DeviceArray((self.max_windows, T, T, T))
        self.windows = []
        self.calibrate = bool(calibrate)
        self.noise_aware = bool(noise_aware)
        self.mc_lo = int(mc_photons_lo)
        self.correct_block = int(correct_block)
        self.scale = 1.0
        self.fallback = False
        self.have_stale = False
        self.judged = False
        self.history = []
        # per-stage device times of predictions [ms], summed while ``profile`` is set; each stage
        # ends with a host wait, so profiling costs a few syncs per epoch and stays off by default
        self.profile = False
        self.stage_ms = {}

    def _mark(self, name, t0):
        if not self.profile:
            return t0
        self.engine._ctx.sync()
        self.net.sync()
        t1 = time.perf_counter()
        self.stage_ms[name] = self.stage_ms.get(name, 0.0) + (t1 - t0) * 1e3
        return t1

    def _predict(self, beam, into_q):
        ctx = self.engine._ctx
        t0 = time.perf_counter()
        ws = ctx.sur_windows(beam, self.margin, self.max_windows)
        nb = len(ws)
        mp = self.net.mc_photons
        lo = mp or self.mc_lo
        t0 = self._mark("windows", t0)

        def low_photon_mc(wait):
            # only a refining network reads the run's fluence; otherwise it scores absorbed
            # weight alone (its Q gives the block totals), with no path-length records to add
            (ctx.mc_run if wait else ctx.mc_launch)(noisy_beam(beam, lo, fluence=bool(mp)))
            ctx.sur_keep_stale()
            if self.mc_lo:
                ctx.sur_keep_aux()

        if lo and mp:
            # a refining network's features hold the low-photon run's fluence: that run first
            low_photon_mc(True)
            t0 = self._mark("mc_lo", t0)
        ctx.sur_features_n(beam, ws, self.x, mp)
        ctx.sync()
        t0 = self._mark("features", t0)
        if lo and not mp:
            # the features do not need the low-photon run: it is queued on the engine's stream and
            # runs beside the forward pass on the network's stream; the composition below follows
            # it on the engine's stream
            low_photon_mc(False)
            t0 = self._mark("mc_lo", t0)
        self.net.forward(self.x.view(0, (nb, T, T, T, CIN)), self.y.view(0, (nb, T, T, T)))
        t0 = self._mark("network", t0)
        ctx.sur_apply_n(beam, ws, self.y, into_q, self.scale)
        if self.mc_lo:
            ctx.sur_correct(self.correct_block, into_q)
        self._mark("compose", t0)
        self.windows = ws
        return ws

    def step_epoch(self) -> dict:
        eng = self.engine
        if eng.power <= 0.0:
            d = eng.step_epoch()
            d['optics_source'] = 'none'
            return d
        # MC on the first epoch (it also supplies the field outside the window), MC and a
        # judgement on the next one, then a judgement every judge_every-th epoch; while
        # falling back every epoch is MC, judged only on those epochs
        self.windows = []
        judging = (not self.judged) or (eng.epoch >= self.next_judge)
        judge = judging or self.fallback
        t0 = time.perf_counter()
        if judge:
            err = mm = None
            compare = self.have_stale and judging
            if compare:
                beam = eng._epoch_beam()
                self._predict(beam, into_q=False)
                if self.noise_aware and not self.mc_lo:
                    # a second MC of the same state: the photon count of the epoch's own run, disjoint ids
                    eng._ctx.mc_run(noisy_beam(beam, beam.n_photons, JUDGE_PHOTON_OFFSET, fluence=False))
                    eng._ctx.sur_keep_aux()
            d = eng.step_epoch()
            if compare:
                j = eng._ctx.sur_judge()
                err = (j[0], j[1])
                k = 1 if self.metric == "block4" else 0
                e = err[k]
                if self.noise_aware:
                    mm = eng._ctx.sur_judge_aux()[k]   # the kept run against the epoch's MC
                    if self.mc_lo:
                        mm = mc_vs_mc_from_lo(mm, eng.n_photons, self.mc_lo)
                self.fallback = not gate_passes(e, mm, self.tol)
                self.judged = True
                e0 = eng.epoch - 1   # the epoch just judged
                if not self.fallback or not self.backoff:
                    self.interval = self.judge_every
                    self.next_judge = (e0 // self.judge_every + 1) * self.judge_every
                else:
                    self.next_judge = e0 + self.interval
                    self.interval = min(2 * self.interval, 8 * self.judge_every)
                if self.calibrate and j[2] > 0.0 and j[3] > 0.0:
                    self.scale = min(2.0, max(0.5, self.scale * j[2] / j[3]))
            eng._ctx.sur_keep_stale()
            self.have_stale = True
            d['optics_source'] = 'mc'
            d['judge_error'] = err
            d['mc_vs_mc'] = mm
        else:
            beam = eng._epoch_beam()
            self._predict(beam, into_q=True)
            t1 = time.perf_counter()
            d = eng.step_epoch_with_q()
            self._mark("epoch_without_transport", t1)
            d['optics_source'] = 'surrogate'
            d['judge_error'] = None
            d['mc_vs_mc'] = None
        d['t_wall_s'] = time.perf_counter() - t0
        d['fallback'] = self.fallback
        d['windows'] = len(self.windows)
        d['scale'] = self.scale
        self.history.append({k: d[k] for k in ('epoch', 'optics_source', 'judge_error', 'mc_vs_mc', 'fallback',
                                               'n_new', 'windows', 'scale')})
        return d

    def close(self):
        self.x.free()
        self.y.free()
        self.net.close()
