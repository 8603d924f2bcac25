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
judge epoch (still every ``judge_every``-th) finds the error back within
``tol``; the fallback epochs in between are plain epochs, with no prediction.

The judged error is the relative L2 over 4x4x4 block sums of the absorbed
power (non-ablated voxels), which averages down the MC's own shot noise; the
per-voxel figure is recorded beside it.

MC-anchored scale (``calibrate``, on by default): each judge epoch also
compares the total absorbed power of the MC field with that of the composed
prediction, and the following predictions are multiplied by the ratio (kept
within [0.5, 2]).  A judged error is always that of a prediction made before
its own epoch's MC, with the scale of the previous judge epoch.
"""

import time

from simulation._bdlib import DeviceArray

from .network import CIN, T, FourierSurrogate


class SurrogateCoupledEngine:
    def __init__(self, engine, weights, judge_every=10, tol=0.05, margin=12, metric="block4", max_windows=4,
                 calibrate=True):
        if engine.plume is not None and engine.plume.model != 'diffusion':
            raise ValueError("the surrogate loop supports the diffusion plume only")
        self.engine = engine
        self.max_windows = int(max_windows)
        self.net = weights if isinstance(weights, FourierSurrogate) else FourierSurrogate.load(
            weights, max_batch=self.max_windows)
        if self.net.max_batch < self.max_windows:
            raise ValueError("the network's max_batch is smaller than max_windows")
        self.judge_every = int(judge_every)
        self.tol = float(tol)
        self.margin = int(margin)
        self.metric = metric
        self.x = DeviceArray((self.max_windows, T, T, T, CIN))
        self.y = DeviceArray((self.max_windows, T, T, T))
        self.windows = []
        self.calibrate = bool(calibrate)
        self.scale = 1.0
        self.fallback = False
        self.have_stale = False
        self.judged = False
        self.history = []

    def _predict(self, beam, into_q):
        ctx = self.engine._ctx
        ws = ctx.sur_windows(beam, self.margin, self.max_windows)
        nb = len(ws)
        ctx.sur_features_n(beam, ws, self.x)
        ctx.sync()
        self.net.forward(self.x.view(0, (nb, T, T, T, CIN)), self.y.view(0, (nb, T, T, T)))
        ctx.sur_apply_n(beam, ws, self.y, into_q, self.scale)
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
        judging = (not self.judged) or (eng.epoch % self.judge_every == 0)
        judge = judging or self.fallback
        t0 = time.perf_counter()
        if judge:
            err = None
            compare = self.have_stale and judging
            if compare:
                beam = eng._epoch_beam()
                self._predict(beam, into_q=False)
            d = eng.step_epoch()
            if compare:
                j = eng._ctx.sur_judge()
                err = (j[0], j[1])
                e = err[1] if self.metric == "block4" else err[0]
                self.fallback = e > self.tol
                self.judged = True
                if self.calibrate and j[2] > 0.0 and j[3] > 0.0:
                    self.scale = min(2.0, max(0.5, self.scale * j[2] / j[3]))
            eng._ctx.sur_keep_stale()
            self.have_stale = True
            d['optics_source'] = 'mc'
            d['judge_error'] = err
        else:
            beam = eng._epoch_beam()
            self._predict(beam, into_q=True)
            d = eng.step_epoch_with_q()
            d['optics_source'] = 'surrogate'
            d['judge_error'] = None
        d['t_wall_s'] = time.perf_counter() - t0
        d['fallback'] = self.fallback
        d['windows'] = len(self.windows)
        d['scale'] = self.scale
        self.history.append({k: d[k] for k in ('epoch', 'optics_source', 'judge_error', 'fallback', 'n_new', 'windows',
                                               'scale')})
        return d

    def close(self):
        self.x.free()
        self.y.free()
        self.net.close()
