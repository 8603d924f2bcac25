"""Device surrogate coupling (bd_sur_*) against the FP64 oracle
(oracle/surrogate_oracle.py), and bd_epoch_q against bd_epoch."""

import math

import numpy as np
import pytest

from oracle import surrogate_oracle as S

pytestmark = pytest.mark.gpu

SP = (0.015, 0.015, 0.015)


def _ctx(props, shape):
    from simulation import _bdlib as B
    c = B.Context(shape, SP)
    for k in ('mu_a', 'mu_s', 'g', 'n'):
        c.upload(k.upper(), props[k].astype(np.float32))
    return c


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def test_features_and_targets_match_oracle_on_a_training_sample():
    from simulation._bdlib import DeviceArray
    from surrogate import data as D
    s = D.random_sample(np.random.default_rng(3))
    c = _ctx(s, (64, 64, 64))
    beam = D.sample_beam(s, 200_000, 11)
    c.mc_run(beam)
    x, t, w = DeviceArray((64, 64, 64, 8)), DeviceArray((64, 64, 64)), DeviceArray((64, 64, 64))
    c.sur_features(beam, (0, 0, 0), x)
    c.sur_targets(beam, (0, 0, 0), t, w)
    c.sync()
    xo = S.features(*(_f32(s[k]) for k in ('mu_a', 'mu_s', 'g', 'n')), SP, (0, 0, 0), s['xc'], s['yc'],
                    s['theta'], s['spot'])
    xg = x.download()
    for ch in range(8):
        np.testing.assert_allclose(xg[..., ch], xo[..., ch], rtol=2e-5, atol=2e-5, err_msg=f"channel {ch}")
    L = c.download('L')
    to, wo = S.targets(L, _f32(s['mu_a']), SP, (0, 0, 0), 200_000, s['spot'])
    np.testing.assert_allclose(t.download(), to, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(w.download(), wo, rtol=1e-4, atol=1e-12)
    assert abs(float(w.download().astype(np.float64).sum()) - 1.0) < 1e-4
    c.close()


def test_mc_fluence_channel_matches_oracle():
    """bd_sur_features_mc: channel 5 carries the last MC run's fluence (the
    target scaling), every other channel is that of bd_sur_features_n."""
    from simulation._bdlib import DeviceArray
    from surrogate import data as D
    s = D.random_sample(np.random.default_rng(5))
    c = _ctx(s, (64, 64, 64))
    beam = D.sample_beam(s, 40_000, 17)
    c.mc_run(beam)
    x0, x1 = DeviceArray((1, 64, 64, 64, 8)), DeviceArray((1, 64, 64, 64, 8))
    c.sur_features_n(beam, [(0, 0, 0)], x0)
    c.sur_features_n(beam, [(0, 0, 0)], x1, mc_photons=40_000)
    c.sync()
    L = c.download('L')
    xo = S.features(*(_f32(s[k]) for k in ('mu_a', 'mu_s', 'g', 'n')), SP, (0, 0, 0), s['xc'], s['yc'],
                    s['theta'], s['spot'], L=L, mc_photons=40_000)
    a, b = x0.download()[0], x1.download()[0]
    for ch in (0, 1, 2, 3, 4, 6, 7):
        np.testing.assert_array_equal(a[..., ch], b[..., ch])
    np.testing.assert_allclose(b[..., 5], xo[..., 5], rtol=2e-5, atol=2e-5)
    assert b[..., 5].max() > 0.5 and b[..., 5].min() >= 0.0
    # exp of the channel recovers Phi/E0 (the target transform)
    to, _ = S.targets(L, _f32(s['mu_a']), SP, (0, 0, 0), 40_000, s['spot'])
    np.testing.assert_allclose((b[..., 5] - 1.0) * S.LN1E4, to, rtol=1e-4, atol=1e-4)
    c.close()


def test_coarse_correct_and_mc_vs_mc_judge_match_oracle():
    """bd_sur_correct (block totals from the kept field) and bd_sur_judge_aux
    against the oracle on a ragged grid with ablated voxels."""
    from simulation import _bdlib as B
    shape = (37, 30, 27)
    rng = np.random.default_rng(12)
    c = B.Context(shape, SP)
    try:
        abl = rng.random(shape) < 0.15
        F = rng.gamma(2.0, 1.0, shape).astype(np.float32)
        R = (F * rng.uniform(0.5, 1.5, shape)).astype(np.float32)
        F[:8, :8, :8] = 0.0
        c.upload('ABLATED', abl)
        for block in (8, 4):
            c.upload('Q', R)
            c.sur_keep_aux()
            c.upload('Q', F)
            c.sur_correct(block, into_q=True)
            got = c.download('Q').astype(np.float64)
            want = S.coarse_correct(_f32(F), _f32(R), abl, block)
            np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-12, err_msg=f"block {block}")
        # MC-vs-MC distance: kept field against Q
        c.upload('Q', R)
        c.sur_keep_aux()
        c.upload('Q', F)
        j = c.sur_judge_aux()
        e0, e1 = S.rel_error(_f32(F), _f32(R), abl)
        assert abs(j[0] - e0) < 1e-9 and abs(j[1] - e1) < 1e-9
        keep = ~abl
        assert abs(j[2] - float(_f32(F)[keep].sum())) < 1e-6 * j[2]
        assert abs(j[3] - float(_f32(R)[keep].sum())) < 1e-6 * j[3]
    finally:
        c.close()


def _stack(shape, rng, air=21):
    props = {k: np.zeros(shape) for k in ('mu_a', 'mu_s', 'g')}
    props['n'] = np.ones(shape)
    z = air
    while z < shape[2]:
        z1 = min(shape[2], z + int(rng.integers(3, 20)))
        props['mu_a'][:, :, z:z1] = rng.uniform(0.01, 20)
        props['mu_s'][:, :, z:z1] = rng.uniform(1, 40)
        props['g'][:, :, z:z1] = rng.uniform(0.7, 0.95)
        props['n'][:, :, z:z1] = rng.uniform(1.0, 1.55)
        z = z1
    return props


def test_window_features_apply_and_error_on_a_larger_domain():
    from simulation._bdlib import DeviceArray
    from simulation.voxel_mc_3d import make_beam
    rng = np.random.default_rng(5)
    shape = (112, 96, 120)
    props = _stack(shape, rng)
    abl = np.zeros(shape, bool)
    abl[50:62, 40:56, 21:45] = True   # a crater under the beam
    for k in ('mu_a', 'mu_s', 'g'):
        props[k][abl] = 0.0
    props['n'][abl] = 1.0
    c = _ctx(props, shape)
    c.upload('ABLATED', abl)
    th = math.radians(22.0)
    xc, yc = 56 * 0.015, 48 * 0.015
    beam = make_beam(40.0, 100_000, 0.5, 22.0, 0.0, (xc, yc, 0.0), 0.0, 0, 'gaussian', 7)
    o = c.sur_window(beam, 12)
    oo, front = S.window(_f32(props['mu_a']), _f32(props['mu_s']), abl, SP, xc, yc, th, margin=12)
    assert tuple(o[:3]) == oo and o[3] == front
    x = DeviceArray((64, 64, 64, 8))
    c.sur_features(beam, o, x)
    c.sync()
    xo = S.features(*(_f32(props[k]) for k in ('mu_a', 'mu_s', 'g', 'n')), SP, oo, xc, yc, th, 0.5)
    np.testing.assert_allclose(x.download(), xo, rtol=2e-5, atol=2e-5)
    # composition over a stale field, into Q and into the judged field
    Qs = rng.uniform(0.0, 50.0, shape).astype(np.float32)
    c.upload('Q', Qs)
    c.sur_keep_stale()
    y = rng.normal(-3.0, 1.0, (64, 64, 64)).astype(np.float32)
    yd = DeviceArray((64, 64, 64)).upload(y)
    c.sur_apply(beam, o, yd, into_q=False)
    want = S.apply(_f32(y), _f32(props['mu_a']), _f32(Qs), oo, 40.0, 0.5)
    e = c.sur_error()
    ew = S.rel_error(_f32(Qs), want, abl)
    assert e[0] == pytest.approx(ew[0], rel=1e-5) and e[1] == pytest.approx(ew[1], rel=1e-5)
    c.sur_apply(beam, o, yd, into_q=True)
    np.testing.assert_allclose(c.download('Q'), want, rtol=2e-6, atol=1e-6)
    c.close()


@pytest.mark.parametrize("phi_deg", [0.0, 180.0])
def test_tiled_windows_features_and_apply_match_oracle(phi_deg):
    """A deep oblique crater: the stacked windows, their features marched from
    the domain top (towards +x and towards -x), and the composition over
    overlapping windows, against the oracle."""
    from simulation._bdlib import DeviceArray
    from simulation.voxel_mc_3d import make_beam
    rng = np.random.default_rng(9)
    shape = (144, 80, 200)
    props = _stack(shape, rng)
    th = math.radians(22.0)
    sd = 1 if phi_deg == 0.0 else -1
    xc, yc = (40.0 if sd > 0 else 104.0) * 0.015, 40.5 * 0.015
    # a crater following the beam axis down to z = 150
    X, Y, Z = np.meshgrid(np.arange(shape[0]) + 0.5, np.arange(shape[1]) + 0.5, np.arange(shape[2]) + 0.5,
                          indexing='ij')
    ax = xc / 0.015 + sd * Z * math.tan(th)
    abl = ((X - ax) ** 2 + (Y - yc / 0.015) ** 2 < 9.0 ** 2) & (Z >= 21) & (Z < 150)
    for k in ('mu_a', 'mu_s', 'g'):
        props[k][abl] = 0.0
    props['n'][abl] = 1.0
    c = _ctx(props, shape)
    c.upload('ABLATED', abl)
    beam = make_beam(40.0, 100_000, 0.5, 22.0, phi_deg, (xc, yc, 0.0), 0.0, 0, 'gaussian', 7)
    ws = c.sur_windows(beam, 12, 4)
    wo, front, surf = S.windows(_f32(props['mu_a']), _f32(props['mu_s']), abl, SP, xc, yc, th,
                                phi=math.radians(phi_deg), margin=12)
    assert [w[:3] for w in ws] == wo and all(w[3] == front for w in ws)
    assert len(ws) >= 3
    x = DeviceArray((len(ws), 64, 64, 64, 8))
    c.sur_features_n(beam, ws, x)
    c.sync()
    xg = x.download()
    for k, o in enumerate(wo):
        xo = S.features(*(_f32(props[n]) for n in ('mu_a', 'mu_s', 'g', 'n')), SP, o, xc, yc, th, 0.5,
                        phi=math.radians(phi_deg))
        np.testing.assert_allclose(xg[k], xo, rtol=2e-5, atol=2e-5, err_msg=f"window {k}")
    Qs = rng.uniform(0.0, 50.0, shape).astype(np.float32)
    c.upload('Q', Qs)
    c.sur_keep_stale()
    ys = rng.normal(-3.0, 1.0, (len(ws), 64, 64, 64)).astype(np.float32)
    yd = DeviceArray(ys.shape).upload(ys)
    c.sur_apply_n(beam, ws, yd, into_q=True)
    want = S.apply_n([_f32(y) for y in ys], _f32(props['mu_a']), _f32(Qs), wo, 40.0, 0.5)
    np.testing.assert_allclose(c.download('Q'), want, rtol=2e-6, atol=1e-6)
    # the scaled composition, and the judge's totals against the composed field
    c.sur_apply_n(beam, ws, yd, into_q=False, scale=1.7)
    want = S.apply_n([_f32(y) for y in ys], _f32(props['mu_a']), _f32(Qs), wo, 40.0, 0.5, scale=1.7)
    j = c.sur_judge()
    q = c.download('Q')   # the unscaled composition written above
    keep = ~abl
    assert j[2] == pytest.approx(float(q[keep].sum()), rel=1e-6)
    assert j[3] == pytest.approx(float(want[keep].sum()), rel=1e-5)
    ew = S.rel_error(q, want, abl)
    assert j[0] == pytest.approx(ew[0], rel=1e-5) and j[1] == pytest.approx(ew[1], rel=1e-5)
    c.close()


def test_epoch_q_reproduces_bd_epoch_given_the_same_q():
    """An epoch fed the MC epoch's own Q through bd_epoch_q evolves the state
    exactly as the MC epoch did (both engines start from the same state; the
    MC tallies themselves are float atomics, so the comparison starts at epoch 0,
    at a power that boils and removes cells in that epoch)."""
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.workloads import layered_stack
    kw = layered_stack(shape=(48, 48, 64), n_photons=20_000, plume_grid=None, seed=2, power=400.0)
    e1 = CoupledAblationEngine(**kw)
    e2 = CoupledAblationEngine(**kw)
    d1 = e1.step_epoch()
    assert d1['n_new'] > 0
    q = e1._ctx.download('Q', dtype=np.float32)
    e2._ctx.upload('Q', q)
    d2 = e2.step_epoch_with_q()
    for name in ('T', 'H_ACC', 'CHI', 'ABLATED', 'DRY'):
        np.testing.assert_array_equal(e1._ctx.download(name), e2._ctx.download(name), err_msg=name)
    assert d1['n_new'] == d2['n_new']
    e1.close()
    e2.close()
