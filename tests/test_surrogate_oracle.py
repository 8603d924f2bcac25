"""Pins of the surrogate-coupling oracle (oracle/surrogate_oracle.py) by closed
forms: Beer-Lambert with the normal-incidence Fresnel factor, the Snell drift of
an oblique beam's centroid, target/weight identities, and composition."""

import math

import numpy as np
import pytest

from oracle import surrogate_oracle as S

SP = (0.015, 0.015, 0.015)


def _slab(n_t=1.4, mu_a=0.5, mu_s=3.0, air=0, shape=(64, 64, 64)):
    ma = np.full(shape, mu_a)
    ms = np.full(shape, mu_s)
    g = np.full(shape, 0.9)
    n = np.full(shape, n_t)
    ma[:, :, :air] = 0.0
    ms[:, :, :air] = 0.0
    n[:, :, :air] = 1.0
    return ma, ms, g, n


def test_beer_lambert_with_fresnel_at_normal_incidence():
    ma, ms, g, n = _slab()
    xc = yc = 32 * 0.015
    B = S.ballistic(ma, ms, n, SP, (0, 0, 0), xc, yc, 0.0, 0.5)
    R = ((1.0 - 1.4) / 2.4) ** 2
    I0 = math.exp(-2.0 * (2 * (0.5 * 0.015) ** 2) / 0.25 ** 2)   # voxel (32, 32) centre is half a voxel off axis
    k = np.arange(64)
    want = I0 * (1 - R) * np.exp(-(0.5 + 3.0) * 0.015 * (k + 1))
    np.testing.assert_allclose(B[32, 32, :], want, rtol=1e-12)


def test_oblique_centroid_drifts_by_snell():
    ma, ms, g, n = _slab(n_t=1.5, mu_a=0.0, mu_s=0.0, air=10)
    th = math.radians(30.0)
    B = S.ballistic(ma, ms, n, SP, (0, 0, 0), 20 * 0.015, 32 * 0.015, th, 0.06)
    x = np.arange(64)
    cen = [(B[:, 32, k] * x).sum() / B[:, 32, k].sum() for k in range(40)]
    d = np.diff(cen)
    assert d[3] == pytest.approx(math.tan(th), rel=1e-9)                      # in air
    tt = math.asin(math.sin(th) / 1.5)
    assert d[25] == pytest.approx(math.tan(tt), rel=1e-9)                     # in the n = 1.5 medium


def test_targets_identities():
    rng = np.random.default_rng(0)
    L = rng.uniform(0.0, 5.0, (70, 66, 64))
    ma = rng.uniform(0.0, 2.0, (70, 66, 64))
    t, w = S.targets(L, ma, SP, (3, 1, 0), 1e6, 0.5)
    assert w.sum() == pytest.approx(1.0, rel=1e-12)
    scale = math.pi * 0.25 ** 2 / 2 / (1e6 * 0.015 ** 3)
    np.testing.assert_allclose(np.exp(t) - 1e-4, L[3:67, 1:65, :] * scale, rtol=1e-10, atol=1e-12)


def test_apply_is_consistent_and_blends_to_the_stale_field():
    rng = np.random.default_rng(1)
    shape = (100, 90, 80)
    ma = rng.uniform(0.1, 2.0, shape)
    Qs = rng.uniform(0.0, 10.0, shape)
    o = (20, 13, 5)
    E0 = 40.0 / (math.pi * 0.25 ** 2 / 2)
    # a prediction that reproduces the stale field leaves it unchanged
    y = np.log(Qs[20:84, 13:77, 5:69] / (E0 * ma[20:84, 13:77, 5:69]) + 1e-4)
    np.testing.assert_allclose(S.apply(y, ma, Qs, o, 40.0, 0.5), Qs, rtol=1e-10)
    # an arbitrary prediction: exact inside, untouched outside
    y2 = rng.standard_normal((64, 64, 64))
    out = S.apply(y2, ma, Qs, o, 40.0, 0.5)
    c = (20 + 32, 13 + 32, 5 + 32)
    assert out[c] == pytest.approx(E0 * ma[c] * max(math.exp(y2[32, 32, 32]) - 1e-4, 0.0), rel=1e-12)
    assert out[0, 0, 0] == Qs[0, 0, 0] and out[99, 89, 79] == Qs[99, 89, 79]


def test_window_finds_the_front():
    ma, ms, g, n = _slab(air=21, shape=(128, 96, 100))
    abl = np.zeros(ma.shape, bool)
    abl[:, :, 21:40] = True
    o, front = S.window(ma, ms, abl, SP, 64 * 0.015, 48 * 0.015, math.radians(22.0), margin=12)
    assert front == 40 and o[2] == 28
    assert o[1] == 16


def test_march_from_the_domain_top_is_beer_lambert_below_the_window_top():
    # a window 40 planes down sees the attenuation of everything above it
    ma, ms, g, n = _slab(shape=(64, 64, 128))
    xc = yc = 32 * 0.015
    B = S.ballistic(ma, ms, n, SP, (0, 0, 40), xc, yc, 0.0, 0.5)
    R = ((1.0 - 1.4) / 2.4) ** 2
    I0 = math.exp(-2.0 * (2 * (0.5 * 0.015) ** 2) / 0.25 ** 2)
    k = np.arange(64)
    np.testing.assert_allclose(B[32, 32, :], I0 * (1 - R) * np.exp(-(0.5 + 3.0) * 0.015 * (40 + k + 1)), rtol=1e-12)
    _, tau = S.ballistic(ma, ms, n, SP, (0, 0, 40), xc, yc, 0.0, 0.5, with_tau=True)
    np.testing.assert_allclose(tau[32, 32, :], (0.5 + 3.0) * 0.015 * (40 + k + 1), rtol=1e-12)


def test_windows_tile_the_crater_from_front_to_surface():
    ma, ms, g, n = _slab(air=21, shape=(128, 96, 200))
    abl = np.zeros(ma.shape, bool)
    abl[54:74, 38:58, 21:150] = True          # a 20-voxel-wide crater down to z = 150
    ws, front, surf = S.windows(ma, ms, abl, SP, 64.5 * 0.015, 48.5 * 0.015, 0.0, margin=12)
    assert front == 150 and surf == 21
    assert [w[2] for w in ws] == [136, 80, 24, 0]
    assert all(w[:2] == (32, 16) for w in ws)
    # consecutive windows overlap by at least the ramp; the stack reaches above the surface
    assert all(a[2] - b[2] <= S.SW - S.RAMP for a, b in zip(ws, ws[1:]))
    assert ws[-1][2] <= surf - S.RAMP
    # a shallow front needs one window
    abl[:] = False
    ws, front, surf = S.windows(ma, ms, abl, SP, 64.5 * 0.015, 48.5 * 0.015, 0.0, margin=12)
    assert front == surf == 21 and [w[2] for w in ws] == [9]


def test_apply_n_is_a_partition_of_unity_in_the_overlap():
    rng = np.random.default_rng(2)
    shape = (80, 70, 150)
    ma = rng.uniform(0.1, 2.0, shape)
    Qs = rng.uniform(0.0, 10.0, shape)
    F = rng.uniform(1.0, 20.0, shape)            # one consistent field predicted by both windows
    E0 = 40.0 / (math.pi * 0.25 ** 2 / 2)
    ws = [(8, 3, 80), (8, 3, 24)]                # overlap z in [80, 88)
    ys = [np.log(F[o[0]:o[0] + 64, o[1]:o[1] + 64, o[2]:o[2] + 64] / (E0 * ma[o[0]:o[0] + 64, o[1]:o[1] + 64,
                                                                            o[2]:o[2] + 64]) + 1e-4) for o in ws]
    out = S.apply_n(ys, ma, Qs, ws, 40.0, 0.5)
    inner = (slice(16, 64), slice(11, 59), slice(32, 136))   # away from the outer faces
    np.testing.assert_allclose(out[inner], F[inner], rtol=1e-10)
    assert out[0, 0, 0] == Qs[0, 0, 0] and out[79, 69, 149] == Qs[79, 69, 149]
    # the scale multiplies the network's part only (F >> 1e-4 E0 mu_a, so the offset is negligible)
    out2 = S.apply_n(ys, ma, Qs, ws, 40.0, 0.5, scale=1.5)
    np.testing.assert_allclose(out2[inner], 1.5 * F[inner], rtol=1e-3)
    assert out2[0, 0, 0] == Qs[0, 0, 0]
    # one window reduces to apply()
    np.testing.assert_allclose(S.apply_n(ys[:1], ma, Qs, ws[:1], 40.0, 0.5), S.apply(ys[0], ma, Qs, ws[0], 40.0, 0.5),
                               rtol=1e-12)


def test_coarse_correct_takes_block_totals_and_keeps_the_shape():
    """Invariants that pin bd_sur_correct's restatement: block totals become R's,
    the shape inside a block is F's, ablated voxels and blocks with an empty
    sum are untouched."""
    rng = np.random.default_rng(4)
    shape, e = (13, 10, 9), 4
    F = rng.uniform(0.0, 2.0, shape)
    R = rng.uniform(0.0, 3.0, shape)
    abl = rng.random(shape) < 0.2
    F[8:, :4, :4] = 0.0            # a block whose composed sum is empty
    R[:4, 4:8, 4:8] = 0.0          # a block whose MC sum is empty
    out = S.coarse_correct(F, R, abl, e)
    np.testing.assert_array_equal(out[abl], F[abl])
    for x in range(0, shape[0], e):
        for y in range(0, shape[1], e):
            for z in range(0, shape[2], e):
                sl = (slice(x, x + e), slice(y, y + e), slice(z, z + e))
                m = ~abl[sl]
                sf, sr = F[sl][m].sum(), R[sl][m].sum()
                if sf > 0 and sr > 0:
                    assert abs(out[sl][m].sum() - sr) <= 1e-12 * sr
                    k = np.flatnonzero(F[sl][m] > 0)
                    ratios = out[sl][m][k] / F[sl][m][k]
                    np.testing.assert_allclose(ratios, ratios[0], rtol=1e-12)
                else:
                    np.testing.assert_array_equal(out[sl], F[sl])
    np.testing.assert_array_equal(out[8:, :4, :4], 0.0)
    np.testing.assert_array_equal(out[:4, 4:8, 4:8][~abl[:4, 4:8, 4:8]], F[:4, 4:8, 4:8][~abl[:4, 4:8, 4:8]])


def test_noise_aware_gate_separates_noise_from_bias():
    """E d_s^2 = e^2 + s^2 and E d_mm^2 = 2 s^2 (e: the candidate's distance from
    the expected field, s: one MC run's noise).  On a synthetic field with
    Poisson-like noise: a candidate that is itself an independent MC run of the
    same noise passes (e = s), one 10 % off in amplitude fails, and the plain
    L2 test fails both once s exceeds the tolerance."""
    from surrogate.coupled import excess, gate_passes, mc_vs_mc_from_lo
    rng = np.random.default_rng(9)
    truth = rng.gamma(2.0, 1.0, 200_000)

    def mc(sig):
        return truth * (1.0 + sig * rng.standard_normal(truth.size))

    def d(a, b):
        return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))

    for sig in (0.02, 0.08):
        m1, m2 = mc(sig), mc(sig)
        d_mm = d(m2, m1)
        assert abs(d_mm - np.sqrt(2) * sig) < 0.05 * np.sqrt(2) * sig
        cand_mc = mc(sig)
        assert gate_passes(d(cand_mc, m1), d_mm, 0.05)
        assert excess(d(cand_mc, m1), d_mm) < 0.6 * sig + 0.01
        biased = 1.10 * truth
        assert not gate_passes(d(biased, m1), d_mm, 0.05)
        assert abs(excess(d(biased, m1), d_mm) - np.sqrt(0.10 ** 2 - sig ** 2)) < 0.02
        # d_mm from an independent run with a quarter of the photons (twice the noise)
        lo = mc(2.0 * sig)
        est = mc_vs_mc_from_lo(d(lo, m1), 4.0, 1.0)
        assert abs(est - np.sqrt(2) * sig) < 0.05 * np.sqrt(2) * sig
    # the plain test rejects an exact copy of the expected field when one run is noisy
    m1 = mc(0.08)
    assert not gate_passes(d(truth, m1), None, 0.05)
