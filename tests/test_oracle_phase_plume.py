"""Pins of the oracle's phase/property update, removal glue and plume kernels (CPU only)."""
import numpy as np
import pytest


def _phase_inputs(rng, shape):
    n = int(np.prod(shape))
    H = np.where(rng.random(shape) < 0.5, rng.uniform(0, 3.0, shape), 0.0)
    hl = np.full(shape, 2.26)
    T = rng.uniform(20.0, 400.0, shape)
    chi = rng.uniform(0, 1, shape)
    dry = (rng.random(shape) < 0.4).astype(np.uint8)
    abl = (rng.random(shape) < 0.1).astype(np.uint8)
    opt = {k: rng.uniform(0.5, 2.0, shape) for k in ("mu_a", "mu_s", "g", "n")}
    k = rng.uniform(0.2e-3, 0.6e-3, shape)
    rc = rng.uniform(2e-3, 4e-3, shape)
    return n, H, hl, T, chi, dry, abl, opt, k, rc


def test_phase_update_truth_table(oracle):
    """coupled_ablation.py:77-180 branch by branch, with scalar dry state and char state."""
    rng = np.random.default_rng(11)
    shape = (5, 6, 7)
    n, H, hl, T, chi, dry, abl, opt, k, rc = _phase_inputs(rng, shape)
    wet = (opt["mu_a"].copy(), k.copy(), rc.copy())
    ds = dict(mu_a=5.0, mu_s=7.0, g=0.6, n=1.3, k=0.25e-3, rho_cp=1.5e-3, t_remove=350.0)
    cs = dict(mu_a=50.0, mu_s=10.0, k=0.8e-3, rho_cp=1.0e-3, t_remove=300.0)
    alpha = 0.35 / (5e-5 * 3 / 0.015 ** 2)
    H0, dry0 = H.copy(), dry.copy()
    rem = np.zeros(shape, np.uint8)
    o = {kk: v.copy() for kk, v in opt.items()}
    oracle.phase_update(H, hl, T, chi, dry, abl, o["mu_a"], o["mu_s"], o["g"], o["n"], k, rc,
                        wet[0], wet[1], wet[2], ds, cs, 0.95, alpha, rem)
    a = abl == 1
    assert not rem[a].any()
    boiled = (~a) & (dry0 == 0) & (H0 >= 2.26)
    boiling = (~a) & (dry0 == 0) & (H0 > 0) & (H0 < 2.26)
    assert np.all(dry[boiled] == 1) and np.all(H[boiled] == 0.0)
    frac = H0[boiling] / 2.26
    np.testing.assert_allclose(o["mu_a"][boiling], wet[0][boiling] + (5.0 - wet[0][boiling]) * frac, rtol=1e-14)
    np.testing.assert_allclose(k[boiling], wet[1][boiling] + (0.25e-3 - wet[1][boiling]) * frac, rtol=1e-14)
    assert not rem[boiling].any()
    nowdry = (~a) & (dry == 1)
    # char interpolation invariant (reference test_carbonization_property_interpolation)
    np.testing.assert_allclose(o["mu_a"][nowdry], 5.0 + chi[nowdry] * 45.0, rtol=1e-13)
    np.testing.assert_allclose(o["mu_s"][nowdry], 7.0 + chi[nowdry] * 3.0, rtol=1e-13)
    np.testing.assert_allclose(k[nowdry], 0.25e-3 + chi[nowdry] * 0.55e-3, rtol=1e-13)
    rc_exp = np.maximum(1.5e-3 + chi[nowdry] * (1.0e-3 - 1.5e-3), k[nowdry] / alpha)
    np.testing.assert_allclose(rc[nowdry], rc_exp, rtol=1e-13)
    np.testing.assert_array_equal(rem[nowdry], ((T[nowdry] >= 300.0) | (chi[nowdry] >= 0.95)).astype(np.uint8))
    # g/n only change on the wet->dry switch
    np.testing.assert_array_equal(o["g"][boiled], 0.6)
    untouched = (~a) & (dry0 == 0) & (H0 <= 0.0)
    for kk in ("mu_a", "mu_s", "g", "n"):
        np.testing.assert_array_equal(o[kk][untouched], opt[kk][untouched])
    assert not rem[untouched].any()


def test_phase_update_without_char_and_per_voxel_dry_state(oracle):
    rng = np.random.default_rng(12)
    shape = (4, 4, 6)
    n, H, hl, T, chi, dry, abl, opt, k, rc = _phase_inputs(rng, shape)
    ds = {kk: rng.uniform(1, 2, shape) for kk in ("mu_a", "mu_s", "g", "n")}
    ds.update(k=rng.uniform(1e-4, 5e-4, shape), rho_cp=rng.uniform(1e-3, 3e-3, shape),
              t_remove=rng.uniform(100, 400, shape))
    rem = np.zeros(shape, np.uint8)
    dry0, H0 = dry.copy(), H.copy()
    oracle.phase_update(H, hl, T, chi, dry, abl, opt["mu_a"], opt["mu_s"], opt["g"], opt["n"], k, rc,
                        opt["mu_a"].copy(), k.copy(), rc.copy(), ds, None, 1.0, 1.0, rem)
    drynow = (abl == 0) & (dry == 1)
    np.testing.assert_array_equal(rem[drynow], (T[drynow] >= ds["t_remove"][drynow]).astype(np.uint8))
    boiled = (abl == 0) & (dry0 == 0) & (H0 >= 2.26)
    np.testing.assert_array_equal(opt["n"][boiled], ds["n"][boiled])
    assert not rem[(abl == 0) & (dry == 0)].any()


def test_single_stage(oracle):
    rng = np.random.default_rng(4)
    shape = (3, 4, 5)
    H = rng.uniform(0, 3, shape)
    hl = np.full(shape, 1.5)
    T = rng.uniform(80, 120, shape)
    abl = (rng.random(shape) < 0.3).astype(np.uint8)
    rem = np.zeros(shape, np.uint8)
    oracle.single_stage(H, hl, T, 100.0, abl, rem)
    np.testing.assert_array_equal(rem, ((abl == 0) & (H >= hl) & (T >= 100.0)).astype(np.uint8))


# ----------------------------------------------------------------- plume
def _diffuse(oracle, C, void, n, dx, dy, dz, dt, Dxy, Dz, vz=0.0, decay=1.0):
    Cn = np.empty_like(C)
    for _ in range(n):
        oracle.plume_diffuse(C, Cn, void, *C.shape, dx, dy, dz, dt, Dxy, Dz, vz, decay)
        C, Cn = Cn, C
    return C


def test_plume_variance_grows_2Dt(oracle):
    """For the explicit scheme away from walls the discrete variance of a point release grows
    by exactly 2 D dt per step along each axis (sum of the stencil's second moment)."""
    N = 41
    C = np.zeros((N, N, N), np.float32)
    C[20, 20, 20] = 1.0
    void = np.ones((N, N, N), np.uint8)
    dx, dz, dt, Dxy, Dz, n = 0.02, 0.015, 1e-5, 1.0, 2.5, 30
    C = _diffuse(oracle, C, void, n, dx, dx, dz, dt, Dxy, Dz)
    x = (np.arange(N) - 20) * dx
    z = (np.arange(N) - 20) * dz
    m = C.astype(np.float64)
    assert m.sum() == pytest.approx(1.0, rel=1e-5)
    vx = float((m.sum(axis=(1, 2)) * x ** 2).sum() / m.sum())
    vz = float((m.sum(axis=(0, 1)) * z ** 2).sum() / m.sum())
    assert vx == pytest.approx(2 * Dxy * dt * n, rel=1e-4)
    assert vz == pytest.approx(2 * Dz * dt * n, rel=1e-4)


def test_plume_confinement_mass_and_decay(oracle):
    """plume.py:49-57,62-71: zero outside the void and on the domain boundary, zero-flux
    walls conserve mass inside a closed cavity, decay multiplies a uniform field exactly."""
    shape = (15, 15, 15)
    void = np.zeros(shape, np.uint8)
    void[5:10, 5:10, 5:10] = 1
    C = np.zeros(shape, np.float32)
    C[7, 7, 7] = 5.0
    C2 = _diffuse(oracle, C.copy(), void, 10, 0.03, 0.03, 0.02, 5e-5, 0.5, 0.5)
    assert np.all(C2[void == 0] == 0.0)
    assert C2[7, 7, 7] < 5.0 and C2[7, 7, 8] > 0.0
    assert float(C2.sum(dtype=np.float64)) == pytest.approx(5.0, rel=1e-6)
    U = np.where(void == 1, np.float32(2.0), np.float32(0.0)).astype(np.float32)
    U2 = _diffuse(oracle, U, void, 1, 0.03, 0.03, 0.02, 5e-5, 0.5, 0.5, decay=float(np.exp(-5e-5 / 0.01)))
    np.testing.assert_allclose(U2[void == 1], 2.0 * np.exp(-5e-5 / 0.01), rtol=1e-6)
    full = np.ones(shape, np.uint8)
    B = _diffuse(oracle, np.ones(shape, np.float32), full, 1, 0.03, 0.03, 0.02, 5e-5, 0.5, 0.5)
    assert np.all(B[0] == 0) and np.all(B[-1] == 0) and np.all(B[:, :, 0] == 0)


def test_plume_upwind_drift_moves_mass_toward_surface(oracle):
    """plume.py:77: -vz (C - C_t)/dz carries mass towards k = 0 at speed vz."""
    N = 31
    C = np.zeros((5, 5, N), np.float32)
    C[2, 2, 20] = 1.0
    void = np.ones((5, 5, N), np.uint8)
    dz, dt, vz, n = 0.02, 1e-5, 50.0, 20
    C = _diffuse(oracle, C, void, n, 0.02, 0.02, dz, dt, 0.0, 0.0, vz=vz)
    m = C[2, 2].astype(np.float64)
    zbar = float((m * np.arange(N)).sum() / m.sum())
    assert zbar == pytest.approx(20 - vz * dt * n / dz, rel=1e-4)


def test_plume_anisotropy_reference_fixture(oracle):
    """tests/test_plume.py:230-266 of the reference (x/y symmetry 1e-4, D_z = 4 D_xy)."""
    shape = (31, 31, 31)
    C = np.zeros(shape, np.float32)
    C[15, 15, 15] = 0.075
    C = _diffuse(oracle, C, np.ones(shape, np.uint8), 20, 0.02, 0.02, 0.02, 2e-5, 1.0, 4.0)
    assert C[17, 15, 15] == pytest.approx(C[15, 17, 15], rel=1e-4)
    assert C[15, 15, 17] > 2.0 * C[17, 15, 15]


def test_separable_convolution_equals_direct(oracle):
    """plume.py:83-139 vs a brute-force 3-D convolution with the outer-product kernel."""
    rng = np.random.default_rng(8)
    raw = np.zeros((14, 13, 12), np.float32)
    imp = (rng.random((3, 4, 2)) < 0.6).astype(np.float32)
    kx = rng.random(5).astype(np.float32)
    ky = rng.random(3).astype(np.float32)
    kz = rng.random(7).astype(np.float32)
    x0, y0, z0 = -1, 4, 5
    oracle.plume_convolve(raw, imp, kx, ky, kz, 0.9, x0, y0, z0)
    ref = np.zeros((14 + 20, 13 + 20, 12 + 20))
    K = kx[:, None, None].astype(np.float64) * ky[None, :, None] * kz[None, None, :]
    for (a, b, c) in zip(*np.nonzero(imp)):
        ref[10 + x0 + a:10 + x0 + a + 5, 10 + y0 + b:10 + y0 + b + 3, 10 + z0 + c:10 + z0 + c + 7] += 0.9 * K
    np.testing.assert_allclose(raw, ref[10:24, 10:23, 10:22], rtol=2e-6, atol=1e-7)


def test_plume_field_kernels_and_injection(oracle):
    """tests/test_plume.py:10-21,110-179 of the reference restated on the oracle plume."""
    p = oracle.OraclePlume(shape=(21, 21, 21), dx_mm=0.03, dy_mm=0.03, dz_mm=0.02, sigma_mm=0.04)
    c = tuple(s // 2 for s in p.kernel.shape)
    assert p.kernel[c] == pytest.approx(1.0) and p.kernel.max() <= 1.0 + 1e-7
    q = oracle.OraclePlume(shape=(15, 15, 15), model='diffusion', delta_mu_per_voxel=5.0, D_plume=0.5,
                           v_drift=10.0, blast_expansion=False)
    newly = np.zeros((15, 15, 15), bool)
    newly[7, 7, 7] = True
    q.inject(newly)
    assert q.raw_field[7, 7, 7] == pytest.approx(5.0) and np.sum(q.raw_field > 0) == 1
    b = oracle.OraclePlume(shape=(25, 25, 25), model='diffusion', delta_mu_per_voxel=1.0, blast_expansion=True)
    void = np.zeros((25, 25, 25), bool)
    void[8:17, 8:17, 5:20] = True
    nb = np.zeros((25, 25, 25), bool)
    nb[12, 12, 15] = True
    b.inject(nb, void_mask=void)
    assert b.raw_field[12, 12, 15] < 0.20
    assert np.sum((b.raw_field > 1e-4) & void) > 10
    assert np.all(b.raw_field[~void] == 0.0)
    # tanh saturation stays below mu_max and approaches it
    s = oracle.OraclePlume(shape=(5, 5, 5), model='diffusion', wavelength_nm=450)
    s.raw_field[:] = 50.0
    s.sync_extinction(np.ones((5, 5, 5), bool))
    assert np.all(s.mu_ext_plume <= s.mu_max) and s.mu_ext_plume.max() == pytest.approx(s.mu_max, rel=1e-6)
