"""Pins of the oracle's voxel Monte Carlo (CPU only).

Closed forms and tables from the reference's verification tier
(knowledge/simulation_lit_comparison/code/verification_benchmarks.py and
consistency_checks.py) plus beam-footprint moments, all with stated
statistical tolerances.  The reference's own recorded engine outputs for the
same cases (tests/golden/reference_engine_values.json) are checked as well,
within the combined standard error.
"""
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_engine_values.json")))["values"]
N_THREADS = 4


def semi_infinite_isotropic_rd(albedo, mu0=1.0, n_quad=400):
    """Chandrasekhar H-function, Rd = 1 - H(mu0) sqrt(1 - a) (verification_benchmarks.py:32-51)."""
    from numpy.polynomial.legendre import leggauss
    x, w = leggauss(n_quad)
    mu = 0.5 * (x + 1.0)
    wt = 0.5 * w
    H = np.ones_like(mu)
    for _ in range(5000):
        Hn = 1.0 / (1.0 - 0.5 * albedo * mu * (wt * H / (mu[:, None] + mu[None, :])).sum(axis=1))
        if np.max(np.abs(Hn - H)) < 1e-14:
            H = Hn
            break
        H = Hn
    Hm = 1.0 / (1.0 - 0.5 * albedo * mu0 * np.sum(wt * H / (mu0 + mu)))
    return float(1.0 - Hm * np.sqrt(1.0 - albedo))


def _batches(oracle, nb, n, **kw):
    return [oracle.run_voxel_mc3d(n_photons=n, seed=1000 + b, n_threads=N_THREADS, **kw) for b in range(nb)]


def _mean_se(v):
    v = np.asarray(v, float)
    return float(v.mean()), float(v.std(ddof=1) / np.sqrt(len(v)))


def test_v1_slab_van_de_hulst(oracle):
    runs = _batches(oracle, 8, 100000, spot_dia=0.01, Lx=20.0, Ly=20.0, Lz=0.2, Nx=20, Ny=20, Nz=10,
                    layers_config=[{'z_end': 0.2, 'mu_a': 1.0, 'mu_s': 9.0, 'g': 0.75, 'n': 1.0}])
    rd, se_r = _mean_se([r['spectro_raw']['R_diff'] for r in runs])
    tt, se_t = _mean_se([r['spectro_raw']['Transmission'] for r in runs])
    assert abs(rd - 0.09739) < 4 * se_r + 1e-5
    assert abs(tt - 0.66096) < 4 * se_t + 1e-5
    g = GOLD["V1a"]["model"]
    assert abs(rd - g) < 4 * np.hypot(se_r, 0.0005)


def test_v1c_fresnel_slab_and_specular(oracle):
    n, mu_a, d = 1.5, 1.0, 0.5
    R = ((n - 1.0) / (n + 1.0)) ** 2
    e1, e2 = np.exp(-mu_a * d), np.exp(-2.0 * mu_a * d)
    T_ref = (1 - R) ** 2 * e1 / (1 - R ** 2 * e2)
    Rd_ref = R * (1 - R) ** 2 * e2 / (1 - R ** 2 * e2)
    N = 400000
    r = oracle.run_voxel_mc3d(n_photons=N, spot_dia=0.01, Lx=2.0, Ly=2.0, Lz=0.55, Nx=4, Ny=4, Nz=11, seed=7,
                              n_threads=N_THREADS,
                              layers_config=[{'z_end': 0.5, 'mu_a': mu_a, 'mu_s': 0.0, 'g': 0.0, 'n': n},
                                             {'z_end': 0.55, 'mu_a': 0.0, 'mu_s': 0.0, 'g': 0.0, 'n': 1.0}])
    s = r['spectro_raw']
    assert abs(s['Transmission'] - T_ref) < 4 * np.sqrt(T_ref * (1 - T_ref) / N)
    assert abs(s['R_diff'] - Rd_ref) < 4 * np.sqrt(Rd_ref * (1 - Rd_ref) / N)
    assert abs(s['R_spec'] - R) < 1e-9
    assert abs(r['energy_closure'] - 1.0) < 1e-9  # no roulette/lateral here: closes exactly


def test_v1f_oblique_refraction(oracle):
    n, mu_a, d, th = 1.5, 1.0, 1.0, np.radians(45.0)
    st = np.sin(th) / n
    ct = np.sqrt(1 - st ** 2)
    ci = np.cos(th)
    R45 = 0.5 * (((ci - n * ct) / (ci + n * ct)) ** 2 + ((ct - n * ci) / (ct + n * ci)) ** 2)
    T_ref = (1 - R45) * np.exp(-mu_a * d / ct)
    N = 200000
    r = oracle.run_voxel_mc3d(n_photons=N, spot_dia=0.2, theta_inc_deg=45.0, Lx=8.0, Ly=8.0, Lz=d, Nx=40, Ny=40,
                              Nz=20, seed=9, n_threads=N_THREADS,
                              layers_config=[{'z_end': d, 'mu_a': mu_a, 'mu_s': 0.0, 'g': 0.0, 'n': n}])
    assert abs(r['spectro_raw']['Transmission'] - T_ref) < 4 * np.sqrt(T_ref * (1 - T_ref) / N)
    assert abs(r['spectro_raw']['R_spec'] - R45) < 1e-9


def test_v2_semi_infinite_h_function(oracle):
    runs = _batches(oracle, 6, 60000, spot_dia=0.01, Lx=20.0, Ly=20.0, Lz=4.0, Nx=20, Ny=20, Nz=8,
                    layers_config=[{'z_end': 4.0, 'mu_a': 1.0, 'mu_s': 9.0, 'g': 0.0, 'n': 1.0}])
    rd, se = _mean_se([r['spectro_raw']['R_diff'] for r in runs])
    ref = semi_infinite_isotropic_rd(0.9)
    assert ref == pytest.approx(GOLD["V2"]["reference_value"], rel=1e-9)
    assert abs(rd - ref) < 4 * se + 2e-4


def test_v2b_mismatched_total_reflectance(oracle):
    runs = _batches(oracle, 6, 60000, spot_dia=0.01, Lx=20.0, Ly=20.0, Lz=4.0, Nx=20, Ny=20, Nz=8,
                    layers_config=[{'z_end': 4.0, 'mu_a': 1.0, 'mu_s': 9.0, 'g': 0.0, 'n': 1.5}])
    tot, se = _mean_se([r['spectro_raw']['R_diff'] + r['spectro_raw']['R_spec'] for r in runs])
    assert abs(tot - 0.2600) < 4 * se + 0.0017


@pytest.mark.parametrize("ang", [0, 30, 60])
def test_c3_oblique_specular_and_mean_depth(oracle, ang):
    """consistency_checks.py:149-164: R_spec exact to 1e-9, mean absorption depth cos(theta_t)/mu_a."""
    n, mu_a, dz, Nz = 1.55, 50.0, 0.002, 250
    r = oracle.run_voxel_mc3d(n_photons=100000, spot_dia=0.2, theta_inc_deg=float(ang), Lx=4.0, Ly=4.0,
                              Lz=Nz * dz, Nx=40, Ny=40, Nz=Nz, seed=11 + ang, n_threads=N_THREADS,
                              layers_config=[{'z_end': Nz * dz, 'mu_a': mu_a, 'mu_s': 0.0, 'g': 0.0, 'n': n}])
    th = np.radians(ang)
    st = np.sin(th) / n
    ct = np.sqrt(1 - st * st)
    ci = np.cos(th)
    R = 0.5 * (((ci - n * ct) / (ci + n * ct)) ** 2 + ((ct - n * ci) / (ct + n * ci)) ** 2)
    assert abs(r['spectro_raw']['R_spec'] - R) < 1e-9
    assert r['spectro_raw']['R_spec'] == pytest.approx(GOLD[f"C3-R-{ang}"]["model"], abs=1e-9)
    A = r['A_3d'].sum(axis=(0, 1))
    zbar = float((A * r['grid_z']).sum() / A.sum())
    assert zbar == pytest.approx(ct / mu_a, rel=0.02)


@pytest.mark.parametrize("profile", ["gaussian", "tophat", "kim_exp_disk", "kim_gauss_disk"])
def test_beam_footprint_moments(oracle, profile):
    """voxel_mc_3d.py:101-120,152-161: photons absorbed in a 1-voxel black surface layer
    map the launch footprint; its mean squared radius has a closed form per profile."""
    N, L, nx = 100000, 2.0, 200
    spot = 0.8
    r = oracle.run_voxel_mc3d(n_photons=N, spot_dia=spot, Lx=L, Ly=L, Lz=0.5, Nx=nx, Ny=nx, Nz=1, seed=21,
                              n_threads=N_THREADS, beam_profile=profile,
                              layers_config=[{'z_end': 0.5, 'mu_a': 1e4, 'mu_s': 0.0, 'g': 0.0, 'n': 1.0}])
    A = r['A_3d'][:, :, 0]
    x = r['grid_x'] - L / 2
    r2 = x[:, None] ** 2 + x[None, :] ** 2
    m2 = float((A * r2).sum() / A.sum())
    w0 = spot / 2.0
    if profile == "gaussian":
        exp_m2 = 2 * (w0 / 2) ** 2
    elif profile == "tophat":
        exp_m2 = w0 ** 2 / 2
    else:
        from scipy.integrate import quad
        f = (lambda s: np.exp(-2 * s)) if profile == "kim_exp_disk" else (lambda s: np.exp(-2 * s * s))
        # rn = sqrt(u) accepted with prob f(rn): radial density ~ rn f(rn) on [0,1]
        num = quad(lambda s: s ** 3 * f(s), 0, 1)[0]
        den = quad(lambda s: s * f(s), 0, 1)[0]
        exp_m2 = w0 ** 2 * num / den
    dxv = L / nx
    assert m2 == pytest.approx(exp_m2 + dxv ** 2 / 6, rel=0.02)


def test_energy_closure_reference_fixture(oracle):
    """tests/test_solvers.py:13-39 of the reference."""
    r = oracle.run_voxel_mc3d(power=2.0, n_photons=2000, spot_dia=0.20, Lx=1.0, Ly=1.0, Lz=0.8, Nx=10, Ny=10,
                              Nz=8, seed=3, n_threads=N_THREADS)
    assert abs(r['energy_closure'] - 1.0) < 0.02
    assert np.all(r['A_3d'] >= 0) and r['A_3d'].shape == (10, 10, 8)
    s = r['spectro_raw']
    assert r['lateral_escape_z'].sum() == pytest.approx(s['Lateral_loss'], rel=1e-12, abs=1e-15)


def test_r_max_cut_at_launch(oracle):
    """A Gaussian beam truncated at r_max loses exp(-r^2 / (2 sigma^2)) of its photons at launch."""
    N, spot, rmax = 200000, 0.8, 0.25
    r = oracle.run_voxel_mc3d(n_photons=N, spot_dia=spot, Lx=4.0, Ly=4.0, Lz=0.2, Nx=20, Ny=20, Nz=2, seed=5,
                              r_max=rmax, n_threads=N_THREADS,
                              layers_config=[{'z_end': 0.2, 'mu_a': 1e3, 'mu_s': 0.0, 'g': 0.0, 'n': 1.0}])
    sig = spot / 4.0
    p = np.exp(-rmax ** 2 / (2 * sig ** 2))
    assert abs(r['spectro_raw']['Lateral_loss'] - p) < 4 * np.sqrt(p * (1 - p) / N)


def test_launch_axis_x_matches_z_statistics(oracle):
    """launch_axis='x' (voxel_mc_3d.py:165-177,538-548) is the same physics rotated."""
    lay = [{'z_end': 0.3, 'mu_a': 0.5, 'mu_s': 10.0, 'g': 0.8, 'n': 1.4},
           {'z_end': 1.0, 'mu_a': 1.0, 'mu_s': 5.0, 'g': 0.9, 'n': 1.33}]
    rz = oracle.run_voxel_mc3d(n_photons=100000, spot_dia=0.3, Lx=4.0, Ly=4.0, Lz=1.0, Nx=20, Ny=20, Nz=25,
                               layers_config=lay, seed=1, n_threads=N_THREADS)
    rx = oracle.run_voxel_mc3d(n_photons=100000, spot_dia=0.3, Lx=1.0, Ly=4.0, Lz=4.0, Nx=25, Ny=20, Nz=20,
                               layers_config=lay, launch_axis='x', seed=2, n_threads=N_THREADS)
    for key in ('R_diff', 'Absorption', 'Transmission'):
        assert abs(rz['spectro_raw'][key] - rx['spectro_raw'][key]) < 0.006
    assert rz['spectro_raw']['R_spec'] == pytest.approx(rx['spectro_raw']['R_spec'], abs=1e-12)


def test_visit_counts_closed_form(oracle):
    """The merged-tally visit counts: through a transparent, index-matched column every photon
    visits every voxel once (path-length adds only) and no collision happens."""
    shape = (1, 1, 12)
    vL, vA = np.zeros(shape), np.zeros(shape)
    props = {'mu_a': np.zeros(shape), 'mu_s': np.zeros(shape), 'g': np.zeros(shape), 'n': np.ones(shape)}
    r = oracle.run_voxel_mc3d(n_photons=5000, spot_dia=1e-4, Lx=0.1, Ly=0.1, Lz=1.2, Nx=1, Ny=1, Nz=12,
                              custom_3d_props=props, seed=2, n_threads=4, visit_grids=(vL, vA))
    assert r['spectro_raw']['Transmission'] == pytest.approx(1.0)
    np.testing.assert_array_equal(vL, np.full(shape, 5000.0))
    assert not vA.any()


def test_tally_concentration_on_the_bench_state(oracle):
    """How concentrated the transport's tally additions are on the 256x256x192 layered stack after
    three epochs (150 000 photons): the share held by the densest voxels and by the best dense box
    a per-block window in LDS could index, and the hits per voxel one 256-photon block makes there.
    Measured: about 163 merged additions per photon; the densest 20 480 voxels (160 KB of float2
    tallies) take 34 % of them, the best box of at most 20 480 voxels 27 %, and one block of 256
    photons adds about 0.7 times per voxel there, so a block-private window would flush about as
    many additions as it saves.  The bounds below keep that finding true."""
    from simulation.workloads import layered_stack
    kw = layered_stack()
    okw = dict(kw)
    pg = okw.pop('plume_grid')
    eng = oracle.OracleEngine(mc_threads=8, plume_grid=pg, **okw)
    for _ in range(3):
        eng.step_epoch()
    opt = {k: v.copy() for k, v in eng.optical.items()}
    opt['mu_a'] = eng.optical['mu_a'] + eng._from_plume(eng.plume.mu_a_plume)
    opt['mu_s'] = eng.optical['mu_s'] + eng._from_plume(eng.plume.mu_s_plume)
    opt['g'][(eng._from_plume(eng.plume.mu_ext_plume) > 0.001) & eng.ablated] = 0.80
    shape = eng.T.shape
    vL, vA = np.zeros(shape), np.zeros(shape)
    N = 150_000
    oracle.run_voxel_mc3d(power=kw['power'], n_photons=N, spot_dia=kw['spot_dia'], theta_inc_deg=kw['theta_inc_deg'],
                          Lx=kw['Lx'], Ly=kw['Ly'], Lz=kw['Lz'], Nx=shape[0], Ny=shape[1], Nz=shape[2],
                          custom_3d_props=opt, beam_center=kw['beam_center'], seed=12345, n_threads=8,
                          visit_grids=(vL, vA))
    v = vL + vA
    tot = float(v.sum())
    cs = np.cumsum(np.sort(v.ravel())[::-1])
    blocks = N / 256.0
    print(f"merged tally additions per photon {tot / N:.1f} (path {vL.sum() / N:.1f}, absorbed {vA.sum() / N:.1f})")
    for k in (4096, 10240, 20480, 65536):
        print(f"densest {k} voxels: share {cs[k - 1] / tot:.3f}, hits per voxel per 256-photon block "
              f"{cs[k - 1] / k / blocks:.2f}")
    S = np.zeros(tuple(n + 1 for n in shape))
    S[1:, 1:, 1:] = v.cumsum(0).cumsum(1).cumsum(2)
    best = (0.0, None)
    for bz in (8, 12, 16, 20, 24, 32, 48, 64):
        for b in (12, 16, 20, 24, 32, 48):
            if b * b * bz > 20480:
                continue
            box = (S[b:, b:, bz:] - S[:-b, b:, bz:] - S[b:, :-b, bz:] - S[b:, b:, :-bz] + S[:-b, :-b, bz:]
                   + S[:-b, b:, :-bz] + S[b:, :-b, :-bz] - S[:-b, :-b, :-bz])
            m = float(box.max())
            if m > best[0]:
                best = (m, (b, b, bz))
    print(f"best box of <= 20480 voxels {best[1]}: share {best[0] / tot:.3f}, hits per voxel per block "
          f"{best[0] / np.prod(best[1]) / blocks:.2f}")
    assert 100 < tot / N < 250
    assert cs[20479] / tot < 0.45 and best[0] / tot < 0.35
    assert cs[20479] / 20480 / blocks < 1.0
