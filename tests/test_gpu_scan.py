"""The scanned-beam liver benchmark (validation_benchmarks.py:83-141) on the device:
the DOP step against the oracle, and one of the reference's recorded cases
(tests/golden/shimojo_reference.json, from its shimojo_cross_sections.json)."""
import json
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.loads((pathlib.Path(__file__).parent / 'golden' / 'shimojo_reference.json').read_text())


def test_damage_optics_matches_oracle():
    """bd_damage_optics against the DOP loop body in float64 on the same stored
    Omega, ablated voxels untouched: equal to the FP32 rounding of the FP64 value
    (one ulp allowed for the device exp)."""
    from oracle.oracle import damage_optics
    from simulation.scan_suite import LIVER, build_engine
    eng, _ = build_engine(10.0, 1.5, n_photons=2000, Lx=3.0, Ly=2.0, Lz=1.5)
    rng = np.random.default_rng(5)
    shape = (eng.Nx, eng.Ny, eng.Nz)
    Om = rng.exponential(1.0, shape) * (rng.random(shape) < 0.6)
    Om[0, 0, :4] = [0.0, np.log(2.0), 60.0, 1e-7]
    eng.Omega = Om
    abl = rng.random(shape) < 0.2
    eng.ablated = abl
    mu_s0 = np.asarray(eng.optical['mu_s'], dtype=np.float64)
    eng.apply_damage_optics(LIVER['musp_native'], LIVER['musp_coag'], LIVER['g'])
    got = np.asarray(eng.optical['mu_s'], dtype=np.float32)
    want = damage_optics(mu_s0, np.asarray(eng.Omega, dtype=np.float32), abl, LIVER['musp_native'],
                         LIVER['musp_coag'], LIVER['g']).astype(np.float32)
    np.testing.assert_array_max_ulp(got, want, maxulp=1)
    assert np.array_equal(got[abl], mu_s0[abl].astype(np.float32))
    with pytest.raises(Exception):
        eng.apply_damage_optics(0.5, 1.8, 1.0)


def test_host_and_device_dop_scans_agree():
    """The reference's host loop body (download Omega, edit optical['mu_s'], upload)
    and the device DOP step drive the same scan to the same state, up to voxels at
    a threshold."""
    from simulation.scan_suite import simulate_scan
    kw = dict(n_photons=4000, Lx=4.0, Ly=3.0, Lz=2.0, scan_margin=0.5, t_off=1.0, dt_sub=0.002, dt_epoch=0.01)
    a = simulate_scan(50.0, 2.0, host_dop=False, **kw)
    b = simulate_scan(50.0, 2.0, host_dop=True, **kw)
    assert a['epochs_on'] == b['epochs_on'] == 100
    # the two mu_s agree to an ulp and the tallies' float atomics sum in any order,
    # so a voxel right at a threshold may flip
    for key in ('xz_coag', 'xz_ablated'):
        diff = int((a[key] ^ b[key]).sum())
        assert diff <= max(2, 0.02 * int((a[key] | b[key]).sum())), (key, diff)
    assert a['xz_coag'].sum() > 0


@pytest.mark.parametrize('power,speed', [(10.0, 1.5), (50.0, 1.0)])
def test_scan_case_matches_reference_record(power, speed):
    """Two of the six recorded cases against the reference's own runs: 10 W at
    1.5 mm/s on the 12 x 12 x 10 mm block (3 334 scan epochs, 10 s off) and 50 W at
    1.0 mm/s on the 24 x 12 x 10 mm block (5 000 scan epochs, 30 s off), each epoch
    20 000 photons with DOP.  All six cases reproduce the recorded metrics exactly
    (``python -m simulation.scan_suite --all``); the bar here is one voxel (0.1 mm)
    on each metric and 0.95 intersection over union of both cross-sections."""
    from simulation.scan_suite import case_key, compare, run_case
    m = run_case(power, speed)
    d = compare(m, GOLD[case_key(power, speed)])
    print({k: round(v, 3) for k, v in d.items()})
    if power < 20.0:
        assert m['vap_depth'] == 0.0
    for k in ('vap_depth', 'vap_width', 'coag_depth', 'coag_width'):
        assert abs(d[k]) <= 0.1 + 1e-9, (k, d)
    assert d['iou_coag'] >= 0.95 and d['iou_ablated'] >= 0.95, d


KIM = json.loads((pathlib.Path(__file__).parent / 'golden' / 'kim_reference.json').read_text())


def test_kim_lesion_growth_matches_reference_record():
    """kim_benchmark.py:46-118: 1 250 epochs of 2 ms on the 160 x 160 x 80 slab of
    25 um voxels, 15 W from a 1.2 mm exp(-2 r/R) disk, against the reference's own
    run.  The fitted exposure times (when the mean 55 C and 200 C lesion depths
    reach the measured means) within 2 %, the depths there within one voxel, the
    55 C top-view area within 10 % and the 200 C area (26-34 voxels, at the noise
    of the peak temperature) within 40 %."""
    from simulation.kim_suite import fit, simulate_kim
    got = fit(simulate_kim(beam_profile='kim_exp_disk'))
    ref = KIM['kim_exp_disk']
    print(got)
    for grp, area_rel in (('coag', 0.10), ('abl', 0.40)):
        assert got[grp]['t_fit_s'] == pytest.approx(ref[grp]['t_fit_s'], rel=0.02), grp
        assert abs(got[grp]['depth'] - ref[grp]['depth']) <= 0.025 + 1e-9, grp
        assert got[grp]['area'] == pytest.approx(ref[grp]['area'], rel=area_rel), grp


KELLER = json.loads((pathlib.Path(__file__).parent / 'golden' / 'keller_reference.json').read_text())


def test_keller_damage_thresholds_match_reference_record():
    """keller_benchmark.py:54-122: the fluence at which the surface Arrhenius integral
    of the water column reaches 1, by 28 bisection steps of a 25 ms / 4.3 ms pulse plus
    60 ms of cooling (about 160 000 device sub-steps per threshold), for both kinetics
    sets, against the reference's recorded thresholds.  Deterministic; the FP32
    temperature moves the crossing by far less than the 1e-3 bar."""
    from simulation.keller_suite import run_all
    got = run_all()
    print(got)
    assert sorted(got) == sorted(KELLER)
    for k, v in KELLER.items():
        assert got[k] == pytest.approx(v, rel=1e-3), k


def test_arrhenius_with_prefactor_beyond_fp32_matches_oracle():
    """Henriques-Moritz skin kinetics (A = 3.1e98 /s, Ea = 6.3e5 J/mol) exceed FP32; the
    device evaluates exp(ln A - Ea/RT) and must agree with the FP64 oracle
    (coupled_ablation.py:33-40), as it does for A within range."""
    from oracle import oracle as O
    from simulation.coupled_ablation import R_GAS, _accumulate_arrhenius
    rng = np.random.default_rng(3)
    shape = (8, 8, 16)
    T = 55.0 + 40.0 * rng.random(shape)
    abl = rng.random(shape) < 0.1
    for A, Ea in ((3.1e98, 6.3e5), (7.39e37, 2.577e5)):
        om_gpu = 0.01 * rng.random(shape)
        om_ref = om_gpu.astype(np.float32).astype(np.float64).copy()
        _accumulate_arrhenius(T, om_gpu, abl, A, Ea / R_GAS, 1e-3, *shape)
        O.arrhenius(T.astype(np.float32).astype(np.float64), om_ref, abl, A, Ea / R_GAS, 1e-3)
        assert np.all(np.isfinite(om_gpu))
        np.testing.assert_allclose(om_gpu, om_ref, rtol=2e-4, atol=1e-7)
        assert np.array_equal(om_gpu[abl], om_ref[abl])


def _char_inputs(shape, seed):
    rng = np.random.default_rng(seed)
    T = 40.0 + 40.0 * rng.random(shape)   # spans the cold cut-off (about 50 C for A = 3.1e98, dt = 50 us)
    dry = rng.random(shape) < 0.6
    abl = (rng.random(shape) < 0.1) & ~dry
    chi = 0.3 * rng.random(shape)
    return T, dry, abl, chi


def test_char_with_prefactor_beyond_fp32_matches_oracle():
    """_accumulate_char (coupled_ablation.py:43-58) with the skin pre-factor A = 3.1e98 /s (beyond
    FP32): exp(ln A - Ea/RT) on the device against the FP64 oracle, over dry cells on both sides of
    the cold cut-off below which rate * dt < 2^-26 leaves chi unchanged in FP32."""
    from oracle import oracle as O
    from simulation.coupled_ablation import R_GAS, _accumulate_char
    shape = (12, 10, 40)
    T, dry, abl, chi = _char_inputs(shape, 5)
    A, EaR, dt = 3.1e98, 6.3e5 / R_GAS, 5e-5
    chi_g = chi.copy()
    chi_o = chi.astype(np.float32).astype(np.float64)
    _accumulate_char(T, chi_g, dry, abl, A, EaR, dt, *shape)
    O.char(T.astype(np.float32).astype(np.float64), chi_o, dry.astype(np.uint8), abl.astype(np.uint8), A, EaR, dt)
    assert np.all(np.isfinite(chi_g))
    moved = np.abs(chi_o - chi.astype(np.float32)) > 1e-3
    assert moved.sum() > 100 and (~moved & dry & ~abl).sum() > 100
    np.testing.assert_allclose(1.0 - chi_g, 1.0 - chi_o, rtol=2e-4, atol=1e-6)
    assert np.array_equal(chi_g[~dry | abl], chi.astype(np.float32)[~dry | abl])


@pytest.mark.parametrize("pairs", ["0", "1"])
def test_fused_char_with_prefactor_beyond_fp32_matches_oracle(pairs, monkeypatch):
    """The fused sub-step's char step with A = 3.1e98 /s, 20 sub-steps, one launch per sub-step
    (pairs 0) and the two-sub-step kernel (pairs 1, the default on this grid): T within 1e-4 K and
    1 - chi within 2e-4 relative of the FP64 chain of module kernels."""
    from oracle import oracle as O
    from simulation import _bdlib as B
    from simulation.coupled_ablation import R_GAS
    from simulation.voxel_thermal_3d import _physics
    monkeypatch.setenv("BUGDEF_SUBSTEP_PAIRS", pairs)
    shape = (20, 18, 33)
    T, dry, abl, chi = _char_inputs(shape, 9)
    rng = np.random.default_rng(10)
    k = rng.uniform(0.2e-3, 0.5e-3, shape)
    rc = rng.uniform(2e-3, 4e-3, shape)
    Q = rng.uniform(0.0, 50.0, shape)
    A, EaR, dt, d = 3.1e98, 6.3e5 / R_GAS, 5e-5, 0.015
    c = B.Context(shape, (d, d, d))
    c.set_physics(_physics(dt=dt, t_ambient=20.0, t_abl=100.0, t_void=100.0, has_char=1, char_A=A,
                           char_Ea_over_R=EaR))
    for name, a in (('T', T), ('Q', Q), ('K', k), ('RHO_CP', rc), ('DRY', dry), ('ABLATED', abl), ('CHI', chi)):
        c.upload(name, a)
    c.fill('H_ACC', 0.0)
    c.substeps(20, 1, B.SUB_PLATEAU | B.SUB_CHAR | B.SUB_MASK_Q)
    Tg, chig = c.download('T'), c.download('CHI')
    c.close()
    f = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    To, Tn, chio, H = f(T), np.empty(shape), f(chi), np.zeros(shape)
    Qm = f(Q)
    Qm[abl] = 0.0
    k, rc = f(k), f(rc)
    a8, d8 = abl.astype(np.uint8), dry.astype(np.uint8)
    for _ in range(20):
        O.thermal_step(To, Tn, Qm, k, rc, *shape, d, d, d, dt, 0.0, 20.0, 1)
        O.plateau(Tn, H, rc, a8, 100.0, *shape, d8, 100.0)
        O.char(Tn, chio, d8, a8, A, EaR, dt)
        To, Tn = Tn, To
    live = (1.0 - chio) > 1e-6
    print("pairs", pairs, "max |dT|", float(np.abs(Tg - To).max()), "max rel d(1-chi) where 1-chi > 1e-6",
          float((np.abs(chig - chio)[live] / (1.0 - chio[live])).max()))
    assert float(np.abs(Tg - To).max()) < 1e-4
    np.testing.assert_allclose(1.0 - chig, 1.0 - chio, rtol=2e-4, atol=1e-6)
    assert (chio > chi + 0.05).sum() > 100
