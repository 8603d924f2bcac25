"""GPU engine (phase update, removal, plume, full epochs) against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _f32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def _phase_inputs(seed, shape):
    rng = np.random.default_rng(seed)
    H = _f32(np.where(rng.random(shape) < 0.5, rng.uniform(0, 3.0, shape), 0.0))
    hl = _f32(np.where(rng.random(shape) < 0.8, 2.26, 0.0))
    T = _f32(rng.uniform(20.0, 400.0, shape))
    chi = _f32(rng.uniform(0, 1, shape))
    chi.flat[:20] = np.float32(0.95)  # exactly at the removal threshold as stored
    dry = rng.random(shape) < 0.4
    abl = rng.random(shape) < 0.1
    opt = {k: _f32(rng.uniform(0.5, 2.0, shape)) for k in ("mu_a", "mu_s", "g", "n")}
    k = _f32(rng.uniform(0.2e-3, 0.6e-3, shape))
    rc = _f32(rng.uniform(2e-3, 4e-3, shape))
    return H, hl, T, chi, dry, abl, opt, k, rc


def test_phase_update_masks_bit_exact(oracle):
    """_update_phase_properties_numba (coupled_ablation.py:77-180): given the same stored fields the
    device's dry and to_remove masks are bit-exact with the FP64 oracle; properties within FP32."""
    from simulation.coupled_ablation import _update_phase_properties_numba
    shape = (23, 19, 31)
    H, hl, T, chi, dry, abl, opt, k, rc = _phase_inputs(3, shape)
    ds = dict(mu_a=5.0, mu_s=7.0, g=0.6, n=1.3, k=0.25e-3, rho_cp=1.5e-3, t_remove=350.0)
    cs = dict(mu_a=50.0, mu_s=10.0, k=0.8e-3, rho_cp=1.0e-3, t_remove=300.0)
    alpha = 0.35 / (5e-5 * 3 / 0.015 ** 2)
    wet = (opt["mu_a"].copy(), k.copy(), rc.copy())
    o = {kk: v.copy() for kk, v in opt.items()}
    Ho, ko, rco, dro = H.copy(), k.copy(), rc.copy(), dry.astype(np.uint8).copy()
    rem_o = np.zeros(shape, np.uint8)
    oracle.phase_update(Ho, hl, T, chi, dro, abl.astype(np.uint8), o["mu_a"], o["mu_s"], o["g"], o["n"], ko, rco,
                        wet[0], wet[1], wet[2], ds, cs, 0.95, alpha, rem_o)
    g = {kk: v.copy() for kk, v in opt.items()}
    Hg, kg, rcg, drg = H.copy(), k.copy(), rc.copy(), dry.copy()
    rem_g = np.zeros(shape, bool)
    arr = lambda v: np.array([v])
    _update_phase_properties_numba(Hg, hl, T, chi, drg, abl, g["mu_a"], g["mu_s"], g["g"], g["n"], kg, rcg, wet[0],
                                   wet[1], wet[2], arr(5.0), arr(7.0), arr(0.6), arr(1.3), arr(0.25e-3),
                                   arr(1.5e-3), arr(350.0), True, True, arr(50.0), arr(10.0), arr(0.8e-3),
                                   arr(1.0e-3), arr(300.0), 0.95, True, True, True, alpha, rem_g)
    np.testing.assert_array_equal(drg, dro.astype(bool))
    np.testing.assert_array_equal(rem_g, rem_o.astype(bool))
    assert rem_o.sum() > 100 and (dro.astype(bool) & ~dry).sum() > 100
    np.testing.assert_array_equal(Hg == 0.0, Ho == 0.0)
    for kk in ("mu_a", "mu_s", "g", "n"):
        np.testing.assert_allclose(g[kk], o[kk], rtol=1e-6)
    np.testing.assert_allclose(kg, ko, rtol=1e-6)
    np.testing.assert_allclose(rcg, rco, rtol=1e-6)


def test_single_stage_and_phase_map_bit_exact(oracle):
    from simulation.coupled_ablation import _update_single_stage_ablation_numba
    from simulation.phase_map import compute_phase_map
    shape = (17, 13, 29)
    H, hl, T, chi, dry, abl, opt, k, rc = _phase_inputs(4, shape)
    rem_o = np.zeros(shape, np.uint8)
    oracle.single_stage(H, hl, T, 100.0, abl.astype(np.uint8), rem_o)
    rem_g = np.zeros(shape, bool)
    _update_single_stage_ablation_numba(H, hl, T, 100.0, abl, rem_g)
    np.testing.assert_array_equal(rem_g, rem_o.astype(bool))
    gap = np.zeros(shape, bool)
    gap[:, :, 5:9] = True
    T[0, 0, :4] = np.float32(99.5)
    for air in (None, gap):
        np.testing.assert_array_equal(compute_phase_map(T, chi, dry, abl, H, air),
                                      oracle.compute_phase_map(T, chi, dry, abl, H, air))


def test_plume_kernels_parity(oracle):
    """plume.py:18-80 diffusion step and :83-139 separable splat vs the oracle."""
    from simulation.plume import _convolve_separable_plume_numba, _diffuse_plume_step_numba
    rng = np.random.default_rng(6)
    shape = (21, 18, 25)
    C = rng.uniform(0, 2.0, shape).astype(np.float32)
    void = rng.random(shape) < 0.7
    Co, Cg = np.empty_like(C), np.empty_like(C)
    args = (*shape, 0.03, 0.025, 0.02, 2e-5, 1.0, 2.0, 15.0, float(np.exp(-2e-5 / 0.01)))
    oracle.plume_diffuse(C, Co, void.astype(np.uint8), *args)
    _diffuse_plume_step_numba(C, Cg, void, *args)
    np.testing.assert_allclose(Cg, Co, rtol=2e-6, atol=1e-7)
    raw_o = rng.uniform(0, 1, shape).astype(np.float32)
    raw_g = raw_o.copy()
    imp = (rng.random((4, 3, 5)) < 0.6).astype(np.float32)
    kx, ky, kz = (rng.random(n).astype(np.float32) for n in (5, 3, 7))
    oracle.plume_convolve(raw_o, imp, kx, ky, kz, 0.9, -1, 4, 5)
    _convolve_separable_plume_numba(raw_g, imp, kx, ky, kz, 0.9, -1, 4, 5)
    np.testing.assert_array_equal(raw_g, raw_o)  # same FP32 accumulation order: bit-exact


def test_paired_plume_rows_match_single_rows(oracle):
    """Rows of a multiple of 128 cells take the paired diffusion kernel (two cells per lane,
    loads issued together, a select for cells that are not void interior ones).  Against the
    one-cell-at-a-time kernel (BUGDEF_PLUME_PAIRS=0, read when the context is made): the same
    bits after 12 steps on a random field and void mask; and one paired step against the
    oracle within FP32 round-off (the bar of test_plume_kernels_parity)."""
    import os
    from simulation.plume import MultispectralPlumeField
    rng = np.random.default_rng(11)
    shape = (13, 11, 256)
    C = rng.uniform(0, 2.0, shape).astype(np.float32)
    void = rng.random(shape) < 0.7
    dt = 2e-5

    def field(pairs):
        old = os.environ.get("BUGDEF_PLUME_PAIRS")
        os.environ["BUGDEF_PLUME_PAIRS"] = pairs
        try:
            p = MultispectralPlumeField(shape=shape, model='diffusion', D_plume=0.5, v_drift=10.0,
                                        blast_expansion=False)
        finally:
            if old is None:
                del os.environ["BUGDEF_PLUME_PAIRS"]
            else:
                os.environ["BUGDEF_PLUME_PAIRS"] = old
        p.raw_field = C
        return p

    paired, single = field("1"), field("0")
    paired.diffuse_substep(void, dt_sub=dt)
    Co = np.empty_like(C)
    decay = float(np.exp(-dt / paired.tau_linger))
    oracle.plume_diffuse(C, Co, void.astype(np.uint8), *shape, paired.dx, paired.dy, paired.dz, dt, paired.D_xy,
                         paired.D_z, paired.v_drift, decay)
    np.testing.assert_allclose(paired.raw_field, Co, rtol=2e-6, atol=1e-7)
    single.diffuse_substep(void, dt_sub=dt)
    for _ in range(11):
        paired.diffuse_substep(void, dt_sub=dt)
        single.diffuse_substep(void, dt_sub=dt)
    a, b = paired.raw_field, single.raw_field
    assert np.any(a > 0.0)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_plume_reference_fixtures_on_gpu():
    """tests/test_plume.py:110-179,230-266 of the reference on the device plume field."""
    from simulation.plume import MultispectralPlumeField
    p = MultispectralPlumeField(shape=(15, 15, 15), model='diffusion', delta_mu_per_voxel=5.0, D_plume=0.5,
                                v_drift=10.0, blast_expansion=False)
    newly = np.zeros((15, 15, 15), bool)
    newly[7, 7, 7] = True
    p.inject(newly)
    raw = p.raw_field
    assert raw[7, 7, 7] == pytest.approx(5.0) and np.sum(raw > 0.0) == 1
    void = np.zeros((15, 15, 15), bool)
    void[5:10, 5:10, 5:10] = True
    for _ in range(10):
        p.diffuse_substep(void, dt_sub=5e-5)
    raw = p.raw_field
    assert raw[7, 7, 7] < 5.0 and raw[7, 7, 8] > 0.0
    assert np.all(raw[~void] == 0.0) and np.all(raw[0] == 0.0) and np.all(raw[-1] == 0.0)
    b = MultispectralPlumeField(shape=(25, 25, 25), model='diffusion', delta_mu_per_voxel=1.0, D_plume=2.0,
                                v_drift=80.0, blast_expansion=True)
    vb = np.zeros((25, 25, 25), bool)
    vb[8:17, 8:17, 5:20] = True
    nb = np.zeros((25, 25, 25), bool)
    nb[12, 12, 15] = True
    b.inject(nb, void_mask=vb)
    rb = b.raw_field
    assert rb[12, 12, 15] < 0.20 and np.sum((rb > 1e-4) & vb) > 10 and np.all(rb[~vb] == 0.0)
    a = MultispectralPlumeField(shape=(31, 31, 31), dx_mm=0.02, dy_mm=0.02, dz_mm=0.02, model='diffusion', D_xy=1.0,
                                D_z=4.0, v_drift=0.0, tau_linger_s=float('inf'), blast_expansion=False)
    nc = np.zeros((31, 31, 31), bool)
    nc[15, 15, 15] = True
    a.inject(nc)
    full = np.ones((31, 31, 31), bool)
    for _ in range(20):
        a.diffuse_substep(full, dt_sub=2e-5)
    ra = a.raw_field
    assert ra[17, 15, 15] == pytest.approx(ra[15, 17, 15], rel=1e-4)
    assert ra[15, 15, 17] > 2.0 * ra[17, 15, 15]


def _small_engine_kwargs(**over):
    N, Nz, voxel = 11, 20, 0.05
    shape = (N, N, Nz)
    kw = dict(optical_props={'mu_a': np.full(shape, 10.0), 'mu_s': np.zeros(shape), 'g': np.zeros(shape),
                             'n': np.ones(shape)},
              k_grid=np.full(shape, 0.5e-3), rho_cp_grid=np.full(shape, 4.0e-3), h_lat_grid=np.full(shape, 2.0),
              Lx=N * voxel, Ly=N * voxel, Lz=Nz * voxel, power=5.0, spot_dia=0.3, t_abl=100.0, t_ambient=25.0,
              dt_epoch=0.002, dt_sub=5e-5, n_photons=10000, seed=5)
    kw.update(over)
    return kw


def test_ablation_extensions_reference_fixtures():
    """tests/test_ablation_extensions.py:72-138 of the reference on the device engine."""
    from simulation.coupled_ablation import CoupledAblationEngine
    eng = CoupledAblationEngine(**_small_engine_kwargs())
    assert eng.char_state is None and np.all(eng.chi == 0.0)
    eng.step_epoch()
    assert np.all(eng.chi == 0.0)
    dry_state = dict(mu_a=5.0, mu_s=0.0, g=0.0, n=1.0, k=0.2e-3, rho_cp=2.0e-3, t_remove=600.0)
    char_state = dict(A=1e8, Ea=60000.0, mu_a=50.0, mu_s=10.0, k=0.8e-3, rho_cp=1.0e-3, t_remove=600.0)
    eng = CoupledAblationEngine(**_small_engine_kwargs(dry_state=dry_state, char_state=char_state, power=15.0))
    for _ in range(25):
        eng.step_epoch()
    dm = eng.dry & ~eng.ablated
    assert np.any(dm)
    np.testing.assert_allclose(eng.optical['mu_a'][dm], 5.0 + eng.chi[dm] * 45.0, rtol=1e-4)
    eng = CoupledAblationEngine(**_small_engine_kwargs(plume_closure=dict(tau_linger_s=0.010, delta_mu_per_voxel=5.0,
                                                                          sigma_mm=0.05), power=20.0))
    hit = False
    for _ in range(30):
        d = eng.step_epoch()
        if d['n_new'] > 0:
            hit = True
            assert d['plume_tau'] > 0.0 and d['plume_transmission'] < 1.0 and d['plume_peak_mu_a'] > 0.0
            assert d['plume_transmission'] == pytest.approx(np.exp(-d['plume_tau']))
            break
    assert hit
    for th in (0.0, 45.0):
        e = CoupledAblationEngine(**_small_engine_kwargs(theta_inc_deg=th))
        assert e.step_epoch()['absorbed_frac'] > 0.0


@pytest.mark.parametrize("dt_clamp", [False, True])
def test_engine_against_oracle_engine(oracle, dt_clamp):
    """Coupled epochs with dry state, char, air gap and diffusion plume: the device engine and
    the FP64 oracle engine trace the same photon streams; compare the state after 6 epochs.
    With dt_clamp the plume takes the Metal solver's diffusion step on both sides."""
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.phase_map import CHAR_STATE
    shape = (32, 32, 40)
    rng = np.random.default_rng(0)
    mu_a = np.zeros(shape)
    mu_s = np.zeros(shape)
    g = np.zeros(shape)
    n = np.ones(shape)
    k = np.full(shape, 0.026e-3)
    rc = np.full(shape, 1.206e-6)
    hl = np.full(shape, 1e12)
    dry = np.zeros(shape, bool)
    abl = np.zeros(shape, bool)
    abl[:, :, :4] = True
    for (a, b), (ma, ms, gg, nn, kk, rr, hh, dd) in (
            ((4, 7), (8.0, 20.0, 0.85, 1.5, 0.3e-3, 2.4e-3, 0.0, True)),
            ((10, 40), (3.0, 15.0, 0.9, 1.37, 0.5e-3, 3.6e-3, 2.26, False))):
        mu_a[:, :, a:b], mu_s[:, :, a:b], g[:, :, a:b], n[:, :, a:b] = ma, ms, gg, nn
        k[:, :, a:b], rc[:, :, a:b], hl[:, :, a:b], dry[:, :, a:b] = kk, rr, hh, dd
    opt = {'mu_a': mu_a, 'mu_s': mu_s, 'g': g, 'n': n}
    ds = {'mu_a': mu_a.copy(), 'mu_s': mu_s.copy(), 'g': g.copy(), 'n': n.copy(), 'k': k.copy(),
          'rho_cp': rc.copy(), 't_remove': 350.0}
    common = dict(optical_props=opt, k_grid=k, rho_cp_grid=rc, h_lat_grid=hl, Lx=0.48, Ly=0.48, Lz=0.6,
                  power=8.0, spot_dia=0.2, t_abl=100.0, t_ambient=20.0, dt_epoch=0.001, dt_sub=5e-5,
                  n_photons=40000, h_eff=15e-6, theta_inc_deg=22.0, char_state=CHAR_STATE,
                  plume_closure={'model': 'diffusion', 'D_xy': 1.0, 'D_z': 2.0, 'v_drift': 0.0,
                                 'tau_linger_s': float('inf'), 'blast_expansion': False, 'dt_clamp': dt_clamp},
                  initial_dry_mask=dry, initial_ablated_mask=abl, air_gap_range=(7, 10), wavelength_nm=532, seed=42)
    eng = CoupledAblationEngine(dry_state={kk: (v.copy() if isinstance(v, np.ndarray) else v)
                                           for kk, v in ds.items()}, **common)
    ref = oracle.OracleEngine(dry_state={kk: (v.copy() if isinstance(v, np.ndarray) else v)
                                         for kk, v in ds.items()}, mc_threads=8, **common)
    dg, dr = [], []
    for _ in range(6):
        dg.append(eng.step_epoch())
        dr.append(ref.step_epoch())
    for a, b in zip(dg, dr):
        assert abs(a['absorbed_frac'] - b['absorbed_frac']) < 0.01
        assert abs(a['mc_energy_closure'] - 1.0) < 0.01
    Tg, Tr = np.asarray(eng.T), ref.T
    rise = float(Tr.max() - 20.0)
    assert rise > 50.0
    assert abs(float(Tg.max()) - float(Tr.max())) < 0.05 * rise
    assert abs(float(Tg.mean()) - float(Tr.mean())) < 0.02 * float(Tr.mean() - 20.0) + 1e-3
    # removed-volume and char agree statistically
    assert abs(int(eng.ablated.sum()) - int(ref.ablated.sum())) <= 0.2 * max(10, int(ref.ablated.sum() - abl.sum()))
    assert eng.T_peak.max() >= float(Tg.max()) - 1e-4
    # the plume masses agree statistically (same photon streams, so nearly the same removals)
    pg = np.asarray(eng.plume.raw_field, dtype=np.float64)
    pr = np.asarray(ref.plume.raw_field, dtype=np.float64)
    assert pg.max() > 0.0
    assert abs(pg.sum() - pr.sum()) <= 0.2 * pr.sum()


def test_graph_epochs_match_eager_epochs():
    """bd_epoch_graph (captured epoch, device-side key and time) against bd_epoch: same seed,
    same photon streams; the only difference is float-atomic summation order in the tallies."""
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.workloads import layered_stack
    kw = layered_stack(shape=(48, 48, 64), n_photons=30000, plume_grid=(24, 24, 32), power=60.0)
    e1 = CoupledAblationEngine(graph=False, **kw)
    e2 = CoupledAblationEngine(graph=True, **layered_stack(shape=(48, 48, 64), n_photons=30000,
                                                           plume_grid=(24, 24, 32), power=60.0))
    for ep in range(8):
        if ep == 6:
            e1.power = e2.power = 0.0  # laser off: a second captured graph
        d1, d2 = e1.step_epoch(), e2.step_epoch()
        assert d1['absorbed_frac'] == pytest.approx(d2['absorbed_frac'], abs=1e-6)
        assert abs(d1['n_new'] - d2['n_new']) <= max(2, 0.02 * d1['n_new'])
    T1, T2 = np.asarray(e1.T), np.asarray(e2.T)
    assert float(np.max(np.abs(T1 - T2))) < 1e-2 * float(T1.max() - 20.0) + 1e-3
    np.testing.assert_allclose(e1.ablation_time[e1.ablated & e2.ablated], e2.ablation_time[e1.ablated & e2.ablated])
    assert e1.t == e2.t and e1.epoch == e2.epoch == 8


def test_graph_epochs_follow_physics_changes():
    """A captured epoch graph holds the physics as kernel arguments: a change of the
    ablation threshold, the surface loss and the dry/char parameters between epochs
    must reach the replayed graph exactly as it reaches the eager epoch."""
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.workloads import layered_stack
    engines = [CoupledAblationEngine(graph=g, **layered_stack(shape=(40, 40, 48), n_photons=20000,
                                                              plume_grid=(20, 20, 24), power=60.0))
               for g in (False, True)]
    for ep in range(6):
        if ep == 3:
            for e in engines:
                e.t_abl = e.t_abl - 40.0
                e.h_eff = 4.0 * e.h_eff + 50.0
        d1, d2 = (e.step_epoch() for e in engines)
        assert d1['absorbed_frac'] == pytest.approx(d2['absorbed_frac'], abs=1e-6)
        assert abs(d1['n_new'] - d2['n_new']) <= max(2, 0.02 * d1['n_new'])
    T1, T2 = (np.asarray(e.T) for e in engines)
    assert float(np.max(np.abs(T1 - T2))) < 1e-2 * float(T1.max() - 20.0) + 1e-3
    assert int(engines[0].ablated.sum()) == pytest.approx(int(engines[1].ablated.sum()), rel=0.02, abs=2)


def _switch(engine, geo2, wl2, p2):
    """The reference's mid-run wavelength switch, verbatim in effect
    (run_dual_permutation_suite.py:257-283): intact-voxel optics and dry-state
    optics replaced in place, wet copies re-taken, dry/char arrays rebuilt,
    plume re-tuned and re-synced, new wavelength and power."""
    intact = ~np.asarray(engine.ablated)
    for key in ('mu_a', 'mu_s', 'g', 'n'):
        engine.optical[key][intact] = geo2['optical'][key][intact]
        engine.dry_state[key][intact] = geo2['dry_state'][key][intact]
    engine._wet = (engine.optical['mu_a'].copy(), engine.k_grid.copy(), engine.rho_cp_grid.copy())
    engine._ds_arrays = None
    engine._init_ds_cs_arrays()
    if engine.plume is not None:
        engine.plume.set_wavelength(wl2)
        engine.plume.sync_extinction(engine.compute_void_mask())
    engine.wavelength_nm = float(wl2)
    engine.power = p2


def test_mid_run_wavelength_switch_matches_oracle(oracle):
    """Three epochs at 532 nm / 30 W, the reference's switch sequence, three
    epochs at 1470 nm / 60 W: device engine against the FP64 oracle engine."""
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.workloads import layered_stack
    kw1 = layered_stack(shape=(32, 32, 56), n_photons=40_000, plume_grid=None, seed=3, power=30.0)
    kw2 = layered_stack(shape=(32, 32, 56), n_photons=40_000, plume_grid=None, seed=5, power=60.0)
    geo2 = {'optical': kw2['optical_props'], 'dry_state': kw2['dry_state']}

    def fresh(kw):
        return {k: ({kk: (vv.copy() if isinstance(vv, np.ndarray) else vv) for kk, vv in v.items()}
                    if isinstance(v, dict) else (v.copy() if isinstance(v, np.ndarray) else v)) for k, v in kw.items()}
    eng = CoupledAblationEngine(**fresh(kw1))
    k_or = fresh(kw1)
    k_or.pop('plume_grid')
    ref = oracle.OracleEngine(mc_threads=8, **k_or)
    for ep in range(6):
        if ep == 3:
            _switch(eng, geo2, 1470.0, 60.0)
            _switch(ref, geo2, 1470.0, 60.0)
            intact = ~np.asarray(eng.ablated)
            np.testing.assert_allclose(np.asarray(eng.optical['mu_a'])[intact], geo2['optical']['mu_a'][intact],
                                       rtol=1e-7)
        a, b = eng.step_epoch(), ref.step_epoch()
        assert abs(a['absorbed_frac'] - b['absorbed_frac']) < 0.01, ep
        assert abs(a['plume_transmission'] - b['plume_transmission']) < 0.02, ep
    Tg, Tr = np.asarray(eng.T), ref.T
    rise = float(Tr.max() - 20.0)
    assert abs(float(Tg.max()) - float(Tr.max())) < 0.05 * rise
    assert eng.plume.wavelength == 1470.0 and eng.wavelength_nm == 1470.0


def _sagi_run():
    """knowledge/simulation_lit_comparison/code/sagi_stages.py:26-55 (run_sagi_stages) on this
    engine: a quasi-1-D wet absorber (McMillan et al. tissue constants, mcmillan_benchmark.py:32-41)
    under a beam ten times wider than the 0.2 mm column, boiled dry and removed at 300 C."""
    from simulation.coupled_ablation import CoupledAblationEngine
    W0, PRO0 = 0.75, 0.25
    rho_wet = 1000.0 / (W0 + 0.649 * PRO0)
    cp_wet = 1000.0 * (4.2 * W0 + 1.09 * PRO0)
    k_wet = rho_wet * (6.28e-4 * W0 + 1.17e-4 * PRO0)
    rho_dry, cp_dry = 1000.0 / 0.649, 1000.0 * 1.09
    k_dry = rho_dry * 1.17e-4
    mu_a_wet, h_lat = 85.0, W0 * rho_wet * 2.26e6 * 1e-9
    I, Ta, t_end, dz, Nz, dt_epoch, n_photons = 0.5, 300.0, 1.2, 0.005, 100, 0.002, 60000
    Nx = Ny = 4
    Lx = 0.2
    w0 = 10.0 * (Lx / 2.0)
    shp = (Nx, Ny, Nz)
    eng = CoupledAblationEngine(
        optical_props={'mu_a': np.full(shp, mu_a_wet), 'mu_s': np.zeros(shp), 'g': np.zeros(shp), 'n': np.ones(shp)},
        k_grid=np.full(shp, k_wet * 1e-3), rho_cp_grid=np.full(shp, rho_wet * cp_wet * 1e-9),
        h_lat_grid=np.full(shp, h_lat), Lx=Lx, Ly=Lx, Lz=Nz * dz,
        power=I * np.pi * w0 ** 2 / 2.0, spot_dia=2 * w0, t_abl=100.0, t_ambient=25.0,
        dt_epoch=dt_epoch, dt_sub=4e-5, n_photons=n_photons, h_eff=0.0, seed=3,
        dry_state=dict(mu_a=0.25 * mu_a_wet, mu_s=0.0, g=0.0, n=1.0, k=k_dry * 1e-3,
                       rho_cp=rho_dry * cp_dry * 1e-9, t_remove=Ta))
    rec = {k: [] for k in ('t_s', 'T_surface', 'depth_mm', 'n_dry_layers')}
    try:
        for _ in range(int(round(t_end / dt_epoch))):
            eng.step_epoch()
            col_abl = np.asarray(eng.ablated).all(axis=(0, 1))
            k_surf = int(np.argmin(col_abl)) if not col_abl.all() else Nz - 1
            rec['t_s'].append(eng.t)
            rec['T_surface'].append(float(np.asarray(eng.T)[:, :, k_surf].mean()))
            rec['depth_mm'].append(float(col_abl.sum() * dz))
            rec['n_dry_layers'].append(int(np.asarray(eng.dry).any(axis=(0, 1)).sum()))
    finally:
        eng.close()
    return {k: np.asarray(v) for k, v in rec.items()}


def test_staged_ablation_series_matches_reference_record():
    """The reference engine's recorded staged-ablation series (sagi_stages.json, 600 epochs):
    heating to boiling, the 100 C plateau, drying, dry heating to 300 C, removal.  The stage
    times, the surface temperature and the crater depth follow the record."""
    import json
    import os
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sagi_stages_reference.json")))
    r = _sagi_run()
    t = np.asarray(ref['t_s'])
    assert np.allclose(r['t_s'], t, atol=1e-9)
    Tr, dr, ndr = (np.asarray(ref[k]) for k in ('T_surface', 'depth_mm', 'n_dry_layers'))
    def first(mask):
        return float(t[np.argmax(mask)]) if mask.any() else float('inf')
    for name, a, b in (("boiling", r['T_surface'] >= 99.9, Tr >= 99.9),
                       ("first dry layer", r['n_dry_layers'] > 0, ndr > 0),
                       ("first removal", r['depth_mm'] > 0, dr > 0)):
        print(name, first(a), first(b))
        assert abs(first(a) - first(b)) <= 0.1 * first(b) + 0.004, name   # the FP64 oracle engine: 0, 8, 2 ms
    dT = np.abs(r['T_surface'] - Tr)
    print("surface T: max |dT|", float(dT.max()), "mean", float(dT.mean()))
    assert float(np.median(dT)) < 1.0 and float(dT.mean()) < 3.0
    dd = np.abs(r['depth_mm'] - dr)
    print("depth: max |d|", float(dd.max()), "at end", float(r['depth_mm'][-1]), float(dr[-1]))
    assert float(dd.max()) <= 0.0101 and abs(r['depth_mm'][-1] - dr[-1]) <= 0.0051


def test_mcmillan_pixel_depths_match_reference_engine():
    """The reference engine's crater depths of the McMillan et al. (2021) pixel runs
    (mcmillan_benchmark.py:46-99: an 81 x 81 x 160 grid of 7.5 um, a pure absorber of
    850 /cm, dry tissue at 25 % of the wet absorption, removal at 500 C, triangular
    pulses of 50-400 mJ at 1.2 ms per mJ, 20 000 photons per 2 ms epoch), recorded as
    the model values of benchmark_results.json (M-<E>-exp)."""
    import json
    import os
    from simulation.coupled_ablation import CoupledAblationEngine
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_engine_values.json")))["values"]
    W0, PRO0 = 0.75, 0.25
    rho_wet = 1000.0 / (W0 + 0.649 * PRO0)
    cp_wet = 1000.0 * (4.2 * W0 + 1.09 * PRO0)
    k_wet = rho_wet * (6.28e-4 * W0 + 1.17e-4 * PRO0)
    rho_dry, cp_dry = 1000.0 / 0.649, 1000.0 * 1.09
    k_dry = rho_dry * 1.17e-4
    mu_a_wet, h_lat = 85.0, W0 * rho_wet * 2.26e6 * 1e-9
    N, voxel, dt = 81, 0.0075, 0.002
    Nz = int(round(1.2 / voxel))
    shp, L = (N, N, Nz), N * voxel
    got = {}
    for E in (50, 100, 150, 200, 250, 300, 350, 400):
        n_ep = int(round(1.2e-3 * E / dt))
        t = (np.arange(n_ep) + 0.5) * dt
        p = np.clip(1.0 - np.abs(2.0 * t / (1.2e-3 * E) - 1.0), 0.0, None)
        P = p * (E * 1e-3) / (p.sum() * dt)
        eng = CoupledAblationEngine(
            optical_props={'mu_a': np.full(shp, mu_a_wet), 'mu_s': np.zeros(shp), 'g': np.zeros(shp),
                           'n': np.ones(shp)},
            k_grid=np.full(shp, k_wet * 1e-3), rho_cp_grid=np.full(shp, rho_wet * cp_wet * 1e-9),
            h_lat_grid=np.full(shp, h_lat), Lx=L, Ly=L, Lz=Nz * voxel, power=1.0, spot_dia=0.25, t_abl=100.0,
            t_ambient=5.0, dt_epoch=dt, dt_sub=4e-5, n_photons=20000, h_eff=0.0, beam_profile='tophat', seed=E,
            dry_state=dict(mu_a=0.25 * mu_a_wet, mu_s=0.0, g=0.0, n=1.0, k=k_dry * 1e-3,
                           rho_cp=rho_dry * cp_dry * 1e-9, t_remove=500.0))
        try:
            for i in range(n_ep):
                eng.power = float(P[i])
                eng.step_epoch()
            got[E] = eng.centerline_depth_mm(N // 2, N // 2)
        finally:
            eng.close()
    print({E: (round(got[E], 4), ref[f"M-{E}-exp"]["model"]) for E in got})
    for E, d in got.items():
        # within one voxel of the reference engine's depth (the runs so far reproduce it exactly;
        # one voxel leaves room for the unseeded reference transport)
        assert abs(d - ref[f"M-{E}-exp"]["model"]) <= voxel + 1e-9, E
