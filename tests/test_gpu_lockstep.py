"""Coupled epochs driven by the same absorbed-power field on the device and on the FP64 oracle.

The transport is the only stochastic step of an epoch, and its photon streams cannot be pinned
bitwise against the reference.  Everything after it is deterministic, so it is checked here to
the FP32 bar: each epoch takes Q from one device transport of the current state and feeds that
same Q to
  * engine A, through the product path (``step_epoch_with_q``: one ``bd_epoch_q`` call);
  * engine B, through the same kernels called one by one (void mask, fused sub-steps with the
    plume, phase update with removal, plume injection and sync), which exposes the fields after
    the sub-steps;
  * the oracle engine, loaded with B's state before each half of the epoch.
Bars: after the sub-steps from one state, |dT| <= 1e-4 K + 4e-7 |T| (the FP32 roundings of
20 sub-steps at up to 350 C; the bar is 1e-4 K below 100 C), |dH|, |dchi| <= 1e-5; after the
phase update from the same fields, dry / ablated / to_remove / the plume void bit-exact,
ablation times, T and H exact, properties within FP32 rounding.  A and B stay bit-identical in
every thermal field and mask, which ties the step-by-step path to the product path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_FIELDS = (('T', 'T'), ('H_acc', 'H_ACC'), ('chi', 'CHI'), ('T_peak', 'T_PEAK'), ('ablation_time', 'ABLATION_TIME'),
           ('k_grid', 'K'), ('rho_cp_grid', 'RHO_CP'), ('mu_a', 'MU_A'), ('mu_s', 'MU_S'), ('g', 'G'), ('n', 'N'))


def _state(ctx):
    s = {name: ctx.download(key) for name, key in _FIELDS}
    s['dry'] = ctx.download('DRY').astype(bool)
    s['ablated'] = ctx.download('ABLATED').astype(bool)
    s['plume_raw'] = ctx.download('PLUME_RAW', dtype=np.float32)
    s['plume_ext'] = ctx.download('PLUME_EXT', dtype=np.float32)
    return s


def _t_bar(T):
    return 1e-4 + 4e-7 * np.abs(T)


def _lockstep(oracle, kw, n_epochs, check_ab=True):
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation import _bdlib as B

    def fresh():
        return {k: ({kk: (vv.copy() if isinstance(vv, np.ndarray) else vv) for kk, vv in v.items()}
                    if isinstance(v, dict) else (v.copy() if isinstance(v, np.ndarray) else v)) for k, v in kw.items()}
    engA = CoupledAblationEngine(**fresh())
    engB = CoupledAblationEngine(**fresh())
    ref = oracle.OracleEngine(mc_threads=16, **fresh())
    ctxA, ctxB = engA._ctx, engB._ctx
    ref.load_state(h_lat=ctxB.download('H_LAT'),
                   wet=(ctxB.download('WET_MU_A'), ctxB.download('WET_K'), ctxB.download('WET_RHO_CP')))
    flags = B.SUB_PLATEAU | B.SUB_MASK_Q | B.SUB_PEAK | B.SUB_CHAR | B.SUB_PLUME
    V = engB.voxel_vol
    worst = {'T': 0.0, 'T_rel_bar': 0.0, 'H': 0.0, 'chi': 0.0}
    seen = {'boiling': 0, 'dry_new': 0, 'removed': 0, 'charred': 0}
    for ep in range(n_epochs):
        # ---- Q of this epoch: one device transport of the current state
        beam = engA._epoch_beam()
        ctxA.mc_run(beam)
        Q = ctxA.download('Q', dtype=np.float32)
        ctxB.upload('Q', Q)
        # ---- A: the product path
        dA = engA.step_epoch_with_q()
        # ---- B and the oracle from the same state: the sub-steps
        s0 = _state(ctxB)
        ref.load_state(**{k: v for k, v in s0.items()})
        engB._epoch_beam()
        ctxB.plume_void()
        ctxB.substeps(engB.n_substeps, 1, flags)
        Q_th = Q.astype(np.float64)
        Q_th[s0['ablated']] = 0.0
        absorbed_J = float(np.sum(Q_th)) * V * ref.dt_epoch
        conducted_J = ref.run_substeps(Q_th)
        s1 = _state(ctxB)
        dT = np.abs(s1['T'] - ref.T)
        worst['T'] = max(worst['T'], float(dT.max()))
        worst['T_rel_bar'] = max(worst['T_rel_bar'], float((dT / _t_bar(ref.T)).max()))
        assert np.all(dT <= _t_bar(ref.T)), (ep, float(dT.max()))
        assert np.all(np.abs(s1['T_peak'] - ref.T_peak) <= _t_bar(ref.T_peak)), ep
        dH = float(np.abs(s1['H_acc'] - ref.H_acc).max())
        dchi = float(np.abs(s1['chi'] - ref.chi).max())
        worst['H'], worst['chi'] = max(worst['H'], dH), max(worst['chi'], dchi)
        assert dH <= 1e-5 and dchi <= 1e-5, (ep, dH, dchi)
        pr = ref.plume.raw_field.astype(np.float64)
        np.testing.assert_allclose(s1['plume_raw'], pr, rtol=1e-5, atol=1e-6 * max(1.0, float(pr.max())))
        # ---- the phase update, removal and plume from identical fields
        ref.load_state(T=s1['T'], H_acc=s1['H_acc'], chi=s1['chi'], T_peak=s1['T_peak'], plume_raw=s1['plume_raw'])
        t_after = engB.t + engB.dt_epoch
        n_new = ctxB.phase_update(t_after)
        ctxB.plume_void()
        if n_new > 0:
            ctxB.plume_inject()
        ctxB.plume_sync()
        tau, trans, peak, mean = ctxB.plume_diag()
        engB.t, engB.epoch = t_after, engB.epoch + 1
        engB._invalidate()
        n_ref, ablated_J, tau_r, trans_r, peak_r, mean_r = ref.finish_epoch()
        s2 = _state(ctxB)
        assert n_new == n_ref == int(dA['n_new']), (ep, n_new, n_ref, dA['n_new'])
        np.testing.assert_array_equal(ctxB.download('TO_REMOVE').astype(bool), ref._to_remove)
        for m in ('dry', 'ablated'):
            np.testing.assert_array_equal(s2[m], getattr(ref, m), err_msg=f"{m} epoch {ep}")
        np.testing.assert_array_equal(ctxB.download('VOID'), ref._to_plume(ref.compute_void_mask()).astype(np.uint8))
        np.testing.assert_array_equal(s2['ablation_time'], ref.ablation_time)
        np.testing.assert_array_equal(s2['T'], ref.T)
        np.testing.assert_array_equal(s2['H_acc'], ref.H_acc)
        np.testing.assert_array_equal(s2['chi'], ref.chi)
        for name in ('mu_a', 'mu_s', 'g', 'n', 'k_grid', 'rho_cp_grid'):
            r = ref.optical[name] if name in ('mu_a', 'mu_s', 'g', 'n') else getattr(ref, name)
            np.testing.assert_allclose(s2[name], r, rtol=1e-6, atol=1e-6 * float(np.abs(r).max()), err_msg=name)
        pr = ref.plume.raw_field.astype(np.float64)
        np.testing.assert_allclose(s2['plume_raw'], pr, rtol=1e-5, atol=1e-6 * max(1.0, float(pr.max())))
        np.testing.assert_allclose(s2['plume_ext'], ref.plume.mu_ext_plume, rtol=1e-5, atol=1e-6)
        for a, b in ((tau, tau_r), (trans, trans_r), (peak, peak_r), (mean, mean_r)):
            assert a == pytest.approx(b, rel=1e-4, abs=1e-9), ep
        # ---- the FP64 ledger of the product path against the oracle's
        L = dA['ledger_J']
        assert L['absorbed'] == pytest.approx(absorbed_J, rel=1e-9), ep
        assert L['conducted'] == pytest.approx(conducted_J, rel=1e-4, abs=1e-12), ep
        assert L['ablated'] == pytest.approx(ablated_J, rel=1e-5, abs=1e-12), ep
        # ---- A (one bd_epoch_q call) and B (the same kernels one by one) are bit-identical
        if check_ab:
            sA = _state(ctxA)
            for name, _ in _FIELDS:
                np.testing.assert_array_equal(sA[name], s2[name], err_msg=f"A vs B {name} epoch {ep}")
            for m in ('dry', 'ablated'):
                np.testing.assert_array_equal(sA[m], s2[m])
            np.testing.assert_allclose(sA['plume_raw'], s2['plume_raw'], rtol=1e-6, atol=1e-7)
        seen['boiling'] = max(seen['boiling'], int(np.sum(s2['H_acc'] > 0.0)))
        seen['dry_new'] += int(np.sum(s2['dry'] & ~s0['dry']))
        seen['removed'] += n_new
        seen['charred'] = max(seen['charred'], int(np.sum(s2['chi'] > 0.01)))
        print(f"epoch {ep}: n_new {n_new} max|dT| {float(dT.max()):.2e} K dH {dH:.1e} dchi {dchi:.1e}", flush=True)
    print("worst", worst, "seen", seen, flush=True)
    for e in (engA, engB):
        e.close()
    return worst, seen


def test_same_q_coupled_epochs_match_oracle(oracle):
    """20 coupled epochs of a layered stack (dry cuticle that chars and ablates at 350 C, an air gap
    that the crater breaches, boiling wet layers, a diffusion plume on a half-resolution grid)."""
    from simulation.workloads import layered_stack
    kw = layered_stack(shape=(40, 40, 56), n_photons=30000, plume_grid=(20, 20, 28), seed=3, power=40.0)
    worst, seen = _lockstep(oracle, kw, 20)
    assert seen['boiling'] > 1000 and seen['removed'] > 1000 and seen['charred'] > 100 and seen['dry_new'] > 0


def test_same_q_coupled_epochs_match_oracle_layered_256x256x192(oracle):
    """The benchmark workload at its own size (256 x 256 x 192 voxels of 15 um, 150 000 photons per
    epoch, plume on 128^3): three epochs, same bars."""
    from simulation.workloads import layered_stack
    worst, seen = _lockstep(oracle, layered_stack(), 3, check_ab=False)
    assert seen['removed'] > 100 and seen['boiling'] > 1000


def test_ledger_balances_state_energy_on_device(oracle):
    """The device's FP64 ledger closes the energy balance of the FP32 state: over 14 epochs of a
    slab that boils and ablates (single-stage model, thermally inert void), each epoch's
    absorbed - conducted - ablated equals the change of sum V (rho_cp (T - T_amb) + H) of the
    intact cells, and matches the oracle engine's ledger for the same Q."""
    from simulation.coupled_ablation import CoupledAblationEngine
    from test_oracle_engine import _ledger_slab, ledger_q
    kw = _ledger_slab(oracle)
    eng = CoupledAblationEngine(**kw)
    ref = oracle.OracleEngine(**_ledger_slab(oracle))
    Q = ledger_q(ref.T.shape, ref.dz).astype(np.float32)

    def energy(e):
        T, H, rc = (np.asarray(a, dtype=np.float64) for a in (e.T, e.H_acc, e.rho_cp_grid))
        live = ~np.asarray(e.ablated)
        return float(np.sum(rc[live] * (T[live] - e.t_ambient) + H[live])) * e.voxel_vol
    E0 = energy(eng)
    removed = 0
    for ep in range(14):
        eng._ctx.upload('Q', Q)
        d = eng.step_epoch_with_q()
        r = ref.step_epoch_with_q(Q.astype(np.float64))
        L, Lr = d['ledger_J'], r['ledger_J']
        E1 = energy(eng)
        bal = L['absorbed'] - L['conducted'] - L['ablated']
        print(f"epoch {ep}: dE {E1 - E0:.6e} balance {bal:.6e} n_new {d['n_new']}", flush=True)
        assert abs((E1 - E0) - bal) <= 2e-5 * L['absorbed'], ep
        assert L['absorbed'] == pytest.approx(Lr['absorbed'], rel=1e-9)
        assert L['conducted'] == pytest.approx(Lr['conducted'], rel=1e-3)
        assert L['ablated'] == pytest.approx(Lr['ablated'], rel=1e-3, abs=1e-12)
        assert d['n_new'] == r['n_new'], ep
        removed += d['n_new']
        E0 = E1
    assert removed > 0
    eng.close()


def test_air_gap_perforation_truth_table_on_device(oracle):
    """bd_phase_update's removal with air-gap perforation (coupled_ablation.py:628-660) against the
    oracle's _remove on the truth-table state of tests/test_oracle_engine.py: a cover cell removed,
    a pending opening, an opening already perforated."""
    from simulation import _bdlib as B
    from simulation.coupled_ablation import CoupledAblationEngine
    from test_oracle_engine import _tiny_engine
    ref = _tiny_engine(oracle, (3, 6))
    shape = ref.T.shape
    kw = dict(optical_props={k: v.copy() for k, v in ref.optical.items()}, k_grid=ref.k_grid.copy(),
              rho_cp_grid=ref.rho_cp_grid.copy(), h_lat_grid=ref.h_lat.copy(), Lx=ref.Lx, Ly=ref.Ly, Lz=ref.Lz,
              power=0.0, spot_dia=0.01, t_abl=100.0, t_ambient=20.0, dt_epoch=1e-3, dt_sub=5e-5, n_photons=1,
              air_gap_range=(3, 6), dry_state={'t_remove': 350.0, 'rho_cp': 2.5e-3}, seed=0)
    eng = CoupledAblationEngine(**kw)
    for e in (ref,):
        e.ablated[1, 1, 2] = True
        e.ablated[2, 0, 2:6] = True
        e.ablation_time[2, 0, 2:6] = 0.001
        e.dry[0, 0, 2] = True
        e.chi[0, 0, 2] = 0.7
        e.T[...] = 42.0
        e.T[0, 0, 2] = 360.0     # the dry cell to be removed (T >= t_remove)
    ctx = eng._ctx
    for name, a in (('ABLATED', ref.ablated), ('ABLATION_TIME', ref.ablation_time), ('DRY', ref.dry),
                    ('CHI', ref.chi), ('T', ref.T)):
        ctx.upload(name, a)
    ref.t = 0.004
    rem = np.zeros(shape, np.uint8)
    oracle.phase_update(ref.H_acc, ref.h_lat, ref.T, ref.chi, ref.dry.view(np.uint8), ref.ablated.view(np.uint8),
                        ref.optical['mu_a'], ref.optical['mu_s'], ref.optical['g'], ref.optical['n'], ref.k_grid,
                        ref.rho_cp_grid, ref._wet[0], ref._wet[1], ref._wet[2], ref._ds, ref._cs,
                        ref._cs_chi_remove, ref.alpha_max_stable, rem)
    assert rem.sum() == 1 and rem[0, 0, 2] == 1
    n_ref = ref._remove(rem.astype(bool))
    n_dev = ctx.phase_update(0.004)
    assert n_dev == n_ref == 1
    np.testing.assert_array_equal(ctx.download('ABLATED').astype(bool), ref.ablated)
    np.testing.assert_array_equal(ctx.download('ABLATION_TIME'), ref.ablation_time)
    np.testing.assert_array_equal(ctx.download('T'), ref.T)
    np.testing.assert_array_equal(ctx.download('DRY').astype(bool), ref.dry)
    assert ref.ablated[1, 1, 3:6].all() and ref.ablated[0, 0, 3:6].all()
    eng.close()


def test_phase_map_truth_table_on_device(oracle):
    """compute_phase_map on the device over the 64-row truth table of tests/test_oracle_engine.py."""
    from simulation.phase_map import compute_phase_map
    from test_oracle_engine import _phase_table
    rows = _phase_table()
    shape = (4, 4, 4)
    f32 = np.float32
    T = np.array([99.5 if r[0] else np.nextafter(f32(99.5), f32(0.0)) for r in rows], f32).reshape(shape)
    H = np.array([1e-30 if r[1] else 0.0 for r in rows], f32).reshape(shape)
    dry = np.array([bool(r[2]) for r in rows]).reshape(shape)
    chi = np.array([0.10 if r[3] else np.nextafter(f32(0.10), f32(0.0)) for r in rows], f32).reshape(shape)
    abl = np.array([bool(r[4]) for r in rows]).reshape(shape)
    gap = np.array([bool(r[5]) for r in rows]).reshape(shape)
    expect = np.array([r[6] for r in rows], dtype=np.uint8).reshape(shape)
    np.testing.assert_array_equal(compute_phase_map(T, chi, dry, abl, H, gap), expect)
