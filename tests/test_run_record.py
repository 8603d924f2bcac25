"""KPI extraction (compute_suite_metrics.py:28-118) on synthetic run records (CPU), and the
device recorder against host-side slicing of the same engine state (GPU)."""
import numpy as np
import pytest


def _synthetic_npz(path, n_ep=250, NX=41, NZ=60, DX=0.015, DZ=0.015):
    grid_x = np.linspace(0, NX * DX, NX)
    grid_z = np.linspace(-0.315, NZ * DZ - 0.315, NZ)
    times = (np.arange(n_ep) + 1) * 1e-3
    depth = np.minimum(times * 4.0, 0.6)  # mm
    ph = np.zeros((n_ep, NX, NZ), np.uint8)
    T = np.full((n_ep, NX, NZ), 20.0, np.float32)
    zum = grid_z * 1000.0
    for e in range(n_ep):
        for iz in range(NZ):
            if 0 <= zum[iz] <= depth[e] * 1000.0:
                ph[e, 15:26, iz] = 3
                ph[e, 12:15, iz] = 2
        T[e] += 100.0 * np.exp(-((np.arange(NX) - NX // 2) ** 2) / 50.0)[:, None]
    np.savez_compressed(path, times=times, grid_x=grid_x, grid_z=grid_z, xz_T=T, xz_phase=ph,
                        crater_depths_mm=depth, DX=DX, DZ=DZ, wavelength_nm=532.0, power_w=40.0)
    return depth


def test_analyze_npz_closed_forms(tmp_path):
    from simulation.run_record import analyze_npz
    p = str(tmp_path / "run.npz")
    depth = _synthetic_npz(p)
    k = analyze_npz(p, pulse_duration_s=0.2)
    assert k['final_depth_um'] == pytest.approx(depth[199] * 1000.0)
    assert k['avg_depth_um'] == pytest.approx(np.mean(depth[:200]) * 1000.0)
    assert k['speed_mm_s'] == pytest.approx(depth[199] / 0.2)
    # every ablated solid slice is 11 voxels wide
    assert k['avg_hole_dia_mm'] == pytest.approx(11 * 0.015)
    # volumes: sum over ablated solid slices of pi r^2 dz (r = 11 dx / 2), boiling annulus 3 voxels
    grid_z = np.linspace(-0.315, 60 * 0.015 - 0.315, 60) * 1000.0
    nsl = np.sum((grid_z >= 0) & (grid_z <= depth[199] * 1000.0) & ~((grid_z >= 165) & (grid_z <= 465)))
    ra, rt = 11 * 0.015 / 2, 14 * 0.015 / 2
    assert k['v_abl_mm3'] == pytest.approx(nsl * np.pi * ra ** 2 * 0.015)
    assert k['v_boil_mm3'] == pytest.approx(nsl * np.pi * (rt ** 2 - ra ** 2) * 0.015)
    assert k['R_pct'] == 19.61 and k['energy_used_j'] == pytest.approx(8.0)


@pytest.mark.gpu
def test_recorder_matches_host_slices(tmp_path):
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.phase_map import compute_phase_map
    from simulation.run_record import RunRecorder, analyze_npz
    from simulation.workloads import layered_stack
    kw = layered_stack(shape=(64, 48, 80), n_photons=20000, plume_grid=(32, 24, 40))
    eng = CoupledAblationEngine(**kw)
    gap = kw['air_gap_range']
    ix_beam = np.clip(np.arange(80) // 4 + 20, 0, 63)
    rec = RunRecorder(eng, nz_air_top=21, air_gap=gap, regions={'dorsal': (21, gap[0]), 'core': (50, 60),
                                                                'ventral': (70, 80)}, ix_beam=ix_beam)
    for _ in range(12):
        rec.record(eng.step_epoch())
    a = rec.arrays()
    j = eng.Ny // 2
    T, Q = np.asarray(eng.T), np.asarray(eng.Q)
    is_gap = np.zeros(eng.T.shape, bool)
    is_gap[:, :, :21] = True
    is_gap[:, :, gap[0]:gap[1]] = True
    ph = compute_phase_map(T, eng.chi, eng.dry, eng.ablated, eng.H_acc, is_air_gap=is_gap)
    np.testing.assert_allclose(a['xz_T'][-1], T[:, j, :], rtol=1e-7)
    np.testing.assert_allclose(a['xz_Q'][-1], Q[:, j, :], rtol=1e-7)
    np.testing.assert_array_equal(a['xz_phase'][-1], ph[:, j, :])
    np.testing.assert_array_equal(a['centerline_phase'][-1], ph[ix_beam, j, np.arange(80)])
    assert a['peak_T_core'][-1] == pytest.approx(float(T[:, j, 50:60].max()), rel=1e-7)
    p = str(tmp_path / "rec.npz")
    rec.save(p, wavelength_nm=532.0, power_w=40.0)
    k = analyze_npz(p, pulse_duration_s=0.012)
    assert k['final_depth_um'] >= 0.0 and np.isfinite(k['core_max_T'])


def test_beetle_model_layout():
    """run_highres_knowledge_suite.py:126-185: 21 slices of ambient air (ablated from
    t = 0), a 165 um dry dorsal shell, the 300 um subelytral air gap on the air
    properties, then the wet body and a 90 um dry ventral shell; optics from
    eval_single_layer at the wavelength; the tilted beam enters the dorsal
    surface at x = 1 mm and reaches x = 2 mm at the ventral face."""
    import math
    from simulation import beetle_suite as S
    from simulation.tissue_stack import BASE_LAYERS, eval_single_layer
    g = S.build_beetle_model(532)
    col = lambda a: a[100, 100, :]
    assert col(g['initial_ablated'])[:21].all() and not col(g['initial_ablated'])[21:].any()
    dry = col(g['dry_mask'])
    assert dry[21:32].all() and not dry[32:178].any() and dry[178:].all()
    k = col(g['k_grid'])
    assert (k[32:52] == 0.026e-3).all() and k[52] == 0.48e-3 and k[90] == 0.50e-3
    ev = eval_single_layer(BASE_LAYERS[6], 532.0)   # hemolymph
    assert col(g['optical']['mu_a'])[73] == float(ev['mu_a']) and col(g['optical']['n'])[89] == float(ev['n'])
    assert abs(S.X_TOP + S.TAN_THETA * S.L_AIR_TOP - 1.0) < 1e-12
    assert abs(S.X_TOP + S.TAN_THETA * S.LZ - 2.0) < 1e-12
    assert abs(S.THETA_INC_DEG - math.degrees(math.atan(1.0 / 2.445))) < 1e-9
