"""GPU parity of the fused sub-step kernel against the FP64 oracle (through the C-ABI).

Tolerances follow the reference's Metal<->CPU contract (tests/test_gpu_thermal.py:49,133-136):
|dT| < 1e-4 K, |dH|, |dOmega|, |dchi| < 1e-5, plus conservation and golden checks.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_engine_values.json")))["values"]


def _ctx(shape, spacing, **phys):
    from simulation import _bdlib as B
    from simulation.voxel_thermal_3d import _physics
    c = B.Context(shape, spacing)
    c.set_physics(_physics(**phys))
    return c


def test_stencil_parity_reference_fixture(oracle):
    """tests/test_gpu_thermal.py:16-50 of the reference: 12x12x10, two strata, 10 steps."""
    from simulation import _bdlib as B
    Nx, Ny, Nz = 12, 12, 10
    dx, dy, dz, dt = 0.05, 0.05, 0.03, 2e-5
    Q = np.zeros((Nx, Ny, Nz))
    Q[4:8, 4:8, 0:3] = 800.0
    k = np.full((Nx, Ny, Nz), 0.25e-3)
    k[:, :, 5:] = 0.50e-3
    rc = np.full((Nx, Ny, Nz), 1.8e-3)
    rc[:, :, 5:] = 3.5e-3
    T = np.full((Nx, Ny, Nz), 25.0)
    Tn = np.empty_like(T)
    for _ in range(10):
        oracle.thermal_step(T, Tn, Q, k, rc, Nx, Ny, Nz, dx, dy, dz, dt, 12.0, 25.0, 1)
        T, Tn = Tn, T
    c = _ctx((Nx, Ny, Nz), (dx, dy, dz), dt=dt, h_eff=12.0, t_ambient=25.0)
    for name, a in (('T', np.full((Nx, Ny, Nz), 25.0)), ('Q', Q), ('K', k), ('RHO_CP', rc)):
        c.upload(name, a)
    c.substeps(10, 1, 0)
    Tg = c.download('T')
    assert float(np.max(np.abs(Tg - T))) < 1e-4
    assert abs(T.max() - Tg.max()) < 1e-3
    # the module function (one call per step) gives the same field
    from simulation.voxel_thermal_3d import _thermal3d_step_explicit
    T2 = np.full((Nx, Ny, Nz), 25.0)
    T2n = np.empty_like(T2)
    for _ in range(10):
        _thermal3d_step_explicit(T2, T2n, Q, k, rc, Nx, Ny, Nz, dx, dy, dz, dt, 12.0, 25.0, 1)
        T2, T2n = T2n, T2
    assert float(np.max(np.abs(T2 - T))) < 1e-4
    assert B.device_count() >= 1


def test_fused_ablation_parity_reference_fixture(oracle):
    """tests/test_gpu_thermal.py:76-136 of the reference: stencil + plateau + Arrhenius + char, 30 steps."""
    Nx, Ny, Nz = 14, 14, 12
    dx, dy, dz, dt = 0.03, 0.03, 0.02, 2.5e-5
    h_eff, t_amb, t_abl, t_void = 15.0e-6, 20.0, 100.0, 20.0
    A, EaR = 1.0e7, 65000.0 / 8.314
    Q = np.zeros((Nx, Ny, Nz))
    Q[5:9, 5:9, 0:3] = 750.0
    k = np.full((Nx, Ny, Nz), 0.3e-3)
    k[:, :, 6:] = 0.5e-3
    rc = np.full((Nx, Ny, Nz), 2.0e-3)
    rc[:, :, 6:] = 3.5e-3
    dry = np.zeros((Nx, Ny, Nz), np.uint8)
    dry[6:8, 6:8, 0:2] = 1
    abl = np.zeros((Nx, Ny, Nz), np.uint8)
    T = np.full((Nx, Ny, Nz), 25.0)
    Tn = np.empty_like(T)
    H, Om, chi = np.zeros_like(T), np.zeros_like(T), np.zeros_like(T)
    for _ in range(30):
        oracle.thermal_step(T, Tn, Q, k, rc, Nx, Ny, Nz, dx, dy, dz, dt, h_eff, t_amb, 1)
        oracle.plateau(Tn, H, rc, abl, t_abl, Nx, Ny, Nz, dry, t_void)
        oracle.arrhenius(Tn, Om, abl, A, EaR, dt)
        oracle.char(Tn, chi, dry, abl, A, EaR, dt)
        T, Tn = Tn, T
    from simulation import _bdlib as B
    c = _ctx((Nx, Ny, Nz), (dx, dy, dz), dt=dt, h_eff=h_eff, t_ambient=t_amb, t_abl=t_abl, t_void=t_void,
             has_arrhenius=1, arr_A=A, arr_Ea_over_R=EaR, has_char=1, char_A=A, char_Ea_over_R=EaR)
    for name, a in (('T', np.full((Nx, Ny, Nz), 25.0)), ('Q', Q), ('K', k), ('RHO_CP', rc), ('DRY', dry),
                    ('ABLATED', abl)):
        c.upload(name, a)
    for name in ('H_ACC', 'OMEGA', 'CHI'):
        c.fill(name, 0.0)
    c.substeps(30, 1, B.SUB_PLATEAU | B.SUB_ARRHENIUS | B.SUB_CHAR)
    assert float(np.max(np.abs(c.download('T') - T))) < 1e-4
    assert float(np.max(np.abs(c.download('H_ACC') - H))) < 1e-5
    assert float(np.max(np.abs(c.download('OMEGA') - Om))) < 1e-5
    assert float(np.max(np.abs(c.download('CHI') - chi))) < 1e-5
    assert H.max() > 0.0 and chi.max() > 0.0  # the fixture exercises both branches


@pytest.mark.parametrize("shape", [(40, 37, 96), (128, 128, 96), (1, 1, 1), (1, 5, 7), (3, 1, 70), (2, 3, 1),
                                   (5, 4, 2), (2, 2, 129)])
def test_multi_tile_parity_with_plateau(oracle, shape):
    """Several z tiles (96 = 64 + 32, a partial tile), partial y tiles, the tuned tile shape
    (>= 1M voxels) and every plateau branch, against the oracle over 20 sub-steps; also
    degenerate grids (one cell, one column, one plane, one row: both boundary fluxes
    on one cell, no x or y neighbours) and a z extent one past two tiles."""
    from simulation import _bdlib as B
    rng = np.random.default_rng(1)
    Nx, Ny, Nz = shape
    dx, dy, dz, dt = 0.015, 0.015, 0.015, 5e-5
    k = rng.uniform(0.1e-3, 0.6e-3, shape)
    rc = rng.uniform(2.0e-3, 4.2e-3, shape)
    T0 = rng.uniform(60.0, 120.0, shape)
    Q = rng.uniform(0.0, 300.0, shape)
    dry = (rng.random(shape) < 0.15).astype(np.uint8)
    abl = ((rng.random(shape) < 0.1) & (dry == 0)).astype(np.uint8)
    H0 = np.where(rng.random(shape) < 0.3, rng.uniform(0, 0.3, shape), 0.0)
    # the device stores FP32: give the oracle the same stored values
    f = lambda a: a.astype(np.float32).astype(np.float64)
    k, rc, T0, Q, H0 = f(k), f(rc), f(T0), f(Q), f(H0)
    T, H = T0.copy(), H0.copy()
    Tn = np.empty_like(T)
    Qm = np.where(abl == 1, 0.0, Q)
    for _ in range(20):
        oracle.thermal_step(T, Tn, Qm, k, rc, Nx, Ny, Nz, dx, dy, dz, dt, 1.5e-5, 20.0, 1)
        oracle.plateau(Tn, H, rc, abl, 100.0, Nx, Ny, Nz, dry, 110.0)
        T, Tn = Tn, T
    c = _ctx(shape, (dx, dy, dz), dt=dt, h_eff=1.5e-5, t_ambient=20.0, t_abl=100.0, t_void=110.0)
    for name, a in (('T', T0), ('Q', Q), ('K', k), ('RHO_CP', rc), ('DRY', dry), ('ABLATED', abl), ('H_ACC', H0)):
        c.upload(name, a)
    c.fill('CHI', 0.0)
    c.substeps(20, 1, B.SUB_PLATEAU | B.SUB_MASK_Q)
    dT = np.abs(c.download('T') - T)
    assert float(dT.max()) < 1e-4, f"max |dT| {dT.max():.3e} at {np.unravel_index(dT.argmax(), shape)}"
    assert float(np.max(np.abs(c.download('H_ACC') - H))) < 1e-5


def test_energy_conservation_insulated_gpu():
    """V3b on the device: with h_eff = 0 stored energy equals deposited energy (FP32: 1e-6)."""
    rng = np.random.default_rng(2)
    shape = (64, 48, 80)
    dx, dt, n = 0.015, 5e-5, 200
    k = rng.uniform(0.1e-3, 0.6e-3, shape)
    rc = rng.uniform(2.0e-3, 4.0e-3, shape).astype(np.float32)
    Q = rng.uniform(0, 200.0, shape).astype(np.float32)
    c = _ctx(shape, (dx, dx, dx), dt=dt, h_eff=0.0, t_ambient=20.0)
    for name, a in (('Q', Q), ('K', k), ('RHO_CP', rc)):
        c.upload(name, a)
    c.fill('T', 20.0)
    c.substeps(n, 1, 0)
    V = dx ** 3
    stored = float(np.sum(rc.astype(np.float64) * (c.download('T') - 20.0)) * V)
    absorbed = float(np.sum(Q.astype(np.float64)) * V * dt * n)
    assert stored == pytest.approx(absorbed, rel=1e-6)


def test_c2_radiant_golden_on_gpu():
    """The reference engine's recorded C2 dark-body series (consistency_checks.py:104-134), 20 001 and
    79 801 sub-steps on the device.  The temperature is stored in FP32 and each step adds a small,
    nearly constant increment, so rounding biases the sum: measured 2.5e-4 relative after 20 001
    steps (the FP64 oracle matches the same series to 1e-9, test_oracle_thermal.py)."""
    N, dx, dt = 10, 0.1, 0.005
    shape = (N, N, N)
    Q = np.zeros(shape)
    Q[:, :, 0] = 0.92 * 0.001 / dx
    c = _ctx(shape, (dx, dx, dx), dt=dt, h_eff=2.05e-5, t_ambient=22.0)
    c.upload('Q', Q)
    c.upload('K', np.full(shape, 0.35e-3))
    c.upload('RHO_CP', np.full(shape, 2.8e-3))
    c.fill('T', 22.0)
    gold = GOLD["C2-dark-end"]["series"]
    # series sampled every 200 steps (1 s); check t = 100 s and t = 399 s
    c.substeps(100 * 200 + 1, 1, 0)
    assert float(c.download('T').mean()) - 22.0 == pytest.approx(gold["dT"][100], rel=5e-4)
    c.substeps(299 * 200, 1, 0)
    assert float(c.download('T').mean()) - 22.0 == pytest.approx(gold["dT"][399], rel=1e-3)


def test_module_kernels_match_oracle(oracle):
    """_apply_latent_heat_plateau, _accumulate_arrhenius, _accumulate_char drop-ins vs the oracle."""
    from simulation.coupled_ablation import _accumulate_arrhenius, _accumulate_char
    from simulation.voxel_thermal_3d import _apply_latent_heat_plateau
    rng = np.random.default_rng(3)
    shape = (9, 8, 7)
    f = lambda a: a.astype(np.float32).astype(np.float64)
    T = f(rng.uniform(60, 140, shape))
    H = f(np.where(rng.random(shape) < 0.5, rng.uniform(0, 0.5, shape), 0.0))
    rc = f(rng.uniform(2e-3, 4e-3, shape))
    abl = rng.random(shape) < 0.2
    dry = (rng.random(shape) < 0.2) & ~abl
    To, Ho = T.copy(), H.copy()
    oracle.plateau(To, Ho, rc, abl.astype(np.uint8), 100.0, *shape, dry.astype(np.uint8), 120.0)
    Tg, Hg = T.copy(), H.copy()
    _apply_latent_heat_plateau(Tg, Hg, rc, abl, 100.0, *shape, dry, 120.0)
    assert np.max(np.abs(Tg - To)) < 1e-4 and np.max(np.abs(Hg - Ho)) < 1e-6
    Om_o, Om_g = np.zeros(shape), np.zeros(shape)
    oracle.arrhenius(T, Om_o, abl.astype(np.uint8), 1e7, 65000 / 8.314, 1e-3)
    _accumulate_arrhenius(T, Om_g, abl, 1e7, 65000 / 8.314, 1e-3, *shape)
    np.testing.assert_allclose(Om_g, Om_o, rtol=2e-5, atol=1e-12)
    chi_o, chi_g = np.full(shape, 0.2), np.full(shape, 0.2)
    oracle.char(T + 200, chi_o, dry.astype(np.uint8), abl.astype(np.uint8), 1e7, 65000 / 8.314, 1e-3)
    _accumulate_char(T + 200, chi_g, dry, abl, 1e7, 65000 / 8.314, 1e-3, *shape)
    np.testing.assert_allclose(chi_g, chi_o, rtol=0, atol=1e-6)


def test_voxel_thermal_solver_gpu_reference_fixture():
    """tests/test_solvers.py:41-75 / test_gpu_thermal.py:52-74 of the reference on the device."""
    from simulation.voxel_thermal_3d import VoxelThermalSolver3D
    solver = VoxelThermalSolver3D(np.linspace(-0.5, 0.5, 8), np.linspace(-0.5, 0.5, 8), np.linspace(0, 0.6, 6),
                                  np.full((8, 8, 6), 0.2e-3), np.full((8, 8, 6), 1.8e-3), h_conv=15.0,
                                  emissivity=0.95, t_ambient=25.0)
    Q = np.zeros((8, 8, 6))
    Q[3:5, 3:5, 0:2] = 20.0
    res = solver.solve(Q_3d=Q, pulse_duration=0.005, total_time=0.015, dt=0.001)
    assert res['engine'] == 'gpu'
    assert res['peak_surface_T'] >= 25.0 and not np.any(np.isnan(res['time_history']))
    assert len(res['time_history']) == 16


@pytest.mark.parametrize("shape", [(37, 41, 125), (160, 136, 192)])
def test_paired_substeps_bit_identical_to_single_substeps(shape, monkeypatch):
    """The two-sub-step kernel (temporal blocking, ping-pong H / flags) against one launch
    per sub-step: every field bit-identical after 22 sub-steps (10 pair launches + 2
    single ones) with the plateau, Arrhenius, char, Q masking and T_peak, on partial
    tiles in every axis and (second shape, above 1M voxels) the autotuned shapes."""
    from simulation import _bdlib as B
    rng = np.random.default_rng(11)
    dx, dt = 0.015, 5e-5
    k = rng.uniform(0.1e-3, 0.6e-3, shape)
    rc = rng.uniform(2.0e-3, 4.2e-3, shape)
    T0 = rng.uniform(60.0, 140.0, shape)
    Q = rng.uniform(0.0, 400.0, shape)
    dry = (rng.random(shape) < 0.2).astype(np.uint8)
    abl = ((rng.random(shape) < 0.1) & (dry == 0)).astype(np.uint8)
    H0 = np.where(rng.random(shape) < 0.3, rng.uniform(0, 0.3, shape), 0.0)
    chi0 = rng.uniform(0.0, 0.5, shape)
    out = []
    for pairs in ("0", "1"):
        monkeypatch.setenv("BUGDEF_SUBSTEP_PAIRS", pairs)
        c = _ctx(shape, (dx, dx, dx), dt=dt, h_eff=1.5e-5, t_ambient=20.0, t_abl=100.0, t_void=110.0,
                 has_arrhenius=1, arr_A=3.1e98, arr_Ea_over_R=6.28e5 / 8.314, has_char=1, char_A=1.0e7,
                 char_Ea_over_R=65000.0 / 8.314)
        for name, a in (('T', T0), ('Q', Q), ('K', k), ('RHO_CP', rc), ('DRY', dry), ('ABLATED', abl),
                        ('H_ACC', H0), ('CHI', chi0)):
            c.upload(name, a)
        c.fill('OMEGA', 0.0)
        c.fill('T_PEAK', 0.0)
        c.substeps(22, 1, B.SUB_PLATEAU | B.SUB_MASK_Q | B.SUB_ARRHENIUS | B.SUB_CHAR | B.SUB_PEAK)
        out.append({n: c.download(n) for n in ('T', 'H_ACC', 'CHI', 'OMEGA', 'T_PEAK', 'ABLATED', 'DRY')})
        c.close()
    for n in out[0]:
        np.testing.assert_array_equal(out[0][n], out[1][n], err_msg=n)


def test_substep_tile_shape_does_not_change_the_result(monkeypatch):
    """Every tile shape evaluates each cell from the same inputs in the same order,
    so the fields after 20 sub-steps are bit-identical across the shapes the
    timing pass can pick (fixed here with BUGDEF_SUBSTEP_TILE)."""
    from simulation import _bdlib as B
    shape = (70, 45, 130)   # partial tiles in every axis
    rng = np.random.default_rng(12)
    dx, dt = 0.015, 5e-5
    k = rng.uniform(0.1e-3, 0.6e-3, shape)
    rc = rng.uniform(2.0e-3, 4.2e-3, shape)
    T0 = rng.uniform(60.0, 140.0, shape)
    Q = rng.uniform(0.0, 400.0, shape)
    dry = (rng.random(shape) < 0.2).astype(np.uint8)
    H0 = np.where(rng.random(shape) < 0.3, rng.uniform(0, 0.3, shape), 0.0)
    chi0 = rng.uniform(0.0, 0.5, shape)
    monkeypatch.setenv("BUGDEF_SUBSTEP_PAIRS", "0")
    out = []
    for tile in ("2x8", "4x16", "8x64", "4x5"):
        monkeypatch.setenv("BUGDEF_SUBSTEP_TILE", tile)
        c = _ctx(shape, (dx, dx, dx), dt=dt, h_eff=1.5e-5, t_ambient=20.0, t_abl=100.0, t_void=110.0,
                 has_char=1, char_A=1.0e7, char_Ea_over_R=65000.0 / 8.314)
        for name, a in (('T', T0), ('Q', Q), ('K', k), ('RHO_CP', rc), ('DRY', dry), ('H_ACC', H0), ('CHI', chi0)):
            c.upload(name, a)
        c.fill('T_PEAK', 0.0)
        c.substeps(20, 1, B.SUB_PLATEAU | B.SUB_MASK_Q | B.SUB_CHAR | B.SUB_PEAK)
        by, xc = (int(v) for v in tile.split("x"))
        t = c.tuning()
        assert (t["substep_tile_y"], t["substep_tile_x"]) == (by, xc)
        out.append({n: c.download(n) for n in ('T', 'H_ACC', 'CHI', 'T_PEAK', 'DRY')})
        c.close()
    for o in out[1:]:
        for n in out[0]:
            np.testing.assert_array_equal(out[0][n], o[n], err_msg=n)


@pytest.mark.parametrize("case", ["reference_fixture", "si_units_with_air"])
def test_voxel_thermal_solver_histories_match_oracle(oracle, case):
    """VoxelThermalSolver3D (voxel_thermal_3d.py:146-309) on the device against the FP64 oracle
    solver: every entry of the centre and surface histories and every snapshot within 1e-4 K.
    The first case is the reference's own fixture (tests/test_solvers.py:41-75); the second passes
    k and rho c in SI units (the unit heuristics) with an air layer that the CFL raise conditions."""
    from simulation.voxel_thermal_3d import VoxelThermalSolver3D
    if case == "reference_fixture":
        n = (8, 8, 6)
        grids = (np.linspace(-0.5, 0.5, 8), np.linspace(-0.5, 0.5, 8), np.linspace(0.0, 0.6, 6))
        k, rc = np.full(n, 0.2e-3), np.full(n, 1.8e-3)
        Q = np.zeros(n)
        Q[3:5, 3:5, 0:2] = 20.0
        kw = dict(h_conv=15.0, emissivity=0.95, t_ambient=25.0)
        run = dict(pulse_duration=0.005, total_time=0.015, dt=0.001)
    else:
        n = (24, 20, 30)
        grids = (np.arange(24) * 0.05, np.arange(20) * 0.05, np.arange(30) * 0.05)
        k, rc = np.full(n, 0.5), np.full(n, 4.0e6)          # W/(m K), J/(m^3 K)
        k[:, :, :4], rc[:, :, :4] = 0.026, 1206.0            # air on top
        Q = np.zeros(n)
        Q[9:15, 8:12, 4:12] = 10.0
        kw = dict(h_conv=25.0, emissivity=0.9, t_ambient=20.0)
        run = dict(pulse_duration=0.03, total_time=0.06, dt=0.001, snapshot_times=[0.0, 0.01, 0.03, 0.045])
    g = VoxelThermalSolver3D(*grids, k_grid=k, rho_cp_grid=rc, **kw).solve(Q_3d=Q, **run)
    o_solver = oracle.OracleThermalSolver3D(*grids, k_grid=k, rho_cp_grid=rc, **kw)
    r = o_solver.solve(Q, **run)
    if case != "reference_fixture":
        assert np.any(o_solver.cfl_conditioned_rho_cp(run['dt']) != o_solver.rho_cp_grid)
    np.testing.assert_array_equal(g['time_history'], r['time_history'])
    for key in ('center_history', 'surface_history'):
        d = float(np.max(np.abs(g[key] - r[key])))
        print(case, key, "max |dT|", d)
        assert d < 1e-4, key
    assert sorted(g['snapshots']) == sorted(r['snapshots'])
    for t in r['snapshots']:
        assert float(np.max(np.abs(g['snapshots'][t] - r['snapshots'][t]))) < 1e-4, t
    assert abs(g['peak_surface_T'] - r['peak_surface_T']) < 1e-4
    assert abs(g['peak_center_T'] - r['peak_center_T']) < 1e-4
    assert max(r['peak_surface_T'], r['peak_center_T']) - kw['t_ambient'] > 1.0
