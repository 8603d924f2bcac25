"""Pins of the oracle's stencil, plateau, Arrhenius and char kernels (CPU only).

Each check is chosen so that a plausible slip in the restatement (a dropped
face, a wrong boundary resistance, an arithmetic-vs-harmonic mean, a sign)
fails: closed forms, conservation laws, an independent ODE solve and the
reference engine's own recorded outputs (tests/golden/reference_engine_values.json).
"""
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_engine_values.json")))["values"]


# ----------------------------------------------------------------- C1: layered stack
def _stack(sandwich, Nz=50):
    k = np.full(Nz, 0.38e-3)
    rc = np.full(Nz, 2.4e-3)
    if sandwich:
        k[8:40] = 0.08e-3
        rc[8:40] = 1.0e-3
    return k, rc


def _engine_layered(O, sandwich, pulse=0.5, total=1.0, T_hot=50.0, T0=25.0, dt=5e-5, dz=0.005):
    """Driver of knowledge/simulation_lit_comparison/code/consistency_checks.py:51-67."""
    Nx = Ny = 4
    Nz = 50
    k1, rc1 = _stack(sandwich, Nz)
    shape = (Nx, Ny, Nz)
    k = np.ascontiguousarray(np.broadcast_to(k1, shape))
    rc = np.ascontiguousarray(np.broadcast_to(rc1, shape))
    T = np.full(shape, T0)
    Tn = np.empty_like(T)
    Q = np.zeros(shape)
    q = 0.0
    for step in range(int(round(total / dt))):
        T[:, :, 0] = T_hot if step * dt < pulse else T0
        O.thermal_step(T, Tn, Q, k, rc, Nx, Ny, Nz, 0.05, 0.05, dz, dt, 1.5e-5, T0, False)
        q += float(k[0, 0, -1] * (Tn[:, :, -2].mean() - T0) / dz) * dt
        Tn[:, :, -1] = T0
        T, Tn = Tn, T
    return q


def _independent_layered(sandwich, pulse=0.5, total=1.0, T_hot=50.0, T0=25.0, dz=0.005):
    """Independent BDF solve of the same finite-volume system (consistency_checks.py:70-92)."""
    from scipy.integrate import solve_ivp
    k, rc = _stack(sandwich)
    kf = 2.0 * k[:-1] * k[1:] / (k[:-1] + k[1:])
    N = len(k)

    def rhs(t, y):
        T = y[:N].copy()
        T[0] = T_hot if t < pulse else T0
        T[-1] = T0
        flux = kf * (T[1:] - T[:-1]) / dz ** 2
        dT = np.zeros(N)
        dT[1:-1] = (flux[1:] - flux[:-1]) / rc[1:-1]
        return np.concatenate([dT, [k[-1] * (T[-2] - T0) / dz]])

    y0 = np.concatenate([np.full(N, T0), [0.0]])
    s1 = solve_ivp(rhs, [0, pulse], y0, method='BDF', rtol=1e-10, atol=1e-12)
    s2 = solve_ivp(rhs, [pulse, total], s1.y[:, -1], method='BDF', rtol=1e-10, atol=1e-12)
    return float(s2.y[-1, -1])


@pytest.mark.parametrize("name,sandwich", [("solid", False), ("sandwich", True)])
def test_c1_layered_stack(oracle, name, sandwich):
    q = _engine_layered(oracle, sandwich) * 1e3  # mJ/mm^2
    # the reference engine's recorded output for the same driver (deterministic)
    assert q == pytest.approx(GOLD[f"C1-{name}"]["model"], rel=1e-9)
    # and the independent ODE solve, to the reference's 0.5 % tolerance
    assert q == pytest.approx(_independent_layered(sandwich) * 1e3, rel=5e-3)


# ----------------------------------------------------------------- C2: radiant heating
@pytest.mark.parametrize("name,alpha", [("dark", 0.92), ("shiny", 0.48)])
def test_c2_lumped_radiant(oracle, name, alpha):
    """consistency_checks.py:104-134: top/bottom series-resistance boundary vs lumped solution."""
    N, dx, dt, duration = 10, 0.1, 0.005, 400.0
    shape = (N, N, N)
    k = np.full(shape, 0.35e-3)
    rc = np.full(shape, 2.8e-3)
    T0, I, h = 22.0, 0.001, 2.05e-5
    T = np.full(shape, T0)
    Tn = np.empty_like(T)
    Q = np.zeros(shape)
    Q[:, :, 0] = alpha * I / dx
    ts, dTs = [], []
    for step in range(int(round(duration / dt))):
        oracle.thermal_step(T, Tn, Q, k, rc, N, N, N, dx, dx, dx, dt, h, T0, True)
        T, Tn = Tn, T
        if step % 200 == 0:
            ts.append(step * dt)
            dTs.append(float(T.mean()) - T0)
    gold = GOLD[f"C2-{name}-end"]["series"]
    np.testing.assert_allclose(dTs, gold["dT"], rtol=1e-9, atol=1e-12)
    tau = 2.8e-3 * 1.0 / (2 * h)
    ss = alpha * I / (2 * h)
    analytic = ss * (1 - np.exp(-np.array(ts) / tau))
    assert dTs[-1] == pytest.approx(analytic[-1], rel=0.03)


# ----------------------------------------------------------------- conservation
def test_energy_conservation_insulated(oracle):
    """V3b (verification_benchmarks.py:231,244-247): with h_eff = 0 every face flux cancels
    pairwise, so sum(rho_cp dT V) equals sum(Q V dt) to round-off (the reference states 1 ppm)."""
    rng = np.random.default_rng(3)
    Nx, Ny, Nz = 9, 7, 11
    shape = (Nx, Ny, Nz)
    k = rng.uniform(0.1e-3, 0.6e-3, shape)
    rc = rng.uniform(1.5e-3, 4.2e-3, shape)
    Q = rng.uniform(0, 500.0, shape)
    T = np.full(shape, 20.0)
    Tn = np.empty_like(T)
    dx, dy, dz, dt, n = 0.02, 0.025, 0.015, 2e-5, 400
    for _ in range(n):
        oracle.thermal_step(T, Tn, Q, k, rc, Nx, Ny, Nz, dx, dy, dz, dt, 0.0, 20.0, True)
        T, Tn = Tn, T
    V = dx * dy * dz
    stored = float(np.sum(rc * (T - 20.0)) * V)
    absorbed = float(np.sum(Q) * V * dt * n)
    assert stored == pytest.approx(absorbed, rel=1e-10)


def test_boundary_flux_matches_series_resistance(oracle):
    """A 1x1x1 column is a lumped body losing heat through two faces with conductance
    h_series = 1 / (0.5 dz / k + 1 / h) per unit area: dT/dt = -2 h_series (T - Tamb) / (rc dz).
    The oracle's single step must equal this to round-off for every (k, h) pair."""
    for kc, h in [(0.3e-3, 1e-5), (0.5e-3, 5e-3), (1e-6, 0.2)]:
        T = np.full((1, 1, 1), 80.0)
        Tn = np.empty_like(T)
        dz, dt, rc = 0.02, 1e-4, 3e-3
        oracle.thermal_step(T, Tn, np.zeros_like(T), np.full_like(T, kc), np.full_like(T, rc), 1, 1, 1,
                            0.05, 0.05, dz, dt, h, 20.0, True)
        hs = 1.0 / (0.5 * dz / kc + 1.0 / h)
        assert Tn[0, 0, 0] == pytest.approx(80.0 - dt * 2 * hs * 60.0 / (rc * dz), rel=1e-12)


def test_reference_kernel_fixture(oracle):
    """tests/test_solvers.py:77-103 of the reference: the centre voxel heats, no NaN/inf."""
    N = 5
    T = np.full((N, N, N), 20.0)
    Tn = np.zeros_like(T)
    Q = np.zeros_like(T)
    Q[2, 2, 2] = 100.0
    oracle.thermal_step(T, Tn, Q, np.full_like(T, 0.2e-3), np.full_like(T, 1.8e-3), N, N, N, 0.1, 0.1, 0.1,
                        0.0005, 1e-5, 20.0, 1)
    assert Tn[2, 2, 2] > 20.0
    assert np.all(np.isfinite(Tn))


# ----------------------------------------------------------------- plateau
def test_plateau_branches_and_enthalpy_invariant(oracle):
    rng = np.random.default_rng(5)
    shape = (6, 5, 4)
    T = rng.uniform(60.0, 140.0, shape)
    H = np.where(rng.random(shape) < 0.5, rng.uniform(0, 0.5, shape), 0.0)
    rc = rng.uniform(2e-3, 4e-3, shape)
    abl = (rng.random(shape) < 0.2).astype(np.uint8)
    dry = ((rng.random(shape) < 0.2) & (abl == 0)).astype(np.uint8)
    T0, H0 = T.copy(), H.copy()
    oracle.plateau(T, H, rc, abl, 100.0, *shape, dry, 120.0)
    wet = (abl == 0) & (dry == 0)
    # enthalpy rc*T + H is conserved for intact wet cells
    np.testing.assert_allclose((rc * T + H)[wet], (rc * T0 + H0)[wet], rtol=1e-14)
    # wet cells never exceed t_abl; banked enthalpy never negative
    assert np.all(T[wet] <= 100.0 + 1e-12) and np.all(H >= 0.0)
    # ablated cells clamped to t_void; dry cells untouched
    a = abl == 1
    np.testing.assert_array_equal(T[a], np.minimum(T0[a], 120.0))
    d = dry == 1
    np.testing.assert_array_equal(T[d], T0[d])
    np.testing.assert_array_equal(H[d | a], H0[d | a])
    # release: a wet cell below t_abl with H>0 rises by at most H/rc and to at most t_abl
    rel = wet & (T0 <= 100.0) & (H0 > 0.0)
    np.testing.assert_allclose(T[rel], np.minimum(100.0, T0[rel] + H0[rel] / rc[rel]), rtol=1e-13)


# ----------------------------------------------------------------- Arrhenius / char
def test_char_c4_golden_and_closed_forms(oracle):
    """consistency_checks.py:170-192 (C4) and tests/test_ablation_extensions.py:46-68."""
    from scipy.integrate import solve_ivp
    A, Ea, R = 1e10, 1.2e5, 8.314
    shape = (2, 2, 2)
    dry = np.ones(shape, np.uint8)
    ab = np.zeros(shape, np.uint8)
    T = np.full(shape, 400.0)
    chi = np.zeros(shape)
    for _ in range(500):
        oracle.char(T, chi, dry, ab, A, Ea / R, 1e-3)
    kr = A * np.exp(-Ea / (R * 673.15))
    assert chi[0, 0, 0] == pytest.approx(GOLD["C4-iso"]["model"], rel=1e-12)
    assert chi[0, 0, 0] == pytest.approx(1 - np.exp(-kr * 0.5), abs=1e-6)
    chi[:] = 0.0
    nst = 2000
    for i in range(nst):
        T[:] = 300.0 + 200.0 * ((i + 0.5) / nst)
        oracle.char(T, chi, dry, ab, A, Ea / R, 2.0 / nst)
    sol = solve_ivp(lambda t, y: (1 - y) * A * np.exp(-Ea / (R * (300.0 + 100.0 * t + 273.15))), [0, 2.0], [0.0],
                    rtol=1e-12, atol=1e-14)
    assert chi[0, 0, 0] == pytest.approx(sol.y[0, -1], abs=1e-3)
    # reference unit fixture: one step at 300 C, rtol 1e-5; non-dry cells untouched
    T = np.full(shape, 300.0)
    chi = np.zeros(shape)
    oracle.char(T, chi, dry, ab, 1.0e6, 8000.0, 1e-4)
    assert np.allclose(chi, 1.0 - np.exp(-1.0e6 * np.exp(-8000.0 / 573.15) * 1e-4), rtol=1e-5)
    before = chi.copy()
    oracle.char(T, chi, np.zeros(shape, np.uint8), ab, 1.0e6, 8000.0, 1e-4)
    np.testing.assert_array_equal(chi, before)


def test_arrhenius_closed_form(oracle):
    shape = (3, 2, 2)
    T = np.full(shape, 70.0)
    Om = np.zeros(shape)
    ab = np.zeros(shape, np.uint8)
    ab[0] = 1
    for _ in range(100):
        oracle.arrhenius(T, Om, ab, 3.1e98, 6.28e5 / 8.314, 1e-4)
    exp_om = 100 * 1e-4 * 3.1e98 * np.exp(-6.28e5 / 8.314 / 343.15)
    np.testing.assert_allclose(Om[1:], exp_om, rtol=1e-12)
    assert np.all(Om[0] == 0.0)


# ------------------------------------------------------------------ VoxelThermalSolver3D (voxel_thermal_3d.py:146-309)
def _solver_grids(n=(8, 8, 6), span=(1.0, 1.0, 0.6)):
    return (np.linspace(-span[0] / 2, span[0] / 2, n[0]), np.linspace(-span[1] / 2, span[1] / 2, n[1]),
            np.linspace(0.0, span[2], n[2]))


def test_oracle_solver_unit_heuristics_and_boundary(oracle):
    """:171-188: k above 0.05 is taken as W/(m K), rho c above 100 as J/(m^3 K); h_eff adds the
    linearised radiation 4 sigma eps T_amb^3.  For h_conv 10 W/(m^2 K), eps 0.9, 22 C:
    4 x 5.670374e-8 x 0.9 x 295.15^3 = 5.24859 W/(m^2 K), so h_eff = 15.24859e-6 W/(mm^2 K)."""
    gx, gy, gz = _solver_grids()
    shape = (8, 8, 6)
    s_si = oracle.OracleThermalSolver3D(gx, gy, gz, np.full(shape, 0.5), np.full(shape, 4.0e6), h_conv=10.0,
                                        emissivity=0.9, t_ambient=22.0)
    s_mm = oracle.OracleThermalSolver3D(gx, gy, gz, np.full(shape, 0.5e-3), np.full(shape, 4.0e-3), h_conv=10.0,
                                        emissivity=0.9, t_ambient=22.0)
    np.testing.assert_allclose(s_si.k_grid, s_mm.k_grid, rtol=1e-15)
    np.testing.assert_allclose(s_si.rho_cp_grid, s_mm.rho_cp_grid, rtol=1e-15)
    assert s_mm.k_grid[0, 0, 0] == 0.5e-3 and s_mm.rho_cp_grid[0, 0, 0] == 4.0e-3
    assert s_si.h_eff == pytest.approx(15.24859e-6, rel=1e-6)
    assert s_si.dx == pytest.approx(1.0 / 7) and s_si.dz == pytest.approx(0.12)


def test_oracle_solver_cfl_raise(oracle):
    """:191-203: after the raise every cell satisfies k / rho_c <= 0.40 / (dt sum 1/d^2), raised cells
    sit exactly on the limit and stable cells keep their rho_c."""
    gx, gy, gz = _solver_grids()
    rng = np.random.default_rng(4)
    shape = (8, 8, 6)
    k = rng.uniform(0.05e-3, 0.6e-3, shape)
    k[:, :, :2] = 0.026e-3        # air-like cells: tiny rho c, large diffusivity
    rc = rng.uniform(2e-3, 4e-3, shape)
    rc[:, :, :2] = 1.2e-6
    s = oracle.OracleThermalSolver3D(gx, gy, gz, k, rc, h_conv=0.0, emissivity=0.0)
    dt = 1e-3
    amax = 0.40 / (dt * (1 / s.dx ** 2 + 1 / s.dy ** 2 + 1 / s.dz ** 2))
    rce = s.cfl_conditioned_rho_cp(dt)
    raised = rce != rc
    assert raised[:, :, :2].all() and not raised[:, :, 2:].any()
    np.testing.assert_allclose((k / rce)[raised], amax, rtol=1e-14)
    assert np.all(k / rce <= amax * (1 + 1e-14))


def test_oracle_solver_conserves_energy_insulated(oracle):
    """With h_conv = 0 and emissivity 0 the boundary is insulated, so the heat stored at the end
    equals the source energy sum Q V dt over the steps the source was on (t < pulse_duration)."""
    gx, gy, gz = _solver_grids()
    shape = (8, 8, 6)
    s = oracle.OracleThermalSolver3D(gx, gy, gz, np.full(shape, 0.2e-3), np.full(shape, 1.8e-3), h_conv=0.0,
                                     emissivity=0.0, t_ambient=25.0)
    Q = np.zeros(shape)
    Q[3:5, 3:5, 0:2] = 20.0
    res = s.solve(Q, pulse_duration=0.005, total_time=0.015, dt=0.001)
    V = s.dx * s.dy * s.dz
    stored = float(np.sum(res['rho_cp_eff'] * (res['T_final'] - 25.0))) * V
    on_steps = sum(1 for t in res['time_history'][:-1] if t < 0.005)
    assert on_steps == 5
    assert stored == pytest.approx(float(Q.sum()) * V * 0.001 * on_steps, rel=1e-12)
    # histories: one entry per step plus the initial state, snapshots at the rounded targets
    assert len(res['time_history']) == 16 and sorted(res['snapshots']) == [0.0, 0.005, 0.015]
    assert res['center_history'][0] == 25.0 and res['peak_surface_T'] >= 25.0


def test_oracle_solver_reference_fixture(oracle):
    """tests/test_solvers.py:41-75 of the reference on the oracle solver."""
    gx, gy, gz = (np.linspace(-0.5, 0.5, 8), np.linspace(-0.5, 0.5, 8), np.linspace(0.0, 0.6, 6))
    s = oracle.OracleThermalSolver3D(gx, gy, gz, np.full((8, 8, 6), 0.2e-3), np.full((8, 8, 6), 1.8e-3),
                                     h_conv=15.0, emissivity=0.95, t_ambient=25.0)
    Q = np.zeros((8, 8, 6))
    Q[3:5, 3:5, 0:2] = 20.0
    res = s.solve(Q, pulse_duration=0.005, total_time=0.015, dt=0.001)
    assert res['peak_surface_T'] >= 25.0 and not np.any(np.isnan(res['time_history']))
    assert len(res['time_history']) > 0 and res['peak_surface_T'] > 25.01
