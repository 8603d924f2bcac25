"""The reference's coupled-engine verification cases V3a, V3b and V4 on the device engine.

verification_benchmarks.py:189-295 runs them through CoupledAblationEngine on a quasi-1-D column
(0.2 x 0.2 mm, insulated sides) under a beam ten times wider than the column.  Each test
restates the set-up, checks the closed form the reference checks (with its tolerance), and
checks the reference engine's recorded model value (tests/golden/reference_engine_values.json,
from benchmark_results.json) within the spread the photon shot noise allows.
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_engine_values.json")


def _recorded():
    return json.load(open(GOLDEN))["values"]


def _quasi_1d(mu_a, k, rho_cp, h_lat, t_abl, Lz, Nz, I_target, dt_epoch, dt_sub, n_photons, seed):
    """verification_benchmarks.py:189-211."""
    from simulation.coupled_ablation import CoupledAblationEngine
    Lx = 0.2
    w0 = 10.0 * (Lx / 2.0)
    shape = (4, 4, Nz)
    return CoupledAblationEngine(
        optical_props={'mu_a': np.full(shape, mu_a), 'mu_s': np.zeros(shape), 'g': np.zeros(shape),
                       'n': np.ones(shape)},
        k_grid=np.full(shape, k), rho_cp_grid=np.full(shape, rho_cp), h_lat_grid=np.full(shape, h_lat), Lx=Lx, Ly=Lx,
        Lz=Lz, power=I_target * np.pi * w0 ** 2 / 2.0, spot_dia=2.0 * w0, t_abl=t_abl, t_ambient=20.0,
        dt_epoch=dt_epoch, dt_sub=dt_sub, n_photons=n_photons, h_eff=0.0, seed=seed)


def _surface_rise(I, mu_a, k, rho_cp, t):
    """verification_benchmarks.py:177-186: insulated semi-infinite medium, Q = I mu_a exp(-mu_a z)."""
    from scipy.special import erfcx
    alpha = k / rho_cp
    x = mu_a * np.sqrt(alpha * t)
    return (I / k) * (2.0 * np.sqrt(alpha * t / np.pi) - (1.0 - erfcx(x)) / mu_a)


def test_v3_beer_lambert_heating_and_energy_conservation():
    """V3a: surface rise at 0.5 s within 2 % of the closed form at the delivered irradiance, and
    within 1 % of the reference engine's 190.862 K.  V3b: stored heat equals absorbed energy
    (insulated column) to 1e-5 relative in FP32 (the reference's FP64 engine: 1 ppm), and both
    within 1 % of the recorded 0.0098745 J."""
    mu_a, k, rho_cp = 10.0, 0.53e-3, 4.18e-3
    eng = _quasi_1d(mu_a, k, rho_cp, 1e9, 1e9, 1.5, 150, 0.5, 0.05, 1e-4, 2_000_000, seed=21)
    I_hist, absorbed_J, ledger = [], 0.0, 0.0
    for _ in range(10):
        d = eng.step_epoch()
        Q = np.asarray(eng.Q)
        I_hist.append(float(Q.sum() * eng.voxel_vol / (eng.Lx * eng.Ly)))
        absorbed_J += float(Q.sum()) * eng.voxel_vol * eng.dt_epoch
        ledger += d['ledger_J']['absorbed']
        assert d['ledger_J']['conducted'] == 0.0
    T = np.asarray(eng.T)
    dT = float(T[:, :, 0].mean() - 20.0)
    I_eff = float(np.mean(I_hist))
    ref = float(_surface_rise(I_eff, mu_a, k, rho_cp, eng.t))
    stored = float(np.sum(rho_cp * (T - 20.0)) * eng.voxel_vol)
    rec = _recorded()
    print(f"V3a: dT {dT:.4f} K closed form {ref:.4f} K recorded {rec['V3a']['model']:.4f} K (I_eff {I_eff:.5f})")
    print(f"V3b: stored {stored:.9e} absorbed {absorbed_J:.9e} ledger {ledger:.9e} "
          f"recorded {rec['V3b']['model']:.9e}")
    assert abs(dT - ref) <= 0.02 * ref
    assert abs(dT - rec['V3a']['model']) <= 0.01 * rec['V3a']['model']
    assert abs(stored - absorbed_J) <= 1e-5 * absorbed_J
    assert ledger == pytest.approx(absorbed_J, rel=1e-9)
    assert abs(stored - rec['V3b']['model']) <= 0.01 * rec['V3b']['model']
    eng.close()


@pytest.mark.parametrize("dt_epoch", [0.002, 0.010])
def test_v4_ablation_front_velocity(dt_epoch):
    """V4 (verification_benchmarks.py:252-295): steady front velocity over the last 0.5 s of a
    1.5 s run against the energy balance v = I / (rho c (Tv - T0) + rho L) (5 %, as the reference;
    with the 10 ms optical update the reference's own engine misses it, so only the recorded value
    is checked there), against the reference engine's recorded velocity (3 %) and its recorded
    depth series (within 2.5 voxels of 5 um at every epoch)."""
    mu_a, k, rho_cp, h_lat = 60.0, 0.53e-3, 4.18e-3, 2.257
    eng = _quasi_1d(mu_a, k, rho_cp, h_lat, 100.0, 0.6, 120, 0.5, dt_epoch, 5e-5, 1_000_000, seed=31)
    n_ep = int(round(1.5 / dt_epoch))
    times, depth, I_hist = [], [], []
    for _ in range(n_ep):
        eng.step_epoch()
        I_hist.append(float(np.asarray(eng.Q).sum() * eng.voxel_vol / (eng.Lx * eng.Ly)))
        times.append(eng.t)
        depth.append(float(np.asarray(eng.ablated).sum() / 16.0 * eng.dz))
    times, depth = np.array(times), np.array(depth)
    I_eff = float(np.mean(I_hist))
    v_ref = I_eff / (rho_cp * 80.0 + h_lat)
    fit = times >= times[-1] - 0.5
    v_model = float(np.polyfit(times[fit], depth[fit], 1)[0])
    key = f"V4-{int(round(dt_epoch * 1000))}ms"
    rec = _recorded()[key]
    d_rec = np.asarray(rec['series']['depth_mm'])
    dd = float(np.max(np.abs(depth - d_rec)))
    print(f"{key}: v {v_model:.5f} mm/s energy balance {v_ref:.5f} recorded {rec['model']:.5f}; "
          f"max |depth - recorded| {dd * 1000:.2f} um")
    if dt_epoch == 0.002:
        assert abs(v_model - v_ref) <= 0.05 * v_ref
    assert abs(v_model - rec['model']) <= 0.03 * rec['model']
    assert np.allclose(times, rec['series']['t_s'], atol=1e-9)
    assert dd <= 2.5 * eng.dz + 1e-12
    eng.close()
