"""The wavelength x power campaign: run set (CPU) and concurrent graph epochs (GPU)."""

import numpy as np
import pytest

from simulation.campaign import GRIDS, POWERS, WAVELENGTHS, campaign_runs


SMALL = ((32, 32, 96), (40, 32, 96), (48, 40, 96))


def test_campaign_run_set():
    assert len(GRIDS) == 3 and GRIDS[-1] == (256, 256, 192)
    runs = campaign_runs(n_photons=1000, grids=SMALL)
    assert len(runs) == len(WAVELENGTHS) * len(POWERS) == 24
    shapes = [kw['optical_props']['mu_a'].shape for kw in runs]
    assert shapes[:6] == list(SMALL) * 2
    assert sorted({kw['wavelength_nm'] for kw in runs}) == sorted(WAVELENGTHS)
    for kw in runs:
        mu_a = kw['optical_props']['mu_a']
        k0, k1 = kw['air_gap_range']
        assert np.all(mu_a[:, :, :21] == 0.0) and np.all(mu_a[:, :, k0:k1] == 0.0)
        assert mu_a[:, :, 21].min() > 0.0 and np.all(kw['rho_cp_grid'] > 0)
    # seeded: the same runs twice
    again = campaign_runs(n_photons=1000, grids=SMALL)
    assert all(np.array_equal(a['optical_props']['mu_s'], b['optical_props']['mu_s']) for a, b in zip(runs, again))
    # water dominates absorption at 1940 nm, melanin at 320 nm: deep soft tissue absorbs far more at 1940
    r1940 = [kw for kw in runs if kw['wavelength_nm'] == 1940.0][0]['optical_props']['mu_a'][0, 0, 80]
    r320 = [kw for kw in runs if kw['wavelength_nm'] == 320.0][0]['optical_props']['mu_a'][0, 0, 80]
    assert r1940 > 5 * r320


@pytest.mark.gpu
def test_concurrent_graph_epochs_match_sequential_runs():
    from simulation.campaign import run_campaign
    runs = campaign_runs(n_photons=20_000, grids=((48, 48, 64), (64, 48, 72)), wavelengths=(532.0, 1470.0),
                         powers=(40.0,))
    w1, s1, _ = run_campaign(runs, 6, concurrent=True)
    w2, s2, _ = run_campaign(runs, 6, concurrent=False)
    for a, b in zip(s1, s2):
        # same photons per epoch (seeded keys); the float-atomic tally order differs, and
        # over several epochs of removal that moves the absorbed fraction at the 1e-3 level
        assert a['absorbed_frac'] == pytest.approx(b['absorbed_frac'], rel=1e-2)
        assert abs(a['removed_total'] - b['removed_total']) <= max(2, 0.01 * b['removed_total'])


@pytest.mark.gpu
def test_gang_transport_epochs_match_sequential_runs():
    """All runs' photon transport in one launch traces the same photons as the
    runs one at a time: the FP64 scalar tallies agree to round-off in the first
    epoch, and the state after several epochs within the float-atomic spread."""
    from simulation.campaign import run_campaign
    from simulation.coupled_ablation import CoupledAblationEngine
    runs = campaign_runs(n_photons=20_000, grids=((48, 48, 64), (64, 48, 72), (40, 56, 64)),
                         wavelengths=(532.0, 1470.0), powers=(40.0,))
    ga = [CoupledAblationEngine(graph=True, **kw) for kw in runs]
    sq = [CoupledAblationEngine(graph=True, **kw) for kw in runs]
    d_g = CoupledAblationEngine.step_epochs_gang(ga)
    d_s = [e.step_epoch() for e in sq]
    for a, b in zip(d_g, d_s):
        assert a['absorbed_frac'] == pytest.approx(b['absorbed_frac'], rel=1e-9)
        assert a['mc_energy_closure'] == pytest.approx(b['mc_energy_closure'], rel=1e-9)
        for k in ('reflected', 'transmitted', 'lateral'):
            assert a['ledger_J'][k] == pytest.approx(b['ledger_J'][k], rel=1e-9, abs=1e-15)
    for e in ga + sq:
        e.close()
    w1, s1, _ = run_campaign(runs, 6, mode="gang")
    w2, s2, _ = run_campaign(runs, 6, mode="sequential")
    for a, b in zip(s1, s2):
        assert a['absorbed_frac'] == pytest.approx(b['absorbed_frac'], rel=1e-2)
        assert abs(a['removed_total'] - b['removed_total']) <= max(2, 0.01 * b['removed_total'])


@pytest.mark.gpu
def test_batch_graph_epochs_match_sequential_runs():
    """Every run's captured epoch graph as a child of one parent graph, one launch per round:
    the same photons and the same sub-steps as the runs one at a time (FP64 scalar tallies equal
    to round-off in the first epoch, the state after several epochs within the float-atomic
    spread); the parent is rebuilt when a run's graph is captured again or the run set changes."""
    from simulation.campaign import run_campaign
    from simulation.coupled_ablation import CoupledAblationEngine
    runs = campaign_runs(n_photons=20_000, grids=((48, 48, 64), (64, 48, 72), (40, 56, 64)),
                         wavelengths=(532.0, 1470.0), powers=(40.0,))
    ba = [CoupledAblationEngine(graph=True, **kw) for kw in runs]
    sq = [CoupledAblationEngine(graph=True, **kw) for kw in runs]
    for _ in range(2):
        d_b = CoupledAblationEngine.step_epochs_batch(ba)
        d_s = [e.step_epoch() for e in sq]
        for a, b in zip(d_b, d_s):
            assert a['absorbed_frac'] == pytest.approx(b['absorbed_frac'], rel=1e-6)
            assert a['ledger_J']['reflected'] == pytest.approx(b['ledger_J']['reflected'], rel=1e-6, abs=1e-15)
            assert a['t_optical_s'] > 0.0 and a['t_thermal_s'] > 0.0
    # a new run in place of the first one: a context the parent graph has never held
    ba[0].close()
    ba[0] = CoupledAblationEngine(graph=True, **runs[0])
    sq[0].close()
    sq[0] = CoupledAblationEngine(graph=True, **runs[0])
    d_b = CoupledAblationEngine.step_epochs_batch(ba)
    d_s = [e.step_epoch() for e in sq]
    assert d_b[0]['absorbed_frac'] == pytest.approx(d_s[0]['absorbed_frac'], rel=1e-9)
    for e in ba + sq:
        e.close()
    w1, s1, _ = run_campaign(runs, 6, mode="batch")
    w2, s2, _ = run_campaign(runs, 6, mode="sequential")
    for a, b in zip(s1, s2):
        assert a['absorbed_frac'] == pytest.approx(b['absorbed_frac'], rel=1e-2)
        assert abs(a['removed_total'] - b['removed_total']) <= max(2, 0.01 * b['removed_total'])
