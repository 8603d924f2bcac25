"""Host-side pins of the scanned-beam liver benchmark: the restated
cross-section metrics against the reference's reported values, and the DOP
oracle against its closed forms."""
import json
import math
import pathlib

import numpy as np
import pytest

from oracle.oracle import damage_optics
from simulation.scan_suite import CASES, case_key, cross_section_metrics, recorded_metrics, unpack_mask

GOLD = json.loads((pathlib.Path(__file__).parent / 'golden' / 'shimojo_reference.json').read_text())


@pytest.mark.parametrize('key', sorted(GOLD))
def test_recorded_masks_give_the_reported_metrics(key):
    """validation_benchmarks.py:59-80 restated: the recorded x-z masks of each case
    reduce to the metrics the reference reported for it (benchmark_results.json)."""
    gold = GOLD[key]
    m = recorded_metrics(gold)
    assert gold['reported']
    for k, v in gold['reported'].items():
        assert m[k] == pytest.approx(v, abs=1e-9), k
    assert unpack_mask(gold['ablated']).shape == (round(gold['Lx'] / 0.1), round(gold['Lz'] / 0.1))


def test_recorded_cases_are_the_scan_table():
    assert sorted(GOLD) == sorted(case_key(p, v) for p, v in CASES)


def test_cross_section_metrics_small_cases():
    """A hand-made crater: 3 columns deep 2, 4 voxels wide at the top; coagulation
    2 voxels below the deepest column's floor and 6 wide one voxel above it."""
    abl = np.zeros((9, 8), dtype=bool)
    abl[3:7, 0] = True
    abl[4:7, 1] = True
    coag = np.zeros_like(abl)
    coag[1:8, 1] = True
    coag[5, 2:4] = True
    coag[:, 7] = True  # far below, not connected to the floor
    m = cross_section_metrics(abl, coag, 0.1, 0.2)
    assert m['vap_depth'] == pytest.approx(0.4)      # 2 voxels of 0.2
    assert m['vap_width'] == pytest.approx(0.4)      # row 0: indices 3..6
    assert m['coag_depth'] == pytest.approx(0.4)     # column 5, floor k=2: coag at 2, 3, then not
    assert m['coag_width'] == pytest.approx(0.7)     # row 1: 1..7 (coag or ablated)
    none = cross_section_metrics(np.zeros_like(abl), coag, 0.1, 0.2)
    assert none['vap_depth'] == 0.0 and none['coag_depth'] == pytest.approx(1.6)
    assert none['coag_width'] == pytest.approx(0.9)


def test_damage_optics_closed_forms():
    musp0, muspc, g = 0.51, 1.85, 0.9
    Om = np.array([0.0, math.log(2.0), 50.0, 1.0, 3.0])
    abl = np.array([False, False, False, False, True])
    out = damage_optics(np.full(5, 7.0), Om, abl, musp0, muspc, g)
    assert out[0] == pytest.approx(musp0 / (1 - g), rel=1e-15)
    assert out[1] == pytest.approx(0.5 * (musp0 + muspc) / (1 - g), rel=1e-14)
    assert out[2] == pytest.approx(muspc / (1 - g), rel=1e-14)
    a = 1.0 - math.exp(-1.0)
    assert out[3] == pytest.approx(((1 - a) * musp0 + a * muspc) / (1 - g), rel=1e-14)
    assert out[4] == 7.0  # ablated voxels keep their value


def test_kim_depth_and_time_fit():
    """kim_benchmark.py:37-44 and :83-118 restated: the mean lesion depth of a hand-made
    section, and the fitted exposure on a made-up series (running maximum, linear
    interpolation, then the first epoch at or after the fitted time)."""
    from simulation.kim_suite import fit, mean_depth
    m = np.zeros((5, 6), dtype=bool)
    m[1, :2] = True
    m[2, :4] = True
    m[3, 3] = True  # deepest voxel counts, not the number of voxels
    assert mean_depth(m, 0.025) == pytest.approx((2 + 4 + 4) / 3 * 0.025)
    assert mean_depth(np.zeros_like(m), 0.025) == 0.0
    t = [0.002 * (i + 1) for i in range(6)]
    sim = {'t_s': t, 'coag_depth': [0.0, 0.3, 0.5, 0.45, 0.6, 0.7], 'coag_area': [0, 1, 2, 3, 4, 5],
           'abl_depth': [0.0] * 6, 'abl_area': [0.0] * 6}
    f = fit(sim)
    # target 0.535 between 0.5 (t=0.006) and 0.5 (running max at t=0.008), then 0.6 at t=0.010
    assert f['coag']['t_fit_s'] == pytest.approx(np.interp(0.535, [0.0, 0.3, 0.5, 0.5, 0.6, 0.7], t))
    assert f['coag']['area'] == 4 and f['coag']['depth'] == 0.6
    assert not np.isfinite(f['abl']['t_fit_s']) and f['abl']['area'] is None
