"""Layer optics (simulation/tissue_stack.py) against the reference's own tests
(tests/test_tissue_optics.py, test_api_endpoints.py:26-34) and the evaluated
table of its tissue model notes (BEETLE_TISSUE_MODEL.md section 3: mu_a of four
dorsal layers at 450 / 1064 / 1470 / 1940 nm, two significant figures)."""

import math

import pytest

from simulation.tissue_stack import BASE_LAYERS, eval_single_layer, water_mua

# layer index in BASE_LAYERS -> mu_a [cm^-1] at (450, 1064, 1470, 1940) nm
TABLE = {1: (67, 3.0, 10.6, 48), 2: (1.8, 0.26, 14, 71), 5: (1.8, 0.22, 10, 50), 7: (1.8, 0.27, 20, 99)}


def test_reference_single_layer_cases():
    cut = {'name': 'exocuticle_sclerotised', 'thickness_um': 100.0,
           'composition': {'chitin': 0.6, 'melanin': 0.3, 'water': 0.1},
           'a_prime': 30.0, 'f_rayleigh': 0.2, 'b_mie': 1.1, 'g500': 0.85}
    r = eval_single_layer(cut, wl=532.0)
    assert r['mu_a'] > 0 and r['mu_s'] > 0 and 0.0 <= r['g'] <= 1.0 and r['n'] > 1.0 and not r['is_air']
    air = eval_single_layer({'name': 'subelytral_air_gap', 'thickness_um': 200.0, 'composition': {'air': 1.0}},
                            wl=532.0)
    assert air['mu_a'] == 0.0 and air['mu_s'] == 0.0 and air['is_air'] and air['n'] == 1.0


def test_base_layers_finite_and_only_the_gap_is_air():
    for lay in BASE_LAYERS:
        ev = eval_single_layer(lay, wl=532.0)
        assert ev['is_air'] == (lay['name'] == 'subelytral_air_gap')
        assert ev['mu_a'] >= 0.0 and ev['rho_cp'] > 0.0


def test_water_absorbs_at_1470_in_hemolymph_more_than_in_the_wax():
    ev = [eval_single_layer(l, wl=1470) for l in BASE_LAYERS]
    assert ev[3]['mu_a'] == 0.0 and ev[6]['mu_a'] > ev[0]['mu_a']


@pytest.mark.parametrize("i", sorted(TABLE))
def test_evaluated_table_of_the_tissue_model(i):
    for wl, want in zip((450, 1064, 1470, 1940), TABLE[i]):
        got = eval_single_layer(BASE_LAYERS[i], wl)['mu_a'] * 10.0      # mm^-1 -> cm^-1
        # two significant figures in the table: half a unit of the last digit
        ulp = 10 ** (math.floor(math.log10(want)) - 1)
        assert abs(got - want) <= 0.5 * ulp + 1e-9, (i, wl, got, want)


def test_water_table_interpolates_log_log_and_is_guarded():
    assert water_mua(1450.0)[0] == pytest.approx(29.0)
    mid = math.exp(0.5 * (math.log(29.0) + math.log(24.0)))   # log-log midpoint of 1450/1470 nm
    assert water_mua(math.exp(0.5 * (math.log(1450) + math.log(1470))))[0] == pytest.approx(mid, rel=1e-12)
    with pytest.raises(ValueError):
        water_mua(150.0)
