"""Layer optics from chromophore composition (drop-in for the reference's
``simulation/tissue_stack.py``: ``BASE_LAYERS``, ``CHROMOPHORE_HEAT_CAPACITIES``,
``eval_single_layer``).

Host-side and deterministic: it turns a layer description (thickness,
volume-fraction composition, scattering power-law parameters) into the
per-layer mu_a, mu_s, g, n and rho_cp that the device engine's property grids
are filled with (per wavelength, once per run or per wavelength switch).

The chromophore basis follows the reference's tissue-optics closure
(knowledge/invertebrate_tissue_optics/insect_tissue_optics/basis.py and
layers.py): Beer-Lambert mixing of absorption over volume fractions plus the
non-pigmented cellular baseline, volume-weighted refractive index, the
two-term Rayleigh/Mie reduced-scattering power law and the wavelength-decaying
anisotropy.  The water table, the power laws and the band parameters are
literature data and are kept identical.  The OMLC hemoglobin extinction table
is not carried: a layer containing hemoglobin takes the same fallback formula
the reference takes when its basis evaluation fails.
"""

from __future__ import annotations

import math

import numpy as np

# ------------------------------------------------------------------ basis spectra (cm^-1)
_WATER_NM = np.array([
    200, 250, 300, 350, 400, 420, 440, 450, 460, 480, 500, 520, 540, 560, 580, 600, 620, 640, 660, 680,
    700, 720, 740, 760, 780, 800, 820, 840, 860, 880, 900, 920, 940, 960, 970, 980, 1000, 1050, 1100, 1150,
    1200, 1250, 1300, 1350, 1400, 1450, 1470, 1500, 1550, 1600, 1650, 1700, 1750, 1800, 1850, 1900, 1930, 1940,
    1950, 2000, 2100, 2200, 2300, 2400, 2500, 2600, 2700, 2800, 2900, 2940, 3000, 3100, 3200, 3400, 3600, 3800,
    4000, 4400, 4800, 5000, 5500, 6100, 6500, 7000, 8000, 9000, 10000, 10600, 11000, 12000], dtype=float)
_WATER_MUA = np.array([
    0.0310, 0.0017, 0.00067, 0.00023, 0.000663, 0.000530, 0.000640, 0.000922, 0.00106, 0.00157,
    0.00257, 0.00429, 0.00565, 0.00619, 0.00896, 0.02224, 0.02760, 0.03385, 0.04150, 0.04580,
    0.00650, 0.01120, 0.01960, 0.02645, 0.02210, 0.02020, 0.02430, 0.03230, 0.04290, 0.05190,
    0.06800, 0.10600, 0.16200, 0.32200, 0.44000, 0.42600, 0.36300, 0.15000, 0.15900, 0.48000,
    1.03000, 0.76800, 1.39000, 5.80000, 12.3000, 29.0000, 24.0000, 20.3000, 11.0000, 6.70000,
    5.30000, 5.70000, 7.40000, 11.5000, 25.0000, 68.0000, 115.000, 120.000, 118.000, 69.0000,
    36.0000, 23.0000, 29.0000, 42.0000, 56.0000, 67.2000, 370.000, 2576.00, 7887.00, 12000.0,
    11390.0, 6200.00, 3630.00, 1360.00, 339.000, 119.000, 107.000, 103.000, 351.000, 312.000,
    461.000, 2760.00, 1348.00, 704.000, 532.000, 617.000, 638.000, 800.000, 826.000, 1032.00], dtype=float)
_WATER_RANGE = (200.0, 12000.0)

N_CHITIN, N_PROTEIN, N_LIPID, N_ARAGONITE, N_MUCIN_SOLIDS = 1.555, 1.550, 1.460, 1.632, 1.500


def _gauss(nm, centre, width, amp):
    return amp * np.exp(-0.5 * ((nm - centre) / width) ** 2)


def water_mua(nm):
    """Pure water: log-log interpolation of the tabulation; outside 200-12000 nm it raises."""
    nm = np.atleast_1d(np.asarray(nm, float))
    if nm.min() < _WATER_RANGE[0] or nm.max() > _WATER_RANGE[1]:
        raise ValueError(f"water absorption is tabulated for {_WATER_RANGE} nm only")
    return np.exp(np.interp(np.log(nm), np.log(_WATER_NM), np.log(_WATER_MUA)))


def water_n(nm):
    nm = np.atleast_1d(np.asarray(nm, float))
    return 1.31279 + 15.762 / nm - 4382.0 / nm ** 2 + 1.1455e6 / nm ** 3


def baseline_mua(nm):
    """Non-pigmented cellular baseline, 1.321 (lambda/500)^-3."""
    return 1.321 * (np.atleast_1d(np.asarray(nm, float)) / 500.0) ** -3.0


def melanin_mua(nm):
    return 519.0 * (np.atleast_1d(np.asarray(nm, float)) / 500.0) ** -3.5


def ommochrome_mua(nm, peak_mua=600.0):
    nm = np.atleast_1d(np.asarray(nm, float))
    return peak_mua * (_gauss(nm, 440.0, 55.0, 1.0) + _gauss(nm, 500.0, 45.0, 0.35) + _gauss(nm, 370.0, 40.0, 0.8))


def plant_pigment_mua(nm, chl_mg_per_ml=0.30, package_factor=0.45):
    nm = np.atleast_1d(np.asarray(nm, float))
    bands = ((435, 22, 1.28), (410, 18, 0.55), (665, 11, 1.00), (615, 16, 0.14), (578, 16, 0.07),
             (460, 20, 0.62), (650, 11, 0.35), (448, 26, 0.70), (475, 24, 0.55), (425, 22, 0.35))
    shape = sum(_gauss(nm, c, w, a) for c, w, a in bands)
    return 96.3 * chl_mg_per_ml * package_factor * shape


def hemocyanin_mua(nm, fu_millimolar=1.0, oxygenation=0.8):
    nm = np.atleast_1d(np.asarray(nm, float))
    c = fu_millimolar * 1e-3
    ln10 = math.log(10.0)
    return (ln10 * c * (16000.0 * oxygenation * _gauss(nm, 346.0, 35.0, 1.0)
                        + 500.0 * oxygenation * _gauss(nm, 578.0, 60.0, 1.0))
            + ln10 * _gauss(nm, 280.0, 25.0, 60.0) * c)


def _ch_overtones(nm):
    return _gauss(nm, 1730.0, 45.0, 3.0) + _gauss(nm, 1200.0, 40.0, 1.5) + _gauss(nm, 2300.0, 60.0, 6.0)


def chitin_mua(nm):
    nm = np.atleast_1d(np.asarray(nm, float))
    return 0.30 * water_mua(nm) + _ch_overtones(nm) + 0.05


def lipid_mua(nm):
    nm = np.atleast_1d(np.asarray(nm, float))
    return 0.02 + _gauss(nm, 1210.0, 30.0, 8.0) + _gauss(nm, 1720.0, 35.0, 12.0) + _gauss(nm, 2300.0, 50.0, 25.0)


def aragonite_mua(nm):
    nm = np.atleast_1d(np.asarray(nm, float))
    return 0.02 + _gauss(nm, 2350.0, 60.0, 15.0)


def aragonite_n(nm):
    nm = np.atleast_1d(np.asarray(nm, float))
    return N_ARAGONITE + 0.012 * (500.0 / nm) ** 2 - 0.0048


def conchiolin_mua(nm):
    nm = np.atleast_1d(np.asarray(nm, float))
    return 45.0 * (nm / 500.0) ** -2.6 + 0.30 * water_mua(nm)


def mus_reduced(nm, a_prime, f_rayleigh, b_mie):
    x = np.atleast_1d(np.asarray(nm, float)) / 500.0
    return a_prime * (f_rayleigh * x ** -4.0 + (1.0 - f_rayleigh) * x ** -b_mie)


def anisotropy(nm, g500, power=0.20):
    nm = np.atleast_1d(np.asarray(nm, float))
    decay = np.where(nm > 500.0, (500.0 / nm) ** power, 1.0 + 0.05 * (1.0 - nm / 500.0))
    return np.clip(g500 * decay, 0.0, 0.99)


# component -> (absorption [cm^-1] at unit fraction, refractive index function or constant)
_COMPONENTS = {
    'water': (water_mua, water_n),
    'chitin': (chitin_mua, N_CHITIN),
    'protein': (chitin_mua, N_PROTEIN),
    'lipid': (lipid_mua, N_LIPID),
    'melanin': (melanin_mua, 1.70),
    'ommochrome': (ommochrome_mua, 1.60),
    'plant_pigment': (plant_pigment_mua, 1.50),
    'hemocyanin': (hemocyanin_mua, 1.55),
    'aragonite': (aragonite_mua, aragonite_n),
    'conchiolin': (conchiolin_mua, 1.56),
    'mucin': (chitin_mua, N_MUCIN_SOLIDS),
    'air': (lambda nm: np.zeros_like(np.atleast_1d(np.asarray(nm, float))), 1.0003),
}


class _BasisUnavailable(Exception):
    pass


def _layer_optics(comp, wl, a_prime, f_ray, b_mie, g500):
    """(mu_a, mu_s', mu_s, g, n) in cm^-1 from the chromophore basis."""
    nm = np.array([float(wl)])
    if abs(sum(comp.values()) - 1.0) > 1e-3:
        raise _BasisUnavailable("fractions do not sum to one")
    mua = np.zeros(1)
    n = np.zeros(1)
    for c, f in comp.items():
        if c not in _COMPONENTS:
            raise _BasisUnavailable(c)     # e.g. hemoglobin: no extinction table here
        if f == 0.0:
            continue
        fa, fn = _COMPONENTS[c]
        if c != 'air':
            mua = mua + f * fa(nm)
        n = n + f * (fn(nm) if callable(fn) else float(fn))
    if comp.get('air', 0.0) < 0.99:
        mua = mua + baseline_mua(nm)
    g = anisotropy(nm, g500)
    musp = mus_reduced(nm, a_prime, f_ray, b_mie)
    mus = musp / (1.0 - np.minimum(g, 0.99))
    return float(mua[0]), float(musp[0]), float(mus[0]), float(g[0]), float(n[0])


# ------------------------------------------------------------------ the reference's layer API
# Armoured beetle body of the 3-D runners (tissue_stack.py:23-41)
BASE_LAYERS = [
    {'name': 'epicuticle_wax', 'thickness_um': 2.0, 'composition': {'lipid': 0.85, 'protein': 0.15}},
    {'name': 'exocuticle_sclerotised', 'thickness_um': 60.0,
     'composition': {'chitin': 0.403, 'protein': 0.33, 'water': 0.18, 'melanin': 0.075, 'ommochrome': 0.012}},
    {'name': 'endocuticle', 'thickness_um': 100.0, 'composition': {'water': 0.42, 'chitin': 0.33, 'protein': 0.25}},
    {'name': 'subelytral_air_gap', 'thickness_um': 300.0, 'composition': {'air': 1.0}},
    {'name': 'epidermis', 'thickness_um': 20.0, 'composition': {'water': 0.70, 'protein': 0.22, 'lipid': 0.08}},
    {'name': 'fat_body', 'thickness_um': 300.0, 'composition': {'lipid': 0.50, 'water': 0.38, 'protein': 0.12}},
    {'name': 'hemolymph', 'thickness_um': 250.0, 'composition': {'water': 0.93, 'protein': 0.07}},
    {'name': 'muscle', 'thickness_um': 1581.0, 'composition': {'water': 0.76, 'protein': 0.21, 'lipid': 0.03}},
    {'name': 'fat_body', 'thickness_um': 300.0, 'composition': {'lipid': 0.50, 'water': 0.38, 'protein': 0.12}},
    {'name': 'epidermis', 'thickness_um': 20.0, 'composition': {'water': 0.70, 'protein': 0.22, 'lipid': 0.08}},
    {'name': 'endocuticle', 'thickness_um': 40.0, 'composition': {'water': 0.42, 'chitin': 0.33, 'protein': 0.25}},
    {'name': 'exocuticle_sclerotised', 'thickness_um': 25.0,
     'composition': {'chitin': 0.43, 'protein': 0.352, 'water': 0.18, 'melanin': 0.03, 'ommochrome': 0.008}},
    {'name': 'epicuticle_wax', 'thickness_um': 2.0, 'composition': {'lipid': 0.85, 'protein': 0.15}},
]

CHROMOPHORE_HEAT_CAPACITIES = {   # J/(mm^3 K)
    'water': 0.004184, 'chitin': 0.0021, 'protein': 0.0022, 'lipid': 0.0019, 'melanin': 0.0024,
    'ommochrome': 0.0022, 'plant_pigment': 0.0035, 'hemocyanin': 0.0038, 'aragonite': 0.00216,
    'conchiolin': 0.0022, 'mucin': 0.0039, 'hemoglobin': 0.0036, 'air': 0.0000012,
}

_ALIASES = {'matrix': 'chitin', 'chlorophyll': 'plant_pigment', 'conchiolin_matrix': 'conchiolin',
            'calcite': 'aragonite', 'blood': 'protein', 'oxyhemoglobin': 'protein'}

_LABELS = {
    'epicuticle_wax': 'Epicuticular Wax Film', 'exocuticle_sclerotised': 'Exocuticle (Sclerotised Armour)',
    'endocuticle': 'Endocuticle (Hydrated Chitin Matrix)',
    'subelytral_air_gap': 'Subelytral Air Cavity (Optothermal Barrier)', 'epidermis': 'Epidermal Epithelium',
    'fat_body': 'Visceral Fat Body', 'hemolymph': 'Hemolymph Vascular Sinus', 'muscle': 'Locomotor Muscle Mass',
}


def _fallback(comp, wl, a_prime, f_ray, g500, b_mie, refined_g):
    """The reference's closed-form branch when the basis cannot be evaluated
    (tissue_stack.py:152-171)."""
    w = float(wl)
    m = comp.get('melanin', 0.0)
    wf = comp.get('water', 0.0)
    mu_a = 0.1812 * (w / 450.0) ** -3.0 + m * 51.5 * (w / 500.0) ** -3.48 + wf * 0.00092
    musp = max(1e-4, a_prime * 0.1 * (f_ray * (w / 500.0) ** -4.0 + (1.0 - f_ray) * (w / 500.0) ** -b_mie))
    if refined_g:
        g = min(0.99, g500 * ((500.0 / w) ** 0.20 if w > 500.0 else (1.0 + 0.05 * (1.0 - w / 500.0))))
    else:
        g = min(0.99, g500)
    return mu_a, musp, musp / max(1.0 - g, 0.01), g, 1.33 * wf + 1.55 * (1.0 - wf)


def eval_single_layer(ld, wl):
    """Per-layer optics in mm^-1 at wavelength wl [nm] (tissue_stack.py:90-196)."""
    name = ld.get('name', 'custom_layer')
    thick_um = float(ld.get('thickness_um', 100.0))
    raw = ld.get('composition', {'water': 0.70, 'protein': 0.30})
    remapped = {}
    for k, v in raw.items():
        ck = _ALIASES.get(k, k)
        remapped[ck] = remapped.get(ck, 0.0) + float(v)
    valid = {k: max(0.0, float(v)) for k, v in remapped.items() if k in CHROMOPHORE_HEAT_CAPACITIES}
    tot = sum(valid.values())
    comp = {k: v / tot for k, v in valid.items()} if tot > 0 else {'water': 1.0}
    a_prime = float(ld.get('a_prime', 40.0))
    f_ray = float(ld.get('f_rayleigh', 0.20))
    b_mie = float(ld.get('b_mie', 1.1))
    g500 = float(ld.get('g500', 0.90))

    if comp.get('air', 0.0) >= 0.99 or name == 'subelytral_air_gap' or 'air' in name:
        mu_a, musp, mu_s, g, n = 0.0, 0.0, 0.0, 0.0, 1.00029
    else:
        try:
            mua_c, musp_c, mus_c, g, n = _layer_optics(comp, wl, a_prime, f_ray, b_mie, min(0.99, g500))
            mu_a, musp, mu_s = mua_c / 10.0, musp_c / 10.0, mus_c / 10.0
        except (_BasisUnavailable, ValueError):
            mu_a, musp, mu_s, g, n = _fallback(comp, wl, a_prime, f_ray, g500, b_mie, refined_g=True)

    rho_cp = sum(f * CHROMOPHORE_HEAT_CAPACITIES.get(c, 0.003) for c, f in comp.items())
    if mu_a > 1e-5 and musp > 1e-5:
        delta = 1.0 / math.sqrt(3.0 * mu_a * (mu_a + musp))
    elif mu_a > 1e-5:
        delta = 1.0 / mu_a
    else:
        delta = 999.0
    return {
        'name': name,
        'label': ld.get('label') or _LABELS.get(name, name.replace('_', ' ').title()),
        'thickness_um': thick_um,
        'thickness_mm': round(thick_um / 1000.0, 4),
        'composition': comp,
        'a_prime': a_prime,
        'f_rayleigh': f_ray,
        'b_mie': b_mie,
        'g500': g500,
        'mu_a': round(mu_a, 4),
        'mu_s': round(mu_s, 3),
        'mu_s_prime': round(musp, 3),
        'g': round(g, 3),
        'n': round(n, 3),
        'rho_cp': round(rho_cp, 6),
        'penetration_depth_mm': round(delta, 3),
        'is_air': comp.get('air', 0) > 0.90,
    }
