"""Writes tests/golden/keller_reference.json from the reference's recorded Keller
et al. (2016) thresholds (knowledge/simulation_lit_comparison/raw_output/
benchmark_results.json, results KE-*): the fluence (J/cm^2) at which the surface
Arrhenius integral of a water column first reaches 1, per kinetics set and pulse
length.  The beam diameter does not enter the 1-D computation, so the 2.5 mm and
5.8 mm rows of one pulse length carry the same value; that is checked here.

    python tests/golden/make_keller_fixture.py
"""

import json
import pathlib

RAW = pathlib.Path('/root/reference/knowledge/simulation_lit_comparison/raw_output')
OUT = pathlib.Path(__file__).with_name('keller_reference.json')


def main():
    out = {}
    for r in json.load(open(RAW / 'benchmark_results.json'))['results']:
        if not r['id'].startswith('KE-'):
            continue
        s = r['series']
        key = f"{s['kinetics'].split()[0]}|{s['t_pulse_ms']}"
        assert out.get(key, r['model']) == r['model'], key
        out[key] = r['model']
    OUT.write_text(json.dumps(out, indent=1) + '\n')
    print(f'wrote {OUT}: {out}')


if __name__ == '__main__':
    main()
