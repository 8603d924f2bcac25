"""Writes tests/golden/shimojo_reference.json from the reference's recorded
scanned-beam liver runs (knowledge/simulation_lit_comparison/raw_output/):

- ``shimojo_cross_sections.json``: the x-z ablated and coagulated masks at
  y = Ly / 2 of the six cases (validation_benchmarks.py:228-233), stored here as
  np.packbits hex with their shapes;
- ``benchmark_results.json``: the metrics the reference reported for them
  (results ``S-<P>W-<v>-<metric>``; the 10 W cases report only the coagulation
  metrics).

    python tests/golden/make_shimojo_fixture.py
"""

import json
import pathlib

import numpy as np

RAW = pathlib.Path('/root/reference/knowledge/simulation_lit_comparison/raw_output')
OUT = pathlib.Path(__file__).with_name('shimojo_reference.json')


def pack(mask):
    a = np.asarray(mask, dtype=bool)
    return {'shape': list(a.shape), 'hex': np.packbits(a.ravel()).tobytes().hex()}


def main():
    secs = json.load(open(RAW / 'shimojo_cross_sections.json'))
    results = json.load(open(RAW / 'benchmark_results.json'))['results']
    reported = {r['id']: r['model'] for r in results if str(r.get('id', '')).startswith('S-')}
    out = {}
    for key, sec in secs.items():
        p, v = key.split('W_')
        metrics = {}
        for m in ('vap_depth', 'vap_width', 'coag_depth', 'coag_width'):
            rid = f'S-{p}W-{v}-{m}'
            if rid in reported:
                metrics[m] = float(reported[rid])
        out[key] = {'power_W': float(p), 'speed_mm_s': float(v), 'Lx': sec['Lx'], 'Lz': sec['Lz'],
                    'reported': metrics, 'ablated': pack(sec['ablated']), 'coag': pack(sec['coag'])}
    OUT.write_text(json.dumps(out, indent=1) + '\n')
    print(f'wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} cases)')


if __name__ == '__main__':
    main()
