"""Writes tests/golden/kim_reference.json from the reference's recorded Kim et al.
(2023) runs (knowledge/simulation_lit_comparison/raw_output/): the exp(-2 r/R)
source from benchmark_results.json (results K-PS*-depth / K-PS*-area and their
series' fitted exposure t_fit_s; PS1/PS2 form the 55 C group, PS3/PS4 the 200 C
group) and the exp(-2 (r/R)^2) source from kim_alt_profile.json.

    python tests/golden/make_kim_fixture.py
"""

import json
import pathlib

RAW = pathlib.Path('/root/reference/knowledge/simulation_lit_comparison/raw_output')
OUT = pathlib.Path(__file__).with_name('kim_reference.json')
GROUP_OF = {'PS1': 'coag', 'PS3': 'abl'}


def main():
    res = {r['id']: r for r in json.load(open(RAW / 'benchmark_results.json'))['results']}
    alt = json.load(open(RAW / 'kim_alt_profile.json'))
    out = {'kim_exp_disk': {}, 'kim_gauss_disk': {}}
    for ps, grp in GROUP_OF.items():
        d, a = res[f'K-{ps}-depth'], res[f'K-{ps}-area']
        assert d['series']['t_fit_s'] == a['series']['t_fit_s']
        out['kim_exp_disk'][grp] = {'t_fit_s': d['series']['t_fit_s'], 'depth': d['model'], 'area': a['model']}
        out['kim_gauss_disk'][grp] = {k: alt[ps][k] for k in ('t_fit_s', 'depth', 'area')}
    OUT.write_text(json.dumps(out, indent=1) + '\n')
    print(f'wrote {OUT}')


if __name__ == '__main__':
    main()
