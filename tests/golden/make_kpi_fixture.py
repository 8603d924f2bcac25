"""Extract the reference's recorded campaign KPIs into a small fixture.

Reads ``knowledge/beetle/kpi_cache.json`` from the reference checkout: the KPIs
its own 200x200x184 beetle runs produced, reduced by compute_suite_metrics.py:
the 21 single-wavelength runs of run_highres_knowledge_suite.py
(``matrix_<wl>nm_<P>w.npz``), and the dual-wavelength and no-plume runs of
run_dual_permutation_suite.py (``matrix_dual_<wl1>nm_<wl2>nm.npz``,
``matrix_<wl>nm_40w_no_plume.npz``).  Run
once in the build container:

    python tests/golden/make_kpi_fixture.py /root/reference

The output ``kpi_highres_reference.json`` is committed; nothing at test time
reads the reference checkout.
"""
import json
import os
import re
import sys

KEYS = ("final_depth_um", "avg_depth_um", "speed_mm_s", "avg_hole_dia_mm", "core_avg_T", "core_max_T",
        "ven_avg_T", "ven_max_T", "v_abl_mm3", "v_boil_mm3", "energy_used_j", "min_trans")


def main(ref_root):
    with open(os.path.join(ref_root, "knowledge", "beetle", "kpi_cache.json")) as f:
        d = json.load(f)
    out = {}
    for name, rec in sorted(d.items()):
        k = rec["kpis"]
        m = re.fullmatch(r"matrix_(\d+)nm_(\d+)w\.npz", name)
        if m:   # run_highres_knowledge_suite.py runs
            out[f"{m.group(1)}nm_{m.group(2)}w"] = {"kind": "single", "wavelength_nm": int(m.group(1)),
                                                   "power_w": float(m.group(2)), **{key: k[key] for key in KEYS}}
            continue
        m = re.fullmatch(r"matrix_dual_(\d+)nm_(\d+)nm\.npz", name)
        if m:   # run_dual_permutation_suite.py, two wavelengths
            out[f"dual_{m.group(1)}nm_{m.group(2)}nm"] = {"kind": "dual", "wavelength_nm": int(m.group(1)),
                                                         "wavelength2_nm": int(m.group(2)),
                                                         **{key: k[key] for key in KEYS}}
            continue
        m = re.fullmatch(r"matrix_(\d+)nm_(\d+)w_no_plume\.npz", name)
        if m:   # run_dual_permutation_suite.py --no-plume
            out[f"{m.group(1)}nm_{m.group(2)}w_no_plume"] = {"kind": "no_plume", "wavelength_nm": int(m.group(1)),
                                                            "power_w": float(m.group(2)),
                                                            **{key: k[key] for key in KEYS}}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kpi_highres_reference.json")
    with open(dst, "w") as f:
        json.dump({"source": "knowledge/beetle/kpi_cache.json (reference checkout)", "runs": out}, f, indent=1,
                  sort_keys=True)
    print(f"{len(out)} runs -> {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
