"""Extract the reference's recorded staged-ablation series into a small fixture.

Reads ``knowledge/simulation_lit_comparison/raw_output/sagi_stages.json`` from
the reference checkout: the per-epoch surface temperature, crater depth and
boiling/dry layer counts its coupled engine produced on the quasi-1-D case of
``knowledge/simulation_lit_comparison/code/sagi_stages.py`` (600 epochs of
2 ms), and keeps every epoch.  Run once in the build container:

    python tests/golden/make_sagi_fixture.py /root/reference

The output ``sagi_stages_reference.json`` is committed; nothing at test time
reads the reference checkout.
"""
import json
import os
import sys


def main(ref_root):
    src = os.path.join(ref_root, "knowledge", "simulation_lit_comparison", "raw_output", "sagi_stages.json")
    with open(src) as f:
        d = json.load(f)
    keep = {k: d[k] for k in ("t_s", "T_surface", "depth_mm", "n_boiling_layers", "n_dry_layers", "I_W_mm2",
                              "Ta", "T_boil", "dz")}
    keep["source"] = "knowledge/simulation_lit_comparison/raw_output/sagi_stages.json (reference checkout)"
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sagi_stages_reference.json")
    with open(dst, "w") as f:
        json.dump(keep, f)
    print(f"{len(keep['t_s'])} epochs -> {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
