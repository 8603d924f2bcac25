"""Extract the reference's own recorded verification outputs into a small fixture.

Reads ``knowledge/simulation_lit_comparison/raw_output/benchmark_results.json``
from the reference checkout (data the reference itself committed: the values its
engine produced for the verification tier) and keeps only the entries the oracle
tests pin against.  Run once in the build container:

    python tests/golden/make_reference_fixtures.py /root/reference

The output ``reference_engine_values.json`` is committed; nothing at test time
reads the reference checkout.
"""
import json
import os
import sys

KEEP = {
    # deterministic (no Monte Carlo inside): pinned to round-off
    "C1-solid", "C1-sandwich", "C2-dark-tau", "C2-dark-end", "C2-shiny-tau", "C2-shiny-end",
    "C4-iso", "C4-ramp",
    # Monte Carlo: statistical
    "V1a", "V1b", "V1c", "V1d", "V1e", "V1f", "V2", "V2b",
    "C3-R-0", "C3-R-15", "C3-R-30", "C3-R-45", "C3-R-60",
    "C3-z-0", "C3-z-15", "C3-z-30", "C3-z-45", "C3-z-60",
    "V3a", "V3b", "V4-2ms", "V4-10ms",
    "V2c-a0.5", "V2c-a0.7", "V2c-a0.9", "V2c-a0.95", "V2c-a0.99",
    "V2d-a0.5", "V2d-a0.7", "V2d-a0.9", "V2d-a0.99",
    # coupled engine: crater depth of the McMillan et al. pixel runs (mcmillan_benchmark.py:60-99)
    "M-50-exp", "M-100-exp", "M-150-exp", "M-200-exp", "M-250-exp", "M-300-exp", "M-350-exp", "M-400-exp",
}


def main(ref_root):
    src = os.path.join(ref_root, "knowledge", "simulation_lit_comparison", "raw_output", "benchmark_results.json")
    with open(src) as f:
        d = json.load(f)
    out = {}
    for r in d["results"]:
        if r["id"] in KEEP:
            rec = {k: r.get(k) for k in ("model", "reference_value", "tolerance_abs", "tolerance_basis", "pass")}
            s = r.get("series") or {}
            if r["id"] in ("C2-dark-end", "C2-shiny-end"):
                rec["series"] = {"t": s["t"], "dT": s["dT"]}
            if r["id"].startswith("V2d-"):
                rec["series"] = {"se_ref": s["se_ref"]}
            if r["id"] in ("V4-2ms", "V4-10ms"):
                rec["series"] = {"t_s": s["t_s"], "depth_mm": s["depth_mm"]}
            out[r["id"]] = rec
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_engine_values.json")
    with open(dst, "w") as f:
        json.dump({"source": "knowledge/simulation_lit_comparison/raw_output/benchmark_results.json",
                   "values": out}, f, indent=1)
    print(f"wrote {len(out)} entries to {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
