"""The drop-in package keeps the reference's callers importable (CPU, no device needed).

A synthetic reference checkout in a temporary directory holds a caller written like
beetle_sim.py:16-18 (it imports the engine, the matrix runner's constants and phase map, and the
layer optics), a stand-in matrix runner that imports only the engine and tissue_stack, as the
reference's does (run_multispectral_matrix.py:27-29), and decoy copies of modules this package
restates.  With BUGDEF_REFERENCE_SIMULATION naming that directory, the caller must get this
package's engine and the checkout's runner, and never a decoy.
"""
import json
import os
import subprocess
import sys

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bug-defense_amd")

_PROBE = r"""
import json, sys
import simulation
import simulation.beetle_caller as caller
import simulation.coupled_ablation as ca
import simulation.run_multispectral_matrix as rmm
import simulation.tissue_stack as ts
import simulation.metal_ablation_solver as mas
import simulation.voxel_thermal_metal as vtm
import simulation.voxel_mc_metal as vmm
print(json.dumps({
    'engine': ca.__file__, 'runner': rmm.__file__, 'tissue': ts.__file__, 'metal': mas.__file__,
    'thermal_metal': vtm.__file__, 'mc_metal': vmm.__file__,
    'same_engine': caller.CoupledAblationEngine is ca.CoupledAblationEngine,
    'runner_origin': rmm.ORIGIN, 'char_state': caller.CHAR_STATE,
    'layers': len(rmm.BASE_LAYERS) > 0,
}))
"""


def _run(code, env_extra):
    env = dict(os.environ)
    env.pop("BUGDEF_REFERENCE_SIMULATION", None)
    env.update(env_extra)
    env["PYTHONPATH"] = PKG
    return subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)


def _checkout(tmp_path):
    sim = tmp_path / "checkout" / "simulation"
    sim.mkdir(parents=True)
    (sim / "__init__.py").write_text('raise RuntimeError("the checkout package init ran")\n')
    for decoy in ("coupled_ablation", "tissue_stack", "voxel_mc_3d", "metal_ablation_solver",
                  "voxel_thermal_metal", "voxel_mc_metal", "plume"):
        (sim / f"{decoy}.py").write_text(f'raise RuntimeError("decoy {decoy} loaded from the checkout")\n')
    (sim / "run_multispectral_matrix.py").write_text(
        "from simulation.coupled_ablation import CoupledAblationEngine, K_AIR, RHO_CP_AIR\n"
        "from simulation.tissue_stack import eval_single_layer\n"
        "from simulation.tissue_stack import BASE_LAYERS\n"
        "ORIGIN = 'checkout'\n"
        "H_LAT_WATER = 2.26\n"
        "CHAR_STATE = {'A': 1.0e7, 'Ea': 65000.0}\n"
        "def compute_phase_map(T, chi, dry, ablated, H_acc, is_air_gap=None):\n"
        "    return None\n")
    (sim / "beetle_caller.py").write_text(
        "from simulation.coupled_ablation import CoupledAblationEngine\n"
        "from simulation.run_multispectral_matrix import CHAR_STATE, H_LAT_WATER, compute_phase_map\n"
        "from simulation.tissue_stack import eval_single_layer\n")
    return sim


def test_reference_callers_resolve_through_the_package_path(tmp_path):
    sim = _checkout(tmp_path)
    r = _run(_PROBE, {"BUGDEF_REFERENCE_SIMULATION": str(sim)})
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    for key in ('engine', 'tissue', 'metal', 'thermal_metal', 'mc_metal'):
        assert os.path.dirname(out[key]) == os.path.join(PKG, "simulation"), (key, out[key])
    assert os.path.dirname(out['runner']) == str(sim)
    assert out['same_engine'] and out['runner_origin'] == 'checkout' and out['layers']
    assert out['char_state'] == {'A': 1.0e7, 'Ea': 65000.0}


def test_without_the_checkout_the_runner_is_absent(tmp_path):
    r = _run("import simulation.run_multispectral_matrix", {})
    assert r.returncode != 0 and "ModuleNotFoundError" in r.stderr
    # a path that does not exist is ignored
    r = _run("import simulation; print(len(simulation.__path__))",
             {"BUGDEF_REFERENCE_SIMULATION": str(tmp_path / "missing")})
    assert r.returncode == 0 and r.stdout.strip() == "1"
