"""The C-ABI library loads and exports every entry point include/*.h declares (CPU only)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for h in sorted(os.listdir(inc)):
        if h.endswith(".h"):
            src = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, h)).read(), flags=re.S)
            names |= set(re.findall(r"\b(bd_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from simulation import _bdlib as B
    lib = B.lib()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    # and the Python binding covers the same set
    assert set(names) == set(B.EXPORTS)


def test_struct_sizes_match_header():
    """The ctypes mirrors must have the C layout (a mismatch corrupts every call)."""
    from simulation import _bdlib as B
    assert ctypes.sizeof(B.Grid) == 3 * 4 + 4 + 3 * 8  # int32 x3, pad, double x3
    assert ctypes.sizeof(B.MCResult) == 10 * 8
    assert ctypes.sizeof(B.EpochDiag) == 8 + 25 * 8
    assert ctypes.sizeof(B.FnoSpec) == 4 * 4


def test_no_device_raises_loudly():
    """Without a HIP device the product refuses to run (no silent CPU fallback)."""
    from simulation import _bdlib as B
    if B.device_count() > 0:
        pytest.skip("a HIP device is visible")
    from simulation.voxel_mc_3d import run_voxel_mc3d
    with pytest.raises(RuntimeError):
        run_voxel_mc3d(n_photons=100, Nx=4, Ny=4, Nz=4)
    with pytest.raises(RuntimeError):
        B.Context((4, 4, 4), (0.1, 0.1, 0.1))


def test_atomic_probe_rejects_bad_arguments():
    """Argument checks run before any device call, so they hold without a GPU."""
    from simulation import _bdlib as B
    ms = ctypes.c_double(0.0)
    for args in ((0, 100, 0, 4096), (0, 100, 101, 4096), (0, 100, 100, 4095), (0, 0, 0, 4096),
                 (0, 100, 100, 0)):
        rc = B.lib().bd_probe_atomic_rate(*args, ctypes.byref(ms))
        assert rc == B.BD_EINVAL
    with pytest.raises(ValueError):
        B.probe_atomic_rate(100, span=200)


def test_cpu_engine_keyword_rejected():
    from simulation.voxel_mc_3d import check_engine
    with pytest.raises(ValueError):
        check_engine('cpu')
    with pytest.raises(ValueError):
        check_engine('metal')
