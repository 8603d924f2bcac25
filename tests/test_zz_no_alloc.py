"""bd_mc_run_noalloc (the allocation-free transport) against bd_mc_run, through the C-ABI.

Both trace the same Philox streams on the same optics; only the order of the float atomic
additions into the tallies differs.  Two calls in a row, the second on a smaller grid with fewer
photons, show that the shared scratch starts each call cleared.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _optics(shape, seed):
    rng = np.random.default_rng(seed)
    Nz = shape[2]
    opt = {k: np.zeros(shape) for k in ('mu_a', 'mu_s', 'g', 'n')}
    cuts = sorted(rng.choice(np.arange(2, Nz - 2), size=2, replace=False))
    for a, b in ((0, cuts[0]), (cuts[0], cuts[1]), (cuts[1], Nz)):
        opt['mu_a'][:, :, a:b] = rng.uniform(0.2, 3.0)
        opt['mu_s'][:, :, a:b] = rng.uniform(2.0, 30.0)
        opt['g'][:, :, a:b] = rng.uniform(0.7, 0.95)
        opt['n'][:, :, a:b] = rng.uniform(1.3, 1.5)
    return opt


def _compare(shape, spacing, n_photons, seed, theta):
    from simulation import _bdlib as B
    from simulation.voxel_mc_3d import make_beam
    L = tuple(n * d for n, d in zip(shape, spacing))
    opt = _optics(shape, seed)
    beam = make_beam(10.0, n_photons, 0.3, theta, 30.0, (L[0] / 2, L[1] / 2, 0.0), 0.0, 0, 'gaussian', 1000 + seed)
    nvox = int(np.prod(shape))
    nd = shape[2]
    # the existing entry point, on a context
    ctx = B.Context(shape, spacing)
    for k in ('mu_a', 'mu_s', 'g', 'n'):
        ctx.upload(k.upper(), opt[k])
    r_ctx = ctx.mc_run(beam)
    A_ctx = ctx.download('A', dtype=np.float32)
    L_ctx = ctx.download('L', dtype=np.float32)
    lat_ctx = ctx.mc_lateral(nd)
    ctx.close()
    # the allocation-free one, on caller-owned device buffers
    packed = np.stack([opt[k].astype(np.float32).ravel() for k in ('mu_a', 'mu_s', 'g', 'n')], axis=1)
    d_opt = B.DeviceArray((nvox, 4)).upload(packed)
    d_al = B.DeviceArray((nvox, 2))
    d_al.fill(123.0)  # stale contents: the call must clear the tallies itself
    grid = B.Grid(*shape, *spacing)
    res = B.MCResult()
    lat = np.full(nd, -7.0)
    rc = B.lib().bd_mc_run_noalloc(ctypes.byref(grid), ctypes.byref(beam), d_opt.ptr, d_al.ptr, None,
                                   ctypes.byref(res), lat.ctypes.data_as(ctypes.c_void_p), nd)
    B.check(rc, "bd_mc_run_noalloc")
    al = d_al.download().reshape(shape + (2,))
    d_opt.free()
    d_al.free()
    # the allocation-free transport adds every path length inline; bd_mc_run defers most of them
    # to its record pass, so the tally additions match in total
    assert res.events == r_ctx.events and res.tally_records == 0.0
    assert res.tally_atomics == r_ctx.tally_atomics + r_ctx.tally_records
    assert res.events > 10 * n_photons
    for key in ('specular', 'diffuse', 'transmitted', 'lateral', 'absorbed'):
        assert getattr(res, key) == pytest.approx(getattr(r_ctx, key), rel=1e-12, abs=1e-12), key
    assert res.kernel_ms == -1.0
    np.testing.assert_allclose(lat, lat_ctx, rtol=1e-12, atol=1e-12)
    for got, ref in ((al[..., 0], A_ctx), (al[..., 1], L_ctx)):
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-6 * float(ref.max()))
        assert float(got.sum(dtype=np.float64)) == pytest.approx(float(ref.sum(dtype=np.float64)), rel=1e-5)
    return res


def test_noalloc_transport_matches_bd_mc_run_on_two_calls():
    r1 = _compare((40, 36, 30), (0.02, 0.02, 0.015), 60_000, 1, 20.0)
    r2 = _compare((24, 20, 16), (0.03, 0.03, 0.02), 20_000, 2, 0.0)   # smaller grid, fewer photons
    assert r2.events < r1.events


def test_noalloc_transport_rejects_a_depth_axis_beyond_its_scratch():
    from simulation import _bdlib as B
    from simulation.voxel_mc_3d import make_beam
    shape = (1, 1, 5000)
    d_opt = B.DeviceArray((5000, 4))
    d_al = B.DeviceArray((5000, 2))
    beam = make_beam(1.0, 100, 0.01, 0.0, 0.0, (0.005, 0.005, 0.0), 0.0, 0, 'gaussian', 3)
    grid = B.Grid(*shape, 0.01, 0.01, 0.01)
    res = B.MCResult()
    rc = B.lib().bd_mc_run_noalloc(ctypes.byref(grid), ctypes.byref(beam), d_opt.ptr, d_al.ptr, None,
                                   ctypes.byref(res), None, 0)
    assert rc == B.BD_EINVAL
    assert "4096" in B.lib().bd_last_error().decode()
    d_opt.free()
    d_al.free()
    # the next valid call is unaffected
    _compare((12, 10, 8), (0.05, 0.05, 0.05), 5_000, 3, 10.0)
