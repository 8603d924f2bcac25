"""The device frame path against host slicing of the same engine state (beetle_sim.py:276-311).

The reference builds each frame from full host grids: T, Q and the plume absorption sliced at
the mid-plane, the phase map evaluated over the whole volume and sliced, core / ventral
figures from the full-resolution plane, the crater depth and the ablated / boiled shares.
``simulation.frames.xz_frame`` builds the same frame from device plane slices; the panels must
be byte-identical and the figures equal.
"""
import base64

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host_frame(engine, oracle, mid_y, sx, sz, is_air, core, ventral, nz_air, tissue):
    """beetle_sim.py:282-305 on host copies of the full grids."""
    enc = lambda a, dt: base64.b64encode(np.ascontiguousarray(a, dtype=dt).tobytes()).decode('ascii')
    T = np.asarray(engine.T)[::sx, mid_y, ::sz]
    Q = np.asarray(engine.Q)[::sx, mid_y, ::sz]
    phase = oracle.compute_phase_map(np.asarray(engine.T), np.asarray(engine.chi), np.asarray(engine.dry),
                                     np.asarray(engine.ablated), np.asarray(engine.H_acc),
                                     is_air_gap=is_air)[::sx, mid_y, ::sz]
    plume = engine.plume.mu_a_plume[::sx, mid_y, ::sz]
    c0, c1 = core
    v0, v1 = ventral
    Tf, abl = np.asarray(engine.T), np.asarray(engine.ablated)
    cc = Tf[:, mid_y, c0:c1][~abl[:, mid_y, c0:c1]]
    vv = Tf[:, mid_y, v0:v1][~abl[:, mid_y, v0:v1]]
    a = abl[:, mid_y, nz_air:]
    n_t = max(1, int(np.sum(tissue)))
    return {'Q': enc(np.maximum(Q, 1e-4).T, '<f4'), 'T': enc(T.T, '<f4'), 'plume': enc(plume.T, '<f4'),
            'phase': enc(phase.T, 'u1'),
            'core_peak': round(float(cc.max()) if cc.size else 20.0, 2),
            'core_mean': round(float(cc.mean()) if cc.size else 20.0, 2),
            'ventral_peak': round(float(vv.max()) if vv.size else 20.0, 2),
            'ventral_mean': round(float(vv.mean()) if vv.size else 20.0, 2),
            'depth_um': round((int(np.where(a)[1].max()) + 1) * engine.dz * 1000.0, 1) if np.any(a) else 0.0,
            'ablated_pct': round(100.0 * float(np.sum((phase == 3) & tissue)) / n_t, 2),
            'boiled_pct': round(100.0 * float(np.sum((phase == 2) & tissue)) / n_t, 2)}


@pytest.mark.parametrize("strides", [(1, 1), (2, 3)])
def test_device_frames_equal_host_frames(oracle, strides):
    """Eight epochs of a layered stack (air on top, a cuticle that chars and ablates, an air gap the
    crater breaches, boiling wet layers, a plume on the thermal grid with the Metal solver's
    stable step); a frame after each of the last four epochs, device path against host path."""
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.frames import xz_frame
    from simulation.workloads import layered_stack
    kw = layered_stack(shape=(40, 36, 56), n_photons=30000, plume_grid=None, seed=3, power=40.0)
    kw['plume_closure'] = dict(kw['plume_closure'], dt_clamp=True)
    eng = CoupledAblationEngine(**kw)
    Nx, Ny, Nz = 40, 36, 56
    k0, k1 = kw['air_gap_range']
    is_air = np.zeros((Nx, Ny, Nz), bool)
    is_air[:, :, :21] = True
    is_air[:, :, k0:k1] = True
    sx, sz = strides
    mid_y = Ny // 2
    tissue = ~is_air[::sx, mid_y, ::sz]
    core, ventral = (k1, k1 + 10), (k1 + 10, Nz)
    for ep in range(8):
        eng.step_epoch()
        if ep < 4:
            continue
        dev = xz_frame(eng, mid_y, sx, sz, is_air_gap=is_air, core=core, ventral=ventral, nz_air=21, tissue=tissue)
        host = _host_frame(eng, oracle, mid_y, sx, sz, is_air, core, ventral, 21, tissue)
        assert set(dev) == set(host)
        for key in host:
            assert dev[key] == host[key], (ep, key)
    assert host['depth_um'] > 0.0 and host['boiled_pct'] > 0.0 and host['ablated_pct'] > 0.0
    assert np.frombuffer(base64.b64decode(host['plume']), '<f4').max() > 0.0
    eng.close()


def test_frame_rejects_an_air_gap_mask_that_varies_in_x_or_y():
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.frames import xz_planes
    from simulation.workloads import layered_stack
    eng = CoupledAblationEngine(**layered_stack(shape=(16, 16, 40), n_photons=1000, plume_grid=None, seed=1))
    bad = np.zeros((16, 16, 40), bool)
    bad[3, :, 25] = True
    with pytest.raises(ValueError):
        xz_planes(eng, 8, bad)
    eng.close()
