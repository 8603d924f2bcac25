"""UI frames of the interactive driver from device-side plane slices.

``beetle_sim.run`` (beetle_sim.py:276-311) builds each animation frame from the x-z mid-plane:
strided T, Q, plume absorption and the four-phase map, plus the core / ventral temperature
figures, the crater depth and the ablated / boiled shares of the tissue.  Done on host arrays
that means pulling six full grids per frame (T, Q, chi, dry, ablated, H_acc) and evaluating the
phase map over the whole volume; here only x-z planes leave the device: stride-1 planes of T,
Q, the plume extinction, the ablated mask and the phase map evaluated on the device
(``bd_download_slice``, ``bd_phase_slice``), which is a few hundred kilobytes per frame.  The
figures are then the reference's own expressions on those planes, so a frame equals the one the
reference computes from the same state.
"""

from __future__ import annotations

import base64

import numpy as np


def _encode(arr, dtype):
    """beetle_sim.py:125-126."""
    return base64.b64encode(np.ascontiguousarray(arr, dtype=dtype).tobytes()).decode('ascii')


def xz_planes(engine, iy, is_air_gap=None):
    """The x-z plane at y index iy of T, Q, plume mu_a, ablated and the phase map, as (Nx, Nz) host
    arrays from device slices.  is_air_gap: None, a per-z profile (Nz,) or a 3-D mask that is
    constant over x and y (the shape the beetle builder produces), marking air-gap voxels."""
    ctx = engine._ctx
    gap_z = None
    if is_air_gap is not None:
        g = np.asarray(is_air_gap, dtype=bool)
        if g.ndim == 3:
            if not np.array_equal(g, np.broadcast_to(g[:1, :1, :], g.shape)):
                raise ValueError("xz_planes: the air-gap mask must be constant over x and y (a per-z profile)")
            g = g[0, 0, :]
        if g.shape != (engine.Nz,):
            raise ValueError("xz_planes: the air-gap profile needs one value per z")
        gap_z = g.astype(np.uint8)
    out = {'T': ctx.slice('T', 1, iy), 'Q': ctx.slice('Q', 1, iy), 'ablated': ctx.slice('ABLATED', 1, iy) != 0,
           'phase': ctx.phase_slice(1, iy, gap_z)}
    if engine.plume is not None:
        pl = engine.plume
        out['plume'] = ctx.slice('PLUME_EXT', 1, iy) * (pl.mu_a_max / max(1e-6, pl.mu_max))
    else:
        out['plume'] = np.zeros_like(out['T'])
    return out


def xz_frame(engine, iy, sx, sz, is_air_gap=None, core=None, ventral=None, nz_air=0, tissue=None,
             t_ambient=20.0):
    """The per-frame panels and figures of beetle_sim.py:276-311 on the device-sliced plane iy:
    base64 '<f4' panels of max(Q, 1e-4), T and plume mu_a and the 'u1' phase panel (each strided
    by sx in x and sz in z and transposed, as the UI expects), the core and ventral temperature
    peak / mean over the non-ablated cells of the z ranges ``core`` and ``ventral`` (full-resolution
    plane), the crater depth [um] below ``nz_air``, and the ablated / boiled shares of the strided
    ``tissue`` mask (strided x-z, True on tissue).  Time, progress and energy fields belong to the
    caller's loop and are not included."""
    p = xz_planes(engine, iy, is_air_gap)
    T = p['T'][::sx, ::sz]
    Q = p['Q'][::sx, ::sz]
    phase = p['phase'][::sx, ::sz]
    plume = p['plume'][::sx, ::sz]
    data = {'Q': _encode(np.maximum(Q, 1e-4).T, '<f4'), 'T': _encode(T.T, '<f4'), 'plume': _encode(plume.T, '<f4'),
            'phase': _encode(phase.T, 'u1')}
    Tf, abl = p['T'].astype(np.float64), p['ablated']
    for name, rng in (('core', core), ('ventral', ventral)):
        if rng is None:
            continue
        c0, c1 = rng
        cells = Tf[:, c0:c1][~abl[:, c0:c1]]
        data[f'{name}_peak'] = round(float(cells.max()) if cells.size else t_ambient, 2)
        data[f'{name}_mean'] = round(float(cells.mean()) if cells.size else t_ambient, 2)
    a = abl[:, nz_air:]
    d_um = engine.dz * 1000.0
    data['depth_um'] = round((int(np.where(a)[1].max()) + 1) * d_um, 1) if np.any(a) else 0.0
    if tissue is not None:
        n_t = max(1, int(np.sum(tissue)))
        data['ablated_pct'] = round(100.0 * float(np.sum((phase == 3) & tissue)) / n_t, 2)
        data['boiled_pct'] = round(100.0 * float(np.sum((phase == 2) & tissue)) / n_t, 2)
    return data
