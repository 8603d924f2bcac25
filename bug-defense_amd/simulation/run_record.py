"""Per-epoch run record and its KPIs, fed from device-side slices.

``RunRecorder`` records what the reference's high-resolution runner stores
each epoch (run_highres_knowledge_suite.py:240-329) and writes the same
``.npz`` keys (:335-367):
- mid-plane x-z slices of T, Q, tissue phase and plume mu_a;
- the centre line along the beam trajectory;
- crater depth;
- dorsal/core/ventral peak temperatures;
- perforation and core-boil times.

The reference pulls six full 3-D fields to the host every epoch to do this.
Here each epoch moves only x-z planes: the phase map and the slices are
evaluated on the device (``bd_download_slice``, ``bd_phase_slice``).

``analyze_npz`` restates compute_suite_metrics.py:28-118, the KPI extraction
the reference runs on those files.
"""

from __future__ import annotations

import numpy as np

from . import _bdlib as B

# compute_suite_metrics.py:19-27
SPECTRO_MAP = {
    320: {'R': 6.23, 'A': 92.96, 'T': 0.00, 'Lat': 0.81},
    450: {'R': 13.54, 'A': 77.92, 'T': 0.13, 'Lat': 8.41},
    532: {'R': 19.61, 'A': 61.33, 'T': 0.60, 'Lat': 18.46},
    1064: {'R': 23.68, 'A': 11.51, 'T': 4.58, 'Lat': 60.23},
    1470: {'R': 9.63, 'A': 86.86, 'T': 0.11, 'Lat': 3.40},
    1940: {'R': 5.00, 'A': 94.86, 'T': 0.00, 'Lat': 0.13},
    10600: {'R': 4.28, 'A': 95.72, 'T': 0.00, 'Lat': 0.00},
}


class RunRecorder:
    """Records one run of a CoupledAblationEngine, epoch by epoch.

    ``nz_air_top``: slices of ambient air above the dorsal surface;
    ``air_gap``: (start, end) slice range of the subelytral gap;
    ``regions``: {'dorsal': (k0, k1), 'core': (k0, k1), 'ventral': (k0, k1)} slice ranges for
    the peak temperatures; ``ix_beam``: x index of the beam axis at each z (the reference's
    tilted trajectory, :237-238).
    """

    def __init__(self, engine, nz_air_top, air_gap, regions, ix_beam, mid_y=None):
        self.eng = engine
        self.nz_air_top = int(nz_air_top)
        self.air_gap = tuple(int(v) for v in air_gap) if air_gap is not None else None
        self.regions = {k: tuple(int(x) for x in v) for k, v in regions.items()}
        self.ix_beam = np.asarray(ix_beam, dtype=np.int64)
        self.mid_y = engine.Ny // 2 if mid_y is None else int(mid_y)
        gap = np.zeros(engine.Nz, dtype=np.uint8)
        gap[:self.nz_air_top] = 1
        if self.air_gap is not None:
            gap[self.air_gap[0]:self.air_gap[1]] = 1
        self.gap_z = gap
        self.rows = {k: [] for k in ('times', 'xz_T', 'xz_Q', 'xz_phase', 'xz_plume', 'centerline_T',
                                     'centerline_Q', 'centerline_phase', 'crater_depths_mm', 'plume_trans',
                                     'peak_T_dorsal', 'peak_T_core', 'peak_T_ventral')}
        self.t_perforation_s = None
        self.t_core_boil_s = None

    def record(self, info: dict):
        e = self.eng
        ctx = e._ctx
        j = self.mid_y
        T = ctx.slice('T', 1, j)
        Q = ctx.slice('Q', 1, j)
        ph = ctx.phase_slice(1, j, self.gap_z)
        abl = ctx.slice('ABLATED', 1, j) > 0.5
        plume = (ctx.slice('PLUME_EXT', 1, j) * (e.plume.mu_a_max / max(1e-6, e.plume.mu_max))
                 if e.plume is not None else np.zeros_like(T))
        k = np.arange(e.Nz)
        r = self.rows
        r['times'].append(e.t)
        r['xz_T'].append(T)
        r['xz_Q'].append(Q)
        r['xz_phase'].append(ph)
        r['xz_plume'].append(plume.astype(np.float32))
        r['centerline_T'].append(T[self.ix_beam, k])
        r['centerline_Q'].append(Q[self.ix_beam, k])
        r['centerline_phase'].append(ph[self.ix_beam, k])
        below = abl[:, self.nz_air_top:]
        depth = (int(np.where(below)[1].max()) + 1) * e.dz if np.any(below) else 0.0
        r['crater_depths_mm'].append(depth)
        r['plume_trans'].append(float(info.get('plume_transmission', 1.0)))
        for name, key in (('dorsal', 'peak_T_dorsal'), ('core', 'peak_T_core'), ('ventral', 'peak_T_ventral')):
            k0, k1 = self.regions[name]
            r[key].append(float(T[:, k0:k1].max()))
        if self.t_perforation_s is None and self.air_gap is not None and np.any(abl[:, self.air_gap[0] - 1]):
            self.t_perforation_s = e.t
        if self.t_core_boil_s is None and r['peak_T_core'][-1] >= 99.5:
            self.t_core_boil_s = e.t

    def arrays(self) -> dict:
        r = self.rows
        out = {
            'times': np.array(r['times']),
            'xz_T': np.array(r['xz_T'], dtype=np.float32), 'xz_Q': np.array(r['xz_Q'], dtype=np.float32),
            'xz_phase': np.array(r['xz_phase'], dtype=np.uint8),
            'xz_plume': np.array(r['xz_plume'], dtype=np.float32),
            'centerline_T': np.array(r['centerline_T'], dtype=np.float32),
            'centerline_Q': np.array(r['centerline_Q'], dtype=np.float32),
            'centerline_phase': np.array(r['centerline_phase'], dtype=np.uint8),
            'crater_depths_mm': np.array(r['crater_depths_mm']), 'plume_trans': np.array(r['plume_trans']),
            'peak_T_dorsal': np.array(r['peak_T_dorsal']), 'peak_T_core': np.array(r['peak_T_core']),
            'peak_T_ventral': np.array(r['peak_T_ventral']),
            't_perforation_s': self.t_perforation_s, 't_core_boil_s': self.t_core_boil_s,
        }
        return out

    def save(self, path: str, **meta):
        """np.savez_compressed with the keys of run_highres_knowledge_suite.py:335-367."""
        e = self.eng
        a = self.arrays()
        a.update(grid_x=np.linspace(0, e.Lx, e.Nx), grid_y=np.linspace(0, e.Ly, e.Ny),
                 grid_z=np.linspace(-self.nz_air_top * e.dz, (e.Nz - self.nz_air_top) * e.dz, e.Nz),
                 NX=e.Nx, NY=e.Ny, NZ=e.Nz, DX=e.dx, DY=e.dy, DZ=e.dz, dt_epoch=e.dt_epoch,
                 theta_inc_deg=e.theta_inc_deg, phi_inc_deg=e.phi_inc_deg)
        a.update(meta)
        for key in ('t_perforation_s', 't_core_boil_s'):
            if a[key] is None:
                a[key] = np.array(None, dtype=object)
        np.savez_compressed(path, **{k: v for k, v in a.items() if not (isinstance(v, np.ndarray)
                                                                     and v.dtype == object)})


def analyze_npz(npz_path: str, pulse_duration_s: float = 0.200) -> dict:
    """compute_suite_metrics.py:28-118: KPIs of one recorded run."""
    d = np.load(npz_path)
    grid_x = d['grid_x'] - (d['grid_x'][-1] / 2.0)
    grid_z = d['grid_z'] * 1000.0
    times = d['times'] * 1000.0
    xz_T = d['xz_T']
    xz_phase = d['xz_phase']
    depths = d['crater_depths_mm'] * 1000.0
    DX = float(d['DX']) if 'DX' in d else 0.015
    DZ = float(d['DZ']) if 'DZ' in d else 0.015
    n_ep = len(times)
    ip = min(int(round(pulse_duration_s * 1000.0)) - 1, n_ep - 1)
    final_depth = float(depths[ip])
    avg_depth = float(np.mean(depths[:ip + 1]))
    speed = float((final_depth / 1000.0) / pulse_duration_s)
    ph_end, T_end = xz_phase[ip], xz_T[ip]
    gap = (grid_z >= 165.0) & (grid_z <= 465.0)
    solid = (grid_z >= 0) & (grid_z <= final_depth) & ~gap
    dia = [np.sum(ph_end[:, iz] == 3) * DX for iz in range(len(grid_z)) if solid[iz] and np.any(ph_end[:, iz] == 3)]
    hole = float(np.mean(dia)) if dia else 0.0
    ixc = np.abs(grid_x) <= 0.30
    out = {}
    for name, (z0, z1) in (('core', (780.0, 1035.0)), ('ven', (2355.0, 2445.0))):
        m = np.outer(ixc, (grid_z >= z0) & (grid_z <= z1))
        valid = m & (ph_end != 3)
        out[name + '_avg_T'] = float(np.mean(T_end[valid])) if np.any(valid) else 20.0
        out[name + '_max_T'] = float(np.max(T_end[m])) if np.any(m) else 20.0
    v_abl = v_boil = 0.0
    for iz in range(len(grid_z)):
        if grid_z[iz] >= 0 and not (165.0 <= grid_z[iz] <= 465.0):
            na = np.sum(ph_end[:, iz] == 3)
            nb = np.sum(ph_end[:, iz] == 2)
            ra, rt = na * DX / 2.0, (na + nb) * DX / 2.0
            v_abl += np.pi * ra ** 2 * DZ
            v_boil += np.pi * max(0.0, rt ** 2 - ra ** 2) * DZ
    wl = int(round(float(d['wavelength_nm']))) if 'wavelength_nm' in d else None
    power = float(d['power_w']) if 'power_w' in d else None
    if 'laser_states' in d and wl is None:
        sp = {'R': 0.0, 'A': 0.0, 'T': 0.0}
    elif wl in SPECTRO_MAP:
        sp = SPECTRO_MAP[wl]
    else:
        sp = {'R': 4.0, 'A': 95.0, 'T': 0.0}
    energy = (4.60 if ('laser_states' in d or 'dual' in npz_path)
              else (float(power * pulse_duration_s) if power is not None else 8.00))
    return {'path': npz_path, 'wavelength_nm': wl, 'power_w': power, 'final_depth_um': final_depth,
            'avg_depth_um': avg_depth, 'speed_mm_s': speed, 'avg_hole_dia_mm': hole, 'R_pct': sp.get('R', 0.0),
            'A_pct': sp.get('A', 0.0), 'T_pct': sp.get('T', 0.0), 'core_avg_T': out['core_avg_T'],
            'core_max_T': out['core_max_T'], 'ven_avg_T': out['ven_avg_T'], 'ven_max_T': out['ven_max_T'],
            'v_abl_mm3': float(v_abl), 'v_boil_mm3': float(v_boil), 'energy_used_j': energy}
