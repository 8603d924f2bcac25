"""The reference's Keller et al. (2016) thermal-damage threshold on this engine.

Follows knowledge/simulation_lit_comparison/code/keller_benchmark.py:44-122: a
water-like column (rho c 4.18 J/cm^3/K, k 0.6 W/m/K, mu_a 850 /cm, 4 x 4 x 133
voxels of 50 x 50 x 3 um, adiabatic, 25 C) under a rectangular pulse of peak
irradiance I with Beer-Lambert absorbed power, explicit sub-steps of 15 us with
the Arrhenius integral accumulated after each, then 60 ms of cooling.  The
threshold fluence is where the surface Omega first reaches 1, found by 28
bisection steps on a geometric bracket of 0.01 to 100 J/cm^2.

The sub-steps run on the device as ``bd_substeps`` with the Arrhenius flag only
(no latent plateau, no char, no removal), the pulse as ``is_active`` on and the
cooling as off; every bisection step restarts the column from 25 C.  No Monte
Carlo is involved, so the thresholds are deterministic.

The reference's recorded thresholds (``benchmark_results.json``, ``KE-*``) are
the golden data.

    python -m simulation.keller_suite [--golden FILE]
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _bdlib as B
from .coupled_ablation import R_GAS
from .voxel_thermal_3d import _physics

MU_A_10P6 = 85.0          # 1/mm
K_WATER = 0.6e-3          # W/(mm K)
RHO_CP_WATER = 4.18e-3    # J/(mm^3 K)
T0 = 25.0
ARRHENIUS = {'Shimojo': (7.39e37, 2.577e5), 'Henriques-Moritz': (3.1e98, 6.3e5)}
PULSES_MS = (25.0, 4.3)


class Column:
    """The 1-D water column of keller_benchmark.py:54-79 on one device context."""

    def __init__(self, A, Ea, mu_a=MU_A_10P6, dz=0.003, Nz=133, dx=0.05, dt=1.5e-5):
        self.shape = (4, 4, Nz)
        self.dt = dt
        self.ctx = B.Context(self.shape, (dx, dx, dz))
        self.ctx.set_physics(_physics(dt=dt, h_eff=0.0, t_ambient=T0, has_arrhenius=1, arr_A=float(A),
                                      arr_Ea_over_R=float(Ea) / R_GAS))
        z = (np.arange(Nz) + 0.5) * dz
        self.profile = mu_a * np.exp(-mu_a * z)
        for name, v in (('K', K_WATER), ('RHO_CP', RHO_CP_WATER), ('H_LAT', 1e30)):
            self.ctx.fill(name, v)
        self.ctx.fill('ABLATED', 0)
        self.ctx.fill('DRY', 0)
        self.ctx.fill('H_ACC', 0.0)

    def omega_after_pulse(self, I_peak, t_pulse, cool=0.06) -> float:
        """Maximum surface Omega after a pulse of peak irradiance I_peak (W/mm^2)."""
        Q = np.broadcast_to(I_peak * self.profile, self.shape)
        self.ctx.upload('Q', np.ascontiguousarray(Q, dtype=np.float64))
        self.ctx.fill('T', T0)
        self.ctx.fill('OMEGA', 0.0)
        n_on, n_cool = int(round(t_pulse / self.dt)), int(round(cool / self.dt))
        self.ctx.substeps(n_on, is_active=1, flags=B.SUB_ARRHENIUS)
        self.ctx.substeps(n_cool, is_active=0, flags=B.SUB_ARRHENIUS)
        return float(self.ctx.download('OMEGA')[:, :, 0].max())


def threshold_fluence(t_pulse_ms, A, Ea, column=None) -> float:
    """keller_benchmark.py:82-93: fluence (J/cm^2) where the surface Omega reaches 1."""
    col = column or Column(A, Ea)
    t = t_pulse_ms * 1e-3
    lo, hi = 1e-4, 1.0
    for _ in range(28):
        mid = np.sqrt(lo * hi)
        if col.omega_after_pulse(mid / t, t) >= 1.0:
            hi = mid
        else:
            lo = mid
    return float(np.sqrt(lo * hi) * 100.0)


def run_all() -> dict:
    out = {}
    for name, (A, Ea) in ARRHENIUS.items():
        col = Column(A, Ea)
        for tp in PULSES_MS:
            out[f'{name}|{tp}'] = threshold_fluence(tp, A, Ea, column=col)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--golden', default=None)
    args = ap.parse_args(argv)
    t0 = time.time()
    res = {'thresholds_J_cm2': run_all(), 'runtime_s': None}
    res['runtime_s'] = round(time.time() - t0, 2)
    if args.golden:
        res['recorded'] = json.load(open(args.golden))
    print(json.dumps(res), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
