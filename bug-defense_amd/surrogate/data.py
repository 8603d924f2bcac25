"""Training data for the surrogate, generated by the device MC itself.

Each sample is a 64^3 property field at 15 um: ambient air on top, a seeded
planar stack drawn like the coupled layered-stack workload (per layer mu_a
U[0.01, 20] /mm, mu_s U[1, 40] /mm, g U[0.7, 0.95], n U[1.0, 1.55], some layers
air gaps), and up to three ellipsoidal craters near the beam axis, emptied to
air and lined with char (mu_a 15 /mm, mu_s 20 /mm, the char state of the
coupled runs); the beam has a random angle in [0, 30] deg and spot in
[0.35, 0.65] mm, positioned as the coupled loop positions its window (axis at
the window centre at mid-depth).  The target is the window's log fluence from
an MC run of n_photons on the device; features and targets are written
straight into device buffers (no host round trip of the 3-D fields).

With mc_photons > 0 the samples are for the refining network: channel 5 of
the features is the fluence of a separate mc_photons-photon MC of the same
state (independent photon streams), and the target is still the n_photons run.
"""

import math

import numpy as np

from simulation import _bdlib as B
from simulation.voxel_mc_3d import make_beam

from .network import CIN, T

VOXEL_MM = 0.015
CHAR_OPTICS = (15.0, 20.0, 0.9, 1.5)   # mu_a, mu_s [1/mm], g, n of the char lining


def random_sample(rng, shape=(T, T, T), voxel=VOXEL_MM):
    """Property grids (float32) and beam parameters of one training sample."""
    Nx, Ny, Nz = shape
    mu_a = np.zeros(shape, np.float32)
    mu_s = np.zeros(shape, np.float32)
    g = np.zeros(shape, np.float32)
    n = np.ones(shape, np.float32)
    n_air = int(rng.integers(4, 25)) if rng.random() < 0.8 else int(rng.integers(25, 48))
    z = n_air
    while z < Nz:
        th = int(rng.integers(1, 31))
        z1 = min(Nz, z + th)
        if rng.random() < 0.12 and z > n_air:
            z = z1  # an air gap
            continue
        mu_a[:, :, z:z1] = rng.uniform(0.01, 20.0)
        mu_s[:, :, z:z1] = rng.uniform(1.0, 40.0)
        g[:, :, z:z1] = rng.uniform(0.7, 0.95)
        n[:, :, z:z1] = rng.uniform(1.0, 1.55)
        z = z1
    theta = math.radians(rng.uniform(0.0, 30.0))
    spot = float(rng.uniform(0.35, 0.65))
    jx, jy = rng.uniform(-4.0, 4.0, 2)
    xc = (0.5 * Nx + jx) * voxel - 0.5 * Nz * voxel * math.tan(theta)
    yc = (0.5 * Ny + jy) * voxel
    # craters along the beam axis, opening at the surface or buried
    X, Y, Z = np.meshgrid(np.arange(Nx) + 0.5, np.arange(Ny) + 0.5, np.arange(Nz) + 0.5, indexing='ij')
    for _ in range(int(rng.integers(0, 4))):
        cz = n_air + rng.uniform(-4.0, 30.0)
        cx = xc / voxel + cz * math.tan(theta) + rng.normal(0.0, 4.0)
        cy = yc / voxel + rng.normal(0.0, 4.0)
        ax, ay, az = rng.uniform(3.0, 22.0), rng.uniform(3.0, 22.0), rng.uniform(3.0, 30.0)
        r2 = ((X - cx) / ax) ** 2 + ((Y - cy) / ay) ** 2 + ((Z - cz) / az) ** 2
        tissue = mu_s > 0
        if rng.random() < 0.6:
            shell = (r2 < 1.35) & (r2 >= 1.0) & tissue
            mu_a[shell], mu_s[shell], g[shell], n[shell] = CHAR_OPTICS
        hole = r2 < 1.0
        mu_a[hole], mu_s[hole], g[hole], n[hole] = 0.0, 0.0, 0.0, 1.0
    return dict(mu_a=mu_a, mu_s=mu_s, g=g, n=n, theta=theta, spot=spot, xc=xc, yc=yc)


def sample_beam(s, n_photons, seed, power=1.0):
    return make_beam(power, n_photons, s['spot'], math.degrees(s['theta']), 0.0, (s['xc'], s['yc'], 0.0), 0.0, 0,
                     'gaussian', seed, score_fluence=True)


NOISY_PHOTON_OFFSET = 1 << 40   # photon ids of the low-photon input run: disjoint from the target run's


def noisy_beam(beam, mc_photons, photon_offset=NOISY_PHOTON_OFFSET, fluence=True):
    """The beam of an extra MC run of the same state: same key, disjoint photon ids.
    With fluence False the run scores absorbed weight only (no path-length tally)."""
    b = B.Beam.from_buffer_copy(beam)
    b.n_photons = int(mc_photons)
    b.photon_offset = int(photon_offset)
    b.score_fluence = 1 if fluence else 0
    return b


class Dataset:
    """N samples resident in device memory: X [N,64,64,64,8], t and w [N,64,64,64]."""

    def __init__(self, n, device=0):
        self.n = int(n)
        self.X = B.DeviceArray((self.n, T, T, T, CIN), device=device)
        self.t = B.DeviceArray((self.n, T, T, T), device=device)
        self.w = B.DeviceArray((self.n, T, T, T), device=device)
        self.meta = []

    def chunk(self, c, nb):
        """Samples [c nb, (c + 1) nb) as one batch of views."""
        return (self.X.view(c, (nb, T, T, T, CIN)), self.t.view(c, (nb, T, T, T)), self.w.view(c, (nb, T, T, T)))


def generate_coupled(n, n_photons=1_000_000, seed=0, shape=(192, 192, 192), epochs_per_run=150, snapshot_every=3,
                     progress=None, device=0, max_windows=4, mc_photons=0, power=None, theta_deg=None,
                     spot_dia=None):
    """Dataset of n windows taken from coupled runs of the layered-stack workload,
    i.e. the states the coupled loop hands the network: every `snapshot_every`-th
    epoch of a plain-MC run (random stack seed, never seed 0; power 20-80 W, angle
    0-30 deg, spot 0.35-0.65 mm unless fixed by power / theta_deg / spot_dia), the
    tiled beam windows of that state (up to max_windows), their features
    (transport optics with the plume), and as targets one fresh n_photons MC over
    the whole domain."""
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.workloads import layered_stack
    rng = np.random.default_rng(seed)
    ds = Dataset(n, device=device)
    i = 0
    while i < n:
        P = float(rng.uniform(20.0, 80.0)) if power is None else float(power)
        th = float(rng.uniform(0.0, 30.0)) if theta_deg is None else float(theta_deg)
        spot = float(rng.uniform(0.35, 0.65)) if spot_dia is None else float(spot_dia)
        kw = layered_stack(shape=shape, seed=int(rng.integers(1, 1 << 30)), power=P, theta_deg=th, spot_dia=spot,
                           plume_grid=tuple(s // 2 for s in shape))
        eng = CoupledAblationEngine(**kw)
        try:
            for ep in range(epochs_per_run):
                eng.step_epoch()
                if (ep + 1) % snapshot_every:
                    continue
                beam = eng._epoch_beam()
                ws = eng._ctx.sur_windows(beam, 12, max_windows)[:n - i]
                if mc_photons:
                    beam.seed = int(rng.integers(1, 1 << 62))
                    eng._ctx.mc_run(noisy_beam(beam, mc_photons))
                for k, o in enumerate(ws):
                    eng._ctx.sur_features_n(beam, [o], ds.X.view(i + k, (1, T, T, T, CIN)), mc_photons)
                beam.n_photons = int(n_photons)
                beam.seed = int(rng.integers(1, 1 << 62))
                beam.score_fluence = 1
                eng._ctx.mc_run(beam)
                for k, o in enumerate(ws):
                    eng._ctx.sur_targets(beam, o, ds.t.view(i + k, (T, T, T)), ds.w.view(i + k, (T, T, T)))
                    ds.meta.append({'theta': math.radians(th), 'spot': spot, 'power': P, 'epoch': ep + 1,
                                    'front': o[3], 'window': k})
                before = i
                i += len(ws)
                if progress and i // progress > before // progress:
                    eng._ctx.sync()
                    print(f"surrogate coupled data: {i}/{n} windows", flush=True)
                if i == n:
                    break
            eng._ctx.sync()
        finally:
            eng.close()
    return ds


def generate(n, n_photons=1_000_000, seed=0, progress=None, device=0, mc_seed_offset=0, mc_photons=0):
    """Dataset of n samples from the device MC (seeded, reproducible); the
    geometry depends on `seed` only, the MC streams also on mc_seed_offset."""
    rng = np.random.default_rng(seed)
    ds = Dataset(n, device=device)
    ctx = B.Context((T, T, T), (VOXEL_MM,) * 3)
    origin = (0, 0, 0)
    try:
        for i in range(n):
            s = random_sample(rng)
            for name in ('mu_a', 'mu_s', 'g', 'n'):
                ctx.upload(name.upper(), s[name])
            beam = sample_beam(s, n_photons, seed * 1_000_003 + i + 1 + mc_seed_offset * 7_777_777_777)
            if mc_photons:
                ctx.mc_run(noisy_beam(beam, mc_photons))
            ctx.sur_features_n(beam, [origin], ds.X.view(i, (1, T, T, T, CIN)), mc_photons)
            ctx.mc_run(beam)
            ctx.sur_targets(beam, origin, ds.t.view(i, (T, T, T)), ds.w.view(i, (T, T, T)))
            ds.meta.append({k: s[k] for k in ('theta', 'spot', 'xc', 'yc')})
            if progress and (i + 1) % progress == 0:
                ctx.sync()
                print(f"surrogate data: {i + 1}/{n} samples", flush=True)
        ctx.sync()
    finally:
        ctx.close()
    return ds
