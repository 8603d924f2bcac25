"""The reference's GPU-context tests, run through this build's Metal-named modules.

tests/test_gpu_thermal.py:16-136, tests/test_plume.py:182-286 and tests/test_gpu_mc.py:14-126 of
the reference drive ``run_thermal3d_substeps_metal``, ``MetalAblationContext`` and
``run_voxel_mc_metal``.  The same calls, fixtures and bars run here on the gfx950 kernels; where
the reference compared against its CPU engine (``engine='cpu'``, which this build does not have)
the FP64 oracle takes its place.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_is_metal_available_reports_the_gpu_engine():
    from simulation.metal_ablation_solver import is_metal_available as a
    from simulation.voxel_mc_metal import is_metal_available as c
    from simulation.voxel_thermal_metal import is_metal_available as b
    assert a() and b() and c()


def test_voxel_thermal_substeps_metal_exact_parity(oracle):
    """tests/test_gpu_thermal.py:16-50: 12x12x10 two-stratum grid, 10 steps; max |dT| < 1e-4 K."""
    from simulation.voxel_thermal_metal import run_thermal3d_substeps_metal
    Nx, Ny, Nz = 12, 12, 10
    dx, dy, dz, dt = 0.05, 0.05, 0.03, 2e-5
    T_cpu = np.full((Nx, Ny, Nz), 25.0)
    T_next = np.empty_like(T_cpu)
    Q = np.zeros_like(T_cpu)
    Q[4:8, 4:8, 0:3] = 800.0
    k = np.full((Nx, Ny, Nz), 0.25e-3)
    k[:, :, 5:] = 0.50e-3
    rc = np.full((Nx, Ny, Nz), 1.8e-3)
    rc[:, :, 5:] = 3.5e-3
    for _ in range(10):
        oracle.thermal_step(T_cpu, T_next, Q, k, rc, Nx, Ny, Nz, dx, dy, dz, dt, 12.0, 25.0, 1)
        T_cpu, T_next = T_next, T_cpu
    T_gpu = run_thermal3d_substeps_metal(np.full((Nx, Ny, Nz), 25.0), Q, k, rc, Nx, Ny, Nz, dx, dy, dz, dt, 10, 12.0,
                                         25.0, 1)
    assert T_gpu.dtype == np.float64
    assert float(np.max(np.abs(T_cpu - T_gpu))) < 1e-4
    assert abs(np.max(T_cpu) - np.max(T_gpu)) < 1e-3


def test_metal_thermal_context_matches_one_shot_call():
    """MetalThermalContext (voxel_thermal_metal.py:288-383): set_T/set_Q/advance/get_T in two
    halves, laser on then off, equals the one-shot function called twice."""
    from simulation.voxel_thermal_metal import MetalThermalContext, run_thermal3d_substeps_metal
    shape = (10, 9, 14)
    rng = np.random.default_rng(2)
    k = rng.uniform(0.2e-3, 0.6e-3, shape)
    rc = rng.uniform(2e-3, 4e-3, shape)
    Q = rng.uniform(0.0, 200.0, shape)
    T0 = rng.uniform(20.0, 60.0, shape)
    ctx = MetalThermalContext(*shape, 0.03, 0.03, 0.02, 2e-5, k, rc, h_eff=5e-3, t_ambient=20.0)
    ctx.set_T(T0)
    ctx.set_Q(Q)
    ctx.advance(7, is_active=1)
    ctx.advance(5, is_active=0)
    a = ctx.get_T()
    b = run_thermal3d_substeps_metal(T0, Q, k, rc, *shape, 0.03, 0.03, 0.02, 2e-5, 7, 5e-3, 20.0, 1)
    b = run_thermal3d_substeps_metal(b, Q, k, rc, *shape, 0.03, 0.03, 0.02, 2e-5, 5, 5e-3, 20.0, 0)
    np.testing.assert_array_equal(a, b)
    assert ctx.get_T_view().dtype == np.float32
    ctx.close()


def _fused_chain(oracle, n_steps, T, Q, k, rho, H, Om, chi, dry, abl, dims, dt, h_eff, t_amb, t_abl, t_void,
                 arr, char):
    Nx, Ny, Nz, dx, dy, dz = dims
    Tn = np.empty_like(T)
    a8, d8 = abl.astype(np.uint8), dry.astype(np.uint8)
    for _ in range(n_steps):
        oracle.thermal_step(T, Tn, Q, k, rho, Nx, Ny, Nz, dx, dy, dz, dt, h_eff, t_amb, 1)
        oracle.plateau(Tn, H, rho, a8, t_abl, Nx, Ny, Nz, d8, t_void)
        if arr is not None:
            oracle.arrhenius(Tn, Om, a8, arr[0], arr[1], dt)
        if char is not None:
            oracle.char(Tn, chi, d8, a8, char[0], char[1], dt)
        T, Tn = Tn, T
    return T


def test_fused_ablation_exact_parity(oracle):
    """tests/test_gpu_thermal.py:76-136 through MetalAblationContext: 14x14x12, 30 fused steps;
    T < 1e-4 K, H / Omega / chi < 1e-5 against the FP64 chain of module kernels."""
    from simulation.metal_ablation_solver import MetalAblationContext, is_metal_available
    assert is_metal_available()
    Nx, Ny, Nz = 14, 14, 12
    dx, dy, dz = 0.03, 0.03, 0.02
    dt, h_eff, t_amb, t_abl, t_void = 2.5e-5, 15.0e-6, 20.0, 100.0, 20.0
    arr = char = (1.0e7, 65000.0 / 8.314)
    shape = (Nx, Ny, Nz)
    Q = np.zeros(shape)
    Q[5:9, 5:9, 0:3] = 750.0
    k = np.full(shape, 0.3e-3)
    k[:, :, 6:] = 0.5e-3
    rho = np.full(shape, 2.0e-3)
    rho[:, :, 6:] = 3.5e-3
    H, Om, chi = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    dry = np.zeros(shape, bool)
    dry[6:8, 6:8, 0:2] = True
    abl = np.zeros(shape, bool)
    T_cpu = _fused_chain(oracle, 30, np.full(shape, 25.0), Q, k, rho, H, Om, chi, dry, abl,
                         (Nx, Ny, Nz, dx, dy, dz), dt, h_eff, t_amb, t_abl, t_void, arr, char)
    ctx = MetalAblationContext(Nx, Ny, Nz, dx, dy, dz, dt, h_eff, t_amb, t_abl, t_void, arrhenius=arr,
                               char_params=char)
    ctx.set_state(np.full(shape, 25.0), np.zeros(shape), np.zeros(shape), np.zeros(shape), dry, abl, Q, k, rho)
    ctx.advance_substeps(30, is_active=1)
    T_gpu, H_gpu, O_gpu, chi_gpu = ctx.get_state()
    assert H.max() > 0.0 and chi.max() > 0.0 and Om.max() > 0.0
    assert float(np.max(np.abs(T_cpu - T_gpu))) < 1e-4
    assert float(np.max(np.abs(H - H_gpu))) < 1e-5
    assert float(np.max(np.abs(Om - O_gpu))) < 1e-5
    assert float(np.max(np.abs(chi - chi_gpu))) < 1e-5
    ctx.close()


def test_plume_diffusion_thermal_independence():
    """tests/test_plume.py:182-227: the plume diffusion in the fused step leaves T, H, Omega and chi
    bit-identical."""
    from simulation.metal_ablation_solver import MetalAblationContext
    Nx, Ny, Nz = 12, 12, 10
    dx, dy, dz = 0.03, 0.03, 0.02
    shape = (Nx, Ny, Nz)
    k = np.full(shape, 0.3e-3, dtype=np.float32)
    rho = np.full(shape, 2.0e-3, dtype=np.float32)
    Q = np.zeros(shape, dtype=np.float32)
    Q[4:8, 4:8, 2:6] = 50.0
    dry = np.zeros(shape, dtype=np.uint8)
    abl = np.zeros(shape, dtype=np.uint8)
    abl[5:7, 5:7, 0:3] = 1
    T0 = np.full(shape, 25.0, dtype=np.float32)
    z = np.zeros(shape, dtype=np.float32)
    kin = dict(arrhenius=(1e4, 5000.0 / 8.314), char_params=(1e3, 4000.0 / 8.314))
    c1 = MetalAblationContext(Nx, Ny, Nz, dx, dy, dz, 5e-5, 10.0, 20.0, 100.0, 300.0, **kin)
    c1.set_state(T0, z, z, z, dry, abl, Q, k, rho)
    c1.advance_substeps(50, is_active=1)
    s1 = c1.get_state()
    c2 = MetalAblationContext(Nx, Ny, Nz, dx, dy, dz, 5e-5, 10.0, 20.0, 100.0, 300.0,
                              plume_params={'D': 0.5, 'vz': 15.0, 'tau_linger': 0.010}, **kin)
    p0 = np.zeros(shape, dtype=np.float32)
    p0[5:7, 5:7, 1] = 10.0
    c2.set_state(T0, z, z, z, dry, abl, Q, k, rho, plume=p0)
    c2.advance_substeps(50, is_active=1)
    s2 = c2.get_state()
    for a, b in zip(s1, s2):
        assert np.max(np.abs(a - b)) == 0.0
    p = c2.get_plume()
    assert p.dtype == np.float32 and p.max() > 0.0 and np.all(p[abl == 0] == 0.0)
    for c in (c1, c2):
        c.close()


def test_plume_anisotropic_diffusion_gpu_ratio():
    """tests/test_plume.py:230-286: D_z = 4 D_xy, 20 steps; the context's z/x spread ratio equals the
    plume field's within 1e-3.  The context takes the Metal solver's diffusion step
    min(dt, 0.45 / (2 d_rate)) (metal_ablation_solver.py:234-236), which binds here
    (d_rate = 15 000 /s gives 15 us < 20 us), so the field it is compared with takes the same
    20 steps of 15 us; the CPU scheme's own 20 steps of 20 us still show the D_z / D_xy elongation."""
    from simulation.metal_ablation_solver import MetalAblationContext
    from simulation.plume import MultispectralPlumeField
    shape = (31, 31, 31)
    dx = dy = dz = 0.02
    plume = MultispectralPlumeField(shape=shape, dx_mm=dx, dy_mm=dy, dz_mm=dz, model='diffusion', D_xy=1.0, D_z=4.0,
                                    v_drift=0.0, tau_linger_s=float('inf'), blast_expansion=False)
    newly = np.zeros(shape, dtype=bool)
    newly[15, 15, 15] = True
    plume.inject(newly)
    full = np.ones(shape, dtype=bool)
    clamp = MultispectralPlumeField(shape=shape, dx_mm=dx, dy_mm=dy, dz_mm=dz, model='diffusion', D_xy=1.0, D_z=4.0,
                                    v_drift=0.0, tau_linger_s=float('inf'), blast_expansion=False)
    clamp.inject(newly)
    d_rate = 1.0 * (1.0 / dx ** 2 + 1.0 / dy ** 2) + 4.0 / dz ** 2
    dt_metal = min(2e-5, 0.45 / (2.0 * d_rate))
    assert dt_metal < 2e-5
    for _ in range(20):
        plume.diffuse_substep(full, dt_sub=2e-5)
        clamp.diffuse_substep(full, dt_sub=dt_metal)
    r = plume.raw_field
    assert r[15, 17, 15] == pytest.approx(r[17, 15, 15], rel=1e-4) and r[15, 15, 17] > 2.0 * r[17, 15, 15]
    r = clamp.raw_field
    val_z, val_x = r[15, 15, 17], r[17, 15, 15]
    ctx = MetalAblationContext(31, 31, 31, dx, dy, dz, 2e-5, 10.0, 20.0, 100.0, 300.0,
                               plume_params={'D_xy': 1.0, 'D_z': 4.0, 'vz': 0.0, 'tau_linger': float('inf')})
    p0 = np.zeros(shape, dtype=np.float32)
    p0[15, 15, 15] = plume.delta_mu
    zf = np.zeros(shape, dtype=np.float32)
    ones = np.ones(shape, dtype=np.uint8)
    ctx.set_state(np.full(shape, 20.0, dtype=np.float32), zf, zf, zf, ones, ones, zf, zf, zf, plume=p0)
    ctx.advance_substeps(20, is_active=1)
    p = ctx.get_plume()
    assert p[15, 15, 17] > 2.0 * p[17, 15, 15]
    assert (p[15, 15, 17] / p[17, 15, 15]) == pytest.approx(val_z / val_x, rel=1e-3)
    ctx.close()


def test_metal_mc_homogeneous_cube_energy_closure():
    """tests/test_gpu_mc.py:14-38."""
    from simulation.voxel_mc_metal import run_voxel_mc_metal
    shape = (20, 20, 25)
    props = {'mu_a': np.full(shape, 0.2), 'mu_s': np.full(shape, 10.0), 'g': np.full(shape, 0.8),
             'n': np.full(shape, 1.4)}
    res = run_voxel_mc_metal(power=5.0, n_photons=50_000, Lx=1.0, Ly=1.0, Lz=1.0, Nx=20, Ny=20, Nz=25,
                             custom_3d_props=props, spot_dia=0.20, base_seed=1234)
    assert abs(res['energy_closure'] - 1.0) < 0.01
    raw = res['spectro_raw']
    assert raw['Absorption'] > 0.0 and raw['R_spec'] > 0.0 and raw['R_diff'] > 0.0
    assert res['engine'] == 'gpu' and res['lateral_escape_z'].shape == (25,)


def test_metal_mc_multilayer_non_normal_gaussian_parity(oracle):
    """tests/test_gpu_mc.py:40-126: 3 layers, 30 deg / 45 deg, 100 000 photons; R_spec within
    0.002 of exact Fresnel, closure < 0.01, A / R_diff / T within 0.015 of the CPU engine (the
    oracle here), depth-profile correlation > 0.995."""
    from simulation.voxel_mc_metal import run_voxel_mc_metal
    shape = (30, 30, 40)
    mu_a, mu_s, g, n = (np.zeros(shape) for _ in range(4))
    for (a, b), (ma, ms, gg, nn) in (((0, 10), (0.60, 16.0, 0.85, 1.52)), ((10, 24), (0.05, 4.0, 0.80, 1.34)),
                                     ((24, 40), (0.70, 18.0, 0.88, 1.42))):
        mu_a[:, :, a:b], mu_s[:, :, a:b], g[:, :, a:b], n[:, :, a:b] = ma, ms, gg, nn
    props = {'mu_a': mu_a, 'mu_s': mu_s, 'g': g, 'n': n}
    common = dict(power=10.0, n_photons=100_000, Lx=2.0, Ly=2.0, Lz=1.0, Nx=30, Ny=30, Nz=40, custom_3d_props=props,
                  theta_inc_deg=30.0, phi_inc_deg=45.0, spot_dia=0.40, beam_center=(1.0, 1.0))
    cpu = oracle.run_voxel_mc3d(beam_profile='gaussian', seed=11, n_threads=8, **common)
    gpu = run_voxel_mc_metal(base_seed=777, **common)
    ci = np.cos(np.radians(30.0))
    st = np.sin(np.radians(30.0)) / 1.52
    ct = np.sqrt(1.0 - st ** 2)
    rs = (ci - 1.52 * ct) / (ci + 1.52 * ct)
    rp = (ct - 1.52 * ci) / (ct + 1.52 * ci)
    assert abs(gpu['spectro_raw']['R_spec'] - 0.5 * (rs ** 2 + rp ** 2)) < 0.002
    assert abs(gpu['energy_closure'] - 1.0) < 0.01
    for key in ('Absorption', 'R_diff', 'Transmission'):
        assert abs(gpu['spectro_raw'][key] - cpu['spectro_raw'][key]) < 0.015, key
    corr = np.corrcoef(cpu['A_3d'].sum(axis=(0, 1)), gpu['A_3d'].sum(axis=(0, 1)))[0, 1]
    assert corr > 0.995
