"""GPU transport kernel against the FP64 oracle and the closed forms (through the C-ABI).

The device and the oracle trace the same Philox streams (same seed, same photon
ids, same draw order), FP32 on the device and FP64 on the CPU, so their
results are strongly correlated; tolerances are still statistical.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAYERS3 = None


def _three_layer():
    Nx, Ny, Nz = 30, 30, 40
    mu_a = np.zeros((Nx, Ny, Nz))
    mu_s = np.zeros_like(mu_a)
    g = np.zeros_like(mu_a)
    n = np.zeros_like(mu_a)
    for (a, b), (ma, ms, gg, nn) in (((0, 10), (0.60, 16.0, 0.85, 1.52)), ((10, 24), (0.05, 4.0, 0.80, 1.34)),
                                     ((24, 40), (0.70, 18.0, 0.88, 1.42))):
        mu_a[:, :, a:b], mu_s[:, :, a:b], g[:, :, a:b], n[:, :, a:b] = ma, ms, gg, nn
    return {'mu_a': mu_a, 'mu_s': mu_s, 'g': g, 'n': n}


def test_homogeneous_cube_closure_reference_fixture():
    """tests/test_gpu_mc.py:14-38 of the reference."""
    from simulation.voxel_mc_3d import run_voxel_mc3d
    shape = (20, 20, 25)
    props = {'mu_a': np.full(shape, 0.2), 'mu_s': np.full(shape, 10.0), 'g': np.full(shape, 0.8),
             'n': np.full(shape, 1.4)}
    res = run_voxel_mc3d(power=5.0, n_photons=50_000, Lx=1.0, Ly=1.0, Lz=1.0, Nx=20, Ny=20, Nz=25,
                         custom_3d_props=props, spot_dia=0.20, seed=1234)
    assert abs(res['energy_closure'] - 1.0) < 0.01
    s = res['spectro_raw']
    assert s['Absorption'] > 0.0 and s['R_spec'] > 0.0 and s['R_diff'] > 0.0
    assert res['engine'] == 'gpu'


def test_multilayer_oblique_parity_with_oracle(oracle):
    """tests/test_gpu_mc.py:40-126 of the reference (theta 30, phi 45, 100k photons),
    against the oracle, with tolerances tightened from 0.015 to 0.005."""
    from simulation.voxel_mc_3d import run_voxel_mc3d
    props = _three_layer()
    kw = dict(power=10.0, n_photons=100_000, Lx=2.0, Ly=2.0, Lz=1.0, Nx=30, Ny=30, Nz=40, custom_3d_props=props,
              theta_inc_deg=30.0, phi_inc_deg=45.0, spot_dia=0.40, beam_center=(1.0, 1.0), seed=777)
    g = run_voxel_mc3d(**kw)
    c = oracle.run_voxel_mc3d(**kw, n_threads=8)
    th = np.radians(30.0)
    n1, n2 = 1.0, 1.52
    ci = np.cos(th)
    ct = np.sqrt(1.0 - ((n1 / n2) * np.sin(th)) ** 2)
    R = 0.5 * (((n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)) ** 2 + ((n1 * ct - n2 * ci) / (n1 * ct + n2 * ci)) ** 2)
    assert abs(g['spectro_raw']['R_spec'] - R) < 1e-6
    assert abs(g['energy_closure'] - 1.0) < 0.01
    for key in ('Absorption', 'R_diff', 'Transmission'):
        assert abs(g['spectro_raw'][key] - c['spectro_raw'][key]) < 0.005, key
    corr = np.corrcoef(g['A_3d'].sum(axis=(0, 1)), c['A_3d'].sum(axis=(0, 1)))[0, 1]
    assert corr > 0.999


def test_per_voxel_three_sigma_parity(oracle):
    """Per-voxel absorbed weight: GPU vs oracle at equal photon counts.  With 16 independent
    batches per side the per-voxel standard error is estimated from the batch scatter; the
    fraction of well-sampled voxels beyond 3 sigma must stay at the Gaussian rate (< 1 %),
    and R, T, A agree within 0.5 % absolute."""
    from simulation.voxel_mc_3d import run_voxel_mc3d
    props = _three_layer()
    kw = dict(power=1.0, n_photons=25_000, Lx=2.0, Ly=2.0, Lz=1.0, Nx=30, Ny=30, Nz=40, custom_3d_props=props,
              theta_inc_deg=22.0, spot_dia=0.5)
    G, C, sg, sc = [], [], [], []
    for b in range(16):
        r = run_voxel_mc3d(seed=100 + b, **kw)
        G.append(r['A_3d'])
        sg.append(r['spectro_raw'])
        q = oracle.run_voxel_mc3d(seed=900 + b, n_threads=8, **kw)  # independent streams
        C.append(q['A_3d'])
        sc.append(q['spectro_raw'])
    G, C = np.array(G), np.array(C)
    mg, mc = G.mean(0), C.mean(0)
    se = np.sqrt(G.var(0, ddof=1) / 16 + C.var(0, ddof=1) / 16)
    ok = (mg > 0) & (mc > 0) & (se > 0) & (np.minimum(mg, mc) > 20 * se / 4)
    z = np.abs(mg - mc)[ok] / se[ok]
    assert ok.sum() > 1000
    assert np.mean(z > 3.0) < 0.01, f"{np.mean(z > 3.0):.4f} of {ok.sum()} voxels beyond 3 sigma"
    for key in ('Absorption', 'R_diff', 'Transmission', 'R_spec'):
        a = np.mean([s[key] for s in sg])
        b = np.mean([s[key] for s in sc])
        assert abs(a - b) < 0.005, key


def test_fresnel_slab_closed_form_gpu():
    """V1c/V1d/V1e (verification_benchmarks.py:102-128) on the device."""
    from simulation.voxel_mc_3d import run_voxel_mc3d
    n, mu_a, d = 1.5, 1.0, 0.5
    R = ((n - 1.0) / (n + 1.0)) ** 2
    e1, e2 = np.exp(-mu_a * d), np.exp(-2.0 * mu_a * d)
    T_ref = (1 - R) ** 2 * e1 / (1 - R ** 2 * e2)
    Rd_ref = R * (1 - R) ** 2 * e2 / (1 - R ** 2 * e2)
    N = 400000
    r = run_voxel_mc3d(n_photons=N, spot_dia=0.01, Lx=2.0, Ly=2.0, Lz=0.55, Nx=4, Ny=4, Nz=11, seed=3,
                       layers_config=[{'z_end': 0.5, 'mu_a': mu_a, 'mu_s': 0.0, 'g': 0.0, 'n': n},
                                      {'z_end': 0.55, 'mu_a': 0.0, 'mu_s': 0.0, 'g': 0.0, 'n': 1.0}])
    s = r['spectro_raw']
    assert abs(s['Transmission'] - T_ref) < 4 * np.sqrt(T_ref * (1 - T_ref) / N)
    assert abs(s['R_diff'] - Rd_ref) < 4 * np.sqrt(Rd_ref * (1 - Rd_ref) / N)
    assert abs(s['R_spec'] - R) < 1e-7


def test_semi_infinite_h_function_gpu():
    """V2 (verification_benchmarks.py:87-99): Rd = 0.41495 (Chandrasekhar H-function)."""
    from simulation.voxel_mc_3d import run_voxel_mc3d
    vals = [run_voxel_mc3d(n_photons=100000, spot_dia=0.01, Lx=20.0, Ly=20.0, Lz=4.0, Nx=20, Ny=20, Nz=8, seed=s,
                           layers_config=[{'z_end': 4.0, 'mu_a': 1.0, 'mu_s': 9.0, 'g': 0.0, 'n': 1.0}]
                           )['spectro_raw']['R_diff'] for s in range(6)]
    m, se = np.mean(vals), np.std(vals, ddof=1) / np.sqrt(6)
    assert abs(m - 0.414947479) < 4 * se + 2e-4


@pytest.mark.parametrize("profile", ["gaussian", "tophat", "kim_exp_disk", "kim_gauss_disk"])
def test_beam_profiles_gpu_vs_oracle(oracle, profile):
    """Launch footprint (voxel_mc_3d.py:101-161): surface-absorber moments vs the oracle."""
    from simulation.voxel_mc_3d import run_voxel_mc3d
    kw = dict(n_photons=100000, spot_dia=0.8, Lx=2.0, Ly=2.0, Lz=0.5, Nx=100, Ny=100, Nz=1, beam_profile=profile,
              theta_inc_deg=20.0, layers_config=[{'z_end': 0.5, 'mu_a': 1e4, 'mu_s': 0.0, 'g': 0.0, 'n': 1.0}])
    g = run_voxel_mc3d(seed=5, **kw)['A_3d'][:, :, 0]
    c = oracle.run_voxel_mc3d(seed=5, n_threads=8, **kw)['A_3d'][:, :, 0]
    x = (np.arange(100) + 0.5) * 0.02 - 1.0
    for w in (x[:, None] ** 2 + 0 * x[None, :], 0 * x[:, None] + x[None, :] ** 2):
        assert (g * w).sum() / g.sum() == pytest.approx((c * w).sum() / c.sum(), rel=0.01)


def test_launch_axis_x_and_r_max_gpu(oracle):
    """launch_axis='x' and the r_max cut (voxel_mc_3d.py:157-177, 209-217, 309-317) vs the oracle."""
    from simulation.voxel_mc_3d import run_voxel_mc3d
    lay = [{'z_end': 0.3, 'mu_a': 0.5, 'mu_s': 10.0, 'g': 0.8, 'n': 1.4},
           {'z_end': 1.0, 'mu_a': 1.0, 'mu_s': 5.0, 'g': 0.9, 'n': 1.33}]
    kw = dict(n_photons=100000, spot_dia=0.3, Lx=1.0, Ly=4.0, Lz=4.0, Nx=25, Ny=20, Nz=20, layers_config=lay,
              launch_axis='x', r_max=0.4)
    g = run_voxel_mc3d(seed=11, **kw)
    c = oracle.run_voxel_mc3d(seed=11, n_threads=8, **kw)
    for key in ('R_diff', 'Absorption', 'Transmission', 'Lateral_loss'):
        assert abs(g['spectro_raw'][key] - c['spectro_raw'][key]) < 0.005, key
    assert g['lateral_escape_z'].sum() == pytest.approx(g['spectro_raw']['Lateral_loss'], rel=1e-5)


def test_tally_atomic_count_and_probe():
    """The transport's count of tally atomics: the same photons with and without
    the path-length tally issue the same absorbed-weight atomics, and the
    path-length atomics (one per voxel visit with a path) number at least as
    many and at most one per event and photon.  The atomic-rate probe runs."""
    from simulation import _bdlib as B
    from simulation.voxel_mc_3d import make_beam
    shape = (40, 40, 30)
    ctx = B.Context(shape, (0.025, 0.025, 0.025))
    try:
        for name, v in (('MU_A', 0.5), ('MU_S', 12.0), ('G', 0.85), ('N', 1.4)):
            ctx.upload(name, np.full(shape, v, dtype=np.float32))
        res = {}
        for sf in (0, 1):
            beam = make_beam(1.0, 20_000, 0.3, 15.0, 0.0, (0.5, 0.5, 0.0), 0.0, 0, 'gaussian', 4242,
                             score_fluence=bool(sf))
            res[sf] = ctx.mc_run(beam)
        n_a = res[0].tally_atomics
        # path-length additions: atomics the kernel issued plus those it deferred as records
        n_l = res[1].tally_atomics + res[1].tally_records - n_a
        assert res[0].tally_records == 0
        assert res[0].events == res[1].events
        assert n_a > 0 and n_l >= n_a
        assert n_l <= res[1].events + 20_000
    finally:
        ctx.close()
    rate = B.probe_atomic_rate(int(np.prod(shape)), n_pairs=4096 * 64)
    assert rate > 1e8


def _mc_context(shape, d, props, defer, records_per_photon=None):
    """A context whose transport defers the path-length tally (or adds it inline), and its
    record buffer size, set by the BUGDEF_MC_DEFER_FLUENCE / BUGDEF_MC_RECORDS_PER_PHOTON knobs
    that bd_create reads."""
    import os
    from simulation import _bdlib as B
    knobs = {"BUGDEF_MC_DEFER_FLUENCE": "1" if defer else "0"}
    if records_per_photon is not None:
        knobs["BUGDEF_MC_RECORDS_PER_PHOTON"] = str(records_per_photon)
    old = {k: os.environ.get(k) for k in knobs}
    os.environ.update(knobs)
    try:
        ctx = B.Context(shape, d)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    for name, key in (('MU_A', 'mu_a'), ('MU_S', 'mu_s'), ('G', 'g'), ('N', 'n')):
        ctx.upload(name, np.ascontiguousarray(props[key], dtype=np.float32))
    return ctx


@pytest.mark.parametrize("case", ["layered", "record_buffer_full"])
def test_deferred_path_length_tally_matches_inline(case):
    """The same photons with the path-length tally deferred to records (the default) and added
    inline by the transport: equal event counts and scalar tallies, and A and L grids equal up
    to the order of the float additions.  'record_buffer_full': a buffer of 4 records per photon
    (the 2^16 minimum) fills, so part of the tally goes inline after the records (both halves
    must add up)."""
    from simulation.voxel_mc_3d import make_beam
    shape, d, props = (30, 30, 40), (2.0 / 30, 2.0 / 30, 1.0 / 40), _three_layer()
    beam = make_beam(1.0, 25_000, 0.5, 22.0, 0.0, (1.0, 1.0, 0.0), 0.0, 0, 'gaussian', 77)
    per_photon = 4 if case == "record_buffer_full" else None
    out = {}
    for defer in (True, False):
        ctx = _mc_context(shape, d, props, defer, per_photon)
        try:
            r = ctx.mc_run(beam)
            out[defer] = (r, ctx.download('A'), ctx.download('L'))
            if defer:
                beam.score_fluence = 0
                n_a = ctx.mc_run(beam).tally_atomics  # absorbed-weight atomics alone
                beam.score_fluence = 1
        finally:
            ctx.close()
    (rd, Ad, Ld), (ri, Ai, Li) = out[True], out[False]
    assert rd.events == ri.events
    assert ri.tally_records == 0 and rd.tally_records > 0
    for key in ('specular', 'diffuse', 'transmitted', 'lateral', 'absorbed'):
        assert getattr(rd, key) == pytest.approx(getattr(ri, key), rel=1e-12, abs=1e-9), key
    # every path-length addition is made once: inline atomics + records against the inline run
    assert rd.tally_atomics + rd.tally_records == ri.tally_atomics
    for g_d, g_i in ((Ad, Ai), (Ld, Li)):
        assert g_d.sum() == pytest.approx(g_i.sum(), rel=1e-5)
        assert np.all(np.abs(g_d - g_i) <= 2e-4 * np.abs(g_i) + 1e-7 * g_i.max())
    n_l = ri.tally_atomics - n_a  # path-length additions of the launch
    if case == "layered":
        # inline only when a lane already has a record waiting: the voxel change and the end of a
        # packet in one event, at most once per photon
        assert n_l - 25_000 <= rd.tally_records < n_l
    else:
        cap = max(25_000 * 4, 1 << 16)
        assert n_l > 4 * cap
        assert 0.5 * cap < rd.tally_records <= cap
        assert rd.tally_atomics - n_a > 0.5 * n_l  # the rest went inline


def _engine_pair(**over):
    """Two engines on the same small layered stack and seeds: one with the path-length tally
    deferred to records (the default), one adding it inline."""
    import os
    from simulation.coupled_ablation import CoupledAblationEngine
    from simulation.workloads import layered_stack
    kw = layered_stack(shape=(40, 40, 56), n_photons=30_000, plume_grid=None, seed=5, power=2.0)
    kw.update(over)
    out = []
    for defer in (True, False):
        old = os.environ.get("BUGDEF_MC_DEFER_FLUENCE")
        os.environ["BUGDEF_MC_DEFER_FLUENCE"] = "1" if defer else "0"
        try:
            out.append(CoupledAblationEngine(**kw))
        finally:
            if old is None:
                del os.environ["BUGDEF_MC_DEFER_FLUENCE"]
            else:
                os.environ["BUGDEF_MC_DEFER_FLUENCE"] = old
    return out


def _tallies_close(ed, ei):
    for name in ('A', 'L'):
        gd, gi = ed._ctx.download(name), ei._ctx.download(name)
        assert gi.sum() > 0.0, name
        assert gd.sum() == pytest.approx(gi.sum(), rel=1e-5), name
        assert np.all(np.abs(gd - gi) <= 2e-4 * np.abs(gi) + 1e-7 * gi.max()), name


@pytest.mark.parametrize("schedule", ["read_every_epoch", "read_after_two", "no_transport_epoch_between"])
def test_epoch_path_length_tally_lags_one_epoch_and_is_complete_on_read(schedule):
    """Eager epochs leave their path-length records to the next epoch, which reduces them beside
    its own transport; any read of the tally first completes the latest transport's.  Against an
    engine that adds path lengths inline (same photons, low power so no cell changes phase): the
    A and L tallies read right after an epoch, after two epochs with no read between (the first
    epoch's records reduced during the second), and after an epoch without transport (which keeps
    the last transport's tally), up to the order of the float additions."""
    ed, ei = _engine_pair()
    try:
        if schedule == "read_every_epoch":
            for _ in range(3):
                ed.step_epoch()
                ei.step_epoch()
                _tallies_close(ed, ei)
        elif schedule == "read_after_two":
            for _ in range(2):
                ed.step_epoch()
                ei.step_epoch()
            assert ed.last_diag.t_mc_records_ms > 0.0  # the first epoch's passes ran in the second
            _tallies_close(ed, ei)
        else:
            ed.step_epoch()
            ei.step_epoch()
            ed.step_epoch_with_q()
            ei.step_epoch_with_q()
            _tallies_close(ed, ei)
            assert np.asarray(ed.Phi).max() > 0.0
        assert not np.any(np.asarray(ed.ablated) != np.asarray(ei.ablated))
    finally:
        ed.close()
        ei.close()


def test_path_length_records_survive_the_blast_plume_splat():
    """Regression: with a blast-expansion plume an epoch that removes cells splats the plume
    through a bounding box kept in the context's counter slots, and the record cursor of the
    epoch's second record buffer (the one every second eager epoch writes, whose passes run on
    the next read) must not share those slots.  Four epochs that remove cells, then a read of the
    tallies, against an engine adding path lengths inline (same photons)."""
    plume = {'model': 'diffusion', 'blast_expansion': True, 'D_plume': 0.5, 'v_drift': 10.0}
    ed, ei = _engine_pair(power=40.0, plume_closure=plume)
    try:
        removed = []
        for _ in range(4):
            ed.step_epoch()
            ei.step_epoch()
            removed.append(int(ed.last_diag.n_new))
        assert removed[1] > 0 and removed[3] > 0, removed  # the splat ran in both even epochs
        assert not np.any(np.asarray(ed.ablated) != np.asarray(ei.ablated))
        _tallies_close(ed, ei)
    finally:
        ed.close()
        ei.close()


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 1, 6), (3, 2, 1), (2, 5, 3)])
def test_degenerate_grids_gpu_vs_oracle(oracle, shape):
    """One voxel, one column, one plane: the entrance split, exits through every
    face and the lateral tally on grids with no interior, against the oracle."""
    from simulation.voxel_mc_3d import run_voxel_mc3d
    Nx, Ny, Nz = shape
    props = {'mu_a': np.full(shape, 0.8), 'mu_s': np.full(shape, 6.0), 'g': np.full(shape, 0.7),
             'n': np.full(shape, 1.33)}
    kw = dict(power=1.0, n_photons=40_000, Lx=0.1 * Nx, Ly=0.1 * Ny, Lz=0.1 * Nz, Nx=Nx, Ny=Ny, Nz=Nz,
              custom_3d_props=props, theta_inc_deg=20.0, spot_dia=0.05 * min(Nx, Ny), seed=31)
    g = run_voxel_mc3d(**kw)
    c = oracle.run_voxel_mc3d(**kw, n_threads=4)
    assert abs(g['energy_closure'] - 1.0) < 0.01
    for key in ('R_spec', 'R_diff', 'Absorption', 'Transmission', 'Lateral_loss'):
        assert abs(g['spectro_raw'][key] - c['spectro_raw'][key]) < 0.01, key
    assert g['A_3d'].shape == shape and np.all(g['A_3d'] >= 0.0)


@pytest.mark.parametrize("albedo", [0.5, 0.7, 0.9, 0.95, 0.99])
def test_semi_infinite_albedo_sweep_gpu(albedo):
    """V2c (verification_benchmarks.py:343-357): diffuse reflectance of a semi-infinite
    isotropic scatterer (n = 1, mu_t = 10 /mm, normal incidence) against the exact
    Chandrasekhar H-function value the reference recorded, at the reference's bar of 3
    standard errors (+2e-4 for the 30 mm depth truncation), over 20 x 100 000 photons so
    that the standard error itself is well estimated; the z-score is printed."""
    import json
    import os
    from simulation.voxel_mc_3d import run_voxel_mc3d
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "reference_engine_values.json")))["values"][f"V2c-a{albedo}"]
    mu_t = 10.0
    vals = [run_voxel_mc3d(n_photons=100000, spot_dia=0.01, Lx=60.0, Ly=60.0, Lz=30.0, Nx=6, Ny=6, Nz=30, seed=11 + s,
                           layers_config=[{'z_end': 30.0, 'mu_a': mu_t * (1 - albedo), 'mu_s': mu_t * albedo,
                                           'g': 0.0, 'n': 1.0}])['spectro_raw']['R_diff'] for s in range(20)]
    m, se = np.mean(vals), np.std(vals, ddof=1) / np.sqrt(len(vals))
    z = (m - gold["reference_value"]) / se
    print(f"V2c albedo {albedo}: Rd {m:.6f} exact {gold['reference_value']:.6f} se {se:.2e} z {z:+.2f}")
    assert abs(m - gold["reference_value"]) < 3 * se + 2e-4, f"albedo {albedo}: z = {z:+.2f} (bar 3 se + 2e-4)"


@pytest.mark.parametrize("albedo", [0.5, 0.7, 0.9, 0.99])
def test_semi_infinite_mismatched_albedo_sweep_gpu(albedo):
    """V2d (verification_benchmarks.py:406-424): total reflectance (specular + diffuse)
    of a semi-infinite isotropic scatterer with n = 1.5 against the reference's
    independent 1-D Monte Carlo (8 x 400 000 photons, its mean and standard error
    recorded), over 20 x 100 000 photons, within the reference's bar of 3 combined standard
    errors + 2e-4; the z-score is printed."""
    import json
    import os
    from simulation.voxel_mc_3d import run_voxel_mc3d
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "reference_engine_values.json")))["values"][f"V2d-a{albedo}"]
    mu_t = 10.0
    vals = []
    for s in range(20):
        r = run_voxel_mc3d(n_photons=100000, spot_dia=0.01, Lx=60.0, Ly=60.0, Lz=30.0, Nx=6, Ny=6, Nz=30,
                           seed=31 + s, layers_config=[{'z_end': 30.0, 'mu_a': mu_t * (1 - albedo),
                                                        'mu_s': mu_t * albedo, 'g': 0.0, 'n': 1.5}])
        vals.append(r['spectro_raw']['R_diff'] + r['spectro_raw']['R_spec'])
    m, se = np.mean(vals), np.std(vals, ddof=1) / np.sqrt(len(vals))
    sc = np.hypot(se, gold["series"]["se_ref"])
    z = (m - gold["reference_value"]) / sc
    print(f"V2d albedo {albedo}: R {m:.6f} reference MC {gold['reference_value']:.6f} combined se {sc:.2e} z {z:+.2f}")
    assert abs(m - gold["reference_value"]) < 3 * sc + 2e-4, f"albedo {albedo}: z = {z:+.2f} (bar 3 se + 2e-4)"
