/*
 * bugdef.h -- C ABI of libbugdef.so, the MI355X (gfx950) implementation of the
 * coupled laser-tissue epoch loop: voxel Monte Carlo photon transport, the
 * explicit bioheat stencil with the enthalpy plateau, Arrhenius damage and
 * char kinetics, the boil/char/ablate property update with void conversion,
 * and void-confined plume diffusion.
 *
 * Conventions
 *  - Every call returns an int status: BD_OK (0) or a negative BD_E* code;
 *    bd_last_error() returns a message for the calling thread's last failure.
 *  - Arrays are C-order (Nx, Ny, Nz): flat index (i*Ny + j)*Nz + k, z fastest,
 *    the layout of the reference's numpy grids.
 *  - The library owns all device memory; the caller owns host buffers.
 *    Uploads and downloads are synchronous with respect to the host.
 *  - Units are the reference engine's (simulation/coupled_ablation.py:219-221):
 *    mm, 1/mm, W/(mm K), J/(mm^3 K), J/mm^3, W/mm^3, degC, s.
 *  - One host thread per context; work runs on the context's own HIP stream.
 */
#ifndef BUGDEF_H
#define BUGDEF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BD_OK 0
#define BD_EINVAL (-1)   /* bad argument: maps to ValueError */
#define BD_ENODEV (-2)   /* no HIP device: maps to RuntimeError */
#define BD_EHIP (-3)     /* HIP runtime failure: maps to RuntimeError */
#define BD_ESTATE (-4)   /* call not valid in the context's current state */

#define BD_F32 0
#define BD_F64 1
#define BD_U8 2

/* Field identifiers (device-resident grids of the context). */
enum bd_field {
    BD_FIELD_T = 0,          /* temperature [degC] (f32) */
    BD_FIELD_H_ACC,          /* banked latent enthalpy [J/mm^3] (f32) */
    BD_FIELD_OMEGA,          /* Arrhenius damage integral (f32) */
    BD_FIELD_CHI,            /* char fraction (f32) */
    BD_FIELD_Q,              /* volumetric source [W/mm^3] (f32) */
    BD_FIELD_K,              /* conductivity (f32) */
    BD_FIELD_RHO_CP,         /* volumetric heat capacity (f32) */
    BD_FIELD_H_LAT,          /* latent enthalpy to ablate (f32) */
    BD_FIELD_DRY,            /* dehydrated mask (u8) */
    BD_FIELD_ABLATED,        /* void mask (u8) */
    BD_FIELD_MU_A, BD_FIELD_MU_S, BD_FIELD_G, BD_FIELD_N,   /* optics (f32) */
    BD_FIELD_WET_MU_A, BD_FIELD_WET_K, BD_FIELD_WET_RHO_CP, /* wet copies (f32) */
    BD_FIELD_T_PEAK,         /* running max of T at epoch ends (f32) */
    BD_FIELD_ABLATION_TIME,  /* time of removal, -1 intact (f64) */
    BD_FIELD_A,              /* absorbed photon weight tally (f32) */
    BD_FIELD_L,              /* photon path-length tally (f32) */
    BD_FIELD_PLUME_RAW,      /* plume raw extinction, plume grid (f32) */
    BD_FIELD_PLUME_EXT,      /* saturated plume extinction, plume grid (f32) */
    BD_FIELD_VOID,           /* plume void mask, plume grid (u8) */
    BD_FIELD_TO_REMOVE,      /* voxels removed by the last phase update (u8) */
    /* dry-state / char-state parameters: upload n == 1 for a scalar,
     * n == Nx*Ny*Nz for a per-voxel array (coupled_ablation.py:123-125) */
    BD_FIELD_DS_MU_A, BD_FIELD_DS_MU_S, BD_FIELD_DS_G, BD_FIELD_DS_N,
    BD_FIELD_DS_K, BD_FIELD_DS_RHO_CP, BD_FIELD_DS_T_REMOVE,
    BD_FIELD_CS_MU_A, BD_FIELD_CS_MU_S, BD_FIELD_CS_K, BD_FIELD_CS_RHO_CP,
    BD_FIELD_CS_T_REMOVE,
    BD_FIELD_COUNT
};

typedef struct bd_ctx bd_ctx;

typedef struct {
    int32_t nx, ny, nz;
    double dx, dy, dz;
} bd_grid;

/* Engine scalars (coupled_ablation.py:184-311). */
typedef struct {
    double dt_sub, h_eff, t_ambient, t_abl, t_void;
    int32_t has_arrhenius;
    double arr_A, arr_Ea_over_R;
    int32_t has_char;
    double char_A, char_Ea_over_R;
    int32_t phase_mode;            /* 0: single-stage (:61-74), 1: dry state (:77-180) */
    int32_t has_cs, has_cs_mu_a, has_cs_mu_s, has_cs_k, has_cs_rho_cp;
    double cs_chi_remove, alpha_max_stable;
    double k_void, rho_cp_void;
    int32_t air_gap_k0, air_gap_k1; /* air_gap_range, or -1 -1 for none */
    int32_t no_removal;            /* 1: bd_phase_update only writes TO_REMOVE (the
                                      module kernels); 0: also applies _remove */
} bd_physics;

/* Plume closure (plume.py:142-268).  The plume grid may equal the thermal
 * grid (the reference layout) or be a coarser grid covering the same box. */
typedef struct {
    int32_t enabled;
    int32_t model;                 /* 0: diffusion, 1: separable (legacy) */
    int32_t nx, ny, nz;            /* plume grid */
    double D_xy, D_z, v_drift, tau_linger;
    double mu_max, mu_a_max, mu_s_max, delta_mu;
    int32_t blast_expansion;
    int32_t dt_clamp;              /* 1: diffusion step min(dt, 0.45 / (2 (D_xy (1/dx^2 + 1/dy^2) + D_z/dz^2))),
                                      the step of the reference's Metal solver (metal_ablation_solver.py:234-236);
                                      0: dt, the CPU scheme (plume.py:18-80) */
} bd_plume_params;

/* Beam and run controls for one Monte Carlo call (voxel_mc_3d.py:425-445). */
typedef struct {
    double power, spot_dia, theta_rad, phi_rad;
    double x_center, y_center, z_center, n_ambient, r_max;
    int32_t beam_profile;          /* 0 gaussian, 1 tophat, 2 kim_exp_disk, 3 kim_gauss_disk */
    int32_t launch_axis;           /* 0: +z from z=0, 1: +x from x=0 */
    int64_t n_photons, photon_offset, max_events;
    uint64_t seed;
    int32_t score_fluence;         /* 1: also tally path length (Phi) */
    int32_t add_plume;             /* 1: add plume extinction to the optics */
} bd_beam;

typedef struct {
    double specular, diffuse, transmitted, lateral, absorbed; /* weight sums */
    double events;                 /* transport events traced */
    double kernel_ms;              /* device time of the transport kernel */
    double tally_atomics;          /* float atomic adds the transport kernel issued into the A/L voxel tallies */
    double tally_records;          /* path-length additions the kernel deferred to the record pass */
    double record_ms;              /* device time of the record pass that adds them (0: none deferred) */
} bd_mc_result;

/* Per-epoch diagnostics (coupled_ablation.py:605-617) plus FP64 ledgers. */
typedef struct {
    int64_t n_new;
    double absorbed_frac, mc_energy_closure;
    double plume_tau, plume_transmission, plume_peak_mu_a, plume_mean_mu_a;
    double t_optical_ms, t_thermal_ms, t_phase_ms;
    double absorbed_J, reflected_J, transmitted_J, lateral_J, conducted_J, ablated_J;
    double R_spec, R_diff, transmission, lateral_loss;  /* MC fractions (voxel_mc_3d.py:665-670) */
    double mc_events;
    double t_substeps_ms;          /* device time of the fused thermal sub-step launches alone */
    double t_mc_kernel_ms;         /* device time of the transport kernel alone */
    double mc_tally_atomics;       /* float atomic adds the transport kernel issued into the A/L tallies */
    double mc_tally_records;       /* path-length additions deferred to the record pass */
    double t_mc_records_ms;        /* device time of the record pass */
} bd_epoch_diag;

/* Device / library */
const char *bd_last_error(void);
int bd_device_count(int *count);
int bd_version(int *major, int *minor);

/* Context lifetime: replaces MetalAblationContext.__init__
 * (simulation/metal_ablation_solver.py:290-379) and the engine's array
 * allocation (coupled_ablation.py:272-307). */
int bd_create(const bd_grid *grid, bd_ctx **out);
int bd_destroy(bd_ctx *ctx);
int bd_set_physics(bd_ctx *ctx, const bd_physics *p);
int bd_set_plume(bd_ctx *ctx, const bd_plume_params *p);
/* Separable splat kernels (plume.py:238-268); which = 0 blast, 1 legacy. */
int bd_set_plume_kernels(bd_ctx *ctx, int which, const float *kx, int nkx, const float *ky, int nky,
                         const float *kz, int nkz, int rx, int ry, int rz);

/* Host <-> device field copies: replace MetalAblationContext.set_state /
 * get_state / get_plume (metal_ablation_solver.py:381-455) and the Metal MC
 * buffer copies (voxel_mc_metal.py:487-506, 535-586). */
int bd_upload(bd_ctx *ctx, int field, int dtype, const void *host, int64_t n);
int bd_download(bd_ctx *ctx, int field, int dtype, void *host, int64_t n);
int bd_fill(bd_ctx *ctx, int field, double value);
/* Sub-volume copy of one field: an x-z plane at j (plane = 1) or a y-z plane
 * at i (plane = 0), every `stride`-th point, as f32 (run_highres_knowledge_suite.py:271-306). */
int bd_download_slice(bd_ctx *ctx, int field, int plane, int index, int stride, float *host, int64_t n);

/* Thermal sub-steps: fused _thermal3d_step_explicit + _apply_latent_heat_plateau
 * + _accumulate_arrhenius + _accumulate_char (voxel_thermal_3d.py:30-143,
 * coupled_ablation.py:33-58, 515-541), plume diffusion included when enabled;
 * replaces MetalAblationContext.advance_substeps (metal_ablation_solver.py:399-437).
 * flags: bit0 plateau, bit1 arrhenius, bit2 char, bit3 zero Q in ablated
 * cells, bit4 track T_peak after the last sub-step, bit5 plume diffusion. */
#define BD_SUB_PLATEAU 1
#define BD_SUB_ARRHENIUS 2
#define BD_SUB_CHAR 4
#define BD_SUB_MASK_Q 8
#define BD_SUB_PEAK 16
#define BD_SUB_PLUME 32
int bd_substeps(bd_ctx *ctx, int n_substeps, int is_active, int flags);

/* Stand-alone module kernels on the context's fields, for callers of the
 * reference's module functions (same arithmetic as bd_substeps). */
int bd_plateau(bd_ctx *ctx);                                      /* voxel_thermal_3d.py:115-143 */
int bd_arrhenius(bd_ctx *ctx, double A, double Ea_over_R, double dt); /* coupled_ablation.py:33-40 */
int bd_char(bd_ctx *ctx, double A, double Ea_over_R, double dt);      /* coupled_ablation.py:43-58 */
int bd_plume_substeps(bd_ctx *ctx, int n, double dt, double decay);   /* plume.py:18-80, 333-350 */

/* Monte Carlo transport: replaces run_voxel_mc_metal / _mc3d_trace_photons
 * (voxel_mc_metal.py:388-586, voxel_mc_3d.py:340-422).  Tallies A and L are
 * zeroed first; Q is set to A * power / (n_photons * voxel volume). */
int bd_mc_run(bd_ctx *ctx, const bd_beam *beam, bd_mc_result *out);
/* bd_mc_run queued on the context's stream without waiting for it; the tallies
 * and Q are ready for the context's later calls.  No reference counterpart: the
 * surrogate loop runs its low-photon transport beside the network's forward pass. */
int bd_mc_launch(bd_ctx *ctx, const bd_beam *beam);
/* Lateral escape per depth index (voxel_mc_3d.py:697), n_depth values. */
int bd_mc_lateral(bd_ctx *ctx, double *host, int64_t n_depth);

/* Measurement helper, no counterpart in the reference: the ceiling for the
 * transport's tally atomics.  One launch of n_pairs independent float atomic
 * add pairs (A and L of one voxel) at voxel indices hashed uniformly over
 * [0, span) of a float2 grid of n_voxels, on device `device`; writes the
 * launch's device time.  n_pairs is a multiple of 4096 (at most 2^32). */
int bd_probe_atomic_rate(int device, int64_t n_voxels, int64_t span, int64_t n_pairs, double *kernel_ms);

/* Measurement helper, no counterpart in the reference: the achievable HBM rate the stencils
 * are set against.  A streaming float4 copy of `bytes` (a multiple of 16) from one device buffer
 * into another on device `device`, timed in six launch shapes (each after an untimed launch);
 * writes the fastest launch's device time.  The copy moves 2 * bytes through HBM. */
int bd_probe_copy_rate(int device, int64_t bytes, double *kernel_ms);

/* Allocation-free transport.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  The photon transport of bd_mc_run (voxel_mc_3d.py:340-422) on caller-owned
 * device buffers, with no hipMalloc, hipFree or allocating library call:
 *   optics  device pointer, Nx*Ny*Nz float4 {mu_a, mu_s, g, n} per voxel, C order;
 *   AL      device pointer, Nx*Ny*Nz float2 {absorbed weight, path length}, cleared by the call;
 *   stream  a hipStream_t, or NULL for the null stream; the call waits for it before returning;
 *   out     the scalar tallies of bd_mc_run (kernel_ms is not measured and reads -1);
 *   lateral host array of n_depth doubles (n_depth = Nz for launch axis z, Nx for x), or NULL.
 * Its scratch (photon work counter, FP64 scalar tallies, lateral loss per depth) is a static
 * device array of 16 + BD_NOALLOC_MAX_DEPTH doubles that each GPU the library runs on holds until
 * the process ends; a depth axis longer than BD_NOALLOC_MAX_DEPTH cells is rejected (BD_EINVAL).
 * Calls on one GPU share that array, so calls that overlap in time on one GPU race. */
#define BD_NOALLOC_MAX_DEPTH 4096
int bd_mc_run_noalloc(const bd_grid *grid, const bd_beam *beam, const void *optics, void *AL, void *stream,
                      bd_mc_result *out, double *lateral, int64_t n_depth);

/* Phase update + removal: _update_phase_properties_numba or
 * _update_single_stage_ablation_numba, then _remove with air-gap perforation
 * (coupled_ablation.py:61-180, 551-583, 628-654).  t_now is the engine time
 * written into ablation_time. */
int bd_phase_update(bd_ctx *ctx, double t_now, int64_t *n_new);

/* Plume bookkeeping (plume.py:292-366, coupled_ablation.py:587-603, 619-626):
 * void mask from the ablated field (with air-gap breach), injection of the
 * voxels removed by the last phase update, tanh saturation, legacy update,
 * centre-column optical depth and peak/mean diagnostics. */
int bd_plume_void(bd_ctx *ctx);
int bd_plume_inject(bd_ctx *ctx);
int bd_plume_sync(bd_ctx *ctx);
int bd_plume_update_separable(bd_ctx *ctx, double dt_epoch);
/* _convolve_separable_plume_numba (plume.py:83-139) on the plume raw field with
 * an explicit float32 impulse box (Lx,Ly,Lz) and output origin (x0,y0,z0);
 * which = 0 blast taps, 1 legacy taps (bd_set_plume_kernels). */
int bd_plume_splat(bd_ctx *ctx, int which, const float *impulse, int Lx, int Ly, int Lz, int x0, int y0, int z0,
                   double delta_mu);
int bd_plume_diag(bd_ctx *ctx, double out[4]); /* tau, transmission, peak, mean */

/* One whole epoch (coupled_ablation.py:440-617): MC when beam->power > 0,
 * n_substeps fused sub-steps, T_peak, phase update and removal, plume
 * injection/sync and diagnostics, with FP64 energy ledgers. */
int bd_epoch(bd_ctx *ctx, const bd_beam *beam, int n_substeps, double t_after, bd_epoch_diag *out);

/* The same epoch replayed from a captured hipGraph: no host wait inside the
 * epoch, one parameter upload and one diagnostics readback per call.  The MC
 * key is SplitMix64(seed_base + (epoch + 1) * 0x9E3779B97F4A7C15) and the
 * removal time t_before + n_substeps * dt_sub, so a run stepped with this call
 * equals one stepped with bd_epoch at beam->seed = that key.  Requires an even
 * n_substeps and a diffusion plume without blast injection (or no plume); the
 * graph is re-captured when the beam or n_substeps change.  Per-phase timings
 * are not recorded: t_thermal_ms holds the whole epoch. */
int bd_epoch_graph(bd_ctx *ctx, const bd_beam *beam, int n_substeps, double t_before, int64_t epoch,
                   uint64_t seed_base, bd_epoch_diag *out);

/* bd_epoch_graph split in two so several contexts (one stream each) can have
 * epochs in flight at once: _launch refreshes the device parameters and
 * replays the graph without waiting; _finish waits and fills the diagnostics.
 * A context takes no other call between the two. */
int bd_epoch_graph_launch(bd_ctx *ctx, const bd_beam *beam, int n_substeps, double t_before, int64_t epoch,
                          uint64_t seed_base);
int bd_epoch_graph_finish(bd_ctx *ctx, bd_epoch_diag *out);

/* One epoch of n contexts (a campaign) with all their photon transport in a
 * single kernel launch: each context's optics pack and tally clears run on its
 * own stream, one launch traces every context's photons (each block traces
 * photons of one context, so the tails of long-photon runs fill with other
 * runs' blocks), and each context's sub-steps, phase update and plume follow
 * on its own stream.  Same per-context inputs as bd_epoch_graph_launch
 * (arrays of n); collect each context with bd_epoch_graph_finish.  The traced
 * photons are those of n separate epochs; only the order of the float tally
 * additions differs.  No reference counterpart: the reference runs a campaign
 * one run after another (run_multispectral_matrix.py). */
int bd_epoch_gang_launch(bd_ctx *const *ctxs, int n, const bd_beam *beams, int n_substeps, const double *t_before,
                         const int64_t *epochs, const uint64_t *seed_bases);

/* One epoch of n contexts (a campaign) as one graph launch: each context's
 * captured epoch graph (as bd_epoch_graph_launch captures it) is an
 * independent child node of one parent graph, launched once on a library
 * stream of the current device after every context's parameter upload; each
 * context's stream then continues after it.  The parent graph is rebuilt only
 * when the contexts or one of their epoch graphs changed.  Same per-context
 * inputs as bd_epoch_graph_launch (arrays of n, contexts on the current
 * device); collect each context with bd_epoch_graph_finish.  Batch calls on
 * one device must not overlap in time.  No reference counterpart: the
 * reference runs a campaign one run after another (run_multispectral_matrix.py). */
int bd_epoch_batch_launch(bd_ctx *const *ctxs, int n, const bd_beam *beams, int n_substeps, const double *t_before,
                          const int64_t *epochs, const uint64_t *seed_bases);

/* Reductions (FP64): op 0 sum, 1 max, 2 min over a field; optional u8 mask
 * field (-1 for none) restricts the reduction to mask != 0. */
int bd_reduce(bd_ctx *ctx, int field, int op, int mask_field, double *out);

/* compute_phase_map (run_multispectral_matrix.py:168-192) on the device;
 * is_air_gap may be NULL.  out is uint8 (Nx,Ny,Nz). */
int bd_phase_map(bd_ctx *ctx, const uint8_t *is_air_gap, uint8_t *out, int64_t n);

/* Damage-dependent scattering of the scan benchmark's DOP option
 * (validation_benchmarks.py:116-121): for every voxel not ablated,
 * alpha = 1 - exp(-Omega), mu_s = ((1 - alpha) musp_native + alpha musp_coag) / (1 - g),
 * in FP64 on the stored Omega.  BD_EINVAL unless the parameters are finite and
 * g < 1; BD_ESTATE before mu_s, Omega and the masks are on the device. */
int bd_damage_optics(bd_ctx *ctx, double musp_native, double musp_coag, double g);

/* compute_phase_map on one plane (plane/index/stride as bd_download_slice),
 * with the air gap given as a per-z profile gap_z[Nz] (may be NULL); this is
 * what the per-epoch recorders need (run_highres_knowledge_suite.py:271-283)
 * without moving the 3-D grids.  out is uint8, n values. */
int bd_phase_slice(bd_ctx *ctx, int plane, int index, int stride, const uint8_t *gap_z, uint8_t *out, int64_t n);

/* bd_epoch with the absorbed-power field already in Q instead of a transport
 * step (coupled_ablation.py:502-617 from the sub-steps on): used when a
 * surrogate supplies Q.  absorbed_J / absorbed_frac come from Q; reflected,
 * transmitted and lateral are zero and mc_energy_closure is NaN. */
int bd_epoch_q(bd_ctx *ctx, const bd_beam *beam, int n_substeps, double t_after, bd_epoch_diag *out);

/* Launch shapes the context's autotuner picked (no reference counterpart; for
 * reports): out[0..5] = sub-step tile rows (y), tile planes (x), pair-kernel
 * tile rows, pair-kernel tile planes, pairs of sub-steps fused in one launch
 * (1/0), tuned yet (1/0).  n is the capacity of out (>= 6). */
int bd_get_tuning(bd_ctx *ctx, int32_t *out, int n);

/* Wait for all queued work of the context. */
int bd_sync(bd_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* BUGDEF_H */
