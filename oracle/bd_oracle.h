/*
 * bd_oracle.h -- FP64 CPU restatement of the reference solver kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP kernels are
 * compared against; it is never linked into, loaded by or called from the
 * product library (bug-defense_amd/).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.
 *
 * Every function cites the reference function it restates.  Arrays are
 * C-order (Nx, Ny, Nz): the flat index of (i, j, k) is (i*Ny + j)*Nz + k, so
 * z is the contiguous axis, exactly as the reference's numpy arrays.
 */
#ifndef BD_ORACLE_H
#define BD_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11). */
void bdo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* simulation/voxel_thermal_3d.py:30-112 (_thermal3d_step_explicit) */
void bdo_thermal_step(const double *T, double *T_new, const double *Q,
                      const double *k, const double *rho_cp,
                      int Nx, int Ny, int Nz, double dx, double dy, double dz,
                      double dt, double h_eff, double t_ambient, int is_active);

/* simulation/voxel_thermal_3d.py:115-143 (_apply_latent_heat_plateau) */
void bdo_plateau(double *T, double *H_acc, const double *rho_cp,
                 const uint8_t *ablated, double t_abl, int Nx, int Ny, int Nz,
                 const uint8_t *dry, double t_void);

/* simulation/coupled_ablation.py:33-40 (_accumulate_arrhenius) */
void bdo_arrhenius(const double *T, double *Omega, const uint8_t *ablated,
                   double A, double Ea_over_R, double dt, int64_t n);

/* simulation/coupled_ablation.py:43-58 (_accumulate_char) */
void bdo_char(const double *T, double *chi, const uint8_t *dry,
              const uint8_t *ablated, double A, double Ea_over_R, double dt,
              int64_t n);

/* simulation/coupled_ablation.py:61-74 (_update_single_stage_ablation_numba) */
void bdo_single_stage(const double *H_acc, const double *h_lat, const double *T,
                      double t_abl, const uint8_t *ablated, uint8_t *to_remove,
                      int64_t n);

/* Per-voxel-or-scalar parameter: n == 1 broadcasts v[0]
 * (simulation/coupled_ablation.py:123-125 "x[0] if x.size == 1 else x[i]"). */
typedef struct { const double *v; int64_t n; } bdo_param;

/* simulation/coupled_ablation.py:77-180 (_update_phase_properties_numba) */
typedef struct {
    bdo_param ds_mu_a, ds_mu_s, ds_g, ds_n, ds_k, ds_rho_cp, ds_t_rem;
    int has_cs, has_cs_mu_a, has_cs_mu_s, has_cs_k, has_cs_rho_cp;
    bdo_param cs_mu_a, cs_mu_s, cs_k, cs_rho_cp, cs_t_rem;
    double cs_chi_rem;
    double alpha_max_stable;
} bdo_phase_params;

void bdo_phase_update(double *H_acc, const double *h_lat, const double *T,
                      const double *chi, uint8_t *dry, const uint8_t *ablated,
                      double *mu_a, double *mu_s, double *g, double *n,
                      double *k, double *rho_cp,
                      const double *wet_mu_a, const double *wet_k,
                      const double *wet_rho_cp, const bdo_phase_params *p,
                      uint8_t *to_remove, int64_t nvox);

/* simulation/plume.py:18-80 (_diffuse_plume_step_numba); float32 field,
 * float64 arithmetic, as the reference's numba typing gives. */
void bdo_plume_diffuse(const float *C, float *C_new, const uint8_t *void_mask,
                       int Nx, int Ny, int Nz, double dx, double dy, double dz,
                       double dt, double D_xy, double D_z, double vz,
                       double decay);

/* simulation/plume.py:83-139 (_convolve_separable_plume_numba) */
void bdo_plume_convolve(float *raw, int Nx, int Ny, int Nz,
                        const float *impulse, int Lx, int Ly, int Lz,
                        const float *kx, int nkx, const float *ky, int nky,
                        const float *kz, int nkz, double delta_mu,
                        int x0_out, int y0_out, int z0_out);

/* simulation/voxel_mc_3d.py:123-422 (_trace_one_photon, _mc3d_trace_photons) */
typedef struct {
    int Nx, Ny, Nz;
    double dx, dy, dz, Lx, Ly, Lz;
    double sigma_axis;      /* w0/2 (voxel_mc_3d.py:617-618) */
    double beam_radius;     /* w0 */
    int beam_profile;       /* 0 gaussian, 1 tophat, 2 kim_exp_disk, 3 kim_gauss_disk */
    double theta_rad, phi_rad;
    double x_center, y_center, z_center;
    double n_ambient;
    double r_max;
    int launch_axis;        /* 0: launch at z=0 along +z; 1: at x=0 along +x */
    int64_t n_photons;
    int64_t photon_offset;  /* first photon id (RNG stream index) */
    uint64_t seed;
    int64_t max_events;     /* voxel_mc_3d.py:30 (_MAX_EVENTS_PER_PHOTON) */
    int n_threads;          /* OpenMP threads (chunks with private tallies) */
} bdo_mc_params;

/* A, L: (Nx,Ny,Nz) accumulated photon weight / path-length tallies;
 * tally[4] = specular, diffuse, transmitted, lateral;
 * lat[n_depth] = lateral escape per depth index (Nz for launch_axis 0,
 * Nx for launch_axis 1).  Outputs are ADDED to (caller zeroes them).
 * Returns the total number of transport events traced. */
/* optional per-voxel counts of merged tally additions (path-length, absorbed); NULL disables */
void bdo_mc_set_visit_grids(double *visit_L, double *visit_A);
int64_t bdo_mc_trace(const bdo_mc_params *p, const double *mu_a,
                     const double *mu_s, const double *g, const double *n,
                     double *A, double *L, double *tally, double *lat);

/* Uniform (0,1) draw used by both the oracle and the device transport:
 * 24 high bits of a Philox word, centred in its bin. */
double bdo_u01(uint32_t x);

#ifdef __cplusplus
}
#endif
#endif
