/*
 * bd_oracle.c -- FP64 CPU restatement of the reference solver kernels.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Never linked into or called by the
 * product library.  Each function names the reference lines it follows; the
 * arithmetic is restated from the reference's CPU (numba) path, not from its
 * Metal siblings.  OpenMP parallelism mirrors the reference's prange loops.
 */
#include "bd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, k) ((((int64_t)(i)) * Ny + (j)) * (int64_t)Nz + (k))

/* ------------------------------------------------------------------------- */
/* Philox4x32-10, as published (Salmon et al. 2011, "Parallel random numbers:
 * as easy as 1, 2, 3").  Round: two 32x32->64 multiplies, xor with the other
 * counter words and the key; key bumped by the Weyl constants between rounds. */
void bdo_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double bdo_u01(uint32_t x) { return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0); }

/* Per-photon stream: key = seed, counter = (block, photon id lo, hi, tag). */
typedef struct {
    uint32_t key[2];
    uint32_t pid_lo, pid_hi;
    uint32_t block;
    uint32_t buf[4];
    int idx;
} rng_t;

#define BD_MC_STREAM_TAG 0x6d63u

static void rng_init(rng_t *r, uint64_t seed, uint64_t pid)
{
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->pid_lo = (uint32_t)pid;
    r->pid_hi = (uint32_t)(pid >> 32);
    r->block = 0;
    r->idx = 4;
}

static double rng_u(rng_t *r)
{
    if (r->idx == 4) {
        uint32_t ctr[4] = {r->block, r->pid_lo, r->pid_hi, BD_MC_STREAM_TAG};
        bdo_philox4x32_10(ctr, r->key, r->buf);
        r->block++;
        r->idx = 0;
    }
    return bdo_u01(r->buf[r->idx++]);
}

/* ------------------------------------------------------------------------- */
/* voxel_thermal_3d.py:30-112 */
void bdo_thermal_step(const double *T, double *T_new, const double *Q,
                      const double *kg, const double *rc,
                      int Nx, int Ny, int Nz, double dx, double dy, double dz,
                      double dt, double h_eff, double t_ambient, int is_active)
{
    const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy), idz2 = 1.0 / (dz * dz);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const int64_t c = IDX(i, j, k);
                const double Tc = T[c], kc = kg[c], cap = rc[c];
                double fw = 0.0, fe = 0.0, fs = 0.0, fn = 0.0, ft = 0.0, fb = 0.0;
                if (i > 0) { double kn = kg[IDX(i - 1, j, k)]; fw = 2.0 * kc * kn / (kc + kn + 1e-12) * (T[IDX(i - 1, j, k)] - Tc) * idx2; }
                if (i < Nx - 1) { double kn = kg[IDX(i + 1, j, k)]; fe = 2.0 * kc * kn / (kc + kn + 1e-12) * (T[IDX(i + 1, j, k)] - Tc) * idx2; }
                if (j > 0) { double kn = kg[IDX(i, j - 1, k)]; fs = 2.0 * kc * kn / (kc + kn + 1e-12) * (T[IDX(i, j - 1, k)] - Tc) * idy2; }
                if (j < Ny - 1) { double kn = kg[IDX(i, j + 1, k)]; fn = 2.0 * kc * kn / (kc + kn + 1e-12) * (T[IDX(i, j + 1, k)] - Tc) * idy2; }
                /* top (k = 0) and bottom (k = Nz-1): half-cell conduction in
                 * series with h_eff (voxel_thermal_3d.py:97,105) */
                const double rser = (h_eff > 0.0)
                    ? dz * (0.5 * dz / fmax(kc, 1e-30) + 1.0 / fmax(h_eff, 1e-30)) : 0.0;
                if (k > 0) { double kn = kg[IDX(i, j, k - 1)]; ft = 2.0 * kc * kn / (kc + kn + 1e-12) * (T[IDX(i, j, k - 1)] - Tc) * idz2; }
                else ft = (h_eff > 0.0) ? (t_ambient - Tc) / rser : 0.0;
                if (k < Nz - 1) { double kn = kg[IDX(i, j, k + 1)]; fb = 2.0 * kc * kn / (kc + kn + 1e-12) * (T[IDX(i, j, k + 1)] - Tc) * idz2; }
                else fb = (h_eff > 0.0) ? (t_ambient - Tc) / rser : 0.0;
                const double div_q = fw + fe + fs + fn + ft + fb;
                const double src = is_active ? Q[c] : 0.0;
                T_new[c] = Tc + dt * (div_q + src) / cap;
            }
}

/* voxel_thermal_3d.py:115-143 */
void bdo_plateau(double *T, double *H, const double *rc, const uint8_t *ablated,
                 double t_abl, int Nx, int Ny, int Nz, const uint8_t *dry, double t_void)
{
    const int64_t n = (int64_t)Nx * Ny * Nz;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        const double t = T[c];
        if (ablated[c]) {
            if (t > t_void) T[c] = t_void;
        } else if (dry[c]) {
            continue;
        } else if (t > t_abl) {
            H[c] += (t - t_abl) * rc[c];
            T[c] = t_abl;
        } else if (H[c] > 0.0) {
            const double cap = rc[c];
            const double rel = fmin(H[c], (t_abl - t) * cap);
            H[c] -= rel;
            T[c] = t + rel / cap;
        }
    }
}

/* coupled_ablation.py:33-40 */
void bdo_arrhenius(const double *T, double *Om, const uint8_t *ablated,
                   double A, double EaR, double dt, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c)
        if (!ablated[c]) Om[c] += dt * A * exp(-EaR / (T[c] + 273.15));
}

/* coupled_ablation.py:43-58 */
void bdo_char(const double *T, double *chi, const uint8_t *dry, const uint8_t *ablated,
              double A, double EaR, double dt, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c)
        if (dry[c] && !ablated[c]) {
            const double tk = T[c] + 273.15;
            const double rate = A * exp(-EaR / tk);
            chi[c] = 1.0 - (1.0 - chi[c]) * exp(-rate * dt);
        }
}

/* coupled_ablation.py:61-74 */
void bdo_single_stage(const double *H, const double *hl, const double *T, double t_abl,
                      const uint8_t *ablated, uint8_t *rem, int64_t n)
{
    for (int64_t c = 0; c < n; ++c)
        rem[c] = ablated[c] ? 0 : (uint8_t)((H[c] >= hl[c]) && (T[c] >= t_abl));
}

static inline double P(const bdo_param *p, int64_t i) { return p->n == 1 ? p->v[0] : p->v[i]; }

/* coupled_ablation.py:77-180 */
void bdo_phase_update(double *H, const double *hl, const double *T, const double *chi,
                      uint8_t *dry, const uint8_t *ablated,
                      double *mu_a, double *mu_s, double *g, double *nn,
                      double *k, double *rc,
                      const double *wet_mu_a, const double *wet_k, const double *wet_rc,
                      const bdo_phase_params *p, uint8_t *rem, int64_t nvox)
{
    for (int64_t i = 0; i < nvox; ++i) {
        if (ablated[i]) { rem[i] = 0; continue; }
        int is_dry = dry[i] != 0;
        const double h_acc = H[i];
        const double h_lat = fmax(hl[i], 1e-30);
        /* 1. boiling: wet -> dry interpolation by banked-enthalpy fraction */
        if (!is_dry) {
            if (h_acc > 0.0) {
                const double frac = fmin(fmax(h_acc / h_lat, 0.0), 1.0);
                const double dmu = P(&p->ds_mu_a, i), dk = P(&p->ds_k, i), drc = P(&p->ds_rho_cp, i);
                mu_a[i] = wet_mu_a[i] + (dmu - wet_mu_a[i]) * frac;
                k[i] = wet_k[i] + (dk - wet_k[i]) * frac;
                rc[i] = wet_rc[i] + (drc - wet_rc[i]) * frac;
                if (h_acc >= hl[i]) {
                    is_dry = 1;
                    dry[i] = 1;
                    H[i] = 0.0;
                    mu_a[i] = dmu;
                    mu_s[i] = P(&p->ds_mu_s, i);
                    g[i] = P(&p->ds_g, i);
                    nn[i] = P(&p->ds_n, i);
                    k[i] = dk;
                    rc[i] = drc;
                }
            }
        }
        /* 2. dry / char state and removal check */
        if (is_dry) {
            const double dmu = P(&p->ds_mu_a, i), dk = P(&p->ds_k, i), drc = P(&p->ds_rho_cp, i);
            if (p->has_cs) {
                const double cv = chi[i];
                if (p->has_cs_mu_a) mu_a[i] = dmu + cv * (P(&p->cs_mu_a, i) - dmu);
                if (p->has_cs_mu_s) {
                    const double dms = P(&p->ds_mu_s, i);
                    mu_s[i] = dms + cv * (P(&p->cs_mu_s, i) - dms);
                }
                if (p->has_cs_k) k[i] = dk + cv * (P(&p->cs_k, i) - dk);
                double rf;
                if (p->has_cs_rho_cp) rf = drc + cv * (P(&p->cs_rho_cp, i) - drc);
                else rf = rc[i];
                const double mn = k[i] / p->alpha_max_stable;
                if (rf < mn) rf = mn;
                rc[i] = rf;
                rem[i] = (uint8_t)((T[i] >= P(&p->cs_t_rem, i)) || (cv >= p->cs_chi_rem));
            } else {
                rem[i] = (uint8_t)(T[i] >= P(&p->ds_t_rem, i));
            }
        } else {
            rem[i] = 0;
        }
    }
}

/* plume.py:18-80 */
void bdo_plume_diffuse(const float *C, float *Cn, const uint8_t *vm,
                       int Nx, int Ny, int Nz, double dx, double dy, double dz,
                       double dt, double Dxy, double Dz, double vz, double decay)
{
    const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy), idz2 = 1.0 / (dz * dz);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const int64_t c = IDX(i, j, k);
                if (!vm[c]) { Cn[c] = 0.0f; continue; }
                if (i == 0 || i == Nx - 1 || j == 0 || j == Ny - 1 || k == 0 || k == Nz - 1) { Cn[c] = 0.0f; continue; }
                const double Cc = C[c];
                const double Cw = vm[IDX(i - 1, j, k)] ? C[IDX(i - 1, j, k)] : Cc;
                const double Ce = vm[IDX(i + 1, j, k)] ? C[IDX(i + 1, j, k)] : Cc;
                const double Cs = vm[IDX(i, j - 1, k)] ? C[IDX(i, j - 1, k)] : Cc;
                const double Cnn = vm[IDX(i, j + 1, k)] ? C[IDX(i, j + 1, k)] : Cc;
                const double Cb = vm[IDX(i, j, k - 1)] ? C[IDX(i, j, k - 1)] : Cc;
                const double Ct = vm[IDX(i, j, k + 1)] ? C[IDX(i, j, k + 1)] : Cc;
                const double diff = Dxy * ((Cw - 2.0 * Cc + Ce) * idx2 + (Cs - 2.0 * Cc + Cnn) * idy2)
                                  + Dz * ((Cb - 2.0 * Cc + Ct) * idz2);
                const double adv = (vz > 0.0) ? -vz * (Cc - Ct) / dz : 0.0;
                const double v = (Cc + dt * (diff + adv)) * decay;
                Cn[c] = (float)fmax(v, 0.0);
            }
}

/* plume.py:83-139: three 1-D passes, float32 intermediates */
void bdo_plume_convolve(float *raw, int Nx, int Ny, int Nz,
                        const float *imp, int Lx, int Ly, int Lz,
                        const float *kx, int nkx, const float *ky, int nky,
                        const float *kz, int nkz, double delta_mu,
                        int x0, int y0, int z0)
{
    const int OX = Lx + nkx - 1, OY = Ly + nky - 1;
    float *t1 = (float *)calloc((size_t)OX * Ly * Lz, sizeof(float));
    float *t2 = (float *)calloc((size_t)OX * OY * Lz, sizeof(float));
    for (int a = 0; a < nkx; ++a) {
        const float w = kx[a];
        for (int ix = 0; ix < Lx; ++ix)
            for (int y = 0; y < Ly; ++y)
                for (int z = 0; z < Lz; ++z)
                    t1[((size_t)(ix + a) * Ly + y) * Lz + z] += imp[((size_t)ix * Ly + y) * Lz + z] * w;
    }
    for (int b = 0; b < nky; ++b) {
        const float w = ky[b];
        for (int ox = 0; ox < OX; ++ox)
            for (int iy = 0; iy < Ly; ++iy)
                for (int z = 0; z < Lz; ++z)
                    t2[((size_t)ox * OY + iy + b) * Lz + z] += t1[((size_t)ox * Ly + iy) * Lz + z] * w;
    }
    for (int c = 0; c < nkz; ++c) {
        const double w = (double)kz[c] * delta_mu;
        for (int ox = 0; ox < OX; ++ox) {
            const int gx = x0 + ox;
            if (gx < 0 || gx >= Nx) continue;
            for (int oy = 0; oy < OY; ++oy) {
                const int gy = y0 + oy;
                if (gy < 0 || gy >= Ny) continue;
                for (int iz = 0; iz < Lz; ++iz) {
                    const int gz = z0 + iz + c;
                    if (gz < 0 || gz >= Nz) continue;
                    float *dst = &raw[((int64_t)gx * Ny + gy) * Nz + gz];
                    *dst = (float)((double)*dst + (double)t2[((size_t)ox * OY + oy) * Lz + iz] * w);
                }
            }
        }
    }
    free(t1);
    free(t2);
}

/* ------------------------------------------------------------------------- */
/* voxel_mc_3d.py:33-48 */
static double fresnel(double n1, double n2, double ci)
{
    if (fabs(n1 - n2) < 1e-9) return 0.0;
    ci = fmin(1.0, fabs(ci));
    if (ci > 0.999999) { const double r = (n1 - n2) / (n1 + n2); return r * r; }
    const double st2 = (n1 / n2) * (n1 / n2) * (1.0 - ci * ci);
    if (st2 >= 1.0) return 1.0;
    const double ct = sqrt(1.0 - st2);
    const double rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct);
    const double rp = (n1 * ct - n2 * ci) / (n1 * ct + n2 * ci);
    return 0.5 * (rs * rs + rp * rp);
}

/* voxel_mc_3d.py:51-69 */
static void refract(double u[3], int axis, double eta)
{
    const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    const double sgn = u[axis] > 0.0 ? 1.0 : -1.0;
    u[a1] *= eta;
    u[a2] *= eta;
    u[axis] = sgn * sqrt(fmax(0.0, 1.0 - u[a1] * u[a1] - u[a2] * u[a2]));
}

typedef struct {
    const bdo_mc_params *p;
    double init_u[3], cos_ph, sin_ph, cos_th_safe, r_max_sq;
    const double *mu_a, *mu_s, *g, *n;
} mc_ctx;

/* voxel_mc_3d.py:101-120 */
static void beam_offsets(const mc_ctx *m, rng_t *r, double *u1, double *u2)
{
    const int prof = m->p->beam_profile;
    const double R = m->p->beam_radius;
    if (prof == 1) {
        const double rr = R * sqrt(rng_u(r));
        const double a = 2.0 * M_PI * rng_u(r);
        *u1 = rr * cos(a); *u2 = rr * sin(a);
        return;
    }
    if (prof == 2 || prof == 3) {
        for (;;) {
            const double rn = sqrt(rng_u(r));
            const double acc = (prof == 2) ? exp(-2.0 * rn) : exp(-2.0 * rn * rn);
            if (rng_u(r) < acc) {
                const double a = 2.0 * M_PI * rng_u(r);
                *u1 = R * rn * cos(a); *u2 = R * rn * sin(a);
                return;
            }
        }
    }
    /* two N(0, sigma) draws: Box-Muller on one uniform pair */
    const double ua = rng_u(r), ub = rng_u(r);
    const double rad = sqrt(-2.0 * log(ua));
    *u1 = m->p->sigma_axis * rad * cos(2.0 * M_PI * ub);
    *u2 = m->p->sigma_axis * rad * sin(2.0 * M_PI * ub);
}

/* Optional per-voxel counts of the tally additions a device transport that merges consecutive
 * contributions to one voxel issues: one path-length add per voxel visit, one absorbed-weight add
 * per visit with a collision.  Off (NULL) unless bdo_mc_set_visit_grids is given two arrays;
 * used to measure how concentrated the tally traffic is. */
static double *g_visit_L = 0, *g_visit_A = 0;
void bdo_mc_set_visit_grids(double *vL, double *vA) { g_visit_L = vL; g_visit_A = vA; }
#define VISIT_FLUSH() do { if (pv >= 0 && g_visit_L) { if (hL) { _Pragma("omp atomic") g_visit_L[pv] += 1.0; } \
    if (hA) { _Pragma("omp atomic") g_visit_A[pv] += 1.0; } } pv = -1; hA = hL = 0; } while (0)
#define VISIT_L(cc) do { if ((cc) != pv) { VISIT_FLUSH(); pv = (cc); } hL = 1; } while (0)

/* voxel_mc_3d.py:123-337 */
static int64_t trace_one(const mc_ctx *m, rng_t *r, double *A, double *L, double *tally, double *lat)
{
    const bdo_mc_params *p = m->p;
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const double d[3] = {p->dx, p->dy, p->dz};
    const int la = p->launch_axis;
    double u1, u2;
    beam_offsets(m, r, &u1, &u2);
    double x[3];
    int ic[3];
    int depth_axis;
    if (la == 0) {
        depth_axis = 2;
        x[0] = p->x_center + (u1 / m->cos_th_safe) * m->cos_ph - u2 * m->sin_ph;
        x[1] = p->y_center + (u1 / m->cos_th_safe) * m->sin_ph + u2 * m->cos_ph;
        x[2] = 0.0;
        if (x[0] < 0.0 || x[0] >= p->Lx || x[1] < 0.0 || x[1] >= p->Ly ||
            (m->r_max_sq > 0.0 && (x[0] - p->x_center) * (x[0] - p->x_center) +
                                  (x[1] - p->y_center) * (x[1] - p->y_center) >= m->r_max_sq)) {
            tally[3] += 1.0; lat[0] += 1.0; return 0;
        }
        ic[0] = (int)(x[0] / p->dx); if (ic[0] > Nx - 1) ic[0] = Nx - 1;
        ic[1] = (int)(x[1] / p->dy); if (ic[1] > Ny - 1) ic[1] = Ny - 1;
        ic[2] = 0;
    } else {
        depth_axis = 0;
        x[0] = 0.0;
        x[1] = p->y_center + (u1 / m->cos_th_safe) * m->cos_ph - u2 * m->sin_ph;
        x[2] = p->z_center + (u1 / m->cos_th_safe) * m->sin_ph + u2 * m->cos_ph;
        if (x[1] < 0.0 || x[1] >= p->Ly || x[2] < 0.0 || x[2] >= p->Lz ||
            (m->r_max_sq > 0.0 && (x[1] - p->y_center) * (x[1] - p->y_center) +
                                  (x[2] - p->z_center) * (x[2] - p->z_center) >= m->r_max_sq)) {
            tally[3] += 1.0; lat[0] += 1.0; return 0;
        }
        ic[0] = 0;
        ic[1] = (int)(x[1] / p->dy); if (ic[1] > Ny - 1) ic[1] = Ny - 1;
        ic[2] = (int)(x[2] / p->dz); if (ic[2] > Nz - 1) ic[2] = Nz - 1;
    }
    double u[3] = {m->init_u[0], m->init_u[1], m->init_u[2]};

    /* entrance face: deterministic specular split + refraction (:183-192) */
    const double n_in = m->n[IDX(ic[0], ic[1], ic[2])];
    const double r_entry = fresnel(p->n_ambient, n_in, u[depth_axis]);
    tally[0] += r_entry;
    double w = 1.0 - r_entry;
    if (w <= 0.0) return 0;
    if (fabs(n_in - p->n_ambient) > 1e-9) refract(u, depth_axis, p->n_ambient / n_in);

    double tau = -log(fmax(rng_u(r), 1e-12));
    int64_t ev;
    int64_t pv = -1;
    int hA = 0, hL = 0;
    for (ev = 0; ev < p->max_events; ++ev) {
        const int64_t c = IDX(ic[0], ic[1], ic[2]);
        const double mua = m->mu_a[c];
        const double mut = mua + m->mu_s[c];
        /* _dist_to_voxel_exit (:72-98) */
        double s[3];
        for (int a = 0; a < 3; ++a) {
            if (u[a] > 1e-15) s[a] = ((ic[a] + 1) * d[a] - x[a]) / u[a];
            else if (u[a] < -1e-15) s[a] = (ic[a] * d[a] - x[a]) / u[a];
            else s[a] = 1e30;
        }
        int axis;
        if (s[0] <= s[1] && s[0] <= s[2]) axis = 0;
        else if (s[1] <= s[2]) axis = 1;
        else axis = 2;
        const double s_exit = fmax(s[axis], 0.0);
        const int sdir = u[axis] > 0.0 ? 1 : -1;

        if (mut > 1e-12 && mut * s_exit >= tau) {
            /* interaction inside the current voxel (:201-265) */
            const double s_col = tau / mut;
            L[c] += w * s_col;
            VISIT_L(c);
            for (int a = 0; a < 3; ++a) x[a] += s_col * u[a];
            if (m->r_max_sq > 0.0) {
                const double rr = (la == 0)
                    ? (x[0] - p->x_center) * (x[0] - p->x_center) + (x[1] - p->y_center) * (x[1] - p->y_center)
                    : (x[1] - p->y_center) * (x[1] - p->y_center) + (x[2] - p->z_center) * (x[2] - p->z_center);
                if (rr >= m->r_max_sq) { tally[3] += w; lat[la == 0 ? ic[2] : ic[0]] += w; VISIT_FLUSH(); return ev + 1; }
            }
            const double dw = w * (mua / mut);
            w -= dw;
            A[c] += dw;
            if (dw != 0.0) hA = 1;
            if (w < 1e-4) {
                if (rng_u(r) < 0.1) w *= 10.0;
                else { VISIT_FLUSH(); return ev + 1; }
            }
            const double gl = m->g[c];
            const double xi = rng_u(r);
            double ct;
            if (fabs(gl) < 1e-4) ct = 2.0 * xi - 1.0;
            else {
                const double tmp = (1.0 - gl * gl) / (1.0 - gl + 2.0 * gl * xi);
                ct = (1.0 + gl * gl - tmp * tmp) / (2.0 * gl);
                ct = fmax(-1.0, fmin(1.0, ct));
            }
            const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
            const double psi = 2.0 * M_PI * rng_u(r);
            const double cp = cos(psi), sp = sin(psi);
            if (fabs(u[2]) > 0.99999) {
                const double sz = u[2] > 0.0 ? 1.0 : -1.0;
                u[0] = st * cp; u[1] = st * sp; u[2] = ct * sz;
            } else {
                const double den = sqrt(fmax(1.0 - u[2] * u[2], 1e-12));
                const double nx = st * (u[0] * u[2] * cp - u[1] * sp) / den + u[0] * ct;
                const double ny = st * (u[1] * u[2] * cp + u[0] * sp) / den + u[1] * ct;
                const double nz = -st * cp * den + u[2] * ct;
                u[0] = nx; u[1] = ny; u[2] = nz;
            }
            const double nrm = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            if (nrm > 1e-12) { u[0] /= nrm; u[1] /= nrm; u[2] /= nrm; }
            tau = -log(fmax(rng_u(r), 1e-12));
            continue;
        }

        /* full chord to the exit face (:267-337) */
        tau -= mut * s_exit;
        L[c] += w * s_exit;
        VISIT_L(c);
        for (int a = 0; a < 3; ++a) x[a] += s_exit * u[a];
        int jc[3] = {ic[0], ic[1], ic[2]};
        jc[axis] += sdir;
        x[axis] = (sdir > 0) ? (ic[axis] + 1) * d[axis] : ic[axis] * d[axis];
        const int nd[3] = {Nx, Ny, Nz};
        if (axis == depth_axis) {
            const int jd = jc[depth_axis];
            if (jd < 0) {
                const double r_out = fresnel(m->n[c], p->n_ambient, u[depth_axis]);
                if (r_out > 0.0 && rng_u(r) < r_out) { u[depth_axis] = -u[depth_axis]; continue; }
                tally[1] += w;
                VISIT_FLUSH();
                return ev + 1;
            }
            if (jd >= nd[depth_axis]) { tally[2] += w; VISIT_FLUSH(); return ev + 1; }
        } else if (jc[0] < 0 || jc[0] >= Nx || jc[1] < 0 || jc[1] >= Ny || jc[2] < 0 || jc[2] >= Nz) {
            tally[3] += w; lat[la == 0 ? ic[2] : ic[0]] += w; VISIT_FLUSH(); return ev + 1;
        }
        if (m->r_max_sq > 0.0) {
            const double rr = (la == 0)
                ? (x[0] - p->x_center) * (x[0] - p->x_center) + (x[1] - p->y_center) * (x[1] - p->y_center)
                : (x[1] - p->y_center) * (x[1] - p->y_center) + (x[2] - p->z_center) * (x[2] - p->z_center);
            if (rr >= m->r_max_sq) { tally[3] += w; lat[la == 0 ? ic[2] : ic[0]] += w; VISIT_FLUSH(); return ev + 1; }
        }
        const double n1 = m->n[c], n2 = m->n[IDX(jc[0], jc[1], jc[2])];
        if (fabs(n1 - n2) > 1e-9) {
            const double r_int = fresnel(n1, n2, u[axis]);
            if (rng_u(r) < r_int) { u[axis] = -u[axis]; continue; }
            refract(u, axis, n1 / n2);
        }
        ic[0] = jc[0]; ic[1] = jc[1]; ic[2] = jc[2];
    }
    VISIT_FLUSH();
    return ev;
}

/* voxel_mc_3d.py:340-422: photons split into chunks with private tallies */
int64_t bdo_mc_trace(const bdo_mc_params *p, const double *mu_a, const double *mu_s,
                     const double *g, const double *n, double *A, double *L,
                     double *tally, double *lat)
{
    mc_ctx m;
    m.p = p;
    const double sth = sin(p->theta_rad), cth = cos(p->theta_rad);
    m.sin_ph = sin(p->phi_rad);
    m.cos_ph = cos(p->phi_rad);
    m.cos_th_safe = fmax(cth, 0.05);
    if (p->launch_axis == 0) {
        m.init_u[0] = sth * m.cos_ph; m.init_u[1] = sth * m.sin_ph; m.init_u[2] = cth;
    } else {
        m.init_u[0] = cth; m.init_u[1] = sth * m.cos_ph; m.init_u[2] = sth * m.sin_ph;
    }
    m.r_max_sq = p->r_max > 0.0 ? p->r_max * p->r_max : 0.0;
    m.mu_a = mu_a; m.mu_s = mu_s; m.g = g; m.n = n;

    const int64_t nvox = (int64_t)p->Nx * p->Ny * p->Nz;
    const int n_depth = p->launch_axis == 0 ? p->Nz : p->Nx;
    int nchunks = p->n_threads > 0 ? p->n_threads : 1;
    if (nchunks > p->n_photons) nchunks = (int)(p->n_photons > 0 ? p->n_photons : 1);
    int64_t events = 0;
    if (nchunks == 1) {
        rng_t r;
        for (int64_t q = 0; q < p->n_photons; ++q) {
            rng_init(&r, p->seed, (uint64_t)(p->photon_offset + q));
            events += trace_one(&m, &r, A, L, tally, lat);
        }
        return events;
    }
    double *Ap = (double *)calloc((size_t)nchunks * nvox, sizeof(double));
    double *Lp = (double *)calloc((size_t)nchunks * nvox, sizeof(double));
    double *tp = (double *)calloc((size_t)nchunks * 4, sizeof(double));
    double *zp = (double *)calloc((size_t)nchunks * n_depth, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1) num_threads(nchunks) reduction(+ : events)
    for (int c = 0; c < nchunks; ++c) {
        const int64_t a = (c * p->n_photons) / nchunks, b = ((c + 1) * p->n_photons) / nchunks;
        rng_t r;
        for (int64_t q = a; q < b; ++q) {
            rng_init(&r, p->seed, (uint64_t)(p->photon_offset + q));
            events += trace_one(&m, &r, Ap + (size_t)c * nvox, Lp + (size_t)c * nvox,
                                tp + 4 * c, zp + (size_t)c * n_depth);
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < nvox; ++v) {
        double sa = 0.0, sl = 0.0;
        for (int c = 0; c < nchunks; ++c) { sa += Ap[(size_t)c * nvox + v]; sl += Lp[(size_t)c * nvox + v]; }
        A[v] += sa;
        L[v] += sl;
    }
    for (int c = 0; c < nchunks; ++c) {
        for (int t = 0; t < 4; ++t) tally[t] += tp[4 * c + t];
        for (int z = 0; z < n_depth; ++z) lat[z] += zp[(size_t)c * n_depth + z];
    }
    free(Ap); free(Lp); free(tp); free(zp);
    return events;
}
