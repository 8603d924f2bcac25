// bd_phase.hip -- once-per-epoch tissue-state update for gfx950.
//
// One pass over the grid fuses
//   coupled_ablation.py:61-74    (_update_single_stage_ablation_numba)   mode 0
//   coupled_ablation.py:77-180   (_update_phase_properties_numba)        mode 1
//   coupled_ablation.py:628-643  (_remove: void conversion of the cells just
//                                 marked, applied to the same cell in place)
// Both steps touch only the voxel itself, so fusing them is exact.  The
// air-gap perforation (coupled_ablation.py:645-654) needs the whole column and
// runs as a second, column-parallel kernel.  Threshold comparisons are done in
// FP64 on the stored values, so the masks are identical to an FP64 evaluation
// of the same stored fields.
#include "bd_internal.hpp"

#include <cmath>

namespace bd {

struct PParam {
    const float *f;   // f32 parameter (mu_a, mu_s, g, n, k, rho_cp)
    const double *d;  // f64 parameter (t_remove)
    int per_voxel;
};

struct PhaseArgs {
    float *__restrict__ H;
    const float *__restrict__ hlat;
    float *__restrict__ T;
    float *__restrict__ chi;
    uint8_t *__restrict__ flags;
    float *__restrict__ mu_a, *__restrict__ mu_s, *__restrict__ g, *__restrict__ n;
    float *__restrict__ k, *__restrict__ rc;
    const float *__restrict__ wet_mu_a, *__restrict__ wet_k, *__restrict__ wet_rc;
    double *__restrict__ abl_time;
    uint8_t *__restrict__ to_remove;
    unsigned long long *__restrict__ n_new;
    double *__restrict__ abl_energy;
    PParam ds[7];
    PParam cs[5];
    int mode, has_cs, has_cs_mu_a, has_cs_mu_s, has_cs_k, has_cs_rc, no_removal;
    double cs_chi_rem, alpha_max, t_abl, t_now, t_amb;
    float k_void, rc_void, t_void;
    double vol;
    int64_t nvox;
    const EpochParams *ep;  // when set, t_now = ep->t + ep->dt_epoch (device-resident epoch time)
};

__device__ __forceinline__ float pf(const PParam &p, int64_t i) { return p.per_voxel ? p.f[i] : p.f[0]; }
__device__ __forceinline__ double pd(const PParam &p, int64_t i) { return p.per_voxel ? p.d[i] : p.d[0]; }

// One intact cell i (flags fl without F_ABL) with its temperature T, banked
// enthalpy h (0 unless F_HPOS) and latent heat hl_raw already loaded: the phase
// update and, when the cell is marked, its conversion to void.  Returns the new
// flags; rem says whether the cell was marked, e_abl gains its removed energy.
__device__ __forceinline__ uint8_t phase_cell(const PhaseArgs &a, int64_t i, uint8_t fl, float T, float h,
                                              float hl_raw, double t_now, bool &rem, double &e_abl)
{
    rem = false;
    if (a.mode == 0) {
        rem = ((double)h >= (double)hl_raw) && ((double)T >= a.t_abl);
    } else {
        bool is_dry = (fl & F_DRY) != 0;
        if (!is_dry && h > 0.f) {
            const float hl = fmaxf(hl_raw, 1e-30f);
            const float frac = fminf(fmaxf(h / hl, 0.f), 1.f);
            const float dmu = pf(a.ds[0], i), dk = pf(a.ds[4], i), drc = pf(a.ds[5], i);
            const float wmu = a.wet_mu_a[i], wk = a.wet_k[i], wrc = a.wet_rc[i];
            a.mu_a[i] = wmu + (dmu - wmu) * frac;
            a.k[i] = wk + (dk - wk) * frac;
            a.rc[i] = wrc + (drc - wrc) * frac;
            if ((double)h >= (double)hl_raw) {
                is_dry = true;
                fl = (uint8_t)((fl | F_DRY) & ~F_HPOS);
                a.H[i] = 0.f;
                h = 0.f;
                a.mu_a[i] = dmu;
                a.mu_s[i] = pf(a.ds[1], i);
                a.g[i] = pf(a.ds[2], i);
                a.n[i] = pf(a.ds[3], i);
                a.k[i] = dk;
                a.rc[i] = drc;
            }
        }
        if (is_dry) {
            if (a.has_cs) {
                const float dmu = pf(a.ds[0], i), dk = pf(a.ds[4], i), drc = pf(a.ds[5], i);
                const float cv = a.chi[i];
                if (a.has_cs_mu_a) a.mu_a[i] = dmu + cv * (pf(a.cs[0], i) - dmu);
                if (a.has_cs_mu_s) {
                    const float dms = pf(a.ds[1], i);
                    a.mu_s[i] = dms + cv * (pf(a.cs[1], i) - dms);
                }
                float kk;
                if (a.has_cs_k) {
                    kk = dk + cv * (pf(a.cs[2], i) - dk);
                    a.k[i] = kk;
                } else {
                    kk = a.k[i];
                }
                float rf = a.has_cs_rc ? drc + cv * (pf(a.cs[3], i) - drc) : a.rc[i];
                const float mn = (float)((double)kk / a.alpha_max);
                if (rf < mn) rf = mn;
                a.rc[i] = rf;
                rem = ((double)T >= pd(a.cs[4], i)) || ((double)cv >= a.cs_chi_rem);
            } else {
                rem = (double)T >= pd(a.ds[6], i);
            }
        }
    }
    if (rem && !a.no_removal) {
        // _remove (coupled_ablation.py:632-643) on this cell
        e_abl += a.vol * ((double)a.rc[i] * ((double)T - a.t_amb) + (double)h);
        fl = (uint8_t)((fl | F_ABL) & ~(F_DRY | F_HPOS));
        a.chi[i] = 0.f;
        if (a.abl_time) a.abl_time[i] = t_now;
        a.mu_a[i] = 0.f;
        a.mu_s[i] = 0.f;
        a.g[i] = 0.f;
        a.n[i] = 1.f;
        a.k[i] = a.k_void;
        a.rc[i] = a.rc_void;
        a.T[i] = a.t_void;
        a.H[i] = 0.f;
    }
    return fl;
}

// Four consecutive cells per thread and iteration: their flags, T and latent heat
// come in one vector load each, issued together before any branch (the banked
// enthalpy only where a cell has some), so a quad waits on memory once.  The
// cells are then updated in order by phase_cell, and the new flags and removal
// marks go out as one 4-byte store each.  Ablated cells keep their flags and get
// a zero mark.  The last nvox % 4 cells are taken one at a time by the first
// threads of block 0.
__global__ __launch_bounds__(256) void k_phase(PhaseArgs a)
{
    unsigned cnt = 0;
    double e_abl = 0.0;
    const double t_now = a.ep ? a.ep->t + a.ep->dt_epoch : a.t_now;
    const int64_t nq = a.nvox >> 2;
    const uint32_t *fl4 = (const uint32_t *)a.flags;
    const float4 *T4 = (const float4 *)a.T;
    const float4 *hl4 = (const float4 *)a.hlat;
    const float4 *H4 = (const float4 *)a.H;
    const uint32_t hpos4 = (uint32_t)F_HPOS * 0x01010101u, abl4 = (uint32_t)F_ABL * 0x01010101u;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nq; q += (int64_t)gridDim.x * blockDim.x) {
        // the three loads are issued together, ahead of the branch on the flags (the empty
        // asm takes their results, so the compiler cannot sink the T and latent-heat loads
        // behind that branch)
        uint32_t fw = fl4[q];
        float4 tq = T4[q];
        float4 lq = hl4[q];
        asm volatile("" : "+v"(fw), "+v"(tq.x), "+v"(lq.x));
        if ((fw & abl4) == abl4) {  // four ablated cells
            ((uint32_t *)a.to_remove)[q] = 0u;
            continue;
        }
        float4 hq = make_float4(0.f, 0.f, 0.f, 0.f);  // banked enthalpy, where a cell has some
        if (fw & hpos4) hq = H4[q];
        const float tv[4] = {tq.x, tq.y, tq.z, tq.w};
        const float lv[4] = {lq.x, lq.y, lq.z, lq.w};
        const float hv[4] = {hq.x, hq.y, hq.z, hq.w};
        uint32_t fo = 0u, ro = 0u;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            uint8_t fl = (uint8_t)(fw >> (8 * e));
            bool rem = false;
            if (!(fl & F_ABL)) {
                fl = phase_cell(a, 4 * q + e, fl, tv[e], (fl & F_HPOS) ? hv[e] : 0.f, lv[e], t_now, rem, e_abl);
                if (rem) ++cnt;
            }
            fo |= (uint32_t)fl << (8 * e);
            ro |= (rem ? 1u : 0u) << (8 * e);
        }
        ((uint32_t *)a.flags)[q] = fo;
        ((uint32_t *)a.to_remove)[q] = ro;
    }
    if (blockIdx.x == 0 && threadIdx.x < (a.nvox & 3)) {
        const int64_t i = 4 * nq + threadIdx.x;
        const uint8_t fl = a.flags[i];
        bool rem = false;
        if (!(fl & F_ABL)) {
            const uint8_t f2 = phase_cell(a, i, fl, a.T[i], (fl & F_HPOS) ? a.H[i] : 0.f, a.hlat[i], t_now, rem,
                                          e_abl);
            if (rem) ++cnt;
            a.flags[i] = f2;
        }
        a.to_remove[i] = rem ? 1 : 0;
    }
    // wave-level count + energy, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        cnt += __shfl_down(cnt, off);
        e_abl += __shfl_down(e_abl, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (cnt) atomicAdd(a.n_new, (unsigned long long)cnt);
        if (e_abl != 0.0) unsafeAtomicAdd(a.abl_energy, e_abl);
    }
}

// coupled_ablation.py:645-654: columns whose cell k0-1 is ablated while k0 is
// not get [k0, k1) marked ablated at time t (only when cells were removed).
__global__ void k_perforate(uint8_t *__restrict__ flags, double *__restrict__ abl_time,
                            const unsigned long long *__restrict__ n_new, int Nx, int Ny, int Nz, int k0, int k1,
                            double t_arg, const EpochParams *ep)
{
    if (*n_new == 0) return;
    const double t_now = ep ? ep->t + ep->dt_epoch : t_arg;
    const int64_t cols = (int64_t)Nx * Ny;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < cols; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t base = c * Nz;
        if ((flags[base + k0 - 1] & F_ABL) && !(flags[base + k0] & F_ABL)) {
            for (int kk = k0; kk < k1; ++kk) {
                flags[base + kk] = (uint8_t)(flags[base + kk] | F_ABL);
                if (abl_time) abl_time[base + kk] = t_now;
            }
        }
    }
}

int launch_phase_enqueue(Ctx *c, double t_now, const EpochParams *ep)
{
    const bd_physics &p = c->phys;
    const size_t f = (size_t)c->nvox * 4;
    BD_TRY(ensure(c, (void **)&c->to_remove, (size_t)c->nvox));
    BD_TRY(ensure(c, (void **)&c->hlat, f));
    BD_TRY(ensure(c, (void **)&c->mu_a, f));
    BD_TRY(ensure(c, (void **)&c->mu_s, f));
    BD_TRY(ensure(c, (void **)&c->gg, f));
    BD_TRY(ensure(c, (void **)&c->nn, f));
    PhaseArgs a{};
    a.mode = p.phase_mode;
    a.no_removal = p.no_removal;
    if (a.mode == 1) {
        for (int q = 0; q < 7; ++q) {
            if (!c->ds[q].ptr) {
                set_error("bd_phase_update: dry-state parameters not uploaded");
                return BD_ESTATE;
            }
            a.ds[q].f = q < 6 ? (const float *)c->ds[q].ptr : nullptr;
            a.ds[q].d = q == 6 ? (const double *)c->ds[q].ptr : nullptr;
            a.ds[q].per_voxel = c->ds[q].n > 1;
        }
        if (!c->wet_mu_a || !c->wet_k || !c->wet_rc) {
            set_error("bd_phase_update: wet copies (mu_a, k, rho_cp) not uploaded");
            return BD_ESTATE;
        }
        if (p.has_cs) {
            const int need[5] = {p.has_cs_mu_a, p.has_cs_mu_s, p.has_cs_k, p.has_cs_rho_cp, 1};
            for (int q = 0; q < 5; ++q) {
                if (need[q] && !c->cs[q].ptr) {
                    set_error("bd_phase_update: char-state parameters not uploaded");
                    return BD_ESTATE;
                }
                a.cs[q].f = q < 4 ? (const float *)c->cs[q].ptr : nullptr;
                a.cs[q].d = q == 4 ? (const double *)c->cs[q].ptr : nullptr;
                a.cs[q].per_voxel = c->cs[q].n > 1;
            }
        }
    }
    a.H = c->H;
    a.hlat = c->hlat;
    a.T = c->T;
    a.chi = c->chi;
    a.flags = c->flags;
    a.mu_a = c->mu_a;
    a.mu_s = c->mu_s;
    a.g = c->gg;
    a.n = c->nn;
    a.k = c->k;
    a.rc = c->rc;
    a.wet_mu_a = c->wet_mu_a;
    a.wet_k = c->wet_k;
    a.wet_rc = c->wet_rc;
    a.abl_time = c->abl_time;
    a.to_remove = c->to_remove;
    a.n_new = c->counter + 1;
    a.abl_energy = c->red + 9;
    a.has_cs = p.has_cs;
    a.has_cs_mu_a = p.has_cs_mu_a;
    a.has_cs_mu_s = p.has_cs_mu_s;
    a.has_cs_k = p.has_cs_k;
    a.has_cs_rc = p.has_cs_rho_cp;
    a.cs_chi_rem = p.cs_chi_remove;
    a.alpha_max = p.alpha_max_stable;
    a.t_abl = p.t_abl;
    a.t_now = t_now;
    a.t_amb = p.t_ambient;
    a.k_void = (float)p.k_void;
    a.rc_void = (float)p.rho_cp_void;
    a.t_void = (float)p.t_void;
    a.vol = c->g.dx * c->g.dy * c->g.dz;
    a.ep = ep;
    a.nvox = c->nvox;
    BD_TRY(launch_zero_f64(c, {{c->counter + 1, 1}, {c->red + 9, 1}}));
    // the quad loads of k_phase need 16-byte aligned fields (hipMalloc gives more)
    BD_CHECK(((uintptr_t)c->T | (uintptr_t)c->H | (uintptr_t)c->hlat | (uintptr_t)c->flags |
              (uintptr_t)c->to_remove) % 16 == 0,
             "bd_phase_update: state fields are not 16-byte aligned");
    const int64_t nq = c->nvox / 4;
    const int nb = nq == 0 ? 1 : (blocks_for(nq, 256) > 8192 ? 8192 : (int)blocks_for(nq, 256));
    hipLaunchKernelGGL(k_phase, dim3(nb), dim3(256), 0, c->stream, a);
    BD_HIP(hipGetLastError());
    if (!p.no_removal && p.air_gap_k0 > 0 && p.air_gap_k1 > p.air_gap_k0) {
        const int64_t cols = (int64_t)c->g.nx * c->g.ny;
        hipLaunchKernelGGL(k_perforate, dim3(blocks_for(cols, 256)), dim3(256), 0, c->stream, c->flags, c->abl_time,
                           c->counter + 1, c->g.nx, c->g.ny, c->g.nz, p.air_gap_k0, p.air_gap_k1, t_now, ep);
        BD_HIP(hipGetLastError());
    }
    return BD_OK;
}

int launch_phase(Ctx *c, double t_now, int64_t *n_new, double *ablated_J)
{
    BD_TRY(launch_phase_enqueue(c, t_now, nullptr));
    if (n_new) {
        BD_HIP(hipMemcpyAsync(c->host_pinned, c->counter + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              c->stream));
        BD_HIP(hipMemcpyAsync((char *)c->host_pinned + 8, c->red + 9, sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        BD_HIP(hipStreamSynchronize(c->stream));
        *n_new = (int64_t)((unsigned long long *)c->host_pinned)[0];
        if (ablated_J) *ablated_J = ((double *)c->host_pinned)[1];
    }
    return BD_OK;
}

// validation_benchmarks.py:116-121 (the DOP option of the scan benchmark): before
// each epoch, every intact voxel's reduced scattering follows its Arrhenius damage,
// alpha = 1 - exp(-Omega), mu_s' = (1 - alpha) mu_s'_native + alpha mu_s'_coag, and
// mu_s = mu_s' / (1 - g).  FP64 arithmetic on the stored Omega, the FP32 result
// stored, as the host formula in float64 gives when its result is uploaded.
__global__ void k_damage_optics(float *__restrict__ mu_s, const float *__restrict__ Om,
                                const uint8_t *__restrict__ flags, double musp_native, double musp_coag,
                                double one_minus_g, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (flags[i] & F_ABL) continue;
        const double alpha = 1.0 - exp(-(double)Om[i]);
        const double musp = (1.0 - alpha) * musp_native + alpha * musp_coag;
        mu_s[i] = (float)(musp / one_minus_g);
    }
}

int launch_damage_optics(Ctx *c, double musp_native, double musp_coag, double g)
{
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_damage_optics, dim3(nb), dim3(256), 0, c->stream, c->mu_s, c->Om, c->flags, musp_native,
                       musp_coag, 1.0 - g, c->nvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

}  // namespace bd

extern "C" int bd_damage_optics(bd_ctx *h, double musp_native, double musp_coag, double g)
{
    using namespace bd;
    BD_CHECK(h, "bd_damage_optics: null argument");
    BD_CHECK(std::isfinite(musp_native) && std::isfinite(musp_coag) && std::isfinite(g) && g < 1.0,
             "bd_damage_optics: parameters must be finite with g < 1");
    Ctx *c = &h->c;
    if (!c->mu_s || !c->Om || !c->flags) {
        set_error("bd_damage_optics: mu_s, Omega and the masks must be on the device first");
        return BD_ESTATE;
    }
    BD_TRY(launch_damage_optics(c, musp_native, musp_coag, g));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}
