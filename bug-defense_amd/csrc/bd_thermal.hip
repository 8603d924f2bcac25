// bd_thermal.hip -- fused bioheat sub-step for gfx950.
//
// One launch = one explicit sub-step of
//   voxel_thermal_3d.py:30-112   (_thermal3d_step_explicit, harmonic-mean faces,
//                                 series-resistance top/bottom boundary)
//   voxel_thermal_3d.py:115-143  (_apply_latent_heat_plateau)
//   coupled_ablation.py:33-58    (_accumulate_arrhenius, _accumulate_char)
// in that order, on FP32 fields, plus the running T_peak on the last sub-step.
//
// Layout: C-order (x slowest, z contiguous).  A block is a 64(z) x BY(y) tile
// that marches along x over XC planes: the x-neighbours of T and k live in
// registers (each value is read from HBM once per sub-step), the y/z
// neighbours come from a double-buffered LDS tile with a one-cell halo, and
// every HBM access is a 64-lane z-contiguous run.  Block ids are remapped so
// that each XCD (its own L2) works on a contiguous band of y tiles, which keeps
// the y-halo re-reads inside one L2.
#include "bd_internal.hpp"

#include <cmath>

namespace bd {

struct SubArgs {
    const float *__restrict__ Tin;
    float *__restrict__ Tout;
    const float *__restrict__ k;
    const float *__restrict__ rc;
    const float *__restrict__ Q;
    uint8_t *__restrict__ flags;
    float *__restrict__ H;
    float *__restrict__ chi;
    float *__restrict__ Om;
    float *__restrict__ Tpeak;
    double *__restrict__ ledger;  // conducted energy accumulator, or null
    int Nx, Ny, Nz;
    int ntz, nty, ntx, xc;
    float idx2, idy2, idz2, dz, inv_h, h_eff;
    float dt, t_amb, t_abl, t_void, qscale;
    float arrA, arrEaR, chA, chEaR;
    int plateau, arr, chr, maskq, peak;
    double vol_dt;
};

__device__ __forceinline__ float hmean(float a, float b) { return 2.0f * a * b / (a + b + 1e-12f); }

template <int BY>
__global__ __launch_bounds__(64 * BY) void k_substep(SubArgs a)
{
    static_assert(BY >= 2, "the y-halo rows are loaded by two different waves");
    const int lane = threadIdx.x;
    const int wy = threadIdx.y;
    // XCD-aware block -> tile bijection (blocks are dealt round-robin to 8 XCDs)
    const int G = gridDim.x, b = blockIdx.x;
    const int xcd = b & 7, local = b >> 3, q8 = G >> 3, r8 = G & 7;
    const int t = xcd * q8 + (xcd < r8 ? xcd : r8) + local;
    const int tx_i = t % a.ntx;
    const int rest = t / a.ntx;
    const int tz_i = rest % a.ntz;
    const int ty_i = rest / a.ntz;
    const int Nx = a.Nx, Ny = a.Ny, Nz = a.Nz;
    const int z = tz_i * 64 + lane;
    const int y = ty_i * BY + wy;
    const int x0 = tx_i * a.xc;
    const int x1 = min(x0 + a.xc, Nx);
    const bool valid = (z < Nz) && (y < Ny);
    const int64_t plane = (int64_t)Ny * Nz;
    const int64_t col = (int64_t)y * Nz + z;

    __shared__ float sT[2][BY + 2][66];
    __shared__ float sK[2][BY + 2][66];

    float Tm = 0.f, Tc = 0.f, Tp = 0.f, km = 0.f, kc = 0.f, kp = 0.f;
    if (valid && x0 < Nx) {
        Tc = a.Tin[x0 * plane + col];
        kc = a.k[x0 * plane + col];
        if (x0 > 0) {
            Tm = a.Tin[(x0 - 1) * plane + col];
            km = a.k[(x0 - 1) * plane + col];
        }
        if (x0 + 1 < Nx) {
            Tp = a.Tin[(x0 + 1) * plane + col];
            kp = a.k[(x0 + 1) * plane + col];
        }
    }
    const int yh_lo = ty_i * BY - 1, yh_hi = ty_i * BY + BY;
    const int zh_lo = tz_i * 64 - 1, zh_hi = tz_i * 64 + 64;
    double econd = 0.0;
    int buf = 0;
    // per-cell inputs of the next plane, loaded one iteration ahead so their
    // latency overlaps the current plane's stencil
    float cap_n = 1.f, q_n = 0.f, h_n = 0.f, chi_n = 0.f;
    uint8_t fl_n = 0;
    if (valid && x0 < x1) {
        const int64_t id0 = x0 * plane + col;
        cap_n = a.rc[id0];
        q_n = a.Q[id0];
        fl_n = a.flags[id0];
        if (a.plateau && (fl_n & F_HPOS)) h_n = a.H[id0];
        if (a.chr && (fl_n & F_DRY)) chi_n = a.chi[id0];
    }
    // halo cells of a plane (y rows by the edge waves, z ends by the edge lanes),
    // also loaded one plane ahead
    const bool hy_lo = wy == 0 && yh_lo >= 0 && z < Nz, hy_hi = wy == BY - 1 && yh_hi < Ny && z < Nz;
    const bool hz_lo = lane == 0 && zh_lo >= 0 && y < Ny, hz_hi = lane == 63 && zh_hi < Nz && y < Ny;
    const int64_t off_y = hy_lo ? (int64_t)yh_lo * Nz + z : (int64_t)yh_hi * Nz + z;
    const int64_t off_z = hz_lo ? (int64_t)y * Nz + zh_lo : (int64_t)y * Nz + zh_hi;
    float hyT = 0.f, hyK = 0.f, hzT = 0.f, hzK = 0.f;
    if (x0 < x1) {
        if (hy_lo || hy_hi) {
            hyT = a.Tin[x0 * plane + off_y];
            hyK = a.k[x0 * plane + off_y];
        }
        if (hz_lo || hz_hi) {
            hzT = a.Tin[x0 * plane + off_z];
            hzK = a.k[x0 * plane + off_z];
        }
    }
    for (int x = x0; x < x1; ++x) {
        const int64_t pbase = x * plane;
        sT[buf][wy + 1][lane + 1] = Tc;
        sK[buf][wy + 1][lane + 1] = kc;
        if (hy_lo) {
            sT[buf][0][lane + 1] = hyT;
            sK[buf][0][lane + 1] = hyK;
        }
        if (hy_hi) {
            sT[buf][BY + 1][lane + 1] = hyT;
            sK[buf][BY + 1][lane + 1] = hyK;
        }
        if (hz_lo) {
            sT[buf][wy + 1][0] = hzT;
            sK[buf][wy + 1][0] = hzK;
        }
        if (hz_hi) {
            sT[buf][wy + 1][65] = hzT;
            sK[buf][wy + 1][65] = hzK;
        }
        if (x + 1 < x1) {
            if (hy_lo || hy_hi) {
                hyT = a.Tin[pbase + plane + off_y];
                hyK = a.k[pbase + plane + off_y];
            }
            if (hz_lo || hz_hi) {
                hzT = a.Tin[pbase + plane + off_z];
                hzK = a.k[pbase + plane + off_z];
            }
        }
        float Tn2 = 0.f, kn2 = 0.f;
        const float cap = cap_n, qv = q_n, h0 = h_n, chi0 = chi_n;
        const uint8_t fl = fl_n;
        const int64_t id = pbase + col;
        if (valid) {
            if (x + 2 < Nx) {
                Tn2 = a.Tin[id + 2 * plane];
                kn2 = a.k[id + 2 * plane];
            }
            if (x + 1 < x1) {
                const int64_t idn = id + plane;
                cap_n = a.rc[idn];
                q_n = a.Q[idn];
                fl_n = a.flags[idn];
                h_n = (a.plateau && (fl_n & F_HPOS)) ? a.H[idn] : 0.f;
                chi_n = (a.chr && (fl_n & F_DRY)) ? a.chi[idn] : 0.f;
            }
        }
        __syncthreads();
        if (valid) {
            // face fluxes in the reference's summation order: w, e, s, n, top, bottom
            float fw = 0.f, fe = 0.f, fs = 0.f, fn = 0.f, ft = 0.f, fb = 0.f;
            if (x > 0) fw = hmean(kc, km) * (Tm - Tc) * a.idx2;
            if (x < Nx - 1) fe = hmean(kc, kp) * (Tp - Tc) * a.idx2;
            if (y > 0) fs = hmean(kc, sK[buf][wy][lane + 1]) * (sT[buf][wy][lane + 1] - Tc) * a.idy2;
            if (y < Ny - 1) fn = hmean(kc, sK[buf][wy + 2][lane + 1]) * (sT[buf][wy + 2][lane + 1] - Tc) * a.idy2;
            float fbound = 0.f;
            if (z > 0) ft = hmean(kc, sK[buf][wy + 1][lane]) * (sT[buf][wy + 1][lane] - Tc) * a.idz2;
            else if (a.h_eff > 0.f) {
                ft = (a.t_amb - Tc) / (a.dz * (0.5f * a.dz / fmaxf(kc, 1e-30f) + a.inv_h));
                fbound += ft;
            }
            if (z < Nz - 1) fb = hmean(kc, sK[buf][wy + 1][lane + 2]) * (sT[buf][wy + 1][lane + 2] - Tc) * a.idz2;
            else if (a.h_eff > 0.f) {
                fb = (a.t_amb - Tc) / (a.dz * (0.5f * a.dz / fmaxf(kc, 1e-30f) + a.inv_h));
                fbound += fb;
            }
            const float div = fw + fe + fs + fn + ft + fb;
            float src = qv * a.qscale;
            if (a.maskq && (fl & F_ABL)) src = 0.f;
            float Tn = Tc + a.dt * (div + src) / cap;
            econd += (double)fbound;
            uint8_t fl2 = fl;
            if (a.plateau) {
                if (fl & F_ABL) {
                    if (Tn > a.t_void) Tn = a.t_void;
                } else if (fl & F_DRY) {
                } else if (Tn > a.t_abl) {
                    float h = (fl & F_HPOS) ? h0 : 0.f;
                    h += (Tn - a.t_abl) * cap;
                    a.H[id] = h;
                    Tn = a.t_abl;
                    fl2 = h > 0.f ? (uint8_t)(fl | F_HPOS) : (uint8_t)(fl & ~F_HPOS);
                } else if (fl & F_HPOS) {
                    float h = h0;
                    const float rel = fminf(h, (a.t_abl - Tn) * cap);
                    h -= rel;
                    a.H[id] = h;
                    Tn = Tn + rel / cap;
                    if (!(h > 0.f)) fl2 = (uint8_t)(fl & ~F_HPOS);
                }
            }
            if (a.arr && !(fl & F_ABL)) a.Om[id] += a.dt * a.arrA * expf(-a.arrEaR / (Tn + 273.15f));
            if (a.chr && (fl & F_DRY) && !(fl & F_ABL)) {
                const float rate = a.chA * expf(-a.chEaR / (Tn + 273.15f));
                a.chi[id] = 1.0f - (1.0f - chi0) * expf(-rate * a.dt);
            }
            a.Tout[id] = Tn;
            if (a.peak) a.Tpeak[id] = fmaxf(a.Tpeak[id], Tn);
            if (fl2 != fl) a.flags[id] = fl2;
        }
        Tm = Tc; km = kc;
        Tc = Tp; kc = kp;
        Tp = Tn2; kp = kn2;
        buf ^= 1;
    }
    if (a.ledger) {
        // block reduction of the boundary energy (only boundary cells contribute)
        __shared__ double red[64 * BY];
        red[wy * 64 + lane] = econd;
        __syncthreads();
        for (int w = 32 * BY; w > 0; w >>= 1) {
            const int tid = wy * 64 + lane;
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        if (wy == 0 && lane == 0 && red[0] != 0.0) unsafeAtomicAdd(a.ledger, red[0] * a.vol_dt);
    }
}

static int launch_one(Ctx *c, SubArgs &a, int by)
{
    const int ntz = (c->g.nz + 63) / 64;
    const int nty = (c->g.ny + by - 1) / by;
    const int ntx = (c->g.nx + c->tile_xc - 1) / c->tile_xc;
    a.ntz = ntz;
    a.nty = nty;
    a.ntx = ntx;
    a.xc = c->tile_xc;
    const int64_t tiles = (int64_t)ntz * nty * ntx;
    if (tiles > 0x7fffffff) { set_error("sub-step grid too large"); return BD_EINVAL; }
    switch (by) {
    case 2: hipLaunchKernelGGL(k_substep<2>, dim3((unsigned)tiles), dim3(64, 2), 0, c->stream, a); break;
    case 8: hipLaunchKernelGGL(k_substep<8>, dim3((unsigned)tiles), dim3(64, 8), 0, c->stream, a); break;
    default: hipLaunchKernelGGL(k_substep<4>, dim3((unsigned)tiles), dim3(64, 4), 0, c->stream, a); break;
    }
    BD_HIP(hipGetLastError());
    return BD_OK;
}

static void fill_args(Ctx *c, SubArgs &a, int is_active, int flags)
{
    const bd_physics &p = c->phys;
    a.k = c->k;
    a.rc = c->rc;
    a.Q = c->Q;
    a.flags = c->flags;
    a.H = c->H;
    a.chi = c->chi;
    a.Om = c->Om;
    a.Tpeak = c->Tpeak;
    a.Nx = c->g.nx;
    a.Ny = c->g.ny;
    a.Nz = c->g.nz;
    a.idx2 = (float)(1.0 / (c->g.dx * c->g.dx));
    a.idy2 = (float)(1.0 / (c->g.dy * c->g.dy));
    a.idz2 = (float)(1.0 / (c->g.dz * c->g.dz));
    a.dz = (float)c->g.dz;
    a.h_eff = (float)p.h_eff;
    a.inv_h = (float)(1.0 / std::fmax(p.h_eff, 1e-30));
    a.dt = (float)p.dt_sub;
    a.t_amb = (float)p.t_ambient;
    a.t_abl = (float)p.t_abl;
    a.t_void = (float)p.t_void;
    a.qscale = is_active ? 1.0f : 0.0f;
    a.arrA = (float)p.arr_A;
    a.arrEaR = (float)p.arr_Ea_over_R;
    a.chA = (float)p.char_A;
    a.chEaR = (float)p.char_Ea_over_R;
    a.plateau = (flags & BD_SUB_PLATEAU) ? 1 : 0;
    a.arr = ((flags & BD_SUB_ARRHENIUS) && p.has_arrhenius) ? 1 : 0;
    a.chr = ((flags & BD_SUB_CHAR) && p.has_char) ? 1 : 0;
    a.maskq = (flags & BD_SUB_MASK_Q) ? 1 : 0;
    a.vol_dt = c->g.dx * c->g.dy * c->g.dz * p.dt_sub;
}

// Pick the tile shape once per context by timing the candidates on this grid
// (a sub-step is idempotent only on scratch buffers, so tuning runs on T2 -> staging).
static int autotune(Ctx *c)
{
    if (c->tuned) return BD_OK;
    c->tuned = true;
    if (c->nvox < (1 << 20)) {  // small grids: fixed shape, nothing to win
        c->tile_by = 4;
        c->tile_xc = 8;
        return BD_OK;
    }
    BD_TRY(ensure_staging(c, (size_t)c->nvox * 8));
    // scratch copies so the tuning launches leave the state untouched
    float *Tscr = (float *)c->staging;
    float *Hscr = (float *)((char *)c->staging + (size_t)c->nvox * 4);
    BD_HIP(hipMemcpyAsync(Tscr, c->T, (size_t)c->nvox * 4, hipMemcpyDeviceToDevice, c->stream));
    BD_HIP(hipMemcpyAsync(Hscr, c->H, (size_t)c->nvox * 4, hipMemcpyDeviceToDevice, c->stream));
    const int bys[] = {2, 4, 8};
    const int xcs[] = {8, 16, 32, 64};
    float best = 1e30f;
    int bb = 4, bx = 16;
    for (int by : bys)
        for (int xc : xcs) {
            c->tile_xc = xc;
            SubArgs a{};
            fill_args(c, a, 1, BD_SUB_PLATEAU);
            a.Tin = Tscr;
            a.Tout = c->T2;
            a.H = Hscr;
            a.flags = nullptr;  // never written: plateau reads flags; give a private copy below
            a.flags = c->flags;
            a.plateau = 0;      // stencil + loads only; the plateau writes would perturb the state
            a.chr = 0;
            a.arr = 0;
            float ms = 0.f;
            BD_TRY(launch_one(c, a, by));  // warm
            BD_HIP(hipEventRecord(c->ev0, c->stream));
            for (int r = 0; r < 3; ++r) BD_TRY(launch_one(c, a, by));
            BD_HIP(hipEventRecord(c->ev1, c->stream));
            BD_HIP(hipEventSynchronize(c->ev1));
            BD_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            if (ms < best) {
                best = ms;
                bb = by;
                bx = xc;
            }
        }
    c->tile_by = bb;
    c->tile_xc = bx;
    return BD_OK;
}

int launch_substeps(Ctx *c, int n, int is_active, int flags, bool ledger)
{
    BD_TRY(autotune(c));
    if (flags & BD_SUB_PEAK) BD_TRY(ensure(c, (void **)&c->Tpeak, (size_t)c->nvox * 4));
    if (flags & BD_SUB_ARRHENIUS) BD_TRY(ensure(c, (void **)&c->Om, (size_t)c->nvox * 4));
    SubArgs a{};
    fill_args(c, a, is_active, flags);
    a.Om = c->Om;
    a.Tpeak = c->Tpeak;
    a.ledger = ledger ? c->red + 8 : nullptr;  // red[8]: conducted energy of this call
    for (int s = 0; s < n; ++s) {
        a.Tin = c->T;
        a.Tout = c->T2;
        a.peak = ((flags & BD_SUB_PEAK) && s == n - 1) ? 1 : 0;
        BD_TRY(launch_one(c, a, c->tile_by));
        float *tmp = c->T;
        c->T = c->T2;
        c->T2 = tmp;
    }
    return BD_OK;
}

// ------------------------------------------------------------------ module kernels
__global__ void k_plateau(float *__restrict__ T, float *__restrict__ H, const float *__restrict__ rc,
                          uint8_t *__restrict__ flags, float t_abl, float t_void, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t fl = flags[i];
        const float t = T[i];
        if (fl & F_ABL) {
            if (t > t_void) T[i] = t_void;
        } else if (fl & F_DRY) {
        } else if (t > t_abl) {
            const float h = ((fl & F_HPOS) ? H[i] : 0.f) + (t - t_abl) * rc[i];
            H[i] = h;
            T[i] = t_abl;
            const uint8_t f2 = h > 0.f ? (uint8_t)(fl | F_HPOS) : (uint8_t)(fl & ~F_HPOS);
            if (f2 != fl) flags[i] = f2;
        } else if (fl & F_HPOS) {
            const float cap = rc[i];
            float h = H[i];
            const float rel = fminf(h, (t_abl - t) * cap);
            h -= rel;
            H[i] = h;
            T[i] = t + rel / cap;
            if (!(h > 0.f)) flags[i] = (uint8_t)(fl & ~F_HPOS);
        }
    }
}

__global__ void k_arrhenius(const float *__restrict__ T, float *__restrict__ Om, const uint8_t *__restrict__ flags,
                            float A, float EaR, float dt, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (!(flags[i] & F_ABL)) Om[i] += dt * A * expf(-EaR / (T[i] + 273.15f));
}

__global__ void k_char(const float *__restrict__ T, float *__restrict__ chi, const uint8_t *__restrict__ flags,
                       float A, float EaR, float dt, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t fl = flags[i];
        if ((fl & F_DRY) && !(fl & F_ABL)) {
            const float rate = A * expf(-EaR / (T[i] + 273.15f));
            chi[i] = 1.0f - (1.0f - chi[i]) * expf(-rate * dt);
        }
    }
}

static int ew_blocks(int64_t n) { return blocks_for(n, 256) > 8192 ? 8192 : blocks_for(n, 256); }

int launch_plateau(Ctx *c)
{
    hipLaunchKernelGGL(k_plateau, dim3(ew_blocks(c->nvox)), dim3(256), 0, c->stream, c->T, c->H, c->rc, c->flags,
                       (float)c->phys.t_abl, (float)c->phys.t_void, c->nvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int launch_arrhenius(Ctx *c, double A, double EaR, double dt)
{
    hipLaunchKernelGGL(k_arrhenius, dim3(ew_blocks(c->nvox)), dim3(256), 0, c->stream, c->T, c->Om, c->flags,
                       (float)A, (float)EaR, (float)dt, c->nvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int launch_char(Ctx *c, double A, double EaR, double dt)
{
    hipLaunchKernelGGL(k_char, dim3(ew_blocks(c->nvox)), dim3(256), 0, c->stream, c->T, c->chi, c->flags, (float)A,
                       (float)EaR, (float)dt, c->nvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

}  // namespace bd
