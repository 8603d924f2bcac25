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

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <utility>

namespace bd {

// Arrhenius rates are evaluated as exp(ln A - Ea/(R T)): published pre-factors
// reach 1e98 (Henriques-Moritz skin kinetics), beyond FP32, while ln A stays
// small.  ln A is taken in FP64 on the host; A = 0 gives -inf, i.e. no damage.
static inline float ln_prefactor(double A)
{
    return A > 0.0 ? (float)std::log(A) : -INFINITY;
}

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
    double *__restrict__ ledger;  // conducted energy: LEDGER_SLOTS FP64 slots LEDGER_STRIDE apart, or null
    int Nx, Ny, Nz;
    int ntz, nty, ntx, xc;
    float idx2, idy2, idz2, dz, inv_h, h_eff;
    float dt, t_amb, t_abl, t_void, qscale;
    float arrLnA, arrEaR, chLnA, chEaR, ch_tcold;  // ln of the pre-factors (they may exceed FP32)
    int plateau, arr, chr, maskq, peak;
    double vol_dt;
    // two-sub-step kernel: H and flags are read from one buffer and written to another
    const uint8_t *__restrict__ flags_in;
    const float *__restrict__ H_in;
};



// Face conductance and 1/rho_cp with the hardware reciprocal (v_rcp_f32, 1 ulp):
// the stencil's divisions dominate its ALU time, and the reference evaluates
// this loop under fastmath (voxel_thermal_3d.py:30, numba fastmath=True).
__device__ __forceinline__ float hmean_fast(float a, float b)
{
    return 2.0f * a * b * __builtin_amdgcn_rcpf(a + b + 1e-12f);
}

// Constants of one sub-step, shared by the one- and two-step kernels.
struct StepK {
    int Nx, Ny, Nz;
    float idx2, idy2, idz2, dz, inv_h, h_eff;
    float dt, t_amb, t_abl, t_void, qscale;
    float arrLnA, arrEaR, chLnA, chEaR, ch_tcold;  // ln of the pre-factors (they may exceed FP32)
    int plateau, arr, chr, maskq;
};

__device__ __forceinline__ StepK step_consts(const SubArgs &a)
{
    StepK k;
    k.Nx = a.Nx; k.Ny = a.Ny; k.Nz = a.Nz;
    k.idx2 = a.idx2; k.idy2 = a.idy2; k.idz2 = a.idz2; k.dz = a.dz; k.inv_h = a.inv_h; k.h_eff = a.h_eff;
    k.dt = a.dt; k.t_amb = a.t_amb; k.t_abl = a.t_abl; k.t_void = a.t_void; k.qscale = a.qscale;
    k.arrLnA = a.arrLnA; k.arrEaR = a.arrEaR; k.chLnA = a.chLnA; k.chEaR = a.chEaR; k.ch_tcold = a.ch_tcold;
    k.plateau = a.plateau; k.arr = a.arr; k.chr = a.chr; k.maskq = a.maskq;
    return k;
}

// The series-resistance flux of a top or bottom boundary cell (voxel_thermal_3d.py:84-98), with
// the hardware reciprocal (1 ulp) like the face conductances: the two IEEE divisions it replaces
// were issued on every plane by each wave holding a boundary lane.
__device__ __forceinline__ float bnd_flux(const StepK &a, float Tc, float kc)
{
#pragma clang fp contract(off)
    return (a.t_amb - Tc) *
           __builtin_amdgcn_rcpf(a.dz * (0.5f * a.dz * __builtin_amdgcn_rcpf(fmaxf(kc, 1e-30f)) + a.inv_h));
}

// Source, explicit update and latent-heat plateau of one cell from its face-flux sum `div`
// (voxel_thermal_3d.py:100-143) on (h, fl).  Every sub-step kernel ends each cell here, with
// contraction off, so they all produce the same bits.
__device__ __forceinline__ float cell_tail(const StepK &a, float Tc, float div, float cap, float qv, uint8_t fl,
                                           float &h, uint8_t &fl2)
{
#pragma clang fp contract(off)
    float src = qv * a.qscale;
    if (a.maskq && (fl & F_ABL)) src = 0.f;
    const float icap = __builtin_amdgcn_rcpf(cap);
    float Tn = Tc + a.dt * (div + src) * icap;
    fl2 = fl;
    if (a.plateau) {
        if (fl & F_ABL) {
            if (Tn > a.t_void) Tn = a.t_void;
        } else if (fl & F_DRY) {
        } else if (Tn > a.t_abl) {
            float hh = (fl & F_HPOS) ? h : 0.f;
            hh += (Tn - a.t_abl) * cap;
            h = hh;
            Tn = a.t_abl;
            fl2 = hh > 0.f ? (uint8_t)(fl | F_HPOS) : (uint8_t)(fl & ~F_HPOS);
        } else if (fl & F_HPOS) {
            float hh = h;
            const float rel = fminf(hh, (a.t_abl - Tn) * cap);
            hh -= rel;
            h = hh;
            Tn = Tn + rel * icap;
            if (!(hh > 0.f)) fl2 = (uint8_t)(fl & ~F_HPOS);
        }
    }
    return Tn;
}

// One explicit sub-step of one cell (voxel_thermal_3d.py:30-143): face fluxes in
// the reference's order (w, e, s, n, top, bottom), source, the latent-heat
// plateau on (h, fl).  Returns the new T; *fbound gets the boundary flux of the cell.
__device__ __forceinline__ float cell_step(const StepK &a, int x, int y, int z, float Tc, float kc, float Tm, float km,
                                           float Tp, float kp, float Ts, float ks, float Tn_, float kn, float Tt,
                                           float kt, float Tb, float kb, float cap, float qv, uint8_t fl, float &h,
                                           uint8_t &fl2, float *fbound)
{
#pragma clang fp contract(off)
    float fw = 0.f, fe = 0.f, fs = 0.f, fn = 0.f, ft = 0.f, fb = 0.f;
    if (x > 0) fw = hmean_fast(kc, km) * (Tm - Tc) * a.idx2;
    if (x < a.Nx - 1) fe = hmean_fast(kc, kp) * (Tp - Tc) * a.idx2;
    if (y > 0) fs = hmean_fast(kc, ks) * (Ts - Tc) * a.idy2;
    if (y < a.Ny - 1) fn = hmean_fast(kc, kn) * (Tn_ - Tc) * a.idy2;
    float fbd = 0.f;
    if (z > 0) ft = hmean_fast(kc, kt) * (Tt - Tc) * a.idz2;
    else if (a.h_eff > 0.f) {
        ft = bnd_flux(a, Tc, kc);
        fbd += ft;
    }
    if (z < a.Nz - 1) fb = hmean_fast(kc, kb) * (Tb - Tc) * a.idz2;
    else if (a.h_eff > 0.f) {
        fb = bnd_flux(a, Tc, kc);
        fbd += fb;
    }
    const float div = fw + fe + fs + fn + ft + fb;
    *fbound = fbd;
    return cell_tail(a, Tc, div, cap, qv, fl, h, fl2);
}

// Char kinetics (coupled_ablation.py:48-58) run on dry, intact cells above
// ch_tcold.  Below it rate dt < 2^-26, where e^{-rate dt} rounds to 1 in FP32
// and the FP32 update 1 - (1 - chi) e^{-rate dt} leaves chi as it is (up to the
// rounding of 1 - (1 - chi)); such a cell keeps its chi, which is then neither
// read nor written.
__device__ __forceinline__ bool char_active(const StepK &a, float Tn, uint8_t fl)
{
    return a.chr && (fl & F_DRY) && !(fl & F_ABL) && Tn > a.ch_tcold;
}

__device__ __forceinline__ float char_step(const StepK &a, float Tn, float chi)
{
#pragma clang fp contract(off)
    const float rate = expf(a.chLnA - a.chEaR * __builtin_amdgcn_rcpf(Tn + 273.15f));
    return 1.0f - (1.0f - chi) * expf(-rate * a.dt);
}

// Arrhenius damage integral (coupled_ablation.py:33-45), on intact cells
__device__ __forceinline__ float arr_step(const StepK &a, float Tn, float om)
{
#pragma clang fp contract(off)
    return om + a.dt * expf(a.arrLnA - a.arrEaR * __builtin_amdgcn_rcpf(Tn + 273.15f));
}

// Arrhenius damage and char kinetics of one cell at its new temperature
// (coupled_ablation.py:33-58)
__device__ __forceinline__ void cell_damage(const StepK &a, float Tn, uint8_t fl, float &om, float &chi)
{
    if (a.arr && !(fl & F_ABL)) om = arr_step(a, Tn, om);
    if (char_active(a, Tn, fl)) chi = char_step(a, Tn, chi);
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Conducted-energy ledger slots of the sub-step launches: a z-face block adds
// into slot t mod LEDGER_SLOTS, the slots a few cache lines apart, and the call
// sums them once after its launches.  A single accumulator took one atomic from
// every z-face block of a launch (two thirds of them at 192 z cells), all queued
// on one L2 line.
constexpr int LEDGER_SLOTS = 64;
constexpr int LEDGER_STRIDE = 64;  // doubles: 512 bytes between slots

// Adds a block's conducted boundary energy to its ledger slot.  Only the cells
// at z == 0 and z == Nz-1 carry a boundary flux, so only the lanes holding those
// two planes can be non-zero: each wave reads them out of its lanes (no LDS
// tree, no barrier in blocks away from the z faces), the block sums its waves
// in wave order and issues at most one atomic.  zlane0 is the z of lane 0;
// lanes below `lmin` are halo lanes whose econd is zero anyway.
template <int BY>
__device__ __forceinline__ void ledger_flush(const SubArgs &a, double econd, int zlane0, int lmin, int wy, int lane,
                                             int t)
{
    const int l_top = -zlane0, l_bot = a.Nz - 1 - zlane0;
    const bool has_top = l_top >= lmin && l_top < 64, has_bot = l_bot >= lmin && l_bot < 64 && l_bot != l_top;
    if (!(has_top || has_bot)) return;  // block-uniform
    double w = 0.0;
    if (has_top) w += readlane_f64(econd, l_top);
    if (has_bot) w += readlane_f64(econd, l_bot);
    __shared__ double red[BY];
    if (lane == 0) red[wy] = w;
    __syncthreads();
    if (wy == 0 && lane == 0) {
        double s = 0.0;
        for (int i = 0; i < BY; ++i) s += red[i];
        if (s != 0.0) unsafeAtomicAdd(&a.ledger[(t % LEDGER_SLOTS) * LEDGER_STRIDE], s * a.vol_dt);
    }
}

template <int BY>
__global__ __launch_bounds__(64 * BY) void k_substep(SubArgs a)
{
    static_assert(BY >= 2, "the y-halo rows are loaded by two different waves");
    const StepK K = step_consts(a);
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
    float cap_n = 1.f, q_n = 0.f, h_n = 0.f;
    uint8_t fl_n = 0;
    if (valid && x0 < x1) {
        const int64_t id0 = x0 * plane + col;
        cap_n = a.rc[id0];
        q_n = a.Q[id0];
        fl_n = a.flags[id0];
        if (a.plateau && (fl_n & F_HPOS)) h_n = a.H[id0];
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
        const float cap = cap_n, qv = q_n, h0 = h_n;
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
            }
        }
        __syncthreads();
        if (valid) {
            float h = h0, fbound = 0.f;
            uint8_t fl2;
            const float Tn = cell_step(K, x, y, z, Tc, kc, Tm, km, Tp, kp, sT[buf][wy][lane + 1],
                                       sK[buf][wy][lane + 1], sT[buf][wy + 2][lane + 1], sK[buf][wy + 2][lane + 1],
                                       sT[buf][wy + 1][lane], sK[buf][wy + 1][lane], sT[buf][wy + 1][lane + 2],
                                       sK[buf][wy + 1][lane + 2], cap, qv, fl, h, fl2, &fbound);
            econd += (double)fbound;
            if (fl2 != fl || h != h0) a.H[id] = h;
            if (a.arr && !(fl & F_ABL)) a.Om[id] = arr_step(K, Tn, a.Om[id]);
            // chi is read only where the char step runs (hot, dry, intact cells)
            if (char_active(K, Tn, fl)) a.chi[id] = char_step(K, Tn, a.chi[id]);
            a.Tout[id] = Tn;
            if (a.peak) a.Tpeak[id] = fmaxf(a.Tpeak[id], Tn);
            if (fl2 != fl) a.flags[id] = fl2;
        }
        Tm = Tc; km = kc;
        Tc = Tp; kc = kp;
        Tp = Tn2; kp = kn2;
        buf ^= 1;
    }
    if (a.ledger) ledger_flush<BY>(a, econd, z - lane, 0, wy, lane, t);
}

// Two sub-steps per launch (temporal blocking).  A block owns a 60(z) x
// (BY-4)(y) tile and marches along x; its 64 lanes x BY rows also cover a
// two-cell halo ring.  Iteration p computes sub-step 1 on plane p for the tile
// and the inner halo ring (from the input T in LDS and registers) and sub-step
// 2 on plane p-1 for the owned cells (from the sub-step-1 values of the
// previous iteration, held in LDS and registers), so the intermediate field
// never goes to HBM: per two sub-steps each field is read once and T, H and
// the flags are written once.  Other blocks still read the input H and flags
// of their halo cells, so those two are read from one buffer and written to
// another (the caller swaps them); chi and Omega only matter to their owner
// and stay in place.  Every cell goes through cell_step, so the result is the
// same as two launches of k_substep.
template <int BY>
__global__ __launch_bounds__(64 * BY) void k_substep2(SubArgs a)
{
    static_assert(BY >= 6, "two halo rows on each side");
    constexpr int IY = BY - 4, IZ = 60;
    const StepK K = step_consts(a);
    const int lane = threadIdx.x, wy = threadIdx.y;
    const int G = gridDim.x, b = blockIdx.x;
    const int xcd = b & 7, local = b >> 3, q8 = G >> 3, r8 = G & 7;
    const int t = xcd * q8 + (xcd < r8 ? xcd : r8) + local;
    const int tx_i = t % a.ntx;
    const int rest = t / a.ntx;
    const int tz_i = rest % a.ntz;
    const int ty_i = rest / a.ntz;
    const int Nx = a.Nx, Ny = a.Ny, Nz = a.Nz;
    const int z = tz_i * IZ - 2 + lane;
    const int y = ty_i * IY - 2 + wy;
    const bool in = z >= 0 && z < Nz && y >= 0 && y < Ny;
    const bool ring = in && lane >= 1 && lane <= 62 && wy >= 1 && wy <= BY - 2;   // sub-step 1 cells
    const bool own = in && lane >= 2 && lane <= 61 && wy >= 2 && wy <= BY - 3;    // sub-step 2 (written) cells
    const int x0 = tx_i * a.xc;
    const int x1 = min(x0 + a.xc, Nx);
    const int64_t plane = (int64_t)Ny * Nz;
    const int64_t col = in ? (int64_t)y * Nz + z : 0;

    // sK holds the k plane read by sub-step 1 of this iteration and by sub-step 2 of
    // the next one, so it needs three slots: the slot written before this
    // iteration's barrier was last read by sub-step 2 two iterations ago
    __shared__ float sT0[2][BY][64], sK[3][BY][64], sT1[2][BY][64];

    auto ldT = [&](int p) { return (in && p >= 0 && p < Nx) ? a.Tin[p * plane + col] : 0.f; };
    auto ldK = [&](int p) { return (in && p >= 0 && p < Nx) ? a.k[p * plane + col] : 0.f; };
    struct Cell {
        float cap, q, h, chi, om;
        uint8_t fl;
    };
    auto ldCell = [&](int p) {
        Cell c{1.f, 0.f, 0.f, 0.f, 0.f, 0};
        if (ring && p >= 0 && p < Nx) {
            const int64_t id = p * plane + col;
            c.cap = a.rc[id];
            c.q = a.Q[id];
            c.fl = a.flags_in[id];
            if (a.plateau && (c.fl & F_HPOS)) c.h = a.H_in[id];
            if (own && p >= x0 && p < x1) {
                if (a.chr && (c.fl & F_DRY)) c.chi = a.chi[id];
                if (a.arr) c.om = a.Om[id];
            }
        }
        return c;
    };

    float T0m = ldT(x0 - 2), T0c = ldT(x0 - 1), T0p = ldT(x0);
    float Kmm = ldK(x0 - 3), Km = ldK(x0 - 2), Kc = ldK(x0 - 1), Kp = ldK(x0);
    float T1m = 0.f, T1c = 0.f;
    Cell cp = ldCell(x0 - 1), cq{1.f, 0.f, 0.f, 0.f, 0.f, 0};
    float h1q = 0.f, chi1q = 0.f, om1q = 0.f;
    uint8_t fl1q = 0;
    double econd = 0.0;
    int buf = 0, kb = 0, kbp = 2;  // sK slot of plane p and of plane p - 1
    for (int p = x0 - 1; p <= x1; ++p) {
        sT0[buf][wy][lane] = T0c;
        sK[kb][wy][lane] = Kc;
        // next iteration's inputs, in flight across the barrier and the two sub-steps
        const float T0n = ldT(p + 2), Kn = ldK(p + 2);
        Cell cn{1.f, 0.f, 0.f, 0.f, 0.f, 0};
        if (p + 1 <= x1) cn = ldCell(p + 1);
        __syncthreads();
        // ---- sub-step 1 on plane p (tile + inner ring)
        float T1p = 0.f, h1p = cp.h, chi1p = cp.chi, om1p = cp.om;
        uint8_t fl1p = cp.fl;
        if (ring && p >= 0 && p < Nx) {
            float fbound = 0.f;
            T1p = cell_step(K, p, y, z, T0c, Kc, T0m, Km, T0p, Kp, sT0[buf][wy - 1][lane], sK[kb][wy - 1][lane],
                            sT0[buf][wy + 1][lane], sK[kb][wy + 1][lane], sT0[buf][wy][lane - 1],
                            sK[kb][wy][lane - 1], sT0[buf][wy][lane + 1], sK[kb][wy][lane + 1], cp.cap, cp.q, cp.fl,
                            h1p, fl1p, &fbound);
            if (own && p >= x0 && p < x1) {
                econd += (double)fbound;
                cell_damage(K, T1p, cp.fl, om1p, chi1p);
            }
        }
        sT1[buf][wy][lane] = T1p;
        // ---- sub-step 2 on plane q = p - 1 (owned cells), from the sub-step-1 plane
        // written in the previous iteration
        const int q = p - 1;
        if (own && q >= x0 && q < x1) {
            const int ob = buf ^ 1;
            float h2 = h1q, fbound = 0.f;
            uint8_t fl2;
            const float T2 = cell_step(K, q, y, z, T1c, Km, T1m, Kmm, T1p, Kc, sT1[ob][wy - 1][lane],
                                       sK[kbp][wy - 1][lane], sT1[ob][wy + 1][lane], sK[kbp][wy + 1][lane],
                                       sT1[ob][wy][lane - 1], sK[kbp][wy][lane - 1], sT1[ob][wy][lane + 1],
                                       sK[kbp][wy][lane + 1], cq.cap, cq.q, fl1q, h2, fl2, &fbound);
            econd += (double)fbound;
            float om2 = om1q, chi2 = chi1q;
            cell_damage(K, T2, fl1q, om2, chi2);
            const int64_t id = q * plane + col;
            a.Tout[id] = T2;
            if (a.peak) a.Tpeak[id] = fmaxf(a.Tpeak[id], T2);
            // the output H buffer must read 0 wherever its flags say nothing is banked:
            // write when banked now or when this buffer's previous contents were banked
            const uint8_t fl_prev_out = a.flags[id];
            if ((fl2 | fl_prev_out) & F_HPOS) a.H[id] = (fl2 & F_HPOS) ? h2 : 0.f;
            a.flags[id] = fl2;
            if (a.arr && !(fl1q & F_ABL)) a.Om[id] = om2;
            if (char_active(K, T1c, cq.fl) || char_active(K, T2, fl1q)) a.chi[id] = chi2;
        }
        // roll the plane registers
        T0m = T0c;
        T0c = T0p;
        T0p = T0n;
        Kmm = Km;
        Km = Kc;
        Kc = Kp;
        Kp = Kn;
        T1m = T1c;
        T1c = T1p;
        cq = cp;
        h1q = h1p;
        fl1q = fl1p;
        chi1q = chi1p;
        om1q = om1p;
        cp = cn;
        buf ^= 1;
        kbp = kb;
        kb = kb == 2 ? 0 : kb + 1;
    }
    if (a.ledger) ledger_flush<BY>(a, econd, z - lane, 2, wy, lane, t);
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

static int launch_two(Ctx *c, SubArgs &a, int by, int xc)
{
    const int ntz = (c->g.nz + 59) / 60;
    const int nty = (c->g.ny + (by - 4) - 1) / (by - 4);
    const int ntx = (c->g.nx + xc - 1) / xc;
    a.ntz = ntz;
    a.nty = nty;
    a.ntx = ntx;
    a.xc = xc;
    const int64_t tiles = (int64_t)ntz * nty * ntx;
    if (tiles > 0x7fffffff) { set_error("sub-step grid too large"); return BD_EINVAL; }
    if (by == 8) hipLaunchKernelGGL(k_substep2<8>, dim3((unsigned)tiles), dim3(64, 8), 0, c->stream, a);
    else hipLaunchKernelGGL(k_substep2<16>, dim3((unsigned)tiles), dim3(64, 16), 0, c->stream, a);
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
    a.arrLnA = ln_prefactor(p.arr_A);
    a.arrEaR = (float)p.arr_Ea_over_R;
    a.chLnA = ln_prefactor(p.char_A);
    a.chEaR = (float)p.char_Ea_over_R;
    // temperature below which one step's rate dt stays under 2^-26:
    // A dt e^{-Ea/(R Tk)} < 2^-26  <=>  Tk < (Ea/R) / ln(A dt 2^26), with 0.1% margin
    {
        const double x = p.char_A * p.dt_sub * std::ldexp(1.0, 26);
        a.ch_tcold = x > 1.0 ? (float)(0.999 * p.char_Ea_over_R / std::log(x) - 273.15) : 3.0e38f;
    }
    a.plateau = (flags & BD_SUB_PLATEAU) ? 1 : 0;
    a.arr = ((flags & BD_SUB_ARRHENIUS) && p.has_arrhenius) ? 1 : 0;
    a.chr = ((flags & BD_SUB_CHAR) && p.has_char) ? 1 : 0;
    a.maskq = (flags & BD_SUB_MASK_Q) ? 1 : 0;
    a.vol_dt = c->g.dx * c->g.dy * c->g.dz * p.dt_sub;
}

// Pick the tile shape once per context by timing the candidates on this grid
// (a sub-step is idempotent only on scratch buffers, so tuning runs on T2 -> staging).
// *acc += the ledger slots (one wave, a fixed order), clearing them.
__global__ __launch_bounds__(64) void k_ledger_sum(double *__restrict__ part, double *__restrict__ acc)
{
    double v = part[threadIdx.x * LEDGER_STRIDE];
    part[threadIdx.x * LEDGER_STRIDE] = 0.0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (threadIdx.x == 0) *acc += v;
}

constexpr int TUNE_REPS = 6;  // timed launches per shape (after one warm-up launch)

static int autotune(Ctx *c)
{
    if (c->tuned) return BD_OK;
    c->tuned = true;
    if (c->tile_pinned) {  // fixed one-step tile, no pairs
        c->use_two = false;
        return BD_OK;
    }
    if (c->nvox < (1 << 20)) {  // small grids: fixed shapes, nothing to win
        c->tile_by = 4;
        c->tile_xc = 8;
        c->tile2_by = 8;
        c->tile2_xc = 16;
        c->use_two = c->two_step_allowed;
        if (c->use_two) {
            BD_TRY(ensure(c, (void **)&c->Hb, (size_t)c->nvox * 4));
            BD_TRY(ensure(c, (void **)&c->flagsb, (size_t)c->nvox));
        }
        return BD_OK;
    }
    // Staging: T [0, 4n), H [4n, 8n), chi [8n, 12n), flags [12n, 13n), the
    // pair kernel's T output [13n, 17n) and Omega [17n, 21n) -- scratch copies, so
    // the tuning launches run the epoch's per-cell work (plateau, char, Arrhenius)
    // and leave the state untouched.  Sized once, before any pointer into it is taken.
    const size_t n = (size_t)c->nvox;
    BD_TRY(ensure_staging(c, n * 21));
    float *Tscr = (float *)c->staging;
    float *Hscr = (float *)((char *)c->staging + n * 4);
    float *Cscr = (float *)((char *)c->staging + n * 8);
    uint8_t *Fscr = (uint8_t *)((char *)c->staging + n * 12);
    float *Oscr = (float *)((char *)c->staging + n * 17);
    const bool chr = c->chi && c->phys.has_char;
    const bool arr = c->Om && c->phys.has_arrhenius;
    BD_HIP(hipMemcpyAsync(Tscr, c->T, n * 4, hipMemcpyDeviceToDevice, c->stream));
    BD_HIP(hipMemcpyAsync(Hscr, c->H, n * 4, hipMemcpyDeviceToDevice, c->stream));
    BD_HIP(hipMemcpyAsync(Fscr, c->flags, n, hipMemcpyDeviceToDevice, c->stream));
    if (chr) BD_HIP(hipMemcpyAsync(Cscr, c->chi, n * 4, hipMemcpyDeviceToDevice, c->stream));
    if (arr) BD_HIP(hipMemcpyAsync(Oscr, c->Om, n * 4, hipMemcpyDeviceToDevice, c->stream));
    const int bys[] = {2, 4, 8};
    const int xcs[] = {8, 16, 32, 64};
    float tms[3][4];
    float best = 1e30f;
    for (int ib = 0; ib < 3; ++ib)
        for (int ix = 0; ix < 4; ++ix) {
            c->tile_xc = xcs[ix];
            SubArgs a{};
            fill_args(c, a, 1, BD_SUB_PLATEAU | BD_SUB_CHAR | BD_SUB_MASK_Q);
            a.Tin = Tscr;
            a.Tout = c->T2;
            a.H = Hscr;
            a.flags = Fscr;
            a.chi = Cscr;
            a.chr = chr ? 1 : 0;
            a.Om = Oscr;
            a.arr = arr ? 1 : 0;
            float ms = 0.f;
            BD_TRY(launch_one(c, a, bys[ib]));  // warm
            BD_HIP(hipEventRecord(c->ev0, c->stream));
            for (int r = 0; r < TUNE_REPS; ++r) BD_TRY(launch_one(c, a, bys[ib]));
            BD_HIP(hipEventRecord(c->ev1, c->stream));
            BD_HIP(hipEventSynchronize(c->ev1));
            BD_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            tms[ib][ix] = ms;
            best = std::min(best, ms);
        }
    // among shapes within 1 % of the fastest, the tallest tile (fewest y-halo
    // re-reads of T and k), then the longest x march: timing noise does not
    // flip the choice between runs
    int bb = 4, bx = 16;
    for (int ib = 2; ib >= 0; --ib) {
        int got = -1;
        for (int ix = 3; ix >= 0; --ix)
            if (tms[ib][ix] <= 1.01f * best) {
                got = ix;
                break;
            }
        if (got >= 0) {
            bb = bys[ib];
            bx = xcs[got];
            break;
        }
    }
    c->tile_by = bb;
    c->tile_xc = bx;
    // two-sub-step kernel: tile shape, and whether it beats two one-step launches
    float *Hscr2 = (float *)c->T2;  // output H of the tuning launches (T2 is scratch between sub-steps)
    BD_HIP(hipMemcpyAsync(Fscr, c->flags, n, hipMemcpyDeviceToDevice, c->stream));
    const int bys2[] = {8, 16};
    const int xcs2[] = {16, 32, 64};
    float best2 = 1e30f;
    int bb2 = 16, bx2 = 32;
    for (int by : bys2)
        for (int xc : xcs2) {
            SubArgs a{};
            fill_args(c, a, 1, BD_SUB_PLATEAU | BD_SUB_CHAR | BD_SUB_MASK_Q);
            a.Tin = Tscr;
            a.Tout = (float *)((char *)c->staging + n * 13);
            a.H_in = Hscr;
            a.H = Hscr2;
            a.flags_in = c->flags;
            a.flags = Fscr;
            a.chi = Cscr;
            a.chr = chr ? 1 : 0;
            a.Om = Oscr;
            a.arr = arr ? 1 : 0;
            float ms = 0.f;
            BD_TRY(launch_two(c, a, by, xc));
            BD_HIP(hipEventRecord(c->ev0, c->stream));
            for (int r = 0; r < TUNE_REPS; ++r) BD_TRY(launch_two(c, a, by, xc));
            BD_HIP(hipEventRecord(c->ev1, c->stream));
            BD_HIP(hipEventSynchronize(c->ev1));
            BD_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            if (c->tune_log) std::fprintf(stderr, "tune pair %dx%d: %.1f us\n", by, xc, ms * 1000.f / TUNE_REPS);
            if (ms < best2) {
                best2 = ms;
                bb2 = by;
                bx2 = xc;
            }
        }
    c->tile2_by = bb2;
    c->tile2_xc = bx2;
    // TUNE_REPS one-step vs TUNE_REPS two-step launches
    c->use_two = c->two_step_allowed && (c->two_step_forced || best2 < 2.f * best);
    if (c->tune_log) {
        for (int ib = 0; ib < 3; ++ib)
            for (int ix = 0; ix < 4; ++ix)
                std::fprintf(stderr, "tune %dx%d: %.1f us\n", bys[ib], xcs[ix], tms[ib][ix] * 1000.f / TUNE_REPS);
        std::fprintf(stderr, "tune pair best %dx%d: %.1f us per launch (two sub-steps)%s\n", bb2, bx2,
                     best2 * 1000.f / TUNE_REPS, c->use_two ? ", used" : "");
    }
    if (c->use_two) {
        // the second H / flags buffers exist before any epoch is captured
        BD_TRY(ensure(c, (void **)&c->Hb, (size_t)c->nvox * 4));
        BD_TRY(ensure(c, (void **)&c->flagsb, (size_t)c->nvox));
    }
    return BD_OK;
}

int launch_substeps(Ctx *c, int n, int is_active, int flags, bool ledger)
{
    // The first timing pass runs on the initial state, where the plateau and
    // char branches hardly fire; after 80 sub-steps (four epochs of 20) the pass
    // runs once more on the developed state, whose per-cell work favours other
    // shapes.  Never inside a graph capture (the pass waits on the host).
    if (c->tuned && !c->retuned && !c->tile_pinned && c->sub_launched >= 80) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        BD_HIP(hipStreamIsCapturing(c->stream, &cs));
        if (cs == hipStreamCaptureStatusNone) {
            c->tuned = false;
            c->retuned = true;
        }
    }
    BD_TRY(autotune(c));
    c->sub_launched += n;
    if (flags & BD_SUB_PEAK) BD_TRY(ensure(c, (void **)&c->Tpeak, (size_t)c->nvox * 4));
    if (flags & BD_SUB_ARRHENIUS) BD_TRY(ensure(c, (void **)&c->Om, (size_t)c->nvox * 4));
    SubArgs a{};
    fill_args(c, a, is_active, flags);
    a.Om = c->Om;
    a.Tpeak = c->Tpeak;
    if (ledger) {
        // the ledger slots (zero between calls), summed into red[8] after the launches
        BD_TRY(ensure(c, (void **)&c->led_part, sizeof(double) * LEDGER_SLOTS * LEDGER_STRIDE));
        a.ledger = c->led_part;
    }
    // pairs of sub-steps in the two-step kernel, an even number of them so the
    // H / flags buffers end where they started; the rest one at a time
    const int n_two = c->use_two ? 2 * (n / 4) : 0;
    int done = 0;
    for (int s = 0; s < n_two; ++s, done += 2) {
        a.Tin = c->T;
        a.Tout = c->T2;
        a.H_in = c->H;
        a.H = c->Hb;
        a.flags_in = c->flags;
        a.flags = c->flagsb;
        a.peak = ((flags & BD_SUB_PEAK) && done + 2 == n) ? 1 : 0;
        BD_TRY(launch_two(c, a, c->tile2_by, c->tile2_xc));
        std::swap(c->T, c->T2);
        std::swap(c->H, c->Hb);
        std::swap(c->flags, c->flagsb);
    }
    a.H = c->H;
    a.flags = c->flags;
    for (; done < n; ++done) {
        a.Tin = c->T;
        a.Tout = c->T2;
        a.peak = ((flags & BD_SUB_PEAK) && done == n - 1) ? 1 : 0;
        BD_TRY(launch_one(c, a, c->tile_by));
        std::swap(c->T, c->T2);
    }
    if (ledger && n > 0) {
        // red[8] += the slots, which are left zero
        hipLaunchKernelGGL(k_ledger_sum, dim3(1), dim3(64), 0, c->stream, c->led_part, c->red + 8);
        BD_HIP(hipGetLastError());
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
            T[i] = t + rel * __builtin_amdgcn_rcpf(cap);
            if (!(h > 0.f)) flags[i] = (uint8_t)(fl & ~F_HPOS);
        }
    }
}

__global__ void k_arrhenius(const float *__restrict__ T, float *__restrict__ Om, const uint8_t *__restrict__ flags,
                            float lnA, float EaR, float dt, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (!(flags[i] & F_ABL)) Om[i] += dt * expf(lnA - EaR * __builtin_amdgcn_rcpf(T[i] + 273.15f));
}

__global__ void k_char(const float *__restrict__ T, float *__restrict__ chi, const uint8_t *__restrict__ flags,
                       float lnA, float EaR, float dt, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t fl = flags[i];
        if ((fl & F_DRY) && !(fl & F_ABL)) {
            const float rate = expf(lnA - EaR * __builtin_amdgcn_rcpf(T[i] + 273.15f));
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
                       ln_prefactor(A), (float)EaR, (float)dt, c->nvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int launch_char(Ctx *c, double A, double EaR, double dt)
{
    hipLaunchKernelGGL(k_char, dim3(ew_blocks(c->nvox)), dim3(256), 0, c->stream, c->T, c->chi, c->flags,
                       ln_prefactor(A), (float)EaR, (float)dt, c->nvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

}  // namespace bd
