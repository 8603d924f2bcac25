// bd_fno.hip -- the absorbed-power surrogate: a 3-D Fourier neural operator on
// 64^3 tiles, channels last, trained and evaluated on the device.
//
//   v0      = x W_lift + b_lift                               (8 -> 32 channels)
//   v_{l+1} = act_l(v_l W_l + b_l + S_l(v_l)),  l = 0..3        (GELU; identity for l = 3)
//   y       = GELU(v_4 P1 + c1) p2 + c2                        (32 -> 64 -> 1)
//
// S_l: truncated transforms along z, y, x with exp(-2 pi i k n / Np), a complex
// 32x32 channel mix per kept mode, inverse transforms with exp(+2 pi i k n / Np),
// real part along z (weight 1 for kz = 0, 2 otherwise), scale 1/Np^3.
//
// Work split:
//  * per-voxel channel mixes (lift, W_l, P1 and their gradients) are dense
//    contractions and run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
//    one wave owns one (b, x, y) row of 64 voxels along z at a time, so the z
//    transform of the next layer (and its adjoint) fuses into the same pass
//    as cross-lane sums;
//  * the y / x transforms and the per-mode mixing touch only the kept modes
//    (a few MB) and run in FP32 on the vector ALUs;
//  * weight gradients are accumulated per wave in registers and reduced in a
//    fixed order (no atomics: bit-reproducible).
// Activations are stored as bf16, gradients with respect to activations as f32.
#include "bd_internal.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace bd {
namespace fno {

constexpr int T = 64, C = 32, H = 64, CIN = 8, L = 4;
constexpr int JMAX = 16;  // 2 * Kz <= 16
constexpr int ROW_THREADS = 256, ROW_WAVES = ROW_THREADS / 64;
constexpr int PS_LAYER = C * C + C;              // dW, db
constexpr int PS_LIFT = CIN * C + C;             // dW_lift, db_lift
constexpr int PS_PROJ = C * H + H + H + 1;       // dP1, dc1, dp2, dc2
constexpr int PS_MAX = PS_PROJ;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 as_bf(uint4 q) { return __builtin_bit_cast(bf16x8, q); }
__device__ __forceinline__ bf16x8 pack8(const float *f)
{
    u16x8 u;
    u[0] = f2bf(f[0]); u[1] = f2bf(f[1]); u[2] = f2bf(f[2]); u[3] = f2bf(f[3]);
    u[4] = f2bf(f[4]); u[5] = f2bf(f[5]); u[6] = f2bf(f[6]); u[7] = f2bf(f[7]);
    return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.39894228040143268f;
}
__device__ __forceinline__ float xsum16_32(float s)
{
    // sum over the four lanes that share (lane & 15); fixed order
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    return s;
}
__device__ __forceinline__ float xsum_lo(float s)
{
    // sum over the sixteen lanes that share (lane >> 4); fixed order
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    s += __shfl_xor(s, 4);
    s += __shfl_xor(s, 8);
    return s;
}

// MFMA 16x16x32 lane layout (checked on gfx950): with lo = lane & 15, hi = lane >> 4,
//   A[i][k]: lane (i, k/8) = (lo, hi), element k % 8      (row i = lo, k = 8 hi + e)
//   B[k][j]: lane (j, k/8) = (lo, hi), element k % 8      (col j = lo, k = 8 hi + e)
//   D[i][j]: lane (j, i/4) = (lo, hi), element i % 4      (col j = lo, row i = 4 hi + e)

// ---------------------------------------------------------------------------
// forward row kernel: lift (LIFT) or one Fourier layer
struct RowFwdArgs {
    const float *x;       // LIFT: [rows][64][CIN] f32
    const uint16_t *v;    // layer input [rows][64][C] bf16
    const float *U;       // inverse-transformed spectral term [rows][J][C] (layer only)
    const float *W;       // [K][C] f32 (K = CIN for the lift, C for a layer)
    const float *bias;    // [C]
    const float *Gz;      // [64][J]
    const float *Fz;      // [J][64]
    uint16_t *vout;       // [rows][64][C] bf16
    float *Znext;         // z transform of vout, [rows][J][C] (null for the last layer)
    int64_t rows;
    int J;
};

template <bool LIFT, bool ACT>
__global__ __launch_bounds__(ROW_THREADS) void k_row_fwd(RowFwdArgs a)
{
    __shared__ float sG[64 * JMAX], sF[JMAX * 64];
    const int J = a.J;
    for (int t = threadIdx.x; t < 64 * J; t += blockDim.x) {
        sG[t] = a.Gz[t];
        sF[t] = a.Fz[t];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
    bf16x8 bw[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        float w[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int i = 8 * hi + e;
            w[e] = LIFT ? (i < CIN ? a.W[i * C + 16 * nt + lo] : 0.f) : a.W[i * C + 16 * nt + lo];
        }
        bw[nt] = pack8(w);
    }
    const float bias0 = a.bias[lo], bias1 = a.bias[16 + lo];
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t r = wave; r < a.rows; r += nwave) {
        f32x4 acc[4][2];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            bf16x8 af;
            const int z = 16 * mt + lo;
            if (LIFT) {
                float xv[8];
                const float *xp = a.x + ((r * 64 + z) * CIN);
#pragma unroll
                for (int e = 0; e < 8; ++e) xv[e] = hi == 0 ? xp[e] : 0.f;
                af = pack8(xv);
            } else {
                af = as_bf(*(const uint4 *)(a.v + ((r * 64 + z) * C + 8 * hi)));
            }
            const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
            acc[mt][0] = mfma(af, bw[0], zero);
            acc[mt][1] = mfma(af, bw[1], zero);
        }
        float u[JMAX][2];
#pragma unroll
        for (int j = 0; j < JMAX; ++j) {
            u[j][0] = u[j][1] = 0.f;
            if (!LIFT && j < J) {
                u[j][0] = a.U[(r * J + j) * C + lo];
                u[j][1] = a.U[(r * J + j) * C + 16 + lo];
            }
        }
        float outv[4][2][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int z = 16 * mt + 4 * hi + e, o = 16 * nt + lo;
                    float p = acc[mt][nt][e] + (nt ? bias1 : bias0);
                    if (!LIFT) {
#pragma unroll
                        for (int j = 0; j < JMAX; ++j)
                            if (j < J) p += sG[z * J + j] * u[j][nt];
                    }
                    const float q = ACT ? gelu(p) : p;
                    const uint16_t qb = f2bf(q);
                    a.vout[(r * 64 + z) * C + o] = qb;
                    outv[mt][nt][e] = bf2f(qb);
                }
        if (a.Znext) {
            for (int j = 0; j < J; ++j)
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) {
                    float s = 0.f;
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                        for (int e = 0; e < 4; ++e) s += sF[j * 64 + 16 * mt + 4 * hi + e] * outv[mt][nt][e];
                    s = xsum16_32(s);
                    if (hi == (j & 3)) a.Znext[(r * J + j) * C + 16 * nt + lo] = s;
                }
        }
    }
}

// ---------------------------------------------------------------------------
// backward row kernel.  MODE 0: GELU layer, 1: identity layer, 2: lift.
// g_out = gvp_next + Fz^T gZ_next is the gradient with respect to this pass's
// output; writes the pointwise-path input gradient gvp = g W^T, the spectral
// gradient gU = Gz^T g, and per-wave partial dW, db.
struct RowBwdArgs {
    const float *x;        // lift input
    const uint16_t *v;     // layer input bf16
    const float *U;        // layer spectral term (to recompute the pre-activation)
    const float *W;
    const float *bias;
    const float *Gz;
    const float *Fz;
    const float *gvp_next; // [rows][64][C]
    const float *gZ_next;  // [rows][J][C] or null
    float *gvp;            // [rows][64][C] (layers)
    float *gU;             // [rows][J][C] (layers)
    float *part;           // [nwave][PS]
    int64_t rows;
    int J;
};

template <int MODE>
__global__ __launch_bounds__(ROW_THREADS) void k_row_bwd(RowBwdArgs a)
{
    constexpr bool LIFT = MODE == 2;
    constexpr int PS = LIFT ? PS_LIFT : PS_LAYER;
    __shared__ float sG[64 * JMAX], sF[JMAX * 64];
    __shared__ __attribute__((aligned(16))) uint16_t sVT[ROW_WAVES][C][64];   // input, transposed [i][z]
    __shared__ __attribute__((aligned(16))) uint16_t sGP[ROW_WAVES][64][C];   // g [z][o]
    __shared__ __attribute__((aligned(16))) uint16_t sGT[ROW_WAVES][C][64];   // g transposed [o][z]
    const int J = a.J;
    for (int t = threadIdx.x; t < 64 * J; t += blockDim.x) {
        sG[t] = a.Gz[t];
        sF[t] = a.Fz[t];
    }
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
    if (LIFT)
        for (int t = lane; t < C * 64; t += 64) (&sVT[w][0][0])[t] = 0;  // rows i >= CIN stay zero
    __syncthreads();
    bf16x8 bw[2], bwt[2];
    float bias0 = 0.f, bias1 = 0.f;
    if (!LIFT) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            float f[8], g[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                f[e] = a.W[(8 * hi + e) * C + 16 * nt + lo];   // B of v W: k = i, col = o
                g[e] = a.W[(16 * nt + lo) * C + 8 * hi + e];   // B of g W^T: k = o, col = i
            }
            bw[nt] = pack8(f);
            bwt[nt] = pack8(g);
        }
        bias0 = a.bias[lo];
        bias1 = a.bias[16 + lo];
    }
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 dW[2][2] = {{zero, zero}, {zero, zero}};
    float db[2] = {0.f, 0.f};
    const int64_t step = (int64_t)gridDim.x * ROW_WAVES;
    for (int64_t base = (int64_t)blockIdx.x * ROW_WAVES; base < a.rows; base += step) {
        const int64_t r = base + w;
        const bool valid = r < a.rows;
        float g[4][2][4];
        if (valid) {
            f32x4 acc[4][2];
            if (LIFT) {
                // lane = z: stage x^T
                const float *xp = a.x + (r * 64 + lane) * CIN;
                for (int i = 0; i < CIN; ++i) sVT[w][i][lane] = f2bf(xp[i]);
            } else {
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    const int z = 16 * mt + lo;
                    const uint4 q = *(const uint4 *)(a.v + ((r * 64 + z) * C + 8 * hi));
                    const u16x8 qs = __builtin_bit_cast(u16x8, q);
#pragma unroll
                    for (int e = 0; e < 8; ++e) sVT[w][8 * hi + e][z] = qs[e];
                    if (MODE == 0) {
                        const bf16x8 af = as_bf(q);
                        acc[mt][0] = mfma(af, bw[0], zero);
                        acc[mt][1] = mfma(af, bw[1], zero);
                    }
                }
            }
            float u[JMAX][2], gz[JMAX][2];
#pragma unroll
            for (int j = 0; j < JMAX; ++j) {
                u[j][0] = u[j][1] = gz[j][0] = gz[j][1] = 0.f;
                if (j < J) {
                    if (MODE == 0) {
                        u[j][0] = a.U[(r * J + j) * C + lo];
                        u[j][1] = a.U[(r * J + j) * C + 16 + lo];
                    }
                    if (a.gZ_next) {
                        gz[j][0] = a.gZ_next[(r * J + j) * C + lo];
                        gz[j][1] = a.gZ_next[(r * J + j) * C + 16 + lo];
                    }
                }
            }
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int z = 16 * mt + 4 * hi + e, o = 16 * nt + lo;
                        float go = a.gvp_next[(r * 64 + z) * C + o];
                        if (a.gZ_next) {
#pragma unroll
                            for (int j = 0; j < JMAX; ++j)
                                if (j < J) go += sF[j * 64 + z] * gz[j][nt];
                        }
                        float gg = go;
                        if (MODE == 0) {
                            float p = acc[mt][nt][e] + (nt ? bias1 : bias0);
#pragma unroll
                            for (int j = 0; j < JMAX; ++j)
                                if (j < J) p += sG[z * J + j] * u[j][nt];
                            gg = go * gelu_d(p);
                        }
                        g[mt][nt][e] = gg;
                        db[nt] += gg;
                        const uint16_t gb = f2bf(gg);
                        sGP[w][z][o] = gb;
                        sGT[w][o][z] = gb;
                    }
            if (!LIFT) {
                for (int j = 0; j < J; ++j)
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
                        float s = 0.f;
#pragma unroll
                        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                            for (int e = 0; e < 4; ++e) s += sG[(16 * mt + 4 * hi + e) * J + j] * g[mt][nt][e];
                        s = xsum16_32(s);
                        if (hi == (j & 3)) a.gU[(r * J + j) * C + 16 * nt + lo] = s;
                    }
            }
        }
        __syncthreads();
        if (valid) {
            // dW[i][o] += sum_z v[z][i] g[z][o]
            const int MT = LIFT ? 1 : 2;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    const bf16x8 af = as_bf(*(const uint4 *)&sVT[w][16 * mt + lo][32 * ks + 8 * hi]);
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
                        const bf16x8 bf = as_bf(*(const uint4 *)&sGT[w][16 * nt + lo][32 * ks + 8 * hi]);
                        dW[mt][nt] = mfma(af, bf, dW[mt][nt]);
                    }
                }
            if (!LIFT) {
                // gvp[z][i] = sum_o g[z][o] W[i][o]
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    const bf16x8 af = as_bf(*(const uint4 *)&sGP[w][16 * mt + lo][8 * hi]);
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
                        const f32x4 d = mfma(af, bwt[nt], zero);
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            a.gvp[(r * 64 + 16 * mt + 4 * hi + e) * C + 16 * nt + lo] = d[e];
                    }
                }
            }
        }
        __syncthreads();
    }
    // per-wave partials: dW (row i = 16 mt + 4 hi + e, col o = 16 nt + lo), then db
    const int64_t wid = (int64_t)blockIdx.x * ROW_WAVES + w;
    float *pp = a.part + wid * PS;
    const int KI = LIFT ? CIN : C;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = 16 * mt + 4 * hi + e;
                if (i < KI) pp[i * C + 16 * nt + lo] = dW[mt][nt][e];
            }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const float s = xsum16_32(db[nt]);
        if (hi == 0) pp[KI * C + 16 * nt + lo] = s;
    }
}

// ---------------------------------------------------------------------------
// projection: y = GELU(v P1 + c1) p2 + c2
struct ProjArgs {
    const uint16_t *v;    // [rows][64][C]
    const float *P1;      // [C][H]
    const float *c1;      // [H]
    const float *p2;      // [H]
    const float *c2;      // [1]
    float *y;             // [rows][64]
    const float *gy;      // backward: [rows][64]
    float *gv;            // backward: [rows][64][C]
    float *part;          // backward: [nwave][PS_PROJ]
    int64_t rows;
};

__global__ __launch_bounds__(ROW_THREADS) void k_proj_fwd(ProjArgs a)
{
    const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
    bf16x8 bp[4];
    float cb[4], pv[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        float f[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) f[e] = a.P1[(8 * hi + e) * H + 16 * nt + lo];
        bp[nt] = pack8(f);
        cb[nt] = a.c1[16 * nt + lo];
        pv[nt] = a.p2[16 * nt + lo];
    }
    const float c2 = a.c2[0];
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t r = wave; r < a.rows; r += nwave) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const bf16x8 af = as_bf(*(const uint4 *)(a.v + ((r * 64 + 16 * mt + lo) * C + 8 * hi)));
            float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const f32x4 d = mfma(af, bp[nt], zero);
#pragma unroll
                for (int e = 0; e < 4; ++e) s[e] += gelu(d[e] + cb[nt]) * pv[nt];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float t = xsum_lo(s[e]);
                if (lo == e) a.y[r * 64 + 16 * mt + 4 * hi + e] = t + c2;
            }
        }
    }
}

constexpr int PROJ_WAVES = 2;  // 20 KB of LDS staging per wave
__global__ __launch_bounds__(64 * PROJ_WAVES) void k_proj_bwd(ProjArgs a)
{
    __shared__ __attribute__((aligned(16))) uint16_t sVT[PROJ_WAVES][C][64];
    __shared__ __attribute__((aligned(16))) uint16_t sGH[PROJ_WAVES][64][H];
    __shared__ __attribute__((aligned(16))) uint16_t sGHT[PROJ_WAVES][H][64];
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63, lo = lane & 15, hi = lane >> 4;
    bf16x8 bp[4], bpt[2][2];
    float cb[4], pv[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        float f[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) f[e] = a.P1[(8 * hi + e) * H + 16 * nt + lo];
        bp[nt] = pack8(f);
        cb[nt] = a.c1[16 * nt + lo];
        pv[nt] = a.p2[16 * nt + lo];
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            float f[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) f[e] = a.P1[(16 * nt + lo) * H + 32 * ks + 8 * hi + e];  // k = n, col = c
            bpt[nt][ks] = pack8(f);
        }
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 dP[2][4];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 4; ++j) dP[i][j] = zero;
    float dc1[4] = {0.f, 0.f, 0.f, 0.f}, dp2[4] = {0.f, 0.f, 0.f, 0.f}, dc2 = 0.f;
    const int64_t step = (int64_t)gridDim.x * PROJ_WAVES;
    for (int64_t base = (int64_t)blockIdx.x * PROJ_WAVES; base < a.rows; base += step) {
        const int64_t r = base + w;
        const bool valid = r < a.rows;
        if (valid) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                const int zr = 16 * mt + lo;
                const uint4 q = *(const uint4 *)(a.v + ((r * 64 + zr) * C + 8 * hi));
                const u16x8 qs = __builtin_bit_cast(u16x8, q);
#pragma unroll
                for (int e = 0; e < 8; ++e) sVT[w][8 * hi + e][zr] = qs[e];
                const bf16x8 af = as_bf(q);
                float gyv[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) gyv[e] = a.gy[r * 64 + 16 * mt + 4 * hi + e];
                if (lo == 0)
#pragma unroll
                    for (int e = 0; e < 4; ++e) dc2 += gyv[e];
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    const f32x4 d = mfma(af, bp[nt], zero);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float hp = d[e] + cb[nt];
                        dp2[nt] += gelu(hp) * gyv[e];
                        const float gh = gyv[e] * pv[nt] * gelu_d(hp);
                        dc1[nt] += gh;
                        const int z = 16 * mt + 4 * hi + e, n = 16 * nt + lo;
                        const uint16_t gb = f2bf(gh);
                        sGH[w][z][n] = gb;
                        sGHT[w][n][z] = gb;
                    }
                }
            }
        }
        __syncthreads();
        if (valid) {
            // dP1[c][n] += sum_z v[z][c] gh[z][n]
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    const bf16x8 af = as_bf(*(const uint4 *)&sVT[w][16 * mt + lo][32 * ks + 8 * hi]);
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt) {
                        const bf16x8 bf = as_bf(*(const uint4 *)&sGHT[w][16 * nt + lo][32 * ks + 8 * hi]);
                        dP[mt][nt] = mfma(af, bf, dP[mt][nt]);
                    }
                }
            // gv[z][c] = sum_n gh[z][n] P1[c][n]
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                f32x4 d0 = zero, d1 = zero;
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    const bf16x8 af = as_bf(*(const uint4 *)&sGH[w][16 * mt + lo][32 * ks + 8 * hi]);
                    d0 = mfma(af, bpt[0][ks], d0);
                    d1 = mfma(af, bpt[1][ks], d1);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t o = (r * 64 + 16 * mt + 4 * hi + e) * C;
                    a.gv[o + lo] = d0[e];
                    a.gv[o + 16 + lo] = d1[e];
                }
            }
        }
        __syncthreads();
    }
    const int64_t wid = (int64_t)blockIdx.x * PROJ_WAVES + w;
    float *pp = a.part + wid * PS_PROJ;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int e = 0; e < 4; ++e) pp[(16 * mt + 4 * hi + e) * H + 16 * nt + lo] = dP[mt][nt][e];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const float s1 = xsum16_32(dc1[nt]), s2 = xsum16_32(dp2[nt]);
        if (hi == 0) {
            pp[C * H + 16 * nt + lo] = s1;
            pp[C * H + H + 16 * nt + lo] = s2;
        }
    }
    float s = xsum16_32(dc2);
    s = xsum_lo(s);
    if (lane == 0) pp[C * H + 2 * H] = s;
}

// fixed-order sum of per-wave partials, out[k] = sum_w part[w * ps + k], in two
// stages: chunk c of RP_CHUNK waves -> mid[c][k] (FP64), then the chunks in order
constexpr int RP_CHUNK = 64;
__global__ void k_reduce_parts1(const float *__restrict__ part, int64_t nwave, int ps, int n, double *__restrict__ mid)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t w0 = (int64_t)blockIdx.y * RP_CHUNK;
    const int64_t w1 = w0 + RP_CHUNK < nwave ? w0 + RP_CHUNK : nwave;
    double s = 0.0;
    for (int64_t w = w0; w < w1; ++w) s += (double)part[w * ps + k];
    mid[(int64_t)blockIdx.y * n + k] = s;
}

__global__ void k_reduce_parts2(const double *__restrict__ mid, int nchunk, int n, float *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += mid[(int64_t)c * n + k];
    out[k] = (float)s;
}

// ---------------------------------------------------------------------------
// truncated transform along one axis:
//   out[o][k][q][ri][e] = sum_n E[k][n] * in[o][n][q][ri][e]   (complex over ri, e < half)
__global__ void k_xform(const float *__restrict__ in, float *__restrict__ out, const float2 *__restrict__ E, int O,
                        int nin, int nout, int Q, int half, int conj)
{
    const int64_t total = (int64_t)O * nout * Q * half;
    const int64_t sn = (int64_t)Q * 2 * half;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(t % half);
        int64_t r = t / half;
        const int q = (int)(r % Q);
        r /= Q;
        const int k = (int)(r % nout);
        const int64_t o = r / nout;
        const float *ip = in + (o * nin * Q + q) * 2 * half + e;
        float sr = 0.f, si = 0.f;
        for (int n = 0; n < nin; ++n) {
            const float2 c = E[k * nin + n];
            const float er = c.x, ei = conj ? -c.y : c.y;
            const float xr = ip[n * sn], xi = ip[n * sn + half];
            sr += er * xr - ei * xi;
            si += er * xi + ei * xr;
        }
        float *op = out + ((o * nout + k) * Q + q) * 2 * half + e;
        op[0] = sr;
        op[half] = si;
    }
}

// per-mode complex channel mix M[b][m][o] = sum_i W[m][i][o] X[b][m][i];
// modes m = (p, q, kz) with X layout [B][Kx][Ky][2][Kz][C]
__global__ void k_mix(const float *__restrict__ X, const float *__restrict__ Wt, float *__restrict__ M, int B,
                      int nPQ, int Kz)
{
    const int half = Kz * C;
    const int64_t total = (int64_t)B * nPQ * half;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(t % C);
        const int kz = (int)((t / C) % Kz);
        const int64_t bpq = t / half;
        const int pq = (int)(bpq % nPQ);
        const float *xp = X + bpq * 2 * half + kz * C;
        const float *wp = Wt + ((int64_t)(pq * Kz + kz) * C * C) * 2 + 2 * o;
        float sr = 0.f, si = 0.f;
        for (int i = 0; i < C; ++i) {
            const float xr = xp[i], xi = xp[half + i];
            const float wr = wp[2 * i * C], wi = wp[2 * i * C + 1];
            sr += wr * xr - wi * xi;
            si += wr * xi + wi * xr;
        }
        float *mp = M + bpq * 2 * half + kz * C + o;
        mp[0] = sr;
        mp[half] = si;
    }
}

// gW[m][i][o] = sum_b conj(X[b][m][i]) gM[b][m][o]   (written as re, im pairs)
__global__ void k_mix_bwd_w(const float *__restrict__ X, const float *__restrict__ gM, float *__restrict__ gW, int B,
                            int nPQ, int Kz)
{
    const int half = Kz * C;
    const int64_t total = (int64_t)nPQ * Kz * C * C;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(t % C);
        const int i = (int)((t / C) % C);
        const int kz = (int)((t / (C * C)) % Kz);
        const int pq = (int)(t / ((int64_t)C * C * Kz));
        float gr = 0.f, gi = 0.f;
        for (int b = 0; b < B; ++b) {
            const int64_t base = ((int64_t)b * nPQ + pq) * 2 * half + kz * C;
            const float xr = X[base + i], xi = X[base + half + i];
            const float mr = gM[base + o], mi = gM[base + half + o];
            gr += xr * mr + xi * mi;
            gi += xr * mi - xi * mr;
        }
        gW[2 * t] = gr;
        gW[2 * t + 1] = gi;
    }
}

// gX[b][m][i] = sum_o conj(W[m][i][o]) gM[b][m][o]
__global__ void k_mix_bwd_x(const float *__restrict__ Wt, const float *__restrict__ gM, float *__restrict__ gX, int B,
                            int nPQ, int Kz)
{
    const int half = Kz * C;
    const int64_t total = (int64_t)B * nPQ * half;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % C);
        const int kz = (int)((t / C) % Kz);
        const int64_t bpq = t / half;
        const int pq = (int)(bpq % nPQ);
        const float *mp = gM + bpq * 2 * half + kz * C;
        const float *wp = Wt + (((int64_t)(pq * Kz + kz) * C + i) * C) * 2;
        float sr = 0.f, si = 0.f;
        for (int o = 0; o < C; ++o) {
            const float wr = wp[2 * o], wi = wp[2 * o + 1];
            const float mr = mp[o], mi = mp[half + o];
            sr += wr * mr + wi * mi;
            si += wr * mi - wi * mr;
        }
        float *xp = gX + bpq * 2 * half + kz * C + i;
        xp[0] = sr;
        xp[half] = si;
    }
}

// ---------------------------------------------------------------------------
// weighted log-space squared error: L = sum w (y - t)^2 / B, gy = 2 w (y - t) / B
// (w normalised by the caller to sum to one per tile)
__global__ void k_loss(const float *__restrict__ y, const float *__restrict__ tg, const float *__restrict__ wt,
                       float *__restrict__ gy, int64_t n, float invB, double *__restrict__ part)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float d = y[i] - tg[i];
        gy[i] = 2.f * wt[i] * d * invB;
        acc += (double)wt[i] * (double)d * (double)d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_loss_final(const double *__restrict__ part, int nb, double invB, double *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < nb; ++i) s += part[i];
        out[0] = s * invB;
    }
}

// AdamW with bias correction, decoupled weight decay
__global__ void k_adamw(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                        int64_t n, float lr, float b1, float b2, float eps, float wd, float bc1, float bc2)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * ((mi / bc1) / (sqrtf(vi / bc2) + eps) + wd * p[i]);
    }
}

__global__ void k_fill(float *__restrict__ p, float v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace fno
}  // namespace bd

// ---------------------------------------------------------------------------
// context and C-ABI
using namespace bd;
using namespace bd::fno;

struct bd_fno {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    int Kx = 0, Ky = 0, Kz = 0, Np = 0, J = 0, Bmax = 0;
    bool train = false;
    int64_t nparam = 0;
    int64_t off_lift = 0, off_spec[L] = {}, off_w[L] = {}, off_proj = 0;
    float *params = nullptr, *grads = nullptr, *am = nullptr, *av = nullptr;
    long long step = 0;
    float *Gz = nullptr, *Fz = nullptr;
    float2 *Ex = nullptr, *ExT = nullptr, *Ey = nullptr, *EyT = nullptr;
    uint16_t *v[L + 1] = {};
    float *U[L] = {}, *Xf[L] = {};
    float *Zf = nullptr, *Yf = nullptr, *Mx = nullptr, *Xi = nullptr;
    float *gvp[2] = {}, *gZ = nullptr, *gU = nullptr, *gA = nullptr;
    float *part = nullptr;
    double *mid = nullptr;   // chunk sums of the partials
    double *dred = nullptr;
    double *hred = nullptr;
    int last_B = 0;
    std::vector<void *> allocs;
    // forward passes repeated on the same buffers (the coupled loop) replay a captured graph
    const float *fw_x = nullptr;
    float *fw_y = nullptr;
    int fw_B = 0, fw_repeats = 0;
    hipGraph_t fw_graph = nullptr;
    hipGraphExec_t fw_exec = nullptr;
};

namespace {

int fno_alloc(bd_fno *f, void **p, size_t bytes)
{
    BD_HIP(hipMalloc(p, bytes));
    f->allocs.push_back(*p);
    BD_HIP(hipMemsetAsync(*p, 0, bytes, f->stream));
    return BD_OK;
}

// Row-kernel grids: up to 2 048 blocks of four waves, i.e. two waves per SIMD on 256 CUs (the
// kernels hold three or four); with 1 024 blocks each wave walked about 2.5 rows of a
// 2.56-window pass one after another, and the forward pass took 16 % longer.
int row_grid(int64_t rows) { return (int)std::min<int64_t>((rows + ROW_WAVES - 1) / ROW_WAVES, 2048); }
int proj_grid(int64_t rows) { return (int)std::min<int64_t>((rows + PROJ_WAVES - 1) / PROJ_WAVES, 2048); }

int xform(bd_fno *f, const float *in, float *out, const float2 *E, int O, int nin, int nout, int Q, bool conj)
{
    const int half = f->Kz * C;
    const int64_t total = (int64_t)O * nout * Q * half;
    const int nb = (int)std::min<int64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_xform, dim3(nb), dim3(256), 0, f->stream, in, out, E, O, nin, nout, Q, half, conj ? 1 : 0);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int reduce_parts(bd_fno *f, int64_t nwave, int ps, int n, float *out)
{
    const int nchunk = (int)((nwave + RP_CHUNK - 1) / RP_CHUNK);
    hipLaunchKernelGGL(k_reduce_parts1, dim3((n + 255) / 256, nchunk), dim3(256), 0, f->stream, f->part, nwave, ps, n,
                       f->mid);
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_parts2, dim3((n + 255) / 256), dim3(256), 0, f->stream, f->mid, nchunk, n, out);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int forward_impl(bd_fno *f, const float *x, int B, float *y)
{
    const int64_t rows = (int64_t)B * T * T;
    const int J = f->J, half = f->Kz * C, nPQ = f->Kx * f->Ky;
    const int g = row_grid(rows);
    const float *P = f->params;
    uint16_t *vbuf[L + 1];
    for (int l = 0; l <= L; ++l) vbuf[l] = f->train ? f->v[l] : f->v[l & 1];
    {
        RowFwdArgs a{};
        a.x = x;
        a.W = P + f->off_lift;
        a.bias = P + f->off_lift + CIN * C;
        a.Gz = f->Gz;
        a.Fz = f->Fz;
        a.vout = vbuf[0];
        a.Znext = f->Zf;
        a.rows = rows;
        a.J = J;
        hipLaunchKernelGGL((k_row_fwd<true, false>), dim3(g), dim3(ROW_THREADS), 0, f->stream, a);
        BD_HIP(hipGetLastError());
    }
    for (int l = 0; l < L; ++l) {
        float *U = f->train ? f->U[l] : f->U[0];
        float *Xf = f->train ? f->Xf[l] : f->Xf[0];
        BD_TRY(xform(f, f->Zf, f->Yf, f->Ey, B * T, T, f->Ky, 1, false));
        BD_TRY(xform(f, f->Yf, Xf, f->Ex, B, T, f->Kx, f->Ky, false));
        {
            const int64_t total = (int64_t)B * nPQ * half;
            const int nb = (int)std::min<int64_t>((total + 255) / 256, 8192);
            hipLaunchKernelGGL(k_mix, dim3(nb), dim3(256), 0, f->stream, Xf, P + f->off_spec[l], f->Mx, B, nPQ, f->Kz);
            BD_HIP(hipGetLastError());
        }
        BD_TRY(xform(f, f->Mx, f->Xi, f->ExT, B, f->Kx, T, f->Ky, true));
        BD_TRY(xform(f, f->Xi, U, f->EyT, B * T, f->Ky, T, 1, true));
        RowFwdArgs a{};
        a.v = vbuf[l];
        a.U = U;
        a.W = P + f->off_w[l];
        a.bias = P + f->off_w[l] + C * C;
        a.Gz = f->Gz;
        a.Fz = f->Fz;
        a.vout = vbuf[l + 1];
        a.Znext = l < L - 1 ? f->Zf : nullptr;
        a.rows = rows;
        a.J = J;
        if (l < L - 1)
            hipLaunchKernelGGL((k_row_fwd<false, true>), dim3(g), dim3(ROW_THREADS), 0, f->stream, a);
        else
            hipLaunchKernelGGL((k_row_fwd<false, false>), dim3(g), dim3(ROW_THREADS), 0, f->stream, a);
        BD_HIP(hipGetLastError());
    }
    ProjArgs pa{};
    pa.v = vbuf[L];
    pa.P1 = P + f->off_proj;
    pa.c1 = pa.P1 + C * H;
    pa.p2 = pa.c1 + H;
    pa.c2 = pa.p2 + H;
    pa.y = y;
    pa.rows = rows;
    hipLaunchKernelGGL(k_proj_fwd, dim3(g), dim3(ROW_THREADS), 0, f->stream, pa);
    BD_HIP(hipGetLastError());
    f->last_B = B;
    return BD_OK;
}

int backward_impl(bd_fno *f, const float *x, const float *gy, int B)
{
    const int64_t rows = (int64_t)B * T * T;
    const int J = f->J, half = f->Kz * C, nPQ = f->Kx * f->Ky;
    const int g = row_grid(rows);
    const float *P = f->params;
    float *G = f->grads;
    {
        ProjArgs pa{};
        pa.v = f->v[L];
        pa.P1 = P + f->off_proj;
        pa.c1 = pa.P1 + C * H;
        pa.p2 = pa.c1 + H;
        pa.c2 = pa.p2 + H;
        pa.gy = gy;
        pa.gv = f->gvp[0];
        pa.part = f->part;
        pa.rows = rows;
        const int gp = proj_grid(rows);
        hipLaunchKernelGGL(k_proj_bwd, dim3(gp), dim3(64 * PROJ_WAVES), 0, f->stream, pa);
        BD_HIP(hipGetLastError());
        BD_TRY(reduce_parts(f, (int64_t)gp * PROJ_WAVES, PS_PROJ, PS_PROJ, G + f->off_proj));
    }
    int cur = 0;
    const float *gZnext = nullptr;
    for (int l = L - 1; l >= 0; --l) {
        RowBwdArgs a{};
        a.v = f->v[l];
        a.U = f->U[l];
        a.W = P + f->off_w[l];
        a.bias = P + f->off_w[l] + C * C;
        a.Gz = f->Gz;
        a.Fz = f->Fz;
        a.gvp_next = f->gvp[cur];
        a.gZ_next = gZnext;
        a.gvp = f->gvp[cur ^ 1];
        a.gU = f->gU;
        a.part = f->part;
        a.rows = rows;
        a.J = J;
        if (l == L - 1)
            hipLaunchKernelGGL((k_row_bwd<1>), dim3(g), dim3(ROW_THREADS), 0, f->stream, a);
        else
            hipLaunchKernelGGL((k_row_bwd<0>), dim3(g), dim3(ROW_THREADS), 0, f->stream, a);
        BD_HIP(hipGetLastError());
        BD_TRY(reduce_parts(f, (int64_t)g * ROW_WAVES, PS_LAYER, PS_LAYER, G + f->off_w[l]));
        cur ^= 1;
        // adjoint chain of the spectral term: gU -> gXi -> gM -> (gW, gXf) -> gYf -> gZ
        BD_TRY(xform(f, f->gU, f->gA, f->Ey, B * T, T, f->Ky, 1, false));
        BD_TRY(xform(f, f->gA, f->Mx, f->Ex, B, T, f->Kx, f->Ky, false));
        {
            const int64_t tw = (int64_t)nPQ * f->Kz * C * C;
            const int nb = (int)std::min<int64_t>((tw + 255) / 256, 8192);
            hipLaunchKernelGGL(k_mix_bwd_w, dim3(nb), dim3(256), 0, f->stream, f->Xf[l], f->Mx, G + f->off_spec[l], B,
                               nPQ, f->Kz);
            BD_HIP(hipGetLastError());
            const int64_t tx = (int64_t)B * nPQ * half;
            const int nb2 = (int)std::min<int64_t>((tx + 255) / 256, 8192);
            hipLaunchKernelGGL(k_mix_bwd_x, dim3(nb2), dim3(256), 0, f->stream, P + f->off_spec[l], f->Mx, f->Yf, B,
                               nPQ, f->Kz);
            BD_HIP(hipGetLastError());
        }
        BD_TRY(xform(f, f->Yf, f->gA, f->ExT, B, f->Kx, T, f->Ky, true));
        BD_TRY(xform(f, f->gA, f->gZ, f->EyT, B * T, f->Ky, T, 1, true));
        gZnext = f->gZ;
    }
    RowBwdArgs a{};
    a.x = x;
    a.Gz = f->Gz;
    a.Fz = f->Fz;
    a.gvp_next = f->gvp[cur];
    a.gZ_next = gZnext;
    a.part = f->part;
    a.rows = rows;
    a.J = J;
    hipLaunchKernelGGL((k_row_bwd<2>), dim3(g), dim3(ROW_THREADS), 0, f->stream, a);
    BD_HIP(hipGetLastError());
    BD_TRY(reduce_parts(f, (int64_t)g * ROW_WAVES, PS_LIFT, PS_LIFT, G + f->off_lift));
    return BD_OK;
}

}  // namespace

extern "C" {

int bd_fno_create(int device, const bd_fno_spec *spec, int max_batch, int train, bd_fno **out)
{
    BD_CHECK(spec && out, "bd_fno_create: null argument");
    *out = nullptr;
    BD_CHECK(spec->kx >= 2 && spec->kx % 2 == 0 && spec->kx <= T, "bd_fno_create: kx must be even, 2..64");
    BD_CHECK(spec->ky >= 2 && spec->ky % 2 == 0 && spec->ky <= T, "bd_fno_create: ky must be even, 2..64");
    BD_CHECK(spec->kz >= 1 && 2 * spec->kz <= JMAX, "bd_fno_create: kz must be 1..8");
    BD_CHECK(spec->np >= T, "bd_fno_create: np (transform length) must be >= 64");
    BD_CHECK(max_batch >= 1 && max_batch <= 64, "bd_fno_create: max_batch must be 1..64");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("bd_fno_create: no HIP device");
        return BD_ENODEV;
    }
    BD_CHECK(device >= 0 && device < ndev, "bd_fno_create: device index out of range");
    BD_HIP(hipSetDevice(device));
    bd_fno *f = new bd_fno();
    f->device = device;
    f->Kx = spec->kx;
    f->Ky = spec->ky;
    f->Kz = spec->kz;
    f->Np = spec->np;
    f->J = 2 * spec->kz;
    f->Bmax = max_batch;
    f->train = train != 0;
    int rc = BD_OK;
    auto fail = [&](int code) {
        bd_fno_destroy(f);
        return code;
    };
    if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("bd_fno_create: hipStreamCreate failed");
        delete f;
        return BD_EHIP;
    }
    // parameter layout (lift | per layer: spec, W, b | projection)
    int64_t o = 0;
    f->off_lift = o;
    o += CIN * C + C;
    const int64_t nspec = (int64_t)f->Kx * f->Ky * f->Kz * C * C * 2;
    for (int l = 0; l < L; ++l) {
        f->off_spec[l] = o;
        o += nspec;
        f->off_w[l] = o;
        o += C * C + C;
    }
    f->off_proj = o;
    o += PS_PROJ;
    f->nparam = o;
    const size_t pb = (size_t)o * sizeof(float);
    if ((rc = fno_alloc(f, (void **)&f->params, pb))) return fail(rc);
    if (f->train) {
        if ((rc = fno_alloc(f, (void **)&f->grads, pb))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->am, pb))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->av, pb))) return fail(rc);
    }
    // transform tables, computed in FP64
    {
        const int J = f->J, Np = f->Np;
        std::vector<float> gz(64 * J), fz(J * 64);
        const double tw = 2.0 * M_PI / Np;
        const double inv3 = 1.0 / ((double)Np * Np * Np);
        for (int z = 0; z < 64; ++z)
            for (int k = 0; k < f->Kz; ++k) {
                const double cz = k == 0 ? 1.0 : 2.0;
                const double ang = tw * k * z;
                gz[z * J + k] = (float)(cz * std::cos(ang) * inv3);
                gz[z * J + f->Kz + k] = (float)(-cz * std::sin(ang) * inv3);
                fz[k * 64 + z] = (float)std::cos(ang);
                fz[(f->Kz + k) * 64 + z] = (float)(-std::sin(ang));
            }
        auto table = [&](int K, std::vector<float2> &E, std::vector<float2> &ET) {
            E.resize((size_t)K * 64);
            ET.resize((size_t)K * 64);
            for (int a = 0; a < K; ++a) {
                const int k = a < K / 2 ? a : a - K;
                for (int n = 0; n < 64; ++n) {
                    const double ang = -tw * k * n;
                    const float2 c = make_float2((float)std::cos(ang), (float)std::sin(ang));
                    E[(size_t)a * 64 + n] = c;
                    ET[(size_t)n * K + a] = c;
                }
            }
        };
        std::vector<float2> ex, ext, ey, eyt;
        table(f->Kx, ex, ext);
        table(f->Ky, ey, eyt);
        if ((rc = fno_alloc(f, (void **)&f->Gz, gz.size() * 4))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->Fz, fz.size() * 4))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->Ex, ex.size() * 8))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->ExT, ext.size() * 8))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->Ey, ey.size() * 8))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->EyT, eyt.size() * 8))) return fail(rc);
        if (hipMemcpy(f->Gz, gz.data(), gz.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(f->Fz, fz.data(), fz.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(f->Ex, ex.data(), ex.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(f->ExT, ext.data(), ext.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(f->Ey, ey.data(), ey.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(f->EyT, eyt.data(), eyt.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("bd_fno_create: table upload failed");
            return fail(BD_EHIP);
        }
    }
    // activations and transform buffers
    const int64_t rows = (int64_t)f->Bmax * T * T;
    const size_t vb = (size_t)rows * 64 * C * 2;
    const size_t ub = (size_t)rows * f->J * C * 4;
    const size_t yb = (size_t)f->Bmax * T * f->Ky * f->J * C * 4;
    const size_t xb = (size_t)f->Bmax * f->Kx * f->Ky * f->J * C * 4;
    const int nv = f->train ? L + 1 : 2, nl = f->train ? L : 1;
    for (int l = 0; l < nv; ++l)
        if ((rc = fno_alloc(f, (void **)&f->v[l], vb))) return fail(rc);
    for (int l = 0; l < nl; ++l) {
        if ((rc = fno_alloc(f, (void **)&f->U[l], ub))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->Xf[l], xb))) return fail(rc);
    }
    if ((rc = fno_alloc(f, (void **)&f->Zf, ub))) return fail(rc);
    if ((rc = fno_alloc(f, (void **)&f->Yf, yb))) return fail(rc);
    if ((rc = fno_alloc(f, (void **)&f->Mx, xb))) return fail(rc);
    if ((rc = fno_alloc(f, (void **)&f->Xi, yb))) return fail(rc);
    if (f->train) {
        const size_t gb = (size_t)rows * 64 * C * 4;
        if ((rc = fno_alloc(f, (void **)&f->gvp[0], gb))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->gvp[1], gb))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->gZ, ub))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->gU, ub))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->gA, yb))) return fail(rc);
        const size_t nw = std::max((size_t)row_grid(rows) * ROW_WAVES, (size_t)proj_grid(rows) * PROJ_WAVES);
        const size_t partb = nw * PS_MAX * 4;
        if ((rc = fno_alloc(f, (void **)&f->part, partb))) return fail(rc);
        const size_t midb = ((nw + RP_CHUNK - 1) / RP_CHUNK) * PS_MAX * sizeof(double);
        if ((rc = fno_alloc(f, (void **)&f->mid, midb))) return fail(rc);
        if ((rc = fno_alloc(f, (void **)&f->dred, (1024 + 8) * sizeof(double)))) return fail(rc);
    }
    if (hipHostMalloc((void **)&f->hred, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        set_error("bd_fno_create: hipHostMalloc failed");
        return fail(BD_EHIP);
    }
    if (hipStreamSynchronize(f->stream) != hipSuccess) {
        set_error("bd_fno_create: stream synchronize failed");
        return fail(BD_EHIP);
    }
    *out = f;
    return BD_OK;
}

int bd_fno_destroy(bd_fno *f)
{
    if (!f) return BD_OK;
    (void)hipSetDevice(f->device);
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    if (f->fw_exec) (void)hipGraphExecDestroy(f->fw_exec);
    if (f->fw_graph) (void)hipGraphDestroy(f->fw_graph);
    for (void *p : f->allocs) (void)hipFree(p);
    if (f->hred) (void)hipHostFree(f->hred);
    if (f->stream && f->own_stream) (void)hipStreamDestroy(f->stream);
    delete f;
    return BD_OK;
}

int bd_fno_num_params(const bd_fno *f, int64_t *n)
{
    BD_CHECK(f && n, "bd_fno_num_params: null argument");
    *n = f->nparam;
    return BD_OK;
}

int bd_fno_set_params(bd_fno *f, const float *host, int64_t n)
{
    BD_CHECK(f && host, "bd_fno_set_params: null argument");
    BD_CHECK(n == f->nparam, "bd_fno_set_params: size mismatch");
    BD_HIP(hipSetDevice(f->device));
    BD_HIP(hipMemcpyAsync(f->params, host, (size_t)n * 4, hipMemcpyHostToDevice, f->stream));
    if (f->train) {
        BD_HIP(hipMemsetAsync(f->am, 0, (size_t)n * 4, f->stream));
        BD_HIP(hipMemsetAsync(f->av, 0, (size_t)n * 4, f->stream));
        f->step = 0;
    }
    BD_HIP(hipStreamSynchronize(f->stream));
    return BD_OK;
}

int bd_fno_get_params(bd_fno *f, float *host, int64_t n)
{
    BD_CHECK(f && host, "bd_fno_get_params: null argument");
    BD_CHECK(n == f->nparam, "bd_fno_get_params: size mismatch");
    BD_HIP(hipSetDevice(f->device));
    BD_HIP(hipMemcpyAsync(host, f->params, (size_t)n * 4, hipMemcpyDeviceToHost, f->stream));
    BD_HIP(hipStreamSynchronize(f->stream));
    return BD_OK;
}

int bd_fno_get_grads(bd_fno *f, float *host, int64_t n)
{
    BD_CHECK(f && host, "bd_fno_get_grads: null argument");
    BD_CHECK(f->train, "bd_fno_get_grads: context was created without training buffers");
    BD_CHECK(n == f->nparam, "bd_fno_get_grads: size mismatch");
    BD_HIP(hipSetDevice(f->device));
    BD_HIP(hipMemcpyAsync(host, f->grads, (size_t)n * 4, hipMemcpyDeviceToHost, f->stream));
    BD_HIP(hipStreamSynchronize(f->stream));
    return BD_OK;
}

int bd_fno_forward(bd_fno *f, const float *x_dev, int B, float *y_dev)
{
    BD_CHECK(f && x_dev && y_dev, "bd_fno_forward: null argument");
    BD_CHECK(B >= 1 && B <= f->Bmax, "bd_fno_forward: batch outside 1..max_batch");
    BD_HIP(hipSetDevice(f->device));
    const bool same = x_dev == f->fw_x && y_dev == f->fw_y && B == f->fw_B;
    f->fw_repeats = same ? f->fw_repeats + 1 : 0;
    f->fw_x = x_dev;
    f->fw_y = y_dev;
    f->fw_B = B;
    if (!same && f->fw_exec) {
        (void)hipGraphExecDestroy(f->fw_exec);
        (void)hipGraphDestroy(f->fw_graph);
        f->fw_exec = nullptr;
        f->fw_graph = nullptr;
    }
    if (f->fw_repeats == 0) return forward_impl(f, x_dev, B, y_dev);
    if (!f->fw_exec) {
        // second pass on the same buffers: capture the launch sequence once
        BD_HIP(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
        const int rc = forward_impl(f, x_dev, B, y_dev);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(f->stream, &g);
        if (rc != BD_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        BD_HIP(e);
        f->fw_graph = g;
        BD_HIP(hipGraphInstantiate(&f->fw_exec, g, nullptr, nullptr, 0));
    }
    BD_HIP(hipGraphLaunch(f->fw_exec, f->stream));
    f->last_B = B;
    return BD_OK;
}

int bd_fno_backward(bd_fno *f, const float *x_dev, const float *gy_dev, int B)
{
    BD_CHECK(f && x_dev && gy_dev, "bd_fno_backward: null argument");
    BD_CHECK(f->train, "bd_fno_backward: context was created without training buffers");
    BD_CHECK(B == f->last_B, "bd_fno_backward: batch differs from the last forward pass");
    BD_HIP(hipSetDevice(f->device));
    return backward_impl(f, x_dev, gy_dev, B);
}

int bd_fno_loss(bd_fno *f, const float *y_dev, const float *t_dev, const float *w_dev, int B, float *gy_dev,
                double *loss)
{
    BD_CHECK(f && y_dev && t_dev && w_dev && gy_dev && loss, "bd_fno_loss: null argument");
    BD_CHECK(f->train, "bd_fno_loss: context was created without training buffers");
    BD_CHECK(B >= 1 && B <= f->Bmax, "bd_fno_loss: batch outside 1..max_batch");
    BD_HIP(hipSetDevice(f->device));
    const int64_t n = (int64_t)B * T * T * T;
    const int nb = 1024;
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, f->stream, y_dev, t_dev, w_dev, gy_dev, n, 1.f / B, f->dred);
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, f->stream, f->dred, nb, 1.0 / B, f->dred + nb);
    BD_HIP(hipGetLastError());
    BD_HIP(hipMemcpyAsync(f->hred, f->dred + nb, sizeof(double), hipMemcpyDeviceToHost, f->stream));
    BD_HIP(hipStreamSynchronize(f->stream));
    *loss = f->hred[0];
    return BD_OK;
}

int bd_fno_adamw(bd_fno *f, double lr, double beta1, double beta2, double eps, double weight_decay)
{
    BD_CHECK(f, "bd_fno_adamw: null argument");
    BD_CHECK(f->train, "bd_fno_adamw: context was created without training buffers");
    BD_CHECK(lr >= 0.0 && beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0 && eps > 0.0,
             "bd_fno_adamw: invalid hyper-parameters");
    BD_HIP(hipSetDevice(f->device));
    f->step += 1;
    const float bc1 = (float)(1.0 - std::pow(beta1, (double)f->step));
    const float bc2 = (float)(1.0 - std::pow(beta2, (double)f->step));
    const int nb = (int)std::min<int64_t>((f->nparam + 255) / 256, 4096);
    hipLaunchKernelGGL(k_adamw, dim3(nb), dim3(256), 0, f->stream, f->params, f->grads, f->am, f->av, f->nparam,
                       (float)lr, (float)beta1, (float)beta2, (float)eps, (float)weight_decay, bc1, bc2);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int bd_dev_alloc(int device, int64_t bytes, void **ptr)
{
    BD_CHECK(ptr, "bd_dev_alloc: null argument");
    BD_CHECK(bytes > 0, "bd_dev_alloc: bytes must be > 0");
    *ptr = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("bd_dev_alloc: no HIP device");
        return BD_ENODEV;
    }
    BD_CHECK(device >= 0 && device < ndev, "bd_dev_alloc: device index out of range");
    BD_HIP(hipSetDevice(device));
    BD_HIP(hipMalloc(ptr, (size_t)bytes));
    return BD_OK;
}

int bd_dev_free(void *ptr)
{
    if (ptr) BD_HIP(hipFree(ptr));
    return BD_OK;
}

int bd_dev_upload(void *dst_dev, const void *src_host, int64_t bytes)
{
    BD_CHECK(dst_dev && src_host && bytes >= 0, "bd_dev_upload: invalid argument");
    if (bytes) BD_HIP(hipMemcpy(dst_dev, src_host, (size_t)bytes, hipMemcpyHostToDevice));
    return BD_OK;
}

int bd_dev_download(void *dst_host, const void *src_dev, int64_t bytes)
{
    BD_CHECK(dst_host && src_dev && bytes >= 0, "bd_dev_download: invalid argument");
    if (bytes) BD_HIP(hipMemcpy(dst_host, src_dev, (size_t)bytes, hipMemcpyDeviceToHost));
    return BD_OK;
}

int bd_dev_copy(void *dst_dev, const void *src_dev, int64_t bytes)
{
    BD_CHECK(dst_dev && src_dev && bytes >= 0, "bd_dev_copy: invalid argument");
    if (bytes) BD_HIP(hipMemcpy(dst_dev, src_dev, (size_t)bytes, hipMemcpyDeviceToDevice));
    return BD_OK;
}

int bd_dev_fill_f32(void *dst_dev, float value, int64_t n)
{
    BD_CHECK(dst_dev && n >= 0, "bd_dev_fill_f32: invalid argument");
    if (!n) return BD_OK;
    const int nb = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(bd::fno::k_fill, dim3(nb), dim3(256), 0, nullptr, (float *)dst_dev, value, n);
    BD_HIP(hipGetLastError());
    BD_HIP(hipDeviceSynchronize());
    return BD_OK;
}

int bd_fno_sync(bd_fno *f)
{
    BD_CHECK(f, "bd_fno_sync: null argument");
    BD_HIP(hipSetDevice(f->device));
    BD_HIP(hipStreamSynchronize(f->stream));
    return BD_OK;
}

}  // extern "C"
