// bd_util.hip -- element-wise helpers: dtype conversion, fills, deterministic
// FP64 reductions, slices and the phase map.  All are HBM-bound single passes.
#include "bd_internal.hpp"

#include <algorithm>
#include <cstring>

namespace bd {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

int ensure(Ctx *c, void **p, size_t bytes)
{
    if (*p) return BD_OK;
    BD_HIP(hipMalloc(p, bytes));
    BD_HIP(hipMemsetAsync(*p, 0, bytes, c->stream));
    return BD_OK;
}

int ensure_staging(Ctx *c, size_t bytes)
{
    if (c->staging_bytes >= bytes) return BD_OK;
    if (c->staging) BD_HIP(hipFree(c->staging));
    c->staging = nullptr;
    BD_HIP(hipMalloc(&c->staging, bytes));
    c->staging_bytes = bytes;
    return BD_OK;
}

int ensure_state(Ctx *c)
{
    const size_t f = (size_t)c->nvox * sizeof(float);
    BD_TRY(ensure(c, (void **)&c->T, f));
    BD_TRY(ensure(c, (void **)&c->T2, f));
    BD_TRY(ensure(c, (void **)&c->H, f));
    BD_TRY(ensure(c, (void **)&c->chi, f));
    BD_TRY(ensure(c, (void **)&c->Q, f));
    BD_TRY(ensure(c, (void **)&c->k, f));
    BD_TRY(ensure(c, (void **)&c->rc, f));
    BD_TRY(ensure(c, (void **)&c->flags, (size_t)c->nvox));
    return BD_OK;
}

// ------------------------------------------------------------------ kernels
__global__ void k_f64_to_f32(const double *__restrict__ s, float *__restrict__ d, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (float)s[i];
}
__global__ void k_f32_to_f64(const float *__restrict__ s, double *__restrict__ d, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (double)s[i];
}
// set / clear one flag bit from a u8 mask
__global__ void k_mask_to_flag(const uint8_t *__restrict__ m, uint8_t *__restrict__ f, uint8_t bit, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        f[i] = m[i] ? (uint8_t)(f[i] | bit) : (uint8_t)(f[i] & ~bit);
}
__global__ void k_flag_to_mask(const uint8_t *__restrict__ f, uint8_t *__restrict__ m, uint8_t bit, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m[i] = (f[i] & bit) ? 1 : 0;
}
__global__ void k_hpos_from_h(const float *__restrict__ H, uint8_t *__restrict__ f, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        f[i] = H[i] > 0.0f ? (uint8_t)(f[i] | F_HPOS) : (uint8_t)(f[i] & ~F_HPOS);
}
__global__ void k_al_get(const float2 *__restrict__ al, float *__restrict__ d, int which, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = which ? al[i].y : al[i].x;
}
__global__ void k_al_set(float2 *__restrict__ al, const float *__restrict__ s, int which, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (which) al[i].y = s[i];
        else al[i].x = s[i];
    }
}
__global__ void k_fill_f32(float *d, float v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = v;
}
__global__ void k_fill_f64(double *d, double v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = v;
}
// Zeroes up to four small 8-byte ranges in one launch (scalar accumulators,
// counters, per-depth tallies): one dependent launch instead of one per range.
struct ZeroSet {
    double *p[4];
    int64_t n[4];
};
__global__ __launch_bounds__(256) void k_zero_f64(ZeroSet z)
{
#pragma unroll
    for (int r = 0; r < 4; ++r)
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < z.n[r]; i += (int64_t)gridDim.x * blockDim.x)
            z.p[r][i] = 0.0;
}
__global__ void k_fill_u8(uint8_t *d, uint8_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = v;
}

constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;

template <int KIND>
__device__ inline double load_as_double(const void *p, int64_t i)
{
    if (KIND == K_F32) return (double)((const float *)p)[i];
    if (KIND == K_F64) return ((const double *)p)[i];
    return (double)((const uint8_t *)p)[i];
}

// op 0 sum, 1 max, 2 min.  Fixed grid, one partial per block, then one final
// block: the FP64 result is bit-reproducible run to run.
template <int KIND>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_partial(const void *__restrict__ src, int64_t n, int op,
                                                                 const uint8_t *__restrict__ mask, uint8_t bits,
                                                                 int invert, double *__restrict__ partial)
{
    __shared__ double s[RED_THREADS];
    double acc = op == 0 ? 0.0 : (op == 1 ? -1.0e308 : 1.0e308);
    int64_t i0 = 0;
    if constexpr (KIND == K_F32) if (((uintptr_t)src & 15) == 0 && ((uintptr_t)mask & 3) == 0) {
        // four elements per load (float4, and the four mask bytes as one word)
        const int64_t n4 = n / 4;
        const float4 *s4 = (const float4 *)src;
        const uint32_t *m4 = (const uint32_t *)mask;
        const uint32_t b4 = (uint32_t)bits * 0x01010101u;
        for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n4; q += (int64_t)gridDim.x * blockDim.x) {
            uint32_t keep = 0xffffffffu;
            if (mask) {
                const uint32_t hit = m4[q] & b4;  // per byte: nonzero where the bit is set
                keep = 0;
                for (int e = 0; e < 4; ++e) {
                    const bool m = ((hit >> (8 * e)) & 0xffu) != 0;
                    if (m != (invert != 0)) keep |= 0xffu << (8 * e);
                }
                if (!keep) continue;
            }
            const float4 v4 = s4[q];
            const float vv[4] = {v4.x, v4.y, v4.z, v4.w};
            for (int e = 0; e < 4; ++e) {
                if (!((keep >> (8 * e)) & 1u)) continue;
                const double v = vv[e];
                acc = op == 0 ? acc + v : (op == 1 ? fmax(acc, v) : fmin(acc, v));
            }
        }
        i0 = n4 * 4;
    }
    for (int64_t i = i0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (mask) {
            const bool m = (mask[i] & bits) != 0;
            if (m == (invert != 0)) continue;
        }
        const double v = load_as_double<KIND>(src, i);
        acc = op == 0 ? acc + v : (op == 1 ? fmax(acc, v) : fmin(acc, v));
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = s[threadIdx.x + w];
            s[threadIdx.x] = op == 0 ? s[threadIdx.x] + o : (op == 1 ? fmax(s[threadIdx.x], o) : fmin(s[threadIdx.x], o));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double *__restrict__ partial, int np, int op,
                                                               double *__restrict__ out)
{
    __shared__ double s[RED_THREADS];
    double acc = op == 0 ? 0.0 : (op == 1 ? -1.0e308 : 1.0e308);
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        const double v = partial[i];
        acc = op == 0 ? acc + v : (op == 1 ? fmax(acc, v) : fmin(acc, v));
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = s[threadIdx.x + w];
            s[threadIdx.x] = op == 0 ? s[threadIdx.x] + o : (op == 1 ? fmax(s[threadIdx.x], o) : fmin(s[threadIdx.x], o));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// Q = A * scale from the absorbed-weight tally (voxel_mc_3d.py:684), and in the same pass the
// block partials of the sum of Q over the cells whose flags do not hold `bits` (invert = 1) or
// do (invert = 0).  The grid, the elements each thread takes, their order and the tree are those
// of k_reduce_partial<K_F32> (sum) on Q, so the partials are the same bits as a separate
// reduction of the Q written here, without reading Q back.
__global__ __launch_bounds__(RED_THREADS) void k_q_sum_partial(const float2 *__restrict__ AL, float *__restrict__ Q,
                                                                float scale, int64_t n,
                                                                const uint8_t *__restrict__ mask, uint8_t bits,
                                                                int invert, double *__restrict__ partial)
{
    __shared__ double s[RED_THREADS];
    double acc = 0.0;
    int64_t i0 = 0;
    if (((uintptr_t)Q & 15) == 0 && ((uintptr_t)mask & 3) == 0 && ((uintptr_t)AL & 15) == 0) {
        const int64_t n4 = n / 4;
        const float4 *a4 = (const float4 *)AL;
        float4 *q4 = (float4 *)Q;
        const uint32_t *m4 = (const uint32_t *)mask;
        const uint32_t b4 = (uint32_t)bits * 0x01010101u;
        for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n4; q += (int64_t)gridDim.x * blockDim.x) {
            const float4 lo = a4[2 * q], hi = a4[2 * q + 1];  // {A, L} of four cells
            const uint32_t hit = m4[q] & b4;
            const float vv[4] = {lo.x * scale, lo.z * scale, hi.x * scale, hi.z * scale};
            q4[q] = make_float4(vv[0], vv[1], vv[2], vv[3]);
            for (int e = 0; e < 4; ++e) {
                const bool m = ((hit >> (8 * e)) & 0xffu) != 0;
                if (m != (invert != 0)) acc = acc + (double)vv[e];
            }
        }
        i0 = n4 * 4;
    }
    for (int64_t i = i0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = AL[i].x * scale;
        Q[i] = v;
        const bool m = (mask[i] & bits) != 0;
        if (m != (invert != 0)) acc = acc + (double)v;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

int q_from_a_sum_enqueue(Ctx *c, float scale, const uint8_t *mask, uint8_t bits, bool invert, double *dev_out)
{
    BD_TRY(ensure(c, (void **)&c->red, sizeof(double) * (64 + RED_BLOCKS)));
    double *partial = c->red + 64;
    hipLaunchKernelGGL(k_q_sum_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream, c->AL, c->Q, scale, c->nvox,
                       mask, bits, invert ? 1 : 0, partial);
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, c->stream, partial, RED_BLOCKS, 0, dev_out);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

static int reduce_enqueue(Ctx *c, const void *src, int kind, int64_t n, int op, const uint8_t *mask, uint8_t bits,
                          int invert, double *dev_out)
{
    // partials live in the staging tail of the reduction scratch
    double *partial = nullptr;
    BD_TRY(ensure(c, (void **)&c->red, sizeof(double) * (64 + RED_BLOCKS)));
    partial = c->red + 64;
    dim3 gr(RED_BLOCKS), bl(RED_THREADS);
    switch (kind) {
    case K_F32: hipLaunchKernelGGL(k_reduce_partial<K_F32>, gr, bl, 0, c->stream, src, n, op, mask, bits, invert, partial); break;
    case K_F64: hipLaunchKernelGGL(k_reduce_partial<K_F64>, gr, bl, 0, c->stream, src, n, op, mask, bits, invert, partial); break;
    default: hipLaunchKernelGGL(k_reduce_partial<K_U8>, gr, bl, 0, c->stream, src, n, op, mask, bits, invert, partial); break;
    }
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_final, dim3(1), bl, 0, c->stream, partial, RED_BLOCKS, op, dev_out);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

static int reduce_impl(Ctx *c, const void *src, int kind, int64_t n, int op, const uint8_t *mask, uint8_t bits,
                       int invert, double *out)
{
    BD_TRY(reduce_enqueue(c, src, kind, n, op, mask, bits, invert, c->red));
    BD_HIP(hipMemcpyAsync(c->host_pinned, c->red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    *out = *(double *)c->host_pinned;
    return BD_OK;
}

int launch_reduce(Ctx *c, const void *src, int kind, int64_t n, int op, const uint8_t *mask, uint8_t bits,
                  double *out)
{
    return reduce_impl(c, src, kind, n, op, mask, bits, 0, out);
}

int reduce_masked_enqueue(Ctx *c, const float *src, int64_t n, const uint8_t *flags, uint8_t bit, bool invert,
                          double *dev_out)
{
    return reduce_enqueue(c, src, K_F32, n, 0, flags, bit, invert ? 1 : 0, dev_out);
}

int reduce_masked(Ctx *c, const float *src, int64_t n, const uint8_t *flags, uint8_t bit, bool invert, double *out)
{
    return reduce_impl(c, src, K_F32, n, 0, flags, bit, invert ? 1 : 0, out);
}

int launch_zero_f64(Ctx *c, std::initializer_list<std::pair<void *, int64_t>> ranges)
{
    BD_CHECK(ranges.size() <= 4, "launch_zero_f64: at most four ranges");
    ZeroSet z{};
    int64_t nmax = 1;
    int r = 0;
    for (const auto &rg : ranges) {
        z.p[r] = (double *)rg.first;
        z.n[r] = rg.second;
        nmax = std::max(nmax, rg.second);
        ++r;
    }
    const int nb = blocks_for(nmax, 256) > 64 ? 64 : blocks_for(nmax, 256);
    hipLaunchKernelGGL(k_zero_f64, dim3(nb), dim3(256), 0, c->stream, z);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int launch_fill(Ctx *c, void *dst, int kind, int64_t n, double v)
{
    const int nb = blocks_for(n, 256) > 4096 ? 4096 : blocks_for(n, 256);
    if (kind == K_F32) hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(256), 0, c->stream, (float *)dst, (float)v, n);
    else if (kind == K_F64) hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(256), 0, c->stream, (double *)dst, v, n);
    else hipLaunchKernelGGL(k_fill_u8, dim3(nb), dim3(256), 0, c->stream, (uint8_t *)dst, (uint8_t)v, n);
    BD_HIP(hipGetLastError());
    return BD_OK;
}


__global__ void k_slice(const void *__restrict__ src, int kind, int Nx, int Ny, int Nz, int plane, int index,
                        int stride, float *__restrict__ out)
{
    // plane 1: x-z plane at j = index -> out[(i/stride) * nzs + k/stride]
    // plane 0: y-z plane at i = index -> out[(j/stride) * nzs + k/stride]
    const int nzs = (Nz + stride - 1) / stride;
    const int na = plane == 1 ? (Nx + stride - 1) / stride : (Ny + stride - 1) / stride;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < na * nzs; t += gridDim.x * blockDim.x) {
        const int a = (t / nzs) * stride, kk = (t % nzs) * stride;
        const int64_t id = plane == 1 ? ((int64_t)a * Ny + index) * Nz + kk : ((int64_t)index * Ny + a) * Nz + kk;
        float v;
        if (kind == K_F32) v = ((const float *)src)[id];
        else if (kind == K_F64) v = (float)((const double *)src)[id];
        else if (kind == K_U8) v = (float)((const uint8_t *)src)[id];
        else if (kind == K_FLAG_ABL) v = (((const uint8_t *)src)[id] & F_ABL) ? 1.0f : 0.0f;
        else if (kind == K_FLAG_DRY) v = (((const uint8_t *)src)[id] & F_DRY) ? 1.0f : 0.0f;
        else if (kind == K_AL_A) v = ((const float2 *)src)[id].x;
        else v = ((const float2 *)src)[id].y;
        out[t] = v;
    }
}

int launch_slice(Ctx *c, const void *src, int kind, int plane, int index, int stride, float *out, int64_t n)
{
    hipLaunchKernelGGL(k_slice, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, src, kind, c->g.nx, c->g.ny,
                       c->g.nz, plane, index, stride, out);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// run_multispectral_matrix.py:168-192: 0 intact, 1 char, 2 boiling, 3 void
__global__ void k_phase_map(const float *__restrict__ T, const float *__restrict__ chi, const float *__restrict__ H,
                            const uint8_t *__restrict__ f, const uint8_t *__restrict__ gap, uint8_t *__restrict__ out,
                            int64_t n, int has_chi)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool g = gap && gap[i];
        const uint8_t fl = f[i];
        uint8_t ph = 0;
        const bool boiling = ((T[i] >= 99.5f) || (H[i] > 0.0f)) && !g;
        const bool charred = (fl & F_DRY) && has_chi && ((double)chi[i] >= 0.10) && !g;
        if (boiling) ph = 2;
        if (charred) ph = 1;
        if ((fl & F_ABL) || g) ph = 3;
        out[i] = ph;
    }
}

// a plume-grid field sampled at the centres of one thermal-grid plane
__global__ void k_slice_plume(const float *__restrict__ src, int Nx, int Ny, int Nz, int Px, int Py, int Pz,
                              int plane, int index, int stride, float *__restrict__ out)
{
    const int nzs = (Nz + stride - 1) / stride;
    const int na = plane == 1 ? (Nx + stride - 1) / stride : (Ny + stride - 1) / stride;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < na * nzs; t += gridDim.x * blockDim.x) {
        const int a = (t / nzs) * stride, kk = (t % nzs) * stride;
        const int i = plane == 1 ? a : index, j = plane == 1 ? index : a;
        const int I = (int)(((int64_t)(2 * i + 1) * Px) / (2 * (int64_t)Nx));
        const int J = (int)(((int64_t)(2 * j + 1) * Py) / (2 * (int64_t)Ny));
        const int K = (int)(((int64_t)(2 * kk + 1) * Pz) / (2 * (int64_t)Nz));
        out[t] = src[((int64_t)I * Py + J) * Pz + K];
    }
}

int launch_slice_plume(Ctx *c, const float *src, int plane, int index, int stride, float *out, int64_t n)
{
    hipLaunchKernelGGL(k_slice_plume, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, src, c->g.nx, c->g.ny,
                       c->g.nz, c->plume.nx, c->plume.ny, c->plume.nz, plane, index, stride, out);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// run_multispectral_matrix.py:168-192 on one plane; air gap as a per-z profile
__global__ void k_phase_slice(const float *__restrict__ T, const float *__restrict__ chi,
                              const float *__restrict__ H, const uint8_t *__restrict__ f,
                              const uint8_t *__restrict__ gap_z, int Nx, int Ny, int Nz, int plane, int index,
                              int stride, uint8_t *__restrict__ out)
{
    const int nzs = (Nz + stride - 1) / stride;
    const int na = plane == 1 ? (Nx + stride - 1) / stride : (Ny + stride - 1) / stride;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < na * nzs; t += gridDim.x * blockDim.x) {
        const int a = (t / nzs) * stride, kk = (t % nzs) * stride;
        const int64_t id = plane == 1 ? ((int64_t)a * Ny + index) * Nz + kk : ((int64_t)index * Ny + a) * Nz + kk;
        const bool g = gap_z && gap_z[kk];
        const uint8_t fl = f[id];
        uint8_t ph = 0;
        if (((T[id] >= 99.5f) || (H[id] > 0.0f)) && !g) ph = 2;
        if ((fl & F_DRY) && ((double)chi[id] >= 0.10) && !g) ph = 1;
        if ((fl & F_ABL) || g) ph = 3;
        out[t] = ph;
    }
}

int launch_phase_slice(Ctx *c, const uint8_t *gap_z_dev, int plane, int index, int stride, uint8_t *out, int64_t n)
{
    hipLaunchKernelGGL(k_phase_slice, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, c->T, c->chi, c->H,
                       c->flags, gap_z_dev, c->g.nx, c->g.ny, c->g.nz, plane, index, stride, out);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int launch_phase_map(Ctx *c, const uint8_t *gap, uint8_t *out)
{
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_phase_map, dim3(nb), dim3(256), 0, c->stream, c->T, c->chi, c->H, c->flags, gap, out,
                       c->nvox, 1);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// exported for the API unit
void convert_f64_to_f32(Ctx *c, const double *s, float *d, int64_t n)
{
    const int nb = blocks_for(n, 256) > 8192 ? 8192 : blocks_for(n, 256);
    hipLaunchKernelGGL(k_f64_to_f32, dim3(nb), dim3(256), 0, c->stream, s, d, n);
}
void convert_f32_to_f64(Ctx *c, const float *s, double *d, int64_t n)
{
    const int nb = blocks_for(n, 256) > 8192 ? 8192 : blocks_for(n, 256);
    hipLaunchKernelGGL(k_f32_to_f64, dim3(nb), dim3(256), 0, c->stream, s, d, n);
}
void mask_to_flag(Ctx *c, const uint8_t *m, uint8_t bit)
{
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_mask_to_flag, dim3(nb), dim3(256), 0, c->stream, m, c->flags, bit, c->nvox);
}
void flag_to_mask(Ctx *c, uint8_t *m, uint8_t bit)
{
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_flag_to_mask, dim3(nb), dim3(256), 0, c->stream, c->flags, m, bit, c->nvox);
}
void hpos_from_h(Ctx *c)
{
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_hpos_from_h, dim3(nb), dim3(256), 0, c->stream, c->H, c->flags, c->nvox);
}
void al_get(Ctx *c, float *d, int which)
{
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_al_get, dim3(nb), dim3(256), 0, c->stream, c->AL, d, which, c->nvox);
}
void al_set(Ctx *c, const float *s, int which)
{
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_al_set, dim3(nb), dim3(256), 0, c->stream, c->AL, s, which, c->nvox);
}

}  // namespace bd

// ------------------------------------------------------------------ copy-rate probe
// The achievable HBM rate the stencils are set against: a streaming float4 copy of one
// buffer into another (each byte read once and written once).  Each lane moves U float4 per
// grid-stride step, all U loads issued before the stores, so U x 16 bytes per lane are in
// flight; six launch shapes (blocks per CU, loads per lane, with and without non-temporal
// hints) are timed, each after one untimed launch, and the fastest counts.
namespace bd {
typedef float f4v __attribute__((ext_vector_type(4)));
template <int U, bool NT>
__global__ __launch_bounds__(256) void k_copy_probe(const f4v *__restrict__ src, f4v *__restrict__ dst, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < n; i += U * stride) {
        f4v v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = NT ? __builtin_nontemporal_load(&src[i + u * stride]) : src[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (NT) __builtin_nontemporal_store(v[u], &dst[i + u * stride]);
            else dst[i + u * stride] = v[u];
        }
    }
    for (; i < n; i += stride) dst[i] = src[i];
}
}  // namespace bd

extern "C" int bd_probe_copy_rate(int device, int64_t bytes, double *kernel_ms)
{
    using namespace bd;
    BD_CHECK(kernel_ms, "bd_probe_copy_rate: null argument");
    BD_CHECK(bytes >= 16 && bytes % 16 == 0 && bytes <= (int64_t(1) << 36),
             "bd_probe_copy_rate: bytes must be a positive multiple of 16, at most 64 GiB");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("bd_probe_copy_rate: no HIP device");
        return BD_ENODEV;
    }
    BD_CHECK(device >= 0 && device < ndev, "bd_probe_copy_rate: device index out of range");
    BD_HIP(hipSetDevice(device));
    f4v *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e[2] = {nullptr, nullptr};
    auto run = [&]() -> int {
        BD_HIP(hipMalloc((void **)&a, (size_t)bytes));
        BD_HIP(hipMalloc((void **)&b, (size_t)bytes));
        BD_HIP(hipStreamCreate(&s));
        BD_HIP(hipEventCreate(&e[0]));
        BD_HIP(hipEventCreate(&e[1]));
        const int64_t n = bytes / 16;
        int ncu = 256;
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
        const int64_t want = (n + 255) / 256;
        float best = -1.f;
        for (int shape = 0; shape < 6; ++shape) {
            const int64_t per_cu = (shape % 3 == 0) ? 8 : (shape % 3 == 1 ? 32 : 128);
            const int64_t cap = (int64_t)ncu * per_cu;
            const unsigned nb = (unsigned)(want < cap ? (want < 1 ? 1 : want) : cap);
            for (int rep = 0; rep < 2; ++rep) {  // rep 0 untimed (first touch / clocks), rep 1 timed
                if (rep == 1) BD_HIP(hipEventRecord(e[0], s));
                if (shape < 3) hipLaunchKernelGGL((k_copy_probe<4, true>), dim3(nb), dim3(256), 0, s, a, b, n);
                else hipLaunchKernelGGL((k_copy_probe<1, false>), dim3(nb), dim3(256), 0, s, a, b, n);
                BD_HIP(hipGetLastError());
            }
            BD_HIP(hipEventRecord(e[1], s));
            BD_HIP(hipStreamSynchronize(s));
            float ms = 0.f;
            BD_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
            if (best < 0.f || ms < best) best = ms;
        }
        *kernel_ms = best;
        return BD_OK;
    };
    const int rc = run();
    if (e[0]) (void)hipEventDestroy(e[0]);
    if (e[1]) (void)hipEventDestroy(e[1]);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
