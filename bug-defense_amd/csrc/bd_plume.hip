// bd_plume.hip -- plume extinction field on the device (plume.py:18-366,
// coupled_ablation.py:587-603, 619-626).
//
// The plume lives on its own grid (Px, Py, Pz) covering the same box as the
// thermal grid.  With Px,Py,Pz == Nx,Ny,Nz every mapping below is the identity
// and the kernels are the reference's arithmetic cell for cell.  A coarser
// plume grid (e.g. 128^3 under a 256x256x192 thermal grid, where the explicit
// plume scheme at the thermal spacing would be unstable) is coupled by
// cell-centre sampling: a plume cell is void when the thermal cell under its
// centre is void, a removed thermal cell deposits delta_mu * (V_thermal /
// V_plume) into the plume cell under its centre (mass conserving), and the
// transport optics read the plume cell under each thermal voxel's centre.
#include "bd_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bd {

__device__ __forceinline__ int cmap(int i, int n_from, int n_to)
{
    return (int)(((int64_t)(2 * i + 1) * n_to) / (2 * (int64_t)n_from));
}

// The same map in 32-bit unsigned arithmetic (W32), exact where every
// (2 i + 1) n_to and every voxel index stays below 2^32 (map32_ok).  A 64-bit
// division by a runtime value is a software routine of some 150 instructions; a
// gather of 16 cells mapped that way kept a wave busy for microseconds.
template <bool W32>
__device__ __forceinline__ int cmapw(int i, int n_from, int n_to)
{
    if (W32) return (int)(((unsigned)(2 * i + 1) * (unsigned)n_to) / (2u * (unsigned)n_from));
    return cmap(i, n_from, n_to);
}

// (i, j, k) of voxel c of an (Nx, Ny, Nz) grid, z fastest
template <bool W32>
__device__ __forceinline__ void split_ijk(int64_t c, int Ny, int Nz, int &i, int &j, int &k)
{
    if (W32) {
        const unsigned u = (unsigned)c, r = u / (unsigned)Nz, ii = r / (unsigned)Ny;
        k = (int)(u - r * (unsigned)Nz);
        j = (int)(r - ii * (unsigned)Ny);
        i = (int)ii;
    } else {
        k = (int)(c % Nz);
        const int64_t r = c / Nz;
        j = (int)(r % Ny);
        i = (int)(r / Ny);
    }
}

// the plume cell under thermal voxel (i, j, k)
template <bool W32>
__device__ __forceinline__ int64_t plume_of(int i, int j, int k, int Nx, int Ny, int Nz, int Px, int Py, int Pz)
{
    return ((int64_t)cmapw<W32>(i, Nx, Px) * Py + cmapw<W32>(j, Ny, Py)) * Pz + cmapw<W32>(k, Nz, Pz);
}

// both directions of the map (thermal -> plume and plume -> thermal) and every voxel index
static bool map32_ok(const Ctx *c)
{
    const int64_t m = std::max({(int64_t)c->g.nx, (int64_t)c->g.ny, (int64_t)c->g.nz, (int64_t)c->plume.nx,
                                (int64_t)c->plume.ny, (int64_t)c->plume.nz});
    return c->nvox < (int64_t(1) << 32) && (2 * m + 1) * m < (int64_t(1) << 32);
}

// One explicit step of the plume diffusion (plume.py:18-80, FP32 arithmetic)
// on the plume grid (Nx, Ny, Nz), z fastest: C in, Cn out, vm the void mask.
struct PlumeStep {
    const float *__restrict__ C;
    float *__restrict__ Cn;
    const uint8_t *__restrict__ vm;
    int Nx, Ny, Nz;
    float idx2, idy2, idz2, inv_dz, dt, Dxy, Dz, vz, decay;
};

// The diffusion update of one void interior cell from its neighbourhood
// (plume.py:18-80, FP32): a neighbour that is not void reflects the cell's own value.
// Contraction is off, so both row kernels below produce the same bits.
__device__ __forceinline__ float plume_cell(const PlumeStep &p, float Cc, float Cw, float Ce, float Cs, float Cq,
                                            float Cb, float Ct)
{
#pragma clang fp contract(off)
    const float diff = p.Dxy * ((Cw - 2.f * Cc + Ce) * p.idx2 + (Cs - 2.f * Cc + Cq) * p.idy2) +
                       p.Dz * ((Cb - 2.f * Cc + Ct) * p.idz2);
    const float adv = p.vz > 0.f ? -p.vz * (Cc - Ct) * p.inv_dz : 0.f;
    const float v = (Cc + p.dt * (diff + adv)) * p.decay;
    return fmaxf(v, 0.f);
}

// The step on the z row (i, j), its cells taken 64 at a time by the lanes of one
// wave, so every neighbour load is one coalesced row access.  Indices are 32-bit
// (plume grids have fewer than 2^31 cells; the launcher checks).
__device__ __forceinline__ void plume_row(const PlumeStep &p, int i, int j, int lane)
{
    const int Ny = p.Ny, Nz = p.Nz;
    const int pl = Ny * Nz;
    const int row = (i * Ny + j) * Nz;
    const bool edge_row = i == 0 || i == p.Nx - 1 || j == 0 || j == Ny - 1;
    for (int k = lane; k < Nz; k += 64) {
        const int c = row + k;
        if (!p.vm[c] || edge_row || k == 0 || k == Nz - 1) {
            p.Cn[c] = 0.f;
            continue;
        }
        const float Cc = p.C[c];
        p.Cn[c] = plume_cell(p, Cc, p.vm[c - pl] ? p.C[c - pl] : Cc, p.vm[c + pl] ? p.C[c + pl] : Cc,
                             p.vm[c - Nz] ? p.C[c - Nz] : Cc, p.vm[c + Nz] ? p.C[c + Nz] : Cc,
                             p.vm[c - 1] ? p.C[c - 1] : Cc, p.vm[c + 1] ? p.C[c + 1] : Cc);
    }
}

// The same step for rows whose length is a multiple of 128: a lane takes the
// cells z = k0 and k0 + 64 together, first their void flags, then both
// neighbourhoods, each set of loads issued at once, so the pair waits on memory
// twice instead of twice per cell.  A cell that is not void or not interior
// gathers its own cell again (a cache hit), and a select, not a branch, writes 0
// for it, so no load is left behind a branch.
__device__ __forceinline__ void plume_row2(const PlumeStep &p, int i, int j, int lane)
{
    const int Ny = p.Ny, Nz = p.Nz;
    const int pl = Ny * Nz;
    const int row = (i * Ny + j) * Nz;
    const bool edge_row = i == 0 || i == p.Nx - 1 || j == 0 || j == Ny - 1;
    for (int k0 = lane; k0 < Nz; k0 += 128) {
        int c[2];
        uint8_t vc[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            c[e] = row + k0 + 64 * e;
            vc[e] = p.vm[c[e]];
        }
        bool act[2];
        uint8_t vn[2][6];
        float cc[2], cn[2][6];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int k = k0 + 64 * e;
            act[e] = vc[e] && !edge_row && k > 0 && k < Nz - 1;
            const int x = c[e];
            const int id[6] = {act[e] ? x - pl : x, act[e] ? x + pl : x, act[e] ? x - Nz : x,
                               act[e] ? x + Nz : x, act[e] ? x - 1 : x,  act[e] ? x + 1 : x};
            cc[e] = p.C[x];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                vn[e][q] = p.vm[id[q]];
                cn[e][q] = p.C[id[q]];
            }
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const float Cc = cc[e];
            float nb[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) nb[q] = vn[e][q] ? cn[e][q] : Cc;
            const float v = plume_cell(p, Cc, nb[0], nb[1], nb[2], nb[3], nb[4], nb[5]);
            p.Cn[c[e]] = act[e] ? v : 0.f;
        }
    }
}

// plume.py:18-80: one diffusion step, PD_ROWS z rows of the plume grid per
// block (one wave per row).  The row comes from the launch shape (blockIdx.y =
// x, blockIdx.x = group of y rows), so no lane divides a linear index.
constexpr int PD_ROWS = 4;

template <bool PAIRS>
__global__ __launch_bounds__(64 * PD_ROWS) void k_plume_diffuse(PlumeStep p)
{
    const int j = blockIdx.x * PD_ROWS + threadIdx.y;
    if (j >= p.Ny) return;
    if (PAIRS) plume_row2(p, blockIdx.y, j, threadIdx.x);
    else plume_row(p, blockIdx.y, j, threadIdx.x);
}

static int ew_blocks(int64_t n) { return blocks_for(n, 256) > 8192 ? 8192 : blocks_for(n, 256); }

static void plume_spacing(Ctx *c, double &dx, double &dy, double &dz)
{
    dx = c->g.dx * c->g.nx / c->plume.nx;
    dy = c->g.dy * c->g.ny / c->plume.ny;
    dz = c->g.dz * c->g.nz / c->plume.nz;
}

static int plume_step_args(Ctx *c, double dt, double decay, PlumeStep &ps)
{
    const bd_plume_params &p = c->plume;
    BD_CHECK(c->pvox < (int64_t(1) << 31) && p.nx <= 65535, "plume grid too large for the diffusion kernel");
    double dx, dy, dz;
    plume_spacing(c, dx, dy, dz);
    if (p.dt_clamp) {
        // the step the reference's Metal solver takes (metal_ablation_solver.py:234-236): stable for any
        // grid, at the cost of diffusing less per sub-step than the nominal coefficients
        const double d_rate = p.D_xy * (1.0 / (dx * dx) + 1.0 / (dy * dy)) + p.D_z / (dz * dz);
        if (d_rate > 1e-6) dt = std::min(dt, 0.45 / (2.0 * d_rate));
    }
    ps = PlumeStep{};
    ps.vm = c->pvoid;
    ps.Nx = p.nx;
    ps.Ny = p.ny;
    ps.Nz = p.nz;
    ps.idx2 = (float)(1.0 / (dx * dx));
    ps.idy2 = (float)(1.0 / (dy * dy));
    ps.idz2 = (float)(1.0 / (dz * dz));
    ps.inv_dz = (float)(1.0 / dz);
    ps.dt = (float)dt;
    ps.Dxy = (float)p.D_xy;
    ps.Dz = (float)p.D_z;
    ps.vz = (float)p.v_drift;
    ps.decay = (float)decay;
    return BD_OK;
}

static int launch_plume_steps(Ctx *c, int n, PlumeStep ps)
{
    const dim3 grid((unsigned)((ps.Ny + PD_ROWS - 1) / PD_ROWS), (unsigned)ps.Nx);
    // rows of a multiple of 128 cells take the paired kernel (BUGDEF_PLUME_PAIRS=0: never)
    const bool pairs = ps.Nz % 128 == 0 && c->plume_pairs;
    for (int s = 0; s < n; ++s) {
        ps.C = c->P;
        ps.Cn = c->P2;
        if (pairs) hipLaunchKernelGGL(k_plume_diffuse<true>, grid, dim3(64, PD_ROWS), 0, c->stream, ps);
        else hipLaunchKernelGGL(k_plume_diffuse<false>, grid, dim3(64, PD_ROWS), 0, c->stream, ps);
        BD_HIP(hipGetLastError());
        std::swap(c->P, c->P2);
    }
    return BD_OK;
}

double plume_decay(const Ctx *c, double dt)
{
    const double tl = c->plume.tau_linger;
    return (tl > 0.0 && std::isfinite(tl)) ? std::exp(-dt / tl) : 1.0;
}

int launch_plume_substeps(Ctx *c, int n, double dt, double decay)
{
    PlumeStep ps;
    BD_TRY(plume_step_args(c, dt, decay, ps));
    return launch_plume_steps(c, n, ps);
}

// any(ablated[:, :, k0-1])  (coupled_ablation.py:624)
__global__ void k_gap_any(const uint8_t *__restrict__ flags, int Nx, int Ny, int Nz, int kk,
                          unsigned long long *__restrict__ out)
{
    const int64_t cols = (int64_t)Nx * Ny;
    int hit = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < cols; c += (int64_t)gridDim.x * blockDim.x)
        hit |= (flags[c * Nz + kk] & F_ABL) ? 1 : 0;
    if (__any(hit) && (threadIdx.x & 63) == 0) atomicOr(out, 1ull);
}

// The plume void mask, one z row of plume cells per wave (blockIdx.y = I, blockIdx.x =
// group of PD_ROWS J rows): the thermal cell under each plume cell's centre is void,
// or the cell is in the air gap once the gap is breached (coupled_ablation.py:588,624).
template <bool W32>
__global__ __launch_bounds__(64 * PD_ROWS) void k_plume_void(const uint8_t *__restrict__ flags,
                                                             const unsigned long long *__restrict__ breach,
                                                             uint8_t *__restrict__ vm, int Nx, int Ny, int Nz, int Px,
                                                             int Py, int Pz, int k0, int k1)
{
    const int I = blockIdx.y, J = blockIdx.x * PD_ROWS + threadIdx.y;
    if (J >= Py) return;
    const bool br = breach && *breach;
    const int64_t col = ((int64_t)cmapw<W32>(I, Px, Nx) * Ny + cmapw<W32>(J, Py, Ny)) * Nz;
    uint8_t *out = vm + ((int64_t)I * Py + J) * Pz;
    for (int K = threadIdx.x; K < Pz; K += 64) {
        const int k = cmapw<W32>(K, Pz, Nz);
        const bool v = (flags[col + k] & F_ABL) || (br && k >= k0 && k < k1);
        out[K] = v ? 1 : 0;
    }
}

int launch_plume_void(Ctx *c)
{
    const bd_physics &ph = c->phys;
    const bool gap = c->has_phys && ph.air_gap_k0 > 0 && ph.air_gap_k1 > ph.air_gap_k0;
    unsigned long long *breach = nullptr;
    if (gap) {
        breach = c->counter + 2;
        BD_TRY(launch_fill(c, breach, K_F64, 1, 0.0));
        const int64_t cols = (int64_t)c->g.nx * c->g.ny;
        hipLaunchKernelGGL(k_gap_any, dim3(blocks_for(cols, 256) > 1024 ? 1024 : blocks_for(cols, 256)), dim3(256), 0,
                           c->stream, c->flags, c->g.nx, c->g.ny, c->g.nz, ph.air_gap_k0 - 1, breach);
        BD_HIP(hipGetLastError());
    }
    BD_CHECK(c->plume.nx <= 65535, "plume grid too large for the void-mask kernel");
    const dim3 grid((unsigned)((c->plume.ny + PD_ROWS - 1) / PD_ROWS), (unsigned)c->plume.nx);
    hipLaunchKernelGGL(map32_ok(c) ? k_plume_void<true> : k_plume_void<false>, grid, dim3(64, PD_ROWS), 0, c->stream,
                       c->flags, breach, c->pvoid,
                       c->g.nx, c->g.ny, c->g.nz, c->plume.nx, c->plume.ny, c->plume.nz,
                       gap ? ph.air_gap_k0 : 0, gap ? ph.air_gap_k1 : 0);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// plume.py:329 point injection of the cells just removed.  The removal marks are
// scanned 16 bytes per load (most chunks hold none); a chunk that holds some adds
// its cells in order.
template <bool W32>
__device__ __forceinline__ void inject_cell(float *__restrict__ P, int64_t c, int Nx, int Ny, int Nz, int Px, int Py,
                                            int Pz, float amount, int same)
{
    if (same) {
        P[c] += amount;  // plume.py:329 raw_field[newly] += delta_mu (one thread per cell)
    } else {
        int i, j, k;
        split_ijk<W32>(c, Ny, Nz, i, j, k);
        unsafeAtomicAdd(&P[plume_of<W32>(i, j, k, Nx, Ny, Nz, Px, Py, Pz)], amount);
    }
}

template <bool W32>
__global__ void k_inject_point(const uint8_t *__restrict__ rem, float *__restrict__ P, int Nx, int Ny, int Nz,
                               int Px, int Py, int Pz, float amount, int same)
{
    const int64_t N = (int64_t)Nx * Ny * Nz;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n16 = N / 16;  // rem is 16-byte aligned (the launcher checks)
    const uint4 *r16 = (const uint4 *)rem;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n16; q += stride) {
        const uint4 w = r16[q];
        if (!(w.x | w.y | w.z | w.w)) continue;
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        for (int e = 0; e < 16; ++e)
            if ((ws[e >> 2] >> (8 * (e & 3))) & 0xffu)
                inject_cell<W32>(P, q * 16 + e, Nx, Ny, Nz, Px, Py, Pz, amount, same);
    }
    for (int64_t c = n16 * 16 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < N; c += stride)
        if (rem[c]) inject_cell<W32>(P, c, Nx, Ny, Nz, Px, Py, Pz, amount, same);
}

__global__ void k_mask_void(float *__restrict__ P, const uint8_t *__restrict__ vm, int64_t n)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        if (!vm[c]) P[c] = 0.f;
}

// bounding box of the removed cells: slots {xmin, xmax, ymin, ymax, zmin, zmax}
__global__ void k_bbox(const uint8_t *__restrict__ rem, int Nx, int Ny, int Nz, unsigned long long *__restrict__ bb)
{
    const int64_t N = (int64_t)Nx * Ny * Nz;
    int mn[3] = {1 << 30, 1 << 30, 1 << 30}, mx[3] = {-1, -1, -1};
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < N; c += (int64_t)gridDim.x * blockDim.x) {
        if (!rem[c]) continue;
        const int k = (int)(c % Nz);
        const int64_t r = c / Nz;
        const int j = (int)(r % Ny), i = (int)(r / Ny);
        mn[0] = min(mn[0], i); mx[0] = max(mx[0], i);
        mn[1] = min(mn[1], j); mx[1] = max(mx[1], j);
        mn[2] = min(mn[2], k); mx[2] = max(mx[2], k);
    }
    for (int a = 0; a < 3; ++a) {
        if (mx[a] >= 0) {
            atomicMin(&bb[2 * a], (unsigned long long)mn[a]);
            atomicMax(&bb[2 * a + 1], (unsigned long long)mx[a]);
        }
    }
}

// plume.py:105-114 as a gather: t1[ox][y][z] = sum_a imp[ox-a][y][z] * kx[a] (a ascending)
// impulse source: the removed-cell mask at box origin (xm, ym, zm), or an
// explicit float32 box when `box` is not null
__global__ void k_splat1(const uint8_t *__restrict__ rem, const float *__restrict__ box, int Ny, int Nz, int xm,
                         int ym, int zm, int Lx, int Ly, int Lz, const float *__restrict__ kx, int nkx,
                         float *__restrict__ t1)
{
    // the reference sums tap by tap with separate multiply and add roundings: no FMA contraction
#pragma clang fp contract(off)
    const int OX = Lx + nkx - 1;
    const int64_t N = (int64_t)OX * Ly * Lz;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(t % Lz);
        const int64_t r = t / Lz;
        const int y = (int)(r % Ly), ox = (int)(r / Ly);
        float s = 0.f;
        for (int a = 0; a < nkx; ++a) {
            const int ix = ox - a;
            if (ix < 0 || ix >= Lx) continue;
            const float imp = box ? box[((int64_t)ix * Ly + y) * Lz + z]
                                  : (rem[((int64_t)(xm + ix) * Ny + (ym + y)) * Nz + (zm + z)] ? 1.f : 0.f);
            s += imp * kx[a];
        }
        t1[t] = s;
    }
}

// plume.py:116-125: t2[ox][oy][z] = sum_b t1[ox][oy-b][z] * ky[b]
__global__ void k_splat2(const float *__restrict__ t1, int OX, int Ly, int Lz, const float *__restrict__ ky, int nky,
                         float *__restrict__ t2)
{
    // the reference sums tap by tap with separate multiply and add roundings: no FMA contraction
#pragma clang fp contract(off)
    const int OY = Ly + nky - 1;
    const int64_t N = (int64_t)OX * OY * Lz;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(t % Lz);
        const int64_t r = t / Lz;
        const int oy = (int)(r % OY), ox = (int)(r / OY);
        float s = 0.f;
        for (int b = 0; b < nky; ++b) {
            const int iy = oy - b;
            if (iy < 0 || iy >= Ly) continue;
            s += t1[((int64_t)ox * Ly + iy) * Lz + z] * ky[b];
        }
        t2[t] = s;
    }
}

// plume.py:127-139: raw[gx][gy][gz] += t2 * (kz[c] * delta_mu), c ascending, FP64
// products rounded to FP32 after each tap exactly as the reference's typing does
__global__ void k_splat3(const float *__restrict__ t2, int OX, int OY, int Lz, const float *__restrict__ kz, int nkz,
                         double delta, int x0, int y0, int z0, float *__restrict__ P, int Nx, int Ny, int Nz)
{
    // the reference sums tap by tap with separate multiply and add roundings: no FMA contraction
#pragma clang fp contract(off)
    const int OZ = Lz + nkz - 1;
    const int64_t N = (int64_t)OX * OY * OZ;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x) {
        const int oz = (int)(t % OZ);
        const int64_t r = t / OZ;
        const int oy = (int)(r % OY), ox = (int)(r / OY);
        const int gx = x0 + ox, gy = y0 + oy, gz = z0 + oz;
        if (gx < 0 || gx >= Nx || gy < 0 || gy >= Ny || gz < 0 || gz >= Nz) continue;
        float *dst = &P[((int64_t)gx * Ny + gy) * Nz + gz];
        float v = *dst;
        for (int cc = 0; cc < nkz; ++cc) {
            const int iz = oz - cc;
            if (iz < 0 || iz >= Lz) continue;
            const double w = (double)kz[cc] * delta;
            v = (float)((double)v + (double)t2[((int64_t)ox * OY + oy) * Lz + iz] * w);
        }
        *dst = v;
    }
}

// three separable passes of an impulse box into the plume raw field; origin of
// the output box is (xo, yo, zo) in plume-grid cells
static int splat_box(Ctx *c, float *const *k3, const int *nk, const float *box, int xm, int ym, int zm, int Lx,
                     int Ly, int Lz, int xo, int yo, int zo, double delta)
{
    const int OX = Lx + nk[0] - 1, OY = Ly + nk[1] - 1, OZ = Lz + nk[2] - 1;
    const size_t need = std::max((size_t)OX * Ly * Lz, (size_t)OX * OY * Lz) * sizeof(float);
    if (c->splat_tmp_bytes < need) {
        if (c->splat_tmp1) BD_HIP(hipFree(c->splat_tmp1));
        if (c->splat_tmp2) BD_HIP(hipFree(c->splat_tmp2));
        c->splat_tmp1 = c->splat_tmp2 = nullptr;
        BD_HIP(hipMalloc((void **)&c->splat_tmp1, need));
        BD_HIP(hipMalloc((void **)&c->splat_tmp2, need));
        c->splat_tmp_bytes = need;
    }
    const int64_t n1 = (int64_t)OX * Ly * Lz, n2 = (int64_t)OX * OY * Lz, n3 = (int64_t)OX * OY * OZ;
    hipLaunchKernelGGL(k_splat1, dim3(ew_blocks(n1)), dim3(256), 0, c->stream, c->to_remove, box, c->g.ny, c->g.nz,
                       xm, ym, zm, Lx, Ly, Lz, k3[0], nk[0], c->splat_tmp1);
    hipLaunchKernelGGL(k_splat2, dim3(ew_blocks(n2)), dim3(256), 0, c->stream, c->splat_tmp1, OX, Ly, Lz, k3[1], nk[1],
                       c->splat_tmp2);
    hipLaunchKernelGGL(k_splat3, dim3(ew_blocks(n3)), dim3(256), 0, c->stream, c->splat_tmp2, OX, OY, Lz, k3[2], nk[2],
                       delta, xo, yo, zo, c->P, c->plume.nx, c->plume.ny, c->plume.nz);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

static int splat(Ctx *c, float *const *k3, const int *nk, const int *rr)
{
    if (!k3[0] || !k3[1] || !k3[2]) {
        set_error("plume splat kernels not uploaded (bd_set_plume_kernels)");
        return BD_ESTATE;
    }
    // the box has slots of its own: a lagged epoch's record cursor lives past this call
    unsigned long long *bb = c->counter + COUNTER_BBOX;
    unsigned long long init[6] = {~0ull, 0ull, ~0ull, 0ull, ~0ull, 0ull};
    BD_HIP(hipMemcpyAsync(bb, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_bbox, dim3(ew_blocks(c->nvox)), dim3(256), 0, c->stream, c->to_remove, c->g.nx, c->g.ny,
                       c->g.nz, bb);
    BD_HIP(hipGetLastError());
    unsigned long long h[6];
    BD_HIP(hipMemcpyAsync(h, bb, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    if (h[0] == ~0ull) return BD_OK;  // nothing removed
    const int xm = (int)h[0], ym = (int)h[2], zm = (int)h[4];
    const int Lx = (int)h[1] - xm + 1, Ly = (int)h[3] - ym + 1, Lz = (int)h[5] - zm + 1;
    return splat_box(c, k3, nk, nullptr, xm, ym, zm, Lx, Ly, Lz, xm - rr[0], ym - rr[1], zm - rr[2],
                     c->plume.delta_mu);
}

int launch_plume_splat_box(Ctx *c, int which, const float *host_box, int Lx, int Ly, int Lz, int x0, int y0, int z0,
                           double delta)
{
    float *const *k3 = which == 0 ? c->kb : c->kl;
    const int *nk = which == 0 ? c->nkb : c->nkl;
    if (!k3[0] || !k3[1] || !k3[2]) {
        set_error("plume splat kernels not uploaded (bd_set_plume_kernels)");
        return BD_ESTATE;
    }
    const size_t bytes = (size_t)Lx * Ly * Lz * sizeof(float);
    BD_TRY(ensure_staging(c, bytes < 64 ? 64 : bytes));
    BD_HIP(hipMemcpyAsync(c->staging, host_box, bytes, hipMemcpyHostToDevice, c->stream));
    BD_TRY(splat_box(c, k3, nk, (const float *)c->staging, 0, 0, 0, Lx, Ly, Lz, x0, y0, z0, delta));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

// plume.py:292-331: injection of the TO_REMOVE cells, then confinement to the
// current VOID mask (the engine computes it first, coupled_ablation.py:657-659)
int launch_plume_inject(Ctx *c)
{
    const bd_plume_params &p = c->plume;
    const bool same = p.nx == c->g.nx && p.ny == c->g.ny && p.nz == c->g.nz;
    if (p.blast_expansion) {
        BD_TRY(splat(c, c->kb, c->nkb, c->rb));
    } else {
        const double ratio = same ? 1.0 : (double)c->pvox / (double)c->nvox;
        BD_CHECK((uintptr_t)c->to_remove % 16 == 0, "plume injection: removal marks are not 16-byte aligned");
        hipLaunchKernelGGL(map32_ok(c) ? k_inject_point<true> : k_inject_point<false>,
                           dim3(ew_blocks((c->nvox + 15) / 16)), dim3(256), 0, c->stream, c->to_remove, c->P, c->g.nx,
                           c->g.ny, c->g.nz, p.nx, p.ny, p.nz, (float)(p.delta_mu * ratio), same ? 1 : 0);
        BD_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mask_void, dim3(ew_blocks(c->pvox)), dim3(256), 0, c->stream, c->P, c->pvoid, c->pvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// plume.py:352-366
__global__ void k_plume_sync(float *__restrict__ P, float *__restrict__ E, const uint8_t *__restrict__ vm, float mu_max,
                             int64_t n)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        if (!vm[c]) {
            P[c] = 0.f;
            E[c] = 0.f;
            continue;
        }
        const float r = P[c];
        E[c] = mu_max > 0.f ? tanhf(r / mu_max) * mu_max : r;
    }
}

int launch_plume_sync(Ctx *c)
{
    hipLaunchKernelGGL(k_plume_sync, dim3(ew_blocks(c->pvox)), dim3(256), 0, c->stream, c->P, c->Pext, c->pvoid,
                       (float)c->plume.mu_max, c->pvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// legacy separable path (plume.py:374-432)
__global__ void k_any_nonzero(const float *__restrict__ a, int64_t n, unsigned long long *__restrict__ out)
{
    int hit = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        hit |= a[c] != 0.f;
    if (__any(hit) && (threadIdx.x & 63) == 0) atomicOr(out, 1ull);
}

__global__ void k_legacy_pre(float *__restrict__ P, const float *__restrict__ E,
                             const unsigned long long *__restrict__ flags2, float decay, int do_decay, int64_t n)
{
    const bool copy = flags2[0] == 0 && flags2[1] != 0;  // raw all zero and ext not
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        float v = copy ? E[c] : P[c];
        if (do_decay) v *= decay;
        P[c] = v;
    }
}

__global__ void k_legacy_post(float *__restrict__ P, float *__restrict__ E, const uint8_t *__restrict__ vm,
                              float mu_max, int64_t n)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const float r = P[c] * (vm[c] ? 1.f : 0.f);
        P[c] = r;
        E[c] = mu_max > 0.f ? tanhf(r / mu_max) * mu_max : r;
    }
}

int launch_plume_update_sep(Ctx *c, double dt_epoch)
{
    const bd_plume_params &p = c->plume;
    unsigned long long *fl = c->counter + 10;
    BD_TRY(launch_fill(c, fl, K_F64, 2, 0.0));
    hipLaunchKernelGGL(k_any_nonzero, dim3(256), dim3(256), 0, c->stream, c->P, c->pvox, fl);
    hipLaunchKernelGGL(k_any_nonzero, dim3(256), dim3(256), 0, c->stream, c->Pext, c->pvox, fl + 1);
    const bool do_decay = p.tau_linger > 0.0 && dt_epoch > 0.0;
    const float decay = do_decay ? (float)std::exp(-dt_epoch / p.tau_linger) : 1.f;
    hipLaunchKernelGGL(k_legacy_pre, dim3(ew_blocks(c->pvox)), dim3(256), 0, c->stream, c->P, c->Pext, fl, decay,
                       do_decay ? 1 : 0, c->pvox);
    BD_HIP(hipGetLastError());
    BD_TRY(splat(c, c->kl, c->nkl, c->rl));
    hipLaunchKernelGGL(k_legacy_post, dim3(ew_blocks(c->pvox)), dim3(256), 0, c->stream, c->P, c->Pext, c->pvoid,
                       (float)p.mu_max, c->pvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// coupled_ablation.py:594-603: centre-column optical depth, peak, mean over ablated cells
// stage 1 of the centre-column / ablated-mean / peak diagnostics
// (coupled_ablation.py:594-603): per-block partial sum and count over ablated
// voxels of the plume value at the voxel's plume cell, and the block's peak bits;
// one write per block, no contended atomics
constexpr int PD_BLOCKS = 2048, PD_THREADS = 256;  // 3 x PD_BLOCKS partials fit the 8192-double scratch
template <bool W32>
__global__ __launch_bounds__(PD_THREADS) void k_plume_diag1(const float *__restrict__ E,
                                                            const uint8_t *__restrict__ flags, int Nx, int Ny,
                                                            int Nz, int Px, int Py, int Pz, float ra,
                                                            double *__restrict__ part)
{
    const int64_t NP = (int64_t)Px * Py * Pz;
    const int64_t N = (int64_t)Nx * Ny * Nz;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    double s = 0.0, cnt = 0.0;
    unsigned pk = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < NP; c += stride)
        pk = max(pk, __float_as_uint(E[c] * ra));
    auto visit = [&](int64_t c) {
        int i, j, k;
        split_ijk<W32>(c, Ny, Nz, i, j, k);
        s += (double)(E[plume_of<W32>(i, j, k, Nx, Ny, Nz, Px, Py, Pz)] * ra);
        cnt += 1.0;
    };
    // the removed-voxel mask, 16 flag bytes per load; most chunks hold none.  In
    // a chunk that does, the 16 plume gathers are issued together (unrolled,
    // predicated) and then summed in voxel order.
    int64_t c0 = 0;
    if (((uintptr_t)flags & 15) == 0) {
        const int64_t n16 = N / 16;
        const uint4 *f16 = (const uint4 *)flags;
        const uint32_t b4 = (uint32_t)F_ABL * 0x01010101u;
        for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n16; q += stride) {
            const uint4 w = f16[q];
            if (!((w.x | w.y | w.z | w.w) & b4)) continue;
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
            int i, j, k;
            split_ijk<W32>(q * 16, Ny, Nz, i, j, k);
            float v[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const bool on = (ws[e >> 2] >> (8 * (e & 3))) & F_ABL;
                v[e] = on ? E[plume_of<W32>(i, j, k, Nx, Ny, Nz, Px, Py, Pz)] * ra : -1.f;
                if (++k == Nz) {
                    k = 0;
                    if (++j == Ny) {
                        j = 0;
                        ++i;
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < 16; ++e)
                if (v[e] >= 0.f) {
                    s += (double)v[e];
                    cnt += 1.0;
                }
        }
        c0 = n16 * 16;
    }
    for (int64_t c = c0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < N; c += stride)
        if (flags[c] & F_ABL) visit(c);
    __shared__ double ss[PD_THREADS], sc[PD_THREADS];
    __shared__ unsigned sp[PD_THREADS];
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = cnt;
    sp[threadIdx.x] = pk;
    __syncthreads();
    for (int h = PD_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            ss[threadIdx.x] += ss[threadIdx.x + h];
            sc[threadIdx.x] += sc[threadIdx.x + h];
            sp[threadIdx.x] = max(sp[threadIdx.x], sp[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = ss[0];
        part[3 * blockIdx.x + 1] = sc[0];
        part[3 * blockIdx.x + 2] = (double)sp[0];  // exact: 32-bit value in a double
    }
}

// stage 2 (one block): fixed-order sums of the partials and the column tau
// summed along z in order; acc = {tau, sum, count, peak bits}
__global__ __launch_bounds__(PD_THREADS) void k_plume_diag2(const float *__restrict__ E, const double *__restrict__ part,
                                                            int nb, int Py, int Pz, float ra, int I, int J,
                                                            double dzp, double *__restrict__ acc)
{
    __shared__ double ss[PD_THREADS], sc[PD_THREADS];
    __shared__ unsigned sp[PD_THREADS];
    double s = 0.0, cnt = 0.0;
    unsigned pk = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        s += part[3 * b];
        cnt += part[3 * b + 1];
        pk = max(pk, (unsigned)part[3 * b + 2]);
    }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = cnt;
    sp[threadIdx.x] = pk;
    __syncthreads();
    for (int h = PD_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            ss[threadIdx.x] += ss[threadIdx.x + h];
            sc[threadIdx.x] += sc[threadIdx.x + h];
            sp[threadIdx.x] = max(sp[threadIdx.x], sp[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float *col = E + ((int64_t)I * Py + J) * Pz;
        double tau = 0.0;
        for (int k = 0; k < Pz; ++k) tau += (double)(col[k] * ra) * dzp;
        acc[0] = tau;
        acc[1] = ss[0];
        acc[2] = sc[0];
        unsigned bits = sp[0];
        double packed = 0.0;
        __builtin_memcpy(&packed, &bits, 4);
        acc[3] = packed;
    }
}

int launch_plume_diag_enqueue(Ctx *c)
{
    const bd_plume_params &p = c->plume;
    double *acc = c->red + 24;
    double *part = c->red + 64;  // the reduction partials (3 x PD_BLOCKS <= 8192 doubles)
    double dx, dy, dz;
    plume_spacing(c, dx, dy, dz);
    const int I = (int)(((int64_t)(2 * (c->g.nx / 2) + 1) * p.nx) / (2 * (int64_t)c->g.nx));
    const int J = (int)(((int64_t)(2 * (c->g.ny / 2) + 1) * p.ny) / (2 * (int64_t)c->g.ny));
    const float ra = (float)(p.mu_a_max / std::fmax(1e-6, p.mu_max));
    hipLaunchKernelGGL(map32_ok(c) ? k_plume_diag1<true> : k_plume_diag1<false>, dim3(PD_BLOCKS), dim3(PD_THREADS), 0,
                       c->stream, c->Pext, c->flags, c->g.nx, c->g.ny, c->g.nz, p.nx, p.ny, p.nz, ra, part);
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_plume_diag2, dim3(1), dim3(PD_THREADS), 0, c->stream, c->Pext, part, PD_BLOCKS, p.ny, p.nz,
                       ra, I, J, dz, acc);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

void plume_diag_from(const double *h, double out[4])
{
    unsigned bits;
    std::memcpy(&bits, (const char *)(h + 3), 4);
    float peak;
    std::memcpy(&peak, &bits, 4);
    out[0] = h[0];
    out[1] = std::exp(-h[0]);
    out[2] = (double)peak;
    out[3] = h[2] > 0.0 ? h[1] / h[2] : 0.0;
}

int launch_plume_diag(Ctx *c, double out[4])
{
    BD_TRY(launch_plume_diag_enqueue(c));
    BD_HIP(hipMemcpyAsync(c->host_pinned, c->red + 24, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    plume_diag_from((const double *)c->host_pinned, out);
    return BD_OK;
}

}  // namespace bd
