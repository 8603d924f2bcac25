// bd_surrogate.hip -- coupling of the surrogate network to an engine context:
// the 64^3 beam window, its input features, training targets from the MC
// tallies, composition of the absorbed-power field from a prediction, the
// judge error against an MC field, and an epoch that consumes a supplied Q.
//
// Windows: the first is laterally centred on the beam axis, its top a few
// voxels above the first non-void voxel the beam axis meets (the ablation
// front), so the beam always enters it through void as it does in the training
// samples.  Further windows are stacked above it along the beam axis, each
// overlapping the one below by RAMP voxels, until one reaches RAMP voxels
// above the tissue surface next to the crater (or the domain top): together
// they cover the crater walls, where most of the power is absorbed once the
// crater is deep.  Where windows overlap, their ramps sum to one.
//
// Features per voxel (channels last, [64][64][64][8]):
//   0 log1p(mu_a)   1 log1p(mu_s / 10)   2 g   3 n - 1
//   4 ballistic: ln(max(B, 1e-6)) / ln(1e6) + 1, B = collimated beam intensity
//     relative to the beam's peak, marched plane by plane from the domain top
//     (planar Snell invariant n sin(theta), Fresnel transmission at index
//     steps, Beer-Lambert attenuation mu_t dz / cos(theta), lateral drift
//     dz tan(theta) by linear interpolation)
//   5 sin(theta_in)   6 spot / 0.5 mm - 1
//   7 optical depth along the beam from the domain top, ln(1 + tau) / ln(101),
//     carried with the same drift as channel 4 (it keeps growing where the
//     beam intensity has fallen below the channel-4 floor)
// Optics are the packed transport optics the MC itself reads (plume included).
//
// Target: t = ln(Phi / E0 + 1e-4) with Phi = L / (N V) * P the fluence rate and
// E0 = P / (pi w0^2 / 2) the beam's peak irradiance; weights
// w = 0.5 Q^2 / sum Q^2 + 0.5 Phi / sum Phi (Q = mu_a Phi), summing to one.
#include "bd_internal.hpp"

#include <algorithm>
#include <cmath>

namespace bd {

constexpr int SW = 64;                    // window edge
constexpr float PHI_EPS = 1e-4f;          // offset of the log target
constexpr int RAMP = 8;                   // blend width at window faces
constexpr int MAXW = 4;                   // windows per beam (4 x 56 + 8 >= 232 planes)
constexpr int FX = 256;                   // x extent of the feature march (window + 192 up-beam)
constexpr int FUP = FX - SW;

__device__ __forceinline__ float fresnel_u(float n1, float n2, float ci)
{
    // unpolarised Fresnel reflectance (voxel_mc_3d.py:33-48)
    if (fabsf(n1 - n2) < 1e-9f) return 0.f;
    ci = fminf(1.f, fabsf(ci));
    if (ci > 0.999999f) {
        const float r = (n1 - n2) / (n1 + n2);
        return r * r;
    }
    const float e = n1 / n2;
    const float st2 = e * e * (1.f - ci * ci);
    if (st2 >= 1.f) return 1.f;
    const float ct = sqrtf(1.f - st2);
    const float rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct);
    const float rp = (n1 * ct - n2 * ci) / (n1 * ct + n2 * ci);
    return 0.5f * (rs * rs + rp * rp);
}

struct SurGeom {
    const float4 *opt;
    int Nx, Ny, Nz;
    float dx, dy, dz;
    int x0, y0, z0;
};

__device__ __forceinline__ bool in_dom(const SurGeom &g, int X, int Y, int Z)
{
    return X >= 0 && X < g.Nx && Y >= 0 && Y < g.Ny && Z >= 0 && Z < g.Nz;
}

// First non-void z along the beam axis (the ablation front) and first non-void
// z over nine probe columns around the axis, +-16 and +-28 voxels in x and y
// (the tissue surface next to the crater); block-parallel over z, min-reduced.
__global__ __launch_bounds__(256) void k_sur_window(const float *__restrict__ mu_a, const float *__restrict__ mu_s,
                                                    const uint8_t *__restrict__ flags, int Nx, int Ny, int Nz,
                                                    float dx, float dy, float dz, float xc, float yc, float tan0,
                                                    int sdir, int *__restrict__ out)
{
    __shared__ int sf[256], ss[256];
    const int Y = (int)floorf(yc / dy);
    int front = Nz, surf = Nz;
    for (int z = threadIdx.x; z < Nz; z += blockDim.x) {
        const float xm = xc + sdir * (z + 0.5f) * dz * tan0;
        const int X = (int)floorf(xm / dx);
        for (int p = 0; p < 9; ++p) {
            const int ox = p == 1 ? 16 : p == 2 ? -16 : p == 5 ? 28 : p == 6 ? -28 : 0;
            const int oy = p == 3 ? 16 : p == 4 ? -16 : p == 7 ? 28 : p == 8 ? -28 : 0;
            const int Xp = X + ox, Yp = Y + oy;
            if (Xp < 0 || Xp >= Nx || Yp < 0 || Yp >= Ny) continue;
            const int64_t v = ((int64_t)Xp * Ny + Yp) * Nz + z;
            if ((flags[v] & F_ABL) == 0 && (mu_a[v] + mu_s[v]) > 0.f) {
                if (p == 0) front = min(front, z);
                surf = min(surf, z);
            }
        }
    }
    sf[threadIdx.x] = front;
    ss[threadIdx.x] = surf;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sf[threadIdx.x] = min(sf[threadIdx.x], sf[threadIdx.x + h]);
            ss[threadIdx.x] = min(ss[threadIdx.x], ss[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sf[0];
        out[1] = min(ss[0], sf[0]);
    }
}

struct FeatArgs {
    SurGeom g;                 // x0/y0/z0 unused: windows below
    int4 win[MAXW];            // x0, y0, z0 of each window
    float sin0, cos0;
    int sdir;
    float xc, yc, w0, n_amb;   // beam axis at the domain top
    float f5, f6;
    const float2 *AL;          // when set, channel 5 is the log fluence of the context's last MC
    float l_scale;             // path-length tally -> fluence relative to the beam's peak irradiance
    float *x;                  // [nwin][64][64][64][8]
};

// One block per (window row j, window).  The march runs from the domain top to
// the window's last plane over x in [x0 - 192, x0 + 64) (mirrored for a beam
// tilted towards -x), so light drifting into the window from up to 192 voxels
// up-beam is carried; optics are loaded 8 planes ahead of the march, and the
// neighbour exchange is double-buffered in LDS (one barrier per plane).
__global__ __launch_bounds__(FX) void k_sur_features(FeatArgs a)
{
    const SurGeom &g = a.g;
    const int j = blockIdx.x, w = blockIdx.y, t = threadIdx.x;
    const int4 o = a.win[w];
    const int xlo = a.sdir > 0 ? o.x - FUP : o.x;   // x of thread 0
    const int tw = o.x - xlo;                       // first window thread
    const int X = xlo + t, Y = o.y + j;
    const bool col_in = X >= 0 && X < g.Nx && Y >= 0 && Y < g.Ny;
    const int64_t colv = col_in ? ((int64_t)X * g.Ny + Y) * g.Nz : 0;
    const int zend = o.z + SW;
    __shared__ float sB[2][FX], sT[2][FX];
    const float xm = (X + 0.5f) * g.dx, ym = (Y + 0.5f) * g.dy;
    const float u = (xm - a.xc) * a.cos0, v = ym - a.yc;
    float B = expf(-2.f * (u * u + v * v) / (a.w0 * a.w0));
    float tau = 0.f;
    float nprev = a.n_amb;
    const float s0 = a.n_amb * a.sin0;
    auto ld = [&](int Z) {
        float4 r = make_float4(0.f, 0.f, 0.f, a.n_amb);
        if (col_in && Z < g.Nz && Z < zend) r = g.opt[colv + Z];
        return r;
    };
    float4 ring[8];
#pragma unroll
    for (int u8 = 0; u8 < 8; ++u8) ring[u8] = ld(u8);
    float *xw = a.x + (int64_t)w * SW * SW * SW * 8;
    int buf = 0;
    for (int z8 = 0; z8 < zend; z8 += 8) {
#pragma unroll
        for (int u8 = 0; u8 < 8; ++u8) {
            const int Z = z8 + u8;
            if (Z >= zend) break;  // block-uniform
            const float4 op = ring[u8];
            ring[u8] = ld(Z + 8);
            sB[buf][t] = B;
            sT[buf][t] = tau;
            __syncthreads();
            const int tn = t - a.sdir;
            const bool nb_in = tn >= 0 && tn < FX;
            const float Bn = nb_in ? sB[buf][tn] : B;
            const float Tn = nb_in ? sT[buf][tn] : tau;
            buf ^= 1;
            const float sk = s0 / op.w;
            float Bk = 0.f, tk = tau + (op.x + op.y) * g.dz;
            if (sk < 1.f) {
                const float ck = sqrtf(1.f - sk * sk);
                const float sh = fminf(g.dz * sk / (ck * g.dx), 1.f);
                const float bi = (1.f - sh) * B + sh * Bn;
                const float sp = s0 / nprev;
                const float cp = sqrtf(fmaxf(0.f, 1.f - sp * sp));
                const float R = fresnel_u(nprev, op.w, cp);
                Bk = bi * (1.f - R) * expf(-(op.x + op.y) * g.dz / ck);
                tk = (1.f - sh) * tau + sh * Tn + (op.x + op.y) * g.dz / ck;
            }
            B = Bk;
            tau = tk;
            nprev = op.w;
            const int k = Z - o.z;
            if (k >= 0 && t >= tw && t < tw + SW) {
                float *xp = xw + (((int64_t)(t - tw) * SW + j) * SW + k) * 8;
                const float4 f0 = make_float4(log1pf(op.x), log1pf(op.y * 0.1f), op.z, op.w - 1.f);
                float f5 = a.f5;
                if (a.AL) {
                    const float phi = (col_in && Z < g.Nz) ? a.AL[colv + Z].y * a.l_scale : 0.f;
                    f5 = logf(phi + 1e-4f) * (1.f / 9.210340371976184f) + 1.f;
                }
                const float4 f1 = make_float4(logf(fmaxf(B, 1e-6f)) * (1.f / 13.815510557964274f) + 1.f, f5, a.f6,
                                              log1pf(tau) * (1.f / 4.61512051684126f));
                reinterpret_cast<float4 *>(xp)[0] = f0;
                reinterpret_cast<float4 *>(xp)[1] = f1;
            }
        }
    }
}

// per-plane sums of Q^2 and Phi over the window (FP64), one block per plane i
__global__ __launch_bounds__(256) void k_sur_tsum(SurGeom g, const float2 *__restrict__ AL, float scale,
                                                  double *__restrict__ part)
{
    __shared__ double sq[256], sp[256];
    const int i = blockIdx.x;
    double q2 = 0.0, ph = 0.0;
    for (int idx = threadIdx.x; idx < SW * SW; idx += blockDim.x) {
        const int j = idx / SW, k = idx % SW;
        const int X = g.x0 + i, Y = g.y0 + j, Z = g.z0 + k;
        if (!in_dom(g, X, Y, Z)) continue;
        const int64_t v = ((int64_t)X * g.Ny + Y) * g.Nz + Z;
        const float phi = AL[v].y * scale;
        const float q = g.opt[v].x * phi;
        q2 += (double)q * q;
        ph += phi;
    }
    sq[threadIdx.x] = q2;
    sp[threadIdx.x] = ph;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sq[threadIdx.x] += sq[threadIdx.x + h];
            sp[threadIdx.x] += sp[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * i] = sq[0];
        part[2 * i + 1] = sp[0];
    }
}

__global__ __launch_bounds__(256) void k_sur_twrite(SurGeom g, const float2 *__restrict__ AL, float scale,
                                                    const double *__restrict__ part, float *__restrict__ tg,
                                                    float *__restrict__ wt)
{
    double tq = 0.0, tp = 0.0;
    for (int i = 0; i < SW; ++i) {
        tq += part[2 * i];
        tp += part[2 * i + 1];
    }
    const float iq = tq > 0.0 ? (float)(0.5 / tq) : 0.f, ip = tp > 0.0 ? (float)(0.5 / tp) : 0.f;
    const float uni = 1.f / (float)(SW * SW * SW);
    const int i = blockIdx.x;
    for (int idx = threadIdx.x; idx < SW * SW; idx += blockDim.x) {
        const int j = idx / SW, k = idx % SW;
        const int X = g.x0 + i, Y = g.y0 + j, Z = g.z0 + k;
        float phi = 0.f, q = 0.f;
        if (in_dom(g, X, Y, Z)) {
            const int64_t v = ((int64_t)X * g.Ny + Y) * g.Nz + Z;
            phi = AL[v].y * scale;
            q = g.opt[v].x * phi;
        }
        const int64_t o = ((int64_t)i * SW + j) * SW + k;
        tg[o] = logf(phi + PHI_EPS);
        wt[o] = (tq > 0.0 ? q * q * iq : 0.5f * uni) + (tp > 0.0 ? phi * ip : 0.5f * uni);
    }
}

__device__ __forceinline__ float ramp(int i, bool lo, bool hi)
{
    float r = 1.f;
    if (lo) r = fminf(r, (i + 0.5f) / RAMP);
    if (hi) r = fminf(r, (SW - i - 0.5f) / RAMP);
    return r;
}

struct ApplyArgs {
    SurGeom g;              // x0/y0/z0: corner of the windows' bounding box
    int4 win[MAXW];         // x0, y0, z0, blend mask of each window
    int nwin;
    int bx, by, bz;         // bounding-box extent
    const float *y;         // [nwin][64^3] predicted log fluence
    const float *Qs;
    float *Qo;
    float E0;
};

// Q over the union of the windows: per voxel, s = sum_k r_k q_k and W = sum_k
// r_k over the windows holding it (r_k: product of the face ramps of window k
// at its faces inside the domain); Q = s / max(W, 1) + (1 - min(W, 1)) Q_stale.
// In an overlap of RAMP voxels the two ramps sum to one, so the windows hand
// over to each other; at the outer faces the prediction blends into the stale
// field.
__global__ void k_sur_apply(ApplyArgs a)
{
    const SurGeom &g = a.g;
    const int64_t nb = (int64_t)a.bx * a.by * a.bz;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nb; t += (int64_t)gridDim.x * blockDim.x) {
        const int Z = g.z0 + (int)(t % a.bz), Y = g.y0 + (int)((t / a.bz) % a.by), X = g.x0 + (int)(t / ((int64_t)a.bz * a.by));
        if (!in_dom(g, X, Y, Z)) continue;
        const int64_t v = ((int64_t)X * g.Ny + Y) * g.Nz + Z;
        float sq = 0.f, W = 0.f;
        const float ma = g.opt[v].x;
        for (int k = 0; k < a.nwin; ++k) {
            const int4 o = a.win[k];
            const int i = X - o.x, j = Y - o.y, l = Z - o.z;
            if (i < 0 || i >= SW || j < 0 || j >= SW || l < 0 || l >= SW) continue;
            const int m = o.w;
            const float r = ramp(i, m & 1, m & 2) * ramp(j, m & 4, m & 8) * ramp(l, m & 16, m & 32);
            const float yv = a.y[(int64_t)k * SW * SW * SW + ((int64_t)i * SW + j) * SW + l];
            sq += r * (a.E0 * ma * fmaxf(expf(yv) - PHI_EPS, 0.f));
            W += r;
        }
        if (W > 0.f) a.Qo[v] = sq / fmaxf(W, 1.f) + (1.f - fminf(W, 1.f)) * a.Qs[v];
    }
}

__global__ void k_copy_f32(const float *__restrict__ s, float *__restrict__ d, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

// Coarse amplitudes from a second field: per block of e^3 voxels (e <= 8), the
// non-ablated voxels of F are scaled by sum R / sum F over the block, when both
// sums are positive (otherwise F is left as it is).  One workgroup per block,
// which owns its voxels, so the sums and the scaling share one pass.
__global__ __launch_bounds__(256) void k_sur_correct(float *__restrict__ F, const float *__restrict__ R,
                                                     const uint8_t *__restrict__ flags, int Nx, int Ny, int Nz,
                                                     int e)
{
    __shared__ double sf[256], sr[256];
    const int nbz = (Nz + e - 1) / e, nby = (Ny + e - 1) / e;
    const int bz = (int)(blockIdx.x % nbz), by = (int)((blockIdx.x / nbz) % nby), bx = (int)(blockIdx.x / (nbz * nby));
    const int n = e * e * e;  // <= 512: two voxels per thread
    int64_t vv[2];
    bool ok[2];
    float f[2], r[2];
    double tf = 0.0, tr = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int l = threadIdx.x + 256 * h;
        const int X = bx * e + l / (e * e), Y = by * e + (l / e) % e, Z = bz * e + l % e;
        ok[h] = l < n && X < Nx && Y < Ny && Z < Nz;
        vv[h] = ok[h] ? ((int64_t)X * Ny + Y) * Nz + Z : 0;
        ok[h] = ok[h] && !(flags[vv[h]] & F_ABL);
        f[h] = ok[h] ? F[vv[h]] : 0.f;
        r[h] = ok[h] ? R[vv[h]] : 0.f;
        tf += f[h];
        tr += r[h];
    }
    sf[threadIdx.x] = tf;
    sr[threadIdx.x] = tr;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sf[threadIdx.x] += sf[threadIdx.x + o];
            sr[threadIdx.x] += sr[threadIdx.x + o];
        }
        __syncthreads();
    }
    const double SF = sf[0], SR = sr[0];
    if (!(SF > 0.0 && SR > 0.0)) return;  // block-uniform
    const float ratio = (float)(SR / SF);
#pragma unroll
    for (int h = 0; h < 2; ++h)
        if (ok[h]) F[vv[h]] = f[h] * ratio;
}

// judge error: per-block partials of sum (a - b)^2, sum a^2 over non-ablated
// voxels, the same over 4^3 block sums (MC shot noise averaged down), and the
// totals sum a, sum b
__global__ __launch_bounds__(256) void k_sur_err(const float *__restrict__ a, const float *__restrict__ b,
                                                 const uint8_t *__restrict__ flags, int Nx, int Ny, int Nz,
                                                 double *__restrict__ part)
{
    __shared__ double s[6][256];
    double d2 = 0.0, a2 = 0.0, bd2 = 0.0, ba2 = 0.0, ta = 0.0, tb = 0.0;
    const int bx = (Nx + 3) / 4, by = (Ny + 3) / 4, bz = (Nz + 3) / 4;
    const int64_t nb = (int64_t)bx * by * bz;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nb; t += (int64_t)gridDim.x * blockDim.x) {
        const int kz = (int)(t % bz), jy = (int)((t / bz) % by), ix = (int)(t / ((int64_t)bz * by));
        double sa = 0.0, sb = 0.0;
        for (int i = 4 * ix; i < 4 * ix + 4 && i < Nx; ++i)
            for (int j = 4 * jy; j < 4 * jy + 4 && j < Ny; ++j)
                for (int k = 4 * kz; k < 4 * kz + 4 && k < Nz; ++k) {
                    const int64_t v = ((int64_t)i * Ny + j) * Nz + k;
                    if (flags[v] & F_ABL) continue;
                    const double x = a[v], w = b[v];
                    d2 += (x - w) * (x - w);
                    a2 += x * x;
                    sa += x;
                    sb += w;
                }
        bd2 += (sa - sb) * (sa - sb);
        ba2 += sa * sa;
        ta += sa;
        tb += sb;
    }
    s[0][threadIdx.x] = d2;
    s[1][threadIdx.x] = a2;
    s[2][threadIdx.x] = bd2;
    s[3][threadIdx.x] = ba2;
    s[4][threadIdx.x] = ta;
    s[5][threadIdx.x] = tb;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 6; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) part[6 * blockIdx.x + q] = s[q][0];
}

__global__ void k_sur_err_final(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    if (threadIdx.x || blockIdx.x) return;
    double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < nb; ++b)
        for (int q = 0; q < 6; ++q) r[q] += part[6 * b + q];
    for (int q = 0; q < 6; ++q) out[q] = r[q];
}

namespace {

struct BeamFrame {
    float sin0, cos0, tan0;
    int sdir;
    float w0;
};

int beam_frame(const bd_beam *b, BeamFrame &f)
{
    const double c = std::cos(b->phi_rad), s = std::sin(b->phi_rad);
    if (b->launch_axis != 0 || std::fabs(s) > 1e-9) {
        set_error("surrogate: only a beam launched along +z with azimuth 0 or pi is supported");
        return BD_EINVAL;
    }
    if (!(b->theta_rad >= 0.0 && b->theta_rad < 1.2) || !(b->spot_dia > 0.0)) {
        set_error("surrogate: beam angle must be in [0, 1.2) rad and the spot > 0");
        return BD_EINVAL;
    }
    f.sin0 = (float)std::sin(b->theta_rad);
    f.cos0 = (float)std::cos(b->theta_rad);
    f.tan0 = (float)std::tan(b->theta_rad);
    f.sdir = c >= 0.0 ? 1 : -1;
    f.w0 = (float)(b->spot_dia / 2.0);
    return BD_OK;
}

SurGeom geom(Ctx *c, const int32_t *o)
{
    SurGeom g{};
    g.opt = c->optics;
    g.Nx = c->g.nx;
    g.Ny = c->g.ny;
    g.Nz = c->g.nz;
    g.dx = (float)c->g.dx;
    g.dy = (float)c->g.dy;
    g.dz = (float)c->g.dz;
    g.x0 = o[0];
    g.y0 = o[1];
    g.z0 = o[2];
    return g;
}

int check_origin(Ctx *c, const int32_t *o)
{
    if (!o) {
        set_error("surrogate: null window origin");
        return BD_EINVAL;
    }
    const int n[3] = {c->g.nx, c->g.ny, c->g.nz};
    for (int d = 0; d < 3; ++d)
        if (o[d] < 0 || (n[d] >= SW && o[d] > n[d] - SW) || (n[d] < SW && o[d] != 0)) {
            set_error("surrogate: window origin outside the domain");
            return BD_EINVAL;
        }
    return BD_OK;
}

}  // namespace
}  // namespace bd

using namespace bd;

extern "C" {

int bd_sur_windows(bd_ctx *h, const bd_beam *beam, int margin, int max_windows, int32_t *origins, int *n_windows)
{
    BD_CHECK(h && beam && origins && n_windows, "bd_sur_windows: null argument");
    BD_CHECK(margin >= 0 && margin < SW, "bd_sur_windows: margin must be in [0, 64)");
    BD_CHECK(max_windows >= 1 && max_windows <= MAXW, "bd_sur_windows: max_windows must be in [1, 4]");
    Ctx *c = &h->c;
    BD_TRY(ensure_state(c));
    BD_CHECK(c->mu_a && c->mu_s, "bd_sur_windows: optics fields are not uploaded");
    BeamFrame f;
    BD_TRY(beam_frame(beam, f));
    int *dev = (int *)(c->counter + 8);
    hipLaunchKernelGGL(k_sur_window, dim3(1), dim3(256), 0, c->stream, c->mu_a, c->mu_s, c->flags, c->g.nx, c->g.ny,
                       c->g.nz, (float)c->g.dx, (float)c->g.dy, (float)c->g.dz, (float)beam->x_center,
                       (float)beam->y_center, f.tan0, f.sdir, dev);
    BD_HIP(hipGetLastError());
    BD_HIP(hipMemcpyAsync(c->host_pinned, dev, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    const int front = ((const int *)c->host_pinned)[0], surf = ((const int *)c->host_pinned)[1];
    const int Nx = c->g.nx, Ny = c->g.ny, Nz = c->g.nz;
    const double tan0 = std::tan(beam->theta_rad);
    int y0 = (int)std::nearbyint(beam->y_center / c->g.dy - 0.5 * SW);
    y0 = std::max(0, std::min(y0, Ny - SW));
    int z0 = std::max(0, std::min(front - margin, Nz - SW));
    int n = 0;
    for (;;) {
        const double xmid = (beam->x_center + f.sdir * (z0 + 0.5 * SW) * c->g.dz * tan0) / c->g.dx;
        int x0 = (int)std::nearbyint(xmid - 0.5 * SW);
        x0 = std::max(0, std::min(x0, Nx - SW));
        origins[4 * n] = x0;
        origins[4 * n + 1] = y0;
        origins[4 * n + 2] = z0;
        origins[4 * n + 3] = front;
        ++n;
        if (z0 <= 0 || z0 <= surf - RAMP || n == max_windows) break;
        z0 = std::max(z0 - (SW - RAMP), 0);
    }
    *n_windows = n;
    return BD_OK;
}

int bd_sur_window(bd_ctx *h, const bd_beam *beam, int margin, int32_t *origin)
{
    BD_CHECK(origin, "bd_sur_window: null argument");
    int32_t o[4 * MAXW];
    int n = 0;
    BD_TRY(bd_sur_windows(h, beam, margin, 1, o, &n));
    for (int d = 0; d < 4; ++d) origin[d] = o[d];
    return BD_OK;
}

int bd_sur_features_mc(bd_ctx *h, const bd_beam *beam, int n, const int32_t *origins, float *x_dev,
                       int64_t mc_photons)
{
    BD_CHECK(h && beam && origins && x_dev, "bd_sur_features: null argument");
    BD_CHECK(n >= 1 && n <= MAXW, "bd_sur_features: window count must be in [1, 4]");
    BD_CHECK(mc_photons >= 0, "bd_sur_features: mc_photons must be >= 0");
    Ctx *c = &h->c;
    BD_CHECK(mc_photons == 0 || c->AL, "bd_sur_features: no MC tallies on the context (run bd_mc_run first)");
    BD_TRY(lag_finalize(c));
    BD_TRY(ensure_state(c));
    for (int k = 0; k < n; ++k) BD_TRY(check_origin(c, origins + 4 * k));
    BeamFrame f;
    BD_TRY(beam_frame(beam, f));
    BD_TRY(launch_pack_optics(c, beam->add_plume != 0));
    FeatArgs a{};
    a.g = geom(c, origins);
    for (int k = 0; k < n; ++k) a.win[k] = make_int4(origins[4 * k], origins[4 * k + 1], origins[4 * k + 2], 0);
    a.sin0 = f.sin0;
    a.cos0 = f.cos0;
    a.sdir = f.sdir;
    a.xc = (float)beam->x_center;
    a.yc = (float)beam->y_center;
    a.w0 = f.w0;
    a.n_amb = (float)beam->n_ambient;
    a.f5 = f.sin0;
    a.f6 = (float)(beam->spot_dia / 0.5 - 1.0);
    if (mc_photons > 0) {
        a.AL = c->AL;
        a.l_scale = (float)(M_PI * (double)f.w0 * f.w0 / 2.0 /
                            ((double)mc_photons * c->g.dx * c->g.dy * c->g.dz));
    }
    a.x = x_dev;
    hipLaunchKernelGGL(k_sur_features, dim3(SW, n), dim3(FX), 0, c->stream, a);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int bd_sur_features_n(bd_ctx *h, const bd_beam *beam, int n, const int32_t *origins, float *x_dev)
{
    return bd_sur_features_mc(h, beam, n, origins, x_dev, 0);
}

int bd_sur_features(bd_ctx *h, const bd_beam *beam, const int32_t *origin, float *x_dev)
{
    return bd_sur_features_mc(h, beam, 1, origin, x_dev, 0);
}

int bd_sur_targets(bd_ctx *h, const bd_beam *beam, const int32_t *origin, float *t_dev, float *w_dev)
{
    BD_CHECK(h && beam && t_dev && w_dev, "bd_sur_targets: null argument");
    Ctx *c = &h->c;
    BD_TRY(check_origin(c, origin));
    BD_CHECK(c->AL && c->optics, "bd_sur_targets: no MC tallies on the context (run bd_mc_run first)");
    BD_TRY(lag_finalize(c));
    BD_CHECK(beam->n_photons > 0, "bd_sur_targets: n_photons must be > 0");
    BeamFrame f;
    BD_TRY(beam_frame(beam, f));
    const double V = c->g.dx * c->g.dy * c->g.dz;
    const float scale = (float)(M_PI * (double)f.w0 * f.w0 / 2.0 / ((double)beam->n_photons * V));
    const SurGeom g = geom(c, origin);
    double *part = c->red + 64;
    hipLaunchKernelGGL(k_sur_tsum, dim3(SW), dim3(256), 0, c->stream, g, c->AL, scale, part);
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sur_twrite, dim3(SW), dim3(256), 0, c->stream, g, c->AL, scale, part, t_dev, w_dev);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int bd_sur_apply_n(bd_ctx *h, const bd_beam *beam, int n, const int32_t *origins, const float *y_dev, int into_q,
                   double scale)
{
    BD_CHECK(scale > 0.0 && std::isfinite(scale), "bd_sur_apply: scale must be positive and finite");
    BD_CHECK(h && beam && origins && y_dev, "bd_sur_apply: null argument");
    BD_CHECK(n >= 1 && n <= MAXW, "bd_sur_apply: window count must be in [1, 4]");
    Ctx *c = &h->c;
    BD_TRY(ensure_state(c));
    for (int k = 0; k < n; ++k) BD_TRY(check_origin(c, origins + 4 * k));
    BD_CHECK(c->optics, "bd_sur_apply: optics are not packed (call bd_sur_features first)");
    BeamFrame f;
    BD_TRY(beam_frame(beam, f));
    BD_TRY(ensure(c, (void **)&c->Qs, (size_t)c->nvox * 4));
    float *dst = c->Q;
    if (!into_q) {
        BD_TRY(ensure(c, (void **)&c->Qsur, (size_t)c->nvox * 4));
        dst = c->Qsur;
    }
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_copy_f32, dim3(nb), dim3(256), 0, c->stream, c->Qs, dst, c->nvox);
    BD_HIP(hipGetLastError());
    ApplyArgs a{};
    const int dims[3] = {c->g.nx, c->g.ny, c->g.nz};
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {0, 0, 0};
    for (int k = 0; k < n; ++k) {
        const int32_t *o = origins + 4 * k;
        int mask = 0;
        for (int d = 0; d < 3; ++d) {
            if (o[d] > 0) mask |= 1 << (2 * d);
            if (o[d] + SW < dims[d]) mask |= 2 << (2 * d);
            lo[d] = std::min(lo[d], o[d]);
            hi[d] = std::max(hi[d], std::min(o[d] + SW, dims[d]));
        }
        a.win[k] = make_int4(o[0], o[1], o[2], mask);
    }
    const int32_t corner[3] = {lo[0], lo[1], lo[2]};
    a.g = geom(c, corner);
    a.nwin = n;
    a.bx = hi[0] - lo[0];
    a.by = hi[1] - lo[1];
    a.bz = hi[2] - lo[2];
    a.y = y_dev;
    a.Qs = c->Qs;
    a.Qo = dst;
    a.E0 = (float)(scale * beam->power / (M_PI * (double)f.w0 * f.w0 / 2.0));
    hipLaunchKernelGGL(k_sur_apply, dim3(2048), dim3(256), 0, c->stream, a);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int bd_sur_apply(bd_ctx *h, const bd_beam *beam, const int32_t *origin, const float *y_dev, int into_q)
{
    return bd_sur_apply_n(h, beam, 1, origin, y_dev, into_q, 1.0);
}

int bd_sur_keep_stale(bd_ctx *h)
{
    BD_CHECK(h, "bd_sur_keep_stale: null argument");
    Ctx *c = &h->c;
    BD_TRY(ensure_state(c));
    BD_TRY(ensure(c, (void **)&c->Qs, (size_t)c->nvox * 4));
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_copy_f32, dim3(nb), dim3(256), 0, c->stream, c->Q, c->Qs, c->nvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int bd_sur_judge(bd_ctx *h, double *out)
{
    BD_CHECK(h && out, "bd_sur_judge: null argument");
    Ctx *c = &h->c;
    BD_CHECK(c->Qsur && c->Q, "bd_sur_judge: no surrogate field to compare (bd_sur_apply with into_q = 0)");
    constexpr int NB = 256;
    double *part = c->red + 64;
    hipLaunchKernelGGL(k_sur_err, dim3(NB), dim3(256), 0, c->stream, c->Q, c->Qsur, c->flags, c->g.nx, c->g.ny,
                       c->g.nz, part);
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sur_err_final, dim3(1), dim3(64), 0, c->stream, part, NB, c->red + 48);
    BD_HIP(hipGetLastError());
    BD_HIP(hipMemcpyAsync(c->host_pinned, c->red + 48, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    const double *r = (const double *)c->host_pinned;
    out[0] = r[1] > 0.0 ? std::sqrt(r[0] / r[1]) : 0.0;
    out[1] = r[3] > 0.0 ? std::sqrt(r[2] / r[3]) : 0.0;
    out[2] = r[4];
    out[3] = r[5];
    return BD_OK;
}

int bd_sur_keep_aux(bd_ctx *h)
{
    BD_CHECK(h, "bd_sur_keep_aux: null argument");
    Ctx *c = &h->c;
    BD_TRY(ensure_state(c));
    BD_TRY(ensure(c, (void **)&c->Qaux, (size_t)c->nvox * 4));
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_copy_f32, dim3(nb), dim3(256), 0, c->stream, c->Q, c->Qaux, c->nvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int bd_sur_correct(bd_ctx *h, int block, int into_q)
{
    BD_CHECK(h, "bd_sur_correct: null argument");
    BD_CHECK(block >= 1 && block <= 8, "bd_sur_correct: block edge must be in [1, 8]");
    Ctx *c = &h->c;
    BD_CHECK(c->Qaux, "bd_sur_correct: no kept MC field (bd_sur_keep_aux)");
    float *F = into_q ? c->Q : c->Qsur;
    BD_CHECK(F && c->flags, "bd_sur_correct: no composed field (bd_sur_apply_n)");
    const int64_t nbx = (c->g.nx + block - 1) / block, nby = (c->g.ny + block - 1) / block,
                  nbz = (c->g.nz + block - 1) / block;
    const int64_t nb = nbx * nby * nbz;
    BD_CHECK(nb < (int64_t(1) << 31), "bd_sur_correct: too many blocks");
    hipLaunchKernelGGL(k_sur_correct, dim3((unsigned)nb), dim3(256), 0, c->stream, F, c->Qaux, c->flags, c->g.nx,
                       c->g.ny, c->g.nz, block);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

int bd_sur_judge_aux(bd_ctx *h, double *out)
{
    BD_CHECK(h && out, "bd_sur_judge_aux: null argument");
    Ctx *c = &h->c;
    BD_CHECK(c->Qaux && c->Q, "bd_sur_judge_aux: no kept MC field (bd_sur_keep_aux)");
    constexpr int NB = 256;
    double *part = c->red + 64;
    hipLaunchKernelGGL(k_sur_err, dim3(NB), dim3(256), 0, c->stream, c->Q, c->Qaux, c->flags, c->g.nx, c->g.ny,
                       c->g.nz, part);
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sur_err_final, dim3(1), dim3(64), 0, c->stream, part, NB, c->red + 48);
    BD_HIP(hipGetLastError());
    BD_HIP(hipMemcpyAsync(c->host_pinned, c->red + 48, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    const double *r = (const double *)c->host_pinned;
    out[0] = r[1] > 0.0 ? std::sqrt(r[0] / r[1]) : 0.0;
    out[1] = r[3] > 0.0 ? std::sqrt(r[2] / r[3]) : 0.0;
    out[2] = r[4];
    out[3] = r[5];
    return BD_OK;
}

int bd_sur_error(bd_ctx *h, double *out)
{
    BD_CHECK(h && out, "bd_sur_error: null argument");
    double r[4];
    BD_TRY(bd_sur_judge(h, r));
    out[0] = r[0];
    out[1] = r[1];
    return BD_OK;
}

}  // extern "C"
