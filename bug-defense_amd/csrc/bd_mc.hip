// bd_mc.hip -- voxel Monte Carlo photon transport for gfx950.
//
// Physics: voxel_mc_3d.py:33-337 (_fresnel_unpolarized, _refract_axis,
// _dist_to_voxel_exit, _sample_beam_offsets, _trace_one_photon) -- optical-depth
// marching voxel by voxel, Henyey-Greenstein scattering, Russian roulette,
// Fresnel/TIR and Snell refraction at every index step and at the entrance and
// exit faces; the CPU semantics (cos_th_safe = max(cos, 0.05), |dn| > 1e-9,
// r_max after every move, lateral escape per depth) are kept.
//
// Execution: a persistent kernel with one photon packet per lane.  When a
// lane's packet terminates, the wave refills every idle lane with one atomic
// on a global photon counter (ballot + prefix rank), so the wave stays full
// while photons of very different lengths finish.  Each photon id owns its
// own Philox4x32-10 stream (key = seed, counter = (draw block, id, tag)), so the
// traced photons do not depend on scheduling; only the float-atomic summation
// order of the voxel tallies does.
//
// Tallies: absorbed weight A and path length L are interleaved per voxel
// (float2) so the two atomics of a collision hit one cache line, and each lane
// coalesces consecutive contributions to the same voxel in registers before
// issuing them (one atomic pair per voxel visit, not per event).  The scalar
// tallies (specular, diffuse, transmitted, lateral, absorbed, events) are kept
// per lane in FP64 and reduced once per wave, with the count of voxel-tally
// atomics issued (the transport's atomic rate is reported against a measured
// ceiling, bd_probe_atomic_rate).
#include "bd_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <string>

namespace bd {

struct McArgs {
    const float4 *__restrict__ opt;   // {mu_a, mu_s, g, n}
    float2 *__restrict__ AL;
    double *__restrict__ acc;         // 0 spec, 1 diffuse, 2 trans, 3 lateral, 4 absorbed, 5 events,
                                      // 6 tally atomics, 7 path-length records
    double *__restrict__ lat;
    unsigned long long *__restrict__ counter;
    int Nx, Ny, Nz;
    float dx, dy, dz, Lx, Ly, Lz;
    float sigma, R;
    int profile, la;
    float iu0, iu1, iu2;
    double cos_in_d;                  // entrance direction cosine, FP64
    float cos_ph, sin_ph, cts;
    float xc, yc, zc;
    double n_amb;
    float rmax2;
    long long n_photons, photon_offset, max_events;
    unsigned k0, k1;
    int score_L;
    const EpochParams *ep;            // when set, the photon key is mix_seed(seed_base, epoch)
    // deferred path-length tally: when rec is set, L contributions go to the record buffer as
    // {voxel, float bits} (REC_NONE marks an unused slot), added afterwards by the record passes
    uint2 *__restrict__ rec;
    unsigned long long *__restrict__ rec_cursor;  // record slots handed out, in chunks of REC_CHUNK
    long long rec_cap;
    // where path lengths added inline go: Lt[voxel * Ls] (the L half of AL, or a separate grid)
    float *__restrict__ Lt;
    int Ls;
};

constexpr unsigned TAG = 0x6d63u;
constexpr unsigned REC_NONE = 0xffffffffu;
constexpr int INTER_EVERY = 4;   // interaction iterations of a burst: every 4th (see mc_trace)
constexpr int REC_CHUNK = 2048;  // record slots a wave takes at a time (one counter atomic)

// A photon's stream is the sequence of Philox blocks (blk = 0, 1, 2, ...), each giving four
// 32-bit draws, consumed in order.  The lane holds the block being consumed (c*, from index idx)
// and the following block (n*), generated ahead.  A draw never runs Philox itself: when the
// current block is used up it takes the next one and marks it for regeneration (need).  The
// only Philox is in rng_top(), called where the wave is converged (top of every event, launch,
// each trial of the rejection sampler); with the next block present that leaves at least four
// draws, the most any one event takes.  Inlined at every draw site, Philox (10 rounds) was
// issued once per site for the whole wave whenever any lane there needed a block.
struct Rng {
    unsigned k0, k1, p0, p1, blk;
    unsigned c0, c1, c2, c3;
    unsigned n0, n1, n2, n3;
    int idx;
    bool need;
};

__device__ __forceinline__ void philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0, unsigned k1,
                                       unsigned &o0, unsigned &o1, unsigned &o2, unsigned &o3)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const unsigned hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const unsigned hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    o0 = c0;
    o1 = c1;
    o2 = c2;
    o3 = c3;
}

__device__ __forceinline__ void rng_top(Rng &r)
{
    if (r.need) {
        philox(r.blk, r.p0, r.p1, TAG, r.k0, r.k1, r.n0, r.n1, r.n2, r.n3);
        r.blk++;
        r.need = false;
    }
}

// A fresh stream for photon id pid: nothing consumed, block 0 still to generate (by rng_top).
__device__ __forceinline__ void rng_start(Rng &r, unsigned k0, unsigned k1, unsigned long long pid)
{
    r.k0 = k0;
    r.k1 = k1;
    r.p0 = (unsigned)pid;
    r.p1 = (unsigned)(pid >> 32);
    r.blk = 0;
    r.idx = 4;
    r.need = true;
}

__device__ __forceinline__ float rng_u(Rng &r)
{
    if (r.idx == 4) {
        r.c0 = r.n0;
        r.c1 = r.n1;
        r.c2 = r.n2;
        r.c3 = r.n3;
        r.idx = 0;
        r.need = true;
    }
    const unsigned x = r.idx == 0 ? r.c0 : (r.idx == 1 ? r.c1 : (r.idx == 2 ? r.c2 : r.c3));
    r.idx++;
    return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float fresnel(float n1, float n2, float ci)
{
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

__device__ __forceinline__ double fresnel_d(double n1, double n2, double ci)
{
    if (fabs(n1 - n2) < 1e-9) return 0.0;
    ci = fmin(1.0, fabs(ci));
    if (ci > 0.999999) {
        const double r = (n1 - n2) / (n1 + n2);
        return r * r;
    }
    const double st2 = (n1 / n2) * (n1 / n2) * (1.0 - ci * ci);
    if (st2 >= 1.0) return 1.0;
    const double ct = sqrt(1.0 - st2);
    const double rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct);
    const double rp = (n1 * ct - n2 * ci) / (n1 * ct + n2 * ci);
    return 0.5 * (rs * rs + rp * rp);
}

__device__ __forceinline__ void refract(float &ua, float &ub, float &uc, float eta)
{
    // ua: component normal to the face; ub, uc tangential
    const float sgn = ua > 0.f ? 1.f : -1.f;
    ub *= eta;
    uc *= eta;
    ua = sgn * sqrtf(fmaxf(0.f, 1.f - ub * ub - uc * uc));
}

struct Photon {
    float x, y, z, ux, uy, uz, w, tau;
    int ix, iy, iz;
    float4 o;
    long long ev;
};

__device__ __forceinline__ void mc_trace(const McArgs &a)
{
    const int lane = threadIdx.x & 63;
    const int Nx = a.Nx, Ny = a.Ny, Nz = a.Nz;
    const int depth_axis = a.la == 0 ? 2 : 0;
    unsigned key0 = a.k0, key1 = a.k1;
    if (a.ep) {
        const unsigned long long key = mix_seed(a.ep->seed_base, a.ep->epoch);
        key0 = (unsigned)key;
        key1 = (unsigned)(key >> 32);
    }
    Photon p;
    Rng rng;
    bool alive = false;
    bool exhausted = false;
    int pv = -1;
    float pA = 0.f, pL = 0.f;
    double t_spec = 0.0, t_diff = 0.0, t_trans = 0.0, t_lat = 0.0, t_abs = 0.0;
    long long t_ev = 0;
    unsigned t_at = 0;   // voxel-tally atomics this lane issued
    unsigned t_rec = 0;  // path-length records this lane appended

    // Deferred path-length tally (a.rec set).  A photon adds a path length to nearly every voxel
    // it crosses, and consecutive crossings are different voxels, so those additions do not
    // coalesce per lane; as atomics they are most of the transport's time.  Instead the lane's
    // contribution to a voxel it leaves waits in (ov, oL) until the wave's next append, which
    // writes every waiting record of the wave with one coalesced store into the wave's chunk of
    // the record buffer.  rec_append() runs only where all 64 lanes are active, so the chunk
    // bounds (wcur, wlim) and `defer` stay identical across the wave.  A lane with a record
    // already waiting (a second voxel change in one event) adds the new one inline.
    bool defer = a.rec != nullptr;
    int ov = -1;
    float oL = 0.f;
    int wcur = 0, wlim = 0;  // record slots; the capacity is at most 2^28

    auto flush = [&]() {
        if (pv >= 0) {
            float *cell = (float *)&a.AL[pv];
            if (pA != 0.f) {
                unsafeAtomicAdd(cell, pA);
                ++t_at;
            }
            if (a.score_L && pL != 0.f) {
                if (defer && ov < 0) {
                    ov = pv;
                    oL = pL;
                } else {
                    unsafeAtomicAdd(a.Lt + (int64_t)pv * a.Ls, pL);
                    ++t_at;
                }
            }
        }
        pv = -1;
        pA = 0.f;
        pL = 0.f;
    };
    auto tally = [&](int v, float dA, float dL) {
        if (v != pv) {
            flush();
            pv = v;
        }
        pA += dA;
        pL += dL;
    };
    auto die_lateral = [&](int depth_idx, float wt) {
        t_lat += (double)wt;
        unsafeAtomicAdd(&a.lat[depth_idx], (double)wt);
        alive = false;
        flush();
    };
    auto rec_append = [&]() {
        if (!defer) return;
        const unsigned long long m = __ballot(ov >= 0);
        if (!m) return;
        const int cnt = __popcll(m);
        if (wcur + cnt > wlim) {
            // close the chunk (its unused slots become REC_NONE) and take the next one
            for (int i = wcur + lane; i < wlim; i += 64) a.rec[i] = make_uint2(REC_NONE, 0u);
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(a.rec_cursor, (unsigned long long)REC_CHUNK);
            base = __shfl(base, 0);
            if ((long long)base + REC_CHUNK > a.rec_cap) {
                // buffer full: its slots from base on become REC_NONE, and this wave adds inline
                for (long long i = (long long)base + lane; i < a.rec_cap; i += 64) a.rec[i] = make_uint2(REC_NONE, 0u);
                defer = false;
                wcur = wlim = 0;
                if (ov >= 0) {
                    unsafeAtomicAdd(a.Lt + (int64_t)ov * a.Ls, oL);
                    ++t_at;
                    ov = -1;
                }
                return;
            }
            wcur = __builtin_amdgcn_readfirstlane((int)base);
            wlim = wcur + REC_CHUNK;
        }
        if (ov >= 0) {
            const int rank = __popcll(m & ((1ull << lane) - 1ull));
            a.rec[wcur + rank] = make_uint2((unsigned)ov, __float_as_uint(oL));
            ov = -1;
            ++t_rec;
        }
        wcur = __builtin_amdgcn_readfirstlane(wcur + cnt);
    };

    // Photons are handed out by the counter, so the lanes of blocks that are resident early
    // take the work of blocks not yet dispatched (a launch of several runs, or concurrent
    // launches, holds more blocks than the device).  The counter only grows: once a refill
    // sees it past the last photon the wave stops asking, and its lanes end as their photons do.
    bool drained = false;
    for (;;) {
        // ---- refill idle lanes
        bool need = !alive && !exhausted;
        if (need && drained) {
            exhausted = true;
            need = false;
        }
        const unsigned long long mask = __ballot(need);
        if (mask) {
            const int cnt = __popcll(mask);
            const int leader = __ffsll((long long)mask) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(a.counter, (unsigned long long)cnt);
            base = __shfl(base, leader);
            const int rank = __popcll(mask & ((1ull << lane) - 1ull));
            const unsigned long long my = base + (unsigned long long)rank;
            if (base + (unsigned long long)cnt >= (unsigned long long)a.n_photons) drained = true;
            if (need) {
                if (my >= (unsigned long long)a.n_photons) {
                    exhausted = true;
                } else {
                    // ---- launch (voxel_mc_3d.py:150-194)
                    const unsigned long long pid = (unsigned long long)a.photon_offset + my;
                    rng_start(rng, key0, key1, pid);
                    rng_top(rng);  // four draws: at most three are taken below outside the sampler loop
                    float u1, u2;
                    if (a.profile == 1) {
                        const float rr = a.R * sqrtf(rng_u(rng));
                        float s, c;
                        __sincosf(6.283185307179586f * rng_u(rng), &s, &c);
                        u1 = rr * c;
                        u2 = rr * s;
                    } else if (a.profile == 2 || a.profile == 3) {
                        for (;;) {
                            rng_top(rng);  // a trial takes at most three draws, leaving one for tau
                            const float rn = sqrtf(rng_u(rng));
                            const float accp = a.profile == 2 ? __expf(-2.f * rn) : __expf(-2.f * rn * rn);
                            if (rng_u(rng) < accp) {
                                float s, c;
                                __sincosf(6.283185307179586f * rng_u(rng), &s, &c);
                                u1 = a.R * rn * c;
                                u2 = a.R * rn * s;
                                break;
                            }
                        }
                    } else {
                        const float ua = rng_u(rng), ub = rng_u(rng);
                        const float rad = sqrtf(-2.f * __logf(ua));
                        float s, c;
                        __sincosf(6.283185307179586f * ub, &s, &c);
                        u1 = a.sigma * rad * c;
                        u2 = a.sigma * rad * s;
                    }
                    bool out;
                    if (a.la == 0) {
                        p.x = a.xc + (u1 / a.cts) * a.cos_ph - u2 * a.sin_ph;
                        p.y = a.yc + (u1 / a.cts) * a.sin_ph + u2 * a.cos_ph;
                        p.z = 0.f;
                        const float ddx = p.x - a.xc, ddy = p.y - a.yc;
                        out = p.x < 0.f || p.x >= a.Lx || p.y < 0.f || p.y >= a.Ly ||
                              (a.rmax2 > 0.f && ddx * ddx + ddy * ddy >= a.rmax2);
                        p.ix = min((int)(p.x / a.dx), Nx - 1);
                        p.iy = min((int)(p.y / a.dy), Ny - 1);
                        p.iz = 0;
                    } else {
                        p.x = 0.f;
                        p.y = a.yc + (u1 / a.cts) * a.cos_ph - u2 * a.sin_ph;
                        p.z = a.zc + (u1 / a.cts) * a.sin_ph + u2 * a.cos_ph;
                        const float ddy = p.y - a.yc, ddz = p.z - a.zc;
                        out = p.y < 0.f || p.y >= a.Ly || p.z < 0.f || p.z >= a.Lz ||
                              (a.rmax2 > 0.f && ddy * ddy + ddz * ddz >= a.rmax2);
                        p.ix = 0;
                        p.iy = min((int)(p.y / a.dy), Ny - 1);
                        p.iz = min((int)(p.z / a.dz), Nz - 1);
                    }
                    if (out) {
                        t_lat += 1.0;
                        unsafeAtomicAdd(&a.lat[0], 1.0);
                    } else {
                        p.ux = a.iu0;
                        p.uy = a.iu1;
                        p.uz = a.iu2;
                        p.o = a.opt[((int64_t)p.ix * Ny + p.iy) * Nz + p.iz];
                        const double r_entry = fresnel_d(a.n_amb, (double)p.o.w, a.cos_in_d);
                        t_spec += r_entry;
                        p.w = (float)(1.0 - r_entry);
                        if (p.w > 0.f) {
                            if (fabs((double)p.o.w - a.n_amb) > 1e-9) {
                                const float eta = (float)(a.n_amb / (double)p.o.w);
                                if (depth_axis == 2) refract(p.uz, p.ux, p.uy, eta);
                                else refract(p.ux, p.uy, p.uz, eta);
                            }
                            p.tau = -__logf(fmaxf(rng_u(rng), 1e-12f));
                            p.ev = 0;
                            alive = true;
                        }
                    }
                }
            }
        }
        if (!__any(alive)) {
            if (__all(exhausted)) break;
            continue;
        }
        // ---- transport a burst of events, then refill
        // (a lane whose packet ended idles to the end of the burst, so that every lane is
        // active at rec_append)
        for (int it = 0; it < 16; ++it) {
            rec_append();
            if (!__any(alive)) break;
            if (!alive) continue;
            rng_top(rng);  // an event takes at most four draws (roulette, polar, azimuth, tau)
            const int v = (p.ix * Ny + p.iy) * Nz + p.iz;
            const float mua = p.o.x;
            const float mut = p.o.x + p.o.y;
            // _dist_to_voxel_exit (voxel_mc_3d.py:72-98), with the hardware reciprocal
            // of the direction (1 ulp; the reference runs this under numba fastmath)
            const float sx = p.ux > 1e-15f ? ((p.ix + 1) * a.dx - p.x) * __builtin_amdgcn_rcpf(p.ux)
                           : (p.ux < -1e-15f ? (p.ix * a.dx - p.x) * __builtin_amdgcn_rcpf(p.ux) : 1e30f);
            const float sy = p.uy > 1e-15f ? ((p.iy + 1) * a.dy - p.y) * __builtin_amdgcn_rcpf(p.uy)
                           : (p.uy < -1e-15f ? (p.iy * a.dy - p.y) * __builtin_amdgcn_rcpf(p.uy) : 1e30f);
            const float sz = p.uz > 1e-15f ? ((p.iz + 1) * a.dz - p.z) * __builtin_amdgcn_rcpf(p.uz)
                           : (p.uz < -1e-15f ? (p.iz * a.dz - p.z) * __builtin_amdgcn_rcpf(p.uz) : 1e30f);
            int axis;
            float s_exit;
            if (sx <= sy && sx <= sz) { axis = 0; s_exit = sx; }
            else if (sy <= sz) { axis = 1; s_exit = sy; }
            else { axis = 2; s_exit = sz; }
            s_exit = fmaxf(s_exit, 0.f);
            // Interactions in batches.  Most events are chords, yet nearly every iteration of a
            // 64-lane wave has some lane at an interaction, so the wave used to run both branch
            // bodies every iteration.  A lane whose event is an interaction now waits (masked,
            // its photon unchanged) until every INTER_EVERY-th iteration of the burst, or until
            // half of the wave's live lanes wait, so most iterations run the chord body alone.
            // Each photon's sequence of events is the same; only the interleaving changes.
            const bool inter = mut > 1e-12f && mut * s_exit >= p.tau;
            const int n_wait = __popcll(__ballot(inter));
            const int n_live = __popcll(__ballot(true));  // every live lane, before any branch
            if (inter && it % INTER_EVERY != INTER_EVERY - 1 && 2 * n_wait < n_live) continue;
            t_ev += 1;
            p.ev++;
            if (inter) {
                // ---- interaction inside the voxel (voxel_mc_3d.py:201-265)
                const float s_col = p.tau / mut;
                tally(v, 0.f, p.w * s_col);
                p.x += s_col * p.ux;
                p.y += s_col * p.uy;
                p.z += s_col * p.uz;
                if (a.rmax2 > 0.f) {
                    const float rr = a.la == 0 ? (p.x - a.xc) * (p.x - a.xc) + (p.y - a.yc) * (p.y - a.yc)
                                               : (p.y - a.yc) * (p.y - a.yc) + (p.z - a.zc) * (p.z - a.zc);
                    if (rr >= a.rmax2) {
                        die_lateral(a.la == 0 ? p.iz : p.ix, p.w);
                        continue;
                    }
                }
                const float dw = p.w * (mua / mut);
                p.w -= dw;
                tally(v, dw, 0.f);
                t_abs += (double)dw;
                if (p.w < 1e-4f) {
                    if (rng_u(rng) < 0.1f) p.w *= 10.f;
                    else {
                        alive = false;
                        flush();
                        continue;
                    }
                }
                const float gl = p.o.z;
                const float xi = rng_u(rng);
                float ct;
                if (fabsf(gl) < 1e-4f) ct = 2.f * xi - 1.f;
                else {
                    const float tmp = (1.f - gl * gl) / (1.f - gl + 2.f * gl * xi);
                    ct = (1.f + gl * gl - tmp * tmp) / (2.f * gl);
                    ct = fmaxf(-1.f, fminf(1.f, ct));
                }
                const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
                float sp, cp;
                __sincosf(6.283185307179586f * rng_u(rng), &sp, &cp);
                if (fabsf(p.uz) > 0.99999f) {
                    const float sg = p.uz > 0.f ? 1.f : -1.f;
                    p.ux = st * cp;
                    p.uy = st * sp;
                    p.uz = ct * sg;
                } else {
                    const float den = sqrtf(fmaxf(1.f - p.uz * p.uz, 1e-12f));
                    const float iden = __builtin_amdgcn_rcpf(den);
                    const float nx = st * (p.ux * p.uz * cp - p.uy * sp) * iden + p.ux * ct;
                    const float ny = st * (p.uy * p.uz * cp + p.ux * sp) * iden + p.uy * ct;
                    const float nz = -st * cp * den + p.uz * ct;
                    p.ux = nx;
                    p.uy = ny;
                    p.uz = nz;
                }
                const float nrm = sqrtf(p.ux * p.ux + p.uy * p.uy + p.uz * p.uz);
                if (nrm > 1e-12f) {
                    const float inrm = __builtin_amdgcn_rcpf(nrm);
                    p.ux *= inrm;
                    p.uy *= inrm;
                    p.uz *= inrm;
                }
                p.tau = -__logf(fmaxf(rng_u(rng), 1e-12f));
            } else {
                // ---- full chord to the exit face (voxel_mc_3d.py:267-337)
                // Optics of the exit voxel, loaded first so that the load's latency overlaps
                // the rest of the chord.  Only this branch loads: on gfx9 loads and the tally
                // atomics share one in-order counter, so a load left in flight on the
                // interaction branch would make that branch wait for it and for its atomics.
                int qx = p.ix, qy = p.iy, qz = p.iz;
                if (axis == 0) qx += p.ux > 0.f ? 1 : -1;
                else if (axis == 1) qy += p.uy > 0.f ? 1 : -1;
                else qz += p.uz > 0.f ? 1 : -1;
                float4 o_next = make_float4(0.f, 0.f, 0.f, 1.f);
                if (qx >= 0 && qx < Nx && qy >= 0 && qy < Ny && qz >= 0 && qz < Nz)
                    o_next = a.opt[((int64_t)qx * Ny + qy) * Nz + qz];
                p.tau -= mut * s_exit;
                tally(v, 0.f, p.w * s_exit);
                p.x += s_exit * p.ux;
                p.y += s_exit * p.uy;
                p.z += s_exit * p.uz;
                int jx = p.ix, jy = p.iy, jz = p.iz;
                int sdir;
                if (axis == 0) {
                    sdir = p.ux > 0.f ? 1 : -1;
                    jx += sdir;
                    p.x = sdir > 0 ? (p.ix + 1) * a.dx : p.ix * a.dx;
                } else if (axis == 1) {
                    sdir = p.uy > 0.f ? 1 : -1;
                    jy += sdir;
                    p.y = sdir > 0 ? (p.iy + 1) * a.dy : p.iy * a.dy;
                } else {
                    sdir = p.uz > 0.f ? 1 : -1;
                    jz += sdir;
                    p.z = sdir > 0 ? (p.iz + 1) * a.dz : p.iz * a.dz;
                }
                bool ended = false, reflected = false;
                if (axis == depth_axis) {
                    const int jd = depth_axis == 2 ? jz : jx;
                    const int nd = depth_axis == 2 ? Nz : Nx;
                    if (jd < 0) {
                        float &ud = depth_axis == 2 ? p.uz : p.ux;
                        const float r_out = fresnel(p.o.w, (float)a.n_amb, ud);
                        if (r_out > 0.f && rng_u(rng) < r_out) {
                            ud = -ud;
                            reflected = true;
                        } else {
                            t_diff += (double)p.w;
                            ended = true;
                        }
                    } else if (jd >= nd) {
                        t_trans += (double)p.w;
                        ended = true;
                    }
                } else if (jx < 0 || jx >= Nx || jy < 0 || jy >= Ny || jz < 0 || jz >= Nz) {
                    die_lateral(a.la == 0 ? p.iz : p.ix, p.w);
                    continue;
                }
                if (ended) {
                    alive = false;
                    flush();
                    continue;
                }
                if (reflected) {
                    if (p.ev >= a.max_events) { alive = false; flush(); }
                    continue;
                }
                if (a.rmax2 > 0.f) {
                    const float rr = a.la == 0 ? (p.x - a.xc) * (p.x - a.xc) + (p.y - a.yc) * (p.y - a.yc)
                                               : (p.y - a.yc) * (p.y - a.yc) + (p.z - a.zc) * (p.z - a.zc);
                    if (rr >= a.rmax2) {
                        die_lateral(a.la == 0 ? p.iz : p.ix, p.w);
                        continue;
                    }
                }
                const float4 o2 = o_next;  // (jx, jy, jz) == (qx, qy, qz), inside the domain here
                const float n1 = p.o.w, n2 = o2.w;
                if (fabsf(n1 - n2) > 1e-9f) {
                    float &un = axis == 0 ? p.ux : (axis == 1 ? p.uy : p.uz);
                    const float r_int = fresnel(n1, n2, un);
                    if (rng_u(rng) < r_int) {
                        un = -un;
                        if (p.ev >= a.max_events) { alive = false; flush(); }
                        continue;
                    }
                    const float eta = n1 / n2;
                    if (axis == 0) refract(p.ux, p.uy, p.uz, eta);
                    else if (axis == 1) refract(p.uy, p.ux, p.uz, eta);
                    else refract(p.uz, p.ux, p.uy, eta);
                }
                p.ix = jx;
                p.iy = jy;
                p.iz = jz;
                p.o = o2;
            }
            if (alive && p.ev >= a.max_events) {
                alive = false;  // voxel_mc_3d.py:196: the event budget ends the packet untallied
                flush();
            }
        }
        rec_append();
    }
    flush();
    rec_append();
    if (defer)
        for (int i = wcur + lane; i < wlim; i += 64) a.rec[i] = make_uint2(REC_NONE, 0u);
    // ---- wave reduction of the FP64 scalar tallies, one atomic per wave
    double vals[8] = {t_spec, t_diff, t_trans, t_lat, t_abs, (double)t_ev, (double)t_at, (double)t_rec};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double v = vals[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (lane == 0 && v != 0.0) unsafeAtomicAdd(&a.acc[q], v);
    }
}

// Three waves per SIMD (at most 168 VGPRs): at the 150k photons of an epoch on 256 CUs that
// keeps every photon's lane resident from the start (2.3 waves per SIMD); the register
// allocator then spills two VGPRs, which costs less than the third wave gains (-4 % kernel time).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_mc(McArgs a) { mc_trace(a); }

// Several contexts' transport in one launch: block b traces photons of context
// block_run[b] (blocks are dealt to contexts in proportion to their photon
// counts, each block to one context), so a context's arguments are uniform
// over a block and sit in scalar registers as in k_mc.  Blocks of all runs
// share the machine: the tails of long-photon runs fill with other runs' work.
__global__ __launch_bounds__(256) void k_mc_gang(const McArgs *__restrict__ args, const int *__restrict__ block_run)
{
    const McArgs a = args[block_run[blockIdx.x]];
    mc_trace(a);
}

// ------------------------------------------------------------------ optics pack
// Transport optics per voxel: base optics plus the plume extinction sampled at
// the voxel centre (coupled_ablation.py:445-450: mu_a += mu_a_plume,
// mu_s += mu_s_plume, g = 0.8 where plume > 1e-3 inside the void).
__global__ void k_pack(const float *__restrict__ mu_a, const float *__restrict__ mu_s, const float *__restrict__ g,
                       const float *__restrict__ n, const float *__restrict__ ext, const uint8_t *__restrict__ flags,
                       float4 *__restrict__ out, int Nx, int Ny, int Nz, int Px, int Py, int Pz, float ra, float rs,
                       int add)
{
    const int64_t N = (int64_t)Nx * Ny * Nz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        float4 o = make_float4(mu_a[i], mu_s[i], g[i], n[i]);
        if (add) {
            const int z = (int)(i % Nz);
            const int64_t r = i / Nz;
            const int y = (int)(r % Ny), x = (int)(r / Ny);
            const int I = (int)(((int64_t)(2 * x + 1) * Px) / (2 * (int64_t)Nx));
            const int J = (int)(((int64_t)(2 * y + 1) * Py) / (2 * (int64_t)Ny));
            const int K = (int)(((int64_t)(2 * z + 1) * Pz) / (2 * (int64_t)Nz));
            const float e = ext[((int64_t)I * Py + J) * Pz + K];
            o.x += e * ra;
            o.y += e * rs;
            if (e > 0.001f && (flags[i] & F_ABL)) o.z = 0.80f;
        }
        out[i] = o;
    }
}

int launch_pack_optics(Ctx *c, bool add_plume)
{
    BD_TRY(ensure(c, (void **)&c->optics, (size_t)c->nvox * sizeof(float4)));
    const bool add = add_plume && c->plume.enabled && c->Pext && c->flags;
    const double mu_max = c->plume.mu_max;
    const float ra = add ? (float)(c->plume.mu_a_max / std::fmax(1e-6, mu_max)) : 0.f;
    const float rs = add ? (float)(c->plume.mu_s_max / std::fmax(1e-6, mu_max)) : 0.f;
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_pack, dim3(nb), dim3(256), 0, c->stream, c->mu_a, c->mu_s, c->gg, c->nn, c->Pext, c->flags,
                       c->optics, c->g.nx, c->g.ny, c->g.nz, add ? c->plume.nx : 1, add ? c->plume.ny : 1,
                       add ? c->plume.nz : 1, ra, rs, add ? 1 : 0);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// Q = A * power / (N * V)  (voxel_mc_3d.py:684)
// ---- the deferred path-length records of a transport launch, added into the L tally: a
// counting sort of the records by 1024-voxel tile (a tile is a range of the linear voxel index),
// then chunks of at most REC_CHUNK_RECS records of one tile are summed in LDS and the sums added
// to L, with plain stores where a tile is one chunk and with atomics (at most one per voxel and
// chunk, far fewer than records) where several chunks share a tile.  The slots below
// min(cursor, capacity) hold records or REC_NONE (the transport closes its chunks).
constexpr int REC_TILE_BITS = 10;
constexpr int REC_TILE = 1 << REC_TILE_BITS;   // voxels per tile: 4 KB of float sums in LDS
constexpr int REC_SLICES = 512;                // record slices of the count and scatter passes
constexpr int REC_MAX_TILES = 16384;           // tile counters in LDS (64 KB): grids up to 2^24 voxels
constexpr int REC_CHUNK_RECS = 32768;          // records per accumulation chunk

__device__ __forceinline__ void rec_slice(const unsigned long long *cursor, long long cap, long long &lo,
                                          long long &hi)
{
    const long long n = (long long)*cursor < cap ? (long long)*cursor : cap;
    const long long per = (n + REC_SLICES - 1) / REC_SLICES;
    lo = (long long)blockIdx.x * per;
    hi = lo + per < n ? lo + per : n;
}

// Pass 1 (REC_SLICES blocks): the slice's records per tile; one returning atomic per touched
// tile reserves the slice's range inside the tile (off) and counts the tile (tcnt).
__global__ __launch_bounds__(256) void k_rec_count(const uint2 *__restrict__ rec,
                                                   const unsigned long long *__restrict__ cursor, long long cap,
                                                   unsigned nvox, int ntiles, unsigned *__restrict__ tcnt,
                                                   unsigned *__restrict__ off)
{
    extern __shared__ unsigned h[];
    for (int t = threadIdx.x; t < ntiles; t += 256) h[t] = 0u;
    __syncthreads();
    long long lo, hi;
    rec_slice(cursor, cap, lo, hi);
    for (long long i = lo + threadIdx.x; i < hi; i += 256) {
        const unsigned v = rec[i].x;
        if (v < nvox) atomicAdd(&h[v >> REC_TILE_BITS], 1u);
    }
    __syncthreads();
    unsigned *o = off + (size_t)blockIdx.x * ntiles;
    for (int t = threadIdx.x; t < ntiles; t += 256)
        if (h[t]) o[t] = atomicAdd(&tcnt[t], h[t]);
}

// Pass 2 (one block of 1024): exclusive scans of the tile counts (tstart, ntiles + 1 entries)
// and of the tiles' chunk counts (cstart), and the chunk -> tile map (cmap); the counts are
// left zeroed for the next launch.
__global__ __launch_bounds__(1024) void k_rec_scan(unsigned *__restrict__ tcnt, int ntiles,
                                                   unsigned *__restrict__ tstart, unsigned *__restrict__ cstart,
                                                   unsigned *__restrict__ cmap)
{
    __shared__ unsigned part[1024], cpart[1024];
    const int per = (ntiles + 1023) / 1024;
    const int lo = threadIdx.x * per, hi = lo + per < ntiles ? lo + per : ntiles;
    unsigned s = 0, cs = 0;
    for (int t = lo; t < hi; ++t) {
        s += tcnt[t];
        cs += (tcnt[t] + REC_CHUNK_RECS - 1) / REC_CHUNK_RECS;
    }
    part[threadIdx.x] = s;
    cpart[threadIdx.x] = cs;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scans of the thread sums
        const unsigned x = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        const unsigned y = threadIdx.x >= d ? cpart[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += x;
        cpart[threadIdx.x] += y;
        __syncthreads();
    }
    unsigned run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    unsigned crun = threadIdx.x ? cpart[threadIdx.x - 1] : 0u;
    for (int t = lo; t < hi; ++t) {
        const unsigned n = tcnt[t];
        tcnt[t] = 0u;  // cleared for the next launch (this pass is the counts' last reader)
        tstart[t] = run;
        cstart[t] = crun;
        const unsigned nch = (n + REC_CHUNK_RECS - 1) / REC_CHUNK_RECS;
        for (unsigned k = 0; k < nch; ++k) cmap[crun + k] = (unsigned)t;
        run += n;
        crun += nch;
    }
    if (threadIdx.x == 1023) {
        tstart[ntiles] = part[1023];
        cstart[ntiles] = cpart[1023];
    }
}

// Pass 3 (REC_SLICES blocks, the slices of pass 1): each record to its place in tile order.
__global__ __launch_bounds__(256) void k_rec_scatter(const uint2 *__restrict__ rec,
                                                     const unsigned long long *__restrict__ cursor, long long cap,
                                                     unsigned nvox, int ntiles, const unsigned *__restrict__ off,
                                                     const unsigned *__restrict__ tstart, uint2 *__restrict__ sorted)
{
    extern __shared__ unsigned cur[];
    const unsigned *o = off + (size_t)blockIdx.x * ntiles;
    long long lo, hi;
    rec_slice(cursor, cap, lo, hi);
    // only tiles this slice touched have an off entry; the others are never indexed below
    for (int t = threadIdx.x; t < ntiles; t += 256) cur[t] = tstart[t] + o[t];
    __syncthreads();
    for (long long i = lo + threadIdx.x; i < hi; i += 256) {
        const uint2 r = rec[i];
        if (r.x < nvox) sorted[atomicAdd(&cur[r.x >> REC_TILE_BITS], 1u)] = r;
    }
}

// Pass 4 (a fixed grid over the chunks): a chunk's records summed in LDS, the sums added to the
// path-length grid Lt[voxel * Ls].
__global__ __launch_bounds__(512) void k_rec_accum(const uint2 *__restrict__ sorted,
                                                   const unsigned *__restrict__ tstart,
                                                   const unsigned *__restrict__ cstart,
                                                   const unsigned *__restrict__ cmap, int ntiles, unsigned nvox,
                                                   float *__restrict__ Lt, int Ls)
{
    __shared__ float acc[REC_TILE];
    const unsigned nchunks = cstart[ntiles];
    for (unsigned j = blockIdx.x; j < nchunks; j += gridDim.x) {
        const unsigned t = cmap[j];
        const unsigned k = j - cstart[t];
        const unsigned s = tstart[t] + k * REC_CHUNK_RECS;
        const unsigned e = min(tstart[t + 1], s + REC_CHUNK_RECS);
        const bool shared_tile = cstart[t + 1] - cstart[t] > 1;
        for (int q = threadIdx.x; q < REC_TILE; q += 512) acc[q] = 0.f;
        __syncthreads();
        for (unsigned i = s + threadIdx.x; i < e; i += 512) {
            const uint2 r = sorted[i];
            atomicAdd(&acc[r.x & (REC_TILE - 1)], __uint_as_float(r.y));
        }
        __syncthreads();
        const unsigned base = t << REC_TILE_BITS;
        for (int q = threadIdx.x; q < REC_TILE; q += 512) {
            const float x = acc[q];
            if (x != 0.f && base + q < nvox) {
                float *dst = Lt + (size_t)(base + q) * Ls;
                if (shared_tile) unsafeAtomicAdd(dst, x);
                else *dst += x;
            }
        }
        __syncthreads();  // acc is cleared for the next chunk only after every thread has read it
    }
}

__global__ void k_q_from_a(const float2 *__restrict__ AL, float *__restrict__ Q, float scale, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        Q[i] = AL[i].x * scale;
}

// The beam and grid arguments of one transport launch (voxel_mc_3d.py:599-625): beam frame,
// incidence vector, launch footprint, Philox key; the buffers are set by the caller.
static void fill_mc_args(const bd_grid &g, const bd_beam *b, McArgs &a)
{
    a = McArgs{};
    a.Nx = g.nx;
    a.Ny = g.ny;
    a.Nz = g.nz;
    a.dx = (float)g.dx;
    a.dy = (float)g.dy;
    a.dz = (float)g.dz;
    a.Lx = (float)(g.dx * g.nx);
    a.Ly = (float)(g.dy * g.ny);
    a.Lz = (float)(g.dz * g.nz);
    const double w0 = b->spot_dia / 2.0;
    a.sigma = (float)(w0 / 2.0);
    a.R = (float)w0;
    a.profile = b->beam_profile;
    a.la = b->launch_axis;
    const double sth = std::sin(b->theta_rad), cth = std::cos(b->theta_rad);
    const double sph = std::sin(b->phi_rad), cph = std::cos(b->phi_rad);
    double iu[3];
    if (b->launch_axis == 0) {
        iu[0] = sth * cph; iu[1] = sth * sph; iu[2] = cth;
    } else {
        iu[0] = cth; iu[1] = sth * cph; iu[2] = sth * sph;
    }
    a.cos_in_d = cth;
    a.iu0 = (float)iu[0];
    a.iu1 = (float)iu[1];
    a.iu2 = (float)iu[2];
    a.cos_ph = (float)cph;
    a.sin_ph = (float)sph;
    a.cts = (float)std::fmax(cth, 0.05);
    a.xc = (float)b->x_center;
    a.yc = (float)b->y_center;
    a.zc = (float)b->z_center;
    a.n_amb = b->n_ambient;
    a.rmax2 = b->r_max > 0.0 ? (float)(b->r_max * b->r_max) : 0.f;
    a.n_photons = b->n_photons;
    a.photon_offset = b->photon_offset;
    a.max_events = b->max_events;
    a.k0 = (unsigned)(b->seed & 0xffffffffu);
    a.k1 = (unsigned)(b->seed >> 32);
    a.score_L = b->score_fluence ? 1 : 0;
}

int join_records(Ctx *c)
{
    if (c->join_pending) {
        BD_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
        c->join_pending = false;
    }
    return BD_OK;
}

// Pack the optics, clear the tallies and fill the kernel arguments of one
// context's transport (no launch).
static int mc_prepare(Ctx *c, const bd_beam *b, const EpochParams *ep, McArgs &a)
{
    BD_TRY(join_records(c));  // the tallies and record buffers are about to be reused
    lag_drop(c);
    BD_TRY(launch_pack_optics(c, b->add_plume != 0));
    BD_TRY(ensure(c, (void **)&c->AL, (size_t)c->nvox * sizeof(float2)));
    BD_TRY(ensure(c, (void **)&c->Q, (size_t)c->nvox * sizeof(float)));
    const int64_t nd = b->launch_axis == 0 ? c->g.nz : c->g.nx;
    if (c->lat_n != nd) {
        if (c->lat) BD_HIP(hipFree(c->lat));
        c->lat = nullptr;
        BD_HIP(hipMalloc((void **)&c->lat, sizeof(double) * nd));
        c->lat_n = nd;
    }
    BD_TRY(launch_fill(c, c->AL, K_F32, c->nvox * 2, 0.0));
    BD_TRY(launch_zero_f64(c, {{c->lat, nd}, {c->red + 16, 8}, {c->counter, 1}}));

    fill_mc_args(c->g, b, a);
    a.opt = c->optics;
    a.AL = c->AL;
    a.Lt = (float *)c->AL + 1;
    a.Ls = 2;
    a.acc = c->red + 16;
    a.lat = c->lat;
    a.counter = c->counter;
    a.ep = ep;
    return BD_OK;
}

// Blocks of k_mc that are resident at once on one CU (its register use sets it), at most 8.
static int mc_resident_per_cu()
{
    static int n = 0;
    if (n == 0) {
        int r = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, (const void *)k_mc, 256, 0) != hipSuccess || r < 1) r = 8;
        n = r < 8 ? r : 8;
    }
    return n;
}

// Blocks for n photons: one photon per lane, at most as many blocks as are resident at once
// (idle lanes refill from the photon counter, so photons beyond the resident lanes go to the
// lanes whose photons end first instead of to blocks that wait for a free slot).
static int64_t mc_blocks(Ctx *c, long long n_photons)
{
    int ncu = 256;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    const int64_t want = (n_photons + 255) / 256;
    int64_t cap = (int64_t)ncu * mc_resident_per_cu();
    if (c->mc_blocks_per_cu > 0) cap = (int64_t)ncu * c->mc_blocks_per_cu;
    return want < cap ? (want < 1 ? 1 : want) : cap;
}

// Absorbed power from the absorbed-weight tally (after the transport).
static int mc_post(Ctx *c, const bd_beam *b, double *q_sum = nullptr)
{
    const float scale = (float)(b->power / ((double)b->n_photons * c->g.dx * c->g.dy * c->g.dz));
    if (q_sum) return q_from_a_sum_enqueue(c, scale, c->flags, F_ABL, true, q_sum);
    const int nb = blocks_for(c->nvox, 256) > 8192 ? 8192 : blocks_for(c->nvox, 256);
    hipLaunchKernelGGL(k_q_from_a, dim3(nb), dim3(256), 0, c->stream, c->AL, c->Q, scale, c->nvox);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// Record buffer of the deferred path-length tally: rec_per_photon records per photon (512 by
// default; a photon of the layered-stack workloads leaves 150-400), between 2^16 and 2^28
// records (0.5 MB - 2 GB).  A launch that fills it adds the rest inline, so the size bounds only
// how much is deferred.
// Whether a launch of mode rec_mode writes path-length records.  By default only lagged launches
// do: their passes run beside the next transport.  After a launch of its own or inside a captured
// epoch the passes cost more than the inline atomics they replace wherever photons are mostly
// scattering events (the campaign's 320-532 nm runs: 1 900 events per photon, the kernel time the
// same with and without records, the passes 1.2-1.7 ms more per epoch).
static bool defers(const Ctx *c, const bd_beam *b, int rec_mode)
{
    const int64_t ntiles = (c->nvox + REC_TILE - 1) >> REC_TILE_BITS;
    if (!b->score_fluence || ntiles > REC_MAX_TILES || c->defer_L == 0) return false;
    return c->defer_L == 1 || rec_mode == REC_LAG;
}

int ensure_records(Ctx *c, const bd_beam *b, int rec_mode)
{
    const int64_t ntiles = (c->nvox + REC_TILE - 1) >> REC_TILE_BITS;
    if (!defers(c, b, rec_mode)) return BD_OK;
    int64_t cap = (int64_t)b->n_photons * c->rec_per_photon;
    cap = std::min<int64_t>(std::max<int64_t>(cap, int64_t(1) << 16), int64_t(1) << 28);
    if (!c->rec_tcnt) {
        BD_HIP(hipMalloc((void **)&c->rec_tcnt, sizeof(unsigned) * (size_t)ntiles));
        BD_HIP(hipMemset(c->rec_tcnt, 0, sizeof(unsigned) * (size_t)ntiles));  // k_rec_scan re-zeroes them
        BD_HIP(hipMalloc((void **)&c->rec_tstart, sizeof(unsigned) * (size_t)(ntiles + 1)));
        BD_HIP(hipMalloc((void **)&c->rec_cstart, sizeof(unsigned) * (size_t)(ntiles + 1)));
        BD_HIP(hipMalloc((void **)&c->rec_off, sizeof(unsigned) * (size_t)REC_SLICES * ntiles));
    }
    if (c->rec_cap >= cap) return BD_OK;
    // larger buffers: nothing may still use the old ones, and lagged records in them are dropped
    if (c->side) BD_HIP(hipStreamSynchronize(c->side));
    BD_HIP(hipStreamSynchronize(c->stream));
    lag_drop(c);
    if (c->rec) BD_HIP(hipFree(c->rec));
    if (c->rec_b) BD_HIP(hipFree(c->rec_b));
    if (c->rec_sorted) BD_HIP(hipFree(c->rec_sorted));
    if (c->rec_cmap) BD_HIP(hipFree(c->rec_cmap));
    c->rec = c->rec_b = c->rec_sorted = nullptr;
    c->rec_cmap = nullptr;
    c->rec_cap = 0;
    BD_HIP(hipMalloc((void **)&c->rec, sizeof(uint2) * (size_t)cap));
    BD_HIP(hipMalloc((void **)&c->rec_sorted, sizeof(uint2) * (size_t)cap));
    // chunks: at most one partial chunk per tile beyond cap / REC_CHUNK_RECS full ones
    BD_HIP(hipMalloc((void **)&c->rec_cmap, sizeof(unsigned) * (size_t)(cap / REC_CHUNK_RECS + ntiles + 1)));
    c->rec_cap = cap;
    return BD_OK;
}

// The record passes over record buffer rec (cursor: its slot count) on stream st, adding into
// the path-length grid Lt[voxel * Ls].
static int launch_record_passes(Ctx *c, hipStream_t st, const uint2 *rec, const unsigned long long *cursor,
                                float *Lt, int Ls)
{
    const int ntiles = (int)((c->nvox + REC_TILE - 1) >> REC_TILE_BITS);
    const size_t lds = sizeof(unsigned) * (size_t)ntiles;
    const unsigned nv = (unsigned)c->nvox;
    hipLaunchKernelGGL(k_rec_count, dim3(REC_SLICES), dim3(256), lds, st, rec, cursor,
                       (long long)c->rec_cap, nv, ntiles, c->rec_tcnt, c->rec_off);
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rec_scan, dim3(1), dim3(1024), 0, st, c->rec_tcnt, ntiles, c->rec_tstart,
                       c->rec_cstart, c->rec_cmap);
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rec_scatter, dim3(REC_SLICES), dim3(256), lds, st, rec, cursor,
                       (long long)c->rec_cap, nv, ntiles, c->rec_off, c->rec_tstart, c->rec_sorted);
    BD_HIP(hipGetLastError());
    int ncu = 256;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    hipLaunchKernelGGL(k_rec_accum, dim3(ncu * 4), dim3(512), 0, st, c->rec_sorted, c->rec_tstart,
                       c->rec_cstart, c->rec_cmap, ntiles, nv, Lt, Ls);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// The second record buffer and the two path-length grids of lagged launches (eager epochs only,
// so never inside a graph capture).
static int ensure_lag_buffers(Ctx *c)
{
    if (c->rec_b && c->Lb[0] && c->side) return BD_OK;
    if (!c->side) BD_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    if (!c->rec_b) BD_HIP(hipMalloc((void **)&c->rec_b, sizeof(uint2) * (size_t)c->rec_cap));
    for (int k = 0; k < 2; ++k)
        if (!c->Lb[k]) BD_HIP(hipMalloc((void **)&c->Lb[k], sizeof(float) * (size_t)c->nvox));
    return BD_OK;
}

int lag_finalize(Ctx *c)
{
    BD_TRY(join_records(c));
    if (c->lag < 0) return BD_OK;
    const int p = c->lag;
    c->lag = -1;
    BD_TRY(launch_record_passes(c, c->stream, p ? c->rec_b : c->rec, c->counter + 3 + p, c->Lb[p], 1));
    al_set(c, c->Lb[p], 1);
    return BD_OK;
}

// The transport launch.  With deferred path lengths the record passes run as rec_mode says:
// REC_MAIN after the kernel on the context stream; REC_LAG (eager epochs) not in this epoch:
// this launch writes the record buffer the previous lagged launch did not, and that launch's
// passes, if still pending, run on the side stream beside this transport.
int launch_mc_enqueue(Ctx *c, const bd_beam *b, const EpochParams *ep, int timed, int rec_mode, double *q_sum)
{
    McArgs a;
    BD_TRY(ensure_records(c, b, rec_mode));  // may drop lagged records (when the buffers grow)
    const int prev = rec_mode == REC_LAG ? c->lag : -1;
    lag_drop(c);  // this transport supersedes the L of a lagged launch (its passes still run below)
    BD_TRY(mc_prepare(c, b, ep, a));
    const bool defer = defers(c, b, rec_mode) && c->rec != nullptr;
    const bool lag = defer && rec_mode == REC_LAG;
    int cur = 0;
    if (lag) {
        BD_TRY(ensure_lag_buffers(c));
        cur = prev == 0 ? 1 : 0;
    }
    if (defer) {
        BD_TRY(launch_fill(c, c->counter + 3 + cur, K_F64, 1, 0.0));
        a.rec = cur ? c->rec_b : c->rec;
        a.rec_cursor = c->counter + 3 + cur;
        a.rec_cap = c->rec_cap;
    }
    if (lag) {
        BD_TRY(launch_fill(c, c->Lb[cur], K_F32, c->nvox, 0.0));
        a.Lt = c->Lb[cur];
        a.Ls = 1;
    }
    if (prev >= 0) {
        // the previous epoch's passes beside this transport (their L is superseded by this
        // launch's; they run so that every epoch's tally is reduced)
        BD_HIP(hipEventRecord(c->ev_fork, c->stream));
        BD_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
        if (timed == 1) BD_HIP(hipEventRecord(c->ev7, c->side));
        BD_TRY(launch_record_passes(c, c->side, prev ? c->rec_b : c->rec, c->counter + 3 + prev, c->Lb[prev], 1));
        if (timed == 1) BD_HIP(hipEventRecord(c->ev6, c->side));
        BD_HIP(hipEventRecord(c->ev_join, c->side));
        c->join_pending = true;
    }
    const int nblk = (int)mc_blocks(c, b->n_photons);
    if (timed == 1) BD_HIP(hipEventRecord(c->ev0, c->stream));
    if (timed == 2) BD_HIP(hipEventRecordWithFlags(c->gev[0], c->stream, hipEventRecordExternal));
    hipLaunchKernelGGL(k_mc, dim3(nblk), dim3(256), 0, c->stream, a);
    BD_HIP(hipGetLastError());
    if (timed == 1) BD_HIP(hipEventRecord(c->ev1, c->stream));
    if (timed == 2) BD_HIP(hipEventRecordWithFlags(c->gev[1], c->stream, hipEventRecordExternal));
    if (lag) {
        c->lag = cur;
        if (prev < 0 && timed == 1) {  // no passes this epoch: a zero-length span
            BD_HIP(hipEventRecord(c->ev7, c->stream));
            BD_HIP(hipEventRecord(c->ev6, c->stream));
        }
        return mc_post(c, b, q_sum);
    }
    if (timed == 1) BD_HIP(hipEventRecord(c->ev7, c->stream));
    if (defer) BD_TRY(launch_record_passes(c, c->stream, c->rec, c->counter + 3, (float *)c->AL + 1, 2));
    if (timed == 1) BD_HIP(hipEventRecord(c->ev6, c->stream));
    if (timed == 2) BD_HIP(hipEventRecordWithFlags(c->gev[6], c->stream, hipEventRecordExternal));
    return mc_post(c, b, q_sum);
}

// ---- gang transport (k_mc_gang): one launch for several contexts' photons.
// Library-wide resources per device: a stream of default priority, an event,
// and the argument / block-map buffers (device and pinned host staging),
// grown on demand.  Gang calls on one device must not overlap in time (they
// share these buffers); each call waits for the previous one's uploads.
namespace {
struct GangRes {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    McArgs *d_args = nullptr, *h_args = nullptr;
    int *d_map = nullptr, *h_map = nullptr;
    int64_t cap_args = 0, cap_map = 0;
};
GangRes g_gang[16];

int gang_res(int dev, int64_t n_args, int64_t n_map, GangRes *&out)
{
    if (dev < 0 || dev >= 16) {
        set_error("gang transport: device index out of range");
        return BD_EINVAL;
    }
    GangRes &g = g_gang[dev];
    if (!g.stream) {
        BD_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        BD_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    }
    BD_HIP(hipStreamSynchronize(g.stream));  // the previous call's uploads read the staging
    if (n_args > g.cap_args) {
        if (g.d_args) BD_HIP(hipFree(g.d_args));
        if (g.h_args) BD_HIP(hipHostFree(g.h_args));
        g.d_args = nullptr;
        g.h_args = nullptr;
        BD_HIP(hipMalloc((void **)&g.d_args, sizeof(McArgs) * n_args));
        BD_HIP(hipHostMalloc((void **)&g.h_args, sizeof(McArgs) * n_args));
        g.cap_args = n_args;
    }
    if (n_map > g.cap_map) {
        if (g.d_map) BD_HIP(hipFree(g.d_map));
        if (g.h_map) BD_HIP(hipHostFree(g.h_map));
        g.d_map = nullptr;
        g.h_map = nullptr;
        BD_HIP(hipMalloc((void **)&g.d_map, sizeof(int) * n_map));
        BD_HIP(hipHostMalloc((void **)&g.h_map, sizeof(int) * n_map));
        g.cap_map = n_map;
    }
    out = &g;
    return BD_OK;
}
}  // namespace

int launch_mc_gang(Ctx *const *cs, const bd_beam *const *beams, int n)
{
    if (n < 1) return BD_OK;
    const int dev = cs[0]->device;
    int64_t nmap = 0;
    for (int i = 0; i < n; ++i) {
        if (cs[i]->device != dev) {
            set_error("gang transport: all contexts must be on one device");
            return BD_EINVAL;
        }
        nmap += mc_blocks(cs[i], beams[i]->n_photons);
    }
    if (nmap > 0x7fffffff) {
        set_error("gang transport: too many blocks");
        return BD_EINVAL;
    }
    GangRes *g = nullptr;
    BD_TRY(gang_res(dev, n, nmap, g));
    int64_t m = 0;
    for (int i = 0; i < n; ++i) {
        Ctx *c = cs[i];
        BD_TRY(mc_prepare(c, beams[i], c->ep, g->h_args[i]));
        const int64_t nb = mc_blocks(c, beams[i]->n_photons);
        for (int64_t b = 0; b < nb; ++b) g->h_map[m++] = i;
        BD_HIP(hipEventRecord(c->ev4, c->stream));
        BD_HIP(hipStreamWaitEvent(g->stream, c->ev4, 0));
    }
    BD_HIP(hipMemcpyAsync(g->d_args, g->h_args, sizeof(McArgs) * n, hipMemcpyHostToDevice, g->stream));
    BD_HIP(hipMemcpyAsync(g->d_map, g->h_map, sizeof(int) * m, hipMemcpyHostToDevice, g->stream));
    for (int i = 0; i < n; ++i) BD_HIP(hipEventRecord(cs[i]->gev[0], g->stream));
    hipLaunchKernelGGL(k_mc_gang, dim3((unsigned)m), dim3(256), 0, g->stream, g->d_args, g->d_map);
    BD_HIP(hipGetLastError());
    for (int i = 0; i < n; ++i) {
        BD_HIP(hipEventRecord(cs[i]->gev[1], g->stream));
        BD_HIP(hipEventRecord(cs[i]->gev[6], g->stream));  // no record pass: the gang transport adds inline
    }
    BD_HIP(hipEventRecord(g->done, g->stream));
    for (int i = 0; i < n; ++i) {
        BD_HIP(hipStreamWaitEvent(cs[i]->stream, g->done, 0));
        BD_TRY(mc_post(cs[i], beams[i]));
    }
    return BD_OK;
}

int launch_mc(Ctx *c, const bd_beam *b, bd_mc_result *out)
{
    BD_TRY(launch_mc_enqueue(c, b, nullptr, 1, REC_MAIN));
    BD_HIP(hipMemcpyAsync(c->host_pinned, c->red + 16, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    const double *r = (const double *)c->host_pinned;
    out->specular = r[0];
    out->diffuse = r[1];
    out->transmitted = r[2];
    out->lateral = r[3];
    out->absorbed = r[4];
    out->events = r[5];
    out->tally_atomics = r[6];
    out->tally_records = r[7];
    float ms = 0.f;
    BD_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    out->kernel_ms = ms;  // the transport kernel alone
    BD_HIP(hipEventElapsedTime(&ms, c->ev7, c->ev6));
    out->record_ms = ms;  // the path-length record passes (0 when the transport adds them inline)
    return BD_OK;
}

}  // namespace bd

// ------------------------------------------------------------------ allocation-free transport
// bd_mc_run_noalloc: the transport of bd_mc_run on caller-owned device buffers.  The launch's
// scratch (photon work counter, FP64 scalar tallies, lateral loss per depth) lives in a
// file-scope device array instead of library allocations, so the call allocates nothing.
namespace bd {
// scratch layout, in doubles: [0] photon work counter (u64), [8, 15) scalar tallies,
// [16, 16 + depth cells) lateral loss per depth
constexpr int NOALLOC_ACC = 8, NOALLOC_LAT = 16;
constexpr int NOALLOC_DOUBLES = NOALLOC_LAT + BD_NOALLOC_MAX_DEPTH;
}  // namespace bd

static __device__ __align__(256) double scratch[bd::NOALLOC_DOUBLES];

namespace bd {
namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
double *g_scratch_dev[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// The device address of `scratch` on device `dev`, looked up once per device with that device
// current.  In host code the array's name is only a stand-in: kernels and copies get this address.
int scratch_address(int dev, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        BD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return BD_OK;
}

// zero the counter, the scalar tallies and the first n_lat depth bins, and the caller's tallies
__global__ void k_noalloc_clear(double *__restrict__ s, int n_lat, float *__restrict__ al, int64_t n_al)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (int64_t i = i0; i < NOALLOC_LAT + n_lat; i += stride) s[i] = 0.0;
    for (int64_t i = i0; i < n_al; i += stride) al[i] = 0.f;
}
}  // namespace
}  // namespace bd

extern "C" int bd_mc_run_noalloc(const bd_grid *grid, const bd_beam *beam, const void *optics, void *AL, void *stream,
                                 bd_mc_result *out, double *lateral, int64_t n_depth)
{
    using namespace bd;
    BD_CHECK(grid && beam && optics && AL && out, "bd_mc_run_noalloc: null argument");
    BD_CHECK(grid->nx > 0 && grid->ny > 0 && grid->nz > 0 && grid->dx > 0.0 && grid->dy > 0.0 && grid->dz > 0.0,
             "bd_mc_run_noalloc: grid dimensions and spacings must be positive");
    const int64_t nvox = (int64_t)grid->nx * grid->ny * grid->nz;
    BD_CHECK(nvox < (int64_t(1) << 31), "bd_mc_run_noalloc: at most 2^31 - 1 voxels");
    BD_CHECK(beam->launch_axis == 0 || beam->launch_axis == 1, "bd_mc_run_noalloc: launch_axis must be 0 or 1");
    BD_CHECK(beam->beam_profile >= 0 && beam->beam_profile <= 3, "bd_mc_run_noalloc: unknown beam profile");
    BD_CHECK(beam->n_photons >= 0 && beam->photon_offset >= 0 && beam->max_events > 0,
             "bd_mc_run_noalloc: photon count, offset and event budget must be non-negative (budget positive)");
    const int64_t nd = beam->launch_axis == 0 ? grid->nz : grid->nx;
    if (nd > BD_NOALLOC_MAX_DEPTH) {
        set_error("bd_mc_run_noalloc: the depth axis has " + std::to_string(nd) +
                  " cells; the allocation-free transport's scratch holds at most " +
                  std::to_string(BD_NOALLOC_MAX_DEPTH) + " (use bd_mc_run)");
        return BD_EINVAL;
    }
    BD_CHECK(!lateral || n_depth == nd, "bd_mc_run_noalloc: n_depth must equal the depth-axis cell count");
    int dev = 0, ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("bd_mc_run_noalloc: no HIP device");
        return BD_ENODEV;
    }
    BD_HIP(hipGetDevice(&dev));
    BD_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "bd_mc_run_noalloc: device index out of range");
    double *s = nullptr;
    BD_TRY(scratch_address(dev, &s));
    hipStream_t st = (hipStream_t)stream;
    const int nb_clear = blocks_for(2 * nvox, 256) > 4096 ? 4096 : blocks_for(2 * nvox, 256);
    hipLaunchKernelGGL(k_noalloc_clear, dim3(nb_clear), dim3(256), 0, st, s, (int)nd, (float *)AL, 2 * nvox);
    BD_HIP(hipGetLastError());
    McArgs a;
    fill_mc_args(*grid, beam, a);
    a.opt = (const float4 *)optics;
    a.AL = (float2 *)AL;
    a.Lt = (float *)AL + 1;
    a.Ls = 2;
    a.acc = s + NOALLOC_ACC;
    a.lat = s + NOALLOC_LAT;
    a.counter = (unsigned long long *)s;
    a.ep = nullptr;
    int ncu = 256;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t want = (beam->n_photons + 255) / 256;
    const int64_t cap = (int64_t)ncu * 8;
    const int nblk = (int)(want < 1 ? 1 : (want < cap ? want : cap));
    hipLaunchKernelGGL(k_mc, dim3(nblk), dim3(256), 0, st, a);
    BD_HIP(hipGetLastError());
    double acc[7];
    BD_HIP(hipMemcpyAsync(acc, s + NOALLOC_ACC, sizeof(acc), hipMemcpyDeviceToHost, st));
    if (lateral) BD_HIP(hipMemcpyAsync(lateral, s + NOALLOC_LAT, sizeof(double) * nd, hipMemcpyDeviceToHost, st));
    BD_HIP(hipStreamSynchronize(st));
    out->specular = acc[0];
    out->diffuse = acc[1];
    out->transmitted = acc[2];
    out->lateral = acc[3];
    out->absorbed = acc[4];
    out->events = acc[5];
    out->tally_atomics = acc[6];
    out->tally_records = 0.0;  // this entry point adds path lengths inline (no record buffer)
    out->record_ms = 0.0;
    out->kernel_ms = -1.0;  // not measured: timing events would be allocations
    return BD_OK;
}

// ------------------------------------------------------------------ atomic-rate probe
namespace bd {

// The ceiling the transport's tally rate is set against: float atomic adds on
// the (A, L) pair of one voxel of a float2-per-voxel grid, at voxel indices
// hashed uniformly over [0, span).  Each lane issues its pairs back to back with
// no dependent load between them, so the launch measures the memory system's
// atomic throughput for that address spread rather than a latency chain.
constexpr int PROBE_PAIRS_PER_LANE = 16;

__global__ __launch_bounds__(256) void k_atomic_probe(float2 *__restrict__ al, uint64_t span, uint32_t seed)
{
    uint32_t h = ((blockIdx.x * 256u + threadIdx.x + 1u) * 0x9E3779B9u) ^ seed;
    if (h == 0u) h = 1u;
#pragma unroll
    for (int i = 0; i < PROBE_PAIRS_PER_LANE; ++i) {
        h ^= h << 13;  // xorshift32
        h ^= h >> 17;
        h ^= h << 5;
        const uint64_t v = ((uint64_t)h * span) >> 32;  // < span
        float *cell = (float *)&al[v];
        unsafeAtomicAdd(cell, 1.0f);
        unsafeAtomicAdd(cell + 1, 1.0f);
    }
}

}  // namespace bd

extern "C" int bd_probe_atomic_rate(int device, int64_t n_voxels, int64_t span, int64_t n_pairs, double *kernel_ms)
{
    using namespace bd;
    BD_CHECK(kernel_ms, "bd_probe_atomic_rate: null argument");
    BD_CHECK(n_voxels > 0 && n_voxels <= (int64_t(1) << 31), "bd_probe_atomic_rate: n_voxels out of range");
    BD_CHECK(span > 0 && span <= n_voxels, "bd_probe_atomic_rate: span must be in [1, n_voxels]");
    const int64_t per_block = 256 * PROBE_PAIRS_PER_LANE;
    BD_CHECK(n_pairs > 0 && n_pairs % per_block == 0 && n_pairs / per_block <= (int64_t(1) << 20),
             "bd_probe_atomic_rate: n_pairs must be a positive multiple of 4096, at most 2^32");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("bd_probe_atomic_rate: no HIP device");
        return BD_ENODEV;
    }
    BD_CHECK(device >= 0 && device < ndev, "bd_probe_atomic_rate: device index out of range");
    BD_HIP(hipSetDevice(device));
    float2 *al = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e[2] = {nullptr, nullptr};
    int rc = BD_OK;
    auto run = [&]() -> int {
        BD_HIP(hipMalloc(&al, sizeof(float2) * (size_t)n_voxels));
        BD_HIP(hipStreamCreate(&s));
        BD_HIP(hipEventCreate(&e[0]));
        BD_HIP(hipEventCreate(&e[1]));
        BD_HIP(hipMemsetAsync(al, 0, sizeof(float2) * (size_t)n_voxels, s));
        const unsigned nb = (unsigned)(n_pairs / per_block);
        // one untimed launch brings the grid into the caches the timed one sees
        hipLaunchKernelGGL(k_atomic_probe, dim3(nb), dim3(256), 0, s, al, (uint64_t)span, 0x2545F491u);
        BD_HIP(hipGetLastError());
        BD_HIP(hipEventRecord(e[0], s));
        hipLaunchKernelGGL(k_atomic_probe, dim3(nb), dim3(256), 0, s, al, (uint64_t)span, 0x9E3779B9u);
        BD_HIP(hipGetLastError());
        BD_HIP(hipEventRecord(e[1], s));
        BD_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        BD_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
        *kernel_ms = ms;
        return BD_OK;
    };
    rc = run();
    if (e[0]) (void)hipEventDestroy(e[0]);
    if (e[1]) (void)hipEventDestroy(e[1]);
    if (s) (void)hipStreamDestroy(s);
    if (al) (void)hipFree(al);
    return rc;
}
