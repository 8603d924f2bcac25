// bd_api.hip -- the extern "C" entry points of bugdef.h: context lifetime,
// field transfer and the per-call orchestration of the kernels.
#include "bd_internal.hpp"

#include <cstdio>

#include <cstdlib>

#include <cmath>
#include <cstring>
#include <new>

namespace bd {
const char *last_error();
void convert_f64_to_f32(Ctx *c, const double *s, float *d, int64_t n);
void convert_f32_to_f64(Ctx *c, const float *s, double *d, int64_t n);
void mask_to_flag(Ctx *c, const uint8_t *m, uint8_t bit);
void flag_to_mask(Ctx *c, uint8_t *m, uint8_t bit);
void hpos_from_h(Ctx *c);
void al_get(Ctx *c, float *d, int which);
void al_set(Ctx *c, const float *s, int which);

namespace {

enum Kind { KF32, KF64, KU8, KFLAG, KAL, KPARAM };

struct Ref {
    Kind kind;
    void **slot;     // device pointer slot (for lazily allocated fields)
    int64_t n;       // expected element count
    uint8_t bit;     // KFLAG
    int which;       // KAL: 0 A, 1 L
    Field *param;    // KPARAM
    bool param_f64;  // KPARAM stored as f64 (removal temperatures)
};

int resolve(Ctx *c, int field, Ref *r)
{
    std::memset(r, 0, sizeof(*r));
    r->n = c->nvox;
    switch (field) {
    case BD_FIELD_T: r->kind = KF32; r->slot = (void **)&c->T; return BD_OK;
    case BD_FIELD_H_ACC: r->kind = KF32; r->slot = (void **)&c->H; return BD_OK;
    case BD_FIELD_OMEGA: r->kind = KF32; r->slot = (void **)&c->Om; return BD_OK;
    case BD_FIELD_CHI: r->kind = KF32; r->slot = (void **)&c->chi; return BD_OK;
    case BD_FIELD_Q: r->kind = KF32; r->slot = (void **)&c->Q; return BD_OK;
    case BD_FIELD_K: r->kind = KF32; r->slot = (void **)&c->k; return BD_OK;
    case BD_FIELD_RHO_CP: r->kind = KF32; r->slot = (void **)&c->rc; return BD_OK;
    case BD_FIELD_H_LAT: r->kind = KF32; r->slot = (void **)&c->hlat; return BD_OK;
    case BD_FIELD_DRY: r->kind = KFLAG; r->bit = F_DRY; r->slot = (void **)&c->flags; return BD_OK;
    case BD_FIELD_ABLATED: r->kind = KFLAG; r->bit = F_ABL; r->slot = (void **)&c->flags; return BD_OK;
    case BD_FIELD_MU_A: r->kind = KF32; r->slot = (void **)&c->mu_a; return BD_OK;
    case BD_FIELD_MU_S: r->kind = KF32; r->slot = (void **)&c->mu_s; return BD_OK;
    case BD_FIELD_G: r->kind = KF32; r->slot = (void **)&c->gg; return BD_OK;
    case BD_FIELD_N: r->kind = KF32; r->slot = (void **)&c->nn; return BD_OK;
    case BD_FIELD_WET_MU_A: r->kind = KF32; r->slot = (void **)&c->wet_mu_a; return BD_OK;
    case BD_FIELD_WET_K: r->kind = KF32; r->slot = (void **)&c->wet_k; return BD_OK;
    case BD_FIELD_WET_RHO_CP: r->kind = KF32; r->slot = (void **)&c->wet_rc; return BD_OK;
    case BD_FIELD_T_PEAK: r->kind = KF32; r->slot = (void **)&c->Tpeak; return BD_OK;
    case BD_FIELD_ABLATION_TIME: r->kind = KF64; r->slot = (void **)&c->abl_time; return BD_OK;
    case BD_FIELD_A: r->kind = KAL; r->which = 0; r->slot = (void **)&c->AL; return BD_OK;
    case BD_FIELD_L: r->kind = KAL; r->which = 1; r->slot = (void **)&c->AL; return BD_OK;
    case BD_FIELD_PLUME_RAW: r->kind = KF32; r->slot = (void **)&c->P; r->n = c->pvox; return BD_OK;
    case BD_FIELD_PLUME_EXT: r->kind = KF32; r->slot = (void **)&c->Pext; r->n = c->pvox; return BD_OK;
    case BD_FIELD_VOID: r->kind = KU8; r->slot = (void **)&c->pvoid; r->n = c->pvox; return BD_OK;
    case BD_FIELD_TO_REMOVE: r->kind = KU8; r->slot = (void **)&c->to_remove; return BD_OK;
    default: break;
    }
    if (field >= BD_FIELD_DS_MU_A && field <= BD_FIELD_DS_T_REMOVE) {
        r->kind = KPARAM;
        r->param = &c->ds[field - BD_FIELD_DS_MU_A];
        r->param_f64 = field == BD_FIELD_DS_T_REMOVE;
        return BD_OK;
    }
    if (field >= BD_FIELD_CS_MU_A && field <= BD_FIELD_CS_T_REMOVE) {
        r->kind = KPARAM;
        r->param = &c->cs[field - BD_FIELD_CS_MU_A];
        r->param_f64 = field == BD_FIELD_CS_T_REMOVE;
        return BD_OK;
    }
    set_error("unknown field id " + std::to_string(field));
    return BD_EINVAL;
}

size_t dsize(int dtype) { return dtype == BD_F64 ? 8 : (dtype == BD_F32 ? 4 : 1); }

// host buffer -> staging (device), returns the staging pointer
int stage_in(Ctx *c, const void *host, size_t bytes)
{
    BD_TRY(ensure_staging(c, bytes < 64 ? 64 : bytes));
    BD_HIP(hipMemcpyAsync(c->staging, host, bytes, hipMemcpyHostToDevice, c->stream));
    return BD_OK;
}

}  // namespace
}  // namespace bd

using namespace bd;

extern "C" {

const char *bd_last_error(void) { return bd::last_error(); }

int bd_version(int *major, int *minor)
{
    if (major) *major = 1;
    if (minor) *minor = 0;
    return BD_OK;
}

int bd_device_count(int *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return BD_OK;
}

int bd_create(const bd_grid *grid, bd_ctx **out)
{
    BD_CHECK(grid && out, "bd_create: null argument");
    BD_CHECK(grid->nx > 0 && grid->ny > 0 && grid->nz > 0, "bd_create: grid dimensions must be positive");
    BD_CHECK(grid->dx > 0 && grid->dy > 0 && grid->dz > 0, "bd_create: voxel spacings must be positive");
    BD_CHECK((int64_t)grid->nx * grid->ny * grid->nz < ((int64_t)1 << 31),
             "bd_create: grid exceeds 2^31 voxels");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("bd_create: no HIP device is visible (the MI355X library needs a GPU)");
        return BD_ENODEV;
    }
    bd_ctx *h = new (std::nothrow) bd_ctx();
    if (!h) { set_error("bd_create: out of host memory"); return BD_EHIP; }
    Ctx *c = &h->c;
    c->g = *grid;
    c->nvox = (int64_t)grid->nx * grid->ny * grid->nz;
    c->pvox = 0;
    BD_HIP(hipGetDevice(&c->device));
    BD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    BD_HIP(hipEventCreate(&c->ev0));
    BD_HIP(hipEventCreate(&c->ev1));
    BD_HIP(hipEventCreate(&c->ev2));
    BD_HIP(hipEventCreate(&c->ev3));
    BD_HIP(hipEventCreate(&c->ev4));
    BD_HIP(hipEventCreate(&c->ev5));
    BD_HIP(hipEventCreate(&c->ev6));
    BD_HIP(hipEventCreate(&c->ev7));
    BD_HIP(hipEventCreate(&c->ev8));
    BD_HIP(hipEventCreate(&c->ev9));
    BD_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    BD_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    for (int i = 0; i < 7; ++i) BD_HIP(hipEventCreate(&c->gev[i]));
    if (const char *e = std::getenv("BUGDEF_MC_BLOCKS_PER_CU")) c->mc_blocks_per_cu = std::max(0, std::atoi(e));
    if (const char *e = std::getenv("BUGDEF_MC_DEFER_FLUENCE")) {
        const int v = std::atoi(e);
        c->defer_L = v == 0 ? 0 : (v == 2 ? 2 : 1);
    }
    if (const char *e = std::getenv("BUGDEF_MC_RECORDS_PER_PHOTON")) c->rec_per_photon = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("BUGDEF_SUBSTEP_PAIRS")) {
        c->two_step_allowed = std::atoi(e) != 0;
        c->two_step_forced = std::atoi(e) == 2;
    }
    if (const char *e = std::getenv("BUGDEF_TUNE_LOG")) c->tune_log = std::atoi(e) != 0;
    if (const char *e = std::getenv("BUGDEF_PLUME_PAIRS")) c->plume_pairs = std::atoi(e) != 0;
    // a fixed one-step tile "BYxXC" (BY 2, 4 or 8; XC >= 1) in place of the timed choice, e.g. for
    // counter runs, where the profiler's per-dispatch overhead distorts the tuning launches
    if (const char *e = std::getenv("BUGDEF_SUBSTEP_TILE")) {
        int by = 0, xc = 0;
        if (std::sscanf(e, "%dx%d", &by, &xc) == 2 && (by == 2 || by == 4 || by == 8) && xc >= 1) {
            c->tile_by = by;
            c->tile_xc = xc;
            c->tile_pinned = true;
        }
    }
    BD_HIP(hipHostMalloc(&c->host_pinned, 64 * sizeof(double), hipHostMallocDefault));
    BD_TRY(ensure(c, (void **)&c->red, sizeof(double) * (64 + 8192)));
    BD_TRY(ensure(c, (void **)&c->counter, sizeof(unsigned long long) * COUNTER_SLOTS));
    *out = h;
    return BD_OK;
}

int bd_destroy(bd_ctx *h)
{
    if (!h) return BD_OK;
    Ctx *c = &h->c;
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *bufs[] = {c->T, c->T2, c->H, c->Om, c->chi, c->Q, c->k, c->rc, c->hlat, c->flags, c->mu_a, c->mu_s,
                    c->gg, c->nn, c->wet_mu_a, c->wet_k, c->wet_rc, c->Tpeak, c->abl_time, c->AL, c->optics,
                    c->P, c->P2, c->Pext, c->pvoid, c->to_remove, c->red, c->lat, c->counter, c->staging,
                    c->splat_tmp1, c->splat_tmp2, c->kb[0], c->kb[1], c->kb[2], c->kl[0], c->kl[1], c->kl[2],
                    c->Qs, c->Qsur, c->Qaux, c->Hb, c->flagsb, c->rec, c->rec_sorted, c->rec_tcnt,
                    c->rec_tstart, c->rec_off, c->rec_cstart, c->rec_cmap, c->rec_b, c->Lb[0], c->Lb[1],
                    c->led_part};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (auto &f : c->ds)
        if (f.ptr) (void)hipFree(f.ptr);
    for (auto &f : c->cs)
        if (f.ptr) (void)hipFree(f.ptr);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->ep) (void)hipFree(c->ep);
    if (c->host_pinned) (void)hipHostFree(c->host_pinned);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->ev3) (void)hipEventDestroy(c->ev3);
    if (c->ev4) (void)hipEventDestroy(c->ev4);
    if (c->ev5) (void)hipEventDestroy(c->ev5);
    if (c->ev6) (void)hipEventDestroy(c->ev6);
    if (c->ev7) (void)hipEventDestroy(c->ev7);
    if (c->ev8) (void)hipEventDestroy(c->ev8);
    if (c->ev9) (void)hipEventDestroy(c->ev9);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    for (int i = 0; i < 7; ++i)
        if (c->gev[i]) (void)hipEventDestroy(c->gev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete h;
    return BD_OK;
}

int bd_set_physics(bd_ctx *h, const bd_physics *p)
{
    BD_CHECK(h && p, "bd_set_physics: null argument");
    BD_CHECK(p->dt_sub > 0.0, "bd_set_physics: dt_sub must be positive");
    BD_CHECK(p->air_gap_k0 < 0 || (p->air_gap_k0 <= p->air_gap_k1 && p->air_gap_k1 <= h->c.g.nz),
             "bd_set_physics: air_gap_range outside the grid");
    // the captured epoch graph holds these values as kernel arguments
    if (!h->c.has_phys || std::memcmp(&h->c.phys, p, sizeof(*p)) != 0) h->c.graph_stale = true;
    h->c.phys = *p;
    h->c.has_phys = true;
    return BD_OK;
}

int bd_set_plume(bd_ctx *h, const bd_plume_params *p)
{
    BD_CHECK(h && p, "bd_set_plume: null argument");
    Ctx *c = &h->c;
    if (std::memcmp(&c->plume, p, sizeof(*p)) != 0) c->graph_stale = true;
    if (!p->enabled) {
        c->plume = *p;
        return BD_OK;
    }
    BD_CHECK(p->nx > 0 && p->ny > 0 && p->nz > 0, "bd_set_plume: plume grid dimensions must be positive");
    const bool same = p->nx == c->g.nx && p->ny == c->g.ny && p->nz == c->g.nz;
    BD_CHECK(same || (p->model == 0 && !p->blast_expansion),
             "bd_set_plume: the separable/blast plume needs the plume grid equal to the thermal grid");
    const int64_t pv = (int64_t)p->nx * p->ny * p->nz;
    if (c->pvox && c->pvox != pv) {
        set_error("bd_set_plume: the plume grid cannot change size after allocation");
        return BD_ESTATE;
    }
    c->plume = *p;
    c->pvox = pv;
    BD_TRY(ensure(c, (void **)&c->P, pv * sizeof(float)));
    BD_TRY(ensure(c, (void **)&c->P2, pv * sizeof(float)));
    BD_TRY(ensure(c, (void **)&c->Pext, pv * sizeof(float)));
    BD_TRY(ensure(c, (void **)&c->pvoid, pv));
    return BD_OK;
}

int bd_set_plume_kernels(bd_ctx *h, int which, const float *kx, int nkx, const float *ky, int nky, const float *kz,
                         int nkz, int rx, int ry, int rz)
{
    BD_CHECK(h && kx && ky && kz, "bd_set_plume_kernels: null argument");
    BD_CHECK(which == 0 || which == 1, "bd_set_plume_kernels: which must be 0 (blast) or 1 (legacy)");
    BD_CHECK(nkx == 2 * rx + 1 && nky == 2 * ry + 1 && nkz == 2 * rz + 1,
             "bd_set_plume_kernels: tap counts must be 2r+1");
    Ctx *c = &h->c;
    float **dst = which == 0 ? c->kb : c->kl;
    int *nk = which == 0 ? c->nkb : c->nkl;
    int *r = which == 0 ? c->rb : c->rl;
    const float *src[3] = {kx, ky, kz};
    const int n[3] = {nkx, nky, nkz}, rr[3] = {rx, ry, rz};
    for (int a = 0; a < 3; ++a) {
        if (dst[a] && nk[a] < n[a]) {
            BD_HIP(hipFree(dst[a]));
            dst[a] = nullptr;
        }
        if (nk[a] != n[a] || r[a] != rr[a]) c->graph_stale = true;
        if (!dst[a]) BD_HIP(hipMalloc((void **)&dst[a], sizeof(float) * n[a]));
        BD_HIP(hipMemcpyAsync(dst[a], src[a], sizeof(float) * n[a], hipMemcpyHostToDevice, c->stream));
        nk[a] = n[a];
        r[a] = rr[a];
    }
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_upload(bd_ctx *h, int field, int dtype, const void *host, int64_t n)
{
    BD_CHECK(h && host, "bd_upload: null argument");
    BD_CHECK(dtype == BD_F32 || dtype == BD_F64 || dtype == BD_U8, "bd_upload: unknown dtype");
    Ctx *c = &h->c;
    Ref r;
    BD_TRY(resolve(c, field, &r));
    if (r.kind == KPARAM) {
        BD_CHECK(n == 1 || n == c->nvox, "bd_upload: a dry/char parameter takes 1 value or one per voxel");
        BD_CHECK(dtype != BD_U8, "bd_upload: parameters are floating point");
        Field &f = *r.param;
        const size_t esz = r.param_f64 ? 8 : 4;
        if (f.ptr && f.n != n) {
            BD_HIP(hipFree(f.ptr));
            f.ptr = nullptr;
        }
        if (!f.ptr) {
            BD_HIP(hipMalloc(&f.ptr, esz * (size_t)n));
            c->graph_stale = true;  // a new pointer (and per-voxel/scalar mode) for the captured kernels
        }
        f.n = n;
        f.bytes = esz * (size_t)n;
        BD_TRY(stage_in(c, host, dsize(dtype) * (size_t)n));
        if (r.param_f64) {
            if (dtype == BD_F64) BD_HIP(hipMemcpyAsync(f.ptr, c->staging, 8 * n, hipMemcpyDeviceToDevice, c->stream));
            else convert_f32_to_f64(c, (const float *)c->staging, (double *)f.ptr, n);
        } else {
            if (dtype == BD_F32) BD_HIP(hipMemcpyAsync(f.ptr, c->staging, 4 * n, hipMemcpyDeviceToDevice, c->stream));
            else convert_f64_to_f32(c, (const double *)c->staging, (float *)f.ptr, n);
        }
        BD_HIP(hipGetLastError());
        BD_HIP(hipStreamSynchronize(c->stream));
        return BD_OK;
    }
    if (r.n == 0) {
        set_error("bd_upload: plume field before bd_set_plume");
        return BD_ESTATE;
    }
    BD_CHECK(n == r.n, "bd_upload: element count does not match the grid");
    switch (r.kind) {
    case KF32: {
        BD_TRY(ensure(c, r.slot, (size_t)n * 4));
        if (dtype == BD_F32) {
            BD_HIP(hipMemcpyAsync(*r.slot, host, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        } else {
            BD_CHECK(dtype == BD_F64, "bd_upload: a float field takes f32 or f64 data");
            BD_TRY(stage_in(c, host, (size_t)n * 8));
            convert_f64_to_f32(c, (const double *)c->staging, (float *)*r.slot, n);
        }
        if (field == BD_FIELD_H_ACC) {
            BD_TRY(ensure(c, (void **)&c->flags, (size_t)c->nvox));
            hpos_from_h(c);
        }
        break;
    }
    case KF64: {
        BD_TRY(ensure(c, r.slot, (size_t)n * 8));
        if (dtype == BD_F64) {
            BD_HIP(hipMemcpyAsync(*r.slot, host, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        } else {
            BD_CHECK(dtype == BD_F32, "bd_upload: a float field takes f32 or f64 data");
            BD_TRY(stage_in(c, host, (size_t)n * 4));
            convert_f32_to_f64(c, (const float *)c->staging, (double *)*r.slot, n);
        }
        break;
    }
    case KU8: {
        BD_CHECK(dtype == BD_U8, "bd_upload: a mask field takes u8 data");
        BD_TRY(ensure(c, r.slot, (size_t)n));
        BD_HIP(hipMemcpyAsync(*r.slot, host, (size_t)n, hipMemcpyHostToDevice, c->stream));
        break;
    }
    case KFLAG: {
        BD_CHECK(dtype == BD_U8, "bd_upload: a mask field takes u8 data");
        BD_TRY(ensure(c, (void **)&c->flags, (size_t)n));
        BD_TRY(stage_in(c, host, (size_t)n));
        mask_to_flag(c, (const uint8_t *)c->staging, r.bit);
        break;
    }
    case KAL: {
        if (r.which == 1) {
            BD_TRY(join_records(c));
            lag_drop(c);  // the uploaded path lengths replace those of a lagged launch
        } else {
            BD_TRY(lag_finalize(c));
        }
        BD_TRY(ensure(c, (void **)&c->AL, (size_t)n * sizeof(float2)));
        BD_TRY(ensure_staging(c, (size_t)n * 12));
        float *tmp = (float *)((char *)c->staging + (size_t)n * 8);
        if (dtype == BD_F32) {
            BD_HIP(hipMemcpyAsync(tmp, host, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        } else {
            BD_CHECK(dtype == BD_F64, "bd_upload: a float field takes f32 or f64 data");
            BD_HIP(hipMemcpyAsync(c->staging, host, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            convert_f64_to_f32(c, (const double *)c->staging, tmp, n);
        }
        al_set(c, tmp, r.which);
        break;
    }
    default: break;
    }
    BD_HIP(hipGetLastError());
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_download(bd_ctx *h, int field, int dtype, void *host, int64_t n)
{
    BD_CHECK(h && host, "bd_download: null argument");
    Ctx *c = &h->c;
    Ref r;
    BD_TRY(resolve(c, field, &r));
    BD_CHECK(r.kind != KPARAM, "bd_download: parameters are write-only");
    if (r.n == 0) {
        set_error("bd_download: plume field before bd_set_plume");
        return BD_ESTATE;
    }
    BD_CHECK(n == r.n, "bd_download: element count does not match the grid");
    if (!*r.slot) {
        set_error("bd_download: field has never been written");
        return BD_ESTATE;
    }
    switch (r.kind) {
    case KF32:
        if (dtype == BD_F32) {
            BD_HIP(hipMemcpyAsync(host, *r.slot, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        } else {
            BD_CHECK(dtype == BD_F64, "bd_download: a float field gives f32 or f64 data");
            BD_TRY(ensure_staging(c, (size_t)n * 8));
            convert_f32_to_f64(c, (const float *)*r.slot, (double *)c->staging, n);
            BD_HIP(hipMemcpyAsync(host, c->staging, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
        }
        break;
    case KF64:
        BD_CHECK(dtype == BD_F64, "bd_download: ablation_time is f64");
        BD_HIP(hipMemcpyAsync(host, *r.slot, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
        break;
    case KU8:
        BD_CHECK(dtype == BD_U8, "bd_download: a mask field gives u8 data");
        BD_HIP(hipMemcpyAsync(host, *r.slot, (size_t)n, hipMemcpyDeviceToHost, c->stream));
        break;
    case KFLAG:
        BD_CHECK(dtype == BD_U8, "bd_download: a mask field gives u8 data");
        BD_TRY(ensure_staging(c, (size_t)n));
        flag_to_mask(c, (uint8_t *)c->staging, r.bit);
        BD_HIP(hipMemcpyAsync(host, c->staging, (size_t)n, hipMemcpyDeviceToHost, c->stream));
        break;
    case KAL: {
        BD_TRY(lag_finalize(c));
        BD_TRY(ensure_staging(c, (size_t)n * 12));
        float *tmp = (float *)((char *)c->staging + (size_t)n * 8);
        al_get(c, tmp, r.which);
        if (dtype == BD_F32) {
            BD_HIP(hipMemcpyAsync(host, tmp, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        } else {
            BD_CHECK(dtype == BD_F64, "bd_download: a float field gives f32 or f64 data");
            convert_f32_to_f64(c, tmp, (double *)c->staging, n);
            BD_HIP(hipMemcpyAsync(host, c->staging, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
        }
        break;
    }
    default: break;
    }
    BD_HIP(hipGetLastError());
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_fill(bd_ctx *h, int field, double value)
{
    BD_CHECK(h, "bd_fill: null argument");
    Ctx *c = &h->c;
    Ref r;
    BD_TRY(resolve(c, field, &r));
    BD_CHECK(r.kind == KF32 || r.kind == KF64 || r.kind == KU8 || r.kind == KFLAG, "bd_fill: field cannot be filled");
    if (r.n == 0) {
        set_error("bd_fill: plume field before bd_set_plume");
        return BD_ESTATE;
    }
    if (r.kind == KFLAG) {
        BD_TRY(ensure(c, (void **)&c->flags, (size_t)c->nvox));
        BD_TRY(ensure_staging(c, (size_t)c->nvox));
        BD_TRY(launch_fill(c, c->staging, K_U8, c->nvox, value != 0.0 ? 1.0 : 0.0));
        mask_to_flag(c, (const uint8_t *)c->staging, r.bit);
    } else {
        const size_t esz = r.kind == KF64 ? 8 : (r.kind == KF32 ? 4 : 1);
        BD_TRY(ensure(c, r.slot, (size_t)r.n * esz));
        BD_TRY(launch_fill(c, *r.slot, r.kind == KF64 ? K_F64 : (r.kind == KF32 ? K_F32 : K_U8), r.n, value));
        if (field == BD_FIELD_H_ACC) {
            BD_TRY(ensure(c, (void **)&c->flags, (size_t)c->nvox));
            hpos_from_h(c);
        }
    }
    BD_HIP(hipGetLastError());
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_download_slice(bd_ctx *h, int field, int plane, int index, int stride, float *host, int64_t n)
{
    BD_CHECK(h && host, "bd_download_slice: null argument");
    Ctx *c = &h->c;
    BD_CHECK(plane == 0 || plane == 1, "bd_download_slice: plane must be 0 (y-z) or 1 (x-z)");
    BD_CHECK(stride >= 1, "bd_download_slice: stride must be >= 1");
    const int lim = plane == 1 ? c->g.ny : c->g.nx;
    BD_CHECK(index >= 0 && index < lim, "bd_download_slice: index outside the grid");
    const int64_t na = plane == 1 ? (c->g.nx + stride - 1) / stride : (c->g.ny + stride - 1) / stride;
    const int64_t expect = na * ((c->g.nz + stride - 1) / stride);
    BD_CHECK(n == expect, "bd_download_slice: output size mismatch");
    Ref r;
    BD_TRY(resolve(c, field, &r));
    const bool plume_field = field == BD_FIELD_PLUME_RAW || field == BD_FIELD_PLUME_EXT;
    BD_CHECK(r.kind != KPARAM && (r.n == c->nvox || plume_field) && field != BD_FIELD_VOID,
             "bd_download_slice: field is not a thermal-grid field");
    if (r.n == 0 || !*r.slot) {
        set_error("bd_download_slice: field has never been written");
        return BD_ESTATE;
    }
    if (r.kind == KAL) BD_TRY(lag_finalize(c));
    BD_TRY(ensure_staging(c, (size_t)n * 4));
    if (plume_field) {
        // plume cells under the thermal plane's cell centres (identity for equal grids)
        BD_TRY(launch_slice_plume(c, (const float *)*r.slot, plane, index, stride, (float *)c->staging, n));
    } else {
        int kind = r.kind == KF32 ? K_F32 : r.kind == KF64 ? K_F64 : r.kind == KU8 ? K_U8
                 : r.kind == KFLAG ? (r.bit == F_ABL ? K_FLAG_ABL : K_FLAG_DRY) : (r.which ? K_AL_L : K_AL_A);
        BD_TRY(launch_slice(c, *r.slot, kind, plane, index, stride, (float *)c->staging, n));
    }
    BD_HIP(hipMemcpyAsync(host, c->staging, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_phase_slice(bd_ctx *h, int plane, int index, int stride, const uint8_t *gap_z, uint8_t *out, int64_t n)
{
    BD_CHECK(h && out, "bd_phase_slice: null argument");
    Ctx *c = &h->c;
    BD_CHECK(plane == 0 || plane == 1, "bd_phase_slice: plane must be 0 (y-z) or 1 (x-z)");
    BD_CHECK(stride >= 1, "bd_phase_slice: stride must be >= 1");
    const int lim = plane == 1 ? c->g.ny : c->g.nx;
    BD_CHECK(index >= 0 && index < lim, "bd_phase_slice: index outside the grid");
    const int64_t na = plane == 1 ? (c->g.nx + stride - 1) / stride : (c->g.ny + stride - 1) / stride;
    BD_CHECK(n == na * ((c->g.nz + stride - 1) / stride), "bd_phase_slice: output size mismatch");
    BD_TRY(ensure_state(c));
    BD_TRY(ensure_staging(c, (size_t)n + (size_t)c->g.nz + 64));
    uint8_t *dev_out = (uint8_t *)c->staging;
    uint8_t *dev_gap = nullptr;
    if (gap_z) {
        dev_gap = dev_out + n;
        BD_HIP(hipMemcpyAsync(dev_gap, gap_z, (size_t)c->g.nz, hipMemcpyHostToDevice, c->stream));
    }
    BD_TRY(launch_phase_slice(c, dev_gap, plane, index, stride, dev_out, n));
    BD_HIP(hipMemcpyAsync(out, dev_out, (size_t)n, hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

static int need_thermal(Ctx *c, const char *who)
{
    if (!c->has_phys) {
        set_error(std::string(who) + ": bd_set_physics has not been called");
        return BD_ESTATE;
    }
    return ensure_state(c);
}

int bd_substeps(bd_ctx *h, int n_substeps, int is_active, int flags)
{
    BD_CHECK(h, "bd_substeps: null argument");
    BD_CHECK(n_substeps >= 0, "bd_substeps: n_substeps must be >= 0");
    Ctx *c = &h->c;
    BD_TRY(need_thermal(c, "bd_substeps"));
    if (flags & BD_SUB_ARRHENIUS) BD_TRY(ensure(c, (void **)&c->Om, (size_t)c->nvox * 4));
    if (flags & BD_SUB_PEAK) BD_TRY(ensure(c, (void **)&c->Tpeak, (size_t)c->nvox * 4));
    if (n_substeps == 0) return BD_OK;
    BD_TRY(launch_substeps(c, n_substeps, is_active, flags, false));
    if ((flags & BD_SUB_PLUME) && c->plume.enabled && c->plume.model == 0)
        BD_TRY(launch_plume_substeps(c, n_substeps, c->phys.dt_sub, plume_decay(c, c->phys.dt_sub)));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_plateau(bd_ctx *h)
{
    BD_CHECK(h, "bd_plateau: null argument");
    BD_TRY(need_thermal(&h->c, "bd_plateau"));
    BD_TRY(launch_plateau(&h->c));
    BD_HIP(hipStreamSynchronize(h->c.stream));
    return BD_OK;
}

int bd_arrhenius(bd_ctx *h, double A, double EaR, double dt)
{
    BD_CHECK(h, "bd_arrhenius: null argument");
    Ctx *c = &h->c;
    BD_TRY(ensure_state(c));
    BD_TRY(ensure(c, (void **)&c->Om, (size_t)c->nvox * 4));
    BD_TRY(launch_arrhenius(c, A, EaR, dt));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_char(bd_ctx *h, double A, double EaR, double dt)
{
    BD_CHECK(h, "bd_char: null argument");
    Ctx *c = &h->c;
    BD_TRY(ensure_state(c));
    BD_TRY(launch_char(c, A, EaR, dt));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_plume_substeps(bd_ctx *h, int n, double dt, double decay)
{
    BD_CHECK(h, "bd_plume_substeps: null argument");
    Ctx *c = &h->c;
    if (!c->plume.enabled || !c->P) {
        set_error("bd_plume_substeps: plume not configured (bd_set_plume)");
        return BD_ESTATE;
    }
    BD_CHECK(n >= 0 && dt >= 0.0, "bd_plume_substeps: n and dt must be >= 0");
    BD_TRY(launch_plume_substeps(c, n, dt, decay));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_mc_run(bd_ctx *h, const bd_beam *b, bd_mc_result *out)
{
    BD_CHECK(h && b && out, "bd_mc_run: null argument");
    Ctx *c = &h->c;
    BD_CHECK(b->n_photons > 0, "bd_mc_run: n_photons must be positive");
    BD_CHECK(b->beam_profile >= 0 && b->beam_profile <= 3, "bd_mc_run: unknown beam profile");
    BD_CHECK(b->launch_axis == 0 || b->launch_axis == 1, "bd_mc_run: launch_axis must be 0 (z) or 1 (x)");
    BD_CHECK(b->max_events > 0, "bd_mc_run: max_events must be positive");
    BD_CHECK(c->mu_a && c->mu_s && c->gg && c->nn, "bd_mc_run: optics (mu_a, mu_s, g, n) not uploaded");
    BD_TRY(launch_mc(c, b, out));
    return BD_OK;
}

int bd_mc_launch(bd_ctx *h, const bd_beam *b)
{
    BD_CHECK(h && b, "bd_mc_launch: null argument");
    Ctx *c = &h->c;
    BD_CHECK(b->n_photons > 0, "bd_mc_launch: n_photons must be positive");
    BD_CHECK(b->beam_profile >= 0 && b->beam_profile <= 3, "bd_mc_launch: unknown beam profile");
    BD_CHECK(b->launch_axis == 0 || b->launch_axis == 1, "bd_mc_launch: launch_axis must be 0 (z) or 1 (x)");
    BD_CHECK(b->max_events > 0, "bd_mc_launch: max_events must be positive");
    BD_CHECK(c->mu_a && c->mu_s && c->gg && c->nn, "bd_mc_launch: optics (mu_a, mu_s, g, n) not uploaded");
    return launch_mc_enqueue(c, b, nullptr, 1, REC_MAIN);
}

int bd_mc_lateral(bd_ctx *h, double *host, int64_t n_depth)
{
    BD_CHECK(h && host, "bd_mc_lateral: null argument");
    Ctx *c = &h->c;
    BD_CHECK(c->lat && n_depth == c->lat_n, "bd_mc_lateral: no MC result of that depth size");
    BD_HIP(hipMemcpyAsync(host, c->lat, sizeof(double) * n_depth, hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_phase_update(bd_ctx *h, double t_now, int64_t *n_new)
{
    BD_CHECK(h && n_new, "bd_phase_update: null argument");
    Ctx *c = &h->c;
    BD_TRY(need_thermal(c, "bd_phase_update"));
    double abl_J = 0.0;
    BD_TRY(launch_phase(c, t_now, n_new, &abl_J));
    return BD_OK;
}

int bd_plume_void(bd_ctx *h)
{
    BD_CHECK(h, "bd_plume_void: null argument");
    if (!h->c.plume.enabled) { set_error("bd_plume_void: plume not configured"); return BD_ESTATE; }
    BD_TRY(ensure_state(&h->c));
    BD_TRY(launch_plume_void(&h->c));
    BD_HIP(hipStreamSynchronize(h->c.stream));
    return BD_OK;
}

int bd_plume_inject(bd_ctx *h)
{
    BD_CHECK(h, "bd_plume_inject: null argument");
    if (!h->c.plume.enabled) { set_error("bd_plume_inject: plume not configured"); return BD_ESTATE; }
    BD_TRY(ensure(&h->c, (void **)&h->c.to_remove, (size_t)h->c.nvox));
    BD_TRY(launch_plume_inject(&h->c));
    BD_HIP(hipStreamSynchronize(h->c.stream));
    return BD_OK;
}

int bd_plume_sync(bd_ctx *h)
{
    BD_CHECK(h, "bd_plume_sync: null argument");
    if (!h->c.plume.enabled) { set_error("bd_plume_sync: plume not configured"); return BD_ESTATE; }
    BD_TRY(launch_plume_sync(&h->c));
    BD_HIP(hipStreamSynchronize(h->c.stream));
    return BD_OK;
}

int bd_plume_update_separable(bd_ctx *h, double dt_epoch)
{
    BD_CHECK(h, "bd_plume_update_separable: null argument");
    if (!h->c.plume.enabled) { set_error("bd_plume_update_separable: plume not configured"); return BD_ESTATE; }
    BD_TRY(ensure(&h->c, (void **)&h->c.to_remove, (size_t)h->c.nvox));
    BD_TRY(launch_plume_update_sep(&h->c, dt_epoch));
    BD_HIP(hipStreamSynchronize(h->c.stream));
    return BD_OK;
}

int bd_plume_splat(bd_ctx *h, int which, const float *impulse, int Lx, int Ly, int Lz, int x0, int y0, int z0,
                   double delta_mu)
{
    BD_CHECK(h && impulse, "bd_plume_splat: null argument");
    BD_CHECK(which == 0 || which == 1, "bd_plume_splat: which must be 0 (blast) or 1 (legacy)");
    BD_CHECK(Lx > 0 && Ly > 0 && Lz > 0, "bd_plume_splat: impulse box must be non-empty");
    if (!h->c.plume.enabled || !h->c.P) { set_error("bd_plume_splat: plume not configured"); return BD_ESTATE; }
    return launch_plume_splat_box(&h->c, which, impulse, Lx, Ly, Lz, x0, y0, z0, delta_mu);
}

int bd_plume_diag(bd_ctx *h, double out[4])
{
    BD_CHECK(h && out, "bd_plume_diag: null argument");
    if (!h->c.plume.enabled) { set_error("bd_plume_diag: plume not configured"); return BD_ESTATE; }
    BD_TRY(ensure_state(&h->c));
    BD_TRY(launch_plume_diag(&h->c, out));
    return BD_OK;
}

int bd_reduce(bd_ctx *h, int field, int op, int mask_field, double *out)
{
    BD_CHECK(h && out, "bd_reduce: null argument");
    BD_CHECK(op >= 0 && op <= 2, "bd_reduce: op must be 0 (sum), 1 (max) or 2 (min)");
    Ctx *c = &h->c;
    Ref r;
    BD_TRY(resolve(c, field, &r));
    BD_CHECK(r.kind == KF32 || r.kind == KF64 || r.kind == KU8 || r.kind == KAL, "bd_reduce: unsupported field");
    if (!*r.slot) { set_error("bd_reduce: field has never been written"); return BD_ESTATE; }
    const uint8_t *mask = nullptr;
    uint8_t bits = 0xFF;
    if (mask_field >= 0) {
        Ref m;
        BD_TRY(resolve(c, mask_field, &m));
        BD_CHECK(m.kind == KFLAG || m.kind == KU8, "bd_reduce: mask must be a mask field");
        BD_CHECK(m.n == r.n, "bd_reduce: mask and field grids differ");
        if (!*m.slot) { set_error("bd_reduce: mask field has never been written"); return BD_ESTATE; }
        mask = (const uint8_t *)*m.slot;
        bits = m.kind == KFLAG ? m.bit : 0xFF;
    }
    if (r.kind == KAL) {
        BD_TRY(lag_finalize(c));
        BD_TRY(ensure_staging(c, (size_t)c->nvox * 4));
        al_get(c, (float *)c->staging, r.which);
        return launch_reduce(c, c->staging, K_F32, r.n, op, mask, bits, out);
    }
    const int kind = r.kind == KF32 ? K_F32 : (r.kind == KF64 ? K_F64 : K_U8);
    return launch_reduce(c, *r.slot, kind, r.n, op, mask, bits, out);
}

int bd_phase_map(bd_ctx *h, const uint8_t *is_air_gap, uint8_t *out, int64_t n)
{
    BD_CHECK(h && out, "bd_phase_map: null argument");
    Ctx *c = &h->c;
    BD_CHECK(n == c->nvox, "bd_phase_map: output size mismatch");
    BD_TRY(ensure_state(c));
    BD_TRY(ensure_staging(c, (size_t)c->nvox * 2));
    uint8_t *gap = nullptr;
    uint8_t *dev_out = (uint8_t *)c->staging;
    if (is_air_gap) {
        gap = dev_out + c->nvox;
        BD_HIP(hipMemcpyAsync(gap, is_air_gap, (size_t)c->nvox, hipMemcpyHostToDevice, c->stream));
    }
    BD_TRY(launch_phase_map(c, gap, dev_out));
    BD_HIP(hipMemcpyAsync(out, dev_out, (size_t)c->nvox, hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    return BD_OK;
}

int bd_sync(bd_ctx *h)
{
    BD_CHECK(h, "bd_sync: null argument");
    BD_TRY(lag_finalize(&h->c));  // every launched transport's path-length tally is complete
    BD_HIP(hipStreamSynchronize(h->c.stream));
    return BD_OK;
}

int bd_get_tuning(bd_ctx *h, int32_t *out, int n)
{
    BD_CHECK(h && out, "bd_get_tuning: null argument");
    BD_CHECK(n >= 6, "bd_get_tuning: out must hold 6 values");
    const Ctx *c = &h->c;
    out[0] = c->tile_by;
    out[1] = c->tile_xc;
    out[2] = c->tile2_by;
    out[3] = c->tile2_xc;
    out[4] = c->use_two ? 1 : 0;
    out[5] = c->tuned ? 1 : 0;
    return BD_OK;
}

}  // extern "C"
