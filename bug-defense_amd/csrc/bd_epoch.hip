// bd_epoch.hip -- one coupled epoch on the device (coupled_ablation.py:440-617).
//
// Order of work, all on the context's stream with the state device-resident:
//   1. power > 0: transport optics (+ plume), MC, Q = A * P / (N V)
//      power == 0: Q = 0
//   2. void mask, n_substeps fused thermal sub-steps (+ plume diffusion),
//      T_peak after the last one, conducted-energy ledger
//   3. phase/property update fused with removal, air-gap perforation
//   4. plume: injection of the removed cells, then tanh sync (diffusion model)
//      or the legacy separable update, then centre-column diagnostics
// The host waits once for the diagnostics at the end (callers receive n_new).
#include "bd_internal.hpp"

#include <atomic>
#include <utility>
#include <vector>

#include <cmath>

namespace bd {

int reduce_masked(Ctx *c, const float *src, int64_t n, const uint8_t *flags, uint8_t bit, bool invert, double *out);

}  // namespace bd

using namespace bd;

static int epoch_impl(bd_ctx *h, const bd_beam *beam, int n_substeps, double t_after, bool run_mc,
                      bd_epoch_diag *out);

extern "C" int bd_epoch(bd_ctx *h, const bd_beam *beam, int n_substeps, double t_after, bd_epoch_diag *out)
{
    return epoch_impl(h, beam, n_substeps, t_after, true, out);
}

// the same epoch with the absorbed-power field already in Q (a surrogate's
// prediction): no transport, so the reflectance/transmission entries are zero
// and the closure is not defined (reported as NaN)
extern "C" int bd_epoch_q(bd_ctx *h, const bd_beam *beam, int n_substeps, double t_after, bd_epoch_diag *out)
{
    return epoch_impl(h, beam, n_substeps, t_after, false, out);
}

static int epoch_impl(bd_ctx *h, const bd_beam *beam, int n_substeps, double t_after, bool run_mc,
                      bd_epoch_diag *out)
{
    BD_CHECK(h && beam && out, "bd_epoch: null argument");
    BD_CHECK(n_substeps >= 0, "bd_epoch: n_substeps must be >= 0");
    Ctx *c = &h->c;
    if (!c->has_phys) {
        set_error("bd_epoch: bd_set_physics has not been called");
        return BD_ESTATE;
    }
    BD_TRY(ensure_state(c));
    *out = bd_epoch_diag{};
    const bd_physics &p = c->phys;
    const double V = c->g.dx * c->g.dy * c->g.dz;
    const double dt_epoch = n_substeps * p.dt_sub;
    const bool plume = c->plume.enabled != 0;
    const bool diff_plume = plume && c->plume.model == 0;
    // The epoch is queued without a host wait until its end (one readback of the scalars), except
    // with the blast or separable plume, whose injection and update take the count of removed
    // cells on the host.
    const bool host_plume = plume && (c->plume.model != 0 || c->plume.blast_expansion);
    const bool mc = beam->power > 0.0 && run_mc;
    hipEvent_t e_start = c->ev2, e_opt = c->ev3;

    // ---- 1. optics
    BD_HIP(hipEventRecord(e_start, c->stream));
    if (beam->power > 0.0 && !run_mc) {
        BD_TRY(reduce_masked_enqueue(c, c->Q, c->nvox, c->flags, F_ABL, true, c->red + 40));
    } else if (mc) {
        // this epoch's record passes run at the next one; the absorbed power over the intact
        // cells goes to red[40] with the conversion to Q
        BD_TRY(launch_mc_enqueue(c, beam, nullptr, 1, REC_LAG, c->red + 40));
    } else {
        BD_TRY(join_records(c));
        lag_drop(c);  // the cleared tally supersedes a lagged launch's
        BD_TRY(launch_fill(c, c->Q, K_F32, c->nvox, 0.0));
        if (c->AL) BD_TRY(launch_fill(c, c->AL, K_F32, 2 * c->nvox, 0.0));
    }
    BD_HIP(hipEventRecord(e_opt, c->stream));

    // ---- 2. thermal sub-steps (+ plume diffusion in the void)
    if (diff_plume) BD_TRY(launch_plume_void(c));
    int flags = BD_SUB_PLATEAU | BD_SUB_MASK_Q | BD_SUB_PEAK;
    if (p.has_arrhenius) flags |= BD_SUB_ARRHENIUS;
    if (p.has_char) flags |= BD_SUB_CHAR;
    BD_TRY(launch_fill(c, c->red + 8, K_F64, 1, 0.0));
    BD_TRY(launch_substeps(c, 0, 1, flags, true));  // tile autotune (first call) outside the timed span
    BD_HIP(hipEventRecord(c->ev4, c->stream));
    BD_TRY(launch_substeps(c, n_substeps, 1, flags, true));
    BD_HIP(hipEventRecord(c->ev5, c->stream));
    if (diff_plume) BD_TRY(launch_plume_substeps(c, n_substeps, p.dt_sub, plume_decay(c, p.dt_sub)));
    BD_HIP(hipEventRecord(c->ev8, c->stream));

    // ---- 3. phase update + removal
    int64_t n_new_host = 0;
    if (host_plume) BD_TRY(launch_phase(c, t_after, &n_new_host, nullptr));
    else BD_TRY(launch_phase_enqueue(c, t_after, nullptr));

    // ---- 4. plume
    if (plume) {
        BD_TRY(launch_plume_void(c));  // void after removal (coupled_ablation.py:588,658)
        if (!host_plume) {
            BD_TRY(launch_plume_inject(c));  // point injection: nothing is added where nothing was removed
            BD_TRY(launch_plume_sync(c));
        } else {
            if (n_new_host > 0) BD_TRY(launch_plume_inject(c));
            if (diff_plume) BD_TRY(launch_plume_sync(c));
            else BD_TRY(launch_plume_update_sep(c, dt_epoch));
        }
        BD_TRY(launch_plume_diag_enqueue(c));
    }
    BD_TRY(join_records(c));
    BD_HIP(hipEventRecord(c->ev9, c->stream));
    BD_HIP(hipMemcpyAsync(c->host_pinned, c->red, 48 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipMemcpyAsync((char *)c->host_pinned + 48 * sizeof(double), c->counter, 2 * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    const double *r = (const double *)c->host_pinned;
    const unsigned long long *cnt = (const unsigned long long *)(r + 48);
    if (beam->power > 0.0 && !run_mc) {
        out->absorbed_J = r[40] * V * dt_epoch;
        out->absorbed_frac = r[40] * V / beam->power;
        out->mc_energy_closure = std::nan("");
    } else if (mc) {
        const double N = (double)beam->n_photons;
        const double *m = r + 16;  // the transport's FP64 scalar tallies
        out->absorbed_frac = m[4] / N;
        out->mc_energy_closure = (m[0] + m[1] + m[2] + m[3] + m[4]) / N;
        out->reflected_J = (m[0] + m[1]) / N * beam->power * dt_epoch;
        out->transmitted_J = m[2] / N * beam->power * dt_epoch;
        out->lateral_J = m[3] / N * beam->power * dt_epoch;
        out->R_spec = m[0] / N;
        out->R_diff = m[1] / N;
        out->transmission = m[2] / N;
        out->lateral_loss = m[3] / N;
        out->mc_events = m[5];
        out->mc_tally_atomics = m[6];
        out->mc_tally_records = m[7];
        out->absorbed_J = r[40] * V * dt_epoch;
    } else {
        out->absorbed_frac = 0.0;
        out->mc_energy_closure = 1.0;
    }
    out->conducted_J = -r[8];
    out->ablated_J = r[9];
    out->n_new = (int64_t)cnt[1];
    if (plume) {
        double d[4];
        plume_diag_from(r + 24, d);
        out->plume_tau = d[0];
        out->plume_transmission = d[1];
        out->plume_peak_mu_a = d[2];
        out->plume_mean_mu_a = d[3];
    } else {
        out->plume_tau = 0.0;
        out->plume_transmission = 1.0;
    }
    float ms = 0.f;
    if (mc) {
        BD_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));  // recorded around the transport kernel
        out->t_mc_kernel_ms = ms;
        BD_HIP(hipEventElapsedTime(&ms, c->ev7, c->ev6));  // the record passes of the previous epoch (0: none)
        out->t_mc_records_ms = ms;
    }
    BD_HIP(hipEventElapsedTime(&ms, e_start, e_opt));
    out->t_optical_ms = ms;
    BD_HIP(hipEventElapsedTime(&ms, e_opt, c->ev8));
    out->t_thermal_ms = ms;
    BD_HIP(hipEventElapsedTime(&ms, c->ev8, c->ev9));
    out->t_phase_ms = ms;
    BD_HIP(hipEventElapsedTime(&ms, c->ev4, c->ev5));
    out->t_substeps_ms = ms;
    return BD_OK;
}

// ---------------------------------------------------------------------------
// Graph-captured epoch.  The same work as bd_epoch, enqueued without any host
// wait and captured once per (power on/off, beam, n_substeps); every per-epoch
// scalar is read from the device EpochParams block, which the host refreshes
// with one 32-byte copy before each replay.  Diagnostics land in the reduction
// scratch and come back with one copy at the end.  Blast (bounding-box)
// injection needs a host round trip and is not graph-capturable.
namespace bd {
void plume_diag_from(const double *h, double out[4]);

// after the transport: the absorbed-power total, or the cleared fields of a
// laser-off epoch
static int enqueue_after_mc(Ctx *c, const bd_beam *beam, bool q_summed = false)
{
    if (beam->power > 0.0) {
        if (!q_summed) BD_TRY(reduce_masked_enqueue(c, c->Q, c->nvox, c->flags, F_ABL, true, c->red + 40));
    } else {
        BD_TRY(join_records(c));
        BD_TRY(launch_fill(c, c->Q, K_F32, c->nvox, 0.0));
        if (c->AL) BD_TRY(launch_fill(c, c->AL, K_F32, 2 * c->nvox, 0.0));
        BD_TRY(launch_fill(c, c->red + 16, K_F64, 8, 0.0));
        BD_TRY(launch_fill(c, c->red + 40, K_F64, 1, 0.0));
    }
    return BD_OK;
}

// everything after the absorbed power: sub-steps, plume, phase update, readback
static int enqueue_epoch_tail(Ctx *c, int n_substeps)
{
    const bd_physics &p = c->phys;
    const bool plume = c->plume.enabled != 0;
    const bool diff_plume = plume && c->plume.model == 0;
    BD_HIP(hipEventRecordWithFlags(c->gev[2], c->stream, hipEventRecordExternal));
    if (diff_plume) BD_TRY(launch_plume_void(c));
    int flags = BD_SUB_PLATEAU | BD_SUB_MASK_Q | BD_SUB_PEAK;
    if (p.has_arrhenius) flags |= BD_SUB_ARRHENIUS;
    if (p.has_char) flags |= BD_SUB_CHAR;
    BD_TRY(launch_fill(c, c->red + 8, K_F64, 1, 0.0));
    BD_HIP(hipEventRecordWithFlags(c->gev[3], c->stream, hipEventRecordExternal));
    BD_TRY(launch_substeps(c, n_substeps, 1, flags, true));
    BD_HIP(hipEventRecordWithFlags(c->gev[4], c->stream, hipEventRecordExternal));
    if (diff_plume) BD_TRY(launch_plume_substeps(c, n_substeps, p.dt_sub, plume_decay(c, p.dt_sub)));
    BD_HIP(hipEventRecordWithFlags(c->gev[5], c->stream, hipEventRecordExternal));
    BD_TRY(launch_phase_enqueue(c, 0.0, c->ep));
    if (plume) {
        BD_TRY(launch_plume_void(c));
        BD_TRY(launch_plume_inject(c));  // point injection: a no-op when nothing was removed
        BD_TRY(launch_plume_sync(c));
        BD_TRY(launch_plume_diag_enqueue(c));
    }
    BD_HIP(hipMemcpyAsync(c->host_pinned, c->red, 48 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipMemcpyAsync((char *)c->host_pinned + 48 * sizeof(double), c->counter, 2 * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, c->stream));
    return BD_OK;
}

static int enqueue_epoch(Ctx *c, const bd_beam *beam, int n_substeps)
{
    // path lengths inline by default (records only with BUGDEF_MC_DEFER_FLUENCE=1, their passes
    // then right after the transport on the context stream: with many contexts' captured epochs
    // in flight, a second stream per context cost more than it overlapped)
    if (beam->power > 0.0) BD_TRY(launch_mc_enqueue(c, beam, c->ep, 2, REC_MAIN, c->red + 40));
    BD_TRY(enqueue_after_mc(c, beam, true));
    return enqueue_epoch_tail(c, n_substeps);
}

static bool same_beam(const bd_beam &a, const bd_beam &b)
{
    return a.power == b.power && a.spot_dia == b.spot_dia && a.theta_rad == b.theta_rad && a.phi_rad == b.phi_rad &&
           a.x_center == b.x_center && a.y_center == b.y_center && a.z_center == b.z_center &&
           a.n_ambient == b.n_ambient && a.r_max == b.r_max && a.beam_profile == b.beam_profile &&
           a.launch_axis == b.launch_axis && a.n_photons == b.n_photons && a.photon_offset == b.photon_offset &&
           a.max_events == b.max_events && a.score_fluence == b.score_fluence && a.add_plume == b.add_plume;
}

}  // namespace bd

namespace bd {
// checks and buffers shared by the graph and gang epochs (no host round trip inside)
static int prepare_enqueued_epoch(Ctx *c, int n_substeps, const char *who)
{
    // the enqueued epoch's transport or cleared tally supersedes the L of a lagged launch
    BD_TRY(join_records(c));
    lag_drop(c);
    if (!c->has_phys) {
        set_error(std::string(who) + ": bd_set_physics has not been called");
        return BD_ESTATE;
    }
    if (c->plume.enabled && (c->plume.model != 0 || c->plume.blast_expansion)) {
        set_error(std::string(who) + ": the blast / separable plume needs host round trips; use bd_epoch");
        return BD_ESTATE;
    }
    if (n_substeps % 2) {
        set_error(std::string(who) + ": an odd sub-step count leaves the ping-pong buffers swapped; use bd_epoch");
        return BD_ESTATE;
    }
    if (c->graph_pending) {
        set_error(std::string(who) + ": the previous epoch was not finished");
        return BD_ESTATE;
    }
    BD_TRY(ensure_state(c));
    BD_TRY(ensure(c, (void **)&c->ep, sizeof(EpochParams)));
    BD_TRY(ensure(c, (void **)&c->Tpeak, (size_t)c->nvox * 4));
    BD_TRY(ensure(c, (void **)&c->to_remove, (size_t)c->nvox));
    if (c->phys.has_arrhenius) BD_TRY(ensure(c, (void **)&c->Om, (size_t)c->nvox * 4));
    return BD_OK;
}

// per-epoch scalars -> device, on the context stream (pinned staging in host_pinned)
static int upload_epoch_params(Ctx *c, int n_substeps, double t_before, int64_t epoch, uint64_t seed_base)
{
    EpochParams *hp = (EpochParams *)((char *)c->host_pinned + 56 * sizeof(double));
    hp->seed_base = seed_base;
    hp->epoch = (long long)epoch;
    hp->t = t_before;
    hp->dt_epoch = n_substeps * c->phys.dt_sub;
    BD_HIP(hipMemcpyAsync(c->ep, hp, sizeof(EpochParams), hipMemcpyHostToDevice, c->stream));
    return BD_OK;
}
}  // namespace bd

namespace bd {
static std::atomic<uint64_t> g_graph_seq{0};

// The context's captured epoch graph for (beam, n_substeps), captured again when the beam, the
// sub-step count, the ping-pong heads, physics / plume values or parameter buffers changed.
static int ensure_epoch_graph(Ctx *c, const bd_beam *beam, int n_substeps)
{
    const bool recapture = !c->graph_exec || c->graph_nsub != n_substeps || !same_beam(c->graph_beam, *beam) ||
                           c->graph_T != c->T || c->graph_P != c->P || c->graph_H != c->H ||
                           c->graph_F != c->flags || c->graph_stale;
    if (!recapture) return BD_OK;
    // allocate every lazily created buffer and tune the tiles before capture
    BD_TRY(launch_substeps(c, 0, 1, BD_SUB_PLATEAU, true));
    if (beam->power > 0.0) BD_TRY(launch_pack_optics(c, beam->add_plume != 0));
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    if (c->graph) {
        (void)hipGraphDestroy(c->graph);
        c->graph = nullptr;
    }
    c->graph_version = ++g_graph_seq;  // unique over all contexts, so a parent never reuses a stale child
    // buffers used inside the graph must exist before capture
    BD_TRY(ensure(c, (void **)&c->AL, (size_t)c->nvox * sizeof(float2)));
    BD_TRY(ensure(c, (void **)&c->optics, (size_t)c->nvox * sizeof(float4)));
    BD_TRY(ensure(c, (void **)&c->hlat, (size_t)c->nvox * 4));
    BD_TRY(ensure_records(c, beam, REC_MAIN));
    const int64_t nd = beam->launch_axis == 0 ? c->g.nz : c->g.nx;
    if (c->lat_n != nd) {
        if (c->lat) BD_HIP(hipFree(c->lat));
        c->lat = nullptr;
        BD_HIP(hipMalloc((void **)&c->lat, sizeof(double) * nd));
        c->lat_n = nd;
    }
    BD_HIP(hipStreamSynchronize(c->stream));
    c->graph_T = c->T;
    c->graph_P = c->P;
    c->graph_H = c->H;
    c->graph_F = c->flags;
    BD_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_epoch(c, beam, n_substeps);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != BD_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    BD_HIP(e);
    c->graph = g;
    BD_HIP(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    c->graph_beam = *beam;
    c->graph_nsub = n_substeps;
    c->graph_stale = false;
    return BD_OK;
}
}  // namespace bd

extern "C" int bd_epoch_graph_launch(bd_ctx *h, const bd_beam *beam, int n_substeps, double t_before,
                                     int64_t epoch, uint64_t seed_base)
{
    BD_CHECK(h && beam, "bd_epoch_graph: null argument");
    BD_CHECK(n_substeps >= 1, "bd_epoch_graph: n_substeps must be >= 1");
    Ctx *c = &h->c;
    BD_TRY(prepare_enqueued_epoch(c, n_substeps, "bd_epoch_graph"));
    BD_TRY(ensure_epoch_graph(c, beam, n_substeps));
    // per-epoch scalars -> device (outside the graph), replay; the readback is bd_epoch_graph_finish
    const double dt_epoch = n_substeps * c->phys.dt_sub;
    BD_HIP(hipEventRecord(c->ev2, c->stream));
    BD_TRY(upload_epoch_params(c, n_substeps, t_before, epoch, seed_base));
    BD_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    BD_HIP(hipEventRecord(c->ev3, c->stream));
    c->graph_pending = true;
    c->pending_beam = *beam;
    c->pending_dt_epoch = dt_epoch;
    return BD_OK;
}

// ---- one launch for an epoch of many contexts: their captured epoch graphs as independent
// child nodes of one parent graph, launched once on a library stream of the device.  The parent
// is rebuilt only when the set of contexts or one of their epoch graphs changed.  Library-wide
// per device, like the gang transport: batch calls on one device must not overlap in time.
// Measured on the 24-run campaign: the runtime runs the child graphs one after another (the
// sequential rate, half that of one stream per run); a graph captured across the contexts'
// streams with the epochs as parallel branches ran the same way.
namespace bd {
namespace {
struct BatchRes {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<const Ctx *, uint64_t>> key;  // (context, its graph's capture id) per child
};
BatchRes g_batch[16];
}  // namespace
}  // namespace bd

extern "C" int bd_epoch_batch_launch(bd_ctx *const *hs, int n, const bd_beam *beams, int n_substeps,
                                     const double *t_before, const int64_t *epochs, const uint64_t *seed_bases)
{
    using namespace bd;
    BD_CHECK(hs && beams && t_before && epochs && seed_bases, "bd_epoch_batch_launch: null argument");
    BD_CHECK(n >= 1 && n <= 4096, "bd_epoch_batch_launch: n must be in [1, 4096]");
    BD_CHECK(n_substeps >= 1, "bd_epoch_batch_launch: n_substeps must be >= 1");
    int dev = 0;
    BD_HIP(hipGetDevice(&dev));
    BD_CHECK(dev >= 0 && dev < 16, "bd_epoch_batch_launch: device index out of range");
    BatchRes &b = g_batch[dev];
    if (!b.stream) {
        BD_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
        BD_HIP(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
    }
    std::vector<std::pair<const Ctx *, uint64_t>> key;
    for (int i = 0; i < n; ++i) {
        BD_CHECK(hs[i], "bd_epoch_batch_launch: null context");
        for (int j = 0; j < i; ++j) BD_CHECK(hs[j] != hs[i], "bd_epoch_batch_launch: a context appears twice");
        Ctx *c = &hs[i]->c;
        BD_CHECK(c->device == dev, "bd_epoch_batch_launch: every context must be on the current device");
        BD_TRY(prepare_enqueued_epoch(c, n_substeps, "bd_epoch_batch_launch"));
        BD_TRY(ensure_epoch_graph(c, &beams[i], n_substeps));
        key.emplace_back(c, c->graph_version);
    }
    if (!b.exec || key != b.key) {
        if (b.exec) (void)hipGraphExecDestroy(b.exec);
        if (b.graph) (void)hipGraphDestroy(b.graph);
        b.exec = nullptr;
        b.graph = nullptr;
        b.key.clear();
        BD_HIP(hipGraphCreate(&b.graph, 0));
        for (int i = 0; i < n; ++i) {
            hipGraphNode_t node;
            BD_HIP(hipGraphAddChildGraphNode(&node, b.graph, nullptr, 0, hs[i]->c.graph));
        }
        BD_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
        b.key = key;
    }
    // each context's per-epoch scalars, queued on its own stream; the batch stream waits for them
    for (int i = 0; i < n; ++i) {
        Ctx *c = &hs[i]->c;
        BD_TRY(upload_epoch_params(c, n_substeps, t_before[i], epochs[i], seed_bases[i]));
        BD_HIP(hipEventRecord(c->ev_fork, c->stream));
        BD_HIP(hipStreamWaitEvent(b.stream, c->ev_fork, 0));
    }
    for (int i = 0; i < n; ++i) BD_HIP(hipEventRecord(hs[i]->c.ev2, b.stream));
    BD_HIP(hipGraphLaunch(b.exec, b.stream));
    BD_HIP(hipEventRecord(b.done, b.stream));
    // every context's stream continues after the launch (bd_epoch_graph_finish waits on it)
    for (int i = 0; i < n; ++i) {
        Ctx *c = &hs[i]->c;
        BD_HIP(hipStreamWaitEvent(c->stream, b.done, 0));
        BD_HIP(hipEventRecord(c->ev3, c->stream));
        c->graph_pending = true;
        c->pending_beam = beams[i];
        c->pending_dt_epoch = n_substeps * c->phys.dt_sub;
    }
    return BD_OK;
}

// One epoch of n contexts whose photon transport runs as a single launch
// (k_mc_gang); every other step is queued on each context's own stream, as in
// the graph epoch.  Collect each context with bd_epoch_graph_finish.
extern "C" int bd_epoch_gang_launch(bd_ctx *const *hs, int n, const bd_beam *beams, int n_substeps,
                                    const double *t_before, const int64_t *epochs, const uint64_t *seed_bases)
{
    BD_CHECK(hs && beams && t_before && epochs && seed_bases, "bd_epoch_gang_launch: null argument");
    BD_CHECK(n >= 1 && n <= 4096, "bd_epoch_gang_launch: n must be in [1, 4096]");
    BD_CHECK(n_substeps >= 1, "bd_epoch_gang_launch: n_substeps must be >= 1");
    std::vector<Ctx *> on;
    std::vector<const bd_beam *> on_beams;
    for (int i = 0; i < n; ++i) {
        BD_CHECK(hs[i], "bd_epoch_gang_launch: null context");
        for (int j = 0; j < i; ++j) BD_CHECK(hs[j] != hs[i], "bd_epoch_gang_launch: a context appears twice");
        Ctx *c = &hs[i]->c;
        BD_TRY(prepare_enqueued_epoch(c, n_substeps, "bd_epoch_gang_launch"));
        BD_TRY(launch_substeps(c, 0, 1, BD_SUB_PLATEAU, true));  // tile tuning (first epoch) before any enqueue
        BD_HIP(hipEventRecord(c->ev2, c->stream));
        BD_TRY(upload_epoch_params(c, n_substeps, t_before[i], epochs[i], seed_bases[i]));
        if (beams[i].power > 0.0) {
            on.push_back(c);
            on_beams.push_back(&beams[i]);
        }
    }
    BD_TRY(launch_mc_gang(on.data(), on_beams.data(), (int)on.size()));
    for (int i = 0; i < n; ++i) {
        Ctx *c = &hs[i]->c;
        BD_TRY(enqueue_after_mc(c, &beams[i]));
        BD_TRY(enqueue_epoch_tail(c, n_substeps));
        BD_HIP(hipEventRecord(c->ev3, c->stream));
        c->graph_pending = true;
        c->pending_beam = beams[i];
        c->pending_dt_epoch = n_substeps * c->phys.dt_sub;
    }
    return BD_OK;
}

// wait for a launched epoch graph and parse its diagnostics
extern "C" int bd_epoch_graph_finish(bd_ctx *h, bd_epoch_diag *out)
{
    BD_CHECK(h && out, "bd_epoch_graph_finish: null argument");
    Ctx *c = &h->c;
    if (!c->graph_pending) {
        set_error("bd_epoch_graph_finish: no epoch graph was launched");
        return BD_ESTATE;
    }
    c->graph_pending = false;
    const bd_beam *beam = &c->pending_beam;
    const double dt_epoch = c->pending_dt_epoch;
    BD_HIP(hipStreamSynchronize(c->stream));
    const double *r = (const double *)c->host_pinned;
    const unsigned long long *cnt = (const unsigned long long *)(r + 48);
    *out = bd_epoch_diag{};
    const double V = c->g.dx * c->g.dy * c->g.dz;
    if (beam->power > 0.0) {
        const double N = (double)beam->n_photons;
        const double *m = r + 16;
        out->absorbed_frac = m[4] / N;
        out->mc_energy_closure = (m[0] + m[1] + m[2] + m[3] + m[4]) / N;
        out->reflected_J = (m[0] + m[1]) / N * beam->power * dt_epoch;
        out->transmitted_J = m[2] / N * beam->power * dt_epoch;
        out->lateral_J = m[3] / N * beam->power * dt_epoch;
        out->R_spec = m[0] / N;
        out->R_diff = m[1] / N;
        out->transmission = m[2] / N;
        out->lateral_loss = m[3] / N;
        out->mc_events = m[5];
        out->mc_tally_atomics = m[6];
        out->mc_tally_records = m[7];
        out->absorbed_J = r[40] * V * dt_epoch;
    } else {
        out->mc_energy_closure = 1.0;
    }
    out->conducted_J = -r[8];
    out->ablated_J = r[9];
    out->n_new = (int64_t)cnt[1];
    if (c->plume.enabled) {
        double d[4];
        plume_diag_from(r + 24, d);
        out->plume_tau = d[0];
        out->plume_transmission = d[1];
        out->plume_peak_mu_a = d[2];
        out->plume_mean_mu_a = d[3];
    } else {
        out->plume_transmission = 1.0;
    }
    // phase spans from the external event nodes recorded inside the graph
    float ms = 0.f;
    BD_HIP(hipEventElapsedTime(&ms, c->ev2, c->gev[2]));
    out->t_optical_ms = ms;
    BD_HIP(hipEventElapsedTime(&ms, c->gev[2], c->gev[5]));
    out->t_thermal_ms = ms;
    BD_HIP(hipEventElapsedTime(&ms, c->gev[5], c->ev3));
    out->t_phase_ms = ms;
    BD_HIP(hipEventElapsedTime(&ms, c->gev[3], c->gev[4]));
    out->t_substeps_ms = ms;
    if (beam->power > 0.0) {
        BD_HIP(hipEventElapsedTime(&ms, c->gev[0], c->gev[1]));
        out->t_mc_kernel_ms = ms;
        BD_HIP(hipEventElapsedTime(&ms, c->gev[1], c->gev[6]));
        out->t_mc_records_ms = ms;
    }
    return BD_OK;
}

extern "C" int bd_epoch_graph(bd_ctx *h, const bd_beam *beam, int n_substeps, double t_before, int64_t epoch,
                              uint64_t seed_base, bd_epoch_diag *out)
{
    BD_CHECK(h && beam && out, "bd_epoch_graph: null argument");
    BD_TRY(bd_epoch_graph_launch(h, beam, n_substeps, t_before, epoch, seed_base));
    return bd_epoch_graph_finish(h, out);
}
