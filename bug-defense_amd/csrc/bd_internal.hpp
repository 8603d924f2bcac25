// bd_internal.hpp -- shared definitions of the gfx950 implementation behind bugdef.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <string>
#include <utility>

#include "../../include/bugdef.h"
#include "../../include/bugdef_fno.h"

namespace bd {

// Flags byte per voxel: ablated / dry / "banked enthalpy > 0".  Keeping the
// three state bits in one byte lets the sub-step read one byte instead of two
// masks and skip the H_acc read for every cell that has nothing banked.
constexpr uint8_t F_ABL = 1;
constexpr uint8_t F_DRY = 2;
constexpr uint8_t F_HPOS = 4;

// Per-epoch scalars kept on the device so a captured epoch graph can replay
// without host updates: the MC key is SplitMix64(seed_base + (epoch+1) * gamma)
// and the removal time is t + dt_epoch, advanced by a kernel at epoch end with
// the same FP64 addition the host engine does (coupled_ablation.py:548).
struct EpochParams {
    unsigned long long seed_base;
    long long epoch;
    double t;
    double dt_epoch;
};

__host__ __device__ inline unsigned long long mix_seed(unsigned long long seed, long long epoch)
{
    unsigned long long z = seed + (unsigned long long)(epoch + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Device buffers of a context.  Unused buffers stay null until first use.
struct Field {
    void *ptr = nullptr;
    size_t bytes = 0;
    int64_t n = 0;   // element count (1 for a broadcast scalar parameter)
};

struct Ctx {
    bd_grid g{};
    int64_t nvox = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr, ev4 = nullptr, ev5 = nullptr;
    hipEvent_t ev6 = nullptr, ev7 = nullptr;  // end / start of the path-length record passes
    hipEvent_t ev8 = nullptr, ev9 = nullptr;  // eager epoch: end of the thermal phase, end of the epoch
    // an eager epoch runs the previous epoch's record passes on `side`, beside its transport;
    // join_records() makes `stream` wait.  Created with the lag buffers, on first use: with
    // few hardware queues per process (4 by default) a stream per context that most contexts
    // never use takes queues from the context streams of concurrent runs.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool join_pending = false;
    // recorded as external event nodes inside the captured epoch graph: MC kernel start/end,
    // optics end, sub-steps start/end, epoch end, end of the path-length record pass
    hipEvent_t gev[7] = {};
    bd_physics phys{};
    bool has_phys = false;
    bd_plume_params plume{};
    int64_t pvox = 0;

    // state (f32 unless noted)
    float *T = nullptr, *T2 = nullptr;   // ping-pong
    float *H = nullptr, *Om = nullptr, *chi = nullptr, *Q = nullptr;
    float *k = nullptr, *rc = nullptr, *hlat = nullptr;
    uint8_t *flags = nullptr;            // F_ABL | F_DRY | F_HPOS
    float *mu_a = nullptr, *mu_s = nullptr, *gg = nullptr, *nn = nullptr;
    float *wet_mu_a = nullptr, *wet_k = nullptr, *wet_rc = nullptr;
    float *Tpeak = nullptr;
    double *abl_time = nullptr;
    float2 *AL = nullptr;                // interleaved absorbed / path tallies
    float4 *optics = nullptr;            // packed transport optics {mu_a, mu_s, g, n}
    float *Qs = nullptr;                 // surrogate coupling: Q of the last MC epoch
    float *Qsur = nullptr;               // surrogate coupling: composed field for judging
    float *Qaux = nullptr;               // surrogate coupling: a second MC field (coarse amplitudes, MC-vs-MC)
    float *P = nullptr, *P2 = nullptr, *Pext = nullptr;  // plume grid
    uint8_t *pvoid = nullptr;
    uint8_t *to_remove = nullptr;
    Field ds[7];                         // dry state: mu_a mu_s g n k rho_cp t_remove
    Field cs[5];                         // char state: mu_a mu_s k rho_cp t_remove
    float *kb[3] = {nullptr, nullptr, nullptr};  // blast splat taps
    float *kl[3] = {nullptr, nullptr, nullptr};  // legacy splat taps
    int nkb[3] = {0, 0, 0}, rb[3] = {0, 0, 0};
    int nkl[3] = {0, 0, 0}, rl[3] = {0, 0, 0};
    float *splat_tmp1 = nullptr, *splat_tmp2 = nullptr;
    size_t splat_tmp_bytes = 0;

    // scratch
    // reduction accumulators (64 doubles) + 8192 doubles of partials from red + 64: [8] conducted
    // energy (sub-step ledger sum), [9] ablated energy (phase), [16, 24) the transport's scalar
    // tallies, [24, 28) plume diagnostics, [40] masked reductions, [48, 54) surrogate error terms
    double *red = nullptr;
    double *lat = nullptr;               // lateral escape per depth
    int64_t lat_n = 0;
    // COUNTER_SLOTS 64-bit slots: [0] photon work counter, [1] n_new, [2] breach, [3] and [4] the
    // record cursors of the two record buffers, [8] surrogate window bounds (two ints), [10] and
    // [11] the legacy plume update's flags, [16, 22) the plume splat's bounding box
    unsigned long long *counter = nullptr;
    uint2 *rec = nullptr;                // deferred path-length records of the transport
    uint2 *rec_sorted = nullptr;         // the records in tile order (record passes)
    int64_t rec_cap = 0;                 // records each of the two buffers holds
    unsigned *rec_tcnt = nullptr, *rec_tstart = nullptr, *rec_off = nullptr;  // per-tile counts, starts, slice offsets
    unsigned *rec_cstart = nullptr, *rec_cmap = nullptr;  // per-tile first chunk, chunk -> tile
    // one-epoch lag of the record passes in eager epochs: the records of an epoch's transport
    // wait in record buffer `lag` (0: rec, cursor counter[3]; 1: rec_b, counter[4]), their inline
    // path lengths in Lb[lag], until the next eager epoch runs their passes beside its own
    // transport (the L they give is superseded by that epoch's), or a read of the path-length
    // tally runs them and copies Lb[lag] into AL (lag_finalize)
    uint2 *rec_b = nullptr;
    float *Lb[2] = {nullptr, nullptr};
    int lag = -1;
    // path lengths to records: 1 (default) in every launch (eager epochs run their passes beside
    // the next transport, other launches right after the kernel); 2 only in eager epochs; 0 never
    // (BUGDEF_MC_DEFER_FLUENCE=1 / 2 / 0)
    int defer_L = 1;
    int rec_per_photon = 512;            // record buffer size per photon (BUGDEF_MC_RECORDS_PER_PHOTON)
    double *led_part = nullptr;          // sub-step ledger slots, zero between calls
    bool plume_pairs = true;             // paired plume diffusion rows where they fit (BUGDEF_PLUME_PAIRS)
    void *staging = nullptr;             // f64 staging for conversions
    size_t staging_bytes = 0;
    void *host_pinned = nullptr;         // small pinned host block (64 doubles)

    // persistent MC grid: blocks of 256 per CU (0: one lane per photon up to 8 blocks per CU)
    int mc_blocks_per_cu = 0;

    // tuned launch shape of the sub-step kernel, and of the two-sub-step kernel
    int tile_by = 4, tile_xc = 16;
    int tile2_by = 16, tile2_xc = 32;
    bool use_two = false;
    bool two_step_allowed = true;      // BUGDEF_SUBSTEP_PAIRS=0 keeps every sub-step in its own launch
    bool two_step_forced = false;      // BUGDEF_SUBSTEP_PAIRS=2 fuses pairs whatever the timing pass finds
    bool tune_log = false;             // BUGDEF_TUNE_LOG=1 prints the timing pass's table to stderr
    float *Hb = nullptr;               // second H_acc buffer (the two-sub-step kernel writes the other one)
    uint8_t *flagsb = nullptr;         // second flags buffer
    bool tuned = false;
    bool tile_pinned = false;  // one-step tile fixed by BUGDEF_SUBSTEP_TILE: no timing pass
    int64_t sub_launched = 0;  // sub-steps launched so far (eager calls)
    bool retuned = false;      // the second timing pass, on a developed state, has run
    double last_ledger_conducted = 0.0;

    // captured epoch graph (bd_epoch_graph) and the beam it was captured for
    EpochParams *ep = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    bd_beam graph_beam{};
    int graph_nsub = -1;
    const float *graph_T = nullptr, *graph_P = nullptr;  // ping-pong heads the graph was captured with
    const float *graph_H = nullptr;
    const uint8_t *graph_F = nullptr;
    bool graph_pending = false;          // launched by bd_epoch_graph_launch, not yet finished
    bool graph_stale = false;            // physics/plume values or parameter buffers changed since capture
    uint64_t graph_version = 0;          // capture id, unique over all contexts (parents holding the graph rebuild)
    bd_beam pending_beam{};
    double pending_dt_epoch = 0.0;
};

void set_error(const std::string &msg);

#define BD_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t _e = (call);                                                          \
        if (_e != hipSuccess) {                                                          \
            ::bd::set_error(std::string(#call) + ": " + hipGetErrorString(_e));          \
            return BD_EHIP;                                                              \
        }                                                                                \
    } while (0)

#define BD_CHECK(cond, msg)                                                              \
    do {                                                                                 \
        if (!(cond)) {                                                                   \
            ::bd::set_error(msg);                                                        \
            return BD_EINVAL;                                                            \
        }                                                                                \
    } while (0)

#define BD_TRY(expr)                                                                     \
    do {                                                                                 \
        int _s = (expr);                                                                 \
        if (_s != BD_OK) return _s;                                                      \
    } while (0)

int ensure(Ctx *c, void **p, size_t bytes);  // lazily allocate (and zero) a buffer
int ensure_state(Ctx *c);                     // T, T2, H, chi, Q, k, rc, flags, ...
int ensure_staging(Ctx *c, size_t bytes);
int ensure_records(Ctx *c, const bd_beam *b, int rec_mode);  // record buffers, when such a launch defers

// launch helpers implemented in the kernel translation units
int launch_substeps(Ctx *c, int n, int is_active, int flags, bool ledger);
int launch_plume_substeps(Ctx *c, int n, double dt, double decay);
// exp(-dt / tau_linger), 1 without lingering decay
double plume_decay(const Ctx *c, double dt);
int launch_plateau(Ctx *c);
int launch_arrhenius(Ctx *c, double A, double EaR, double dt);
int launch_char(Ctx *c, double A, double EaR, double dt);
// where the record passes of a transport launch run: after it on the context stream, or at the
// next eager epoch, on the side stream beside that epoch's transport
enum RecMode { REC_MAIN = 0, REC_LAG = 2 };
constexpr int COUNTER_SLOTS = 32;
constexpr int COUNTER_BBOX = 16;  // six slots
int launch_mc(Ctx *c, const bd_beam *b, bd_mc_result *out);  // bd_mc_run: passes after the kernel
int join_records(Ctx *c);   // the context stream waits for record passes left on the side stream
int lag_finalize(Ctx *c);   // the lagged record passes (if any) run and their L lands in AL
void al_get(Ctx *c, float *d, int which);        // one half of AL (0: A, 1: L) into a float grid
void al_set(Ctx *c, const float *s, int which);  // a float grid into one half of AL
inline void lag_drop(Ctx *c) { c->lag = -1; }  // a new transport or a cleared tally supersedes them
// enqueue only (no host wait); with ep != null the photon key comes from the device params
// timed: 0 no events, 1 ev0/ev1 around the transport kernel, 2 graph-external gev[0]/gev[1]
// q_sum: when set, the conversion of the absorbed weight to Q also sums Q over the intact cells
// (flags without F_ABL) into *q_sum on the device, in the order of reduce_masked_enqueue
int launch_mc_enqueue(Ctx *c, const bd_beam *b, const EpochParams *ep, int timed, int rec_mode = REC_MAIN,
                      double *q_sum = nullptr);
// Q = A * scale and the masked sum of Q into *dev_out (one pass over the tally)
int q_from_a_sum_enqueue(Ctx *c, float scale, const uint8_t *mask, uint8_t bits, bool invert, double *dev_out);
// One transport launch for n contexts (each with power on, EpochParams in c->ep):
// prepares each on its own stream, traces all photons in one kernel on a
// library stream, then queues each context's absorbed-power conversion on its
// stream after it.
int launch_mc_gang(Ctx *const *cs, const bd_beam *const *beams, int n);
int launch_phase(Ctx *c, double t_now, int64_t *n_new, double *ablated_J);
int launch_phase_enqueue(Ctx *c, double t_now, const EpochParams *ep);
int launch_plume_diag_enqueue(Ctx *c);
void plume_diag_from(const double *h, double out[4]);  // tau, transmission, peak, mean from the accumulators
int reduce_masked_enqueue(Ctx *c, const float *src, int64_t n, const uint8_t *flags, uint8_t bit, bool invert,
                          double *dev_out);
int launch_plume_void(Ctx *c);
int launch_plume_inject(Ctx *c);
int launch_plume_sync(Ctx *c);
int launch_plume_update_sep(Ctx *c, double dt_epoch);
int launch_plume_diag(Ctx *c, double out[4]);
int launch_plume_splat_box(Ctx *c, int which, const float *host_box, int Lx, int Ly, int Lz, int x0, int y0, int z0,
                           double delta);
int launch_pack_optics(Ctx *c, bool add_plume);
int launch_reduce(Ctx *c, const void *src, int kind, int64_t n, int op, const uint8_t *mask, uint8_t mask_bits,
                  double *out);
int launch_fill(Ctx *c, void *dst, int kind, int64_t n, double v);
// zeroes up to four ranges of 8-byte words (doubles or 64-bit counters) in one launch
int launch_zero_f64(Ctx *c, std::initializer_list<std::pair<void *, int64_t>> ranges);
int launch_phase_map(Ctx *c, const uint8_t *air_gap_dev, uint8_t *out_dev);
int launch_slice(Ctx *c, const void *src, int kind, int plane, int index, int stride, float *out, int64_t n);
int launch_slice_plume(Ctx *c, const float *src, int plane, int index, int stride, float *out, int64_t n);
int launch_phase_slice(Ctx *c, const uint8_t *gap_z_dev, int plane, int index, int stride, uint8_t *out, int64_t n);

// element kinds of device buffers
constexpr int K_F32 = 0, K_F64 = 1, K_U8 = 2, K_FLAG_ABL = 3, K_FLAG_DRY = 4, K_AL_A = 5, K_AL_L = 6;

inline int blocks_for(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    return (int)(b < 1 ? 1 : (b > 1048576 ? 1048576 : b));
}

}  // namespace bd

struct bd_ctx {
    bd::Ctx c;
};
