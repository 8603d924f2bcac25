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

#include <cmath>

namespace bd {

int reduce_masked(Ctx *c, const float *src, int64_t n, const uint8_t *flags, uint8_t bit, bool invert, double *out);

}  // namespace bd

using namespace bd;

extern "C" int bd_epoch(bd_ctx *h, const bd_beam *beam, int n_substeps, double t_after, bd_epoch_diag *out)
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
    hipEvent_t e_start = c->ev2, e_opt = c->ev3;

    // ---- 1. optics
    BD_HIP(hipEventRecord(e_start, c->stream));
    if (beam->power > 0.0) {
        bd_mc_result mc{};
        BD_TRY(launch_mc(c, beam, &mc));
        const double N = (double)beam->n_photons;
        out->absorbed_frac = mc.absorbed / N;
        out->mc_energy_closure = (mc.specular + mc.diffuse + mc.transmitted + mc.lateral + mc.absorbed) / N;
        out->reflected_J = (mc.specular + mc.diffuse) / N * beam->power * dt_epoch;
        out->transmitted_J = mc.transmitted / N * beam->power * dt_epoch;
        out->lateral_J = mc.lateral / N * beam->power * dt_epoch;
        out->R_spec = mc.specular / N;
        out->R_diff = mc.diffuse / N;
        out->transmission = mc.transmitted / N;
        out->lateral_loss = mc.lateral / N;
        out->mc_events = mc.events;
        out->t_mc_kernel_ms = mc.kernel_ms;
        double qsum = 0.0;
        BD_TRY(reduce_masked(c, c->Q, c->nvox, c->flags, F_ABL, true, &qsum));
        out->absorbed_J = qsum * V * dt_epoch;
    } else {
        BD_TRY(launch_fill(c, c->Q, K_F32, c->nvox, 0.0));
        if (c->AL) BD_TRY(launch_fill(c, c->AL, K_F32, 2 * c->nvox, 0.0));
        out->absorbed_frac = 0.0;
        out->mc_energy_closure = 1.0;
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
    if (diff_plume) {
        const double tl = c->plume.tau_linger;
        const double decay = (tl > 0.0 && std::isfinite(tl)) ? std::exp(-p.dt_sub / tl) : 1.0;
        BD_TRY(launch_plume_substeps(c, n_substeps, p.dt_sub, decay));
    }
    BD_HIP(hipEventRecord(c->ev0, c->stream));

    // ---- 3. phase update + removal
    int64_t n_new = 0;
    double abl_J = 0.0;
    BD_TRY(launch_phase(c, t_after, &n_new, &abl_J));
    out->n_new = n_new;
    out->ablated_J = abl_J;

    // ---- 4. plume
    if (plume) {
        BD_TRY(launch_plume_void(c));  // void after removal (coupled_ablation.py:588,658)
        if (n_new > 0) BD_TRY(launch_plume_inject(c));
        if (diff_plume) BD_TRY(launch_plume_sync(c));
        else BD_TRY(launch_plume_update_sep(c, dt_epoch));
        double d[4];
        BD_TRY(launch_plume_diag(c, d));
        out->plume_tau = d[0];
        out->plume_transmission = d[1];
        out->plume_peak_mu_a = d[2];
        out->plume_mean_mu_a = d[3];
    } else {
        out->plume_tau = 0.0;
        out->plume_transmission = 1.0;
    }
    BD_HIP(hipEventRecord(c->ev1, c->stream));
    BD_HIP(hipMemcpyAsync(c->host_pinned, c->red + 8, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BD_HIP(hipStreamSynchronize(c->stream));
    out->conducted_J = -*(double *)c->host_pinned;
    float ms = 0.f;
    BD_HIP(hipEventElapsedTime(&ms, e_start, e_opt));
    out->t_optical_ms = ms;
    BD_HIP(hipEventElapsedTime(&ms, e_opt, c->ev0));
    out->t_thermal_ms = ms;
    BD_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    out->t_phase_ms = ms;
    BD_HIP(hipEventElapsedTime(&ms, c->ev4, c->ev5));
    out->t_substeps_ms = ms;
    return BD_OK;
}
