// dl_milne_api.cpp -- the MilneDevice integrator behind dl_milne_step
// (include/dl_gpu.h): attempt launch sequence, buffers and the captured
// attempt graph.  The kernels are in dl_milne.hip; the force evaluations are
// the strut/joint kernels in force mode (dl_kernels.hip) and the contact
// step (dl_contact.hip), run once on the start state and once on the
// predicted state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "dl_ctx.h"

static int milne_buffers(dl_ctx *c, int64_t n_attempts) {
  MilneState &M = c->milne;
  const int64_t N = c->d.N;
  int rc;
  if (!M.ready) {
    MilneDev &m = M.m;
    // partial sums: one per node block (node kernels), or one per strut
    // block plus one per joint block (fused corrector)
    const int64_t nb_node = (N + 255) / 256;
    const int64_t nb_fused = (c->d.NL + 255) / 256 + (c->d.NJ + 255) / 256;
    M.part_cap = std::max<int64_t>(nb_node, nb_fused);
    m.u0 = c->d.u;
    m.r0 = c->d.rot;
    m.v0 = c->d.v;
    if ((rc = dalloc(c, &m.u1, 3 * N))) return rc;
    if ((rc = dalloc(c, &m.r1, 4 * N))) return rc;
    if ((rc = dalloc(c, &m.v1, 6 * N))) return rc;
    if ((rc = dalloc(c, &m.part, 2 * (size_t)M.part_cap))) return rc;
    if ((rc = dalloc(c, &m.dt, 1))) return rc;
    if ((rc = dalloc(c, &m.tacc, 1))) return rc;
    if ((rc = dalloc(c, &m.acc, 1))) return rc;
    if ((rc = dalloc(c, &m.nacc, 1))) return rc;
    M.ready = true;
  }
  if (n_attempts > M.log_cap) {
    dfree(c, M.m.log);
    M.m.log = nullptr;
    M.log_cap = 0;
    const int64_t cap = std::max<int64_t>(n_attempts, 16);
    if ((rc = dalloc(c, &M.m.log, 4 * cap))) return rc;
    M.log_cap = cap;
    c->gen++; // the log pointer is captured in attempt graphs
  }
  return DL_OK;
}

// events per timed attempt: start, after contact/strut/joint/predictor of
// the start-state evaluation, after contact/strut/joint of the predicted one, end
#define MILNE_EVENTS 9

// one attempt, kernels only (capturable); ev: MILNE_EVENTS events or null.
// The predictor and corrector run inside the strut / joint kernels of the
// two force evaluations (the nodal force stays in registers); with ImpactBC
// groups, whose accelerations need the group sums first, they run as node
// kernels over fout.
static int launch_attempt(dl_ctx *c, const MilneCtl &mc, hipEvent_t *ev) {
  const DevData &d = c->d;
  MilneDev m = c->milne.m;
  hipStream_t s = c->stream;
  const StepParams sp{0.0, 1, 0};
  const bool fused = c->imp.ngroups == 0;
  m.ls = mc.len_scale;
  if (fused) {
    m.jofs = (int32_t)((d.NL + 255) / 256);
    m.nblk = m.jofs + (int32_t)((d.NJ + 255) / 256);
  } else {
    m.jofs = 0;
    m.nblk = (int32_t)((d.N + 255) / 256);
  }
  if (ev) HIPCHK(hipEventRecord(ev[0], s));
  // updateWhen: the mass of the attempt, at its start state
  launch_mass_update(d, c->mu, c->so3, s);
  // boundary values at the end time of the attempt; the clock moves on accept
  if (c->pd.nprog > 0 || c->pd.clock_on) {
    launch_bc_program(c->pd, 0.0, m.dt, s);
    m.clock_t = c->pd.clock_t;
    m.clock_k = c->pd.clock_k;
  }
  // forces at the start state; contact pairs counted in the attempt's row
  if (c->contact) {
    int rc = contact_step(c->cs, d, s, c->hist_nc, c->hist_ncand, c->step_ctr);
    if (rc) return rc;
  }
  // the previous attempt's GroupOutput row, at this attempt's start state
  // (timed with the contact phase)
  launch_react(d, c->so3, c->plastic, c->contact, s);
  if (ev) HIPCHK(hipEventRecord(ev[1], s));
  launch_strut_pp(d, sp, c->pp, fused ? MODE_MPRED : MODE_FORCE, c->so3, c->mat_kind, c->contact,
                  s, &m);
  if (ev) HIPCHK(hipEventRecord(ev[2], s));
  launch_joint(d, sp, fused ? MODE_MPRED : MODE_FORCE, c->so3, s, &m);
  if (!fused) {
    // ImpactBC: the groups' reactions at the start state set their accelerations
    launch_impact_sum(c->imp, d, s);
    m.imp_ready = c->imp.ready;
    m.imp_dof = c->imp.dof;
  }
  if (ev) HIPCHK(hipEventRecord(ev[3], s));
  if (!fused) launch_milne_pred(d, m, c->so3, s);
  if (ev) HIPCHK(hipEventRecord(ev[4], s));
  // forces at the predicted state
  DevData dp = d;
  dp.u = m.u1;
  dp.rot = m.r1;
  dp.v = m.v1;
  if (c->contact) {
    int rc = contact_step(c->cs, dp, s, nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  if (ev) HIPCHK(hipEventRecord(ev[5], s));
  launch_strut_pp(dp, sp, c->pp, fused ? MODE_MCORR : MODE_FORCE, c->so3, c->mat_kind, c->contact,
                  s, &m);
  if (ev) HIPCHK(hipEventRecord(ev[6], s));
  launch_joint(dp, sp, fused ? MODE_MCORR : MODE_FORCE, c->so3, s, &m);
  if (ev) HIPCHK(hipEventRecord(ev[7], s));
  if (!fused) launch_milne_corr(d, m, c->so3, s);
  launch_milne_ctl(m, mc, c->step_ctr, s);
  launch_milne_restore(d, m, c->so3, s);
  // the material commits the predicted-state update (the last evaluation);
  // the elastic material keeps no state that anything reads back
  if (c->plastic) launch_milne_commit_pl(d, m, s);
  launch_step_end(d, c->rgptr, c->n_groups, false, 0, nullptr, c->hist_react, c->hist_diss,
                  c->step_ctr, s);
  if (ev) HIPCHK(hipEventRecord(ev[8], s));
  return DL_OK;
}

// attempts with event timing: phase sums over both evaluations per attempt
// ("node": predictor, corrector, control, commit)
static int timed_attempts(dl_ctx *c, int64_t n, const MilneCtl &mc) {
  std::vector<hipEvent_t> ev(MILNE_EVENTS * (size_t)n);
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  int rc = DL_OK;
  for (int64_t k = 0; k < n && !rc; k++) rc = launch_attempt(c, mc, &ev[MILNE_EVENTS * k]);
  if (!rc) {
    HIPCHK(hipStreamSynchronize(c->stream));
    c->t_step = c->t_contact = c->t_strut = c->t_joint = c->t_node = 0.0;
    for (int64_t k = 0; k < n; k++) {
      const hipEvent_t *e = &ev[MILNE_EVENTS * k];
      float t[MILNE_EVENTS - 1];
      for (int i = 0; i + 1 < MILNE_EVENTS; i++) HIPCHK(hipEventElapsedTime(&t[i], e[i], e[i + 1]));
      c->t_contact += t[0] + t[4];
      c->t_strut += t[1] + t[5];
      c->t_joint += t[2] + t[6];
      c->t_node += t[3] + t[7];
      for (int i = 0; i + 1 < MILNE_EVENTS; i++) c->t_step += t[i];
    }
    c->t_count = n;
  }
  for (auto &e : ev) hipEventDestroy(e);
  return rc;
}

static bool same_ctl(const MilneCtl &a, const MilneCtl &b) {
  return std::memcmp(&a, &b, sizeof(MilneCtl)) == 0;
}

static int milne_attempts(dl_ctx *c, int64_t n, const MilneCtl &mc) {
  MilneState &M = c->milne;
  const int64_t K = c->graph_steps;
  int64_t done = 0;
  if (K > 0 && n >= K) {
    if (!M.gexec || M.graph_gen != c->gen || !same_ctl(M.graph_ctl, mc)) {
      if (M.gexec) {
        hipGraphExecDestroy(M.gexec);
        M.gexec = nullptr;
      }
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = DL_OK;
      for (int64_t k = 0; k < K && !rc; k++) rc = launch_attempt(c, mc, nullptr);
      hipError_t ec = hipStreamEndCapture(c->stream, &g);
      if (rc) return rc;
      if (ec != hipSuccess)
        return fail(DL_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
      ec = hipGraphInstantiate(&M.gexec, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (ec != hipSuccess)
        return fail(DL_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ec));
      M.graph_gen = c->gen;
      M.graph_ctl = mc;
    }
    for (; done + K <= n; done += K) HIPCHK(hipGraphLaunch(M.gexec, c->stream));
  }
  for (; done < n; done++) {
    int rc = launch_attempt(c, mc, nullptr);
    if (rc) return rc;
  }
  return DL_OK;
}

extern "C" int dl_milne_step(dl_ctx *c, int64_t n_attempts, const dl_milne_opts *o,
                             double *attempts, dl_milne_out *out) {
  if (!c) return fail(DL_ERR_INPUT, "NULL context");
  if (!c->initialized) return fail(DL_ERR_STATE, "MilneDevice attempts before INIT");
  if (!o) return fail(DL_ERR_INPUT, "NULL options");
  if (n_attempts < 0) return fail(DL_ERR_INPUT, "n_attempts >= 0 required");
  if (!(o->dt > 0.0) || !(o->precision > 0.0) || !(o->len_scale > 0.0) || !(o->dt_min >= 0.0) ||
      !(o->dt_max >= o->dt_min) || !(o->safety > 0.0) || !(o->incr > 0.0) || !(o->decr > 0.0))
    return fail(DL_ERR_INPUT, "dt, precision, len_scale, safety, incr, decr > 0 and "
                              "0 <= dt_min <= dt_max required");
  if (c->d.ndt) return fail(DL_ERR_STATE, "MilneDevice with per-node time steps");
  hipSetDevice(c->device);
  c->adj.valid = false;
  int rc;
  if ((rc = milne_buffers(c, n_attempts))) return rc;
  if ((rc = prepare_history(c, n_attempts))) return rc;
  MilneDev &m = c->milne.m;
  launch_set_f64(m.dt, o->dt, c->stream);
  launch_set_f64(m.tacc, 0.0, c->stream);
  HIPCHK(hipMemsetAsync(m.nacc, 0, sizeof(int64_t), c->stream));
  const MilneCtl mc{o->dt_min, o->dt_max, o->precision, o->safety, o->incr, o->decr,
                    o->len_scale, 6 * c->d.N};
  if ((rc = c->timing ? timed_attempts(c, n_attempts, mc) : milne_attempts(c, n_attempts, mc)))
    return rc;
  if (n_attempts > 0 && (rc = launch_reaction_tail(c))) return rc;
  HIPCHK(hipGetLastError());
  c->hist_n = n_attempts;
  if ((rc = check_err(c))) return rc;
  if (attempts && n_attempts > 0)
    HIPCHK(hipMemcpy(attempts, m.log, 4 * n_attempts * sizeof(double), hipMemcpyDeviceToHost));
  if (out) {
    out->attempts = n_attempts;
    HIPCHK(hipMemcpy(&out->accepted, m.nacc, sizeof(int64_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&out->time, m.tacc, sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&out->dt_next, m.dt, sizeof(double), hipMemcpyDeviceToHost));
  }
  return DL_OK;
}
