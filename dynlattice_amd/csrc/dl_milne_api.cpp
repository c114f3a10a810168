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
    m.nblk = (int32_t)((N + 255) / 256);
    if ((rc = dalloc(c, &m.u1, 3 * N))) return rc;
    if ((rc = dalloc(c, &m.r1, 4 * N))) return rc;
    if ((rc = dalloc(c, &m.v1, 6 * N))) return rc;
    if ((rc = dalloc(c, &m.part, 2 * (size_t)m.nblk))) return rc;
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

// one attempt, kernels only (capturable)
static int launch_attempt(dl_ctx *c, const MilneCtl &mc) {
  const DevData &d = c->d;
  const MilneDev &m = c->milne.m;
  hipStream_t s = c->stream;
  const StepParams sp{0.0, 1, 0};
  // forces at the start state; contact pairs counted in the attempt's row
  if (c->contact) {
    int rc = contact_step(c->cs, d, s, c->hist_nc, c->hist_ncand, c->step_ctr);
    if (rc) return rc;
  }
  launch_strut_pp(d, sp, c->pp, MODE_FORCE, c->so3, c->mat_kind, c->contact, s);
  launch_joint(d, sp, MODE_FORCE, c->so3, s);
  launch_milne_pred(d, m, c->so3, s);
  // forces at the predicted state
  DevData dp = d;
  dp.u = m.u1;
  dp.rot = m.r1;
  dp.v = m.v1;
  if (c->contact) {
    int rc = contact_step(c->cs, dp, s, nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  launch_strut_pp(dp, sp, c->pp, MODE_FORCE, c->so3, c->mat_kind, c->contact, s);
  launch_joint(dp, sp, MODE_FORCE, c->so3, s);
  launch_milne_corr(d, m, mc.len_scale, c->so3, s);
  launch_milne_ctl(m, mc, c->step_ctr, s);
  launch_milne_commit(d, m, c->so3, s);
  // the material commits the predicted-state update (the last evaluation)
  if (c->plastic) launch_commit_pl(d, true, nullptr, nullptr, s, m.acc);
  launch_step_end(d, c->rgptr, c->n_groups, false, 0, nullptr, c->hist_react, c->hist_diss,
                  c->step_ctr, s);
  return DL_OK;
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
      for (int64_t k = 0; k < K && !rc; k++) rc = launch_attempt(c, mc);
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
    int rc = launch_attempt(c, mc);
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
  if ((rc = milne_attempts(c, n_attempts, mc))) return rc;
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
