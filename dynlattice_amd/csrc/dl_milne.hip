// dl_milne.hip -- gfx950 kernels of the MilneDevice predictor-corrector
// integrator with adaptive step size (FP64).
//
// One attempt (MilneDeviceModule.cpp:41-112 with ExplicitModule::commit,
// ExplicitModule.cpp:290-335) is, on one stream and without a host round
// trip:
//   forces at the start state (u, v)        k_strut/k_joint in force mode
//   k_milne_pred   a = M(u)^-1 (f_ext - f - gyro); v_pre = v + dt a;
//                  u_pre = u (+) dt v            -> (u1, r1, v1)
//   forces at (u_pre, v_pre)                 k_strut/k_joint on (u1, r1)
//   k_milne_corr   a_pre with the start state's mass; v_new = v + dt a_pre;
//                  u_new = u (+) dt v_pre        -> (u, rot, v), and the
//                  start state -> (u1, r1, v1); per-block sums of the
//                  squared differences pre - new
//   k_milne_ctl    error, accept or reject, next step size (one block)
//   k_milne_restore  on reject the start state comes back from (u1, r1, v1)
//   k_milne_commit_pl  on accept the material commits the predicted-state
//                  evaluation
// Accepted attempts (the usual case) thus move no state beyond the
// corrector's own writes.  The step size lives in device memory, so a
// sequence of attempts can be replayed from a captured graph.
#include <hip/hip_runtime.h>

#include "dl_internal.h"
#include "dl_milne_node.h"

#define DL_BLOCK 256

namespace {

template <bool SO3>
__global__ __launch_bounds__(DL_BLOCK) void k_milne_pred(DevData d, MilneDev m) {
  const int64_t n = (int64_t)blockIdx.x * DL_BLOCK + threadIdx.x;
  if (n >= d.N) return;
  const int64_t N = d.N;
  double f[6];
#pragma unroll
  for (int c = 0; c < 6; c++) f[c] = d.fout[c * N + n];
  milne_pred_node<SO3>(d, m, n, f);
}

template <bool SO3>
__global__ __launch_bounds__(DL_BLOCK) void k_milne_corr(DevData d, MilneDev m) {
  const int64_t n = (int64_t)blockIdx.x * DL_BLOCK + threadIdx.x;
  const int64_t N = d.N;
  double e[2] = {0.0, 0.0};
  if (n < N) {
    double f[6];
#pragma unroll
    for (int c = 0; c < 6; c++) f[c] = d.fout[c * N + n];
    milne_corr_node<SO3>(d, m, n, f, e);
  }
  milne_block_partials(m, e, blockIdx.x);
}

// error = (rms(u_pre - u_new) + dt rms(v_pre - v_new)) / 2 over all dofs
// (getQuality, ExplicitModule.cpp:490-499; MilneDeviceModule.cpp:105-110),
// then the step-size rule of ExplicitModule::commit (order 1)
__global__ __launch_bounds__(1024) void k_milne_ctl(MilneDev m, MilneCtl mc, const int32_t *step) {
  __shared__ double red[2][1024];
  double su = 0.0, sv = 0.0;
  for (int32_t i = threadIdx.x; i < m.nblk; i += 1024) {
    su += m.part[i];
    sv += m.part[m.nblk + i];
  }
  red[0][threadIdx.x] = su;
  red[1][threadIdx.x] = sv;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double dt = *m.dt;
  const double sn = sqrt((double)mc.ndof);
  double err = 0.0;
  err += sqrt(red[0][0]) / sn;
  err += (sqrt(red[1][0]) / sn) * dt;
  err = 0.5 * err;
  const double dt_opt = dt * pow(mc.prec / err, 0.5);
  const bool accept = err <= mc.prec || dt <= mc.dt_min;
  double t = mc.safety * dt_opt;
  if (accept) {
    if (mc.incr * dt < t) t = mc.incr * dt;
    if (mc.dt_max < t) t = mc.dt_max;
  }
  if (t < mc.decr * dt) t = mc.decr * dt;
  if (t < mc.dt_min) t = mc.dt_min;
  const int64_t k = *step;
  double *row = m.log + 4 * k;
  row[0] = dt;
  row[1] = err;
  row[2] = accept ? 1.0 : 0.0;
  row[3] = t;
  *m.acc = accept ? 1 : 0;
  if (accept) {
    *m.tacc += dt;
    *m.nacc += 1;
    if (m.clock_t) {
      *m.clock_t += dt;
      *m.clock_k += 1;
    }
    if (m.imp_ready) *m.imp_ready = 1;
  }
  *m.dt = t;
}

// rejected attempt: the start state comes back (the next attempt
// evaluates contact there, so the list trigger checks it)
template <bool SO3>
__global__ __launch_bounds__(DL_BLOCK) void k_milne_restore(DevData d, MilneDev m) {
  if (*m.acc) return;
  const int64_t n = (int64_t)blockIdx.x * DL_BLOCK + threadIdx.x;
  if (n >= d.N) return;
  const int64_t N = d.N;
  double u[3], r[4], v[6];
  Quat q;
  load_node<SO3>(m.u1, m.r1, m.v1, N, n, u, r, v, q);
  store_node<SO3>(d.u, d.rot, d.v, N, n, u, r, v);
  list_trigger(d, n, u);
}

// accepted attempt: applyDeform of the predicted-state evaluation (the last
// one).  The trial plastic state was written only where the return map
// corrected it (active); the committed total strain follows the trial one.
__global__ __launch_bounds__(DL_BLOCK) void k_milne_commit_pl(DevData d, const int32_t *acc) {
  if (!*acc) return;
  const int64_t e = (int64_t)blockIdx.x * DL_BLOCK + threadIdx.x;
  if (e >= d.E) return;
  const int64_t E = d.E;
  double ec[6];
  rec6_load(d.ecur, e, ec);
  rec6_store(d.eo, e, ec);
  if (d.active[e]) {
    double ep[6], al[7];
    pst_load(d.pst_cur, d.aiso_cur, e, true, ep, al);
    pst_store(d.pst, d.aiso, e, true, ep, al);
    const uint8_t fl = d.eflag[e];
    if (!(fl & PST_FLAG)) d.eflag[e] = (uint8_t)(fl | PST_FLAG);
  }
}

__global__ void k_set_f64(double *p, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + DL_BLOCK - 1) / DL_BLOCK); }

} // namespace

void launch_milne_pred(const DevData &d, const MilneDev &m, bool so3, hipStream_t s) {
  if (so3) hipLaunchKernelGGL(k_milne_pred<true>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, m);
  else hipLaunchKernelGGL(k_milne_pred<false>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, m);
}

void launch_milne_corr(const DevData &d, const MilneDev &m, bool so3, hipStream_t s) {
  if (so3) hipLaunchKernelGGL(k_milne_corr<true>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, m);
  else hipLaunchKernelGGL(k_milne_corr<false>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, m);
}

void launch_milne_ctl(const MilneDev &m, const MilneCtl &mc, const int32_t *step, hipStream_t s) {
  hipLaunchKernelGGL(k_milne_ctl, dim3(1), dim3(1024), 0, s, m, mc, step);
}

void launch_milne_restore(const DevData &d, const MilneDev &m, bool so3, hipStream_t s) {
  if (so3) hipLaunchKernelGGL(k_milne_restore<true>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, m);
  else hipLaunchKernelGGL(k_milne_restore<false>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, m);
}

void launch_milne_commit_pl(const DevData &d, const MilneDev &m, hipStream_t s) {
  if (d.E == 0) return;
  hipLaunchKernelGGL(k_milne_commit_pl, dim3(nblk(d.E)), dim3(DL_BLOCK), 0, s, d,
                     (const int32_t *)m.acc);
}

void launch_set_f64(double *p, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(1), 0, s, p, v);
}
