// dl_milne_node.h -- the MilneDevice predictor and corrector at one node,
// shared by the node kernels of dl_milne.hip and the strut / joint kernels
// of dl_kernels.hip (which run them on the nodes they finish, with the
// nodal force still in registers).
#ifndef DL_MILNE_NODE_H
#define DL_MILNE_NODE_H

#include "dl_internal.h"

// boundary record of node n: constrained dofs, prescribed acceleration,
// external force.  An ImpactBC dof is free until the first accepted attempt
// (no group reaction yet).
__device__ __forceinline__ void milne_node_bc(const DevData &d, const MilneDev &m, int32_t n,
                                              unsigned &mask, double ac[6], double fe[6]) {
  const int32_t b = d.bc[n];
  mask = 0u;
#pragma unroll
  for (int c = 0; c < 6; c++) { ac[c] = 0.0; fe[c] = 0.0; }
  if (b < 0) return;
  const int64_t nb = d.nbc;
  mask = d.bc_mask[b];
  if (d.bc_imp && d.bc_imp[b] >= 0 && !*m.imp_ready) mask &= ~(1u << m.imp_dof[d.bc_imp[b]]);
#pragma unroll
  for (int c = 0; c < 6; c++) {
    ac[c] = d.bc_a[c * nb + b];
    fe[c] = d.bc_f[c * nb + b];
  }
}

template <bool SO3>
__device__ __forceinline__ void load_node(const double *u, const double *rot, const double *v,
                                          int64_t N, int64_t n, double uu[3], double rr[4],
                                          double vv[6], Quat &q) {
#pragma unroll
  for (int c = 0; c < 3; c++) uu[c] = u[c * N + n];
#pragma unroll
  for (int c = 0; c < 6; c++) vv[c] = v[c * N + n];
  if (SO3) {
#pragma unroll
    for (int c = 0; c < 4; c++) rr[c] = rot[c * N + n];
    q.w = rr[0]; q.x = rr[1]; q.y = rr[2]; q.z = rr[3];
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) rr[c] = rot[c * N + n];
    rr[3] = 0.0;
    q = qexp(rr[0], rr[1], rr[2]);
  }
}

template <bool SO3>
__device__ __forceinline__ void store_node(double *u, double *rot, double *v, int64_t N, int64_t n,
                                           const double uu[3], const double rr[4],
                                           const double vv[6]) {
#pragma unroll
  for (int c = 0; c < 3; c++) u[c * N + n] = uu[c];
#pragma unroll
  for (int c = 0; c < 6; c++) v[c * N + n] = vv[c];
#pragma unroll
  for (int c = 0; c < (SO3 ? 4 : 3); c++) rot[c * N + n] = rr[c];
}

// contact candidate lists: flag a rebuild when the node has left the skin
__device__ __forceinline__ void list_trigger(const DevData &d, int64_t n, const double u[3]) {
  if (!d.uref) return;
  const int64_t N = d.N;
  bool moved = false;
#pragma unroll
  for (int c = 0; c < 3; c++) moved = moved || fabs(u[c] - d.uref[c * N + n]) > d.ctrig;
  if (moved) *d.cgate = 1;
}

// predictor (MilneDeviceModule.cpp:64-96) at node n with internal force f
// at the start state (m.u0, m.r0, m.v0): the predicted state into (u1, r1,
// v1)
template <bool SO3>
__device__ __forceinline__ void milne_pred_node(const DevData &d, const MilneDev &m, int64_t n,
                                                double f[6]) {
  const int64_t N = d.N;
  double ac[6], fe[6];
  unsigned mask;
  milne_node_bc(d, m, (int32_t)n, mask, ac, fe);
  double u[3], r[4], v[6];
  Quat q;
  load_node<SO3>(m.u0, m.r0, m.v0, N, n, u, r, v, q);
  double J0[6];
  const double *J0d = d.Js ? d.Js : d.J0;
#pragma unroll
  for (int c = 0; c < 6; c++) J0[c] = J0d[c * N + n];
  double up[3], rp[4], vp[6];
  node_milne_pred<SO3>(f, q, u, v, r, d.mt[n], J0, SO3 ? d.gm[n] : 0.0, mask, ac, fe, *m.dt,
                       up, vp, rp, d.Js != nullptr);
  store_node<SO3>(m.u1, m.r1, m.v1, N, n, up, rp, vp);
  list_trigger(d, n, up);
}

// corrector (MilneDeviceModule.cpp:98-112) at node n with internal force f
// at the predicted state: the corrected state into (u0, r0, v0), the start
// state kept in (u1, r1, v1) for a rejection (each caller touches its own
// node only); e[0], e[1] += the node's squared differences pre - new.
// The same arithmetic as node_milne_corr, streamed quantity by quantity
// (velocities, translations, rotations) so that few values are live at once.
template <bool SO3>
__device__ __forceinline__ void milne_corr_node(const DevData &d, const MilneDev &m, int64_t n,
                                                double f[6], double e[2]) {
  const int64_t N = d.N;
  const double dt = *m.dt, ls = m.ls, rs = SO3 ? 1.0 : m.ls;
  double ac[6], fe[6];
  unsigned mask;
  milne_node_bc(d, m, (int32_t)n, mask, ac, fe);
  double r[4];
  Quat q;
  if (SO3) {
#pragma unroll
    for (int c = 0; c < 4; c++) r[c] = m.r0[c * N + n];
    q.w = r[0]; q.x = r[1]; q.y = r[2]; q.z = r[3];
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) r[c] = m.r0[c * N + n];
    r[3] = 0.0;
    q = qexp(r[0], r[1], r[2]);
  }
  double vp[6], a[6];
#pragma unroll
  for (int c = 0; c < 6; c++) vp[c] = m.v1[c * N + n];
  {
    double J0[6];
    const double *J0d = d.Js ? d.Js : d.J0;
#pragma unroll
    for (int c = 0; c < 6; c++) J0[c] = J0d[c * N + n];
    node_accel<SO3>(f, q, vp, d.mt[n], J0, SO3 ? d.gm[n] : 0.0, mask, ac, fe, a, d.Js != nullptr);
  }
  // velocities: v_new = v + dt a_pre
  double dvel[6];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    const double v = m.v0[c * N + n];
    const double vn = v + dt * a[c];
    dvel[c] = (vp[c] - vn) / (c < 3 ? ls : rs);
    m.v0[c * N + n] = vn;
    m.v1[c * N + n] = v;
  }
  // translations: u_new = u + dt v_pre
  double du[3], un[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double u = m.u0[c * N + n], up = m.u1[c * N + n];
    un[c] = u + dt * vp[c];
    du[c] = (up - un[c]) / ls;
    m.u0[c * N + n] = un[c];
    m.u1[c * N + n] = u;
  }
  // rotations: r_new = r (+) dt w_pre, compared as rotation vectors
  double rn[4], rp[4], tp[3], tn[3];
#pragma unroll
  for (int c = 0; c < 4; c++) rn[c] = r[c];
  rot_advance<SO3>(rn, q, vp + 3, dt);
#pragma unroll
  for (int c = 0; c < (SO3 ? 4 : 3); c++) rp[c] = m.r1[c * N + n];
  if (!SO3) rp[3] = 0.0;
  rot_vector<SO3>(rp, tp);
  rot_vector<SO3>(rn, tn);
#pragma unroll
  for (int c = 0; c < (SO3 ? 4 : 3); c++) {
    m.r0[c * N + n] = rn[c];
    m.r1[c * N + n] = r[c];
  }
  double eu = 0.0, ev = 0.0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double dr = (tp[c] - tn[c]) / rs;
    eu += du[c] * du[c] + dr * dr;
    ev += dvel[c] * dvel[c] + dvel[3 + c] * dvel[3 + c];
  }
  e[0] += eu;
  e[1] += ev;
  list_trigger(d, n, un);
}

// per-block sums of a 256-thread block in a fixed tree order into
// m.part[slot] (displacement) and m.part[m.nblk + slot] (velocity)
__device__ __forceinline__ void milne_block_partials(const MilneDev &m, const double e[2],
                                                     int64_t slot) {
  __shared__ double red[2][256];
  red[0][threadIdx.x] = e[0];
  red[1][threadIdx.x] = e[1];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    m.part[slot] = red[0][0];
    m.part[m.nblk + slot] = red[1][0];
  }
}

#endif
