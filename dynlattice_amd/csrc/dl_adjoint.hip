// dl_adjoint.hip -- reverse sweep of one explicit step (SO(3) mode, elastic
// or elastoplastic material), as gfx950 kernels plus the contact adjoint
// (dl_contact.hip):
//   k_adj_node    one thread per node: gather of the element ends of the step
//                 reversed before, then the VJP of the node update (mass solve,
//                 gyroscopic term, constraints, SO(3) composition) from the
//                 taped start state; gives the adjoint of the node's internal
//                 force (plus the objective's weight on the reaction), the
//                 adjoints of its start velocity and rotation, and the mass
//                 adjoints.
//   k_adj_elem    one thread per element: VJP of the element internal force
//                 (quaternion kinematics, strains, elastic stress resultants)
//                 into per-element-end buffers and the stiffness adjoint.
//   k_adj_elem_pl the same for the elastoplastic material: the cutting-plane
//                 return map reversed through every correction it took, the
//                 output load case of the reaction rows, and the adjoint of
//                 the committed plastic state carried one step back.
//   The node kernel of the next reversed step first sums the node's element
//   ends (CSR, fixed order) into its displacement and rotation adjoints.
// Like the forward step, assembly is a gather: no floating-point atomics,
// the result is bitwise reproducible.
#include <hip/hip_runtime.h>

#include "dl_adjoint.h"
#include "dl_internal.h"

#define AB 256

__global__ __launch_bounds__(AB) void k_adj_node(DevData d, AdjDev a, double dt_all, int gather) {
  const int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (n >= d.N) return;
  const int64_t N = d.N, E = d.E;
  const double *t = a.tape;
  if (d.uref) {
    // contact candidate lists: stale for this taped state?
    bool moved = false;
#pragma unroll
    for (int c = 0; c < 3; c++) moved = moved || fabs(t[c * N + n] - d.uref[c * N + n]) > d.ctrig;
    if (moved) *d.cgate = 1;
  }
  Quat q;
  q.w = t[3 * N + n]; q.x = t[4 * N + n]; q.y = t[5 * N + n]; q.z = t[6 * N + n];
  double v[6], f[6];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    v[c] = t[(7 + c) * N + n];
    f[c] = t[(13 + c) * N + n];
  }
  const int32_t b = d.bc[n];
  unsigned mask = 0u;
  double ac[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double fe[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int32_t rs = -1;
  if (b >= 0) {
    const int64_t nb = d.nbc;
    mask = d.bc_mask[b];
#pragma unroll
    for (int c = 0; c < 6; c++) {
      ac[c] = d.bc_a[c * nb + b];
      fe[c] = d.bc_f[c * nb + b];
    }
    rs = d.bc_rslot[b];
  }
  double J0[6];
#pragma unroll
  for (int c = 0; c < 6; c++) J0[c] = d.J0[c * N + n];
  const double dt = d.ndt ? d.ndt[n] : dt_all;
  double ubn[3], vbn[6];
  Quat qbn;
#pragma unroll
  for (int c = 0; c < 3; c++) ubn[c] = a.ubar[c * N + n];
  qbn.w = a.qbar[n]; qbn.x = a.qbar[N + n]; qbn.y = a.qbar[2 * N + n]; qbn.z = a.qbar[3 * N + n];
  if (gather) {
    // element ends of the step reversed before (the next one in time): their
    // displacement and rotation adjoints, summed in CSR order
    double ub[3] = {0.0, 0.0, 0.0}, qb[4] = {0.0, 0.0, 0.0, 0.0};
    const int32_t t1 = a.nptr[n + 1];
    for (int32_t k = a.nptr[n]; k < t1; k++) {
      const int32_t code = a.nlist[k];
      const int64_t e = code >> 1;
      const int o = (code & 1) ? 7 : 0;
#pragma unroll
      for (int c = 0; c < 3; c++) ub[c] += a.ebuf[(o + c) * E + e];
#pragma unroll
      for (int c = 0; c < 4; c++) qb[c] += a.ebuf[(o + 3 + c) * E + e];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) ubn[c] += ub[c];
    qbn.w += qb[0]; qbn.x += qb[1]; qbn.y += qb[2]; qbn.z += qb[3];
  }
#pragma unroll
  for (int c = 0; c < 6; c++) vbn[c] = a.vbar[c * N + n];
  NodeVJP r;
  node_advance_vjp(f, q, v, d.mt[n], J0, d.gm[n], mask, ac, fe, dt, ubn, qbn, vbn, r);
  // the recorded reaction is the internal force of the constrained dof
  const double rw = (rs >= 0 && a.rbar) ? a.rbar[a.slot_group[rs]] : 0.0;
  if (a.rnode) {
    // elastoplastic: the reaction is the output load case's force, which
    // differs from the step's where an element yields; the element kernel
    // takes the two apart
    double fs = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) fs = (c == d.react_dof) ? r.f[c] : fs;
    a.rnode[n] = rw;
    a.fstep_r[n] = fs;
  }
#pragma unroll
  for (int c = 0; c < 6; c++) r.f[c] += (c == d.react_dof) ? rw : 0.0;
#pragma unroll
  for (int c = 0; c < 6; c++) {
    a.fbar[c * N + n] = r.f[c];
    a.vbar[c * N + n] = r.v[c];
  }
  if (gather) {
#pragma unroll
    for (int c = 0; c < 3; c++) a.ubar[c * N + n] = r.u[c];
  }
  a.qbar[n] = r.q.w; a.qbar[N + n] = r.q.x; a.qbar[2 * N + n] = r.q.y; a.qbar[3 * N + n] = r.q.z;
  a.mtbar[n] += r.mt;
#pragma unroll
  for (int c = 0; c < 9; c++) a.J0bar[c * N + n] += r.J0[c];
}

__global__ __launch_bounds__(AB) void k_adj_elem(DevData d, AdjDev a) {
  const int64_t e = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (e >= d.E) return;
  const int64_t N = d.N, E = d.E, R = d.R;
  const int32_t na = d.en0[e], nb = d.en1[e], s = d.es[e];
  const double *t = a.tape;
  double Xa[3], ua[3], Xb[3], ub[3], fba[6], fbb[6], eps0[6], C[6];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    Xa[i] = d.X0[i * N + na]; ua[i] = t[i * N + na];
    Xb[i] = d.X0[i * N + nb]; ub[i] = t[i * N + nb];
  }
  Quat qna, qnb, qNa, qNb;
  qna.w = t[3 * N + na]; qna.x = t[4 * N + na]; qna.y = t[5 * N + na]; qna.z = t[6 * N + na];
  qnb.w = t[3 * N + nb]; qnb.x = t[4 * N + nb]; qnb.y = t[5 * N + nb]; qnb.z = t[6 * N + nb];
  qNa.w = d.qN[e]; qNa.x = d.qN[E + e]; qNa.y = d.qN[2 * E + e]; qNa.z = d.qN[3 * E + e];
  qNb.w = d.qN[4 * E + e]; qNb.x = d.qN[5 * E + e]; qNb.y = d.qN[6 * E + e]; qNb.z = d.qN[7 * E + e];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    fba[i] = a.fbar[i * N + na];
    fbb[i] = a.fbar[i * N + nb];
    eps0[i] = d.eps0[i * E + e];
    C[i] = d.C[i * R + s];
  }
  ElemVJP r;
  elem_force_vjp(Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, C, fba, fbb, r);
  double *eb = a.ebuf;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    eb[i * E + e] = r.ua[i];
    eb[(7 + i) * E + e] = r.ub[i];
  }
  eb[3 * E + e] = r.qa.w; eb[4 * E + e] = r.qa.x; eb[5 * E + e] = r.qa.y; eb[6 * E + e] = r.qa.z;
  eb[10 * E + e] = r.qb.w; eb[11 * E + e] = r.qb.x; eb[12 * E + e] = r.qb.y; eb[13 * E + e] = r.qb.z;
#pragma unroll
  for (int i = 0; i < 6; i++) a.Cbar[i * E + e] += r.C[i];
}

// Element force of the elastoplastic material reversed (one thread per
// element): the return map from the committed state taped for the step,
// and, at reaction-group nodes, the output load case (elastic predictor from
// that state) that the GroupOutput rows record; with a.seed (the closing
// evaluation at the final state) only the latter.  The adjoint of the
// committed plastic state moves one step back in a.pbar.
template <int EXPK>
__global__ __launch_bounds__(AB) void k_adj_elem_pl(DevData d, AdjDev a, PlasticParams pp) {
  const int64_t e = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (e >= d.E) return;
  const int64_t N = d.N, E = d.E, R = d.R;
  const int32_t na = d.en0[e], nb = d.en1[e], s = d.es[e];
  const double *t = a.tape;
  double Xa[3], ua[3], Xb[3], ub[3], fba[6], fbb[6], rba[6], rbb[6], eps0[6], C[6], w[6], Hk[6];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    Xa[i] = d.X0[i * N + na]; ua[i] = t[i * N + na];
    Xb[i] = d.X0[i * N + nb]; ub[i] = t[i * N + nb];
  }
  Quat qna, qnb, qNa, qNb;
  qna.w = t[3 * N + na]; qna.x = t[4 * N + na]; qna.y = t[5 * N + na]; qna.z = t[6 * N + na];
  qnb.w = t[3 * N + nb]; qnb.x = t[4 * N + nb]; qnb.y = t[5 * N + nb]; qnb.z = t[6 * N + nb];
  qNa.w = d.qN[e]; qNa.x = d.qN[E + e]; qNa.y = d.qN[2 * E + e]; qNa.z = d.qN[3 * E + e];
  qNb.w = d.qN[4 * E + e]; qNb.x = d.qN[5 * E + e]; qNb.y = d.qN[6 * E + e]; qNb.z = d.qN[7 * E + e];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    fba[i] = a.fbar[i * N + na];
    fbb[i] = a.fbar[i * N + nb];
    eps0[i] = d.eps0[i * E + e];
    C[i] = d.C[i * R + s];
    w[i] = d.pw[i * R + s];
    Hk[i] = kin_modulus(pp, i, C[i]);
  }
  bool has_r = false;
  if (a.seed) {
    // the closing row: the whole force adjoint is the output load case's
#pragma unroll
    for (int i = 0; i < 6; i++) {
      rba[i] = fba[i];
      rbb[i] = fbb[i];
    }
    has_r = true;
  } else {
    const double wa = a.rnode[na], wb = a.rnode[nb];
    const double sa = a.fstep_r[na], sb = a.fstep_r[nb];
#pragma unroll
    for (int i = 0; i < 6; i++) {
      const bool rd = i == d.react_dof;
      rba[i] = rd ? wa : 0.0;
      rbb[i] = rd ? wb : 0.0;
      fba[i] = (rd && wa != 0.0) ? sa : fba[i];
      fbb[i] = (rd && wb != 0.0) ? sb : fbb[i];
    }
    has_r = wa != 0.0 || wb != 0.0;
  }
  // committed plastic state at the step's start (the taped records)
  double ep0[6], al0[7], epb[6], alb[7];
  const double *rec = a.ptape + 12 * e;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    ep0[i] = rec[i];
    al0[1 + i] = rec[6 + i];
  }
  al0[0] = pp.has_iso ? a.ptape[12 * E + e] : 0.0;
  if (!a.seed) {
#pragma unroll
    for (int i = 0; i < 6; i++) epb[i] = a.pbar[i * E + e];
#pragma unroll
    for (int i = 0; i < 7; i++) alb[i] = a.pbar[(6 + i) * E + e];
  }
  ElemVJP r;
  MatBar mb;
#pragma unroll
  for (int i = 0; i < 6; i++) mb.C[i] = mb.w[i] = mb.Hk[i] = 0.0;
  const int it = elem_force_vjp_pl<EXPK>(Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, pp, C, w, Hk,
                                         ep0, al0, a.seed != 0, fba, fbb, has_r, rba, rbb, epb,
                                         alb, r, mb);
  if (it < 0) *d.err = -3;
  double *eb = a.ebuf;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    eb[i * E + e] = r.ua[i];
    eb[(7 + i) * E + e] = r.ub[i];
  }
  eb[3 * E + e] = r.qa.w; eb[4 * E + e] = r.qa.x; eb[5 * E + e] = r.qa.y; eb[6 * E + e] = r.qa.z;
  eb[10 * E + e] = r.qb.w; eb[11 * E + e] = r.qb.x; eb[12 * E + e] = r.qb.y; eb[13 * E + e] = r.qb.z;
#pragma unroll
  for (int i = 0; i < 6; i++) a.pbar[i * E + e] = epb[i];
#pragma unroll
  for (int i = 0; i < 7; i++) a.pbar[(6 + i) * E + e] = alb[i];
  // H_kin = kin_const + kin_ratio C
#pragma unroll
  for (int i = 0; i < 6; i++) {
    a.Cbar[i * E + e] += mb.C[i] + (pp.has_kin ? pp.kin_ratio * mb.Hk[i] : 0.0);
    a.Wbar[i * E + e] += mb.w[i];
  }
}

// the closing reaction row's weight as the adjoint of the nodal internal
// force at the final state: rbar of the node's group on the reaction dof
__global__ __launch_bounds__(AB) void k_adj_seed(DevData d, AdjDev a) {
  const int64_t n = (int64_t)blockIdx.x * AB + threadIdx.x;
  if (n >= d.N) return;
  const int32_t b = d.bc[n];
  const int32_t rs = b >= 0 ? d.bc_rslot[b] : -1;
  const double rw = rs >= 0 ? a.rbar[a.slot_group[rs]] : 0.0;
#pragma unroll
  for (int c = 0; c < 6; c++) a.fbar[c * d.N + n] = (c == d.react_dof) ? rw : 0.0;
}

static unsigned nblk_a(int64_t n) { return (unsigned)((n + AB - 1) / AB); }

void launch_adj_seed(const DevData &d, const AdjDev &a, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_seed, dim3(nblk_a(d.N)), dim3(AB), 0, s, d, a);
}

void launch_adj_node(const DevData &d, const AdjDev &a, double dt, bool gather, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_node, dim3(nblk_a(d.N)), dim3(AB), 0, s, d, a, dt, (int)gather);
}
void launch_adj_elem(const DevData &d, const AdjDev &a, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_elem, dim3(nblk_a(d.E)), dim3(AB), 0, s, d, a);
}
void launch_adj_elem_pl(const DevData &d, const AdjDev &a, const PlasticParams &pp, int expk,
                        hipStream_t s) {
  if (expk == 1)
    hipLaunchKernelGGL(k_adj_elem_pl<1>, dim3(nblk_a(d.E)), dim3(AB), 0, s, d, a, pp);
  else if (expk == 2)
    hipLaunchKernelGGL(k_adj_elem_pl<2>, dim3(nblk_a(d.E)), dim3(AB), 0, s, d, a, pp);
  else
    hipLaunchKernelGGL(k_adj_elem_pl<0>, dim3(nblk_a(d.E)), dim3(AB), 0, s, d, a, pp);
}
