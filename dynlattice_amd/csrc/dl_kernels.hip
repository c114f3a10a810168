// dl_kernels.hip -- gfx950 kernels of the explicit rod-lattice step (FP64).
//
// One explicit step is two launches (plus contact when enabled):
//   k_strut  one lane per element; a strut's elements (a chain of at most
//            64) sit in consecutive lanes of one wavefront.  Computes the
//            element's internal force (kinematics, strains, material return
//            map, XI^T sigma), sums the two element contributions of each
//            private (single-strut) node in registers (the left neighbour's
//            by a lane shuffle) and immediately applies the node update
//            (mass, gyro, constraints, central difference, SO(3)
//            composition).  Contributions to shared (joint) nodes go to a
//            slot buffer.
//   k_joint  one thread per shared node.  Sums its slots in a fixed order
//            (CSR, strut order) and applies the same node update.
// Assembly is therefore a deterministic gather: no floating-point atomics.
// The reference evaluates the same quantities element by element and
// scatters into the global vector (SpecialCosseratRodModel.cpp:899-949),
// then updates every DOF (ExplicitModule.cpp:339-392, LeapFrogModule.cpp:44-84).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dl_internal.h"
#include "dl_milne_node.h"

#define DL_BLOCK 256

// ------------------------------------------------------------------
// node helpers
// ------------------------------------------------------------------
template <bool SO3>
__device__ __forceinline__ void load_rot(const DevData &d, int32_t n, double r[4], Quat &q) {
  const int64_t N = d.N;
  if (SO3) {
    q.w = d.rot[n]; q.x = d.rot[N + n]; q.y = d.rot[2 * N + n]; q.z = d.rot[3 * N + n];
    r[0] = q.w; r[1] = q.x; r[2] = q.y; r[3] = q.z;
  } else {
    r[0] = d.rot[n]; r[1] = d.rot[N + n]; r[2] = d.rot[2 * N + n]; r[3] = 0.0;
    q = qexp(r[0], r[1], r[2]);
  }
}

// a = M^-1 (f_ext - f) with constraints, then v += dt a, u += dt v,
// R <- exp(dt w) R (SO(3)) or theta += dt w (additive).
// f: element (+ contact) internal force of the node at the step's start state.
// With a tape, the node's start state and f are recorded for the adjoint.
// The rotation is (re)loaded here rather than kept live from the element
// part: fewer registers across the material update (the line is in cache).
// Nodes of ImpactBC groups are deferred: their force goes to fout and the
// impact kernel updates them once the group's reaction is known
// (impact_pass, with free_mask clearing the impact dof before the first
// output).
// The node's own data, loaded ahead of its force where registers allow
// (the joint kernel: its loads overlap the slot loads).
struct NodeData {
  double v[6], u[3], rr[4], J0[6], mt, gm;
  int32_t b;
};

template <bool SO3>
__device__ __forceinline__ void load_node_data(const DevData &d, int32_t n, NodeData &nd) {
  const int64_t N = d.N;
  nd.b = d.bc[n];
#pragma unroll
  for (int c = 0; c < 6; c++) nd.v[c] = d.v[c * N + n];
#pragma unroll
  for (int c = 0; c < 3; c++) nd.u[c] = d.u[c * N + n];
#pragma unroll
  for (int c = 0; c < (SO3 ? 4 : 3); c++) nd.rr[c] = d.rot[c * N + n];
  if (!SO3) nd.rr[3] = 0.0;
  const double *J0 = d.Js ? d.Js : d.J0;
#pragma unroll
  for (int c = 0; c < 6; c++) nd.J0[c] = J0[c * N + n];
  nd.mt = d.mt[n];
  nd.gm = SO3 ? d.gm[n] : 0.0;
}

template <bool SO3>
__device__ __forceinline__ void node_update(const DevData &d, const StepParams &sp, int32_t n,
                                            double f[6], unsigned free_mask = 0u,
                                            bool impact_pass = false,
                                            const NodeData *pre = nullptr) {
  const int64_t N = d.N;
  const int32_t b = pre ? pre->b : d.bc[n];
  if (!impact_pass && b >= 0 && d.bc_imp && d.bc_imp[b] >= 0) {
#pragma unroll
    for (int c = 0; c < 6; c++) d.fout[c * N + n] = f[c];
    return;
  }
  unsigned mask = 0u;
  double ac[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double fe[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (b >= 0) {
    const int64_t nb = d.nbc;
    mask = d.bc_mask[b] & ~free_mask;
#pragma unroll
    for (int c = 0; c < 6; c++) {
      ac[c] = d.bc_a[c * nb + b];
      fe[c] = d.bc_f[c * nb + b];
    }
  }
  double v[6], u[3], rr[4];
  Quat q;
  if (pre) {
#pragma unroll
    for (int c = 0; c < 4; c++) rr[c] = pre->rr[c];
    if (SO3) {
      q.w = rr[0]; q.x = rr[1]; q.y = rr[2]; q.z = rr[3];
    } else {
      q = qexp(rr[0], rr[1], rr[2]);
    }
#pragma unroll
    for (int c = 0; c < 6; c++) v[c] = pre->v[c];
#pragma unroll
    for (int c = 0; c < 3; c++) u[c] = pre->u[c];
  } else {
    // compiler barrier: reload the rotation instead of reusing the element
    // part's copy (which would stay live in registers)
    asm volatile("" ::: "memory");
    load_rot<SO3>(d, n, rr, q);
#pragma unroll
    for (int c = 0; c < 6; c++) v[c] = d.v[c * N + n];
#pragma unroll
    for (int c = 0; c < 3; c++) u[c] = d.u[c * N + n];
  }
  if (d.tape) {
    // adjoint tape row: u(3) rot(4) v(6) f(6), component-major
    double *t = d.tape;
#pragma unroll
    for (int c = 0; c < 3; c++) t[c * N + n] = u[c];
#pragma unroll
    for (int c = 0; c < 4; c++) t[(3 + c) * N + n] = rr[c];
#pragma unroll
    for (int c = 0; c < 6; c++) t[(7 + c) * N + n] = v[c];
#pragma unroll
    for (int c = 0; c < 6; c++) t[(13 + c) * N + n] = f[c];
  }
  const double mt = pre ? pre->mt : d.mt[n];
  double J0[6];
  const double *J0d = d.Js ? d.Js : d.J0;
#pragma unroll
  for (int c = 0; c < 6; c++) J0[c] = pre ? pre->J0[c] : J0d[c * N + n];
  const double dt = d.ndt ? d.ndt[n] : sp.dt;
  node_advance<SO3>(f, q, u, v, rr, mt, J0, pre ? pre->gm : (SO3 ? d.gm[n] : 0.0), mask, ac, fe,
                    dt, d.Js != nullptr);
#pragma unroll
  for (int c = 0; c < 6; c++) d.v[c * N + n] = v[c];
#pragma unroll
  for (int c = 0; c < 3; c++) d.u[c * N + n] = u[c];
  if (d.uref) {
    bool moved = false;
#pragma unroll
    for (int c = 0; c < 3; c++) moved = moved || fabs(u[c] - d.uref[c * N + n]) > d.ctrig;
    if (moved) *d.cgate = 1;
  }
  if (SO3) {
#pragma unroll
    for (int c = 0; c < 4; c++) d.rot[c * N + n] = rr[c];
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) d.rot[c * N + n] = rr[c];
  }
}

// ------------------------------------------------------------------
// INIT: initial material strain at zero displacement (initStrain_,
// SpecialCosseratRodModel.cpp:536-569)
// ------------------------------------------------------------------
__global__ __launch_bounds__(DL_BLOCK) void k_init_strain(DevData d) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.E) return;
  const int64_t E = d.E, N = d.N;
  const int32_t na = d.en0[e], nb = d.en1[e];
  double Xa[3], Xb[3];
  for (int i = 0; i < 3; i++) { Xa[i] = d.X0[i * N + na]; Xb[i] = d.X0[i * N + nb]; }
  Quat qa, qb;
  qa.w = d.qN[e]; qa.x = d.qN[E + e]; qa.y = d.qN[2 * E + e]; qa.z = d.qN[3 * E + e];
  qb.w = d.qN[4 * E + e]; qb.x = d.qN[5 * E + e]; qb.y = d.qN[6 * E + e]; qb.z = d.qN[7 * E + e];
  double eps0[6];
  elem_eps0(Xa, Xb, qa, qb, eps0, (d.eflag[e] & 4u) != 0);
  for (int i = 0; i < 6; i++) d.eps0[i * E + e] = eps0[i];
}

// ------------------------------------------------------------------
// element-chain kernel: one thread (lane) per element
//
// The elements of a chain (a strut, or a <= 64-element piece of a longer
// strut) sit in consecutive lanes of one 64-wide wavefront.  Lane k owns the
// chain node p_k (its element's node a): the node's internal force is
// fb(k-1) + fa(k), with fb(k-1) taken from the left neighbour lane by a
// shuffle.  The last lane of a chain also owns p_m.  A private node is
// updated in place right away; a shared node's contribution goes to its
// slot and the joint kernel finishes it.
// ------------------------------------------------------------------
// MAT: 0 elastic; 1, 2, 3 elastoplastic with yield exponents all 1, all 2, general
// GroupOutput rows of an element with an end at a reaction-group node
// (GroupOutputModule.cpp:100-116, SpecialCosseratRodModel.cpp:762): the
// output load case at the step's start state is the elastic predictor from
// the committed plastic state -- the trial stress st of this step -- plus
// the element's contact row; the same operations as k_react.
template <bool CONTACT>
__device__ __forceinline__ void react_rows(const DevData &d, const ElemKin &k, const double st[6],
                                           int32_t e) {
  const int64_t E = d.E;
  const int dof = d.react_dof;
  double va = 0.0, vb = 0.0;
  switch (dof) {
  case 0: elem_force_dof<0>(k, st, &va, &vb); break;
  case 1: elem_force_dof<1>(k, st, &va, &vb); break;
  case 2: elem_force_dof<2>(k, st, &va, &vb); break;
  case 3: elem_force_dof<3>(k, st, &va, &vb); break;
  case 4: elem_force_dof<4>(k, st, &va, &vb); break;
  default: elem_force_dof<5>(k, st, &va, &vb); break;
  }
  const bool cb = CONTACT && dof < 3 && ((d.cbits[e >> 6] >> (e & 63)) & 1ull);
  const int32_t ka = d.rslot_e[e], kb = d.rslot_e[E + e];
  if (ka >= 0) {
    if (cb) va += d.fcs[dof * E + e];
    d.rbuf[ka] = va;
  }
  if (kb >= 0) {
    if (cb) vb += d.fcs[(3 + dof) * E + e];
    d.rbuf[kb] = vb;
  }
}

template <int MODE, bool SO3, int MAT, bool CONTACT>
__global__ __launch_bounds__(DL_BLOCK, MAT == 3 ? 1 : 3) void k_strut(DevData d, StepParams sp, PlasticParams pp,
                                                    MilneDev m) {
  constexpr bool PLASTIC = MAT != 0;
  const int64_t L = (int64_t)blockIdx.x * DL_BLOCK + threadIdx.x;
  // element of the lane: identity packing (every chain fills whole
  // wavefronts) saves a dependent load before the element data
  const int32_t e = d.lane_ident ? (L < d.E ? (int32_t)L : -1) : ((L < d.NL) ? d.lane_e[L] : -1);
  const int64_t N = d.N, E = d.E, R = d.R;
  double fa[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, fb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int32_t na = 0, nb = 0;
  uint32_t fl = 0;
  double diss = 0.0;
  // LDS of the block: the plastic material's stash (21 doubles per lane),
  // later the forces of the wavefront joints (6 per lane); each lane writes
  // its own column
  __shared__ double lds[PLASTIC ? 21 : 6][DL_BLOCK];
  const int tx = threadIdx.x;
  if (e >= 0) {
    fl = d.eflag[e];
    na = d.en0[e];
    nb = d.en1[e];
    const int32_t s = d.es[e];
    double C[6];
#pragma unroll
    for (int i = 0; i < 6; i++) C[i] = d.C[i * R + s];
    double Xa[3], ua[3], Xb[3], ub[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      Xa[i] = d.X0[i * N + na]; ua[i] = d.u[i * N + na];
      Xb[i] = d.X0[i * N + nb]; ub[i] = d.u[i * N + nb];
    }
    double ra[4], rb[4];
    Quat qna, qnb;
    load_rot<SO3>(d, na, ra, qna);
    load_rot<SO3>(d, nb, rb, qnb);
    // end-a frame; a lane that is not its chain's last takes its end-b frame
    // from the next lane (same chain, so that lane is active here too)
    Quat qNa, qNb;
    qNa.w = d.qN[e]; qNa.x = d.qN[E + e]; qNa.y = d.qN[2 * E + e]; qNa.z = d.qN[3 * E + e];
    qNb.w = __shfl_down(qNa.w, 1, 64);
    qNb.x = __shfl_down(qNa.x, 1, 64);
    qNb.y = __shfl_down(qNa.y, 1, 64);
    qNb.z = __shfl_down(qNa.z, 1, 64);
    const uint64_t ends = __ballot(fl & 2u);
    if (!d.qN_share) {
      qNb.w = d.qN[4 * E + e]; qNb.x = d.qN[5 * E + e]; qNb.y = d.qN[6 * E + e]; qNb.z = d.qN[7 * E + e];
    } else if (fl & 2u) {
      const int lane = threadIdx.x & 63;
      const int64_t ne = d.qNe_wave[L >> 6] + __popcll(ends & ((1ull << lane) - 1ull));
      qNb.w = d.qNe[ne]; qNb.x = d.qNe[d.nqe + ne]; qNb.y = d.qNe[2 * d.nqe + ne]; qNb.z = d.qNe[3 * d.nqe + ne];
    }
    double strain[6], sm[6];
    elem_eps0(Xa, Xb, qNa, qNb, strain);
    ElemKin k;
    elem_kinematics(Xa, ua, Xb, ub, qmul(qna, qNa), qmul(qnb, qNb), k);
    if (fl & 4u) { // frames -I (strut along -e3)
      elem_mirror(k);
#pragma unroll
      for (int i = 0; i < 3; i++) strain[i] = -strain[i];
    }
#pragma unroll
    for (int i = 0; i < 6; i++) strain[i] = k.eps[i] - strain[i];
    if (!PLASTIC) {
#pragma unroll
      for (int i = 0; i < 6; i++) sm[i] = C[i] * strain[i];
      if (MODE == MODE_STEP && (fl & 8u)) react_rows<CONTACT>(d, k, sm, e);
      if (MODE != MODE_STEP) rec6_store(d.ecur, e, strain);
    } else {
      double pw[6], Hk[6];
#pragma unroll
      for (int i = 0; i < 6; i++) { pw[i] = d.pw[i * R + s]; Hk[i] = kin_modulus(pp, i, C[i]); }
      // committed plastic state: its record only where it may be nonzero;
      // the isotropic parameter only when in use
      double ep[6], al[7];
      if (fl & PST_FLAG) pst_load(d.pst, d.aiso, e, pp.has_iso, ep, al);
      else pst_zero(ep, al);
      double dfl;
      // The kinematics needed after the return map (Lambda, phi', weights)
      // and the strain wait in LDS meanwhile: the return map's working set
      // then fits 168 VGPRs, i.e. 3 waves per SIMD instead of 2.  The
      // compiler barriers keep the values from being forwarded in registers.
      double (*stash)[DL_BLOCK] = lds;
#pragma unroll
      for (int i = 0; i < 9; i++) stash[i][tx] = k.R[i];
#pragma unroll
      for (int i = 0; i < 3; i++) stash[9 + i][tx] = k.phiP[i];
#pragma unroll
      for (int i = 0; i < 6; i++) stash[12 + i][tx] = strain[i];
      stash[18][tx] = k.w;
      stash[19][tx] = k.g0;
      stash[20][tx] = k.g1;
      asm volatile("" ::: "memory");
      const int it = return_map<MAT == 3 ? 0 : MAT>(pp, C, pw, Hk, strain, ep, al, sm, &dfl,
                                                     sp.inelastic);
      asm volatile("" ::: "memory");
#pragma unroll
      for (int i = 0; i < 9; i++) k.R[i] = stash[i][tx];
#pragma unroll
      for (int i = 0; i < 3; i++) k.phiP[i] = stash[9 + i][tx];
#pragma unroll
      for (int i = 0; i < 6; i++) strain[i] = stash[12 + i][tx];
      k.w = stash[18][tx];
      k.g0 = stash[19][tx];
      k.g1 = stash[20][tx];
      if (MODE == MODE_STEP && (fl & 8u)) {
        // the reaction rows from the trial stress: the committed plastic
        // strain is still the one in memory here
        double st[6], epc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (fl & PST_FLAG) {
#pragma unroll
          for (int i = 0; i < 6; i++) epc[i] = d.pst[12 * (int64_t)e + i];
        }
#pragma unroll
        for (int i = 0; i < 6; i++) st[i] = C[i] * (strain[i] - epc[i]);
        react_rows<CONTACT>(d, k, st, e);
      }
      if (it < 0) *d.err = -3;
      if (MODE == MODE_STEP) {
        if (sp.track_diss) {
          // applyDeform dissipation (ElastoPlasticRodMaterial.cpp:285-296)
          if (fabs(dfl) > 2.220446049250313e-16) {
            double eo[6], epo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            rec6_load(d.eo, e, eo);
            if (fl & PST_FLAG) { // committed plastic strain (updated below)
#pragma unroll
              for (int i = 0; i < 6; i++) epo[i] = d.pst[12 * (int64_t)e + i];
            }
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < 6; i++) {
              const double so = C[i] * (eo[i] - epo[i]);
              t += ((so + sm[i]) / 2.0) * (ep[i] - epo[i]);
            }
            diss = k.w * t;
          }
          rec6_store(d.eo, e, strain);
        }
        if (it > 0) {
          pst_store(d.pst, d.aiso, e, pp.has_iso, ep, al);
          if (!(fl & PST_FLAG)) d.eflag[e] = (uint8_t)(fl | PST_FLAG);
        }
      } else {
        // trial plastic state: written only where the return map corrected
        // it (elsewhere it equals the committed state, and COMMIT reads the
        // committed one through the active flag)
        if (it > 0) pst_store(d.pst_cur, d.aiso_cur, e, true, ep, al);
        rec6_store(d.ecur, e, strain);
        d.dflow[e] = dfl;
      }
      d.active[e] = (uint8_t)(it > 0);
    }
    elem_forces(k, sm, fa, fb);
    if (CONTACT) {
      // rod-rod contact force on this element's nodes (RodContactModel.cpp:729)
      if ((d.cbits[e >> 6] >> (e & 63)) & 1ull) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
          fa[i] += d.fcs[i * E + e];
          fb[i] += d.fcs[(3 + i) * E + e];
        }
      }
    }
  }
  // node a of lane k receives fb of lane k-1 (same chain, same wavefront)
  double em[2] = {0.0, 0.0};  // corrector error sums of this lane's nodes
  double f[6];
#pragma unroll
  for (int i = 0; i < 6; i++) f[i] = __shfl_up(fb[i], 1, 64);
  int32_t slot = -1;
  if (e >= 0) {
    if (fl & 1u) {
#pragma unroll
      for (int i = 0; i < 6; i++) f[i] = fa[i];
    } else {
#pragma unroll
      for (int i = 0; i < 6; i++) f[i] += fa[i];
    }
    slot = d.sa[e];
  }
  // Joints whose slots are all node-a forces of this wavefront (a BCC cell
  // centre's eight chains): every contributing lane puts its force in its
  // LDS column, and the joint's lowest lane sums the columns of its
  // contributors in lane order -- the joint's slot order, summed as the
  // joint kernel sums slots -- and then updates the node.  The exchange
  // stays within the wavefront, whose LDS writes and the reads after them
  // execute in order.  (Columns 0..5 of the block's LDS; each lane writes
  // its own.)
  const int me = tx & 63;
  bool own_a = slot == -1; // this lane updates node a
  const uint64_t jm = __ballot(slot <= -2); // the wavefront's contributors
  if (jm) {
    const int own = -2 - slot;
    if (slot <= -2) {
#pragma unroll
      for (int i = 0; i < 6; i++) lds[i][tx] = f[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // each owner's contributors: one ballot per owner of the wavefront
    uint64_t mine = 0ull;
    for (uint64_t om = __ballot(slot <= -2 && own == me); om; om &= om - 1ull) {
      const int o = __ffsll((long long)om) - 1;
      const uint64_t mo = __ballot(slot <= -2 && own == o);
      if (me == o) mine = mo;
    }
    if (mine) {
      own_a = true;
      const int base = tx - me; // the wavefront's first column
      // the owner's own force is in its column too: start from zero, add in
      // lane order, two columns' loads ahead of their additions
#pragma unroll
      for (int i = 0; i < 6; i++) f[i] = 0.0;
      while (mine) {
        const int c0 = base + __ffsll((long long)mine) - 1;
        mine &= mine - 1ull;
        const bool two = mine != 0ull;
        const int c1 = two ? base + __ffsll((long long)mine) - 1 : c0;
        if (two) mine &= mine - 1ull;
        double v0[6], v1[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
          v0[i] = lds[i][c0];
          v1[i] = lds[i][c1];
        }
#pragma unroll
        for (int i = 0; i < 6; i++) f[i] += v0[i];
        if (two) {
#pragma unroll
          for (int i = 0; i < 6; i++) f[i] += v1[i];
        }
      }
    }
  }
  if (e >= 0) {
    if (slot >= 0) {
#pragma unroll
      for (int i = 0; i < 6; i++) d.sbuf[(int64_t)i * d.nslot + slot] = f[i];
    } else if (!own_a) {
      // a joint of this wavefront summed by another lane
    } else if (MODE == MODE_FORCE) {
#pragma unroll
      for (int i = 0; i < 6; i++) d.fout[i * N + na] = f[i];
    } else if (MODE == MODE_STEP) {
      node_update<SO3>(d, sp, na, f);
    } else if (MODE == MODE_MPRED) {
      milne_pred_node<SO3>(d, m, na, f);
    } else {
      milne_corr_node<SO3>(d, m, na, f, em);
    }
    if (fl & 2u) { // last element of the chain: also node b
      const int32_t sl = d.sb[e];
      if (sl >= 0) {
#pragma unroll
        for (int i = 0; i < 6; i++) d.sbuf[(int64_t)i * d.nslot + sl] = fb[i];
      } else if (MODE == MODE_FORCE) {
#pragma unroll
        for (int i = 0; i < 6; i++) d.fout[i * N + nb] = fb[i];
      } else if (MODE == MODE_STEP) {
        node_update<SO3>(d, sp, nb, fb);
      } else if (MODE == MODE_MPRED) {
        milne_pred_node<SO3>(d, m, nb, fb);
      } else {
        milne_corr_node<SO3>(d, m, nb, fb, em);
      }
    }
  }
  if (MODE == MODE_MCORR) milne_block_partials(m, em, blockIdx.x);
  if (PLASTIC && MODE == MODE_STEP) {
    if (sp.track_diss) {
      // per-block partial of the dissipation increment, fixed tree order
      __shared__ double red[DL_BLOCK];
      red[threadIdx.x] = diss;
      __syncthreads();
      for (int w = DL_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) d.dbuf[blockIdx.x] = red[0];
    }
  }
}

// ------------------------------------------------------------------
// joint kernel: shared nodes
// ------------------------------------------------------------------
// internal force of shared node j: its chains' slots summed in slot order,
// eight slots per iteration (a BCC joint's eight struts in one round) with
// their loads in flight together; the additions keep the order, so the sum
// is the same bits as one by one.  The joint kernel has few wavefronts, so
// the registers this takes cost no occupancy.
__device__ __forceinline__ void joint_force(const DevData &d, int32_t j, double f[6]) {
  const int32_t t0 = d.jptr[j], t1 = d.jptr[j + 1];
#pragma unroll
  for (int i = 0; i < 6; i++) f[i] = 0.0;
  for (int32_t t = t0; t < t1; t += 8) {
    double g[8][6];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int64_t tq = (t + q < t1) ? t + q : t;
#pragma unroll
      for (int i = 0; i < 6; i++) g[q][i] = (t + q < t1) ? d.sbuf[i * (int64_t)d.nslot + tq] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; q++)
#pragma unroll
      for (int i = 0; i < 6; i++)
        if (t + q < t1) f[i] += g[q][i];
  }
}

// ------------------------------------------------------------------
// end of step: per-group reaction sums (the reaction of the state the step
// started from, i.e. the previous step's GroupOutput row) and the
// dissipation increment, in a fixed tree order (deterministic)
// ------------------------------------------------------------------
// one sum (reaction group g < ngroups, else the dissipation partials)
// Sums g0 .. g0 + P - 1 (those below nsums) side by side: their loads are
// in flight together and they share the tree's barriers.  Each sum keeps
// its own fixed order.
// rrow: history row of the reaction sums (< 0: not written); drow: row of
// the dissipation sum
#define SE_NT 256 // threads of a step-end block
template <int P>
__device__ __forceinline__ void step_end_sums(int g0, int nsums, const double *rbuf,
                                              const int32_t *rgptr, int32_t ngroups,
                                              const double *dbuf, int64_t ndb,
                                              double *diss_total, double *hist_react,
                                              double *hist_diss, int64_t rrow, int64_t drow,
                                              double (*red)[SE_NT]) {
  const double *in[P];
  int64_t n[P];
#pragma unroll
  for (int q = 0; q < P; q++) {
    const int g = g0 + q;
    if (g >= nsums) {
      in[q] = dbuf;
      n[q] = 0;
    } else if (g < ngroups) {
      in[q] = rbuf + rgptr[g];
      n[q] = rgptr[g + 1] - rgptr[g];
    } else {
      in[q] = dbuf;
      n[q] = ndb;
    }
  }
  // thread t sums in[t], in[t + SE_NT], ... in that order, eight (then four)
  // loads per sum in flight at a time
  double s[P];
#pragma unroll
  for (int q = 0; q < P; q++) {
    s[q] = 0.0;
    int64_t i = threadIdx.x;
    for (; i + 7 * SE_NT < n[q]; i += 8 * SE_NT) {
      double a[8];
#pragma unroll
      for (int k = 0; k < 8; k++) a[k] = in[q][i + k * SE_NT];
#pragma unroll
      for (int k = 0; k < 8; k++) s[q] += a[k];
    }
    for (; i + 3 * SE_NT < n[q]; i += 4 * SE_NT) {
      const double a0 = in[q][i], a1 = in[q][i + SE_NT], a2 = in[q][i + 2 * SE_NT],
                   a3 = in[q][i + 3 * SE_NT];
      s[q] += a0;
      s[q] += a1;
      s[q] += a2;
      s[q] += a3;
    }
    for (; i < n[q]; i += SE_NT) s[q] += in[q][i];
    red[q][threadIdx.x] = s[q];
  }
  // tree: red[t] += red[t + w] for w = SE_NT / 2 .. 1; the levels below 64
  // run in the first wavefront as shuffles (same pairs, same order)
  __syncthreads();
  for (int w = SE_NT / 2; w >= 64; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int q = 0; q < P; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 64) {
#pragma unroll
    for (int q = 0; q < P; q++) {
      double r = red[q][threadIdx.x];
      for (int w = 32; w > 0; w >>= 1) r += __shfl_down(r, w, 64);
      const int g = g0 + q;
      if (threadIdx.x == 0 && g < nsums) {
        if (g < ngroups) {
          if (rrow >= 0) hist_react[rrow * ngroups + g] = r;
        } else {
          const double t = *diss_total + r;
          *diss_total = t;
          hist_diss[drow] = t;
        }
      }
    }
  }
  __syncthreads();
}

// several sums, one block each (the step counter moves in k_step_inc)
__global__ __launch_bounds__(SE_NT) void k_step_end(const double *rbuf, const int32_t *rgptr,
                                                    int32_t ngroups, const double *dbuf, int64_t ndb,
                                                    double *diss_total, double *hist_react,
                                                    double *hist_diss, const int32_t *step) {
  __shared__ double red[1][SE_NT];
  const int64_t st = *step;
  step_end_sums<1>(blockIdx.x, blockIdx.x + 1, rbuf, rgptr, ngroups, dbuf, ndb, diss_total,
                   hist_react, hist_diss, st - 1, st, red);
}

// few sums: one block does them two at a time and then moves the step
// counter; tail: the reaction row only.  Run by k_step_end1 or by the first
// block of the step's k_joint launch (no other block of it reads the counter)
__device__ __forceinline__ void step_end_block(const StepEnd &se, double (*red)[SE_NT]) {
  const int64_t st = *se.step;
  for (int g = 0; g < se.nsums; g += 2)
    step_end_sums<2>(g, se.nsums, se.rbuf, se.rgptr, se.ngroups, se.dbuf, se.ndb, se.diss_total,
                     se.hist_react, se.hist_diss, st - 1, st, red);
  __syncthreads();
  if (threadIdx.x == 0 && !se.tail) *se.step = (int32_t)(st + 1);
}

// one sum per block (more than 8 sums); the last block to finish moves the
// counter -- every other block has read it by then
__device__ __forceinline__ void step_end_one(const StepEnd &se, int g, double (*red)[SE_NT]) {
  const int64_t st = *se.step;
  step_end_sums<1>(g, g + 1, se.rbuf, se.rgptr, se.ngroups, se.dbuf, se.ndb, se.diss_total,
                   se.hist_react, se.hist_diss, st - 1, st, red);
  if (threadIdx.x == 0 && atomicAdd(se.done, 1) == se.nblocks - 1) {
    *se.done = 0;
    if (!se.tail) *se.step = (int32_t)(st + 1);
  }
}

__global__ __launch_bounds__(SE_NT) void k_step_end1(StepEnd se) {
  __shared__ double red[2][SE_NT];
  step_end_block(se, red);
}

template <int MODE, bool SO3, bool SE>
__global__ __launch_bounds__(DL_BLOCK) void k_joint(DevData d, StepParams sp, MilneDev m, StepEnd se) {
  static_assert(DL_BLOCK == SE_NT, "a step-end block is a k_joint block");
  if constexpr (SE) {
    // the first se.nblocks blocks: the step's end sums (their inputs come
    // from k_strut).  Dispatched first, their serial sums overlap the joint
    // blocks instead of trailing them.
    const int g = (int)blockIdx.x;
    if (g < se.nblocks) {
      __shared__ double red[2][SE_NT];
      if (se.nblocks == 1) step_end_block(se, red);
      else step_end_one(se, g, red);
      return;
    }
  }
  const int32_t j = ((int32_t)blockIdx.x - (SE ? se.nblocks : 0)) * blockDim.x + threadIdx.x;
  if (MODE == MODE_MCORR) {
    // every thread joins the block's error partial sums
    double em[2] = {0.0, 0.0};
    if (j < d.NJ) {
      double f[6];
      joint_force(d, j, f);
      milne_corr_node<SO3>(d, m, d.jnode[j], f, em);
    }
    milne_block_partials(m, em, m.jofs + blockIdx.x);
    return;
  }
  if (j >= d.NJ) return;
  const int32_t n = d.jnode[j];
  NodeData nd;
  if (MODE == MODE_STEP) load_node_data<SO3>(d, n, nd);
  double f[6];
  joint_force(d, j, f);
  if (MODE == MODE_FORCE) {
#pragma unroll
    for (int i = 0; i < 6; i++) d.fout[i * d.N + n] = f[i];
    return;
  }
  if (MODE == MODE_MPRED) {
    milne_pred_node<SO3>(d, m, n, f);
    return;
  }
  node_update<SO3>(d, sp, n, f, 0u, false, &nd);
}

// ------------------------------------------------------------------
// action path: gyroscopic term added to f_out (GET_INT_VECTOR with MATRIX2)
// ------------------------------------------------------------------
template <bool SO3>
__global__ __launch_bounds__(DL_BLOCK) void k_force_extras(DevData d, int gyro) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.N) return;
  const int64_t N = d.N;
  if (gyro) {
    double r[4];
    Quat q;
    load_rot<SO3>(d, (int32_t)n, r, q);
    // the gyroscopic term with the mass in use (the matrix passed with
    // GET_INT_VECTOR: the last update's when updateWhen is set)
    double J[9], J0[6], w[3];
    const double *J0d = d.Js ? d.Js : d.J0;
    for (int c = 0; c < 6; c++) J0[c] = J0d[c * N + n];
    for (int c = 0; c < 3; c++) w[c] = d.v[(3 + c) * N + n];
    node_inertia(q, J0, d.Js != nullptr, J);
    const double jw0 = J[0] * w[0] + J[1] * w[1] + J[2] * w[2];
    const double jw1 = J[3] * w[0] + J[4] * w[1] + J[5] * w[2];
    const double jw2 = J[6] * w[0] + J[7] * w[1] + J[8] * w[2];
    const double gm = d.gm[n];
    d.fout[3 * N + n] += gm * (w[1] * jw2 - w[2] * jw1);
    d.fout[4 * N + n] += gm * (w[2] * jw0 - w[0] * jw2);
    d.fout[5 * N + n] += gm * (w[0] * jw1 - w[1] * jw0);
  }
}

// ------------------------------------------------------------------
// GroupOutput reaction (GroupOutputModule.cpp:100-116,162): the model's
// internal force with LOAD_CASE "output" at the current state -- the rod
// elements with the elastic predictor from the committed plastic state
// (SpecialCosseratRodModel.cpp:762, getStress with inelastic = false) and
// the contact forces of this state's contact search (RodContactModel.cpp:
// 153-166; the default update condition searches again) -- summed over a
// node group's dof.  One thread per entry (element, end) at a group node:
// the entry's share of the dof goes to rbuf, and the step-end kernel sums
// each group's entries in a fixed order.  It runs right after the contact
// pass of a state (the next step's start, or the closing pass of a call),
// so the contact rows it adds are this state's.
// ------------------------------------------------------------------
template <bool SO3, bool PLASTIC, bool CONTACT>
__global__ __launch_bounds__(DL_BLOCK) void k_react(DevData d) {
  const int64_t k = (int64_t)blockIdx.x * DL_BLOCK + threadIdx.x;
  if (k >= d.nrent) return;
  const int32_t code = d.rent[k], e = code >> 1, end = code & 1;
  const int64_t N = d.N, E = d.E, R = d.R;
  const int32_t na = d.en0[e], nb = d.en1[e], s = d.es[e];
  double Xa[3], ua[3], Xb[3], ub[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    Xa[i] = d.X0[i * N + na]; ua[i] = d.u[i * N + na];
    Xb[i] = d.X0[i * N + nb]; ub[i] = d.u[i * N + nb];
  }
  double ra[4], rb[4];
  Quat qna, qnb, qNa, qNb;
  load_rot<SO3>(d, na, ra, qna);
  load_rot<SO3>(d, nb, rb, qnb);
  qNa.w = d.qN[e]; qNa.x = d.qN[E + e]; qNa.y = d.qN[2 * E + e]; qNa.z = d.qN[3 * E + e];
  qNb.w = d.qN[4 * E + e]; qNb.x = d.qN[5 * E + e]; qNb.y = d.qN[6 * E + e]; qNb.z = d.qN[7 * E + e];
  double strain[6], sm[6], fa[6], fb[6];
  const uint8_t fl = d.eflag[e];
  const bool mir = (fl & 4u) != 0;
  const bool pst = PLASTIC && (fl & PST_FLAG);
  elem_eps0(Xa, Xb, qNa, qNb, strain, mir);
  ElemKin kin;
  elem_kinematics(Xa, ua, Xb, ub, qmul(qna, qNa), qmul(qnb, qNb), kin);
  if (mir) elem_mirror(kin);
#pragma unroll
  for (int i = 0; i < 6; i++) {
    strain[i] = kin.eps[i] - strain[i];
    if (PLASTIC) strain[i] = strain[i] - (pst ? d.pst[12 * (int64_t)e + i] : 0.0);
    sm[i] = d.C[i * R + s] * strain[i];
  }
  elem_forces(kin, sm, fa, fb);
  const int dof = d.react_dof;
  double v = end ? fb[dof] : fa[dof];
  if (CONTACT && dof < 3 && ((d.cbits[e >> 6] >> (e & 63)) & 1ull))
    v += d.fcs[(3 * end + dof) * E + e];
  d.rbuf[k] = v;
}

__global__ void k_step_inc(int32_t *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

__global__ void k_set_i32(int32_t *p, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ __launch_bounds__(DL_BLOCK) void k_zero_f64(double *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * DL_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * DL_BLOCK)
    p[i] = 0.0;
}

// ------------------------------------------------------------------
// boundary programs: the scale of every program at (clock time + dt, steps
// taken), then every program-driven record value = its base x its scale.
// One block; programs and values are few (loaded groups, supports).
// ------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bc_program(ProgDev P, double dt, const double *dtp) {
  __shared__ double scale[DL_PROG_MAXPROGS];
  const double t = *P.clock_t + (dtp ? *dtp : dt);
  const double k = (double)*P.clock_k;
  for (int32_t q = threadIdx.x; q < P.nprog; q += blockDim.x) {
    const int32_t c0 = P.cptr[q];
    scale[q] = prog_scale(P.kind[q], P.abc + 3 * q, P.code + c0, P.cptr[q + 1] - c0, t, k);
  }
  __syncthreads();
  const int64_t nb = P.nbc;
  for (int32_t j = threadIdx.x; j < P.nent; j += blockDim.x) {
    const int32_t dof = P.ent_dof[j];
    const double v = P.ent_base[j] * scale[P.ent_prog[j]];
    const int64_t at = (int64_t)(dof & 7) * nb + P.ent_rec[j];
    if (dof & 8) P.bc_a[at] = v;
    else P.bc_f[at] = v;
  }
}

// ------------------------------------------------------------------
// ImpactBC: one block per group.  R = the group's internal-force sum on its
// dof (fixed tree order), acceleration -R / w into the group's records,
// then the deferred node updates (dof free before the first output).
// ------------------------------------------------------------------
__device__ __forceinline__ double impact_group_sum(const DevData &d, const ImpDev &I, int32_t g) {
  __shared__ double red[256];
  const int32_t t0 = I.ptr[g], t1 = I.ptr[g + 1], dof = I.dof[g];
  double s = 0.0;
  for (int32_t j = t0 + (int32_t)threadIdx.x; j < t1; j += 256) s += d.fout[(int64_t)dof * d.N + I.nodes[j]];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double R = red[0];
  const double a = -R / I.w[g];
  for (int32_t j = t0 + (int32_t)threadIdx.x; j < t1; j += 256)
    I.bc_a[(int64_t)dof * I.nbc + d.bc[I.nodes[j]]] = a;
  if (threadIdx.x == 0) I.R[g] = R;
  return R;
}

template <bool SO3>
__global__ __launch_bounds__(256) void k_impact_step(DevData d, StepParams sp, ImpDev I) {
  const int32_t g = blockIdx.x;
  impact_group_sum(d, I, g);
  __syncthreads();
  const unsigned free_mask = *I.ready ? 0u : (1u << I.dof[g]);
  const int64_t N = d.N;
  for (int32_t j = I.ptr[g] + (int32_t)threadIdx.x; j < I.ptr[g + 1]; j += 256) {
    const int32_t n = I.nodes[j];
    double f[6];
#pragma unroll
    for (int c = 0; c < 6; c++) f[c] = d.fout[c * N + n];
    node_update<SO3>(d, sp, n, f, free_mask, true);
  }
}

__global__ __launch_bounds__(256) void k_impact_sum(ImpDev I, DevData d) {
  impact_group_sum(d, I, blockIdx.x);
}

__global__ void k_clock_advance(double *t, int64_t *k, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *t += dt;
    *k += 1;
  }
}

// ------------------------------------------------------------------
// COMMIT (applyDeform): per-element potential energy and dissipation
// increment, old <- curr
// ------------------------------------------------------------------
template <bool PLASTIC>
__global__ __launch_bounds__(DL_BLOCK) void k_commit(DevData d, double *epot, double *ediss) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.E) return;
  const int64_t E = d.E;
  const int32_t lo = d.es[e];
  const int64_t R = d.R;
  double C[6];
  for (int i = 0; i < 6; i++) C[i] = d.C[i * R + lo];
  double ec[6];
  rec6_load(d.ecur, e, ec);
  if (!PLASTIC) {
    double t = 0.0;
    for (int i = 0; i < 6; i++) {
      const double c = ec[i];
      t += c * (C[i] * c);
    }
    epot[e] = 0.5 * t;
    ediss[e] = 0.0;
    rec6_store(d.eo, e, ec);
    return;
  }
  // the trial plastic state is in pst_cur only where the last evaluation
  // corrected it (active); elsewhere it is the committed one
  const bool act = d.active[e] != 0;
  const uint8_t fl = d.eflag[e];
  double eo[6], epo[6], alo[7], epc[6], alc[7];
  rec6_load(d.eo, e, eo);
  if (fl & PST_FLAG) pst_load(d.pst, d.aiso, e, true, epo, alo);
  else pst_zero(epo, alo);
  if (act) pst_load(d.pst_cur, d.aiso_cur, e, true, epc, alc);
  double t = 0.0, ds = 0.0;
  for (int i = 0; i < 6; i++) {
    const double epn = act ? epc[i] : epo[i];
    const double oel = eo[i] - epo[i];
    const double cel = ec[i] - epn;
    const double dep = epn - epo[i];
    const double os = C[i] * oel, cs = C[i] * cel;
    t += cel * cs;
    ds += ((os + cs) / 2.0) * dep;
  }
  epot[e] = 0.5 * t;
  ediss[e] = (fabs(d.dflow[e]) > 2.220446049250313e-16) ? ds : 0.0;
  rec6_store(d.eo, e, ec);
  if (act) {
    pst_store(d.pst, d.aiso, e, true, epc, alc);
    if (!(fl & PST_FLAG)) d.eflag[e] = (uint8_t)(fl | PST_FLAG);
  }
  d.dflow[e] = 0.0;
}

// ------------------------------------------------------------------
// updateWhen: the mass update at the start of a step / attempt
template <bool SO3>
__global__ __launch_bounds__(DL_BLOCK) void k_mass_update(DevData d, MassUpd mu) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.N) return;
  const bool upd = *mu.force != 0 ||
                   prog_eval_expr(mu.code, mu.nops, *mu.clock_t, (double)*mu.clock_k) != 0.0;
  if (!upd) return;
  const int64_t N = d.N;
  double r[4], R[9], J[9], J0[6];
  Quat q;
  load_rot<SO3>(d, (int32_t)n, r, q);
  for (int c = 0; c < 6; c++) J0[c] = d.J0[c * N + n];
  qmat(q, R);
  spatial_inertia(R, J0, J);
  const double pk[6] = {J[0], J[4], J[8], J[1], J[2], J[5]};
  for (int c = 0; c < 6; c++) mu.Js[c * N + n] = pk[c];
}

// GET_MATRIX2: 6x6 block per node, row-major [N][36]
// ------------------------------------------------------------------
template <bool SO3>
__global__ __launch_bounds__(DL_BLOCK) void k_mass_blocks(DevData d, int lumped, double *blk) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.N) return;
  const int64_t N = d.N;
  double r[4];
  Quat q;
  load_rot<SO3>(d, (int32_t)n, r, q);
  double R[9], J[9], J0[6];
  for (int c = 0; c < 6; c++) J0[c] = d.J0[c * N + n];
  qmat(q, R);
  spatial_inertia(R, J0, J);
  double *B = blk + 36 * n;
  for (int i = 0; i < 36; i++) B[i] = 0.0;
  for (int i = 0; i < 3; i++) B[7 * i] = d.mt[n];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (!lumped || i == j) B[6 * (3 + i) + 3 + j] = J[3 * i + j];
}

// ------------------------------------------------------------------
// state conversion between SoA device state and DOF vectors [N][6]
// ------------------------------------------------------------------
template <bool SO3>
__global__ __launch_bounds__(DL_BLOCK) void k_state_io(DevData d, int to_dof, double *disp,
                                                       double *velo) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.N) return;
  const int64_t N = d.N;
  if (to_dof) {
    if (disp) {
      for (int c = 0; c < 3; c++) disp[6 * n + c] = d.u[c * N + n];
      if (SO3) {
        Quat q;
        q.w = d.rot[n]; q.x = d.rot[N + n]; q.y = d.rot[2 * N + n]; q.z = d.rot[3 * N + n];
        double t[3];
        qlog(q, t);
        for (int c = 0; c < 3; c++) disp[6 * n + 3 + c] = t[c];
      } else {
        for (int c = 0; c < 3; c++) disp[6 * n + 3 + c] = d.rot[c * N + n];
      }
    }
    if (velo)
      for (int c = 0; c < 6; c++) velo[6 * n + c] = d.v[c * N + n];
  } else {
    if (disp) {
      for (int c = 0; c < 3; c++) d.u[c * N + n] = disp[6 * n + c];
      if (SO3) {
        const Quat q = qexp(disp[6 * n + 3], disp[6 * n + 4], disp[6 * n + 5]);
        d.rot[n] = q.w; d.rot[N + n] = q.x; d.rot[2 * N + n] = q.y; d.rot[3 * N + n] = q.z;
      } else {
        for (int c = 0; c < 3; c++) d.rot[c * N + n] = disp[6 * n + 3 + c];
      }
    }
    if (velo)
      for (int c = 0; c < 6; c++) d.v[c * N + n] = velo[6 * n + c];
  }
}

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------
static inline unsigned nblk(int64_t n) { return (unsigned)((n + DL_BLOCK - 1) / DL_BLOCK); }

void launch_init_strain(const DevData &d, hipStream_t s) {
  if (d.E == 0) return;
  hipLaunchKernelGGL(k_init_strain, dim3(nblk(d.E)), dim3(DL_BLOCK), 0, s, d);
}

template <int MODE, bool SO3, int MAT, bool CONTACT>
static void ls(const DevData &d, const StepParams &sp, const PlasticParams &pp, const MilneDev &m,
               hipStream_t s) {
  hipLaunchKernelGGL((k_strut<MODE, SO3, MAT, CONTACT>), dim3(nblk(d.NL)), dim3(DL_BLOCK), 0, s,
                     d, sp, pp, m);
}

template <int MODE, bool SO3, int MAT>
static void ls_c(const DevData &d, const StepParams &sp, const PlasticParams &pp,
                 const MilneDev &m, bool contact, hipStream_t s) {
  if (contact) ls<MODE, SO3, MAT, true>(d, sp, pp, m, s);
  else ls<MODE, SO3, MAT, false>(d, sp, pp, m, s);
}

template <int MODE, bool SO3>
static void ls_m(const DevData &d, const StepParams &sp, const PlasticParams &pp,
                 const MilneDev &m, int mat, bool contact, hipStream_t s) {
  switch (mat) {
  case 0: ls_c<MODE, SO3, 0>(d, sp, pp, m, contact, s); break;
  case 1: ls_c<MODE, SO3, 1>(d, sp, pp, m, contact, s); break;
  case 2: ls_c<MODE, SO3, 2>(d, sp, pp, m, contact, s); break;
  default: ls_c<MODE, SO3, 3>(d, sp, pp, m, contact, s); break;
  }
}

template <int MODE>
static void ls_s(const DevData &d, const StepParams &sp, const PlasticParams &pp,
                 const MilneDev &m, bool so3, int mat, bool contact, hipStream_t s) {
  if (so3) ls_m<MODE, true>(d, sp, pp, m, mat, contact, s);
  else ls_m<MODE, false>(d, sp, pp, m, mat, contact, s);
}

// mat: 0 elastic, 1/2/3 elastoplastic with yield exponents all 1 / all 2 / general
void launch_strut_pp(const DevData &d, const StepParams &sp, const PlasticParams &pp, int mode,
                     bool so3, int mat, bool contact, hipStream_t s, const MilneDev *m) {
  if (d.NL == 0) return;
  const MilneDev mm = m ? *m : MilneDev{};
  switch (mode) {
  case MODE_STEP: ls_s<MODE_STEP>(d, sp, pp, mm, so3, mat, contact, s); break;
  case MODE_FORCE: ls_s<MODE_FORCE>(d, sp, pp, mm, so3, mat, contact, s); break;
  case MODE_MPRED: ls_s<MODE_MPRED>(d, sp, pp, mm, so3, mat, contact, s); break;
  default: ls_s<MODE_MCORR>(d, sp, pp, mm, so3, mat, contact, s); break;
  }
}

template <int MODE>
static void lj(const DevData &d, const StepParams &sp, const MilneDev &m, bool so3, hipStream_t s) {
  const dim3 g(nblk(d.NJ)), b(DL_BLOCK);
  const StepEnd se{};
  if (so3) hipLaunchKernelGGL((k_joint<MODE, true, false>), g, b, 0, s, d, sp, m, se);
  else hipLaunchKernelGGL((k_joint<MODE, false, false>), g, b, 0, s, d, sp, m, se);
}

// the step's joint kernel with the end-of-step sums as its first block(s)
static void lj_end(const DevData &d, const StepParams &sp, bool so3, const StepEnd &se,
                   hipStream_t s) {
  const dim3 g(nblk(d.NJ) + se.nblocks), b(DL_BLOCK);
  const MilneDev m{};
  if (so3) hipLaunchKernelGGL((k_joint<MODE_STEP, true, true>), g, b, 0, s, d, sp, m, se);
  else hipLaunchKernelGGL((k_joint<MODE_STEP, false, true>), g, b, 0, s, d, sp, m, se);
}

void launch_joint(const DevData &d, const StepParams &sp, int mode, bool so3, hipStream_t s,
                  const MilneDev *m, const StepEnd *se) {
  if (d.NJ == 0) return;
  const MilneDev mm = m ? *m : MilneDev{};
  if (se && mode == MODE_STEP) {
    lj_end(d, sp, so3, *se, s);
    return;
  }
  switch (mode) {
  case MODE_STEP: lj<MODE_STEP>(d, sp, mm, so3, s); break;
  case MODE_FORCE: lj<MODE_FORCE>(d, sp, mm, so3, s); break;
  case MODE_MPRED: lj<MODE_MPRED>(d, sp, mm, so3, s); break;
  default: lj<MODE_MCORR>(d, sp, mm, so3, s); break;
  }
}

void launch_force_extras(const DevData &d, bool so3, bool gyro, hipStream_t s) {
  if (!gyro) return;
  if (so3) hipLaunchKernelGGL(k_force_extras<true>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, 1);
  else hipLaunchKernelGGL(k_force_extras<false>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, 1);
}

StepEnd make_step_end(const DevData &d, const int32_t *rgptr, int32_t ngroups, bool diss,
                      int64_t ndb, double *diss_total, double *hist_react, double *hist_diss,
                      int32_t *step, int32_t *done, bool tail) {
  if (tail) diss = false;
  const int32_t nsums = ngroups + (diss ? 1 : 0);
  return StepEnd{(const double *)d.rbuf, rgptr, ngroups, nsums, (const double *)d.dbuf, ndb,
                 diss_total, hist_react, hist_diss, step, tail ? 1 : 0,
                 nsums <= 8 ? 1 : nsums, done};
}

void launch_step_end(const DevData &d, const int32_t *rgptr, int32_t ngroups, bool diss,
                     int64_t ndb, double *diss_total, double *hist_react, double *hist_diss,
                     int32_t *step, hipStream_t s, bool tail) {
  if (tail) diss = false;
  const int nb = ngroups + (diss ? 1 : 0);
  if (tail && nb == 0) return;
  if (nb <= 8) {
    const StepEnd se = make_step_end(d, rgptr, ngroups, diss, ndb, diss_total, hist_react,
                                     hist_diss, step, nullptr, tail);
    hipLaunchKernelGGL(k_step_end1, dim3(1), dim3(SE_NT), 0, s, se);
    return;
  }
  hipLaunchKernelGGL(k_step_end, dim3(nb), dim3(SE_NT), 0, s, (const double *)d.rbuf, rgptr,
                     ngroups, (const double *)d.dbuf, ndb, diss_total, hist_react, hist_diss,
                     (const int32_t *)step);
  if (!tail) hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(1), 0, s, step);
}

template <bool SO3, bool PLASTIC>
static void lr(const DevData &d, bool contact, hipStream_t s) {
  const dim3 g(nblk(d.nrent)), b(DL_BLOCK);
  if (contact) hipLaunchKernelGGL((k_react<SO3, PLASTIC, true>), g, b, 0, s, d);
  else hipLaunchKernelGGL((k_react<SO3, PLASTIC, false>), g, b, 0, s, d);
}

void launch_react(const DevData &d, bool so3, bool plastic, bool contact, hipStream_t s) {
  if (d.nrent == 0) return;
  if (so3) {
    if (plastic) lr<true, true>(d, contact, s);
    else lr<true, false>(d, contact, s);
  } else {
    if (plastic) lr<false, true>(d, contact, s);
    else lr<false, false>(d, contact, s);
  }
}

void launch_set_i32(int32_t *p, int32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, s, p, v);
}

void launch_zero_f64(double *p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + DL_BLOCK - 1) / DL_BLOCK, 1024);
  hipLaunchKernelGGL(k_zero_f64, dim3(g), dim3(DL_BLOCK), 0, s, p, n);
}

void launch_bc_program(const ProgDev &P, double dt, const double *dtp, hipStream_t s) {
  if (P.nprog == 0) return;
  hipLaunchKernelGGL(k_bc_program, dim3(1), dim3(256), 0, s, P, dt, dtp);
}

void launch_impact_step(const DevData &d, const StepParams &sp, const ImpDev &I, bool so3,
                        hipStream_t s) {
  if (I.ngroups == 0) return;
  if (so3) hipLaunchKernelGGL(k_impact_step<true>, dim3(I.ngroups), dim3(256), 0, s, d, sp, I);
  else hipLaunchKernelGGL(k_impact_step<false>, dim3(I.ngroups), dim3(256), 0, s, d, sp, I);
  hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, s, I.ready, 1);
}

void launch_impact_sum(const ImpDev &I, const DevData &d, hipStream_t s) {
  if (I.ngroups == 0) return;
  hipLaunchKernelGGL(k_impact_sum, dim3(I.ngroups), dim3(256), 0, s, I, d);
}

void launch_clock_advance(const ProgDev &P, double dt, hipStream_t s) {
  if (P.nprog == 0 && !P.clock_on) return;
  hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(1), 0, s, P.clock_t, P.clock_k, dt);
}

void launch_mass_update(const DevData &d, const MassUpd &mu, bool so3, hipStream_t s) {
  if (!mu.Js) return;
  if (so3) hipLaunchKernelGGL(k_mass_update<true>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, mu);
  else hipLaunchKernelGGL(k_mass_update<false>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, mu);
  hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, s, mu.force, 0);
}

void launch_commit_pl(const DevData &d, bool plastic, double *epot, double *ediss, hipStream_t s) {
  if (d.E == 0) return;
  if (plastic) hipLaunchKernelGGL(k_commit<true>, dim3(nblk(d.E)), dim3(DL_BLOCK), 0, s, d, epot, ediss);
  else hipLaunchKernelGGL(k_commit<false>, dim3(nblk(d.E)), dim3(DL_BLOCK), 0, s, d, epot, ediss);
}

void launch_mass_blocks(const DevData &d, bool so3, bool lumped, double *blocks, hipStream_t s) {
  if (so3) hipLaunchKernelGGL(k_mass_blocks<true>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d,
                              (int)lumped, blocks);
  else hipLaunchKernelGGL(k_mass_blocks<false>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d,
                          (int)lumped, blocks);
}

void launch_state_io(const DevData &d, bool so3, bool to_dof, double *disp, double *velo,
                     hipStream_t s) {
  if (so3) hipLaunchKernelGGL(k_state_io<true>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d,
                              (int)to_dof, disp, velo);
  else hipLaunchKernelGGL(k_state_io<false>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d,
                          (int)to_dof, disp, velo);
}
