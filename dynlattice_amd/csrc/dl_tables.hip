// dl_tables.hip -- GET_TABLE-style observables at the current state (FP64):
// per-element strain / stress tables and potential energy, per-node kinetic
// energy.  Not on the step's path: one pass over the elements or nodes per
// request.
//
// Element tables (SpecialCosseratRodModel.cpp:221-262, 426-531):
//   mat_strain  (Gamma, K) - initial strain, material frame
//   strain      PI (Gamma, K)                      PI = blockdiag(L_ip, L_ip)
//   mat_stress  material stress: elastic C eps; elastoplastic the return map
//               from the committed plastic state (the default load case of
//               getStresses_, SpecialCosseratRodModel.h:288-295), without
//               changing the committed state
//   stress      PI mat_stress
// Potential energy per element: w 1/2 (eps - eps_p) . sigma of that
// evaluation (ElastoPlasticRodMaterial.cpp:291-292 at the current state).
// Kinetic energy per node: 1/2 (m_t |v|^2 + w . J w), J = R J0 R^T
// (the block-lumped mass; additive mode: its diagonal).
#include <hip/hip_runtime.h>

#include "dl_internal.h"

#define DL_BLOCK 256

namespace {

template <bool SO3>
__device__ __forceinline__ Quat node_rot(const DevData &d, int32_t n) {
  const int64_t N = d.N;
  if (SO3) {
    Quat q;
    q.w = d.rot[n]; q.x = d.rot[N + n]; q.y = d.rot[2 * N + n]; q.z = d.rot[3 * N + n];
    return q;
  }
  return qexp(d.rot[n], d.rot[N + n], d.rot[2 * N + n]);
}

// kind: 0 strain, 1 mat_strain, 2 stress, 3 mat_stress; out [E][6] or null;
// epot [E] or null
template <bool SO3, int MAT>
__global__ __launch_bounds__(DL_BLOCK) void k_elem_table(DevData d, PlasticParams pp, int kind,
                                                         double *out, double *epot) {
  const int64_t e = (int64_t)blockIdx.x * DL_BLOCK + threadIdx.x;
  if (e >= d.E) return;
  const int64_t N = d.N, E = d.E, R = d.R;
  const int32_t na = d.en0[e], nb = d.en1[e], s = d.es[e];
  double Xa[3], ua[3], Xb[3], ub[3];
  for (int i = 0; i < 3; i++) {
    Xa[i] = d.X0[i * N + na]; ua[i] = d.u[i * N + na];
    Xb[i] = d.X0[i * N + nb]; ub[i] = d.u[i * N + nb];
  }
  Quat qNa, qNb;
  qNa.w = d.qN[e]; qNa.x = d.qN[E + e]; qNa.y = d.qN[2 * E + e]; qNa.z = d.qN[3 * E + e];
  qNb.w = d.qN[4 * E + e]; qNb.x = d.qN[5 * E + e]; qNb.y = d.qN[6 * E + e]; qNb.z = d.qN[7 * E + e];
  ElemKin k;
  elem_kinematics(Xa, ua, Xb, ub, qmul(node_rot<SO3>(d, na), qNa), qmul(node_rot<SO3>(d, nb), qNb), k);
  if (d.eflag[e] & 4u) elem_mirror(k); // frames -I (strut along -e3)
  double strain[6], C[6], sm[6], ep[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < 6; i++) {
    strain[i] = k.eps[i] - d.eps0[i * E + e];
    C[i] = d.C[i * R + s];
  }
  if (MAT == 0) {
    for (int i = 0; i < 6; i++) sm[i] = C[i] * strain[i];
  } else {
    double pw[6], Hk[6], al[7], dfl;
    for (int i = 0; i < 6; i++) {
      pw[i] = d.pw[i * R + s];
      Hk[i] = kin_modulus(pp, i, C[i]);
    }
    if (d.eflag[e] & PST_FLAG) pst_load(d.pst, d.aiso, e, pp.has_iso, ep, al);
    else pst_zero(ep, al);
    if (return_map<MAT == 3 ? 0 : MAT>(pp, C, pw, Hk, strain, ep, al, sm, &dfl, 1) < 0) *d.err = -3;
  }
  if (epot) {
    double t = 0.0;
    for (int i = 0; i < 6; i++) t += (strain[i] - ep[i]) * sm[i];
    epot[e] = k.w * (0.5 * t);
  }
  if (!out) return;
  const double *v = (kind <= 1) ? strain : sm;
  double *o = out + 6 * e;
  if (kind == 1 || kind == 3) {
    for (int i = 0; i < 6; i++) o[i] = v[i];
    return;
  }
  for (int i = 0; i < 3; i++) {
    o[i] = k.R[3 * i] * v[0] + k.R[3 * i + 1] * v[1] + k.R[3 * i + 2] * v[2];
    o[3 + i] = k.R[3 * i] * v[3] + k.R[3 * i + 1] * v[4] + k.R[3 * i + 2] * v[5];
  }
}

template <bool SO3>
__global__ __launch_bounds__(DL_BLOCK) void k_node_ekin(DevData d, double *ekin) {
  const int64_t n = (int64_t)blockIdx.x * DL_BLOCK + threadIdx.x;
  if (n >= d.N) return;
  const int64_t N = d.N;
  double R[9], J[9], J0[6], v[6];
  for (int c = 0; c < 6; c++) {
    J0[c] = d.J0[c * N + n];
    v[c] = d.v[c * N + n];
  }
  qmat(node_rot<SO3>(d, (int32_t)n), R);
  spatial_inertia(R, J0, J);
  double t = d.mt[n] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (SO3) {
    for (int i = 0; i < 3; i++) {
      const double jw = J[3 * i] * v[3] + J[3 * i + 1] * v[4] + J[3 * i + 2] * v[5];
      t += v[3 + i] * jw;
    }
  } else {
    for (int i = 0; i < 3; i++) t += v[3 + i] * (J[4 * i] * v[3 + i]);
  }
  ekin[n] = 0.5 * t;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + DL_BLOCK - 1) / DL_BLOCK); }

template <bool SO3>
void elem_table_m(const DevData &d, const PlasticParams &pp, int mat, int kind, double *out,
                  double *epot, hipStream_t s) {
  const dim3 g(nblk(d.E)), b(DL_BLOCK);
  switch (mat) {
  case 0: hipLaunchKernelGGL((k_elem_table<SO3, 0>), g, b, 0, s, d, pp, kind, out, epot); break;
  case 1: hipLaunchKernelGGL((k_elem_table<SO3, 1>), g, b, 0, s, d, pp, kind, out, epot); break;
  case 2: hipLaunchKernelGGL((k_elem_table<SO3, 2>), g, b, 0, s, d, pp, kind, out, epot); break;
  default: hipLaunchKernelGGL((k_elem_table<SO3, 3>), g, b, 0, s, d, pp, kind, out, epot); break;
  }
}

} // namespace

void launch_elem_table(const DevData &d, const PlasticParams &pp, bool so3, int mat, int kind,
                       double *out, double *epot, hipStream_t s) {
  if (d.E == 0) return;
  if (so3) elem_table_m<true>(d, pp, mat, kind, out, epot, s);
  else elem_table_m<false>(d, pp, mat, kind, out, epot, s);
}

void launch_node_ekin(const DevData &d, bool so3, double *ekin, hipStream_t s) {
  if (d.N == 0) return;
  if (so3) hipLaunchKernelGGL(k_node_ekin<true>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, ekin);
  else hipLaunchKernelGGL(k_node_ekin<false>, dim3(nblk(d.N)), dim3(DL_BLOCK), 0, s, d, ekin);
}
