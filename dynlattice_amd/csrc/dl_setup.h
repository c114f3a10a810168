// dl_setup.h -- host-side set-up math shared by dl_create, dl_set_sections
// and the adjoint's mapping of its results back to the section data:
// per-(strut, node) reference frames (initRotation_,
// SpecialCosseratRodModel.cpp:574-658), element geometry (Line3D.cpp:91-115)
// and the block-lumped mass (assembleM_, SpecialCosseratRodModel.cpp:969-1031).
#ifndef DL_SETUP_H
#define DL_SETUP_H

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dl_ctx.h"
#include "dl_device.h"

// ------------------------------------------------------------------
// host math for set-up (row-major 3x3)
// ------------------------------------------------------------------
static inline void mm3(const double A[9], const double B[9], double C[9]) {
  double T[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = A[3 * i] * B[j];
      s += A[3 * i + 1] * B[3 + j];
      s += A[3 * i + 2] * B[6 + j];
      T[3 * i + j] = s;
    }
  std::memcpy(C, T, sizeof T);
}

// rotation matrix -> unit quaternion (Shepperd's method)
static inline Quat mat2quat(const double R[9]) {
  const double tr = R[0] + R[4] + R[8];
  Quat q;
  if (tr >= R[0] && tr >= R[4] && tr >= R[8]) {
    const double s = std::sqrt(1.0 + tr) * 2.0;
    q.w = 0.25 * s;
    q.x = (R[7] - R[5]) / s;
    q.y = (R[2] - R[6]) / s;
    q.z = (R[3] - R[1]) / s;
  } else if (R[0] >= R[4] && R[0] >= R[8]) {
    const double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2.0;
    q.w = (R[7] - R[5]) / s;
    q.x = 0.25 * s;
    q.y = (R[1] + R[3]) / s;
    q.z = (R[2] + R[6]) / s;
  } else if (R[4] >= R[8]) {
    const double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2.0;
    q.w = (R[2] - R[6]) / s;
    q.x = (R[1] + R[3]) / s;
    q.y = 0.25 * s;
    q.z = (R[5] + R[7]) / s;
  } else {
    const double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2.0;
    q.w = (R[3] - R[1]) / s;
    q.x = (R[2] + R[6]) / s;
    q.y = (R[5] + R[7]) / s;
    q.z = 0.25 * s;
  }
  if (q.w < 0) { q.w = -q.w; q.x = -q.x; q.y = -q.y; q.z = -q.z; }
  return qnormalize(q);
}

// initRotation_ frame for a unit tangent d: the rotation taking e3 to d
// (SpecialCosseratRodModel.cpp:641-655), -I for the 180 degree turn.
static inline void frame_from_tangent(const double d[3], double Rm[9]) {
  const double v[3] = {-1 * d[1], 1 * d[0], 0.0};
  const double cc = d[0] * 0. + d[1] * 0. + d[2] * 1.;
  for (int i = 0; i < 9; i++) Rm[i] = (i % 4 == 0) ? 1.0 : 0.0;
  if (std::fabs(cc + 1.0) > 2.220446049250313e-16) {
    const double Sv[9] = {0.0, -1 * v[2], 1 * v[1], 1 * v[2], 0.0, -1 * v[0], -1 * v[1], 1 * v[0], 0.0};
    double SS[9];
    mm3(Sv, Sv, SS);
    const double f = 1 / (1 + cc);
    for (int i = 0; i < 9; i++) Rm[i] += Sv[i] + f * SS[i];
  } else {
    for (int i = 0; i < 9; i++) Rm[i] *= -1.;
  }
}

// Gauss1 Jacobian J = L0 / 2 of the Line2 element (a, b) in arclength
// coordinates, in the reference's operation order (Line3D.cpp:91-115)
static inline double elem_jac(const double *X, int32_t a, int32_t b) {
  double dd[3] = {X[3 * b] - X[3 * a], X[3 * b + 1] - X[3 * a + 1], X[3 * b + 2] - X[3 * a + 2]};
  double ll = dd[0] * dd[0];
  ll += dd[1] * dd[1];
  ll += dd[2] * dd[2];
  const double Lc = std::sqrt(ll) + 0.;
  return -0.5 * 0. + 0.5 * Lc;
}

// Walks every strut's elements with the frames of their two end nodes:
// a strut's rod nodes get the averaged element tangent and the frame turning
// e3 onto it; fn(s, e, a, b, La, Lb) is called per element in strut order.
template <class F>
static void walk_elements(const MeshHost &m, F fn) {
  std::vector<double> dirs, lam;
  std::vector<int32_t> lnode;
  const double *X = m.X.data();
  const int32_t *en = m.en.data();
  for (int32_t s = 0; s < m.R; s++) {
    const int64_t c0 = m.eptr[s] + s, c1 = m.eptr[s + 1] + s;
    lnode.clear();
    for (int64_t p = c0; p <= c1; p++)
      if (std::find(lnode.begin(), lnode.end(), m.cnode[p]) == lnode.end()) lnode.push_back(m.cnode[p]);
    const size_t nl = lnode.size();
    dirs.assign(3 * nl, 0.0);
    lam.assign(9 * nl, 0.0);
    auto lid = [&](int32_t n) { return (size_t)(std::find(lnode.begin(), lnode.end(), n) - lnode.begin()); };
    for (int64_t e = m.eptr[s]; e < m.eptr[s + 1]; e++) {
      const int32_t a = en[2 * e], b = en[2 * e + 1];
      const size_t la = lid(a), lb = lid(b);
      for (int k = 0; k < 3; k++) {
        const double t = X[3 * b + k] - X[3 * a + k];
        dirs[3 * la + k] += t;
        dirs[3 * lb + k] += t;
      }
    }
    for (size_t k = 0; k < nl; k++) {
      double *dd = &dirs[3 * k];
      const double nn = std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
      dd[0] = dd[0] / nn; dd[1] = dd[1] / nn; dd[2] = dd[2] / nn;
      frame_from_tangent(dd, &lam[9 * k]);
    }
    for (int64_t e = m.eptr[s]; e < m.eptr[s + 1]; e++) {
      const int32_t a = en[2 * e], b = en[2 * e + 1];
      fn(s, e, a, b, (const double *)&lam[9 * lid(a)], (const double *)&lam[9 * lid(b)]);
    }
  }
}

// Gauss2 points of the translational mass row sums
static inline void mass_shapes(double sh[2][2]) {
  const double xi[2] = {-1. / std::sqrt(3.), 1. / std::sqrt(3.)};
  for (int ip = 0; ip < 2; ip++) {
    sh[0][ip] = 0.5 * (1. - xi[ip]);
    sh[1][ip] = 0.5 * (1. + xi[ip]);
  }
}

// per-strut mass densities Mm = rho (A, A, A, c I1, c I2, c J), c = inertia correction
static inline void mass_density(const SecHost &sc, int32_t s, double Mm[6]) {
  const double rho = sc.rho[s], A = sc.area[s];
  Mm[0] = rho * A; Mm[1] = rho * A; Mm[2] = rho * A;
  Mm[3] = rho * sc.I1[s]; Mm[4] = rho * sc.I2[s]; Mm[5] = rho * sc.J[s];
  if (sc.inertia_correct != 1.0)
    for (int i = 3; i < 6; i++) Mm[i] *= sc.inertia_correct;
}

// One element's lumped mass (assembleM_): translational row sums into mt of
// both end nodes, rotational L diag(Mr) L^T / 2 into the packed J0 (xx, yy,
// zz, xy, xz, yz; [6][N]) of both end nodes.
static inline void add_elem_mass(const SecHost &sc, int32_t s, double Jc, int32_t a, int32_t b,
                                 const double *La, const double *Lb, int64_t N, double *mt,
                                 double *J0) {
  double Mm[6];
  mass_density(sc, s, Mm);
  const double wts = 1. * Jc;
  const double l = (wts + wts) / (2 - 1);
  double sh[2][2];
  mass_shapes(sh);
  const int32_t nodes2[2] = {a, b};
  for (int in = 0; in < 2; in++) {
    double acc = 0.0;
    for (int jn = 0; jn < 2; jn++)
      for (int ip = 0; ip < 2; ip++) acc += wts * sh[in][ip] * sh[jn][ip] * Mm[0];
    mt[nodes2[in]] += acc;
  }
  double Mr[3] = {Mm[3] * l, Mm[4] * l, Mm[5] * l};
  const double add = sc.area[s] * sc.rho[s] * std::pow(l, 3) / 12.;
  Mr[0] += add;
  Mr[1] += add;
  for (int in = 0; in < 2; in++) {
    const double *L = in == 0 ? La : Lb;
    double Jm[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        Jm[3 * i + j] = (L[3 * i] * Mr[0] * L[3 * j] + L[3 * i + 1] * Mr[1] * L[3 * j + 1] +
                         L[3 * i + 2] * Mr[2] * L[3 * j + 2]) / 2.;
    const int32_t n = nodes2[in];
    J0[0 * N + n] += Jm[0];
    J0[1 * N + n] += Jm[4];
    J0[2 * N + n] += Jm[8];
    J0[3 * N + n] += Jm[1];
    J0[4 * N + n] += Jm[2];
    J0[5 * N + n] += Jm[5];
  }
}

#endif
