// dl_device.h -- per-element / per-node FP64 math of the explicit rod step.
//
// Everything here is __host__ __device__ inline so the kernels in
// dl_kernels.hip and host-side set-up code share one definition.
//
// Rotations are unit quaternions (w, x, y, z) with R(q1*q2) = R(q1) R(q2).
// The element follows the reference's 2-node, 1-Gauss-point Simo-Reissner
// element (SpecialCosseratRodModel.cpp:694-771,899-949; Line3D.cpp:102-273)
// with the Crisfield-Jelenic interpolation evaluated in closed form:
//   Lambda_k   = exp(theta_k) LambdaN_k                         (:794-796)
//   phi_ab     = log(Lambda_a^T Lambda_b)                       (Line3D.cpp:258)
//   Lambda_r   = Lambda_a exp(phi_ab / 2)                       (Line3D.cpp:260-262)
//   psi_a,b    = log(Lambda_r^T Lambda_a,b) = -+ phi_ab / 2      (Line3D.cpp:271-272)
//   Lambda_ip  = Lambda_r exp(N . psi) = Lambda_r               (psi_ip = 0)
//   Lambda'    = Lambda_r skew(psi')  (|psi_ip| < TINY branch of expVecP, helpers.cpp:163-166)
// so Gamma = Lambda_r^T phi' - Gamma0, K = psi' - K0 with psi' = -phi/2 g_a + phi/2 g_b.
#ifndef DL_DEVICE_H
#define DL_DEVICE_H

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define DL_HD __host__ __device__ __forceinline__

struct Quat {
  double w, x, y, z;
};

DL_HD Quat qmul(const Quat a, const Quat b) {
  Quat r;
  r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
  r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
  return r;
}

// conj(a) * b
DL_HD Quat qcmul(const Quat a, const Quat b) {
  Quat r;
  r.w = a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z;
  r.x = a.w * b.x - a.x * b.w - a.y * b.z + a.z * b.y;
  r.y = a.w * b.y + a.x * b.z - a.y * b.w - a.z * b.x;
  r.z = a.w * b.z - a.x * b.y + a.y * b.x - a.z * b.w;
  return r;
}

DL_HD Quat qnormalize(Quat q) {
  const double s = 1.0 / sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  q.w *= s; q.x *= s; q.y *= s; q.z *= s;
  return q;
}

// exp of a rotation vector (expVec, helpers.cpp:139-157) as a quaternion
DL_HD Quat qexp(const double tx, const double ty, const double tz) {
  const double th = sqrt(tx * tx + ty * ty + tz * tz);
  Quat q;
  if (th == 0.0) {
    q.w = 1.0; q.x = 0.0; q.y = 0.0; q.z = 0.0;
    return q;
  }
  double s, c;
  sincos(0.5 * th, &s, &c);
  const double f = s / th;
  q.w = c; q.x = f * tx; q.y = f * ty; q.z = f * tz;
  return q;
}

// log of a quaternion as a rotation vector with angle in [0, pi]
// (logMat, helpers.cpp:99-111, picks the same branch)
DL_HD void qlog(Quat q, double v[3]) {
  if (q.w < 0.0) { q.w = -q.w; q.x = -q.x; q.y = -q.y; q.z = -q.z; }
  const double s = sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
  if (s == 0.0) { v[0] = v[1] = v[2] = 0.0; return; }
  const double f = 2.0 * atan2(s, q.w) / s;
  v[0] = f * q.x; v[1] = f * q.y; v[2] = f * q.z;
}

DL_HD void qmat(const Quat q, double R[9]) {
  const double xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
  const double xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
  const double wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
  R[0] = 1.0 - 2.0 * (yy + zz); R[1] = 2.0 * (xy - wz);       R[2] = 2.0 * (xz + wy);
  R[3] = 2.0 * (xy + wz);       R[4] = 1.0 - 2.0 * (xx + zz); R[5] = 2.0 * (yz - wx);
  R[6] = 2.0 * (xz - wy);       R[7] = 2.0 * (yz + wx);       R[8] = 1.0 - 2.0 * (xx + yy);
}

// ---------------------------------------------------------------------
// element kinematics (material strain), shared by INIT and the step
// ---------------------------------------------------------------------
struct ElemKin {
  double g0, g1, w;   // shape gradients and Gauss1 weight
  double phiP[3];     // phi' (spatial)
  double R[9];        // Lambda_ip
  double eps[6];      // material strain before subtracting the initial strain
};

// Shape gradients of Line2 with Gauss1 in arclength coordinates
// (Line3D.cpp:102-115): J = dN0/du * 0 + dN1/du * L0, g = dN/du / J, w = 2 J.
DL_HD void elem_kinematics(const double Xa[3], const double ua[3], const double Xb[3],
                           const double ub[3], const Quat qa, const Quat qb, ElemKin &k) {
#pragma clang fp contract(off)
  const double d0 = Xb[0] - Xa[0], d1 = Xb[1] - Xa[1], d2 = Xb[2] - Xa[2];
  double ll = d0 * d0;
  ll += d1 * d1;
  ll += d2 * d2;
  const double L = sqrt(ll) + 0.;
  const double J = -0.5 * 0. + 0.5 * L;
  k.g0 = -0.5 / J;
  k.g1 = 0.5 / J;
  k.w = 2. * J;
  // phi' = (X0 + u) . grads (SpecialCosseratRodModel.cpp:720-722)
  for (int i = 0; i < 3; i++) {
    double s = (Xa[i] + ua[i]) * k.g0;
    s += (Xb[i] + ub[i]) * k.g1;
    k.phiP[i] = s;
  }
  // relative rotation and its half (getRefRot_)
  Quat qr = qcmul(qa, qb);
  if (qr.w < 0.0) { qr.w = -qr.w; qr.x = -qr.x; qr.y = -qr.y; qr.z = -qr.z; }
  const double s2 = qr.x * qr.x + qr.y * qr.y + qr.z * qr.z;
  const double s = sqrt(s2);
  double phi[3] = {0.0, 0.0, 0.0};
  if (s > 0.0) {
    const double f = 2.0 * atan2(s, qr.w) / s;
    phi[0] = f * qr.x; phi[1] = f * qr.y; phi[2] = f * qr.z;
  }
  const double hn = 1.0 / sqrt(2.0 * (1.0 + qr.w));
  Quat qh;
  qh.w = (1.0 + qr.w) * hn; qh.x = qr.x * hn; qh.y = qr.y * hn; qh.z = qr.z * hn;
  const Quat qm = qmul(qa, qh);
  qmat(qm, k.R);
  // Gamma = Lambda^T phi'
  for (int i = 0; i < 3; i++) {
    double t = k.R[i] * k.phiP[0];
    t += k.R[3 + i] * k.phiP[1];
    t += k.R[6 + i] * k.phiP[2];
    k.eps[i] = t;
  }
  // K = psi' = psi_a g0 + psi_b g1, psi_a = -phi/2, psi_b = phi/2
  for (int i = 0; i < 3; i++) {
    double t = (-0.5 * phi[i]) * k.g0;
    t += (0.5 * phi[i]) * k.g1;
    k.eps[3 + i] = t;
  }
}

// A strut along -e3 gets the frame -I from the reference's construction
// ("180 deg turn == point mirroring", SpecialCosseratRodModel.cpp:651-654),
// which is no rotation.  Its elements keep identity frame quaternions and a
// flag; with Lambda = -R the kinematics give Gamma = Lambda^T phi' = -R^T phi'
// and spatial resultants Lambda sigma = -R sigma, while K = unskew(Lambda^T
// Lambda') is unchanged -- exactly the negated R and Gamma below.
DL_HD void elem_mirror(ElemKin &k) {
  for (int i = 0; i < 9; i++) k.R[i] = -k.R[i];
  for (int i = 0; i < 3; i++) k.eps[i] = -k.eps[i];
}

// Initial material strain (initStrain_, SpecialCosseratRodModel.cpp:536-569):
// the strain of the reference configuration, u = 0 and node rotations at
// the identity.  The element kernels recompute it from the frames and
// coordinates they load anyway instead of reading a stored copy (48 B per
// element per evaluation); the INIT kernel stores the same bits for the
// tables and the adjoint.
DL_HD void elem_eps0(const double Xa[3], const double Xb[3], const Quat qNa, const Quat qNb,
                     double eps0[6], bool mirror = false) {
  const double z[3] = {0.0, 0.0, 0.0};
  const Quat id = {1.0, 0.0, 0.0, 0.0};
  ElemKin k;
  elem_kinematics(Xa, z, Xb, z, qmul(id, qNa), qmul(id, qNb), k);
  for (int i = 0; i < 6; i++) eps0[i] = (mirror && i < 3) ? -k.eps[i] : k.eps[i];
}

// internal nodal forces of the element from the material stress
// fint[Idofs] += w XI sigma_spatial (SpecialCosseratRodModel.cpp:945, Line3D.cpp:141-162)
DL_HD void elem_forces(const ElemKin &k, const double sm[6], double fa[6], double fb[6]) {
  double n[3], m[3];
  for (int i = 0; i < 3; i++) {
    n[i] = k.R[3 * i] * sm[0] + k.R[3 * i + 1] * sm[1] + k.R[3 * i + 2] * sm[2];
    m[i] = k.R[3 * i] * sm[3] + k.R[3 * i + 1] * sm[4] + k.R[3 * i + 2] * sm[5];
  }
  double c[3];
  c[0] = k.phiP[1] * n[2] - k.phiP[2] * n[1];
  c[1] = k.phiP[2] * n[0] - k.phiP[0] * n[2];
  c[2] = k.phiP[0] * n[1] - k.phiP[1] * n[0];
  for (int i = 0; i < 3; i++) {
    fa[i] = k.w * (k.g0 * n[i]);
    fb[i] = k.w * (k.g1 * n[i]);
    fa[3 + i] = k.w * (-0.5 * c[i] + k.g0 * m[i]);
    fb[3 + i] = k.w * (-0.5 * c[i] + k.g1 * m[i]);
  }
}

// component DOF of elem_forces's fa and fb, by the same operations (the
// reaction rows of the step, which need one component only)
template <int DOF>
DL_HD void elem_force_dof(const ElemKin &k, const double sm[6], double *va, double *vb) {
  if (DOF < 3) {
    constexpr int i = DOF < 3 ? DOF : 0;
    const double n = k.R[3 * i] * sm[0] + k.R[3 * i + 1] * sm[1] + k.R[3 * i + 2] * sm[2];
    *va = k.w * (k.g0 * n);
    *vb = k.w * (k.g1 * n);
  } else {
    constexpr int i = DOF < 3 ? 0 : DOF - 3, i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    const double n1 = k.R[3 * i1] * sm[0] + k.R[3 * i1 + 1] * sm[1] + k.R[3 * i1 + 2] * sm[2];
    const double n2 = k.R[3 * i2] * sm[0] + k.R[3 * i2 + 1] * sm[1] + k.R[3 * i2 + 2] * sm[2];
    const double m = k.R[3 * i] * sm[3] + k.R[3 * i + 1] * sm[4] + k.R[3 * i + 2] * sm[5];
    const double c = k.phiP[i1] * n2 - k.phiP[i2] * n1;
    *va = k.w * (-0.5 * c + k.g0 * m);
    *vb = k.w * (-0.5 * c + k.g1 * m);
  }
}

// ---------------------------------------------------------------------
// elastoplastic rod material, power-sum yield family
//   f = sum_i |w_i (s_i + ck_i q_i)|^p_i - Y (1 + ci q_0),  q = -H alpha
// (ElastoPlasticRodMaterial.cpp:198-269, cutting plane)
// ---------------------------------------------------------------------
struct PlasticParams {
  double yield_level;
  double kin_coupling[6];
  double iso_coupling;
  double iso_modulus;
  double precision;
  double exponent[6];
  double kin_const[6];  // H_kin = kin_const + kin_ratio * C
  double kin_ratio;
  int32_t has_iso, has_kin, max_iter;
};

// kinematic hardening modulus of component i: the diagonal of the
// "kinematicTensor" hardening matrix (ElastoPlasticRodMaterial.cpp:76-106),
// given as kin_const + kin_ratio * C (kin_ratio = 1/100: "H = E/100 on the
// resultants"); unfused, like the oracle's set-up of the same table
DL_HD double kin_modulus(const PlasticParams &P, int i, double C) {
#pragma clang fp contract(off)
  return P.kin_const[i] + P.kin_ratio * C;
}

DL_HD double dl_sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }
// |x|^p and d|x|^p/d|x| of the power-sum yield family.  EXPK selects the
// exponent kind uniformly for all six components: 1 (sum-abs), 2 (quadratic)
// or 0 (general p_i through pow).
template <int EXPK>
DL_HD double yield_pterm(double x, double p) {
  const double ax = fabs(x);
  if (EXPK == 1) return ax;
  if (EXPK == 2) return ax * ax;
  if (p == 1.0) return ax;
  if (p == 2.0) return ax * ax;
  return pow(ax, p);
}
template <int EXPK>
DL_HD double yield_pderiv(double x, double p) {
  const double ax = fabs(x);
  if (EXPK == 1) return 1.0;
  if (EXPK == 2) return 2.0 * ax;
  if (p == 1.0) return 1.0;
  if (p == 2.0) return 2.0 * ax;
  return p * pow(ax, p - 1.0);
}

// Returns the number of cutting-plane corrections taken, or -1 when the
// iteration cap is hit.  ep[6], al[7] (iso, kin x6) in/out; stress out.
template <int EXPK>
DL_HD int return_map(const PlasticParams &P, const double C[6], const double w[6],
                     const double Hk[6], const double strain[6], double ep[6], double al[7],
                     double stress[6], double *dflow_out, int inelastic) {
  double dflow = 0.0;
  int liter = 0;
  for (;;) {
    for (int i = 0; i < 6; i++) stress[i] = C[i] * (strain[i] - ep[i]);
    double q[7];
    q[0] = -1.0 * (P.has_iso ? P.iso_modulus * al[0] : 0.0);
    for (int i = 0; i < 6; i++) q[1 + i] = -1.0 * (P.has_kin ? Hk[i] * al[1 + i] : 0.0);
    if (!inelastic) break;
    double a[6];
    double f = 0.0;
    for (int i = 0; i < 6; i++) {
      a[i] = stress[i] + (P.has_kin ? P.kin_coupling[i] * q[1 + i] : 0.0);
      f += yield_pterm<EXPK>(w[i] * a[i], P.exponent[i]);
    }
    f -= P.yield_level * (1.0 + (P.has_iso ? P.iso_coupling * q[0] : 0.0));
    if (!(liter < P.max_iter)) { *dflow_out = dflow; return -1; }
    if (f < P.precision) break;
    double gs[6], gq[7];
    for (int i = 0; i < 6; i++) {
      const double dd = yield_pderiv<EXPK>(w[i] * a[i], P.exponent[i]) * w[i] * dl_sgn(a[i]);
      gs[i] = dd;
      gq[1 + i] = P.has_kin ? dd * P.kin_coupling[i] : 0.0;
      if (fabs(stress[i]) < 2.220446049250313e-16) gs[i] = 0.0;
    }
    gq[0] = P.has_iso ? -P.yield_level * P.iso_coupling : 0.0;
    double den = 0.0;
    for (int i = 0; i < 6; i++) den += gs[i] * (C[i] * gs[i]);
    double b = 0.0;
    if (P.has_iso) b += gq[0] * (P.iso_modulus * gq[0]);
    if (P.has_kin)
      for (int i = 0; i < 6; i++) b += gq[1 + i] * (Hk[i] * gq[1 + i]);
    den += b;
    const double ddl = f / den;
    for (int i = 0; i < 6; i++) ep[i] += ddl * gs[i];
    if (P.has_iso) al[0] += ddl * gq[0];
    if (P.has_kin)
      for (int i = 0; i < 6; i++) al[1 + i] += ddl * gq[1 + i];
    dflow += ddl;
    liter++;
  }
  *dflow_out = dflow;
  return liter;
}

// ---------------------------------------------------------------------
// node: mass, gyro, acceleration with constraints
// ---------------------------------------------------------------------
// spatial rotational inertia J = R J0 R^T, J0 packed (xx, yy, zz, xy, xz, yz)
DL_HD void spatial_inertia(const double R[9], const double J0p[6], double J[9]) {
  const double J0[9] = {J0p[0], J0p[3], J0p[4], J0p[3], J0p[1], J0p[5], J0p[4], J0p[5], J0p[2]};
  double T[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      T[3 * i + j] = R[3 * i] * J0[j] + R[3 * i + 1] * J0[3 + j] + R[3 * i + 2] * J0[6 + j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      J[3 * i + j] = T[3 * i] * R[3 * j] + T[3 * i + 1] * R[3 * j + 1] + T[3 * i + 2] * R[3 * j + 2];
}

// rotational inertia of a node: J0 turned by the node rotation q, or, when
// the mass is only updated now and then (updateWhen), the spatial inertia
// of the last update, given packed like J0 (Jsp)
DL_HD void node_inertia(const Quat q, const double J0p[6], bool Jsp, double J[9]) {
  if (Jsp) {
    J[0] = J0p[0]; J[1] = J0p[3]; J[2] = J0p[4];
    J[3] = J0p[3]; J[4] = J0p[1]; J[5] = J0p[5];
    J[6] = J0p[4]; J[7] = J0p[5]; J[8] = J0p[2];
  } else {
    double R[9];
    qmat(q, R);
    spatial_inertia(R, J0p, J);
  }
}

// solve J a = f for the free components (mask bit i set = constrained with
// value c[i]) -- the reduced system of a Dirichlet-constrained block solve
DL_HD void solve3_masked(const double J[9], const double f[3], unsigned mask, const double c[3],
                         double a[3]) {
  double A[9], b[3];
  for (int i = 0; i < 3; i++) {
    b[i] = f[i];
    for (int j = 0; j < 3; j++) A[3 * i + j] = J[3 * i + j];
  }
  for (int j = 0; j < 3; j++) {
    if (!(mask & (1u << j))) continue;
    for (int i = 0; i < 3; i++) {
      if (mask & (1u << i)) continue;
      b[i] -= A[3 * i + j] * c[j];
    }
    for (int i = 0; i < 3; i++) { A[3 * i + j] = 0.0; A[3 * j + i] = 0.0; }
    A[3 * j + j] = 1.0;
    b[j] = c[j];
  }
  const double c00 = A[4] * A[8] - A[5] * A[7];
  const double c01 = A[5] * A[6] - A[3] * A[8];
  const double c02 = A[3] * A[7] - A[4] * A[6];
  const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
  const double id = 1.0 / det;
  const double c10 = A[2] * A[7] - A[1] * A[8];
  const double c11 = A[0] * A[8] - A[2] * A[6];
  const double c12 = A[1] * A[6] - A[0] * A[7];
  const double c20 = A[1] * A[5] - A[2] * A[4];
  const double c21 = A[2] * A[3] - A[0] * A[5];
  const double c22 = A[0] * A[4] - A[1] * A[3];
  a[0] = (c00 * b[0] + c10 * b[1] + c20 * b[2]) * id;
  a[1] = (c01 * b[0] + c11 * b[1] + c21 * b[2]) * id;
  a[2] = (c02 * b[0] + c12 * b[1] + c22 * b[2]) * id;
  for (int j = 0; j < 3; j++)
    if (mask & (1u << j)) a[j] = c[j];
}

// One node's central-difference update (ExplicitModule.cpp:339-392,
// LeapFrogModule.cpp:44-84): a = M^-1 (f_ext - f - gyro) with constraints,
// v += dt a, u += dt v, R <- exp(dt w) R (SO(3)) or theta += dt w (additive).
//   f   in: element (+ contact) internal force; out: plus the gyroscopic term
//   u, v, r in/out (r: quaternion w,x,y,z in SO(3) mode, rotation vector else)
//   q   the node rotation as a quaternion (SO(3): r itself; additive: exp(r))
// Nodal acceleration a = M^-1 (f_ext - f - gyro) with constraints
// (getAcce, ExplicitModule.cpp:337-354): mass at rotation q, gyroscopic term
// with velocity v.  f: internal force in, plus the gyroscopic term out.
template <bool SO3>
DL_HD void node_accel(double f[6], const Quat q, const double v[6], double mt,
                      const double J0[6], double gm, unsigned mask, const double ac[6],
                      const double fe[6], double a[6], bool Jsp = false) {
  double J[9];
  node_inertia(q, J0, Jsp, J);
  if (SO3) {
    // assembleGyro_ (SpecialCosseratRodModel.cpp:951-967): w x (J w), once per strut holding the node
    const double jw0 = J[0] * v[3] + J[1] * v[4] + J[2] * v[5];
    const double jw1 = J[3] * v[3] + J[4] * v[4] + J[5] * v[5];
    const double jw2 = J[6] * v[3] + J[7] * v[4] + J[8] * v[5];
    f[3] += gm * (v[4] * jw2 - v[5] * jw1);
    f[4] += gm * (v[5] * jw0 - v[3] * jw2);
    f[5] += gm * (v[3] * jw1 - v[4] * jw0);
  }
  double fr[6];
  for (int c = 0; c < 6; c++) fr[c] = fe[c] - f[c];
  const double imt = 1.0 / mt;
  for (int c = 0; c < 3; c++) a[c] = (mask & (1u << c)) ? ac[c] : imt * fr[c];
  if (SO3) {
    solve3_masked(J, fr + 3, mask >> 3, ac + 3, a + 3);
  } else {
    for (int c = 0; c < 3; c++) a[3 + c] = (mask & (8u << c)) ? ac[3 + c] : (1.0 / J[4 * c]) * fr[3 + c];
  }
}

// r <- r (+) dt w: SO(3) composition exp(dt w) q (quaternion r), or the
// additive rotation vector update
template <bool SO3>
DL_HD void rot_advance(double r[4], const Quat q, const double w[3], double dt) {
  if (SO3) {
    const Quat dq = qexp(dt * w[0], dt * w[1], dt * w[2]);
    const Quat qn = qnormalize(qmul(dq, q));
    r[0] = qn.w; r[1] = qn.x; r[2] = qn.y; r[3] = qn.z;
  } else {
    for (int c = 0; c < 3; c++) r[c] = r[c] + dt * w[c];
  }
}

template <bool SO3>
DL_HD void node_advance(double f[6], const Quat q, double u[3], double v[6], double r[4],
                        double mt, const double J0[6], double gm, unsigned mask,
                        const double ac[6], const double fe[6], double dt, bool Jsp = false) {
  double J[9];
  node_inertia(q, J0, Jsp, J);
  if (SO3) {
    // assembleGyro_ (SpecialCosseratRodModel.cpp:951-967): w x (J w), once per strut holding the node
    const double jw0 = J[0] * v[3] + J[1] * v[4] + J[2] * v[5];
    const double jw1 = J[3] * v[3] + J[4] * v[4] + J[5] * v[5];
    const double jw2 = J[6] * v[3] + J[7] * v[4] + J[8] * v[5];
    f[3] += gm * (v[4] * jw2 - v[5] * jw1);
    f[4] += gm * (v[5] * jw0 - v[3] * jw2);
    f[5] += gm * (v[3] * jw1 - v[4] * jw0);
  }
  double fr[6];
  for (int c = 0; c < 6; c++) fr[c] = fe[c] - f[c];
  double a[6];
  const double imt = 1.0 / mt;
  for (int c = 0; c < 3; c++) a[c] = (mask & (1u << c)) ? ac[c] : imt * fr[c];
  if (SO3) {
    solve3_masked(J, fr + 3, mask >> 3, ac + 3, a + 3);
  } else {
    for (int c = 0; c < 3; c++) a[3 + c] = (mask & (8u << c)) ? ac[3 + c] : (1.0 / J[4 * c]) * fr[3 + c];
  }
  for (int c = 0; c < 6; c++) v[c] = v[c] + dt * a[c];
  for (int c = 0; c < 3; c++) u[c] = u[c] + dt * v[c];
  if (SO3) {
    const Quat dq = qexp(dt * v[3], dt * v[4], dt * v[5]);
    const Quat qn = qnormalize(qmul(dq, q));
    r[0] = qn.w; r[1] = qn.x; r[2] = qn.y; r[3] = qn.z;
  } else {
    for (int c = 0; c < 3; c++) r[c] = r[c] + dt * v[3 + c];
  }
}

// MilneDevice predictor at one node (MilneDeviceModule.cpp:64-96):
//   a = M(q)^-1 (f_ext - f - gyro(v));  v_pre = v + dt a;  u_pre = u + dt v;
//   r_pre = r (+) dt w.  Outputs in *_p; the current state is read only.
template <bool SO3>
DL_HD void node_milne_pred(double f[6], const Quat q, const double u[3], const double v[6],
                           const double r[4], double mt, const double J0[6], double gm,
                           unsigned mask, const double ac[6], const double fe[6], double dt,
                           double u_p[3], double v_p[6], double r_p[4], bool Jsp = false) {
  double a[6];
  node_accel<SO3>(f, q, v, mt, J0, gm, mask, ac, fe, a, Jsp);
  for (int c = 0; c < 6; c++) v_p[c] = v[c] + dt * a[c];
  for (int c = 0; c < 3; c++) u_p[c] = u[c] + dt * v[c];
  for (int c = 0; c < 4; c++) r_p[c] = r[c];
  rot_advance<SO3>(r_p, q, v + 3, dt);
}

// rotation vector of a node rotation (the reference's state vector holds
// log(R), angle in [0, pi])
template <bool SO3>
DL_HD void rot_vector(const double r[4], double t[3]) {
  if (SO3) {
    Quat q;
    q.w = r[0]; q.x = r[1]; q.y = r[2]; q.z = r[3];
    qlog(q, t);
  } else {
    t[0] = r[0]; t[1] = r[1]; t[2] = r[2];
  }
}

// MilneDevice corrector at one node (MilneDeviceModule.cpp:98-112) with the
// node's contribution to the error measure (getQuality,
// ExplicitModule.cpp:490-499):
//   a = M(q)^-1 (f_ext - f_pre - gyro(v_pre))  (mass of the start state);
//   v_new = v + dt a;  u_new = u + dt v_pre;  r_new = r (+) dt w_pre.
//   e[0] = sum of squared (u_pre - u_new) over the node's dofs, e[1] the same
//   for v; translational dofs (all dofs in the additive mode) divided by the
//   length scale.  Outputs overwrite the predictor values in place.
template <bool SO3>
DL_HD void node_milne_corr(double f[6], const Quat q, const double u[3], const double v[6],
                           const double r[4], double mt, const double J0[6], double gm,
                           unsigned mask, const double ac[6], const double fe[6], double dt,
                           double ls, double u_p[3], double v_p[6], double r_p[4],
                           double e[2]) {
  double a[6], un[3], vn[6], rn[4];
  node_accel<SO3>(f, q, v_p, mt, J0, gm, mask, ac, fe, a);
  for (int c = 0; c < 6; c++) vn[c] = v[c] + dt * a[c];
  for (int c = 0; c < 3; c++) un[c] = u[c] + dt * v_p[c];
  for (int c = 0; c < 4; c++) rn[c] = r[c];
  rot_advance<SO3>(rn, q, v_p + 3, dt);
  double tp[3], tn[3];
  rot_vector<SO3>(r_p, tp);
  rot_vector<SO3>(rn, tn);
  const double rs = SO3 ? 1.0 : ls;
  double eu = 0.0, ev = 0.0;
  for (int c = 0; c < 3; c++) {
    const double du = (u_p[c] - un[c]) / ls, dr = (tp[c] - tn[c]) / rs;
    const double dv = (v_p[c] - vn[c]) / ls, dw = (v_p[3 + c] - vn[3 + c]) / rs;
    eu += du * du + dr * dr;
    ev += dv * dv + dw * dw;
  }
  e[0] = eu;
  e[1] = ev;
  for (int c = 0; c < 3; c++) u_p[c] = un[c];
  for (int c = 0; c < 6; c++) v_p[c] = vn[c];
  for (int c = 0; c < 4; c++) r_p[c] = rn[c];
}

#endif
