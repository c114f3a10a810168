// dl_adjoint.h -- reverse-mode (vector-Jacobian) products of the explicit
// step, written by hand: element internal force, quaternion maps, node
// update (block mass solve, gyroscopic term, SO(3) composition).  Used by the
// adjoint kernels (dl_adjoint.hip); __host__ __device__ so that the same code
// is checked against finite differences on the host.
//
// Quaternion adjoint rules (Hamilton product, L/R multiplication matrices):
//   p = a * b   ->   abar += pbar * conj(b),   bbar += conj(a) * pbar
//   p = conj(a) * b -> abar += b * conj(pbar), bbar += a * pbar
#ifndef DL_ADJOINT_H
#define DL_ADJOINT_H

#include "dl_contact_math.h"
#include "dl_device.h"
#include "dl_dual.h"

DL_HD Quat qconj(const Quat a) {
  Quat r;
  r.w = a.w; r.x = -a.x; r.y = -a.y; r.z = -a.z;
  return r;
}
DL_HD Quat qadd(const Quat a, const Quat b) {
  Quat r;
  r.w = a.w + b.w; r.x = a.x + b.x; r.y = a.y + b.y; r.z = a.z + b.z;
  return r;
}

// R = qmat(q): qbar += dR/dq : Rbar
DL_HD Quat qmat_vjp(const Quat q, const double Rb[9]) {
  const double w = q.w, x = q.x, y = q.y, z = q.z;
  Quat g;
  g.w = 2.0 * (-z * Rb[1] + y * Rb[2] + z * Rb[3] - x * Rb[5] - y * Rb[6] + x * Rb[7]);
  g.x = 2.0 * (y * Rb[1] + z * Rb[2] + y * Rb[3] - 2.0 * x * Rb[4] - w * Rb[5] + z * Rb[6] +
               w * Rb[7] - 2.0 * x * Rb[8]);
  g.y = 2.0 * (-2.0 * y * Rb[0] + x * Rb[1] + w * Rb[2] + x * Rb[3] + z * Rb[5] - w * Rb[6] +
               z * Rb[7] - 2.0 * y * Rb[8]);
  g.z = 2.0 * (-2.0 * z * Rb[0] - w * Rb[1] + x * Rb[2] + w * Rb[3] - 2.0 * z * Rb[4] +
               y * Rb[5] + x * Rb[6] + y * Rb[7]);
  return g;
}

// q = qexp(t): tbar from qbar
DL_HD void qexp_vjp(const double t[3], const Quat qb, double tb[3]) {
  const double th2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
  const double th = sqrt(th2);
  double f, g, s0; // f = sin(th/2)/th, g = f'(th)/th, s0 = dq0/dth / th = -sin(th/2)/(2 th)
  if (th < 1e-4) {
    f = 0.5 - th2 / 48.0;
    g = -1.0 / 24.0 + th2 / 960.0;
    s0 = -0.25 + th2 / 96.0;
  } else {
    double sh, ch;
    sincos(0.5 * th, &sh, &ch);
    f = sh / th;
    g = (0.5 * ch * th - sh) / (th2 * th);
    s0 = -0.5 * sh / th;
  }
  const double dv = t[0] * qb.x + t[1] * qb.y + t[2] * qb.z;
  const double qv[3] = {qb.x, qb.y, qb.z};
#pragma unroll
  for (int i = 0; i < 3; i++) tb[i] = qb.w * s0 * t[i] + f * qv[i] + g * dv * t[i];
}

// element internal force (same computation as elem_kinematics + elem_forces)
// and its VJP, in three parts so that the elastic and the elastoplastic
// material share them: the forward intermediates (elem_vjp_forward), the
// adjoint of one material stress from the adjoint of the nodal force it
// makes (elem_vjp_stress; the force is linear in the stress, so several
// stress sources add up), and the adjoint of the strain back to the node
// displacements and quaternions (elem_vjp_strain).
struct ElemVJP {
  double ua[3], ub[3];
  Quat qa, qb;
  double C[6];
};

struct ElemFwd {
  ElemKin k;
  Quat qa, qb, qr, qh, qm;
  double sgn, s, s2, hn;
  double strain[6]; // material strain minus the initial strain
};

DL_HD void elem_vjp_forward(const double Xa[3], const double ua[3], const Quat qna, const Quat qNa,
                            const double Xb[3], const double ub[3], const Quat qnb, const Quat qNb,
                            const double eps0[6], ElemFwd &F) {
  F.qa = qmul(qna, qNa);
  F.qb = qmul(qnb, qNb);
  elem_kinematics(Xa, ua, Xb, ub, F.qa, F.qb, F.k);
  // intermediate quantities of the rotation interpolation
  Quat qr = qcmul(F.qa, F.qb);
  F.sgn = (qr.w < 0.0) ? -1.0 : 1.0;
  qr.w *= F.sgn; qr.x *= F.sgn; qr.y *= F.sgn; qr.z *= F.sgn;
  F.qr = qr;
  F.s2 = qr.x * qr.x + qr.y * qr.y + qr.z * qr.z;
  F.s = sqrt(F.s2);
  F.hn = 1.0 / sqrt(2.0 * (1.0 + qr.w));
  F.qh.w = (1.0 + qr.w) * F.hn; F.qh.x = qr.x * F.hn; F.qh.y = qr.y * F.hn; F.qh.z = qr.z * F.hn;
  F.qm = qmul(F.qa, F.qh);
#pragma unroll
  for (int i = 0; i < 6; i++) F.strain[i] = F.k.eps[i] - eps0[i];
}

// force (fa, fb) = XI^T (R sm) w: given its adjoint (fba, fbb), the adjoint
// of the material stress sm (smb) and the accumulations into the adjoints
// of Lambda (Rb) and phi' (pPb)
DL_HD void elem_vjp_stress(const ElemFwd &F, const double sm[6], const double fba[6],
                           const double fbb[6], double Rb[9], double pPb[3], double smb[6]) {
  const ElemKin &k = F.k;
  const double *R = k.R;
  double n[3];
#pragma unroll
  for (int i = 0; i < 3; i++) n[i] = R[3 * i] * sm[0] + R[3 * i + 1] * sm[1] + R[3 * i + 2] * sm[2];
  double nb[3], mb[3], cb[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    nb[i] = k.w * k.g0 * fba[i] + k.w * k.g1 * fbb[i];
    mb[i] = k.w * k.g0 * fba[3 + i] + k.w * k.g1 * fbb[3 + i];
    cb[i] = -0.5 * k.w * (fba[3 + i] + fbb[3 + i]);
  }
  // c = phiP x n
  pPb[0] += n[1] * cb[2] - n[2] * cb[1];
  pPb[1] += n[2] * cb[0] - n[0] * cb[2];
  pPb[2] += n[0] * cb[1] - n[1] * cb[0];
  nb[0] += cb[1] * k.phiP[2] - cb[2] * k.phiP[1];
  nb[1] += cb[2] * k.phiP[0] - cb[0] * k.phiP[2];
  nb[2] += cb[0] * k.phiP[1] - cb[1] * k.phiP[0];
  // n = R sm_t, m = R sm_r
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Rb[3 * i + j] += nb[i] * sm[j] + mb[i] * sm[3 + j];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    smb[j] = R[j] * nb[0] + R[3 + j] * nb[1] + R[6 + j] * nb[2];
    smb[3 + j] = R[j] * mb[0] + R[3 + j] * mb[1] + R[6 + j] * mb[2];
  }
}

// strain = (R^T phiP, phi cL) - eps0: from its adjoint stb (and Rb, pPb
// accumulated by elem_vjp_stress) the adjoints of the node displacements
// and quaternions (out.ua, ub, qa, qb)
DL_HD void elem_vjp_strain(const ElemFwd &F, const double stb[6], double Rb[9], double pPb[3],
                           const Quat qNa, const Quat qNb, ElemVJP &out) {
  const ElemKin &k = F.k;
  const double *R = k.R;
  const Quat qr = F.qr;
  const double hn = F.hn, s = F.s, s2 = F.s2;
  // eps_t = R^T phiP
#pragma unroll
  for (int kk = 0; kk < 3; kk++)
#pragma unroll
    for (int i = 0; i < 3; i++) Rb[3 * kk + i] += k.phiP[kk] * stb[i];
#pragma unroll
  for (int kk = 0; kk < 3; kk++)
    pPb[kk] += R[3 * kk] * stb[0] + R[3 * kk + 1] * stb[1] + R[3 * kk + 2] * stb[2];
  // eps_r = phi * cL, cL = -0.5 g0 + 0.5 g1
  const double cL = -0.5 * k.g0 + 0.5 * k.g1;
  double phib[3] = {cL * stb[3], cL * stb[4], cL * stb[5]};
  // R = qmat(qm); qm = qa * qh
  const Quat qmb = qmat_vjp(F.qm, Rb);
  Quat qab = qmul(qmb, qconj(F.qh));
  const Quat qhb = qmul(qconj(F.qa), qmb);
  // qh = ((1+w) hn, v hn), hn = (2(1+w))^-1/2
  double wb = qhb.w * (0.5 * hn) - (qhb.x * qr.x + qhb.y * qr.y + qhb.z * qr.z) * hn * hn * hn;
  double vb[3] = {hn * qhb.x, hn * qhb.y, hn * qhb.z};
  // phi = f(s, w) v, f = 2 atan2(s, w) / s
  {
    const double w = qr.w;
    const double r2 = s2 + w * w;
    double f, gs; // gs = (df/ds) / s
    if (s < 1e-4 * fabs(w)) {
      const double x2 = s2 / (w * w);
      f = 2.0 / w * (1.0 - x2 / 3.0);
      gs = (-4.0 / 3.0 + 1.6 * x2) / (w * w * w);
    } else {
      const double at = atan2(s, w);
      f = 2.0 * at / s;
      gs = 2.0 * (s * w / r2 - at) / (s2 * s);
    }
    const double dfdw = -2.0 / r2;
    const double vdp = qr.x * phib[0] + qr.y * phib[1] + qr.z * phib[2];
    vb[0] += f * phib[0] + gs * vdp * qr.x;
    vb[1] += f * phib[1] + gs * vdp * qr.y;
    vb[2] += f * phib[2] + gs * vdp * qr.z;
    wb += vdp * dfdw;
  }
  // qr = sgn * conj(qa) * qb
  Quat qrb;
  qrb.w = F.sgn * wb; qrb.x = F.sgn * vb[0]; qrb.y = F.sgn * vb[1]; qrb.z = F.sgn * vb[2];
  qab = qadd(qab, qmul(F.qb, qconj(qrb)));
  const Quat qbb = qmul(F.qa, qrb);
  // qa = qna * qNa
  out.qa = qmul(qab, qconj(qNa));
  out.qb = qmul(qbb, qconj(qNb));
#pragma unroll
  for (int i = 0; i < 3; i++) {
    out.ua[i] = k.g0 * pPb[i];
    out.ub[i] = k.g1 * pPb[i];
  }
}

// elastic material: inputs node displacements, node quaternions (the nodal
// rotation state), constant frames/strain; given fbar_a, fbar_b (6 each)
// returns ubar_a/b, qbar_a/b (node quaternions) and Cbar (material stiffness)
DL_HD void elem_force_vjp(const double Xa[3], const double ua[3], const Quat qna, const Quat qNa,
                          const double Xb[3], const double ub[3], const Quat qnb, const Quat qNb,
                          const double eps0[6], const double C[6], const double fba[6],
                          const double fbb[6], ElemVJP &out) {
  ElemFwd F;
  elem_vjp_forward(Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, F);
  double sm[6], smb[6], stb[6];
#pragma unroll
  for (int i = 0; i < 6; i++) sm[i] = C[i] * F.strain[i];
  double Rb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, pPb[3] = {0, 0, 0};
  elem_vjp_stress(F, sm, fba, fbb, Rb, pPb, smb);
#pragma unroll
  for (int i = 0; i < 6; i++) {
    out.C[i] = smb[i] * F.strain[i];
    stb[i] = C[i] * smb[i];
  }
  elem_vjp_strain(F, stb, Rb, pPb, qNa, qNb, out);
}

// ---------------------------------------------------------------------
// elastoplastic material: the cutting plane of return_map reversed
// ---------------------------------------------------------------------
// second derivative g''(|x|) of the yield family's term g(|x|) = |x|^p
template <int EXPK>
DL_HD double yield_pderiv2(double x, double p) {
  const double ax = fabs(x);
  if (EXPK == 1) return 0.0;
  if (EXPK == 2) return 2.0;
  if (p == 1.0) return 0.0;
  if (p == 2.0) return 2.0;
  return ax > 0.0 ? p * (p - 1.0) * pow(ax, p - 2.0) : 0.0;
}

// Parameter adjoints of the material: stiffness, yield weights, kinematic
// moduli (accumulated).
struct MatBar {
  double C[6], w[6], Hk[6];
};

// One correction of return_map<EXPK> taken at the plastic state (ep, al),
// reversed.  epb, alb: adjoints of the corrected state on entry, of (ep, al)
// on return; strb and the parameter adjoints accumulate.  The forward part
// repeats return_map's operations for one pass of its loop.
template <int EXPK>
DL_HD void rm_correction_vjp(const PlasticParams &P, const double C[6], const double w[6],
                             const double Hk[6], const double strain[6], const double ep[6],
                             const double al[7], double epb[6], double alb[7], double strb[6],
                             MatBar &mb) {
  const double tiny = 2.220446049250313e-16;
  double st[6], q[7], a[6];
  q[0] = -1.0 * (P.has_iso ? P.iso_modulus * al[0] : 0.0);
  double f = 0.0;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    st[i] = C[i] * (strain[i] - ep[i]);
    q[1 + i] = -1.0 * (P.has_kin ? Hk[i] * al[1 + i] : 0.0);
    a[i] = st[i] + (P.has_kin ? P.kin_coupling[i] * q[1 + i] : 0.0);
    f += yield_pterm<EXPK>(w[i] * a[i], P.exponent[i]);
  }
  f -= P.yield_level * (1.0 + (P.has_iso ? P.iso_coupling * q[0] : 0.0));
  double gs[6], gq[7];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const double dd = yield_pderiv<EXPK>(w[i] * a[i], P.exponent[i]) * w[i] * dl_sgn(a[i]);
    gs[i] = dd;
    gq[1 + i] = P.has_kin ? dd * P.kin_coupling[i] : 0.0;
    if (fabs(st[i]) < tiny) gs[i] = 0.0;
  }
  gq[0] = P.has_iso ? -P.yield_level * P.iso_coupling : 0.0;
  double den = 0.0;
#pragma unroll
  for (int i = 0; i < 6; i++) den += gs[i] * (C[i] * gs[i]);
  double b = 0.0;
  if (P.has_iso) b += gq[0] * (P.iso_modulus * gq[0]);
  if (P.has_kin)
#pragma unroll
    for (int i = 0; i < 6; i++) b += gq[1 + i] * (Hk[i] * gq[1 + i]);
  den += b;
  const double ddl = f / den;
  // ---- reverse: ep' = ep + ddl gs, al' = al + ddl gq
  double lb = 0.0;
#pragma unroll
  for (int i = 0; i < 6; i++) lb += epb[i] * gs[i];
  if (P.has_iso) lb += alb[0] * gq[0];
  if (P.has_kin)
#pragma unroll
    for (int i = 0; i < 6; i++) lb += alb[1 + i] * gq[1 + i];
  // ddl = f / den
  const double fb = lb / den, denb = -lb * ddl / den;
  double q0b = P.has_iso ? -P.yield_level * P.iso_coupling * fb : 0.0;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    // den = sum gs C gs + sum gq Hk gq (+ the isotropic constant)
    const double gsb = ddl * epb[i] + 2.0 * C[i] * gs[i] * denb;
    mb.C[i] += gs[i] * gs[i] * denb;
    double gqb = 0.0;
    if (P.has_kin) {
      gqb = ddl * alb[1 + i] + 2.0 * Hk[i] * gq[1 + i] * denb;
      mb.Hk[i] += gq[1 + i] * gq[1 + i] * denb;
    }
    // gs = dd (zeroed where |stress| is tiny), gq = dd ck;
    // dd = g'(|x|) w sgn(a), x = w a, d|x| = sgn(x) (w da + a dw)
    const double ddb = (fabs(st[i]) < tiny ? 0.0 : gsb) + (P.has_kin ? P.kin_coupling[i] * gqb : 0.0);
    const double x = w[i] * a[i], sx = dl_sgn(x), sa = dl_sgn(a[i]);
    const double g1 = yield_pderiv<EXPK>(x, P.exponent[i]);
    const double g2 = yield_pderiv2<EXPK>(x, P.exponent[i]);
    // f = sum g(|x|) - ...
    const double ab = ddb * g2 * sx * w[i] * w[i] * sa + fb * g1 * sx * w[i];
    mb.w[i] += ddb * (g2 * sx * a[i] * w[i] * sa + g1 * sa) + fb * g1 * sx * a[i];
    // a = st + ck q_kin, q_kin = -Hk al_kin, st = C (strain - ep)
    if (P.has_kin) {
      const double qkb = P.kin_coupling[i] * ab;
      alb[1 + i] += -Hk[i] * qkb;
      mb.Hk[i] += -al[1 + i] * qkb;
    }
    mb.C[i] += ab * (strain[i] - ep[i]);
    strb[i] += C[i] * ab;
    epb[i] -= C[i] * ab;
  }
  if (P.has_iso) alb[0] += -P.iso_modulus * q0b;
}

// Reverse of return_map<EXPK>(..., inelastic = 1) from the committed state
// (ep0, al0) that took n >= 0 corrections and ended at plastic strain epn:
// given the adjoint of the returned stress (sb) and of the returned state
// (epb, alb; on return the adjoints of ep0, al0), accumulates the adjoints
// of the strain and of the material parameters.  The state before
// correction k is recomputed by return_map itself capped at k corrections,
// so it has the forward's bits.
template <int EXPK>
DL_HD void return_map_vjp(const PlasticParams &P, const double C[6], const double w[6],
                          const double Hk[6], const double strain[6], const double ep0[6],
                          const double al0[7], int n, const double epn[6], const double sb[6],
                          double epb[6], double alb[7], double strb[6], MatBar &mb) {
  // stress = C (strain - ep_n)
#pragma unroll
  for (int i = 0; i < 6; i++) {
    mb.C[i] += sb[i] * (strain[i] - epn[i]);
    strb[i] += C[i] * sb[i];
    epb[i] -= C[i] * sb[i];
  }
  PlasticParams Pk = P;
#pragma unroll 1
  for (int k = n - 1; k >= 0; k--) {
    double ep[6], al[7], st[6], dfl;
#pragma unroll
    for (int i = 0; i < 6; i++) ep[i] = ep0[i];
#pragma unroll
    for (int i = 0; i < 7; i++) al[i] = al0[i];
    Pk.max_iter = k;
    return_map<EXPK>(Pk, C, w, Hk, strain, ep, al, st, &dfl, 1);
    rm_correction_vjp<EXPK>(P, C, w, Hk, strain, ep, al, epb, alb, strb, mb);
  }
}

// Element force of the elastoplastic material, reversed.  Two stress
// sources share the kinematics:
//  - the step's internal force (unless seed): the return map from the
//    committed state (ep0, al0), with force adjoint (fba, fbb);
//  - the GroupOutput rows (when has_r): the output load case, the elastic
//    predictor C (strain - ep0), with force adjoint (rba, rbb).
// epb, alb: on entry the adjoints of the plastic state the step commits
// (ignored with seed), on return those of the committed state (ep0, al0).
// Returns the number of corrections, or -1 at the iteration cap.
template <int EXPK>
DL_HD int elem_force_vjp_pl(const double Xa[3], const double ua[3], const Quat qna,
                            const Quat qNa, const double Xb[3], const double ub[3],
                            const Quat qnb, const Quat qNb, const double eps0[6],
                            const PlasticParams &P, const double C[6], const double w[6],
                            const double Hk[6], const double ep0[6], const double al0[7],
                            bool seed, const double fba[6], const double fbb[6], bool has_r,
                            const double rba[6], const double rbb[6], double epb[6],
                            double alb[7], ElemVJP &out, MatBar &mb) {
  ElemFwd F;
  elem_vjp_forward(Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, F);
  double Rb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, pPb[3] = {0, 0, 0};
  double stb[6] = {0, 0, 0, 0, 0, 0};
  int n = 0;
  if (seed) {
#pragma unroll
    for (int i = 0; i < 6; i++) epb[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 7; i++) alb[i] = 0.0;
  } else {
    double ep[6], al[7], sm[6], smb[6], dfl;
#pragma unroll
    for (int i = 0; i < 6; i++) ep[i] = ep0[i];
#pragma unroll
    for (int i = 0; i < 7; i++) al[i] = al0[i];
    n = return_map<EXPK>(P, C, w, Hk, F.strain, ep, al, sm, &dfl, 1);
    if (n < 0) return n;
    elem_vjp_stress(F, sm, fba, fbb, Rb, pPb, smb);
    return_map_vjp<EXPK>(P, C, w, Hk, F.strain, ep0, al0, n, ep, smb, epb, alb, stb, mb);
  }
  if (has_r) {
    double so[6], sob[6];
#pragma unroll
    for (int i = 0; i < 6; i++) so[i] = C[i] * (F.strain[i] - ep0[i]);
    elem_vjp_stress(F, so, rba, rbb, Rb, pPb, sob);
#pragma unroll
    for (int i = 0; i < 6; i++) {
      mb.C[i] += sob[i] * (F.strain[i] - ep0[i]);
      stb[i] += C[i] * sob[i];
      epb[i] -= C[i] * sob[i];
    }
  }
  elem_vjp_strain(F, stb, Rb, pPb, qNa, qNb, out);
#pragma unroll
  for (int i = 0; i < 6; i++) out.C[i] = 0.0;
  return n;
}

// VJP of node_advance<true> (SO(3) mode).  Inputs are those of the forward
// call (f: element + contact internal force before the gyroscopic term);
// given the adjoints of the updated u, q, v returns the adjoints of u, q, v,
// f and of the mass data mt and J0 (J0 as the full 3x3 matrix it packs).
struct NodeVJP {
  double u[3];
  Quat q;
  double v[6];
  double f[6];
  double mt;
  double J0[9];
};

DL_HD void node_advance_vjp(const double f[6], const Quat q, const double v[6], double mt,
                            const double J0p[6], double gm, unsigned mask, const double ac[6],
                            const double fe[6], double dt, const double ubn[3], const Quat qbn,
                            const double vbn[6], NodeVJP &out) {
  // ---- forward
  double R[9], J[9];
  qmat(q, R);
  spatial_inertia(R, J0p, J);
  const double w[3] = {v[3], v[4], v[5]};
  double h[3];
#pragma unroll
  for (int i = 0; i < 3; i++) h[i] = J[3 * i] * w[0] + J[3 * i + 1] * w[1] + J[3 * i + 2] * w[2];
  double fr[6];
#pragma unroll
  for (int c = 0; c < 6; c++) fr[c] = fe[c] - f[c];
  fr[3] -= gm * (w[1] * h[2] - w[2] * h[1]);
  fr[4] -= gm * (w[2] * h[0] - w[0] * h[2]);
  fr[5] -= gm * (w[0] * h[1] - w[1] * h[0]);
  double a[6];
  const double imt = 1.0 / mt;
#pragma unroll
  for (int c = 0; c < 3; c++) a[c] = (mask & (1u << c)) ? ac[c] : imt * fr[c];
  solve3_masked(J, fr + 3, mask >> 3, ac + 3, a + 3);
  double vn[6];
#pragma unroll
  for (int c = 0; c < 6; c++) vn[c] = v[c] + dt * a[c];
  const double th[3] = {dt * vn[3], dt * vn[4], dt * vn[5]};
  const Quat dq = qexp(th[0], th[1], th[2]);
  const Quat P = qmul(dq, q);
  const double Pn = sqrt(P.w * P.w + P.x * P.x + P.y * P.y + P.z * P.z);
  Quat qn;
  qn.w = P.w / Pn; qn.x = P.x / Pn; qn.y = P.y / Pn; qn.z = P.z / Pn;
  // ---- reverse
  double vnb[6];
#pragma unroll
  for (int c = 0; c < 6; c++) vnb[c] = vbn[c];
  // q' = normalize(P), P = dq * q
  const double dqq = qn.w * qbn.w + qn.x * qbn.x + qn.y * qbn.y + qn.z * qbn.z;
  Quat Pb;
  Pb.w = (qbn.w - qn.w * dqq) / Pn; Pb.x = (qbn.x - qn.x * dqq) / Pn;
  Pb.y = (qbn.y - qn.y * dqq) / Pn; Pb.z = (qbn.z - qn.z * dqq) / Pn;
  Quat qb = qmul(qconj(dq), Pb);
  const Quat dqb = qmul(Pb, qconj(q));
  double thb[3];
  qexp_vjp(th, dqb, thb);
#pragma unroll
  for (int c = 0; c < 3; c++) vnb[3 + c] += dt * thb[c];
  // u' = u + dt v'_t
#pragma unroll
  for (int c = 0; c < 3; c++) {
    out.u[c] = ubn[c];
    vnb[c] += dt * ubn[c];
  }
  // v' = v + dt a
  double ab[6];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    out.v[c] = vnb[c];
    ab[c] = dt * vnb[c];
  }
  // translational acceleration
  double frb[6] = {0, 0, 0, 0, 0, 0};
  out.mt = 0.0;
#pragma unroll
  for (int c = 0; c < 3; c++)
    if (!(mask & (1u << c))) {
      frb[c] = imt * ab[c];
      out.mt -= imt * imt * fr[c] * ab[c];
    }
  // rotational: a_F = J_FF^-1 (fr_F - J_FC a_C)
  double lam[3];
  const double zero[3] = {0.0, 0.0, 0.0};
  double abm[3] = {ab[3], ab[4], ab[5]};
#pragma unroll
  for (int c = 0; c < 3; c++)
    if (mask & (8u << c)) abm[c] = 0.0;
  solve3_masked(J, abm, mask >> 3, zero, lam);
  double Jb[9];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    frb[3 + i] = lam[i];
#pragma unroll
    for (int j = 0; j < 3; j++) Jb[3 * i + j] = -lam[i] * a[3 + j];
  }
  // fr = fe - f - gyro
#pragma unroll
  for (int c = 0; c < 6; c++) out.f[c] = -frb[c];
  double gb[3] = {-frb[3], -frb[4], -frb[5]};
  // g = gm w x h, h = J w
  double wb[3], hb[3];
  wb[0] = gm * (h[1] * gb[2] - h[2] * gb[1]);
  wb[1] = gm * (h[2] * gb[0] - h[0] * gb[2]);
  wb[2] = gm * (h[0] * gb[1] - h[1] * gb[0]);
  hb[0] = gm * (gb[1] * w[2] - gb[2] * w[1]);
  hb[1] = gm * (gb[2] * w[0] - gb[0] * w[2]);
  hb[2] = gm * (gb[0] * w[1] - gb[1] * w[0]);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    wb[i] += J[i] * hb[0] + J[3 + i] * hb[1] + J[6 + i] * hb[2];
#pragma unroll
    for (int j = 0; j < 3; j++) Jb[3 * i + j] += hb[i] * w[j];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) out.v[3 + c] += wb[c];
  // J = R J0 R^T
  const double J0[9] = {J0p[0], J0p[3], J0p[4], J0p[3], J0p[1], J0p[5], J0p[4], J0p[5], J0p[2]};
  double S[9], T[9], Rb[9];
#pragma unroll
  for (int i = 0; i < 9; i++) S[i] = Jb[i] + Jb[3 * (i % 3) + i / 3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      T[3 * i + j] = R[3 * i] * J0[j] + R[3 * i + 1] * J0[3 + j] + R[3 * i + 2] * J0[6 + j];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      Rb[3 * i + j] = S[3 * i] * T[j] + S[3 * i + 1] * T[3 + j] + S[3 * i + 2] * T[6 + j];
  // J0bar = R^T Jb R
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double t = 0.0;
#pragma unroll
      for (int a1 = 0; a1 < 3; a1++)
#pragma unroll
        for (int b1 = 0; b1 < 3; b1++) t += R[3 * a1 + i] * Jb[3 * a1 + b1] * R[3 * b1 + j];
      out.J0[3 * i + j] = t;
    }
  out.q = qadd(qb, qmat_vjp(q, Rb));
}

// Contact pair VJP from one element's side.  X is one element of a candidate
// pair (X, Y) with A = lower strut id (XisA) as in the step.  The pair's
// penalty forces on all four nodes depend on the positions pX, pY and the
// contact distance R; with fbX, fbY the adjoints of the nodal forces of X's
// and Y's nodes, returns the gradient of fbX . F_X + fbY . F_Y with respect
// to X's six coordinates (out[0..5]) and R (out[6]).  Returns whether the
// pair is closed (out untouched otherwise).  Forward mode, one direction per
// pass, keeps the live set small enough for the device's registers.
DL_HD bool contact_pair_vjp(const double pX[2][3], const double pY[2][3], bool XisA, double R,
                            double kS, double kN, const double fbX[6], const double fbY[6],
                            double out[7]) {
  typedef Dual<1> D;
  const bool closed = XisA ? cm_narrowphase<false, 0>(pX, pY, R, kS, kN, (double *)nullptr,
                                                      (double *)nullptr)
                           : cm_narrowphase<false, 0>(pY, pX, R, kS, kN, (double *)nullptr,
                                                      (double *)nullptr);
  if (!closed) return false;
#pragma unroll 1
  for (int j = 0; j < 7; j++) {
    D dX[2][3], dY[2][3];
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        dX[k][i] = (j == 3 * k + i) ? D::var(pX[k][i], 0) : D(pX[k][i]);
        dY[k][i] = D(pY[k][i]);
      }
    const D dR = (j == 6) ? D::var(R, 0) : D(R);
    D fX0[3] = {}, fX1[3] = {}, fY0[3] = {}, fY1[3] = {};
    if (XisA) {
      cm_narrowphase<true, 0>(dX, dY, dR, kS, kN, fX0, fX1);
      cm_narrowphase<true, 1>(dX, dY, dR, kS, kN, fY0, fY1);
    } else {
      cm_narrowphase<true, 1>(dY, dX, dR, kS, kN, fX0, fX1);
      cm_narrowphase<true, 0>(dY, dX, dR, kS, kN, fY0, fY1);
    }
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      t += fbX[i] * fX0[i].d[0] + fbX[3 + i] * fX1[i].d[0];
      t += fbY[i] * fY0[i].d[0] + fbY[3 + i] * fY1[i].d[0];
    }
    // static indices only (a dynamic one would put out in scratch memory)
#pragma unroll
    for (int q = 0; q < 7; q++) out[q] = (q == j) ? t : out[q];
  }
  return true;
}

#endif
