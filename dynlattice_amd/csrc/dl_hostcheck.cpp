// dl_hostcheck.cpp -- host instantiations of the device math used by the
// step and the adjoint (dl_device.h, dl_contact_math.h, dl_adjoint.h), for
// unit tests on machines without a GPU: the vector-Jacobian products are
// checked against central finite differences of the forward functions.
// Built into its own library (libdlhost.so); the product library does not
// contain or call any of this.
#include "dl_adjoint.h"

static Quat q4(const double *p) {
  Quat q;
  q.w = p[0]; q.x = p[1]; q.y = p[2]; q.z = p[3];
  return q;
}
static void put4(const Quat q, double *p) {
  p[0] = q.w; p[1] = q.x; p[2] = q.y; p[3] = q.z;
}

// element internal force (elastic): f[12] = (fa 6, fb 6)
extern "C" void hc_elem_force(const double *Xa, const double *ua, const double *qna,
                              const double *qNa, const double *Xb, const double *ub,
                              const double *qnb, const double *qNb, const double *eps0,
                              const double *C, double *f) {
  ElemKin k;
  elem_kinematics(Xa, ua, Xb, ub, qmul(q4(qna), q4(qNa)), qmul(q4(qnb), q4(qNb)), k);
  double sm[6];
  for (int i = 0; i < 6; i++) sm[i] = C[i] * (k.eps[i] - eps0[i]);
  elem_forces(k, sm, f, f + 6);
}

// out[20] = ubar_a 3, qbar_a 4, ubar_b 3, qbar_b 4, Cbar 6
extern "C" void hc_elem_force_vjp(const double *Xa, const double *ua, const double *qna,
                                  const double *qNa, const double *Xb, const double *ub,
                                  const double *qnb, const double *qNb, const double *eps0,
                                  const double *C, const double *fbar, double *out) {
  ElemVJP r;
  elem_force_vjp(Xa, ua, q4(qna), q4(qNa), Xb, ub, q4(qnb), q4(qNb), eps0, C, fbar, fbar + 6, r);
  for (int i = 0; i < 3; i++) { out[i] = r.ua[i]; out[7 + i] = r.ub[i]; }
  put4(r.qa, out + 3);
  put4(r.qb, out + 10);
  for (int i = 0; i < 6; i++) out[14 + i] = r.C[i];
}

// SO(3) node update: out[13] = u' 3, q' 4, v' 6
extern "C" void hc_node_advance(const double *f, const double *q, const double *u,
                                const double *v, double mt, const double *J0, double gm,
                                unsigned mask, const double *ac, const double *fe, double dt,
                                double *out) {
  double ff[6], uu[3], vv[6], rr[4];
  for (int i = 0; i < 6; i++) { ff[i] = f[i]; vv[i] = v[i]; }
  for (int i = 0; i < 3; i++) uu[i] = u[i];
  for (int i = 0; i < 4; i++) rr[i] = q[i];
  node_advance<true>(ff, q4(q), uu, vv, rr, mt, J0, gm, mask, ac, fe, dt);
  for (int i = 0; i < 3; i++) out[i] = uu[i];
  for (int i = 0; i < 4; i++) out[3 + i] = rr[i];
  for (int i = 0; i < 6; i++) out[7 + i] = vv[i];
}

// out[29] = ubar 3, qbar 4, vbar 6, fbar 6, mtbar 1, J0bar 9
extern "C" void hc_node_advance_vjp(const double *f, const double *q, const double *v, double mt,
                                    const double *J0, double gm, unsigned mask, const double *ac,
                                    const double *fe, double dt, const double *ubn,
                                    const double *qbn, const double *vbn, double *out) {
  NodeVJP r;
  node_advance_vjp(f, q4(q), v, mt, J0, gm, mask, ac, fe, dt, ubn, q4(qbn), vbn, r);
  for (int i = 0; i < 3; i++) out[i] = r.u[i];
  put4(r.q, out + 3);
  for (int i = 0; i < 6; i++) { out[7 + i] = r.v[i]; out[13 + i] = r.f[i]; }
  out[19] = r.mt;
  for (int i = 0; i < 9; i++) out[20 + i] = r.J0[i];
}

// all four nodal forces of a candidate pair: f[12] = A0, A1, B0, B1; returns closed
extern "C" int hc_pair_forces(const double *pA, const double *pB, double R, double kS, double kN,
                              double *f) {
  double a[2][3], b[2][3];
  for (int k = 0; k < 2; k++)
    for (int i = 0; i < 3; i++) { a[k][i] = pA[3 * k + i]; b[k][i] = pB[3 * k + i]; }
  for (int i = 0; i < 12; i++) f[i] = 0.0;
  const bool c = cm_narrowphase<true, 0>(a, b, R, kS, kN, f, f + 3);
  cm_narrowphase<true, 1>(a, b, R, kS, kN, f + 6, f + 9);
  return c ? 1 : 0;
}

extern "C" int hc_pair_vjp(const double *pX, const double *pY, int XisA, double R, double kS,
                           double kN, const double *fbX, const double *fbY, double *out) {
  double x[2][3], y[2][3];
  for (int k = 0; k < 2; k++)
    for (int i = 0; i < 3; i++) { x[k][i] = pX[3 * k + i]; y[k][i] = pY[3 * k + i]; }
  return contact_pair_vjp(x, y, XisA != 0, R, kS, kN, fbX, fbY, out) ? 1 : 0;
}

// ---------------------------------------------------------------- plastic
// material parameters as a flat array: yield_level, kin_coupling 6,
// iso_coupling, iso_modulus, precision, exponent 6, kin_const 6, kin_ratio,
// has_iso, has_kin, max_iter (26)
static PlasticParams pp_of(const double *a) {
  PlasticParams P;
  P.yield_level = a[0];
  for (int i = 0; i < 6; i++) P.kin_coupling[i] = a[1 + i];
  P.iso_coupling = a[7];
  P.iso_modulus = a[8];
  P.precision = a[9];
  for (int i = 0; i < 6; i++) {
    P.exponent[i] = a[10 + i];
    P.kin_const[i] = a[16 + i];
  }
  P.kin_ratio = a[22];
  P.has_iso = (int32_t)a[23];
  P.has_kin = (int32_t)a[24];
  P.max_iter = (int32_t)a[25];
  return P;
}

template <int K>
static int rm_fwd(const PlasticParams &P, const double *C, const double *w, const double *Hk,
                  const double *strain, double *ep, double *al, double *st) {
  double dfl;
  return return_map<K>(P, C, w, Hk, strain, ep, al, st, &dfl, 1);
}

// return map from (ep, al) (in/out); stress out; returns the corrections
extern "C" int hc_return_map(int expk, const double *pa, const double *C, const double *w,
                             const double *Hk, const double *strain, double *ep, double *al,
                             double *st) {
  const PlasticParams P = pp_of(pa);
  if (expk == 1) return rm_fwd<1>(P, C, w, Hk, strain, ep, al, st);
  if (expk == 2) return rm_fwd<2>(P, C, w, Hk, strain, ep, al, st);
  return rm_fwd<0>(P, C, w, Hk, strain, ep, al, st);
}

template <int K>
static int rm_vjp(const PlasticParams &P, const double *C, const double *w, const double *Hk,
                  const double *strain, const double *ep0, const double *al0, const double *sb,
                  double *epb, double *alb, double *strb, MatBar &mb) {
  double ep[6], al[7], st[6], dfl;
  for (int i = 0; i < 6; i++) ep[i] = ep0[i];
  for (int i = 0; i < 7; i++) al[i] = al0[i];
  const int n = return_map<K>(P, C, w, Hk, strain, ep, al, st, &dfl, 1);
  if (n < 0) return n;
  return_map_vjp<K>(P, C, w, Hk, strain, ep0, al0, n, ep, sb, epb, alb, strb, mb);
  return n;
}

// out[37] = strain_bar 6, ep0_bar 6, al0_bar 7, C_bar 6, w_bar 6, Hk_bar 6,
// given the adjoints of the returned stress (sb), plastic strain (epb) and
// hardening parameters (alb)
extern "C" int hc_return_map_vjp(int expk, const double *pa, const double *C, const double *w,
                                 const double *Hk, const double *strain, const double *ep0,
                                 const double *al0, const double *sb, const double *epb_in,
                                 const double *alb_in, double *out) {
  const PlasticParams P = pp_of(pa);
  double epb[6], alb[7], strb[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 6; i++) epb[i] = epb_in[i];
  for (int i = 0; i < 7; i++) alb[i] = alb_in[i];
  MatBar mb = {};
  int n;
  if (expk == 1) n = rm_vjp<1>(P, C, w, Hk, strain, ep0, al0, sb, epb, alb, strb, mb);
  else if (expk == 2) n = rm_vjp<2>(P, C, w, Hk, strain, ep0, al0, sb, epb, alb, strb, mb);
  else n = rm_vjp<0>(P, C, w, Hk, strain, ep0, al0, sb, epb, alb, strb, mb);
  for (int i = 0; i < 6; i++) {
    out[i] = strb[i];
    out[6 + i] = epb[i];
    out[19 + i] = mb.C[i];
    out[25 + i] = mb.w[i];
    out[31 + i] = mb.Hk[i];
  }
  for (int i = 0; i < 7; i++) out[12 + i] = alb[i];
  return n;
}

// elastoplastic element: out[37] = the step's force (fa, fb: 12, return map
// from (ep0, al0)), the output load case's force (12, elastic predictor),
// the returned plastic strain 6 and hardening parameters 7
extern "C" int hc_elem_force_pl(int expk, const double *pa, const double *Xa, const double *ua,
                                const double *qna, const double *qNa, const double *Xb,
                                const double *ub, const double *qnb, const double *qNb,
                                const double *eps0, const double *C, const double *w,
                                const double *Hk, const double *ep0, const double *al0,
                                double *out) {
  const PlasticParams P = pp_of(pa);
  ElemKin k;
  elem_kinematics(Xa, ua, Xb, ub, qmul(q4(qna), q4(qNa)), qmul(q4(qnb), q4(qNb)), k);
  double strain[6], ep[6], al[7], sm[6], so[6];
  for (int i = 0; i < 6; i++) {
    strain[i] = k.eps[i] - eps0[i];
    ep[i] = ep0[i];
    so[i] = C[i] * (strain[i] - ep0[i]);
  }
  for (int i = 0; i < 7; i++) al[i] = al0[i];
  int n;
  if (expk == 1) n = rm_fwd<1>(P, C, w, Hk, strain, ep, al, sm);
  else if (expk == 2) n = rm_fwd<2>(P, C, w, Hk, strain, ep, al, sm);
  else n = rm_fwd<0>(P, C, w, Hk, strain, ep, al, sm);
  elem_forces(k, sm, out, out + 6);
  elem_forces(k, so, out + 12, out + 18);
  for (int i = 0; i < 6; i++) out[24 + i] = ep[i];
  for (int i = 0; i < 7; i++) out[30 + i] = al[i];
  return n;
}

template <int K>
static int efp_vjp(const PlasticParams &P, const double *Xa, const double *ua, const Quat qna,
                   const Quat qNa, const double *Xb, const double *ub, const Quat qnb,
                   const Quat qNb, const double *eps0, const double *C, const double *w,
                   const double *Hk, const double *ep0, const double *al0, bool seed,
                   const double *fb, bool has_r, const double *rb, double *epb, double *alb,
                   ElemVJP &r, MatBar &mb) {
  return elem_force_vjp_pl<K>(Xa, ua, qna, qNa, Xb, ub, qnb, qNb, eps0, P, C, w, Hk, ep0, al0,
                              seed, fb, fb + 6, has_r, rb, rb + 6, epb, alb, r, mb);
}

// out[45] = ubar_a 3, qbar_a 4, ubar_b 3, qbar_b 4, ep0_bar 6, al0_bar 7,
// C_bar 6, w_bar 6, Hk_bar 6; fbar[12] the step force's adjoint (unused
// with seed), rbar[12] the output load case's (or NULL), epb[6], alb[7]
// the adjoints of the returned plastic state (unused with seed)
extern "C" int hc_elem_force_vjp_pl(int expk, const double *pa, const double *Xa,
                                    const double *ua, const double *qna, const double *qNa,
                                    const double *Xb, const double *ub, const double *qnb,
                                    const double *qNb, const double *eps0, const double *C,
                                    const double *w, const double *Hk, const double *ep0,
                                    const double *al0, int seed, const double *fbar,
                                    const double *rbar, const double *epb_in,
                                    const double *alb_in, double *out) {
  const PlasticParams P = pp_of(pa);
  double epb[6], alb[7];
  for (int i = 0; i < 6; i++) epb[i] = epb_in[i];
  for (int i = 0; i < 7; i++) alb[i] = alb_in[i];
  const double zero[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double *rb = rbar ? rbar : zero;
  ElemVJP r;
  MatBar mb = {};
  int n;
  if (expk == 1)
    n = efp_vjp<1>(P, Xa, ua, q4(qna), q4(qNa), Xb, ub, q4(qnb), q4(qNb), eps0, C, w, Hk, ep0, al0,
                   seed != 0, fbar, rbar != nullptr, rb, epb, alb, r, mb);
  else if (expk == 2)
    n = efp_vjp<2>(P, Xa, ua, q4(qna), q4(qNa), Xb, ub, q4(qnb), q4(qNb), eps0, C, w, Hk, ep0, al0,
                   seed != 0, fbar, rbar != nullptr, rb, epb, alb, r, mb);
  else
    n = efp_vjp<0>(P, Xa, ua, q4(qna), q4(qNa), Xb, ub, q4(qnb), q4(qNb), eps0, C, w, Hk, ep0, al0,
                   seed != 0, fbar, rbar != nullptr, rb, epb, alb, r, mb);
  for (int i = 0; i < 3; i++) { out[i] = r.ua[i]; out[7 + i] = r.ub[i]; }
  put4(r.qa, out + 3);
  put4(r.qb, out + 10);
  for (int i = 0; i < 6; i++) {
    out[14 + i] = epb[i];
    out[27 + i] = mb.C[i];
    out[33 + i] = mb.w[i];
    out[39 + i] = mb.Hk[i];
  }
  for (int i = 0; i < 7; i++) out[20 + i] = alb[i];
  return n;
}

// material strain minus the initial strain of an element (out[6])
extern "C" void hc_elem_strain(const double *Xa, const double *ua, const double *qna,
                               const double *qNa, const double *Xb, const double *ub,
                               const double *qnb, const double *qNb, const double *eps0,
                               double *out) {
  ElemKin k;
  elem_kinematics(Xa, ua, Xb, ub, qmul(q4(qna), q4(qNa)), qmul(q4(qnb), q4(qNb)), k);
  for (int i = 0; i < 6; i++) out[i] = k.eps[i] - eps0[i];
}
