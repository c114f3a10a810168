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
