// dl_contact_math.h -- closest-point narrowphase of the rod-rod penalty
// contact, shared by the device kernels (dl_contact.hip).
//
// Restates RodContactModel::computeContacts_ for one candidate pair of
// 2-node elements (RodContactModel.cpp:451-738) with findClosestPoints_
// (:953-991), getClosestPoint_ (:996-1026), computeSTS_ (:1031-1152, force
// part) and computeNTS_ (:1157-1208, force part), in the same operation
// order (no FMA contraction) so that the closed/open decision of every pair
// is the same bit pattern as the oracle's.
//
// Every function is a template on the scalar type T: the step instantiates
// it with double; the adjoint instantiates it with forward-mode dual numbers
// (dl_dual.h), whose value parts follow the same operations bit for bit, so
// both see the same branch of every pair.
#ifndef DL_CONTACT_MATH_H
#define DL_CONTACT_MATH_H

#include "dl_device.h"

DL_HD double dl_val(double x) { return x; }

// StdLine::containsLocalPoint(u): -1 <= u <= 1 (NaN is outside)
template <typename T>
DL_HD bool cm_contains(const T &u) { return u >= -1.0 && u <= 1.0; }

template <typename T>
DL_HD T cm_dot(const T a[3], const T b[3]) {
#pragma clang fp contract(off)
  T s = a[0] * b[0];
  s += a[1] * b[1];
  s += a[2] * b[2];
  return s;
}

template <typename T>
DL_HD void cm_point(const T &u, const T p0[3], const T p1[3], T x[3]) {
#pragma clang fp contract(off)
  const T N0 = 0.5 * (1. - u), N1 = 0.5 * (1. + u);
  for (int i = 0; i < 3; i++) {
    T s = p0[i] * N0;
    s += p1[i] * N1;
    x[i] = s;
  }
}

template <typename T>
DL_HD void cm_closest_points(const T pA[2][3], const T pB[2][3], T *uA, T *uB) {
#pragma clang fp contract(off)
  T bA[3], bB[3], tA[3], tB[3], d[3], vA[3], vB[3];
  for (int i = 0; i < 3; i++) {
    bA[i] = pA[1][i] + pA[0][i];
    bB[i] = pB[1][i] + pB[0][i];
    tA[i] = pA[1][i] - pA[0][i];
    tB[i] = pB[1][i] - pB[0][i];
    d[i] = bB[i] - bA[i];
  }
  const T tBtA = cm_dot(tB, tA), tBtB = cm_dot(tB, tB), tAtA = cm_dot(tA, tA);
  for (int i = 0; i < 3; i++) {
    vA[i] = tB[i] * tBtA - tA[i] * tBtB;
    vB[i] = tA[i] * tBtA - tB[i] * tAtA;
  }
  const T den = tBtB * tAtA - tBtA * tBtA;
  *uA = -1. * cm_dot(d, vA) / den;
  *uB = +1. * cm_dot(d, vB) / den;
}

template <typename T>
DL_HD T cm_closest_point(const T s[3], const T pM[2][3]) {
#pragma clang fp contract(off)
  T bM[3], tM[3];
  for (int i = 0; i < 3; i++) {
    bM[i] = pM[1][i] + pM[0][i];
    tM[i] = pM[1][i] - pM[0][i];
  }
  return (2 * cm_dot(s, tM) - cm_dot(tM, bM)) / cm_dot(tM, tM);
}

// jem::isTiny(d): |d| < DBL_MIN
template <typename T>
DL_HD bool cm_tiny(const T &d) { return fabs(dl_val(d)) < 2.2250738585072014e-308; }

// segment-to-segment; f[12] += (A0, A1, B0, B1) when f != nullptr
template <typename T>
DL_HD bool cm_sts(const T pA[2][3], const T pB[2][3], const T &uA, const T &uB, const T &R,
                  double kap, T *f) {
#pragma clang fp contract(off)
  T xA[3], xB[3], d[3];
  cm_point(uA, pA[0], pA[1], xA);
  cm_point(uB, pB[0], pB[1], xB);
  for (int i = 0; i < 3; i++) d[i] = xB[i] - xA[i];
  const T dist = sqrt(cm_dot(d, d));
  if (cm_tiny(dist)) return false;
  if (dist > R) return false;
  if (!f) return true;
  T n[3];
  for (int i = 0; i < 3; i++) n[i] = d[i] / dist;
  const T NA0 = 0.5 * (1. - uA), NA1 = 0.5 * (1. + uA);
  const T NB0 = 0.5 * (1. - uB), NB1 = 0.5 * (1. + uB);
  const T sc = kap * (dist - R);
  for (int i = 0; i < 3; i++) {
    f[0 + i] += sc * ((-1. * NA0) * n[i]);
    f[3 + i] += sc * ((-1. * NA1) * n[i]);
    f[6 + i] += sc * (NB0 * n[i]);
    f[9 + i] += sc * (NB1 * n[i]);
  }
  return true;
}

// node-to-segment; f[9] += (slave, M0, M1) when f != nullptr
template <typename T>
DL_HD bool cm_nts(const T s[3], const T pM[2][3], const T &uM, const T &R, double kap, T *f) {
#pragma clang fp contract(off)
  T xM[3], d[3];
  cm_point(uM, pM[0], pM[1], xM);
  for (int i = 0; i < 3; i++) d[i] = xM[i] - s[i];
  const T dist = sqrt(cm_dot(d, d));
  if (cm_tiny(dist)) return false;
  if (dist > R) return false;
  if (!f) return true;
  T n[3];
  for (int i = 0; i < 3; i++) n[i] = (s[i] - xM[i]) / dist;
  const T sc = kap * (dist - R);
  const T a = -.5 * (1. - uM), b = -.5 * (1. + uM);
  for (int i = 0; i < 3; i++) {
    f[i] += sc * n[i];
    f[3 + i] += sc * (a * n[i]);
    f[6 + i] += sc * (b * n[i]);
  }
  return true;
}

// One speculative node-to-segment contact of cm_narrowphase: end node NODE
// of A onto B (FROM_A) or of B onto A, its contributions added to the rows
// that SIDE selects (see cm_narrowphase).  Returns whether it is closed.
template <bool F, int SIDE, int NODE, bool FROM_A, typename T>
DL_HD bool cm_spec_nts(const T pA[2][3], const T pB[2][3], const T &R, double kN, T f0[3],
                       T f1[3], T *h0, T *h1) {
  constexpr bool WA = SIDE != 1, WB = SIDE != 0;
  const T *node = FROM_A ? pA[NODE] : pB[NODE];
  const T u = FROM_A ? cm_closest_point(node, pB) : cm_closest_point(node, pA);
  if (!cm_contains(u)) return false;
  T g[9] = {};
  if (!(FROM_A ? cm_nts(node, pB, u, R, kN, F ? g : nullptr)
               : cm_nts(node, pA, u, R, kN, F ? g : nullptr)))
    return false;
  if (F)
    for (int i = 0; i < 3; i++) {
      // the node's force goes to its own element's row NODE, the segment's
      // to the other element's two rows
      if (WA) {
        if (FROM_A) (NODE == 0 ? f0 : f1)[i] += g[i];
        else {
          f0[i] += g[3 + i];
          f1[i] += g[6 + i];
        }
      }
      if (WB) {
        if (FROM_A) {
          (SIDE == 2 ? h0 : f0)[i] += g[3 + i];
          (SIDE == 2 ? h1 : f1)[i] += g[6 + i];
        } else if (NODE == 0) (SIDE == 2 ? h0 : f0)[i] += g[i];
        else (SIDE == 2 ? h1 : f1)[i] += g[i];
      }
    }
  return true;
}

// computeContacts_ per candidate pair (A, B) after the shared-node and
// blacklist filters.  Returns whether the contact is closed.  When F is true
// the translational contributions are added: SIDE 0 -> (A0, A1) into (f0, f1),
// SIDE 1 -> (B0, B1) into (f0, f1), SIDE 2 -> (A0, A1) into (f0, f1) and
// (B0, B1) into (h0, h1).  Each side's sums are the same operations in the
// same order for every SIDE, so one SIDE 2 evaluation gives both sides the
// bits of a SIDE 0 and a SIDE 1 evaluation.  All indices are static so
// nothing spills.
template <bool F, int SIDE, typename T>
DL_HD bool cm_narrowphase(const T pA[2][3], const T pB[2][3], const T &R, double kS, double kN,
                          T f0[3], T f1[3], T *h0 = nullptr, T *h1 = nullptr) {
  constexpr bool WA = SIDE != 1, WB = SIDE != 0; // the sides written
  // B's rows: (h0, h1) for SIDE 2, else (f0, f1)
#define CM_B0(i) (SIDE == 2 ? h0[i] : f0[i])
#define CM_B1(i) (SIDE == 2 ? h1[i] : f1[i])
  T uA, uB;
  cm_closest_points(pA, pB, &uA, &uB);
  const bool inA = cm_contains(uA), inB = cm_contains(uB);
  if (inA && inB) {
    T g[12] = {};
    if (!cm_sts(pA, pB, uA, uB, R, kS, F ? g : nullptr)) return false;
    if (F)
      for (int i = 0; i < 3; i++) {
        if (WA) {
          f0[i] += g[i];
          f1[i] += g[3 + i];
        }
        if (WB) {
          CM_B0(i) += g[6 + i];
          CM_B1(i) += g[9 + i];
        }
      }
    return true;
  }
  if (!inA && inB) {
    const bool i1 = !(uA < -1.);
    // RodContactModel.cpp:575 projects node iA of B (not of A) onto A
    const T sB[3] = {i1 ? pB[1][0] : pB[0][0], i1 ? pB[1][1] : pB[0][1],
                          i1 ? pB[1][2] : pB[0][2]};
    const T sA[3] = {i1 ? pA[1][0] : pA[0][0], i1 ? pA[1][1] : pA[0][1],
                          i1 ? pA[1][2] : pA[0][2]};
    const T u = cm_closest_point(sB, pA);
    if (!cm_contains(u)) return false;
    T g[9] = {};
    if (!cm_nts(sA, pB, u, R, kN, F ? g : nullptr)) return false;
    if (F)
      for (int i = 0; i < 3; i++) {
        if (WA) {
          f0[i] += i1 ? T(0.0) : g[i];
          f1[i] += i1 ? g[i] : T(0.0);
        }
        if (WB) {
          CM_B0(i) += g[3 + i];
          CM_B1(i) += g[6 + i];
        }
      }
    return true;
  }
  if (!inB && inA) {
    const bool i1 = !(uB < -1.);
    const T sB[3] = {i1 ? pB[1][0] : pB[0][0], i1 ? pB[1][1] : pB[0][1],
                          i1 ? pB[1][2] : pB[0][2]};
    const T u = cm_closest_point(sB, pA);
    if (!cm_contains(u)) return false;
    T g[9] = {};
    if (!cm_nts(sB, pA, u, R, kN, F ? g : nullptr)) return false;
    if (F)
      for (int i = 0; i < 3; i++) {
        if (WB) {
          CM_B0(i) += i1 ? T(0.0) : g[i];
          CM_B1(i) += i1 ? g[i] : T(0.0);
        }
        if (WA) {
          f0[i] += g[3 + i];
          f1[i] += g[6 + i];
        }
      }
    return true;
  }
  // speculative NTS from all four end nodes, A's then B's (the order of
  // the sums)
  bool closed = cm_spec_nts<F, SIDE, 0, true>(pA, pB, R, kN, f0, f1, h0, h1);
  closed = cm_spec_nts<F, SIDE, 1, true>(pA, pB, R, kN, f0, f1, h0, h1) || closed;
  closed = cm_spec_nts<F, SIDE, 0, false>(pA, pB, R, kN, f0, f1, h0, h1) || closed;
  closed = cm_spec_nts<F, SIDE, 1, false>(pA, pB, R, kN, f0, f1, h0, h1) || closed;
  return closed;
#undef CM_B0
#undef CM_B1
}

#endif
