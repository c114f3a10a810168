// dl_dual.h -- forward-mode dual numbers with ND tangent directions.
//
// Used to differentiate the closest-point contact pair force (the templated
// functions of dl_contact_math.h) with respect to one element's six nodal
// coordinates and the contact distance.  The value part of every operation is
// the plain double operation, evaluated without FMA contraction, so a dual
// evaluation takes exactly the branches that the double evaluation takes.
#ifndef DL_DUAL_H
#define DL_DUAL_H

#include "dl_device.h"

template <int ND>
struct Dual {
  double v;
  double d[ND];
  Dual() = default;
  DL_HD Dual(double x) : v(x) {
    for (int i = 0; i < ND; i++) d[i] = 0.0;
  }
  // independent variable number k
  static DL_HD Dual var(double x, int k) {
    Dual r(x);
    r.d[k] = 1.0;
    return r;
  }
};

template <int ND>
DL_HD double dl_val(const Dual<ND> &a) { return a.v; }

template <int ND>
DL_HD Dual<ND> operator+(const Dual<ND> &a, const Dual<ND> &b) {
#pragma clang fp contract(off)
  Dual<ND> r;
  r.v = a.v + b.v;
  for (int i = 0; i < ND; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int ND>
DL_HD Dual<ND> operator-(const Dual<ND> &a, const Dual<ND> &b) {
#pragma clang fp contract(off)
  Dual<ND> r;
  r.v = a.v - b.v;
  for (int i = 0; i < ND; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int ND>
DL_HD Dual<ND> operator*(const Dual<ND> &a, const Dual<ND> &b) {
#pragma clang fp contract(off)
  Dual<ND> r;
  r.v = a.v * b.v;
  for (int i = 0; i < ND; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int ND>
DL_HD Dual<ND> operator/(const Dual<ND> &a, const Dual<ND> &b) {
#pragma clang fp contract(off)
  Dual<ND> r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
  for (int i = 0; i < ND; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int ND>
DL_HD Dual<ND> operator+(double a, const Dual<ND> &b) { return Dual<ND>(a) + b; }
template <int ND>
DL_HD Dual<ND> operator+(const Dual<ND> &a, double b) { return a + Dual<ND>(b); }
template <int ND>
DL_HD Dual<ND> operator-(double a, const Dual<ND> &b) { return Dual<ND>(a) - b; }
template <int ND>
DL_HD Dual<ND> operator-(const Dual<ND> &a, double b) { return a - Dual<ND>(b); }
template <int ND>
DL_HD Dual<ND> operator*(double a, const Dual<ND> &b) {
#pragma clang fp contract(off)
  Dual<ND> r;
  r.v = a * b.v;
  for (int i = 0; i < ND; i++) r.d[i] = a * b.d[i];
  return r;
}
template <int ND>
DL_HD Dual<ND> operator*(const Dual<ND> &a, double b) {
#pragma clang fp contract(off)
  Dual<ND> r;
  r.v = a.v * b;
  for (int i = 0; i < ND; i++) r.d[i] = a.d[i] * b;
  return r;
}
template <int ND>
DL_HD Dual<ND> operator/(const Dual<ND> &a, double b) { return a / Dual<ND>(b); }
template <int ND>
DL_HD Dual<ND> &operator+=(Dual<ND> &a, const Dual<ND> &b) {
  a = a + b;
  return a;
}
template <int ND>
DL_HD Dual<ND> sqrt(const Dual<ND> &a) {
#pragma clang fp contract(off)
  Dual<ND> r;
  r.v = sqrt(a.v);
  const double h = 0.5 / r.v;
  for (int i = 0; i < ND; i++) r.d[i] = a.d[i] * h;
  return r;
}
template <int ND>
DL_HD bool operator<(const Dual<ND> &a, double b) { return a.v < b; }
template <int ND>
DL_HD bool operator>(const Dual<ND> &a, double b) { return a.v > b; }
template <int ND>
DL_HD bool operator<=(const Dual<ND> &a, double b) { return a.v <= b; }
template <int ND>
DL_HD bool operator>=(const Dual<ND> &a, double b) { return a.v >= b; }
template <int ND>
DL_HD bool operator>(const Dual<ND> &a, const Dual<ND> &b) { return a.v > b.v; }
template <int ND>
DL_HD bool operator<(const Dual<ND> &a, const Dual<ND> &b) { return a.v < b.v; }

#endif
