// dual.h -- forward-mode dual numbers for the oracle's tangent mode.
//
// TEST INFRASTRUCTURE ONLY (part of the parity oracle).  oracle.cpp is a
// template on its scalar type; instantiated with CDual<N> it carries, next to
// every value the plain oracle computes (bit for bit the same operations),
// its derivatives along N seeded parameters.  That gives an FP64 CPU
// reference gradient of any output (here the wall-reaction history) with
// respect to the seeded parameters, with every branch following the values.
#ifndef ORACLE_DUAL_H
#define ORACLE_DUAL_H

#include <cmath>

template <int N> struct CDual {
  double v;
  double d[N];
  CDual() : v(0.0) {
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
  CDual(double x) : v(x) {
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
};

inline double sval(double x) { return x; }
template <int N> inline double sval(const CDual<N> &x) { return x.v; }

template <int N> inline CDual<N> operator+(const CDual<N> &a, const CDual<N> &b) {
  CDual<N> r;
  r.v = a.v + b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N> inline CDual<N> operator-(const CDual<N> &a, const CDual<N> &b) {
  CDual<N> r;
  r.v = a.v - b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N> inline CDual<N> operator-(const CDual<N> &a) {
  CDual<N> r;
  r.v = -a.v;
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N> inline CDual<N> operator*(const CDual<N> &a, const CDual<N> &b) {
  CDual<N> r;
  r.v = a.v * b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int N> inline CDual<N> operator/(const CDual<N> &a, const CDual<N> &b) {
  CDual<N> r;
  r.v = a.v / b.v;
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
template <int N> inline CDual<N> operator+(const CDual<N> &a, double b) { return a + CDual<N>(b); }
template <int N> inline CDual<N> operator+(double a, const CDual<N> &b) { return CDual<N>(a) + b; }
template <int N> inline CDual<N> operator-(const CDual<N> &a, double b) { return a - CDual<N>(b); }
template <int N> inline CDual<N> operator-(double a, const CDual<N> &b) { return CDual<N>(a) - b; }
template <int N> inline CDual<N> operator*(const CDual<N> &a, double b) {
  CDual<N> r;
  r.v = a.v * b;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int N> inline CDual<N> operator*(double a, const CDual<N> &b) {
  CDual<N> r;
  r.v = a * b.v;
  for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int N> inline CDual<N> operator/(const CDual<N> &a, double b) {
  CDual<N> r;
  r.v = a.v / b;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
  return r;
}
template <int N> inline CDual<N> operator/(double a, const CDual<N> &b) { return CDual<N>(a) / b; }
template <int N> inline CDual<N> &operator+=(CDual<N> &a, const CDual<N> &b) { return a = a + b; }
template <int N> inline CDual<N> &operator-=(CDual<N> &a, const CDual<N> &b) { return a = a - b; }
template <int N> inline CDual<N> &operator*=(CDual<N> &a, const CDual<N> &b) { return a = a * b; }
template <int N> inline CDual<N> &operator+=(CDual<N> &a, double b) { return a = a + b; }
template <int N> inline CDual<N> &operator-=(CDual<N> &a, double b) { return a = a - b; }
template <int N> inline CDual<N> &operator*=(CDual<N> &a, double b) { return a = a * b; }

template <int N> inline bool operator<(const CDual<N> &a, const CDual<N> &b) { return a.v < b.v; }
template <int N> inline bool operator<(const CDual<N> &a, double b) { return a.v < b; }
template <int N> inline bool operator<(double a, const CDual<N> &b) { return a < b.v; }
template <int N> inline bool operator>(const CDual<N> &a, const CDual<N> &b) { return a.v > b.v; }
template <int N> inline bool operator>(const CDual<N> &a, double b) { return a.v > b; }
template <int N> inline bool operator<=(const CDual<N> &a, const CDual<N> &b) { return a.v <= b.v; }
template <int N> inline bool operator<=(const CDual<N> &a, double b) { return a.v <= b; }

template <int N> inline CDual<N> sqrt(const CDual<N> &a) {
  CDual<N> r;
  r.v = std::sqrt(a.v);
  const double s = r.v > 0 ? 0.5 / r.v : 0.0;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
  return r;
}
template <int N> inline CDual<N> fabs(const CDual<N> &a) {
  CDual<N> r;
  r.v = std::fabs(a.v);
  const double s = a.v > 0 ? 1.0 : (a.v < 0 ? -1.0 : 0.0);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
  return r;
}
template <int N> inline CDual<N> asin(const CDual<N> &a) {
  CDual<N> r;
  r.v = std::asin(a.v);
  const double q = 1.0 - a.v * a.v;
  const double s = q > 0 ? 1.0 / std::sqrt(q) : 0.0;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
  return r;
}

#endif
