// pg_dual.h -- forward-mode dual numbers for the adjoint kernels.
//
// Dual<N> carries a value and N directional derivatives.  The value part of
// every operation is the same IEEE operation the time step performs on
// doubles (the files are compiled with -ffp-contract=off), so a force
// function evaluated on Dual<N> takes exactly the branches the time step
// took, and the derivative parts give the force's Jacobian along the N
// seeded directions.
#ifndef PG_DUAL_H
#define PG_DUAL_H
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

namespace pg {

// the double overloads stay visible next to the Dual ones below
using ::asin;
using ::fabs;
using ::sqrt;

template <int N> struct Dual {
  double v;
  double d[N];
  __host__ __device__ Dual() {}
  __host__ __device__ Dual(double x) : v(x) {
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
  // x with derivative 1 along direction `dir` (dir < 0: a constant)
  __host__ __device__ static Dual seed(double x, int dir) {
    Dual r(x);
    if (dir >= 0) r.d[dir] = 1.0;
    return r;
  }
};

// value of a scalar of either kind (branch tests)
__host__ __device__ inline double sval(double x) { return x; }
template <int N> __host__ __device__ inline double sval(const Dual<N> &x) { return x.v; }

template <int N> __host__ __device__ inline Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N> &a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N> __host__ __device__ inline Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int N> __host__ __device__ inline Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
// mixed forms: a double operand is a constant
template <int N> __host__ __device__ inline Dual<N> operator+(const Dual<N> &a, double b) { return a + Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator+(double a, const Dual<N> &b) { return Dual<N>(a) + b; }
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N> &a, double b) { return a - Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator-(double a, const Dual<N> &b) { return Dual<N>(a) - b; }
template <int N> __host__ __device__ inline Dual<N> operator*(const Dual<N> &a, double b) {
  Dual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int N> __host__ __device__ inline Dual<N> operator*(double a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int N> __host__ __device__ inline Dual<N> operator/(const Dual<N> &a, double b) {
  Dual<N> r;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
  return r;
}
template <int N> __host__ __device__ inline Dual<N> operator/(double a, const Dual<N> &b) { return Dual<N>(a) / b; }
template <int N> __host__ __device__ inline Dual<N> &operator+=(Dual<N> &a, const Dual<N> &b) { return a = a + b; }
template <int N> __host__ __device__ inline Dual<N> &operator-=(Dual<N> &a, const Dual<N> &b) { return a = a - b; }
template <int N> __host__ __device__ inline Dual<N> &operator*=(Dual<N> &a, const Dual<N> &b) { return a = a * b; }
template <int N> __host__ __device__ inline Dual<N> &operator+=(Dual<N> &a, double b) { return a = a + b; }
template <int N> __host__ __device__ inline Dual<N> &operator-=(Dual<N> &a, double b) { return a = a - b; }
template <int N> __host__ __device__ inline Dual<N> &operator*=(Dual<N> &a, double b) { return a = a * b; }

// comparisons act on values
template <int N> __host__ __device__ inline bool operator<(const Dual<N> &a, const Dual<N> &b) { return a.v < b.v; }
template <int N> __host__ __device__ inline bool operator<(const Dual<N> &a, double b) { return a.v < b; }
template <int N> __host__ __device__ inline bool operator<(double a, const Dual<N> &b) { return a < b.v; }
template <int N> __host__ __device__ inline bool operator>(const Dual<N> &a, double b) { return a.v > b; }
template <int N> __host__ __device__ inline bool operator>(const Dual<N> &a, const Dual<N> &b) { return a.v > b.v; }
template <int N> __host__ __device__ inline bool operator<=(const Dual<N> &a, double b) { return a.v <= b; }
template <int N> __host__ __device__ inline bool operator<=(const Dual<N> &a, const Dual<N> &b) { return a.v <= b.v; }

// elementary functions: value as the double function, derivative by the
// chain rule (zero where the derivative does not exist)
template <int N> __device__ inline Dual<N> sqrt(const Dual<N> &a) {
  Dual<N> r;
  r.v = ::sqrt(a.v);
  const double s = r.v > 0 ? 0.5 / r.v : 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
  return r;
}
template <int N> __device__ inline Dual<N> fabs(const Dual<N> &a) {
  Dual<N> r;
  r.v = ::fabs(a.v);
  const double s = a.v > 0 ? 1.0 : (a.v < 0 ? -1.0 : 0.0);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
  return r;
}
template <int N> __device__ inline Dual<N> asin(const Dual<N> &a) {
  Dual<N> r;
  r.v = ::asin(a.v);
  const double q = 1.0 - a.v * a.v;
  const double s = q > 0 ? 1.0 / ::sqrt(q) : 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
  return r;
}

} // namespace pg

#endif
