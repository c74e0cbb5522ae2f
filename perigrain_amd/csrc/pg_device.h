// pg_device.h -- device helpers shared by the step kernels and the adjoint
// kernels: exact-rounding division/square root, record accessors, grain
// boxes, and the contact and wall force functions.  The force functions are
// templates on the scalar type: double for the time step, Dual<N> (pg_dual.h)
// for the adjoint, which evaluates the same operations in the same order and
// so takes the same branches.
#ifndef PG_DEVICE_H
#define PG_DEVICE_H
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "pg_dual.h"
#include "pg_internal.h"

namespace pg {

// ---------------------------------------------------------------- helpers

template <int D> __device__ __forceinline__ double red(const double *a, const double *b);
template <> __device__ __forceinline__ double red<2>(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1];
}
template <> __device__ __forceinline__ double red<3>(const double *a, const double *b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
// the same sums on dual numbers
template <int D, int N> __device__ __forceinline__ Dual<N> red(const Dual<N> *a, const Dual<N> *b) {
  if (D == 2) return a[0] * b[0] + a[1] * b[1];
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// ---- correctly rounded division and square root without range fix-ups.
//
// The compiler lowers an IEEE binary64 a/b on gfx950 to
//   v_div_scale (den), v_rcp, two Newton steps r += r*(1-b*r), v_div_scale
//   (num), q = a*r, rem = a - b*q (fma), v_div_fmas q + rem*r, v_div_fixup,
// and sqrt(x) to a scale-by-2^256 test, v_rsq, a Goldschmidt/Newton pair of
// corrections and a special-value select.  For operands well inside the
// normal range the scale steps leave the operands unchanged and the fix-ups
// pass the result through, so the sequences below -- the same operations
// minus those steps -- return bit-identical, correctly rounded results
// there.  The reciprocal depends on the denominator only, so one serves
// several numerators.  Callers check the range (fast_*_ok) and otherwise use
// the IEEE operators.
__device__ __forceinline__ double recip_nr(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  return r;
}
__device__ __forceinline__ double div_nr(double a, double b, double r) {
  const double q = a * r;
  const double rem = fma(-b, q, a);
  return fma(rem, r, q);
}
__device__ __forceinline__ double sqrt_nr(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  double g = x * r;
  double h = r * 0.5;
  const double e = fma(-h, g, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  g = fma(d, h, g);
  return g;
}
// sqrt_nr with the seed taken at max(x, 2^-1000): the same for x >= 2^-1000,
// and exactly +0 for x = +0 (where rsq(0) = inf would give a NaN).
__device__ __forceinline__ double sqrt_nr0(double x) {
  const double r = __builtin_amdgcn_rsq(fmax(x, 0x1p-1000));
  double g = x * r;
  double h = r * 0.5;
  const double e = fma(-h, g, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  g = fma(d, h, g);
  return g;
}
// Ranges where the shortened sequences equal the IEEE ones: denominator in
// [2^-300, 2^300]; numerator zero (handled by div_r) or in
// [2^-300, 2^300]; square-root argument in [2^-600, 2^600].
__device__ __forceinline__ bool fast_den_ok(double b) { return b >= 0x1p-300 && b <= 0x1p300; }
__device__ __forceinline__ bool fast_num_ok(double a) {
  const double m = fabs(a);
  return a == 0.0 || (m >= 0x1p-300 && m <= 0x1p300);
}
__device__ __forceinline__ bool fast_sqrt_ok(double x) { return x >= 0x1p-600 && x <= 0x1p600; }
// x in [2^-600, 2^601) for x >= 0, from the high word: biased exponent in
// [423, 1623] (zero, NaN and infinities fail)
__device__ __forceinline__ bool sqrt_arg_in_range(double x) {
  const uint32_t hi = (uint32_t)((unsigned long long)__double_as_longlong(x) >> 32);
  return hi - 0x1A700000u < 0x4B100000u;
}
// "Nice" coordinates make the bond loop's per-bond range test unnecessary.
// A double x is nice if it is zero or 2^-240 <= |x| <= 2^240.  Nice values
// are integer multiples of 2^-292 (their ulp is at least 2^-240-52), so if
// every component of X and u is nice:
//   xi_q = X_jq - X_iq and u_jq - u_iq are zero or multiples of 2^-292 --
//   the exact difference is such a multiple, and rounding keeps it one
//   (below 2^-239 it is exact, above, the ulp is at least 2^-291) -- and so
//   is eta_q = (u_jq - u_iq) + xi_q by the same argument.  Each eta_q is then
//   zero or at least 2^-292 >= 2^-300 in magnitude; |eta|^2 is zero or at
//   least 2^-584 >= 2^-600 (and |xi|^2 likewise, nonzero for a bond of
//   nonzero length); and all of them stay below 2^601 since every
//   coordinate is at most 4 * 2^240 in magnitude.
// That is the fast range of bond_geom, except for eta = 0.
__host__ __device__ __forceinline__ bool nice(double x) {
  const double m = fabs(x);
  return m == 0.0 || (m >= 0x1p-240 && m <= 0x1p240);
}

// a/b through a precomputed reciprocal; a zero numerator returns itself
// (the IEEE quotient of +-0 by a positive number).
__device__ __forceinline__ double div_r(double a, double b, double r) {
  return a == 0.0 ? a : div_nr(a, b, r);
}
// Stand-alone forms (range-checked), used by the diagnostics kernel.
__device__ __forceinline__ double div_rn(double a, double b) {
  return (fast_den_ok(b) && fast_num_ok(a)) ? div_r(a, b, recip_nr(b)) : a / b;
}
__device__ __forceinline__ double sqrt_rn(double x) {
  return fast_sqrt_ok(x) ? sqrt_nr(x) : sqrt(x);
}

// Logical block index that gives each XCD (blocks are dealt round-robin to
// the 8 XCDs) one contiguous range of nodes, so neighbour gathers of a block
// hit the L2 of the XCD that already holds its neighbours.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t q = nb >> 3, r = nb & 7;
  const uint32_t x = b & 7, idx = b >> 3;
  return (x < r) ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

__device__ __forceinline__ bool grain_movable(const GrainDev &gp, const InstDev &ip,
                                              uint32_t g, uint32_t t) {
  return (gp.flags & GF_MOVABLE) ||
         (ip.set_movable_grain == (int32_t)g && t >= ip.set_movable_step);
}
__device__ __forceinline__ bool grain_stoppable(const GrainDev &gp, const InstDev &ip,
                                                uint32_t g, uint32_t t) {
  return (gp.flags & GF_STOPPABLE) ||
         (ip.set_stoppable_grain == (int32_t)g && t >= ip.set_stoppable_step);
}

// Step index of a launch: the argument itself, or, for launches replayed
// from a captured graph, the argument added to the device-side base step.
// The value is the same in every lane; readfirstlane lets the compiler keep
// it, and what derives from it (record array bases), in scalar registers.
__device__ __forceinline__ uint32_t step_index(const StepArgs &A, uint32_t t_arg) {
  return __builtin_amdgcn_readfirstlane(A.tbase ? *A.tbase + t_arg : t_arg);
}

typedef double __attribute__((ext_vector_type(2))) dvec2;
// Word 3 of a raw (untyped) buffer descriptor on gfx9-family GPUs: data
// format 32 bits, no index stride, no thread-id addressing.
constexpr int kBufWord3 = 0x00020000;

// Record accessors (records are [D][ns][2], rec_index).
template <int D>
__device__ __forceinline__ void rec_X(const StepArgs &A, uint32_t i, double *X) {
#pragma unroll
  for (int k = 0; k < D; ++k) X[k] = A.rec[0][rec_index(A.ns, i, k)]; // X: both parities
}
// The reference position `pos` of node i of grain g: X, plus the y shift of
// reset_partzero_y on the shifted grain (timeloop.h:646-648 adds it to pos in
// place).  Positions (CurrPos, contact, walls, boxes, the self-contact
// reference distance) use it; bonds keep X, since the reference does not
// regenerate xi after the shift.
template <int D>
__device__ __forceinline__ void pos_X(const StepArgs &A, const InstDev &ip, uint32_t g,
                                      uint32_t i, double *X) {
  rec_X<D>(A, i, X);
  if ((int32_t)g == ip.yshift_grain) X[1] = X[1] + ip.yshift;
}
template <int D>
__device__ __forceinline__ void rec_u(const double *R, uint32_t ns, uint32_t i, double *u) {
#pragma unroll
  for (int k = 0; k < D; ++k) u[k] = R[rec_index(ns, i, D + k)];
}
template <int D>
__device__ __forceinline__ void rec_set_u(double *R, uint32_t ns, uint32_t i, const double *u) {
#pragma unroll
  for (int k = 0; k < D; ++k) R[rec_index(ns, i, D + k)] = u[k];
}

template <int D>
__device__ __forceinline__ void load_vec(const double *p, uint32_t ns, uint32_t i, double *v) {
#pragma unroll
  for (int k = 0; k < D; ++k) v[k] = p[(size_t)k * ns + i];
}
template <int D>
__device__ __forceinline__ void store_vec(double *p, uint32_t ns, uint32_t i, const double *v) {
#pragma unroll
  for (int k = 0; k < D; ++k) p[(size_t)k * ns + i] = v[k];
}

// bbox key arrays are [grain][2*D]: mins then maxes.
template <int D>
__device__ __forceinline__ void load_box(const unsigned long long *bb, uint32_t g, double *b) {
#pragma unroll
  for (int k = 0; k < 2 * D; ++k) b[k] = dkey_inv(bb[(size_t)g * 2 * D + k]);
}

// intersects_box_fext (contact.h:983-1020).  Boxes are stored as
// [min_0..min_{D-1}, max_0..max_{D-1}]; in 2D the reference's layout is
// [xmin,xmax,ymax,ymin], so its A_Left/A_Right/A_Top/A_Bottom are
// A[0]/A[2]/A[3]/A[1] here.  In 3D the two layouts coincide.
template <int D>
__device__ __forceinline__ bool boxes_meet(const double *A, const double *B, double c) {
  if (D == 2) {
    const double AL = A[0] - c, AR = A[2] + c, AT = A[3] + c, AB = A[1] - c;
    return (AL <= B[2]) && (AR >= B[0]) && (AT >= B[1]) && (AB <= B[3]);
  } else {
    const double ax0 = A[0] - c, ay0 = A[1] - c, az0 = A[2] - c;
    const double ax1 = A[3] + c, ay1 = A[4] + c, az1 = A[5] + c;
    return (ax0 <= B[3]) && (ax1 >= B[0]) && (ay1 >= B[1]) && (ay0 <= B[4]) &&
           (az0 <= B[5]) && (az1 >= B[2]);
  }
}

// within_interior_fext (contact.h:1066-1080) on a [min..,max..] box `bx`
// and the walls in lrtb() order.  The reference indexes the first four
// entries of get_minmax's output whatever the dimension: 2D
// [xmin,xmax,ymax,ymin] vs [left,right,top,bottom]; 3D [xmin,ymin,zmin,xmax]
// vs [x_min,y_min,z_min,x_max] (so in 3D the test is practically never
// true and the walls are always evaluated).
template <int D>
__device__ __forceinline__ bool within_interior(const double *bx, const double *w, double c) {
  double A0, A1, A2, A3;
  if (D == 2) {
    A0 = bx[0]; A1 = bx[2]; A2 = bx[3]; A3 = bx[1];
  } else {
    A0 = bx[0]; A1 = bx[1]; A2 = bx[2]; A3 = bx[3];
  }
  const double AL = A0 - c, AR = A1 + c, AT = A2 + c, AB = A3 - c;
  return (w[0] < AL) && (AR < w[1]) && (w[2] > AT) && (w[3] < AB);
}

// segment_mass<dim> (contact.h:16-41), returns area, writes centroid.  T is
// double (time step) or Dual<N> (adjoint); R is a constant.
template <int D, class T>
__device__ __forceinline__ T segment_mass(double R, T d, T &centroid) {
  T Area, I;
  if (D == 2) {
    Area = M_PI * (R * R) / 2 - (R * R) * asin(d / R) - sqrt(R * R - d * d) * d;
    I = T(0.0); // 2 / 3 * pow(..., 1.5) with integer 2/3 == 0 and a finite pow
  } else {
    Area = M_PI / 3 * ((R - d) * (R - d)) * (2 * R - d);
    I = M_PI / 4 * ((R * R - d * d) * (R * R - d * d));
  }
  centroid = I / Area;
  return Area;
}

// wall_side_force_update (contact.h:310-431).  Returns true if the node is
// within contact distance of this side.  Kn and mu are the instance's normal
// stiffness and friction coefficient, passed as T so that the adjoint can
// differentiate along them.  If t4 is given it receives the reaction terms
// {normal, damping, friction t1, friction t2} (each already vol-scaled).
template <int D, class T>
__device__ __forceinline__ bool wall_side(T dist, int dir, const InstDev &ip, T Kn, T mu,
                                          double rho, double vol, double wspeed, const T *v,
                                          T *F, double *t4) {
  if (!(fabs(dist) <= ip.Rc)) return false;
  T cen;
  const T area = segment_mass<D, T>(ip.Rc, fabs(dist), cen);
  const T contrib = (ip.Rc - cen);
  const T pos = (contrib > 0 ? contrib : T(0));
  const T Fw = (Kn * pos * area);
  const double sgn = (dist > 0 ? 1.0 : -1.0);
  F[dir] += (sgn * Fw);
  if (t4) {
    t4[0] = (vol * sgn * sval(Fw));
    t4[1] = 0.0;
    t4[2] = 0.0;
    t4[3] = 0.0;
  }
  T rel[D];
#pragma unroll
  for (int k = 0; k < D; ++k) rel[k] = v[k];
  rel[dir] -= wspeed;
  if (ip.flags & IF_FRICTION) {
    if (D == 2) {
      const int tg = (dir + 1) % 2;
      if (fabs(rel[tg]) > 0) {
        const double sg = (rel[tg] > 0 ? 1.0 : -1.0);
        T ff = (-mu) * fabs(Fw) * sg;
        if (fabs(rel[tg]) < mu * fabs(Fw) * ip.dt / rho) ff = -rel[tg] * rho / ip.dt;
        F[tg] += ff;
        if (t4) t4[2] = (vol * sval(ff));
      }
    } else {
      const int t1 = (dir + 1) % 3, t2 = (dir + 2) % 3;
      const T vp1 = rel[t1], vp2 = rel[t2];
      const T vpn = sqrt(vp1 * vp1 + vp2 * vp2);
      if (vpn > 0) {
        const T u1 = vp1 / vpn, u2 = vp2 / vpn;
        T ff1 = (-mu) * fabs(Fw) * u1;
        T ff2 = (-mu) * fabs(Fw) * u2;
        if (vpn < mu * fabs(Fw) * ip.dt / rho) {
          ff1 = -vpn / rho / ip.dt * u1;
          ff2 = -vpn / rho / ip.dt * u2;
        }
        F[t1] += ff1;
        F[t2] += ff2;
        if (t4) {
          t4[2] = (vol * sval(ff1));
          t4[3] = (vol * sval(ff2));
        }
      }
    }
  }
  if (ip.flags & IF_DAMPING) {
    const T cd = 2 * ip.zeta * sqrt(Kn * rho) * sqrt(fabs(area));
    const T fd = (-cd) * rel[dir];
    F[dir] += fd;
    if (t4) t4[1] = (vol * sval(fd));
  }
  return true;
}

// Normal, damping and friction force of one contact pair
// (contact.h:735-808).  d = x_j - x_i, r = |d|, R/K the pair's radius and
// stiffness, vj the partner's volume, w = v_i - v_j, mu the friction
// coefficient.
template <int D, class T>
__device__ __forceinline__ void pair_force(const T *d, T r, double R, T K, double vj, const T *w,
                                           double rho, T mu, const InstDev &ip, T *F) {
  T e[D], fc[D];
#pragma unroll
  for (int k = 0; k < D; ++k) e[k] = d[k] / r;
  const T g = -K * (R - r) * vj;
#pragma unroll
  for (int k = 0; k < D; ++k) fc[k] = g * e[k];
  const T fcn = sqrt(red<D>(fc, fc));
#pragma unroll
  for (int k = 0; k < D; ++k) F[k] += fc[k];
  if (ip.flags & IF_DAMPING) {
    const T p = red<D>(w, e);
    const T cd = 2 * ip.zeta * sqrt(K * rho) * sqrt(vj);
#pragma unroll
    for (int k = 0; k < D; ++k) F[k] += (-cd) * (p * e[k]);
  }
  if (ip.flags & IF_FRICTION) {
    const T p = red<D>(w, e);
    T tt[D];
#pragma unroll
    for (int k = 0; k < D; ++k) tt[k] = w[k] - p * e[k];
    const T tau = sqrt(red<D>(tt, tt));
    if (tau > 0) {
      T ut[D];
#pragma unroll
      for (int k = 0; k < D; ++k) ut[k] = tt[k] / tau;
      T s = -mu * fcn;
      if (tau < mu * fcn * ip.dt / rho) s = -tau / rho / ip.dt;
#pragma unroll
      for (int k = 0; k < D; ++k) F[k] += s * ut[k];
    }
  }
}

// pair_force's three terms without the sums, for callers that add them
// later in pair_force's order: F += fc, then F += fd, then F += ff.  A term
// pair_force does not add is +0 here; adding +0 leaves those sums unchanged
// (they start at +0, so they are never -0).  The operations are pair_force's
// on doubles.
template <int D>
__device__ __forceinline__ void pair_terms(const double *d, double r, double R, double K,
                                           double vj, const double *w, double rho, double mu,
                                           const InstDev &ip, double *fc, double *fd, double *ff) {
  double e[D];
#pragma unroll
  for (int k = 0; k < D; ++k) e[k] = d[k] / r;
  const double g = -K * (R - r) * vj;
#pragma unroll
  for (int k = 0; k < D; ++k) fc[k] = g * e[k];
  const double fcn = sqrt(red<D>(fc, fc));
#pragma unroll
  for (int k = 0; k < D; ++k) {
    fd[k] = 0.0;
    ff[k] = 0.0;
  }
  if (ip.flags & IF_DAMPING) {
    const double p = red<D>(w, e);
    const double cd = 2 * ip.zeta * sqrt(K * rho) * sqrt(vj);
#pragma unroll
    for (int k = 0; k < D; ++k) fd[k] = (-cd) * (p * e[k]);
  }
  if (ip.flags & IF_FRICTION) {
    const double p = red<D>(w, e);
    double tt[D];
#pragma unroll
    for (int k = 0; k < D; ++k) tt[k] = w[k] - p * e[k];
    const double tau = sqrt(red<D>(tt, tt));
    if (tau > 0) {
      double ut[D];
#pragma unroll
      for (int k = 0; k < D; ++k) ut[k] = tt[k] / tau;
      double s = -mu * fcn;
      if (tau < mu * fcn * ip.dt / rho) s = -tau / rho / ip.dt;
#pragma unroll
      for (int k = 0; k < D; ++k) ff[k] = s * ut[k];
    }
  }
}

} // namespace pg

#endif
