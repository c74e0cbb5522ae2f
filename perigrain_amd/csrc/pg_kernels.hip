// pg_kernels.hip -- gfx950 kernels of libpg (HIP, FP64).
//
// Arithmetic contract: every floating-point expression below is written in
// the operation order of the reference (file:line cited per function) and
// the file is compiled with -ffp-contract=off, so each +,-,*,/ is a single
// IEEE binary64 rounding and no FMA is formed.  Division and square root are
// correctly rounded (the IEEE operators, or the shortened sequences of
// pg_device.h where they are provably identical).  3-component norms and dot products
// sum as (a0*b0 + a1*b1) + a2*b2, the order of Eigen 3.3.9's vectorised
// redux that the reference's norm()/dot() use.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "pg_internal.h"
#include "pg_kernels.h"

// Waves per SIMD the one-lane step kernel is compiled for: 4 caps it at 128
// VGPRs, the residency the bond loop's FP64 chains need.  The kernels with
// several lanes per node run at most two wavefronts per SIMD (64k nodes
// times two lanes, 16k times four), so they are compiled for two: up to 256
// VGPRs, which the inputs they load up front take without spilling.
#ifndef PG_STEP_WAVES_PER_EU
#define PG_STEP_WAVES_PER_EU 4
#endif
// 1: bond loop uses the shortened division/square-root sequences
#ifndef PG_FAST_DIVSQRT
#define PG_FAST_DIVSQRT 1
#endif

#include "pg_device.h"

namespace pg {

// Partner data a bond needs: reference position, volume, displacement.
template <int D> struct Nbr {
  double X[D];
  double V;
  double u[D];
};

// Length |xi|, stretch s = (|eta| - |xi|)/|xi| and direction e = eta/|eta| of
// one bond (particle2.h:134-146), each a correctly rounded IEEE result.
// They go through the shortened division and square-root sequences of
// pg_device.h, which are exact when xin2 = |xi|^2 and n2 = |eta|^2 lie in
// [2^-600, 2^601) (so n, |xi| lie in [2^-300, 2^301)) and each eta_q is zero
// or of magnitude >= 2^-300: then every numerator is zero or of magnitude in
// [2^-352, 2^301] (at most n or max(n, |xi|)) over a denominator in
// [2^-300, 2^301), where the sequences need no scaling.  The numerator of s,
// n - |xi|, needs no test: n and |xi| are >= 2^-300, so both are multiples of
// 2^-352 and their difference is zero or at least 2^-352 in magnitude.  A
// zero numerator gives a zero quotient whose sign may differ from the IEEE
// one; a zero of either sign adds nothing to a force and compares equal in
// the break test, so the results are the same.
//
// CHECK: test that range for this bond (bond_fast_range) and use the IEEE
// operators outside it; only live bonds take the IEEE path (dead slots'
// values are discarded).  !CHECK: the caller has established the range for
// every bond of the step (every coordinate of X and u is nice(), no bond has
// xi = 0) except eta = 0 (coincident nodes), where the IEEE results are
// n = 0, s = -1 exactly and e = 0/0, a NaN.  The square root of |eta|^2
// seeds from rsq(max(|eta|^2, 2^-1000)) (sqrt_nr0), which changes nothing in
// the range and returns n = 0 for |eta|^2 = 0, so s = -|xi|/|xi| = -1
// exactly and e is a NaN (NaN payloads are not compared: the host's and the
// device's default NaNs differ anyway).
__device__ __forceinline__ bool bond_fast_range(double xin2, double n2, const double *eta, int D) {
  // frexp exponent: zero -> 0, |x| >= 2^-300 -> >= -299
  int emin = __builtin_amdgcn_frexp_exp(eta[0]);
  for (int q = 1; q < D; ++q) emin = min(emin, __builtin_amdgcn_frexp_exp(eta[q]));
  return sqrt_arg_in_range(xin2) && sqrt_arg_in_range(n2) && emin >= -299;
}
template <int D, bool CHECK>
__device__ __forceinline__ void bond_geom(const double *xi, const double *eta, bool live,
                                          double &s, double *e) {
  const double xin2 = red<D>(xi, xi);
  const double n2 = red<D>(eta, eta);
#if PG_FAST_DIVSQRT
  double xin = sqrt_nr(xin2), n = CHECK ? sqrt_nr(n2) : sqrt_nr0(n2);
  const double dn = n - xin;
  s = div_nr(dn, xin, recip_nr(xin));
  const double rn = recip_nr(n);
#pragma unroll
  for (int q = 0; q < D; ++q) e[q] = div_nr(eta[q], n, rn);
  if (CHECK) {
    if (live && !bond_fast_range(xin2, n2, eta, D)) { // outside the fast range: IEEE operators
      xin = sqrt(xin2);
      n = sqrt(n2);
      s = (n - xin) / xin;
#pragma unroll
      for (int q = 0; q < D; ++q) e[q] = eta[q] / n;
    }
  }
#else
  const double xin = sqrt(xin2);
  const double n = sqrt(n2);
  s = (n - xin) / xin;
#pragma unroll
  for (int q = 0; q < D; ++q) e[q] = eta[q] / n;
#endif
}

// One bond of get_peridynamic_force (particle2.h:125-149) and its break test
// (remove_bonds, particle2.h:207), evaluated for every slot; broken and
// padding slots (ghost partner, finite values) get a zero weight and add a
// zero of either sign (the sum starts at +0.0 and can never become -0.0, so
// adding a zero leaves it bit-identical).  A bond whose stretch exceeds s0
// gets the broken flag in its column word.
template <int D, bool CHECK>
__device__ __forceinline__ void bond_term(uint32_t cw, const Nbr<D> &nb, const double *Xo,
                                          const double *ui, double cnot, double snot, bool brk,
                                          uint32_t *slot, double *F, const StepArgs &A,
                                          uint32_t t, bool &broke) {
  const bool live = !(cw & BCOL_BROKEN);
  double xi[D], eta[D], e[D], s;
#pragma unroll
  for (int q = 0; q < D; ++q) xi[q] = nb.X[q] - Xo[q];
#pragma unroll
  for (int q = 0; q < D; ++q) eta[q] = (nb.u[q] - ui[q]) + xi[q];
  bond_geom<D, CHECK>(xi, eta, live, s, e);
  const double w = live ? (cnot * s) * nb.V : 0.0;
#pragma unroll
  for (int q = 0; q < D; ++q) F[q] = F[q] + w * e[q];
  if (live && brk && s > snot) {
    *slot = cw | BCOL_BROKEN;
    broke = true;
    if (A.break_step) A.break_step[slot - A.bcol] = t;
  }
}

// bond_term's arithmetic without the sum: p = w*e, the term bond_term adds
// to F (F[q] + p[q] is F[q] + w*e[q] with the same two roundings).
template <int D, bool CHECK>
__device__ __forceinline__ bool bond_prod(uint32_t cw, const Nbr<D> &nb, const double *Xo,
                                          const double *ui, double cnot, double snot, bool brk,
                                          double *p) {
  const bool live = !(cw & BCOL_BROKEN);
  double xi[D], eta[D], e[D], s;
#pragma unroll
  for (int q = 0; q < D; ++q) xi[q] = nb.X[q] - Xo[q];
#pragma unroll
  for (int q = 0; q < D; ++q) eta[q] = (nb.u[q] - ui[q]) + xi[q];
  bond_geom<D, CHECK>(xi, eta, live, s, e);
  const double w = live ? (cnot * s) * nb.V : 0.0;
#pragma unroll
  for (int q = 0; q < D; ++q) p[q] = w * e[q];
  return live && brk && s > snot;
}

// A bond whose stretch exceeded s0 gets the broken flag in its column word
// (and, in a recorded run, its break step); bond_term does the same inline.
__device__ __forceinline__ void bond_break(uint32_t cw, uint32_t *slot, const StepArgs &A,
                                           uint32_t t) {
  *slot = cw | BCOL_BROKEN;
  if (A.break_step) A.break_step[slot - A.bcol] = t;
}

// Contact gates of a list entry (the gates of timeloop.h:698-747 and the
// self-contact rules of contact.h:686-733): partner j of grain gj, reference
// position Xj, bond slot bs (-1: no bond, or one known to be broken).
// Returns whether the pair is evaluated at step t, with its contact radius R
// and stiffness K.
template <int D>
__device__ __forceinline__ bool contact_gate(const StepArgs &A, uint32_t t, uint32_t gj,
                                            int32_t bs, uint32_t g, const GrainDev &gp,
                                            const InstDev &ip, const double *Xc,
                                            const double *Xj, bool selfc, bool stop,
                                            const double *boxg, double &R, double &K) {
  R = ip.Rc;
  K = ip.Kn;
  if (gj == g) {
    if (!selfc) return false;
    if (bs >= 0 && !(A.bcol[bs] & BCOL_BROKEN)) return false; // intact bond
    double dX[D];
#pragma unroll
    for (int k = 0; k < D; ++k) dX[k] = Xc[k] - Xj[k];
    const double rd = sqrt(red<D>(dX, dX));
    if (rd < ip.Rc) {
      R = rd;
      K = gp.cnot / rd;
    } else {
      R = ip.Rc;
      K = gp.cnot / ip.Rc;
    }
    return true;
  }
  if (!stop) return false;
  double boxj[2 * D];
  load_box<D>(A.bbox[t % 3], gj, boxj);
  return boxes_meet<D>(boxg, boxj, ip.Rc);
}

// Relative position d = x_j - x_i and distance r of a pair; true if r <= R.
template <int D>
__device__ __forceinline__ bool pair_within(const double *Xj, const double *uj, const double *xc,
                                           double R, double *d, double &r) {
#pragma unroll
  for (int k = 0; k < D; ++k) d[k] = (Xj[k] + uj[k]) - xc[k];
  r = sqrt(red<D>(d, d));
  return r <= R;
}

// Wavefront min/max of D coordinates grouped by `grp`, combined into
// keys[grp][2*D] with 64-bit atomics: one shuffle reduction and one atomic per
// key for each distinct group in the wavefront (a wavefront that straddles
// two grains makes two rounds), so that no key sees more than one atomic per
// wavefront.  All 64 lanes call it together.
template <int D>
__device__ __forceinline__ void wave_box_atomic(unsigned long long *keys, uint32_t grp,
                                                bool valid, const double *x) {
  unsigned long long key[D];
#pragma unroll
  for (int k = 0; k < D; ++k) key[k] = dkey(x[k]);
  bool todo = valid && grp != 0xFFFFFFFFu;
  for (;;) {
    const unsigned long long m = __ballot(todo);
    if (m == 0) break;
    const uint32_t g0 = __builtin_amdgcn_readlane(grp, (int)__builtin_ctzll(m));
    const bool in = todo && grp == g0;
    unsigned long long kmin[D], kmax[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      kmin[k] = in ? key[k] : ~0ull;
      kmax[k] = in ? key[k] : 0ull;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const unsigned long long a = __shfl_xor(kmin[k], off, 64);
        const unsigned long long b = __shfl_xor(kmax[k], off, 64);
        kmin[k] = a < kmin[k] ? a : kmin[k];
        kmax[k] = b > kmax[k] ? b : kmax[k];
      }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        atomicMin(&keys[(size_t)g0 * 2 * D + k], kmin[k]);
        atomicMax(&keys[(size_t)g0 * 2 * D + D + k], kmax[k]);
      }
    }
    todo = todo && !in;
  }
}

// Grain boxes combined per workgroup first: the boxes of the workgroup's
// first BOX_SLOTS grains (nodes are grain-ordered, so a workgroup's grains
// are consecutive from its first node's grain g0) collect in LDS, and the
// workgroup's last wavefront to finish adds them to the global keys, so a
// key sees one global atomic per workgroup instead of one per wavefront.
// Grains beyond the slots go straight to the global keys.
constexpr uint32_t BOX_SLOTS = 4;
struct BlockBox {
  unsigned long long key[BOX_SLOTS][6];
  uint32_t g0;   // grain of the workgroup's first node (0xFFFFFFFF: none)
  uint32_t done; // wavefronts that have added their boxes
};

// Set up by every thread of the workgroup, followed by a barrier.
template <int D>
__device__ __forceinline__ void block_box_init(BlockBox &bb, const StepArgs &A,
                                               uint32_t first_node) {
  if (threadIdx.x < BOX_SLOTS * 2 * D) {
    const uint32_t k = threadIdx.x % (2 * D);
    bb.key[threadIdx.x / (2 * D)][k] = k < (uint32_t)D ? ~0ull : 0ull;
  }
  if (threadIdx.x == 0) {
    bb.g0 = first_node < A.n_nodes ? A.node_grain[first_node] : 0xFFFFFFFFu;
    bb.done = 0;
  }
}

// wave_box_atomic's reduction, then the workgroup combination; all lanes of
// every wavefront of the workgroup call it, `calls` calls per workgroup in
// all (the last one adds the workgroup's boxes to the global keys).
template <int D>
__device__ __forceinline__ void wave_box_block(unsigned long long *keys, uint32_t grp, bool valid,
                                               const double *x, BlockBox &bb, uint32_t calls) {
  unsigned long long key[D];
#pragma unroll
  for (int k = 0; k < D; ++k) key[k] = dkey(x[k]);
  const bool lane0 = (threadIdx.x & 63) == 0;
  bool todo = valid && grp != 0xFFFFFFFFu;
  for (;;) {
    const unsigned long long m = __ballot(todo);
    if (m == 0) break;
    const uint32_t g0 = __builtin_amdgcn_readlane(grp, (int)__builtin_ctzll(m));
    const bool in = todo && grp == g0;
    unsigned long long kmin[D], kmax[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      kmin[k] = in ? key[k] : ~0ull;
      kmax[k] = in ? key[k] : 0ull;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const unsigned long long a = __shfl_xor(kmin[k], off, 64);
        const unsigned long long b = __shfl_xor(kmax[k], off, 64);
        kmin[k] = a < kmin[k] ? a : kmin[k];
        kmax[k] = b > kmax[k] ? b : kmax[k];
      }
    }
    if (lane0) {
      const uint32_t slot = g0 - bb.g0; // g0 >= bb.g0: the first node's grain is the smallest
      if (slot < BOX_SLOTS) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
          atomicMin(&bb.key[slot][k], kmin[k]);
          atomicMax(&bb.key[slot][D + k], kmax[k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < D; ++k) {
          atomicMin(&keys[(size_t)g0 * 2 * D + k], kmin[k]);
          atomicMax(&keys[(size_t)g0 * 2 * D + D + k], kmax[k]);
        }
      }
    }
    todo = todo && !in;
  }
  // this wavefront's LDS atomics are done before it counts itself; the
  // wavefront that counts last sees every other wavefront's
  __threadfence_block();
  uint32_t prev = 0;
  if (lane0) prev = atomicAdd(&bb.done, 1u);
  prev = __builtin_amdgcn_readfirstlane(prev);
  if (prev + 1 == calls) {
    __threadfence_block();
    const uint32_t l = threadIdx.x & 63;
    if (l < BOX_SLOTS * 2 * D && bb.g0 != 0xFFFFFFFFu) {
      const uint32_t slot = l / (2 * D), k = l % (2 * D);
      const unsigned long long v = bb.key[slot][k];
      // an untouched slot holds the identity, whose atomic would change nothing
      if (k < (uint32_t)D) {
        if (v != ~0ull) atomicMin(&keys[(size_t)(bb.g0 + slot) * 2 * D + k], v);
      } else {
        if (v != 0ull) atomicMax(&keys[(size_t)(bb.g0 + slot) * 2 * D + k], v);
      }
    }
  }
}

// The PMB bond loop of node i (particle2.h:112-152): its row's slots in
// order, software-pipelined -- the partner data of the next two bonds and the
// column words of the two after those are in flight while two bonds are
// evaluated.
template <int D, bool CHECK, class RS>
__device__ __forceinline__ void bond_loop(const StepArgs &A, uint32_t t, uint32_t i, const RS &rs,
                                          const double *Xo, const double *ui, double cnot,
                                          double snot, double gvol, bool brk, double *F,
                                          bool &broke) {
  const uint32_t sl = i >> 6, lane = i & 63;
  const uint32_t base = A.slice_base[sl], width = A.slice_width[sl];
  // column words of this lane: group q (slots 4q..4q+3) at bc + 256*q
  uint32_t *bc = A.bcol + base + BGROUP * lane;
  // They are read by buffer loads over the slice (a descriptor with the
  // slice's first slot as base, so offsets stay small; its range check is
  // left open, bounds come from BCOL_TAIL as for the other loads): on the
  // 3D compression workload 1% faster than the same loads through bc.
  // (the wavefront's lanes share the slice, so the base is uniform and the
  // descriptor stays in scalar registers)
  const __amdgpu_buffer_rsrc_t brs = __builtin_amdgcn_make_buffer_rsrc(
      static_cast<void *>(A.bcol + __builtin_amdgcn_readfirstlane(base)), (short)0, 0x7FFFFFFF,
      kBufWord3);
  // the column array has BCOL_TAIL spare slots after the last slice, so
  // reading two groups past a row's end stays in bounds
  auto colg = [&](uint32_t q) -> uint4 {
    typedef unsigned int uv4 __attribute__((ext_vector_type(4)));
    const uv4 v = __builtin_bit_cast(
        uv4, __builtin_amdgcn_raw_buffer_load_b128(
                 brs, (BGROUP * lane + (SLICE * BGROUP) * q) * (uint32_t)sizeof(uint32_t), 0, 0));
    const uint4 w = make_uint4(v.x, v.y, v.z, v.w);
    return (BGROUP * q < width) ? w : make_uint4(BCOL_PAD, BCOL_PAD, BCOL_PAD, BCOL_PAD);
  };
  Nbr<D> ra0, ra1, rb0, rb1;
  uint4 cg = colg(0);
  uint4 cn = colg(1);
  rs.load(cg.x, i, gvol, ra0);
  rs.load(cg.y, i, gvol, ra1);
  // Issue order matters: loads complete in order, so the next column
  // group (no dependences) goes out before the record loads that need
  // the group loaded one iteration earlier.
  for (uint32_t q = 0; BGROUP * q < width; ++q) {
    uint32_t *sq = bc + (SLICE * BGROUP) * q;
    rs.load(cg.z, i, gvol, rb0);
    rs.load(cg.w, i, gvol, rb1);
    bond_term<D, CHECK>(cg.x, ra0, Xo, ui, cnot, snot, brk, sq + 0, F, A, t, broke);
    bond_term<D, CHECK>(cg.y, ra1, Xo, ui, cnot, snot, brk, sq + 1, F, A, t, broke);
    const uint4 cnn = colg(q + 2);
    rs.load(cn.x, i, gvol, ra0);
    rs.load(cn.y, i, gvol, ra1);
    bond_term<D, CHECK>(cg.z, rb0, Xo, ui, cnot, snot, brk, sq + 2, F, A, t, broke);
    bond_term<D, CHECK>(cg.w, rb1, Xo, ui, cnot, snot, brk, sq + 3, F, A, t, broke);
    cg = cn;
    cn = cnn;
  }
}

// Lane j of a node's LPN lanes (LPN = 2 or 4, neighbouring lanes of one
// quad) as a DPP source: every lane of the node's group reads lane j's x.
template <int LPN>
__device__ __forceinline__ double node_lane_d(double x, int j) {
  // quad_perm patterns: LPN 4 reads lane j of the quad; LPN 2 lane j of
  // each pair
  const long long b = __double_as_longlong(x);
  int lo, hi;
#define PG_QP(c)                                                                      \
  lo = __builtin_amdgcn_update_dpp(0, (int)b, (c), 0xF, 0xF, false);                 \
  hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), (c), 0xF, 0xF, false)
  if (LPN == 4) {
    switch (j) {
    case 0: PG_QP(0x00); break;
    case 1: PG_QP(0x55); break;
    case 2: PG_QP(0xAA); break;
    default: PG_QP(0xFF); break;
    }
  } else {
    if (j == 0) { PG_QP(0xA0); } // lanes 0,0,2,2
    else { PG_QP(0xF5); }        // lanes 1,1,3,3
  }
#undef PG_QP
  return __longlong_as_double(((long long)hi << 32) | (long long)(uint32_t)lo);
}

// The bond loop of node i shared by its LPN lanes (small workloads, where one
// lane per node would leave most of the GPU idle): lane `sub` evaluates slots
// j = sub, sub + LPN, ... of each four-slot group, and every lane of the node
// then adds the group's four terms to F in slot order, each read from the lane
// that evaluated it -- the same additions in the same order as bond_loop, so
// all LPN lanes hold bond_loop's sum.  Column words go out two groups ahead,
// partner records one group ahead.
template <int D, bool CHECK, int LPN, class RS>
__device__ __forceinline__ void bond_loop_lanes(const StepArgs &A, uint32_t t, uint32_t i,
                                                uint32_t sub, const RS &rs, const double *Xo,
                                                const double *ui, double cnot, double snot,
                                                double gvol, bool brk, double *F, bool &broke) {
  constexpr int PER = BGROUP / LPN; // slots per lane per group
  const uint32_t sl = i >> 6, lane = i & 63;
  const uint32_t base = A.slice_base[sl], width = A.slice_width[sl];
  uint32_t *bc = A.bcol + base + BGROUP * lane;
  // column words by buffer loads over the slice, as in bond_loop (a
  // wavefront's nodes lie in one slice)
  const __amdgpu_buffer_rsrc_t brs = __builtin_amdgcn_make_buffer_rsrc(
      static_cast<void *>(A.bcol + __builtin_amdgcn_readfirstlane(base)), (short)0, 0x7FFFFFFF,
      kBufWord3);
  // the column array has BCOL_TAIL spare slots after the last slice, so
  // reading two groups past a row's end stays in bounds
  auto colw = [&](uint32_t q, int p) -> uint32_t {
    const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(
        brs, ((SLICE * BGROUP) * q + BGROUP * lane + sub + LPN * p) * (uint32_t)sizeof(uint32_t),
        0, 0);
    return (BGROUP * q < width) ? w : BCOL_PAD;
  };
  uint32_t c0[PER], c1[PER], c2[PER];
  Nbr<D> r0[PER], r1[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    c0[p] = colw(0, p);
    c1[p] = colw(1, p);
  }
#pragma unroll
  for (int p = 0; p < PER; ++p) rs.load(c0[p], i, gvol, r0[p]);
  for (uint32_t q = 0; BGROUP * q < width; ++q) {
#pragma unroll
    for (int p = 0; p < PER; ++p) c2[p] = colw(q + 2, p);
#pragma unroll
    for (int p = 0; p < PER; ++p) rs.load(c1[p], i, gvol, r1[p]);
    double pr[PER][D];
    bool bk = false, bks[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      bks[p] = bond_prod<D, CHECK>(c0[p], r0[p], Xo, ui, cnot, snot, brk, pr[p]);
      bk = bk || bks[p];
    }
    // slot j of the group: lane j % LPN, its p = j / LPN
#pragma unroll
    for (int j = 0; j < BGROUP; ++j)
#pragma unroll
      for (int k = 0; k < D; ++k) F[k] = F[k] + node_lane_d<LPN>(pr[j / LPN][k], j % LPN);
    if (bk) { // rare
      uint32_t *sq = bc + (SLICE * BGROUP) * q;
#pragma unroll
      for (int p = 0; p < PER; ++p)
        if (bks[p]) bond_break(c0[p], sq + sub + LPN * p, A, t);
      broke = true;
    }
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      c0[p] = c1[p];
      c1[p] = c2[p];
      r0[p] = r1[p];
    }
  }
}

// ------------------------------------------------------------ step kernel

// One full time step t for every node (timeloop.h:608-878):
//   PMB bond force + bond breaking        particle2.h:112-152, 195-230
//   self contact, inter-grain contact     timeloop.h:709-747, contact.h:667-812
//   wall contact                          timeloop.h:753-769, contact.h:899-964
//   external force, torque                timeloop.h:774-824
//   velocity Verlet + clamped nodes       timeloop.h:834-863
//   wall motion                           timeloop.h:866-878
// followed by the position update of step t+1 (timeloop.h:661-669), the
// grain bounding boxes of step t+1 (get_minmax, contact.h:1023-1062) and
// the Verlet-skin check.  Reads u^t from u[t&1] and v^{t-1} from
// vel[(t+1)&1]; writes u^{t+1} to u[(t+1)&1] and v^t to vel[t&1].
// Where the bond loop reads partner records from: the global record array
// of step t, or a workgroup's window of it staged in LDS.
// The partner data of a column word: one 16-byte load from each record part
// (3 in 3D, 2 in 2D), all at the byte offset 16*j (node indices stay below
// 2^28, so the offset fits 32 bits).  The loads are buffer loads, one
// descriptor per part with a uniform base and the part's 16*ns bytes as its
// range (an index out of range would read zeros instead of faulting); on the
// 3D compression workload they ran 0.8% faster than global loads from scalar
// bases.  Broken and padding slots load the ghost record (a point far from
// every node, so their discarded term stays finite), so the bond loop has no
// branches around its loads.  UVOL: the grain's nodes share one volume,
// passed in gvol.
template <int D, bool UVOL> struct GlobalRecs {
  __amdgpu_buffer_rsrc_t R[D]; // the record parts of step t
  const double *V;
  uint32_t ghost;
  __device__ __forceinline__ GlobalRecs(const double *U, uint32_t ns, const double *V_,
                                        uint32_t ghost_)
      : V(V_), ghost(ghost_) {
    const uint64_t bytes = (uint64_t)ns * 16u;
    const int range = (int)(uint32_t)(bytes < 0xFFFFFFFFull ? bytes : 0xFFFFFFFFull);
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const uint64_t a = reinterpret_cast<uint64_t>(U + (size_t)q * ns * 2);
      const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
      const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
      R[q] = __builtin_amdgcn_make_buffer_rsrc(
          reinterpret_cast<void *>(((uint64_t)hi << 32) | lo), (short)0, range, kBufWord3);
    }
  }
  __device__ __forceinline__ dvec2 part(int q, uint32_t off) const {
    return __builtin_bit_cast(dvec2, __builtin_amdgcn_raw_buffer_load_b128(R[q], off, 0, 0));
  }
  __device__ __forceinline__ void load(uint32_t cw, uint32_t, double gvol, Nbr<D> &nb) const {
    const uint32_t j = (cw & BCOL_BROKEN) ? ghost : cw;
    const uint32_t off = j * 16u;
    const dvec2 a = part(0, off);
    const dvec2 b = part(1, off);
    if constexpr (D == 3) {
      const dvec2 c = part(2, off);
      nb.X[0] = a.x;
      nb.X[1] = a.y;
      nb.X[2] = b.x;
      nb.u[0] = b.y;
      nb.u[1] = c.x;
      nb.u[2] = c.y;
    } else {
      nb.X[0] = a.x;
      nb.X[1] = a.y;
      nb.u[0] = b.x;
      nb.u[1] = b.y;
    }
    nb.V = UVOL ? gvol : V[j];
  }
};
template <int D, bool UVOL> struct LdsRecs {
  const double *W; // LDS: records of nodes lo .. lo+cnt-1, then the ghost
  uint32_t lo, cnt, ghost;
  const double *V;
  // Broken and padding slots read the ghost record (after the window)
  __device__ __forceinline__ void load(uint32_t cw, uint32_t, double gvol, Nbr<D> &nb) const {
    const uint32_t off = (cw & BCOL_BROKEN) ? cnt : cw - lo;
    const double2 *p = reinterpret_cast<const double2 *>(W + off * (2 * D));
    if constexpr (D == 3) {
      const double2 a = p[0], b = p[1], c = p[2];
      nb.X[0] = a.x;
      nb.X[1] = a.y;
      nb.X[2] = b.x;
      nb.u[0] = b.y;
      nb.u[1] = c.x;
      nb.u[2] = c.y;
    } else {
      const double2 a = p[0], b = p[1];
      nb.X[0] = a.x;
      nb.X[1] = a.y;
      nb.u[0] = b.x;
      nb.u[1] = b.y;
    }
    nb.V = UVOL ? gvol : V[(cw & BCOL_BROKEN) ? ghost : cw];
  }
};

// Per-launch housekeeping, spread over all threads of the grid: clear the box
// buffer of step t+2 and move the walls.
template <int D>
__device__ __forceinline__ void step_housekeeping(const StepArgs &A, uint32_t t, uint32_t gi,
                                                  uint32_t nthreads) {
  unsigned long long *bz = A.bbox[(t + 2) % 3];
  const uint32_t nz = A.n_grains * 2 * D;
  for (uint32_t q = gi; q < nz; q += nthreads) bz[q] = ((q % (2 * D)) < D) ? ~0ull : 0ull;
  if (gi == 0) A.unice[(t + 2) % 3] = 0; // step t+1 sets it if some u^{t+2} is not nice
  if (gi < A.n_insts * 6) {
    const InstDev &ip = A.insts[gi / 6];
    A.wallpos[(t + 1) & 1][gi] = A.wallpos[t & 1][gi] + ip.speed[gi % 6] * ip.dt;
  }
}

// A node's own step inputs that depend on its index alone: reference
// position and displacement (record of step t), previous velocity,
// acceleration, external force, position at the list build, flags, contact
// key and its contact-list range.  None of them is written during the step
// before its use (brk_new[i] only by node i's own bond loop, when `broke`
// already says the same).
template <int D> struct NodeIn {
  double X[D], u[D], v[D], a[D], ex[D], xb[D];
  uint32_t key, cl0, cl1, nfast;
  uint8_t info, brk_new;
  __device__ __forceinline__ void load(const StepArgs &A, uint32_t t, uint32_t i) {
    rec_X<D>(A, i, X);
    rec_u<D>(A.rec[t & 1], A.ns, i, u);
    load_vec<D>(A.vel[(t + 1) & 1], A.ns, i, v);
    load_vec<D>(A.acc, A.ns, i, a);
    load_vec<D>(A.ext, A.ns, i, ex);
    load_vec<D>(A.xb, A.ns, i, xb);
    info = A.node_info[i];
    key = A.ckey[i];
    cl0 = A.cl_ptr[i];
    cl1 = A.cl_ptr[i + 1];
    nfast = A.cl_nfast[i];
    brk_new = A.brk_new[i];
  }
};

// The step of node i (all lanes of a wavefront call it together: the grain
// boxes are combined across the wavefront).
template <int D, bool RECORD, bool UVOL, int LPN = 1, class RS>
__device__ __forceinline__ void step_node(const StepArgs &A, const uint32_t t, const uint32_t i,
                                          const RS &rs, BlockBox *bb = nullptr,
                                          uint32_t sub = 0) {
  const uint32_t ns = A.ns;
  const bool valid = i < A.n_nodes;
  uint32_t g = 0xFFFFFFFFu;
  double xn[D];
#pragma unroll
  for (int k = 0; k < D; ++k) xn[k] = 0.0;
  bool viol = false;

  // With several lanes per node (small workloads) a step lasts as long as
  // one wavefront's chain of dependent loads, so there the node's inputs
  // that depend on i alone are requested together, up front.  With one lane
  // per node they are loaded where they are used, which leaves the
  // registers to the bond loop.
  constexpr bool EARLY = LPN > 1;
  if (valid) {
    g = A.node_grain[i];
    NodeIn<D> early;
    if constexpr (EARLY) early.load(A, t, i);
    const double *U = A.rec[t & 1];
    double F[D];
#pragma unroll
    for (int k = 0; k < D; ++k) F[k] = 0.0;

    bool broke = false; // a bond of node i broke in this step
    // ---- (a) PMB bond force with bond breaking.  Only what the bond loop
    // needs is live here; everything else is loaded after it.  The loop is
    // software-pipelined: the partner data of the next two bonds and the
    // column words of the two after those are in flight while two bonds are
    // evaluated.
    {
      const GrainDev *gq = A.grains + g;
      const int32_t gflags = gq->flags;
      const InstDev &iq = A.insts[gq->inst];
      const bool mov_a = (gflags & GF_MOVABLE) ||
                         (iq.set_movable_grain == (int32_t)g && t >= iq.set_movable_step);
      if (mov_a) {
        double Xo[D], ui[D];
        if constexpr (EARLY) {
#pragma unroll
          for (int k = 0; k < D; ++k) {
            Xo[k] = early.X[k];
            ui[k] = early.u[k];
          }
        } else {
          rec_X<D>(A, i, Xo);
          rec_u<D>(U, A.ns, i, ui);
        }
        const double cnot = gq->cnot, snot = gq->snot;
        const double gvol = UVOL ? gq->vol : 0.0;
        const bool brk = (iq.flags & IF_FRACTURE) && (gflags & GF_BREAKABLE);
        if constexpr (LPN == 1) {
          if (A.unice[t % 3] == 0 && A.xi_ok) // the same for every lane of the launch
            bond_loop<D, false>(A, t, i, rs, Xo, ui, cnot, snot, gvol, brk, F, broke);
          else
            bond_loop<D, true>(A, t, i, rs, Xo, ui, cnot, snot, gvol, brk, F, broke);
        } else {
          if (A.unice[t % 3] == 0 && A.xi_ok)
            bond_loop_lanes<D, false, LPN>(A, t, i, sub, rs, Xo, ui, cnot, snot, gvol, brk, F,
                                           broke);
          else
            bond_loop_lanes<D, true, LPN>(A, t, i, sub, rs, Xo, ui, cnot, snot, gvol, brk, F,
                                          broke);
          // a break in any of the node's lanes (the node's lanes are all
          // active here: they share its conditions)
          const unsigned long long m = __ballot(broke);
          const uint32_t l0 = (threadIdx.x & 63) & ~(uint32_t)(LPN - 1);
          broke = ((m >> l0) & ((1ull << LPN) - 1)) != 0;
        }
        if (broke) A.brk_new[i] = 1;
      }
    }

    // ---- everything after the bond sum
    const GrainDev gp = A.grains[g];
    const InstDev &ip = A.insts[gp.inst];
    double Xc[D], ui[D], xc[D];
    if constexpr (EARLY) {
#pragma unroll
      for (int k = 0; k < D; ++k) ui[k] = early.u[k];
    } else {
      rec_u<D>(U, A.ns, i, ui);
    }
    pos_X<D>(A, ip, g, i, Xc);
#pragma unroll
    for (int k = 0; k < D; ++k) xc[k] = Xc[k] + ui[k]; // CurrPos = pos + disp
    const uint8_t info = EARLY ? early.info : A.node_info[i];
    const bool mov = grain_movable(gp, ip, g, t);
    const double *vprev = A.vel[(t + 1) & 1];
    double *vcur = A.vel[t & 1];
    double vi[D], ai[D];
    if constexpr (EARLY) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        vi[k] = early.v[k];
        ai[k] = early.a[k];
      }
    } else {
      load_vec<D>(vprev, ns, i, vi);
      load_vec<D>(A.acc, ns, i, ai);
    }
    const uint32_t key = EARLY ? early.key : A.ckey[i];
    uint32_t wmask = 0; // bit s: the node touched wall side s
    double wt[6][4];
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
      for (int q = 0; q < 4; ++q) wt[s][q] = 0.0;

    if (mov) {
      const bool brk = (ip.flags & IF_FRACTURE) && (gp.flags & GF_BREAKABLE);

      // ---- (b) contact from the sorted Verlet list
      const bool selfc = brk && (ip.flags & IF_SELF_CONTACT);
      const bool stop = (ip.flags & IF_UNGATED) || grain_stoppable(gp, ip, g, t);
      double boxg[2 * D];
      load_box<D>(A.bbox[t % 3], g, boxg);
      if ((info & NI_CONTACT) && (selfc || stop)) {
        // the fast list while no bond of the node has broken since the list
        // build (its other entries are intact bonds, which stay excluded)
        const uint32_t c0 = EARLY ? early.cl0 : A.cl_ptr[i];
        const bool full = broke || (EARLY ? early.brk_new : A.brk_new[i]);
        const uint32_t c1 = full ? (EARLY ? early.cl1 : A.cl_ptr[i + 1])
                                 : c0 + (EARLY ? early.nfast : A.cl_nfast[i]);
        // the next fast-list entry is loaded while the current one is
        // evaluated, and a partner's data is loaded before the gates decide
        // whether it is needed (one round trip to memory per entry)
        if constexpr (LPN > 1) {
          // the node's lanes take consecutive entries; each lane's three
          // terms are then added in entry order on every lane.  A bond
          // broken in this step's bond loop by another of the node's lanes
          // is visible here: its store is complete before these loads.
          if (broke) __builtin_amdgcn_s_waitcnt(0);
          for (uint32_t cb = c0; cb < c1; cb += LPN) {
            const uint32_t c = cb + sub;
            double fc[D], fd[D], ff[D];
#pragma unroll
            for (int k = 0; k < D; ++k) fc[k] = fd[k] = ff[k] = 0.0;
            if (c < c1) {
              uint32_t j, gj;
              int32_t bs;
              if (full) {
                j = A.cl_j[c];
                gj = A.node_grain[j];
                bs = A.cl_slot[c];
              } else {
                const uint2 en = A.cl_fast[c];
                j = en.x;
                gj = en.y;
                bs = -1;
              }
              double Xj[D], uj[D], vj[D];
              pos_X<D>(A, ip, gj, j, Xj);
              rec_u<D>(U, A.ns, j, uj);
              load_vec<D>(vprev, ns, j, vj);
              const double Vj = A.V[j];
              double d[D], r, R, K;
              if (contact_gate<D>(A, t, gj, bs, g, gp, ip, Xc, Xj, selfc, stop, boxg, R, K) &&
                  pair_within<D>(Xj, uj, xc, R, d, r)) {
                double w[D];
#pragma unroll
                for (int k = 0; k < D; ++k) w[k] = vi[k] - vj[k];
                pair_terms<D>(d, r, R, K, Vj, w, gp.rho, ip.mu, ip, fc, fd, ff);
              }
            }
#pragma unroll
            for (int jl = 0; jl < LPN; ++jl)
#pragma unroll
              for (int k = 0; k < D; ++k) {
                F[k] = F[k] + node_lane_d<LPN>(fc[k], jl);
                F[k] = F[k] + node_lane_d<LPN>(fd[k], jl);
                F[k] = F[k] + node_lane_d<LPN>(ff[k], jl);
              }
          }
        } else {
        uint2 nxt = (!full && c0 < c1) ? A.cl_fast[c0] : make_uint2(0u, 0u);
        for (uint32_t c = c0; c < c1; ++c) {
          uint32_t j, gj;
          int32_t bs;
          if (full) {
            j = A.cl_j[c];
            gj = A.node_grain[j];
            bs = A.cl_slot[c];
          } else {
            j = nxt.x;
            gj = nxt.y;
            bs = -1;
            if (c + 1 < c1) nxt = A.cl_fast[c + 1];
          }
          double Xj[D], uj[D], vj[D];
          pos_X<D>(A, ip, gj, j, Xj);
          rec_u<D>(U, A.ns, j, uj);
          load_vec<D>(vprev, ns, j, vj);
          const double Vj = A.V[j];
          double d[D], r, R, K;
          if (!contact_gate<D>(A, t, gj, bs, g, gp, ip, Xc, Xj, selfc, stop, boxg, R, K) ||
              !pair_within<D>(Xj, uj, xc, R, d, r))
            continue;
          double w[D];
#pragma unroll
          for (int k = 0; k < D; ++k) w[k] = vi[k] - vj[k];
          pair_force<D, double>(d, r, R, K, Vj, w, gp.rho, ip.mu, ip, F);
        }
        }
      }

      // ---- (c) walls
      if (stop && (ip.flags & IF_WALL) && (info & NI_CONTACT)) {
        const double *wp = A.wallpos[t & 1] + (size_t)gp.inst * 6;
        if (!within_interior<D>(boxg, wp, ip.Rc)) {
          const double vol = A.V[i];
          if (D == 2) {
            wmask |= wall_side<D, double>(xc[1] - wp[3], 1, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[3], vi, F, wt[3]) ? (1u << 3) : 0u;
            wmask |= wall_side<D, double>(xc[0] - wp[0], 0, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[0], vi, F, wt[0]) ? (1u << 0) : 0u;
            wmask |= wall_side<D, double>(xc[0] - wp[1], 0, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[1], vi, F, wt[1]) ? (1u << 1) : 0u;
            wmask |= wall_side<D, double>(xc[1] - wp[2], 1, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[2], vi, F, wt[2]) ? (1u << 2) : 0u;
          } else {
            wmask |= wall_side<D, double>(xc[0] - wp[0], 0, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[0], vi, F, wt[0]) ? (1u << 0) : 0u;
            wmask |= wall_side<D, double>(xc[0] - wp[3], 0, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[3], vi, F, wt[3]) ? (1u << 3) : 0u;
            wmask |= wall_side<D, double>(xc[1] - wp[1], 1, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[1], vi, F, wt[1]) ? (1u << 1) : 0u;
            wmask |= wall_side<D, double>(xc[1] - wp[4], 1, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[4], vi, F, wt[4]) ? (1u << 4) : 0u;
            wmask |= wall_side<D, double>(xc[2] - wp[2], 2, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[2], vi, F, wt[2]) ? (1u << 2) : 0u;
            wmask |= wall_side<D, double>(xc[2] - wp[5], 2, ip, ip.Kn, ip.mu, gp.rho, vol, ip.speed[5], vi, F, wt[5]) ? (1u << 5) : 0u;
          }
        }
      }

      // ---- external force (timeloop.h:776-786) and torque (:790-824)
      double ex[D];
      if constexpr (EARLY) {
#pragma unroll
        for (int k = 0; k < D; ++k) ex[k] = early.ex[k];
      } else {
        load_vec<D>(A.ext, ns, i, ex);
      }
      if (ip.flags & IF_WORM) {
        // timeloop.h:1183: extforce times (t_num + 1) % 2, i.e. only in the
        // phases with the front nodes clamped (adding 0*extforce leaves F)
        if (((t / 1000u) % 3u) == 0u) {
#pragma unroll
          for (int k = 0; k < D; ++k) F[k] += ex[k];
        }
      } else if (ip.flags & IF_GRAD_EXT) {
        double sc = (double)t / (double)ip.extforce_maxstep;
        if (sc > 1) sc = 1;
#pragma unroll
        for (int k = 0; k < D; ++k) F[k] += sc * ex[k];
      } else {
#pragma unroll
        for (int k = 0; k < D; ++k) F[k] += ex[k];
      }
      if (ip.flags & IF_TORQUE) {
        const double *cm = A.gmean + (size_t)g * 3;
        const uint32_t ta = gp.torque_axis;
        double rv[3] = {0, 0, 0}, rp[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < D; ++k) rv[k] = xc[k] - cm[k];
        if (D == 3) rv[ta] = 0;
        rp[(ta + 1) % 3] = -rv[(ta + 2) % 3];
        rp[(ta + 2) % 3] = rv[(ta + 1) % 3];
        const double pn = sqrt(red<D>(rp, rp));
        if (pn > 0) {
#pragma unroll
          for (int k = 0; k < D; ++k) rp[k] = rp[k] / pn;
        } else {
          rp[(ta + 1) % 3] = 0;
          rp[(ta + 2) % 3] = 0;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) F[k] += (gp.torque_val * rp[k]);
      }

      // ---- (d) velocity Verlet with clamped nodes (timeloop.h:834-863)
      double an[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        an[k] = (1 / gp.rho) * F[k];
        double tmp = ai[k] + an[k];
        tmp *= ip.half_dt;
        vi[k] = vi[k] + tmp;
        ai[k] = an[k];
      }
      if (info & NI_CLAMPED) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
          ui[k] = 0.0;
          vi[k] = 0.0;
          ai[k] = 0.0;
        }
      }
      if (ip.flags & IF_WORM) {
        // timeloop.h:1106-1116, 1215-1245: in phases of 1000 steps the front
        // nodes (0, 3) are held in every third phase, the back nodes (1, 2)
        // in the others; nodes 0-3 never move in y
        const bool front = ((t / 1000u) % 3u) == 0u;
        if ((front && (info & NI_WORM_FRONT)) || (!front && (info & NI_WORM_BACK))) {
#pragma unroll
          for (int k = 0; k < D; ++k) {
            vi[k] = 0.0;
            ai[k] = 0.0;
          }
        }
        if (info & (NI_WORM_FRONT | NI_WORM_BACK)) {
          vi[1] = 0.0;
          ai[1] = 0.0;
        }
      }
      store_vec<D>(A.acc, ns, i, ai);
    }
    store_vec<D>(vcur, ns, i, vi);

    // ---- position update of step t+1 (timeloop.h:661-669)
    double un[D];
    if ((ip.flags & IF_UNGATED) || grain_movable(gp, ip, g, t + 1)) { // worm, compact: every grain
#pragma unroll
      for (int k = 0; k < D; ++k) un[k] = ui[k] + ((ip.dt * vi[k]) + (ip.h * ai[k]));
    } else {
#pragma unroll
      for (int k = 0; k < D; ++k) un[k] = ui[k];
    }
    rec_set_u<D>(A.rec[(t + 1) & 1], A.ns, i, un);
    bool unice = true;
#pragma unroll
    for (int k = 0; k < D; ++k) unice = unice && nice(un[k]);
    if (!unice) atomicOr(&A.unice[(t + 1) % 3], 1u); // rare: the next step checks each bond
#pragma unroll
    for (int k = 0; k < D; ++k) xn[k] = Xc[k] + un[k];
    if (info & NI_CONTACT) {
      double dd[D];
#pragma unroll
      for (int k = 0; k < D; ++k) dd[k] = xn[k] - (EARLY ? early.xb[k] : A.xb[(size_t)k * ns + i]);
      viol = red<D>(dd, dd) > ip.skin_half2;
    }
    if (RECORD) {
      store_vec<D>(A.force_out, ns, i, F);
      if (key != 0xFFFFFFFFu) {
        A.wflag[key] = (uint8_t)wmask;
        const int nsd = (D == 2) ? 4 : 6;
        for (int s = 0; s < nsd; ++s)
          if ((wmask >> s) & 1u)
            for (int q = 0; q < 4; ++q) A.wterm[((size_t)key * 6 + s) * 4 + q] = wt[s][q];
      }
    }
  }

  if (viol) atomicMin(A.violation, t + 1);
  if (bb) wave_box_block<D>(A.bbox[(t + 1) % 3], g, valid, xn, *bb, blockDim.x / 64);
  else wave_box_atomic<D>(A.bbox[(t + 1) % 3], g, valid, xn);
}

// LPN lanes per node (1, 2 or 4): with more than one, the node's lanes share
// its bond loop (bond_loop_lanes) and repeat the rest of its step, which
// writes the same values and whose atomics (min, max, or) are idempotent.
template <int D, bool RECORD, bool UVOL, int LPN>
__global__ __launch_bounds__(256, LPN > 1 ? 2 : PG_STEP_WAVES_PER_EU) void k_step(
    const StepArgs A, const uint32_t t_arg) {
  const uint32_t t = step_index(A, t_arg);
  if (*A.violation <= t) return; // neighbour list stale since an earlier step
  const uint32_t nb = gridDim.x;
  const uint32_t lb = xcd_block(blockIdx.x, nb);
  const uint32_t gt = lb * blockDim.x + threadIdx.x;
  const uint32_t i = gt / LPN;
  __shared__ BlockBox bb;
  block_box_init<D>(bb, A, lb * blockDim.x / LPN);
  __syncthreads();
  step_housekeeping<D>(A, t, gt, nb * blockDim.x);
  step_node<D, RECORD, UVOL, LPN>(A, t, i,
                                  GlobalRecs<D, UVOL>(A.rec[t & 1], A.ns, A.V, A.ghost), &bb,
                                  gt % LPN);
}

// The same step with each workgroup's partner window staged in LDS: block b
// owns A.blk_slices consecutive 64-node slices, and every bond partner of its
// nodes lies in [win_lo[b], win_lo[b] + win_cnt[b]), which it copies from the
// record array of step t with coalesced loads before the bond loops read it.
template <int D, bool RECORD, bool UVOL>
__global__ __launch_bounds__(1024) void k_step_lds(const StepArgs A, const uint32_t t_arg) {
  extern __shared__ double lds_rec[];
  const uint32_t t = step_index(A, t_arg);
  if (*A.violation <= t) return; // uniform: no launch writes a value <= t
  const uint32_t nb = gridDim.x;
  const uint32_t lb = xcd_block(blockIdx.x, nb);
  step_housekeeping<D>(A, t, lb * blockDim.x + threadIdx.x, nb * blockDim.x);
  const uint32_t lo = A.win_lo[lb], cnt = A.win_cnt[lb];
  {
    // LDS holds whole records [cnt + 1][D] double2 (the ghost last), copied
    // part by part from the record parts of nodes lo .. lo + cnt - 1
    const double2 *src = reinterpret_cast<const double2 *>(A.rec[t & 1]);
    double2 *dst = reinterpret_cast<double2 *>(lds_rec);
    const uint32_t n2 = cnt * D;
    for (uint32_t q = threadIdx.x; q < n2; q += blockDim.x) {
      const uint32_t part = q / cnt, r = q - part * cnt;
      dst[r * D + part] = src[(size_t)part * A.ns + lo + r];
    }
    if (threadIdx.x < D) // the ghost record after the window
      dst[n2 + threadIdx.x] = src[(size_t)threadIdx.x * A.ns + A.ghost];
  }
  __syncthreads();
  const uint32_t S = A.blk_slices;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  const LdsRecs<D, UVOL> rs{lds_rec, lo, cnt, A.ghost, A.V};
  for (uint32_t sl = wave; sl < S; sl += nw)
    step_node<D, RECORD, UVOL>(A, t, (lb * S + sl) * 64 + lane, rs);
}

// Position update of step 1 from the initial state (timeloop.h:661-669),
// plus the grain boxes of step 1.
template <int D>
__global__ __launch_bounds__(256) void k_prologue(const StepArgs A) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < A.n_nodes;
  const uint32_t ns = A.ns;
  uint32_t g = 0xFFFFFFFFu;
  double xn[D];
#pragma unroll
  for (int k = 0; k < D; ++k) xn[k] = 0.0;
  if (valid) {
    g = A.node_grain[i];
    const GrainDev gp = A.grains[g];
    const InstDev &ip = A.insts[gp.inst];
    double Xc[D], un[D], u0[D];
    pos_X<D>(A, ip, g, i, Xc);
    rec_u<D>(A.rec[0], A.ns, i, u0);
    const bool mov = (ip.flags & IF_UNGATED) || grain_movable(gp, ip, g, 1);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double v = A.vel[0][(size_t)k * ns + i];
      const double a = A.acc[(size_t)k * ns + i];
      un[k] = mov ? u0[k] + ((ip.dt * v) + (ip.h * a)) : u0[k];
      xn[k] = Xc[k] + un[k];
    }
    rec_set_u<D>(A.rec[1], A.ns, i, un);
  }
  wave_box_atomic<D>(A.bbox[1], g, valid, xn);
}

// Grain boxes of step t (get_minmax, contact.h:1023-1062) from the positions
// of step t, into a cleared box buffer: after reset_partzero_y moves a grain
// between the step that computed them and the step that uses them.
template <int D>
__global__ __launch_bounds__(256) void k_boxes(const StepArgs A, const uint32_t t,
                                               unsigned long long *keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < A.n_nodes;
  uint32_t g = 0xFFFFFFFFu;
  double x[D];
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = 0.0;
  if (valid) {
    g = A.node_grain[i];
    const InstDev &ip = A.insts[A.grains[g].inst];
    double Xc[D], u[D];
    pos_X<D>(A, ip, g, i, Xc);
    rec_u<D>(A.rec[t & 1], A.ns, i, u);
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = Xc[k] + u[k];
  }
  wave_box_atomic<D>(keys, g, valid, x);
}

// Per-grain post-processing of the completed step t (the quantities of
// read_sim_output.py extract_bulk and read_force_val): one block per grain,
// each thread summing a strided share of the grain's nodes, then a block
// reduction.  out[g - g0][5D+1] as documented at pg_grain_stats.
template <int D>
__global__ __launch_bounds__(256) void k_grain_stats(const StepArgs A, const uint32_t t,
                                                     const uint32_t g0, double *out) {
  constexpr int NV = 5 * D + 1;
  __shared__ double red_s[NV][256];
  const uint32_t g = g0 + blockIdx.x;
  const GrainDev gp = A.grains[g];
  const InstDev &ip = A.insts[gp.inst];
  const uint32_t ns = A.ns;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (uint32_t e = gp.node_begin + threadIdx.x; e < gp.node_end; e += blockDim.x) {
    const uint32_t n = A.ext2int[e];
    double X[D], u[D];
    pos_X<D>(A, ip, g, n, X);
    rec_u<D>(A.rec[t & 1], ns, n, u);
    const double vr = A.V[n] * gp.rho;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double a = A.acc[(size_t)k * ns + n];
      acc[k] += A.force_out[(size_t)k * ns + n];
      acc[D + k] += A.vel[t & 1][(size_t)k * ns + n];
      acc[2 * D + k] += a;
      acc[3 * D + k] += X[k] + u[k];
      acc[4 * D + 1 + k] += vr * a;
    }
    // bonds at creation: the row's non-padding slots; intact: those without
    // the broken flag
    const uint32_t sl = n >> 6, lane = n & 63;
    const uint32_t base = A.slice_base[sl], width = A.slice_width[sl];
    uint32_t orig = 0, now = 0;
    for (uint32_t k = 0; k < width; ++k) {
      const uint32_t w = A.bcol[bslot(base, lane, k)];
      if (w == BCOL_PAD) break;
      ++orig;
      now += (w & BCOL_BROKEN) ? 0u : 1u;
    }
    acc[4 * D] += (double)(orig - now) / (double)orig;
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) red_s[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < NV; ++k) red_s[k][threadIdx.x] += red_s[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nn = (double)(gp.node_end - gp.node_begin);
    double *o = out + (size_t)blockIdx.x * NV;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const bool mean = (k >= D && k < 4 * D + 1);
      o[k] = mean ? red_s[k][0] / nn : red_s[k][0];
    }
  }
}

// mean_CurrPos of every grain (particle2.h:89-98), summed sequentially in the
// reference node order.
template <int D>
__global__ void k_grain_mean(const StepArgs A, const uint32_t t_arg, double *gmean) {
  const uint32_t t = step_index(A, t_arg);
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.n_grains) return;
  const GrainDev gp = A.grains[g];
  const InstDev &ipg = A.insts[gp.inst];
  const double *U = A.rec[t & 1];
  double m[D];
#pragma unroll
  for (int k = 0; k < D; ++k) m[k] = 0.0;
  for (uint32_t e = gp.node_begin; e < gp.node_end; ++e) {
    const uint32_t n = A.ext2int[e];
    double Xc[D], un[D];
    pos_X<D>(A, ipg, g, n, Xc);
    rec_u<D>(U, A.ns, n, un);
#pragma unroll
    for (int k = 0; k < D; ++k) m[k] += (Xc[k] + un[k]);
  }
  const uint32_t nn = gp.node_end - gp.node_begin;
#pragma unroll
  for (int k = 0; k < D; ++k) gmean[(size_t)g * 3 + k] = m[k] / (double)nn;
}

// Wall reaction of recorded steps, summed in the reference's sequential
// order (grain, node in contact order, then side; contact.h:339,374,
// 411-412,428 via timeloop.h:753-769), in two launches.
//
// k_wall_compact: the keys whose node touched a wall, compacted in key
// order, in parallel over chunks of REACT_CHUNK keys: chunk b of recorded
// step y writes its hits, ascending, to hits[y][b*REACT_CHUNK ...] -- each
// the key's offset in the chunk with the side mask in bits 16-23 -- and
// their number to chunk_cnt[y][b].  One block per (chunk, step).
constexpr uint32_t REACT_CHUNK = 4096;           // keys per chunk: 16 per thread
static_assert(REACT_CHUNK == (uint32_t)kWflagSpare, "flag loads run one chunk past the keys");

__global__ __launch_bounds__(256) void k_wall_compact(ReactionArgs R) {
  const uint32_t slot = (R.first + blockIdx.y) % R.slots;
  const uint8_t *wflag = R.wflag + slot * R.fstride;
  uint32_t *hits = R.hits + blockIdx.y * R.hstride + (size_t)blockIdx.x * REACT_CHUNK;
  __shared__ uint32_t wtot[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t k0 = blockIdx.x * REACT_CHUNK + tid * 16;
  // 16 flags in one load (the flag array has a chunk of zero bytes past the
  // last key, and ring slots start on 16-byte boundaries)
  const uint4 f4 = *reinterpret_cast<const uint4 *>(wflag + k0);
  const uint32_t w[4] = {f4.x, f4.y, f4.z, f4.w};
  uint32_t cnt = 0;
#pragma unroll
  for (int q = 0; q < 16; ++q) cnt += ((w[q >> 2] >> (8 * (q & 3))) & 0xFFu) ? 1u : 0u;
  uint32_t v = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(v, off, 64);
    if (lane >= (uint32_t)off) v += u;
  }
  if (lane == 63) wtot[wave] = v;
  __syncthreads();
  uint32_t pos = v - cnt, total = 0;
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q) {
    if (q < wave) pos += wtot[q];
    total += wtot[q];
  }
#pragma unroll
  for (int q = 0; q < 16; ++q)
    if (const uint32_t m = (w[q >> 2] >> (8 * (q & 3))) & 0xFFu) // key within the chunk | sides
      hits[pos++] = (tid * 16 + q) | (m << 16);
  if (tid == 0) R.chunk_cnt[blockIdx.y * R.cstride + blockIdx.x] = total;
}

// k_reaction: one block per recorded step.  The compacted hits are taken in
// pieces of T, one per thread, T being the block size: 512 for one step's
// reaction, whose wall may hold a thousand hits, 256 for a ring of recorded
// steps, one block per step with a few hits each.  The block stages a piece:
// the four terms of every side each hit touched -- normal, damping, friction
// along the side's two tangents -- go into one list per side, in key order
// (ballots give the positions).  Each (instance, side, component) is a chain
// of subtractions in key order: terms 0 and 1 of every entry along the
// side's direction, term 2 or 3 across it.  A node that did not touch a side
// adds nothing to it (in the reference that side is not updated; here it is
// not visited).  Running sums stay in LDS between pieces (the first
// REACT_LDS_SUMS of them; the rest in the output).  Each thread loads its hit
// of the next piece (key, sides, terms) before the current piece is summed,
// so those loads overlap the sums.
//
// A chain of at most REACT_SEQ_MAX terms is summed by one thread in order.
// A longer one -- one wall pressed by many nodes -- would be a long chain of
// dependent FP64 subtractions, so a wavefront evaluates it in 64 chunks at
// once (react_chain_wave), each from a guessed start value, and corrects the
// guesses until every chunk starts exactly where the one before it ends:
// then every chunk repeated the sequential subtractions from the sequential
// start value, and the last chunk's end is the sequential result, bit for
// bit.
constexpr uint32_t REACT_SEQ_MAX = 64;     // longest chain summed by one thread
constexpr uint32_t REACT_LONG_MAX = 192;   // longer chains of a piece handed to wavefronts
constexpr uint32_t REACT_LDS_SUMS = 18 * 64; // running sums held in LDS
constexpr int REACT_BATCH = 8;             // terms read ahead of a chain

template <int NS> struct ReactHit {
  uint32_t key, mask; // mask bit s: the node touched side s
  double2 t[NS][2];   // terms (0, 1), (2, 3) of each side it touched
};

// Hit e of the chunk group that starts at chunk g0 (s_off: the group's chunk
// offsets in the list of hits, s_off[ngroup] = total): the last chunk whose
// offset is <= e holds it.
template <int NS>
__device__ __forceinline__ void react_fetch(const uint32_t *hits, const double *wterm,
                                            const uint32_t *s_off, uint32_t ngroup, uint32_t g0,
                                            uint32_t total, uint32_t e, ReactHit<NS> &h) {
  h.key = 0;
  h.mask = 0;
  if (e >= total) return;
  uint32_t lo = 0, hi = ngroup;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s_off[mid] <= e) lo = mid;
    else hi = mid;
  }
  const uint32_t w = hits[(size_t)(g0 + lo) * REACT_CHUNK + (e - s_off[lo])];
  h.key = (g0 + lo) * REACT_CHUNK + (w & 0xFFFFu);
  h.mask = w >> 16;
  const double2 *src = reinterpret_cast<const double2 *>(wterm + (size_t)h.key * 24);
#pragma unroll
  for (int sd = 0; sd < NS; ++sd)
    if ((h.mask >> sd) & 1u) {
      h.t[sd][0] = src[2 * sd];
      h.t[sd][1] = src[2 * sd + 1];
    }
}

// Cross-lane moves of a double by DPP (two 32-bit halves): lanes whose
// source lies outside the pattern, or whose row ROW_MASK leaves out, get +0.0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(uint32_t)lo);
}
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114,
              DPP_ROW_SHR8 = 0x118, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143,
              DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
// lane l gets lane l+1's x (lane 63: +0.0), and lane l-1's (lane 0: +0.0)
__device__ __forceinline__ double lane_next(double x) { return dpp_d<DPP_WAVE_SHL1, 0xF>(x); }
__device__ __forceinline__ double lane_prev(double x) { return dpp_d<DPP_WAVE_SHR1, 0xF>(x); }

// Exclusive prefix sum over the wavefront's lanes, in tree order (all lanes
// active).  Used only for guesses: the results are checked, not trusted.
__device__ __forceinline__ double wave_excl_sum(double v) {
  double x = v;
  x = x + dpp_d<DPP_ROW_SHR1, 0xF>(x); // sums within rows of 16 lanes
  x = x + dpp_d<DPP_ROW_SHR2, 0xF>(x);
  x = x + dpp_d<DPP_ROW_SHR4, 0xF>(x);
  x = x + dpp_d<DPP_ROW_SHR8, 0xF>(x);
  x = x + dpp_d<DPP_ROW_BCAST15, 0xA>(x); // row 0's total into row 1, row 2's into row 3
  x = x + dpp_d<DPP_ROW_BCAST31, 0xC>(x); // lane 31's total into rows 2 and 3
  return lane_prev(x);
}
__device__ __forceinline__ double read_lane_d(double x, uint32_t l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, (int)l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), (int)l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(uint32_t)lo);
}

// The terms of one chain are base[idx(m)], m = 0 .. n-1: entry lo + m/2,
// terms 0 and 1, when the component is the side's direction (two); entry
// lo + m, term o1, otherwise.  Indices past `end` read the all-zero entry
// (x - (+0.0) == x for every x, zeros and NaNs included), so reads run in
// batches with no select in the chain.
struct ChainTerms {
  const double *base;
  uint32_t lo, o1, zero;
  bool two;
  __device__ __forceinline__ uint32_t idx(uint32_t m, uint32_t end) const {
    if (m >= end) return zero * 4;
    return two ? (lo + (m >> 1)) * 4 + (m & 1u) : (lo + m) * 4 + o1;
  }
  // x minus terms m0 .. m1-1, in order
  __device__ __forceinline__ double run(double x, uint32_t m0, uint32_t m1) const {
    for (uint32_t m = m0; m < m1; m += REACT_BATCH) {
      double v[REACT_BATCH];
#pragma unroll
      for (int k = 0; k < REACT_BATCH; ++k) v[k] = base[idx(m + (uint32_t)k, m1)];
#pragma unroll
      for (int k = 0; k < REACT_BATCH; ++k) x = x - v[k];
    }
    return x;
  }
};

// s0 minus the n terms of a chain in order, by one wavefront (all lanes call
// it; every lane returns the result).  Lane c takes chunk c, terms
// [c*L, (c+1)*L); chunks 0 .. C-1 are not empty.  Each round evaluates every
// chunk from its guess g_c; g_0 = s0 always.  If every g_{c+1} equals chunk
// c's end bit for bit, all chunks started from the sequential values (by
// induction from g_0) and chunk C-1's end is the answer.  Otherwise, with cs
// the first chunk whose end differs from the next guess, chunks 0 .. cs
// started exactly, so g_{cs+1} becomes chunk cs's end -- one more exact start
// per round, at most C + 1 rounds -- and the later guesses move by the
// differences summed from cs on, which makes them exact as well whenever a
// chunk's end shifts by its start's shift (the usual case: no rounding to a
// coarser binade inside the chunk).  The first guesses are s0 minus the
// chunk sums of the chunks before.  Chunks of up to REACT_REG_TERMS terms
// (every chain of a reaction piece) stay in registers across the rounds.
constexpr uint32_t REACT_REG_TERMS = 16;
__device__ __forceinline__ double react_chain_wave(const ChainTerms &ct, uint32_t n, double s0,
                                                   uint32_t lane) {
  if (n == 0) return s0;
  const uint32_t L = (n + 63) / 64, C = (n + L - 1) / L;
  const uint32_t m0 = min(lane * L, n), m1 = min(m0 + L, n);
  const bool regs = L <= REACT_REG_TERMS; // the same on every lane
  double v[REACT_REG_TERMS];
  if (regs) {
#pragma unroll
    for (uint32_t k = 0; k < REACT_REG_TERMS; ++k) v[k] = ct.base[ct.idx(m0 + k, m1)];
  }
  auto chunk = [&](double x) -> double {
    if (regs) {
#pragma unroll
      for (uint32_t k = 0; k < REACT_REG_TERMS; ++k) x = x - v[k];
      return x;
    }
    return ct.run(x, m0, m1);
  };
  const double part = chunk(0.0);
  const double pre = wave_excl_sum(part);
  double g = (lane == 0) ? s0 : s0 + pre;
  double end = g;
  for (uint32_t round = 0; round <= C; ++round) {
    end = chunk(g);
    const double gn = lane_next(g);
    const bool bad =
        lane + 1 < C && __double_as_longlong(gn) != __double_as_longlong(end);
    const unsigned long long mb = __ballot(bad);
    if (mb == 0) break;
    const uint32_t cs = (uint32_t)__builtin_ctzll(mb);
    const double dx = wave_excl_sum((lane >= cs && lane + 1 < C) ? end - gn : 0.0);
    const double prev_end = lane_prev(end);
    if (lane == cs + 1) g = prev_end;
    else if (lane > cs + 1 && lane < C) g = g + dx;
  }
  return read_lane_d(end, C - 1);
}

template <int D, uint32_t T>
__global__ __launch_bounds__(T) void k_reaction(ReactionArgs R) {
  const uint32_t slot = (R.first + blockIdx.y) % R.slots;
  const double *wterm = R.wterm + slot * R.wstride;
  const uint32_t *hits = R.hits + blockIdx.y * R.hstride;
  const uint32_t *ccnt = R.chunk_cnt + blockIdx.y * R.cstride;
  double *reaction = R.reaction + blockIdx.y * R.rstride;
  constexpr int NS = (D == 2) ? 4 : 6;
  constexpr uint32_t PER_INST = NS * D;
  constexpr uint32_t NW = T / 64;
  constexpr uint32_t ZERO = T; // index of each side list's all-zero entry
  __shared__ uint32_t s_off[T + 1];
  __shared__ uint32_t wtot[NW];
  __shared__ uint32_t s_wcnt[NW][NS]; // hits of a piece per wavefront and side
  __shared__ uint32_t s_key[NS][T];
  __shared__ double2 s_side[NS][T + 1][2]; // terms (0, 1), (2, 3) of one side
  __shared__ double s_acc[REACT_LDS_SUMS];
  __shared__ uint32_t s_long[REACT_LONG_MAX][3]; // assignment, first entry, entry end
  __shared__ uint32_t s_nlong;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t n_assign = R.n_insts * PER_INST;
  auto out_at = [&](uint32_t a) {
    return reaction + (size_t)(a / PER_INST) * 18 + ((a / D) % NS) * 3 + a % D;
  };
  // the terms of assignment a's chain among entries [lo, e1) of its side
  auto chain_of = [&](uint32_t a, uint32_t lo) {
    const uint32_t sd = (a / D) % NS, cp = a % D;
    const int dir = (D == 2) ? ((sd < 2) ? 0 : 1) : (int)(sd % 3);
    ChainTerms ct;
    ct.base = reinterpret_cast<const double *>(&s_side[sd][0][0]);
    ct.lo = lo;
    ct.two = (int)cp == dir;
    ct.o1 = ct.two ? 0u : (((int)cp == (dir + 1) % D) ? 2u : 3u);
    ct.zero = ZERO;
    return ct;
  };
  // a thread's short-chain sums are the assignments a = tid (mod
  // T): it alone clears, updates and writes them; a long chain's
  // sum is read and written by a wavefront between two barriers
  for (uint32_t a = tid; a < n_assign; a += T) {
    if (a < REACT_LDS_SUMS) s_acc[a] = 0.0;
    else *out_at(a) = 0.0;
  }
  if (tid < 2 * NS) s_side[tid >> 1][ZERO][tid & 1] = make_double2(0.0, 0.0);
  for (uint32_t g0 = 0; g0 < R.nchunks; g0 += T) {
    // offsets of this group's chunks in the key-ordered list of hits
    const uint32_t c = (g0 + tid < R.nchunks) ? ccnt[g0 + tid] : 0u;
    uint32_t v = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t u = __shfl_up(v, off, 64);
      if (lane >= (uint32_t)off) v += u;
    }
    if (lane == 63) wtot[wave] = v;
    __syncthreads();
    uint32_t excl = v - c, total = 0;
#pragma unroll
    for (uint32_t q = 0; q < NW; ++q) {
      if (q < wave) excl += wtot[q];
      total += wtot[q];
    }
    s_off[tid] = excl;
    if (tid == 0) s_off[T] = total;
    __syncthreads();
    ReactHit<NS> h;
    react_fetch<NS>(hits, wterm, s_off, T, g0, total, tid, h);
    for (uint32_t p0 = 0; p0 < total; p0 += T) {
      // positions in the per-side lists: per wavefront, the hits of lower
      // lanes (ballot and bit count); across wavefronts, their totals
      uint32_t wpos[NS];
#pragma unroll
      for (int sd = 0; sd < NS; ++sd) {
        const unsigned long long b = __ballot((h.mask >> sd) & 1u);
        wpos[sd] = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                             __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
        if (lane == 0) s_wcnt[wave][sd] = (uint32_t)__popcll(b);
      }
      if (tid == 0) s_nlong = 0;
      __syncthreads();
#pragma unroll
      for (int sd = 0; sd < NS; ++sd)
        if ((h.mask >> sd) & 1u) {
          uint32_t ps = wpos[sd];
          for (uint32_t q = 0; q < wave; ++q) ps += s_wcnt[q][sd];
          s_key[sd][ps] = h.key;
          s_side[sd][ps][0] = h.t[sd][0];
          s_side[sd][ps][1] = h.t[sd][1];
        }
      __syncthreads();
      // the next piece's loads overlap the sums below
      react_fetch<NS>(hits, wterm, s_off, T, g0, total, p0 + T + tid, h);
      for (uint32_t a = tid; a < n_assign; a += T) {
        const uint32_t inst = a / PER_INST, sd = (a / D) % NS;
        uint32_t nside = 0;
        for (uint32_t q = 0; q < NW; ++q) nside += s_wcnt[q][sd];
        // this instance's entries of the side's list (keys ascend, instances
        // own consecutive key ranges)
        uint32_t lo = 0, e1 = nside;
        if (R.n_insts > 1) {
          const uint32_t kb = R.insts[inst].ckey_begin, ke = R.insts[inst].ckey_end;
          uint32_t hi = nside;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_key[sd][mid] < kb) lo = mid + 1;
            else hi = mid;
          }
          e1 = lo;
          uint32_t e2 = nside;
          while (e1 < e2) {
            const uint32_t mid = (e1 + e2) >> 1;
            if (s_key[sd][mid] < ke) e1 = mid + 1;
            else e2 = mid;
          }
        }
        if (lo == e1) continue;
        const ChainTerms ct = chain_of(a, lo);
        const uint32_t n = (e1 - lo) * (ct.two ? 2u : 1u);
        if (n > REACT_SEQ_MAX) {
          const uint32_t q = atomicAdd(&s_nlong, 1u);
          if (q < REACT_LONG_MAX) {
            s_long[q][0] = a;
            s_long[q][1] = lo;
            s_long[q][2] = e1;
            continue;
          }
        }
        double acc = (a < REACT_LDS_SUMS) ? s_acc[a] : *out_at(a);
        acc = ct.run(acc, 0, n);
        if (a < REACT_LDS_SUMS) s_acc[a] = acc;
        else *out_at(a) = acc;
      }
      __syncthreads();
      const uint32_t nlong = min(s_nlong, REACT_LONG_MAX);
      for (uint32_t q = wave; q < nlong; q += NW) {
        const uint32_t a = s_long[q][0], lo = s_long[q][1], e1 = s_long[q][2];
        const ChainTerms ct = chain_of(a, lo);
        const uint32_t n = (e1 - lo) * (ct.two ? 2u : 1u);
        const double s0 = (a < REACT_LDS_SUMS) ? s_acc[a] : *out_at(a);
        const double r = react_chain_wave(ct, n, s0, lane);
        if (lane == 0) {
          if (a < REACT_LDS_SUMS) s_acc[a] = r;
          else *out_at(a) = r;
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t a = tid; a < n_assign && a < REACT_LDS_SUMS; a += T)
    *out_at(a) = s_acc[a];
}

// Diagnostics: the wall reaction's ordered subtraction on given chains, so
// tests can compare it with the host's sequential loop.  Chain c holds
// entries [eoff[c], eoff[c+1]) of `ent` (4 doubles each; entry `zero` is all
// zero) and takes terms 0 and 1 of every entry (two[c] != 0) or term 2.  One
// wavefront per chain writes react_chain_wave's result to wave[c] and one
// lane's plain loop to seq[c].
__global__ __launch_bounds__(64) void k_chain_check(const double *ent, const uint64_t *eoff,
                                                    const int32_t *two, const double *s0,
                                                    uint32_t zero, double *wave, double *seq) {
  const uint32_t c = blockIdx.x, lane = threadIdx.x;
  ChainTerms ct;
  ct.base = ent;
  ct.lo = (uint32_t)eoff[c];
  ct.two = two[c] != 0;
  ct.o1 = ct.two ? 0u : 2u;
  ct.zero = zero;
  const uint32_t n = (uint32_t)(eoff[c + 1] - eoff[c]) * (ct.two ? 2u : 1u);
  const double r = react_chain_wave(ct, n, s0[c], lane);
  if (lane == 0) {
    wave[c] = r;
    seq[c] = ct.run(s0[c], 0, n);
  }
}

// ------------------------------------------------------- neighbour list

template <int D>
__device__ __forceinline__ bool contact_x(const StepArgs &A, uint32_t t, uint32_t i,
                                          double *x, uint32_t &inst) {
  if (!(A.node_info[i] & NI_CONTACT)) return false;
  const uint32_t gi = A.node_grain[i];
  const GrainDev gp = A.grains[gi];
  double Xc[D], un[D];
  pos_X<D>(A, A.insts[gp.inst], gi, i, Xc);
  rec_u<D>(A.rec[t & 1], A.ns, i, un);
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = Xc[k] + un[k];
  inst = gp.inst;
  return true;
}

// Per-instance box of contact nodes at step t (list build).
// Workgroups of 1024 nodes combine their wavefronts' boxes in LDS first, so
// an instance's keys see one global atomic per workgroup, not per wavefront.
template <int D>
__global__ __launch_bounds__(1024) void k_inst_box(const StepArgs A, const uint32_t t,
                                                   unsigned long long *ibox) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ BlockBox bb;
  if (threadIdx.x < BOX_SLOTS * 2 * D) {
    const uint32_t k = threadIdx.x % (2 * D);
    bb.key[threadIdx.x / (2 * D)][k] = k < (uint32_t)D ? ~0ull : 0ull;
  }
  if (threadIdx.x == 0) {
    // instance of the workgroup's first node: the smallest of its nodes'
    const uint32_t f = blockIdx.x * blockDim.x;
    bb.g0 = f < A.n_nodes ? A.grains[A.node_grain[f]].inst : 0xFFFFFFFFu;
    bb.done = 0;
  }
  __syncthreads();
  bool valid = false;
  uint32_t inst = 0xFFFFFFFFu;
  double x[D];
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = 0.0;
  if (i < A.n_nodes) valid = contact_x<D>(A, t, i, x, inst);
  if (!valid) inst = 0xFFFFFFFFu;
  wave_box_block<D>(ibox, inst, valid, x, bb, blockDim.x / 64);
}

template <int D>
__device__ __forceinline__ void cell_coords(const GridDev &gd, const double *x, int *c) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double f = floor((x[k] - gd.origin[k]) * gd.inv_cell);
    int v = (f < 0.0) ? 0 : (f > (double)(gd.dims[k] - 1) ? gd.dims[k] - 1 : (int)f);
    c[k] = v;
  }
}
template <int D> __device__ __forceinline__ uint32_t cell_lin(const GridDev &gd, const int *c) {
  if (D == 2) return gd.cell_begin + (uint32_t)c[1] * gd.dims[0] + c[0];
  return gd.cell_begin + ((uint32_t)c[2] * gd.dims[1] + c[1]) * gd.dims[0] + c[0];
}

template <int D>
__global__ void k_cell_count(const StepArgs A, const uint32_t t, const GridDev *grids,
                             uint32_t *cell_of, uint32_t *cell_cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n_nodes) return;
  double x[D];
  uint32_t inst;
  if (!contact_x<D>(A, t, i, x, inst)) {
    cell_of[i] = 0xFFFFFFFFu;
    return;
  }
  int c[D];
  const GridDev gd = grids[inst];
  cell_coords<D>(gd, x, c);
  const uint32_t cl = cell_lin<D>(gd, c);
  cell_of[i] = cl;
  atomicAdd(&cell_cnt[cl], 1u);
}

__global__ void k_cell_fill(uint32_t n, const uint32_t *cell_of, const uint32_t *cell_start,
                            uint32_t *cell_cur, uint32_t *cell_nodes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t cl = cell_of[i];
  if (cl == 0xFFFFFFFFu) return;
  const uint32_t p = cell_start[cl] + atomicAdd(&cell_cur[cl], 1u);
  cell_nodes[p] = i;
}

// Count (FILL=false) or write (FILL=true) the Verlet-list entries of node i:
// every other contact node of the same instance within the list radius at
// build time.
// The fill pass also writes the node's fast list: its entries without an
// intact bond, as (partner, partner grain), and their number.
template <int D, bool FILL>
__global__ void k_list(const StepArgs A, const uint32_t t, const GridDev *grids,
                       const uint32_t *cell_start, const uint32_t *cell_nodes, uint32_t *cnt,
                       const uint32_t *ptr, uint32_t *lj, int32_t *lslot, uint2 *lfast,
                       uint32_t *nfast, double *xb) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n_nodes) return;
  double x[D];
  uint32_t inst;
  if (!contact_x<D>(A, t, i, x, inst)) {
    if (!FILL) cnt[i] = 0;
    else nfast[i] = 0;
    return;
  }
  const GridDev gd = grids[inst];
  const double r2 = gd.r2;
  int c[D];
  cell_coords<D>(gd, x, c);
  uint32_t n = 0;
  const uint32_t out0 = FILL ? ptr[i] : 0;
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < D; ++k) {
    lo[k] = c[k] > 0 ? c[k] - 1 : 0;
    hi[k] = c[k] < gd.dims[k] - 1 ? c[k] + 1 : gd.dims[k] - 1;
  }
  int cc[3];
  for (cc[2] = lo[2]; cc[2] <= hi[2]; ++cc[2])
    for (cc[1] = lo[1]; cc[1] <= hi[1]; ++cc[1])
      for (cc[0] = lo[0]; cc[0] <= hi[0]; ++cc[0]) {
        const uint32_t cl = cell_lin<D>(gd, cc);
        for (uint32_t p = cell_start[cl]; p < cell_start[cl + 1]; ++p) {
          const uint32_t j = cell_nodes[p];
          if (j == i) continue;
          double y[D];
          uint32_t ij;
          contact_x<D>(A, t, j, y, ij);
          double d[D];
#pragma unroll
          for (int k = 0; k < D; ++k) d[k] = y[k] - x[k];
          if (red<D>(d, d) <= r2) {
            if (FILL) lj[out0 + n] = j;
            ++n;
          }
        }
      }
  if (!FILL) {
    cnt[i] = n;
    return;
  }
  store_vec<D>(xb, A.ns, i, x);
  // sort by the reference's accumulation order: own grain first, then by
  // contact key (grain, then node or boundary-list position)
  const uint32_t g = A.node_grain[i];
  for (uint32_t a = 1; a < n; ++a) {
    const uint32_t v = lj[out0 + a];
    const unsigned long long kv =
        ((unsigned long long)(A.node_grain[v] == g ? 0 : 1) << 32) | A.ckey[v];
    uint32_t b = a;
    while (b > 0) {
      const uint32_t u = lj[out0 + b - 1];
      const unsigned long long ku =
          ((unsigned long long)(A.node_grain[u] == g ? 0 : 1) << 32) | A.ckey[u];
      if (ku <= kv) break;
      lj[out0 + b] = u;
      --b;
    }
    lj[out0 + b] = v;
  }
  // bond slot of same-grain partners, for the intact-bond exclusion
  const uint32_t sl = i >> 6, lane = i & 63;
  const uint32_t base = A.slice_base[sl], width = A.slice_width[sl];
  uint32_t nf = 0;
  for (uint32_t a = 0; a < n; ++a) {
    const uint32_t j = lj[out0 + a];
    const uint32_t gj = A.node_grain[j];
    int32_t slot = -1;
    bool intact = false;
    if (gj == g) {
      for (uint32_t k = 0; k < width; ++k) {
        const uint32_t e = (uint32_t)bslot(base, lane, k);
        const uint32_t w = A.bcol[e];
        if (w == BCOL_PAD) break;
        if ((w & BCOL_INDEX) == j) {
          slot = (int32_t)e;
          intact = !(w & BCOL_BROKEN);
          break;
        }
      }
    }
    lslot[out0 + a] = slot;
    if (!intact) lfast[out0 + nf++] = make_uint2(j, gj);
  }
  nfast[i] = nf;
}

// Contact pairs of the completed step t (readback for parity checks).
template <int D, bool FILL>
__global__ void k_pairs(const StepArgs A, const uint32_t t, uint32_t *cnt, const uint32_t *ptr,
                        uint32_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n_nodes) return;
  uint32_t n = 0;
  const uint32_t g = A.node_grain[i];
  const GrainDev gp = A.grains[g];
  const InstDev &ip = A.insts[gp.inst];
  const uint8_t info = A.node_info[i];
  const bool mov = grain_movable(gp, ip, g, t);
  const bool brk = (ip.flags & IF_FRACTURE) && (gp.flags & GF_BREAKABLE);
  const bool selfc = brk && (ip.flags & IF_SELF_CONTACT);
  const bool stop = (ip.flags & IF_UNGATED) || grain_stoppable(gp, ip, g, t);
  if (mov && (info & NI_CONTACT) && (selfc || stop)) {
    const double *U = A.rec[t & 1];
    double Xc[D], xc[D], box[2 * D], ui[D];
    pos_X<D>(A, ip, g, i, Xc);
    rec_u<D>(U, A.ns, i, ui);
#pragma unroll
    for (int k = 0; k < D; ++k) xc[k] = Xc[k] + ui[k];
    load_box<D>(A.bbox[t % 3], g, box);
    const uint32_t o = FILL ? ptr[i] : 0;
    for (uint32_t c = A.cl_ptr[i]; c < A.cl_ptr[i + 1]; ++c) {
      const uint32_t j = A.cl_j[c];
      double Xj[D], uj[D], d[D], r, R, K;
      pos_X<D>(A, ip, A.node_grain[j], j, Xj);
      rec_u<D>(U, A.ns, j, uj);
      if (!contact_gate<D>(A, t, A.node_grain[j], A.cl_slot[c], g, gp, ip, Xc, Xj, selfc, stop,
                           box, R, K) ||
          !pair_within<D>(Xj, uj, xc, R, d, r))
        continue;
      if (FILL) {
        out[2 * (o + n)] = i;
        out[2 * (o + n) + 1] = j;
      }
      ++n;
    }
  }
  if (!FILL) cnt[i] = n;
}

// Up to four device-to-device copies of 8-byte words in one launch (the
// per-step snapshots of a differentiable run's recompute).
__global__ void k_copy4(CopySet cs) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int r = 0; r < 4; ++r)
    for (size_t i = i0; i < cs.n[r]; i += stride) cs.dst[r][i] = cs.src[r][i];
}

hipError_t launch_copy4(const CopySet &cs, hipStream_t s) {
  size_t n = 0;
  for (int r = 0; r < 4; ++r) n = cs.n[r] > n ? cs.n[r] : n;
  if (n == 0) return hipSuccess;
  const uint32_t nb = (uint32_t)std::min<size_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_copy4, dim3(nb), dim3(256), 0, s, cs);
  return hipGetLastError();
}

__global__ void k_fill_u64(unsigned long long *p, size_t n, unsigned long long a,
                           unsigned long long b, uint32_t period, uint32_t split) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = ((i % period) < split) ? a : b;
}

// Diagnostics: the division and square root the step kernels use (the
// shortened sequences inside their range, the IEEE operators outside),
// applied elementwise, so tests can compare their rounding with the host's.
__global__ void k_arith(uint64_t n, const double *a, const double *b, double *q, double *r) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#if PG_FAST_DIVSQRT
  q[i] = div_rn(a[i], b[i]);
  r[i] = sqrt_rn(a[i]);
#else
  q[i] = a[i] / b[i];
  r[i] = sqrt(a[i]);
#endif
}

// Diagnostics: bond_geom itself -- the bond loop's length, stretch and
// direction arithmetic, its range test (CHECK) or the nice-input shortcut
// (!CHECK) -- on given xi and eta vectors ([n][D]), so tests can compare its
// rounding with the host's IEEE operators at the edges of its ranges.
template <int D, bool CHECK>
__global__ void k_bond_geom(uint64_t n, const double *xi, const double *eta, double *s,
                            double *e) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[D], y[D], ee[D], ss;
#pragma unroll
  for (int q = 0; q < D; ++q) {
    x[q] = xi[i * D + q];
    y[q] = eta[i * D + q];
  }
  bond_geom<D, CHECK>(x, y, true, ss, ee);
  s[i] = ss;
#pragma unroll
  for (int q = 0; q < D; ++q) e[i * D + q] = ee[q];
}

// ------------------------------------------------------------- launchers

static inline uint32_t nblk(size_t n, int b) { return (uint32_t)((n + b - 1) / b); }

template <int D, bool RECORD>
static void launch_step_t(bool uvol, const StepArgs &A, uint32_t t, uint32_t nb, int block,
                          hipStream_t s) {
  if (A.win_lo) {
    const dim3 g(A.lds_blocks), b(A.lds_threads);
    if (uvol) hipLaunchKernelGGL((k_step_lds<D, RECORD, true>), g, b, A.lds_bytes, s, A, t);
    else hipLaunchKernelGGL((k_step_lds<D, RECORD, false>), g, b, A.lds_bytes, s, A, t);
    return;
  }
  // lanes per node for small workloads, where one lane per node leaves most
  // SIMDs without a wavefront: four up to 16k nodes (at most one wavefront
  // per SIMD), two up to 64k.  Measured against one lane per node: 1.3x on
  // 1.4k and 15k-node 2D runs and 1.7x on a 3.9k-node 3D run; at 122k nodes
  // (two wavefronts per SIMD already) two lanes per node were 9% slower.
  const uint32_t lpn = A.n_nodes <= 16384u ? 4u : (A.n_nodes <= 65536u ? 2u : 1u);
  const uint32_t nbl = (uint32_t)nblk((uint64_t)A.n_nodes * lpn, block);
  const dim3 g(nbl > 0 ? nbl : 1), b(block);
  if (lpn == 4) {
    if (uvol) hipLaunchKernelGGL((k_step<D, RECORD, true, 4>), g, b, 0, s, A, t);
    else hipLaunchKernelGGL((k_step<D, RECORD, false, 4>), g, b, 0, s, A, t);
  } else if (lpn == 2) {
    if (uvol) hipLaunchKernelGGL((k_step<D, RECORD, true, 2>), g, b, 0, s, A, t);
    else hipLaunchKernelGGL((k_step<D, RECORD, false, 2>), g, b, 0, s, A, t);
  } else {
    if (uvol) hipLaunchKernelGGL((k_step<D, RECORD, true, 1>), dim3(nb), b, 0, s, A, t);
    else hipLaunchKernelGGL((k_step<D, RECORD, false, 1>), dim3(nb), b, 0, s, A, t);
  }
}

template <int D>
static hipError_t lds_attr_t(uint32_t bytes) {
  hipError_t e = hipSuccess;
  const void *f[4] = {(const void *)k_step_lds<D, false, false>, (const void *)k_step_lds<D, false, true>,
                      (const void *)k_step_lds<D, true, false>, (const void *)k_step_lds<D, true, true>};
  for (int q = 0; q < 4 && e == hipSuccess; ++q)
    e = hipFuncSetAttribute(f[q], hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  return e;
}

// Allow the LDS step kernels `bytes` of dynamic LDS per workgroup.
hipError_t set_step_lds_limit(int dim, uint32_t bytes) {
  return dim == 2 ? lds_attr_t<2>(bytes) : lds_attr_t<3>(bytes);
}

hipError_t launch_step(int dim, bool record, const StepArgs &A, uint32_t t, int block,
                       hipStream_t s) {
  const uint32_t nb = nblk(A.n_nodes, block) > 0 ? nblk(A.n_nodes, block) : 1;
  const bool uvol = A.uniform_vol != 0;
  if (dim == 2) {
    if (record) launch_step_t<2, true>(uvol, A, t, nb, block, s);
    else launch_step_t<2, false>(uvol, A, t, nb, block, s);
  } else {
    if (record) launch_step_t<3, true>(uvol, A, t, nb, block, s);
    else launch_step_t<3, false>(uvol, A, t, nb, block, s);
  }
  return hipGetLastError();
}

hipError_t launch_prologue(int dim, const StepArgs &A, hipStream_t s) {
  const uint32_t nb = nblk(A.n_nodes, 256) > 0 ? nblk(A.n_nodes, 256) : 1;
  if (dim == 2) hipLaunchKernelGGL(k_prologue<2>, dim3(nb), dim3(256), 0, s, A);
  else hipLaunchKernelGGL(k_prologue<3>, dim3(nb), dim3(256), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_boxes(int dim, const StepArgs &A, uint32_t t, unsigned long long *keys,
                        hipStream_t s) {
  const uint32_t nb = nblk(A.n_nodes, 256) > 0 ? nblk(A.n_nodes, 256) : 1;
  if (dim == 2) hipLaunchKernelGGL(k_boxes<2>, dim3(nb), dim3(256), 0, s, A, t, keys);
  else hipLaunchKernelGGL(k_boxes<3>, dim3(nb), dim3(256), 0, s, A, t, keys);
  return hipGetLastError();
}

hipError_t launch_grain_stats(int dim, const StepArgs &A, uint32_t t, uint32_t g0,
                              uint32_t ng, double *out, hipStream_t s) {
  if (ng == 0) return hipSuccess;
  if (dim == 2) hipLaunchKernelGGL(k_grain_stats<2>, dim3(ng), dim3(256), 0, s, A, t, g0, out);
  else hipLaunchKernelGGL(k_grain_stats<3>, dim3(ng), dim3(256), 0, s, A, t, g0, out);
  return hipGetLastError();
}

hipError_t launch_grain_mean(int dim, const StepArgs &A, uint32_t t, double *gmean,
                             hipStream_t s) {
  const uint32_t nb = nblk(A.n_grains, 64);
  if (dim == 2) hipLaunchKernelGGL(k_grain_mean<2>, dim3(nb), dim3(64), 0, s, A, t, gmean);
  else hipLaunchKernelGGL(k_grain_mean<3>, dim3(nb), dim3(64), 0, s, A, t, gmean);
  return hipGetLastError();
}

hipError_t launch_reactions(int dim, const ReactionArgs &R, uint32_t nsteps, hipStream_t s) {
  if (nsteps == 0) return hipSuccess;
  if (R.nchunks > 0)
    hipLaunchKernelGGL(k_wall_compact, dim3(R.nchunks, nsteps), dim3(256), 0, s, R);
  // one step: pieces of 512; a ring of steps: one block of 256 per step
  if (nsteps == 1) {
    if (dim == 2) hipLaunchKernelGGL((k_reaction<2, 512>), dim3(1, 1), dim3(512), 0, s, R);
    else hipLaunchKernelGGL((k_reaction<3, 512>), dim3(1, 1), dim3(512), 0, s, R);
  } else {
    if (dim == 2) hipLaunchKernelGGL((k_reaction<2, 256>), dim3(1, nsteps), dim3(256), 0, s, R);
    else hipLaunchKernelGGL((k_reaction<3, 256>), dim3(1, nsteps), dim3(256), 0, s, R);
  }
  return hipGetLastError();
}

hipError_t launch_fill_box(unsigned long long *p, size_t n, int dim, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(nblk(n, 256)), dim3(256), 0, s, p, n, ~0ull, 0ull,
                     (uint32_t)(2 * dim), (uint32_t)dim);
  return hipGetLastError();
}

hipError_t launch_inst_box(int dim, const StepArgs &A, uint32_t t, unsigned long long *ibox,
                           hipStream_t s) {
  const uint32_t nb = nblk(A.n_nodes, 1024) > 0 ? nblk(A.n_nodes, 1024) : 1;
  if (dim == 2) hipLaunchKernelGGL(k_inst_box<2>, dim3(nb), dim3(1024), 0, s, A, t, ibox);
  else hipLaunchKernelGGL(k_inst_box<3>, dim3(nb), dim3(1024), 0, s, A, t, ibox);
  return hipGetLastError();
}

hipError_t launch_cell_count(int dim, const StepArgs &A, uint32_t t, const GridDev *grids,
                             uint32_t *cell_of, uint32_t *cell_cnt, hipStream_t s) {
  const uint32_t nb = nblk(A.n_nodes, 256);
  if (nb == 0) return hipSuccess;
  if (dim == 2)
    hipLaunchKernelGGL(k_cell_count<2>, dim3(nb), dim3(256), 0, s, A, t, grids, cell_of, cell_cnt);
  else
    hipLaunchKernelGGL(k_cell_count<3>, dim3(nb), dim3(256), 0, s, A, t, grids, cell_of, cell_cnt);
  return hipGetLastError();
}

hipError_t launch_cell_fill(uint32_t n, const uint32_t *cell_of, const uint32_t *cell_start,
                            uint32_t *cell_cur, uint32_t *cell_nodes, hipStream_t s) {
  const uint32_t nb = nblk(n, 256);
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_fill, dim3(nb), dim3(256), 0, s, n, cell_of, cell_start, cell_cur,
                     cell_nodes);
  return hipGetLastError();
}

hipError_t launch_list(int dim, bool fill, const StepArgs &A, uint32_t t, const GridDev *grids,
                       const uint32_t *cell_start, const uint32_t *cell_nodes, uint32_t *cnt,
                       const uint32_t *ptr, uint32_t *lj, int32_t *lslot, uint2 *lfast,
                       uint32_t *nfast, double *xb,
                       hipStream_t s) {
  const uint32_t nb = nblk(A.n_nodes, 128);
  if (nb == 0) return hipSuccess;
  if (dim == 2) {
    if (fill)
      hipLaunchKernelGGL((k_list<2, true>), dim3(nb), dim3(128), 0, s, A, t, grids, cell_start,
                         cell_nodes, cnt, ptr, lj, lslot, lfast, nfast, xb);
    else
      hipLaunchKernelGGL((k_list<2, false>), dim3(nb), dim3(128), 0, s, A, t, grids, cell_start,
                         cell_nodes, cnt, ptr, lj, lslot, lfast, nfast, xb);
  } else {
    if (fill)
      hipLaunchKernelGGL((k_list<3, true>), dim3(nb), dim3(128), 0, s, A, t, grids, cell_start,
                         cell_nodes, cnt, ptr, lj, lslot, lfast, nfast, xb);
    else
      hipLaunchKernelGGL((k_list<3, false>), dim3(nb), dim3(128), 0, s, A, t, grids, cell_start,
                         cell_nodes, cnt, ptr, lj, lslot, lfast, nfast, xb);
  }
  return hipGetLastError();
}

hipError_t launch_pairs(int dim, bool fill, const StepArgs &A, uint32_t t, uint32_t *cnt,
                        const uint32_t *ptr, uint32_t *out, hipStream_t s) {
  const uint32_t nb = nblk(A.n_nodes, 128);
  if (nb == 0) return hipSuccess;
  if (dim == 2) {
    if (fill) hipLaunchKernelGGL((k_pairs<2, true>), dim3(nb), dim3(128), 0, s, A, t, cnt, ptr, out);
    else hipLaunchKernelGGL((k_pairs<2, false>), dim3(nb), dim3(128), 0, s, A, t, cnt, ptr, out);
  } else {
    if (fill) hipLaunchKernelGGL((k_pairs<3, true>), dim3(nb), dim3(128), 0, s, A, t, cnt, ptr, out);
    else hipLaunchKernelGGL((k_pairs<3, false>), dim3(nb), dim3(128), 0, s, A, t, cnt, ptr, out);
  }
  return hipGetLastError();
}

hipError_t launch_arith(uint64_t n, const double *a, const double *b, double *q, double *r,
                        hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_arith, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, n, a, b, q, r);
  return hipGetLastError();
}

hipError_t launch_chain_check(uint32_t nchains, const double *ent, const uint64_t *eoff,
                              const int32_t *two, const double *s0, uint32_t zero, double *wave,
                              double *seq, hipStream_t s) {
  if (nchains == 0) return hipSuccess;
  hipLaunchKernelGGL(k_chain_check, dim3(nchains), dim3(64), 0, s, ent, eoff, two, s0, zero, wave,
                     seq);
  return hipGetLastError();
}

hipError_t launch_bond_geom(int dim, bool check, uint64_t n, const double *xi, const double *eta,
                            double *s, double *e, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const dim3 g((uint32_t)((n + 255) / 256)), b(256);
  if (dim == 2) {
    if (check) hipLaunchKernelGGL((k_bond_geom<2, true>), g, b, 0, st, n, xi, eta, s, e);
    else hipLaunchKernelGGL((k_bond_geom<2, false>), g, b, 0, st, n, xi, eta, s, e);
  } else {
    if (check) hipLaunchKernelGGL((k_bond_geom<3, true>), g, b, 0, st, n, xi, eta, s, e);
    else hipLaunchKernelGGL((k_bond_geom<3, false>), g, b, 0, st, n, xi, eta, s, e);
  }
  return hipGetLastError();
}

hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, void *tmp,
                              size_t *tmp_bytes, hipStream_t s) {
  return hipcub::DeviceScan::ExclusiveSum(tmp, *tmp_bytes, in, out, (int)n, s);
}

} // namespace pg
