// pg_adjoint.hip -- adjoint (reverse-mode) kernels of the time step for the
// differentiable form.  No reference counterpart: the reference has no
// derivatives.  The forward map being differentiated is the step of
// pg_kernels.hip (timeloop.h:608-878):
//   F^t  = F(u^t, v^{t-1}; c, Kn, mu)            forces (bonds, contact, walls)
//   a^t  = (1/rho) F^t
//   v^t  = v^{t-1} + ((a^{t-1} + a^t) * (dt/2))
//   u^{t+1} = u^t + ((dt v^t) + (dt^2/2) a^t)
//   R^t  = -sum_i V_i f_wall,i                   wall reaction (the loss reads it)
// with bond status, contact pairs and every branch frozen to what the forward
// pass did.  Going backwards through step t with P_u = dL/du^{t+1},
// P_v = dL/dv^t (partial) and P_a = dL/da^t (partial):
//   lv = P_v + dt P_u,  la = P_a + h P_u + (dt/2) lv,  lF = (1/rho) la
//   P_u <- P_u + (dF/du^t)^T lF + (dR/du^t)^T lR
//   P_v <- lv  + (dF/dv^{t-1})^T lF + (dR/dv^{t-1})^T lR
//   P_a <- (dt/2) lv
// Bond Jacobians are written out (the stretch/direction derivative matrix);
// contact and wall Jacobians come from the force functions evaluated on
// Dual<2*dim+2> numbers seeded along (x_i, v_i, K, mu).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "pg_device.h"
#include "pg_internal.h"
#include "pg_kernels.h"

namespace pg {

namespace {

template <int D>
__device__ __forceinline__ void load_cols(const double *p, uint32_t ns, uint32_t i, double *v) {
#pragma unroll
  for (int k = 0; k < D; ++k) v[k] = p[(size_t)k * ns + i];
}

// lF of a node from its cotangents (the first lines of the step's adjoint)
template <int D>
__device__ __forceinline__ void force_cotangent(const InstDev &ip, double rho, const double *Pu,
                                                const double *Pv, const double *Pa, double *lv,
                                                double *lF) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    lv[k] = Pv[k] + ip.dt * Pu[k];
    const double la = (Pa[k] + ip.h * Pu[k]) + ip.half_dt * lv[k];
    lF[k] = (1 / rho) * la;
  }
}

// Contact gate of node n's grain toward a partner (timeloop.h:698-747 and
// contact.h:686-733) for step t: returns false if no force is computed, else
// the pair's radius R and stiffness K (K is c/rd or c/Rc for self contact).
template <int D>
__device__ __forceinline__ bool pair_gate(const AdjArgs &B, uint32_t t, const GrainDev &gn,
                                          uint32_t gn_id, const InstDev &ip, uint32_t gm_id,
                                          const double *Xn, const double *Xm, int32_t bslot_nm,
                                          double &R, double &K, bool &self) {
  const StepArgs &A = B.A;
  const bool brk = (ip.flags & IF_FRACTURE) && (gn.flags & GF_BREAKABLE);
  const bool selfc = brk && (ip.flags & IF_SELF_CONTACT);
  const bool stop = (gn.flags & GF_STOPPABLE) != 0;
  R = ip.Rc;
  K = ip.Kn;
  self = (gm_id == gn_id);
  if (self) {
    if (!selfc) return false;
    // bonded at step t: the bond has not broken at step t or earlier
    if (bslot_nm >= 0 && A.break_step[bslot_nm] > t) return false;
    double dX[D];
#pragma unroll
    for (int k = 0; k < D; ++k) dX[k] = Xn[k] - Xm[k];
    const double rd = sqrt(red<D>(dX, dX));
    if (rd < ip.Rc) {
      R = rd;
      K = gn.cnot / rd;
    } else {
      R = ip.Rc;
      K = gn.cnot / ip.Rc;
    }
    return true;
  }
  if (!stop) return false;
  double bn[2 * D], bm[2 * D];
  load_box<D>(B.sbox, gn_id, bn);
  load_box<D>(B.sbox, gm_id, bm);
  return boxes_meet<D>(bn, bm, ip.Rc);
}

} // namespace

// One node per L = ADJ_LANES consecutive lanes: lane r of node i takes slot
// r % BGROUP of every (L/BGROUP)-th column group starting at group
// r / BGROUP (slots r, r + L, ...), contact entries r, r + L, ... of i's
// list and wall sides r, r + L, ...; the lanes' partial
// cotangents are summed with shuffles and lane 0 writes the node's result.
// The adjoint has no reference order to keep, so the sums run in any fixed
// order.
constexpr uint32_t ADJ_BLOCK = 256; // threads per workgroup of k_adj_step

template <int D>
__global__ __launch_bounds__(ADJ_BLOCK) void k_adj_step(const AdjArgs B, const uint32_t t) {
  constexpr int N = 2 * D + 2; // dual directions: x_i (D), v_i (D), K, mu
  constexpr uint32_t L = ADJ_LANES;
  static_assert(L % BGROUP == 0 && (L & (L - 1)) == 0, "whole column groups per pass");
  using DN = Dual<N>;
  const StepArgs &A = B.A;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = gid / L, r = gid % L;
  const uint32_t ns = A.ns;
  double gc = 0.0, gk = 0.0, gm = 0.0;
  double au[D], av[D], Pu[D], lv[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    au[k] = 0.0;
    av[k] = 0.0;
    Pu[k] = 0.0;
    lv[k] = 0.0;
  }
  bool act = false; // node i exists and its grain moves (the same in all L lanes)
  uint32_t g = 0;
  if (i < A.n_nodes) {
    g = A.node_grain[i];
    act = (A.grains[g].flags & GF_MOVABLE) != 0;
  }
  if (act) {
    const GrainDev gp = A.grains[g];
    const InstDev &ip = A.insts[gp.inst];
    double Pv[D], Pa[D], lF[D];
    load_cols<D>(B.Pu, ns, i, Pu);
    load_cols<D>(B.Pv, ns, i, Pv);
    load_cols<D>(B.Pa, ns, i, Pa);
    force_cotangent<D>(ip, gp.rho, Pu, Pv, Pa, lv, lF);
    double Xi[D], ui[D], xi[D], vi[D];
    rec_X<D>(A, i, Xi);
    rec_u<D>(B.su, A.ns, i, ui);
    load_cols<D>(B.sv, ns, i, vi);
#pragma unroll
    for (int k = 0; k < D; ++k) xi[k] = Xi[k] + ui[k];
    const double Vi = A.V[i];
    const double c = gp.cnot;

    // ---- bonds: F_i = sum_j c V_j s e,  dF_i/du_j = c V_j M,  M symmetric.
    // Padding slots and bonds removed after an earlier step read the ghost
    // (finite values) and their terms are selected away.
    {
      const uint32_t sl = i >> 6, lane = i & 63;
      const uint32_t base = A.slice_base[sl], width = A.slice_width[sl];
      const uint32_t *bc = A.bcol + base + BGROUP * lane + r % BGROUP;
      const uint32_t *bk = A.break_step + base + BGROUP * lane + r % BGROUP;
      for (uint32_t k0 = BGROUP * (r / BGROUP); k0 < width; k0 += L) {
        const uint32_t cw = bc[SLICE * k0], bks = bk[SLICE * k0];
        const bool live = cw != BCOL_PAD && bks >= t; // a bond broken at t is still used at t
        const uint32_t j = live ? (cw & BCOL_INDEX) : A.ghost;
        double Xj[D], uj[D], lFj[D];
        rec_X<D>(A, j, Xj);
        rec_u<D>(B.su, A.ns, j, uj);
        load_cols<D>(B.lf_in, ns, j, lFj);
        const double Vj = A.V[j];
        double xb[D], eta[D], e[D], q[D];
#pragma unroll
        for (int m = 0; m < D; ++m) {
          xb[m] = Xj[m] - Xi[m];
          eta[m] = (uj[m] - ui[m]) + xb[m];
        }
        const double xn = sqrt(red<D>(xb, xb));
        const double n = sqrt(red<D>(eta, eta));
        // reciprocals by two Newton steps and products in place of the IEEE
        // divisions: the adjoint keeps no reference rounding, and these are
        // within an ulp or two for the bond lengths of any live bond
        const double rxn = recip_nr(xn), rn = recip_nr(n);
        const double s = (n - xn) * rxn;
        double eq = 0.0, el = 0.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
          e[m] = eta[m] * rn;
          q[m] = c * Vi * lFj[m] - c * Vj * lF[m];
          eq += e[m] * q[m];
          el += e[m] * lF[m];
        }
#pragma unroll
        for (int m = 0; m < D; ++m) {
          const double du = e[m] * eq * rxn + s * (q[m] - e[m] * eq) * rn;
          au[m] += live ? du : 0.0;
        }
        const double dc = c * Vj * s * el;
        gc += live ? dc : 0.0;
      }
    }

    const uint8_t info = A.node_info[i];
    const bool stop = (gp.flags & GF_STOPPABLE) != 0;

    // ---- node-node contact over the segment's candidate list (i's own
    // terms are gated by i's grain, the partners' terms by theirs)
    if (info & NI_CONTACT) {
      const uint32_t c0 = A.cl_ptr[i], c1 = A.cl_ptr[i + 1];
      for (uint32_t cc = c0 + r; cc < c1; cc += L) {
        // a same-grain pair still bonded at step t has no contact in either
        // direction (its two slots break in the same step): skip it before
        // loading the partner
        const int32_t bs = A.cl_slot[cc];
        if (bs >= 0 && A.break_step[bs] > t) continue;
        const uint32_t j = A.cl_j[cc];
        const uint32_t gj = A.node_grain[j];
        const GrainDev gpj = A.grains[gj];
        double Xj[D], uj[D], xj[D], vj[D];
        rec_X<D>(A, j, Xj);
        rec_u<D>(B.su, A.ns, j, uj);
        load_cols<D>(B.sv, ns, j, vj);
#pragma unroll
        for (int k = 0; k < D; ++k) xj[k] = Xj[k] + uj[k];
        // i's force from j: seeds on x_i, v_i, K, mu
        double R, K;
        bool self;
        if (pair_gate<D>(B, t, gp, g, ip, gj, Xi, Xj, bs, R, K, self)) {
          DN d[D], w[D], F[D];
#pragma unroll
          for (int k = 0; k < D; ++k) {
            d[k] = DN(Xj[k] + uj[k]) - DN::seed(xi[k], k);
            w[k] = DN::seed(vi[k], D + k) - DN(vj[k]);
            F[k] = DN(0.0);
          }
          const DN rr = sqrt(red<D, N>(d, d));
          if (rr.v <= R) {
            pair_force<D, DN>(d, rr, R, DN::seed(K, 2 * D), A.V[j], w, gp.rho,
                              DN::seed(ip.mu, 2 * D + 1), ip, F);
            double gK = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
#pragma unroll
              for (int m = 0; m < D; ++m) {
                au[m] += F[k].d[m] * lF[k];
                av[m] += F[k].d[D + m] * lF[k];
              }
              gK += F[k].d[2 * D] * lF[k];
              gm += F[k].d[2 * D + 1] * lF[k];
            }
            if (self) gc += gK * K; // K = c/rd: d/d(c scale) = K
            else gk += gK;
          }
        }
        // j's force from i (computed by j when j's grain moves): seeds on
        // x_i, v_i only; j accounts for its own parameter derivatives
        if ((gpj.flags & GF_MOVABLE)) {
          double Rj, Kj;
          bool selfj;
          if (pair_gate<D>(B, t, gpj, gj, ip, g, Xj, Xi, bs, Rj, Kj, selfj)) {
            DN d[D], w[D], F[D];
#pragma unroll
            for (int k = 0; k < D; ++k) {
              d[k] = DN::seed(xi[k], k) - DN(xj[k]);
              w[k] = DN(vj[k]) - DN::seed(vi[k], D + k);
              F[k] = DN(0.0);
            }
            const DN rr = sqrt(red<D, N>(d, d));
            if (rr.v <= Rj) {
              double lFj[D];
              load_cols<D>(B.lf_in, ns, j, lFj);
              pair_force<D, DN>(d, rr, Rj, DN(Kj), Vi, w, gpj.rho, DN(ip.mu), ip, F);
#pragma unroll
              for (int k = 0; k < D; ++k)
#pragma unroll
                for (int m = 0; m < D; ++m) {
                  au[m] += F[k].d[m] * lFj[k];
                  av[m] += F[k].d[D + m] * lFj[k];
                }
            }
          }
        }
      }
    }

    // ---- walls: cotangent of side s's force on i is lF - V_i lR_s
    if (stop && (ip.flags & IF_WALL) && (info & NI_CONTACT)) {
      const double *wp = B.swall + (size_t)gp.inst * 6;
      double boxg[2 * D];
      load_box<D>(B.sbox, g, boxg);
      if (!within_interior<D>(boxg, wp, ip.Rc)) {
        const double *lR = B.lR + (size_t)gp.inst * 6 * D;
        constexpr int NS = (D == 2) ? 4 : 6;
        const int sides2[4] = {3, 0, 1, 2}, dirs2[4] = {1, 0, 0, 1};
        const int sides3[6] = {0, 3, 1, 4, 2, 5}, dirs3[6] = {0, 0, 1, 1, 2, 2};
#pragma unroll
        for (int qq = 0; qq < NS; ++qq) {
          if ((uint32_t)qq % L != r) continue;
          const int side = (D == 2) ? sides2[qq] : sides3[qq];
          const int dir = (D == 2) ? dirs2[qq] : dirs3[qq];
          DN v[D], F[D];
#pragma unroll
          for (int k = 0; k < D; ++k) {
            v[k] = DN::seed(vi[k], D + k);
            F[k] = DN(0.0);
          }
          const DN dist = DN::seed(xi[dir], dir) - DN(wp[side]);
          if (wall_side<D, DN>(dist, dir, ip, DN::seed(ip.Kn, 2 * D),
                               DN::seed(ip.mu, 2 * D + 1), gp.rho, Vi, ip.speed[side], v, F,
                               nullptr)) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
              const double cot = lF[k] - Vi * lR[side * D + k];
#pragma unroll
              for (int m = 0; m < D; ++m) {
                au[m] += F[k].d[m] * cot;
                av[m] += F[k].d[D + m] * cot;
              }
              gk += F[k].d[2 * D] * cot;
              gm += F[k].d[2 * D + 1] * cot;
            }
          }
        }
      }
    }
  }

  // the node's L lanes are consecutive and share `act`: sum their partial
  // cotangents (every lane of the wavefront takes part in the shuffles)
#pragma unroll
  for (uint32_t o = 1; o < L; o <<= 1)
#pragma unroll
    for (int k = 0; k < D; ++k) {
      au[k] += __shfl_xor(au[k], (int)o);
      av[k] += __shfl_xor(av[k], (int)o);
    }

  // ---- cotangents of step t-1
  if (act && r == 0) {
    const GrainDev gp = A.grains[g];
    const InstDev &ip = A.insts[gp.inst];
    double nPu[D], nPv[D], nPa[D], nlv[D], nlF[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      nPu[k] = Pu[k] + au[k];
      nPv[k] = lv[k] + av[k];
      nPa[k] = ip.half_dt * lv[k];
      B.Pu[(size_t)k * ns + i] = nPu[k];
      B.Pv[(size_t)k * ns + i] = nPv[k];
      B.Pa[(size_t)k * ns + i] = nPa[k];
    }
    force_cotangent<D>(ip, gp.rho, nPu, nPv, nPa, nlv, nlF);
#pragma unroll
    for (int k = 0; k < D; ++k) B.lf_out[(size_t)k * ns + i] = nlF[k];
  }

  // deterministic block sums of the parameter partials: each wavefront's by
  // shuffles, then the four wavefronts' in a fixed order
  static_assert(ADJ_BLOCK == 256, "four wavefronts per workgroup");
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    gc += __shfl_xor(gc, off);
    gk += __shfl_xor(gk, off);
    gm += __shfl_xor(gm, off);
  }
  __shared__ double red_s[3][4];
  if ((threadIdx.x & 63) == 0) {
    red_s[0][threadIdx.x >> 6] = gc;
    red_s[1][threadIdx.x >> 6] = gk;
    red_s[2][threadIdx.x >> 6] = gm;
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      B.gpart[(size_t)blockIdx.x * 3 + q] += (red_s[q][0] + red_s[q][1]) + (red_s[q][2] + red_s[q][3]);
}

// Largest displacement of a contact node over `steps` records of the segment
// buffer, relative to the first (as the order-preserving key of a
// non-negative double, combined with an atomic max).
template <int D>
__global__ void k_adj_maxdisp(const StepArgs A, const double *su, uint32_t steps,
                              unsigned long long *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n_nodes || !(A.node_info[i] & NI_CONTACT)) return;
  const size_t stride = (size_t)A.ns * 2 * D;
  double u0[D];
  rec_u<D>(su, A.ns, i, u0);
  double m = 0.0;
  for (uint32_t s = 1; s < steps; ++s) {
    double u[D], dd[D];
    rec_u<D>(su + s * stride, A.ns, i, u);
#pragma unroll
    for (int k = 0; k < D; ++k) dd[k] = u[k] - u0[k];
    const double r = sqrt(red<D>(dd, dd));
    m = r > m ? r : m;
  }
  atomicMax(out, dkey(m));
}

// Bond status of the state after step t, from the recorded break steps.
__global__ void k_adj_restore_bcol(uint32_t *bcol, const uint32_t *break_step, uint64_t n,
                                   uint32_t t) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t w = bcol[e];
  if (w == BCOL_PAD) return;
  bcol[e] = (w & BCOL_INDEX) | (break_step[e] <= t ? BCOL_BROKEN : 0u);
}

// Break steps at the start of a recorded pass: 0 for bonds already broken,
// "never" for the rest.
__global__ void k_adj_init_break(const uint32_t *bcol, uint32_t *break_step, uint64_t n) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t w = bcol[e];
  break_step[e] = (w != BCOL_PAD && (w & BCOL_BROKEN)) ? 0u : 0xFFFFFFFFu;
}

hipError_t launch_adj_step(int dim, const AdjArgs &B, uint32_t t, uint32_t blocks,
                           hipStream_t s) {
  if (dim == 2) hipLaunchKernelGGL(k_adj_step<2>, dim3(blocks), dim3(ADJ_BLOCK), 0, s, B, t);
  else hipLaunchKernelGGL(k_adj_step<3>, dim3(blocks), dim3(ADJ_BLOCK), 0, s, B, t);
  return hipGetLastError();
}

hipError_t launch_adj_maxdisp(int dim, const StepArgs &A, const double *su, uint32_t steps,
                              unsigned long long *out, hipStream_t s) {
  const uint32_t nb = (A.n_nodes + 255) / 256;
  if (nb == 0) return hipSuccess;
  if (dim == 2) hipLaunchKernelGGL(k_adj_maxdisp<2>, dim3(nb), dim3(256), 0, s, A, su, steps, out);
  else hipLaunchKernelGGL(k_adj_maxdisp<3>, dim3(nb), dim3(256), 0, s, A, su, steps, out);
  return hipGetLastError();
}

hipError_t launch_adj_restore_bcol(uint32_t *bcol, const uint32_t *break_step, uint64_t n,
                                   uint32_t t, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_restore_bcol, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, bcol,
                     break_step, n, t);
  return hipGetLastError();
}

hipError_t launch_adj_init_break(const uint32_t *bcol, uint32_t *break_step, uint64_t n,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_init_break, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, bcol,
                     break_step, n);
  return hipGetLastError();
}

} // namespace pg
