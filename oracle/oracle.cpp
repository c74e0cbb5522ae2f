// oracle.cpp -- FP64 CPU restatement of the reference PeriDEM time loop.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU
// baseline ("port") of the project.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it; the product (libpg) never
// links, loads or calls it.
//
// It restates, operation by operation and in the same loop and summation
// order, the reference's per-step algorithm:
//   run_timeloop            src/particle/timeloop.h:500-1015
//   ParticleN               src/particle/particle2.h:28-264
//   get_peridynamic_force   src/particle/particle2.h:102-188
//   remove_bonds            src/particle/particle2.h:191-232
//   gen_xi                  src/particle/particle2.h:236-252
//   conn2NArr               src/read/load_hdf5.h:28-38
//   NbdArr2conn             src/particle/timeloop.h:25-49
//   segment_mass            src/particle/contact.h:16-41
//   wall_side_force_update  src/particle/contact.h:310-431
//   update_contact_force_by_boundary       src/particle/contact.h:667-812
//   update_wall_contact_force_by_boundary  src/particle/contact.h:899-964
//   intersects_box_fext / get_minmax / within_interior_fext
//                                          src/particle/contact.h:983-1080
//   vector operators        src/compat/overloads.h:35-159
// It is built with g++ -O2 -ffp-contract=off and no -march, like the
// reference's baseline x86-64 build (makefile:4,23), so no FMA is formed.
//
// Eigen reductions: the reference evaluates norm()/dot() of fixed-size row
// vectors with Eigen 3.3.9 (not in the reference tree).  For 2 components
// that is a0*b0 + a1*b1; for 3 components Eigen's SSE2 linear-vectorised,
// fully unrolled redux gives (a0*b0 + a1*b1) + a2*b2, which is what red3()
// below computes.  Eigen is absent here, so the 3-component order is
// restated from Eigen's published redux_impl (LinearVectorizedTraversal,
// CompleteUnrolling), not pinned by running it.
//
// Departures that do not change results:
//  - grains are passed by reference, not deep-copied per call
//    (contact.h:668-669,901 copy them by value);
//  - each grain's bounding box is computed once per step instead of once per
//    grain pair (get_minmax is a pure function of CurrPos);
//  - the wall reaction is accumulated in the sequential order (grain, node,
//    side) also when the force loop runs on several threads: each grain
//    records its reaction terms and they are subtracted in grain order after
//    the loop.  The reference's threaded accumulation is a data race
//    (contact.h:339 via timeloop.h:685).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "dual.h"
#include "perigrain.h"

namespace {

thread_local std::string g_err;

// The scalar type T is double (the oracle) or CDual<N> (its tangent mode,
// dual.h): same operations, same order, derivatives carried alongside.
using std::asin;
using std::fabs;
using std::sqrt;

template <unsigned D, class T> struct Vec {
  T c[D];
  T &operator()(unsigned k) { return c[k]; }
  const T &operator()(unsigned k) const { return c[k]; }
  static Vec zero() {
    Vec v;
    for (unsigned k = 0; k < D; ++k) v.c[k] = T(0.0);
    return v;
  }
};

template <unsigned D, class T> Vec<D, T> operator+(const Vec<D, T> &a, const Vec<D, T> &b) {
  Vec<D, T> r;
  for (unsigned k = 0; k < D; ++k) r.c[k] = a.c[k] + b.c[k];
  return r;
}
template <unsigned D, class T> Vec<D, T> operator-(const Vec<D, T> &a, const Vec<D, T> &b) {
  Vec<D, T> r;
  for (unsigned k = 0; k < D; ++k) r.c[k] = a.c[k] - b.c[k];
  return r;
}
template <unsigned D, class T, class S> Vec<D, T> operator*(const S &s, const Vec<D, T> &a) {
  Vec<D, T> r;
  for (unsigned k = 0; k < D; ++k) r.c[k] = s * a.c[k];
  return r;
}
template <unsigned D, class T, class S> Vec<D, T> operator/(const Vec<D, T> &a, const S &s) {
  Vec<D, T> r;
  for (unsigned k = 0; k < D; ++k) r.c[k] = a.c[k] / s;
  return r;
}
template <unsigned D, class T> Vec<D, T> &operator+=(Vec<D, T> &a, const Vec<D, T> &b) {
  for (unsigned k = 0; k < D; ++k) a.c[k] = a.c[k] + b.c[k];
  return a;
}

// Eigen 3.3.9 sum of a 2- or 3-term product expression (see header).
template <class T> inline T red(const Vec<2, T> &a, const Vec<2, T> &b) {
  return a.c[0] * b.c[0] + a.c[1] * b.c[1];
}
template <class T> inline T red(const Vec<3, T> &a, const Vec<3, T> &b) {
  return (a.c[0] * b.c[0] + a.c[1] * b.c[1]) + a.c[2] * b.c[2];
}
template <unsigned D, class T> inline T dot(const Vec<D, T> &a, const Vec<D, T> &b) {
  return red(a, b);
}
template <unsigned D, class T> inline T norm(const Vec<D, T> &a) {
  return sqrt(red(a, a));
}

// ParticleN<dim> (particle2.h:28-86) plus slot bookkeeping used only to
// report which original bond slots are still intact.
template <unsigned D, class T> struct Grain {
  using V = Vec<D, T>;
  unsigned nnodes = 0;
  std::vector<V> pos, disp, CurrPos, vel, acc, force, extforce;
  std::vector<double> vol;
  std::vector<unsigned> boundary_nodes, clamped_nodes;
  std::vector<std::vector<unsigned>> NbdArr;
  std::vector<unsigned> total_neighbors;
  std::vector<std::vector<V>> xi;
  std::vector<std::vector<T>> xi_norm, stretch;
  std::vector<std::vector<uint64_t>> slot; // original slot id of each entry
  double delta = 0, rho = 0, snot = 0;
  T cnot = T(0.0);
  int breakable = 1, movable = 1, stoppable = 1;
  bool break_bonds = false;
  unsigned torque_axis = 2;
  double torque_val = 0;

  // mean_CurrPos (particle2.h:89-98)
  V mean_CurrPos() const {
    V mean = V::zero();
    for (unsigned i = 0; i < nnodes; i++) mean += (pos[i] + disp[i]);
    for (unsigned k = 0; k < D; ++k) mean.c[k] = mean.c[k] / (double)nnodes;
    return mean;
  }

  // get_peridynamic_force (particle2.h:102-188)
  std::vector<V> get_peridynamic_force() {
    std::vector<V> tot(nnodes);
    for (unsigned i = 0; i < nnodes; ++i) {
      V v = V::zero();
      for (unsigned j = 0; j < total_neighbors[i]; ++j) {
        unsigned idx_nbr = NbdArr[i][j];
        V eta_p_xi = (disp[idx_nbr] - disp[i]) + xi[i][j];
        T eta_p_xi_norm = norm(eta_p_xi);
        V unit_dir = eta_p_xi / eta_p_xi_norm;
        T xi_norm_here = xi_norm[i][j];
        T diff_norm = (eta_p_xi_norm - xi_norm_here);
        T str = diff_norm / xi_norm_here;
        V v_more = ((cnot * str) * vol[idx_nbr]) * unit_dir;
        v += v_more;
        stretch[i][j] = str;
      }
      tot[i] = v;
    }
    return tot;
  }

  // remove_bonds (particle2.h:191-232): erase, in order, every bond whose
  // stretch exceeds snot (positive stretch only).
  void remove_bonds(std::vector<uint8_t> &intact) {
    if (!break_bonds) return;
    for (unsigned i = 0; i < nnodes; ++i) {
      size_t w = 0;
      size_t n = stretch[i].size();
      for (size_t j = 0; j < n; ++j) {
        if (stretch[i][j] > snot) {
          intact[slot[i][j]] = 0;
          --total_neighbors[i];
          continue;
        }
        if (w != j) {
          NbdArr[i][w] = NbdArr[i][j];
          xi[i][w] = xi[i][j];
          xi_norm[i][w] = xi_norm[i][j];
          stretch[i][w] = stretch[i][j];
          slot[i][w] = slot[i][j];
        }
        ++w;
      }
      NbdArr[i].resize(w);
      xi[i].resize(w);
      xi_norm[i].resize(w);
      stretch[i].resize(w);
      slot[i].resize(w);
    }
  }

  // gen_xi (particle2.h:236-252)
  void gen_xi() {
    for (unsigned i = 0; i < nnodes; ++i) {
      unsigned ss = (unsigned)NbdArr[i].size();
      total_neighbors[i] = ss;
      xi_norm[i].resize(ss);
      xi[i].resize(ss);
      stretch[i].resize(ss);
      for (unsigned j = 0; j < ss; ++j) {
        unsigned idx_nbr = NbdArr[i][j];
        xi[i][j] = pos[idx_nbr] - pos[i];
        xi_norm[i][j] = norm(xi[i][j]);
      }
    }
  }
};

template <class T> struct ContactP { // Contact (contact.h:43-112)
  double contact_rad, damping_ratio;
  T normal_stiffness, friction_coefficient;
  bool allow_friction, allow_damping, nl_bdry_only, self_contact;
};

// One reaction term: reaction[side](comp) -= value (contact.h:339,374,...).
template <class T> struct RTerm {
  uint8_t side, comp;
  T value;
};

// segment_mass<dim> (contact.h:16-41): {centroid, area}.
template <unsigned D, class T> void segment_mass(double R, T d, T &centroid, T &area) {
  T Area, I;
  if (D == 2) {
    Area = M_PI * (R * R) / 2 - (R * R) * asin(d / R) -
           sqrt(R * R - d * d) * d;
    I = T(0.0 * std::pow(sval(R * R - d * d), 1.5)); // 2/3 == 0 (integer division)
  } else {
    Area = M_PI / 3 * ((R - d) * (R - d)) * (2 * R - d);
    I = M_PI / 4 * ((R * R - d * d) * (R * R - d * d));
  }
  centroid = I / Area;
  area = Area;
}

// wall_side_force_update (contact.h:310-431).  `side` indexes the reaction.
template <unsigned D, class T>
void wall_side_force_update(T dist, const ContactP<T> &CN, unsigned dir,
                            const Vec<D, T> &vel_i, double vol_i, double rho,
                            Vec<D, T> &f, double dt, double wall_side_speed,
                            unsigned side, std::vector<RTerm<T>> *terms) {
  if (fabs(dist) <= CN.contact_rad) {
    T c0, c1;
    segment_mass<D, T>(CN.contact_rad, fabs(dist), c0, c1);
    T contact_rad_contrib = (CN.contact_rad - c0);
    T positive = (contact_rad_contrib > 0 ? contact_rad_contrib : T(0.0));
    T this_contact_force = (CN.normal_stiffness * positive * c1);
    double c_sign = (dist > 0 ? 1 : -1);
    f(dir) += (c_sign * this_contact_force);
    if (terms) terms->push_back({(uint8_t)side, (uint8_t)dir, vol_i * c_sign * this_contact_force});

    Vec<D, T> rel_vel_i = vel_i;
    rel_vel_i(dir) -= wall_side_speed;

    if (CN.allow_friction) {
      if (D == 2) {
        unsigned dir_tang = (dir + 1) % 2;
        if (fabs(rel_vel_i(dir_tang)) > 0) {
          double sgn = (rel_vel_i(dir_tang) > 0 ? 1 : -1);
          T friction_force =
              (-CN.friction_coefficient) * fabs(this_contact_force) * sgn;
          if (fabs(rel_vel_i(dir_tang)) <
              CN.friction_coefficient * fabs(this_contact_force) * dt / rho) {
            friction_force = -rel_vel_i(dir_tang) * rho / dt;
          }
          f((dir + 1) % 2) += friction_force;
          if (terms) terms->push_back({(uint8_t)side, (uint8_t)((dir + 1) % 2), vol_i * friction_force});
        }
      } else {
        unsigned t1 = (dir + 1) % 3, t2 = (dir + 2) % 3;
        T vp1 = rel_vel_i(t1), vp2 = rel_vel_i(t2);
        T vpn = sqrt(vp1 * vp1 + vp2 * vp2);
        if (vpn > 0) {
          T u1 = vp1 / vpn, u2 = vp2 / vpn;
          T ff1 = (-CN.friction_coefficient) * fabs(this_contact_force) * u1;
          T ff2 = (-CN.friction_coefficient) * fabs(this_contact_force) * u2;
          if (vpn < CN.friction_coefficient * fabs(this_contact_force) * dt / rho) {
            ff1 = -vpn / rho / dt * u1;
            ff2 = -vpn / rho / dt * u2;
          }
          f(t1) += ff1;
          f(t2) += ff2;
          if (terms) {
            terms->push_back({(uint8_t)side, (uint8_t)t1, vol_i * ff1});
            terms->push_back({(uint8_t)side, (uint8_t)t2, vol_i * ff2});
          }
        }
      }
    }
    if (CN.allow_damping) {
      T damping_coefficient = 2 * CN.damping_ratio *
                              sqrt(CN.normal_stiffness * rho) *
                              sqrt(fabs(c1));
      T damping_force = (-damping_coefficient) * rel_vel_i(dir);
      f(dir) += damping_force;
      if (terms) terms->push_back({(uint8_t)side, (uint8_t)dir, vol_i * damping_force});
    }
  }
}

// update_contact_force_by_boundary (contact.h:667-812).  C and N are the same
// object for self contact.  Pairs are appended to `pairs` (grain-local ids
// offset by the given global offsets) when non-null.
template <unsigned D, class T>
void contact_by_boundary(const Grain<D, T> &P_C, const Grain<D, T> &P_N,
                         const ContactP<T> &CN, double dt, std::vector<Vec<D, T>> &F,
                         bool is_self, std::vector<uint32_t> *pairs,
                         uint32_t off_C, uint32_t off_N) {
  unsigned tn_i = (CN.nl_bdry_only ? (unsigned)P_C.boundary_nodes.size() : P_C.nnodes);
  unsigned tn_j = (CN.nl_bdry_only ? (unsigned)P_N.boundary_nodes.size() : P_N.nnodes);
  for (unsigned p = 0; p < tn_i; ++p) {
    unsigned i = (CN.nl_bdry_only ? P_C.boundary_nodes[p] : p);
    for (unsigned q = 0; q < tn_j; ++q) {
      unsigned j = (CN.nl_bdry_only ? P_N.boundary_nodes[q] : q);
      Vec<D, T> dir = P_N.CurrPos[j] - P_C.CurrPos[i];
      T dir_norm = norm(dir);
      bool do_compute = true;
      double R_c = CN.contact_rad;
      T K_n = CN.normal_stiffness;
      if (is_self) {
        if (i == j) {
          do_compute = false;
        } else {
          const auto &row = P_C.NbdArr[i];
          if (std::find(row.begin(), row.end(), j) != row.end()) {
            do_compute = false;
          } else {
            double ref_dist = sval(norm(P_C.pos[i] - P_C.pos[j]));
            if (ref_dist < CN.contact_rad) {
              R_c = ref_dist;
              K_n = P_C.cnot / ref_dist;
            } else {
              R_c = CN.contact_rad;
              K_n = P_C.cnot / CN.contact_rad;
            }
          }
        }
      }
      if (do_compute && (dir_norm <= R_c)) {
        if (pairs) {
          pairs->push_back(off_C + i);
          pairs->push_back(off_N + j);
        }
        Vec<D, T> unit_dir = dir / dir_norm;
        T contact_rad_contrib = (R_c - dir_norm);
        double N_vol = P_N.vol[j];
        Vec<D, T> f_contact = ((-K_n * contact_rad_contrib) * N_vol) * unit_dir;
        T contact_force_norm = norm(f_contact);
        F[i] += f_contact;
        if (CN.allow_damping) {
          Vec<D, T> rel_v = P_C.vel[i] - P_N.vel[j];
          T rel_v_proj = dot(rel_v, unit_dir);
          Vec<D, T> rel_v_normal_comp = rel_v_proj * unit_dir;
          T damping_coefficient = 2 * CN.damping_ratio *
                                  sqrt(K_n * P_C.rho) *
                                  sqrt(N_vol);
          Vec<D, T> f_damping = (-damping_coefficient) * rel_v_normal_comp;
          F[i] += f_damping;
        }
        if (CN.allow_friction) {
          Vec<D, T> rel_v = P_C.vel[i] - P_N.vel[j];
          Vec<D, T> rel_v_tangential_comp = rel_v - dot(rel_v, unit_dir) * unit_dir;
          T rel_v_tangential_norm = norm(rel_v_tangential_comp);
          if (rel_v_tangential_norm > 0) {
            Vec<D, T> unit_t = rel_v_tangential_comp / norm(rel_v_tangential_comp);
            Vec<D, T> f_friction =
                (-CN.friction_coefficient * contact_force_norm) * unit_t;
            if (rel_v_tangential_norm <
                CN.friction_coefficient * contact_force_norm * dt / P_C.rho) {
              f_friction = (-rel_v_tangential_norm / P_C.rho / dt) * unit_t;
            }
            F[i] += f_friction;
          }
        }
      }
    }
  }
}

// get_minmax (contact.h:1023-1062): 2D [xmin,xmax,ymax,ymin],
// 3D [xmin,ymin,zmin,xmax,ymax,zmax].
template <unsigned D, class T> void get_minmax(const std::vector<Vec<D, T>> &P, double *out) {
  if (D == 2) {
    out[0] = sval(P[0](0)); out[1] = sval(P[0](0)); out[2] = sval(P[0](1)); out[3] = sval(P[0](1));
    for (size_t i = 0; i < P.size(); i++) {
      out[0] = std::min(out[0], sval(P[i](0)));
      out[1] = std::max(out[1], sval(P[i](0)));
      out[2] = std::max(out[2], sval(P[i](1)));
      out[3] = std::min(out[3], sval(P[i](1)));
    }
  } else {
    for (unsigned k = 0; k < 3; ++k) { out[k] = sval(P[0](k)); out[3 + k] = sval(P[0](k)); }
    for (size_t i = 0; i < P.size(); i++) {
      for (unsigned k = 0; k < 3; ++k) {
        out[k] = std::min(out[k], sval(P[i](k)));
        out[3 + k] = std::max(out[3 + k], sval(P[i](k)));
      }
    }
  }
}

// intersects_box_fext (contact.h:983-1020)
template <unsigned D> bool intersects_box_fext(const double *A, const double *B, double c) {
  if (D == 2) {
    double AL = A[0] - c, AR = A[1] + c, AT = A[2] + c, AB = A[3] - c;
    return ((AL <= B[1]) && (AR >= B[0]) && (AT >= B[3]) && (AB <= B[2]));
  } else {
    double ax0 = A[0] - c, ay0 = A[1] - c, az0 = A[2] - c;
    double ax1 = A[3] + c, ay1 = A[4] + c, az1 = A[5] + c;
    return ((ax0 <= B[3]) && (ax1 >= B[0]) && (ay1 >= B[1]) && (ay0 <= B[4]) &&
            (az0 <= B[5]) && (az1 >= B[2]));
  }
}

// within_interior_fext (contact.h:1066-1080): uses the first four entries of
// both arrays whatever the dimension.
bool within_interior_fext(const double *A, const double *B, double c) {
  double AL = A[0] - c, AR = A[1] + c, AT = A[2] + c, AB = A[3] - c;
  return ((B[0] < AL) && (AR < B[1]) && (B[2] > AT) && (B[3] < AB));
}

struct OracleBase {
  virtual ~OracleBase() {}
  virtual int step(uint32_t n, int threads, int record_contacts) = 0;
  virtual int read_nodes(int field, double *dst) = 0;
  virtual int read_wall(double *wall_info, double *reaction) = 0;
  virtual int read_bond_intact(uint8_t *dst) = 0;
  virtual int read_conn(int grain, uint32_t *dst, uint64_t cap, uint64_t *rows) = 0;
  virtual int read_contacts(uint32_t *dst, uint64_t cap, uint64_t *n) = 0;
  virtual int sizes(uint64_t *nn, uint64_t *ng, uint64_t *ns) = 0;
  virtual int read_wall_tangent(double *reaction, double *dreaction) = 0;
};

// Seeds of the tangent mode (CDual<3>): d/d(common scale of every grain's
// cnot, at 1), d/d(normal_stiffness), d/d(friction_coefficient).
inline void seed_param(double v, double &x) { x = v; }
template <int N> inline void seed_param(double v, CDual<N> &x) {
  x = CDual<N>(v);
}
inline void seed_dir(double, int, double &) {}
template <int N> inline void seed_dir(double scale, int dir, CDual<N> &x) {
  if (dir < N) x.d[dir] = scale;
}
inline void deriv(const double &, double *, int) {}
template <int N> inline void deriv(const CDual<N> &x, double *out, int stride) {
  for (int k = 0; k < N; ++k) out[k * stride] = x.d[k];
}

template <unsigned D, class T> struct Oracle : OracleBase {
  using V = Vec<D, T>;
  std::vector<Grain<D, T>> PArr;
  std::vector<uint32_t> node_off;
  ContactP<T> CN;
  pg_params TL;
  bool allow_wall = false;
  double wall[6] = {0, 0, 0, 0, 0, 0};
  double speed[6] = {0, 0, 0, 0, 0, 0};
  T reaction[6][3];
  std::vector<uint8_t> intact;
  uint64_t n_slots = 0;
  uint32_t t_done = 0;
  std::vector<uint32_t> contacts;

  int init(const pg_setup *s, const pg_params *p) {
    TL = *p;
    int G = s->n_grains;
    PArr.resize(G);
    node_off.resize(G + 1);
    uint64_t slot_base = 0;
    for (int g = 0; g < G; ++g) {
      Grain<D, T> &P = PArr[g];
      int64_t a = s->node_offset[g], b = s->node_offset[g + 1];
      unsigned n = (unsigned)(b - a);
      node_off[g] = (uint32_t)a;
      P.nnodes = n;
      P.pos.resize(n); P.disp.resize(n); P.CurrPos.resize(n); P.vel.resize(n);
      P.acc.resize(n); P.force.resize(n); P.extforce.resize(n); P.vol.resize(n);
      P.NbdArr.resize(n); P.total_neighbors.resize(n); P.xi.resize(n);
      P.xi_norm.resize(n); P.stretch.resize(n); P.slot.resize(n);
      for (unsigned i = 0; i < n; ++i) {
        for (unsigned k = 0; k < D; ++k) {
          size_t q = (size_t)(a + i) * D + k;
          P.pos[i](k) = s->pos[q];
          P.disp[i](k) = s->disp[q];
          P.vel[i](k) = s->vel[q];
          P.acc[i](k) = s->acc[q];
          P.extforce[i](k) = s->extforce[q];
          P.force[i](k) = 0.0;
        }
        P.vol[i] = s->vol[a + i];
      }
      for (int64_t r = s->bdry_offset[g]; r < s->bdry_offset[g + 1]; ++r)
        P.boundary_nodes.push_back(s->bdry_nodes[r]);
      for (int64_t r = s->clamped_offset[g]; r < s->clamped_offset[g + 1]; ++r)
        P.clamped_nodes.push_back(s->clamped_nodes[r]);
      // conn2NArr (load_hdf5.h:28-38)
      for (int64_t r = s->conn_offset[g]; r < s->conn_offset[g + 1]; ++r) {
        uint32_t i = s->conn[2 * r], j = s->conn[2 * r + 1];
        if (i >= n || j >= n) { g_err = "oracle: connectivity index out of range"; return PG_ERR_INVALID; }
        P.NbdArr[i].push_back(j);
        P.NbdArr[j].push_back(i);
      }
      for (unsigned i = 0; i < n; ++i) {
        P.slot[i].resize(P.NbdArr[i].size());
        for (size_t k = 0; k < P.NbdArr[i].size(); ++k) P.slot[i][k] = slot_base++;
      }
      P.gen_xi();
      P.delta = s->delta[g]; P.rho = s->rho[g]; P.snot = s->snot[g];
      seed_param(s->cnot[g], P.cnot);
      seed_dir(s->cnot[g], 0, P.cnot);
      P.movable = s->movable[g]; P.breakable = s->breakable[g]; P.stoppable = s->stoppable[g];
      if (p->enable_torque) {
        P.torque_axis = s->torque_axis ? s->torque_axis[g] : 2;
        P.torque_val = s->torque_val ? s->torque_val[g] : 0.0;
      }
      P.break_bonds = true; // simulate2d.cpp:76-78
    }
    node_off[G] = (uint32_t)s->node_offset[G];
    n_slots = slot_base;
    intact.assign(n_slots, 1);
    CN.contact_rad = s->contact_radius;
    seed_param(s->normal_stiffness, CN.normal_stiffness);
    seed_dir(1.0, 1, CN.normal_stiffness);
    seed_param(s->friction_coefficient, CN.friction_coefficient);
    seed_dir(1.0, 2, CN.friction_coefficient);
    CN.damping_ratio = s->damping_ratio;
    CN.allow_friction = p->allow_friction != 0;
    CN.allow_damping = p->allow_damping != 0;
    CN.nl_bdry_only = p->nl_bdry_only != 0;
    CN.self_contact = p->self_contact != 0;
    allow_wall = s->allow_wall != 0;
    for (int k = 0; k < 6; ++k) { wall[k] = s->wall[k]; speed[k] = p->wall_speed[k]; }
    zero_reaction();
    return PG_OK;
  }

  void zero_reaction() {
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 3; ++b) reaction[a][b] = T(0.0);
  }

  // One iteration of the loop body at timeloop.h:608-878.
  void one_step(uint32_t t, int threads, bool rec_contacts) {
    const unsigned G = (unsigned)PArr.size();
    zero_reaction(); // setzero_reaction
    // run_timeloop_worm (timeloop.h:1017-1369) and run_timeloop_compact
    // (:335-498): the same step except for the lines marked "ungated" (both)
    // and "worm"; neither has timed toggles or torque, the worm loop no ramp
    const bool ungated = TL.timeloop_variant != 0;
    const bool worm = TL.timeloop_variant == 1;
    // worm: clamp the front (0, 3) or the back (1, 2) nodes (timeloop.h:1106-1116)
    const int t_num = ((t / 1000) % (2 + 1) == 0) ? 0 : 1;
    if (!ungated && TL.set_movable_index != -1) {
      unsigned pi = (unsigned)TL.set_movable_index, ts = (unsigned)TL.set_movable_timestep;
      if (ts == t && pi < G) PArr[pi].movable = 1;
    }
    if (!ungated && TL.set_stoppable_index != -1) {
      unsigned pi = (unsigned)TL.set_stoppable_index, ts = (unsigned)TL.set_stoppable_timestep;
      if (ts == t && pi < G) PArr[pi].stoppable = 1;
    }
    if (!ungated && TL.reset_partzero_y && G > 1) {
      if (TL.reset_partzero_y_timestep == t) {
        double max_bulk_y = sval(PArr[1].pos[0](1) + PArr[1].disp[1](1));
        for (unsigned i = 1; i < G; ++i)
          for (unsigned j = 0; j < PArr[i].nnodes; j++) {
            double now_y = sval(PArr[i].pos[j](1) + PArr[i].disp[j](1));
            if (now_y > max_bulk_y) max_bulk_y = now_y;
          }
        double dest = max_bulk_y + TL.wheel_rad + CN.contact_rad;
        double to_move_by = dest - sval(PArr[0].mean_CurrPos()(1));
        for (unsigned j = 0; j < PArr[0].nnodes; j++) PArr[0].pos[j](1) += to_move_by;
      }
    }
    const double dt = TL.dt;
#pragma omp parallel for num_threads(threads) if (threads > 1)
    for (unsigned i = 0; i < G; ++i) {
      Grain<D, T> &P = PArr[i];
      if (P.movable || ungated) { // ungated: every grain (timeloop.h:364-366, 1120-1124)
        const double h = (dt * dt * 0.5);
        for (unsigned n = 0; n < P.nnodes; ++n) {
          V inc = (dt * P.vel[n]) + (h * P.acc[n]);
          P.disp[n] += inc;
        }
      }
      for (unsigned n = 0; n < P.nnodes; ++n) P.CurrPos[n] = P.pos[n] + P.disp[n];
    }

    std::vector<double> bbox(G * 6);
    for (unsigned i = 0; i < G; ++i) get_minmax<D>(PArr[i].CurrPos, &bbox[i * 6]);

    std::vector<std::vector<RTerm<T>>> terms(G);
    std::vector<std::vector<uint32_t>> gpairs(rec_contacts ? G : 0);

#pragma omp parallel for schedule(dynamic) num_threads(threads) if (threads > 1)
    for (unsigned i = 0; i < G; ++i) {
      Grain<D, T> &P = PArr[i];
      std::vector<V> ft(P.nnodes, V::zero());
      std::vector<uint32_t> *pairs = rec_contacts ? &gpairs[i] : nullptr;
      if (P.movable) {
        ft = P.get_peridynamic_force();
        if (TL.enable_fracture) {
          if (P.breakable) {
            P.remove_bonds(intact);
            if (CN.self_contact)
              contact_by_boundary<D, T>(P, P, CN, dt, ft, true, pairs, node_off[i], node_off[i]);
          }
        }
        if (P.stoppable || ungated) { // ungated: contact not gated (timeloop.h:404, 1160)
          const double *C = &bbox[i * 6];
          for (unsigned j = 0; j < G; j++) {
            if (j != i) {
              if (intersects_box_fext<D>(C, &bbox[j * 6], CN.contact_rad))
                contact_by_boundary<D, T>(P, PArr[j], CN, dt, ft, false, pairs, node_off[i], node_off[j]);
            }
          }
          if (allow_wall) {
            if (!within_interior_fext(C, wall, CN.contact_rad))
              wall_force(P, dt, ft, &terms[i]);
          }
        }
        if (worm) { // timeloop.h:1183
          const double fac = (double)((t_num + 1) % 2);
          for (unsigned n = 0; n < P.nnodes; ++n) ft[n] += fac * P.extforce[n];
        } else if (TL.gradient_extforce) {
          double extf_sc = (double)t / (double)TL.extforce_maxstep;
          if (extf_sc > 1) extf_sc = 1;
          for (unsigned n = 0; n < P.nnodes; ++n) ft[n] += extf_sc * P.extforce[n];
        } else {
          for (unsigned n = 0; n < P.nnodes; ++n) ft[n] += P.extforce[n];
        }
        if (!ungated && TL.enable_torque) {
          V c = P.mean_CurrPos();
          unsigned taxis = P.torque_axis;
          for (unsigned nn = 0; nn < P.nnodes; nn++) {
            V r = P.CurrPos[nn] - c;
            V rp = V::zero();
            if (D == 3) r((taxis) % D) = T(0.0);
            rp((taxis + 1) % 3) = -r((taxis + 2) % 3);
            rp((taxis + 2) % 3) = r((taxis + 1) % 3);
            T pn = norm(rp);
            if (pn > 0) {
              rp = rp / pn;
            } else {
              rp((taxis + 1) % 3) = T(0.0);
              rp((taxis + 2) % 3) = T(0.0);
            }
            ft[nn] += (P.torque_val * rp);
          }
        }
      }
      P.force = ft;
    }
    for (unsigned i = 0; i < G; ++i)
      for (const RTerm<T> &r : terms[i]) reaction[r.side][r.comp] -= r.value;
    if (rec_contacts) {
      contacts.clear();
      // pairs of grain i are appended in processing order; a stable sort by
      // node gives each node's accumulation order.
      std::vector<std::pair<uint32_t, uint32_t>> all;
      for (unsigned i = 0; i < G; ++i)
        for (size_t k = 0; k < gpairs[i].size(); k += 2)
          all.push_back({gpairs[i][k], gpairs[i][k + 1]});
      std::stable_sort(all.begin(), all.end(),
                       [](const std::pair<uint32_t, uint32_t> &a,
                          const std::pair<uint32_t, uint32_t> &b) { return a.first < b.first; });
      for (auto &pr : all) { contacts.push_back(pr.first); contacts.push_back(pr.second); }
    }

#pragma omp parallel for num_threads(threads) if (threads > 1)
    for (unsigned i = 0; i < G; ++i) {
      Grain<D, T> &P = PArr[i];
      if (P.movable) {
        for (unsigned n = 0; n < P.nnodes; ++n) {
          V temp_acc = P.acc[n];
          P.acc[n] = (1 / P.rho) * P.force[n];
          temp_acc += P.acc[n];
          for (unsigned k = 0; k < D; ++k) temp_acc(k) *= (0.5 * dt);
          P.vel[n] += temp_acc;
        }
        for (unsigned cn = 0; cn < P.clamped_nodes.size(); cn++) {
          unsigned cid = P.clamped_nodes[cn];
          P.disp[cid] = V::zero();
          P.vel[cid] = V::zero();
          P.acc[cid] = V::zero();
        }
        if (worm) { // timeloop.h:1215-1245
          const unsigned held[2][2] = {{0, 3}, {1, 2}};
          for (unsigned k = 0; k < 2; ++k) {
            const unsigned cid = held[t_num][k];
            P.vel[cid] = V::zero();
            P.acc[cid] = V::zero();
          }
          for (unsigned cid = 0; cid < 4; ++cid) {
            P.vel[cid](1) = T(0.0);
            P.acc[cid](1) = T(0.0);
          }
        }
      }
    }
    // wall boundary update (timeloop.h:866-878)
    int ns = (D == 2) ? 4 : 6;
    for (int k = 0; k < ns; ++k) wall[k] += speed[k] * dt;
  }

  // update_wall_contact_force_by_boundary (contact.h:899-964)
  void wall_force(const Grain<D, T> &P, double dt, std::vector<V> &F,
                  std::vector<RTerm<T>> *terms) {
    unsigned tn = (CN.nl_bdry_only ? (unsigned)P.boundary_nodes.size() : P.nnodes);
    for (unsigned p = 0; p < tn; p++) {
      unsigned i = (CN.nl_bdry_only ? P.boundary_nodes[p] : p);
      const V &x = P.CurrPos[i];
      if (D == 2) {
        // sides in lrtb() order: 0 left, 1 right, 2 top, 3 bottom
        T yb = x(1) - wall[3], xl = x(0) - wall[0];
        T yt = x(1) - wall[2], xr = x(0) - wall[1];
        wall_side_force_update<D, T>(yb, CN, 1, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[3], 3, terms);
        wall_side_force_update<D, T>(xl, CN, 0, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[0], 0, terms);
        wall_side_force_update<D, T>(xr, CN, 0, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[1], 1, terms);
        wall_side_force_update<D, T>(yt, CN, 1, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[2], 2, terms);
      } else {
        // sides: 0 x_min 1 y_min 2 z_min 3 x_max 4 y_max 5 z_max
        T d0 = x(0) - wall[0], d1 = x(1) - wall[1], d2 = x(2) - wall[2];
        T d3 = x(0) - wall[3], d4 = x(1) - wall[4], d5 = x(2) - wall[5];
        wall_side_force_update<D, T>(d0, CN, 0, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[0], 0, terms);
        wall_side_force_update<D, T>(d3, CN, 0, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[3], 3, terms);
        wall_side_force_update<D, T>(d1, CN, 1, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[1], 1, terms);
        wall_side_force_update<D, T>(d4, CN, 1, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[4], 4, terms);
        wall_side_force_update<D, T>(d2, CN, 2, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[2], 2, terms);
        wall_side_force_update<D, T>(d5, CN, 2, P.vel[i], P.vol[i], P.rho, F[i], dt, speed[5], 5, terms);
      }
    }
  }

  int step(uint32_t n, int threads, int record_contacts) override {
    for (uint32_t k = 0; k < n; ++k) {
      ++t_done;
      one_step(t_done, threads < 1 ? 1 : threads, record_contacts && k + 1 == n);
    }
    return PG_OK;
  }

  int sizes(uint64_t *nn, uint64_t *ng, uint64_t *ns) override {
    if (nn) *nn = node_off.back();
    if (ng) *ng = PArr.size();
    if (ns) *ns = n_slots;
    return PG_OK;
  }

  int read_nodes(int field, double *dst) override {
    for (size_t g = 0; g < PArr.size(); ++g) {
      const Grain<D, T> &P = PArr[g];
      const std::vector<V> *src = nullptr;
      switch (field) {
        case PG_FIELD_CURRPOS: src = &P.CurrPos; break;
        case PG_FIELD_VEL: src = &P.vel; break;
        case PG_FIELD_ACC: src = &P.acc; break;
        case PG_FIELD_FORCE: src = &P.force; break;
        case PG_FIELD_DISP: src = &P.disp; break;
        default: g_err = "oracle: bad field"; return PG_ERR_INVALID;
      }
      for (unsigned i = 0; i < P.nnodes; ++i)
        for (unsigned k = 0; k < D; ++k) dst[(size_t)(node_off[g] + i) * D + k] = sval((*src)[i](k));
    }
    return PG_OK;
  }

  int read_wall(double *wall_info, double *react) override {
    if (wall_info) for (int k = 0; k < 6; ++k) wall_info[k] = wall[k];
    if (react) {
      // get_reaction() order equals the lrtb() side order used here
      int ns = (D == 2) ? 4 : 6;
      for (int s = 0; s < ns; ++s)
        for (unsigned k = 0; k < D; ++k) react[s * D + k] = sval(reaction[s][k]);
    }
    return PG_OK;
  }

  // Reaction of the last step and its derivatives along the seeded
  // parameters: dreaction[(s*D + k)*3 + q] (zero for the plain oracle).
  int read_wall_tangent(double *react, double *dreact) override {
    const int ns = (D == 2) ? 4 : 6;
    for (int s = 0; s < ns; ++s)
      for (unsigned k = 0; k < D; ++k) {
        if (react) react[s * D + k] = sval(reaction[s][k]);
        if (dreact) {
          for (int q = 0; q < 3; ++q) dreact[(s * D + k) * 3 + q] = 0.0;
          deriv(reaction[s][k], dreact + (s * D + k) * 3, 1);
        }
      }
    return PG_OK;
  }

  int read_bond_intact(uint8_t *dst) override {
    std::memcpy(dst, intact.data(), intact.size());
    return PG_OK;
  }

  // NbdArr2conn (timeloop.h:25-49)
  int read_conn(int grain, uint32_t *dst, uint64_t cap, uint64_t *rows) override {
    if (grain < 0 || grain >= (int)PArr.size()) { g_err = "oracle: bad grain"; return PG_ERR_INVALID; }
    const Grain<D, T> &P = PArr[grain];
    uint64_t r = 0;
    for (unsigned i = 0; i < P.nnodes; i++)
      for (unsigned q : P.NbdArr[i])
        if (i < q) {
          if (dst && r < cap) { dst[2 * r] = i; dst[2 * r + 1] = q; }
          ++r;
        }
    if (rows) *rows = r;
    return PG_OK;
  }

  int read_contacts(uint32_t *dst, uint64_t cap, uint64_t *n) override {
    uint64_t np = contacts.size() / 2;
    if (n) *n = np;
    if (dst) std::memcpy(dst, contacts.data(), std::min(cap, np) * 2 * sizeof(uint32_t));
    return PG_OK;
  }
};

} // namespace

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

int oracle_create(const pg_setup *s, const pg_params *p, void **out) {
  try {
    if (!s || !p || !out) { g_err = "oracle: null argument"; return PG_ERR_INVALID; }
    if (s->dim == 2) {
      auto *o = new Oracle<2, double>();
      int rc = o->init(s, p);
      if (rc) { delete o; return rc; }
      *out = static_cast<OracleBase *>(o);
    } else if (s->dim == 3) {
      auto *o = new Oracle<3, double>();
      int rc = o->init(s, p);
      if (rc) { delete o; return rc; }
      *out = static_cast<OracleBase *>(o);
    } else {
      g_err = "oracle: dim must be 2 or 3";
      return PG_ERR_INVALID;
    }
    return PG_OK;
  } catch (const std::exception &e) {
    g_err = e.what();
    return PG_ERR_ALLOC;
  }
}

// Tangent mode: the same restatement on CDual<3> numbers, seeded along
// (cnot scale, normal_stiffness, friction_coefficient).
int oracle_create_tangent(const pg_setup *s, const pg_params *p, void **out) {
  try {
    if (!s || !p || !out) { g_err = "oracle: null argument"; return PG_ERR_INVALID; }
    if (s->dim == 2) {
      auto *o = new Oracle<2, CDual<3>>();
      int rc = o->init(s, p);
      if (rc) { delete o; return rc; }
      *out = static_cast<OracleBase *>(o);
    } else if (s->dim == 3) {
      auto *o = new Oracle<3, CDual<3>>();
      int rc = o->init(s, p);
      if (rc) { delete o; return rc; }
      *out = static_cast<OracleBase *>(o);
    } else {
      g_err = "oracle: dim must be 2 or 3";
      return PG_ERR_INVALID;
    }
    return PG_OK;
  } catch (const std::exception &e) {
    g_err = e.what();
    return PG_ERR_ALLOC;
  }
}

int oracle_read_wall_tangent(void *h, double *reaction, double *dreaction) {
  return static_cast<OracleBase *>(h)->read_wall_tangent(reaction, dreaction);
}

int oracle_destroy(void *h) {
  delete static_cast<OracleBase *>(h);
  return PG_OK;
}

int oracle_step(void *h, uint32_t n, int threads, int record_contacts) {
  return static_cast<OracleBase *>(h)->step(n, threads, record_contacts);
}
int oracle_sizes(void *h, uint64_t *nn, uint64_t *ng, uint64_t *ns) {
  return static_cast<OracleBase *>(h)->sizes(nn, ng, ns);
}
int oracle_read_nodes(void *h, int field, double *dst) {
  return static_cast<OracleBase *>(h)->read_nodes(field, dst);
}
int oracle_read_wall(void *h, double *wall_info, double *reaction) {
  return static_cast<OracleBase *>(h)->read_wall(wall_info, reaction);
}
int oracle_read_bond_intact(void *h, uint8_t *dst) {
  return static_cast<OracleBase *>(h)->read_bond_intact(dst);
}
int oracle_read_conn(void *h, int grain, uint32_t *dst, uint64_t cap, uint64_t *rows) {
  return static_cast<OracleBase *>(h)->read_conn(grain, dst, cap, rows);
}
int oracle_read_contacts(void *h, uint32_t *dst, uint64_t cap, uint64_t *n) {
  return static_cast<OracleBase *>(h)->read_contacts(dst, cap, n);
}
int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Scalar kernels exposed for known-answer tests.
void oracle_segment_mass(int dim, double R, double d, double *centroid, double *area) {
  if (dim == 2) segment_mass<2, double>(R, d, *centroid, *area);
  else segment_mass<3, double>(R, d, *centroid, *area);
}

} // extern "C"
