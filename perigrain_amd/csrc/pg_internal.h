// pg_internal.h -- device data layout shared by libpg's host code and its
// gfx950 kernels.
//
// Layout in HBM (one handle may pack several independent instances):
//  * nodes are packed instance by instance, grain by grain, in the order of
//    the reference's P_%05u groups, so node n of grain g sits at
//    grain.node_begin + n;
//  * NodeRec (64 B, two copies for the two step parities): the reference
//    position X (pos), nodal volume V and the displacement u (disp) in one
//    aligned half cache line, because the bond and contact gathers need all
//    three for each partner;
//  * vel is double-buffered (partners read the previous step's velocity
//    while their own is being updated); acc and extforce are single arrays;
//  * bonds are stored as sliced ELL with slices of 64 rows (one wavefront):
//    entry k of row i lives at slice_base[i/64] + 64*k + i%64, so a
//    wavefront's loads of bond k are contiguous.  Row order is the
//    reference's NbdArr order (conn2NArr, load_hdf5.h:28-38), which fixes the
//    summation order of the bond force;
//  * contact candidates are a Verlet list built with a skin and kept sorted
//    per node in the order the reference accumulates contact forces.
#ifndef PG_INTERNAL_H
#define PG_INTERNAL_H

#include <stdint.h>

namespace pg {

// Bond slots hold one 32-bit column word: the partner's node index, with the
// top bit set once the bond has broken; padding slots are all ones.
constexpr uint32_t BCOL_BROKEN = 0x80000000u;
constexpr uint32_t BCOL_INDEX = 0x7FFFFFFFu;
constexpr uint32_t BCOL_PAD = 0xFFFFFFFFu;

constexpr int SLICE = 64; // rows per sliced-ELL slice (= wavefront width)
constexpr int BCOL_TAIL = 8; // padding slots past the last slice (bond-loop read-ahead)

// grain flag bits
constexpr int32_t GF_MOVABLE = 1;
constexpr int32_t GF_BREAKABLE = 2;
constexpr int32_t GF_STOPPABLE = 4;
constexpr int32_t GF_YSHIFT = 8; // pos(1) carries a reset_partzero_y shift

// instance flag bits
constexpr int32_t IF_FRACTURE = 1;
constexpr int32_t IF_BDRY_ONLY = 2;
constexpr int32_t IF_DAMPING = 4;
constexpr int32_t IF_FRICTION = 8;
constexpr int32_t IF_SELF_CONTACT = 16;
constexpr int32_t IF_WALL = 32;
constexpr int32_t IF_GRAD_EXT = 64;
constexpr int32_t IF_TORQUE = 128;

// node info bits
constexpr uint8_t NI_CLAMPED = 1;
constexpr uint8_t NI_CONTACT = 2; // takes part in node-node and wall contact

struct alignas(64) NodeRec {
  double X[4]; // pos[0..dim-1]; X[dim] = vol (3D: X[3]; 2D: X[2], X[3] = 0)
  double u[4]; // disp[0..dim-1]
};

struct GrainDev {
  double rho, cnot, snot, yshift;
  int32_t flags;
  int32_t inst;
  uint32_t node_begin, node_end;
  uint32_t torque_axis;
  int32_t pad;
  double torque_val;
};

struct InstDev {
  double dt, h, half_dt; // dt, (dt*dt)*0.5, 0.5*dt
  double Rc, Kn, zeta, mu;
  double speed[6];
  int32_t flags;
  uint32_t extforce_maxstep;
  int32_t set_movable_grain;  // packed grain index or -1
  uint32_t set_movable_step;
  int32_t set_stoppable_grain;
  uint32_t set_stoppable_step;
  uint32_t grain_begin, grain_end;
  uint32_t node_begin, node_end;
  uint32_t ckey_begin, ckey_end; // contact-order keys of this instance
  int32_t nsides;
  int32_t pad;
  double skin_half2; // rebuild when a contact node moved more than this (squared)
};

// Cell grid of one instance for the Verlet-list build.
struct GridDev {
  double origin[3];
  double inv_cell;
  int32_t dims[3];
  uint32_t cell_begin;
  double r2; // squared Verlet-list radius (contact radius + skin, with margin)
};

// Everything the per-step kernel touches (passed by value as kernarg).
struct StepArgs {
  NodeRec *rec[2];
  double *vel[2];
  double *acc;
  const double *ext;
  const uint32_t *node_grain;
  const uint8_t *node_info;
  const uint32_t *ckey;
  const uint32_t *ext2int; // reference node index -> device node index
  const GrainDev *grains;
  const InstDev *insts;
  const uint32_t *slice_base;
  const uint32_t *slice_width;
  uint32_t *bcol; // column words (partner index | BCOL_BROKEN), BCOL_PAD
  const uint32_t *cl_ptr;
  const uint32_t *cl_j;
  const int32_t *cl_slot;
  const double *xb;      // positions at the last list build
  unsigned long long *bbox[3]; // ordered keys, [grains][2*dim] (min..., max...)
  double *wallpos[2];    // [insts][6]
  const double *gmean;   // [grains][3] mean CurrPos (torque)
  uint32_t *violation;   // first step whose positions broke the skin
  double *force_out;     // record steps
  double *wterm;         // record steps: [ckeys][6][4]
  uint8_t *wflag;        // record steps: [ckeys]
  uint32_t n_nodes;
  uint32_t n_grains;
  uint32_t n_insts;
  uint32_t any_torque;
};

// Order-preserving map of an IEEE double onto uint64 (for atomic min/max).
__host__ __device__ inline unsigned long long dkey(double x) {
  union {
    double d;
    unsigned long long u;
  } c;
  c.d = x;
  return (c.u >> 63) ? ~c.u : (c.u | 0x8000000000000000ull);
}
__host__ __device__ inline double dkey_inv(unsigned long long k) {
  union {
    double d;
    unsigned long long u;
  } c;
  c.u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return c.d;
}

} // namespace pg

#endif
