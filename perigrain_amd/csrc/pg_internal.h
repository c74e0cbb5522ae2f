// pg_internal.h -- device data layout shared by libpg's host code and its
// gfx950 kernels.
//
// Layout in HBM (one handle may pack several independent instances):
//  * nodes are packed instance by instance and grain by grain, in the
//    reference's node order unless a degree-sort window is set (see
//    pg_api.hip); perm/inv map between the device order and the reference's;
//  * what a bond gathers from its partner -- reference position X and
//    displacement u, 2*dim doubles, X then u -- is a record split into dim
//    parts of two doubles: part q of node i is elements 2q, 2q+1 at
//    rec[(q*ns + i)*2] (3D: (X0,X1) (X2,u0) (u1,u2); 2D: (X0,X1) (u0,u1)).
//    A gather is dim 16-byte loads at the same byte offset 16*i from the
//    part bases; each load of a wavefront covers the 16-byte slots of 64
//    partners, which share cache lines far more often than 48-byte
//    records do.  The partner volume comes from the grain when all its
//    nodes share one volume (lattice grains), else from V.  Records are
//    double-buffered by step parity (X is stored in both): partners read u
//    of step t while a node writes u of step t+1;
//  * the other node data are structure-of-arrays, component-major with a
//    padded stride `ns` (vel[k*ns + i]), vel double-buffered like u;
//  * bonds are sliced ELL with slices of 64 rows (one wavefront) and rows
//    padded to a multiple of 4 slots; slot k of row i lives at
//    slice_base[i/64] + 256*(k/4) + 4*(i%64) + k%4, so one 16-byte load gives
//    a lane four consecutive slots and a wavefront's load is contiguous.  Row
//    order is the reference's NbdArr order (conn2NArr, load_hdf5.h:28-38),
//    which fixes the bond-force summation order;
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

constexpr int SLICE = 64;     // rows per sliced-ELL slice (= wavefront width)
constexpr int BGROUP = 4;     // slots per lane per column load
constexpr int ADJ_LANES = 4;  // lanes per node in the adjoint step kernel (8 measured no faster)
constexpr int BCOL_TAIL = 512; // padding slots past the last slice (bond-loop read-ahead)

// Slot k of the row in lane `lane` of a slice starting at `base`.
__host__ __device__ inline uint64_t bslot(uint64_t base, uint32_t lane, uint32_t k) {
  return base + (uint64_t)(k >> 2) * (SLICE * BGROUP) + lane * BGROUP + (k & 3);
}

// grain flag bits
constexpr int32_t GF_MOVABLE = 1;
constexpr int32_t GF_BREAKABLE = 2;
constexpr int32_t GF_STOPPABLE = 4;
constexpr int32_t GF_UNIFORM_VOL = 16; // every node of the grain has volume GrainDev::vol

// instance flag bits
constexpr int32_t IF_FRACTURE = 1;
constexpr int32_t IF_BDRY_ONLY = 2;
constexpr int32_t IF_DAMPING = 4;
constexpr int32_t IF_FRICTION = 8;
constexpr int32_t IF_SELF_CONTACT = 16;
constexpr int32_t IF_WALL = 32;
constexpr int32_t IF_GRAD_EXT = 64;
constexpr int32_t IF_TORQUE = 128;
// run_timeloop_worm and run_timeloop_compact (timeloop.h:1017-1369, 335-498):
// every grain advances its position, contact is not gated by stoppable
constexpr int32_t IF_UNGATED = 256;
constexpr int32_t IF_WORM = 512; // the worm loop's phases (timeloop.h:1106-1245)

// node info bits
constexpr uint8_t NI_CLAMPED = 1;
constexpr uint8_t NI_CONTACT = 2; // takes part in node-node and wall contact
constexpr uint8_t NI_WORM_FRONT = 4; // node 0 or 3 of its grain (run_timeloop_worm)
constexpr uint8_t NI_WORM_BACK = 8;  // node 1 or 2 of its grain

struct GrainDev {
  double rho, cnot, snot;
  int32_t flags;
  int32_t inst;
  uint32_t node_begin, node_end;
  uint32_t torque_axis;
  int32_t pad;
  double torque_val;
  double vol; // the common nodal volume when GF_UNIFORM_VOL is set
};

struct InstDev {
  double dt, h, half_dt; // dt, (dt*dt)*0.5, 0.5*dt
  double Rc, Kn, zeta, mu;
  double speed[6];
  int32_t flags;
  uint32_t extforce_maxstep;
  int32_t set_movable_grain; // packed grain index or -1
  uint32_t set_movable_step;
  int32_t set_stoppable_grain;
  uint32_t set_stoppable_step;
  uint32_t grain_begin, grain_end;
  uint32_t node_begin, node_end;
  uint32_t ckey_begin, ckey_end; // contact-order keys of this instance
  int32_t nsides;
  // reset_partzero_y (timeloop.h:629-647): once it happened, the positions
  // `pos` of this packed grain (the instance's first) are X with yshift added
  // to the y component; -1 before.  Bonds keep X.
  int32_t yshift_grain;
  double skin_half2; // rebuild when a contact node moved more than this (squared)
  double yshift;
};

// Cell grid of one instance for the Verlet-list build.
struct GridDev {
  double origin[3];
  double inv_cell;
  int32_t dims[3];
  uint32_t cell_begin;
  double r2; // squared Verlet-list radius (contact radius + skin, with margin)
};

// Element m (0..2*dim-1: X then u) of node i's record in a record array.
__host__ __device__ inline size_t rec_index(uint32_t ns, size_t i, int m) {
  return ((size_t)(m >> 1) * ns + i) * 2 + (m & 1);
}

// Everything the kernels touch (passed by value as kernarg).  Records are
// [dim][ns][2] (parts of X then u, rec_index); the other node arrays are
// [dim][ns].
struct StepArgs {
  double *rec[2];     // [dim][ns][2] pos and disp, by step parity
  const double *V;    // [ns] nodal volumes
  double *vel[2];     // [dim][ns] velocity, by step parity
  double *acc;        // [dim][ns]
  const double *ext;  // [dim][ns] extforce (force density)
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
  const int32_t *cl_slot;     // bond slot of a same-grain partner in the node's row, or -1
  // Entries cl_ptr[i] .. cl_ptr[i] + cl_nfast[i] - 1 of cl_fast are the
  // node's list entries without an intact bond at the list build, in list
  // order, as (partner, partner grain).  A node reads them instead of its
  // full list while none of its bonds has broken since the build (brk_new).
  const uint2 *cl_fast;
  const uint32_t *cl_nfast;
  uint8_t *brk_new;           // [ns] a bond of the node broke since the last list build
  const double *xb;            // [dim][ns] positions at the last list build
  unsigned long long *bbox[3]; // ordered keys, [grains][2*dim] (min..., max...)
  double *wallpos[2];          // [insts][6]
  const double *gmean;         // [grains][3] mean CurrPos (torque)
  uint32_t *violation;         // first step whose positions broke the skin
  const uint32_t *tbase;       // graph replays: base step added to the launch's step argument
  uint32_t *break_step;        // differentiable runs: step at which each slot's bond broke
  // [3]: unice[t % 3] == 0 means every displacement component of u^t is
  // nice() (pg_device.h); step t-1 sets it otherwise, step t-2 clears it
  uint32_t *unice;
  uint32_t xi_ok;              // every X component is nice() and no bond has length zero
  // LDS windows (k_step_lds): block b covers blk_slices slices and reads the
  // records of nodes win_lo[b] .. win_lo[b] + win_cnt[b] - 1
  const uint32_t *win_lo;
  const uint32_t *win_cnt;
  uint32_t blk_slices;
  uint32_t lds_blocks, lds_threads, lds_bytes;
  double *force_out;           // [dim][ns], record steps
  double *wterm;               // record steps: [ckeys][6][4]
  uint8_t *wflag;              // record steps: [ckeys]
  uint32_t ns;                 // node stride (n_nodes rounded up)
  uint32_t n_nodes;
  uint32_t n_grains;
  uint32_t n_insts;
  uint32_t any_torque;
  uint32_t uniform_vol; // every grain has GF_UNIFORM_VOL
  uint32_t ghost;       // index of the ghost record (partner of broken and padding slots)
};

// Adjoint of one time step (differentiable runs).  State of step t as the
// forward pass saw it, the cotangents carried backwards, and the parameter
// partial sums.  Node vectors are [dim][ns] unless noted.
struct AdjArgs {
  StepArgs A;                  // static data; A.cl_* is the segment's candidate list
  const double *su;            // records holding u^t
  const double *sv;            // v^{t-1}
  const double *swall;         // [insts][6] wall positions of step t
  const unsigned long long *sbox; // [grains][2*dim] grain boxes of step t
  const double *lR;            // [insts][6][dim] cotangent of the wall reaction of step t
  const double *lf_in;         // cotangent of the force density of step t
  double *lf_out;              // ... of step t-1 (written)
  double *Pu, *Pv, *Pa;        // cotangents of u^{t+1}, v^t, a^t (partial), updated to step t-1
  double *gpart;               // [blocks][3]: d/d(c scale), d/dKn, d/dmu partial sums
};

// Zero bytes after the last wall flag: the wall-reaction kernel reads the
// flags in whole chunks of this many keys.
constexpr size_t kWflagSpare = 4096;

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
