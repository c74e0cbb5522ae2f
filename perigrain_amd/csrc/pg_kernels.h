// pg_kernels.h -- host-side launchers of the gfx950 kernels in pg_kernels.hip.
#ifndef PG_KERNELS_H
#define PG_KERNELS_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pg_internal.h"

namespace pg {

hipError_t launch_step(int dim, bool record, const StepArgs &A, uint32_t t, int block,
                       hipStream_t s);
hipError_t set_step_lds_limit(int dim, uint32_t bytes);
hipError_t launch_prologue(int dim, const StepArgs &A, hipStream_t s);
// Up to four copies of 8-byte words (unused entries: n = 0) in one launch.
struct CopySet {
  unsigned long long *dst[4];
  const unsigned long long *src[4];
  size_t n[4];
};
hipError_t launch_copy4(const CopySet &cs, hipStream_t s);
// Per-grain post-processing of step t for grains g0 .. g0+ng-1 (pg_grain_stats).
hipError_t launch_grain_stats(int dim, const StepArgs &A, uint32_t t, uint32_t g0,
                              uint32_t ng, double *out, hipStream_t s);
// Grain boxes of step t from its positions, into a cleared buffer.
hipError_t launch_boxes(int dim, const StepArgs &A, uint32_t t, unsigned long long *keys,
                        hipStream_t s);
hipError_t launch_grain_mean(int dim, const StepArgs &A, uint32_t t, double *gmean,
                             hipStream_t s);
// Wall reaction of recorded steps (k_wall_compact, then k_reaction).  Step
// y (0 .. nsteps-1) reads ring slot (first + y) % slots of the wall terms
// (wstride doubles apart) and flags (fstride bytes apart, a multiple of 16)
// and writes reaction + y * rstride; hits (hstride = nchunks * 4096 keys per
// step) and chunk_cnt (cstride per step) are its scratch.
struct ReactionArgs {
  const InstDev *insts;
  uint32_t n_insts;
  uint32_t nchunks; // ceil(contact keys / 4096)
  const double *wterm;
  const uint8_t *wflag;
  double *reaction;
  uint32_t *hits;
  uint32_t *chunk_cnt;
  uint32_t first, slots;
  size_t wstride, fstride, rstride, hstride, cstride;
};
hipError_t launch_reactions(int dim, const ReactionArgs &R, uint32_t nsteps, hipStream_t s);
hipError_t launch_fill_box(unsigned long long *p, size_t n, int dim, hipStream_t s);
hipError_t launch_inst_box(int dim, const StepArgs &A, uint32_t t, unsigned long long *ibox,
                           hipStream_t s);
hipError_t launch_cell_count(int dim, const StepArgs &A, uint32_t t, const GridDev *grids,
                             uint32_t *cell_of, uint32_t *cell_cnt, hipStream_t s);
hipError_t launch_cell_fill(uint32_t n, const uint32_t *cell_of, const uint32_t *cell_start,
                            uint32_t *cell_cur, uint32_t *cell_nodes, hipStream_t s);
hipError_t launch_list(int dim, bool fill, const StepArgs &A, uint32_t t, const GridDev *grids,
                       const uint32_t *cell_start, const uint32_t *cell_nodes, uint32_t *cnt,
                       const uint32_t *ptr, uint32_t *lj, int32_t *lslot, uint2 *lfast,
                       uint32_t *nfast, double *xb,
                       hipStream_t s);
hipError_t launch_pairs(int dim, bool fill, const StepArgs &A, uint32_t t, uint32_t *cnt,
                        const uint32_t *ptr, uint32_t *out, hipStream_t s);
// adjoint (pg_adjoint.hip)
hipError_t launch_adj_step(int dim, const AdjArgs &B, uint32_t t, uint32_t blocks, hipStream_t s);
hipError_t launch_adj_maxdisp(int dim, const StepArgs &A, const double *su, uint32_t steps,
                              unsigned long long *out, hipStream_t s);
hipError_t launch_adj_restore_bcol(uint32_t *bcol, const uint32_t *break_step, uint64_t n,
                                   uint32_t t, hipStream_t s);
hipError_t launch_adj_init_break(const uint32_t *bcol, uint32_t *break_step, uint64_t n,
                                 hipStream_t s);
hipError_t launch_bond_geom(int dim, bool check, uint64_t n, const double *xi, const double *eta,
                            double *s, double *e, hipStream_t st);
hipError_t launch_arith(uint64_t n, const double *a, const double *b, double *q, double *r,
                        hipStream_t s);
hipError_t launch_chain_check(uint32_t nchains, const double *ent, const uint64_t *eoff,
                              const int32_t *two, const double *s0, uint32_t zero, double *wave,
                              double *seq, hipStream_t s);
// hipcub exclusive sum; call with tmp=nullptr to size *tmp_bytes.
hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, void *tmp,
                              size_t *tmp_bytes, hipStream_t s);

} // namespace pg

#endif
