// pg_api.hip -- host implementation of the libpg C ABI (include/perigrain.h).
//
// A handle owns the device state of one or more packed instances and drives
// one fused step kernel per time step on its own HIP stream.  The contact
// neighbour list is a Verlet list with a skin: each step kernel checks how
// far contact nodes moved since the list was built and, the first time one
// moved more than half the skin, records that step; that kernel and all
// later ones of the same batch return at once without touching state.  The
// host checks the record after each batch of launches, rebuilds the list
// from the positions of the first unexecuted step and relaunches from there,
// so no step ever runs with a stale list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "perigrain.h"
#include "pg_internal.h"
#include "pg_kernels.h"

using namespace pg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// Body of an exported entry point: no C++ exception leaves the C ABI; an
// exception becomes an error code and message like any other failure.
template <class F> int guarded(F &&f) {
  try {
    return f();
  } catch (const std::bad_alloc &) {
    return fail(PG_ERR_ALLOC, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(PG_ERR_INVALID, std::string("unexpected error: ") + e.what());
  } catch (...) {
    return fail(PG_ERR_INVALID, "unexpected error");
  }
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(PG_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Host staging of one node (reference position, volume, displacement).
struct HostRec {
  double X[4];
  double u[4];
};

struct InstHost {
  uint32_t node_begin, node_end, grain_begin, grain_end;
  uint64_t slot_begin, slot_end;
  uint32_t ckey_begin, ckey_end;
  double Rc, skin;
  bool contact;
  int nsides;
  // reset_partzero_y: the step at which the first grain is moved (0: never),
  // the config's wheel_rad, and whether it happened
  uint32_t reset_step = 0;
  double wheel_rad = 0.0;
  bool reset_done = false;
};

} // namespace

// Staging array of pg_grain_stats_noalloc: the per-grain results go here
// instead of into a buffer allocated at call time.  Each device holds its own
// copy; host code never passes the array's name to a kernel or a copy (on the
// host it is only a stand-in), but the device address that
// hipGetSymbolAddress returns on that device, kept in g_scratch_dev.
static __device__ __align__(256) double scratch[PG_NOALLOC_STATS_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};
} // namespace

struct pg_handle {
  int dim = 0, device = 0, block = 256;
  hipStream_t stream = nullptr;
  uint32_t N = 0, G = 0, I = 0, S = 0, CK = 0;
  uint64_t E = 0;
  uint32_t t = 0;
  uint32_t check_every = 64;
  double skin_factor = 0.5;
  bool any_torque = false, any_contact = false;
  std::vector<InstHost> inst;
  std::vector<InstDev> inst_dev;
  std::vector<uint32_t> deg, slice_base, slice_width, bcol, grain_node_begin;
  std::vector<uint64_t> slot0;
  std::vector<uint32_t> perm, inv; // device index <-> reference index
  // nodes are degree-sorted within windows of this many nodes of a grain
  // (0: reference order, which measured fastest: it keeps the gathers local)
  uint32_t sort_window = 0;
  bool uniform_vol = false;
  std::vector<void *> allocs;
  // device
  uint32_t ns = 0; // node stride of the component-major arrays
  double *rec[2] = {nullptr, nullptr}, *V = nullptr; // [ns][2*dim] records, [ns] volumes
  double *vel[2] = {nullptr, nullptr};
  double *acc = nullptr, *ext = nullptr, *xb = nullptr, *gmean = nullptr;
  uint32_t *node_grain = nullptr, *ckey = nullptr, *ext2int = nullptr;
  uint8_t *node_info = nullptr;
  GrainDev *grains = nullptr;
  InstDev *insts = nullptr;
  uint32_t *d_slice_base = nullptr, *d_slice_width = nullptr, *d_bcol = nullptr;
  uint32_t *cl_ptr = nullptr, *cl_j = nullptr, *cl_nfast = nullptr;
  int32_t *cl_slot = nullptr;
  uint2 *cl_fast = nullptr;
  uint8_t *brk_new = nullptr;
  size_t cl_cap = 0;
  unsigned long long *bbox[3] = {nullptr, nullptr, nullptr};
  double *wallpos[2] = {nullptr, nullptr};
  uint32_t *violation = nullptr;
  // niceness flags of u by step (StepArgs::unice) and whether X and the bond
  // lengths allow the unchecked bond loop (StepArgs::xi_ok)
  uint32_t *unice = nullptr;
  uint32_t unice_host[3] = {1, 1, 0};
  bool xi_ok = false;
  double *force_out = nullptr, *wterm = nullptr, *reaction = nullptr;
  uint8_t *wflag = nullptr;
  // wall-reaction scratch: compacted hit keys and per-chunk counts, one set
  // for the plain recording step (rhits, rcnt) and one per ring slot of a
  // recorded run (rhits_ring, rcnt_ring)
  uint32_t *rhits = nullptr, *rcnt = nullptr, *rhits_ring = nullptr, *rcnt_ring = nullptr;
  // list build scratch
  GridDev *grids = nullptr;
  unsigned long long *ibox = nullptr;
  uint32_t *cell_of = nullptr, *cell_cnt = nullptr, *cell_start = nullptr, *cell_cur = nullptr,
           *cell_nodes = nullptr, *cl_cnt = nullptr;
  size_t cell_cap = 0;
  void *scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  uint32_t *h_pin = nullptr;
  // graph replay of whole chunks of step launches (pg_set_graph)
  bool use_graph = false;
  uint32_t *tbase = nullptr; // device-side base step of a replayed chunk
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint64_t args_gen = 0, graph_gen = 0; // graph is valid while graph_gen == args_gen
  hipEvent_t ev[8] = {};
  hipEvent_t kev[2] = {}; // step-kernel timing (pg_kernel_time)
  double kernel_ms = 0.0;
  uint64_t kernel_launches = 0;
  uint64_t rebuilds = 0, skipped = 0, list_entries = 0;
  // differentiable runs (pg_adjoint_*)
  bool any_clamped = false;
  double *hist = nullptr; // while recording: [steps][insts][18] wall reaction of every step
  uint32_t hist_t0 = 0;
  // while recording: wall terms and flags of the last check_every steps, one
  // ring slot per step, so that one launch sums a whole chunk's reactions
  double *wring = nullptr;
  uint8_t *fring = nullptr;
  double *d_stats = nullptr; // pg_grain_stats output
  size_t stats_cap = 0;
  struct Adj {
    bool recorded = false;
    // the recorded pass kept every step's state in the segment slots (one
    // segment: checkpoint_every >= nsteps), so the backward pass recomputes
    // nothing
    bool stored = false;
    uint32_t t0 = 0, T = 0, K = 0, nck = 0;
    uint32_t *break_step = nullptr; // [E + tail] step at which each slot's bond broke
    double *hist = nullptr;         // [T][18]
    // checkpoint k: state after step t0 + k*K (the last one: after t0 + T)
    double *ck_rec = nullptr, *ck_vel = nullptr, *ck_acc = nullptr, *ck_wall = nullptr;
    unsigned long long *ck_box = nullptr;
    // snapshots of one segment, one per step
    double *su = nullptr, *sv = nullptr, *swall = nullptr;
    unsigned long long *sbox = nullptr;
    double *Pu = nullptr, *Pv = nullptr, *Pa = nullptr, *lf[2] = {nullptr, nullptr};
    double *gpart = nullptr, *lR = nullptr;
    unsigned long long *dmax = nullptr;
    // what the readers see after the recorded pass: u^T, force, reaction
    double *fin_rec = nullptr, *fin_force = nullptr, *fin_reaction = nullptr;
    unsigned long long *fin_box = nullptr;
    size_t cap_ck = 0, cap_seg = 0, cap_T = 0;
  } adj;
  // LDS windows of the step kernel (k_step_lds); off when a window would not fit
  bool lds_on = false;
  uint32_t lds_slices = 0, lds_blocks = 0, lds_threads = 0, lds_bytes = 0;
  uint32_t *d_win_lo = nullptr, *d_win_cnt = nullptr;
  // host copy of the bond table, valid for step bcol_host_t
  std::vector<uint32_t> bcol_host;
  uint32_t bcol_host_t = 0xFFFFFFFFu;
  // snapshot slots (pg_snapshot_*): device copies of a step's snapshot
  // state, read back on rstream by another host thread
  struct Snap {
    bool alloc = false, taken = false;
    uint32_t t = 0;
    double *rec = nullptr, *vel = nullptr, *acc = nullptr, *force = nullptr;
    double *wall = nullptr, *reaction = nullptr;
    uint32_t *bcol = nullptr;
    hipEvent_t ev = nullptr;
    std::vector<int32_t> yshift_grain; // per instance, as at the take
    std::vector<double> yshift;
    std::vector<uint32_t> bcol_host; // filled by the first Connectivity read
    bool bcol_ready = false;
  } snap[PG_SNAPSHOT_SLOTS];
  hipStream_t rstream = nullptr;
  StepArgs A{};
};

namespace {

// chunks of the wall-reaction compaction for ck contact keys
size_t react_chunks(size_t ck) { return (ck + kWflagSpare - 1) / kWflagSpare; }

template <class T> int dalloc(pg_handle *h, T *&p, size_t n) {
  if (n == 0) n = 1;
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, n * sizeof(T));
  if (e != hipSuccess)
    return fail(PG_ERR_ALLOC, std::string("hipMalloc ") + std::to_string(n * sizeof(T)) +
                                  " bytes: " + hipGetErrorString(e));
  h->allocs.push_back(q);
  p = static_cast<T *>(q);
  return PG_OK;
}

template <class T> int dfree(pg_handle *h, T *&p) {
  if (!p) return PG_OK;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), (void *)p);
  if (it != h->allocs.end()) h->allocs.erase(it);
  hipError_t e = hipFree(p);
  p = nullptr;
  if (e != hipSuccess) return fail(PG_ERR_DEVICE, std::string("hipFree: ") + hipGetErrorString(e));
  return PG_OK;
}

template <class T> int h2d(pg_handle *h, T *dst, const T *src, size_t n) {
  if (n == 0) return PG_OK;
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
  return PG_OK;
}
template <class T> int d2h_on(hipStream_t st, T *dst, const T *src, size_t n) {
  if (n == 0) return PG_OK;
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return PG_OK;
}
template <class T> int d2h(pg_handle *h, T *dst, const T *src, size_t n) {
  return d2h_on(h->stream, dst, src, n);
}

#define RC(x)                  \
  do {                         \
    int rc_ = (x);             \
    if (rc_ != PG_OK) return rc_; \
  } while (0)

void fill_args(pg_handle *h) {
  StepArgs &A = h->A;
  h->args_gen++;
  A.tbase = nullptr;
  A.rec[0] = h->rec[0];
  A.rec[1] = h->rec[1];
  A.V = h->V;
  A.ns = h->ns;
  A.vel[0] = h->vel[0];
  A.vel[1] = h->vel[1];
  A.acc = h->acc;
  A.ext = h->ext;
  A.node_grain = h->node_grain;
  A.node_info = h->node_info;
  A.ckey = h->ckey;
  A.ext2int = h->ext2int;
  A.grains = h->grains;
  A.insts = h->insts;
  A.slice_base = h->d_slice_base;
  A.slice_width = h->d_slice_width;
  A.bcol = h->d_bcol;
  A.cl_ptr = h->cl_ptr;
  A.cl_j = h->cl_j;
  A.cl_slot = h->cl_slot;
  A.cl_fast = h->cl_fast;
  A.cl_nfast = h->cl_nfast;
  A.brk_new = h->brk_new;
  A.xb = h->xb;
  for (int k = 0; k < 3; ++k) A.bbox[k] = h->bbox[k];
  A.wallpos[0] = h->wallpos[0];
  A.wallpos[1] = h->wallpos[1];
  A.gmean = h->gmean;
  A.violation = h->violation;
  A.unice = h->unice;
  A.xi_ok = h->xi_ok ? 1 : 0;
  A.force_out = h->force_out;
  A.wterm = h->wterm;
  A.wflag = h->wflag;
  A.n_nodes = h->N;
  A.n_grains = h->G;
  A.n_insts = h->I;
  A.any_torque = h->any_torque ? 1 : 0;
  A.uniform_vol = h->uniform_vol ? 1 : 0;
  A.ghost = h->N;
  A.break_step = h->adj.break_step;
  A.win_lo = h->lds_on ? h->d_win_lo : nullptr;
  A.win_cnt = h->lds_on ? h->d_win_cnt : nullptr;
  A.blk_slices = h->lds_slices;
  A.lds_blocks = h->lds_blocks;
  A.lds_threads = h->lds_threads;
  A.lds_bytes = h->lds_bytes;
}

int scan(pg_handle *h, const uint32_t *in, uint32_t *out, size_t n) {
  size_t need = 0;
  HIPCHK(exclusive_scan_u32(in, out, n, nullptr, &need, h->stream));
  if (need > h->scan_tmp_bytes) {
    if (h->scan_tmp) {
      auto it = std::find(h->allocs.begin(), h->allocs.end(), h->scan_tmp);
      if (it != h->allocs.end()) h->allocs.erase(it);
      HIPCHK(hipFree(h->scan_tmp));
      h->scan_tmp = nullptr;
    }
    char *p = nullptr;
    RC(dalloc(h, p, need));
    h->scan_tmp = p;
    h->scan_tmp_bytes = need;
  }
  size_t bytes = h->scan_tmp_bytes;
  HIPCHK(exclusive_scan_u32(in, out, n, h->scan_tmp, &bytes, h->stream));
  return PG_OK;
}

// Build the contact Verlet list from the positions of step ts (u^ts in
// rec[ts & 1], or in rec_override when given).
int build_list(pg_handle *h, uint32_t ts, const double *rec_override = nullptr) {
  const int D = h->dim;
  auto args = [&]() {
    StepArgs A2 = h->A;
    if (rec_override) A2.rec[ts & 1] = const_cast<double *>(rec_override);
    return A2;
  };
  if (!h->any_contact) {
    HIPCHK(hipMemsetAsync(h->cl_ptr, 0, (h->N + 1) * sizeof(uint32_t), h->stream));
    HIPCHK(hipMemsetAsync(h->cl_nfast, 0, h->N * sizeof(uint32_t), h->stream));
    HIPCHK(hipMemsetAsync(h->brk_new, 0, h->N, h->stream));
    HIPCHK(hipMemsetAsync(h->violation, 0xFF, sizeof(uint32_t), h->stream));
    return PG_OK;
  }
  // 1. per-instance box of contact nodes
  HIPCHK(launch_fill_box(h->ibox, (size_t)h->I * 2 * D, D, h->stream));
  HIPCHK(launch_inst_box(D, args(), ts, h->ibox, h->stream));
  std::vector<unsigned long long> kb((size_t)h->I * 2 * D);
  RC(d2h(h, kb.data(), h->ibox, kb.size()));
  // 2. per-instance grids
  std::vector<GridDev> gd(h->I);
  uint64_t cells = 0;
  for (uint32_t s = 0; s < h->I; ++s) {
    const InstHost &ih = h->inst[s];
    GridDev &g = gd[s];
    std::memset(&g, 0, sizeof(g));
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool empty = !ih.contact;
    for (int k = 0; k < D; ++k) {
      lo[k] = dkey_inv(kb[(size_t)s * 2 * D + k]);
      hi[k] = dkey_inv(kb[(size_t)s * 2 * D + D + k]);
      if (!(lo[k] <= hi[k]) || !std::isfinite(lo[k]) || !std::isfinite(hi[k])) empty = true;
    }
    const double rs = (ih.Rc + ih.skin) * (1.0 + 1e-9);
    g.r2 = rs * rs;
    double L = rs;
    const uint32_t ncontact = ih.node_end - ih.node_begin;
    const double cap = std::max<double>(4096.0, 4.0 * ncontact);
    int dims[3] = {1, 1, 1};
    if (!empty && rs > 0) {
      for (int it = 0; it < 64; ++it) {
        double tot = 1;
        for (int k = 0; k < D; ++k) {
          double n = std::floor((hi[k] - lo[k]) / L) + 1;
          if (n < 1) n = 1;
          dims[k] = (int)std::min(n, 1e9);
          tot *= n;
        }
        if (tot <= cap) break;
        L *= (D == 3) ? std::cbrt(tot / cap) * 1.01 : std::sqrt(tot / cap) * 1.01;
      }
    }
    for (int k = 0; k < 3; ++k) g.origin[k] = (k < D && !empty) ? lo[k] : 0.0;
    g.inv_cell = (L > 0 && std::isfinite(L)) ? 1.0 / L : 0.0;
    for (int k = 0; k < 3; ++k) g.dims[k] = k < D ? dims[k] : 1;
    g.cell_begin = (uint32_t)cells;
    cells += (uint64_t)g.dims[0] * g.dims[1] * g.dims[2];
  }
  if (cells + 1 >= 0xFFFFFFFFull) return fail(PG_ERR_INVALID, "cell grid too large");
  RC(h2d(h, h->grids, gd.data(), gd.size()));
  if (cells + 1 > h->cell_cap) {
    RC(dfree(h, h->cell_cnt));
    RC(dfree(h, h->cell_start));
    RC(dfree(h, h->cell_cur));
    h->cell_cap = (size_t)((cells + 1) * 1.25) + 64;
    RC(dalloc(h, h->cell_cnt, h->cell_cap));
    RC(dalloc(h, h->cell_start, h->cell_cap));
    RC(dalloc(h, h->cell_cur, h->cell_cap));
  }
  // 3. counting sort of contact nodes into cells
  HIPCHK(hipMemsetAsync(h->cell_cnt, 0, (cells + 1) * sizeof(uint32_t), h->stream));
  HIPCHK(launch_cell_count(D, args(), ts, h->grids, h->cell_of, h->cell_cnt, h->stream));
  RC(scan(h, h->cell_cnt, h->cell_start, cells + 1));
  HIPCHK(hipMemsetAsync(h->cell_cur, 0, (cells + 1) * sizeof(uint32_t), h->stream));
  HIPCHK(launch_cell_fill(h->N, h->cell_of, h->cell_start, h->cell_cur, h->cell_nodes, h->stream));
  // 4. count, scan, fill the per-node lists
  HIPCHK(hipMemsetAsync(h->cl_cnt, 0, (h->N + 1) * sizeof(uint32_t), h->stream));
  HIPCHK(launch_list(D, false, args(), ts, h->grids, h->cell_start, h->cell_nodes, h->cl_cnt,
                     nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, h->stream));
  RC(scan(h, h->cl_cnt, h->cl_ptr, (size_t)h->N + 1));
  uint32_t total = 0;
  RC(d2h(h, &total, h->cl_ptr + h->N, 1));
  if (total > h->cl_cap) {
    RC(dfree(h, h->cl_j));
    RC(dfree(h, h->cl_slot));
    RC(dfree(h, h->cl_fast));
    h->cl_cap = (size_t)(total * 1.5) + 1024;
    RC(dalloc(h, h->cl_j, h->cl_cap));
    RC(dalloc(h, h->cl_slot, h->cl_cap));
    RC(dalloc(h, h->cl_fast, h->cl_cap));
    fill_args(h);
  }
  HIPCHK(launch_list(D, true, args(), ts, h->grids, h->cell_start, h->cell_nodes, nullptr,
                     h->cl_ptr, h->cl_j, h->cl_slot, h->cl_fast, h->cl_nfast, h->xb,
                     h->stream));
  // the fast lists hold the bond states of the build: no bond broken since
  HIPCHK(hipMemsetAsync(h->brk_new, 0, h->N, h->stream));
  HIPCHK(hipMemsetAsync(h->violation, 0xFF, sizeof(uint32_t), h->stream));
  h->list_entries = total;
  h->rebuilds += 1;
  return PG_OK;
}

// Bond table of the current step on the host (one copy per step, shared by
// the per-grain readers).
int bond_status(pg_handle *h, const std::vector<uint32_t> *&st) {
  if (h->bcol_host_t != h->t || h->bcol_host.size() != h->E) {
    h->bcol_host.resize(h->E);
    HIPCHK(hipSetDevice(h->device));
    RC(d2h(h, h->bcol_host.data(), h->d_bcol, h->E));
    h->bcol_host_t = h->t;
  }
  st = &h->bcol_host;
  return PG_OK;
}

// Niceness flags for a run that continues with step t+1 from state that
// no step kernel checked: u^{t+1} unknown (per-bond tests), the flag of
// u^{t+2} cleared for step t+1 to set.
int set_unice(pg_handle *h, uint32_t t) {
  h->unice_host[(t + 1) % 3] = 1;
  h->unice_host[(t + 2) % 3] = 0;
  h->unice_host[t % 3] = 1;
  HIPCHK(hipMemcpyAsync(h->unice, h->unice_host, sizeof(h->unice_host), hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream)); // the host array is reused
  return PG_OK;
}

int check_setup(const pg_setup *s, const pg_params *p, int dim_expected) {
  if (!s || !p) return fail(PG_ERR_INVALID, "null setup or params");
  if (s->dim != 2 && s->dim != 3) return fail(PG_ERR_INVALID, "dim must be 2 or 3");
  if (dim_expected && s->dim != dim_expected)
    return fail(PG_ERR_INVALID, "all instances must share one dimension");
  if (s->n_grains < 1) return fail(PG_ERR_INVALID, "n_grains must be >= 1");
  if (!s->node_offset || !s->pos || !s->vol || !s->disp || !s->vel || !s->acc ||
      !s->extforce || !s->conn_offset || !s->bdry_offset || !s->clamped_offset || !s->delta ||
      !s->rho || !s->cnot || !s->snot || !s->movable || !s->breakable || !s->stoppable)
    return fail(PG_ERR_INVALID, "missing setup array");
  if (s->node_offset[0] != 0 || s->conn_offset[0] != 0 || s->bdry_offset[0] != 0 ||
      s->clamped_offset[0] != 0)
    return fail(PG_ERR_INVALID, "offset arrays must start at 0");
  for (int g = 0; g < s->n_grains; ++g) {
    if (s->node_offset[g + 1] <= s->node_offset[g])
      return fail(PG_ERR_INVALID, "grain " + std::to_string(g) + " has no nodes");
    if (s->conn_offset[g + 1] < s->conn_offset[g] || s->bdry_offset[g + 1] < s->bdry_offset[g] ||
        s->clamped_offset[g + 1] < s->clamped_offset[g])
      return fail(PG_ERR_INVALID, "offset arrays must be non-decreasing");
    if (p->enable_torque) {
      if (!s->torque_axis || !s->torque_val)
        return fail(PG_ERR_INVALID, "enable_torque needs torque_axis and torque_val");
      if (s->torque_axis[g] > 2 || (s->dim == 2 && s->torque_axis[g] != 2))
        return fail(PG_ERR_INVALID, "torque_axis must be 2 in 2D and < 3 in 3D");
    }
  }
  if (s->node_offset[s->n_grains] >= 0x7FFFFF00ll)
    return fail(PG_ERR_INVALID, "too many nodes");
  if (!(p->dt > 0)) return fail(PG_ERR_INVALID, "dt must be > 0");
  if (p->reset_partzero_y && p->reset_partzero_y_timestep > 0 &&
      (s->n_grains < 2 || s->node_offset[2] - s->node_offset[1] < 2))
    return fail(PG_ERR_INVALID,
                "reset_partzero_y needs a second grain with at least two nodes (timeloop.h:632)");
  auto chk_idx = [&](int32_t idx, int32_t ts) -> bool {
    return idx == -1 || ts <= 0 || (idx >= 0 && idx < s->n_grains);
  };
  if (!chk_idx(p->set_movable_index, p->set_movable_timestep) ||
      !chk_idx(p->set_stoppable_index, p->set_stoppable_timestep))
    return fail(PG_ERR_INVALID, "set_movable_index/set_stoppable_index out of range");
  if (p->timeloop_variant < PG_TIMELOOP_STANDARD || p->timeloop_variant > PG_TIMELOOP_COMPACT)
    return fail(PG_ERR_INVALID, "timeloop_variant must be 0, 1 or 2");
  return PG_OK;
}

int create_impl(const pg_setup *setups, const pg_params *params, int n, int device,
                pg_handle *h) {
  if (n < 1) return fail(PG_ERR_INVALID, "need at least one instance");
  const int D = setups[0].dim;
  for (int s = 0; s < n; ++s) RC(check_setup(&setups[s], &params[s], D));
  h->dim = D;
  h->device = device;
  {
    // PG_SORT_WINDOW=n overrides the degree-sort window (0: reference order)
    const char *w = std::getenv("PG_SORT_WINDOW");
    if (w && *w) h->sort_window = (uint32_t)std::strtoul(w, nullptr, 10);
  }

  // ---- sizes and per-instance ranges (host-side validation and staging
  // come first, so a bad setup is reported before any device work)
  uint64_t N = 0, G = 0, CK = 0, slots = 0;
  h->inst.resize(n);
  for (int s = 0; s < n; ++s) {
    const pg_setup &st = setups[s];
    InstHost &ih = h->inst[s];
    ih.node_begin = (uint32_t)N;
    ih.grain_begin = (uint32_t)G;
    N += (uint64_t)st.node_offset[st.n_grains];
    G += (uint64_t)st.n_grains;
    ih.node_end = (uint32_t)N;
    ih.grain_end = (uint32_t)G;
    ih.Rc = st.contact_radius;
    ih.contact = st.contact_radius > 0;
    ih.skin = h->skin_factor * (st.contact_radius > 0 ? st.contact_radius : 0.0);
    ih.nsides = D == 2 ? 4 : 6;
    const pg_params &pp = params[s];
    ih.reset_step = (pp.reset_partzero_y && pp.timeloop_variant == PG_TIMELOOP_STANDARD)
                        ? pp.reset_partzero_y_timestep : 0;
    ih.wheel_rad = pp.wheel_rad;
  }
  // record gathers use 32-bit byte offsets of 16 bytes per node
  if (N >= (1ull << 28) - 64) return fail(PG_ERR_INVALID, "too many nodes in batch (limit 2^28 - 64)");
  h->N = (uint32_t)N;
  h->G = (uint32_t)G;
  h->I = (uint32_t)n;

  // ---- host staging of node data
  std::vector<HostRec> rec(N);
  std::vector<double> vel(N * D), acc(N * D), ext(N * D);
  std::vector<uint32_t> node_grain(N), ckey(N, 0xFFFFFFFFu);
  std::vector<uint8_t> node_info(N, 0);
  std::vector<GrainDev> grains(G);
  h->inst_dev.resize(n);
  h->deg.assign(N, 0);
  h->grain_node_begin.resize(G + 1);
  std::vector<uint64_t> row_ptr(N + 1, 0);

  for (int s = 0; s < n; ++s) {
    const pg_setup &st = setups[s];
    const pg_params &pp = params[s];
    InstHost &ih = h->inst[s];
    const bool bdry_only = pp.nl_bdry_only != 0;
    // nodes
    for (int g = 0; g < st.n_grains; ++g) {
      const uint32_t gg = ih.grain_begin + g;
      const int64_t a = st.node_offset[g], b = st.node_offset[g + 1];
      h->grain_node_begin[gg] = ih.node_begin + (uint32_t)a;
      for (int64_t q = a; q < b; ++q) {
        const uint64_t i = ih.node_begin + q;
        HostRec &r = rec[i];
        std::memset(&r, 0, sizeof(r));
        for (int k = 0; k < D; ++k) {
          r.X[k] = st.pos[q * D + k];
          r.u[k] = st.disp[q * D + k];
          vel[i * D + k] = st.vel[q * D + k];
          acc[i * D + k] = st.acc[q * D + k];
          ext[i * D + k] = st.extforce[q * D + k];
        }
        r.X[D] = st.vol[q];
        node_grain[i] = gg;
        if (!bdry_only && ih.contact) node_info[i] |= NI_CONTACT;
      }
      if (pp.timeloop_variant == PG_TIMELOOP_WORM) { // nodes 0, 3 (front), 1, 2 (back)
        if (b - a < 4) return fail(PG_ERR_INVALID, "the worm loop needs four nodes in every grain");
        node_info[ih.node_begin + a + 0] |= NI_WORM_FRONT;
        node_info[ih.node_begin + a + 3] |= NI_WORM_FRONT;
        node_info[ih.node_begin + a + 1] |= NI_WORM_BACK;
        node_info[ih.node_begin + a + 2] |= NI_WORM_BACK;
      }
      for (int64_t c = st.clamped_offset[g]; c < st.clamped_offset[g + 1]; ++c) {
        const uint32_t cn = st.clamped_nodes[c];
        if (cn >= (uint64_t)(b - a)) return fail(PG_ERR_INVALID, "clamped node out of range");
        node_info[ih.node_begin + a + cn] |= NI_CLAMPED;
        h->any_clamped = true;
      }
      GrainDev &gd = grains[gg];
      std::memset(&gd, 0, sizeof(gd));
      gd.rho = st.rho[g];
      gd.cnot = st.cnot[g];
      gd.snot = st.snot[g];
      gd.flags = (st.movable[g] ? GF_MOVABLE : 0) | (st.breakable[g] ? GF_BREAKABLE : 0) |
                 (st.stoppable[g] ? GF_STOPPABLE : 0);
      gd.inst = s;
      gd.node_begin = ih.node_begin + (uint32_t)a;
      gd.node_end = ih.node_begin + (uint32_t)b;
      gd.torque_axis = (pp.enable_torque && st.torque_axis) ? st.torque_axis[g] : 2;
      gd.torque_val = (pp.enable_torque && st.torque_val) ? st.torque_val[g] : 0.0;
    }
    // contact order keys: node order, or boundary-list order (contact.h:672-681)
    ih.ckey_begin = (uint32_t)CK;
    if (!bdry_only) {
      for (uint32_t i = ih.node_begin; i < ih.node_end; ++i) ckey[i] = (uint32_t)(CK++);
    } else {
      for (int g = 0; g < st.n_grains; ++g) {
        const int64_t a = st.node_offset[g], nn = st.node_offset[g + 1] - a;
        for (int64_t c = st.bdry_offset[g]; c < st.bdry_offset[g + 1]; ++c) {
          const uint32_t bn = st.bdry_nodes[c];
          if (bn >= (uint64_t)nn) return fail(PG_ERR_INVALID, "boundary node out of range");
          const uint64_t i = ih.node_begin + a + bn;
          if (ckey[i] != 0xFFFFFFFFu)
            return fail(PG_ERR_INVALID, "duplicate entry in bdry_nodes of grain " + std::to_string(g));
          ckey[i] = (uint32_t)(CK++);
          if (ih.contact) node_info[i] |= NI_CONTACT;
        }
      }
    }
    ih.ckey_end = (uint32_t)CK;
    // degrees (conn2NArr, load_hdf5.h:28-38)
    for (int g = 0; g < st.n_grains; ++g) {
      const int64_t a = st.node_offset[g], nn = st.node_offset[g + 1] - a;
      for (int64_t r = st.conn_offset[g]; r < st.conn_offset[g + 1]; ++r) {
        const uint32_t ci = st.conn[2 * r], cj = st.conn[2 * r + 1];
        if (ci >= (uint64_t)nn || cj >= (uint64_t)nn)
          return fail(PG_ERR_INVALID, "Connectivity index out of range in grain " + std::to_string(g));
        if (ci == cj) return fail(PG_ERR_INVALID, "Connectivity has a self bond");
        h->deg[ih.node_begin + a + ci]++;
        h->deg[ih.node_begin + a + cj]++;
      }
    }
    // instance constants
    InstDev &id = h->inst_dev[s];
    std::memset(&id, 0, sizeof(id));
    id.dt = pp.dt;
    id.h = (pp.dt * pp.dt * 0.5);
    id.half_dt = (0.5 * pp.dt);
    id.Rc = st.contact_radius;
    id.Kn = st.normal_stiffness;
    id.zeta = st.damping_ratio;
    id.mu = st.friction_coefficient;
    for (int k = 0; k < 6; ++k) id.speed[k] = pp.wall_speed[k];
    id.flags = (pp.enable_fracture ? IF_FRACTURE : 0) | (bdry_only ? IF_BDRY_ONLY : 0) |
               (pp.allow_damping ? IF_DAMPING : 0) | (pp.allow_friction ? IF_FRICTION : 0) |
               (pp.self_contact ? IF_SELF_CONTACT : 0) | (st.allow_wall ? IF_WALL : 0) |
               (pp.gradient_extforce ? IF_GRAD_EXT : 0) | (pp.enable_torque ? IF_TORQUE : 0);
    // the worm and compact loops have no torque or timed toggles; the worm
    // loop has no extforce ramp either
    const bool worm = pp.timeloop_variant != PG_TIMELOOP_STANDARD;
    if (worm) id.flags = (id.flags & ~IF_TORQUE) | IF_UNGATED;
    if (pp.timeloop_variant == PG_TIMELOOP_WORM) id.flags = (id.flags & ~IF_GRAD_EXT) | IF_WORM;
    id.extforce_maxstep = pp.extforce_maxstep;
    id.set_movable_grain = (!worm && pp.set_movable_index >= 0 && pp.set_movable_timestep > 0)
                               ? (int32_t)(ih.grain_begin + pp.set_movable_index) : -1;
    id.set_movable_step = (uint32_t)std::max(pp.set_movable_timestep, 0);
    id.set_stoppable_grain = (!worm && pp.set_stoppable_index >= 0 && pp.set_stoppable_timestep > 0)
                                 ? (int32_t)(ih.grain_begin + pp.set_stoppable_index) : -1;
    id.set_stoppable_step = (uint32_t)std::max(pp.set_stoppable_timestep, 0);
    id.grain_begin = ih.grain_begin;
    id.grain_end = ih.grain_end;
    id.node_begin = ih.node_begin;
    id.node_end = ih.node_end;
    id.ckey_begin = ih.ckey_begin;
    id.ckey_end = ih.ckey_end;
    id.nsides = ih.nsides;
    id.yshift_grain = -1;
    id.yshift = 0.0;
    const double sh = 0.5 * ih.skin * (1.0 - 1e-9);
    id.skin_half2 = sh * sh;
    if (pp.enable_torque && !worm) h->any_torque = true;
    if (ih.contact) h->any_contact = true;
  }
  h->grain_node_begin[G] = (uint32_t)N;
  h->CK = (uint32_t)CK;
  // grains whose nodes all carry bit-identical volumes (lattice grains) let
  // the bond loop take the partner volume from the grain instead of a gather
  h->uniform_vol = true;
  for (uint32_t gg = 0; gg < G; ++gg) {
    const uint32_t a = h->grain_node_begin[gg], b = h->grain_node_begin[gg + 1];
    bool same = true;
    for (uint32_t q = a + 1; q < b && same; ++q)
      same = std::memcmp(&rec[q].X[D], &rec[a].X[D], sizeof(double)) == 0;
    if (same) {
      grains[gg].flags |= GF_UNIFORM_VOL;
      grains[gg].vol = rec[a].X[D];
    } else {
      h->uniform_vol = false;
    }
  }

  // ---- adjacency rows in the reference's NbdArr order
  for (uint64_t i = 0; i < N; ++i) row_ptr[i + 1] = row_ptr[i] + h->deg[i];
  slots = row_ptr[N];
  h->slot0.assign(row_ptr.begin(), row_ptr.end());
  std::vector<uint32_t> rows(slots);
  {
    std::vector<uint64_t> cur(row_ptr.begin(), row_ptr.end() - 1);
    for (int s = 0; s < n; ++s) {
      const pg_setup &st = setups[s];
      const InstHost &ih = h->inst[s];
      for (int g = 0; g < st.n_grains; ++g) {
        const uint32_t base = ih.node_begin + (uint32_t)st.node_offset[g];
        for (int64_t r = st.conn_offset[g]; r < st.conn_offset[g + 1]; ++r) {
          const uint32_t ci = base + st.conn[2 * r], cj = base + st.conn[2 * r + 1];
          rows[cur[ci]++] = cj;
          rows[cur[cj]++] = ci;
        }
      }
    }
  }
  {
    std::vector<uint32_t> tmp;
    for (uint64_t i = 0; i < N; ++i) {
      tmp.assign(rows.begin() + row_ptr[i], rows.begin() + row_ptr[i + 1]);
      std::sort(tmp.begin(), tmp.end());
      if (std::adjacent_find(tmp.begin(), tmp.end()) != tmp.end())
        return fail(PG_ERR_INVALID, "Connectivity lists a bond twice (node " + std::to_string(i) + ")");
    }
  }
  {
    // every coordinate of X nice (zero or 2^-240 <= |x| <= 2^240, see
    // nice() in pg_device.h) and no bond of length zero: then the bond loop
    // needs no per-bond range test while u is nice too
    auto nice_h = [](double x) {
      const double m = std::fabs(x);
      return m == 0.0 || (m >= 0x1p-240 && m <= 0x1p240);
    };
    bool ok = true;
    for (uint64_t i = 0; i < N && ok; ++i)
      for (int k = 0; k < D; ++k) ok = ok && nice_h(rec[i].X[k]);
    for (uint64_t i = 0; i < N && ok; ++i)
      for (uint64_t r = row_ptr[i]; r < row_ptr[i + 1] && ok; ++r) {
        bool same = true;
        for (int k = 0; k < D; ++k) same = same && rec[rows[r]].X[k] == rec[i].X[k];
        ok = !same;
      }
    h->xi_ok = ok;
  }
  for (int s = 0; s < n; ++s) {
    h->inst[s].slot_begin = row_ptr[h->inst[s].node_begin];
    h->inst[s].slot_end = row_ptr[h->inst[s].node_end];
  }
  // ---- device node order: within each grain, nodes sorted by degree
  // (descending, stable) inside windows of sort_window nodes, so the 64 rows
  // of a slice have nearly equal lengths while neighbouring nodes stay close
  // in memory.  perm maps device index -> reference index, inv the reverse.
  // Everything order-sensitive (row order, contact order keys, the torque
  // mean) keeps the reference order.
  std::vector<uint32_t> &perm = h->perm, &inv = h->inv;
  perm.resize(N);
  inv.resize(N);
  for (uint32_t gg = 0; gg < G; ++gg) {
    const uint32_t a = h->grain_node_begin[gg], b = h->grain_node_begin[gg + 1];
    for (uint32_t q = a; q < b; ++q) perm[q] = q;
    if (h->sort_window > 1)
      for (uint32_t w0 = a; w0 < b; w0 += h->sort_window)
        std::stable_sort(perm.begin() + w0, perm.begin() + std::min(b, w0 + h->sort_window),
                         [&](uint32_t x, uint32_t y) { return h->deg[x] > h->deg[y]; });
  }
  for (uint64_t i = 0; i < N; ++i) inv[perm[i]] = (uint32_t)i;
  auto permute = [&](auto &v, size_t w) {
    auto out = v;
    for (uint64_t i = 0; i < N; ++i)
      for (size_t k = 0; k < w; ++k) out[i * w + k] = v[(uint64_t)perm[i] * w + k];
    v.swap(out);
  };
  permute(rec, 1);
  permute(vel, D);
  permute(acc, D);
  permute(ext, D);
  permute(node_grain, 1);
  permute(node_info, 1);
  permute(ckey, 1);

  // ---- sliced ELL over the device order
  const uint32_t S = (uint32_t)((N + SLICE - 1) / SLICE);
  h->S = S;
  h->slice_base.assign(S, 0);
  h->slice_width.assign(S, 0);
  uint64_t E = 0;
  for (uint32_t sl = 0; sl < S; ++sl) {
    uint32_t w = 0;
    for (uint64_t i = (uint64_t)sl * SLICE; i < std::min<uint64_t>(N, (uint64_t)(sl + 1) * SLICE); ++i)
      w = std::max(w, h->deg[perm[i]]);
    w = (w + BGROUP - 1) / BGROUP * BGROUP; // whole slot groups
    h->slice_base[sl] = (uint32_t)E;
    h->slice_width[sl] = w;
    E += (uint64_t)w * SLICE;
    if (E >= 0xFFFFFFFFull) return fail(PG_ERR_INVALID, "bond table exceeds 2^32 entries");
  }
  h->E = E;
  h->bcol.assign(E, 0);
  std::vector<uint32_t> bcol_init(E + BCOL_TAIL, BCOL_PAD);
  for (uint64_t i = 0; i < N; ++i) {
    const uint32_t sl = (uint32_t)(i / SLICE), lane = (uint32_t)(i % SLICE);
    const uint32_t r = perm[i];
    for (uint32_t k = 0; k < h->deg[r]; ++k) {
      const uint64_t e = bslot(h->slice_base[sl], lane, k);
      h->bcol[e] = inv[rows[row_ptr[r] + k]];
      bcol_init[e] = h->bcol[e];
    }
  }

  // ---- device allocations
  HIPCHK(hipSetDevice(device));

  // ---- LDS windows: blocks of S consecutive slices (S chosen so that the
  // blocks just cover the CUs once when they can), each with the node range
  // holding all bond partners of its nodes
  std::vector<uint32_t> win_lo, win_cnt;
  {
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (cus < 1) cus = 1;
    // PG_LDS=1 selects the LDS kernel; otherwise the gathering one runs. The
    // gathering kernel measured faster at every size, also for small handles
    // (at most one 64-node slice per CU: 17.9 against 18.5 us per step on the
    // two-disk case, 15.0 against 15.7 on a 15k-node 2D compression).
    const char *env = std::getenv("PG_LDS");
    const bool force = env && env[0] == '1';
    const bool want = force;
    constexpr uint32_t kLdsLimit = 160u * 1024u; // gfx950: 160 KiB per CU
    const uint32_t rec_bytes = 2u * D * 8u;
    uint32_t spb = std::max<uint32_t>(1, (h->S + (uint32_t)cus - 1) / (uint32_t)cus);
    if (!force && spb > 16) spb = 16; // one slice per wavefront; more blocks
    spb = std::min<uint32_t>(spb, 32);
    for (; want && spb >= 1; spb = (spb > 1) ? (spb + 1) / 2 : 0) {
      const uint32_t nb = (h->S + spb - 1) / spb;
      win_lo.assign(nb, 0);
      win_cnt.assign(nb, 0);
      uint64_t worst = 0;
      for (uint32_t b = 0; b < nb; ++b) {
        const uint64_t i0 = (uint64_t)b * spb * SLICE, i1 = std::min<uint64_t>(N, (uint64_t)(b + 1) * spb * SLICE);
        uint64_t lo = i0, hi = i1 > i0 ? i1 - 1 : i0;
        for (uint32_t sl = b * spb; sl < std::min<uint32_t>(h->S, (b + 1) * spb); ++sl)
          for (uint64_t e = h->slice_base[sl]; e < (uint64_t)h->slice_base[sl] + (uint64_t)h->slice_width[sl] * SLICE; ++e) {
            const uint32_t w = bcol_init[e];
            if (w == BCOL_PAD) continue;
            const uint64_t j = w & BCOL_INDEX;
            lo = std::min(lo, j);
            hi = std::max(hi, j);
          }
        win_lo[b] = (uint32_t)lo;
        win_cnt[b] = (uint32_t)(hi - lo + 1);
        worst = std::max<uint64_t>(worst, (uint64_t)(hi - lo + 2) * rec_bytes); // + ghost
      }
      // blocks resident per CU (LDS and 16 wavefronts of 128 VGPRs), rounds
      const uint32_t waves = std::min<uint32_t>(spb, 16);
      const uint64_t per_cu = std::max<uint64_t>(
          1, std::min<uint64_t>(worst ? kLdsLimit / worst : 16, 16 / waves));
      const double rounds = (double)nb / ((double)cus * (double)per_cu);
      const double waste = rounds > 1.0 ? std::ceil(rounds) / rounds : 1.0;
      const bool balanced = waste <= 1.10 && spb <= 16;
      if (worst <= kLdsLimit && (force || balanced)) {
        h->lds_on = true;
        h->lds_slices = spb;
        h->lds_blocks = nb;
        h->lds_threads = std::min<uint32_t>(spb, 16) * SLICE;
        h->lds_bytes = (uint32_t)worst;
        break;
      }
      if (spb == 1 || (worst <= kLdsLimit && !balanced)) break;
    }
    if (h->lds_on) HIPCHK(set_step_lds_limit(D, h->lds_bytes));
  }
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  for (int k = 0; k < 8; ++k) HIPCHK(hipEventCreate(&h->ev[k]));
  for (int k = 0; k < 2; ++k) HIPCHK(hipEventCreate(&h->kev[k]));
  // one spare record at index N at least: the ghost partner of broken and
  // padding bond slots
  h->ns = (uint32_t)(((N + 1 + SLICE - 1) / SLICE) * SLICE);
  if (h->ns == 0) h->ns = SLICE;
  const size_t ND = (size_t)h->ns * D; // component-major [D][ns]
  RC(dalloc(h, h->rec[0], 2 * ND));
  RC(dalloc(h, h->rec[1], 2 * ND));
  RC(dalloc(h, h->V, h->ns));
  RC(dalloc(h, h->vel[0], ND));
  RC(dalloc(h, h->vel[1], ND));
  RC(dalloc(h, h->acc, ND));
  RC(dalloc(h, h->ext, ND));
  RC(dalloc(h, h->xb, ND));
  RC(dalloc(h, h->force_out, ND));
  RC(dalloc(h, h->node_grain, N));
  RC(dalloc(h, h->node_info, N));
  RC(dalloc(h, h->ckey, N));
  RC(dalloc(h, h->ext2int, N));
  RC(dalloc(h, h->grains, G));
  RC(dalloc(h, h->insts, n));
  RC(dalloc(h, h->d_slice_base, S));
  RC(dalloc(h, h->d_slice_width, S));
  RC(dalloc(h, h->d_bcol, E + BCOL_TAIL));
  if (h->lds_on) {
    RC(dalloc(h, h->d_win_lo, win_lo.size()));
    RC(dalloc(h, h->d_win_cnt, win_cnt.size()));
  }
  RC(dalloc(h, h->cl_ptr, (size_t)N + 1));
  RC(dalloc(h, h->cl_cnt, (size_t)N + 1));
  RC(dalloc(h, h->cl_nfast, (size_t)N));
  RC(dalloc(h, h->brk_new, (size_t)N));
  h->cl_cap = 1024;
  RC(dalloc(h, h->cl_j, h->cl_cap));
  RC(dalloc(h, h->cl_slot, h->cl_cap));
  RC(dalloc(h, h->cl_fast, h->cl_cap));
  for (int k = 0; k < 3; ++k) RC(dalloc(h, h->bbox[k], (size_t)G * 2 * D));
  RC(dalloc(h, h->wallpos[0], (size_t)n * 6));
  RC(dalloc(h, h->wallpos[1], (size_t)n * 6));
  RC(dalloc(h, h->gmean, h->any_torque ? (size_t)G * 3 : 1));
  RC(dalloc(h, h->violation, 1));
  RC(dalloc(h, h->unice, 3));
  RC(dalloc(h, h->tbase, 1));
  RC(dalloc(h, h->wterm, (size_t)(CK ? CK : 1) * 24));
  RC(dalloc(h, h->wflag, (size_t)CK + kWflagSpare)); // zero spare: the reaction reads whole chunks
  RC(dalloc(h, h->rhits, react_chunks(CK) * kWflagSpare));
  RC(dalloc(h, h->rcnt, react_chunks(CK)));
  RC(dalloc(h, h->reaction, (size_t)n * 18));
  RC(dalloc(h, h->grids, n));
  RC(dalloc(h, h->ibox, (size_t)n * 2 * D));
  RC(dalloc(h, h->cell_of, N));
  RC(dalloc(h, h->cell_nodes, N));
  {
    void *p = nullptr;
    HIPCHK(hipHostMalloc(&p, 64 * sizeof(uint32_t), hipHostMallocDefault));
    h->h_pin = static_cast<uint32_t *>(p);
  }

  // ---- uploads (records in parts, the rest transposed to component-major)
  {
    const uint32_t nsd = h->ns;
    std::vector<double> sR(2 * ND, 0.0), sV(nsd, 0.0), sVel(ND, 0.0), sAcc(ND, 0.0),
        sExt(ND, 0.0);
    // the ghost: far from every node (1e30 in each coordinate), volume 1
    for (int k = 0; k < D; ++k) sR[rec_index(nsd, N, k)] = 1e30;
    sV[N] = 1.0;
    for (uint64_t i = 0; i < N; ++i) {
      sV[i] = rec[i].X[D];
      for (int k = 0; k < D; ++k) {
        sR[rec_index(nsd, i, k)] = rec[i].X[k];
        sR[rec_index(nsd, i, D + k)] = rec[i].u[k];
        sVel[(size_t)k * nsd + i] = vel[i * D + k];
        sAcc[(size_t)k * nsd + i] = acc[i * D + k];
        sExt[(size_t)k * nsd + i] = ext[i * D + k];
      }
    }
    RC(h2d(h, h->rec[0], sR.data(), 2 * ND));
    RC(h2d(h, h->rec[1], sR.data(), 2 * ND));
    RC(h2d(h, h->V, sV.data(), (size_t)nsd));
    RC(h2d(h, h->vel[0], sVel.data(), ND));
    RC(h2d(h, h->vel[1], sVel.data(), ND));
    RC(h2d(h, h->acc, sAcc.data(), ND));
    RC(h2d(h, h->ext, sExt.data(), ND));
    HIPCHK(hipStreamSynchronize(h->stream)); // staging buffers go out of scope
  }
  RC(h2d(h, h->node_grain, node_grain.data(), N));
  RC(h2d(h, h->node_info, node_info.data(), N));
  RC(h2d(h, h->ckey, ckey.data(), N));
  RC(h2d(h, h->ext2int, h->inv.data(), N));
  RC(h2d(h, h->grains, grains.data(), G));
  RC(h2d(h, h->insts, h->inst_dev.data(), (size_t)n));
  RC(h2d(h, h->d_slice_base, h->slice_base.data(), S));
  RC(h2d(h, h->d_slice_width, h->slice_width.data(), S));
  RC(h2d(h, h->d_bcol, bcol_init.data(), bcol_init.size()));
  if (h->lds_on) {
    RC(h2d(h, h->d_win_lo, win_lo.data(), win_lo.size()));
    RC(h2d(h, h->d_win_cnt, win_cnt.size() ? win_cnt.data() : nullptr, win_cnt.size()));
  }
  std::vector<double> walls((size_t)n * 6);
  for (int s = 0; s < n; ++s)
    for (int k = 0; k < 6; ++k) walls[(size_t)s * 6 + k] = setups[s].wall[k];
  RC(h2d(h, h->wallpos[0], walls.data(), walls.size()));
  RC(h2d(h, h->wallpos[1], walls.data(), walls.size()));
  HIPCHK(hipMemsetAsync(h->force_out, 0, ND * sizeof(double), h->stream));
  HIPCHK(hipMemsetAsync(h->xb, 0, ND * sizeof(double), h->stream));
  HIPCHK(hipMemsetAsync(h->reaction, 0, (size_t)n * 18 * sizeof(double), h->stream));
  HIPCHK(hipMemsetAsync(h->wflag, 0, (size_t)CK + kWflagSpare, h->stream));
  HIPCHK(hipMemsetAsync(h->wterm, 0, (size_t)(CK ? CK : 1) * 24 * sizeof(double), h->stream));
  HIPCHK(hipMemsetAsync(h->cl_ptr, 0, ((size_t)N + 1) * sizeof(uint32_t), h->stream));
  HIPCHK(hipMemsetAsync(h->gmean, 0, (h->any_torque ? (size_t)G * 3 : 1) * sizeof(double), h->stream));
  for (int k = 0; k < 3; ++k) HIPCHK(launch_fill_box(h->bbox[k], (size_t)G * 2 * D, D, h->stream));
  HIPCHK(hipMemsetAsync(h->violation, 0xFF, sizeof(uint32_t), h->stream));
  RC(set_unice(h, 0));
  fill_args(h);

  // ---- position update of step 1 and the first neighbour list
  HIPCHK(launch_prologue(D, h->A, h->stream));
  RC(build_list(h, 1));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->t = 0;
  return PG_OK;
}

// Teardown in reverse order of dependence: wait for the stream's work; drop
// the captured graph, whose kernel nodes hold pointers into the handle's
// buffers, while those buffers still exist; destroy the stream before the
// events recorded on it and the pinned buffer its copies used; free the
// device memory last.
void destroy_impl(pg_handle *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  if (h->stream) (void)hipStreamDestroy(h->stream);
  h->stream = nullptr;
  if (h->rstream) {
    (void)hipStreamSynchronize(h->rstream);
    (void)hipStreamDestroy(h->rstream);
    h->rstream = nullptr;
  }
  for (auto &sn : h->snap)
    if (sn.ev) (void)hipEventDestroy(sn.ev);
  for (int k = 0; k < 8; ++k)
    if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
  for (int k = 0; k < 2; ++k)
    if (h->kev[k]) (void)hipEventDestroy(h->kev[k]);
  if (h->h_pin) (void)hipHostFree(h->h_pin);
  h->h_pin = nullptr;
  for (void *p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  delete h;
}

// Capture one chunk of check_every step launches (no recording step) whose
// step index is the device-side base plus 1..check_every.
int ensure_graph(pg_handle *h) {
  if (h->gexec && h->graph_gen == h->args_gen) return PG_OK;
  if (h->gexec) HIPCHK(hipGraphExecDestroy(h->gexec));
  if (h->graph) HIPCHK(hipGraphDestroy(h->graph));
  h->gexec = nullptr;
  h->graph = nullptr;
  StepArgs G = h->A;
  G.tbase = h->tbase;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (uint32_t k = 1; k <= h->check_every && e == hipSuccess; ++k) {
    if (h->any_torque) e = launch_grain_mean(h->dim, G, k, h->gmean, h->stream);
    if (e == hipSuccess) e = launch_step(h->dim, false, G, k, h->block, h->stream);
  }
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(h->stream, &g);
  if (e != hipSuccess) return fail(PG_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(PG_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
  h->graph = g;
  HIPCHK(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
  h->graph_gen = h->args_gen;
  return PG_OK;
}

size_t rec_len(const pg_handle *h) { return (size_t)h->ns * 2 * h->dim; }
size_t vec_len(const pg_handle *h) { return (size_t)h->ns * h->dim; }
size_t box_len(const pg_handle *h) { return (size_t)h->G * 2 * h->dim; }
size_t wring_stride(const pg_handle *h) { return (size_t)(h->CK ? h->CK : 1) * 24; }
size_t fring_stride(const pg_handle *h) { return ((size_t)h->CK + kWflagSpare + 15) / 16 * 16; }

// Wall reaction of the steps whose terms and flags lie in ring slots
// first .. first + nsteps - 1 (modulo slots), into out (rstride per step).
int reactions(pg_handle *h, const double *wterm, const uint8_t *wflag, uint32_t *hits,
              uint32_t *cnt, uint32_t nsteps, uint32_t first, uint32_t slots, double *out,
              size_t rstride) {
  ReactionArgs R{};
  R.insts = h->insts;
  R.n_insts = h->I;
  R.nchunks = (uint32_t)react_chunks(h->CK);
  R.wterm = wterm;
  R.wflag = wflag;
  R.reaction = out;
  R.hits = hits;
  R.chunk_cnt = cnt;
  R.first = first;
  R.slots = slots;
  R.wstride = wring_stride(h);
  R.fstride = fring_stride(h);
  R.rstride = rstride;
  R.hstride = (size_t)R.nchunks * kWflagSpare;
  R.cstride = R.nchunks;
  HIPCHK(launch_reactions(h->dim, R, nsteps, h->stream));
  return PG_OK;
}

// reset_partzero_y of instance s before step tr (timeloop.h:629-647): the
// largest y among the other grains' nodes of pos + disp of step tr-1
// (starting, as the reference does, from pos of node 0 plus disp of node 1 of
// the second grain) plus wheel_rad plus the contact radius is where the
// first grain's mean CurrPos y goes; the difference is added to that grain's
// pos y.  Computed on the host from the device state, once per run.
int do_reset(pg_handle *h, uint32_t s, uint32_t tr) {
  const int D = h->dim;
  InstHost &ih = h->inst[s];
  const size_t n = ih.node_end - ih.node_begin;
  // record parts holding element 1 (X y) and element D+1 (u y, of step tr-1)
  const int mu = D + 1;
  std::vector<double> px(2 * n), pu(2 * n);
  RC(d2h(h, px.data(), h->rec[0] + rec_index(h->ns, ih.node_begin, 0), 2 * n));
  RC(d2h(h, pu.data(), h->rec[(tr - 1) & 1] + rec_index(h->ns, ih.node_begin, mu & ~1), 2 * n));
  auto dev = [&](uint32_t e) -> size_t { return h->inv[e] - ih.node_begin; };
  auto y = [&](size_t a_pos, size_t a_disp) { return px[2 * a_pos + 1] + pu[2 * a_disp + (mu & 1)]; };
  const uint32_t g0 = ih.grain_begin, g1 = ih.grain_begin + 1;
  double max_y = y(dev(h->grain_node_begin[g1]), dev(h->grain_node_begin[g1] + 1));
  for (uint32_t e = h->grain_node_begin[g1]; e < h->grain_node_begin[ih.grain_end]; ++e) {
    const double v = y(dev(e), dev(e));
    if (v > max_y) max_y = v;
  }
  double mean = 0.0;
  for (uint32_t e = h->grain_node_begin[g0]; e < h->grain_node_begin[g0 + 1]; ++e)
    mean += y(dev(e), dev(e));
  mean = mean / (double)(h->grain_node_begin[g0 + 1] - h->grain_node_begin[g0]);
  const double dest = max_y + ih.wheel_rad + ih.Rc;
  InstDev &id = h->inst_dev[s];
  id.yshift_grain = (int32_t)g0;
  id.yshift = dest - mean;
  RC(h2d(h, h->insts + s, &id, 1));
  ih.reset_done = true;
  return PG_OK;
}

// First step after t at which some instance's reset_partzero_y is due, or 0.
uint32_t next_reset(const pg_handle *h) {
  uint32_t r = 0;
  for (const InstHost &ih : h->inst)
    if (ih.reset_step > h->t && !ih.reset_done && (r == 0 || ih.reset_step < r)) r = ih.reset_step;
  return r;
}

template <class T> int d2d(pg_handle *h, T *dst, const T *src, size_t n) {
  if (n == 0) return PG_OK;
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, h->stream));
  return PG_OK;
}

// Step-indexed state of steps c+1 .. c+n of the differentiable form: what
// step s reads (u^s with X, v^{s-1}, the wall positions and grain boxes of
// step s) lives in slot s-c-1 of the segment buffers, and each step writes
// its results straight into the next slot, so keeping them costs no copies.
// Past the segment the handle's own buffers take the writes.  The slots' X
// parts must hold X (fill_slot_X).
struct SegMap {
  uint32_t c, n;
  bool in(uint32_t s) const { return s > c && s - c - 1 < n; }
  double *R(pg_handle *h, uint32_t s) const {
    return in(s) ? h->adj.su + (size_t)(s - c - 1) * rec_len(h) : h->rec[s & 1];
  }
  double *V(pg_handle *h, uint32_t s) const {
    return in(s) ? h->adj.sv + (size_t)(s - c - 1) * vec_len(h) : h->vel[(s + 1) & 1];
  }
  double *W(pg_handle *h, uint32_t s) const {
    return in(s) ? h->adj.swall + (size_t)(s - c - 1) * h->I * 6 : h->wallpos[s & 1];
  }
  unsigned long long *B(pg_handle *h, uint32_t s) const {
    return in(s) ? h->adj.sbox + (size_t)(s - c - 1) * box_len(h) : h->bbox[s % 3];
  }
  // the launch arguments of step t
  void remap(pg_handle *h, uint32_t t, StepArgs &As) const {
    As.rec[t & 1] = R(h, t);
    As.rec[(t + 1) & 1] = R(h, t + 1);
    As.vel[(t + 1) & 1] = V(h, t);
    As.vel[t & 1] = V(h, t + 1);
    As.wallpos[t & 1] = W(h, t);
    As.wallpos[(t + 1) & 1] = W(h, t + 1);
    As.bbox[t % 3] = B(h, t);
    As.bbox[(t + 1) % 3] = B(h, t + 1);
    As.bbox[(t + 2) % 3] = B(h, t + 2);
  }
  // slot 0 from the handle's state after step c (h->t == c), and the cleared
  // box buffer that step c+1 fills for step c+2
  int seed(pg_handle *h) const {
    CopySet cs;
    auto w8 = [](const void *p) { return (unsigned long long *)const_cast<void *>(p); };
    cs.dst[0] = w8(R(h, c + 1));
    cs.src[0] = w8(h->rec[(c + 1) & 1]);
    cs.n[0] = rec_len(h);
    cs.dst[1] = w8(V(h, c + 1));
    cs.src[1] = w8(h->vel[c & 1]);
    cs.n[1] = vec_len(h);
    cs.dst[2] = w8(W(h, c + 1));
    cs.src[2] = w8(h->wallpos[(c + 1) & 1]);
    cs.n[2] = (size_t)h->I * 6;
    cs.dst[3] = w8(B(h, c + 1));
    cs.src[3] = w8(h->bbox[(c + 1) % 3]);
    cs.n[3] = box_len(h);
    HIPCHK(launch_copy4(cs, h->stream));
    if (B(h, c + 2) != h->bbox[(c + 2) % 3]) RC(d2d(h, B(h, c + 2), h->bbox[(c + 2) % 3], box_len(h)));
    return PG_OK;
  }
  // after steps up to c+n ran on the slots: the handle's buffers of step c+n
  // (u^{c+n}, v^{c+n-1}, its walls and boxes) from the last slot, so that
  // the handle reads and steps on as after unmapped steps
  int settle(pg_handle *h) const {
    const uint32_t s = c + n;
    RC(d2d(h, h->rec[s & 1], R(h, s), rec_len(h)));
    RC(d2d(h, h->vel[(s + 1) & 1], V(h, s), vec_len(h)));
    RC(d2d(h, h->wallpos[s & 1], W(h, s), (size_t)h->I * 6));
    RC(d2d(h, h->bbox[s % 3], B(h, s), box_len(h)));
    return PG_OK;
  }
};

// The X parts of the first n record slots (steps write only u): every slot
// takes the whole record of rec[0], by copies that double the filled range.
int fill_slot_X(pg_handle *h, uint32_t n) {
  if (n == 0) return PG_OK;
  RC(d2d(h, h->adj.su, h->rec[0], rec_len(h)));
  for (uint32_t done = 1; done < n;) {
    const uint32_t m = std::min(done, n - done);
    RC(d2d(h, h->adj.su + (size_t)done * rec_len(h), h->adj.su, (size_t)m * rec_len(h)));
    done += m;
  }
  return PG_OK;
}

// Steps h->t+1 .. h->t+nsteps.  map: the steps of a segment of the
// differentiable form read and write its slots (no graph replay then).
int run_steps(pg_handle *h, uint32_t nsteps, const SegMap *map = nullptr) {
  if (nsteps == 0) return PG_OK;
  const uint32_t t_end = h->t + nsteps;
  if (t_end < h->t || t_end >= 0xFFFFFFF0u) return fail(PG_ERR_INVALID, "step counter overflow");
  bool react_queued = false; // the last step's reaction is in the stream behind it
  while (h->t < t_end) {
    uint32_t tr = next_reset(h);
    if (tr == h->t + 1) {
      // the moved grain's positions change: boxes of step tr and a new list
      for (uint32_t s = 0; s < h->I; ++s)
        if (h->inst[s].reset_step == tr && !h->inst[s].reset_done) RC(do_reset(h, s, tr));
      HIPCHK(launch_fill_box(h->bbox[tr % 3], box_len(h), h->dim, h->stream));
      HIPCHK(launch_boxes(h->dim, h->A, tr, h->bbox[tr % 3], h->stream));
      RC(build_list(h, tr));
      tr = next_reset(h); // the next one due, after tr
    }
    // a chunk ends before the next reset
    uint32_t chunk = std::min<uint32_t>(t_end - h->t, h->check_every);
    if (tr > h->t + 1) chunk = std::min<uint32_t>(chunk, tr - 1 - h->t);
    // a whole chunk without the recording step can be replayed from a graph
    const bool replay =
        !map && h->use_graph && !h->hist && chunk == h->check_every && h->t + chunk < t_end;
    if (replay) {
      RC(ensure_graph(h));
      h->h_pin[8] = h->t; // read by the copy below, before the chunk's final sync
      HIPCHK(hipMemcpyAsync(h->tbase, &h->h_pin[8], sizeof(uint32_t), hipMemcpyHostToDevice,
                            h->stream));
    }
    HIPCHK(hipEventRecord(h->kev[0], h->stream));
    if (replay) {
      HIPCHK(hipGraphLaunch(h->gexec, h->stream));
    } else {
      for (uint32_t s = h->t + 1; s <= h->t + chunk; ++s) {
        StepArgs As = h->A;
        if (map) map->remap(h, s, As);
        if (h->any_torque) HIPCHK(launch_grain_mean(h->dim, As, s, h->gmean, h->stream));
        if (h->hist) { // recording: the step's wall terms go to its ring slot
          const uint32_t slot = (s - 1) % h->check_every;
          As.wterm = h->wring + (size_t)slot * wring_stride(h);
          As.wflag = h->fring + (size_t)slot * fring_stride(h);
          HIPCHK(launch_step(h->dim, true, As, s, h->block, h->stream));
        } else {
          HIPCHK(launch_step(h->dim, s == t_end, As, s, h->block, h->stream));
        }
      }
    }
    HIPCHK(hipEventRecord(h->kev[1], h->stream));
    // a plain run's last chunk: its recorded step's reaction goes into the
    // stream now, so the wait below covers it; if the chunk turns out to
    // have stopped early, the redone steps queue it again
    react_queued = !h->hist && h->t + chunk == t_end;
    if (react_queued)
      RC(reactions(h, h->wterm, h->wflag, h->rhits, h->rcnt, 1, 0, 1, h->reaction,
                   (size_t)h->I * 18));
    HIPCHK(hipMemcpyAsync(h->h_pin, h->violation, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const uint32_t v = h->h_pin[0];
    if (h->hist) { // wall reactions of the steps that ran, in one launch
      const uint32_t ran = v > h->t ? std::min<uint32_t>(v - 1, h->t + chunk) - h->t : 0;
      RC(reactions(h, h->wring, h->fring, h->rhits_ring, h->rcnt_ring, ran,
                   h->t % h->check_every, h->check_every,
                   h->hist + (size_t)(h->t - h->hist_t0) * h->I * 18, (size_t)h->I * 18));
    }
    if (v > h->t + chunk) { // no launch of this run returned early
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, h->kev[0], h->kev[1]));
      h->kernel_ms += ms;
      h->kernel_launches += chunk;
    }
    if (v <= h->t + chunk) {
      if (v <= h->t) return fail(PG_ERR_STATE, "neighbour list invalid before the batch started");
      h->skipped += (uint64_t)(h->t + chunk) - v + 1;
      h->t = v - 1;
      react_queued = false;
      RC(build_list(h, v, map ? map->R(h, v) : nullptr));
      continue;
    }
    h->t += chunk;
  }
  if (h->hist) // the last step's reaction is the history's last entry
    HIPCHK(hipMemcpyAsync(h->reaction, h->hist + (size_t)(t_end - h->hist_t0 - 1) * h->I * 18,
                          (size_t)h->I * 18 * sizeof(double), hipMemcpyDeviceToDevice,
                          h->stream));
  else if (!react_queued)
    RC(reactions(h, h->wterm, h->wflag, h->rhits, h->rcnt, 1, 0, 1, h->reaction, (size_t)h->I * 18));
  return PG_OK;
}

// ------------------------------------------------------------ adjoint


// Save the state after step t (everything step t+1 reads) as checkpoint k.
int ck_save(pg_handle *h, uint32_t k, uint32_t t) {
  auto &a = h->adj;
  RC(d2d(h, a.ck_rec + k * rec_len(h), h->rec[(t + 1) & 1], rec_len(h)));
  RC(d2d(h, a.ck_vel + k * vec_len(h), h->vel[t & 1], vec_len(h)));
  RC(d2d(h, a.ck_acc + k * vec_len(h), h->acc, vec_len(h)));
  RC(d2d(h, a.ck_wall + (size_t)k * h->I * 6, h->wallpos[(t + 1) & 1], (size_t)h->I * 6));
  RC(d2d(h, a.ck_box + k * box_len(h), h->bbox[(t + 1) % 3], box_len(h)));
  return PG_OK;
}

// Restore checkpoint k (the state after step t) into the live buffers: bond
// status from the recorded break steps, a fresh box buffer for step t+2 and
// a new neighbour list.
int ck_restore(pg_handle *h, uint32_t k, uint32_t t) {
  auto &a = h->adj;
  RC(d2d(h, h->rec[(t + 1) & 1], a.ck_rec + k * rec_len(h), rec_len(h)));
  RC(d2d(h, h->vel[t & 1], a.ck_vel + k * vec_len(h), vec_len(h)));
  RC(d2d(h, h->acc, a.ck_acc + k * vec_len(h), vec_len(h)));
  RC(d2d(h, h->wallpos[(t + 1) & 1], a.ck_wall + (size_t)k * h->I * 6, (size_t)h->I * 6));
  RC(d2d(h, h->bbox[(t + 1) % 3], a.ck_box + k * box_len(h), box_len(h)));
  HIPCHK(launch_fill_box(h->bbox[(t + 2) % 3], box_len(h), h->dim, h->stream));
  HIPCHK(launch_adj_restore_bcol(h->d_bcol, a.break_step, h->E, t, h->stream));
  RC(set_unice(h, t));
  h->t = t;
  h->bcol_host_t = 0xFFFFFFFFu;
  return build_list(h, t + 1);
}

int adj_alloc(pg_handle *h, uint32_t T, uint32_t K) {
  auto &a = h->adj;
  const uint32_t nck = (T + K - 1) / K;
  if (!a.break_step) {
    RC(dalloc(h, a.break_step, h->E + BCOL_TAIL));
    fill_args(h);
  }
  if ((size_t)nck + 1 > a.cap_ck) {
    RC(dfree(h, a.ck_rec));
    RC(dfree(h, a.ck_vel));
    RC(dfree(h, a.ck_acc));
    RC(dfree(h, a.ck_wall));
    RC(dfree(h, a.ck_box));
    a.cap_ck = (size_t)nck + 1;
    RC(dalloc(h, a.ck_rec, a.cap_ck * rec_len(h)));
    RC(dalloc(h, a.ck_vel, a.cap_ck * vec_len(h)));
    RC(dalloc(h, a.ck_acc, a.cap_ck * vec_len(h)));
    RC(dalloc(h, a.ck_wall, a.cap_ck * h->I * 6));
    RC(dalloc(h, a.ck_box, a.cap_ck * box_len(h)));
  }
  const uint32_t seg = std::min(K, T); // slots of one segment
  if (seg > a.cap_seg) {
    RC(dfree(h, a.su));
    RC(dfree(h, a.sv));
    RC(dfree(h, a.swall));
    RC(dfree(h, a.sbox));
    a.cap_seg = seg;
    RC(dalloc(h, a.su, a.cap_seg * rec_len(h)));
    RC(dalloc(h, a.sv, a.cap_seg * vec_len(h)));
    RC(dalloc(h, a.swall, a.cap_seg * h->I * 6));
    RC(dalloc(h, a.sbox, a.cap_seg * box_len(h)));
  }
  if (T > a.cap_T) {
    RC(dfree(h, a.hist));
    RC(dfree(h, a.lR));
    a.cap_T = T;
    RC(dalloc(h, a.hist, a.cap_T * h->I * 18));
    RC(dalloc(h, a.lR, a.cap_T * h->I * 6 * h->dim));
  }
  if (!a.Pu) {
    RC(dalloc(h, a.Pu, vec_len(h)));
    RC(dalloc(h, a.Pv, vec_len(h)));
    RC(dalloc(h, a.Pa, vec_len(h)));
    RC(dalloc(h, a.lf[0], vec_len(h)));
    RC(dalloc(h, a.lf[1], vec_len(h)));
    RC(dalloc(h, a.gpart, (((size_t)h->N * ADJ_LANES + 255) / 256 + 1) * 3));
    RC(dalloc(h, a.dmax, 1));
    RC(dalloc(h, a.fin_rec, rec_len(h)));
    RC(dalloc(h, a.fin_force, vec_len(h)));
    RC(dalloc(h, a.fin_reaction, (size_t)h->I * 18));
    RC(dalloc(h, a.fin_box, box_len(h)));
  }
  a.nck = nck;
  return PG_OK;
}

// Forward recompute of steps c+1 .. c+n from checkpoint state, keeping in
// the segment slots what the adjoint of each step t reads: u^t, v^{t-1}, the
// wall positions and the grain boxes of step t.
int adj_recompute(pg_handle *h, uint32_t c, uint32_t n) {
  const uint32_t end = c + n;
  const SegMap m{c, n};
  RC(m.seed(h));
  while (h->t < end) {
    for (uint32_t t = h->t + 1; t <= end; ++t) {
      StepArgs As = h->A;
      m.remap(h, t, As);
      HIPCHK(launch_step(h->dim, false, As, t, h->block, h->stream));
    }
    HIPCHK(hipMemcpyAsync(h->h_pin, h->violation, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const uint32_t v = h->h_pin[0];
    if (v <= end) {
      if (v <= h->t) return fail(PG_ERR_STATE, "neighbour list invalid in the recompute");
      h->t = v - 1;
      RC(build_list(h, v, m.R(h, v)));
      continue;
    }
    h->t = end;
  }
  return PG_OK;
}

// Node field `field` of instance inst in the reference's order and layout
// ([n][dim]) from the given device arrays (records holding u of the step,
// velocity, acceleration, force density), copied on stream st.  ysg/ysh: the
// grain moved by reset_partzero_y and its y shift (-1: none).
int read_nodes_from(pg_handle *h, int inst, int field, const double *rec, const double *vel,
                    const double *acc, const double *force, int32_t ysg, double ysh,
                    hipStream_t st, double *dst) {
  const int D = h->dim;
  const InstHost &ih = h->inst[inst];
  const size_t n = ih.node_end - ih.node_begin;
  // device order -> reference order (the permutation stays inside the range)
  auto at = [&](size_t e) -> size_t { return h->inv[ih.node_begin + e] - ih.node_begin; };
  if (field == PG_FIELD_CURRPOS || field == PG_FIELD_DISP) {
    // record parts [D][ns][2] of the instance's node range; CurrPos = pos + disp
    std::vector<double> r((size_t)D * n * 2);
    for (int q = 0; q < D; ++q)
      RC(d2h_on(st, r.data() + (size_t)q * n * 2, rec + rec_index(h->ns, ih.node_begin, 2 * q),
                n * 2));
    auto el = [&](size_t e, int m) { return r[((size_t)(m >> 1) * n + e) * 2 + (m & 1)]; };
    // pos of the grain moved by reset_partzero_y carries the y shift
    size_t s0 = 0, s1 = 0;
    if (ysg >= 0) {
      s0 = h->grain_node_begin[ysg] - ih.node_begin;
      s1 = h->grain_node_begin[ysg + 1] - ih.node_begin;
    }
    for (size_t e = 0; e < n; ++e) {
      const size_t a = at(e);
      for (int k = 0; k < D; ++k) {
        double X = el(a, k);
        if (k == 1 && e >= s0 && e < s1) X = X + ysh;
        dst[e * D + k] = field == PG_FIELD_CURRPOS ? X + el(a, D + k) : el(a, D + k);
      }
    }
    return PG_OK;
  }
  const double *src = nullptr;
  switch (field) {
    case PG_FIELD_VEL: src = vel; break;
    case PG_FIELD_ACC: src = acc; break;
    case PG_FIELD_FORCE: src = force; break;
    default: return fail(PG_ERR_INVALID, "unknown field");
  }
  std::vector<double> v(n);
  for (int k = 0; k < D; ++k) {
    RC(d2h_on(st, v.data(), src + (size_t)k * h->ns + ih.node_begin, n));
    for (size_t e = 0; e < n; ++e) dst[e * D + k] = v[at(e)];
  }
  return PG_OK;
}

// Intact bonds of grain `grain` of instance inst as NbdArr2conn writes them,
// from a host copy of the bond table.
int read_conn_from(pg_handle *h, int inst, int grain, const std::vector<uint32_t> &st,
                   uint32_t *dst, uint64_t cap_rows, uint64_t *rows) {
  const InstHost &ih = h->inst[inst];
  if (grain < 0 || grain >= (int)(ih.grain_end - ih.grain_begin))
    return fail(PG_ERR_INVALID, "bad grain");
  const uint32_t gg = ih.grain_begin + grain;
  const uint32_t a = h->grain_node_begin[gg], b = h->grain_node_begin[gg + 1];
  uint64_t r = 0;
  for (uint32_t ge = a; ge < b; ++ge) { // reference node order
    const uint32_t i = h->inv[ge];
    const uint32_t sl = i / SLICE, lane = i % SLICE;
    for (uint32_t k = 0; k < h->deg[ge]; ++k) {
      const uint64_t e = bslot(h->slice_base[sl], lane, k);
      if (st[e] & BCOL_BROKEN) continue;
      const uint32_t q = h->perm[h->bcol[e]] - a, li = ge - a;
      if (li < q) {
        if (dst && r < cap_rows) {
          dst[2 * r] = li;
          dst[2 * r + 1] = q;
        }
        ++r;
      }
    }
  }
  if (rows) *rows = r;
  return PG_OK;
}

} // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

const char *pg_last_error(void) { return g_err.c_str(); }

int pg_version(char *buf, size_t cap) {
  return guarded([&]() -> int {
    const char *v = "libpg 0.1 (gfx950, FP64)";
    if (!buf || cap == 0) return fail(PG_ERR_INVALID, "null buffer");
    std::snprintf(buf, cap, "%s", v);
    return PG_OK;
  });
}

int pg_batch_create(const pg_setup *setups, const pg_params *params, int n, int device,
                    pg_handle **out) {
  if (!out || !setups || !params) return fail(PG_ERR_INVALID, "null argument");
  *out = nullptr;
  pg_handle *h = nullptr;
  try {
    h = new pg_handle();
    int rc = create_impl(setups, params, n, device, h);
    if (rc != PG_OK) {
      std::string keep = g_err;
      destroy_impl(h);
      g_err = keep;
      return rc;
    }
  } catch (const std::bad_alloc &) {
    destroy_impl(h);
    return fail(PG_ERR_ALLOC, "host allocation failed");
  } catch (const std::exception &e) {
    destroy_impl(h);
    return fail(PG_ERR_INVALID, e.what());
  }
  *out = h;
  return PG_OK;
}

int pg_create(const pg_setup *setup, const pg_params *params, int device, pg_handle **out) {
  return pg_batch_create(setup, params, 1, device, out);
}

int pg_destroy(pg_handle *h) {
  return guarded([&]() -> int {
    destroy_impl(h);
    return PG_OK;
  });
}

int pg_step(pg_handle *h, uint32_t nsteps) {
  return guarded([&]() -> int {
    if (!h) return fail(PG_ERR_INVALID, "null handle");
    HIPCHK(hipSetDevice(h->device));
    return run_steps(h, nsteps);
  });
}

int pg_sync(pg_handle *h) {
  return guarded([&]() -> int {
    if (!h) return fail(PG_ERR_INVALID, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return PG_OK;
  });
}

int pg_current_step(pg_handle *h, uint32_t *t) {
  return guarded([&]() -> int {
    if (!h || !t) return fail(PG_ERR_INVALID, "null argument");
    *t = h->t;
    return PG_OK;
  });
}

int pg_sizes(pg_handle *h, int inst, uint64_t *n_nodes, uint64_t *n_grains, uint64_t *n_slots) {
  return guarded([&]() -> int {
    if (!h || inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad handle or instance");
    const InstHost &ih = h->inst[inst];
    if (n_nodes) *n_nodes = ih.node_end - ih.node_begin;
    if (n_grains) *n_grains = ih.grain_end - ih.grain_begin;
    if (n_slots) *n_slots = ih.slot_end - ih.slot_begin;
    return PG_OK;
  });
}

int pg_read_nodes(pg_handle *h, int inst, int field, double *dst) {
  return guarded([&]() -> int {
    if (!h || !dst || inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    HIPCHK(hipSetDevice(h->device));
    const InstDev &id = h->inst_dev[inst];
    return read_nodes_from(h, inst, field, h->rec[h->t & 1], h->vel[h->t & 1], h->acc,
                           h->force_out, id.yshift_grain, id.yshift, h->stream, dst);
  });
}

int pg_read_wall(pg_handle *h, int inst, double *wall_info, double *reaction) {
  return guarded([&]() -> int {
    if (!h || inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    const int D = h->dim;
    if (wall_info) RC(d2h(h, wall_info, h->wallpos[(h->t + 1) & 1] + (size_t)inst * 6, 6));
    if (reaction) {
      double r[18];
      RC(d2h(h, r, h->reaction + (size_t)inst * 18, 18));
      for (int s = 0; s < PG_MAX_SIDES; ++s)
        for (int k = 0; k < D; ++k) reaction[s * D + k] = r[s * 3 + k];
    }
    return PG_OK;
  });
}

int pg_grain_stats(pg_handle *h, int inst, double *out) {
  return guarded([&]() -> int {
    if (!h || !out || inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    const int D = h->dim;
    const InstHost &ih = h->inst[inst];
    const uint32_t ng = ih.grain_end - ih.grain_begin;
    const size_t nv = (size_t)PG_GRAIN_STATS(D) * ng;
    HIPCHK(hipSetDevice(h->device));
    if (nv > h->stats_cap) {
      RC(dfree(h, h->d_stats));
      RC(dalloc(h, h->d_stats, nv));
      h->stats_cap = nv;
    }
    HIPCHK(launch_grain_stats(D, h->A, h->t, ih.grain_begin, ng, h->d_stats, h->stream));
    RC(d2h(h, out, h->d_stats, nv));
    return PG_OK;
  });
}

int pg_grain_stats_noalloc(pg_handle *h, int inst, double *out) {
  return guarded([&]() -> int {
    if (!h || !out || inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    const int D = h->dim;
    const InstHost &ih = h->inst[inst];
    const uint32_t ng = ih.grain_end - ih.grain_begin;
    const size_t nv = (size_t)PG_GRAIN_STATS(D) * ng;
    if (nv > (size_t)PG_NOALLOC_STATS_DOUBLES)
      return fail(PG_ERR_INVALID, "pg_grain_stats_noalloc: the instance has " + std::to_string(ng) +
                                      " grains; its static staging array holds " +
                                      std::to_string(PG_NOALLOC_STATS_DOUBLES / PG_GRAIN_STATS(D)) +
                                      " (pg_grain_stats takes any number)");
    if (h->device < 0 || h->device >= kMaxDevices)
      return fail(PG_ERR_INVALID, "pg_grain_stats_noalloc: device index beyond " +
                                      std::to_string(kMaxDevices - 1));
    HIPCHK(hipSetDevice(h->device));
    double *stage = nullptr;
    {
      // once per device, with that device current: the device address of
      // this device's copy of the array
      std::lock_guard<std::mutex> lock(g_scratch_mu);
      if (!g_scratch_dev[h->device]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PG_ERR_DEVICE, "hipGetSymbolAddress returned no address");
        g_scratch_dev[h->device] = static_cast<double *>(p);
      }
      stage = g_scratch_dev[h->device];
    }
    // every value of the output range is written by the kernel: nothing to clear
    HIPCHK(launch_grain_stats(D, h->A, h->t, ih.grain_begin, ng, stage, h->stream));
    RC(d2h(h, out, stage, nv));
    return PG_OK;
  });
}

int pg_read_bond_intact(pg_handle *h, int inst, uint8_t *dst) {
  return guarded([&]() -> int {
    if (!h || !dst || inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    const std::vector<uint32_t> *stp = nullptr;
    RC(bond_status(h, stp));
    const std::vector<uint32_t> &st = *stp;
    const InstHost &ih = h->inst[inst];
    for (uint32_t e = ih.node_begin; e < ih.node_end; ++e) { // reference order
      const uint32_t i = h->inv[e];
      const uint32_t sl = i / SLICE, lane = i % SLICE;
      for (uint32_t k = 0; k < h->deg[e]; ++k)
        dst[h->slot0[e] - ih.slot_begin + k] =
            (st[bslot(h->slice_base[sl], lane, k)] & BCOL_BROKEN) ? 0 : 1;
    }
    return PG_OK;
  });
}

int pg_read_conn(pg_handle *h, int inst, int grain, uint32_t *dst, uint64_t cap_rows,
                 uint64_t *rows) {
  return guarded([&]() -> int {
    if (!h || inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    const std::vector<uint32_t> *stp = nullptr;
    RC(bond_status(h, stp));
    return read_conn_from(h, inst, grain, *stp, dst, cap_rows, rows);
  });
}

int pg_read_contacts(pg_handle *h, int inst, uint32_t *dst, uint64_t cap_pairs, uint64_t *pairs) {
  return guarded([&]() -> int {
    if (!h || inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    if (h->t == 0) return fail(PG_ERR_STATE, "no step has been taken");
    const int D = h->dim;
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipMemsetAsync(h->cl_cnt, 0, ((size_t)h->N + 1) * sizeof(uint32_t), h->stream));
    HIPCHK(launch_pairs(D, false, h->A, h->t, h->cl_cnt, nullptr, nullptr, h->stream));
    uint32_t *ptr = nullptr;
    RC(dalloc(h, ptr, (size_t)h->N + 1));
    int rc = scan(h, h->cl_cnt, ptr, (size_t)h->N + 1);
    if (rc) {
      dfree(h, ptr);
      return rc;
    }
    uint32_t total = 0;
    rc = d2h(h, &total, ptr + h->N, 1);
    uint32_t *out = nullptr;
    if (!rc) rc = dalloc(h, out, (size_t)2 * (total ? total : 1));
    if (!rc) {
      hipError_t e = launch_pairs(D, true, h->A, h->t, nullptr, ptr, out, h->stream);
      if (e != hipSuccess) rc = fail(PG_ERR_DEVICE, hipGetErrorString(e));
    }
    std::vector<uint32_t> all((size_t)2 * total);
    if (!rc) rc = d2h(h, all.data(), out, all.size());
    dfree(h, ptr);
    dfree(h, out);
    if (rc) return rc;
    const InstHost &ih = h->inst[inst];
    // to reference indices, ordered by node and then by accumulation order
    std::vector<std::pair<uint32_t, uint32_t>> pr;
    for (uint64_t q = 0; q < total; ++q) {
      const uint32_t i = h->perm[all[2 * q]], j = h->perm[all[2 * q + 1]];
      if (i < ih.node_begin || i >= ih.node_end) continue;
      pr.push_back({i - ih.node_begin, j - ih.node_begin});
    }
    std::stable_sort(pr.begin(), pr.end(),
                     [](const std::pair<uint32_t, uint32_t> &x,
                        const std::pair<uint32_t, uint32_t> &y) { return x.first < y.first; });
    const uint64_t np = pr.size();
    if (dst)
      for (uint64_t q = 0; q < std::min(np, cap_pairs); ++q) {
        dst[2 * q] = pr[q].first;
        dst[2 * q + 1] = pr[q].second;
      }
    if (pairs) *pairs = np;
    return PG_OK;
  });
}

int pg_stream(pg_handle *h, void **stream) {
  return guarded([&]() -> int {
    if (!h || !stream) return fail(PG_ERR_INVALID, "null argument");
    *stream = (void *)h->stream;
    return PG_OK;
  });
}

int pg_stats(pg_handle *h, uint64_t *rebuilds, uint64_t *list_entries, uint64_t *skipped_steps) {
  return guarded([&]() -> int {
    if (!h) return fail(PG_ERR_INVALID, "null handle");
    if (rebuilds) *rebuilds = h->rebuilds;
    if (list_entries) *list_entries = h->list_entries;
    if (skipped_steps) *skipped_steps = h->skipped;
    return PG_OK;
  });
}

int pg_set_tuning(pg_handle *h, double skin_factor, int block) {
  return guarded([&]() -> int {
    if (!h) return fail(PG_ERR_INVALID, "null handle");
    if (!(skin_factor > 0) || skin_factor > 16)
      return fail(PG_ERR_INVALID, "skin_factor must be in (0, 16]");
    if (block != 64 && block != 128 && block != 256)
      return fail(PG_ERR_INVALID, "block must be 64, 128 or 256");
    HIPCHK(hipSetDevice(h->device));
    h->block = block;
    h->skin_factor = skin_factor;
    h->args_gen++; // a captured chunk holds the old block size
    for (uint32_t s = 0; s < h->I; ++s) {
      InstHost &ih = h->inst[s];
      ih.skin = skin_factor * (ih.Rc > 0 ? ih.Rc : 0.0);
      const double sh = 0.5 * ih.skin * (1.0 - 1e-9);
      h->inst_dev[s].skin_half2 = sh * sh;
    }
    RC(h2d(h, h->insts, h->inst_dev.data(), h->I));
    // the next step to run is t+1: rebuild from its positions
    return build_list(h, h->t + 1);
  });
}

int pg_autotune(pg_handle *h, uint32_t trial_steps, double *skin_factor, int *block) {
  return guarded([&]() -> int {
    if (!h || trial_steps == 0) return fail(PG_ERR_INVALID, "null handle or zero trial steps");
    static const double skins[] = {0.25, 0.5, 1.0};
    static const int blocks[] = {64, 128, 256};
    // the workload and device the choice is valid for
    hipDeviceProp_t prop{};
    HIPCHK(hipGetDeviceProperties(&prop, h->device));
    char key[512];
    std::snprintf(key, sizeof key, "%s/%d dim=%d nodes=%u slots=%llu grains=%u insts=%u keys=%u",
                  prop.gcnArchName, prop.multiProcessorCount, h->dim, h->N,
                  (unsigned long long)h->E, h->G, h->I, h->CK);
    const char *cache = std::getenv("PG_AUTOTUNE_CACHE");
    if (cache && *cache) {
      if (FILE *f = std::fopen(cache, "r")) {
        char line[640];
        double best_skin = 0.0;
        int best_block = 0;
        while (std::fgets(line, sizeof line, f)) {
          const char *bar = std::strstr(line, " | ");
          if (!bar || (size_t)(bar - line) != std::strlen(key) || std::strncmp(line, key, bar - line))
            continue;
          double sk = 0.0;
          int bl = 0;
          if (std::sscanf(bar + 3, "%lf %d", &sk, &bl) == 2) {
            best_skin = sk;
            best_block = bl;
          }
        }
        std::fclose(f);
        if (best_block) { // the last entry for this workload wins
          RC(pg_set_tuning(h, best_skin, best_block));
          if (skin_factor) *skin_factor = best_skin;
          if (block) *block = best_block;
          return PG_OK;
        }
      }
    }
    // one untimed run first, so the first candidate does not pay for cold
    // caches and first-launch costs
    RC(run_steps(h, trial_steps));
    HIPCHK(hipStreamSynchronize(h->stream));
    double best = 1e300, best_skin = h->skin_factor;
    int best_block = h->block;
    for (double sk : skins)
      for (int bl : blocks) {
        RC(pg_set_tuning(h, sk, bl));
        HIPCHK(hipStreamSynchronize(h->stream));
        const auto t0 = std::chrono::steady_clock::now();
        RC(run_steps(h, trial_steps));
        HIPCHK(hipStreamSynchronize(h->stream));
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (dt < best) {
          best = dt;
          best_skin = sk;
          best_block = bl;
        }
      }
    RC(pg_set_tuning(h, best_skin, best_block));
    if (cache && *cache) {
      if (FILE *f = std::fopen(cache, "a")) {
        std::fprintf(f, "%s | %.17g %d\n", key, best_skin, best_block);
        std::fclose(f);
      }
    }
    if (skin_factor) *skin_factor = best_skin;
    if (block) *block = best_block;
    return PG_OK;
  });
}

int pg_adjoint_forward(pg_handle *h, uint32_t nsteps, uint32_t checkpoint_every,
                       double *reaction_history) {
  return guarded([&]() -> int {
    if (!h || !reaction_history) return fail(PG_ERR_INVALID, "null argument");
    if (nsteps == 0 || checkpoint_every == 0) return fail(PG_ERR_INVALID, "nsteps and checkpoint_every must be > 0");
    if (h->I != 1) return fail(PG_ERR_INVALID, "the differentiable form takes one instance per handle");
    const InstDev &id = h->inst_dev[0];
    if (h->any_torque || id.set_movable_grain >= 0 || id.set_stoppable_grain >= 0 || h->any_clamped ||
        h->inst[0].reset_step != 0 || (id.flags & IF_UNGATED))
      return fail(PG_ERR_INVALID,
                  "the adjoint supports no torque, no timed movable/stoppable toggles, no clamped "
                  "nodes, no reset_partzero_y and only the standard loop");
    HIPCHK(hipSetDevice(h->device));
    auto &a = h->adj;
    RC(adj_alloc(h, nsteps, checkpoint_every));
    if (!h->wring) {
      RC(dalloc(h, h->wring, (size_t)h->check_every * wring_stride(h)));
      RC(dalloc(h, h->fring, (size_t)h->check_every * fring_stride(h)));
      RC(dalloc(h, h->rhits_ring, (size_t)h->check_every * react_chunks(h->CK) * kWflagSpare));
      RC(dalloc(h, h->rcnt_ring, (size_t)h->check_every * react_chunks(h->CK)));
      HIPCHK(hipMemsetAsync(h->fring, 0, (size_t)h->check_every * fring_stride(h), h->stream));
    }
    HIPCHK(launch_adj_init_break(h->d_bcol, a.break_step, h->E, h->stream));
    a.t0 = h->t;
    a.T = nsteps;
    a.K = checkpoint_every;
    a.recorded = false;
    a.stored = false;
    RC(ck_save(h, 0, h->t));
    h->hist = a.hist;
    h->hist_t0 = a.t0;
    int rc = PG_OK;
    if (a.nck == 1) {
      // one segment: the steps run on its slots, which then hold every
      // step's state for the backward pass
      const SegMap m{a.t0, a.T};
      rc = fill_slot_X(h, a.T);
      if (rc == PG_OK) rc = m.seed(h);
      if (rc == PG_OK) rc = run_steps(h, a.T, &m);
      if (rc == PG_OK) rc = m.settle(h);
      if (rc == PG_OK) rc = ck_save(h, 1, h->t);
      a.stored = rc == PG_OK;
    } else {
      for (uint32_t k = 0; k < a.nck && rc == PG_OK; ++k) {
        const uint32_t n = std::min(a.K, a.T - k * a.K);
        rc = run_steps(h, n);
        if (rc == PG_OK) rc = ck_save(h, k + 1, h->t);
      }
    }
    h->hist = nullptr;
    RC(rc);
    RC(d2d(h, a.fin_rec, h->rec[h->t & 1], rec_len(h)));
    RC(d2d(h, a.fin_force, h->force_out, vec_len(h)));
    RC(d2d(h, a.fin_reaction, h->reaction, (size_t)h->I * 18));
    RC(d2d(h, a.fin_box, h->bbox[h->t % 3], box_len(h)));
    const int D = h->dim;
    std::vector<double> hh((size_t)nsteps * 18);
    RC(d2h(h, hh.data(), a.hist, hh.size()));
    for (size_t t = 0; t < nsteps; ++t)
      for (int sd = 0; sd < PG_MAX_SIDES; ++sd)
        for (int k = 0; k < D; ++k)
          reaction_history[(t * PG_MAX_SIDES + sd) * D + k] = hh[t * 18 + sd * 3 + k];
    a.recorded = true;
    return PG_OK;
  });
}

int pg_adjoint_backward(pg_handle *h, const double *reaction_bar, double *grad) {
  return guarded([&]() -> int {
    if (!h || !reaction_bar || !grad) return fail(PG_ERR_INVALID, "null argument");
    auto &a = h->adj;
    if (!a.recorded) return fail(PG_ERR_STATE, "no recorded forward pass (pg_adjoint_forward)");
    HIPCHK(hipSetDevice(h->device));
    const int D = h->dim;
    RC(h2d(h, a.lR, reaction_bar, (size_t)a.T * PG_MAX_SIDES * D));
    HIPCHK(hipMemsetAsync(a.Pu, 0, vec_len(h) * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(a.Pv, 0, vec_len(h) * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(a.Pa, 0, vec_len(h) * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(a.lf[0], 0, vec_len(h) * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(a.lf[1], 0, vec_len(h) * sizeof(double), h->stream));
    const uint32_t blocks = (uint32_t)(((size_t)h->N * ADJ_LANES + 255) / 256);
    HIPCHK(hipMemsetAsync(a.gpart, 0, (size_t)blocks * 3 * sizeof(double), h->stream));
    // X parts of the segment's record slots (the recompute writes only u)
    if (!a.stored) RC(fill_slot_X(h, std::min(a.K, a.T)));
    InstHost &ih = h->inst[0];
    const double skin0 = ih.skin;
    int rc = PG_OK;
    for (int64_t k = (int64_t)a.nck - 1; k >= 0 && rc == PG_OK; --k) {
      const uint32_t c = a.t0 + (uint32_t)k * a.K;
      const uint32_t n = std::min(a.K, a.T - (uint32_t)k * a.K);
      if (!a.stored) { // the segment's states: from its checkpoint, recomputed
        rc = ck_restore(h, (uint32_t)k, c);
        if (rc == PG_OK) rc = adj_recompute(h, c, n);
        if (rc != PG_OK) break;
      }
      // one candidate list for the whole segment: pairs within the contact
      // radius at any of its steps are within radius + twice the largest
      // displacement at its first step
      const unsigned long long zero = 0;
      HIPCHK(hipMemcpyAsync(a.dmax, &zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
      HIPCHK(launch_adj_maxdisp(D, h->A, a.su, n, a.dmax, h->stream));
      unsigned long long dk = 0;
      RC(d2h(h, &dk, a.dmax, 1));
      const double dmax = dkey_inv(dk);
      ih.skin = 2.0 * dmax * (1.0 + 1e-6) + 1e-12 * ih.Rc;
      rc = build_list(h, c + 1, a.su);
      ih.skin = skin0;
      if (rc != PG_OK) break;
      for (uint32_t t = c + n; t >= c + 1; --t) {
        const size_t q = t - c - 1;
        AdjArgs B;
        B.A = h->A;
        B.su = a.su + q * rec_len(h);
        B.sv = a.sv + q * vec_len(h);
        B.swall = a.swall + q * h->I * 6;
        B.sbox = a.sbox + q * box_len(h);
        B.lR = a.lR + (size_t)(t - a.t0 - 1) * PG_MAX_SIDES * D;
        B.lf_in = a.lf[t & 1];
        B.lf_out = a.lf[(t - 1) & 1];
        B.Pu = a.Pu;
        B.Pv = a.Pv;
        B.Pa = a.Pa;
        B.gpart = a.gpart;
        hipError_t e = launch_adj_step(D, B, t, blocks, h->stream);
        if (e != hipSuccess) {
          rc = fail(PG_ERR_DEVICE, std::string("adjoint step: ") + hipGetErrorString(e));
          break;
        }
      }
    }
    ih.skin = skin0;
    RC(rc);
    std::vector<double> gp((size_t)blocks * 3);
    RC(d2h(h, gp.data(), a.gpart, gp.size()));
    grad[0] = grad[1] = grad[2] = 0.0;
    for (uint32_t b = 0; b < blocks; ++b)
      for (int q = 0; q < 3; ++q) grad[q] += gp[(size_t)b * 3 + q];
    // back to the state after the recorded pass
    RC(ck_restore(h, a.nck, a.t0 + a.T));
    RC(d2d(h, h->rec[h->t & 1], a.fin_rec, rec_len(h)));
    RC(d2d(h, h->force_out, a.fin_force, vec_len(h)));
    RC(d2d(h, h->reaction, a.fin_reaction, (size_t)h->I * 18));
    RC(d2d(h, h->bbox[h->t % 3], a.fin_box, box_len(h)));
    HIPCHK(hipStreamSynchronize(h->stream));
    return PG_OK;
  });
}

int pg_set_graph(pg_handle *h, int use_graph) {
  return guarded([&]() -> int {
    if (!h) return fail(PG_ERR_INVALID, "null handle");
    h->use_graph = use_graph != 0;
    return PG_OK;
  });
}

// Events on the handle's stream, for timing by the caller.
int pg_event_record(pg_handle *h, int slot) {
  return guarded([&]() -> int {
    if (!h || slot < 0 || slot >= 8) return fail(PG_ERR_INVALID, "bad event slot");
    HIPCHK(hipEventRecord(h->ev[slot], h->stream));
    return PG_OK;
  });
}

int pg_event_elapsed(pg_handle *h, int a, int b, float *ms) {
  return guarded([&]() -> int {
    if (!h || !ms || a < 0 || a >= 8 || b < 0 || b >= 8) return fail(PG_ERR_INVALID, "bad event slot");
    HIPCHK(hipEventSynchronize(h->ev[b]));
    HIPCHK(hipEventElapsedTime(ms, h->ev[a], h->ev[b]));
    return PG_OK;
  });
}

int pg_kernel_time(pg_handle *h, double *ms, uint64_t *launches) {
  return guarded([&]() -> int {
    if (!h) return fail(PG_ERR_INVALID, "null handle");
    if (ms) *ms = h->kernel_ms;
    if (launches) *launches = h->kernel_launches;
    return PG_OK;
  });
}

// ---------------------------------------------------------- snapshots

int pg_snapshot_take(pg_handle *h, int slot) {
  return guarded([&]() -> int {
    if (!h || slot < 0 || slot >= PG_SNAPSHOT_SLOTS) return fail(PG_ERR_INVALID, "bad snapshot slot");
    HIPCHK(hipSetDevice(h->device));
    auto &sn = h->snap[slot];
    if (!sn.alloc) {
      if (!h->rstream) HIPCHK(hipStreamCreateWithFlags(&h->rstream, hipStreamNonBlocking));
      RC(dalloc(h, sn.rec, rec_len(h)));
      RC(dalloc(h, sn.vel, vec_len(h)));
      RC(dalloc(h, sn.acc, vec_len(h)));
      RC(dalloc(h, sn.force, vec_len(h)));
      RC(dalloc(h, sn.wall, (size_t)h->I * 6));
      RC(dalloc(h, sn.reaction, (size_t)h->I * 18));
      RC(dalloc(h, sn.bcol, (size_t)h->E));
      HIPCHK(hipEventCreateWithFlags(&sn.ev, hipEventDisableTiming));
      sn.alloc = true;
    }
    // what pg_read_nodes / pg_read_wall / pg_read_conn read after step t
    RC(d2d(h, sn.rec, h->rec[h->t & 1], rec_len(h)));
    RC(d2d(h, sn.vel, h->vel[h->t & 1], vec_len(h)));
    RC(d2d(h, sn.acc, h->acc, vec_len(h)));
    RC(d2d(h, sn.force, h->force_out, vec_len(h)));
    RC(d2d(h, sn.wall, h->wallpos[(h->t + 1) & 1], (size_t)h->I * 6));
    RC(d2d(h, sn.reaction, h->reaction, (size_t)h->I * 18));
    RC(d2d(h, sn.bcol, h->d_bcol, (size_t)h->E));
    HIPCHK(hipEventRecord(sn.ev, h->stream));
    sn.t = h->t;
    sn.yshift_grain.resize(h->I);
    sn.yshift.resize(h->I);
    for (uint32_t q = 0; q < h->I; ++q) {
      sn.yshift_grain[q] = h->inst_dev[q].yshift_grain;
      sn.yshift[q] = h->inst_dev[q].yshift;
    }
    sn.bcol_ready = false;
    sn.taken = true;
    return PG_OK;
  });
}

namespace {
// The slot's copy, once the handle's stream has made it.
int snap_ready(pg_handle *h, int slot, pg_handle::Snap *&out) {
  if (!h || slot < 0 || slot >= PG_SNAPSHOT_SLOTS) return fail(PG_ERR_INVALID, "bad snapshot slot");
  auto &sn = h->snap[slot];
  if (!sn.taken) return fail(PG_ERR_STATE, "snapshot slot was never taken");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamWaitEvent(h->rstream, sn.ev, 0));
  out = &sn;
  return PG_OK;
}
} // namespace

int pg_snapshot_step(pg_handle *h, int slot, uint32_t *t) {
  return guarded([&]() -> int {
    if (!h || !t || slot < 0 || slot >= PG_SNAPSHOT_SLOTS || !h->snap[slot].taken)
      return fail(PG_ERR_INVALID, "bad snapshot slot");
    *t = h->snap[slot].t;
    return PG_OK;
  });
}

int pg_snapshot_read_nodes(pg_handle *h, int slot, int inst, int field, double *dst) {
  return guarded([&]() -> int {
    pg_handle::Snap *sn = nullptr;
    RC(snap_ready(h, slot, sn));
    if (!dst || inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    return read_nodes_from(h, inst, field, sn->rec, sn->vel, sn->acc, sn->force,
                           sn->yshift_grain[inst], sn->yshift[inst], h->rstream, dst);
  });
}

int pg_snapshot_read_conn(pg_handle *h, int slot, int inst, int grain, uint32_t *dst,
                          uint64_t cap_rows, uint64_t *rows) {
  return guarded([&]() -> int {
    pg_handle::Snap *sn = nullptr;
    RC(snap_ready(h, slot, sn));
    if (inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    if (!sn->bcol_ready) {
      sn->bcol_host.resize(h->E);
      RC(d2h_on(h->rstream, sn->bcol_host.data(), sn->bcol, (size_t)h->E));
      sn->bcol_ready = true;
    }
    return read_conn_from(h, inst, grain, sn->bcol_host, dst, cap_rows, rows);
  });
}

int pg_snapshot_read_wall(pg_handle *h, int slot, int inst, double *wall_info, double *reaction) {
  return guarded([&]() -> int {
    pg_handle::Snap *sn = nullptr;
    RC(snap_ready(h, slot, sn));
    if (inst < 0 || inst >= (int)h->I) return fail(PG_ERR_INVALID, "bad argument");
    const int D = h->dim;
    if (wall_info) RC(d2h_on(h->rstream, wall_info, sn->wall + (size_t)inst * 6, 6));
    if (reaction) {
      double r[18];
      RC(d2h_on(h->rstream, r, sn->reaction + (size_t)inst * 18, 18));
      for (int sd = 0; sd < PG_MAX_SIDES; ++sd)
        for (int k = 0; k < D; ++k) reaction[sd * D + k] = r[sd * 3 + k];
    }
    return PG_OK;
  });
}

} // extern "C"

extern "C" int pg_device_count(int *n) {
  return guarded([&]() -> int {
    if (!n) return fail(PG_ERR_INVALID, "null argument");
    *n = 0;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e == hipErrorNoDevice) return PG_OK;
    if (e != hipSuccess) return fail(PG_ERR_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    *n = c;
    return PG_OK;
  });
}

extern "C" int pg_check_bond(int device, int dim, int check, uint64_t n, const double *xi,
                             const double *eta, double *s, double *e) {
  return guarded([&]() -> int {
    if (!xi || !eta || !s || !e || (dim != 2 && dim != 3)) return fail(PG_ERR_INVALID, "bad argument");
    HIPCHK(hipSetDevice(device));
    double *d = nullptr;
    const size_t nv = n * (size_t)dim, tot = 3 * nv + n;
    HIPCHK(hipMalloc(&d, (tot ? tot : 1) * sizeof(double)));
    hipError_t err = hipMemcpy(d, xi, nv * sizeof(double), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(d + nv, eta, nv * sizeof(double), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = launch_bond_geom(dim, check != 0, n, d, d + nv, d + 2 * nv, d + 2 * nv + n, nullptr);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err == hipSuccess) err = hipMemcpy(s, d + 2 * nv, n * sizeof(double), hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(e, d + 2 * nv + n, nv * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (err != hipSuccess) return fail(PG_ERR_DEVICE, std::string("pg_check_bond: ") + hipGetErrorString(err));
    return PG_OK;
  });
}

extern "C" int pg_check_ordered_sum(int device, uint32_t nchains, const uint64_t *entry_off,
                                    const int32_t *two, const double *s0, const double *entries,
                                    double *wave, double *seq) {
  return guarded([&]() -> int {
    if (!entry_off || !two || !s0 || !entries || !wave || !seq)
      return fail(PG_ERR_INVALID, "null argument");
    if (entry_off[0] != 0) return fail(PG_ERR_INVALID, "entry_off[0] must be 0");
    for (uint32_t c = 0; c < nchains; ++c)
      if (entry_off[c + 1] < entry_off[c]) return fail(PG_ERR_INVALID, "entry_off decreases");
    const uint64_t ne = entry_off[nchains];
    if (ne > (1ull << 26)) return fail(PG_ERR_INVALID, "more than 2^26 entries");
    if (nchains == 0) return PG_OK;
    HIPCHK(hipSetDevice(device));
    // device layout: entries (+1 zero entry), s0, wave, seq, offsets, flags
    const size_t nd = (ne + 1) * 4 + 3 * (size_t)nchains;
    const size_t bytes = nd * sizeof(double) + (nchains + 1) * sizeof(uint64_t) +
                         nchains * sizeof(int32_t);
    char *d = nullptr;
    HIPCHK(hipMalloc(&d, bytes));
    double *dent = reinterpret_cast<double *>(d), *ds0 = dent + (ne + 1) * 4;
    double *dwave = ds0 + nchains, *dseq = dwave + nchains;
    uint64_t *doff = reinterpret_cast<uint64_t *>(dseq + nchains);
    int32_t *dtwo = reinterpret_cast<int32_t *>(doff + nchains + 1);
    hipError_t err = hipMemcpy(dent, entries, ne * 4 * sizeof(double), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(dent + ne * 4, 0, 4 * sizeof(double));
    if (err == hipSuccess) err = hipMemcpy(ds0, s0, nchains * sizeof(double), hipMemcpyHostToDevice);
    if (err == hipSuccess)
      err = hipMemcpy(doff, entry_off, (nchains + 1) * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(dtwo, two, nchains * sizeof(int32_t), hipMemcpyHostToDevice);
    if (err == hipSuccess)
      err = launch_chain_check(nchains, dent, doff, dtwo, ds0, (uint32_t)ne, dwave, dseq, nullptr);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err == hipSuccess) err = hipMemcpy(wave, dwave, nchains * sizeof(double), hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(seq, dseq, nchains * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (err != hipSuccess)
      return fail(PG_ERR_DEVICE, std::string("pg_check_ordered_sum: ") + hipGetErrorString(err));
    return PG_OK;
  });
}

extern "C" int pg_check_arith(int device, uint64_t n, const double *a, const double *b, double *q,
                              double *r) {
  return guarded([&]() -> int {
    if (!a || !b || !q || !r) return fail(PG_ERR_INVALID, "null argument");
    HIPCHK(hipSetDevice(device));
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, 4 * (n ? n : 1) * sizeof(double)));
    const size_t bytes = n * sizeof(double);
    hipError_t e = hipMemcpy(d, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_arith(n, d, d + n, d + 2 * n, d + 3 * n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(q, d + 2 * n, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(r, d + 3 * n, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(PG_ERR_DEVICE, std::string("pg_check_arith: ") + hipGetErrorString(e));
    return PG_OK;
  });
}
