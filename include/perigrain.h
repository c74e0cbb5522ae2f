/*
 * perigrain.h -- C ABI of libpg, the MI355X (gfx950) PeriDEM time stepper.
 *
 * The reference simulator exposes no FFI: its hot path is the C++ template
 *   template<unsigned dim> void run_timeloop(vector<ParticleN<dim>>& PArr,
 *        Timeloop TL, Contact CN, RectWall<dim> Wall)
 *   (src/particle/timeloop.h:500-502), driven by run/simulate2d.cpp:31-94.
 * libpg replaces that call.  The host (simulate2d/simulate3d drop-ins, the
 * Python binding) flattens what load_particles/load_contact/load_wall
 * (src/read/load_hdf5.h:41-201) produce, after the apply_config overrides
 * (src/particle/timeloop.h:174-202, src/particle/contact.h:85-108,254-301),
 * into a pg_setup + pg_params pair, and drives the loop through pg_step.
 *
 * Conventions
 *  - All entry points return PG_OK (0) or a PG_ERR_* code; no C++ exception
 *    crosses the ABI.  pg_last_error() returns a thread-local message.
 *  - Arrays passed in are caller-owned host memory; they are copied.
 *  - Node arrays are row-major [node][component], nodes of grain g occupying
 *    [node_offset[g], node_offset[g+1]) (the reference's P_%05u order).
 *  - All physics is IEEE binary64.  Indices are uint32 (grain-local in the
 *    setup, as in the reference's Connectivity datasets).
 *  - One host thread per handle; handles are independent.
 */
#ifndef PERIGRAIN_H
#define PERIGRAIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_OK 0
#define PG_ERR_INVALID 1 /* bad argument or inconsistent setup */
#define PG_ERR_DEVICE 2  /* HIP runtime error */
#define PG_ERR_ALLOC 3   /* device or host allocation failed */
#define PG_ERR_STATE 4   /* call not valid in the handle's state */
#define PG_ERR_IO 5      /* file I/O (HDF5) error */

/* Wall sides.  2D order follows RectWall::lrtb() (contact.h:139-143):
 * left, right, top, bottom.  3D order follows contact.h:145-150:
 * x_min, y_min, z_min, x_max, y_max, z_max. */
#define PG_MAX_SIDES 6

/* One simulation's initial state (what data/hdf5/all.h5 holds, after the
 * config overrides of Contact::apply_config and RectWall::apply_config). */
typedef struct pg_setup {
  int32_t dim;                /* 2 or 3 */
  int32_t n_grains;           /* total_particles_univ */
  const int64_t *node_offset; /* [n_grains+1] */
  const double *pos;          /* [N*dim]  P_xxxxx/Pos */
  const double *vol;          /* [N]      P_xxxxx/Vol */
  const double *disp;         /* [N*dim]  P_xxxxx/disp */
  const double *vel;          /* [N*dim]  P_xxxxx/vel */
  const double *acc;          /* [N*dim]  P_xxxxx/acc */
  const double *extforce;     /* [N*dim]  P_xxxxx/extforce (force density) */
  const int64_t *conn_offset; /* [n_grains+1] rows of Connectivity per grain */
  const uint32_t *conn;       /* [rows*2] grain-local (i,j) pairs, file order */
  const int64_t *bdry_offset; /* [n_grains+1] */
  const uint32_t *bdry_nodes; /* grain-local, file order (nl_bdry_only) */
  const int64_t *clamped_offset; /* [n_grains+1] */
  const uint32_t *clamped_nodes; /* grain-local */
  const double *delta;        /* [n_grains] */
  const double *rho;          /* [n_grains] */
  const double *cnot;         /* [n_grains] PMB micromodulus c */
  const double *snot;         /* [n_grains] critical stretch s0 */
  const int32_t *movable;     /* [n_grains] */
  const int32_t *breakable;   /* [n_grains] */
  const int32_t *stoppable;   /* [n_grains] */
  const uint32_t *torque_axis; /* [n_grains] or NULL (read only if enable_torque) */
  const double *torque_val;   /* [n_grains] or NULL */
  double contact_radius;      /* Contact::contact_rad */
  double normal_stiffness;    /* Contact::normal_stiffness (after override) */
  double damping_ratio;       /* after override */
  double friction_coefficient;/* after override */
  int32_t allow_wall;
  double wall[PG_MAX_SIDES];  /* lrtb() order, after RectWall::apply_config */
} pg_setup;

/* Time-loop and contact toggles (Timeloop/Contact/RectWall after
 * apply_config; keys of config/main.conf, src/read/read_config.h:298-446). */
typedef struct pg_params {
  double dt;
  int32_t enable_fracture;
  int32_t nl_bdry_only;
  int32_t allow_damping;
  int32_t allow_friction;
  int32_t self_contact;
  int32_t gradient_extforce;
  uint32_t extforce_maxstep;
  int32_t enable_torque;
  int32_t set_movable_index;     /* -1 = unused */
  int32_t set_movable_timestep;
  int32_t set_stoppable_index;   /* -1 = unused */
  int32_t set_stoppable_timestep;
  /* timeloop.h:629-647: at step reset_partzero_y_timestep, move the first
   * grain's pos in y so that its mean CurrPos y is the other grains' top y
   * plus wheel_rad plus the contact radius (bonds keep their reference
   * vectors).  Needs a second grain with at least two nodes; not with the
   * adjoint entry points. */
  int32_t reset_partzero_y;
  uint32_t reset_partzero_y_timestep;
  double wheel_rad;
  double wall_speed[PG_MAX_SIDES]; /* same order as pg_setup.wall */
  /* Which of the reference's loops the steps follow (PG_TIMELOOP_*):
   * run_timeloop (timeloop.h:500-1015), run_timeloop_worm (:1017-1369) or
   * run_timeloop_compact (:335-498).  The worm and compact loops advance
   * every grain's position, do not gate contact by stoppable, and have no
   * timed toggles or torque.  The worm loop also runs phases of 1000 steps:
   * the front nodes 0 and 3 of every grain are held in every third phase
   * and extforce acts only then, the back nodes 1 and 2 are held in the
   * others, and nodes 0-3 never move in y (every grain needs four nodes);
   * it has no extforce ramp.  Neither works with the adjoint entry points. */
  int32_t timeloop_variant;
} pg_params;

#define PG_TIMELOOP_STANDARD 0
#define PG_TIMELOOP_WORM 1
#define PG_TIMELOOP_COMPACT 2

/* Fields readable with pg_read_nodes. */
#define PG_FIELD_CURRPOS 0 /* CurrPos of the last completed step t */
#define PG_FIELD_VEL 1     /* vel after step t */
#define PG_FIELD_ACC 2     /* acc after step t */
#define PG_FIELD_FORCE 3   /* force density of step t */
#define PG_FIELD_DISP 4    /* disp used by step t (CurrPos = pos + disp) */

typedef struct pg_handle pg_handle;

/* Thread-local description of the last error on this thread. */
const char *pg_last_error(void);

/* Library/device information: writes a short NUL-terminated string. */
int pg_version(char *buf, size_t cap);

/* Number of HIP devices visible to the library. */
int pg_device_count(int *n);

/* Diagnostics (no reference counterpart): q[i] = a[i] / b[i] and
 * r[i] = sqrt(a[i]) computed on `device` with the operations the step
 * kernels use, so callers can check their rounding against the host's. */
int pg_check_arith(int device, uint64_t n, const double *a, const double *b, double *q,
                   double *r);

/* Diagnostics (no reference counterpart): the bond loop's own arithmetic on
 * given vectors xi and eta ([n][dim]): s[i] = (|eta| - |xi|)/|xi| and
 * e[i] = eta/|eta| as the step kernel computes them (particle2.h:134-146).
 * check = 1: with its per-bond range test and IEEE fallback (any input);
 * check = 0: the shortcut the kernel takes when every coordinate of the
 * step is "nice" (zero or of magnitude in [2^-240, 2^240]), valid only for
 * such inputs. */
int pg_check_bond(int device, int dim, int check, uint64_t n, const double *xi, const double *eta,
                  double *s, double *e);

/* Diagnostics (no reference counterpart): the wall reaction's ordered
 * subtraction (contact.h:339,374,411-412,428 accumulate one wall side's
 * reaction this way).  Chain c holds entries entry_off[c] .. entry_off[c+1]-1
 * of `entries` ([entry_off[nchains]][4] doubles, at most 2^26) and its terms
 * are terms 0 and 1 of every entry (two[c] != 0) or term 2; the result is
 * s0[c] minus its terms one at a time, in order.  wave[c] is that result as
 * the reaction kernel evaluates a long chain (one wavefront, chunks from
 * guessed start values corrected until exact); seq[c] is one device
 * thread's plain loop. */
int pg_check_ordered_sum(int device, uint32_t nchains, const uint64_t *entry_off,
                         const int32_t *two, const double *s0, const double *entries,
                         double *wave, double *seq);

/* Create a stepper for one simulation on HIP device `device`.
 * Replaces the state set-up of run_timeloop (timeloop.h:504-557; gen_xi at
 * particle2.h:236-252 and conn2NArr at load_hdf5.h:28-38) without the file
 * I/O.  The step counter starts at t=0 (run_timeloop's loop starts at t=1
 * even on resume, timeloop.h:608). */
int pg_create(const pg_setup *setup, const pg_params *params, int device,
              pg_handle **out);

/* Batched form: n independent simulations packed into one handle and
 * advanced together (one fused launch per step).  Each instance keeps its
 * own walls, contact constants and toggles; instances never interact.
 * All instances must share dim and dt. No reference counterpart. */
int pg_batch_create(const pg_setup *setups, const pg_params *params, int n,
                    int device, pg_handle **out);

int pg_destroy(pg_handle *h);

/* Advance every instance by nsteps time steps (the body of the loop at
 * timeloop.h:608-878 without the snapshot write).  Force density and the
 * wall reaction are recorded for the last of these steps.  Stream-ordered:
 * returns once the work is enqueued; the read functions synchronise. */
int pg_step(pg_handle *h, uint32_t nsteps);

/* Block until all enqueued work on the handle's stream is done. */
int pg_sync(pg_handle *h);

/* Index t of the last completed step (0 before the first pg_step). */
int pg_current_step(pg_handle *h, uint32_t *t);

/* Sizes of instance `inst`: node count, grain count, directed bond slots. */
int pg_sizes(pg_handle *h, int inst, uint64_t *n_nodes, uint64_t *n_grains,
             uint64_t *n_slots);

/* Copy a node field of instance `inst` ([N*dim] doubles) to host memory. */
int pg_read_nodes(pg_handle *h, int inst, int field, double *dst);

/* Wall positions after step t's motion (wall_xxxxx.h5:wall_info) and the
 * reaction of step t (wall_xxxxx.h5:reaction, get_reaction() order,
 * contact.h:185-202), [PG_MAX_SIDES] and [PG_MAX_SIDES*dim] doubles; only
 * the first 4 / 4*dim entries are meaningful in 2D. */
int pg_read_wall(pg_handle *h, int inst, double *wall_info, double *reaction);

/* Per directed bond slot, in the reference's NbdArr order (grain, node,
 * position in the node's row as built by conn2NArr): 1 = intact, 0 = broken.
 * [n_slots] bytes. */
int pg_read_bond_intact(pg_handle *h, int inst, uint8_t *dst);

/* Intact bonds of grain `grain` as NbdArr2conn (timeloop.h:25-49) writes
 * them: rows (i,q) with i<q in adjacency order.  Call with dst=NULL to get
 * the row count in *rows. */
int pg_read_conn(pg_handle *h, int inst, int grain, uint32_t *dst,
                 uint64_t cap_rows, uint64_t *rows);

/* Node-node contact pairs that produced a contact force in step t, as
 * (node i, partner j) pairs of instance-global node indices, ordered by i
 * and then by the order in which the reference accumulates them
 * (contact.h:675-811 called from timeloop.h:709-747).  dst=NULL returns the
 * count only. */
int pg_read_contacts(pg_handle *h, int inst, uint32_t *dst, uint64_t cap_pairs,
                     uint64_t *pairs);

/* On-device post-processing of the last completed step t, per grain of
 * instance inst in the reference's grain order: the per-particle quantities
 * of read_sim_output.py extract_bulk (:143-266) and the force sum of
 * read_force_val (:287-316, used by gen_wall_reaction.py), without writing
 * a snapshot.  out holds PG_GRAIN_STATS(dim) doubles per grain:
 *   [0, dim)        sum of force density over the grain's nodes
 *   [dim, 2dim)     mean velocity
 *   [2dim, 3dim)    mean acceleration
 *   [3dim, 4dim)    mean CurrPos
 *   4dim            damage: mean over nodes of (bonds at creation - intact
 *                   bonds) / bonds at creation (NaN for a node without bonds,
 *                   as numpy gives)
 *   [4dim+1, 5dim+1) sum of vol * rho * acc
 * The sums are parallel reductions, so they agree with numpy's to rounding,
 * not bit for bit.  Damage counts against the bonds the handle was created
 * with (a run resumed from a snapshot starts from that snapshot's bonds). */
#define PG_GRAIN_STATS(dim) (5 * (dim) + 1)
int pg_grain_stats(pg_handle *h, int inst, double *out);

/* pg_grain_stats without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory at the time of the
 * call (tight serving loops that sample the per-grain series every few
 * steps, harnesses that forbid hipMalloc).  Same kernel and launch, same
 * output, bit for bit; the per-grain results are staged in a static device
 * array of PG_NOALLOC_STATS_DOUBLES doubles (2 MiB) that the library holds on
 * each device it runs on until the process ends, instead of in a buffer it
 * allocates.  Instances with more than PG_NOALLOC_STATS_DOUBLES /
 * PG_GRAIN_STATS(dim) grains are rejected with PG_ERR_INVALID.  Every call
 * on a device shares that device's array: calls that overlap on one device
 * (from several host threads) race and must be serialised by the caller.
 * Each call writes every value it returns, so nothing is left from an
 * earlier call. */
#define PG_NOALLOC_STATS_DOUBLES (1 << 18)
int pg_grain_stats_noalloc(pg_handle *h, int inst, double *out);

/* Stream (hipStream_t) the handle enqueues its kernels on. */
int pg_stream(pg_handle *h, void **stream);

/* Statistics of the contact neighbour list: rebuild count, entries. */
int pg_stats(pg_handle *h, uint64_t *rebuilds, uint64_t *list_entries,
             uint64_t *skipped_steps);

/* Tuning knobs (used by the autotuner).  skin_factor: Verlet skin as a
 * multiple of the contact radius; block: threads per block of the step
 * kernel (64, 128 or 256).  The contact list is rebuilt at once. */
int pg_set_tuning(pg_handle *h, double skin_factor, int block);

/* Autotuner (no reference counterpart): runs the next trial_steps steps once
 * untimed (warm-up), then once per candidate configuration -- Verlet skin
 * 0.25, 0.5 or 1.0 times the contact radius, step-kernel blocks of 64, 128
 * or 256 threads -- timing each on the host clock including neighbour-list
 * rebuilds, and keeps the fastest.  The trials are ordinary steps of the
 * simulation (10*trial_steps in total): skin and block size change which
 * candidate pairs are listed and how threads are grouped, never a result.
 * If the environment variable PG_AUTOTUNE_CACHE names a file, a choice
 * recorded there for the same workload (dimension, node, slot, grain,
 * instance and contact-key counts) and device (architecture, CU count) is
 * applied without trials (and no steps run), and a new choice is appended
 * to it. */
int pg_autotune(pg_handle *h, uint32_t trial_steps, double *skin_factor, int *block);

/* Replay whole chunks of step launches (64 steps that do not include the
 * recording step) from a captured hipGraph (1), or launch each step (0).
 * The step index of a replayed launch comes from a device-side base that
 * the host sets before each replay.  Results are the same either way. */
int pg_set_graph(pg_handle *h, int use_graph);

/* Differentiable form (no reference counterpart).  pg_adjoint_forward runs
 * nsteps steps of a one-instance handle, recording the wall reaction of
 * every step into reaction_history ([nsteps][PG_MAX_SIDES][dim], the
 * get_reaction() layout of each step) and checkpointing the state every
 * checkpoint_every steps.  pg_adjoint_backward takes the cotangent of that
 * history (dL/dR, same layout) and returns in grad[3] the derivatives of L
 * with respect to
 *   grad[0]: a common scale of every grain's micromodulus cnot (at 1), which
 *            also scales the self-contact stiffness cnot/r;
 *   grad[1]: the normal contact stiffness (node-node and wall contact);
 *   grad[2]: the friction coefficient.
 * Bond status, contact pairs and every branch are frozen to the forward
 * pass; segments between checkpoints are recomputed from the checkpoints.
 * With checkpoint_every >= nsteps (one segment) the forward pass keeps every
 * step's state instead (node records, velocities, walls and boxes: about
 * 24*dim bytes per node and step) and the backward pass recomputes nothing.
 * Requires no torque, no timed toggles and no clamped nodes.  After the
 * backward pass the handle is back in the state after the recorded steps. */
int pg_adjoint_forward(pg_handle *h, uint32_t nsteps, uint32_t checkpoint_every,
                       double *reaction_history);
int pg_adjoint_backward(pg_handle *h, const double *reaction_bar, double *grad);

/* HIP events on the handle's stream (slots 0..7), for timing by callers:
 * record an event, then read the elapsed milliseconds between two slots
 * (waits for the later one). */
int pg_event_record(pg_handle *h, int slot);
int pg_event_elapsed(pg_handle *h, int a, int b, float *ms);

/* Device time spent in step-kernel launches since the handle was created:
 * pg_step brackets each run of back-to-back step launches with HIP events on
 * the handle's stream and adds up the elapsed times of runs that no
 * neighbour-list rebuild cut short.  *launches counts the step launches of
 * those runs, so *ms / *launches is the average step-kernel duration
 * (including any per-step grain-mean launches when torque is on). */
int pg_kernel_time(pg_handle *h, double *ms, uint64_t *launches);

/* Snapshots that overlap with stepping.  The reference writes a snapshot
 * between steps (timeloop.h:881-951); these calls let a caller write it from
 * another thread while the steps go on.  pg_snapshot_take enqueues on the
 * handle's stream, after the steps enqueued so far, a device copy of what a
 * snapshot of the last completed step holds -- CurrPos, vel, acc, force,
 * the bond status, the wall positions and reaction -- into snapshot slot
 * `slot` (0 .. PG_SNAPSHOT_SLOTS-1, device memory the handle allocates on the
 * slot's first use) and returns at once.  The pg_snapshot_* readers give the
 * same values as pg_read_nodes / pg_read_conn / pg_read_wall would have given
 * right after that step: they wait for the slot's copy, then copy to host
 * memory on a stream of their own, so one other host thread may call them
 * while the handle's own thread keeps calling pg_step.  A slot must not be
 * taken again until its reader is done with it. */
#define PG_SNAPSHOT_SLOTS 2
int pg_snapshot_take(pg_handle *h, int slot);
int pg_snapshot_step(pg_handle *h, int slot, uint32_t *t);
int pg_snapshot_read_nodes(pg_handle *h, int slot, int inst, int field, double *dst);
int pg_snapshot_read_conn(pg_handle *h, int slot, int inst, int grain, uint32_t *dst,
                          uint64_t cap_rows, uint64_t *rows);
int pg_snapshot_read_wall(pg_handle *h, int slot, int inst, double *wall_info,
                          double *reaction);

#ifdef __cplusplus
}
#endif

#endif /* PERIGRAIN_H */
