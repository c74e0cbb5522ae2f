/*
 * pgio.h -- C ABI of libpgio, the file side of the drop-in boundary.
 *
 * The reference's executables (run/simulate2d.cpp; bin/simulate2d and
 * bin/simulate3d, makefile:59-65) read config/main.conf and
 * data/hdf5/all.h5 and write output/hdf5/{plotinfo.h5, tc_%05u.h5,
 * wall_%05u.h5, run_time.h5}.  libpgio reads and writes exactly those files
 * and turns them into the pg_setup / pg_params pair that libpg
 * (perigrain.h) takes.  The perigrain_amd simulate2d / simulate3d
 * executables are built on these entry points; tools and tests use them to
 * produce and inspect the same files.
 *
 * Conventions as in perigrain.h: PG_OK or a PG_ERR_* code, message from
 * pgio_last_error(), no exception crosses the ABI.
 */
#ifndef PGIO_H
#define PGIO_H

#include <stddef.h>
#include <stdint.h>

#include "perigrain.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Every key of config/main.conf (ConfigVal, src/read/read_config.h:14-120),
 * with the reference's defaults (read_config.h:61-120). */
typedef struct pgio_config {
  uint32_t timesteps, modulo;
  double dt;
  int32_t is_parallel;
  int32_t do_resume, wall_resume;
  uint32_t resume_ind;
  int32_t save_file, enable_fracture, nl_bdry_only;
  int32_t allow_damping, allow_friction;
  double normal_stiffness, damping_ratio, friction_coefficient; /* -1: keep file value */
  int32_t self_contact;
  double self_contact_rad;
  int32_t enable_torque;
  int32_t set_movable_index, set_movable_timestep;
  int32_t set_stoppable_index, set_stoppable_timestep;
  int32_t reset_partzero_y;
  uint32_t reset_partzero_y_timestep;
  double wheel_rad;
  /* -999: keep file value */
  double wall_left, wall_right, wall_top, wall_bottom;
  double wall_x_min, wall_y_min, wall_z_min, wall_x_max, wall_y_max, wall_z_max;
  double speed_wall_left, speed_wall_right, speed_wall_top, speed_wall_bottom;
  double speed_wall_x_min, speed_wall_y_min, speed_wall_z_min;
  double speed_wall_x_max, speed_wall_y_max, speed_wall_z_max;
  int32_t gradient_extforce;
  uint32_t extforce_maxstep;
  uint32_t unknown_keys; /* count of names read_file reports as wrong */
} pgio_config;

const char *pgio_last_error(void);

/* ConfigVal::ConfigVal() defaults (read_config.h:61-120). */
int pgio_config_defaults(pgio_config *cfg);

/* ConfigVal::read_file (read_config.h:124-290): whitespace is removed from
 * every line, lines starting with '#' and empty lines are skipped, the rest
 * is name=value.  Unknown names are reported on stderr and counted; a
 * missing file is reported on stderr and leaves the defaults (as in the
 * reference, which continues). */
int pgio_read_config(const char *path, pgio_config *cfg);

/* One simulation loaded from all.h5 (load_particles, load_contact,
 * load_wall: src/read/load_hdf5.h:41-201) with the config applied
 * (Timeloop::apply_config timeloop.h:174-202, Contact::apply_config
 * contact.h:85-108, RectWall::apply_config contact.h:254-301) and
 * break_bonds forced on (run/simulate2d.cpp:76-78). Owns its arrays. */
typedef struct pgio_setup pgio_setup;

int pgio_load_setup(const char *all_h5, int dim, const pgio_config *cfg, pgio_setup **out);

/* Views into a loaded setup, valid until it is freed or changed. */
int pgio_setup_view(pgio_setup *s, pg_setup *setup, pg_params *params);

/* Resume from a snapshot (run_timeloop, timeloop.h:508-557):
 * disp = CurrPos - pos, vel and acc from tc_h5; if load_connectivity, the
 * bond lists are rebuilt from its Connectivity rows (conn2NArr); if wall_h5
 * is not NULL the wall positions are set from its wall_info. */
int pgio_resume(pgio_setup *s, const char *tc_h5, int load_connectivity, const char *wall_h5);

int pgio_setup_free(pgio_setup *s);

/* Writers of the output files (timeloop.h:561-582, 903-949, 1010-1014).
 * Node arrays are [N*dim] over all grains; conn holds the grains'
 * Connectivity rows back to back, conn_offset[g] .. conn_offset[g+1]. */
int pgio_write_plotinfo(const char *path, int dim, uint32_t first_counter, uint32_t last_counter,
                        double dt, uint32_t modulo, int32_t allow_wall, const double *wall_info);
int pgio_write_tc(const char *path, int dim, int n_grains, const int64_t *node_offset,
                  const double *currpos, const double *vel, const double *acc,
                  const double *force, const int64_t *conn_offset, const uint32_t *conn);
int pgio_write_wall(const char *path, int dim, const double *wall_info, const double *reaction);
int pgio_write_runtime(const char *path, uint64_t n, const double *run_time, const double *t_ind);

/* Generic dataset access, for tools and tests (the role of rw_hdf5.h's
 * store_col / store_rowvec / load_col / load_rowvecs / load_rank1).
 * type: PGIO_F64, PGIO_I64, PGIO_U32, PGIO_I32 (memory type; HDF5
 * converts from the file type on read).  pgio_h5_write creates the file if
 * it does not exist (create_new truncates it first) and any missing groups
 * on the path. rank 1 writes a [rows] dataset, rank 2 a [rows][cols] one. */
#define PGIO_F64 0
#define PGIO_I64 1
#define PGIO_U32 2
#define PGIO_I32 3
int pgio_h5_write(const char *path, int create_new, const char *dataset, int type, int rank,
                  uint64_t rows, uint64_t cols, const void *data);
int pgio_h5_shape(const char *path, const char *dataset, int *rank, uint64_t *dims2);
int pgio_h5_read(const char *path, const char *dataset, int type, void *dst, uint64_t cap_elems);

#ifdef __cplusplus
}
#endif

#endif /* PGIO_H */
