// pgio.cpp -- libpgio: config/main.conf parsing, data/hdf5/all.h5 loading
// and the output/hdf5 writers of the drop-in boundary (include/pgio.h).
//
// Follows the reference's file handling:
//   ConfigVal                      src/read/read_config.h:14-290
//   load_particles / load_contact / load_wall   src/read/load_hdf5.h:41-201
//   load_col / load_rowvecs / load_rank1, store_col / store_rowvec
//                                  src/read/rw_hdf5.h:48-330
//   Timeloop/Contact/RectWall::apply_config
//                                  timeloop.h:174-202, contact.h:85-108, 254-301
//   resume                         timeloop.h:508-557
//   plotinfo / tc / wall / run_time writers   timeloop.h:561-582, 903-949, 1010-1014
// Uses the HDF5 C API directly.
#include <hdf5.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgio.h"

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

struct Quiet {
  Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};

// RAII for HDF5 identifiers.
struct Hid {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid() = default;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Hid(const Hid &) = delete;
  Hid &operator=(const Hid &) = delete;
  ~Hid() {
    if (id >= 0 && close) close(id);
  }
  bool ok() const { return id >= 0; }
};

hid_t mem_type(int t) {
  switch (t) {
    case PGIO_F64: return H5T_NATIVE_DOUBLE;
    case PGIO_I64: return H5T_NATIVE_INT64;
    case PGIO_U32: return H5T_NATIVE_UINT;
    case PGIO_I32: return H5T_NATIVE_INT;
  }
  return -1;
}

// A dataset read whole into memory of type T, with its shape.
template <class T> struct Arr {
  std::vector<T> v;
  int rank = 0;
  hsize_t dims[2] = {0, 0};
  hsize_t rows() const { return rank >= 1 ? dims[0] : 1; }
  hsize_t cols() const { return rank >= 2 ? dims[1] : 1; }
};

template <class T> int type_of();
template <> int type_of<double>() { return PGIO_F64; }
template <> int type_of<uint32_t>() { return PGIO_U32; }
template <> int type_of<int32_t>() { return PGIO_I32; }

template <class T> int read_arr(hid_t file, const std::string &name, Arr<T> &a) {
  Hid ds(H5Dopen2(file, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(PG_ERR_IO, "cannot open dataset " + name);
  Hid sp(H5Dget_space(ds.id), H5Sclose);
  const int rank = H5Sget_simple_extent_ndims(sp.id);
  if (rank < 0 || rank > 2) return fail(PG_ERR_IO, "dataset " + name + " has rank > 2");
  hsize_t d[2] = {1, 1};
  if (rank > 0) H5Sget_simple_extent_dims(sp.id, d, nullptr);
  a.rank = rank;
  a.dims[0] = rank >= 1 ? d[0] : 1;
  a.dims[1] = rank >= 2 ? d[1] : 1;
  const hsize_t n = a.dims[0] * a.dims[1];
  a.v.assign(n, T{});
  if (n && H5Dread(ds.id, mem_type(type_of<T>()), H5S_ALL, H5S_ALL, H5P_DEFAULT, a.v.data()) < 0)
    return fail(PG_ERR_IO, "cannot read dataset " + name);
  return PG_OK;
}

#define RC(x)                      \
  do {                             \
    int rc_ = (x);                 \
    if (rc_ != PG_OK) return rc_;  \
  } while (0)

// load_col<T>: first column of an [n][c] dataset (a rank-1 dataset is taken
// as one column).
template <class T> int load_col(hid_t f, const std::string &name, std::vector<T> &out) {
  Arr<T> a;
  RC(read_arr(f, name, a));
  out.resize(a.rows());
  for (hsize_t i = 0; i < a.rows(); ++i) out[i] = a.v[i * a.cols()];
  return PG_OK;
}
template <class T> int load_scalar(hid_t f, const std::string &name, T &out) {
  std::vector<T> v;
  RC(load_col(f, name, v));
  if (v.empty()) return fail(PG_ERR_IO, "dataset " + name + " is empty");
  out = v[0];
  return PG_OK;
}
// load_rank1<T>: every element, in storage order.
template <class T> int load_rank1(hid_t f, const std::string &name, std::vector<T> &out) {
  Arr<T> a;
  RC(read_arr(f, name, a));
  out = a.v;
  return PG_OK;
}
// load_rowvecs<T, dim>: the first dim columns of an [n][c] dataset, appended.
template <class T>
int load_rowvecs(hid_t f, const std::string &name, int dim, std::vector<T> &out, uint64_t *rows) {
  Arr<T> a;
  RC(read_arr(f, name, a));
  if ((int)a.cols() < dim)
    return fail(PG_ERR_IO, "dataset " + name + " has " + std::to_string(a.cols()) +
                               " columns, need " + std::to_string(dim));
  for (hsize_t i = 0; i < a.rows(); ++i)
    for (int k = 0; k < dim; ++k) out.push_back(a.v[i * a.cols() + k]);
  if (rows) *rows = a.rows();
  return PG_OK;
}

// Create every missing group on the path of `name` (everything before the
// last '/').
int make_groups(hid_t file, const std::string &name) {
  size_t pos = 0;
  while ((pos = name.find('/', pos + 1)) != std::string::npos) {
    const std::string g = name.substr(0, pos);
    if (g.empty() || g == "/") continue;
    htri_t e = H5Lexists(file, g.c_str(), H5P_DEFAULT);
    if (e > 0) continue;
    Hid gr(H5Gcreate2(file, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!gr.ok()) return fail(PG_ERR_IO, "cannot create group " + g);
  }
  return PG_OK;
}

int write_ds(hid_t file, const std::string &name, int type, int rank, uint64_t rows, uint64_t cols,
             const void *data) {
  RC(make_groups(file, name));
  hsize_t d[2] = {rows, cols};
  Hid sp(H5Screate_simple(rank, d, nullptr), H5Sclose);
  if (!sp.ok()) return fail(PG_ERR_IO, "cannot create dataspace for " + name);
  Hid ds(H5Dcreate2(file, name.c_str(), mem_type(type), sp.id, H5P_DEFAULT, H5P_DEFAULT,
                    H5P_DEFAULT),
         H5Dclose);
  if (!ds.ok()) return fail(PG_ERR_IO, "cannot create dataset " + name);
  if ((rows * cols) != 0 && H5Dwrite(ds.id, mem_type(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(PG_ERR_IO, "cannot write dataset " + name);
  return PG_OK;
}

hid_t create_file(const char *path) {
  return H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
}

std::string grain_name(unsigned i) {
  char b[32];
  std::snprintf(b, sizeof b, "P_%05u", i);
  return b;
}

} // namespace

struct pgio_setup {
  int dim = 0;
  std::vector<int64_t> node_offset{0}, conn_offset{0}, bdry_offset{0}, clamped_offset{0};
  std::vector<double> pos, vol, disp, vel, acc, ext, delta, rho, cnot, snot, torque_val;
  std::vector<uint32_t> conn, bdry, clamped, torque_axis;
  std::vector<int32_t> movable, breakable, stoppable;
  double contact_radius = 0, normal_stiffness = 0, damping_ratio = 0, friction_coefficient = 0;
  int32_t allow_wall = 0;
  double wall[PG_MAX_SIDES] = {0, 0, 0, 0, 0, 0};
  pg_params params{};
};

extern "C" {

const char *pgio_last_error(void) { return g_err.c_str(); }

int pgio_config_defaults(pgio_config *c) {
  return guarded([&]() -> int {
    if (!c) return fail(PG_ERR_INVALID, "null config");
    std::memset(c, 0, sizeof(*c));
    c->timesteps = 8000;
    c->modulo = 100;
    c->dt = 0.02 / 1e5;
    c->is_parallel = 1;
    c->do_resume = 0;
    c->wall_resume = 0;
    c->resume_ind = 50;
    c->save_file = 1;
    c->enable_fracture = 0;
    c->nl_bdry_only = 0;
    c->allow_damping = 1;
    c->allow_friction = 1;
    c->self_contact = 1;
    c->self_contact_rad = -1;
    c->enable_torque = 0;
    c->normal_stiffness = -1;
    c->damping_ratio = -1;
    c->friction_coefficient = -1;
    c->set_movable_index = -1;
    c->set_movable_timestep = -1;
    c->set_stoppable_index = -1;
    c->set_stoppable_timestep = -1;
    c->reset_partzero_y = 0;
    c->reset_partzero_y_timestep = 0;
    c->wheel_rad = 0;
    c->wall_left = c->wall_right = c->wall_top = c->wall_bottom = -999;
    c->wall_x_min = c->wall_y_min = c->wall_z_min = -999;
    c->wall_x_max = c->wall_y_max = c->wall_z_max = -999;
    c->gradient_extforce = 0;
    c->extforce_maxstep = 0;
    return PG_OK;
  });
}

int pgio_read_config(const char *path, pgio_config *c) {
  if (!path || !c) return fail(PG_ERR_INVALID, "null argument");
  std::ifstream in(path);
  if (!in.is_open()) {
    std::cerr << "Couldn't open config file for reading.\n";
    return PG_OK;
  }
  std::cout << "Reading config file: " << path << std::endl;
  std::string line;
  try {
    while (std::getline(in, line)) {
      std::string s;
      for (char ch : line)
        if (!std::isspace((unsigned char)ch)) s.push_back(ch);
      if (s.empty() || s[0] == '#') continue;
      const size_t eq = s.find('=');
      const std::string name = s.substr(0, eq);
      const std::string value = eq == std::string::npos ? s : s.substr(eq + 1);
      auto I = [&]() { return std::stoi(value); };
      auto D = [&]() { return std::stod(value); };
      // integer keys the reference stores in unsigned/bool members
      if (name == "timesteps") c->timesteps = (uint32_t)I();
      else if (name == "modulo") c->modulo = (uint32_t)I();
      else if (name == "dt") c->dt = D();
      else if (name == "is_parallel") c->is_parallel = I() != 0;
      else if (name == "do_resume") c->do_resume = I() != 0;
      else if (name == "wall_resume") c->wall_resume = I() != 0;
      else if (name == "resume_ind") c->resume_ind = (uint32_t)I();
      else if (name == "save_file") c->save_file = I() != 0;
      else if (name == "enable_fracture") c->enable_fracture = I() != 0;
      else if (name == "nl_bdry_only") c->nl_bdry_only = I() != 0;
      else if (name == "allow_damping") c->allow_damping = I() != 0;
      else if (name == "allow_friction") c->allow_friction = I() != 0;
      else if (name == "damping_ratio") c->damping_ratio = D();
      else if (name == "friction_coefficient") c->friction_coefficient = D();
      else if (name == "normal_stiffness") c->normal_stiffness = D();
      else if (name == "self_contact") c->self_contact = I() != 0;
      else if (name == "self_contact_rad") c->self_contact_rad = D();
      else if (name == "enable_torque") c->enable_torque = I() != 0;
      else if (name == "set_movable_index") c->set_movable_index = I();
      else if (name == "set_movable_timestep") c->set_movable_timestep = I();
      else if (name == "set_stoppable_index") c->set_stoppable_index = I();
      else if (name == "set_stoppable_timestep") c->set_stoppable_timestep = I();
      else if (name == "reset_partzero_y") c->reset_partzero_y = I() != 0;
      else if (name == "reset_partzero_y_timestep") c->reset_partzero_y_timestep = (uint32_t)I();
      else if (name == "wheel_rad") c->wheel_rad = D();
      else if (name == "wall_left") c->wall_left = D();
      else if (name == "wall_right") c->wall_right = D();
      else if (name == "wall_top") c->wall_top = D();
      else if (name == "wall_bottom") c->wall_bottom = D();
      else if (name == "speed_wall_left") c->speed_wall_left = D();
      else if (name == "speed_wall_right") c->speed_wall_right = D();
      else if (name == "speed_wall_top") c->speed_wall_top = D();
      else if (name == "speed_wall_bottom") c->speed_wall_bottom = D();
      else if (name == "wall_x_min") c->wall_x_min = D();
      else if (name == "wall_y_min") c->wall_y_min = D();
      else if (name == "wall_z_min") c->wall_z_min = D();
      else if (name == "wall_x_max") c->wall_x_max = D();
      else if (name == "wall_y_max") c->wall_y_max = D();
      else if (name == "wall_z_max") c->wall_z_max = D();
      else if (name == "speed_wall_x_min") c->speed_wall_x_min = D();
      else if (name == "speed_wall_y_min") c->speed_wall_y_min = D();
      else if (name == "speed_wall_z_min") c->speed_wall_z_min = D();
      else if (name == "speed_wall_x_max") c->speed_wall_x_max = D();
      else if (name == "speed_wall_y_max") c->speed_wall_y_max = D();
      else if (name == "speed_wall_z_max") c->speed_wall_z_max = D();
      else if (name == "gradient_extforce") c->gradient_extforce = I() != 0;
      else if (name == "extforce_maxstep") c->extforce_maxstep = (uint32_t)I();
      else {
        std::cerr << "Error: Wrong config name: " << name << " !!\n";
        c->unknown_keys++;
      }
    }
  } catch (const std::exception &e) {
    // std::stoi/stod throw on a malformed value; the reference terminates
    return fail(PG_ERR_INVALID, std::string("config value of line '") + line + "': " + e.what());
  }
  return PG_OK;
}

int pgio_load_setup(const char *all_h5, int dim, const pgio_config *cfg, pgio_setup **out) {
  return guarded([&]() -> int {
    static Quiet q;
    if (!all_h5 || !cfg || !out) return fail(PG_ERR_INVALID, "null argument");
    if (dim != 2 && dim != 3) return fail(PG_ERR_INVALID, "dim must be 2 or 3");
    *out = nullptr;
    std::unique_ptr<pgio_setup> S(new pgio_setup());
    S->dim = dim;
    Hid f(H5Fopen(all_h5, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(PG_ERR_IO, std::string("cannot open ") + all_h5);
    uint32_t total = 0;
    RC(load_scalar(f.id, "total_particles_univ", total));
    std::cout << "Reading total particles: " << total << std::endl;
    for (uint32_t i = 0; i < total; ++i) {
      const std::string g = grain_name(i + 1) + "/"; // 1-based in all.h5
      uint64_t n = 0, n2 = 0;
      RC(load_rowvecs(f.id, g + "Pos", dim, S->pos, &n));
      std::vector<double> vol;
      RC(load_col(f.id, g + "Vol", vol));
      if (vol.size() != n) return fail(PG_ERR_IO, g + "Vol has the wrong length");
      S->vol.insert(S->vol.end(), vol.begin(), vol.end());
      std::vector<uint32_t> b, cl;
      RC(load_col(f.id, g + "bdry_nodes", b));
      RC(load_rank1(f.id, g + "clamped_nodes", cl));
      S->bdry.insert(S->bdry.end(), b.begin(), b.end());
      S->clamped.insert(S->clamped.end(), cl.begin(), cl.end());
      RC(load_rowvecs(f.id, g + "disp", dim, S->disp, &n2));
      if (n2 != n) return fail(PG_ERR_IO, g + "disp has the wrong length");
      RC(load_rowvecs(f.id, g + "vel", dim, S->vel, &n2));
      if (n2 != n) return fail(PG_ERR_IO, g + "vel has the wrong length");
      RC(load_rowvecs(f.id, g + "acc", dim, S->acc, &n2));
      if (n2 != n) return fail(PG_ERR_IO, g + "acc has the wrong length");
      RC(load_rowvecs(f.id, g + "extforce", dim, S->ext, &n2));
      if (n2 != n) return fail(PG_ERR_IO, g + "extforce has the wrong length");
      uint64_t m = 0;
      RC(load_rowvecs(f.id, g + "Connectivity", 2, S->conn, &m));
      double delta, rho, cnot, snot;
      RC(load_scalar(f.id, g + "delta", delta));
      RC(load_scalar(f.id, g + "rho", rho));
      RC(load_scalar(f.id, g + "cnot", cnot));
      RC(load_scalar(f.id, g + "snot", snot));
      int32_t mov, brk, stp;
      RC(load_scalar(f.id, g + "movable", mov));
      RC(load_scalar(f.id, g + "breakable", brk));
      RC(load_scalar(f.id, g + "stoppable", stp));
      uint32_t tax = 2;
      double tval = 0;
      if (cfg->enable_torque) {
        RC(load_scalar(f.id, g + "torque_axis", tax));
        RC(load_scalar(f.id, g + "torque_val", tval));
      }
      S->node_offset.push_back(S->node_offset.back() + (int64_t)n);
      S->conn_offset.push_back(S->conn_offset.back() + (int64_t)m);
      S->bdry_offset.push_back(S->bdry_offset.back() + (int64_t)b.size());
      S->clamped_offset.push_back(S->clamped_offset.back() + (int64_t)cl.size());
      S->delta.push_back(delta);
      S->rho.push_back(rho);
      S->cnot.push_back(cnot);
      S->snot.push_back(snot);
      S->movable.push_back(mov);
      S->breakable.push_back(brk);
      S->stoppable.push_back(stp);
      S->torque_axis.push_back(tax);
      S->torque_val.push_back(tval);
    }
    std::cout << "Done reading particles." << std::endl;
    // load_contact (load_hdf5.h:131-151)
    RC(load_scalar(f.id, "pairwise/contact_radius", S->contact_radius));
    RC(load_scalar(f.id, "pairwise/normal_stiffness", S->normal_stiffness));
    RC(load_scalar(f.id, "pairwise/friction_coefficient", S->friction_coefficient));
    RC(load_scalar(f.id, "pairwise/damping_ratio", S->damping_ratio));
    // load_wall (load_hdf5.h:172-201)
    uint32_t aw = 0;
    RC(load_scalar(f.id, "wall/allow_wall", aw));
    S->allow_wall = aw != 0;
    std::vector<double> ci;
    RC(load_col(f.id, "wall/geom_wall_info", ci));
    const int nside = dim == 2 ? 4 : 6;
    if (ci.size() != 1) {
      if ((int)ci.size() < nside) return fail(PG_ERR_IO, "wall/geom_wall_info is too short");
      for (int k = 0; k < nside; ++k) S->wall[k] = ci[k];
    }

    // Contact::apply_config (contact.h:85-108)
    if (cfg->damping_ratio != -1) S->damping_ratio = cfg->damping_ratio;
    if (cfg->friction_coefficient != -1) S->friction_coefficient = cfg->friction_coefficient;
    if (cfg->normal_stiffness != -1) S->normal_stiffness = cfg->normal_stiffness;
    // RectWall::apply_config (contact.h:254-301); lrtb order: 2D left, right,
    // top, bottom; 3D x_min, y_min, z_min, x_max, y_max, z_max
    double *w = S->wall;
    double *sp = S->params.wall_speed;
    if (dim == 2) {
      if (cfg->wall_left != -999) w[0] = cfg->wall_left;
      if (cfg->wall_right != -999) w[1] = cfg->wall_right;
      if (cfg->wall_top != -999) w[2] = cfg->wall_top;
      if (cfg->wall_bottom != -999) w[3] = cfg->wall_bottom;
      sp[0] = cfg->speed_wall_left;
      sp[1] = cfg->speed_wall_right;
      sp[2] = cfg->speed_wall_top;
      sp[3] = cfg->speed_wall_bottom;
    } else {
      // the 2D keys wall_top / wall_right also set z_max / y_max in 3D
      if (cfg->wall_top != -999) w[5] = cfg->wall_top;
      if (cfg->wall_right != -999) w[4] = cfg->wall_right;
      if (cfg->wall_x_min != -999) w[0] = cfg->wall_x_min;
      if (cfg->wall_y_min != -999) w[1] = cfg->wall_y_min;
      if (cfg->wall_z_min != -999) w[2] = cfg->wall_z_min;
      if (cfg->wall_x_max != -999) w[3] = cfg->wall_x_max;
      if (cfg->wall_y_max != -999) w[4] = cfg->wall_y_max;
      if (cfg->wall_z_max != -999) w[5] = cfg->wall_z_max;
      sp[0] = cfg->speed_wall_x_min;
      sp[1] = cfg->speed_wall_y_min;
      sp[2] = cfg->speed_wall_z_min;
      sp[3] = cfg->speed_wall_x_max;
      sp[4] = cfg->speed_wall_y_max;
      sp[5] = cfg->speed_wall_z_max;
    }
    // Timeloop::apply_config (timeloop.h:174-202) and Contact toggles
    pg_params &p = S->params;
    p.dt = cfg->dt;
    p.enable_fracture = cfg->enable_fracture;
    p.nl_bdry_only = cfg->nl_bdry_only;
    p.allow_damping = cfg->allow_damping;
    p.allow_friction = cfg->allow_friction;
    p.self_contact = cfg->self_contact;
    p.gradient_extforce = cfg->gradient_extforce;
    p.extforce_maxstep = cfg->extforce_maxstep;
    p.enable_torque = cfg->enable_torque;
    p.set_movable_index = cfg->set_movable_index;
    p.set_movable_timestep = cfg->set_movable_timestep;
    p.set_stoppable_index = cfg->set_stoppable_index;
    p.set_stoppable_timestep = cfg->set_stoppable_timestep;
    p.reset_partzero_y = cfg->reset_partzero_y;
    p.reset_partzero_y_timestep = cfg->reset_partzero_y_timestep;
    p.wheel_rad = cfg->wheel_rad;
    *out = S.release();
    return PG_OK;
  });
}

int pgio_setup_view(pgio_setup *S, pg_setup *s, pg_params *p) {
  return guarded([&]() -> int {
    if (!S || !s) return fail(PG_ERR_INVALID, "null argument");
    std::memset(s, 0, sizeof(*s));
    auto ptr = [](auto &v) { return v.empty() ? nullptr : v.data(); };
    // never hand out NULL for arrays the stepper requires
    static const uint32_t zero_u32[2] = {0, 0};
    s->dim = S->dim;
    s->n_grains = (int32_t)S->delta.size();
    s->node_offset = S->node_offset.data();
    s->pos = ptr(S->pos);
    s->vol = ptr(S->vol);
    s->disp = ptr(S->disp);
    s->vel = ptr(S->vel);
    s->acc = ptr(S->acc);
    s->extforce = ptr(S->ext);
    s->conn_offset = S->conn_offset.data();
    s->conn = S->conn.empty() ? zero_u32 : S->conn.data();
    s->bdry_offset = S->bdry_offset.data();
    s->bdry_nodes = S->bdry.empty() ? zero_u32 : S->bdry.data();
    s->clamped_offset = S->clamped_offset.data();
    s->clamped_nodes = S->clamped.empty() ? zero_u32 : S->clamped.data();
    s->delta = ptr(S->delta);
    s->rho = ptr(S->rho);
    s->cnot = ptr(S->cnot);
    s->snot = ptr(S->snot);
    s->movable = ptr(S->movable);
    s->breakable = ptr(S->breakable);
    s->stoppable = ptr(S->stoppable);
    s->torque_axis = ptr(S->torque_axis);
    s->torque_val = ptr(S->torque_val);
    s->contact_radius = S->contact_radius;
    s->normal_stiffness = S->normal_stiffness;
    s->damping_ratio = S->damping_ratio;
    s->friction_coefficient = S->friction_coefficient;
    s->allow_wall = S->allow_wall;
    for (int k = 0; k < PG_MAX_SIDES; ++k) s->wall[k] = S->wall[k];
    if (p) *p = S->params;
    return PG_OK;
  });
}

int pgio_resume(pgio_setup *S, const char *tc_h5, int load_connectivity, const char *wall_h5) {
  return guarded([&]() -> int {
    static Quiet q;
    if (!S || !tc_h5) return fail(PG_ERR_INVALID, "null argument");
    const int dim = S->dim;
    Hid f(H5Fopen(tc_h5, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(PG_ERR_IO, std::string("cannot open ") + tc_h5);
    std::cout << "Loading from file: " << tc_h5 << std::endl;
    const size_t G = S->delta.size();
    std::vector<uint32_t> conn;
    std::vector<int64_t> conn_offset{0};
    for (size_t i = 0; i < G; ++i) {
      const std::string g = grain_name((unsigned)i) + "/"; // 0-based in tc files
      const int64_t a = S->node_offset[i], n = S->node_offset[i + 1] - a;
      std::vector<double> cp, v, ac;
      uint64_t r1 = 0, r2 = 0, r3 = 0;
      RC(load_rowvecs(f.id, g + "CurrPos", dim, cp, &r1));
      RC(load_rowvecs(f.id, g + "vel", dim, v, &r2));
      RC(load_rowvecs(f.id, g + "acc", dim, ac, &r3));
      if ((int64_t)r1 != n || (int64_t)r2 != n || (int64_t)r3 != n)
        return fail(PG_ERR_IO, g + " node count differs from all.h5");
      for (int64_t q = 0; q < n * dim; ++q) {
        // disp = CurrPos - pos (timeloop.h:525-527)
        S->disp[a * dim + q] = cp[q] - S->pos[a * dim + q];
        S->vel[a * dim + q] = v[q];
        S->acc[a * dim + q] = ac[q];
      }
      if (load_connectivity) {
        uint64_t m = 0;
        RC(load_rowvecs(f.id, g + "Connectivity", 2, conn, &m));
        conn_offset.push_back(conn_offset.back() + (int64_t)m);
      }
    }
    if (load_connectivity) {
      S->conn.swap(conn);
      S->conn_offset.swap(conn_offset);
    }
    if (wall_h5) {
      Hid fw(H5Fopen(wall_h5, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
      if (!fw.ok()) return fail(PG_ERR_IO, std::string("cannot open ") + wall_h5);
      std::cout << "Loading wall from file: " << wall_h5 << std::endl;
      std::vector<double> vv;
      RC(load_col(fw.id, "wall_info", vv));
      const size_t nside = dim == 2 ? 4 : 6;
      if (vv.size() < nside) return fail(PG_ERR_IO, "wall_info is too short");
      for (size_t k = 0; k < nside; ++k) S->wall[k] = vv[k];
    }
    return PG_OK;
  });
}

int pgio_setup_free(pgio_setup *S) {
  return guarded([&]() -> int {
    delete S;
    return PG_OK;
  });
}

int pgio_write_plotinfo(const char *path, int dim, uint32_t first_counter, uint32_t last_counter,
                        double dt, uint32_t modulo, int32_t allow_wall, const double *wall_info) {
  return guarded([&]() -> int {
    static Quiet q;
    if (!path || !wall_info) return fail(PG_ERR_INVALID, "null argument");
    Hid f(create_file(path), H5Fclose);
    if (!f.ok()) return fail(PG_ERR_IO, std::string("cannot create ") + path);
    const uint32_t fl[2] = {first_counter, last_counter};
    const uint32_t d = (uint32_t)dim;
    RC(write_ds(f.id, "f_l_counter", PGIO_U32, 2, 2, 1, fl));
    RC(write_ds(f.id, "dimension", PGIO_U32, 2, 1, 1, &d));
    RC(write_ds(f.id, "dt", PGIO_F64, 2, 1, 1, &dt));
    RC(write_ds(f.id, "modulo", PGIO_U32, 2, 1, 1, &modulo));
    RC(write_ds(f.id, "/wall/allow_wall", PGIO_I32, 2, 1, 1, &allow_wall));
    RC(write_ds(f.id, "/wall/geom_wall_info", PGIO_F64, 2, dim == 2 ? 4 : 6, 1, wall_info));
    return PG_OK;
  });
}

int pgio_write_tc(const char *path, int dim, int n_grains, const int64_t *node_offset,
                  const double *currpos, const double *vel, const double *acc,
                  const double *force, const int64_t *conn_offset, const uint32_t *conn) {
  return guarded([&]() -> int {
    static Quiet q;
    if (!path || !node_offset || !currpos || !vel || !acc || !force || !conn_offset)
      return fail(PG_ERR_INVALID, "null argument");
    Hid f(create_file(path), H5Fclose);
    if (!f.ok()) return fail(PG_ERR_IO, std::string("cannot create ") + path);
    for (int i = 0; i < n_grains; ++i) {
      const std::string g = grain_name((unsigned)i);
      const int64_t a = node_offset[i], n = node_offset[i + 1] - a;
      const int64_t c0 = conn_offset[i], m = conn_offset[i + 1] - c0;
      RC(write_ds(f.id, g + "/CurrPos", PGIO_F64, 2, n, dim, currpos + a * dim));
      RC(write_ds(f.id, g + "/vel", PGIO_F64, 2, n, dim, vel + a * dim));
      RC(write_ds(f.id, g + "/acc", PGIO_F64, 2, n, dim, acc + a * dim));
      RC(write_ds(f.id, g + "/force", PGIO_F64, 2, n, dim, force + a * dim));
      RC(write_ds(f.id, g + "/Connectivity", PGIO_U32, 2, m, 2, conn ? conn + 2 * c0 : nullptr));
    }
    return PG_OK;
  });
}

int pgio_write_wall(const char *path, int dim, const double *wall_info, const double *reaction) {
  return guarded([&]() -> int {
    static Quiet q;
    if (!path || !wall_info || !reaction) return fail(PG_ERR_INVALID, "null argument");
    Hid f(create_file(path), H5Fclose);
    if (!f.ok()) return fail(PG_ERR_IO, std::string("cannot create ") + path);
    const int ns = dim == 2 ? 4 : 6;
    RC(write_ds(f.id, "wall_info", PGIO_F64, 2, ns, 1, wall_info));
    RC(write_ds(f.id, "reaction", PGIO_F64, 2, ns, dim, reaction));
    return PG_OK;
  });
}

int pgio_write_runtime(const char *path, uint64_t n, const double *run_time, const double *t_ind) {
  return guarded([&]() -> int {
    static Quiet q;
    if (!path || (n && (!run_time || !t_ind))) return fail(PG_ERR_INVALID, "null argument");
    Hid f(create_file(path), H5Fclose);
    if (!f.ok()) return fail(PG_ERR_IO, std::string("cannot create ") + path);
    RC(write_ds(f.id, "run_time", PGIO_F64, 2, n, 1, run_time));
    RC(write_ds(f.id, "t_ind", PGIO_F64, 2, n, 1, t_ind));
    return PG_OK;
  });
}

int pgio_h5_write(const char *path, int create_new, const char *dataset, int type, int rank,
                  uint64_t rows, uint64_t cols, const void *data) {
  return guarded([&]() -> int {
    static Quiet q;
    if (!path || !dataset || ((rows * cols) != 0 && !data)) return fail(PG_ERR_INVALID, "null argument");
    if (mem_type(type) < 0) return fail(PG_ERR_INVALID, "unknown type");
    if (rank != 1 && rank != 2) return fail(PG_ERR_INVALID, "rank must be 1 or 2");
    hid_t id = -1;
    if (!create_new) id = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
    if (id < 0) id = create_file(path);
    Hid f(id, H5Fclose);
    if (!f.ok()) return fail(PG_ERR_IO, std::string("cannot open or create ") + path);
    return write_ds(f.id, dataset, type, rank, rows, rank == 1 ? 1 : cols, data);
  });
}

int pgio_h5_shape(const char *path, const char *dataset, int *rank, uint64_t *dims2) {
  return guarded([&]() -> int {
    static Quiet q;
    if (!path || !dataset || !rank || !dims2) return fail(PG_ERR_INVALID, "null argument");
    Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(PG_ERR_IO, std::string("cannot open ") + path);
    Hid ds(H5Dopen2(f.id, dataset, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(PG_ERR_IO, std::string("cannot open dataset ") + dataset);
    Hid sp(H5Dget_space(ds.id), H5Sclose);
    const int r = H5Sget_simple_extent_ndims(sp.id);
    if (r < 0 || r > 2) return fail(PG_ERR_IO, "rank > 2");
    hsize_t d[2] = {1, 1};
    if (r > 0) H5Sget_simple_extent_dims(sp.id, d, nullptr);
    *rank = r;
    dims2[0] = r >= 1 ? d[0] : 1;
    dims2[1] = r >= 2 ? d[1] : 1;
    return PG_OK;
  });
}

int pgio_h5_read(const char *path, const char *dataset, int type, void *dst, uint64_t cap_elems) {
  return guarded([&]() -> int {
    static Quiet q;
    if (!path || !dataset) return fail(PG_ERR_INVALID, "null argument");
    if (mem_type(type) < 0) return fail(PG_ERR_INVALID, "unknown type");
    int rank = 0;
    uint64_t d[2];
    RC(pgio_h5_shape(path, dataset, &rank, d));
    const uint64_t n = d[0] * d[1];
    if (n > cap_elems) return fail(PG_ERR_INVALID, "destination too small");
    if (n == 0) return PG_OK;
    if (!dst) return fail(PG_ERR_INVALID, "null destination");
    Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    Hid ds(H5Dopen2(f.id, dataset, H5P_DEFAULT), H5Dclose);
    if (H5Dread(ds.id, mem_type(type), H5S_ALL, H5S_ALL, H5P_DEFAULT, dst) < 0)
      return fail(PG_ERR_IO, std::string("cannot read dataset ") + dataset);
    return PG_OK;
  });
}

} // extern "C"
