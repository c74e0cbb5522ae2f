// simulate.cpp -- drop-in for the reference's bin/simulate2d and
// bin/simulate3d (run/simulate2d.cpp:31-94, makefile:59-65).
//
// Run from a directory holding config/main.conf and data/hdf5/all.h5; writes
// output/hdf5/plotinfo.h5, tc_%05u.h5, wall_%05u.h5 and run_time.h5 with the
// reference's counters and layout (run_timeloop, timeloop.h:500-1015).  The
// time stepping runs in libpg on a HIP device (PG_DEVICE, default 0).
// Built twice, with PG_DIM=2 and PG_DIM=3.
#include <sys/stat.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <deque>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "perigrain.h"
#include "pgio.h"

#ifndef PG_DIM
#error "PG_DIM must be 2 or 3"
#endif

namespace {

[[noreturn]] void die(const std::string &what, const char *msg) {
  std::cerr << "Error: " << what << ": " << msg << std::endl;
  std::exit(1);
}
void pg_ok(int rc, const char *what) {
  if (rc != PG_OK) die(what, pg_last_error());
}
void io_ok(int rc, const char *what) {
  if (rc != PG_OK) die(what, pgio_last_error());
}

std::string counter_file(const char *stem, unsigned c) {
  char b[64];
  std::snprintf(b, sizeof b, "output/hdf5/%s_%05u.h5", stem, c);
  return b;
}

// Snapshot writer: takes the tc/wall files of a step off the stepping
// thread.  The stepping thread copies a step's snapshot state into one of
// libpg's snapshot slots (pg_snapshot_take, a device-side copy in stream
// order) and queues the slot; this thread reads it back and writes the files
// (timeloop.h:881-951) while the steps go on.  A slot is taken again only
// after its files are written.
class SnapshotWriter {
 public:
  SnapshotWriter(pg_handle *h, int dim, const pg_setup &s) : h_(h), dim_(dim), s_(s) {
    th_ = std::thread([this] { run(); });
  }
  ~SnapshotWriter() { finish(); }

  // Wait until `slot` may be taken again; false if writing failed.
  bool wait_free(int slot) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return !busy_[slot] || !err_.empty(); });
    return err_.empty();
  }
  void push(unsigned counter, int slot) {
    std::lock_guard<std::mutex> lk(mu_);
    busy_[slot] = true;
    q_.push_back({counter, slot});
    cv_.notify_all();
  }
  // Write what is queued, stop the thread; the first error, or "".
  std::string finish() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      cv_.notify_all();
    }
    if (th_.joinable()) th_.join();
    return err_;
  }

 private:
  struct Job {
    unsigned counter;
    int slot;
  };
  void run() {
    const int G = s_.n_grains;
    const size_t N = (size_t)s_.node_offset[G];
    std::vector<double> cp(N * dim_), vel(N * dim_), acc(N * dim_), frc(N * dim_);
    std::vector<uint32_t> conn;
    std::vector<int64_t> conn_off(G + 1);
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !q_.empty() || stop_; });
        if (q_.empty()) return;
        j = q_.front();
        q_.pop_front();
      }
      std::string e = write(j, cp, vel, acc, frc, conn, conn_off);
      std::lock_guard<std::mutex> lk(mu_);
      busy_[j.slot] = false;
      if (!e.empty() && err_.empty()) err_ = e;
      cv_.notify_all();
      if (!err_.empty()) return;
    }
  }
  std::string write(const Job &j, std::vector<double> &cp, std::vector<double> &vel,
                    std::vector<double> &acc, std::vector<double> &frc,
                    std::vector<uint32_t> &conn, std::vector<int64_t> &conn_off) {
    auto pg = [](int rc, const char *what) {
      return rc == PG_OK ? std::string() : std::string(what) + ": " + pg_last_error();
    };
    std::string e;
    const int G = s_.n_grains;
    if ((e = pg(pg_snapshot_read_nodes(h_, j.slot, 0, PG_FIELD_CURRPOS, cp.data()), "read CurrPos")).size() ||
        (e = pg(pg_snapshot_read_nodes(h_, j.slot, 0, PG_FIELD_VEL, vel.data()), "read vel")).size() ||
        (e = pg(pg_snapshot_read_nodes(h_, j.slot, 0, PG_FIELD_ACC, acc.data()), "read acc")).size() ||
        (e = pg(pg_snapshot_read_nodes(h_, j.slot, 0, PG_FIELD_FORCE, frc.data()), "read force")).size())
      return e;
    conn.clear();
    conn_off[0] = 0;
    for (int g = 0; g < G; ++g) {
      uint64_t rows = 0, got = 0;
      if ((e = pg(pg_snapshot_read_conn(h_, j.slot, 0, g, nullptr, 0, &rows), "read Connectivity")).size())
        return e;
      const size_t at = conn.size();
      conn.resize(at + 2 * rows);
      if ((e = pg(pg_snapshot_read_conn(h_, j.slot, 0, g, conn.data() + at, rows, &got),
                  "read Connectivity")).size())
        return e;
      conn_off[g + 1] = conn_off[g] + (int64_t)rows;
    }
    const std::string tc = counter_file("tc", j.counter);
    if (pgio_write_tc(tc.c_str(), dim_, G, s_.node_offset, cp.data(), vel.data(), acc.data(),
                      frc.data(), conn_off.data(), conn.data()) != PG_OK)
      return tc + ": " + pgio_last_error();
    double wall_info[PG_MAX_SIDES] = {}, r[PG_MAX_SIDES * 3] = {}, reaction[PG_MAX_SIDES * 3] = {};
    if ((e = pg(pg_snapshot_read_wall(h_, j.slot, 0, wall_info, r), "read wall")).size()) return e;
    const int nside = dim_ == 2 ? 4 : 6;
    for (int q = 0; q < nside * dim_; ++q) reaction[q] = r[q];
    const std::string wf = counter_file("wall", j.counter);
    if (pgio_write_wall(wf.c_str(), dim_, wall_info, reaction) != PG_OK)
      return wf + ": " + pgio_last_error();
    return "";
  }

  pg_handle *h_;
  int dim_;
  const pg_setup &s_;
  std::thread th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Job> q_;
  bool busy_[PG_SNAPSHOT_SLOTS] = {};
  bool stop_ = false;
  std::string err_;
};

} // namespace

int main() {
  constexpr int dim = PG_DIM;
  using clock = std::chrono::system_clock;

  pgio_config cfg;
  pgio_config_defaults(&cfg);
  io_ok(pgio_read_config("config/main.conf", &cfg), "config/main.conf");

  pgio_setup *S = nullptr;
  io_ok(pgio_load_setup("data/hdf5/all.h5", dim, &cfg, &S), "data/hdf5/all.h5");
  std::cout << "extforce_gradient" << cfg.gradient_extforce << std::endl;

  // counters (timeloop.h:504-557)
  if (cfg.modulo == 0) die("config", "modulo must be > 0");
  unsigned counter, first_counter, last_counter;
  if (cfg.do_resume) {
    counter = cfg.resume_ind + 1;
    first_counter = counter;
    last_counter = cfg.resume_ind + cfg.timesteps / cfg.modulo;
    const std::string tc = counter_file("tc", cfg.resume_ind);
    const std::string wall = counter_file("wall", cfg.resume_ind);
    io_ok(pgio_resume(S, tc.c_str(), cfg.enable_fracture, cfg.wall_resume ? wall.c_str() : nullptr),
          "resume");
    std::cout << "Resuming from counter " << counter << std::endl;
  } else {
    counter = 1;
    first_counter = counter;
    last_counter = cfg.timesteps / cfg.modulo;
  }

  pg_setup s;
  pg_params p;
  io_ok(pgio_setup_view(S, &s, &p), "setup");
  ::mkdir("output", 0755);
  ::mkdir("output/hdf5", 0755);
  io_ok(pgio_write_plotinfo("output/hdf5/plotinfo.h5", dim, first_counter, last_counter, cfg.dt,
                            cfg.modulo, s.allow_wall, s.wall),
        "plotinfo.h5");

  const char *dev_env = std::getenv("PG_DEVICE");
  const int device = dev_env ? std::atoi(dev_env) : 0;
  auto t_run0 = clock::now();
  pg_handle *h = nullptr;
  pg_ok(pg_create(&s, &p, device, &h), "pg_create");

  std::vector<double> run_time, t_ind;
  SnapshotWriter writer(h, dim, s);
  int slot = 0;

  std::time_t now = std::time(nullptr);
  std::cout << "----------------------------------------------------------------------\n"
            << "Starting time loop (libpg, device " << device << "): " << std::ctime(&now);
  const int w = 12;
  char cnt[64];
  std::snprintf(cnt, sizeof cnt, "count[%u:%u]", first_counter, last_counter);
  std::cout << std::left << std::setw(w) << "t " << std::setw(w) << cnt << std::setw(w)
            << "duration(s)" << std::setw(w + 10) << "date" << std::endl;

  auto start = clock::now();
  unsigned t = 0;
  const unsigned T = cfg.timesteps;
  while (t < T) {
    // run to the next snapshot step (or to the end)
    unsigned next = T;
    if (cfg.save_file) {
      const unsigned k = (t / cfg.modulo + 1) * cfg.modulo;
      if (k <= T) next = k;
    }
    pg_ok(pg_step(h, next - t), "pg_step");
    t = next;
    if (!cfg.save_file || t % cfg.modulo != 0) continue;

    // snapshot (timeloop.h:881-951): the files are written by the writer
    // thread while the next steps run
    if (!writer.wait_free(slot)) break;
    pg_ok(pg_snapshot_take(h, slot), "pg_snapshot_take");
    writer.push(counter, slot);
    slot = (slot + 1) % PG_SNAPSHOT_SLOTS;

    auto stop = clock::now();
    const double ms =
        (double)std::chrono::duration_cast<std::chrono::milliseconds>(stop - start).count();
    start = stop;
    now = std::time(nullptr);
    char ts[80];
    std::strftime(ts, sizeof ts, "%F-%T", std::localtime(&now));
    std::cout << std::left << std::setw(w) << t << std::setw(w) << counter << std::setw(w)
              << ms / 1000 << std::setw(w + 10) << ts << std::endl;
    run_time.push_back(ms);
    t_ind.push_back(t);
    ++counter;
  }
  const std::string werr = writer.finish();
  if (!werr.empty()) die("snapshot", werr.c_str());
  pg_ok(pg_sync(h), "pg_sync");
  pg_destroy(h);
  pgio_setup_free(S);
  io_ok(pgio_write_runtime("output/hdf5/run_time.h5", run_time.size(), run_time.data(),
                           t_ind.data()),
        "run_time.h5");
  const auto secs =
      std::chrono::duration_cast<std::chrono::seconds>(clock::now() - t_run0).count();
  std::cout << "Runtime: " << secs << "s" << std::endl;
  return 0;
}
