// epi_amd.hpp -- C++ facade of the epi5_MT call surface over the C ABI in
// epi_amd.h.  Header-only; link against libepi_amd.so.
//
// Code written against epi5_MT's driver (proc.cpp:90-173) keeps its calls:
//
//   connect_cell(cman);                      cell_connection.h:3
//   cell_interaction(cman);                  cell_interaction.h:5
//   pos_copy(cman);                          cellmanager.h:117
//   cell_state_renew(cman);                  cell_state_renew.h:3
//   cell_pos_periodic_fix(cman);             cellmanager.h:111
//   double zz = calc_zzmax(cman);            proc.cpp:34
//   setup_map_lat(cman, cmap1, cmap2);       map.h:5-7
//   calc_ext_stim(ext_stim, cmap1, cmap2, zz);            ext_stim.h:5
//   initialize_sc(cman, zz);                 cell_state_renew.h:4
//   calc_ca2p(cman, ATP, ext_stim.first(), cmap1, cmap2, zz);  ca2p.h:7
//   cman.should_calc_ca(); cman.ca_calc_condition_reset();     cellmanager.h:97-103
//   cman_init(cman, deck, use_last, uvp, w);                   cell_init.cpp:132-142
//
// What differs: the cells and the four grids live in GPU memory owned by
// the CellManager, so the grid objects are handles obtained from it
// (cman.atp(), cman.ext_stim(), cman.cell_map1(), cman.cell_map2()) rather
// than objects the caller allocates; the voxel map holds cell indices
// (-1 = none) where the reference holds Cell pointers; the random draws of
// genrand_real come from a replay stream (set_random_stream or
// seed_random).  Errors that the reference handles by printf + exit(1) or
// assert (a bad deck, a list or bucket overflow, a force beyond its
// sanity bound) throw epi_amd::Error, whose code is the EPI_ERR_* status.
#pragma once
#include <cstdint>
#include <cstdio>
#include <exception>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "epi_amd.h"

namespace epi_amd {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
  int code;
};

// cell states (define.h:36-47)
enum CELL_STATE : uint8_t { ALIVE = EPI_ALIVE, DEAD = EPI_DEAD, DISA = EPI_DISA, UNUSED = EPI_UNUSED,
                            FIX = EPI_FIX, BLANK = EPI_BLANK, DER = EPI_DER, MUSUME = EPI_MUSUME,
                            AIR = EPI_AIR, MEMB = EPI_MEMB };
using cmask_ty = uint8_t;
struct Cell;   // voxel maps hold cell indices; the type only names the map

inline void check(const epi_ctx* ctx, int rc, const char* what) {
  if (rc != EPI_OK) {
    std::string msg = std::string(what) + ": status " + std::to_string(rc);
    if (ctx) msg += std::string(" (") + epi_last_error(ctx) + ")";
    throw Error(rc, msg);
  }
}

class CellManager;

// Grid handles.  kind: 0 ATP, 1 ext_stim (B), 2 cmap1, 3 cmap2;
// buf: 0 = first() (the current values), 1 = second().
template <class T>
class FArr3D {
 public:
  FArr3D() = default;
  FArr3D(CellManager* owner, int kind, int buf) : owner_(owner), kind_(kind), buf_(buf) {}
  CellManager* owner() const { return owner_; }
  int kind() const { return kind_; }
  int buffer() const { return buf_; }
  // host copy in the reference's [x][y][z] order (Field.h:9-53)
  std::vector<T> read() const;
  // Field::output_binary (Field.h:30-40): raw [x][y][z] array
  void output_binary(const std::string& filename) const {
    const std::vector<T> v = read();
    std::ofstream ofs(filename, std::ios_base::out | std::ios_base::binary | std::ios_base::trunc);
    if (!ofs) throw Error(EPI_ERR_INVALID, "Field data output error. Filename:" + filename);
    ofs.write(reinterpret_cast<const char*>(v.data()), sizeof(T) * v.size());
  }

 private:
  CellManager* owner_ = nullptr;
  int kind_ = -1, buf_ = 0;
};

template <class F>
class SwapData;

template <class T>
class SwapData<FArr3D<T>> {
 public:
  SwapData() = default;
  SwapData(CellManager* owner, int kind) : owner_(owner), kind_(kind) {}
  FArr3D<T> first() const { return FArr3D<T>(owner_, kind_, 0); }
  FArr3D<T> second() const { return FArr3D<T>(owner_, kind_, 1); }
  CellManager* owner() const { return owner_; }
  // Field::read_binary of both buffers (proc.cpp:107-110)
  void read_binary(const std::string& filename);

 private:
  CellManager* owner_ = nullptr;
  int kind_ = -1;
};

// Host view of the cells (a download of epi_cells).
struct CellData {
  int32_t n = 0, nmemb = 0, nder = 0, sw = 0;
  std::vector<uint8_t> state;
  std::vector<int32_t> fix_origin, rest_div_times, pair, dermis, lat, conn_ptr, conn;
  std::vector<double> x, y, z, radius, ca2p, ip3, ca2p_avg, ex_inert, agek, ageb, ex_fat, in_fat,
      spr_nat_len, gj;
  std::vector<int64_t> uid;
};

// CellManager::output's per-cell record (cellmanager.cpp:451-538)
struct OutRecord {
  int32_t n = 0;
  std::vector<uint8_t> state, touch;
  std::vector<int32_t> rest_div_times, pair, fix_origin;
  std::vector<double> radius, ageb, agek, ca2p, x, y, z, ca2p_avg, ex_fat, in_fat, spr_nat_len;
};

// the file CellManager::output writes: 17 text columns per cell, or the
// binary records of cellmanager.cpp:496-513
inline void write_output(const OutRecord& d, const std::string& filename, bool binary_mode) {
  FILE* f = std::fopen(filename.c_str(), binary_mode ? "wb" : "w");
  if (!f) throw Error(EPI_ERR_INVALID, "Output file creation error. Filename:" + filename);
  for (int i = 0; i < d.n; i++) {
    if (binary_mode) {
      const int32_t idx = i, st = d.state[i], rdt = d.rest_div_times[i], pr = d.pair[i], fo = d.fix_origin[i];
      const double a[8] = {d.radius[i], d.ageb[i], d.agek[i], d.ca2p[i], d.x[i], d.y[i], d.z[i], d.ca2p_avg[i]};
      const double b[2] = {d.ex_fat[i], d.in_fat[i]};
      const char touch = (char)d.touch[i];
      std::fwrite(&idx, 4, 1, f); std::fwrite(&st, 4, 1, f); std::fwrite(a, 8, 8, f);
      std::fwrite(&rdt, 4, 1, f); std::fwrite(b, 8, 2, f); std::fwrite(&touch, 1, 1, f);
      std::fwrite(&d.spr_nat_len[i], 8, 1, f); std::fwrite(&pr, 4, 1, f); std::fwrite(&fo, 4, 1, f);
    } else {
      std::fprintf(f, "%d %d %.15f %.15f %.15f %.15f %.15f %.15f %.15f %.15f %d %.15f %.15f %d %.15f %d %d\n",
                   i, (int)d.state[i], d.radius[i], d.ageb[i], d.agek[i], d.ca2p[i], d.x[i], d.y[i], d.z[i],
                   d.ca2p_avg[i], d.rest_div_times[i], d.ex_fat[i], d.in_fat[i], (int)d.touch[i], d.spr_nat_len[i],
                   d.pair[i], d.fix_origin[i]);
    }
  }
  std::fclose(f);
}

class CellManager {
 public:
  // geometry: define.h:72-113 as runtime values
  explicit CellManager(const epi_geometry& g) : geo_(g) {
    check(nullptr, epi_create(1, &geo_, &ctx_), "epi_create");
  }
  ~CellManager() {
    try {
      flush_outputs();
    } catch (...) {
    }
    epi_destroy(ctx_);
  }
  CellManager(const CellManager&) = delete;
  CellManager& operator=(const CellManager&) = delete;

  epi_ctx* ctx() const { return ctx_; }
  const epi_geometry& geometry() const { return geo_; }
  epi_status status() const {
    epi_status s{};
    check(ctx_, epi_get_status(ctx_, 0, &s), "epi_get_status");
    return s;
  }
  size_t size() const { return (size_t)status().n; }                   // cellmanager.h:92
  bool should_calc_ca() const {                                         // cellmanager.h:97
    int32_t f = 0;
    check(ctx_, epi_should_calc_ca(ctx_, &f), "should_calc_ca");
    return f != 0;
  }
  void ca_calc_condition_reset() {                                      // cellmanager.h:101
    check(ctx_, epi_ca_calc_condition_reset(ctx_), "ca_calc_condition_reset");
  }
  // clean_up (cellmanager.cpp:543-557) erases stale gj entries; stale
  // entries are never kept here, so there is nothing to do
  void clean_up() {}
  // the step the driver is in (proc.cpp:132): cell_state_renew records it in
  // the cells' cell_stat
  unsigned long long current_timestep = 0;
  // push current_timestep to the device when the device's count differs
  void sync_timestep() {
    if ((long long)current_timestep != device_ts_) {
      check(ctx_, epi_set_timestep(ctx_, 0, (int64_t)current_timestep), "epi_set_timestep");
      device_ts_ = (long long)current_timestep;
    }
  }
  void renewed() { device_ts_++; }   // a renewal advanced the device's timestep

  // append_stat_data (cellmanager.cpp:559-567): one line "k_aging_start
  // k_cornified k_disap" per DEAD cell removed since the last call, appended
  // to filename; the store is then emptied.  A field the run never set is
  // written as -1 (the reference leaves it uninitialised).
  void append_stat_data(const std::string& filename) {
    const size_t cap = (size_t)geo_.max_cells;
    std::vector<int64_t> r(3 * cap);
    int32_t k = 0;
    check(ctx_, epi_take_cell_stats(ctx_, 0, r.data(), (int32_t)cap, &k), "epi_take_cell_stats");
    if (k == 0) return;
    FILE* f = std::fopen(filename.c_str(), "a");
    if (!f) throw Error(EPI_ERR_INVALID, "stat data append error: " + filename);
    for (int32_t i = 0; i < k; i++)
      std::fprintf(f, "%lld %lld %lld\n", (long long)r[3 * i], (long long)r[3 * i + 1], (long long)r[3 * i + 2]);
    std::fclose(f);
  }

  // runtime parameters: the keys of epi_intg's CalcParams
  // (epi_intg/calc/CalcParams.cpp:51-162), epi5_MT's values by default.
  // Set them before cman_init for the load to use agki_max, fac and div_max.
  epi_params params() const {
    epi_params p;
    check(ctx_, epi_get_params(ctx_, 0, &p), "epi_get_params");
    return p;
  }
  void set_params(const epi_params& p) { check(ctx_, epi_set_params(ctx_, 0, &p), "epi_set_params"); }
  // Params::load over parse_paramtext (epi_intg/Params.cpp:91-96,
  // parser.cpp:30-58): every key must be in the file unless require_all is
  // false; the grid and box are this manager's
  void load_params(const std::string& path, bool require_all = true) {
    epi_params p = params();
    char err[512];
    const int rc = epi_params_load(path.c_str(), require_all ? 1 : 0, &p, err, (int32_t)sizeof err);
    if (rc != EPI_OK) throw Error(rc, err);
    set_params(p);
  }
  // Params::generate_paramfile (Params.cpp:98-110)
  void save_params(const std::string& path) const {
    const epi_params p = params();
    if (epi_params_save(path.c_str(), &p) != EPI_OK) throw Error(EPI_ERR_INVALID, "cannot write " + path);
  }

  // grids owned by this manager
  SwapData<FArr3D<double>> atp() { return SwapData<FArr3D<double>>(this, 0); }
  SwapData<FArr3D<double>> ext_stim() { return SwapData<FArr3D<double>>(this, 1); }
  FArr3D<const Cell*> cell_map1() { return FArr3D<const Cell*>(this, 2, 0); }
  FArr3D<cmask_ty> cell_map2() { return FArr3D<cmask_ty>(this, 3, 0); }

  // genrand_real replay stream (Random_gen.cpp:14-54)
  void set_random_stream(const std::vector<double>& draws) {
    check(ctx_, epi_set_rng(ctx_, 0, draws.data(), (int64_t)draws.size()), "epi_set_rng");
  }
  void seed_random(uint64_t seed, int64_t n_draws) {
    std::vector<double> d((size_t)n_draws);
    check(ctx_, epi_rng_mt19937_64(seed, d.data(), n_draws), "epi_rng_mt19937_64");
    set_random_stream(d);
  }

  CellData download() const {
    const epi_status s = status();
    CellData d;
    const size_t n = (size_t)s.n;
    d.n = s.n;
    d.state.resize(n); d.fix_origin.resize(n); d.rest_div_times.resize(n); d.pair.resize(n);
    d.dermis.resize(n); d.lat.resize(3 * n); d.conn_ptr.resize(n + 1);
    for (auto* v : {&d.x, &d.y, &d.z, &d.radius, &d.ca2p, &d.ip3, &d.ca2p_avg, &d.ex_inert, &d.agek,
                    &d.ageb, &d.ex_fat, &d.in_fat, &d.spr_nat_len})
      v->resize(n);
    d.uid.resize(n);
    const size_t cap = n * (size_t)(geo_.row_stride > 0 ? geo_.row_stride : 400);
    d.conn.resize(cap); d.gj.resize(cap);
    epi_cells o{};
    o.state = d.state.data(); o.fix_origin = d.fix_origin.data();
    o.x = d.x.data(); o.y = d.y.data(); o.z = d.z.data(); o.radius = d.radius.data();
    o.ca2p = d.ca2p.data(); o.ip3 = d.ip3.data(); o.ca2p_avg = d.ca2p_avg.data();
    o.ex_inert = d.ex_inert.data(); o.agek = d.agek.data(); o.ageb = d.ageb.data();
    o.ex_fat = d.ex_fat.data(); o.in_fat = d.in_fat.data(); o.spr_nat_len = d.spr_nat_len.data();
    o.rest_div_times = d.rest_div_times.data(); o.pair = d.pair.data(); o.dermis = d.dermis.data();
    o.lat = d.lat.data(); o.uid = d.uid.data();
    o.conn_ptr = d.conn_ptr.data(); o.conn = d.conn.data(); o.gj = d.gj.data();
    o.conn_cap = (int32_t)cap;
    check(ctx_, epi_get_cells(ctx_, 0, &o), "epi_get_cells");
    d.nmemb = o.nmemb; d.nder = o.nder; d.sw = o.sw;
    d.conn.resize((size_t)d.conn_ptr[n]); d.gj.resize((size_t)d.conn_ptr[n]);
    return d;
  }

  // CellManager::output (cellmanager.cpp:451-538): the 17 deck columns per
  // cell, is_touch from check_localization (:430-445: an ALIVE cell with a
  // DEAD neighbour).  Text: doubles in fixed notation with 15 decimals, as
  // the reference's `fixed << setprecision(15)`; the state is written as an
  // integer where the reference streams the raw state byte, which its own
  // loader cannot read back.  Binary: per cell, in this order and without
  // padding, int32 index, int32 state, radius, ageb, agek, ca2p, x, y, z,
  // ca2p_avg (doubles), int32 rest_div_times, ex_fat, in_fat (doubles), char
  // is_touch, double spr_nat_len, int32 pair (-1 = none), int32 fix_origin
  // (109 bytes; :496-513).
  void output(const std::string& filename, bool binary_mode = false) const {
    const CellData d = download();
    OutRecord r;
    r.n = d.n;
    r.touch.assign((size_t)d.n, 0);
    for (int i = 0; i < d.n; i++)
      if (d.state[i] == ALIVE)
        for (int e = d.conn_ptr[i]; e < d.conn_ptr[i + 1]; e++)
          if (d.state[d.conn[e]] == DEAD) { r.touch[i] = 1; break; }
    r.state = d.state; r.rest_div_times = d.rest_div_times; r.pair = d.pair; r.fix_origin = d.fix_origin;
    r.radius = d.radius; r.ageb = d.ageb; r.agek = d.agek; r.ca2p = d.ca2p; r.x = d.x; r.y = d.y; r.z = d.z;
    r.ca2p_avg = d.ca2p_avg; r.ex_fat = d.ex_fat; r.in_fat = d.in_fat; r.spr_nat_len = d.spr_nat_len;
    write_output(r, filename, binary_mode);
  }

  // output() without stopping the step loop, for the snapshot every CUT
  // steps (proc.cpp:139-141): the record is packed on the device behind the
  // queued work and downloaded on a copy stream (epi_snapshot_begin) while
  // later steps run; at the next call on the same slot (the slots alternate)
  // or at flush_outputs() it is copied out and a writer thread writes the
  // file.  The files are those output() writes at the same point.
  void output_async(const std::string& filename, bool binary_mode = false) {
    const int slot = next_slot_;
    next_slot_ = (next_slot_ + 1) % EPI_SNAPSHOT_SLOTS;
    collect(slot);
    check(ctx_, epi_snapshot_begin(ctx_, 0, slot), "epi_snapshot_begin");
    pending_[slot] = Pending{filename, binary_mode, true};
  }
  // every file of output_async written; rethrows a writer's error
  void flush_outputs() {
    for (int s = 0; s < EPI_SNAPSHOT_SLOTS; s++) collect(s);
    for (auto& w : writers_) w.join();
    writers_.clear();
    std::exception_ptr e;
    {
      std::lock_guard<std::mutex> g(err_mu_);
      std::swap(e, writer_err_);
    }
    if (e) std::rethrow_exception(e);
  }

 private:
  struct Pending {
    std::string filename;
    bool binary = false, active = false;
  };
  // the slot's landed snapshot, copied out and handed to a writer thread
  void collect(int slot) {
    Pending& p = pending_[slot];
    if (!p.active) return;
    p.active = false;
    epi_snapshot s{};
    check(ctx_, epi_snapshot_wait(ctx_, slot, &s), "epi_snapshot_wait");
    OutRecord r;
    const size_t n = (size_t)s.n;
    r.n = s.n;
    r.state.assign(s.state, s.state + n); r.touch.assign(s.touch, s.touch + n);
    r.rest_div_times.assign(s.rest_div_times, s.rest_div_times + n);
    r.pair.assign(s.pair, s.pair + n); r.fix_origin.assign(s.fix_origin, s.fix_origin + n);
    for (auto [dst, src] : {std::make_pair(&r.radius, s.radius), std::make_pair(&r.ageb, s.ageb),
                            std::make_pair(&r.agek, s.agek), std::make_pair(&r.ca2p, s.ca2p),
                            std::make_pair(&r.x, s.x), std::make_pair(&r.y, s.y), std::make_pair(&r.z, s.z),
                            std::make_pair(&r.ca2p_avg, s.ca2p_avg), std::make_pair(&r.ex_fat, s.ex_fat),
                            std::make_pair(&r.in_fat, s.in_fat), std::make_pair(&r.spr_nat_len, s.spr_nat_len)})
      dst->assign(src, src + n);
    writers_.emplace_back([this, r = std::move(r), f = p.filename, b = p.binary]() {
      try {
        write_output(r, f, b);
      } catch (...) {
        std::lock_guard<std::mutex> g(err_mu_);
        if (!writer_err_) writer_err_ = std::current_exception();
      }
    });
  }
  Pending pending_[EPI_SNAPSHOT_SLOTS];
  int next_slot_ = 0;
  std::vector<std::thread> writers_;
  std::mutex err_mu_;
  std::exception_ptr writer_err_;

  epi_geometry geo_;
  epi_ctx* ctx_ = nullptr;
  long long device_ts_ = 0;
};

template <class T>
std::vector<T> FArr3D<T>::read() const {
  const epi_geometry& g = owner_->geometry();
  const size_t n = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  std::vector<T> out(n);
  epi_ctx* c = owner_->ctx();
  if constexpr (sizeof(T) == sizeof(double) && !std::is_pointer<T>::value) {
    std::vector<double> tmp(n);
    double* p[4] = {nullptr, nullptr, nullptr, nullptr};
    p[kind_ * 2 + buf_] = tmp.data();
    check(c, epi_get_fields(c, 0, p[0], p[1], p[2], p[3], nullptr, nullptr), "epi_get_fields");
    for (size_t i = 0; i < n; i++) out[i] = (T)tmp[i];
  } else if constexpr (std::is_pointer<T>::value) {
    std::vector<int32_t> m(n);
    check(c, epi_get_fields(c, 0, nullptr, nullptr, nullptr, nullptr, m.data(), nullptr), "epi_get_fields");
    // indices stand in for Cell pointers: -1 becomes nullptr, i becomes (T)(i + 1)
    for (size_t i = 0; i < n; i++) out[i] = m[i] < 0 ? nullptr : reinterpret_cast<T>((uintptr_t)m[i] + 1);
  } else {
    std::vector<uint8_t> m(n);
    check(c, epi_get_fields(c, 0, nullptr, nullptr, nullptr, nullptr, nullptr, m.data()), "epi_get_fields");
    for (size_t i = 0; i < n; i++) out[i] = (T)m[i];
  }
  return out;
}

template <class T>
void SwapData<FArr3D<T>>::read_binary(const std::string& filename) {
  const epi_geometry& g = owner_->geometry();
  const size_t n = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  std::vector<double> v(n);
  std::ifstream ifs(filename, std::ios_base::in | std::ios_base::binary);
  if (!ifs) throw Error(EPI_ERR_INVALID, "Field data input error. Filename:" + filename);
  ifs.read(reinterpret_cast<char*>(v.data()), sizeof(double) * n);
  if ((size_t)ifs.gcount() != sizeof(double) * n) throw Error(EPI_ERR_INVALID, "short field file: " + filename);
  epi_ctx* c = owner_->ctx();
  if (kind_ == 0)
    check(c, epi_set_fields(c, 0, v.data(), v.data(), nullptr, nullptr), "epi_set_fields");
  else
    check(c, epi_set_fields(c, 0, nullptr, nullptr, v.data(), v.data()), "epi_set_fields");
}

namespace detail {
inline void same_owner(const CellManager& cman, const CellManager* o, const char* what) {
  if (o != &cman) throw Error(EPI_ERR_INVALID, std::string(what) + ": grid of another CellManager");
}
}  // namespace detail

// ---- the call surface ------------------------------------------------------
inline void connect_cell(CellManager& cman) { check(cman.ctx(), epi_connect_cell(cman.ctx()), "connect_cell"); }
inline void cell_interaction(CellManager& cman) {
  check(cman.ctx(), epi_cell_interaction(cman.ctx()), "cell_interaction");
}
inline void pos_copy(CellManager& cman) { check(cman.ctx(), epi_pos_copy(cman.ctx()), "pos_copy"); }
inline void cell_state_renew(CellManager& cman) {
  check(cman.ctx(), epi_cell_state_renew(cman.ctx()), "cell_state_renew");
}
inline void cell_pos_periodic_fix(CellManager& cman) {
  check(cman.ctx(), epi_cell_pos_periodic_fix(cman.ctx()), "cell_pos_periodic_fix");
}
inline double calc_zzmax(CellManager& cman) {
  double z = 0;
  check(cman.ctx(), epi_calc_zzmax(cman.ctx(), &z), "calc_zzmax");
  return z;
}
inline void setup_map_lat(CellManager& cman, FArr3D<const Cell*>& cmap1, FArr3D<cmask_ty>& cmap2) {
  detail::same_owner(cman, cmap1.owner(), "setup_map_lat");
  detail::same_owner(cman, cmap2.owner(), "setup_map_lat");
  check(cman.ctx(), epi_setup_map_lat(cman.ctx()), "setup_map_lat");
}
inline void calc_ext_stim(SwapData<FArr3D<double>>& ext_stim, const FArr3D<const Cell*>& cmap1,
                          const FArr3D<cmask_ty>& cmap2, double zzmax) {
  CellManager* o = ext_stim.owner();
  if (!o || cmap1.owner() != o || cmap2.owner() != o)
    throw Error(EPI_ERR_INVALID, "calc_ext_stim: grids of different CellManagers");
  check(o->ctx(), epi_calc_ext_stim(o->ctx(), &zzmax), "calc_ext_stim");
}
inline void initialize_sc(CellManager& cman, double zzmax) {
  check(cman.ctx(), epi_initialize_sc(cman.ctx(), &zzmax), "initialize_sc");
}
inline void calc_ca2p(CellManager& cman, SwapData<FArr3D<double>>& ATP, const FArr3D<double>& ext_stim_first,
                      const FArr3D<const Cell*>& cmap1, const FArr3D<cmask_ty>& cmap2, double zzmax) {
  detail::same_owner(cman, ATP.owner(), "calc_ca2p");
  detail::same_owner(cman, ext_stim_first.owner(), "calc_ca2p");
  detail::same_owner(cman, cmap1.owner(), "calc_ca2p");
  detail::same_owner(cman, cmap2.owner(), "calc_ca2p");
  check(cman.ctx(), epi_calc_ca2p(cman.ctx(), &zzmax), "calc_ca2p");
}

// ---- loading ---------------------------------------------------------------
// CellManager::_load_from_file (cellmanager.cpp:198-323): 17 columns per
// row, reading stops at the first BLANK row, rows come MEMB, DER, rest.
inline void load_deck(CellManager& cman, const std::string& path) {
  std::ifstream in(path);
  if (!in) throw Error(EPI_ERR_INVALID, "load failed: " + path);
  std::vector<uint8_t> st, mal;
  std::vector<int32_t> fo, rdt, pr;
  std::vector<double> x, y, z, r, ca, cav, ip3, exi, agek, ageb, exf, inf, spr;
  std::string line;
  int nmemb = 0, nder = 0, phase = 0;
  while (std::getline(in, line)) {
    int idx, state, div_times, touch, pair, stem;
    double rad, ab, ak, c2, px, py, pz, cavg, ef, fat, sl;
    if (std::sscanf(line.c_str(), "%d %d %lf %lf %lf %lf %lf %lf %lf %lf %d %lf %lf %d %lf %d %d", &idx, &state,
                    &rad, &ab, &ak, &c2, &px, &py, &pz, &cavg, &div_times, &ef, &fat, &touch, &sl, &pair,
                    &stem) != 17)
      throw Error(EPI_ERR_INVALID, "bad deck row: " + line);
    if (state == BLANK) break;
    if (state == DER && rad != 1.4) throw Error(EPI_ERR_INVALID, "radii of DER not consistent");
    if (state == MEMB && rad != 1.0) throw Error(EPI_ERR_INVALID, "radii of MEMB not consistent");
    if (phase == 0 && state != MEMB) phase = 1;
    if (phase == 1 && state != DER) phase = 2;
    if (phase > 0 && state == MEMB) throw Error(EPI_ERR_INVALID, "non phase0 memb");
    if (phase > 1 && state == DER) throw Error(EPI_ERR_INVALID, "non phase1 der");
    nmemb += state == MEMB;
    nder += state == DER;
    st.push_back((uint8_t)state); fo.push_back(stem); rdt.push_back(div_times); pr.push_back(pair);
    x.push_back(px); y.push_back(py); z.push_back(pz); r.push_back(rad); ca.push_back(c2);
    cav.push_back(cavg); ip3.push_back(0.0); exi.push_back(0.97); agek.push_back(ak); ageb.push_back(ab);
    exf.push_back(ef); inf.push_back(fat); spr.push_back(sl);
    mal.push_back(stem < 0 ? 1 : 0);   // stem_orig_id < MALIG_NUM (define.h:92)
  }
  const int n = (int)st.size();
  for (int i = 0; i < n; i++)
    if (pr[i] >= n) pr[i] = -1;
  epi_cell_input ci{};
  ci.n = n; ci.nmemb = nmemb; ci.nder = nder;
  ci.state = st.data(); ci.fix_origin = fo.data(); ci.x = x.data(); ci.y = y.data(); ci.z = z.data();
  ci.radius = r.data(); ci.ca2p = ca.data(); ci.ca2p_avg = cav.data(); ci.ip3 = ip3.data();
  ci.ex_inert = exi.data(); ci.agek = agek.data(); ci.ageb = ageb.data(); ci.ex_fat = exf.data();
  ci.in_fat = inf.data(); ci.spr_nat_len = spr.data(); ci.rest_div_times = rdt.data();
  ci.is_malignant = mal.data(); ci.pair = pr.data();
  check(cman.ctx(), epi_load_cells(cman.ctx(), 0, &ci), "epi_load_cells");
}

// load_uvp and load_w (cell_init.cpp:47-109) in one pass: per-cell
// "idx ca2p ex_inert IP3" rows, then per cell "idx len" and len rows
// "neighbour gj"
inline void load_last(CellManager& cman, const std::string& ld_uvp, const std::string& ld_w) {
  const CellData d = cman.download();
  const size_t n = (size_t)d.n;
  std::vector<double> ca(d.ca2p), ip3(d.ip3), exi(d.ex_inert), gj(d.gj);
  {
    std::ifstream f(ld_uvp);
    if (!f) throw Error(EPI_ERR_INVALID, "Last uvp data open error. Filename:" + ld_uvp);
    std::string ln;
    size_t cnt = 0;
    while (std::getline(f, ln)) {
      int i;
      double a, b, c;
      if (std::sscanf(ln.c_str(), "%d %lf %lf %lf", &i, &a, &b, &c) != 4 || i < 0 || (size_t)i >= n)
        throw Error(EPI_ERR_INVALID, "bad uvp row: " + ln);
      ca[i] = a; exi[i] = b; ip3[i] = c;
      cnt++;
    }
    if (cnt != n) throw Error(EPI_ERR_INVALID, "Num of cell mismatch in " + ld_uvp);
  }
  if (!ld_w.empty()) {
    std::ifstream f(ld_w);
    if (!f) throw Error(EPI_ERR_INVALID, "Last w data open error. Filename:" + ld_w);
    std::string ln;
    size_t cnt = 0;
    while (std::getline(f, ln)) {
      int i, len;
      if (std::sscanf(ln.c_str(), "%d %d", &i, &len) != 2 || i < 0 || (size_t)i >= n)
        throw Error(EPI_ERR_INVALID, "bad GJ data format (index)");
      for (int k = 0; k < len; k++) {
        int o;
        double v;
        if (!std::getline(f, ln) || std::sscanf(ln.c_str(), "%d %lf", &o, &v) != 2)
          throw Error(EPI_ERR_INVALID, "bad GJ data format");
        // an entry for a current neighbour sets that edge; others are the
        // stale entries the reference keeps at their value until clean_up
        for (int e = d.conn_ptr[i]; e < d.conn_ptr[i + 1]; e++)
          if (d.conn[e] == o) gj[e] = v;
      }
      cnt++;
    }
    if (cnt != n) throw Error(EPI_ERR_INVALID, "Num of cell mismatch in " + ld_w);
  }
  check(cman.ctx(), epi_set_cell_values(cman.ctx(), 0, ca.data(), ip3.data(), exi.data(), nullptr, gj.data()),
        "epi_set_cell_values");
}

// cman_init (cell_init.cpp:132-142): load the deck, build the lists, then
// set the start values (_cman_value_init) or the last run's values
inline void cman_init(CellManager& cells, const std::string& init_data_path, bool use_last = false,
                      const std::string& ld_uvp = "", const std::string& ld_w = "") {
  load_deck(cells, init_data_path);
  connect_cell(cells);
  if (use_last)
    load_last(cells, ld_uvp, ld_w);
  else
    check(cells.ctx(), epi_value_init(cells.ctx(), 0), "epi_value_init");
}

// ---- saving the last state (proc.cpp:53-84) ---------------------------------
// the file names of define.h:219-224
constexpr const char* last_data_uvp_name = "last_data_uvp";
constexpr const char* last_data_cell_name = "last_data_cell";
constexpr const char* last_data_w_name = "last_data_w_alt";
constexpr const char* last_data_a_name = "last_data_a_alt";
constexpr const char* last_data_B_name = "last_data_B_alt";

// output_uvp (proc.cpp:53-60): "index ca2p ex_inert IP3" per cell, the rows
// load_last reads
inline void output_uvp(const CellManager& cman, const std::string& filename) {
  const CellData d = cman.download();
  FILE* f = std::fopen(filename.c_str(), "w");
  if (!f) throw Error(EPI_ERR_INVALID, "uvp output error. Filename:" + filename);
  for (int i = 0; i < d.n; i++) std::fprintf(f, "%d %1.14e %1.14e %1.14e\n", i, d.ca2p[i], d.ex_inert[i], d.ip3[i]);
  std::fclose(f);
}

// output_w (proc.cpp:63-72): per cell "index count", then count rows
// "neighbour gj".  The reference walks its per-cell gj map, which also holds
// entries of past neighbours (reset to gj_init at every rebuild until
// clean_up erases them); here the entries are the current neighbours, in
// list order.  A past neighbour's entry would be read back as gj_init and
// reset to it at the next rebuild, so the two files load to the same state.
inline void output_w(const CellManager& cman, const std::string& filename) {
  const CellData d = cman.download();
  FILE* f = std::fopen(filename.c_str(), "w");
  if (!f) throw Error(EPI_ERR_INVALID, "w output error. Filename:" + filename);
  for (int i = 0; i < d.n; i++) {
    std::fprintf(f, "%d %d\n", i, d.conn_ptr[i + 1] - d.conn_ptr[i]);
    for (int e = d.conn_ptr[i]; e < d.conn_ptr[i + 1]; e++) std::fprintf(f, "%d %1.14e\n", d.conn[e], d.gj[e]);
  }
  std::fclose(f);
}

// output_field_data (proc.cpp:74-84): the cells as text, uvp, w and the first
// buffers of ATP and ext_stim as raw [x][y][z] doubles, under dir (the
// reference writes them to the working directory)
inline void output_field_data(CellManager& cman, const FArr3D<double>& ATP_first,
                              const FArr3D<double>& ext_stim_first, const std::string& dir = ".") {
  detail::same_owner(cman, ATP_first.owner(), "output_field_data");
  detail::same_owner(cman, ext_stim_first.owner(), "output_field_data");
  const std::string p = dir.empty() ? std::string() : dir + "/";
  cman.output(p + last_data_cell_name);
  output_uvp(cman, p + last_data_uvp_name);
  output_w(cman, p + last_data_w_name);
  ATP_first.output_binary(p + last_data_a_name);
  ext_stim_first.output_binary(p + last_data_B_name);
}

}  // namespace epi_amd
