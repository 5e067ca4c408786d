// proc.cpp -- the epi5_MT driver loop (proc.cpp:90-173 of the reference)
// written against the C++ facade include/epi_amd.hpp: the same calls in the
// same order, each one a launch sequence on the GPU.
//
//   proc <deck> <steps> [--seed S] [--force-cornif] [--ca-every K]
//        [--geometry nmx nmy nx ny nz lx ly lz max_cells] [--dump FILE]
//        [--output FILE] [--output-binary FILE] [--snapshots DIR [--cut N]]
//        [--save-last DIR] [--use-last DIR] [--params FILE] [--params-out FILE]
//   proc --init-gen FILE [--nfix N] [--der D] [--params FILE]
//
// Defaults: define.h's geometry (NMX 100, NMY_tri 114, 100x100x200 grid on
// 50x50x100), seed 1, the reference's Ca gating only.  --dump writes the
// final n, then x, y, z, ca2p, agek as raw doubles; --output writes the
// 17-column text of CellManager::output.  --snapshots writes that text to
// DIR/<i / N> before every step i with i % N == 0 (output_cell_data,
// proc.cpp:45-51, 139-141; N defaults to CUT = 1000): the numbered series the
// reference's viewers read (epi_vis_lightweight/main.cpp:493-520).  Those
// snapshots are asynchronous (CellManager::output_async): the record is
// packed on the GPU and downloaded and written while later steps run.
// --output-binary writes CellManager::output's binary form.  --save-last
// writes output_field_data's five files (last_data_cell, last_data_uvp,
// last_data_w_alt, last_data_a_alt, last_data_B_alt) and appends the
// removed cells' records to DIR/cell_stat (append_stat_data) every 10 N steps
// (proc.cpp:142-152) and after the last step; --use-last starts from such a
// set, with DIR/last_data_cell given as the deck: load_uvp / load_w and both
// buffers of ATP and ext_stim from the binaries (cman_init with use_last,
// proc.cpp:98-110).  --params reads a CalcParams key=value file (every key,
// as epi_intg/Params.cpp:91-96 requires; epi_intg/ppsss is one) whose grid
// and box become the run's unless --geometry is given; the driver then uses
// its DT_Cell, T_TURNOVER and NUM_SC_INIT.  --params-out writes the run's
// parameters in the same form.  --init-gen writes the deck of epi_intg's
// init_gen(N = 16, D = 8) under the defaults or a --params file.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/epi_amd.hpp"

using namespace epi_amd;

namespace {

void cell_dynamics(CellManager& cellset) {   // proc.cpp:22-32
  cell_interaction(cellset);
  pos_copy(cellset);
  cell_state_renew(cellset);
  cell_pos_periodic_fix(cellset);
  connect_cell(cellset);
}

int usage() {
  std::fprintf(stderr,
               "usage: proc <deck> <steps> [--seed S] [--force-cornif] [--ca-every K]\n"
               "            [--geometry nmx nmy nx ny nz lx ly lz max_cells] [--dump FILE] [--output FILE]\n"
               "            [--output-binary FILE] [--snapshots DIR [--cut N]] [--save-last DIR]\n"
               "            [--use-last DIR] [--params FILE] [--params-out FILE]\n"
               "       proc --init-gen FILE [--nfix N] [--der D] [--params FILE]\n");
  return 2;
}

// epi_intg's "i" mode (main.cpp:103-112): init_gen_output(FILE, 16, 8), the
// cells written as CellManager::output's text (calc/cpu2/CellManager.cpp:
// 210-229: fixed, 15 decimals)
int init_gen_mode(int argc, char** argv) {
  const std::string out = argv[2];
  int nfix = 16, der = 8;
  std::string pin;
  for (int i = 3; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--nfix" && i + 1 < argc) nfix = std::atoi(argv[++i]);
    else if (a == "--der" && i + 1 < argc) der = std::atoi(argv[++i]);
    else if (a == "--params" && i + 1 < argc) pin = argv[++i];
    else return usage();
  }
  epi_params p;
  epi_params_default(&p);
  char err[512];
  if (!pin.empty() && epi_params_load(pin.c_str(), 1, &p, err, (int32_t)sizeof err) != EPI_OK) {
    std::fprintf(stderr, "%s\n", err);
    return 1;
  }
  epi_gen_cells g{};
  int rc = epi_init_gen(&p, nfix, der, 0, &g, err, (int32_t)sizeof err);
  if (rc != EPI_OK && rc != EPI_ERR_CAPACITY) {
    std::fprintf(stderr, "%s\n", err);
    return 1;
  }
  const int n = g.n;
  std::vector<uint8_t> st(n);
  std::vector<int32_t> fo(n);
  std::vector<double> x(n), y(n), z(n), r(n);
  g.state = st.data(); g.fix_origin = fo.data();
  g.x = x.data(); g.y = y.data(); g.z = z.data(); g.radius = r.data();
  rc = epi_init_gen(&p, nfix, der, n, &g, err, (int32_t)sizeof err);
  if (rc != EPI_OK) {
    std::fprintf(stderr, "%s\n", err);
    return 1;
  }
  FILE* f = std::fopen(out.c_str(), "w");
  if (!f) {
    std::fprintf(stderr, "Output file creation error. Filename:%s\n", out.c_str());
    return 1;
  }
  for (int i = 0; i < n; i++)
    std::fprintf(f, "%d %d %.15f %.15f %.15f %.15f %.15f %.15f %.15f %.15f %d %.15f %.15f %d %.15f %d %d\n", i,
                 (int)st[i], r[i], 0.0, 0.0, 0.0, x[i], y[i], z[i], 0.0, 0, 0.0, 0.0, 0, 0.0, -1, (int)fo[i]);
  std::fclose(f);
  std::printf("init_gen: %d cells (sheet %d x %d, %d dermis, %d FIX) -> %s\n", n, g.nmx, g.nmy, g.nder, nfix,
              out.c_str());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return usage();
  if (std::string(argv[1]) == "--init-gen") return init_gen_mode(argc, argv);
  const std::string deck = argv[1];
  const long steps = std::atol(argv[2]);
  epi_geometry g{100, 114, 100, 100, 200, 50.0, 50.0, 100.0, 100 * 114 + 50000, 0, 0};
  unsigned long long seed = 1;
  bool force_cornif = false;
  long ca_every = 0;
  std::string dump, output, output_bin, snapdir, save_dir, last_dir, params_in, params_out;
  bool geometry_given = false;
  long cut = 1000;   // CUT (define.h)
  for (int i = 3; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--seed" && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
    else if (a == "--force-cornif") force_cornif = true;
    else if (a == "--ca-every" && i + 1 < argc) ca_every = std::atol(argv[++i]);
    else if (a == "--dump" && i + 1 < argc) dump = argv[++i];
    else if (a == "--output" && i + 1 < argc) output = argv[++i];
    else if (a == "--output-binary" && i + 1 < argc) output_bin = argv[++i];
    else if (a == "--save-last" && i + 1 < argc) save_dir = argv[++i];
    else if (a == "--use-last" && i + 1 < argc) last_dir = argv[++i];
    else if (a == "--params" && i + 1 < argc) params_in = argv[++i];
    else if (a == "--params-out" && i + 1 < argc) params_out = argv[++i];
    else if (a == "--snapshots" && i + 1 < argc) snapdir = argv[++i];
    else if (a == "--cut" && i + 1 < argc) cut = std::atol(argv[++i]);
    else if (a == "--geometry" && i + 9 < argc) {
      geometry_given = true;
      g.nmx = std::atoi(argv[++i]); g.nmy = std::atoi(argv[++i]);
      g.nx = std::atoi(argv[++i]); g.ny = std::atoi(argv[++i]); g.nz = std::atoi(argv[++i]);
      g.lx = std::atof(argv[++i]); g.ly = std::atof(argv[++i]); g.lz = std::atof(argv[++i]);
      g.max_cells = std::atoi(argv[++i]);
    } else return usage();
  }
  try {
    epi_params prm;
    epi_params_default(&prm);
    if (!params_in.empty()) {
      // a CalcParams file (epi_intg/ppsss): every key must be there; its grid
      // and box are the run's (unless --geometry set them)
      char err[512];
      const int rc = epi_params_load(params_in.c_str(), 1, &prm, err, (int32_t)sizeof err);
      if (rc != EPI_OK) throw Error(rc, err);
      if (!geometry_given) {
        g.nx = prm.NX; g.ny = prm.NY; g.nz = prm.NZ;
        g.lx = prm.LX; g.ly = prm.LY; g.lz = prm.LZ;
      }
    }
    CellManager cman(g);
    {
      epi_params p = cman.params();
      const epi_params geo = p;
      p = prm;
      p.LX = geo.LX; p.LY = geo.LY; p.LZ = geo.LZ;
      p.NX = geo.NX; p.NY = geo.NY; p.NZ = geo.NZ;
      cman.set_params(p);
      prm = p;
    }
    if (!params_out.empty()) cman.save_params(params_out);
    if (last_dir.empty()) {
      cman_init(cman, deck);
    } else {
      cman_init(cman, deck, true, last_dir + "/" + last_data_uvp_name, last_dir + "/" + last_data_w_name);
      cman.atp().read_binary(last_dir + "/" + last_data_a_name);
      cman.ext_stim().read_binary(last_dir + "/" + last_data_B_name);
    }
    // genrand_real replay: a generous stream for the run
    cman.seed_random(seed, (long long)(steps + 64) * ((long long)cman.size() / 4 + 64) + 100000);
    auto ATP = cman.atp();
    auto ext_stim = cman.ext_stim();
    auto cell_map1 = cman.cell_map1();
    auto cell_map2 = cman.cell_map2();
    bool flg_forced_sc = force_cornif;
    int num_sc = 0;
    double zzmax = 0;
    std::printf("current cell num:%zu\n", cman.size());
    const auto start = std::chrono::steady_clock::now();
    long ca_solves = 0;
    for (long i = 0; i < steps; i++) {   // proc.cpp:131-173
      cman.current_timestep = (unsigned long long)i;
      if (i % 100 == 0) {
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        std::printf("loop:%ld elapsed[sec]:%lf\n", i, s);
        std::fflush(stdout);
      }
      // taken on the device and written by a writer thread while the loop goes on
      if (!snapdir.empty() && cut > 0 && i % cut == 0) cman.output_async(snapdir + "/" + std::to_string(i / cut));
      if (!save_dir.empty() && cut > 0 && i % (cut * 10) == 0 && i > 0) {   // proc.cpp:142-152
        cman.clean_up();
        output_field_data(cman, ATP.first(), ext_stim.first(), save_dir);
        cman.append_stat_data(save_dir + "/cell_stat");
      }
      cell_dynamics(cman);
      zzmax = calc_zzmax(cman);
      setup_map_lat(cman, cell_map1, cell_map2);
      calc_ext_stim(ext_stim, cell_map1, cell_map2, zzmax);
      if (i * prm.DT_Cell > prm.T_TURNOVER && flg_forced_sc) {   // proc.cpp:158-163
        flg_forced_sc = false;
        std::printf("forced cornif\n");
        initialize_sc(cman, zzmax);
        num_sc = (int)prm.NUM_SC_INIT;
      }
      const bool forced = ca_every > 0 && (i + 1) % ca_every == 0;
      if (cman.should_calc_ca() || num_sc > 0 || forced) {
        calc_ca2p(cman, ATP, ext_stim.first(), cell_map1, cell_map2, zzmax);
        ca_solves++;
        if (num_sc > 0) num_sc--;
        cman.ca_calc_condition_reset();
      }
    }
    cman.flush_outputs();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::printf("done: %ld steps, %zu cells, %ld Ca solves, %.3f s\n", steps, cman.size(), ca_solves, s);
    if (!output.empty()) cman.output(output);
    if (!output_bin.empty()) cman.output(output_bin, true);
    if (!save_dir.empty()) {
      output_field_data(cman, ATP.first(), ext_stim.first(), save_dir);
      cman.append_stat_data(save_dir + "/cell_stat");
    }
    if (!dump.empty()) {
      const CellData d = cman.download();
      FILE* f = std::fopen(dump.c_str(), "wb");
      if (!f) throw Error(EPI_ERR_INVALID, "cannot write " + dump);
      const double n = d.n;
      std::fwrite(&n, sizeof(double), 1, f);
      for (const auto* v : {&d.x, &d.y, &d.z, &d.ca2p, &d.agek}) std::fwrite(v->data(), sizeof(double), v->size(), f);
      std::fclose(f);
    }
  } catch (const Error& e) {
    std::fprintf(stderr, "epi_amd error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
