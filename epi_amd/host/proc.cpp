// proc.cpp -- the epi5_MT driver loop (proc.cpp:90-173 of the reference)
// written against the C++ facade include/epi_amd.hpp: the same calls in the
// same order, each one a launch sequence on the GPU.
//
//   proc <deck> <steps> [--seed S] [--force-cornif] [--ca-every K]
//        [--geometry nmx nmy nx ny nz lx ly lz max_cells] [--dump FILE]
//        [--output FILE] [--output-binary FILE] [--snapshots DIR [--cut N]]
//        [--save-last DIR] [--use-last DIR]
//
// Defaults: define.h's geometry (NMX 100, NMY_tri 114, 100x100x200 grid on
// 50x50x100), seed 1, the reference's Ca gating only.  --dump writes the
// final n, then x, y, z, ca2p, agek as raw doubles; --output writes the
// 17-column text of CellManager::output.  --snapshots writes that text to
// DIR/<i / N> before every step i with i % N == 0 (output_cell_data,
// proc.cpp:45-51, 139-141; N defaults to CUT = 1000): the numbered series the
// reference's viewers read (epi_vis_lightweight/main.cpp:493-520).
// --output-binary writes CellManager::output's binary form.  --save-last
// writes output_field_data's five files (last_data_cell, last_data_uvp,
// last_data_w_alt, last_data_a_alt, last_data_B_alt) and appends the
// removed cells' records to DIR/cell_stat (append_stat_data) every 10 N steps
// (proc.cpp:142-152) and after the last step; --use-last starts from such a
// set, with DIR/last_data_cell given as the deck: load_uvp / load_w and both
// buffers of ATP and ext_stim from the binaries (cman_init with use_last,
// proc.cpp:98-110).
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
               "            [--use-last DIR]\n");
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return usage();
  const std::string deck = argv[1];
  const long steps = std::atol(argv[2]);
  epi_geometry g{100, 114, 100, 100, 200, 50.0, 50.0, 100.0, 100 * 114 + 50000, 0, 0};
  unsigned long long seed = 1;
  bool force_cornif = false;
  long ca_every = 0;
  std::string dump, output, output_bin, snapdir, save_dir, last_dir;
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
    else if (a == "--snapshots" && i + 1 < argc) snapdir = argv[++i];
    else if (a == "--cut" && i + 1 < argc) cut = std::atol(argv[++i]);
    else if (a == "--geometry" && i + 9 < argc) {
      g.nmx = std::atoi(argv[++i]); g.nmy = std::atoi(argv[++i]);
      g.nx = std::atoi(argv[++i]); g.ny = std::atoi(argv[++i]); g.nz = std::atoi(argv[++i]);
      g.lx = std::atof(argv[++i]); g.ly = std::atof(argv[++i]); g.lz = std::atof(argv[++i]);
      g.max_cells = std::atoi(argv[++i]);
    } else return usage();
  }
  try {
    CellManager cman(g);
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
      if (!snapdir.empty() && cut > 0 && i % cut == 0) cman.output(snapdir + "/" + std::to_string(i / cut));
      if (!save_dir.empty() && cut > 0 && i % (cut * 10) == 0 && i > 0) {   // proc.cpp:142-152
        cman.clean_up();
        output_field_data(cman, ATP.first(), ext_stim.first(), save_dir);
        cman.append_stat_data(save_dir + "/cell_stat");
      }
      cell_dynamics(cman);
      zzmax = calc_zzmax(cman);
      setup_map_lat(cman, cell_map1, cell_map2);
      calc_ext_stim(ext_stim, cell_map1, cell_map2, zzmax);
      if (i * 0.01 > 6000.0 && flg_forced_sc) {   // i*DT_Cell > T_TURNOVER
        flg_forced_sc = false;
        std::printf("forced cornif\n");
        initialize_sc(cman, zzmax);
        num_sc = 1;   // NUM_SC_INIT
      }
      const bool forced = ca_every > 0 && (i + 1) % ca_every == 0;
      if (cman.should_calc_ca() || num_sc > 0 || forced) {
        calc_ca2p(cman, ATP, ext_stim.first(), cell_map1, cell_map2, zzmax);
        ca_solves++;
        if (num_sc > 0) num_sc--;
        cman.ca_calc_condition_reset();
      }
    }
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
