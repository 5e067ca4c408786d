// launch.h -- host-side launchers of the kernel families (one per .hip file)
// and the optional per-kernel event timer.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

namespace epi {
struct Tis;
struct Status;

enum KernelId {
  K_FORCES, K_FORCES_LONG, K_RENEW_FLAGS, K_RENEW_SCAN, K_RENEW_APPLY, K_REMOVE, K_PAIR_DISPERSE,
  K_INIT_SC, K_PERIODIC_FIX, K_BUCKET_INSERT, K_BUCKET_SORT, K_SWEEP,
  K_FINISH, K_SNAPSHOT, K_ZZ_SET, K_ZZMAX, K_MAP_CLEAR, K_MAP_STAMP, K_MAP_HALO,
  K_EXT_STIM, K_CA_BEGIN, K_CA_VOX, K_CA_CELLS, K_ATP, K_CA_END, K_SW_RESET,
  K_RD_SNAPSHOT, K_RD_RESTORE, K_RD_PUT_GJ, K_RD_STATE, K_RD_SEED, K_RD_STEP,
  K_RD_GJ0, K_RD_OWN, K_MECH_STATE, K_MECH_SEED, K_MECH_BWD_BEGIN, K_MECH_BWD_PAIRS, K_MECH_BWD_MEMB,
  K_CA_FUSED, K_CA_CORNERS, K_SWEEP_WAVE, K_WL, K_CA_SWAP, K_OUT_PACK, K_COUNT
};

// ints of a context's work-list buffer: the offsets and per-tissue iz_bound
// of the B stencil's list, the rest for the Ca solve's lists
constexpr long long WL_INTS(int ntis) { return 8LL * (ntis + 1); }

// Launch shapes that the autotuner may change (values checked by the host).
struct Tuning {
  int b_zb = 4, b_rows = 4;   // k_ext_stim: z planes per tile, 64-lane rows per block (points per tile)
  int a_zb = 4, a_rows = 4;   // k_atp: the same (a fused solve's tiles are 128 points)
  int ca_group = 16;          // k_ca_cells: lanes per cell
  int xcd = 0;                // k_ca_cells, k_atp: XCD-contiguous workgroup order
  int b_xcd = 0;              // k_ext_stim: its tile list in one contiguous run per XCD
  int sweep_wave = 0;         // k_sweep: 1, one wave per cell (k_sweep_wave)
  int side_fork = 0;          // 1: the side stream forks after the whole build
  int rd_group = 8;           // k_rd_step: lanes per cell
  int ca_fuse = 1;            // Ca solves: ATP of a sub-step and cells of the next in one launch
                              // (0 never, 1 always, 2 for solves of at most 131072 cells)
  int ca_group_small = 8;     // k_ca_cells lanes per cell in the solves that ca_fuse 2 fuses
  int map_lanes = 1;          // k_map_stamp: lanes per cell (1, 4 or 16; a box slice split between them)
  int step_graph = 0;         // epi_step: 1 replays the renewal and build from a captured graph
};

// Event timer: when the bit of a kernel id is set, its launches carry a
// start and a stop event that the dispatch itself records
// (hipExtLaunchKernelGGL), so the elapsed time is the kernel's own duration
// without the cost of separate event packets (see host.hip).
bool prof_on(int k, bool* counted);
void prof_events(hipEvent_t* a, hipEvent_t* b);
void prof_add(int k, hipEvent_t a, hipEvent_t b, bool counted);

#define EPI_LAUNCH(kid, kernel, grid, block, shmem, stream, ...)                              \
  do {                                                                                        \
    bool cnt_ = false;                                                                        \
    if (::epi::prof_on(kid, &cnt_)) {                                                         \
      hipEvent_t a_ = nullptr, b_ = nullptr;                                                  \
      ::epi::prof_events(&a_, &b_);                                                           \
      hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, a_, b_, 0, __VA_ARGS__);      \
      ::epi::prof_add(kid, a_, b_, cnt_);                                                     \
    } else {                                                                                  \
      hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                    \
    }                                                                                         \
  } while (0)

// Rows of up to FORCES_ROW_CAP entries go to the cooperative kernel (whose
// LDS slots per cell, 40, 48 or 64, follow max_len), longer ones to a
// one-lane-per-cell kernel, launched only when max_len exceeds the slots.
constexpr int FORCES_ROW_CAP = 64;
void launch_forces(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int max_memb, int max_other, int max_len);
// the two halves of launch_forces: the cooperative kernel for rows of up to
// the returned number of slots (chosen from max_len), and the long rows
int launch_forces_rows(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int max_other, int max_len);
void launch_forces_long(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int max_other, int cap);
void launch_renew(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int cpar, int max_n);
void launch_initialize_sc(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int pb, int max_n);
void launch_periodic_fix(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n);
// the bucket insertion (with the periodic fix) and the rest of the build,
// for callers that run the voxel maps beside the rest
void launch_connect_insert(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n, int fix);
// with h_status, k_snapshot also copies every tissue's status there
void launch_connect_rest(hipStream_t s, const Tis* dT, int ntis, int pb, int nb, int max_n, int sweep_wave,
                         Status* h_status);
void launch_connect(hipStream_t s, const Tis* dT, int ntis, int pb, int nb, int max_n, int max_buckets, int fix,
                    int sweep_wave);
void launch_zzmax(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n);
// A snapshot's output record (CellManager::output's fields): a header of
// OUT_HDR bytes whose first int is the device's cell count, then m >= n
// slots per array: the doubles radius, ageb, agek, ca2p, x, y, z, ca2p_avg,
// ex_fat, in_fat, spr_nat_len, the ints rest_div_times, pair, fix_origin and
// the bytes state, touch.  m is a multiple of 32, so every array is aligned.
constexpr int OUT_NF64 = 11, OUT_NI32 = 3, OUT_NU8 = 2;
constexpr long long OUT_HDR = 256;
constexpr long long out_record_bytes(long long m) { return OUT_HDR + m * (8 * OUT_NF64 + 4 * OUT_NI32 + OUT_NU8); }

// the output record of one tissue's first n cells into out, positions from
// buffer pb, Ca2+ from cp, lists from cb
void launch_out_pack(hipStream_t s, const Tis* dT, int tissue, int pb, int cp, int cb, int n, long long m, char* out);
void launch_zz_set(hipStream_t s, const Tis* dT, int ntis, double v);
void launch_maps(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n, long long max_gsize, int max_halo,
                 const Tuning& tu);
// the B stencil over a work list of (tissue, tile) built on the device from
// zzmax (d_wl: 2 ntis + 1 ints); tiles_est sizes the grid (any value covers
// the list: the blocks walk it grid-stride)
void launch_ext_stim(hipStream_t s, const Tis* dT, int ntis, int bp, long long tiles_est, int* d_wl,
                     const Tuning& tu);
void launch_ca_begin(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int cb, int max_n, long long max_box);
struct RdTis;
// A Ca solve's work, flattened over its tissues (built on the host once per
// solve, device pointers): blocks of the cell kernels map to (tissue, block
// of that tissue, blocks of that tissue) through cmap (ncb entries); the ATP
// grid is a list of tiles of `abs` plane points by a_zb planes, aoff[i]
// being the first tile of the list's i-th tissue (aoff[ntis] = atiles), with
// its iz_bound aizb[i] and tissue id atis[i].
// A list of one tissue (a single-tissue solve) also carries that tissue id
// and iz_bound by value (t0, izb0): the kernels then read no list memory.
struct CaList {
  const int* aoff;
  const int* aizb;
  const int* atis;
  const int4* cmap;
  int ntis, ncb, atiles, abs;
  int t0, izb0;
};
// points per ATP tile of a solve: the fused kernel's block, or the ATP
// kernel's 64 a_rows
int ca_list_bs(const Tuning& tu, bool fused);
// sv: a reaction-diffusion window's reverse-pass storage; the cell kernel
// then also saves the state before the sub-step into its segment slot
// `slot`; cells_only skips the ATP update (the last state of a segment)
void launch_ca_substep(hipStream_t s, const Tis* dT, const CaList& cl, int cb, int cp, int ap, int bp,
                       const Tuning& tu, const RdTis* sv = nullptr, int slot = 0, bool cells_only = false);
// n_sub sub-steps (fused when tu.ca_fuse), flipping cp and ap after each
void launch_ca_substeps(hipStream_t s, const Tis* dT, const CaList& cl, int cb, int& cp, int& ap, int bp,
                        int n_sub, const Tuning& tu);
void launch_ca_touch(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int cb, int max_n, long long max_box);
void launch_ca_diffu_zero(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int max_n);
void launch_ca_end(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int max_n);
// the listed tissues' two Ca, IP3 and ATP buffers trade contents (after a
// solve of part of a context with an odd sub-step count, whose buffer
// parities are shared by every tissue)
void launch_ca_swap(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int max_n, long long max_gsize);
// frozen reaction-diffusion window (k_adjoint.hip)
void launch_rd_snapshot(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int cb, int cp, int ap, long long items);
void launch_rd_own(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int max_n);
void launch_rd_restore(hipStream_t s, const Tis* dT, RdTis* dR, const double* gj0, int ntis, int cb, int cp, int ap,
                       int use_kpa, long long items);
void launch_rd_put_gj(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int cb, long long items);
void launch_rd_state(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int slot, int dir, int cb, int cp, int ap,
                     long long items);
void launch_rd_seed(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int cp, int cb, int q, long long items);
void launch_rd_backward_substep(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int slot, int cb, int bp, int q,
                                int qa, int max_n, long long max_box, int rg);
// frozen mechanics window (k_mech.hip)
struct MechTis;
void launch_mech_state(hipStream_t s, const Tis* dT, const MechTis* dM, int ntis, int slot, int dir, int pb,
                       int max_n);
void launch_mech_seed(hipStream_t s, const Tis* dT, MechTis* dM, int ntis, int pb, int q, int max_n);
void launch_mech_backward_step(hipStream_t s, const Tis* dT, MechTis* dM, int ntis, int slot, int pb, int cb, int q,
                               int max_n);
void launch_rd_gj0(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int cb, int q, long long items);
}  // namespace epi
