// epi_dev.h -- device-side data layout and the scalar laws of the epi5_MT hot
// path, for CDNA4 (gfx950).
//
// Layout: one Tis descriptor per tissue in device memory; kernels take the
// descriptor array and use blockIdx.y as the tissue index, so a single
// launch advances every tissue of a batch.  Cells are SoA, neighbour lists
// are fixed-stride rows (stride = ns entries per cell), grids are stored
// z-major / x-fastest with a padded x pitch.
//
// Every FP64 expression keeps the operand order of the reference so that a
// build without FP contraction (-ffp-contract=off) reproduces it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "params.h"

namespace epi {

// cell states, epi5_MT/define.h:36-47
enum : uint8_t { ALIVE = 0, DEAD = 1, DISA = 2, UNUSED = 3, FIX = 4, BLANK = 5,
                 DER = 6, MUSUME = 7, AIR = 8, MEMB = 9 };

// status codes, mirrored in include/epi_amd.h
enum { OK = 0, ERR_CAPACITY = 1, ERR_STRONG_INTERACTION = 2, ERR_NO_DERMIS = 3,
       ERR_OUT_OF_GRID = 4, ERR_REMOVE_ALIAS = 5, ERR_RNG_EXHAUSTED = 6,
       ERR_INVALID = 7, ERR_HIP = 8, ERR_PAIR_REMOVED = 9 };

// ---- constants -----------------------------------------------------------
// The physics constants are runtime values, one record per tissue (Tis::prm,
// params.h).  What stays here fixes the shape of the build.
constexpr int MEMB_ADJ = 6;                // define.h:147-153 (TRI_MEMB)
constexpr int MAX_CONN = 400;              // define.h:80, cell_connection.cpp:30
constexpr int N3 = 200;                    // cell_connection.cpp:28
// slots between two buckets: the capacity rounded up to a multiple of 4, so
// the sweep can read a bucket's indices four at a time (16-byte aligned)
__host__ __device__ constexpr int bucket_stride(int bcap) { return (bcap + 3) & ~3; }
constexpr double AREA_GRID = 2.0 + 1e-7;   // cell_connection.cpp:16-19 (the bucket edge)

// force laws (cell_interaction.cpp:68-76)
enum : int { LAW_NONE = 0, LAW_OTHER, LAW_DER_DER, LAW_FIXMU, LAW_FIX_MEMB,
             LAW_MU_MEMB, LAW_ALAIRDE, LAW_MEMB_MEMB, LAW_PAIR };

// ---- per-tissue device status (changes on device) ---------------------
struct Status {
  int n;                  // live cells
  int n_old;              // cells before the last state renewal
  int err;                // first error code
  int sw;                 // cornifications since the last Ca solve (uint8 wrap)
  int ndiv;               // divisions in the last renewal
  int nrm;                // removals in the last renewal
  int maxlen;             // longest neighbour list of the last build
  int stamp_z;            // highest z plane ever stamped into the voxel maps
  long long rng_pos;      // replay cursor
  long long next_uid;
  unsigned long long zzmax_bits;  // calc_zzmax result (non-negative double bits)
  double zzmax;
  long long timestep;     // CellManager::current_timestep; k_remove advances it once per renewal
  int nstat;              // records in cst_store
  int pad_st;
  long long grid_b;       // sum of nx*ny*iz_bound over the B sweeps (k_wl_b)
};

// ---- per-tissue descriptor (constant between allocations) ---------------
struct Tis {
  // geometry
  int nmemb, nder, cap, ns;   // ns = neighbour row stride
  int NX, NY, NZ, XP;         // XP = padded x pitch of grid rows
  long long PL;               // plane pitch (XP*(NY+1))
  long long gsize;            // XP*(NY+1)*(NZ+1)
  double LX, LY, LZ, dx, dy, dz, inv_dx, inv_dy, inv_dz;
  int ANX, ANY, ANZ, bcap;
  int irx, iry, irz, pad1;
  // ca2p_du (ca2p.cpp:183), dp (ca2p.cpp:17), eps_m, K_TOTAL, Kspring and
  // KBEND (cell_interaction.cpp:21-28) as the kernels use them: prm's
  // values, or a window's while one is open
  double ca_du, ca_dp;
  double m_eps, m_kadh, m_kspring, m_kbend;
  Status* st;
  // cells
  uint8_t* state; uint8_t* is_malignant; uint8_t* touch_air;
  int* fix_origin; int* rest_div_times; int* pair; int* dermis; int* lat;  // lat: 3/cell
  double* pos[2][3];          // [buffer][axis]
  double *radius, *div_age_thresh;
  double* ca2p[2]; double* ip3[2];
  double *ca2p_avg, *ex_inert, *agek, *ageb, *ex_fat, *in_fat, *spr_nat_len, *diff_u;
  double *ca_th, *ca_kpa, *ca_iag;   // per-solve constants of the Ca step (agek and state are frozen)
  // per list entry during a solve: neighbour << 2 | kind (0 skip, 1 coupled,
  // 3 coupled ALIVE), one word for the index and the kind
  int* ca_ent;
  // the coupled entries alone (kind 1 or 3), in list order: (word, slot of
  // the entry in the row, where its gj lives), and their number per cell
  int2* ca_ce;
  int* ca_cn;
  // (ca2p, IP3) of buffer b beside the two arrays during a solve, so that a
  // list entry reads its neighbour's pair with one 16-byte gather; every
  // kernel that writes ca2p or IP3 inside a solve or window writes it too
  double2* ca_up[2];
  // diff_u of the Ca sub-steps at parity 1 (parity 0 writes diff_u itself)
  double* ca_dfu;
  // per AFM cell and grid_avg8 ATP corner, during a solve: the corner's voxel
  // word and the masks of its six stencil neighbours (bits 0-5; bit 31 set
  // when a mask is not 0/1 and the corner reads the map instead)
  uint2* ca_cq;
  long long* uid; long long* old_uid; int* src_row;
  // cell_stat (define.h:198-202): per cell k_aging_start, k_cornified,
  // k_disap (-1 = never set); cst_store (cellmanager.h:106): the records of
  // removed DEAD cells, up to cap of them, Status::nstat in use
  long long* cst; long long* cst_store;
  // neighbour rows (two buffers: current and previous build)
  int* nlen[2]; int* col[2]; double* gj[2];
  int* nlo;                   // entries of the row below the cell index
  int* btmp; int* bcnt;       // upper partners gathered during the sweep
  int* memb6;                 // triangular adjacency, 6 per membrane cell
  // buckets
  int* bucket_of; int* bslot; int* bucket_n; int* bucket_cells;
  int* bucket_sorted;   // the buckets in ascending cell index, written by k_bucket_sort for the sweep
  // (x, y, z, radius) of every cell after the periodic fix, written by the
  // bucket insertion: the sweep and the row assembly gather a candidate's
  // position and radius with one 32-byte load instead of four
  double4* pk;
  // state renewal
  int* div_slot; double* div_dir;  // per division: parent, direction
  int* div_parent; int* rm_list;
  uint8_t* rflag; int* chunk_any;   // renewal flags and per-1024-cell chunk marks
  // grids
  double* atp[2]; double* bf[2];
  int* cmap1; uint8_t* cmap2; unsigned* vox;
  // replay stream
  const double* rng; long long rng_n;
  // the tissue's runtime parameters (epi_set_params)
  Params prm;
};

__device__ __forceinline__ void set_err(Status* st, int e) { atomicCAS(&st->err, 0, e); }

// ---- geometry helpers: utils.h:25-53 ----------------------------------
__device__ __forceinline__ double p_diff_x(const Tis& t, double x1, double x2) {
  const double diff = x1 - x2;
  if (diff > 0.5 * t.LX) return diff - t.LX;
  if (diff <= -0.5 * t.LX) return diff + t.LX;
  return diff;
}
__device__ __forceinline__ double p_diff_y(const Tis& t, double y1, double y2) {
  const double diff = y1 - y2;
  if (diff > 0.5 * t.LY) return diff - t.LY;
  if (diff <= -0.5 * t.LY) return diff + t.LY;
  return diff;
}
__device__ __forceinline__ double min0(double a) { return a > 0 ? a : 0; }

#define POW6(x) ((x) * (x) * (x) * (x) * (x) * (x))
#define POW3(x) ((x) * (x) * (x))

struct P3 { double x, y, z; };

__device__ __forceinline__ P3 ld_pos(const Tis& t, int pb, int c) {
  return P3{t.pos[pb][0][c], t.pos[pb][1][c], t.pos[pb][2][c]};
}

__device__ __forceinline__ double dist_sq(const Tis& t, P3 a, P3 b) {
  const double diffx = p_diff_x(t, a.x, b.x);
  const double diffy = p_diff_y(t, a.y, b.y);
  const double diffz = a.z - b.z;
  return diffx * diffx + diffy * diffy + diffz * diffz;
}

// grid index of (x, y, z) lattice point in the device layout
__device__ __forceinline__ long long gidx(const Tis& t, int j, int k, int l) {
  return (long long)l * t.PL + (long long)k * t.XP + j;
}

__device__ __forceinline__ bool is_afm(uint8_t s) { return s == ALIVE || s == FIX || s == MUSUME; }

}  // namespace epi
