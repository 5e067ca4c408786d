// k_field.hip -- setup_map_lat (map.cpp:18-167), calc_ext_stim
// (ext_stim.cpp:18-69) and calc_ca2p (ca2p.cpp:14-402) on CDNA4.
//
// Grids are z-major with x fastest (x pitch padded to 16 doubles), so a
// wavefront reads 64 consecutive x points of one (y, z) row.  The stencils
// map threads to (x, y) columns and march z in chunks of ZB planes (a
// template parameter; the y rows per block are a launch parameter, both set
// by the autotuner); the number of active planes (iz_bound) is derived on the
// device from zzmax and the z chunks are grid-strided, so no host round trip
// is needed.
#include <algorithm>
#include <type_traits>

#include "epi_dev.h"
#include "launch.h"
#include "epi_group.h"
#include "ca_laws.h"
#include "rd_dev.h"

namespace epi {

// ---- voxel maps ---------------------------------------------------------
// Only planes 0..stamp_z can hold stamps (stamp_z never decreases), so only
// they are cleared; the rest of the maps stays empty from the first build.
// Four voxels per thread: a plane holds PL = XP (NY + 1) voxels with XP a
// multiple of 16, and the maps start 256-byte aligned, so every quad is one
// 16-byte and one 4-byte store.
__global__ void __launch_bounds__(256) k_map_clear(const Tis* __restrict__ T) {
  const Tis& t = T[blockIdx.y];
  const int planes = min(t.st->stamp_z, t.NZ) + 1;
  const long long g4 = (long long)planes * t.PL / 4;
  int4* __restrict__ m1 = reinterpret_cast<int4*>(t.cmap1);
  unsigned* __restrict__ m2 = reinterpret_cast<unsigned*>(t.cmap2);
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g4;
       q += (long long)gridDim.x * blockDim.x) {
    m1[q] = make_int4(-1, -1, -1, -1);
    m2[q] = 0u;
  }
}

__device__ __forceinline__ int lat_wrap(long v, int N, Status* st) {
  // precalc_lat_* hold i%N for i in [0, 3N) (utils.cpp:18-30)
  const long i = v + N;
  if (i < 0 || i >= 3L * N) { set_err(st, ERR_OUT_OF_GRID); return (int)(((i % N) + N) % N); }
  return (int)(i % N);
}

// set_lattice (:18-22) and the sphere stamping (:39-155); the last writer of
// the serial loop is the highest cell index, hence atomicMax on cmap1.  One
// lane per cell.
// MS lanes per cell (1, 4 or 16, a launch shape) split each x slice of the
// stamp box, its (y, z) voxels taken z fastest, round-robin: the lanes of a
// cell touch voxels a plane apart.  The result does not depend on MS (a max
// and idempotent stores).
template <int MS>
__global__ void __launch_bounds__(128) k_map_stamp(const Tis* __restrict__ T, int pb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n = st->n;
  const int zl = threadIdx.x & (MS - 1);
  const double FAC_MAP = t.prm.FAC_MAP;
  int ztop = 0;
  for (int c = (blockIdx.x * blockDim.x + threadIdx.x) / MS; c < n; c += gridDim.x * blockDim.x / MS) {
    const P3 p = ld_pos(t, pb, c);
    const int lx = lat_wrap((long)(int)(p.x * t.inv_dx), t.NX, st);
    const int ly = lat_wrap((long)(int)(p.y * t.inv_dy), t.NY, st);
    const int lz = lat_wrap((long)(int)(p.z * t.inv_dz), t.NZ, st);
    if (zl == 0) { t.lat[3 * c] = lx; t.lat[3 * c + 1] = ly; t.lat[3 * c + 2] = lz; }
    const uint8_t s = t.state[c];
    const bool afm = is_afm(s);
    const int irx = afm ? 2 * t.irx : t.irx, iry = afm ? 2 * t.iry : t.iry, irz = afm ? 2 * t.irz : t.irz;
    const double r = t.radius[c];
    const double normal_radSq = r * r;
    const double crad = FAC_MAP * r;
    const double cradSq = crad * crad;
    const int z_b = lz - irz, z_u = lz + irz;
    const int zmin = z_b >= 1 ? z_b : 1;
    const int zmax = z_u < t.NZ ? z_u : t.NZ - 1;
    ztop = max(ztop, zmax);
    const int nzr = zmax - zmin + 1;
    if (nzr <= 0) {   // no plane to stamp: the box's x and y range is still checked
      if (zl == 0) {
        for (int k = lx - irx; k <= lx + irx; k++) lat_wrap(k, t.NX, st);
        for (int l = ly - iry; l <= ly + iry; l++) lat_wrap(l, t.NY, st);
      }
      continue;
    }
    // this lane's first (y, z) voxel of a slice: the zl-th, z fastest
    const int l_first = ly - iry + zl / nzr, m_first = zmin + zl % nzr;
    for (int k = lx - irx; k <= lx + irx; k++) {
      const double mx = k * t.dx;
      const int ipx = lat_wrap(k, t.NX, st);
      const double diffx = p_diff_x(t, mx, p.x);
      const double diffxSq = diffx * diffx;
      int m = m_first;
      for (int l = l_first; l <= ly + iry;) {
        const double my = l * t.dy;
        const int ipy = lat_wrap(l, t.NY, st);
        const double diffy = p_diff_y(t, my, p.y);
        const double tmps = diffxSq + diffy * diffy;
        for (; m <= zmax; m += MS) {
          const double mz = m * t.dz;
          const double diffz = mz - p.z;
          const double distSq = tmps + diffz * diffz;
          const long long v = gidx(t, ipx, ipy, m);
          if (!afm) {
            if (distSq < normal_radSq) atomicMax(&t.cmap1[v], c);
          } else if (distSq < cradSq) {
            t.cmap2[v] = 1;
            if (distSq < normal_radSq) atomicMax(&t.cmap1[v], c);
          }
        }
        // carry the stride into the next rows
        while (m > zmax) { m -= nzr; l++; }
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) ztop = max(ztop, __shfl_down(ztop, o, 64));
  if ((threadIdx.x & 63) == 0 && ztop > __hip_atomic_load(&st->stamp_z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(&st->stamp_z, ztop);
}

// periodic planes x = NX and y = NY for l < NZ (:156-166)
__global__ void __launch_bounds__(256) k_map_halo(const Tis* __restrict__ T) {
  const Tis& t = T[blockIdx.y];
  const int NX = t.NX, NY = t.NY;
  const int per_l = (NX + 1) + (NY + 1);
  const long long tot = (long long)per_l * min(t.st->stamp_z + 1, t.NZ);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
    const int l = (int)(i / per_l), r = (int)(i % per_l);
    long long dst, src;
    if (r <= NX) {   // y = NY row
      const int j = r;
      dst = gidx(t, j, NY, l);
      src = gidx(t, j == NX ? 0 : j, 0, l);
    } else {         // x = NX column
      const int k = r - (NX + 1);
      dst = gidx(t, NX, k, l);
      src = gidx(t, 0, k == NY ? 0 : k, l);
    }
    t.cmap1[dst] = t.cmap1[src];
    t.cmap2[dst] = t.cmap2[src];
  }
}

// ---- ext_stim B: calc_ext_stim (ext_stim.cpp:18-69) ------------------------
struct BK {   // fB's constants, copied out of the record before any store (see CaK)
  double THRESH_DEAD, DUR_ALIVE, DUR_DEAD, kb;
};
__device__ __forceinline__ BK b_k(const Params& pm) { return BK{pm.THRESH_DEAD, pm.DUR_ALIVE, pm.DUR_DEAD, pm.kb}; }
__device__ __forceinline__ double fB(const BK& k, double age, double B, bool cornif) {   // :9-16
  return (cornif && age > k.THRESH_DEAD - k.DUR_ALIVE && age <= k.THRESH_DEAD + k.DUR_DEAD ? 1 : 0) - k.kb * B;
}

// ---- flattened work lists ------------------------------------------------
// A launch over tissues of different sizes numbers the tiles of all of them
// consecutively, tissue by tissue: off[i] is the first tile of the launch's
// i-th tissue and off[nt] the total.  A tile is a chunk of `bs` consecutive
// points of the tissue's (x, y) plane (x fastest, so a wave's lanes read
// consecutive x of one or two rows) times a chunk of ZB z planes, plane
// chunks fastest.  No lane covers a point outside its tissue except in the
// last plane chunk, and no block is launched for a tissue that has no work.

// exclusive prefix sum of v over a block of blockDim.x (a multiple of 64,
// at most 1024) threads; returns the block total in *total
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
  __shared__ int s_w[16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wv] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int i = 0; i < nw; i++) {
    const int w = s_w[i];
    if (i < wv) before += w;
    all += w;
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// The B stencil's list, once zzmax is on the device: per tissue
// iz_bound = (int)((zzmax + FAC_MAP * R_max) * inv_dz) (ext_stim.cpp:22) into
// izb[i], and ceil(NX NY / bs) ceil(iz_bound / zb) tiles.  A tissue whose
// iz_bound lies outside [0, NZ] gets none and reports ERR_OUT_OF_GRID.  The
// sweep's NX NY iz_bound points are added to the tissue's count.  One block.
__global__ void __launch_bounds__(256) k_wl_b(const Tis* __restrict__ T, int ntis, int bs, int zb,
                                              int* __restrict__ off, int* __restrict__ izb_out) {
  int carry = 0;
  for (int base = 0; base < ntis; base += blockDim.x) {
    const int i = base + threadIdx.x;
    int cnt = 0;
    if (i < ntis) {
      const Tis& t = T[i];
      const double zzmax = __longlong_as_double((long long)t.st->zzmax_bits);
      int izb = (int)((zzmax + t.prm.iz_margin) * t.inv_dz);
      if (izb > t.NZ || izb < 0) {
        set_err(t.st, ERR_OUT_OF_GRID);
        izb = 0;
      }
      izb_out[i] = izb;
      t.st->grid_b += (long long)t.NX * t.NY * izb;
      cnt = ((t.NX * t.NY + bs - 1) / bs) * ((izb + zb - 1) / zb);
    }
    int tot;
    const int ex = block_exclusive_scan(cnt, &tot);
    if (i < ntis) off[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) off[ntis] = carry;
}

// Tiles of a work list for this block: a grid-stride loop.  With `xcd` (and
// a grid that is a multiple of 8) the list is cut into 8 contiguous runs,
// one per XCD (workgroups go to the XCDs round-robin by id), so that tiles
// that share halo rows share an L2.
struct TileWalk {
  int w, end, stride;
  // block b of G walking the list (b = blockIdx.x, G = gridDim.x unless the
  // launch shares its grid with other work)
  __device__ TileWalk(int total, int xcd, int b, int G) {
    if (xcd && G >= 8 && (G & 7) == 0) {
      const int seg = (total + 7) / 8, x = b & 7;
      w = x * seg + (b >> 3);
      end = min(total, (x + 1) * seg);
      stride = G >> 3;
    } else {
      w = b;
      end = total;
      stride = G;
    }
  }
};

// the launch's tissue of tile w: the last i in [lo, nt) with off[i] <= w
template <class F>
__device__ __forceinline__ int wl_find(F off, int lo, int nt, int w) {
  int hi = nt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off(mid) <= w) lo = mid; else hi = mid - 1;
  }
  return lo;
}

constexpr int WL_LDS_T = 1024;   // tissues whose list a block stages in LDS

// One (x, y) column of a tile: planes l0 .. l1 - 1, at most ZB of them.
// Every load of the column is issued before any value is used: the column's
// own planes l0 - 1 .. l1 (mirror prev_z(0) = 1, self next_z(NZ - 1) =
// NZ - 1, utils.cpp:32-72), then each plane's four lateral neighbours, masks
// and voxel cell, then the cells' states and ages.  So a tile waits for three
// memory round trips, not three per plane.  The arithmetic is
// ext_stim.cpp:36-57's, in its order.
template <int ZB>
__device__ __forceinline__ void b_column(const Tis& t, int j, int k, int l0, int l1, int bp) {
  const int NX = t.NX, NY = t.NY, NZ = t.NZ;
  const BK bk = b_k(t.prm);
  const double DB = t.prm.DB, DT_Ca = t.prm.DT_Ca;
  const double* __restrict__ carr = t.bf[bp];
  double* __restrict__ narr = t.bf[1 - bp];
  const uint8_t* __restrict__ cm2 = t.cmap2;
  const int prev_x = j == 0 ? NX - 1 : j - 1, next_x = j == NX - 1 ? 0 : j + 1;
  const int prev_y = k == 0 ? NY - 1 : k - 1, next_y = k == NY - 1 ? 0 : k + 1;
  const long long PL = t.PL;
  const long long cxy = gidx(t, j, k, 0);
  const int nl = l1 - l0;
  double cz[ZB + 2];
  uint8_t mz[ZB + 2];
#pragma unroll
  for (int q = 0; q < ZB + 2; q++) {
    int l = q == 0 ? (l0 == 0 ? 1 : l0 - 1) : l0 + q - 1;
    if (l > NZ - 1) l = NZ - 1;
    cz[q] = 0;
    mz[q] = 0;
    if (q <= nl + 1) {
      cz[q] = carr[cxy + l * PL];
      mz[q] = cm2[cxy + l * PL];
    }
  }
  double xm[ZB], xp[ZB], ym[ZB], yp[ZB];
  uint8_t mxm[ZB], mxp[ZB], mym[ZB], myp[ZB];
  int cell[ZB];
#pragma unroll
  for (int q = 0; q < ZB; q++) {
    xm[q] = xp[q] = ym[q] = yp[q] = 0;
    mxm[q] = mxp[q] = mym[q] = myp[q] = 0;
    cell[q] = -1;
    if (q < nl) {
      const int l = l0 + q;
      const long long vxm = gidx(t, prev_x, k, l), vym = gidx(t, j, prev_y, l);
      const long long vyp = gidx(t, j, next_y, l), vxp = gidx(t, next_x, k, l);
      xm[q] = carr[vxm]; ym[q] = carr[vym]; yp[q] = carr[vyp]; xp[q] = carr[vxp];
      mxm[q] = cm2[vxm]; mym[q] = cm2[vym]; myp[q] = cm2[vyp]; mxp[q] = cm2[vxp];
      cell[q] = t.cmap1[cxy + l * PL];
    }
  }
  // a voxel's state and age are read together (the age of a cell that is
  // neither DEAD nor ALIVE is then dropped), one round trip after the map
  uint8_t cs[ZB];
  double age[ZB];
#pragma unroll
  for (int q = 0; q < ZB; q++) {
    cs[q] = MEMB;
    age[q] = 0.0;
    if (cell[q] >= 0) {
      cs[q] = t.state[cell[q]];
      age[q] = t.agek[cell[q]];
    }
  }
#pragma unroll
  for (int q = 0; q < ZB; q++) age[q] = (cs[q] == DEAD || cs[q] == ALIVE) ? age[q] : 0.0;
  const double inv_dz = t.inv_dz;
#pragma unroll
  for (int q = 0; q < ZB; q++) {
    if (q < nl) {
      const int l = l0 + q;
      const bool flg = cs[q] == DEAD || cs[q] == ALIVE;
      const double cext = cz[q + 1];
      const double nv = cext + DT_Ca * (DB *
          ((double)mxm[q] * (xm[q] - cext)
           + (double)mym[q] * (ym[q] - cext)
           + (double)mz[q] * (cz[q] - cext)
           + (double)mz[q + 2] * (cz[q + 2] - cext)
           + (double)myp[q] * (yp[q] - cext)
           + (double)mxp[q] * (xp[q] - cext))
          * inv_dz * inv_dz + fB(bk, age[q], cext, flg));
      narr[cxy + l * PL] = nv;
      // periodic copies (:63-66)
      if (k == 0) narr[gidx(t, j, NY, l)] = nv;
      if (j == 0) narr[gidx(t, NX, k, l)] = nv;
      if (j == 0 && k == 0) narr[gidx(t, NX, NY, l)] = nv;
    }
  }
}

// The same column with the z column rolling through registers, one plane
// after the other: three values and masks live at a time, so a tile of many
// planes keeps the registers (and occupancy) of one.  Large tiles, where the
// sweep is bandwidth- and not latency-bound, take this form.  The voxel's
// cell is read one plane ahead and its state and age together, so a plane
// waits for one memory round trip, not for the chain map -> state -> age.
__device__ __forceinline__ void b_column_roll(const Tis& t, int j, int k, int l0, int l1, int bp) {
  const int NX = t.NX, NY = t.NY, NZ = t.NZ;
  const BK bk = b_k(t.prm);
  const double DB = t.prm.DB, DT_Ca = t.prm.DT_Ca;
  const double* __restrict__ carr = t.bf[bp];
  double* __restrict__ narr = t.bf[1 - bp];
  const uint8_t* __restrict__ cm2 = t.cmap2;
  const int* __restrict__ cm1 = t.cmap1;
  const uint8_t* __restrict__ cst = t.state;
  const double* __restrict__ cage = t.agek;
  const int prev_x = j == 0 ? NX - 1 : j - 1, next_x = j == NX - 1 ? 0 : j + 1;
  const int prev_y = k == 0 ? NY - 1 : k - 1, next_y = k == NY - 1 ? 0 : k + 1;
  const double inv_dz = t.inv_dz;
  const long long PL = t.PL;
  const long long cxy = gidx(t, j, k, 0);
  const int pz = l0 == 0 ? 1 : l0 - 1;   // mirror z (utils.cpp:32-72)
  double czm = carr[cxy + pz * PL];
  int mzm = cm2[cxy + pz * PL];
  double cc = carr[cxy + l0 * PL];
  int mc = cm2[cxy + l0 * PL];
  int cell = l0 < l1 ? cm1[cxy + l0 * PL] : -1;
  for (int l = l0; l < l1; l++) {
    const int next_z = l + 1 == NZ ? NZ - 1 : l + 1;
    const long long v = cxy + l * PL;
    const long long vzp = cxy + next_z * PL;
    const double czp = carr[vzp];
    const int mzp = cm2[vzp];
    const int cell_next = l + 1 < l1 ? cm1[v + PL] : -1;
    uint8_t s = MEMB;
    double a = 0;
    if (cell >= 0) {
      s = cst[cell];
      a = cage[cell];
    }
    const bool flg = s == DEAD || s == ALIVE;
    const double dum_age = flg ? a : 0.0;
    const double cext = cc;
    const long long vxm = gidx(t, prev_x, k, l), vym = gidx(t, j, prev_y, l);
    const long long vyp = gidx(t, j, next_y, l), vxp = gidx(t, next_x, k, l);
    const double nv = cext + DT_Ca * (DB *
        ((double)cm2[vxm] * (carr[vxm] - cext)
         + (double)cm2[vym] * (carr[vym] - cext)
         + (double)mzm * (czm - cext)
         + (double)mzp * (czp - cext)
         + (double)cm2[vyp] * (carr[vyp] - cext)
         + (double)cm2[vxp] * (carr[vxp] - cext))
        * inv_dz * inv_dz + fB(bk, dum_age, cext, flg));
    narr[v] = nv;
    if (k == 0) narr[gidx(t, j, NY, l)] = nv;
    if (j == 0) narr[gidx(t, NX, k, l)] = nv;
    if (j == 0 && k == 0) narr[gidx(t, NX, NY, l)] = nv;
    czm = cc; mzm = mc;
    cc = czp; mc = mzp;
    cell = cell_next;
  }
}

// The B stencil over its tile list.
// One thread per (x, y) point of a tile: tiles of up to 2 planes load the
// column first (b_column), taller ones roll it (b_column_roll).
template <int ZB>
__global__ void __launch_bounds__(256) k_ext_stim(const Tis* __restrict__ T, const int* __restrict__ woff,
                                                  const int* __restrict__ wizb, int ntis, int bp, int xcd) {
  __shared__ int s_off[WL_LDS_T + 1];
  __shared__ int s_izb[WL_LDS_T];
  const bool lds = ntis <= WL_LDS_T;
  if (lds) {
    for (int i = threadIdx.x; i <= ntis; i += blockDim.x) {
      s_off[i] = woff[i];
      if (i < ntis) s_izb[i] = wizb[i];
    }
    __syncthreads();
  }
  auto off = [&](int i) { return lds ? s_off[i] : woff[i]; };
  const int bs = blockDim.x;
  int ti = 0;
  for (TileWalk tw(off(ntis), xcd, blockIdx.x, gridDim.x); tw.w < tw.end; tw.w += tw.stride) {
    // the tile, hence its tissue, is the block's: held in scalar registers,
    // so the tissue's record is read with scalar loads
    ti = __builtin_amdgcn_readfirstlane(wl_find(off, ti, ntis, tw.w));
    const Tis& t = T[ti];
    const int iz_bound = __builtin_amdgcn_readfirstlane(lds ? s_izb[ti] : wizb[ti]);
    const int NX = t.NX, NY = t.NY, NZ = t.NZ;
    const int plane = NX * NY;
    const int npt = (plane + bs - 1) / bs;
    const int local = tw.w - __builtin_amdgcn_readfirstlane(off(ti));
    const int p = (local % npt) * bs + (int)threadIdx.x;
    if (p >= plane) continue;
    const int j = p % NX, k = p / NX;
    const int l0 = (local / npt) * ZB;
    const int l1 = min(l0 + ZB, iz_bound);
    if (ZB >= 4) b_column_roll(t, j, k, l0, l1, bp);
    else b_column<ZB>(t, j, k, l0, l1, bp);
  }
}

// ---- calc_ca2p ------------------------------------------------------------
// init_ca2p_avg (:20-27) and the per-cell "touches AIR" flag used by
// init_ca2p_map (:262-268)
__global__ void __launch_bounds__(256) k_ca_begin(const Tis* __restrict__ T, const int* tlist, int cb, int zero_avg) {
  const Tis& t = T[tlist ? tlist[blockIdx.y] : blockIdx.y];
  const int n = t.st->n;
  const int start = t.nmemb + t.nder;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int* row = t.col[cb] + (size_t)c * t.ns;
    const int len = t.nlen[cb][c];
    uint8_t air = 0;
    for (int e = 0; e < len; e++)
      if (t.state[row[e]] == AIR) { air = 1; break; }
    t.touch_air[c] = air;
    const uint8_t s = t.state[c];
    t.ca_up[0][c] = make_double2(t.ca2p[0][c], t.ip3[0][c]);
    t.ca_up[1][c] = make_double2(t.ca2p[1][c], t.ip3[1][c]);
    if (c >= start && (s == DEAD || is_afm(s))) {
      // which entries take part in the gap-junction sums (supra_calc
      // :201-209; dead_IP3_calc :160-165 counts the ALIVE ones)
      int* ent = t.ca_ent + (size_t)c * t.ns;
      int2* ce = t.ca_ce + (size_t)c * t.ns;
      int k = 0;
      for (int e = 0; e < len; e++) {
        const int o = row[e];
        const uint8_t so = t.state[o];
        const int w = (o << 2) | ((so == ALIVE) ? 3 : (so == DEAD || so == FIX || so == MUSUME) ? 1 : 0);
        ent[e] = w;
        if (w & 3) ce[k++] = make_int2(w, e);
      }
      t.ca_cn[c] = k;
    }
    if (c >= start && is_afm(s)) {
      // calc_th, calc_Kpa, calc_IAG (ca2p.cpp:90-126): agek and state do
      // not change during a solve
      const bool is_alive = s == ALIVE;
      const double agek = t.agek[c];
      // the constants are read before this cell's stores (see CaK)
      const Params& pm = t.prm;
      const double thgra = pm.thgra, thpri = pm.thpri, delta_th = pm.delta_th, TSP = pm.THRESH_SP;
      const double kpa = pm.kpa, Kpri = pm.Kpri, delta_K = pm.delta_K, delta_I = pm.delta_I, ik = pm.iage_kitei;
      if (zero_avg) t.ca2p_avg[c] = 0;
      t.ca_th[c] = is_alive ? thgra + ((thpri - thgra) * 0.5) * (1.0 + tanh((TSP - agek) / delta_th)) : thpri;
      t.ca_kpa[c] = is_alive ? kpa + ((Kpri - kpa) * 0.5) * (1.0 + tanh((TSP - agek) / delta_K)) : Kpri;
      t.ca_iag[c] = is_alive ? 0.5 * (1.0 + tanh((agek - TSP) / delta_I)) : ik;
    }
  }
}

// init_ca2p_map (:238-278): per voxel of the active box the AFM owner
// (+1, 0 = the dummy zero) and the air flag in bit 31
__global__ void __launch_bounds__(256) k_ca_vox(const Tis* __restrict__ T, const int* tlist) {
  const Tis& t = T[tlist ? tlist[blockIdx.y] : blockIdx.y];
  const int izb = ca_iz_bound(t);
  if (izb > t.NZ || izb < 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) set_err(t.st, ERR_OUT_OF_GRID);
    return;
  }
  const long long tot = (long long)t.NX * t.NY * izb;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(i % t.NX);
    const long long r = i / t.NX;
    const int k = (int)(r % t.NY), l = (int)(r / t.NY);
    const long long v = gidx(t, j, k, l);
    const int c = t.cmap1[v];
    unsigned w = 0;
    if (c >= 0) w = (is_afm(t.state[c]) ? (unsigned)(c + 1) : 0u) | ((unsigned)t.touch_air[c] << 31);
    t.vox[v] = w;
  }
}

// Workgroups go to the 8 XCDs round-robin by linear id.  With `on`, the
// linear id is renumbered so that each XCD runs one contiguous run of logical
// ids (a bijection for any grid size): neighbouring cells and grid rows then
// share an L2.  Results do not depend on it.
__device__ __forceinline__ unsigned xcd_order(unsigned L, unsigned G, int on) {
  if (!on) return L;
  const unsigned q = G >> 3, r = G & 7, x = L & 7;
  return x * q + min(x, r) + (L >> 3);
}

// ATP_refresh's update of one point (ca2p.cpp:302-338) from its value, its
// six neighbours (mirror z, periodic x and y) and their masks.  Every path
// that advances an ATP point goes through this one expression.
__device__ __forceinline__ double atp_new(const AtpK& k, double catp, double cxm, double cxp, double cym,
                                          double cyp, double czm, double czp, double mxm, double mxp, double mym,
                                          double myp, double mzm, double mzp, double diffu, double air, double inv_dx) {
  const double fa = k.STIM11 * min0(diffu) - catp * k.Kaa;
  return catp + k.DT_Ca * (k.Da * (mxm * (cxm - catp) + mxp * (cxp - catp) + mym * (cym - catp) + myp * (cyp - catp)
                                   + mzm * (czm - catp) + mzp * (czp - catp))
                           * inv_dx * inv_dx + fa + air * k.AIR_STIM);
}

// diff_u of a sub-step by Ca parity: the sub-step at parity cp writes it and
// its ATP update reads it, so a fused launch (ATP of one sub-step beside the
// cells of the next) reads one buffer while it writes the other
__device__ __forceinline__ double* dfu_buf(const Tis& t, int cp) { return cp ? t.ca_dfu : t.diff_u; }

// The voxel map, the masks and iz_bound do not change during a solve, so
// each AFM cell's eight ATP corners keep their voxel word and neighbour masks
// for all of its sub-steps (used when the corners are computed, not read).
// One thread per (cell, corner), after k_ca_vox.
__global__ void __launch_bounds__(256) k_ca_corners(const Tis* __restrict__ T, const int* tlist) {
  const Tis& t = T[tlist ? tlist[blockIdx.y] : blockIdx.y];
  const int n = t.st->n;
  const int start = t.nmemb + t.nder;
  const int izb = ca_iz_bound(t);
  const int NX = t.NX, NY = t.NY, NZ = t.NZ;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < 8LL * n;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i >> 3), k = (int)(i & 7);
    if (c < start || !is_afm(t.state[c])) continue;
    int j = t.lat[3 * c] + ((k == 1 || k == 4 || k == 5 || k == 7) ? 1 : 0);
    int kk = t.lat[3 * c + 1] + ((k == 2 || k == 4 || k == 6 || k == 7) ? 1 : 0);
    const int l = t.lat[3 * c + 2] + ((k == 3 || k == 5 || k == 6 || k == 7) ? 1 : 0);
    if (j == NX) j = 0;
    if (kk == NY) kk = 0;
    uint2 pre = make_uint2(0u, 0x80000000u);
    if (l < izb && izb <= NZ) {
      const int px = j == 0 ? NX - 1 : j - 1, nx = j == NX - 1 ? 0 : j + 1;
      const int py = kk == 0 ? NY - 1 : kk - 1, ny = kk == NY - 1 ? 0 : kk + 1;
      const int pz = l == 0 ? 1 : l - 1, nz = l + 1 == NZ ? NZ - 1 : l + 1;
      const uint8_t m[6] = {t.cmap2[gidx(t, px, kk, l)], t.cmap2[gidx(t, nx, kk, l)], t.cmap2[gidx(t, j, py, l)],
                            t.cmap2[gidx(t, j, ny, l)], t.cmap2[gidx(t, j, kk, pz)], t.cmap2[gidx(t, j, kk, nz)]};
      unsigned bits = 0;
      for (int b = 0; b < 6; b++) {
        if (m[b] > 1) bits |= 0x80000000u;
        bits |= (unsigned)(m[b] & 1) << b;
      }
      pre = make_uint2(t.vox[gidx(t, j, kk, l)], bits);
    }
    t.ca_cq[i] = pre;
  }
}

// The new value of interior point (j, k, l), l < iz_bound, from buffer carr
// and the sub-step's diff_u, with the point's voxel word and neighbour masks
// from k_ca_corners.  The seven ATP loads do not depend on them; a point whose
// top bit is set (a mask byte above 1) reads the word and masks itself.
__device__ __forceinline__ double atp_point_pre(const Tis& t, const AtpK& ak, const double* __restrict__ carr,
                                                const double* __restrict__ dfu, int j, int k, int l, uint2 pre) {
  const int NX = t.NX, NY = t.NY, NZ = t.NZ;
  const int px = j == 0 ? NX - 1 : j - 1, nx = j == NX - 1 ? 0 : j + 1;
  const int py = k == 0 ? NY - 1 : k - 1, ny = k == NY - 1 ? 0 : k + 1;
  const int pz = l == 0 ? 1 : l - 1, nz = l + 1 == NZ ? NZ - 1 : l + 1;
  const long long v = gidx(t, j, k, l);
  const long long vxm = gidx(t, px, k, l), vxp = gidx(t, nx, k, l);
  const long long vym = gidx(t, j, py, l), vyp = gidx(t, j, ny, l);
  const long long vzm = gidx(t, j, k, pz), vzp = gidx(t, j, k, nz);
  const double c0 = carr[v], cxm = carr[vxm], cxp = carr[vxp], cym = carr[vym], cyp = carr[vyp], czm = carr[vzm],
               czp = carr[vzp];
  unsigned w;
  double m[6];
  if (pre.y & 0x80000000u) {
    const uint8_t* __restrict__ cm2 = t.cmap2;
    w = t.vox[v];
    m[0] = cm2[vxm]; m[1] = cm2[vxp]; m[2] = cm2[vym]; m[3] = cm2[vyp]; m[4] = cm2[vzm]; m[5] = cm2[vzp];
  } else {
    w = pre.x;
#pragma unroll
    for (int b = 0; b < 6; b++) m[b] = (double)((pre.y >> b) & 1);
  }
  const unsigned cell1 = w & 0x7fffffffu;
  const double diffu = cell1 ? dfu[cell1 - 1] : 0.0;
  return atp_new(ak, c0, cxm, cxp, cym, cyp, czm, czp, m[0], m[1], m[2], m[3], m[4], m[5], diffu, (double)(w >> 31),
                 t.inv_dx);
}

// dead_IP3_calc (:156-173) and supra_calc (:182-222).  A group of CG lanes
// (4, 8 or 16, set by the autotuner) owns one cell: lanes evaluate the
// gap-junction terms of the list entries (and update gj toward ALIVE
// neighbours) 16 (CG 16: 32) entries at a time and load the 16 grid corners
// of the two grid_avg8; the leader sums them in list / corner order.  Each
// group takes one cell, so the grid must cover every cell past the membrane
// and dermis (a launch that does not is reported as ERR_INVALID); a loop over
// cells would keep loop-invariant constants of the math in registers for
// the whole kernel and cost a wave per SIMD.  Every lane issues
// all loads of its entries of a chunk (list word, gj, then the neighbours'
// Ca2+ and IP3 pair) before it uses any of them.  With carr_prev set, the ATP
// corners are not read but computed from the previous sub-step's buffer
// (atp_point_pre), the ATP update of that sub-step running beside this one.
// With sv set (a reaction-diffusion window's reverse pass), the cell also
// writes what the reverse sub-step reads of the state before it into segment
// slot `slot`: ca2p and IP3 (AFM and DEAD cells), ex_inert, the ATP
// grid_avg8 and the gj of its coupled entries (AFM cells).
constexpr int C_BLOCK = 128;

template <int CG, int NT>
__device__ __forceinline__ void ca_cells_part(const Tis& t, unsigned bx, unsigned nbx, int cb, int cp, int ap, int bp,
                                              const double* __restrict__ carr_prev,
                                              const double* __restrict__ dfu_prev,
                                              const RdTis* __restrict__ sv = nullptr, int slot = 0) {
  constexpr int C_GROUPS = NT / CG;
  constexpr int EPL = CG >= 8 ? 2 : 4;   // entries per lane per chunk
  constexpr int CCAP = EPL * CG;    // list entries per chunk
  constexpr int NQ = 16 / CG;       // grid corners per lane
  // per-entry terms for the leader; rows padded by 16 bytes so that the
  // leaders of a wave's groups read different LDS banks
  __shared__ double2 s_a[C_GROUPS][CCAP + 1];
  __shared__ double s_g[C_GROUPS][17];
  const int lane = threadIdx.x & (CG - 1);
  const int grp = threadIdx.x / CG;
  const double2* __restrict__ up = t.ca_up[cp];
  double2* __restrict__ up_next = t.ca_up[1 - cp];
  double* __restrict__ dfu_out = dfu_buf(t, cp);
  const int start = t.nmemb + t.nder;
  const int c = start + bx * C_GROUPS + grp;
  if (c >= t.cap) return;
  const CaK ck = ca_k(t.prm);
  const AtpK ak = atp_k(t.prm);
  {
    // The cell's own words are loaded before the cell count and iz_bound
    // (device values) arrive, in one round trip with them: a cell past the
    // count reads allocated, stale slots and leaves once the count is in.
    const uint8_t s = t.state[c];
    // the coupled entries only, in list order: an entry that is skipped
    // would add +0.0 to the sums, which leaves them as they are
    const int len = t.ca_cn[c];
    const double2 uc = up[c];
    const double u = uc.x, p = uc.y;
    // the first chunk's entries do not wait for len: a row holds ns >= CCAP
    // slots, and slots past len are masked once len has arrived
    const int2* ent = t.ca_ce + (size_t)c * t.ns;
    double* g = t.gj[cb] + (size_t)c * t.ns;
    int2 en0[EPL];
#pragma unroll
    for (int i = 0; i < EPL; i++) en0[i] = ent[lane + i * CG];
    // the lattice position is fetched now too, so the corner gathers start
    // one round trip earlier
    const int lx = t.lat[3 * c], ly = t.lat[3 * c + 1], lz = t.lat[3 * c + 2];
    // the corners' per-solve words depend on c only
    uint2 pre[NQ];
#pragma unroll
    for (int j = 0; j < NQ; j++) {
      const int q = lane + j * CG;
      pre[j] = q >= 8 ? t.ca_cq[8 * (size_t)c + (q & 7)] : make_uint2(0u, 0u);
    }
    const int n = t.st->n;
    const int izb = carr_prev ? ca_iz_bound(t) : 0;
    const bool recompute = carr_prev && izb <= t.NZ;
    if (bx == 0 && threadIdx.x == 0 && start + (long long)nbx * C_GROUPS < n) set_err(t.st, ERR_INVALID);
    if (c >= n) return;
    const bool dead = s == DEAD;
    if (!dead && !is_afm(s)) return;
    if (dead) {
      int count = 0;
      double tmp = 0;
      for (int e0 = 0; e0 < len; e0 += CCAP) {
        const int m = min(CCAP, len - e0);
        int en[EPL];
        double po[EPL];
#pragma unroll
        for (int i = 0; i < EPL; i++) {
          const int e = lane + i * CG;
          en[i] = e < m ? (e0 == 0 ? en0[i] : ent[e0 + e]).x : 0;
        }
#pragma unroll
        for (int i = 0; i < EPL; i++) po[i] = (en[i] & 3) == 3 ? up[en[i] >> 2].y : 0.0;
#pragma unroll
        for (int i = 0; i < EPL; i++) {
          const int e = lane + i * CG;
          if (e < m) s_a[grp][e] = make_double2(po[i], (en[i] & 3) == 3 ? 1.0 : 0.0);
        }
        wave_sync();
        // adding the +0.0 of an entry that is not ALIVE leaves tmp as it is
        // (tmp starts at +0.0, so it is never -0.0)
        if (lane == 0) {
          int e = 0;
          for (; e + 4 <= m; e += 4) {
            double2 a[4];
#pragma unroll
            for (int j = 0; j < 4; j++) a[j] = s_a[grp][e + j];
#pragma unroll
            for (int j = 0; j < 4; j++) { tmp += a[j].x; count += a[j].y != 0.0; }
          }
          for (; e < m; e++) { const double2 a = s_a[grp][e]; tmp += a.x; count += a.y != 0.0; }
        }
        wave_sync();
      }
      if (lane == 0) {
        if (sv) {
          sv->su[slot * (size_t)t.cap + c] = u;
          sv->sp[slot * (size_t)t.cap + c] = p;
        }
        const double np = p + ck.DT_Ca * (t.ca_dp * (tmp - count * p) - ck.Kpp * p);
        t.ip3[1 - cp][c] = np;
        reinterpret_cast<double*>(up_next + c)[1] = np;   // ca2p of a DEAD cell stays as it is
      }
      return;
    }
    // the corners of grid_avg8 (utils.h:18-23; slots 0-7 B, 8-15 ATP) are
    // fetched now, beside the list loads
    double cq[NQ];
    {
#pragma unroll
      for (int j = 0; j < NQ; j++) {
        const int q = lane + j * CG;
        const int k = q & 7;
        const int ix = lx + ((k == 1 || k == 4 || k == 5 || k == 7) ? 1 : 0);
        const int iy = ly + ((k == 2 || k == 4 || k == 6 || k == 7) ? 1 : 0);
        const int iz = lz + ((k == 3 || k == 5 || k == 6 || k == 7) ? 1 : 0);
        if (q < 8) {
          cq[j] = t.bf[bp][gidx(t, ix, iy, iz)];
        } else if (recompute && iz < izb) {
          // the periodic copies x = NX, y = NY hold the values of x = 0, y = 0
          cq[j] = atp_point_pre(t, ak, carr_prev, dfu_prev, ix == t.NX ? 0 : ix, iy == t.NY ? 0 : iy, iz, pre[j]);
        } else if (recompute) {
          // plane iz_bound is never updated inside a solve; its copies are
          // written beside this launch, so read the value they copy
          cq[j] = t.atp[ap][gidx(t, ix == t.NX ? 0 : ix, iy == t.NY ? 0 : iy, iz)];
        } else {
          cq[j] = t.atp[ap][gidx(t, ix, iy, iz)];
        }
      }
    }
    double tmp_diffu = 0, tmp_IP3 = 0;
    for (int e0 = 0; e0 < len; e0 += CCAP) {
      const int m = min(CCAP, len - e0);
      int en[EPL], sl[EPL];
      double gv[EPL];
      double2 w[EPL];
#pragma unroll
      for (int i = 0; i < EPL; i++) {
        const int e = lane + i * CG;
        const int2 x = e < m ? (e0 == 0 ? en0[i] : ent[e0 + e]) : make_int2(0, 0);
        en[i] = x.x;
        sl[i] = x.y;
      }
#pragma unroll
      for (int i = 0; i < EPL; i++) {
        gv[i] = (en[i] & 3) ? g[sl[i]] : 0.0;
        w[i] = (en[i] & 3) ? up[en[i] >> 2] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int i = 0; i < EPL; i++) {
        const int e = lane + i * CG;
        if (e < m) {
          const int kd = en[i] & 3;
          double a1 = 0, a2 = 0;
          if (kd) {
            if (sv) sv->sg[(slot * (size_t)t.cap + c) * t.ns + sl[i]] = gv[i];
            a1 = gv[i] * (w[i].x - u);
            a2 = gv[i] * (w[i].y - p);
            if (kd == 3) g[sl[i]] = gv[i] + ck.DT_Ca * fw_term(ck.wd, fabs(w[i].x - u), gv[i]);
          }
          s_a[grp][e] = make_double2(a1, a2);
        }
      }
      wave_sync();
      // the leader adds in list order, four entries' LDS reads ahead of their
      // adds.  A skipped entry adds +0.0, which leaves the sums as they are
      // (they start at +0.0, so neither is ever -0.0).
      if (lane == 0) {
        int e = 0;
        for (; e + 4 <= m; e += 4) {
          double2 a[4];
#pragma unroll
          for (int j = 0; j < 4; j++) a[j] = s_a[grp][e + j];
#pragma unroll
          for (int j = 0; j < 4; j++) { tmp_diffu += a[j].x; tmp_IP3 += a[j].y; }
        }
        for (; e < m; e++) { const double2 a = s_a[grp][e]; tmp_diffu += a.x; tmp_IP3 += a.y; }
      }
      wave_sync();
    }
#pragma unroll
    for (int j = 0; j < NQ; j++) s_g[grp][lane + j * CG] = cq[j];
    wave_sync();
    if (lane == 0) {
      const double exi = t.ex_inert[c], avg = t.ca2p_avg[c], _th = t.ca_th[c], _Kpa = t.ca_kpa[c],
                   IAGv = t.ca_iag[c];
      const double ca2p_du = t.ca_du;
      const double* q = s_g[grp];
      const double Bavg = 0.125 * (q[0] + q[1] + q[2] + q[3] + q[4] + q[5] + q[6] + q[7]);
      const double Aavg = 0.125 * (q[8] + q[9] + q[10] + q[11] + q[12] + q[13] + q[14] + q[15]);
      if (sv) {
        const size_t k = slot * (size_t)t.cap + c;
        sv->su[k] = u;
        sv->sp[k] = p;
        sv->sh[k] = exi;
        sv->sAv[k] = Aavg;
      }
      const double diff_u = ca2p_reaction_factor(ck, u, exi, p, Bavg) + ca2p_du * IAGv * tmp_diffu;
      dfu_out[c] = diff_u;
      const double H0 = ck.H0, DT_Ca = ck.DT_Ca;
      const double np = p + DT_Ca * ((_Kpa * Aavg / (H0 + Aavg) - ck.Kpp * p) + t.ca_dp * IAGv * tmp_IP3);
      t.ip3[1 - cp][c] = np;
      const double cs = u + DT_Ca * diff_u;
      t.ca2p[1 - cp][c] = cs;
      up_next[c] = make_double2(cs, np);
      t.ca2p_avg[c] = avg + cs;
      const double para_k2 = ck.para_k2;
      t.ex_inert[c] = exi + DT_Ca * ((para_k2 * para_k2 / (para_k2 * para_k2 + u * u) - exi) / _th);
    }
    wave_sync();
  }
}

template <int CG>
__global__ void __launch_bounds__(C_BLOCK) k_ca_cells(const Tis* __restrict__ T, CaList cl, int cb, int cp,
                                                      int ap, int bp, int xcd, const RdTis* __restrict__ sv,
                                                      int slot) {
  const unsigned b = xcd_order(blockIdx.x, gridDim.x, xcd);
  const int4 e = cl.ntis == 1 ? make_int4(cl.t0, (int)b, cl.ncb, 0) : cl.cmap[b];
  const Tis& t = T[e.x];
  ca_cells_part<CG, C_BLOCK>(t, e.y, e.z, cb, cp, ap, bp, nullptr, nullptr, sv ? sv + e.x : nullptr, slot);
}

// ATP_refresh (:302-338) with the periodic copies of planes 0..iz_bound,
// one tile of a solve's list: blockDim.x consecutive points of the (x, y)
// plane times ZB planes; the z column rolls through registers.
// One (x, y) column of an ATP tile, planes l0 .. l1 - 1 (at most ZB): as
// b_column, every load first (the column, the lateral neighbours and masks,
// each plane's voxel word), then the owners' diff_u, then ATP_refresh's
// point update (atp_new) per plane.
template <int ZB>
__device__ __forceinline__ void atp_column(const Tis& t, int j, int k, int l0, int l1, int ap, int cp) {
  const int NX = t.NX, NY = t.NY, NZ = t.NZ;
  const AtpK ak = atp_k(t.prm);
  const double* __restrict__ carr = t.atp[ap];
  double* __restrict__ narr = t.atp[1 - ap];
  const double* __restrict__ dfu = dfu_buf(t, cp);
  const uint8_t* __restrict__ cm2 = t.cmap2;
  const int prev_x = j == 0 ? NX - 1 : j - 1, next_x = j == NX - 1 ? 0 : j + 1;
  const int prev_y = k == 0 ? NY - 1 : k - 1, next_y = k == NY - 1 ? 0 : k + 1;
  const long long PL = t.PL;
  const long long cxy = gidx(t, j, k, 0);
  const int nl = l1 - l0;
  double cz[ZB + 2];
  uint8_t mz[ZB + 2];
#pragma unroll
  for (int q = 0; q < ZB + 2; q++) {
    int l = q == 0 ? (l0 == 0 ? 1 : l0 - 1) : l0 + q - 1;   // mirror z boundary (utils.cpp:32-72)
    if (l > NZ - 1) l = NZ - 1;
    cz[q] = 0;
    mz[q] = 0;
    if (q <= nl + 1) {
      cz[q] = carr[cxy + l * PL];
      mz[q] = cm2[cxy + l * PL];
    }
  }
  double xm[ZB], xp[ZB], ym[ZB], yp[ZB];
  uint8_t mxm[ZB], mxp[ZB], mym[ZB], myp[ZB];
  unsigned w[ZB];
#pragma unroll
  for (int q = 0; q < ZB; q++) {
    xm[q] = xp[q] = ym[q] = yp[q] = 0;
    mxm[q] = mxp[q] = mym[q] = myp[q] = 0;
    w[q] = 0;
    if (q < nl) {
      const int l = l0 + q;
      const long long vxm = gidx(t, prev_x, k, l), vxp = gidx(t, next_x, k, l);
      const long long vym = gidx(t, j, prev_y, l), vyp = gidx(t, j, next_y, l);
      xm[q] = carr[vxm]; xp[q] = carr[vxp]; ym[q] = carr[vym]; yp[q] = carr[vyp];
      mxm[q] = cm2[vxm]; mxp[q] = cm2[vxp]; mym[q] = cm2[vym]; myp[q] = cm2[vyp];
      w[q] = t.vox[cxy + l * PL];
    }
  }
  double diffu[ZB];
#pragma unroll
  for (int q = 0; q < ZB; q++) {
    const unsigned cell1 = w[q] & 0x7fffffffu;
    diffu[q] = cell1 ? dfu[cell1 - 1] : 0.0;
  }
  const double inv_dx = t.inv_dx;
#pragma unroll
  for (int q = 0; q < ZB; q++) {
    if (q < nl) {
      const int l = l0 + q;
      const double nv = atp_new(ak, cz[q + 1], xm[q], xp[q], ym[q], yp[q], cz[q], cz[q + 2], (double)mxm[q],
                                (double)mxp[q], (double)mym[q], (double)myp[q], (double)mz[q], (double)mz[q + 2],
                                diffu[q], (double)(w[q] >> 31), inv_dx);
      narr[cxy + l * PL] = nv;
      if (k == 0) narr[gidx(t, j, NY, l)] = nv;
      if (j == 0) narr[gidx(t, NX, k, l)] = nv;
      if (j == 0 && k == 0) narr[gidx(t, NX, NY, l)] = nv;
    }
  }
}

// The same column rolling through registers plane by plane (as
// b_column_roll), for tall tiles; the voxel word is read one plane ahead, so
// a plane waits for one round trip (its owner's diff_u), not two.
__device__ __forceinline__ void atp_column_roll(const Tis& t, int j, int k, int l0, int l1, int ap, int cp) {
  const int NX = t.NX, NY = t.NY, NZ = t.NZ;
  const AtpK ak = atp_k(t.prm);
  const double* __restrict__ carr = t.atp[ap];
  double* __restrict__ narr = t.atp[1 - ap];
  const double* __restrict__ dfu = dfu_buf(t, cp);
  const uint8_t* __restrict__ cm2 = t.cmap2;
  const int prev_x = j == 0 ? NX - 1 : j - 1, next_x = j == NX - 1 ? 0 : j + 1;
  const int prev_y = k == 0 ? NY - 1 : k - 1, next_y = k == NY - 1 ? 0 : k + 1;
  const double inv_dx = t.inv_dx;
  const long long PL = t.PL;
  const long long cxy = gidx(t, j, k, 0);
  const int pz = l0 == 0 ? 1 : l0 - 1;   // mirror z (utils.cpp:32-72)
  double czm = carr[cxy + pz * PL];
  int mzm = cm2[cxy + pz * PL];
  double cc = carr[cxy + l0 * PL];
  int mc = cm2[cxy + l0 * PL];
  // x - 1 and x + 1 lie in the lanes beside this one (the lanes of a wave
  // hold consecutive points of a row) except at a row's ends and the wave's
  // edges, where they are loaded.  (The same exchange in b_column_roll
  // measured 3 % slower; here 1.6 % faster on the 256-tissue batch.)
  const int wl = threadIdx.x & 63;
  const bool lane_xm = j > 0 && wl != 0, lane_xp = j < NX - 1 && wl != 63;
  const unsigned* __restrict__ vox = t.vox;
  unsigned w = l0 < l1 ? vox[cxy + l0 * PL] : 0u;
  for (int l = l0; l < l1; l++) {
    const int next_z = l + 1 == NZ ? NZ - 1 : l + 1;
    const long long v = cxy + l * PL;
    const long long vzp = cxy + next_z * PL;
    const double czp = carr[vzp];
    const int mzp = cm2[vzp];
    const unsigned w_next = l + 1 < l1 ? vox[v + PL] : 0u;
    const unsigned cell1 = w & 0x7fffffffu;
    const double diffu = cell1 ? dfu[cell1 - 1] : 0.0;
    const long long vxm = gidx(t, prev_x, k, l), vxp = gidx(t, next_x, k, l);
    const long long vym = gidx(t, j, prev_y, l), vyp = gidx(t, j, next_y, l);
    const double sxm = __shfl_up(cc, 1, 64), sxp = __shfl_down(cc, 1, 64);
    const int smxm = __shfl_up(mc, 1, 64), smxp = __shfl_down(mc, 1, 64);
    const double axm = lane_xm ? sxm : carr[vxm], axp = lane_xp ? sxp : carr[vxp];
    const int mxm = lane_xm ? smxm : cm2[vxm], mxp = lane_xp ? smxp : cm2[vxp];
    const double nv = atp_new(ak, cc, axm, axp, carr[vym], carr[vyp], czm, czp, (double)mxm, (double)mxp,
                              (double)cm2[vym], (double)cm2[vyp], (double)mzm, (double)mzp, diffu,
                              (double)(w >> 31), inv_dx);
    narr[v] = nv;
    if (k == 0) narr[gidx(t, j, NY, l)] = nv;
    if (j == 0) narr[gidx(t, NX, k, l)] = nv;
    if (j == 0 && k == 0) narr[gidx(t, NX, NY, l)] = nv;
    czm = cc; mzm = mc;
    cc = czp; mc = mzp;
    w = w_next;
  }
}

// ATP_refresh (:302-338) with the periodic copies of planes 0..iz_bound,
// one tile of a solve's list: blockDim.x consecutive points of the (x, y)
// plane times ZB planes (atp_column up to 2 planes, atp_column_roll above).
template <int ZB>
__device__ __forceinline__ void atp_tile(const Tis& t, int local, int izb, int ap, int cp) {
  const int NX = t.NX, NY = t.NY;
  const int bs = blockDim.x;
  const int plane = NX * NY;
  const int npt = (plane + bs - 1) / bs;
  const int p = (local % npt) * bs + (int)threadIdx.x;
  if (p >= plane) return;
  const int j = p % NX, k = p / NX;
  const int l0 = (local / npt) * ZB;
  if (ZB >= 4) atp_column_roll(t, j, k, l0, min(l0 + ZB, izb), ap, cp);
  else atp_column<ZB>(t, j, k, l0, min(l0 + ZB, izb), ap, cp);
  // plane l = iz_bound: copies of values left in the next buffer (:334-337);
  // the corner (NX, NY) ends as a copy of (0, 0)
  if (l0 == 0) {
    double* __restrict__ narr = t.atp[1 - ap];
    if (k == 0) narr[gidx(t, j, NY, izb)] = narr[gidx(t, j, 0, izb)];
    if (j == 0) narr[gidx(t, NX, k, izb)] = narr[gidx(t, 0, k, izb)];
    if (j == 0 && k == 0) narr[gidx(t, NX, NY, izb)] = narr[gidx(t, 0, 0, izb)];
  }
}

// the ATP tiles of a solve's list, walked grid-stride by `nb` blocks of
// which this is block `b` (XCD runs with xcd): the tile offsets, iz_bound
// and tissue ids of up to CA_LDS_T tissues are staged in LDS (kept small: the
// fused kernel's cell blocks hold LDS of their own)
constexpr int CA_LDS_T = 256;
template <int ZB>
__device__ __forceinline__ void atp_walk(const Tis* __restrict__ T, const CaList& cl, int b, int nb, int xcd,
                                         int ap, int cp) {
  const int nt = cl.ntis;
  if (nt == 1) {   // one tissue: its id and iz_bound come with the list
    const Tis& t = T[cl.t0];
    for (TileWalk tw(cl.atiles, xcd, b, nb); tw.w < tw.end; tw.w += tw.stride) atp_tile<ZB>(t, tw.w, cl.izb0, ap, cp);
    return;
  }
  __shared__ int s_off[CA_LDS_T + 1];
  __shared__ int2 s_ti[CA_LDS_T];
  const bool lds = nt <= CA_LDS_T;
  if (lds) {
    for (int i = threadIdx.x; i <= nt; i += blockDim.x) {
      s_off[i] = cl.aoff[i];
      if (i < nt) s_ti[i] = make_int2(cl.atis[i], cl.aizb[i]);
    }
    __syncthreads();
  }
  auto off = [&](int i) { return lds ? s_off[i] : cl.aoff[i]; };
  int ti = 0;
  for (TileWalk tw(off(nt), xcd, b, nb); tw.w < tw.end; tw.w += tw.stride) {
    // uniform over the block (see k_ext_stim)
    ti = __builtin_amdgcn_readfirstlane(wl_find(off, ti, nt, tw.w));
    const int2 q = lds ? s_ti[ti] : make_int2(cl.atis[ti], cl.aizb[ti]);
    atp_tile<ZB>(T[__builtin_amdgcn_readfirstlane(q.x)], tw.w - __builtin_amdgcn_readfirstlane(off(ti)),
                 __builtin_amdgcn_readfirstlane(q.y), ap, cp);
  }
}

template <int ZB>
__global__ void __launch_bounds__(256) k_atp(const Tis* __restrict__ T, CaList cl, int ap, int cp, int xcd) {
  atp_walk<ZB>(T, cl, blockIdx.x, gridDim.x, xcd, ap, cp);
}

// one launch for the ATP update of sub-step i (parities cp, ap) and the cells
// of sub-step i + 1: blocks [0, ncb) take the cells (one entry of the list's
// cell-block map each), the rest walk the ATP tiles.  With xcd (the grid then
// holds 8 ceil(ncb / 8) cell blocks, the last ones idle, and a multiple of 8
// ATP blocks) each XCD (block id mod 8) walks one contiguous run of the ATP
// tiles, so that the tiles above and below a tile, which read its plane,
// share its L2; the cell blocks stay dealt round-robin (contiguous runs of
// cells are runs of one layer, whose cells cost alike, and left the XCDs
// unevenly loaded).  Results do not depend on it.
constexpr int F_BLOCK = 128;

template <int CG, int ZB>
__global__ void __launch_bounds__(F_BLOCK) k_ca_fused(const Tis* __restrict__ T, CaList cl, int cb, int cp,
                                                      int ap, int bp, int ncb, int xcd) {
  int cbk = -1;   // this block's cell block, or -1 for an ATP block
  int ab = 0, an = 0;
  if (xcd) {
    const int cblocks = (ncb + 7) & ~7;   // a multiple of 8: ATP block ids keep their XCD
    if ((int)blockIdx.x < cblocks) {
      cbk = (int)blockIdx.x;
      if (cbk >= ncb) return;
    } else {
      ab = (int)blockIdx.x - cblocks;
      an = (int)gridDim.x - cblocks;
    }
  } else if ((int)blockIdx.x < ncb) {
    cbk = (int)blockIdx.x;
  } else {
    ab = (int)blockIdx.x - ncb;
    an = (int)gridDim.x - ncb;
  }
  if (cbk >= 0) {
    const int4 e = cl.ntis == 1 ? make_int4(cl.t0, cbk, ncb, 0) : cl.cmap[cbk];
    const Tis& t = T[e.x];
    ca_cells_part<CG, F_BLOCK>(t, e.y, e.z, cb, 1 - cp, 1 - ap, bp, t.atp[ap], dfu_buf(t, cp));
  } else {
    atp_walk<ZB>(T, cl, ab, an, xcd, ap, cp);
  }
}

__global__ void __launch_bounds__(256) k_ca_end(const Tis* __restrict__ T, const int* tlist, int full) {
  const Tis& t = T[tlist ? tlist[blockIdx.y] : blockIdx.y];
  const int n = t.st->n;
  const int Ca_ITR = t.prm.Ca_ITR;
  for (int c = t.nmemb + t.nder + blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    if (!full) {
    } else if (is_afm(t.state[c])) {
      t.ca2p_avg[c] = t.ca2p_avg[c] / Ca_ITR;
    } else {
      t.ca2p_avg[c] = 0;
      t.ca2p[0][c] = 0;
      t.ca2p[1][c] = 0;
    }
    t.diff_u[c] = 0;
  }
}

__global__ void __launch_bounds__(256) k_ca_swap(const Tis* __restrict__ T, const int* __restrict__ tlist) {
  const Tis& t = T[tlist[blockIdx.y]];
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long long i = i0; i < t.gsize; i += stride) {
    const double a = t.atp[0][i];
    t.atp[0][i] = t.atp[1][i];
    t.atp[1][i] = a;
  }
  const int n = t.st->n;
  for (long long c = i0; c < n; c += stride) {
    const double u = t.ca2p[0][c], p = t.ip3[0][c];
    const double2 w = t.ca_up[0][c];
    t.ca2p[0][c] = t.ca2p[1][c];
    t.ca2p[1][c] = u;
    t.ip3[0][c] = t.ip3[1][c];
    t.ip3[1][c] = p;
    t.ca_up[0][c] = t.ca_up[1][c];
    t.ca_up[1][c] = w;
  }
}

__global__ void k_sw_reset(const Tis* __restrict__ T, const int* tlist) {
  const Tis& t = T[tlist ? tlist[blockIdx.x] : blockIdx.x];
  if (threadIdx.x == 0) t.st->sw = 0;
}

void launch_ca_swap(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int max_n, long long max_gsize) {
  const long long items = std::max<long long>(max_gsize, max_n);
  const long long g = std::min<long long>((items + 255) / 256, 4096);
  EPI_LAUNCH(K_CA_SWAP, k_ca_swap, dim3((unsigned)g, ntis), dim3(256), 0, s, dT, tlist);
}

void launch_maps(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n, long long max_gsize, int max_halo,
                 const Tuning& tu) {
  const long long q = (max_gsize / 4 + 255) / 256;
  dim3 gg((unsigned)(q < 1 ? 1 : q < 65535 ? q : 65535), ntis);
  EPI_LAUNCH(K_MAP_CLEAR, k_map_clear, gg, dim3(256), 0, s, dT);
  if (tu.map_lanes == 16)
    EPI_LAUNCH(K_MAP_STAMP, k_map_stamp<16>, dim3((16 * max_n + 127) / 128, ntis), dim3(128), 0, s, dT, pb);
  else if (tu.map_lanes == 4)
    EPI_LAUNCH(K_MAP_STAMP, k_map_stamp<4>, dim3((4 * max_n + 127) / 128, ntis), dim3(128), 0, s, dT, pb);
  else
    EPI_LAUNCH(K_MAP_STAMP, k_map_stamp<1>, dim3((max_n + 127) / 128, ntis), dim3(128), 0, s, dT, pb);
  EPI_LAUNCH(K_MAP_HALO, k_map_halo, dim3((max_halo + 255) / 256, ntis), dim3(256), 0, s, dT);
}

namespace {
template <class F>
void by_zb(int zb, F f) {
  switch (zb) {
    case 1: f(std::integral_constant<int, 1>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    case 8: f(std::integral_constant<int, 8>()); break;
    case 16: f(std::integral_constant<int, 16>()); break;
    default: f(std::integral_constant<int, 4>()); break;
  }
}
}  // namespace

void launch_ext_stim(hipStream_t s, const Tis* dT, int ntis, int bp, long long tiles_est, int* d_wl,
                     const Tuning& tu) {
  const int bs = 64 * tu.b_rows;
  // a grid-stride walk over the list: the grid is sized from the host's
  // estimate (the last known zzmax), one block per tile (a cap at the blocks
  // the CUs hold at once left a tail of blocks waiting: 376-383 against
  // 328-338 us on the 256-tissue batch), and a multiple of 8 so that the XCD
  // split applies
  const long long cap = 1LL << 24;
  long long g = std::min(std::max(tiles_est, 1LL), cap);
  g = (g + 7) / 8 * 8;
  by_zb(tu.b_zb, [&](auto zbc) {
    constexpr int zb = decltype(zbc)::value;
    EPI_LAUNCH(K_WL, k_wl_b, dim3(1), dim3(256), 0, s, dT, ntis, bs, zb, d_wl, d_wl + ntis + 1);
    EPI_LAUNCH(K_EXT_STIM, k_ext_stim<zb>, dim3((unsigned)g), dim3(bs), 0, s, dT, d_wl, d_wl + ntis + 1, ntis, bp,
               tu.b_xcd);
  });
}

static void ca_maps(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int cb, int max_n, long long max_box,
                    int zero_avg) {
  EPI_LAUNCH(K_CA_BEGIN, k_ca_begin, dim3((max_n + 255) / 256, ntis), dim3(256), 0, s, dT, tlist, cb, zero_avg);
  unsigned gb = (unsigned)((max_box + 255) / 256);
  if (gb < 1) gb = 1;
  if (gb > 65535) gb = 65535;
  EPI_LAUNCH(K_CA_VOX, k_ca_vox, dim3(gb, ntis), dim3(256), 0, s, dT, tlist);
  EPI_LAUNCH(K_CA_CORNERS, k_ca_corners, dim3((8 * max_n + 255) / 256, ntis), dim3(256), 0, s, dT, tlist);
}

void launch_ca_begin(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int cb, int max_n, long long max_box) {
  ca_maps(s, dT, tlist, ntis, cb, max_n, max_box, 1);
}

void launch_ca_touch(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int cb, int max_n, long long max_box) {
  ca_maps(s, dT, tlist, ntis, cb, max_n, max_box, 0);
}

void launch_ca_diffu_zero(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int max_n) {
  EPI_LAUNCH(K_CA_END, k_ca_end, dim3((max_n + 255) / 256, ntis), dim3(256), 0, s, dT, tlist, 0);
}

namespace {
void launch_cells(hipStream_t s, const Tis* dT, const CaList& cl, int cb, int cp, int ap, int bp, const Tuning& tu,
                  const RdTis* sv = nullptr, int slot = 0) {
  const int cg = tu.ca_group;
  const dim3 gc((unsigned)std::max(cl.ncb, 1));
  if (cg == 4)
    EPI_LAUNCH(K_CA_CELLS, k_ca_cells<4>, gc, dim3(C_BLOCK), 0, s, dT, cl, cb, cp, ap, bp, tu.xcd, sv, slot);
  else if (cg == 8)
    EPI_LAUNCH(K_CA_CELLS, k_ca_cells<8>, gc, dim3(C_BLOCK), 0, s, dT, cl, cb, cp, ap, bp, tu.xcd, sv, slot);
  else
    EPI_LAUNCH(K_CA_CELLS, k_ca_cells<16>, gc, dim3(C_BLOCK), 0, s, dT, cl, cb, cp, ap, bp, tu.xcd, sv, slot);
}

// blocks that walk a list of `tiles` tiles: one per tile (as launch_ext_stim),
// a multiple of 8 (the XCD split)
unsigned walk_blocks(long long tiles, int bs) {
  (void)bs;
  const long long cap = 1LL << 24;
  long long g = std::min(std::max(tiles, 1LL), cap);
  return (unsigned)((g + 7) / 8 * 8);
}

void launch_atp(hipStream_t s, const Tis* dT, const CaList& cl, int cp, int ap, const Tuning& tu) {
  by_zb(tu.a_zb, [&](auto zbc) {
    constexpr int zb = decltype(zbc)::value;
    EPI_LAUNCH(K_ATP, k_atp<zb>, dim3(walk_blocks(cl.atiles, cl.abs)), dim3(cl.abs), 0, s, dT, cl, ap, cp, tu.xcd);
  });
}

template <int CG>
void launch_fused_cg(hipStream_t s, const Tis* dT, const CaList& cl, int cb, int cp, int ap, int bp,
                     const Tuning& tu) {
  by_zb(tu.a_zb, [&](auto zbc) {
    constexpr int zb = decltype(zbc)::value;
    const unsigned cblocks = tu.xcd ? (unsigned)(cl.ncb + 7) / 8 * 8 : (unsigned)cl.ncb;
    const dim3 g(cblocks + walk_blocks(cl.atiles, F_BLOCK));
    const auto kfn = k_ca_fused<CG, zb>;
    EPI_LAUNCH(K_CA_FUSED, kfn, g, dim3(F_BLOCK), 0, s, dT, cl, cb, cp, ap, bp, cl.ncb, tu.xcd);
  });
}
}  // namespace

int ca_list_bs(const Tuning& tu, bool fused) { return fused ? F_BLOCK : 64 * tu.a_rows; }

void launch_ca_substep(hipStream_t s, const Tis* dT, const CaList& cl, int cb, int cp, int ap, int bp,
                       const Tuning& tu, const RdTis* sv, int slot, bool cells_only) {
  launch_cells(s, dT, cl, cb, cp, ap, bp, tu, sv, slot);
  if (!cells_only) launch_atp(s, dT, cl, cp, ap, tu);
}

void launch_ca_substeps(hipStream_t s, const Tis* dT, const CaList& cl, int cb, int& cp, int& ap, int bp,
                        int n_sub, const Tuning& tu) {
  if (!tu.ca_fuse || n_sub < 2) {
    for (int i = 0; i < n_sub; i++) {
      launch_ca_substep(s, dT, cl, cb, cp, ap, bp, tu);
      cp ^= 1;
      ap ^= 1;
    }
    return;
  }
  // cells of sub-step 0; then (ATP of i, cells of i + 1); then ATP of the last
  launch_cells(s, dT, cl, cb, cp, ap, bp, tu);
  for (int i = 0; i + 1 < n_sub; i++) {
    if (tu.ca_group == 4)
      launch_fused_cg<4>(s, dT, cl, cb, cp, ap, bp, tu);
    else if (tu.ca_group == 8)
      launch_fused_cg<8>(s, dT, cl, cb, cp, ap, bp, tu);
    else
      launch_fused_cg<16>(s, dT, cl, cb, cp, ap, bp, tu);
    cp ^= 1;
    ap ^= 1;
  }
  launch_atp(s, dT, cl, cp, ap, tu);
  cp ^= 1;
  ap ^= 1;
}

void launch_ca_end(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int max_n) {
  EPI_LAUNCH(K_CA_END, k_ca_end, dim3((max_n + 255) / 256, ntis), dim3(256), 0, s, dT, tlist, 1);
  EPI_LAUNCH(K_SW_RESET, k_sw_reset, dim3(ntis), dim3(64), 0, s, dT, tlist);
}

}  // namespace epi
