// k_connect.hip -- connect_cell (cell_connection.cpp:38-203) and calc_zzmax
// (proc.cpp:34-43) on CDNA4.
//
// Neighbour rows are rebuilt every step in the reference's canonical order:
//   [6 membrane neighbours if MEMB] ++ [partners o < X found by X's own
//   5x5x5 bucket sweep, in sweep order] ++ [partners c > X, ascending c].
// k_bucket_insert  places cells in fixed-capacity buckets (grid_init :38-62),
// k_bucket_sort    ranks each cell in its bucket by cell index (the serial
//                  insertion order) and writes it to that slot of a second
//                  copy of the buckets,
// k_sweep          runs connect_proc (:70-129) for every non-membrane cell:
//                  the lower part of its own row is written directly, the
//                  upper parts of its partners are appended with atomics,
// k_finish         sorts the upper parts, carries gap-junction
//                  values over from the previous rows by cell uid
//                  (gj_refresh :133-152) and picks the dermis (:158-185).
#include "epi_dev.h"
#include "launch.h"
#include "epi_group.h"
#include "gj_carry.h"

namespace epi {

// grid_init (:38-62); with fix != 0 it first applies cell_pos_periodic_fix
// (cellmanager.cpp:395-414) to the cell, which touches only that cell
__global__ void __launch_bounds__(256) k_bucket_insert(const Tis* __restrict__ T, int pb, int fix) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n = st->n;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    double x = t.pos[pb][0][c], y = t.pos[pb][1][c];
    const double z = t.pos[pb][2][c];
    if (fix) {
      if (x > t.LX) { x = x - t.LX; t.pos[pb][0][c] = x; }
      else if (x < 0) { x = x + t.LX; t.pos[pb][0][c] = x; }
      if (y > t.LY) { y = y - t.LY; t.pos[pb][1][c] = y; }
      else if (y < 0) { y = y + t.LY; t.pos[pb][1][c] = y; }
    }
    const long aix = (long)((0.5 * t.LX - p_diff_x(t, 0.5 * t.LX, x)) / AREA_GRID);
    const long aiy = (long)((0.5 * t.LY - p_diff_y(t, 0.5 * t.LY, y)) / AREA_GRID);
    const long aiz = (long)((min0(z)) / AREA_GRID);
    t.pk[c] = make_double4(x, y, z, t.radius[c]);
    t.bcnt[c] = 0;
    t.nlo[c] = 0;
    t.bslot[c] = -1;
    if (aix >= t.ANX || aiy >= t.ANY || aiz >= t.ANZ || aix < 0 || aiy < 0 || aiz < 0) {
      set_err(st, ERR_OUT_OF_GRID);
      t.bucket_of[c] = -1;
      continue;
    }
    const int b = (int)((aix * t.ANY + aiy) * t.ANZ + aiz);
    t.bucket_of[c] = b;
    const int slot = atomicAdd(&t.bucket_n[b], 1);
    // the reference asserts fewer than N3 = 200 cells per bucket (:59)
    if (slot + 1 >= N3 || slot >= t.bcap) { set_err(st, ERR_CAPACITY); continue; }
    t.bucket_cells[(size_t)b * bucket_stride(t.bcap) + slot] = c;
    t.bslot[c] = slot;
  }
}

// each bucket in ascending cell index: every cell counts the cells of its
// bucket with a smaller index (its rank, kept in bslot) and writes itself to
// that slot of bucket_sorted, the copy the sweep reads.  A bucket holds
// distinct cells, so the ranks of its m cells are 0 .. m - 1.
__global__ void __launch_bounds__(256) k_bucket_sort(const Tis* __restrict__ T) {
  const Tis& t = T[blockIdx.y];
  const int n = t.st->n;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    if (t.bslot[c] < 0) continue;
    const int b = t.bucket_of[c];
    const int m = min(t.bucket_n[b], t.bcap);
    const size_t base = (size_t)b * bucket_stride(t.bcap);
    const int* a = t.bucket_cells + base;
    int rank = 0;
    // four slots per 16-byte load (a bucket's slots start 16-byte aligned and
    // its stride is a multiple of 4); slots from m on are not counted
    for (int i = 0; i < m; i += 4) {
      const int4 q = *reinterpret_cast<const int4*>(a + i);
      rank += (q.x < c) + (i + 1 < m && q.y < c) + (i + 2 < m && q.z < c) + (i + 3 < m && q.w < c);
    }
    t.bslot[c] = rank;
    t.bucket_sorted[base + rank] = c;
  }
}

// connect_proc (:70-129): the 5x5x5 sweep, periodic in x, y and z, buckets
// in ascending (j, k, l) with cells in ascending index inside a bucket.
// A group of 32 lanes takes one cell; lane g < 25 sweeps the column
// (j, k) = (g / 5, g % 5) over its 5 z buckets.  Lanes count their partners
// (keeping the first SWEEP_H of them in LDS), an exclusive scan over the
// group gives each lane its place in the row, and the lane writes its
// partners there -- the row comes out in sweep order.  A lane with more than
// SWEEP_H partners sweeps its column a second time.
constexpr int SWEEP_G = 32;
constexpr int SWEEP_H = 8;

__device__ __forceinline__ bool sweep_hit(const Tis& t, const P3& pc, double rc, int o) {
  const double4 po = t.pk[o];
  const double diffx = p_diff_x(t, pc.x, po.x);
  const double diffy = p_diff_y(t, pc.y, po.y);
  const double diffz = pc.z - po.z;
  const double rad_sum = rc + po.w;
  const double lj = t.prm.LJ_THRESH;
  return diffx * diffx + diffy * diffy + diffz * diffz <= lj * lj * rad_sum * rad_sum;
}

__global__ void __launch_bounds__(256) k_sweep(const Tis* __restrict__ T, int pb, int nb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n = st->n;
  const int ANX = t.ANX, ANY = t.ANY, ANZ = t.ANZ;
  const int lane = threadIdx.x & (SWEEP_G - 1);
  const int groups = gridDim.x * (blockDim.x / SWEEP_G);
  __shared__ int s_hit[256][SWEEP_H];
  int* const hits = s_hit[threadIdx.x];
  int* const col = t.col[nb];
  for (int c = t.nmemb + blockIdx.x * (blockDim.x / SWEEP_G) + threadIdx.x / SWEEP_G; c < n; c += groups) {
    const double4 pq = t.pk[c];
    const P3 pc{pq.x, pq.y, pq.z};
    const double rc = pq.w;
    const long anx = (long)(pc.x / AREA_GRID);
    const long any = (long)(pc.y / AREA_GRID);
    const long anz = (long)(pc.z / AREA_GRID);
    const bool bad = anx >= ANX || any >= ANY || anz >= ANZ || anx < 0 || any < 0 || anz < 0;
    if (bad) {
      if (lane == 0) set_err(st, ERR_OUT_OF_GRID);
      continue;
    }
    const bool act = lane < 25;
    const long aix = (anx - 2 + lane / 5 + ANX) % ANX;
    const long aiy = (any - 2 + lane % 5 + ANY) % ANY;
    const long bxy = (aix * ANY + aiy) * ANZ;
    int cnt = 0;
    if (act) {
      int sz[5];
#pragma unroll
      for (int i = 0; i < 5; i++) sz[i] = min(t.bucket_n[bxy + (anz - 2 + i + ANZ) % ANZ], t.bcap);
      // a bucket's indices four at a time (one 16-byte load: a bucket's
      // slots start 16-byte aligned and its stride is a multiple of 4), then those
      // candidates' positions together; the walk stops at the first index
      // >= c (ascending within the bucket)
#pragma unroll
      for (int i = 0; i < 5; i++) {
        const int* cells = t.bucket_sorted + (size_t)(bxy + (anz - 2 + i + ANZ) % ANZ) * bucket_stride(t.bcap);
        bool done = false;
        for (int m0 = 0; m0 < sz[i] && !done; m0 += 4) {
          const int4 q = *reinterpret_cast<const int4*>(cells + m0);
          const int o4[4] = {q.x, q.y, q.z, q.w};
          bool hit[4];
#pragma unroll
          for (int u = 0; u < 4; u++) hit[u] = m0 + u < sz[i] && o4[u] < c && sweep_hit(t, pc, rc, o4[u]);
#pragma unroll
          for (int u = 0; u < 4; u++) {
            if (done || m0 + u >= sz[i]) continue;
            if (c <= o4[u]) { done = true; continue; }
            if (hit[u]) {
              if (cnt < SWEEP_H) hits[cnt] = o4[u];
              cnt++;
            }
          }
        }
      }
    }
    int total;
    int off = group_exclusive_scan<SWEEP_G>(cnt, lane, &total);
    if (cnt) {
      int* const row = col + (size_t)c * t.ns;
      if (cnt <= SWEEP_H) {
        for (int i = 0; i < cnt; i++, off++) {
          const int o = hits[i];
          if (off < t.ns) row[off] = o;
          const int slot = atomicAdd(&t.bcnt[o], 1);
          if (slot < t.ns) t.btmp[(size_t)o * t.ns + slot] = c;
        }
      } else {
        for (long l = anz - 2 + ANZ; l <= anz + 2 + ANZ; l++) {
          const long b = bxy + l % ANZ;
          const int sz = min(t.bucket_n[b], t.bcap);
          const int* cells = t.bucket_sorted + (size_t)b * bucket_stride(t.bcap);
          for (int m = 0; m < sz; ++m) {
            const int o = cells[m];
            if (c <= o) break;
            if (sweep_hit(t, pc, rc, o)) {
              if (off < t.ns) row[off] = o;
              off++;
              const int slot = atomicAdd(&t.bcnt[o], 1);
              if (slot < t.ns) t.btmp[(size_t)o * t.ns + slot] = c;
            }
          }
        }
      }
    }
    if (lane == 0) t.nlo[c] = total;
  }
}

// The same sweep with one wave per cell, for small tissues, whose few cells
// leave the GPU mostly idle: lane g takes the buckets 2g and 2g + 1 of the
// 125 in sweep order (bucket q is (j, k, l) = (q / 25, q / 5 % 5, q % 5)).
// A lane issues both bucket counts, then both buckets' first four indices,
// then those eight candidates' positions, so a sweep waits for three memory
// round trips; buckets of more than four cells walk on four at a time.  The
// wave's scan places each lane's partners, so the row is the same.  Twice
// the lanes per cell cost more where the GPU is full (a launch shape chosen
// by the tuner).
constexpr int SWEEP_W = 64;

__global__ void __launch_bounds__(256) k_sweep_wave(const Tis* __restrict__ T, int pb, int nb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n = st->n;
  const int ANX = t.ANX, ANY = t.ANY, ANZ = t.ANZ;
  const int lane = threadIdx.x & (SWEEP_W - 1);
  const int groups = gridDim.x * (blockDim.x / SWEEP_W);
  __shared__ int s_hit[256][SWEEP_H];
  int* const hits = s_hit[threadIdx.x];
  int* const col = t.col[nb];
  const int bstride = bucket_stride(t.bcap);
  for (int c = t.nmemb + blockIdx.x * (blockDim.x / SWEEP_W) + threadIdx.x / SWEEP_W; c < n; c += groups) {
    const double4 pq = t.pk[c];
    const P3 pc{pq.x, pq.y, pq.z};
    const double rc = pq.w;
    const long anx = (long)(pc.x / AREA_GRID);
    const long any = (long)(pc.y / AREA_GRID);
    const long anz = (long)(pc.z / AREA_GRID);
    const bool bad = anx >= ANX || any >= ANY || anz >= ANZ || anx < 0 || any < 0 || anz < 0;
    if (bad) {
      if (lane == 0) set_err(st, ERR_OUT_OF_GRID);
      continue;
    }
    long bk[2];
    int sz[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int q = 2 * lane + u;
      bk[u] = ((((anx - 2 + q / 25) + ANX) % ANX) * ANY + ((any - 2 + q / 5 % 5) + ANY) % ANY) * ANZ +
              ((anz - 2 + q % 5) + ANZ) % ANZ;
      sz[u] = q < 125 ? min(t.bucket_n[bk[u]], t.bcap) : 0;
    }
    int4 q4[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      q4[u] = make_int4(0x7fffffff, 0x7fffffff, 0x7fffffff, 0x7fffffff);
      if (sz[u] > 0) q4[u] = *reinterpret_cast<const int4*>(t.bucket_sorted + (size_t)bk[u] * bstride);
    }
    bool hit[2][4];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int o4[4] = {q4[u].x, q4[u].y, q4[u].z, q4[u].w};
#pragma unroll
      for (int v = 0; v < 4; v++) hit[u][v] = v < sz[u] && o4[v] < c && sweep_hit(t, pc, rc, o4[v]);
    }
    int cnt = 0;
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int o4[4] = {q4[u].x, q4[u].y, q4[u].z, q4[u].w};
#pragma unroll
      for (int v = 0; v < 4; v++)
        if (hit[u][v]) {
          if (cnt < SWEEP_H) hits[cnt] = o4[v];
          cnt++;
        }
      // the rest of a bucket of more than four cells whose first four lie below c
      if (sz[u] > 4 && q4[u].w < c) {
        const int* cells = t.bucket_sorted + (size_t)bk[u] * bstride;
        bool done = false;
        for (int m0 = 4; m0 < sz[u] && !done; m0 += 4) {
          const int4 q = *reinterpret_cast<const int4*>(cells + m0);
          const int o4b[4] = {q.x, q.y, q.z, q.w};
          bool h[4];
#pragma unroll
          for (int v = 0; v < 4; v++) h[v] = m0 + v < sz[u] && o4b[v] < c && sweep_hit(t, pc, rc, o4b[v]);
#pragma unroll
          for (int v = 0; v < 4; v++) {
            if (done || m0 + v >= sz[u]) continue;
            if (c <= o4b[v]) { done = true; continue; }
            if (h[v]) {
              if (cnt < SWEEP_H) hits[cnt] = o4b[v];
              cnt++;
            }
          }
        }
      }
    }
    int total;
    int off = group_exclusive_scan<SWEEP_W>(cnt, lane, &total);
    if (cnt) {
      int* const row = col + (size_t)c * t.ns;
      if (cnt <= SWEEP_H) {
        for (int i = 0; i < cnt; i++, off++) {
          const int o = hits[i];
          if (off < t.ns) row[off] = o;
          const int slot = atomicAdd(&t.bcnt[o], 1);
          if (slot < t.ns) t.btmp[(size_t)o * t.ns + slot] = c;
        }
      } else {
        for (int u = 0; u < 2; u++) {
          const int* cells = t.bucket_sorted + (size_t)bk[u] * bstride;
          for (int m = 0; m < sz[u]; ++m) {
            const int o = cells[m];
            if (c <= o) break;
            if (sweep_hit(t, pc, rc, o)) {
              if (off < t.ns) row[off] = o;
              off++;
              const int slot = atomicAdd(&t.bcnt[o], 1);
              if (slot < t.ns) t.btmp[(size_t)o * t.ns + slot] = c;
            }
          }
        }
      }
    }
    if (lane == 0) t.nlo[c] = total;
  }
}

// Row assembly, gj carry-over and dermis.  A group of KG lanes owns one
// cell.  Rows of at most KCAP entries are assembled in LDS: the lower part
// (written by k_sweep in sweep order) and the upper partners (gathered in
// k_sweep in arbitrary order, placed here by rank since they are distinct).
// The previous row's uids and gj values are staged in LDS too and each lane resolves the
// gap junction of its entries there (gj_refresh :133-152: a directed edge
// keeps its value while it stays in the list, a new edge starts at gj_init).
// The dermis (find_dermis :158-172: the first MEMB entry of minimal distance
// in list order, strictly below LX^2) is a lexicographic (distance, entry)
// minimum over the group.  Longer rows take the same steps through global
// memory.
constexpr int KG = 16;
constexpr int KCAP = 64;   // rows (and previous rows) longer than this take the global-memory path
constexpr int K_BLOCK = 128;
constexpr int K_GROUPS = K_BLOCK / KG;

__global__ void __launch_bounds__(K_BLOCK) k_finish(const Tis* __restrict__ T, int pb, int nb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n = st->n;
  const int ob = 1 - nb;
  __shared__ int s_row[K_GROUPS][KCAP];
  __shared__ int s_bt[K_GROUPS][KCAP];
  __shared__ long long s_ou[K_GROUPS][KCAP];
  __shared__ double s_og[K_GROUPS][KCAP];
  const int lane = threadIdx.x & (KG - 1);
  const int grp = threadIdx.x / KG;
  const double gj_init = t.prm.gj_init;   // read before any store (a store could alias the descriptor)
  int mylen = 0;
  const int stride = gridDim.x * K_GROUPS;
  for (int c = blockIdx.x * K_GROUPS + grp; c < n; c += stride) {
    const bool memb = c < t.nmemb;
    const int base = memb ? MEMB_ADJ : 0;
    const int na = memb ? 0 : t.nlo[c];
    const int nbh = t.bcnt[c];
    const int src = t.src_row[c];
    const int len = base + na + nbh;
    // the reference asserts a list shorter than N2 = 400 (:122)
    if (len >= MAX_CONN || len > t.ns) {
      if (lane == 0) { set_err(st, ERR_CAPACITY); t.nlen[nb][c] = 0; }
      continue;
    }
    mylen = max(mylen, len);
    const int olen = src >= 0 ? t.nlen[ob][src] : 0;
    const int* orow = src >= 0 ? t.col[ob] + (size_t)src * t.ns : nullptr;
    const double* og = src >= 0 ? t.gj[ob] + (size_t)src * t.ns : nullptr;
    int* const row = t.col[nb] + (size_t)c * t.ns;
    const int* bt = t.btmp + (size_t)c * t.ns;
    const bool fits = len <= KCAP;
    const bool staged = olen <= KCAP;
    int* const r = fits ? s_row[grp] : row;
    if (lane == 0) t.nlen[nb][c] = len;
    if (memb && lane < MEMB_ADJ) r[lane] = t.memb6[6 * c + lane];
    if (fits) {
      for (int i = lane; i < na; i += KG) r[i] = row[i];
      for (int i = lane; i < nbh; i += KG) s_bt[grp][i] = bt[i];
    }
    if (staged)
      for (int f = lane; f < olen; f += KG) {
        s_ou[grp][f] = t.old_uid[orow[f]];
        s_og[grp][f] = og[f];
      }
    wave_sync();
    const int* const b = fits ? s_bt[grp] : bt;
    for (int i = lane; i < nbh; i += KG) {
      const int v = b[i];
      int rk = 0;
      for (int k = 0; k < nbh; k++) rk += b[k] < v;
      r[base + na + rk] = v;
    }
    wave_sync();
    const uint8_t s = t.state[c];
    const bool want_dermis = !memb && c >= t.nmemb + t.nder && (s == FIX || s == MUSUME);
    const P3 pc = ld_pos(t, pb, c);
    double bd = t.LX * t.LX;
    int be = 0x7fffffff;
    double* const g = t.gj[nb] + (size_t)c * t.ns;
    for (int e = lane; e < len; e += KG) {
      const int o = r[e];
      if (fits && (e < base || e >= base + na)) row[e] = o;
      const long long u = t.uid[o];
      const long long* const ou = s_ou[grp];
      const int q = gj_carry_slot(e, olen, u, [&](int k) { return staged ? ou[k] : t.old_uid[orow[k]]; });
      g[e] = q < 0 ? gj_init : staged ? s_og[grp][q] : og[q];
      if (want_dermis && t.state[o] == MEMB) {
        const double d = dist_sq(t, pc, ld_pos(t, pb, o));
        if (d < bd) { bd = d; be = e; }   // entries of a lane ascend
      }
    }
    if (want_dermis) {
#pragma unroll
      for (int o = 1; o < KG; o <<= 1) {
        const double d2 = __shfl_xor(bd, o, KG);
        const int e2 = __shfl_xor(be, o, KG);
        if (d2 < bd || (d2 == bd && e2 < be)) { bd = d2; be = e2; }
      }
      if (lane == 0) t.dermis[c] = be == 0x7fffffff ? -1 : r[be];
    }
    wave_sync();
  }
  // maxlen only grows: a wave that cannot raise it skips the atomic, which
  // would otherwise serialise every wave of the launch on one address
  for (int o = 32; o > 0; o >>= 1) mylen = max(mylen, __shfl_down(mylen, o, 64));
  if ((threadIdx.x & 63) == 0 && mylen > __hip_atomic_load(&st->maxlen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(&st->maxlen, mylen);
}

// after a build: every cell's row is its own, uids snapshot the numbering,
// and the buckets are emptied for the next build
// With h_out (pinned host memory), block 0 of each tissue first copies the
// tissue's status there: the status is final once the build's row assembly
// is done, so the step's read-back needs no copy launch of its own.
__global__ void __launch_bounds__(256) k_snapshot(const Tis* __restrict__ T, int pb, Status* h_out) {
  (void)pb;
  const Tis& t = T[blockIdx.y];
  static_assert(sizeof(Status) % 8 == 0 && sizeof(Status) / 8 <= 256, "status copied in 8-byte words");
  if (h_out && blockIdx.x == 0 && threadIdx.x < sizeof(Status) / 8)
    reinterpret_cast<unsigned long long*>(h_out + blockIdx.y)[threadIdx.x] =
        reinterpret_cast<const unsigned long long*>(t.st)[threadIdx.x];
  const int n = t.st->n;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    t.src_row[c] = c;
    t.old_uid[c] = t.uid[c];
    const int b = t.bucket_of[c];
    if (b >= 0) t.bucket_n[b] = 0;
  }
}

// calc_zzmax (proc.cpp:34-43): z values that matter are positive, where the
// IEEE bit order is the numeric order
__global__ void __launch_bounds__(256) k_zzmax(const Tis* __restrict__ T, int pb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n = st->n;
  double m = 0;
  for (int c = t.nmemb + t.nder + blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const uint8_t s = t.state[c];
    if (s == DEAD || s == ALIVE || s == MUSUME || s == FIX) {
      const double z = t.pos[pb][2][c];
      if (m < z) m = z;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_down(m, o, 64);
    if (m < v) m = v;
  }
  // the value was cleared by k_zz_set and only grows during this launch
  const unsigned long long mb = (unsigned long long)__double_as_longlong(m);
  if ((threadIdx.x & 63) == 0 && m > 0 &&
      mb > __hip_atomic_load(&st->zzmax_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(&st->zzmax_bits, mb);
}

__global__ void k_zz_set(const Tis* __restrict__ T, int use_value, double v) {
  const Tis& t = T[blockIdx.x];
  if (threadIdx.x == 0) t.st->zzmax_bits = use_value ? (unsigned long long)__double_as_longlong(v) : 0ull;
}

// CellManager::output's per-cell record (cellmanager.cpp:451-538) packed for
// an asynchronous download: one thread per cell; touch is an ALIVE cell with
// a DEAD cell in its list, as output computes it
__global__ void __launch_bounds__(256) k_out_pack(const Tis* __restrict__ T, int tissue, int pb, int cp, int cb,
                                                  int n, long long m, char* __restrict__ out) {
  const Tis& t = T[tissue];
  double* __restrict__ f = reinterpret_cast<double*>(out + OUT_HDR);
  int* __restrict__ iv = reinterpret_cast<int*>(out + OUT_HDR + 8 * OUT_NF64 * m);
  uint8_t* __restrict__ bv = reinterpret_cast<uint8_t*>(out + OUT_HDR + (8 * OUT_NF64 + 4 * OUT_NI32) * m);
  if (blockIdx.x == 0 && threadIdx.x == 0) *reinterpret_cast<int*>(out) = t.st->n;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const uint8_t s = t.state[c];
    uint8_t touch = 0;
    if (s == ALIVE) {
      const int len = t.nlen[cb][c];
      const int* row = t.col[cb] + (size_t)c * t.ns;
      for (int e = 0; e < len; e++)
        if (t.state[row[e]] == DEAD) {
          touch = 1;
          break;
        }
    }
    const double v[OUT_NF64] = {t.radius[c], t.ageb[c], t.agek[c], t.ca2p[cp][c], t.pos[pb][0][c],
                                t.pos[pb][1][c], t.pos[pb][2][c], t.ca2p_avg[c], t.ex_fat[c], t.in_fat[c],
                                t.spr_nat_len[c]};
#pragma unroll
    for (int k = 0; k < OUT_NF64; k++) f[k * m + c] = v[k];
    iv[c] = t.rest_div_times[c];
    iv[m + c] = t.pair[c];
    iv[2 * m + c] = t.fix_origin[c];
    bv[c] = s;
    bv[m + c] = touch;
  }
}

void launch_out_pack(hipStream_t s, const Tis* dT, int tissue, int pb, int cp, int cb, int n, long long m,
                     char* out) {
  const unsigned g = (unsigned)std::max(1, std::min((n + 255) / 256, 4096));
  EPI_LAUNCH(K_OUT_PACK, k_out_pack, dim3(g), dim3(256), 0, s, dT, tissue, pb, cp, cb, n, m, out);
}

void launch_connect_insert(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n, int fix) {
  dim3 gc((max_n + 255) / 256, ntis);
  EPI_LAUNCH(K_BUCKET_INSERT, k_bucket_insert, gc, dim3(256), 0, s, dT, pb, fix);
}

void launch_connect_rest(hipStream_t s, const Tis* dT, int ntis, int pb, int nb, int max_n, int sweep_wave,
                         Status* h_status) {
  dim3 gc((max_n + 255) / 256, ntis);
  EPI_LAUNCH(K_BUCKET_SORT, k_bucket_sort, gc, dim3(256), 0, s, dT);
  if (sweep_wave) {
    EPI_LAUNCH(K_SWEEP_WAVE, k_sweep_wave, dim3((max_n + 3) / 4, ntis), dim3(256), 0, s, dT, pb, nb);
  } else {
    EPI_LAUNCH(K_SWEEP, k_sweep, dim3((max_n + 7) / 8, ntis), dim3(256), 0, s, dT, pb, nb);
  }
  dim3 gk((max_n + K_GROUPS - 1) / K_GROUPS, ntis);
  EPI_LAUNCH(K_FINISH, k_finish, gk, dim3(K_BLOCK), 0, s, dT, pb, nb);
  EPI_LAUNCH(K_SNAPSHOT, k_snapshot, gc, dim3(256), 0, s, dT, pb, h_status);
}

void launch_connect(hipStream_t s, const Tis* dT, int ntis, int pb, int nb, int max_n, int max_buckets, int fix,
                    int sweep_wave) {
  (void)max_buckets;
  launch_connect_insert(s, dT, ntis, pb, max_n, fix);
  launch_connect_rest(s, dT, ntis, pb, nb, max_n, sweep_wave, nullptr);
}

void launch_zzmax(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n) {
  EPI_LAUNCH(K_ZZ_SET, k_zz_set, dim3(ntis), dim3(64), 0, s, dT, 0, 0.0);
  dim3 gc((max_n + 255) / 256, ntis);
  EPI_LAUNCH(K_ZZMAX, k_zzmax, gc, dim3(256), 0, s, dT, pb);
}

void launch_zz_set(hipStream_t s, const Tis* dT, int ntis, double v) {
  EPI_LAUNCH(K_ZZ_SET, k_zz_set, dim3(ntis), dim3(64), 0, s, dT, 1, v);
}

}  // namespace epi
