// k_connect.hip -- connect_cell (cell_connection.cpp:38-203) and calc_zzmax
// (proc.cpp:34-43) on CDNA4.
//
// Neighbour rows are rebuilt every step in the reference's canonical order:
//   [6 membrane neighbours if MEMB] ++ [partners o < X found by X's own
//   5x5x5 bucket sweep, in sweep order] ++ [partners c > X, ascending c].
// k_bucket_insert  places cells in fixed-capacity buckets (grid_init :38-62),
// k_bucket_sort    orders each bucket by cell index (the serial insertion
//                  order),
// k_sweep          runs connect_proc (:70-129) for every non-membrane cell:
//                  the lower part of its own row is written directly, the
//                  upper parts of its partners are appended with atomics,
// k_finish         sorts the upper parts, keeps a sorted copy of the lower
//                  part (the force gather needs it), carries gap-junction
//                  values over from the previous rows by cell uid
//                  (gj_refresh :133-152) and picks the dermis (:158-185).
#include "epi_dev.h"
#include "launch.h"

namespace epi {

__global__ void __launch_bounds__(256) k_bucket_clear(const Tis* __restrict__ T) {
  const Tis& t = T[blockIdx.y];
  const int nb = t.ANX * t.ANY * t.ANZ;
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) t.bucket_n[b] = 0;
}

// grid_init (:38-62)
__global__ void __launch_bounds__(256) k_bucket_insert(const Tis* __restrict__ T, int pb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n = st->n;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const double x = t.pos[pb][0][c], y = t.pos[pb][1][c], z = t.pos[pb][2][c];
    const long aix = (long)((0.5 * t.LX - p_diff_x(t, 0.5 * t.LX, x)) / AREA_GRID);
    const long aiy = (long)((0.5 * t.LY - p_diff_y(t, 0.5 * t.LY, y)) / AREA_GRID);
    const long aiz = (long)((min0(z)) / AREA_GRID);
    t.bcnt[c] = 0;
    t.nlo[c] = 0;
    if (aix >= t.ANX || aiy >= t.ANY || aiz >= t.ANZ || aix < 0 || aiy < 0 || aiz < 0) {
      set_err(st, ERR_OUT_OF_GRID);
      t.bucket_of[c] = -1;
      continue;
    }
    const int b = (int)((aix * t.ANY + aiy) * t.ANZ + aiz);
    const int slot = atomicAdd(&t.bucket_n[b], 1);
    // the reference asserts fewer than N3 = 200 cells per bucket (:59)
    if (slot + 1 >= N3 || slot >= t.bcap) { set_err(st, ERR_CAPACITY); continue; }
    t.bucket_cells[(size_t)b * t.bcap + slot] = c;
    t.bucket_of[c] = b;
  }
}

__global__ void __launch_bounds__(256) k_bucket_sort(const Tis* __restrict__ T) {
  const Tis& t = T[blockIdx.y];
  const int nb = t.ANX * t.ANY * t.ANZ;
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
    const int m = min(t.bucket_n[b], t.bcap);
    if (m < 2) continue;
    int* a = t.bucket_cells + (size_t)b * t.bcap;
    for (int i = 1; i < m; i++) {
      const int v = a[i];
      int j = i - 1;
      while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; j--; }
      a[j + 1] = v;
    }
  }
}

// connect_proc (:70-129): the 5x5x5 sweep, periodic in x, y and z, buckets
// in ascending (j, k, l) with cells in ascending index inside a bucket
__global__ void __launch_bounds__(256) k_sweep(const Tis* __restrict__ T, int pb, int nb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n = st->n;
  const int ANX = t.ANX, ANY = t.ANY, ANZ = t.ANZ;
  int* const col = t.col[nb];
  for (int c = t.nmemb + blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const P3 pc = ld_pos(t, pb, c);
    const double rc = t.radius[c];
    const long anx = (long)(pc.x / AREA_GRID);
    const long any = (long)(pc.y / AREA_GRID);
    const long anz = (long)(pc.z / AREA_GRID);
    if (anx >= ANX || any >= ANY || anz >= ANZ || anx < 0 || any < 0 || anz < 0) {
      set_err(st, ERR_OUT_OF_GRID);
      continue;
    }
    int* const row = col + (size_t)c * t.ns;
    int na = 0;
    for (long j = anx - 2 + ANX; j <= anx + 2 + ANX; j++) {
      const long aix = j % ANX;
      for (long k = any - 2 + ANY; k <= any + 2 + ANY; k++) {
        const long aiy = k % ANY;
        const long bxy = (aix * ANY + aiy) * ANZ;
        for (long l = anz - 2 + ANZ; l <= anz + 2 + ANZ; l++) {
          const long b = bxy + l % ANZ;
          const int sz = min(t.bucket_n[b], t.bcap);
          const int* cells = t.bucket_cells + (size_t)b * t.bcap;
          for (int m = 0; m < sz; ++m) {
            const int o = cells[m];
            if (c <= o) break;   // ascending within the bucket
            const P3 po = ld_pos(t, pb, o);
            const double diffx = p_diff_x(t, pc.x, po.x);
            const double diffy = p_diff_y(t, pc.y, po.y);
            const double diffz = pc.z - po.z;
            const double rad_sum = rc + t.radius[o];
            if (diffx * diffx + diffy * diffy + diffz * diffz <= LJ_THRESH * LJ_THRESH * rad_sum * rad_sum) {
              if (na < t.ns) row[na] = o;
              na++;
              const int slot = atomicAdd(&t.bcnt[o], 1);
              if (slot < t.ns) t.btmp[(size_t)o * t.ns + slot] = c;
            }
          }
        }
      }
    }
    t.nlo[c] = na;
  }
}

// Row assembly, gj carry-over and dermis.
__global__ void __launch_bounds__(256) k_finish(const Tis* __restrict__ T, int pb, int nb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n = st->n;
  const int ob = 1 - nb;
  int mylen = 0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const bool memb = c < t.nmemb;
    const int base = memb ? MEMB_ADJ : 0;
    const int na = memb ? 0 : t.nlo[c];
    const int nbh = t.bcnt[c];
    const int len = base + na + nbh;
    // the reference asserts a list shorter than N2 = 400 (:122)
    if (len >= MAX_CONN || len > t.ns) {
      set_err(st, ERR_CAPACITY);
      t.nlen[nb][c] = 0;
      continue;
    }
    mylen = max(mylen, len);
    int* const row = t.col[nb] + (size_t)c * t.ns;
    if (memb)
      for (int i = 0; i < MEMB_ADJ; i++) row[i] = t.memb6[6 * c + i];
    // upper partners, ascending
    const int* bt = t.btmp + (size_t)c * t.ns;
    int* const up = row + base + na;
    for (int i = 0; i < nbh; i++) {
      const int v = bt[i];
      int j = i - 1;
      while (j >= 0 && up[j] > v) { up[j + 1] = up[j]; j--; }
      up[j + 1] = v;
    }
    // sorted copy of the lower part
    int* const ls = t.lo_sorted + (size_t)c * t.ns;
    for (int i = 0; i < na; i++) {
      const int v = row[i];
      int j = i - 1;
      while (j >= 0 && ls[j] > v) { ls[j + 1] = ls[j]; j--; }
      ls[j + 1] = v;
    }
    t.nlen[nb][c] = len;
    // gj_refresh: a directed edge keeps its value while it stays in the list.
    // Rows change little between builds, so the search starts after the
    // previous match and wraps around once.
    double* const g = t.gj[nb] + (size_t)c * t.ns;
    const int src = t.src_row[c];
    const int* orow = src >= 0 ? t.col[ob] + (size_t)src * t.ns : nullptr;
    const double* og = src >= 0 ? t.gj[ob] + (size_t)src * t.ns : nullptr;
    const int olen = src >= 0 ? t.nlen[ob][src] : 0;
    int hint = 0;
    for (int e = 0; e < len; e++) {
      const long long u = t.uid[row[e]];
      double v = gj_init;
      for (int f = 0; f < olen; f++) {
        int q = hint + f;
        if (q >= olen) q -= olen;
        if (t.old_uid[orow[q]] == u) { v = og[q]; hint = q + 1; break; }
      }
      g[e] = v;
    }
    // set_dermis / find_dermis (:158-185)
    if (!memb && c >= t.nmemb + t.nder) {
      const uint8_t s = t.state[c];
      if (s == FIX || s == MUSUME) {
        double d1Sq = t.LX * t.LX;
        int dermis = -1;
        const P3 pc = ld_pos(t, pb, c);
        for (int e = 0; e < len; e++) {
          const int o = row[e];
          if (t.state[o] == MEMB) {
            const double distanceSq = dist_sq(t, pc, ld_pos(t, pb, o));
            if (distanceSq < d1Sq) { d1Sq = distanceSq; dermis = o; }
          }
        }
        t.dermis[c] = dermis;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) mylen = max(mylen, __shfl_down(mylen, o, 64));
  if ((threadIdx.x & 63) == 0 && mylen > 0) atomicMax(&st->maxlen, mylen);
}

// after a build: every cell's row is its own, uids snapshot the numbering
__global__ void __launch_bounds__(256) k_snapshot(const Tis* __restrict__ T) {
  const Tis& t = T[blockIdx.y];
  const int n = t.st->n;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    t.src_row[c] = c;
    t.old_uid[c] = t.uid[c];
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
  if ((threadIdx.x & 63) == 0 && m > 0)
    atomicMax(&st->zzmax_bits, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_zz_set(const Tis* __restrict__ T, int use_value, double v) {
  const Tis& t = T[blockIdx.x];
  if (threadIdx.x == 0) t.st->zzmax_bits = use_value ? (unsigned long long)__double_as_longlong(v) : 0ull;
}

void launch_connect(hipStream_t s, const Tis* dT, int ntis, int pb, int nb, int max_n, int max_buckets) {
  dim3 gb((max_buckets + 255) / 256, ntis), gc((max_n + 255) / 256, ntis);
  EPI_LAUNCH(K_BUCKET_CLEAR, k_bucket_clear, gb, dim3(256), 0, s, dT);
  EPI_LAUNCH(K_BUCKET_INSERT, k_bucket_insert, gc, dim3(256), 0, s, dT, pb);
  EPI_LAUNCH(K_BUCKET_SORT, k_bucket_sort, gb, dim3(256), 0, s, dT);
  EPI_LAUNCH(K_SWEEP, k_sweep, gc, dim3(256), 0, s, dT, pb, nb);
  EPI_LAUNCH(K_FINISH, k_finish, gc, dim3(256), 0, s, dT, pb, nb);
  EPI_LAUNCH(K_SNAPSHOT, k_snapshot, gc, dim3(256), 0, s, dT);
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
