// k_renew.hip -- cell_state_renew (cell_state_renew.cpp:400-429) on CDNA4.
//
// The reference walks the non-membrane cells in one loop; each cell only
// changes itself and the daughter it appends, but the replayed random draws
// are consumed in loop order and a division consumes extra draws.  So:
//   k_renew_scan   one workgroup per tissue walks the cells in 1024-cell
//                  chunks, compacts the division-ready cells and the removal
//                  queue in index order with wave ballots, and decides the
//                  divisions 64 ready cells at a time (a ballot finds the
//                  first division of a batch; draws after it shift by its
//                  direction draws);
//   k_renew_apply  every cell applies its own rule in parallel; daughters are
//                  written at n_old + division rank (parent order);
//   k_remove       remove_exec's swap-with-tail, one workgroup per tissue;
//   k_pair_disperse pair_disperse from the higher-index partner.
#include "epi_dev.h"
#include "launch.h"

namespace epi {

// cell_state_renew.cpp:17-119: the rates and thresholds are runtime values
// (Tis::prm); these two are fixed by the reference's code
constexpr double stoch_corr_coef = 1;
constexpr double PI_D = 3.14159265358979323846;

// The constants of the renewal laws, copied out of the tissue's record once,
// before a kernel's first store: the compiler cannot tell a cell store from
// the descriptor, so a record field read after a store would be loaded from
// memory again.
struct RenewK {
  double DT_Cell, THRESH_DEAD, THRESH_SP, ca2p_init, stoch_div_time_ratio, S2, accel_div, eps_kb, alpha_b, S0, eps_ks,
      accel_diff, alpha_k, eps_kk, S1, ubar, delta_lipid, delta_sig_r1, lipid_rel, lipid, ADHE_CONST, DISA;
};
__device__ __forceinline__ RenewK renew_k(const Params& pm) {
  return RenewK{pm.DT_Cell, pm.THRESH_DEAD, pm.THRESH_SP, pm.ca2p_init, pm.stoch_div_time_ratio, pm.S2, pm.accel_div,
                pm.eps_kb, pm.alpha_b, pm.S0, pm.eps_ks, pm.accel_diff, pm.alpha_k, pm.eps_kk, pm.S1, pm.ubar,
                pm.delta_lipid, pm.delta_sig_r1, pm.lipid_rel, pm.lipid, pm.ADHE_CONST, pm.DISA_conn_num_thresh};
}

__device__ __forceinline__ double weighted_eps_kb(const RenewK& k, uint8_t malig) {
  return malig ? k.accel_div * k.eps_kb : k.eps_kb;
}

// is_divide_ready without the draw (:147-155), with the MUSUME/FIX guards of
// _MUSUME_state_renew (:282-298) and _FIX_state_renew (:308-318)
__device__ __forceinline__ bool needs_draw(const Tis& t, const RenewK& k, int c) {
  const uint8_t s = t.state[c];
  if (s == FIX) {
    if (t.dermis[c] < 0) return false;
  } else if (s == MUSUME) {
    if (t.dermis[c] < 0 && t.pair[c] < 0) return false;
    if (!(t.rest_div_times[c] > 0)) return false;
  } else {
    return false;
  }
  return t.pair[c] < 0 && (t.ageb[c] >= t.div_age_thresh[c] * (1.0 - k.stoch_div_time_ratio));
}

__device__ __forceinline__ bool to_remove(const Tis& t, const RenewK& k, int cb, int c) {
  const uint8_t s = t.state[c];
  return (s == DEAD || s == AIR) && t.agek[c] >= k.ADHE_CONST && (double)t.nlen[cb][c] <= k.DISA;
}

// Flags of every non-membrane cell (bit 0: needs a draw, bit 1: removal)
// and, per 1024-cell chunk, whether anything needs the ordered walk.
constexpr int RCHUNK = 1024;

__global__ void __launch_bounds__(256) k_renew_flags(const Tis* __restrict__ T, int cb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const RenewK rk = renew_k(t.prm);
  const int start = t.nmemb + t.nder;
  const int n = st->n;
  // blocks are 256 cells and chunks 1024, so a wave never spans two chunks
  for (int c = start + blockIdx.x * blockDim.x + threadIdx.x; c - (int)(threadIdx.x & 63) < n;
       c += gridDim.x * blockDim.x) {
    uint8_t f = 0;
    if (c < n) {
      f = (needs_draw(t, rk, c) ? 1 : 0) | (to_remove(t, rk, cb, c) ? 2 : 0);
      t.rflag[c] = f;
    }
    const unsigned long long any = __ballot(f != 0);
    if (any && (threadIdx.x & 63) == 0) atomicOr(&t.chunk_any[(c - start) / RCHUNK], 1);
  }
}

__global__ void __launch_bounds__(1024) k_renew_scan(const Tis* __restrict__ T, int pb, int cb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  __shared__ int s_ready[RCHUNK];
  __shared__ double s_thr[RCHUNK], s_rhs[RCHUNK];
  __shared__ double s_rng[2 * RCHUNK];
  __shared__ int s_wcnt[16], s_wrm[16];
  __shared__ int s_nrm_base, s_ndiv;
  __shared__ long long s_cursor;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int start = t.nmemb + t.nder;
  const int n_old = st->n;
  const RenewK rk = renew_k(t.prm);
  if (tid == 0) { s_cursor = st->rng_pos; s_ndiv = 0; s_nrm_base = 0; st->n_old = n_old; }
  __syncthreads();
  for (int base = start; base < n_old; base += RCHUNK) {
    const int chunk = (base - start) / RCHUNK;
    if (!t.chunk_any[chunk]) continue;           // uniform for the block
    const int c = base + tid;
    const uint8_t f = c < n_old ? t.rflag[c] : 0;
    const bool rd = f & 1, rm = f & 2;
    const unsigned long long mr = __ballot(rd), mm = __ballot(rm);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (lane == 0) { s_wcnt[wv] = __popcll(mr); s_wrm[wv] = __popcll(mm); }
    // replay window for this chunk's decisions
    const long long cur0 = s_cursor;
    for (int i = tid; i < 2 * RCHUNK; i += RCHUNK) {
      const long long q = cur0 + i;
      s_rng[i] = q < t.rng_n ? t.rng[q] : 2.0;
    }
    __syncthreads();
    int off_r = 0, off_m = 0, tot_r = 0, tot_m = 0;
    for (int w = 0; w < 16; w++) {
      if (w < wv) { off_r += s_wcnt[w]; off_m += s_wrm[w]; }
      tot_r += s_wcnt[w]; tot_m += s_wrm[w];
    }
    if (rd) {
      const int k = off_r + __popcll(mr & below);
      s_ready[k] = c;
      s_thr[k] = t.div_age_thresh[c] * rk.stoch_div_time_ratio;
      s_rhs[k] = stoch_corr_coef * rk.DT_Cell * weighted_eps_kb(rk, t.is_malignant[c]) * rk.S2;
    }
    if (rm) t.rm_list[s_nrm_base + off_m + __popcll(mm & below)] = c;
    __syncthreads();
    if (wv == 0) {
      // decide divisions in index order, 64 ready cells per round
      long long cur = cur0;
      int ndiv = s_ndiv;
      int i = 0;
      while (i < tot_r) {
        const int idx = i + lane;
        const bool act = idx < tot_r;
        const int cell = act ? s_ready[idx] : 0;
        bool dv = false;
        if (act) {
          const long long dpos = cur + lane;
          if (dpos >= t.rng_n) {
            set_err(st, ERR_RNG_EXHAUSTED);
          } else {
            const long long wi = dpos - cur0;
            const double u = wi < 2 * RCHUNK ? s_rng[wi] : t.rng[dpos];
            dv = u * s_thr[idx] <= s_rhs[idx];
          }
        }
        const unsigned long long md = __ballot(dv);
        const int k = min(64, tot_r - i);
        if (md == 0ull) {
          if (act) t.div_slot[cell] = -1;
          cur += k;
          i += k;
        } else {
          const int f0 = __ffsll((long long)md) - 1;
          if (act && lane < f0) t.div_slot[cell] = -1;
          int att = 0;
          if (lane == f0) {
            // div_direction (:192-217) with draws after the decision draw
            const int der = t.dermis[cell];
            const P3 me = ld_pos(t, pb, cell), dm = ld_pos(t, pb, der);
            const double nvx = p_diff_x(t, me.x, dm.x);
            const double nvy = p_diff_y(t, me.y, dm.y);
            const double nvz = me.z - dm.z;
            const double norm = sqrt(nvx * nvx + nvy * nvy + nvz * nvz);
            const double nx = nvx / norm, ny = nvy / norm, nz = nvz / norm;
            double ox, oy, oz, sum;
            long long dp = cur + f0 + 1;
            do {
              if (dp + 1 >= t.rng_n) { set_err(st, ERR_RNG_EXHAUSTED); ox = 1; oy = 0; oz = 0; att++; break; }
              const double rand_theta = PI_D * t.rng[dp];
              const double cr1 = cos(rand_theta);
              const double sr1 = sin(rand_theta);
              const double rand_phi = 2 * PI_D * t.rng[dp + 1];
              const double cr2 = cos(rand_phi);
              const double sr2 = sin(rand_phi);
              const double rvx = sr1 * cr2;
              const double rvy = sr1 * sr2;
              const double rvz = cr1;
              ox = ny * rvz - nz * rvy;
              oy = nz * rvx - nx * rvz;
              oz = nx * rvy - ny * rvx;
              dp += 2;
              att++;
            } while ((sum = ox * ox + oy * oy + oz * oz) < 1.0e-14);
            sum = sqrt(sum);
            t.div_slot[cell] = ndiv;
            t.div_parent[ndiv] = cell;
            t.div_dir[3 * ndiv + 0] = ox / sum;
            t.div_dir[3 * ndiv + 1] = oy / sum;
            t.div_dir[3 * ndiv + 2] = oz / sum;
          }
          att = __shfl(att, f0, 64);
          ndiv++;
          cur += f0 + 1 + 2 * att;
          i += f0 + 1;
        }
      }
      if (lane == 0) { s_cursor = cur; s_ndiv = ndiv; s_nrm_base += tot_m; }
    }
    __syncthreads();
    if (tid == 0) t.chunk_any[chunk] = 0;
  }
  if (tid == 0) {
    st->rng_pos = s_cursor;
    st->ndiv = s_ndiv;
    st->nrm = s_nrm_base;
  }
}

// ageing and lipid laws (:31-119)
__device__ __forceinline__ double ageb_const(const RenewK& k, double ca2p_avg, uint8_t malig) {
  return weighted_eps_kb(k, malig) * (k.S2 + k.alpha_b * min0(ca2p_avg - k.ca2p_init));
}
__device__ __forceinline__ double agek_ALIVE_const(const RenewK& k, double ca2p_avg, uint8_t malig) {
  return (malig ? k.accel_diff * k.eps_ks : k.eps_ks) * (k.S0 + k.alpha_k * min0(ca2p_avg - k.ca2p_init));
}
__device__ __forceinline__ double k_lipid_release(const RenewK& k, double ca2p_avg, double agek) {
  return 0.25 * k.lipid_rel * (1 + tanh((ca2p_avg - k.ubar) / k.delta_sig_r1)) *
         (1 + tanh((agek - k.THRESH_SP) / k.delta_lipid));
}
__device__ __forceinline__ double k_lipid(const RenewK& k, double ca2p_avg, double agek) {
  return 0.25 * k.lipid * (1 + tanh((k.ubar - ca2p_avg) / k.delta_sig_r1)) *
         (1 + tanh((agek - k.THRESH_SP) / k.delta_lipid));
}

// cell_divide (:227-270) for division rank k of parent c
__device__ void divide(const Tis& t, int pb, int cpar, int c, int k, int n_old, long long next_uid) {
  const int d = n_old + k;
  if (d >= t.cap) { set_err(t.st, ERR_CAPACITY); return; }
  const P3 p = ld_pos(t, pb, c);
  const double divx = t.div_dir[3 * k], divy = t.div_dir[3 * k + 1], divz = t.div_dir[3 * k + 2];
  const double cav = t.ca2p[cpar][c];
  const double ipv = t.ip3[cpar][c];
  const int rdt = t.rest_div_times[c];
  const bool mus = t.state[c] == MUSUME;
  // daughter: CellManager::create(MUSUME, ...) (cellmanager.cpp:366-390)
  t.state[d] = MUSUME;
  t.fix_origin[d] = t.fix_origin[c];
  t.radius[d] = t.radius[c];
  t.div_age_thresh[d] = t.prm.agki_max;
  t.ca2p[0][d] = cav; t.ca2p[1][d] = cav;
  t.ca2p_avg[d] = cav;
  t.ip3[0][d] = ipv; t.ip3[1][d] = ipv;
  t.ex_inert[d] = t.ex_inert[c];
  t.agek[d] = 0; t.ageb[d] = 0; t.ex_fat[d] = 0; t.in_fat[d] = 0;
  const double delta_L = t.prm.delta_L;   // :231
  t.spr_nat_len[d] = delta_L;
  t.rest_div_times[d] = mus ? rdt - 1 : rdt;
  t.is_malignant[d] = t.is_malignant[c];
  t.pair[d] = c;
  t.dermis[d] = -1;
  t.diff_u[d] = 0;
  t.uid[d] = next_uid + k;
  t.cst[3 * d] = -1; t.cst[3 * d + 1] = -1; t.cst[3 * d + 2] = -1;
  t.src_row[d] = -1;
  t.touch_air[d] = 0;
  t.lat[3 * d] = 0; t.lat[3 * d + 1] = 0; t.lat[3 * d + 2] = 0;
  t.pos[pb][0][d] = p.x - divx * 0.5 * delta_L;
  t.pos[pb][1][d] = p.y - divy * 0.5 * delta_L;
  t.pos[pb][2][d] = p.z - divz * 0.5 * delta_L;
  // parent
  t.pair[c] = d;
  t.spr_nat_len[c] = delta_L;
  t.ageb[c] = 0;
  if (mus) t.rest_div_times[c] = rdt - 1;
  t.pos[pb][0][c] = p.x + divx * 0.5 * delta_L;
  t.pos[pb][1][c] = p.y + divy * 0.5 * delta_L;
  t.pos[pb][2][c] = p.z + divz * 0.5 * delta_L;
}

__global__ void __launch_bounds__(256) k_renew_apply(const Tis* __restrict__ T, int pb, int cb, int cpar) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int n_old = st->n_old;
  const long long next_uid = st->next_uid;
  const long long ts = st->timestep;
  const RenewK pm = renew_k(t.prm);
  const double DT_Cell = pm.DT_Cell;
  int ncorn = 0;
  for (int c = t.nmemb + t.nder + blockIdx.x * blockDim.x + threadIdx.x; c < n_old;
       c += gridDim.x * blockDim.x) {
    switch (t.state[c]) {
      case ALIVE: {   // _ALIVE_state_renew :347-358
        const double agek = t.agek[c];
        if (agek >= pm.THRESH_DEAD) {
          t.state[c] = DEAD;   // cornificate (cellmanager.cpp:48-53)
          t.cst[3 * c + 1] = ts;
          ncorn++;
        } else {
          const double cav = t.ca2p_avg[c];
          const double in_fat = t.in_fat[c];
          const double tmp = k_lipid_release(pm, cav, agek) * in_fat;
          t.in_fat[c] = in_fat + DT_Cell * (k_lipid(pm, cav, agek) * (1.0 - in_fat) - tmp);
          t.ex_fat[c] = t.ex_fat[c] + DT_Cell * tmp;
          t.agek[c] = agek + DT_Cell * agek_ALIVE_const(pm, cav, t.is_malignant[c]);
        }
        break;
      }
      case DEAD: case AIR:   // _DEAD_AIR_state_renew :330-341
        if (!to_remove(t, pm, cb, c)) t.agek[c] = t.agek[c] + DT_Cell * (pm.eps_kk * pm.S1);
        break;
      case FIX: {   // _FIX_state_renew :308-324
        if (t.dermis[c] < 0) { set_err(st, ERR_NO_DERMIS); break; }
        const bool rd = needs_draw(t, pm, c);
        const int k = rd ? t.div_slot[c] : -1;
        if (k >= 0) divide(t, pb, cpar, c, k, n_old, next_uid);
        else t.ageb[c] = t.ageb[c] + DT_Cell * ageb_const(pm, t.ca2p_avg[c], t.is_malignant[c]);
        break;
      }
      case MUSUME: {   // _MUSUME_state_renew :282-302 (SYSTEM == WHOLE)
        if (t.dermis[c] < 0 && t.pair[c] < 0) {
          t.state[c] = ALIVE;
          t.cst[3 * c] = ts;   // :287
          break;
        }
        const bool rd = needs_draw(t, pm, c);
        const int k = rd ? t.div_slot[c] : -1;
        if (k >= 0) divide(t, pb, cpar, c, k, n_old, next_uid);
        else t.ageb[c] = t.ageb[c] + DT_Cell * ageb_const(pm, t.ca2p_avg[c], t.is_malignant[c]);
        break;
      }
      default:
        break;
    }
  }
  if (ncorn) atomicAdd(&st->sw, ncorn);
}

// CellManager::remove_exec (cellmanager.cpp:338-352): queued slot q_i takes
// the cell at the current tail N-1-i.  A queued slot at or past the current
// end means a queued cell had already been moved (the reference deletes a
// live cell there): ERR_REMOVE_ALIAS.  A removed DEAD cell's cell_stat goes
// to cst_store with k_disap = the current timestep (:343-346), in queue
// order; the timestep then advances (one renewal per step).
__global__ void __launch_bounds__(256) k_remove(const Tis* __restrict__ T, int pb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  const int ndiv = st->ndiv, nrm = st->nrm;
  const int N = st->n_old + ndiv;
  if (threadIdx.x == 0) {
    const long long ts = st->timestep;
    int ns = st->nstat;
    for (int i = 0; i < nrm; i++) {
      const int q = t.rm_list[i];
      if (q >= N - i || t.state[q] != DEAD) continue;
      if (ns >= t.cap) { set_err(st, ERR_CAPACITY); break; }
      long long* o = t.cst_store + 3 * (size_t)ns++;
      o[0] = t.cst[3 * q]; o[1] = t.cst[3 * q + 1]; o[2] = ts;
    }
    st->nstat = ns;
    st->timestep = ts + 1;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nrm; i += blockDim.x) {
    const int q = t.rm_list[i];
    if (q >= N - i) { set_err(st, ERR_REMOVE_ALIAS); continue; }
    if (t.pair[q] >= 0) set_err(st, ERR_PAIR_REMOVED);
    const int s = N - 1 - i;
    if (s == q) continue;
    t.state[q] = t.state[s]; t.is_malignant[q] = t.is_malignant[s]; t.touch_air[q] = t.touch_air[s];
    t.fix_origin[q] = t.fix_origin[s]; t.rest_div_times[q] = t.rest_div_times[s];
    t.pair[q] = t.pair[s]; t.dermis[q] = t.dermis[s];
    for (int a = 0; a < 3; a++) {
      t.lat[3 * q + a] = t.lat[3 * s + a];
      t.pos[pb][a][q] = t.pos[pb][a][s];
      t.cst[3 * q + a] = t.cst[3 * s + a];
    }
    t.radius[q] = t.radius[s]; t.div_age_thresh[q] = t.div_age_thresh[s];
    t.ca2p[0][q] = t.ca2p[0][s]; t.ca2p[1][q] = t.ca2p[1][s];
    t.ip3[0][q] = t.ip3[0][s]; t.ip3[1][q] = t.ip3[1][s];
    t.ca2p_avg[q] = t.ca2p_avg[s]; t.ex_inert[q] = t.ex_inert[s]; t.agek[q] = t.agek[s];
    t.ageb[q] = t.ageb[s]; t.ex_fat[q] = t.ex_fat[s]; t.in_fat[q] = t.in_fat[s];
    t.spr_nat_len[q] = t.spr_nat_len[s]; t.diff_u[q] = t.diff_u[s];
    t.uid[q] = t.uid[s]; t.src_row[q] = t.src_row[s];
  }
  __syncthreads();
  // pair pointers follow the moved cells (Cell* identity in the reference)
  for (int i = threadIdx.x; i < nrm; i += blockDim.x) {
    const int q = t.rm_list[i];
    const int s = N - 1 - i;
    if (q >= N - i || s == q) continue;
    const int p = t.pair[q];
    if (p < 0) continue;
    int np = p;
    if (p >= N - nrm) {
      const int j = N - 1 - p;
      np = t.rm_list[j];
      if (np == p || np >= N - j) { set_err(st, ERR_PAIR_REMOVED); continue; }
    } else {
      t.pair[p] = q;
    }
    t.pair[q] = np;
  }
  if (threadIdx.x == 0) {
    st->n = N - nrm;
    st->next_uid += ndiv;
  }
}

// pair_disperse (:364-385), from the higher-index partner (:422-428)
__global__ void __launch_bounds__(256) k_pair_disperse(const Tis* __restrict__ T, int pb) {
  const Tis& t = T[blockIdx.y];
  const int n = t.st->n;
  for (int c = t.nmemb + t.nder + blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int p = t.pair[c];
    if (p < 0 || !(c > p)) continue;
    const double eps_L = t.prm.eps_L, unpair_dist_coef = t.prm.unpair_dist_coef;
    const double rad_sum = t.radius[c] + t.radius[p];
    const double unpair_th = unpair_dist_coef * rad_sum;
    const double snl = t.spr_nat_len[c];
    if (snl < 2.0 * t.radius[c]) {
      const double v = snl + t.prm.DT_Cell * eps_L;
      t.spr_nat_len[c] = v;
      t.spr_nat_len[p] = v;
    } else if (dist_sq(t, ld_pos(t, pb, c), ld_pos(t, pb, p)) > unpair_th * unpair_th) {
      t.spr_nat_len[c] = 0;
      t.spr_nat_len[p] = 0;
      t.pair[p] = -1;
      t.pair[c] = -1;
    }
  }
}

// initialize_sc (:435-443)
__global__ void __launch_bounds__(256) k_initialize_sc(const Tis* __restrict__ T, const int* tlist, int pb) {
  const Tis& t = T[tlist ? tlist[blockIdx.y] : blockIdx.y];
  const int n = t.st->n;
  const double zzmax = __longlong_as_double((long long)t.st->zzmax_bits);
  for (int c = t.nmemb + t.nder + blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const double THRESH_DEAD = t.prm.THRESH_DEAD;
    if (t.state[c] == ALIVE && zzmax - t.pos[pb][2][c] < 8 * t.prm.R_max) {
      t.agek[c] = t.agek[c] > THRESH_DEAD ? t.agek[c] : THRESH_DEAD;
      t.state[c] = AIR;
    }
  }
}

// cell_pos_periodic_fix (cellmanager.cpp:395-414)
__global__ void __launch_bounds__(256) k_periodic_fix(const Tis* __restrict__ T, int pb) {
  const Tis& t = T[blockIdx.y];
  const int n = t.st->n;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const double x = t.pos[pb][0][c], y = t.pos[pb][1][c];
    if (x > t.LX) t.pos[pb][0][c] = x - t.LX;
    else if (x < 0) t.pos[pb][0][c] = x + t.LX;
    if (y > t.LY) t.pos[pb][1][c] = y - t.LY;
    else if (y < 0) t.pos[pb][1][c] = y + t.LY;
  }
}

void launch_renew(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int cpar, int max_n) {
  dim3 g((max_n + 255) / 256, ntis);
  EPI_LAUNCH(K_RENEW_FLAGS, k_renew_flags, g, dim3(256), 0, s, dT, cb);
  EPI_LAUNCH(K_RENEW_SCAN, k_renew_scan, dim3(1, ntis), dim3(1024), 0, s, dT, pb, cb);
  EPI_LAUNCH(K_RENEW_APPLY, k_renew_apply, g, dim3(256), 0, s, dT, pb, cb, cpar);
  EPI_LAUNCH(K_REMOVE, k_remove, dim3(1, ntis), dim3(256), 0, s, dT, pb);
  EPI_LAUNCH(K_PAIR_DISPERSE, k_pair_disperse, g, dim3(256), 0, s, dT, pb);
}

void launch_initialize_sc(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int pb, int max_n) {
  dim3 g((max_n + 255) / 256, ntis);
  EPI_LAUNCH(K_INIT_SC, k_initialize_sc, g, dim3(256), 0, s, dT, tlist, pb);
}

void launch_periodic_fix(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n) {
  dim3 g((max_n + 255) / 256, ntis);
  EPI_LAUNCH(K_PERIODIC_FIX, k_periodic_fix, g, dim3(256), 0, s, dT, pb);
}

}  // namespace epi
