// k_adjoint.hip -- the frozen reaction-diffusion window (the calc_ca2p
// sub-step, ca2p.cpp:392-399, with positions, lists, states and voxel maps
// frozen) and its reverse pass.
//
// Forward: the window restores a snapshot of the tissue and runs the plain
// sub-step kernels (k_ca_cells, k_atp) with the runtime coefficients in Tis.
// Reverse: for the loss L = mean over AFM cells of (ca2p_T - target)^2 each
// sub-step is differentiated with respect to its state (ca2p, IP3, ex_inert,
// per-edge gj, ATP) and the parameters.  Values a sub-step never writes
// (ca2p of DEAD cells, ATP planes from iz_bound up) are constants.  The
// reverse pass scatters with FP64 atomics, so gradients are reproducible to
// rounding, not bit for bit.
#include "epi_dev.h"
#include "launch.h"
#include "ca_laws.h"
#include "rd_dev.h"

namespace epi {

namespace {
constexpr int RL = 16;   // lanes per cell for row-wise copies

// every (cell, list entry) of the rows of buffer cb, RL lanes per cell
#define RD_FOR_ROW_ENTRIES(t, cb, n, c, i)                                                      \
  for (int c = (int)((blockIdx.x * (long long)blockDim.x + threadIdx.x) / RL); c < (n);          \
       c += (int)((long long)gridDim.x * blockDim.x / RL))                                       \
    for (long long i = (long long)c * (t).ns + (threadIdx.x & (RL - 1)),                         \
                   i##_end = (long long)c * (t).ns + (t).nlen[cb][c];                            \
         i < i##_end; i += RL)

__device__ __forceinline__ long long box_end(const Tis& t, const RdTis& r) { return (long long)r.izb * t.PL; }

// grid point that a grid_avg8 corner reads: copy planes x = NX, y = NY hold
// the values of x = 0, y = 0; planes from iz_bound up are constants (-1)
__device__ __forceinline__ long long corner_var(const Tis& t, const RdTis& r, int j, int k, int l) {
  if (l >= r.izb) return -1;
  if (j == t.NX) j = 0;
  if (k == t.NY) k = 0;
  return gidx(t, j, k, l);
}
}  // namespace

// restore the snapshot with the window's starting conductance (and kpa)
__global__ void __launch_bounds__(256) k_rd_restore(const Tis* __restrict__ T, RdTis* __restrict__ R,
                                                     const double* __restrict__ gj0, int cb, int cp, int ap,
                                                     int use_kpa) {
  const Tis& t = T[blockIdx.y];
  const RdTis& r = R[blockIdx.y];
  const int n = t.st->n;
  const double g0 = gj0[blockIdx.y];
  const long long stride = (long long)gridDim.x * blockDim.x;
  RD_FOR_ROW_ENTRIES(t, cb, n, c, i) t.gj[cb][i] = g0;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
    t.ca2p[cp][c] = r.u0[0][c]; t.ca2p[1 - cp][c] = r.u0[1][c];
    t.ip3[cp][c] = r.p0[0][c]; t.ip3[1 - cp][c] = r.p0[1][c];
    t.ca_up[cp][c] = make_double2(r.u0[0][c], r.p0[0][c]);
    t.ca_up[1 - cp][c] = make_double2(r.u0[1][c], r.p0[1][c]);
    t.ex_inert[c] = r.h0[c]; t.ca2p_avg[c] = r.avg0[c]; t.diff_u[c] = 0;
    t.ca_kpa[c] = use_kpa ? r.kpa_in[c] : r.kpa0[c];
  }
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < t.gsize; v += stride) {
    t.atp[ap][v] = r.A0[0][v]; t.atp[1 - ap][v] = r.A0[1][v];
  }
}

// take the snapshot (the reverse of k_rd_restore; gj rows are kept as they are)
__global__ void __launch_bounds__(256) k_rd_snapshot(const Tis* __restrict__ T, RdTis* __restrict__ R, int cb,
                                                      int cp, int ap) {
  const Tis& t = T[blockIdx.y];
  RdTis& r = R[blockIdx.y];
  const int n = t.st->n;
  const int start = t.nmemb + t.nder;
  const long long stride = (long long)gridDim.x * blockDim.x;
  int nafm = 0;
  RD_FOR_ROW_ENTRIES(t, cb, n, c, i) r.g0[i] = t.gj[cb][i];
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
    r.u0[0][c] = t.ca2p[cp][c]; r.u0[1][c] = t.ca2p[1 - cp][c];
    r.p0[0][c] = t.ip3[cp][c]; r.p0[1][c] = t.ip3[1 - cp][c];
    r.h0[c] = t.ex_inert[c]; r.avg0[c] = t.ca2p_avg[c]; r.kpa0[c] = t.ca_kpa[c];
    nafm += (c >= start && is_afm(t.state[c])) ? 1 : 0;
  }
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < t.gsize; v += stride) {
    r.A0[0][v] = t.atp[ap][v]; r.A0[1][v] = t.atp[1 - ap][v];
  }
  for (int o = 32; o > 0; o >>= 1) nafm += __shfl_down(nafm, o, 64);
  if ((threadIdx.x & 63) == 0 && nafm) atomicAdd(&r.nafm, nafm);
}

// put back the gj rows of the snapshot (end of a window)
__global__ void __launch_bounds__(256) k_rd_put_gj(const Tis* __restrict__ T, RdTis* __restrict__ R, int cb) {
  const Tis& t = T[blockIdx.y];
  const RdTis& r = R[blockIdx.y];
  const int n = t.st->n;
  RD_FOR_ROW_ENTRIES(t, cb, n, c, i) t.gj[cb][i] = r.g0[i];
}

// copy the state between the tissue and a segment slot (dir 0: save,
// 1: restore; of the ATP grid a slot keeps only the cells' grid_avg8, so a
// restore from it leaves the grid as it is) or a checkpoint (both buffers;
// dir 2: save, 3: restore)
__global__ void __launch_bounds__(256) k_rd_state(const Tis* __restrict__ T, RdTis* __restrict__ R, int slot,
                                                   int dir, int cb, int cp, int ap) {
  const Tis& t = T[blockIdx.y];
  const RdTis& r = R[blockIdx.y];
  const int n = t.st->n;
  const long long cap = t.cap, ns = t.ns;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const bool ck = dir >= 2, save = (dir & 1) == 0;
  double* gs = (ck ? r.cg : r.sg) + slot * cap * ns;
  RD_FOR_ROW_ENTRIES(t, cb, n, c, i) {
    if (save) gs[i] = t.gj[cb][i];
    else t.gj[cb][i] = gs[i];
  }
  const int nb = ck ? 2 : 1;
  const int start = t.nmemb + t.nder;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
    double* hs = (ck ? r.ch : r.sh) + slot * cap;
    if (save) hs[c] = t.ex_inert[c];
    else t.ex_inert[c] = hs[c];
    if (!ck && save) {
      // grid_avg8 of the ATP buffer at the cell's lattice corner
      double av = 0;
      if (c >= start && is_afm(t.state[c])) {
        const double* A = t.atp[ap];
        const long long b0 = gidx(t, t.lat[3 * c], t.lat[3 * c + 1], t.lat[3 * c + 2]);
        av = 0.125 * (A[b0] + A[b0 + 1] + A[b0 + t.XP] + A[b0 + t.PL] + A[b0 + t.XP + 1] + A[b0 + t.PL + 1]
                      + A[b0 + t.PL + t.XP] + A[b0 + t.PL + t.XP + 1]);
      }
      r.sAv[slot * cap + c] = av;
    }
    for (int b = 0; b < nb; b++) {
      const int buf = b ? 1 - cp : cp;
      double* us = (ck ? r.cu + (2 * slot + b) * cap : r.su + slot * cap);
      double* ps = (ck ? r.cp_ + (2 * slot + b) * cap : r.sp + slot * cap);
      if (save) { us[c] = t.ca2p[buf][c]; ps[c] = t.ip3[buf][c]; }
      else { t.ca2p[buf][c] = us[c]; t.ip3[buf][c] = ps[c]; t.ca_up[buf][c] = make_double2(us[c], ps[c]); }
    }
  }
  if (!ck) return;   // a segment state keeps grid_avg8 per cell, not the grid
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < r.box; v += stride) {
    for (int b = 0; b < nb; b++) {
      const int buf = b ? 1 - ap : ap;
      double* as = r.cA + (2 * slot + b) * r.box;
      if (save) as[v] = t.atp[buf][v];
      else t.atp[buf][v] = as[v];
    }
  }
}

// the loss and the seed of the reverse pass.  Every adjoint buffer starts at
// zero; the sub-steps keep them so (a cell's group clears the cell adjoints
// it has read, the grid part the ATP buffer the next sub-step scatters into).
__global__ void __launch_bounds__(256) k_rd_seed(const Tis* __restrict__ T, RdTis* __restrict__ R, int cp, int cb,
                                                  int q) {
  const Tis& t = T[blockIdx.y];
  const RdTis& r = R[blockIdx.y];
  const int n = t.st->n;
  const int start = t.nmemb + t.nder;
  const long long stride = (long long)gridDim.x * blockDim.x;
  double part = 0;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
    double s = 0;
    if (c >= start && is_afm(t.state[c])) {
      const double d = t.ca2p[cp][c] - r.target[c];
      part += d * d;
      s = 2.0 * d / r.nafm;
    }
    r.lu[q][c] = s; r.lp[q][c] = 0; r.lh[q][c] = 0; r.gkpa[c] = 0;
    r.lu[1 - q][c] = 0; r.lp[1 - q][c] = 0; r.lh[1 - q][c] = 0;
  }
  RD_FOR_ROW_ENTRIES(t, cb, n, c, i) r.lg[q][i] = 0;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < r.box; v += stride) {
    r.lA[0][v] = 0; r.lA[1][v] = 0; r.lA[2][v] = 0;
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
  if ((threadIdx.x & 63) == 0 && part != 0) atomicAdd(&r.scal[0], part / r.nafm);
}

// reverse of dead_IP3_calc and supra_calc for one sub-step, the state before
// the sub-step from slot `slot`.  A group of RG lanes (8 or 16, a launch
// shape) owns one cell: lanes take the coupled list entries (the per-solve
// compact (word, slot) list; gj and its adjoint live at the slot), their partial sums are
// reduced over the group (a gradient need not follow the forward's summation
// order), and each lane scatters its entries' contributions.  A lane keeps
// its first 32 / RG entries (and their neighbour values) in registers for
// both passes, every load of them issued up front; entries past 32 take the
// same steps from memory.  The ATP adjoint of the cell's fa source is
// gathered from the voxels the cell owns (k_rd_own's list).
template <int RG>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = 1; o < RG; o <<= 1) v += __shfl_xor(v, o, RG);
  return v;
}

namespace {
// the contribution of one list entry of an AFM cell to the adjoints before
// the sub-step (a lane's share of the second pass)
struct RdEntry {
  double abar_du_iag;   // abar * du * IAG
  double lp_dp_iag;     // lp * DT_Ca * dp * IAG
  double uc, pc;
  double DT_Ca, wd;
};

__device__ __forceinline__ void rd_entry_bwd(const RdEntry& k, int kd, int o, double gv, double uo, double po,
                                             double lg_in, uint8_t so, double* ng_e, double* nu, double* np,
                                             double& own_u) {
  double to_u = k.abar_du_iag * gv;
  double lge = lg_in;
  if (kd == 3) {
    // d fw / d |d| = -(1 - tanh^2 y) / (2 * 0.1) with y = (wd - |d|) / 0.1,
    // and 1 - tanh^2 y = 4 E / (1 + E)^2 with E = e^(2y)
    const double d = uo - k.uc;
    const double DT_Ca = k.DT_Ca;
    const double E = exp(2.0 * ((k.wd - fabs(d)) / 0.1));
    const double fwd = -5.0 * (4.0 * E / ((1.0 + E) * (1.0 + E)));
    const double sg = d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0);
    to_u += lg_in * DT_Ca * fwd * sg;
    own_u -= lg_in * DT_Ca * fwd * sg;
    lge = lg_in * (1.0 - DT_Ca);
  }
  *ng_e = lge + k.abar_du_iag * (uo - k.uc) + k.lp_dp_iag * (po - k.pc);
  if (is_afm(so)) atomicAdd(&nu[o], to_u);
  atomicAdd(&np[o], k.lp_dp_iag * gv);
}
}  // namespace

template <int RG>
__device__ __forceinline__ void rd_cells_part(const Tis& t, const RdTis& r, int bid, int nblk, int slot, int cb,
                                              int bp, int q, int qa) {
  constexpr int REP = 32 / RG;
  const int start = t.nmemb + t.nder;
  const long long cap = t.cap, ns = t.ns;
  const int lane = threadIdx.x & (RG - 1);
  const int groups = nblk * (blockDim.x / RG);
  const double* u = r.su + slot * cap;
  const double* p = r.sp + slot * cap;
  const double* h = r.sh + slot * cap;
  const double* g = r.sg + slot * cap * ns;
  const double *lu = r.lu[q], *lp = r.lp[q], *lh = r.lh[q], *lg = r.lg[q];
  double *nu = r.lu[1 - q], *np = r.lp[1 - q], *nh = r.lh[1 - q], *ng = r.lg[1 - q];
  const double* lA = r.lA[qa];
  double* nA = r.lA[qa == 2 ? 0 : qa + 1];
  const double du = t.ca_du, dp = t.ca_dp;
  const Params& pm = t.prm;
  const CaK ck = ca_k(pm);
  const double DT_Ca = ck.DT_Ca, Kpp = ck.Kpp, STIM11 = pm.STIM11;
  double pdu = 0, pdp = 0;
  double* const lu_w = r.lu[q];
  double* const lp_w = r.lp[q];
  double* const lh_w = r.lh[q];
  // one cell per group (the grid covers every cell past the membrane and
  // dermis; a launch that does not is reported as ERR_INVALID): a loop over
  // cells would hold the math's loop-invariant constants in registers
  const int c = start + bid * (blockDim.x / RG) + threadIdx.x / RG;
  // the cell's own words are loaded in one round trip with the cell count (a
  // cell past the count reads allocated, stale slots and leaves)
  const int cc = c < cap ? c : (int)cap - 1;
  const uint8_t s0 = t.state[cc];
  const int len0 = t.ca_cn[cc];
  const double lpc0 = lp[cc], uc0 = u[cc], pc0 = p[cc];
  const int n = t.st->n;
  if (bid == 0 && threadIdx.x == 0 && start + (long long)groups < n) set_err(t.st, ERR_INVALID);
  if (c < n) do {   // "continue" below leaves the cell
    const uint8_t s = s0;
    // the coupled entries only, (word, slot of the entry in the row)
    const int len = len0;
    const double lpc = lpc0;
    const double uc = uc0, pc = pc0;
    const int2* ent = t.ca_ce + (size_t)c * ns;
    if (s == DEAD) {
      double cnt = 0, sum = 0;
      for (int e = lane; e < len; e += RG)
        if ((ent[e].x & 3) == 3) {
          const int o = ent[e].x >> 2;
          sum += p[o];
          cnt += 1;
          atomicAdd(&np[o], lpc * DT_Ca * dp);
        }
      cnt = group_sum<RG>(cnt);
      sum = group_sum<RG>(sum);
      if (lane == 0) {
        atomicAdd(&np[c], lpc * (1.0 - DT_Ca * dp * cnt - DT_Ca * Kpp));
        pdp += lpc * DT_Ca * (sum - cnt * pc);
        // consumed: cleared for the sub-step after next, which scatters here
        lu_w[c] = 0; lp_w[c] = 0; lh_w[c] = 0;
      }
      continue;
    }
    if (!is_afm(s)) continue;
    const double* gr = g + (size_t)c * ns;
    const double* lgr = lg + (size_t)c * ns;
    // the lane's first REP entries: packed neighbour and kind, gj, edge
    // adjoint, then the neighbours' state, Ca2+ and IP3, all in flight together
    const int m0 = min(len, RG * REP);
    int o[REP], sl[REP];
    uint8_t kd[REP], so[REP];
    double gv[REP], lgv[REP], uo[REP], po[REP];
    // every load unconditional (a row holds ns >= 32 slots; a masked entry
    // reads slot 0 and cell 0), masked with integer masks afterwards, so that
    // no load sits in a branch that waits for it
    int2 x[REP];
#pragma unroll
    for (int i = 0; i < REP; i++) x[i] = ent[lane + i * RG];
#pragma unroll
    for (int i = 0; i < REP; i++) {
      const int msk = -(int)(lane + i * RG < m0);
      o[i] = (x[i].x & msk) >> 2;
      kd[i] = (uint8_t)(x[i].x & msk & 3);
      sl[i] = x[i].y & msk;
    }
#pragma unroll
    for (int i = 0; i < REP; i++) {
      gv[i] = gr[sl[i]];
      lgv[i] = lgr[sl[i]];
      uo[i] = u[o[i]];
      po[i] = p[o[i]];
      so[i] = t.state[o[i]];
    }
    const int ix = t.lat[3 * c], iy = t.lat[3 * c + 1], iz = t.lat[3 * c + 2];
    const double hc = h[c], luc = lu[c], lhc = lh[c];
    const double iag = t.ca_iag[c], kpa = t.ca_kpa[c], th = t.ca_th[c];
    // ATP adjoint of the voxels the cell owns (after the sub-step)
    // four voxels per lane at a time: their indices, then their values
    double acc_own = 0;
    {
      const int no = r.own_n[c];
      const int* ov = r.own + (size_t)c * r.oc;
      for (int i0 = 0; i0 < no; i0 += 4 * RG) {
        int v[4];
        double a[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int i = i0 + lane + k * RG;
          v[k] = i < no ? ov[i] : -1;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) a[k] = v[k] >= 0 ? lA[v[k]] : 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) acc_own += a[k];
      }
    }
    double su = 0, sp = 0, gsum = 0;
#pragma unroll
    for (int i = 0; i < REP; i++)
      if (kd[i]) {
        su += gv[i] * (uo[i] - uc);
        sp += gv[i] * (po[i] - pc);
        gsum += gv[i];
      }
    for (int e = RG * REP + lane; e < len; e += RG) {
      const int2 en = ent[e];
      const int oe = en.x >> 2;
      const double ge = gr[en.y];
      su += ge * (u[oe] - uc);
      sp += ge * (p[oe] - pc);
      gsum += ge;
    }
    su = group_sum<RG>(su);
    sp = group_sum<RG>(sp);
    gsum = group_sum<RG>(gsum);
    acc_own = group_sum<RG>(acc_own);
    const double Bavg = grid_avg8(t, t.bf[bp], ix, iy, iz);
    const double Aavg = r.sAv[slot * cap + c];
    const double diffu = ca2p_reaction_factor(ck, uc, hc, pc, Bavg) + du * iag * su;
    const double abar = DT_Ca * luc + (diffu > 0 ? acc_own * DT_Ca * STIM11 : 0.0);
    const RdEntry k{abar * du * iag, lpc * DT_Ca * dp * iag, uc, pc, DT_Ca, ck.wd};
    double own_u = 0;
#pragma unroll
    for (int i = 0; i < REP; i++)
      if (kd[i]) rd_entry_bwd(k, kd[i], o[i], gv[i], uo[i], po[i], lgv[i], so[i], &ng[(size_t)c * ns + sl[i]],
                              nu, np, own_u);
    for (int e = RG * REP + lane; e < len; e += RG) {
      const int2 en = ent[e];
      const int oe = en.x >> 2;
      rd_entry_bwd(k, en.x & 3, oe, gr[en.y], u[oe], p[oe], lgr[en.y], t.state[oe], &ng[(size_t)c * ns + en.y], nu,
                   np, own_u);
    }
    own_u = group_sum<RG>(own_u);
    if (lane < 8) {   // the eight grid_avg8 corners of the IP3 production term
      const double H0 = ck.H0;
      const double wA = lpc * DT_Ca * kpa * H0 / ((H0 + Aavg) * (H0 + Aavg)) * 0.125;
      const long long vv = corner_var(t, r, ix + (lane & 1), iy + ((lane >> 1) & 1), iz + ((lane >> 2) & 1));
      if (vv >= 0) atomicAdd(&nA[vv], wA);
    }
    if (lane == 0) {
      const double kflux = ck.k_flux, mu0 = ck.mu0, mu1 = ck.mu1, kmu = ck.kmu, pb = ck.para_b, pbb = ck.para_bb,
                   pk1 = ck.para_k1;
      const double kg = ck.kg, gam = ck.gamma, k2 = ck.para_k2, H0 = ck.H0;
      const double e1 = mu0 + mu1 * pc / (pc + kmu), e2 = pb + pbb * uc / (pk1 + uc);
      const double dRdu = kflux * e1 * (pbb * pk1 / ((pk1 + uc) * (pk1 + uc))) * hc - gam * kg / ((kg + uc) * (kg + uc));
      const double dRdh = kflux * e1 * e2;
      const double dRdp = kflux * (mu1 * kmu / ((pc + kmu) * (pc + kmu))) * e2 * hc;
      const double qq = k2 * k2 + uc * uc;
      atomicAdd(&nu[c], own_u + luc + abar * (dRdu - du * iag * gsum)
                            + lhc * DT_Ca * (-2.0 * uc * k2 * k2 / (qq * qq)) / th);
      atomicAdd(&np[c], lpc * (1.0 - DT_Ca * Kpp - DT_Ca * dp * iag * gsum) + abar * dRdp);
      atomicAdd(&nh[c], lhc * (1.0 - DT_Ca / th) + abar * dRdh);
      pdu += abar * iag * su;
      pdp += lpc * DT_Ca * iag * sp;
      r.gkpa[c] += lpc * DT_Ca * Aavg / (H0 + Aavg);
      lu_w[c] = 0; lp_w[c] = 0; lh_w[c] = 0;
    }
  } while (false);
  for (int o = 32; o > 0; o >>= 1) {
    pdu += __shfl_down(pdu, o, 64);
    pdp += __shfl_down(pdp, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (pdu != 0) atomicAdd(&r.scal[1], pdu);
    if (pdp != 0) atomicAdd(&r.scal[2], pdp);
  }
}

// reverse of ATP_refresh: each point of the box gathers from the points
// whose update read it (periodic x, y; mirror z: prev_z(0) = 1,
// next_z(NZ - 1) = NZ - 1) and adds it to the ATP adjoint before the
// sub-step, beside the grid_avg8 terms the cell part scatters there (both
// with atomics, so the two parts run side by side).  The dead ATP buffer is
// cleared for the next sub-step.
__device__ __forceinline__ void rd_grid_part(const Tis& t, const RdTis& r, int bid, int nblk, int qa) {
  const long long end = box_end(t, r);
  const double* lA = r.lA[qa];
  double* nA = r.lA[qa == 2 ? 0 : qa + 1];
  double* dead = r.lA[qa == 0 ? 2 : qa - 1];
  const uint8_t* m = t.cmap2;
  const int NX = t.NX, NY = t.NY, NZ = t.NZ, izb = r.izb;
  const double DT_Ca = t.prm.DT_Ca, Kaa = t.prm.Kaa;
  const double cD = DT_Ca * t.prm.Da * t.inv_dx * t.inv_dx;
  const long long stride = (long long)nblk * blockDim.x;
  for (long long v = (long long)bid * blockDim.x + threadIdx.x; v < r.box; v += stride) {
    dead[v] = 0;
    if (v >= end) continue;
    const int j = (int)(v % t.XP), k = (int)((v / t.XP) % (NY + 1)), l = (int)(v / t.PL);
    if (j >= NX || k >= NY) continue;
    const int px = j == 0 ? NX - 1 : j - 1, nx = j == NX - 1 ? 0 : j + 1;
    const int py = k == 0 ? NY - 1 : k - 1, ny = k == NY - 1 ? 0 : k + 1;
    const int pz = l == 0 ? 1 : l - 1, nz = l + 1 == NZ ? NZ - 1 : l + 1;
    const double msum = (double)m[gidx(t, px, k, l)] + m[gidx(t, nx, k, l)] + m[gidx(t, j, py, l)]
                        + m[gidx(t, j, ny, l)] + m[gidx(t, j, k, pz)] + m[gidx(t, j, k, nz)];
    const double mv = (double)m[v];
    double acc = lA[v] * (1.0 - cD * msum - DT_Ca * Kaa);
    // in-plane readers of v: its four periodic neighbours
    acc += (lA[gidx(t, px, k, l)] + lA[gidx(t, nx, k, l)] + lA[gidx(t, j, py, l)] + lA[gidx(t, j, ny, l)]) * cD * mv;
    // z readers: w = l - 1 reads v as next_z, w = l + 1 reads v as prev_z;
    // w = 0 reads plane 1 twice; w = NZ - 1 reads itself as next_z
    if (l >= 1) acc += lA[gidx(t, j, k, l - 1)] * cD * mv;
    if (l + 1 < izb) acc += lA[gidx(t, j, k, l + 1)] * cD * mv;
    if (l == 1) acc += lA[gidx(t, j, k, 0)] * cD * mv;
    if (l == NZ - 1) acc += lA[v] * cD * mv;
    if (acc != 0) atomicAdd(&nA[v], acc);
  }
}

// one reverse sub-step: blocks [0, ncb) take the cells, the rest the box
template <int RG>
__global__ void __launch_bounds__(256) k_rd_step(const Tis* __restrict__ T, RdTis* __restrict__ R, int slot, int cb,
                                                  int bp, int q, int qa, int ncb) {
  const Tis& t = T[blockIdx.y];
  const RdTis& r = R[blockIdx.y];
  if ((int)blockIdx.x < ncb) rd_cells_part<RG>(t, r, blockIdx.x, ncb, slot, cb, bp, q, qa);
  else rd_grid_part(t, r, blockIdx.x - ncb, gridDim.x - ncb, qa);
}

// dL/dgj0: the adjoints of every coupled edge at the window's start
__global__ void __launch_bounds__(256) k_rd_gj0(const Tis* __restrict__ T, RdTis* __restrict__ R, int cb, int q) {
  const Tis& t = T[blockIdx.y];
  RdTis& r = R[blockIdx.y];
  const int n = t.st->n;
  const int start = t.nmemb + t.nder;
  double part = 0;
  RD_FOR_ROW_ENTRIES(t, cb, n, c, i)
    if (c >= start && is_afm(t.state[c]) && (t.ca_ent[i] & 3)) part += r.lg[q][i];
  for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
  if ((threadIdx.x & 63) == 0 && part != 0) atomicAdd(&r.scal[3], part);
}

// the voxels each AFM cell owns in the frozen voxel map of a window: they
// lie in the box its sphere stamp covers (k_map_stamp: lat +- 2 ir, z from 1
// to NZ - 1, periodic x and y) and below iz_bound.  One thread per cell,
// stamp-box order; more than oc voxels is an error (the host sizes oc as the
// whole box).
__global__ void __launch_bounds__(256) k_rd_own(const Tis* __restrict__ T, RdTis* __restrict__ R) {
  const Tis& t = T[blockIdx.y];
  const RdTis& r = R[blockIdx.y];
  const int n = t.st->n;
  const int start = t.nmemb + t.nder;
  const int bx = 4 * t.irx + 1, by = 4 * t.iry + 1;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    int m = 0;
    if (c >= start && is_afm(t.state[c])) {
      const unsigned me = (unsigned)c + 1;
      const int ix = t.lat[3 * c], iy = t.lat[3 * c + 1], iz = t.lat[3 * c + 2];
      const int z0 = max(iz - 2 * t.irz, 1), z1 = min(min(iz + 2 * t.irz, t.NZ - 1), r.izb - 1);
      int* ov = r.own + (size_t)c * r.oc;
      for (int l = z0; l <= z1; l++)
        for (int b = 0; b < bx * by; b++) {
          const int jx = (ix - 2 * t.irx + b % bx + t.NX) % t.NX;
          const int ky = (iy - 2 * t.iry + b / bx + t.NY) % t.NY;
          const long long v = gidx(t, jx, ky, l);
          if ((t.vox[v] & 0x7fffffffu) != me) continue;
          if (m < r.oc) ov[m] = (int)v;
          m++;
        }
      if (m > r.oc) { set_err(t.st, ERR_CAPACITY); m = r.oc; }
    }
    r.own_n[c] = m;
  }
}

// ---- launchers -------------------------------------------------------------
namespace {
dim3 cells_grid(long long items, int ntis) {
  long long b = (items + 255) / 256;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return dim3((unsigned)b, ntis);
}
}  // namespace

void launch_rd_snapshot(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int cb, int cp, int ap, long long items) {
  EPI_LAUNCH(K_RD_SNAPSHOT, k_rd_snapshot, cells_grid(items, ntis), dim3(256), 0, s, dT, dR, cb, cp, ap);
}
void launch_rd_own(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int max_n) {
  EPI_LAUNCH(K_RD_OWN, k_rd_own, cells_grid(max_n, ntis), dim3(256), 0, s, dT, dR);
}
void launch_rd_restore(hipStream_t s, const Tis* dT, RdTis* dR, const double* gj0, int ntis, int cb, int cp, int ap,
                       int use_kpa, long long items) {
  EPI_LAUNCH(K_RD_RESTORE, k_rd_restore, cells_grid(items, ntis), dim3(256), 0, s, dT, dR, gj0, cb, cp, ap, use_kpa);
}
void launch_rd_put_gj(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int cb, long long items) {
  EPI_LAUNCH(K_RD_PUT_GJ, k_rd_put_gj, cells_grid(items, ntis), dim3(256), 0, s, dT, dR, cb);
}
void launch_rd_state(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int slot, int dir, int cb, int cp, int ap,
                     long long items) {
  EPI_LAUNCH(K_RD_STATE, k_rd_state, cells_grid(items, ntis), dim3(256), 0, s, dT, dR, slot, dir, cb, cp, ap);
}
void launch_rd_seed(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int cp, int cb, int q, long long items) {
  EPI_LAUNCH(K_RD_SEED, k_rd_seed, cells_grid(items, ntis), dim3(256), 0, s, dT, dR, cp, cb, q);
}
void launch_rd_backward_substep(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int slot, int cb, int bp, int q,
                                int qa, int max_n, long long max_box, int rg) {
  const int ncb = (int)(((long long)max_n * rg + 255) / 256);   // one cell per group: no cap
  const dim3 g(ncb + cells_grid(max_box, ntis).x, ntis);
  if (rg == 16)
    EPI_LAUNCH(K_RD_STEP, k_rd_step<16>, g, dim3(256), 0, s, dT, dR, slot, cb, bp, q, qa, ncb);
  else
    EPI_LAUNCH(K_RD_STEP, k_rd_step<8>, g, dim3(256), 0, s, dT, dR, slot, cb, bp, q, qa, ncb);
}
void launch_rd_gj0(hipStream_t s, const Tis* dT, RdTis* dR, int ntis, int cb, int q, long long items) {
  EPI_LAUNCH(K_RD_GJ0, k_rd_gj0, cells_grid(items, ntis), dim3(256), 0, s, dT, dR, cb, q);
}

}  // namespace epi
