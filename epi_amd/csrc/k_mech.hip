// k_mech.hip -- cell_interaction (cell_interaction.cpp:648-685) on CDNA4.
//
// The reference scatters every pair force into both cells with CAS atomics
// (cell_interaction.cpp:48-66, atomics.h:42-47), so its result depends on the
// order of the additions.  Here every cell gathers its own contributions and
// adds them in the order of the serial loop: bending, wall, its own handler
// over its list, lower-index handlers before and higher-index handlers after
// (cell_interaction.cpp:650-679).  In k_forces a group of FG = 16 lanes owns
// one cell: the lanes compute the cell's terms (neighbour deltas and the 12
// bend dihedrals) into LDS and the group's leader adds them in that order;
// k_forces_long, for rows longer than the LDS slots, gives one lane one
// cell.  No atomics: the FP64 result is the serial one bit for bit.
#include "epi_dev.h"
#include "launch.h"
#include "epi_group.h"
#include "mech_dev.h"

namespace epi {

struct CV {           // what a force law reads of a cell
  P3 p;
  double r, agek, snl;
  int pair, dermis, rdt;
  uint8_t s;
};

__device__ __forceinline__ CV load_cv(const Tis& t, int pb, int c) {
  CV v;
  v.p = ld_pos(t, pb, c);
  v.r = t.radius[c];
  v.agek = t.agek[c];
  v.snl = t.spr_nat_len[c];
  v.pair = t.pair[c];
  v.dermis = t.dermis[c];
  v.rdt = t.rest_div_times[c];
  v.s = t.state[c];
  return v;
}

// The law cell c's handler applies to its list entry q, or LAW_NONE.
// _DER_ :549-561, _FIX_ :581-604, _MUSUME_ :613-634, _AL_AIR_DE_ :563-579.
__device__ __forceinline__ int proc_law(const CV& c, int ci, const CV& q, int qi) {
  switch (c.s) {
    case DER:
      return (q.s == DER && ci > qi) ? LAW_DER_DER : LAW_OTHER;
    case FIX:
      if (q.s == FIX) return (c.pair != qi && ci > qi) ? LAW_FIXMU : LAW_NONE;
      if (q.s == MUSUME) return (c.pair != qi) ? LAW_FIXMU : LAW_NONE;
      if (q.s == MEMB) return (c.dermis == qi) ? LAW_FIX_MEMB : LAW_OTHER;
      return LAW_NONE;
    case MUSUME:
      if (q.s == MEMB) return (c.dermis == qi) ? LAW_MU_MEMB : LAW_OTHER;
      if (q.s == MUSUME && c.pair != qi && ci > qi) return LAW_FIXMU;
      return LAW_NONE;
    case ALIVE: case AIR: case DEAD:
      if (q.s == ALIVE || q.s == AIR || q.s == DEAD) return ci > qi ? LAW_ALAIRDE : LAW_NONE;
      if (q.s == DER) return LAW_NONE;
      return LAW_ALAIRDE;
    default:
      return LAW_NONE;
  }
}

__device__ __forceinline__ double ljmain(const Tis& t, const CV& a, const CV& b) {
  const double distSq = dist_sq(t, a.p, b.p);
  const double rad_sum = a.r + b.r;
  const double LJ6 = POW6(rad_sum) / POW3(distSq);
  return 4.0 * t.m_eps * LJ6 * (LJ6 - 1) / distSq;
}
__device__ __forceinline__ bool is_near(const Tis& t, const CV& a, const CV& b) {
  return dist_sq(t, a.p, b.p) < (a.r + b.r) * (a.r + b.r);
}
__device__ __forceinline__ double adhesion(const Tis& t, const CV& a, const CV& b, double spring_const) {
  const double distlj = sqrt(dist_sq(t, a.p, b.p));
  const double rad_sum = a.r + b.r;
  const double LJ2_m1 = (distlj / rad_sum) - 1;
  return (LJ2_m1 + 1 > t.prm.LJ_THRESH ? 0.0
          : -(spring_const / distlj) * LJ2_m1 *
                (1 - LJ2_m1 * LJ2_m1 / ((t.prm.LJ_THRESH - 1.0) * (t.prm.LJ_THRESH - 1.0))));
}

// Coefficient of law `law` with c1 = a, c2 = b (cell_interaction.cpp:77-285).
__device__ double coef_law(const Tis& t, int law, const CV& a, const CV& b, Status* st) {
  switch (law) {
    case LAW_OTHER:
      return ljmain(t, a, b);
    case LAW_DER_DER: {
      const double dist = sqrt(dist_sq(t, a.p, b.p));
      if (dist + t.prm.delta_R < a.r + b.r) {
        const double dist2 = dist + t.prm.delta_R;
        const double rad_sum = a.r + b.r;
        const double LJ6 = POW6(rad_sum) / POW6(dist2);
        return 4.0 * t.m_eps * LJ6 * (LJ6 - 1) / (dist * dist2) + t.prm.para_ljp2;
      } else if (is_near(t, a, b)) {
        return t.prm.para_ljp2;
      } else {
        const double dist2 = dist + t.prm.delta_R;
        const double rad_sum = a.r + b.r;
        const double LJ6 = POW6(rad_sum) / POW6(dist2);
        return 4.0 * t.m_eps * LJ6 * (LJ6 - 1) / (dist * dist) + t.prm.para_ljp2;
      }
    }
    case LAW_ALAIRDE:
      if (is_near(t, a, b)) return ljmain(t, a, b);
      else {
        double spf = (t.m_kadh * t.prm.K_DESMOSOME_RATIO);
        if (a.agek > t.prm.THRESH_SP && b.agek > t.prm.THRESH_SP) spf = t.m_kadh;
        return adhesion(t, a, b, spf);
      }
    case LAW_FIXMU:
      if (is_near(t, a, b)) return ljmain(t, a, b);
      return adhesion(t, a, b, (t.m_kadh * t.prm.K_DESMOSOME_RATIO));
    case LAW_MU_MEMB: {
      if (is_near(t, a, b)) return ljmain(t, a, b);
      // spring_force_to_memb, set at the top of _MUSUME_interaction (:613-620)
      double spring = 0;
      if (a.pair >= 0 && t.state[a.pair] == FIX) spring = t.m_kspring;
      else if (a.rdt > 0) spring = t.prm.Kspring_d;
      return adhesion(t, a, b, spring);
    }
    case LAW_FIX_MEMB:
      if (is_near(t, a, b)) return ljmain(t, a, b);
      else {
        const double distlj = sqrt(dist_sq(t, a.p, b.p));
        const double LJ2 = distlj / (a.r + b.r);
        return -(t.m_kspring / distlj) * (LJ2 - 1.0);
      }
    case LAW_MEMB_MEMB: {
      const double rad_sum = a.r + b.r;
      const double rad_sum_sq = rad_sum * rad_sum;
      double cr_dist_sq;
      const double distSq = dist_sq(t, a.p, b.p);
      double tmp;
      if (distSq < (cr_dist_sq = rad_sum_sq * t.prm.P_MEMB * t.prm.P_MEMB)) {
        const double LJ6 = POW3(cr_dist_sq) / POW3(distSq);
        tmp = 4.0 * t.m_eps * LJ6 * (LJ6 - 1) / distSq;
      } else if (distSq < rad_sum_sq) {
        const double distlj = sqrt(distSq);
        const double cr_dist = rad_sum * t.prm.P_MEMB;
        tmp = -(t.prm.DER_DER_CONST / distlj) * (distlj / cr_dist - 1.0);
      } else {
        const double distlj = sqrt(distSq);
        const double lambda_dist = (1.0 + t.prm.P_MEMB) * rad_sum;
        const double cr_dist = rad_sum * t.prm.P_MEMB;
        const double LJ6 = POW6(cr_dist) / POW6(lambda_dist - distlj);
        tmp = -(t.prm.DER_DER_CONST / rad_sum) * ((1.0 - t.prm.P_MEMB) / t.prm.P_MEMB)
              - 4 * t.m_eps * (LJ6 * (LJ6 - 1.0)) / ((lambda_dist - distlj) * distlj);
      }
      if (fabs(tmp) > 100) set_err(st, ERR_STRONG_INTERACTION);
      return tmp;
    }
    case LAW_PAIR: {
      const double dist = sqrt(dist_sq(t, a.p, b.p));
      const double force = t.prm.Kspring_division * (dist - a.snl);
      return -force / dist;
    }
    default:
      return 0.0;
  }
}

// cell_interaction_apply (:48-66): the displacement added to c1, subtracted from c2
__device__ __forceinline__ P3 delta(const Tis& t, int law, const CV& a, const CV& b, Status* st) {
  const double coef = coef_law(t, law, a, b, st);
  if (fabs(coef) > 1000) set_err(st, ERR_STRONG_INTERACTION);
  const double d = t.prm.DT_Cell * coef;
  return P3{d * p_diff_x(t, a.p.x, b.p.x), d * p_diff_y(t, a.p.y, b.p.y), d * (a.p.z - b.p.z)};
}

// wall_interaction (:287-294): the z increment
__device__ __forceinline__ double wall_dz(const Tis& t, const CV& c) {
  const double distlj = 2.0 * c.p.z;
  const double LJ6 = POW6(c.r) / POW6(c.p.z);
  const double ljm = 4.0 * t.m_eps * LJ6 * (LJ6 - 1.0) / (distlj * distlj);
  return t.prm.DT_Cell * ljm * 2.0 * c.p.z;
}

// ---- triangular bending (:451-499) ---------------------------------------
__device__ __forceinline__ P3 v_cross(P3 a, P3 b) {
  return P3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double v_dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double v_norm(P3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ P3 v_normalize(P3& a) {   // vec3::normalize mutates
  double nm = v_norm(a);
  a.x /= nm; a.y /= nm; a.z /= nm;
  return a;
}
__device__ __forceinline__ P3 v_scale(double c, P3 v) { return P3{c * v.x, c * v.y, c * v.z}; }
__device__ __forceinline__ P3 v_div(P3 v, double c) { return P3{v.x / c, v.y / c, v.z / c}; }
__device__ __forceinline__ P3 v_sub(P3 a, P3 b) { return P3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ P3 v_add(P3 a, P3 b) { return P3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ P3 vpsub(const Tis& t, P3 a, P3 b) {
  return P3{p_diff_x(t, a.x, b.x), p_diff_y(t, a.y, b.y), a.z - b.z};
}
__device__ __forceinline__ double bend_factor(double cdot) {   // :349-360
  if (cdot >= 0) return (1.0 - cdot);
  return 1.0 / ((1.0 + cdot) * (1.0 + cdot));
}
__device__ __forceinline__ P3 tri_rho(P3 n1, P3 b1) { return v_sub(b1, v_scale(v_dot(n1, b1), n1)); }

__device__ P3 tri_bend_1(const Tis& t, P3 r1v, P3 a1v, P3 b1v, P3 c1v) {   // :468-474
  P3 norm1_pre = v_cross(vpsub(t, r1v, b1v), vpsub(t, a1v, b1v));
  P3 norm1 = v_normalize(norm1_pre);
  P3 n2 = v_cross(vpsub(t, a1v, b1v), vpsub(t, c1v, b1v));
  P3 norm2 = v_normalize(n2);
  double dot1 = v_dot(norm1, norm2);
  return v_scale(bend_factor(dot1), v_div(v_cross(vpsub(t, a1v, b1v), tri_rho(norm1, norm2)), v_norm(norm1_pre)));
}
__device__ P3 tri_bend_2(const Tis& t, P3 r1v, P3 a1v, P3 b1v, P3 c1v) {   // :476-486
  P3 norm1_pre = v_cross(vpsub(t, r1v, c1v), vpsub(t, b1v, c1v));
  P3 norm2_pre = v_cross(vpsub(t, r1v, b1v), vpsub(t, a1v, b1v));
  P3 norm1 = v_normalize(norm1_pre), norm2 = v_normalize(norm2_pre);
  double dot1 = v_dot(norm1, norm2);
  return v_scale(bend_factor(dot1),
                 v_add(v_div(v_cross(vpsub(t, b1v, c1v), tri_rho(norm1, norm2)), v_norm(norm1_pre)),
                       v_div(v_cross(vpsub(t, a1v, b1v), tri_rho(norm2, norm1)), v_norm(norm2_pre))));
}

// ---- serial form: one lane does a whole cell --------------------------------
// Used for rows longer than the LDS slots of the cooperative kernel.
__device__ void forces_serial(const Tis& t, int pb, int cb, int X, Status* st) {
  const CV cx = load_cv(t, pb, X);
  P3 acc = cx.p;
  const int* row = t.col[cb] + (size_t)X * t.ns;
  const int len = t.nlen[cb][X];
  if (X < t.nmemb) {
    const int* md = t.memb6 + 6 * X;
    P3 dr = P3{0, 0, 0};
    for (int i = 0; i < 6; i++) {
      const int i1 = (i + 1) % 6;
      dr = v_add(dr, tri_bend_1(t, cx.p, ld_pos(t, pb, md[i]), ld_pos(t, pb, md[i1]),
                                ld_pos(t, pb, t.memb6[6 * md[i1] + i])));
    }
    for (int i = 0; i < 6; i++)
      dr = v_add(dr, tri_bend_2(t, cx.p, ld_pos(t, pb, md[i]), ld_pos(t, pb, md[(i + 1) % 6]),
                                ld_pos(t, pb, md[(i + 2) % 6])));
    acc.x += t.prm.DT_Cell * t.m_kbend * dr.x;
    acc.y += t.prm.DT_Cell * t.m_kbend * dr.y;
    acc.z += t.prm.DT_Cell * t.m_kbend * dr.z;
    if (cx.p.z < cx.r) acc.z += wall_dz(t, cx);
    int hi[6], nhi = 0;
    for (int i = 0; i < 6; i++) {
      const int q = md[i];
      if (t.state[q] != MEMB) continue;
      if (X > q) {
        const P3 d = delta(t, LAW_MEMB_MEMB, cx, load_cv(t, pb, q), st);
        acc.x += d.x; acc.y += d.y; acc.z += d.z;
      } else if (q > X) {
        int k = nhi++;
        while (k > 0 && hi[k - 1] > q) { hi[k] = hi[k - 1]; k--; }
        hi[k] = q;
      }
    }
    for (int i = 0; i < nhi; i++) {
      const P3 d = delta(t, LAW_MEMB_MEMB, load_cv(t, pb, hi[i]), cx, st);
      acc.x -= d.x; acc.y -= d.y; acc.z -= d.z;
    }
    for (int i = MEMB_ADJ; i < len; i++) {
      const int p = row[i];
      const CV cp = load_cv(t, pb, p);
      const int law = proc_law(cp, p, cx, X);
      if (law != LAW_NONE) {
        const P3 d = delta(t, law, cp, cx, st);
        acc.x -= d.x; acc.y -= d.y; acc.z -= d.z;
      }
    }
  } else {
    const int nlo = t.nlo[X];
    // lower-index handlers in ascending order (selection over the lower part)
    int last = -1;
    for (int r = 0; r < nlo; r++) {
      int best = 0x7fffffff;
      for (int i = 0; i < nlo; i++) { const int v = row[i]; if (v > last && v < best) best = v; }
      last = best;
      if (t.state[best] == MEMB) continue;
      const CV cp = load_cv(t, pb, best);
      const int law = proc_law(cp, best, cx, X);
      if (law != LAW_NONE) {
        const P3 d = delta(t, law, cp, cx, st);
        acc.x -= d.x; acc.y -= d.y; acc.z -= d.z;
      }
    }
    if (cx.s == DER && cx.p.z < cx.r) acc.z += wall_dz(t, cx);
    for (int i = 0; i < len; i++) {
      const int q = row[i];
      const CV cq = load_cv(t, pb, q);
      const int law = proc_law(cx, X, cq, q);
      if (law != LAW_NONE) {
        const P3 d = delta(t, law, cx, cq, st);
        acc.x += d.x; acc.y += d.y; acc.z += d.z;
      }
    }
    if (cx.pair >= 0 && X > cx.pair) {
      const P3 d = delta(t, LAW_PAIR, cx, load_cv(t, pb, cx.pair), st);
      acc.x += d.x; acc.y += d.y; acc.z += d.z;
    }
    int P = cx.pair > X ? cx.pair : -1;
    for (int i = nlo; i < len; i++) {
      const int p = row[i];
      if (P >= 0 && P < p) {
        const P3 d = delta(t, LAW_PAIR, load_cv(t, pb, P), cx, st);
        acc.x -= d.x; acc.y -= d.y; acc.z -= d.z;
        P = -1;
      }
      const CV cp = load_cv(t, pb, p);
      const int law = proc_law(cp, p, cx, X);
      if (law != LAW_NONE) {
        const P3 d = delta(t, law, cp, cx, st);
        acc.x -= d.x; acc.y -= d.y; acc.z -= d.z;
      }
      if (p == P) {
        const P3 d = delta(t, LAW_PAIR, cp, cx, st);
        acc.x -= d.x; acc.y -= d.y; acc.z -= d.z;
        P = -1;
      }
    }
    if (P >= 0) {
      const P3 d = delta(t, LAW_PAIR, load_cv(t, pb, P), cx, st);
      acc.x -= d.x; acc.y -= d.y; acc.z -= d.z;
    }
  }
  t.pos[1 - pb][0][X] = acc.x;
  t.pos[1 - pb][1][X] = acc.y;
  t.pos[1 - pb][2][X] = acc.z;
}

// ---- cooperative form --------------------------------------------------------
// A group of FG lanes owns one cell.  Every list entry e is handled by lane
// e % FG, which evaluates both displacements the pair can contribute to the
// cell: "own" (the cell's handler processes the entry, added) and "oth" (the
// entry's handler processes the cell, subtracted).  The group leader then
// adds them in the serial order of cell_interaction (:650-679): bending,
// wall, lower-index handlers ascending, the cell's own list, its pair,
// higher-index handlers ascending with a higher-index partner's spring
// after that partner's list.  Lane 0..11 evaluate the 12 dihedral terms of a
// membrane particle.
//
// An absent "own" term is stored as -0.0 and an absent "oth" term as +0.0:
// x + (-0.0) and x - (+0.0) are x for every x, so the leader adds and
// subtracts every slot without flags or branches, four slots' LDS reads
// ahead of their adds, in the same order and with the same results.
//
// Each group takes one cell, so the grid must cover all cells (a launch that
// does not is reported as ERR_INVALID); a loop over cells would keep the
// math's loop-invariant constants in registers and cost occupancy.
//
// FCAP, the LDS slots per cell, is chosen per launch from the longest list
// built so far: 40 keeps a 128-lane block's LDS small enough for 4 waves per
// SIMD, 48 and 64 take longer lists at lower occupancy, and rows beyond 64 go
// to k_forces_long.
constexpr int FG = 16;
constexpr int F_BLOCK = 128;
constexpr int F_GROUPS = F_BLOCK / FG;

// a group's LDS slots, padded by 8 bytes so that the leaders of a wave's
// groups read different banks
template <int FCAP>
struct FRow {
  P3 v[FCAP];
  double pad;
};

__device__ __forceinline__ void p3_add(P3& a, const P3& d) { a.x += d.x; a.y += d.y; a.z += d.z; }
__device__ __forceinline__ void p3_sub(P3& a, const P3& d) { a.x -= d.x; a.y -= d.y; a.z -= d.z; }

template <int FCAP>
__global__ void __launch_bounds__(F_BLOCK) k_forces(const Tis* __restrict__ T, int pb, int cb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  __shared__ FRow<FCAP> s_own[F_GROUPS];
  __shared__ FRow<FCAP> s_oth[F_GROUPS];
  __shared__ int s_idx[F_GROUPS][FCAP];
  __shared__ uint8_t s_flag[F_GROUPS][FCAP];
  __shared__ uint8_t s_perm[F_GROUPS][FCAP];
  __shared__ P3 s_bend[F_GROUPS][12];
  const int lane = threadIdx.x & (FG - 1);
  const int grp = threadIdx.x / FG;
  const int X = blockIdx.x * F_GROUPS + grp;
  if (X >= t.cap) return;
  // the cell's own words (and its lane's first list word) are loaded before
  // the cell count arrives, in one round trip with it: a cell past the count
  // reads allocated, stale slots and leaves once the count is in
  const int len = t.nlen[cb][X];
  const CV cx = load_cv(t, pb, X);
  const bool memb = X < t.nmemb;
  const int nA = memb ? 0 : t.nlo[X];
  const int* row = t.col[cb] + (size_t)X * t.ns;
  const int q0 = row[lane];   // a row holds ns >= 64 > FG slots
  const int n = st->n;
  if (blockIdx.x == 0 && threadIdx.x == 0 && (long long)gridDim.x * F_GROUPS < n) set_err(st, ERR_INVALID);
  if (X >= n) return;
  if (len > FCAP) return;   // k_forces_long
  for (int e = lane; e < len; e += FG) {
    const int q = e == lane ? q0 : row[e];
    const CV cq = load_cv(t, pb, q);
    uint8_t f = 0;
    P3 own = P3{-0.0, -0.0, -0.0}, oth = P3{0, 0, 0};
    if (memb) {
      if (e < MEMB_ADJ) {
        if (cq.s == MEMB) {
          if (X > q) { own = delta(t, LAW_MEMB_MEMB, cx, cq, st); f |= 1; }
          else if (q > X) { oth = delta(t, LAW_MEMB_MEMB, cq, cx, st); f |= 2; }
        }
      } else {
        const int law2 = proc_law(cq, q, cx, X);
        if (law2 != LAW_NONE) { oth = delta(t, law2, cq, cx, st); f |= 2; }
      }
    } else {
      const int law = proc_law(cx, X, cq, q);
      if (law != LAW_NONE) { own = delta(t, law, cx, cq, st); f |= 1; }
      if (e >= nA || cq.s != MEMB) {
        const int law2 = proc_law(cq, q, cx, X);
        if (law2 != LAW_NONE) { oth = delta(t, law2, cq, cx, st); f |= 2; }
      }
    }
    s_idx[grp][e] = q;
    s_flag[grp][e] = f;
    s_own[grp].v[e] = own;
    s_oth[grp].v[e] = oth;
  }
  if (memb && lane < 12) {
    const int* md = t.memb6 + 6 * X;
    const int i = lane < 6 ? lane : lane - 6;
    const P3 a = ld_pos(t, pb, md[i]);
    const P3 b = ld_pos(t, pb, md[(i + 1) % 6]);
    if (lane < 6)
      s_bend[grp][lane] = tri_bend_1(t, cx.p, a, b, ld_pos(t, pb, t.memb6[6 * md[(i + 1) % 6] + i]));
    else
      s_bend[grp][lane] = tri_bend_2(t, cx.p, a, b, ld_pos(t, pb, md[(i + 2) % 6]));
  }
  wave_sync();
  // rank of each lower-part entry by cell index
  for (int e = lane; e < nA; e += FG) {
    const int v = s_idx[grp][e];
    int r = 0;
    for (int k = 0; k < nA; k++) r += s_idx[grp][k] < v;
    s_perm[grp][r] = (uint8_t)e;
  }
  wave_sync();
  if (lane != 0) return;
  const P3* own = s_own[grp].v;
  const P3* oth = s_oth[grp].v;
  P3 acc = cx.p;
  if (memb) {
    P3 dr = P3{0, 0, 0};
    for (int i = 0; i < 12; i++) dr = v_add(dr, s_bend[grp][i]);
    acc.x += t.prm.DT_Cell * t.m_kbend * dr.x;
    acc.y += t.prm.DT_Cell * t.m_kbend * dr.y;
    acc.z += t.prm.DT_Cell * t.m_kbend * dr.z;
    if (cx.p.z < cx.r) acc.z += wall_dz(t, cx);
    int hi[6], nhi = 0;
    for (int e = 0; e < MEMB_ADJ; e++) {
      p3_add(acc, own[e]);
      if (s_flag[grp][e] & 2) {
        int k = nhi++;
        const int q = s_idx[grp][e];
        while (k > 0 && s_idx[grp][hi[k - 1]] > q) { hi[k] = hi[k - 1]; k--; }
        hi[k] = e;
      }
    }
    for (int i = 0; i < nhi; i++) p3_sub(acc, oth[hi[i]]);
    int e = MEMB_ADJ;
#pragma unroll 1
    for (; e + 4 <= len; e += 4) {
      P3 d[4];
#pragma unroll
      for (int j = 0; j < 4; j++) d[j] = oth[e + j];
#pragma unroll
      for (int j = 0; j < 4; j++) p3_sub(acc, d[j]);
    }
    for (; e < len; e++) p3_sub(acc, oth[e]);
  } else {
    {
      int r = 0;
#pragma unroll 1
      for (; r + 4 <= nA; r += 4) {
        int k[4];
        P3 d[4];
#pragma unroll
        for (int j = 0; j < 4; j++) k[j] = s_perm[grp][r + j];
#pragma unroll
        for (int j = 0; j < 4; j++) d[j] = oth[k[j]];
#pragma unroll
        for (int j = 0; j < 4; j++) p3_sub(acc, d[j]);
      }
      for (; r < nA; r++) p3_sub(acc, oth[s_perm[grp][r]]);
    }
    if (cx.s == DER && cx.p.z < cx.r) acc.z += wall_dz(t, cx);
    {
      int e = 0;
#pragma unroll 1
      for (; e + 4 <= len; e += 4) {
        P3 d[4];
#pragma unroll
        for (int j = 0; j < 4; j++) d[j] = own[e + j];
#pragma unroll
        for (int j = 0; j < 4; j++) p3_add(acc, d[j]);
      }
      for (; e < len; e++) p3_add(acc, own[e]);
    }
    if (cx.pair >= 0 && X > cx.pair) p3_add(acc, delta(t, LAW_PAIR, cx, load_cv(t, pb, cx.pair), st));
    int P = cx.pair > X ? cx.pair : -1;
    if (P < 0) {
      int e = nA;
#pragma unroll 1
      for (; e + 4 <= len; e += 4) {
        P3 d[4];
#pragma unroll
        for (int j = 0; j < 4; j++) d[j] = oth[e + j];
#pragma unroll
        for (int j = 0; j < 4; j++) p3_sub(acc, d[j]);
      }
      for (; e < len; e++) p3_sub(acc, oth[e]);
    } else {
      for (int e = nA; e < len; e++) {
        const int p = s_idx[grp][e];
        if (P >= 0 && P < p) {
          p3_sub(acc, delta(t, LAW_PAIR, load_cv(t, pb, P), cx, st));
          P = -1;
        }
        p3_sub(acc, oth[e]);
        if (p == P) {
          p3_sub(acc, delta(t, LAW_PAIR, load_cv(t, pb, P), cx, st));
          P = -1;
        }
      }
      if (P >= 0) p3_sub(acc, delta(t, LAW_PAIR, load_cv(t, pb, P), cx, st));
    }
  }
  t.pos[1 - pb][0][X] = acc.x;
  t.pos[1 - pb][1][X] = acc.y;
  t.pos[1 - pb][2][X] = acc.z;
}

// cells whose rows exceed the LDS slots (cap) of k_forces, one lane per cell
__global__ void __launch_bounds__(256) k_forces_long(const Tis* __restrict__ T, int pb, int cb, int cap) {
  const Tis& t = T[blockIdx.y];
  const int n = t.st->n;
  for (int X = blockIdx.x * blockDim.x + threadIdx.x; X < n; X += gridDim.x * blockDim.x)
    if (t.nlen[cb][X] > cap) forces_serial(t, pb, cb, X, t.st);
}

int launch_forces_rows(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int max_other, int max_len) {
  dim3 g((max_other + F_GROUPS - 1) / F_GROUPS, ntis);
  int cap;
  if (max_len <= 40) {
    cap = 40;
    EPI_LAUNCH(K_FORCES, k_forces<40>, g, dim3(F_BLOCK), 0, s, dT, pb, cb);
  } else if (max_len <= 48) {
    cap = 48;
    EPI_LAUNCH(K_FORCES, k_forces<48>, g, dim3(F_BLOCK), 0, s, dT, pb, cb);
  } else {
    cap = FORCES_ROW_CAP;
    EPI_LAUNCH(K_FORCES, k_forces<FORCES_ROW_CAP>, g, dim3(F_BLOCK), 0, s, dT, pb, cb);
  }
  return cap;
}

void launch_forces_long(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int max_other, int cap) {
  dim3 gl((max_other + 255) / 256, ntis);
  EPI_LAUNCH(K_FORCES_LONG, k_forces_long, gl, dim3(256), 0, s, dT, pb, cb, cap);
}

void launch_forces(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int max_memb, int max_other,
                   int max_len) {
  (void)max_memb;
  const int cap = launch_forces_rows(s, dT, ntis, pb, cb, max_other, max_len);
  if (max_len > cap) launch_forces_long(s, dT, ntis, pb, cb, max_other, cap);
}

// ==== frozen mechanics window: the reverse pass ==============================
// The window is the plain k_forces step repeated with lists, states, pairs
// and dermis frozen.  Which interactions apply depends only on that frozen
// data, so the reverse pass enumerates them from each handler cell (its own
// list, its pair, the membrane slots), evaluates the derivative of the
// active branch of each law at the step's positions, and scatters
// lam_c1 - lam_c2 through the pair's Jacobian with FP64 atomics.

namespace {
struct LawD { double coef, ds, deps, dadh, dspr; };

__device__ LawD ljmain_d(double s, double R, double eps) {
  const double L = POW6(R) / POW3(s);
  return LawD{4.0 * eps * L * (L - 1) / s, 4.0 * eps * (4.0 * L - 7.0 * L * L) / (s * s), 4.0 * L * (L - 1) / s, 0, 0};
}
__device__ LawD adhesion_d(double s, double R, double K, double lj) {
  LawD r{0, 0, 0, 0, 0};
  const double d = sqrt(s);
  const double x = d / R - 1;
  if (x + 1 > lj) return r;
  const double c2 = (lj - 1.0) * (lj - 1.0);
  const double g = x * (1 - x * x / c2);
  r.coef = -(K / d) * g;
  const double dg = (1 - 3 * x * x / c2) / R;
  r.ds = (K * g / (d * d) - (K / d) * dg) / (2 * d);
  return r;
}

// coef of `law` (c1 = a, c2 = b) and its derivatives with respect to the
// squared distance and to eps_m, K_TOTAL, Kspring
__device__ LawD coef_law_d(const Tis& t, int law, const CV& a, const CV& b) {
  const double s = dist_sq(t, a.p, b.p);
  const double R = a.r + b.r;
  const double eps = t.m_eps;
  LawD r{0, 0, 0, 0, 0};
  switch (law) {
    case LAW_OTHER:
      return ljmain_d(s, R, eps);
    case LAW_DER_DER: {
      const double d = sqrt(s), D = d + t.prm.delta_R;
      const double L = POW6(R) / POW6(D);
      const double dL = -6.0 * L / D;
      double g, dgd;
      if (D < R) {
        g = L * (L - 1) / (d * D);
        dgd = (2 * L - 1) * dL / (d * D) - L * (L - 1) * (D + d) / ((d * D) * (d * D));
      } else if (s < R * R) {
        return r;
      } else {
        g = L * (L - 1) / (d * d);
        dgd = (2 * L - 1) * dL / (d * d) - 2 * L * (L - 1) / (d * d * d);
      }
      return LawD{4.0 * eps * g + t.prm.para_ljp2, 4.0 * eps * dgd / (2 * d), 4.0 * g, 0, 0};
    }
    case LAW_ALAIRDE:
      if (s < R * R) return ljmain_d(s, R, eps);
      {
        const bool strong = a.agek > t.prm.THRESH_SP && b.agek > t.prm.THRESH_SP;
        r = adhesion_d(s, R, strong ? t.m_kadh : t.m_kadh * t.prm.K_DESMOSOME_RATIO, t.prm.LJ_THRESH);
        r.dadh = r.coef / t.m_kadh;
        return r;
      }
    case LAW_FIXMU:
      if (s < R * R) return ljmain_d(s, R, eps);
      r = adhesion_d(s, R, t.m_kadh * t.prm.K_DESMOSOME_RATIO, t.prm.LJ_THRESH);
      r.dadh = r.coef / t.m_kadh;
      return r;
    case LAW_MU_MEMB: {
      if (s < R * R) return ljmain_d(s, R, eps);
      const bool fixpair = a.pair >= 0 && t.state[a.pair] == FIX;
      const double spring = fixpair ? t.m_kspring : (a.rdt > 0 ? t.prm.Kspring_d : 0.0);
      r = adhesion_d(s, R, spring, t.prm.LJ_THRESH);
      if (fixpair) r.dspr = r.coef / t.m_kspring;
      return r;
    }
    case LAW_FIX_MEMB:
      if (s < R * R) return ljmain_d(s, R, eps);
      {
        const double d = sqrt(s);
        return LawD{-(t.m_kspring / d) * (d / R - 1.0), -t.m_kspring / (2 * d * d * d), 0, 0,
                    -(1.0 / d) * (d / R - 1.0)};
      }
    case LAW_MEMB_MEMB: {
      const double cr = R * t.prm.P_MEMB, cr2 = cr * cr;
      if (s < cr2) {
        const double L = POW3(cr2) / POW3(s);
        return LawD{4.0 * eps * L * (L - 1) / s, 4.0 * eps * (4.0 * L - 7.0 * L * L) / (s * s),
                    4.0 * L * (L - 1) / s, 0, 0};
      } else if (s < R * R) {
        const double d = sqrt(s);
        return LawD{-(t.prm.DER_DER_CONST / d) * (d / cr - 1.0), -t.prm.DER_DER_CONST / (2 * d * d * d), 0, 0, 0};
      } else {
        const double d = sqrt(s);
        const double lam = (1.0 + t.prm.P_MEMB) * R;
        const double m = lam - d;
        const double L = POW6(cr) / POW6(m);
        const double dL = 6.0 * L / m;
        const double q = L * (L - 1.0) / (m * d);
        const double dq = (2 * L - 1) * dL / (m * d) + L * (L - 1.0) * (d - m) / ((m * d) * (m * d));
        return LawD{-(t.prm.DER_DER_CONST / R) * ((1.0 - t.prm.P_MEMB) / t.prm.P_MEMB) - 4 * eps * q, -4 * eps * dq / (2 * d), -4 * q,
                    0, 0};
      }
    }
    case LAW_PAIR: {
      const double d = sqrt(s);
      return LawD{-(t.prm.Kspring_division * (d - a.snl)) / d, -t.prm.Kspring_division * a.snl / (2 * d * d * d), 0, 0, 0};
    }
    default:
      return r;
  }
}

__device__ __forceinline__ void atomic_add3(double* lam, long long cap, int c, P3 g) {
  atomicAdd(&lam[c], g.x);
  atomicAdd(&lam[cap + c], g.y);
  atomicAdd(&lam[2 * cap + c], g.z);
}
__device__ __forceinline__ P3 ld3(const double* lam, long long cap, int c) {
  return P3{lam[c], lam[cap + c], lam[2 * cap + c]};
}

struct ParAcc { double eps, adh, spr; };

// scatters the pair's term to ib and returns the one for ia (the caller's
// cell, whose lanes reduce their terms before one atomic)
__device__ __forceinline__ P3 pair_bwd(const Tis& t, int law, const CV& a, int ia, const CV& b, int ib,
                                       const double* lam, double* nl, long long cap, ParAcc& pa) {
  const LawD ld = coef_law_d(t, law, a, b);
  const P3 r{p_diff_x(t, a.p.x, b.p.x), p_diff_y(t, a.p.y, b.p.y), a.p.z - b.p.z};
  const P3 la = ld3(lam, cap, ia), lb = ld3(lam, cap, ib);
  const P3 av{la.x - lb.x, la.y - lb.y, la.z - lb.z};
  const double ar = v_dot(av, r);
  const double k2 = 2 * ld.ds * ar;
  const P3 g{t.prm.DT_Cell * (ld.coef * av.x + k2 * r.x), t.prm.DT_Cell * (ld.coef * av.y + k2 * r.y),
             t.prm.DT_Cell * (ld.coef * av.z + k2 * r.z)};
  atomic_add3(nl, cap, ib, P3{-g.x, -g.y, -g.z});
  pa.eps += t.prm.DT_Cell * ld.deps * ar;
  pa.adh += t.prm.DT_Cell * ld.dadh * ar;
  pa.spr += t.prm.DT_Cell * ld.dspr * ar;
  return g;
}

// adjoint of n = N / |N|
__device__ __forceinline__ P3 unit_bar(P3 N, P3 nb) {
  const double nn = v_norm(N);
  const P3 n = v_div(N, nn);
  return v_div(v_sub(nb, v_scale(v_dot(nb, n), n)), nn);
}
__device__ __forceinline__ double bend_factor_d(double c) {
  if (c >= 0) return -1.0;
  return -2.0 / ((1.0 + c) * (1.0 + c) * (1.0 + c));
}
// y = scale f(d) (v x (n2 - d n1)), d = n1.n2
__device__ __forceinline__ void bend_pair_bar(P3 n1, P3 n2, P3 v, double scale, P3 ob, P3& vb, P3& n1b, P3& n2b,
                                              double& db) {
  const double d = v_dot(n1, n2);
  const P3 rho = v_sub(n2, v_scale(d, n1));
  const P3 oy = v_scale(scale * bend_factor(d), ob);
  db += scale * bend_factor_d(d) * v_dot(ob, v_cross(v, rho));
  vb = v_add(vb, v_cross(rho, oy));
  const P3 rb = v_cross(oy, v);
  n2b = v_add(n2b, rb);
  db += -v_dot(rb, n1);
  n1b = v_add(n1b, v_scale(-d, rb));
}
// reverse of tri_bend_1(r, a, b, c); bar[0..3] get the adjoints of r, a, b, c
__device__ void tri_bend_1_bar(const Tis& t, P3 rv, P3 avp, P3 bv, P3 cv, P3 ob, P3* bar) {
  const P3 u = vpsub(t, rv, bv), v = vpsub(t, avp, bv), w = vpsub(t, cv, bv);
  P3 N1 = v_cross(u, v), N2 = v_cross(v, w);
  P3 t1 = N1, t2 = N2;
  const P3 n1 = v_normalize(t1), n2 = v_normalize(t2);
  const double div = v_norm(t1);
  P3 vb{0, 0, 0}, n1b{0, 0, 0}, n2b{0, 0, 0};
  double db = 0;
  bend_pair_bar(n1, n2, v, 1.0 / div, ob, vb, n1b, n2b, db);
  n1b = v_add(n1b, v_scale(db, n2));
  n2b = v_add(n2b, v_scale(db, n1));
  const P3 N1b = unit_bar(N1, n1b), N2b = unit_bar(N2, n2b);
  const P3 ub = v_cross(v, N1b);
  vb = v_add(vb, v_add(v_cross(N1b, u), v_cross(w, N2b)));
  const P3 wb = v_cross(N2b, v);
  bar[0] = v_add(bar[0], ub);
  bar[1] = v_add(bar[1], vb);
  bar[3] = v_add(bar[3], wb);
  bar[2] = v_sub(bar[2], v_add(ub, v_add(vb, wb)));
}
// reverse of tri_bend_2(r, a, b, c)
__device__ void tri_bend_2_bar(const Tis& t, P3 rv, P3 avp, P3 bv, P3 cv, P3 ob, P3* bar) {
  const P3 p = vpsub(t, rv, cv), q = vpsub(t, bv, cv);
  const P3 u = vpsub(t, rv, bv), v = vpsub(t, avp, bv);
  P3 N1 = v_cross(p, q), N2 = v_cross(u, v);
  P3 t1 = N1, t2 = N2;
  const P3 n1 = v_normalize(t1), n2 = v_normalize(t2);
  const double div1 = v_norm(t1), div2 = v_norm(t2);
  P3 qb{0, 0, 0}, vb{0, 0, 0}, n1b{0, 0, 0}, n2b{0, 0, 0};
  double db = 0;
  bend_pair_bar(n1, n2, q, 1.0 / div1, ob, qb, n1b, n2b, db);
  bend_pair_bar(n2, n1, v, 1.0 / div2, ob, vb, n2b, n1b, db);
  n1b = v_add(n1b, v_scale(db, n2));
  n2b = v_add(n2b, v_scale(db, n1));
  const P3 N1b = unit_bar(N1, n1b), N2b = unit_bar(N2, n2b);
  const P3 pb = v_cross(q, N1b);
  qb = v_add(qb, v_cross(N1b, p));
  const P3 ub = v_cross(v, N2b);
  vb = v_add(vb, v_cross(N2b, u));
  bar[0] = v_add(bar[0], v_add(pb, ub));
  bar[1] = v_add(bar[1], vb);
  bar[2] = v_add(bar[2], v_sub(qb, v_add(ub, vb)));
  bar[3] = v_sub(bar[3], v_add(pb, qb));
}
}  // namespace

// positions between the tissue's current buffer and the window's storage:
// dir 0 save to slot, 1 restore from slot, 2 save checkpoint, 3 restore
// checkpoint, 4 restore the snapshot (both buffers), 5 take the snapshot
__global__ void __launch_bounds__(256) k_mech_state(const Tis* __restrict__ T, const MechTis* __restrict__ M,
                                                     int slot, int dir, int pb) {
  const Tis& t = T[blockIdx.y];
  const MechTis& m = M[blockIdx.y];
  const int n = t.st->n;
  const long long cap = t.cap;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    for (int a = 0; a < 3; a++) {
      if (dir <= 3) {
        double* st = (dir < 2 ? m.slot : m.ck) + (long long)slot * 3 * cap + a * cap;
        if (dir == 0 || dir == 2) st[c] = t.pos[pb][a][c];
        else t.pos[pb][a][c] = st[c];
      } else if (dir == 4) {
        t.pos[pb][a][c] = m.x0[0][a * cap + c];
        t.pos[1 - pb][a][c] = m.x0[1][a * cap + c];
      } else {
        m.x0[0][a * cap + c] = t.pos[pb][a][c];
        m.x0[1][a * cap + c] = t.pos[1 - pb][a][c];
      }
    }
  }
}

// loss = mean over cells of |x - target|^2 (minimum image in x, y) and the seed
__global__ void __launch_bounds__(256) k_mech_seed(const Tis* __restrict__ T, MechTis* __restrict__ M, int pb,
                                                    int q) {
  const Tis& t = T[blockIdx.y];
  MechTis& m = M[blockIdx.y];
  const int n = t.st->n;
  const long long cap = t.cap;
  double part = 0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const double dx = p_diff_x(t, t.pos[pb][0][c], m.target[c]);
    const double dy = p_diff_y(t, t.pos[pb][1][c], m.target[cap + c]);
    const double dz = t.pos[pb][2][c] - m.target[2 * cap + c];
    part += dx * dx + dy * dy + dz * dz;
    m.lam[q][c] = 2 * dx / n;
    m.lam[q][cap + c] = 2 * dy / n;
    m.lam[q][2 * cap + c] = 2 * dz / n;
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
  if ((threadIdx.x & 63) == 0 && part != 0) atomicAdd(&m.scal[0], part / n);
}

// restore the state of slot `slot` and start the next adjoint at the identity
__global__ void __launch_bounds__(256) k_mech_bwd_begin(const Tis* __restrict__ T, MechTis* __restrict__ M, int slot,
                                                         int pb, int q) {
  const Tis& t = T[blockIdx.y];
  MechTis& m = M[blockIdx.y];
  const int n = t.st->n;
  const long long cap = t.cap;
  const double* st = m.slot + (long long)slot * 3 * cap;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    for (int a = 0; a < 3; a++) {
      t.pos[pb][a][c] = st[a * cap + c];
      m.lam[1 - q][a * cap + c] = m.lam[q][a * cap + c];
    }
}

// The reverse kernels take one lane per cell.  A cell's own terms are summed
// by its lane and added with one atomic; the terms of other cells are
// scattered with atomics and the parameter sums are reduced at the end, so
// no order is kept (a gradient need not follow the forward's).  (Lane groups
// per cell measured slower: the laws' branches and the dihedral terms' two
// code paths then run one after the other within a group.)

// every interaction a cell's handler applies (cell_interaction.cpp:529-641)
__global__ void __launch_bounds__(256) k_mech_bwd_pairs(const Tis* __restrict__ T, MechTis* __restrict__ M, int pb,
                                                         int cb, int q) {
  const Tis& t = T[blockIdx.y];
  MechTis& m = M[blockIdx.y];
  const int n = t.st->n;
  const long long cap = t.cap;
  const double* lam = m.lam[q];
  double* nl = m.lam[1 - q];
  ParAcc pa{0, 0, 0};
  for (int X = blockIdx.x * blockDim.x + threadIdx.x; X < n; X += gridDim.x * blockDim.x) {
    const CV cx = load_cv(t, pb, X);
    P3 gx{0, 0, 0};
    if (X < t.nmemb) {
      for (int i = 0; i < MEMB_ADJ; i++) {
        const int o = t.memb6[6 * X + i];
        if (t.state[o] == MEMB && X > o)
          gx = v_add(gx, pair_bwd(t, LAW_MEMB_MEMB, cx, X, load_cv(t, pb, o), o, lam, nl, cap, pa));
      }
    } else {
      const int* row = t.col[cb] + (size_t)X * t.ns;
      const int len = t.nlen[cb][X];
      for (int e = 0; e < len; e++) {
        const int o = row[e];
        const CV co = load_cv(t, pb, o);
        const int law = proc_law(cx, X, co, o);
        if (law != LAW_NONE) gx = v_add(gx, pair_bwd(t, law, cx, X, co, o, lam, nl, cap, pa));
      }
      if (cx.pair >= 0 && X > cx.pair)
        gx = v_add(gx, pair_bwd(t, LAW_PAIR, cx, X, load_cv(t, pb, cx.pair), cx.pair, lam, nl, cap, pa));
    }
    atomic_add3(nl, cap, X, gx);
  }
  for (int o = 32; o > 0; o >>= 1) {
    pa.eps += __shfl_down(pa.eps, o, 64);
    pa.adh += __shfl_down(pa.adh, o, 64);
    pa.spr += __shfl_down(pa.spr, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (pa.eps != 0) atomicAdd(&m.scal[1], pa.eps);
    if (pa.adh != 0) atomicAdd(&m.scal[2], pa.adh);
    if (pa.spr != 0) atomicAdd(&m.scal[3], pa.spr);
  }
}

// the wall (membrane and DER cells, :287-294) and the bending of every
// membrane particle (:488-499)
__global__ void __launch_bounds__(256) k_mech_bwd_memb(const Tis* __restrict__ T, MechTis* __restrict__ M, int pb,
                                                        int q) {
  const Tis& t = T[blockIdx.y];
  MechTis& m = M[blockIdx.y];
  const int n = t.st->n;
  const long long cap = t.cap;
  const double* lam = m.lam[q];
  double* nl = m.lam[1 - q];
  double peps = 0, pbend = 0;
  for (int X = blockIdx.x * blockDim.x + threadIdx.x; X < n; X += gridDim.x * blockDim.x) {
    const bool memb = X < t.nmemb;
    if (!memb && t.state[X] != DER) continue;
    const CV cx = load_cv(t, pb, X);
    if (cx.p.z < cx.r) {
      const double z = cx.p.z;
      const double L = POW6(cx.r) / POW6(z);
      atomicAdd(&nl[2 * cap + X], lam[2 * cap + X] * t.prm.DT_Cell * 2.0 * t.m_eps * (7.0 * L - 13.0 * L * L) / (z * z));
      peps += lam[2 * cap + X] * t.prm.DT_Cell * 2.0 * L * (L - 1.0) / z;
    }
    if (!memb) continue;
    const P3 lm = ld3(lam, cap, X);
    const P3 ob = v_scale(t.prm.DT_Cell * t.m_kbend, lm);
    const int* md = t.memb6 + 6 * X;
    P3 dr{0, 0, 0};
    for (int i = 0; i < 6; i++) {
      const int i1 = (i + 1) % 6;
      const int idx[4] = {X, md[i], md[i1], t.memb6[6 * md[i1] + i]};
      const P3 pr = ld_pos(t, pb, idx[1]), pbv = ld_pos(t, pb, idx[2]), pc = ld_pos(t, pb, idx[3]);
      dr = v_add(dr, tri_bend_1(t, cx.p, pr, pbv, pc));
      P3 bar[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      tri_bend_1_bar(t, cx.p, pr, pbv, pc, ob, bar);
      for (int k = 0; k < 4; k++) atomic_add3(nl, cap, idx[k], bar[k]);
    }
    for (int i = 0; i < 6; i++) {
      const int idx[4] = {X, md[i], md[(i + 1) % 6], md[(i + 2) % 6]};
      const P3 pa = ld_pos(t, pb, idx[1]), pbv = ld_pos(t, pb, idx[2]), pc = ld_pos(t, pb, idx[3]);
      dr = v_add(dr, tri_bend_2(t, cx.p, pa, pbv, pc));
      P3 bar[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      tri_bend_2_bar(t, cx.p, pa, pbv, pc, ob, bar);
      for (int k = 0; k < 4; k++) atomic_add3(nl, cap, idx[k], bar[k]);
    }
    pbend += t.prm.DT_Cell * v_dot(lm, dr);
  }
  for (int o = 32; o > 0; o >>= 1) {
    peps += __shfl_down(peps, o, 64);
    pbend += __shfl_down(pbend, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (peps != 0) atomicAdd(&m.scal[1], peps);
    if (pbend != 0) atomicAdd(&m.scal[4], pbend);
  }
}

void launch_mech_state(hipStream_t s, const Tis* dT, const MechTis* dM, int ntis, int slot, int dir, int pb,
                       int max_n) {
  EPI_LAUNCH(K_MECH_STATE, k_mech_state, dim3((max_n + 255) / 256, ntis), dim3(256), 0, s, dT, dM, slot, dir, pb);
}
void launch_mech_seed(hipStream_t s, const Tis* dT, MechTis* dM, int ntis, int pb, int q, int max_n) {
  EPI_LAUNCH(K_MECH_SEED, k_mech_seed, dim3((max_n + 255) / 256, ntis), dim3(256), 0, s, dT, dM, pb, q);
}
void launch_mech_backward_step(hipStream_t s, const Tis* dT, MechTis* dM, int ntis, int slot, int pb, int cb, int q,
                               int max_n) {
  const dim3 g((max_n + 255) / 256, ntis);
  EPI_LAUNCH(K_MECH_BWD_BEGIN, k_mech_bwd_begin, g, dim3(256), 0, s, dT, dM, slot, pb, q);
  EPI_LAUNCH(K_MECH_BWD_PAIRS, k_mech_bwd_pairs, g, dim3(256), 0, s, dT, dM, pb, cb, q);
  EPI_LAUNCH(K_MECH_BWD_MEMB, k_mech_bwd_memb, g, dim3(256), 0, s, dT, dM, pb, q);
}

}  // namespace epi
