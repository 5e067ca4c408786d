// k_mech.hip -- cell_interaction (cell_interaction.cpp:648-685) on CDNA4.
//
// The reference scatters every pair force into both cells with CAS atomics
// (cell_interaction.cpp:48-66, atomics.h:42-47), so its result depends on the
// order of the additions.  Here every cell gathers its own contributions and
// adds them in the order of the serial loop: bending, wall, its own handler
// over its list, lower-index handlers before and higher-index handlers after
// (cell_interaction.cpp:650-679).  One thread owns one cell: no atomics, the
// FP64 result is the serial one bit for bit.
#include "epi_dev.h"
#include "launch.h"
#include "epi_group.h"

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
  return 4.0 * eps_m * LJ6 * (LJ6 - 1) / distSq;
}
__device__ __forceinline__ bool is_near(const Tis& t, const CV& a, const CV& b) {
  return dist_sq(t, a.p, b.p) < (a.r + b.r) * (a.r + b.r);
}
__device__ __forceinline__ double adhesion(const Tis& t, const CV& a, const CV& b, double spring_const) {
  const double distlj = sqrt(dist_sq(t, a.p, b.p));
  const double rad_sum = a.r + b.r;
  const double LJ2_m1 = (distlj / rad_sum) - 1;
  return (LJ2_m1 + 1 > LJ_THRESH ? 0.0
          : -(spring_const / distlj) * LJ2_m1 *
                (1 - LJ2_m1 * LJ2_m1 / ((LJ_THRESH - 1.0) * (LJ_THRESH - 1.0))));
}

// Coefficient of law `law` with c1 = a, c2 = b (cell_interaction.cpp:77-285).
__device__ double coef_law(const Tis& t, int law, const CV& a, const CV& b, Status* st) {
  switch (law) {
    case LAW_OTHER:
      return ljmain(t, a, b);
    case LAW_DER_DER: {
      const double dist = sqrt(dist_sq(t, a.p, b.p));
      if (dist + delta_R < a.r + b.r) {
        const double dist2 = dist + delta_R;
        const double rad_sum = a.r + b.r;
        const double LJ6 = POW6(rad_sum) / POW6(dist2);
        return 4.0 * eps_m * LJ6 * (LJ6 - 1) / (dist * dist2) + para_ljp2;
      } else if (is_near(t, a, b)) {
        return para_ljp2;
      } else {
        const double dist2 = dist + delta_R;
        const double rad_sum = a.r + b.r;
        const double LJ6 = POW6(rad_sum) / POW6(dist2);
        return 4.0 * eps_m * LJ6 * (LJ6 - 1) / (dist * dist) + para_ljp2;
      }
    }
    case LAW_ALAIRDE:
      if (is_near(t, a, b)) return ljmain(t, a, b);
      else {
        double spf = K_DESMOSOME;
        if (a.agek > THRESH_SP && b.agek > THRESH_SP) spf = K_TOTAL;
        return adhesion(t, a, b, spf);
      }
    case LAW_FIXMU:
      if (is_near(t, a, b)) return ljmain(t, a, b);
      return adhesion(t, a, b, K_DESMOSOME);
    case LAW_MU_MEMB: {
      if (is_near(t, a, b)) return ljmain(t, a, b);
      // spring_force_to_memb, set at the top of _MUSUME_interaction (:613-620)
      double spring = 0;
      if (a.pair >= 0 && t.state[a.pair] == FIX) spring = Kspring;
      else if (a.rdt > 0) spring = Kspring_d;
      return adhesion(t, a, b, spring);
    }
    case LAW_FIX_MEMB:
      if (is_near(t, a, b)) return ljmain(t, a, b);
      else {
        const double distlj = sqrt(dist_sq(t, a.p, b.p));
        const double LJ2 = distlj / (a.r + b.r);
        return -(Kspring / distlj) * (LJ2 - 1.0);
      }
    case LAW_MEMB_MEMB: {
      const double rad_sum = a.r + b.r;
      const double rad_sum_sq = rad_sum * rad_sum;
      double cr_dist_sq;
      const double distSq = dist_sq(t, a.p, b.p);
      double tmp;
      if (distSq < (cr_dist_sq = rad_sum_sq * P_MEMB * P_MEMB)) {
        const double LJ6 = POW3(cr_dist_sq) / POW3(distSq);
        tmp = 4.0 * eps_m * LJ6 * (LJ6 - 1) / distSq;
      } else if (distSq < rad_sum_sq) {
        const double distlj = sqrt(distSq);
        const double cr_dist = rad_sum * P_MEMB;
        tmp = -(DER_DER_CONST / distlj) * (distlj / cr_dist - 1.0);
      } else {
        const double distlj = sqrt(distSq);
        const double lambda_dist = (1.0 + P_MEMB) * rad_sum;
        const double cr_dist = rad_sum * P_MEMB;
        const double LJ6 = POW6(cr_dist) / POW6(lambda_dist - distlj);
        tmp = -(DER_DER_CONST / rad_sum) * ((1.0 - P_MEMB) / P_MEMB)
              - 4 * eps_m * (LJ6 * (LJ6 - 1.0)) / ((lambda_dist - distlj) * distlj);
      }
      if (fabs(tmp) > 100) set_err(st, ERR_STRONG_INTERACTION);
      return tmp;
    }
    case LAW_PAIR: {
      const double dist = sqrt(dist_sq(t, a.p, b.p));
      const double force = Kspring_division * (dist - a.snl);
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
  const double d = DT_Cell * coef;
  return P3{d * p_diff_x(t, a.p.x, b.p.x), d * p_diff_y(t, a.p.y, b.p.y), d * (a.p.z - b.p.z)};
}

// wall_interaction (:287-294): the z increment
__device__ __forceinline__ double wall_dz(const CV& c) {
  const double distlj = 2.0 * c.p.z;
  const double LJ6 = POW6(c.r) / POW6(c.p.z);
  const double ljm = 4.0 * eps_m * LJ6 * (LJ6 - 1.0) / (distlj * distlj);
  return DT_Cell * ljm * 2.0 * c.p.z;
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
    acc.x += DT_Cell * KBEND * dr.x;
    acc.y += DT_Cell * KBEND * dr.y;
    acc.z += DT_Cell * KBEND * dr.z;
    if (cx.p.z < cx.r) acc.z += wall_dz(cx);
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
    if (cx.s == DER && cx.p.z < cx.r) acc.z += wall_dz(cx);
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
constexpr int FG = 16;
constexpr int FCAP = 48;
constexpr int F_BLOCK = 128;
constexpr int F_GROUPS = F_BLOCK / FG;

__global__ void __launch_bounds__(F_BLOCK) k_forces(const Tis* __restrict__ T, int pb, int cb) {
  const Tis& t = T[blockIdx.y];
  Status* st = t.st;
  __shared__ P3 s_own[F_GROUPS][FCAP];
  __shared__ P3 s_oth[F_GROUPS][FCAP];
  __shared__ int s_idx[F_GROUPS][FCAP];
  __shared__ uint8_t s_flag[F_GROUPS][FCAP];
  __shared__ uint8_t s_perm[F_GROUPS][FCAP];
  __shared__ P3 s_bend[F_GROUPS][12];
  const int lane = threadIdx.x & (FG - 1);
  const int grp = threadIdx.x / FG;
  const int n = st->n;
  const int stride = gridDim.x * F_GROUPS;
  for (int X = blockIdx.x * F_GROUPS + grp; X < n; X += stride) {
    const int len = t.nlen[cb][X];
    if (len > FCAP) continue;   // k_forces_long
    const CV cx = load_cv(t, pb, X);
    const bool memb = X < t.nmemb;
    const int nA = memb ? 0 : t.nlo[X];
    const int* row = t.col[cb] + (size_t)X * t.ns;
    for (int e = lane; e < len; e += FG) {
      const int q = row[e];
      const CV cq = load_cv(t, pb, q);
      uint8_t f = 0;
      P3 own = P3{0, 0, 0}, oth = P3{0, 0, 0};
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
      s_own[grp][e] = own;
      s_oth[grp][e] = oth;
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
    if (lane == 0) {
      P3 acc = cx.p;
      if (memb) {
        P3 dr = P3{0, 0, 0};
        for (int i = 0; i < 12; i++) dr = v_add(dr, s_bend[grp][i]);
        acc.x += DT_Cell * KBEND * dr.x;
        acc.y += DT_Cell * KBEND * dr.y;
        acc.z += DT_Cell * KBEND * dr.z;
        if (cx.p.z < cx.r) acc.z += wall_dz(cx);
        int hi[6], nhi = 0;
        for (int e = 0; e < MEMB_ADJ; e++) {
          const uint8_t f = s_flag[grp][e];
          if (f & 1) { const P3 d = s_own[grp][e]; acc.x += d.x; acc.y += d.y; acc.z += d.z; }
          if (f & 2) {
            int k = nhi++;
            const int q = s_idx[grp][e];
            while (k > 0 && s_idx[grp][hi[k - 1]] > q) { hi[k] = hi[k - 1]; k--; }
            hi[k] = e;
          }
        }
        for (int i = 0; i < nhi; i++) {
          const P3 d = s_oth[grp][hi[i]];
          acc.x -= d.x; acc.y -= d.y; acc.z -= d.z;
        }
        for (int e = MEMB_ADJ; e < len; e++)
          if (s_flag[grp][e] & 2) { const P3 d = s_oth[grp][e]; acc.x -= d.x; acc.y -= d.y; acc.z -= d.z; }
      } else {
        for (int r = 0; r < nA; r++) {
          const int e = s_perm[grp][r];
          if (s_flag[grp][e] & 2) { const P3 d = s_oth[grp][e]; acc.x -= d.x; acc.y -= d.y; acc.z -= d.z; }
        }
        if (cx.s == DER && cx.p.z < cx.r) acc.z += wall_dz(cx);
        for (int e = 0; e < len; e++)
          if (s_flag[grp][e] & 1) { const P3 d = s_own[grp][e]; acc.x += d.x; acc.y += d.y; acc.z += d.z; }
        if (cx.pair >= 0 && X > cx.pair) {
          const P3 d = delta(t, LAW_PAIR, cx, load_cv(t, pb, cx.pair), st);
          acc.x += d.x; acc.y += d.y; acc.z += d.z;
        }
        int P = cx.pair > X ? cx.pair : -1;
        for (int e = nA; e < len; e++) {
          const int p = s_idx[grp][e];
          if (P >= 0 && P < p) {
            const P3 d = delta(t, LAW_PAIR, load_cv(t, pb, P), cx, st);
            acc.x -= d.x; acc.y -= d.y; acc.z -= d.z;
            P = -1;
          }
          if (s_flag[grp][e] & 2) { const P3 d = s_oth[grp][e]; acc.x -= d.x; acc.y -= d.y; acc.z -= d.z; }
          if (p == P) {
            const P3 d = delta(t, LAW_PAIR, load_cv(t, pb, P), cx, st);
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
    wave_sync();
  }
}

// cells whose rows exceed the LDS slots of k_forces, one lane per cell
__global__ void __launch_bounds__(256) k_forces_long(const Tis* __restrict__ T, int pb, int cb) {
  const Tis& t = T[blockIdx.y];
  const int n = t.st->n;
  for (int X = blockIdx.x * blockDim.x + threadIdx.x; X < n; X += gridDim.x * blockDim.x)
    if (t.nlen[cb][X] > FCAP) forces_serial(t, pb, cb, X, t.st);
}

void launch_forces(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int max_memb, int max_other) {
  (void)max_memb;
  dim3 g((max_other + F_GROUPS - 1) / F_GROUPS, ntis);
  EPI_LAUNCH(K_FORCES, k_forces, g, dim3(F_BLOCK), 0, s, dT, pb, cb);
  dim3 gl((max_other + 255) / 256, ntis);
  EPI_LAUNCH(K_FORCES_LONG, k_forces_long, gl, dim3(256), 0, s, dT, pb, cb);
}

}  // namespace epi
