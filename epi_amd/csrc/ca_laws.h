// ca_laws.h -- the scalar laws of calc_ca2p (ca2p.cpp:14-144) shared by the
// forward kernels (k_field.hip) and the adjoint (k_adjoint.hip).
#pragma once
#include "epi_dev.h"

namespace epi {

// ca2p_du (ca2p.cpp:183) and dp (ca2p.cpp:17) are read from Tis::ca_du and
// Tis::ca_dp, the tissue's parameters unless a reaction-diffusion window sets
// them; every other constant of the laws comes from the tissue's record.
//
// A kernel copies the constants it uses out of the record once, before its
// first store: the compiler cannot tell a store to a grid or cell array from
// the descriptor, so a record field read after a store would be loaded from
// memory again, once per store.
struct CaK {   // the per-cell update of a sub-step (ca2p.cpp:36-144, 156-222)
  double DT_Ca, Kpp, H0, para_k2, wd;
  double mu0, mu1, kmu, para_b, para_bb, para_k1, k_flux, gamma, kg, kbc, Cout, Hb, leak_from_storage;
};
__device__ __forceinline__ CaK ca_k(const Params& pm) {
  return CaK{pm.DT_Ca, pm.Kpp, pm.H0, pm.para_k2, pm.wd, pm.mu0, pm.mu1, pm.kmu, pm.para_b, pm.para_bb, pm.para_k1,
             pm.k_flux, pm.gamma, pm.kg, pm.kbc, pm.Cout, pm.Hb, pm.leak_from_storage};
}
struct AtpK {   // ATP_refresh (ca2p.cpp:302-338)
  double DT_Ca, Da, STIM11, Kaa, AIR_STIM;
};
__device__ __forceinline__ AtpK atp_k(const Params& pm) {
  return AtpK{pm.DT_Ca, pm.Da, pm.STIM11, pm.Kaa, pm.AIR_STIM};
}

__device__ __forceinline__ double ca2p_reaction_factor(const CaK& k, double u, double ER3, double p,
                                                       double B) {   // :36-87
  const double er1 = (k.mu0 + k.mu1 * p / (p + k.kmu));
  const double er2 = (k.para_b + k.para_bb * u / (k.para_k1 + u));
  const double ca2p_from_storage = k.k_flux * er1 * er2 * ER3;
  const double into = k.gamma * u / (k.kg + u);
  const double by_ext = k.kbc * B * B * k.Cout / (k.Hb + B * B);
  return ca2p_from_storage - into + k.leak_from_storage + by_ext;
}

__device__ __forceinline__ double grid_avg8(const Tis& t, const double* g, int ix, int iy, int iz) {   // utils.h:18-23
  return 0.125 * (g[gidx(t, ix, iy, iz)] + g[gidx(t, ix + 1, iy, iz)] + g[gidx(t, ix, iy + 1, iz)]
                  + g[gidx(t, ix, iy, iz + 1)] + g[gidx(t, ix + 1, iy + 1, iz)] + g[gidx(t, ix + 1, iy, iz + 1)]
                  + g[gidx(t, ix, iy + 1, iz + 1)] + g[gidx(t, ix + 1, iy + 1, iz + 1)]);
}

__device__ __forceinline__ int ca_iz_bound(const Tis& t) {   // :386
  const double zzmax = __longlong_as_double((long long)t.st->zzmax_bits);
  return (int)((zzmax + t.prm.iz_margin) / t.dz);
}

// fw (:138-144): (1 - w) + (-1 + tanh(y)) / 2 with y = (wd - diff) / 0.1.
// (-1 + tanh y) / 2 equals -1 / (1 + e^(2y)); that form costs one exp and one
// division (a double tanh costs about three times as much) and does not
// cancel where tanh y is near 1.  A large y makes e^(2y) infinite and the term
// -0, the limit.
__device__ __forceinline__ double fw_term(double wd, double diff, double w) {
  return (1. - w) - 1. / (1. + exp(2. * ((wd - diff) / 0.1)));
}

}  // namespace epi
