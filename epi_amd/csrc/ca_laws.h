// ca_laws.h -- the scalar laws of calc_ca2p (ca2p.cpp:14-144) shared by the
// forward kernels (k_field.hip) and the adjoint (k_adjoint.hip).
#pragma once
#include "epi_dev.h"

namespace epi {

constexpr double Kpp = 0.3;
// ca2p_du (ca2p.cpp:183) and dp (ca2p.cpp:17) are per-tissue values
// (Tis::ca_du, Tis::ca_dp), 0.01 and 0.1 unless a reaction-diffusion window
// sets them

__device__ __forceinline__ double ca2p_reaction_factor(double u, double ER3, double p, double B) {   // :36-87
  const double kbc = 0.4 * 1.2, Hb = 0.01, Cout = 1.0;
  const double kg = 0.1, gamma = 2.0;
  const double mu0 = 0.567, mu1 = 0.1, kmu = 0.05;
  const double para_b = 0.11, para_bb = 0.89, para_k1 = 0.7;
  const double k_flux = 8.1, beta_zero = 0.02, CA_OUT = 1.0;
  const double leak_from_storage = CA_OUT * beta_zero;
  const double er1 = (mu0 + mu1 * p / (p + kmu));
  const double er2 = (para_b + para_bb * u / (para_k1 + u));
  const double ca2p_from_storage = k_flux * er1 * er2 * ER3;
  const double into = gamma * u / (kg + u);
  const double by_ext = kbc * B * B * Cout / (Hb + B * B);
  return ca2p_from_storage - into + leak_from_storage + by_ext;
}

__device__ __forceinline__ double grid_avg8(const Tis& t, const double* g, int ix, int iy, int iz) {   // utils.h:18-23
  return 0.125 * (g[gidx(t, ix, iy, iz)] + g[gidx(t, ix + 1, iy, iz)] + g[gidx(t, ix, iy + 1, iz)]
                  + g[gidx(t, ix, iy, iz + 1)] + g[gidx(t, ix + 1, iy + 1, iz)] + g[gidx(t, ix + 1, iy, iz + 1)]
                  + g[gidx(t, ix, iy + 1, iz + 1)] + g[gidx(t, ix + 1, iy + 1, iz + 1)]);
}

__device__ __forceinline__ int ca_iz_bound(const Tis& t) {   // :386
  const double zzmax = __longlong_as_double((long long)t.st->zzmax_bits);
  return (int)((zzmax + FAC_MAP * R_max) / t.dz);
}

// fw (:138-144): (1 - w) + (-1 + tanh(y)) / 2 with y = (0.1 - diff) / 0.1.
// (-1 + tanh y) / 2 equals -1 / (1 + e^(2y)); that form costs one exp and one
// division (a double tanh costs about three times as much) and does not
// cancel where tanh y is near 1.  A large y makes e^(2y) infinite and the term
// -0, the limit.
__device__ __forceinline__ double fw_term(double diff, double w) {
  return (1. - w) - 1. / (1. + exp(2. * ((0.1 - diff) / 0.1)));
}

}  // namespace epi
