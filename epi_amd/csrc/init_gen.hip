// init_gen.hip -- the initial-state generator of epi_intg (init_gen,
// epi_intg/calc/init_gen.cpp:23-106) over the runtime parameters: a
// membrane sheet, dermis layers and a ring of FIX cells, each FIX lowered
// onto the sheet.  Host code: the rows it writes are a deck for
// epi_load_cells (MEMB, then DER, then FIX).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/epi_amd.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

void set_err(char* err, int err_len, const std::string& msg) {
  if (err && err_len > 0) {
    std::strncpy(err, msg.c_str(), (size_t)err_len - 1);
    err[err_len - 1] = 0;
  }
}

// utils.cpp:8-32 over the parameters' box
double p_diff(double a, double b, double L) {
  const double diff = a - b;
  if (diff > 0.5 * L) return diff - L;
  if (diff <= -0.5 * L) return diff + L;
  return diff;
}

}  // namespace

extern "C" int epi_sheet_size(const epi_params* p, int32_t* nmx, int32_t* nmy) {
  if (!p || !nmx || !nmy || !(p->COMPRESS_FACTOR > 0) || !(p->R_memb > 0)) return EPI_ERR_INVALID;
  // CalcParams::calculate_param (CalcParams.cpp:20-26)
  const double P_MEMB = 1.0 / p->COMPRESS_FACTOR;
  const double pratio = 2.0 * P_MEMB;
  unsigned x = (unsigned)(p->LX / (p->R_memb * pratio));
  unsigned y = (unsigned)(p->LY / (p->R_memb * pratio));
  if (p->USE_TRI_MEMB) {
    const double y_tri_comp_ratio = 2.0 / std::sqrt(3.0);   // CalcParams.cpp:277
    y = (uint32_t)(y * y_tri_comp_ratio) & ~1u;               // evenify (:13-15)
  }
  *nmx = (int32_t)x;
  *nmy = (int32_t)y;
  return EPI_OK;
}

extern "C" int epi_init_gen(const epi_params* p, int32_t nfix, int32_t der, int32_t cap, epi_gen_cells* out,
                            char* err, int32_t err_len) {
  if (!p || !out || nfix < 1 || der < 0) {
    set_err(err, err_len, "invalid arguments");
    return EPI_ERR_INVALID;
  }
  int32_t nmx = 0, nmy = 0;
  if (epi_sheet_size(p, &nmx, &nmy) != EPI_OK || !(p->R_der > 0)) {
    set_err(err, err_len, "invalid parameters");
    return EPI_ERR_INVALID;
  }
  const double pratio = 2.0 * (1.0 / p->COMPRESS_FACTOR);
  const double Z_LEVEL = der * 2.0 * p->R_der;
  const unsigned NX_DER = (unsigned)(p->LX / (2.0 * p->R_der));
  const unsigned NY_DER = (unsigned)(p->LY / (2.0 * p->R_der));
  const long long nmemb = (long long)nmx * nmy, nder = (long long)der * NY_DER * NX_DER;
  const long long n = nmemb + nder + nfix;
  out->nmx = nmx;
  out->nmy = nmy;
  out->nmemb = (int32_t)nmemb;
  out->nder = (int32_t)nder;
  out->n = n > INT32_MAX ? INT32_MAX : (int32_t)n;
  if (n > cap) {
    set_err(err, err_len, "capacity " + std::to_string(cap) + " < " + std::to_string(n) + " cells");
    return EPI_ERR_CAPACITY;
  }
  if (!out->state || !out->fix_origin || !out->x || !out->y || !out->z || !out->radius) {
    set_err(err, err_len, "null output array");
    return EPI_ERR_INVALID;
  }
  long long i = 0;
  auto put = [&](uint8_t s, int fo, double x, double y, double z, double r) {
    out->state[i] = s; out->fix_origin[i] = fo;
    out->x[i] = x; out->y[i] = y; out->z[i] = z; out->radius[i] = r;
    i++;
  };
  const double Rm = p->R_memb;
  // the sheet (:59-77)
  if (!p->USE_TRI_MEMB) {
    for (int k = 0; k < nmy; k++)
      for (int j = 0; j < nmx; j++)
        put(EPI_MEMB, 0, 0.5 * Rm + (unsigned)j * Rm * pratio, 0.5 * Rm + (unsigned)k * Rm * pratio, Z_LEVEL + Rm, Rm);
  } else {
    const double y_tri_comp_ratio = 2.0 / std::sqrt(3.0);
    for (int k = 0; k < nmy; k++)
      for (int j = 0; j < nmx; j++)
        put(EPI_MEMB, 0, 0.5 * Rm + (unsigned)j * Rm * pratio + (k % 2 == 0 ? 0.0 : Rm * pratio * 0.5),
            0.5 * Rm + (unsigned)k * Rm * pratio / y_tri_comp_ratio, Z_LEVEL + Rm, Rm);
  }
  // dermis layers (:80-91)
  const double Rd = p->R_der;
  for (int l = 0; l < der; l++)
    for (unsigned k = 0; k < NY_DER; k++)
      for (unsigned j = 0; j < NX_DER; j++) put(EPI_DER, 0, Rd + 2.0 * j * Rd, Rd + 2.0 * k * Rd, Rd + 2.0 * l * Rd, Rd);
  // the FIX ring (:92-97)
  for (int l = 0; l < nfix; l++)
    put(EPI_FIX, l, p->LX / 2.0 + (p->LX / 4.0) * std::cos(2.0 * kPi * l / nfix),
        p->LY / 2.0 + (p->LY / 4.0) * std::sin(2.0 * kPi * l / nfix), p->R_max + Rm + Z_LEVEL + Rm, p->R_max);
  // relaxation (:23-46): each FIX goes down by 0.001 until a membrane cell is
  // within 1.1 (radius + R_der); the nearest distance is kept over the
  // descent, as the reference keeps distmin
  const long long NB = n - nfix;
  for (long long j = NB; j < n; j++) {
    double distmin = 10000.0;
    for (;;) {
      for (long long l = 0; l < NB; l++) {
        if (out->state[l] != EPI_MEMB) break;
        const double dx = p_diff(out->x[j], out->x[l], p->LX), dy = p_diff(out->y[j], out->y[l], p->LY);
        const double dz = out->z[j] - out->z[l];
        const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (d < distmin) distmin = d;
      }
      if (distmin < 1.1 * (out->radius[j] + p->R_der)) break;
      if (out->z[j] < Z_LEVEL - p->R_max) {
        set_err(err, err_len, "zz too low: zz=" + std::to_string(out->z[j]) +
                                  " zz-zlevel=" + std::to_string(out->z[j] - Z_LEVEL));
        return EPI_ERR_INVALID;
      }
      out->z[j] = out->z[j] - 0.001;
    }
  }
  set_err(err, err_len, "");
  return EPI_OK;
}
