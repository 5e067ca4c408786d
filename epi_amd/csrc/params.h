// params.h -- the runtime parameters on the device: the caller's record
// (epi_params) with the derived quantities of CalcParams::calculate_param
// (epi_intg/calc/CalcParams.cpp:16-45).  Every tissue descriptor holds one
// (Tis::prm), and the kernels read every physics constant through it.
#pragma once
#include "../../include/epi_amd.h"

namespace epi {

struct Params : epi_params {
  double P_MEMB;              // 1 / COMPRESS_FACTOR
  double agki_max_fix;        // fac * agki_max
  double leak_from_storage;   // CA_OUT * beta_zero
  double K_DESMOSOME;         // K_TOTAL * K_DESMOSOME_RATIO
  double delta_R;             // delta_R_coef * R_der
  double delta_L;             // delta_L_coef * R_max
  double iz_margin;           // FAC_MAP * R_max (ca2p.cpp:386, ext_stim.cpp:22)
  int Ca_ITR;                 // (int)(Ca_avg_time / DT_Ca)
  int pad_p;
};

void derive_params(const epi_params& in, Params& P);
// why `in` cannot be a context's parameters (a key that fixes the build
// differs from `ref`, or a value is out of range), or nullptr
const char* params_refusal(const epi_params& in, const epi_params& ref);

}  // namespace epi
