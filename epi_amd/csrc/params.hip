// params.hip -- the runtime parameter record (epi_params in include/epi_amd.h):
// epi5_MT's values, the ppsss key=value reader and writer of epi_intg's
// Params (epi_intg/Params.cpp:91-110, epi_intg/parser.cpp:30-58), and the
// derived quantities of CalcParams::calculate_param
// (epi_intg/calc/CalcParams.cpp:16-45) that the kernels read.
#include <cerrno>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>

#include "params.h"

namespace epi {
namespace {

enum Kind { F64, U32, BOOL, STR };
struct Key {
  const char* name;
  const char* alias;   // a second name the reader accepts (gpa in CalcParams.cpp:53-136)
  Kind kind;
  size_t off;
};

#define KD(n) {#n, nullptr, F64, offsetof(epi_params, n)}
#define KU(n) {#n, nullptr, U32, offsetof(epi_params, n)}
#define KB(n) {#n, nullptr, BOOL, offsetof(epi_params, n)}
#define KA(n, a, k) {#n, a, k, offsetof(epi_params, n)}

// the keys in the order of CalcParams::init's piset (CalcParams.cpp:51-162),
// which is the order Params::generate_paramfile writes
const Key kKeys[] = {
    KU(MAX_CONNECT_CELL_NUM), KD(DT_Cell), KD(DT_Ca), KA(MALIG_NUM, "MALIGNANT", U32), KD(LX), KD(LY), KD(LZ),
    KU(NX), KU(NY), KU(NZ), KD(R_max), KD(R_der), KD(R_memb), KD(COMPRESS_FACTOR), KU(MEMB_ADHE_RANGE),
    KD(THRESH_SP), KD(LJ_THRESH), KD(THRESH_DEAD), KA(NUM_ITR, "LOOP", U32),
    KA(Ca_avg_time, "Ca2P_avg_time", F64), KD(ca2p_init), KD(ex_inert_init), KD(IP3_init), KD(gj_init),
    KD(ATP_init), KD(ext_stim_init), KD(FAC_MAP), KU(SW_THRESH), KA(T_TURNOVER, "TURNOVER", F64),
    KU(NUM_SC_INIT), KB(USE_TRI_MEMB), KD(agki_max), KD(fac), KU(div_max), KD(Kpp), KD(dp), KD(kbc), KD(Hb),
    KD(Cout), KD(kg), KD(gamma), KD(mu0), KD(mu1), KD(kmu), KD(para_b), KD(para_bb), KD(para_k1), KD(k_flux),
    KD(beta_zero), KD(CA_OUT), KD(thgra), KD(delta_th), KD(thpri), KD(kpa), KD(delta_K), KD(Kpri), KD(delta_I),
    KD(iage_kitei), KD(para_k2), KD(H0), KD(wd), KD(ca2p_du), KD(STIM11), KD(Kaa), KD(Da), KD(AIR_STIM),
    KD(AREA_GRID_ORIGINAL), KU(N2), KU(N3), KD(eps_m), KD(DER_DER_CONST), KD(K_TOTAL), KD(K_DESMOSOME_RATIO),
    KD(Kspring), KD(Kspring_d), KD(KBEND), KD(delta_R_coef), KD(para_ljp2), KD(Kspring_division),
    KD(stoch_div_time_ratio), KD(S2), KD(accel_div), KD(eps_kb), KD(alpha_b), KD(S0), KD(eps_ks), KD(accel_diff),
    KD(alpha_k), KD(eps_kk), KD(S1), KD(ubar), KD(delta_lipid), KD(delta_sig_r1), KD(lipid_rel), KD(lipid),
    KD(delta_L_coef), KD(ADHE_CONST), KD(DISA_conn_num_thresh), KD(eps_L), KD(unpair_dist_coef), KD(kb),
    KD(DUR_ALIVE), KD(DUR_DEAD), KD(DB), KU(SYSTEM), KU(CUT), KB(NEW_BEND_POT), KB(STOCHASTIC),
    KB(force_cornification), {"outputdir", nullptr, STR, offsetof(epi_params, outputdir)},
};
#undef KD
#undef KU
#undef KB
#undef KA

// remove_LR_space (parser.cpp:27-29): no leading or trailing spaces, runs of
// spaces inside become one
std::string trim(const std::string& s) {
  std::string out;
  size_t i = 0, n = s.size();
  while (i < n && s[i] == ' ') i++;
  while (n > i && s[n - 1] == ' ') n--;
  for (; i < n; i++) {
    if (s[i] == ' ' && !out.empty() && out.back() == ' ') continue;
    out += s[i];
  }
  return out;
}

void set_err(char* err, int err_len, const std::string& msg) {
  if (err && err_len > 0) {
    std::strncpy(err, msg.c_str(), (size_t)err_len - 1);
    err[err_len - 1] = 0;
  }
}

// str_cast (Params.cpp:16-58): numbers as std::stod / stoi / stoul read
// them (a leading number, the rest ignored), a bool is 1 when the text reads
// as the integer 1, strings in double or single quotes.  An unsigned key
// refuses a negative value (stoul would wrap it).
bool read_value(const Key& k, const std::string& v, epi_params* p) {
  char* base = reinterpret_cast<char*>(p) + k.off;
  if (k.kind == STR) {
    if (v.size() < 2 || (v.front() != '"' && v.front() != '\'') || v.back() != v.front()) return false;
    const std::string s = v.substr(1, v.size() - 2);
    if (s.size() >= sizeof(p->outputdir)) return false;
    std::strcpy(base, s.c_str());
    return true;
  }
  const char* c = v.c_str();
  char* end = nullptr;
  errno = 0;
  if (k.kind == F64) {
    const double d = std::strtod(c, &end);
    if (end == c || errno == ERANGE) return false;
    std::memcpy(base, &d, sizeof d);
  } else {
    const long long l = std::strtoll(c, &end, 10);
    if (end == c || errno == ERANGE || l < INT32_MIN || l > INT32_MAX) return false;
    if (k.kind == U32 && l < 0) return false;
    const int32_t i = k.kind == BOOL ? (l == 1 ? 1 : 0) : (int32_t)l;
    std::memcpy(base, &i, sizeof i);
  }
  return true;
}

}  // namespace

// CalcParams::calculate_param (CalcParams.cpp:16-45) and the device record
void derive_params(const epi_params& in, Params& P) {
  static_cast<epi_params&>(P) = in;
  P.P_MEMB = 1.0 / in.COMPRESS_FACTOR;
  P.agki_max_fix = in.fac * in.agki_max;
  P.leak_from_storage = in.CA_OUT * in.beta_zero;
  P.K_DESMOSOME = in.K_TOTAL * in.K_DESMOSOME_RATIO;
  P.delta_R = in.delta_R_coef * in.R_der;
  P.delta_L = in.delta_L_coef * in.R_max;
  P.iz_margin = in.FAC_MAP * in.R_max;
  P.Ca_ITR = (int)(in.Ca_avg_time / in.DT_Ca);
  P.pad_p = 0;
}

const char* params_refusal(const epi_params& in, const epi_params& ref) {
  // the keys that fix the build: geometry, capacities, bucket edge, membrane
  // topology and the code paths this implementation has
  if (in.LX != ref.LX || in.LY != ref.LY || in.LZ != ref.LZ) return "LX, LY and LZ are the context's geometry";
  if (in.NX != ref.NX || in.NY != ref.NY || in.NZ != ref.NZ) return "NX, NY and NZ are the context's geometry";
  if (in.MAX_CONNECT_CELL_NUM != 400 || in.N2 != 400 || in.N3 != 200)
    return "MAX_CONNECT_CELL_NUM, N2 and N3 are fixed at 400, 400 and 200";
  if (in.AREA_GRID_ORIGINAL != 2.0) return "AREA_GRID_ORIGINAL (the bucket edge) is fixed at 2.0";
  if (in.USE_TRI_MEMB != 1 || in.NEW_BEND_POT != 1 || in.STOCHASTIC != 1 || in.SYSTEM != 0)
    return "USE_TRI_MEMB, NEW_BEND_POT and STOCHASTIC must be 1, SYSTEM 0";
  if (!(in.DT_Cell > 0) || !(in.DT_Ca > 0)) return "DT_Cell and DT_Ca must be positive";
  if (!(in.Ca_avg_time / in.DT_Ca >= 1.0) || !(in.Ca_avg_time / in.DT_Ca < 2147483647.0))
    return "Ca_avg_time / DT_Ca must be at least 1";
  return nullptr;
}

}  // namespace epi

using epi::kKeys;

extern "C" void epi_params_default(epi_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  // define.h:80-230
  p->MAX_CONNECT_CELL_NUM = 400;
  p->DT_Cell = 0.01;
  p->DT_Ca = 0.01;
  p->MALIG_NUM = 0;
  p->LX = 50.0; p->LY = 50.0; p->LZ = 100.0;
  p->NX = 100; p->NY = 100; p->NZ = 200;
  p->R_max = 1.4; p->R_der = 1.4; p->R_memb = 1.0;
  p->COMPRESS_FACTOR = 4;
  p->MEMB_ADHE_RANGE = 0;
  p->THRESH_SP = 3.0;
  p->LJ_THRESH = 1.2;
  p->THRESH_DEAD = 22.0;
  p->NUM_ITR = 4000000;
  p->Ca_avg_time = 10.0;
  p->ca2p_init = 0.122; p->ex_inert_init = 0.97; p->IP3_init = 0.0; p->gj_init = 0.99;
  p->ATP_init = 0.0; p->ext_stim_init = 0.0;
  p->FAC_MAP = 2.0;
  p->SW_THRESH = 20;
  p->T_TURNOVER = 6000.0;
  p->NUM_SC_INIT = 1;
  p->USE_TRI_MEMB = 1;
  // cell.h:23-32
  p->agki_max = 6.0; p->fac = 1.0; p->div_max = 15;
  // ca2p.cpp:16-144, 183, 286-304
  p->Kpp = 0.3; p->dp = 0.1;
  p->kbc = 0.4 * 1.2; p->Hb = 0.01; p->Cout = 1.0;
  p->kg = 0.1; p->gamma = 2.0;
  p->mu0 = 0.567; p->mu1 = 0.1; p->kmu = 0.05;
  p->para_b = 0.11; p->para_bb = 0.89; p->para_k1 = 0.7;
  p->k_flux = 8.1; p->beta_zero = 0.02; p->CA_OUT = 1.0;
  p->thgra = 0.2; p->delta_th = 1.0; p->thpri = 1.0;
  p->kpa = 4.0; p->delta_K = 1.0; p->Kpri = 6.0; p->delta_I = 1.5; p->iage_kitei = 0.0;
  p->para_k2 = 0.7; p->H0 = 0.5; p->wd = 0.1; p->ca2p_du = 0.01;
  p->STIM11 = 0.002; p->Kaa = 0.5; p->Da = 1.0; p->AIR_STIM = 0.1;
  // cell_connection.cpp:16-30
  p->AREA_GRID_ORIGINAL = 2.0; p->N2 = 400; p->N3 = 200;
  // cell_interaction.cpp:21-45
  p->eps_m = 0.01; p->DER_DER_CONST = 0.2; p->K_TOTAL = 3.0; p->K_DESMOSOME_RATIO = 0.01;
  p->Kspring = 25.0; p->Kspring_d = 5.0; p->KBEND = 0.5; p->delta_R_coef = 0.4;
  p->para_ljp2 = 0.005; p->Kspring_division = 5.0;
  // cell_state_renew.cpp:17-119, 231, 331-332, 366-367
  p->stoch_div_time_ratio = 0.25; p->S2 = 0.1; p->accel_div = 1.0; p->eps_kb = 0.12; p->alpha_b = 5.0;
  p->S0 = 0.1 * 0.2; p->eps_ks = 0.10 * 0.5; p->accel_diff = 1.0; p->alpha_k = 2.0; p->eps_kk = 0.10 * 0.5;
  p->S1 = 0.1; p->ubar = 0.45; p->delta_lipid = 0.1; p->delta_sig_r1 = 0.1; p->lipid_rel = 0.05 * 4.0;
  p->lipid = 0.6; p->delta_L_coef = 0.01; p->ADHE_CONST = 31.3; p->DISA_conn_num_thresh = 11;
  p->eps_L = 0.14; p->unpair_dist_coef = 0.9;
  // ext_stim.cpp:9-20
  p->kb = 0.025; p->DUR_ALIVE = 0.5; p->DUR_DEAD = 2.0; p->DB = 0.0009;
  // define.h:205-230
  p->SYSTEM = 0; p->CUT = 1000; p->NEW_BEND_POT = 1; p->STOCHASTIC = 1; p->force_cornification = 1;
  std::strcpy(p->outputdir, "output");
}

extern "C" int epi_params_load(const char* path, int32_t require_all, epi_params* p, char* err, int32_t err_len) {
  if (!path || !p) return EPI_ERR_INVALID;
  std::ifstream f(path);
  if (!f) {
    epi::set_err(err, err_len, std::string("Parse failed.") + path);
    return EPI_ERR_INVALID;
  }
  std::map<std::string, std::string> kv;
  std::string line;
  while (std::getline(f, line)) {
    if (line.rfind("#", 0) == 0 || line.rfind("//", 0) == 0) continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    kv.insert({epi::trim(line.substr(0, eq)), epi::trim(line.substr(eq + 1))});   // the first line of a key wins
  }
  epi_params q = *p;
  for (const auto& k : kKeys) {
    auto it = kv.find(k.name);
    if (it == kv.end() && k.alias) it = kv.find(k.alias);
    if (it == kv.end()) {
      if (require_all) {
        epi::set_err(err, err_len, std::string(k.name) + " must be specified.");
        return EPI_ERR_INVALID;
      }
      continue;
    }
    if (!epi::read_value(k, it->second, &q)) {
      epi::set_err(err, err_len, std::string("Failed to read the value of ") + k.name + "(='" + it->second + "')");
      return EPI_ERR_INVALID;
    }
  }
  *p = q;
  epi::set_err(err, err_len, "");
  return EPI_OK;
}

extern "C" int epi_params_save(const char* path, const epi_params* p) {
  if (!path || !p) return EPI_ERR_INVALID;
  FILE* f = std::fopen(path, "w");
  if (!f) return EPI_ERR_INVALID;
  for (const auto& k : kKeys) {
    const char* base = reinterpret_cast<const char*>(p) + k.off;
    if (k.kind == epi::F64) {
      double d;
      std::memcpy(&d, base, sizeof d);
      std::fprintf(f, "%s=%.17g\n", k.name, d);
    } else if (k.kind != epi::STR) {
      int32_t i;
      std::memcpy(&i, base, sizeof i);
      std::fprintf(f, "%s=%d\n", k.name, i);
    } else {
      std::fprintf(f, "%s=\"%s\"\n", k.name, base);
    }
  }
  return std::fclose(f) == 0 ? EPI_OK : EPI_ERR_INVALID;
}
