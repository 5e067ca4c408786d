// noalloc.hip -- calc_ext_stim (ext_stim.cpp:18-69) over caller-held device
// grids, allocating nothing: epi_calc_ext_stim_noalloc (include/epi_amd.h).
//
// The context path keeps its grid descriptors (Tis), per-tissue status and
// the B stencil's tile work list in memory the library allocates.  Here they
// live in one static device array, `scratch`, of a fixed size per GPU: a
// first kernel writes one descriptor and status per caller view into it,
// then the context path's own kernels (k_wl_b, k_ext_stim) run on them, and
// a last kernel copies each grid's status out.  The array's device address
// is looked up once per device (hipGetSymbolAddress) and that address, not
// the symbol, is what host code hands to kernels.
#include <mutex>

#include "epi_dev.h"
#include "launch.h"
#include "../../include/epi_amd.h"

namespace epi {
namespace {

constexpr size_t al256(size_t v) { return (v + 255) / 256 * 256; }
constexpr int NA_MAX = EPI_NOALLOC_MAX_GRIDS;
constexpr size_t NA_TIS = al256(sizeof(Tis) * NA_MAX);         // descriptors
constexpr size_t NA_ST = al256(sizeof(Status) * NA_MAX);       // per-grid status
constexpr size_t NA_WL = al256(sizeof(int) * (2 * NA_MAX + 1));  // tile list: offsets, iz_bound
constexpr size_t NA_BYTES = NA_TIS + NA_ST + NA_WL;

}  // namespace
}  // namespace epi

// Scratch of epi_calc_ext_stim_noalloc: every call on a GPU shares that GPU's
// copy, which keeps the last call's contents; each call rewrites what it
// reads (descriptors, statuses, list) before reading it.
static __device__ __align__(256) unsigned char scratch[epi::NA_BYTES];

namespace epi {
namespace {

constexpr int NA_DEVICES = 64;
std::mutex g_na_mu;
unsigned char* g_na_ptr[NA_DEVICES] = {};

// the current device's address of `scratch`, looked up on the first call per
// device with that device current
int na_scratch(unsigned char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_DEVICES) return EPI_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return EPI_ERR_HIP;
    g_na_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_na_ptr[dev];
  return EPI_OK;
}

// one descriptor and status per view: the fields k_wl_b and k_ext_stim read
// (the same expressions as epi_create's for dz and inv_dz), with epi5_MT's
// parameters
__global__ void __launch_bounds__(256) k_na_setup(const epi_bfield_view* __restrict__ views, int n,
                                                  Tis* __restrict__ T, Status* __restrict__ S, const Params prm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const epi_bfield_view v = views[i];
  Status s = {};
  s.zzmax_bits = (unsigned long long)__double_as_longlong(v.zzmax);
  s.zzmax = v.zzmax;
  S[i] = s;
  Tis t = {};
  t.NX = v.nx; t.NY = v.ny; t.NZ = v.nz; t.XP = v.xp;
  t.PL = v.pl;
  t.gsize = v.pl * (v.nz + 1);
  t.LZ = v.lz;
  t.dz = v.lz / v.nz;
  t.inv_dz = v.nz / v.lz;
  t.st = S + i;
  t.bf[0] = const_cast<double*>(v.b_in);
  t.bf[1] = v.b_out;
  t.cmap1 = const_cast<int*>(v.cmap1);
  t.cmap2 = const_cast<uint8_t*>(v.cmap2);
  t.state = const_cast<uint8_t*>(v.state);
  t.agek = const_cast<double*>(v.agek);
  t.prm = prm;
  T[i] = t;
}

__global__ void __launch_bounds__(256) k_na_finish(const Status* __restrict__ S, int n, int32_t* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) err[i] = S[i].err;
}

}  // namespace
}  // namespace epi

extern "C" int64_t epi_noalloc_scratch_bytes(void) { return (int64_t)epi::NA_BYTES; }

extern "C" int epi_calc_ext_stim_noalloc(const epi_bfield_view* views, int32_t ngrids, int32_t* err, void* stream) {
  using namespace epi;
  if (!views || !err || ngrids < 1 || ngrids > NA_MAX) return EPI_ERR_INVALID;
  unsigned char* base = nullptr;
  const int rc = na_scratch(&base);
  if (rc != EPI_OK) return rc;
  Tis* T = reinterpret_cast<Tis*>(base);
  Status* S = reinterpret_cast<Status*>(base + NA_TIS);
  int* wl = reinterpret_cast<int*>(base + NA_TIS + NA_ST);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned g = (unsigned)((ngrids + 255) / 256);
  static const Params prm = [] {
    epi_params d;
    epi_params_default(&d);
    Params p;
    derive_params(d, p);
    return p;
  }();
  hipLaunchKernelGGL(k_na_setup, dim3(g), dim3(256), 0, s, views, ngrids, T, S, prm);
  // the grid sizes are in device memory: the blocks walk the list grid-stride
  // from a grid sized for 256 tiles per view
  launch_ext_stim(s, T, ngrids, 0, 256LL * ngrids, wl, Tuning{});
  hipLaunchKernelGGL(k_na_finish, dim3(g), dim3(256), 0, s, S, ngrids, err);
  return hipGetLastError() == hipSuccess ? EPI_OK : EPI_ERR_HIP;
}
