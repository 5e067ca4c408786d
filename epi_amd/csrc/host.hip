// host.hip -- context, memory, the C ABI (include/epi_amd.h) and the step
// driver of proc.cpp:131-173.
//
// A context owns one device arena holding every tissue; per-tissue
// descriptors (Tis) live in device memory once, so each kernel launch covers
// all tissues (blockIdx.y).  Double buffers are selected by parities kept on
// the host and passed as kernel arguments; all tissues of a context step in
// lock-step, so the parities are shared.  One status read-back per step
// feeds the Ca gating (proc.cpp:158-172) and error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/epi_amd.h"
#include "epi_dev.h"
#include "launch.h"
#include "rd_dev.h"
#include "mech_dev.h"

using namespace epi;

// ---- per-kernel event timer ----------------------------------------------
namespace epi {
namespace {
struct ProfRec { int k; hipEvent_t a, b; bool counted; };
struct Prof {
  unsigned long long mask = 0;
  // sampling: with period > 1, runs of `run` consecutive launches (of any
  // kernel) carry events, one run starting every `period` launches; the first
  // launch of a run only gives the second a bracketed predecessor and is not
  // counted
  long long period = 1, run = 1, launches = 0;
  std::vector<ProfRec> pending;
  std::vector<hipEvent_t> pool;
  double ms[K_COUNT] = {};
  long long cnt[K_COUNT] = {};
  long long launched[K_COUNT] = {};   // every launch of an enabled kernel, timed or not
  bool capturing = false;             // launches captured into a graph carry no events
};
Prof g_prof;
hipEvent_t take_event() {
  if (!g_prof.pool.empty()) {
    hipEvent_t e = g_prof.pool.back();
    g_prof.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
}  // namespace
bool prof_on(int k, bool* counted) {
  if (!g_prof.mask || g_prof.capturing) return false;
  const bool want = (g_prof.mask >> k) & 1ull;
  if (want) g_prof.launched[k]++;
  if (g_prof.period <= 1) {
    *counted = want;
    return want;
  }
  const long long ph = g_prof.launches++ % g_prof.period;
  if (ph >= g_prof.run) return false;
  *counted = want && ph > 0;
  return true;
}
void prof_events(hipEvent_t* a, hipEvent_t* b) {
  *a = take_event();
  *b = take_event();
}
void prof_add(int k, hipEvent_t a, hipEvent_t b, bool counted) { g_prof.pending.push_back({k, a, b, counted}); }
}  // namespace epi

namespace {

const char* kKernelNames[K_COUNT] = {
    "k_forces", "k_forces_long", "k_renew_flags", "k_renew_scan", "k_renew_apply", "k_remove", "k_pair_disperse",
    "k_initialize_sc", "k_periodic_fix", "k_bucket_insert", "k_bucket_sort", "k_sweep",
    "k_finish", "k_snapshot", "k_zz_set", "k_zzmax", "k_map_clear", "k_map_stamp", "k_map_halo",
    "k_ext_stim", "k_ca_begin", "k_ca_vox", "k_ca_cells", "k_atp", "k_ca_end", "k_sw_reset",
    "k_rd_snapshot", "k_rd_restore", "k_rd_put_gj", "k_rd_state", "k_rd_seed",
    "k_rd_step", "k_rd_gj0", "k_rd_own", "k_mech_state", "k_mech_seed", "k_mech_bwd_begin",
    "k_mech_bwd_pairs", "k_mech_bwd_memb", "k_ca_fused", "k_ca_corners", "k_sweep_wave", "k_wl", "k_ca_swap",
    "k_out_pack"};

// with ready_only, only the pairs whose kernels have finished (the rest stay
// pending for a later read-back)
void prof_resolve(bool ready_only = false) {
  size_t keep = 0;
  for (size_t i = 0; i < g_prof.pending.size(); i++) {
    auto& r = g_prof.pending[i];
    if (ready_only && hipEventQuery(r.b) != hipSuccess) {
      g_prof.pending[keep++] = r;
      continue;
    }
    float ms = 0;
    if (r.counted && hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      g_prof.ms[r.k] += ms;
      g_prof.cnt[r.k]++;
    }
    g_prof.pool.push_back(r.a);
    g_prof.pool.push_back(r.b);
  }
  g_prof.pending.resize(keep);
}

struct TH {
  epi_geometry g{};
  int nmemb = 0, nder = 0, cap = 0, ns = 0, bcap = 0;
  int NX = 0, NY = 0, NZ = 0, XP = 0;
  long long PL = 0, gsize = 0;
  int nbuckets = 0;
  int n_est = 0;
  double* rng_d = nullptr;
  long long rng_n = 0;
  bool loaded = false;
};

}  // namespace

struct epi_ctx {
  int T = 0;
  std::vector<TH> th;
  std::vector<Tis> tis;
  Tis* d_tis = nullptr;
  Status* d_st = nullptr;
  Status* h_st = nullptr;
  double* h_zz = nullptr;
  int* d_tlist = nullptr;
  int* h_tlist = nullptr;
  // flattened work lists of the launches over a batch (offsets and per-tissue
  // iz_bound, written on the device): WL_INTS(T) ints
  int* d_wl = nullptr;
  // a Ca solve's list (CaList): built in pinned h_cal, uploaded to d_cal
  // (cal_ints ints) on the stream, ev_cal recorded after the upload; kept
  // while its key (tissues, cell counts, iz_bound, shapes) holds
  int* d_cal = nullptr;
  int* h_cal = nullptr;
  long long cal_ints = 0;
  hipEvent_t ev_cal = nullptr;
  std::vector<long long> cal_key;
  CaList cal{};
  char* arena = nullptr;
  size_t arena_bytes = 0;
  hipStream_t stream = nullptr;
  bool len_known = false;   // h_st maxlen read back since the last list build
  // a step builds the voxel maps on side_stream beside the connection build
  // (both only need the positions that bucket insertion fixes); the fork and
  // join events order it with the main stream
  hipStream_t side_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_st = nullptr;   // the status copy of a step, waited on alone
  int pb = 0, cb = 0, cp = 0, ap = 0, bp = 0;
  bool pos_pending = false;
  Tuning tu;
  // with STEP_GRAPH: a step's renewal and build, and whole Ca solves, as
  // instantiated graphs, one per launch key (parities, launch sizes, launch
  // shapes, the solve's list), least recently used evicted
  struct StepGraph {
    std::vector<long long> key;
    hipGraphExec_t exec = nullptr;
    unsigned long long used = 0;
  };
  std::vector<StepGraph> graphs;
  unsigned long long graph_tick = 0;
  // frozen reaction-diffusion window (epi_rd_*)
  struct Rd {
    bool active = false;
    std::vector<RdTis> h;
    RdTis* d = nullptr;
    char* arena = nullptr;      // snapshot and adjoint arrays
    char* work = nullptr;       // segment slots and checkpoints
    size_t work_bytes = 0;
    double* d_gj0 = nullptr;
    std::vector<int> izb;
    std::vector<double> zz;
  } rd;
  // frozen mechanics window (epi_mech_*)
  struct Mech {
    bool active = false;
    std::vector<MechTis> h;
    MechTis* d = nullptr;
    char* arena = nullptr;
    char* work = nullptr;
    size_t work_bytes = 0;
  } mech;
  // asynchronous snapshots (epi_snapshot_*): per slot the packed record on
  // the device, its pinned host copy (downloaded on copy_stream) and events
  struct Snap {
    char* dev = nullptr;
    char* host = nullptr;
    size_t bytes = 0;
    hipEvent_t ev_pack = nullptr, ev_done = nullptr;
    int tissue = -1, n = 0;
    long long m = 0;
    bool pending = false, landed = false;
  } snap[EPI_SNAPSHOT_SLOTS];
  hipStream_t copy_stream = nullptr;
  std::vector<long long> ca_solves;
  // each tissue's device timestep (Status::timestep) as the host last set or
  // advanced it
  std::vector<long long> dev_ts;
  // throughput accounting of epi_step
  std::vector<long long> steps, cell_steps, grid_ca;
  std::string err;
};

namespace {

int fail(epi_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                               \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(ctx, EPI_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// the coefficients a window may vary, back to the tissue's parameters
void window_values_from_params(Tis& t) {
  t.ca_du = t.prm.ca2p_du;
  t.ca_dp = t.prm.dp;
  t.m_eps = t.prm.eps_m;
  t.m_kadh = t.prm.K_TOTAL;
  t.m_kspring = t.prm.Kspring;
  t.m_kbend = t.prm.KBEND;
}

// proc.cpp:158: i * DT_Cell > T_TURNOVER && flg_forced_sc
bool sc_turnover(const Params& p, const epi_driver& d) {
  return (double)d.step * p.DT_Cell > p.T_TURNOVER && d.force_cornif;
}

struct Carver {
  char* base;
  size_t off = 0;
  bool dry;
  template <class T>
  T* take(size_t count) {
    off = align_up(off, 256);
    T* p = dry ? nullptr : reinterpret_cast<T*>(base + off);
    off += sizeof(T) * (count ? count : 1);
    return p;
  }
};

void carve_tissue(Carver& cv, TH& h, Tis& t) {
  const size_t cap = h.cap, ns = h.ns;
  t.state = cv.take<uint8_t>(cap);
  t.is_malignant = cv.take<uint8_t>(cap);
  t.touch_air = cv.take<uint8_t>(cap);
  t.fix_origin = cv.take<int>(cap);
  t.rest_div_times = cv.take<int>(cap);
  t.pair = cv.take<int>(cap);
  t.dermis = cv.take<int>(cap);
  t.lat = cv.take<int>(3 * cap);
  for (int b = 0; b < 2; b++)
    for (int a = 0; a < 3; a++) t.pos[b][a] = cv.take<double>(cap);
  t.radius = cv.take<double>(cap);
  t.div_age_thresh = cv.take<double>(cap);
  for (int b = 0; b < 2; b++) { t.ca2p[b] = cv.take<double>(cap); t.ip3[b] = cv.take<double>(cap); }
  t.ca2p_avg = cv.take<double>(cap);
  t.ex_inert = cv.take<double>(cap);
  t.agek = cv.take<double>(cap);
  t.ageb = cv.take<double>(cap);
  t.ex_fat = cv.take<double>(cap);
  t.in_fat = cv.take<double>(cap);
  t.spr_nat_len = cv.take<double>(cap);
  t.diff_u = cv.take<double>(cap);
  t.ca_th = cv.take<double>(cap);
  t.ca_kpa = cv.take<double>(cap);
  t.ca_iag = cv.take<double>(cap);
  t.ca_ent = cv.take<int>(cap * ns);
  t.ca_ce = cv.take<int2>(cap * ns);
  t.ca_cn = cv.take<int>(cap);
  for (int b = 0; b < 2; b++) t.ca_up[b] = cv.take<double2>(cap);
  t.ca_dfu = cv.take<double>(cap);
  t.ca_cq = cv.take<uint2>(cap * 8);
  t.uid = cv.take<long long>(cap);
  t.old_uid = cv.take<long long>(cap);
  t.cst = cv.take<long long>(3 * cap);
  t.cst_store = cv.take<long long>(3 * cap);
  t.src_row = cv.take<int>(cap);
  for (int b = 0; b < 2; b++) {
    t.nlen[b] = cv.take<int>(cap);
    t.col[b] = cv.take<int>(cap * ns);
    t.gj[b] = cv.take<double>(cap * ns);
  }
  t.nlo = cv.take<int>(cap);
  t.btmp = cv.take<int>(cap * ns);
  t.bcnt = cv.take<int>(cap);
  t.memb6 = cv.take<int>((size_t)h.nmemb * 6);
  t.bucket_of = cv.take<int>(cap);
  t.bslot = cv.take<int>(cap);
  t.bucket_n = cv.take<int>(h.nbuckets);
  t.bucket_cells = cv.take<int>((size_t)h.nbuckets * bucket_stride(h.bcap));
  t.bucket_sorted = cv.take<int>((size_t)h.nbuckets * bucket_stride(h.bcap));
  t.pk = cv.take<double4>(cap);
  t.div_slot = cv.take<int>(cap);
  t.div_dir = cv.take<double>(3 * cap);
  t.div_parent = cv.take<int>(cap);
  t.rm_list = cv.take<int>(cap);
  t.rflag = cv.take<uint8_t>(cap);
  t.chunk_any = cv.take<int>(cap / 1024 + 2);
  for (int b = 0; b < 2; b++) { t.atp[b] = cv.take<double>(h.gsize); t.bf[b] = cv.take<double>(h.gsize); }
  t.cmap1 = cv.take<int>(h.gsize);
  t.cmap2 = cv.take<uint8_t>(h.gsize);
  t.vox = cv.take<unsigned>(h.gsize);
}

int check_tissue(epi_ctx* c, int t) {
  if (!c) return EPI_ERR_INVALID;
  if (t < 0 || t >= c->T) return fail(c, EPI_ERR_INVALID, "tissue index out of range");
  return EPI_OK;
}

// the status read-back, split so that work issued after the copy can run
// while the host waits for the copy alone (sync_status_end with the stream
// still busy leaves the timing events for a later read-back to resolve)
// after issue_build: its k_snapshot has copied the statuses into h_st
int sync_status_begin(epi_ctx* c) {
  HIPCHK(c, hipEventRecord(c->ev_st, c->stream));
  return EPI_OK;
}
int status_check(epi_ctx* c);
int sync_status_end(epi_ctx* c) {
  HIPCHK(c, hipEventSynchronize(c->ev_st));
  if (!g_prof.pending.empty()) prof_resolve(true);
  return status_check(c);
}

int sync_status(epi_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->h_st, c->d_st, sizeof(Status) * c->T, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!g_prof.pending.empty()) prof_resolve();
  return status_check(c);
}

int status_check(epi_ctx* c) {
  for (int t = 0; t < c->T; t++) {
    c->th[t].n_est = c->h_st[t].n;
    // the device holds zzmax as bits (k_zzmax, set_zz); the host's double
    // follows them, so the Ca lists' iz_bound is the device's
    memcpy(&c->h_st[t].zzmax, &c->h_st[t].zzmax_bits, sizeof(double));
  }
  c->len_known = true;
  for (int t = 0; t < c->T; t++)
    if (c->h_st[t].err) {
      char buf[128];
      snprintf(buf, sizeof buf, "tissue %d: device error code %d", t, c->h_st[t].err);
      return fail(c, c->h_st[t].err, buf);
    }
  return EPI_OK;
}

int put_status(epi_ctx* c, int t) {
  HIPCHK(c, hipMemcpyAsync(c->d_st + t, c->h_st + t, sizeof(Status), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EPI_OK;
}

// the longest neighbour list any build has made (maxlen only grows), or
// INT_MAX when a build has run since the last status read
int max_list_len(const epi_ctx* c) {
  if (!c->len_known) return INT_MAX;
  int m = 0;
  for (int t = 0; t < c->T; t++) m = std::max(m, c->h_st[t].maxlen);
  return m;
}

// launch bounds over all tissues
int max_n(const epi_ctx* c) {
  int m = 1;
  for (auto& h : c->th) m = std::max(m, h.n_est + 64);
  return m;
}
int max_memb(const epi_ctx* c) {
  int m = 0;
  for (auto& h : c->th) m = std::max(m, h.nmemb);
  return m;
}
int max_nb(const epi_ctx* c) {
  int m = 1;
  for (auto& h : c->th) m = std::max(m, h.nbuckets);
  return m;
}
long long max_gsize(const epi_ctx* c) {
  long long m = 1;
  for (auto& h : c->th) m = std::max(m, h.gsize);
  return m;
}
int max_halo(const epi_ctx* c) {
  int m = 1;
  for (auto& h : c->th) m = std::max(m, (h.NX + 1 + h.NY + 1) * h.NZ);
  return m;
}
int max_nx(const epi_ctx* c) {
  int m = 1;
  for (auto& h : c->th) m = std::max(m, h.NX);
  return m;
}
int max_ny(const epi_ctx* c) {
  int m = 1;
  for (auto& h : c->th) m = std::max(m, h.NY);
  return m;
}
// z planes the stencils will likely touch: from the last known zzmax
int izb_est(const epi_ctx* c, bool ca) {
  int m = 1;
  for (int t = 0; t < c->T; t++) {
    const Tis& ti = c->tis[t];
    double zz = c->h_st ? c->h_st[t].zzmax : 0.0;
    int v = ca ? (int)((zz + ti.prm.iz_margin) / ti.dz) : (int)((zz + ti.prm.iz_margin) * ti.inv_dz);
    m = std::max(m, std::min(v + 2, c->th[t].NZ + 1));
  }
  return m;
}

// tiles of the B stencil's work list (k_ext_stim) from the last known
// zzmax, one plane more than it gives, for sizing the grid
long long b_tiles_est(const epi_ctx* c) {
  const int bs = 64 * c->tu.b_rows, zb = c->tu.b_zb;
  long long n = 0;
  for (int t = 0; t < c->T; t++) {
    const Tis& ti = c->tis[t];
    const double zz = c->h_st ? c->h_st[t].zzmax : 0.0;
    const int izb = std::min((int)((zz + ti.prm.iz_margin) * ti.inv_dz) + 1, c->th[t].NZ);
    n += (long long)((ti.NX * ti.NY + bs - 1) / bs) * ((std::max(izb, 1) + zb - 1) / zb);
  }
  return n;
}

long long dev_index(const TH& h, int j, int k, int l) { return (long long)l * h.PL + (long long)k * h.XP + j; }

// [x][y][z] (z fastest) <-> device layout
void to_device_layout(const TH& h, const double* src, std::vector<double>& dst) {
  dst.assign(h.gsize, 0.0);
  for (int j = 0; j <= h.NX; j++)
    for (int k = 0; k <= h.NY; k++)
      for (int l = 0; l <= h.NZ; l++)
        dst[dev_index(h, j, k, l)] = src[((size_t)j * (h.NY + 1) + k) * (h.NZ + 1) + l];
}
template <class T>
void from_device_layout(const TH& h, const std::vector<T>& src, T* dst) {
  for (int j = 0; j <= h.NX; j++)
    for (int k = 0; k <= h.NY; k++)
      for (int l = 0; l <= h.NZ; l++)
        dst[((size_t)j * (h.NY + 1) + k) * (h.NZ + 1) + l] = src[dev_index(h, j, k, l)];
}

template <class T>
int upload(epi_ctx* c, T* dst, const T* src, size_t n) {
  if (!src || !n) return EPI_OK;
  HIPCHK(c, hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EPI_OK;
}
template <class T>
int download(epi_ctx* c, T* dst, const T* src, size_t n) {
  if (!dst || !n) return EPI_OK;
  HIPCHK(c, hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EPI_OK;
}

#define RCHK(x)                  \
  do {                           \
    int r_ = (x);                \
    if (r_ != EPI_OK) return r_; \
  } while (0)

int set_zz(epi_ctx* c, const double* zzmax) {
  if (!zzmax) return EPI_OK;
  for (int t = 0; t < c->T; t++) {
    c->h_zz[t] = zzmax[t];
    HIPCHK(c, hipMemcpyAsync(&c->d_st[t].zzmax_bits, &c->h_zz[t], 8, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int t = 0; t < c->T; t++) c->h_st[t].zzmax = zzmax[t];
  return EPI_OK;
}

// The list of a Ca solve (or window) over tissues ids[0..ntis) (nullptr:
// all of them), for cg lanes per cell and ATP tiles of bs points by zb
// planes, from what the host holds: the cell counts of the last status
// read-back (with a margin of 64 cells; a launch that falls short reports
// ERR_INVALID) and the zzmax of set_zz or of the read-back, whose iz_bound
// (ca2p.cpp:386) is the device's.  Reused while its key holds.
int ca_list(epi_ctx* c, const int* ids, int ntis, int cg, int bs, int zb, CaList* out) {
  const int cpb = 128 / cg;   // cells per block of the Ca cell kernels
  std::vector<long long> key = {ntis, cg, bs, zb};
  std::vector<int> nb(ntis), izb(ntis);
  long long ncb = 0, tiles = 0;
  for (int i = 0; i < ntis; i++) {
    const int t = ids ? ids[i] : i;
    const TH& h = c->th[t];
    // with step graphs, blocks for the cells up to the next multiple of 1024:
    // the list (and a captured solve) then outlives the divisions of many steps
    const int upto = c->tu.step_graph ? std::min((h.n_est + 64 + 1023) / 1024 * 1024, h.cap) : h.n_est + 64;
    const int cells = std::max(upto - h.nmemb - h.nder, 1);
    nb[i] = (cells + cpb - 1) / cpb;
    izb[i] = (int)((c->h_st[t].zzmax + c->tis[t].prm.iz_margin) / c->tis[t].dz);
    if (izb[i] < 0 || izb[i] > h.NZ) izb[i] = 0;   // k_ca_vox reports it
    ncb += nb[i];
    tiles += (long long)((h.NX * h.NY + bs - 1) / bs) * ((izb[i] + zb - 1) / zb);
    key.insert(key.end(), {t, nb[i], izb[i]});
  }
  const long long hdr = (3LL * ntis + 1 + 3) / 4 * 4;
  if (hdr + 4 * ncb > c->cal_ints || tiles > INT32_MAX) return fail(c, EPI_ERR_INVALID, "Ca work list capacity");
  if (key != c->cal_key) {
    HIPCHK(c, hipEventSynchronize(c->ev_cal));   // the last upload has read h_cal
    int* aoff = c->h_cal;
    int* aizb = aoff + ntis + 1;
    int* atis = aizb + ntis;
    int* cm = c->h_cal + hdr;
    int o = 0, b = 0;
    for (int i = 0; i < ntis; i++) {
      const int t = ids ? ids[i] : i;
      const TH& h = c->th[t];
      aoff[i] = o;
      aizb[i] = izb[i];
      atis[i] = t;
      o += ((h.NX * h.NY + bs - 1) / bs) * ((izb[i] + zb - 1) / zb);
      for (int k = 0; k < nb[i]; k++, b++) {
        cm[4 * b] = t;
        cm[4 * b + 1] = k;
        cm[4 * b + 2] = nb[i];
        cm[4 * b + 3] = i;
      }
    }
    aoff[ntis] = o;
    HIPCHK(c, hipMemcpyAsync(c->d_cal, c->h_cal, sizeof(int) * (hdr + 4 * ncb), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(c->ev_cal, c->stream));
    c->cal_key = key;
    c->cal = CaList{c->d_cal, c->d_cal + ntis + 1, c->d_cal + 2 * ntis + 1,
                    reinterpret_cast<const int4*>(c->d_cal + hdr), ntis, (int)ncb, (int)tiles, bs,
                    ids ? ids[0] : 0, izb[0]};
  }
  *out = c->cal;
  return EPI_OK;
}

// Replays the graph of `key`, or captures what `issue` launches on the
// stream into a new one (kept with the others, least recently used evicted)
// and launches it.  Captured launches carry no timing events.
template <class F>
int graph_run(epi_ctx* c, const std::vector<long long>& key, F issue) {
  for (auto& g : c->graphs)
    if (g.key == key) {
      g.used = ++c->graph_tick;
      HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
      return EPI_OK;
    }
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  g_prof.capturing = true;
  const hipError_t e1 = issue();
  g_prof.capturing = false;
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(c, EPI_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t e3 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e3 != hipSuccess) return fail(c, EPI_ERR_HIP, std::string("graph instantiation: ") + hipGetErrorString(e3));
  constexpr size_t MAX_GRAPHS = 12;
  if (c->graphs.size() >= MAX_GRAPHS) {
    auto old = std::min_element(c->graphs.begin(), c->graphs.end(),
                                [](const epi_ctx::StepGraph& a, const epi_ctx::StepGraph& b) { return a.used < b.used; });
    (void)hipGraphExecDestroy(old->exec);
    c->graphs.erase(old);
  }
  c->graphs.push_back({key, exec, ++c->graph_tick});
  HIPCHK(c, hipGraphLaunch(exec, c->stream));
  return EPI_OK;
}

// calc_ca2p (ca2p.cpp:380-402) for the tissues in tlist (nullptr = all)
// (with d_tlist, c->h_tlist holds the same ntis ids), n_sub sub-steps.  The
// list covers the solving tissues only: a gated solve of one small tissue in
// a large batch launches that tissue's blocks, not the batch's largest.
int run_ca_n(epi_ctx* c, const int* d_tlist, int ntis, int n_sub) {
  int mn = 1;
  long long box = 1;
  for (int i = 0; i < ntis; i++) {
    const int t = d_tlist ? c->h_tlist[i] : i;
    const TH& h = c->th[t];
    const double zz = c->h_st[t].zzmax;
    const int iz = std::min((int)((zz + c->tis[t].prm.iz_margin) / c->tis[t].dz) + 2, h.NZ + 1);
    mn = std::max(mn, h.n_est + 64);
    box = std::max(box, (long long)h.NX * h.NY * iz);
  }
  // ca_fuse 2: fuse the solves of at most FUSE_AUTO_CELLS cells (a gated solve
  // of a few tissues is latency-bound, where one launch per sub-step wins; a
  // solve of a whole large batch is throughput-bound, where two leaner ones do)
  constexpr long long FUSE_AUTO_CELLS = 131072;
  long long cells = 0;
  for (int i = 0; i < ntis; i++) cells += c->th[d_tlist ? c->h_tlist[i] : i].n_est;
  Tuning tu = c->tu;
  if (tu.ca_fuse == 2 && cells <= FUSE_AUTO_CELLS) {
    // a small, latency-bound solve: fused, with its own lanes per cell, and
    // one-plane ATP tiles (the planes per thread chosen for the batch-wide
    // grid kernel roll 8 or 16 planes through one thread, which a
    // latency-bound solve waits for)
    tu.ca_fuse = 1;
    tu.ca_group = tu.ca_group_small;
    if (tu.a_zb > 2) tu.a_zb = 1;
  } else {
    tu.ca_fuse = tu.ca_fuse == 1;
  }
  CaList cl;
  RCHK(ca_list(c, d_tlist ? c->h_tlist : nullptr, ntis, tu.ca_group, ca_list_bs(tu, tu.ca_fuse && n_sub >= 2),
               tu.a_zb, &cl));
  constexpr unsigned long long seg = (1ull << K_CA_BEGIN) | (1ull << K_CA_VOX) | (1ull << K_CA_CORNERS) |
                                     (1ull << K_CA_CELLS) | (1ull << K_ATP) | (1ull << K_CA_FUSED) |
                                     (1ull << K_CA_END) | (1ull << K_SW_RESET);
  if (c->tu.step_graph && !(g_prof.mask & seg)) {
    // the whole solve (set-up, n_sub sub-steps, averages) as one graph per
    // launch key; its grids cover the cells up to the next multiple of 1024
    int mcap = 1;
    for (int i = 0; i < ntis; i++) mcap = std::max(mcap, c->th[d_tlist ? c->h_tlist[i] : i].cap);
    const int mr = std::max(mn, std::min((mn + 1023) / 1024 * 1024, mcap));
    std::vector<long long> key = {2, d_tlist != nullptr, ntis, n_sub, mr, box, c->cb, c->cp, c->ap, c->bp,
                                  tu.ca_fuse, tu.ca_group, tu.a_zb, tu.a_rows, tu.xcd};
    key.insert(key.end(), c->cal_key.begin(), c->cal_key.end());
    RCHK(graph_run(c, key, [&] {
      int cp = c->cp, ap = c->ap;
      launch_ca_begin(c->stream, c->d_tis, d_tlist, ntis, c->cb, mr, box);
      launch_ca_substeps(c->stream, c->d_tis, cl, c->cb, cp, ap, c->bp, n_sub, tu);
      launch_ca_end(c->stream, c->d_tis, d_tlist, ntis, mr);
      return hipGetLastError();
    }));
    c->cp ^= n_sub & 1;
    c->ap ^= n_sub & 1;
  } else {
    launch_ca_begin(c->stream, c->d_tis, d_tlist, ntis, c->cb, mn, box);
    launch_ca_substeps(c->stream, c->d_tis, cl, c->cb, c->cp, c->ap, c->bp, n_sub, tu);
    launch_ca_end(c->stream, c->d_tis, d_tlist, ntis, mn);
  }
  if ((n_sub & 1) && d_tlist && ntis < c->T) {
    // the parities name the same buffer for every tissue: an odd count
    // flipped them for all, but advanced only the listed tissues.  They go
    // back, and the listed tissues' buffers trade contents instead.
    c->cp ^= 1;
    c->ap ^= 1;
    long long gmax = 1;
    for (int i = 0; i < ntis; i++) gmax = std::max(gmax, c->th[c->h_tlist[i]].gsize);
    launch_ca_swap(c->stream, c->d_tis, d_tlist, ntis, mn, gmax);
  }
  HIPCHK(c, hipGetLastError());
  return EPI_OK;
}

// calc_ca2p with each tissue's own sub-step count, Ca_ITR (ca2p.cpp:14,
// 392).  Tissues that share it solve together; a list that mixes counts
// (tissues given different Ca_avg_time / DT_Ca) solves one count at a time.
int run_ca(epi_ctx* c, const int* d_tlist, int ntis) {
  auto id = [&](int i) { return d_tlist ? c->h_tlist[i] : i; };
  const int first = c->tis[id(0)].prm.Ca_ITR;
  bool same = true;
  for (int i = 1; i < ntis; i++) same = same && c->tis[id(i)].prm.Ca_ITR == first;
  if (same) return run_ca_n(c, d_tlist, ntis, first);
  std::map<int, std::vector<int>> groups;
  for (int i = 0; i < ntis; i++) groups[c->tis[id(i)].prm.Ca_ITR].push_back(id(i));
  for (const auto& g : groups) {
    // h_tlist may still feed an earlier copy: wait for the stream first
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::copy(g.second.begin(), g.second.end(), c->h_tlist);
    HIPCHK(c, hipMemcpy(c->d_tlist, c->h_tlist, sizeof(int) * g.second.size(), hipMemcpyHostToDevice));
    RCHK(run_ca_n(c, c->d_tlist, (int)g.second.size(), g.first));
  }
  return EPI_OK;
}

void memb_init_host(int nmx, int nmemb, std::vector<int>& m) {   // cellmanager.cpp:66-113
  m.assign((size_t)nmemb * 6, 0);
  for (int j = 0; j < nmemb; j++) {
    int jj = j % nmx;
    m[6 * j + 3] = (jj == 0) ? j + nmx - 1 : j - 1;
    m[6 * j + 0] = (jj == nmx - 1) ? j - (nmx - 1) : j + 1;
  }
  for (int j = 0; j < nmemb; j++) {
    int kk = j / nmx;
    int top = (j - nmx + nmemb) % nmemb;
    int bot = (j + nmx) % nmemb;
    if (kk % 2 == 0) {
      m[6 * j + 1] = top;
      m[6 * j + 2] = m[6 * top + 3];
      m[6 * j + 4] = m[6 * bot + 3];
      m[6 * j + 5] = bot;
    } else {
      m[6 * j + 1] = m[6 * top + 0];
      m[6 * j + 2] = top;
      m[6 * j + 4] = bot;
      m[6 * j + 5] = m[6 * bot + 0];
    }
  }
}

}  // namespace

extern "C" {

int epi_create(int32_t ntissues, const epi_geometry* geos, epi_ctx** out) {
  if (!out || ntissues <= 0 || !geos) return EPI_ERR_INVALID;
  *out = nullptr;
  epi_ctx* c = new epi_ctx();
  c->T = ntissues;
  c->th.resize(ntissues);
  c->tis.resize(ntissues);
  c->ca_solves.assign(ntissues, 0);
  c->dev_ts.assign(ntissues, 0);
  c->steps.assign(ntissues, 0);
  c->cell_steps.assign(ntissues, 0);
  c->grid_ca.assign(ntissues, 0);
  for (int t = 0; t < ntissues; t++) {
    const epi_geometry& g = geos[t];
    TH& h = c->th[t];
    Tis& ti = c->tis[t];
    if (g.nmx < 3 || g.nmy < 4 || (g.nmy % 2) || g.nx < 2 || g.ny < 2 || g.nz < 2 || g.max_cells <= g.nmx * g.nmy ||
        !(g.lx > 0) || !(g.ly > 0) || !(g.lz > 0)) {
      delete c;
      return EPI_ERR_INVALID;
    }
    h.g = g;
    h.nmemb = g.nmx * g.nmy;
    h.cap = g.max_cells;
    // a row holds at least 64 slots: the Ca cell kernel reads its first list
    // chunk (up to 64 entries) before it knows the length
    h.ns = g.row_stride > 0 ? std::max(64, std::min(g.row_stride, MAX_CONN)) : MAX_CONN;
    h.bcap = g.bucket_cap > 0 ? std::min(g.bucket_cap, N3) : N3;
    h.NX = g.nx; h.NY = g.ny; h.NZ = g.nz;
    h.XP = (int)align_up(g.nx + 1, 16);
    h.PL = (long long)h.XP * (g.ny + 1);
    h.gsize = h.PL * (g.nz + 1);
    memset(&ti, 0, sizeof(Tis));
    ti.nmemb = h.nmemb; ti.nder = 0; ti.cap = h.cap; ti.ns = h.ns;
    ti.NX = g.nx; ti.NY = g.ny; ti.NZ = g.nz; ti.XP = h.XP; ti.PL = h.PL; ti.gsize = h.gsize;
    ti.LX = g.lx; ti.LY = g.ly; ti.LZ = g.lz;
    // define.h:111-113
    ti.dx = g.lx / g.nx; ti.inv_dx = g.nx / g.lx;
    ti.dy = g.ly / g.ny; ti.inv_dy = g.ny / g.ly;
    ti.dz = g.lz / g.nz; ti.inv_dz = g.nz / g.lz;
    // cell_connection.cpp:22-26
    ti.ANX = (int)((double)g.lx / 2.0 + 0.5);
    ti.ANY = (int)((double)g.ly / 2.0 + 0.5);
    ti.ANZ = (int)((double)g.lz / 2.0);
    if (ti.ANX < 5 || ti.ANY < 5 || ti.ANZ < 5) { delete c; return EPI_ERR_INVALID; }
    h.nbuckets = ti.ANX * ti.ANY * ti.ANZ;
    ti.bcap = h.bcap;
    // map.cpp:27-31
    const double Rad = 1.4;
    ti.irx = (int)(Rad * g.nx / g.lx);
    ti.iry = (int)(Rad * g.ny / g.ly);
    ti.irz = (int)(Rad * g.nz / g.lz);
    // epi5_MT's parameters with this geometry's grid and box
    epi_params dp;
    epi_params_default(&dp);
    dp.LX = g.lx; dp.LY = g.ly; dp.LZ = g.lz;
    dp.NX = g.nx; dp.NY = g.ny; dp.NZ = g.nz;
    derive_params(dp, ti.prm);
    window_values_from_params(ti);
  }
  // the Ca list: 3 ints per tissue + 1, padding, and one int4 per cell block
  // (8 cells per block at the most lanes per cell)
  {
    long long blocks = 0;
    for (int t = 0; t < ntissues; t++) blocks += (c->th[t].cap + 64 + 7) / 8;
    c->cal_ints = 3LL * ntissues + 8 + 4 * blocks;
  }
  // arena
  Carver dry{nullptr, 0, true};
  for (int t = 0; t < ntissues; t++) carve_tissue(dry, c->th[t], c->tis[t]);
  c->arena_bytes = align_up(dry.off, 256);
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return EPI_ERR_HIP; }
  if (hipStreamCreate(&c->side_stream) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_st, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_cal, hipEventDisableTiming) != hipSuccess) {
    epi_destroy(c);
    return EPI_ERR_HIP;
  }
  if (hipMalloc(&c->arena, c->arena_bytes) != hipSuccess ||
      hipMalloc(&c->d_tis, sizeof(Tis) * ntissues) != hipSuccess ||
      hipMalloc(&c->d_st, sizeof(Status) * ntissues) != hipSuccess ||
      hipMalloc(&c->d_tlist, sizeof(int) * ntissues) != hipSuccess ||
      hipMalloc(&c->d_wl, sizeof(int) * WL_INTS(ntissues)) != hipSuccess ||
      hipMalloc(&c->d_cal, sizeof(int) * c->cal_ints) != hipSuccess ||
      hipHostMalloc(&c->h_cal, sizeof(int) * c->cal_ints, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&c->h_st, sizeof(Status) * ntissues, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&c->h_zz, sizeof(double) * ntissues, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&c->h_tlist, sizeof(int) * ntissues, hipHostMallocDefault) != hipSuccess) {
    epi_destroy(c);
    return EPI_ERR_HIP;
  }
  if (hipMemsetAsync(c->arena, 0, c->arena_bytes, c->stream) != hipSuccess ||
      hipMemsetAsync(c->d_st, 0, sizeof(Status) * ntissues, c->stream) != hipSuccess) {
    epi_destroy(c);
    return EPI_ERR_HIP;
  }
  memset(c->h_st, 0, sizeof(Status) * ntissues);
  Carver real{c->arena, 0, false};
  for (int t = 0; t < ntissues; t++) {
    carve_tissue(real, c->th[t], c->tis[t]);
    c->tis[t].st = c->d_st + t;
  }
  if (hipMemcpyAsync(c->d_tis, c->tis.data(), sizeof(Tis) * ntissues, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    epi_destroy(c);
    return EPI_ERR_HIP;
  }
  // cmap1 starts empty (-1)
  std::vector<int> neg;
  for (int t = 0; t < ntissues; t++) {
    neg.assign(c->th[t].gsize, -1);
    if (hipMemcpy(c->tis[t].cmap1, neg.data(), sizeof(int) * neg.size(), hipMemcpyHostToDevice) != hipSuccess) {
      epi_destroy(c);
      return EPI_ERR_HIP;
    }
  }
  *out = c;
  return EPI_OK;
}

void epi_destroy(epi_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  for (auto& S : c->snap) {
    if (S.dev) (void)hipFree(S.dev);
    if (S.host) (void)hipHostFree(S.host);
    if (S.ev_pack) (void)hipEventDestroy(S.ev_pack);
    if (S.ev_done) (void)hipEventDestroy(S.ev_done);
  }
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->rd.arena) (void)hipFree(c->rd.arena);
  if (c->mech.arena) (void)hipFree(c->mech.arena);
  if (c->mech.work) (void)hipFree(c->mech.work);
  if (c->mech.d) (void)hipFree(c->mech.d);
  if (c->rd.work) (void)hipFree(c->rd.work);
  if (c->rd.d) (void)hipFree(c->rd.d);
  if (c->rd.d_gj0) (void)hipFree(c->rd.d_gj0);
  for (auto& h : c->th)
    if (h.rng_d) (void)hipFree(h.rng_d);
  if (c->arena) (void)hipFree(c->arena);
  if (c->d_tis) (void)hipFree(c->d_tis);
  if (c->d_st) (void)hipFree(c->d_st);
  if (c->d_tlist) (void)hipFree(c->d_tlist);
  if (c->d_wl) (void)hipFree(c->d_wl);
  if (c->d_cal) (void)hipFree(c->d_cal);
  if (c->h_cal) (void)hipHostFree(c->h_cal);
  if (c->ev_cal) (void)hipEventDestroy(c->ev_cal);
  if (c->h_st) (void)hipHostFree(c->h_st);
  if (c->h_zz) (void)hipHostFree(c->h_zz);
  if (c->h_tlist) (void)hipHostFree(c->h_tlist);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->side_stream) (void)hipStreamDestroy(c->side_stream);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_st) (void)hipEventDestroy(c->ev_st);
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  delete c;
}

const char* epi_last_error(const epi_ctx* c) { return c ? c->err.c_str() : "null context"; }
int32_t epi_num_tissues(const epi_ctx* c) { return c ? c->T : 0; }

int epi_load_cells(epi_ctx* c, int32_t t, const epi_cell_input* in) {
  RCHK(check_tissue(c, t));
  TH& h = c->th[t];
  Tis& ti = c->tis[t];
  if (!in || in->n <= 0 || in->n > h.cap) return fail(c, EPI_ERR_CAPACITY, "cell count exceeds capacity");
  if (in->nmemb != h.nmemb) return fail(c, EPI_ERR_INVALID, "nmemb must equal nmx*nmy (cellmanager.cpp:318-321)");
  const int n = in->n;
  // phase order MEMB, DER, rest (cellmanager.cpp:266-282)
  for (int i = 0; i < n; i++) {
    const uint8_t s = in->state[i];
    const bool want_memb = i < in->nmemb, want_der = !want_memb && i < in->nmemb + in->nder;
    if ((s == MEMB) != want_memb || (s == DER) != want_der) return fail(c, EPI_ERR_INVALID, "rows must be MEMB, then DER, then the rest");
  }
  const Params& pm = ti.prm;
  std::vector<uint8_t> st(n), mal(n);
  std::vector<int> fo(n), rdt(n), pair(n, -1), derm(n, -1), src(n, -1), nl(n, 0);
  std::vector<double> dat(n), ca(n), ip(n);
  std::vector<long long> uid(n);
  for (int i = 0; i < n; i++) {
    st[i] = in->state[i];
    mal[i] = in->is_malignant ? in->is_malignant[i] : 0;
    fo[i] = (st[i] == DER || st[i] == MEMB) ? -1 : in->fix_origin[i];           // cell.cpp:67
    rdt[i] = st[i] == FIX ? (int)pm.div_max : in->rest_div_times[i];            // cell.cpp:18-20
    dat[i] = st[i] == FIX ? pm.agki_max_fix : st[i] == MUSUME ? pm.agki_max : 0.0;  // cell.cpp:11-15
    uid[i] = i;
    if (in->pair && in->pair[i] >= 0) {
      if (in->pair[i] >= n) return fail(c, EPI_ERR_INVALID, "pair out of range");
      pair[i] = in->pair[i];
    }
  }
  for (int i = 0; i < n; i++)
    if (pair[i] >= 0 && pair[pair[i]] != i) return fail(c, EPI_ERR_INVALID, "pairs must be mutual");
  std::vector<int> m6;
  memb_init_host(h.g.nmx, h.nmemb, m6);
  // the gather in k_forces_memb needs a symmetric adjacency of distinct cells
  for (int j = 0; j < h.nmemb; j++)
    for (int a = 0; a < 6; a++) {
      const int q = m6[6 * j + a];
      int cnt = 0, back = 0;
      for (int b = 0; b < 6; b++) { cnt += m6[6 * j + b] == q; back += m6[6 * q + b] == j; }
      if (cnt != 1 || back != 1 || q == j) return fail(c, EPI_ERR_INVALID, "membrane adjacency is not symmetric");
    }
  h.nder = in->nder;
  ti.nder = in->nder;
  HIPCHK(c, hipMemcpy(c->d_tis + t, &ti, sizeof(Tis), hipMemcpyHostToDevice));
  RCHK(upload(c, ti.state, st.data(), n));
  RCHK(upload(c, ti.is_malignant, mal.data(), n));
  RCHK(upload(c, ti.fix_origin, fo.data(), n));
  RCHK(upload(c, ti.rest_div_times, rdt.data(), n));
  RCHK(upload(c, ti.pair, pair.data(), n));
  RCHK(upload(c, ti.dermis, derm.data(), n));
  RCHK(upload(c, ti.src_row, src.data(), n));
  RCHK(upload(c, ti.nlen[0], nl.data(), n));
  RCHK(upload(c, ti.nlen[1], nl.data(), n));
  RCHK(upload(c, ti.uid, uid.data(), n));
  RCHK(upload(c, ti.old_uid, uid.data(), n));
  for (int b = 0; b < 2; b++) {
    RCHK(upload(c, ti.pos[b][0], in->x, n));
    RCHK(upload(c, ti.pos[b][1], in->y, n));
    RCHK(upload(c, ti.pos[b][2], in->z, n));
    RCHK(upload(c, ti.ca2p[b], in->ca2p, n));
    RCHK(upload(c, ti.ip3[b], in->ip3, n));
  }
  RCHK(upload(c, ti.radius, in->radius, n));
  RCHK(upload(c, ti.div_age_thresh, dat.data(), n));
  RCHK(upload(c, ti.ca2p_avg, in->ca2p_avg, n));
  RCHK(upload(c, ti.ex_inert, in->ex_inert, n));
  RCHK(upload(c, ti.agek, in->agek, n));
  RCHK(upload(c, ti.ageb, in->ageb, n));
  RCHK(upload(c, ti.ex_fat, in->ex_fat, n));
  RCHK(upload(c, ti.in_fat, in->in_fat, n));
  RCHK(upload(c, ti.spr_nat_len, in->spr_nat_len, n));
  RCHK(upload(c, ti.memb6, m6.data(), m6.size()));
  const std::vector<long long> unset(3 * (size_t)n, -1);   // cell_stat fields never set
  RCHK(upload(c, ti.cst, unset.data(), unset.size()));
  Status s{};
  s.n = n;
  s.next_uid = n;
  s.stamp_z = h.NZ;   // the first map build clears every plane
  c->h_st[t] = s;
  RCHK(put_status(c, t));
  c->dev_ts[t] = 0;
  h.n_est = n;
  h.loaded = true;
  return EPI_OK;
}

int epi_value_init(epi_ctx* c, int32_t t) {
  RCHK(check_tissue(c, t));
  RCHK(sync_status(c));
  const Tis& ti = c->tis[t];
  const TH& h = c->th[t];
  const int n = c->h_st[t].n;
  std::vector<uint8_t> st(n);
  RCHK(download(c, st.data(), ti.state, n));
  const Params& pm = ti.prm;   // cell_init.cpp:23-28
  std::vector<double> ca(n, pm.ca2p_init), avg(n, pm.ca2p_init), exi(n, pm.ex_inert_init), ip(n, pm.IP3_init);
  for (int i = h.nmemb + h.nder; i < n; i++)
    if (st[i] == DEAD) { avg[i] = 0; ca[i] = 0; }
  for (int b = 0; b < 2; b++) {
    RCHK(upload(c, ti.ca2p[b], ca.data(), n));
    RCHK(upload(c, ti.ip3[b], ip.data(), n));
  }
  RCHK(upload(c, ti.ca2p_avg, avg.data(), n));
  RCHK(upload(c, ti.ex_inert, exi.data(), n));
  return EPI_OK;
}

int epi_set_cell_values(epi_ctx* c, int32_t t, const double* ca2p, const double* ip3, const double* ex_inert,
                        const double* ca2p_avg, const double* gj) {
  RCHK(check_tissue(c, t));
  RCHK(sync_status(c));
  const Tis& ti = c->tis[t];
  const TH& h = c->th[t];
  const int n = c->h_st[t].n;
  for (int b = 0; b < 2; b++) {
    RCHK(upload(c, ti.ca2p[b], ca2p, n));
    RCHK(upload(c, ti.ip3[b], ip3, n));
  }
  RCHK(upload(c, ti.ex_inert, ex_inert, n));
  RCHK(upload(c, ti.ca2p_avg, ca2p_avg, n));
  if (gj) {
    std::vector<int> nl(n);
    RCHK(download(c, nl.data(), ti.nlen[c->cb], n));
    std::vector<double> g((size_t)n * h.ns, 0.0);
    size_t p = 0;
    for (int i = 0; i < n; i++)
      for (int e = 0; e < nl[i]; e++) g[(size_t)i * h.ns + e] = gj[p++];
    RCHK(upload(c, ti.gj[c->cb], g.data(), g.size()));
  }
  return EPI_OK;
}

int epi_set_fields(epi_ctx* c, int32_t t, const double* atp, const double* atp_next, const double* b,
                   const double* b_next) {
  RCHK(check_tissue(c, t));
  const Tis& ti = c->tis[t];
  const TH& h = c->th[t];
  std::vector<double> tmp;
  const double* srcs[4] = {atp, atp_next, b, b_next};
  double* dsts[4] = {ti.atp[c->ap], ti.atp[1 - c->ap], ti.bf[c->bp], ti.bf[1 - c->bp]};
  for (int i = 0; i < 4; i++) {
    if (!srcs[i]) continue;
    to_device_layout(h, srcs[i], tmp);
    RCHK(upload(c, dsts[i], tmp.data(), tmp.size()));
  }
  return EPI_OK;
}

int epi_set_rng(epi_ctx* c, int32_t t, const double* draws, int64_t n) {
  RCHK(check_tissue(c, t));
  if (n < 0 || (n > 0 && !draws)) return fail(c, EPI_ERR_INVALID, "bad stream");
  RCHK(sync_status(c));
  TH& h = c->th[t];
  if (h.rng_d) { (void)hipFree(h.rng_d); h.rng_d = nullptr; }
  if (n > 0) {
    HIPCHK(c, hipMalloc(&h.rng_d, sizeof(double) * n));
    RCHK(upload(c, h.rng_d, draws, n));
  }
  h.rng_n = n;
  c->tis[t].rng = h.rng_d;
  c->tis[t].rng_n = n;
  HIPCHK(c, hipMemcpy(c->d_tis + t, &c->tis[t], sizeof(Tis), hipMemcpyHostToDevice));
  c->h_st[t].rng_pos = 0;
  return put_status(c, t);
}

int epi_rng_mt19937_64(uint64_t seed, double* out, int64_t n) {
  if (n < 0 || (n > 0 && !out)) return EPI_ERR_INVALID;
  std::mt19937_64 mt(seed);
  std::uniform_real_distribution<double> dist(0.0, 1.0);
  for (int64_t i = 0; i < n; i++) out[i] = dist(mt);
  return EPI_OK;
}

int epi_cell_interaction(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  launch_forces(c->stream, c->d_tis, c->T, c->pb, c->cb, max_memb(c), max_n(c), max_list_len(c));
  HIPCHK(c, hipGetLastError());
  c->pos_pending = true;
  return EPI_OK;
}

int epi_pos_copy(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  if (c->pos_pending) {
    c->pb ^= 1;
    c->pos_pending = false;
  }
  return EPI_OK;
}

int epi_cell_state_renew(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  if (c->pos_pending) return fail(c, EPI_ERR_INVALID, "pos_copy must follow cell_interaction");
  launch_renew(c->stream, c->d_tis, c->T, c->pb, c->cb, c->cp, max_n(c));
  HIPCHK(c, hipGetLastError());
  for (auto& v : c->dev_ts) v++;   // k_remove advances each tissue's timestep
  return EPI_OK;
}

int epi_cell_pos_periodic_fix(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  launch_periodic_fix(c->stream, c->d_tis, c->T, c->pb, max_n(c));
  HIPCHK(c, hipGetLastError());
  return EPI_OK;
}

int epi_connect_cell(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  const int nb = 1 - c->cb;
  launch_connect(c->stream, c->d_tis, c->T, c->pb, nb, max_n(c), max_nb(c), 0, c->tu.sweep_wave);
  c->len_known = false;
  HIPCHK(c, hipGetLastError());
  c->cb = nb;
  return EPI_OK;
}

int epi_calc_zzmax(epi_ctx* c, double* zzmax) {
  if (!c) return EPI_ERR_INVALID;
  launch_zzmax(c->stream, c->d_tis, c->T, c->pb, max_n(c));
  HIPCHK(c, hipGetLastError());
  RCHK(sync_status(c));
  for (int t = 0; t < c->T; t++) {
    double v;
    unsigned long long b = c->h_st[t].zzmax_bits;
    memcpy(&v, &b, 8);
    c->h_st[t].zzmax = v;
    if (zzmax) zzmax[t] = v;
  }
  return EPI_OK;
}

int epi_setup_map_lat(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  launch_maps(c->stream, c->d_tis, c->T, c->pb, max_n(c), max_gsize(c), max_halo(c), c->tu);
  HIPCHK(c, hipGetLastError());
  return EPI_OK;
}

int epi_calc_ext_stim(epi_ctx* c, const double* zzmax) {
  if (!c) return EPI_ERR_INVALID;
  RCHK(set_zz(c, zzmax));
  launch_ext_stim(c->stream, c->d_tis, c->T, c->bp, b_tiles_est(c), c->d_wl, c->tu);
  HIPCHK(c, hipGetLastError());
  c->bp ^= 1;
  return EPI_OK;
}

int epi_initialize_sc(epi_ctx* c, const double* zzmax) {
  if (!c) return EPI_ERR_INVALID;
  RCHK(set_zz(c, zzmax));
  launch_initialize_sc(c->stream, c->d_tis, nullptr, c->T, c->pb, max_n(c));
  HIPCHK(c, hipGetLastError());
  return EPI_OK;
}

int epi_calc_ca2p(epi_ctx* c, const double* zzmax) {
  if (!c) return EPI_ERR_INVALID;
  RCHK(set_zz(c, zzmax));
  RCHK(run_ca(c, nullptr, c->T));
  return sync_status(c);
}

int epi_calc_ca2p_n(epi_ctx* c, const double* zzmax, int32_t n_sub) {
  if (!c || n_sub < 0) return EPI_ERR_INVALID;
  RCHK(set_zz(c, zzmax));
  RCHK(run_ca_n(c, nullptr, c->T, n_sub));
  return sync_status(c);
}

int epi_should_calc_ca(epi_ctx* c, int32_t* flags) {
  if (!c) return EPI_ERR_INVALID;
  RCHK(sync_status(c));
  for (int t = 0; t < c->T; t++) flags[t] = (c->h_st[t].sw & 255) >= c->tis[t].prm.SW_THRESH;
  return EPI_OK;
}

int epi_ca_calc_condition_reset(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  RCHK(sync_status(c));
  for (int t = 0; t < c->T; t++) {
    c->h_st[t].sw = 0;
    HIPCHK(c, hipMemcpyAsync(&c->d_st[t].sw, &c->h_st[t].sw, sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EPI_OK;
}

int epi_set_timestep(epi_ctx* c, int32_t t, int64_t step) {
  RCHK(check_tissue(c, t));
  RCHK(sync_status(c));
  c->h_st[t].timestep = step;
  HIPCHK(c, hipMemcpyAsync(&c->d_st[t].timestep, &c->h_st[t].timestep, sizeof(long long), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->dev_ts[t] = step;
  return EPI_OK;
}

int epi_take_cell_stats(epi_ctx* c, int32_t t, int64_t* records, int32_t max_records, int32_t* n_records) {
  RCHK(check_tissue(c, t));
  if (!n_records || max_records < 0 || (max_records > 0 && !records)) return EPI_ERR_INVALID;
  RCHK(sync_status(c));
  const int k = c->h_st[t].nstat;
  *n_records = k;
  if (k > max_records) return fail(c, EPI_ERR_CAPACITY, "more cell_stat records than max_records");
  if (k) {
    std::vector<long long> r(3 * (size_t)k);
    RCHK(download(c, r.data(), c->tis[t].cst_store, r.size()));
    for (size_t i = 0; i < r.size(); i++) records[i] = r[i];
  }
  c->h_st[t].nstat = 0;
  HIPCHK(c, hipMemcpyAsync(&c->d_st[t].nstat, &c->h_st[t].nstat, sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EPI_OK;
}

int epi_ca_substep(epi_ctx* c, int32_t iz_bound) {
  if (!c) return EPI_ERR_INVALID;
  // a zzmax whose calc_ca2p bound is iz_bound
  std::vector<double> zz(c->T);
  for (int t = 0; t < c->T; t++) zz[t] = iz_bound * c->tis[t].dz - c->tis[t].prm.iz_margin + 0.5 * c->tis[t].dz;
  RCHK(set_zz(c, zz.data()));
  for (int t = 0; t < c->T; t++)
    if ((int)((zz[t] + c->tis[t].prm.iz_margin) / c->tis[t].dz) != iz_bound)
      return fail(c, EPI_ERR_INVALID, "iz_bound");
  const int mn = max_n(c);
  long long box = 1;
  for (auto& h : c->th) box = std::max(box, (long long)h.NX * h.NY * (iz_bound + 1));
  CaList cl;
  RCHK(ca_list(c, nullptr, c->T, c->tu.ca_group, ca_list_bs(c->tu, false), c->tu.a_zb, &cl));
  launch_ca_touch(c->stream, c->d_tis, nullptr, c->T, c->cb, mn, box);
  launch_ca_substep(c->stream, c->d_tis, cl, c->cb, c->cp, c->ap, c->bp, c->tu);
  c->cp ^= 1;
  c->ap ^= 1;
  launch_ca_diffu_zero(c->stream, c->d_tis, nullptr, c->T, mn);
  HIPCHK(c, hipGetLastError());
  return sync_status(c);
}

// cell_state_renew (proc.cpp:153), the bucket insertion with the periodic
// fix and the rest of connect_cell (:154), and, on the side stream forked
// after the insertion, calc_zzmax, setup_map_lat and calc_ext_stim
// (:155-157), which read only the fixed positions, states and radii (and the
// B field) that the rest of the build leaves alone.  The build, the longer
// chain, is issued first.
// With join false the main stream does not wait for the side stream's maps
// and B sweep: epi_step joins it after issuing the next step's forces, which
// read neither.
hipError_t issue_build(epi_ctx* c, int mn, int nb, int izb, bool join = true) {
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
  launch_renew(c->stream, c->d_tis, c->T, c->pb, c->cb, c->cp, mn);
  launch_connect_insert(c->stream, c->d_tis, c->T, c->pb, mn, 1);
  // the side stream forks after the insertion (SIDE_FORK 0: the maps run
  // beside the sweep) or after the whole build (1: the sweep runs alone and
  // the maps beside the next forces)
  if (!c->tu.side_fork) chk(hipEventRecord(c->ev_fork, c->stream));
  launch_connect_rest(c->stream, c->d_tis, c->T, c->pb, nb, mn, c->tu.sweep_wave, c->h_st);
  if (c->tu.side_fork) chk(hipEventRecord(c->ev_fork, c->stream));
  chk(hipStreamWaitEvent(c->side_stream, c->ev_fork, 0));
  launch_zzmax(c->side_stream, c->d_tis, c->T, c->pb, mn);
  launch_maps(c->side_stream, c->d_tis, c->T, c->pb, mn, max_gsize(c), max_halo(c), c->tu);
  launch_ext_stim(c->side_stream, c->d_tis, c->T, c->bp, b_tiles_est(c), c->d_wl, c->tu);
  chk(hipEventRecord(c->ev_join, c->side_stream));
  if (join) chk(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  chk(hipGetLastError());
  return e;
}

// issue_build launched one kernel at a time, or replayed from a graph
// captured once per launch key.  A replay's grids cover the cells up to the
// next multiple of 1024 (at most the largest capacity): every kernel of the
// chain skips cells past the tissue's count, and the rounding lets a key
// outlive the divisions of many steps.  Launches whose kernels are timed go
// one by one, so that their events exist.
// *joined tells whether the main stream already waits for the side stream.
int run_build(epi_ctx* c, int mn, int nb, int izb, bool* joined) {
  constexpr unsigned long long seg =
      (1ull << K_RENEW_FLAGS) | (1ull << K_RENEW_SCAN) | (1ull << K_RENEW_APPLY) | (1ull << K_REMOVE) |
      (1ull << K_PAIR_DISPERSE) | (1ull << K_BUCKET_INSERT) | (1ull << K_BUCKET_SORT) | (1ull << K_SWEEP_WAVE) |
      (1ull << K_SWEEP) | (1ull << K_FINISH) | (1ull << K_SNAPSHOT) | (1ull << K_ZZ_SET) | (1ull << K_ZZMAX) |
      (1ull << K_MAP_CLEAR) | (1ull << K_MAP_STAMP) | (1ull << K_MAP_HALO) | (1ull << K_EXT_STIM);
  if (!c->tu.step_graph || (g_prof.mask & seg)) {
    *joined = false;
    HIPCHK(c, issue_build(c, mn, nb, izb, false));
    return EPI_OK;
  }
  *joined = true;   // a captured graph rejoins its forked stream
  int mcap = 1;
  for (auto& h : c->th) mcap = std::max(mcap, h.cap);
  const int mr = std::max(mn, std::min((mn + 1023) / 1024 * 1024, mcap));
  const Tuning& tu = c->tu;
  const std::vector<long long> key = {1, c->pb, c->cb, c->cp, c->bp, nb, mr, izb, max_gsize(c), max_halo(c),
                                      max_nx(c), max_ny(c), tu.b_zb, tu.b_rows, tu.map_lanes, tu.sweep_wave, tu.side_fork};
  return graph_run(c, key, [&] { return issue_build(c, mr, nb, izb); });
}

int epi_step(epi_ctx* c, int64_t nsteps, epi_driver* drv) {
  if (!c || !drv) return EPI_ERR_INVALID;
  // current_timestep (proc.cpp:132) of each tissue's renewals: the driver's
  // step; the device advances it once per step from here
  for (int t = 0; t < c->T; t++)
    if (c->dev_ts[t] != drv[t].step) {
      c->h_st[t].timestep = drv[t].step;
      HIPCHK(c, hipMemcpyAsync(&c->d_st[t].timestep, &c->h_st[t].timestep, sizeof(long long), hipMemcpyHostToDevice,
                               c->stream));
      c->dev_ts[t] = drv[t].step;
    }
  // The next step's cell_interaction reads positions, radii, ages, states,
  // pairs, dermis and the lists; a calc_ca2p solve writes none of them (and
  // reads no position). So, unless initialize_sc (which changes states)
  // fires, the next step's force kernel is issued before this step's status
  // read-back, over a grid that covers every cell a tissue can hold, and runs
  // while the host waits and decides on the solve; the solve follows it.
  bool forces_issued = false;
  int forces_cap = 0, forces_pb = 0;
  for (int64_t it = 0; it < nsteps; it++) {
    const int mn = max_n(c);
    // cell_dynamics (proc.cpp:22-32)
    if (!forces_issued) {
      launch_forces(c->stream, c->d_tis, c->T, c->pb, c->cb, max_memb(c), mn, max_list_len(c));
      c->pb ^= 1;
    }
    forces_issued = false;
    const int nb = 1 - c->cb;
    bool joined = true;
    RCHK(run_build(c, mn, nb, izb_est(c, false), &joined));
    c->len_known = false;
    c->cb = nb;
    c->bp ^= 1;
    bool sc_due = false;   // proc.cpp:158-163, decided from host state alone
    for (int t = 0; t < c->T; t++)
      if (sc_turnover(c->tis[t].prm, drv[t])) sc_due = true;
    if (it + 1 < nsteps && !sc_due) {
      // the rows are those of this step's build, whose longest list is known
      // only after the read-back: the slots follow the last known maxlen,
      // and rows beyond them go to k_forces_long once it is known
      int cap_n = 1, known_len = 0;
      for (int t = 0; t < c->T; t++) {
        cap_n = std::max(cap_n, c->th[t].cap);
        known_len = std::max(known_len, c->h_st[t].maxlen);
      }
      // the read-back follows this step's main chain only: the maps and the
      // B sweep (side stream) run on beside the next forces, which read
      // neither, and what the main stream issues after those forces (the
      // next renewal, a solve) waits for them.  Their zzmax reaches the host
      // with a later read-back; a solve reads it again first.
      RCHK(sync_status_begin(c));
      forces_pb = c->pb;
      forces_cap = launch_forces_rows(c->stream, c->d_tis, c->T, c->pb, c->cb, cap_n, known_len);
      c->pb ^= 1;
      forces_issued = true;
      if (!joined) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
      HIPCHK(c, hipGetLastError());
      RCHK(sync_status_end(c));
    } else {
      if (!joined) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
      HIPCHK(c, hipGetLastError());
      RCHK(sync_status(c));
    }
    if (forces_issued && max_list_len(c) > forces_cap)
      launch_forces_long(c->stream, c->d_tis, c->T, forces_pb, c->cb, max_n(c), forces_cap);
    int nsc = 0, nca = 0;
    std::vector<int> ca_list;
    for (int t = 0; t < c->T; t++) {
      const Tis& ti = c->tis[t];
      c->steps[t]++;
      c->cell_steps[t] += c->h_st[t].n;
      epi_driver& d = drv[t];
      // proc.cpp:158-163
      if (sc_turnover(ti.prm, d)) {
        d.force_cornif = 0;
        c->h_tlist[nsc++] = t;
        d.num_sc = (int32_t)ti.prm.NUM_SC_INIT;
      }
    }
    if (nsc) {
      HIPCHK(c, hipMemcpyAsync(c->d_tlist, c->h_tlist, sizeof(int) * nsc, hipMemcpyHostToDevice, c->stream));
      launch_initialize_sc(c->stream, c->d_tis, c->d_tlist, nsc, c->pb, mn);
    }
    for (int t = 0; t < c->T; t++) {
      epi_driver& d = drv[t];
      const bool forced = d.ca_every > 0 && ((d.step + 1) % d.ca_every) == 0;
      const Tis& ti = c->tis[t];
      if ((c->h_st[t].sw & 255) >= ti.prm.SW_THRESH || d.num_sc > 0 || forced) {
        ca_list.push_back(t);
        if (d.num_sc > 0) d.num_sc--;
        c->ca_solves[t]++;
        c->grid_ca[t] += (long long)ti.prm.Ca_ITR * ti.NX * ti.NY *
                         (int)((c->h_st[t].zzmax + ti.prm.iz_margin) / ti.dz);
      }
      d.step++;
      c->dev_ts[t]++;
    }
    nca = (int)ca_list.size();
    if (nca) {
      // the solve's work lists need this step's zzmax, which an early
      // read-back may have missed
      if (forces_issued) RCHK(sync_status(c));
      if (nsc) HIPCHK(c, hipStreamSynchronize(c->stream));
      for (int i = 0; i < nca; i++) c->h_tlist[i] = ca_list[i];
      HIPCHK(c, hipMemcpyAsync(c->d_tlist, c->h_tlist, sizeof(int) * nca, hipMemcpyHostToDevice, c->stream));
      RCHK(run_ca(c, c->d_tlist, nca));
      RCHK(sync_status(c));
    }
  }
  return EPI_OK;
}

int epi_synchronize(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  return sync_status(c);
}

int epi_set_params(epi_ctx* c, int32_t t, const epi_params* p) {
  RCHK(check_tissue(c, t));
  if (!p) return fail(c, EPI_ERR_INVALID, "null parameters");
  if (c->rd.active || c->mech.active) return fail(c, EPI_ERR_INVALID, "parameters are fixed while a window is open");
  if (const char* why = params_refusal(*p, c->tis[t].prm)) return fail(c, EPI_ERR_INVALID, why);
  Tis& ti = c->tis[t];
  derive_params(*p, ti.prm);
  ti.prm.outputdir[sizeof(ti.prm.outputdir) - 1] = 0;
  window_values_from_params(ti);
  HIPCHK(c, hipMemcpy(c->d_tis + t, &ti, sizeof(Tis), hipMemcpyHostToDevice));
  return EPI_OK;
}

int epi_get_params(epi_ctx* c, int32_t t, epi_params* p) {
  RCHK(check_tissue(c, t));
  if (!p) return fail(c, EPI_ERR_INVALID, "null parameters");
  *p = static_cast<const epi_params&>(c->tis[t].prm);
  return EPI_OK;
}

int epi_get_status(epi_ctx* c, int32_t t, epi_status* out) {
  RCHK(check_tissue(c, t));
  int rc = sync_status(c);
  const Status& s = c->h_st[t];
  out->n = s.n; out->err = s.err; out->sw = s.sw & 255; out->ndiv = s.ndiv; out->nrm = s.nrm;
  out->maxlen = s.maxlen; out->rng_pos = s.rng_pos; out->next_uid = s.next_uid;
  unsigned long long b = s.zzmax_bits;
  memcpy(&out->zzmax, &b, 8);
  out->ca_solves = c->ca_solves[t];
  out->steps = c->steps[t];
  out->cell_steps = c->cell_steps[t];
  out->grid_points_b = s.grid_b;
  out->grid_points_ca = c->grid_ca[t];
  return rc;
}

int epi_profile_enable(epi_ctx* c, uint64_t mask) {
  if (!c) return EPI_ERR_INVALID;
  RCHK(sync_status(c));
  g_prof.mask = mask;
  return EPI_OK;
}

int epi_profile_sample(epi_ctx* c, int32_t period, int32_t run) {
  if (!c) return EPI_ERR_INVALID;
  if (period < 1 || (period > 1 && (run < 2 || run > period)))
    return fail(c, EPI_ERR_INVALID, "profile sampling needs period >= 1 and 2 <= run <= period");
  RCHK(sync_status(c));
  g_prof.period = period;
  g_prof.run = period > 1 ? run : 1;
  g_prof.launches = 0;
  return EPI_OK;
}

int epi_profile_read(epi_ctx* c, int32_t kernel, const char** name, double* total_ms, int64_t* count) {
  if (!c) return EPI_ERR_INVALID;
  if (kernel < 0 || kernel >= K_COUNT) return fail(c, EPI_ERR_INVALID, "kernel id");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  prof_resolve();
  if (name) *name = kKernelNames[kernel];
  if (total_ms) *total_ms = g_prof.ms[kernel];
  if (count) *count = g_prof.cnt[kernel];
  return EPI_OK;
}

int epi_profile_launches(epi_ctx* c, int32_t kernel, int64_t* launches) {
  if (!c || !launches) return EPI_ERR_INVALID;
  if (kernel < 0 || kernel >= K_COUNT) return fail(c, EPI_ERR_INVALID, "kernel id");
  *launches = g_prof.launched[kernel];
  return EPI_OK;
}

int epi_profile_reset(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  prof_resolve();
  for (int k = 0; k < K_COUNT; k++) { g_prof.ms[k] = 0; g_prof.cnt[k] = 0; g_prof.launched[k] = 0; }
  return EPI_OK;
}

int32_t epi_profile_kernel_count(void) { return K_COUNT; }

namespace {
int* tuning_slot(Tuning& tu, int key) {
  switch (key) {
    case EPI_TUNE_B_ZB: return &tu.b_zb;
    case EPI_TUNE_B_ROWS: return &tu.b_rows;
    case EPI_TUNE_ATP_ZB: return &tu.a_zb;
    case EPI_TUNE_ATP_ROWS: return &tu.a_rows;
    case EPI_TUNE_CA_GROUP: return &tu.ca_group;
    case EPI_TUNE_XCD: return &tu.xcd;
    case EPI_TUNE_RD_GROUP: return &tu.rd_group;
    case EPI_TUNE_CA_FUSE: return &tu.ca_fuse;
    case EPI_TUNE_MAP_LANES: return &tu.map_lanes;
    case EPI_TUNE_CA_GROUP_SMALL: return &tu.ca_group_small;
    case EPI_TUNE_STEP_GRAPH: return &tu.step_graph;
    case EPI_TUNE_B_XCD: return &tu.b_xcd;
    case EPI_TUNE_SWEEP_WAVE: return &tu.sweep_wave;
    case EPI_TUNE_SIDE_FORK: return &tu.side_fork;
    default: return nullptr;
  }
}
}  // namespace

int epi_set_tuning(epi_ctx* c, int32_t key, int32_t value) {
  if (!c) return EPI_ERR_INVALID;
  int* slot = tuning_slot(c->tu, key);
  if (!slot) return fail(c, EPI_ERR_INVALID, "unknown tuning key");
  bool ok = false;
  switch (key) {
    case EPI_TUNE_B_ZB: case EPI_TUNE_ATP_ZB: ok = value == 1 || value == 2 || value == 4 || value == 8 || value == 16; break;
    case EPI_TUNE_B_ROWS: case EPI_TUNE_ATP_ROWS: ok = value == 1 || value == 2 || value == 4; break;
    case EPI_TUNE_CA_GROUP: ok = value == 4 || value == 8 || value == 16; break;
    case EPI_TUNE_XCD: ok = value == 0 || value == 1; break;
    case EPI_TUNE_RD_GROUP: ok = value == 8 || value == 16; break;
    case EPI_TUNE_CA_FUSE: ok = value >= 0 && value <= 2; break;
    case EPI_TUNE_MAP_LANES: ok = value == 1 || value == 4 || value == 16; break;
    case EPI_TUNE_CA_GROUP_SMALL: ok = value == 4 || value == 8 || value == 16; break;
    case EPI_TUNE_STEP_GRAPH: ok = value == 0 || value == 1; break;
    case EPI_TUNE_B_XCD: ok = value == 0 || value == 1; break;
    case EPI_TUNE_SWEEP_WAVE: ok = value == 0 || value == 1; break;
    case EPI_TUNE_SIDE_FORK: ok = value == 0 || value == 1; break;
  }
  if (!ok) return fail(c, EPI_ERR_INVALID, "tuning value out of range");
  *slot = value;
  return EPI_OK;
}

int epi_get_tuning(epi_ctx* c, int32_t key, int32_t* value) {
  if (!c || !value) return EPI_ERR_INVALID;
  int* slot = tuning_slot(c->tu, key);
  if (!slot) return fail(c, EPI_ERR_INVALID, "unknown tuning key");
  *value = *slot;
  return EPI_OK;
}

int epi_get_cells(epi_ctx* c, int32_t t, epi_cells* o) {
  RCHK(check_tissue(c, t));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->h_st, c->d_st, sizeof(Status) * c->T, hipMemcpyDeviceToHost));
  const Tis& ti = c->tis[t];
  const TH& h = c->th[t];
  const int n = c->h_st[t].n;
  o->n = n; o->nmemb = h.nmemb; o->nder = h.nder; o->sw = c->h_st[t].sw & 255;
  const int p1 = c->pos_pending ? 1 - c->pb : c->pb;
  RCHK(download(c, o->state, ti.state, n));
  RCHK(download(c, o->fix_origin, ti.fix_origin, n));
  RCHK(download(c, o->x, ti.pos[c->pb][0], n));
  RCHK(download(c, o->y, ti.pos[c->pb][1], n));
  RCHK(download(c, o->z, ti.pos[c->pb][2], n));
  RCHK(download(c, o->xn, ti.pos[p1][0], n));
  RCHK(download(c, o->yn, ti.pos[p1][1], n));
  RCHK(download(c, o->zn, ti.pos[p1][2], n));
  RCHK(download(c, o->radius, ti.radius, n));
  RCHK(download(c, o->ca2p, ti.ca2p[c->cp], n));
  RCHK(download(c, o->ca2p_next, ti.ca2p[1 - c->cp], n));
  RCHK(download(c, o->ip3, ti.ip3[c->cp], n));
  RCHK(download(c, o->ip3_next, ti.ip3[1 - c->cp], n));
  RCHK(download(c, o->ca2p_avg, ti.ca2p_avg, n));
  RCHK(download(c, o->ex_inert, ti.ex_inert, n));
  RCHK(download(c, o->agek, ti.agek, n));
  RCHK(download(c, o->ageb, ti.ageb, n));
  RCHK(download(c, o->ex_fat, ti.ex_fat, n));
  RCHK(download(c, o->in_fat, ti.in_fat, n));
  RCHK(download(c, o->spr_nat_len, ti.spr_nat_len, n));
  RCHK(download(c, o->diff_u, ti.diff_u, n));
  RCHK(download(c, o->rest_div_times, ti.rest_div_times, n));
  RCHK(download(c, o->pair, ti.pair, n));
  RCHK(download(c, o->dermis, ti.dermis, n));
  RCHK(download(c, o->lat, ti.lat, 3 * (size_t)n));
  RCHK(download(c, o->is_malignant, ti.is_malignant, n));
  if (o->uid) {
    std::vector<long long> u(n);
    RCHK(download(c, u.data(), ti.uid, n));
    for (int i = 0; i < n; i++) o->uid[i] = u[i];
  }
  if (o->memb_adj) RCHK(download(c, o->memb_adj, ti.memb6, 6 * (size_t)h.nmemb));
  if (o->cst) {
    std::vector<long long> v(3 * (size_t)n);
    RCHK(download(c, v.data(), ti.cst, v.size()));
    for (size_t i = 0; i < v.size(); i++) o->cst[i] = v[i];
  }
  if (o->conn_ptr || o->conn || o->gj) {
    std::vector<int> nl(n);
    RCHK(download(c, nl.data(), ti.nlen[c->cb], n));
    long long nnz = 0;
    for (int i = 0; i < n; i++) nnz += nl[i];
    if (o->conn_ptr) {
      int acc = 0;
      for (int i = 0; i < n; i++) { o->conn_ptr[i] = acc; acc += nl[i]; }
      o->conn_ptr[n] = acc;
    }
    if (o->conn || o->gj) {
      if (nnz > o->conn_cap) return fail(c, EPI_ERR_CAPACITY, "conn_cap too small");
      std::vector<int> col((size_t)n * h.ns);
      std::vector<double> g((size_t)n * h.ns);
      RCHK(download(c, col.data(), ti.col[c->cb], col.size()));
      RCHK(download(c, g.data(), ti.gj[c->cb], g.size()));
      size_t p = 0;
      for (int i = 0; i < n; i++)
        for (int e = 0; e < nl[i]; e++, p++) {
          if (o->conn) o->conn[p] = col[(size_t)i * h.ns + e];
          if (o->gj) o->gj[p] = g[(size_t)i * h.ns + e];
        }
    }
  }
  return EPI_OK;
}

int epi_snapshot_begin(epi_ctx* c, int32_t t, int32_t slot) {
  RCHK(check_tissue(c, t));
  if (slot < 0 || slot >= EPI_SNAPSHOT_SLOTS) return fail(c, EPI_ERR_INVALID, "snapshot slot out of range");
  if (c->rd.active || c->mech.active) return fail(c, EPI_ERR_INVALID, "no snapshot while a window is active");
  auto& S = c->snap[slot];
  if (S.pending) {   // the slot's last download still reads its buffers
    HIPCHK(c, hipEventSynchronize(S.ev_done));
    S.pending = false;
  }
  S.landed = false;
  const int n = c->h_st[t].n;   // current after epi_step's status read-back
  const long long m = std::max(32LL, ((long long)n + 31) / 32 * 32);
  const size_t bytes = (size_t)out_record_bytes(m);
  if (bytes > S.bytes) {
    if (S.dev) HIPCHK(c, hipFree(S.dev));
    if (S.host) HIPCHK(c, hipHostFree(S.host));
    S.dev = nullptr;
    S.host = nullptr;
    S.bytes = 0;
    HIPCHK(c, hipMalloc(&S.dev, bytes));
    HIPCHK(c, hipHostMalloc(reinterpret_cast<void**>(&S.host), bytes, hipHostMallocDefault));
    S.bytes = bytes;
  }
  if (!S.ev_pack) HIPCHK(c, hipEventCreateWithFlags(&S.ev_pack, hipEventDisableTiming));
  if (!S.ev_done) HIPCHK(c, hipEventCreateWithFlags(&S.ev_done, hipEventDisableTiming));
  if (!c->copy_stream) HIPCHK(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  launch_out_pack(c->stream, c->d_tis, t, c->pb, c->cp, c->cb, n, m, S.dev);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(S.ev_pack, c->stream));
  HIPCHK(c, hipStreamWaitEvent(c->copy_stream, S.ev_pack, 0));
  HIPCHK(c, hipMemcpyAsync(S.host, S.dev, (size_t)out_record_bytes(m), hipMemcpyDeviceToHost, c->copy_stream));
  HIPCHK(c, hipEventRecord(S.ev_done, c->copy_stream));
  S.tissue = t;
  S.n = n;
  S.m = m;
  S.pending = true;
  return EPI_OK;
}

int epi_snapshot_wait(epi_ctx* c, int32_t slot, epi_snapshot* o) {
  if (!c || !o) return EPI_ERR_INVALID;
  if (slot < 0 || slot >= EPI_SNAPSHOT_SLOTS) return fail(c, EPI_ERR_INVALID, "snapshot slot out of range");
  auto& S = c->snap[slot];
  if (!S.pending && !S.landed) return fail(c, EPI_ERR_INVALID, "no snapshot in this slot");
  if (S.pending) {
    HIPCHK(c, hipEventSynchronize(S.ev_done));
    S.pending = false;
    S.landed = true;
  }
  if (*reinterpret_cast<const int*>(S.host) != S.n)
    return fail(c, EPI_ERR_INVALID, "the cell count changed after the host last read it: take snapshots between epi_step calls");
  const long long m = S.m;
  const char* base = S.host + OUT_HDR;
  const double* f = reinterpret_cast<const double*>(base);
  const int* iv = reinterpret_cast<const int*>(base + 8 * OUT_NF64 * m);
  const uint8_t* bv = reinterpret_cast<const uint8_t*>(base + (8 * OUT_NF64 + 4 * OUT_NI32) * m);
  o->tissue = S.tissue;
  o->n = S.n;
  o->radius = f; o->ageb = f + m; o->agek = f + 2 * m; o->ca2p = f + 3 * m;
  o->x = f + 4 * m; o->y = f + 5 * m; o->z = f + 6 * m; o->ca2p_avg = f + 7 * m;
  o->ex_fat = f + 8 * m; o->in_fat = f + 9 * m; o->spr_nat_len = f + 10 * m;
  o->rest_div_times = iv; o->pair = iv + m; o->fix_origin = iv + 2 * m;
  o->state = bv; o->touch = bv + m;
  return EPI_OK;
}

int epi_get_fields(epi_ctx* c, int32_t t, double* atp, double* atp_next, double* b, double* b_next, int32_t* cmap1,
                   uint8_t* cmap2) {
  RCHK(check_tissue(c, t));
  const Tis& ti = c->tis[t];
  const TH& h = c->th[t];
  std::vector<double> tmp(h.gsize);
  double* dsts[4] = {atp, atp_next, b, b_next};
  const double* srcs[4] = {ti.atp[c->ap], ti.atp[1 - c->ap], ti.bf[c->bp], ti.bf[1 - c->bp]};
  for (int i = 0; i < 4; i++) {
    if (!dsts[i]) continue;
    RCHK(download(c, tmp.data(), srcs[i], tmp.size()));
    from_device_layout(h, tmp, dsts[i]);
  }
  if (cmap1) {
    std::vector<int> m(h.gsize);
    RCHK(download(c, m.data(), ti.cmap1, m.size()));
    from_device_layout(h, m, cmap1);
  }
  if (cmap2) {
    std::vector<uint8_t> m(h.gsize);
    RCHK(download(c, m.data(), ti.cmap2, m.size()));
    from_device_layout(h, m, cmap2);
  }
  return EPI_OK;
}

int epi_set_positions(epi_ctx* c, int32_t t, const double* x, const double* y, const double* z) {
  RCHK(check_tissue(c, t));
  RCHK(sync_status(c));
  const Tis& ti = c->tis[t];
  const int n = c->h_st[t].n;
  for (int b = 0; b < 2; b++) {
    RCHK(upload(c, ti.pos[b][0], x, n));
    RCHK(upload(c, ti.pos[b][1], y, n));
    RCHK(upload(c, ti.pos[b][2], z, n));
  }
  c->pos_pending = false;
  return EPI_OK;
}

int epi_set_state(epi_ctx* c, int32_t t, const uint8_t* state, const double* agek) {
  RCHK(check_tissue(c, t));
  RCHK(sync_status(c));
  const Tis& ti = c->tis[t];
  const int n = c->h_st[t].n;
  RCHK(upload(c, ti.state, state, n));
  RCHK(upload(c, ti.agek, agek, n));
  return EPI_OK;
}

int epi_set_lists(epi_ctx* c, int32_t t, const int32_t* ptr, const int32_t* conn, const double* gj) {
  RCHK(check_tissue(c, t));
  RCHK(sync_status(c));
  const Tis& ti = c->tis[t];
  const TH& h = c->th[t];
  const int n = c->h_st[t].n;
  std::vector<int> col((size_t)n * h.ns, 0), nl(n), nlo(n, 0), srow(n);
  std::vector<double> g((size_t)n * h.ns, c->tis[t].prm.gj_init);
  std::vector<long long> u(n);
  RCHK(download(c, u.data(), ti.uid, n));
  for (int i = 0; i < n; i++) {
    const int len = ptr[i + 1] - ptr[i];
    if (len > h.ns || len >= MAX_CONN) return fail(c, EPI_ERR_CAPACITY, "list too long");
    nl[i] = len;
    srow[i] = i;
    int k = i < h.nmemb ? MEMB_ADJ : 0;
    int lo = 0;
    for (int e = 0; e < len; e++) {
      const int q = conn[ptr[i] + e];
      col[(size_t)i * h.ns + e] = q;
      if (gj) g[(size_t)i * h.ns + e] = gj[ptr[i] + e];
    }
    if (i >= h.nmemb) {
      while (lo < len && col[(size_t)i * h.ns + lo] < i) lo++;
      k = lo;
    }
    for (int e = k; e < len; e++)
      if (col[(size_t)i * h.ns + e] <= i || (e > k && col[(size_t)i * h.ns + e] < col[(size_t)i * h.ns + e - 1]))
        return fail(c, EPI_ERR_INVALID, "list is not in canonical order");
    nlo[i] = lo;
  }
  RCHK(upload(c, ti.col[c->cb], col.data(), col.size()));
  RCHK(upload(c, ti.gj[c->cb], g.data(), g.size()));
  RCHK(upload(c, ti.nlen[c->cb], nl.data(), n));
  RCHK(upload(c, ti.nlo, nlo.data(), n));
  RCHK(upload(c, ti.src_row, srow.data(), n));
  RCHK(upload(c, ti.old_uid, u.data(), n));
  return EPI_OK;
}

int epi_set_maps(epi_ctx* c, int32_t t, const int32_t* cmap1, const uint8_t* cmap2, const int32_t* lat) {
  RCHK(check_tissue(c, t));
  RCHK(sync_status(c));
  const Tis& ti = c->tis[t];
  const TH& h = c->th[t];
  if (cmap1) {
    std::vector<int> m(h.gsize, -1);
    for (int j = 0; j <= h.NX; j++)
      for (int k = 0; k <= h.NY; k++)
        for (int l = 0; l <= h.NZ; l++) m[dev_index(h, j, k, l)] = cmap1[((size_t)j * (h.NY + 1) + k) * (h.NZ + 1) + l];
    RCHK(upload(c, ti.cmap1, m.data(), m.size()));
  }
  if (cmap2) {
    std::vector<uint8_t> m(h.gsize, 0);
    for (int j = 0; j <= h.NX; j++)
      for (int k = 0; k <= h.NY; k++)
        for (int l = 0; l <= h.NZ; l++) m[dev_index(h, j, k, l)] = cmap2[((size_t)j * (h.NY + 1) + k) * (h.NZ + 1) + l];
    RCHK(upload(c, ti.cmap2, m.data(), m.size()));
  }
  if (lat) RCHK(upload(c, ti.lat, lat, 3 * (size_t)c->h_st[t].n));
  return EPI_OK;
}

}  // extern "C"

// ---- frozen reaction-diffusion window and its adjoint ----------------------
namespace {

// threads for the window's copy kernels: 16 lanes per cell for the rows,
// one per cell, one per grid point
long long rd_items(const epi_ctx* c) {
  long long m = 1;
  for (int t = 0; t < c->T; t++) m = std::max({m, (long long)(c->th[t].n_est + 64) * 16, c->th[t].gsize});
  return m;
}

long long rd_max_box(const epi_ctx* c) {
  long long m = 1;
  for (auto& r : c->rd.h) m = std::max(m, r.box);
  return m;
}

// du/dp into the device descriptors, kpa (or its snapshot) and gj0 staged
int rd_params(epi_ctx* c, const epi_rd_params* prm) {
  std::vector<double> g0(c->T);
  for (int t = 0; t < c->T; t++) {
    c->tis[t].ca_du = prm[t].du;
    c->tis[t].ca_dp = prm[t].dp;
    g0[t] = prm[t].gj0;
    const RdTis& r = c->rd.h[t];
    const size_t n = (size_t)c->h_st[t].n;
    if (prm[t].kpa)
      HIPCHK(c, hipMemcpyAsync(r.kpa_in, prm[t].kpa, 8 * n, hipMemcpyHostToDevice, c->stream));
    else
      HIPCHK(c, hipMemcpyAsync(r.kpa_in, r.kpa0, 8 * n, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(c->d_tis, c->tis.data(), sizeof(Tis) * c->T, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->rd.d_gj0, g0.data(), 8 * c->T, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return EPI_OK;
}

// one window sub-step; with sv the state before it is saved into segment
// slot `slot` by the cell kernel (cells_only: only that, no ATP update)
int rd_substep(epi_ctx* c, const RdTis* sv = nullptr, int slot = 0, bool cells_only = false) {
  CaList cl;
  RCHK(ca_list(c, nullptr, c->T, c->tu.ca_group, ca_list_bs(c->tu, false), c->tu.a_zb, &cl));
  launch_ca_substep(c->stream, c->d_tis, cl, c->cb, c->cp, c->ap, c->bp, c->tu, sv, slot, cells_only);
  c->cp ^= 1;
  c->ap ^= 1;
  return EPI_OK;
}

}  // namespace

int epi_rd_begin(epi_ctx* c, const double* zzmax) {
  if (!c || !zzmax) return EPI_ERR_INVALID;
  if (c->rd.active) return fail(c, EPI_ERR_INVALID, "a reaction-diffusion window is already open");
  RCHK(sync_status(c));
  auto& rd = c->rd;
  rd.izb.assign(c->T, 0);
  rd.zz.assign(zzmax, zzmax + c->T);
  for (int t = 0; t < c->T; t++) {
    rd.izb[t] = (int)((zzmax[t] + c->tis[t].prm.iz_margin) / c->tis[t].dz);   // ca2p.cpp:386
    if (rd.izb[t] < 0 || rd.izb[t] > c->tis[t].NZ) return fail(c, EPI_ERR_OUT_OF_GRID, "iz_bound outside the grid");
    // the reverse pass gathers a cell's voxels from its stamp box, which must
    // not wrap onto itself in x or y
    if (c->tis[t].NX < 4 * c->tis[t].irx + 1 || c->tis[t].NY < 4 * c->tis[t].iry + 1)
      return fail(c, EPI_ERR_INVALID, "grid too small for a window (stamp box wraps onto itself)");
  }
  // every AFM cell's grid_avg8 corners must lie in planes 0..iz_bound (the
  // planes a saved state holds)
  for (int t = 0; t < c->T; t++) {
    const int n = c->h_st[t].n;
    std::vector<uint8_t> st(n);
    std::vector<int> lat(3 * (size_t)n);
    HIPCHK(c, hipMemcpy(st.data(), c->tis[t].state, n, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(lat.data(), c->tis[t].lat, 12 * (size_t)n, hipMemcpyDeviceToHost));
    for (int i = c->tis[t].nmemb + c->tis[t].nder; i < n; i++)
      if ((st[i] == ALIVE || st[i] == FIX || st[i] == MUSUME) && lat[3 * i + 2] + 1 > rd.izb[t])
        return fail(c, EPI_ERR_INVALID, "zzmax lies below a cell of the window");
  }
  RCHK(set_zz(c, zzmax));
  // per-solve maps: entry kinds, th/Kpa/IAG, voxel owners (init_ca2p_map)
  long long box = 1;
  for (int t = 0; t < c->T; t++) box = std::max(box, (long long)c->th[t].NX * c->th[t].NY * (rd.izb[t] + 1));
  launch_ca_touch(c->stream, c->d_tis, nullptr, c->T, c->cb, max_n(c), box);
  // arena: snapshot + adjoint arrays per tissue
  rd.h.assign(c->T, RdTis{});
  Carver dry{nullptr, 0, true};
  auto carve = [&](Carver& cv) {
    for (int t = 0; t < c->T; t++) {
      RdTis& r = rd.h[t];
      const size_t cap = c->th[t].cap, ns = c->th[t].ns, gs = c->th[t].gsize;
      r.izb = rd.izb[t];
      r.box = (long long)(rd.izb[t] + 1) * c->th[t].PL;
      for (int b = 0; b < 2; b++) {
        r.u0[b] = cv.take<double>(cap); r.p0[b] = cv.take<double>(cap); r.A0[b] = cv.take<double>(gs);
        r.lu[b] = cv.take<double>(cap); r.lp[b] = cv.take<double>(cap); r.lh[b] = cv.take<double>(cap);
        r.lg[b] = cv.take<double>(cap * ns);
      }
      r.h0 = cv.take<double>(cap); r.avg0 = cv.take<double>(cap); r.kpa0 = cv.take<double>(cap);
      r.g0 = cv.take<double>(cap * ns);
      for (int b = 0; b < 3; b++) r.lA[b] = cv.take<double>((size_t)r.box);
      r.gkpa = cv.take<double>(cap); r.target = cv.take<double>(cap);
      r.kpa_in = cv.take<double>(cap); r.scal = cv.take<double>(4);
      const Tis& ti = c->tis[t];
      r.oc = (4 * ti.irx + 1) * (4 * ti.iry + 1) * (4 * ti.irz + 1);
      r.own = cv.take<int>(cap * (size_t)r.oc); r.own_n = cv.take<int>(cap);
      r.nafm = 0;
    }
  };
  carve(dry);
  if (hipMalloc(&rd.arena, align_up(dry.off, 256)) != hipSuccess ||
      hipMalloc(&rd.d, sizeof(RdTis) * c->T) != hipSuccess || hipMalloc(&rd.d_gj0, 8 * c->T) != hipSuccess)
    return fail(c, EPI_ERR_HIP, "window allocation failed");
  Carver real{rd.arena, 0, false};
  carve(real);
  HIPCHK(c, hipMemcpyAsync(rd.d, rd.h.data(), sizeof(RdTis) * c->T, hipMemcpyHostToDevice, c->stream));
  launch_rd_snapshot(c->stream, c->d_tis, rd.d, c->T, c->cb, c->cp, c->ap, rd_items(c));
  launch_rd_own(c->stream, c->d_tis, rd.d, c->T, max_n(c));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(rd.h.data(), rd.d, sizeof(RdTis) * c->T, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  rd.active = true;
  return EPI_OK;
}

int epi_rd_forward(epi_ctx* c, const epi_rd_params* prm, int32_t n_sub) {
  if (!c || !prm || n_sub < 0) return EPI_ERR_INVALID;
  if (!c->rd.active) return fail(c, EPI_ERR_INVALID, "no reaction-diffusion window is open");
  RCHK(rd_params(c, prm));
  launch_rd_restore(c->stream, c->d_tis, c->rd.d, c->rd.d_gj0, c->T, c->cb, c->cp, c->ap, 1, rd_items(c));
  for (int s = 0; s < n_sub; s++) RCHK(rd_substep(c));
  HIPCHK(c, hipGetLastError());
  return sync_status(c);
}

int epi_rd_get_ca2p(epi_ctx* c, int32_t t, double* out) {
  RCHK(check_tissue(c, t));
  if (!out) return EPI_ERR_INVALID;
  HIPCHK(c, hipMemcpy(out, c->tis[t].ca2p[c->cp], 8 * (size_t)c->h_st[t].n, hipMemcpyDeviceToHost));
  return EPI_OK;
}

int epi_rd_adjoint(epi_ctx* c, const epi_rd_params* prm, int32_t n_sub, int32_t every, const double* const* targets,
                   double* loss, epi_rd_grad* grads) {
  if (!c || !prm || !targets || n_sub < 1 || every < 1) return EPI_ERR_INVALID;
  auto& rd = c->rd;
  if (!rd.active) return fail(c, EPI_ERR_INVALID, "no reaction-diffusion window is open");
  const int K = std::min(every, n_sub);
  const int nck = (n_sub + K - 1) / K;
  // segment slots (K states) and checkpoints (nck, both buffers)
  size_t need = 0;
  {
    Carver dry{nullptr, 0, true};
    for (int t = 0; t < c->T; t++) {
      const size_t cap = c->th[t].cap, ns = c->th[t].ns, box = (size_t)rd.h[t].box;
      dry.take<double>(K * cap); dry.take<double>(K * cap); dry.take<double>(K * cap);
      dry.take<double>(K * cap * ns); dry.take<double>(K * box);
      dry.take<double>(2 * nck * cap); dry.take<double>(2 * nck * cap); dry.take<double>(nck * cap);
      dry.take<double>(nck * cap * ns); dry.take<double>(2 * nck * box);
    }
    need = align_up(dry.off, 256);
  }
  if (need > rd.work_bytes) {
    if (rd.work) (void)hipFree(rd.work);
    rd.work = nullptr;
    rd.work_bytes = 0;
    if (hipMalloc(&rd.work, need) != hipSuccess) return fail(c, EPI_ERR_HIP, "window work allocation failed");
    rd.work_bytes = need;
  }
  Carver cv{rd.work, 0, false};
  for (int t = 0; t < c->T; t++) {
    RdTis& r = rd.h[t];
    const size_t cap = c->th[t].cap, ns = c->th[t].ns, box = (size_t)r.box;
    r.su = cv.take<double>(K * cap); r.sp = cv.take<double>(K * cap); r.sh = cv.take<double>(K * cap);
    r.sg = cv.take<double>(K * cap * ns); r.sAv = cv.take<double>(K * cap);
    r.cu = cv.take<double>(2 * nck * cap); r.cp_ = cv.take<double>(2 * nck * cap); r.ch = cv.take<double>(nck * cap);
    r.cg = cv.take<double>(nck * cap * ns); r.cA = cv.take<double>(2 * nck * box);
    HIPCHK(c, hipMemcpyAsync(r.target, targets[t], 8 * (size_t)c->h_st[t].n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(r.scal, 0, 4 * sizeof(double), c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(rd.d, rd.h.data(), sizeof(RdTis) * c->T, hipMemcpyHostToDevice, c->stream));
  RCHK(rd_params(c, prm));
  const long long items = rd_items(c);
  launch_rd_restore(c->stream, c->d_tis, rd.d, rd.d_gj0, c->T, c->cb, c->cp, c->ap, 1, items);
  // forward, checkpoint every K sub-steps; the last segment saves its states
  // straight into the slots, so the reverse pass starts without a recompute
  std::vector<int> ck_cp(nck), ck_ap(nck);
  const int last0 = (nck - 1) * K;
  for (int s = 0; s < n_sub; s++) {
    if (s >= last0) {
      RCHK(rd_substep(c, rd.d, s - last0));
      continue;
    }
    if (s % K == 0) {
      ck_cp[s / K] = c->cp;
      ck_ap[s / K] = c->ap;
      launch_rd_state(c->stream, c->d_tis, rd.d, c->T, s / K, 2, c->cb, c->cp, c->ap, items);
    }
    RCHK(rd_substep(c));
  }
  int q = 0, qa = 0;
  launch_rd_seed(c->stream, c->d_tis, rd.d, c->T, c->cp, c->cb, q, items);
  const int mn = max_n(c);
  const long long mb = rd_max_box(c);
  for (int k = nck - 1; k >= 0; k--) {
    const int s0 = k * K, s1 = std::min(n_sub, s0 + K);
    if (k < nck - 1) {
      c->cp = ck_cp[k];
      c->ap = ck_ap[k];
      launch_rd_state(c->stream, c->d_tis, rd.d, c->T, k, 3, c->cb, c->cp, c->ap, items);
      // the states before sub-steps s0 .. s1 - 1, saved by the sub-steps'
      // cell kernels; the state after s1 - 1 is the next checkpoint, so the
      // last sub-step runs its cell kernel only, for the save
      for (int s = s0; s < s1; s++) RCHK(rd_substep(c, rd.d, s - s0, s + 1 == s1));
    }
    for (int s = s1 - 1; s >= s0; s--) {
      launch_rd_backward_substep(c->stream, c->d_tis, rd.d, c->T, s - s0, c->cb, c->bp, q, qa, mn, mb, c->tu.rd_group);
      q ^= 1;
      qa = qa == 2 ? 0 : qa + 1;
    }
  }
  launch_rd_gj0(c->stream, c->d_tis, rd.d, c->T, c->cb, q, items);
  HIPCHK(c, hipGetLastError());
  RCHK(sync_status(c));
  for (int t = 0; t < c->T; t++) {
    double sc[4];
    HIPCHK(c, hipMemcpy(sc, rd.h[t].scal, sizeof(sc), hipMemcpyDeviceToHost));
    if (loss) loss[t] = sc[0];
    if (grads) {
      grads[t].du = sc[1];
      grads[t].dp = sc[2];
      grads[t].gj0 = sc[3];
      if (grads[t].kpa)
        HIPCHK(c, hipMemcpy(grads[t].kpa, rd.h[t].gkpa, 8 * (size_t)c->h_st[t].n, hipMemcpyDeviceToHost));
    }
  }
  return EPI_OK;
}

int epi_rd_end(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  auto& rd = c->rd;
  if (!rd.active) return EPI_OK;
  std::vector<double> g0(c->T);
  for (int t = 0; t < c->T; t++) {
    window_values_from_params(c->tis[t]);
    g0[t] = c->tis[t].prm.gj_init;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_tis, c->tis.data(), sizeof(Tis) * c->T, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(rd.d_gj0, g0.data(), 8 * c->T, hipMemcpyHostToDevice, c->stream));
  const long long items = rd_items(c);
  launch_rd_restore(c->stream, c->d_tis, rd.d, rd.d_gj0, c->T, c->cb, c->cp, c->ap, 0, items);
  launch_rd_put_gj(c->stream, c->d_tis, rd.d, c->T, c->cb, items);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(rd.arena); rd.arena = nullptr;
  if (rd.work) (void)hipFree(rd.work);
  rd.work = nullptr; rd.work_bytes = 0;
  (void)hipFree(rd.d); rd.d = nullptr;
  (void)hipFree(rd.d_gj0); rd.d_gj0 = nullptr;
  rd.active = false;
  return EPI_OK;
}

// ---- frozen mechanics window and its adjoint --------------------------------
namespace {
int mech_params(epi_ctx* c, const epi_mech_params* prm) {
  for (int t = 0; t < c->T; t++) {
    c->tis[t].m_eps = prm[t].eps;
    c->tis[t].m_kadh = prm[t].k_adh;
    c->tis[t].m_kspring = prm[t].k_spring;
    c->tis[t].m_kbend = prm[t].k_bend;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_tis, c->tis.data(), sizeof(Tis) * c->T, hipMemcpyHostToDevice, c->stream));
  return EPI_OK;
}
int mech_step(epi_ctx* c) {
  launch_forces(c->stream, c->d_tis, c->T, c->pb, c->cb, max_memb(c), max_n(c), max_list_len(c));
  c->pb ^= 1;
  return EPI_OK;
}
}  // namespace

int epi_mech_begin(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  auto& m = c->mech;
  if (m.active) return fail(c, EPI_ERR_INVALID, "a mechanics window is already open");
  if (c->pos_pending) return fail(c, EPI_ERR_INVALID, "pos_copy pending: open the window between steps");
  RCHK(sync_status(c));
  m.h.assign(c->T, MechTis{});
  Carver dry{nullptr, 0, true};
  auto carve = [&](Carver& cv) {
    for (int t = 0; t < c->T; t++) {
      MechTis& r = m.h[t];
      const size_t cap = c->th[t].cap;
      r.x0[0] = cv.take<double>(3 * cap); r.x0[1] = cv.take<double>(3 * cap);
      r.lam[0] = cv.take<double>(3 * cap); r.lam[1] = cv.take<double>(3 * cap);
      r.target = cv.take<double>(3 * cap);
      r.scal = cv.take<double>(5);
    }
  };
  carve(dry);
  if (hipMalloc(&m.arena, align_up(dry.off, 256)) != hipSuccess || hipMalloc(&m.d, sizeof(MechTis) * c->T) != hipSuccess)
    return fail(c, EPI_ERR_HIP, "window allocation failed");
  Carver real{m.arena, 0, false};
  carve(real);
  HIPCHK(c, hipMemcpyAsync(m.d, m.h.data(), sizeof(MechTis) * c->T, hipMemcpyHostToDevice, c->stream));
  launch_mech_state(c->stream, c->d_tis, m.d, c->T, 0, 5, c->pb, max_n(c));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  m.active = true;
  return EPI_OK;
}

int epi_mech_forward(epi_ctx* c, const epi_mech_params* prm, int32_t n_steps) {
  if (!c || !prm || n_steps < 0) return EPI_ERR_INVALID;
  if (!c->mech.active) return fail(c, EPI_ERR_INVALID, "no mechanics window is open");
  RCHK(mech_params(c, prm));
  launch_mech_state(c->stream, c->d_tis, c->mech.d, c->T, 0, 4, c->pb, max_n(c));
  for (int s = 0; s < n_steps; s++) RCHK(mech_step(c));
  HIPCHK(c, hipGetLastError());
  return sync_status(c);
}

int epi_mech_get_positions(epi_ctx* c, int32_t t, double* xyz) {
  RCHK(check_tissue(c, t));
  if (!xyz) return EPI_ERR_INVALID;
  const size_t n = (size_t)c->h_st[t].n;
  for (int a = 0; a < 3; a++)
    HIPCHK(c, hipMemcpy(xyz + a * n, c->tis[t].pos[c->pb][a], 8 * n, hipMemcpyDeviceToHost));
  return EPI_OK;
}

int epi_mech_adjoint(epi_ctx* c, const epi_mech_params* prm, int32_t n_steps, int32_t every,
                     const double* const* targets, double* loss, epi_mech_grad* grads) {
  if (!c || !prm || !targets || n_steps < 1 || every < 1) return EPI_ERR_INVALID;
  auto& m = c->mech;
  if (!m.active) return fail(c, EPI_ERR_INVALID, "no mechanics window is open");
  const int K = std::min(every, n_steps);
  const int nck = (n_steps + K - 1) / K;
  size_t need = 0;
  for (int t = 0; t < c->T; t++) need += align_up(8 * 3 * (size_t)c->th[t].cap * (K + nck), 256) + 512;
  if (need > m.work_bytes) {
    if (m.work) (void)hipFree(m.work);
    m.work = nullptr;
    m.work_bytes = 0;
    if (hipMalloc(&m.work, need) != hipSuccess) return fail(c, EPI_ERR_HIP, "window work allocation failed");
    m.work_bytes = need;
  }
  Carver cv{m.work, 0, false};
  for (int t = 0; t < c->T; t++) {
    MechTis& r = m.h[t];
    const size_t cap = c->th[t].cap, n = (size_t)c->h_st[t].n;
    r.slot = cv.take<double>(3 * cap * K);
    r.ck = cv.take<double>(3 * cap * nck);
    for (int a = 0; a < 3; a++)
      HIPCHK(c, hipMemcpyAsync(r.target + a * cap, targets[t] + a * n, 8 * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(r.scal, 0, 5 * sizeof(double), c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(m.d, m.h.data(), sizeof(MechTis) * c->T, hipMemcpyHostToDevice, c->stream));
  RCHK(mech_params(c, prm));
  const int mn = max_n(c);
  launch_mech_state(c->stream, c->d_tis, m.d, c->T, 0, 4, c->pb, mn);
  // checkpoint every K steps; the last segment saves its states straight
  // into the slots, so the reverse pass starts without a recompute
  std::vector<int> ck_pb(nck);
  const int last0 = (nck - 1) * K;
  for (int s = 0; s < n_steps; s++) {
    if (s >= last0) {
      launch_mech_state(c->stream, c->d_tis, m.d, c->T, s - last0, 0, c->pb, mn);
    } else if (s % K == 0) {
      ck_pb[s / K] = c->pb;
      launch_mech_state(c->stream, c->d_tis, m.d, c->T, s / K, 2, c->pb, mn);
    }
    RCHK(mech_step(c));
  }
  int q = 0;
  launch_mech_seed(c->stream, c->d_tis, m.d, c->T, c->pb, q, mn);
  for (int k = nck - 1; k >= 0; k--) {
    const int s0 = k * K, s1 = std::min(n_steps, s0 + K);
    if (k < nck - 1) {
      c->pb = ck_pb[k];
      launch_mech_state(c->stream, c->d_tis, m.d, c->T, k, 3, c->pb, mn);
      // the states before steps s0 .. s1 - 1; the one after s1 - 1 is the
      // next checkpoint and is not recomputed
      for (int s = s0; s < s1; s++) {
        launch_mech_state(c->stream, c->d_tis, m.d, c->T, s - s0, 0, c->pb, mn);
        if (s + 1 < s1) RCHK(mech_step(c));
      }
    }
    for (int s = s1 - 1; s >= s0; s--) {
      launch_mech_backward_step(c->stream, c->d_tis, m.d, c->T, s - s0, c->pb, c->cb, q, mn);
      q ^= 1;
    }
  }
  HIPCHK(c, hipGetLastError());
  RCHK(sync_status(c));
  for (int t = 0; t < c->T; t++) {
    double sc[5];
    HIPCHK(c, hipMemcpy(sc, m.h[t].scal, sizeof(sc), hipMemcpyDeviceToHost));
    if (loss) loss[t] = sc[0];
    if (grads) { grads[t].eps = sc[1]; grads[t].k_adh = sc[2]; grads[t].k_spring = sc[3]; grads[t].k_bend = sc[4]; }
  }
  return EPI_OK;
}

int epi_mech_end(epi_ctx* c) {
  if (!c) return EPI_ERR_INVALID;
  auto& m = c->mech;
  if (!m.active) return EPI_OK;
  for (int t = 0; t < c->T; t++) window_values_from_params(c->tis[t]);
  HIPCHK(c, hipMemcpyAsync(c->d_tis, c->tis.data(), sizeof(Tis) * c->T, hipMemcpyHostToDevice, c->stream));
  launch_mech_state(c->stream, c->d_tis, m.d, c->T, 0, 4, c->pb, max_n(c));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(m.arena); m.arena = nullptr;
  if (m.work) (void)hipFree(m.work);
  m.work = nullptr; m.work_bytes = 0;
  (void)hipFree(m.d); m.d = nullptr;
  m.active = false;
  return EPI_OK;
}
