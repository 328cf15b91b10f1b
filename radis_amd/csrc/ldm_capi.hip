// Host side of the C ABI declared in include/radis_ldm.h: device context,
// device-resident line lists, FFT planning, per-request workspace, the job
// table and the launch sequence of the LDM pipeline.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "radis_ldm.h"
#include "ldm_job.h"
#include "ldm_kernels.h"
#include "ldm_fftc.h"
#include "ldm_slit.h"
#include "ldm_wrap.h"
#include "ldm_setup.h"

namespace {

// ---- constants: radis/phys/constants.py:22-78 ------------------------------
const double K_H = 6.62607015e-34;
const double K_C = 2.99792458e8;
const double K_KB = 1.3806490e-23;
const double K_NA = 6.02214076e23;
const double K_KB_CGS = 1.380648813e-16;
const double HC_K = K_H * K_C / K_KB * 100;  // constants.py:69
const double ATM_PER_BAR = 1.01325;          // lbl/broadening.py:1063

const int STAT_BLOCK = 4096;             // lines per host-side statistics block
const int FILL_LDS_TARGET = 44 * 1024;   // preferred LDS bytes of one fill tile (FP64 cells; 3 tiles per CU)
const int FILL_LDS_MAX = 60 * 1024;      // hard limit (plus ~2.2 KB of static axis tables, < 64 KB)
const int MIN_TILE_W = 16, MAX_TILE_W = 4096;
const int N1_MAX = 1024, N2_MAX = 8192;

// last error of the calling thread (a fixed buffer: reporting an error
// allocates nothing)
thread_local char g_err[1024];

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(RADIS_LDM_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int ctz(long long x) {
  int n = 0;
  while (x > 0 && !(x & 1)) { x >>= 1; ++n; }
  return n;
}
int floor_log2(long long x) {
  int n = -1;
  while (x > 0) { x >>= 1; ++n; }
  return n;
}

}  // namespace

// Scratch array of radis_ldm_run_noalloc: job table, axes, work items,
// twiddle tables and workspace of one request.  Every device the library
// runs on holds its own copy until the process ends.  In host code the name
// is a stand-in, not a device address: kernels get the address that
// hipGetSymbolAddress returns for the current device (scratch_address).
static __device__ __align__(256) unsigned char scratch[RADIS_LDM_NOALLOC_BYTES];

namespace {

const int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char* g_scratch[SCRATCH_MAX_DEVICES] = {};

// The scratch array's address on `device` (which must be current), taken
// once per device.
int scratch_address(int device, char** out) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES)
    return fail(RADIS_LDM_EINVAL, "device %d beyond the %d supported", device, SCRATCH_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(RADIS_LDM_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch[device] = (char*)p;
  }
  *out = g_scratch[device];
  return 0;
}

}  // namespace

// ============================================================================
// objects
// ============================================================================
// Contexts and line lists are reference counted by the objects built on
// them: destroying a parent while children are alive only marks it closed,
// and the last child to go releases it.
struct radis_ldm_ctx {
  int device = 0;
  int n_cu = 256;       // compute units (persistent pass A: one workgroup each)
  int64_t budget = 0;
  int64_t in_use = 0;   // bytes of live line lists and batch workspaces
  int refs = 0;
  bool closed = false;
  // Released device / pinned-host blocks kept for reuse, by size: batch
  // set-up and tear-down in a serving loop then never call hipMalloc or
  // hipFree (which can synchronise the device).
  std::multimap<size_t, void*> dev_cache, host_cache;
  size_t dev_cached = 0, host_cached = 0;
  // this device's copy of the no-allocation entry point's scratch array
  // (its address from hipGetSymbolAddress, taken once per device)
  char* scratch = nullptr;
  ~radis_ldm_ctx();
};

struct radis_ldm_db {
  radis_ldm_ctx* ctx = nullptr;
  int refs = 0;
  bool closed = false;
  int64_t n = 0;
  void* mem = nullptr;
  size_t bytes = 0;
  LdmDb dev{};
  bool has_iso = false, has_nself = false, has_delta = false, has_evib = false;
  double molar_mass[LDM_MAX_ISO] = {0};
  int n_iso_slots = 0;
  unsigned iso_mask = 0;
  double max_abs_delta = 0.0;
  std::vector<double> h_nu0;
  // per-block statistics used for workspace and scale bounds
  std::vector<float> ga_min, ga_max, gs_min, gs_max, na_min, na_max, ns_min, ns_max, e_min, e_max, s_max;
  std::vector<float> dv_min, dv_max;  // Evib_u - Evib_l (two-temperature bound)
  bool has_sky = false;                // Lorentzian-width candidates per block (k_axis)
  std::vector<int> h_sky_off;          // host copy of the candidate offsets
  // the whole list's Pareto sets for the 4 directions (the host's bound on N_L)
  std::vector<float4> h_gsky[4];
  // partition-function splines (radis_ldm_db_set_qspline): device slots and host copies
  LdmQSpline* qs_mem = nullptr;
  std::vector<LdmQSpline> h_qs;
};

namespace {

void ctx_release(radis_ldm_ctx* ctx) {
  if (--ctx->refs == 0 && ctx->closed) delete ctx;
}

// Round up to the cache granularity (256 B below 1 MiB, 1 MiB above).
size_t cache_round(size_t n) {
  return n < (1u << 20) ? align_up(std::max<size_t>(n, 256), 256) : align_up(n, (size_t)1 << 20);
}

// Take a cached block of at least n bytes (and at most 2n) or allocate one.
void* cache_take(std::multimap<size_t, void*>& cache, size_t& cached, size_t n, bool host) {
  n = cache_round(n);
  auto it = cache.lower_bound(n);
  if (it != cache.end() && it->first <= 2 * n) {
    void* p = it->second;
    cached -= it->first;
    cache.erase(it);
    return p;
  }
  void* p = nullptr;
  hipError_t e = host ? hipHostMalloc(&p, n) : hipMalloc(&p, n);
  if (e != hipSuccess && !cache.empty()) {
    // release the cache and retry once
    for (auto& kv : cache) (void)(host ? hipHostFree(kv.second) : hipFree(kv.second));
    cache.clear();
    cached = 0;
    e = host ? hipHostMalloc(&p, n) : hipMalloc(&p, n);
  }
  return e == hipSuccess ? p : nullptr;
}

// Block sizes are tracked by the owners; the cache keys blocks by the size
// they were taken with (rounded the same way).
void cache_give(std::multimap<size_t, void*>& cache, size_t& cached, void* p, size_t n) {
  if (!p) return;
  n = cache_round(n);
  cache.emplace(n, p);
  cached += n;
}

void* dev_alloc(radis_ldm_ctx* c, size_t n) { return cache_take(c->dev_cache, c->dev_cached, n, false); }
void dev_release(radis_ldm_ctx* c, void* p, size_t n) { cache_give(c->dev_cache, c->dev_cached, p, n); }
void* host_alloc(radis_ldm_ctx* c, size_t n) { return cache_take(c->host_cache, c->host_cached, n, true); }
void host_release(radis_ldm_ctx* c, void* p, size_t n) { cache_give(c->host_cache, c->host_cached, p, n); }

void db_free(radis_ldm_db* db) {
  if (db->mem) {
    (void)hipFree(db->mem);
    db->ctx->in_use -= (int64_t)db->bytes;
  }
  if (db->qs_mem) {
    (void)hipFree(db->qs_mem);
    db->ctx->in_use -= (int64_t)(sizeof(LdmQSpline) * LDM_MAX_ISO);
  }
  radis_ldm_ctx* ctx = db->ctx;
  delete db;
  if (ctx) ctx_release(ctx);
}

void db_release(radis_ldm_db* db) {
  if (--db->refs == 0 && db->closed) db_free(db);
}

struct FftPlan {
  int64_t L = 0;
  int n1 = 0, n2 = 0;
  bool fc = false;  // compile-time-sized passes (ldm_fftc.h) for this n1, n2
  int nr1 = 0, nr2 = 0;
  int r1[LDM_MAX_RAD] = {0}, r2[LDM_MAX_RAD] = {0};
  int wrap_k = 0;   // bins of the circular-image correction (0: L is the spectrum's period)
};

// Sizes with compile-time-sized passes: n1 (passes A and C) and n2 (pass B).
const int FC_N1[] = {256, 320, 384, 512, 640, 768, 1024};
const int FC_N2[] = {256, 512, 1024, 2048};
// A compiled split is preferred while its length is within this factor of
// the smallest generic (7-smooth) length: its passes run ~3x faster.
const double FC_SLACK = 1.35;

struct Req {
  radis_ldm_request r{};
  radis_ldm_derived derived{};  // all null: none
  int64_t nv = 0;
  double delta_w = 0;
  FftPlan plan;
  int64_t i0 = 0, i1 = 0;
  int m_need = 0;      // bound for the current conditions
  int m_cap = 0;       // allocated rows
  int wrap_rows = 0;   // L-axis nodes of the image-correction table (N_L <= M/2: N_G >= 2)
  int tile_w = 0, n_tiles = 0;
  int64_t stride = 0;
  int scale_log2 = 0;
  int bA = 0, bB = 0, bC = 0;  // log2 of the pass A / C column blocks and pass B row blocks
  int n_axp = 0;       // axis partials (items of k_axis)
  bool sky_whole = false;  // width extremes from the whole list's Pareto sets in k_finalize (no k_axis items)
  size_t off_ldm = 0, off_y = 0, off_b = 0, off_rows = 0, off_axp = 0, off_wrap = 0;
  const float2 *tw1 = nullptr, *tw2 = nullptr, *twl_hi = nullptr, *twl_lo = nullptr;
};

}  // namespace

// One launch of an FFT pass: items of one kernel variant are contiguous;
// cls = the compiled size (n1 for A and C, n2 for B), 0 = generic (pass B:
// 0 = generic with small tiles, -1 = generic with large tiles).
struct LaunchGroup { int cls, first, count; };

struct radis_ldm_batch {
  radis_ldm_ctx* ctx = nullptr;
  radis_ldm_db* db = nullptr;
  int n = 0;
  std::vector<Req> reqs;
  // two pinned job tables / result slots so that run k only waits for run k-2
  LdmJob* jobs_h[2] = {nullptr, nullptr};
  LdmJob* jobs_d = nullptr;
  LdmAxes* axes_d = nullptr;
  LdmResult* res_h[2] = {nullptr, nullptr};
  // the job-table bytes last copied to jobs_d: an unchanged table is not
  // copied again (the device side rewrites only what k_finalize recomputes)
  std::vector<char> jobs_last;
  // res_d holds results of a run not yet copied to res_h[last_slot]
  bool res_stale = false;
  LdmResult* res_d = nullptr;
  int next_slot = 0, last_slot = 0;
  void* work = nullptr;
  size_t work_bytes = 0;
  void* twmem = nullptr;
  size_t tw_bytes = 0;
  int4* items_d = nullptr;
  size_t items_bytes = 0;
  int n_axis = 0, n_wrap = 0, n_fill = 0, n_a = 0, n_b = 0, n_c = 0;
  // launches of the FFT passes (see LaunchGroup)
  using Group = LaunchGroup;
  std::vector<Group> grp_a, grp_b, grp_c;
  bool tiles_valid = false;  // fill tile bounds computed for the current layout
  size_t fill_lds = 0;
  std::vector<float*> outs;        // outputs of the last eager run
  std::vector<float*> graph_outs;  // outputs bound to the captured graph
  hipEvent_t ev_done[2] = {nullptr, nullptr};
  bool pending[2] = {false, false};
  bool timing = false;
  bool slot_timed[2] = {false, false};
  // start/stop events of every kernel of a timed run (tile bounds counted in the fill stage)
  static constexpr int N_KERNELS = RADIS_LDM_N_STAGES;
  hipEvent_t st_ev[2][2 * N_KERNELS] = {};
  bool ev_used[2][N_KERNELS] = {};
  double stage_sum[RADIS_LDM_N_STAGES] = {0};
  int stage_count = 0;
  hipStream_t own_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
};

namespace {

// ---- FFT planning ----------------------------------------------------------
// 16*a with a 7-smooth, <= maxv, ascending (fixed storage: planning
// allocates nothing)
struct SmoothList {
  int v[512];
  int n = 0;
  explicit SmoothList(int maxv) {
    for (int a = 1; 16 * a <= maxv && n < 512; ++a) {
      int x = a;
      for (int p : {2, 3, 5, 7})
        while (x % p == 0) x /= p;
      if (x == 1) v[n++] = 16 * a;
    }
  }
  const int* begin() const { return v; }
  const int* end() const { return v + n; }
};

bool radices_of(int n, int* r, int& cnt) {
  cnt = 0;
  for (int p : {16, 8, 4, 2, 7, 5, 3}) {
    while (n % p == 0) {
      if (cnt == LDM_MAX_RAD) return false;
      r[cnt++] = p;
      n /= p;
    }
  }
  return n == 1;
}

// Bins of the circular-image correction for a transform of length L of a
// spectrum with period need = 2 N_v (LDM_WRAP_* in ldm_job.h): 0 when
// L == need, -1 when the window over the gap would need more than
// LDM_WRAP_KMAX bins (a gap narrower than ~2.5 % of L, or L above ~20 N_v).
int wrap_bins(int64_t L, int64_t need) {
  if (L == need) return 0;
  if (L < need) return -1;
  const double w = std::min((double)(L - need), 0.25 * (double)need);
  const double k = std::ceil(LDM_WRAP_CK * LDM_WRAP_NSIG / (2 * 3.14159265358979323846) * (double)L / w);
  return k <= LDM_WRAP_KMAX ? (int)k : -1;
}

// Smallest transform length above need whose gap is wide enough.
int64_t wrap_floor(int64_t need) {
  const double c = LDM_WRAP_CK * LDM_WRAP_NSIG / (2 * 3.14159265358979323846);
  int64_t L = (int64_t)std::ceil((double)need * LDM_WRAP_KMAX / (LDM_WRAP_KMAX - c));
  while (L <= 40 * need && wrap_bins(L, need) < 0) ++L;
  return L;
}

// Search L = n1*n2 >= 2*N_v (or == want) with 16 | n1 <= 1024 and
// 16 | n2 <= n2_max, 7-smooth cofactors; smallest L, then n1 closest to 512.
int64_t search_fft(int64_t need, int64_t want, int n2_max, int& b1, int& b2, int only_n1 = 0) {
  static const SmoothList c1(N1_MAX);
  static const SmoothList c2(N2_MAX);
  int64_t bestL = -1;
  double bestScore = 1e300;
  for (int n1 : c1) {
    if (only_n1 && n1 != only_n1) continue;
    int n2 = 0;
    if (want > 0) {
      if (want % n1) continue;
      const int64_t q = want / n1;
      if (q > n2_max || !std::binary_search(c2.begin(), c2.end(), (int)q)) continue;
      n2 = (int)q;
    } else {
      const int64_t n2min = (need + n1 - 1) / n1;
      if (n2min > n2_max) continue;
      auto itn = std::lower_bound(c2.begin(), c2.end(), (int)n2min);
      if (itn == c2.end() || *itn > n2_max) continue;
      n2 = *itn;
    }
    const int64_t L = (int64_t)n1 * n2;
    const double score = std::fabs(std::log2((double)n1) - 9.0);
    if (bestL < 0 || L < bestL || (L == bestL && score < bestScore)) {
      bestL = L; bestScore = score; b1 = n1; b2 = n2;
    }
  }
  return bestL;
}

// Plan the transform of a spectrum of period need = 2 N_v.  With `period`
// (RADIS_LDM_PERIOD_L) the transform length is the spectrum's period, any
// L >= need; otherwise L is need itself or at least wrap_floor(need), and
// k_wrap corrects the images (the result is RADIS's at L = 2 N_v).
int plan_fft(int64_t nv, int64_t want, FftPlan& out, int only_n1 = 0, bool allow_fc = true, bool period = false) {
  const int64_t need = 2 * nv;
  if (want > 0 && want < need) return fail(RADIS_LDM_EINVAL, "fft_len %lld < 2*N_v = %lld", (long long)want, (long long)need);
  if (want > 0 && !period && wrap_bins(want, need) < 0)
    return fail(RADIS_LDM_ERANGE,
                "fft_len %lld is too close to 2*N_v = %lld for the circular-image correction (need L = 2*N_v or "
                "L >= %lld)", (long long)want, (long long)need, (long long)wrap_floor(need));
  // smallest admissible length among the candidates of `search` (L == need
  // itself, else from `lo` on)
  const int64_t lo = period ? need : wrap_floor(need);
  auto pick = [&](int n2_max, int& b1, int& b2) -> int64_t {
    if (want > 0) return search_fft(need, want, n2_max, b1, b2, only_n1);
    int e1 = 0, e2 = 0;
    const int64_t exact = search_fft(need, need, n2_max, e1, e2, only_n1);
    if (exact > 0) { b1 = e1; b2 = e2; return exact; }
    return search_fft(lo, 0, n2_max, b1, b2, only_n1);
  };
  int a1 = 0, a2 = 0, s1 = 0, s2 = 0;
  const int64_t any = pick(N2_MAX, a1, a2);
  const int64_t small = pick(LDM_CAP_B, s1, s2);  // pass B with many small tiles
  if (any < 0) {
    if (want > 0)
      return fail(RADIS_LDM_ERANGE, "fft_len %lld is not n1*n2 with 16|n1<=%d, 16|n2<=%d and 7-smooth cofactors",
                  (long long)want, N1_MAX, N2_MAX);
    return fail(RADIS_LDM_ERANGE, "grid of %lld points needs an FFT above %lld points", (long long)nv,
                (long long)N1_MAX * N2_MAX);
  }
  const bool use_small = small > 0 && (double)small <= 1.03 * (double)any;
  out.L = use_small ? small : any;
  out.n1 = use_small ? s1 : a1;
  out.n2 = use_small ? s2 : a2;
  out.fc = false;
  // compiled split: smallest admissible L = n1*n2 (or == want), then n1 closest to 512
  int64_t bestc = -1;
  int c1 = 0, c2 = 0;
  for (int n1 : FC_N1) {
    if (!allow_fc) break;
    if (only_n1 && n1 != only_n1) continue;
    for (int n2 : FC_N2) {
      const int64_t L = (int64_t)n1 * n2;
      if (want > 0 ? L != want : (L != need && L < lo)) continue;
      if (bestc < 0 || L < bestc || (L == bestc && std::abs(n1 - 512) < std::abs(c1 - 512))) {
        bestc = L; c1 = n1; c2 = n2;
      }
    }
  }
  if (bestc > 0 && (double)bestc <= FC_SLACK * (double)out.L) {
    out.L = bestc;
    out.n1 = c1;
    out.n2 = c2;
    out.fc = true;
  }
  if (!radices_of(out.n1, out.r1, out.nr1) || !radices_of(out.n2, out.r2, out.nr2))
    return fail(RADIS_LDM_ERANGE, "radix factorisation failed");
  out.wrap_k = period ? 0 : wrap_bins(out.L, need);
  if (out.wrap_k < 0)
    return fail(RADIS_LDM_ERANGE,
                "grid of %lld points is too short for the circular-image correction of the smallest transform "
                "(L = %lld); request RADIS_LDM_PERIOD_L", (long long)nv, (long long)out.L);
  return 0;
}

// numpy.arange(wmin, wmax + wstep, wstep): length ceil((stop-start)/step),
// values start + i*delta with delta = (start+step) - start.
int grid_of(double wmin, double wmax, double wstep, int64_t& nv, double& delta) {
  if (!(wstep > 0) || !(wmax > wmin) || !std::isfinite(wmin) || !std::isfinite(wmax))
    return fail(RADIS_LDM_EINVAL, "invalid grid wmin=%g wmax=%g wstep=%g", wmin, wmax, wstep);
  const double stop = wmax + wstep;
  const double len = std::ceil((stop - wmin) / wstep);
  if (!(len >= 2) || len > 1e9) return fail(RADIS_LDM_EINVAL, "grid length %g out of range", len);
  nv = (int64_t)len;
  delta = (wmin + wstep) - wmin;
  return 0;
}

int validate_request(const radis_ldm_request& r) {
  if (!(r.Tgas > 0) || !(r.Tref > 0)) return fail(RADIS_LDM_EINVAL, "temperatures must be positive");
  if (!(r.pressure_bar > 0)) return fail(RADIS_LDM_EINVAL, "pressure must be positive");
  if (!(r.mole_fraction >= 0) || !(r.mole_fraction <= 1)) return fail(RADIS_LDM_EINVAL, "mole_fraction must be in [0, 1]");
  if (!(r.dxG > 0) || !(r.dxL > 0)) return fail(RADIS_LDM_EINVAL, "dxG and dxL must be positive");
  if (!(r.folding_thresh > 0)) return fail(RADIS_LDM_EINVAL, "folding_thresh must be positive");
  if ((r.flags & RADIS_LDM_NONEQ) && (!(r.Tvib > 0) || !(r.Trot > 0)))
    return fail(RADIS_LDM_EINVAL, "RADIS_LDM_NONEQ needs positive Tvib and Trot");
  return 0;
}

double dopp_factor(double T, double M) {
  // sqrt((2 * Na * k_b_CGS * Tgas * ln(2)) / molar_mass), lbl/broadening.py:170-172
  return std::sqrt((2 * K_NA * K_KB_CGS * T * std::log(2.0)) / M);
}

// Upper bound of N_G*N_L and the intensity scale for the request's
// conditions, from per-block statistics of the window's lines.
int bound_request(const radis_ldm_db& db, Req& R) {
  const radis_ldm_request& r = R.r;
  if ((r.flags & RADIS_LDM_NONEQ) && !db.has_evib)
    return fail(RADIS_LDM_EINVAL, "RADIS_LDM_NONEQ needs a line list with Evib_l and Evib_u");
  if (r.flags & RADIS_LDM_DEVICE_Q) {
    if (r.flags & RADIS_LDM_NONEQ)
      return fail(RADIS_LDM_EINVAL, "RADIS_LDM_DEVICE_Q takes equilibrium partition functions (not with RADIS_LDM_NONEQ)");
    for (int s = 0; s < db.n_iso_slots; ++s) {
      if (!(db.iso_mask & (1u << s))) continue;
      if (!(db.dev.qs_mask & (1u << s)))
        return fail(RADIS_LDM_EINVAL, "RADIS_LDM_DEVICE_Q: isotopologue slot %d has lines but no partition-function spline", s);
      const LdmQSpline& q = db.h_qs[s];
      // PartFuncCO2_CDSDtab._inrange, radis/levels/partfunc_cdsd.py:101-103
      for (double T : {r.Tgas, r.Tref})
        if (!(T >= q.lo && T <= q.hi))
          return fail(RADIS_LDM_ERANGE, "temperature %g K outside the partition table of slot %d [%g, %g]", T, s, q.lo, q.hi);
    }
  }
  if (R.i1 <= R.i0) {
    R.m_need = 1;
    R.scale_log2 = 0;
    return 0;
  }
  const int b0 = (int)(R.i0 / STAT_BLOCK), b1 = (int)((R.i1 - 1) / STAT_BLOCK);
  double ga0 = 1e300, ga1 = -1e300, gs0 = 1e300, gs1 = -1e300, na0 = 1e300, na1 = -1e300;
  double ns0 = 1e300, ns1 = -1e300, e0 = 1e300, e1 = -1e300, smax = 0;
  for (int b = b0; b <= b1; ++b) {
    ga0 = std::min(ga0, (double)db.ga_min[b]); ga1 = std::max(ga1, (double)db.ga_max[b]);
    gs0 = std::min(gs0, (double)db.gs_min[b]); gs1 = std::max(gs1, (double)db.gs_max[b]);
    na0 = std::min(na0, (double)db.na_min[b]); na1 = std::max(na1, (double)db.na_max[b]);
    ns0 = std::min(ns0, (double)db.ns_min[b]); ns1 = std::max(ns1, (double)db.ns_max[b]);
    e0 = std::min(e0, (double)db.e_min[b]); e1 = std::max(e1, (double)db.e_max[b]);
    smax = std::max(smax, (double)db.s_max[b]);
  }
  const double p_atm = r.pressure_bar / ATM_PER_BAR;
  const double lr = std::log(r.Tref / r.Tgas);
  auto rlo = [&](double a, double b) { return std::min(std::exp(a * lr), std::exp(b * lr)); };
  auto rhi = [&](double a, double b) { return std::max(std::exp(a * lr), std::exp(b * lr)); };
  const double x = r.mole_fraction;
  double wl_lo = 2 * p_atm * (rlo(na0, na1) * ga0 * (1 - x) + rlo(ns0, ns1) * gs0 * x) * (1 - 1e-9);
  double wl_hi = 2 * p_atm * (rhi(na0, na1) * ga1 * (1 - x) + rhi(ns0, ns1) * gs1 * x) * (1 + 1e-9);
  if (!db.h_gsky[0].empty()) {
    // tighter: the width's extremes over the whole list's Pareto sets (the
    // same directions k_axis takes; a superset of the window, so still
    // bounds); the per-block statistics above pair extremes of g and n
    // that no single line has
    auto wl = [&](const float4& q) {
      return 2 * p_atm * (std::exp((double)q.z * lr) * q.x * (1 - x) + std::exp((double)q.w * lr) * q.y * x);
    };
    const int dmx = lr > 0.0 ? 2 : 0, dmn = dmx + 1;
    double hi = 0, lo = 1e300;
    for (const float4& q : db.h_gsky[dmx]) hi = std::max(hi, wl(q));
    for (const float4& q : db.h_gsky[dmn]) lo = std::min(lo, wl(q));
    wl_hi = std::min(wl_hi, hi * (1 + 1e-9));
    wl_lo = std::max(wl_lo, lo * (1 - 1e-9));
  }
  if (!(wl_lo > 0)) return fail(RADIS_LDM_EINVAL, "zero Lorentzian width in the window (check g_air/g_self and mole fraction)");
  double dmin = 1e300, dmax = 0;
  for (int s = 0; s < db.n_iso_slots; ++s) {
    if (!(db.iso_mask & (1u << s))) continue;
    const double d = dopp_factor(r.Tgas, db.molar_mass[s]);
    dmin = std::min(dmin, d);
    dmax = std::max(dmax, d);
  }
  const double wg_lo = 2 * (db.h_nu0[R.i0] / 2.99792458e10) * dmin * (1 - 1e-9);
  const double wg_hi = 2 * (db.h_nu0[R.i1 - 1] / 2.99792458e10) * dmax * (1 + 1e-9);
  if (!(wg_lo > 0)) return fail(RADIS_LDM_EINVAL, "non-positive Gaussian width (nu0 <= 0 in window)");
  // k_axis takes N = ceil((log(w_max + 1e-4) - log(w_min)) / dx) + 1 from the
  // exact extremes; the bounds above enclose them with a 1e-9 relative
  // margin, far above the rounding of either side's logarithm, so the same
  // expression on them is at least the device's N
  const double ngf = std::ceil((std::log(wg_hi + 1e-4) - std::log(wg_lo)) / r.dxG) + 1;
  const double nlf = std::ceil((std::log(wl_hi + 1e-4) - std::log(wl_lo)) / r.dxL) + 1;
  if (ngf > LDM_MAX_AX || nlf > LDM_MAX_AX)
    return fail(RADIS_LDM_ERANGE, "LDM axes too long (N_G<=%g, N_L<=%g; limit %d): increase dxG/dxL", ngf, nlf, LDM_MAX_AX);
  R.m_need = (int)(ngf * nlf);
  // intensity scale: log2 of an upper bound of S(T) * density
  double lb = -1e300;
  if (smax > 0) {
    const bool neq = (r.flags & RADIS_LDM_NONEQ) != 0;
    double qmax = 0;
    const bool dq = (r.flags & RADIS_LDM_DEVICE_Q) != 0;
    for (int s = 0; s < db.n_iso_slots; ++s)
      if (db.iso_mask & (1u << s))
        // device-evaluated ratios: the same spline evaluated here (for the bound only)
        qmax = std::max(qmax, dq ? qspline_eval(db.h_qs[s], r.Tref) / qspline_eval(db.h_qs[s], r.Tgas) : r.q_ratio[s]);
    const double numin = std::max(db.h_nu0[R.i0], 1e-3);
    double boltz, stim;
    if (!neq) {
      const double cb = -HC_K * (1 / r.Tgas - 1 / r.Tref);
      boltz = std::max(cb * e0, cb * e1);
      stim = (r.Tgas >= r.Tref) ? 1.0 : 1.0 / (1.0 - std::exp(-HC_K * numin / r.Tref));
    } else {
      // exponent -c2(Evl/Tv + (El-Evl)/Tr - El/Tref) is linear in Evl in
      // [0, El]: extreme at Evl = 0 or El; the stimulated-emission numerator
      // 1 - exp(-c2(dEv/Tv + (nu0-dEv)/Tr)) is below 1 unless that exponent
      // is positive (inversion), where it is below exp(-c2 X_min) over the
      // block's dEv = Evib_u - Evib_l range
      const double cv = -HC_K * (1 / r.Tvib - 1 / r.Tref), cr = -HC_K * (1 / r.Trot - 1 / r.Tref);
      boltz = std::max(std::max(cv * e0, cv * e1), std::max(cr * e0, cr * e1));
      double dv0 = 1e300, dv1 = -1e300;
      for (int b = b0; b <= b1; ++b) { dv0 = std::min(dv0, (double)db.dv_min[b]); dv1 = std::max(dv1, (double)db.dv_max[b]); }
      const double k = 1 / r.Tvib - 1 / r.Trot;
      const double xmin = HC_K * (std::min(dv0 * k, dv1 * k) + numin / r.Trot);
      const double num = xmin < 0 ? std::exp(-xmin) : 1.0;
      stim = num / (1.0 - std::exp(-HC_K * numin / r.Tref));
    }
    const double dens = r.mole_fraction * ((r.pressure_bar * 1e5) / (K_KB * r.Tgas)) * 1e-6;
    if (qmax > 0 && dens > 0)
      lb = std::log2(smax) + std::log2(qmax) + boltz / std::log(2.0) + std::log2(stim) + std::log2(dens);
  }
  R.scale_log2 = (lb > 80 || (lb < -80 && lb > -1e299)) ? (int)std::lround(lb) - 40 : 0;
  return 0;
}

size_t fill_tile_bytes(int m_cap, int w) { return (size_t)m_cap * (size_t)(w + LDM_FILL_PAD) * sizeof(double); }

int choose_tile(int m_cap) {
  int w = MAX_TILE_W;
  while (w > MIN_TILE_W && fill_tile_bytes(m_cap, w) > (size_t)FILL_LDS_TARGET) w >>= 1;
  return w;
}

void fill_job(const radis_ldm_db& db, const Req& R, char* work, float* out, LdmJob& J) {
  const radis_ldm_request& r = R.r;
  std::memset(&J, 0, sizeof J);
  J.wmin = r.wmin;
  J.delta_w = R.delta_w;
  J.inv_delta_w = 1.0 / R.delta_w;
  {
    // w[j+1] - w[j] is delta_w up to the rounding of the two grid values
    // (a few ulp of the largest |w|); below one step minus that and 1e-7 of
    // a step, x - w[j] >= 0 puts x inside [w[j], w[j+1]) with NumPy's
    // position rounding to j (grid_index)
    const double wlast = r.wmin + (double)(R.nv - 1) * R.delta_w;
    const double wabs = std::max(std::fabs(r.wmin), std::fabs(wlast));
    J.grid_fast_hi = R.delta_w * (1.0 - 1e-7) - 8.0 * 2.220446049250313e-16 * wabs;
    J.inv_delta_f = (float)(1.0 / R.delta_w);
  }
  J.wstep = r.wstep;
  J.nv = R.nv;
  J.v = R.nv + 1;
  J.L = R.plan.L;
  J.i0 = R.i0;
  J.i1 = R.i1;
  J.p_atm = r.pressure_bar / ATM_PER_BAR;
  J.x_self = r.mole_fraction;
  J.x_air = 1 - r.mole_fraction;
  J.log_r = std::log(r.Tref / r.Tgas);
  J.c_boltz = -HC_K * (1 / r.Tgas - 1 / r.Tref);
  const double dens = r.mole_fraction * ((r.pressure_bar * 1e5) / (K_KB * r.Tgas)) * 1e-6;  // factory.py:908
  J.i_scale = std::ldexp(dens, -R.scale_log2);
  J.i_fast = 1;
  for (int s = 0; s < LDM_MAX_ISO; ++s) {
    J.dopp[s] = (s < db.n_iso_slots && db.molar_mass[s] > 0) ? dopp_factor(r.Tgas, db.molar_mass[s]) : 0.0;
    J.qratio[s] = r.q_ratio[s];
    const double qi = r.q_ratio[s] * J.i_scale;
    J.qi_f[s] = (float)qi;
    // S_ref (>= ~1e-45) times qi must stay in FP32 range before the
    // Boltzmann power of two is applied (k_finalize decides it for the
    // slots whose ratio it computes)
    const bool dev_q = (r.flags & RADIS_LDM_DEVICE_Q) && (db.dev.qs_mask & (1u << s));
    if ((db.iso_mask & (1u << s)) && !dev_q && !(qi > 0x1p-60 && qi < 0x1p+100)) J.i_fast = 0;
  }
  J.dxG = r.dxG;
  J.dxL = r.dxL;
  J.fold_thresh = r.folding_thresh;
  J.shift_margin = (db.has_delta ? db.max_abs_delta * J.p_atm : 0.0) + 2 * R.delta_w;
  J.log_r_f = (float)J.log_r;
  J.pa_air_f = (float)(J.p_atm * J.x_air);
  J.pa_self_f = (float)(J.p_atm * J.x_self);
  for (int s = 0; s < LDM_MAX_ISO; ++s) J.lgc_f[s] = J.dopp[s] > 0 ? (float)std::log2(2 * J.dopp[s] / 2.99792458e10) : 0.0f;
  J.fp64_widths = (r.flags & RADIS_LDM_FP64_WIDTHS) ? 1 : 0;
  J.min_rms = (r.flags & RADIS_LDM_MIN_RMS) ? 1 : 0;
  J.noneq = (r.flags & RADIS_LDM_NONEQ) ? 1 : 0;
  J.device_q = (r.flags & RADIS_LDM_DEVICE_Q) ? 1 : 0;
  J.Tref = r.Tref;
  J.c_vib = J.noneq ? -HC_K / r.Tvib : 0.0;
  J.c_rot = J.noneq ? -HC_K / r.Trot : 0.0;
  J.c_ref = HC_K / r.Tref;
  {
    // constants of the min-RMS corrections, radis/lbl/broadening.py:2271-2276
    const double pi = 3.14159265358979323846, ln2 = std::log(2.0);
    J.mr_dv = (float)r.wstep;
    J.mr_dxG = (float)r.dxG;
    J.mr_dxL = (float)r.dxL;
    J.mr_c1gg = (float)std::pow((6 * pi - 16) / (15 * pi - 32), 1 / 1.50);
    J.mr_c1lg = (float)std::pow((6 * pi - 16) / 3 * std::pow(ln2 / (2 * pi), 0.5), 1 / 2.25);
    J.mr_c2gg = (float)std::pow(2 * ln2 / 15, 1 / 1.50);
    J.mr_c2lg = (float)std::pow((2 * ln2) * (2 * ln2) / 15, 1 / 2.25);
  }
  J.c_stim_T = (float)(-HC_K / r.Tgas);
  J.c_stim_R = (float)(-HC_K / r.Tref);
  J.out_scale = (float)(1.0 / ((double)R.plan.L * r.wstep));
  J.inv_Lws = (float)(1.0 / ((double)R.plan.L * r.wstep));
  J.half_inv_ws = (float)(0.5 / r.wstep);
  J.single_iso = db.has_iso ? 0 : 1;
  J.n1 = R.plan.n1;
  J.n2 = R.plan.n2;
  J.nrad1 = R.plan.nr1;
  J.nrad2 = R.plan.nr2;
  for (int i = 0; i < J.nrad1; ++i) J.rad1[i] = R.plan.r1[i];
  for (int i = 0; i < J.nrad2; ++i) J.rad2[i] = R.plan.r2[i];
  J.bA = R.bA;
  J.bB = R.bB;
  J.m_cap = R.m_cap;
  J.tile_w = R.tile_w;
  J.n_tiles = R.n_tiles;
  J.ldm_stride = R.stride;
  J.ldm = (float*)(work + R.off_ldm);
  J.ybuf = (float2*)(work + R.off_y);
  J.bbuf = (float2*)(work + R.off_b);
  J.out = out;
  J.d_abs = R.derived.absorbance;
  J.d_trans = R.derived.transmittance;
  J.d_emis = R.derived.emissivity;
  J.d_rad = R.derived.radiance;
  J.derived = (J.d_abs || J.d_trans || J.d_emis || J.d_rad) ? 1 : 0;
  J.path_length = R.derived.path_length;
  J.Tgas = r.Tgas;
  J.scale_log2 = R.scale_log2;
  J.nfold = (int*)(work + R.off_rows);
  J.row_a = (float*)(work + R.off_rows + (size_t)4 * R.m_cap);
  J.row_b = (float*)(work + R.off_rows + (size_t)8 * R.m_cap);
  J.row_inv_f0 = (float*)(work + R.off_rows + (size_t)12 * R.m_cap);
  J.ax_part = (unsigned long long*)(work + R.off_axp);
  J.n_ax_part = R.n_axp;
  J.wrap_k = R.plan.wrap_k;
  J.wrap_rows = R.wrap_rows;
  J.wrap_h = R.plan.wrap_k ? (float*)(work + R.off_wrap) : nullptr;
  J.wrap_ls = 2 * R.nv;
  J.sky_whole = R.sky_whole ? 1 : 0;
  J.tw1 = R.tw1;
  J.tw2 = R.tw2;
  J.twl_hi = R.twl_hi;
  J.twl_lo = R.twl_lo;
  J.wl_min_bits = ~0ull;
  J.wl_max_bits = 0ull;
  J.wg_min_bits = ~0ull;
  J.wg_max_bits = 0ull;
  J.status = 0;
}

void free_layout(radis_ldm_batch* b) {
  if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  b->graph_valid = false;
  if (b->work) { dev_release(b->ctx, b->work, b->work_bytes); b->ctx->in_use -= (int64_t)b->work_bytes; b->work = nullptr; b->work_bytes = 0; }
  if (b->twmem) { dev_release(b->ctx, b->twmem, b->tw_bytes); b->ctx->in_use -= (int64_t)b->tw_bytes; b->twmem = nullptr; b->tw_bytes = 0; }
  if (b->items_d) { dev_release(b->ctx, b->items_d, b->items_bytes); b->items_d = nullptr; }
}

// Allocate workspace, twiddle tables and work-item lists for every request.
// Workspace of one request (tile width, pass splits and the offsets of its
// buffers from `total`, which it advances); shared by the batch layout and
// radis_ldm_request_bytes.
int layout_request(Req& R, size_t& total) {
  R.tile_w = R.r.tile_w ? R.r.tile_w : choose_tile(R.m_cap);
  if (fill_tile_bytes(R.m_cap, R.tile_w) > (size_t)FILL_LDS_MAX)
    return fail(RADIS_LDM_ERANGE, "LDM of %d rows does not fit a fill tile of width %d (max %d rows at width %d)",
                R.m_cap, R.tile_w, (int)(FILL_LDS_MAX / fill_tile_bytes(1, MIN_TILE_W)), MIN_TILE_W);
  const int64_t V = R.nv + 1;
  R.n_tiles = (int)((V + R.tile_w - 1) / R.tile_w);
  R.stride = (int64_t)align_up((size_t)V, 64);
  const int n1 = R.plan.n1, n2 = R.plan.n2;
  if (R.plan.fc) {
    R.bA = floor_log2(fc_cols_a(n1));
    R.bB = floor_log2(fc_rows_b(n2));
    R.bC = floor_log2(fc_cols_c(n1));
  } else {
    R.bA = std::min({6, floor_log2(LDM_CAP / n1), ctz(n2)});
    R.bB = std::min(floor_log2((n2 <= LDM_CAP_B ? LDM_CAP_B : LDM_CAP) / n2), ctz(n1));
    R.bC = R.bA;
  }
  const int pcap = (R.m_cap + 1) / 2;
  R.off_ldm = total; total = align_up(total + (size_t)R.m_cap * R.stride * 4, 256);
  R.off_y = total;   total = align_up(total + (size_t)pcap * R.plan.L * 8, 256);
  R.off_b = total;   total = align_up(total + (size_t)R.plan.L * 8, 256);
  R.off_rows = total; total = align_up(total + (size_t)R.m_cap * 16, 256);
  // axis partials: at most one per chunk, plus the two partial blocks at the window edges
  R.off_axp = total; total = align_up(total + (size_t)((R.i1 - R.i0 + LDM_AX_CHUNK - 1) / LDM_AX_CHUNK + 2) * 32, 256);
  // circular-image correction: wrap_k floats per L-axis node (N_L <= M/2, as N_G >= 2)
  R.wrap_rows = R.plan.wrap_k ? std::max(1, std::min(R.m_cap / 2, LDM_MAX_AX)) : 0;
  R.off_wrap = total; total = align_up(total + (size_t)R.wrap_rows * R.plan.wrap_k * 4, 256);
  return 0;
}

// Work-item plan of one request (after layout_request): axis items (edge
// chunks scanned line by line, whole blocks through their width candidates
// in groups; none when the window is a single-isotopologue list's every
// line, whose Pareto sets k_finalize evaluates itself), image-correction
// node pairs, fill tiles and the FFT passes' blocks.  Sets R.sky_whole and
// R.n_axp (the axis partials).
LdmItemPlan item_plan(const radis_ldm_db& db, Req& R) {
  LdmItemPlan P{};
  P.i0 = R.i0;
  const bool scan_all = (R.r.flags & RADIS_LDM_SCAN_AXIS) != 0;
  R.sky_whole = db.has_sky && !db.has_iso && R.i0 == 0 && R.i1 == db.n && R.i1 > R.i0 && !scan_all;
  auto chunks = [](int64_t a, int64_t e) { return e > a ? (int)((e - a + LDM_AX_CHUNK - 1) / LDM_AX_CHUNK) : 0; };
  const int64_t fb0 = (R.i0 + LDM_SKY_BLOCK - 1) / LDM_SKY_BLOCK, fb1 = R.i1 / LDM_SKY_BLOCK;
  if (R.sky_whole) {
  } else if (db.has_sky && fb1 > fb0 && !scan_all) {
    P.lo_a = R.i0;
    P.lo_e = fb0 * LDM_SKY_BLOCK;
    P.n_lo = chunks(P.lo_a, P.lo_e);
    // at least LDM_SKY_GROUP_MIN blocks per workgroup (each partial costs
    // the reduction a read), and at most about LDM_SKY_WGS workgroups
    P.G = std::max<int64_t>(LDM_SKY_GROUP_MIN, (fb1 - fb0 + LDM_SKY_WGS - 1) / LDM_SKY_WGS);
    P.g0 = fb0;
    P.fb1 = fb1;
    P.n_g = (int)((fb1 - fb0 + P.G - 1) / P.G);
    P.hi_a = fb1 * LDM_SKY_BLOCK;
    P.hi_e = R.i1;
    P.n_hi = chunks(P.hi_a, P.hi_e);
  } else {
    P.lo_a = R.i0;
    P.lo_e = R.i1;
    P.n_lo = chunks(R.i0, R.i1);
  }
  R.n_axp = P.n_lo + P.n_g + P.n_hi;
  P.n_wrap = (R.wrap_rows + 1) / 2;
  P.n_fill = R.n_tiles;
  P.pcap = (R.m_cap + 1) / 2;
  P.n_colA = R.plan.n2 >> R.bA;
  P.n_rowB = R.plan.n1 >> R.bB;
  P.n_colC = R.plan.n2 >> R.bC;
  return P;
}

int build_layout(radis_ldm_batch* b) {
  free_layout(b);
  b->jobs_last.clear();
  b->tiles_valid = false;
  // --- workspace
  size_t total = 0;
  size_t fill_lds = 0;
  for (Req& R : b->reqs) {
    int rc = layout_request(R, total);
    if (rc) return rc;
    fill_lds = std::max(fill_lds, fill_tile_bytes(R.m_cap, R.tile_w));
  }
  // --- twiddle tables, deduplicated by size
  std::map<int, size_t> sub_off;
  std::map<int64_t, size_t> l_off;
  size_t twsz = 0;
  for (Req& R : b->reqs) {
    for (int N : {R.plan.n1, R.plan.n2})
      if (!sub_off.count(N)) { sub_off[N] = twsz; twsz = align_up(twsz + (size_t)N * 8, 256); }
    if (!l_off.count(R.plan.L)) {
      l_off[R.plan.L] = twsz;
      const size_t nhi = (size_t)((R.plan.L + 1023) >> LDM_TW_LO_BITS);
      twsz = align_up(twsz + (nhi + 1024) * 8, 256);
    }
  }
  if (b->ctx->budget > 0 && b->ctx->in_use + (int64_t)(total + twsz) > b->ctx->budget)
    return fail(RADIS_LDM_ENOMEM, "workspace of %zu bytes exceeds the device budget (%lld in use of %lld)",
                total + twsz, (long long)b->ctx->in_use, (long long)b->ctx->budget);
  b->work = dev_alloc(b->ctx, std::max<size_t>(total, 256));
  if (!b->work) return fail(RADIS_LDM_ENOMEM, "device allocation of %zu workspace bytes failed", total);
  b->work_bytes = std::max<size_t>(total, 256);
  b->ctx->in_use += (int64_t)b->work_bytes;
  b->twmem = dev_alloc(b->ctx, twsz);
  if (!b->twmem) return fail(RADIS_LDM_ENOMEM, "device allocation of %zu twiddle bytes failed", twsz);
  b->tw_bytes = twsz;
  b->ctx->in_use += (int64_t)twsz;
  // tables computed on the device (k_twiddles; the no-allocation entry
  // point builds its tables the same way)
  for (auto& kv : sub_off)
    hipLaunchKernelGGL(k_twiddles, dim3((kv.first + 255) / 256), dim3(256), 0, nullptr,
                       (float2*)((char*)b->twmem + kv.second), kv.first, (float2*)nullptr, 0, (float2*)nullptr, 0ll);
  for (auto& kv : l_off) {
    const long long nt = ((kv.first + 1023) >> LDM_TW_LO_BITS) + 1024;
    hipLaunchKernelGGL(k_twiddles, dim3((unsigned)std::min<long long>((nt + 255) / 256, 4096)), dim3(256), 0, nullptr,
                       (float2*)nullptr, 0, (float2*)nullptr, 0, (float2*)((char*)b->twmem + kv.second),
                       (long long)kv.first);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(nullptr));
  for (Req& R : b->reqs) {
    char* base = (char*)b->twmem;
    R.tw1 = (const float2*)(base + sub_off[R.plan.n1]);
    R.tw2 = (const float2*)(base + sub_off[R.plan.n2]);
    R.twl_hi = (const float2*)(base + l_off[R.plan.L]);
    R.twl_lo = R.twl_hi + (size_t)((R.plan.L + 1023) >> LDM_TW_LO_BITS);
  }
  // --- work items (FFT passes grouped by kernel variant), per request in
  // the order plan_item gives them
  std::vector<int4> ax, fi, wr;
  std::map<int, std::vector<int4>> ma, mb, mc;
  for (int j = 0; j < b->n; ++j) {
    Req& R = b->reqs[j];
    const LdmItemPlan P = item_plan(*b->db, R);
    const int nax = P.n_lo + P.n_g + P.n_hi, bw = nax, bf = bw + P.n_wrap, ba = bf + P.n_fill;
    const int bb = ba + P.pcap * P.n_colA, bc = bb + P.n_rowB, be = bc + P.n_colC;
    const int ca_ = R.plan.fc ? R.plan.n1 : 0;
    const int cb_ = R.plan.fc ? R.plan.n2 : (R.plan.n2 <= LDM_CAP_B ? 0 : -1);
    auto put = [&](std::vector<int4>& v, int a, int e) {
      for (int i = a; i < e; ++i) {
        int4 t = plan_item(P, i);
        t.x = j;
        v.push_back(t);
      }
    };
    put(ax, 0, nax);
    put(wr, bw, bf);
    put(fi, bf, ba);
    put(ma[ca_], ba, bb);
    put(mb[cb_], bb, bc);
    put(mc[ca_], bc, be);
  }
  std::vector<int4> fa, fb, fc;
  auto flatten = [](std::map<int, std::vector<int4>>& m, std::vector<int4>& v,
                    std::vector<radis_ldm_batch::Group>& g) {
    g.clear();
    for (auto& kv : m) {
      g.push_back({kv.first, (int)v.size(), (int)kv.second.size()});
      v.insert(v.end(), kv.second.begin(), kv.second.end());
    }
  };
  flatten(ma, fa, b->grp_a);
  flatten(mb, fb, b->grp_b);
  flatten(mc, fc, b->grp_c);
  b->n_axis = (int)ax.size();
  b->n_wrap = (int)wr.size();
  b->n_fill = (int)fi.size();
  b->n_a = (int)fa.size();
  b->n_b = (int)fb.size();
  b->n_c = (int)fc.size();
  std::vector<int4> all;
  for (auto* v : {&ax, &wr, &fi, &fa, &fb, &fc}) all.insert(all.end(), v->begin(), v->end());
  if (all.empty()) all.push_back(make_int4(0, 0, 0, 0));
  b->items_bytes = all.size() * sizeof(int4);
  b->items_d = (int4*)dev_alloc(b->ctx, b->items_bytes);
  if (!b->items_d) return fail(RADIS_LDM_ENOMEM, "device allocation of work items failed");
  HIPCHK(hipMemcpy(b->items_d, all.data(), all.size() * sizeof(int4), hipMemcpyHostToDevice));
  b->fill_lds = std::max<size_t>(fill_lds, 4);
  return 0;
}

int plan_one(const radis_ldm_db& db, const radis_ldm_request& rq, Req& R) {
  int rc = validate_request(rq);
  if (rc) return rc;
  R.r = rq;
  rc = grid_of(rq.wmin, rq.wmax, rq.wstep, R.nv, R.delta_w);
  if (rc) return rc;
  if (rq.n_v > 0) {
    if (rq.n_v < 2 || rq.n_v > 1000000000) return fail(RADIS_LDM_EINVAL, "n_v %lld out of range", (long long)rq.n_v);
    R.nv = rq.n_v;
  }
  if (rq.fft_n1 && (rq.fft_n1 < 16 || rq.fft_n1 > N1_MAX || rq.fft_n1 % 16))
    return fail(RADIS_LDM_EINVAL, "fft_n1 %d: need 16 | fft_n1 <= %d", rq.fft_n1, N1_MAX);
  if (rq.tile_w && (rq.tile_w < MIN_TILE_W || rq.tile_w > MAX_TILE_W || (rq.tile_w & (rq.tile_w - 1))))
    return fail(RADIS_LDM_EINVAL, "tile_w %d: need a power of 2 in [%d, %d]", rq.tile_w, MIN_TILE_W, MAX_TILE_W);
  rc = plan_fft(R.nv, rq.fft_len, R.plan, rq.fft_n1, !(rq.flags & RADIS_LDM_GENERIC_FFT),
                (rq.flags & RADIS_LDM_PERIOD_L) != 0);
  if (rc) return rc;
  const bool lw = rq.line_hi > rq.line_lo;
  const double lo = lw ? rq.line_lo : rq.wmin, hi = lw ? rq.line_hi : rq.wmax;
  R.i0 = std::lower_bound(db.h_nu0.begin(), db.h_nu0.end(), lo) - db.h_nu0.begin();
  R.i1 = std::upper_bound(db.h_nu0.begin(), db.h_nu0.end(), hi) - db.h_nu0.begin();
  if (R.i1 > R.i0 && R.i1 - R.i0 > (int64_t)2000000000)
    return fail(RADIS_LDM_ERANGE, "window holds more than 2e9 lines");
  return bound_request(db, R);
}

// Host-to-device upload of the job table (outside any graph: a captured
// copy from pinned memory is not guaranteed to re-read the host buffer).
int upload_jobs(radis_ldm_batch* b, hipStream_t s, int slot) {
  const size_t bytes = sizeof(LdmJob) * b->n;
  const char* h = (const char*)b->jobs_h[slot];
  if (b->jobs_last.size() == bytes && std::memcmp(b->jobs_last.data(), h, bytes) == 0) return 0;
  HIPCHK(hipMemcpyAsync(b->jobs_d, h, bytes, hipMemcpyHostToDevice, s));
  b->jobs_last.assign(h, h + bytes);
  return 0;
}

// Results (axis sizes, status, width extremes) are copied back when they are
// asked for, after the runs have completed, rather than after every run.
int fetch_results(radis_ldm_batch* b) {
  if (!b->res_stale) return 0;
  HIPCHK(hipMemcpy(b->res_h[b->last_slot], b->res_d, sizeof(LdmResult) * b->n, hipMemcpyDeviceToHost));
  b->res_stale = false;
  return 0;
}

// Launch one kernel; in a timed run with per-dispatch start/stop events,
// which the runtime stamps on the kernel's own packet.
// One launch of a stage made of several launches: the stage's start event
// is recorded by its first launch and its stop event by its last.
template <typename K, typename... Args>
void launch_part(radis_ldm_batch* b, bool timed, int slot, int kid, bool first, bool last, K kernel, dim3 grid,
                 dim3 block, size_t shmem, hipStream_t s, Args... args) {
  if (timed && b) {
    hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)shmem, s, first ? b->st_ev[slot][2 * kid] : nullptr,
                          last ? b->st_ev[slot][2 * kid + 1] : nullptr, 0, args...);
    b->ev_used[slot][kid] = true;
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
  }
}

template <typename K, typename... Args>
void launch(radis_ldm_batch* b, bool timed, int slot, int kid, K kernel, dim3 grid, dim3 block, size_t shmem,
            hipStream_t s, Args... args) {
  launch_part(b, timed, slot, kid, true, true, kernel, grid, block, shmem, s, args...);
}

// stage of each kernel id: axis, finalize, tile bounds, fill, fft_a, fft_b, fft_c, wrap
const int KERNEL_STAGE[RADIS_LDM_N_STAGES] = {0, 1, 2, 3, 4, 5, 6, 7};

// Device state and work items one launch sequence runs over: a batch's,
// or the no-allocation entry point's single request in its scratch array.
struct LaunchSet {
  LdmJob* jobs;
  LdmAxes* axes;
  LdmResult* res;
  int4* items;  // axis, image-correction, fill, pass A, pass B, pass C items in this order
  int n_req;
  int n_axis, n_wrap, n_fill, n_a, n_b, n_c;
  const LaunchGroup *ga, *gb, *gc;
  int nga, ngb, ngc;
  size_t fill_lds;
  bool tile_bounds;  // run k_tile_bounds (else the items hold valid tile line ranges)
  int n_cu;
};

// The kernel sequence of one run (what a graph captures).  Timed runs (b
// non-null) stamp per-kernel events on the launch stream.
int launch_set(const LaunchSet& X, const LdmDb& dbd, radis_ldm_batch* b, hipStream_t s, bool timed, int slot) {
  const int4* items = X.items;
  const LdmJob* cj = X.jobs;
  const LdmAxes* ca = X.axes;
  if (X.n_axis) launch(b, timed, slot, 0, k_axis, dim3(X.n_axis), dim3(256), 0, s, X.jobs, items, dbd);
  launch(b, timed, slot, 1, k_finalize, dim3(X.n_req), dim3(256), 0, s, X.jobs, X.axes, X.res, dbd);
  items += X.n_axis;
  if (X.n_wrap) launch(b, timed, slot, 7, k_wrap, dim3(X.n_wrap), dim3(256), 0, s, cj, ca, items);
  items += X.n_wrap;
  if (X.n_fill) {
    if (X.tile_bounds)
      launch(b, timed, slot, 2, k_tile_bounds, dim3((X.n_fill + 3) / 4), dim3(256), 0, s, X.jobs,
             X.items + (items - X.items), X.n_fill, dbd);
    launch(b, timed, slot, 3, k_fill, dim3(X.n_fill), dim3(LDM_FILL_T), X.fill_lds, s, cj, ca, items, dbd);
  }
  items += X.n_fill;
  for (int g = 0; g < X.nga; ++g) {
    const auto& G = X.ga[g];
    const int4* it = items + G.first;
    const bool f = g == 0, l = g == X.nga - 1;
    const dim3 grid(G.count);
    // compiled pass A is persistent: one workgroup per CU loops over the column blocks
    const dim3 grid_p(std::min(G.count, X.n_cu));
#define FC_A(N) \
  case N: launch_part(b, timed, slot, 4, f, l, k_fft_a_c<N>, grid_p, dim3(FC_T_A), 0, s, cj, ca, it, G.count); break;
    switch (G.cls) {
      FC_A(256) FC_A(320) FC_A(384) FC_A(512) FC_A(640) FC_A(768) FC_A(1024)
      default: launch_part(b, timed, slot, 4, f, l, k_fft_a, grid, dim3(LDM_FFT_T), 0, s, cj, ca, it, G.count);
    }
#undef FC_A
  }
  items += X.n_a;
  for (int g = 0; g < X.ngb; ++g) {
    const auto& G = X.gb[g];
    const int4* it = items + G.first;
    const bool f = g == 0, l = g == X.ngb - 1;
    const dim3 grid(G.count);
#define FC_B(N) \
  case N: launch_part(b, timed, slot, 5, f, l, k_fft_b_c<N>, grid, dim3(FC_T_B), 0, s, cj, ca, it); break;
    switch (G.cls) {
      FC_B(256) FC_B(512) FC_B(1024) FC_B(2048)
      case 0:
        launch_part(b, timed, slot, 5, f, l, k_fft_b<LDM_CAP_B, LDM_FFT_T_B>, grid, dim3(LDM_FFT_T_B), 0, s, cj, ca,
                    it);
        break;
      default:
        launch_part(b, timed, slot, 5, f, l, k_fft_b<LDM_CAP, LDM_FFT_T>, grid, dim3(LDM_FFT_T), 0, s, cj, ca, it);
    }
#undef FC_B
  }
  items += X.n_b;
  for (int g = 0; g < X.ngc; ++g) {
    const auto& G = X.gc[g];
    const int4* it = items + G.first;
    const bool f = g == 0, l = g == X.ngc - 1;
    const dim3 grid(G.count);
#define FC_C(N) \
  case N: launch_part(b, timed, slot, 6, f, l, k_fft_c_c<N>, grid, dim3(FC_T), 0, s, cj, it, G.count); break;
    switch (G.cls) {
      FC_C(256) FC_C(320) FC_C(384) FC_C(512) FC_C(640) FC_C(768) FC_C(1024)
      default: launch_part(b, timed, slot, 6, f, l, k_fft_c, grid, dim3(LDM_FFT_T), 0, s, cj, it, G.count);
    }
#undef FC_C
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// A batch's launch sequence.
int launch_kernels(radis_ldm_batch* b, hipStream_t s, bool timed, int slot, bool capturing = false) {
  const radis_ldm_db& db = *b->db;
  for (int k = 0; k < radis_ldm_batch::N_KERNELS; ++k) b->ev_used[slot][k] = false;
  LaunchSet X{};
  X.jobs = b->jobs_d;
  X.axes = b->axes_d;
  X.res = b->res_d;
  X.items = b->items_d;
  X.n_req = b->n;
  X.n_axis = b->n_axis;
  X.n_wrap = b->n_wrap;
  X.n_fill = b->n_fill;
  X.n_a = b->n_a;
  X.n_b = b->n_b;
  X.n_c = b->n_c;
  X.ga = b->grp_a.data();
  X.nga = (int)b->grp_a.size();
  X.gb = b->grp_b.data();
  X.ngb = (int)b->grp_b.size();
  X.gc = b->grp_c.data();
  X.ngc = (int)b->grp_c.size();
  X.fill_lds = b->fill_lds;
  // Tile line bounds depend on the grid and, through the pressure-shift
  // margin, on p; without a shift column they are fixed per layout.
  X.tile_bounds = db.has_delta || !b->tiles_valid;
  X.n_cu = b->ctx->n_cu;
  const int rc = launch_set(X, db.dev, b, s, timed, slot);
  // a captured launch has not run yet: only an eager run validates the bounds
  if (!rc && X.n_fill && X.tile_bounds && !capturing) b->tiles_valid = !db.has_delta;
  return rc;
}

int enqueue(radis_ldm_batch* b, hipStream_t s, bool timed, int slot) {
  int rc = upload_jobs(b, s, slot);
  if (!rc) rc = launch_kernels(b, s, timed, slot);
  if (!rc) b->res_stale = true;
  return rc;
}

// Wait for the run that last used `slot` and fold its stage times into the
// accumulators.
int wait_slot(radis_ldm_batch* b, int slot) {
  if (!b->pending[slot]) return 0;
  HIPCHK(hipEventSynchronize(b->ev_done[slot]));
  b->pending[slot] = false;
  if (b->slot_timed[slot]) {
    for (int k = 0; k < radis_ldm_batch::N_KERNELS; ++k) {
      if (!b->ev_used[slot][k]) continue;
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, b->st_ev[slot][2 * k], b->st_ev[slot][2 * k + 1]));
      b->stage_sum[KERNEL_STAGE[k]] += ms;
    }
    b->stage_count++;
    b->slot_timed[slot] = false;
  }
  return 0;
}

int wait_pending(radis_ldm_batch* b) {
  int rc = wait_slot(b, 0);
  if (!rc) rc = wait_slot(b, 1);
  return rc;
}

void refresh_jobs(radis_ldm_batch* b, int slot, const std::vector<float*>& outs) {
  for (int j = 0; j < b->n; ++j) fill_job(*b->db, b->reqs[j], (char*)b->work, outs[j], b->jobs_h[slot][j]);
}

}  // namespace

radis_ldm_ctx::~radis_ldm_ctx() {
  for (auto& kv : dev_cache) (void)hipFree(kv.second);
  for (auto& kv : host_cache) (void)hipHostFree(kv.second);
}

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char* radis_ldm_last_error(void) { return g_err; }
int radis_ldm_abi_version(void) { return RADIS_LDM_ABI_VERSION; }

int radis_ldm_ctx_create(int device, radis_ldm_ctx** out) {
  if (!out) return fail(RADIS_LDM_EINVAL, "null out");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(RADIS_LDM_EINVAL, "device %d not in [0, %d)", device, n);
  HIPCHK(hipSetDevice(device));
  int n_cu = 0;
  HIPCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
  char* scr = nullptr;
  const int rc = scratch_address(device, &scr);
  if (rc) return rc;
  auto* c = new radis_ldm_ctx();
  c->device = device;
  c->scratch = scr;
  if (n_cu > 0) c->n_cu = n_cu;
  *out = c;
  return 0;
}

void radis_ldm_ctx_destroy(radis_ldm_ctx* ctx) {
  if (!ctx) return;
  ctx->closed = true;
  if (ctx->refs == 0) delete ctx;
}

int radis_ldm_ctx_set_budget(radis_ldm_ctx* ctx, int64_t bytes) {
  if (!ctx || bytes < 0) return fail(RADIS_LDM_EINVAL, "bad budget");
  ctx->budget = bytes;
  return 0;
}

int radis_ldm_ctx_bytes_in_use(const radis_ldm_ctx* ctx, int64_t* bytes) {
  if (!ctx || !bytes) return fail(RADIS_LDM_EINVAL, "null argument");
  *bytes = ctx->in_use;
  return 0;
}

int radis_ldm_plan_grid(double wmin, double wmax, double wstep, int64_t* n_v, int64_t* fft_len) {
  int64_t nv;
  double d;
  int rc = grid_of(wmin, wmax, wstep, nv, d);
  if (rc) return rc;
  FftPlan p;
  rc = plan_fft(nv, 0, p);
  if (rc) return rc;
  if (n_v) *n_v = nv;
  if (fft_len) *fft_len = p.L;
  return 0;
}

namespace {

// Pareto-maximal lines of one block for monotonicity direction d (see
// k_axis): d0 maximises (g_air, g_self, -n_air, -n_self), d1 (-g, +n),
// d2 (+g, +n), d3 (-g, -n).  Sort-filter-skyline: visit the lines by
// decreasing coordinate sum (a dominating line never comes later, up to
// rounding of the sum, which can only keep a dominated line too) and keep
// those no kept line dominates.
void block_skyline(const radis_ldm_lines* L, int64_t i0, int cnt, int d, std::vector<float4>& out) {
  const float sg = (d == 0 || d == 2) ? 1.0f : -1.0f;
  const float sn = (d == 1 || d == 2) ? 1.0f : -1.0f;
  struct P { float c[4]; double key; int64_t i; };
  std::vector<P> v((size_t)cnt);
  for (int k = 0; k < cnt; ++k) {
    const int64_t i = i0 + k;
    const float na = L->n_air[i], ns = L->n_self ? L->n_self[i] : na;
    P& p = v[(size_t)k];
    p.c[0] = sg * L->g_air[i]; p.c[1] = sg * L->g_self[i]; p.c[2] = sn * na; p.c[3] = sn * ns;
    p.key = (double)p.c[0] + (double)p.c[1] + (double)p.c[2] + (double)p.c[3];
    p.i = i;
  }
  std::sort(v.begin(), v.end(), [](const P& a, const P& b) { return a.key > b.key; });
  std::vector<const P*> sky;
  for (const P& p : v) {
    bool dom = false;
    for (const P* q : sky)
      if (q->c[0] >= p.c[0] && q->c[1] >= p.c[1] && q->c[2] >= p.c[2] && q->c[3] >= p.c[3]) { dom = true; break; }
    if (!dom) sky.push_back(&p);
  }
  for (const P* q : sky) {
    const int64_t i = q->i;
    const float na = L->n_air[i];
    out.push_back(make_float4(L->g_air[i], L->g_self[i], na, L->n_self ? L->n_self[i] : na));
  }
}

// The Pareto set of points (same directions and dominance as block_skyline).
void point_skyline(std::vector<float4> pts, int d, std::vector<float4>& out) {
  const float sg = (d == 0 || d == 2) ? 1.0f : -1.0f;
  const float sn = (d == 1 || d == 2) ? 1.0f : -1.0f;
  auto key = [&](const float4& q) {
    return (double)(sg * q.x) + (double)(sg * q.y) + (double)(sn * q.z) + (double)(sn * q.w);
  };
  std::sort(pts.begin(), pts.end(), [&](const float4& a, const float4& b) { return key(a) > key(b); });
  out.clear();
  for (const float4& p : pts) {
    bool dom = false;
    for (const float4& q : out)
      if (sg * q.x >= sg * p.x && sg * q.y >= sg * p.y && sn * q.z >= sn * p.z && sn * q.w >= sn * p.w) {
        dom = true;
        break;
      }
    if (!dom) out.push_back(p);
  }
}

}  // namespace

int radis_ldm_db_upload(radis_ldm_ctx* ctx, const radis_ldm_lines* L, radis_ldm_db** out) {
  if (!ctx || !L || !out) return fail(RADIS_LDM_EINVAL, "null argument");
  if (L->n <= 0) return fail(RADIS_LDM_EINVAL, "empty line list");
  // the fill addresses line columns with 32-bit byte offsets (8-byte nu0)
  if (L->n > ((int64_t)1 << 29)) return fail(RADIS_LDM_ERANGE, "at most 2^29 lines per database (got %lld)", (long long)L->n);
  if (!L->nu0 || !L->S_ref || !L->El || !L->g_air || !L->g_self || !L->n_air || !L->molar_mass)
    return fail(RADIS_LDM_EINVAL, "missing required column");
  if (L->n_iso_slots < 2 || L->n_iso_slots > RADIS_LDM_MAX_ISO)
    return fail(RADIS_LDM_EINVAL, "n_iso_slots must be in [2, %d]", RADIS_LDM_MAX_ISO);
  const int64_t n = L->n;
  for (int64_t i = 1; i < n; ++i)
    if (!(L->nu0[i - 1] <= L->nu0[i])) return fail(RADIS_LDM_EINVAL, "nu0 not sorted ascending at line %lld", (long long)i);
  if (!std::isfinite(L->nu0[0]) || !std::isfinite(L->nu0[n - 1])) return fail(RADIS_LDM_EINVAL, "non-finite nu0");
  HIPCHK(hipSetDevice(ctx->device));
  auto* db = new radis_ldm_db();
  db->ctx = nullptr;  // attached (and counted) once the device copy exists
  db->n = n;
  db->has_iso = L->iso != nullptr;
  db->has_nself = L->n_self != nullptr;
  db->has_delta = L->delta != nullptr;
  if ((L->Evib_l == nullptr) != (L->Evib_u == nullptr)) {
    delete db;
    return fail(RADIS_LDM_EINVAL, "Evib_l and Evib_u come together");
  }
  db->has_evib = L->Evib_l != nullptr;
  db->n_iso_slots = L->n_iso_slots;
  for (int s = 0; s < L->n_iso_slots; ++s) db->molar_mass[s] = L->molar_mass[s];
  if (db->has_iso) {
    for (int64_t i = 0; i < n; ++i) {
      const int s = L->iso[i];
      if (s <= 0 || s >= L->n_iso_slots || !(L->molar_mass[s] > 0)) {
        delete db;
        return fail(RADIS_LDM_EINVAL, "line %lld: iso slot %d has no molar mass", (long long)i, s);
      }
      db->iso_mask |= 1u << s;
    }
  } else {
    if (!(L->molar_mass[1] > 0)) { delete db; return fail(RADIS_LDM_EINVAL, "molar_mass[1] must be positive"); }
    db->iso_mask = 2u;
  }
  if (db->has_delta)
    for (int64_t i = 0; i < n; ++i) db->max_abs_delta = std::max(db->max_abs_delta, (double)std::fabs(L->delta[i]));
  db->h_nu0.assign(L->nu0, L->nu0 + n);
  const int nb = (int)((n + STAT_BLOCK - 1) / STAT_BLOCK);
  for (auto* v : {&db->ga_min, &db->gs_min, &db->na_min, &db->ns_min, &db->e_min}) v->assign(nb, 3.4e38f);
  for (auto* v : {&db->ga_max, &db->gs_max, &db->na_max, &db->ns_max, &db->e_max}) v->assign(nb, -3.4e38f);
  db->s_max.assign(nb, 0.0f);
  for (int64_t i = 0; i < n; ++i) {
    const int b = (int)(i / STAT_BLOCK);
    const float ns = L->n_self ? L->n_self[i] : L->n_air[i];
    db->ga_min[b] = std::min(db->ga_min[b], L->g_air[i]); db->ga_max[b] = std::max(db->ga_max[b], L->g_air[i]);
    db->gs_min[b] = std::min(db->gs_min[b], L->g_self[i]); db->gs_max[b] = std::max(db->gs_max[b], L->g_self[i]);
    db->na_min[b] = std::min(db->na_min[b], L->n_air[i]); db->na_max[b] = std::max(db->na_max[b], L->n_air[i]);
    db->ns_min[b] = std::min(db->ns_min[b], ns); db->ns_max[b] = std::max(db->ns_max[b], ns);
    db->e_min[b] = std::min(db->e_min[b], L->El[i]); db->e_max[b] = std::max(db->e_max[b], L->El[i]);
    db->s_max[b] = std::max(db->s_max[b], L->S_ref[i]);
  }
  if (db->has_evib) {
    db->dv_min.assign(nb, 3.4e38f);
    db->dv_max.assign(nb, -3.4e38f);
    for (int64_t i = 0; i < n; ++i) {
      const int b = (int)(i / STAT_BLOCK);
      const float el = L->El[i], evl = L->Evib_l[i], evu = L->Evib_u[i];
      if (!(evl >= 0 && evl <= el * (1 + 1e-6f) + 1e-3f && evu >= 0 &&
            evu <= (el + (float)L->nu0[i]) * (1 + 1e-6f) + 1e-3f)) {
        delete db;
        return fail(RADIS_LDM_EINVAL, "line %lld: need 0 <= Evib_l <= El and 0 <= Evib_u <= El + nu0", (long long)i);
      }
      const float d = evu - evl;
      db->dv_min[b] = std::min(db->dv_min[b], d);
      db->dv_max[b] = std::max(db->dv_max[b], d);
    }
  }
  // Lorentzian-width candidates of the whole blocks (k_axis).  Only for
  // finite, non-negative widths and exponents: otherwise the monotonicity the
  // candidates rest on fails, and every line is scanned.
  bool sky_ok = true;
  for (int64_t i = 0; i < n && sky_ok; ++i) {
    const float na = L->n_air[i], ns = L->n_self ? L->n_self[i] : na;
    sky_ok = std::isfinite(na) && std::isfinite(ns) && L->g_air[i] >= 0 && L->g_self[i] >= 0 &&
             std::isfinite(L->g_air[i]) && std::isfinite(L->g_self[i]) && L->nu0[i] >= 0;
  }
  std::vector<float4> sky_vals;
  std::vector<int> sky_off;
  std::vector<int2> iso_rng;
  const int nfull = (int)(n / LDM_SKY_BLOCK);
  if (sky_ok && nfull > 0) {
    std::vector<std::vector<float4>> per((size_t)nfull * 4);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; ++t)
      th.emplace_back([&, t] {
        for (int b = (int)t; b < nfull; b += (int)nth)
          for (int d = 0; d < 4; ++d) block_skyline(L, (int64_t)b * LDM_SKY_BLOCK, LDM_SKY_BLOCK, d, per[(size_t)b * 4 + d]);
      });
    for (auto& x : th) x.join();
    sky_off.assign((size_t)4 * (nfull + 1), 0);
    for (int d = 0; d < 4; ++d) {
      sky_off[(size_t)d * (nfull + 1)] = (int)sky_vals.size();
      for (int b = 0; b < nfull; ++b) {
        const auto& v = per[(size_t)b * 4 + d];
        sky_vals.insert(sky_vals.end(), v.begin(), v.end());
        sky_off[(size_t)d * (nfull + 1) + b + 1] = (int)sky_vals.size();
      }
    }
    // the whole list's Pareto sets: every globally Pareto-maximal line is
    // maximal in its block, so the blocks' sets plus the tail's hold them
    std::vector<std::thread> tg;
    for (int d = 0; d < 4; ++d)
      tg.emplace_back([&, d] {
        std::vector<float4> pts(sky_vals.begin() + sky_off[(size_t)d * (nfull + 1)],
                                sky_vals.begin() + sky_off[(size_t)d * (nfull + 1) + nfull]);
        if ((int64_t)nfull * LDM_SKY_BLOCK < n)
          block_skyline(L, (int64_t)nfull * LDM_SKY_BLOCK, (int)(n - (int64_t)nfull * LDM_SKY_BLOCK), d, pts);
        point_skyline(std::move(pts), d, db->h_gsky[d]);
      });
    for (auto& x : tg) x.join();
    for (int d = 0; d < 4; ++d) {
      db->dev.gsky_off[d] = (int)sky_vals.size();
      sky_vals.insert(sky_vals.end(), db->h_gsky[d].begin(), db->h_gsky[d].end());
    }
    db->dev.gsky_off[4] = (int)sky_vals.size();
    if (db->has_iso) {
      iso_rng.assign((size_t)nfull * L->n_iso_slots, make_int2(-1, -1));
      for (int64_t i = 0; i < (int64_t)nfull * LDM_SKY_BLOCK; ++i) {
        int2& r = iso_rng[(size_t)(i / LDM_SKY_BLOCK) * L->n_iso_slots + L->iso[i]];
        if (r.x < 0) r.x = (int)i;
        r.y = (int)i;
      }
    }
    db->has_sky = true;
    db->h_sky_off = sky_off;
  }
  // device arrays: one allocation
  size_t off[15], tot = 0;
  const size_t sz[15] = {(size_t)n * 8, (size_t)n * 4, (size_t)n * 4, (size_t)n * 4, (size_t)n * 4, (size_t)n * 4,
                         db->has_nself ? (size_t)n * 4 : 0, db->has_delta ? (size_t)n * 4 : 0,
                         db->has_iso ? (size_t)n : 0, 8,
                         db->has_evib ? (size_t)n * 4 : 0, db->has_evib ? (size_t)n * 4 : 0,
                         sky_vals.size() * sizeof(float4), sky_off.size() * sizeof(int), iso_rng.size() * sizeof(int2)};
  for (int k = 0; k < 15; ++k) { off[k] = tot; tot = align_up(tot + sz[k], 256); }
  if (ctx->budget > 0 && ctx->in_use + (int64_t)tot > ctx->budget) {
    delete db;
    return fail(RADIS_LDM_ENOMEM, "line list of %zu bytes exceeds the device budget", tot);
  }
  if (hipMalloc(&db->mem, tot) != hipSuccess) {
    delete db;
    return fail(RADIS_LDM_ENOMEM, "hipMalloc of %zu bytes for the line list failed", tot);
  }
  db->bytes = tot;
  db->ctx = ctx;
  ctx->refs++;
  ctx->in_use += (int64_t)tot;
  char* m = (char*)db->mem;
  // defaults block: a float 0 (missing delta column) and a byte 1 (missing iso column)
  unsigned char dflt[8] = {0, 0, 0, 0, 1, 0, 0, 0};
  const void* src[15] = {L->nu0, L->S_ref, L->El, L->g_air, L->g_self, L->n_air, L->n_self, L->delta, L->iso, dflt,
                         L->Evib_l, L->Evib_u, sky_vals.data(), sky_off.data(), iso_rng.data()};
  for (int k = 0; k < 15; ++k)
    if (sz[k]) {
      hipError_t e = hipMemcpy(m + off[k], src[k], sz[k], hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        radis_ldm_db_destroy(db);
        return fail(RADIS_LDM_EHIP, "line list upload: %s", hipGetErrorString(e));
      }
    }
  db->dev.nu0 = (const double*)(m + off[0]);
  db->dev.S_ref = (const float*)(m + off[1]);
  db->dev.El = (const float*)(m + off[2]);
  db->dev.g_air = (const float*)(m + off[3]);
  db->dev.g_self = (const float*)(m + off[4]);
  db->dev.n_air = (const float*)(m + off[5]);
  db->dev.n_self = db->has_nself ? (const float*)(m + off[6]) : nullptr;
  db->dev.delta = db->has_delta ? (const float*)(m + off[7]) : nullptr;
  db->dev.iso = db->has_iso ? (const uint8_t*)(m + off[8]) : nullptr;
  db->dev.dflt = (const float*)(m + off[9]);
  db->dev.evib_l = db->has_evib ? (const float*)(m + off[10]) : nullptr;
  db->dev.evib_u = db->has_evib ? (const float*)(m + off[11]) : nullptr;
  db->dev.sky = db->has_sky ? (const float4*)(m + off[12]) : nullptr;
  db->dev.sky_off = db->has_sky ? (const int*)(m + off[13]) : nullptr;
  db->dev.sky_nb = db->has_sky ? nfull : 0;
  db->dev.n_iso_slots = L->n_iso_slots;
  db->dev.iso_rng = (db->has_sky && db->has_iso) ? (const int2*)(m + off[14]) : nullptr;
  *out = db;
  return 0;
}

void radis_ldm_db_destroy(radis_ldm_db* db) {
  if (!db) return;
  db->closed = true;
  if (db->refs == 0) db_free(db);
}

int radis_ldm_db_size(const radis_ldm_db* db, int64_t* n_lines) {
  if (!db || !n_lines) return fail(RADIS_LDM_EINVAL, "null argument");
  *n_lines = db->n;
  return 0;
}

int radis_ldm_db_set_qspline(radis_ldm_db* db, int iso, const double* knots, int32_t n_knots, const double* coefs,
                             int32_t degree, double T_lo, double T_hi) {
  if (!db || !knots || !coefs) return fail(RADIS_LDM_EINVAL, "null argument");
  if (iso < 0 || iso >= LDM_MAX_ISO) return fail(RADIS_LDM_EINVAL, "isotopologue slot %d not in [0, %d)", iso, LDM_MAX_ISO);
  if (degree < 1 || degree > LDM_QS_MAXDEG) return fail(RADIS_LDM_EINVAL, "spline degree %d not in [1, %d]", degree, LDM_QS_MAXDEG);
  if (n_knots < 2 * (degree + 1) || n_knots > LDM_QS_MAXK)
    return fail(RADIS_LDM_EINVAL, "%d knots: need %d to %d for degree %d", n_knots, 2 * (degree + 1), LDM_QS_MAXK, degree);
  if (!(T_lo < T_hi) || !std::isfinite(T_lo) || !std::isfinite(T_hi) || !(T_lo > 0))
    return fail(RADIS_LDM_EINVAL, "bad temperature range [%g, %g]", T_lo, T_hi);
  LdmQSpline q{};
  q.n = n_knots;
  q.k = degree;
  q.lo = T_lo;
  q.hi = T_hi;
  for (int i = 0; i < n_knots; ++i) {
    if (!std::isfinite(knots[i]) || (i > 0 && knots[i] < knots[i - 1]))
      return fail(RADIS_LDM_EINVAL, "knots must be finite and non-decreasing (knot %d)", i);
    q.t[i] = knots[i];
  }
  if (!(knots[degree + 1] > knots[degree]) || !(knots[n_knots - degree - 1] > knots[n_knots - degree - 2]))
    return fail(RADIS_LDM_EINVAL, "empty first or last knot interval");
  for (int i = 0; i < n_knots - degree - 1; ++i) {
    if (!std::isfinite(coefs[i])) return fail(RADIS_LDM_EINVAL, "coefficient %d is not finite", i);
    q.c[i] = coefs[i];
  }
  HIPCHK(hipSetDevice(db->ctx->device));
  if (!db->qs_mem) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, sizeof(LdmQSpline) * LDM_MAX_ISO));
    db->qs_mem = (LdmQSpline*)p;
    db->ctx->in_use += (int64_t)(sizeof(LdmQSpline) * LDM_MAX_ISO);
    db->h_qs.assign(LDM_MAX_ISO, LdmQSpline{});
  }
  HIPCHK(hipMemcpy(db->qs_mem + iso, &q, sizeof q, hipMemcpyHostToDevice));
  db->h_qs[iso] = q;
  db->dev.qs = db->qs_mem;
  db->dev.qs_mask |= 1u << iso;
  return 0;
}

int radis_ldm_qspline_eval(const radis_ldm_db* db, int iso, const double* T, int64_t n, double* Q, void* stream) {
  if (!db) return fail(RADIS_LDM_EINVAL, "null database");
  if (iso < 0 || iso >= LDM_MAX_ISO || !(db->dev.qs_mask & (1u << iso)))
    return fail(RADIS_LDM_EINVAL, "isotopologue slot %d has no partition-function spline", iso);
  if (n < 0 || (n > 0 && (!T || !Q))) return fail(RADIS_LDM_EINVAL, "bad arguments");
  if (n == 0) return 0;
  HIPCHK(hipSetDevice(db->ctx->device));
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_qspline_eval, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const LdmQSpline*)db->qs_mem,
                     iso, T, (long long)n, Q);
  HIPCHK(hipGetLastError());
  return 0;
}

int radis_ldm_request_bytes(const radis_ldm_db* db, const radis_ldm_request* req, int64_t* bytes,
                            radis_ldm_request_info* info) {
  if (!db || !req || !bytes) return fail(RADIS_LDM_EINVAL, "null argument");
  Req R;
  int rc = plan_one(*db, *req, R);
  if (rc) return rc;
  R.m_cap = R.m_need;
  size_t total = 0;
  rc = layout_request(R, total);
  if (rc) return rc;
  // the batch's work items and job table (16 B per item, a few per tile and block)
  int64_t b = (int64_t)total + (int64_t)R.n_tiles * 16;
  // twiddle tables, plus alignment slack: an upper bound of what a batch of
  // this one request allocates
  b += (int64_t)(R.plan.n1 + R.plan.n2 + 1024 + (R.plan.L >> LDM_TW_LO_BITS) + 1) * 8 + 64 * 1024;
  *bytes = b;
  if (info) {
    std::memset(info, 0, sizeof *info);
    info->n_v = R.nv;
    info->fft_len = R.plan.L;
    info->n1 = R.plan.n1;
    info->n2 = R.plan.n2;
    info->compiled_fft = R.plan.fc ? 1 : 0;
    info->line_begin = R.i0;
    info->line_end = R.i1;
    info->m_cap = R.m_cap;
    info->tile_w = R.tile_w;
    info->scale_log2 = R.scale_log2;
  }
  return 0;
}

int radis_ldm_batch_create(radis_ldm_ctx* ctx, const radis_ldm_db* db, int n, const radis_ldm_request* reqs,
                           radis_ldm_batch** out) {
  if (!ctx || !db || !reqs || !out || n <= 0) return fail(RADIS_LDM_EINVAL, "bad batch arguments");
  if (db->ctx != ctx) return fail(RADIS_LDM_EINVAL, "line list belongs to another context");
  HIPCHK(hipSetDevice(ctx->device));
  auto* b = new radis_ldm_batch();
  b->ctx = ctx;
  b->db = const_cast<radis_ldm_db*>(db);
  ctx->refs++;
  b->db->refs++;
  b->n = n;
  b->reqs.resize(n);
  b->outs.assign(n, nullptr);
  int rc = 0;
  for (int j = 0; j < n && !rc; ++j) {
    rc = plan_one(*db, reqs[j], b->reqs[j]);
    b->reqs[j].m_cap = b->reqs[j].m_need;
  }
  if (!rc) rc = build_layout(b);
  if (!rc) {
    b->jobs_d = (LdmJob*)dev_alloc(ctx, sizeof(LdmJob) * n);
    b->axes_d = (LdmAxes*)dev_alloc(ctx, sizeof(LdmAxes) * n);
    b->res_d = (LdmResult*)dev_alloc(ctx, sizeof(LdmResult) * n);
    bool ok = b->jobs_d && b->axes_d && b->res_d;
    for (int k = 0; k < 2 && ok; ++k) {
      b->jobs_h[k] = (LdmJob*)host_alloc(ctx, sizeof(LdmJob) * n);
      b->res_h[k] = (LdmResult*)host_alloc(ctx, sizeof(LdmResult) * n);
      ok = b->jobs_h[k] && b->res_h[k] &&
           hipEventCreateWithFlags(&b->ev_done[k], hipEventDisableTiming) == hipSuccess;
      if (ok) std::memset(b->res_h[k], 0, sizeof(LdmResult) * n);
      for (auto& e : b->st_ev[k])
        if (ok && hipEventCreate(&e) != hipSuccess) ok = false;
    }
    if (!ok) rc = fail(RADIS_LDM_ENOMEM, "batch bookkeeping allocation failed");
  }
  if (rc) {
    char msg[sizeof g_err];
    std::memcpy(msg, g_err, sizeof msg);
    radis_ldm_batch_destroy(b);
    std::memcpy(g_err, msg, sizeof msg);
    return rc;
  }
  *out = b;
  return 0;
}

int radis_ldm_batch_set_request(radis_ldm_batch* b, int i, const radis_ldm_request* req) {
  if (!b || !req || i < 0 || i >= b->n) return fail(RADIS_LDM_EINVAL, "bad arguments");
  Req& R = b->reqs[i];
  if (req->wmin != R.r.wmin || req->wmax != R.r.wmax || req->wstep != R.r.wstep ||
      (req->fft_len != 0 && req->fft_len != R.plan.L) || req->n_v != R.r.n_v ||
      req->line_lo != R.r.line_lo || req->line_hi != R.r.line_hi || req->fft_n1 != R.r.fft_n1 ||
      req->tile_w != R.r.tile_w)
    return fail(RADIS_LDM_EINVAL, "set_request cannot change the grid; create a new batch");
  // flags fixed by the plan and the work items (FFT variant, axis items, period)
  const int plan_flags = RADIS_LDM_GENERIC_FFT | RADIS_LDM_SCAN_AXIS | RADIS_LDM_PERIOD_L;
  if ((req->flags ^ R.r.flags) & plan_flags)
    return fail(RADIS_LDM_EINVAL, "set_request cannot change RADIS_LDM_GENERIC_FFT, _SCAN_AXIS or _PERIOD_L; "
                                  "create a new batch");
  int rc = validate_request(*req);
  if (rc) return rc;
  rc = wait_pending(b);
  if (rc) return rc;
  Req T = R;
  T.r = *req;
  T.r.fft_len = R.plan.L;
  rc = bound_request(*b->db, T);
  if (rc) return rc;
  const bool grow = T.m_need > R.m_cap;
  R = T;
  if (grow) {
    R.m_cap = std::max(T.m_need, (int)std::ceil(1.5 * R.m_cap));  // vulkan_compute_lib.py:1308-1313
    rc = build_layout(b);
    if (rc) return rc;
  }
  return 0;
}

int radis_ldm_batch_run(radis_ldm_batch* b, float* const* out_dev, void* stream) {
  if (!b || !out_dev) return fail(RADIS_LDM_EINVAL, "null argument");
  for (int j = 0; j < b->n; ++j)
    if (!out_dev[j]) return fail(RADIS_LDM_EINVAL, "null output pointer for request %d", j);
  HIPCHK(hipSetDevice(b->ctx->device));
  const int slot = b->next_slot;
  int rc = wait_slot(b, slot);
  if (rc) return rc;
  for (int j = 0; j < b->n; ++j) b->outs[j] = out_dev[j];
  refresh_jobs(b, slot, b->outs);
  hipStream_t s = (hipStream_t)stream;
  rc = enqueue(b, s, b->timing, slot);
  if (rc) return rc;
  b->slot_timed[slot] = b->timing;
  HIPCHK(hipEventRecord(b->ev_done[slot], s));
  b->pending[slot] = true;
  b->last_slot = slot;
  b->next_slot = slot ^ 1;
  return 0;
}

int radis_ldm_batch_set_derived(radis_ldm_batch* b, int i, const radis_ldm_derived* d) {
  if (!b) return fail(RADIS_LDM_EINVAL, "null batch");
  if (i < 0 || i >= b->n) return fail(RADIS_LDM_EINVAL, "request index %d out of range [0, %d)", i, b->n);
  if (d && !(d->path_length >= 0.0)) return fail(RADIS_LDM_EINVAL, "path_length must be >= 0");
  b->reqs[i].derived = d ? *d : radis_ldm_derived{};
  return 0;
}

int radis_ldm_batch_graph_capture(radis_ldm_batch* b, float* const* out_dev, void* stream) {
  if (!b || !out_dev) return fail(RADIS_LDM_EINVAL, "null argument");
  HIPCHK(hipSetDevice(b->ctx->device));
  int rc = wait_pending(b);
  if (rc) return rc;
  if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  b->graph_outs.assign(out_dev, out_dev + b->n);
  hipStream_t s = (hipStream_t)stream;
  if (!s) {
    if (!b->own_stream) HIPCHK(hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking));
    s = b->own_stream;
  }
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  rc = launch_kernels(b, s, false, 0, true);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(RADIS_LDM_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  b->graph = g;
  HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
  b->graph_valid = true;
  return 0;
}

int radis_ldm_batch_graph_replay(radis_ldm_batch* b, void* stream) {
  if (!b) return fail(RADIS_LDM_EINVAL, "null batch");
  if (!b->graph_valid || !b->gexec) return fail(RADIS_LDM_EINVAL, "no valid captured graph (capture again after a re-plan)");
  HIPCHK(hipSetDevice(b->ctx->device));
  const int slot = b->next_slot;
  int rc = wait_slot(b, slot);
  if (rc) return rc;
  refresh_jobs(b, slot, b->graph_outs);
  hipStream_t s = (hipStream_t)stream;
  rc = upload_jobs(b, s, slot);
  if (rc) return rc;
  HIPCHK(hipGraphLaunch(b->gexec, s));
  b->res_stale = true;
  b->slot_timed[slot] = false;
  HIPCHK(hipEventRecord(b->ev_done[slot], s));
  b->pending[slot] = true;
  b->last_slot = slot;
  b->next_slot = slot ^ 1;
  return 0;
}

int radis_ldm_batch_info(radis_ldm_batch* b, int i, radis_ldm_request_info* info, void* stream) {
  (void)stream;
  if (!b || !info || i < 0 || i >= b->n) return fail(RADIS_LDM_EINVAL, "bad arguments");
  int rc = wait_pending(b);
  if (rc) return rc;
  const Req& R = b->reqs[i];
  info->n_v = R.nv;
  info->fft_len = R.plan.L;
  info->n1 = R.plan.n1;
  info->n2 = R.plan.n2;
  info->compiled_fft = R.plan.fc ? 1 : 0;
  info->line_begin = R.i0;
  info->line_end = R.i1;
  info->m_cap = R.m_cap;
  info->tile_w = R.tile_w;
  info->scale_log2 = R.scale_log2;
  rc = fetch_results(b);
  if (rc) return rc;
  const LdmResult& res = b->res_h[b->last_slot][i];
  info->n_g = res.ng;
  info->n_l = res.nl;
  info->status = res.status;
  info->wg_min = res.wg_min;
  info->wg_max = res.wg_max;
  info->wl_min = res.wl_min;
  info->wl_max = res.wl_max;
  // what k_axis reads for this request: scanned lines and width candidates
  {
    const radis_ldm_db& db = *b->db;
    const int64_t fb0 = (R.i0 + LDM_SKY_BLOCK - 1) / LDM_SKY_BLOCK, fb1 = R.i1 / LDM_SKY_BLOCK;
    const int dmax = std::log(R.r.Tref / R.r.Tgas) > 0.0 ? 2 : 0;
    if (R.sky_whole) {
      info->axis_candidates = db.dev.gsky_off[dmax + 2] - db.dev.gsky_off[dmax];
      info->axis_scanned = 0;
    } else if (db.has_sky && fb1 > fb0 && !(R.r.flags & RADIS_LDM_SCAN_AXIS)) {
      const int nb1 = db.dev.sky_nb + 1;
      int64_t cand = 0;
      for (int d = dmax; d <= dmax + 1; ++d) cand += db.h_sky_off[(size_t)d * nb1 + fb1] - db.h_sky_off[(size_t)d * nb1 + fb0];
      info->axis_candidates = cand;
      info->axis_scanned = (R.i1 - R.i0) - (fb1 - fb0) * LDM_SKY_BLOCK;
    } else {
      info->axis_candidates = 0;
      info->axis_scanned = R.i1 - R.i0;
    }
  }
  return 0;
}

int radis_ldm_batch_workspace_bytes(const radis_ldm_batch* b, int64_t* bytes) {
  if (!b || !bytes) return fail(RADIS_LDM_EINVAL, "null argument");
  *bytes = (int64_t)(b->work_bytes + b->tw_bytes);
  return 0;
}

int radis_ldm_batch_enable_timing(radis_ldm_batch* b, int enable) {
  if (!b) return fail(RADIS_LDM_EINVAL, "null batch");
  b->timing = enable != 0;
  return 0;
}

int radis_ldm_batch_stage_ms(radis_ldm_batch* b, float* ms_out) {
  if (!b || !ms_out) return fail(RADIS_LDM_EINVAL, "null argument");
  int rc = wait_pending(b);
  if (rc) return rc;
  if (b->stage_count == 0) return fail(RADIS_LDM_EINVAL, "no timed run available (enable timing and run)");
  for (int k = 0; k < RADIS_LDM_N_STAGES; ++k) {
    ms_out[k] = (float)(b->stage_sum[k] / b->stage_count);
    b->stage_sum[k] = 0;
  }
  b->stage_count = 0;
  return 0;
}

int radis_ldm_batch_line_terms(radis_ldm_batch* b, int i, int64_t* k0, int32_t* l0, int32_t* m0, float* tv,
                               float* tg, float* tl, float* I, void* stream) {
  if (!b || i < 0 || i >= b->n || !k0 || !l0 || !m0 || !tv || !tg || !tl || !I)
    return fail(RADIS_LDM_EINVAL, "bad arguments");
  if (!b->outs[i]) return fail(RADIS_LDM_EINVAL, "run the batch first");
  HIPCHK(hipSetDevice(b->ctx->device));
  const Req& R = b->reqs[i];
  const int64_t n = R.i1 - R.i0;
  if (n <= 0) return 0;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_line_terms, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const LdmJob*)b->jobs_d,
                     (const LdmAxes*)b->axes_d, i, b->db->dev, (long long*)k0, (int*)l0, (int*)m0, tv, tg, tl, I);
  HIPCHK(hipGetLastError());
  return 0;
}

int radis_ldm_batch_copy_ldm(radis_ldm_batch* b, int i, float* dst, void* stream) {
  if (!b || i < 0 || i >= b->n || !dst) return fail(RADIS_LDM_EINVAL, "bad arguments");
  int rc = wait_pending(b);
  if (rc) return rc;
  const Req& R = b->reqs[i];
  rc = fetch_results(b);
  if (rc) return rc;
  const int rows = b->res_h[b->last_slot][i].ng * b->res_h[b->last_slot][i].nl;
  if (rows == 0) return 0;
  const size_t cols = (size_t)(R.nv + 1);
  HIPCHK(hipMemcpy2DAsync(dst, cols * 4, (char*)b->work + R.off_ldm, (size_t)R.stride * 4, cols * 4, rows,
                          hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return 0;
}

static_assert(sizeof(radis_ldm_result) == sizeof(LdmResult), "radis_ldm_result mirrors LdmResult");
static_assert(sizeof(LdmJob) <= 2048, "the job table travels as a kernel argument");

int radis_ldm_run_noalloc(radis_ldm_ctx* ctx, const radis_ldm_db* db, const radis_ldm_request* req, float* out_dev,
                          radis_ldm_result* result_dev, radis_ldm_request_info* info, void* stream) {
  if (!ctx || !db || !req || !out_dev) return fail(RADIS_LDM_EINVAL, "null argument");
  if (db->ctx != ctx) return fail(RADIS_LDM_EINVAL, "line list belongs to another context");
  if (!ctx->scratch) return fail(RADIS_LDM_EINVAL, "context has no scratch array");
  HIPCHK(hipSetDevice(ctx->device));
  // plan (host, nothing allocated): sizes, workspace layout, work items
  Req R;
  int rc = plan_one(*db, *req, R);
  if (rc) return rc;
  R.m_cap = R.m_need;
  size_t work = 0;
  rc = layout_request(R, work);
  if (rc) return rc;
  const LdmItemPlan P = item_plan(*db, R);
  const int n_items = plan_items(P);
  // scratch layout
  size_t off = 0;
  auto take = [&](size_t n) {
    const size_t o = off;
    off = align_up(off + n, 256);
    return o;
  };
  const size_t o_job = take(sizeof(LdmJob)), o_axes = take(sizeof(LdmAxes)), o_res = take(sizeof(LdmResult));
  const size_t o_items = take((size_t)n_items * sizeof(int4));
  const long long nhi = (R.plan.L + 1023) >> LDM_TW_LO_BITS;
  const size_t o_tw1 = take((size_t)R.plan.n1 * 8), o_tw2 = take((size_t)R.plan.n2 * 8);
  const size_t o_twl = take((size_t)(nhi + 1024) * 8);
  const size_t o_work = take(work);
  if (off > (size_t)RADIS_LDM_NOALLOC_BYTES)
    return fail(RADIS_LDM_ERANGE, "request needs %zu bytes of scratch, above the %lld of radis_ldm_run_noalloc "
                                  "(use a batch)", off, (long long)RADIS_LDM_NOALLOC_BYTES);
  char* const base = ctx->scratch;
  float2* tw1 = (float2*)(base + o_tw1);
  float2* tw2 = (float2*)(base + o_tw2);
  float2* twl = (float2*)(base + o_twl);
  R.tw1 = tw1;
  R.tw2 = tw2;
  R.twl_hi = twl;
  R.twl_lo = twl + nhi;
  LdmJob J;
  fill_job(*db, R, base + o_work, out_dev, J);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_put_job, dim3(1), dim3(64), 0, s, J, (LdmJob*)(base + o_job));
  const long long n_tw = R.plan.n1 + R.plan.n2 + nhi + 1024;
  hipLaunchKernelGGL(k_twiddles, dim3((unsigned)std::min<long long>((n_tw + 255) / 256, 4096)), dim3(256), 0, s, tw1,
                     R.plan.n1, tw2, R.plan.n2, twl, (long long)R.plan.L);
  hipLaunchKernelGGL(k_make_items, dim3((n_items + 255) / 256), dim3(256), 0, s, (int4*)(base + o_items), P, n_items);
  HIPCHK(hipGetLastError());
  const int cls_ac = R.plan.fc ? R.plan.n1 : 0;
  const int cls_b = R.plan.fc ? R.plan.n2 : (R.plan.n2 <= LDM_CAP_B ? 0 : -1);
  const LaunchGroup ga{cls_ac, 0, P.pcap * P.n_colA}, gb{cls_b, 0, P.n_rowB}, gc{cls_ac, 0, P.n_colC};
  LaunchSet X{};
  X.jobs = (LdmJob*)(base + o_job);
  X.axes = (LdmAxes*)(base + o_axes);
  X.res = result_dev ? (LdmResult*)result_dev : (LdmResult*)(base + o_res);
  X.items = (int4*)(base + o_items);
  X.n_req = 1;
  X.n_axis = P.n_lo + P.n_g + P.n_hi;
  X.n_wrap = P.n_wrap;
  X.n_fill = P.n_fill;
  X.n_a = ga.count;
  X.n_b = gb.count;
  X.n_c = gc.count;
  X.ga = &ga;
  X.nga = 1;
  X.gb = &gb;
  X.ngb = 1;
  X.gc = &gc;
  X.ngc = 1;
  X.fill_lds = std::max<size_t>(fill_tile_bytes(R.m_cap, R.tile_w), 4);
  X.tile_bounds = true;
  X.n_cu = ctx->n_cu;
  rc = launch_set(X, db->dev, nullptr, s, false, 0);
  if (rc) return rc;
  if (info) {
    std::memset(info, 0, sizeof *info);
    info->n_v = R.nv;
    info->fft_len = R.plan.L;
    info->n1 = R.plan.n1;
    info->n2 = R.plan.n2;
    info->compiled_fft = R.plan.fc ? 1 : 0;
    info->line_begin = R.i0;
    info->line_end = R.i1;
    info->m_cap = R.m_cap;
    info->tile_w = R.tile_w;
    info->scale_log2 = R.scale_log2;
  }
  return 0;
}

int radis_ldm_slit_convolve(const float* in, int64_t n, const float* slit, int32_t k, double scale, float* out,
                            void* stream) {
  if (!in || !slit || !out) return fail(RADIS_LDM_EINVAL, "null pointer");
  if (n <= 0 || k <= 0) return fail(RADIS_LDM_EINVAL, "need n > 0 and k > 0 (n=%lld, k=%d)", (long long)n, k);
  if (!std::isfinite(scale)) return fail(RADIS_LDM_EINVAL, "non-finite scale");
  const char *a0 = (const char*)in, *a1 = a0 + n * 4, *b0 = (const char*)out, *b1 = b0 + n * 4;
  const char *s0 = (const char*)slit, *s1 = s0 + (int64_t)k * 4;
  if ((a0 < b1 && b0 < a1) || (s0 < b1 && b0 < s1)) return fail(RADIS_LDM_EINVAL, "output overlaps the input or the slit");
  const int64_t grid = (n + SLIT_OUT - 1) / SLIT_OUT;
  if (grid > INT32_MAX) return fail(RADIS_LDM_ERANGE, "grid too long");
  hipLaunchKernelGGL(k_slit_conv, dim3((unsigned)grid), dim3(SLIT_T), 0, (hipStream_t)stream, in, (long long)n, slit, k,
                     scale, out);
  HIPCHK(hipGetLastError());
  return 0;
}

void radis_ldm_batch_destroy(radis_ldm_batch* b) {
  if (!b) return;
  for (int k = 0; k < 2; ++k)
    if (b->pending[k] && b->ev_done[k]) (void)hipEventSynchronize(b->ev_done[k]);
  free_layout(b);
  for (int k = 0; k < 2; ++k) {
    host_release(b->ctx, b->jobs_h[k], sizeof(LdmJob) * b->n);
    host_release(b->ctx, b->res_h[k], sizeof(LdmResult) * b->n);
    if (b->ev_done[k]) (void)hipEventDestroy(b->ev_done[k]);
    for (auto& e : b->st_ev[k])
      if (e) (void)hipEventDestroy(e);
  }
  dev_release(b->ctx, b->jobs_d, sizeof(LdmJob) * b->n);
  dev_release(b->ctx, b->axes_d, sizeof(LdmAxes) * b->n);
  dev_release(b->ctx, b->res_d, sizeof(LdmResult) * b->n);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  radis_ldm_ctx* ctx = b->ctx;
  radis_ldm_db* db = b->db;
  delete b;
  if (db) db_release(db);
  if (ctx) ctx_release(ctx);
}

}  // extern "C"
