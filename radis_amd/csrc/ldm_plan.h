// Host planning of the LDM pipeline: request validation, grid and FFT
// plans, bounds on the LDM rows and the intensity scale from the line
// list's statistics, per-request workspace layout, work-item plans and the
// job table.  Host code only (no kernels and no HIP runtime calls), so it
// also builds with a host compiler alone, sanitisers included
// (tests/host/plan_check.cpp).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "radis_ldm.h"
#include "ldm_job.h"

struct radis_ldm_ctx;

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
const int FILL_GROUP_TILE_W = 32;      // tile width of fills in row groups
const int N1_MAX = 1024, N2_MAX = 8192;
// grids beyond N1_MAX * N2_MAX: run-time-sized passes A and C with n1 up to
// the LDS capacity (one to a few columns per workgroup: slow, but any grid
// up to N1_BIG * N2_MAX = 2^26 points transforms)
const int N1_BIG = LDM_CAP;

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

// Line list: device columns (LdmDb) and the host-side copies and
// statistics the planner reads.
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
const int FC_N2[] = {64, 128, 256, 512, 1024, 2048};
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
  int fill_rows = 0, fill_groups = 1;  // row groups of a fill tile (an LDM taller than one tile's LDS)
  int64_t stride = 0;
  int scale_log2 = 0;
  int bA = 0, bB = 0, bC = 0;  // log2 of the pass A / C column blocks and pass B row blocks
  int split_b = 1;     // workgroups per pass B row block, each taking a share of the pairs
  int n_axp = 0;       // axis partials (items of k_axis)
  bool sky_whole = false;  // width extremes from the whole list's Pareto sets in k_finalize (no k_axis items)
  size_t off_ldm = 0, off_y = 0, off_b = 0, off_b2 = 0, off_rows = 0, off_axp = 0, off_wrap = 0;
  const float2 *tw1 = nullptr, *tw2 = nullptr, *twl_hi = nullptr, *twl_lo = nullptr, *tw_wrap = nullptr;
};

// Host arrays of a line list that go to the device with its columns: the
// per-block Lorentzian-width candidates (k_axis) and their offsets, the
// isotopologue ranges per block.
struct DbHostArrays {
  std::vector<float4> sky_vals;
  std::vector<int> sky_off;
  std::vector<int2> iso_rng;
  int nfull = 0;  // whole candidate blocks
};

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


// Validate a host line list and compute everything the planner and the
// device need besides its columns: flags, molar masses, per-block statistics
// (widths, exponents, energies, intensities), the width-candidate Pareto sets
// per block and of the whole list.  Fills the host fields of db (db->dev
// only gsky_off) and A; allocates no device memory.
int db_prepare(const radis_ldm_lines* L, radis_ldm_db* db, DbHostArrays& A) {
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
  db->n = n;
  db->has_iso = L->iso != nullptr;
  db->has_nself = L->n_self != nullptr;
  db->has_delta = L->delta != nullptr;
  if ((L->Evib_l == nullptr) != (L->Evib_u == nullptr)) {
    return fail(RADIS_LDM_EINVAL, "Evib_l and Evib_u come together");
  }
  db->has_evib = L->Evib_l != nullptr;
  db->n_iso_slots = L->n_iso_slots;
  for (int s = 0; s < L->n_iso_slots; ++s) db->molar_mass[s] = L->molar_mass[s];
  if (db->has_iso) {
    for (int64_t i = 0; i < n; ++i) {
      const int s = L->iso[i];
      if (s <= 0 || s >= L->n_iso_slots || !(L->molar_mass[s] > 0)) {
        return fail(RADIS_LDM_EINVAL, "line %lld: iso slot %d has no molar mass", (long long)i, s);
      }
      db->iso_mask |= 1u << s;
    }
  } else {
    if (!(L->molar_mass[1] > 0)) { return fail(RADIS_LDM_EINVAL, "molar_mass[1] must be positive"); }
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
  std::vector<float4>& sky_vals = A.sky_vals;
  std::vector<int>& sky_off = A.sky_off;
  std::vector<int2>& iso_rng = A.iso_rng;
  const int nfull = (int)(n / LDM_SKY_BLOCK);
  A.nfull = nfull;
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
  return 0;
}

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

// Search L = n1*n2 >= 2*N_v (or == want) with 16 | n1 <= 1024 (n1_big:
// <= N1_BIG) and 16 | n2 <= n2_max, 7-smooth cofactors; smallest L, then n1
// closest to 512.
int64_t search_fft(int64_t need, int64_t want, int n2_max, int& b1, int& b2, int only_n1 = 0, bool n1_big = false) {
  static const SmoothList c1(N1_MAX);
  static const SmoothList c1b(N1_BIG);
  static const SmoothList c2(N2_MAX);
  int64_t bestL = -1;
  double bestScore = 1e300;
  for (int n1 : n1_big ? c1b : c1) {
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
// the image-correction items (wrap_item) correct the images (the result is
// RADIS's at L = 2 N_v).
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
  int64_t any = pick(N2_MAX, a1, a2);
  const int64_t small = pick(LDM_CAP_B, s1, s2);  // pass B with many small tiles
  if (any < 0 && !only_n1 && want == 0) {
    // beyond N1_MAX * N2_MAX: a run-time-sized split with a longer n1
    const int64_t exact = search_fft(need, need, N2_MAX, a1, a2, 0, true);
    any = exact > 0 ? exact : search_fft(lo, 0, N2_MAX, a1, a2, 0, true);
  }
  if (any < 0) {
    if (want > 0)
      return fail(RADIS_LDM_ERANGE, "fft_len %lld is not n1*n2 with 16|n1<=%d, 16|n2<=%d and 7-smooth cofactors",
                  (long long)want, N1_MAX, N2_MAX);
    return fail(RADIS_LDM_ERANGE, "grid of %lld points needs an FFT above %lld points", (long long)nv,
                (long long)N1_BIG * N2_MAX);
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

size_t fill_tile_bytes(int m_cap, int w) { return (size_t)m_cap * (size_t)(w + ldm_fill_pad(w)) * sizeof(double); }

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
  J.ratio_T = r.Tref / r.Tgas;
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
  // a line's index can round up across a node only from within a few ulp
  // below it: 1e-3 of a step covers that; each step of margin would add
  // 1/tile_w of the tile's line reads
  J.shift_margin = (db.has_delta ? db.max_abs_delta * J.p_atm : 0.0) + 1e-3 * R.delta_w;
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
  J.fill_rows = R.fill_rows;
  J.fill_groups = R.fill_groups;
  J.ldm_stride = R.stride;
  J.ldm = (float*)(work + R.off_ldm);
  J.ybuf = (float2*)(work + R.off_y);
  J.bbuf = (float2*)(work + R.off_b);
  J.bbuf2 = R.split_b > 1 ? (float2*)(work + R.off_b2) : nullptr;
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
  J.tw_wrap = R.tw_wrap;
  J.wl_min_bits = ~0ull;
  J.wl_max_bits = 0ull;
  J.wg_min_bits = ~0ull;
  J.wg_max_bits = 0ull;
  J.status = 0;
}

// Workgroups per pass B row block (Req::split_b) for requests reqs[0..n)
// that run together.  A compiled pass B workgroup walks every pair of its
// row block; when a launch has few row blocks (a 500,001-point grid at
// L = 512 x 2048: 512 workgroups, two per CU) the device runs them at two
// waves per SIMD, latency-bound.  Splitting the pairs over two workgroups
// (pass C adds the two partial spectra) doubles the waves when the doubled
// launch still fits FC_B_SPLIT_WGS_PER_CU workgroups per CU at once; it
// costs an inverse transform per share and L*8 bytes written and read.
// Measured on that grid: pass B 44.2 -> 37.7 us; on L = 640 x 2048 with 72
// pairs (1280 workgroups, more than fit at once) the split was slower.
constexpr int FC_B_SPLIT_WGS_PER_CU = 4;
constexpr int FC_B_SPLIT_MIN_PAIRS = 4;
void choose_split_b(Req* reqs, int n, int n_cu) {
  std::map<int, long long> blocks;  // compiled pass B class (N2) -> row blocks of the launch
  for (int i = 0; i < n; ++i)
    if (reqs[i].plan.fc) blocks[reqs[i].plan.n2] += reqs[i].plan.n1 / fc_rows_b(reqs[i].plan.n2);
  for (int i = 0; i < n; ++i) {
    Req& R = reqs[i];
    R.split_b = 1;
    if (R.plan.fc && (R.m_cap + 1) / 2 >= FC_B_SPLIT_MIN_PAIRS &&
        2 * blocks[R.plan.n2] <= (long long)FC_B_SPLIT_WGS_PER_CU * n_cu)
      R.split_b = 2;
  }
}

// Workspace of one request (tile width, pass splits and the offsets of its
// buffers from `total`, which it advances); shared by the batch layout and
// radis_ldm_request_bytes.
int layout_request(Req& R, size_t& total) {
  R.tile_w = R.r.tile_w ? R.r.tile_w : choose_tile(R.m_cap);
  R.fill_rows = R.m_cap;
  R.fill_groups = 1;
  if (fill_tile_bytes(R.m_cap, R.tile_w) > (size_t)FILL_LDS_MAX) {
    // An LDM taller than one tile's LDS: every tile is filled in row groups
    // of an even number of rows (whole interleaved pairs), each group its own
    // workgroup reading the tile's lines again and keeping the corners that
    // fall in its rows.  32-point tiles (128-B flush segments per pair).
    if (!R.r.tile_w) R.tile_w = FILL_GROUP_TILE_W;
    const int rows = (int)((size_t)FILL_LDS_TARGET / fill_tile_bytes(1, R.tile_w)) & ~1;
    if (rows < 2)
      return fail(RADIS_LDM_ERANGE, "fill tile of width %d does not fit two LDM rows", R.tile_w);
    R.fill_rows = rows;
    R.fill_groups = (R.m_cap + rows - 1) / rows;
  }
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
  // LDM: pcap pairs of rows, rows 2p and 2p+1 interleaved (element n of row
  // 2p+h at float 2n+h of pair p): pass A reads each element pair as one float2
  R.off_ldm = total; total = align_up(total + (size_t)pcap * 2 * R.stride * 4, 256);
  R.off_y = total;   total = align_up(total + (size_t)pcap * R.plan.L * 8, 256);
  R.off_b = total;   total = align_up(total + (size_t)R.plan.L * 8, 256);
  // pass B's second share of the pairs (split pass B)
  R.off_b2 = total;  if (R.split_b > 1) total = align_up(total + (size_t)R.plan.L * 8, 256);
  R.off_rows = total; total = align_up(total + (size_t)R.m_cap * 16, 256);
  // axis partials: at most one per chunk, plus the two partial blocks at the window edges
  R.off_axp = total; total = align_up(total + (size_t)((R.i1 - R.i0 + LDM_AX_CHUNK - 1) / LDM_AX_CHUNK + 2) * 32, 256);
  // circular-image correction: wrap_k floats per L-axis node (N_L <= M/2, as N_G >= 2)
  R.wrap_rows = R.plan.wrap_k ? std::max(1, std::min(R.m_cap / 2, LDM_MAX_AX)) : 0;
  R.off_wrap = total; total = align_up(total + (size_t)R.wrap_rows * R.plan.wrap_k * 4, 256);
  return 0;
}

// Every buffer of request R inside a workspace of `total` bytes (the layout
// guarantees it; checked after each layout, before any job table points
// into the workspace).
int check_layout(const Req& R, size_t total) {
  const size_t V = (size_t)R.nv + 1;
  const size_t parts = (size_t)((R.i1 - R.i0 + LDM_AX_CHUNK - 1) / LDM_AX_CHUNK + 2);
  const bool ok = (size_t)R.stride >= V && R.off_ldm + (size_t)((R.m_cap + 1) / 2) * 2 * R.stride * 4 <= total &&
                  R.off_y + (size_t)((R.m_cap + 1) / 2) * R.plan.L * 8 <= total &&
                  R.off_b + (size_t)R.plan.L * 8 <= total &&
                  (R.split_b == 1 || R.off_b2 + (size_t)R.plan.L * 8 <= total) && R.off_rows + (size_t)R.m_cap * 16 <= total &&
                  R.off_axp + parts * 32 <= total && (size_t)R.n_axp <= parts &&
                  R.off_wrap + (size_t)R.wrap_rows * R.plan.wrap_k * 4 <= total;
  return ok ? 0 : fail(RADIS_LDM_EINVAL, "internal: request workspace layout exceeds its %zu bytes", total);
}

}  // namespace

// The work items of one request (job 0), in the order the batch layout
// builds them: axis items (edge chunks scanned, whole blocks in groups of G
// through their candidates), image-correction node pairs, fill tiles,
// pass A (pair, column block), pass B row blocks, pass C column blocks.
struct LdmItemPlan {
  long long i0;                 // window start (axis chunk offsets are relative to it)
  long long lo_a, lo_e;         // lines [lo_a, lo_e) scanned in LDM_AX_CHUNK chunks
  long long g0, fb1, G;         // whole blocks [g0, fb1) in groups of G
  long long hi_a, hi_e;         // lines [hi_a, hi_e) scanned in chunks
  int n_lo, n_g, n_hi;          // items of the three axis parts
  int n_wrap, n_fill, pcap, n_colA, n_rowB, n_colC;
  int split_b;                  // pass B items per row block (n_rowB counts them all)
};

// Item i of the request (job field 0), i < plan_items(P).
__host__ __device__ inline int plan_items(const LdmItemPlan& P) {
  return P.n_lo + P.n_g + P.n_hi + P.n_wrap + P.n_fill + P.pcap * P.n_colA + P.n_rowB + P.n_colC;
}
__host__ __device__ inline int4 plan_item(const LdmItemPlan& P, int i) {
  int k = i;
  if (k < P.n_lo + P.n_g + P.n_hi) {
    // axis items: y = the item's partial index
    if (k < P.n_lo) {
      const long long c = P.lo_a + (long long)k * LDM_AX_CHUNK;
      const long long n = P.lo_e - c < LDM_AX_CHUNK ? P.lo_e - c : LDM_AX_CHUNK;
      return make_int4(0, k, (int)(c - P.i0), (int)n);
    }
    if (k < P.n_lo + P.n_g) {
      const long long g = P.g0 + (long long)(k - P.n_lo) * P.G;
      const long long n = P.fb1 - g < P.G ? P.fb1 - g : P.G;
      return make_int4(0, k, (int)g, -(int)n);
    }
    const long long c = P.hi_a + (long long)(k - P.n_lo - P.n_g) * LDM_AX_CHUNK;
    const long long n = P.hi_e - c < LDM_AX_CHUNK ? P.hi_e - c : LDM_AX_CHUNK;
    return make_int4(0, k, (int)(c - P.i0), (int)n);
  }
  k -= P.n_lo + P.n_g + P.n_hi;
  if (k < P.n_wrap) return make_int4(0, k, 0, 0);                 // L-axis node pair
  if ((k -= P.n_wrap) < P.n_fill) return make_int4(0, k, 0, 0);  // fill tile
  if ((k -= P.n_fill) < P.pcap * P.n_colA) return make_int4(0, k / P.n_colA, k % P.n_colA, 0);  // (pair, column block)
  if ((k -= P.pcap * P.n_colA) < P.n_rowB) return make_int4(0, k / P.split_b, k % P.split_b, 0);  // pass B (row block, share)
  return make_int4(0, 0, k - P.n_rowB, 0);                        // pass C column block
}

namespace {

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
  P.n_fill = R.n_tiles * R.fill_groups;  // item y = tile * fill_groups + group
  P.pcap = (R.m_cap + 1) / 2;
  P.n_colA = R.plan.n2 >> R.bA;
  P.split_b = R.split_b;
  P.n_rowB = (R.plan.n1 >> R.bB) * R.split_b;
  P.n_colC = R.plan.n2 >> R.bC;
  return P;
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

}  // namespace
