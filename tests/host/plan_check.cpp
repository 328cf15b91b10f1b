// Host planning of the LDM library under AddressSanitizer and
// UndefinedBehaviorSanitizer: line-list preparation, request planning,
// workspace layout, work items and job tables of batches, and the re-plan a
// condition change can trigger (LDM row growth), as the batch entry points
// run them, minus the device.  Every host buffer is sized exactly as the
// library sizes it, so a write past one aborts the run.  Prints "ok" and
// exits 0 when every check holds.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "ldm_plan.h"

namespace {

int g_fail = 0;
#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      std::fprintf(stderr, "check failed: %s: ", #cond);  \
      std::fprintf(stderr, __VA_ARGS__);                  \
      std::fprintf(stderr, "\n");                         \
      ++g_fail;                                           \
    }                                                     \
  } while (0)

struct Lines {
  std::vector<double> nu0;
  std::vector<float> S, El, ga, gs, na, ns, dl;
  std::vector<uint8_t> iso;
  std::vector<double> mm;
  radis_ldm_lines L{};
};

void make_lines(Lines& X, std::mt19937_64& rng, int64_t n, bool iso, bool nself, bool delta) {
  std::uniform_real_distribution<double> U(0.0, 1.0);
  X.nu0.resize(n);
  for (auto& v : X.nu0) v = 2000.0 + 500.0 * U(rng);
  std::sort(X.nu0.begin(), X.nu0.end());
  auto fill = [&](std::vector<float>& v, double a, double b) {
    v.resize(n);
    for (auto& x : v) x = (float)(a + (b - a) * U(rng));
  };
  X.S.resize(n);
  for (auto& x : X.S) x = (float)std::pow(10.0, -30.0 + 11.0 * U(rng));
  fill(X.El, 0.0, 8000.0);
  fill(X.ga, 0.03, 0.08);
  fill(X.gs, 0.05, 0.10);
  fill(X.na, 0.5, 0.8);
  if (nself) fill(X.ns, 0.4, 0.9); else X.ns.clear();
  if (delta) fill(X.dl, -0.01, 0.01); else X.dl.clear();
  X.mm = {0.0, 44.0, 45.0, 46.0};
  if (iso) {
    X.iso.resize(n);
    for (auto& s : X.iso) s = (uint8_t)(1 + rng() % 3);
  } else {
    X.iso.clear();
  }
  radis_ldm_lines& L = X.L;
  L = radis_ldm_lines{};
  L.nu0 = X.nu0.data();
  L.S_ref = X.S.data();
  L.El = X.El.data();
  L.g_air = X.ga.data();
  L.g_self = X.gs.data();
  L.n_air = X.na.data();
  L.n_self = nself ? X.ns.data() : nullptr;
  L.delta = delta ? X.dl.data() : nullptr;
  L.iso = iso ? X.iso.data() : nullptr;
  L.n = n;
  L.molar_mass = X.mm.data();
  L.n_iso_slots = 4;
}

radis_ldm_request make_request(std::mt19937_64& rng, bool whole) {
  std::uniform_real_distribution<double> U(0.0, 1.0);
  radis_ldm_request r{};
  const double W = whole ? 520.0 : 5.0 + 200.0 * U(rng);
  r.wmin = whole ? 1990.0 : 1990.0 + (520.0 - W) * U(rng);
  r.wmax = r.wmin + W;
  const double ws[] = {0.001, 0.002, 0.005, 0.01};
  r.wstep = ws[rng() % 4];
  r.Tgas = 300.0 + 2700.0 * U(rng);
  r.pressure_bar = std::exp(std::log(0.1) + (std::log(20.0) - std::log(0.1)) * U(rng));
  r.mole_fraction = 0.01 + 0.99 * U(rng);
  r.Tref = 296.0;
  r.dxG = std::log(std::sqrt(0.01 / 0.46) + 1);
  r.dxL = std::log(std::sqrt(0.01 / 0.20) + 1);
  if (rng() % 4 == 0) {  // fine axes: tall LDMs, filled in row groups
    r.dxG = 0.01;
    r.dxL = 0.02;
  }
  r.folding_thresh = 1e-6;
  for (int s = 0; s < RADIS_LDM_MAX_ISO; ++s) r.q_ratio[s] = 0.5 + U(rng);
  r.flags = (rng() % 4 == 0 ? RADIS_LDM_GENERIC_FFT : 0) | (rng() % 5 == 0 ? RADIS_LDM_PERIOD_L : 0) |
            (rng() % 7 == 0 ? RADIS_LDM_SCAN_AXIS : 0);
  return r;
}

// The host side of a batch: the requests' plans, one workspace laid out
// over all of them, the work-item list and the job table, each sized as
// build_layout / batch_create size them.
struct HostBatch {
  std::vector<Req> reqs;
  size_t total = 0;
  std::vector<int4> items;
  std::vector<LdmJob> jobs;
};

char g_work_base[1];
int g_splits = 0;  // requests laid out with a split pass B
int g_grouped = 0;  // requests laid out with fill row groups  // job tables hold offsets from a workspace base; never dereferenced here

void layout(const radis_ldm_db& db, HostBatch& B, int n_cu) {
  B.total = 0;
  choose_split_b(B.reqs.data(), (int)B.reqs.size(), n_cu);
  for (Req& R : B.reqs) {
    g_splits += R.split_b == 2;
    g_grouped += R.fill_groups > 1;
    const int rc = layout_request(R, B.total);
    CHECK(rc == 0, "layout_request: %s", g_err);
  }
  B.items.clear();
  for (size_t j = 0; j < B.reqs.size(); ++j) {
    Req& R = B.reqs[j];
    const LdmItemPlan P = item_plan(db, R);
    const int n = plan_items(P);
    const int nax = P.n_lo + P.n_g + P.n_hi;
    // axis partials fit the partial buffer the layout reserved
    const int64_t cap = (R.i1 - R.i0 + LDM_AX_CHUNK - 1) / LDM_AX_CHUNK + 2;
    CHECK(nax <= cap, "axis partials %d above %lld", nax, (long long)cap);
    CHECK(nax == R.n_axp, "n_axp %d vs %d", R.n_axp, nax);
    for (int i = 0; i < n; ++i) {
      int4 t = plan_item(P, i);
      if (i < nax) {
        CHECK(t.y == i, "axis partial index");
        if (t.w > 0) CHECK(t.z >= 0 && t.z + t.w <= R.i1 - R.i0 && t.w <= LDM_AX_CHUNK, "scan item out of the window");
        else CHECK(t.z >= 0 && (int64_t)(t.z - t.w) * LDM_SKY_BLOCK <= db.n, "candidate group out of the list");
      } else if (i < nax + P.n_wrap) {
        CHECK(2 * t.y < R.wrap_rows, "wrap pair beyond the table");
      } else if (i < nax + P.n_wrap + P.n_fill) {
        CHECK(t.y >= 0 && t.y < R.n_tiles * R.fill_groups, "tile / row group");
      } else if (i >= nax + P.n_wrap + P.n_fill + P.pcap * P.n_colA && i < n - P.n_colC) {
        // pass B: (row block, share), each once, in order
        const int k = i - (nax + P.n_wrap + P.n_fill + P.pcap * P.n_colA);
        CHECK(t.y == k / R.split_b && t.z == k % R.split_b && t.y < (R.plan.n1 >> R.bB), "pass B item");
      }
      t.x = (int)j;
      B.items.push_back(t);
    }
    // every buffer of the request lies inside the workspace (the library's
    // own check, then each buffer here)
    CHECK(check_layout(R, B.total) == 0, "check_layout: %s", g_err);
    const size_t V = (size_t)R.nv + 1;
    // LDM: interleaved row pairs, ceil(m_cap / 2) pairs of 2 * stride floats
    CHECK(R.off_ldm + (size_t)((R.m_cap + 1) / 2) * 2 * R.stride * 4 <= B.total, "LDM");
    CHECK((size_t)R.stride >= V, "LDM stride below N_v + 1");
    CHECK(R.off_y + (size_t)((R.m_cap + 1) / 2) * R.plan.L * 8 <= B.total, "Y");
    CHECK(R.off_b + (size_t)R.plan.L * 8 <= B.total, "B");
    CHECK(R.split_b == 1 || (R.split_b == 2 && R.plan.fc && R.off_b2 >= R.off_b + (size_t)R.plan.L * 8 &&
                             R.off_b2 + (size_t)R.plan.L * 8 <= B.total),
          "split pass B: second share's buffer");
    CHECK(R.off_rows + (size_t)R.m_cap * 16 <= B.total, "rows");
    CHECK(R.off_axp + (size_t)cap * 32 <= B.total, "axis partials");
    CHECK(R.off_wrap + (size_t)R.wrap_rows * R.plan.wrap_k * 4 <= B.total, "wrap table");
    CHECK(R.plan.L >= 2 * R.nv, "L below 2 N_v");
    // the compiled pass B corrects bins k2 = 0 and N2 - 1 only: K <= N1 there
    // (the run-time-sized pass B checks every bin)
    CHECK(R.plan.wrap_k == 0 || (R.plan.wrap_k <= LDM_WRAP_KMAX && (!R.plan.fc || R.plan.wrap_k <= R.plan.n1) &&
                                 R.plan.L > 2 * R.nv),
          "wrap bins %d, n1 %d, compiled %d", R.plan.wrap_k, R.plan.n1, (int)R.plan.fc);
    CHECK((R.plan.wrap_k == 0) == (R.plan.L == 2 * R.nv || (R.r.flags & RADIS_LDM_PERIOD_L)), "wrap iff L != 2 N_v");
    CHECK(fill_tile_bytes(R.fill_rows, R.tile_w) <= (size_t)FILL_LDS_MAX, "fill tile LDS");
    CHECK(R.fill_groups >= 1 && (long long)R.fill_rows * R.fill_groups >= R.m_cap &&
              (R.fill_groups == 1 ? R.fill_rows == R.m_cap : R.fill_rows % 2 == 0),
          "fill row groups: %d x %d rows for %d", R.fill_groups, R.fill_rows, R.m_cap);
  }
  // the job table: exactly one LdmJob per request
  B.jobs.assign(B.reqs.size(), LdmJob{});
  for (size_t j = 0; j < B.reqs.size(); ++j) fill_job(db, B.reqs[j], g_work_base, nullptr, B.jobs[j]);
  for (size_t j = 0; j < B.reqs.size(); ++j) {
    const LdmJob& J = B.jobs[j];
    const Req& R = B.reqs[j];
    CHECK(J.n_ax_part == R.n_axp && J.m_cap == R.m_cap && J.L == R.plan.L && J.nv == R.nv, "job fields");
    CHECK(J.wrap_k == R.plan.wrap_k && J.wrap_rows == R.wrap_rows && J.wrap_ls == 2 * R.nv, "job wrap fields");
    CHECK((J.bbuf2 != nullptr) == (R.split_b == 2), "job second pass B share");
    CHECK(J.nrad1 == R.plan.nr1 && J.nrad2 == R.plan.nr2 && J.nrad1 <= LDM_MAX_RAD && J.nrad2 <= LDM_MAX_RAD,
          "radix plan");
  }
}

}  // namespace

int main() {
  std::mt19937_64 rng(20261203);
  int planned = 0, refused = 0, replans = 0;
  for (int trial = 0; trial < 24; ++trial) {
    Lines X;
    const int64_t sizes[] = {1, 37, 4096, 5000, 20000, 70000};
    make_lines(X, rng, sizes[trial % 6], trial % 3 == 1, trial % 4 == 2, trial % 5 == 3);
    radis_ldm_db db;
    DbHostArrays A;
    int rc = db_prepare(&X.L, &db, A);
    CHECK(rc == 0, "db_prepare: %s", g_err);
    if (rc) continue;
    CHECK(db.has_sky == (A.nfull > 0), "candidate blocks");
    CHECK(A.sky_off.size() == (db.has_sky ? (size_t)4 * (A.nfull + 1) : 0), "candidate offsets");
    // a batch of requests, including a whole-list window
    HostBatch B;
    std::vector<radis_ldm_request> rq;
    for (int k = 0; k < 7; ++k) {
      radis_ldm_request r = make_request(rng, k == 0);
      Req R;
      rc = plan_one(db, r, R);
      if (rc) {
        ++refused;  // a planning error (e.g. an axis too long) is reported, never a fault
        CHECK(g_err[0] != 0, "error without a message");
        continue;
      }
      R.m_cap = R.m_need;
      B.reqs.push_back(R);
      rq.push_back(r);
      ++planned;
    }
    if (B.reqs.empty()) continue;
    // compute units: pass B split on every eligible request, on some, or on none
    const int n_cu = trial % 3 == 0 ? (1 << 20) : (trial % 3 == 1 ? 256 : 1);
    layout(db, B, n_cu);
    // condition changes (set_request): new T, p, x on the same grid; the
    // rows grow by 1.5x when the bound rises above the allocation, and the
    // whole batch is laid out again (build_layout)
    for (int round = 0; round < 3; ++round) {
      bool grow = false;
      for (size_t j = 0; j < B.reqs.size(); ++j) {
        radis_ldm_request r = rq[j];
        r.pressure_bar = rq[j].pressure_bar * (2.0 + round);
        r.mole_fraction = 1.0 - 0.5 * rq[j].mole_fraction;
        r.Tgas = 250.0 + 300.0 * round;
        Req T = B.reqs[j];
        T.r = r;
        T.r.fft_len = B.reqs[j].plan.L;
        rc = bound_request(db, T);
        if (rc) continue;
        const bool g = T.m_need > B.reqs[j].m_cap;
        const int cap = B.reqs[j].m_cap;
        B.reqs[j] = T;
        if (g) {
          B.reqs[j].m_cap = std::max(T.m_need, (int)std::ceil(1.5 * cap));
          grow = true;
        }
      }
      if (grow) {
        ++replans;
        layout(db, B, n_cu);
      } else {
        B.jobs.assign(B.reqs.size(), LdmJob{});
        for (size_t j = 0; j < B.reqs.size(); ++j) fill_job(db, B.reqs[j], g_work_base, nullptr, B.jobs[j]);
      }
    }
  }
  // planner rules on their own: the transform length is 2 N_v or leaves the
  // image correction its gap, for every grid length up to 2^25 (beyond
  // 1024 x 8192 points the split takes a longer n1, up to the LDS capacity)
  int big = 0;
  for (int64_t nv = 2; nv < (1 << 25); nv = nv * 9 / 8 + 1) {
    FftPlan p;
    const int rc = plan_fft(nv, 0, p);
    if (rc) {
      CHECK(nv < 1000 || wrap_floor(2 * nv) > (int64_t)N1_BIG * N2_MAX, "nv %lld refused: %s", (long long)nv, g_err);
      continue;  // grids too short for the correction, or beyond the longest split, report an error
    }
    CHECK(p.L == 2 * nv || (p.wrap_k > 0 && p.wrap_k <= LDM_WRAP_KMAX && p.L >= wrap_floor(2 * nv)),
          "nv %lld: L %lld wrap %d", (long long)nv, (long long)p.L, p.wrap_k);
    CHECK(p.n1 <= (p.L > (int64_t)N1_MAX * N2_MAX ? N1_BIG : N1_MAX) && p.n2 <= N2_MAX && p.L == (int64_t)p.n1 * p.n2,
          "nv %lld: split %d x %d", (long long)nv, p.n1, p.n2);
    big += p.n1 > N1_MAX;
  }
  CHECK(big > 0, "no grid took a split with n1 above %d", N1_MAX);
  CHECK(g_splits > 0, "no request took the split pass B");
  CHECK(g_grouped > 0, "no request took fill row groups");
  std::printf("planned %d requests, %d refused, %d re-plans, %d split pass B layouts, %d row-group layouts, "
              "%d failed checks\n", planned, refused, replans, g_splits, g_grouped, g_fail);
  if (g_fail) return 1;
  std::printf("ok\n");
  return 0;
}
