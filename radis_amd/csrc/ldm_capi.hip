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
#include "ldm_plan.h"
#include "ldm_kernels.h"
#include "ldm_fftc.h"
#include "ldm_slit.h"
#include "ldm_wrap.h"
#include "ldm_setup.h"

namespace {

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(RADIS_LDM_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

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
  // non-blocking stream for set-up copies and result read-back: they wait
  // for nothing but themselves (the null stream would wait for every
  // blocking stream's work, i.e. for other batches in flight)
  hipStream_t side = nullptr;
  ~radis_ldm_ctx();
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

}  // namespace

// One launch of an FFT pass: items of one kernel variant are contiguous;
// cls = the compiled size (n1 for A and C, n2 for B), 0 = generic (pass B:
// 0 = generic with small tiles, -1 = generic with large tiles).
struct LaunchGroup {
  int cls, first, count;
  int rows;  // pass B: the rows' coefficients staged in LDS (the largest m_cap of the group, at most FC_B_ROWS_LDS)
  int wh;    // pass B: L-axis nodes of the image-correction values staged in LDS (largest wrap_rows, 0: none)
};

struct radis_ldm_batch {
  radis_ldm_ctx* ctx = nullptr;
  radis_ldm_db* db = nullptr;
  int n = 0;
  std::vector<Req> reqs;
  // two pinned job tables / result slots so that run k only waits for run k-2
  // (each n entries, n fixed at creation)
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
  bool fill_grouped = false;  // some request fills its tiles in row groups
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

void free_layout(radis_ldm_batch* b) {
  if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  b->graph_valid = false;
  if (b->work) { dev_release(b->ctx, b->work, b->work_bytes); b->ctx->in_use -= (int64_t)b->work_bytes; b->work = nullptr; b->work_bytes = 0; }
  if (b->twmem) { dev_release(b->ctx, b->twmem, b->tw_bytes); b->ctx->in_use -= (int64_t)b->tw_bytes; b->twmem = nullptr; b->tw_bytes = 0; }
  if (b->items_d) { dev_release(b->ctx, b->items_d, b->items_bytes); b->items_d = nullptr; }
}

int build_layout(radis_ldm_batch* b) {
  free_layout(b);
  b->jobs_last.clear();
  b->tiles_valid = false;
  // --- workspace
  size_t total = 0;
  size_t fill_lds = 0;
  bool fill_grouped = false;
  choose_split_b(b->reqs.data(), (int)b->reqs.size(), b->ctx->n_cu);
  for (Req& R : b->reqs) {
    int rc = layout_request(R, total);
    if (rc) return rc;
    fill_lds = std::max(fill_lds, fill_tile_bytes(R.fill_rows, R.tile_w));
    fill_grouped = fill_grouped || R.fill_groups > 1;
  }
  // --- twiddle tables, deduplicated by size
  std::map<int, size_t> sub_off;
  std::map<int64_t, size_t> l_off;
  size_t twsz = 0;
  sub_off[LDM_WRAP_LC] = twsz;  // the image correction's transform
  twsz = align_up(twsz + (size_t)LDM_WRAP_LC * 8, 256);
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
    hipLaunchKernelGGL(k_twiddles, dim3((kv.first + 255) / 256), dim3(256), 0, b->ctx->side,
                       (float2*)((char*)b->twmem + kv.second), kv.first, (float2*)nullptr, 0, (float2*)nullptr, 0ll);
  for (auto& kv : l_off) {
    const long long nt = ((kv.first + 1023) >> LDM_TW_LO_BITS) + 1024;
    hipLaunchKernelGGL(k_twiddles, dim3((unsigned)std::min<long long>((nt + 255) / 256, 4096)), dim3(256), 0, b->ctx->side,
                       (float2*)nullptr, 0, (float2*)nullptr, 0, (float2*)((char*)b->twmem + kv.second),
                       (long long)kv.first);
  }
  HIPCHK(hipGetLastError());
  for (Req& R : b->reqs) {
    char* base = (char*)b->twmem;
    R.tw1 = (const float2*)(base + sub_off[R.plan.n1]);
    R.tw2 = (const float2*)(base + sub_off[R.plan.n2]);
    R.twl_hi = (const float2*)(base + l_off[R.plan.L]);
    R.twl_lo = R.twl_hi + (size_t)((R.plan.L + 1023) >> LDM_TW_LO_BITS);
    R.tw_wrap = (const float2*)(base + sub_off[LDM_WRAP_LC]);
  }
  // --- work items (FFT passes grouped by kernel variant), per request in
  // the order plan_item gives them
  std::vector<int4> ax, fi, wr;
  std::map<int, std::vector<int4>> ma, mb, mc;
  std::map<int, int> mb_rows, mb_wh;  // pass B class -> largest m_cap, largest wrap_rows
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
    mb_rows[cb_] = std::max(mb_rows[cb_], R.m_cap);
    mb_wh[cb_] = std::max(mb_wh[cb_], R.wrap_rows);
    put(mc[ca_ | (R.split_b > 1 ? FC_C_SUM : 0)], bc, be);  // split pass B: pass C adds two shares
  }
  std::vector<int4> fa, fb, fc;
  auto flatten = [](std::map<int, std::vector<int4>>& m, std::vector<int4>& v,
                    std::vector<radis_ldm_batch::Group>& g) {
    g.clear();
    for (auto& kv : m) {
      g.push_back({kv.first, (int)v.size(), (int)kv.second.size(), 0, 0});
      v.insert(v.end(), kv.second.begin(), kv.second.end());
    }
  };
  flatten(ma, fa, b->grp_a);
  flatten(mb, fb, b->grp_b);
  for (auto& G : b->grp_b) {
    G.rows = std::min(mb_rows[G.cls], FC_B_ROWS_LDS);
    G.wh = mb_wh[G.cls];
  }
  flatten(mc, fc, b->grp_c);
  b->n_axis = (int)ax.size();
  b->n_wrap = (int)wr.size();
  b->n_fill = (int)fi.size();
  b->n_a = (int)fa.size();
  b->n_b = (int)fb.size();
  b->n_c = (int)fc.size();
  for (const Req& R : b->reqs) {
    const int rc = check_layout(R, total);
    if (rc) return rc;
  }
  std::vector<int4> all;
  for (auto* v : {&ax, &wr, &fi, &fa, &fb, &fc}) all.insert(all.end(), v->begin(), v->end());
  if (all.size() != (size_t)(b->n_axis + b->n_wrap + b->n_fill + b->n_a + b->n_b + b->n_c) || all.size() > INT32_MAX)
    return fail(RADIS_LDM_EINVAL, "internal: %zu work items for the launch counts", all.size());
  if (all.empty()) all.push_back(make_int4(0, 0, 0, 0));
  b->items_bytes = all.size() * sizeof(int4);
  b->items_d = (int4*)dev_alloc(b->ctx, b->items_bytes);
  if (!b->items_d) return fail(RADIS_LDM_ENOMEM, "device allocation of work items failed");
  HIPCHK(hipMemcpyAsync(b->items_d, all.data(), all.size() * sizeof(int4), hipMemcpyHostToDevice, b->ctx->side));
  // the twiddle tables and the items are in place before any run is enqueued
  // (and `all` outlives the copy)
  HIPCHK(hipStreamSynchronize(b->ctx->side));
  b->fill_lds = std::max<size_t>(fill_lds, 4);
  b->fill_grouped = fill_grouped;
  return 0;
}


// Host-to-device upload of the job table (outside any graph: a captured
// copy from pinned memory is not guaranteed to re-read the host buffer).
int upload_jobs(radis_ldm_batch* b, hipStream_t s, int slot) {
  if (slot < 0 || slot > 1 || !b->jobs_h[slot] || (int)b->reqs.size() != b->n)
    return fail(RADIS_LDM_EINVAL, "internal: job table slot %d of a batch of %d", slot, b->n);
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
  if (b->last_slot < 0 || b->last_slot > 1 || !b->res_h[b->last_slot])
    return fail(RADIS_LDM_EINVAL, "internal: result slot %d", b->last_slot);
  // the runs have completed (wait_pending); a copy on the side stream waits
  // for nothing else
  HIPCHK(hipMemcpyAsync(b->res_h[b->last_slot], b->res_d, sizeof(LdmResult) * b->n, hipMemcpyDeviceToHost,
                        b->ctx->side));
  HIPCHK(hipStreamSynchronize(b->ctx->side));
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
  bool fill_grouped;  // the fill kernel with row groups (some request's LDM is taller than a tile)
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
  // The image-correction items run as the first workgroups of the fill
  // launch, concurrently with the tiles (as a launch of their own they held
  // the device for ~11 us; a fork onto a stream of its own, joined before
  // pass B, cost more in cross-stream events than the kernel).
  const int4* wrap_fill = items;  // wrap items, then fill items
  items += X.n_wrap;
  if (X.n_fill && X.tile_bounds)
    launch(b, timed, slot, 2, k_tile_bounds, dim3((X.n_fill + 3) / 4), dim3(256), 0, s, X.jobs,
           X.items + (items - X.items), X.n_fill, dbd);
  if (X.n_wrap + X.n_fill) {
    const size_t lds = X.n_wrap ? std::max(X.fill_lds, (size_t)LDM_WRAP_LDS * sizeof(float2)) : X.fill_lds;
    const dim3 grid(X.n_wrap + X.n_fill);
    const bool plain = !dbd.iso && !dbd.delta && !dbd.n_self;
    if (X.fill_grouped) {
      if (plain)
        launch(b, timed, slot, 3, k_fill<true, true>, grid, dim3(LDM_FILL_T), lds, s, cj, ca, wrap_fill, X.n_wrap, dbd);
      else
        launch(b, timed, slot, 3, k_fill<false, true>, grid, dim3(LDM_FILL_T), lds, s, cj, ca, wrap_fill, X.n_wrap, dbd);
    } else if (plain) {
      launch(b, timed, slot, 3, k_fill<true, false>, grid, dim3(LDM_FILL_T), lds, s, cj, ca, wrap_fill, X.n_wrap, dbd);
    } else {
      launch(b, timed, slot, 3, k_fill<false, false>, grid, dim3(LDM_FILL_T), lds, s, cj, ca, wrap_fill, X.n_wrap, dbd);
    }
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
  case N:                                                                                                    \
    launch_part(b, timed, slot, 5, f, l, k_fft_b_c<N>, grid, dim3(FC_T_B),                                   \
                (size_t)G.rows * sizeof(float4) + (size_t)2 * fc_rows_b(N) * G.wh * sizeof(float), s, cj, ca, it,  \
                G.rows, G.wh);                                                                                 \
    break;
    switch (G.cls) {
      FC_B(64) FC_B(128) FC_B(256) FC_B(512) FC_B(1024) FC_B(2048)
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
#define FC_C(N)                                                                                                \
  case N: launch_part(b, timed, slot, 6, f, l, k_fft_c_c<N, false>, grid, dim3(FC_T), 0, s, cj, it, G.count); break; \
  case N | FC_C_SUM:                                                                                                  \
    launch_part(b, timed, slot, 6, f, l, k_fft_c_c<N, true>, grid, dim3(FC_T), 0, s, cj, it, G.count);               \
    break;
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
  X.fill_grouped = b->fill_grouped;
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

// The job table of `slot` from the requests and their output pointers
// (n of each; the host buffers were sized for n at creation).
int refresh_jobs(radis_ldm_batch* b, int slot, const std::vector<float*>& outs) {
  if ((int)outs.size() != b->n || (int)b->reqs.size() != b->n || slot < 0 || slot > 1 || !b->jobs_h[slot])
    return fail(RADIS_LDM_EINVAL, "internal: %zu outputs, %zu requests for a batch of %d", outs.size(),
                b->reqs.size(), b->n);
  for (int j = 0; j < b->n; ++j) fill_job(*b->db, b->reqs[j], (char*)b->work, outs[j], b->jobs_h[slot][j]);
  return 0;
}

}  // namespace

radis_ldm_ctx::~radis_ldm_ctx() {
  if (side) (void)hipStreamDestroy(side);
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
  hipStream_t side = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  auto* c = new radis_ldm_ctx();
  c->device = device;
  c->scratch = scr;
  c->side = side;
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


int radis_ldm_db_upload(radis_ldm_ctx* ctx, const radis_ldm_lines* L, radis_ldm_db** out) {
  if (!ctx || !L || !out) return fail(RADIS_LDM_EINVAL, "null argument");
  auto* db = new radis_ldm_db();
  db->ctx = nullptr;  // attached (and counted) once the device copy exists
  DbHostArrays A;
  int rc = db_prepare(L, db, A);
  if (rc) {
    delete db;
    return rc;
  }
  const int64_t n = db->n;
  const int nfull = A.nfull;
  const std::vector<float4>& sky_vals = A.sky_vals;
  const std::vector<int>& sky_off = A.sky_off;
  const std::vector<int2>& iso_rng = A.iso_rng;
  if (hipSetDevice(ctx->device) != hipSuccess) {
    delete db;
    return fail(RADIS_LDM_EHIP, "hipSetDevice(%d) failed", ctx->device);
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
  choose_split_b(&R, 1, db->ctx->n_cu);
  size_t total = 0;
  rc = layout_request(R, total);
  if (rc) return rc;
  // the batch's work items and job table (16 B per item, a few per tile and block)
  int64_t b = (int64_t)total + (int64_t)R.n_tiles * R.fill_groups * 16;
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
  rc = refresh_jobs(b, slot, b->outs);
  if (rc) return rc;
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
  rc = refresh_jobs(b, slot, b->graph_outs);
  if (rc) return rc;
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
  const long long cols = R.nv + 1;
  const long long n = (long long)rows * cols;
  hipLaunchKernelGGL(k_copy_ldm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float*)((char*)b->work + R.off_ldm), (long long)R.stride, rows, cols, dst);
  HIPCHK(hipGetLastError());
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
  choose_split_b(&R, 1, ctx->n_cu);
  size_t work = 0;
  rc = layout_request(R, work);
  if (rc) return rc;
  const LdmItemPlan P = item_plan(*db, R);
  const int n_items = plan_items(P);
  rc = check_layout(R, work);
  if (rc) return rc;
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
  const size_t o_twc = take((size_t)LDM_WRAP_LC * 8);
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
  R.tw_wrap = (float2*)(base + o_twc);
  LdmJob J;
  fill_job(*db, R, base + o_work, out_dev, J);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_put_job, dim3(1), dim3(64), 0, s, J, (LdmJob*)(base + o_job));
  const long long n_tw = R.plan.n1 + R.plan.n2 + nhi + 1024;
  hipLaunchKernelGGL(k_twiddles, dim3((unsigned)std::min<long long>((n_tw + 255) / 256, 4096)), dim3(256), 0, s, tw1,
                     R.plan.n1, tw2, R.plan.n2, twl, (long long)R.plan.L);
  hipLaunchKernelGGL(k_twiddles, dim3(LDM_WRAP_LC / 256), dim3(256), 0, s, (float2*)(base + o_twc), LDM_WRAP_LC,
                     (float2*)nullptr, 0, (float2*)nullptr, 0ll);
  hipLaunchKernelGGL(k_make_items, dim3((n_items + 255) / 256), dim3(256), 0, s, (int4*)(base + o_items), P, n_items);
  HIPCHK(hipGetLastError());
  const int cls_ac = R.plan.fc ? R.plan.n1 : 0;
  const int cls_b = R.plan.fc ? R.plan.n2 : (R.plan.n2 <= LDM_CAP_B ? 0 : -1);
  const LaunchGroup ga{cls_ac, 0, P.pcap * P.n_colA, 0, 0},
      gb{cls_b, 0, P.n_rowB, std::min(R.m_cap, FC_B_ROWS_LDS), R.wrap_rows},
      gc{cls_ac | (R.split_b > 1 ? FC_C_SUM : 0), 0, P.n_colC, 0, 0};
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
  X.fill_lds = std::max<size_t>(fill_tile_bytes(R.fill_rows, R.tile_w), 4);
  X.fill_grouped = R.fill_groups > 1;
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
