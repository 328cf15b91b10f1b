// Set-up kernels: FFT twiddle tables, and for the no-allocation entry point
// the job table and work items of one request written into its scratch
// buffer (no host copies, so the sequence is also capturable).
#pragma once
#include "ldm_fft.h"

// Twiddle tables of one size set: W_n1^e (e < n1), W_n2^e (e < n2) and the
// two-level W_L table: hi[q] = W_L^((q * 1024) mod L), q < ceil(L / 1024),
// then lo[e] = W_L^e, e < 1024.  W_N^e = exp(-2 pi i e / N), evaluated as
// sincospi(-2 e / N) in FP64 and rounded to FP32.  Null tables are skipped.
__global__ __launch_bounds__(256) void k_twiddles(float2* __restrict__ tw1, int n1, float2* __restrict__ tw2, int n2,
                                                  float2* __restrict__ twl, long long L) {
  const long long nhi = (L + (1 << LDM_TW_LO_BITS) - 1) >> LDM_TW_LO_BITS;
  const long long total = (tw1 ? n1 : 0) + (tw2 ? n2 : 0) + (twl ? nhi + (1 << LDM_TW_LO_BITS) : 0);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    long long k = i;
    float2* dst;
    double e, N;
    if (tw1 && k < n1) {
      dst = tw1 + k; e = (double)k; N = (double)n1;
    } else {
      if (tw1) k -= n1;
      if (tw2 && k < n2) {
        dst = tw2 + k; e = (double)k; N = (double)n2;
      } else {
        if (tw2) k -= n2;
        dst = twl + k;
        N = (double)L;
        e = k < nhi ? (double)((k << LDM_TW_LO_BITS) % L) : (double)(k - nhi);
      }
    }
    double sn, cs;
    sincospi(-2.0 * e / N, &sn, &cs);
    *dst = make_float2((float)cs, (float)sn);
  }
}

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
  if ((k -= P.pcap * P.n_colA) < P.n_rowB) return make_int4(0, k, 0, 0);  // pass B row block
  return make_int4(0, 0, k - P.n_rowB, 0);                        // pass C column block
}

__global__ __launch_bounds__(256) void k_make_items(int4* __restrict__ items, LdmItemPlan P, int n_items) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_items) items[i] = plan_item(P, i);
}

// The job table of one request, passed by value (a kernel argument, so a
// captured launch holds its own copy).
__global__ __launch_bounds__(64) void k_put_job(LdmJob job, LdmJob* __restrict__ dst) {
  static_assert(sizeof(LdmJob) % 4 == 0, "job table copied as 32-bit words");
  const unsigned* s = (const unsigned*)&job;
  unsigned* d = (unsigned*)dst;
  for (int w = threadIdx.x; w < (int)(sizeof(LdmJob) / 4); w += blockDim.x) d[w] = s[w];
}
