// Set-up kernels: FFT twiddle tables, and for the no-allocation entry point
// the job table and work items of one request written into its scratch
// buffer (no host copies, so the sequence is also capturable).
#pragma once
#include "ldm_fft.h"
#include "ldm_plan.h"

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

// The first `rows` LDM rows of a request de-interleaved into dst[rows][cols]
// (the LDM stores rows 2p and 2p+1 interleaved; a diagnostic copy).
__global__ __launch_bounds__(256) void k_copy_ldm(const float* __restrict__ ldm, long long stride, int rows,
                                                  long long cols, float* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * cols) return;
  const int r = (int)(i / cols);
  const long long n = i - (long long)r * cols;
  dst[i] = ldm[(long long)(r >> 1) * 2 * stride + 2 * n + (r & 1)];
}
