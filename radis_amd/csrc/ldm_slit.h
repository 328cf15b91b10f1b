// Instrumental slit convolution on the device: RADIS convolves the
// *_noslit quantities with the slit function sampled on the spectrum grid
// (radis/tools/slit.py:477-757, scipy.signal.oaconvolve(I, slit, 'same') *
// wstep).  Here the sum is direct:
//   out[i] = scale * sum_j slit[j] * in[i + c - j],  c = (k - 1) / 2,
// with in[] outside [0, n) read as zero.  A workgroup takes SLIT_OUT
// consecutive outputs; the slit is consumed in LDS chunks of SLIT_KCH taps
// together with the input window those taps reach.  Lanes read consecutive
// input words and the same slit word (a broadcast), so the LDS reads are
// conflict-free.  Slit widths are a few hundred to a few thousand grid
// points against grids of 10^5-10^6, so this is a small kernel next to the
// synthesis.  Products are FP32 and the running sums FP64, so the error is
// that of one FP32 product per tap, whatever the slit length.
#pragma once

#define SLIT_T 256
#define SLIT_PER 4                    // outputs per thread
#define SLIT_OUT (SLIT_T * SLIT_PER)  // outputs per workgroup
#define SLIT_KCH 2048                 // slit taps per LDS chunk

__global__ __launch_bounds__(SLIT_T) void k_slit_conv(const float* __restrict__ in, long long n,
                                                    const float* __restrict__ slit, int k, double scale,
                                                    float* __restrict__ out) {
  __shared__ float s_in[SLIT_OUT + SLIT_KCH];
  __shared__ float s_k[SLIT_KCH];
  const long long o0 = (long long)blockIdx.x * SLIT_OUT;
  const int c = (k - 1) / 2;
  double acc[SLIT_PER];
#pragma unroll
  for (int q = 0; q < SLIT_PER; ++q) acc[q] = 0.0;
  for (int j0 = 0; j0 < k; j0 += SLIT_KCH) {
    const int kc = (k - j0 < SLIT_KCH) ? k - j0 : SLIT_KCH;
    __syncthreads();  // the previous chunk has been consumed
    for (int e = threadIdx.x; e < kc; e += SLIT_T) s_k[e] = slit[j0 + e];
    // taps j0..j0+kc-1 of outputs o0..o0+SLIT_OUT-1 read in[base .. base+SLIT_OUT+kc-2]
    const long long base = o0 + c - j0 - kc + 1;
    for (int e = threadIdx.x; e < SLIT_OUT + kc - 1; e += SLIT_T) {
      const long long x = base + e;
      s_in[e] = (x >= 0 && x < n) ? in[x] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SLIT_PER; ++q) {
      const int lo = threadIdx.x + q * SLIT_T;  // in[o0 + lo + c - j] = s_in[lo + kc - 1 - (j - j0)]
      const float* w = s_in + lo + kc - 1;
      double a = acc[q];
      for (int jj = 0; jj < kc; ++jj) a += (double)(s_k[jj] * w[-jj]);
      acc[q] = a;
    }
  }
#pragma unroll
  for (int q = 0; q < SLIT_PER; ++q) {
    const long long o = o0 + threadIdx.x + q * SLIT_T;
    if (o < n) out[o] = (float)(acc[q] * (double)scale);
  }
}
