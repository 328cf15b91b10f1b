// Compile-time-sized LDS FFT and the four-step passes built on it (gfx950).
//
// The generic transform (ldm_fft.h) takes N and its radices at run time; here
// N is a template parameter, so the radix plan, every loop bound, every
// division and the LDS offsets of a butterfly's inputs and outputs are
// constants: an R-point butterfly reads and writes LDS through one base
// address plus immediate offsets.  Stage twiddles come from a table of
// W_N^e (e < N) staged in LDS once per workgroup.
//
// Plan for N = 2^a * f, f in {1, 3, 5}: the odd factor f first (no twiddles),
// then radix-16 stages, then one radix-2/4/8 stage for the remaining power.
// Columns: B = 2^bs independent columns of length N, element i of column b at
// LDS slot pad_idx(i*B + b) (the same layout as the generic transform).
#pragma once
#include "ldm_fft.h"

constexpr int fc_odd(int n) { return n % 3 == 0 ? 3 : (n % 5 == 0 ? 5 : 1); }
constexpr int fc_log2(int n) { return n <= 1 ? 0 : 1 + fc_log2(n >> 1); }
constexpr bool fc_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
// number of stages and radix of stage s
constexpr int fc_nstages(int n) {
  const int f = fc_odd(n);
  const int a = fc_log2(n / f);
  return (f > 1 ? 1 : 0) + a / 4 + (a % 4 ? 1 : 0);
}
constexpr int fc_radix(int n, int s) {
  const int f = fc_odd(n);
  const int a = fc_log2(n / f);
  if (f > 1) {
    if (s == 0) return f;
    --s;
  }
  if (s < a / 4) return 16;
  return 1 << (a % 4);
}
// product of the radices before stage s
constexpr int fc_ns(int n, int s) { return s == 0 ? 1 : fc_ns(n, s - 1) * fc_radix(n, s - 1); }
// the sizes with a compiled transform
constexpr bool fc_supported(int n) {
  const int f = fc_odd(n);
  return n % f == 0 && fc_pow2(n / f) && n / f >= 64;
}
constexpr int fc_pow2_floor(int n) { return n < 2 ? 1 : 2 * fc_pow2_floor(n >> 1); }

// padded LDS offset of a constant element offset c (valid as an addition to
// pad_idx(e) when c is a multiple of 32, or when e mod 32 + c mod 32 < 32)
constexpr int pad_c(int c) { return c + ((c >> 5) << 1); }

// One Stockham DIT stage (see ldm_fft.h for the recursion) with compile-time
// N, radix R, Ns, B and T.
template <int N, int B, int T, int R, int NS, bool FWD>
__device__ __forceinline__ void fc_stage(float2* __restrict__ buf, const float2* __restrict__ tw) {
  constexpr int NR = N / R;
  constexpr int ITEMS = NR * B;
  constexpr int IT = (ITEMS + T - 1) / T;
  constexpr int LB = fc_log2(B);
  // read offsets q*NR*B pad linearly when NR*B is a multiple of 32
  constexpr bool RD_LIN = (NR * B) % 32 == 0;
  // write offsets q*NS*B: linear when NS*B is a multiple of 32, or when NS*B
  // divides 32 and the group base (j - k)*R*B is a multiple of 32
  constexpr bool WR_LIN = (NS * B) % 32 == 0 || (32 % (NS * B) == 0 && (NS * R * B) % 32 == 0);
  float2 v[IT][R];
#pragma unroll
  for (int s = 0; s < IT; ++s) {
    const int it = threadIdx.x + s * T;
    if (ITEMS % T == 0 || it < ITEMS) {
      if constexpr (RD_LIN) {
        const float2* src = buf + pad_idx(it);
#pragma unroll
        for (int q = 0; q < R; ++q) v[s][q] = src[pad_c(q * NR * B)];
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) v[s][q] = buf[pad_idx(it + q * NR * B)];
      }
      if constexpr (NS > 1) {
        const int j = it >> LB;
        const int k = j % NS;
        const int e = k * (N / (NS * R));
#pragma unroll
        for (int q = 1; q < R; ++q) {
          const float2 w = tw[q * e];
          v[s][q] = FWD ? cmul(v[s][q], w) : cmulc(v[s][q], w);
        }
      }
      dft<R, FWD>(v[s]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < IT; ++s) {
    const int it = threadIdx.x + s * T;
    if (ITEMS % T == 0 || it < ITEMS) {
      const int j = it >> LB, b = it & (B - 1);
      const int k = j % NS;
      const int e0 = (((j - k) * R + k) << LB) + b;
      if constexpr (WR_LIN) {
        float2* dst = buf + pad_idx(e0);
#pragma unroll
        for (int q = 0; q < R; ++q) dst[pad_c(q * NS * B)] = v[s][q];
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) buf[pad_idx(e0 + q * NS * B)] = v[s][q];
      }
    }
  }
  __syncthreads();
}

template <int N, int B, int T, bool FWD, int S = 0>
__device__ __forceinline__ void fc_fft(float2* __restrict__ buf, const float2* __restrict__ tw) {
  if constexpr (S < fc_nstages(N)) {
    fc_stage<N, B, T, fc_radix(N, S), fc_ns(N, S), FWD>(buf, tw);
    fc_fft<N, B, T, FWD, S + 1>(buf, tw);
  }
}

// stage the forward twiddle table W_N^e, e < N, into LDS
template <int N, int T>
__device__ __forceinline__ void fc_load_tw(float2* __restrict__ s_tw, const float2* __restrict__ tw) {
#pragma unroll
  for (int e = threadIdx.x; e < N; e += T) s_tw[e] = tw[e];
}

// ---------------------------------------------------------------------------
// Pass A with N1 fixed (see k_fft_a): B = fc_cols_a(N1) columns per workgroup.
// ---------------------------------------------------------------------------
#define FC_CAP 8192   // complex elements of a pass A / C tile (at most)
#define FC_T 512
#define FC_CAP_B 2048 // complex elements of a pass B tile
#define FC_T_B 256
// columns per workgroup of pass A / C for N1, and rows per workgroup of pass B for N2
constexpr int fc_cols_a(int n1) { return fc_pow2_floor(FC_CAP / n1); }
constexpr int fc_rows_b(int n2) { return FC_CAP_B / n2; }

template <int N1>
__global__ __launch_bounds__(FC_T, 4) void k_fft_a_c(const LdmJob* __restrict__ jobs, const LdmAxes* __restrict__ axes,
                                                     const int4* __restrict__ items, int n_items) {
  constexpr int B = fc_cols_a(N1), LB = fc_log2(B), CAPE = N1 * B;
  static_assert(CAPE % FC_T == 0, "tile is a whole number of thread passes");
  __shared__ float2 buf[LdsSize<CAPE>::value];
  __shared__ float2 s_tw[N1];
  const int4 it = items[xcd_block(blockIdx.x, n_items)];
  const LdmJob& J = jobs[it.x];
  const int p = it.y;
  if (J.status || p >= axes[it.x].p) return;
  const int N2 = J.n2;
  const int n2_0 = it.z * B;
  const float* ra = J.ldm + (long long)(2 * p) * J.ldm_stride;
  const bool has_b = (2 * p + 1) < axes[it.x].m;
  const float* rb = ra + J.ldm_stride;
  const int V = (int)J.v;
  fc_load_tw<N1, FC_T>(s_tw, J.tw1);
#pragma unroll
  for (int s = 0; s < CAPE / FC_T; ++s) {
    const int e = threadIdx.x + s * FC_T;
    const int n = n2_0 + (e & (B - 1)) + N2 * (e >> LB);
    float xa = 0.0f, xb = 0.0f;
    if (n < V) {
      xa = ra[n];
      if (has_b) xb = rb[n];
    }
    buf[pad_idx(e)] = make_float2(xa, xb);
  }
  __syncthreads();
  fc_fft<N1, B, FC_T, true>(buf, s_tw);
  float2* Y = J.ybuf + (long long)p * J.L;
#pragma unroll
  for (int s = 0; s < CAPE / FC_T; ++s) {
    const int e = threadIdx.x + s * FC_T;
    const int k1 = e >> LB, n2 = n2_0 + (e & (B - 1));
    Y[(long long)k1 * N2 + n2] = cmul(buf[pad_idx(e)], twiddle_L(J, (long long)n2 * k1));
  }
}

// ---------------------------------------------------------------------------
// Pass B with N2 fixed (see k_fft_b): B = FC_CAP_B / N2 rows k1 per workgroup.
// ---------------------------------------------------------------------------
template <int N2>
__global__ __launch_bounds__(FC_T_B, 2) void k_fft_b_c(const LdmJob* __restrict__ jobs,
                                                      const LdmAxes* __restrict__ axes,
                                                      const int4* __restrict__ items) {
  constexpr int CAP = FC_CAP_B, T = FC_T_B;
  constexpr int B = fc_rows_b(N2), LB = fc_log2(B), LN2 = fc_log2(N2);
  constexpr int NACC = CAP / T;
  static_assert(fc_pow2(N2), "pass B rows are powers of two");
  __shared__ float2 buf[LdsSize<CAP>::value];
  __shared__ float2 s_tw[N2];
  const int4 it = items[blockIdx.x];
  const LdmJob& J = jobs[it.x];
  if (J.status) return;
  const int N1 = J.n1;
  const int k1_0 = it.y * B;
  const long long L = J.L;
  const int P = axes[it.x].p, M = axes[it.x].m;
  fc_load_tw<N2, T>(s_tw, J.tw2);
  float2 acc[NACC];
#pragma unroll
  for (int s = 0; s < NACC; ++s) acc[s] = make_float2(0.0f, 0.0f);
  // element e of a pair's block: row b = e >> log2(N2), column n2 (fastest
  // in global memory), stored transposed at LDS [n2][b]; the next pair's
  // block is fetched into registers while this one is transformed
  const float2* Yb = J.ybuf + (long long)k1_0 * N2;
  float2 pre[NACC];
  if (P > 0) {
#pragma unroll
    for (int s = 0; s < NACC; ++s) pre[s] = Yb[threadIdx.x + s * T];
  }
  for (int p = 0; p < P; ++p) {
#pragma unroll
    for (int s = 0; s < NACC; ++s) {
      const int e = threadIdx.x + s * T;
      buf[pad_idx(((e & (N2 - 1)) << LB) + (e >> LN2))] = pre[s];
    }
    __syncthreads();
    if (p + 1 < P) {
      const float2* Yn = Yb + (long long)(p + 1) * L;
#pragma unroll
      for (int s = 0; s < NACC; ++s) pre[s] = Yn[threadIdx.x + s * T];
    }
    fc_fft<N2, B, T, true>(buf, s_tw);
    const int ra = 2 * p, rb = 2 * p + 1;
#pragma unroll
    for (int s = 0; s < NACC; ++s) {
      const int e = threadIdx.x + s * T;
      const int k2 = e >> LB, b = e & (B - 1);
      const long long k = (long long)(k1_0 + b) + (long long)N1 * k2;
      const long long kk = (k <= L - k) ? k : (L - k);
      const float fa = lineshape_ft(J, ra, kk);
      const float fb = (rb < M) ? lineshape_ft(J, rb, kk) : 0.0f;
      const float2 z = buf[pad_idx(e)];
      acc[s].x += z.x * fa + z.y * fb;
      acc[s].y += z.y * fa - z.x * fb;
    }
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < NACC; ++s) buf[pad_idx(threadIdx.x + s * T)] = acc[s];
  __syncthreads();
  fc_fft<N2, B, T, false>(buf, s_tw);
  float2* Bo = J.bbuf + (long long)k1_0 * N2;
#pragma unroll
  for (int s = 0; s < NACC; ++s) {
    const int e = threadIdx.x + s * T;
    const int b = e >> LN2, n2 = e & (N2 - 1);
    Bo[e] = cmulc(buf[pad_idx((n2 << LB) + b)], twiddle_L(J, (long long)n2 * (k1_0 + b)));
  }
}

// ---------------------------------------------------------------------------
// Pass C with N1 fixed (see k_fft_c).
// ---------------------------------------------------------------------------
template <int N1>
__global__ __launch_bounds__(FC_T, 4) void k_fft_c_c(const LdmJob* __restrict__ jobs, const int4* __restrict__ items,
                                                     int n_items) {
  constexpr int B = fc_cols_a(N1), LB = fc_log2(B), CAPE = N1 * B;
  static_assert(CAPE % FC_T == 0, "tile is a whole number of thread passes");
  __shared__ float2 buf[LdsSize<CAPE>::value];
  __shared__ float2 s_tw[N1];
  const int4 it = items[xcd_block(blockIdx.x, n_items)];
  const LdmJob& J = jobs[it.x];
  const int N2 = J.n2;
  const int n2_0 = it.z * B;
  const int nv = (int)J.nv;
  if (J.status) {
#pragma unroll
    for (int s = 0; s < CAPE / FC_T; ++s) {
      const int e = threadIdx.x + s * FC_T;
      const int n = n2_0 + (e & (B - 1)) + N2 * (e >> LB);
      if (n < nv) J.out[n] = 0.0f;
    }
    return;
  }
  fc_load_tw<N1, FC_T>(s_tw, J.tw1);
#pragma unroll
  for (int s = 0; s < CAPE / FC_T; ++s) {
    const int e = threadIdx.x + s * FC_T;
    buf[pad_idx(e)] = J.bbuf[(long long)(e >> LB) * N2 + n2_0 + (e & (B - 1))];
  }
  __syncthreads();
  fc_fft<N1, B, FC_T, false>(buf, s_tw);
  const float sc = J.out_scale;
#pragma unroll
  for (int s = 0; s < CAPE / FC_T; ++s) {
    const int e = threadIdx.x + s * FC_T;
    const int n = n2_0 + (e & (B - 1)) + N2 * (e >> LB);
    if (n < nv) J.out[n] = buf[pad_idx(e)].x * sc;
  }
}
