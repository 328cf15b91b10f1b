// LDS-staged mixed-radix Stockham FFT for CDNA4 (gfx950).
//
// A workgroup of T threads transforms B independent columns of length N
// held in LDS as buf[pad_idx(i*B + b)] (element i of column b), N*B <= CAP.
// The column index is the lane-fastest dimension and the LDS index is
// padded, so the strided Stockham accesses spread over the LDS banks.
// Radices 2, 3, 4, 5, 7, 8, 16 are supported.
//
// Stage s (radix R, Ns = product of the earlier radices) of the autosort
// Stockham DIT recursion:
//   for j in [0, N/R):  k = j mod Ns
//     v[q] = in[j + q*N/R] * W_{Ns*R}^(q*k)       q in [0, R)
//     v    = DFT_R(v)
//     out[(j - k)*R + k + q*Ns] = v[q]
// Each stage reads all of its inputs into registers before the barrier and
// writes after it, so one LDS buffer suffices.
#pragma once
#include <hip/hip_runtime.h>
#include "ldm_job.h"

// Global-memory (address space 1) view of a device pointer.  Pointers read
// from the job table are generic to the compiler, and generic (flat) loads
// and stores also count against the LDS wait counter: every LDS wait of an
// FFT stage would then also wait for the global loads in flight.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> glb(T* p) {
  return (gptr<T>)p;
}
// element at a uniform base plus a 32-bit byte offset (float2 goes through
// a native 2-vector: the float2 class's own copy and assignment operators
// take generic pointers)
typedef float gf2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float gld(const float* p, unsigned byte_off) {
  return *(gptr<const float>)((const char*)p + byte_off);
}
__device__ __forceinline__ float2 gld(const float2* p, unsigned byte_off) {
  const gf2 v = *(gptr<const gf2>)((const char*)p + byte_off);
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ int gld(const int* p, unsigned byte_off) {
  return *(gptr<const int>)((const char*)p + byte_off);
}
__device__ __forceinline__ void gst(float* p, unsigned byte_off, float v) {
  *(gptr<float>)((char*)p + byte_off) = v;
}
__device__ __forceinline__ void gst(float2* p, unsigned byte_off, float2 v) {
  gf2 w;
  w.x = v.x;
  w.y = v.y;
  *(gptr<gf2>)((char*)p + byte_off) = w;
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by W_4 = -i (forward) or +i (inverse)
template <bool FWD>
__device__ __forceinline__ float2 mul_w4(float2 z) {
  return FWD ? make_float2(z.y, -z.x) : make_float2(-z.y, z.x);
}

template <bool FWD>
__device__ __forceinline__ void dft2(float2* v) {
  float2 t = v[0];
  v[0] = cadd(t, v[1]);
  v[1] = csub(t, v[1]);
}

template <bool FWD>
__device__ __forceinline__ void dft4(float2* v) {
  float2 a0 = cadd(v[0], v[2]);
  float2 a1 = csub(v[0], v[2]);
  float2 a2 = cadd(v[1], v[3]);
  float2 a3 = mul_w4<FWD>(csub(v[1], v[3]));
  v[0] = cadd(a0, a2);
  v[2] = csub(a0, a2);
  v[1] = cadd(a1, a3);
  v[3] = csub(a1, a3);
}

template <bool FWD>
__device__ __forceinline__ void dft8(float2* v) {
  const float c = 0.70710678118654752440f;
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<FWD>(e);
  dft4<FWD>(o);
  float2 o1, o3;
  if (FWD) {
    o1 = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
    o3 = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
  } else {
    o1 = make_float2(c * (o[1].x - o[1].y), c * (o[1].x + o[1].y));
    o3 = make_float2(-c * (o[3].x + o[3].y), c * (o[3].x - o[3].y));
  }
  float2 o2 = mul_w4<FWD>(o[2]);
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// cos/sin(2*pi*j/16), j = 0..15
__device__ __forceinline__ float2 w16(int j, bool fwd) {
  const float cs[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                        0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                        -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
                        0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
  const float sn = cs[(j + 12) & 15];  // sin(x) = cos(x - pi/2)
  return make_float2(cs[j & 15], fwd ? -sn : sn);
}

template <bool FWD>
__device__ __forceinline__ void dft16(float2* v) {
  // 16 = 4 x 4: n = b + 4a, k = k1 + 4*k2
  float2 t[4][4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    float2 u[4] = {v[b], v[b + 4], v[b + 8], v[b + 12]};
    dft4<FWD>(u);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) t[b][k1] = (b * k1 == 0) ? u[k1] : cmul(u[k1], w16(b * k1, FWD));
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 u[4] = {t[0][k1], t[1][k1], t[2][k1], t[3][k1]};
    dft4<FWD>(u);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = u[k2];
  }
}

// Direct DFT for the odd radices 3, 5, 7 with exact constant tables.
template <int R>
struct OddTab;
template <>
struct OddTab<3> {
  static __device__ __forceinline__ float c(int j) { return j == 0 ? 1.0f : -0.5f; }
  static __device__ __forceinline__ float s(int j) {
    return j == 0 ? 0.0f : (j == 1 ? 0.86602540378443865f : -0.86602540378443865f);
  }
};
template <>
struct OddTab<5> {
  static __device__ __forceinline__ float c(int j) {
    const float t[5] = {1.0f, 0.30901699437494742f, -0.80901699437494742f, -0.80901699437494742f,
                        0.30901699437494742f};
    return t[j];
  }
  static __device__ __forceinline__ float s(int j) {
    const float t[5] = {0.0f, 0.95105651629515357f, 0.58778525229247313f, -0.58778525229247313f,
                        -0.95105651629515357f};
    return t[j];
  }
};
template <>
struct OddTab<7> {
  static __device__ __forceinline__ float c(int j) {
    const float t[7] = {1.0f, 0.62348980185873353f, -0.22252093395631440f, -0.90096886790241913f,
                        -0.90096886790241913f, -0.22252093395631440f, 0.62348980185873353f};
    return t[j];
  }
  static __device__ __forceinline__ float s(int j) {
    const float t[7] = {0.0f, 0.78183148246802981f, 0.97492791218182361f, 0.43388373911755812f,
                        -0.43388373911755812f, -0.97492791218182361f, -0.78183148246802981f};
    return t[j];
  }
};

template <int R, bool FWD>
__device__ __forceinline__ void dft_odd(float2* v) {
  float2 out[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    float2 acc = v[0];
#pragma unroll
    for (int n = 1; n < R; ++n) {
      const int j = (n * k) % R;
      const float2 w = make_float2(OddTab<R>::c(j), FWD ? -OddTab<R>::s(j) : OddTab<R>::s(j));
      acc = cadd(acc, cmul(v[n], w));
    }
    out[k] = acc;
  }
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = out[k];
}

template <int R, bool FWD>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) dft2<FWD>(v);
  else if constexpr (R == 4) dft4<FWD>(v);
  else if constexpr (R == 8) dft8<FWD>(v);
  else if constexpr (R == 16) dft16<FWD>(v);
  else dft_odd<R, FWD>(v);
}

// LDS slot of element e: two complex of padding after every 32 break the
// power-of-two strides of the Stockham stages across the LDS banks.
__device__ __forceinline__ int pad_idx(int e) { return e + ((e >> 5) << 1); }
template <int CAP>
struct LdsSize {
  static constexpr int value = CAP + (CAP >> 4);
};

template <int R, bool FWD, int CAP, int T>
__device__ __forceinline__ void stockham_stage_body(float2* __restrict__ buf, int N, int bshift, int Ns,
                                                    const float2* __restrict__ tw) {
  constexpr int MAXIT = (CAP + R * T - 1) / (R * T);
  const int tid = threadIdx.x;
  const int NR = N / R;
  const int bmask = (1 << bshift) - 1;
  const int nb = NR << bshift;
  const int twstride = N / (Ns * R);
  float2 v[MAXIT][R];
#pragma unroll
  for (int s = 0; s < MAXIT; ++s) {
    const int it = tid + s * T;
    if (it < nb) {
      const int j = it >> bshift;
      const int b = it & bmask;
#pragma unroll
      for (int q = 0; q < R; ++q) v[s][q] = buf[pad_idx(((j + q * NR) << bshift) + b)];
      if (Ns > 1) {
        // W_{Ns*R}^(q*k) as powers of one table entry: w^2m = (w^m)^2,
        // w^(2m+1) = w^2m * w (at most six rounding steps for R = 16)
        const int k = j % Ns;
        float2 wp[R];
        wp[1] = tw[k * twstride];
#pragma unroll
        for (int q = 2; q < R; ++q) wp[q] = (q & 1) ? cmul(wp[q - 1], wp[1]) : cmul(wp[q >> 1], wp[q >> 1]);
#pragma unroll
        for (int q = 1; q < R; ++q) v[s][q] = FWD ? cmul(v[s][q], wp[q]) : cmulc(v[s][q], wp[q]);
      }
      dft<R, FWD>(v[s]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < MAXIT; ++s) {
    const int it = tid + s * T;
    if (it < nb) {
      const int j = it >> bshift;
      const int b = it & bmask;
      const int k = j % Ns;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) buf[pad_idx(((base + q * Ns) << bshift) + b)] = v[s][q];
    }
  }
  __syncthreads();
}

// One out-of-line function per radix keeps the register allocation of the
// radix variants separate (inlined together they spill).
template <int R, bool FWD, int CAP, int T>
__device__ __noinline__ void stockham_stage_call(float2* __restrict__ buf, int N, int bshift, int Ns,
                                                 const float2* __restrict__ tw) {
  stockham_stage_body<R, FWD, CAP, T>(buf, N, bshift, Ns, tw);
}

template <int R, bool FWD, int CAP, int T>
__device__ __forceinline__ void stockham_stage(float2* __restrict__ buf, int N, int bshift, int Ns,
                                               const float2* __restrict__ tw) {
  stockham_stage_call<R, FWD, CAP, T>(buf, N, bshift, Ns, tw);
}

// Full transform of the B = 2^bshift columns in buf (forward: exp(-2 pi i nk/N)),
// N*B <= CAP, executed by T threads.
template <bool FWD, int CAP, int T>
__device__ void lds_fft(float2* __restrict__ buf, int N, int bshift, const int* rad, int nrad,
                        const float2* __restrict__ tw) {
  int Ns = 1;
  for (int s = 0; s < nrad; ++s) {
    const int R = rad[s];
    switch (R) {
      case 16: stockham_stage<16, FWD, CAP, T>(buf, N, bshift, Ns, tw); break;
      case 8: stockham_stage<8, FWD, CAP, T>(buf, N, bshift, Ns, tw); break;
      case 4: stockham_stage<4, FWD, CAP, T>(buf, N, bshift, Ns, tw); break;
      case 2: stockham_stage<2, FWD, CAP, T>(buf, N, bshift, Ns, tw); break;
      case 3: stockham_stage<3, FWD, CAP, T>(buf, N, bshift, Ns, tw); break;
      case 5: stockham_stage<5, FWD, CAP, T>(buf, N, bshift, Ns, tw); break;
      case 7: stockham_stage<7, FWD, CAP, T>(buf, N, bshift, Ns, tw); break;
      default: break;
    }
    Ns *= R;
  }
}
