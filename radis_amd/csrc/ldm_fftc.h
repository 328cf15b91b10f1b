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
// then radix-16 stages, then one radix-2/4/8 stage for the remaining power
// (pass B at N = 256 runs a variant with radix-4 end stages, fc_pv_ends4).
// Columns: B = 2^bs independent columns of length N, element i of column b at
// LDS slot pad_idx(i*B + b) (the same layout as the generic transform).
#pragma once
#include "ldm_fft.h"

constexpr int fc_odd(int n) { return n % 3 == 0 ? 3 : (n % 5 == 0 ? 5 : 1); }
constexpr int fc_log2(int n) { return n <= 1 ? 0 : 1 + fc_log2(n >> 1); }
constexpr bool fc_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
// Plan variant pv = 1 at N = 256 (pass B): radices 4, 16, 4 instead of
// 16, 16.  Pass B keeps its first and last stages' elements in registers
// (the next pair's prefetch; the accumulated spectrum and the lineshape
// FTs), so a radix-16 end stage over 8 rows of 256 held 16 elements per
// thread on half the threads (187 VGPRs); radix-4 ends hold 2 x 4 on all.
constexpr bool fc_pv_ends4(int n, int pv) { return pv == 1 && n == 256; }
// number of stages and radix of stage s
constexpr int fc_nstages(int n, int pv = 0) {
  if (fc_pv_ends4(n, pv)) return 3;
  const int f = fc_odd(n);
  const int a = fc_log2(n / f);
  return (f > 1 ? 1 : 0) + a / 4 + (a % 4 ? 1 : 0);
}
constexpr int fc_radix(int n, int s, int pv = 0) {
  if (fc_pv_ends4(n, pv)) return s == 1 ? 16 : 4;
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
constexpr int fc_ns(int n, int s, int pv = 0) { return s == 0 ? 1 : fc_ns(n, s - 1, pv) * fc_radix(n, s - 1, pv); }
// the sizes with a compiled transform
constexpr bool fc_supported(int n) {
  const int f = fc_odd(n);
  return n % f == 0 && fc_pow2(n / f) && n / f >= 64;
}

// padded LDS offset of a constant element offset c (valid as an addition to
// pad_idx(e) when c is a multiple of 32, or when e mod 32 + c mod 32 < 32)
constexpr int pad_c(int c) { return c + ((c >> 5) << 1); }

// LDS paddings: PD = 5 is pad_idx (2 slots after every 32 elements); PD = 4
// puts 1 slot after every 16, which makes the stride-16 writes of a radix-16
// stage over one column (B = 1) conflict-free.  Both add N/16 slots.
template <int PD>
__device__ __forceinline__ int pad_e(int e) {
  return PD == 5 ? pad_idx(e) : e + (e >> 4);
}
template <int PD>
constexpr int pad_ce(int c) {
  return PD == 5 ? pad_c(c) : c + (c >> 4);
}

// Per-stage twiddle tables (fc_load_tw_stages): stage s >= 1 (radix R, NS)
// holds W^(q k), q in [1, R), k < NS, as (q-1)*NS + k; a stage with
// NS >= 64 holds only q = 1 (NS entries) and its butterflies form the
// powers W^(q k) = W^(a k) W^(b k), a + b = q, a = q/2 (products at most
// log2(R) deep).  fc_tw_off is a stage's offset in the concatenated table.
constexpr bool fc_tw_compact(int n, int s, int pv = 0) { return s > 0 && fc_ns(n, s, pv) >= 64; }
constexpr int fc_tw_size(int n, int s, int pv = 0) {
  return s == 0 ? 0 : (fc_tw_compact(n, s, pv) ? fc_ns(n, s, pv) : fc_ns(n, s, pv) * (fc_radix(n, s, pv) - 1));
}
constexpr int fc_tw_off(int n, int s, int pv = 0) {
  return s <= 1 ? 0 : fc_tw_off(n, s - 1, pv) + fc_tw_size(n, s - 1, pv);
}

// One Stockham DIT stage (see ldm_fft.h for the recursion) with compile-time
// N, radix R, Ns, B and T.  Element e = i*B + b is element i of column b.
// IN_REG: the stage's inputs come from ld(s, q, e) instead of LDS (the first
// stage, fed from registers or global memory; its LDS writes need the
// buffer free, which the caller guarantees).  OUT_REG: its outputs go to
// st(s, q, e, value) instead of LDS (the last stage; no trailing barrier).
// (s, q) = (thread pass, butterfly leg), the same for a given e in every
// call, so callers can keep per-element state in registers indexed by them.
// threadIdx.x behind an empty asm barrier: the stages' per-thread LDS and
// twiddle addresses are then computed where they are used, instead of being
// hoisted out of a caller's loop and held in registers across it
__device__ __forceinline__ int fc_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// PD: the LDS padding (pad_e).  TWO >= 0: stage twiddles from a per-stage
// table (fc_load_tw_stages) at tw[TWO + (q-1)*NS + k] (TWC: only q = 1,
// powers formed here), so lanes with consecutive k read consecutive slots;
// TWO < 0: from the W_N table at tw[q*k*N/(NS*R)], whose power-of-two
// strides put a wave's lanes on a few LDS banks when B is small.
// OOP: the stage writes obuf, a buffer other than the one it reads (buf),
// so no barrier is needed between its reads and its writes.
template <int N, int B, int T, int R, int NS, bool FWD, bool IN_REG, bool OUT_REG, bool OPQ, int PD, int TWO, bool TWC,
          bool OOP, class LD, class ST>
__device__ __forceinline__ void fc_stage_io(const float2* buf, float2* obuf, const float2* __restrict__ tw, LD& ld,
                                            ST& st) {
  const int tid = OPQ ? fc_tid() : (int)threadIdx.x;
  constexpr int NR = N / R;
  constexpr int ITEMS = NR * B;
  constexpr int IT = (ITEMS + T - 1) / T;
  constexpr int LB = fc_log2(B);
  constexpr int G = 1 << PD;  // padding group
  // read offsets q*NR*B pad linearly when NR*B is a multiple of G
  constexpr bool RD_LIN = (NR * B) % G == 0;
  // write offsets q*NS*B: linear when NS*B is a multiple of G, or when NS*B
  // divides G and the group base (j - k)*R*B is a multiple of G
  constexpr bool WR_LIN = (NS * B) % G == 0 || (G % (NS * B) == 0 && (NS * R * B) % G == 0);
  float2 v[IT][R];
#pragma unroll
  for (int s = 0; s < IT; ++s) {
    const int it = tid + s * T;
    if (ITEMS % T == 0 || it < ITEMS) {
      if constexpr (IN_REG) {
#pragma unroll
        for (int q = 0; q < R; ++q) v[s][q] = ld(s, q, it + q * NR * B);
      } else if constexpr (RD_LIN) {
        const float2* src = buf + pad_e<PD>(it);
#pragma unroll
        for (int q = 0; q < R; ++q) v[s][q] = src[pad_ce<PD>(q * NR * B)];
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) v[s][q] = buf[pad_e<PD>(it + q * NR * B)];
      }
      if constexpr (NS > 1) {
        const int j = it >> LB;
        const int k = j % NS;
        const int e = k * (N / (NS * R));
        if constexpr (TWO >= 0 && TWC) {
          float2 w[R];
          w[1] = tw[TWO + k];
#pragma unroll
          for (int q = 2; q < R; ++q) w[q] = cmul(w[q / 2], w[q - q / 2]);
#pragma unroll
          for (int q = 1; q < R; ++q) v[s][q] = FWD ? cmul(v[s][q], w[q]) : cmulc(v[s][q], w[q]);
        } else {
#pragma unroll
          for (int q = 1; q < R; ++q) {
            const float2 w = TWO >= 0 ? tw[TWO + (q - 1) * NS + k] : tw[q * e];
            v[s][q] = FWD ? cmul(v[s][q], w) : cmulc(v[s][q], w);
          }
        }
      }
      dft<R, FWD>(v[s]);
    }
  }
  // LDS inputs must all be read before any output overwrites them
  if constexpr (!IN_REG && !OUT_REG && !OOP) __syncthreads();
#pragma unroll
  for (int s = 0; s < IT; ++s) {
    const int it = tid + s * T;
    if (ITEMS % T == 0 || it < ITEMS) {
      const int j = it >> LB, b = it & (B - 1);
      const int k = j % NS;
      const int e0 = (((j - k) * R + k) << LB) + b;
      if constexpr (OUT_REG) {
#pragma unroll
        for (int q = 0; q < R; ++q) st(s, q, e0 + q * NS * B, v[s][q]);
      } else if constexpr (WR_LIN) {
        float2* dst = obuf + pad_e<PD>(e0);
#pragma unroll
        for (int q = 0; q < R; ++q) dst[pad_ce<PD>(q * NS * B)] = v[s][q];
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) obuf[pad_e<PD>(e0 + q * NS * B)] = v[s][q];
      }
    }
  }
  if constexpr (!OUT_REG) __syncthreads();
}

struct FcNone {
  __device__ float2 operator()(int, int, int) const { return make_float2(0.0f, 0.0f); }
  __device__ void operator()(int, int, int, float2) const {}
};

// Transform of the B columns in buf.  IN_REG: the first stage reads its
// inputs through ld (then calls after_first()); OUT_REG: the last stage
// hands its outputs to st (after before_last()).  A transform with both set
// touches LDS only between its stages.
// OPQ: the stages take the thread index through fc_tid() (for callers that
// run the transform inside a loop).
// PD: the LDS padding (pad_e).  TWS: tw is the per-stage twiddle table
// (fc_load_tw_stages) instead of W_N.  Ping-pong (b1 != b0): stage S writes
// b0 for even S and b1 for odd S and reads the other one, so no stage needs
// a barrier between its reads and its writes; b1 == b0: in place.
template <int N, int B, int T, bool FWD, bool IN_REG, bool OUT_REG, bool OPQ, int PD = 5, bool TWS = false,
          bool PP = false, int PV = 0, int S = 0, class LD, class ST, class AF, class BL>
__device__ __forceinline__ void fc_fft_pp(float2* b0, float2* b1, const float2* __restrict__ tw, LD& ld, ST& st,
                                          AF& after_first, BL& before_last) {
  constexpr int NST = fc_nstages(N, PV);
  if constexpr (S < NST) {
    if constexpr (S == NST - 1) before_last();
    constexpr int TWO = (TWS && S > 0) ? fc_tw_off(N, S, PV) : -1;
    float2* ib = (PP && (S & 1) == 0) ? b1 : b0;  // stage S reads what stage S-1 wrote
    float2* ob = (PP && (S & 1)) ? b1 : b0;
    fc_stage_io<N, B, T, fc_radix(N, S, PV), fc_ns(N, S, PV), FWD, IN_REG && S == 0, OUT_REG && S == NST - 1, OPQ, PD,
                TWO, fc_tw_compact(N, S, PV), PP>(ib, ob, tw, ld, st);
    if constexpr (S == 0) after_first();
    fc_fft_pp<N, B, T, FWD, IN_REG, OUT_REG, OPQ, PD, TWS, PP, PV, S + 1>(b0, b1, tw, ld, st, after_first,
                                                                          before_last);
  }
}
template <int N, int B, int T, bool FWD, bool IN_REG, bool OUT_REG, bool OPQ, int PD = 5, bool TWS = false, class LD,
          class ST, class AF, class BL>
__device__ __forceinline__ void fc_fft_io(float2* __restrict__ buf, const float2* __restrict__ tw, LD& ld, ST& st,
                                          AF& after_first, BL& before_last) {
  fc_fft_pp<N, B, T, FWD, IN_REG, OUT_REG, OPQ, PD, TWS, false>(buf, buf, tw, ld, st, after_first, before_last);
}
template <int N, int B, int T, bool FWD, bool IN_REG, bool OUT_REG, class LD, class ST, class AF>
__device__ __forceinline__ void fc_fft_io(float2* __restrict__ buf, const float2* __restrict__ tw, LD& ld, ST& st,
                                          AF& after_first) {
  auto nop = [] {};
  fc_fft_io<N, B, T, FWD, IN_REG, OUT_REG, false>(buf, tw, ld, st, after_first, nop);
}

template <int N, int B, int T, bool FWD>
__device__ __forceinline__ void fc_fft(float2* __restrict__ buf, const float2* __restrict__ tw) {
  FcNone none;
  auto nop = [] {};
  fc_fft_io<N, B, T, FWD, false, false>(buf, tw, none, none, nop);
}

// (thread pass, leg) extents of a transform's first and last stages, for
// callers that keep per-element values in registers
template <int N, int B, int T, int PV = 0>
struct FcShape {
  static constexpr int R0 = fc_radix(N, 0, PV), IT0 = (N / R0 * B + T - 1) / T;
  static constexpr int RL = fc_radix(N, fc_nstages(N, PV) - 1, PV), ITL = (N / RL * B + T - 1) / T;
  static constexpr int ITEMS0 = N / R0 * B, ITEMSL = N / RL * B;
};

// stage the forward twiddle table W_N^e, e < N, into LDS
template <int N, int T>
__device__ __forceinline__ void fc_load_tw(float2* __restrict__ s_tw, const float2* __restrict__ tw) {
#pragma unroll
  for (int e = threadIdx.x; e < N; e += T) s_tw[e] = gld(tw, 8u * (unsigned)e);
}

// Per-stage twiddle tables from the W_N table tw (global): stage
// S >= 1 (radix R, NS) holds W_N^(q k N/(NS R)) for k < NS and q in
// [1, R) (q = 1 only when fc_tw_compact) at fc_tw_off(N, S) + (q-1)*NS + k.
// The tables add up to fewer than N entries (fc_tw_off(N, fc_nstages(N))).
template <int N, int T, int PV = 0, int S = 1>
__device__ __forceinline__ void fc_load_tw_stages(float2* __restrict__ s_tw, const float2* __restrict__ tw) {
  if constexpr (S < fc_nstages(N, PV)) {
    constexpr int NS = fc_ns(N, S, PV), R = fc_radix(N, S, PV), LNS = fc_log2(NS);
    constexpr int O = fc_tw_off(N, S, PV), STEP = N / (NS * R), SZ = fc_tw_size(N, S, PV);
    static_assert(fc_pow2(NS), "stages after the first have power-of-two NS");
    for (int i = threadIdx.x; i < SZ; i += T) {
      const int q = (i >> LNS) + 1, k = i & (NS - 1);
      s_tw[O + i] = gld(tw, 8u * (unsigned)(q * k * STEP));
    }
    fc_load_tw_stages<N, T, PV, S + 1>(s_tw, tw);
  }
}

// Twiddles W_L^(m * (j0 + s*DS + q*DQ)) of one thread's last-stage outputs
// (s < IT thread passes, q < R legs; m fixed per thread) as
// base * A^s * C^q from three table values, so an output costs one complex
// product instead of two table gathers.  The longest product chain has
// IT + R - 2 factors (~1e-7 rounding each).  Every exponent is below L.
template <int IT, int R>
struct TwProg {
  float2 ws[IT], wq[R];
  __device__ __forceinline__ void init(const LdmJob& J, long long m, long long j0, long long DS, long long DQ) {
    set(twiddle_L(J, m * j0), twiddle_L(J, m * DS), twiddle_L(J, m * DQ));
  }
  // from the three table values W^(m j0), W^(m DS), W^(m DQ)
  __device__ __forceinline__ void set(float2 base, float2 A, float2 C) {
    ws[0] = base;
#pragma unroll
    for (int s = 1; s < IT; ++s) ws[s] = cmul(ws[s - 1], A);
    wq[0] = make_float2(1.0f, 0.0f);
#pragma unroll
    for (int q = 1; q < R; ++q) wq[q] = cmul(wq[q - 1], C);
  }
  __device__ __forceinline__ float2 operator()(int s, int q) const { return cmul(ws[s], wq[q]); }
};

// ---------------------------------------------------------------------------
// Pass A with N1 fixed (see k_fft_a): B = fc_cols_a(N1) columns per workgroup.
// ---------------------------------------------------------------------------
// Pass A reads B consecutive elements of each of a column block's N1 rows
// (stride N2); an element is the float2 of the pair's two interleaved LDM
// rows, so a row segment is 8 B bytes.  Segments of 128 B or more stream at
// close to the full HBM rate, 64 B ones at about 0.6 of it and 32 B ones at
// about 0.35, so pass A takes the widest tile that fits LDS (one
// 1024-thread workgroup per CU).
// (tile sizes FC_CAP_A, FC_CAP, FC_CAP_B and fc_cols_a / fc_cols_c /
// fc_rows_b are in ldm_job.h: the host planner sizes the work items with them)

template <int N1>
__global__ __launch_bounds__(FC_T_A) void k_fft_a_c(const LdmJob* __restrict__ jobs, const LdmAxes* __restrict__ axes,
                                                   const int4* __restrict__ items, int n_items) {
  constexpr int B = fc_cols_a(N1), LB = fc_log2(B), CAPE = N1 * B, T = FC_T_A;
  static_assert(CAPE % T == 0, "tile is a whole number of thread passes");
  using SH = FcShape<N1, B, T>;
  constexpr int NR0 = N1 / SH::R0;
  constexpr int NSL = fc_ns(N1, fc_nstages(N1) - 1);
  __shared__ float2 buf[LdsSize<CAPE>::value];
  __shared__ float2 s_tw[N1];
  // Persistent: workgroup g takes the column blocks g', g' + G, g' + 2G, ...
  // (G = gridDim.x <= n_items).  Blocks of fewer than 32 columns take
  // g' = xcd_block(g, G): the workgroups of one XCD then take consecutive
  // blocks.  That was worth 339 -> 330 us at N1 = 640 when a 16-column
  // block read 64-B row segments; with the LDM's interleaved row pairs it
  // reads 128 B per row and the remap measures neutral.  32-column blocks
  // keep g' = g (the remap cost them 2.6 us at N1 = 512).  A block's first-stage inputs
  // sit in v; the next block's loads go into v right after the first stage
  // has consumed them, so they are in flight during the LDS stages and the
  // stores.
  float2 v[SH::IT0][SH::R0];
  // issue the loads of block idx into v; false (for the whole workgroup)
  // when the block has no work
  auto fetch = [&](int idx) -> bool {
    if (idx >= n_items) return false;
    const int4 it = items[idx];
    const LdmJob& J = jobs[it.x];
    const int p = it.y;
    if (J.status || p >= axes[it.x].p) return false;
    const int N2 = J.n2, n2_0 = it.z * B, V = (int)J.v;
    // pair p's rows, interleaved: z = (x_2p, x_2p+1) is one float2
    const float2* rz = reinterpret_cast<const float2*>(J.ldm) + (long long)p * J.ldm_stride;
    const bool has_b = (2 * p + 1) < axes[it.x].m;
    // L >= 2 N_v, so V = N_v + 1 <= L/2 unless L = 2 N_v exactly: then the
    // upper half of every column (n1 >= N1/2, n >= L/2) is zero and its
    // loads are skipped
    const bool upper_zero = (long long)V <= J.L / 2;
    // the two LDM rows, zero beyond V (and row b zero when the pair is odd)
#pragma unroll
    for (int s = 0; s < SH::IT0; ++s) {
      const int it0 = threadIdx.x + s * T;
#pragma unroll
      for (int q = 0; q < SH::R0; ++q) {
        // s*(T/B) + q*NR0 is the smallest row n1 = e >> LB of this (s, q)
        if (upper_zero && s * (T >> LB) + q * NR0 >= N1 / 2) {
          v[s][q] = make_float2(0.0f, 0.0f);
          continue;
        }
        const int e = it0 + q * NR0 * B;
        const int n = n2_0 + (e & (B - 1)) + N2 * (e >> LB);
        const bool in = (SH::ITEMS0 % T == 0 || it0 < SH::ITEMS0) && n < V;
        const unsigned o = in ? 8u * (unsigned)n : 0u;  // n < V < 2^28
        const float2 z = gld(rz, o);
        v[s][q] = make_float2(in ? z.x : 0.0f, (in && has_b) ? z.y : 0.0f);
      }
    }
    return true;
  };
  // W_N1 table: the same for every request of this N1
  const int first = B < 32 ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  fc_load_tw<N1, T>(s_tw, jobs[items[first].x].tw1);
  bool live = fetch(first);
  for (int idx = first; idx < n_items; idx += gridDim.x) {
    if (!live) {
      live = fetch(idx + gridDim.x);
      continue;
    }
    const int4 it = items[idx];
    const LdmJob& J = jobs[it.x];
    const int N2 = J.n2, n2_0 = it.z * B;
    float2* Y = J.ybuf + (long long)it.y * J.L;
    // output k1 of this thread's last-stage element (s, q): j0 + s*(T/B) + q*NS;
    // the table halves of the progression's three factors load now and are
    // multiplied out just before the last stage
    float2 twr[3][2];
    {
      const unsigned m = n2_0 + (threadIdx.x & (B - 1));
      const unsigned ex[3] = {m * (threadIdx.x >> LB), m * (T >> LB), m * NSL};  // < L <= 2^21
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        twr[t][0] = gld(J.twl_hi, 8u * (ex[t] >> LDM_TW_LO_BITS));
        twr[t][1] = gld(J.twl_lo, 8u * (ex[t] & ((1u << LDM_TW_LO_BITS) - 1)));
      }
    }
    TwProg<SH::ITL, SH::RL> tw;
    auto before_last = [&] {
      tw.set(cmul(twr[0][0], twr[0][1]), cmul(twr[1][0], twr[1][1]), cmul(twr[2][0], twr[2][1]));
    };
    auto ld = [&](int s, int q, int) -> float2 { return v[s][q]; };
    auto st = [&](int s, int q, int e, float2 z) {
      const unsigned k1 = (unsigned)e >> LB, n2 = n2_0 + (e & (B - 1));
      // byte offsets are below 8 L <= 2^24
      gst(Y, 8u * (k1 * (unsigned)N2 + n2), cmul(z, tw(s, q)));
    };
    bool next = false;
    auto prefetch = [&] { next = fetch(idx + gridDim.x); };
    // first stage from v, last stage straight to Y with the W_L twiddle
    fc_fft_io<N1, B, T, true, true, true, true>(buf, s_tw, ld, st, prefetch, before_last);
    live = next;
    // the last stage read LDS; the next block's first stage writes it
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Pass B with N2 fixed (see k_fft_b): B = FC_CAP_B / N2 rows k1 per workgroup.
// Per pair the forward transform's first stage reads the pair's block from
// registers (fetched during the previous pair) and its last stage multiplies
// by the lineshape FTs and accumulates in registers; only the accumulated
// spectrum goes through LDS into the inverse transform, whose last stage
// writes B with the W_L twiddle.  Item (job, row block, share): with
// J.bbuf2 set, share 0 takes the first ceil(P/2) pairs into B and share 1
// the rest into bbuf2 (pass C adds the two).  The dynamic LDS (sized per
// launch) holds the rows' lineshape coefficients (n_rows_lds float4), then
// the image-correction values of the block's rows at its two corrected bins
// (2 x B x n_wh floats, n_wh >= N_L of every request with a correction).
// ---------------------------------------------------------------------------
template <int N2>
__global__ __launch_bounds__(FC_T_B, (N2 >= 1024 || N2 == 256) ? 4 : 2) void k_fft_b_c(const LdmJob* __restrict__ jobs,
                                                      const LdmAxes* __restrict__ axes,
                                                      const int4* __restrict__ items, int n_rows_lds, int n_wh) {
  constexpr int CAP = FC_CAP_B, T = FC_T_B;
  constexpr int B = fc_rows_b(N2), LB = fc_log2(B);
  static_assert(fc_pow2(N2), "pass B rows are powers of two");
  constexpr int PV = N2 == 256 ? 1 : 0;  // plan variant (fc_pv_ends4)
  using SH = FcShape<N2, B, T, PV>;
  constexpr int NR0 = N2 / SH::R0;
  constexpr int NSL = fc_ns(N2, fc_nstages(N2, PV) - 1, PV);
  // Two transform buffers, used ping-pong (stage s writes one, stage s+1
  // reads it and writes the other): no stage needs a barrier between its
  // reads and its writes, and the next pair's first stage writes a buffer
  // whose last readers have passed a stage barrier.  With an even number of
  // stages the last stage reads the first stage's buffer, so the pairs
  // alternate the two roles.
  __shared__ float2 bufs[2][LdsSize<CAP>::value];
  constexpr bool SWAP = fc_nstages(N2, PV) % 2 == 0;
  __shared__ float2 s_tw[fc_tw_off(N2, fc_nstages(N2, PV), PV)];  // per-stage twiddle tables
  const int4 it = items[blockIdx.x];
  const LdmJob& J = jobs[it.x];
  if (J.status) return;
  const int N1 = J.n1;
  const int k1_0 = it.y * B;
  const long long L = J.L;
  const int P = axes[it.x].p, M = axes[it.x].m;
  const int p_mid = J.bbuf2 ? (P + 1) / 2 : P;
  const int p0 = it.z ? p_mid : 0, p1 = it.z ? P : p_mid;  // this workgroup's pairs
  fc_load_tw_stages<N2, T, PV>(s_tw, J.tw2);
  // The rows' lineshape coefficients (a_l, b_m, 1/F(0), fold count) staged
  // in LDS.  Read from global memory they are scalar loads, and the LDS
  // waits of the transform stages then also wait for them.
  extern __shared__ float4 s_row[];
  const bool rows_lds = M <= n_rows_lds;
  if (rows_lds) {
    for (int r = threadIdx.x; r < M; r += T) {
      const unsigned o = 4u * (unsigned)r;
      s_row[r] = make_float4(gld(J.row_a, o), gld(J.row_b, o), gld(J.row_inv_f0, o), __int_as_float(gld(J.nfold, o)));
    }
  }
  __syncthreads();
  auto row_c = [&](int r) -> float4 {
    if (rows_lds) return s_row[r];
    const unsigned o = 4u * (unsigned)r;
    return make_float4(gld(J.row_a, o), gld(J.row_b, o), gld(J.row_inv_f0, o), __int_as_float(gld(J.nfold, o)));
  };
  float2 acc[SH::ITL][SH::RL];
#pragma unroll
  for (int s = 0; s < SH::ITL; ++s)
#pragma unroll
    for (int q = 0; q < SH::RL; ++q) acc[s][q] = make_float2(0.0f, 0.0f);
  // pair p's block: rows k1_0 .. k1_0+B-1 of Y_p, element e = n2*B + b
  const float2* Yb = J.ybuf + (long long)k1_0 * N2;
  float2 pre[SH::IT0][SH::R0];
  auto prefetch = [&](int p) {
    const float2* Yp = Yb + (long long)p * L;
#pragma unroll
    for (int s = 0; s < SH::IT0; ++s) {
      const int t = threadIdx.x + s * T;
      if (SH::ITEMS0 % T == 0 || t < SH::ITEMS0) {
#pragma unroll
        for (int q = 0; q < SH::R0; ++q) {
          const int e = t + q * NR0 * B;
          pre[s][q] = gld(Yp, 8u * (unsigned)((e & (B - 1)) * N2 + (e >> LB)));
        }
      }
    }
  };
  // frequencies of this thread's bins (forward last-stage element (s, q):
  // k2 = j + q*NS, j = (tid + s*T) >> LB), folded to [0, L/2]
  float xs[SH::ITL][SH::RL];
#pragma unroll
  for (int s = 0; s < SH::ITL; ++s) {
    const int t = threadIdx.x + s * T;
#pragma unroll
    for (int q = 0; q < SH::RL; ++q) {
      const long long k = (long long)(k1_0 + (t & (B - 1))) + (long long)N1 * ((t >> LB) + q * NSL);
      const long long kk = (k <= L - k) ? k : (L - k);
      xs[s][q] = (float)kk * J.inv_Lws;
    }
  }
  // Lineshape FTs of this thread's bins.  F_r(x) = exp(-a_l x^2) exp(-b_m x)
  // for row r = l*N_L + m (plus alias terms where the row folds).  The
  // Gaussian factor depends on l only and is kept per bin while consecutive
  // rows share l (refilled N_G times).  The Lorentzian factor is geometric
  // along a thread's bins: for fixed s the RL bins are kk = kk0 + i*L/RL in
  // two runs of RL/2 (below and above L/2, folded), so exp(-b x) takes one
  // exponential per run and a product per bin.  That is about 3 exponentials
  // per row and thread instead of RL*ITL.
  const int NLr = axes[it.x].nl;
  const float dlt = (float)(L / SH::RL) * J.inv_Lws;
  float u0[SH::ITL][2];
#pragma unroll
  for (int s = 0; s < SH::ITL; ++s) {
    const int t = threadIdx.x + s * T;
    const long long ks = (long long)(k1_0 + (t & (B - 1))) + (long long)N1 * (t >> LB);  // bin of q = 0
    u0[s][0] = (float)ks * J.inv_Lws;                    // q = 0, 1, ..., RL/2-1
    u0[s][1] = (float)(L / SH::RL - ks) * J.inv_Lws;     // q = RL-1, RL-2, ..., RL/2
  }
  // bins below the circular-image cutoff (kk < wrap_k <= N1, wrap_item): with
  // k = k1 + N1 k2 only k2 = 0 (kk = k1) and k2 = N2 - 1 (kk = N1 - k1)
  // qualify.  The first is slot (0, 0) of thread b < B (t >> LB == 0); the
  // second is slot (s, RL - 1) of the thread with t >> LB == NSL - 1.  The
  // workgroup's table values (both bins of its B rows, every L-axis node)
  // are staged in LDS here, so the pair loop reads no global memory for
  // them (a global load there would wait behind the next pair's prefetch).
  // w_lo / w_hi hold the row b of those bins (-1: none), w_s the second's s.
  float* const s_wh = reinterpret_cast<float*>(s_row + n_rows_lds);  // [2][B][n_wh]
  auto wh = [&](int c, int b, int m) -> float& { return s_wh[(c * B + b) * n_wh + m]; };
  const int WK = J.wrap_k;
  int w_lo = -1, w_hi = -1, w_s = 0;
  if (WK) {
    const int NLw = axes[it.x].nl;
    for (int i = threadIdx.x; i < 2 * B * NLw; i += T) {
      const int c = i / (B * NLw), b = (i / NLw) % B, m = i % NLw;
      const int k1 = k1_0 + b, kk = c ? N1 - k1 : k1;
      wh(c, b, m) = kk < WK ? J.wrap_h[m * WK + kk] : 0.0f;
    }
    if ((int)threadIdx.x < B && k1_0 + (int)threadIdx.x < WK) w_lo = threadIdx.x;
#pragma unroll
    for (int s = 0; s < SH::ITL; ++s) {
      const int t = threadIdx.x + s * T;
      const int b = t & (B - 1);
      if ((t >> LB) == NSL - 1 && N1 - (k1_0 + b) < WK) { w_hi = b; w_s = s; }
    }
    __syncthreads();  // WK is the job's: uniform over the workgroup
  }
  float gcache[SH::ITL][SH::RL];
  int g_l = -1;
  auto gauss_fill = [&](int r) {
    const float a = row_c(r).x;
#pragma unroll
    for (int s = 0; s < SH::ITL; ++s)
#pragma unroll
      for (int q = 0; q < SH::RL; ++q) gcache[s][q] = __expf(-(a * xs[s][q] * xs[s][q]));
  };
  auto row_ft = [&](int r, float (&F)[SH::ITL][SH::RL]) {
    const float4 c = row_c(r);
    const float bco = c.y;
    const float E = __expf(-bco * dlt);
#pragma unroll
    for (int s = 0; s < SH::ITL; ++s)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float H = __expf(-bco * u0[s][h]);
#pragma unroll
        for (int i = 0; i < SH::RL / 2; ++i) {
          const int q = h ? SH::RL - 1 - i : i;
          F[s][q] = gcache[s][q] * H;
          H *= E;
        }
      }
    const int nf = __float_as_int(c.w);
    if (nf > 0) {
      const float a = c.x;
#pragma unroll
      for (int s = 0; s < SH::ITL; ++s)
#pragma unroll
        for (int q = 0; q < SH::RL; ++q)
          for (int n = 1; n <= nf; ++n) {
            // the folded frequency's mirror L/2 - kk (alias terms are rare: recomputed)
            const float xr = (float)(L / 2) * J.inv_Lws - xs[s][q];
            const float xa = (n & 1) ? ((float)n * J.half_inv_ws + xr) : ((float)n * J.half_inv_ws + xs[s][q]);
            F[s][q] += __expf(-(a * xa * xa) - bco * xa);
          }
    }
    if (w_lo >= 0 || w_hi >= 0) {
      const int m = r % NLr;
      if (w_lo >= 0) F[0][0] += wh(0, w_lo, m);
      if (w_hi >= 0) {
        const float hv = wh(1, w_hi, m);
#pragma unroll
        for (int s = 0; s < SH::ITL; ++s)
          if (s == w_s) F[s][SH::RL - 1] += hv;
      }
    }
    const float f0 = c.z;
#pragma unroll
    for (int s = 0; s < SH::ITL; ++s)
#pragma unroll
      for (int q = 0; q < SH::RL; ++q) F[s][q] *= f0;
  };
  if (p1 > p0) prefetch(p0);
  for (int p = p0; p < p1; ++p) {
    const int ra = 2 * p, rb = 2 * p + 1;
    // the pair's lineshape FTs, evaluated just before the last stage that
    // uses them (they are then not live across the earlier stages)
    float fa[SH::ITL][SH::RL], fb[SH::ITL][SH::RL];
    auto lineshapes = [&] {
      if (ra / NLr != g_l) { g_l = ra / NLr; gauss_fill(ra); }
      row_ft(ra, fa);
      if (rb < M) {
        if (rb / NLr != g_l) { g_l = rb / NLr; gauss_fill(rb); }
        row_ft(rb, fb);
      } else {
#pragma unroll
        for (int s = 0; s < SH::ITL; ++s)
#pragma unroll
          for (int q = 0; q < SH::RL; ++q) fb[s][q] = 0.0f;
      }
    };
    auto ld = [&](int s, int q, int) -> float2 { return pre[s][q]; };
    auto st = [&](int s, int q, int, float2 z) {
      acc[s][q].x += z.x * fa[s][q] + z.y * fb[s][q];
      acc[s][q].y += z.y * fa[s][q] - z.x * fb[s][q];
    };
    auto next = [&] {
      if (p + 1 < p1) prefetch(p + 1);
    };
    float2* const x0 = bufs[SWAP ? (p & 1) : 0];
    float2* const x1 = bufs[SWAP ? 1 - (p & 1) : 1];
    fc_fft_pp<N2, B, T, true, true, true, false, 4, true, true, PV>(x0, x1, s_tw, ld, st, next, lineshapes);
  }
  __syncthreads();
  // the accumulated spectrum into LDS (bufs[1]), in the forward last stage's element order
#pragma unroll
  for (int s = 0; s < SH::ITL; ++s) {
    const int t = threadIdx.x + s * T;
    if (SH::ITEMSL % T == 0 || t < SH::ITEMSL) {
      const int j = t >> LB, b = t & (B - 1);
      const int k = j % NSL;
      const int e0 = (((j - k) * SH::RL + k) << LB) + b;
#pragma unroll
      for (int q = 0; q < SH::RL; ++q) bufs[1][pad_e<4>(e0 + q * NSL * B)] = acc[s][q];
    }
  }
  __syncthreads();
  float2* Bo = (it.z ? J.bbuf2 : J.bbuf) + (long long)k1_0 * N2;
  // output n2 of this thread's last-stage element (s, q): j0 + s*(T/B) + q*NS, row k1 fixed
  TwProg<SH::ITL, SH::RL> tw;
  tw.init(J, k1_0 + (threadIdx.x & (B - 1)), threadIdx.x >> LB, T >> LB, NSL);
  auto st_out = [&](int s, int q, int e, float2 z) {
    const int b = e & (B - 1), n2 = e >> LB;
    gst(Bo, 8u * (unsigned)(b * N2 + n2), cmulc(z, tw(s, q)));
  };
  FcNone none;
  auto nop = [] {};
  // first stage reads bufs[1] (ping-pong)
  fc_fft_pp<N2, B, T, false, false, true, false, 4, true, true, PV>(bufs[0], bufs[1], s_tw, none, st_out, nop, nop);
}

// ---------------------------------------------------------------------------
// Pass C with N1 fixed (see k_fft_c): first stage straight from B, last
// stage straight to the output.  SUM: B is J.bbuf + J.bbuf2 (pass B split
// its pairs over two workgroups per row block).
// ---------------------------------------------------------------------------
template <int N1, bool SUM>
__global__ __launch_bounds__(FC_T, 4) void k_fft_c_c(const LdmJob* __restrict__ jobs, const int4* __restrict__ items,
                                                     int n_items) {
  constexpr int B = fc_cols_c(N1), LB = fc_log2(B), CAPE = N1 * B;
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
      if (n < nv) gst(J.out, 4u * (unsigned)n, 0.0f);
    }
    return;
  }
  fc_load_tw<N1, FC_T>(s_tw, J.tw1);
  const float sc = J.out_scale;
  auto nop = [] {};
  auto off = [&](int e) { return 8u * (unsigned)((e >> LB) * N2 + n2_0 + (e & (B - 1))); };  // < 8 L <= 2^24
  auto ld = [&](int, int, int e) -> float2 {
    float2 z = gld(J.bbuf, off(e));
    if constexpr (SUM) {
      const float2 w = gld(J.bbuf2, off(e));
      z.x += w.x;
      z.y += w.y;
    }
    return z;
  };
  if (J.derived) {  // uniform per workgroup: the plain store path stays as lean as without the epilogue
    auto st = [&](int, int, int e, float2 z) {
      const int n = n2_0 + (e & (B - 1)) + N2 * (e >> LB);
      if (n < nv) store_derived(J, n, z.x * sc);
    };
    fc_fft_io<N1, B, FC_T, false, true, true>(buf, s_tw, ld, st, nop);
  } else {
    auto st = [&](int, int, int e, float2 z) {
      const int n = n2_0 + (e & (B - 1)) + N2 * (e >> LB);
      if (n < nv) gst(J.out, 4u * (unsigned)n, z.x * sc);
    };
    fc_fft_io<N1, B, FC_T, false, true, true>(buf, s_tw, ld, st, nop);
  }
}
