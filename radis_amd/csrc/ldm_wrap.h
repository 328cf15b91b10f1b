// wrap_item: the circular-image correction that makes a transform of length
// L' > 2 N_v return RADIS's spectrum at L = 2 N_v (radis/lbl/broadening.py:
// 1900-1902 rfftfreq(2*len(w)), 2340-2346 the LDM of 2*len(w) rows).
//
// With the folded FT templates, the real-space kernel of an LDM row at
// transform length L is the L-periodised, grid-sampled Voigt profile.  An
// output j < N_v takes LDM column k' <= N_v through that kernel at offset
// d = j - k' in [-N_v, N_v]; periods L and L' differ only in which periodic
// images fall there, all at |d + nL| >= N_v grid steps from the centre,
// where the Gaussian part is gone (relative weight (sigma_G / N_v)^2).  So
//   g_L(d) - g_L'(d) = P_L(d) - P_L'(d) = c(d)   for |d| <= N_v,
// with P_L the L-periodised sampled Lorentzian of the row's L-axis node,
//   P_L(d) = (1/L) sinh(a) / (cosh(a) - cos(2 pi d / L)),  a = 2 pi gamma / L
// (gamma = HWHM in grid steps; the Fourier series of the periodised
// Lorentzian).  The L' - 2 N_v offsets past N_v are never read; there c is
// taken smoothly to zero, c_ext(d) = phi(d) c(d) + phi(L' - d) c(L' - d),
// phi = 1 on [0, N_v], an erfc step over the next w = min(L' - 2 N_v, N_v/2)
// offsets (c is smooth there: its nearest singular peak is at d = 2 N_v),
// sigma = w / LDM_WRAP_NSIG.  c_ext is even, so its DFT H_m is real, and it
// is negligible beyond K = wrap_k bins: pass B adds H_m[kk] to the
// lineshape FT of every row on node m at its bins kk < K.  H_m comes from
// LC >= 8 K samples of c_ext (512, 1024 or LDM_WRAP_LC; the sampled sum
// equals the full one up to aliases of H_m beyond LC - K, far below 1e-9 of
// it).
//
// One workgroup per (request, pair of L-axis nodes): the two real
// sequences are the real and imaginary parts of one complex transform (the
// stage twiddles come from a table computed with the batch's other
// twiddles).  The items run as the first workgroups of the fill launch
// (k_fill below), 256 threads each, concurrently with the fill tiles: pass
// B, two launches later, is the first reader of the table.
#pragma once
#include "ldm_fftc.h"
#include "ldm_kernels.h"

struct WrapNode {
  double omE, omE2, E;  // 1 - e^-a, 1 - e^-2a, e^-a for the period's a = 2 pi gamma / period
};

__device__ __forceinline__ WrapNode wrap_node(double gamma, double period) {
  const double a = 6.283185307179586 * gamma / period;
  WrapNode w;
  w.omE = -expm1(-a);
  w.omE2 = -expm1(-2.0 * a);
  w.E = exp(-a);
  return w;
}

// P(d) of one period in the overflow-free form (1/L)(1 - E^2) / ((1 - E)^2 +
// 4 E sin^2(pi d / L)) (sinh and cosh multiplied by 2 e^-a); s2 = sin^2(pi d / L)
__device__ __forceinline__ double lor_periodic(const WrapNode& w, double s2, double inv_period) {
  return inv_period * w.omE2 / (w.omE * w.omE + 4.0 * w.E * s2);
}

// LDS of one item: the transform buffer and its twiddle table (float2s)
constexpr int LDM_WRAP_LDS = LdsSize<LDM_WRAP_LC>::value + LDM_WRAP_LC;

// The table of one node pair from LC samples (LC >= 8 K: 512, 1024 or 2048)
// by T threads.
template <int LC, int T>
__device__ __forceinline__ void wrap_pair(const LdmJob& J, const LdmAxes& A, int ma, float2* __restrict__ buf,
                                          float2* __restrict__ s_tw) {
  constexpr int STRIDE = LDM_WRAP_LC / LC;
  const int K = J.wrap_k;
  const int mb = ma + 1;
  const bool has_b = mb < A.nl;
  // W_LC^e = W_LDM_WRAP_LC^(e * STRIDE)
  for (int e = threadIdx.x; e < LC; e += T) s_tw[e] = gld(J.tw_wrap, 8u * (unsigned)(e * STRIDE));
  const double Lp = (double)J.L, Ls = (double)J.wrap_ls;
  const double nv = 0.5 * Ls;
  const double w = fmin(Lp - Ls, 0.5 * nv);
  const double d_mid = nv + 0.5 * w;                                    // centre of the step
  const double isig = LDM_WRAP_NSIG / (w * 1.4142135623730951);         // 1 / (sigma sqrt 2)
  const double z_cut = 7.0 / 1.4142135623730951;                         // phi < 1e-12 beyond
  const double inv_s = 1.0 / Ls, inv_p = 1.0 / Lp;
  // HWHM of the two nodes in grid steps: FWHM / (2 wstep)
  const double h = 0.5 / J.wstep;
  const double ga = exp(A.laL[ma]) * h, gb = has_b ? exp(A.laL[mb]) * h : ga;
  const WrapNode sa = wrap_node(ga, Ls), pa = wrap_node(ga, Lp);
  const WrapNode sb = wrap_node(gb, Ls), pb = wrap_node(gb, Lp);
  // phi(x) c(x) for both nodes (x in [0, L')): the two nodes share x's
  // sines; phi needs no more than FP32 (it is 1 - 3e-7 at x = N_v)
  auto side = [&](double x, double& va, double& vb) {
    const double z = (x - d_mid) * isig;
    if (z >= z_cut) return;
    const double phi = 0.5 * (double)erfcf((float)z);
    const double ss = sinpi(x * inv_s), sp = sinpi(x * inv_p);
    const double s2s = ss * ss, s2p = sp * sp;
    va += phi * (lor_periodic(sa, s2s, inv_s) - lor_periodic(pa, s2p, inv_p));
    if (has_b) vb += phi * (lor_periodic(sb, s2s, inv_s) - lor_periodic(pb, s2p, inv_p));
  };
  const double step = Lp / (double)LC;
  for (int j = threadIdx.x; j < LC; j += T) {
    const double d = (double)j * step;
    double ca = 0.0, cb = 0.0;
    side(d, ca, cb);
    side(Lp - d, ca, cb);
    buf[pad_idx(j)] = make_float2((float)ca, (float)cb);
  }
  __syncthreads();
  fc_fft<LC, 1, T, true>(buf, s_tw);
  // sum over the L' integer offsets = (L' / LC) * the sampled sum
  const float sc = (float)step;
  float* H = J.wrap_h;
  for (int k = threadIdx.x; k < K; k += T) {
    const float2 z = buf[pad_idx(k)];
    H[ma * K + k] = z.x * sc;
    if (has_b) H[mb * K + k] = z.y * sc;
  }
}

// One image-correction item (request it.x, nodes 2 it.y and 2 it.y + 1)
// with T threads; lds holds LDM_WRAP_LDS float2s.
template <int T>
__device__ __forceinline__ void wrap_item(const LdmJob* __restrict__ jobs, const LdmAxes* __restrict__ axes,
                                          const int4 it, float2* lds) {
  const LdmJob& J = jobs[it.x];
  const LdmAxes& A = axes[it.x];
  const int K = J.wrap_k;
  const int ma = 2 * it.y;
  if (J.status || K == 0 || ma >= A.nl) return;
  float2* buf = lds;
  float2* s_tw = lds + LdsSize<LDM_WRAP_LC>::value;
  // fewest samples with LC >= 8 K (uniform per workgroup)
  if (8 * K <= 512) wrap_pair<512, T>(J, A, ma, buf, s_tw);
  else if (8 * K <= 1024) wrap_pair<1024, T>(J, A, ma, buf, s_tw);
  else wrap_pair<LDM_WRAP_LC, T>(J, A, ma, buf, s_tw);
}

// The fill launch: workgroups below n_wrap take the image-correction items
// (items[0, n_wrap)), the others the fill tiles that follow them in the
// item list.  One kernel per line-list variant (PLAIN: no isotopologue,
// shift or n_self column; the host picks), so each gets its own register
// allocation: the plain kernel needs the fewest; GROUPED (fill_tile) only
// when some request's LDM is taller than a tile.  The dynamic LDS holds the
// larger of a fill tile (or row group) and an image-correction item.
template <bool PLAIN, bool GROUPED>
__global__ __launch_bounds__(LDM_FILL_T, 4) void k_fill(const LdmJob* __restrict__ jobs,
                                                     const LdmAxes* __restrict__ axes,
                                                     const int4* __restrict__ items, int n_wrap, LdmDb db) {
  extern __shared__ __align__(16) double acc[];
  const int4 it = items[blockIdx.x];
  if ((int)blockIdx.x < n_wrap) {
    wrap_item<LDM_FILL_T>(jobs, axes, it, reinterpret_cast<float2*>(acc));
    return;
  }
  const LdmJob& J = jobs[it.x];
  const LdmAxes& A = axes[it.x];
  if (A.m == 0 || J.status) return;
  if (J.noneq) fill_tile<true, PLAIN, GROUPED>(J, A, it, db, acc);
  else fill_tile<false, PLAIN, GROUPED>(J, A, it, db, acc);
}
