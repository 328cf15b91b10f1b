// HIP kernels of the DIT/LDM pipeline (gfx950).
//
//   k_axis        per-request min/max of the line widths (LDM axis limits)
//   k_finalize    LDM axes, N_G, N_L, per-row alias-fold counts and norms
//   k_tile_bounds line range of every fill tile (binary search on nu0)
//   k_fill        line parameters + bit-exact np.interp indices + 2x2x2
//                 weights, scattered into an LDS tile [M][W] and flushed once
//   k_fft_a       four-step pass 1: N1-point DFTs over strided columns of two
//                 LDM rows packed as one complex row, times W_L^(n2 k1)
//   k_fft_b       four-step pass 2 fused with the lineshape multiply and the
//                 (l, m) reduction, then the first inverse pass
//   k_fft_c       last inverse pass, real part, 1/(L wstep) scale and crop
//
// Semantics follow RADIS's CPU LDM-fft path; file:line citations are to the
// radis 0.17.1 sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ldm_fft.h"
#include "ldm_job.h"

#define LDM_C_CGS 2.99792458e10       /* radis/phys/constants.py:78 */

// ---------------------------------------------------------------------------
// FP64 arithmetic that must round exactly like NumPy: no FMA contraction.
// ---------------------------------------------------------------------------

// Position of x on the grid w[i] = wmin + i*delta (numpy.arange fill rule),
// following numpy.interp(x, w, arange(nv)) bit for bit:
//   x < w[0] -> 0; x >= w[nv-1] -> nv-1; w[j] == x -> j;
//   else (1/(w[j+1]-w[j]))*(x-w[j]) + j with w[j] <= x < w[j+1].
// radis/lbl/broadening.py:2151-2154, 2253.
__device__ __forceinline__ double grid_position(double x, double wmin, double delta, double inv_delta, int nv) {
#pragma clang fp contract(off)
  // grids are shorter than 2^31 points (the FFT length caps them far below),
  // so 32-bit indices convert to and from FP64 in one instruction
  const double wlast = wmin + (double)(nv - 1) * delta;
  if (x < wmin) return 0.0;
  if (!(x < wlast)) return (double)(nv - 1);
  // First guess: (x - wmin)/delta has a relative error of a few ulp, i.e.
  // ~1e-9 grid steps for grids below 1e7 points, so the largest j with
  // w[j] <= x is the guess or one of its neighbours; one exact comparison
  // each way settles it.
  int j = (int)((x - wmin) * inv_delta);  // x > wmin: truncation is floor
  j = j > nv - 2 ? nv - 2 : j;
  double wj = wmin + (double)j * delta;
  if (wj > x && j > 0) {
    --j;
    wj = wmin + (double)j * delta;
  }
  double wj1 = wmin + (double)(j + 1) * delta;
  if (wj1 <= x && j < nv - 2) {
    ++j;
    wj = wj1;
    wj1 = wmin + (double)(j + 1) * delta;
  }
  if (wj == x) return (double)j;
  const double slope = 1.0 / (wj1 - wj);
  return slope * (x - wj) + (double)j;
}

// The same as (index, fraction) = (int(pos), pos - int(pos)) for the fill.
// NumPy's (1/(w[j+1]-w[j]))*(x-w[j]) + j reaches j+1 only when (x-w[j]) is
// within a few ulp of j's spacing below the step (the sum's half-ulp is
// below 2e-9 for j < 2^24), so away from that the index is j and the
// fraction, which only feeds FP32 weights, takes an FP32 reciprocal instead
// of the FP64 division.
// Common case: 0 <= x - w[j] < fast_hi (one step less a margin above the
// rounding of w[j+1] and 1e-7 of a step), so the first guess j is x's
// interval and the position rounds to j; the fraction then takes the FP32
// reciprocal of the step (w[j+1] - w[j] differs from it by rounding only).
__device__ __forceinline__ void grid_index(double x, double wmin, double delta, double inv_delta, int nv,
                                           double fast_hi, float inv_delta_f, int& k0, float& tv) {
#pragma clang fp contract(off)
  const double wlast = wmin + (double)(nv - 1) * delta;
  if (x < wmin) { k0 = 0; tv = 0.0f; return; }
  if (!(x < wlast)) { k0 = nv - 1; tv = 0.0f; return; }
  int j = (int)((x - wmin) * inv_delta);
  j = j > nv - 2 ? nv - 2 : j;
  double wj = wmin + (double)j * delta;
  const double d0 = x - wj;
  if (d0 >= 0.0 && d0 < fast_hi) {
    k0 = j;
    tv = (float)d0 * inv_delta_f;
    return;
  }
  if (wj > x && j > 0) {
    --j;
    wj = wmin + (double)j * delta;
  }
  double wj1 = wmin + (double)(j + 1) * delta;
  if (wj1 <= x && j < nv - 2) {
    ++j;
    wj = wj1;
    wj1 = wmin + (double)(j + 1) * delta;
  }
  const double d = x - wj, dl = wj1 - wj;
  if (d < dl * (1.0 - 1e-7)) {
    k0 = j;
    tv = (float)d * __builtin_amdgcn_rcpf((float)dl);
    return;
  }
  const double pos = (1.0 / dl) * d + (double)j;
  k0 = (int)pos;
  tv = (float)(pos - (double)k0);
}

// Same rule on a small tabulated increasing axis la[0..n) with the slopes
// sl[j] = 1/(la[j+1]-la[j]) precomputed (NumPy computes the same quotient).
__device__ __forceinline__ double axis_position(double x, const double* la, const double* sl, int n) {
#pragma clang fp contract(off)
  if (x < la[0]) return 0.0;
  if (!(x < la[n - 1])) return (double)(n - 1);
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (la[mid] <= x) lo = mid; else hi = mid;
  }
  if (la[lo] == x) return (double)lo;
  return sl[lo] * (x - la[lo]) + (double)lo;
}

// Gaussian FWHM, radis/lbl/broadening.py:170-172 (x2 at 1855):
// ((nu0 / c_CGS) * sqrt((2 Na k_b T ln2)/M)) * 2
__device__ __forceinline__ double fwhm_gauss(double nu0, double dopp) {
#pragma clang fp contract(off)
  return ((nu0 / LDM_C_CGS) * dopp) * 2.0;
}

// Lorentzian FWHM, radis/lbl/broadening.py:341-383 (x2 at 1854), with
// diluent {'air': 1-x}; (Tref/T)**n as pow, as NumPy evaluates it (the
// width extremes the LDM axes start from).  A called function: k_axis and
// k_finalize evaluate it at several sites, and one copy of its two FP64 pow
// routines keeps their code short (short kernels that fetch their code cold
// once per spectrum).
__device__ __noinline__ double fwhm_lorentz(const LdmJob& J, double g_air, double g_self, double n_air,
                                               double n_self) {
#pragma clang fp contract(off)
  const double ra = pow(J.ratio_T, n_air);
  const double rs = (n_self == n_air) ? ra : pow(J.ratio_T, n_self);
  const double air = ((ra * g_air) * J.p_atm) * J.x_air;
  const double slf = rs * ((g_self * J.p_atm) * J.x_self);
  return (air + slf) * 2.0;
}

// FP32 Lorentzian FWHM (relative error ~1e-6), for filtering and fast paths.
__device__ __forceinline__ float fwhm_lorentz_f32(const LdmJob& J, float ga, float gs, float na, float ns) {
  const float ra = __expf(na * J.log_r_f);
  const float rs = (ns == na) ? ra : __expf(ns * J.log_r_f);
  return 2.0f * (ra * ga * J.pa_air_f + rs * gs * J.pa_self_f);
}

__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long double_to_bits(double d) {
  return (unsigned long long)__double_as_longlong(d);
}

// ---------------------------------------------------------------------------
// k_axis: min/max of the Lorentzian (and, for several isotopologues, the
// Gaussian) FWHM over each request's window; radis/lbl/broadening.py:1823-1852.
// Positive doubles order like their bit patterns, so the extremes are reduced
// as unsigned 64-bit integers: one partial per item, combined by k_finalize.
// Item = (job, partial index, z, w): w > 0 scans the w lines from J.i0 + z;
// w < 0 takes the width candidates of the -w whole blocks from block z.
//
// The Lorentzian FWHM 2 p [(Tref/T)^n_air g_air (1-x) + (Tref/T)^n_self g_self x]
// is non-decreasing in g_air and g_self and monotone in n_air and n_self, in
// the direction of log(Tref/T).  Its maximum (minimum) over a block is
// therefore reached on the block's Pareto-maximal lines for (g up, n in that
// direction) ((g down, n against it)), and FP64 evaluation keeps that order
// (rounding is monotone, and different float inputs move exp's argument by
// far more than an ulp), so the extremes over the candidates are bit-identical
// to those over every line.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_axis(LdmJob* __restrict__ jobs, const int4* __restrict__ items,
                                              LdmDb db) {
  const int4 it = items[blockIdx.x];
  LdmJob& J = jobs[it.x];
  unsigned long long lmin = ~0ull, lmax = 0ull, gmin = ~0ull, gmax = 0ull;
  if (it.w < 0) {
    const int gb0 = it.z, G = -it.w;
    const int dmax = J.log_r > 0.0 ? 2 : 0, dmin = dmax + 1;
    const int stride = db.sky_nb + 1;
    const int* offa = db.sky_off + dmax * stride;
    const int* offb = db.sky_off + dmin * stride;
    const int a0 = offa[gb0], na = offa[gb0 + G] - a0;
    const int b0 = offb[gb0], nb = offb[gb0 + G] - b0;
    // Gaussian FWHM = nu0 * const(iso): per isotopologue the block's first
    // and last line (nu0 sorted)
    const int ns = db.n_iso_slots;
    const int nc = J.single_iso ? 0 : G * ns;
    // one pass over the two candidate lists and the isotopologue ranges, so
    // that their loads are in flight together
    const int nmax = max(max(na, nb), nc);
    for (int c = threadIdx.x; c < nmax; c += blockDim.x) {
      float4 qa = make_float4(0.f, 0.f, 0.f, 0.f), qb = qa;
      int2 r = make_int2(-1, -1);
      if (c < na) qa = db.sky[a0 + c];
      if (c < nb) qb = db.sky[b0 + c];
      if (c < nc) r = db.iso_rng[(long long)(gb0 + c / ns) * ns + (c - (c / ns) * ns)];
      if (c < na) {
        const unsigned long long bl = double_to_bits(fwhm_lorentz(J, qa.x, qa.y, qa.z, qa.w));
        lmax = bl > lmax ? bl : lmax;
      }
      if (c < nb) {
        const unsigned long long bl = double_to_bits(fwhm_lorentz(J, qb.x, qb.y, qb.z, qb.w));
        lmin = bl < lmin ? bl : lmin;
      }
      if (r.x >= 0) {
        const int s = c - (c / ns) * ns;
        const unsigned long long g0 = double_to_bits(fwhm_gauss(db.nu0[r.x], J.dopp[s]));
        const unsigned long long g1 = double_to_bits(fwhm_gauss(db.nu0[r.y], J.dopp[s]));
        gmin = g0 < gmin ? g0 : gmin;
        gmax = g1 > gmax ? g1 : gmax;
      }
    }
  } else {
    const long long a = J.i0 + it.z;
    const long long e = a + it.w;
    // FP32 widths of this lane's lines stay in registers; only the FP32 argmin
    // and argmax are re-evaluated in FP64.  If another of the lane's lines is
    // within the FP32 error bound of an extreme (a near tie), every line of the
    // lane is evaluated in FP64 (rare), so the exact FP64 extremes are found.
    constexpr int NPER = LDM_AX_CHUNK / 256;
    float wf[NPER];
    float fmin = 3.0e38f, fmax = -1.0f;
    int kmin = -1, kmax = -1;
#pragma unroll
    for (int k = 0; k < NPER; ++k) {
      const long long i = a + threadIdx.x + (long long)k * 256;
      wf[k] = -1.0f;
      if (i < e) {
        const float na = db.n_air[i];
        wf[k] = fwhm_lorentz_f32(J, db.g_air[i], db.g_self[i], na, db.n_self ? db.n_self[i] : na);
        if (wf[k] < fmin) { fmin = wf[k]; kmin = k; }
        if (wf[k] > fmax) { fmax = wf[k]; kmax = k; }
      }
    }
    bool tie = false;
#pragma unroll
    for (int k = 0; k < NPER; ++k) {
      if (wf[k] >= 0.0f && k != kmin && wf[k] <= fmin * (1.0f + 2e-5f)) tie = true;
      if (wf[k] >= 0.0f && k != kmax && wf[k] >= fmax * (1.0f - 2e-5f)) tie = true;
    }
    auto exact = [&](int k) {
      const long long i = a + threadIdx.x + (long long)k * 256;
      const double na = db.n_air[i];
      return double_to_bits(fwhm_lorentz(J, db.g_air[i], db.g_self[i], na, db.n_self ? (double)db.n_self[i] : na));
    };
    if (kmin >= 0) {
      if (!tie) {
        lmin = exact(kmin);
        lmax = exact(kmax);
      } else {
        for (int k = 0; k < NPER; ++k)
          if (wf[k] >= 0.0f) {
            const unsigned long long bl = exact(k);
            lmin = bl < lmin ? bl : lmin;
            lmax = bl > lmax ? bl : lmax;
          }
      }
    }
    if (!J.single_iso) {
      for (long long i = a + threadIdx.x; i < e; i += blockDim.x) {
        const int iso = db.iso ? db.iso[i] : 1;
        const unsigned long long bg = double_to_bits(fwhm_gauss(db.nu0[i], J.dopp[iso]));
        gmin = bg < gmin ? bg : gmin;
        gmax = bg > gmax ? bg : gmax;
      }
    }
  }
  // wave reduction
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o;
    o = __shfl_xor(lmin, off); lmin = o < lmin ? o : lmin;
    o = __shfl_xor(lmax, off); lmax = o > lmax ? o : lmax;
    o = __shfl_xor(gmin, off); gmin = o < gmin ? o : gmin;
    o = __shfl_xor(gmax, off); gmax = o > gmax ? o : gmax;
  }
  __shared__ unsigned long long red[4][4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][w] = lmin; red[1][w] = lmax; red[2][w] = gmin; red[3][w] = gmax; }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int k = 1; k < nw; ++k) {
      lmin = red[0][k] < lmin ? red[0][k] : lmin;
      lmax = red[1][k] > lmax ? red[1][k] : lmax;
      gmin = red[2][k] < gmin ? red[2][k] : gmin;
      gmax = red[3][k] > gmax ? red[3][k] : gmax;
    }
    // per-chunk partials, reduced by k_finalize (same-address global atomics
    // from every workgroup would serialise in L2)
    unsigned long long* P = J.ax_part + 4 * (long long)it.y;
    P[0] = lmin;
    P[1] = lmax;
    P[2] = gmin;
    P[3] = gmax;
  }
}


// Q(T) of one slot at n temperatures (radis_ldm_qspline_eval)
__global__ __launch_bounds__(256) void k_qspline_eval(const LdmQSpline* __restrict__ qs, int iso,
                                                      const double* __restrict__ T, long long n,
                                                      double* __restrict__ Q) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    Q[i] = qspline_eval(qs[iso], T[i]);
}

// Voigt FT at x for FWHM (wG, wL): exp(-(pi x wG)^2/(4 ln2)) * exp(-pi x wL)
// (radis/lbl/broadening.py:252, 604, 850-853 with hwhm = fwhm/2); a called
// function, one copy for k_finalize's two sites.
__device__ __noinline__ double voigt_ft64(double x, double wG, double wL) {
  const double PI = 3.141592653589793;
  const double g = PI * x * wG;
  return exp(-(g * g) / (4.0 * 0.6931471805599453)) * exp(-PI * x * wL);
}

// FP64 exp and log as called functions in k_finalize (one copy each of
// routines it runs at several sites; same routines, same results)
__device__ __noinline__ double exp_call(double x) { return exp(x); }
__device__ __noinline__ double log_call(double x) { return log(x); }

// ---------------------------------------------------------------------------
// k_finalize: one workgroup per request.  Builds the two log-spaced axes
// exactly as _init_w_axis (radis/lbl/broadening.py:1823-1837), the per-row
// alias-fold counts and normalisations of the FT templates (1895-1926).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_finalize(LdmJob* __restrict__ jobs, LdmAxes* __restrict__ axes,
                                                 LdmResult* __restrict__ res, LdmDb db) {
  LdmJob& J = jobs[blockIdx.x];
  LdmAxes& A = axes[blockIdx.x];
  __shared__ double wgax[LDM_MAX_AX], wlax[LDM_MAX_AX], lgax[LDM_MAX_AX], llax[LDM_MAX_AX];
  __shared__ double s_wg_min, s_wl_min, s_wg_max, s_wl_max;
  __shared__ int s_ng, s_nl, s_status;
  __shared__ unsigned long long s_w[4][4];
  // This kernel is a chain of dependent memory round trips, so everything it
  // reads is loaded up front: the job's values, the partial extremes and,
  // for a single isotopologue, the window's end lines (its Gaussian width
  // extremes).  The axis nodes and logs stay in LDS, and the status is
  // written once at the end (it is this run's own: the job table is not
  // re-copied when the host has nothing new for it).
  const long long nch = J.n_ax_part;
  const unsigned long long* part = J.ax_part;
  const long long i0 = J.i0, i1 = J.i1;
  const bool single = J.single_iso;
  const double dopp1 = J.dopp[1], dxG = J.dxG, dxL = J.dxL, wstep = J.wstep, fold_thresh = J.fold_thresh;
  const long long L = J.L;
  const int m_cap = J.m_cap;
  int* const nfold = J.nfold;
  float* const row_a = J.row_a;
  float* const row_b = J.row_b;
  float* const row_inv_f0 = J.row_inv_f0;
  const bool ends = threadIdx.x == 0 && single && i1 > i0;
  const double nu_a = ends ? db.nu0[i0] : 0.0, nu_b = ends ? db.nu0[i1 - 1] : 0.0;
  // Q(Tref)/Q(T) of the slots with a partition-function spline (the fill
  // reads these after this kernel): calc_linestrength_eq's Qref/Q,
  // radis/lbl/base.py:2605-2697
  if (J.device_q && threadIdx.x < LDM_MAX_ISO && ((db.qs_mask >> threadIdx.x) & 1u)) {
    const int s = threadIdx.x;
    const double q = qspline_eval(db.qs[s], J.Tref) / qspline_eval(db.qs[s], J.Tgas);
    J.qratio[s] = q;
    const double qi = q * J.i_scale;
    J.qi_f[s] = (float)qi;
    // as the host decides it for host-given ratios
    if (!(qi > 0x1p-60 && qi < 0x1p+100)) J.i_fast = 0;
  }
  {
    // reduce the per-item width extremes written by k_axis, or, when the
    // window is the whole single-isotopologue list, evaluate the list's two
    // Pareto sets here (the extremes k_axis would find, bit for bit: same
    // candidates' superset property and same FP64 expression)
    unsigned long long lmin = ~0ull, lmax = 0ull, gmin = ~0ull, gmax = 0ull;
    if (J.sky_whole) {
      const int dmax = J.log_r > 0.0 ? 2 : 0;
      const int a1 = db.gsky_off[dmax + 1], b1 = db.gsky_off[dmax + 2];
      for (int c = db.gsky_off[dmax] + threadIdx.x; c < b1; c += blockDim.x) {
        const float4 q = db.sky[c];
        const unsigned long long bl = double_to_bits(fwhm_lorentz(J, q.x, q.y, q.z, q.w));
        if (c < a1) lmax = bl > lmax ? bl : lmax;
        else lmin = bl < lmin ? bl : lmin;
      }
    } else {
      for (long long c = threadIdx.x; c < nch; c += blockDim.x) {
        const unsigned long long* P = part + 4 * c;
        lmin = P[0] < lmin ? P[0] : lmin;
        lmax = P[1] > lmax ? P[1] : lmax;
        gmin = P[2] < gmin ? P[2] : gmin;
        gmax = P[3] > gmax ? P[3] : gmax;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      unsigned long long o;
      o = __shfl_xor(lmin, off); lmin = o < lmin ? o : lmin;
      o = __shfl_xor(lmax, off); lmax = o > lmax ? o : lmax;
      o = __shfl_xor(gmin, off); gmin = o < gmin ? o : gmin;
      o = __shfl_xor(gmax, off); gmax = o > gmax ? o : gmax;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { s_w[0][w] = lmin; s_w[1][w] = lmax; s_w[2][w] = gmin; s_w[3][w] = gmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
        lmin = s_w[0][k] < lmin ? s_w[0][k] : lmin;
        lmax = s_w[1][k] > lmax ? s_w[1][k] : lmax;
        gmin = s_w[2][k] < gmin ? s_w[2][k] : gmin;
        gmax = s_w[3][k] > gmax ? s_w[3][k] : gmax;
      }
      int status = 0, ng = 0, nl = 0;
      double wg_min = 0, wl_min = 0, wg_max = 0, wl_max = 0;
      if (i1 > i0) {
        if (single) {
          wg_min = fwhm_gauss(nu_a, dopp1);
          wg_max = fwhm_gauss(nu_b, dopp1);
        } else {
          wg_min = bits_to_double(gmin);
          wg_max = bits_to_double(gmax);
        }
        wl_min = bits_to_double(lmin);
        wl_max = bits_to_double(lmax);
        if (!(wg_min > 0.0) || !(wl_min > 0.0)) {
          status |= LDM_ST_WIDTH;
        } else {
          // N = ceil((log(w_max + 1e-4) - log(w_min)) / dx) + 1, radis/lbl/broadening.py:1835-1836
          ng = (int)ceil((log_call(wg_max + 1e-4) - log_call(wg_min)) / dxG) + 1;
          nl = (int)ceil((log_call(wl_max + 1e-4) - log_call(wl_min)) / dxL) + 1;
          if (ng > LDM_MAX_AX || nl > LDM_MAX_AX || ng < 2 || nl < 2) status |= LDM_ST_AXIS;
          else if (ng * nl > m_cap) status |= LDM_ST_CAP;
        }
      }
      if (status) { ng = 0; nl = 0; }
      s_ng = ng; s_nl = nl; s_status = status;
      s_wg_min = wg_min; s_wl_min = wl_min;
      s_wg_max = wg_max; s_wl_max = wl_max;
      A.ng = ng;
      A.nl = nl;
      A.m = ng * nl;
      A.p = (ng * nl + 1) / 2;
    }
    __syncthreads();
  }
  // axis nodes w_min * exp(dx * i) and their logs, radis/lbl/broadening.py:1837
  for (int i = threadIdx.x; i < s_ng; i += blockDim.x) {
    wgax[i] = s_wg_min * exp_call(dxG * (double)i);
    lgax[i] = log_call(wgax[i]);
    A.laG[i] = lgax[i];
  }
  for (int i = threadIdx.x; i < s_nl; i += blockDim.x) {
    wlax[i] = s_wl_min * exp_call(dxL * (double)i);
    llax[i] = log_call(wlax[i]);
    A.laL[i] = llax[i];
  }
  __syncthreads();
  for (int i = threadIdx.x; i + 1 < s_ng; i += blockDim.x) A.slG[i] = 1.0 / (lgax[i + 1] - lgax[i]);
  for (int i = threadIdx.x; i + 1 < s_nl; i += blockDim.x) A.slL[i] = 1.0 / (llax[i + 1] - llax[i]);
  if (threadIdx.x == 0 && s_ng > 0) {
    // FP32 fast positions work in log2 units (the hardware logarithm)
    const double LOG2E = 1.4426950408889634, LN2 = 0.6931471805599453;
    J.laG0_f = (float)(lgax[0] * LOG2E);
    J.laL0_f = (float)(llax[0] * LOG2E);
    J.inv_dxG_f = (float)(LN2 / dxG);
    J.inv_dxL_f = (float)(LN2 / dxL);
  }
  const int ng = s_ng, nl = s_nl;
  int status = 0;
  const double PI = 3.141592653589793;
  const double h = 0.5 / wstep;                      // n/(2 wstep) step
  const double xrev0 = (double)(L / 2) * (1.0 / ((double)L * wstep));  // rfftfreq(L, wstep)[-1]
  for (int r = threadIdx.x; r < ng * nl; r += blockDim.x) {
    const int l = r / nl, m = r - (r / nl) * nl;
    const double wG = wgax[l], wL = wlax[m];
    int n = 1;
    while (n < LDM_NFOLD_MAX && voigt_ft64(n * h, wG, wL) >= fold_thresh) ++n;
    if (n >= LDM_NFOLD_MAX) status |= LDM_ST_FOLD;
    const int nf = n - 1;
    double f0 = 1.0;
    for (int q = 1; q <= nf; ++q) f0 += voigt_ft64(q * h + ((q & 1) ? xrev0 : 0.0), wG, wL);
    nfold[r] = nf;
    row_a[r] = (float)((PI * wG) * (PI * wG) / (4.0 * 0.6931471805599453));
    row_b[r] = (float)(PI * wL);
    row_inv_f0[r] = (float)(1.0 / f0);
  }
  if (status) atomicOr(&s_status, status);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int st = s_status;
    J.status = st;
    res[blockIdx.x].ng = ng;
    res[blockIdx.x].nl = nl;
    res[blockIdx.x].status = st;
    res[blockIdx.x].wg_min = s_wg_min;
    res[blockIdx.x].wg_max = s_wg_max;
    res[blockIdx.x].wl_min = s_wl_min;
    res[blockIdx.x].wl_max = s_wl_max;
  }
}

// first index in [lo, hi) with nu0[i] >= v
__device__ __forceinline__ long long lower_bound_nu(const double* __restrict__ nu0, long long lo, long long hi,
                                                    double v) {
  while (lo < hi) {
    const long long mid = lo + ((hi - lo) >> 1);
    if (nu0[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Same result computed by a whole wave: 64 probes per round narrow the range
// 64-fold (ballot + popcount), so ~log64(n) dependent loads instead of
// log2(n).  Every lane of the wave must call it with the same arguments.
__device__ __forceinline__ long long lower_bound_nu_wave(const double* __restrict__ nu0, long long lo, long long hi,
                                                         double v) {
  const int lane = threadIdx.x & 63;
  // invariant: the answer lies in [lo, hi]
  while (hi - lo > 64) {
    const long long step = (hi - lo + 63) >> 6;
    const long long idx = lo + (long long)(lane + 1) * step - 1;
    const bool lt = idx < hi && nu0[idx] < v;
    // probes are sorted: the c probes below v are lanes 0..c-1
    const long long c = __popcll(__ballot(lt));
    const long long nhi = lo + (c + 1) * step - 1;
    lo += c * step;
    hi = nhi < hi ? nhi : hi;
  }
  const long long idx = lo + lane;
  const bool lt = idx < hi && nu0[idx] < v;
  return lo + __popcll(__ballot(lt));
}

// ---------------------------------------------------------------------------
// k_tile_bounds: for every fill tile [k_lo, k_hi) the line range that can
// put a corner inside it: k0 in [k_lo-1, k_hi-1].  A line's np.interp
// position is monotone in its shifted centre, so a search on nu0 widened by
// the largest pressure shift (plus 1e-3 of a grid step for rounding) is a
// superset; k_fill filters each line by its exact k0.  Item = (job, tile).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tile_bounds(LdmJob* __restrict__ jobs, int4* __restrict__ items,
                                                     int n_items, LdmDb db) {
  // one wave per item
  const int g = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (g >= n_items) return;
  const int4 it = items[g];
  LdmJob& J = jobs[it.x];
  const long long t = it.y / J.fill_groups;  // item y = tile * fill_groups + row group
  const long long k_lo = t * J.tile_w;
  long long k_hi = k_lo + J.tile_w;
  if (k_hi > J.v) k_hi = J.v;
  long long lo, hi;
  if (k_lo <= 1) lo = J.i0;
  else lo = lower_bound_nu_wave(db.nu0, J.i0, J.i1, J.wmin + (double)(k_lo - 1) * J.delta_w - J.shift_margin);
  if (k_hi >= J.nv) hi = J.i1;
  else hi = lower_bound_nu_wave(db.nu0, lo, J.i1, J.wmin + (double)k_hi * J.delta_w + J.shift_margin);
  if (hi < lo) hi = lo;
  // the range goes into the tile's own work item (offsets from the window
  // start; windows hold < 2^31 lines), so the fill reads it with the item
  if ((threadIdx.x & 63) == 0) {
    items[g].z = (int)(lo - J.i0);
    items[g].w = (int)(hi - J.i0);
  }
}

// Everything the fill needs from one line:
//   shifted centre        radis/lbl/base.py:2351-2360                 (FP64)
//   np.interp indices     radis/lbl/broadening.py:2151-2154, 2253-2263
//   FWHM_G, FWHM_L        radis/lbl/broadening.py:170-172, 341-383
//   S(T)                  radis/lbl/base.py:2664-2674 (Boltzmann and Q ratio FP64)
// I = S(T) * number density * 2^-scale in FP32.
//
// The grid index and its fraction are always FP64 (np.interp rule).  The two
// log-width positions are first evaluated in FP32 from hardware base-2
// logarithms: log2 w within LDM_TIE_LOG2 (a few ulp of |log2 w| <= 32: the
// logarithms, the sum with the isotopologue term, the FP32 axis origin),
// i.e. within LDM_TIE_LOG2 / dx2 axis steps for an axis step of dx2 in log2
// units.  Only when a position lies within the margin tie = max(LDM_TIE_EPS,
// LDM_TIE_LOG2 / dx2) of an axis node, or J.fp64_widths is set, are they
// recomputed in FP64 with NumPy's rule, so the integer indices always equal
// the FP64 ones.  (RADIS's default axes, dx2 ~ 0.2 and 0.3, keep the margin
// at LDM_TIE_EPS; fine axes, dxG = 0.005, widen it to ~2e-3 of a step.)
#define LDM_TIE_EPS 2e-4f
#define LDM_TIE_LOG2 1.5e-5f

// s * exp(x) rounded to FP32 for |x| up to ~700: the exponent x (FP64) is
// reduced in FP64, x = n ln2 + r with |r| <= ln2/2, so the FP32 exp of r
// (relative error ~1e-7) and the exact power of two keep the full relative
// accuracy that an FP32 exponent of magnitude ~100 would lose.
__device__ __forceinline__ float boltzmann_scale(double s, double x) {
#pragma clang fp contract(off)
  const double n = rint(x * 1.4426950408889634);
  const double r = (x - n * 0.6931471805599453) - n * 2.3190468138462996e-17;
  return (float)ldexp(s * (double)__expf((float)r), (int)n);
}

// Same with the prefactor s already in FP32 range (the power of two is
// applied in FP32; results below FP32's range are negligible here).
__device__ __forceinline__ float boltzmann_scale32(float s, double x) {
#pragma clang fp contract(off)
  const double n = rint(x * 1.4426950408889634);
  const double r = (x - n * 0.6931471805599453) - n * 2.3190468138462996e-17;
  return ldexpf(s * __expf((float)r), (int)n);
}

struct LineTerms {
  int k0;
  int l0, m0;
  float tv, tg, tl, I;
};

// One line as stored in the device columns.
struct LineRaw {
  double nu0;
  float dl, ga, gs, na, ns, S, El;
  int iso;
  float evl, evu;  // vibrational energies of the lower/upper state (two-temperature requests)
};

// Column pointers with the optional columns made unconditional: a missing
// delta column reads 0.0f and a missing iso column reads 1 (both at index 0
// of the database's defaults block, selected by an index mask of 0), a
// missing n_self column reads n_air.  Every load is then valid for any line
// index, so loads of several lines can be in flight with no branch between.
struct LineCols {
  const double* nu0;
  const float *S, *El, *ga, *gs, *na, *ns, *dl;
  const uint8_t* iso;
  const float *evl, *evu;
  unsigned dmask, imask;
  __device__ explicit LineCols(const LdmDb& db)
      : nu0(db.nu0), S(db.S_ref), El(db.El), ga(db.g_air), gs(db.g_self), na(db.n_air),
        ns(db.n_self ? db.n_self : db.n_air), dl(db.delta ? db.delta : db.dflt),
        iso(db.iso ? db.iso : reinterpret_cast<const uint8_t*>(db.dflt + 1)), evl(db.evib_l), evu(db.evib_u),
        dmask(db.delta ? ~0u : 0u), imask(db.iso ? ~0u : 0u) {}
  // Line i's columns.  The upload caps a database at 2^29 lines, so every
  // byte offset fits 32 bits: each load is a uniform base plus one 32-bit
  // offset (no 64-bit address arithmetic per column).
  template <typename T>
  __device__ __forceinline__ static T at(const T* p, unsigned byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + byte_off);
  }
  // NONEQ: also the vibrational energies (the request checked they exist).
  // PLAIN: the database has no iso, delta or n_self column (iso 1, no
  // shift, n_self = n_air), so those are not loaded.
  template <bool NONEQ = false, bool PLAIN = false>
  __device__ __forceinline__ void load(unsigned i, LineRaw& r) const {
    const unsigned o4 = i * 4u;
    if (NONEQ) {
      r.evl = at(evl, o4);
      r.evu = at(evu, o4);
    }
    if (PLAIN) {
      r.iso = 1;
      r.dl = 0.0f;
    } else {
      r.iso = at(iso, i & imask);  // first: the per-isotopologue lookups wait on it alone
      r.dl = at(dl, o4 & (dmask << 2));
    }
    r.nu0 = at(nu0, i * 8u);
    r.ga = at(ga, o4);
    r.gs = at(gs, o4);
    r.na = at(na, o4);
    r.ns = PLAIN ? r.na : at(ns, o4);
    r.S = at(S, o4);
    r.El = at(El, o4);
  }
};

// Per-isotopologue values of a line, read from the job table in global
// memory: vector loads, whose completion counter is independent of the LDS
// atomics (an LDS table read would make every line wait for the previous
// line's atomics to retire).
struct IsoVals {
  float lgc;   // log2(2*dopp_iso/c_CGS)
  float qi;    // Q(Tref)/Q(T) * number density * 2^-scale_log2 (when in FP32 range)
  double q;    // Q(Tref)/Q(T)
};

__device__ __forceinline__ void iso_vals(const LdmJob& J, int iso, IsoVals& v) {
  v.lgc = J.lgc_f[iso];
  v.qi = J.qi_f[iso];
  v.q = J.qratio[iso];
}

// Request constants of the per-line computation, read once per workgroup;
// the per-isotopologue values live in LDS (indexed per lane).
struct LineConst {
  double wmin, delta_w, inv_delta_w, grid_fast_hi;
  float inv_delta_f;
  double p_atm, x_air, x_self, log_r;
  double c_boltz, i_scale;
  float c_stim_T, c_stim_R, log_r_f, pa_air_f, pa_self_f;
  float laG0_f, inv_dxG_f, laL0_f, inv_dxL_f;
  float tie_g, tie_l;  // tie margins of the FP32 axis positions (axis steps)
  int nv, fp64_widths, has_delta, min_rms, i_fast;
  double c_vib, c_rot, c_ref;
  const LdmJob* job;
  const LdmAxes* axes;  // the axis tables of the rare exact path (global memory)
  __device__ explicit LineConst(const LdmJob& J, const LdmAxes& A, const LdmDb& db)
      : wmin(J.wmin), delta_w(J.delta_w), inv_delta_w(J.inv_delta_w), grid_fast_hi(J.grid_fast_hi),
        inv_delta_f(J.inv_delta_f), p_atm(J.p_atm), x_air(J.x_air),
        x_self(J.x_self), log_r(J.log_r), c_boltz(J.c_boltz), i_scale(J.i_scale), c_stim_T(J.c_stim_T),
        c_stim_R(J.c_stim_R), log_r_f(J.log_r_f), pa_air_f(J.pa_air_f), pa_self_f(J.pa_self_f),
        laG0_f(J.laG0_f), inv_dxG_f(J.inv_dxG_f), laL0_f(J.laL0_f), inv_dxL_f(J.inv_dxL_f),
        tie_g(fmaxf(LDM_TIE_EPS, LDM_TIE_LOG2 * J.inv_dxG_f)), tie_l(fmaxf(LDM_TIE_EPS, LDM_TIE_LOG2 * J.inv_dxL_f)),
        nv((int)J.nv),
        fp64_widths(J.fp64_widths), has_delta(db.delta != nullptr), min_rms(J.min_rms), i_fast(J.i_fast), c_vib(J.c_vib),
        c_rot(J.c_rot), c_ref(J.c_ref), job(&J), axes(&A) {}
};

// min-RMS weights, radis/lbl/broadening.py:2265-2303: the Gaussian and
// Lorentzian fractions get second-order corrections in the three fractions,
// with coefficients from the width ratio alpha = wL/wG (FWHM) and
// dv/wG.  FP32: the corrections feed the FP32 weights either way.
__device__ __forceinline__ void min_rms_weights(const LdmJob& J, float log_wg, float wl, LineTerms& t) {
  const float wg = __expf(log_wg);
  const float dxv = J.mr_dv / wg;
  const float la = __logf(wl) - log_wg;  // log(alpha)
  const float dxG = J.mr_dxG, dxL = J.mr_dxL;
  const float ln2 = 0.69314718055994531f;
  const float r_gv = 8.0f * ln2;
  const float r_gg = 2.0f - 1.0f / __powf(J.mr_c1gg + J.mr_c2gg * __expf(la * (2.0f / 1.50f)), 1.50f);
  const float r_gl = -2.0f * ln2 * __expf(2.0f * la);
  const float r_lg = 1.0f / __powf(J.mr_c1lg * __expf(la * (1.0f / 2.25f)) + J.mr_c2lg * __expf(la * (4.0f / 2.25f)),
                                   2.25f);
  const float qv = t.tv * (t.tv - 1.0f), qg = t.tg * (t.tg - 1.0f), ql = t.tl * (t.tl - 1.0f);
  const float ag = t.tg + (r_gv * qv * dxv * dxv + r_gg * qg * dxG * dxG + r_gl * ql * dxL * dxL) / (2.0f * dxG);
  const float al = t.tl + (r_lg * qg * dxG * dxG + ql * dxL * dxL) / (2.0f * dxL);
  t.tg = ag;
  t.tl = al;
}

__device__ __forceinline__ bool near_node(float p, int n, float eps) {
  const float f = p - floorf(p);
  return f < eps || f > 1.0f - eps || p < eps || p > (float)(n - 1) - eps;
}

// FP64 Lorentzian FWHM from the request constants (same expression as
// fwhm_lorentz).
__device__ __forceinline__ double fwhm_lorentz_c(const LineConst& C, double g_air, double g_self, double n_air,
                                                 double n_self) {
#pragma clang fp contract(off)
  // exp(n log(Tref/T)) here (pow, as in fwhm_lorentz, would push the fill's
  // registers into spills): within 2 ulp of NumPy's pow; it only places
  // lines within 2e-4 of an axis node
  const double ra = exp(n_air * C.log_r);
  const double rs = (n_self == n_air) ? ra : exp(n_self * C.log_r);
  const double air = ((ra * g_air) * C.p_atm) * C.x_air;
  const double slf = rs * ((g_self * C.p_atm) * C.x_self);
  return (air + slf) * 2.0;
}

template <bool NONEQ = false>
__device__ __forceinline__ void line_terms(const LineConst& C, const LineRaw& r, const IsoVals& iv,
                                           int ng, int nl, LineTerms& t) {
  const double nu0 = r.nu0;
  double nu = nu0;
  if (C.has_delta) {
#pragma clang fp contract(off)
    nu = nu0 + ((double)r.dl * C.p_atm);
  }
  grid_index(nu, C.wmin, C.delta_w, C.inv_delta_w, C.nv, C.grid_fast_hi, C.inv_delta_f, t.k0, t.tv);
  const int iso = r.iso;
  const float nuf = (float)nu0;
  // FP32 positions on the log axes
  const float ra = __expf(r.na * C.log_r_f);
  const float rs = (r.ns == r.na) ? ra : __expf(r.ns * C.log_r_f);
  const float wl_f = 2.0f * (ra * r.ga * C.pa_air_f + rs * r.gs * C.pa_self_f);  // as fwhm_lorentz_f32
  // positions on the log axes from base-2 logarithms (v_log_f32; ~1 ulp,
  // far inside the tie margin that sends a line to the exact FP64 path)
  const float l2wg = __builtin_amdgcn_logf(nuf) + iv.lgc;  // log2 FWHM_G
  const float l2wl = __builtin_amdgcn_logf(wl_f);          // log2 FWHM_L
  const float pg = (l2wg - C.laG0_f) * C.inv_dxG_f;
  const float pl = (l2wl - C.laL0_f) * C.inv_dxL_f;
  const bool exact_g = C.fp64_widths || near_node(pg, ng, C.tie_g);
  const bool exact_l = C.fp64_widths || near_node(pl, nl, C.tie_l);
  if (exact_g) {
    // the exact path is rare (positions within 2e-4 of a node): its tables
    // stay in global memory, where they cost the fill no LDS
    const double p = axis_position(log(fwhm_gauss(nu0, C.job->dopp[iso])), C.axes->laG, C.axes->slG, ng);
    t.l0 = (int)p;
    t.tg = (float)(p - (double)t.l0);
  } else {
    const float f = floorf(pg);
    t.l0 = (int)f;
    t.tg = pg - f;
  }
  if (exact_l) {
    const double p = axis_position(log(fwhm_lorentz_c(C, r.ga, r.gs, r.na, r.ns)), C.axes->laL, C.axes->slL, nl);
    t.m0 = (int)p;
    t.tl = (float)(p - (double)t.m0);
  } else {
    const float f = floorf(pl);
    t.m0 = (int)f;
    t.tl = pl - f;
  }
  if (t.l0 > ng - 2) { t.l0 = ng - 2; t.tg = 1.0f; }
  if (t.m0 > nl - 2) { t.m0 = nl - 2; t.tl = 1.0f; }
  if (C.min_rms) min_rms_weights(*C.job, l2wg * 0.69314718055994531f, wl_f, t);
  const double s64 = (double)r.S * iv.q * C.i_scale;
  // (1 - exp(-c2 nu/T)) / (1 - exp(-c2 nu/Tref)), radis/lbl/base.py:2664-2674:
  // for |x| >= 0.1 the difference 1 - e^x loses under 4 bits, so the fast
  // exponential and division (~2 ulp) keep ~1e-6 relative accuracy; smaller
  // arguments (far-infrared lines, hot gas) take expm1f
  if (NONEQ) {
    // two-temperature populations, radis/lbl/base.py:3479-3530 and
    // 3937-3947: S0/Q_neq * [e^{-c2(Evl/Tv + Erl/Tr)} - e^{-c2(Evu/Tv + Eru/Tr)}]
    // with Er = E - Ev and E_u = E_l + nu0, written as the equilibrium form
    // S_ref Q(Tref)/Q_neq e^{x} (1 - e^{y}) / (1 - e^{-c2 nu0/Tref}), where
    // x = -c2(Evl/Tv + Erl/Tr - El/Tref) and
    // y = -c2((Evu-Evl)/Tv + (nu0-(Evu-Evl))/Tr)  (y > 0: population inversion, S < 0)
#pragma clang fp contract(off)
    const double el = (double)r.El, evl = (double)r.evl, dev = (double)r.evu - evl;
    const double x = (C.c_vib * evl + C.c_rot * (el - evl)) + C.c_ref * el;
    const float y = (float)(C.c_vib * dev + C.c_rot * (nu0 - dev));
    const float xR = C.c_stim_R * nuf;
    const float stim = expm1f(y) / expm1f(xR);
    t.I = boltzmann_scale(s64, x) * stim;
  } else {
    const float xT = C.c_stim_T * nuf, xR = C.c_stim_R * nuf;
    float stim;
    if (xT < -0.1f && xR < -0.1f) stim = (1.0f - __expf(xT)) * __builtin_amdgcn_rcpf(1.0f - __expf(xR));
    else stim = expm1f(xT) / expm1f(xR);
    const double x = C.c_boltz * (double)r.El;
    t.I = (C.i_fast ? boltzmann_scale32(r.S * iv.qi, x) : boltzmann_scale(s64, x)) * stim;
  }
}

// ---------------------------------------------------------------------------
// k_fill: one workgroup per (request, tile of W grid points).  Every line
// whose 2x2x2 stencil touches the tile gets its LineTerms, then adds the
// weights (radis/lbl/broadening.py:2305-2318) of its corners that fall in
// the tile into the LDS tile with FP64 atomics.  The tile is written to HBM
// once, so the LDM needs no clearing and no global atomics.  Row index
// r = l*N_L + m; LDS row stride W + ldm_fill_pad(W) doubles (neighbouring
// rows on different banks).  The tile accumulates in FP64 with ds_add_f64: on gfx950
// LDS ds_add_f32 retires far slower than the 64-bit and integer LDS atomics
// (measured ~3 cycles per lane per CU, which bounded this kernel), and the
// FP64 sums are also more accurate.
// Consecutive lines go to consecutive threads, so loads are coalesced.  The
// raw columns of a thread's next two lines are loaded while the current one
// is computed, and no LDS read
// sits on the common path (only atomics), so atomics retire in the background.
// ---------------------------------------------------------------------------

// GROUPED: the LDM is taller than one tile's LDS (J.fill_groups > 1 for some
// request of the launch); item y = tile * fill_groups + g, and the workgroup
// keeps rows [g * fill_rows, g * fill_rows + MR) only: each of its 8 atomics
// checks its corner's row.  A request with fill_groups = 1 has one group of
// every row.
template <bool NONEQ, bool PLAIN, bool GROUPED>
__device__ __forceinline__ void fill_tile(const LdmJob& J, const LdmAxes& A, const int4 it, const LdmDb& db,
                                          double* acc) {
  const int M = A.m;
  const int ng = A.ng, nl = A.nl;
  const int W = J.tile_w, Wp = W + ldm_fill_pad(W);
  const int tile = GROUPED ? it.y / J.fill_groups : it.y;
  const int r0 = GROUPED ? (it.y - tile * J.fill_groups) * J.fill_rows : 0;  // even
  const int MR = GROUPED ? min(J.fill_rows, M - r0) : M;                     // rows of this group
  if (GROUPED && MR <= 0) return;  // uniform: before any barrier
  const long long k_lo = (long long)tile * W;
  long long k_hi = k_lo + W;
  if (k_hi > J.v) k_hi = J.v;
  const LineConst C(J, A, db);
  const LineCols cols(db);
  const unsigned lo = (unsigned)(J.i0 + it.z), hi = (unsigned)(J.i0 + it.w);  // written by k_tile_bounds
  // Thread t takes lines lo + t, lo + t + T, ...: every load instruction of
  // a wave reads 64 consecutive lines.  (Giving each lane its own run of
  // lines spreads a wave's loads over as many cache lines and pages as it has
  // lanes, which measured up to 10x slower on dense lists; collisions of
  // neighbouring lines on one LDS address cost far less.)
  const unsigned hlast = hi - 1;
  // first two lines in flight before the LDS set-up (indices clamped to the
  // tile's last line, which is valid whenever the thread has a line)
  LineRaw n1, n2;
  IsoVals v1;
  const unsigned first = lo + threadIdx.x;
  // a plain database has one isotopologue: its values once, for every line
  if (PLAIN) iso_vals(J, 1, v1);
  if (first < hi) {
    cols.load<NONEQ, PLAIN>(first, n1);
    if (!PLAIN) iso_vals(J, n1.iso, v1);
    const unsigned l1 = first + LDM_FILL_T;
    cols.load<NONEQ, PLAIN>(l1 < hlast ? l1 : hlast, n2);
  }
  {
    // whole 16-byte pairs, and the last cell alone when MR * Wp is odd
    const int n = MR * Wp, n2 = n >> 1;
    double2* a2 = reinterpret_cast<double2*>(acc);
    for (int e = threadIdx.x; e < n2; e += blockDim.x) a2[e] = make_double2(0.0, 0.0);
    if ((n & 1) && threadIdx.x == 0) acc[n - 1] = 0.0;
  }
  __syncthreads();
  const int wlen = (int)(k_hi - k_lo);
  const int sg = nl * Wp;
  // Corner order per lane.  The 8 atomics of a line go to its stencil's
  // cells (column c0+dc, G row l0+dl, L row m0+dm); lanes whose lines share
  // a stencil (on a dense list most of a wave: one grid column, one G row,
  // a dozen L rows) would hit the same cell in the same instruction, and
  // same-address lanes serialise.  Lane bits (x0, x1, x2) map the
  // instruction for corner (dc, dl, dm) to the cell (dc^x0, dl^x1, dm^x2),
  // so such lanes spread over 8 cells.  Each cell still receives exactly
  // the product it received before (same operands, same order).
  const int x0 = threadIdx.x & 1, x1 = (threadIdx.x >> 1) & 1, x2 = (threadIdx.x >> 2) & 1;
  const int s_c = x0 ? -1 : 1, s_g = x1 ? -sg : sg, s_l = x2 ? -Wp : Wp;  // steps from the lane's first cell
  const int x_off = x0 + (x1 ? sg : 0) + (x2 ? Wp : 0);                  // the lane's first cell
  for (unsigned i = first; i < hi; i += LDM_FILL_T) {
    const LineRaw cur = n1;
    const IsoVals cv = v1;
    n1 = n2;
    if (!PLAIN) iso_vals(J, n1.iso, v1);
    const unsigned i2 = i + 2 * LDM_FILL_T;
    cols.load<NONEQ, PLAIN>(i2 < hlast ? i2 : hlast, n2);
    LineTerms t;
    line_terms<NONEQ>(C, cur, cv, ng, nl, t);
    const int c0 = t.k0 - (int)k_lo;  // column of corner k0 in the tile, in [-1, W-1] if touching
    if (c0 < -1 || c0 > wlen - 1) continue;
    // the stencil's cells are base + {0, 1} + {0, nl*Wp} + {0, Wp}; the
    // weight of cell (c, l, m) is fc[c] * (fg[l] * fl[m])
    const int rb = t.l0 * nl + t.m0 - r0;  // the stencil's first row in the group
    double* base = acc + rb * Wp + c0 + x_off;
    // GROUPED: which of this lane's corner rows (its first, + L step, + G
    // step, both) fall in the group
    bool q00 = true, q01 = true, q10 = true, q11 = true;
    if constexpr (GROUPED) {
      const int ra = rb + (x1 ? nl : 0) + (x2 ? 1 : 0), sl = x2 ? -1 : 1, sgr = x1 ? -nl : nl;
      q00 = (unsigned)ra < (unsigned)MR;
      q01 = (unsigned)(ra + sl) < (unsigned)MR;
      q10 = (unsigned)(ra + sgr) < (unsigned)MR;
      q11 = (unsigned)(ra + sgr + sl) < (unsigned)MR;
    }
    const float ug = 1.0f - t.tg, ul = 1.0f - t.tl;
    const float fg0 = x1 ? t.tg : ug, fg1 = x1 ? ug : t.tg;  // this lane's corner dl = 0, 1
    const float fl0 = x2 ? t.tl : ul, fl1 = x2 ? ul : t.tl;
    const float w00 = fg0 * fl0, w01 = fg0 * fl1, w10 = fg1 * fl0, w11 = fg1 * fl1;
    const float v0 = t.I * (1.0f - t.tv), v1 = t.I * t.tv;
    // columns c0 (needs c0 >= 0) and c0 + 1 (needs c0 + 1 < wlen)
    if (x0 ? (c0 + 1 < wlen) : (c0 >= 0)) {
      const float v = x0 ? v1 : v0;
      if (q00) atomicAdd(base, (double)(v * w00));
      if (q01) atomicAdd(base + s_l, (double)(v * w01));
      if (q10) atomicAdd(base + s_g, (double)(v * w10));
      if (q11) atomicAdd(base + s_g + s_l, (double)(v * w11));
    }
    if (x0 ? (c0 >= 0) : (c0 + 1 < wlen)) {
      const float v = x0 ? v0 : v1;
      double* b1 = base + s_c;
      if (q00) atomicAdd(b1, (double)(v * w00));
      if (q01) atomicAdd(b1 + s_l, (double)(v * w01));
      if (q10) atomicAdd(b1 + s_g, (double)(v * w10));
      if (q11) atomicAdd(b1 + s_g + s_l, (double)(v * w11));
    }
  }
  __syncthreads();
  // flush: the LDM interleaves rows 2p and 2p+1 (element n of row 2p+h at
  // float 2n+h of pair p).  A lane writes 2 columns of one pair (16 bytes)
  // per step: (2p, c), (2p+1, c), (2p, c+1), (2p+1, c+1); the second row of
  // an odd last pair is written as zeros.  W is a power of two; even rows
  // start 16-byte aligned.
  const int P2 = (MR + 1) >> 1;
  const int W2 = W >> 1, sh2 = __builtin_ctz(W2);
  const long long pstride = 2 * J.ldm_stride;                   // floats per pair
  float* const out = J.ldm + 2 * k_lo + (long long)(r0 >> 1) * pstride;  // column k_lo of the group's first pair
  for (int w = threadIdx.x; w < (P2 << sh2); w += LDM_FILL_T) {
    const int p = w >> sh2, c = 2 * (w & (W2 - 1));
    if (c >= wlen) continue;
    const double* src = acc + (2 * p) * Wp + c;
    const double2 a = *reinterpret_cast<const double2*>(src);
    // row 2p starts on an even cell (16-byte aligned), row 2p+1 on an odd one when Wp is odd
    const double2 b = (2 * p + 1 < MR) ? ((Wp & 1) ? make_double2(src[Wp], src[Wp + 1])
                                                   : *reinterpret_cast<const double2*>(src + Wp))
                                      : make_double2(0.0, 0.0);
    float* dst = out + p * pstride + 2 * c;
    if (c + 2 <= wlen) *reinterpret_cast<float4*>(dst) = make_float4((float)a.x, (float)b.x, (float)a.y, (float)b.y);
    else *reinterpret_cast<float2*>(dst) = make_float2((float)a.x, (float)b.x);
  }
}

// (the fill kernel, k_fill, is in ldm_wrap.h: its launch also runs the
// image-correction items)

// four-step twiddle W_L^e (forward sign)
__device__ __forceinline__ float2 twiddle_L(const LdmJob& J, long long e) {
  const float2 h = gld(J.twl_hi, 8u * (unsigned)(e >> LDM_TW_LO_BITS));  // e < L <= 2^23
  const float2 l = gld(J.twl_lo, 8u * (unsigned)(e & ((1 << LDM_TW_LO_BITS) - 1)));
  return cmul(h, l);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Map hardware block h to a logical block so that consecutive logical
// blocks (neighbouring column blocks, which share cache lines) run on the
// same XCD at about the same time.
__device__ __forceinline__ int xcd_block(int h, int G) {
  const int per = G >> 3, rem = G & 7;
  const int x = h & 7, s = h >> 3;
  return x * per + (x < rem ? x : rem) + s;
}

// ---------------------------------------------------------------------------
// k_fft_a: item = (job, pair p, column block cb).  Packs LDM rows 2p and
// 2p+1 as z = x_2p + i x_2p+1 (rows are zero beyond V), runs the N1-point
// DFTs of columns n2 in [cb*B, cb*B+B) and writes
//   Y_p[k1][n2] = W_L^(n2 k1) * sum_n1 z[n2 + N2 n1] W_N1^(n1 k1).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(LDM_FFT_T, 4) void k_fft_a(const LdmJob* __restrict__ jobs,
                                                     const LdmAxes* __restrict__ axes,
                                                     const int4* __restrict__ items, int n_items) {
  __shared__ float2 buf[LdsSize<LDM_CAP>::value];
  const int4 it = items[xcd_block(blockIdx.x, n_items)];
  const LdmJob& J = jobs[it.x];
  const int p = it.y;
  if (J.status || p >= axes[it.x].p) return;
  const int N1 = J.n1, N2 = J.n2, bs = J.bA, B = 1 << bs;
  const long long n2_0 = (long long)it.z * B;
  const float2* rz = reinterpret_cast<const float2*>(J.ldm) + (long long)p * J.ldm_stride;  // rows 2p, 2p+1
  const bool has_b = (2 * p + 1) < axes[it.x].m;
  const long long V = J.v;
  for (int e = threadIdx.x; e < (N1 << bs); e += LDM_FFT_T) {
    const int n1 = e >> bs, b = e & (B - 1);
    const long long n = n2_0 + b + (long long)N2 * n1;
    float xa = 0.0f, xb = 0.0f;
    if (n < V) {
      const float2 z = rz[n];
      xa = z.x;
      if (has_b) xb = z.y;
    }
    buf[pad_idx(e)] = make_float2(xa, xb);
  }
  __syncthreads();
  lds_fft<true, LDM_CAP, LDM_FFT_T>(buf, N1, bs, J.rad1, J.nrad1, J.tw1);
  float2* Y = J.ybuf + (long long)p * J.L;
  for (int e = threadIdx.x; e < (N1 << bs); e += LDM_FFT_T) {
    const int k1 = e >> bs, b = e & (B - 1);
    const long long n2 = n2_0 + b;
    Y[(long long)k1 * N2 + n2] = cmul(buf[pad_idx(e)], twiddle_L(J, n2 * k1));
  }
}

// Folded, normalised FT lineshape of LDM row r (N_L = nl) at bin k (|k|
// folded to [0, L/2]); radis/lbl/broadening.py:1895-1926.  Bins below
// wrap_k add the circular-image correction of the row's L-axis node
// (wrap_item) before the normalisation, as the image terms belong to the
// kernel that RADIS divides by its folded value at x = 0.
__device__ __forceinline__ float lineshape_ft(const LdmJob& J, int r, int nl, long long kk) {
  const float a = J.row_a[r], bco = J.row_b[r];
  const float x = (float)kk * J.inv_Lws;
  float v = __expf(-(a * x * x) - bco * x);
  const int nf = J.nfold[r];
  if (nf > 0) {
    const float xr = (float)(J.L / 2 - kk) * J.inv_Lws;
    for (int n = 1; n <= nf; ++n) {
      const float xa = (n & 1) ? ((float)n * J.half_inv_ws + xr) : ((float)n * J.half_inv_ws + x);
      v += __expf(-(a * xa * xa) - bco * xa);
    }
  }
  if (kk < J.wrap_k) v += J.wrap_h[(r % nl) * J.wrap_k + (int)kk];
  return v * J.row_inv_f0[r];
}

// ---------------------------------------------------------------------------
// k_fft_b: item = (job, block of B2 rows k1).  For every pair p:
//   Z_p[k1 + N1 k2] = sum_n2 Y_p[k1][n2] W_N2^(n2 k2)
//   A[k] += Z_p[k] * (F_2p[k] - i F_2p+1[k])
// and then B[k1][n2'] = W_L^(-n2' k1) * sum_k2 A[k1 + N1 k2] W_N2^(-n2' k2).
// With real LDM rows x_a, x_b and real symmetric F, Re(IFFT(A)) equals
// IFFT(X_a F_a + X_b F_b): the pair packing needs no real-FFT unpacking.
// CAP/T select a small tile (many workgroups, N2 <= 2048) or a large one.
// ---------------------------------------------------------------------------
template <int CAP, int T>
__global__ __launch_bounds__(T, 768 / T) void k_fft_b(const LdmJob* __restrict__ jobs, const LdmAxes* __restrict__ axes,
                                             const int4* __restrict__ items) {
  __shared__ float2 buf[LdsSize<CAP>::value];
  const int4 it = items[blockIdx.x];
  const LdmJob& J = jobs[it.x];
  if (J.status) return;
  const int N1 = J.n1, N2 = J.n2, bs = J.bB, B = 1 << bs;
  const int k1_0 = it.y * B;
  const long long L = J.L;
  const int P = axes[it.x].p, M = axes[it.x].m, NLr = axes[it.x].nl;
  const int total = N2 << bs;
  constexpr int NACC = CAP / T;
  float2 acc[NACC];
#pragma unroll
  for (int s = 0; s < NACC; ++s) acc[s] = make_float2(0.0f, 0.0f);
  // element e of a pair's block: row b = e / N2, column n2 = e % N2 (n2
  // fastest in global memory, transposed into [n2][b] in LDS); the next
  // pair's block is fetched into registers while this one is transformed
  float2 pre[NACC];
#pragma unroll
  for (int s = 0; s < NACC; ++s) {
    const int e = threadIdx.x + s * T;
    if (e < total && P > 0) {
      const int b = e / N2, n2 = e - b * N2;
      pre[s] = J.ybuf[(long long)(k1_0 + b) * N2 + n2];
    }
  }
  for (int p = 0; p < P; ++p) {
#pragma unroll
    for (int s = 0; s < NACC; ++s) {
      const int e = threadIdx.x + s * T;
      if (e < total) {
        const int b = e / N2, n2 = e - b * N2;
        buf[pad_idx((n2 << bs) + b)] = pre[s];
      }
    }
    __syncthreads();
    if (p + 1 < P) {
      const float2* Yn = J.ybuf + (long long)(p + 1) * L;
#pragma unroll
      for (int s = 0; s < NACC; ++s) {
        const int e = threadIdx.x + s * T;
        if (e < total) {
          const int b = e / N2, n2 = e - b * N2;
          pre[s] = Yn[(long long)(k1_0 + b) * N2 + n2];
        }
      }
    }
    lds_fft<true, CAP, T>(buf, N2, bs, J.rad2, J.nrad2, J.tw2);
    const int ra = 2 * p, rb = 2 * p + 1;
#pragma unroll
    for (int s = 0; s < NACC; ++s) {
      const int e = threadIdx.x + s * T;
      if (e < total) {
        const int k2 = e >> bs, b = e & (B - 1);
        const long long k = (long long)(k1_0 + b) + (long long)N1 * k2;
        const long long kk = (k <= L - k) ? k : (L - k);
        const float fa = lineshape_ft(J, ra, NLr, kk);
        const float fb = (rb < M) ? lineshape_ft(J, rb, NLr, kk) : 0.0f;
        const float2 z = buf[pad_idx(e)];
        acc[s].x += z.x * fa + z.y * fb;
        acc[s].y += z.y * fa - z.x * fb;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < NACC; ++s) {
    const int e = threadIdx.x + s * T;
    if (e < total) buf[pad_idx(e)] = acc[s];
  }
  __syncthreads();
  lds_fft<false, CAP, T>(buf, N2, bs, J.rad2, J.nrad2, J.tw2);
  for (int e = threadIdx.x; e < total; e += T) {
    const int b = e / N2, n2 = e - (e / N2) * N2;
    const int k1 = k1_0 + b;
    J.bbuf[(long long)k1 * N2 + n2] = cmulc(buf[pad_idx((n2 << bs) + b)], twiddle_L(J, (long long)n2 * k1));
  }
}

// abscoeff (times 2^-scale_log2) at grid point n and the requested derived
// quantities (FP64): radis/lbl/factory.py:1289-1297,
// radis/spectrum/equations.py:21-44, radis/phys/blackbody.py:35-74.
#define LDM_H_SI 6.62607015e-34      /* radis/phys/constants.py:22,33,39 */
#define LDM_C_SI 2.99792458e8
#define LDM_KB_SI 1.3806490e-23
__device__ __forceinline__ void store_derived(const LdmJob& J, int n, float v) {
#pragma clang fp contract(off)
  J.out[n] = v;
  const double a_scaled = (double)v * J.path_length;   // absorbance times 2^-scale_log2
  const double A = ldexp(a_scaled, J.scale_log2);
  const double emis = -expm1(-A);
  if (J.d_abs) J.d_abs[n] = (float)a_scaled;
  if (J.d_emis) J.d_emis[n] = (float)emis;
  if (J.d_trans) J.d_trans[n] = (float)(1.0 - emis);
  if (J.d_rad) {
    const double nu = J.wmin + (double)n * J.delta_w;   // grid value (numpy.arange rule)
    const double lbd = (1.0 / nu * 1e9 / 100.0) * 1e-9;  // m (radis/phys/convert.py:120-122)
    const double pl = (2.0 * LDM_H_SI * (LDM_C_SI * LDM_C_SI) / (lbd * lbd * lbd * lbd * lbd)) *
                      (1.0 / (exp(LDM_H_SI * LDM_C_SI / (lbd * LDM_KB_SI * J.Tgas)) - 1.0));
    J.d_rad[n] = (float)(emis * (pl * 1e-10));
  }
}

// ---------------------------------------------------------------------------
// k_fft_c: item = (job, column block cb).  y[n2 + N2 n1] =
// sum_k1 B[k1][n2] W_N1^(-n1 k1); abscoeff[n] = Re(y[n]) / (L wstep) for
// n < N_v (radis/lbl/broadening.py:2529-2530 crop and /wstep).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(LDM_FFT_T, 4) void k_fft_c(const LdmJob* __restrict__ jobs, const int4* __restrict__ items,
                                                     int n_items) {
  __shared__ float2 buf[LdsSize<LDM_CAP>::value];
  const int4 it = items[xcd_block(blockIdx.x, n_items)];
  const LdmJob& J = jobs[it.x];
  const int N1 = J.n1, N2 = J.n2, bs = J.bA, B = 1 << bs;
  const long long n2_0 = (long long)it.z * B;
  if (J.status) {
    for (int e = threadIdx.x; e < (N1 << bs); e += LDM_FFT_T) {
      const long long n = n2_0 + (e & (B - 1)) + (long long)N2 * (e >> bs);
      if (n < J.nv) J.out[n] = 0.0f;
    }
    return;
  }
  for (int e = threadIdx.x; e < (N1 << bs); e += LDM_FFT_T) {
    const int k1 = e >> bs, b = e & (B - 1);
    buf[pad_idx(e)] = J.bbuf[(long long)k1 * N2 + n2_0 + b];
  }
  __syncthreads();
  lds_fft<false, LDM_CAP, LDM_FFT_T>(buf, N1, bs, J.rad1, J.nrad1, J.tw1);
  const float sc = J.out_scale;
  const bool derived = J.derived != 0;
  for (int e = threadIdx.x; e < (N1 << bs); e += LDM_FFT_T) {
    const int n1 = e >> bs, b = e & (B - 1);
    const long long n = n2_0 + b + (long long)N2 * n1;
    if (n < J.nv) {
      if (derived) store_derived(J, (int)n, buf[pad_idx(e)].x * sc);
      else J.out[n] = buf[pad_idx(e)].x * sc;
    }
  }
}

// ---------------------------------------------------------------------------
// k_line_terms: the per-line values k_fill uses, for every line of one
// request's window, written out for validation against NumPy.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_line_terms(const LdmJob* __restrict__ jobs, const LdmAxes* __restrict__ axes,
                                                    int job, LdmDb db, long long* __restrict__ k0, int* __restrict__ l0,
                                                    int* __restrict__ m0, float* __restrict__ tv, float* __restrict__ tg,
                                                    float* __restrict__ tl, float* __restrict__ I) {
  const LdmJob& J = jobs[job];
  const LdmAxes& A = axes[job];
  if (A.m == 0 || J.status) return;
  const LineConst C(J, A, db);
  const LineCols cols(db);
  const long long n = J.i1 - J.i0;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    LineRaw r;
    IsoVals iv;
    LineTerms t;
    if (J.noneq) {
      cols.load<true>((unsigned)(J.i0 + q), r);
      iso_vals(J, r.iso, iv);
      line_terms<true>(C, r, iv, A.ng, A.nl, t);
    } else {
      cols.load((unsigned)(J.i0 + q), r);
      iso_vals(J, r.iso, iv);
      line_terms(C, r, iv, A.ng, A.nl, t);
    }
    k0[q] = t.k0; l0[q] = t.l0; m0[q] = t.m0;
    tv[q] = t.tv; tg[q] = t.tg; tl[q] = t.tl; I[q] = t.I;
  }
}
