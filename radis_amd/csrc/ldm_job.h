// Device-side job descriptor shared by the host planner and the kernels.
//
// One LdmJob describes one spectrum request: its grid, its window of lines
// in the device line list, its physical conditions and its workspace.  The
// host fills it before every run (one H2D copy of the whole job table); the
// kernels write the LDM axes and sizes into a separate LdmAxes array, so no
// host round trip sits inside the pipeline and the same launch sequence
// serves one request, a ragged batch, or a replayed hipGraph.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LDM_MAX_AX 128     // maximum N_G or N_L
#define LDM_MAX_ISO 16     // isotopologue slots
#define LDM_MAX_RAD 12     // radix stages per sub-FFT
#define LDM_CAP 8192       // complex elements held in LDS by a pass A/C workgroup (and a large pass B one)
#define LDM_FFT_T 512      // threads of a pass A/C workgroup
#define LDM_CAP_B 2048     // complex elements held by a small pass B workgroup
#define LDM_FFT_T_B 256    // threads of a small pass B workgroup
#define LDM_FILL_T 256     // threads of a fill workgroup
// LDS row stride of a fill tile: tile_w + ldm_fill_pad(tile_w) doubles.  The
// stride sets how a wave's FP64 atomics (lanes on neighbouring L rows, G rows
// and columns) spread over the LDS banks; measured best: + 2 for 32-point
// tiles (dense lists: 100 M lines on 600,001 points), + 3 for wider ones
// (+ 1, + 3, + 5, + 7 measured slower or equal in each case).
constexpr int ldm_fill_pad(int tile_w) { return tile_w >= 64 ? 3 : 2; }
#define LDM_AX_CHUNK 4096  // lines per axis-reduction workgroup (16 per lane)
#define LDM_SKY_BLOCK 4096  // lines per width-candidate block
#define LDM_SKY_GROUP_MIN 8  // fewest candidate blocks per k_axis workgroup
#define LDM_SKY_WGS 600       // most candidate workgroups of k_axis per request (about)
static_assert(LDM_AX_CHUNK == LDM_SKY_BLOCK, "a partial block is one scan chunk");
#define LDM_TW_LO_BITS 10  // four-step twiddle split W_L^e = hi[e>>10] * lo[e&1023]
#define LDM_NFOLD_MAX 100000
// Circular-image correction (wrap_item, ldm_wrap.h): a transform of length L' > 2 N_v
// reproduces RADIS's L = 2 N_v result when the lineshape FT of every LDM row
// gains, at its lowest K bins, the DFT (length L') of the difference of the
// L- and L'-periodised Lorentzian over offsets |d| <= N_v, continued
// smoothly to zero over the gap of L' - 2 N_v offsets no output reads: an
// erfc window of width w = min(gap, N_v / 2) and sigma = w / LDM_WRAP_NSIG.
// Its spectrum is a Gaussian of width L'/(2 pi sigma) bins, kept up to
// LDM_WRAP_CK of those widths:
// K = ceil(LDM_WRAP_CK * LDM_WRAP_NSIG / (2 pi) * L' / w) <= LDM_WRAP_KMAX
// (so gap >= ~2.5 % of L', and L' <= ~20 N_v).  K <= 256 <= N1 of every
// compiled split keeps the corrected bins at k2 = 0 and k2 = N2 - 1 of the
// four-step order.  The table is computed from LDM_WRAP_LC >= 8 K samples.
#define LDM_WRAP_NSIG 10.0
#define LDM_WRAP_CK 4.0
#define LDM_WRAP_KMAX 256
#define LDM_WRAP_LC 2048

// Compile-time-sized FFT passes (ldm_fftc.h): tile sizes and threads.
#define FC_CAP_A 16384 // complex elements of a pass A tile (at most)
#define FC_T_A 1024
#define FC_CAP 8192    // complex elements of a pass C tile (at most)
#define FC_T 512
#define FC_CAP_B 2048  // complex elements of a pass B tile
#define FC_T_B 256
#define FC_C_SUM (1 << 20)  // pass C launch-class flag: B is the sum of two pass B shares
#define FC_B_ROWS_LDS 512  // rows whose lineshape coefficients pass B stages in LDS (more: read from global memory)
constexpr int fc_pow2_floor(int n) { return n < 2 ? 1 : 2 * fc_pow2_floor(n >> 1); }
// columns per workgroup of pass A and of pass C for N1, and rows per
// workgroup of pass B for N2
constexpr int fc_cols_a(int n1) { return fc_pow2_floor(FC_CAP_A / n1); }
constexpr int fc_cols_c(int n1) { return fc_pow2_floor(FC_CAP / n1); }
constexpr int fc_rows_b(int n2) { return FC_CAP_B / n2; }

// device status codes (LdmJob::status)
#define LDM_ST_OK 0
#define LDM_ST_AXIS 1      // N_G or N_L above LDM_MAX_AX
#define LDM_ST_CAP 2       // N_G*N_L above the allocated LDM rows
#define LDM_ST_FOLD 4      // alias folding did not converge
#define LDM_ST_WIDTH 8     // non-positive broadening width

// Tabulated partition function of one isotopologue slot as a B-spline of
// Q(T) in scipy's splrep form: knots t[0..n), coefficients c[0..n-k-1),
// degree k; valid (as the reference's table allows) for T in [lo, hi].
#define LDM_QS_MAXK 256
#define LDM_QS_MAXDEG 5
struct LdmQSpline {
  double t[LDM_QS_MAXK], c[LDM_QS_MAXK];
  int n, k;
  double lo, hi;
};

// Q(T) from a partition-function B-spline, evaluated as FITPACK's splev
// does (scipy.interpolate.splev with ext=0; the reference interpolates its
// tabulated Q(T) with splrep/splev, radis/levels/partfunc_cdsd.py:84-117):
// the knot interval l is the largest l in [k, n-k-2] with t[l] <= x (so the
// end pieces extrapolate), the k+1 B-splines that are nonzero there come
// from fpbspl's Cox-de Boor recursion, and Q = sum_j c[l-k+j] B_j.  Shared
// by the device (k_finalize, k_qspline_eval) and the host (scale bound).
__host__ __device__ inline double qspline_eval(const LdmQSpline& S, double x) {
#pragma clang fp contract(off)
  const int k = S.k;
  int l = k, r = S.n - k - 2;
  while (l < r) {
    const int mid = (l + r + 1) >> 1;
    if (S.t[mid] <= x) l = mid;
    else r = mid - 1;
  }
  double h[LDM_QS_MAXDEG + 1], hh[LDM_QS_MAXDEG];
  h[0] = 1.0;
  for (int j = 1; j <= k; ++j) {
    for (int i = 0; i < j; ++i) hh[i] = h[i];
    h[0] = 0.0;
    for (int i = 1; i <= j; ++i) {
      const int li = l + i, lj = li - j;
      if (S.t[li] == S.t[lj]) {
        h[i] = 0.0;
        continue;
      }
      const double f = hh[i - 1] / (S.t[li] - S.t[lj]);
      h[i - 1] = h[i - 1] + f * (S.t[li] - x);
      h[i] = f * (x - S.t[lj]);
    }
  }
  double q = 0.0;
  for (int j = 0; j <= k; ++j) q = q + S.c[l - k + j] * h[j];
  return q;
}

struct LdmDb {
  const double* nu0;
  const float* S_ref;
  const float* El;
  const float* g_air;
  const float* g_self;
  const float* n_air;
  const float* n_self;   // may be null
  const float* delta;    // may be null
  const uint8_t* iso;    // may be null
  const float* dflt;     // device block {0.0f, then byte 1}: what a missing delta / iso column reads
  const float* evib_l;   // lower/upper-state vibrational energies (cm-1), may be null
  const float* evib_u;
  // Lorentzian-width candidates of every LDM_SKY_BLOCK-line block: for each of
  // the 4 monotonicity directions, the block's Pareto-maximal lines as
  // (g_air, g_self, n_air, n_self); list d of block b is
  // sky[sky_off[d*(sky_nb+1) + b] .. sky_off[d*(sky_nb+1) + b + 1]).  Null if
  // the database has none (then every line is scanned).
  const float4* sky;
  const int* sky_off;
  int sky_nb;
  // the whole list's Pareto sets, direction d: sky[gsky_off[d] .. gsky_off[d + 1])
  int gsky_off[5];
  int n_iso_slots;
  const int2* iso_rng;   // [sky_nb][n_iso_slots]: first/last line of each iso in the block (-1: none)
  const LdmQSpline* qs;  // [LDM_MAX_ISO] partition-function splines (slots in qs_mask), may be null
  unsigned qs_mask;
};

struct LdmJob {
  double wmin, delta_w, wstep;   // grid w[i] = wmin + i*delta_w
  double inv_delta_w;            // 1/delta_w (first guess of a grid index only)
  double grid_fast_hi;           // x - w[j] below this: j is x's interval and NumPy cannot round to j+1
  float inv_delta_f;             // 1/delta_w in FP32 (fraction of the common case)
  long long nv, v;               // N_v and V = N_v + 1 (written LDM length)
  long long L;                   // FFT length = n1*n2
  long long i0, i1;              // window line range
  double p_atm, x_self, x_air;   // pressure (atm), mole fractions
  double log_r;                  // log(Tref/T)
  double ratio_T;                // Tref/T (the base of the FP64 width powers)
  double c_boltz;                // -hc_k*(1/T - 1/Tref)
  double i_scale;                // number density * 2^-scale_log2
  double dopp[LDM_MAX_ISO];      // sqrt((2 Na k_b T ln2)/M_iso)
  double qratio[LDM_MAX_ISO];    // Q(Tref)/Q(T)
  float qi_f[LDM_MAX_ISO];       // qratio * i_scale in FP32 (used when i_fast)
  int i_fast;                    // 1: every qi_f is well inside FP32's range
  double dxG, dxL, fold_thresh;
  double shift_margin;           // max|delta_air|*p_atm + 1e-3*delta_w
  float c_stim_T, c_stim_R;      // -hc_k/T, -hc_k/Tref (stimulated emission, FP32)
  float log_r_f;                 // log(Tref/T), FP32
  float pa_air_f, pa_self_f;     // p_atm*(1-x), p_atm*x, FP32
  float lgc_f[LDM_MAX_ISO];      // log2(2*dopp_iso/c_CGS): log2 FWHM_G = log2(nu0) + lgc
  float laG0_f, inv_dxG_f;       // first G-axis node (log2) and ln2/dxG, FP32 (set by k_finalize)
  float laL0_f, inv_dxL_f;       // same for the L axis (set by k_finalize)
  int fp64_widths;               // 1: log-width positions always in FP64
  int min_rms;                   // 1: min-RMS weights
  int noneq;                     // 1: two-temperature (Tvib, Trot) line strengths
  int device_q;                  // 1: qratio/qi_f of the database's spline slots computed by k_finalize
  double Tref;                   // reference temperature of S_ref (device Q ratio)
  double c_vib, c_rot, c_ref;    // -hc_k/Tvib, -hc_k/Trot, hc_k/Tref
  float mr_dv, mr_dxG, mr_dxL;   // wstep, dxG, dxL of the min-RMS corrections
  float mr_c1gg, mr_c2gg, mr_c1lg, mr_c2lg;
  float out_scale;               // 1/(L*wstep)
  float inv_Lws;                 // 1/(L*wstep) (frequency step of the FT grid)
  float half_inv_ws;             // 1/(2*wstep)
  int single_iso;                // 1: every line is iso slot 1 and the G axis comes from the window ends
  int n1, n2;                    // four-step split
  int nrad1, nrad2;
  int rad1[LDM_MAX_RAD], rad2[LDM_MAX_RAD];
  int bA, bB;                    // columns per workgroup in passes A/C and pass B
  int m_cap, tile_w, n_tiles;
  int fill_rows, fill_groups;    // LDM rows per fill row group, groups per tile (1: the tile holds every row)
  long long ldm_stride;          // elements per LDM row
  float* ldm;                    // [ceil(m_cap/2)][ldm_stride] float2: rows 2p, 2p+1 interleaved
  float2* ybuf;                  // [(m_cap+1)/2][L]
  float2* bbuf;                  // [L]
  float2* bbuf2;                 // [L] pass B's second share of the pairs (null: not split); pass C adds it
  float* out;                    // [nv]
  // derived outputs (null = not wanted): absorbance, transmittance, emissivity, radiance
  float *d_abs, *d_trans, *d_emis, *d_rad;
  double path_length, Tgas;
  int derived, scale_log2;
  int* nfold;                    // [m_cap]
  float* row_a;                  // [m_cap] (pi*wG_l)^2/(4 ln2)
  float* row_b;                  // [m_cap] pi*wL_m
  float* row_inv_f0;             // [m_cap] 1/F_lm(0)
  const float2* tw1;             // exp(-2 pi i e/n1), e < n1
  const float2* tw2;             // exp(-2 pi i e/n2), e < n2
  const float2* twl_hi;          // W_L^(h*1024)
  const float2* twl_lo;          // W_L^(e), e < 1024
  const float2* tw_wrap;         // exp(-2 pi i e/LDM_WRAP_LC), e < LDM_WRAP_LC (the image correction's transform)
  unsigned long long* ax_part;   // [n_ax_part][4]: per-item width extremes (wl min, wl max, wg min, wg max) as bits
  float* wrap_h;                 // [wrap_rows][wrap_k]: image correction of the lineshape FT at bins kk < wrap_k, per L-axis node
  int wrap_k;                    // 0: the transform length is the period (L = 2 N_v, or the request asked for L's own period)
  int wrap_rows;                 // L-axis nodes the table has room for
  long long wrap_ls;             // the period the spectrum is computed for (2 N_v)
  int n_ax_part;
  int sky_whole;                 // 1: the window is a single-isotopologue list's every line: k_finalize takes
                                 // the Lorentzian width extremes from the list's Pareto sets (no k_axis items)
  // reduction targets and status, initialised by the host every run
  unsigned long long wl_min_bits, wl_max_bits, wg_min_bits, wg_max_bits;
  int status;
  int pad0_;
};

// Device-computed LDM axes of one request (written by k_finalize).
struct LdmAxes {
  int ng, nl, m, p;                          // N_G, N_L, rows M = N_G*N_L, pairs P = ceil(M/2)
  double laG[LDM_MAX_AX], slG[LDM_MAX_AX];   // log of G-axis nodes, 1/(la[j+1]-la[j])
  double laL[LDM_MAX_AX], slL[LDM_MAX_AX];
};

// Per-request results copied back to the host after every run.
struct LdmResult {
  int ng, nl, status, pad_;
  double wg_min, wg_max, wl_min, wl_max;  // FWHM extremes the axes were built from
};
