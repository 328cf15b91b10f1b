/*
 * radis_ldm.h — C ABI of the MI355X-native DIT/LDM spectral-synthesis path.
 *
 * Plain C: pointers and sizes only, no C++ or torch types.  Every function
 * returns 0 on success and a negative RADIS_LDM_E* code on failure; the
 * message of the last failure on the calling thread is returned by
 * radis_ldm_last_error().  No exception crosses this boundary.
 *
 * Object model (mirrors the reference's gpuApp life cycle,
 * radis/gpu/gpu.py:35-380):
 *
 *   ctx    one HIP device                         (GPUApplication, gpu.py:102)
 *   db     a device-resident, nu0-sorted line list (database_d upload, gpu.py:190-203)
 *   batch  one or more spectrum requests over a db with their device
 *          workspace, job table and optional captured hipGraph
 *          (command buffer recording, gpu.py:255-284; iterate 302-370)
 *
 * Semantics follow RADIS's CPU LDM-fft path (SpectrumFactory.eq_spectrum
 * with broadening_method='fft', optimization='simple'): p in bar converted
 * to atm for widths and shifts, exact per-request LDM axes from the lines'
 * min/max widths, np.interp grid indices (bit-exact), alias folding of the
 * Voigt Fourier templates.  See the library's design notes for every
 * departure from the reference GPU shaders.
 */
#ifndef RADIS_LDM_H
#define RADIS_LDM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RADIS_LDM_ABI_VERSION 6

enum {
    RADIS_LDM_OK = 0,
    RADIS_LDM_EINVAL = -1,   /* invalid argument (shape, range, unsorted nu0 ...) */
    RADIS_LDM_EHIP = -2,     /* HIP runtime error */
    RADIS_LDM_ENOMEM = -3,   /* device allocation failed or exceeds the byte budget */
    RADIS_LDM_ERANGE = -4,   /* grid or LDM exceeds a compiled limit (LDM axes of more than 128
                                nodes, transforms above 8192 x 8192 points, 2^29 lines) */
    RADIS_LDM_EDEVICE = -5,  /* device-side check failed (see radis_ldm_batch_status) */
};

typedef struct radis_ldm_ctx radis_ldm_ctx;
typedef struct radis_ldm_db radis_ldm_db;
typedef struct radis_ldm_batch radis_ldm_batch;

/* Host-side line list, structure of arrays.  nu0 must be sorted ascending.
 * Optional columns may be NULL.  Replaces the database arrays handed to
 * gpuApp (radis/gpu/gpu.py:36-54, 190-203): v0, da, na, S0, El, gamma_arr,
 * iso, Mm_arr.  Intensities are tabulated S(Tref) as used by the CPU path
 * (radis/lbl/base.py:2664-2674). */
typedef struct {
    const double *nu0;        /* [n] line centre, cm-1 (sorted ascending)            */
    const float *S_ref;       /* [n] line intensity at Tref, cm/molecule               */
    const float *El;          /* [n] lower-state energy, cm-1                          */
    const float *g_air;       /* [n] air-broadening HWHM at Tref, cm-1/atm             */
    const float *g_self;      /* [n] self-broadening HWHM at Tref, cm-1/atm            */
    const float *n_air;       /* [n] temperature exponent of g_air                    */
    const float *n_self;      /* [n] or NULL: temperature exponent of g_self (NULL: n_air) */
    const float *delta;       /* [n] or NULL: air pressure shift, cm-1/atm (NULL: none) */
    const uint8_t *iso;       /* [n] or NULL: isotopologue slot (NULL: all slot 1)      */
    int64_t n;                /* number of lines                                        */
    const double *molar_mass; /* [n_iso_slots] g/mol, indexed by iso (slot 0 unused)    */
    int32_t n_iso_slots;      /* <= RADIS_LDM_MAX_ISO                                   */
    const float *Evib_l;      /* [n] or NULL: vibrational energy of the lower state, cm-1
                                 (0 <= Evib_l <= El); needed by RADIS_LDM_NONEQ requests */
    const float *Evib_u;      /* [n] or NULL: vibrational energy of the upper state, cm-1
                                 (0 <= Evib_u <= El + nu0); with Evib_l                   */
} radis_ldm_lines;

#define RADIS_LDM_MAX_ISO 16

/* One spectrum request.  Replaces the per-iteration state of gpuApp.iterate
 * (radis/gpu/gpu.py:302-357, params.py:120-155) and the grid of
 * SpectrumFactory (radis/lbl/factory.py:2560-2617). */
typedef struct {
    double wmin, wmax, wstep; /* grid = arange(wmin, wmax+wstep, wstep); lines with wmin<=nu0<=wmax */
    double Tgas;              /* K                                                      */
    double pressure_bar;      /* bar                                                    */
    double mole_fraction;     /* self mole fraction; air = 1 - x                        */
    double Tref;              /* K (reference temperature of S_ref and widths)           */
    double dxG, dxL;          /* LDM log-steps (RADIS Parameters.dxG/dxL)               */
    double folding_thresh;    /* alias folding threshold (RADIS default 1e-6)           */
    double q_ratio[RADIS_LDM_MAX_ISO]; /* Q(Tref)/Q(Tgas) per iso slot (host evaluates Q) */
    int64_t fft_len;          /* 0 = planner choice: 2*N_v when it has a supported split,
                                 else the smallest supported L with room for the
                                 circular-image correction (see RADIS_LDM_PERIOD_L) */
    int64_t n_v;              /* 0 = len(arange(wmin, wmax+wstep, wstep)); else the grid
                                 is wmin + i*((wmin+wstep)-wmin), i < n_v (gpuApp's Nv) */
    double line_lo, line_hi;  /* if line_hi > line_lo: lines with line_lo<=nu0<=line_hi,
                                 otherwise wmin<=nu0<=wmax (lbl/loader.py:2707)          */
    int32_t flags;            /* RADIS_LDM_FP64_WIDTHS: log-width axis positions and
                                 fractions always in FP64 (default: FP32 with an exact
                                 FP64 recomputation next to axis nodes);
                                 RADIS_LDM_GENERIC_FFT: run-time-sized FFT passes even
                                 where compile-time-sized ones exist;
                                 RADIS_LDM_MIN_RMS: min-RMS LDM weights
                                 (optimization="min-RMS", lbl/broadening.py:2265-2303)
                                 instead of the simple ones (2305-2318)                  */
    int32_t fft_n1;           /* 0 = planner; else the four-step split L = fft_n1 * n2
                                 (16 | fft_n1 <= 1024, 7-smooth cofactor)                */
    int32_t tile_w;           /* 0 = planner; else the fill tile width (power of 2)      */
    double Tvib, Trot;        /* RADIS_LDM_NONEQ: vibrational and rotational temperatures
                                 (K) of the level populations; Tgas is then the
                                 translational temperature of the widths, and q_ratio
                                 holds Q(Tref)/Q_neq(Tvib, Trot) per iso slot             */
} radis_ldm_request;

#define RADIS_LDM_FP64_WIDTHS 1
#define RADIS_LDM_GENERIC_FFT 2
#define RADIS_LDM_MIN_RMS 4
/* Two-temperature line strengths (Boltzmann vibrational and rotational
 * distributions, radis/lbl/base.py:3479-3530 populations and 3937-3947
 * calc_linestrength_noneq); the line list must carry Evib_l and Evib_u. */
#define RADIS_LDM_NONEQ 8
/* Validation: find the width extremes by scanning every window line instead
 * of the per-block candidates (same result, bit for bit). */
#define RADIS_LDM_SCAN_AXIS 16
/* Q(Tref)/Q(Tgas) of every isotopologue slot with lines from the database's
 * partition-function splines (radis_ldm_db_set_qspline), evaluated on the
 * device; q_ratio is then ignored for those slots.  Tgas and Tref must lie
 * in each table's range.  Equilibrium requests only. */
#define RADIS_LDM_DEVICE_Q 32
/* The spectrum is the circular convolution at the transform length fft_len
 * (the reference GPU's semantics at next_fast_len_even(2*N_v),
 * radis/gpu/gpu.py:22-29, 133-154) instead of RADIS's CPU result at
 * L = 2*N_v (radis/lbl/broadening.py:1900-1902, 2340-2346).  Without this
 * flag a transform longer than 2*N_v has its circular images corrected
 * (image-correction items in the fill launch), so the output is the L = 2*N_v
 * result whatever fft_len is; an
 * explicit fft_len must then be 2*N_v itself or leave a gap the correction
 * can span (about 0.5 % of L). */
#define RADIS_LDM_PERIOD_L 64

/* Planned sizes of a request (filled by radis_ldm_batch_create / _info). */
typedef struct {
    int64_t n_v;      /* grid points (output length)              */
    int64_t fft_len;  /* L                                         */
    int32_t n1, n2;   /* four-step split L = n1*n2                */
    int64_t line_begin, line_end; /* window line range in the db   */
    int32_t m_cap;    /* LDM rows allocated (upper bound of N_G*N_L) */
    int32_t tile_w;   /* fill tile width in grid points            */
    int32_t scale_log2; /* output = stored value * 2**scale_log2   */
    int32_t n_g, n_l; /* device-computed LDM axis sizes (valid after a run has completed) */
    int32_t status;   /* device status of the last completed run (0 = ok) */
    int32_t compiled_fft; /* 1: the FFT passes are the compile-time-sized kernels for (n1, n2) */
    double wg_min, wg_max, wl_min, wl_max; /* Gaussian / Lorentzian FWHM extremes over the window
                                 that the LDM axes of the last completed run start from
                                 (radis/lbl/broadening.py:1823-1837)                       */
    int64_t axis_scanned;     /* window lines the width reduction reads in full          */
    int64_t axis_candidates;  /* per-block width candidates it reads (16 B each) instead  */
} radis_ldm_request_info;

/* --- context -------------------------------------------------------------- */
/* Replaces GPUApplication.__init__ device selection (vulkan_compute_lib.py:26-72). */
int radis_ldm_ctx_create(int device, radis_ldm_ctx **out);
void radis_ldm_ctx_destroy(radis_ldm_ctx *ctx);
/* Device-memory budget for workspace admission (bytes; 0 = no limit). */
int radis_ldm_ctx_set_budget(radis_ldm_ctx *ctx, int64_t bytes);
int radis_ldm_ctx_bytes_in_use(const radis_ldm_ctx *ctx, int64_t *bytes);

/* --- line database -------------------------------------------------------- */
/* Replaces the database_d upload of gpuApp.__init__ (gpu.py:174-203).
 * At most 2^29 lines (536 M) per database; larger lists are split into
 * several databases (RADIS's own chunking of the line list does the same). */
int radis_ldm_db_upload(radis_ldm_ctx *ctx, const radis_ldm_lines *lines, radis_ldm_db **out);
void radis_ldm_db_destroy(radis_ldm_db *db);
int radis_ldm_db_size(const radis_ldm_db *db, int64_t *n_lines);

/* Tabulated partition function Q(T) of isotopologue slot iso as a B-spline
 * in scipy's splrep form: n_knots knots (non-decreasing, at most 256),
 * n_knots - degree - 1 coefficients, degree 1..5, valid for T in
 * [T_lo, T_hi].  The reference builds the same spline from a Q(T) table
 * (PartFuncCO2_CDSDtab: splrep, splev, 5 % range margin,
 * radis/levels/partfunc_cdsd.py:84-117) and hands gpuApp host Q callables
 * (radis/lbl/factory.py:1188-1199, radis/gpu/params.py:154-155); with a
 * spline per slot, requests flagged RADIS_LDM_DEVICE_Q need no host Q.
 * Synchronous; do not call while a run that reads the database is in
 * flight. */
int radis_ldm_db_set_qspline(radis_ldm_db *db, int iso, const double *knots, int32_t n_knots, const double *coefs,
                             int32_t degree, double T_lo, double T_hi);
/* Q(T) of slot iso at n temperatures: T and Q are device FP64 arrays
 * (splev semantics, extrapolating past the knots).  Enqueued on stream. */
int radis_ldm_qspline_eval(const radis_ldm_db *db, int iso, const double *T, int64_t n, double *Q, void *stream);

/* --- planning helpers ----------------------------------------------------- */
/* Grid length of arange(wmin, wmax+wstep, wstep) and the planner's FFT length. */
int radis_ldm_plan_grid(double wmin, double wmax, double wstep, int64_t *n_v, int64_t *fft_len);

/* Device workspace a request would need in a batch of its own (bytes), and
 * its planned sizes, without allocating anything: admission control by HBM
 * budget in the serving loop. */
int radis_ldm_request_bytes(const radis_ldm_db *db, const radis_ldm_request *req, int64_t *bytes,
                            radis_ldm_request_info *info);

/* --- batches -------------------------------------------------------------- */
/* Plan n requests over db and allocate their workspace.  Replaces gpuApp's
 * buffer set-up and command recording (gpu.py:205-297).  A request's
 * conditions may later be changed with radis_ldm_batch_set_request as long
 * as the grid (wmin, wmax, wstep, fft_len) is unchanged; the workspace grows
 * by 1.5x when the new LDM size needs it (vulkan_compute_lib.py:1308-1323). */
int radis_ldm_batch_create(radis_ldm_ctx *ctx, const radis_ldm_db *db, int n,
                           const radis_ldm_request *reqs, radis_ldm_batch **out);
int radis_ldm_batch_set_request(radis_ldm_batch *b, int i, const radis_ldm_request *req);
/* Enqueue the whole pipeline for every request on `stream` (a hipStream_t,
 * NULL = the null stream).  out_dev[i] is a device pointer to n_v floats.
 * Nothing is synchronised; the call allocates nothing in steady state.
 * Replaces gpuApp.iterate's run() (gpu.py:352-357). */
int radis_ldm_batch_run(radis_ldm_batch *b, float *const *out_dev, void *stream);
/* Capture the pipeline into a hipGraph bound to out_dev and replay it.
 * Conditions changed with _set_request are picked up by later replays. */
int radis_ldm_batch_graph_capture(radis_ldm_batch *b, float *const *out_dev, void *stream);
int radis_ldm_batch_graph_replay(radis_ldm_batch *b, void *stream);
/* Sizes and device results of request i.  n_g/n_l/status are read back
 * from the device (synchronises `stream`).  Replaces gpuApp.get_griddims
 * (gpu.py:372-374). */
int radis_ldm_batch_info(radis_ldm_batch *b, int i, radis_ldm_request_info *info, void *stream);
int radis_ldm_batch_workspace_bytes(const radis_ldm_batch *b, int64_t *bytes);
void radis_ldm_batch_destroy(radis_ldm_batch *b);

/* Device time per stage (axis, finalize, tile_bounds, fill, fft_a, fft_b,
 * fft_c, wrap) in milliseconds, averaged over the runs made with timing enabled
 * since the previous call; each stage is bracketed by start/stop events
 * (hipExtLaunchKernelGGL) on the launch stream.  A stage whose kernel did not
 * run (tile bounds reused) reports 0; the image-correction items (wrap) run
 * inside the fill launch, so wrap reports 0 and fill includes them.
 * Replaces GPUApplication.get_timestamps (vulkan_compute_lib.py:973-1024). */
#define RADIS_LDM_N_STAGES 8
int radis_ldm_batch_enable_timing(radis_ldm_batch *b, int enable);
int radis_ldm_batch_stage_ms(radis_ldm_batch *b, float *ms_out);

/* Derived spectral quantities written by the last pass next to abscoeff
 * (device pointers of n_v floats, NULL = not wanted), with the same
 * arithmetic as radis/lbl/factory.py:1289-1297 and
 * radis/spectrum/equations.py:21-44 (radiance from
 * radis/phys/blackbody.py:35-74), evaluated in FP64 per grid point:
 *   absorbance    = abscoeff * path_length   (stored with the request's
 *                   2**-scale_log2 factor, like abscoeff)
 *   emissivity    = -expm1(-absorbance)
 *   transmittance = 1 - emissivity
 *   radiance      = emissivity * Planck(1e7/nu nm, Tgas)   mW/cm2/sr/nm
 * The reference computes these on the host after the GPU run. */
typedef struct {
    float *absorbance, *transmittance, *emissivity, *radiance;
    double path_length; /* cm */
} radis_ldm_derived;

/* Set (or clear with NULL) request i's derived outputs for the following
 * runs and graph replays.  Replaces the host post-processing of
 * eq_spectrum_gpu (radis/lbl/factory.py:1284-1297). */
int radis_ldm_batch_set_derived(radis_ldm_batch *b, int i, const radis_ldm_derived *d);

/* --- diagnostics ---------------------------------------------------------- */
/* Per-line values the fill kernel scatters, for every line of request i's
 * window (line_end - line_begin entries, device pointers), computed with the
 * LDM axes of the last run: grid index k0 and fraction tv, log-wG index l0
 * and fraction tg, log-wL index m0 and fraction tl, and the intensity
 * I = S(T) * number density * 2**-scale_log2.  Counterpart of the index and
 * weight arrays of _apply_lineshape_LDM (radis/lbl/broadening.py:2253-2318). */
int radis_ldm_batch_line_terms(radis_ldm_batch *b, int i, int64_t *k0, int32_t *l0, int32_t *m0, float *tv,
                               float *tg, float *tl, float *I, void *stream);
/* Copy request i's LDM as filled by the last run into dst (device,
 * row-major [n_g*n_l][n_v+1], rows r = l*N_L + m); the counterpart of the
 * LDM array of radis/lbl/broadening.py:2340-2346 (transposed). */
int radis_ldm_batch_copy_ldm(radis_ldm_batch *b, int i, float *dst, void *stream);

/* --- no-allocation entry point --------------------------------------------- */
/* Device results of one request: LDM axis sizes, device status (0 = ok, bits
 * as radis_ldm_request_info.status) and the FWHM extremes the axes start
 * from. */
typedef struct {
    int32_t n_g, n_l, status, pad_;
    double wg_min, wg_max, wl_min, wl_max;
} radis_ldm_result;

/* Bytes of the per-device scratch array of radis_ldm_run_noalloc. */
#define RADIS_LDM_NOALLOC_BYTES (256ll << 20)

/* One request run entirely in a fixed per-device scratch array: no device
 * or host allocation, no host-to-device copy from host memory (the job table
 * travels as a kernel argument), so it can run under hipGraph capture and in
 * loops that must not allocate.  It has no counterpart in the reference; it
 * is the variant of radis_ldm_batch_create + radis_ldm_batch_run for one
 * request, for callers that cannot allocate.  Everything goes on `stream`
 * (a hipStream_t, NULL = null stream); nothing is synchronised.
 *   out_dev     device array of n_v floats (abscoeff times 2**-scale_log2)
 *   result_dev  NULL, or a device radis_ldm_result the run fills
 *   info        NULL, or host memory for the planned sizes (scale_log2,
 *               n_v, fft_len, ...; n_g, n_l and status are in result_dev)
 * A request whose workspace exceeds RADIS_LDM_NOALLOC_BYTES is refused
 * (RADIS_LDM_ERANGE, nothing launched).  Every call on one device shares
 * that device's scratch array: calls whose runs overlap on a device (other
 * threads, or streams not ordered with each other) race, and a run reads
 * nothing a previous call left there.  Derived quantities and timing are not
 * offered. */
int radis_ldm_run_noalloc(radis_ldm_ctx *ctx, const radis_ldm_db *db, const radis_ldm_request *req,
                          float *out_dev, radis_ldm_result *result_dev, radis_ldm_request_info *info,
                          void *stream);

const char *radis_ldm_last_error(void);
int radis_ldm_abi_version(void);

/* --- instrumental slit ------------------------------------------------------ */
/* Convolution with a slit function sampled on the spectrum grid, on the
 * device: out[i] = scale * sum_j slit[j] * in[i + (k-1)/2 - j], with in[]
 * zero outside [0, n) -- scipy.signal.oaconvolve(in, slit, mode='same') *
 * wstep as in convolve_with_slit (radis/tools/slit.py:477-757; pass the grid
 * step as scale).  in, slit and out are device arrays of n, k and n floats;
 * out must not overlap in or slit.  The slit is built and cropped by the caller
 * ('valid' mode: radis/tools/slit.py:1005-1092).  Replaces the host
 * convolution behind Spectrum.apply_slit (radis/spectrum/spectrum.py:3127),
 * which the reference's GPU iterate leaves undone (radis/gpu/gpu.py:308-311). */
int radis_ldm_slit_convolve(const float *in, int64_t n, const float *slit, int32_t k, double scale, float *out,
                            void *stream);

#ifdef __cplusplus
}
#endif
#endif /* RADIS_LDM_H */
