/*
 * n4_oracle.h -- CPU FP64 restatement of the reference N4 algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity reference for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links against it.
 *
 * Semantics of record: N4.py (reference file), run in FP64 throughout.
 * Deviations from N4.py (each one is a deliberate, documented choice):
 *   - N4.py:75 stores the histogram numerator as complex64 and numpy >= 2
 *     FFTs it in single precision; here it is FP64 like every other step.
 *   - N4.py:292 fixes the initial lattice at 4x4x4 and N4.py:307-311 fixes
 *     bins/noise/FWHM; here they come from n4o_params (same defaults as the
 *     reference's N4Param, cudaN4.h:24-94).
 *   - N4.py:152-154,161 clips with the x-axis count on every axis; here the
 *     clip bound is per axis (identical for isotropic lattices).
 *   - Shrink (absent from both reference implementations): block mean of the
 *     log intensity over the masked voxels of each s*s*s block; a working
 *     voxel is foreground when 2*n_masked >= n_block.  Shrink 1 reduces
 *     exactly to N4.py.
 *   - A constant masked image (max == min) makes N4.py raise; here the
 *     sharpening step returns the input unchanged (zero residual).
 *   - Confidence weights (n4o_params.weighted; the reference takes binary
 *     masks only, cudaN4.h:346): foreground where the weight is > 0 and
 *     finite; each foreground voxel's numerator and denominator terms of the
 *     fit (N4.py:230-236) are multiplied by its weight, the way ITK/ANTs N4
 *     passes a confidence image to its B-spline fit; histogram and
 *     convergence measure stay unweighted.  With a shrink, a working voxel's
 *     weight is the mean weight of its block's foreground voxels.  Weights
 *     all 1 reduce exactly to the binary case.
 *
 * Layout: C order [x][y][z], z fastest (cudaN4.cu:476,562).
 */
#ifndef N4_ORACLE_H
#define N4_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define N4O_MAX_LEVELS 16

typedef struct n4o_params {
    int32_t nbins;                 /* cudaN4.h:30   (200)   */
    double wiener_noise;           /* cudaN4.h:38   (0.01)  */
    double fwhm;                   /* cudaN4.h:44   (0.15)  */
    int32_t ncpt[3];               /* cudaN4.h:60-62 (4,4,4) */
    int32_t levels;                /* cudaN4.h:71   (4)     */
    int32_t max_iter[N4O_MAX_LEVELS]; /* cudaN4.h:75 / N4.py:278 (50) */
    double conv_threshold;         /* cudaN4.h:84   (0.001) */
    double low_value;              /* cudaN4.h:85   (10)    */
    int32_t shrink;                /* build-defined (1)     */
    double fit_eps;                /* N4.py:186     (1e-3)  */
    int32_t nthreads;              /* OpenMP threads for fit/evaluate (1) */
    int32_t weighted;              /* mask_f holds confidence weights (0)  */
} n4o_params;

typedef struct n4o_report {
    int32_t iterations[N4O_MAX_LEVELS];   /* iterations executed per level   */
    double last_cv[N4O_MAX_LEVELS];       /* last convergence value per level */
    int64_t n_mask_full;                  /* foreground voxels after threshold */
    int64_t n_mask_work;                  /* foreground working voxels         */
    int32_t work_dims[3];
    int32_t final_ncpt[3];
} n4o_report;

void n4o_default_params(n4o_params *p);

/* N4.py:174-181 */
void n4o_cubic_bspline(double t, double w[4]);

/* Per-axis span table, N4.py:192,210-217: span = domain/(ncpt-3);
 * b[i] = floor((i+0.5)/span); t = ((i+0.5) - b*span)/span; w = B(t). */
void n4o_axis_table(int32_t npos, double domain, int32_t ncpt,
                    int32_t *b, double *w /* npos*4 */);

/* Padded FFT length and offset, N4.py:31-37 (integer log2). */
void n4o_padded_size(int32_t nbins, int32_t *P, int32_t *offset);

/* In-place complex FFT, radix 2, n a power of two.  inverse=1 divides by n
 * (numpy.fft.ifft convention). data = interleaved re,im. */
void n4o_fft(double *data, int32_t n, int32_t inverse);

/* Masked min/max, N4.py:14. Returns number of masked voxels. */
int64_t n4o_masked_minmax(const double *x, const uint8_t *mask, int64_t n,
                          double *mn, double *mx);

/* Fractional histogram, N4.py:14-29.  H has nbins entries (overwritten).
 * idx_out (nullable) receives floor(cidx) per voxel, -1 where unmasked. */
void n4o_histogram(const double *x, const uint8_t *mask, int64_t n,
                   int32_t nbins, double bin_min, double slope,
                   double *H, int32_t *idx_out);

/* Wiener deconvolution + E table, N4.py:31-101.  E has nbins entries. */
void n4o_wiener_etable(const double *H, int32_t nbins, double bin_min,
                       double slope, double noise, double fwhm, double *E);

/* Full sharpenImage, N4.py:9-118.  Outputs sharpened (0 where unmasked);
 * optional H, E (nbins each) and bin_min/slope (nullable). */
void n4o_sharpen(const double *x, const uint8_t *mask, int64_t n,
                 int32_t nbins, double noise, double fwhm, double *out,
                 double *H_out, double *E_out, double *minmax_slope_out);

/* FitBspline3D, N4.py:186-237 (direct form). lattice: ncpt product. */
void n4o_fit_bspline3d(const double *field, const uint8_t *mask,
                       const int32_t dims[3], const int32_t ncpt[3],
                       double eps, int32_t nthreads, double *lattice);

/* Numerator and denominator of the fit (before division), for checks. */
void n4o_fit_numden(const double *field, const uint8_t *mask,
                    const int32_t dims[3], const int32_t ncpt[3],
                    int32_t nthreads, double *num, double *den);

/* The same with per-voxel confidence weights conf (nullable: all 1). */
void n4o_fit_numden_w(const double *field, const uint8_t *mask,
                      const double *conf, const int32_t dims[3],
                      const int32_t ncpt[3], int32_t nthreads, double *num,
                      double *den);

/* EvaluateBspline3D, N4.py:241-273, generalised to a domain length per axis
 * (domain[d] == dims[d] reproduces N4.py). */
void n4o_evaluate_bspline3d(const double *lattice, const int32_t ncpt[3],
                            const int32_t dims[3], const double domain[3],
                            int32_t nthreads, double *out);

/* CalculateConvergenceMeasurement, N4.py:121-126 (ddof 0). */
double n4o_convergence(const double *f1, const double *f2,
                       const uint8_t *mask, int64_t n);

/* UpsampleLattice3D, N4.py:132-165 (per-axis clip bounds). */
void n4o_upsample_lattice3d(const double *lattice, const int32_t ncpt[3],
                            double *out);

/* Threshold + log + shrink (build-defined shrink; s=1 is N4.py:280-287).
 * mask_f may be NULL (all foreground).  Outputs working-grid arrays. */
void n4o_threshold_log_shrink(const float *im, const float *mask_f,
                              const int32_t fdims[3], double low_value,
                              int32_t shrink, int32_t wdims[3],
                              double *imlog_w, uint8_t *mask_w,
                              int64_t *n_mask_full, int64_t *n_mask_work);

/* The same with mask_f read as confidence weights when weighted != 0;
 * conf_w (nullable) receives each working voxel's mean foreground weight
 * (0 for background working voxels; 1 where weighted == 0). */
void n4o_threshold_log_shrink_w(const float *im, const float *mask_f,
                                int32_t weighted, const int32_t fdims[3],
                                double low_value, int32_t shrink,
                                int32_t wdims[3], double *imlog_w,
                                uint8_t *mask_w, double *conf_w,
                                int64_t *n_mask_full, int64_t *n_mask_work);

/* Whole N4, N4.py:276-332 with the shrink extension.
 *   im, mask_f: full resolution float32, [x][y][z]. mask_f nullable.
 *   out_norm, out_logbias: full resolution (nullable).
 *   lattice_out: final-level lattice (nullable).
 *   cv_trace: optional, capacity cv_cap, filled in iteration order.
 * Returns 0 on success, nonzero on bad arguments. */
int32_t n4o_n4(const float *im, const float *mask_f, const int32_t fdims[3],
               const n4o_params *p, double *out_norm, double *out_logbias,
               double *lattice_out, double *cv_trace, int32_t cv_cap,
               n4o_report *rep);

#ifdef __cplusplus
}
#endif
#endif
