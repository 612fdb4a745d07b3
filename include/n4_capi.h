/*
 * n4_capi.h -- C ABI of the MI355X (gfx950) N4 bias-field-correction library
 * (libn4amd.so).  This is the drop-in boundary: plain pointers and sizes,
 * device pointers unless a name says host, status codes instead of exit().
 *
 * Layout of every volume: float32, C order [x][y][z] with z fastest, x the
 * slowest axis (the reference's device layout, cudaN4.cu:476,562; its CLI
 * passes ITK (size[2], size[1], size[0]) as (fx, fy, fz), main.cpp:153-155).
 *
 * Each entry point below names the reference interface it replaces.
 */
#ifndef N4_CAPI_H
#define N4_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define N4_MAX_LEVELS 16
#define N4_MAX_BINS 512

typedef enum n4_status {
    N4_OK = 0,
    N4_ERR_INVALID_ARG = 1, /* bad size, null pointer, parameter out of range */
    N4_ERR_HIP = 2,         /* a HIP runtime call failed */
    N4_ERR_NO_MEMORY = 3,   /* device allocation failed */
    N4_ERR_UNSUPPORTED = 4, /* valid request this build does not implement */
    N4_ERR_INTERNAL = 5
} n4_status;

typedef enum n4_mask_kind {
    N4_MASK_F32 = 0,  /* float32, nonzero = foreground (cudaN4.h:346) */
    N4_MASK_U8 = 1,   /* uint8, nonzero = foreground */
    N4_MASK_NONE = 2  /* every voxel foreground (main.cpp:137-140) */
} n4_mask_kind;

typedef enum n4_precision {
    N4_PREC_FP32 = 0,       /* FP32 direct fit / reconstruction (the drop-in) */
    N4_PREC_BF16 = 1,       /* separable MFMA contractions, BF16 in, FP32 acc */
    N4_PREC_SPLIT_BF16 = 2  /* hi+lo split BF16 on MFMA, FP32 accumulate */
} n4_precision;

/* Flags for n4_params.flags */
#define N4_FLAG_FIXED_ITERATIONS 0x1u /* ignore the convergence threshold */
/* F32 masks are confidence weights (the reference takes binary masks only,
 * cudaN4.h:346): a voxel is foreground where its weight is > 0 and finite;
 * its fit contributions (numerator and denominator, N4.py:230-236) are
 * scaled by the weight, as ITK/ANTs N4 uses a confidence image; histogram
 * and convergence measure stay unweighted over the foreground.  With a
 * shrink factor the working voxel's weight is the mean weight of the
 * foreground voxels of its block.  U8 and absent masks stay binary.  The
 * fit and evaluation then run the FP32 kernels whatever the precision. */
#define N4_FLAG_WEIGHTED_MASK 0x2u

/* Parameters.  Replaces struct N4Param (cudaN4.h:24-94); N4.py:276-279,
 * 292, 307-311 for the per-level iteration list and fixed constants.
 * Floating-point parameters are double so that 0.15 / 0.01 / 0.001 are the
 * same numbers N4.py uses. */
typedef struct n4_params {
    uint32_t number_of_histogram_bins;          /* cudaN4.h:30  (200)   */
    double wiener_filter_noise;                 /* cudaN4.h:38  (0.01)  */
    double bias_field_fwhm;                     /* cudaN4.h:44  (0.15)  */
    uint32_t ncpt[3];                           /* cudaN4.h:60-62 (4,4,4) x,y,z */
    uint32_t number_of_fitting_levels;          /* cudaN4.h:71  (4)     */
    uint32_t max_iterations[N4_MAX_LEVELS];     /* cudaN4.h:75  (50)    */
    double convergence_threshold;               /* cudaN4.h:84  (0.001) */
    double low_value;                           /* cudaN4.h:85  (10)    */
    uint32_t shrink_factor;                     /* working grid = ceil(full/s) (1) */
    double fit_eps;                             /* N4.py:236 denominator eps (1e-3) */
    uint32_t precision;                         /* n4_precision (FP32)  */
    uint32_t flags;                             /* N4_FLAG_*            */
} n4_params;

/* Per-volume run record (no counterpart in the reference, which only prints
 * "Level %d, iter %d", cudaN4.cu:1530). */
typedef struct n4_report {
    uint32_t iterations[N4_MAX_LEVELS];   /* iterations executed per level */
    double last_convergence[N4_MAX_LEVELS];
    uint64_t n_mask_full;                 /* foreground voxels after threshold */
    uint64_t n_mask_work;                 /* foreground voxels of the working grid */
    uint32_t work_dims[3];
    uint32_t final_ncpt[3];
    float elapsed_ms;                     /* device time of the whole call */
} n4_report;

/* One volume of a batch (no counterpart in the reference). */
typedef struct n4_volume {
    const float *d_im;           /* input image, dims[0]*dims[1]*dims[2] */
    const void *d_mask;          /* mask (kind below), may be NULL with N4_MASK_NONE */
    uint32_t mask_kind;          /* n4_mask_kind */
    uint32_t dims[3];            /* fx, fy, fz */
    uint32_t shrink_factor;      /* 0 = use n4_params.shrink_factor */
    float *d_im_normalized;      /* output, required */
    float *d_lattice;            /* output final lattice, nullable */
    float *d_bias_field;         /* output exp(log bias), nullable */
    float *d_log_bias_field;     /* output log bias, nullable */
} n4_volume;

typedef struct n4_handle n4_handle;

/* ---- library / handle ------------------------------------------------- */
const char *n4_version(void);
const char *n4_status_string(n4_status s);
/* Message of the last failed call on this thread ("" if none). */
const char *n4_last_error(void);
/* Default parameters: N4Param's defaults (cudaN4.h:24-94), shrink 1, eps 1e-3. */
n4_status n4_params_default(n4_params *p);
/* Final-level control points per axis: (ncpt-3)*2^(levels-1)+3
 * (main.cpp:83-92, cudaN4.cu:1778-1788). */
n4_status n4_final_ncpt(const n4_params *p, uint32_t out_ncpt[3]);
/* Handle = device + workspace + stream-ordered state.  Replaces
 * N4Data::assign (cudaN4.h:132-174): no FFTW plans, nothing on the host. */
n4_status n4_create(n4_handle **h, int device);
/* Replaces N4Data::clean (cudaN4.h:177-193). */
n4_status n4_destroy(n4_handle *h);

/* ---- whole correction -------------------------------------------------- */
/* Replaces N4() (cudaN4.h:343-365, cudaN4.cu:1418-1747).  Inputs are not
 * modified (the reference thresholds d_im/d_mask in place, cudaN4.cu:1459).
 * d_lattice (nullable) receives the final-level lattice, n4_final_ncpt.
 * stream: a hipStream_t, or NULL for the null stream.  The call returns
 * after the work completes (the report needs the iteration counts). */
n4_status n4_correct(n4_handle *h, const float *d_im, const void *d_mask,
                     uint32_t mask_kind, uint32_t fx, uint32_t fy, uint32_t fz,
                     const n4_params *p, float *d_im_normalized,
                     float *d_lattice, float *d_bias_field,
                     float *d_log_bias_field, n4_report *report, void *stream);

/* n4_correct without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and results as n4_correct; the plan and working
 * buffers live in a file-scope device array of 64 MiB per device (held
 * until the process ends) instead of the handle's arena, so the working
 * grid is limited (about 6.5 M working voxels at the default parameters);
 * larger requests fail with N4_ERR_NO_MEMORY.  It runs the step kernels on
 * `stream` with every step enqueued up front (no progress words in pinned
 * memory, no events, no graph instantiation; kernel timing is ignored) and
 * returns after the work completes.  Calls that overlap on one device race
 * on the shared array.  Call on the handle's device. */
n4_status n4_correct_noalloc(n4_handle *h, const float *d_im, const void *d_mask,
                             uint32_t mask_kind, uint32_t fx, uint32_t fy, uint32_t fz,
                             const n4_params *p, float *d_im_normalized,
                             float *d_lattice, float *d_bias_field,
                             float *d_log_bias_field, n4_report *report, void *stream);

/* Replaces testN4 (cudaN4.h:367-387, cudaN4.cu:1751-1847): host in/out. */
n4_status n4_correct_host(n4_handle *h, const float *h_im, const float *h_mask,
                          uint32_t fx, uint32_t fy, uint32_t fz,
                          const n4_params *p, float *h_im_normalized,
                          float *h_lattice, float *h_bias_field,
                          n4_report *report);

/* Many volumes of uneven size in one launch per stage (no counterpart in
 * the reference).  reports: nvol entries or NULL. */
n4_status n4_correct_batch(n4_handle *h, const n4_volume *vols, uint32_t nvol,
                           const n4_params *p, n4_report *reports, void *stream);

/* Run records of the last correction call on this handle (n4_correct,
 * n4_correct_noalloc, n4_correct_host, n4_correct_batch), volume `volume`
 * of that call.  The reference prints "Level %d, iter %d" per iteration
 * (cudaN4.cu:1530) and drops the lattice (main.cpp:122); these keep both.
 *   n4_last_cv_trace: the convergence measurement (N4.py:121-126) of every
 *     iteration, level after level (report.iterations[l] values for level
 *     l); *count = the number of values, of which the first `capacity` are
 *     written.  Iterations after the first 4096 of a call are not recorded.
 *   n4_last_level_lattice: the FP32 lattice of fitting level `level` as it
 *     stood at the end of that level, [cx][cy][cz] with cz fastest,
 *     (ncpt[d] - 3) * 2^level + 3 control points along axis d; capacity in
 *     floats.
 * Both fail with N4_ERR_INVALID_ARG once the record is gone: any later call
 * on the handle (stage calls included) reuses its memory, and so does, for
 * a record of n4_correct_noalloc, any later allocation-free call on the
 * device.  Synchronous. */
n4_status n4_last_cv_trace(n4_handle *h, uint32_t volume, double *out,
                           uint32_t capacity, uint32_t *count);
n4_status n4_last_level_lattice(n4_handle *h, uint32_t volume, uint32_t level,
                                float *out, uint64_t capacity);

/* ---- stage entry points (device pointers, synchronous) ----------------- */
/* Masked min/max + fractional histogram (cudaN4.cu:703-731).
 * d_mask: uint8.  out_minmax_slope (host, 3 doubles), d_hist (nbins doubles,
 * device), d_bin_index (device int32 per voxel, -1 where unmasked; nullable). */
n4_status n4_histogram(n4_handle *h, const float *d_x, const uint8_t *d_mask,
                       uint64_t n, uint32_t nbins, double *out_minmax_slope,
                       double *d_hist, int32_t *d_bin_index);

/* Replaces sharpenImage (cudaN4.h:198-234): histogram, Wiener deconvolution
 * (in-LDS FP64 FFT, no FFTW) and E[u|v] remap.  d_after = 0 where unmasked.
 * d_etable (nullable, device, nbins doubles). */
n4_status n4_sharpen_image(n4_handle *h, const float *d_before,
                           const uint8_t *d_mask, uint64_t n, uint32_t nbins,
                           double wiener_noise, double fwhm, float *d_after,
                           double *d_etable);

/* Replaces FitBspline3D (cudaN4.h:313-325) with N4.py semantics
 * (lattice = num/(den+eps), N4.py:236).  d_mask: uint8. */
n4_status n4_fit_bspline3d(n4_handle *h, const float *d_field,
                           const uint8_t *d_mask, const uint32_t ncpt[3],
                           const uint32_t dims[3], double eps,
                           float *d_lattice);

/* Replaces EvaluateBspline3D (cudaN4.h:288-298). */
n4_status n4_evaluate_bspline3d(n4_handle *h, const float *d_lattice,
                                const uint32_t ncpt[3], const uint32_t dims[3],
                                float *d_out);

/* Replaces calculateConvergenceMeasurement (cudaN4.h:249-257); ddof 0
 * (N4.py:125).  Result written to *out (host). */
n4_status n4_convergence_measurement(n4_handle *h, const float *d_field1,
                                     const float *d_field2,
                                     const uint8_t *d_mask, uint64_t n,
                                     double *out);

/* The reference's float masks (nonzero = foreground, cudaN4.h:346) as the
 * uint8 masks the stage entry points take: d_out[i] = d_mask[i] != 0, on
 * the device (n values; synchronous).  No counterpart in the reference,
 * whose kernels read the float mask directly. */
n4_status n4_mask_to_u8(n4_handle *h, const float *d_mask, uint64_t n,
                        uint8_t *d_out);

/* Replaces upsampleLattice3D_gpu (cudaN4.h:269-276): (n+3) -> (2n+3). */
n4_status n4_upsample_lattice3d(n4_handle *h, const float *d_lattice,
                                const uint32_t ncpt[3], float *d_out);

/* Replace Reducer::reduce_{sum,max,min}_wrapper (reducer.h:7-9); FP64
 * accumulation, wave64, single launch, result to host. */
n4_status n4_reduce_sum(n4_handle *h, uint64_t n, const float *d_in, double *out);
n4_status n4_reduce_max(n4_handle *h, uint64_t n, const float *d_in, float *out);
n4_status n4_reduce_min(n4_handle *h, uint64_t n, const float *d_in, float *out);

/* ---- launch schedule and autotuner (no counterpart in the reference) --- */
/* How n4_correct / n4_correct_batch lay out their launches.  Two paths run
 * the level/iteration loop (cudaN4.cu:1517-1657):
 *   steps: the host enqueues identical steps of five launches (histogram,
 *     Wiener, fit, lattice update, evaluation) per volume group, each
 *     volume advancing its own state machine on the device;
 *   teams: one persistent launch in which a team of workgroups per volume
 *     runs the volume's whole loop, synchronising at team barriers (every
 *     workgroup of the launch is resident at once).
 * Both run the same per-tile arithmetic, so they give the same bits.
 * Results do not depend on the schedule except through the summation order
 * of the FP64 convergence partials (one partial per tile).  A field left 0 takes the
 * environment override (N4_GROUPS, N4_STEP_TILE_VOX, N4_GRAPH_STEPS) if set,
 * else the built-in default. */
#define N4_GRAPH_OFF 0xFFFFFFFFu
typedef struct n4_tuning {
    uint32_t groups;           /* volume groups, one stream each; default 4 for
                                  >= 8 volumes, 2 for >= 4, else 1 (max 8) */
    uint32_t step_tile_voxels; /* working voxels per workgroup tile of the
                                  step kernels (rounded to whole rows); 8192 */
    uint32_t graph_steps;      /* iteration steps per captured hipGraph,
                                  1..64; N4_GRAPH_OFF = direct launches */
    uint32_t eval_kernel;      /* field-increment kernel of the MFMA
                                  precisions: N4_EVAL_FP32 (default) or
                                  N4_EVAL_MFMA; FP32 runs use FP32 */
    uint32_t path;             /* N4_PATH_AUTO (default: the step kernels,
                                  the faster path on every workload measured);
                                  N4_PATH_STEPS; N4_PATH_TEAMS: the persistent
                                  team kernel (N4_ERR_UNSUPPORTED where the
                                  batch is not eligible) */
    uint32_t team_wgs_per_cu;  /* persistent path: workgroups per compute
                                  unit (0 = as many as fit, at most 4) */
    float ms;                  /* n4_autotune_batch: wall time of one batch */
} n4_tuning;
#define N4_EVAL_FP32 1u
#define N4_EVAL_MFMA 2u
#define N4_PATH_AUTO 0u
#define N4_PATH_STEPS 1u
#define N4_PATH_TEAMS 2u
/* The handle's schedule as set (0 fields = defaults). */
n4_status n4_get_tuning(n4_handle *h, n4_tuning *out);
/* Sets the handle's schedule; NULL restores the defaults. */
n4_status n4_set_tuning(n4_handle *h, const n4_tuning *t);
/* Times n4_correct_batch on these volumes (outputs are written) under each
 * candidate schedule -- groups {1,2,4,8} up to nvol, tiles {4096, 8192,
 * 16384} voxels, graphs {off, 8, 16} steps and, for the MFMA precisions,
 * both field-increment kernels -- keeps the fastest on the handle and
 * returns it in *best (ms = its best wall time). */
n4_status n4_autotune_batch(n4_handle *h, const n4_volume *vols, uint32_t nvol,
                            const n4_params *p, n4_tuning *best, void *stream);

/* ---- kernel timing (no counterpart in the reference) ------------------- */
/* Kernels of the correction, for n4_profile_*. */
typedef enum n4_kernel_id {
    N4_K_PREP = 0,    /* threshold + log + shrink (full resolution) */
    N4_K_LEVEL = 1,   /* level setup / lattice refinement / commit */
    N4_K_HIST = 2,    /* convergence decision + fractional histogram */
    N4_K_FIT_DEN = 3, /* fit denominator partials (once per level) */
    N4_K_FIT = 4,     /* remap + residual + fit numerator partials */
    N4_K_EVAL = 5,    /* field + convergence partials + logU update */
    N4_K_RECON = 6,   /* full-resolution field + normalisation */
    N4_K_WIENER = 7,  /* Wiener deconvolution + E table (one block/volume) */
    N4_K_LATTICE = 8, /* mode-x contraction + lattice update (stage calls) */
    N4_K_UPDATE = 9,  /* lattice update of a correction step */
    N4_K_TEAMS = 10,  /* persistent path: every volume's level/iteration loop */
    /* Phases of the persistent path, timed inside the launch by each team's
     * leader workgroup (device clock; total over the volumes' leaders,
     * launches = iteration steps).  Enabled with the N4_K_TEAMS bit. */
    N4_K_T_WAIT = 11,   /* team barrier waits */
    N4_K_T_REDUCE = 12, /* partial reduction and step decision */
    N4_K_T_HIST = 13,   /* histogram of the own tiles */
    N4_K_T_WIENER = 14, /* Wiener step (E table) */
    N4_K_T_FIT = 15,    /* fit of the own tiles */
    N4_K_T_UPDATE = 16, /* owned control points, increment planes */
    N4_K_T_EVAL = 17,   /* evaluation of the own tiles */
    N4_K_T_TRANS = 18,  /* level transitions (refinement, denominator) */
    N4_K_COUNT = 19
} n4_kernel_id;
/* Time every launch of the kernels in kernel_mask (bit = n4_kernel_id) with
 * a pair of HIP events recorded on the launch stream around it (step graphs
 * are not used meanwhile); clears the accumulators. 0 disables.  The
 * interval includes the event markers and the dispatch of the launch, so
 * for kernels of tens of microseconds it exceeds the execution time a
 * profiler reports. */
n4_status n4_profile_enable(n4_handle *h, uint32_t kernel_mask);
/* Accumulated device time and launch count of one kernel since enable. */
n4_status n4_profile_read(n4_handle *h, uint32_t kernel, double *total_ms, uint64_t *launches);

/* ---- device memory plumbing (for hosts without their own HIP runtime) -- */
n4_status n4_device_count(int *count);
n4_status n4_malloc(n4_handle *h, void **d_ptr, size_t bytes);
n4_status n4_free(n4_handle *h, void *d_ptr);
n4_status n4_memcpy_htod(n4_handle *h, void *d_dst, const void *h_src, size_t bytes);
n4_status n4_memcpy_dtoh(n4_handle *h, void *h_dst, const void *d_src, size_t bytes);
n4_status n4_memcpy_dtod(n4_handle *h, void *d_dst, const void *d_src, size_t bytes);
n4_status n4_memset(n4_handle *h, void *d_dst, int value, size_t bytes);
n4_status n4_synchronize(n4_handle *h);

#ifdef __cplusplus
}
#endif
#endif
