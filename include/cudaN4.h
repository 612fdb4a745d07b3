/*
 * cudaN4.h -- source-compatible C++ API of the reference's correction path,
 * implemented on the MI355X library (include/n4_capi.h).  A program written
 * against the reference header (cudaN4.h:24-387) compiles against this one
 * with the changes listed below and links against libn4compat.so.
 *
 * Differences from the reference header, all deliberate:
 *   - no CUDA or FFTW includes: N4Data holds an n4_handle instead of FFTW
 *     plans and host scratch (cudaN4.h:98-195); sharpenImage takes N4Data&
 *     in place of its 18 FFTW/scratch arguments (cudaN4.h:198-234);
 *   - N4() and the stage functions leave d_im, d_mask and param unchanged
 *     (the reference thresholds them in place, cudaN4.cu:1459, and writes
 *     the last convergence value into param.ConvergenceThreshold,
 *     cudaN4.cu:1643);
 *   - semantics follow N4.py: the fit divides by den + 1e-3 (N4.py:236),
 *     the convergence measure uses ddof 0 (N4.py:125) and the level loop
 *     stops once it is <= ConvergenceThreshold (N4.py:301-302);
 *   - extra fields at the end of N4Param: ShrinkFactor, FitEps,
 *     IterationsPerLevel (per-level maxima, N4.py:278), Precision,
 *     FixedIterations (run every level to its maximum, the reference's
 *     effective behaviour, cudaN4.cu:1643), WeightedMask (float masks are
 *     confidence weights; the reference takes binary masks, cudaN4.h:346).
 * Errors throw n4_error (std::runtime_error) instead of exiting.
 */
#ifndef CUDAN4_H
#define CUDAN4_H

#include <math.h>
#include <stdexcept>
#include <string>
#include <vector>

#include "n4_capi.h"

struct n4_error : std::runtime_error {
    explicit n4_error(const std::string& m) : std::runtime_error(m) {}
};

/* Parameters (cudaN4.h:24-94). */
struct N4Param {
    unsigned int NumberOfHistogramBins = 200;                                    /* cudaN4.h:30 */
    float exponent = ceil(log((float)NumberOfHistogramBins) / log(2.0)) + 1.;    /* cudaN4.h:31 */
    unsigned int paddedHistogramSize = (unsigned int)(pow(2.0, exponent) + 0.5); /* cudaN4.h:32 */
    unsigned int histogramOffset = (unsigned int)(0.5 * (paddedHistogramSize - NumberOfHistogramBins));
    float WienerFilterNoise = 0.01;                  /* cudaN4.h:38 */
    float BiasFieldFullWidthAtHalfMaximum = 0.15;    /* cudaN4.h:44 */
    unsigned int NumberOfControlPoints = 4;          /* cudaN4.h:58 */
    unsigned int NumberOfControlPoints_x = 4;        /* cudaN4.h:60-62 */
    unsigned int NumberOfControlPoints_y = 4;
    unsigned int NumberOfControlPoints_z = 4;
    unsigned int NumberOfFittingLevels = 4;          /* cudaN4.h:71 */
    unsigned int MaximumNumberOfIterations = 50;     /* cudaN4.h:75 */
    float ConvergenceThreshold = 0.001;              /* cudaN4.h:84 */
    float low_value = 10.;                           /* cudaN4.h:85 */
    std::string inputImFile;                         /* cudaN4.h:89-92 */
    std::string outputImFile;
    std::string maskFile;
    std::string outBiasFieldFile;
    /* extensions */
    unsigned int ShrinkFactor = 1;
    double FitEps = 1e-3;
    std::vector<unsigned int> IterationsPerLevel;    /* empty: MaximumNumberOfIterations */
    unsigned int Precision = N4_PREC_FP32;
    bool FixedIterations = false;                    /* N4_FLAG_FIXED_ITERATIONS */
    bool WeightedMask = false;                       /* N4_FLAG_WEIGHTED_MASK */
};

/* Fills C-ABI parameters from N4Param.  Float fields are widened through
 * their shortest decimal form, so 0.15f becomes the double 0.15 (the value
 * N4.py uses) rather than 0.15000000596... */
n4_params N4ParamToC(const N4Param& p);

/* Device handle and workspace; replaces the FFTW plans and host scratch of
 * the reference (cudaN4.h:98-195). */
struct N4Data {
    n4_handle* handle = nullptr;
    int device = 0;
    /* cudaN4.h:132-174: arguments kept for source compatibility. */
    void assign(const unsigned int NumberOfHistogramBins, const unsigned int paddedHistogramSize);
    /* cudaN4.h:177-193 */
    void clean();
};

/* cudaN4.h:198-234, FFTW arguments replaced by N4Data&. */
void sharpenImage(float* d_before_sharpen, const float* d_mask, const unsigned int n,
                  const unsigned int NumberOfHistogramBins, const unsigned int paddedHistogramSize,
                  const unsigned int histogramOffset, const float WienerFilterNoise,
                  const float BiasFieldFullWidthAtHalfMaximum, N4Data& data, float* d_after_sharpen);

/* cudaN4.h:236-247 */
void testsharpenImage(const float* h_before_sharpen, const float* h_mask, const unsigned int numberOfPixels,
                      N4Param& param, N4Data& data, float* h_before_sharpen_log, float* h_after_sharpen,
                      float& binMin_out, float& binMax_out);

/* cudaN4.h:249-257 (numberOfForeground is recomputed from the mask). */
void calculateConvergenceMeasurement(const float* d_field1, const float* d_field2, const float* d_mask,
                                     const unsigned int numberOfPixels, const float numberOfForeground,
                                     float& convergence);

/* cudaN4.h:260-285 */
void upsampleLattice3D_cpu(const float* h_lattice, const unsigned int ncpt_x, const unsigned int ncpt_y,
                           const unsigned int ncpt_z, float* h_lattice_upsample);
void upsampleLattice3D_gpu(const float* d_lattice, const unsigned int ncpt_x, const unsigned int ncpt_y,
                           const unsigned int ncpt_z, float* d_lattice_upsample);
void testUpsampleLattice3D_gpu(const float* h_lattice, const unsigned int ncpt_x, const unsigned int ncpt_y,
                               const unsigned int ncpt_z, float* h_lattice_upsample);

/* cudaN4.h:288-310 */
void EvaluateBspline3D(const float* d_lattice, const unsigned int ncpt_x, const unsigned int ncpt_y,
                       const unsigned int ncpt_z, const unsigned int fx, const unsigned int fy,
                       const unsigned int fz, float* d_fitted);
void testEvaluateBspline3D(const float* h_lattice, const unsigned int ncpt_x, const unsigned int ncpt_y,
                           const unsigned int ncpt_z, const unsigned int fx, const unsigned int fy,
                           const unsigned int fz, float* h_fitted);

/* cudaN4.h:313-340 */
void FitBspline3D(const float* d_field, const float* d_mask, const unsigned int ncpt_x, const unsigned int ncpt_y,
                  const unsigned int ncpt_z, const unsigned int fx, const unsigned int fy, const unsigned int fz,
                  float* d_lattice);
void testFitBspline3D(const float* h_field, const float* h_mask, const unsigned int ncpt_x,
                      const unsigned int ncpt_y, const unsigned int ncpt_z, const unsigned int fx,
                      const unsigned int fy, const unsigned int fz, float* h_lattice);

/* cudaN4.h:343-365 */
void N4(float* d_im, float* d_mask, const unsigned int fx, const unsigned int fy, const unsigned int fz,
        N4Param& param, N4Data& data, float* d_im_normalized, float* d_lattice, float* d_biasField = NULL);

/* cudaN4.h:367-387 */
void testN4(float* h_im, float* h_mask, const unsigned int fx, const unsigned int fy, const unsigned int fz,
            N4Param& param, N4Data& data, float* h_im_normalized, float* h_lattice, float* h_biasField = NULL);

/* Per-call run record of the last testN4/N4 on this thread (extension). */
const n4_report& N4LastReport();

/* Convergence value of every iteration of the last testN4/N4 on `data`,
 * one vector per level (extension; n4_last_cv_trace). */
std::vector<std::vector<double>> N4LastCVTrace(N4Data& data);

/* Lattice of fitting level `level` at that level's end, of the last
 * testN4/N4 on `data`, [cx][cy][cz] (extension; n4_last_level_lattice). */
std::vector<float> N4LastLevelLattice(N4Data& data, unsigned level);

#endif
