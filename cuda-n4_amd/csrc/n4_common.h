// n4_common.h -- structures shared by the gfx950 kernels and the host driver.
//
// Device memory layout (per volume, all in one arena owned by the handle):
//   full resolution (caller's buffers): im f32, mask f32|u8, outputs f32
//   working grid W = ceil(F/s) per axis, C order [x][y][z]:
//     imlog f32, logU f32, wmask u8, wconf f32 (weighted masks only)
//   lattice: lat[2] f64 (ping-pong across levels), den f64, latf f32,
//            lresf f32 (largest level size)
//   per-slab partials (one slab = one working x index):
//     part  f32 [W0][Cy*Cz]   fit contraction over (y, z) of the slab
//     spart f64 [W0][4]       CV sums / counts and logU min / max
//   hist  u64 [nbins]         fixed-point (2^-32) fractional histogram
//   etab f64 [nbins], state (VolState)
// Axis tables (AxisEntry) are computed on the host in FP64 per level and per
// axis and uploaded once per call.
#pragma once
#include <stdint.h>

namespace n4 {

constexpr int kBlock = 256;
constexpr int kHistLanes = 8;   // LDS histogram copies per workgroup (power of two)
constexpr int kMaxCopies = 16;  // most global copies of a volume's histogram / fit numerator
// logU and the working mask are allocated with kQuadPad extra elements, so
// 16-byte (4-voxel) loads of aligned quads may run past a tile's end.
constexpr int kQuadPad = 16;
struct alignas(16) F4 {
    float v[4];
};
constexpr int kMaxLevels = 16;
constexpr int kMaxBins = 512;
constexpr int kChunkVox = 4096;   // working voxels staged in LDS per row chunk
constexpr int kMfmaMaxTiles = 4;  // MFMA kernels: up to 4 x 16 control points per axis

// One position along one axis at one level (A11/A12/A14 of the B-spline).
struct AxisEntry {
    float w[4];     // cubic B-spline weights B0..B3 at t (N4.py:174-181)
    float p[4];     // w^3 / sum_a(w_a^2)  -- separable fit-numerator factor
    float q[4];     // w^2                 -- separable fit-denominator factor
    int32_t b;      // span index floor((i+0.5)/span)  (N4.py:210)
    int32_t pad[3];
};

struct LevelDesc {
    int32_t C[3];            // control points per axis
    int32_t L;               // C0*C1*C2
    const AxisEntry* ax[3];  // W[d] entries (working grid)
    const int32_t* sup[3];   // 2*C[d]: [lo, hi) positions whose support holds c
    uint64_t mCz;            // magic divisor for C[2]
};

// Each volume runs its own level/iteration state machine inside the same
// launches ("steps": hist -> wiener -> fit -> eval).  At the start of a step
// the histogram kernel turns the state (mode, lvl, it) into the step's kind;
// the Wiener kernel (one block per volume) advances the state for the next
// step.  Kinds:
//   ITER(lvl, it)   histogram, Wiener step, fit numerator, lattice update
//   TRANS(lvl)      enter level lvl: refine the lattice (lvl > 0), fit the
//                   level's denominator
//   FINAL           all levels done: final FP32 lattice
//   SKIP            finished volume
enum StateMode : int32_t { kModeTrans = 0, kModeIter = 1, kModeFinished = 2 };
enum StepKind : int32_t { kStepSkip = 0, kStepIter = 1, kStepTrans = 2, kStepFinal = 3 };

struct VolState {
    double bin_min, bin_max, slope;   // histogram range of logU over the mask
    double cv;                        // last convergence measurement
    int32_t degenerate;               // empty mask or max == min: no sharpening
    int32_t done;                     // all levels finished
    int32_t mode, lvl, it;            // state at the start of the next step
    int32_t s_kind, s_lvl, s_it;      // the current step (written by the histogram kernel)
    int32_t passes;                   // histogram passes run (steps before the finish)
    uint64_t n_mask_full, n_mask_work;
    int32_t iters[kMaxLevels];
    double lastcv[kMaxLevels];
};

struct VolDesc {
    // full resolution
    const float* im;
    const void* mask;
    int32_t mask_kind;   // n4_mask_kind
    int32_t s;           // shrink factor
    float* out_norm;
    float* out_bias;
    float* out_logbias;
    float* out_lattice;
    int32_t F[3];
    int32_t W[3];
    int64_t nvox_w;      // W0*W1*W2 (flat stage calls may use fewer)
    // working grid
    float* imlog;
    float* logU;
    uint8_t* wmask;
    float* wconf;        // confidence weight per working voxel (weighted masks; else null)
    // lattice
    double* lat[2];
    double* den;
    float* latf;
    float* lresf;
    // partials
    float* part;               // per-slab fit contraction (DEN and FIELD modes)
    uint64_t* numfx;           // num_copies x lmax fixed-point lattice numerator copies
    uint64_t* hist;            // hist_copies x nbins fixed-point bins (summed by the Wiener step)
    double* spart;
    double* etab;
    VolState* st;
    uint64_t* progress;        // progress word in pinned host memory (see progress_word)
    double* cvtrace;           // convergence value of every iteration, level-major (null: none)
    int32_t trace_cap;         // entries of cvtrace (later iterations are not recorded)
    float* lvlat;              // each level's lattice at the level's end, level-major (null: none)
    LevelDesc lev[kMaxLevels];
    const AxisEntry* fax[3];   // full-res axis tables for the final lattice
    int32_t nlevels;
    int32_t max_iter[kMaxLevels];
    int32_t part_stride;       // floats per slab in part (max Cy*Cz)
    int32_t prep_rows;         // working rows per prep tile
    int32_t prep_tiles_x;      // prep tiles per working slab
    int32_t recon_rows;        // full-resolution rows per reconstruction tile
    int32_t wtile_rows;        // working rows per tile of the step kernels
    int32_t wtiles_x;          // such tiles per working slab (spart has W0*wtiles_x entries)
    int32_t hist_copies;       // tile t adds its histogram into copy t % hist_copies
    int32_t num_copies;        // tile t adds its fit numerator into copy t % num_copies
    int32_t lmax;              // control points of the finest level (numfx copy stride)
    double* ppart;             // prep partials [W0 * prep_tiles_x][4]
    uint64_t mW2;              // magic divisor for the working z extent
    uint64_t mF2;              // magic divisor for the full-resolution z extent
    uint32_t* team_ctr;        // persistent path: [0] the team's barrier counter, [1] abort flag
    uint64_t* tphase;          // persistent path (nullable): leader's device-clock ticks per phase (kPhases)
};

// Phases of the persistent path timed by a team's leader (tphase slots).
enum TeamPhase : int32_t {
    kPhWait = 0, kPhReduce, kPhHist, kPhWiener, kPhFit, kPhUpdate, kPhEval, kPhTrans, kPhSteps, kPhases
};
constexpr double kClockHz = 100.0e6;  // s_memrealtime

struct D2 {
    double x, y;
};

struct RunParams {
    const D2* twiddle;  // exp(-2 pi i k / P), k < P/2 (host-computed)
    int32_t nbins;
    int32_t P;        // padded FFT length (power of two)
    int32_t log2P;
    int32_t offset;   // histogram offset in the padded array
    double noise;     // Wiener filter noise
    double fwhm;      // bias field FWHM
    double thr;       // convergence threshold
    double eps;       // fit denominator eps
    double low;       // low value threshold
    float low_f;      // largest float <= low: x > low_f  <=>  (double)x > low
    int32_t fixed_iters;
    int32_t prec;     // n4_precision: 0 FP32, 1 BF16 MFMA, 2 split-BF16 MFMA
    // Call-wide values the step kernels would otherwise load through the
    // volume descriptor after the state (kernel arguments arrive in scalar
    // registers at launch): the volumes' states (volume v at states + v,
    // the same address as its VolDesc::st), the level count and the
    // per-level iteration limits.  Kernels launched over a group's volumes
    // (a sub-range of the descriptor array `descs`) find a volume's state at
    // states + (its descriptor - descs).
    VolState* states;
    const VolDesc* descs;
    int32_t nlevels;
    int32_t max_iter[kMaxLevels];
};

// Persistent path: one workgroup of a volume's team.  The team's
// workgroups are consecutive blocks; workgroup `rank` owns the volume's
// tiles [t0, t1) of the slab list (contiguous in (x, y0) order) for the
// whole correction.
struct PWork {
    int32_t vol;
    int32_t rank;   // 0 = the team's leader
    int32_t nteam;  // workgroups in the team
    int32_t t0, t1;
    int32_t vt0;    // the volume's first tile in the slab list
};

// One workgroup's tile: rows [y0, y0+ny) of slab x of volume vol.
struct Slab {
    int32_t vol;
    int32_t x;
    int32_t y0;
    int32_t ny;
};

// Progress word of one volume, written by the device into pinned host
// memory (one 64-bit store, so the fields are always read consistently):
// bits 40.. level + 1, bits 1..39 histogram passes run (pass k runs only
// after step k-2 has finished), bit 0 done.
constexpr uint64_t progress_word(int level, int passes, int done) {
    return ((uint64_t)(level + 1) << 40) | ((uint64_t)passes << 1) | (uint64_t)(done ? 1 : 0);
}
constexpr int progress_level(uint64_t w) { return (int)(w >> 40) - 1; }
constexpr int progress_passes(uint64_t w) { return (int)((w >> 1) & ((1ull << 39) - 1)); }
constexpr int progress_done(uint64_t w) { return (int)(w & 1); }

// Exact floor(v / d) for v * d < 2^40 as one 64-bit multiply and shift:
// m = floor(2^40 / d) + 1 (computed on the host).
constexpr int kDivShift = 40;

// Fit modes.  NUM and DEN_SCATTER add the slab's contraction, weighted by
// the x factor, into the fixed-point numerator buffers (the correction
// loop); DEN and FIELD write per-slab partials reduced by the lattice
// kernel (stage calls).  STEP takes the mode, level and iteration from the
// volume's current step (ITER -> NUM, TRANS -> DEN_SCATTER, else nothing).
enum FitMode : int32_t { kFitNum = 0, kFitDen = 1, kFitField = 2, kFitDenScatter = 3, kFitStep = 4 };

// Fixed-point scale of the lattice numerator accumulated by the fit's
// integer atomics (order-independent, so bitwise reproducible): 2^-30
// resolution, range +-8.6e9.
constexpr double kNumScale = 1073741824.0;  // 2^30

}  // namespace n4
