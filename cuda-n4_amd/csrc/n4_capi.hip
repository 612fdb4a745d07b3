// n4_capi.hip -- host driver behind the C ABI (include/n4_capi.h).
//
// A call plans the batch on the host (working grids, per-level control
// point counts, FP64 axis tables, slab lists, arena offsets), uploads the
// plan with one copy, then enqueues the level loop on the caller's stream:
//   prep -> for each level: level setup, fit(DEN) + lattice(DEN),
//   iterations x {hist, wiener, fit, lattice, eval} -> recon.
// Batches are split into volume groups that run on their own streams.
// Nothing is allocated inside the loop: the handle's arena grows only when a
// larger batch arrives.  Under N4_EMULATE the same driver runs the kernel
// bodies on the host (CPU unit tests of the kernel logic only).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <map>
#include <thread>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/n4_capi.h"
#include "n4_kernels.h"

#ifndef N4_EMULATE
#include <hip/hip_runtime.h>
#endif

using namespace n4;

// ----------------------------------------------------------- error state --
static thread_local std::string g_err;

static n4_status fail(n4_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

// --------------------------------------------------- device abstraction --
#ifndef N4_EMULATE
#define N4_CHECK_HIP(expr)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(N4_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));         \
    } while (0)
static int dev_alloc(void** p, size_t n) { return hipMalloc(p, n) == hipSuccess ? 0 : 1; }
static void dev_free(void* p) {
    if (p) (void)hipFree(p);
}
#define H2D(dst, src, n, st) N4_CHECK_HIP(hipMemcpyAsync((dst), (src), (n), hipMemcpyHostToDevice, (hipStream_t)(st)))
#define D2H(dst, src, n, st) N4_CHECK_HIP(hipMemcpyAsync((dst), (src), (n), hipMemcpyDeviceToHost, (hipStream_t)(st)))
#define D2D(dst, src, n, st) N4_CHECK_HIP(hipMemcpyAsync((dst), (src), (n), hipMemcpyDeviceToDevice, (hipStream_t)(st)))
#define DSET(dst, v, n, st) N4_CHECK_HIP(hipMemsetAsync((dst), (v), (n), (hipStream_t)(st)))
#define SYNC(st) N4_CHECK_HIP(hipStreamSynchronize((hipStream_t)(st)))
#define LAUNCH(expr)                                                                         \
    do {                                                                                     \
        int e_ = (expr);                                                                     \
        if (e_ != 0)                                                                         \
            return fail(N4_ERR_HIP, "kernel launch failed: %s", hipGetErrorString((hipError_t)e_)); \
    } while (0)
#else
static int dev_alloc(void** p, size_t n) {
    *p = aligned_alloc(256, (n + 255) & ~size_t(255));
    if (*p) memset(*p, 0, n);
    return *p ? 0 : 1;
}
static void dev_free(void* p) { free(p); }
#define H2D(dst, src, n, st) memcpy((dst), (src), (n))
#define D2H(dst, src, n, st) memcpy((dst), (src), (n))
#define D2D(dst, src, n, st) memcpy((dst), (src), (n))
#define DSET(dst, v, n, st) memset((dst), (v), (n))
#define SYNC(st) \
    do {         \
    } while (0)
#define LAUNCH(expr) (void)(expr)
#endif

// ---------------------------------------------------------------- handle --
static const int kMaxGroups = 8;

struct Plan;  // a correction's device layout (defined below)

struct n4_handle {
    int device = 0;
    void* arena = nullptr;
    size_t arena_bytes = 0;
    // Host image of a plan's upload region (stage): pinned host memory
    // (direct DMA) for ordinary calls, the pageable vector for the
    // allocation-free entry point (it allocates no pinned memory) and the
    // host emulation.  stage_ev marks the last upload from pinned memory:
    // the image is rewritten only after it.
    unsigned char* stage = nullptr;
    std::vector<unsigned char> staging;
    unsigned char* stage_pinned = nullptr;
    size_t stage_pinned_cap = 0;
#ifndef N4_EMULATE
    hipEvent_t stage_ev = nullptr;
    bool stage_ev_pending = false;
#endif
    uint64_t* pinned_flags = nullptr;    // flag_cap progress words (pinned host)
    uint64_t* dev_flags = nullptr;       // the same memory as seen by the device
    int flag_cap = 0;
#ifndef N4_EMULATE
    bool events_init = false;
    hipStream_t gstream[kMaxGroups] = {};  // streams of batch groups 1..
    hipEvent_t fork_ev = nullptr, join_ev[kMaxGroups] = {};
    std::vector<hipEvent_t> prof_ev;       // event pool for kernel timing
    std::vector<int> prof_pending;         // kernel id per recorded pair
#endif
    uint32_t prof_mask = 0;
    double prof_ms[N4_K_COUNT] = {0};
    uint64_t prof_n[N4_K_COUNT] = {0};
    n4_tuning tune{};  // launch schedule; 0 fields = defaults
    void* scratch = nullptr;  // the device's allocation-free scratch array (device address)
    // Run records of the last correction call (n4_last_cv_trace,
    // n4_last_level_lattice): device addresses in the arena (or the
    // scratch array) and the iterations per level, per volume in the
    // caller's order.  Any later plan reuses that memory and clears `valid`.
    struct LastRun {
        bool valid = false;
        bool in_scratch = false;     // allocation-free call: valid while the scratch generation holds
        uint64_t scratch_gen = 0;
        int nlevels = 0;
        std::vector<int> max_iter;   // per level
        std::vector<const double*> cvtrace;
        std::vector<int> trace_cap;
        std::vector<const float*> lvlat;
        std::vector<std::array<uint32_t, kMaxLevels>> iters;
        std::vector<std::array<int32_t, 3 * kMaxLevels>> ncpt;
    } last;
    // The last correction call's plan, reused by a call with the same key
    // (volume shapes, shrink factors, mask kinds, parameters, tiling, team
    // sizes, progress words, kernel timing) while no other plan has been
    // built since: the stage then still holds its upload image, and
    // only the caller's buffer addresses in the descriptors change.
    std::vector<unsigned char> plan_key;
    std::unique_ptr<Plan> plan_cached;
#ifndef N4_EMULATE
    // Captured chunks of iteration steps, one per group stream, reused while
    // their launch arguments (the key) stay the same.
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<unsigned char> key;
    } sgraph[kMaxGroups];
    hipStream_t cap_stream = nullptr;
#endif
};

// ------------------------------------------------- allocation-free scratch --
// n4_correct_noalloc keeps its plan and working buffers in this file-scope
// array (each device that runs the library holds its own copy until the
// process ends) instead of the handle's arena.  Host code never uses the
// array's name as an address: the device address is resolved once per
// device (hipGetSymbolAddress, at n4_create with the device current) and
// kept; calls that overlap on one device race on it.
static const size_t kScratchBytes = (size_t)64 << 20;
#ifndef N4_EMULATE
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
#else
alignas(256) static unsigned char scratch[kScratchBytes];
#endif
static const int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
// allocation-free calls started per device (a run record in the scratch is
// valid while this count is unchanged)
static std::atomic<uint64_t> g_scratch_gen[kMaxDevices];

static uint64_t scratch_generation(int device) {
    return (device >= 0 && device < kMaxDevices) ? g_scratch_gen[device].load() : 0;
}

// The current device's scratch address (resolved on first use).
static n4_status resolve_scratch(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return fail(N4_ERR_INVALID_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
#ifndef N4_EMULATE
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(N4_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
#else
        g_scratch_ptr[device] = scratch;
#endif
    }
    *out = g_scratch_ptr[device];
    return N4_OK;
}

// Early exit: each histogram pass writes the volume's progress word (level,
// passes run, done) straight into pinned host memory.  The host keeps at
// most kPipeline iterations enqueued ahead of the slowest unfinished volume
// and stops enqueueing once every volume is done at the current level; the
// iteration stream carries no events.
static const int kPipeline = 3;

static n4_status ensure_arena(n4_handle* h, size_t bytes) {
    if (bytes <= h->arena_bytes) return N4_OK;
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipSetDevice(h->device));
    N4_CHECK_HIP(hipDeviceSynchronize());
#endif
    dev_free(h->arena);
    h->arena = nullptr;
    h->arena_bytes = 0;
    size_t want = bytes + bytes / 8 + (1u << 20);
    if (dev_alloc(&h->arena, want))
        return fail(N4_ERR_NO_MEMORY, "device allocation of %zu bytes failed", want);
    h->arena_bytes = want;
    return N4_OK;
}

// ---------------------------------------------------------------- tables --
static void cubic_bspline(double t, double w[4]) {  // N4.py:174-181
    const double t2 = t * t, t3 = t * t * t, om = 1.0 - t;
    w[0] = om * om * om / 6.0;
    w[1] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
    w[2] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
    w[3] = t3 / 6.0;
}

// Span index / weights per position (N4.py:192, 210-217) and, per control
// point, the contiguous range of positions whose 4-support contains it.
static void axis_table(int npos, double domain, int ncpt, AxisEntry* out, int32_t* sup) {
    const double span = domain / (double)(ncpt - 3);
    for (int c = 0; c < ncpt; ++c) {
        sup[2 * c] = npos;
        sup[2 * c + 1] = 0;
    }
    for (int i = 0; i < npos; ++i) {
        const double pos = (double)i + 0.5;
        int b = (int)floor(pos / span);
        const double t = (pos - (double)b * span) / span;
        if (b < 0) b = 0;
        if (b > ncpt - 4) b = ncpt - 4;
        double w[4];
        cubic_bspline(t, w);
        const double S = w[0] * w[0] + w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
        AxisEntry& e = out[i];
        memset(&e, 0, sizeof e);
        for (int a = 0; a < 4; ++a) {
            e.w[a] = (float)w[a];
            e.p[a] = (float)(w[a] * w[a] * w[a] / S);
            e.q[a] = (float)(w[a] * w[a]);
            const int c = b + a;
            if (i < sup[2 * c]) sup[2 * c] = i;
            if (i + 1 > sup[2 * c + 1]) sup[2 * c + 1] = i + 1;
        }
        e.b = b;
    }
    for (int c = 0; c < ncpt; ++c)
        if (sup[2 * c] >= sup[2 * c + 1]) sup[2 * c] = sup[2 * c + 1] = 0;
}

static int padded_size(int nbins, int* log2P) {  // N4.py:31-37
    int e = 0;
    while ((1 << e) < nbins) ++e;
    *log2P = e + 1;
    return 1 << (e + 1);
}

static int level_ncpt(int c0, int level) { return (c0 - 3) * (1 << level) + 3; }

// The MFMA kernels hold up to 16 * kMfmaMaxTiles control points along y and
// z and take binary masks; a call whose final lattice has more, or that
// passes confidence weights, runs the FP32 kernels instead (in every level:
// the launches of one step mix volumes at different levels).
static bool mfma_covers(const n4_params* p) {
    if (p->flags & N4_FLAG_WEIGHTED_MASK) return false;
    const int L = (int)p->number_of_fitting_levels;
    for (int d = 1; d < 3; ++d)
        if (level_ncpt((int)p->ncpt[d], L - 1) > 16 * kMfmaMaxTiles) return false;
    return true;
}

static RunParams make_run_params(const n4_params* p) {
    RunParams rp;
    memset(&rp, 0, sizeof rp);
    rp.nbins = (int)p->number_of_histogram_bins;
    rp.P = padded_size(rp.nbins, &rp.log2P);
    rp.offset = (int)(0.5 * (double)(rp.P - rp.nbins));
    rp.noise = p->wiener_filter_noise;
    rp.fwhm = p->bias_field_fwhm;
    rp.thr = p->convergence_threshold;
    rp.eps = p->fit_eps;
    rp.low = p->low_value;
    float f = (float)p->low_value;
    if ((double)f > p->low_value) f = nextafterf(f, -INFINITY);
    rp.low_f = f;
    rp.fixed_iters = (p->flags & N4_FLAG_FIXED_ITERATIONS) ? 1 : 0;
    rp.prec = mfma_covers(p) ? (int)p->precision : 0;
    rp.nlevels = (int)p->number_of_fitting_levels;
    for (int l = 0; l < kMaxLevels; ++l) rp.max_iter[l] = (int)p->max_iterations[l];
    return rp;
}

// ----------------------------------------------------------------- plans --
// Bump layout of one arena: offsets first, pointers after allocation.
struct Layout {
    size_t off = 0;
    size_t take(size_t bytes, size_t align = 256) {
        off = (off + align - 1) / align * align;
        size_t o = off;
        off += bytes;
        return o;
    }
};

struct VolIn {  // one volume of a request
    const float* im;
    const void* mask;
    int mask_kind;
    int F[3];
    int s;
    float* out_norm;
    float* out_bias;
    float* out_logbias;
    float* out_lattice;
};

struct Plan {
    int nvol = 0;
    int nlevels = 0;
    RunParams rp;
    std::vector<VolDesc> desc;  // host copies (pointers into the arena)
    size_t upload_bytes = 0;    // [0, upload_bytes) of the arena comes from the stage
    size_t total_bytes = 0;
    size_t off_desc = 0, off_wslab = 0, off_fslab = 0, off_pslab = 0, off_state = 0, off_tw = 0;
    int n_wslab = 0, n_fslab = 0, n_pslab = 0;
    std::vector<int> max_iter;  // per level (max over volumes)
    std::vector<int> recon_tiles;  // recon tiles per volume
    size_t lds_fit_max[kMaxLevels] = {0}, lds_eval_max[kMaxLevels] = {0}, lds_eval32_max[kMaxLevels] = {0};
    int cx_max[kMaxLevels] = {0};
    int mt[kMaxLevels] = {0}, nt[kMaxLevels] = {0};  // MFMA tiles along y, z per level
    size_t lds_hist = 0, lds_wiener = 0, lds_recon = 0;
    // persistent path
    size_t off_work = 0;     // PWork[n_work] in the upload region
    int n_work = 0;
    size_t lds_persist = 0;  // dynamic LDS of the persistent kernel (largest volume)
    unsigned char* base = nullptr;  // device arena the plan lives in
};


// Row tiles of the step kernels on a working grid W: rows per tile (whole
// multiples of 16 rows, one MFMA row tile per wave) and tiles per slab.
// tiled == false (stage calls): whole slabs.
static void step_tiles(const int W[3], bool tiled, int tile_vox, int* rows, int* tiles_x) {
    int r = W[1];
    if (tiled && (size_t)W[1] * W[2] > (size_t)tile_vox) {
        const int t = (tile_vox / std::max(W[2], 1) + 15) / 16 * 16;
        r = std::min(W[1], std::max(t, 16));
    }
    *rows = r;
    *tiles_x = (W[1] + r - 1) / r;
}


// Magic multiplier for floor(v / d) (see fdiv in the kernels).
static uint64_t magic_div(uint32_t d) { return ((uint64_t)1 << kDivShift) / d + 1; }

// Working voxels per prep tile and full-resolution voxels per recon tile:
// enough workgroups to fill the device even for one volume.
static const int kPrepTileVox = 512;
// Default working voxels per tile of the step kernels.
static const int kStepTileVox = 8192;
// Convergence values recorded per volume and call (iterations after the
// first kTraceCap of a call are not recorded).
static const int kTraceCap = 4096;
static const int kReconTileVox = 8192;
// Default iteration steps per captured graph, and the largest chunk.
static const int kGraphSteps = 8;
static const int kMaxGraphSteps = 64;

// Launch schedule of one correction call (n4_tuning resolved).
struct Schedule {
    int groups;       // volume groups (streams)
    int tile_vox;     // working voxels per step-kernel tile
    int graph_steps;  // steps per captured graph; 0 = direct launches
    bool mfma_eval;   // MFMA field-increment kernel (MFMA precisions)
    int path;         // N4_PATH_*
    int team_per_cu;  // persistent path: workgroups per compute unit (0 = occupancy limit)
};

static int env_int(const char* name, int dflt) {
    if (const char* e = getenv(name)) {
        char* end = nullptr;
        const long v = strtol(e, &end, 10);
        if (end != e) return (int)v;
    }
    return dflt;
}

// Handle setting, else environment, else default.  Groups: each group runs
// its own chain of steps on its own stream (group 0 on the caller's
// stream), so the small per-volume kernels of one group overlap the slab
// kernels of the others and a group whose volumes converge stops without
// waiting for the rest.
static Schedule resolve_schedule(const n4_handle* h, int nvol) {
    Schedule s;
    int g = h->tune.groups ? (int)h->tune.groups : env_int("N4_GROUPS", nvol >= 8 ? 4 : (nvol >= 4 ? 2 : 1));
    g = std::min(std::min(g, kMaxGroups), nvol);
    s.groups = std::max(g, 1);
    const int t = h->tune.step_tile_voxels ? (int)std::min<uint32_t>(h->tune.step_tile_voxels, 1u << 30)
                                           : env_int("N4_STEP_TILE_VOX", kStepTileVox);
    s.tile_vox = std::max(t, 16);
    int k;
    if (h->tune.graph_steps == N4_GRAPH_OFF)
        k = 0;
    else if (h->tune.graph_steps)
        k = (int)std::min<uint32_t>(h->tune.graph_steps, kMaxGraphSteps);
    else
        k = env_int("N4_GRAPH_STEPS", kGraphSteps);
    s.graph_steps = std::min(std::max(k, 0), kMaxGraphSteps);
    // the FP32 evaluation measured faster than the MFMA one on MI355X at
    // every working-grid size tried, so it is the default in every mode
    const uint32_t ek = h->tune.eval_kernel ? h->tune.eval_kernel
                                            : (uint32_t)env_int("N4_EVAL_KERNEL", (int)N4_EVAL_FP32);
    s.mfma_eval = ek == N4_EVAL_MFMA;
    s.path = h->tune.path ? (int)h->tune.path : env_int("N4_PATH", (int)N4_PATH_AUTO);
    // the step kernels measured faster than the team kernel on every
    // workload tried (per-tile memory latency is exposed when a workgroup
    // walks its tiles one after another), so AUTO runs the steps
    if (s.path == (int)N4_PATH_AUTO) s.path = (int)N4_PATH_STEPS;
    s.team_per_cu = h->tune.team_wgs_per_cu ? (int)h->tune.team_wgs_per_cu : env_int("N4_TEAM_WGS_PER_CU", 0);
    return s;
}

// Copies of each volume's global histogram: the tiles of a correction
// spread their atomic flushes over up to kHistCopies copies (stage calls
// keep one, which they read back directly).
static const int kHistCopies = kMaxCopies;
static int hist_copies(int tiles, bool tiled) { return tiled ? std::max(1, std::min(kHistCopies, tiles)) : 1; }
// Copies of the fixed-point fit numerator, the same way (summed by the
// lattice update of each step).
static const int kNumCopies = kMaxCopies;
static int num_copies(int tiles, bool tiled) { return tiled ? std::max(1, std::min(kNumCopies, tiles)) : 1; }

static int rows_for(int rows, int cols, int target) {
    int r = target / (cols > 0 ? cols : 1);
    if (r < 1) r = 1;
    return r < rows ? r : rows;
}

// The last upload from the stage has completed.
static void stage_wait(n4_handle* h) {
#ifndef N4_EMULATE
    if (h->stage_ev_pending) {
        (void)hipEventSynchronize(h->stage_ev);
        h->stage_ev_pending = false;
    }
#else
    (void)h;
#endif
}

// A stage of n bytes, pinned when allowed (growing the pinned buffer if
// needed; should pinned memory be refused, the pageable vector serves).
// Its content is left as it is: build_plan writes every byte a kernel reads
// and zeroes the volume states.
static void stage_reset(n4_handle* h, size_t n, bool pinned_ok) {
    stage_wait(h);
#ifndef N4_EMULATE
    if (pinned_ok && h->stage_pinned_cap < n) {
        if (h->stage_pinned) (void)hipHostFree(h->stage_pinned);
        h->stage_pinned = nullptr;
        h->stage_pinned_cap = 0;
        void* q = nullptr;
        if (hipHostMalloc(&q, n, hipHostMallocDefault) == hipSuccess) {
            h->stage_pinned = (unsigned char*)q;
            h->stage_pinned_cap = n;
        } else {
            (void)hipGetLastError();
        }
    }
    if (pinned_ok && h->stage_pinned) {
        h->stage = h->stage_pinned;
        return;
    }
#else
    (void)pinned_ok;
#endif
    if (h->staging.size() < n) h->staging.resize(n);
    h->stage = h->staging.data();
}

// Enqueues a copy of n bytes at offset off of the stage to dst on stream st
// and, for a pinned stage, marks it (stage_wait).
static n4_status stage_upload(n4_handle* h, void* dst, size_t off, size_t n, void* st) {
    H2D(dst, h->stage + off, n, st);
#ifndef N4_EMULATE
    if (h->stage == h->stage_pinned) {
        if (!h->stage_ev) N4_CHECK_HIP(hipEventCreateWithFlags(&h->stage_ev, hipEventDisableTiming));
        N4_CHECK_HIP(hipEventRecord(h->stage_ev, (hipStream_t)st));
        h->stage_ev_pending = true;
    }
#endif
    return N4_OK;
}

// Builds the plan and fills h->stage with the upload region.  Pointers in
// the descriptors are arena-relative until `rebase` adds the arena base.
// tiled: cut working slabs into row tiles for the step kernels (the
// correction); stage calls keep whole slabs (their partials are per slab).

// teams (nullable, persistent path): workgroups per volume; the plan then
// holds the PWork list (consecutive ranks per volume, each rank a
// contiguous share of the volume's tiles), a barrier slot per volume and
// histogram copies sized to the team.
// noalloc: the plan lives in the device's file-scope scratch array (no
// allocation; N4_ERR_NO_MEMORY beyond its size) and carries no progress
// words (pinned host memory).
static n4_status build_plan(n4_handle* h, const std::vector<VolIn>& in, const n4_params* p, Plan* P,
                            bool need_tables_full = true, bool tiled = false, int tile_vox = kStepTileVox,
                            const std::vector<int>* teams = nullptr, bool noalloc = false) {
    const int nv = (int)in.size();
    const int L = (int)p->number_of_fitting_levels;
    h->last.valid = false;  // the plan reuses the memory of the last run's records
    h->plan_cached.reset();  // the stage and the arena are rewritten
    P->nvol = nv;
    P->nlevels = L;
    P->rp = make_run_params(p);
    P->max_iter.assign(L, 0);
    for (int l = 0; l < L; ++l) P->max_iter[l] = (int)p->max_iterations[l];

    // ---- sizes
    struct VInfo {
        int W[3];
        int C[kMaxLevels][3];
        size_t Lmax, part_stride;
        int prep_rows, prep_tiles_x, recon_rows, recon_tiles_x, wtile_rows, wtiles_x;
    };
    std::vector<VInfo> vi(nv);
    int nws = 0, nfs = 0, nps = 0;
    for (int v = 0; v < nv; ++v) {
        VInfo& I = vi[v];
        for (int d = 0; d < 3; ++d) I.W[d] = (in[v].F[d] + in[v].s - 1) / in[v].s;
        for (int l = 0; l < L; ++l)
            for (int d = 0; d < 3; ++d) I.C[l][d] = level_ncpt((int)p->ncpt[d], l);
        const int* Cf = I.C[L - 1];
        I.Lmax = (size_t)Cf[0] * Cf[1] * Cf[2];
        I.part_stride = (size_t)Cf[1] * Cf[2];
        I.prep_rows = rows_for(I.W[1], I.W[2], kPrepTileVox);
        I.prep_tiles_x = (I.W[1] + I.prep_rows - 1) / I.prep_rows;
        I.recon_rows = rows_for(in[v].F[1], in[v].F[2], kReconTileVox);
        I.recon_tiles_x = (in[v].F[1] + I.recon_rows - 1) / I.recon_rows;
        step_tiles(I.W, tiled, tile_vox, &I.wtile_rows, &I.wtiles_x);
        nws += I.W[0] * I.wtiles_x;
        nps += I.W[0] * I.prep_tiles_x;
        nfs += in[v].F[0] * I.recon_tiles_x;
    }
    P->n_wslab = nws;
    P->n_pslab = nps;
    P->recon_tiles.assign(nv, 0);
    for (int v = 0; v < nv; ++v) P->recon_tiles[v] = need_tables_full ? in[v].F[0] * vi[v].recon_tiles_x : 0;
    P->n_fslab = need_tables_full ? nfs : 0;

    Layout lay;
    P->off_desc = lay.take(sizeof(VolDesc) * nv);
    P->off_wslab = lay.take(sizeof(Slab) * (nws > 0 ? nws : 1));
    P->off_fslab = lay.take(sizeof(Slab) * (P->n_fslab > 0 ? P->n_fslab : 1));
    P->off_pslab = lay.take(sizeof(Slab) * (nps > 0 ? nps : 1));
    P->off_state = lay.take(sizeof(VolState) * nv);
    P->off_tw = lay.take(sizeof(D2) * P->rp.P);
    P->n_work = 0;
    if (teams)
        for (int v = 0; v < nv; ++v) P->n_work += (*teams)[v];
    P->off_work = lay.take(sizeof(PWork) * std::max(P->n_work, 1));
    // axis tables: identical ones (same positions, domain and control
    // points) are laid out and computed once per plan and shared by every
    // volume and level using them; working-grid tables carry their support
    // ranges, full-resolution tables do not
    struct AxisJob {
        int npos;
        double domain;
        int ncpt;
        size_t oa, os;  // entries; support ranges (working-grid tables)
        bool full;
    };
    std::vector<AxisJob> ax_jobs;
    std::map<std::tuple<bool, int, double, int>, size_t> ax_index;
    auto axis_slot = [&](bool full, int npos, double domain, int ncpt) -> const AxisJob& {
        const auto key = std::make_tuple(full, npos, domain, ncpt);
        auto it = ax_index.find(key);
        if (it == ax_index.end()) {
            AxisJob j{npos, domain, ncpt, lay.take(sizeof(AxisEntry) * npos, 64),
                      full ? 0 : lay.take(sizeof(int32_t) * 2 * ncpt, 16), full};
            it = ax_index.emplace(key, ax_jobs.size()).first;
            ax_jobs.push_back(j);
        }
        return ax_jobs[it->second];
    };
    std::vector<size_t> off_ax(nv * kMaxLevels * 3), off_sup(nv * kMaxLevels * 3), off_fax(nv * 3);
    for (int v = 0; v < nv; ++v)
        for (int l = 0; l < L; ++l)
            for (int d = 0; d < 3; ++d) {
                const AxisJob& j = axis_slot(false, vi[v].W[d], (double)vi[v].W[d], vi[v].C[l][d]);
                off_ax[(v * kMaxLevels + l) * 3 + d] = j.oa;
                off_sup[(v * kMaxLevels + l) * 3 + d] = j.os;
            }
    if (need_tables_full)
        for (int v = 0; v < nv; ++v)
            for (int d = 0; d < 3; ++d)
                off_fax[v * 3 + d] =
                    axis_slot(true, in[v].F[d], (double)vi[v].W[d] * (double)in[v].s, vi[v].C[L - 1][d]).oa;
    P->upload_bytes = lay.take(0);

    // per-volume working buffers
    struct Offs {
        size_t imlog, logU, wmask, wconf, lat0, lat1, den, latf, lresf, part, numfx, hist, spart, etab, ppart, team,
            trace, lvlat;
    };
    // histogram copies: per tile (step kernels) or per team member (teams)
    auto hcopies = [&](int v) {
        if (teams) return std::max(1, std::min(kHistCopies, ((*teams)[v] + 7) / 8));
        return hist_copies(vi[v].W[0] * vi[v].wtiles_x, tiled);
    };
    int trace_cap = 0;
    for (int l = 0; l < L; ++l) trace_cap += (int)std::min<uint32_t>(p->max_iterations[l], (uint32_t)kTraceCap);
    trace_cap = std::min(trace_cap, kTraceCap);
    // confidence weights of the working grid: F32 masks under N4_FLAG_WEIGHTED_MASK
    auto weighted = [&](int v) { return (p->flags & N4_FLAG_WEIGHTED_MASK) && in[v].mask_kind == N4_MASK_F32; };
    std::vector<Offs> off(nv);
    const int nb = P->rp.nbins;
    for (int v = 0; v < nv; ++v) {
        const VInfo& I = vi[v];
        const size_t nw = (size_t)I.W[0] * I.W[1] * I.W[2];
        Offs& o = off[v];
        o.imlog = lay.take(sizeof(float) * nw);
        o.logU = lay.take(sizeof(float) * (nw + kQuadPad));
        o.wmask = lay.take(nw + kQuadPad);
        o.wconf = weighted(v) ? lay.take(sizeof(float) * nw) : 0;
        o.lat0 = lay.take(sizeof(double) * I.Lmax);
        o.lat1 = lay.take(sizeof(double) * I.Lmax);
        o.den = lay.take(sizeof(double) * I.Lmax);
        o.latf = lay.take(sizeof(float) * I.Lmax);
        o.lresf = lay.take(sizeof(float) * I.Lmax);
        o.part = lay.take(sizeof(float) * I.W[0] * I.part_stride);
        o.numfx = lay.take(sizeof(uint64_t) * I.Lmax * num_copies(I.W[0] * I.wtiles_x, tiled));
        o.hist = lay.take(sizeof(uint64_t) * nb * hcopies(v));
        o.spart = lay.take(sizeof(double) * 4 * I.W[0] * I.wtiles_x);
        o.ppart = lay.take(sizeof(double) * 4 * I.W[0] * I.prep_tiles_x);
        o.etab = lay.take(sizeof(double) * nb);
        o.team = teams ? lay.take(64 + sizeof(uint64_t) * kPhases) : 0;
        // run records of a correction (tiled plans): CV trace, level lattices
        o.trace = tiled ? lay.take(sizeof(double) * trace_cap) : 0;
        size_t lsum = 0;
        for (int l = 0; l < L; ++l) lsum += (size_t)I.C[l][0] * I.C[l][1] * I.C[l][2];
        o.lvlat = tiled ? lay.take(sizeof(float) * lsum) : 0;
    }
    P->total_bytes = lay.take(0);

    unsigned char* base = nullptr;
    if (noalloc) {
        if (!h->scratch) return fail(N4_ERR_INTERNAL, "allocation-free scratch of device %d not resolved", h->device);
        if (P->total_bytes > kScratchBytes)
            return fail(N4_ERR_NO_MEMORY, "the request needs %zu bytes of device scratch; the allocation-free entry "
                        "point has %zu (use n4_correct)", P->total_bytes, (size_t)kScratchBytes);
        base = (unsigned char*)h->scratch;
        if (h->device >= 0 && h->device < kMaxDevices) g_scratch_gen[h->device]++;
    } else {
        n4_status s = ensure_arena(h, P->total_bytes);
        if (s != N4_OK) return s;
        base = (unsigned char*)h->arena;
    }
    P->base = base;
    P->rp.states = (VolState*)(base + P->off_state);
    P->rp.descs = (const VolDesc*)(base + P->off_desc);

    // ---- host image of the upload region
    stage_reset(h, P->upload_bytes, !noalloc);
    unsigned char* hb = h->stage;
    P->desc.assign(nv, VolDesc());
    Slab* ws = (Slab*)(hb + P->off_wslab);
    Slab* fs = (Slab*)(hb + P->off_fslab);
    Slab* ps = (Slab*)(hb + P->off_pslab);
    int iws = 0, ifs = 0, ips = 0, iwk = 0;
    {
        D2* tw = (D2*)(hb + P->off_tw);
        for (int m = 0; m < P->rp.P; ++m) {
            const double ang = 2.0 * (double)m / (double)P->rp.P;  // in units of pi
            tw[m] = D2{cos(M_PI * ang), -sin(M_PI * ang)};
        }
        P->rp.twiddle = (const D2*)(base + P->off_tw);
    }
    // the volume states start at zero (the rest of the image is written
    // below wherever a kernel reads it)
    memset(hb + P->off_state, 0, sizeof(VolState) * nv);
    // the distinct axis tables, on up to 8 host threads when there are many
    // (each thread takes the next table; the calling thread works too, so
    // every table is computed even if no other thread could be started)
    {
        const int nj = (int)ax_jobs.size();
        std::atomic<int> next{0};
        auto run_jobs = [&]() {
            std::vector<int32_t> sup_tmp;
            for (int k = next.fetch_add(1); k < nj; k = next.fetch_add(1)) {
                const AxisJob& j = ax_jobs[k];
                int32_t* sup = (int32_t*)(hb + j.os);
                if (j.full) {
                    sup_tmp.assign(2 * (size_t)j.ncpt, 0);
                    sup = sup_tmp.data();
                }
                axis_table(j.npos, j.domain, j.ncpt, (AxisEntry*)(hb + j.oa), sup);
            }
        };
        size_t entries = 0;
        for (const AxisJob& j : ax_jobs) entries += (size_t)j.npos;
        const int nt = entries >= 8192 ? std::min(8, std::max(1, (int)std::thread::hardware_concurrency())) : 1;
        std::vector<std::thread> pool;
        for (int k = 1; k < nt; ++k) {
            try {
                pool.emplace_back(run_jobs);
            } catch (...) {
                break;
            }
        }
        run_jobs();
        for (std::thread& th : pool) th.join();
    }
    for (int v = 0; v < nv; ++v) {
        const VInfo& I = vi[v];
        VolDesc D;
        memset(&D, 0, sizeof D);
        D.im = in[v].im;
        D.mask = in[v].mask;
        D.mask_kind = in[v].mask_kind;
        D.s = in[v].s;
        D.out_norm = in[v].out_norm;
        D.out_bias = in[v].out_bias;
        D.out_logbias = in[v].out_logbias;
        D.out_lattice = in[v].out_lattice;
        for (int d = 0; d < 3; ++d) {
            D.F[d] = in[v].F[d];
            D.W[d] = I.W[d];
        }
        D.nvox_w = (int64_t)I.W[0] * I.W[1] * I.W[2];
        const Offs& o = off[v];
        D.imlog = (float*)(base + o.imlog);
        D.logU = (float*)(base + o.logU);
        D.wmask = (uint8_t*)(base + o.wmask);
        D.wconf = weighted(v) ? (float*)(base + o.wconf) : nullptr;
        D.lat[0] = (double*)(base + o.lat0);
        D.lat[1] = (double*)(base + o.lat1);
        D.den = (double*)(base + o.den);
        D.latf = (float*)(base + o.latf);
        D.lresf = (float*)(base + o.lresf);
        D.part = (float*)(base + o.part);
        D.numfx = (uint64_t*)(base + o.numfx);
        D.num_copies = num_copies(I.W[0] * I.wtiles_x, tiled);
        D.lmax = (int32_t)I.Lmax;
        D.hist = (uint64_t*)(base + o.hist);
        D.spart = (double*)(base + o.spart);
        D.ppart = (double*)(base + o.ppart);
        D.prep_rows = I.prep_rows;
        D.prep_tiles_x = I.prep_tiles_x;
        D.recon_rows = I.recon_rows;
        D.wtile_rows = I.wtile_rows;
        D.wtiles_x = I.wtiles_x;
        D.hist_copies = hcopies(v);
        D.team_ctr = teams ? (uint32_t*)(base + o.team) : nullptr;
        D.tphase = (teams && (h->prof_mask >> N4_K_TEAMS & 1u)) ? (uint64_t*)(base + o.team + 64) : nullptr;
        D.mW2 = magic_div((uint32_t)I.W[2]);
        D.mF2 = magic_div((uint32_t)in[v].F[2]);
        D.etab = (double*)(base + o.etab);
        D.st = (VolState*)(base + P->off_state) + v;
        D.progress = (!noalloc && h->dev_flags && v < h->flag_cap) ? h->dev_flags + v : nullptr;
        D.cvtrace = tiled ? (double*)(base + o.trace) : nullptr;
        D.trace_cap = tiled ? trace_cap : 0;
        D.lvlat = tiled ? (float*)(base + o.lvlat) : nullptr;
        D.nlevels = L;
        D.part_stride = (int32_t)I.part_stride;
        for (int l = 0; l < L; ++l) {
            D.max_iter[l] = (int)p->max_iterations[l];
            LevelDesc& LD = D.lev[l];
            for (int d = 0; d < 3; ++d) {
                LD.C[d] = I.C[l][d];
                LD.ax[d] = (const AxisEntry*)(base + off_ax[(v * kMaxLevels + l) * 3 + d]);
                LD.sup[d] = (const int32_t*)(base + off_sup[(v * kMaxLevels + l) * 3 + d]);
            }
            LD.L = LD.C[0] * LD.C[1] * LD.C[2];
            LD.mCz = magic_div((uint32_t)LD.C[2]);
        }
        if (need_tables_full)
            for (int d = 0; d < 3; ++d) D.fax[d] = (const AxisEntry*)(base + off_fax[v * 3 + d]);
        P->desc[v] = D;
        if (teams) {
            PWork* wk = (PWork*)(hb + P->off_work);
            const int nt = (*teams)[v], ntl = I.W[0] * I.wtiles_x;
            for (int r = 0; r < nt; ++r)
                wk[iwk++] = PWork{v, r, nt, iws + (int)((int64_t)ntl * r / nt), iws + (int)((int64_t)ntl * (r + 1) / nt),
                                  iws};
            P->lds_persist = std::max(P->lds_persist, lds_persist(D, nb, P->rp.P, P->rp.prec));
        }
        for (int x = 0; x < I.W[0]; ++x) {
            for (int y0 = 0; y0 < I.W[1]; y0 += I.wtile_rows)
                ws[iws++] = Slab{v, x, y0, std::min(I.wtile_rows, I.W[1] - y0)};
            for (int y0 = 0; y0 < I.W[1]; y0 += I.prep_rows)
                ps[ips++] = Slab{v, x, y0, std::min(I.prep_rows, I.W[1] - y0)};
        }
        if (need_tables_full)
            for (int x = 0; x < in[v].F[0]; ++x)
                for (int y0 = 0; y0 < in[v].F[1]; y0 += I.recon_rows)
                    fs[ifs++] = Slab{v, x, y0, std::min(I.recon_rows, in[v].F[1] - y0)};
        const int prec = P->rp.prec;
        for (int l = 0; l < L; ++l) {
            P->lds_fit_max[l] = std::max(P->lds_fit_max[l], lds_fit(nb, D, l, prec));
            P->lds_eval_max[l] = std::max(P->lds_eval_max[l], lds_eval(D, l, prec));
            P->lds_eval32_max[l] = std::max(P->lds_eval32_max[l], lds_eval(D, l, 0));
            P->cx_max[l] = std::max(P->cx_max[l], D.lev[l].C[0]);
            P->mt[l] = mfma_tiles(D.lev[l].C[1]);
            P->nt[l] = mfma_tiles(D.lev[l].C[2]);
        }
        if (need_tables_full) P->lds_recon = std::max(P->lds_recon, lds_recon(D));
    }
    P->lds_hist = lds_hist(nb);
    P->lds_wiener = lds_wiener(P->rp.P);
    memcpy(hb + P->off_desc, P->desc.data(), sizeof(VolDesc) * nv);
    // VolState region stays zero: counters start at 0.
    return N4_OK;
}

// LDS of every kernel a grid plan launches must fit one workgroup.
static const size_t kLdsCap = 64 * 1024;
static n4_status check_lds(const Plan& P, bool grid_kernels, bool recon) {
    size_t worst = std::max(P.lds_hist, P.lds_wiener);
    if (grid_kernels)
        for (int l = 0; l < P.nlevels; ++l) worst = std::max(worst, std::max(P.lds_fit_max[l], P.lds_eval_max[l]));
    if (recon) worst = std::max(worst, P.lds_recon);
    if (worst > kLdsCap)
        return fail(N4_ERR_UNSUPPORTED, "volume/lattice too large for one workgroup's LDS (%zu bytes)", worst);
    return N4_OK;
}

static inline VolDesc* dev_desc(n4_handle* h, const Plan& P) {
    return (VolDesc*)(P.base + P.off_desc);
}
static inline Slab* dev_wslab(n4_handle* h, const Plan& P) {
    return (Slab*)(P.base + P.off_wslab);
}
static inline Slab* dev_fslab(n4_handle* h, const Plan& P) {
    return (Slab*)(P.base + P.off_fslab);
}
static inline Slab* dev_pslab(n4_handle* h, const Plan& P) {
    return (Slab*)(P.base + P.off_pslab);
}
static inline VolState* dev_state(n4_handle* h, const Plan& P) {
    return (VolState*)(P.base + P.off_state);
}

// --------------------------------------------------------------- checks --
static n4_status check_params(const n4_params* p) {
    if (!p) return fail(N4_ERR_INVALID_ARG, "params is NULL");
    if (p->number_of_histogram_bins < 2 || p->number_of_histogram_bins > N4_MAX_BINS ||
        p->number_of_histogram_bins > (uint32_t)kMaxBins)
        return fail(N4_ERR_INVALID_ARG, "number_of_histogram_bins must be in [2, %d]", kMaxBins);
    if (p->number_of_fitting_levels < 1 || p->number_of_fitting_levels > N4_MAX_LEVELS)
        return fail(N4_ERR_INVALID_ARG, "number_of_fitting_levels must be in [1, %d]", N4_MAX_LEVELS);
    for (int d = 0; d < 3; ++d)
        if (p->ncpt[d] < 4 || p->ncpt[d] > 4096)
            return fail(N4_ERR_INVALID_ARG, "ncpt[%d] = %u: need >= 4 control points (spline order 3)", d,
                        p->ncpt[d]);
    for (uint32_t l = 0; l < p->number_of_fitting_levels; ++l)
        if (p->max_iterations[l] > 100000) return fail(N4_ERR_INVALID_ARG, "max_iterations[%u] too large", l);
    if (!(p->bias_field_fwhm > 0.0)) return fail(N4_ERR_INVALID_ARG, "bias_field_fwhm must be > 0");
    if (!(p->wiener_filter_noise >= 0.0)) return fail(N4_ERR_INVALID_ARG, "wiener_filter_noise must be >= 0");
    if (!(p->fit_eps >= 0.0)) return fail(N4_ERR_INVALID_ARG, "fit_eps must be >= 0");
    if (p->shrink_factor < 1 || p->shrink_factor > 64) return fail(N4_ERR_INVALID_ARG, "shrink_factor must be in [1, 64]");
    if (p->precision != N4_PREC_FP32 && p->precision != N4_PREC_BF16 && p->precision != N4_PREC_SPLIT_BF16)
        return fail(N4_ERR_INVALID_ARG, "precision %u is not an n4_precision value", p->precision);
    if (p->flags & ~(N4_FLAG_FIXED_ITERATIONS | N4_FLAG_WEIGHTED_MASK))
        return fail(N4_ERR_INVALID_ARG, "flags 0x%x: unknown bits", p->flags);
#ifdef N4_EMULATE
    if (p->precision != N4_PREC_FP32)
        return fail(N4_ERR_UNSUPPORTED, "precision %u: the host emulation build runs the FP32 kernels only",
                    p->precision);
#endif
    // final lattice size must stay sane
    uint64_t Lf = 1;
    for (int d = 0; d < 3; ++d) {
        uint64_t c = (uint64_t)(p->ncpt[d] - 3) << (p->number_of_fitting_levels - 1);
        if (c > 4096) return fail(N4_ERR_INVALID_ARG, "final lattice too large on axis %d", d);
        Lf *= c + 3;
    }
    if (Lf > (1u << 24)) return fail(N4_ERR_INVALID_ARG, "final lattice too large");
    return N4_OK;
}

static n4_status check_volume(const VolIn& v) {
    for (int d = 0; d < 3; ++d)
        if (v.F[d] < 1 || v.F[d] > 65536) return fail(N4_ERR_INVALID_ARG, "dimension %d = %d out of range", d, v.F[d]);
    if ((uint64_t)v.F[0] * v.F[1] * v.F[2] > (1ull << 34)) return fail(N4_ERR_INVALID_ARG, "volume too large");
    if (!v.im) return fail(N4_ERR_INVALID_ARG, "input image pointer is NULL");
    if (!v.out_norm) return fail(N4_ERR_INVALID_ARG, "normalized output pointer is NULL");
    if (v.mask_kind < 0 || v.mask_kind > 2) return fail(N4_ERR_INVALID_ARG, "mask_kind %d invalid", v.mask_kind);
    if (v.mask_kind != N4_MASK_NONE && !v.mask) return fail(N4_ERR_INVALID_ARG, "mask pointer is NULL");
    if (v.s < 1 || v.s > 64) return fail(N4_ERR_INVALID_ARG, "shrink factor %d out of range", v.s);
    return N4_OK;
}

// --------------------------------------------------------------- engine --
// Brackets one launch with an event pair when its kernel is profiled.
struct ProfScope {
    n4_handle* h;
    int id;
    void* stream;
    bool on;
    ProfScope(n4_handle* h_, int id_, void* s_) : h(h_), id(id_), stream(s_), on((h_->prof_mask >> id_) & 1u) {
#ifndef N4_EMULATE
        if (on) record();
#endif
    }
    ~ProfScope() {
#ifndef N4_EMULATE
        if (on) {
            record();
            h->prof_pending.push_back(id);
        }
#endif
    }
#ifndef N4_EMULATE
    // Timing events without the system-scope fence an event record
    // performs by default (a cache flush that would stretch the measured
    // interval and slow the kernel after it).
    void record() {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
        (void)hipEventRecord(e, (hipStream_t)stream);
        h->prof_ev.push_back(e);
    }
#endif
};

// Folds recorded event pairs into the accumulators (after a stream sync).
static void prof_collect(n4_handle* h) {
#ifndef N4_EMULATE
    const size_t npairs = std::min(h->prof_pending.size(), h->prof_ev.size() / 2);
    for (size_t i = 0; i < npairs; ++i) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, h->prof_ev[2 * i], h->prof_ev[2 * i + 1]) == hipSuccess) {
            h->prof_ms[h->prof_pending[i]] += ms;
            h->prof_n[h->prof_pending[i]] += 1;
        }
    }
    for (hipEvent_t e : h->prof_ev) (void)hipEventDestroy(e);
    h->prof_ev.clear();
    h->prof_pending.clear();
#else
    (void)h;
#endif
}

static n4_status ensure_flags(n4_handle* h, int nvol) {
    if (nvol <= h->flag_cap) return N4_OK;
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipDeviceSynchronize());
    if (h->pinned_flags) (void)hipHostFree(h->pinned_flags);
    h->pinned_flags = nullptr;
    N4_CHECK_HIP(hipHostMalloc((void**)&h->pinned_flags, sizeof(uint64_t) * nvol, hipHostMallocDefault));
    N4_CHECK_HIP(hipHostGetDevicePointer((void**)&h->dev_flags, h->pinned_flags, 0));
    if (!h->events_init) {
        N4_CHECK_HIP(hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming));
        for (int g = 0; g < kMaxGroups; ++g) N4_CHECK_HIP(hipEventCreateWithFlags(&h->join_ev[g], hipEventDisableTiming));
        h->events_init = true;
    }
#else
    free(h->pinned_flags);
    h->pinned_flags = (uint64_t*)calloc((size_t)nvol, sizeof(uint64_t));
    h->dev_flags = h->pinned_flags;
#endif
    h->flag_cap = nvol;
    return N4_OK;
}

// Non-blocking check of a group (volumes [v0, v1)) that has iterations
// [0, it) of `level` enqueued on `stream`: *ready once every volume has
// finished iteration it - depth - 1 (its histogram pass it - depth + 1
// has run) or is done; *all_done once every volume is done.  Progress only
// moves forward within a level, so a newer state than the one asked for is
// harmless.  With `query`, a drained stream makes the words final (this
// also ends the wait if a stream stopped early).
static n4_status check_progress(n4_handle* h, int v0, int v1, int level, int it, int depth, void* stream,
                                bool query, bool* ready, bool* all_done) {
    const int need = it - depth + 1;
    const volatile uint64_t* f = h->pinned_flags;
    bool rdy = true, all = true;
    for (int v = v0; v < v1; ++v) {
        const uint64_t w = f[v];
        const bool cur = progress_level(w) == level;
        const bool done = cur && progress_done(w);
        all = all && done;
        rdy = rdy && (done || need <= 0 || (cur && progress_passes(w) >= need));
    }
#ifndef N4_EMULATE
    if (!rdy && query) {
        const hipError_t q = hipStreamQuery((hipStream_t)stream);
        if (q == hipSuccess)
            rdy = true;
        else if (q != hipErrorNotReady)
            return fail(N4_ERR_HIP, "HIP error %d (%s) while polling progress", (int)q, hipGetErrorString(q));
    }
#else
    (void)stream;
    (void)query;
    rdy = true;  // host emulation: every enqueued kernel has run
#endif
    *ready = rdy;
    *all_done = all;
    return N4_OK;
}

// Identity of a captured step chunk: everything its launches read from the
// host side (pointers, sizes, run parameters, chunk length).
struct GraphKey {
    std::vector<unsigned char> b;
    template <class T>
    void add(const T& v) {
        const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
        b.insert(b.end(), p, p + sizeof v);
    }
    void add_rp(const RunParams& r) {
        add(r.twiddle);
        add(r.nbins);
        add(r.P);
        add(r.log2P);
        add(r.offset);
        add(r.noise);
        add(r.fwhm);
        add(r.thr);
        add(r.eps);
        add(r.low);
        add(r.low_f);
        add(r.fixed_iters);
        add(r.prec);
        add(r.states);
        add(r.descs);
        add(r.nlevels);
        for (int l = 0; l < kMaxLevels; ++l) add(r.max_iter[l]);
    }
};

#ifndef N4_EMULATE
// Launches group g's chunk graph on `stream`, capturing it first (on the
// handle's capture stream, through `enqueue`) when the key changed.
template <class F>
static n4_status launch_step_graph(n4_handle* h, int g, const GraphKey& key, F&& enqueue, void* stream) {
    n4_handle::StepGraph& sg = h->sgraph[g];
    if (!sg.exec || sg.key != key.b) {
        if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
        sg.exec = nullptr;
        sg.key.clear();
        if (!h->cap_stream) N4_CHECK_HIP(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
        N4_CHECK_HIP(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
        const n4_status s = enqueue((void*)h->cap_stream);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(h->cap_stream, &graph);
        if (s != N4_OK || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return s != N4_OK ? s : fail(N4_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
        }
        const hipError_t ei = hipGraphInstantiate(&sg.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            sg.exec = nullptr;
            return fail(N4_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
        }
        sg.key = key.b;
    }
    N4_CHECK_HIP(hipGraphLaunch(sg.exec, (hipStream_t)stream));
    return N4_OK;
}
#endif

// Shrink classes of the prep kernels, in the order a group's volumes are
// laid out (each class is one prep launch per group).
static const int kPrepClasses[5] = {1, 2, 3, 4, 0};

struct GroupRun {
    int v0 = 0, v1 = 0;    // volumes
    int ws0 = 0, ws1 = 0;  // working slabs
    int fs0 = 0, fs1 = 0;  // recon tiles
    int ps0 = 0, ps1 = 0;  // prep tiles
    // prep tiles by shrink class (kPrepClasses order): [pc[i], pc[i + 1])
    int pc[6] = {0, 0, 0, 0, 0, 0};
    void* stream = nullptr;
    int it = 0;            // steps enqueued
    bool finished = false;
};

// Persistent path: team sizes (workgroups per volume) for a batch, or
// false with the reason when the batch is not eligible.  Every workgroup
// of the launch must be resident at once, so the teams share the device's
// resident-workgroup capacity for the kernel's LDS size, in proportion to
// the volumes' working voxels (at least one workgroup, at most one per
// tile, per volume).
static bool plan_teams(const std::vector<VolIn>& in, const n4_params* p, const Schedule& sch, std::vector<int>* teams,
                       std::string* why) {
    const int nv = (int)in.size();
    const int L = (int)p->number_of_fitting_levels;
#ifdef N4_EMULATE
    if (p->precision != N4_PREC_FP32) {
        *why = "the host emulation runs the FP32 fit";
        return false;
    }
#endif
    const RunParams rp = make_run_params(p);
    if (rp.prec != 0)
        for (int d = 1; d < 3; ++d)
            if (level_ncpt((int)p->ncpt[d], L - 1) > 16) {
                *why = "the MFMA fit of the persistent path covers up to 16 control points along y and z";
                return false;
            }
    size_t lds = 0;
    std::vector<int> tiles(nv);
    std::vector<double> w(nv);
    for (int v = 0; v < nv; ++v) {
        VolDesc D;
        memset(&D, 0, sizeof D);
        for (int d = 0; d < 3; ++d) D.W[d] = (in[v].F[d] + in[v].s - 1) / in[v].s;
        D.nlevels = L;
        for (int l = 0; l < L; ++l) {
            for (int d = 0; d < 3; ++d) D.lev[l].C[d] = level_ncpt((int)p->ncpt[d], l);
            D.lev[l].L = D.lev[l].C[0] * D.lev[l].C[1] * D.lev[l].C[2];
        }
        int rows, tx;
        step_tiles(D.W, true, sch.tile_vox, &rows, &tx);
        D.wtile_rows = rows;
        lds = std::max(lds, lds_persist(D, rp.nbins, rp.P, rp.prec));
        tiles[v] = D.W[0] * tx;
        w[v] = (double)D.W[0] * D.W[1] * D.W[2];
    }
    if (lds > kLdsCap) {
        *why = "a volume's tables exceed one workgroup's LDS";
        return false;
    }
    int per_cu = 0, cus = 0;
    if (persist_capacity(lds, rp.prec, &per_cu, &cus) != 0 || per_cu < 1 || cus < 1) {
        *why = "occupancy query failed";
        return false;
    }
    per_cu = std::min(per_cu, sch.team_per_cu > 0 ? sch.team_per_cu : 4);
    const int cap = per_cu * cus;
    if (nv > cap) {
        *why = "more volumes than resident workgroups";
        return false;
    }
    double tot = 0.0;
    for (int v = 0; v < nv; ++v) tot += w[v];
    teams->assign(nv, 1);
    const int spare = cap - nv;
    for (int v = 0; v < nv; ++v) {
        const int extra = tot > 0.0 ? (int)floor((double)spare * w[v] / tot) : 0;
        (*teams)[v] = std::max(1, std::min(tiles[v], 1 + extra));
    }
    return true;
}

#define GLAUNCH(id, expr)                  \
    do {                                   \
        ProfScope ps_(h, (id), gs);        \
        LAUNCH(expr);                      \
    } while (0)

// Identity of a correction plan apart from the caller's buffer addresses
// (see n4_handle::plan_key): every input of build_plan that shapes the
// layout or the descriptors.
static std::vector<unsigned char> correction_plan_key(const n4_handle* h, const std::vector<VolIn>& in,
                                                      const n4_params* p, int tile_vox,
                                                      const std::vector<int>* teams) {
    std::vector<unsigned char> k;
    auto add = [&](const auto& x) {
        const unsigned char* b = reinterpret_cast<const unsigned char*>(&x);
        k.insert(k.end(), b, b + sizeof x);
    };
    add(in.size());
    for (const VolIn& v : in) {
        add(v.mask_kind);
        add(v.s);
        for (int d = 0; d < 3; ++d) add(v.F[d]);
    }
    add(p->number_of_histogram_bins);
    add(p->wiener_filter_noise);
    add(p->bias_field_fwhm);
    for (int d = 0; d < 3; ++d) add(p->ncpt[d]);
    add(p->number_of_fitting_levels);
    for (int l = 0; l < N4_MAX_LEVELS; ++l) add(p->max_iterations[l]);
    add(p->convergence_threshold);
    add(p->low_value);
    add(p->shrink_factor);
    add(p->fit_eps);
    add(p->precision);
    add(p->flags);
    add(tile_vox);
    add(teams ? teams->size() : (size_t)0);
    if (teams)
        for (int t : *teams) add(t);
    add(h->dev_flags);
    add(h->flag_cap);
    add(h->prof_mask);
    add(h->arena);
    return k;
}

// noalloc (n4_correct_noalloc): the plan lives in the device's scratch
// array; one group on the caller's stream, every step enqueued up front
// (no progress polling, so no pinned host words), direct launches (no graph
// instantiation), no events, no kernel timing.
static n4_status run_batch(n4_handle* h, const std::vector<VolIn>& in_user, const n4_params* p, n4_report* reports,
                           void* stream, bool noalloc = false) {
    g_err.clear();
    const auto t_host0 = std::chrono::steady_clock::now();
    n4_status s = check_params(p);
    if (s != N4_OK) return s;
    for (const VolIn& v : in_user) {
        s = check_volume(v);
        if (s != N4_OK) return s;
    }
    if (in_user.empty()) return N4_OK;
    const int nv = (int)in_user.size();
    const Schedule sch = resolve_schedule(h, nv);
    std::vector<int> teams_user;
    bool teams_path = false;
    if (sch.path != (int)N4_PATH_STEPS && !noalloc) {
        std::string why;
        teams_path = plan_teams(in_user, p, sch, &teams_user, &why);
        if (!teams_path && sch.path == (int)N4_PATH_TEAMS)
            return fail(N4_ERR_UNSUPPORTED, "persistent team path not available for this batch: %s", why.c_str());
    }
    const int G = (teams_path || noalloc) ? 1 : sch.groups;
    // kernel timing creates events: off for the allocation-free path
    struct ProfMaskGuard {
        n4_handle* h;
        uint32_t saved;
        ~ProfMaskGuard() { h->prof_mask = saved; }
    } prof_guard{h, h->prof_mask};
    if (noalloc) h->prof_mask = 0;
    // order volumes by group: largest first (working voxels), each onto the
    // group with the least work so far
    std::vector<int> order(nv), gid(nv, 0);
    {
        std::vector<std::pair<double, int>> sz(nv);
        for (int v = 0; v < nv; ++v) {
            const VolIn& a = in_user[v];
            double w = 1.0;
            for (int d = 0; d < 3; ++d) w *= (double)((a.F[d] + a.s - 1) / a.s);
            sz[v] = {-w, v};
        }
        std::sort(sz.begin(), sz.end());
        std::vector<double> load(G, 0.0);
        for (auto& e : sz) {
            const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            gid[e.second] = g;
            load[g] += -e.first;
        }
        // within a group, volumes of one prep shrink class are contiguous
        int k = 0;
        for (int g = 0; g < G; ++g)
            for (int c : kPrepClasses)
                for (int v = 0; v < nv; ++v)
                    if (gid[v] == g && prep_class(in_user[v].s) == c) order[k++] = v;
    }
    std::vector<VolIn> in(nv);
    std::vector<int> teams(teams_path ? nv : 0);
    for (int k = 0; k < nv; ++k) {
        in[k] = in_user[order[k]];
        if (teams_path) teams[k] = teams_user[order[k]];
    }
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipSetDevice(h->device));
#endif
    if (!noalloc) {
        s = ensure_flags(h, nv);
        if (s != N4_OK) return s;
        for (int v = 0; v < nv; ++v) h->pinned_flags[v] = 0;
    }
#ifndef N4_EMULATE
    for (int g = 1; g < G; ++g)
        if (!h->gstream[g]) N4_CHECK_HIP(hipStreamCreateWithFlags(&h->gstream[g], hipStreamNonBlocking));
    struct EvPair {
        hipEvent_t a = nullptr, b = nullptr;
        ~EvPair() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    if (!noalloc) {
        N4_CHECK_HIP(hipEventCreate(&ev.a));
        N4_CHECK_HIP(hipEventCreate(&ev.b));
        N4_CHECK_HIP(hipEventRecord(ev.a, (hipStream_t)stream));
    }
#endif
    Plan P;
    std::vector<unsigned char> pkey;
    if (!noalloc) pkey = correction_plan_key(h, in, p, sch.tile_vox, teams_path ? &teams : nullptr);
    const bool plan_reused = !noalloc && h->plan_cached && pkey == h->plan_key;
    if (plan_reused) {
        // the plan of the last call, whose upload image is still in the stage:
        // only the caller's buffers change
        P = *h->plan_cached;
        h->last.valid = false;  // its run records are reused
        for (int v = 0; v < nv; ++v) {
            VolDesc& D = P.desc[v];
            D.im = in[v].im;
            D.mask = in[v].mask;
            D.out_norm = in[v].out_norm;
            D.out_bias = in[v].out_bias;
            D.out_logbias = in[v].out_logbias;
            D.out_lattice = in[v].out_lattice;
        }
        stage_wait(h);
        memcpy(h->stage + P.off_desc, P.desc.data(), sizeof(VolDesc) * nv);
    } else {
        s = build_plan(h, in, p, &P, true, true, sch.tile_vox, teams_path ? &teams : nullptr, noalloc);
        if (s != N4_OK) return s;
        if (!noalloc) {
            h->plan_cached.reset(new Plan(P));
            h->plan_key = pkey;
        }
    }
    s = check_lds(P, !teams_path, true);
    if (s != N4_OK) return s;
    if (plan_reused) {
        // the arena still holds the rest of the upload region (slab lists,
        // twiddles, axis tables; no kernel writes them): descriptors, and the
        // volume states back to their initial (zero) image
        s = stage_upload(h, P.base + P.off_desc, P.off_desc, sizeof(VolDesc) * nv, stream);
        if (s == N4_OK) s = stage_upload(h, P.base + P.off_state, P.off_state, sizeof(VolState) * nv, stream);
    } else {
        s = stage_upload(h, P.base, 0, P.upload_bytes, stream);
    }
    if (s != N4_OK) return s;
    VolDesc* dd = dev_desc(h, P);
    Slab* wsl = dev_wslab(h, P);
    Slab* fsl = dev_fslab(h, P);
    const RunParams rp = P.rp;
    // group ranges (volumes and their working slabs, prep tiles and recon
    // tiles are contiguous)
    std::vector<GroupRun> gr(G);
    {
        int v = 0, ws = 0, fs = 0, ps = 0;
        for (int g = 0; g < G; ++g) {
            GroupRun& R = gr[g];
            R.v0 = v;
            R.ws0 = ws;
            R.fs0 = fs;
            R.ps0 = ps;
            int ci = 0;
            R.pc[0] = ps;
            while (v < nv && gid[order[v]] == g) {
                while (ci < 4 && prep_class(P.desc[v].s) != kPrepClasses[ci]) R.pc[++ci] = ps;
                ws += P.desc[v].W[0] * P.desc[v].wtiles_x;
                fs += P.recon_tiles[v];
                ps += P.desc[v].W[0] * P.desc[v].prep_tiles_x;
                ++v;
            }
            while (ci < 5) R.pc[++ci] = ps;
            R.v1 = v;
            R.ws1 = ws;
            R.fs1 = fs;
            R.ps1 = ps;
#ifndef N4_EMULATE
            R.stream = g == 0 ? stream : (void*)h->gstream[g];
#else
            R.stream = stream;
#endif
        }
    }
    // Every volume runs its own level/iteration state machine on the device;
    // the host enqueues identical steps (hist -> wiener -> fit -> eval) per
    // group, round-robin, in chunks of `chunk` steps (one captured graph
    // each, or direct launches).  A group whose pipeline is full is skipped
    // until its progress allows more, the host spins only when no group can
    // advance, and a group ends once all its volumes are finished.  Every
    // volume needs at most sum(max_iter + 1) + 1 steps; steps past a
    // volume's end are skips.  Graphs are not used while kernels are timed
    // with events.
#ifndef N4_EMULATE
    const bool use_graph = sch.graph_steps > 0 && h->prof_mask == 0 && !noalloc;
#else
    const bool use_graph = false;
#endif
    const int chunk = sch.graph_steps > 0 ? sch.graph_steps : 1;
    const int depth = chunk + kPipeline - 1;
    int max_steps = 1;
    for (int l = 0; l < P.nlevels; ++l) max_steps += P.max_iter[l] + 1;
    size_t lds_fit_all = 0, lds_eval_all = 0;
    for (int l = 0; l < P.nlevels; ++l) {
        lds_fit_all = std::max(lds_fit_all, P.lds_fit_max[l]);
        lds_eval_all = std::max(lds_eval_all, sch.mfma_eval ? P.lds_eval_max[l] : P.lds_eval32_max[l]);
    }
    const int mt = P.mt[P.nlevels - 1], nt = P.nt[P.nlevels - 1];
    RunParams rp_eval = rp;  // the evaluation kernel's precision
    if (!sch.mfma_eval) rp_eval.prec = 0;
    int lmax_all = 1;
    for (int v = 0; v < nv; ++v) lmax_all = std::max(lmax_all, P.desc[v].lmax);
    const int upd_chunks = (lmax_all + kBlock - 1) / kBlock;
#ifndef N4_EMULATE
    if (G > 1) {
        N4_CHECK_HIP(hipEventRecord(h->fork_ev, (hipStream_t)stream));
        for (int g = 1; g < G; ++g) N4_CHECK_HIP(hipStreamWaitEvent(h->gstream[g], h->fork_ev, 0));
    }
#endif
    // each group prepares its own volumes on its own stream, so the
    // full-resolution pass of one group overlaps the steps of the others
    for (int g = 0; g < G; ++g) {
        GroupRun& R = gr[g];
        void* gs = R.stream;
        for (int c = 0; c < 5; ++c)
            if (R.pc[c + 1] > R.pc[c])
                GLAUNCH(N4_K_PREP,
                        launch_prep(dd, dev_pslab(h, P) + R.pc[c], R.pc[c + 1] - R.pc[c], kPrepClasses[c], rp, gs));
        GLAUNCH(N4_K_LEVEL, launch_level(dd + R.v0, R.v1 - R.v0, 0, rp, gs));
    }
    if (teams_path) {
        // one launch runs every volume's level/iteration loop
        void* gs = stream;
        GLAUNCH(N4_K_TEAMS, launch_persist(dd, wsl, (const PWork*)(P.base + P.off_work), P.n_work,
                                           rp, P.lds_persist, gs));
        GLAUNCH(N4_K_RECON, launch_recon(dd, fsl, P.n_fslab, P.nlevels - 1, rp, P.lds_recon, gs));
        for (GroupRun& R : gr) R.finished = true;
    }
    // One group's next move: wait (0), enqueue a chunk of steps (1), or,
    // once every volume of the group has finished, its reconstruction and
    // join event (2); < 0: failed (the status is in *err).
    auto advance_group = [&](int g, bool query, n4_status* err) -> int {
        GroupRun& R = gr[g];
        void* gs = R.stream;
        bool end = R.it >= max_steps;
        n4_status s2 = N4_OK;
        if (R.it > 0 && !end && !noalloc) {
            bool ready = false, all_done = false;
            s2 = check_progress(h, R.v0, R.v1, 0, R.it, depth, gs, query, &ready, &all_done);
            if (s2 != N4_OK) return *err = s2, -1;
            if (!ready) return 0;
            end = all_done;
        }
        const int nvg = R.v1 - R.v0, nws = R.ws1 - R.ws0;
        const VolDesc* gv = dd + R.v0;
        const Slab* gsl = wsl + R.ws0;
        auto body = [&]() -> n4_status {
            if (!end) {
                auto enqueue = [&](void* gs) -> n4_status {
                    for (int k = 0; k < chunk; ++k) {
                        GLAUNCH(N4_K_HIST, launch_hist(dd, gsl, nws, 0, rp, nullptr, P.lds_hist, gs));
                        GLAUNCH(N4_K_WIENER, launch_wiener(gv, nvg, rp, P.lds_wiener, gs));
                        GLAUNCH(N4_K_FIT, launch_fit(dd, gsl, nws, 0, 0, kFitStep, rp, lds_fit_all, mt, nt, gs));
                        GLAUNCH(N4_K_UPDATE, launch_update(gv, nvg, upd_chunks, rp, gs));
                        GLAUNCH(N4_K_EVAL, launch_eval(dd, gsl, nws, rp_eval, lds_eval_all, mt, nt, gs));
                    }
                    return N4_OK;
                };
                n4_status s3 = N4_OK;
                if (use_graph) {
#ifndef N4_EMULATE
                    GraphKey key;
                    key.add(dd);
                    key.add(gv);  // the group's first volume (k_wiener, k_update)
                    key.add(gsl);
                    key.add(nws);
                    key.add(nvg);
                    key.add(P.lds_hist);
                    key.add(P.lds_wiener);
                    key.add(lds_fit_all);
                    key.add(lds_eval_all);
                    key.add(mt);
                    key.add(nt);
                    key.add(chunk);
                    key.add(upd_chunks);
                    key.add(rp_eval.prec);
                    key.add_rp(rp);
                    s3 = launch_step_graph(h, g, key, enqueue, gs);
#endif
                } else {
                    s3 = enqueue(gs);
                }
                if (s3 != N4_OK) return s3;
                R.it += chunk;
                return N4_OK;
            }
            // (the final step wrote each volume's output lattice)
            GLAUNCH(N4_K_RECON, launch_recon(dd, fsl + R.fs0, R.fs1 - R.fs0, P.nlevels - 1, rp, P.lds_recon, gs));
#ifndef N4_EMULATE
            if (g > 0) {
                N4_CHECK_HIP(hipEventRecord(h->join_ev[g], (hipStream_t)gs));
                N4_CHECK_HIP(hipStreamWaitEvent((hipStream_t)stream, h->join_ev[g], 0));
            }
#endif
            R.finished = true;
            return N4_OK;
        };
        s2 = body();
        if (s2 != N4_OK) return *err = s2, -1;
        return R.finished ? 2 : 1;
    };
    // one host thread drives every group, round robin (a thread per group
    // measured no faster: the host's launch cost is not the bound)
    int remaining = teams_path ? 0 : G;
    uint64_t idle_rounds = 0;
    while (remaining > 0) {
        bool advanced = false;
        for (int g = 0; g < G; ++g) {
            if (gr[g].finished) continue;
            n4_status e = N4_OK;
            const int r = advance_group(g, (idle_rounds & 63) == 63, &e);
            if (r < 0) return e;
            if (r == 0) continue;
            advanced = true;
            if (r == 2) --remaining;
        }
        if (advanced) {
            idle_rounds = 0;
        } else {
            ++idle_rounds;
#if defined(__x86_64__) && !defined(N4_EMULATE)
            __builtin_ia32_pause();
#endif
        }
    }
    std::vector<VolState> st(nv);
    D2H(st.data(), dev_state(h, P), sizeof(VolState) * nv, stream);
    std::vector<uint32_t> tctr(teams_path ? 2 * nv : 0);
    std::vector<uint64_t> tph(teams_path && P.desc[0].tphase ? (size_t)kPhases * nv : 0);
    if (teams_path)
        for (int v = 0; v < nv; ++v) {
            D2H(&tctr[2 * v], P.desc[v].team_ctr, 2 * sizeof(uint32_t), stream);
            if (!tph.empty()) D2H(&tph[(size_t)kPhases * v], P.desc[v].tphase, sizeof(uint64_t) * kPhases, stream);
        }
    float ms = 0.0f;
#ifndef N4_EMULATE
    if (!noalloc) {
        N4_CHECK_HIP(hipEventRecord(ev.b, (hipStream_t)stream));
        N4_CHECK_HIP(hipEventSynchronize(ev.b));
        N4_CHECK_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
    } else {
        SYNC(stream);
        ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();
    }
    N4_CHECK_HIP(hipGetLastError());
#else
    (void)t_host0;
#endif
    prof_collect(h);
    if (!tph.empty()) {
        // the leaders' phase clocks: total over the volumes, launches = steps
        static const int kPhaseId[kPhSteps] = {N4_K_T_WAIT, N4_K_T_REDUCE, N4_K_T_HIST, N4_K_T_WIENER,
                                               N4_K_T_FIT, N4_K_T_UPDATE, N4_K_T_EVAL, N4_K_T_TRANS};
        for (int v = 0; v < nv; ++v)
            for (int k = 0; k < kPhSteps; ++k) {
                h->prof_ms[kPhaseId[k]] += 1e3 * (double)tph[(size_t)kPhases * v + k] / kClockHz;
                h->prof_n[kPhaseId[k]] += tph[(size_t)kPhases * v + kPhSteps];
            }
    }
    for (int v = 0; v < (int)tctr.size() / 2; ++v)
        if (tctr[2 * v + 1])
            return fail(N4_ERR_INTERNAL, "persistent path: a team barrier of volume %d timed out", order[v]);
    {
        n4_handle::LastRun& R = h->last;
        R.nlevels = P.nlevels;
        R.max_iter = P.max_iter;
        R.cvtrace.assign(nv, nullptr);
        R.trace_cap.assign(nv, 0);
        R.lvlat.assign(nv, nullptr);
        R.iters.assign(nv, {});
        R.ncpt.assign(nv, {});
        for (int k = 0; k < nv; ++k) {
            const int v = order[k];
            R.cvtrace[v] = P.desc[k].cvtrace;
            R.trace_cap[v] = P.desc[k].trace_cap;
            R.lvlat[v] = P.desc[k].lvlat;
            for (int l = 0; l < P.nlevels; ++l) {
                R.iters[v][l] = (uint32_t)st[k].iters[l];
                for (int d = 0; d < 3; ++d) R.ncpt[v][3 * l + d] = P.desc[k].lev[l].C[d];
            }
        }
        R.in_scratch = noalloc;
        R.scratch_gen = noalloc ? scratch_generation(h->device) : 0;
        R.valid = true;
    }
    if (reports)
        for (int k = 0; k < nv; ++k) {
            const int v = order[k];
            n4_report& r = reports[v];
            memset(&r, 0, sizeof r);
            for (int l = 0; l < P.nlevels; ++l) {
                r.iterations[l] = (uint32_t)st[k].iters[l];
                r.last_convergence[l] = st[k].lastcv[l];
            }
            r.n_mask_full = st[k].n_mask_full;
            r.n_mask_work = st[k].n_mask_work;
            for (int d = 0; d < 3; ++d) {
                r.work_dims[d] = (uint32_t)P.desc[k].W[d];
                r.final_ncpt[d] = (uint32_t)P.desc[k].lev[P.nlevels - 1].C[d];
            }
            r.elapsed_ms = ms;
        }
    return N4_OK;
}

// ================================================================ C ABI ==
extern "C" {

const char* n4_version(void) {
#ifdef N4_EMULATE
    return "n4amd 0.1.0 (host emulation of the kernel bodies; test use only)";
#else
    return "n4amd 0.1.0 (gfx950)";
#endif
}

const char* n4_status_string(n4_status s) {
    switch (s) {
        case N4_OK: return "ok";
        case N4_ERR_INVALID_ARG: return "invalid argument";
        case N4_ERR_HIP: return "HIP runtime error";
        case N4_ERR_NO_MEMORY: return "out of device memory";
        case N4_ERR_UNSUPPORTED: return "unsupported";
        case N4_ERR_INTERNAL: return "internal error";
    }
    return "unknown status";
}

const char* n4_last_error(void) { return g_err.c_str(); }

n4_status n4_params_default(n4_params* p) {
    if (!p) return fail(N4_ERR_INVALID_ARG, "params is NULL");
    memset(p, 0, sizeof *p);
    p->number_of_histogram_bins = 200;
    p->wiener_filter_noise = 0.01;
    p->bias_field_fwhm = 0.15;
    p->ncpt[0] = p->ncpt[1] = p->ncpt[2] = 4;
    p->number_of_fitting_levels = 4;
    for (int l = 0; l < N4_MAX_LEVELS; ++l) p->max_iterations[l] = 50;
    p->convergence_threshold = 0.001;
    p->low_value = 10.0;
    p->shrink_factor = 1;
    p->fit_eps = 1e-3;
    p->precision = N4_PREC_FP32;
    p->flags = 0;
    return N4_OK;
}

n4_status n4_final_ncpt(const n4_params* p, uint32_t out[3]) {
    if (!p || !out) return fail(N4_ERR_INVALID_ARG, "NULL argument");
    if (p->number_of_fitting_levels < 1 || p->number_of_fitting_levels > N4_MAX_LEVELS)
        return fail(N4_ERR_INVALID_ARG, "number_of_fitting_levels out of range");
    for (int d = 0; d < 3; ++d) {
        if (p->ncpt[d] < 4) return fail(N4_ERR_INVALID_ARG, "ncpt < 4");
        out[d] = (uint32_t)level_ncpt((int)p->ncpt[d], (int)p->number_of_fitting_levels - 1);
    }
    return N4_OK;
}

n4_status n4_device_count(int* count) {
    if (!count) return fail(N4_ERR_INVALID_ARG, "count is NULL");
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipGetDeviceCount(count));
#else
    *count = 1;
#endif
    return N4_OK;
}

n4_status n4_create(n4_handle** out, int device) {
    if (!out) return fail(N4_ERR_INVALID_ARG, "handle pointer is NULL");
    *out = nullptr;
#ifndef N4_EMULATE
    int n = 0;
    N4_CHECK_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(N4_ERR_INVALID_ARG, "device %d not present (%d devices)", device, n);
    N4_CHECK_HIP(hipSetDevice(device));
#endif
    n4_handle* h = new n4_handle();
    h->device = device;
    {
        // the device's allocation-free scratch, resolved with the device current
        const n4_status ss = resolve_scratch(device, &h->scratch);
        if (ss != N4_OK) {
            delete h;
            return ss;
        }
    }
#ifndef N4_EMULATE
    // HIP maps streams onto the process's few hardware queues in creation
    // order: the group streams are created first, in group order, so that
    // the first groups (with the null stream) own a queue each; the capture
    // stream, which never executes work, comes last.
    for (int g = 1; g < kMaxGroups; ++g) {
        const hipError_t e = hipStreamCreateWithFlags(&h->gstream[g], hipStreamNonBlocking);
        if (e != hipSuccess) {
            n4_destroy(h);
            return fail(N4_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
        }
    }
    const hipError_t e = hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        n4_destroy(h);
        return fail(N4_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
    }
#endif
    *out = h;
    return N4_OK;
}

n4_status n4_destroy(n4_handle* h) {
    if (!h) return N4_OK;
#ifndef N4_EMULATE
    (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
#endif
    dev_free(h->arena);
#ifndef N4_EMULATE
    if (h->pinned_flags) (void)hipHostFree(h->pinned_flags);
    if (h->stage_pinned) (void)hipHostFree(h->stage_pinned);
    if (h->stage_ev) (void)hipEventDestroy(h->stage_ev);
    if (h->events_init) {
        (void)hipEventDestroy(h->fork_ev);
        for (int g = 0; g < kMaxGroups; ++g) (void)hipEventDestroy(h->join_ev[g]);
    }
    for (int g = 0; g < kMaxGroups; ++g)
        if (h->gstream[g]) (void)hipStreamDestroy(h->gstream[g]);
    for (int g = 0; g < kMaxGroups; ++g)
        if (h->sgraph[g].exec) (void)hipGraphExecDestroy(h->sgraph[g].exec);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
#else
    free(h->pinned_flags);
#endif
    delete h;
    return N4_OK;
}

n4_status n4_correct(n4_handle* h, const float* d_im, const void* d_mask, uint32_t mask_kind, uint32_t fx,
                     uint32_t fy, uint32_t fz, const n4_params* p, float* d_im_normalized, float* d_lattice,
                     float* d_bias_field, float* d_log_bias_field, n4_report* report, void* stream) {
    if (!h) return fail(N4_ERR_INVALID_ARG, "handle is NULL");
    if (!p) return fail(N4_ERR_INVALID_ARG, "params is NULL");
    VolIn v{d_im, d_mask, (int)mask_kind, {(int)fx, (int)fy, (int)fz}, (int)p->shrink_factor,
            d_im_normalized, d_bias_field, d_log_bias_field, d_lattice};
    return run_batch(h, std::vector<VolIn>{v}, p, report, stream);
}

n4_status n4_correct_noalloc(n4_handle* h, const float* d_im, const void* d_mask, uint32_t mask_kind, uint32_t fx,
                             uint32_t fy, uint32_t fz, const n4_params* p, float* d_im_normalized, float* d_lattice,
                             float* d_bias_field, float* d_log_bias_field, n4_report* report, void* stream) {
    if (!h) return fail(N4_ERR_INVALID_ARG, "handle is NULL");
    if (!p) return fail(N4_ERR_INVALID_ARG, "params is NULL");
    VolIn v{d_im, d_mask, (int)mask_kind, {(int)fx, (int)fy, (int)fz}, (int)p->shrink_factor,
            d_im_normalized, d_bias_field, d_log_bias_field, d_lattice};
    return run_batch(h, std::vector<VolIn>{v}, p, report, stream, true);
}

// The run record of volume `volume` of the last correction call, checked.
static n4_status last_run(n4_handle* h, uint32_t volume, const n4_handle::LastRun** out) {
    if (!h) return fail(N4_ERR_INVALID_ARG, "handle is NULL");
    const n4_handle::LastRun& R = h->last;
    if (!R.valid)
        return fail(N4_ERR_INVALID_ARG, "no run record: no correction ran on this handle since its last plan "
                    "(stage calls reuse the same memory)");
    if (R.in_scratch && R.scratch_gen != scratch_generation(h->device))
        return fail(N4_ERR_INVALID_ARG, "no run record: another allocation-free call ran on device %d since",
                    h->device);
    if (volume >= R.cvtrace.size())
        return fail(N4_ERR_INVALID_ARG, "volume %u: the last call had %zu", volume, R.cvtrace.size());
    *out = &R;
    return N4_OK;
}

n4_status n4_last_cv_trace(n4_handle* h, uint32_t volume, double* out, uint32_t capacity, uint32_t* count) {
    const n4_handle::LastRun* R = nullptr;
    n4_status s = last_run(h, volume, &R);
    if (s != N4_OK) return s;
    if (!count) return fail(N4_ERR_INVALID_ARG, "count is NULL");
    if (capacity > 0 && !out) return fail(N4_ERR_INVALID_ARG, "out is NULL");
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipSetDevice(h->device));
#endif
    // level l's values sit at the sum of the earlier levels' maxima
    uint32_t n = 0;
    int off = 0;
    for (int l = 0; l < R->nlevels; ++l) {
        const uint32_t it = R->iters[volume][l];
        const int avail = std::max(0, std::min<int>((int)it, R->trace_cap[volume] - off));
        const uint32_t take = std::min<uint32_t>((uint32_t)avail, capacity > n ? capacity - n : 0u);
        if (take > 0) {
#ifndef N4_EMULATE
            N4_CHECK_HIP(hipMemcpy(out + n, R->cvtrace[volume] + off, sizeof(double) * take, hipMemcpyDeviceToHost));
#else
            memcpy(out + n, R->cvtrace[volume] + off, sizeof(double) * take);
#endif
        }
        n += (uint32_t)avail;
        off += R->max_iter[l];
    }
    *count = n;
    return N4_OK;
}

n4_status n4_last_level_lattice(n4_handle* h, uint32_t volume, uint32_t level, float* out, uint64_t capacity) {
    const n4_handle::LastRun* R = nullptr;
    n4_status s = last_run(h, volume, &R);
    if (s != N4_OK) return s;
    if (level >= (uint32_t)R->nlevels)
        return fail(N4_ERR_INVALID_ARG, "level %u: the last call had %d", level, R->nlevels);
    size_t off = 0, L = 0;
    for (uint32_t l = 0; l <= level; ++l) {
        const int32_t* c = &R->ncpt[volume][3 * l];
        L = (size_t)c[0] * c[1] * c[2];
        if (l < level) off += L;
    }
    if (capacity < L) return fail(N4_ERR_INVALID_ARG, "level %u lattice has %zu values, capacity %llu", level, L,
                                  (unsigned long long)capacity);
    if (!out) return fail(N4_ERR_INVALID_ARG, "out is NULL");
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipSetDevice(h->device));
    N4_CHECK_HIP(hipMemcpy(out, R->lvlat[volume] + off, sizeof(float) * L, hipMemcpyDeviceToHost));
#else
    memcpy(out, R->lvlat[volume] + off, sizeof(float) * L);
#endif
    return N4_OK;
}

n4_status n4_correct_batch(n4_handle* h, const n4_volume* vols, uint32_t nvol, const n4_params* p,
                           n4_report* reports, void* stream) {
    if (!h) return fail(N4_ERR_INVALID_ARG, "handle is NULL");
    if (!p) return fail(N4_ERR_INVALID_ARG, "params is NULL");
    if (nvol > 0 && !vols) return fail(N4_ERR_INVALID_ARG, "volume list is NULL");
    std::vector<VolIn> in;
    in.reserve(nvol);
    for (uint32_t i = 0; i < nvol; ++i) {
        const n4_volume& a = vols[i];
        VolIn v{a.d_im, a.d_mask, (int)a.mask_kind, {(int)a.dims[0], (int)a.dims[1], (int)a.dims[2]},
                (int)(a.shrink_factor ? a.shrink_factor : p->shrink_factor), a.d_im_normalized, a.d_bias_field,
                a.d_log_bias_field, a.d_lattice};
        in.push_back(v);
    }
    return run_batch(h, in, p, reports, stream);
}

n4_status n4_get_tuning(n4_handle* h, n4_tuning* out) {
    if (!h || !out) return fail(N4_ERR_INVALID_ARG, "NULL argument");
    *out = h->tune;
    return N4_OK;
}

n4_status n4_set_tuning(n4_handle* h, const n4_tuning* t) {
    if (!h) return fail(N4_ERR_INVALID_ARG, "handle is NULL");
    if (!t) {
        h->tune = n4_tuning{};
        return N4_OK;
    }
    if (t->groups > (uint32_t)kMaxGroups) return fail(N4_ERR_INVALID_ARG, "groups > %d", kMaxGroups);
    if (t->step_tile_voxels != 0 && t->step_tile_voxels < 16)
        return fail(N4_ERR_INVALID_ARG, "step_tile_voxels < 16");
    if (t->graph_steps != N4_GRAPH_OFF && t->graph_steps > (uint32_t)kMaxGraphSteps)
        return fail(N4_ERR_INVALID_ARG, "graph_steps > %d", kMaxGraphSteps);
    if (t->eval_kernel > N4_EVAL_MFMA) return fail(N4_ERR_INVALID_ARG, "unknown eval_kernel %u", t->eval_kernel);
    if (t->path > N4_PATH_TEAMS) return fail(N4_ERR_INVALID_ARG, "unknown path %u", t->path);
    if (t->team_wgs_per_cu > 32) return fail(N4_ERR_INVALID_ARG, "team_wgs_per_cu > 32");
    h->tune = *t;
    h->tune.ms = 0.0f;
    return N4_OK;
}

// Exhaustive search over a small schedule grid, timed on the caller's own
// volumes: one untimed run per candidate (captures its graphs), then the
// best of two timed runs (host wall clock around the synchronous call).
n4_status n4_autotune_batch(n4_handle* h, const n4_volume* vols, uint32_t nvol, const n4_params* p,
                            n4_tuning* best, void* stream) {
    if (!h || !p || !best) return fail(N4_ERR_INVALID_ARG, "NULL argument");
    if (nvol == 0 || !vols) return fail(N4_ERR_INVALID_ARG, "no volumes to tune on");
    const n4_tuning saved = h->tune;
    std::vector<n4_tuning> cand;
    for (uint32_t g : {1u, 2u, 4u, 8u}) {
        if (g > nvol || g > (uint32_t)kMaxGroups) continue;
        for (uint32_t t : {4096u, 8192u, 16384u})
            for (uint32_t k : {N4_GRAPH_OFF, 8u, 16u}) {
                cand.push_back(n4_tuning{g, t, k, N4_EVAL_FP32, N4_PATH_STEPS, 0, 0.0f});
                if (p->precision != N4_PREC_FP32) cand.push_back(n4_tuning{g, t, k, N4_EVAL_MFMA, N4_PATH_STEPS, 0, 0.0f});
            }
    }
    // best wall time of `reps` timed runs after one untimed run; N4_AUTOTUNE_LOG=1
    // prints every measurement to stderr
    const bool log = env_int("N4_AUTOTUNE_LOG", 0) != 0;
    auto time_it = [&](n4_tuning& c, int reps) -> n4_status {
        h->tune = c;
        h->tune.ms = 0.0f;
        double ms_best = -1.0;
        for (int rep = 0; rep <= reps; ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            const n4_status s = n4_correct_batch(h, vols, nvol, p, nullptr, stream);
            const auto t1 = std::chrono::steady_clock::now();
            if (s != N4_OK) return s;
            const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            if (rep > 0 && (ms_best < 0.0 || ms < ms_best)) ms_best = ms;
        }
        c.ms = (float)ms_best;
        if (log)
            fprintf(stderr, "[n4 autotune] groups %u tile %u graph %d eval %s: %.3f ms (best of %d)\n", c.groups,
                    c.step_tile_voxels, c.graph_steps == N4_GRAPH_OFF ? 0 : (int)c.graph_steps,
                    c.eval_kernel == N4_EVAL_MFMA ? "mfma" : "fp32", ms_best, reps);
        return N4_OK;
    };
    for (n4_tuning& c : cand) {
        const n4_status s = time_it(c, 2);
        if (s != N4_OK) {
            h->tune = saved;
            return s;
        }
    }
    // the three fastest again, with more runs, against timing noise
    std::stable_sort(cand.begin(), cand.end(), [](const n4_tuning& a, const n4_tuning& b) { return a.ms < b.ms; });
    const size_t nfin = std::min<size_t>(3, cand.size());
    for (size_t i = 0; i < nfin; ++i) {
        const n4_status s = time_it(cand[i], 5);
        if (s != N4_OK) {
            h->tune = saved;
            return s;
        }
    }
    const n4_tuning top = *std::min_element(cand.begin(), cand.begin() + nfin,
                                            [](const n4_tuning& a, const n4_tuning& b) { return a.ms < b.ms; });
    h->tune = top;
    h->tune.ms = 0.0f;
    *best = top;
    return N4_OK;
}

n4_status n4_correct_host(n4_handle* h, const float* h_im, const float* h_mask, uint32_t fx, uint32_t fy,
                          uint32_t fz, const n4_params* p, float* h_im_normalized, float* h_lattice,
                          float* h_bias_field, n4_report* report) {
    if (!h || !p || !h_im || !h_im_normalized) return fail(N4_ERR_INVALID_ARG, "NULL argument");
    uint32_t nc[3];
    n4_status s = n4_final_ncpt(p, nc);
    if (s != N4_OK) return s;
    const size_t n = (size_t)fx * fy * fz;
    const size_t L = (size_t)nc[0] * nc[1] * nc[2];
    void *d_im = nullptr, *d_mask = nullptr, *d_out = nullptr, *d_lat = nullptr, *d_bias = nullptr;
    s = N4_OK;
    if ((s = n4_malloc(h, &d_im, n * 4)) != N4_OK) goto done;
    if (h_mask && (s = n4_malloc(h, &d_mask, n * 4)) != N4_OK) goto done;
    if ((s = n4_malloc(h, &d_out, n * 4)) != N4_OK) goto done;
    if ((s = n4_malloc(h, &d_lat, L * 4)) != N4_OK) goto done;
    if (h_bias_field && (s = n4_malloc(h, &d_bias, n * 4)) != N4_OK) goto done;
    if ((s = n4_memcpy_htod(h, d_im, h_im, n * 4)) != N4_OK) goto done;
    if (h_mask && (s = n4_memcpy_htod(h, d_mask, h_mask, n * 4)) != N4_OK) goto done;
    s = n4_correct(h, (const float*)d_im, d_mask, h_mask ? N4_MASK_F32 : N4_MASK_NONE, fx, fy, fz, p,
                   (float*)d_out, (float*)d_lat, (float*)d_bias, nullptr, report, nullptr);
    if (s != N4_OK) goto done;
    if ((s = n4_memcpy_dtoh(h, h_im_normalized, d_out, n * 4)) != N4_OK) goto done;
    if (h_lattice && (s = n4_memcpy_dtoh(h, h_lattice, d_lat, L * 4)) != N4_OK) goto done;
    if (h_bias_field && (s = n4_memcpy_dtoh(h, h_bias_field, d_bias, n * 4)) != N4_OK) goto done;
done:
    n4_free(h, d_im);
    n4_free(h, d_mask);
    n4_free(h, d_out);
    n4_free(h, d_lat);
    n4_free(h, d_bias);
    return s;
}

// ----------------------------------------------------------- stage calls --
// Flat stage plans reuse the engine's kernels on a 1-D array cut into
// kChunkVox-voxel slabs.
// hist: the histogram kernel will read x and m; it takes 16-byte loads of
// aligned voxel quads, so the plan's own padded buffers get copies of them.
// The flat kernels read x, m directly (m may be NULL: no mask).
static n4_status flat_plan(n4_handle* h, uint64_t n, int nbins, Plan* P, const float* x, const uint8_t* m,
                           bool hist = false) {
    n4_params p;
    n4_params_default(&p);
    p.number_of_histogram_bins = (uint32_t)nbins;
    p.number_of_fitting_levels = 1;
    const int nslab = (int)((n + kChunkVox - 1) / kChunkVox);
    VolIn v{x, m, N4_MASK_U8, {nslab, 1, kChunkVox}, 1, nullptr, nullptr, nullptr, nullptr};
    n4_status s = build_plan(h, std::vector<VolIn>{v}, &p, P, false);
    if (s != N4_OK) return s;
    VolDesc& D = P->desc[0];
    D.nvox_w = (int64_t)n;
    if (!hist) {
        D.logU = const_cast<float*>(x);
        D.wmask = const_cast<uint8_t*>(m);
    }
    memcpy(h->stage + P->off_desc, &D, sizeof D);
    s = stage_upload(h, P->base, 0, P->upload_bytes, nullptr);
    if (s != N4_OK) return s;
    if (hist) {
        const size_t nw = (size_t)nslab * kChunkVox;
        D2D(D.logU, x, sizeof(float) * n, nullptr);
        D2D(D.wmask, m, n, nullptr);
        DSET(D.wmask + n, 0, nw + kQuadPad - n, nullptr);
    }
    return N4_OK;
}

// Host-side reduction of the flat kernels' slab partials.
static n4_status read_partials(n4_handle* h, const Plan& P, std::vector<double>* sp) {
    sp->assign((size_t)4 * P.n_wslab, 0.0);
    D2H(sp->data(), P.desc[0].spart, sizeof(double) * sp->size(), nullptr);
    SYNC(nullptr);
    (void)h;
    return N4_OK;
}

// Min/max + histogram (+ Wiener E table) of a flat array: flat min/max
// partials -> level 0 setup (mask count, cleared accumulators) -> hist ->
// wiener.
static n4_status flat_hist(n4_handle* h, const Plan& P, const float* d_x, RunParams rp, int32_t* d_bin_index) {
    VolDesc* dd = dev_desc(h, P);
    Slab* sl = dev_wslab(h, P);
    LAUNCH(launch_flat(dd, sl, P.n_wslab, kFlatMinMax, d_x, nullptr, nullptr, rp, nullptr));
    LAUNCH(launch_level(dd, 1, 0, rp, nullptr));
    LAUNCH(launch_hist(dd, sl, P.n_wslab, -1, rp, d_bin_index, P.lds_hist, nullptr));
    return N4_OK;
}

n4_status n4_histogram(n4_handle* h, const float* d_x, const uint8_t* d_mask, uint64_t n, uint32_t nbins,
                       double* out_mms, double* d_hist, int32_t* d_bin_index) {
    g_err.clear();
    if (!h || !d_x || !d_mask || n == 0) return fail(N4_ERR_INVALID_ARG, "NULL argument or empty input");
    if (nbins < 2 || nbins > (uint32_t)kMaxBins) return fail(N4_ERR_INVALID_ARG, "nbins out of range");
    Plan P;
    n4_status s = flat_plan(h, n, (int)nbins, &P, d_x, d_mask, true);
    if (s != N4_OK) return s;
    s = flat_hist(h, P, d_x, P.rp, d_bin_index);
    if (s != N4_OK) return s;
    VolState st;
    D2H(&st, dev_state(h, P), sizeof st, nullptr);
    std::vector<uint64_t> hb(nbins);
    D2H(hb.data(), P.desc[0].hist, hb.size() * 8, nullptr);
    SYNC(nullptr);
    std::vector<double> H(nbins, 0.0);
    if (!st.degenerate)
        for (uint32_t b = 0; b < nbins; ++b) H[b] = (double)hb[b] * 0x1p-32;
    if (out_mms) {
        out_mms[0] = st.bin_min;
        out_mms[1] = st.bin_max;
        out_mms[2] = st.slope;
    }
    if (d_hist) H2D(d_hist, H.data(), sizeof(double) * nbins, nullptr);
    SYNC(nullptr);
    return N4_OK;
}

n4_status n4_sharpen_image(n4_handle* h, const float* d_before, const uint8_t* d_mask, uint64_t n, uint32_t nbins,
                           double wiener_noise, double fwhm, float* d_after, double* d_etable) {
    g_err.clear();
    if (!h || !d_before || !d_mask || !d_after || n == 0) return fail(N4_ERR_INVALID_ARG, "NULL argument or empty input");
    if (nbins < 2 || nbins > (uint32_t)kMaxBins) return fail(N4_ERR_INVALID_ARG, "nbins out of range");
    if (!(fwhm > 0.0) || !(wiener_noise >= 0.0)) return fail(N4_ERR_INVALID_ARG, "bad Wiener parameters");
    Plan P;
    n4_status s = flat_plan(h, n, (int)nbins, &P, d_before, d_mask, true);
    if (s != N4_OK) return s;
    RunParams rp = P.rp;
    rp.noise = wiener_noise;
    rp.fwhm = fwhm;
    s = flat_hist(h, P, d_before, rp, nullptr);
    if (s != N4_OK) return s;
    VolDesc* dd = dev_desc(h, P);
    Slab* sl = dev_wslab(h, P);
    LAUNCH(launch_wiener(dd, 1, rp, P.lds_wiener, nullptr));
    LAUNCH(launch_flat(dd, sl, P.n_wslab, kFlatRemap, d_before, nullptr, d_after, rp, nullptr));
    if (d_etable) D2D(d_etable, P.desc[0].etab, sizeof(double) * nbins, nullptr);
    SYNC(nullptr);
    return N4_OK;
}

n4_status n4_fit_bspline3d(n4_handle* h, const float* d_field, const uint8_t* d_mask, const uint32_t ncpt[3],
                           const uint32_t dims[3], double eps, float* d_lattice) {
    g_err.clear();
    if (!h || !d_field || !d_mask || !ncpt || !dims || !d_lattice) return fail(N4_ERR_INVALID_ARG, "NULL argument");
    n4_params p;
    n4_params_default(&p);
    p.number_of_fitting_levels = 1;
    p.fit_eps = eps;
    for (int d = 0; d < 3; ++d) p.ncpt[d] = ncpt[d];
    n4_status s = check_params(&p);
    if (s != N4_OK) return s;
    VolIn v{d_field, d_mask, N4_MASK_U8, {(int)dims[0], (int)dims[1], (int)dims[2]}, 1, nullptr, nullptr, nullptr, nullptr};
    for (int d = 0; d < 3; ++d)
        if (v.F[d] < 1) return fail(N4_ERR_INVALID_ARG, "empty dimension");
    Plan P;
    s = build_plan(h, std::vector<VolIn>{v}, &p, &P, false);
    if (s != N4_OK) return s;
    s = check_lds(P, true, false);
    if (s != N4_OK) return s;
    VolDesc& D = P.desc[0];
    s = stage_upload(h, P.base, 0, P.upload_bytes, nullptr);
    if (s != N4_OK) return s;
    // the fit kernel takes 16-byte loads of aligned voxel quads: it reads
    // the plan's own padded copies of the field and the mask
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    D2D(D.logU, d_field, sizeof(float) * nvox, nullptr);
    D2D(D.wmask, d_mask, nvox, nullptr);
    VolDesc* dd = dev_desc(h, P);
    Slab* sl = dev_wslab(h, P);
    LAUNCH(launch_level(dd, 1, 0, P.rp, nullptr));
    LAUNCH(launch_fit(dd, sl, P.n_wslab, 0, 0, kFitDen, P.rp, P.lds_fit_max[0], P.mt[0], P.nt[0], nullptr));
    LAUNCH(launch_lattice(dd, 1, P.cx_max[0], 0, kFitDen, P.rp, nullptr));
    LAUNCH(launch_fit(dd, sl, P.n_wslab, 0, 0, kFitField, P.rp, P.lds_fit_max[0], P.mt[0], P.nt[0], nullptr));
    LAUNCH(launch_lattice(dd, 1, P.cx_max[0], 0, kFitField, P.rp, nullptr));
    D2D(d_lattice, D.lresf, sizeof(float) * D.lev[0].L, nullptr);
    SYNC(nullptr);
    return N4_OK;
}

n4_status n4_evaluate_bspline3d(n4_handle* h, const float* d_lattice, const uint32_t ncpt[3], const uint32_t dims[3],
                                float* d_out) {
    g_err.clear();
    if (!h || !d_lattice || !ncpt || !dims || !d_out) return fail(N4_ERR_INVALID_ARG, "NULL argument");
    n4_params p;
    n4_params_default(&p);
    p.number_of_fitting_levels = 1;
    for (int d = 0; d < 3; ++d) p.ncpt[d] = ncpt[d];
    n4_status s = check_params(&p);
    if (s != N4_OK) return s;
    VolIn v{nullptr, nullptr, N4_MASK_NONE, {(int)dims[0], (int)dims[1], (int)dims[2]}, 1, nullptr, nullptr, d_out, nullptr};
    for (int d = 0; d < 3; ++d)
        if (v.F[d] < 1) return fail(N4_ERR_INVALID_ARG, "empty dimension");
    Plan P;
    s = build_plan(h, std::vector<VolIn>{v}, &p, &P, true);
    if (s != N4_OK) return s;
    s = check_lds(P, false, true);
    if (s != N4_OK) return s;
    VolDesc& D = P.desc[0];
    D.latf = const_cast<float*>(d_lattice);
    memcpy(h->stage + P.off_desc, &D, sizeof D);
    s = stage_upload(h, P.base, 0, P.upload_bytes, nullptr);
    if (s != N4_OK) return s;
    LAUNCH(launch_recon(dev_desc(h, P), dev_fslab(h, P), P.n_fslab, 0, P.rp, P.lds_recon, nullptr));
    SYNC(nullptr);
    return N4_OK;
}

n4_status n4_convergence_measurement(n4_handle* h, const float* d_f1, const float* d_f2, const uint8_t* d_mask,
                                     uint64_t n, double* out) {
    g_err.clear();
    if (!h || !d_f1 || !d_f2 || !d_mask || !out || n == 0) return fail(N4_ERR_INVALID_ARG, "NULL argument or empty input");
    Plan P;
    n4_status s = flat_plan(h, n, 2, &P, d_f1, d_mask);
    if (s != N4_OK) return s;
    LAUNCH(launch_flat(dev_desc(h, P), dev_wslab(h, P), P.n_wslab, kFlatCV, d_f1, d_f2, nullptr, P.rp, nullptr));
    std::vector<double> sp;
    s = read_partials(h, P, &sp);
    if (s != N4_OK) return s;
    double cnt = 0.0, s1 = 0.0, s2 = 0.0;
    for (int x = 0; x < P.n_wslab; ++x) {
        cnt += sp[4 * x];
        s1 += sp[4 * x + 1];
        s2 += sp[4 * x + 2];
    }
    double cvv = 0.0;
    if (cnt > 0.0) {
        const double mu = s1 / cnt;
        double var = s2 / cnt - mu * mu;
        if (var < 0.0) var = 0.0;
        cvv = sqrt(var) / (1.0 + mu);
    }
    *out = cvv;
    return N4_OK;
}

n4_status n4_mask_to_u8(n4_handle* h, const float* d_mask, uint64_t n, uint8_t* d_out) {
    g_err.clear();
    if (!h || !d_mask || !d_out || n == 0) return fail(N4_ERR_INVALID_ARG, "NULL argument or empty input");
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipSetDevice(h->device));
#endif
    LAUNCH(launch_mask_u8(d_mask, d_out, n, nullptr));
    SYNC(nullptr);
    return N4_OK;
}

n4_status n4_upsample_lattice3d(n4_handle* h, const float* d_lattice, const uint32_t ncpt[3], float* d_out) {
    g_err.clear();
    if (!h || !d_lattice || !ncpt || !d_out) return fail(N4_ERR_INVALID_ARG, "NULL argument");
    n4_params p;
    n4_params_default(&p);
    p.number_of_fitting_levels = 2;
    for (int d = 0; d < 3; ++d) p.ncpt[d] = ncpt[d];
    n4_status s = check_params(&p);
    if (s != N4_OK) return s;
    // A 1x1x1 dummy grid: only the lattice buffers of the plan are used.
    VolIn v{nullptr, nullptr, N4_MASK_NONE, {1, 1, 1}, 1, nullptr, nullptr, nullptr, nullptr};
    Plan P;
    s = build_plan(h, std::vector<VolIn>{v}, &p, &P, false);
    if (s != N4_OK) return s;
    s = stage_upload(h, P.base, 0, P.upload_bytes, nullptr);
    if (s != N4_OK) return s;
    const VolDesc& D = P.desc[0];
    const size_t L0 = (size_t)D.lev[0].L, L1 = (size_t)D.lev[1].L;
    std::vector<float> lf(L0);
    std::vector<double> ld(L0);
    D2H(lf.data(), d_lattice, L0 * 4, nullptr);
    SYNC(nullptr);
    for (size_t i = 0; i < L0; ++i) ld[i] = lf[i];
    H2D(D.lat[0], ld.data(), L0 * 8, nullptr);
    LAUNCH(launch_level(dev_desc(h, P), 1, 1, P.rp, nullptr));
    D2D(d_out, D.latf, L1 * 4, nullptr);
    SYNC(nullptr);
    return N4_OK;
}

static n4_status reduce_flat(n4_handle* h, uint64_t n, const float* d_in, double* sum, double* mn, double* mx) {
    if (!h || !d_in || n == 0) return fail(N4_ERR_INVALID_ARG, "NULL argument or empty input");
    Plan P;
    n4_status s = flat_plan(h, n, 2, &P, d_in, nullptr);
    if (s != N4_OK) return s;
    LAUNCH(launch_flat(dev_desc(h, P), dev_wslab(h, P), P.n_wslab, kFlatSum, d_in, nullptr, nullptr, P.rp, nullptr));
    std::vector<double> sp;
    s = read_partials(h, P, &sp);
    if (s != N4_OK) return s;
    double a = 0.0, lo = INFINITY, hi = -INFINITY;
    for (int x = 0; x < P.n_wslab; ++x) {
        a += sp[4 * x];
        lo = std::min(lo, sp[4 * x + 2]);
        hi = std::max(hi, sp[4 * x + 3]);
    }
    *sum = a;
    *mn = lo;
    *mx = hi;
    return N4_OK;
}

n4_status n4_reduce_sum(n4_handle* h, uint64_t n, const float* d_in, double* out) {
    g_err.clear();
    double a, lo, hi;
    n4_status s = reduce_flat(h, n, d_in, &a, &lo, &hi);
    if (s == N4_OK && out) *out = a;
    return s;
}

n4_status n4_reduce_max(n4_handle* h, uint64_t n, const float* d_in, float* out) {
    g_err.clear();
    double a, lo, hi;
    n4_status s = reduce_flat(h, n, d_in, &a, &lo, &hi);
    if (s == N4_OK && out) *out = (float)hi;
    return s;
}

n4_status n4_reduce_min(n4_handle* h, uint64_t n, const float* d_in, float* out) {
    g_err.clear();
    double a, lo, hi;
    n4_status s = reduce_flat(h, n, d_in, &a, &lo, &hi);
    if (s == N4_OK && out) *out = (float)lo;
    return s;
}

n4_status n4_profile_enable(n4_handle* h, uint32_t mask) {
    if (!h) return fail(N4_ERR_INVALID_ARG, "handle is NULL");
    h->prof_mask = mask;
    for (int k = 0; k < N4_K_COUNT; ++k) {
        h->prof_ms[k] = 0.0;
        h->prof_n[k] = 0;
    }
    return N4_OK;
}

n4_status n4_profile_read(n4_handle* h, uint32_t kernel, double* total_ms, uint64_t* launches) {
    if (!h || kernel >= N4_K_COUNT) return fail(N4_ERR_INVALID_ARG, "bad handle or kernel id");
    if (total_ms) *total_ms = h->prof_ms[kernel];
    if (launches) *launches = h->prof_n[kernel];
    return N4_OK;
}

// ------------------------------------------------------- memory plumbing --
n4_status n4_malloc(n4_handle* h, void** d_ptr, size_t bytes) {
    if (!h || !d_ptr) return fail(N4_ERR_INVALID_ARG, "NULL argument");
    *d_ptr = nullptr;
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipSetDevice(h->device));
#endif
    if (bytes == 0) return N4_OK;
    if (dev_alloc(d_ptr, bytes)) return fail(N4_ERR_NO_MEMORY, "allocation of %zu bytes failed", bytes);
    return N4_OK;
}

n4_status n4_free(n4_handle* h, void* d_ptr) {
    (void)h;
    dev_free(d_ptr);
    return N4_OK;
}

n4_status n4_memcpy_htod(n4_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h || (bytes && (!dst || !src))) return fail(N4_ERR_INVALID_ARG, "NULL argument");
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
#else
    memcpy(dst, src, bytes);
#endif
    return N4_OK;
}

n4_status n4_memcpy_dtoh(n4_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h || (bytes && (!dst || !src))) return fail(N4_ERR_INVALID_ARG, "NULL argument");
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
#else
    memcpy(dst, src, bytes);
#endif
    return N4_OK;
}

n4_status n4_memcpy_dtod(n4_handle* h, void* dst, const void* src, size_t bytes) {
    if (!h || (bytes && (!dst || !src))) return fail(N4_ERR_INVALID_ARG, "NULL argument");
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
#else
    memcpy(dst, src, bytes);
#endif
    return N4_OK;
}

n4_status n4_memset(n4_handle* h, void* dst, int value, size_t bytes) {
    if (!h || (bytes && !dst)) return fail(N4_ERR_INVALID_ARG, "NULL argument");
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipMemset(dst, value, bytes));
#else
    memset(dst, value, bytes);
#endif
    return N4_OK;
}

n4_status n4_synchronize(n4_handle* h) {
    if (!h) return fail(N4_ERR_INVALID_ARG, "handle is NULL");
#ifndef N4_EMULATE
    N4_CHECK_HIP(hipSetDevice(h->device));
    N4_CHECK_HIP(hipDeviceSynchronize());
#endif
    return N4_OK;
}

}  // extern "C"
