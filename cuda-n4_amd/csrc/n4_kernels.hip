// n4_kernels.hip -- gfx950 kernels of the N4 fitting iteration.
//
// Kernel bodies are written as barrier-separated phases (BLOCK_FOR ...
// BLOCK_SYNC): on the GPU a phase is the straight-line code of one thread;
// in the host emulation build (N4_EMULATE, used only by CPU unit tests of
// the kernel logic) a phase is a loop over the block's threads.  All state
// that crosses a barrier lives in LDS, so both readings are the same program.
//
// Reference stages (cudaN4.cu / N4.py) and where they live here:
//   lowthreshold + log_kernel (+ build-defined shrink) -> body_prep
//   set_background + reduce_min/max + CV reductions    -> per-slab partials,
//                                                         reduced in body_hist
//   histogramGPU_kernel (cudaN4.cu:153-220)            -> body_hist
//   host FFTW Wiener block (cudaN4.cu:733-820)         -> body_wiener (LDS FP64 FFT)
//   histogramMapping + subtract + accumulate_WC2phic_and_WC2
//                                                      -> body_fit
//   divide_kernel + sum_inplace_kernel                 -> body_lattice
//   evaluate_bspline_kernel_3D + subtract_and_exp + subtract
//                                                      -> body_eval
//   upsample_lattice_kernel_3D (cudaN4.cu:257-339)     -> body_level
//   evaluate at full res + exp_and_divide + exp        -> body_recon
#include "n4_kernels.h"

#ifdef N4_EMULATE
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>
#define __host__
#define __device__
#else
#include <hip/hip_runtime.h>
#endif

#include <math.h>

#include <type_traits>

#if defined(__HIP_DEVICE_COMPILE__)
#define N4_DEVICE_PASS 1
#endif

#ifdef N4_DEVICE_PASS
#define BLOCK_FOR(t) { const int t = (int)threadIdx.x; (void)t;
#define BLOCK_END }
#define BLOCK_SYNC() __syncthreads()
#define UNROLL _Pragma("unroll")
#else
#define BLOCK_FOR(t) for (int t = 0; t < kBlock; ++t) {
#define BLOCK_END }
#define BLOCK_SYNC() do { } while (0)
#define UNROLL
#endif

#define HD __host__ __device__ inline

// LDS view: every pointer into a workgroup's LDS carries this address space.
#ifdef N4_DEVICE_PASS
#define N4_LDS __attribute__((address_space(3)))
#else
#define N4_LDS
#endif

namespace n4 {

// ---------------------------------------------------------------- helpers --

HD D2 cmul(D2 a, D2 b) { return D2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

HD void lds_add_u64(N4_LDS unsigned long long* p, unsigned long long v) {
#ifdef N4_DEVICE_PASS
    atomicAdd(p, v);
#else
    *p += v;
#endif
}

HD void global_add_u64(uint64_t* p, unsigned long long v) {
#ifdef N4_DEVICE_PASS
    atomicAdd(reinterpret_cast<unsigned long long*>(p), v);
#else
    __atomic_fetch_add(p, (uint64_t)v, __ATOMIC_SEQ_CST);
#endif
}

HD void sincospi_d(double x, double* s, double* c) {
#ifdef N4_DEVICE_PASS
    sincospi(x, s, c);
#else
    *s = sin(M_PI * x);
    *c = cos(M_PI * x);
#endif
}

HD size_t align16(size_t v) { return (v + 15) & ~size_t(15); }

// Global-memory view of a descriptor pointer.  A pointer loaded from a
// descriptor is generic to the compiler, which then emits flat loads and
// stores; those also count against the LDS wait counter, so every LDS read
// after them waits for the outstanding memory traffic.  Accesses through
// gp(p) are global_load / global_store.
#ifdef N4_DEVICE_PASS
#define N4_GLOBAL __attribute__((address_space(1)))
#else
#define N4_GLOBAL
#endif
template <class T>
HD N4_GLOBAL T* gp(T* p) {
    return (N4_GLOBAL T*)p;
}

// Four consecutive floats from any 4-byte-aligned address, in one vector
// load (global loads need only dword alignment).
HD F4 load_run4(const float* p) {
#ifdef N4_DEVICE_PASS
    typedef float v4u __attribute__((ext_vector_type(4), aligned(4)));
    const v4u x = *reinterpret_cast<const N4_GLOBAL v4u*>(gp(p));
    return F4{{x[0], x[1], x[2], x[3]}};
#else
    F4 r;
    memcpy(r.v, p, sizeof r.v);
    return r;
#endif
}

// One 16-byte global load of a voxel quad (p 16-byte aligned).
HD F4 load_quad(const float* p) {
#ifdef N4_DEVICE_PASS
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f x = *reinterpret_cast<const N4_GLOBAL v4f*>(gp(p));
    return F4{{x[0], x[1], x[2], x[3]}};
#else
    return *reinterpret_cast<const F4*>(p);
#endif
}

// One 16-byte global store of a voxel quad (p 16-byte aligned).
HD void store_quad(float* p, const F4& x) {
#ifdef N4_DEVICE_PASS
    typedef float v4f __attribute__((ext_vector_type(4)));
    v4f y = {x.v[0], x.v[1], x.v[2], x.v[3]};
    *reinterpret_cast<N4_GLOBAL v4f*>(gp(p)) = y;
#else
    *reinterpret_cast<F4*>(p) = x;
#endif
}

// Data shared between the workgroups of a persistent team (partials,
// histogram and numerator copies, lattice, increments, barrier counter) is
// written only with agent-scope atomics / stores and read only with
// agent-scope loads: coherent across the XCDs' L2 caches without cache
// write-backs or invalidations.
template <class T>
HD T ld_agent(const T* p) {
#ifdef N4_DEVICE_PASS
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
    T v;
    __atomic_load(p, &v, __ATOMIC_SEQ_CST);
    return v;
#endif
}
template <class T>
HD void st_agent(T* p, T v) {
#ifdef N4_DEVICE_PASS
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
    __atomic_store(p, &v, __ATOMIC_SEQ_CST);
#endif
}

// Index of a step-kernel tile within its volume (its slot in spart).
HD int tile_id(const VolDesc& V, const Slab& sl) {
    return sl.x * V.wtiles_x + (V.wtiles_x > 1 ? sl.y0 / V.wtile_rows : 0);
}
HD int round_up(int v, int m) { return (v + m - 1) / m * m; }

// floor(v / d) with the host-computed magic m = floor(2^40/d) + 1
// (exact while v * d < 2^40).
HD uint32_t fdiv(uint32_t v, uint64_t m) { return (uint32_t)(((uint64_t)v * m) >> kDivShift); }

// Bump allocator over the block's dynamic LDS; on the host, with base ==
// nullptr, it only measures (host-side sizing uses the same code as the
// device carve).  Every carve-out is typed as LDS (N4_LDS), so the compiler
// rejects passing one where a generic pointer is expected: LDS is only ever
// read and written with ds_* instructions, never with flat ones.  On the
// device an LDS pointer is never tested against null: the dynamic LDS
// starts at offset 0, which compares equal to the LDS null pointer.
struct Carve {
    N4_LDS unsigned char* base;
    size_t off;
    template <class T>
    HD N4_LDS T* take(size_t n) {
        off = align16(off);
#ifdef N4_DEVICE_PASS
        N4_LDS T* p = reinterpret_cast<N4_LDS T*>(base + off);
#else
        N4_LDS T* p = base ? reinterpret_cast<N4_LDS T*>(base + off) : nullptr;
#endif
        off += n * sizeof(T);
        return p;
    }
    // the region at offset o of the same block (measuring: nullptr)
    HD N4_LDS unsigned char* at(size_t o) const {
#ifdef N4_DEVICE_PASS
        return base + o;
#else
        return base ? base + o : nullptr;
#endif
    }
};

// Progress word in pinned host memory, read by the host's early-exit poll
// (no events in the iteration stream).
HD void publish_progress(const VolDesc& V, int level, int passes, int done) {
    if (!V.progress) return;
    const uint64_t w = progress_word(level, passes, done);
#ifdef N4_DEVICE_PASS
    __hip_atomic_store(V.progress, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
#else
    *V.progress = w;
#endif
}

HD double red_apply(int op, double a, double b) {
    return op == 0 ? a + b : (op == 1 ? fmin(a, b) : fmax(a, b));
}

// Reduction of four kBlock-long LDS arrays (op: 0 sum, 1 min, 2 max);
// results in red[0], red[kBlock], red[2*kBlock], red[3*kBlock].  Each
// wave64 reduces its lanes with an xor butterfly (every lane ends with the
// same value: the ops are commutative), then the wave results combine in
// wave order.  The order of operations is fixed, so every block reducing
// the same data gets the same bits; the host emulation performs the same
// operations in the same order.
HD void tree_reduce4(N4_LDS double* red, int o0, int o1, int o2, int o3) {
    BLOCK_SYNC();
    const int ops[4] = {o0, o1, o2, o3};
    constexpr int kWaves = kBlock / 64;
#ifdef N4_DEVICE_PASS
    const int t = (int)threadIdx.x, lane = t & 63, w = t >> 6;
    double v[4];
    UNROLL for (int k = 0; k < 4; ++k) v[k] = red[k * kBlock + t];
    UNROLL for (int sh = 32; sh > 0; sh >>= 1)
        UNROLL for (int k = 0; k < 4; ++k) v[k] = red_apply(ops[k], v[k], __shfl_xor(v[k], sh));
    __syncthreads();
    if (lane == 0)
        UNROLL for (int k = 0; k < 4; ++k) red[k * kBlock + w] = v[k];
    __syncthreads();
    if (t == 0) {
        UNROLL for (int k = 0; k < 4; ++k) {
            double r = red[k * kBlock];
            for (int i = 1; i < kWaves; ++i) r = red_apply(ops[k], r, red[k * kBlock + i]);
            red[k * kBlock] = r;
        }
    }
    __syncthreads();
#else
    for (int k = 0; k < 4; ++k) {
        double wres[kWaves];
        for (int w = 0; w < kWaves; ++w) {
            double v[64], nv[64];
            for (int l = 0; l < 64; ++l) v[l] = red[k * kBlock + 64 * w + l];
            for (int sh = 32; sh > 0; sh >>= 1) {
                for (int l = 0; l < 64; ++l) nv[l] = red_apply(ops[k], v[l], v[l ^ sh]);
                for (int l = 0; l < 64; ++l) v[l] = nv[l];
            }
            wres[w] = v[0];
        }
        double r = wres[0];
        for (int i = 1; i < kWaves; ++i) r = red_apply(ops[k], r, wres[i]);
        red[k * kBlock] = r;
    }
#endif
}

#ifdef N4_DEVICE_PASS
// tree_reduce4's results (red[k * kBlock], its order, so the same bits)
// from each thread's four values in registers: wave butterflies, the waves'
// results through LDS, combined in wave order by thread 0; no LDS round
// trip for the threads' values.  The wave results overwrite slots of red:
// no thread may still read red when the block enters.
__device__ void reduce4_regs(double a, double b, double c, double d, N4_LDS double* red, int o0, int o1, int o2,
                             int o3) {
    const int t = (int)threadIdx.x;
    const int ops[4] = {o0, o1, o2, o3};
    double v[4] = {a, b, c, d};
    UNROLL for (int sh = 32; sh > 0; sh >>= 1)
        UNROLL for (int k = 0; k < 4; ++k) v[k] = red_apply(ops[k], v[k], __shfl_xor(v[k], sh));
    if ((t & 63) == 0)
        UNROLL for (int k = 0; k < 4; ++k) red[k * kBlock + (t >> 6)] = v[k];
    __syncthreads();
    if (t == 0) {
        UNROLL for (int k = 0; k < 4; ++k) {
            double r = red[k * kBlock];
            for (int i = 1; i < kBlock / 64; ++i) r = red_apply(ops[k], r, red[k * kBlock + i]);
            red[k * kBlock] = r;
        }
    }
    __syncthreads();
}
#endif

// Reduces n partial quadruples part[i] = (s0, s1, min, max).  Thread t
// folds partials t, t + kBlock, ... in that order; their loads are issued
// eight at a time (two 16-byte loads per partial) before they are folded.
struct alignas(16) D2x {
    double a, b;
};
HD void reduce_partials(const double* part, int nslab, N4_LDS double* red) {
    BLOCK_FOR(t)
    double r0 = 0.0, r1 = 0.0, r2 = INFINITY, r3 = -INFINITY;
    for (int x0 = t; x0 < nslab; x0 += 8 * kBlock) {
        D2x lo[8], hi[8];
        UNROLL for (int u = 0; u < 8; ++u) {
            const int xx = x0 + u * kBlock;
            const D2x* sp = reinterpret_cast<const D2x*>(part + (size_t)(xx < nslab ? xx : 0) * 4);
            lo[u] = sp[0];
            hi[u] = sp[1];
        }
        UNROLL for (int u = 0; u < 8; ++u) {
            if (x0 + u * kBlock >= nslab) break;
            r0 += lo[u].a;
            r1 += lo[u].b;
            r2 = fmin(r2, hi[u].a);
            r3 = fmax(r3, hi[u].b);
        }
    }
#ifdef N4_DEVICE_PASS
    reduce4_regs(r0, r1, r2, r3, red, 0, 0, 1, 2);  // red is not read before
#else
    red[t] = r0;
    red[kBlock + t] = r1;
    red[2 * kBlock + t] = r2;
    red[3 * kBlock + t] = r3;
#endif
    BLOCK_END
#ifndef N4_DEVICE_PASS
    tree_reduce4(red, 0, 0, 1, 2);
#endif
}

// Coefficient of variation of exp(old - new) over n voxels from
// s1 = sum(u), s2 = sum(u^2), u = exp(old - new) - 1 (N4.py:121-126, ddof 0).
HD double cv_from_sums(double s1, double s2, double n) {
    if (!(n > 0.0)) return 0.0;
    const double mu = s1 / n;
    double var = s2 / n - mu * mu;
    if (var < 0.0) var = 0.0;
    return sqrt(var) / (1.0 + mu);
}

// Rows of the evaluation kernel's chunk: its staging is only Cz floats per
// row, so a chunk covers a whole default tile.
HD int rows_per_eval_chunk(int Y, int Z) {
    int r = 4 * kChunkVox / (Z > 0 ? Z : 1);
    if (r < 1) r = 1;
    return r < Y ? r : Y;
}

HD int rows_per_chunk(int Y, int Z) {
    int r = kChunkVox / (Z > 0 ? Z : 1);
    if (r < 1) r = 1;
    return r < Y ? r : Y;
}

// ------------------------------------------------------------------ prep --
// Threshold + log + shrink (N4.py:280-287; cudaN4.cu:131-147, 639-652) for
// one working slab: each thread owns working voxels (y, z) and reads their
// s*s*s full-resolution blocks (lanes on consecutive z read consecutive
// s-float runs).  Writes imlog, logU (= imlog), wmask and the slab partial
// (masked full voxels, foreground working voxels, min, max of imlog).
// S > 0: shrink factor known at compile time (unrolled); S == 0: any.
// MK: the mask's kind (n4_mask_kind F32 0, U8 1, none 2) or 3, an F32
// confidence mask (N4_FLAG_WEIGHTED_MASK): foreground where the weight is
// > 0 and finite, and the block's mean foreground weight goes to wconf.
template <int MK>
using PrepMaskT = typename std::conditional<MK == 3, float, bool>::type;

template <int MK>
HD PrepMaskT<MK> prep_mask(const VolDesc& V, size_t fv) {
    if constexpr (MK == 3) return gp(reinterpret_cast<const float*>(V.mask))[fv];
    if (MK == 0) return gp(reinterpret_cast<const float*>(V.mask))[fv] != 0.0f;
    if (MK == 1) return gp(reinterpret_cast<const uint8_t*>(V.mask))[fv] != 0;
    return true;
}

// Foreground test of one full voxel (N4.py:280-287: mask and x > low_value).
template <int MK>
HD bool prep_fg(PrepMaskT<MK> m, float xv, const RunParams& rp) {
    bool in;
    if constexpr (MK == 3)
        in = m > 0.0f && m < INFINITY;
    else
        in = m;
    return in && xv > rp.low_f && isfinite(xv);
}

template <int MK>
HD void prep_fold(PrepMaskT<MK> m, float xv, const RunParams& rp, int& nm, float& acc, float& wsum) {
    if (prep_fg<MK>(m, xv, rp)) {
        ++nm;
        acc += logf(xv);
        if constexpr (MK == 3) wsum += m;
    }
}

template <int S, int MK>
HD void prep_voxel(const VolDesc& V, const RunParams& rp, int x, int J, int K, int& nm, int& nb, float& acc,
                   float& wsum) {
    const int F0 = V.F[0], F1 = V.F[1], F2 = V.F[2];
    // Unrolled interior blocks for shrink 3 and 4 (shrink 1 and 2 fold their
    // interior blocks in body_prep_t and come here for the border only).
    // Confidence masks take the generic loop with a run-time block size
    // (their float mask values would raise every shrink class's registers).
    constexpr bool kUnrolled = S > 2 && MK != 3;
    const int ss = (S > 0 && MK != 3) ? S : V.s;
    const int i0 = x * ss, j0 = J * ss, k0 = K * ss;
    const bool interior = (i0 + ss <= F0) && (j0 + ss <= F1) && (k0 + ss <= F2);
    if (kUnrolled && interior) {
        constexpr int SS = S > 0 ? S : 1;
        UNROLL for (int a = 0; a < SS; ++a) {
            float xv[SS * SS];
            PrepMaskT<MK> mv[SS * SS];
            UNROLL for (int b = 0; b < SS; ++b) {
                UNROLL for (int c = 0; c < SS; ++c) {
                    const size_t fv = ((size_t)(i0 + a) * F1 + (j0 + b)) * F2 + (k0 + c);
                    xv[b * SS + c] = gp(V.im)[fv];
                    mv[b * SS + c] = prep_mask<MK>(V, fv);
                }
            }
            // (written out rather than through prep_fold: this form keeps
            // the shrink-3 and -4 kernels at 49 and 64 VGPRs)
            UNROLL for (int q = 0; q < SS * SS; ++q) {
                ++nb;
                if (mv[q] && xv[q] > rp.low_f && isfinite(xv[q])) {
                    ++nm;
                    acc += logf(xv[q]);
                }
            }
        }
        return;
    }
    const int i1 = (i0 + ss < F0) ? i0 + ss : F0;
    const int j1 = (j0 + ss < F1) ? j0 + ss : F1;
    const int k1 = (k0 + ss < F2) ? k0 + ss : F2;
    for (int i = i0; i < i1; ++i)
        for (int j = j0; j < j1; ++j) {
            const size_t row = ((size_t)i * F1 + j) * F2;
            for (int k = k0; k < k1; ++k) {
                ++nb;
                prep_fold<MK>(prep_mask<MK>(V, row + k), gp(V.im)[row + k], rp, nm, acc, wsum);
            }
        }
}

// Writes one working voxel (block of nb full voxels, nm of them masked with
// log sum acc and, for confidence masks, weight sum wsum) and folds it into
// the thread's partials.
template <int MK>
HD void prep_commit(const VolDesc& V, int x, int J, int K, int nm, int nb, float acc, float wsum, double& nfull,
                    double& nwork, double& mn, double& mx) {
    const size_t w = ((size_t)x * V.W[1] + J) * V.W[2] + K;
    const bool fg = nm > 0 && 2 * nm >= nb;
    const float lg = nm > 0 ? acc / (float)nm : 0.0f;
    gp(V.imlog)[w] = lg;
    gp(V.logU)[w] = lg;
    gp(V.wmask)[w] = fg ? 1 : 0;
    if constexpr (MK == 3) gp(V.wconf)[w] = fg ? wsum / (float)nm : 0.0f;
    nfull += nm;
    if (fg) {
        nwork += 1.0;
        mn = fmin(mn, (double)lg);
        mx = fmax(mx, (double)lg);
    }
}

template <int S, int MK>
HD void body_prep_t(const VolDesc& V, const Slab sl, const RunParams& rp, N4_LDS unsigned char* smem) {
    Carve cv{smem, 0};
    N4_LDS double* red = cv.take<double>(4 * kBlock);
    const int W2 = V.W[2];
    const int x = sl.x;
    // Shrink 1 and 2: UN working voxels per thread have all their s^3 image
    // and mask loads in flight before any is folded (same folding order as
    // prep_voxel).  Larger shrinks keep one working voxel per iteration
    // (their blocks already hold 27+ loads; the shrink cases share one
    // kernel, whose register count is that of its largest case).
    constexpr bool kBatched = S >= 1 && S <= 2 && MK != 3;
    BLOCK_FOR(t)
    double nfull = 0.0, nwork = 0.0;
    double mn = INFINITY, mx = -INFINITY;
    if constexpr (kBatched) {
        constexpr int UN = S == 1 ? 8 : 2;
        constexpr int SB = kBatched ? S : 1;
        constexpr int NV = SB * SB * SB;
        const int F0 = V.F[0], F1 = V.F[1], F2 = V.F[2];
        const int n = sl.ny * W2;
        for (int v0 = t; v0 < n; v0 += UN * kBlock) {
            float xv[UN][NV];
            PrepMaskT<MK> mv[UN][NV];
            bool in[UN];
            UNROLL for (int u = 0; u < UN; ++u) {
                const int v = v0 + u * kBlock;
                in[u] = false;
                if (v < n) {
                    const int jr = (int)fdiv((uint32_t)v, V.mW2);
                    const int i0 = x * SB, j0 = (sl.y0 + jr) * SB, k0 = (v - jr * W2) * SB;
                    in[u] = (i0 + SB <= F0) && (j0 + SB <= F1) && (k0 + SB <= F2);
                    if (in[u]) {
                        UNROLL for (int a = 0; a < SB; ++a) UNROLL for (int b = 0; b < SB; ++b) UNROLL for (int c = 0; c < SB; ++c) {
                            const size_t fv = ((size_t)(i0 + a) * F1 + (j0 + b)) * F2 + (k0 + c);
                            xv[u][(a * SB + b) * SB + c] = gp(V.im)[fv];
                            mv[u][(a * SB + b) * SB + c] = prep_mask<MK>(V, fv);
                        }
                    }
                }
            }
            UNROLL for (int u = 0; u < UN; ++u) {
                const int v = v0 + u * kBlock;
                if (v >= n) break;
                const int jr = (int)fdiv((uint32_t)v, V.mW2);
                const int J = sl.y0 + jr, K = v - jr * W2;
                int nm = 0, nb = 0;
                float acc = 0.0f, wsum = 0.0f;
                if (in[u]) {
                    UNROLL for (int q = 0; q < NV; ++q) {
                        ++nb;
                        prep_fold<MK>(mv[u][q], xv[u][q], rp, nm, acc, wsum);
                    }
                } else {
                    prep_voxel<S, MK>(V, rp, x, J, K, nm, nb, acc, wsum);
                }
                prep_commit<MK>(V, x, J, K, nm, nb, acc, wsum, nfull, nwork, mn, mx);
            }
        }
    } else {
        for (int v = t; v < sl.ny * W2; v += kBlock) {
            const int jr = (int)fdiv((uint32_t)v, V.mW2);
            const int J = sl.y0 + jr, K = v - jr * W2;
            int nm = 0, nb = 0;
            float acc = 0.0f, wsum = 0.0f;
            prep_voxel<S, MK>(V, rp, x, J, K, nm, nb, acc, wsum);
            prep_commit<MK>(V, x, J, K, nm, nb, acc, wsum, nfull, nwork, mn, mx);
        }
    }
    red[t] = nfull;
    red[kBlock + t] = nwork;
    red[2 * kBlock + t] = mn;
    red[3 * kBlock + t] = mx;
    BLOCK_END
    tree_reduce4(red, 0, 0, 1, 2);
    BLOCK_FOR(t)
    if (t == 0) {
        double* sp = V.ppart + ((size_t)x * V.prep_tiles_x + sl.y0 / V.prep_rows) * 4;
        sp[0] = red[0];
        sp[1] = red[kBlock];
        sp[2] = red[2 * kBlock];
        sp[3] = red[3 * kBlock];
    }
    BLOCK_END
}

template <int S>
HD void body_prep_s(const VolDesc& V, const Slab sl, const RunParams& rp, N4_LDS unsigned char* smem) {
    if (V.wconf)
        body_prep_t<S, 3>(V, sl, rp, smem);
    else if (V.mask_kind == 0)
        body_prep_t<S, 0>(V, sl, rp, smem);
    else if (V.mask_kind == 1)
        body_prep_t<S, 1>(V, sl, rp, smem);
    else
        body_prep_t<S, 2>(V, sl, rp, smem);
}

HD void body_prep(const VolDesc* vols, const Slab* slabs, RunParams rp, int bid, N4_LDS unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    switch (V.s) {
        case 1: body_prep_s<1>(V, sl, rp, smem); break;
        case 2: body_prep_s<2>(V, sl, rp, smem); break;
        case 3: body_prep_s<3>(V, sl, rp, smem); break;
        case 4: body_prep_s<4>(V, sl, rp, smem); break;
        default: body_prep_s<0>(V, sl, rp, smem); break;
    }
}

// ------------------------------------------------------------ level setup --
// Dyadic refinement weights (N4.py:134-136): even child {1/8, 6/8, 1/8},
// odd child {0, 1/2, 1/2} over parents i-1, i, i+1.
HD double refine_w(int odd, int a) {
    return odd ? (a == 0 ? 0.0 : 0.5) : (a == 1 ? 0.75 : 0.125);
}

// Run records of a correction (both nullable in VolDesc): the convergence
// value of iteration it (1-based) of level lvl, at index sum of the earlier
// levels' maxima + it - 1 of cvtrace, and the FP32 lattice of each level at
// its end, at offset sum of the earlier levels' sizes of lvlat.
HD void record_cv(const VolDesc& V, int lvl, int it, double cvv) {
    if (!V.cvtrace) return;
    int o = it - 1;
    for (int l = 0; l < lvl; ++l) o += V.max_iter[l];
    if (o < V.trace_cap) V.cvtrace[o] = cvv;
}

HD float* level_lattice_record(const VolDesc& V, int lvl) {
    if (!V.lvlat) return nullptr;
    size_t o = 0;
    for (int l = 0; l < lvl; ++l) o += (size_t)V.lev[l].L;
    return V.lvlat + o;
}

// Lattice of `level` (> 0) as the dyadic refinement of level-1's FP64
// lattice (N4.py:132-165, cudaN4.cu:257-339) with per-axis clip bounds;
// also its FP32 copy.  Runs inside a thread phase (thread t of kBlock).
// TEAM (persistent path, run by the team's leader): the lattices are read
// and written with agent-scope accesses (the team's owners update them).
template <bool TEAM>
HD void refine_lattice(const VolDesc& V, int level, int t) {
    const LevelDesc& LD = V.lev[level];
    const LevelDesc& PD = V.lev[level - 1];
    const double* src = V.lat[(level - 1) & 1];
    double* dst = V.lat[level & 1];
    if (float* rec = level_lattice_record(V, level - 1))
        for (int l = t; l < PD.L; l += kBlock) rec[l] = (float)(TEAM ? ld_agent(src + l) : src[l]);
    for (int l = t; l < LD.L; l += kBlock) {
        int c[3];
        c[0] = l / (LD.C[1] * LD.C[2]);
        c[1] = (l / LD.C[2]) % LD.C[1];
        c[2] = l % LD.C[2];
        int par[3], odd[3];
        for (int d = 0; d < 3; ++d) {
            const int ic = c[d] - 1;  // child index in [-1, 2n+1]
            const int i = (ic >= 0) ? ic / 2 : -((-ic + 1) / 2);
            par[d] = i;
            odd[d] = ic - 2 * i;
        }
        double acc = 0.0;
        for (int a = 0; a < 3; ++a) {
            int X = par[0] - 1 + a;
            const int nx = PD.C[0] - 3;
            X = X < -1 ? -1 : (X > nx + 1 ? nx + 1 : X);
            for (int b = 0; b < 3; ++b) {
                int Y = par[1] - 1 + b;
                const int ny = PD.C[1] - 3;
                Y = Y < -1 ? -1 : (Y > ny + 1 ? ny + 1 : Y);
                for (int e = 0; e < 3; ++e) {
                    int Z = par[2] - 1 + e;
                    const int nz = PD.C[2] - 3;
                    Z = Z < -1 ? -1 : (Z > nz + 1 ? nz + 1 : Z);
                    const double wgt = refine_w(odd[0], a) * refine_w(odd[1], b) * refine_w(odd[2], e);
                    const double* sp = src + ((size_t)(X + 1) * PD.C[1] + (Y + 1)) * PD.C[2] + (Z + 1);
                    acc += wgt * (TEAM ? ld_agent(sp) : *sp);
                }
            }
        }
        if (TEAM) {
            st_agent(dst + l, acc);
        } else {
            dst[l] = acc;
            V.latf[l] = (float)acc;
        }
    }
}

// One block per volume.
//  level == 0: start of a correction (or of a stage call): reduces the prep
//    partials to the mask counts, seeds the slab partials with the logU
//    range, clears lattice and accumulators and puts the volume in state
//    TRANS(0).
//  level > 0: refines the level-1 lattice (lattice refinement stage call).
HD void body_level(const VolDesc* vols, int level, RunParams rp, int bid, N4_LDS unsigned char* smem) {
    const VolDesc& V = vols[bid];
    VolState* st = V.st;
    if (level > 0) {
        BLOCK_FOR(t)
        refine_lattice<false>(V, level, t);
        BLOCK_END
        return;
    }
    Carve cv{smem, 0};
    N4_LDS double* red = cv.take<double>(4 * kBlock);
    reduce_partials(V.ppart, V.W[0] * V.prep_tiles_x, red);
    const LevelDesc& LD = V.lev[0];
    int Lmax = 0;
    for (int l = 0; l < V.nlevels; ++l) Lmax = V.lev[l].L > Lmax ? V.lev[l].L : Lmax;
    BLOCK_FOR(t)
    for (int xx = t; xx < V.W[0] * V.wtiles_x; xx += kBlock) {
        double* sp = V.spart + (size_t)xx * 4;
        sp[0] = 0.0;
        sp[1] = 0.0;
        sp[2] = xx == 0 ? red[2 * kBlock] : INFINITY;
        sp[3] = xx == 0 ? red[3 * kBlock] : -INFINITY;
    }
    for (int l = t; l < LD.L; l += kBlock) {
        V.lat[0][l] = 0.0;
        V.latf[l] = 0.0f;
        V.lresf[l] = 0.0f;
    }
    for (int l = t; l < Lmax * V.num_copies; l += kBlock) V.numfx[l] = 0;
    for (int i = t; i < rp.nbins * V.hist_copies; i += kBlock) V.hist[i] = 0;
    if (t < 2 && V.team_ctr) V.team_ctr[t] = 0;
    if (t < kPhases && V.tphase) V.tphase[t] = 0;
    if (t == 0) {
        st->n_mask_full = (uint64_t)red[0];
        st->n_mask_work = (uint64_t)red[kBlock];
        st->mode = kModeTrans;
        st->lvl = 0;
        st->it = 0;
        st->s_kind = kStepSkip;
        st->s_lvl = 0;
        st->s_it = 0;
        st->passes = 0;
        st->done = 0;
        st->cv = 1e10;
        for (int l = 0; l < kMaxLevels; ++l) {
            st->iters[l] = 0;
            st->lastcv[l] = 1e10;
        }
        publish_progress(V, 0, 0, 0);
    }
    BLOCK_END
}

// Where the histogram copies come from and the table goes:
//   kWienerLocal  the copies were completed by earlier launches: plain
//                 loads, cleared; table to V.etab
//   kWienerTeam   persistent path: agent-scope loads, left for the team's
//                 leader to clear; table to Ef (FP32, LDS: the remap's copy)
enum WienerMode : int { kWienerLocal = 0, kWienerTeam = 2 };

// ------------------------------------------------------------- histogram --
// Fractional histogram of one tile into kHistLanes lane-strided LDS copies
// of the bins (odd stride hs: the copies of one bin fall in different
// banks), N4.py:19-29: bin index c = (x - min) / slope in FP64 exactly as
// N4.py computes it, contributions as 2^-32 fixed point (integer adds:
// order-independent).  Aligned quads of voxels cover the tile: one 16-byte
// logU load and one 4-byte mask load per quad, four quads in flight per
// thread.  dbg_idx (nullable): per-voxel bin index, -1 where unmasked.
HD void hist_zero(N4_LDS unsigned long long* h, int hs) {
    BLOCK_FOR(t)
    for (int i = t; i < kHistLanes * hs; i += kBlock) h[i] = 0ull;
    BLOCK_END
}

HD void hist_tile(const VolDesc& V, const Slab sl, double mn, double slope, int nb, N4_LDS unsigned long long* h,
                  int hs, int32_t* dbg_idx) {
    const int YZ = sl.ny * V.W[2];
    const size_t base = ((size_t)sl.x * V.W[1] + sl.y0) * V.W[2];
    const size_t vend = std::min(base + (size_t)YZ, (size_t)V.nvox_w);
    const size_t q0 = base & ~(size_t)3;
    const int nq = (int)((vend - q0 + 3) >> 2);
    BLOCK_FOR(t)
    N4_LDS unsigned long long* hl = h + (size_t)(t & (kHistLanes - 1)) * hs;
    for (int i0 = 0; i0 < nq; i0 += 4 * kBlock) {
        F4 xs[4];
        uint32_t ms[4];
        UNROLL for (int u = 0; u < 4; ++u) {
            const int qi = i0 + u * kBlock + t;
            const size_t w = q0 + 4 * (size_t)qi;
            const bool ok = qi < nq;
            ms[u] = ok ? *reinterpret_cast<const N4_GLOBAL uint32_t*>(gp(V.wmask) + w) : 0u;
            xs[u] = ok ? load_quad(V.logU + w) : F4{{0.0f, 0.0f, 0.0f, 0.0f}};
        }
        UNROLL for (int u = 0; u < 4; ++u) {
            const size_t wq = q0 + 4 * (size_t)(i0 + u * kBlock + t);
            UNROLL for (int j = 0; j < 4; ++j) {
                const size_t w = wq + j;
                if (w < base || w >= vend) continue;
                if (!((ms[u] >> (8 * j)) & 0xffu)) {
                    if (dbg_idx) dbg_idx[w] = -1;
                    continue;
                }
                const double c = ((double)xs[u].v[j] - mn) / slope;
                const double fl = floor(c);
                const int idx = (fl >= 0.0 && fl < (double)nb) ? (int)fl : -1;
                if (dbg_idx) dbg_idx[w] = idx;
                if (idx < 0) continue;
                const double off = c - fl;
                if (off == 0.0) {
                    lds_add_u64(&hl[idx], 1ull << 32);
                } else if (idx < nb - 1) {
                    const unsigned long long q = (unsigned long long)rint(off * 4294967296.0);
                    lds_add_u64(&hl[idx], (1ull << 32) - q);
                    lds_add_u64(&hl[idx + 1], q);
                }
            }
        }
    }
    BLOCK_END
}

// Merges the LDS copies and adds them to one global copy of the bins.
HD void hist_flush(const N4_LDS unsigned long long* h, int hs, int nb, uint64_t* dst) {
    BLOCK_FOR(t)
    for (int i = t; i < nb; i += kBlock) {
        unsigned long long s = 0;
        for (int c = 0; c < kHistLanes; ++c) s += h[(size_t)c * hs + i];
        if (s) global_add_u64(&dst[i], s);
    }
    BLOCK_END
}


// Prologue (every block of the volume computes identical bits): reduce the
// eval partials of the previous step -> logU min/max (N4.py:14) and the
// convergence measurement of the previous iteration, then the step's kind
// from the volume's state:
//   TRANS(l)            -> TRANS(l)
//   ITER(l, it), it > 0 and CV <= threshold (N4.py:301-302), or it at the
//   level's maximum     -> the level ends after `it` iterations: TRANS(l+1),
//                          or FINAL after the last level
//   ITER(l, it) else    -> ITER(l, it)
//   finished            -> SKIP
// The block of slab 0 commits the level record and the step and publishes
// the volume's progress.  step < 0 (stage calls) forces ITER(0, 0).  Then,
// for ITER steps, the fractional (linear-split) histogram of logU over the
// mask, N4.py:19-29 (cudaN4.cu:153-220): bin index in FP64 exactly as N4.py
// computes it; contributions accumulate as 2^-32 fixed point in LDS-private
// u64 bins, then in the volume's global u64 bins (integer adds:
// order-independent, so bitwise deterministic).
HD void body_hist(const VolDesc* vols, const Slab* slabs, int step, RunParams rp, int32_t* dbg_idx, int bid,
                  N4_LDS unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    VolState* st = rp.states + sl.vol;
    const int nb = rp.nbins;
    // The partial reduction's scratch and the bins share the block's LDS
    // (the bins are cleared once the reduction's results are in registers),
    // so the kernel's LDS is the larger of the two and 8 workgroups fit a CU.
    N4_LDS double* red = Carve{smem, 0}.take<double>(4 * kBlock);
    // kHistLanes lane-strided copies of the bins (odd stride: the copies of
    // one bin fall in different LDS banks), merged before the global flush
    const int hs = nb | 1;
    N4_LDS unsigned long long* h = Carve{smem, 0}.take<unsigned long long>((size_t)kHistLanes * hs);
    const int mode = step < 0 ? kModeIter : st->mode;
    const int lvl = step < 0 ? 0 : st->lvl, it = step < 0 ? 0 : st->it;
    if (mode == kModeFinished) {
        if (sl.x == 0 && sl.y0 == 0) {
            BLOCK_FOR(t)
            if (t == 0) st->s_kind = kStepSkip;
            BLOCK_END
        }
        return;
    }
    const double n = (double)st->n_mask_work;  // loaded with the state, before the reduction's barriers
    reduce_partials(V.spart, V.W[0] * V.wtiles_x, red);
    const double mn = red[2 * kBlock], mx = red[3 * kBlock];
    double cvv = 0.0;
    int kind = kStepTrans, s_lvl = lvl;
    bool level_end = false;
    if (mode == kModeIter) {
        if (it > 0) cvv = cv_from_sums(red[0], red[kBlock], n);
        const bool stop = it > 0 && !rp.fixed_iters && !(cvv > rp.thr);
        level_end = stop || it >= rp.max_iter[lvl];
        if (!level_end) {
            kind = kStepIter;
        } else if (lvl + 1 < rp.nlevels) {
            kind = kStepTrans;
            s_lvl = lvl + 1;
        } else {
            kind = kStepFinal;
        }
    }
    const double slope = (mx - mn) / (double)(nb - 1);
    const bool degenerate = !(n > 0.0) || !(slope > 0.0);
    if (sl.x == 0 && sl.y0 == 0) {
        BLOCK_FOR(t)
        if (t == 0) {
            if (mode == kModeIter && it > 0 && step >= 0) record_cv(V, lvl, it, cvv);
            if (level_end) {
                st->iters[lvl] = it;
                if (it > 0) {
                    st->lastcv[lvl] = cvv;
                    st->cv = cvv;
                }
            }
            st->s_kind = kind;
            st->s_lvl = s_lvl;
            st->s_it = it;
            st->bin_min = mn;
            st->bin_max = mx;
            st->slope = slope;
            st->degenerate = degenerate ? 1 : 0;
            if (step >= 0) {
                st->passes += 1;
                publish_progress(V, 0, st->passes, kind == kStepFinal ? 1 : 0);
            }
        }
        BLOCK_END
    }
    if (kind == kStepIter && !degenerate) {
        BLOCK_SYNC();  // every thread has read the reduction's results (the bins overlay them)
        hist_zero(h, hs);
        BLOCK_SYNC();
        hist_tile(V, sl, mn, slope, nb, h, hs, dbg_idx);
        BLOCK_SYNC();
        // tiles spread their flush over the copies: fewer atomics per address
        hist_flush(h, hs, nb, V.hist + (size_t)(tile_id(V, sl) % V.hist_copies) * nb);
    }
}

// ---------------------------------------------------------------- Wiener --
// In-register DFT of R = 2, 4 or 8 points (natural order in and out):
// radix-2 decimation in time on the bit-reversed input.  R is a template
// parameter so every index is a compile-time constant (registers only).
template <int R>
HD void dft_small(D2 (&v)[R], double sg) {
    const double c = 0.70710678118654752440;  // sqrt(2)/2
    D2 u[R];
    if constexpr (R == 8) {
        u[0] = v[0]; u[1] = v[4]; u[2] = v[2]; u[3] = v[6];
        u[4] = v[1]; u[5] = v[5]; u[6] = v[3]; u[7] = v[7];
    } else if constexpr (R == 4) {
        u[0] = v[0]; u[1] = v[2]; u[2] = v[1]; u[3] = v[3];
    } else {
        u[0] = v[0]; u[1] = v[1];
    }
    UNROLL for (int i = 0; i < R; i += 2) {
        const D2 a = u[i], b = u[i + 1];
        u[i] = D2{a.x + b.x, a.y + b.y};
        u[i + 1] = D2{a.x - b.x, a.y - b.y};
    }
    if constexpr (R >= 4) {
        UNROLL for (int g = 0; g < R; g += 4) {
            const D2 a0 = u[g], a1 = u[g + 1], b0 = u[g + 2], b1 = u[g + 3];
            const D2 tb1{-sg * b1.y, sg * b1.x};  // b1 * exp(sg i pi/2)
            u[g] = D2{a0.x + b0.x, a0.y + b0.y};
            u[g + 2] = D2{a0.x - b0.x, a0.y - b0.y};
            u[g + 1] = D2{a1.x + tb1.x, a1.y + tb1.y};
            u[g + 3] = D2{a1.x - tb1.x, a1.y - tb1.y};
        }
    }
    if constexpr (R == 8) {
        // W8^j = exp(sg i pi j / 4)
        const D2 w[4] = {D2{1.0, 0.0}, D2{c, sg * c}, D2{0.0, sg}, D2{-c, sg * c}};
        UNROLL for (int j = 0; j < 4; ++j) {
            const D2 t = cmul(u[j + 4], w[j]);
            const D2 a = u[j];
            u[j] = D2{a.x + t.x, a.y + t.y};
            u[j + 4] = D2{a.x - t.x, a.y - t.y};
        }
    }
    UNROLL for (int i = 0; i < R; ++i) v[i] = u[i];
}

// One Stockham pass of radix R over P points (span Ns): butterfly j takes
// a[j + r*P/R], twiddles them by exp(-+2 pi i r k / (Ns R)) (k = j mod Ns)
// and writes b[(j - k) R + k + r Ns].  Thread t of the block takes
// butterflies t, t + kBlock, ...  P and the direction are compile-time
// constants: with every pass unrolled and its index arithmetic folded, a
// 512-point transform takes 1.3 us on MI355X against 2.3 us for the same
// code with a run-time size and direction.
template <int R, int P, bool INV>
HD void block_fft_pass(const N4_LDS D2* a, N4_LDS D2* b, const N4_LDS D2* tw, int Ns) {
    constexpr int nbf = P / R;
    const int tstep = P / (Ns * R);
    constexpr double sg = INV ? 1.0 : -1.0;
    BLOCK_FOR(t)
    for (int j = t; j < nbf; j += kBlock) {
        const int k = j & (Ns - 1);
        D2 v[R];
        UNROLL for (int r = 0; r < R; ++r) v[r] = a[j + r * nbf];
        if (Ns > 1) {
            UNROLL for (int r = 1; r < R; ++r) {
                D2 w = tw[r * k * tstep];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small<R>(v, sg);
        const int o = (j - k) * R + k;
        UNROLL for (int r = 0; r < R; ++r) b[o + r * Ns] = v[r];
    }
    BLOCK_END
}

// LDS FP64 Stockham FFT over P = 2^k points (tw[m] = exp(-2 pi i m / P), m <
// P, in LDS): radix-8 passes (radix 4 or 2 for the last), a block barrier
// between passes.  Returns the buffer holding the result (the other one is
// free).  (Radix 2 and 4 over all 256 threads, and radix 8 in one
// wavefront with wave-level ordering, measured equal or slower on MI355X.)
constexpr int kFftRadix = 8;
template <int P, bool INV>
HD N4_LDS D2* block_fft(N4_LDS D2* a, N4_LDS D2* b, const N4_LDS D2* tw) {
    UNROLL for (int Ns = 1; Ns < P;) {
        const int rem = P / Ns;
        const int R = rem >= kFftRadix ? kFftRadix : rem;
        if (R == 8)
            block_fft_pass<8, P, INV>(a, b, tw, Ns);
        else if (R == 4)
            block_fft_pass<4, P, INV>(a, b, tw, Ns);
        else
            block_fft_pass<2, P, INV>(a, b, tw, Ns);
        BLOCK_SYNC();
        N4_LDS D2* tmp = a;
        a = b;
        b = tmp;
        Ns *= R;
    }
    return a;
}

// Wiener deconvolution of the fractional histogram and the E[u|v] table,
// N4.py:31-101 (cudaN4.cu:733-820 did this on the host with FFTW).  Two real
// transforms share one complex FFT (V in the real part, F in the
// imaginary).  Reads and clears the volume's histogram.
template <int WM, int P>
HD void wiener_table_p(const VolDesc& V, double bmin, double slope, RunParams rp, N4_LDS unsigned char* smem,
                       N4_LDS float* Ef) {
    constexpr bool TEAM = WM == kWienerTeam;
    constexpr bool AGENT = WM != kWienerLocal;
    constexpr int half = P / 2;
    const int off = rp.offset, nb = rp.nbins;
    Carve cv{smem, 0};
    N4_LDS D2* z = cv.take<D2>(P);
    N4_LDS D2* y = cv.take<D2>(P);
    N4_LDS D2* ff = cv.take<D2>(P);
    N4_LDS D2* tw = cv.take<D2>(P);
    const double scaledFWHM = rp.fwhm / slope;
    const double expF = 4.0 * log(2.0) / (scaledFWHM * scaledFWHM);
    const double scale = 2.0 * sqrt(log(2.0) / M_PI) / scaledFWHM;
    BLOCK_FOR(t)
    for (int j = t; j < P; j += kBlock) {
        const int jj = (j <= half) ? j : P - j;
        const double Fj = (j == 0) ? scale : scale * exp((double)(-(int64_t)jj * jj) * expF);
        double Vj = 0.0;
        if (j >= off && j < off + nb) {
            // exact: the copies are fixed-point; all loads first, then the clears
            uint64_t* hp = V.hist + (j - off);
            const int nc = V.hist_copies;
            uint64_t c[kMaxCopies];
            UNROLL for (int k = 0; k < kMaxCopies; ++k) {
                if (AGENT)
                    c[k] = k < nc ? ld_agent(hp + (size_t)k * nb) : 0;
                else
                    c[k] = k < nc ? hp[(size_t)k * nb] : 0;
            }
            uint64_t s = 0;
            UNROLL for (int k = 0; k < kMaxCopies; ++k) s += c[k];
            if (WM == kWienerLocal)
                UNROLL for (int k = 0; k < kMaxCopies; ++k)
                    if (k < nc) hp[(size_t)k * nb] = 0;

            Vj = (double)s * 0x1p-32;
        }
        z[j] = D2{Vj, Fj};
        tw[j] = rp.twiddle[j];
    }
    BLOCK_END
    BLOCK_SYNC();
    // the chain of transforms and pointwise steps, block-wide
    N4_LDS D2* Z = block_fft<P, false>(z, y, tw);
    N4_LDS D2* U = (Z == z) ? y : z;
    // Split Vf and Ff, Gf = conj(Ff)/(conj(Ff)Ff + noise), Uf = Vf*Re(Gf)
    BLOCK_FOR(t)
    for (int k = t; k < P; k += kBlock) {
        const D2 a = Z[k];
        const D2 m = Z[(P - k) & (P - 1)];
        const D2 Vf{0.5 * (a.x + m.x), 0.5 * (a.y - m.y)};
        const D2 Ff{0.5 * (a.y + m.y), -0.5 * (a.x - m.x)};
        ff[k] = Ff;
        const double g = Ff.x / ((Ff.x * Ff.x + Ff.y * Ff.y) + rp.noise);
        U[k] = D2{Vf.x * g, Vf.y * g};
    }
    BLOCK_END
    BLOCK_SYNC();
    N4_LDS D2* Ur = block_fft<P, true>(U, Z, tw);
    N4_LDS D2* Wb = (Ur == U) ? Z : U;
    const double invP = 1.0 / (double)P;
    BLOCK_FOR(t)
    for (int j = t; j < P; j += kBlock) {
        double u = Ur[j].x * invP;
        if (u < 0.0) u = 0.0;
        Wb[j] = D2{(bmin + (double)(j - off) * slope) * u, u};
    }
    BLOCK_END
    BLOCK_SYNC();
    N4_LDS D2* Wf = block_fft<P, false>(Wb, Ur, tw);
    N4_LDS D2* Wo = (Wf == Wb) ? Ur : Wb;
    BLOCK_FOR(t)
    for (int k = t; k < P; k += kBlock) Wf[k] = cmul(Wf[k], ff[k]);
    BLOCK_END
    BLOCK_SYNC();
    N4_LDS D2* R = block_fft<P, true>(Wf, Wo, tw);
    BLOCK_FOR(t)
    for (int i = t; i < nb; i += kBlock) {
        const double nu = R[off + i].x * invP;
        const double de = R[off + i].y * invP;
        const double e = (de != 0.0) ? nu / de : 0.0;
        if (TEAM)
            Ef[i] = (float)e;
        else
            V.etab[i] = e;
    }
    BLOCK_END
}

// wiener_table_p for the run's padded size P = 2^(ceil(log2 nbins) + 1),
// 4 <= P <= 1024 (nbins in [2, 512]).
template <int WM>
HD void wiener_table(const VolDesc& V, double bmin, double slope, RunParams rp, N4_LDS unsigned char* smem,
                     N4_LDS float* Ef = nullptr) {
    switch (rp.P) {
        case 4: wiener_table_p<WM, 4>(V, bmin, slope, rp, smem, Ef); break;
        case 8: wiener_table_p<WM, 8>(V, bmin, slope, rp, smem, Ef); break;
        case 16: wiener_table_p<WM, 16>(V, bmin, slope, rp, smem, Ef); break;
        case 32: wiener_table_p<WM, 32>(V, bmin, slope, rp, smem, Ef); break;
        case 64: wiener_table_p<WM, 64>(V, bmin, slope, rp, smem, Ef); break;
        case 128: wiener_table_p<WM, 128>(V, bmin, slope, rp, smem, Ef); break;
        case 256: wiener_table_p<WM, 256>(V, bmin, slope, rp, smem, Ef); break;
        case 512: wiener_table_p<WM, 512>(V, bmin, slope, rp, smem, Ef); break;
        default: wiener_table_p<WM, 1024>(V, bmin, slope, rp, smem, Ef); break;
    }
}

// A volume's per-step work and the state for the next step (one block):
//   ITER   Wiener step (unless degenerate); next: ITER(l, it+1)
//   TRANS  refine the lattice into level l (l > 0); next: ITER(l, 0)
//   FINAL  FP32 copy of the final lattice (and the caller's output lattice);
//          next: finished
// WM: where the histogram copies come from (WienerMode).
template <int WM>
HD void volume_step(const VolDesc& V, int kind, int s_lvl, int s_it, bool degenerate, double bmin, double slope,
                    RunParams rp, N4_LDS unsigned char* smem) {
    VolState* st = V.st;
    if (kind == kStepIter) {
        if (!degenerate) wiener_table<WM>(V, bmin, slope, rp, smem);
        BLOCK_FOR(t)
        if (t == 0) {
            st->mode = kModeIter;
            st->it = s_it + 1;
        }
        BLOCK_END
    } else if (kind == kStepTrans) {
        const int l = s_lvl;
        BLOCK_FOR(t)
        if (l > 0) refine_lattice<false>(V, l, t);
        if (t == 0) {
            st->mode = kModeIter;
            st->lvl = l;
            st->it = 0;
            st->cv = 1e10;
        }
        BLOCK_END
    } else if (kind == kStepFinal) {
        const LevelDesc& FD = V.lev[V.nlevels - 1];
        const double* fin = V.lat[(V.nlevels - 1) & 1];
        float* rec = level_lattice_record(V, V.nlevels - 1);
        BLOCK_FOR(t)
        for (int l = t; l < FD.L; l += kBlock) {
            const float f = (float)fin[l];
            V.latf[l] = f;
            if (rec) rec[l] = f;
            if (V.out_lattice) V.out_lattice[l] = f;
        }
        if (t == 0) {
            st->mode = kModeFinished;
            st->done = 1;
        }
        BLOCK_END
    }
}

// One block per volume: the step as the histogram kernel committed it.
HD void body_wiener(const VolDesc* vols, RunParams rp, int bid, N4_LDS unsigned char* smem) {
    const VolDesc& V = vols[bid];
    const VolState* st = rp.states + (&V - rp.descs);
    // the state fields of the step, in one round of loads
    const int kind = st->s_kind, degenerate = st->degenerate, s_it = st->s_it, s_lvl = st->s_lvl;
    const double bmin = st->bin_min, slope = st->slope;
    volume_step<kWienerLocal>(V, kind, s_lvl, s_it, degenerate != 0, bmin, slope, rp, smem);
}

// ---------------------------------------------------------- lattice update --
// The fit of a tile of slab x adds its (y, z) contraction T2, weighted by
// the x factor Px[x, cx], to the numerator of the 4 control planes cx =
// b(x) .. b(x)+3 it touches, as fixed-point integer atomics into copy
// tile % num_copies of the numerator (integer sums do not depend on order,
// so the result is bitwise reproducible; the copies spread the atomics of
// many tiles over more addresses): num[cx,p] = sum_x Px[x,cx] T2_x[p]
// (N4.py:186-237).  The denominator of a level (mask contracted with the
// w^2 factors) is accumulated the same way in the level's first step.
// Zero terms (control points outside the tile's rows) are skipped.
HD void scatter_numerator(const VolDesc& V, const LevelDesc& LD, const AxisEntry& ex, int copy, bool q, int pr,
                          float t2) {
    const int CyCz = LD.C[1] * LD.C[2];
    uint64_t* acc = V.numfx + (size_t)copy * V.lmax;
    UNROLL for (int k = 0; k < 4; ++k) {
        const double c = (double)(q ? ex.q[k] : ex.p[k]) * (double)t2;
        const long long q = (long long)rint(c * kNumScale);
        if (q != 0) global_add_u64(acc + (size_t)(ex.b + k) * CyCz + pr, (unsigned long long)q);
    }
}

// Lattice update of control point l (N4.py:236, 317): num = the exact sum
// of the numerator copies (2^-30 fixed point), which are cleared for the
// next scatter; TRANS: den = num (the level's denominator); ITER:
// lres = num / (den + eps), lat += lres, lresf = lres (FP32 for the
// evaluation).  AGENT_IN: the copies were filled in the same launch (or
// team) and are read and cleared with agent-scope accesses; AGENT_OUT: den,
// the lattice and lresf are shared with other workgroups of the launch.
template <bool AGENT_IN, bool AGENT_OUT>
HD void update_cp(const VolDesc& V, int l, bool trans, int level, double eps) {
    uint64_t* p = V.numfx + l;
    const int nc = V.num_copies;
    // all loads first (independent), then the clears
    uint64_t c[kMaxCopies];
    UNROLL for (int k = 0; k < kMaxCopies; ++k) {
        if (AGENT_IN)
            c[k] = k < nc ? ld_agent(p + (size_t)k * V.lmax) : 0;
        else
            c[k] = k < nc ? p[(size_t)k * V.lmax] : 0;
    }
    uint64_t sum = 0;
    UNROLL for (int k = 0; k < kMaxCopies; ++k) sum += c[k];
    UNROLL for (int k = 0; k < kMaxCopies; ++k) {
        if (k >= nc) continue;
        if (AGENT_IN)
            st_agent(p + (size_t)k * V.lmax, (uint64_t)0);
        else
            p[(size_t)k * V.lmax] = 0;
    }
    const double num = (double)(long long)sum / kNumScale;
    double* lat = V.lat[level & 1] + l;
    if (AGENT_OUT) {
        if (trans) {
            st_agent(V.den + l, num);
        } else {
            const double lres = num / (ld_agent(V.den + l) + eps);
            st_agent(lat, ld_agent(lat) + lres);
            st_agent(V.lresf + l, (float)lres);
        }
    } else {
        if (trans) {
            V.den[l] = num;
        } else {
            const double lres = num / (V.den[l] + eps);
            *lat += lres;
            V.lresf[l] = (float)lres;
        }
    }
}

// Lattice update between the fit and the evaluation of a step,
// one thread per control point l of the volume's current level (block
// `chunk` of the volume covers l = chunk*kBlock ..).
HD void body_update(const VolDesc* vols, RunParams rp, int chunk, int vol) {
    const VolDesc& V = vols[vol];
    const VolState* st = rp.states + (&V - rp.descs);
    const int kind = st->s_kind;
    if (kind != kStepIter && kind != kStepTrans) return;
    const int level = st->s_lvl;
    const int L = V.lev[level].L;
    BLOCK_FOR(t)
    const int l = chunk * kBlock + t;
    if (l < L) update_cp<false, false>(V, l, kind == kStepTrans, level, rp.eps);
    BLOCK_END
}

// ------------------------------------------------------------------- fit --
// Remap (N4.py:103-115) + residual (N4.py:313) + the (y, z) part of the
// separable B-spline fit (N4.py:186-237 as mode-n contractions, exact
// because the normaliser sum(w^2) factorises per axis):
//   T1[y,cz] = sum_z g[y,z] Pz[z,cz]      (LDS, per row chunk)
//   T2[cy,cz] += sum_y Py[y,cy] T1[y,cz]  (LDS, accumulated over the slab)
//   part[x] = T2                          (global, per slab)
// Modes: NUM (g = residual, P = w^3/S), DEN (g = mask, Q = w^2; once per
// level), FIELD (g = the logU buffer itself; stage API).
struct FitLds {
    N4_LDS float* E;
    N4_LDS float* zwc;  // [Cz][zmax]  factor of control point cz at z = lo(cz) + i
    N4_LDS int* zs;
    N4_LDS int* ys;
    N4_LDS float* g;
    N4_LDS float* T1;
    N4_LDS float* T2;
    N4_LDS float* ywc;  // [Cy][ymax]  factor of control point cy at row y = lo(cy) + i
    N4_LDS AxisEntry* ex;  // the slab's x factors (copied in the set-up round)
};

// Longest support (in grid points) of one cubic B-spline basis function over
// n points with C control points: 4 spans of at most ceil(n / (C - 3)) + 1.
HD int support_len(int n, int C) {
    const int nint = C - 3 > 0 ? C - 3 : 1;
    const int m = 4 * ((n + nint - 1) / nint + 1);
    return m < n ? m : n;
}

// Resolves kFitStep from the volume's current step; false: nothing to fit.
HD bool fit_step_mode(const VolState* st, int* mode, int* level, int* it) {
    if (*mode != kFitStep) return true;
    if (st->s_kind == kStepIter) {
        *mode = kFitNum;
        *level = st->s_lvl;
        *it = st->s_it;
        return true;
    }
    if (st->s_kind == kStepTrans) {
        *mode = kFitDenScatter;
        *level = st->s_lvl;
        *it = 0;
        return true;
    }
    return false;
}

// The remap of a step in FP32: the E table (rounded from the FP64 Wiener
// output), the histogram origin and the reciprocal bin width.  FP32 here
// moves the residual by about one FP32 ulp of logU (~3e-7 at |logU| ~ 5),
// the rounding logU itself carries; the histogram bin index, which must be
// exact, stays FP64 in the histogram kernel.
struct Remap {
    const N4_LDS float* E;
    float bmin, islope;
    int nb;
};

HD Remap make_remap(const N4_LDS float* E, double bmin, double slope, int nb) {
    return Remap{E, (float)bmin, (float)(1.0 / slope), nb};
}

// Contracted quantity g of one working voxel: the residual logU - E[u|v]
// (N4.py:103-115 remap with the cudaN4.cu:242 clamp, then N4.py:313) for
// NUM, the mask for DEN, logU itself for FIELD; 0 outside the mask.  With
// confidence weights (V.wconf), the fit kernels scale g of a foreground
// voxel by its weight (conf_scale), which scales both its numerator and its
// denominator contribution.
HD float fit_input(int mode, uint8_t m, float xv, bool degenerate, const Remap& R) {
    if (!m) return 0.0f;
    if (mode == kFitDen) return 1.0f;
    if (mode == kFitField) return xv;
    if (degenerate) return 0.0f;
    float c = (xv - R.bmin) * R.islope;
    c = fminf(fmaxf(c, 0.0f), (float)R.nb);
    const int idx = (int)c;  // c >= 0: truncation is floor
    const float fr = c - (float)idx;
    const float sh = (idx < R.nb - 1) ? R.E[idx] + (R.E[idx + 1] - R.E[idx]) * fr : R.E[R.nb - 1];
    return xv - sh;
}

HD float conf_scale(const VolDesc& V, uint8_t m, size_t w, float g) {
    return (V.wconf && m) ? g * gp(V.wconf)[w] : g;
}

HD size_t carve_fit(N4_LDS unsigned char* base, int nbins, int Y, int Z, int R, int Cy, int Cz, FitLds* o) {
    Carve c{base, 0};
    FitLds L;
    L.E = c.take<float>(nbins);
    L.zwc = c.take<float>((size_t)Cz * support_len(Z, Cz));
    L.zs = c.take<int>((size_t)2 * Cz);
    L.ys = c.take<int>((size_t)2 * Cy);
    L.g = c.take<float>((size_t)R * Z);
    L.T1 = c.take<float>((size_t)R * Cz);
    L.T2 = c.take<float>((size_t)Cy * Cz);
    L.ywc = c.take<float>((size_t)Cy * support_len(Y, Cy));
    L.ex = c.take<AxisEntry>(1);
    if (o) *o = L;
    return align16(c.off);
}

// Thread t's share of a fit's factor tables: per control point, its
// factors (w^2 for the denominator, else w^3/sum w^2) over its support as
// one contiguous run (the contraction loops then read two independent LDS
// streams), and the support ranges.
HD void fit_tables_t(const LevelDesc& LD, const FitLds& S, int Y, int Z, bool den, int t) {
    const int Cy = LD.C[1], Cz = LD.C[2];
    const int zmax = support_len(Z, Cz), ymax = support_len(Y, Cy);
    const AxisEntry* axz = LD.ax[2];
    const AxisEntry* axy = LD.ax[1];
    for (int e = t; e < Cz * zmax; e += kBlock) {
        const int c = e / zmax, i = e - c * zmax;
        const int z = LD.sup[2][2 * c] + i;
        float w = 0.0f;
        if (z < LD.sup[2][2 * c + 1]) {
            const AxisEntry& a = axz[z];
            const int k = c - a.b;
            w = den ? a.q[k] : a.p[k];
        }
        S.zwc[e] = w;
    }
    for (int e = t; e < Cy * ymax; e += kBlock) {
        const int c = e / ymax, i = e - c * ymax;
        const int y = LD.sup[1][2 * c] + i;
        float w = 0.0f;
        if (y < LD.sup[1][2 * c + 1]) {
            const AxisEntry& a = axy[y];
            const int k = c - a.b;
            w = den ? a.q[k] : a.p[k];
        }
        S.ywc[e] = w;
    }
    for (int i = t; i < 2 * Cz; i += kBlock) S.zs[i] = LD.sup[2][i];
    for (int i = t; i < 2 * Cy; i += kBlock) S.ys[i] = LD.sup[1][i];
}

// The (y, z) contraction of one tile into S.T2 (zeroed by the caller):
// per chunk of R rows, g = fit_input of each voxel (aligned quads, 16-byte
// logU loads), T1[r,cz] = sum_z g Pz over the control point's contiguous
// support run, T2[cy,cz] += sum_y Py T1.  Tables (zwc, ywc, zs, ys) and, for
// NUM, the E table must be in LDS.
HD void fit_chunks(const VolDesc& V, const LevelDesc& LD, const Slab sl, const FitLds& S, int gmode, bool den,
                   bool degenerate, const Remap& rm) {
    const int Y = V.W[1], Z = V.W[2];
    const int Cy = LD.C[1], Cz = LD.C[2];
    const int R = rows_per_chunk(Y, Z);
    const int zmax = support_len(Z, Cz), ymax = support_len(Y, Cy);
    const int x = sl.x;
    const int yend = sl.y0 + sl.ny;
    for (int y0 = sl.y0; y0 < yend; y0 += R) {
        const int ny = (yend - y0 < R) ? yend - y0 : R;
        const size_t base = ((size_t)x * Y + y0) * Z;
        BLOCK_FOR(t)
        // aligned voxel quads covering the chunk (16-byte logU loads)
        const size_t vend = base + (size_t)ny * Z, q0 = base & ~(size_t)3;
        const int nq = (int)((vend - q0 + 3) >> 2);
        for (int i0 = 0; i0 < nq; i0 += 4 * kBlock) {
            F4 xs[4];
            uint32_t ms[4];
            UNROLL for (int u = 0; u < 4; ++u) {
                const int qi = i0 + u * kBlock + t;
                const size_t w = q0 + 4 * (size_t)qi;
                const bool ok = qi < nq;
                ms[u] = ok ? *reinterpret_cast<const N4_GLOBAL uint32_t*>(gp(V.wmask) + w) : 0u;
                xs[u] = (ok && !den) ? load_quad(V.logU + w) : F4{{0.0f, 0.0f, 0.0f, 0.0f}};
            }
            UNROLL for (int u = 0; u < 4; ++u) {
                const size_t wq = q0 + 4 * (size_t)(i0 + u * kBlock + t);
                UNROLL for (int j = 0; j < 4; ++j) {
                    const size_t w = wq + j;
                    if (w < base || w >= vend) continue;
                    const uint8_t m = (uint8_t)((ms[u] >> (8 * j)) & 0xffu);
                    S.g[w - base] = conf_scale(V, m, w, fit_input(gmode, m, xs[u].v[j], degenerate, rm));
                }
            }
        }
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        for (int pr = t; pr < ny * Cz; pr += kBlock) {
            const int r = (int)fdiv((uint32_t)pr, LD.mCz), c = pr - r * Cz;
            const int lo = S.zs[2 * c], n = S.zs[2 * c + 1] - lo;
            const N4_LDS float* gr = S.g + (size_t)r * Z + lo;
            const N4_LDS float* wc = S.zwc + (size_t)c * zmax;
            float acc = 0.0f;
            for (int i = 0; i < n; ++i) acc += gr[i] * wc[i];
            S.T1[pr] = acc;
        }
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        for (int pr = t; pr < Cy * Cz; pr += kBlock) {
            const int cy = (int)fdiv((uint32_t)pr, LD.mCz), c = pr - cy * Cz;
            const int ylo = S.ys[2 * cy];
            int lo = ylo, hi = S.ys[2 * cy + 1];
            lo = lo > y0 ? lo : y0;
            hi = hi < y0 + ny ? hi : y0 + ny;
            const N4_LDS float* wc = S.ywc + (size_t)cy * ymax;  // wc[i]: row ylo + i
            float acc = S.T2[pr];
            for (int yy = lo; yy < hi; ++yy) acc += wc[yy - ylo] * S.T1[(yy - y0) * Cz + c];
            S.T2[pr] = acc;
        }
        BLOCK_END
        BLOCK_SYNC();
    }
}

// Adds a tile's T2, weighted by the slab's x factors, to the numerator
// (scatter: fixed-point atomics into copy tile % num_copies) or writes it
// as the slab's partial (stage calls).
HD void fit_scatter(const VolDesc& V, const LevelDesc& LD, const Slab sl, const FitLds& S, bool scatter, bool den) {
    const int CyCz = LD.C[1] * LD.C[2];
    BLOCK_FOR(t)
    const AxisEntry ex = *S.ex;
    for (int pr = t; pr < CyCz; pr += kBlock) {
        if (scatter)
            scatter_numerator(V, LD, ex, tile_id(V, sl) % V.num_copies, den, pr, S.T2[pr]);
        else
            V.part[(size_t)sl.x * V.part_stride + pr] = S.T2[pr];
    }
    BLOCK_END
}

HD void body_fit(const VolDesc* vols, const Slab* slabs, int level, int it, int mode, RunParams rp, int bid,
                 N4_LDS unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    const VolState* st = rp.states + sl.vol;
    if (!fit_step_mode(st, &mode, &level, &it)) return;
    const LevelDesc& LD = V.lev[level];
    const int Y = V.W[1], Z = V.W[2];
    const int Cy = LD.C[1], Cz = LD.C[2];
    const int nb = rp.nbins;
    FitLds S;
    carve_fit(smem, nb, Y, Z, rows_per_chunk(Y, Z), Cy, Cz, &S);
    const bool num = (mode == kFitNum);
    const bool den = (mode == kFitDen || mode == kFitDenScatter);
    const bool scatter = (mode == kFitNum || mode == kFitDenScatter);
    const int gmode = den ? (int)kFitDen : mode;
    const bool degenerate = num && st->degenerate;
    const double bmin = st->bin_min, slope = st->slope;
    const Remap rm = make_remap(S.E, bmin, slope, nb);
    const int x = sl.x;
    BLOCK_FOR(t)
    if (num)
        for (int i = t; i < nb; i += kBlock) S.E[i] = (float)V.etab[i];
    fit_tables_t(LD, S, Y, Z, den, t);
    for (int i = t; i < Cy * Cz; i += kBlock) S.T2[i] = 0.0f;
    // the slab's x factors for the final scatter, in this round of loads
    if (t < (int)(sizeof(AxisEntry) / 4))
        reinterpret_cast<N4_LDS int32_t*>(S.ex)[t] = reinterpret_cast<const int32_t*>(&LD.ax[0][x])[t];
    BLOCK_END
    BLOCK_SYNC();
    fit_chunks(V, LD, sl, S, gmode, den, degenerate, rm);
    fit_scatter(V, LD, sl, S, scatter, den);
}

// --------------------------------------------------------------- lattice --
// Mode-x contraction of the slab partials and the lattice update; one block
// per (control-point plane cx, volume):
//   num[cx,p] = sum_x Px[x,cx] part[x][p]   (FP64, fixed summation order)
//   NUM  : lres = num/(den+eps); lat += lres  (N4.py:236, 317)
//   DEN  : den = num
//   FIELD: lres = num/(den+eps)
// Each column p is summed by G = kBlock/CyCz threads over interleaved slabs,
// then combined in a fixed order.
HD void body_lattice(const VolDesc* vols, int level, int mode, RunParams rp, int cx, int vol,
                     N4_LDS unsigned char* smem) {
    const VolDesc& V = vols[vol];
    const LevelDesc& LD = V.lev[level];
    if (cx >= LD.C[0]) return;
    if (mode == kFitNum && V.st->done) return;
    Carve cv{smem, 0};
    N4_LDS double* red = cv.take<double>(kBlock);
    const int CyCz = LD.C[1] * LD.C[2];
    const int lo = LD.sup[0][2 * cx], hi = LD.sup[0][2 * cx + 1];
    const AxisEntry* axx = LD.ax[0];
    const int cols = (CyCz <= kBlock) ? CyCz : kBlock;
    const int G = kBlock / cols;
    double* lat = V.lat[level & 1];
    for (int p0 = 0; p0 < CyCz; p0 += cols) {
        BLOCK_FOR(t)
        double acc = 0.0;
        const int g = t / cols, p = p0 + (t - (t / cols) * cols);
        if (g < G && p < CyCz) {
            for (int xx = lo + g; xx < hi; xx += G) {
                const AxisEntry& e = axx[xx];
                const float wx = (mode == kFitDen) ? e.q[cx - e.b] : e.p[cx - e.b];
                acc += (double)wx * (double)V.part[(size_t)xx * V.part_stride + p];
            }
        }
        red[t] = acc;
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        const int p = p0 + t;
        if (t < cols && p < CyCz) {
            double numv = 0.0;
            for (int g = 0; g < G; ++g) numv += red[g * cols + t];
            const size_t l = (size_t)cx * CyCz + p;
            if (mode == kFitDen) {
                V.den[l] = numv;
            } else {
                const double lres = numv / (V.den[l] + rp.eps);
                if (mode == kFitNum) {
                    const double nl = lat[l] + lres;
                    lat[l] = nl;
                    V.latf[l] = (float)nl;
                }
                V.lresf[l] = (float)lres;
            }
        }
        BLOCK_END
        BLOCK_SYNC();
    }
}

// ------------------------------------------------------------------ eval --
// Field increment on the working grid, convergence sums (N4.py:121-126) and
// logU update (N4.py:318, 322), fused.  The field is linear in the lattice,
// so new field = old field + d with d the evaluation of the increment, and
//   Ar[cy,cz] = sum_a wx[a] lres[bx+a,cy,cz]
//   Br[y,cz]  = sum_b wy[b] Ar[by+b,cz]
//   d(y,z)    = sum_c wz[c] Br[y,bz+c]
//   logU -= d  (= imlog - new field);  u = exp(old-new) - 1 = expm1(-d)
// Per slab: spart[x] = (sum u, sum u^2 in FP64, min, max of logU).
struct EvalLds {
    N4_LDS float* Ar;
    N4_LDS float* Br;
    N4_LDS float* zw;
    N4_LDS int* zb;
    N4_LDS float* yw;  // [tile rows][4] y factors of the tile's rows
    N4_LDS int* yb;
    N4_LDS double* red;
};

HD size_t carve_eval(N4_LDS unsigned char* base, int Z, int R, int Yt, int Cy, int Cz, EvalLds* o) {
    Carve c{base, 0};
    EvalLds L;
    L.Ar = c.take<float>((size_t)Cy * Cz);
    L.Br = c.take<float>((size_t)R * Cz);
    L.zw = c.take<float>((size_t)4 * Z);
    L.zb = c.take<int>(Z);
    L.yw = c.take<float>((size_t)4 * Yt);
    L.yb = c.take<int>(Yt);
    L.red = c.take<double>((size_t)4 * kBlock);
    if (o) *o = L;
    return align16(c.off);
}

// Field increment of one tile from its x-contracted increment planes
// S.Ar (and the tile-relative y tables S.yw, S.yb): per chunk of rows
// Br[r,cz] = sum_b wy Ar[by+b,cz], then per voxel d = sum_c wz Br (z tables
// S.zw, S.zb), logU -= d, FP64 sums of u = expm1(-d) and FP32 min/max of
// the new logU accumulated in S.red (initialised by the caller: 0, 0, +inf,
// -inf), then reduced in a fixed order: red[0] sum u, red[kBlock] sum u^2,
// red[2 kBlock] min, red[3 kBlock] max.
HD void eval_chunks(const VolDesc& V, const LevelDesc& LD, const Slab sl, const EvalLds& S) {
    const int Y = V.W[1], Z = V.W[2];
    const int Cz = LD.C[2];
    const int R = rows_per_eval_chunk(Y, Z);
    const int x = sl.x;
    const int yend = sl.y0 + sl.ny;
    for (int y0 = sl.y0; y0 < yend; y0 += R) {
        const int ny = (yend - y0 < R) ? yend - y0 : R;
        const size_t base = ((size_t)x * Y + y0) * Z;
        BLOCK_FOR(t)
        for (int pr = t; pr < ny * Cz; pr += kBlock) {
            const int r = (int)fdiv((uint32_t)pr, LD.mCz), c = pr - r * Cz;
            const int tr = y0 - sl.y0 + r;  // tile row
            const N4_LDS float* yw = S.yw + 4 * tr;
            const int b0 = S.yb[tr];
            float br = 0.0f;
            for (int k = 0; k < 4; ++k) br += yw[k] * S.Ar[(b0 + k) * Cz + c];
            S.Br[pr] = br;
        }
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        double s1 = S.red[t], s2 = S.red[kBlock + t];
        // min/max of FP32 values are exact in FP32
        float mn = (float)S.red[2 * kBlock + t], mx = (float)S.red[3 * kBlock + t];
        // aligned voxel quads (one 16-byte logU load and one 4-byte mask
        // load per 4 voxels: a quarter of the memory instructions of
        // per-voxel accesses), kEvalQuads quads in flight per thread, logU
        // read whatever the mask; a quad wholly inside the chunk is stored
        // with one 16-byte store, a quad straddling its ends voxel by voxel
        // (the voxels outside belong to other tiles)
        constexpr int kEvalQuads = 2;
        const size_t vend = base + (size_t)ny * Z;
        const size_t q0 = base & ~(size_t)3;
        const int nqd = (int)((vend - q0 + 3) >> 2);
        for (int i0 = 0; i0 < nqd; i0 += kEvalQuads * kBlock) {
            F4 xs[kEvalQuads];
            uint32_t ms[kEvalQuads];
            UNROLL for (int u = 0; u < kEvalQuads; ++u) {
                const int qi = i0 + u * kBlock + t;
                const size_t w = q0 + 4 * (size_t)qi;
                const bool ok = qi < nqd;
                ms[u] = ok ? *reinterpret_cast<const N4_GLOBAL uint32_t*>(gp(V.wmask) + w) : 0u;
                xs[u] = ok ? load_quad(V.logU + w) : F4{{0.0f, 0.0f, 0.0f, 0.0f}};
            }
            UNROLL for (int u = 0; u < kEvalQuads; ++u) {
                if (!ms[u]) continue;
                const size_t wq = q0 + 4 * (size_t)(i0 + u * kBlock + t);
                const bool full = wq >= base && wq + 4 <= vend;
                F4 o = xs[u];
                UNROLL for (int j = 0; j < 4; ++j) {
                    const size_t w = wq + j;
                    if (w < base || w >= vend || !((ms[u] >> (8 * j)) & 0xffu)) continue;
                    const int v = (int)(w - base);
                    const int r = (int)fdiv((uint32_t)v, V.mW2), z = v - r * Z;
                    const N4_LDS float* wz = S.zw + 4 * z;
                    const N4_LDS float* brr = S.Br + r * Cz + S.zb[z];
                    const float d = wz[0] * brr[0] + wz[1] * brr[1] + wz[2] * brr[2] + wz[3] * brr[3];
                    const float lg = xs[u].v[j] - d;
                    o.v[j] = lg;
                    if (!full) gp(V.logU)[w] = lg;
                    const double uu = (double)expm1f(-d);
                    s1 += uu;
                    s2 += uu * uu;
                    mn = fminf(mn, lg);
                    mx = fmaxf(mx, lg);
                }
                if (full) store_quad(V.logU + wq, o);
            }
        }
        S.red[t] = s1;
        S.red[kBlock + t] = s2;
        S.red[2 * kBlock + t] = mn;
        S.red[3 * kBlock + t] = mx;
        BLOCK_END
        BLOCK_SYNC();
    }
    tree_reduce4(S.red, 0, 0, 1, 2);
}

HD void body_eval(const VolDesc* vols, const Slab* slabs, RunParams rp, int bid, N4_LDS unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    const VolState* st = rp.states + sl.vol;
    if (st->s_kind != kStepIter) return;
    const int level = st->s_lvl;
    const LevelDesc& LD = V.lev[level];
    const int Y = V.W[1], Z = V.W[2];
    const int Cy = LD.C[1], Cz = LD.C[2];
    const int CyCz = Cy * Cz;
    const int R = rows_per_eval_chunk(Y, Z);
    EvalLds S;
    carve_eval(smem, Z, R, V.wtile_rows, Cy, Cz, &S);
    const int x = sl.x;
    const AxisEntry ex = LD.ax[0][x];
    const float* LR = V.lresf + (size_t)ex.b * CyCz;  // the 4 increment planes slab x touches
    BLOCK_FOR(t)
    for (int pr = t; pr < CyCz; pr += kBlock) {
        float ar = 0.0f;
        for (int k = 0; k < 4; ++k) ar += ex.w[k] * LR[k * CyCz + pr];
        S.Ar[pr] = ar;
    }
    for (int z = t; z < Z; z += kBlock) {
        const AxisEntry& e = LD.ax[2][z];
        for (int k = 0; k < 4; ++k) S.zw[4 * z + k] = e.w[k];
        S.zb[z] = e.b;
    }
    // y factors of the tile's rows, in the same round of loads
    for (int r = t; r < sl.ny; r += kBlock) {
        const AxisEntry& e = LD.ax[1][sl.y0 + r];
        for (int k = 0; k < 4; ++k) S.yw[4 * r + k] = e.w[k];
        S.yb[r] = e.b;
    }
    S.red[t] = 0.0;
    S.red[kBlock + t] = 0.0;
    S.red[2 * kBlock + t] = INFINITY;
    S.red[3 * kBlock + t] = -INFINITY;
    BLOCK_END
    BLOCK_SYNC();
    eval_chunks(V, LD, sl, S);
    BLOCK_FOR(t)
    if (t == 0) {
        double* sp = V.spart + (size_t)tile_id(V, sl) * 4;
        sp[0] = S.red[0];
        sp[1] = S.red[kBlock];
        sp[2] = S.red[2 * kBlock];
        sp[3] = S.red[3 * kBlock];
    }
    BLOCK_END
}

// ----------------------------------------------------------------- recon --
// Final field at full resolution from the last lattice (N4.py:329-330,
// cudaN4.cu:1700-1729): separable evaluation, then normalised image
// (thresholded input / exp(field)), optional exp(field) and field.
struct ReconLds {
    N4_LDS float* A;
    N4_LDS float* B;
    N4_LDS float* zw;
    N4_LDS int* zb;
    N4_LDS float* yw;  // [tile rows][4] y factors of the tile's rows
    N4_LDS int* yb;
};

// Rows per y-contracted chunk of a reconstruction tile: the whole tile when
// its rows of Cz floats fit kReconChunkFloats, so a tile streams with no
// barrier between its voxels.
constexpr int kReconChunkFloats = 8192;
HD int recon_chunk_rows(int tile_rows, int Cz) {
    int r = kReconChunkFloats / (Cz > 0 ? Cz : 1);
    if (r < 1) r = 1;
    return r < tile_rows ? r : tile_rows;
}

HD size_t carve_recon(N4_LDS unsigned char* base, int Z, int R, int Yt, int Cy, int Cz, ReconLds* o) {
    Carve c{base, 0};
    ReconLds L;
    L.A = c.take<float>((size_t)Cy * Cz);
    L.B = c.take<float>((size_t)R * Cz);
    L.zw = c.take<float>((size_t)4 * Z);
    L.zb = c.take<int>(Z);
    L.yw = c.take<float>((size_t)4 * Yt);
    L.yb = c.take<int>(Yt);
    if (o) *o = L;
    return align16(c.off);
}

HD void body_recon(const VolDesc* vols, const Slab* slabs, int level, RunParams rp, int bid,
                   N4_LDS unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    const LevelDesc& LD = V.lev[level];
    const int Y = V.F[1], Z = V.F[2];
    const int Cy = LD.C[1], Cz = LD.C[2];
    const int CyCz = Cy * Cz;
    const int R = recon_chunk_rows(V.recon_rows, Cz);
    ReconLds S;
    carve_recon(smem, Z, R, V.recon_rows, Cy, Cz, &S);
    const int x = sl.x;
    const AxisEntry ex = V.fax[0][x];
    // set-up: x-contracted planes, and the z and (tile) y factor tables, all
    // from one round of global loads
    BLOCK_FOR(t)
    for (int pr = t; pr < CyCz; pr += kBlock) {
        float a = 0.0f;
        for (int k = 0; k < 4; ++k) a += ex.w[k] * V.latf[(size_t)(ex.b + k) * CyCz + pr];
        S.A[pr] = a;
    }
    for (int z = t; z < Z; z += kBlock) {
        const AxisEntry& e = V.fax[2][z];
        for (int k = 0; k < 4; ++k) S.zw[4 * z + k] = e.w[k];
        S.zb[z] = e.b;
    }
    for (int r = t; r < sl.ny; r += kBlock) {
        const AxisEntry& e = V.fax[1][sl.y0 + r];
        for (int k = 0; k < 4; ++k) S.yw[4 * r + k] = e.w[k];
        S.yb[r] = e.b;
    }
    BLOCK_END
    BLOCK_SYNC();
    const int yend = sl.y0 + sl.ny;
    for (int y0 = sl.y0; y0 < yend; y0 += R) {
        const int ny = (yend - y0 < R) ? yend - y0 : R;
        const size_t base = ((size_t)x * Y + y0) * Z;
        BLOCK_FOR(t)
        for (int pr = t; pr < ny * Cz; pr += kBlock) {
            const int r = (int)fdiv((uint32_t)pr, LD.mCz), c = pr - r * Cz;
            const int tr = y0 - sl.y0 + r;  // tile row
            const N4_LDS float* yw = S.yw + 4 * tr;
            const int b0 = S.yb[tr];
            float b = 0.0f;
            for (int k = 0; k < 4; ++k) b += yw[k] * S.A[(b0 + k) * Cz + c];
            S.B[pr] = b;
        }
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        const N4_GLOBAL float* im = gp(V.im);
        N4_GLOBAL float* o_lb = gp(V.out_logbias);
        N4_GLOBAL float* o_b = gp(V.out_bias);
        N4_GLOBAL float* o_n = gp(V.out_norm);
        // consecutive lanes on consecutive voxels, kReconLoads image loads
        // in flight per thread (8 measured 9% faster than 16 once the
        // division became a product)
        constexpr int kReconLoads = 8;
        for (int v0 = 0; v0 < ny * Z; v0 += kReconLoads * kBlock) {
            float xin[kReconLoads];
            UNROLL for (int u = 0; u < kReconLoads; ++u) {
                const int v = v0 + u * kBlock + t;
                xin[u] = (o_n && v < ny * Z) ? im[base + v] : 0.0f;
            }
            UNROLL for (int u = 0; u < kReconLoads; ++u) {
                const int v = v0 + u * kBlock + t;
                if (v >= ny * Z) continue;
                const int r = (int)fdiv((uint32_t)v, V.mF2), z = v - r * Z;
                const N4_LDS float* wz = S.zw + 4 * z;
                const N4_LDS float* br = S.B + r * Cz + S.zb[z];
                const float f = wz[0] * br[0] + wz[1] * br[1] + wz[2] * br[2] + wz[3] * br[3];
                if (o_lb) o_lb[base + v] = f;
                if (o_b) o_b[base + v] = expf(f);
                if (o_n) {
                    // thresholded image / exp(f) as a product with exp(-f):
                    // no IEEE division per voxel (within 2 ulp of the quotient)
                    const float xo = (xin[u] <= rp.low_f) ? 0.0f : xin[u];  // N4.py:285
                    o_n[base + v] = xo * expf(-f);
                }
            }
        }
        BLOCK_END
        BLOCK_SYNC();
    }
}

// ------------------------------------------------------------------ flat --
// 1-D helpers for the stage entry points.  Slab x covers voxels
// [x*kChunkVox, min((x+1)*kChunkVox, n)) of V.logU / V.wmask.  The
// reductions write spart[x] (reduced on the host):
//   minmax (count, count, min, max); sum (sum, count, min, max);
//   cv (count, sum u, sum u^2, -) with u = exp(a-b) - 1; remap writes out.
HD void body_flat(const VolDesc* vols, const Slab* slabs, int op, const float* a, const float* b,
                  float* out, RunParams rp, int bid, N4_LDS unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    Carve cv{smem, 0};
    N4_LDS double* red = cv.take<double>(4 * kBlock);
    N4_LDS double* E = cv.take<double>(rp.nbins);
    const int64_t n = V.nvox_w;
    const int64_t v0 = (int64_t)sl.x * kChunkVox;
    const int nb = rp.nbins;
    if (op == kFlatRemap) {
        const double bmin = V.st->bin_min, slope = V.st->slope;
        BLOCK_FOR(t)
        for (int i = t; i < nb; i += kBlock) E[i] = V.etab[i];
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        for (int k = t; k < kChunkVox; k += kBlock) {
            const int64_t v = v0 + k;
            if (v >= n) break;
            const bool m = V.wmask ? (V.wmask[v] != 0) : true;
            float o = 0.0f;
            if (m) {
                if (V.st->degenerate) {
                    o = a[v];
                } else {
                    double c = ((double)a[v] - bmin) / slope;
                    c = fmin(fmax(c, 0.0), (double)nb);
                    const double fl = floor(c);
                    const int idx = (int)fl;
                    o = (float)((idx < nb - 1) ? E[idx] + (E[idx + 1] - E[idx]) * (c - fl) : E[nb - 1]);
                }
            }
            out[v] = o;
        }
        BLOCK_END
        return;
    }
    BLOCK_FOR(t)
    double r0 = 0.0, r1 = 0.0, r2 = (op == kFlatCV) ? 0.0 : INFINITY, r3 = -INFINITY;
    for (int k = t; k < kChunkVox; k += kBlock) {
        const int64_t v = v0 + k;
        if (v >= n) break;
        const bool m = V.wmask ? (V.wmask[v] != 0) : true;
        if (!m) continue;
        if (op == kFlatCV) {
            const double u = (double)expm1f(a[v] - b[v]);
            r0 += 1.0;
            r1 += u;
            r2 += u * u;
        } else {
            const double xv = (double)a[v];
            if (op == kFlatSum) {
                r0 += xv;
                r1 += 1.0;
            } else {
                r0 += 1.0;
                r1 += 1.0;
            }
            r2 = fmin(r2, xv);
            r3 = fmax(r3, xv);
        }
    }
    red[t] = r0;
    red[kBlock + t] = r1;
    red[2 * kBlock + t] = r2;
    red[3 * kBlock + t] = r3;
    BLOCK_END
    if (op == kFlatCV)
        tree_reduce4(red, 0, 0, 0, 2);
    else
        tree_reduce4(red, 0, 0, 1, 2);
    BLOCK_FOR(t)
    if (t == 0) {
        // min/max partials play the role of the prep partials (level 0 setup)
        double* sp = (op == kFlatMinMax ? V.ppart : V.spart) + (size_t)sl.x * 4;
        sp[0] = red[0];
        sp[1] = red[kBlock];
        sp[2] = red[2 * kBlock];
        sp[3] = red[3 * kBlock];
    }
    BLOCK_END
}

}  // namespace n4

#ifndef N4_EMULATE
#include "n4_mfma.h"
#endif

namespace n4 {

// ============================================================ persistent ==
// The whole level/iteration loop of every volume in one launch (between the
// prep + level-0 set-up launches and the reconstruction launch), in place
// of the host-paced chain of five launches per step.  Each volume has a
// team of consecutive workgroups; workgroup `rank` owns the volume's tiles
// [t0, t1) for the whole correction and runs them through the same tile
// code as the step kernels, so the results are bitwise those of the step
// path.  Every workgroup of a team runs the volume's state machine itself:
// the decisions (step kind, convergence, early exit, N4.py:296-322) come
// from the same reduction of the same partials in the same order, so they
// agree without being broadcast.  Per iteration step (reference loop
// cudaN4.cu:1517-1657):
//   B1  eval partials of the previous step are in: reduce (min, max, CV) ->
//       step kind
//       histogram of the own tiles -> one fixed-point copy of the bins
//   B2  histogram complete: each workgroup sums the copies and runs the
//       Wiener step itself (E table in its LDS); fit numerator of the own
//       tiles -> fixed-point copies
//   B3  numerator complete: each workgroup updates the control points it
//       owns (a 1/nteam share: lres = num/(den+eps), lattice += lres, lresf;
//       copies cleared); the leader clears the histogram copies
//   B4  increments complete: each workgroup loads the increment planes its
//       tiles touch, evaluates its tiles (logU -= d) and publishes the tiles'
//       partials
// A level transition refines the lattice (leader), scatters the level's
// denominator (own tiles) and, after a barrier, moves it to den (owners).
constexpr uint32_t kSpinLimit = 1u << 23;  // polls of a team barrier before it gives up (seconds)

// Phase timer of a team's leader: device clock (100 MHz), thread 0 only,
// accumulated into V.tphase when the caller asked for it.
HD int threadIdx_x() {
#ifdef N4_DEVICE_PASS
    return (int)threadIdx.x;
#else
    return 0;
#endif
}

struct PhaseTimer {
    uint64_t* slots;
    uint64_t last;
    HD static uint64_t now() {
#ifdef N4_DEVICE_PASS
        return __builtin_amdgcn_s_memrealtime();
#else
        return 0;
#endif
    }
    HD void mark(int phase) {
#ifdef N4_DEVICE_PASS
        if (!slots || threadIdx.x != 0) return;
        const uint64_t t = now();
        slots[phase] += t - last;
        last = t;
#else
        (void)phase;
#endif
    }
};

HD void spin_pause() {
#if defined(N4_DEVICE_PASS)
    __builtin_amdgcn_s_sleep(2);
#elif defined(N4_EMULATE)
    std::this_thread::yield();
#endif
}

// Team barrier number gen + 1.  Every wave first waits for its own memory
// operations to complete (every cross-workgroup write is an agent-scope
// atomic or store, so completion is visibility); the workgroup then
// synchronises and thread 0 arrives on the team's counter and polls it
// until all nteam arrivals are in.  Bounded: after kSpinLimit polls, or
// once a teammate has given up, it raises the team's abort flag and
// returns false (the workgroup then leaves; the host reports the error).
HD bool team_barrier(const VolDesc& V, int nteam, uint32_t& gen, N4_LDS int* bc) {
#ifdef N4_DEVICE_PASS
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
#endif
    BLOCK_SYNC();
    BLOCK_FOR(t)
    if (t == 0) {
        int ok = 1;
        if (nteam > 1) {
#ifdef N4_DEVICE_PASS
            __hip_atomic_fetch_add(V.team_ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
            __atomic_fetch_add(V.team_ctr, 1u, __ATOMIC_SEQ_CST);
#endif
            const uint32_t target = (gen + 1u) * (uint32_t)nteam;
            for (uint32_t spins = 1; ld_agent(V.team_ctr) < target; ++spins) {
                spin_pause();
                if ((spins & 63u) == 0 && (spins > kSpinLimit || ld_agent(V.team_ctr + 1) != 0)) {
                    ok = 0;
                    break;
                }
            }
            if (!ok) st_agent(V.team_ctr + 1, 1u);
        } else if (ld_agent(V.team_ctr + 1) != 0) {
            ok = 0;
        }
        bc[0] = ok;
    }
    BLOCK_END
    BLOCK_SYNC();
    ++gen;
    return bc[0] != 0;
}

// Reduction of the tiles' partial quadruples, as reduce_partials (same
// fold order, same results) but with agent-scope loads.
HD void reduce_partials_team(const double* part, int nslab, N4_LDS double* red) {
    BLOCK_FOR(t)
    double r0 = 0.0, r1 = 0.0, r2 = INFINITY, r3 = -INFINITY;
    for (int x0 = t; x0 < nslab; x0 += 8 * kBlock) {
        double v[8][4];
        UNROLL for (int u = 0; u < 8; ++u) {
            const int xx = x0 + u * kBlock;
            const double* sp = part + (size_t)(xx < nslab ? xx : 0) * 4;
            UNROLL for (int k = 0; k < 4; ++k) v[u][k] = ld_agent(sp + k);
        }
        UNROLL for (int u = 0; u < 8; ++u) {
            if (x0 + u * kBlock >= nslab) break;
            r0 += v[u][0];
            r1 += v[u][1];
            r2 = fmin(r2, v[u][2]);
            r3 = fmax(r3, v[u][3]);
        }
    }
    red[t] = r0;
    red[kBlock + t] = r1;
    red[2 * kBlock + t] = r2;
    red[3 * kBlock + t] = r3;
    BLOCK_END
    tree_reduce4(red, 0, 0, 1, 2);
}

// The control points a team member owns at a level: [o0, o1).
HD void owned_range(int L, int rank, int nteam, int* o0, int* o1) {
    *o0 = (int)((int64_t)L * rank / nteam);
    *o1 = (int)((int64_t)L * (rank + 1) / nteam);
}

// Owners' share of a step's lattice update (update_cp with agent-scope
// accesses throughout: the copies were filled by the team, and den, the
// lattice and lresf are read by other members).
HD void owner_update(const VolDesc& V, int level, bool trans, double eps, int o0, int o1) {
    BLOCK_FOR(t)
    for (int l = o0 + t; l < o1; l += kBlock) update_cp<true, true>(V, l, trans, level, eps);
    BLOCK_END
}

struct PersistLds {
    N4_LDS double* red;  // [4 kBlock]
    N4_LDS int* bc;      // barrier broadcast
    N4_LDS float* lres;  // (union) increment planes [pa, pa + np) x Cy*Cz of the current level
    FitLds F;            // E, zwc, zs, ys, ywc, ex persist; g, T1, T2 in the union
    EvalLds Ev;          // zw, zb, yw, yb (whole y extent) persist; Ar, Br in the union
    N4_LDS unsigned long long* h;  // (union) histogram copies
    N4_LDS unsigned char* u;       // the union: histogram | Wiener | fit chunk | increments + eval chunk
#ifndef N4_EMULATE
    FitMfmaLds M;  // MFMA fit (prec > 0): E, Pz, ex persist; Py (tile rows), red, Gh, Gl in the union
#endif
};

// Largest level value of C[d] * support_len(n, C[d]) (the fit's per-level
// factor runs).
HD size_t max_support_run(const VolDesc& V, int d, int n) {
    size_t m = 0;
    for (int l = 0; l < V.nlevels; ++l) {
        const size_t r = (size_t)V.lev[l].C[d] * support_len(n, V.lev[l].C[d]);
        m = r > m ? r : m;
    }
    return m;
}

// prec > 0: the fit runs on MFMA (one 16-wide tile of control points along
// y and z) instead of the FP32 factor runs.
HD size_t carve_persist(N4_LDS unsigned char* base, const VolDesc& V, int nbins, int P, int prec, PersistLds* o) {
    const int Y = V.W[1], Z = V.W[2];
    const LevelDesc& LF = V.lev[V.nlevels - 1];
    const int Cy = LF.C[1], Cz = LF.C[2];
    Carve c{base, 0};
    PersistLds L;
    L.red = c.take<double>((size_t)4 * kBlock);
    L.bc = c.take<int>(4);
    L.F.E = c.take<float>(nbins);
    L.F.ex = c.take<AxisEntry>(1);
    if (prec == 0) {
        L.F.zwc = c.take<float>(max_support_run(V, 2, Z));
        L.F.zs = c.take<int>((size_t)2 * Cz);
        L.F.ys = c.take<int>((size_t)2 * Cy);
        L.F.ywc = c.take<float>(max_support_run(V, 1, Y));
    }
#ifndef N4_EMULATE
    if (prec > 0) {
        L.M.E = L.F.E;
        L.M.ex = L.F.ex;
        const int zs = round_up(Z, 32) + 8;
        L.M.Pz.stride = zs;
        L.M.Pz.has_lo = true;
        L.M.Pz.hi = c.take<unsigned short>((size_t)16 * zs);
        L.M.Pz.lo = c.take<unsigned short>((size_t)16 * zs);
    }
#endif
    L.Ev.zw = c.take<float>((size_t)4 * Z);
    L.Ev.zb = c.take<int>(Z);
    L.Ev.yw = c.take<float>((size_t)4 * Y);
    L.Ev.yb = c.take<int>(Y);
    L.Ev.red = L.red;
    c.off = align16(c.off);
    const size_t u0 = c.off;
    L.u = c.at(u0);
    size_t um = 0;
    {  // histogram
        Carve h{L.u, 0};
        L.h = h.take<unsigned long long>((size_t)kHistLanes * (nbins | 1));
        um = h.off > um ? h.off : um;
    }
    {  // Wiener (carved again by wiener_table from the union base)
        const size_t w = align16(4 * (size_t)P * sizeof(D2));
        um = w > um ? w : um;
    }
    if (prec == 0) {  // fit chunk
        Carve f{L.u, 0};
        const int R = rows_per_chunk(Y, Z);
        L.F.g = f.take<float>((size_t)R * Z);
        L.F.T1 = f.take<float>((size_t)R * Cz);
        L.F.T2 = f.take<float>((size_t)Cy * Cz);
        um = f.off > um ? f.off : um;
    }
#ifndef N4_EMULATE
    if (prec > 0) {  // MFMA fit of a tile: Py over its rows, wave combine, g staging
        Carve f{L.u, 0};
        const int ys = round_up(V.wtile_rows, 16) + 4;
        L.M.Py.stride = ys;
        L.M.Py.has_lo = true;
        L.M.Py.hi = f.take<unsigned short>((size_t)16 * ys);
        L.M.Py.lo = f.take<unsigned short>((size_t)16 * ys);
        L.M.red = f.take<float>(4 * 256);  // a slot per wave (fit_mfma_tile<PREC, 1, 1>)
        L.M.Gh = f.take<unsigned short>((size_t)4 * 16 * kGStride);
        L.M.Gl = prec == 2 ? f.take<unsigned short>((size_t)4 * 16 * kGStride) : nullptr;
        um = f.off > um ? f.off : um;
    }
#endif
    {  // eval: the increment planes the tiles touch (at most all of them), then the chunk
        Carve e{L.u, 0};
        L.lres = e.take<float>((size_t)LF.L);
        L.Ev.Ar = e.take<float>((size_t)Cy * Cz);
        L.Ev.Br = e.take<float>((size_t)rows_per_eval_chunk(Y, Z) * Cz);
        um = e.off > um ? e.off : um;
    }
    if (o) *o = L;
    return align16(u0 + um);
}

// The level's tables that persist across its iterations: the fit's factor
// runs (w^2 for the denominator, else w^3/sum w^2) and the evaluation's z
// and y weights.
template <int PREC>
HD void persist_tables(const VolDesc& V, const LevelDesc& LD, const PersistLds& S, bool den) {
    const int Y = V.W[1], Z = V.W[2];
    BLOCK_FOR(t)
    if constexpr (PREC == 0) {
        fit_tables_t(LD, S.F, Y, Z, den, t);
    } else {
#ifndef N4_EMULATE
        fit_mfma_tables_t<1, 1>(LD, S.M, Z, 0, 0, den ? 2 : 1, true, false, t);
#endif
    }
    if (!den) {
        for (int z = t; z < Z; z += kBlock) {
            const AxisEntry& e = LD.ax[2][z];
            for (int k = 0; k < 4; ++k) S.Ev.zw[4 * z + k] = e.w[k];
            S.Ev.zb[z] = e.b;
        }
        for (int y = t; y < Y; y += kBlock) {
            const AxisEntry& e = LD.ax[1][y];
            for (int k = 0; k < 4; ++k) S.Ev.yw[4 * y + k] = e.w[k];
            S.Ev.yb[y] = e.b;
        }
    }
    BLOCK_END
    BLOCK_SYNC();
}

// Fit of the own tiles (NUM: residual with the LDS E table; DEN: mask with
// the w^2 factors), scattered into the numerator copies as the step
// kernels do.
template <int PREC>
HD void persist_fit(const VolDesc& V, const LevelDesc& LD, const Slab* my, int nmy, const PersistLds& S, bool den,
                    bool degenerate, const Remap& rm) {
    const int CyCz = LD.C[1] * LD.C[2];
    const int gmode = den ? (int)kFitDen : (int)kFitNum;
#ifndef N4_EMULATE
    if constexpr (PREC > 0) {
        const int t = (int)threadIdx.x, wv = t >> 6;
        for (int i = 0; i < nmy; ++i) {
            const Slab sl = my[i];
            uint8_t ms[4 * kFitRuns];
            float xs[4 * kFitRuns];
            fit_mfma_load(V, sl, 16 * wv, 0, den, ms, xs);
            fit_mfma_tables_t<1, 1>(LD, S.M, V.W[2], sl.y0, sl.ny, den ? 2 : 1, false, true, t);
            if (t < (int)(sizeof(AxisEntry) / 4))
                reinterpret_cast<N4_LDS int32_t*>(S.M.ex)[t] = reinterpret_cast<const int32_t*>(&LD.ax[0][sl.x])[t];
            __syncthreads();
            fit_mfma_tile<PREC, 1, 1>(V, LD, sl, S.M, gmode, den, true, degenerate, rm, ms, xs);
            __syncthreads();
        }
        return;
    }
#endif
    for (int i = 0; i < nmy; ++i) {
        const Slab sl = my[i];
        BLOCK_FOR(t)
        for (int k = t; k < CyCz; k += kBlock) S.F.T2[k] = 0.0f;
        if (t < (int)(sizeof(AxisEntry) / 4))
            reinterpret_cast<N4_LDS int32_t*>(S.F.ex)[t] = reinterpret_cast<const int32_t*>(&LD.ax[0][sl.x])[t];
        BLOCK_END
        BLOCK_SYNC();
        fit_chunks(V, LD, sl, S.F, gmode, den, degenerate, rm);
        fit_scatter(V, LD, sl, S.F, true, den);
        BLOCK_SYNC();
    }
}

template <int PREC>
HD void body_persist(const VolDesc* vols, const Slab* slabs, const PWork* work, RunParams rp, int bid,
                     N4_LDS unsigned char* smem) {
    const PWork W = work[bid];
    const VolDesc& V = vols[W.vol];
    VolState* st = V.st;
    const bool leader = W.rank == 0;
    const int nb = rp.nbins;
    const int ntiles = V.W[0] * V.wtiles_x;
    const Slab* my = slabs + W.t0;
    const int nmy = W.t1 - W.t0;
    PersistLds S;
    carve_persist(smem, V, nb, rp.P, PREC, &S);
    const double n = (double)st->n_mask_work;  // written by the level-0 set-up launch
    uint32_t gen = 0;
    int mode = kModeTrans, lvl = 0, it = 0;
    int pa = 0;  // first increment plane the own tiles touch (current level)
    bool need_b1 = false;
    PhaseTimer pt{leader ? V.tphase : nullptr, PhaseTimer::now()};
    for (;;) {
        if (need_b1 && !team_barrier(V, W.nteam, gen, S.bc)) return;  // B1
        pt.mark(kPhWait);
        reduce_partials_team(V.spart, ntiles, S.red);
        const double mn = S.red[2 * kBlock], mx = S.red[3 * kBlock];
        double cvv = 0.0;
        int kind = kStepTrans, s_lvl = lvl;
        bool level_end = false;
        if (mode == kModeIter) {
            if (it > 0) cvv = cv_from_sums(S.red[0], S.red[kBlock], n);
            const bool stop = it > 0 && !rp.fixed_iters && !(cvv > rp.thr);
            level_end = stop || it >= V.max_iter[lvl];
            if (!level_end) {
                kind = kStepIter;
            } else if (lvl + 1 < V.nlevels) {
                kind = kStepTrans;
                s_lvl = lvl + 1;
            } else {
                kind = kStepFinal;
            }
        }
        const double slope = (mx - mn) / (double)(nb - 1);
        const bool degenerate = !(n > 0.0) || !(slope > 0.0);
        BLOCK_SYNC();  // S.red is read above before any later phase rewrites it
        if (leader) {
            BLOCK_FOR(t)
            if (t == 0) {
                if (mode == kModeIter && it > 0) record_cv(V, lvl, it, cvv);
                if (level_end) {
                    st->iters[lvl] = it;
                    if (it > 0) {
                        st->lastcv[lvl] = cvv;
                        st->cv = cvv;
                    }
                }
                st->s_kind = kind;
                st->s_lvl = s_lvl;
                st->s_it = it;
                st->bin_min = mn;
                st->bin_max = mx;
                st->slope = slope;
                st->degenerate = degenerate ? 1 : 0;
                st->passes += 1;
            }
            BLOCK_END
            pt.mark(kPhReduce);
            if (pt.slots && threadIdx_x() == 0) pt.slots[kPhSteps] += 1;
        }
        if (kind == kStepFinal) {
            if (leader) {
                const LevelDesc& FD = V.lev[V.nlevels - 1];
                const double* fin = V.lat[(V.nlevels - 1) & 1];
                float* rec = level_lattice_record(V, V.nlevels - 1);
                BLOCK_FOR(t)
                for (int l = t; l < FD.L; l += kBlock) {
                    const float f = (float)ld_agent(fin + l);
                    V.latf[l] = f;
                    if (rec) rec[l] = f;
                    if (V.out_lattice) V.out_lattice[l] = f;
                }
                if (t == 0) {
                    st->mode = kModeFinished;
                    st->done = 1;
                }
                BLOCK_END
            }
            return;
        }
        if (kind == kStepTrans) {
            const int l = s_lvl;
            const LevelDesc& LD = V.lev[l];
            if (l > 0 && leader) {
                BLOCK_FOR(t)
                refine_lattice<true>(V, l, t);
                BLOCK_END
            }
            persist_tables<PREC>(V, LD, S, true);
            persist_fit<PREC>(V, LD, my, nmy, S, true, false, Remap{S.F.E, 0.0f, 0.0f, nb});
            pt.mark(kPhTrans);
            if (!team_barrier(V, W.nteam, gen, S.bc)) return;
            pt.mark(kPhWait);
            int o0, o1;
            owned_range(LD.L, W.rank, W.nteam, &o0, &o1);
            owner_update(V, l, true, rp.eps, o0, o1);
            persist_tables<PREC>(V, LD, S, false);
            pa = nmy > 0 ? LD.ax[0][my[0].x].b : 0;
            pt.mark(kPhTrans);
            mode = kModeIter;
            lvl = l;
            it = 0;
            need_b1 = false;  // the partials are unchanged
            continue;
        }
        // ---- ITER(lvl, it)
        const LevelDesc& LD = V.lev[lvl];
        const int CyCz = LD.C[1] * LD.C[2];
        const int hs = nb | 1;
        if (!degenerate) {
            hist_zero(S.h, hs);
            BLOCK_SYNC();
            for (int i = 0; i < nmy; ++i) hist_tile(V, my[i], mn, slope, nb, S.h, hs, nullptr);
            BLOCK_SYNC();
            hist_flush(S.h, hs, nb, V.hist + (size_t)(W.rank % V.hist_copies) * nb);
        }
        pt.mark(kPhHist);
        if (!team_barrier(V, W.nteam, gen, S.bc)) return;  // B2
        pt.mark(kPhWait);
        if (!degenerate) {
            wiener_table<kWienerTeam>(V, mn, slope, rp, S.u, S.F.E);
            BLOCK_SYNC();
        }
        pt.mark(kPhWiener);
        persist_fit<PREC>(V, LD, my, nmy, S, false, degenerate, make_remap(S.F.E, mn, slope, nb));
        pt.mark(kPhFit);
        if (!team_barrier(V, W.nteam, gen, S.bc)) return;  // B3
        pt.mark(kPhWait);
        if (leader) {
            BLOCK_FOR(t)
            for (int i = t; i < V.hist_copies * nb; i += kBlock) st_agent(V.hist + i, (uint64_t)0);
            BLOCK_END
        }
        {
            int o0, o1;
            owned_range(LD.L, W.rank, W.nteam, &o0, &o1);
            owner_update(V, lvl, false, rp.eps, o0, o1);
        }
        pt.mark(kPhUpdate);
        if (!team_barrier(V, W.nteam, gen, S.bc)) return;  // B4
        pt.mark(kPhWait);
        if (nmy > 0) {
            const int pb = LD.ax[0][my[nmy - 1].x].b + 3;
            BLOCK_FOR(t)
            for (int i = t; i < (pb - pa + 1) * CyCz; i += kBlock)
                S.lres[i] = ld_agent(V.lresf + (size_t)pa * CyCz + i);
            BLOCK_END
            BLOCK_SYNC();
        }
        pt.mark(kPhUpdate);
        for (int i = 0; i < nmy; ++i) {
            const Slab sl = my[i];
            const AxisEntry ex = LD.ax[0][sl.x];
            const N4_LDS float* LR = S.lres + (size_t)(ex.b - pa) * CyCz;
            EvalLds E = S.Ev;
            E.yw = S.Ev.yw + 4 * sl.y0;  // tile-relative y tables
            E.yb = S.Ev.yb + sl.y0;
            BLOCK_FOR(t)
            for (int pr = t; pr < CyCz; pr += kBlock) {
                float ar = 0.0f;
                for (int k = 0; k < 4; ++k) ar += ex.w[k] * LR[k * CyCz + pr];
                E.Ar[pr] = ar;
            }
            E.red[t] = 0.0;
            E.red[kBlock + t] = 0.0;
            E.red[2 * kBlock + t] = INFINITY;
            E.red[3 * kBlock + t] = -INFINITY;
            BLOCK_END
            BLOCK_SYNC();
            eval_chunks(V, LD, sl, E);
            BLOCK_FOR(t)
            if (t == 0) {
                double* sp = V.spart + (size_t)tile_id(V, sl) * 4;
                st_agent(sp, E.red[0]);
                st_agent(sp + 1, E.red[kBlock]);
                st_agent(sp + 2, E.red[2 * kBlock]);
                st_agent(sp + 3, E.red[3 * kBlock]);
            }
            BLOCK_END
            BLOCK_SYNC();
        }
        pt.mark(kPhEval);
        it += 1;
        need_b1 = true;
    }
}


// ---------------------------------------------------------------- sizing --

// MFMA tiles of 16 control points along one axis (0: not an MFMA launch).
static int cp_tiles(int c) { return (c + 15) / 16; }

size_t lds_prep() { return align16(4 * kBlock * sizeof(double)); }
size_t lds_level() { return align16(4 * kBlock * sizeof(double)); }
size_t lds_hist(int nbins) {
    return std::max(align16(4 * kBlock * sizeof(double)), align16((size_t)kHistLanes * (nbins | 1) * sizeof(uint64_t)));
}
size_t lds_wiener(int P) { return align16(4 * (size_t)P * sizeof(D2)); }
size_t lds_lattice() { return align16(kBlock * sizeof(double)); }
size_t lds_flat(int nbins) { return align16(4 * kBlock * sizeof(double)) + align16(nbins * sizeof(double)); }

size_t lds_fit(int nbins, const VolDesc& v, int level, int prec) {
    const LevelDesc& L = v.lev[level];
#ifndef N4_EMULATE
    if (prec)
        return carve_fit_mfma(nullptr, nbins, v.wtile_rows, v.W[2], cp_tiles(L.C[1]), cp_tiles(L.C[2]), prec,
                              nullptr);
#endif
    (void)prec;
    return carve_fit(nullptr, nbins, v.W[1], v.W[2], rows_per_chunk(v.W[1], v.W[2]), L.C[1], L.C[2], nullptr);
}

size_t lds_eval(const VolDesc& v, int level, int prec) {
    const LevelDesc& L = v.lev[level];
#ifndef N4_EMULATE
    if (prec)
        return carve_eval_mfma(nullptr, v.wtile_rows, v.W[2], L.C[1], L.C[2], cp_tiles(L.C[1]), cp_tiles(L.C[2]),
                               prec, nullptr);
#endif
    (void)prec;
    return carve_eval(nullptr, v.W[2], rows_per_eval_chunk(v.W[1], v.W[2]), v.wtile_rows, L.C[1], L.C[2], nullptr);
}

size_t lds_recon(const VolDesc& v) {
    const LevelDesc& L = v.lev[v.nlevels - 1];
    return carve_recon(nullptr, v.F[2], recon_chunk_rows(v.recon_rows, L.C[2]), v.recon_rows, L.C[1], L.C[2],
                       nullptr);
}

int mfma_tiles(int c) { return cp_tiles(c); }

size_t lds_persist(const VolDesc& v, int nbins, int P, int prec) {
    return carve_persist(nullptr, v, nbins, P, prec, nullptr);
}

// --------------------------------------------------------------- launches --
#ifndef N4_EMULATE

#define SMEM() ((N4_LDS unsigned char*)(smem_))

// One instance per shrink factor 1..4 (S = 0: any factor), so each keeps
// the register count of its own case.
template <int S>
__global__ __launch_bounds__(kBlock) void k_prep(const VolDesc* __restrict__ vols, const Slab* __restrict__ slabs,
                                                 RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    const Slab sl = slabs[blockIdx.x];
    body_prep_s<S>(vols[sl.vol], sl, rp, SMEM());
}

__global__ __launch_bounds__(kBlock) void k_level(const VolDesc* __restrict__ vols, int level, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_level(vols, level, rp, blockIdx.x, SMEM());
}

__global__ __launch_bounds__(kBlock) void k_hist(const VolDesc* __restrict__ vols, const Slab* __restrict__ slabs, int step, RunParams rp,
                                                 int32_t* dbg_idx) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_hist(vols, slabs, step, rp, dbg_idx, blockIdx.x, SMEM());
}

__global__ __launch_bounds__(kBlock) void k_wiener(const VolDesc* __restrict__ vols, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_wiener(vols, rp, blockIdx.x, SMEM());
}

__global__ __launch_bounds__(kBlock) void k_fit(const VolDesc* __restrict__ vols, const Slab* __restrict__ slabs, int level, int it,
                                                int mode, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_fit(vols, slabs, level, it, mode, rp, blockIdx.x, SMEM());
}

__global__ __launch_bounds__(kBlock) void k_lattice(const VolDesc* __restrict__ vols, int level, int mode, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_lattice(vols, level, mode, rp, blockIdx.x, blockIdx.y, SMEM());
}

__global__ __launch_bounds__(kBlock) void k_update(const VolDesc* __restrict__ vols, RunParams rp) {
    body_update(vols, rp, (int)blockIdx.x, (int)blockIdx.y);
}
// The evaluation and the reconstruction are held to 5 and 4 waves/SIMD
// (96 and 128 VGPRs), the occupancy their memory streams need.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_eval(const VolDesc* __restrict__ vols, const Slab* __restrict__ slabs, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_eval(vols, slabs, rp, blockIdx.x, SMEM());
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_recon(const VolDesc* __restrict__ vols, const Slab* __restrict__ slabs, int level,
                                                  RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_recon(vols, slabs, level, rp, blockIdx.x, SMEM());
}

__global__ __launch_bounds__(kBlock) void k_flat(const VolDesc* __restrict__ vols, const Slab* __restrict__ slabs, int op,
                                                 const float* a, const float* b, float* out, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_flat(vols, slabs, op, a, b, out, rp, blockIdx.x, SMEM());
}

// The one-tile instance (up to 16 control points along y and z) is held to
// 80 VGPRs in BF16 (6 waves per SIMD: 6 workgroups per CU, as many as its
// LDS allows for tiles of up to 72 x 72 voxels) and to 96 in split BF16
// (at 80 it spills).
template <int PREC, int MT, int NT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MT * NT == 1 ? (PREC == 1 ? 6 : 5) : 1))) void k_fit_mfma(const VolDesc* __restrict__ vols, const Slab* __restrict__ slabs, int level, int it,
                                                     int mode, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_fit_mfma<PREC, MT, NT>(vols, slabs, level, it, mode, rp, blockIdx.x, SMEM());
}

template <int PREC, int MT, int NT>
__global__ __launch_bounds__(kBlock) void k_eval_mfma(const VolDesc* __restrict__ vols, const Slab* __restrict__ slabs, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_eval_mfma<PREC, MT, NT>(vols, slabs, rp, blockIdx.x, SMEM());
}

// Launchers of every (precision, Cy tiles, Cz tiles) instance, indexed
// [prec-1][MT-1][NT-1].
typedef void (*FitLaunch)(int, size_t, hipStream_t, const VolDesc*, const Slab*, int, int, int, RunParams);
typedef void (*EvalLaunch)(int, size_t, hipStream_t, const VolDesc*, const Slab*, int, RunParams);

template <int PREC, int MT, int NT>
static void fit_mfma_launch(int n, size_t lds, hipStream_t s, const VolDesc* v, const Slab* sl, int level, int it,
                            int mode, RunParams rp) {
    hipLaunchKernelGGL((k_fit_mfma<PREC, MT, NT>), dim3(n), dim3(kBlock), lds, s, v, sl, level, it, mode, rp);
}
template <int PREC, int MT, int NT>
static void eval_mfma_launch(int n, size_t lds, hipStream_t s, const VolDesc* v, const Slab* sl, int,
                             RunParams rp) {
    hipLaunchKernelGGL((k_eval_mfma<PREC, MT, NT>), dim3(n), dim3(kBlock), lds, s, v, sl, rp);
}
#define N4_ROW(F, P, M) {F<P, M, 1>, F<P, M, 2>, F<P, M, 3>, F<P, M, 4>}
#define N4_TABLE(F) \
    {{N4_ROW(F, 1, 1), N4_ROW(F, 1, 2), N4_ROW(F, 1, 3), N4_ROW(F, 1, 4)}, \
     {N4_ROW(F, 2, 1), N4_ROW(F, 2, 2), N4_ROW(F, 2, 3), N4_ROW(F, 2, 4)}}
static const FitLaunch kFitMfma[2][kMfmaMaxTiles][kMfmaMaxTiles] = N4_TABLE(fit_mfma_launch);
static const EvalLaunch kEvalMfma[2][kMfmaMaxTiles][kMfmaMaxTiles] = N4_TABLE(eval_mfma_launch);
#undef N4_TABLE
#undef N4_ROW

static bool mfma_ok(const RunParams& rp, int mt, int nt) {
    return (rp.prec == 1 || rp.prec == 2) && mt >= 1 && mt <= kMfmaMaxTiles && nt >= 1 && nt <= kMfmaMaxTiles;
}

static int last_launch_status() { return (int)hipGetLastError(); }

// At least 3 waves per SIMD (at most 168 VGPRs, some spilled): the
// kernel holds every phase's body, and without the bound the compiler
// takes 209-256 VGPRs (1-2 waves per SIMD, fewer team members, more tiles
// each); measured 12% (FP32) and 38% (BF16) faster on a 64-volume batch.
template <int PREC>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_persist(const VolDesc* __restrict__ vols, const Slab* __restrict__ slabs,
                                                    const PWork* __restrict__ work, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem_);
    body_persist<PREC>(vols, slabs, work, rp, blockIdx.x, SMEM());
}

int persist_capacity(size_t lds, int prec, int* per_cu, int* cus) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) {
        if (prec == 1)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_persist<1>, kBlock, lds);
        else if (prec == 2)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_persist<2>, kBlock, lds);
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_persist<0>, kBlock, lds);
    }
    return (int)e;
}

int launch_persist(const VolDesc* vols, const Slab* slabs, const PWork* work, int nwg, RunParams rp, size_t lds,
                   void* stream) {
    if (nwg <= 0) return 0;
    const dim3 g(nwg), b(kBlock);
    const hipStream_t s = (hipStream_t)stream;
    if (rp.prec == 1)
        hipLaunchKernelGGL(k_persist<1>, g, b, lds, s, vols, slabs, work, rp);
    else if (rp.prec == 2)
        hipLaunchKernelGGL(k_persist<2>, g, b, lds, s, vols, slabs, work, rp);
    else
        hipLaunchKernelGGL(k_persist<0>, g, b, lds, s, vols, slabs, work, rp);
    return last_launch_status();
}

int launch_prep(const VolDesc* vols, const Slab* slabs, int nslab, int shrink, RunParams rp, void* stream) {
    if (nslab <= 0) return 0;
    const dim3 g(nslab), b(kBlock);
    const hipStream_t s = (hipStream_t)stream;
    switch (prep_class(shrink)) {
        case 1: hipLaunchKernelGGL(k_prep<1>, g, b, lds_prep(), s, vols, slabs, rp); break;
        case 2: hipLaunchKernelGGL(k_prep<2>, g, b, lds_prep(), s, vols, slabs, rp); break;
        case 3: hipLaunchKernelGGL(k_prep<3>, g, b, lds_prep(), s, vols, slabs, rp); break;
        case 4: hipLaunchKernelGGL(k_prep<4>, g, b, lds_prep(), s, vols, slabs, rp); break;
        default: hipLaunchKernelGGL(k_prep<0>, g, b, lds_prep(), s, vols, slabs, rp); break;
    }
    return last_launch_status();
}

int launch_level(const VolDesc* vols, int nvol, int level, RunParams rp, void* stream) {
    if (nvol <= 0) return 0;
    hipLaunchKernelGGL(k_level, dim3(nvol), dim3(kBlock), lds_level(), (hipStream_t)stream, vols, level, rp);
    return last_launch_status();
}

int launch_hist(const VolDesc* vols, const Slab* slabs, int nslab, int step, RunParams rp, int32_t* dbg_idx,
                size_t lds, void* stream) {
    if (nslab <= 0) return 0;
    hipLaunchKernelGGL(k_hist, dim3(nslab), dim3(kBlock), lds, (hipStream_t)stream, vols, slabs, step, rp, dbg_idx);
    return last_launch_status();
}

int launch_wiener(const VolDesc* vols, int nvol, RunParams rp, size_t lds, void* stream) {
    if (nvol <= 0) return 0;
    hipLaunchKernelGGL(k_wiener, dim3(nvol), dim3(kBlock), lds, (hipStream_t)stream, vols, rp);
    return last_launch_status();
}

int launch_fit(const VolDesc* vols, const Slab* slabs, int nslab, int level, int it, int mode, RunParams rp,
               size_t lds, int mt, int nt, void* stream) {
    if (nslab <= 0) return 0;
    if (rp.prec) {
        if (!mfma_ok(rp, mt, nt)) return (int)hipErrorInvalidValue;
        kFitMfma[rp.prec - 1][mt - 1][nt - 1](nslab, lds, (hipStream_t)stream, vols, slabs, level, it, mode, rp);
        return last_launch_status();
    }
    hipLaunchKernelGGL(k_fit, dim3(nslab), dim3(kBlock), lds, (hipStream_t)stream, vols, slabs, level, it, mode,
                       rp);
    return last_launch_status();
}

int launch_lattice(const VolDesc* vols, int nvol, int cx_max, int level, int mode, RunParams rp, void* stream) {
    if (nvol <= 0 || cx_max <= 0) return 0;
    hipLaunchKernelGGL(k_lattice, dim3(cx_max, nvol), dim3(kBlock), lds_lattice(), (hipStream_t)stream, vols, level,
                       mode, rp);
    return last_launch_status();
}

int launch_update(const VolDesc* vols, int nvol, int chunks, RunParams rp, void* stream) {
    if (nvol <= 0 || chunks <= 0) return 0;
    hipLaunchKernelGGL(k_update, dim3(chunks, nvol), dim3(kBlock), 0, (hipStream_t)stream, vols, rp);
    return last_launch_status();
}

int launch_eval(const VolDesc* vols, const Slab* slabs, int nslab, RunParams rp, size_t lds, int mt, int nt,
                void* stream) {
    if (nslab <= 0) return 0;
    if (rp.prec) {
        if (!mfma_ok(rp, mt, nt)) return (int)hipErrorInvalidValue;
        kEvalMfma[rp.prec - 1][mt - 1][nt - 1](nslab, lds, (hipStream_t)stream, vols, slabs, 0, rp);
        return last_launch_status();
    }
    hipLaunchKernelGGL(k_eval, dim3(nslab), dim3(kBlock), lds, (hipStream_t)stream, vols, slabs, rp);
    return last_launch_status();
}

int launch_recon(const VolDesc* vols, const Slab* slabs, int nslab, int level, RunParams rp, size_t lds,
                 void* stream) {
    if (nslab <= 0) return 0;
    hipLaunchKernelGGL(k_recon, dim3(nslab), dim3(kBlock), lds, (hipStream_t)stream, vols, slabs, level, rp);
    return last_launch_status();
}

// Float mask -> uint8 mask (nonzero = foreground, cudaN4.h:346),
// grid-stride.  vec (in 16-byte and out 4-byte aligned): four voxels per
// thread and iteration, one 16-byte load and one 4-byte store, the last
// n % 4 voxels by block 0; else one voxel per thread and iteration.
__global__ __launch_bounds__(kBlock) void k_mask_u8(const float* __restrict__ in, uint8_t* __restrict__ out,
                                                    uint64_t n, int vec) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t t0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (!vec) {
        for (uint64_t i = t0; i < n; i += stride) gp(out)[i] = gp(in)[i] != 0.0f ? 1 : 0;
        return;
    }
    const uint64_t nq = n >> 2;
    for (uint64_t q = t0; q < nq; q += stride) {
        const F4 v = *reinterpret_cast<const N4_GLOBAL F4*>(gp(in) + 4 * q);
        const uint32_t w = (v.v[0] != 0.0f ? 1u : 0u) | (v.v[1] != 0.0f ? 1u << 8 : 0u) |
                           (v.v[2] != 0.0f ? 1u << 16 : 0u) | (v.v[3] != 0.0f ? 1u << 24 : 0u);
        *reinterpret_cast<N4_GLOBAL uint32_t*>(gp(out) + 4 * q) = w;
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
        const uint64_t i = 4 * nq + threadIdx.x;
        gp(out)[i] = gp(in)[i] != 0.0f ? 1 : 0;
    }
}

int launch_mask_u8(const float* in, uint8_t* out, uint64_t n, void* stream) {
    if (n == 0) return 0;
    const int vec = ((uintptr_t)in % 16 == 0) && ((uintptr_t)out % 4 == 0);
    const uint64_t items = vec ? std::max<uint64_t>(n >> 2, 1) : n;
    const uint64_t blocks = std::min<uint64_t>((items + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_mask_u8, dim3((unsigned)blocks), dim3(kBlock), 0, (hipStream_t)stream, in, out, n, vec);
    return last_launch_status();
}

int launch_flat(const VolDesc* vols, const Slab* slabs, int nslab, int op, const float* a, const float* b,
                float* out, RunParams rp, void* stream) {
    if (nslab <= 0) return 0;
    hipLaunchKernelGGL(k_flat, dim3(nslab), dim3(kBlock), lds_flat(rp.nbins), (hipStream_t)stream, vols, slabs, op,
                       a, b, out, rp);
    return last_launch_status();
}

#else  // N4_EMULATE: host loops over blocks, used only by CPU unit tests

static unsigned char* emu_lds(size_t n) {
    static std::vector<unsigned char> buf;
    if (buf.size() < n + 64) buf.assign(n + 64, 0);
    return buf.data();
}

int launch_prep(const VolDesc* vols, const Slab* slabs, int nslab, int shrink, RunParams rp, void*) {
    unsigned char* b = emu_lds(lds_prep());
    for (int i = 0; i < nslab; ++i)
        if (prep_class(vols[slabs[i].vol].s) != prep_class(shrink)) return 1;  // a launch of mixed classes
    for (int i = 0; i < nslab; ++i) body_prep(vols, slabs, rp, i, b);
    return 0;
}

int launch_level(const VolDesc* vols, int nvol, int level, RunParams rp, void*) {
    unsigned char* b = emu_lds(lds_level());
    for (int i = 0; i < nvol; ++i) body_level(vols, level, rp, i, b);
    return 0;
}

int launch_hist(const VolDesc* vols, const Slab* slabs, int nslab, int step, RunParams rp, int32_t* dbg_idx,
                size_t lds, void*) {
    unsigned char* b = emu_lds(lds);
    for (int i = 0; i < nslab; ++i) body_hist(vols, slabs, step, rp, dbg_idx, i, b);
    return 0;
}

int launch_wiener(const VolDesc* vols, int nvol, RunParams rp, size_t lds, void*) {
    unsigned char* b = emu_lds(lds);
    for (int i = 0; i < nvol; ++i) body_wiener(vols, rp, i, b);
    return 0;
}

int launch_fit(const VolDesc* vols, const Slab* slabs, int nslab, int level, int it, int mode, RunParams rp,
               size_t lds, int, int, void*) {
    if (rp.prec) return 1;
    unsigned char* b = emu_lds(lds);
    for (int i = 0; i < nslab; ++i) body_fit(vols, slabs, level, it, mode, rp, i, b);
    return 0;
}

int launch_lattice(const VolDesc* vols, int nvol, int cx_max, int level, int mode, RunParams rp, void*) {
    unsigned char* b = emu_lds(lds_lattice());
    for (int v = 0; v < nvol; ++v)
        for (int cx = 0; cx < cx_max; ++cx) body_lattice(vols, level, mode, rp, cx, v, b);
    return 0;
}

int launch_update(const VolDesc* vols, int nvol, int chunks, RunParams rp, void*) {
    for (int v = 0; v < nvol; ++v)
        for (int c = 0; c < chunks; ++c) body_update(vols, rp, c, v);
    return 0;
}

int launch_eval(const VolDesc* vols, const Slab* slabs, int nslab, RunParams rp, size_t lds, int, int, void*) {
    if (rp.prec) return 1;
    unsigned char* b = emu_lds(lds);
    for (int i = 0; i < nslab; ++i) body_eval(vols, slabs, rp, i, b);
    return 0;
}

int launch_recon(const VolDesc* vols, const Slab* slabs, int nslab, int level, RunParams rp, size_t lds, void*) {
    unsigned char* b = emu_lds(lds);
    for (int i = 0; i < nslab; ++i) body_recon(vols, slabs, level, rp, i, b);
    return 0;
}

// Workgroups of the persistent kernel wait for each other at team
// barriers, so each runs on a host thread of its own.
int persist_capacity(size_t, int, int* per_cu, int* cus) {
    *per_cu = 4;
    *cus = 16;
    return 0;
}

int launch_persist(const VolDesc* vols, const Slab* slabs, const PWork* work, int nwg, RunParams rp, size_t lds,
                   void*) {
    std::vector<std::thread> th;
    th.reserve(nwg);
    for (int b = 0; b < nwg; ++b)
        th.emplace_back([=] {
            std::vector<unsigned char> buf(lds + 64, 0);
            body_persist<0>(vols, slabs, work, rp, b, buf.data());
        });
    for (std::thread& x : th) x.join();
    return 0;
}

int launch_flat(const VolDesc* vols, const Slab* slabs, int nslab, int op, const float* a, const float* b,
                float* out, RunParams rp, void*) {
    unsigned char* buf = emu_lds(lds_flat(rp.nbins));
    for (int i = 0; i < nslab; ++i) body_flat(vols, slabs, op, a, b, out, rp, i, buf);
    return 0;
}

int launch_mask_u8(const float* in, uint8_t* out, uint64_t n, void*) {
    for (uint64_t i = 0; i < n; ++i) out[i] = in[i] != 0.0f ? 1 : 0;
    return 0;
}

}  // namespace n4

// Test hook of the emulation build only: the FP32 remap of the correction
// loop (fit_input, NUM mode: residual x - E[u|x] over the masked voxels,
// 0 elsewhere) on n values, with the FP64 E table narrowed to FP32 as the
// fit kernels stage it in LDS.  The CPU tests compare it with the oracle's
// FP64 residual at bin edges, duplicated maxima and values below bmin.
extern "C" int n4emu_remap_residual(const float* x, const uint8_t* m, int64_t n, const double* E, int nb,
                                    double bmin, double slope, int degenerate, float* out) {
    if (nb < 2 || nb > n4::kMaxBins) return 1;
    float Ef[n4::kMaxBins];
    for (int i = 0; i < nb; ++i) Ef[i] = (float)E[i];
    const n4::Remap rm = n4::make_remap(Ef, bmin, slope, nb);
    for (int64_t i = 0; i < n; ++i) out[i] = n4::fit_input(n4::kFitNum, m[i], x[i], degenerate != 0, rm);
    return 0;
}

namespace n4 {

#endif

}  // namespace n4
