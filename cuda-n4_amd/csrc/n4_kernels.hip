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
//   set_background + reduce_min/max                   -> fused into prep / eval tails
//   histogramGPU_kernel (cudaN4.cu:153-220)            -> body_hist
//   host FFTW Wiener block (cudaN4.cu:733-820)         -> wiener_tail (LDS FP64 FFT)
//   histogramMapping + subtract + accumulate_WC2phic_and_WC2 + divide + sum_inplace
//                                                      -> body_fit (+ fit_tail)
//   evaluate_bspline_kernel_3D + subtract_and_exp + CV + subtract
//                                                      -> body_eval (+ eval_tail)
//   upsample_lattice_kernel_3D (cudaN4.cu:257-339)     -> body_level
//   evaluate at full res + exp_and_divide + exp        -> body_recon
#include "n4_kernels.h"

#ifdef N4_EMULATE
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#define __host__
#define __device__
#else
#include <hip/hip_runtime.h>
#endif

#include <math.h>

#if defined(__HIP_DEVICE_COMPILE__)
#define N4_DEVICE_PASS 1
#endif

#ifdef N4_DEVICE_PASS
#define BLOCK_FOR(t) { const int t = (int)threadIdx.x; (void)t;
#define BLOCK_END }
#define BLOCK_SYNC() __syncthreads()
#else
#define BLOCK_FOR(t) for (int t = 0; t < kBlock; ++t) {
#define BLOCK_END }
#define BLOCK_SYNC() do { } while (0)
#endif

#define HD __host__ __device__ inline

namespace n4 {

// ---------------------------------------------------------------- helpers --

struct D2 {
    double x, y;
};

HD D2 cmul(D2 a, D2 b) { return D2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

HD void lds_add_u64(unsigned long long* p, unsigned long long v) {
#ifdef N4_DEVICE_PASS
    atomicAdd(p, v);
#else
    *p += v;
#endif
}

HD void fence_release() {
#ifdef N4_DEVICE_PASS
    __threadfence();
#endif
}

HD void fence_acquire() {
#ifdef N4_DEVICE_PASS
    __threadfence();
#endif
}

HD uint32_t counter_arrive(uint32_t* c) {
#ifdef N4_DEVICE_PASS
    return atomicAdd(c, 1u);
#else
    return (*c)++;
#endif
}

HD void counter_reset(uint32_t* c) {
#ifdef N4_DEVICE_PASS
    atomicExch(c, 0u);
#else
    *c = 0;
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

// Bump allocator over the block's dynamic LDS; with base == nullptr it only
// measures (host-side sizing uses the same code as the device carve).
struct Carve {
    unsigned char* base;
    size_t off;
    template <class T>
    HD T* take(size_t n) {
        off = align16(off);
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += n * sizeof(T);
        return p;
    }
};

HD bool mask_at(const VolDesc& V, size_t i) {
    if (V.mask_kind == 0) return reinterpret_cast<const float*>(V.mask)[i] != 0.0f;
    if (V.mask_kind == 1) return reinterpret_cast<const uint8_t*>(V.mask)[i] != 0;
    return true;
}

// Level-tagged done flag in pinned host memory (level+1 when the volume has
// finished the level, 0 otherwise), read by the host's early-exit poll.
HD void publish_done(const VolDesc& V, int level, int done) {
    if (!V.done_flag) return;
#ifdef N4_DEVICE_PASS
    __hip_atomic_store(V.done_flag, done ? level + 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
#else
    *V.done_flag = done ? level + 1 : 0;
#endif
}

// Histogram range for the next sharpening (N4.py:14-15).
HD void set_range(VolState* st, double mn, double mx, uint64_t n, int nbins) {
    st->bin_min = mn;
    st->bin_max = mx;
    double slope = (mx - mn) / (double)(nbins - 1);
    st->slope = slope;
    st->degenerate = (n == 0 || !(slope > 0.0)) ? 1 : 0;
}

HD double red_apply(int op, double a, double b) {
    return op == 0 ? a + b : (op == 1 ? fmin(a, b) : fmax(a, b));
}

// Tree reduction of four kBlock-long LDS arrays (op: 0 sum, 1 min, 2 max);
// results in red[0], red[kBlock], red[2*kBlock], red[3*kBlock].
HD void tree_reduce4(double* red, int o0, int o1, int o2, int o3) {
    BLOCK_SYNC();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        BLOCK_FOR(t)
        if (t < s) {
            red[t] = red_apply(o0, red[t], red[t + s]);
            red[kBlock + t] = red_apply(o1, red[kBlock + t], red[kBlock + t + s]);
            red[2 * kBlock + t] = red_apply(o2, red[2 * kBlock + t], red[2 * kBlock + t + s]);
            red[3 * kBlock + t] = red_apply(o3, red[3 * kBlock + t], red[3 * kBlock + t + s]);
        }
        BLOCK_END
        BLOCK_SYNC();
    }
}

// Counts this slab as finished; true in the last slab of the volume to
// finish (all partials of the volume are then visible to it).
HD bool slab_arrive(uint32_t* counter, uint32_t total, int* flag) {
    BLOCK_FOR(t)
    fence_release();
    BLOCK_END
    BLOCK_SYNC();
    BLOCK_FOR(t)
    if (t == 0) {
        uint32_t prev = counter_arrive(counter);
        int last = (prev + 1 == total) ? 1 : 0;
        if (last) counter_reset(counter);
        *flag = last;
    }
    BLOCK_END
    BLOCK_SYNC();
    bool last = (*flag != 0);
    if (last) {
        BLOCK_FOR(t)
        fence_acquire();
        BLOCK_END
    }
    return last;
}

// ------------------------------------------------------------------ prep --
// Threshold + log + shrink (N4.py:280-287; cudaN4.cu:131-147, 639-652) for
// one working slab: each thread owns working voxels (y, z) and reads their
// s*s*s full-resolution blocks.  Writes imlog, logU (= imlog), wmask; the
// last slab reduces counts and the initial logU min/max.
HD void body_prep(const VolDesc* vols, const Slab* slabs, RunParams rp, int bid, unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    Carve cv{smem, 0};
    double* red = cv.take<double>(4 * kBlock);
    int* flag = cv.take<int>(1);
    const int s = V.s;
    const int F0 = V.F[0], F1 = V.F[1], F2 = V.F[2];
    const int W1 = V.W[1], W2 = V.W[2];
    const int x = sl.x;
    BLOCK_FOR(t)
    double nfull = 0.0, nwork = 0.0;
    double mn = INFINITY, mx = -INFINITY;
    const int i0 = x * s, i1 = (i0 + s < F0) ? i0 + s : F0;
    for (int v = t; v < W1 * W2; v += kBlock) {
        const int J = v / W2, K = v - (v / W2) * W2;
        const int j0 = J * s, j1 = (j0 + s < F1) ? j0 + s : F1;
        const int k0 = K * s, k1 = (k0 + s < F2) ? k0 + s : F2;
        int nm = 0, nb = 0;
        float acc = 0.0f;
        for (int i = i0; i < i1; ++i)
            for (int j = j0; j < j1; ++j) {
                const size_t row = ((size_t)i * F1 + j) * F2;
                for (int k = k0; k < k1; ++k) {
                    ++nb;
                    const float xv = V.im[row + k];
                    if (mask_at(V, row + k) && xv > rp.low_f && isfinite(xv)) {
                        ++nm;
                        acc += logf(xv);
                    }
                }
            }
        const size_t w = ((size_t)x * W1 + J) * W2 + K;
        const bool fg = nm > 0 && 2 * nm >= nb;
        const float lg = nm > 0 ? acc / (float)nm : 0.0f;
        V.imlog[w] = lg;
        V.logU[w] = lg;
        V.wmask[w] = fg ? 1 : 0;
        nfull += nm;
        if (fg) {
            nwork += 1.0;
            mn = fmin(mn, (double)lg);
            mx = fmax(mx, (double)lg);
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
        double* sp = V.spart + (size_t)x * 4;
        sp[0] = red[0];
        sp[1] = red[kBlock];
        sp[2] = red[2 * kBlock];
        sp[3] = red[3 * kBlock];
    }
    BLOCK_END
    if (!slab_arrive(&V.st->cnt[0], (uint32_t)V.W[0], flag)) return;
    BLOCK_FOR(t)
    if (t == 0) {
        double nf = 0.0, nw = 0.0, mn = INFINITY, mx = -INFINITY;
        for (int xx = 0; xx < V.W[0]; ++xx) {
            const double* sp = V.spart + (size_t)xx * 4;
            nf += sp[0];
            nw += sp[1];
            mn = fmin(mn, sp[2]);
            mx = fmax(mx, sp[3]);
        }
        VolState* st = V.st;
        st->n_mask_full = (uint64_t)nf;
        st->n_mask_work = (uint64_t)nw;
        set_range(st, mn, mx, (uint64_t)nw, rp.nbins);
        st->level = 0;
        st->iter = 0;
        st->done = 0;
        st->cv = 1e10;
    }
    BLOCK_END
}

// ------------------------------------------------------------ level setup --
// One block per volume.  Level 0: zero lattice.  Level l > 0: dyadic
// refinement of the previous level's FP64 lattice (N4.py:132-165,
// cudaN4.cu:257-339), per-axis clip bounds.
HD void body_level(const VolDesc* vols, int level, int bid) {
    const VolDesc& V = vols[bid];
    const LevelDesc& LD = V.lev[level];
    double* dst = V.lat[level & 1];
    if (level == 0) {
        BLOCK_FOR(t)
        for (int l = t; l < LD.L; l += kBlock) {
            dst[l] = 0.0;
            V.latf[l] = 0.0f;
            V.lresf[l] = 0.0f;
        }
        BLOCK_END
    } else {
        const LevelDesc& PD = V.lev[level - 1];
        const double* src = V.lat[(level - 1) & 1];
        const double bw[2][3] = {{1.0 / 8, 6.0 / 8, 1.0 / 8}, {0.0, 1.0 / 2, 1.0 / 2}};
        BLOCK_FOR(t)
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
                        const double wgt = bw[odd[0]][a] * bw[odd[1]][b] * bw[odd[2]][e];
                        acc += wgt * src[((size_t)(X + 1) * PD.C[1] + (Y + 1)) * PD.C[2] + (Z + 1)];
                    }
                }
            }
            dst[l] = acc;
            V.latf[l] = (float)acc;
            V.lresf[l] = 0.0f;
        }
        BLOCK_END
    }
    BLOCK_FOR(t)
    if (t == 0) {
        VolState* st = V.st;
        st->level = level;
        st->iter = 0;
        st->cv = 1e10;
        st->done = (V.max_iter[level] <= 0) ? 1 : 0;
        publish_done(V, level, st->done);
        st->iters[level] = 0;
        st->lastcv[level] = 1e10;
    }
    BLOCK_END
}

// ----------------------------------------------------------- Wiener tail --
// LDS FP64 Stockham radix-2 FFT over P points; returns the buffer holding
// the result (the other one is free).
HD D2* lds_fft(D2* a, D2* b, const D2* tw, int P, bool inverse) {
    const int half = P >> 1;
    for (int Ns = 1; Ns < P; Ns <<= 1) {
        BLOCK_FOR(t)
        for (int j = t; j < half; j += kBlock) {
            const D2 v0 = a[j];
            const D2 v1 = a[j + half];
            const int k = j & (Ns - 1);
            D2 w = tw[k * (half / Ns)];
            if (inverse) w.y = -w.y;
            const D2 tt = cmul(v1, w);
            const int o = ((j - k) << 1) + k;
            b[o] = D2{v0.x + tt.x, v0.y + tt.y};
            b[o + Ns] = D2{v0.x - tt.x, v0.y - tt.y};
        }
        BLOCK_END
        BLOCK_SYNC();
        D2* tmp = a;
        a = b;
        b = tmp;
    }
    return a;
}

// Wiener deconvolution of the fractional histogram and the E[u|v] table,
// N4.py:31-101 (cudaN4.cu:733-820 did this on the host with FFTW).  Two real
// transforms share one complex FFT (V in the real part, F in the imaginary).
HD void wiener_tail(const VolDesc& V, const RunParams& rp, unsigned char* smem) {
    const int P = rp.P, half = P / 2, off = rp.offset, nb = rp.nbins;
    Carve cv{smem, 0};
    D2* z = cv.take<D2>(P);
    D2* y = cv.take<D2>(P);
    D2* ff = cv.take<D2>(P);
    D2* tw = cv.take<D2>(half);
    const double bmin = V.st->bin_min, slope = V.st->slope;
    const double scaledFWHM = rp.fwhm / slope;
    const double expF = 4.0 * log(2.0) / (scaledFWHM * scaledFWHM);
    const double scale = 2.0 * sqrt(log(2.0) / M_PI) / scaledFWHM;
    const int nslab = V.W[0];
    BLOCK_FOR(t)
    for (int j = t; j < P; j += kBlock) {
        const int jj = (j <= half) ? j : P - j;
        const double Fj = (j == 0) ? scale : scale * exp((double)(-(int64_t)jj * jj) * expF);
        double Vj = 0.0;
        if (j >= off && j < off + nb) {
            unsigned long long acc = 0;
            for (int xx = 0; xx < nslab; ++xx) acc += V.hpart[(size_t)xx * nb + (j - off)];
            Vj = (double)acc * 0x1p-32;
        }
        z[j] = D2{Vj, Fj};
    }
    for (int k = t; k < half; k += kBlock) {
        double sn, cs;
        sincospi_d(-2.0 * (double)k / (double)P, &sn, &cs);
        tw[k] = D2{cs, sn};
    }
    BLOCK_END
    BLOCK_SYNC();
    D2* Z = lds_fft(z, y, tw, P, false);
    D2* U = (Z == z) ? y : z;
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
    D2* Ur = lds_fft(U, Z, tw, P, true);
    D2* Wb = (Ur == U) ? Z : U;
    const double invP = 1.0 / (double)P;
    BLOCK_FOR(t)
    for (int j = t; j < P; j += kBlock) {
        double u = Ur[j].x * invP;
        if (u < 0.0) u = 0.0;
        Wb[j] = D2{(bmin + (double)(j - off) * slope) * u, u};
    }
    BLOCK_END
    BLOCK_SYNC();
    D2* Wf = lds_fft(Wb, Ur, tw, P, false);
    D2* Wo = (Wf == Wb) ? Ur : Wb;
    BLOCK_FOR(t)
    for (int k = t; k < P; k += kBlock) Wf[k] = cmul(Wf[k], ff[k]);
    BLOCK_END
    BLOCK_SYNC();
    D2* R = lds_fft(Wf, Wo, tw, P, true);
    BLOCK_FOR(t)
    for (int i = t; i < nb; i += kBlock) {
        const double nu = R[off + i].x * invP;
        const double de = R[off + i].y * invP;
        V.etab[i] = (de != 0.0) ? nu / de : 0.0;
    }
    BLOCK_END
}

// ------------------------------------------------------------- histogram --
// Fractional (linear-split) histogram of logU over the mask, N4.py:19-29
// (cudaN4.cu:153-220).  Bin index in FP64 exactly as N4.py computes it;
// contributions accumulate as 2^-32 fixed point in LDS-private u64 bins
// (integer adds: order-independent, so bitwise deterministic).  The last
// slab of a volume runs the Wiener tail.
HD void body_hist(const VolDesc* vols, const Slab* slabs, RunParams rp, int32_t* dbg_idx, int bid,
                  unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    if (V.st->done || V.st->degenerate) return;
    const int nb = rp.nbins;
    Carve cv{smem, 0};
    unsigned long long* h = cv.take<unsigned long long>(nb);
    int* flag = cv.take<int>(1);
    const double bmin = V.st->bin_min, slope = V.st->slope;
    const int YZ = V.W[1] * V.W[2];
    const size_t base = (size_t)sl.x * YZ;
    BLOCK_FOR(t)
    for (int i = t; i < nb; i += kBlock) h[i] = 0ull;
    BLOCK_END
    BLOCK_SYNC();
    BLOCK_FOR(t)
    for (int v = t; v < YZ; v += kBlock) {
        const size_t w = base + v;
        if ((int64_t)w >= V.nvox_w) break;
        if (!V.wmask[w]) {
            if (dbg_idx) dbg_idx[w] = -1;
            continue;
        }
        const double c = ((double)V.logU[w] - bmin) / slope;
        const double fl = floor(c);
        const int idx = (fl >= 0.0 && fl < (double)nb) ? (int)fl : -1;
        if (dbg_idx) dbg_idx[w] = idx;
        if (idx < 0) continue;
        const double off = c - fl;
        if (off == 0.0) {
            lds_add_u64(&h[idx], 1ull << 32);
        } else if (idx < nb - 1) {
            const unsigned long long q = (unsigned long long)rint(off * 4294967296.0);
            lds_add_u64(&h[idx], (1ull << 32) - q);
            lds_add_u64(&h[idx + 1], q);
        }
    }
    BLOCK_END
    BLOCK_SYNC();
    BLOCK_FOR(t)
    for (int i = t; i < nb; i += kBlock) V.hpart[(size_t)sl.x * nb + i] = h[i];
    BLOCK_END
    if (!slab_arrive(&V.st->cnt[1], (uint32_t)V.W[0], flag)) return;
    BLOCK_SYNC();
    wiener_tail(V, rp, smem);
}

// ------------------------------------------------------------------- fit --
// Remap (N4.py:103-115) + residual (N4.py:313) + separable B-spline fit
// (N4.py:186-237 rewritten as mode-n contractions, exact because the
// normaliser sum(w^2) factorises per axis):
//   T1[y,cz] = sum_z g[y,z] Pz[z,cz]      (LDS, per row chunk)
//   T2[cy,cz] += sum_y Py[y,cy] T1[y,cz]  (LDS, accumulated over the slab)
//   part[x] = T2                          (global, per slab)
// tail: num[cx,cy,cz] = sum_x Px[x,cx] part[x][cy,cz] in FP64, then
//   NUM  : lres = num/(den+eps); lat += lres  (N4.py:236, 317)
//   DEN  : den = num (g = mask, Q = w^2 weights; once per level)
//   FIELD: lres = num/(den+eps) of the logU buffer itself (stage API)
struct FitLds {
    double* E;
    float* zw;
    int* zb;
    int* zs;
    int* ys;
    float* g;
    float* T1;
    float* T2;
    float* yw;
    int* yb;
    int* flag;
};

HD size_t carve_fit(unsigned char* base, int nbins, int Z, int R, int Cy, int Cz, FitLds* o) {
    Carve c{base, 0};
    FitLds L;
    L.E = c.take<double>(nbins);
    L.zw = c.take<float>((size_t)4 * Z);
    L.zb = c.take<int>(Z);
    L.zs = c.take<int>((size_t)2 * Cz);
    L.ys = c.take<int>((size_t)2 * Cy);
    L.g = c.take<float>((size_t)R * Z);
    L.T1 = c.take<float>((size_t)R * Cz);
    L.T2 = c.take<float>((size_t)Cy * Cz);
    L.yw = c.take<float>((size_t)4 * R);
    L.yb = c.take<int>(R);
    L.flag = c.take<int>(1);
    if (o) *o = L;
    return align16(c.off);
}

HD int rows_per_chunk(int Y, int Z) {
    int r = kChunkVox / (Z > 0 ? Z : 1);
    if (r < 1) r = 1;
    return r < Y ? r : Y;
}

HD void body_fit(const VolDesc* vols, const Slab* slabs, int level, int mode, RunParams rp, int bid,
                 unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    VolState* st = V.st;
    if (mode == kFitNum && st->done) return;
    const LevelDesc& LD = V.lev[level];
    const int Y = V.W[1], Z = V.W[2];
    const int Cy = LD.C[1], Cz = LD.C[2];
    const int R = rows_per_chunk(Y, Z);
    const int nb = rp.nbins;
    FitLds S;
    carve_fit(smem, nb, Z, R, Cy, Cz, &S);
    const bool num = (mode == kFitNum);
    const bool degenerate = num && st->degenerate;
    const double bmin = st->bin_min, slope = st->slope;
    const AxisEntry* axz = LD.ax[2];
    const AxisEntry* axy = LD.ax[1];
    const int x = sl.x;
    BLOCK_FOR(t)
    if (num)
        for (int i = t; i < nb; i += kBlock) S.E[i] = V.etab[i];
    for (int z = t; z < Z; z += kBlock) {
        const AxisEntry& e = axz[z];
        for (int a = 0; a < 4; ++a) S.zw[4 * z + a] = (mode == kFitDen) ? e.q[a] : e.p[a];
        S.zb[z] = e.b;
    }
    for (int i = t; i < 2 * Cz; i += kBlock) S.zs[i] = LD.sup[2][i];
    for (int i = t; i < 2 * Cy; i += kBlock) S.ys[i] = LD.sup[1][i];
    for (int i = t; i < Cy * Cz; i += kBlock) S.T2[i] = 0.0f;
    BLOCK_END
    BLOCK_SYNC();
    for (int y0 = 0; y0 < Y; y0 += R) {
        const int ny = (Y - y0 < R) ? Y - y0 : R;
        BLOCK_FOR(t)
        for (int r = t; r < ny; r += kBlock) {
            const AxisEntry& e = axy[y0 + r];
            for (int a = 0; a < 4; ++a) S.yw[4 * r + a] = (mode == kFitDen) ? e.q[a] : e.p[a];
            S.yb[r] = e.b;
        }
        for (int v = t; v < ny * Z; v += kBlock) {
            const size_t w = ((size_t)x * Y + y0) * Z + v;
            float gv = 0.0f;
            if (V.wmask[w]) {
                if (mode == kFitDen) {
                    gv = 1.0f;
                } else if (mode == kFitField) {
                    gv = V.logU[w];
                } else if (!degenerate) {
                    const float xv = V.logU[w];
                    double c = ((double)xv - bmin) / slope;
                    c = fmin(fmax(c, 0.0), (double)nb);  // cudaN4.cu:242 clamp
                    const double fl = floor(c);
                    const int idx = (int)fl;
                    const double sh = (idx < nb - 1)
                                          ? S.E[idx] + (S.E[idx + 1] - S.E[idx]) * (c - fl)
                                          : S.E[nb - 1];
                    gv = (float)((double)xv - sh);
                }
            }
            S.g[v] = gv;
        }
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        for (int pr = t; pr < ny * Cz; pr += kBlock) {
            const int r = pr / Cz, c = pr - (pr / Cz) * Cz;
            const int lo = S.zs[2 * c], hi = S.zs[2 * c + 1];
            const float* gr = S.g + (size_t)r * Z;
            float acc = 0.0f;
            for (int z = lo; z < hi; ++z) acc += gr[z] * S.zw[4 * z + (c - S.zb[z])];
            S.T1[pr] = acc;
        }
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        for (int pr = t; pr < Cy * Cz; pr += kBlock) {
            const int cy = pr / Cz, c = pr - (pr / Cz) * Cz;
            int lo = S.ys[2 * cy], hi = S.ys[2 * cy + 1];
            lo = lo > y0 ? lo : y0;
            hi = hi < y0 + ny ? hi : y0 + ny;
            float acc = S.T2[pr];
            for (int yy = lo; yy < hi; ++yy) {
                const int r = yy - y0;
                acc += S.yw[4 * r + (cy - S.yb[r])] * S.T1[r * Cz + c];
            }
            S.T2[pr] = acc;
        }
        BLOCK_END
        BLOCK_SYNC();
    }
    BLOCK_FOR(t)
    for (int pr = t; pr < Cy * Cz; pr += kBlock) V.part[(size_t)x * V.part_stride + pr] = S.T2[pr];
    BLOCK_END
    if (!slab_arrive(&st->cnt[2], (uint32_t)V.W[0], S.flag)) return;
    // tail: mode-x contraction in FP64 and the lattice update
    const AxisEntry* axx = LD.ax[0];
    const int* supx = LD.sup[0];
    const int CyCz = Cy * Cz;
    double* lat = V.lat[level & 1];
    BLOCK_FOR(t)
    for (int l = t; l < LD.L; l += kBlock) {
        const int cx = l / CyCz, rest = l - (l / CyCz) * CyCz;
        double acc = 0.0;
        for (int xx = supx[2 * cx]; xx < supx[2 * cx + 1]; ++xx) {
            const AxisEntry& e = axx[xx];
            const float wx = (mode == kFitDen) ? e.q[cx - e.b] : e.p[cx - e.b];
            acc += (double)wx * (double)V.part[(size_t)xx * V.part_stride + rest];
        }
        if (mode == kFitDen) {
            V.den[l] = acc;
        } else {
            const double lres = acc / (V.den[l] + rp.eps);
            if (mode == kFitNum) {
                const double nl = lat[l] + lres;
                lat[l] = nl;
                V.latf[l] = (float)nl;
            }
            V.lresf[l] = (float)lres;
        }
    }
    BLOCK_END
}

// ------------------------------------------------------------------ eval --
// New field on the working grid (N4.py:318), convergence sums (N4.py:121-126)
// and logU update (N4.py:322), fused:
//   A[cy,cz]  = sum_a wx[a] L[bx+a,cy,cz]       (also for the increment Lres)
//   B[y,cz]   = sum_b wy[b] A[by+b,cz]
//   f(y,z)    = sum_c wz[c] B[y,bz+c];  d = same with Lres = new - old field
//   logU = imlog - f;  u = exp(old-new) - 1 = expm1(-d)
// Per slab: sum u, sum u^2 (FP64), min/max of logU; the last slab computes
// CV = std/mean (ddof 0), the stop flag and the next histogram range.
struct EvalLds {
    float* A;
    float* Ar;
    float* B;
    float* Br;
    float* zw;
    int* zb;
    double* red;
    int* flag;
};

HD size_t carve_eval(unsigned char* base, int Z, int R, int Cy, int Cz, EvalLds* o) {
    Carve c{base, 0};
    EvalLds L;
    L.A = c.take<float>((size_t)Cy * Cz);
    L.Ar = c.take<float>((size_t)Cy * Cz);
    L.B = c.take<float>((size_t)R * Cz);
    L.Br = c.take<float>((size_t)R * Cz);
    L.zw = c.take<float>((size_t)4 * Z);
    L.zb = c.take<int>(Z);
    L.red = c.take<double>((size_t)4 * kBlock);
    L.flag = c.take<int>(1);
    if (o) *o = L;
    return align16(c.off);
}

HD void body_eval(const VolDesc* vols, const Slab* slabs, int level, RunParams rp, int bid,
                  unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    VolState* st = V.st;
    if (st->done) return;
    const LevelDesc& LD = V.lev[level];
    const int Y = V.W[1], Z = V.W[2];
    const int Cy = LD.C[1], Cz = LD.C[2];
    const int CyCz = Cy * Cz;
    const int R = rows_per_chunk(Y, Z);
    EvalLds S;
    carve_eval(smem, Z, R, Cy, Cz, &S);
    const int x = sl.x;
    const AxisEntry ex = LD.ax[0][x];
    BLOCK_FOR(t)
    for (int pr = t; pr < CyCz; pr += kBlock) {
        float a = 0.0f, ar = 0.0f;
        for (int k = 0; k < 4; ++k) {
            const size_t li = (size_t)(ex.b + k) * CyCz + pr;
            a += ex.w[k] * V.latf[li];
            ar += ex.w[k] * V.lresf[li];
        }
        S.A[pr] = a;
        S.Ar[pr] = ar;
    }
    for (int z = t; z < Z; z += kBlock) {
        const AxisEntry& e = LD.ax[2][z];
        for (int k = 0; k < 4; ++k) S.zw[4 * z + k] = e.w[k];
        S.zb[z] = e.b;
    }
    S.red[t] = 0.0;
    S.red[kBlock + t] = 0.0;
    S.red[2 * kBlock + t] = INFINITY;
    S.red[3 * kBlock + t] = -INFINITY;
    BLOCK_END
    BLOCK_SYNC();
    for (int y0 = 0; y0 < Y; y0 += R) {
        const int ny = (Y - y0 < R) ? Y - y0 : R;
        BLOCK_FOR(t)
        for (int pr = t; pr < ny * Cz; pr += kBlock) {
            const int r = pr / Cz, c = pr - (pr / Cz) * Cz;
            const AxisEntry& e = LD.ax[1][y0 + r];
            float b = 0.0f, br = 0.0f;
            for (int k = 0; k < 4; ++k) {
                b += e.w[k] * S.A[(e.b + k) * Cz + c];
                br += e.w[k] * S.Ar[(e.b + k) * Cz + c];
            }
            S.B[pr] = b;
            S.Br[pr] = br;
        }
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        double s1 = S.red[t], s2 = S.red[kBlock + t];
        double mn = S.red[2 * kBlock + t], mx = S.red[3 * kBlock + t];
        for (int v = t; v < ny * Z; v += kBlock) {
            const size_t w = ((size_t)x * Y + y0) * Z + v;
            if (!V.wmask[w]) continue;
            const int r = v / Z, z = v - (v / Z) * Z;
            const float* wz = S.zw + 4 * z;
            const float* br = S.B + r * Cz + S.zb[z];
            const float* brr = S.Br + r * Cz + S.zb[z];
            const float f = wz[0] * br[0] + wz[1] * br[1] + wz[2] * br[2] + wz[3] * br[3];
            const float d = wz[0] * brr[0] + wz[1] * brr[1] + wz[2] * brr[2] + wz[3] * brr[3];
            const float lg = V.imlog[w] - f;
            V.logU[w] = lg;
            const double u = (double)expm1f(-d);
            s1 += u;
            s2 += u * u;
            mn = fmin(mn, (double)lg);
            mx = fmax(mx, (double)lg);
        }
        S.red[t] = s1;
        S.red[kBlock + t] = s2;
        S.red[2 * kBlock + t] = mn;
        S.red[3 * kBlock + t] = mx;
        BLOCK_END
        BLOCK_SYNC();
    }
    tree_reduce4(S.red, 0, 0, 1, 2);
    BLOCK_FOR(t)
    if (t == 0) {
        double* sp = V.spart + (size_t)x * 4;
        sp[0] = S.red[0];
        sp[1] = S.red[kBlock];
        sp[2] = S.red[2 * kBlock];
        sp[3] = S.red[3 * kBlock];
    }
    BLOCK_END
    if (!slab_arrive(&st->cnt[3], (uint32_t)V.W[0], S.flag)) return;
    BLOCK_FOR(t)
    if (t == 0) {
        double s1 = 0.0, s2 = 0.0, mn = INFINITY, mx = -INFINITY;
        for (int xx = 0; xx < V.W[0]; ++xx) {
            const double* sp = V.spart + (size_t)xx * 4;
            s1 += sp[0];
            s2 += sp[1];
            mn = fmin(mn, sp[2]);
            mx = fmax(mx, sp[3]);
        }
        const double n = (double)st->n_mask_work;
        double cvv = 0.0;
        if (n > 0.0) {
            const double mu = s1 / n;
            double var = s2 / n - mu * mu;
            if (var < 0.0) var = 0.0;
            cvv = sqrt(var) / (1.0 + mu);
        }
        st->iter += 1;
        st->cv = cvv;
        st->iters[level] = st->iter;
        st->lastcv[level] = cvv;
        // N4.py:301-302: continue while CV > threshold and iterations remain
        if ((!rp.fixed_iters && !(cvv > rp.thr)) || st->iter >= V.max_iter[level]) st->done = 1;
        publish_done(V, level, st->done);
        set_range(st, mn, mx, st->n_mask_work, rp.nbins);
    }
    BLOCK_END
}

// ----------------------------------------------------------------- recon --
// Final field at full resolution from the last lattice (N4.py:329-330,
// cudaN4.cu:1700-1729): separable evaluation, then normalised image
// (thresholded input / exp(field)), optional exp(field) and field.
struct ReconLds {
    float* A;
    float* B;
    float* zw;
    int* zb;
};

HD size_t carve_recon(unsigned char* base, int Z, int R, int Cy, int Cz, ReconLds* o) {
    Carve c{base, 0};
    ReconLds L;
    L.A = c.take<float>((size_t)Cy * Cz);
    L.B = c.take<float>((size_t)R * Cz);
    L.zw = c.take<float>((size_t)4 * Z);
    L.zb = c.take<int>(Z);
    if (o) *o = L;
    return align16(c.off);
}

HD void body_recon(const VolDesc* vols, const Slab* slabs, int level, RunParams rp, int bid,
                   unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    const LevelDesc& LD = V.lev[level];
    const int Y = V.F[1], Z = V.F[2];
    const int Cy = LD.C[1], Cz = LD.C[2];
    const int CyCz = Cy * Cz;
    const int R = rows_per_chunk(Y, Z);
    ReconLds S;
    carve_recon(smem, Z, R, Cy, Cz, &S);
    const int x = sl.x;
    const AxisEntry ex = V.fax[0][x];
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
    BLOCK_END
    BLOCK_SYNC();
    for (int y0 = 0; y0 < Y; y0 += R) {
        const int ny = (Y - y0 < R) ? Y - y0 : R;
        BLOCK_FOR(t)
        for (int pr = t; pr < ny * Cz; pr += kBlock) {
            const int r = pr / Cz, c = pr - (pr / Cz) * Cz;
            const AxisEntry& e = V.fax[1][y0 + r];
            float b = 0.0f;
            for (int k = 0; k < 4; ++k) b += e.w[k] * S.A[(e.b + k) * Cz + c];
            S.B[pr] = b;
        }
        BLOCK_END
        BLOCK_SYNC();
        BLOCK_FOR(t)
        for (int v = t; v < ny * Z; v += kBlock) {
            const int r = v / Z, z = v - (v / Z) * Z;
            const size_t fv = ((size_t)x * Y + y0) * Z + v;
            const float* wz = S.zw + 4 * z;
            const float* br = S.B + r * Cz + S.zb[z];
            const float f = wz[0] * br[0] + wz[1] * br[1] + wz[2] * br[2] + wz[3] * br[3];
            if (V.out_logbias) V.out_logbias[fv] = f;
            if (V.out_norm || V.out_bias) {
                const float ef = expf(f);
                if (V.out_bias) V.out_bias[fv] = ef;
                if (V.out_norm) {
                    const float xin = V.im[fv];
                    const float xo = (xin <= rp.low_f) ? 0.0f : xin;  // N4.py:285
                    V.out_norm[fv] = xo / ef;
                }
            }
        }
        BLOCK_END
        BLOCK_SYNC();
    }
}

// ------------------------------------------------------------------ flat --
// 1-D helpers for the stage entry points.  Slab x covers voxels
// [x*kChunkVox, min((x+1)*kChunkVox, n)) of V.logU / V.wmask.
HD void body_flat(const VolDesc* vols, const Slab* slabs, int op, const float* a, const float* b,
                  float* out, RunParams rp, int bid, unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    Carve cv{smem, 0};
    double* red = cv.take<double>(4 * kBlock);
    double* E = cv.take<double>(kMaxBins);
    int* flag = cv.take<int>(1);
    const int64_t n = V.nvox_w;
    const int64_t v0 = (int64_t)sl.x * kChunkVox;
    const int nb = rp.nbins;
    const double bmin = V.st->bin_min, slope = V.st->slope;
    if (op == kFlatRemap) {
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
    // slots: minmax (count, -, min, max); sum (sum, count, min, max);
    //        cv (count, sum u, sum u^2, -)   with u = exp(a-b) - 1
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
        double* sp = V.spart + (size_t)sl.x * 4;
        sp[0] = red[0];
        sp[1] = red[kBlock];
        sp[2] = red[2 * kBlock];
        sp[3] = red[3 * kBlock];
    }
    BLOCK_END
    const uint32_t nslab = (uint32_t)((n + kChunkVox - 1) / kChunkVox);
    if (!slab_arrive(&V.st->cnt[0], nslab, flag)) return;
    BLOCK_FOR(t)
    if (t == 0) {
        double a0 = 0.0, a1 = 0.0, a2 = (op == kFlatCV) ? 0.0 : INFINITY, a3 = -INFINITY;
        for (uint32_t xx = 0; xx < nslab; ++xx) {
            const double* sp = V.spart + (size_t)xx * 4;
            a0 += sp[0];
            a1 += sp[1];
            a2 = (op == kFlatCV) ? a2 + sp[2] : fmin(a2, sp[2]);
            a3 = fmax(a3, sp[3]);
        }
        VolState* st = V.st;
        if (op == kFlatMinMax) {
            st->n_mask_work = (uint64_t)a0;
            set_range(st, a2, a3, (uint64_t)a0, nb);
        } else if (op == kFlatSum) {
            st->cv = a0;
            st->n_mask_work = (uint64_t)a1;
            st->bin_min = a2;
            st->bin_max = a3;
        } else {
            double cvv = 0.0;
            if (a0 > 0.0) {
                const double mu = a1 / a0;
                double var = a2 / a0 - mu * mu;
                if (var < 0.0) var = 0.0;
                cvv = sqrt(var) / (1.0 + mu);
            }
            st->cv = cvv;
            st->n_mask_work = (uint64_t)a0;
        }
    }
    BLOCK_END
}

// ---------------------------------------------------------------- sizing --

size_t lds_prep() { return align16(4 * kBlock * sizeof(double) + 16); }

size_t lds_hist(int nbins, int P) {
    const size_t h = align16(nbins * sizeof(unsigned long long)) + 16;
    const size_t w = (3 * (size_t)P + P / 2) * sizeof(D2) + 64;
    return align16(h > w ? h : w);
}

size_t lds_fit(int nbins, const VolDesc& v, int level) {
    const LevelDesc& L = v.lev[level];
    return carve_fit(nullptr, nbins, v.W[2], rows_per_chunk(v.W[1], v.W[2]), L.C[1], L.C[2], nullptr);
}

size_t lds_eval(const VolDesc& v, int level) {
    const LevelDesc& L = v.lev[level];
    return carve_eval(nullptr, v.W[2], rows_per_chunk(v.W[1], v.W[2]), L.C[1], L.C[2], nullptr);
}

size_t lds_recon(const VolDesc& v) {
    const LevelDesc& L = v.lev[v.nlevels - 1];
    return carve_recon(nullptr, v.F[2], rows_per_chunk(v.F[1], v.F[2]), L.C[1], L.C[2], nullptr);
}

size_t lds_flat() { return align16((4 * kBlock + kMaxBins) * sizeof(double) + 16); }

// --------------------------------------------------------------- launches --
#ifndef N4_EMULATE

__global__ __launch_bounds__(kBlock) void k_prep(const VolDesc* vols, const Slab* slabs, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem);
    body_prep(vols, slabs, rp, blockIdx.x, reinterpret_cast<unsigned char*>(smem));
}

__global__ __launch_bounds__(kBlock) void k_level(const VolDesc* vols, int level) {
    body_level(vols, level, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_hist(const VolDesc* vols, const Slab* slabs, RunParams rp,
                                                 int32_t* dbg_idx) {
    HIP_DYNAMIC_SHARED(double, smem);
    body_hist(vols, slabs, rp, dbg_idx, blockIdx.x, reinterpret_cast<unsigned char*>(smem));
}

__global__ __launch_bounds__(kBlock) void k_fit(const VolDesc* vols, const Slab* slabs, int level, int mode,
                                                RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem);
    body_fit(vols, slabs, level, mode, rp, blockIdx.x, reinterpret_cast<unsigned char*>(smem));
}

__global__ __launch_bounds__(kBlock) void k_eval(const VolDesc* vols, const Slab* slabs, int level,
                                                 RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem);
    body_eval(vols, slabs, level, rp, blockIdx.x, reinterpret_cast<unsigned char*>(smem));
}

__global__ __launch_bounds__(kBlock) void k_recon(const VolDesc* vols, const Slab* slabs, int level,
                                                  RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem);
    body_recon(vols, slabs, level, rp, blockIdx.x, reinterpret_cast<unsigned char*>(smem));
}

__global__ __launch_bounds__(kBlock) void k_flat(const VolDesc* vols, const Slab* slabs, int op,
                                                 const float* a, const float* b, float* out, RunParams rp) {
    HIP_DYNAMIC_SHARED(double, smem);
    body_flat(vols, slabs, op, a, b, out, rp, blockIdx.x, reinterpret_cast<unsigned char*>(smem));
}

static int last_launch_status() { return (int)hipGetLastError(); }

int launch_prep(const VolDesc* vols, const Slab* slabs, int nslab, RunParams rp, void* stream) {
    if (nslab <= 0) return 0;
    hipLaunchKernelGGL(k_prep, dim3(nslab), dim3(kBlock), lds_prep(), (hipStream_t)stream, vols, slabs, rp);
    return last_launch_status();
}

int launch_level(const VolDesc* vols, int nvol, int level, RunParams, void* stream) {
    if (nvol <= 0) return 0;
    hipLaunchKernelGGL(k_level, dim3(nvol), dim3(kBlock), 0, (hipStream_t)stream, vols, level);
    return last_launch_status();
}

int launch_hist(const VolDesc* vols, const Slab* slabs, int nslab, RunParams rp, int32_t* dbg_idx, size_t lds,
                void* stream) {
    if (nslab <= 0) return 0;
    hipLaunchKernelGGL(k_hist, dim3(nslab), dim3(kBlock), lds, (hipStream_t)stream, vols, slabs, rp, dbg_idx);
    return last_launch_status();
}

int launch_fit(const VolDesc* vols, const Slab* slabs, int nslab, int level, int mode, RunParams rp, size_t lds,
               void* stream) {
    if (nslab <= 0) return 0;
    hipLaunchKernelGGL(k_fit, dim3(nslab), dim3(kBlock), lds, (hipStream_t)stream, vols, slabs, level, mode, rp);
    return last_launch_status();
}

int launch_eval(const VolDesc* vols, const Slab* slabs, int nslab, int level, RunParams rp, size_t lds,
                void* stream) {
    if (nslab <= 0) return 0;
    hipLaunchKernelGGL(k_eval, dim3(nslab), dim3(kBlock), lds, (hipStream_t)stream, vols, slabs, level, rp);
    return last_launch_status();
}

int launch_recon(const VolDesc* vols, const Slab* slabs, int nslab, int level, RunParams rp, size_t lds,
                 void* stream) {
    if (nslab <= 0) return 0;
    hipLaunchKernelGGL(k_recon, dim3(nslab), dim3(kBlock), lds, (hipStream_t)stream, vols, slabs, level, rp);
    return last_launch_status();
}

int launch_flat(const VolDesc* vols, const Slab* slabs, int nslab, int op, const float* a, const float* b,
                float* out, RunParams rp, void* stream) {
    if (nslab <= 0) return 0;
    hipLaunchKernelGGL(k_flat, dim3(nslab), dim3(kBlock), lds_flat(), (hipStream_t)stream, vols, slabs, op, a,
                       b, out, rp);
    return last_launch_status();
}

#else  // N4_EMULATE: host loops over blocks, used only by CPU unit tests

static std::vector<unsigned char>& emu_lds(size_t n) {
    static std::vector<unsigned char> buf;
    if (buf.size() < n + 64) buf.assign(n + 64, 0);
    return buf;
}

int launch_prep(const VolDesc* vols, const Slab* slabs, int nslab, RunParams rp, void*) {
    auto& b = emu_lds(lds_prep());
    for (int i = 0; i < nslab; ++i) body_prep(vols, slabs, rp, i, b.data());
    return 0;
}

int launch_level(const VolDesc* vols, int nvol, int level, RunParams, void*) {
    for (int i = 0; i < nvol; ++i) body_level(vols, level, i);
    return 0;
}

int launch_hist(const VolDesc* vols, const Slab* slabs, int nslab, RunParams rp, int32_t* dbg_idx, size_t lds,
                void*) {
    auto& b = emu_lds(lds);
    for (int i = 0; i < nslab; ++i) body_hist(vols, slabs, rp, dbg_idx, i, b.data());
    return 0;
}

int launch_fit(const VolDesc* vols, const Slab* slabs, int nslab, int level, int mode, RunParams rp, size_t lds,
               void*) {
    auto& b = emu_lds(lds);
    for (int i = 0; i < nslab; ++i) body_fit(vols, slabs, level, mode, rp, i, b.data());
    return 0;
}

int launch_eval(const VolDesc* vols, const Slab* slabs, int nslab, int level, RunParams rp, size_t lds, void*) {
    auto& b = emu_lds(lds);
    for (int i = 0; i < nslab; ++i) body_eval(vols, slabs, level, rp, i, b.data());
    return 0;
}

int launch_recon(const VolDesc* vols, const Slab* slabs, int nslab, int level, RunParams rp, size_t lds, void*) {
    auto& b = emu_lds(lds);
    for (int i = 0; i < nslab; ++i) body_recon(vols, slabs, level, rp, i, b.data());
    return 0;
}

int launch_flat(const VolDesc* vols, const Slab* slabs, int nslab, int op, const float* a, const float* b,
                float* out, RunParams rp, void*) {
    auto& buf = emu_lds(lds_flat());
    for (int i = 0; i < nslab; ++i) body_flat(vols, slabs, op, a, b, out, rp, i, buf.data());
    return 0;
}

#endif

}  // namespace n4
