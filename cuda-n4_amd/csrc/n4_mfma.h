// n4_mfma.h -- BF16 and split-BF16 MFMA forms of the separable B-spline
// contractions (device build only; included by n4_kernels.hip).
//
// The per-voxel B-spline normaliser factorises per axis, so the fit
// numerator/denominator and the field evaluation are chains of mode-n
// products with banded per-axis factor matrices (weights w, w^3/sum w^2,
// w^2 of the axis tables).  Here each chain runs on gfx950 matrix cores:
//
//   fit   T1[y,cz]  = sum_z g[y,z] Pz[z,cz]          v_mfma_f32_16x16x32_bf16
//         T2[cy,cz] = sum_y Py[y,cy] T1[y,cz]        v_mfma_f32_16x16x16_bf16
//   eval  Bt[c,y]   = sum_cy A[cy,c] Wy[y,cy]        v_mfma_f32_16x16x16_bf16
//         d[z,y]    = sum_c Wz[z,c] Bt[c,y]          v_mfma_f32_16x16x16_bf16
//
// The accumulator layout of one product is the operand layout of the next
// (lane l holds rows 4*(l/16)+i of column l%16), so the chains never leave
// registers.  The voxel-side operand (residual, x-contracted increment) is
// rounded to BF16 (PREC 1) or split into hi + lo BF16 halves (PREC 2); the
// B-spline basis factors are always hi + lo (their rounding would otherwise
// be systematic); products dh*bh + dh*bl (+ dl*bh); accumulation is FP32
// throughout.  The voxel-side work (remap, masks, exp, FP64 CV sums)
// is identical to the FP32 kernels.
//
// A wave owns 16 rows (y) of the slab at a time; 4 waves per workgroup.
// Control-point counts per axis are padded to 16 (MT/NT tiles of 16).
#pragma once

namespace n4 {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// float -> BF16 bits, round to nearest even (finite inputs).
__device__ inline unsigned short bf16_rn(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf16_to_f32(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

// f ~= hi + lo with both halves BF16.
__device__ inline void bf16_split(float f, unsigned short& hi, unsigned short& lo) {
    hi = bf16_rn(f);
    lo = bf16_rn(f - bf16_to_f32(hi));
}

__device__ inline f32x4 mma_k32(u16x8 a, u16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}
__device__ inline f32x4 mma_k16(i16x4 a, i16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// c += data * basis.  The B-spline basis factors (weights, w^3/S, w^2) are
// the same for every voxel, so their rounding would be a systematic error:
// they are always carried as hi + lo.  The voxel-side operand is one BF16
// (PREC 1) or hi + lo (PREC 2).  Products: dh*bh + dh*bl (+ dl*bh).
// DATA_A: the data operand is the instruction's A operand (else B).
template <int PREC, bool DATA_A>
__device__ inline f32x4 mma_k32p(u16x8 dh, u16x8 dl, u16x8 bh, u16x8 bl, f32x4 c) {
    if constexpr (PREC == 2) c = DATA_A ? mma_k32(dl, bh, c) : mma_k32(bh, dl, c);
    c = DATA_A ? mma_k32(dh, bl, c) : mma_k32(bl, dh, c);
    return DATA_A ? mma_k32(dh, bh, c) : mma_k32(bh, dh, c);
}
template <int PREC, bool DATA_A>
__device__ inline f32x4 mma_k16p(i16x4 dh, i16x4 dl, i16x4 bh, i16x4 bl, f32x4 c) {
    if constexpr (PREC == 2) c = DATA_A ? mma_k16(dl, bh, c) : mma_k16(bh, dl, c);
    c = DATA_A ? mma_k16(dh, bl, c) : mma_k16(bl, dh, c);
    return DATA_A ? mma_k16(dh, bh, c) : mma_k16(bh, dh, c);
}

// Four accumulator floats as a K=16 operand (hi / lo).
template <int PREC>
__device__ inline void to_op4(f32x4 v, i16x4& hi, i16x4& lo) {
    UNROLL for (int i = 0; i < 4; ++i) {
        unsigned short h, l = 0;
        if constexpr (PREC == 2)
            bf16_split(v[i], h, l);
        else
            h = bf16_rn(v[i]);
        hi[i] = (short)h;
        lo[i] = (short)l;
    }
}

// BF16 factor table in LDS: rows of `stride` halves, hi and (has_lo) lo.
// (An LDS pointer is never tested against null: offset 0 is a valid
// carve-out and compares equal to the LDS null pointer.)
struct Tab {
    N4_LDS unsigned short* hi;
    N4_LDS unsigned short* lo;
    int stride;
    bool has_lo;
    __device__ void put(int r, int c, float v) const {
        unsigned short h, l;
        bf16_split(v, h, l);
        hi[r * stride + c] = h;
        if (has_lo) lo[r * stride + c] = l;
    }
    __device__ i16x4 get4(const N4_LDS unsigned short* t, int r, int c) const {
        return *reinterpret_cast<const N4_LDS i16x4*>(t + r * stride + c);
    }
    __device__ u16x8 get8(const N4_LDS unsigned short* t, int r, int c) const {
        return *reinterpret_cast<const N4_LDS u16x8*>(t + r * stride + c);
    }
};

// One position of an axis (ep; nullptr for padding) in a factor table: 0
// for every control point except its 4-point support, whose factors are w
// (k 0), w^3/sum w^2 (k 1) or w^2 (k 2).  cp_rows: control points index the
// table's rows (else its columns).
__device__ inline void fill_axis(const Tab& T, bool cp_rows, int pos, int ncp_pad, int C, const AxisEntry* ep,
                                 int k) {
    for (int c = 0; c < ncp_pad; ++c) {
        if (cp_rows)
            T.put(c, pos, 0.0f);
        else
            T.put(pos, c, 0.0f);
    }
    if (!ep) return;
    const AxisEntry e = *ep;
    UNROLL for (int a = 0; a < 4; ++a) {
        const int c = e.b + a;
        const float v = k == 0 ? e.w[a] : (k == 1 ? e.p[a] : e.q[a]);
        if (c >= 0 && c < C) {
            if (cp_rows)
                T.put(c, pos, v);
            else
                T.put(pos, c, v);
        }
    }
}

// ------------------------------------------------------------------- fit --
struct FitMfmaLds {
    N4_LDS float* E;
    Tab Pz;   // [NT*16][Zp + 8]  Pz^T: row = control point, column = z
    Tab Py;   // [MT*16][Yp + 4]  Py^T: row = control point, column = y
    N4_LDS float* red;
    N4_LDS unsigned short* Gh;  // [4 waves][16][kGStride]  staged g, BF16 hi
    N4_LDS unsigned short* Gl;  // same, lo (PREC 2)
    N4_LDS AxisEntry* ex;       // the slab's x factors (copied in the set-up round)
};

// g staging: a wave loads 16 rows x one z chunk of at most 64 (4-voxel runs
// spread over the lanes, see fit_mfma_load), remaps, and parks the BF16
// halves here; the MFMA A operands (8 consecutive z per lane) are then 16-B
// LDS reads.
constexpr int kGStride = 72;  // halves per staged row (144 B: rows stay 16-B aligned)
constexpr int kFitRuns = 4;   // 4-voxel runs per lane: 16 rows x 64 z over 64 lanes
HD int fit_chunk_width(int Z, int zc) { return Z - zc > 64 ? 64 : Z - zc; }

HD size_t carve_fit_mfma(N4_LDS unsigned char* base, int nbins, int Y, int Z, int MT, int NT, int prec, FitMfmaLds* o) {
    Carve c{base, 0};
    FitMfmaLds L;
    L.E = c.take<float>(nbins);
    const int zs = round_up(Z, 32) + 8, ys = round_up(Y, 16) + 4;
    L.Pz.stride = zs;
    L.Pz.has_lo = true;
    L.Pz.hi = c.take<unsigned short>((size_t)NT * 16 * zs);
    L.Pz.lo = c.take<unsigned short>((size_t)NT * 16 * zs);
    L.Py.stride = ys;
    L.Py.has_lo = true;
    L.Py.hi = c.take<unsigned short>((size_t)MT * 16 * ys);
    L.Py.lo = c.take<unsigned short>((size_t)MT * 16 * ys);
    // one slot per wave for up to two tiles of control points (a single
    // barrier before the combine), else one running slot
    L.red = c.take<float>((size_t)MT * NT * 256 * (MT * NT <= 2 ? 4 : 1));
    L.Gh = c.take<unsigned short>((size_t)4 * 16 * kGStride);
    L.Gl = prec == 2 ? c.take<unsigned short>((size_t)4 * 16 * kGStride) : nullptr;
    L.ex = c.take<AxisEntry>(1);
    if (o) *o = L;
    return align16(c.off);
}

// Thread t's share of the fit's BF16 factor tables: Pz^T over the z extent
// (NT*16 rows) and Py^T over rows [ylo, ylo + ny) of the y axis (MT*16
// rows, columns tile-local); kf 2: w^2 (denominator), else w^3/sum w^2.
// One axis position per thread (z positions first, then y): its entry is
// loaded once (one round of independent global loads for the whole
// block) and the thread fills the position's column, zeros outside the
// 4-point support and in the padding; consecutive threads write
// consecutive columns of a control point's row.
template <int MT, int NT>
__device__ void fit_mfma_tables_t(const LevelDesc& LD, const FitMfmaLds& S, int Z, int ylo, int ny, int kf, bool pz,
                                  bool py, int t) {
    const int Zp = pz ? round_up(Z, 32) : 0, Yp = py ? round_up(ny, 16) : 0;
    for (int e = t; e < Zp + Yp; e += kBlock) {
        const bool isz = e < Zp;
        const int pos = isz ? e : e - Zp;
        const bool live = isz ? pos < Z : pos < ny;
        const int C = isz ? LD.C[2] : LD.C[1];
        float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        int b = 0;
        if (live) {
            const AxisEntry& a = isz ? LD.ax[2][pos] : LD.ax[1][ylo + pos];
            b = a.b;
            UNROLL for (int k = 0; k < 4; ++k) f[k] = kf == 2 ? a.q[k] : a.p[k];
        }
        const Tab& T = isz ? S.Pz : S.Py;
        const int rows = isz ? NT * 16 : MT * 16;
        for (int c = 0; c < rows; ++c) {
            const int k = c - b;
            float v = 0.0f;
            UNROLL for (int a = 0; a < 4; ++a)
                if (k == a) v = f[a];
            T.put(c, pos, (live && c < C) ? v : 0.0f);
        }
    }
}

// The voxel block of one wave: 16 tile-local rows from y0, the z chunk
// from zc (fit_chunk_width).  Its 16 x nq 4-voxel runs (nq runs per row
// of the chunk) are spread over the lanes: lane l takes runs q = l + 64 i
// (i < kFitRuns, q < 16 nq), run q being row q / nq at z = zc + 4 (q % nq),
// so a lane issues ceil(nq / 4) loads of each kind whatever the row length.
// Value 4i + j is z + j of run i: its logU from one 16-byte load
// (dword-aligned), its mask bytes one by one; runs cut by the row's end
// load voxel by voxel; runs outside the tile hold zeros.
struct FitRun {
    int row, col;  // block row, chunk-local z of the run's first voxel
    bool live;     // q < 16 nq
};
__device__ inline FitRun fit_run(int nq, float rq, int i) {
    const int q = ((int)threadIdx.x & 63) + 64 * i;
    // q / nq for q < 256, nq <= 16: the quotient's fractional part is at
    // least 1/32 from an integer, far above the float error
    const int row = (int)(((float)q + 0.5f) * rq);
    return FitRun{row, 4 * (q - row * nq), q < 16 * nq};
}
__device__ inline void fit_mfma_load(const VolDesc& V, const Slab sl, int y0, int zc, bool den,
                                     uint8_t (&ms)[4 * kFitRuns], float (&xs)[4 * kFitRuns]) {
    const int Y = V.W[1], Z = V.W[2];
    const int nq = (fit_chunk_width(Z, zc) + 3) >> 2;
    const float rq = 1.0f / (float)nq;
    UNROLL for (int i = 0; i < kFitRuns; ++i) {
        const FitRun u = fit_run(nq, rq, i);
        const int zq = zc + u.col;  // < Z for a live run
        const bool rok = u.live && y0 + u.row < sl.ny;
        const size_t v = ((size_t)sl.x * Y + sl.y0 + (rok ? y0 + u.row : 0)) * Z + (rok ? zq : 0);
        F4 x4 = F4{{0.0f, 0.0f, 0.0f, 0.0f}};
        if (rok && !den) {
            if (zq + 4 <= Z) {
                x4 = load_run4(V.logU + v);
            } else {
                UNROLL for (int j = 0; j < 4; ++j)
                    if (zq + j < Z) x4.v[j] = gp(V.logU)[v + j];
            }
        }
        UNROLL for (int j = 0; j < 4; ++j) {
            ms[4 * i + j] = (rok && zq + j < Z) ? gp(V.wmask)[v + j] : 0;
            xs[4 * i + j] = x4.v[j];
        }
    }
}

// The (y, z) contraction of one tile on MFMA (tables, E for NUM, and the
// slab's x factors in S.ex ready; ms/xs hold this wave's first block,
// fit_mfma_load(16 * wave, 0)), then the scatter (or the slab partial for
// stage calls).
template <int PREC, int MT, int NT>
__device__ void fit_mfma_tile(const VolDesc& V, const LevelDesc& LD, const Slab sl, const FitMfmaLds& S, int gmode,
                              bool den, bool scatter, bool degenerate, const Remap& rm,
                              uint8_t (&ms)[4 * kFitRuns], float (&xs)[4 * kFitRuns]) {
    const int Z = V.W[2];
    const int Cy = LD.C[1], Cz = LD.C[2];
    const int x = sl.x;
    const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6, r16 = lane & 15, g4 = lane >> 4;
    const int Yt = sl.ny;  // the tile's rows sl.y0 .. sl.y0 + Yt (Py columns are tile-local)
    f32x4 acc[MT][NT];
    UNROLL for (int m = 0; m < MT; ++m) UNROLL for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0, 0, 0, 0};

    N4_LDS unsigned short* Gh = S.Gh + wv * 16 * kGStride;
    N4_LDS unsigned short* Gl = (PREC == 2) ? S.Gl + wv * 16 * kGStride : nullptr;
    // The staged block's columns past a chunk's runs (up to its last K
    // step) meet zero rows of Pz but are read: they start as zeros (the
    // LDS may hold anything) and later hold only staged finite values.
    // A wave's LDS accesses run in order, so its own block needs no barrier.
    {
        constexpr int kQ = 16 * kGStride / 8;  // 16-byte words of the wave's block
        UNROLL for (int k = lane; k < kQ; k += 64) {
            reinterpret_cast<N4_LDS u16x8*>(Gh)[k] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
            if constexpr (PREC == 2) reinterpret_cast<N4_LDS u16x8*>(Gl)[k] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
        }
    }
    // every wave runs the same trip counts (barriers inside)
    for (int rb = 0; rb < Yt; rb += 64) {
        const int y0 = rb + 16 * wv;  // this wave's 16 tile-local rows (past Yt: zeros)
        f32x4 t1[NT];
        UNROLL for (int n = 0; n < NT; ++n) t1[n] = f32x4{0, 0, 0, 0};
        for (int zc = 0; zc < Z;) {
            const int zw = fit_chunk_width(Z, zc);
            if (rb > 0 || zc > 0) fit_mfma_load(V, sl, y0, zc, den, ms, xs);  // (the first block is already in registers)
            const int nq = (zw + 3) >> 2;
            const float rq = 1.0f / (float)nq;
            UNROLL for (int i = 0; i < kFitRuns; ++i) {
                const FitRun u = fit_run(nq, rq, i);
                if (u.live) {
                    // the run's 4 halves in one 8-byte store (rows are 144 B, runs 8 B aligned)
                    uint32_t h2[2], l2[2] = {0u, 0u};
                    UNROLL for (int j = 0; j < 4; j += 2) {
                        unsigned short h0, h1, l0 = 0, l1 = 0;
                        const float g0 = fit_input(gmode, ms[4 * i + j], xs[4 * i + j], degenerate, rm);
                        const float g1 = fit_input(gmode, ms[4 * i + j + 1], xs[4 * i + j + 1], degenerate, rm);
                        if constexpr (PREC == 2) {
                            bf16_split(g0, h0, l0);
                            bf16_split(g1, h1, l1);
                        } else {
                            h0 = bf16_rn(g0);
                            h1 = bf16_rn(g1);
                        }
                        h2[j >> 1] = (uint32_t)h0 | ((uint32_t)h1 << 16);
                        l2[j >> 1] = (uint32_t)l0 | ((uint32_t)l1 << 16);
                    }
                    const int slot = u.row * kGStride + u.col;
                    *reinterpret_cast<N4_LDS u32x2*>(Gh + slot) = u32x2{h2[0], h2[1]};
                    if constexpr (PREC == 2) *reinterpret_cast<N4_LDS u32x2*>(Gl + slot) = u32x2{l2[0], l2[1]};
                }
            }
            __syncthreads();
            UNROLL for (int ks = 0; ks < 2; ++ks) {
                const int z0 = zc + 32 * ks;
                if (32 * ks < zw) {
                    const u16x8 ah = *reinterpret_cast<const N4_LDS u16x8*>(Gh + r16 * kGStride + 32 * ks + 8 * g4);
                    const u16x8 al =
                        (PREC == 2) ? *reinterpret_cast<const N4_LDS u16x8*>(Gl + r16 * kGStride + 32 * ks + 8 * g4) : ah;
                    UNROLL for (int n = 0; n < NT; ++n) {
                        const u16x8 bh = S.Pz.get8(S.Pz.hi, n * 16 + r16, z0 + 8 * g4);
                        const u16x8 bl = S.Pz.get8(S.Pz.lo, n * 16 + r16, z0 + 8 * g4);
                        t1[n] = mma_k32p<PREC, true>(ah, al, bh, bl, t1[n]);
                    }
                }
            }
            __syncthreads();
            zc += zw;
        }
        if (y0 >= Yt) continue;  // no rows: nothing to add (and no Py rows)
        // t1[n]: lane holds T1[y0 + 4*g4 + i][16n + r16] = the K=16 operand (k = y)
        UNROLL for (int n = 0; n < NT; ++n) {
            i16x4 bh, bl;
            to_op4<PREC>(t1[n], bh, bl);
            UNROLL for (int m = 0; m < MT; ++m) {
                const i16x4 a_h = S.Py.get4(S.Py.hi, m * 16 + r16, y0 + 4 * g4);
                const i16x4 a_l = S.Py.get4(S.Py.lo, m * 16 + r16, y0 + 4 * g4);
                acc[m][n] = mma_k16p<PREC, false>(bh, bl, a_h, a_l, acc[m][n]);
            }
        }
    }
    // waves combine in a fixed order (bitwise reproducible): with a slot
    // per wave, one barrier and the sum over the slots in wave order;
    // else the waves add into one slot in turn
    constexpr int kE = MT * NT * 256;
    constexpr bool kSlots = MT * NT <= 2;
    if constexpr (kSlots) {
        UNROLL for (int m = 0; m < MT; ++m) UNROLL for (int n = 0; n < NT; ++n) UNROLL for (int i = 0; i < 4; ++i)
            S.red[wv * kE + (((m * NT + n) * 4 + i) << 6) + lane] = acc[m][n][i];
        __syncthreads();
    } else {
        for (int w = 0; w < 4; ++w) {
            if (wv == w) {
                UNROLL for (int m = 0; m < MT; ++m) UNROLL for (int n = 0; n < NT; ++n) UNROLL for (int i = 0; i < 4; ++i) {
                    const int idx = (((m * NT + n) * 4 + i) << 6) + lane;
                    S.red[idx] = (w == 0 ? 0.0f : S.red[idx]) + acc[m][n][i];
                }
            }
            __syncthreads();
        }
    }
    float* out = V.part + (size_t)x * V.part_stride;
    const AxisEntry ex = *S.ex;
    for (int e = t; e < kE; e += kBlock) {
        const int ln = e & 63, i = (e >> 6) & 3, tile = e >> 8;
        const int m = tile / NT, n = tile - (tile / NT) * NT;
        const int cy = m * 16 + 4 * (ln >> 4) + i, cz = n * 16 + (ln & 15);
        if (cy < Cy && cz < Cz) {
            const float r = kSlots ? ((S.red[e] + S.red[kE + e]) + S.red[2 * kE + e]) + S.red[3 * kE + e] : S.red[e];
            if (scatter)
                scatter_numerator(V, LD, ex, tile_id(V, sl) % V.num_copies, den, cy * Cz + cz, r);
            else
                out[cy * Cz + cz] = r;
        }
    }
}

template <int PREC, int MT, int NT>
__device__ void body_fit_mfma(const VolDesc* vols, const Slab* slabs, int level, int it, int mode, RunParams rp,
                              int bid, N4_LDS unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    const VolState* st = rp.states + sl.vol;
    if (!fit_step_mode(st, &mode, &level, &it)) return;
    const LevelDesc& LD = V.lev[level];
    const int Z = V.W[2];
    const int nb = rp.nbins;
    FitMfmaLds S;
    carve_fit_mfma(smem, nb, V.wtile_rows, Z, MT, NT, PREC, &S);
    const bool num = (mode == kFitNum);
    const bool den = (mode == kFitDen || mode == kFitDenScatter);
    const bool scatter = (mode == kFitNum || mode == kFitDenScatter);
    const int gmode = den ? (int)kFitDen : mode;
    const bool degenerate = num && st->degenerate;
    const double bmin = st->bin_min, slope = st->slope;
    const int kf = den ? 2 : 1;  // q = w^2 for the denominator, p = w^3/S otherwise
    const int t = (int)threadIdx.x, wv = t >> 6;
    const Remap rm = make_remap(S.E, bmin, slope, nb);
    // the first block of voxels (row batch 0, z chunk 0) is loaded before
    // the table set-up, so the two global round trips overlap
    uint8_t ms[4 * kFitRuns];
    float xs[4 * kFitRuns];
    fit_mfma_load(V, sl, 16 * wv, 0, den, ms, xs);
    if (num)
        for (int i = t; i < nb; i += kBlock) S.E[i] = (float)V.etab[i];
    fit_mfma_tables_t<MT, NT>(LD, S, Z, sl.y0, sl.ny, kf, true, true, t);
    // the slab's x factors for the final scatter, in this round of loads
    if (t < (int)(sizeof(AxisEntry) / 4))
        reinterpret_cast<N4_LDS int32_t*>(S.ex)[t] = reinterpret_cast<const int32_t*>(&LD.ax[0][sl.x])[t];
    __syncthreads();
    fit_mfma_tile<PREC, MT, NT>(V, LD, sl, S, gmode, den, scatter, degenerate, rm, ms, xs);
}

// ------------------------------------------------------------------ eval --
// Working-grid evaluation of the lattice increment (logU update and CV
// sums).  Four control planes along x feed each slab: A[cy,cz] =
// sum_k wx[k] plane_k[cy,cz] (FP32), then the (y, z) chain runs on MFMA.
// (The final full-resolution reconstruction streams 12 B per voxel from
// HBM and stays on the FP32 kernel in every mode: an MFMA form with its
// BF16 tables in LDS ran 5x slower there, limited by occupancy.)
struct EvalMfmaLds {
    N4_LDS float* A;    // [Cy*Cz]     x-contracted planes
    Tab At;      // [NT*16][MT*16 + 4]  A^T: row = cz, column = cy
    Tab Wy;      // [Yrows][MT*16 + 4]  row = y (slab-local), column = cy
    Tab Wz;      // [Zp16][NT*16 + 4]   row = z, column = cz
    N4_LDS double* red; // [4*kBlock]
};

HD size_t carve_eval_mfma(N4_LDS unsigned char* base, int Yrows, int Z, int Cy, int Cz, int MT, int NT, int prec,
                          EvalMfmaLds* o) {
    Carve c{base, 0};
    EvalMfmaLds L;
    L.A = c.take<float>((size_t)Cy * Cz);
    const int as = MT * 16 + 4, zs = NT * 16 + 4;
    L.At.stride = as;
    L.At.has_lo = prec == 2;
    L.At.hi = c.take<unsigned short>((size_t)NT * 16 * as);
    L.At.lo = prec == 2 ? c.take<unsigned short>((size_t)NT * 16 * as) : nullptr;
    const int yr = round_up(Yrows, 16);
    L.Wy.stride = as;
    L.Wy.has_lo = true;
    L.Wy.hi = c.take<unsigned short>((size_t)yr * as);
    L.Wy.lo = c.take<unsigned short>((size_t)yr * as);
    const int zr = round_up(Z, 16);
    L.Wz.stride = zs;
    L.Wz.has_lo = true;
    L.Wz.hi = c.take<unsigned short>((size_t)zr * zs);
    L.Wz.lo = c.take<unsigned short>((size_t)zr * zs);
    L.red = c.take<double>((size_t)4 * kBlock);
    if (o) *o = L;
    return align16(c.off);
}

// A from the 4 planes (plane k at planes + k*Cy*Cz) and the BF16 tables.
// Rows [ylo, ylo + yrows) of axis table ay are the slab's rows.
template <int MT, int NT>
__device__ void eval_mfma_setup(const EvalMfmaLds& S, const AxisEntry ex, const float* planes, const AxisEntry* ay,
                                int ylo, int yrows, const AxisEntry* az, int Z, int Cy, int Cz) {
    const int t = (int)threadIdx.x;
    const int CyCz = Cy * Cz;
    for (int pr = t; pr < CyCz; pr += kBlock) {
        float a = 0.0f;
        UNROLL for (int k = 0; k < 4; ++k) a += ex.w[k] * planes[(size_t)k * CyCz + pr];
        S.A[pr] = a;
    }
    const int yr = round_up(yrows, 16);
    for (int r = t; r < yr; r += kBlock) fill_axis(S.Wy, false, r, MT * 16, Cy, r < yrows ? ay + ylo + r : nullptr, 0);
    const int zr = round_up(Z, 16);
    for (int z = t; z < zr; z += kBlock) fill_axis(S.Wz, false, z, NT * 16, Cz, z < Z ? az + z : nullptr, 0);
    __syncthreads();
    for (int e = t; e < NT * 16 * MT * 16; e += kBlock) {
        const int cz = e / (MT * 16), cy = e - cz * (MT * 16);
        S.At.put(cz, cy, (cz < Cz && cy < Cy) ? S.A[cy * Cz + cz] : 0.0f);
    }
    __syncthreads();
}

// Bt[n] operands for the 16 rows starting at slab-local row r0:
// lane holds Bt[c = 16n + 4*g4 + i][y = r0 + r16] as a K=16 operand.
template <int PREC, int MT, int NT>
__device__ inline void eval_mfma_rows(const EvalMfmaLds& S, int r0, int r16, int g4, i16x4 (&bh)[NT],
                                      i16x4 (&bl)[NT]) {
    UNROLL for (int n = 0; n < NT; ++n) {
        f32x4 b = f32x4{0, 0, 0, 0};
        UNROLL for (int k = 0; k < MT; ++k) {
            const i16x4 a_h = S.At.get4(S.At.hi, n * 16 + r16, k * 16 + 4 * g4);
            const i16x4 a_l = (PREC == 2) ? S.At.get4(S.At.lo, n * 16 + r16, k * 16 + 4 * g4) : a_h;
            const i16x4 w_h = S.Wy.get4(S.Wy.hi, r0 + r16, k * 16 + 4 * g4);
            const i16x4 w_l = S.Wy.get4(S.Wy.lo, r0 + r16, k * 16 + 4 * g4);
            b = mma_k16p<PREC, true>(a_h, a_l, w_h, w_l, b);
        }
        to_op4<PREC>(b, bh[n], bl[n]);
    }
}

// f[z = z0 + 4*g4 + i][y = r16 row] for one 16x16 (z, y) tile.
template <int PREC, int NT>
__device__ inline f32x4 eval_mfma_tile(const EvalMfmaLds& S, int z0, int r16, int g4, const i16x4 (&bh)[NT],
                                       const i16x4 (&bl)[NT]) {
    f32x4 f = f32x4{0, 0, 0, 0};
    UNROLL for (int n = 0; n < NT; ++n) {
        const i16x4 w_h = S.Wz.get4(S.Wz.hi, z0 + r16, n * 16 + 4 * g4);
        const i16x4 w_l = S.Wz.get4(S.Wz.lo, z0 + r16, n * 16 + 4 * g4);
        f = mma_k16p<PREC, false>(bh[n], bl[n], w_h, w_l, f);
    }
    return f;
}

// Field increment d on the working grid, logU -= d, CV sums and logU
// min/max per slab (same contract as the FP32 evaluation kernel).
template <int PREC, int MT, int NT>
__device__ void body_eval_mfma(const VolDesc* vols, const Slab* slabs, RunParams rp, int bid, N4_LDS unsigned char* smem) {
    const Slab sl = slabs[bid];
    const VolDesc& V = vols[sl.vol];
    const VolState* st = rp.states + sl.vol;
    if (st->s_kind != kStepIter) return;
    const int level = st->s_lvl;
    const LevelDesc& LD = V.lev[level];
    const int Y = V.W[1], Z = V.W[2];
    const int Cy = LD.C[1], Cz = LD.C[2];
    EvalMfmaLds S;
    carve_eval_mfma(smem, V.wtile_rows, Z, Cy, Cz, MT, NT, PREC, &S);
    const int x = sl.x;
    const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6, r16 = lane & 15, g4 = lane >> 4;
    const AxisEntry ex = LD.ax[0][x];
    const int Yt = sl.ny;  // the tile's rows (Wy rows are tile-local)
    // the 4 increment planes slab x touches (FP32, written by the update kernel)
    eval_mfma_setup<MT, NT>(S, ex, V.lresf + (size_t)ex.b * Cy * Cz, LD.ax[1], sl.y0, Yt, LD.ax[2], Z, Cy, Cz);
    double s1 = 0.0, s2 = 0.0;
    float mn = INFINITY, mx = -INFINITY;  // exact in FP32
    for (int y0 = 16 * wv; y0 < Yt; y0 += 64) {
        i16x4 bh[NT], bl[NT];
        eval_mfma_rows<PREC, MT, NT>(S, y0, r16, g4, bh, bl);
        const bool yok = y0 + r16 < Yt;
        const size_t row = ((size_t)x * Y + sl.y0 + (yok ? y0 + r16 : 0)) * Z;
        for (int z0 = 0; z0 < Z; z0 += 16) {
            const f32x4 d = eval_mfma_tile<PREC, NT>(S, z0, r16, g4, bh, bl);
            const int zb = z0 + 4 * g4;
            uint8_t ms[4];
            float lu[4];
            UNROLL for (int i = 0; i < 4; ++i) {
                const bool ok = yok && (zb + i) < Z;
                ms[i] = ok ? gp(V.wmask)[row + zb + i] : 0;
                lu[i] = ok ? gp(V.logU)[row + zb + i] : 0.0f;
            }
            UNROLL for (int i = 0; i < 4; ++i) {
                if (!ms[i]) continue;
                const float lg = lu[i] - d[i];
                gp(V.logU)[row + zb + i] = lg;
                const double uu = (double)expm1f(-d[i]);
                s1 += uu;
                s2 += uu * uu;
                mn = fminf(mn, lg);
                mx = fmaxf(mx, lg);
            }
        }
    }
    S.red[t] = s1;
    S.red[kBlock + t] = s2;
    S.red[2 * kBlock + t] = mn;
    S.red[3 * kBlock + t] = mx;
    tree_reduce4(S.red, 0, 0, 1, 2);
    if (t == 0) {
        double* sp = V.spart + (size_t)tile_id(V, sl) * 4;
        sp[0] = S.red[0];
        sp[1] = S.red[kBlock];
        sp[2] = S.red[2 * kBlock];
        sp[3] = S.red[3 * kBlock];
    }
}

}  // namespace n4
