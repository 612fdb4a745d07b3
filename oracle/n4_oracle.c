/*
 * n4_oracle.c -- CPU FP64 restatement of the reference N4 (N4.py).
 *
 * TEST INFRASTRUCTURE ONLY: the parity reference for the HIP product path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Every function cites the N4.py lines it restates; deviations are
 * listed in n4_oracle.h.  Built with -ffp-contract=off so that every FP64
 * operation rounds exactly as written (the histogram bin index in particular).
 */
#include "n4_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void n4o_default_params(n4o_params *p) {
    memset(p, 0, sizeof(*p));
    p->nbins = 200;
    p->wiener_noise = 0.01;
    p->fwhm = 0.15;
    p->ncpt[0] = p->ncpt[1] = p->ncpt[2] = 4;
    p->levels = 4;
    for (int i = 0; i < N4O_MAX_LEVELS; ++i) p->max_iter[i] = 50;
    p->conv_threshold = 0.001;
    p->low_value = 10.0;
    p->shrink = 1;
    p->fit_eps = 1e-3;
    p->nthreads = 1;
}

/* N4.py:174-181 */
void n4o_cubic_bspline(double t, double w[4]) {
    double t2 = t * t;
    double t3 = t * t * t;
    double om = 1.0 - t;
    w[0] = om * om * om / 6.0;
    w[1] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
    w[2] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
    w[3] = t3 / 6.0;
}

/* N4.py:192 (spans = size/numberOfSpans), :210-217 (b, t) */
void n4o_axis_table(int32_t npos, double domain, int32_t ncpt, int32_t *b,
                    double *w) {
    double span = domain / (double)(ncpt - 3);
    for (int32_t i = 0; i < npos; ++i) {
        double pos = (double)i + 0.5;
        int32_t bi = (int32_t)floor(pos / span);
        double t = (pos - (double)bi * span) / span;
        b[i] = bi;
        n4o_cubic_bspline(t, w + 4 * i);
    }
}

/* N4.py:31-37: exponent = ceil(log2(nbins)) + 1, P = 2^exponent,
 * offset = int(0.5*(P - nbins)).  log2 is taken in integers. */
void n4o_padded_size(int32_t nbins, int32_t *P, int32_t *offset) {
    int32_t e = 0;
    while ((1 << e) < nbins) ++e;
    *P = 1 << (e + 1);
    *offset = (int32_t)(0.5 * (double)(*P - nbins));
}

void n4o_fft(double *data, int32_t n, int32_t inverse) {
    /* bit reversal */
    for (int32_t i = 1, j = 0; i < n; ++i) {
        int32_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = data[2 * i], ti = data[2 * i + 1];
            data[2 * i] = data[2 * j];
            data[2 * i + 1] = data[2 * j + 1];
            data[2 * j] = tr;
            data[2 * j + 1] = ti;
        }
    }
    const double sgn = inverse ? 1.0 : -1.0;
    for (int32_t len = 2; len <= n; len <<= 1) {
        int32_t half = len >> 1;
        for (int32_t k = 0; k < half; ++k) {
            double ang = sgn * 2.0 * M_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (int32_t s = 0; s < n; s += len) {
                double *a = data + 2 * (s + k);
                double *c = data + 2 * (s + k + half);
                double xr = c[0] * wr - c[1] * wi;
                double xi = c[0] * wi + c[1] * wr;
                c[0] = a[0] - xr;
                c[1] = a[1] - xi;
                a[0] = a[0] + xr;
                a[1] = a[1] + xi;
            }
        }
    }
    if (inverse) {
        double inv = 1.0 / (double)n;
        for (int32_t i = 0; i < 2 * n; ++i) data[i] *= inv;
    }
}

/* N4.py:14 */
int64_t n4o_masked_minmax(const double *x, const uint8_t *mask, int64_t n,
                          double *mn, double *mx) {
    int64_t cnt = 0;
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        if (!mask[i]) continue;
        ++cnt;
        if (x[i] < lo) lo = x[i];
        if (x[i] > hi) hi = x[i];
    }
    *mn = lo;
    *mx = hi;
    return cnt;
}

/* N4.py:19-29 */
void n4o_histogram(const double *x, const uint8_t *mask, int64_t n,
                   int32_t nbins, double bin_min, double slope, double *H,
                   int32_t *idx_out) {
    for (int32_t i = 0; i < nbins; ++i) H[i] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (!mask[i]) {
            if (idx_out) idx_out[i] = -1;
            continue;
        }
        double cidx = (x[i] - bin_min) / slope;
        int32_t idx = (int32_t)floor(cidx);
        double offset = cidx - (double)idx;
        if (idx_out) idx_out[i] = idx;
        if (offset == 0.0) {
            H[idx] += 1.0;
        } else if (idx < nbins - 1) {
            H[idx] += 1.0 - offset;
            H[idx + 1] += offset;
        }
    }
}

/* N4.py:31-101 */
void n4o_wiener_etable(const double *H, int32_t nbins, double bin_min,
                       double slope, double noise, double fwhm, double *E) {
    int32_t P, off;
    n4o_padded_size(nbins, &P, &off);
    double *V = calloc((size_t)2 * P, sizeof(double));
    double *F = calloc((size_t)2 * P, sizeof(double));
    double *U = calloc((size_t)2 * P, sizeof(double));
    double *Nm = calloc((size_t)2 * P, sizeof(double));
    double *Dn = calloc((size_t)2 * P, sizeof(double));
    double *Fr = calloc((size_t)P, sizeof(double));

    /* N4.py:39-43: padded histogram, Vf = fft(V) */
    for (int32_t i = 0; i < nbins; ++i) V[2 * (off + i)] = H[i];
    n4o_fft(V, P, 0);

    /* N4.py:46-60: Gaussian kernel */
    double scaledFWHM = fwhm / slope;
    double expFactor = 4.0 * log(2.0) / (scaledFWHM * scaledFWHM);
    double scaleFactor = 2.0 * sqrt(log(2.0) / M_PI) / scaledFWHM;
    int32_t halfSize = (int32_t)(0.5 * (double)P);
    Fr[0] = scaleFactor;
    for (int32_t i = 1; i <= halfSize; ++i) {
        double v = scaleFactor * exp((double)(-(int64_t)i * i) * expFactor);
        Fr[i] = v;
        Fr[P - i] = v;
    }
    if (P % 2 == 0)
        Fr[halfSize] = scaleFactor * exp(-0.25 * P * P * expFactor);
    for (int32_t i = 0; i < P; ++i) F[2 * i] = Fr[i];
    n4o_fft(F, P, 0); /* Ff, N4.py:62 */

    /* N4.py:64-74: Gf = conj(Ff)/(conj(Ff)Ff + noise); Uf = Vf*Re(Gf) */
    for (int32_t i = 0; i < P; ++i) {
        double ar = F[2 * i], ai = F[2 * i + 1];
        double d = (ar * ar + ai * ai) + noise;
        double g = ar / d;
        U[2 * i] = V[2 * i] * g;
        U[2 * i + 1] = V[2 * i + 1] * g;
    }
    n4o_fft(U, P, 1);
    for (int32_t i = 0; i < P; ++i) {
        double u = U[2 * i];
        if (u < 0) u = 0;
        U[2 * i] = u;
        U[2 * i + 1] = 0.0;
    }

    /* N4.py:76-89: numerator = bin centre * U, smoothed by Ff; same for U */
    for (int32_t i = 0; i < P; ++i) {
        Nm[2 * i] = (bin_min + (double)(i - off) * slope) * U[2 * i];
        Nm[2 * i + 1] = 0.0;
        Dn[2 * i] = U[2 * i];
        Dn[2 * i + 1] = 0.0;
    }
    n4o_fft(Nm, P, 0);
    n4o_fft(Dn, P, 0);
    for (int32_t i = 0; i < P; ++i) {
        double fr = F[2 * i], fi = F[2 * i + 1];
        double nr = Nm[2 * i], ni = Nm[2 * i + 1];
        Nm[2 * i] = nr * fr - ni * fi;
        Nm[2 * i + 1] = nr * fi + ni * fr;
        double dr = Dn[2 * i], di = Dn[2 * i + 1];
        Dn[2 * i] = dr * fr - di * fi;
        Dn[2 * i + 1] = dr * fi + di * fr;
    }
    n4o_fft(Nm, P, 1);
    n4o_fft(Dn, P, 1);

    /* N4.py:91-101 */
    for (int32_t i = 0; i < nbins; ++i) {
        double nu = Nm[2 * (off + i)];
        double de = Dn[2 * (off + i)];
        E[i] = (de != 0.0) ? nu / de : 0.0;
    }
    free(V);
    free(F);
    free(U);
    free(Nm);
    free(Dn);
    free(Fr);
}

/* N4.py:9-118 */
void n4o_sharpen(const double *x, const uint8_t *mask, int64_t n,
                 int32_t nbins, double noise, double fwhm, double *out,
                 double *H_out, double *E_out, double *mms_out) {
    double mn, mx;
    int64_t cnt = n4o_masked_minmax(x, mask, n, &mn, &mx);
    double *H = calloc((size_t)nbins, sizeof(double));
    double *E = calloc((size_t)nbins, sizeof(double));
    double slope = 0.0;
    if (cnt == 0) {
        for (int64_t i = 0; i < n; ++i) out[i] = 0.0;
    } else {
        slope = (mx - mn) / (double)(nbins - 1);
        if (slope == 0.0) {
            /* degenerate: identity sharpening (N4.py would raise here) */
            for (int64_t i = 0; i < n; ++i) out[i] = mask[i] ? x[i] : 0.0;
        } else {
            n4o_histogram(x, mask, n, nbins, mn, slope, H, NULL);
            n4o_wiener_etable(H, nbins, mn, slope, noise, fwhm, E);
            /* N4.py:103-115 */
            for (int64_t i = 0; i < n; ++i) {
                if (!mask[i]) {
                    out[i] = 0.0;
                    continue;
                }
                double cidx = (x[i] - mn) / slope;
                int32_t idx = (int32_t)floor(cidx);
                if (idx < nbins - 1)
                    out[i] = E[idx] + (E[idx + 1] - E[idx]) * (cidx - (double)idx);
                else
                    out[i] = E[nbins - 1];
            }
        }
    }
    if (H_out) memcpy(H_out, H, sizeof(double) * nbins);
    if (E_out) memcpy(E_out, E, sizeof(double) * nbins);
    if (mms_out) {
        mms_out[0] = mn;
        mms_out[1] = mx;
        mms_out[2] = slope;
    }
    free(H);
    free(E);
}

typedef struct axis_tab {
    int32_t *b;
    double *w;
} axis_tab;

static void make_tabs(axis_tab t[3], const int32_t dims[3],
                      const double domain[3], const int32_t ncpt[3]) {
    for (int d = 0; d < 3; ++d) {
        t[d].b = malloc(sizeof(int32_t) * dims[d]);
        t[d].w = malloc(sizeof(double) * 4 * dims[d]);
        n4o_axis_table(dims[d], domain[d], ncpt[d], t[d].b, t[d].w);
    }
}

static void free_tabs(axis_tab t[3]) {
    for (int d = 0; d < 3; ++d) {
        free(t[d].b);
        free(t[d].w);
    }
}

/* N4.py:186-237 direct form; accumulates numerator and denominator, each
 * voxel's terms scaled by its confidence weight conf[v] (NULL: 1). */
void n4o_fit_numden_w(const double *field, const uint8_t *mask,
                      const double *conf, const int32_t dims[3],
                      const int32_t ncpt[3], int32_t nthreads, double *num,
                      double *den) {
    const int32_t X = dims[0], Y = dims[1], Z = dims[2];
    const int64_t L = (int64_t)ncpt[0] * ncpt[1] * ncpt[2];
    double dom[3] = {X, Y, Z};
    axis_tab t[3];
    make_tabs(t, dims, dom, ncpt);
    int nt = nthreads > 0 ? nthreads : 1;
    double *pn = calloc((size_t)L * nt, sizeof(double));
    double *pd = calloc((size_t)L * nt, sizeof(double));
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int32_t i = 0; i < X; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *an = pn + (size_t)L * tid, *ad = pd + (size_t)L * tid;
        const int32_t bx = t[0].b[i];
        const double *wx = t[0].w + 4 * i;
        for (int32_t j = 0; j < Y; ++j) {
            const int32_t by = t[1].b[j];
            const double *wy = t[1].w + 4 * j;
            for (int32_t k = 0; k < Z; ++k) {
                int64_t v = ((int64_t)i * Y + j) * Z + k;
                if (mask[v] == 0) continue;
                const int32_t bz = t[2].b[k];
                const double *wz = t[2].w + 4 * k;
                double wc[64];
                double wc_sum = 0.0;
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b)
                        for (int c = 0; c < 4; ++c) {
                            double w = wx[a] * wy[b] * wz[c];
                            wc[a * 16 + b * 4 + c] = w;
                            wc_sum += w * w;
                        }
                double f = field[v];
                const double cw = conf ? conf[v] : 1.0;
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b)
                        for (int c = 0; c < 4; ++c) {
                            double w = wc[a * 16 + b * 4 + c];
                            double phi = f * w / wc_sum;
                            int64_t li = ((int64_t)(bx + a) * ncpt[1] + (by + b)) * ncpt[2] + (bz + c);
                            an[li] += cw * ((w * w) * phi);
                            ad[li] += cw * (w * w);
                        }
            }
        }
    }
    for (int64_t l = 0; l < L; ++l) {
        double sn = 0.0, sd = 0.0;
        for (int k = 0; k < nt; ++k) {
            sn += pn[(size_t)L * k + l];
            sd += pd[(size_t)L * k + l];
        }
        num[l] = sn;
        den[l] = sd;
    }
    free(pn);
    free(pd);
    free_tabs(t);
}

void n4o_fit_numden(const double *field, const uint8_t *mask,
                    const int32_t dims[3], const int32_t ncpt[3],
                    int32_t nthreads, double *num, double *den) {
    n4o_fit_numden_w(field, mask, NULL, dims, ncpt, nthreads, num, den);
}

/* N4.py:236: lattice = wc2_phic/(wc2 + eps) */
static void fit_bspline3d_w(const double *field, const uint8_t *mask,
                            const double *conf, const int32_t dims[3],
                            const int32_t ncpt[3], double eps,
                            int32_t nthreads, double *lattice) {
    const int64_t L = (int64_t)ncpt[0] * ncpt[1] * ncpt[2];
    double *num = malloc(sizeof(double) * L);
    double *den = malloc(sizeof(double) * L);
    n4o_fit_numden_w(field, mask, conf, dims, ncpt, nthreads, num, den);
    for (int64_t l = 0; l < L; ++l) lattice[l] = num[l] / (den[l] + eps);
    free(num);
    free(den);
}

void n4o_fit_bspline3d(const double *field, const uint8_t *mask,
                       const int32_t dims[3], const int32_t ncpt[3],
                       double eps, int32_t nthreads, double *lattice) {
    fit_bspline3d_w(field, mask, NULL, dims, ncpt, eps, nthreads, lattice);
}

/* N4.py:241-273 */
void n4o_evaluate_bspline3d(const double *lattice, const int32_t ncpt[3],
                            const int32_t dims[3], const double domain[3],
                            int32_t nthreads, double *out) {
    const int32_t X = dims[0], Y = dims[1], Z = dims[2];
    axis_tab t[3];
    make_tabs(t, dims, domain, ncpt);
    int nt = nthreads > 0 ? nthreads : 1;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int32_t i = 0; i < X; ++i) {
        const int32_t bx = t[0].b[i];
        const double *wx = t[0].w + 4 * i;
        for (int32_t j = 0; j < Y; ++j) {
            const int32_t by = t[1].b[j];
            const double *wy = t[1].w + 4 * j;
            for (int32_t k = 0; k < Z; ++k) {
                const int32_t bz = t[2].b[k];
                const double *wz = t[2].w + 4 * k;
                double s = 0.0;
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b)
                        for (int c = 0; c < 4; ++c) {
                            int64_t li = ((int64_t)(bx + a) * ncpt[1] + (by + b)) * ncpt[2] + (bz + c);
                            s += lattice[li] * (wx[a] * wy[b] * wz[c]);
                        }
                out[((int64_t)i * Y + j) * Z + k] = s;
            }
        }
    }
    free_tabs(t);
}

/* N4.py:121-126 */
double n4o_convergence(const double *f1, const double *f2,
                       const uint8_t *mask, int64_t n) {
    int64_t cnt = 0;
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i)
        if (mask[i]) {
            s += exp(f1[i] - f2[i]);
            ++cnt;
        }
    if (cnt == 0) return 0.0;
    double mu = s / (double)cnt;
    double ss = 0.0;
    for (int64_t i = 0; i < n; ++i)
        if (mask[i]) {
            double d = exp(f1[i] - f2[i]) - mu;
            ss += d * d;
        }
    double sigma = sqrt(ss / (double)cnt);
    return sigma / mu;
}

/* N4.py:132-165: dyadic refinement (n+3) -> (2n+3) per axis. */
void n4o_upsample_lattice3d(const double *lat, const int32_t ncpt[3],
                            double *out) {
    static const double bw[2][3] = {{1.0 / 8, 6.0 / 8, 1.0 / 8},
                                    {0.0, 1.0 / 2, 1.0 / 2}};
    const int32_t n[3] = {ncpt[0] - 3, ncpt[1] - 3, ncpt[2] - 3};
    const int32_t m[3] = {2 * n[0] + 3, 2 * n[1] + 3, 2 * n[2] + 3};
    for (int32_t i = -1; i <= n[0] + 1; ++i)
        for (int32_t j = -1; j <= n[1] + 1; ++j)
            for (int32_t k = -1; k <= n[2] + 1; ++k) {
                double piece[27];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int c = 0; c < 3; ++c) {
                            int32_t X = i - 1 + a, Y = j - 1 + b, Z = k - 1 + c;
                            X = X < -1 ? -1 : (X > n[0] + 1 ? n[0] + 1 : X);
                            Y = Y < -1 ? -1 : (Y > n[1] + 1 ? n[1] + 1 : Y);
                            Z = Z < -1 ? -1 : (Z > n[2] + 1 ? n[2] + 1 : Z);
                            piece[a * 9 + b * 3 + c] =
                                lat[((int64_t)(X + 1) * ncpt[1] + (Y + 1)) * ncpt[2] + (Z + 1)];
                        }
                for (int xx = 0; xx < 2; ++xx)
                    for (int yy = 0; yy < 2; ++yy)
                        for (int zz = 0; zz < 2; ++zz) {
                            int32_t ix = 2 * i + xx, iy = 2 * j + yy, iz = 2 * k + zz;
                            if (ix < -1 || ix > 2 * n[0] + 1) continue;
                            if (iy < -1 || iy > 2 * n[1] + 1) continue;
                            if (iz < -1 || iz > 2 * n[2] + 1) continue;
                            double s = 0.0;
                            for (int a = 0; a < 3; ++a)
                                for (int b = 0; b < 3; ++b)
                                    for (int c = 0; c < 3; ++c)
                                        s += (bw[xx][a] * bw[yy][b] * bw[zz][c]) * piece[a * 9 + b * 3 + c];
                            out[((int64_t)(ix + 1) * m[1] + (iy + 1)) * m[2] + (iz + 1)] = s;
                        }
            }
}

/* N4.py:280-287 (threshold, log) + the build-defined shrink; with
 * weighted != 0, mask_f holds confidence weights (foreground where > 0 and
 * finite) and conf_w the working voxels' mean foreground weights. */
void n4o_threshold_log_shrink_w(const float *im, const float *mask_f,
                                int32_t weighted, const int32_t fdims[3],
                                double low_value, int32_t s, int32_t wdims[3],
                                double *imlog_w, uint8_t *mask_w,
                                double *conf_w, int64_t *n_mask_full,
                                int64_t *n_mask_work) {
    const int32_t FX = fdims[0], FY = fdims[1], FZ = fdims[2];
    for (int d = 0; d < 3; ++d) wdims[d] = (fdims[d] + s - 1) / s;
    int64_t nfull = 0, nwork = 0;
    for (int32_t I = 0; I < wdims[0]; ++I)
        for (int32_t J = 0; J < wdims[1]; ++J)
            for (int32_t K = 0; K < wdims[2]; ++K) {
                int64_t nm = 0, nb = 0;
                double acc = 0.0, wsum = 0.0;
                for (int32_t i = I * s; i < (I + 1) * s && i < FX; ++i)
                    for (int32_t j = J * s; j < (J + 1) * s && j < FY; ++j)
                        for (int32_t k = K * s; k < (K + 1) * s && k < FZ; ++k) {
                            int64_t v = ((int64_t)i * FY + j) * FZ + k;
                            ++nb;
                            float x = im[v];
                            int in = 1;
                            if (mask_f && weighted)
                                in = mask_f[v] > 0.0f && isfinite(mask_f[v]);
                            else if (mask_f)
                                in = mask_f[v] != 0.0f;
                            int m = in && ((double)x > low_value) && isfinite(x);
                            if (!m) continue;
                            ++nm;
                            acc += log((double)x);
                            wsum += (mask_f && weighted) ? (double)mask_f[v] : 1.0;
                        }
                int64_t w = ((int64_t)I * wdims[1] + J) * wdims[2] + K;
                nfull += nm;
                int fg = (nm > 0) && (2 * nm >= nb);
                mask_w[w] = (uint8_t)fg;
                imlog_w[w] = nm > 0 ? acc / (double)nm : 0.0;
                if (conf_w) conf_w[w] = fg ? wsum / (double)nm : 0.0;
                nwork += fg;
            }
    *n_mask_full = nfull;
    *n_mask_work = nwork;
}

void n4o_threshold_log_shrink(const float *im, const float *mask_f,
                              const int32_t fdims[3], double low_value,
                              int32_t s, int32_t wdims[3], double *imlog_w,
                              uint8_t *mask_w, int64_t *n_mask_full,
                              int64_t *n_mask_work) {
    n4o_threshold_log_shrink_w(im, mask_f, 0, fdims, low_value, s, wdims,
                               imlog_w, mask_w, NULL, n_mask_full,
                               n_mask_work);
}

/* N4.py:276-332 */
int32_t n4o_n4(const float *im, const float *mask_f, const int32_t fdims[3],
               const n4o_params *p, double *out_norm, double *out_logbias,
               double *lattice_out, double *cv_trace, int32_t cv_cap,
               n4o_report *rep) {
    if (p->levels < 1 || p->levels > N4O_MAX_LEVELS || p->shrink < 1) return 1;
    for (int d = 0; d < 3; ++d)
        if (p->ncpt[d] < 4 || fdims[d] < 1) return 1;
    if (p->nbins < 2) return 1;
    n4o_report r;
    memset(&r, 0, sizeof(r));

    int32_t W[3];
    for (int d = 0; d < 3; ++d) W[d] = (fdims[d] + p->shrink - 1) / p->shrink;
    const int64_t nw = (int64_t)W[0] * W[1] * W[2];
    double *imlog = malloc(sizeof(double) * nw);
    uint8_t *mask = malloc(nw);
    const int weighted = p->weighted && mask_f;
    double *conf = weighted ? malloc(sizeof(double) * nw) : NULL;
    n4o_threshold_log_shrink_w(im, mask_f, weighted, fdims, p->low_value,
                               p->shrink, W, imlog, mask, conf,
                               &r.n_mask_full, &r.n_mask_work);

    double *logU = malloc(sizeof(double) * nw);
    double *field = calloc((size_t)nw, sizeof(double));
    double *newf = malloc(sizeof(double) * nw);
    double *sharp = malloc(sizeof(double) * nw);
    double *resid = malloc(sizeof(double) * nw);
    memcpy(logU, imlog, sizeof(double) * nw);

    int32_t nc[3] = {p->ncpt[0], p->ncpt[1], p->ncpt[2]};
    int64_t L = (int64_t)nc[0] * nc[1] * nc[2];
    double *lat = calloc((size_t)L, sizeof(double));
    double *lres = malloc(sizeof(double) * L);
    const double wdom[3] = {W[0], W[1], W[2]};
    int32_t ncv = 0;

    for (int32_t lev = 0; lev < p->levels; ++lev) {
        double cv = 1e10;
        int32_t it = 0;
        while (cv > p->conv_threshold && it < p->max_iter[lev]) {
            ++it;
            n4o_sharpen(logU, mask, nw, p->nbins, p->wiener_noise, p->fwhm,
                        sharp, NULL, NULL, NULL);
            for (int64_t v = 0; v < nw; ++v) resid[v] = logU[v] - sharp[v];
            fit_bspline3d_w(resid, mask, conf, W, nc, p->fit_eps, p->nthreads, lres);
            for (int64_t l = 0; l < L; ++l) lat[l] += lres[l];
            n4o_evaluate_bspline3d(lat, nc, W, wdom, p->nthreads, newf);
            cv = n4o_convergence(field, newf, mask, nw);
            if (cv_trace && ncv < cv_cap) cv_trace[ncv] = cv;
            ++ncv;
            double *tmp = field;
            field = newf;
            newf = tmp;
            for (int64_t v = 0; v < nw; ++v) logU[v] = imlog[v] - field[v];
        }
        r.iterations[lev] = it;
        r.last_cv[lev] = cv;
        if (lev != p->levels - 1) {
            int32_t nn[3] = {2 * (nc[0] - 3) + 3, 2 * (nc[1] - 3) + 3, 2 * (nc[2] - 3) + 3};
            int64_t LN = (int64_t)nn[0] * nn[1] * nn[2];
            double *up = malloc(sizeof(double) * LN);
            n4o_upsample_lattice3d(lat, nc, up);
            free(lat);
            free(lres);
            lat = up;
            lres = malloc(sizeof(double) * LN);
            memcpy(nc, nn, sizeof(nc));
            L = LN;
        }
    }

    /* N4.py:329-330: final field at full resolution, normalisation */
    const int64_t nf = (int64_t)fdims[0] * fdims[1] * fdims[2];
    double fdom[3];
    for (int d = 0; d < 3; ++d) fdom[d] = (double)W[d] * (double)p->shrink;
    double *lb = malloc(sizeof(double) * nf);
    n4o_evaluate_bspline3d(lat, nc, fdims, fdom, p->nthreads, lb);
    if (out_logbias) memcpy(out_logbias, lb, sizeof(double) * nf);
    if (out_norm)
        for (int64_t v = 0; v < nf; ++v) {
            double x = (double)im[v];
            if (!((double)im[v] > p->low_value)) x = 0.0;
            out_norm[v] = x / exp(lb[v]);
        }
    if (lattice_out) memcpy(lattice_out, lat, sizeof(double) * L);
    memcpy(r.work_dims, W, sizeof(W));
    memcpy(r.final_ncpt, nc, sizeof(nc));
    if (rep) *rep = r;

    free(lb);
    free(imlog);
    free(mask);
    free(conf);
    free(logU);
    free(field);
    free(newf);
    free(sharp);
    free(resid);
    free(lat);
    free(lres);
    return 0;
}
