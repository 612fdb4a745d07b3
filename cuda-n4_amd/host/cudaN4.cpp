// cudaN4.cpp -- the reference's C++ API (include/cudaN4.h, include/reducer.h)
// on top of the MI355X C ABI (include/n4_capi.h).  Host-only code: every
// computation runs in libn4amd.so; this file converts arguments, owns
// temporaries and turns status codes into exceptions.
#include "cudaN4.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "reducer.h"

namespace {

void check(n4_status s, const char* what) {
    if (s != N4_OK)
        throw n4_error(std::string(what) + ": " + n4_status_string(s) + ": " + n4_last_error());
}

// Handle for calls that carry no N4Data (stage functions, Reducer): one per
// thread, created on first use, released at thread exit.
struct ThreadHandle {
    n4_handle* h = nullptr;
    ~ThreadHandle() {
        if (h) n4_destroy(h);
    }
};

n4_handle* thread_handle() {
    static thread_local ThreadHandle th;
    if (!th.h) check(n4_create(&th.h, 0), "n4_create");
    return th.h;
}

thread_local n4_report g_last_report;
thread_local unsigned g_last_levels = 0;  // fitting levels of that call

// Device buffer owned for the duration of a call.
struct DevBuf {
    n4_handle* h;
    void* p = nullptr;
    DevBuf(n4_handle* h_, size_t bytes) : h(h_) { check(n4_malloc(h, &p, bytes), "n4_malloc"); }
    ~DevBuf() {
        if (p) n4_free(h, p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// float mask (nonzero = foreground, cudaN4.h:346) -> uint8 mask, on the device.
void mask_to_u8(n4_handle* h, const float* d_mask, size_t n, void* d_out) {
    check(n4_mask_to_u8(h, d_mask, n, (uint8_t*)d_out), "n4_mask_to_u8");
}

double widen(float f) {
    char buf[32];
    snprintf(buf, sizeof buf, "%.7g", (double)f);
    return strtod(buf, nullptr);
}

}  // namespace

n4_params N4ParamToC(const N4Param& q) {
    n4_params p;
    check(n4_params_default(&p), "n4_params_default");
    p.number_of_histogram_bins = q.NumberOfHistogramBins;
    p.wiener_filter_noise = widen(q.WienerFilterNoise);
    p.bias_field_fwhm = widen(q.BiasFieldFullWidthAtHalfMaximum);
    p.ncpt[0] = q.NumberOfControlPoints_x;
    p.ncpt[1] = q.NumberOfControlPoints_y;
    p.ncpt[2] = q.NumberOfControlPoints_z;
    p.number_of_fitting_levels = q.NumberOfFittingLevels;
    for (unsigned l = 0; l < N4_MAX_LEVELS; ++l)
        p.max_iterations[l] = q.IterationsPerLevel.empty()
                                  ? q.MaximumNumberOfIterations
                                  : q.IterationsPerLevel[std::min<size_t>(l, q.IterationsPerLevel.size() - 1)];
    p.convergence_threshold = widen(q.ConvergenceThreshold);
    p.low_value = widen(q.low_value);
    p.shrink_factor = q.ShrinkFactor;
    p.fit_eps = q.FitEps;
    p.precision = q.Precision;
    p.flags = (q.FixedIterations ? N4_FLAG_FIXED_ITERATIONS : 0u) | (q.WeightedMask ? N4_FLAG_WEIGHTED_MASK : 0u);
    return p;
}

void N4Data::assign(const unsigned int, const unsigned int) {
    if (!handle) check(n4_create(&handle, device), "n4_create");
}

void N4Data::clean() {
    if (handle) n4_destroy(handle);
    handle = nullptr;
}

static n4_handle* data_handle(N4Data& data) {
    if (!data.handle) data.assign(0, 0);
    return data.handle;
}

void sharpenImage(float* d_before_sharpen, const float* d_mask, const unsigned int n,
                  const unsigned int NumberOfHistogramBins, const unsigned int, const unsigned int,
                  const float WienerFilterNoise, const float BiasFieldFullWidthAtHalfMaximum, N4Data& data,
                  float* d_after_sharpen) {
    n4_handle* h = data_handle(data);
    DevBuf m(h, n);
    mask_to_u8(h, d_mask, n, m.p);
    check(n4_sharpen_image(h, d_before_sharpen, (const uint8_t*)m.p, n, NumberOfHistogramBins,
                           widen(WienerFilterNoise), widen(BiasFieldFullWidthAtHalfMaximum), d_after_sharpen,
                           nullptr),
          "n4_sharpen_image");
}

void testsharpenImage(const float* h_before_sharpen, const float* h_mask, const unsigned int n, N4Param& param,
                      N4Data& data, float* h_before_sharpen_log, float* h_after_sharpen, float& binMin_out,
                      float& binMax_out) {
    // cudaN4.cu:842-937: threshold (in a copy), masked log, then sharpen
    n4_handle* h = data_handle(data);
    const double low = widen(param.low_value);
    std::vector<float> lg(n);
    std::vector<uint8_t> m(n);
    for (unsigned i = 0; i < n; ++i) {
        const bool fg = h_mask[i] != 0.0f && (double)h_before_sharpen[i] > low;
        m[i] = fg;
        lg[i] = fg ? logf(h_before_sharpen[i]) : 0.0f;
    }
    if (h_before_sharpen_log) memcpy(h_before_sharpen_log, lg.data(), sizeof(float) * n);
    DevBuf dx(h, sizeof(float) * n), dm(h, n), dout(h, sizeof(float) * n);
    check(n4_memcpy_htod(h, dx.p, lg.data(), sizeof(float) * n), "upload");
    check(n4_memcpy_htod(h, dm.p, m.data(), n), "upload");
    double mms[3] = {0, 0, 0};
    check(n4_histogram(h, (const float*)dx.p, (const uint8_t*)dm.p, n, param.NumberOfHistogramBins, mms, nullptr,
                       nullptr),
          "n4_histogram");
    check(n4_sharpen_image(h, (const float*)dx.p, (const uint8_t*)dm.p, n, param.NumberOfHistogramBins,
                           widen(param.WienerFilterNoise), widen(param.BiasFieldFullWidthAtHalfMaximum),
                           (float*)dout.p, nullptr),
          "n4_sharpen_image");
    check(n4_memcpy_dtoh(h, h_after_sharpen, dout.p, sizeof(float) * n), "download");
    binMin_out = (float)mms[0];
    binMax_out = (float)mms[1];
}

void calculateConvergenceMeasurement(const float* d_field1, const float* d_field2, const float* d_mask,
                                     const unsigned int n, const float, float& convergence) {
    n4_handle* h = thread_handle();
    DevBuf m(h, n);
    mask_to_u8(h, d_mask, n, m.p);
    double cv = 0.0;
    check(n4_convergence_measurement(h, d_field1, d_field2, (const uint8_t*)m.p, n, &cv),
          "n4_convergence_measurement");
    convergence = (float)cv;
}

// cudaN4.cu:995-1084 / N4.py:132-165 on the host, per-axis clip bounds.
void upsampleLattice3D_cpu(const float* lat, const unsigned int cx, const unsigned int cy, const unsigned int cz,
                           float* out) {
    const int n[3] = {(int)cx - 3, (int)cy - 3, (int)cz - 3};
    const int m[3] = {2 * n[0] + 3, 2 * n[1] + 3, 2 * n[2] + 3};
    auto bw = [](int odd, int a) { return odd ? (a == 0 ? 0.0 : 0.5) : (a == 1 ? 0.75 : 0.125); };
    auto clip = [](int v, int hi) { return v < -1 ? -1 : (v > hi ? hi : v); };
    for (int X = 0; X < m[0]; ++X)
        for (int Y = 0; Y < m[1]; ++Y)
            for (int Z = 0; Z < m[2]; ++Z) {
                const int c[3] = {X - 1, Y - 1, Z - 1};
                int par[3], odd[3];
                for (int d = 0; d < 3; ++d) {
                    par[d] = c[d] >= 0 ? c[d] / 2 : -((-c[d] + 1) / 2);
                    odd[d] = c[d] - 2 * par[d];
                }
                double acc = 0.0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int e = 0; e < 3; ++e) {
                            const int i = clip(par[0] - 1 + a, n[0] + 1) + 1;
                            const int j = clip(par[1] - 1 + b, n[1] + 1) + 1;
                            const int k = clip(par[2] - 1 + e, n[2] + 1) + 1;
                            acc += bw(odd[0], a) * bw(odd[1], b) * bw(odd[2], e) *
                                   (double)lat[((size_t)i * cy + j) * cz + k];
                        }
                out[((size_t)X * m[1] + Y) * m[2] + Z] = (float)acc;
            }
}

void upsampleLattice3D_gpu(const float* d_lattice, const unsigned int cx, const unsigned int cy,
                           const unsigned int cz, float* d_out) {
    const uint32_t nc[3] = {cx, cy, cz};
    check(n4_upsample_lattice3d(thread_handle(), d_lattice, nc, d_out), "n4_upsample_lattice3d");
}

void testUpsampleLattice3D_gpu(const float* h_lattice, const unsigned int cx, const unsigned int cy,
                               const unsigned int cz, float* h_out) {
    n4_handle* h = thread_handle();
    const size_t L = (size_t)cx * cy * cz;
    const size_t L2 = (size_t)(2 * cx - 3) * (2 * cy - 3) * (2 * cz - 3);
    DevBuf a(h, L * 4), b(h, L2 * 4);
    check(n4_memcpy_htod(h, a.p, h_lattice, L * 4), "upload");
    upsampleLattice3D_gpu((const float*)a.p, cx, cy, cz, (float*)b.p);
    check(n4_memcpy_dtoh(h, h_out, b.p, L2 * 4), "download");
}

void EvaluateBspline3D(const float* d_lattice, const unsigned int cx, const unsigned int cy, const unsigned int cz,
                       const unsigned int fx, const unsigned int fy, const unsigned int fz, float* d_fitted) {
    const uint32_t nc[3] = {cx, cy, cz}, dims[3] = {fx, fy, fz};
    check(n4_evaluate_bspline3d(thread_handle(), d_lattice, nc, dims, d_fitted), "n4_evaluate_bspline3d");
}

void testEvaluateBspline3D(const float* h_lattice, const unsigned int cx, const unsigned int cy,
                           const unsigned int cz, const unsigned int fx, const unsigned int fy,
                           const unsigned int fz, float* h_fitted) {
    n4_handle* h = thread_handle();
    const size_t L = (size_t)cx * cy * cz, n = (size_t)fx * fy * fz;
    DevBuf a(h, L * 4), b(h, n * 4);
    check(n4_memcpy_htod(h, a.p, h_lattice, L * 4), "upload");
    EvaluateBspline3D((const float*)a.p, cx, cy, cz, fx, fy, fz, (float*)b.p);
    check(n4_memcpy_dtoh(h, h_fitted, b.p, n * 4), "download");
}

void FitBspline3D(const float* d_field, const float* d_mask, const unsigned int cx, const unsigned int cy,
                  const unsigned int cz, const unsigned int fx, const unsigned int fy, const unsigned int fz,
                  float* d_lattice) {
    n4_handle* h = thread_handle();
    const size_t n = (size_t)fx * fy * fz;
    DevBuf m(h, n);
    mask_to_u8(h, d_mask, n, m.p);
    const uint32_t nc[3] = {cx, cy, cz}, dims[3] = {fx, fy, fz};
    check(n4_fit_bspline3d(h, d_field, (const uint8_t*)m.p, nc, dims, 1e-3, d_lattice), "n4_fit_bspline3d");
}

void testFitBspline3D(const float* h_field, const float* h_mask, const unsigned int cx, const unsigned int cy,
                      const unsigned int cz, const unsigned int fx, const unsigned int fy, const unsigned int fz,
                      float* h_lattice) {
    n4_handle* h = thread_handle();
    const size_t n = (size_t)fx * fy * fz, L = (size_t)cx * cy * cz;
    DevBuf f(h, n * 4), m(h, n * 4), o(h, L * 4);
    check(n4_memcpy_htod(h, f.p, h_field, n * 4), "upload");
    check(n4_memcpy_htod(h, m.p, h_mask, n * 4), "upload");
    FitBspline3D((const float*)f.p, (const float*)m.p, cx, cy, cz, fx, fy, fz, (float*)o.p);
    check(n4_memcpy_dtoh(h, h_lattice, o.p, L * 4), "download");
}

void N4(float* d_im, float* d_mask, const unsigned int fx, const unsigned int fy, const unsigned int fz,
        N4Param& param, N4Data& data, float* d_im_normalized, float* d_lattice, float* d_biasField) {
    const n4_params p = N4ParamToC(param);
    check(n4_correct(data_handle(data), d_im, d_mask, d_mask ? N4_MASK_F32 : N4_MASK_NONE, fx, fy, fz, &p,
                     d_im_normalized, d_lattice, d_biasField, nullptr, &g_last_report, nullptr),
          "n4_correct");
    g_last_levels = p.number_of_fitting_levels;
}

void testN4(float* h_im, float* h_mask, const unsigned int fx, const unsigned int fy, const unsigned int fz,
            N4Param& param, N4Data& data, float* h_im_normalized, float* h_lattice, float* h_biasField) {
    const n4_params p = N4ParamToC(param);
    check(n4_correct_host(data_handle(data), h_im, h_mask, fx, fy, fz, &p, h_im_normalized, h_lattice,
                          h_biasField, &g_last_report),
          "n4_correct_host");
    g_last_levels = p.number_of_fitting_levels;
}

const n4_report& N4LastReport() { return g_last_report; }

std::vector<std::vector<double>> N4LastCVTrace(N4Data& data) {
    n4_handle* h = data_handle(data);
    uint32_t n = 0;
    check(n4_last_cv_trace(h, 0, nullptr, 0, &n), "n4_last_cv_trace");
    std::vector<double> all(n);
    if (n) check(n4_last_cv_trace(h, 0, all.data(), n, &n), "n4_last_cv_trace");
    std::vector<std::vector<double>> per;
    size_t k = 0;
    for (unsigned l = 0; l < g_last_levels; ++l) {
        const size_t it = std::min<size_t>(g_last_report.iterations[l], all.size() - k);
        per.emplace_back(all.begin() + k, all.begin() + std::min(all.size(), k + it));
        k += it;
    }
    return per;
}

std::vector<float> N4LastLevelLattice(N4Data& data, unsigned level) {
    // level l of L has (final - 3) / 2^(L-1-l) + 3 control points per axis
    if (level >= g_last_levels) throw n4_error("N4LastLevelLattice: level out of range");
    size_t n = 1;
    for (int d = 0; d < 3; ++d) n *= ((g_last_report.final_ncpt[d] - 3) >> (g_last_levels - 1 - level)) + 3;
    std::vector<float> out(n);
    check(n4_last_level_lattice(data_handle(data), 0, level, out.data(), n), "n4_last_level_lattice");
    return out;
}

float Reducer::reduce_sum_wrapper(int n, float* d_idata, float*) {
    double s = 0.0;
    check(n4_reduce_sum(thread_handle(), (uint64_t)n, d_idata, &s), "n4_reduce_sum");
    return (float)s;
}

float Reducer::reduce_max_wrapper(int n, float* d_idata, float*) {
    float v = 0.0f;
    check(n4_reduce_max(thread_handle(), (uint64_t)n, d_idata, &v), "n4_reduce_max");
    return v;
}

float Reducer::reduce_min_wrapper(int n, float* d_idata, float*) {
    float v = 0.0f;
    check(n4_reduce_min(thread_handle(), (uint64_t)n, d_idata, &v), "n4_reduce_min");
    return v;
}
