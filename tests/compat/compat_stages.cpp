// compat_stages.cpp -- test driver of the reference's C++ API on the MI355X
// library: a program written against include/cudaN4.h and include/reducer.h
// (the reference's cudaN4.h:198-387 and reducer.h:3-12) that runs every
// stage function on inputs the test wrote and writes the outputs back, for
// the test to compare with the oracle.
//
// Usage: compat_stages DIR
//   DIR/spec.txt   X Y Z CX CY CZ NBINS LEVELS ITERS
//   DIR/*.f32      raw float32 inputs, C order [x][y][z]:
//                    logim (X*Y*Z)   masked log image for sharpenImage
//                    mask  (X*Y*Z)   float mask, nonzero = foreground
//                    field (X*Y*Z)   field for FitBspline3D
//                    lat   (CX*CY*CZ) lattice for EvaluateBspline3D and the upsample
//                    f1, f2 (X*Y*Z)  fields for the convergence measurement and Reducer
//                    im    (X*Y*Z)   intensities for testsharpenImage and testN4
// Writes DIR/out_*.f32 and DIR/values.txt ("name value" lines).
#include <stdio.h>
#include <stdlib.h>

#include <fstream>
#include <string>
#include <vector>

#include "cudaN4.h"
#include "n4_capi.h"
#include "reducer.h"

namespace {

std::string g_dir;

std::vector<float> load(const char* name, size_t n) {
    std::vector<float> v(n);
    FILE* f = fopen((g_dir + "/" + name + ".f32").c_str(), "rb");
    if (!f || fread(v.data(), sizeof(float), n, f) != n) throw std::runtime_error(std::string("cannot read ") + name);
    fclose(f);
    return v;
}

void save(const char* name, const float* p, size_t n) {
    FILE* f = fopen((g_dir + "/out_" + name + ".f32").c_str(), "wb");
    if (!f || fwrite(p, sizeof(float), n, f) != n) throw std::runtime_error(std::string("cannot write ") + name);
    fclose(f);
}

void ok(n4_status s, const char* what) {
    if (s != N4_OK) throw std::runtime_error(std::string(what) + ": " + n4_last_error());
}

// Device copy of a host array (owned by the handle until freed).
struct Dev {
    n4_handle* h;
    float* p = nullptr;
    size_t n;
    Dev(n4_handle* h_, size_t n_) : h(h_), n(n_) { ok(n4_malloc(h, (void**)&p, n * sizeof(float)), "n4_malloc"); }
    Dev(n4_handle* h_, const std::vector<float>& v) : Dev(h_, v.size()) {
        ok(n4_memcpy_htod(h, p, v.data(), n * sizeof(float)), "upload");
    }
    ~Dev() { n4_free(h, p); }
    std::vector<float> get() const {
        std::vector<float> v(n);
        ok(n4_memcpy_dtoh(h, v.data(), p, n * sizeof(float)), "download");
        return v;
    }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: compat_stages DIR\n");
        return 2;
    }
    g_dir = argv[1];
    try {
        unsigned X, Y, Z, CX, CY, CZ, NB, LV, IT;
        std::ifstream sp(g_dir + "/spec.txt");
        if (!(sp >> X >> Y >> Z >> CX >> CY >> CZ >> NB >> LV >> IT)) throw std::runtime_error("bad spec.txt");
        const size_t n = (size_t)X * Y * Z, L = (size_t)CX * CY * CZ;
        const size_t L2 = (size_t)(2 * CX - 3) * (2 * CY - 3) * (2 * CZ - 3);
        const std::vector<float> logim = load("logim", n), mask = load("mask", n), field = load("field", n),
                                 lat = load("lat", L), f1 = load("f1", n), f2 = load("f2", n), im = load("im", n);
        n4_handle* h = nullptr;
        ok(n4_create(&h, 0), "n4_create");
        std::ofstream vals(g_dir + "/values.txt");
        vals.precision(17);
        N4Param param;
        param.NumberOfHistogramBins = NB;
        N4Data data;
        data.assign(param.NumberOfHistogramBins, param.paddedHistogramSize);
        {
            // device-pointer stage functions
            Dev dx(h, logim), dm(h, mask), dsh(h, n), dfield(h, field), dfit(h, L), dlat(h, lat), dev(h, n),
                dup(h, L2), d1(h, f1), d2(h, f2);
            sharpenImage(dx.p, dm.p, (unsigned)n, NB, param.paddedHistogramSize, param.histogramOffset,
                         param.WienerFilterNoise, param.BiasFieldFullWidthAtHalfMaximum, data, dsh.p);
            save("sharpen", dsh.get().data(), n);
            FitBspline3D(dfield.p, dm.p, CX, CY, CZ, X, Y, Z, dfit.p);
            save("fit", dfit.get().data(), L);
            EvaluateBspline3D(dlat.p, CX, CY, CZ, X, Y, Z, dev.p);
            save("evaluate", dev.get().data(), n);
            upsampleLattice3D_gpu(dlat.p, CX, CY, CZ, dup.p);
            save("upsample", dup.get().data(), L2);
            float cv = 0.0f;
            calculateConvergenceMeasurement(d1.p, d2.p, dm.p, (unsigned)n, 0.0f, cv);
            vals << "convergence " << cv << "\n";
            vals << "reduce_sum " << Reducer::reduce_sum_wrapper((int)n, d1.p, nullptr) << "\n";
            vals << "reduce_max " << Reducer::reduce_max_wrapper((int)n, d1.p, nullptr) << "\n";
            vals << "reduce_min " << Reducer::reduce_min_wrapper((int)n, d1.p, nullptr) << "\n";
        }
        {
            // host-pointer test* variants
            std::vector<float> o(std::max(n, L2));
            upsampleLattice3D_cpu(lat.data(), CX, CY, CZ, o.data());
            save("upsample_cpu", o.data(), L2);
            testUpsampleLattice3D_gpu(lat.data(), CX, CY, CZ, o.data());
            save("upsample_host", o.data(), L2);
            testEvaluateBspline3D(lat.data(), CX, CY, CZ, X, Y, Z, o.data());
            save("evaluate_host", o.data(), n);
            testFitBspline3D(field.data(), mask.data(), CX, CY, CZ, X, Y, Z, o.data());
            save("fit_host", o.data(), L);
            std::vector<float> lg(n), sh(n);
            float bmin = 0.0f, bmax = 0.0f;
            testsharpenImage(im.data(), mask.data(), (unsigned)n, param, data, lg.data(), sh.data(), bmin, bmax);
            save("tsharpen_log", lg.data(), n);
            save("tsharpen", sh.data(), n);
            vals << "bin_min " << bmin << "\nbin_max " << bmax << "\n";
        }
        {
            // the whole correction through testN4 and its run records
            N4Param p4;
            p4.NumberOfFittingLevels = LV;
            p4.MaximumNumberOfIterations = IT;
            p4.NumberOfControlPoints_x = 4;
            p4.NumberOfControlPoints_y = 5;
            p4.NumberOfControlPoints_z = 6;
            const size_t Lf = (size_t)(((4 - 3) << (LV - 1)) + 3) * (((5 - 3) << (LV - 1)) + 3) * (((6 - 3) << (LV - 1)) + 3);
            std::vector<float> norm(n), lt(Lf), bias(n);
            std::vector<float> im2 = im, m2 = mask;
            testN4(im2.data(), m2.data(), X, Y, Z, p4, data, norm.data(), lt.data(), bias.data());
            save("n4_norm", norm.data(), n);
            save("n4_bias", bias.data(), n);
            save("n4_lattice", lt.data(), Lf);
            const n4_report& r = N4LastReport();
            vals << "n4_iterations";
            for (unsigned l = 0; l < LV; ++l) vals << " " << r.iterations[l];
            vals << "\nn4_trace";
            for (const auto& lv : N4LastCVTrace(data))
                for (double c : lv) vals << " " << c;
            vals << "\n";
            const std::vector<float> l0 = N4LastLevelLattice(data, 0);
            save("n4_level0", l0.data(), l0.size());
        }
        data.clean();
        n4_destroy(h);
    } catch (const std::exception& e) {
        fprintf(stderr, "compat_stages: %s\n", e.what());
        return 1;
    }
    return 0;
}
