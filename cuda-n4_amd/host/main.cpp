// main.cpp -- CUDAN4 command-line driver on the MI355X library.
//
// Same flags and flow as the reference driver (main.cpp:19-186): parse
// -i/-o/-m/-b/-n/-t/-c into N4Param, read the volumes, testN4() with the
// axis order flipped (fx = slowest file axis, main.cpp:153-155), write the
// outputs.  Extensions (the parameter surface N4.py and N4Param expose but
// the reference CLI does not): -c i,j,k (per file axis, cudaN4.h:60-62), -s
// shrink, --bins, --fwhm, --noise, --threshold, --low, --iters a,b,c (per
// level), --eps, --precision, --fixed-iterations, --weighted-mask,
// --lattice file, --level-lattices prefix, --report file.json (with the
// per-iteration convergence trace).  Volumes are NIfTI-1 (.nii/.nii.gz),
// read and written with the input's header (orientation kept).
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cudaN4.h"
#include "nifti_io.h"

namespace {

int usage() {
    printf("CUDAN4: N4 bias field correction on AMD Instinct MI355X (gfx950).\n");
    printf("Usage:\n  CUDAN4 [options]\n");
    printf("Required options:\n");
    printf("  -i input.nii.gz            : input image\n");
    printf("  -o norm.nii.gz             : bias corrected input image\n");
    printf("Additional options:\n");
    printf("  -m mask.nii.gz             : mask of the region used for the fit\n");
    printf("  -n N                       : maximum iterations per level (default 50)\n");
    printf("  -t levels                  : number of fitting levels (default 4)\n");
    printf("  -c control_points          : control points per axis at the first level (default 4),\n");
    printf("                               or i,j,k per file axis (i fastest)\n");
    printf("  -b bias.nii.gz             : also write the bias field exp(log bias)\n");
    printf("Extensions:\n");
    printf("  -s shrink                  : fit on a grid shrunk by this factor per axis (default 1)\n");
    printf("  --bins N --fwhm F --noise F --threshold F --low F --eps F\n");
    printf("  --iters a,b,c              : maximum iterations per level\n");
    printf("  --precision fp32|bf16|split_bf16 : B-spline contractions (default fp32)\n");
    printf("  --fixed-iterations         : run every level to its maximum (no convergence test)\n");
    printf("  --weighted-mask            : mask values are confidence weights (fit weights, > 0 = foreground)\n");
    printf("  --lattice out.nii.gz       : write the final B-spline lattice\n");
    printf("  --level-lattices prefix    : write each level's lattice to prefix<level>.nii.gz\n");
    printf("  --report out.json          : write a JSON run record (with the convergence trace)\n");
    return -1;
}

struct Args {
    int argc;
    char** argv;
    int i = 1;
    std::string cmd;
    bool at_end() const { return i >= argc; }
    std::string next() {
        if (i >= argc) throw std::runtime_error("Unexpected end of command line arguments.");
        return argv[i++];
    }
    std::string command() {
        cmd = next();
        if (cmd.empty() || cmd[0] != '-')
            throw std::runtime_error("Expected a command at position " + std::to_string(i) + ", instead got '" + cmd + "'.");
        return cmd;
    }
    std::string existing_file() {
        std::string f = next();
        if (!file_exists(f)) throw std::runtime_error("File '" + f + "' does not exist");
        return f;
    }
    long integer() {
        std::string a = next();
        errno = 0;
        char* end = nullptr;
        long v = strtol(a.c_str(), &end, 10);
        if (errno || *end) throw std::runtime_error("Expected an integer as parameter to '" + cmd + "', instead got '" + a + "'");
        return v;
    }
    double real() {
        std::string a = next();
        errno = 0;
        char* end = nullptr;
        double v = strtod(a.c_str(), &end);
        if (errno || *end) throw std::runtime_error("Expected a number as parameter to '" + cmd + "', instead got '" + a + "'");
        return v;
    }
};

// Comma-separated unsigned integers ("4" or "4,5,6").
std::vector<unsigned> uint_list(const std::string& s, const std::string& cmd) {
    std::vector<unsigned> v;
    size_t pos = 0;
    while (pos <= s.size()) {
        size_t e = s.find(',', pos);
        if (e == std::string::npos) e = s.size();
        const std::string t = s.substr(pos, e - pos);
        char* end = nullptr;
        errno = 0;
        const unsigned long x = strtoul(t.c_str(), &end, 10);
        if (t.empty() || errno || *end)
            throw std::runtime_error("Expected integers separated by commas as parameter to '" + cmd +
                                     "', instead got '" + s + "'");
        v.push_back((unsigned)x);
        pos = e + 1;
    }
    return v;
}

void write_report(const std::string& path, const N4Param& p, const n4_report& r, unsigned levels,
                  const std::vector<std::vector<double>>& trace) {
    std::ofstream f(path);
    f.precision(17);
    f << "{\"iterations\": [";
    for (unsigned l = 0; l < levels; ++l) f << (l ? ", " : "") << r.iterations[l];
    f << "], \"last_convergence\": [";
    for (unsigned l = 0; l < levels; ++l) f << (l ? ", " : "") << r.last_convergence[l];
    f << "], \"convergence_trace\": [";
    for (size_t l = 0; l < trace.size(); ++l) {
        f << (l ? ", " : "") << "[";
        for (size_t i = 0; i < trace[l].size(); ++i) f << (i ? ", " : "") << trace[l][i];
        f << "]";
    }
    f << "], \"n_mask_full\": " << r.n_mask_full << ", \"n_mask_work\": " << r.n_mask_work
      << ", \"work_dims\": [" << r.work_dims[0] << ", " << r.work_dims[1] << ", " << r.work_dims[2] << "]"
      << ", \"final_ncpt\": [" << r.final_ncpt[0] << ", " << r.final_ncpt[1] << ", " << r.final_ncpt[2] << "]"
      << ", \"elapsed_ms\": " << r.elapsed_ms << ", \"shrink\": " << p.ShrinkFactor
      << ", \"ncpt\": [" << p.NumberOfControlPoints_x << ", " << p.NumberOfControlPoints_y << ", "
      << p.NumberOfControlPoints_z << "], \"fixed_iterations\": " << (p.FixedIterations ? "true" : "false")
      << ", \"weighted_mask\": " << (p.WeightedMask ? "true" : "false") << "}\n";
}

// A lattice [cx][cy][cz] (cz fastest) as a NIfTI volume of dim (cz, cy, cx)
// with the image's header otherwise.
void write_lattice(const std::string& file, const NiftiVolume& im, const std::vector<float>& lat, const uint32_t nc[3]) {
    NiftiVolume lv = im;
    lv.dim[0] = (int)nc[2];
    lv.dim[1] = (int)nc[1];
    lv.dim[2] = (int)nc[0];
    int16_t d16[4] = {3, (int16_t)nc[2], (int16_t)nc[1], (int16_t)nc[0]};
    memcpy(lv.header.data() + 40, d16, sizeof d16);
    nifti_write(file, lv, lat.data());
}

int run(int argc, char** argv) {
    if (argc < 2) return usage();
    N4Param param;
    std::string lattice_file, report_file, level_prefix;
    Args cl{argc, argv};
    while (!cl.at_end()) {
        const std::string arg = cl.command();
        if (arg == "-i") {
            param.inputImFile = cl.existing_file();
        } else if (arg == "-o") {
            param.outputImFile = cl.next();
        } else if (arg == "-m") {
            param.maskFile = cl.existing_file();
        } else if (arg == "-b") {
            param.outBiasFieldFile = cl.next();
        } else if (arg == "-n") {
            param.MaximumNumberOfIterations = (unsigned)cl.integer();
        } else if (arg == "-t") {
            param.NumberOfFittingLevels = (unsigned)cl.integer();
        } else if (arg == "-c") {
            // "N" (every axis, the reference's form) or "i,j,k" in the file's
            // axis order; the library's x is the file's slowest axis k
            // (main.cpp:153-155), so x = k, y = j, z = i
            const std::vector<unsigned> c = uint_list(cl.next(), arg);
            if (c.size() != 1 && c.size() != 3)
                throw std::runtime_error("Expected N or i,j,k as parameter to '-c'");
            param.NumberOfControlPoints = c[0];
            param.NumberOfControlPoints_x = c[c.size() == 3 ? 2 : 0];
            param.NumberOfControlPoints_y = c[c.size() == 3 ? 1 : 0];
            param.NumberOfControlPoints_z = c[0];
        } else if (arg == "-s") {
            param.ShrinkFactor = (unsigned)cl.integer();
        } else if (arg == "--bins") {
            param.NumberOfHistogramBins = (unsigned)cl.integer();
        } else if (arg == "--fwhm") {
            param.BiasFieldFullWidthAtHalfMaximum = (float)cl.real();
        } else if (arg == "--noise") {
            param.WienerFilterNoise = (float)cl.real();
        } else if (arg == "--threshold") {
            param.ConvergenceThreshold = (float)cl.real();
        } else if (arg == "--low") {
            param.low_value = (float)cl.real();
        } else if (arg == "--eps") {
            param.FitEps = cl.real();
        } else if (arg == "--iters") {
            param.IterationsPerLevel = uint_list(cl.next(), arg);
        } else if (arg == "--fixed-iterations") {
            param.FixedIterations = true;
        } else if (arg == "--weighted-mask") {
            param.WeightedMask = true;
        } else if (arg == "--level-lattices") {
            level_prefix = cl.next();
        } else if (arg == "--precision") {
            const std::string v = cl.next();
            if (v == "fp32")
                param.Precision = N4_PREC_FP32;
            else if (v == "bf16")
                param.Precision = N4_PREC_BF16;
            else if (v == "split_bf16")
                param.Precision = N4_PREC_SPLIT_BF16;
            else
                throw std::runtime_error("Expected fp32, bf16 or split_bf16 as parameter to '--precision', instead got '" +
                                         v + "'");
        } else if (arg == "--lattice") {
            lattice_file = cl.next();
        } else if (arg == "--report") {
            report_file = cl.next();
        } else {
            fprintf(stderr, "Unknown option %s\n", arg.c_str());
            return -1;
        }
    }
    if (param.inputImFile.empty() || param.outputImFile.empty()) return usage();

    NiftiVolume im = nifti_read(param.inputImFile);
    const int* size = im.dim;
    NiftiVolume mask;
    if (!param.maskFile.empty()) {
        NiftiVolume mh = nifti_read_header(param.maskFile);
        if (mh.dim[0] != size[0] || mh.dim[1] != size[1] || mh.dim[2] != size[2]) {
            printf("Size mismatch. im:(%d, %d, %d), mask:(%d, %d, %d) \n", size[0], size[1], size[2], mh.dim[0],
                   mh.dim[1], mh.dim[2]);
            return -1;
        }
        mask = nifti_read(param.maskFile);
    } else {
        mask.data.assign(im.data.size(), 1.0f);
    }
    const size_t n = im.data.size();
    n4_params cp = N4ParamToC(param);
    uint32_t nc[3];
    if (n4_final_ncpt(&cp, nc) != N4_OK) throw std::runtime_error(n4_last_error());
    std::vector<float> norm(n), lattice((size_t)nc[0] * nc[1] * nc[2]), bias;
    if (!param.outBiasFieldFile.empty()) bias.resize(n);

    N4Data data;
    data.assign(param.NumberOfHistogramBins, param.paddedHistogramSize);
    testN4(im.data.data(), mask.data.data(), (unsigned)size[2], (unsigned)size[1], (unsigned)size[0], param, data,
           norm.data(), lattice.data(), bias.empty() ? nullptr : bias.data());
    const n4_report rep = N4LastReport();
    nifti_write(param.outputImFile, im, norm.data());
    if (!bias.empty()) nifti_write(param.outBiasFieldFile, im, bias.data());
    if (!lattice_file.empty()) write_lattice(lattice_file, im, lattice, nc);
    if (!level_prefix.empty())
        for (unsigned l = 0; l < param.NumberOfFittingLevels; ++l) {
            const uint32_t c[3] = {((nc[0] - 3) >> (param.NumberOfFittingLevels - 1 - l)) + 3,
                                   ((nc[1] - 3) >> (param.NumberOfFittingLevels - 1 - l)) + 3,
                                   ((nc[2] - 3) >> (param.NumberOfFittingLevels - 1 - l)) + 3};
            write_lattice(level_prefix + std::to_string(l) + ".nii.gz", im, N4LastLevelLattice(data, l), c);
        }
    if (!report_file.empty())
        write_report(report_file, param, rep, param.NumberOfFittingLevels, N4LastCVTrace(data));
    data.clean();
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        fprintf(stderr, "CUDAN4: %s\n", e.what());
        return -1;
    }
}
