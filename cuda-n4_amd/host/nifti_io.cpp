// nifti_io.cpp -- NIfTI-1 single-file reader/writer (see nifti_io.h).
#include "nifti_io.h"

#include <string.h>
#include <sys/stat.h>
#include <zlib.h>

#include <stdexcept>

namespace {

// NIfTI-1 header field offsets (bytes)
constexpr int kSizeofHdr = 0;
constexpr int kDim = 40;        // int16[8]
constexpr int kDatatype = 70;   // int16
constexpr int kBitpix = 72;     // int16
constexpr int kPixdim = 76;     // float[8]
constexpr int kVoxOffset = 108; // float
constexpr int kSclSlope = 112;  // float
constexpr int kSclInter = 116;  // float
constexpr int kQformCode = 252; // int16
constexpr int kSformCode = 254; // int16
constexpr int kMagic = 344;     // char[4]
constexpr int kHdrSize = 348;

template <class T>
T get(const unsigned char* h, int off) {
    T v;
    memcpy(&v, h + off, sizeof v);
    return v;
}

template <class T>
void put(unsigned char* h, int off, T v) {
    memcpy(h + off, &v, sizeof v);
}

void swap_bytes(void* p, int n) {
    unsigned char* b = (unsigned char*)p;
    for (int i = 0; i < n / 2; ++i) std::swap(b[i], b[n - 1 - i]);
}

// Byte-swaps every multi-byte header field we use (foreign-endian files).
void swap_header(unsigned char* h) {
    swap_bytes(h + kSizeofHdr, 4);
    for (int i = 0; i < 8; ++i) swap_bytes(h + kDim + 2 * i, 2);
    swap_bytes(h + kDatatype, 2);
    swap_bytes(h + kBitpix, 2);
    for (int i = 0; i < 8; ++i) swap_bytes(h + kPixdim + 4 * i, 4);
    swap_bytes(h + kVoxOffset, 4);
    swap_bytes(h + kSclSlope, 4);
    swap_bytes(h + kSclInter, 4);
}

struct GzFile {
    gzFile f = nullptr;
    explicit GzFile(const std::string& p, const char* mode) : f(gzopen(p.c_str(), mode)) {
        if (!f) throw std::runtime_error("cannot open '" + p + "'");
    }
    ~GzFile() {
        if (f) gzclose(f);
    }
    void read(void* dst, size_t n, const std::string& p) {
        unsigned char* d = (unsigned char*)dst;
        while (n > 0) {
            const unsigned chunk = n > (1u << 30) ? (1u << 30) : (unsigned)n;
            const int r = gzread(f, d, chunk);
            if (r <= 0) throw std::runtime_error("truncated NIfTI file '" + p + "'");
            d += r;
            n -= (size_t)r;
        }
    }
    void write(const void* src, size_t n, const std::string& p) {
        const unsigned char* s = (const unsigned char*)src;
        while (n > 0) {
            const unsigned chunk = n > (1u << 30) ? (1u << 30) : (unsigned)n;
            const int w = gzwrite(f, s, chunk);
            if (w <= 0) throw std::runtime_error("write failed for '" + p + "'");
            s += w;
            n -= (size_t)w;
        }
    }
};

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

NiftiVolume parse_header(GzFile& g, const std::string& path, bool* swapped) {
    NiftiVolume v;
    v.header.resize(kHdrSize);
    g.read(v.header.data(), kHdrSize, path);
    unsigned char* h = v.header.data();
    *swapped = false;
    if (get<int32_t>(h, kSizeofHdr) != kHdrSize) {
        swap_header(h);
        *swapped = true;
        if (get<int32_t>(h, kSizeofHdr) != kHdrSize) throw std::runtime_error("'" + path + "' is not a NIfTI-1 file");
        v.foreign_endian = true;
    }
    if (memcmp(h + kMagic, "n+1", 3) != 0)
        throw std::runtime_error("'" + path + "': only single-file NIfTI-1 (.nii/.nii.gz) is supported");
    const int nd = get<int16_t>(h, kDim);
    if (nd < 1 || nd > 7) throw std::runtime_error("'" + path + "': bad dim[0]");
    for (int d = 0; d < 3; ++d) {
        const int n = d < nd ? get<int16_t>(h, kDim + 2 * (d + 1)) : 1;
        v.dim[d] = n < 1 ? 1 : n;
        v.spacing[d] = get<float>(h, kPixdim + 4 * (d + 1));
    }
    for (int d = 3; d < nd; ++d)
        if (get<int16_t>(h, kDim + 2 * (d + 1)) > 1)
            throw std::runtime_error("'" + path + "': only 3-D volumes are supported");
    return v;
}

}  // namespace

bool file_exists(const std::string& path) {
    struct stat st;
    return stat(path.c_str(), &st) == 0;
}

NiftiVolume nifti_read_header(const std::string& path) {
    GzFile g(path, "rb");
    bool sw;
    return parse_header(g, path, &sw);
}

NiftiVolume nifti_read(const std::string& path) {
    GzFile g(path, "rb");
    bool sw;
    NiftiVolume v = parse_header(g, path, &sw);
    const unsigned char* h = v.header.data();
    const size_t off = (size_t)get<float>(h, kVoxOffset);
    if (off > kHdrSize) {
        std::vector<unsigned char> skip(off - kHdrSize);
        g.read(skip.data(), skip.size(), path);
    }
    const size_t n = (size_t)v.dim[0] * v.dim[1] * v.dim[2];
    const int dt = get<int16_t>(h, kDatatype);
    int bytes;
    switch (dt) {
        case 2: case 256: bytes = 1; break;    // uint8, int8
        case 4: case 512: bytes = 2; break;    // int16, uint16
        case 8: case 768: case 16: bytes = 4; break;  // int32, uint32, float32
        case 64: bytes = 8; break;             // float64
        default: throw std::runtime_error("'" + path + "': unsupported NIfTI datatype " + std::to_string(dt));
    }
    std::vector<unsigned char> raw(n * bytes);
    g.read(raw.data(), raw.size(), path);
    v.data.resize(n);
    float slope = get<float>(h, kSclSlope), inter = get<float>(h, kSclInter);
    const bool scale = slope != 0.0f && !(slope == 1.0f && inter == 0.0f);
    for (size_t i = 0; i < n; ++i) {
        unsigned char* p = raw.data() + i * bytes;
        if (sw && bytes > 1) swap_bytes(p, bytes);
        double x;
        switch (dt) {
            case 2: x = *(uint8_t*)p; break;
            case 256: x = *(int8_t*)p; break;
            case 4: x = get<int16_t>(p, 0); break;
            case 512: x = get<uint16_t>(p, 0); break;
            case 8: x = get<int32_t>(p, 0); break;
            case 768: x = get<uint32_t>(p, 0); break;
            case 16: x = get<float>(p, 0); break;
            default: x = get<double>(p, 0); break;
        }
        v.data[i] = (float)(scale ? x * slope + inter : x);
    }
    return v;
}

void nifti_write(const std::string& path, const NiftiVolume& like, const float* data) {
    std::vector<unsigned char> h = like.header;
    if (h.size() != (size_t)kHdrSize) throw std::runtime_error("nifti_write: template header missing");
    put<int16_t>(h.data(), kDatatype, 16);
    put<int16_t>(h.data(), kBitpix, 32);
    put<float>(h.data(), kVoxOffset, 352.0f);
    put<float>(h.data(), kSclSlope, 1.0f);
    put<float>(h.data(), kSclInter, 0.0f);
    if (like.foreign_endian) {  // only the fields parsed above were swapped
        put<int16_t>(h.data(), kQformCode, 0);
        put<int16_t>(h.data(), kSformCode, 0);
    }
    memcpy(h.data() + kMagic, "n+1\0", 4);
    const size_t n = (size_t)like.dim[0] * like.dim[1] * like.dim[2];
    const char* mode = ends_with(path, ".gz") ? "wb6" : "wbT";
    GzFile g(path, mode);
    g.write(h.data(), h.size(), path);
    const unsigned char ext[4] = {0, 0, 0, 0};
    g.write(ext, 4, path);
    g.write(data, n * sizeof(float), path);
}
