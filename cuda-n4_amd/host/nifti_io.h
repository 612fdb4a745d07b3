// nifti_io.h -- minimal NIfTI-1 single-file (.nii / .nii.gz) volume I/O for
// the CLI, in place of the reference's ITK wrappers (itkLoadVolume.h:13-43,
// 78-123, 172-225), which are absent from this image.
//
// Voxels are returned as float32 in file order (i fastest), which is the
// reference's host packing c + r*C + d*C*R (itkLoadVolume.h:116).  The
// 348-byte header of the input is kept and reused for the outputs, so
// orientation and origin survive (the reference keeps spacing only,
// itkLoadVolume.h:193-198).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

struct NiftiVolume {
    int dim[3] = {0, 0, 0};          // ITK size[0..2] = NIfTI dim[1..3]
    float spacing[3] = {1, 1, 1};
    std::vector<unsigned char> header;  // original 348-byte header (native byte order)
    bool foreign_endian = false;        // input was byte-swapped (orientation dropped on write)
    std::vector<float> data;
};

// Reads header only (dims, spacing).  Throws std::runtime_error on failure.
NiftiVolume nifti_read_header(const std::string& path);
// Reads header and voxels (any common datatype, scaled by scl_slope/inter).
NiftiVolume nifti_read(const std::string& path);
// Writes float32 voxels with `like`'s header (dims from `like`); a path
// ending in .gz is compressed.
void nifti_write(const std::string& path, const NiftiVolume& like, const float* data);
bool file_exists(const std::string& path);
