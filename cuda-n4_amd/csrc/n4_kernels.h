// n4_kernels.h -- launch interface of the gfx950 N4 kernels.
//
// Every kernel is one workgroup of kBlock (256) threads per "slab": one x
// index of one volume's working (or full-resolution) grid.  A batch of
// volumes of different sizes is one launch: the slab list names
// (volume, x) pairs across all volumes.  Cross-slab reductions are
// deterministic: each slab writes a partial, and the last slab of a volume
// to finish (an agent-scope counter) reduces the partials in slab order and
// runs the volume's tail work (Wiener step, lattice update, convergence).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "n4_common.h"

namespace n4 {

// LDS bytes each kernel needs for one volume (host-side sizing).
size_t lds_prep();
size_t lds_hist(int nbins, int P);
size_t lds_fit(int nbins, const VolDesc& v, int level);
size_t lds_eval(const VolDesc& v, int level);
size_t lds_recon(const VolDesc& v);
size_t lds_flat();

// Launches; `stream` is a hipStream_t (void* keeps this header HIP-free).
// grid = number of slabs (or volumes for the level kernel).
int launch_prep(const VolDesc* vols, const Slab* slabs, int nslab, RunParams rp, void* stream);
int launch_level(const VolDesc* vols, int nvol, int level, RunParams rp, void* stream);
int launch_hist(const VolDesc* vols, const Slab* slabs, int nslab, RunParams rp, int32_t* dbg_idx,
                size_t lds, void* stream);
int launch_fit(const VolDesc* vols, const Slab* slabs, int nslab, int level, int mode, RunParams rp,
               size_t lds, void* stream);
int launch_eval(const VolDesc* vols, const Slab* slabs, int nslab, int level, RunParams rp,
                size_t lds, void* stream);
int launch_recon(const VolDesc* vols, const Slab* slabs, int nslab, int level, RunParams rp,
                 size_t lds, void* stream);

// Flat (1-D) stage helpers used by the stage entry points.  A flat array of
// n voxels is cut into slabs of kChunkVox voxels.
enum FlatOp : int32_t { kFlatMinMax = 0, kFlatSum = 1, kFlatCV = 2, kFlatRemap = 3 };
int launch_flat(const VolDesc* vols, const Slab* slabs, int nslab, int op, const float* a,
                const float* b, float* out, RunParams rp, void* stream);

}  // namespace n4
