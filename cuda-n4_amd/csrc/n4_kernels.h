// n4_kernels.h -- launch interface of the gfx950 N4 kernels.
//
// Grid-wide kernels run one workgroup of kBlock (256) threads per "slab":
// one x index of one volume's working (or full-resolution) grid.  A batch
// of volumes of different sizes is one launch: the slab list names
// (volume, x) pairs across all volumes.  Per-volume kernels (level setup,
// Wiener step, lattice update) run one workgroup per volume (or per volume
// and control-point plane).  No kernel communicates across its own
// workgroups except through integer atomics: every cross-slab reduction is
// written as per-slab partials and reduced in a fixed order by a later
// kernel, so results are bitwise reproducible.
//
// One correction step = hist -> wiener -> fit -> eval, each volume at its
// own level and iteration (see VolState).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "n4_common.h"

namespace n4 {

// LDS bytes each kernel needs (host-side sizing, same carve as the device).
size_t lds_prep();
size_t lds_level();
size_t lds_hist(int nbins);
size_t lds_wiener(int P);
// prec (RunParams::prec) selects the FP32 or the MFMA kernels.
size_t lds_fit(int nbins, const VolDesc& v, int level, int prec);
size_t lds_lattice();
size_t lds_eval(const VolDesc& v, int level, int prec);
size_t lds_recon(const VolDesc& v);
// MFMA tiles (16 control points each) for c control points along an axis.
int mfma_tiles(int c);
size_t lds_flat(int nbins);

// `stream` is a hipStream_t (void* keeps this header HIP-free).  Return 0
// on success, else a hipError_t value.
// Prep kernels are instantiated per shrink class: 1..4, and 0 for any
// other factor (the class-0 kernel handles every factor).  All slabs of one
// launch_prep call must belong to volumes of class prep_class(shrink).
inline int prep_class(int shrink) { return (shrink >= 1 && shrink <= 4) ? shrink : 0; }
int launch_prep(const VolDesc* vols, const Slab* slabs, int nslab, int shrink, RunParams rp, void* stream);
// level 0: start of a correction (counts, cleared accumulators, state
// TRANS(0)); level > 0: refinement of level-1's lattice (stage call).
int launch_level(const VolDesc* vols, int nvol, int level, RunParams rp, void* stream);
// One correction step is hist -> wiener -> fit(kFitStep) -> eval; each
// volume's step kind comes from its own state, and the launches of every
// step are identical (a chunk of steps can be captured once and replayed).
// step >= 0: a correction step (its pass count, kept in the state, is
// published as progress); step < 0 (stage calls) forces a histogram.
int launch_hist(const VolDesc* vols, const Slab* slabs, int nslab, int step, RunParams rp, int32_t* dbg_idx,
                size_t lds, void* stream);
int launch_wiener(const VolDesc* vols, int nvol, RunParams rp, size_t lds, void* stream);
// mt, nt: MFMA tiles along y and z of the level's lattice (used when
// rp.prec != 0; the MFMA kernels exist for 1..4 tiles per axis).
// it: iteration within the level (selects the numerator buffer by parity).
int launch_fit(const VolDesc* vols, const Slab* slabs, int nslab, int level, int it, int mode, RunParams rp,
               size_t lds, int mt, int nt, void* stream);
int launch_lattice(const VolDesc* vols, int nvol, int cx_max, int level, int mode, RunParams rp, void* stream);
// Lattice update of a step (numerator copies -> lattice increment or level
// denominator); chunks = ceil(largest level size / kBlock).
int launch_update(const VolDesc* vols, int nvol, int chunks, RunParams rp, void* stream);
int launch_eval(const VolDesc* vols, const Slab* slabs, int nslab, RunParams rp, size_t lds, int mt, int nt,
                void* stream);
// The final full-resolution reconstruction: FP32 in every precision mode.
int launch_recon(const VolDesc* vols, const Slab* slabs, int nslab, int level, RunParams rp, size_t lds,
                 void* stream);

// Persistent path: the level/iteration loop of every volume in one launch
// of nwg workgroups (work[b]: team, rank and tiles of block b; a volume's
// team members are consecutive blocks).  Every workgroup must be resident
// at once: nwg <= per_cu * cus from persist_capacity for the same LDS size.
// prec (RunParams::prec): FP32 fit, or the MFMA fit for up to 16 control
// points along y and z (persist_eligible).
size_t lds_persist(const VolDesc& v, int nbins, int P, int prec);
int persist_capacity(size_t lds, int prec, int* per_cu, int* cus);
int launch_persist(const VolDesc* vols, const Slab* slabs, const PWork* work, int nwg, RunParams rp, size_t lds,
                   void* stream);

// Flat (1-D) stage helpers used by the stage entry points.  A flat array of
// n voxels is cut into slabs of kChunkVox voxels; each slab writes
// spart[x] = 4 doubles (reduced on the host by the stage call).
enum FlatOp : int32_t { kFlatMinMax = 0, kFlatSum = 1, kFlatCV = 2, kFlatRemap = 3 };
int launch_flat(const VolDesc* vols, const Slab* slabs, int nslab, int op, const float* a, const float* b,
                float* out, RunParams rp, void* stream);
// out[i] = in[i] != 0 (a float mask as the uint8 mask of the stage calls);
// vector loads where `in` is 16-byte and `out` 4-byte aligned.
int launch_mask_u8(const float* in, uint8_t* out, uint64_t n, void* stream);

}  // namespace n4
