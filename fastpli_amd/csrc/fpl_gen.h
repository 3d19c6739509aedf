// Host-side entry of tissue generation shared by the one-shot entry points
// (fpl_generate, fpl_generate_fields) and the serving loop.
#pragma once

#include "fpl_internal.h"

namespace fpl {
namespace gen {

// Where the voxelizer's final pass writes a voxel: the API's C-order fields
// (label[x][y][z], vector[x][y][z][3]) and / or the march's [x][z][y] records
// (fpl_records.h), each optional (at least the labels or the records).
struct Outputs {
   int32_t *label;
   float *vector;
   void *rec_a;
   uint32_t *rec_b;
   int32_t rec_storage;
};

// PliGenerator::RunTissueGeneration (generator.cpp:165-261) on `st`; returns
// after the outputs are complete.  props (host, n_props x 2) receives the
// property list; label_max (may be null) an upper bound of the labels written
// (0: every voxel is 0).  fbs->points / fiber_ptr are device pointers.
Status run(const fpl_volume *vol, const fpl_fiber_bundles *fbs, const Outputs &out,
           double *props, uint32_t *warn, int32_t *label_max, hipStream_t st);

// An upper estimate of the stream workspace `run` takes for a model of
// n_points points, n_fibers fibers and n_bundles bundles in a volume of dim
// voxels (both binnings, so a scene that falls back to the radix sort fits).
size_t workspace_bytes(int64_t n_points, int64_t n_fibers, int64_t n_bundles,
                       const int64_t dim[3]);

} // namespace gen
} // namespace fpl
