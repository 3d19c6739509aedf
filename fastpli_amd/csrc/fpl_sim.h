// Host-side pieces of the ray march shared by the one-shot entry point
// (fpl_simulate) and the serving loop.
#pragma once

#include <vector>

#include "fpl_internal.h"

namespace fpl {
namespace sim {

// The parameters of one simulation laid out as the march kernel reads them
// from device memory: Params, the tilts, the layer properties (256-byte
// aligned parts).  Built on the host from an fpl_setup, the property table and
// the tilt list.
struct Block {
   std::vector<char> bytes;
   size_t off_t = 0, off_pr = 0;
   int64_t dim[3] = {0, 0, 0};
   int64_t n_tilts = 0, n_props = 0;
   int n_rho = 0, storage = 0;
};

// Bytes per voxel of the packed field storage (the angle storage has a
// second 4-byte array).
size_t record_bytes(int storage);

// setup::Setup::Check + CheckInput (setup.hpp:48-71, simulator.cpp:169-175)
// and the per-tilt host constants; fills B.
Status prepare(const fpl_setup *setup, const int64_t dim[3], const double *props, int64_t n_props,
               const double *tilts, int64_t n_tilts, Block &B);

// Resets the warning word and the label min / max pair.
void clear_flags(uint32_t *d_flags, int32_t *d_mm, hipStream_t st);

// Packs label + axis into the march's [x][z][y] records and reduces the label
// range into d_mm (min, max).
void pack(int storage, const int64_t dim[3], const int32_t *label, const float *vector,
          void *packed, uint32_t *packed_b, int32_t *d_mm, hipStream_t st);

// Whether the march holds B's property table in LDS (a kernel variant).
bool props_in_lds(const Block &B);

// All tilts of B in one launch; d_block holds B.bytes on the device.
void march(const Block &B, const void *d_block, const void *packed, const uint32_t *packed_b,
           double *images, uint32_t *d_flags, hipStream_t st);

// The reference's label checks (simulator.cpp:131-168) on a reduced range.
Status check_labels(int32_t lmin, int32_t lmax, int64_t n_props, int storage);

// Packs C-order fields into march fields on `st` and records their label
// range (returns after the packing is complete).
Status pack_fields(const int64_t dim[3], const int32_t *label, const float *vector,
                   fpl_fields *Fd, hipStream_t st);

} // namespace sim
} // namespace fpl
