// The march's field storage: the label and optical-axis fields repacked in
// [x][z][y] order, so that adjacent lanes (adjacent rays in y) read adjacent
// records and every load of a wave is one contiguous run.  Written by the
// voxelizer's final pass (generation straight into march records) and by the
// pack kernel (fields the caller hands in).  Three storages:
//   ST_F32      16-byte record: int32 label, float axis x, y, z
//   ST_F16       8-byte record: uint16 label, fp16 axis x, y, z; the axis is
//                canonicalised to z >= 0 (the retardance and the interpolation
//                are invariant under v -> -v) and the sign bit of z carries the
//                reference's "not normalized" test of the source axis
//                (simulator.cpp:298-302); axes that pass it are renormalised
//                after decoding
//   ST_ANGLE16   6 bytes per voxel in two arrays: uint16 label (bit 15 = "not
//                normalized") and the axis as 16-bit polar / azimuth codes on the
//                upper hemisphere, polar code 0xffff = zero axis
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fpl {
namespace rec {

enum Storage { ST_F32 = 0, ST_F16 = 1, ST_ANGLE16 = 2 };

constexpr double ANG_THETA = 1.5707963267948966 / 65534.0; // polar step
constexpr double ANG_PHI = 6.283185307179586 / 65536.0;    // azimuth step

// Bytes per voxel of the first array (the angle storage has a second 4-byte one).
inline size_t record_bytes(int storage) {
   return storage == ST_F32 ? 16 : (storage == ST_F16 ? 8 : 2);
}
inline size_t second_bytes(int storage) { return storage == ST_ANGLE16 ? 4 : 0; }

// Largest label a storage holds (labels are checked against it before packing).
inline int64_t label_limit(int storage) {
   return storage == ST_F32 ? 0x7fffffffLL : (storage == ST_F16 ? 0xffffLL : 0x7fffLL);
}

// One voxel in storage ST, as (first-array word(s), second-array word) in an int4.
// Identical arithmetic in every translation unit that includes this header.
__device__ __forceinline__ void encode(int ST, int32_t l, float x, float y, float z, int4 &out) {
#pragma clang fp contract(off)
   if (ST == ST_F32) {
      out = make_int4(l, __float_as_int(x), __float_as_int(y), __float_as_int(z));
      return;
   }
   // "optical_axis not normalized" test of the reference, in double
   // (simulator.cpp:298): length2 of the float components minus one > 1e-5
   const double n2 = (double)x * x + (double)y * y + (double)z * z;
   const bool nn = n2 - 1 > 1e-5;
   if (z < 0.0f) { // canonical orientation: z >= 0
      x = -x;
      y = -y;
      z = -z;
   }
   z = fabsf(z);
   if (ST == ST_F16) {
      const uint32_t hx = __half_as_ushort(__float2half_rn(x));
      const uint32_t hy = __half_as_ushort(__float2half_rn(y));
      const uint32_t hz = __half_as_ushort(__float2half_rn(z)) | (nn ? 0x8000u : 0u);
      out = make_int4((int)(((uint32_t)l & 0xffffu) | (hx << 16)), (int)(hy | (hz << 16)), 0, 0);
      return;
   }
   uint32_t code;
   if (x == 0.0f && y == 0.0f && z == 0.0f) {
      code = 0xffffu;
   } else {
      const double th = atan2(sqrt((double)x * x + (double)y * y), (double)z); // [0, pi/2]
      const double ph = atan2((double)y, (double)x);                          // [-pi, pi]
      uint32_t tc = (uint32_t)rint(th / ANG_THETA);
      tc = tc > 65534u ? 65534u : tc;
      const uint32_t pcode = ((uint32_t)rint((ph + 3.141592653589793) / ANG_PHI)) & 0xffffu;
      code = tc | (pcode << 16);
   }
   out = make_int4((int)(((uint32_t)l & 0x7fffu) | (nn ? 0x8000u : 0u)), (int)code, 0, 0);
}

// Stores an encoded voxel at record index o.
// (streaming stores: the writers, the voxelizer and the pack, never read the
// records back)
__device__ __forceinline__ void store(int ST, void *a, uint32_t *b, int64_t o, const int4 &e) {
   typedef int v4i __attribute__((ext_vector_type(4)));
   typedef uint32_t v2u __attribute__((ext_vector_type(2)));
   if (ST == ST_F32) {
      const v4i v = {e.x, e.y, e.z, e.w};
      __builtin_nontemporal_store(v, (v4i *)a + o);
   } else if (ST == ST_F16) {
      const v2u v = {(uint32_t)e.x, (uint32_t)e.y};
      __builtin_nontemporal_store(v, (v2u *)a + o);
   } else {
      __builtin_nontemporal_store((uint16_t)e.x, (uint16_t *)a + o);
      __builtin_nontemporal_store((uint32_t)e.y, b + o);
   }
}

} // namespace rec
} // namespace fpl
