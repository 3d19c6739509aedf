// Tissue generation (voxelization of layered fiber polylines) for gfx950.
//
// Semantics: fastpli PliGenerator::RunTissueGeneration / FillVoxelsAroundFiberSegment
// (src/simulation/generator.cpp:165-383) -- for every voxel, the segments whose
// integer bounding box contains the voxel are visited in global
// (bundle, fiber, segment) order; a segment updates the voxel when its
// float-rounded FP64 squared distance is within the outer layer and not larger
// than the stored one (later segment wins ties).
//
// GPU structure (no atomics on the volume, every voxel written exactly once):
//   k_fiber_boxes one wave per fiber: translate/resize (fiber.cpp:128-150) and
//                 the fiber AABB (fiber.cpp:83-99) as a wave min/max reduction
//   k_bundle_boxes one workgroup per bundle: union of its fiber boxes
//                 (fiber_bundle.cpp:147-154); min/max is exact in any order
//   k_fibers      one thread per fiber: bundle/fiber culling (generator.cpp:208-221)
//   k_segments    one thread per segment: FP64 record + integer bbox
//                 (generator.cpp:293-305) + the 8x8x32 voxel tiles of the bbox that
//                 can hold a candidate (tile_box): count and mask, and every
//                 (tile, segment) pair placed in its tile's fixed-size list by
//                 per-tile atomic cursors (reserved per workgroup through an
//                 LDS table)
//   k_tile_sort   one wave per tile: its list in global segment order (bitonic),
//                 its range and the bundles at its two ends
//   (lists longer than 256: segment scan, k_pad, k_emit, a stable radix sort
//   by tile and k_ranges instead)
//   k_voxelize    one workgroup per tile, the tile's segment records staged in
//                 LDS; each wave owns every fourth z layer of the tile (8x8x8
//                 voxels), whose state lives in LDS.
//                 A float lower-bound test (conservative by construction)
//                 queues candidates; the exact FP64 test runs on full batches of
//                 64 queued candidates; per voxel the reference's update rule is
//                 applied in segment order.
//
// This file must be compiled with -ffp-contract=off: the distance test must
// round every FP64 operation exactly as the reference's scalar code.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "fpl_gen.h"
#include "fpl_internal.h"
#include "fpl_records.h"

namespace fpl {
namespace gen {

constexpr int TX = 8, TY = 8, TZ = 32; // voxel tile of one workgroup
constexpr int WG = 256;                // 4 waves; wave w owns the layers z = w + 4 l, l < 8
constexpr int SZ = 8;                  // layers of a wave
constexpr int SLOTS = 64 * SZ;         // voxels of a wave: slot = l * 64 + column
constexpr int CHUNK = 40;              // segment records staged in LDS at a time (LDS per workgroup
                                       // 27 KB in the 16-bit build)
constexpr int QN = 128;                // candidate ring of a wave (never holds 128 entries)
constexpr int LSQ_LDS = 64;            // layer tables up to this many rows are read from LDS
constexpr uint32_t SHORT_LIST = 65535; // longest tile list of the 16-bit winner indices

struct SegRec {        // 128 bytes, one per segment (indexed by its first point)
   double p[3];        // start point, voxel units
   double q[3];        // end point
   double rp, rq;      // radii at p and q
   int32_t lo[3];      // integer bbox [lo, hi) after intersection with the volume
   int32_t hi[3];
   float bl[3];        // min(p, q) - 0.5 rounded down and max(p, q) - 0.5 rounded up: the
   float bh[3];        // segment's box in the frame where voxel i has its centre at i
   float thr;          // f_last * float(max(rp, rq)^2), the outer-layer test at the larger
                       // radius, widened by the stage A margins
   float f_last;       // float(last scale^2) of the bundle
   int32_t layer_off;  // the bundle's rows in the (sentinel-padded) layer tables
   int32_t bundle;
};
static_assert(sizeof(SegRec) == 128, "SegRec layout");

struct BundleDev {
   int32_t layer_off;  // into the layer tables
   int32_t n_layers;
   int32_t label_base; // 1 + b * max_layer
   float f_last_sq;    // float(scale_sq.back())
   double sqrt_last;   // sqrt(scale_sq.back())
   int64_t fiber_begin, fiber_end;
};

struct Params {
   int64_t nx, ny, nz;
   int64_t n_points, n_fibers, n_bundles;
   double off[3];  // -origin
   int32_t shift;  // origin != 0
   double scale;   // 1.0 / voxel_size
   int32_t max_layer;
   int32_t ntx, nty, ntz;
   int32_t n_lsq; // rows of the layer tables
   int32_t n_props;
   int32_t single_free; // tiles of one bundle use the free mode (see k_voxelize)
};

// aabb::AABB(p, q) unsorted constructor (aabb.hpp:54-71)
__device__ __forceinline__ void box_of(const double p[3], const double q[3], double lo[3],
                                       double hi[3]) {
   for (int i = 0; i < 3; i++) {
      if (p[i] < q[i]) {
         lo[i] = p[i];
         hi[i] = q[i];
      } else {
         lo[i] = q[i];
         hi[i] = p[i];
      }
   }
}

// aabb::AABB::Unite (aabb.hpp:90-97); min/max is exact and order independent,
// so the union can be reduced in any order
__device__ __forceinline__ void box_unite(double lo[3], double hi[3], const double a[3],
                                          const double b[3]) {
   for (int k = 0; k < 3; k++) {
      if (a[k] < lo[k]) lo[k] = a[k];
      if (b[k] > hi[k]) hi[k] = b[k];
   }
}

__device__ __forceinline__ void box_empty(double lo[3], double hi[3]) {
   for (int k = 0; k < 3; k++) {
      lo[k] = INFINITY;
      hi[k] = -INFINITY;
   }
}

__device__ __forceinline__ void wave_box_reduce(double lo[3], double hi[3]) {
   for (int off = 32; off > 0; off >>= 1) {
      for (int k = 0; k < 3; k++) {
         const double a = __shfl_xor(lo[k], off), b = __shfl_xor(hi[k], off);
         if (a < lo[k]) lo[k] = a;
         if (b > hi[k]) hi[k] = b;
      }
   }
}

__device__ __forceinline__ int find_bundle(const BundleDev *bd, int64_t nb, int64_t f) {
   int lo = 0, hi = (int)nb;
   while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (bd[mid].fiber_begin <= f)
         lo = mid;
      else
         hi = mid;
   }
   return lo;
}

__device__ __forceinline__ int64_t find_fiber(const int64_t *fiber_ptr, int64_t n_fibers,
                                              int64_t i) {
   // largest f with fiber_ptr[f] <= i
   int64_t lo = 0, hi = n_fibers; // answer in [lo, hi)
   while (hi - lo > 1) {
      int64_t mid = (lo + hi) >> 1;
      if (fiber_ptr[mid] <= i)
         lo = mid;
      else
         hi = mid;
   }
   return lo;
}

__device__ __forceinline__ void load_point(const double *P, int64_t i, const Params &pr,
                                           double p[3], double &r) {
   double x = P[4 * i + 0], y = P[4 * i + 1], z = P[4 * i + 2];
   r = P[4 * i + 3];
   if (pr.shift) { // Fiber::Translate (fiber.cpp:128-133)
      x = x + pr.off[0];
      y = y + pr.off[1];
      z = z + pr.off[2];
   }
   // Fiber::Resize (fiber.cpp:144-150)
   p[0] = x * pr.scale;
   p[1] = y * pr.scale;
   p[2] = z * pr.scale;
   r = r * pr.scale;
}

// Fiber::CalculateVoi (fiber.cpp:83-99) after translate/resize: one wave per
// fiber; box layout lo x,y,z then hi x,y,z.  A fiber without points keeps the
// default (+inf, -inf) box.
__global__ void k_fiber_boxes(Params pr, const double *__restrict__ P,
                              const int64_t *__restrict__ fiber_ptr, double *fiber_box) {
   const int64_t f = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   if (f >= pr.n_fibers) return; // uniform per wave
   const int64_t beg = fiber_ptr[f], end = fiber_ptr[f + 1];
   double lo[3], hi[3], tl[3], th[3];
   box_empty(lo, hi);
   for (int64_t i = beg + lane; i < end; i += 64) {
      double p[3], r;
      load_point(P, i, pr, p, r);
      if (i == beg) { // AABB(p0 - r0, p0 + r0) (fiber.cpp:90-91)
         const double a[3] = {p[0] - r, p[1] - r, p[2] - r};
         const double c[3] = {p[0] + r, p[1] + r, p[2] + r};
         box_of(a, c, tl, th);
         box_unite(lo, hi, tl, th);
      }
      if (i + 1 < end) { // segment box grown by max(r_i, r_i+1) (fiber.cpp:92-98)
         double q[3], rq;
         load_point(P, i + 1, pr, q, rq);
         const double rm = (r < rq) ? rq : r; // std::max
         box_of(p, q, tl, th);
         for (int k = 0; k < 3; k++) {
            tl[k] -= rm;
            th[k] += rm;
         }
         box_unite(lo, hi, tl, th);
      }
   }
   wave_box_reduce(lo, hi);
   if (lane < 6) fiber_box[6 * f + lane] = lane < 3 ? lo[lane] : hi[lane - 3];
}

// Bundle::CalculateVoi (fiber_bundle.cpp:147-154): union of its fiber boxes,
// one workgroup per bundle
__global__ void k_bundle_boxes(Params pr, const BundleDev *__restrict__ bd,
                               const double *__restrict__ fiber_box, double *bundle_box,
                               uint32_t *flags) {
   __shared__ double red[6][4];
   const int b = blockIdx.x;
   if (b == 0 && threadIdx.x == 0) {
      flags[0] = 0; // warnings
      flags[1] = 0; // voxelizer: a tile list too long for 16-bit positions
      flags[2] = 0; // binning: bit 0 radix pairs beyond the assumed count, bit 1 a long list
      flags[3] = 0; // voxelizer: a label >= n_props was written
   }
   double lo[3], hi[3];
   box_empty(lo, hi);
   for (int64_t f = bd[b].fiber_begin + threadIdx.x; f < bd[b].fiber_end; f += blockDim.x)
      box_unite(lo, hi, fiber_box + 6 * f, fiber_box + 6 * f + 3);
   wave_box_reduce(lo, hi);
   const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
   if (lane == 0)
      for (int k = 0; k < 3; k++) {
         red[k][w] = lo[k];
         red[3 + k][w] = hi[k];
      }
   __syncthreads();
   if (threadIdx.x < 6) {
      double v = red[threadIdx.x][0];
      for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
         const double o = red[threadIdx.x][i];
         if (threadIdx.x < 3 ? (o < v) : (o > v)) v = o;
      }
      bundle_box[6 * b + threadIdx.x] = v;
   }
}

__device__ __forceinline__ bool overlaps_volume(const double *box, const Params &pr) {
   // aabb::Overlap(volume_bb, box), strict (aabb.hpp:128-137)
   const double vhi[3] = {(double)pr.nx, (double)pr.ny, (double)pr.nz};
   for (int k = 0; k < 3; k++)
      if (0.0 >= box[3 + k] || vhi[k] <= box[k]) return false;
   return true;
}

__global__ void k_fibers(Params pr, const int64_t *__restrict__ fiber_ptr,
                         const BundleDev *__restrict__ bd,
                         const double *__restrict__ fiber_box,
                         const double *__restrict__ bundle_box, uint8_t *fiber_ok,
                         uint32_t *flags) {
   int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (f >= pr.n_fibers) return;
   const int b = find_bundle(bd, pr.n_bundles, f);
   const int64_t n = fiber_ptr[f + 1] - fiber_ptr[f];
   const bool bundle_ok = overlaps_volume(bundle_box + 6 * b, pr) && bd[b].n_layers > 0;
   uint8_t ok = 0;
   if (bundle_ok) {
      if (n <= 1)
         atomicOr(&flags[0], FPL_WARN_SINGLE_POINT);
      else if (overlaps_volume(fiber_box + 6 * f, pr))
         ok = 1;
   }
   fiber_ok[f] = ok;
}

// vm::unit (vemath.hpp:390-402) rounded to float
__device__ __forceinline__ void unit_to_float(double x, double y, double z, float *out) {
   double n2 = 0.0;
   n2 = n2 + x * x;
   n2 = n2 + y * y;
   n2 = n2 + z * z;
   if (n2 > 0) {
      const double inv = 1.0 / sqrt(n2);
      x = x * inv;
      y = y * inv;
      z = z * inv;
   }
   out[0] = (float)x;
   out[1] = (float)y;
   out[2] = (float)z;
}

// The slab of a segment (free mode only; see the key layout at k_voxelize).
// With u the unit direction of segment k = [P_k, P_k+1], a(v) = (v - P_k).u
// and R_out = sqrt(f_last) max(r_k, r_k+1) (no voxel farther from the segment
// can pass its radius test), a voxel with a(v) > L_k + m + d_e is beyond the
// segment's end, and by the choice of d_e its projection on segment k+1 is at
// least m inside it: segment k+1 is at least as close (exactly: |v - P_k+1|^2
// minus the squared projection) and, with r_k+2 >= r_k+1, passes its radius
// test whenever k does, so the later segment k+1 (or one it loses to) wins
// over k.  Likewise a voxel with a(v) < -(m + d_s) is at least m^2 closer to
// segment k-1 (strictly, also after rounding to float: m >= R_out 2^-9),
// which wins when r_k-1 >= r_k.  d = (m (1 - c) + s R_out) / c with c, s the
// cosine and sine of the bend to the neighbour (no clip for bends beyond
// 60 deg, short segments, missing neighbours or the radius conditions
// failing).  m also covers the float rounding of a(v) in the clip test.
// seg_clip[2 i] = (u, a_lo), seg_clip[2 i + 1] = (a_hi, P_k - 0.5 as float)
__device__ void segment_clip(const Params &pr, const double *P, int64_t i, bool has_prev,
                             bool has_next, const double p[3], const double q[3], double rp,
                             double rq, double r_out, float4 *clip) {
   const double inf = __builtin_huge_val();
   double d[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
   const double L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
   const double m = fmax(0x1p-5, r_out * 0x1p-9) +
                    (double)(pr.nx + pr.ny + pr.nz) * 0x1p-19;
   double lo = -inf, hi = inf;
   double u[3] = {0.0, 0.0, 0.0};
   if (L >= 4.0 * m) {
      for (int k = 0; k < 3; k++) u[k] = d[k] / L;
      auto delta = [&](const double e[3], double Le) {
         double c = 0.0;
         for (int k = 0; k < 3; k++) c += u[k] * e[k] / Le;
         if (!(c > 0.5)) return inf;
         const double sn = sqrt(fmax(0.0, 1.0 - c * c));
         return (m * (1.0 - c) + sn * r_out) / c;
      };
      if (has_next) {
         double pn[3], rn;
         load_point(P, i + 2, pr, pn, rn);
         const double e[3] = {pn[0] - q[0], pn[1] - q[1], pn[2] - q[2]};
         const double Le = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
         if (Le >= 2.0 * m && rn >= rq) hi = L + m + delta(e, Le);
      }
      if (has_prev) {
         double pp[3], rpp;
         load_point(P, i - 1, pr, pp, rpp);
         const double e[3] = {p[0] - pp[0], p[1] - pp[1], p[2] - pp[2]};
         const double Le = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
         if (Le >= 2.0 * m && rpp >= rp) lo = -(m + delta(e, Le));
      }
   }
   clip[0] = make_float4((float)u[0], (float)u[1], (float)u[2], (float)lo);
   clip[1] = make_float4((float)hi, (float)(p[0] - 0.5), (float)(p[1] - 0.5), (float)(p[2] - 0.5));
}

// A record's candidate region in the voxel tile at (xt, yt, zt): its integer
// box within the tile, and the z range [za, zb) narrowed by the tile's nearest
// in-box columns (empty: za >= zb).  The in-box columns nearest to the
// segment's box bound every voxel's float ex and ey from below (float
// rounding is monotone), so a voxel that passes the float radius test
// RN(ez^2 + exy) <= thr has ez^2 <= rem + ulp(thr), rem = thr - exy_min >= 0;
// h covers that with margin (the hardware square root is within 1 ulp), and
// the range ends are computed in double, exact for any float box and voxel
// index.  A tile where the range is empty holds no candidate of the segment:
// k_segments and k_emit leave it out of the tile lists.
struct TileBox {
   int cx0, cx1, cy0, cy1, za, zb;
};
__device__ __forceinline__ TileBox tile_box(const SegRec &R, int xt, int yt, int zt) {
   TileBox b;
   const int z0 = max(R.lo[2], zt), z1 = min(R.hi[2], zt + TZ);
   b.cx0 = max(R.lo[0], xt);
   b.cx1 = min(R.hi[0], xt + TX);
   b.cy0 = max(R.lo[1], yt);
   b.cy1 = min(R.hi[1], yt + TY);
   b.za = 0;
   b.zb = 0;
   if (z0 < z1 && b.cx0 < b.cx1 && b.cy0 < b.cy1) {
      const float thr = R.thr;
      const float exm =
          fmaxf(fmaxf(R.bl[0] - (float)(b.cx1 - 1), (float)b.cx0 - R.bh[0]), 0.0f);
      const float eym =
          fmaxf(fmaxf(R.bl[1] - (float)(b.cy1 - 1), (float)b.cy0 - R.bh[1]), 0.0f);
      const float rem = thr - fmaf(eym, eym, exm * exm);
      if (rem >= 0.0f) {
         const float h = fmaf(__builtin_amdgcn_sqrtf(fmaf(thr, 0x1p-20f, rem)),
                              1.0f + 0x1p-10f, 0x1p-10f);
         b.za = max(z0, (int)ceil((double)R.bl[2] - (double)h));
         b.zb = min(z1, (int)floor((double)R.bh[2] + (double)h) + 1);
      }
   }
   return b;
}

// Per-tile counting of a workgroup's (tile, segment) pairs in LDS: the
// segments of a workgroup are consecutive pieces of a few fibers and share
// most of their tiles, so one global atomic per distinct tile replaces one
// per pair (same-address global atomics serialise in L2).  Open addressing
// with a bounded probe: a tile that finds no slot takes the global atomic
// itself.  A tile is either in the table or finds no slot for every pair of
// the workgroup (slots are never freed).
constexpr int HASH_N = 2048;           // slots of a workgroup's table
constexpr uint32_t HASH_EMPTY = 0xffffffffu;
constexpr int HASH_PROBES = 32;

struct TileHash {
   uint32_t key[HASH_N];
   uint32_t cnt[HASH_N];
};

__device__ __forceinline__ uint32_t hash_start(uint32_t t) {
   return (t * 2654435761u) >> (32 - 11); // 11 bits: HASH_N slots
}

// slot of tile t (inserted if absent), or -1 when its probe run is full
__device__ __forceinline__ int hash_insert(TileHash &H, uint32_t t) {
   uint32_t h = hash_start(t);
   for (int p = 0; p < HASH_PROBES; p++) {
      const uint32_t prev = atomicCAS(&H.key[h], HASH_EMPTY, t);
      if (prev == HASH_EMPTY || prev == t) return (int)h;
      h = (h + 1) & (HASH_N - 1);
   }
   return -1;
}

// slot of tile t after every insertion of the workgroup, or -1
__device__ __forceinline__ int hash_find(const TileHash &H, uint32_t t) {
   uint32_t h = hash_start(t);
   for (int p = 0; p < HASH_PROBES; p++) {
      const uint32_t k = H.key[h];
      if (k == t) return (int)h;
      if (k == HASH_EMPTY) return -1;
      h = (h + 1) & (HASH_N - 1);
   }
   return -1;
}

__device__ __forceinline__ void hash_clear(TileHash &H) {
   for (int k = threadIdx.x; k < HASH_N; k += blockDim.x) {
      H.key[k] = HASH_EMPTY;
      H.cnt[k] = 0;
   }
}

// Tile lists of the binning by atomics: SORT_MAX entries per tile, tile t's at
// list[t * SORT_MAX]; a tile with more segments makes the host bin the volume
// again with the radix sort (its list is longer than the per-tile sort takes).
constexpr uint32_t SORT_MAX = 256;

// One thread per segment: its record, direction and slab data, the kept tiles
// of its box (count, and a mask for k_emit), and with `tile_count` set the
// binning by atomics: every kept (tile, segment) pair goes into the tile's
// list at the tile's cursor (tile_count[t], which ends as the tile's segment
// count), in no particular order (k_tile_sort orders the lists).  The
// segments of a workgroup are consecutive pieces of a few fibers and share
// most of their tiles: the workgroup counts its pairs per tile in an LDS
// table, reserves one range per distinct tile with one global atomic, and
// places its pairs inside those ranges (same-address global atomics
// serialise in L2).  A tile without a table slot takes a global atomic per
// pair.
__global__ void k_segments(Params pr, const double *__restrict__ P,
                           const int64_t *__restrict__ fiber_ptr,
                           const BundleDev *__restrict__ bd, const uint8_t *__restrict__ fiber_ok,
                           SegRec *segs, float4 *seg_dir, float4 *seg_clip, int64_t *seg_count,
                           uint32_t *tile_count, uint32_t *list) {
   __shared__ TileHash H;
   __shared__ uint32_t base[HASH_N];
   if (tile_count) {
      hash_clear(H);
      __syncthreads();
   }
   int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   const bool seg = i < pr.n_points;
   const int64_t f = seg ? find_fiber(fiber_ptr, pr.n_fibers, i) : 0;
   int64_t count = 0;
   SegRec R;
   uint32_t mask = 0;
   if (seg && fiber_ok[f] && i + 1 < fiber_ptr[f + 1]) {
      const int b = find_bundle(bd, pr.n_bundles, f);
      double rp, rq;
      load_point(P, i, pr, R.p, rp);
      load_point(P, i + 1, pr, R.q, rq);
      R.rp = rp;
      R.rq = rq;
      // generator.cpp:295-305
      const double rmax = (rp < rq) ? rq : rp; // std::max
      const double max_r = rmax * bd[b].sqrt_last;
      double lo[3], hi[3];
      box_of(R.p, R.q, lo, hi);
      for (int k = 0; k < 3; k++) {
         R.bl[k] = __double2float_rd(lo[k] - 0.5);
         R.bh[k] = __double2float_ru(hi[k] - 0.5);
      }
      const double vhi[3] = {(double)pr.nx, (double)pr.ny, (double)pr.nz};
      bool empty = false;
      int64_t ilo[3], ihi[3];
      for (int k = 0; k < 3; k++) {
         lo[k] -= max_r;
         hi[k] += max_r;
         if (0.0 > lo[k]) lo[k] = 0.0; // Intersect (aabb.hpp:108-115)
         if (vhi[k] < hi[k]) hi[k] = vhi[k];
         ilo[k] = (int64_t)lo[k];
         ihi[k] = (int64_t)(hi[k] + 0.5);
         if (!(ilo[k] < ihi[k])) empty = true;
      }
      if (!empty) {
         for (int k = 0; k < 3; k++) {
            R.lo[k] = (int32_t)ilo[k];
            R.hi[k] = (int32_t)ihi[k];
         }
         R.f_last = bd[b].f_last_sq;
         // outer-layer threshold with the stage A margins (relative 2^-12, absolute 2^-6)
         R.thr = fmaf(R.f_last * (float)(rmax * rmax), 1.0f + 0x1p-12f, 0x1p-6f);
         R.layer_off = bd[b].layer_off;
         R.bundle = b;
         segs[i] = R;
         // the 'p' orientation of the segment, unit(q - p) as the voxel update
         // computes it (generator.cpp:365-378)
         float d[3];
         unit_to_float(R.q[0] - R.p[0], R.q[1] - R.p[1], R.q[2] - R.p[2], d);
         seg_dir[i] = make_float4(d[0], d[1], d[2], 0.0f);
         segment_clip(pr, P, i, i > fiber_ptr[f], i + 2 < fiber_ptr[f + 1], R.p, R.q, rp, rq,
                      max_r, seg_clip + 2 * i);
         // the tiles of the box that hold candidates of the segment; for boxes
         // of up to 31 tiles their mask (bit k: the k-th tile in the loop
         // order below) goes into the direction entry's w, else bit 31 asks
         // for the tiles to be tested again
         const int tx0 = R.lo[0] / TX, tx1 = (R.hi[0] - 1) / TX;
         const int ty0 = R.lo[1] / TY, ty1 = (R.hi[1] - 1) / TY;
         const int tz0 = R.lo[2] / TZ, tz1 = (R.hi[2] - 1) / TZ;
         int k = 0;
         for (int tx = tx0; tx <= tx1; tx++)
            for (int ty = ty0; ty <= ty1; ty++)
               for (int tz = tz0; tz <= tz1; tz++, k++) {
                  const TileBox b = tile_box(R, tx * TX, ty * TY, tz * TZ);
                  if (b.za < b.zb) {
                     count++;
                     if (k < 31) mask |= 1u << k;
                     if (tile_count) {
                        const uint32_t t = (uint32_t)((tx * pr.nty + ty) * pr.ntz + tz);
                        const int h = hash_insert(H, t);
                        if (h >= 0) atomicAdd(&H.cnt[h], 1u);
                     }
                  }
               }
         if (k > 31) mask = 0x80000000u;
         seg_dir[i].w = __uint_as_float(mask);
      }
   }
   if (seg) seg_count[i] = count;
   if (!tile_count) return;
   // one reservation per tile of the table, then every pair its place
   __syncthreads();
   for (int k = threadIdx.x; k < HASH_N; k += blockDim.x) {
      if (H.key[k] != HASH_EMPTY) {
         base[k] = atomicAdd(&tile_count[H.key[k]], H.cnt[k]);
         H.cnt[k] = 0;
      }
   }
   __syncthreads();
   if (count == 0) return;
   const int tx0 = R.lo[0] / TX, tx1 = (R.hi[0] - 1) / TX;
   const int ty0 = R.lo[1] / TY, ty1 = (R.hi[1] - 1) / TY;
   const int tz0 = R.lo[2] / TZ, tz1 = (R.hi[2] - 1) / TZ;
   int k = 0;
   for (int tx = tx0; tx <= tx1; tx++)
      for (int ty = ty0; ty <= ty1; ty++)
         for (int tz = tz0; tz <= tz1; tz++, k++) {
            bool keep = (mask >> k) & 1u;
            if (mask & 0x80000000u) {
               const TileBox b = tile_box(R, tx * TX, ty * TY, tz * TZ);
               keep = b.za < b.zb;
            }
            if (!keep) continue;
            const uint32_t t = (uint32_t)((tx * pr.nty + ty) * pr.ntz + tz);
            const int h = hash_find(H, t);
            const uint32_t r =
                h >= 0 ? base[h] + atomicAdd(&H.cnt[h], 1u) : atomicAdd(&tile_count[t], 1u);
            if (r < SORT_MAX) list[(size_t)t * SORT_MAX + r] = (uint32_t)i;
         }
}

__global__ void k_emit(Params pr, const SegRec *__restrict__ segs,
                       const float4 *__restrict__ seg_dir,
                       const int64_t *__restrict__ seg_count, const int64_t *__restrict__ seg_off,
                       int64_t cap, uint32_t *keys, uint32_t *vals) {
   int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= pr.n_points) return;
   if (seg_count[i] == 0) return;
   const SegRec &R = segs[i];
   int64_t o = seg_off[i];
   const int tx0 = R.lo[0] / TX, tx1 = (R.hi[0] - 1) / TX;
   const int ty0 = R.lo[1] / TY, ty1 = (R.hi[1] - 1) / TY;
   const int tz0 = R.lo[2] / TZ, tz1 = (R.hi[2] - 1) / TZ;
   const uint32_t mask = __float_as_uint(seg_dir[i].w); // k_segments' tile mask
   int k = 0;
   for (int tx = tx0; tx <= tx1; tx++)
      for (int ty = ty0; ty <= ty1; ty++)
         for (int tz = tz0; tz <= tz1; tz++, k++) {
            bool keep = (mask >> k) & 1u;
            if (mask & 0x80000000u) { // as k_segments counts
               const TileBox b = tile_box(R, tx * TX, ty * TY, tz * TZ);
               keep = b.za < b.zb;
            }
            if (keep) {
               if (o < cap) { // pairs past the capacity: k_pad reports them
                  keys[o] = (uint32_t)(((int64_t)tx * pr.nty + ty) * pr.ntz + tz);
                  vals[o] = (uint32_t)i;
               }
               o++;
            }
         }
}

// The pair count (the scan's total) into flags[4..5] for the host's next
// call; when the host sized the pair arrays without reading it (cap, the
// previous call's count), pads [count, cap) with keys past the last tile, or
// sets flags[2] when the count exceeds cap.
__global__ void k_pad(const int64_t *__restrict__ count, int64_t cap, uint32_t pad_key,
                      uint32_t *keys, uint32_t *vals, uint32_t *flags) {
   const int64_t n = *count;
   const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (p == 0) {
      flags[4] = (uint32_t)n;
      flags[5] = (uint32_t)((uint64_t)n >> 32);
      if (n > cap) flags[2] = 1;
   }
   if (p >= n && p < cap) {
      keys[p] = pad_key;
      vals[p] = 0;
   }
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
   int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = v;
}

// Each tile's [begin, end) in the sorted list, and the bundles of its first
// and last segment (the list is in global segment order, bundle-major)
__global__ void k_ranges(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                         int64_t n, uint32_t n_tiles, const SegRec *__restrict__ segs,
                         uint32_t *range, int2 *tile_bundle) {
   int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (p >= n) return;
   const uint32_t t = keys[p];
   if (t >= n_tiles) return; // padding (k_pad)
   if (p == 0 || keys[p - 1] != t) {
      range[2 * (int64_t)t] = (uint32_t)p;
      tile_bundle[t].x = segs[vals[p]].bundle;
   }
   if (p == n - 1 || keys[p + 1] != t) {
      range[2 * (int64_t)t + 1] = (uint32_t)(p + 1);
      tile_bundle[t].y = segs[vals[p]].bundle;
   }
}

// ---- binning by atomics: k_segments places the (tile, segment) pairs in the
// tiles' fixed-size lists (in no particular order) and k_tile_sort puts every
// list in global segment order, one wave per tile.  A longer list makes the
// host bin the volume again with the radix sort above.

// Bitonic sort of the n <= 64 E values at p (ascending), element e * 64 + lane
// in register e of the lane; the padding sorts last.
template <int E>
__device__ __forceinline__ void wave_sort(uint32_t *p, uint32_t n, int lane, uint32_t &first,
                                          uint32_t &last) {
   uint32_t v[E];
#pragma unroll
   for (int e = 0; e < E; e++) {
      const uint32_t idx = (uint32_t)(e * 64 + lane);
      v[e] = idx < n ? p[idx] : 0xffffffffu;
   }
#pragma unroll
   for (int k = 2; k <= 64 * E; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
         if (j >= 64) { // the partner is another register of this lane
            const int je = j >> 6;
#pragma unroll
            for (int e = 0; e < E; e++) {
               if (e & je) continue;
               const bool up = (((e * 64 + lane) & k) == 0);
               const uint32_t lo = min(v[e], v[e | je]), hi = max(v[e], v[e | je]);
               v[e] = up ? lo : hi;
               v[e | je] = up ? hi : lo;
            }
         } else {
            const bool lower = (lane & j) == 0;
#pragma unroll
            for (int e = 0; e < E; e++) {
               const uint32_t o = (uint32_t)__shfl_xor((int)v[e], j);
               const bool up = (((e * 64 + lane) & k) == 0);
               v[e] = (up == lower) ? min(v[e], o) : max(v[e], o);
            }
         }
      }
   }
#pragma unroll
   for (int e = 0; e < E; e++) {
      const uint32_t idx = (uint32_t)(e * 64 + lane);
      if (idx < n) p[idx] = v[e];
   }
   const uint32_t el = (n - 1) >> 6;
   uint32_t w = v[0];
#pragma unroll
   for (int e = 1; e < E; e++)
      if (el == (uint32_t)e) w = v[e];
   first = (uint32_t)__shfl((int)v[0], 0);
   last = (uint32_t)__shfl((int)w, (int)((n - 1) & 63));
}

// One wave per tile: its list in global segment order, its range [begin, end)
// and the bundles of its first and last segment.  A tile with more than
// SORT_MAX segments sets flags[2] bit 1 (the host bins again); its range
// stays inside its list, so the voxelizer still reads in bounds.
__global__ void __launch_bounds__(256) k_tile_sort(uint32_t n_tiles,
                                                   const uint32_t *__restrict__ tile_count,
                                                   uint32_t *list,
                                                   const SegRec *__restrict__ segs,
                                                   uint32_t *range, int2 *tile_bundle,
                                                   uint32_t *flags) {
   const uint32_t t = blockIdx.x * 4u + (threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   if (t >= n_tiles) return;
   const uint32_t c = tile_count[t];
   const uint32_t n = c < SORT_MAX ? c : SORT_MAX;
   const uint32_t b = t * SORT_MAX;
   if (lane == 0) {
      range[2 * (int64_t)t] = b;
      range[2 * (int64_t)t + 1] = b + n;
   }
   if (n == 0) return;
   if (c > SORT_MAX) {
      if (lane == 0) atomicOr(&flags[2], 2u);
      return;
   }
   uint32_t first, last;
   if (n <= 64)
      wave_sort<1>(list + b, n, lane, first, last);
   else if (n <= 128)
      wave_sort<2>(list + b, n, lane, first, last);
   else
      wave_sort<4>(list + b, n, lane, first, last);
   if (lane == 0) tile_bundle[t] = make_int2(segs[first].bundle, segs[last].bundle);
}

struct SegExact { // what the FP64 distance test reads of a segment
   double p[3], v[3], q[3], c2;
};

// v = q - p and c2 = dot(v, v) rounded as generator.cpp:283-289 computes them
__device__ __forceinline__ SegExact seg_exact(const double p[3], const double q[3]) {
   SegExact E;
   for (int k = 0; k < 3; k++) {
      E.p[k] = p[k];
      E.q[k] = q[k];
      E.v[k] = q[k] - p[k];
   }
   double c2 = 0.0;
   c2 = c2 + E.v[0] * E.v[0];
   c2 = c2 + E.v[1] * E.v[1];
   c2 = c2 + E.v[2] * E.v[2];
   E.c2 = c2;
   return E;
}

struct Closest {
   double c[3];
   double t;
};

// ShortestPointToLineSegmentVecCalculation (generator.cpp:475-494)
__device__ __forceinline__ Closest closest(const double pt[3], const SegExact &R) {
   const double w0 = pt[0] - R.p[0], w1 = pt[1] - R.p[1], w2 = pt[2] - R.p[2];
   double c1 = 0.0;
   c1 = c1 + w0 * R.v[0];
   c1 = c1 + w1 * R.v[1];
   c1 = c1 + w2 * R.v[2];
   Closest r;
   if (c1 < 0) {
      r.c[0] = R.p[0];
      r.c[1] = R.p[1];
      r.c[2] = R.p[2];
      r.t = 0.0;
   } else if (R.c2 <= c1) {
      r.c[0] = R.q[0];
      r.c[1] = R.q[1];
      r.c[2] = R.q[2];
      r.t = 1.0;
   } else {
      const double b = c1 / R.c2;
      r.c[0] = R.p[0] + R.v[0] * b;
      r.c[1] = R.p[1] + R.v[1] * b;
      r.c[2] = R.p[2] + R.v[2] * b;
      r.t = b;
   }
   return r;
}

// the parameter t of ShortestPointToLineSegmentVecCalculation
// (generator.cpp:475-494), with the operations of seg_exact and closest
__device__ __forceinline__ double closest_t(const double pt[3], const double p[3],
                                           const double q[3]) {
   const SegExact E = seg_exact(p, q);
   const double w0 = pt[0] - p[0], w1 = pt[1] - p[1], w2 = pt[2] - p[2];
   double c1 = 0.0;
   c1 = c1 + w0 * E.v[0];
   c1 = c1 + w1 * E.v[1];
   c1 = c1 + w2 * E.v[2];
   if (c1 < 0) return 0.0;
   if (E.c2 <= c1) return 1.0;
   return c1 / E.c2;
}

struct WinRec { // what the final pass reads of a winning segment
   double p[3], q[3], rp, rq;
   int32_t layer_off, bundle;
   float dir[3]; // unit(q - p) as float, the 'p' orientation
};

__device__ __forceinline__ WinRec load_winner(const SegRec &R, const float4 &d) {
   WinRec o;
   for (int k = 0; k < 3; k++) {
      o.p[k] = R.p[k];
      o.q[k] = R.q[k];
   }
   o.rp = R.rp;
   o.rq = R.rq;
   o.layer_off = R.layer_off;
   o.bundle = R.bundle;
   o.dir[0] = d.x;
   o.dir[1] = d.y;
   o.dir[2] = d.z;
   return o;
}

__device__ __forceinline__ int tile_of_block(int bid, int n) {
   // keep spatially adjacent tiles (which share segment records) on one XCD:
   // the hardware deals consecutive workgroups round-robin over 8 XCDs
   const int full = (n / 8) * 8;
   if (bid >= full) return bid;
   const int per = full / 8;
   return (bid % 8) * per + bid / 8;
}

// orders the LDS accesses of one wave: its lanes exchange data through LDS
__device__ __forceinline__ void wave_sync() {
   __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
   __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int wave_min(int v) {
   for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
   return v;
}

__device__ __forceinline__ int wave_max(int v) {
   for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
   return v;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Inclusive prefix sum over the wave's 64 lanes: DPP row shifts within each
// row of 16 lanes, then the broadcasts of lane 15 to rows 1 and 3 and of lane
// 31 to rows 2 and 3.  Lanes with no source lane add 0.
__device__ __forceinline__ int wave_scan_add(int x) {
   x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false); // row_shr:1
   x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false); // row_shr:2
   x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false); // row_shr:4
   x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false); // row_shr:8
   x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false); // row_bcast:15
   x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false); // row_bcast:31
   return x;
}

// Inclusive prefix maximum of non-negative values, as wave_scan_add.
__device__ __forceinline__ int wave_scan_max(int x) {
   x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));
   x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));
   x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));
   x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));
   x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));
   x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));
   return x;
}

// Load-balancing search over a 64-lane window [wb, wb + 64) of a run of
// items (the pairs of a chunk's records, the candidates of a group's pairs):
// returns the owner of the window's item at this lane, the last owner i (a
// lane) with items whose first item start[i] (non-decreasing in i) lies at or
// before that position.  Owners with items that start inside the window mark
// their start with 1 + i (a byte), a prefix maximum carries the marks
// forward, and `carry` (1 + the owner of the item before the window, 0 at the
// first window) covers the positions before the first mark; `carry` is
// updated for the next window.
__device__ __forceinline__ int lane_owner(uint8_t *own, int lane, bool has_items, int start,
                                          int wb, int &carry) {
   own[lane] = 0;
   wave_sync();
   if (has_items && start >= wb && start < wb + 64) own[start - wb] = (uint8_t)(lane + 1);
   wave_sync();
   const int o = max(wave_scan_max((int)own[lane]), carry);
   carry = __builtin_amdgcn_readlane(o, 63);
   wave_sync(); // the marks are rewritten for the next window
   return o - 1;
}

// Shrinks [xa, xb) to the x for which some y in [ya, yb) has
// a_lo <= ux x + uy y + a0 <= a_hi, conservatively: the bounds are widened by
// 1/64 in a and in x, which covers the float rounding for coordinates below
// 2^16 (the segment clip's margin covers the rest).
__device__ __forceinline__ void clip_axis(float ux, float uy, float a0, float alo, float ahi,
                                          int &xa, int &xb, int ya, int yb) {
   const float y0 = uy * (float)ya, y1 = uy * (float)(yb - 1);
   const float lo_a = alo - a0 - fmaxf(y0, y1) - 0x1p-6f; // ux x must reach this
   const float hi_a = ahi - a0 - fminf(y0, y1) + 0x1p-6f; // and stay below this
   if (fabsf(ux) < 0x1p-12f) {
      // ux x barely varies over the columns: keep all or none
      const float r0 = ux * (float)xa, r1 = ux * (float)(xb - 1);
      if (!(fmaxf(r0, r1) >= lo_a && fminf(r0, r1) <= hi_a)) xb = xa;
      return;
   }
   // rcp is within 1 ulp of 1 / ux: the quotients are within 1.5 * 2^-23 of
   // their exact values, below 2^-6.4 for |x| < 2^16
   const float inv = __builtin_amdgcn_rcpf(ux);
   float lo = lo_a * inv, hi = hi_a * inv;
   if (ux < 0.0f) {
      const float t = lo;
      lo = hi;
      hi = t;
   }
   lo = fmaxf(lo - 0x1p-6f, -0x1p30f);
   hi = fminf(hi + 0x1p-6f, 0x1p30f);
   xa = max(xa, (int)ceilf(lo));
   xb = min(xb, (int)floorf(hi) + 1);
}

// A voxel's state is one 64-bit key: the float bits of its stored squared
// distance (high word; non-negative floats order as integers) and a low word
// that breaks ties.
//
// Ordered mode (the tile's candidates applied in global segment order, which
// the "objects overlap" warning needs): at rest the low word is 0x80000000 |
// label (32-bit winner indices, separate winner array) or 0x80000000 |
// label15 << 16 | position + 1 (16-bit list positions, no winner array;
// label15 0x7FFF = "label later").  A candidate of a batch enters with the
// low word 63 - lane, so the minimum key over the stored state and the batch
// is the smallest distance with ties going to the latest candidate -- the
// state the reference's sequential rule "update if stored >= new" ends in.
//
// Free mode (16-bit positions; the warning already raised or the tile's
// segments all from one bundle, so it cannot be raised here): the low word is
// 0xFFFF - position.  The minimum key is then the smallest distance with ties
// going to the latest segment whatever the order the candidates come in, and
// candidates that can never be that minimum may be left out: each segment's
// candidates are clipped to its slab (segment_clip).
constexpr unsigned long long KEY_EMPTY = (0x7F800000ull << 32) | 0x80000000ull; // +inf, label 0
constexpr int32_t LABEL_LATER = 0x7FFFFFFF; // label left to the final pass
constexpr uint32_t LABEL15_LATER = 0x7FFF;  // the same in the packed low word

// std::lower_bound over the bundle's double scale^2 list with the float ratio
// (generator.cpp:345-349) -> label ly + 1 + bundle * max_layer (:351); the
// list ends with a +inf sentinel, so the search stops at n_layers at the latest
__device__ __forceinline__ int32_t layer_label(float dist_sq, float r_sq, const double *lsq,
                                               int32_t layer_off, int32_t bundle, int max_layer) {
   const double ratio = (double)(dist_sq / r_sq);
   const double *L = lsq + layer_off;
   int ly = 0;
   while (L[ly] < ratio) ly++;
   return ly + 1 + bundle * max_layer;
}

// SegT: uint16_t keeps the winner's position in the tile list inside the key
// (LDS of a workgroup below 27 KB: 6 workgroups per CU); int32_t stores its
// global index (-1: none) in a separate array, for tiles with longer lists or
// labels that do not fit 15 bits
template <typename SegT> struct WaveState {
   unsigned long long key[SLOTS];
   SegT seg[SLOTS];
   uint32_t queue[QN];  // candidates (chunk index << 9 | slot), ring
   uint2 tab[64];       // per chunk record: candidate prefix | box << 16, division magics
   uint8_t own[64];     // owner marks of a 64-lane window (lane_owner)
};
template <> struct WaveState<uint16_t> {
   unsigned long long key[SLOTS];
   uint32_t queue[QN];
   uint2 tab[64];
   uint8_t own[64];
};

struct alignas(16) LdsSeg { // 144-byte stride spreads records over the LDS banks
   SegRec r;
   int32_t pad[4];
};

// Exact FP64 test of up to 64 queued candidates (one per lane) and their
// ordered application to the wave's voxel state.
template <typename SegT>
__device__ __forceinline__ void run_batch(WaveState<SegT> &W, const LdsSeg *chunk,
                                          const uint32_t *chunk_gidx, uint32_t pos1, int n,
                                          int q_head, int lane, int xt, int yt, int zw,
                                          int max_layer,
                                          const double *lsq, const uint8_t *lmodel,
                                          uint32_t *flags, bool &ov, bool free_mode) {
   constexpr bool packed = sizeof(SegT) == 2;
   uint32_t free_lo = 0; // free mode: the key's low word of an accepted candidate
   const bool valid = lane < n;
   const uint32_t e = valid ? W.queue[(q_head + lane) & (QN - 1)] : 0u;
   const int c = (int)(e >> 9), slot = (int)(e & (SLOTS - 1));
   // the winner index a candidate leaves: its list position + 1 (16-bit
   // build) or its global segment index
   const uint32_t widx = packed ? pos1 + (uint32_t)c : (valid ? chunk_gidx[c] : 0u);
   bool acc = false;
   float dist_sq = 0.0f;
   int32_t new_label = 0;
   if (valid) {
      const SegRec &R = chunk[c].r;
      const double p[3] = {R.p[0], R.p[1], R.p[2]};
      const double q[3] = {R.q[0], R.q[1], R.q[2]};
      const double rp = R.rp, rq = R.rq;
      const float f_last = R.f_last;
      const SegExact E = seg_exact(p, q);
      const int col = slot & 63;
      const double pt[3] = {(double)(xt + (col >> 3)) + 0.5, (double)(yt + (col & 7)) + 0.5,
                            (double)(zw + 4 * (slot >> 6)) + 0.5};
      // FillVoxelsAroundFiberSegment inner loop (generator.cpp:322-349)
      const Closest cp = closest(pt, E);
      const double d0 = cp.c[0] - pt[0], d1 = cp.c[1] - pt[1], d2v = cp.c[2] - pt[2];
      double dd = 0.0;
      dd = dd + d0 * d0;
      dd = dd + d1 * d1;
      dd = dd + d2v * d2v;
      dist_sq = (float)dd;
      const double rr = rp * (1 - cp.t) + rq * cp.t;
      const float r_sq = (float)(rr * rr);
      acc = !(dist_sq > f_last * r_sq);
      if (packed && free_mode) {
         // free mode keeps the winner's label and layer model in its key, so
         // the final pass needs no record of it: (0x10000 - position - 1) << 16
         // | model << 14 | label (model 0 'b', 1 'p', 2 'r'; labels < 2^14)
         if (acc) {
            const int32_t lab = layer_label(dist_sq, r_sq, lsq, R.layer_off, R.bundle, max_layer);
            const int32_t layer = lab - 1 - R.bundle * max_layer;
            // the past-the-end layer of a bundle with max_layer layers is
            // max_layer labels from its layer 0: replacing one by the other is
            // an overlap for the reference (generator.cpp:353-359), which the
            // order-free single-bundle mode does not follow; the host reruns
            // the volume in order (bit 2 of flags[3])
            if (!ov && layer == max_layer) atomicOr(&flags[3], 4u);
            const uint8_t m = lmodel[R.layer_off + layer];
            const uint32_t m2 = m == 'p' ? 1u : (m == 'r' ? 2u : 0u);
            free_lo = (0x10000u - widx) << 16 | m2 << 14 | (uint32_t)lab;
         }
      } else {
         // the label is needed here only while the overlap warning is
         // undecided; otherwise the final pass computes it for the winner alone
         new_label = LABEL_LATER;
         if (acc && !ov)
            new_label = layer_label(dist_sq, r_sq, lsq, R.layer_off, R.bundle, max_layer);
      }
   }
   if (packed && free_mode) {
      // order-free minimum: smallest distance, then latest list position
      if (acc)
         atomicMin(&W.key[slot], ((unsigned long long)__float_as_uint(dist_sq) << 32) | free_lo);
      wave_sync();
      return;
   }
   // the low word a winner keeps at rest
   uint32_t rest_lo = 0x80000000u | (uint32_t)new_label;
   if (packed)
      rest_lo = 0x80000000u |
                (new_label == LABEL_LATER ? LABEL15_LATER : (uint32_t)new_label) << 16 |
                widx;
   auto label_of = [&](unsigned long long key) -> int {
      if (packed) return (int)((uint32_t)(key >> 16) & 0x7FFFu);
      return (int)(uint32_t)(key & 0x7FFFFFFFull);
   };
   // The overlap warning (generator.cpp:353-359) depends on the order of the
   // updates.  Unless it is already raised, a batch whose labels (new ones and
   // the ones they may replace) all lie within max_layer of each other cannot
   // raise it; any other batch is applied one candidate at a time.
   bool serial = false;
   if (!ov) {
      unsigned long long old = KEY_EMPTY;
      if (acc) old = W.key[slot];
      int mn = INT_MAX, mx = INT_MIN;
      if (acc) {
         mn = mx = new_label;
         const int ol = label_of(old);
         if (ol != 0) {
            mn = min(mn, ol);
            mx = max(mx, ol);
         }
      }
      mn = wave_min(mn);
      mx = wave_max(mx);
      serial = uniform((int)(mx != INT_MIN && (long long)mx - mn >= max_layer)) != 0;
   }
   wave_sync();
   const unsigned long long rest = ((unsigned long long)__float_as_uint(dist_sq) << 32) | rest_lo;
   if (!serial) {
      const unsigned long long k =
          ((unsigned long long)__float_as_uint(dist_sq) << 32) | (unsigned long long)(63 - lane);
      if (acc) atomicMin(&W.key[slot], k);
      wave_sync();
      if (acc && W.key[slot] == k) {
         W.key[slot] = rest;
         if constexpr (!packed) W.seg[slot] = (SegT)widx;
      }
   } else {
      uint32_t overlap = 0;
      for (int i = 0; i < n; i++) {
         if (lane == i && acc) {
            const unsigned long long cur = W.key[slot];
            if (__uint_as_float((uint32_t)(cur >> 32)) >= dist_sq) {
               const int ol = label_of(cur);
               if (ol != 0 && abs(ol - new_label) >= max_layer) overlap = 1;
               W.key[slot] = rest;
               if constexpr (!packed) W.seg[slot] = (SegT)widx;
            }
         }
         wave_sync();
      }
      if (__ballot(overlap != 0) != 0) {
         if (lane == 0) atomicOr(&flags[0], FPL_WARN_OVERLAP);
         ov = true;
      }
   }
   wave_sync();
}

// the 16-bit storages (a double atan2 for the angle codes) out of line
__device__ __forceinline__ void store_record_short(const Outputs &o, int64_t r, int32_t lab,
                                                float x, float y, float z) {
   int4 e;
   rec::encode(o.rec_storage, lab, x, y, z, e);
   rec::store(o.rec_storage, o.rec_a, o.rec_b, r, e);
}

// bit 1 of flags[3]: some label is not 0 (the "All labels are 0" warning of
// Simpli.generate_tissue without a pass over the labels); one load per wave
__device__ __forceinline__ void note_nonzero(uint64_t any, uint32_t *flags) {
   if (any != 0 && (threadIdx.x & 63) == 0 &&
       (__atomic_load_n(&flags[3], __ATOMIC_RELAXED) & 2u) == 0)
      atomicOr(&flags[3], 2u);
}

// REC16: the records are in a 16-bit storage (a kernel variant of its own:
// the encoders' registers would make the f32 build spill)
// streaming (nontemporal) stores: nothing here reads the outputs again, and
// L2 keeps the segment records the tiles stage instead (the arrays and the
// records of 600 x 600 x 300 voxels, 3.5 GB, would otherwise pass through it)
__device__ __forceinline__ void write_arrays(const Outputs &o, const Params &pr, int vx, int vy,
                                             int vz, int32_t lab, const float vo[3],
                                             uint32_t *flags) {
   // the past-the-end layer index of the last bundle gives label n_props
   // (beyond the property table, which CheckInput rejects): bit 0 of flags[3]
   if (lab >= pr.n_props) atomicOr(&flags[3], 1u);
   if (o.label || o.vector) {
      const int64_t ind = ((int64_t)vx * pr.ny + vy) * pr.nz + vz;
      if (o.label) __builtin_nontemporal_store(lab, o.label + ind);
      if (o.vector) {
         typedef float v3f __attribute__((ext_vector_type(3), aligned(4)));
         const v3f e = {vo[0], vo[1], vo[2]};
         __builtin_nontemporal_store(e, (v3f *)(o.vector + 3 * ind));
      }
   }
}

template <bool REC16>
__device__ __forceinline__ void write_record(const Outputs &o, const Params &pr, int vx, int vy,
                                             int vz, int32_t lab, const float vo[3]) {
   const int64_t r = ((int64_t)vx * pr.nz + vz) * pr.ny + vy;
   if constexpr (REC16)
      store_record_short(o, r, lab, vo[0], vo[1], vo[2]);
   else
      rec::store(rec::ST_F32, o.rec_a, nullptr, r,
                 make_int4(lab, __float_as_int(vo[0]), __float_as_int(vo[1]),
                           __float_as_int(vo[2])));
}

template <bool REC16>
__device__ __forceinline__ void write_voxel(const Outputs &o, const Params &pr, int vx, int vy,
                                            int vz, int32_t lab, const float vo[3],
                                            uint32_t *flags) {
   write_arrays(o, pr, vx, vy, vz, lab, vo, flags);
   if (o.rec_a) write_record<REC16>(o, pr, vx, vy, vz, lab, vo);
}

// One workgroup per 8x8x32 tile, wave w owning the layers z = 4 l + w, l < 8
// (slot = l * 64 + column, column = x * 8 + y within the tile): a segment's
// candidates spread evenly over the waves, which meet at every chunk.  The tile's
// segments (global order) are staged in LDS in chunks of CHUNK.  Stage A: one
// lane per record computes the record's box within the wave's layers; the candidate
// voxels of all records are then enumerated 64 at a time in record order, and
// each is tested with a float lower bound of its squared distance (distance
// to the segment's box) against the outer-layer threshold and the voxel's
// stored distance, with margins that keep every candidate the exact test
// could accept; survivors enter the wave's queue in order.  Stage B
// (run_batch) runs the exact FP64 test on 64 queued candidates at a time.
// LSQ_IN_LDS: the layer tables fit the LDS copy; a compile-time choice keeps
// the layer searches plain LDS reads
template <typename SegT, bool LSQ_IN_LDS, bool REC16>
__global__ void __launch_bounds__(WG, sizeof(SegT) == 2 ? 6 : 4)
    k_voxelize(Params pr, const SegRec *__restrict__ segs, const float4 *__restrict__ seg_dir,
               const float4 *__restrict__ seg_clip,
               const uint32_t *__restrict__ list,
               const uint32_t *__restrict__ range, const int2 *__restrict__ tile_bundle,
               const BundleDev *__restrict__ bd,
               const double *__restrict__ layer_sq, const uint8_t *__restrict__ layer_model,
               Outputs out, uint32_t *flags) {
   __shared__ WaveState<SegT> wstate[WG / 64];
   __shared__ LdsSeg chunk[CHUNK];
   __shared__ uint32_t chunk_gidx[CHUNK]; // global segment index of each record

   const int n_tiles = pr.ntx * pr.nty * pr.ntz;
   const int t = tile_of_block((int)blockIdx.x, n_tiles);
   const int tz = t % pr.ntz;
   const int ty = (t / pr.ntz) % pr.nty;
   const int tx = t / (pr.ntz * pr.nty);
   // the wave index is uniform: keep it (and all loop bounds derived from it) scalar
   const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
   const int lane = threadIdx.x & 63;
   const int xt = tx * TX, yt = ty * TY;
   const int zt = tz * TZ, zw = zt + wave; // the wave's layers: z = zw + 4 * (0..7)
   const int max_layer = pr.max_layer;
   // float margins of the stage A test: relative 2^-12, absolute 2^-6
   const float rel = 1.0f + 0x1p-12f, absm = 0x1p-6f;

   // the layer tables in LDS when small (read through a generic pointer)
   __shared__ double lsq_lds[LSQ_LDS];
   __shared__ uint8_t lmodel_lds[LSQ_LDS];
   if (LSQ_IN_LDS)
      for (int i = threadIdx.x; i < pr.n_lsq; i += WG) {
         lsq_lds[i] = layer_sq[i];
         lmodel_lds[i] = layer_model[i];
      }
   const double *lsq = LSQ_IN_LDS ? (const double *)lsq_lds : layer_sq;
   const uint8_t *lmodel = LSQ_IN_LDS ? (const uint8_t *)lmodel_lds : layer_model;

   const uint32_t beg = range[2 * t], end = range[2 * t + 1];
   constexpr bool short_seg = sizeof(SegT) == 2;
   if (short_seg && end - beg > SHORT_LIST) {
      // this tile's list does not fit 16-bit positions: the host reruns the
      // volume with 32-bit winner indices (uniform over the workgroup)
      if (threadIdx.x == 0) atomicOr(&flags[1], 1u);
      return;
   }
   // the list entries of this thread's (at most two) staging pieces of the
   // next chunk, read one chunk ahead (the first chunk's before the set-up
   // below): staging then waits for the record loads only
   static_assert(CHUNK * 8 <= 2 * WG, "two staging pieces per thread");
   uint32_t ahead0 = 0, ahead1 = 0;
   auto read_ahead = [&](uint32_t nb) {
      const int nn = nb < end ? (int)min(end - nb, (uint32_t)CHUNK) : 0;
      if ((int)threadIdx.x < nn * 8) ahead0 = list[nb + (threadIdx.x >> 3)];
      if ((int)threadIdx.x + WG < nn * 8) ahead1 = list[nb + ((threadIdx.x + WG) >> 3)];
   };
   read_ahead(beg);
   WaveState<SegT> &W = wstate[wave];
   for (int s = lane; s < SLOTS; s += 64) {
      W.key[s] = KEY_EMPTY;
      if constexpr (!short_seg) W.seg[s] = (SegT)-1;
   }
   // one read of the warning for the workgroup: the mode below must be the
   // same in all its waves, whose keys the final pass reads across waves
   __shared__ uint32_t wg_warn;
   if (threadIdx.x == 0) wg_warn = __atomic_load_n(&flags[0], __ATOMIC_RELAXED);
   __syncthreads();
   bool ov = uniform((int)((wg_warn & FPL_WARN_OVERLAP) != 0)) != 0;
   // free mode (see the key layout above): the warning is raised already, or
   // the list (global segment order, bundle-major) starts and ends in one
   // bundle, so no update here can replace another bundle's label (but for
   // the bundle's past-the-end label: stage B flags it, and the host reruns
   // the volume with pr.single_free = 0)
   bool free_mode = false;
   if (short_seg) {
      bool single = true;
      if (end > beg) {
         const int2 tb = tile_bundle[t];
         single = tb.x == tb.y;
      }
      free_mode = uniform((int)(ov || (single && pr.single_free))) != 0;
   }
   int q_head = 0, q_tail = 0;
   wave_sync();

   for (uint32_t base = beg; base < end; base += CHUNK) {
      const int n = (int)((end - base) < (uint32_t)CHUNK ? (end - base) : (uint32_t)CHUNK);
      const uint32_t pos1 = base - beg + 1u; // list position + 1 of the chunk's first record
      __syncthreads(); // every wave is done with the previous chunk
      for (int e = threadIdx.x, r = 0; e < n * 8; e += WG, r++) {
         const int c = e >> 3, piece = e & 7;
         const uint32_t sidx = r == 0 ? ahead0 : ahead1;
         ((int4 *)&chunk[c].r)[piece] = ((const int4 *)&segs[sidx])[piece];
         if (piece == 0) chunk_gidx[c] = sidx;
      }
      read_ahead(base + CHUNK);
      __syncthreads();

      // Stage A set-up, one lane per record of the chunk: the record's candidate
      // region in the tile (tile_box) restricted to the wave's layers.  (Each
      // wave repeats the set-up: one lane per record, off the chunk's critical
      // path.)
      int nl = 0, la = 0;
      if (lane < n) {
         const TileBox tb = tile_box(chunk[lane].r, xt, yt, zt);
         // the wave's layers in [za, zb): z = zw + 4 * l, l in [la, la + nl)
         la = (tb.za - zw + 3) >> 2;
         nl = max(((tb.zb - zw + 3) >> 2) - la, 0);
         if (free_mode && nl > 0) {
            // no voxel of the record's box in this tile lies in the segment's
            // slab (a is linear: its extremes are at corners)
            const uint32_t gi = chunk_gidx[lane];
            const float4 c0 = seg_clip[2 * gi], c1 = seg_clip[2 * gi + 1];
            float amin = __builtin_huge_valf(), amax = -__builtin_huge_valf();
#pragma unroll
            for (int k = 0; k < 8; k++) {
               const float x = (float)((k & 1) ? tb.cx1 - 1 : tb.cx0) - c1.y;
               const float y = (float)((k & 2) ? tb.cy1 - 1 : tb.cy0) - c1.z;
               const float z = (float)((k & 4) ? tb.zb - 1 : tb.za) - c1.w;
               const float a = fmaf(x, c0.x, fmaf(y, c0.y, z * c0.z));
               amin = fminf(amin, a);
               amax = fmaxf(amax, a);
            }
            if (amax < c0.w - 0x1p-6f || amin > c1.x + 0x1p-6f) nl = 0;
         }
      }
      // (record, layer) pairs in record order: exclusive prefix of the counts
      const int incl = wave_scan_add(nl);
      const int PL = incl - nl;
      const int n_pairs = __builtin_amdgcn_readlane(incl, 63);
      const int rec_info = PL << 3 | la;
      int rcarry = 0; // 1 + the record holding the pair before the group

      for (int gb = 0; gb < n_pairs; gb += 64) {
         // Pair set-up, one lane per (record, layer) pair: the columns of the
         // layer that can pass, a rectangle.  A voxel that passes the float
         // test (lb = RN(ez^2 + RN(ey^2 + RN(ex^2))) <= thr, all terms >= 0)
         // has ex^2 + ey^2 + ez^2 <= thr (1 + 2^-22), so ex^2 <= Rx =
         // thr (1 + 2^-20) - eym^2 - ez^2 (computed in double: the squares of
         // floats are exact, the rest is covered by the margin), with eym the
         // least ey of the tile's in-box rows; likewise for ey.  hx covers
         // sqrt(Rx) with a relative margin that also covers the float rounding
         // of ex = RN(bl - x) or RN(x - bh).
         const int p = gb + lane;
         const int c = lane_owner(W.own, lane, nl > 0, PL, gb, rcarry);
         const int ri = __shfl(rec_info, c);
         const int rc = c; // the record in the chunk
         int cnt = 0;
         uint32_t geo = 0, magic = 0;
         if (p < n_pairs) {
            // the slab of the record, requested before the rectangle's arithmetic
            float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0;
            if (free_mode) {
               const uint32_t gi = chunk_gidx[rc];
               c0 = seg_clip[2 * gi];
               c1 = seg_clip[2 * gi + 1];
            }
            const SegRec &R = chunk[rc].r;
            const int l = (ri & 7) + (p - (ri >> 3));
            const int cx0 = max(R.lo[0], xt), cx1 = min(R.hi[0], xt + TX);
            const int cy0 = max(R.lo[1], yt), cy1 = min(R.hi[1], yt + TY);
            const float exm =
                fmaxf(fmaxf(R.bl[0] - (float)(cx1 - 1), (float)cx0 - R.bh[0]), 0.0f);
            const float eym =
                fmaxf(fmaxf(R.bl[1] - (float)(cy1 - 1), (float)cy0 - R.bh[1]), 0.0f);
            const float zf = (float)(zw + 4 * l);
            const float ez = fmaxf(fmaxf(R.bl[2] - zf, zf - R.bh[2]), 0.0f);
            const double tb = (double)R.thr * (1.0 + 0x1p-20) - (double)ez * (double)ez;
            const double rx = tb - (double)eym * (double)eym;
            const double ry = tb - (double)exm * (double)exm;
            if (rx >= 0.0 && ry >= 0.0) {
               const float hx =
                   fmaf(__builtin_amdgcn_sqrtf((float)rx), 1.0f + 0x1p-10f, 0x1p-10f);
               const float hy =
                   fmaf(__builtin_amdgcn_sqrtf((float)ry), 1.0f + 0x1p-10f, 0x1p-10f);
               int xa = max(cx0, (int)ceil((double)R.bl[0] - (double)hx));
               int xb = min(cx1, (int)floor((double)R.bh[0] + (double)hx) + 1);
               int ya = max(cy0, (int)ceil((double)R.bl[1] - (double)hy));
               int yb = min(cy1, (int)floor((double)R.bh[1] + (double)hy) + 1);
               if (free_mode && xa < xb && ya < yb) {
                  // the columns and rows whose voxels can lie in the segment's slab
                  // a in [a_lo, a_hi] (above); a = u.x x + u.y y + a0 is linear
                  const float a0 = (zf - c1.w) * c0.z - c1.y * c0.x - c1.z * c0.y;
                  clip_axis(c0.x, c0.y, a0, c0.w, c1.x, xa, xb, ya, yb);
                  clip_axis(c0.y, c0.x, a0, c0.w, c1.x, ya, yb, xa, xb);
               }
               if (xa < xb && ya < yb) {
                  const int ncy = yb - ya;
                  cnt = (xb - xa) * ncy;
                  geo = (uint32_t)(xa - xt) | (uint32_t)(ya - yt) << 3 | (uint32_t)l << 6 |
                        (uint32_t)(ncy - 1) << 9;
                  // floor(i / n) = (i * ceil(2^15 / n)) >> 15 for i < 512, n <= 8;
                  // 2^15 rcp(n) is within 2^15 / n * 2^-23 < 0.002 of 2^15 / n, which
                  // is an integer (n = 1, 2, 4, 8) or 0.14 or more from one
                  const uint32_t cdiv = (uint32_t)ceilf(
                      32768.0f * __builtin_amdgcn_rcpf((float)ncy));
                  magic = cdiv | (uint32_t)rc << 16;
               }
            }
         }
         // candidates of the group in pair order
         const int ci = wave_scan_add(cnt);
         const int P = ci - cnt;
         const int total = __builtin_amdgcn_readlane(ci, 63);
         W.tab[lane] = make_uint2((uint32_t)P | geo << 13, magic);
         wave_sync();

         // Stage A test, 64 candidates per window
         int kcarry = 0; // 1 + the pair holding the entry before the window
         for (int wb = 0; wb < total; wb += 64) {
            const int e = wb + lane;
            const int k = lane_owner(W.own, lane, cnt > 0, P, wb, kcarry);
            bool pass = false;
            uint32_t entry = 0;
            if (e < total) {
               const uint2 tbk = W.tab[k];
               const uint32_t g = tbk.x >> 13;
               const uint32_t li = (uint32_t)e - (tbk.x & 0x1fffu);
               const uint32_t ix = __umul24(li, tbk.y & 0xffffu) >> 15;
               const uint32_t iy = li - __umul24(ix, ((g >> 9) & 7) + 1);
               const uint32_t lx = (g & 7) + ix, ly = ((g >> 3) & 7) + iy, lz = (g >> 6) & 7;
               const uint32_t slot = lz << 6 | lx << 3 | ly;
               const uint32_t rc = tbk.y >> 16;
               const SegRec &R =
                   *(const SegRec *)((const char *)chunk + __umul24(rc, (uint32_t)sizeof(LdsSeg)));
               const float xf = (float)(xt + (int)lx), yf = (float)(yt + (int)ly);
               const float zf = (float)(zw + 4 * (int)lz);
               const float ex = fmaxf(fmaxf(R.bl[0] - xf, xf - R.bh[0]), 0.0f);
               const float ey = fmaxf(fmaxf(R.bl[1] - yf, yf - R.bh[1]), 0.0f);
               const float ez = fmaxf(fmaxf(R.bl[2] - zf, zf - R.bh[2]), 0.0f);
               const float lb = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
               const float stored = __uint_as_float((uint32_t)(W.key[slot] >> 32));
               pass = (lb <= R.thr) & (lb <= fmaf(stored, rel, absm));
               entry = rc << 9 | slot;
            }
            const unsigned long long m = __ballot(pass);
            if (m) {
               const int pos = q_tail + (int)__builtin_amdgcn_mbcnt_hi(
                                            (uint32_t)(m >> 32),
                                            __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
               if (pass) W.queue[pos & (QN - 1)] = entry;
               q_tail += __popcll(m);
               if (q_tail - q_head >= 64) {
                  wave_sync();
                  run_batch(W, chunk, chunk_gidx, pos1, 64, q_head, lane, xt, yt, zw,
                            max_layer, lsq, lmodel, flags, ov, free_mode);
                  q_head += 64;
               }
            }
         }
         wave_sync(); // the table is rewritten by the next group
      }
      // drain: the queue refers to this chunk's records
      wave_sync();
      while (q_tail > q_head) {
         const int nb = min(64, q_tail - q_head);
         run_batch(W, chunk, chunk_gidx, pos1, nb, q_head, lane, xt, yt, zw, max_layer, lsq, lmodel,
                   flags, ov, free_mode);
         q_head += nb;
      }
      if (!ov && !free_mode) // (the free mode does not look at the warning)
         ov = uniform((int)((__atomic_load_n(&flags[0], __ATOMIC_RELAXED) & FPL_WARN_OVERLAP) !=
                            0)) != 0;
   }

   // the results: wave w writes the layers z = zt + 8 w + (0..7) of the tile
   // (lane: y = lane >> 3, z offset lane & 7), each voxel's key read from the
   // wave that owns its layer, so that a column's eight labels and axes are
   // contiguous; the keys are final once every wave is here
   __syncthreads();
   const int zo = 8 * wave + (lane & 7);
   const WaveState<SegT> &WS = wstate[zo & 3];
   const int lbase = (zo >> 2) * 64 + (lane >> 3);
   const bool need_vec = out.vector != nullptr || out.rec_a != nullptr;
   uint64_t nonzero = 0; // wave-uniform: some voxel of the wave got a label != 0
   if (short_seg && free_mode) {
      // Results of the free mode.  The label / axis arrays ([x][y][z]) want
      // whole z columns in one store, the march records ([x][z][y]) whole y
      // rows: wave w resolves the two x rows xt + 2 w, xt + 2 w + 1 of the tile
      // with lane = z (0..31) + 32 (y & 1), eight voxels per lane, and stores
      // labels and axes as 128-byte runs; then, after a barrier (no key is
      // read again), each x row goes through the wave's own key array (4 KB)
      // to lane = y (0..7) + 8 (z & 7) and the records are stored as whole
      // 128-byte rows.  The label and layer model come from the key; a 'p'
      // voxel reads its winner's direction (list entry, then table), an 'r'
      // voxel its points.
      const int lz = lane & 31, yp = lane >> 5;
      const WaveState<SegT> &KS = wstate[lz & 3]; // the wave that owns layer lz
      int32_t sgs[8];
      uint32_t los[8];
#pragma unroll
      for (int i = 0; i < 8; i++) { // voxel (2 wave + (i & 1), 2 (i >> 1) + yp, lz)
         const int lx = 2 * wave + (i & 1), ly = 2 * (i >> 1) + yp;
         const unsigned long long key = KS.key[(lz >> 2) * 64 + lx * 8 + ly];
         const bool hit = (uint32_t)(key >> 32) != 0x7F800000u;
         los[i] = hit ? (uint32_t)key : 0u;
         sgs[i] = -1;
         if (need_vec && ((los[i] >> 14) & 3u) != 0u)
            sgs[i] = (int32_t)list[beg + (0x10000u - (los[i] >> 16)) - 1u];
      }
      // every load before the first store: on this part a load issued after a
      // store also waits for the store's completion
      int32_t labs[8];
      float vos[8][3];
      const int vz = zt + lz;
#pragma unroll
      for (int i = 0; i < 8; i++) {
         const int vx = xt + 2 * wave + (i & 1), vy = yt + 2 * (i >> 1) + yp;
         const uint32_t lo = los[i];
         labs[i] = (int32_t)(lo & 0x3FFFu);
         float *vo = vos[i];
         vo[0] = vo[1] = vo[2] = 0.0f;
         if (vx < pr.nx && vy < pr.ny && vz < pr.nz && need_vec) {
            const uint32_t m2 = (lo >> 14) & 3u;
            if (m2 == 1u) {
               const float4 d = seg_dir[sgs[i]];
               vo[0] = d.x;
               vo[1] = d.y;
               vo[2] = d.z;
            } else if (m2 == 2u) {
               const SegRec &R = segs[sgs[i]];
               const double pt[3] = {vx + 0.5, vy + 0.5, vz + 0.5};
               const double pp[3] = {R.p[0], R.p[1], R.p[2]};
               const double qq[3] = {R.q[0], R.q[1], R.q[2]};
               const SegExact E = seg_exact(pp, qq);
               const Closest cp = closest(pt, E);
               unit_to_float(pt[0] - cp.c[0], pt[1] - cp.c[1], pt[2] - cp.c[2], vo);
            }
         }
      }
      __syncthreads(); // every key is read: each wave's key array is free
#pragma unroll
      for (int i = 0; i < 8; i++) {
         const int vx = xt + 2 * wave + (i & 1), vy = yt + 2 * (i >> 1) + yp;
         if (vx < pr.nx && vy < pr.ny && vz < pr.nz) {
            write_arrays(out, pr, vx, vy, vz, labs[i], vos[i], flags);
            nonzero |= __ballot(labs[i] != 0);
         }
      }
      note_nonzero(nonzero, flags);
      if (out.rec_a) {
         // y-row order through LDS: voxel (y, z) of an x row at 16-byte slot
         // 8 z + (y ^ (z & 7)) (the xor spreads the column writes over the banks)
         uint4 *T = (uint4 *)wstate[wave].key;
         static_assert(sizeof(wstate[0].key) >= 256 * sizeof(uint4), "transpose buffer");
         const int ry = lane & 7, rz = lane >> 3;
#pragma unroll
         for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
               const int i = 2 * k + h, ly = 2 * k + yp;
               T[8 * lz + (ly ^ (lz & 7))] =
                   make_uint4((uint32_t)labs[i], __float_as_uint(vos[i][0]),
                              __float_as_uint(vos[i][1]), __float_as_uint(vos[i][2]));
            }
            wave_sync();
            const int vx = xt + 2 * wave + h, vy = yt + ry;
#pragma unroll
            for (int zb = 0; zb < 4; zb++) {
               const int z = 8 * zb + rz, vz2 = zt + z;
               const uint4 e = T[8 * z + (ry ^ (z & 7))];
               if (vx < pr.nx && vy < pr.ny && vz2 < pr.nz) {
                  const float vo[3] = {__uint_as_float(e.y), __uint_as_float(e.z),
                                       __uint_as_float(e.w)};
                  write_record<REC16>(out, pr, vx, vy, vz2, (int32_t)e.x, vo);
               }
            }
            wave_sync(); // the next x row overwrites the buffer
         }
      }
      return;
   }
   // Results of the ordered mode, eight columns of eight voxels per pass.  The
   // winners' records are fetched ahead: the global segment index (a load from
   // the tile list for 16-bit positions) two passes ahead, the record one pass
   // ahead.
   auto winner = [&](int j) -> int32_t {
      const int slot = lbase + j * 8;
      if constexpr (short_seg) {
         const unsigned long long key = WS.key[slot];
         const uint32_t w = (uint32_t)key & 0xFFFFu; // list position + 1 (0: none)
         return w ? (int32_t)list[beg + w - 1u] : -1;
      } else {
         return (int32_t)WS.seg[slot];
      }
   };
   int32_t sg = winner(0), sg1 = winner(1);
   WinRec cur, nxt;
   if (sg >= 0) cur = load_winner(segs[sg], seg_dir[sg]);
   for (int j = 0; j < 8; j++) {
      const int slot = lbase + j * 8;
      const int32_t sg2 = j < 6 ? winner(j + 2) : -1;
      if (j < 7 && sg1 >= 0) nxt = load_winner(segs[sg1], seg_dir[sg1]);
      const int vx = xt + j, vy = yt + (lane >> 3), vz = zt + zo;
      if (vx < pr.nx && vy < pr.ny && vz < pr.nz) {
         const unsigned long long key = WS.key[slot];
         int32_t lab = (int32_t)(uint32_t)(key & 0x7FFFFFFFull);
         if (short_seg) {
            const uint32_t l15 = (uint32_t)(key >> 16) & 0x7FFFu;
            lab = l15 == LABEL15_LATER ? LABEL_LATER : (int32_t)l15;
         }
         const double pt[3] = {vx + 0.5, vy + 0.5, vz + 0.5};
         if (sg >= 0 && lab == LABEL_LATER) {
            // the winner's label from its exact test: the stored float distance
            // and the closest point's parameter, computed as in that test
            const double t = closest_t(pt, cur.p, cur.q);
            const double rr = cur.rp * (1 - t) + cur.rq * t;
            lab = layer_label(__uint_as_float((uint32_t)(key >> 32)), (float)(rr * rr), lsq,
                              cur.layer_off, cur.bundle, max_layer);
         }
         float vo[3] = {0.0f, 0.0f, 0.0f};
         if (need_vec) {
            if (sg >= 0) {
               // a float ratio just above the last scale^2 gives the past-the-end
               // layer index; the padded model table repeats the outermost
               // layer's orientation there
               const int ly = lab - (1 + cur.bundle * max_layer);
               const uint8_t m = lmodel[cur.layer_off + ly];
               if (m == 'p') {
                  vo[0] = cur.dir[0];
                  vo[1] = cur.dir[1];
                  vo[2] = cur.dir[2];
               } else if (m == 'r') {
                  const SegExact E = seg_exact(cur.p, cur.q);
                  const Closest cp = closest(pt, E);
                  unit_to_float(pt[0] - cp.c[0], pt[1] - cp.c[1], pt[2] - cp.c[2], vo);
               }
            }
         }
         write_voxel<REC16>(out, pr, vx, vy, vz, lab, vo, flags);
         nonzero |= __ballot(lab != 0);
      }
      cur = nxt;
      sg = sg1;
      sg1 = sg2;
   }
   note_nonzero(nonzero, flags);
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// A scene hint with this bit: its tile lists are too long for the per-tile sort
constexpr int64_t LONG_LISTS = (int64_t)1 << 62;

// One generation.  use_hint: the radix sort sizes its (tile, segment) pair
// arrays by the previous call's pair count on this stream instead of reading
// the count back in the middle of the call (a host round trip that leaves the
// GPU idle); force_radix: bin with the radix sort.  `overflow`: the radix
// sort's pairs exceeded the count this call assumed; `long_lists`: a tile list
// exceeded the fixed-size lists of the binning by atomics.  The caller runs
// again in either case (the outputs are incomplete), with the hint updated.
Status run_once(const fpl_volume *vol, const fpl_fiber_bundles *fbs, const Outputs &out,
                double *props, uint32_t *warn, int32_t *label_max, hipStream_t st, bool use_hint,
                bool force_radix, bool &overflow, bool &long_lists) {
   overflow = false;
   long_lists = false;
   if (!vol || !fbs) return make_status(FPL_E_INVALID, "null argument");
   const int64_t nx = vol->dim[0], ny = vol->dim[1], nz = vol->dim[2];
   if (nx <= 0 || ny <= 0 || nz <= 0)
      return make_status(FPL_E_INVALID, "dim.global[any] <= 0: [%lld,%lld,%lld]", (long long)nx,
                         (long long)ny, (long long)nz);
   if (nx >= (1 << 24) || ny >= (1 << 24) || nz >= (1 << 24))
      return make_status(FPL_E_INVALID, "volume dimension >= 2^24 is not supported");
   if (!(vol->voxel_size > 0))
      return make_status(FPL_E_INVALID, "voxel_size <= 0: %f", vol->voxel_size);
   if (!out.label && !out.rec_a) return make_status(FPL_E_INVALID, "output pointers are null");
   if (fbs->n_bundles < 0 || fbs->n_fibers < 0 || fbs->n_points < 0)
      return make_status(FPL_E_INVALID, "negative sizes");
   if (fbs->n_points >= ((int64_t)1 << 31))
      return make_status(FPL_E_INVALID, "too many fiber points: %lld", (long long)fbs->n_points);
   if (fbs->bundle_ptr[fbs->n_bundles] != fbs->n_fibers)
      return make_status(FPL_E_INVALID, "fbs and prs not the same size");

   // ---- host tables: bundles, sorted layers, property list
   const int64_t nb = fbs->n_bundles;
   int32_t max_layer = 0;
   std::vector<BundleDev> bdev(nb > 0 ? nb : 1);
   std::vector<double> lsq;
   std::vector<uint8_t> lmodel;
   std::vector<double> ldn, lmu;
   for (int64_t b = 0; b < nb; b++) {
      std::vector<int64_t> order;
      for (int64_t l = fbs->layer_ptr[b]; l < fbs->layer_ptr[b + 1]; l++) order.push_back(l);
      // fiber::Bundle sorts layers by scale (fiber_bundle.cpp:83-86)
      std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t c) {
         return fbs->layer_scale[a] < fbs->layer_scale[c];
      });
      BundleDev &B = bdev[b];
      B.layer_off = (int32_t)lsq.size();
      B.n_layers = (int32_t)order.size();
      for (int64_t l : order) {
         const uint8_t m = fbs->layer_model[l];
         if (m != 'b' && m != 'p' && m != 'r')
            return make_status(FPL_E_INVALID, "Orientation must be \"b\", \"p\", or \"r\"");
         lsq.push_back(fbs->layer_scale[l] * fbs->layer_scale[l]);
         lmodel.push_back(m);
         ldn.push_back(fbs->layer_dn[l]);
         lmu.push_back(fbs->layer_mu[l]);
      }
      if (B.n_layers > max_layer) max_layer = B.n_layers;
      const double last = B.n_layers ? lsq.back() : 0.0;
      B.f_last_sq = (float)last;
      B.sqrt_last = std::sqrt(last);
      B.fiber_begin = fbs->bundle_ptr[b];
      B.fiber_end = fbs->bundle_ptr[b + 1];
      // one padding row per bundle: a +inf scale^2 ends the layer search and
      // the past-the-end layer index reads the outermost layer's orientation
      lsq.push_back(INFINITY);
      lmodel.push_back(B.n_layers ? lmodel.back() : (uint8_t)'b');
      ldn.push_back(0.0);
      lmu.push_back(0.0);
   }
   for (int64_t b = 0; b < nb; b++) bdev[b].label_base = (int32_t)(1 + b * max_layer);
   if (lsq.empty()) {
      lsq.push_back(0.0);
      lmodel.push_back('b');
   }
   // GetPropertyList (generator.cpp:496-526)
   const int64_t n_props = nb * max_layer + 1;
   if (n_props >= 0x7FFFFFFF) // labels are int32; 0x7FFFFFFF is an internal marker
      return make_status(FPL_E_INVALID, "too many labels: %lld", (long long)n_props);
   if (props) {
      for (int64_t i = 0; i < 2 * n_props; i++) props[i] = 0.0;
      for (int64_t b = 0; b < nb; b++)
         for (int l = 0; l < bdev[b].n_layers; l++) {
            const int64_t id = 1 + l + b * max_layer;
            const int64_t src = bdev[b].layer_off + l;
            props[2 * id + 1] = lmu[src];
            props[2 * id + 0] = lmodel[src] == 'b' ? 0.0 : ldn[src];
         }
   }

   Params pr;
   pr.nx = nx;
   pr.ny = ny;
   pr.nz = nz;
   pr.n_points = fbs->n_points;
   pr.n_fibers = fbs->n_fibers;
   pr.n_bundles = nb;
   pr.shift = (vol->origin[0] != 0 || vol->origin[1] != 0 || vol->origin[2] != 0);
   for (int k = 0; k < 3; k++) pr.off[k] = -vol->origin[k];
   pr.scale = 1.0 / vol->voxel_size;
   pr.max_layer = max_layer;
   pr.n_lsq = (int32_t)lsq.size();
   pr.n_props = (int32_t)n_props;
   pr.single_free = 1;
   pr.ntx = (int32_t)((nx + TX - 1) / TX);
   pr.nty = (int32_t)((ny + TY - 1) / TY);
   pr.ntz = (int32_t)((nz + TZ - 1) / TZ);
   const int64_t n_tiles = (int64_t)pr.ntx * pr.nty * pr.ntz;
   if (n_tiles >= ((int64_t)1 << 31))
      return make_status(FPL_E_INVALID, "volume too large: %lld tiles", (long long)n_tiles);

   // ---- small tables to the device
   const size_t sz_bd = sizeof(BundleDev) * bdev.size();
   const size_t sz_lsq = sizeof(double) * lsq.size();
   const size_t off_lsq = (sz_bd + 255) & ~size_t(255);
   const size_t off_lm = (off_lsq + sz_lsq + 255) & ~size_t(255);
   const size_t small_bytes = off_lm + lmodel.size();
   void *small = nullptr, *stage = nullptr;
   FPL_TRY(ws_get(st, WS_SMALL, small_bytes, &small));
   FPL_TRY(host_stage(st, small_bytes + 64, &stage));
   FPL_HIP(hipStreamSynchronize(st)); // stage may still be read by a previous copy
   memcpy((char *)stage, bdev.data(), sz_bd);
   memcpy((char *)stage + off_lsq, lsq.data(), sz_lsq);
   memcpy((char *)stage + off_lm, lmodel.data(), lmodel.size());
   FPL_HIP(hipMemcpyAsync(small, stage, small_bytes, hipMemcpyHostToDevice, st));
   const BundleDev *d_bd = (const BundleDev *)small;
   const double *d_lsq = (const double *)((char *)small + off_lsq);
   const uint8_t *d_lm = (const uint8_t *)((char *)small + off_lm);

   // ---- scratch
   const int64_t np = fbs->n_points, nf = fbs->n_fibers;
   void *boxes, *fok, *segs, *cnt, *off, *flags;
   FPL_TRY(ws_get(st, WS_BOXES, sizeof(double) * 6 * (nf + nb + 1), &boxes));
   FPL_TRY(ws_get(st, WS_FIBER_OK, nf + 1, &fok));
   FPL_TRY(ws_get(st, WS_SEGS, (sizeof(SegRec) + 3 * sizeof(float4)) * (np + 1), &segs));
   float4 *seg_dir = (float4 *)((SegRec *)segs + (np + 1));
   float4 *seg_clip = seg_dir + (np + 1);
   FPL_TRY(ws_get(st, WS_SEG_COUNT, sizeof(int64_t) * (np + 1), &cnt));
   FPL_TRY(ws_get(st, WS_SEG_OFF, sizeof(int64_t) * (np + 2), &off));
   FPL_TRY(ws_get(st, WS_FLAGS, 64, &flags));
   double *fiber_box = (double *)boxes;
   double *bundle_box = fiber_box + 6 * nf;
   uint32_t *d_flags = (uint32_t *)flags;

   const int B = 256;
   // The binning by atomics writes fixed-size tile lists (SORT_MAX entries per
   // tile: no count to read back); a scene whose hint has bit 62 set (its
   // tile lists are too long for them), or too many tiles for 32-bit list
   // positions, is binned by the radix sort, whose pair arrays are sized by
   // the previous call's pair count on this stream (the rest of the hint)
   // instead of a read-back of the count in the middle of the call; a hint
   // above 64 tiles per point (a much larger previous scene) would make this
   // call pad and sort mostly padding: it reads its count back instead.  The
   // workspace of the binning is taken before the first launch, so that the
   // small kernels go out back to back.
   const std::array<int64_t, 5> scene{fbs->n_points, fbs->n_fibers, nx, ny, nz};
   const int64_t raw_hint = pair_hint(st, scene);
   const bool radix = force_radix || (raw_hint & LONG_LISTS) != 0 ||
                      n_tiles > ((int64_t)1 << 31) / SORT_MAX;
   int64_t hint = use_hint ? (raw_hint & ~LONG_LISTS) : 0;
   if (hint > 64 * fbs->n_points) hint = 0;
   void *ka, *va, *kb, *vb, *rng;
   uint32_t *tile_count = nullptr; // the binning by atomics: per-tile cursors
   if (!radix) {
      FPL_TRY(ws_get(st, WS_KEYS_A, sizeof(uint32_t) * n_tiles, &ka));
      FPL_TRY(ws_get(st, WS_VALS_B, sizeof(uint32_t) * SORT_MAX * n_tiles, &vb));
      tile_count = (uint32_t *)ka;
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n_tiles, B)), dim3(B), 0, st, tile_count,
                         n_tiles, 0u);
   }
   FPL_TRY(ws_get(st, WS_TILE_RANGE, sizeof(uint32_t) * 4 * n_tiles, &rng));
   if (nf > 0)
      hipLaunchKernelGGL(k_fiber_boxes, dim3(grid_for(nf, B / 64)), dim3(B), 0, st, pr,
                         fbs->points, fbs->fiber_ptr, fiber_box);
   hipLaunchKernelGGL(k_bundle_boxes, dim3((unsigned)(nb > 0 ? nb : 1)), dim3(B), 0, st, pr,
                      d_bd, fiber_box, bundle_box, d_flags);
   if (nf > 0) {
      hipLaunchKernelGGL(k_fibers, dim3(grid_for(nf, B)), dim3(B), 0, st, pr, fbs->fiber_ptr, d_bd,
                         fiber_box, bundle_box, (uint8_t *)fok, d_flags);
   }
   if (np > 0)
      hipLaunchKernelGGL(k_segments, dim3(grid_for(np, B)), dim3(B), 0, st, pr, fbs->points,
                         fbs->fiber_ptr, d_bd, (const uint8_t *)fok, (SegRec *)segs,
                         seg_dir, seg_clip, (int64_t *)cnt, tile_count, (uint32_t *)vb);
   int2 *tile_bundle = (int2 *)((uint32_t *)rng + 2 * n_tiles); // set with the ranges
   if (!radix) {
      if (np > 0) {
         hipLaunchKernelGGL(k_tile_sort, dim3((unsigned)((n_tiles + 3) / 4)), dim3(B), 0, st,
                            (uint32_t)n_tiles, (const uint32_t *)tile_count, (uint32_t *)vb,
                            (const SegRec *)segs, (uint32_t *)rng, tile_bundle, d_flags);
      } else {
         hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2 * n_tiles, B)), dim3(B), 0, st,
                            (uint32_t *)rng, 2 * n_tiles, 0u);
      }
   } else {
      // exclusive scan of the per-segment counts: the segments' pair positions
      int64_t total = 0;
      if (np > 0) {
         size_t tmp_bytes = 0;
         void *tmp;
         FPL_HIP(hipMemsetAsync((int64_t *)cnt + np, 0, sizeof(int64_t), st));
         FPL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, (int64_t *)cnt,
                                                  (int64_t *)off, (int)(np + 1), st));
         FPL_TRY(ws_get(st, WS_CUB_TEMP, tmp_bytes, &tmp));
         FPL_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, (int64_t *)cnt, (int64_t *)off,
                                                  (int)(np + 1), st));
         if (hint > 0) {
            total = hint; // k_pad reports a larger count
         } else {
            FPL_HIP(hipMemcpyAsync(stage, (int64_t *)off + np, sizeof(int64_t),
                                   hipMemcpyDeviceToHost, st));
            FPL_HIP(hipStreamSynchronize(st));
            total = *(int64_t *)stage;
         }
      }
      if (total >= ((int64_t)1 << 31))
         return make_status(FPL_E_INVALID, "too many segment/tile pairs: %lld", (long long)total);
      FPL_TRY(ws_get(st, WS_VALS_B, sizeof(uint32_t) * (total + 1), &vb));
      FPL_TRY(ws_get(st, WS_KEYS_A, sizeof(uint32_t) * (total + 1), &ka));
      FPL_TRY(ws_get(st, WS_VALS_A, sizeof(uint32_t) * (total + 1), &va));
      FPL_TRY(ws_get(st, WS_KEYS_B, sizeof(uint32_t) * (total + 1), &kb));
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2 * n_tiles, B)), dim3(B), 0, st,
                         (uint32_t *)rng, 2 * n_tiles, 0u);
      if (np > 0)
         hipLaunchKernelGGL(k_pad, dim3(grid_for(total > 0 ? total : 1, B)), dim3(B), 0, st,
                            (const int64_t *)off + np, total, (uint32_t)n_tiles, (uint32_t *)ka,
                            (uint32_t *)va, d_flags);
      if (total > 0) {
         hipLaunchKernelGGL(k_emit, dim3(grid_for(np, B)), dim3(B), 0, st, pr,
                            (const SegRec *)segs, (const float4 *)seg_dir, (const int64_t *)cnt,
                            (const int64_t *)off, total, (uint32_t *)ka, (uint32_t *)va);
         // keys up to n_tiles (the padding key) take part in the sort
         int nbits = 1;
         while (((int64_t)1 << nbits) <= n_tiles) nbits++;
         size_t tmp_bytes = 0;
         FPL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (uint32_t *)ka,
                                                    (uint32_t *)kb, (uint32_t *)va,
                                                    (uint32_t *)vb, (int)total, 0, nbits, st));
         void *tmp;
         FPL_TRY(ws_get(st, WS_CUB_TEMP, tmp_bytes, &tmp));
         FPL_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, (uint32_t *)ka,
                                                    (uint32_t *)kb, (uint32_t *)va,
                                                    (uint32_t *)vb, (int)total, 0, nbits, st));
         hipLaunchKernelGGL(k_ranges, dim3(grid_for(total, B)), dim3(B), 0, st, (uint32_t *)kb,
                            (const uint32_t *)vb, total, (uint32_t)n_tiles,
                            (const SegRec *)segs, (uint32_t *)rng, tile_bundle);
      }
   }
   const bool lsq_lds = pr.n_lsq <= LSQ_LDS;
   const bool rec16 = out.rec_a && out.rec_storage != rec::ST_F32;
   // the voxelizer variant: 16- or 32-bit winner positions, layer tables in LDS
   // or not, 16-bit record storage or not
   auto voxelize = [&](bool short_pos) {
#define FPL_VOX(SEG, LDS, R16)                                                                    \
   hipLaunchKernelGGL((k_voxelize<SEG, LDS, R16>), dim3((unsigned)n_tiles), dim3(WG), 0, st, pr,  \
                      (const SegRec *)segs, seg_dir, seg_clip, (const uint32_t *)vb,               \
                      (const uint32_t *)rng, tile_bundle, d_bd, d_lsq, d_lm, out, d_flags)
#define FPL_VOX2(SEG, LDS)                                                                        \
   do {                                                                                           \
      if (rec16) FPL_VOX(SEG, LDS, true); else FPL_VOX(SEG, LDS, false);                          \
   } while (0)
      if (short_pos) {
         if (lsq_lds) FPL_VOX2(uint16_t, true); else FPL_VOX2(uint16_t, false);
      } else {
         if (lsq_lds) FPL_VOX2(int32_t, true); else FPL_VOX2(int32_t, false);
      }
#undef FPL_VOX2
#undef FPL_VOX
   };
   // the 16-bit build packs a winner's label into 14 bits of its key (labels
   // reach n_props when a float ratio rounds past the last layer)
   const bool short_labels = n_props <= 0x3FFF;
   bool rerun = !short_labels;
   // the binning's verdict, read with the voxelizer's flags (false: the
   // outputs are incomplete and the caller runs again): a tile with more
   // segments than the fixed-size lists hold, or radix-sort pairs beyond the
   // count this call assumed.  The radix sort's pair count is the scene's hint
   // for the next call.
   auto binned = [&](const uint32_t *fl) -> bool {
      if (np == 0) return true;
      long_lists = (fl[2] & 2u) != 0;
      overflow = (fl[2] & 1u) != 0;
      const int64_t pairs = radix ? (int64_t)((uint64_t)fl[4] | (uint64_t)fl[5] << 32) : 0;
      set_pair_hint(st, scene, pairs | ((radix || long_lists) ? LONG_LISTS : 0));
      return fl[2] == 0;
   };
   if (short_labels) {
      KernelTimer kt("voxelize", st);
      voxelize(true);
      kt.stop();
      FPL_HIP(hipGetLastError());
      FPL_HIP(hipMemcpyAsync(stage, d_flags, 6 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      FPL_HIP(hipStreamSynchronize(st));
      kt.collect();
      const uint32_t *f = (const uint32_t *)stage;
      if (!binned(f)) return Status();
      rerun = (f[1] & 1u) != 0;
      if (!rerun && (f[3] & 4u) && !(f[0] & FPL_WARN_OVERLAP)) {
         // a past-the-end label in a single-bundle tile run without order:
         // the overlap warning needs the update order, so the volume runs
         // again with those tiles in order (same voxels; warnings are an OR)
         pr.single_free = 0;
         KernelTimer kt3("voxelize", st);
         voxelize(true);
         kt3.stop();
         FPL_HIP(hipGetLastError());
         FPL_HIP(hipMemcpyAsync(stage, d_flags, 6 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
         FPL_HIP(hipStreamSynchronize(st));
         kt3.collect();
      }
   }
   if (rerun) {
      // a tile list longer than 65535 (or labels beyond 15 bits): the whole
      // volume with 32-bit winner indices (every voxel is rewritten; warnings
      // are an OR)
      KernelTimer kt2("voxelize", st);
      voxelize(false);
      kt2.stop();
      FPL_HIP(hipGetLastError());
      FPL_HIP(hipMemcpyAsync(stage, d_flags, 6 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      FPL_HIP(hipStreamSynchronize(st));
      kt2.collect();
      if (!binned((const uint32_t *)stage)) return Status();
   }
   if (warn) *warn = *(uint32_t *)stage;
   if (label_max) {
      // an upper bound of the labels written: n_props only in the past-the-end
      // case, 0 when every voxel is 0
      const uint32_t f3 = ((const uint32_t *)stage)[3];
      *label_max = (int32_t)((f3 & 1u) ? n_props : ((f3 & 2u) ? n_props - 1 : 0));
   }
   return Status();
}

size_t workspace_bytes(int64_t n_points, int64_t n_fibers, int64_t n_bundles,
                       const int64_t dim[3]) {
   const size_t np = (size_t)(n_points > 0 ? n_points : 0) + 2;
   const size_t n_tiles = (size_t)((dim[0] + TX - 1) / TX) * (size_t)((dim[1] + TY - 1) / TY) *
                          (size_t)((dim[2] + TZ - 1) / TZ);
   // segment records, directions, slab data, counts and offsets
   size_t b = (sizeof(SegRec) + 3 * sizeof(float4) + 2 * sizeof(int64_t)) * np;
   // fixed-size tile lists and cursors; per-tile ranges and end bundles
   b += (SORT_MAX + 1) * sizeof(uint32_t) * n_tiles + 4 * sizeof(uint32_t) * n_tiles;
   // the radix sort's four pair arrays and its temporary storage (about 8
   // pairs per segment)
   b += 8 * 24 * np;
   // fiber and bundle boxes, culling flags, small tables, scan temporaries
   b += 56 * (size_t)((n_fibers > 0 ? n_fibers : 0) + (n_bundles > 0 ? n_bundles : 0) + 1) +
        ((size_t)1 << 20);
   return b;
}

Status run(const fpl_volume *vol, const fpl_fiber_bundles *fbs, const Outputs &out,
           double *props, uint32_t *warn, int32_t *label_max, hipStream_t st) {
   // the binning by atomics first (unless the scene's hint says its lists are
   // too long); a call with a list longer than the fixed-size lists runs
   // again with the radix sort, a radix call whose pairs exceeded the count it
   // assumed runs again with its exact count (both leave the hint for the
   // next call)
   bool overflow = false, long_lists = false;
   FPL_TRY(run_once(vol, fbs, out, props, warn, label_max, st, true, false, overflow, long_lists));
   if (overflow)
      FPL_TRY(run_once(vol, fbs, out, props, warn, label_max, st, false, false, overflow,
                       long_lists));
   if (long_lists)
      FPL_TRY(run_once(vol, fbs, out, props, warn, label_max, st, true, true, overflow,
                       long_lists));
   if (overflow)
      FPL_TRY(run_once(vol, fbs, out, props, warn, label_max, st, false, true, overflow,
                       long_lists));
   return Status();
}

} // namespace gen
} // namespace fpl

extern "C" int fpl_generate(const fpl_volume *vol, const fpl_fiber_bundles *fbs, int only_label,
                            int32_t *label, float *vector, double *props, uint32_t *warn,
                            void *stream, char *msg, size_t msg_len) {
   if (warn) *warn = 0;
   fpl::Status s;
   if (!label || (!only_label && !vector)) {
      s = fpl::make_status(FPL_E_INVALID, "output pointers are null");
   } else {
      const fpl::gen::Outputs out{label, only_label ? nullptr : vector, nullptr, nullptr, 0};
      s = fpl::gen::run(vol, fbs, out, props, warn, nullptr, (hipStream_t)stream);
   }
   return fpl::finish(s, msg, msg_len);
}

extern "C" int fpl_generate_fields(const fpl_volume *vol, const fpl_fiber_bundles *fbs,
                                   int32_t *label, float *vector, double *props,
                                   fpl_fields *fields, uint32_t *warn, void *stream, char *msg,
                                   size_t msg_len) {
   if (warn) *warn = 0;
   fpl::Status s;
   do {
      if (!vol || !fields) {
         s = fpl::make_status(FPL_E_INVALID, "null argument");
         break;
      }
      const int32_t st = fields->storage;
      if (st < 0 || st > 2) {
         s = fpl::make_status(FPL_E_INVALID,
                              "field_storage must be 0 (f32), 1 (f16) or 2 (angle16)");
         break;
      }
      if (fields->records && st == fpl::rec::ST_ANGLE16 && !fields->codes) {
         s = fpl::make_status(FPL_E_INVALID, "the angle16 storage needs the codes array");
         break;
      }
      if (fbs && fields->records) {
         // labels reach n_props when a float ratio rounds past the last layer
         const int64_t np = fpl_num_properties(fbs);
         if (np > fpl::rec::label_limit(st)) {
            s = fpl::make_status(FPL_E_INVALID, "label %lld does not fit the %s field storage",
                                 (long long)np, st == fpl::rec::ST_F16 ? "16-bit" : "15-bit");
            break;
         }
      }
      for (int k = 0; k < 3; k++) fields->dim[k] = vol->dim[k];
      const fpl::gen::Outputs out{label, vector, fields->records, fields->codes, st};
      int32_t lmax = 0;
      s = fpl::gen::run(vol, fbs, out, props, warn, &lmax, (hipStream_t)stream);
      fields->label_min = 0;
      fields->label_max = lmax;
   } while (false);
   return fpl::finish(s, msg, msg_len);
}
