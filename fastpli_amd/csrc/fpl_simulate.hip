// Ray march (Stokes vector propagation through per-voxel retarders) for gfx950.
//
// Semantics: fastpli PliSimulator::RunSimulation (src/simulation/simulator.cpp:182-411).
// One thread per ray (CCD pixel); all tilts of a call run in one launch, the
// interpolating tilts of one ray tile in consecutive workgroups, the
// nearest-neighbour tilts (short work) after them.
//
// Formulation (exact in exact arithmetic, derived from simulator.cpp:207-214,
// 322-343, 355-360):
//   * the per-voxel update for rotation rho is R3(2 rho) M0 R3(-2 rho), R3 a
//     rotation of the (s1, s2) plane, and the input state (I0/2, 0, 0, -I0/2)
//     is R3-invariant, so s(rho) = R3(2 rho) s(0): one state per ray gives all
//     n_rho intensities I(rho) = 1/2 (s0 - cos 2rho s1 + sin 2rho s2);
//   * attenuation is a common scalar: s = A * (I0/2) * (1, u), A = exp(-sum mu t),
//     u a unit vector on the Poincare sphere rotated by angle -delta about the
//     in-plane axis n = (cos 2phi, -sin 2phi, 0) at every birefringent voxel;
//   * cos^2(alpha) = (vx^2 + vy^2) / |v|^2 and (cos 2phi, sin 2phi) =
//     (vx^2 - vy^2, 2 vx vy) / (vx^2 + vy^2) replace asin/atan2/cos/pow.
// Precision: u is accumulated in float-float (TwoSum) from FP32 increments
// (Rodrigues form, 1 - cos d without cancellation), sum mu t likewise; ray positions
// are stepped in FP64 exactly as the reference so that the visited voxels are
// identical.
#include <algorithm>
#include <mutex>
#include <map>
#include <cmath>
#include <cstring>
#include <vector>

#include <hip/hip_fp16.h>

#include "fpl_internal.h"
#include "fpl_records.h"
#include "fpl_sim.h"

namespace fpl {
namespace sim {

constexpr int WG = 256;
constexpr int64_t MAX_RAY_STEPS = (int64_t)1 << 24; // z-steps of one ray (prepare rejects more)
constexpr int MAX_LDS_PROPS = 128; // property tables up to this size are held in LDS
constexpr int MAX_RHO = 64;

struct PropDev {
   float K;     // pi/2 * dn * 4 * thickness / lambda  (delta per unit |v| cos^2 alpha)
   float mu_t;  // mu * 1e3 * thickness
   int32_t bg;  // dn == 0 or label 0 -> attenuation only
   int32_t pad;
};

struct TiltDev {
   double step[3];       // light_step = dir * step_size
   double half_step[3];  // light_step * 0.5
   double half_shift[2]; // 0.5 * shift (untilted sensor start)
   double shift[2];      // shift (tilted sensor start)
   double pc[3], pt[3], e1[3], e2[3], rinv[9];
   double a, b, c, d, D;
   float R[9];           // Mat3RotZYZ(phi, theta, -phi) for the optical axis
   int32_t rotate;       // theta != 0
   int32_t centers;      // every sample point is a voxel centre: no interpolation
   int32_t out;          // index of the tilt in the caller's order (image slot)
   int32_t trips;        // z-steps of every ray (tilt_steps)
};

// z-march trip count of one tilt (host)
int64_t tilt_steps(const TiltDev &T, int64_t nz);

struct Params {
   int64_t nx, ny, nz;
   int64_t n_props;
   int32_t n_tilts, n_rho;
   int32_t n_heavy;      // tilts [0, n_heavy) interpolate; they are launched first
   int32_t pad_;
   int32_t interp;       // 0 NN, 1 Lerp, 2 Slerp
   int32_t untilt;       // untilt_sensor_view
   int32_t flip_z;
   double quarter_I0;    // light_intensity / 4
   double c2r[MAX_RHO], s2r[MAX_RHO];
};

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ void two_sum(float &hi, float &lo, float b) {
   const float s = hi + b;
   const float bb = s - hi;
   const float err = (hi - (s - bb)) + (b - bb);
   hi = s;
   lo += err;
}

// The march reads the fields as [x][z][y] records (the storages of
// fpl_records.h): the 64 rays of a wave read one contiguous run per z-step.
using rec::ST_F32;
using rec::ST_F16;
using rec::ST_ANGLE16;
using rec::ANG_THETA;
using rec::ANG_PHI;

struct Vox {
   int32_t lab;
   float3 v;
   bool nn; // source axis failed |v|^2 - 1 <= 1e-5 (precision storages only)
};


template <typename Idx, int ST> struct Field {
   const void *a; // ST_F32: int4[], ST_F16: uint2[], ST_ANGLE16: uint16 labels
   const uint32_t *b; // ST_ANGLE16: packed angle codes
   int32_t nx, ny, nz;
   int32_t flip_z;

   __device__ __forceinline__ bool inside(int32_t x, int32_t y, int32_t z) const {
      return x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
   }
   // record of the voxel holding the sample point (px, py, pz) whose z floor
   // `fz` is inside [0, nz) when `live`; outside x / y it is background
   // (simulator.cpp:430-432).  One branch: all tests are evaluated.
   __device__ __forceinline__ Vox fetch(double px, double py, double fz, bool live) const;
   // The storage words of an in-range voxel, undecoded: ST_F32 (label, x, y, z),
   // ST_F16 (label | x << 16, y | z << 16), ST_ANGLE16 (label word, angle
   // codes).  flip_z mirrors z as GetLabel / GetVec do (simulator.cpp:427-428,
   // 559-560).  The march decodes the axis only where a step needs it.
   __device__ __forceinline__ uint4 raw(int32_t x, int32_t y, int32_t z) const {
      if (flip_z) z = nz - 1 - z;
      return load(((Idx)x * (Idx)nz + (Idx)z) * (Idx)ny + (Idx)y);
   }
   // the words of record i ([x][z][y] order)
   __device__ __forceinline__ uint4 load(Idx i) const {
      if (ST == ST_F32) {
         const int4 r = ((const int4 *)a)[i];
         return make_uint4((uint32_t)r.x, (uint32_t)r.y, (uint32_t)r.z, (uint32_t)r.w);
      } else if (ST == ST_F16) {
         const uint2 r = ((const uint2 *)a)[i];
         return make_uint4(r.x, r.y, 0u, 0u);
      }
      return make_uint4((uint32_t)((const uint16_t *)a)[i], b[i], 0u, 0u);
   }
   // Loads raw() of the voxel holding the sample point (px, py, pz) whose z
   // floor `fz` is inside [0, nz) when `live`, and returns whether the point is
   // inside the volume; outside x / y it is background (simulator.cpp:430-432)
   // and `r` holds voxel (0, 0, 0): the caller takes label 0 for it.  The load
   // is issued in every case and nothing is selected on its result here: a
   // load under a branch, or a value selected from it before the loop's next
   // turn, makes the compiler wait for every load in flight, which serialises
   // the loads the march keeps in flight.
   __device__ __forceinline__ bool fetch_raw(double px, double py, double fz, bool live,
                                             uint4 &r) const;
   __device__ static __forceinline__ int32_t label_of(const uint4 &r) {
      return ST == ST_F32 ? (int32_t)r.x : (int32_t)(r.x & (ST == ST_F16 ? 0xffffu : 0x7fffu));
   }
   // the precision storages' "not normalized" bit of the source axis
   __device__ static __forceinline__ bool nn_of(const uint4 &r) {
      return ST == ST_F32 ? false
                          : (ST == ST_F16 ? (r.y & 0x80000000u) != 0 : (r.x & 0x8000u) != 0);
   }
   // The optical axis.  ST_F16 with `renorm`: an axis that passed the
   // reference's normalisation test is scaled back to unit length after the
   // fp16 rounding.
   __device__ static __forceinline__ float3 axis_of(const uint4 &r, bool renorm) {
      if (ST == ST_F32) return make_float3(__uint_as_float(r.y), __uint_as_float(r.z),
                                           __uint_as_float(r.w));
      if (ST == ST_F16) {
         // z is stored >= 0; its sign bit is the "not normalized" bit
         float3 v = make_float3(__half2float(__ushort_as_half((unsigned short)(r.x >> 16))),
                                __half2float(__ushort_as_half((unsigned short)(r.y & 0xffffu))),
                                fabsf(__half2float(__ushort_as_half((unsigned short)(r.y >> 16)))));
         if (renorm && !nn_of(r) && !(v.x == 0.0f && v.y == 0.0f && v.z == 0.0f)) {
            const float inv = rsqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
            v = make_float3(v.x * inv, v.y * inv, v.z * inv);
         }
         return v;
      }
      const uint32_t tc = r.y & 0xffffu, pc = r.y >> 16;
      if (tc == 0xffffu) return make_float3(0.0f, 0.0f, 0.0f);
      const float th = (float)(tc * ANG_THETA), ph = (float)(pc * ANG_PHI) - 3.14159265f;
      const float st = __sinf(th);
      return make_float3(st * __cosf(ph), st * __sinf(ph), __cosf(th));
   }
   // decoded record of an in-range voxel
   __device__ __forceinline__ Vox at(int32_t x, int32_t y, int32_t z) const {
      const uint4 r = raw(x, y, z);
      return Vox{label_of(r), axis_of(r, true), nn_of(r)};
   }
};

// floor of a ray coordinate as int32; values far outside the volume saturate
// (they only ever select "outside -> background")
__device__ __forceinline__ int32_t ifloor(double v) {
   return (int32_t)fmin(fmax(floor(v), -1073741824.0), 1073741824.0);
}

template <typename Idx, int ST>
__device__ __forceinline__ Vox Field<Idx, ST>::fetch(double px, double py, double fz,
                                                     bool live) const {
   const int32_t jx = ifloor(px), jy = ifloor(py);
   const bool in = live & ((uint32_t)jx < (uint32_t)nx) & ((uint32_t)jy < (uint32_t)ny);
   Vox r{0, make_float3(0.0f, 0.0f, 0.0f), false};
   if (in) r = at(jx, jy, (int32_t)fz);
   return r;
}

template <typename Idx, int ST>
__device__ __forceinline__ bool Field<Idx, ST>::fetch_raw(double px, double py, double fz,
                                                         bool live, uint4 &r) const {
   const int32_t jx = ifloor(px), jy = ifloor(py);
   const bool in = live & ((uint32_t)jx < (uint32_t)nx) & ((uint32_t)jy < (uint32_t)ny);
   r = raw(in ? jx : 0, in ? jy : 0, in ? (int32_t)fz : (flip_z ? nz - 1 : 0));
   return in;
}

__device__ __forceinline__ bool is_zero(const float3 &v) {
   return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f;
}
__device__ __forceinline__ float dot3(const float3 &a, const float3 &b) {
   return a.x * b.x + a.y * b.y + a.z * b.z;
}

// LinearOrientationInterpolation (simulator.cpp:449-471)
__device__ __forceinline__ float3 lerp_orient(const float3 &v, const float3 &u, float t) {
   if (is_zero(v)) return u;
   if (is_zero(u)) return v;
   const float3 a = make_float3(v.x * (1.0f - t), v.y * (1.0f - t), v.z * (1.0f - t));
   const float3 b = make_float3(u.x * t, u.y * t, u.z * t);
   if (dot3(a, b) >= 0.0f) return make_float3(a.x + b.x, a.y + b.y, a.z + b.z);
   return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}

// sin(a theta) / sin(theta) = a 2F1(1 - a, 1 + a; 3/2; z), z = sin^2(theta/2),
// theta in [0, pi/2].  Term ratio: (k^2 - a^2) z / ((k + 1/2) k), at most z in
// magnitude for a in [0, 1], so K terms leave less than z^(K+1) / (1 - z).
// Both weights of a slerp, a = 1 - t and a = t, share z: the Horner terms
// k^2 z / ((k + 1/2) k) are computed once.  Tiers, chosen for the whole wave
// (lanes then take the same path): z <= 0.025 (theta <= 18 deg) four terms
// (< 1.1e-8), z <= 0.1 (theta <= 37 deg) eight terms (< 1.2e-9); beyond, the
// series is summed until a term drops below 1e-8 (at most 27 terms).
template <int K>
__device__ __forceinline__ void slerp_horner(float a, float b, float z, float &wa, float &wb) {
   const float qa = z * (a * a), qb = z * (b * b);
   float acc_a = 1.0f, acc_b = 1.0f;
#pragma unroll
   for (int k = K; k >= 1; k--) {
      const float c = 1.0f / ((k + 0.5f) * (float)k);
      const float pk = z * ((float)(k * k) * c);
      acc_a = fmaf(fmaf(-qa, c, pk), acc_a, 1.0f);
      acc_b = fmaf(fmaf(-qb, c, pk), acc_b, 1.0f);
   }
   wa = a * acc_a;
   wb = b * acc_b;
}

__device__ __forceinline__ float slerp_series(float a, float z) {
   const float a2 = a * a;
   float term = 1.0f, sum = 1.0f;
   for (int k = 1; k <= 40; k++) {
      const float kf = (float)k;
      term *= (kf * kf - a2) * z / ((kf + 0.5f) * kf);
      sum += term;
      if (fabsf(term) < 1e-8f) break;
   }
   return a * sum;
}

__device__ __forceinline__ void slerp_weights(float t, float z, float &wa, float &wb) {
   const float a = 1.0f - t;
   if (__ballot(z > 0.025f) == 0) {
      slerp_horner<4>(a, t, z, wa, wb);
   } else if (z <= 0.1f) {
      slerp_horner<8>(a, t, z, wa, wb);
   } else {
      wa = slerp_series(a, z);
      wb = slerp_series(t, z);
   }
}

// sin x and 1 - cos x for the per-voxel retardance; odd / even Taylor
// polynomials: two terms when every lane of the wave has |x| <= 2^-6
// (truncation below 5e-10 relative; per-voxel retardances of thin voxels are
// ~1e-2), four for |x| <= 0.25 (truncation < 1e-11), sincosf beyond
__device__ __forceinline__ void sin_omc(float x, float &s, float &omc) {
   if (__ballot(fabsf(x) > 0.015625f) == 0) {
      const float x2 = x * x;
      s = x * fmaf(x2, -1.0f / 6.0f, 1.0f);
      omc = x2 * fmaf(x2, -1.0f / 24.0f, 0.5f);
   } else if (fabsf(x) <= 0.25f) {
      const float x2 = x * x;
      s = x * fmaf(x2, fmaf(x2, fmaf(x2, -1.0f / 5040.0f, 1.0f / 120.0f), -1.0f / 6.0f), 1.0f);
      omc = x2 * fmaf(x2, fmaf(x2, fmaf(x2, -1.0f / 40320.0f, 1.0f / 720.0f), -1.0f / 24.0f), 0.5f);
   } else {
      float sh, ch;
      sincosf(0.5f * x, &sh, &ch);
      s = 2.0f * sh * ch;
      omc = 2.0f * sh * sh; // 1 - cos x without cancellation
   }
}

// SphericalInterpolation (simulator.cpp:473-494) without trigonometric calls.
// Every lane computes the blend and the cases the reference returns early
// (a zero argument, arguments within 1e-6 of each other, t = 0) are selected
// at the end: the wave runs the blend for its other lanes anyway, and early
// exits cost mask bookkeeping.  A zero argument gives c = fminf(NaN, 1) = 1,
// z = 0, so such lanes never raise the wave's series tier.
__device__ __forceinline__ float3 slerp_orient(const float3 &v, const float3 &u0, float t) {
   const bool zv = is_zero(v), zu = is_zero(u0);
   const float d0 = dot3(v, u0);
   const bool flip = d0 < 0.0f;
   const float3 u = flip ? make_float3(-u0.x, -u0.y, -u0.z) : u0;
   const float d = flip ? -d0 : d0;
   const float3 df = make_float3(v.x - u.x, v.y - u.y, v.z - u.z);
   const bool same = dot3(df, df) < 1e-12f || t == 0.0f;
   const float c = fminf(d * rsqrtf(dot3(v, v) * dot3(u, u)), 1.0f);
   const float z = 0.5f * (1.0f - c);
   float wa, wb;
   slerp_weights(t, z, wa, wb);
   const float3 r = make_float3(v.x * wa + u.x * wb, v.y * wa + u.y * wb, v.z * wa + u.z * wb);
   return zv ? u0 : ((zu || same) ? v : r);
}

__device__ __forceinline__ int32_t clampi(int32_t v, int32_t n) {
   return v < 0 ? 0 : (v > n - 1 ? n - 1 : v);
}

// PliSimulator::GetVec Lerp/Slerp (simulator.cpp:496-555), any field storage.
// The position lies inside the volume (label `lab` != 0).  x0 = floor(x - 0.5)
// and x1 = ceil(x - 0.5), both clamped to the volume; x1 - x0 is 1 whenever
// the indices differ, so the weight (x - x0 - 0.5) / (x1 - x0) is
// (x - x0) - 0.5 exactly (0 when they coincide).  Corners of another label
// are zeroed (simulator.cpp:533-540); the interpolation runs x, y, z (4, 2, 1
// blends); a level whose two corner indices coincide is skipped: with weight
// 0 and equal arguments both interpolations return their first argument
// exactly.  When all corners coincide they are the voxel holding the point,
// read like a corner.  ST_F16 under Slerp: the corners are used as decoded,
// and unless a corner used carries the "not normalized" bit the result counts
// as a unit axis, `unit` (a slerp of unit axes is one; only its direction is
// used); under Lerp each corner is scaled to unit length (the chord's length
// enters the retardance).  ST_ANGLE16 decodes unit axes.
//
// Every corner load is issued before any corner is used: a load issued after
// a branch on an earlier load's data waits for that load, one memory round
// trip per level.  A view whose light spreads over at most one of x and y in
// the wave (the axis "a"; a tilt along x or y) reads its four corners
// (a0, a1) x (z0, z1) at once and interpolates a, then z: for such a view the
// x, y, z order reduces to that, because the other axis' level returns its
// first argument.  A view spreading over both reads all eight.
template <typename Idx, int ST>
__device__ float3 get_vec_queued(const Field<Idx, ST> &F, double px, double py, double pz,
                                 int32_t lab, int slerp, bool &unit) {
   const double fx = px - 0.5, fy = py - 0.5, fz = pz - 0.5;
   const double flx = floor(fx), fly = floor(fy), flz = floor(fz);
   const int32_t x0 = clampi((int32_t)flx, F.nx), x1 = clampi((int32_t)flx + (fx != flx), F.nx);
   const int32_t y0 = clampi((int32_t)fly, F.ny), y1 = clampi((int32_t)fly + (fy != fly), F.ny);
   const int32_t z0 = clampi((int32_t)flz, F.nz), z1 = clampi((int32_t)flz + (fz != flz), F.nz);
   const bool dx = x0 != x1, dy = y0 != y1, dz = z0 != z1;
   // indices that differ are unclamped (x0 = floor(x - 0.5)), and then
   // (x - x0) - 0.5 = (x - 0.5) - floor(x - 0.5), both exact in FP64
   const float xd = dx ? (float)(fx - flx) : 0.0f;
   const float yd = dy ? (float)(fy - fly) : 0.0f;
   const float zd = dz ? (float)(fz - flz) : 0.0f;
   using Fd = Field<Idx, ST>;
   const bool late = ST == ST_F16 && slerp;
   // corners of another label are zeroed; `keep` collects the "not
   // normalized" bits of the corners used (a corner is decoded where it is
   // used: its record has arrived by then)
   bool keep = false;
   auto corner = [&](const uint4 &r) {
      if (Fd::label_of(r) != lab) return make_float3(0.0f, 0.0f, 0.0f);
      if (late) keep |= Fd::nn_of(r);
      return Fd::axis_of(r, !late);
   };
   auto ip = [&](const float3 &a, const float3 &b, float t) {
      return slerp ? slerp_orient(a, b, t) : lerp_orient(a, b, t);
   };
   // record indices ([x][z][y]): i(x, y, z) = (x nz + z') ny + y, z' the
   // flipped z under flip_z
   const Idx ny = (Idx)F.ny, nzny = (Idx)F.nz * (Idx)F.ny;
   const int32_t z0f = F.flip_z ? F.nz - 1 - z0 : z0, z1f = F.flip_z ? F.nz - 1 - z1 : z1;
   const Idx i00 = (Idx)x0 * nzny + (Idx)z0f * ny + (Idx)y0;
   const Idx i01 = (Idx)x0 * nzny + (Idx)z1f * ny + (Idx)y0;
   const bool xs = __ballot(dx) != 0, ys = __ballot(dy) != 0;
   float3 v;
   if (!(xs && ys)) {
      // the spreading axis a (x, else y; with neither, a = y repeats y0)
      const Idx da_idx = xs ? (Idx)(x1 - x0) * nzny : (Idx)(y1 - y0);
      const bool da = xs ? dx : dy;
      const float ad = xs ? xd : yd;
      const uint4 r00 = F.load(i00), r01 = F.load(i01), r10 = F.load(i00 + da_idx),
                  r11 = F.load(i01 + da_idx);
      // an empty asm statement reading the four records keeps the loads
      // ahead of the branches below (the compiler would sink a load into
      // the branch that uses it); words x and y: ST_ANGLE16 reads them with
      // two loads
      __asm__ volatile("" ::"v"(r00.x), "v"(r01.x), "v"(r10.x), "v"(r11.x), "v"(r00.y),
                       "v"(r01.y), "v"(r10.y), "v"(r11.y));
      float3 c0 = corner(r00);
      if (da) c0 = ip(c0, corner(r10), ad);
      v = c0;
      if (dz) {
         float3 c1 = corner(r01);
         if (da) c1 = ip(c1, corner(r11), ad);
         v = ip(c0, c1, zd);
      }
   } else {
      const Idx dxi = (Idx)(x1 - x0) * nzny, dyi = (Idx)(y1 - y0);
      const uint4 r000 = F.load(i00), r001 = F.load(i01), r100 = F.load(i00 + dxi),
                  r101 = F.load(i01 + dxi), r010 = F.load(i00 + dyi), r011 = F.load(i01 + dyi),
                  r110 = F.load(i00 + dxi + dyi), r111 = F.load(i01 + dxi + dyi);
      __asm__ volatile("" ::"v"(r000.x), "v"(r001.x), "v"(r100.x), "v"(r101.x), "v"(r010.x),
                       "v"(r011.x), "v"(r110.x), "v"(r111.x));
      __asm__ volatile("" ::"v"(r000.y), "v"(r001.y), "v"(r100.y), "v"(r101.y), "v"(r010.y),
                       "v"(r011.y), "v"(r110.y), "v"(r111.y));
      // level x
      float3 c00 = corner(r000);
      if (dx) c00 = ip(c00, corner(r100), xd);
      float3 c01 = c00, c10 = c00, c11 = c00;
      if (dz) {
         c01 = corner(r001);
         if (dx) c01 = ip(c01, corner(r101), xd);
      }
      if (dy) {
         c10 = corner(r010);
         if (dx) c10 = ip(c10, corner(r110), xd);
         if (dz) {
            c11 = corner(r011);
            if (dx) c11 = ip(c11, corner(r111), xd);
         }
      }
      // level y
      float3 c0 = c00, c1 = c01;
      if (dy) {
         c0 = ip(c00, c10, yd);
         if (dz) c1 = ip(c01, c11, yd);
      }
      // level z
      v = dz ? ip(c0, c1, zd) : c0;
   }
   unit = (late && !keep) || (ST == ST_ANGLE16 && slerp);
   return v;
}

// GetVec on f32 fields when the axis `own` of the voxel holding the point is
// at hand (the adjoint's forward steps): `own` when all corners coincide,
// else get_vec_queued.
template <typename Idx>
__device__ float3 get_vec_interp(const Field<Idx, ST_F32> &F, double px, double py, double pz,
                                 int32_t lab, const float3 &own, int slerp) {
   const double fx = px - 0.5, fy = py - 0.5, fz = pz - 0.5;
   const int32_t x0 = clampi((int32_t)floor(fx), F.nx), x1 = clampi((int32_t)ceil(fx), F.nx);
   const int32_t y0 = clampi((int32_t)floor(fy), F.ny), y1 = clampi((int32_t)ceil(fy), F.ny);
   const int32_t z0 = clampi((int32_t)floor(fz), F.nz), z1 = clampi((int32_t)ceil(fz), F.nz);
   if (x0 == x1 && y0 == y1 && z0 == z1) return own;
   bool unit;
   return get_vec_queued<Idx, ST_F32>(F, px, py, pz, lab, slerp, unit);
}

// CalcStartingLightPositionsTilted (simulator.cpp:659-703) for one pixel,
// every FP64 operation rounded separately as in the reference.
__device__ void tilted_start(const TiltDev &T, int64_t i, int64_t j, double &tx, double &ty) {
#pragma clang fp contract(off)
   const double e = ((double)i + 0.5 - T.pc[0]) - T.pt[0];
   const double f = ((double)j + 0.5 - T.pc[1]) - T.pt[1];
   const double Ds = e * T.d - f * T.b;
   const double Dt = T.a * f - T.c * e;
   const double s = Ds / T.D;
   const double t = Dt / T.D;
   double p[3];
   for (int k = 0; k < 3; k++) p[k] = ((T.pc[k] + T.pt[k]) + T.e1[k] * s) + T.e2[k] * t;
   const double u0 = p[0] - T.pc[0], u1 = p[1] - T.pc[1], u2 = p[2] - T.pc[2];
   double acc0 = 0.0, acc1 = 0.0;
   acc0 = acc0 + u0 * T.rinv[0];
   acc0 = acc0 + u1 * T.rinv[1];
   acc0 = acc0 + u2 * T.rinv[2];
   acc1 = acc1 + u0 * T.rinv[3];
   acc1 = acc1 + u1 * T.rinv[4];
   acc1 = acc1 + u2 * T.rinv[5];
   tx = (acc0 + T.pc[0]) - T.shift[0];
   ty = (acc1 + T.pc[1]) - T.shift[1];
}

// ------------------------------------------------------------------ kernel
// A workgroup marches WG consecutive rays of the row-major (x, y) pixel grid:
// the 64 rays of a wave are consecutive y (one contiguous record run per
// z-step, two at a row end), and no lanes idle when ny is not a multiple of 64.
__host__ __device__ inline int64_t ray_tiles(int64_t nx, int64_t ny) {
   return (nx * ny + WG - 1) / WG;
}

// Workgroup b of a launch over n_tiles ray tiles x nh tilts -> (tile, tilt).
// Workgroups go to the 8 XCDs round-robin (b mod 8), each XCD with its own
// L2: the nh tilts of a tile, which read overlapping columns at the same
// time, get workgroup numbers equal mod 8 (one XCD), consecutive on that XCD;
// tiles stay interleaved over the XCDs, which keeps their loads even.  The
// last n_tiles mod 8 tiles take the plain order.
__device__ __forceinline__ void tile_tilt(unsigned b, unsigned nh, unsigned n_tiles, int &tilt,
                                          int64_t &tile) {
   const unsigned full = n_tiles / 8u * 8u * nh;
   if (b < full) {
      const unsigned g = b / (8u * nh), r = b - g * (8u * nh);
      tilt = (int)(r >> 3);
      tile = (int64_t)g * 8 + (r & 7u);
   } else {
      const unsigned c = b - full;
      tilt = (int)(c % nh);
      tile = (int64_t)(full / nh + c / nh);
   }
}

// Wave-cooperative march of interpolating views.  A quarter of the voxels of
// a fiber tissue are birefringent, scattered over the 64 rays of a wave; one
// thread per ray would run the costly part of a step (corner gather +
// interpolation, axis rotation, retarder angles) for the whole wave whenever
// one ray needs it.  Instead every ray walks its voxels one step per
// iteration and queues each birefringent step in a per-wave LDS ring (ballot
// + mbcnt, no atomics); attenuation-only steps are summed at once (they
// commute with the retarders).  When 64 steps are queued (or no ray can go
// on) the wave evaluates them, one per lane, and writes every step's retarder
// (cos 2phi, sin 2phi, sin delta, 1 - cos delta, mu t) into its entry; each
// ray then applies its own evaluated steps to its Poincare state in its own
// order.  A ray holds at most QDEPTH unapplied steps: at the next
// birefringent voxel it waits (attenuation-only voxels it still passes).
constexpr int QCAP = 128;  // ring entries per wave: < 64 waiting + 64 new
constexpr int QDEPTH = 4;  // queued steps per ray (8-bit sequence numbers in one word)

__device__ __forceinline__ void wave_sync() {
   __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
   __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double pack_double(uint32_t lo, uint32_t hi) {
   return __hiloint2double((int)hi, (int)lo);
}


// The retarder of one birefringent step (simulator.cpp:292-320) in the
// rho-invariant form: (cos 2phi, sin 2phi) of the tilted axis, sin delta and
// 1 - cos delta.  False for a zero axis (skipped without attenuation,
// simulator.cpp:292-297).  `nn`: the precision storages' "not normalized" bit.
// The tilt's rotation of the optical axis, kept in registers
struct AxisRot {
   float R[9];
   bool on; // theta != 0
};

// `unit`: v is a unit axis by construction (16-bit storages), so |v| = 1 in
// d_rel (simulator.cpp:309); every other use of v is invariant to its length.
template <int ST>
__device__ __forceinline__ bool retarder(const AxisRot &T, float K, float3 v, bool nn, bool unit,
                                         uint32_t &warn, float &C, float &S, float &sd,
                                         float &omc) {
   C = 1.0f;
   S = 0.0f;
   sd = 0.0f;
   omc = 0.0f;
   if (is_zero(v)) {
      warn |= FPL_WARN_ZERO_VEC;
      return false;
   }
   if (ST == ST_F32 ? (dot3(v, v) - 1.0f > 1e-5f) : nn) warn |= FPL_WARN_NOT_NORM;
   const float dr = (ST != ST_F32 && unit) ? K : K * sqrtf(dot3(v, v));
   if (T.on) {
      const float3 w = v;
      v.x = T.R[0] * w.x + T.R[1] * w.y + T.R[2] * w.z;
      v.y = T.R[3] * w.x + T.R[4] * w.y + T.R[5] * w.z;
      v.z = T.R[6] * w.x + T.R[7] * w.y + T.R[8] * w.z;
   }
   const float xx = v.x * v.x, yy = v.y * v.y;
   const float rxy = xx + yy;
   const float rr = rxy + v.z * v.z;
   const float delta = dr * (rxy / rr);
   if (rxy > 0.0f) {
      const float inv = 1.0f / rxy;
      C = (xx - yy) * inv;
      S = 2.0f * v.x * v.y * inv;
   }
   sin_omc(delta, sd, omc);
   return true;
}

// u <- rotation of u by -delta about n = (cos 2phi, -sin 2phi, 0), the
// increment added in float-float.  The component along n is unchanged; the
// components w = m . u along m = (sin 2phi, cos 2phi, 0) and u3 turn by delta:
// (w, u3) += (sin d u3 - (1 - cos d) w, -sin d w - (1 - cos d) u3), and the
// change of w goes back along m (Rodrigues' formula with |n| = 1)
__device__ __forceinline__ void apply_retarder(float C, float S, float sd, float omc, float &h1,
                                               float &h2, float &h3, float &l1, float &l2,
                                               float &l3) {
   const float w = S * h1 + C * h2;
   const float g = sd * h3 - omc * w;
   const float d3 = -(sd * w + omc * h3);
   two_sum(h1, l1, S * g);
   two_sum(h2, l2, C * g);
   two_sum(h3, l3, d3);
}

// At most 80 VGPRs (6 waves per SIMD): the queue evaluation's interpolation
// needs them, and spilling costs more than the lost occupancy.
// PROPS_LDS: the whole property table fits the LDS copy (else it is read from
// HBM through the caches); a compile-time choice keeps every lookup a plain
// LDS read.
template <typename Idx, int ST, bool PROPS_LDS>
__global__ void __launch_bounds__(WG, 6)
    k_march(const Params *__restrict__ prm, const TiltDev *__restrict__ tilts,
            const PropDev *__restrict__ props, const void *__restrict__ rec_a,
            const uint32_t *__restrict__ rec_b, double *__restrict__ images, uint32_t *flags) {
   __shared__ PropDev sp[MAX_LDS_PROPS];
   // queue entry words 0-3 / 4-7.  Queued: px, py, pz as doubles, label,
   // "not normalized" flag of the ray's voxel.  Evaluated: cos 2phi, sin 2phi,
   // sin delta, 1 - cos delta, mu t.
   __shared__ uint4 qa[WG / 64][QCAP];
   __shared__ uint4 qb[WG / 64][QCAP];
   const Params &P = *prm;
   const int64_t n_props = P.n_props;
   // a label outside [0, n_props) as one unsigned compare
   const uint32_t n_labels = (uint32_t)(n_props < 0x7fffffff ? n_props : 0x7fffffff);
   if (PROPS_LDS) {
      for (int k = threadIdx.x; k < (int)n_props; k += WG) sp[k] = props[k];
      __syncthreads();
   }
   auto prop = [&](int32_t l) -> PropDev { return PROPS_LDS ? sp[l] : props[l]; };

   // the interpolating tilts of one ray tile run on one XCD (tile_tilt): they
   // march through overlapping columns at the same time, so re-reads hit its
   // L2.  Tilts without interpolation (a third of the work per ray) come last
   // and fill the end of the launch, when the long rays are finishing.
   int tilt;
   int64_t tile;
   {
      const unsigned nh = (unsigned)P.n_heavy, nt = (unsigned)P.n_tilts;
      const unsigned n_tiles = (unsigned)ray_tiles(P.nx, P.ny);
      const unsigned split = nh * n_tiles;
      const unsigned b = blockIdx.x;
      if (b < split) {
         tile_tilt(b, nh, n_tiles, tilt, tile);
      } else {
         const unsigned c = b - split;
         tilt = (int)(nh + c % (nt - nh));
         tile = c / (nt - nh);
      }
   }
   const int lane = (int)(threadIdx.x & 63u);
   uint4 *const QA = qa[threadIdx.x >> 6];
   uint4 *const QB = qb[threadIdx.x >> 6];
   const int64_t ray = tile * WG + threadIdx.x;
   const int64_t i = ray / P.ny, j = ray - i * P.ny;
   // rays past the grid's end stay in the loop (their lanes evaluate queued steps)
   const bool valid = i < P.nx;
   const TiltDev &T = tilts[tilt];
   const int64_t nxy = P.nx * P.ny;
   const Field<Idx, ST> F{rec_a, rec_b, (int32_t)P.nx, (int32_t)P.ny, (int32_t)P.nz, P.flip_z};
   const bool interp = !(P.interp == 0 || T.centers);
   const int slerp = P.interp == 2;
   AxisRot rot;
   for (int k = 0; k < 9; k++) rot.R[k] = T.R[k];
   rot.on = T.rotate != 0;

   double px = 0.0, py = 0.0, pz = 0.0;
   if (P.untilt) {
      px = ((double)i + 0.5) - T.half_shift[0];
      py = ((double)j + 0.5) - T.half_shift[1];
   } else if (valid) {
      tilted_start(T, i, j, px, py);
   }
   px = px + T.half_step[0];
   py = py + T.half_step[1];
   pz = pz + T.half_step[2];
   const double sx = T.step[0], sy = T.step[1], sz = T.step[2];
   const double nzd = (double)P.nz;

   // Poincare state u = (u1, u2, u3) in float-float, starts at (0, 0, -1)
   float h1 = 0.0f, h2 = 0.0f, h3 = -1.0f, l1 = 0.0f, l2 = 0.0f, l3 = 0.0f;
   float mh = 0.0f, ml = 0.0f; // sum of mu * t
   uint32_t warn = 0, err = 0;
   if (!interp) {
      // voxel axes used as they are (nearest neighbour, or every sample point
      // a voxel centre): a step is cheap, so every ray runs its own steps, with
      // the records of its next three steps in flight while it processes one (one
      // load in flight leaves the loop waiting on memory latency).  Every ray
      // takes the same number of steps, T.trips: the reference's loop test
      // (simulator.cpp:259-261) looks at z alone, and every ray starts at
      // z = step_z / 2.  Four slots hold steps k .. k + 3; a slot is processed,
      // then refilled with the step four ahead in the same registers (refilling
      // first would need a register copy, which waits for the load).
      const int64_t trips = T.trips;
      auto process = [&](const uint4 &r, bool in) {
         int32_t lab = in ? F.label_of(r) : 0;
         if ((uint32_t)lab >= n_labels) {
            err = 1;
            lab = 0;
         }
         const PropDev pd = prop(lab);
         if (pd.bg) {
            if (pd.mu_t != 0.0f) two_sum(mh, ml, pd.mu_t);
         } else {
            float C, S, sd, omc;
            // 16-bit storages: an axis that passed the normalisation test counts
            // as unit length (no rescaling after decoding)
            const bool nn = F.nn_of(r);
            if (retarder<ST>(rot, pd.K, F.axis_of(r, false), nn, ST == ST_ANGLE16 || !nn, warn,
                             C, S, sd, omc)) {
               apply_retarder(C, S, sd, omc, h1, h2, h3, l1, l2, l3);
               if (pd.mu_t != 0.0f) two_sum(mh, ml, pd.mu_t);
            }
         }
      };
      auto run = [&](auto refill) {
         uint4 r0, r1, r2, r3;
         bool n0, n1, n2, n3;
         refill(r0, n0);
         refill(r1, n1);
         refill(r2, n2);
         refill(r3, n3);
         // rays past the grid's end run along (label 0, nothing written)
         for (int64_t k = 0; k < trips; k += 4) {
            process(r0, n0);
            refill(r0, n0);
            if (k + 1 >= trips) break;
            process(r1, n1);
            refill(r1, n1);
            if (k + 2 >= trips) break;
            process(r2, n2);
            refill(r2, n2);
            if (k + 3 >= trips) break;
            process(r3, n3);
            refill(r3, n3);
         }
      };
      int64_t kf = 0; // the next step to fetch
      if (sx == 0.0 && sy == 0.0) {
         // the light runs along z (untilted views): the ray stays in one voxel
         // column, x and y never change, and step k's record is the column's
         // base plus floor(z_k) rows of ny records
         const int32_t jx = ifloor(px), jy = ifloor(py);
         const bool in = valid & ((uint32_t)jx < (uint32_t)P.nx) & ((uint32_t)jy < (uint32_t)P.ny);
         const Idx nyi = (Idx)P.ny;
         const Idx base = in ? ((Idx)jx * (Idx)P.nz) * nyi + (Idx)jy : (Idx)0;
         run([&](uint4 &r, bool &rin) {
            int32_t z = (int32_t)floor(pz);
            rin = in & (kf < trips);
            z = rin ? z : 0;
            if (P.flip_z) z = (int32_t)P.nz - 1 - z;
            r = F.load(base + (Idx)z * nyi);
            pz = pz + sz;
            kf++;
         });
      } else {
         run([&](uint4 &r, bool &rin) {
            rin = F.fetch_raw(px, py, floor(pz), valid & (kf < trips), r);
            px = px + sx;
            py = py + sy;
            pz = pz + sz;
            kf++;
         });
      }
   } else {
      // interpolating views: the wave-cooperative queue described above
      uint32_t pend = 0; // sequence numbers (mod 256) of the ray's queued steps, oldest lowest
      int npend = 0;
      uint32_t tail = 0, head = 0; // wave: steps queued / evaluated so far

      // The record of the current step is loaded one iteration ahead.  A ray
      // takes T.trips steps (the reference's loop test, simulator.cpp:259-261,
      // looks at z alone).  A coordinate the light does not move in any lane
      // (a view tilted along the other axis: adding the step leaves it as it
      // is, so it stays so) is neither stepped nor floored again.
      const int32_t trips = T.trips;
      const bool xfix = __ballot(px + sx != px) == 0, yfix = __ballot(py + sy != py) == 0;
      int32_t jx = ifloor(px), jy = ifloor(py); // the current step's voxel column
      int32_t ks = 0;                           // steps taken
      bool live = valid & (trips > 0);
      auto fetch = [&](uint4 &rec) -> bool {
         const bool in = live & ((uint32_t)jx < (uint32_t)F.nx) & ((uint32_t)jy < (uint32_t)F.ny);
         int32_t jz = (int32_t)floor(pz);
         if (F.flip_z) jz = F.nz - 1 - jz;
         // the load is issued in every case (record 0 for a point outside)
         const Idx i = ((Idx)jx * (Idx)F.nz + (Idx)jz) * (Idx)F.ny + (Idx)jy;
         rec = F.load(in ? i : (Idx)0);
         return in;
      };
      // the next step's position; the empty asm statements keep the
      // wave-uniform branches (the compiler would otherwise compute both sides)
      auto advance = [&]() {
         if (!xfix) {
            __asm__ volatile("");
            px = px + sx;
            jx = ifloor(px);
         }
         if (!yfix) {
            __asm__ volatile("");
            py = py + sy;
            jy = ifloor(py);
         }
         pz = pz + sz;
         ks++;
         live = ks < trips;
      };
      uint4 r;
      bool rin = fetch(r);
      for (;;) {
         int32_t lab = rin ? F.label_of(r) : 0;
         if ((uint32_t)lab >= n_labels) {
            err = 1;
            lab = 0;
         }
         const PropDev pd = prop(lab);
         const bool bire = live && !pd.bg;
         const bool go = live && (!bire || npend < QDEPTH);
         const uint64_t gomask = __ballot(go);
         if (gomask == 0 && tail == head) break; // every ray ended, every step applied
         if (gomask != 0) {
            const uint64_t qmask = __ballot(go && bire);
            if (go) {
               if (bire) {
                  const uint32_t seq =
                      tail + __builtin_amdgcn_mbcnt_hi((uint32_t)(qmask >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)qmask, 0u));
                  const uint32_t slot = seq & (QCAP - 1);
                  QA[slot] = make_uint4(__double2loint(px), __double2hiint(px), __double2loint(py),
                                        __double2hiint(py));
                  QB[slot] = make_uint4(__double2loint(pz), __double2hiint(pz), (uint32_t)lab,
                                        F.nn_of(r) ? 1u : 0u);
                  pend |= (seq & 255u) << (8 * npend);
                  npend++;
               } else if (pd.mu_t != 0.0f) {
                  two_sum(mh, ml, pd.mu_t);
               }
               advance();
               rin = fetch(r);
            }
            tail += (uint32_t)__popcll(qmask);
         }
         const uint32_t queued = tail - head;
         if (queued >= 64 || gomask == 0) {
            // the evaluation issues ahead of walking waves of the same SIMD
            // (a walking wave mostly waits on its next record): 5-tilt launch
            // -1.5 %
            __builtin_amdgcn_s_setprio(2);
            const uint32_t n = queued < 64 ? queued : 64;
            wave_sync();
            if ((uint32_t)lane < n) {
               const uint32_t slot = (head + (uint32_t)lane) & (QCAP - 1);
               const uint4 a = QA[slot], b = QB[slot];
               const int32_t elab = (int32_t)b.z;
               const PropDev epd = prop(elab);
               bool unit;
               const float3 v = get_vec_queued<Idx, ST>(F, pack_double(a.x, a.y),
                                                        pack_double(a.z, a.w),
                                                        pack_double(b.x, b.y), elab, slerp, unit);
               float C, S, sd, omc, mu = 0.0f;
               if (retarder<ST>(rot, epd.K, v, b.w != 0, unit, warn, C, S, sd, omc))
                  mu = epd.mu_t;
               QA[slot] = make_uint4(__float_as_uint(C), __float_as_uint(S), __float_as_uint(sd),
                                     __float_as_uint(omc));
               QB[slot].x = __float_as_uint(mu);
            }
            wave_sync();
            // every ray applies its steps of this batch, oldest first: a
            // prefix of its queued steps (their sequence numbers increase)
            int k = 0;
#pragma unroll
            for (int q = 0; q < QDEPTH; q++)
               k += (q < npend) & (((((pend >> (8 * q)) & 255u) - head) & 255u) < n);
            for (int q = 0; q < k; q++) {
               const uint32_t slot = (pend >> (8 * q)) & (QCAP - 1);
               const uint4 e = QA[slot];
               const float mu = __uint_as_float(QB[slot].x);
               const float C = __uint_as_float(e.x), S = __uint_as_float(e.y);
               const float sd = __uint_as_float(e.z), omc = __uint_as_float(e.w);
               apply_retarder(C, S, sd, omc, h1, h2, h3, l1, l2, l3);
               two_sum(mh, ml, mu); // exact no-op for mu = 0
            }
            pend = (uint32_t)((uint64_t)pend >> (8 * k));
            npend -= k;
            head += n;
            __builtin_amdgcn_s_setprio(0);
         }
      }
   }

   if (valid) {
      const double A = exp(-((double)mh + (double)ml));
      const double u1 = (double)h1 + (double)l1, u2 = (double)h2 + (double)l2;
      const double base = P.quarter_I0 * A;
      double *out = images + ((int64_t)T.out * nxy + ray) * P.n_rho;
      bool neg = false; // the caller's "intensity < 0" check (simpli.py:184-185)
      for (int k = 0; k < P.n_rho; k++) {
         const double I = base * (1.0 - (P.c2r[k] * u1 - P.s2r[k] * u2));
         out[k] = I;
         neg |= I < 0.0;
      }
      if (neg) warn |= FPL_FLAG_NEGATIVE;
   }
   const uint32_t f = warn | (err ? 0x80000000u : 0u);
   if (f) atomicOr(&flags[0], f);
}

// ------------------------------------------------------------------ adjoint
// Reverse mode of the march: the gradient of L = sum(g * images) with respect
// to the property table (dn, mu of every label) and, for nearest-neighbour
// views, the per-voxel optical axis.  The reference has no derivative; the
// forward it differentiates is the march above (the reference's per-rho
// chain in the rho-invariant form).
//
// Per ray (one thread, tilt-interleaved workgroups):
//   forward   the march, keeping a checkpoint (FP64 position, Poincare state
//             u) before every CK-th step;
//   seeds     I_r = b (1 - c_r u1 + s_r u2), b = I0/4 exp(-M):
//             dL/du = b (-sum g_r c_r, sum g_r s_r, 0), dL/dM = -sum g_r I_r;
//   reverse   segment by segment from the last: replay the CK steps from the
//             checkpoint (the states before each step go to LDS), then walk
//             them backwards.  A retarder step is u' = (I - sin d K + (1 - cos d)
//             K^2) u with K = [n]x, n = (C, -S, 0); its adjoint is
//             ubar <- (I + sin d K + (1 - cos d) K^2) ubar, and
//             dbar = ubar . du'/dd, Cbar = ubar . du'/dC, Sbar = ubar . du'/dS give
//             dL/dK = dbar |v| cos^2(alpha) and, through C, S and d as functions of
//             the axis, dL/dv.  Every step that attenuates adds dL/dM to mu.
// Per-label sums run along the ray while the label stays the same, then go to
// LDS (FP64 atomics) and once per workgroup to the device table.
constexpr int CK = 8;              // steps per checkpoint segment
constexpr int MAX_ADJ_PROPS = 256; // labels with LDS accumulators

struct Ckpt { // state before the first step of a segment
   double p[3];
   float u[3];
   float pad;
};

// the record of the voxel holding (px, py, pz) (background outside) and its index
template <typename Idx>
__device__ __forceinline__ Vox fetch_own(const Field<Idx, ST_F32> &F, double px, double py,
                                         double pz, int32_t &ix, int32_t &iy, int32_t &iz) {
   ix = ifloor(px);
   iy = ifloor(py);
   iz = ifloor(floor(pz));
   Vox r{0, make_float3(0.0f, 0.0f, 0.0f), false};
   if (F.inside(ix, iy, iz)) r = F.at(ix, iy, iz);
   return r;
}

// What one z-step of the march does at a position (k_march's arithmetic).
struct StepEval {
   int kind;   // 0: zero axis, skipped; 1: attenuation only; 2: retarder
   int32_t lab;
   float K, sd, omc, C, S, g; // g = |v| cos^2(alpha), delta = K g
   float3 v;                  // the axis after the tilt rotation
   int32_t vx, vy, vz;        // the voxel read (flip_z applied to vz)
};

template <typename Idx, bool PROPS_LDS>
__device__ __forceinline__ StepEval eval_step(const Params &P, const TiltDev &T,
                                              const PropDev *sp, const PropDev *props,
                                              const Field<Idx, ST_F32> &F, double px, double py,
                                              double pz) {
   StepEval E;
   const int32_t ix = ifloor(px), iy = ifloor(py), iz = ifloor(floor(pz));
   Vox r{0, make_float3(0.0f, 0.0f, 0.0f), false};
   if (F.inside(ix, iy, iz)) r = F.at(ix, iy, iz);
   const int32_t lab = (r.lab < 0 || r.lab >= P.n_props) ? 0 : r.lab;
   const PropDev pd = PROPS_LDS ? sp[lab] : props[lab];
   E.lab = lab;
   E.K = pd.K;
   E.vx = ix;
   E.vy = iy;
   E.vz = F.flip_z ? F.nz - 1 - iz : iz;
   if (pd.bg) {
      E.kind = 1;
      return E;
   }
   float3 v = (P.interp == 0 || T.centers)
                  ? r.v
                  : get_vec_interp(F, px, py, pz, lab, r.v, P.interp == 2);
   if (is_zero(v)) {
      E.kind = 0;
      return E;
   }
   const float n2 = dot3(v, v);
   const float nrm = sqrtf(n2);
   if (T.rotate) {
      const float3 w = v;
      v.x = T.R[0] * w.x + T.R[1] * w.y + T.R[2] * w.z;
      v.y = T.R[3] * w.x + T.R[4] * w.y + T.R[5] * w.z;
      v.z = T.R[6] * w.x + T.R[7] * w.y + T.R[8] * w.z;
   }
   const float xx = v.x * v.x, yy = v.y * v.y;
   const float rxy = xx + yy;
   const float rr = rxy + v.z * v.z;
   E.g = nrm * (rxy / rr);
   const float delta = (pd.K * nrm) * (rxy / rr);
   E.C = 1.0f;
   E.S = 0.0f;
   if (rxy > 0.0f) {
      const float inv = 1.0f / rxy;
      E.C = (xx - yy) * inv;
      E.S = 2.0f * v.x * v.y * inv;
   }
   sin_omc(delta, E.sd, E.omc);
   E.v = v;
   E.kind = 2;
   return E;
}

// W(a, z) = sin(a theta) / sin(theta) (z = sin^2(theta/2)) and dW/dz, from
// the series of slerp_weights and its term-wise derivative
__device__ void slerp_weight_dz(float a, float z, float &W, float &dW) {
   const float a2 = a * a;
   float A = 1.0f, zk1 = 1.0f, P = 1.0f, dP = 0.0f; // A_k, z^(k-1)
   for (int k = 1; k <= 40; k++) {
      A *= ((float)(k * k) - a2) / ((k + 0.5f) * (float)k);
      const float termd = (float)k * A * zk1;
      zk1 *= z;
      const float term = A * zk1;
      P += term;
      dP += termd;
      if (fabsf(term) < 1e-9f && fabsf(termd) < 1e-9f) break;
   }
   W = a * P;
   dW = a * dP;
}

// Reverse of one orientation interpolation s = interp(v, u, t) (lerp_orient /
// slerp_orient, with the same branches): sb = dL/ds -> vb, ub.
__device__ void interp_back(const float3 &v, const float3 &u, float t, int slerp,
                            const float3 &sb, float3 &vb, float3 &ub) {
   vb = make_float3(0.0f, 0.0f, 0.0f);
   ub = vb;
   if (is_zero(v)) {
      ub = sb;
      return;
   }
   if (is_zero(u)) {
      vb = sb;
      return;
   }
   if (!slerp) {
      const float3 a = make_float3(v.x * (1.0f - t), v.y * (1.0f - t), v.z * (1.0f - t));
      const float3 b = make_float3(u.x * t, u.y * t, u.z * t);
      const float sg = dot3(a, b) >= 0.0f ? t : -t;
      vb = make_float3(sb.x * (1.0f - t), sb.y * (1.0f - t), sb.z * (1.0f - t));
      ub = make_float3(sb.x * sg, sb.y * sg, sb.z * sg);
      return;
   }
   float d = dot3(v, u), sg = 1.0f;
   float3 w = u;
   if (d < 0.0f) {
      w = make_float3(-u.x, -u.y, -u.z);
      d = -d;
      sg = -1.0f;
   }
   const float3 df = make_float3(v.x - w.x, v.y - w.y, v.z - w.z);
   if (dot3(df, df) < 1e-12f || t == 0.0f) {
      vb = sb;
      return;
   }
   const float nv2 = dot3(v, v), nw2 = dot3(w, w);
   const float inv = rsqrtf(nv2 * nw2);
   const float craw = d * inv, c = fminf(craw, 1.0f);
   const float z = 0.5f * (1.0f - c);
   float wa, dwa, wb, dwb;
   slerp_weight_dz(1.0f - t, z, wa, dwa);
   slerp_weight_dz(t, z, wb, dwb);
   // s = wa v + wb w, wa and wb functions of z = (1 - c) / 2, c = v.w / (|v||w|)
   vb = make_float3(wa * sb.x, wa * sb.y, wa * sb.z);
   float3 wbar = make_float3(wb * sb.x, wb * sb.y, wb * sb.z);
   if (craw < 1.0f) {
      const float gc = -0.5f * (dot3(sb, v) * dwa + dot3(sb, w) * dwb);
      const float cv = c / nv2, cw = c / nw2;
      vb.x += gc * (w.x * inv - v.x * cv);
      vb.y += gc * (w.y * inv - v.y * cv);
      vb.z += gc * (w.z * inv - v.z * cv);
      wbar.x += gc * (v.x * inv - w.x * cw);
      wbar.y += gc * (v.y * inv - w.y * cw);
      wbar.z += gc * (v.z * inv - w.z * cw);
   }
   ub = make_float3(sg * wbar.x, sg * wbar.y, sg * wbar.z);
}

// Reverse of get_vec_interp: vb = dL/d(the interpolated axis) is
// distributed over the corners, each adding to its voxel's axis gradient
// (API layout [x][y][z][3], z mirrored under flip_z).  A corner of another
// label entered the interpolation as 0 and gets no gradient.
template <typename Idx>
__device__ void get_vec_interp_back(const Field<Idx, ST_F32> &F, double px, double py,
                                    double pz, int32_t lab, int slerp, const float3 &vb,
                                    float *__restrict__ g_vec) {
   const double fx = px - 0.5, fy = py - 0.5, fz = pz - 0.5;
   const double flx = floor(fx), fly = floor(fy), flz = floor(fz);
   const int32_t x0 = clampi((int32_t)flx, F.nx), x1 = clampi((int32_t)flx + (fx != flx), F.nx);
   const int32_t y0 = clampi((int32_t)fly, F.ny), y1 = clampi((int32_t)fly + (fy != fly), F.ny);
   const int32_t z0 = clampi((int32_t)flz, F.nz), z1 = clampi((int32_t)flz + (fz != flz), F.nz);
   auto add = [&](int32_t x, int32_t y, int32_t z, int32_t l, const float3 &g) {
      if (l != lab) return;
      const int32_t zz = F.flip_z ? F.nz - 1 - z : z;
      float *o = g_vec + 3 * (((int64_t)x * F.ny + y) * F.nz + zz);
      unsafeAtomicAdd(o + 0, g.x);
      unsafeAtomicAdd(o + 1, g.y);
      unsafeAtomicAdd(o + 2, g.z);
   };
   const Vox r000 = F.at(x0, y0, z0);
   if (x0 == x1 && y0 == y1 && z0 == z1) { // the voxel's own axis
      add(x0, y0, z0, r000.lab, vb);
      return;
   }
   const bool dx = x0 != x1, dy = y0 != y1, dz = z0 != z1;
   const float xd = dx ? (float)((px - (double)x0) - 0.5) : 0.0f;
   const float yd = dy ? (float)((py - (double)y0) - 0.5) : 0.0f;
   const float zd = dz ? (float)((pz - (double)z0) - 0.5) : 0.0f;
   const Vox r100 = F.at(x1, y0, z0), r010 = F.at(x0, y1, z0), r110 = F.at(x1, y1, z0);
   const Vox r001 = F.at(x0, y0, z1), r101 = F.at(x1, y0, z1);
   const Vox r011 = F.at(x0, y1, z1), r111 = F.at(x1, y1, z1);
   auto corner = [&](const Vox &r) {
      return r.lab != lab ? make_float3(0.0f, 0.0f, 0.0f) : r.v;
   };
   auto ip = [&](const float3 &a, const float3 &b, float t) {
      return slerp ? slerp_orient(a, b, t) : lerp_orient(a, b, t);
   };
   const float3 m000 = corner(r000), m100 = corner(r100), m010 = corner(r010),
                m110 = corner(r110), m001 = corner(r001), m101 = corner(r101),
                m011 = corner(r011), m111 = corner(r111);
   // forward, keeping the intermediate values (get_vec_interp's tree)
   float3 c00 = m000;
   if (dx) c00 = ip(m000, m100, xd);
   float3 c01 = c00, c10 = c00, c11 = c00;
   if (dz) {
      c01 = m001;
      if (dx) c01 = ip(m001, m101, xd);
   }
   if (dy) {
      c10 = m010;
      if (dx) c10 = ip(m010, m110, xd);
      if (dz) {
         c11 = m011;
         if (dx) c11 = ip(m011, m111, xd);
      }
   }
   float3 c0 = c00, c1 = c01;
   if (dy) {
      c0 = ip(c00, c10, yd);
      if (dz) c1 = ip(c01, c11, yd);
   }
   // backward
   const float3 zero = make_float3(0.0f, 0.0f, 0.0f);
   float3 c0b = vb, c1b = zero;
   if (dz) interp_back(c0, c1, zd, slerp, vb, c0b, c1b);
   float3 c00b = c0b, c01b = c1b, c10b = zero, c11b = zero;
   if (dy) {
      interp_back(c00, c10, yd, slerp, c0b, c00b, c10b);
      c01b = zero;
      if (dz) interp_back(c01, c11, yd, slerp, c1b, c01b, c11b);
   }
   float3 g0, g1;
   if (dx) {
      interp_back(m000, m100, xd, slerp, c00b, g0, g1);
      add(x0, y0, z0, r000.lab, g0);
      add(x1, y0, z0, r100.lab, g1);
   } else {
      add(x0, y0, z0, r000.lab, c00b);
   }
   if (dz) {
      if (dx) {
         interp_back(m001, m101, xd, slerp, c01b, g0, g1);
         add(x0, y0, z1, r001.lab, g0);
         add(x1, y0, z1, r101.lab, g1);
      } else {
         add(x0, y0, z1, r001.lab, c01b);
      }
   }
   if (dy) {
      if (dx) {
         interp_back(m010, m110, xd, slerp, c10b, g0, g1);
         add(x0, y1, z0, r010.lab, g0);
         add(x1, y1, z0, r110.lab, g1);
      } else {
         add(x0, y1, z0, r010.lab, c10b);
      }
      if (dz) {
         if (dx) {
            interp_back(m011, m111, xd, slerp, c11b, g0, g1);
            add(x0, y1, z1, r011.lab, g0);
            add(x1, y1, z1, r111.lab, g1);
         } else {
            add(x0, y1, z1, r011.lab, c11b);
         }
      }
   }
}

// PROPS_LDS: the property table fits the LDS copy (every lookup a plain LDS read)
template <typename Idx, bool PROPS_LDS>
__global__ void __launch_bounds__(WG)
    k_adjoint(const Params *__restrict__ prm, const TiltDev *__restrict__ tilts,
              const PropDev *__restrict__ props, const void *__restrict__ rec_a,
              const double *__restrict__ grad_images, Ckpt *__restrict__ ck, int64_t ck_stride,
              double kfac, double mfac, double *__restrict__ g_props, float *__restrict__ g_vec,
              int tilt0, int n_sub) {
   __shared__ PropDev sp[MAX_ADJ_PROPS];
   __shared__ double acc[MAX_ADJ_PROPS][2];
   __shared__ float ub[CK * 3][WG]; // states before the steps of the current segment
   const Params &P = *prm;
   const int np_lds = P.n_props < MAX_ADJ_PROPS ? (int)P.n_props : MAX_ADJ_PROPS;
   for (int k = threadIdx.x; k < np_lds; k += WG) {
      sp[k] = props[k];
      acc[k][0] = 0.0;
      acc[k][1] = 0.0;
   }
   __syncthreads();

   // the tilts [tilt0, tilt0 + n_sub) of the launch
   int tilt;
   int64_t tile;
   tile_tilt(blockIdx.x, (unsigned)n_sub, (unsigned)ray_tiles(P.nx, P.ny), tilt, tile);
   tilt += tilt0;
   const int64_t ray = tile * WG + threadIdx.x;
   const int64_t i = ray / P.ny, j = ray - i * P.ny;
   const int tid = threadIdx.x;
   // no early return: the workgroup meets again to flush its label sums
   if (i < P.nx) {
      const TiltDev &T = tilts[tilt];
      const int64_t nxy = P.nx * P.ny;
      const Field<Idx, ST_F32> F{rec_a, nullptr, (int32_t)P.nx, (int32_t)P.ny, (int32_t)P.nz,
                                 P.flip_z};
      Ckpt *my = ck + ((int64_t)tilt * nxy + ray) * ck_stride;

      double px, py, pz = 0.0;
      if (P.untilt) {
         px = ((double)i + 0.5) - T.half_shift[0];
         py = ((double)j + 0.5) - T.half_shift[1];
      } else {
         tilted_start(T, i, j, px, py);
      }
      px = px + T.half_step[0];
      py = py + T.half_step[1];
      pz = pz + T.half_step[2];
      const double sx = T.step[0], sy = T.step[1], sz = T.step[2];
      const double nzd = (double)P.nz;

      // forward with checkpoints
      float h1 = 0.0f, h2 = 0.0f, h3 = -1.0f, l1 = 0.0f, l2 = 0.0f, l3 = 0.0f;
      float mh = 0.0f, ml = 0.0f;
      int64_t n = 0;
      for (;; n++) {
         const double fz = floor(pz);
         if (!(fz >= 0.0 && fz < nzd)) break;
         if (n % CK == 0) {
            Ckpt c;
            c.p[0] = px;
            c.p[1] = py;
            c.p[2] = pz;
            c.u[0] = h1 + l1;
            c.u[1] = h2 + l2;
            c.u[2] = h3 + l3;
            c.pad = 0.0f;
            my[n / CK] = c;
         }
         const StepEval E = eval_step<Idx, PROPS_LDS>(P, T, sp, props, F, px, py, pz);
         if (E.kind == 1) {
            const float mt = PROPS_LDS ? sp[E.lab].mu_t : props[E.lab].mu_t;
            if (mt != 0.0f) two_sum(mh, ml, mt);
         } else if (E.kind == 2) {
            const float nv = E.C * h1 - E.S * h2;
            const float d1 = E.sd * E.S * h3 + E.omc * (E.C * nv - h1);
            const float d2 = E.sd * E.C * h3 + E.omc * (-E.S * nv - h2);
            const float d3 = -E.sd * (E.C * h2 + E.S * h1) - E.omc * h3;
            two_sum(h1, l1, d1);
            two_sum(h2, l2, d2);
            two_sum(h3, l3, d3);
            const float mt = PROPS_LDS ? sp[E.lab].mu_t : props[E.lab].mu_t;
            if (mt != 0.0f) two_sum(mh, ml, mt);
         }
         px = px + sx;
         py = py + sy;
         pz = pz + sz;
      }

      // seeds from the intensities
      const double A = exp(-((double)mh + (double)ml));
      const double u1 = (double)h1 + (double)l1, u2 = (double)h2 + (double)l2;
      const double base = P.quarter_I0 * A;
      const double *g = grad_images + ((int64_t)T.out * nxy + ray) * P.n_rho;
      double gc = 0.0, gs = 0.0, mbar = 0.0;
      for (int k = 0; k < P.n_rho; k++) {
         const double I = base * (1.0 - (P.c2r[k] * u1 - P.s2r[k] * u2));
         mbar -= g[k] * I;
         gc += g[k] * P.c2r[k];
         gs += g[k] * P.s2r[k];
      }
      float b1 = (float)(-base * gc), b2 = (float)(base * gs), b3 = 0.0f;

      // the sums of the current label run
      int32_t run_lab = -1;
      double run_k = 0.0;
      int64_t run_n = 0;
      auto flush = [&]() {
         if (run_lab < 0) return;
         const double dk = run_k * kfac, dm = mbar * (double)run_n * mfac;
         if (run_lab < MAX_ADJ_PROPS) {
            atomicAdd(&acc[run_lab][0], dk);
            atomicAdd(&acc[run_lab][1], dm);
         } else {
            atomicAdd(&g_props[2 * run_lab], dk);
            atomicAdd(&g_props[2 * run_lab + 1], dm);
         }
      };

      for (int64_t sgm = (n - 1) / CK; sgm >= 0 && n > 0; sgm--) {
         const Ckpt c = my[sgm];
         const int len = (int)min((int64_t)CK, n - sgm * CK);
         // replay: the states before each step of the segment
         double qx = c.p[0], qy = c.p[1], qz = c.p[2];
         float a1 = c.u[0], a2 = c.u[1], a3 = c.u[2], e1 = 0.0f, e2 = 0.0f, e3 = 0.0f;
         for (int t = 0; t < len; t++) {
            ub[3 * t + 0][tid] = a1 + e1;
            ub[3 * t + 1][tid] = a2 + e2;
            ub[3 * t + 2][tid] = a3 + e3;
            const StepEval E = eval_step<Idx, PROPS_LDS>(P, T, sp, props, F, qx, qy, qz);
            if (E.kind == 2) {
               const float nv = E.C * a1 - E.S * a2;
               const float d1 = E.sd * E.S * a3 + E.omc * (E.C * nv - a1);
               const float d2 = E.sd * E.C * a3 + E.omc * (-E.S * nv - a2);
               const float d3 = -E.sd * (E.C * a2 + E.S * a1) - E.omc * a3;
               two_sum(a1, e1, d1);
               two_sum(a2, e2, d2);
               two_sum(a3, e3, d3);
            }
            qx = qx + sx;
            qy = qy + sy;
            qz = qz + sz;
         }
         // backwards through the segment
         for (int t = len - 1; t >= 0; t--) {
            double rx = c.p[0], ry = c.p[1], rz = c.p[2];
            for (int q = 0; q < t; q++) {
               rx = rx + sx;
               ry = ry + sy;
               rz = rz + sz;
            }
            const StepEval E = eval_step<Idx, PROPS_LDS>(P, T, sp, props, F, rx, ry, rz);
            if (E.kind == 0) continue;
            if (E.lab != run_lab) {
               flush();
               run_lab = E.lab;
               run_k = 0.0;
               run_n = 0;
            }
            run_n++;
            if (E.kind == 1) continue;
            const float x1 = ub[3 * t + 0][tid], x2 = ub[3 * t + 1][tid], x3 = ub[3 * t + 2][tid];
            const float sd = E.sd, omc = E.omc, co = 1.0f - omc, C = E.C, S = E.S;
            const float nv = C * x1 - S * x2;
            // derivatives of the state after the step
            const float dd1 = co * S * x3 + sd * (C * nv - x1);
            const float dd2 = co * C * x3 + sd * (-S * nv - x2);
            const float dd3 = -co * (C * x2 + S * x1) - sd * x3;
            const float dbar = b1 * dd1 + b2 * dd2 + b3 * dd3;
            const float cbar = b1 * (omc * (nv + C * x1)) + b2 * (sd * x3 - omc * S * x1) +
                               b3 * (-sd * x2);
            const float sbar = b1 * (sd * x3 - omc * C * x2) + b2 * (omc * (S * x2 - nv)) +
                               b3 * (-sd * x1);
            run_k += (double)(dbar * E.g);
            if (g_vec) {
               const float vx = E.v.x, vy = E.v.y, vz = E.v.z;
               const float rxy = vx * vx + vy * vy;
               if (rxy > 0.0f) {
                  const float n2 = rxy + vz * vz, nr = sqrtf(n2);
                  const float in3 = 1.0f / (n2 * nr), ir2 = 1.0f / (rxy * rxy);
                  // delta = K rxy / |v|
                  const float ddx = E.K * (2.0f * vx / nr - rxy * vx * in3);
                  const float ddy = E.K * (2.0f * vy / nr - rxy * vy * in3);
                  const float ddz = -E.K * rxy * vz * in3;
                  // C = (x^2 - y^2) / rxy, S = 2 x y / rxy
                  const float dcx = 4.0f * vx * vy * vy * ir2, dcy = -4.0f * vy * vx * vx * ir2;
                  const float dsx = 2.0f * vy * (vy * vy - vx * vx) * ir2;
                  const float dsy = 2.0f * vx * (vx * vx - vy * vy) * ir2;
                  float w0 = dbar * ddx + cbar * dcx + sbar * dsx;
                  float w1 = dbar * ddy + cbar * dcy + sbar * dsy;
                  float w2 = dbar * ddz;
                  if (T.rotate) { // back through v' = R v
                     const float a = T.R[0] * w0 + T.R[3] * w1 + T.R[6] * w2;
                     const float b = T.R[1] * w0 + T.R[4] * w1 + T.R[7] * w2;
                     const float cc = T.R[2] * w0 + T.R[5] * w1 + T.R[8] * w2;
                     w0 = a;
                     w1 = b;
                     w2 = cc;
                  }
                  if (P.interp == 0 || T.centers) { // the step's axis is the voxel's own
                     float *o = g_vec + 3 * (((int64_t)E.vx * P.ny + E.vy) * P.nz + E.vz);
                     unsafeAtomicAdd(o + 0, w0);
                     unsafeAtomicAdd(o + 1, w1);
                     unsafeAtomicAdd(o + 2, w2);
                  } else { // through the interpolation to its corners
                     get_vec_interp_back(F, rx, ry, rz, E.lab, P.interp == 2,
                                         make_float3(w0, w1, w2), g_vec);
                  }
               }
            }
            // ubar <- (I + sin d K + (1 - cos d) K^2) ubar
            const float nu = C * b1 - S * b2;
            const float k1 = -S * b3, k2 = -C * b3, k3 = S * b1 + C * b2;
            const float q1 = C * nu - b1, q2 = -S * nu - b2, q3 = -b3;
            b1 = b1 + sd * k1 + omc * q1;
            b2 = b2 + sd * k2 + omc * q2;
            b3 = b3 + sd * k3 + omc * q3;
         }
      }
      flush();
   }
   __syncthreads();
   for (int k = threadIdx.x; k < np_lds; k += WG) {
      if (acc[k][0] != 0.0) atomicAdd(&g_props[2 * k], acc[k][0]);
      if (acc[k][1] != 0.0) atomicAdd(&g_props[2 * k + 1], acc[k][1]);
   }
}

// ------------------------------------------------------- adjoint, queued
// The adjoint of the interpolating views (Lerp / Slerp), built on the march's
// wave-cooperative queue.  One thread per ray again, but the costly parts of
// a birefringent step run one step per lane:
//   forward   the queued march of k_march; the evaluation of each queued step
//             also stores its result (cos 2phi, sin 2phi, sin delta,
//             1 - cos delta, g = |v| cos^2(alpha), the tilted axis v) in the
//             ray's row of `res` (index = the ray's count of birefringent
//             steps), and a checkpoint (position, Poincare state, that count)
//             is kept before every CKQ-th step: the position when the ray
//             passes it, the state when the ray applies its first step at or
//             after it;
//   reverse   segment by segment from the last (the rays of a wave take the
//             same number of steps, so they walk the segments together): the
//             labels of the segment's steps are read, the states before each
//             step rebuilt from the checkpoint with the stored results (no
//             interpolation), and the steps walked backwards as in k_adjoint;
//             the gradient of a step's interpolated axis goes to a second
//             per-wave queue, and 64 queued steps at a time take it back
//             through the interpolation to the corners (get_vec_interp_back).
constexpr int CKQ = 4; // steps per checkpoint segment of the queued adjoint

struct StepRes {      // a birefringent step's evaluation, stored by the forward
   float C, S, sd, omc;
   float g;           // |v| cos^2(alpha); -1: zero axis (skipped, no attenuation)
   float vx, vy, vz;  // the axis after the tilt rotation
};

// the reverse's axis-gradient queue: ng queued steps, one per lane, back
// through the interpolation to the corners (out of line: called from the
// unrolled walk back)
template <typename Idx>
__device__ __noinline__ void grad_drain(Field<Idx, ST_F32> F, const uint4 *QA, const uint4 *QB,
                                        const uint4 *QC, uint32_t ghead, uint32_t ng, int lane,
                                        int slerp, float *g_vec) {
   if ((uint32_t)lane < ng) {
      const uint32_t slot = (ghead + (uint32_t)lane) & (QCAP - 1);
      const uint4 a = QA[slot], b = QB[slot], w = QC[slot];
      get_vec_interp_back(F, pack_double(a.x, a.y), pack_double(a.z, a.w), pack_double(b.x, b.y),
                          (int32_t)b.z, slerp,
                          make_float3(__uint_as_float(w.x), __uint_as_float(w.y),
                                      __uint_as_float(w.z)),
                          g_vec);
   }
}

template <typename Idx, bool PROPS_LDS>
__global__ void __launch_bounds__(WG)
    k_adjoint_queue(const Params *__restrict__ prm, const TiltDev *__restrict__ tilts,
                    const PropDev *__restrict__ props, const void *__restrict__ rec_a,
                    const double *__restrict__ grad_images, Ckpt *__restrict__ ck,
                    int64_t ck_stride, StepRes *__restrict__ res, int64_t res_stride,
                    double kfac, double mfac, double *__restrict__ g_props,
                    float *__restrict__ g_vec, int tilt0, int n_sub) {
   __shared__ PropDev sp[MAX_ADJ_PROPS];
   __shared__ double acc[MAX_ADJ_PROPS][2];
   // forward: the march queue (qc: count index, step, owner lane); reverse:
   // the axis-gradient queue (qa, qb: position, label; qc: the gradient)
   __shared__ uint4 qa[WG / 64][QCAP];
   __shared__ uint4 qb[WG / 64][QCAP];
   __shared__ uint4 qc[WG / 64][QCAP];
   const Params &P = *prm;
   const int np_lds = P.n_props < MAX_ADJ_PROPS ? (int)P.n_props : MAX_ADJ_PROPS;
   for (int k = threadIdx.x; k < np_lds; k += WG) {
      sp[k] = props[k];
      acc[k][0] = 0.0;
      acc[k][1] = 0.0;
   }
   __syncthreads();
   auto prop = [&](int32_t l) -> PropDev { return PROPS_LDS ? sp[l] : props[l]; };

   int tilt;
   int64_t tile;
   tile_tilt(blockIdx.x, (unsigned)n_sub, (unsigned)ray_tiles(P.nx, P.ny), tilt, tile);
   tilt += tilt0;
   const int lane = (int)(threadIdx.x & 63u);
   uint4 *const QA = qa[threadIdx.x >> 6];
   uint4 *const QB = qb[threadIdx.x >> 6];
   uint4 *const QC = qc[threadIdx.x >> 6];
   const int64_t ray = tile * WG + threadIdx.x;
   const int64_t wave_ray0 = ray - lane;
   const int64_t i = ray / P.ny, j = ray - i * P.ny;
   const bool valid = i < P.nx;
   const TiltDev &T = tilts[tilt];
   const int64_t nxy = P.nx * P.ny;
   const Field<Idx, ST_F32> F{rec_a, nullptr, (int32_t)P.nx, (int32_t)P.ny, (int32_t)P.nz,
                              P.flip_z};
   const int slerp = P.interp == 2;
   AxisRot rot;
   for (int k = 0; k < 9; k++) rot.R[k] = T.R[k];
   rot.on = T.rotate != 0;
   Ckpt *const my = ck + ((int64_t)tilt * nxy + (valid ? ray : 0)) * ck_stride;
   StepRes *const res_t = res + (int64_t)tilt * nxy * res_stride;

   double px = 0.0, py = 0.0, pz = 0.0;
   if (P.untilt) {
      px = ((double)i + 0.5) - T.half_shift[0];
      py = ((double)j + 0.5) - T.half_shift[1];
   } else if (valid) {
      tilted_start(T, i, j, px, py);
   }
   px = px + T.half_step[0];
   py = py + T.half_step[1];
   pz = pz + T.half_step[2];
   const double sx = T.step[0], sy = T.step[1], sz = T.step[2];
   const double nzd = (double)P.nz;

   // ---------------------------------------------------------------- forward
   float h1 = 0.0f, h2 = 0.0f, h3 = -1.0f, l1 = 0.0f, l2 = 0.0f, l3 = 0.0f;
   float mh = 0.0f, ml = 0.0f;
   int64_t n = 0;        // steps taken
   uint32_t nb = 0;      // birefringent steps queued
   int64_t ck_next = 0;  // first checkpoint whose state is not written yet
   {
      uint32_t pend = 0;
      int npend = 0;
      uint32_t tail = 0, head = 0;
      const double fz0 = floor(pz);
      bool live = valid & (fz0 >= 0.0) & (fz0 < nzd);
      Vox r = F.fetch(px, py, fz0, live);
      for (;;) {
         int32_t lab = r.lab;
         if (lab < 0 || lab >= P.n_props) lab = 0; // ranges are checked before the launch
         const PropDev pd = prop(lab);
         const bool bire = live && !pd.bg;
         const bool go = live && (!bire || npend < QDEPTH);
         const uint64_t gomask = __ballot(go);
         if (gomask == 0 && tail == head) break;
         if (gomask != 0) {
            const uint64_t qmask = __ballot(go && bire);
            if (go) {
               if (n % CKQ == 0) {
                  Ckpt &c = my[n / CKQ];
                  c.p[0] = px;
                  c.p[1] = py;
                  c.p[2] = pz;
                  c.pad = __uint_as_float(nb);
               }
               if (bire) {
                  const uint32_t seq =
                      tail + __builtin_amdgcn_mbcnt_hi(
                                 (uint32_t)(qmask >> 32),
                                 __builtin_amdgcn_mbcnt_lo((uint32_t)qmask, 0u));
                  const uint32_t slot = seq & (QCAP - 1);
                  QA[slot] = make_uint4(__double2loint(px), __double2hiint(px),
                                        __double2loint(py), __double2hiint(py));
                  QB[slot] = make_uint4(__double2loint(pz), __double2hiint(pz), (uint32_t)lab,
                                        0u);
                  QC[slot] = make_uint4(nb, (uint32_t)n, (uint32_t)lane, 0u);
                  pend |= (seq & 255u) << (8 * npend);
                  npend++;
                  nb++;
               } else if (pd.mu_t != 0.0f) {
                  two_sum(mh, ml, pd.mu_t);
               }
               n++;
               px = px + sx;
               py = py + sy;
               pz = pz + sz;
               const double fz = floor(pz);
               live = (fz >= 0.0) & (fz < nzd);
               r = F.fetch(px, py, fz, live);
            }
            tail += (uint32_t)__popcll(qmask);
         }
         const uint32_t queued = tail - head;
         if (queued >= 64 || gomask == 0) {
            const uint32_t nq = queued < 64 ? queued : 64;
            wave_sync();
            if ((uint32_t)lane < nq) {
               const uint32_t slot = (head + (uint32_t)lane) & (QCAP - 1);
               const uint4 a = QA[slot], b = QB[slot], cq = QC[slot];
               const int32_t elab = (int32_t)b.z;
               const PropDev epd = prop(elab);
               const double ex = pack_double(a.x, a.y), ey = pack_double(a.z, a.w);
               const double ez = pack_double(b.x, b.y);
               bool unit;
               float3 v = get_vec_queued<Idx, ST_F32>(F, ex, ey, ez, elab, slerp, unit);
               StepRes o;
               o.C = 1.0f;
               o.S = 0.0f;
               o.sd = 0.0f;
               o.omc = 0.0f;
               o.g = -1.0f;
               o.vx = o.vy = o.vz = 0.0f;
               float mu = 0.0f;
               if (!is_zero(v)) {
                  const float nrm = sqrtf(dot3(v, v));
                  if (rot.on) {
                     const float3 w = v;
                     v.x = rot.R[0] * w.x + rot.R[1] * w.y + rot.R[2] * w.z;
                     v.y = rot.R[3] * w.x + rot.R[4] * w.y + rot.R[5] * w.z;
                     v.z = rot.R[6] * w.x + rot.R[7] * w.y + rot.R[8] * w.z;
                  }
                  const float xx = v.x * v.x, yy = v.y * v.y;
                  const float rxy = xx + yy;
                  const float rr = rxy + v.z * v.z;
                  o.g = nrm * (rxy / rr);
                  if (rxy > 0.0f) {
                     const float inv = 1.0f / rxy;
                     o.C = (xx - yy) * inv;
                     o.S = 2.0f * v.x * v.y * inv;
                  }
                  sin_omc((epd.K * nrm) * (rxy / rr), o.sd, o.omc);
                  o.vx = v.x;
                  o.vy = v.y;
                  o.vz = v.z;
                  mu = epd.mu_t;
               }
               res_t[(wave_ray0 + (int64_t)cq.z) * res_stride + cq.x] = o;
               QA[slot] = make_uint4(__float_as_uint(o.C), __float_as_uint(o.S),
                                     __float_as_uint(o.sd), __float_as_uint(o.omc));
               QB[slot].x = __float_as_uint(mu);
            }
            wave_sync();
            int k = 0;
#pragma unroll
            for (int q = 0; q < QDEPTH; q++)
               k += (q < npend) & (((((pend >> (8 * q)) & 255u) - head) & 255u) < nq);
            for (int q = 0; q < k; q++) {
               const uint32_t slot = (pend >> (8 * q)) & (QCAP - 1);
               // the state before this step closes the checkpoints up to it
               const int64_t step = (int64_t)QC[slot].y;
               while (ck_next * CKQ <= step) {
                  Ckpt &c = my[ck_next];
                  c.u[0] = h1 + l1;
                  c.u[1] = h2 + l2;
                  c.u[2] = h3 + l3;
                  ck_next++;
               }
               const uint4 e = QA[slot];
               const float mu = __uint_as_float(QB[slot].x);
               apply_retarder(__uint_as_float(e.x), __uint_as_float(e.y), __uint_as_float(e.z),
                              __uint_as_float(e.w), h1, h2, h3, l1, l2, l3);
               two_sum(mh, ml, mu);
            }
            pend = (uint32_t)((uint64_t)pend >> (8 * k));
            npend -= k;
            head += nq;
         }
      }
      while (ck_next * CKQ < n) { // no step after these checkpoints changes the state
         Ckpt &c = my[ck_next];
         c.u[0] = h1 + l1;
         c.u[1] = h2 + l2;
         c.u[2] = h3 + l3;
         ck_next++;
      }
   }

   // ------------------------------------------------------------------ seeds
   float b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;
   double mbar = 0.0;
   if (valid) {
      const double A = exp(-((double)mh + (double)ml));
      const double u1 = (double)h1 + (double)l1, u2 = (double)h2 + (double)l2;
      const double base = P.quarter_I0 * A;
      const double *g = grad_images + ((int64_t)T.out * nxy + ray) * P.n_rho;
      double gc = 0.0, gs = 0.0;
      for (int k = 0; k < P.n_rho; k++) {
         const double I = base * (1.0 - (P.c2r[k] * u1 - P.s2r[k] * u2));
         mbar -= g[k] * I;
         gc += g[k] * P.c2r[k];
         gs += g[k] * P.s2r[k];
      }
      b1 = (float)(-base * gc);
      b2 = (float)(base * gs);
   }

   // ---------------------------------------------------------------- reverse
   int32_t run_lab = -1;
   double run_k = 0.0;
   int64_t run_n = 0;
   auto flush = [&]() {
      if (run_lab < 0) return;
      const double dk = run_k * kfac, dm = mbar * (double)run_n * mfac;
      if (run_lab < MAX_ADJ_PROPS) {
         atomicAdd(&acc[run_lab][0], dk);
         atomicAdd(&acc[run_lab][1], dm);
      } else {
         atomicAdd(&g_props[2 * run_lab], dk);
         atomicAdd(&g_props[2 * run_lab + 1], dm);
      }
   };
   StepRes *const mr = res_t + (valid ? ray : 0) * res_stride;
   // every valid ray of the wave took the same number of steps
   int64_t nw = n;
   {
      const uint64_t vm = __ballot(valid);
      const int first = vm ? __builtin_ctzll(vm) : 0;
      nw = (int64_t)__builtin_amdgcn_readlane((int)n, first);
   }
   uint32_t gtail = 0, ghead = 0; // the axis-gradient queue
   auto gdrain = [&](uint32_t ng) {
      wave_sync();
      grad_drain<Idx>(F, QA, QB, QC, ghead, ng, lane, slerp, g_vec);
      wave_sync();
      ghead += ng;
   };
   for (int64_t sgm = (nw - 1) / CKQ; sgm >= 0 && nw > 0; sgm--) {
      // the segment: labels of its steps, and the states before them
      const int len = (int)min((int64_t)CKQ, nw - sgm * CKQ);
      double q0x = 0.0, q0y = 0.0, q0z = 0.0;
      int32_t lab[CKQ];
      bool bi[CKQ];
      float ux[CKQ], uy[CKQ], uz[CKQ];
      uint32_t ri = 0; // the count index of the segment's first birefringent step
      if (valid) {
         const Ckpt c = my[sgm];
         q0x = c.p[0];
         q0y = c.p[1];
         q0z = c.p[2];
         ri = __float_as_uint(c.pad);
         float a1 = c.u[0], a2 = c.u[1], a3 = c.u[2], e1 = 0.0f, e2 = 0.0f, e3 = 0.0f;
         double qx = q0x, qy = q0y, qz = q0z;
         uint32_t rk = ri;
#pragma unroll
         for (int t = 0; t < CKQ; t++) {
            ux[t] = a1 + e1;
            uy[t] = a2 + e2;
            uz[t] = a3 + e3;
            lab[t] = 0;
            bi[t] = false;
            if (t < len) {
               int32_t ix, iy, iz;
               const Vox rr = fetch_own(F, qx, qy, qz, ix, iy, iz);
               const int32_t l = (rr.lab < 0 || rr.lab >= P.n_props) ? 0 : rr.lab;
               lab[t] = l;
               bi[t] = !prop(l).bg;
               if (bi[t]) {
                  const StepRes o = mr[rk++];
                  if (o.g >= 0.0f) {
                     const float w = o.S * a1 + o.C * a2;
                     const float g = o.sd * a3 - o.omc * w;
                     const float d3 = -(o.sd * w + o.omc * a3);
                     two_sum(a1, e1, o.S * g);
                     two_sum(a2, e2, o.C * g);
                     two_sum(a3, e3, d3);
                  }
               }
               qx = qx + sx;
               qy = qy + sy;
               qz = qz + sz;
            }
         }
      }
      // backwards through the segment (uniform over the wave)
      uint32_t rk = ri;
#pragma unroll
      for (int t = 0; t < CKQ; t++)
         rk += (valid && t < len && bi[t]) ? 1u : 0u;
#pragma unroll
      for (int t = CKQ - 1; t >= 0; t--) {
         bool qg = false;
         float w0 = 0.0f, w1 = 0.0f, w2 = 0.0f;
         double rx = q0x, ry = q0y, rz = q0z;
         if (valid && t < len) {
            for (int q = 0; q < t; q++) {
               rx = rx + sx;
               ry = ry + sy;
               rz = rz + sz;
            }
            StepRes o;
            o.g = 0.0f;
            if (bi[t]) o = mr[--rk];
            const bool skipped = bi[t] && o.g < 0.0f; // zero axis: no attenuation either
            if (!skipped) {
               if (lab[t] != run_lab) {
                  flush();
                  run_lab = lab[t];
                  run_k = 0.0;
                  run_n = 0;
               }
               run_n++;
            }
            if (bi[t] && !skipped) {
               const float x1 = ux[t], x2 = uy[t], x3 = uz[t];
               const float sd = o.sd, omc = o.omc, co = 1.0f - omc, C = o.C, S = o.S;
               const float nv = C * x1 - S * x2;
               const float dd1 = co * S * x3 + sd * (C * nv - x1);
               const float dd2 = co * C * x3 + sd * (-S * nv - x2);
               const float dd3 = -co * (C * x2 + S * x1) - sd * x3;
               const float dbar = b1 * dd1 + b2 * dd2 + b3 * dd3;
               const float cbar = b1 * (omc * (nv + C * x1)) + b2 * (sd * x3 - omc * S * x1) +
                                  b3 * (-sd * x2);
               const float sbar = b1 * (sd * x3 - omc * C * x2) + b2 * (omc * (S * x2 - nv)) +
                                  b3 * (-sd * x1);
               run_k += (double)(dbar * o.g);
               if (g_vec) {
                  const float vx = o.vx, vy = o.vy, vz = o.vz;
                  const float rxy = vx * vx + vy * vy;
                  if (rxy > 0.0f) {
                     const float K = prop(lab[t]).K;
                     const float n2 = rxy + vz * vz, nr = sqrtf(n2);
                     const float in3 = 1.0f / (n2 * nr), ir2 = 1.0f / (rxy * rxy);
                     const float ddx = K * (2.0f * vx / nr - rxy * vx * in3);
                     const float ddy = K * (2.0f * vy / nr - rxy * vy * in3);
                     const float ddz = -K * rxy * vz * in3;
                     const float dcx = 4.0f * vx * vy * vy * ir2, dcy = -4.0f * vy * vx * vx * ir2;
                     const float dsx = 2.0f * vy * (vy * vy - vx * vx) * ir2;
                     const float dsy = 2.0f * vx * (vx * vx - vy * vy) * ir2;
                     w0 = dbar * ddx + cbar * dcx + sbar * dsx;
                     w1 = dbar * ddy + cbar * dcy + sbar * dsy;
                     w2 = dbar * ddz;
                     if (rot.on) {
                        const float a = rot.R[0] * w0 + rot.R[3] * w1 + rot.R[6] * w2;
                        const float b = rot.R[1] * w0 + rot.R[4] * w1 + rot.R[7] * w2;
                        const float cc = rot.R[2] * w0 + rot.R[5] * w1 + rot.R[8] * w2;
                        w0 = a;
                        w1 = b;
                        w2 = cc;
                     }
                     qg = true;
                  }
               }
               const float nu = C * b1 - S * b2;
               const float k1 = -S * b3, k2 = -C * b3, k3 = S * b1 + C * b2;
               const float q1 = C * nu - b1, q2 = -S * nu - b2, q3 = -b3;
               b1 = b1 + sd * k1 + omc * q1;
               b2 = b2 + sd * k2 + omc * q2;
               b3 = b3 + sd * k3 + omc * q3;
            }
         }
         // the step's axis gradient to the corners, 64 steps at a time
         const uint64_t gm = __ballot(qg);
         if (qg) {
            const uint32_t gslot =
                (gtail + __builtin_amdgcn_mbcnt_hi((uint32_t)(gm >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)gm, 0u))) &
                (QCAP - 1);
            QA[gslot] = make_uint4(__double2loint(rx), __double2hiint(rx), __double2loint(ry),
                                   __double2hiint(ry));
            QB[gslot] = make_uint4(__double2loint(rz), __double2hiint(rz), (uint32_t)lab[t], 0u);
            QC[gslot] = make_uint4(__float_as_uint(w0), __float_as_uint(w1),
                                   __float_as_uint(w2), 0u);
         }
         gtail += (uint32_t)__popcll(gm);
         if (gtail - ghead >= 64) gdrain(64);
      }
   }
   while (gtail != ghead) gdrain(gtail - ghead < 64 ? gtail - ghead : 64);
   flush();
   __syncthreads();
   for (int k = threadIdx.x; k < np_lds; k += WG) {
      if (acc[k][0] != 0.0) atomicAdd(&g_props[2 * k], acc[k][0]);
      if (acc[k][1] != 0.0) atomicAdd(&g_props[2 * k + 1], acc[k][1]);
   }
}

// Packs label[x][y][z] + vector[x][y][z][3] into the march's [x][z][y]
// storage through a 32 x 32 (y, z) LDS tile, and reduces the label min / max
// of CheckInput (simulator.cpp:149-167) on the way.
template <int ST>
__global__ void __launch_bounds__(256)
    k_pack(const int32_t *__restrict__ label, const float *__restrict__ vec, int64_t nx,
           int64_t ny, int64_t nz, void *__restrict__ rec_a, uint32_t *__restrict__ rec_b,
           int32_t *mm) {
   // one workgroup per (x, 32 rows of y), walking the whole z extent: the
   // source rows (1200 B of labels, 3600 B of axes for nz = 300) are not
   // 128-byte aligned, and a row split over workgroups would fetch its
   // boundary lines twice
   __shared__ int4 tile[32][33];
   const int64_t nyb = (ny + 31) / 32;
   const int64_t x = blockIdx.x / nyb;
   const int64_t y0 = (blockIdx.x % nyb) * 32;
   const int t = threadIdx.x;
   int32_t lo = 0x7fffffff, hi = (int32_t)0x80000000;
   for (int64_t z0 = 0; z0 < nz; z0 += 32) {
      for (int r = 0; r < 4; r++) {
         const int yy = (t >> 5) + 8 * r, zz = t & 31;
         const int64_t y = y0 + yy, z = z0 + zz;
         if (y < ny && z < nz) {
            const int64_t i = (x * ny + y) * nz + z;
            const int32_t l = label[i];
            lo = l < lo ? l : lo;
            hi = l > hi ? l : hi;
            const float3 v = *(const float3 *)(vec + 3 * i);
            rec::encode(ST, l, v.x, v.y, v.z, tile[zz][yy]);
         }
      }
      __syncthreads();
      for (int r = 0; r < 4; r++) {
         const int yy = t & 31, zz = (t >> 5) + 8 * r;
         const int64_t y = y0 + yy, z = z0 + zz;
         if (y < ny && z < nz) {
            const int64_t o = (x * nz + z) * ny + y;
            rec::store(ST, rec_a, rec_b, o, tile[zz][yy]);
         }
      }
      __syncthreads();
   }
   for (int o = 32; o > 0; o >>= 1) {
      const int32_t a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
      lo = a < lo ? a : lo;
      hi = b > hi ? b : hi;
   }
   // only waves that improve the running min / max touch the shared words
   if ((t & 63) == 0) {
      if (lo < __atomic_load_n(&mm[0], __ATOMIC_RELAXED)) atomicMin(&mm[0], lo);
      if (hi > __atomic_load_n(&mm[1], __ATOMIC_RELAXED)) atomicMax(&mm[1], hi);
   }
}

__global__ void k_init_flags(uint32_t *flags, int32_t *mm) {
   flags[0] = 0;
   mm[0] = 0x7fffffff;
   mm[1] = (int32_t)0x80000000;
}

// ------------------------------------------------------------------ host
double pi_sin(double v) { // vm::rot_pi_cases::sin (vemath.hpp:499-508)
   v = std::fmod(v, 2 * M_PI);
   if (v == 0 || v == M_PI || v == -M_PI) return 0;
   if (v == M_PI_2 || v == -(M_PI + M_PI_2)) return 1;
   if (v == M_PI + M_PI_2 || v == -M_PI_2) return -1;
   return std::sin(v);
}
double pi_cos(double v) { // vm::rot_pi_cases::cos (vemath.hpp:510-520)
   v = std::fmod(v, 2 * M_PI);
   if (v == 0) return 1;
   if (v == M_PI_2 || v == M_PI + M_PI_2 || v == -M_PI_2 || v == -(M_PI + M_PI_2)) return 0;
   if (v == M_PI || v == -M_PI) return -1;
   return std::cos(v);
}

struct M3 {
   double a[9];
};
M3 mat_mul(const M3 &A, const M3 &B) {
   M3 R;
   for (int m = 0; m < 3; m++)
      for (int n = 0; n < 3; n++) {
         double acc = 0.0;
         for (int k = 0; k < 3; k++) acc += A.a[3 * m + k] * B.a[3 * k + n];
         R.a[3 * m + n] = acc;
      }
   return R;
}
void mat_vec(const M3 &A, const double v[3], double out[3]) {
   for (int m = 0; m < 3; m++) {
      double acc = 0.0;
      acc = acc + v[0] * A.a[3 * m + 0];
      acc = acc + v[1] * A.a[3 * m + 1];
      acc = acc + v[2] * A.a[3 * m + 2];
      out[m] = acc;
   }
}
// vm::rot_pi_cases::Mat3RotZYZ (vemath.hpp:527-544)
M3 rot_zyz(double g, double b, double a) {
   auto rz = [](double t) {
      const double s = pi_sin(t), c = pi_cos(t);
      return M3{{c, -s, 0, s, c, 0, 0, 0, 1}};
   };
   auto ry = [](double t) {
      const double s = pi_sin(t), c = pi_cos(t);
      return M3{{c, 0, s, 0, 1, 0, -s, 0, c}};
   };
   return mat_mul(rz(g), mat_mul(ry(b), rz(a)));
}

TiltDev make_tilt(const fpl_setup &S, const int64_t dim[3], double theta, double phi) {
   TiltDev T;
   memset(&T, 0, sizeof(T));
   // LightDirectionUnitVector (simulator.cpp:585-590)
   const double dir[3] = {pi_cos(phi) * pi_sin(theta), pi_sin(phi) * pi_sin(theta),
                          pi_cos(theta)};
   for (int k = 0; k < 3; k++) {
      T.step[k] = dir[k] * S.step_size;
      T.half_step[k] = T.step[k] * 0.5;
   }
   const double ct = std::cos(theta);
   const double shift[3] = {dir[0] / ct * (double)dim[2], dir[1] / ct * (double)dim[2],
                            dir[2] / ct * (double)dim[2]};
   T.half_shift[0] = 0.5 * shift[0];
   T.half_shift[1] = 0.5 * shift[1];
   T.shift[0] = shift[0];
   T.shift[1] = shift[1];
   // CalcStartingLightPositionsTilted constants (simulator.cpp:628-656)
   const double ang = std::asin(std::sin(-theta) * S.tissue_refrection);
   const M3 rot = rot_zyz(phi, ang, -phi);
   const M3 rinv = rot_zyz(phi, -ang, -phi);
   const double pc[3] = {dim[0] * 0.5, dim[1] * 0.5, dim[2] * 0.5};
   const double pt0[3] = {-pc[0], -pc[1], pc[2]};
   const double e10[3] = {1, 0, 0}, e20[3] = {0, 1, 0};
   mat_vec(rot, pt0, T.pt);
   mat_vec(rot, e10, T.e1);
   mat_vec(rot, e20, T.e2);
   for (int k = 0; k < 3; k++) T.pc[k] = pc[k];
   for (int k = 0; k < 9; k++) T.rinv[k] = rinv.a[k];
   T.a = T.e1[0];
   T.b = T.e2[0];
   T.c = T.e1[1];
   T.d = T.e2[1];
   T.D = T.a * T.d - T.c * T.b;
   const M3 R = rot_zyz(phi, theta, -phi);
   for (int k = 0; k < 9; k++) T.R[k] = (float)R.a[k];
   T.rotate = theta != 0;
   // theta = 0 and step 1 with the untilted start: points (i + 1/2, j + 1/2,
   // k + 1/2), where all corner indices of GetVec coincide (simulator.cpp:516-521)
   T.centers = (theta == 0 && S.step_size == 1.0 && S.untilt_sensor_view) ? 1 : 0;
   return T;
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

size_t record_bytes(int storage) { return rec::record_bytes(storage); }

Status prepare(const fpl_setup *setup, const int64_t dim[3], const double *props, int64_t n_props,
               const double *tilts, int64_t n_tilts, Block &B) {
   if (!setup || !dim || !props || !tilts) return make_status(FPL_E_INVALID, "null argument");
   const fpl_setup &S = *setup;
   // setup::Setup::Check (setup.hpp:48-71)
   if (S.step_size <= 0) return make_status(FPL_E_INVALID, "step_size <= 0: %f", S.step_size);
   if (S.light_intensity < 0)
      return make_status(FPL_E_INVALID, "light intensity < 0: %f", S.light_intensity);
   if (S.voxel_size <= 0) return make_status(FPL_E_INVALID, "voxel_size <= 0: %f", S.voxel_size);
   if (S.wavelength <= 0) return make_status(FPL_E_INVALID, "wavelength <= 0: %f", S.wavelength);
   if (S.tissue_refrection < 1)
      return make_status(FPL_E_INVALID, "tissue_refrection < 1: %f", S.tissue_refrection);
   if (S.n_rho <= 0 || !S.filter_rotations)
      return make_status(FPL_E_INVALID, "filter_rotations is empty: []");
   if (S.n_rho > MAX_RHO)
      return make_status(FPL_E_INVALID, "at most %d filter_rotations are supported", MAX_RHO);
   if (S.interpolate < 0 || S.interpolate > 2)
      return make_status(FPL_E_INVALID, "Only NN, Lerp or Slerp are supported");
   if (S.field_storage < 0 || S.field_storage > 2)
      return make_status(FPL_E_INVALID, "field_storage must be 0 (f32), 1 (f16) or 2 (angle16)");
   if (dim[0] <= 0 || dim[1] <= 0 || dim[2] <= 0)
      return make_status(FPL_E_INVALID, "global_dim[any] <= 0: [%lld,%lld,%lld]",
                         (long long)dim[0], (long long)dim[1], (long long)dim[2]);
   if (dim[0] >= (1 << 30) || dim[1] >= (1 << 30) || dim[2] >= (1 << 30))
      return make_status(FPL_E_INVALID, "volume dimension too large");
   if (n_props <= 0) return make_status(FPL_E_INVALID, "properties are empty");
   if (n_tilts <= 0 || n_tilts > 65535) return make_status(FPL_E_INVALID, "bad number of tilts");
   // CheckInput (simulator.cpp:169-175)
   if (props[0] != 0) return make_status(FPL_E_BG_DN, "background birefringence has to be 0");
   for (int64_t l = 0; l < n_props; l++)
      if (props[2 * l + 1] < 0)
         return make_status(FPL_E_INVALID, "mu < 0: %f", props[2 * l + 1]);
   {
      const int64_t nyb = (dim[1] + 31) / 32, nzb = (dim[2] + 31) / 32;
      if (dim[0] * nyb >= ((int64_t)1 << 31) || nzb > 65535)
         return make_status(FPL_E_INVALID, "volume too large");
      const int64_t n_tiles = ray_tiles(dim[0], dim[1]);
      if (n_tiles * n_tilts >= ((int64_t)1 << 31))
         return make_status(FPL_E_INVALID, "too many rays");
   }

   const double lambda = S.wavelength * 1e-9;
   const double thickness = S.voxel_size * 1e-6 * S.step_size;

   std::vector<PropDev> pdev(n_props);
   for (int64_t l = 0; l < n_props; l++) {
      const double dn = props[2 * l], mu = props[2 * l + 1] * 1e3;
      PropDev &p = pdev[l];
      p.bg = (dn == 0 || l == 0);
      p.K = (float)(M_PI_2 * (dn * 4.0 * thickness / lambda));
      p.mu_t = (float)(mu * thickness);
      p.pad = 0;
   }
   Params prm;
   memset(&prm, 0, sizeof(prm));
   prm.nx = dim[0];
   prm.ny = dim[1];
   prm.nz = dim[2];
   prm.n_props = n_props;
   prm.n_tilts = (int32_t)n_tilts;
   prm.n_rho = S.n_rho;
   prm.interp = S.interpolate;
   prm.untilt = S.untilt_sensor_view != 0;
   prm.flip_z = S.flip_z != 0;
   prm.quarter_I0 = 0.25 * S.light_intensity;
   for (int k = 0; k < S.n_rho; k++) {
      prm.c2r[k] = std::cos(2.0 * S.filter_rotations[k]);
      prm.s2r[k] = std::sin(2.0 * S.filter_rotations[k]);
   }
   std::vector<TiltDev> tdev(n_tilts);
   for (int64_t t = 0; t < n_tilts; t++) {
      tdev[t] = make_tilt(S, dim, tilts[2 * t], tilts[2 * t + 1]);
      // a light direction parallel to the tissue plane never leaves the volume
      // (the reference loops forever here); reject it
      if (tdev[t].step[2] == 0.0)
         return make_status(FPL_E_INVALID, "tilt theta = %f gives no progress along z",
                            tilts[2 * t]);
      // a direction just short of parallel (theta within ~1e-7 of 90 degrees)
      // ends, but after up to ~1e10 steps per ray: an effectively endless
      // kernel.  Reject trip counts above MAX_RAY_STEPS (nz / |step_z| bounds it)
      const double trips = (double)dim[2] / std::fabs(tdev[t].step[2]);
      if (!(trips <= (double)MAX_RAY_STEPS))
         return make_status(FPL_E_INVALID,
                            "tilt theta = %f needs %.3g z-steps per ray, more than %lld",
                            tilts[2 * t], trips, (long long)MAX_RAY_STEPS);
      tdev[t].trips = (int32_t)tilt_steps(tdev[t], dim[2]);
   }
   // tilts that interpolate (long rays of work) first, so that the short
   // nearest-neighbour workgroups fill the end of the launch
   for (int64_t t = 0; t < n_tilts; t++) tdev[t].out = (int32_t)t;
   std::stable_partition(tdev.begin(), tdev.end(), [&](const TiltDev &T) {
      return prm.interp != 0 && !T.centers;
   });
   prm.n_heavy = 0;
   for (int64_t t = 0; t < n_tilts; t++) prm.n_heavy += (prm.interp != 0 && !tdev[t].centers);
   if (prm.n_heavy == 0 || prm.n_heavy == n_tilts) prm.n_heavy = (int32_t)n_tilts;
   B.off_t = (sizeof(Params) + 255) & ~size_t(255);
   B.off_pr = (B.off_t + sizeof(TiltDev) * n_tilts + 255) & ~size_t(255);
   B.bytes.assign(B.off_pr + sizeof(PropDev) * n_props, 0);
   memcpy(B.bytes.data(), &prm, sizeof(Params));
   memcpy(B.bytes.data() + B.off_t, tdev.data(), sizeof(TiltDev) * n_tilts);
   memcpy(B.bytes.data() + B.off_pr, pdev.data(), sizeof(PropDev) * n_props);
   B.dim[0] = dim[0];
   B.dim[1] = dim[1];
   B.dim[2] = dim[2];
   B.n_tilts = n_tilts;
   B.n_props = n_props;
   B.n_rho = S.n_rho;
   B.storage = S.field_storage;
   return Status();
}

void clear_flags(uint32_t *d_flags, int32_t *d_mm, hipStream_t st) {
   hipLaunchKernelGGL(k_init_flags, dim3(1), dim3(1), 0, st, d_flags, d_mm);
}

void pack(int storage, const int64_t dim[3], const int32_t *label, const float *vector,
          void *packed, uint32_t *packed_b, int32_t *d_mm, hipStream_t st) {
   const dim3 grid((unsigned)(dim[0] * ((dim[1] + 31) / 32)));
   if (storage == ST_F32)
      hipLaunchKernelGGL(k_pack<ST_F32>, grid, dim3(256), 0, st, label, vector, dim[0], dim[1],
                         dim[2], packed, packed_b, d_mm);
   else if (storage == ST_F16)
      hipLaunchKernelGGL(k_pack<ST_F16>, grid, dim3(256), 0, st, label, vector, dim[0], dim[1],
                         dim[2], packed, packed_b, d_mm);
   else
      hipLaunchKernelGGL(k_pack<ST_ANGLE16>, grid, dim3(256), 0, st, label, vector, dim[0],
                         dim[1], dim[2], packed, packed_b, d_mm);
}

bool props_in_lds(const Block &B) { return B.n_props <= MAX_LDS_PROPS; }

void march(const Block &B, const void *d_block, const void *packed, const uint32_t *packed_b,
           double *images, uint32_t *d_flags, hipStream_t st) {
   const int64_t n_tiles = ray_tiles(B.dim[0], B.dim[1]);
   const Params *d_p = (const Params *)d_block;
   const TiltDev *d_t = (const TiltDev *)((const char *)d_block + B.off_t);
   const PropDev *d_pr = (const PropDev *)((const char *)d_block + B.off_pr);
   const dim3 mgrid((unsigned)(n_tiles * B.n_tilts));
   const bool small_idx = B.dim[0] * B.dim[1] * B.dim[2] < ((int64_t)1 << 32);
   const bool props_lds = props_in_lds(B);
#define FPL_MARCH(IDX, STV)                                                                       \
   do {                                                                                           \
      if (props_lds)                                                                              \
         hipLaunchKernelGGL((k_march<IDX, STV, true>), mgrid, dim3(WG), 0, st, d_p, d_t, d_pr,    \
                            packed, packed_b, images, d_flags);                                   \
      else                                                                                        \
         hipLaunchKernelGGL((k_march<IDX, STV, false>), mgrid, dim3(WG), 0, st, d_p, d_t, d_pr,   \
                            packed, packed_b, images, d_flags);                                   \
   } while (0)
   if (B.storage == ST_F32) {
      if (small_idx) FPL_MARCH(uint32_t, ST_F32); else FPL_MARCH(int64_t, ST_F32);
   } else if (B.storage == ST_F16) {
      if (small_idx) FPL_MARCH(uint32_t, ST_F16); else FPL_MARCH(int64_t, ST_F16);
   } else {
      if (small_idx) FPL_MARCH(uint32_t, ST_ANGLE16); else FPL_MARCH(int64_t, ST_ANGLE16);
   }
#undef FPL_MARCH
}

Status check_labels(int32_t lmin, int32_t lmax, int64_t n_props, int storage) {
   if (lmin < 0) return make_status(FPL_E_LABEL_NEG, "label < 0 detected");
   if ((int64_t)lmax >= n_props) return make_status(FPL_E_LABEL_MAX, "label exceed properties.size()");
   if ((storage == ST_F16 && lmax > 0xffff) || (storage == ST_ANGLE16 && lmax > 0x7fff))
      return make_status(FPL_E_INVALID, "label %d does not fit the %s field storage", lmax,
                         storage == ST_F16 ? "16-bit" : "15-bit");
   return Status();
}

Status run(const fpl_setup *setup, const int64_t dim[3], const int32_t *label, const float *vector,
           const double *props, int64_t n_props, const double *tilts, int64_t n_tilts,
           double *images, uint32_t *warn, hipStream_t st) {
   Block B;
   FPL_TRY(prepare(setup, dim, props, n_props, tilts, n_tilts, B));
   if (!label || !vector || !images) return make_status(FPL_E_INVALID, "null array");

   void *small, *stage, *flags;
   FPL_TRY(ws_get(st, WS_SMALL, B.bytes.size(), &small));
   FPL_TRY(ws_get(st, WS_FLAGS, 64, &flags));
   FPL_TRY(host_stage(st, B.bytes.size() + 64, &stage));
   FPL_HIP(hipStreamSynchronize(st));
   memcpy(stage, B.bytes.data(), B.bytes.size());
   FPL_HIP(hipMemcpyAsync(small, stage, B.bytes.size(), hipMemcpyHostToDevice, st));
   uint32_t *d_flags = (uint32_t *)flags;
   int32_t *d_mm = (int32_t *)((char *)flags + 16);

   const int64_t n_vox = dim[0] * dim[1] * dim[2];
   void *packed, *packed_b = nullptr;
   FPL_TRY(ws_get(st, WS_PACKED, record_bytes(B.storage) * (size_t)n_vox, &packed));
   if (B.storage == ST_ANGLE16) FPL_TRY(ws_get(st, WS_PACKED_B, 4 * (size_t)n_vox, &packed_b));
   clear_flags(d_flags, d_mm, st);
   KernelTimer kp("pack", st);
   pack(B.storage, dim, label, vector, packed, (uint32_t *)packed_b, d_mm, st);
   kp.stop();
   KernelTimer kt("march", st);
   march(B, small, packed, (const uint32_t *)packed_b, images, d_flags, st);
   kt.stop();
   FPL_HIP(hipGetLastError());
   FPL_HIP(hipMemcpyAsync(stage, flags, 32, hipMemcpyDeviceToHost, st));
   FPL_HIP(hipStreamSynchronize(st));
   kt.collect();
   kp.collect();
   const uint32_t f = *(uint32_t *)stage;
   const int32_t lmin = ((int32_t *)((char *)stage + 16))[0];
   const int32_t lmax = ((int32_t *)((char *)stage + 16))[1];
   FPL_TRY(check_labels(lmin, lmax, n_props, B.storage));
   if (warn) *warn = f & 0x7fffffffu;
   return Status();
}

// ------------------------------------------------------------ no allocation
// fpl_simulate_noalloc: fpl_simulate whose scratch (the packed records, the
// parameter block and the flag words) is a static device array instead of the
// per-stream workspace, for callers that cannot allocate device memory.  Each
// device holds its own copy from the library's load until the process ends;
// every call on a device shares it (keeping the last call's contents), so
// calls that overlap on one device race.
constexpr size_t NOALLOC_FIELDS = (size_t)64 << 20; // records (+ angle codes)
constexpr size_t NOALLOC_BLOCK = (size_t)64 << 10;  // parameter block, then 64 B of flags
__device__ __align__(256) unsigned char scratch[NOALLOC_FIELDS + NOALLOC_BLOCK];

// The device address of `scratch` on the current device, looked up once per
// device (the array's name in host code is only a stand-in for the symbol).
Status scratch_address(void **out) {
   static std::mutex mu;
   static std::map<int, void *> addr;
   int dev = 0;
   FPL_HIP(hipGetDevice(&dev));
   std::lock_guard<std::mutex> lk(mu);
   auto it = addr.find(dev);
   if (it == addr.end()) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
         return make_status(FPL_E_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                            hipGetErrorString(e));
      it = addr.emplace(dev, p).first;
   }
   *out = it->second;
   return Status();
}

Status run_noalloc(const fpl_setup *setup, const int64_t dim[3], const int32_t *label,
                   const float *vector, const double *props, int64_t n_props,
                   const double *tilts, int64_t n_tilts, double *images, uint32_t *warn,
                   hipStream_t st) {
   Block B;
   FPL_TRY(prepare(setup, dim, props, n_props, tilts, n_tilts, B));
   if (!label || !vector || !images) return make_status(FPL_E_INVALID, "null array");
   const size_t n_vox = (size_t)(dim[0] * dim[1] * dim[2]);
   const size_t off_b = (n_vox * rec::record_bytes(B.storage) + 255) & ~(size_t)255;
   const size_t fields = off_b + n_vox * rec::second_bytes(B.storage);
   if (fields > NOALLOC_FIELDS)
      return make_status(FPL_E_INVALID,
                         "fpl_simulate_noalloc: the packed fields need %zu bytes, the static "
                         "scratch holds %zu (at most %zu voxels in this storage)",
                         fields, NOALLOC_FIELDS,
                         NOALLOC_FIELDS /
                             (rec::record_bytes(B.storage) + rec::second_bytes(B.storage)));
   if (B.bytes.size() > NOALLOC_BLOCK - 64)
      return make_status(FPL_E_INVALID,
                         "fpl_simulate_noalloc: the parameter block needs %zu bytes, the static "
                         "scratch holds %zu (fewer tilts or labels)",
                         B.bytes.size(), NOALLOC_BLOCK - 64);
   void *base = nullptr;
   FPL_TRY(scratch_address(&base));
   char *records = (char *)base;
   uint32_t *codes = B.storage == ST_ANGLE16 ? (uint32_t *)(records + off_b) : nullptr;
   char *block = records + NOALLOC_FIELDS;
   uint32_t *d_flags = (uint32_t *)(block + NOALLOC_BLOCK - 64);
   int32_t *d_mm = (int32_t *)(d_flags + 4);
   FPL_HIP(hipMemcpyAsync(block, B.bytes.data(), B.bytes.size(), hipMemcpyHostToDevice, st));
   clear_flags(d_flags, d_mm, st); // a kernel: the words start from the last call's values
   pack(B.storage, dim, label, vector, records, codes, d_mm, st);
   march(B, block, records, codes, images, d_flags, st);
   FPL_HIP(hipGetLastError());
   uint32_t host[8] = {};
   FPL_HIP(hipMemcpyAsync(host, d_flags, sizeof(host), hipMemcpyDeviceToHost, st));
   FPL_HIP(hipStreamSynchronize(st));
   FPL_TRY(check_labels((int32_t)host[4], (int32_t)host[5], n_props, B.storage));
   if (warn) *warn = host[0] & 0x7fffffffu;
   return Status();
}

// fpl_simulate on march fields (fpl_records.h layout): the setup checks, then
// the label range against the property table (CheckInput before the march, as
// the reference), then the march.
Status run_fields(const fpl_setup *setup, const fpl_fields *Fd, const double *props,
                  int64_t n_props, const double *tilts, int64_t n_tilts, double *images,
                  uint32_t *warn, hipStream_t st) {
   if (!setup || !Fd) return make_status(FPL_E_INVALID, "null argument");
   fpl_setup S = *setup;
   S.field_storage = Fd->storage;
   Block B;
   FPL_TRY(prepare(&S, Fd->dim, props, n_props, tilts, n_tilts, B));
   if (!Fd->records || !images || (B.storage == ST_ANGLE16 && !Fd->codes))
      return make_status(FPL_E_INVALID, "null array");
   FPL_TRY(check_labels(Fd->label_min, Fd->label_max, n_props, B.storage));

   void *small, *stage, *flags;
   FPL_TRY(ws_get(st, WS_SMALL, B.bytes.size(), &small));
   FPL_TRY(ws_get(st, WS_FLAGS, 64, &flags));
   FPL_TRY(host_stage(st, B.bytes.size() + 64, &stage));
   FPL_HIP(hipStreamSynchronize(st));
   memcpy(stage, B.bytes.data(), B.bytes.size());
   FPL_HIP(hipMemcpyAsync(small, stage, B.bytes.size(), hipMemcpyHostToDevice, st));
   uint32_t *d_flags = (uint32_t *)flags;
   clear_flags(d_flags, (int32_t *)((char *)flags + 16), st);
   KernelTimer kt("march", st);
   march(B, small, Fd->records, Fd->codes, images, d_flags, st);
   kt.stop();
   FPL_HIP(hipGetLastError());
   FPL_HIP(hipMemcpyAsync(stage, flags, 4, hipMemcpyDeviceToHost, st));
   FPL_HIP(hipStreamSynchronize(st));
   kt.collect();
   const uint32_t f = *(uint32_t *)stage;
   // the range was checked: a label outside it means the fields changed since
   if (f & 0x80000000u) return make_status(FPL_E_LABEL_MAX, "label exceed properties.size()");
   if (warn) *warn = f & 0x7fffffffu;
   return Status();
}

// Packs caller-owned C-order fields into march fields and records their
// label range (the min / max of CheckInput, simulator.cpp:149-167).
Status pack_fields(const int64_t dim[3], const int32_t *label, const float *vector,
                   fpl_fields *Fd, hipStream_t st) {
   if (!dim || !Fd) return make_status(FPL_E_INVALID, "null argument");
   if (Fd->storage < 0 || Fd->storage > 2)
      return make_status(FPL_E_INVALID, "field_storage must be 0 (f32), 1 (f16) or 2 (angle16)");
   for (int k = 0; k < 3; k++)
      if (dim[k] <= 0 || dim[k] >= (1 << 30))
         return make_status(FPL_E_INVALID, "bad field dimensions");
   if (!label || !vector || !Fd->records || (Fd->storage == ST_ANGLE16 && !Fd->codes))
      return make_status(FPL_E_INVALID, "null array");
   void *stage, *flags;
   FPL_TRY(ws_get(st, WS_FLAGS, 64, &flags));
   FPL_TRY(host_stage(st, 64, &stage));
   uint32_t *d_flags = (uint32_t *)flags;
   int32_t *d_mm = (int32_t *)((char *)flags + 16);
   clear_flags(d_flags, d_mm, st);
   KernelTimer kp("pack", st);
   pack(Fd->storage, dim, label, vector, Fd->records, Fd->codes, d_mm, st);
   kp.stop();
   FPL_HIP(hipGetLastError());
   FPL_HIP(hipMemcpyAsync(stage, d_mm, 8, hipMemcpyDeviceToHost, st));
   FPL_HIP(hipStreamSynchronize(st));
   kp.collect();
   for (int k = 0; k < 3; k++) Fd->dim[k] = dim[k];
   Fd->label_min = ((int32_t *)stage)[0];
   Fd->label_max = ((int32_t *)stage)[1];
   return Status();
}

// z-march trip count of one tilt (simulator.cpp:259-261): the same for every
// ray, which all start at z = 0 + step_z / 2
int64_t tilt_steps(const TiltDev &T, int64_t nz) {
   double z = 0.0;
   z = z + T.half_step[2];
   int64_t n = 0;
   while (std::floor(z) >= 0 && std::floor(z) < (double)nz && n < ((int64_t)1 << 40)) {
      n++;
      z = z + T.step[2];
   }
   return n;
}

Status adjoint(const fpl_setup *setup, const int64_t dim[3], const int32_t *label,
               const float *vector, const double *props, int64_t n_props, const double *tilts,
               int64_t n_tilts, const double *grad_images, double *grad_props, float *grad_vector,
               hipStream_t st) {
   Block B;
   FPL_TRY(prepare(setup, dim, props, n_props, tilts, n_tilts, B));
   if (!label || !vector || !grad_images || !grad_props)
      return make_status(FPL_E_INVALID, "null array");
   if (B.storage != ST_F32)
      return make_status(FPL_E_INVALID, "the adjoint reads the f32 field storage");
   const TiltDev *tdev = (const TiltDev *)(B.bytes.data() + B.off_t);
   int64_t max_steps = 0;
   for (int64_t t = 0; t < n_tilts; t++) max_steps = std::max(max_steps, tilt_steps(tdev[t], dim[2]));
   // the interpolating tilts (first in the block's order) run the queued
   // adjoint, the others k_adjoint; both index checkpoints by the tilt
   int n_iq = 0;
   for (int64_t t = 0; t < n_tilts; t++) n_iq += (setup->interpolate != 0 && !tdev[t].centers);
   const int64_t ck_stride = std::max<int64_t>(1, (max_steps + CKQ - 1) / CKQ);
   const int64_t res_stride = std::max<int64_t>(1, max_steps);
   const double lambda = setup->wavelength * 1e-9;
   const double thickness = setup->voxel_size * 1e-6 * setup->step_size;
   const double kfac = M_PI_2 * (4.0 * thickness / lambda); // dK / d dn
   const double mfac = 1e3 * thickness;                       // d(mu t) / d mu

   void *small, *stage, *flags, *scratch;
   FPL_TRY(ws_get(st, WS_SMALL, B.bytes.size(), &small));
   FPL_TRY(ws_get(st, WS_FLAGS, 64, &flags));
   const size_t gp_bytes = sizeof(double) * 2 * (size_t)n_props;
   FPL_TRY(host_stage(st, std::max(B.bytes.size() + 64, gp_bytes), &stage));
   const int64_t n_vox = dim[0] * dim[1] * dim[2];
   const size_t ck_bytes = sizeof(Ckpt) * (size_t)n_tilts * (size_t)(dim[0] * dim[1]) *
                           (size_t)ck_stride;
   // the queued adjoint's stored step results: a row of max_steps per ray
   const size_t res_bytes = sizeof(StepRes) * (size_t)n_iq * (size_t)(dim[0] * dim[1]) *
                            (size_t)res_stride;
   FPL_TRY(ws_get(st, WS_ADJOINT, ck_bytes + res_bytes + gp_bytes, &scratch));
   Ckpt *d_ck = (Ckpt *)scratch;
   StepRes *d_res = (StepRes *)((char *)scratch + ck_bytes);
   double *d_gp = (double *)((char *)scratch + ck_bytes + res_bytes);
   FPL_HIP(hipStreamSynchronize(st));
   memcpy(stage, B.bytes.data(), B.bytes.size());
   FPL_HIP(hipMemcpyAsync(small, stage, B.bytes.size(), hipMemcpyHostToDevice, st));
   uint32_t *d_flags = (uint32_t *)flags;
   int32_t *d_mm = (int32_t *)((char *)flags + 16);

   void *packed;
   FPL_TRY(ws_get(st, WS_PACKED, record_bytes(B.storage) * (size_t)n_vox, &packed));
   clear_flags(d_flags, d_mm, st);
   pack(B.storage, dim, label, vector, packed, nullptr, d_mm, st);
   FPL_HIP(hipMemcpyAsync(stage, flags, 32, hipMemcpyDeviceToHost, st));
   FPL_HIP(hipStreamSynchronize(st));
   {
      const int32_t lmin = ((int32_t *)((char *)stage + 16))[0];
      const int32_t lmax = ((int32_t *)((char *)stage + 16))[1];
      FPL_TRY(check_labels(lmin, lmax, n_props, B.storage));
   }
   FPL_HIP(hipMemsetAsync(d_gp, 0, gp_bytes, st));
   if (grad_vector) FPL_HIP(hipMemsetAsync(grad_vector, 0, sizeof(float) * 3 * n_vox, st));

   const int64_t n_tiles = ray_tiles(dim[0], dim[1]);
   const Params *d_p = (const Params *)small;
   const TiltDev *d_t = (const TiltDev *)((const char *)small + B.off_t);
   const PropDev *d_pr = (const PropDev *)((const char *)small + B.off_pr);
   KernelTimer kt("adjoint", st);
   const int n_nq = (int)n_tilts - n_iq;
   const dim3 grid_q((unsigned)(n_tiles * (n_iq > 0 ? n_iq : 1)));
   const dim3 grid((unsigned)(n_tiles * (n_nq > 0 ? n_nq : 1)));
#define FPL_ADJ(IDX, PL)                                                                          \
   do {                                                                                           \
      if (n_iq > 0)                                                                               \
         hipLaunchKernelGGL((k_adjoint_queue<IDX, PL>), grid_q, dim3(WG), 0, st, d_p, d_t, d_pr,  \
                            packed, grad_images, d_ck, ck_stride, d_res, res_stride, kfac, mfac,  \
                            d_gp, grad_vector, 0, n_iq);                                          \
      if (n_nq > 0)                                                                               \
         hipLaunchKernelGGL((k_adjoint<IDX, PL>), grid, dim3(WG), 0, st, d_p, d_t, d_pr, packed,  \
                            grad_images, d_ck, ck_stride, kfac, mfac, d_gp, grad_vector, n_iq,    \
                            n_nq);                                                                \
   } while (0)
   const bool props_lds = n_props <= MAX_ADJ_PROPS;
   if (n_vox < ((int64_t)1 << 32)) {
      if (props_lds) FPL_ADJ(uint32_t, true); else FPL_ADJ(uint32_t, false);
   } else {
      if (props_lds) FPL_ADJ(int64_t, true); else FPL_ADJ(int64_t, false);
   }
#undef FPL_ADJ
   kt.stop();
   FPL_HIP(hipGetLastError());
   FPL_HIP(hipMemcpyAsync(stage, d_gp, gp_bytes, hipMemcpyDeviceToHost, st));
   FPL_HIP(hipStreamSynchronize(st));
   kt.collect();
   memcpy(grad_props, stage, gp_bytes);
   return Status();
}

} // namespace sim
} // namespace fpl

extern "C" int fpl_simulate(const fpl_setup *setup, const int64_t dim[3], const int32_t *label,
                            const float *vector, const double *props, int64_t n_props,
                            const double *tilts, int64_t n_tilts, double *images, uint32_t *warn,
                            void *stream, char *msg, size_t msg_len) {
   if (warn) *warn = 0;
   fpl::Status s = fpl::sim::run(setup, dim, label, vector, props, n_props, tilts, n_tilts, images,
                                 warn, (hipStream_t)stream);
   return fpl::finish(s, msg, msg_len);
}

extern "C" int fpl_simulate_adjoint(const fpl_setup *setup, const int64_t dim[3],
                                    const int32_t *label, const float *vector,
                                    const double *props, int64_t n_props, const double *tilts,
                                    int64_t n_tilts, const double *grad_images,
                                    double *grad_props, float *grad_vector, void *stream,
                                    char *msg, size_t msg_len) {
   fpl::Status s = fpl::sim::adjoint(setup, dim, label, vector, props, n_props, tilts, n_tilts,
                                     grad_images, grad_props, grad_vector, (hipStream_t)stream);
   return fpl::finish(s, msg, msg_len);
}

extern "C" int fpl_simulate_fields(const fpl_setup *setup, const fpl_fields *fields,
                                   const double *props, int64_t n_props, const double *tilts,
                                   int64_t n_tilts, double *images, uint32_t *warn, void *stream,
                                   char *msg, size_t msg_len) {
   if (warn) *warn = 0;
   fpl::Status s = fpl::sim::run_fields(setup, fields, props, n_props, tilts, n_tilts, images,
                                        warn, (hipStream_t)stream);
   return fpl::finish(s, msg, msg_len);
}

extern "C" int fpl_pack_fields(const int64_t dim[3], const int32_t *label, const float *vector,
                               fpl_fields *fields, void *stream, char *msg, size_t msg_len) {
   fpl::Status s = fpl::sim::pack_fields(dim, label, vector, fields, (hipStream_t)stream);
   return fpl::finish(s, msg, msg_len);
}

extern "C" int fpl_simulate_noalloc(const fpl_setup *setup, const int64_t dim[3],
                                    const int32_t *label, const float *vector,
                                    const double *props, int64_t n_props, const double *tilts,
                                    int64_t n_tilts, double *images, uint32_t *warn, void *stream,
                                    char *msg, size_t msg_len) {
   if (warn) *warn = 0;
   fpl::Status s = fpl::sim::run_noalloc(setup, dim, label, vector, props, n_props, tilts,
                                         n_tilts, images, warn, (hipStream_t)stream);
   return fpl::finish(s, msg, msg_len);
}
