// FP64 CPU oracle for the fastpli simulation hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: tests/, the smoke entry
// point and bench.py's cpu_baseline leg load it.  The product path
// (fastpli_amd) never links, loads or calls it.
//
// It restates, in plain C++17 + OpenMP and in the reference's own precision
// (FP64 geometry, float distance buffer, FP64 Stokes vectors), the two
// native hot loops of fastpli:
//
//   * tissue generation  - src/simulation/generator.cpp:165-261 (RunTissueGeneration),
//                          :267-383 (FillVoxelsAroundFiberSegment),
//                          :475-494 (closest point on a segment), :496-526 (property list),
//                          src/objects/fiber.cpp:83-99,115-150 (fiber AABB, radius, resize),
//                          src/simulation/fiber_bundle.cpp:65-92,147-154 (layer sort, scale^2),
//                          src/include/aabb.hpp:54-137 (AABB ctor / intersect / overlap),
//                          src/include/vemath.hpp:342-410 (dot = left fold from 0, unit).
//   * simulation         - src/simulation/simulator.cpp:182-411 (RunSimulation),
//                          :28-56 (polarizer / retarder matrices), :420-579 (label /
//                          vector lookup, NN/Lerp/Slerp), :585-773 (light start
//                          positions), src/include/vemath.hpp:498-545 (rot_pi_cases).
//
// Build with -ffp-contract=off so that no multiply-add is fused: every
// floating point operation below rounds exactly once, in the order written.
//
// Pins (tests/test_oracle*.py): the reference's golden fixture (labels, axes,
// 5 Slerp tilts), closed forms and brute force for generation and for NN /
// Slerp / flip_z / step_size / background absorption.  The Lerp blend and the
// tilted-sensor start positions have no reference vectors; they are pinned by
// a library routine and by geometry: Lerp on fields whose blends keep
// non-negative dot products equals scipy's trilinear interpolation, negating
// corners changes the result by its sign only, corners of other labels hand
// the blend to the other corner; each tilted-sensor entry point, moved up the
// light direction and turned by scipy's ZYZ rotation, lies on its pixel's
// vertical line (oracle_light_starts / oracle_get_vec expose the two pieces).
//
// Parallelism mirrors the reference so that the timing is a fair CPU
// baseline: generation lets every thread walk all segments and own the
// x-slabs x % n_threads == tid (generator.cpp:202-205,315-317); simulation
// is a guided parallel-for over rays (simulator.cpp:236-237).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include <omp.h>

namespace {

// ---------------------------------------------------------------- helpers ---

struct V3 {
   double x, y, z;
};

inline V3 sub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(const V3 &a, const V3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scale(const V3 &a, double f) { return {a.x * f, a.y * f, a.z * f}; }
// std::inner_product starting from T{} (vemath.hpp:342-344, 404-406)
inline double dot(const V3 &a, const V3 &b) {
   double acc = 0.0;
   acc = acc + a.x * b.x;
   acc = acc + a.y * b.y;
   acc = acc + a.z * b.z;
   return acc;
}
inline double len2(const V3 &a) { return dot(a, a); }
inline double len(const V3 &a) { return std::sqrt(len2(a)); }
// vm::unit: multiply by 1/sqrt(|u|^2) when |u|^2 > 0 (vemath.hpp:390-402)
inline V3 unit(const V3 &u) {
   double n2 = len2(u);
   if (n2 > 0) {
      double inv = 1.0 / std::sqrt(n2);
      return {u.x * inv, u.y * inv, u.z * inv};
   }
   return u;
}
inline bool is_zero(const V3 &a) { return a.x == 0 && a.y == 0 && a.z == 0; }

struct Box {
   double lo[3] = {std::numeric_limits<double>::infinity(),
                   std::numeric_limits<double>::infinity(),
                   std::numeric_limits<double>::infinity()};
   double hi[3] = {-std::numeric_limits<double>::infinity(),
                   -std::numeric_limits<double>::infinity(),
                   -std::numeric_limits<double>::infinity()};
};
// aabb::AABB(p, q) unsorted constructor (aabb.hpp:54-71)
inline Box box_of(const V3 &p, const V3 &q) {
   Box b;
   const double pp[3] = {p.x, p.y, p.z}, qq[3] = {q.x, q.y, q.z};
   for (int i = 0; i < 3; i++) {
      if (pp[i] < qq[i]) {
         b.lo[i] = pp[i];
         b.hi[i] = qq[i];
      } else {
         b.lo[i] = qq[i];
         b.hi[i] = pp[i];
      }
   }
   return b;
}
inline void unite(Box &a, const Box &b) {
   for (int i = 0; i < 3; i++) {
      if (b.lo[i] < a.lo[i]) a.lo[i] = b.lo[i];
      if (b.hi[i] > a.hi[i]) a.hi[i] = b.hi[i];
   }
}
inline void intersect(Box &a, const Box &b) {
   for (int i = 0; i < 3; i++) {
      if (b.lo[i] > a.lo[i]) a.lo[i] = b.lo[i];
      if (b.hi[i] < a.hi[i]) a.hi[i] = b.hi[i];
   }
}
// aabb::Overlap: strict (aabb.hpp:128-137)
inline bool overlap(const Box &a, const Box &b) {
   for (int i = 0; i < 3; i++)
      if (a.lo[i] >= b.hi[i] || a.hi[i] <= b.lo[i]) return false;
   return true;
}

// ------------------------------------------------------------- generation ---

struct Fib {
   std::vector<V3> p;
   std::vector<double> r;
   Box box;
};

// Fiber::CalculateVoi (fiber.cpp:83-99)
void fiber_box(Fib &f) {
   f.box = Box();
   if (f.p.empty()) return;
   const V3 r0 = {f.r[0], f.r[0], f.r[0]};
   f.box = box_of(sub(f.p[0], r0), add(f.p[0], r0));
   for (size_t i = 0; i + 1 < f.p.size(); i++) {
      double rm = std::max(f.r[i], f.r[i + 1]);
      Box t = box_of(f.p[i], f.p[i + 1]);
      for (int k = 0; k < 3; k++) {
         t.lo[k] -= rm;
         t.hi[k] += rm;
      }
      unite(f.box, t);
   }
}

struct Bundle {
   std::vector<Fib> fibers;
   std::vector<double> scale_sq; // sorted ascending by scale
   std::vector<double> dn, mu;
   std::vector<char> orient;
   Box box;
};

struct ClosestPoint {
   V3 p;
   double t;
};

// PliGenerator::ShortestPointToLineSegmentVecCalculation (generator.cpp:475-494)
inline ClosestPoint closest_on_segment(const V3 &pt, const V3 &s0, const V3 &s1) {
   const V3 v = sub(s1, s0);
   const V3 w = sub(pt, s0);
   const double c1 = dot(w, v);
   if (c1 < 0) return {s0, 0.0};
   const double c2 = dot(v, v);
   if (c2 <= c1) return {s1, 1.0};
   const double b = c1 / c2;
   return {add(s0, scale(v, b)), b};
}

} // namespace

extern "C" {

// Warning bits shared with the product library (same meaning, same values).
enum {
   ORACLE_WARN_OVERLAP = 1u << 0,      // "objects overlap"
   ORACLE_WARN_SINGLE_POINT = 1u << 1, // "fiber object with only single point"
   ORACLE_WARN_ZERO_VEC = 1u << 2,     // "optical_axis is 0 for tissue > 0"
   ORACLE_WARN_NOT_NORM = 1u << 3,     // "optical_axis not normalized"
   ORACLE_WARN_NO_LIGHT = 1u << 4,     // "light_positions is empty"
};

// Returns 0 on success, negative on bad input.
//  points: (n_points, 4) x, y, z, r in micro meter, bundles -> fibers -> points
//  bundle_fiber_ptr[n_bundles+1], fiber_point_ptr[n_fibers+1]
//  bundle_layer_ptr[n_bundles+1] indexes the layer arrays (any order; they are
//  sorted by scale per bundle exactly as fiber::Bundle does)
//  label_out[nx*ny*nz] int32, vector_out[nx*ny*nz*3] float or NULL,
//  props_out[(n_bundles*max_layer+1)*2]
int oracle_generate(int64_t nx, int64_t ny, int64_t nz, const double *origin,
                    double voxel_size, int64_t n_bundles,
                    const int64_t *bundle_fiber_ptr,
                    const int64_t *fiber_point_ptr, const double *points,
                    const int64_t *bundle_layer_ptr, const double *layer_scale,
                    const double *layer_dn, const double *layer_mu,
                    const char *layer_orient, int only_label, int n_threads,
                    int32_t *label_out, float *vector_out, double *props_out,
                    uint32_t *warn_out) {
   if (nx <= 0 || ny <= 0 || nz <= 0 || voxel_size <= 0) return -1;
   uint32_t warn = 0;

   // ---- bundles (fiber_bundle.cpp:76-92): sort layers by scale, store scale^2
   std::vector<Bundle> bundles(n_bundles);
   int64_t max_layer = 0;
   for (int64_t b = 0; b < n_bundles; b++) {
      Bundle &B = bundles[b];
      std::vector<int64_t> order;
      for (int64_t l = bundle_layer_ptr[b]; l < bundle_layer_ptr[b + 1]; l++)
         order.push_back(l);
      std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t c) {
         return layer_scale[a] < layer_scale[c];
      });
      for (int64_t l : order) {
         char o = layer_orient[l];
         if (o != 'b' && o != 'p' && o != 'r') return -2;
         B.scale_sq.push_back(layer_scale[l] * layer_scale[l]);
         B.dn.push_back(layer_dn[l]);
         B.mu.push_back(layer_mu[l]);
         B.orient.push_back(o);
      }
      max_layer = std::max<int64_t>(max_layer, (int64_t)B.scale_sq.size());
      for (int64_t f = bundle_fiber_ptr[b]; f < bundle_fiber_ptr[b + 1]; f++) {
         Fib F;
         for (int64_t i = fiber_point_ptr[f]; i < fiber_point_ptr[f + 1]; i++) {
            F.p.push_back({points[4 * i + 0], points[4 * i + 1], points[4 * i + 2]});
            F.r.push_back(points[4 * i + 3]);
         }
         B.fibers.push_back(std::move(F));
      }
   }

   // ---- property list (generator.cpp:496-526)
   const int64_t n_props = n_bundles * max_layer + 1;
   for (int64_t i = 0; i < 2 * n_props; i++) props_out[i] = 0.0;
   for (int64_t b = 0; b < n_bundles; b++)
      for (size_t l = 0; l < bundles[b].scale_sq.size(); l++) {
         int64_t id = 1 + (int64_t)l + b * max_layer;
         props_out[2 * id + 1] = bundles[b].mu[l];
         props_out[2 * id + 0] = bundles[b].orient[l] == 'b' ? 0.0 : bundles[b].dn[l];
      }

   // ---- translate (only if origin != 0) and resize (generator.cpp:185-197)
   const bool shift = origin[0] != 0 || origin[1] != 0 || origin[2] != 0;
   const V3 off = {-origin[0], -origin[1], -origin[2]};
   const double f = 1.0 / voxel_size;
   for (auto &B : bundles) {
      for (auto &F : B.fibers) {
         if (shift)
            for (auto &p : F.p) p = add(p, off);
         for (auto &p : F.p) p = scale(p, f);
         for (auto &r : F.r) r *= f;
         fiber_box(F);
      }
      B.box = Box();
      if (!B.fibers.empty()) B.box = B.fibers.front().box;
      for (auto &F : B.fibers) unite(B.box, F.box);
   }

   Box vol;
   vol.lo[0] = vol.lo[1] = vol.lo[2] = 0.0;
   vol.hi[0] = (double)nx;
   vol.hi[1] = (double)ny;
   vol.hi[2] = (double)nz;

   const int64_t n_vox = nx * ny * nz;
   std::vector<float> dist(n_vox, std::numeric_limits<float>::infinity());
   for (int64_t i = 0; i < n_vox; i++) label_out[i] = 0;
   if (!only_label)
      for (int64_t i = 0; i < 3 * n_vox; i++) vector_out[i] = 0.0f;

   if (n_threads <= 0) n_threads = 1;
   int overlap_flag = 0;

   // single point warnings are raised once by the caller; detect them here
   for (int64_t b = 0; b < n_bundles; b++) {
      if (!overlap(vol, bundles[b].box)) continue;
      for (auto &F : bundles[b].fibers)
         if (F.p.size() <= 1) warn |= ORACLE_WARN_SINGLE_POINT;
   }

#pragma omp parallel num_threads(n_threads) reduction(| : overlap_flag)
   {
      const int tid = omp_get_thread_num();
      const int nth = omp_get_num_threads();
      for (int64_t b = 0; b < n_bundles; b++) {
         const Bundle &B = bundles[b];
         if (!overlap(vol, B.box)) continue;
         if (B.scale_sq.empty()) continue;
         const double last_sq = B.scale_sq.back();
         const float f_last_sq = (float)last_sq;
         for (const Fib &F : B.fibers) {
            if (F.p.size() <= 1) continue;
            if (!overlap(vol, F.box)) continue;
            for (size_t s = 0; s + 1 < F.p.size(); s++) {
               // FillVoxelsAroundFiberSegment (generator.cpp:267-383)
               const V3 &p = F.p[s];
               const V3 &q = F.p[s + 1];
               const double max_r = std::max(F.r[s], F.r[s + 1]) * std::sqrt(last_sq);
               Box sb = box_of(p, q);
               for (int k = 0; k < 3; k++) {
                  sb.lo[k] -= max_r;
                  sb.hi[k] += max_r;
               }
               intersect(sb, vol);
               int64_t lo[3], hi[3];
               for (int k = 0; k < 3; k++) {
                  lo[k] = (int64_t)sb.lo[k];
                  hi[k] = (int64_t)(sb.hi[k] + 0.5);
               }
               for (int64_t x = lo[0]; x < hi[0]; x++) {
                  if (x % nth != tid) continue;
                  for (int64_t y = lo[1]; y < hi[1]; y++) {
                     for (int64_t z = lo[2]; z < hi[2]; z++) {
                        const V3 pt = {x + 0.5, y + 0.5, z + 0.5};
                        const ClosestPoint cp = closest_on_segment(pt, p, q);
                        const float d2 = (float)len2(sub(cp.p, pt));
                        const double rr = F.r[s] * (1 - cp.t) + F.r[s + 1] * cp.t;
                        const float r2 = (float)(rr * rr);
                        if (d2 > f_last_sq * r2) continue;
                        const int64_t ind = x * ny * nz + y * nz + z;
                        if (dist[ind] >= d2) {
                           // std::lower_bound on the double scale^2 list with a
                           // float ratio (generator.cpp:345-349)
                           const double ratio = (double)(d2 / r2);
                           int ly = 0;
                           while (ly < (int)B.scale_sq.size() && B.scale_sq[ly] < ratio) ly++;
                           const int32_t new_label = (int32_t)(ly + 1 + b * max_layer);
                           const int32_t old = label_out[ind];
                           if (old != 0 && std::abs(old - new_label) >= max_layer)
                              overlap_flag = 1;
                           dist[ind] = d2;
                           label_out[ind] = new_label;
                           if (!only_label) {
                              // past-the-end layer (ratio above the last scale^2 after
                              // float rounding) is read as the outermost layer
                              const int lyo = std::min<int>(ly, (int)B.scale_sq.size() - 1);
                              V3 v = {0, 0, 0};
                              if (B.orient[lyo] == 'r')
                                 v = unit(sub(pt, cp.p));
                              else if (B.orient[lyo] == 'p')
                                 v = unit(sub(q, p));
                              vector_out[3 * ind + 0] = (float)v.x;
                              vector_out[3 * ind + 1] = (float)v.y;
                              vector_out[3 * ind + 2] = (float)v.z;
                           }
                        }
                     }
                  }
               }
            }
         }
      }
   }
   if (overlap_flag) warn |= ORACLE_WARN_OVERLAP;
   if (warn_out) *warn_out = warn;
   return 0;
}

} // extern "C"

// ------------------------------------------------------------- simulation ---
namespace {

// vm::rot_pi_cases::sin / cos (vemath.hpp:498-520): exact values at multiples
// of pi/2 after fmod(v, 2 pi)
double pi_sin(double v) {
   v = std::fmod(v, 2 * M_PI);
   if (v == 0 || v == M_PI || v == -M_PI) return 0;
   if (v == M_PI_2 || v == -(M_PI + M_PI_2)) return 1;
   if (v == M_PI + M_PI_2 || v == -M_PI_2) return -1;
   return std::sin(v);
}
double pi_cos(double v) {
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
V3 mat_vec(const M3 &A, const V3 &v) {
   V3 r;
   double o[3];
   for (int m = 0; m < 3; m++) {
      double acc = 0.0;
      acc = acc + v.x * A.a[3 * m + 0];
      acc = acc + v.y * A.a[3 * m + 1];
      acc = acc + v.z * A.a[3 * m + 2];
      o[m] = acc;
   }
   r.x = o[0];
   r.y = o[1];
   r.z = o[2];
   return r;
}
M3 rot_z(double a) {
   double s = pi_sin(a), c = pi_cos(a);
   return {{c, -s, 0, s, c, 0, 0, 0, 1}};
}
M3 rot_y(double a) {
   double s = pi_sin(a), c = pi_cos(a);
   return {{c, 0, s, 0, 1, 0, -s, 0, c}};
}
// vm::rot_pi_cases::Mat3RotZYZ (vemath.hpp:542-544)
M3 rot_zyz(double g, double b, double a) { return mat_mul(rot_z(g), mat_mul(rot_y(b), rot_z(a))); }

struct Sim {
   int64_t nx, ny, nz;
   const int32_t *label;
   const float *vec;
   int flip_z;
   int interp; // 0 NN, 1 Lerp, 2 Slerp

   int label_at(int64_t x, int64_t y, int64_t z) const {
      if (flip_z) z = nz - 1 - z;
      if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) return 0;
      return label[x * ny * nz + y * nz + z];
   }
   V3 vec_at(int64_t x, int64_t y, int64_t z) const {
      if (flip_z) z = nz - 1 - z;
      const int64_t i = x * ny * nz * 3 + y * nz * 3 + z * 3;
      return {(double)vec[i], (double)vec[i + 1], (double)vec[i + 2]};
   }

   // LinearOrientationInterpolation (simulator.cpp:449-471)
   static V3 lerp(const V3 &v, const V3 &u, double t) {
      if (is_zero(v)) return u;
      if (is_zero(u)) return v;
      const V3 a = scale(v, 1 - t), b = scale(u, t);
      if (dot(a, b) >= 0) return add(a, b);
      return sub(a, b);
   }
   // SphericalInterpolation (simulator.cpp:473-494)
   static V3 slerp(const V3 &v, V3 u, double t) {
      if (is_zero(v)) return u;
      if (is_zero(u)) return v;
      if (dot(v, u) < 0) u = {-u.x, -u.y, -u.z};
      if (len2(sub(v, u)) < 1e-12) return v;
      const double th = std::acos(dot(v, u) / (len(v) * len(u)));
      const double sa = std::sin((1 - t) * th), sb = std::sin(t * th), st = std::sin(th);
      const V3 a = {v.x * sa / st, v.y * sa / st, v.z * sa / st};
      const V3 b = {u.x * sb / st, u.y * sb / st, u.z * sb / st};
      return add(a, b);
   }

   // PliSimulator::GetVec (simulator.cpp:496-579)
   V3 get_vec(const V3 &pos) const {
      const int64_t fx = (int64_t)std::floor(pos.x), fy = (int64_t)std::floor(pos.y),
                    fz = (int64_t)std::floor(pos.z);
      if (interp == 0) return vec_at(fx, fy, fz);
      const int lab = label_at(fx, fy, fz);
      if (lab == 0) return {0, 0, 0};
      auto clampi = [](int64_t v, int64_t n) { return std::min(std::max(v, (int64_t)0), n - 1); };
      const int64_t x0 = clampi((int64_t)std::floor(pos.x - 0.5), nx);
      const int64_t y0 = clampi((int64_t)std::floor(pos.y - 0.5), ny);
      const int64_t z0 = clampi((int64_t)std::floor(pos.z - 0.5), nz);
      const int64_t x1 = clampi((int64_t)std::ceil(pos.x - 0.5), nx);
      const int64_t y1 = clampi((int64_t)std::ceil(pos.y - 0.5), ny);
      const int64_t z1 = clampi((int64_t)std::ceil(pos.z - 0.5), nz);
      if (x0 == x1 && y0 == y1 && z0 == z1) return vec_at(x0, y0, z0);
      double xd = (pos.x - x0 - 0.5) / (x1 - x0);
      double yd = (pos.y - y0 - 0.5) / (y1 - y0);
      double zd = (pos.z - z0 - 0.5) / (z1 - z0);
      if (x0 == x1) xd = 0;
      if (y0 == y1) yd = 0;
      if (z0 == z1) zd = 0;
      auto corner = [&](int64_t x, int64_t y, int64_t z) {
         V3 v = vec_at(x, y, z);
         const double m = (label_at(x, y, z) == lab) ? 1.0 : 0.0;
         return V3{v.x * m, v.y * m, v.z * m};
      };
      const V3 c000 = corner(x0, y0, z0), c100 = corner(x1, y0, z0);
      const V3 c010 = corner(x0, y1, z0), c110 = corner(x1, y1, z0);
      const V3 c001 = corner(x0, y0, z1), c101 = corner(x1, y0, z1);
      const V3 c011 = corner(x0, y1, z1), c111 = corner(x1, y1, z1);
      auto ip = [&](const V3 &a, const V3 &b, double t) {
         return interp == 2 ? slerp(a, b, t) : lerp(a, b, t);
      };
      const V3 c00 = ip(c000, c100, xd), c01 = ip(c001, c101, xd);
      const V3 c10 = ip(c010, c110, xd), c11 = ip(c011, c111, xd);
      const V3 c0 = ip(c00, c10, yd), c1 = ip(c01, c11, yd);
      return ip(c0, c1, zd);
   }
};

struct Ray {
   V3 pos;
   int64_t cx, cy;
};

// CalcStartingLightPositions (simulator.cpp:604-773): the entry point of each
// pixel's ray on the bottom plane z = 0 (before the half step)
std::vector<Ray> light_starts(int64_t nx, int64_t ny, int64_t nz, double theta, double phi,
                              double tissue_refrection, int untilt_sensor_view) {
   const V3 dir = {pi_cos(phi) * pi_sin(theta), pi_sin(phi) * pi_sin(theta), pi_cos(theta)};
   std::vector<Ray> rays;
   rays.reserve(nx * ny);
   const V3 shift = scale(V3{dir.x / std::cos(theta), dir.y / std::cos(theta), dir.z / std::cos(theta)},
                          (double)nz);
   if (untilt_sensor_view) {
      for (int64_t i = 0; i < nx; i++)
         for (int64_t j = 0; j < ny; j++)
            rays.push_back({{(i + 0.5) - 0.5 * shift.x, (j + 0.5) - 0.5 * shift.y, 0.0}, i, j});
   } else {
      const double ang = std::asin(std::sin(-theta) * tissue_refrection);
      const M3 r1 = rot_zyz(phi, ang, -phi);
      const M3 r1_inv = rot_zyz(phi, -ang, -phi);
      const V3 pc = {nx * 0.5, ny * 0.5, nz * 0.5};
      V3 pt = {-pc.x, -pc.y, pc.z};
      V3 e1 = {1, 0, 0}, e2 = {0, 1, 0};
      pt = mat_vec(r1, pt);
      e1 = mat_vec(r1, e1);
      e2 = mat_vec(r1, e2);
      const double a = e1.x, b = e2.x, c = e1.y, d = e2.y;
      for (int64_t i = 0; i < nx; i++)
         for (int64_t j = 0; j < ny; j++) {
            const double e = i + 0.5 - pc.x - pt.x;
            const double ff = j + 0.5 - pc.y - pt.y;
            const double D = a * d - c * b;
            const double Ds = e * d - ff * b;
            const double Dt = a * ff - c * e;
            const double s = Ds / D, t = Dt / D;
            const V3 p = add(add(add(pc, pt), scale(e1, s)), scale(e2, t));
            const V3 tis = add(mat_vec(r1_inv, sub(p, pc)), pc);
            rays.push_back({{tis.x - shift.x, tis.y - shift.y, 0.0}, i, j});
         }
   }
   return rays;
}

} // namespace

extern "C" {

// Test access to two pieces of the march: the rays' entry points on the bottom
// plane (out: (nx, ny, 2) x, y before the half step) and GetVec at one
// position (out: 3 doubles; interp 0 NN, 1 Lerp, 2 Slerp).
void oracle_light_starts(int64_t nx, int64_t ny, int64_t nz, double theta, double phi,
                         double tissue_refrection, int untilt_sensor_view, double *out) {
   const std::vector<Ray> rays =
       light_starts(nx, ny, nz, theta, phi, tissue_refrection, untilt_sensor_view);
   for (const Ray &r : rays) {
      out[2 * (r.cx * ny + r.cy) + 0] = r.pos.x;
      out[2 * (r.cx * ny + r.cy) + 1] = r.pos.y;
   }
}

void oracle_get_vec(int64_t nx, int64_t ny, int64_t nz, const int32_t *label, const float *vec,
                    int flip_z, int interp, double x, double y, double z, double *out) {
   Sim S{nx, ny, nz, label, vec, flip_z, interp};
   const V3 v = S.get_vec(V3{x, y, z});
   out[0] = v.x;
   out[1] = v.y;
   out[2] = v.z;
}

// interp: 0 NN, 1 Lerp, 2 Slerp.  props: (n_props, 2) dn, mu.
// out: (nx, ny, n_rho) float64, C order.  Returns 0, or a negative code:
//  -1 setup invalid, -2 label < 0, -3 label >= n_props, -4 background dn != 0,
//  -5 mu < 0, -6 bad dims.
int oracle_simulate(int64_t nx, int64_t ny, int64_t nz, const int32_t *label,
                    const float *vec, const double *props, int64_t n_props,
                    double step_size, double light_intensity, double voxel_size,
                    double wavelength, double tissue_refrection, int interp,
                    int untilt_sensor_view, int flip_z, const double *rho,
                    int64_t n_rho, double theta, double phi, int n_threads,
                    double *out, uint32_t *warn_out) {
   // setup::Setup::Check (setup.hpp:48-71)
   if (step_size <= 0 || light_intensity < 0 || voxel_size <= 0 || wavelength <= 0 ||
       tissue_refrection < 1 || n_rho <= 0)
      return -1;
   if (nx <= 0 || ny <= 0 || nz <= 0) return -6;
   const int64_t n_vox = nx * ny * nz;
   // CheckInput (simulator.cpp:131-176)
   int32_t lmin = std::numeric_limits<int32_t>::max(), lmax = std::numeric_limits<int32_t>::min();
   for (int64_t i = 0; i < n_vox; i++) {
      lmin = std::min(lmin, label[i]);
      lmax = std::max(lmax, label[i]);
   }
   if (lmin < 0) return -2;
   if ((int64_t)lmax >= n_props) return -3;
   if (props[0] != 0) return -4;
   for (int64_t i = 0; i < n_props; i++)
      if (props[2 * i + 1] < 0) return -5;

   uint32_t warn = 0;
   Sim S{nx, ny, nz, label, vec, flip_z, interp};

   const double lambda = wavelength * 1e-9;
   const double thickness = voxel_size * 1e-6 * step_size;

   // S_in = RetarderMatrix(pi/2, -pi/2) . PolX(1) . (I0, 0, 0, 0)
   // (simulator.cpp:28-56, 207-214); PolX(1) gives (I0/2, I0/2, 0, 0) and
   // the quarter wave plate with beta = pi/2, ret = -pi/2 (exact pi-cases)
   // maps it to (I0/2, 0, 0, -I0/2).
   double s_in[4];
   {
      const double half = 0.5;
      const double p0 = (1 * light_intensity + 1 * 0.0) * half; // row 0 of PolX(1)
      const double p1 = (1 * light_intensity + 1 * 0.0) * half; // row 1
      const double cb = pi_cos(M_PI_2), sbt = pi_sin(M_PI_2);
      const double cr = pi_cos(-M_PI_2), sr = pi_sin(-M_PI_2);
      const double R[16] = {1, 0, 0, 0,
                            0, cb * cb + cr * sbt * sbt, (1 - cr) * sbt * cb, -sr * sbt,
                            0, (1 - cr) * sbt * cb, sbt * sbt + cr * cb * cb, sr * cb,
                            0, sr * sbt, -sr * cb, cr};
      const double in[4] = {p0, p1, 0.0, 0.0};
      for (int m = 0; m < 4; m++) {
         double acc = 0.0;
         for (int k = 0; k < 4; k++) acc = acc + in[k] * R[4 * m + k];
         s_in[m] = acc;
      }
   }

   const M3 rot = rot_zyz(phi, theta, -phi);
   const V3 dir = {pi_cos(phi) * pi_sin(theta), pi_sin(phi) * pi_sin(theta), pi_cos(theta)};
   const V3 step = scale(dir, step_size);

   // starting positions (simulator.cpp:604-773)
   std::vector<Ray> rays = light_starts(nx, ny, nz, theta, phi, tissue_refrection,
                                        untilt_sensor_view);
   if (rays.empty()) warn |= ORACLE_WARN_NO_LIGHT;
   for (auto &r : rays) r.pos = add(r.pos, scale(step, 0.5));

   for (int64_t i = 0; i < nx * ny * n_rho; i++) out[i] = -std::numeric_limits<double>::infinity();
   if (n_threads <= 0) n_threads = 1;
   uint32_t warn_loop = 0;

#pragma omp parallel for schedule(guided) num_threads(n_threads) reduction(| : warn_loop)
   for (size_t r = 0; r < rays.size(); r++) {
      V3 pos = rays[r].pos;
      std::vector<double> s(4 * n_rho);
      for (int64_t k = 0; k < n_rho; k++)
         for (int c = 0; c < 4; c++) s[4 * k + c] = s_in[c];
      for (; std::floor(pos.z) >= 0 && std::floor(pos.z) < nz; pos = add(pos, step)) {
         const int lab = S.label_at((int64_t)std::floor(pos.x), (int64_t)std::floor(pos.y),
                                    (int64_t)std::floor(pos.z));
         const double dn = props[2 * lab + 0];
         const double mu = props[2 * lab + 1] * 1e3;
         const double att = std::exp(-mu * thickness);
         if (dn == 0 || lab == 0) {
            if (mu == 0) continue;
            for (int64_t i = 0; i < 4 * n_rho; i++) s[i] *= att;
            continue;
         }
         V3 v = S.get_vec(pos);
         if (is_zero(v)) {
            warn_loop |= ORACLE_WARN_ZERO_VEC;
            continue;
         }
         if (len2(v) - 1 > 1e-5) warn_loop |= ORACLE_WARN_NOT_NORM;
         const double d_rel = dn * 4.0 * thickness / lambda * len(v);
         if (theta != 0) v = mat_vec(rot, v);
         const double alpha = std::asin(v.z / len(v));
         const double ca = std::cos(alpha);
         const double ret = M_PI_2 * d_rel * (ca * ca);
         const double sin_r = std::sin(ret), cos_r = std::cos(ret);
         const double ph = std::atan2(v.y, v.x);
         for (int64_t k = 0; k < n_rho; k++) {
            const double beta = 2 * (rho[k] - ph);
            const double sb = std::sin(-beta), cb = std::cos(-beta);
            double *sv = &s[4 * k];
            const double a1 = sv[1] * cb - sv[2] * sb;
            const double c1 = sv[1] * sb + sv[2] * cb;
            const double b1 = c1 * cos_r + sv[3] * sin_r;
            const double n0 = sv[0];
            const double n1 = a1 * cb + b1 * sb;
            const double n2 = -a1 * sb + b1 * cb;
            const double n3 = -c1 * sin_r + sv[3] * cos_r;
            sv[0] = n0 * att;
            sv[1] = n1 * att;
            sv[2] = n2 * att;
            sv[3] = n3 * att;
         }
      }
      // PolY(1) . s -> row 0 = 0.5 * (s0 - s1)
      const int64_t ccd = rays[r].cx * ny + rays[r].cy;
      for (int64_t k = 0; k < n_rho; k++) {
         const double *sv = &s[4 * k];
         double acc = 0.0;
         acc = acc + sv[0] * 0.5;
         acc = acc + sv[1] * -0.5;
         acc = acc + sv[2] * 0.0;
         acc = acc + sv[3] * 0.0;
         out[ccd * n_rho + k] = acc;
      }
   }
   warn |= warn_loop;
   if (warn_out) *warn_out = warn;
   return 0;
}

// Number of z-march iterations of one ray (simulator.cpp:259-261), summed
// over all rays of a tilt: the unit of the "Mueller voxel-steps" metric is
// this count times n_rho.
int64_t oracle_count_steps(int64_t nx, int64_t ny, int64_t nz, double step_size,
                           double theta, double phi) {
   const V3 dir = {pi_cos(phi) * pi_sin(theta), pi_sin(phi) * pi_sin(theta), pi_cos(theta)};
   const V3 step = scale(dir, step_size);
   (void)nx;
   (void)ny;
   double z = 0.0 + step.z * 0.5;
   int64_t n = 0;
   for (; std::floor(z) >= 0 && std::floor(z) < nz; z = z + step.z) n++;
   return n * nx * ny;
}

} // extern "C"
