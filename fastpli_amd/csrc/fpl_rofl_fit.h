// The per-pixel ROFL fit (the reference's local fit, see fpl_rofl.hip), for
// the kernel and for host builds of the same code.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

#include "fpl_lmdif.h"

namespace fpl {
namespace rofl {

constexpr int MAX_RHO = 64;

struct Const {
   double M[4][9];                     // tilt matrices, psi = 0, pi/2, pi, 3pi/2
   double rho[MAX_RHO];                // linspace(0, pi, n + 1)[:-1] + dir_offset
   double es[MAX_RHO], ec[MAX_RHO];    // sin / cos of 2 linspace(0, pi, n, False): EPA direction
   double incl[6], trel[6];            // brute-force grid
   double cos_tau, gain, dir_offset;
   int32_t n_rho, pad;
};

// numpy's float % (result has the sign of the divisor)
__host__ __device__ inline double pymod(double a, double b) {
   double r = fmod(a, b);
   if (r != 0 && ((r < 0) != (b < 0))) r += b;
   return r;
}

// _symmetrize_angles (_ROFL_with_jacobi.py:60-69)
__host__ __device__ inline void symmetrize(double &d, double &a) {
   const double s = -pymod(floor(d / M_PI), 2.0) + 0.5;
   a = (pymod(a + M_PI / 2., M_PI) - M_PI / 2.) * (s > 0 ? 1.0 : (s < 0 ? -1.0 : 0.0));
   d = pymod(d, M_PI);
}

// numpy's sum of a contiguous row (pairwise: below 8 values in order, else
// eight interleaved partial sums combined as a tree, then the rest in order)
template <typename Get> __host__ __device__ inline double np_sum(int n, Get v) {
   if (n < 8) {
      double s = 0.0;
      for (int i = 0; i < n; i++) s += v(i);
      return s;
   }
   double r[8];
   for (int k = 0; k < 8; k++) r[k] = v(k);
   int i = 8;
   for (; i < n - (n % 8); i += 8)
      for (int k = 0; k < 8; k++) r[k] += v(i + k);
   double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
   for (; i < n; i++) s += v(i);
   return s;
}

// The tilted views' angles of the fiber (phi, alpha, t) for view j
// (_calculate_rotated_fiber_params, :164-193), not yet symmetrised
__host__ __device__ inline void tilted(const Const &K, int j, double phi, double alpha, double t,
                                       double &pj, double &aj, double &tj) {
   if (j == 0) {
      pj = phi;
      aj = alpha;
      tj = t;
      return;
   }
   const double v[3] = {cos(alpha) * cos(phi), cos(alpha) * sin(phi), sin(alpha)};
   const double *M = K.M[j - 1];
   double w[3];
   for (int k = 0; k < 3; k++) w[k] = M[3 * k] * v[0] + M[3 * k + 1] * v[1] + M[3 * k + 2] * v[2];
   pj = atan(w[1] / w[0]);
   aj = asin(w[2]);
   tj = t / K.cos_tau;
}

// The weighted residual of one pixel (the function lmdif minimises)
template <int MAXM> struct Residual {
   const Const *K;
   double centered[MAXM], weights[MAXM];
   __host__ __device__ void operator()(const double x[3], double *f) const {
      const int n = K->n_rho;
      for (int j = 0; j < 5; j++) {
         double pj, aj, tj;
         tilted(*K, j, x[0], x[1], x[2], pj, aj, tj);
         symmetrize(pj, aj);
         tj = fabs(tj);
         const double ca = cos(aj);
         const double A = sin(M_PI / 2 * tj * (ca * ca));
         for (int r = 0; r < n; r++) {
            const int k = j * n + r;
            f[k] = weights[k] * (A * sin(2 * (K->rho[r] - pj)) - centered[k]);
         }
      }
   }
   // the reference's Dfun, _weighted_jacobi (:383-400): per view, the
   // derivatives of its intensities w.r.t. the view's own (symmetrised)
   // angles and |t| (_calc_Jacobi, :226-270), weighted; columns = parameters
   __host__ __device__ void jacobian(const double x[3], double *J) const {
      const int n = K->n_rho, m = 5 * n;
      for (int j = 0; j < 5; j++) {
         double pj, aj, tj;
         tilted(*K, j, x[0], x[1], x[2], pj, aj, tj);
         symmetrize(pj, aj);
         tj = fabs(tj);
         const double ca = cos(aj), sa = sin(aj);
         const double arg = M_PI / 2 * tj * (ca * ca);
         const double sA = sin(arg), cA = cos(arg);
         for (int r = 0; r < n; r++) {
            const int k = j * n + r;
            const double s2 = sin(2 * (K->rho[r] - pj)), c2 = cos(2 * (K->rho[r] - pj));
            J[k] = weights[k] * (-2 * c2 * sA);
            J[k + m] = weights[k] * (-M_PI * tj * ca * sa * cA * s2);
            J[k + 2 * m] = weights[k] * (M_PI / 2 * (ca * ca) * s2 * cA);
         }
      }
   }
};

template <typename F> __host__ __device__ inline double sum_sq(int m, const F &f) {
   return np_sum(m, [&](int i) { return f[i] * f[i]; });
}

// One pixel: in(j, r) = the intensity of tilt j, rotation r; out[8] =
// direction, inclination, t_rel, their deviations, chi^2, evaluations
template <int MAXM, bool GRAD, typename In>
__host__ __device__ void fit_pixel(const Const &K, const In &in, double out[8]) {
   const int n = K.n_rho, m = 5 * n;
   Residual<MAXM> R;
   R.K = &K;
   double mean[5];
   for (int j = 0; j < 5; j++) {
      mean[j] = np_sum(n, [&](int r) { return in(j, r); }) / n;
      const double tr = mean[j];
      for (int r = 0; r < n; r++) {
         const double I = in(j, r);
         R.centered[j * n + r] = I / tr - 1.0;
         const double sigma = sqrt(K.gain * (I / (tr * tr) + (I * I) / (pow(tr, 3.0) * n)));
         R.weights[j * n + r] = 1.0 / sigma;
      }
   }
   // EPA direction of the untilted view (epa.py:48-75)
   const double a1 = 2 * np_sum(n, [&](int r) { return in(0, r) * K.es[r]; }) / n;
   const double b1 = 2 * np_sum(n, [&](int r) { return in(0, r) * K.ec[r]; }) / n;
   const double phi0 = pymod(0.5 * atan2(-b1, a1) + M_PI, M_PI) + K.dir_offset;
   // brute-force start: the first grid minimum, inclination major
   double x[3] = {phi0, K.incl[0], K.trel[0]};
   double fvec[MAXM], wa4[MAXM], fjac[3 * MAXM];
   double best = INFINITY;
   for (int a = 0; a < 6; a++)
      for (int b = 0; b < 6; b++) {
         const double xg[3] = {phi0, K.incl[a], K.trel[b]};
         R(xg, fvec);
         const double c = sum_sq(m, fvec);
         if (c < best) {
            best = c;
            x[1] = K.incl[a];
            x[2] = K.trel[b];
         }
      }
   const lm::Result res = lm::minimize<GRAD>(R, m, x, fvec, fjac, wa4, lm::Options());
   const double fvalue = sum_sq(m, fvec);
   double dir = x[0], incl = x[1];
   symmetrize(dir, incl);
   const double trel = fabs(x[2]);
   // deviations (_ROFL_with_jacobi.py:489-522): the per-tilt Jacobian of
   // _calc_Jacobi at the symmetrised result, weighted, J J^T inverted as the
   // reference's _invertmatrix does
   double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
   for (int j = 0; j < 5; j++) {
      double pj, aj, tj;
      tilted(K, j, dir, incl, trel, pj, aj, tj);
      symmetrize(pj, aj);
      tj = fabs(tj);
      const double ca = cos(aj), sa = sin(aj);
      const double arg = M_PI / 2 * tj * (ca * ca);
      const double sA = sin(arg), cA = cos(arg);
      for (int r = 0; r < n; r++) {
         const double w = R.weights[j * n + r];
         const double rho2 = 2 * (K.rho[r] - pj);
         const double s2 = sin(rho2), c2 = cos(rho2);
         const double g[3] = {w * (-2 * c2 * sA), w * (-M_PI * tj * ca * sa * cA * s2),
                              w * (M_PI / 2 * (ca * ca) * s2 * cA)};
         for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) S[3 * a + b] += g[a] * g[b];
      }
   }
   const double det = S[0] * S[4] * S[8] + S[3] * S[7] * S[2] + S[6] * S[5] -
                      S[0] * S[7] * S[5] - S[6] * S[4] * S[2] - S[3] * S[1] * S[8];
   const double inv_diag[3] = {(S[4] * S[8] - S[5] * S[7]) / det,
                               (S[0] * S[8] - S[2] * S[6]) / det,
                               (S[0] * S[4] - S[1] * S[3]) / det};
   const double sigma_sq = fvalue / (double)(m - 3);
   out[0] = dir;
   out[1] = incl;
   out[2] = trel;
   for (int k = 0; k < 3; k++) out[3 + k] = sqrt(fabs(inv_diag[k] * sigma_sq));
   out[6] = fvalue;
   out[7] = (double)res.nfev;
}

// _create_rotation_matrix (_ROFL_with_jacobi.py:72-100)
inline void rotation(double psi, double tau, double *M) {
   const double c = std::cos(psi), s = std::sin(psi), ct = std::cos(tau), st = std::sin(tau);
   const double R[9] = {ct * (c * c) + s * s, (ct - 1) * s * c,     c * st,
                        (ct - 1) * s * c,     ct * (s * s) + c * c, s * st,
                        -c * st,              -s * st,              ct};
   for (int k = 0; k < 9; k++) M[k] = R[k];
}

// The constants of a stack: tilt matrices, rotations, the brute-force grid
inline void make_const(int64_t n_rho, double tilt_angle, double gain, double dir_offset, Const &K) {
   memset(&K, 0, sizeof(K));
   // np.linspace(0, 2 pi, 5)[:-1]
   for (int k = 0; k < 4; k++) rotation(k * (2 * M_PI / 4), tilt_angle, K.M[k]);
   for (int r = 0; r < n_rho; r++) {
      // np.linspace(0, pi, n + 1)[:-1] + dir_offset
      K.rho[r] = r * (M_PI / n_rho) + dir_offset;
      const double e = 2 * (r * (M_PI / n_rho));
      K.es[r] = std::sin(e);
      K.ec[r] = std::cos(e);
   }
   for (int k = 0; k < 6; k++) { // linspace(a, b, 8)[1:-1]
      K.incl[k] = (k + 1) * (M_PI / 7) + -M_PI / 2;
      K.trel[k] = (k + 1) * (0.9 / 7) + 0.0;
   }
   K.cos_tau = std::cos(tilt_angle);
   K.gain = gain;
   K.dir_offset = dir_offset;
   K.n_rho = (int32_t)n_rho;
}

} // namespace rofl
} // namespace fpl
