// ROFL tilt analysis (Schmitz et al., Front. Neuroanat. 2018,
// doi:10.3389/fnana.2018.00075) of a 5-tilt x n_rho stack per pixel:
// fastpli's analysis.rofl.rofl (src/fastpli/analysis/rofl.py:17-60) with the
// local fit of src/fastpli/analysis/_ROFL_with_jacobi.py:403-522.
//
// One thread per pixel, FP64:
//   * centered intensities I / mean_tilt - 1 and weights
//     1 / sqrt(gain (I / mean^2 + I^2 / (mean^3 n_rho)))  (:47-57);
//   * start: the EPA direction of the untilted view (rofl.py:55) and the first
//     minimum of the weighted squared residual over the 6 x 6 grid of
//     inclinations linspace(-pi/2, pi/2, 8)[1:-1] and t_rel
//     linspace(0, 0.9, 8)[1:-1] (:313-370);
//   * model (:164-223): fiber vector v(phi, alpha) rotated by the four tilt
//     matrices (:72-100), tilted angles (atan(w1 / w0), asin(w2)), t_rel /
//     cos(tau) for the tilted views; I = sin(pi/2 |t| cos^2 alpha)
//     sin(2 (rho - phi));
//   * minimisation: Levenberg-Marquardt with the exact chain-rule Jacobian,
//     run to a relative step of 1e-14 (the reference's MINPACK lmdif stops at
//     1.49e-8 with finite differences, so results agree to that level);
//   * angles symmetrised as :60-69, |t_rel|, and the parameter deviations from
//     the reference's per-tilt Jacobian (:226-270) and its 3 x 3 inverse
//     (:14-40), computed as the reference computes them (its determinant's
//     third term is m20 * m12, without m01).
#include <cmath>
#include <vector>

#include "fpl_internal.h"

namespace fpl {
namespace rofl {

constexpr int MAX_RHO = 64;

struct Const {
   double M[4][9];                     // tilt matrices, psi = 0, pi/2, pi, 3pi/2
   double rho[MAX_RHO];                // linspace(0, pi, n + 1)[:-1] + dir_offset
   double s2r[MAX_RHO], c2r[MAX_RHO];  // sin / cos 2 rho
   double es[MAX_RHO], ec[MAX_RHO];    // sin / cos of 2 linspace(0, pi, n, False): EPA direction
   double incl[6], trel[6];            // brute-force grid
   double inv_cos_tau, gain, dir_offset;
   int32_t n_rho, pad;
};

// numpy's float % (result has the sign of the divisor)
__device__ __forceinline__ double pymod(double a, double b) {
   double r = fmod(a, b);
   if (r != 0 && ((r < 0) != (b < 0))) r += b;
   return r;
}

// _symmetrize_angles (_ROFL_with_jacobi.py:60-69)
__device__ __forceinline__ void symmetrize(double &d, double &a) {
   const double s = -pymod(floor(d / M_PI), 2.0) + 0.5;
   a = (pymod(a + M_PI / 2., M_PI) - M_PI / 2.) * (s > 0 ? 1.0 : (s < 0 ? -1.0 : 0.0));
   d = pymod(d, M_PI);
}

// tilted angles of view j and their derivatives w.r.t. (phi, alpha)
struct Tilted {
   double phi, alpha, t, dphi[2], dalpha[2], dt;
};

__device__ __forceinline__ Tilted tilted(const Const &K, int j, double phi, double alpha,
                                         double t) {
   Tilted o;
   const double ca = cos(alpha), sa = sin(alpha), cf = cos(phi), sf = sin(phi);
   if (j == 0) {
      o.phi = phi;
      o.alpha = alpha;
      o.t = t;
      o.dphi[0] = 1;
      o.dphi[1] = 0;
      o.dalpha[0] = 0;
      o.dalpha[1] = 1;
      o.dt = 1;
      return o;
   }
   const double v[3] = {ca * cf, ca * sf, sa};
   const double dvf[3] = {-ca * sf, ca * cf, 0.0};
   const double dva[3] = {-sa * cf, -sa * sf, ca};
   const double *M = K.M[j - 1];
   double w[3], wf[3], wa[3];
   for (int k = 0; k < 3; k++) {
      w[k] = M[3 * k] * v[0] + M[3 * k + 1] * v[1] + M[3 * k + 2] * v[2];
      wf[k] = M[3 * k] * dvf[0] + M[3 * k + 1] * dvf[1] + M[3 * k + 2] * dvf[2];
      wa[k] = M[3 * k] * dva[0] + M[3 * k + 1] * dva[1] + M[3 * k + 2] * dva[2];
   }
   o.phi = atan(w[1] / w[0]);
   o.alpha = asin(w[2]);
   o.t = t * K.inv_cos_tau;
   const double q = w[0] * w[0] + w[1] * w[1];
   o.dphi[0] = (w[0] * wf[1] - w[1] * wf[0]) / q;
   o.dphi[1] = (w[0] * wa[1] - w[1] * wa[0]) / q;
   const double rs = 1.0 / sqrt(fmax(1.0 - w[2] * w[2], 1e-300));
   o.dalpha[0] = wf[2] * rs;
   o.dalpha[1] = wa[2] * rs;
   o.dt = K.inv_cos_tau;
   return o;
}

// weighted squared residual at x; with jac, also J^T J (xx, xy, xz, yy, yz,
// zz) and J^T r
__device__ double evaluate(const Const &K, const double *in, int64_t n_pix, int64_t p,
                           const double mean[5], const double x[3], bool jac, double JtJ[6],
                           double Jtr[3]) {
   const int n = K.n_rho;
   double cost = 0.0;
   if (jac)
      for (int k = 0; k < 6; k++) JtJ[k] = 0.0;
   if (jac) Jtr[0] = Jtr[1] = Jtr[2] = 0.0;
   for (int j = 0; j < 5; j++) {
      const Tilted T = tilted(K, j, x[0], x[1], x[2]);
      const double ca = cos(T.alpha), sa = sin(T.alpha);
      const double at = fabs(T.t), sg = T.t < 0 ? -1.0 : 1.0;
      const double A = M_PI / 2 * at * ca * ca;
      const double sA = sin(A), cA = cos(A);
      const double c2f = cos(2 * T.phi), s2f = sin(2 * T.phi);
      const double m_ = mean[j];
      const double *row = in + ((int64_t)j * n_pix + p) * n;
      for (int r = 0; r < n; r++) {
         const double I = row[r];
         const double s2 = K.s2r[r] * c2f - K.c2r[r] * s2f; // sin 2 (rho - phi)
         const double w = 1.0 / sqrt(K.gain * (I / (m_ * m_) + I * I / (m_ * m_ * m_ * n)));
         const double res = w * (sA * s2 - (I / m_ - 1.0));
         cost += res * res;
         if (jac) {
            const double c2 = K.c2r[r] * c2f + K.s2r[r] * s2f; // cos 2 (rho - phi)
            const double dmf = -2.0 * sA * c2;
            const double dma = -M_PI * at * ca * sa * cA * s2;
            const double dmt = M_PI / 2 * ca * ca * s2 * cA * sg;
            const double g0 = w * (dmf * T.dphi[0] + dma * T.dalpha[0]);
            const double g1 = w * (dmf * T.dphi[1] + dma * T.dalpha[1]);
            const double g2 = w * dmt * T.dt;
            JtJ[0] += g0 * g0;
            JtJ[1] += g0 * g1;
            JtJ[2] += g0 * g2;
            JtJ[3] += g1 * g1;
            JtJ[4] += g1 * g2;
            JtJ[5] += g2 * g2;
            Jtr[0] += g0 * res;
            Jtr[1] += g1 * res;
            Jtr[2] += g2 * res;
         }
      }
   }
   return cost;
}

__device__ __forceinline__ double det3(const double m[9]) {
   return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
          m[2] * (m[3] * m[7] - m[4] * m[6]);
}

__global__ void k_rofl(Const K, const double *__restrict__ in, int64_t n_pix,
                       const uint8_t *__restrict__ mask, double *__restrict__ out) {
   const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (p >= n_pix) return;
   if (mask && !mask[p]) {
      for (int k = 0; k < 8; k++) out[k * n_pix + p] = NAN;
      return;
   }
   const int n = K.n_rho;
   double mean[5];
   for (int j = 0; j < 5; j++) {
      const double *row = in + ((int64_t)j * n_pix + p) * n;
      double s = 0.0;
      for (int r = 0; r < n; r++) s += row[r];
      mean[j] = s / n;
   }
   // EPA direction of the untilted view (epa.py:48-75)
   double a1 = 0.0, b1 = 0.0;
   {
      const double *row = in + p * n;
      for (int r = 0; r < n; r++) {
         a1 += row[r] * K.es[r];
         b1 += row[r] * K.ec[r];
      }
      a1 = 2 * a1 / n;
      b1 = 2 * b1 / n;
   }
   const double phi0 = pymod(0.5 * atan2(-b1, a1) + M_PI, M_PI) + K.dir_offset;
   // brute-force start: the first grid minimum, inclination major
   double x[3] = {phi0, K.incl[0], K.trel[0]};
   double best = INFINITY;
   double dummy6[6], dummy3[3];
   for (int a = 0; a < 6; a++)
      for (int b = 0; b < 6; b++) {
         const double xg[3] = {phi0, K.incl[a], K.trel[b]};
         const double c = evaluate(K, in, n_pix, p, mean, xg, false, dummy6, dummy3);
         if (c < best) {
            best = c;
            x[1] = K.incl[a];
            x[2] = K.trel[b];
         }
      }
   // Levenberg-Marquardt
   double JtJ[6], Jtr[3];
   double cost = evaluate(K, in, n_pix, p, mean, x, true, JtJ, Jtr);
   int n_eval = 37;
   double lambda = 1e-3;
   for (int it = 0; it < 200; it++) {
      double A[9] = {JtJ[0], JtJ[1], JtJ[2], JtJ[1], JtJ[3], JtJ[4], JtJ[2], JtJ[4], JtJ[5]};
      A[0] += lambda * fmax(JtJ[0], 1e-300);
      A[4] += lambda * fmax(JtJ[3], 1e-300);
      A[8] += lambda * fmax(JtJ[5], 1e-300);
      const double D = det3(A);
      if (!(D != 0) || !isfinite(D)) break;
      // Cramer's rule for A d = -J^T r
      double d[3];
      for (int c = 0; c < 3; c++) {
         double B[9];
         for (int k = 0; k < 9; k++) B[k] = A[k];
         for (int r = 0; r < 3; r++) B[3 * r + c] = -Jtr[r];
         d[c] = det3(B) / D;
      }
      const double xn[3] = {x[0] + d[0], x[1] + d[1], x[2] + d[2]};
      double JtJn[6], Jtrn[3];
      const double cn = evaluate(K, in, n_pix, p, mean, xn, true, JtJn, Jtrn);
      n_eval++;
      if (cn < cost) {
         const double drop = cost - cn;
         for (int k = 0; k < 3; k++) x[k] = xn[k];
         for (int k = 0; k < 6; k++) JtJ[k] = JtJn[k];
         for (int k = 0; k < 3; k++) Jtr[k] = Jtrn[k];
         cost = cn;
         lambda = fmax(lambda * 0.1, 1e-15);
         const double step = fabs(d[0]) + fabs(d[1]) + fabs(d[2]);
         const double size = fabs(x[0]) + fabs(x[1]) + fabs(x[2]);
         if (step <= 1e-14 * (size + 1e-14) || drop <= 1e-16 * cost) break;
      } else {
         lambda *= 10.0;
         if (lambda > 1e16) break;
      }
   }
   double dir = x[0], incl = x[1];
   symmetrize(dir, incl);
   const double trel = fabs(x[2]);
   // deviations (_ROFL_with_jacobi.py:489-522): the per-tilt Jacobian of
   // _calc_Jacobi at the symmetrised result, weighted, J J^T inverted as the
   // reference's _invertmatrix does
   double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
   for (int j = 0; j < 5; j++) {
      const Tilted T = tilted(K, j, dir, incl, trel);
      double pj = T.phi, aj = T.alpha;
      symmetrize(pj, aj);
      const double tj = fabs(T.t);
      const double ca = cos(aj), sa = sin(aj);
      const double arg = M_PI / 2 * tj * ca * ca;
      const double sA = sin(arg), cA = cos(arg);
      const double m_ = mean[j];
      const double *row = in + ((int64_t)j * n_pix + p) * n;
      for (int r = 0; r < n; r++) {
         const double I = row[r];
         const double w = 1.0 / sqrt(K.gain * (I / (m_ * m_) + I * I / (m_ * m_ * m_ * n)));
         const double rho2 = 2 * (K.rho[r] - pj);
         const double s2 = sin(rho2), c2 = cos(rho2);
         const double g[3] = {w * (-2 * c2 * sA), w * (-M_PI * tj * ca * sa * cA * s2),
                              w * (M_PI / 2 * ca * ca * s2 * cA)};
         for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) S[3 * a + b] += g[a] * g[b];
      }
   }
   const double det = S[0] * S[4] * S[8] + S[3] * S[7] * S[2] + S[6] * S[5] -
                      S[0] * S[7] * S[5] - S[6] * S[4] * S[2] - S[3] * S[1] * S[8];
   const double inv_diag[3] = {(S[4] * S[8] - S[5] * S[7]) / det,
                               (S[0] * S[8] - S[2] * S[6]) / det,
                               (S[0] * S[4] - S[1] * S[3]) / det};
   const double sigma_sq = cost / (double)(5 * n - 3);
   out[0 * n_pix + p] = dir;
   out[1 * n_pix + p] = incl;
   out[2 * n_pix + p] = trel;
   for (int k = 0; k < 3; k++) out[(3 + k) * n_pix + p] = sqrt(fabs(inv_diag[k] * sigma_sq));
   out[6 * n_pix + p] = cost;
   out[7 * n_pix + p] = (double)n_eval;
}

// _create_rotation_matrix (_ROFL_with_jacobi.py:72-100)
void rotation(double psi, double tau, double *M) {
   const double c = std::cos(psi), s = std::sin(psi), ct = std::cos(tau), st = std::sin(tau);
   const double R[9] = {ct * c * c + s * s, (ct - 1) * s * c,   c * st,
                        (ct - 1) * s * c,   ct * s * s + c * c, s * st,
                        -c * st,            -s * st,            ct};
   for (int k = 0; k < 9; k++) M[k] = R[k];
}

Status run(const double *in, int64_t n_pix, int64_t n_rho, double tilt_angle, double gain,
           double dir_offset, const uint8_t *mask, double *out, hipStream_t st) {
   if (!in || !out) return make_status(FPL_E_INVALID, "null array");
   if (n_rho <= 3) return make_status(FPL_E_INVALID, "data needs at least 3 equidistant rotations");
   if (n_rho > MAX_RHO)
      return make_status(FPL_E_INVALID, "at most %d rotations are supported", MAX_RHO);
   if (gain <= 1) return make_status(FPL_E_INVALID, "rofl gain <= 1");
   if (n_pix < 0) return make_status(FPL_E_INVALID, "bad sizes");
   if (n_pix == 0) return Status();
   Const K;
   memset(&K, 0, sizeof(K));
   // np.linspace(0, 2 pi, 5)[:-1]
   for (int k = 0; k < 4; k++) rotation(k * (2 * M_PI / 4), tilt_angle, K.M[k]);
   for (int r = 0; r < n_rho; r++) {
      // np.linspace(0, pi, n + 1)[:-1] + dir_offset
      const double rho = r * (M_PI / n_rho) + dir_offset;
      K.rho[r] = rho;
      K.s2r[r] = std::sin(2 * rho);
      K.c2r[r] = std::cos(2 * rho);
      const double e = 2 * (r * (M_PI / n_rho));
      K.es[r] = std::sin(e);
      K.ec[r] = std::cos(e);
   }
   for (int k = 0; k < 6; k++) { // linspace(a, b, 8)[1:-1]
      K.incl[k] = -M_PI / 2 + (k + 1) * (M_PI / 7);
      K.trel[k] = (k + 1) * (0.9 / 7);
   }
   K.inv_cos_tau = 1.0 / std::cos(tilt_angle);
   K.gain = gain;
   K.dir_offset = dir_offset;
   K.n_rho = (int32_t)n_rho;
   hipLaunchKernelGGL(k_rofl, dim3((unsigned)((n_pix + 127) / 128)), dim3(128), 0, st, K, in,
                      n_pix, mask, out);
   FPL_HIP(hipGetLastError());
   FPL_HIP(hipStreamSynchronize(st));
   return Status();
}

} // namespace rofl
} // namespace fpl

extern "C" int fpl_rofl(const double *in, int64_t n_pix, int64_t n_rho, double tilt_angle,
                        double gain, double dir_offset, const uint8_t *mask, double *out,
                        void *stream, char *msg, size_t msg_len) {
   return fpl::finish(fpl::rofl::run(in, n_pix, n_rho, tilt_angle, gain, dir_offset, mask, out,
                                     (hipStream_t)stream),
                      msg, msg_len);
}
