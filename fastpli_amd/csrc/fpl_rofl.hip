// ROFL tilt analysis (Schmitz et al., Front. Neuroanat. 2018,
// doi:10.3389/fnana.2018.00075) of a 5-tilt x n_rho stack per pixel:
// fastpli's analysis.rofl.rofl (src/fastpli/analysis/rofl.py:17-60) with the
// local fit of src/fastpli/analysis/_ROFL_with_jacobi.py:403-522.
//
// One thread per pixel, FP64, in the reference's steps:
//   * centered intensities I / mean_tilt - 1 and weights
//     1 / sqrt(gain (I / mean^2 + I^2 / (mean^3 n_rho)))  (:47-57), the means
//     summed as numpy sums a row (eight partial sums from 8 values on);
//   * start: the EPA direction of the untilted view (rofl.py:55) and the first
//     minimum of the squared weighted residual over the 6 x 6 grid of
//     inclinations linspace(-pi/2, pi/2, 8)[1:-1] and t_rel
//     linspace(0, 0.9, 8)[1:-1] (:313-370);
//   * the residual w (model - centered) (:383-410): the fiber vector
//     v(phi, alpha) rotated by the four tilt matrices (:72-142), tilted angles
//     atan(w1 / w0), asin(w2), t_rel / cos(tau) (:164-193), symmetrised, |t|,
//     model sin(pi/2 |t| cos^2 alpha) sin(2 (rho - phi)) (:196-223);
//   * minimisation: MINPACK as scipy.optimize.leastsq runs it (fpl_lmdif.h,
//     maxfev 150): lmdif with forward differences, or in grad mode lmder with
//     the reference's Dfun, the weighted per-view Jacobian (:383-393,
//     :467-475), so the fitted angles and the evaluation count (itermap =
//     nfev) are the reference's;
//   * angles symmetrised as :60-69, |t_rel|, chi^2 of the final residual, and
//     the parameter deviations from the reference's per-tilt Jacobian
//     (:226-270) and its 3 x 3 inverse (:14-40), computed as the reference
//     computes them (its determinant's third term is m20 * m12, without m01).
#include <cmath>
#include <cstring>
#include <vector>

#include "fpl_internal.h"
#include "fpl_rofl_fit.h"

namespace fpl {
namespace rofl {

template <int MAXM, bool GRAD>
__global__ void __launch_bounds__(64) k_rofl(Const K, const double *__restrict__ in, int64_t n_pix,
                                             const uint8_t *__restrict__ mask,
                                             double *__restrict__ out) {
   const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (p >= n_pix) return;
   double o[8];
   if (mask && !mask[p]) {
      for (int k = 0; k < 8; k++) out[k * n_pix + p] = NAN;
      return;
   }
   const int n = K.n_rho;
   auto get = [&](int j, int r) { return in[((int64_t)j * n_pix + p) * n + r]; };
   fit_pixel<MAXM, GRAD>(K, get, o);
   for (int k = 0; k < 8; k++) out[k * n_pix + p] = o[k];
}

Status check(int64_t n_pix, int64_t n_rho, double gain) {
   if (n_rho <= 3) return make_status(FPL_E_INVALID, "data needs at least 3 equidistant rotations");
   if (n_rho > MAX_RHO)
      return make_status(FPL_E_INVALID, "at most %d rotations are supported", MAX_RHO);
   if (gain <= 1) return make_status(FPL_E_INVALID, "rofl gain <= 1");
   if (n_pix < 0) return make_status(FPL_E_INVALID, "bad sizes");
   return Status();
}

Status run(const double *in, int64_t n_pix, int64_t n_rho, double tilt_angle, double gain,
           double dir_offset, int grad_mode, const uint8_t *mask, double *out, hipStream_t st) {
   if (!in || !out) return make_status(FPL_E_INVALID, "null array");
   FPL_TRY(check(n_pix, n_rho, gain));
   if (n_pix == 0) return Status();
   Const K;
   make_const(n_rho, tilt_angle, gain, dir_offset, K);
   const dim3 grid((unsigned)((n_pix + 63) / 64));
   // residual arrays sized for the rotation count (per-thread scratch)
   auto kernel = n_rho <= 16 ? (grad_mode ? k_rofl<5 * 16, true> : k_rofl<5 * 16, false>)
                             : (grad_mode ? k_rofl<5 * MAX_RHO, true> : k_rofl<5 * MAX_RHO, false>);
   hipLaunchKernelGGL(kernel, grid, dim3(64), 0, st, K, in, n_pix, mask, out);
   FPL_HIP(hipGetLastError());
   FPL_HIP(hipStreamSynchronize(st));
   return Status();
}

} // namespace rofl
} // namespace fpl

extern "C" int fpl_rofl(const double *in, int64_t n_pix, int64_t n_rho, double tilt_angle,
                        double gain, double dir_offset, int grad_mode, const uint8_t *mask,
                        double *out, void *stream, char *msg, size_t msg_len) {
   return fpl::finish(fpl::rofl::run(in, n_pix, n_rho, tilt_angle, gain, dir_offset, grad_mode,
                                     mask, out, (hipStream_t)stream),
                      msg, msg_len);
}
