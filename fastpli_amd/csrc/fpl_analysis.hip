// Image-side steps that follow the ray march in fastpli's pipeline
// (src/fastpli/simulation/simpli.py:354-389):
//   * optic resampling: Gaussian blur of every rotation image followed by a box
//     resampling to the CCD pixel grid (src/fastpli/simulation/optic.py:36-55,
//     85-120, 169-187; src/fastpli/simulation/__simpli.py:491-549);
//   * EPA, the per-pixel Fourier analysis of the rotation stack into
//     transmittance, direction and retardation (src/fastpli/analysis/epa.py:9-45).
// Both work on the float64 image stacks the march writes, in float64.
//
// The blur is the separable Gaussian of scipy.ndimage.gaussian_filter (the
// reference's call): kernel radius int(4 sigma + 1/2), weights
// exp(-x^2 / (2 sigma^2)) normalised to 1, axis 0 then axis 1, "reflect"
// boundary (d c b a | a b c d | d c b a).  Compiled without FMA contraction;
// the weights come from the C library's exp, so results agree with scipy to a
// few ulp rather than bit for bit.
#include <cmath>
#include <vector>

#include "fpl_internal.h"

namespace fpl {
namespace ana {

constexpr int MAX_RADIUS = 1024;

// index of a "reflect" boundary (half-sample symmetric), period 2n
__device__ __forceinline__ int64_t reflect(int64_t i, int64_t n) {
   const int64_t p = 2 * n;
   i %= p;
   if (i < 0) i += p;
   return i < n ? i : p - 1 - i;
}

// one 1-D pass along axis `ax` (0: x, 1: y) of an (nx, ny, nr) stack
__global__ void k_gauss1d(const double *__restrict__ in, double *__restrict__ out, int64_t nx,
                          int64_t ny, int64_t nr, int ax, const double *__restrict__ w,
                          int radius) {
   const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= nx * ny * nr) return;
   const int64_t r = e % nr, y = (e / nr) % ny, x = e / (nr * ny);
   const int64_t n = ax == 0 ? nx : ny, c = ax == 0 ? x : y;
   // scipy's symmetric correlate1d order: centre, then pairs from the outside in
   double acc = w[0] * in[e];
   for (int k = radius; k >= 1; k--) {
      const int64_t a = reflect(c - k, n), b = reflect(c + k, n);
      const int64_t ia = ax == 0 ? (a * ny + y) * nr + r : (x * ny + a) * nr + r;
      const int64_t ib = ax == 0 ? (b * ny + y) * nr + r : (x * ny + b) * nr + r;
      acc += w[k] * (in[ia] + in[ib]);
   }
   out[e] = acc;
}

// box resampling (optic.py:85-120): source pixel (i, j) goes to
// (floor((i + 1/2) s), floor((j + 1/2) s)); mean of the contributions
__global__ void k_resample(const double *__restrict__ in, double *__restrict__ out, int64_t nx,
                           int64_t ny, int64_t nr, int64_t mx, int64_t my, double scale) {
   const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= mx * my * nr) return;
   const int64_t r = e % nr, jj = (e / nr) % my, ii = e / (nr * my);
   // source rows whose target row is ii: floor((i + 0.5) * scale) == ii
   const int64_t i0 = (int64_t)ceil(ii / scale - 0.5) - 1, i1 = (int64_t)floor((ii + 1) / scale) + 1;
   const int64_t j0 = (int64_t)ceil(jj / scale - 0.5) - 1, j1 = (int64_t)floor((jj + 1) / scale) + 1;
   double acc = 0.0;
   int64_t cnt = 0;
   for (int64_t i = i0 < 0 ? 0 : i0; i <= i1 && i < nx; i++) {
      if ((int64_t)floor((i + 0.5) * scale) != ii) continue;
      for (int64_t j = j0 < 0 ? 0 : j0; j <= j1 && j < ny; j++) {
         if ((int64_t)floor((j + 0.5) * scale) != jj) continue;
         acc += in[(i * ny + j) * nr + r];
         cnt++;
      }
   }
   out[e] = acc / (double)cnt; // 0/0 -> nan as numpy's divide for empty pixels
}

// epa (epa.py:9-45) for float64 data
__global__ void k_epa(const double *__restrict__ in, int64_t n_pix, int nr,
                      const double *__restrict__ sc, double *__restrict__ t,
                      double *__restrict__ d, double *__restrict__ r) {
   const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (p >= n_pix) return;
   const double *v = in + p * nr;
   double s0 = 0.0, s1 = 0.0, s2 = 0.0;
   for (int k = 0; k < nr; k++) {
      s0 += v[k];
      s1 += v[k] * sc[2 * k];
      s2 += v[k] * sc[2 * k + 1];
   }
   double a0 = s0 / nr;
   const double a1 = 2 * s1 / nr, b1 = 2 * s2 / nr;
   t[p] = 2 * a0;
   const double dir = 0.5 * atan2(-b1, a1) + M_PI;
   if (a0 < 2.220446049250313e-16 * 100) a0 = 1;
   r[p] = sqrt(a1 * a1 + b1 * b1) / a0;
   d[p] = fmod(dir, M_PI); // dir > 0, so fmod equals numpy's % here
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

Status optic(const double *in, int64_t nx, int64_t ny, int64_t nr, double sigma, double scale,
             double *out, hipStream_t st) {
   if (!in || !out) return make_status(FPL_E_INVALID, "null array");
   if (nx <= 0 || ny <= 0 || nr <= 0) return make_status(FPL_E_INVALID, "empty image");
   if (!(scale > 0)) return make_status(FPL_E_INVALID, "scale <= 0");
   if (sigma < 0) return make_status(FPL_E_INVALID, "optical_sigma is < 0");
   const int64_t mx = (int64_t)std::nearbyint(nx * scale), my = (int64_t)std::nearbyint(ny * scale);
   if (mx <= 0 || my <= 0) return make_status(FPL_E_INVALID, "optical image size is 0");
   const int64_t n = nx * ny * nr;
   void *tmp;
   FPL_TRY(ws_get(st, WS_ANALYSIS, sizeof(double) * 2 * (size_t)n, &tmp));
   double *a = (double *)tmp, *b = a + n;
   const double *src = in;
   const int B = 256;
   if (sigma > 0) { // optic.filter: sigma == 0 leaves the image unchanged
      // scipy.ndimage._gaussian_kernel1d
      const int radius = (int)(4.0 * sigma + 0.5);
      if (radius > MAX_RADIUS) return make_status(FPL_E_INVALID, "optical_sigma too large");
      std::vector<double> w(radius + 1);
      double sum = 0.0;
      std::vector<double> full(2 * radius + 1);
      for (int k = -radius; k <= radius; k++) {
         full[k + radius] = std::exp(-0.5 / (sigma * sigma) * (double)(k * k));
         sum += full[k + radius];
      }
      for (int k = 0; k <= radius; k++) w[k] = full[k + radius] / sum;
      void *small, *stage;
      FPL_TRY(ws_get(st, WS_SMALL, sizeof(double) * w.size(), &small));
      FPL_TRY(host_stage(st, sizeof(double) * w.size(), &stage));
      FPL_HIP(hipStreamSynchronize(st));
      memcpy(stage, w.data(), sizeof(double) * w.size());
      FPL_HIP(hipMemcpyAsync(small, stage, sizeof(double) * w.size(), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_gauss1d, dim3(grid_for(n, B)), dim3(B), 0, st, in, a, nx, ny, nr, 0,
                         (const double *)small, radius);
      hipLaunchKernelGGL(k_gauss1d, dim3(grid_for(n, B)), dim3(B), 0, st, a, b, nx, ny, nr, 1,
                         (const double *)small, radius);
      src = b;
   }
   hipLaunchKernelGGL(k_resample, dim3(grid_for(mx * my * nr, B)), dim3(B), 0, st, src, out, nx,
                      ny, nr, mx, my, scale);
   FPL_HIP(hipGetLastError());
   FPL_HIP(hipStreamSynchronize(st));
   return Status();
}

Status epa(const double *in, int64_t n_pix, int64_t nr, double *t, double *d, double *r,
           hipStream_t st) {
   if (!in || !t || !d || !r) return make_status(FPL_E_INVALID, "null array");
   if (n_pix < 0 || nr <= 0 || nr > 4096) return make_status(FPL_E_INVALID, "bad sizes");
   if (n_pix == 0) return Status();
   // rho_2 = 2 * linspace(0, pi, n, endpoint=False) (epa.py:29)
   std::vector<double> sc(2 * nr);
   for (int64_t k = 0; k < nr; k++) {
      const double rho = k * (M_PI / nr);
      sc[2 * k] = std::sin(2 * rho);
      sc[2 * k + 1] = std::cos(2 * rho);
   }
   void *small, *stage;
   FPL_TRY(ws_get(st, WS_SMALL, sizeof(double) * sc.size(), &small));
   FPL_TRY(host_stage(st, sizeof(double) * sc.size(), &stage));
   FPL_HIP(hipStreamSynchronize(st));
   memcpy(stage, sc.data(), sizeof(double) * sc.size());
   FPL_HIP(hipMemcpyAsync(small, stage, sizeof(double) * sc.size(), hipMemcpyHostToDevice, st));
   hipLaunchKernelGGL(k_epa, dim3(grid_for(n_pix, 256)), dim3(256), 0, st, in, n_pix, (int)nr,
                      (const double *)small, t, d, r);
   FPL_HIP(hipGetLastError());
   FPL_HIP(hipStreamSynchronize(st));
   return Status();
}

} // namespace ana
} // namespace fpl

extern "C" int fpl_optic_resample(const double *in, int64_t nx, int64_t ny, int64_t n_rho,
                                  double sigma, double scale, double *out, void *stream,
                                  char *msg, size_t msg_len) {
   return fpl::finish(fpl::ana::optic(in, nx, ny, n_rho, sigma, scale, out, (hipStream_t)stream),
                      msg, msg_len);
}

extern "C" int fpl_epa(const double *in, int64_t n_pix, int64_t n_rho, double *transmittance,
                       double *direction, double *retardation, void *stream, char *msg,
                       size_t msg_len) {
   return fpl::finish(fpl::ana::epa(in, n_pix, n_rho, transmittance, direction, retardation,
                                    (hipStream_t)stream),
                      msg, msg_len);
}
