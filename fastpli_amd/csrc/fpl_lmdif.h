// Levenberg-Marquardt least squares: the algorithm of MINPACK's lmdif
// (forward-difference Jacobian) and lmder (the caller's Jacobian) (Moré,
// Garbow, Hillstrom, "User Guide for MINPACK-1", ANL-80-74, 1980: lmdif,
// lmder, fdjac2, qrfac, lmpar, qrsolv, enorm), which scipy.optimize.leastsq
// runs without / with Dfun.  The ROFL fit of the reference calls it so
// (src/fastpli/analysis/_ROFL_with_jacobi.py:460-475, maxfev 150, ftol = xtol
// = 1.49012e-8, gtol 0, factor 100, scaling by the column norms); its stopping
// point and evaluation count (itermap = nfev, function calls only in lmder)
// depend on the steps below, which are restated in MINPACK's order for
// N = 3 unknowns and M residuals.
//
// Usable on the device (one thread per fit) and on the host (a test build).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace fpl {
namespace lm {

constexpr int N = 3;

struct Options {
   double ftol = 1.49012e-08, xtol = 1.49012e-08, gtol = 0.0;
   double epsfcn = DBL_EPSILON; // scipy's default (finfo(float64).eps)
   double factor = 100.0;
   int maxfev = 150;
};

struct Result {
   int info = 0;
   int nfev = 0;
   double fnorm = 0.0; // Euclidean norm of the final residual
};

// enorm: the Euclidean norm with MINPACK's scaling for tiny and huge
// components (sums in index order)
template <typename Vec>
__host__ __device__ inline double enorm(int n, const Vec &x) {
   const double rdwarf = 3.834e-20, rgiant = 1.304e19;
   double s1 = 0, s2 = 0, s3 = 0, x1max = 0, x3max = 0;
   const double agiant = rgiant / (double)n;
   for (int i = 0; i < n; i++) {
      const double xabs = fabs(x[i]);
      if (xabs > rdwarf && xabs < agiant) {
         s2 += xabs * xabs;
      } else if (xabs <= rdwarf) {
         if (xabs > x3max) {
            const double r = x3max / xabs;
            s3 = 1.0 + s3 * r * r;
            x3max = xabs;
         } else if (xabs != 0.0) {
            const double r = xabs / x3max;
            s3 += r * r;
         }
      } else {
         if (xabs > x1max) {
            const double r = x1max / xabs;
            s1 = 1.0 + s1 * r * r;
            x1max = xabs;
         } else {
            const double r = xabs / x1max;
            s1 += r * r;
         }
      }
   }
   if (s1 != 0.0) return x1max * sqrt(s1 + (s2 / x1max) / x1max);
   if (s2 != 0.0) {
      if (s2 >= x3max) return sqrt(s2 * (1.0 + (x3max / s2) * (x3max * s3)));
      return sqrt(x3max * ((s2 / x3max) + (x3max * s3)));
   }
   return x3max * sqrt(s3);
}

// a column of an M x N column-major matrix from row `from` on
struct Col {
   const double *p;
   __host__ __device__ double operator[](int i) const { return p[i]; }
};

// qrfac with column pivoting: a (m x N, column-major, leading dimension lda)
// becomes R above the diagonal and the Householder vectors from the
// diagonal down; rdiag = diag(R), acnorm = the input column norms
__host__ __device__ inline void qrfac(int m, double *a, int lda, int ipvt[N], double rdiag[N],
                                      double acnorm[N]) {
   const double epsmch = DBL_EPSILON;
   double wa[N];
   for (int j = 0; j < N; j++) {
      acnorm[j] = enorm(m, Col{a + j * lda});
      rdiag[j] = acnorm[j];
      wa[j] = rdiag[j];
      ipvt[j] = j;
   }
   const int minmn = m < N ? m : N;
   for (int j = 0; j < minmn; j++) {
      int kmax = j;
      for (int k = j; k < N; k++)
         if (rdiag[k] > rdiag[kmax]) kmax = k;
      if (kmax != j) {
         for (int i = 0; i < m; i++) {
            const double t = a[i + j * lda];
            a[i + j * lda] = a[i + kmax * lda];
            a[i + kmax * lda] = t;
         }
         rdiag[kmax] = rdiag[j];
         wa[kmax] = wa[j];
         const int k = ipvt[j];
         ipvt[j] = ipvt[kmax];
         ipvt[kmax] = k;
      }
      double ajnorm = enorm(m - j, Col{a + j + j * lda});
      if (ajnorm != 0.0) {
         if (a[j + j * lda] < 0.0) ajnorm = -ajnorm;
         for (int i = j; i < m; i++) a[i + j * lda] /= ajnorm;
         a[j + j * lda] += 1.0;
         for (int k = j + 1; k < N; k++) {
            double sum = 0.0;
            for (int i = j; i < m; i++) sum += a[i + j * lda] * a[i + k * lda];
            const double temp = sum / a[j + j * lda];
            for (int i = j; i < m; i++) a[i + k * lda] -= temp * a[i + j * lda];
            if (rdiag[k] != 0.0) {
               const double t = a[j + k * lda] / rdiag[k];
               const double r = 1.0 - t * t;
               rdiag[k] *= sqrt(r > 0.0 ? r : 0.0);
               const double q = rdiag[k] / wa[k];
               if (0.05 * (q * q) <= epsmch) {
                  rdiag[k] = enorm(m - j - 1, Col{a + (j + 1) + k * lda});
                  wa[k] = rdiag[k];
               }
            }
         }
      }
      rdiag[j] = -ajnorm;
   }
}

// qrsolv: solves (R^T R + D^2) x = R^T qtb in the least squares sense with
// Givens rotations; r (N x N, column-major, ldr) keeps its upper triangle,
// its strict lower triangle receives the triangular factor S^T
__host__ __device__ inline void qrsolv(double *r, int ldr, const int ipvt[N],
                                       const double diag[N], const double qtb[N], double x[N],
                                       double sdiag[N]) {
   double wa[N];
   for (int j = 0; j < N; j++) {
      for (int i = j; i < N; i++) r[i + j * ldr] = r[j + i * ldr];
      x[j] = r[j + j * ldr];
      wa[j] = qtb[j];
   }
   for (int j = 0; j < N; j++) {
      const int l = ipvt[j];
      if (diag[l] != 0.0) {
         for (int k = j; k < N; k++) sdiag[k] = 0.0;
         sdiag[j] = diag[l];
         double qtbpj = 0.0;
         for (int k = j; k < N; k++) {
            if (sdiag[k] == 0.0) continue;
            double sn, cs;
            if (fabs(r[k + k * ldr]) < fabs(sdiag[k])) {
               const double cotan = r[k + k * ldr] / sdiag[k];
               sn = 0.5 / sqrt(0.25 + 0.25 * (cotan * cotan));
               cs = sn * cotan;
            } else {
               const double tn = sdiag[k] / r[k + k * ldr];
               cs = 0.5 / sqrt(0.25 + 0.25 * (tn * tn));
               sn = cs * tn;
            }
            r[k + k * ldr] = cs * r[k + k * ldr] + sn * sdiag[k];
            const double temp = cs * wa[k] + sn * qtbpj;
            qtbpj = -sn * wa[k] + cs * qtbpj;
            wa[k] = temp;
            for (int i = k + 1; i < N; i++) {
               const double t2 = cs * r[i + k * ldr] + sn * sdiag[i];
               sdiag[i] = -sn * r[i + k * ldr] + cs * sdiag[i];
               r[i + k * ldr] = t2;
            }
         }
      }
      sdiag[j] = r[j + j * ldr];
      r[j + j * ldr] = x[j];
   }
   int nsing = N;
   for (int j = 0; j < N; j++) {
      if (sdiag[j] == 0.0 && nsing == N) nsing = j;
      if (nsing < N) wa[j] = 0.0;
   }
   for (int k = 1; k <= nsing; k++) {
      const int j = nsing - k;
      double sum = 0.0;
      for (int i = j + 1; i < nsing; i++) sum += r[i + j * ldr] * wa[i];
      wa[j] = (wa[j] - sum) / sdiag[j];
   }
   for (int j = 0; j < N; j++) x[ipvt[j]] = wa[j];
}

// lmpar: the Levenberg-Marquardt parameter par for the trust radius delta
// and the step x; r as qrfac left it (upper triangle of R)
__host__ __device__ inline void lmpar(double *r, int ldr, const int ipvt[N],
                                      const double diag[N], const double qtb[N], double delta,
                                      double &par, double x[N], double sdiag[N]) {
   const double dwarf = DBL_MIN, p1 = 0.1, p001 = 0.001;
   double wa1[N], wa2[N];
   int nsing = N;
   for (int j = 0; j < N; j++) {
      wa1[j] = qtb[j];
      if (r[j + j * ldr] == 0.0 && nsing == N) nsing = j;
      if (nsing < N) wa1[j] = 0.0;
   }
   for (int k = 1; k <= nsing; k++) {
      const int j = nsing - k;
      wa1[j] /= r[j + j * ldr];
      const double temp = wa1[j];
      for (int i = 0; i < j; i++) wa1[i] -= r[i + j * ldr] * temp;
   }
   for (int j = 0; j < N; j++) x[ipvt[j]] = wa1[j];
   int iter = 0;
   for (int j = 0; j < N; j++) wa2[j] = diag[j] * x[j];
   double dxnorm = enorm(N, wa2);
   double fp = dxnorm - delta;
   if (fp <= p1 * delta) {
      par = 0.0;
      return;
   }
   double parl = 0.0;
   if (nsing >= N) {
      for (int j = 0; j < N; j++) {
         const int l = ipvt[j];
         wa1[j] = diag[l] * (wa2[l] / dxnorm);
      }
      for (int j = 0; j < N; j++) {
         double sum = 0.0;
         for (int i = 0; i < j; i++) sum += r[i + j * ldr] * wa1[i];
         wa1[j] = (wa1[j] - sum) / r[j + j * ldr];
      }
      const double temp = enorm(N, wa1);
      parl = ((fp / delta) / temp) / temp;
   }
   for (int j = 0; j < N; j++) {
      double sum = 0.0;
      for (int i = 0; i <= j; i++) sum += r[i + j * ldr] * qtb[i];
      wa1[j] = sum / diag[ipvt[j]];
   }
   const double gnorm = enorm(N, wa1);
   double paru = gnorm / delta;
   if (paru == 0.0) paru = dwarf / (delta < p1 ? delta : p1);
   par = par > parl ? par : parl;
   par = par < paru ? par : paru;
   if (par == 0.0) par = gnorm / dxnorm;
   for (;;) {
      iter++;
      if (par == 0.0) par = dwarf > p001 * paru ? dwarf : p001 * paru;
      const double temp = sqrt(par);
      for (int j = 0; j < N; j++) wa1[j] = temp * diag[j];
      qrsolv(r, ldr, ipvt, wa1, qtb, x, sdiag);
      for (int j = 0; j < N; j++) wa2[j] = diag[j] * x[j];
      dxnorm = enorm(N, wa2);
      const double fp_old = fp;
      fp = dxnorm - delta;
      if (fabs(fp) <= p1 * delta || (parl == 0.0 && fp <= fp_old && fp_old < 0.0) || iter == 10)
         break;
      for (int j = 0; j < N; j++) {
         const int l = ipvt[j];
         wa1[j] = diag[l] * (wa2[l] / dxnorm);
      }
      for (int j = 0; j < N; j++) {
         wa1[j] /= sdiag[j];
         const double t = wa1[j];
         for (int i = j + 1; i < N; i++) wa1[i] -= r[i + j * ldr] * t;
      }
      const double tn = enorm(N, wa1);
      const double parc = ((fp / delta) / tn) / tn;
      if (fp > 0.0) parl = parl > par ? parl : par;
      if (fp < 0.0) paru = paru < par ? paru : par;
      par = parl > par + parc ? parl : par + parc;
   }
}

// lmdif (ANALYTIC false) / lmder (true): minimises the sum of squares of the
// m residuals fcn(x, fvec) over x[N] from the start x; lmder takes the
// Jacobian from fcn.jacobian(x, fjac).  Work arrays: fvec, wa4 (m each) and
// fjac (m x N, column-major); fvec holds the residuals at the returned x.
template <bool ANALYTIC, typename Fcn>
__host__ __device__ inline Result minimize(Fcn &fcn, int m, double x[N], double *fvec,
                                           double *fjac, double *wa4, const Options &o) {
   const double p1 = 0.1, p5 = 0.5, p25 = 0.25, p75 = 0.75, p0001 = 1.0e-4;
   const double epsmch = DBL_EPSILON;
   Result res;
   if (m < N || o.ftol < 0 || o.xtol < 0 || o.gtol < 0 || o.maxfev <= 0 || o.factor <= 0)
      return res;
   double diag[N], qtf[N], wa1[N], wa2[N], wa3[N];
   int ipvt[N];
   fcn(x, fvec);
   res.nfev = 1;
   double fnorm = enorm(m, Col{fvec});
   double par = 0.0, delta = 0.0, xnorm = 0.0;
   int iter = 1;
   for (;;) {
      if (ANALYTIC) {
         fcn.jacobian(x, fjac);
      } else {
         // fdjac2: forward differences
         const double eps = sqrt(o.epsfcn > epsmch ? o.epsfcn : epsmch);
         for (int j = 0; j < N; j++) {
            const double temp = x[j];
            double h = eps * fabs(temp);
            if (h == 0.0) h = eps;
            x[j] = temp + h;
            fcn(x, wa4);
            x[j] = temp;
            for (int i = 0; i < m; i++) fjac[i + j * m] = (wa4[i] - fvec[i]) / h;
         }
         res.nfev += N;
      }
      qrfac(m, fjac, m, ipvt, wa1, wa2);
      if (iter == 1) {
         for (int j = 0; j < N; j++) diag[j] = wa2[j] == 0.0 ? 1.0 : wa2[j];
         for (int j = 0; j < N; j++) wa3[j] = diag[j] * x[j];
         xnorm = enorm(N, wa3);
         delta = o.factor * xnorm;
         if (delta == 0.0) delta = o.factor;
      }
      // Q^T fvec, the first N components into qtf
      for (int i = 0; i < m; i++) wa4[i] = fvec[i];
      for (int j = 0; j < N; j++) {
         if (fjac[j + j * m] != 0.0) {
            double sum = 0.0;
            for (int i = j; i < m; i++) sum += fjac[i + j * m] * wa4[i];
            const double temp = -sum / fjac[j + j * m];
            for (int i = j; i < m; i++) wa4[i] += fjac[i + j * m] * temp;
         }
         fjac[j + j * m] = wa1[j];
         qtf[j] = wa4[j];
      }
      // the norm of the scaled gradient
      double gnorm = 0.0;
      if (fnorm != 0.0) {
         for (int j = 0; j < N; j++) {
            const int l = ipvt[j];
            if (wa2[l] != 0.0) {
               double sum = 0.0;
               for (int i = 0; i <= j; i++) sum += fjac[i + j * m] * (qtf[i] / fnorm);
               const double g = fabs(sum / wa2[l]);
               gnorm = gnorm > g ? gnorm : g;
            }
         }
      }
      if (gnorm <= o.gtol) res.info = 4;
      if (res.info != 0) break;
      for (int j = 0; j < N; j++) diag[j] = diag[j] > wa2[j] ? diag[j] : wa2[j];
      double ratio = 0.0;
      do {
         lmpar(fjac, m, ipvt, diag, qtf, delta, par, wa1, wa2);
         for (int j = 0; j < N; j++) {
            wa1[j] = -wa1[j];
            wa2[j] = x[j] + wa1[j];
            wa3[j] = diag[j] * wa1[j];
         }
         const double pnorm = enorm(N, wa3);
         if (iter == 1) delta = delta < pnorm ? delta : pnorm;
         fcn(wa2, wa4);
         res.nfev++;
         const double fnorm1 = enorm(m, Col{wa4});
         double actred = -1.0;
         if (p1 * fnorm1 < fnorm) {
            const double q = fnorm1 / fnorm;
            actred = 1.0 - q * q;
         }
         // the predicted reduction and the directional derivative
         for (int j = 0; j < N; j++) {
            wa3[j] = 0.0;
            const double temp = wa1[ipvt[j]];
            for (int i = 0; i <= j; i++) wa3[i] += fjac[i + j * m] * temp;
         }
         const double temp1 = enorm(N, wa3) / fnorm;
         const double temp2 = (sqrt(par) * pnorm) / fnorm;
         const double prered = temp1 * temp1 + temp2 * temp2 / p5;
         const double dirder = -(temp1 * temp1 + temp2 * temp2);
         ratio = 0.0;
         if (prered != 0.0) ratio = actred / prered;
         // update the step bound
         if (ratio <= p25) {
            double temp = actred >= 0.0 ? p5 : p5 * dirder / (dirder + p5 * actred);
            if (p1 * fnorm1 >= fnorm || temp < p1) temp = p1;
            const double lim = pnorm / p1;
            delta = temp * (delta < lim ? delta : lim);
            par = par / temp;
         } else if (par == 0.0 || ratio >= p75) {
            delta = pnorm / p5;
            par = p5 * par;
         }
         if (ratio >= p0001) { // a successful step
            for (int j = 0; j < N; j++) {
               x[j] = wa2[j];
               wa2[j] = diag[j] * x[j];
            }
            for (int i = 0; i < m; i++) fvec[i] = wa4[i];
            xnorm = enorm(N, wa2);
            fnorm = fnorm1;
            iter++;
         }
         // convergence
         if (fabs(actred) <= o.ftol && prered <= o.ftol && p5 * ratio <= 1.0) res.info = 1;
         if (delta <= o.xtol * xnorm) res.info = 2;
         if (fabs(actred) <= o.ftol && prered <= o.ftol && p5 * ratio <= 1.0 && res.info == 2)
            res.info = 3;
         if (res.info != 0) break;
         // termination
         if (res.nfev >= o.maxfev) res.info = 5;
         if (fabs(actred) <= epsmch && prered <= epsmch && p5 * ratio <= 1.0) res.info = 6;
         if (delta <= epsmch * xnorm) res.info = 7;
         if (gnorm <= epsmch) res.info = 8;
         if (res.info != 0) break;
      } while (ratio < p0001);
      if (res.info != 0) break;
   }
   res.fnorm = fnorm;
   return res;
}

} // namespace lm
} // namespace fpl
