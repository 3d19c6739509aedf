// Host build of the library's per-pixel ROFL fit (fastpli_amd/csrc/
// fpl_rofl_fit.h, the code the fpl_rofl kernel runs on the device), so the
// CPU tests can compare it with the oracle (scipy's leastsq) pixel by pixel.
// Test infrastructure: built by tests/test_rofl_fit_host.py.
#include "../../fastpli_amd/csrc/fpl_rofl_fit.h"

extern "C" int rofl_fit_host(const double *stack, int64_t n_rho, double tilt_angle, double gain,
                             double dir_offset, int grad_mode, double *out) {
   if (n_rho <= 3 || n_rho > fpl::rofl::MAX_RHO) return -1;
   fpl::rofl::Const K;
   fpl::rofl::make_const(n_rho, tilt_angle, gain, dir_offset, K);
   auto get = [&](int j, int r) { return stack[j * n_rho + r]; };
   if (grad_mode)
      fpl::rofl::fit_pixel<5 * fpl::rofl::MAX_RHO, true>(K, get, out);
   else
      fpl::rofl::fit_pixel<5 * fpl::rofl::MAX_RHO, false>(K, get, out);
   return 0;
}
