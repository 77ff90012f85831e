// fb_hostmath.hip -- host build of the library's hand-written device math,
// for CPU accuracy tests (tests/test_device_math.py).  Test infrastructure:
// the product library does not link it.
//   exp_nonpos               (fb_common.h) runs on every covariance element
//   exp_cr                   (fb_common.h) gives every lengthscale
//   goldschmidt_sqrt_rsqrt   (fb_tiles.h) refines every Cholesky pivot's
//                            square root and its reciprocal
#include "../../ferebus-v7_amd/csrc/fb_tiles.h"

extern "C" {

void fbh_exp_nonpos(int n, const double *x, double *y)
{
    double c[fb::kExpConsts];
    fb::exp_constants(c);
    for (int i = 0; i < n; ++i) y[i] = fb::exp_nonpos(x[i], c);
}

void fbh_sqrt_rsqrt(int n, const double *x, const double *y0, double *sq, double *rsq)
{
    for (int i = 0; i < n; ++i) fb::goldschmidt_sqrt_rsqrt(x[i], y0[i], sq[i], rsq[i]);
}

}  // extern "C"

extern "C" void fbh_exp_cr(int n, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) y[i] = fb::exp_cr(x[i]);
}
