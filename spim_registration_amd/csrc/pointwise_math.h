// Per-voxel arithmetic shared by every gather and epilogue: the mirror-single
// out-of-bounds rule, the Richardson-Lucy quotient and the Tikhonov-regularised
// update, evaluated with the reference's float/double cast points and with FMA
// contraction off so that every kernel rounds the same way.
#pragma once

#include <hip/hip_runtime.h>

namespace mvd {

// Mirror-single extension of the block copy (Views.extendMirrorSingle,
// deconv/MVDeconFFTThreads.java:44,78): period 2n - 2, edges not repeated.
__device__ __forceinline__ int mirror_single(int p, int n) {
  if (p >= 0 && p < n) return p;
  if (n == 1) return 0;
  const int period = 2 * n - 2;
  p %= period;
  if (p < 0) p += period;
  return p < n ? p : period - p;
}

// MVDeconvolution.computeNextValue + tikhonov (deconv/MVDeconvolution.java:692-726).
// inv_lambda = 1 / lambda, computed once by the caller: the division by lambda
// becomes a multiplication (the double result may move by one ulp; the float
// it is rounded to practically never does).
__device__ __forceinline__ float rl_next_value(float last, float integral, float weight, double lambda,
                                               double inv_lambda) {
#pragma clang fp contract(off)
  const float minv = 0.0001f;
  const float value = last * integral;
  float adjusted;
  if (value > 0.f) {
    if (lambda > 0.0)
      adjusted = (float)((sqrt(1.0 + 2.0 * lambda * (double)value) - 1.0) * inv_lambda);
    else
      adjusted = value;
  } else {
    adjusted = minv;
  }
  const float next = isnan(adjusted) ? minv : fmaxf(minv, adjusted);
  return last + ((next - last) * weight);
}

__device__ __forceinline__ double rl_inv_lambda(double lambda) { return lambda > 0.0 ? 1.0 / lambda : 0.0; }

// Quotient of the observed view by the blurred estimate (MVDeconvolution.computeQuotient, deconv/MVDeconvolution.java:494-546).
__device__ __forceinline__ float rl_quotient(float blurred, float img) {
#pragma clang fp contract(off)
  return img > 0.f ? img / blurred : 1.f;
}

}  // namespace mvd
