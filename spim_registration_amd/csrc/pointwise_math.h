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

// t / lambda in double, correctly rounded, from inv = 1 / lambda (correctly
// rounded, computed once): q = t inv is within one ulp of the quotient, the
// remainder t - q lambda is exact with one fused multiply-add, and q + rem inv
// rounds to the correctly rounded quotient (Markstein's theorem; checked equal
// to the division on 5e8 Tikhonov numerators over ten values of lambda).  Three
// FP64 operations instead of the hardware division sequence.
__device__ __forceinline__ double div_by_lambda(double t, double lambda, double inv) {
  const double q = t * inv;
  const double rem = fma(-q, lambda, t);
  return fma(rem, inv, q);
}

__device__ __forceinline__ double rl_inv_lambda(double lambda) { return lambda > 0.0 ? 1.0 / lambda : 0.0; }

// MVDeconvolution.computeNextValue + tikhonov (deconv/MVDeconvolution.java:692-726),
// including the division by lambda in double; inv_lambda = rl_inv_lambda(lambda).
__device__ __forceinline__ float rl_next_value(float last, float integral, float weight, double lambda,
                                               double inv_lambda) {
#pragma clang fp contract(off)
  const float minv = 0.0001f;
  const float value = last * integral;
  float adjusted;
  if (value > 0.f) {
    if (lambda > 0.0)
      adjusted = (float)div_by_lambda(sqrt(1.0 + 2.0 * lambda * (double)value) - 1.0, lambda, inv_lambda);
    else
      adjusted = value;
  } else {
    adjusted = minv;
  }
  const float next = isnan(adjusted) ? minv : fmaxf(minv, adjusted);
  return last + ((next - last) * weight);
}

// Quotient of the observed view by the blurred estimate (MVDeconvolution.computeQuotient, deconv/MVDeconvolution.java:494-546).
__device__ __forceinline__ float rl_quotient(float blurred, float img) {
#pragma clang fp contract(off)
  return img > 0.f ? img / blurred : 1.f;
}

// Poisson negative log-likelihood term w (b - y ln b) of one voxel with b > 0,
// for a sum kept in FP64.  The log is taken in FP32: its rounding, about one
// ulp of y ln b per voxel and of either sign, stays near 1e-9 of the sum over a
// volume (the FP64 oracle's NLL agrees to 1e-6), while an FP64 log is some
// sixty FP64 instructions per voxel, which made the residual x transform
// VALU-bound.
__device__ __forceinline__ double nll_term(float w, float b, float y) {
  return (double)w * ((double)b - (double)y * (double)logf(b));
}

// Residual w (1 - y / b) of the likelihood gradient for b > 0, in FP32.
__device__ __forceinline__ float nll_residual(float w, float b, float y) {
#pragma clang fp contract(off)
  return w * (1.f - y / b);
}

// java.lang.Math.max / min on floats: NaN if either argument is NaN (fmaxf and
// fminf return the other argument instead).  The per-view change statistics
// (MVDeconvolution.java:647,677,454) and the OSEM clamp
// (ProcessForDeconvolution.java:382) keep NaN this way.
__device__ __forceinline__ float java_maxf(float a, float b) {
  return a != a ? a : b != b ? b : fmaxf(a, b);
}
__device__ __forceinline__ float java_minf(float a, float b) {
  return a != a ? a : b != b ? b : fminf(a, b);
}
// java.lang.Math.min on doubles (NormalizingRandomAccess.get, weights/NormalizingRandomAccess.java:62).
__device__ __forceinline__ double java_min(double a, double b) {
  return a != a ? a : b != b ? b : fmin(a, b);
}

}  // namespace mvd
