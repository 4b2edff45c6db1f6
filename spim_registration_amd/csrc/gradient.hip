// Helper kernels of the likelihood gradient (the differentiation extension of
// the deconvolution path): spectrum products, extraction of the PSF gradient
// from a correlation box, and the fold of the adjoint convolution over the
// mirror-single extension back onto the image.
#include "mvd_internal.h"

namespace mvd {

constexpr int kGThreads = 256;

__global__ __launch_bounds__(kGThreads) void k_spec_product(float2* out, const float2* a, const float2* b, int conj_a,
                                                           int conj_b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGThreads) {
    float2 x = a[i], y = b[i];
    if (conj_a) x.y = -x.y;
    if (conj_b) y.y = -y.y;
    out[i] = cmul(x, y);
  }
}

hipError_t launch_spec_product(float2* out, const float2* a, const float2* b, int conj_a, int conj_b, int64_t n,
                               hipStream_t stream) {
  const int64_t blocks = std::min<int64_t>((n + kGThreads - 1) / kGThreads, 16384);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_spec_product, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, out, a, b, conj_a, conj_b,
                     n);
  return hipGetLastError();
}

// grad[j] = scale * box[(c - j) mod q]: the correlation corr(R, W)(s) evaluated
// at s = c - j for every kernel voxel j.
__global__ __launch_bounds__(kGThreads) void k_extract_psf_grad(const float* box, Int3 q, Int3 kd, float scale,
                                                               float* grad) {
  const int64_t n = (int64_t)kd.x * kd.y * kd.z;
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGThreads) {
    const int x = (int)(i % kd.x), y = (int)((i / kd.x) % kd.y), z = (int)(i / ((int64_t)kd.x * kd.y));
    int sx = kd.x / 2 - x, sy = kd.y / 2 - y, sz = kd.z / 2 - z;
    if (sx < 0) sx += q.x;
    if (sy < 0) sy += q.y;
    if (sz < 0) sz += q.z;
    grad[i] = box[sx + (int64_t)q.x * (sy + (int64_t)q.y * sz)] * scale;
  }
}

hipError_t launch_extract_psf_grad(const float* box, Int3 q, Int3 kd, float scale, float* grad, hipStream_t stream) {
  const int64_t n = (int64_t)kd.x * kd.y * kd.z;
  const int64_t blocks = std::min<int64_t>((n + kGThreads - 1) / kGThreads, 4096);
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_extract_psf_grad, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, box, q, kd, scale, grad);
  return hipGetLastError();
}

// Window positions along one axis whose mirror-single image is p: with period
// P = 2n - 2 these are u = p + kP and u = -p + kP inside the window [-c, n + c)
// (the two families coincide for p = 0 and p = n - 1).  Requires n >= 2 and
// c <= 2 (n - 1), which the host checks.  Returns the window indices u + c.
constexpr int kMaxPre = 16;
__device__ __forceinline__ int preimages(int p, int n, int c, int* w) {
  const int P = 2 * n - 2;
  int k = 0;
  for (int kk = -3; kk <= 3; ++kk) {
    const int u1 = p + kk * P;
    if (u1 >= -c && u1 < n + c && k < kMaxPre) w[k++] = u1 + c;
    if (p != 0 && p != n - 1) {
      const int u2 = -p + kk * P;
      if (u2 >= -c && u2 < n + c && k < kMaxPre) w[k++] = u2 + c;
    }
  }
  return k;
}

__global__ __launch_bounds__(kGThreads) void k_fold_adjoint(const float* win, Int3 wd, Int3 n, Int3 c, double* acc) {
  const int64_t total = (int64_t)n.x * n.y * n.z;
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kGThreads) {
    const int x = (int)(i % n.x), y = (int)((i / n.x) % n.y), z = (int)(i / ((int64_t)n.x * n.y));
    int ux[kMaxPre], uy[kMaxPre], uz[kMaxPre];
    const int nx = preimages(x, n.x, c.x, ux), ny = preimages(y, n.y, c.y, uy), nz = preimages(z, n.z, c.z, uz);
    double s = 0.0;
    for (int a = 0; a < nz; ++a)
      for (int b = 0; b < ny; ++b)
        for (int d = 0; d < nx; ++d) s += (double)win[ux[d] + (int64_t)wd.x * (uy[b] + (int64_t)wd.y * uz[a])];
    acc[i] += s;
  }
}

hipError_t launch_fold_adjoint(const float* win, Int3 wdim, Int3 n, Int3 c, double* acc, hipStream_t stream) {
  const int64_t total = (int64_t)n.x * n.y * n.z;
  const int64_t blocks = std::min<int64_t>((total + kGThreads - 1) / kGThreads, 16384);
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_fold_adjoint, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, win, wdim, n, c, acc);
  return hipGetLastError();
}

__global__ __launch_bounds__(kGThreads) void k_to_float(const double* src, float* dst, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGThreads)
    dst[i] = (float)src[i];
}

hipError_t launch_to_float(const double* src, float* dst, int64_t n, hipStream_t stream) {
  const int64_t blocks = std::min<int64_t>((n + kGThreads - 1) / kGThreads, 16384);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_to_float, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, src, dst, n);
  return hipGetLastError();
}

__global__ __launch_bounds__(kGThreads) void k_zero_double(double* dst, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGThreads)
    dst[i] = 0.0;
}

hipError_t launch_zero_double(double* dst, int64_t n, hipStream_t stream) {
  const int64_t blocks = std::min<int64_t>((n + kGThreads - 1) / kGThreads, 16384);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_zero_double, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, dst, n);
  return hipGetLastError();
}

}  // namespace mvd
