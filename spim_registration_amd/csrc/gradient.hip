// Helper kernels of the likelihood gradient (the differentiation extension of
// the deconvolution path): spectrum products, extraction of the PSF gradient
// from a correlation box, and the fold of the adjoint convolution over the
// mirror-single extension back onto the image.
#include "mvd_internal.h"
#include "pointwise_math.h"

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

// out (+)= conj(a) b: the PSF correlation spectra of the blocks summed in the
// spectral domain (the correlation is linear and every block's window frame
// gives the same lags), written on the first block, added afterwards.
__global__ __launch_bounds__(kGThreads) void k_spec_corr_acc(float2* out, const float2* a, const float2* b, int first,
                                                            int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGThreads) {
    float2 x = a[i];
    x.y = -x.y;
    const float2 p = cmul(x, b[i]);
    out[i] = first ? p : cadd(out[i], p);
  }
}

hipError_t launch_spec_corr_acc(float2* out, const float2* a, const float2* b, int first, int64_t n,
                                hipStream_t stream) {
  const int64_t blocks = std::min<int64_t>((n + kGThreads - 1) / kGThreads, 16384);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_spec_corr_acc, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, out, a, b, first, n);
  return hipGetLastError();
}

// grad[j] (+)= scale * box[(c - j) mod q]: the correlation corr(R, W)(s)
// evaluated at s = c - j for every kernel voxel j; accumulate adds the block's
// share to the view's gradient.
__global__ __launch_bounds__(kGThreads) void k_extract_psf_grad(const float* box, Int3 q, Int3 kd, float scale,
                                                               float* grad, int accumulate) {
  const int64_t n = (int64_t)kd.x * kd.y * kd.z;
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGThreads) {
    const int x = (int)(i % kd.x), y = (int)((i / kd.x) % kd.y), z = (int)(i / ((int64_t)kd.x * kd.y));
    int sx = kd.x / 2 - x, sy = kd.y / 2 - y, sz = kd.z / 2 - z;
    if (sx < 0) sx += q.x;
    if (sy < 0) sy += q.y;
    if (sz < 0) sz += q.z;
    const float v = box[sx + (int64_t)q.x * (sy + (int64_t)q.y * sz)] * scale;
    grad[i] = accumulate ? grad[i] + v : v;
  }
}

hipError_t launch_extract_psf_grad(const float* box, Int3 q, Int3 kd, float scale, float* grad, int accumulate,
                                   hipStream_t stream) {
  const int64_t n = (int64_t)kd.x * kd.y * kd.z;
  const int64_t blocks = std::min<int64_t>((n + kGThreads - 1) / kGThreads, 4096);
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_extract_psf_grad, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, box, q, kd, scale, grad,
                     accumulate);
  return hipGetLastError();
}

// dst (dense, dims size) = src[origin + (x, y, z)] of an image with dims n: the
// residual of one block's effective region.
__global__ __launch_bounds__(kGThreads) void k_crop(float* dst, const float* src, Int3 n, Int3 origin, Int3 size) {
  const int64_t total = (int64_t)size.x * size.y * size.z;
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kGThreads) {
    const int x = (int)(i % size.x), y = (int)((i / size.x) % size.y), z = (int)(i / ((int64_t)size.x * size.y));
    dst[i] = src[(int64_t)(origin.x + x) + (int64_t)n.x * ((origin.y + y) + (int64_t)n.y * (origin.z + z))];
  }
}

hipError_t launch_crop(float* dst, const float* src, Int3 n, Int3 origin, Int3 size, hipStream_t stream) {
  const int64_t total = (int64_t)size.x * size.y * size.z;
  const int64_t blocks = std::min<int64_t>((total + kGThreads - 1) / kGThreads, 16384);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_crop, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, dst, src, n, origin, size);
  return hipGetLastError();
}

// Adjoint of the mirror-single window gather: every window position u in
// [lo, lo + wd) adds its value to the image voxel mirror(u).  Positions of one
// window can share a voxel near the borders, hence the FP64 atomic adds
// (accumulation order, and with it the last bits, is not fixed).
__global__ __launch_bounds__(kGThreads) void k_fold_window(const float* win, Int3 wd, Int3 lo, Int3 n, double* acc) {
  const int64_t total = (int64_t)wd.x * wd.y * wd.z;
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kGThreads) {
    const int x = (int)(i % wd.x), y = (int)((i / wd.x) % wd.y), z = (int)(i / ((int64_t)wd.x * wd.y));
    const int px = mirror_single(lo.x + x, n.x), py = mirror_single(lo.y + y, n.y), pz = mirror_single(lo.z + z, n.z);
    unsafeAtomicAdd(acc + (px + (int64_t)n.x * (py + (int64_t)n.y * pz)), (double)win[i]);
  }
}

hipError_t launch_fold_window(const float* win, Int3 wd, Int3 lo, Int3 n, double* acc, hipStream_t stream) {
  const int64_t total = (int64_t)wd.x * wd.y * wd.z;
  const int64_t blocks = std::min<int64_t>((total + kGThreads - 1) / kGThreads, 16384);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_fold_window, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, win, wd, lo, n, acc);
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
