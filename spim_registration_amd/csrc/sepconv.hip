// The reference's second native hook, CUDASeparableConvolution
// (spim/process/cuda/CUDASeparableConvolution.java:30-43), used by the
// difference-of-Gaussian interest points (DifferenceOfGaussianCUDA.java:163-191
// through CUDASeparableConvolutionFunctions.gauss, :123-217): in-place
// separable convolution of a w x h x d float image (x fastest) with one kernel
// of odd size N in {7, 15, 31, 63, 127} per axis, centred at N/2, each axis
// optional; out of bounds 0 = zero, 1 = a given value, 2 = the nearest border
// pixel.  out(x) = sum_j in(x - j) k(N/2 + j), along x, then y, then z.
//
// One pass per axis through HBM (read + write 8 bytes per voxel), the line
// segment and its halo staged in LDS, the kernel in LDS, FP32 fused
// multiply-adds in kernel order.  The x pass tiles 256 outputs of one row; the
// y and z passes tile 64 consecutive x by 64 outputs along the axis (every
// global access of a wavefront is 256 contiguous bytes).  Large kernels make
// the passes VALU-bound (2 N flops per voxel and axis against 8 bytes).
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mvdecon.h"
#include "mvd_internal.h"

namespace mvd {

namespace {

constexpr int kSepThreads = 256;
constexpr int kXTile = 256;               // x pass: outputs per workgroup
constexpr int kCols = 64, kAxisTile = 64;  // y / z pass: 64 columns x 64 outputs

__device__ __forceinline__ float oob_value(const float* __restrict__ in, int64_t base, int64_t stride, int i, int n,
                                           int oob, float value) {
  if (i >= 0 && i < n) return gmem(in)[base + (int64_t)i * stride];
  if (oob == 2) return gmem(in)[base + (int64_t)(i < 0 ? 0 : n - 1) * stride];
  return oob == 1 ? value : 0.f;
}

// Grids are one-dimensional (workgroup indices decoded below): the y and z
// grid dimensions are too small for large images.
template <int R>
__global__ __launch_bounds__(kSepThreads) void k_sep_x(const float* __restrict__ in, float* __restrict__ out,
                                                       const float* __restrict__ kern, int w, int chunks, int oob,
                                                       float value) {
  __shared__ float s[kXTile + 2 * R];
  __shared__ float k[2 * R + 1];
  const int64_t row = blockIdx.x / chunks;  // y + h * z
  const int x0 = (int)(blockIdx.x - row * chunks) * kXTile;
  const int64_t base = row * w;
  for (int i = threadIdx.x; i < 2 * R + 1; i += kSepThreads) k[i] = gmem(kern)[i];
  for (int i = threadIdx.x; i < kXTile + 2 * R; i += kSepThreads) s[i] = oob_value(in, base, 1, x0 + i - R, w, oob, value);
  __syncthreads();
  const int x = x0 + threadIdx.x;
  if (x >= w) return;
  // out(x) = sum_j in(x - j) k(R + j): in(x - j) is s[threadIdx.x + R - j]
  float acc = 0.f;
#pragma unroll 8
  for (int j = -R; j <= R; ++j) acc = fmaf(s[threadIdx.x + R - j], k[R + j], acc);
  gmem(out)[base + x] = acc;
}

// y or z pass: axis length n, axis stride (in elements) astride; the other
// coordinates: x in [0, w) and `outer` rows of stride ostride.
template <int R>
__global__ __launch_bounds__(kSepThreads) void k_sep_axis(const float* __restrict__ in, float* __restrict__ out,
                                                          const float* __restrict__ kern, int w, int n,
                                                          int64_t astride, int64_t ostride, int gx, int ga, int oob,
                                                          float value) {
  __shared__ float s[(kAxisTile + 2 * R) * kCols];
  __shared__ float k[2 * R + 1];
  const int cx = threadIdx.x % kCols, cr = threadIdx.x / kCols;  // column, row group (0..3)
  const int64_t outer = blockIdx.x / ((int64_t)gx * ga);
  const int rem = (int)(blockIdx.x - outer * gx * ga);
  const int x = (rem % gx) * kCols + cx;
  const int a0 = (rem / gx) * kAxisTile;
  const int64_t base = outer * ostride + (x < w ? x : w - 1);
  for (int i = threadIdx.x; i < 2 * R + 1; i += kSepThreads) k[i] = gmem(kern)[i];
  for (int i = cr; i < kAxisTile + 2 * R; i += kSepThreads / kCols)
    s[i * kCols + cx] = oob_value(in, base, astride, a0 + i - R, n, oob, value);
  __syncthreads();
  if (x >= w) return;
#pragma unroll 1
  for (int o = cr; o < kAxisTile; o += kSepThreads / kCols) {
    const int a = a0 + o;
    if (a >= n) break;
    float acc = 0.f;
#pragma unroll 8
    for (int j = -R; j <= R; ++j) acc = fmaf(s[(o + R - j) * kCols + cx], k[R + j], acc);
    gmem(out)[base + (int64_t)a * astride] = acc;
  }
}

template <int R>
hipError_t run_passes(float* a, float* b, const float* kx, const float* ky, const float* kz, int w, int h, int d,
                      int oob, float value, float** result) {
  float* cur = a;
  float* nxt = b;
  const dim3 block(kSepThreads);
  const int chunks = (w + kXTile - 1) / kXTile, gx = (w + kCols - 1) / kCols;
  if (kx) {
    hipLaunchKernelGGL(k_sep_x<R>, dim3((unsigned)((int64_t)chunks * h * d)), block, 0, nullptr, cur, nxt, kx, w,
                       chunks, oob, value);
    std::swap(cur, nxt);
  }
  if (ky) {
    const int ga = (h + kAxisTile - 1) / kAxisTile;
    hipLaunchKernelGGL(k_sep_axis<R>, dim3((unsigned)((int64_t)gx * ga * d)), block, 0, nullptr, cur, nxt, ky, w, h,
                       (int64_t)w, (int64_t)w * h, gx, ga, oob, value);
    std::swap(cur, nxt);
  }
  if (kz) {
    const int ga = (d + kAxisTile - 1) / kAxisTile;
    hipLaunchKernelGGL(k_sep_axis<R>, dim3((unsigned)((int64_t)gx * ga * h)), block, 0, nullptr, cur, nxt, kz, w, d,
                       (int64_t)w * h, (int64_t)w, gx, ga, oob, value);
    std::swap(cur, nxt);
  }
  *result = cur;
  return hipGetLastError();
}

// Host side of every convolve_N entry: no error channel (Java boolean), so
// failures are printed and the image is left unchanged.
int sep_convolve(int size, float* image, const float* kx, const float* ky, const float* kz, int w, int h, int d,
                 int cx, int cy, int cz, int oob, float value, int dev) {
  const char* what = nullptr;
  if (!image || w <= 0 || h <= 0 || d <= 0) what = "invalid image";
  else if (oob < 0 || oob > 2) what = "invalid out-of-bounds strategy";
  else if ((cx && !kx) || (cy && !ky) || (cz && !kz)) what = "missing kernel";
  else if ((int64_t)((w + kCols - 1) / kCols) * ((std::max(h, d) + kAxisTile - 1) / kAxisTile) * std::max(h, d) >
               2147483647 ||
           (int64_t)((w + kXTile - 1) / kXTile) * h * d > 2147483647)
    what = "image too large for one launch";
  int ndev = 0;
  if (!what && (hipGetDeviceCount(&ndev) != hipSuccess || dev < 0 || dev >= ndev)) what = "invalid device";
  if (what) {
    std::fprintf(stderr, "convolve_%d: %s\n", size, what);
    set_error(what);
    return 0;
  }
  if (!cx && !cy && !cz) return 1;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(dev);
  const int64_t n = (int64_t)w * h * d;
  float *a = nullptr, *b = nullptr, *k = nullptr, *res = nullptr;
  hipError_t e = hipMalloc(&a, sizeof(float) * n);
  if (e == hipSuccess) e = hipMalloc(&b, sizeof(float) * n);
  if (e == hipSuccess) e = hipMalloc(&k, sizeof(float) * 3 * size);
  if (e == hipSuccess) e = staged_h2d(a, image, sizeof(float) * n, nullptr);  // Java array: pageable
  if (e == hipSuccess && cx) e = hipMemcpy(k, kx, sizeof(float) * size, hipMemcpyHostToDevice);
  if (e == hipSuccess && cy) e = hipMemcpy(k + size, ky, sizeof(float) * size, hipMemcpyHostToDevice);
  if (e == hipSuccess && cz) e = hipMemcpy(k + 2 * size, kz, sizeof(float) * size, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    (void)hipGetLastError();
    const float* dkx = cx ? k : nullptr;
    const float* dky = cy ? k + size : nullptr;
    const float* dkz = cz ? k + 2 * size : nullptr;
    switch (size) {
      case 7: e = run_passes<3>(a, b, dkx, dky, dkz, w, h, d, oob, value, &res); break;
      case 15: e = run_passes<7>(a, b, dkx, dky, dkz, w, h, d, oob, value, &res); break;
      case 31: e = run_passes<15>(a, b, dkx, dky, dkz, w, h, d, oob, value, &res); break;
      case 63: e = run_passes<31>(a, b, dkx, dky, dkz, w, h, d, oob, value, &res); break;
      case 127: e = run_passes<63>(a, b, dkx, dky, dkz, w, h, d, oob, value, &res); break;
      default: e = hipErrorInvalidValue;
    }
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = staged_d2h(image, res, sizeof(float) * n, nullptr);
  for (float* p : {a, b, k})
    if (p) (void)hipFree(p);
  (void)hipSetDevice(prev);
  if (e != hipSuccess) {
    std::fprintf(stderr, "convolve_%d: %s\n", size, hipGetErrorString(e));
    set_error(std::string("separable convolution failed: ") + hipGetErrorString(e));
    return 0;
  }
  return 1;
}

}  // namespace

}  // namespace mvd

using namespace mvd;

#define MVD_SEP_ENTRY(N)                                                                                            \
  extern "C" int convolve_##N(float* image, float* kernelX, float* kernelY, float* kernelZ, int imageW, int imageH, \
                              int imageD, int convolveX, int convolveY, int convolveZ, int outofbounds,             \
                              float outofboundsvalue, int devCUDA) {                                                \
    return sep_convolve(N, image, kernelX, kernelY, kernelZ, imageW, imageH, imageD, convolveX, convolveY,          \
                        convolveZ, outofbounds, outofboundsvalue, devCUDA);                                        \
  }
MVD_SEP_ENTRY(7)
MVD_SEP_ENTRY(15)
MVD_SEP_ENTRY(31)
MVD_SEP_ENTRY(63)
MVD_SEP_ENTRY(127)

extern "C" void convolutionCPU(float* image, float* kernelX, float* kernelY, float* kernelZ, int kernelRX, int kernelRY,
                               int kernelRZ, int imageW, int imageH, int imageD, int outofbounds,
                               float outofboundsvalue) {
  // the interface's "CPU" entry runs the same GPU passes on device 0: one
  // kernel size for all axes, the smallest supported size holding 2 R + 1
  const int need = 2 * std::max(kernelRX, std::max(kernelRY, kernelRZ)) + 1;
  int size = 0;
  for (int s : {7, 15, 31, 63, 127})
    if (!size && need <= s) size = s;
  if (!size || kernelRX < 0 || kernelRY < 0 || kernelRZ < 0) {
    std::fprintf(stderr, "convolutionCPU: kernel radius out of range\n");
    return;
  }
  // kernels of 2 R + 1 taps, centred in `size` zeros
  std::vector<float> k(3 * (size_t)size, 0.f);
  const float* src[3] = {kernelX, kernelY, kernelZ};
  const int rad[3] = {kernelRX, kernelRY, kernelRZ};
  for (int a = 0; a < 3; ++a)
    if (src[a])
      for (int i = 0; i <= 2 * rad[a]; ++i) k[a * size + size / 2 - rad[a] + i] = src[a][i];
  (void)sep_convolve(size, image, kernelX ? k.data() : nullptr, kernelY ? k.data() + size : nullptr,
                     kernelZ ? k.data() + 2 * size : nullptr, imageW, imageH, imageD, kernelX != nullptr,
                     kernelY != nullptr, kernelZ != nullptr, outofbounds, outofboundsvalue, 0);
}
