// Input preparation on the GPU: one view resampled into the fused bounding box
// and its blending weight evaluated at the same position
// (deconv/TransformInputAndWeights.java:100-135, deconv/TransformInput.java:69-95,
//  spim/process/fusion/weights/BlendingRealRandomAccess.java:44-121).
//
// Threads run along x (the image and weight stores of a wavefront are 256
// contiguous bytes each), four output rows per thread.  The arithmetic
// keeps the reference's types and order so that results are bit-identical to
// the CPU restatement: the inverse affine map in double, evaluated left to
// right without fused multiply-adds and rounded to float; trilinear weights in
// double; each corner term rounded to float and summed in float in the
// interpolator's corner order; blending distances in float and the lookup
// table in double.  The inputs are read through the texture-less vector path:
// the eight corners of neighbouring voxels overlap, so they hit L1/L2 and the
// HBM traffic is about one read of the raw view plus the two outputs.
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mvdecon.h"
#include "mvd_internal.h"

namespace mvd {

namespace {

constexpr int kTfThreads = 256;
constexpr float kMinValueF = 0.0001f;  // deconv/MVDeconvolution.java:70

// BlendingRealRandomAccess.java:47-53: d accumulated in double in steps of
// 0.001 while d <= 1.0001; entry round(d * 1000) = (cos((1 - d) pi) + 1) / 2.
std::vector<double> blend_lut_host() {
  std::vector<double> t(kBlendLutSize, 0.0);
  for (double d = 0; d <= 1.0001; d = d + 0.001) t[(int)std::floor(d * 1000.0 + 0.5)] = (std::cos((1 - d) * M_PI) + 1) / 2;
  return t;
}

std::mutex g_lut_mu;
std::vector<double*> g_lut;  // per device, held until the process ends

__device__ __forceinline__ float blend_weight(const float t[3], const int size[3], const float border[3],
                                              const float range[3], const double* __restrict__ lut) {
#pragma clang fp contract(off)
  float m = 1.f;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float l = t[d];
    const float a = l - border[d], b = (float)(size[d] - 1) - l - border[d];
    // Math.max(0, Math.min(a, b)), NaN propagating as in Java
    const float lo = a != a ? a : (b != b ? b : (a < b ? a : b));
    const float dist = (lo != lo || lo > 0.f) ? lo : 0.f;
    if (dist == 0.f) return 0.f;
    // dist >= range gives a rounded quotient >= 1 (no factor): the division
    // is only needed inside the blending band
    if (!(dist >= range[d])) {
      const float rel = __fdiv_rn(dist, range[d]);
      if (rel < 1.f) {
        const double x = (double)rel * 1000.0;
        m = (float)((double)m * gmem(lut)[(int)floor(x + 0.5)]);
      }
    }
  }
  return m;
}

// Trilinear interpolation over the mirror-single extension, for 0 <= t < size
// (the caller checks): the lower corner is in range and the upper one only
// leaves it at the far face, where the mirror maps n to n - 2 (n = 1: 0).
// kOff32: the raw view spans fewer than 2^32 bytes, so corner addresses are
// 32-bit byte offsets from the (uniform) base.
template <bool kOff32>
__device__ __forceinline__ float trilinear(const float* __restrict__ raw, const int size[3], const float t[3]) {
#pragma clang fp contract(off)
  const double p0 = t[0], p1 = t[1], p2 = t[2];
  const double f0 = floor(p0), f1 = floor(p1), f2 = floor(p2);
  const double w0 = p0 - f0, w0i = 1.0 - w0;
  const double w1 = p1 - f1, w1i = 1.0 - w1;
  const double w2 = p2 - f2, w2i = 1.0 - w2;
  // NLinearInterpolator3D weights: products in the order x, y, z
  const double a00 = w0i * w1i, a10 = w0 * w1i, a01 = w0i * w1, a11 = w0 * w1;
  const double w000 = a00 * w2i, w100 = a10 * w2i, w010 = a01 * w2i, w110 = a11 * w2i;
  const double w001 = a00 * w2, w101 = a10 * w2, w011 = a01 * w2, w111 = a11 * w2;
  const int x0 = (int)f0, y0 = (int)f1, z0 = (int)f2;
  const int x1 = x0 + 1 < size[0] ? x0 + 1 : (size[0] > 1 ? size[0] - 2 : 0);
  const int y1 = y0 + 1 < size[1] ? y0 + 1 : (size[1] > 1 ? size[1] - 2 : 0);
  const int z1 = z0 + 1 < size[2] ? z0 + 1 : (size[2] > 1 ? size[2] - 2 : 0);
  float v000, v100, v110, v010, v011, v111, v101, v001;
  if constexpr (kOff32) {
    const uint32_t sx = (uint32_t)size[0], sxy = (uint32_t)size[0] * (uint32_t)size[1];
    const uint32_t r00 = (y0 * sx + z0 * sxy) * 4u, r10 = (y1 * sx + z0 * sxy) * 4u;
    const uint32_t r01 = (y0 * sx + z1 * sxy) * 4u, r11 = (y1 * sx + z1 * sxy) * 4u;
    const uint32_t c0 = (uint32_t)x0 * 4u, c1 = (uint32_t)x1 * 4u;
    v000 = *gbyte(raw, r00 + c0);
    v100 = *gbyte(raw, r00 + c1);
    v110 = *gbyte(raw, r10 + c1);
    v010 = *gbyte(raw, r10 + c0);
    v011 = *gbyte(raw, r11 + c0);
    v111 = *gbyte(raw, r11 + c1);
    v101 = *gbyte(raw, r01 + c1);
    v001 = *gbyte(raw, r01 + c0);
  } else {
    const int64_t sx = size[0], sxy = (int64_t)size[0] * size[1];
    const int64_t r00 = y0 * sx + z0 * sxy, r10 = y1 * sx + z0 * sxy;
    const int64_t r01 = y0 * sx + z1 * sxy, r11 = y1 * sx + z1 * sxy;
    const auto g = gmem(raw);
    v000 = g[r00 + x0];
    v100 = g[r00 + x1];
    v110 = g[r10 + x1];
    v010 = g[r10 + x0];
    v011 = g[r11 + x0];
    v111 = g[r11 + x1];
    v101 = g[r01 + x1];
    v001 = g[r01 + x0];
  }
  // FloatType accumulation in the interpolator's corner order
  float acc = (float)((double)v000 * w000);
  acc = acc + (float)((double)v100 * w100);
  acc = acc + (float)((double)v110 * w110);
  acc = acc + (float)((double)v010 * w010);
  acc = acc + (float)((double)v011 * w011);
  acc = acc + (float)((double)v111 * w111);
  acc = acc + (float)((double)v101 * w101);
  acc = acc + (float)((double)v001 * w001);
  return acc;
}

// A workgroup covers kTfThreads consecutive x of kTfVox consecutive output rows
// (same z); each thread walks its column down the rows, so the raw rows that
// two neighbouring output rows share are fetched once into this CU's cache
// (every store instruction of a wavefront still writes 256 contiguous bytes).
constexpr int kTfVox = 4;

template <bool kOff32>
__global__ __launch_bounds__(kTfThreads) void k_transform_view(TransformArgs a, int chunks, int ygroups) {
#pragma clang fp contract(off)
  const int zy = blockIdx.x / chunks;  // z * ygroups + y group
  const int x = (blockIdx.x - zy * chunks) * kTfThreads + threadIdx.x;
  if (x >= a.nx) return;
  const int z = zy / ygroups, y0 = (zy - z * ygroups) * kTfVox;
  // cursor position (float) + offset (int): TransformInput.java:79-83;
  // AffineTransform3D.applyInverse: (float)(m0*s0 + m1*s1 + m2*s2 + m3), left
  // to right in double -- m0*s0 and m2*s2 are the column's terms
  const double s0 = (double)((float)x + a.off[0]), s2 = (double)((float)z + a.off[2]);
  double ra[3], rc[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    ra[r] = a.inv[4 * r] * s0;
    rc[r] = a.inv[4 * r + 2] * s2;
  }
#pragma unroll 1
  for (int k = 0; k < kTfVox; ++k) {
    const int y = y0 + k;
    if (y >= a.ny) break;
    const double s1 = (double)((float)y + a.off[1]);
    float t[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) t[r] = (float)(((ra[r] + a.inv[4 * r + 1] * s1) + rc[r]) + a.inv[4 * r + 3]);
    const int64_t o = ((int64_t)z * a.ny + y) * a.nx + x;
    if (a.out_img) {
      float v = 0.f;  // outside the view: the reference's new image keeps 0
      // FusionHelper.intersects
      if (t[0] >= 0.f && t[1] >= 0.f && t[2] >= 0.f && (double)t[0] < (double)a.rsize[0] &&
          (double)t[1] < (double)a.rsize[1] && (double)t[2] < (double)a.rsize[2]) {
        const float r = trilinear<kOff32>(a.raw, a.rsize, t);
        v = r != r ? r : (r > kMinValueF ? r : kMinValueF);  // Math.max(minValue, r)
      }
      gmem(a.out_img)[o] = v;
    }
    if (a.out_w) gmem(a.out_w)[o] = blend_weight(t, a.rsize, a.border, a.range, a.lut);
  }
}

}  // namespace

const double* blend_lut_device(int device) {
  std::lock_guard<std::mutex> lk(g_lut_mu);
  if ((int)g_lut.size() <= device) g_lut.resize(device + 1, nullptr);
  if (!g_lut[device]) {
    const auto t = blend_lut_host();
    double* p = nullptr;
    if (hipMalloc(&p, sizeof(double) * kBlendLutSize) != hipSuccess) return nullptr;
    if (hipMemcpy(p, t.data(), sizeof(double) * kBlendLutSize, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(p);
      return nullptr;
    }
    g_lut[device] = p;
  }
  return g_lut[device];
}

const char* check_transform_args(const TransformArgs& a) {
  if (a.nx <= 0 || a.ny <= 0 || a.nz <= 0) return "output size must be positive";
  if ((int64_t)((a.ny + kTfVox - 1) / kTfVox) * a.nz * ((a.nx + kTfThreads - 1) / kTfThreads) >= (int64_t(1) << 31))
    return "output too large for one launch";
  if (!a.out_img && !a.out_w) return "no output requested";
  if (a.rsize[0] <= 0 || a.rsize[1] <= 0 || a.rsize[2] <= 0) return "raw view size must be positive";
  if (a.out_img && !a.raw) return "raw view missing";
  if ((int64_t)a.rsize[0] * a.rsize[1] * a.rsize[2] >= (int64_t(1) << 40)) return "raw view too large";
  if (a.out_w) {
    if (!a.lut) return "blending table missing";
    for (int d = 0; d < 3; ++d)
      if (!(a.range[d] > 0.f)) return "blending range must be positive";
  }
  for (int d = 0; d < 12; ++d)
    if (!std::isfinite(a.inv[d])) return "affine map must be finite";
  return nullptr;
}

hipError_t launch_transform_view(const TransformArgs& a, hipStream_t stream) {
  const int chunks = (a.nx + kTfThreads - 1) / kTfThreads;
  const int ygroups = (a.ny + kTfVox - 1) / kTfVox;
  const int64_t wgs = (int64_t)ygroups * a.nz * chunks;
  const bool off32 = (int64_t)a.rsize[0] * a.rsize[1] * a.rsize[2] * 4 < (int64_t(1) << 32);
  (void)hipGetLastError();
  if (off32)
    hipLaunchKernelGGL(k_transform_view<true>, dim3((unsigned)wgs), dim3(kTfThreads), 0, stream, a, chunks,
                       ygroups);
  else
    hipLaunchKernelGGL(k_transform_view<false>, dim3((unsigned)wgs), dim3(kTfThreads), 0, stream, a, chunks,
                       ygroups);
  return hipGetLastError();
}

}  // namespace mvd

using namespace mvd;

extern "C" int mvd_transform_view(int device, const float* raw, int rx, int ry, int rz, const double* inverse_affine,
                                  const long long* offset, int nx, int ny, int nz, const float* border,
                                  const float* range, float* out_img, float* out_weight) {
  if (!inverse_affine || !offset) {
    set_error("NULL affine map or offset");
    return -1;
  }
  if ((border == nullptr) != (range == nullptr) || (out_weight && !border)) {
    set_error("weights need both border and range");
    return -1;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    set_error("invalid device");
    return -1;
  }
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (hipSetDevice(device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return -2;
  }
  TransformArgs a{};
  a.rsize[0] = rx;
  a.rsize[1] = ry;
  a.rsize[2] = rz;
  for (int i = 0; i < 12; ++i) a.inv[i] = inverse_affine[i];
  for (int d = 0; d < 3; ++d) a.off[d] = (float)offset[d];
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  if (border)
    for (int d = 0; d < 3; ++d) {
      a.border[d] = border[d];
      a.range[d] = range[d];
    }
  a.lut = out_weight ? blend_lut_device(device) : nullptr;
  const int64_t n = (int64_t)nx * ny * nz, nr = (int64_t)rx * ry * rz;
  float *d_raw = nullptr, *d_img = nullptr, *d_w = nullptr;
  int rc = 0;
  auto fail = [&](const std::string& m, int code) {
    set_error(m);
    rc = code;
  };
  a.raw = out_img ? raw : nullptr;
  a.out_img = out_img;
  a.out_w = out_weight;
  if (const char* msg = check_transform_args(a)) {
    fail(msg, -1);
  } else if (out_img && hipMalloc(&d_raw, sizeof(float) * nr) != hipSuccess) {
    fail("out of device memory (raw view)", -2);
  } else if (out_img && hipMalloc(&d_img, sizeof(float) * n) != hipSuccess) {
    fail("out of device memory (image)", -2);
  } else if (out_weight && hipMalloc(&d_w, sizeof(float) * n) != hipSuccess) {
    fail("out of device memory (weight)", -2);
  } else if (out_img && staged_h2d(d_raw, raw, sizeof(float) * nr, nullptr) != hipSuccess) {
    fail("copy of the raw view failed", -2);
  } else {
    a.raw = d_raw;
    a.out_img = d_img;
    a.out_w = d_w;
    hipError_t e = launch_transform_view(a, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && out_img) e = staged_d2h(out_img, d_img, sizeof(float) * n, nullptr);
    if (e == hipSuccess && out_weight) e = staged_d2h(out_weight, d_w, sizeof(float) * n, nullptr);
    if (e != hipSuccess) fail(std::string("transform failed: ") + hipGetErrorString(e), -2);
  }
  if (d_raw) (void)hipFree(d_raw);
  if (d_img) (void)hipFree(d_img);
  if (d_w) (void)hipFree(d_w);
  (void)hipSetDevice(prev);
  return rc;
}

extern "C" int mvd_transform_view_device(int device, const float* raw, int rx, int ry, int rz,
                                         const double* inverse_affine, const long long* offset, int nx, int ny,
                                         int nz, const float* border, const float* range, float* out_img,
                                         float* out_weight, void* stream) {
  if (!inverse_affine || !offset) {
    set_error("NULL affine map or offset");
    return -1;
  }
  if ((border == nullptr) != (range == nullptr) || (out_weight && !border)) {
    set_error("weights need both border and range");
    return -1;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    set_error("invalid device");
    return -1;
  }
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (hipSetDevice(device) != hipSuccess) {
    set_error("hipSetDevice failed");
    return -2;
  }
  TransformArgs a{};
  a.raw = raw;
  a.rsize[0] = rx;
  a.rsize[1] = ry;
  a.rsize[2] = rz;
  for (int i = 0; i < 12; ++i) a.inv[i] = inverse_affine[i];
  for (int d = 0; d < 3; ++d) a.off[d] = (float)offset[d];
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  if (border)
    for (int d = 0; d < 3; ++d) {
      a.border[d] = border[d];
      a.range[d] = range[d];
    }
  a.out_img = out_img;
  a.out_w = out_weight;
  a.lut = out_weight ? blend_lut_device(device) : nullptr;
  int rc = 0;
  if (const char* msg = check_transform_args(a)) {
    set_error(msg);
    rc = -1;
  } else {
    const hipError_t e = launch_transform_view(a, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
      set_error(std::string("transform launch failed: ") + hipGetErrorString(e));
      rc = -2;
    }
  }
  (void)hipSetDevice(prev);
  return rc;
}
