// PSF extraction from bead locations on the GPU
// (deconv/ExtractPSF.java:277-458: extractNextImg, extractPSFLocal, normalize,
//  transformPSF, transform).
//
// One thread per PSF voxel sums the image interpolated at that voxel's
// position around every bead, in bead order (the reference's float sum, so the
// result is bit-identical to the CPU restatement); one workgroup normalises
// the PSF to [0, 1] (double min/max); one thread per output voxel resamples
// the PSF into the view's transformed frame (zero out of bounds).  The
// interpolation keeps imglib2's NLinearInterpolator3D arithmetic: double
// weights, corner terms rounded to float and summed in float in its corner
// order.  These are small volumes (a PSF is 10^4-10^6 voxels); the work is
// beads x PSF voxels interpolations.
#include <cmath>
#include <string>

#include "../../include/mvdecon.h"
#include "mvd_internal.h"

namespace mvd {

namespace {

constexpr int kPsfThreads = 256;
constexpr int kNormThreads = 1024;

__device__ __forceinline__ int64_t wrap_index(int64_t p, int64_t n) {
  if (p >= 0 && p < n) return p;
  const int64_t q = p % n;
  return q < 0 ? q + n : q;
}

// NLinearInterpolator3D.get at a double position; PERIODIC: extendPeriodic,
// else extendZero.
template <bool PERIODIC>
__device__ __forceinline__ float interp3(const float* __restrict__ img, const int n[3], const double p[3]) {
#pragma clang fp contract(off)
  const double f0 = floor(p[0]), f1 = floor(p[1]), f2 = floor(p[2]);
  const double w0 = p[0] - f0, w0i = 1.0 - w0;
  const double w1 = p[1] - f1, w1i = 1.0 - w1;
  const double w2 = p[2] - f2, w2i = 1.0 - w2;
  const double a00 = w0i * w1i, a10 = w0 * w1i, a01 = w0i * w1, a11 = w0 * w1;
  const double w000 = a00 * w2i, w100 = a10 * w2i, w010 = a01 * w2i, w110 = a11 * w2i;
  const double w001 = a00 * w2, w101 = a10 * w2, w011 = a01 * w2, w111 = a11 * w2;
  // positions stay within a PSF size of the image (|p| < 2^31)
  const int64_t x0 = (int)f0, y0 = (int)f1, z0 = (int)f2;
  const auto g = gmem(img);
  auto at = [&](int i, int j, int k) -> float {
    int64_t x = x0 + i, y = y0 + j, z = z0 + k;
    if (PERIODIC) {
      x = wrap_index(x, n[0]);
      y = wrap_index(y, n[1]);
      z = wrap_index(z, n[2]);
    } else if (x < 0 || y < 0 || z < 0 || x >= n[0] || y >= n[1] || z >= n[2]) {
      return 0.f;
    }
    return g[x + (int64_t)n[0] * (y + (int64_t)n[1] * z)];
  };
  float acc = (float)((double)at(0, 0, 0) * w000);
  acc = acc + (float)((double)at(1, 0, 0) * w100);
  acc = acc + (float)((double)at(1, 1, 0) * w110);
  acc = acc + (float)((double)at(0, 1, 0) * w010);
  acc = acc + (float)((double)at(0, 1, 1) * w011);
  acc = acc + (float)((double)at(1, 1, 1) * w111);
  acc = acc + (float)((double)at(1, 0, 1) * w101);
  acc = acc + (float)((double)at(0, 0, 1) * w001);
  return acc;
}

struct ExtractArgs {
  const float* img;
  int n[3];
  const double* loc;  // nloc x (x, y, z)
  int nloc;
  int size[3];
  float* psf;
};

// extractPSFLocal (ExtractPSF.java:374-415)
__global__ __launch_bounds__(kPsfThreads) void k_extract_psf(ExtractArgs a) {
#pragma clang fp contract(off)
  const int64_t m = (int64_t)a.size[0] * a.size[1] * a.size[2];
  const int64_t i = (int64_t)blockIdx.x * kPsfThreads + threadIdx.x;
  if (i >= m) return;
  const int x = (int)(i % a.size[0]), y = (int)((i / a.size[0]) % a.size[1]), z = (int)(i / ((int64_t)a.size[0] * a.size[1]));
  const double dx = (double)(x - a.size[0] / 2), dy = (double)(y - a.size[1] / 2), dz = (double)(z - a.size[2] / 2);
  const auto loc = gmem(a.loc);
  float acc = 0.f;
  // unrolled so that the gathers of several beads are in flight together; the
  // sum still runs in bead order
#pragma unroll 4
  for (int b = 0; b < a.nloc; ++b) {
    const double p[3] = {dx + loc[3 * b], dy + loc[3 * b + 1], dz + loc[3 * b + 2]};
    acc = acc + interp3<true>(a.img, a.n, p);
  }
  gmem(a.psf)[i] = acc;
}

// normalize (ExtractPSF.java:298-317): one workgroup, min and max in double,
// (v - min) / (max - min) in double, stored as float.
__global__ __launch_bounds__(kNormThreads) void k_normalize_psf(const float* __restrict__ in, float* out, int64_t m) {
  __shared__ double smin[kNormThreads / 64], smax[kNormThreads / 64];
  double mn = 1.7976931348623157e308, mx = -1.7976931348623157e308;
  const auto g = gmem(in);
  for (int64_t i = threadIdx.x; i < m; i += kNormThreads) {
    const double v = g[i];
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double a = __shfl_xor(mn, o), b = __shfl_xor(mx, o);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  if ((threadIdx.x & 63) == 0) {
    smin[threadIdx.x >> 6] = mn;
    smax[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  mn = smin[0];
  mx = smax[0];
  for (int w = 1; w < kNormThreads / 64; ++w) {
    mn = smin[w] < mn ? smin[w] : mn;
    mx = smax[w] > mx ? smax[w] : mx;
  }
  const double range = mx - mn;
  for (int64_t i = threadIdx.x; i < m; i += kNormThreads) gmem(out)[i] = (float)(((double)g[i] - mn) / range);
}

struct PsfTransformArgs {
  const float* psf;
  int size[3];
  double inv[12];
  int out_size[3];
  double offset[3];
  float* out;
};

// transform (ExtractPSF.java:417-458)
__global__ __launch_bounds__(kPsfThreads) void k_transform_psf(PsfTransformArgs a) {
#pragma clang fp contract(off)
  const int64_t m = (int64_t)a.out_size[0] * a.out_size[1] * a.out_size[2];
  const int64_t i = (int64_t)blockIdx.x * kPsfThreads + threadIdx.x;
  if (i >= m) return;
  const int x = (int)(i % a.out_size[0]), y = (int)((i / a.out_size[0]) % a.out_size[1]);
  const int z = (int)(i / ((int64_t)a.out_size[0] * a.out_size[1]));
  const double s0 = (double)x + a.offset[0], s1 = (double)y + a.offset[1], s2 = (double)z + a.offset[2];
  double t[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) t[r] = ((a.inv[4 * r] * s0 + a.inv[4 * r + 1] * s1) + a.inv[4 * r + 2] * s2) + a.inv[4 * r + 3];
  gmem(a.out)[i] = interp3<false>(a.psf, a.size, t);
}

void apply_affine(const double* m, const double s[3], double t[3]) {
  for (int r = 0; r < 3; ++r) t[r] = ((m[4 * r] * s[0] + m[4 * r + 1] * s[1]) + m[4 * r + 2] * s[2]) + m[4 * r + 3];
}

}  // namespace

// transformPSF geometry (ExtractPSF.java:325-372): the 8 corners of
// [0, size - 1] mapped by the forward model give the extent; the output size
// is (int) extent + 1, made odd; the offset puts the mapped centre size / 2 at
// the output centre.
void psf_transform_geometry(const int size[3], const double* model, int new_size[3], double offset[3]) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; ++c) {
    const double s[3] = {(c & 1) ? (double)(size[0] - 1) : 0.0, (c & 2) ? (double)(size[1] - 1) : 0.0,
                         (c & 4) ? (double)(size[2] - 1) : 0.0};
    double t[3];
    apply_affine(model, s, t);
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::fmin(lo[d], t[d]);
      hi[d] = std::fmax(hi[d], t[d]);
    }
  }
  const double centre[3] = {(double)(size[0] / 2), (double)(size[1] / 2), (double)(size[2] / 2)};
  double tc[3];
  apply_affine(model, centre, tc);
  for (int d = 0; d < 3; ++d) {
    new_size[d] = (int)(hi[d] - lo[d]) + 1;
    if (new_size[d] % 2 == 0) ++new_size[d];
    offset[d] = tc[d] - (double)(new_size[d] / 2);
  }
}

}  // namespace mvd

using namespace mvd;

extern "C" int mvd_psf_transform_size(const int* psf_size, const double* model, int* new_size) {
  if (!psf_size || !model || !new_size) {
    set_error("NULL argument");
    return -1;
  }
  for (int d = 0; d < 3; ++d)
    if (psf_size[d] <= 0) {
      set_error("PSF size must be positive");
      return -1;
    }
  for (int i = 0; i < 12; ++i)
    if (!std::isfinite(model[i])) {
      set_error("model must be finite");
      return -1;
    }
  double off[3];
  psf_transform_geometry(psf_size, model, new_size, off);
  for (int d = 0; d < 3; ++d)
    if (new_size[d] <= 0 || new_size[d] > 4096) {
      set_error("transformed PSF size out of range");
      return -1;
    }
  return 0;
}

extern "C" int mvd_extract_psf(int device, const float* img, int nx, int ny, int nz, const double* locations,
                               int nloc, const int* psf_size, const double* model, const double* inverse_model,
                               float* psf_out, float* transformed_out) {
  if (!img || !psf_size || (nloc > 0 && !locations) || !psf_out) {
    set_error("NULL argument");
    return -1;
  }
  if (transformed_out && (!model || !inverse_model)) {
    set_error("the transformed PSF needs the model and its inverse");
    return -1;
  }
  if (nx <= 0 || ny <= 0 || nz <= 0 || nloc < 0) {
    set_error("invalid image size or bead count");
    return -1;
  }
  for (int d = 0; d < 3; ++d)
    if (psf_size[d] <= 0 || psf_size[d] > 4096) {
      set_error("PSF size out of range");
      return -1;
    }
  int new_size[3] = {0, 0, 0};
  double offset[3] = {0, 0, 0};
  if (transformed_out) {
    for (int i = 0; i < 12; ++i)
      if (!std::isfinite(model[i]) || !std::isfinite(inverse_model[i])) {
        set_error("model and inverse must be finite");
        return -1;
      }
    psf_transform_geometry(psf_size, model, new_size, offset);
    for (int d = 0; d < 3; ++d)
      if (new_size[d] <= 0 || new_size[d] > 4096) {
        set_error("transformed PSF size out of range");
        return -1;
      }
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
  const int64_t n = (int64_t)nx * ny * nz, m = (int64_t)psf_size[0] * psf_size[1] * psf_size[2];
  const int64_t mt = (int64_t)new_size[0] * new_size[1] * new_size[2];
  float *d_img = nullptr, *d_raw = nullptr, *d_psf = nullptr, *d_out = nullptr;
  double* d_loc = nullptr;
  hipError_t e = hipMalloc(&d_img, sizeof(float) * n);
  if (e == hipSuccess) e = hipMalloc(&d_raw, sizeof(float) * m);
  if (e == hipSuccess) e = hipMalloc(&d_psf, sizeof(float) * m);
  if (e == hipSuccess && nloc > 0) e = hipMalloc(&d_loc, sizeof(double) * 3 * nloc);
  if (e == hipSuccess && transformed_out) e = hipMalloc(&d_out, sizeof(float) * mt);
  if (e == hipSuccess) e = staged_h2d(d_img, img, sizeof(float) * n, nullptr);
  if (e == hipSuccess && nloc > 0) e = hipMemcpy(d_loc, locations, sizeof(double) * 3 * nloc, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    ExtractArgs a{};
    a.img = d_img;
    a.n[0] = nx;
    a.n[1] = ny;
    a.n[2] = nz;
    a.loc = d_loc;
    a.nloc = nloc;
    for (int d = 0; d < 3; ++d) a.size[d] = psf_size[d];
    a.psf = d_raw;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_extract_psf, dim3((unsigned)((m + kPsfThreads - 1) / kPsfThreads)), dim3(kPsfThreads), 0,
                       nullptr, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_normalize_psf, dim3(1), dim3(kNormThreads), 0, nullptr, d_raw, d_psf, m);
    e = hipGetLastError();
  }
  if (e == hipSuccess && transformed_out) {
    PsfTransformArgs t{};
    t.psf = d_psf;
    for (int d = 0; d < 3; ++d) {
      t.size[d] = psf_size[d];
      t.out_size[d] = new_size[d];
      t.offset[d] = offset[d];
    }
    for (int i = 0; i < 12; ++i) t.inv[i] = inverse_model[i];
    t.out = d_out;
    hipLaunchKernelGGL(k_transform_psf, dim3((unsigned)((mt + kPsfThreads - 1) / kPsfThreads)), dim3(kPsfThreads), 0,
                       nullptr, t);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(psf_out, d_psf, sizeof(float) * m, hipMemcpyDeviceToHost);
  if (e == hipSuccess && transformed_out)
    e = hipMemcpy(transformed_out, d_out, sizeof(float) * mt, hipMemcpyDeviceToHost);
  for (void* p : {(void*)d_img, (void*)d_raw, (void*)d_psf, (void*)d_out, (void*)d_loc})
    if (p) (void)hipFree(p);
  (void)hipSetDevice(prev);
  if (e != hipSuccess) {
    set_error(std::string("PSF extraction failed: ") + hipGetErrorString(e));
    return -2;
  }
  return 0;
}
