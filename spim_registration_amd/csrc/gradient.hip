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

__global__ __launch_bounds__(kGThreads) void k_extract_psf_grad_shifted(const float* box, Int3 kd, float scale,
                                                                       float* grad) {
  const int64_t n = (int64_t)kd.x * kd.y * kd.z;
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kGThreads) {
    const int x = (int)(i % kd.x), y = (int)((i / kd.x) % kd.y), z = (int)(i / ((int64_t)kd.x * kd.y));
    grad[i] = box[(kd.x - 1 - x) + (int64_t)kd.x * ((kd.y - 1 - y) + (int64_t)kd.y * (kd.z - 1 - z))] * scale;
  }
}

hipError_t launch_extract_psf_grad_shifted(const float* box, Int3 kd, float scale, float* grad, hipStream_t stream) {
  const int64_t n = (int64_t)kd.x * kd.y * kd.z;
  const int64_t blocks = std::min<int64_t>((n + kGThreads - 1) / kGThreads, 4096);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_extract_psf_grad_shifted, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, box, kd, scale,
                     grad);
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
// [lo, lo + wd) adds its value to the image voxel mirror(u).  Positions inside
// the image map one to one onto image voxels: they are added with plain
// read-add-writes over the direct box [da, db) (window coordinates).  The
// positions outside the image on some axis (the mirrored shell of a window
// that crosses an image face) can share voxels with each other and with the
// direct box: a second launch, after the first on the same stream, adds them
// with atomics (their order, and with it the last bits, is not fixed).  This
// path serves windows that reach past an image face by n - 1 or more (several
// reflections); the others take the gather below.
__global__ __launch_bounds__(kGThreads) void k_fold_direct(const float* win, Int3 wd, Int3 lo, Int3 n, Int3 da,
                                                          Int3 db, float* acc) {
  const int bx = db.x - da.x, by = db.y - da.y;
  const int64_t total = (int64_t)bx * by * (db.z - da.z);
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kGThreads) {
    const int x = da.x + (int)(i % bx), y = da.y + (int)((i / bx) % by), z = da.z + (int)(i / ((int64_t)bx * by));
    const int64_t p = (lo.x + x) + (int64_t)n.x * ((lo.y + y) + (int64_t)n.y * (lo.z + z));
    acc[p] += win[x + (int64_t)wd.x * (y + (int64_t)wd.y * z)];
  }
}

// The shell as up to six boxes of window coordinates (below and above the
// direct box along z, then along y inside its z range, then along x inside its
// y and z ranges); prefix[k] counts the positions of the boxes before k.
struct ShellBoxes {
  Int3 a[6], b[6];
  int64_t prefix[7];
  int count;
};

__global__ __launch_bounds__(kGThreads) void k_fold_mirrored(const float* win, Int3 wd, Int3 lo, Int3 n,
                                                            ShellBoxes sb, float* acc) {
  const int64_t total = sb.prefix[sb.count];
  for (int64_t i = blockIdx.x * (int64_t)kGThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kGThreads) {
    int k = 0;
    while (k + 1 < sb.count && sb.prefix[k + 1] <= i) ++k;
    const int64_t j = i - sb.prefix[k];
    const Int3 a = sb.a[k], b = sb.b[k];
    const int bx = b.x - a.x, by = b.y - a.y;
    const int x = a.x + (int)(j % bx), y = a.y + (int)((j / bx) % by), z = a.z + (int)(j / ((int64_t)bx * by));
    const int px = mirror_single(lo.x + x, n.x), py = mirror_single(lo.y + y, n.y), pz = mirror_single(lo.z + z, n.z);
    unsafeAtomicAdd(acc + (px + (int64_t)n.x * (py + (int64_t)n.y * pz)), win[x + (int64_t)wd.x * (y + (int64_t)wd.y * z)]);
  }
}

hipError_t launch_fold_window(const float* win, Int3 wd, Int3 lo, Int3 n, float* acc, hipStream_t stream) {
  const int64_t total = (int64_t)wd.x * wd.y * wd.z;
  if (total <= 0) return hipSuccess;
  const Int3 da = {std::max(0, -lo.x), std::max(0, -lo.y), std::max(0, -lo.z)};
  const Int3 db = {std::min(wd.x, n.x - lo.x), std::min(wd.y, n.y - lo.y), std::min(wd.z, n.z - lo.z)};
  const bool direct = db.x > da.x && db.y > da.y && db.z > da.z;
  ShellBoxes sb{};
  auto add = [&sb](Int3 a, Int3 b) {
    if (b.x > a.x && b.y > a.y && b.z > a.z) {
      sb.a[sb.count] = a;
      sb.b[sb.count] = b;
      sb.prefix[sb.count + 1] = sb.prefix[sb.count] + (int64_t)(b.x - a.x) * (b.y - a.y) * (b.z - a.z);
      ++sb.count;
    }
  };
  if (!direct) {
    add({0, 0, 0}, wd);  // no window position inside the image
  } else {
    add({0, 0, 0}, {wd.x, wd.y, da.z});
    add({0, 0, db.z}, wd);
    add({0, 0, da.z}, {wd.x, da.y, db.z});
    add({0, db.y, da.z}, {wd.x, wd.y, db.z});
    add({0, da.y, da.z}, {da.x, db.y, db.z});
    add({db.x, da.y, da.z}, {wd.x, db.y, db.z});
  }
  (void)hipGetLastError();
  if (direct) {
    const int64_t nd = (int64_t)(db.x - da.x) * (db.y - da.y) * (db.z - da.z);
    const int64_t blocks = std::min<int64_t>((nd + kGThreads - 1) / kGThreads, 16384);
    hipLaunchKernelGGL(k_fold_direct, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, win, wd, lo, n, da, db,
                       acc);
  }
  if (sb.count > 0) {
    const int64_t ns = sb.prefix[sb.count];
    const int64_t blocks = std::min<int64_t>((ns + kGThreads - 1) / kGThreads, 16384);
    hipLaunchKernelGGL(k_fold_mirrored, dim3((unsigned)blocks), dim3(kGThreads), 0, stream, win, wd, lo, n, sb, acc);
  }
  return hipGetLastError();
}

// out = sum of the partials part[0, n) in double (one workgroup, a fixed order).
__global__ __launch_bounds__(kGThreads) void k_reduce_sum(const double* part, int n, double* out) {
  __shared__ double rs[kGThreads];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kGThreads) v += part[i];
  rs[threadIdx.x] = v;
  __syncthreads();
  for (int o = kGThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) rs[threadIdx.x] += rs[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = rs[0];
}

hipError_t launch_reduce_sum(const double* part, int n, double* out, hipStream_t stream) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kGThreads), 0, stream, part, n, out);
  return hipGetLastError();
}

hipError_t launch_sum_to(const float* a, int64_t n, double* part, int nparts, double* out, hipStream_t stream) {
  hipError_t e = launch_sum(a, n, part, nparts, stream);
  if (e != hipSuccess) return e;
  return launch_reduce_sum(part, nparts, out, stream);
}

// spec[0] = (float) *v: the zero-frequency value of a real line set from its
// double sum.
__global__ void k_set_dc(float2* spec, const double* v) { spec[0] = make_float2((float)*v, 0.f); }

hipError_t launch_set_dc(float2* spec, const double* v, hipStream_t stream) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_set_dc, dim3(1), dim3(1), 0, stream, spec, v);
  return hipGetLastError();
}

// The same fold as a gather over the image voxels the window covers (its
// direct box [a, b)): per voxel, the window positions whose mirror-single
// image it is -- the direct one and, near a face the window crosses, the one
// reflected at that face -- are summed, then added to acc (add) or written
// (!add).  No atomics: every voxel is written by one thread.  Needs a window
// that extends past each image face by less than n - 1 (one reflection per
// face).  One workgroup per 256 voxels of an image row: the y and z
// preimages are the workgroup's.
__device__ __forceinline__ int fold_preimages(int p, int lo, int wd, int n, int* w) {
  int c = 0;
  w[c++] = p - lo;  // the direct position (inside [0, wd) on the direct box)
  const int w1 = -p - lo, w2 = 2 * (n - 1) - p - lo;
  if (p > 0 && w1 >= 0 && w1 < wd) w[c++] = w1;
  if (p < n - 1 && w2 >= 0 && w2 < wd) w[c++] = w2;
  return c;
}

__global__ __launch_bounds__(kGThreads) void k_fold_gather(const float* win, Int3 wd, Int3 lo, Int3 n, Int3 a, Int3 b,
                                                          float* acc, int add) {
  const int px = a.x + blockIdx.x * kGThreads + threadIdx.x;
  const int py = a.y + blockIdx.y, pz = a.z + blockIdx.z;
  if (px >= b.x) return;
  int wx[3], wy[3], wz[3];
  const int cy = fold_preimages(py, lo.y, wd.y, n.y, wy);
  const int cz = fold_preimages(pz, lo.z, wd.z, n.z, wz);
  const int cx = fold_preimages(px, lo.x, wd.x, n.x, wx);
  float s = 0.f;
  for (int k = 0; k < cz; ++k)
    for (int j = 0; j < cy; ++j) {
      const float* row = win + (int64_t)wd.x * (wy[j] + (int64_t)wd.y * wz[k]);
      for (int l = 0; l < cx; ++l) s += row[wx[l]];
    }
  float* p = acc + px + (int64_t)n.x * (py + (int64_t)n.y * pz);
  *p = add ? *p + s : s;
}

bool fold_gather_ok(Int3 wd, Int3 lo, Int3 n) {
  auto axis = [](int w, int l, int m) {
    const int below = std::max(0, -l), above = std::max(0, l + w - m);
    return m >= 2 && below < m - 1 && above < m - 1 && l < m && l + w > 0;
  };
  return axis(wd.x, lo.x, n.x) && axis(wd.y, lo.y, n.y) && axis(wd.z, lo.z, n.z);
}

bool fold_gather_covers(Int3 wd, Int3 lo, Int3 n) {
  return lo.x <= 0 && lo.y <= 0 && lo.z <= 0 && lo.x + wd.x >= n.x && lo.y + wd.y >= n.y && lo.z + wd.z >= n.z;
}

hipError_t launch_fold_gather(const float* win, Int3 wd, Int3 lo, Int3 n, float* acc, bool add, hipStream_t stream) {
  if (!fold_gather_ok(wd, lo, n)) return hipErrorInvalidValue;
  const Int3 a = {std::max(0, lo.x), std::max(0, lo.y), std::max(0, lo.z)};
  const Int3 b = {std::min(n.x, lo.x + wd.x), std::min(n.y, lo.y + wd.y), std::min(n.z, lo.z + wd.z)};
  if (b.x <= a.x || b.y <= a.y || b.z <= a.z) return hipSuccess;
  if (b.y - a.y > 65535 || b.z - a.z > 65535) return hipErrorInvalidValue;
  (void)hipGetLastError();
  const dim3 grid((b.x - a.x + kGThreads - 1) / kGThreads, b.y - a.y, b.z - a.z);
  hipLaunchKernelGGL(k_fold_gather, grid, dim3(kGThreads), 0, stream, win, wd, lo, n, a, b, acc, add ? 1 : 0);
  return hipGetLastError();
}

}  // namespace mvd
