// Oracle: CPU restatement of the reference's input preparation for the
// deconvolution -- each view resampled into the fused bounding box and its
// blending weight evaluated at the same position.  TEST INFRASTRUCTURE ONLY:
// loaded by tests/ and by bench.py's cpu_baseline leg as the checker and the
// CPU baseline; the product library never links or calls it.
//
// Reference (paths relative to src/main/java/spim/process/fusion/):
//  * deconvolution/TransformInputAndWeights.java:100-135 (image + weights),
//    deconvolution/TransformInput.java:69-95 (image only),
//    deconvolution/TransformWeights.java:53-91 (weights only);
//  * deconvolution/ProcessForDeconvolution.java:187-277 (output = bounding
//    box, offset = its min, portions of numThreads*4, new images are zero);
//  * weights/BlendingRealRandomAccess.java:44-121 (cosine blending, 1001-entry
//    lookup table);
//  * FusionHelper.java:75-81 (intersects).
//
// Third-party arithmetic not in the reference tree, restated from its
// published source (parity unpinned at the rounding level; pinned here by
// closed forms -- identity and integer shifts reproduce samples, linear
// functions are reproduced by the interpolation, the mirror rule at the far
// face -- in tests/test_transform_cpu.py):
//  * imglib2 AffineTransform3D.applyInverse(float[] source, float[] target):
//    source_i = (float)(m_i0*t0 + m_i1*t1 + m_i2*t2 + m_i3), evaluated left to
//    right in double with the inverse matrix m (the caller passes the inverse
//    as AffineTransform3D.inverse().getRowPackedCopy() gives it);
//  * imglib2 NLinearInterpolator3D.get over Views.extendMirrorSingle: weights
//    w_d = pos_d - floor(pos_d) in double, corner weights as products in the
//    order (x, y, z), corners accumulated in FloatType in the traversal order
//    000, 100, 110, 010, 011, 111, 101, 001, each term (float)(v * w) and the
//    running sum in float.
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

const float kMinValue = 0.0001f;  // deconv/MVDeconvolution.java:70

struct Portion {
  int64_t start, size;
};
std::vector<Portion> portions(int64_t n, int k) {  // FusionHelper.java:257-280
  std::vector<Portion> p;
  const int64_t chunk = n / k, mod = n % k;
  for (int i = 0; i < k; ++i) p.push_back({i * chunk, i == k - 1 ? chunk + mod : chunk});
  return p;
}

// BlendingRealRandomAccess.java:44-53: d runs 0, 0.001, ... (accumulated in
// double) while d <= 1.0001; entry round(d * 1000) = (cos((1 - d) pi) + 1) / 2.
struct BlendLut {
  double v[1001];
  BlendLut() {
    for (double& x : v) x = 0;
    for (double d = 0; d <= 1.0001; d = d + 0.001) v[(int)std::floor(d * 1000.0 + 0.5)] = (std::cos((1 - d) * M_PI) + 1) / 2;
  }
};
const BlendLut& lut() {
  static BlendLut l;
  return l;
}

// BlendingRealRandomAccess.computeWeight (:91-121) for an interval with min 0.
float blend_weight(const float t[3], const int64_t size[3], const float border[3], const float range[3]) {
  const BlendLut& L = lut();
  float m = 1;
  for (int d = 0; d < 3; ++d) {
    const float l = t[d];
    const float dim_minus_1 = (float)(size[d] - 1);
    const float a = l - border[d], b = dim_minus_1 - l - border[d];
    // Math.max(0, Math.min(a, b)), NaN propagating as in Java
    const float lo = std::isnan(a) ? a : (std::isnan(b) ? b : (a < b ? a : b));
    const float dist = (std::isnan(lo) || lo > 0.f) ? lo : 0.f;
    if (dist == 0) return 0;
    const float rel = dist / range[d];
    if (rel < 1) m = (float)((double)m * L.v[(int)std::floor((double)rel * 1000.0 + 0.5)]);
  }
  return m;
}

// OutOfBoundsMirrorSingleBoundary: -1 -> 1, n -> n - 2, period 2n - 2 (n = 1: 0).
inline int64_t mirror_single(int64_t p, int64_t n) {
  if (n == 1) return 0;
  const int64_t period = 2 * n - 2;
  int64_t q = p % period;
  if (q < 0) q += period;
  return q < n ? q : period - q;
}

float trilinear(const float* img, const int64_t rs[3], const float t[3]) {
  const double p0 = t[0], p1 = t[1], p2 = t[2];
  const int64_t f0 = (int64_t)std::floor(p0), f1 = (int64_t)std::floor(p1), f2 = (int64_t)std::floor(p2);
  const double w0 = p0 - (double)f0, w0i = 1.0 - w0;
  const double w1 = p1 - (double)f1, w1i = 1.0 - w1;
  const double w2 = p2 - (double)f2, w2i = 1.0 - w2;
  double w[8];
  w[0] = w0i * w1i * w2i;
  w[1] = w0 * w1i * w2i;
  w[2] = w0i * w1 * w2i;
  w[3] = w0 * w1 * w2i;
  w[4] = w0i * w1i * w2;
  w[5] = w0 * w1i * w2;
  w[6] = w0i * w1 * w2;
  w[7] = w0 * w1 * w2;
  const int64_t x[2] = {mirror_single(f0, rs[0]), mirror_single(f0 + 1, rs[0])};
  const int64_t y[2] = {mirror_single(f1, rs[1]), mirror_single(f1 + 1, rs[1])};
  const int64_t z[2] = {mirror_single(f2, rs[2]), mirror_single(f2 + 1, rs[2])};
  auto at = [&](int i, int j, int k) { return img[x[i] + rs[0] * (y[j] + rs[1] * z[k])]; };
  // corner (i, j, k) has weight index i + 2 j + 4 k; traversal order of the
  // interpolator: 000, 100, 110, 010, 011, 111, 101, 001
  static const int order[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, {0, 0, 1}};
  float acc = 0;
  for (int c = 0; c < 8; ++c) {
    const int i = order[c][0], j = order[c][1], k = order[c][2];
    const float term = (float)((double)at(i, j, k) * w[i + 2 * j + 4 * k]);
    acc = c == 0 ? term : acc + term;
  }
  return acc;
}

}  // namespace

extern "C" {

// The blending lookup table (1001 doubles).
void orc_blend_lut(double* out) {
  for (int i = 0; i < 1001; ++i) out[i] = lut().v[i];
}

// Blending weight of a view of `size` (x, y, z) at one position.
float orc_blend_weight(const float* t, const int64_t* size, const float* border, const float* range) {
  return blend_weight(t, size, border, range);
}

// One view into the bounding box: out (nx, ny, nz) voxel s = (x, y, z) +
// offset maps to t = inv(s) in the raw view (rx, ry, rz).  out_img (nullable)
// receives max(minValue, interpolated value) where t lies inside the view and
// is left unchanged elsewhere (the reference's freshly created image holds 0
// there); out_w (nullable) receives the blending weight at t everywhere.
// Threads take the reference's portions (threads * 4 of the output).
int orc_transform_view(const float* raw, const int64_t* rsize, const double* inv, const int64_t* offset,
                       const int64_t* osize, const float* border, const float* range, float* out_img, float* out_w,
                       int threads) {
  if (!inv || !osize || !rsize || (!out_img && !out_w) || (out_img && !raw)) return -1;
  if (out_w && (!border || !range)) return -1;
  const int64_t nx = osize[0], ny = osize[1], n = osize[0] * osize[1] * osize[2];
  if (threads < 1) threads = 1;
  auto work = [&](Portion p) {
    for (int64_t i = p.start; i < p.start + p.size; ++i) {
      const int64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
      // s = cursor position (float) + offset (int), TransformInput.java:79-83
      const float s[3] = {(float)x + (float)offset[0], (float)y + (float)offset[1], (float)z + (float)offset[2]};
      float t[3];
      for (int r = 0; r < 3; ++r) {
        const double a = inv[4 * r] * (double)s[0];
        const double b = inv[4 * r + 1] * (double)s[1];
        const double c = inv[4 * r + 2] * (double)s[2];
        t[r] = (float)(((a + b) + c) + inv[4 * r + 3]);
      }
      if (out_img) {
        // FusionHelper.intersects (double comparisons)
        if ((double)t[0] >= 0 && (double)t[1] >= 0 && (double)t[2] >= 0 && (double)t[0] < (double)rsize[0] &&
            (double)t[1] < (double)rsize[1] && (double)t[2] < (double)rsize[2]) {
          const float v = trilinear(raw, rsize, t);
          out_img[i] = std::isnan(v) ? v : (v > kMinValue ? v : kMinValue);  // Math.max(minValue, v)
        }
      }
      if (out_w) out_w[i] = blend_weight(t, rsize, border, range);
    }
  };
  std::vector<std::thread> pool;
  auto ps = portions(n, threads * 4);
  std::vector<Portion> todo(ps.begin(), ps.end());
  std::atomic<size_t> next{0};
  for (int k = 0; k < threads; ++k)
    pool.emplace_back([&] {
      for (size_t j; (j = next.fetch_add(1)) < todo.size();) work(todo[j]);
    });
  for (auto& th : pool) th.join();
  return 0;
}

}  // extern "C"
