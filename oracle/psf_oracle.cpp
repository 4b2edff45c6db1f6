// Oracle: CPU restatement of the reference's PSF extraction from bead
// locations.  TEST INFRASTRUCTURE ONLY: loaded by tests/ and by bench.py's
// cpu_baseline leg as the checker and the CPU baseline; the product library
// never links or calls it.
//
// Reference (src/main/java/spim/process/fusion/deconvolution/ExtractPSF.java):
//  * extractNextImg (:277-296): extract, normalise, transform;
//  * extractPSFLocal (:374-415): for every bead location in order, every PSF
//    voxel p (x fastest) adds the image interpolated at p - size/2 + location
//    (periodic out-of-bounds) to a float PSF;
//  * normalize (:298-317): (v - min) / (max - min) in double, stored as float;
//  * transformPSF (:325-372): output size (int)(extent of the transformed
//    interval) + 1, made odd; offset so that the transformed centre size/2 is
//    the output centre;
//  * transform (:417-458): every output voxel p takes the PSF interpolated at
//    inverse(p + offset) (zero out of bounds).
//
// Third-party arithmetic not in the reference tree (imglib2), restated from its
// published source (parity unpinned at the rounding level, pinned here by
// closed forms in tests/test_psf_cpu.py): AffineTransform3D.apply on double[]
// (m0*s0 + m1*s1 + m2*s2 + m3 left to right), estimateBounds (the 8 corners of
// [0, size - 1]), NLinearInterpolator3D (double weights, corners accumulated in
// float in the order 000, 100, 110, 010, 011, 111, 101, 001).
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace {

inline int64_t wrap(int64_t p, int64_t n) {
  int64_t q = p % n;
  return q < 0 ? q + n : q;
}

// NLinearInterpolator3D.get at a double position; oob 0 = periodic, 1 = zero.
float interp(const float* img, const int64_t n[3], const double p[3], int oob) {
  const double f0 = std::floor(p[0]), f1 = std::floor(p[1]), f2 = std::floor(p[2]);
  const double w0 = p[0] - f0, w0i = 1.0 - w0;
  const double w1 = p[1] - f1, w1i = 1.0 - w1;
  const double w2 = p[2] - f2, w2i = 1.0 - w2;
  double w[8];
  w[0] = w0i * w1i * w2i;
  w[1] = w0 * w1i * w2i;
  w[2] = w0i * w1 * w2i;
  w[3] = w0 * w1 * w2i;
  w[4] = w0i * w1i * w2;
  w[5] = w0 * w1i * w2;
  w[6] = w0i * w1 * w2;
  w[7] = w0 * w1 * w2;
  const int64_t b[3] = {(int64_t)f0, (int64_t)f1, (int64_t)f2};
  auto at = [&](int i, int j, int k) -> float {
    int64_t x = b[0] + i, y = b[1] + j, z = b[2] + k;
    if (oob == 0) {
      x = wrap(x, n[0]);
      y = wrap(y, n[1]);
      z = wrap(z, n[2]);
    } else if (x < 0 || y < 0 || z < 0 || x >= n[0] || y >= n[1] || z >= n[2]) {
      return 0.f;
    }
    return img[x + n[0] * (y + n[1] * z)];
  };
  static const int order[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, {0, 0, 1}};
  float acc = 0;
  for (int c = 0; c < 8; ++c) {
    const int i = order[c][0], j = order[c][1], k = order[c][2];
    const float term = (float)((double)at(i, j, k) * w[i + 2 * j + 4 * k]);
    acc = c == 0 ? term : acc + term;
  }
  return acc;
}

void apply(const double* m, const double s[3], double t[3]) {
  for (int r = 0; r < 3; ++r) t[r] = ((m[4 * r] * s[0] + m[4 * r + 1] * s[1]) + m[4 * r + 2] * s[2]) + m[4 * r + 3];
}

}  // namespace

extern "C" {

// extractPSFLocal + normalize.  img: n = (nx, ny, nz); loc: nloc x 3 doubles
// (x, y, z); size: PSF size (x, y, z); out: the PSF before (raw, nullable) and
// after normalisation.
int orc_extract_psf(const float* img, const int64_t* n, const double* loc, int64_t nloc, const int64_t* size,
                    float* raw_out, float* out) {
  const int64_t m = size[0] * size[1] * size[2];
  const int64_t half[3] = {size[0] / 2, size[1] / 2, size[2] / 2};
  std::vector<float> psf(m, 0.f);
  for (int64_t b = 0; b < nloc; ++b) {
    for (int64_t i = 0; i < m; ++i) {
      const int64_t x = i % size[0], y = (i / size[0]) % size[1], z = i / (size[0] * size[1]);
      const double p[3] = {(double)(x - half[0]) + loc[3 * b], (double)(y - half[1]) + loc[3 * b + 1],
                           (double)(z - half[2]) + loc[3 * b + 2]};
      psf[i] = psf[i] + interp(img, n, p, 0);
    }
  }
  if (raw_out)
    for (int64_t i = 0; i < m; ++i) raw_out[i] = psf[i];
  double mn = std::numeric_limits<double>::max(), mx = -std::numeric_limits<double>::max();
  for (float v : psf) {
    if ((double)v < mn) mn = v;
    if ((double)v > mx) mx = v;
  }
  for (int64_t i = 0; i < m; ++i) out[i] = (float)(((double)psf[i] - mn) / (mx - mn));
  return 0;
}

// transformPSF's output size and offset for a PSF of `size` and the forward
// map `model` (row-packed 3x4).
void orc_psf_transform_geometry(const int64_t* size, const double* model, int64_t* new_size, double* offset) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; ++c) {
    const double s[3] = {(c & 1) ? (double)(size[0] - 1) : 0.0, (c & 2) ? (double)(size[1] - 1) : 0.0,
                         (c & 4) ? (double)(size[2] - 1) : 0.0};
    double t[3];
    apply(model, s, t);
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::fmin(lo[d], t[d]);
      hi[d] = std::fmax(hi[d], t[d]);
    }
  }
  const double centre[3] = {(double)(size[0] / 2), (double)(size[1] / 2), (double)(size[2] / 2)};
  double tc[3];
  apply(model, centre, tc);
  for (int d = 0; d < 3; ++d) {
    new_size[d] = (int64_t)(int)(hi[d] - lo[d]) + 1;
    if (new_size[d] % 2 == 0) ++new_size[d];
    offset[d] = tc[d] - (double)(new_size[d] / 2);
  }
}

// transform: out (new_size) voxel p takes psf (size) interpolated with zero
// out-of-bounds at inverse(p + offset).
void orc_psf_transform(const float* psf, const int64_t* size, const double* inverse, const int64_t* new_size,
                       const double* offset, float* out) {
  const int64_t m = new_size[0] * new_size[1] * new_size[2];
  for (int64_t i = 0; i < m; ++i) {
    const int64_t x = i % new_size[0], y = (i / new_size[0]) % new_size[1], z = i / (new_size[0] * new_size[1]);
    const double s[3] = {(double)x + offset[0], (double)y + offset[1], (double)z + offset[2]};
    double t[3];
    apply(inverse, s, t);
    out[i] = interp(psf, size, t, 1);
  }
}

}  // extern "C"
