// Seeded synthetic multiview data (host only): Gaussian beads, rotated
// anisotropic Gaussian PSFs and Poisson-sampled views.  Deterministic for a
// given seed, independent of the number of threads (every chunk of voxels
// draws from its own stream).
//
// RNG: xoshiro256** (Blackman & Vigna, 2018) seeded through splitmix64 of
// (seed, stream id).  Poisson sampling: inversion by sequential search for
// lambda < 10, the PTRS transformed-rejection method of W. Hoermann, "The
// transformed rejection method for generating Poisson random variables",
// Insurance: Mathematics and Economics 12 (1993) 39-45, for lambda >= 10.
//
// Views are the continuous convolution of the ground truth with the view's
// PSF sampled at voxel centres: a bead of peak amplitude A and width s
// convolved with a normalised Gaussian PSF of covariance S gives
// A s^3 / sqrt(det(s^2 I + S)) exp(-d^T (s^2 I + S)^-1 d / 2); the constant
// background is unchanged by a normalised PSF.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Xoshiro {
  uint64_t s[4];
  Xoshiro(uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (auto& v : s) v = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform double in [0, 1)
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

int64_t poisson(Xoshiro& rng, double lam) {
  if (!(lam > 0)) return 0;
  if (lam < 10) {
    // inversion by sequential search
    double p = std::exp(-lam), f = p, u = rng.uniform();
    int64_t k = 0;
    while (u > f && k < 1000) {
      ++k;
      p *= lam / (double)k;
      f += p;
    }
    return k;
  }
  // PTRS
  const double slam = std::sqrt(lam), loglam = std::log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2);
  for (;;) {
    const double U = rng.uniform() - 0.5;
    const double V = rng.uniform();
    const double us = 0.5 - std::fabs(U);
    const int64_t k = (int64_t)std::floor((2 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <= -lam + k * loglam - std::lgamma(k + 1.0))
      return k;
  }
}

template <class F>
void parallel_for(int64_t n, int threads, F&& f) {
  threads = std::max(1, threads);
  if (threads == 1 || n < 2 * threads) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}

// Covariance of a Gaussian with sigmas (sxy, sxy, sz) rotated by `deg` about y.
void rotated_cov(double sxy, double sz, double deg, double c[3][3]) {
  const double t = deg * M_PI / 180.0, ct = std::cos(t), st = std::sin(t);
  // R = rotation about y: x' = ct x + st z, z' = -st x + ct z
  const double R[3][3] = {{ct, 0, st}, {0, 1, 0}, {-st, 0, ct}};
  const double D[3] = {sxy * sxy, sxy * sxy, sz * sz};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += R[i][k] * D[k] * R[j][k];
      c[i][j] = s;
    }
}

bool invert3(const double a[3][3], double inv[3][3], double* det) {
  const double d = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                   a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                   a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  if (!(std::fabs(d) > 0)) return false;
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / d;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / d;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / d;
  inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / d;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / d;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / d;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / d;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / d;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / d;
  *det = d;
  return true;
}

inline double quad(const double q[3][3], double dx, double dy, double dz) {
  const double v[3] = {dx, dy, dz};
  double s = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s += v[i] * q[i][j] * v[j];
  return s;
}

}  // namespace

extern "C" {

// Beads: n records {x, y, z, amplitude}; positions uniform in the volume,
// amplitudes uniform in [amin, amax].
int syn_beads(int n, int nx, int ny, int nz, double amin, double amax, uint64_t seed, double* out) {
  Xoshiro r(seed, 0xBEADull);
  for (int i = 0; i < n; ++i) {
    out[4 * i + 0] = r.uniform() * nx;
    out[4 * i + 1] = r.uniform() * ny;
    out[4 * i + 2] = r.uniform() * nz;
    out[4 * i + 3] = amin + (amax - amin) * r.uniform();
  }
  return 0;
}

// Normalised anisotropic Gaussian PSF (kx, ky, kz), sigmas (sxy, sxy, sz),
// rotated by `deg` about y, centred at index k/2; normalised in double,
// stored as float.
int syn_psf(int kx, int ky, int kz, double sxy, double sz, double deg, float* out) {
  double c[3][3], ci[3][3], det;
  rotated_cov(sxy, sz, deg, c);
  if (!invert3(c, ci, &det)) return -1;
  std::vector<double> v((size_t)kx * ky * kz);
  double sum = 0;
  for (int z = 0; z < kz; ++z)
    for (int y = 0; y < ky; ++y)
      for (int x = 0; x < kx; ++x) {
        const double e = std::exp(-0.5 * quad(ci, x - kx / 2, y - ky / 2, z - kz / 2));
        v[x + (size_t)kx * (y + (size_t)ky * z)] = e;
        sum += e;
      }
  for (size_t i = 0; i < v.size(); ++i) out[i] = (float)(v[i] / sum);
  return 0;
}

// Render beads blurred by a Gaussian of covariance (bead sigma^2 I + PSF cov)
// plus background plus an optional smooth field sum_i a cos(...), then
// Poisson-sample when `poisson_seed` != 0.  `psf_sxy` = 0 renders the ground
// truth itself.  field: nfield records {amplitude, fx, fy, fz, phase}.
int syn_render(int nx, int ny, int nz, const double* beads, int nbeads, double bead_sigma, double psf_sxy,
               double psf_sz, double psf_deg, double background, const double* field, int nfield,
               uint64_t poisson_seed, int threads, float* out) {
  double c[3][3], ci[3][3], det;
  if (psf_sxy > 0) {
    rotated_cov(psf_sxy, psf_sz, psf_deg, c);
  } else {
    std::memset(c, 0, sizeof(c));
  }
  for (int i = 0; i < 3; ++i) c[i][i] += bead_sigma * bead_sigma;
  if (!invert3(c, ci, &det)) return -1;
  const double peak = bead_sigma * bead_sigma * bead_sigma / std::sqrt(det);
  // bounding half-widths: 5 standard deviations of the marginal along each axis
  const int hw[3] = {(int)std::ceil(5 * std::sqrt(c[0][0])), (int)std::ceil(5 * std::sqrt(c[1][1])),
                     (int)std::ceil(5 * std::sqrt(c[2][2]))};
  const int64_t n = (int64_t)nx * ny * nz;
  std::vector<double> acc(n, background);
  if (nfield > 0) {
    parallel_for(nz, threads, [&](int64_t zb, int64_t ze) {
      for (int64_t z = zb; z < ze; ++z)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x) {
            double s = 0;
            for (int f = 0; f < nfield; ++f) {
              const double* p = field + 5 * f;
              s += p[0] * std::cos(2 * M_PI * (p[1] * x / nx + p[2] * y / ny + p[3] * z / nz) + p[4]);
            }
            acc[x + (int64_t)nx * (y + (int64_t)ny * z)] += s;
          }
    });
  }
  // beads, split by z slabs so that threads never write the same voxel
  parallel_for(nz, threads, [&](int64_t zb, int64_t ze) {
    for (int b = 0; b < nbeads; ++b) {
      const double* p = beads + 4 * b;
      const int z0 = std::max<int64_t>(zb, (int64_t)std::floor(p[2] - hw[2]));
      const int z1 = std::min<int64_t>(ze - 1, (int64_t)std::ceil(p[2] + hw[2]));
      if (z0 > z1) continue;
      const int y0 = std::max(0, (int)std::floor(p[1] - hw[1])), y1 = std::min(ny - 1, (int)std::ceil(p[1] + hw[1]));
      const int x0 = std::max(0, (int)std::floor(p[0] - hw[0])), x1 = std::min(nx - 1, (int)std::ceil(p[0] + hw[0]));
      for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
          for (int x = x0; x <= x1; ++x) {
            const double q = quad(ci, x - p[0], y - p[1], z - p[2]);
            if (q > 50) continue;
            acc[x + (int64_t)nx * (y + (int64_t)ny * z)] += p[3] * peak * std::exp(-0.5 * q);
          }
    }
  });
  constexpr int64_t kChunk = 1 << 16;
  const int64_t nchunks = (n + kChunk - 1) / kChunk;
  parallel_for(nchunks, threads, [&](int64_t cb, int64_t ce) {
    for (int64_t ch = cb; ch < ce; ++ch) {
      const int64_t b = ch * kChunk, e = std::min(n, b + kChunk);
      if (poisson_seed) {
        Xoshiro r(poisson_seed, 0x9000000000ull + (uint64_t)ch);
        for (int64_t i = b; i < e; ++i) out[i] = (float)poisson(r, acc[i]);
      } else {
        for (int64_t i = b; i < e; ++i) out[i] = (float)acc[i];
      }
    }
  });
  return 0;
}

// Uniform doubles for tests of the generator.
int syn_uniform(uint64_t seed, uint64_t stream, int n, double* out) {
  Xoshiro r(seed, stream);
  for (int i = 0; i < n; ++i) out[i] = r.uniform();
  return 0;
}

int syn_poisson(uint64_t seed, uint64_t stream, double lam, int n, double* out) {
  Xoshiro r(seed, stream);
  for (int i = 0; i < n; ++i) out[i] = (double)poisson(r, lam);
  return 0;
}

}  // extern "C"
