// Oracle: CPU restatement of the reference's efficient-Bayesian multiview
// deconvolution (fiji/SPIM_Registration, Java).  TEST INFRASTRUCTURE ONLY:
// it is loaded by tests/, by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg as the checker and the CPU baseline; the product library
// never links or calls it.
//
// Two builds of the same templates: T = float restates the reference's FP32
// path (float image data, double sums, the reference's cast points and
// thread portioning); T = double is the FP64 twin used to report rounding
// error.  Each function names the reference lines it restates.  Paths are
// relative to src/main/java/ of the reference; deconv/ = spim/process/fusion/
// deconvolution/, cuda/ = spim/process/cuda/.
//
// Third-party arithmetic that is not in the reference tree and is restated
// from its published behaviour (parity unpinned at the rounding level):
//  * imglib2-algorithm FFTConvolution (version managed by pom-scijava 34.1.0):
//    image extended by the given out-of-bounds strategy, padded centred to a
//    fast size >= n+k-1, kernel zero-padded and wrapped so its centre k/2
//    sits at index 0, real FFT, complex product, inverse, centred unpad.
//    Here the FFT is fft.hpp's and the fast sizes are 5-smooth.
//  * imglib2 OutOfBoundsMirrorSingleBoundary: reflection without repeating
//    the edge sample, period 2n-2.
//  * mpicbg.util.RealSum (AdjustInput) and net.imglib2.util.RealSum
//    (FirstIteration): restated as pairwise (binary-counter) and Kahan
//    summation respectively.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fft.hpp"

namespace orc {

struct Dims {
  int64_t x, y, z;
  int64_t n() const { return x * y * z; }
};

enum Oob { OOB_MIRROR = 0, OOB_ONE = 1, OOB_ZERO = 2 };
enum PsfType { OPTIMIZATION_II = 0, OPTIMIZATION_I = 1, EFFICIENT_BAYESIAN = 2, INDEPENDENT = 3 };

// minValue, deconv/MVDeconvolution.java:70 (a float constant).
static const float kMinValue = 0.0001f;

// FusionHelper.divideIntoPortions, spim/process/fusion/FusionHelper.java:257-280:
// k chunks of floor(n/k), the last one also takes the remainder.
struct Portion {
  int64_t start, size;
};
static std::vector<Portion> divide_into_portions(int64_t n, int k) {
  std::vector<Portion> p;
  const int64_t chunk = n / k, mod = n % k;
  for (int i = 0; i < k; ++i) p.push_back({i * chunk, i == k - 1 ? chunk + mod : chunk});
  return p;
}

// mpicbg.util.RealSum: binary-counter pairwise summation of doubles.
class PairwiseSum {
 public:
  void add(double a) {
    size_t i = 0;
    double s = a;
    while (i < flags_.size() && flags_[i]) {
      flags_[i] = false;
      s += sums_[i];
      sums_[i] = 0.0;
      ++i;
    }
    if (i == flags_.size()) {
      flags_.push_back(false);
      sums_.push_back(0.0);
    }
    flags_[i] = true;
    sums_[i] = s;
  }
  double sum() const {
    double s = 0;
    for (double v : sums_) s += v;
    return s;
  }

 private:
  std::vector<bool> flags_;
  std::vector<double> sums_;
};

// net.imglib2.util.RealSum: compensated (Kahan) summation.
class KahanSum {
 public:
  void add(double v) {
    const double y = v - c_;
    const double t = s_ + y;
    c_ = (t - s_) - y;
    s_ = t;
  }
  double sum() const { return s_; }

 private:
  double s_ = 0, c_ = 0;
};

// AdjustInput.sumImg, deconv/AdjustInput.java:62-100.  2*numThreads portions;
// the final sum adds sums[0] once before the loop over all portions, so
// portion 0 is counted twice (AdjustInput.java:93-97).
template <class T>
double sum_img(const T* k, int64_t n, int num_threads) {
  auto portions = divide_into_portions(n, 2 * num_threads);
  std::vector<double> part;
  for (auto& p : portions) {
    PairwiseSum s;
    for (int64_t j = 0; j < p.size; ++j) s.add((double)k[p.start + j]);
    part.push_back(s.sum());
  }
  PairwiseSum total;
  total.add(part[0]);
  for (double v : part) total.add(v);
  return total.sum();
}

// AdjustInput.normImg, deconv/AdjustInput.java:50-56: t = (T)((double)t / sum).
template <class T>
void norm_img(T* k, int64_t n, int num_threads) {
  const double s = sum_img(k, n, num_threads);
  for (int64_t i = 0; i < n; ++i) k[i] = (T)((double)k[i] / s);
}

// Mirror.mirror, deconv/Mirror.java:52-129, sequential semantics: every
// position p <= n/2 along `axis` is swapped with n-1-p.  For even n the pair
// (n/2-1, n/2) is swapped twice and therefore stays in place.
template <class T>
void mirror(T* img, Dims d, int axis) {
  const int64_t len = axis == 0 ? d.x : axis == 1 ? d.y : d.z;
  const int64_t half = len / 2, maxm = len - 1;
  for (int64_t z = 0; z < d.z; ++z)
    for (int64_t y = 0; y < d.y; ++y)
      for (int64_t x = 0; x < d.x; ++x) {
        int64_t p[3] = {x, y, z};
        if (p[axis] > half) continue;
        p[axis] = maxm - p[axis];
        const int64_t a = x + d.x * (y + d.y * z);
        const int64_t b = p[0] + d.x * (p[1] + d.y * p[2]);
        std::swap(img[a], img[b]);
      }
}

// MVDeconFFT.computeInvertedKernel, deconv/MVDeconFFT.java:336-344.
template <class T>
std::vector<T> inverted(const std::vector<T>& k, Dims d) {
  std::vector<T> r = k;
  for (int a = 0; a < 3; ++a) mirror(r.data(), d, a);
  return r;
}

// imglib2 OutOfBoundsMirrorSingleBoundary index map.
inline int64_t mirror_index(int64_t p, int64_t n) {
  if (n == 1) return 0;
  const int64_t period = 2 * n - 2;
  p %= period;
  if (p < 0) p += period;
  return p < n ? p : period - p;
}

// Value of E(in) at integer position (x, y, z).
template <class T>
inline T extended(const T* in, Dims d, Oob oob, int64_t x, int64_t y, int64_t z) {
  const bool inside = x >= 0 && x < d.x && y >= 0 && y < d.y && z >= 0 && z < d.z;
  if (inside) return in[x + d.x * (y + d.y * z)];
  if (oob == OOB_ONE) return (T)1;
  if (oob == OOB_ZERO) return (T)0;
  return in[mirror_index(x, d.x) + d.x * (mirror_index(y, d.y) + d.y * mirror_index(z, d.z))];
}

// FFTConvolution (imglib2-algorithm) as called by the reference:
// out(x) = sum_j E(in)(x + k/2 - j) * ker(j), out has in's size.  The kernel
// spectrum for a given input size is computed once and reused, as
// FFTConvolution keeps its kernel FFT between calls (MVDeconFFT.java:392-399,
// 413-419).
template <class T>
struct KernelFFT {
  Dims din{0, 0, 0};
  int64_t qx = 0, qy = 0, qz = 0, lx = 0, ly = 0, lz = 0;
  std::unique_ptr<RealFFT3D<T>> fft;
  std::vector<std::complex<T>> spec;
};

template <class T>
KernelFFT<T> make_kernel_fft(Dims din, const T* ker, Dims dk, int threads) {
  KernelFFT<T> kf;
  kf.din = din;
  kf.qx = next_fast_size(din.x + dk.x - 1, true);
  kf.qy = next_fast_size(din.y + dk.y - 1, false);
  kf.qz = next_fast_size(din.z + dk.z - 1, false);
  kf.lx = (kf.qx - din.x) / 2;
  kf.ly = (kf.qy - din.y) / 2;
  kf.lz = (kf.qz - din.z) / 2;
  const int64_t qx = kf.qx, qy = kf.qy, qz = kf.qz;
  std::vector<T> kpad(qx * qy * qz, (T)0);
  const int64_t cx = dk.x / 2, cy = dk.y / 2, cz = dk.z / 2;
  for (int64_t z = 0; z < dk.z; ++z)
    for (int64_t y = 0; y < dk.y; ++y)
      for (int64_t x = 0; x < dk.x; ++x) {
        const int64_t hx = ((x - cx) % qx + qx) % qx, hy = ((y - cy) % qy + qy) % qy,
                      hz = ((z - cz) % qz + qz) % qz;
        kpad[hx + qx * (hy + qy * hz)] = ker[x + dk.x * (y + dk.y * z)];
      }
  kf.fft = std::make_unique<RealFFT3D<T>>((int)qx, (int)qy, (int)qz, threads);
  kf.spec.resize(kf.fft->spectrum_size());
  kf.fft->forward(kpad.data(), kf.spec.data());
  return kf;
}

template <class T>
void fft_convolve(const T* in, Dims din, Oob oob, const KernelFFT<T>& kf, T* out, int threads) {
  using C = std::complex<T>;
  const int64_t qx = kf.qx, qy = kf.qy, qz = kf.qz, lx = kf.lx, ly = kf.ly, lz = kf.lz;
  const int64_t q = qx * qy * qz;
  std::vector<T> pad(q);
  parallel_for(qz, threads, [&](int64_t zb, int64_t ze) {
    for (int64_t z = zb; z < ze; ++z)
      for (int64_t y = 0; y < qy; ++y)
        for (int64_t x = 0; x < qx; ++x)
          pad[x + qx * (y + qy * z)] = extended(in, din, oob, x - lx, y - ly, z - lz);
  });
  std::vector<C> si(kf.fft->spectrum_size());
  kf.fft->forward(pad.data(), si.data());
  const std::vector<C>& sk = kf.spec;
  parallel_for((int64_t)si.size(), threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const C a = si[i], h = sk[i];
      si[i] = C(a.real() * h.real() - a.imag() * h.imag(), a.real() * h.imag() + a.imag() * h.real());
    }
  });
  kf.fft->inverse(si.data(), pad.data());
  const T scale = (T)(1.0L / (long double)q);
  parallel_for(din.z, threads, [&](int64_t zb, int64_t ze) {
    for (int64_t z = zb; z < ze; ++z)
      for (int64_t y = 0; y < din.y; ++y)
        for (int64_t x = 0; x < din.x; ++x)
          out[x + din.x * (y + din.y * z)] = pad[(x + lx) + qx * ((y + ly) + qy * (z + lz))] * scale;
  });
}

template <class T>
void fft_convolve(const T* in, Dims din, Oob oob, const T* ker, Dims dk, T* out, int threads) {
  KernelFFT<T> kf = make_kernel_fft(din, ker, dk, threads);
  fft_convolve(in, din, oob, kf, out, threads);
}

// Brute-force FP64 direct convolution with the same definition.
void direct_convolve(const double* in, Dims din, Oob oob, const double* ker, Dims dk, double* out,
                     int threads) {
  const int64_t cx = dk.x / 2, cy = dk.y / 2, cz = dk.z / 2;
  parallel_for(din.z, threads, [&](int64_t zb, int64_t ze) {
    for (int64_t z = zb; z < ze; ++z)
      for (int64_t y = 0; y < din.y; ++y)
        for (int64_t x = 0; x < din.x; ++x) {
          double acc = 0;
          for (int64_t j = 0; j < dk.z; ++j)
            for (int64_t i = 0; i < dk.y; ++i)
              for (int64_t h = 0; h < dk.x; ++h)
                acc += extended(in, din, oob, x + cx - h, y + cy - i, z + cz - j) *
                       ker[h + dk.x * (i + dk.y * j)];
          out[x + din.x * (y + din.y * z)] = acc;
        }
  });
}

// BlockGeneratorFixedSizePrecise.divideIntoBlocks, cuda/BlockGeneratorFixedSizePrecise.java:46-122.
struct BlockGeo {
  int64_t offset[3], eff_size[3], eff_offset[3], eff_local[3];
};
static std::vector<BlockGeo> divide_into_blocks(Dims img, Dims ker, Dims bs) {
  const int64_t is[3] = {img.x, img.y, img.z}, ks[3] = {ker.x, ker.y, ker.z}, b[3] = {bs.x, bs.y, bs.z};
  int64_t eff[3], nb[3];
  for (int d = 0; d < 3; ++d) {
    eff[d] = b[d] - ks[d] + 1;
    if (eff[d] <= 0) throw std::invalid_argument("block size smaller than kernel");
    nb[d] = is[d] / eff[d] + (is[d] % eff[d] != 0);
  }
  std::vector<BlockGeo> out;
  for (int64_t k = 0; k < nb[2]; ++k)
    for (int64_t j = 0; j < nb[1]; ++j)
      for (int64_t i = 0; i < nb[0]; ++i) {
        const int64_t cur[3] = {i, j, k};
        BlockGeo g;
        for (int d = 0; d < 3; ++d) {
          g.eff_size[d] = eff[d];
          g.eff_offset[d] = cur[d] * eff[d];
          g.offset[d] = g.eff_offset[d] - ks[d] / 2;
          g.eff_local[d] = ks[d] / 2;
          if (g.eff_offset[d] + g.eff_size[d] > is[d]) g.eff_size[d] = is[d] - g.eff_offset[d];
        }
        out.push_back(g);
      }
  return out;
}

// MVDeconFFT.convolve1/convolve2 on the CPU (deconv/MVDeconFFT.java:384-436,
// 477-526; deconv/MVDeconFFTThreads.java:39-71; cuda/Block.java:134-215).
// Without blocks the whole image is convolved with E = oob.  With blocks each
// block is copied from E(in) (Block.copyBlock), convolved with the block's own
// mirror-single extension (FFTConvolution(block, kernel)), and its effective
// region pasted into out (Block.pasteBlock).
template <class T>
void convolve(const T* in, Dims d, Oob oob, const KernelFFT<T>& kf, Dims dk, const Dims* bs, T* out, int threads) {
  if (!bs) {
    fft_convolve(in, d, oob, kf, out, threads);
    return;
  }
  auto blocks = divide_into_blocks(d, dk, *bs);
  std::vector<T> blk(bs->n()), res(bs->n());
  for (auto& g : blocks) {
    parallel_for(bs->z, threads, [&](int64_t zb, int64_t ze) {
      for (int64_t z = zb; z < ze; ++z)
        for (int64_t y = 0; y < bs->y; ++y)
          for (int64_t x = 0; x < bs->x; ++x)
            blk[x + bs->x * (y + bs->y * z)] =
                extended(in, d, oob, x + g.offset[0], y + g.offset[1], z + g.offset[2]);
    });
    fft_convolve(blk.data(), *bs, OOB_MIRROR, kf, res.data(), threads);
    for (int64_t z = 0; z < g.eff_size[2]; ++z)
      for (int64_t y = 0; y < g.eff_size[1]; ++y)
        for (int64_t x = 0; x < g.eff_size[0]; ++x)
          out[(x + g.eff_offset[0]) + d.x * ((y + g.eff_offset[1]) + d.y * (z + g.eff_offset[2]))] =
              res[(x + g.eff_local[0]) + bs->x * ((y + g.eff_local[1]) + bs->y * (z + g.eff_local[2]))];
  }
}

template <class T>
void convolve(const T* in, Dims d, Oob oob, const T* ker, Dims dk, const Dims* bs, T* out, int threads) {
  KernelFFT<T> kf = make_kernel_fft(bs ? *bs : d, ker, dk, threads);
  convolve(in, d, oob, kf, dk, bs, out, threads);
}

// MVDeconFFT.init, deconv/MVDeconFFT.java:183-324, for all views in order
// (MVDeconInput.init, deconv/MVDeconInput.java:62-68).  k1[v] is normalised in
// place when view v is initialised; view v reads the current k1 of the other
// views, which are normalised only if they come earlier.
template <class T>
void init_kernels(std::vector<std::vector<T>>& k1, const std::vector<Dims>& kd, PsfType type,
                  int norm_threads, int threads, std::vector<std::vector<T>>& k2) {
  const int nv = (int)k1.size();
  k2.assign(nv, {});
  for (int v = 0; v < nv; ++v) {
    norm_img(k1[v].data(), kd[v].n(), norm_threads);
    if (nv == 1 || type == INDEPENDENT) {
      k2[v] = inverted(k1[v], kd[v]);
    } else if (type == EFFICIENT_BAYESIAN) {
      // MVDeconFFT.java:202-265
      std::vector<T> tmp = inverted(k1[v], kd[v]);
      std::vector<T> a(kd[v].n()), b(kd[v].n());
      for (int w = 0; w < nv; ++w) {
        if (w == v) continue;
        std::vector<T> inv_v = inverted(k1[v], kd[v]);
        fft_convolve(inv_v.data(), kd[v], OOB_ZERO, k1[w].data(), kd[w], a.data(), threads);
        std::vector<T> inv_w = inverted(k1[w], kd[w]);
        fft_convolve(a.data(), kd[v], OOB_ZERO, inv_w.data(), kd[w], b.data(), threads);
        for (int64_t i = 0; i < kd[v].n(); ++i) tmp[i] = b[i] * tmp[i];
      }
      norm_img(tmp.data(), kd[v].n(), norm_threads);
      k2[v] = tmp;
    } else if (type == OPTIMIZATION_I) {
      // MVDeconFFT.java:266-312
      std::vector<T> tmp = k1[v];
      std::vector<T> a(kd[v].n());
      for (int w = 0; w < nv; ++w) {
        if (w == v) continue;
        std::vector<T> inv_w = inverted(k1[w], kd[w]);
        fft_convolve(k1[v].data(), kd[v], OOB_ZERO, inv_w.data(), kd[w], a.data(), threads);
        for (int64_t i = 0; i < kd[v].n(); ++i) tmp[i] = a[i] * tmp[i];
      }
      norm_img(tmp.data(), kd[v].n(), norm_threads);
      k2[v] = inverted(tmp, kd[v]);
    } else {
      // OPTIMIZATION_II, MVDeconFFT.java:313-354: float power by repeated product.
      std::vector<T> e = k1[v];
      for (auto& f : e) {
        T r = f;
        for (int i = 1; i < nv; ++i) r *= f;
        f = r;
      }
      norm_img(e.data(), kd[v].n(), norm_threads);
      k2[v] = inverted(e, kd[v]);
    }
  }
}

// MVDeconvolution.fuseFirstIteration + FirstIteration.call, deconv/
// MVDeconvolution.java:213-256 and deconv/FirstIteration.java:84-192.
// Writes the per-voxel count of views with data into `counts` and returns
// the average over voxels with data of the per-voxel mean.
template <class T>
double first_iteration(const std::vector<const T*>& imgs, int64_t n, int norm_threads, T* counts) {
  auto portions = divide_into_portions(n, 2 * norm_threads);
  KahanSum s;
  int64_t count = 0;
  for (auto& p : portions) {
    KahanSum ps;
    int64_t pc = 0;
    for (int64_t i = p.start; i < p.start + p.size; ++i) {
      double sum = 0;
      int c = 0;
      for (auto* im : imgs) {
        const double v = (double)im[i];
        if (v > 0) {
          sum += v;
          ++c;
        }
      }
      if (c > 0) {
        ps.add(sum / c);
        ++pc;
      }
      counts[i] = (T)c;
    }
    s.add(ps.sum());
    count += pc;
  }
  return s.sum() / (double)count;
}

// java.lang.Math.max(double, double): NaN when either argument is NaN
// (std::max returns the first argument instead).  The change statistics use
// it (deconv/MVDeconvolution.java:647,677,454).
inline double java_max(double a, double b) { return a != a ? a : b != b ? b : std::max(a, b); }

// MVDeconvolution.computeQuotient, deconv/MVDeconvolution.java:494-546.
template <class T>
inline T quotient(T blurred, T img) {
  return img > 0 ? img / blurred : (T)1;
}

// MVDeconvolution.computeNextValue + tikhonov, deconv/MVDeconvolution.java:692-726.
template <class T>
inline T next_value(T last, T integral, T weight, double lambda) {
  const T minv = (T)kMinValue;
  const T value = last * integral;
  T adjusted;
  if (value > 0) {
    if (lambda > 0)
      adjusted = (T)((std::sqrt(1.0 + 2.0 * lambda * (double)value) - 1.0) / lambda);
    else
      adjusted = value;
  } else {
    adjusted = minv;
  }
  T next;
  if (std::isnan(adjusted))
    next = minv;
  else
    next = std::max(minv, adjusted);
  return last + ((next - last) * weight);
}

struct RunArgs {
  int nviews;
  Dims img;
  std::vector<Dims> kdims;
  PsfType type;
  int iterations;
  double lambda;
  bool use_blocks;
  Dims block;
  int norm_threads;
  int threads;
};

// MVDeconvolution.<init> + runIteration, deconv/MVDeconvolution.java:94-211,
// 354-465.  stats receives, per iteration and view, {sum change, max change}.
template <class T>
void deconvolve(const RunArgs& a, const std::vector<const T*>& imgs, const std::vector<const T*>& weights,
                std::vector<std::vector<T>> k1, T* psi, double* stats, const T* psi0,
                std::vector<std::vector<T>>* k2_out) {
  const int64_t n = a.img.n();
  std::vector<std::vector<T>> k2;
  init_kernels(k1, a.kdims, a.type, a.norm_threads, a.threads, k2);
  if (k2_out) *k2_out = k2;

  std::vector<T> tmp1(n), tmp2(n);
  if (psi0) {
    std::copy(psi0, psi0 + n, psi);
  } else {
    double avg = first_iteration(imgs, n, a.norm_threads, psi);
    if (std::isnan(avg)) avg = 0.5;  // MVDeconvolution.java:138-142
    const T v = (T)avg;
    std::fill(psi, psi + n, v);
  }
  const Dims* bs = a.use_blocks ? &a.block : nullptr;
  auto portions = divide_into_portions(n, 2 * a.norm_threads);
  std::vector<KernelFFT<T>> kf1, kf2;
  for (int v = 0; v < a.nviews; ++v) {
    kf1.push_back(make_kernel_fft(bs ? *bs : a.img, k1[v].data(), a.kdims[v], a.threads));
    kf2.push_back(make_kernel_fft(bs ? *bs : a.img, k2[v].data(), a.kdims[v], a.threads));
  }
  for (int it = 0; it < a.iterations; ++it) {
    for (int v = 0; v < a.nviews; ++v) {
      convolve(psi, a.img, OOB_MIRROR, kf1[v], a.kdims[v], bs, tmp1.data(), a.threads);
      const T* img = imgs[v];
      parallel_for(n, a.threads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) tmp1[i] = quotient(tmp1[i], img[i]);
      });
      convolve(tmp1.data(), a.img, OOB_ONE, kf2[v], a.kdims[v], bs, tmp2.data(), a.threads);
      const T* w = weights[v];
      std::vector<double> psum(portions.size()), pmax(portions.size());
      parallel_for((int64_t)portions.size(), a.threads, [&](int64_t pb, int64_t pe) {
        for (int64_t p = pb; p < pe; ++p) {
          double sum_change = 0, max_change = -1;
          for (int64_t i = portions[p].start; i < portions[p].start + portions[p].size; ++i) {
            const T last = psi[i];
            const T next = next_value(last, tmp2[i], w ? w[i] : (T)1, a.lambda);
            psi[i] = next;
            const T change = std::abs(next - last);
            sum_change += change;
            max_change = java_max(max_change, (double)change);
          }
          psum[p] = sum_change;
          pmax[p] = max_change;
        }
      });
      double sum_change = 0, max_change = -1;
      for (size_t p = 0; p < portions.size(); ++p) {
        sum_change += psum[p];
        max_change = java_max(max_change, pmax[p]);
      }
      if (stats) {
        stats[(it * a.nviews + v) * 2 + 0] = sum_change;
        stats[(it * a.nviews + v) * 2 + 1] = max_change;
      }
    }
  }
  // MVDeconvolution.java:201-208: zero where no view has data.
  first_iteration(imgs, n, a.norm_threads, tmp1.data());
  for (int64_t i = 0; i < n; ++i)
    if (tmp1[i] == 0) psi[i] = 0;
}


// ---------------------------------------------------------------------------
// Poisson likelihood of the multiview forward model and its gradient (the
// differentiation extension of the path; the reference has no gradient, the
// forward model is its convolve1: b_v = E_mirror(psi) * P_v, MVDeconFFT.java:384-470).
//
//   L = sum_v sum_x w_v(x) [ b_v(x) - y_v(x) ln b_v(x) ]      (x in the image)
//   r_v(x)        = w_v(x) (1 - y_v(x) / b_v(x))
//   dL/dP_v(j)    = sum_x r_v(x) E(psi)(x + c - j)
//   dL/dpsi(p)    = sum_v sum_x r_v(x) sum_j P_v(j) [m(x + c - j) = p]
// with m the mirror-single index map and c = k/2.  Voxels with b <= 0 are
// skipped (contribute nothing).  Sums in double.

// Full-window correlation / adjoint helpers on a padded box of size q >=
// n + k - 1 (window origin -c).
template <class T>
struct Box {
  int64_t qx, qy, qz;
  std::unique_ptr<RealFFT3D<T>> fft;
};

template <class T>
Box<T> make_box(Dims d, Dims dk, int threads) {
  Box<T> b;
  b.qx = next_fast_size(d.x + dk.x - 1, true);
  b.qy = next_fast_size(d.y + dk.y - 1, false);
  b.qz = next_fast_size(d.z + dk.z - 1, false);
  b.fft = std::make_unique<RealFFT3D<T>>((int)b.qx, (int)b.qy, (int)b.qz, threads);
  return b;
}

// `variant` != 0 plants a known mistake for the tests that show the checks of
// this function can fail: 1 drops the mirrored shell of the fold (window
// positions outside the image contribute nothing to dL/dpsi).
template <class T>
double nll_grad(const std::vector<const T*>& imgs, const std::vector<const T*>& weights, const T* psi, Dims d,
                const std::vector<const T*>& psfs, const std::vector<Dims>& kd, T* grad_psi,
                const std::vector<T*>& grad_psf, int threads, int variant = 0) {
  using C = std::complex<T>;
  const int64_t n = d.n();
  std::vector<double> gpsi(n, 0.0);
  double total = 0.0;
  for (size_t v = 0; v < imgs.size(); ++v) {
    const Dims k = kd[v];
    const int64_t cx = k.x / 2, cy = k.y / 2, cz = k.z / 2;
    std::vector<T> blur(n);
    fft_convolve(psi, d, OOB_MIRROR, psfs[v], k, blur.data(), threads);
    std::vector<T> r(n);
    double part = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double b = (double)blur[i];
      const double w = weights.empty() || !weights[v] ? 1.0 : (double)weights[v][i];
      if (!(b > 0)) {
        r[i] = 0;
        continue;
      }
      const double y = (double)imgs[v][i];
      part += w * (b - y * std::log(b));
      r[i] = (T)(w * (1.0 - y / b));
    }
    total += part;
    // window boxes: W(q) = E(psi)(q - c), R(q) = r(q - c) (zero outside the image)
    Box<T> box = make_box<T>(d, k, threads);
    const int64_t qx = box.qx, qy = box.qy, qz = box.qz, q = qx * qy * qz;
    std::vector<T> W(q, (T)0), R(q, (T)0), H(q, (T)0);
    for (int64_t z = 0; z < qz; ++z)
      for (int64_t y = 0; y < qy; ++y)
        for (int64_t x = 0; x < qx; ++x) {
          const int64_t i = x + qx * (y + qy * z);
          if (x < d.x + k.x - 1 && y < d.y + k.y - 1 && z < d.z + k.z - 1)
            W[i] = extended(psi, d, OOB_MIRROR, x - cx, y - cy, z - cz);
          const int64_t px = x - cx, py = y - cy, pz = z - cz;
          if (px >= 0 && px < d.x && py >= 0 && py < d.y && pz >= 0 && pz < d.z)
            R[i] = r[px + d.x * (py + d.y * pz)];
        }
    for (int64_t z = 0; z < k.z; ++z)
      for (int64_t y = 0; y < k.y; ++y)
        for (int64_t x = 0; x < k.x; ++x) {
          const int64_t hx = ((x - cx) % qx + qx) % qx, hy = ((y - cy) % qy + qy) % qy,
                        hz = ((z - cz) % qz + qz) % qz;
          H[hx + qx * (hy + qy * hz)] = psfs[v][x + k.x * (y + k.y * z)];
        }
    const int64_t hs = box.fft->spectrum_size();
    std::vector<C> sw(hs), sr(hs), sh(hs), prod(hs);
    box.fft->forward(W.data(), sw.data());
    box.fft->forward(R.data(), sr.data());
    box.fft->forward(H.data(), sh.data());
    const T scale = (T)(1.0L / (long double)q);
    // dL/dP(j) = corr(R, W)(c - j),  corr(R, W) = IFFT(conj(R^) W^)
    for (int64_t i = 0; i < hs; ++i) prod[i] = std::conj(sr[i]) * sw[i];
    std::vector<T> corr(q);
    box.fft->inverse(prod.data(), corr.data());
    for (int64_t z = 0; z < k.z; ++z)
      for (int64_t y = 0; y < k.y; ++y)
        for (int64_t x = 0; x < k.x; ++x) {
          const int64_t sx = ((cx - x) % qx + qx) % qx, sy = ((cy - y) % qy + qy) % qy,
                        sz = ((cz - z) % qz + qz) % qz;
          grad_psf[v][x + k.x * (y + k.y * z)] = corr[sx + qx * (sy + qy * sz)] * scale;
        }
    // adjoint on the window: G(q) = IFFT(R^ conj(H^)); fold window position q
    // (image position q - c, mirror-extended) back onto the image
    for (int64_t i = 0; i < hs; ++i) prod[i] = sr[i] * std::conj(sh[i]);
    box.fft->inverse(prod.data(), corr.data());
    for (int64_t z = 0; z < d.z + k.z - 1; ++z)
      for (int64_t y = 0; y < d.y + k.y - 1; ++y)
        for (int64_t x = 0; x < d.x + k.x - 1; ++x) {
          const int64_t ix = x - cx, iy = y - cy, iz = z - cz;
          if (variant == 1 && (ix < 0 || ix >= d.x || iy < 0 || iy >= d.y || iz < 0 || iz >= d.z)) continue;
          const int64_t px = mirror_index(ix, d.x), py = mirror_index(iy, d.y), pz = mirror_index(iz, d.z);
          gpsi[px + d.x * (py + d.y * pz)] += (double)(corr[x + qx * (y + qy * z)] * scale);
        }
  }
  for (int64_t i = 0; i < n; ++i) grad_psi[i] = (T)gpsi[i];
  return total;
}

// L for one view in double with the blur given (b = blur + h * shift_j psi):
// central differences of single PSF voxels, each an O(n) rank-one update.
// L(+) - L(-) is summed voxel by voxel as w (2 h s - y (log1p(h s / b) -
// log1p(-h s / b))), the same difference without subtracting two full sums
// (which loses about log10(L / (h dL)) digits); a voxel where either blur is
// <= 0 falls back to its two separate terms, each skipped when <= 0.
double nll_view_fd(const double* img, const double* weight, const double* blur, const double* psi, Dims d,
                   Dims k, const int64_t* voxels, int nvox, double h, double* out, int threads) {
  const int64_t cx = k.x / 2, cy = k.y / 2, cz = k.z / 2;
  for (int t = 0; t < nvox; ++t) {
    const int64_t j = voxels[t];
    const int64_t jx = j % k.x, jy = (j / k.x) % k.y, jz = j / (k.x * k.y);
    // per-z-slice partial sums, each slice written by one thread, summed in order
    std::vector<double> sp(d.z, 0.0);
    parallel_for(d.z, threads, [&](int64_t zb, int64_t ze) {
      for (int64_t z = zb; z < ze; ++z) {
        double diff = 0;
        for (int64_t y = 0; y < d.y; ++y)
          for (int64_t x = 0; x < d.x; ++x) {
            const int64_t i = x + d.x * (y + d.y * z);
            const double s = extended(psi, d, OOB_MIRROR, x + cx - jx, y + cy - jy, z + cz - jz);
            const double w = weight ? weight[i] : 1.0;
            const double b = blur[i], hs = h * s, bp = b + hs, bm = b - hs;
            if (bp > 0 && bm > 0) {
              diff += w * (2.0 * hs - img[i] * (std::log1p(hs / b) - std::log1p(-hs / b)));
            } else {
              if (bp > 0) diff += w * (bp - img[i] * std::log(bp));
              if (bm > 0) diff -= w * (bm - img[i] * std::log(bm));
            }
          }
        sp[z] = diff;
      }
    });
    double dl = 0;
    for (int64_t z = 0; z < d.z; ++z) dl += sp[z];
    out[t] = dl / (2 * h);
  }
  return 0;
}

}  // namespace orc

using namespace orc;

static thread_local std::string g_err;

#define ORC_TRY(...)                \
  try {                             \
    __VA_ARGS__;                    \
    return 0;                       \
  } catch (const std::exception& e) { \
    g_err = e.what();               \
    return -1;                      \
  }

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Portions of FusionHelper.divideIntoPortions: start[k], size[k].
int orc_divide_portions(int64_t n, int k, int64_t* start, int64_t* size) {
  ORC_TRY({
    auto p = divide_into_portions(n, k);
    for (int i = 0; i < k; ++i) {
      start[i] = p[i].start;
      size[i] = p[i].size;
    }
  })
}

// Block geometry: returns the number of blocks, fills up to `cap` records of
// 12 int64 (offset[3], eff_size[3], eff_offset[3], eff_local[3]).
int orc_divide_blocks(const int64_t* img, const int64_t* ker, const int64_t* bs, int64_t* out, int cap) {
  try {
    auto g = divide_into_blocks({img[0], img[1], img[2]}, {ker[0], ker[1], ker[2]}, {bs[0], bs[1], bs[2]});
    for (int i = 0; i < (int)g.size() && i < cap; ++i)
      for (int d = 0; d < 3; ++d) {
        out[i * 12 + d] = g[i].offset[d];
        out[i * 12 + 3 + d] = g[i].eff_size[d];
        out[i * 12 + 6 + d] = g[i].eff_offset[d];
        out[i * 12 + 9 + d] = g[i].eff_local[d];
      }
    return (int)g.size();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int orc_direct_convolve_f64(const double* in, const int64_t* dims, int oob, const double* ker,
                            const int64_t* kdims, double* out, int threads) {
  ORC_TRY(direct_convolve(in, {dims[0], dims[1], dims[2]}, (Oob)oob, ker, {kdims[0], kdims[1], kdims[2]}, out,
                          threads))
}

#define ORC_TYPED(T, SUF)                                                                              \
  double orc_sum_img_##SUF(const T* k, int64_t n, int norm_threads) {                                 \
    return sum_img(k, n, norm_threads);                                                                \
  }                                                                                                    \
  int orc_norm_img_##SUF(T* k, int64_t n, int norm_threads) { ORC_TRY(norm_img(k, n, norm_threads)) } \
  int orc_mirror_##SUF(T* img, const int64_t* dims, int axis) {                                       \
    ORC_TRY(mirror(img, {dims[0], dims[1], dims[2]}, axis))                                           \
  }                                                                                                    \
  int orc_fft_convolve_##SUF(const T* in, const int64_t* dims, int oob, const T* ker, const int64_t* kdims, \
                             T* out, int threads) {                                                    \
    ORC_TRY(fft_convolve(in, {dims[0], dims[1], dims[2]}, (Oob)oob, ker, {kdims[0], kdims[1], kdims[2]}, \
                         out, threads))                                                                \
  }                                                                                                    \
  int orc_convolve_##SUF(const T* in, const int64_t* dims, int oob, const T* ker, const int64_t* kdims, \
                         const int64_t* block, T* out, int threads) {                                  \
    ORC_TRY({                                                                                          \
      Dims bs{0, 0, 0};                                                                                \
      if (block) bs = {block[0], block[1], block[2]};                                                  \
      convolve(in, {dims[0], dims[1], dims[2]}, (Oob)oob, ker, {kdims[0], kdims[1], kdims[2]},         \
               block ? &bs : nullptr, out, threads);                                                   \
    })                                                                                                 \
  }                                                                                                    \
  /* kernels: nviews arrays of kdims[3*v..] each; k2 receives kernel2 */                               \
  int orc_init_kernels_##SUF(int nviews, T* const* k1, const int64_t* kdims, int psf_type,            \
                             int norm_threads, int threads, T* const* k2) {                            \
    ORC_TRY({                                                                                          \
      std::vector<std::vector<T>> kk(nviews), out;                                                     \
      std::vector<Dims> kd(nviews);                                                                    \
      for (int v = 0; v < nviews; ++v) {                                                               \
        kd[v] = {kdims[3 * v], kdims[3 * v + 1], kdims[3 * v + 2]};                                    \
        kk[v].assign(k1[v], k1[v] + kd[v].n());                                                        \
      }                                                                                                \
      init_kernels(kk, kd, (PsfType)psf_type, norm_threads, threads, out);                             \
      for (int v = 0; v < nviews; ++v) {                                                               \
        std::copy(kk[v].begin(), kk[v].end(), k1[v]);                                                  \
        std::copy(out[v].begin(), out[v].end(), k2[v]);                                                \
      }                                                                                                \
    })                                                                                                 \
  }                                                                                                    \
  double orc_first_iteration_##SUF(int nviews, const T* const* imgs, int64_t n, int norm_threads,     \
                                   T* counts) {                                                        \
    std::vector<const T*> v(imgs, imgs + nviews);                                                      \
    return first_iteration(v, n, norm_threads, counts);                                                \
  }                                                                                                    \
  int orc_quotient_##SUF(T* blurred, const T* img, int64_t n) {                                       \
    for (int64_t i = 0; i < n; ++i) blurred[i] = quotient(blurred[i], img[i]);                         \
    return 0;                                                                                          \
  }                                                                                                    \
  int orc_next_value_##SUF(const T* last, const T* integral, const T* weight, int64_t n, double lambda, \
                           T* out) {                                                                   \
    for (int64_t i = 0; i < n; ++i) out[i] = next_value(last[i], integral[i], weight[i], lambda);      \
    return 0;                                                                                          \
  }                                                                                                    \
  int orc_deconvolve_##SUF(int nviews, const int64_t* dims, const T* const* imgs, const T* const* weights, \
                           const T* const* kernels, const int64_t* kdims, int psf_type, int iterations, \
                           double lambda, const int64_t* block, int norm_threads, int threads,         \
                           const T* psi0, T* psi, double* stats, T* const* k2_out) {                   \
    ORC_TRY({                                                                                          \
      RunArgs a;                                                                                       \
      a.nviews = nviews;                                                                               \
      a.img = {dims[0], dims[1], dims[2]};                                                             \
      std::vector<std::vector<T>> k1(nviews);                                                          \
      for (int v = 0; v < nviews; ++v) {                                                               \
        a.kdims.push_back({kdims[3 * v], kdims[3 * v + 1], kdims[3 * v + 2]});                         \
        k1[v].assign(kernels[v], kernels[v] + a.kdims[v].n());                                         \
      }                                                                                                \
      a.type = (PsfType)psf_type;                                                                      \
      a.iterations = iterations;                                                                       \
      a.lambda = lambda;                                                                               \
      a.use_blocks = block != nullptr;                                                                 \
      a.block = block ? Dims{block[0], block[1], block[2]} : Dims{0, 0, 0};                            \
      a.norm_threads = norm_threads;                                                                   \
      a.threads = threads;                                                                             \
      std::vector<const T*> im(imgs, imgs + nviews), w(nviews, nullptr);                               \
      if (weights)                                                                                     \
        for (int v = 0; v < nviews; ++v) w[v] = weights[v];                                            \
      std::vector<std::vector<T>> k2;                                                                  \
      deconvolve(a, im, w, k1, psi, stats, psi0, k2_out ? &k2 : nullptr);                              \
      if (k2_out)                                                                                      \
        for (int v = 0; v < nviews; ++v) std::copy(k2[v].begin(), k2[v].end(), k2_out[v]);             \
    })                                                                                                 \
  }


int orc_nll_view_fd_f64(const double* img, const double* weight, const double* blur, const double* psi,
                        const int64_t* dims, const int64_t* kdims, const int64_t* voxels, int nvox, double h,
                        double* out, int threads) {
  ORC_TRY(nll_view_fd(img, weight, blur, psi, {dims[0], dims[1], dims[2]}, {kdims[0], kdims[1], kdims[2]}, voxels,
                      nvox, h, out, threads))
}

#define ORC_GRAD(T, SUF)                                                                               \
  int orc_nll_grad_##SUF(int nviews, const int64_t* dims, const T* const* imgs, const T* const* weights,  \
                         const T* psi, const T* const* psfs, const int64_t* kdims, T* grad_psi,          \
                         T* const* grad_psf, double* nll, int threads, int variant) {                    \
    ORC_TRY({                                                                                          \
      std::vector<const T*> im(imgs, imgs + nviews), w, ps(psfs, psfs + nviews);                       \
      if (weights) w.assign(weights, weights + nviews);                                                \
      std::vector<Dims> kd;                                                                            \
      for (int v = 0; v < nviews; ++v) kd.push_back({kdims[3 * v], kdims[3 * v + 1], kdims[3 * v + 2]}); \
      std::vector<T*> gp(grad_psf, grad_psf + nviews);                                                 \
      *nll = nll_grad(im, w, psi, {dims[0], dims[1], dims[2]}, ps, kd, grad_psi, gp, threads, variant); \
    })                                                                                                 \
  }
ORC_GRAD(float, f32)
ORC_GRAD(double, f64)

ORC_TYPED(float, f32)
ORC_TYPED(double, f64)

}  // extern "C"
