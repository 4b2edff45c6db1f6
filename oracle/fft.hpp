// CPU mixed-radix FFT used by the oracle (test infrastructure only).
//
// The reference computes its FFT convolutions in the third-party library
// imglib2-algorithm (net.imglib2.algorithm.fft2.FFTConvolution, called from
// deconv/MVDeconFFT.java:394,415,487,508 and the compound-kernel builder at
// deconv/MVDeconFFT.java:222-248,285-295), which in turn uses the Mines JTK
// real/complex FFTs.  Neither library is in the reference tree, so this file
// is a textbook restatement: a Stockham autosort decimation-in-time FFT
// (radix 2/3/4/5 butterflies plus a generic odd-prime DFT) evaluated in the
// arithmetic type T.  Twiddles are computed in long double and rounded once
// to T.  Transforms are unnormalised: forward uses exp(-2*pi*i*jk/n), inverse
// exp(+2*pi*i*jk/n).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <thread>
#include <vector>

namespace orc {

// Run f(begin, end) over [0, n) split into `threads` contiguous chunks.
template <class F>
inline void parallel_for(int64_t n, int threads, F&& f) {
  if (threads <= 1 || n < 2 * threads) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  int64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}

// True if n has no prime factor above 5.
inline bool is_5smooth(int64_t n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

// Smallest 5-smooth size >= n (even if `even`).
inline int64_t next_fast_size(int64_t n, bool even) {
  int64_t m = std::max<int64_t>(n, 1);
  for (;; ++m)
    if (is_5smooth(m) && (!even || m % 2 == 0)) return m;
}

template <class T>
class FFT1D {
 public:
  using C = std::complex<T>;

  explicit FFT1D(int n) : n_(n) {
    if (n < 1) throw std::invalid_argument("fft length < 1");
    int m = n;
    for (int r : {4, 2, 3, 5}) {
      while (m % r == 0) {
        radix_.push_back(r);
        m /= r;
      }
    }
    for (int p = 7; m > 1; p += 2) {
      while (m % p == 0) {
        radix_.push_back(p);
        m /= p;
      }
    }
    tw_.resize(n);
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int k = 0; k < n; ++k) {
      long double a = -two_pi * (long double)k / (long double)n;
      tw_[k] = C((T)cosl(a), (T)sinl(a));
    }
    for (int r : radix_) {
      if (r <= 4 || (int)dft_.size() > r) continue;
      dft_.resize(r + 1);
      dft_[r].resize((size_t)r);
      for (int q = 0; q < r; ++q) {
        long double a = -two_pi * (long double)q / (long double)r;
        dft_[r][q] = C((T)cosl(a), (T)sinl(a));
      }
    }
  }

  int size() const { return n_; }

  // In-place transform of `data` (length n); `scratch` must hold n values.
  // sign = -1 forward, +1 inverse (unnormalised).
  void run(C* data, C* scratch, int sign) const {
    if (n_ == 1) return;
    C* in = data;
    C* out = scratch;
    int ns = 1;
    for (int r : radix_) {
      pass(in, out, r, ns, sign);
      std::swap(in, out);
      ns *= r;
    }
    if (in != data) std::copy(in, in + n_, data);
  }

 private:
  C twiddle(int64_t idx, int sign) const {
    C w = tw_[idx % n_];
    return sign < 0 ? w : std::conj(w);
  }

  void pass(const C* in, C* out, int r, int ns, int sign) const {
    const int m = n_ / r;          // butterflies per pass
    const int stride = n_ / (ns * r);  // twiddle table stride
    C v[64];
    std::vector<C> big;
    C* vb = v;
    if (r > 64) {
      big.resize(r);
      vb = big.data();
    }
    for (int j = 0; j < m; ++j) {
      const int k = j % ns;
      for (int t = 0; t < r; ++t) {
        C x = in[j + t * m];
        if (t && k) x *= twiddle((int64_t)t * k * stride, sign);
        vb[t] = x;
      }
      butterfly(vb, r, sign);
      const int base = (j / ns) * ns * r + k;
      for (int t = 0; t < r; ++t) out[base + t * ns] = vb[t];
    }
  }

  void butterfly(C* v, int r, int sign) const {
    const T s = (T)sign;
    switch (r) {
      case 2: {
        C a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
        return;
      }
      case 4: {
        C a0 = v[0] + v[2], a1 = v[0] - v[2];
        C a2 = v[1] + v[3], a3 = v[1] - v[3];
        // multiply a3 by sign*i
        C ja3(-s * a3.imag(), s * a3.real());
        v[0] = a0 + a2;
        v[2] = a0 - a2;
        v[1] = a1 + ja3;
        v[3] = a1 - ja3;
        return;
      }
      case 3: {
        const T c1 = (T)-0.5L;
        const T s1 = (T)(0.86602540378443864676372317075294L) * s;
        C a = v[0], b = v[1], c = v[2];
        C sum = b + c, dif = b - c;
        C m1 = a + c1 * sum;
        C jd(-s1 * dif.imag(), s1 * dif.real());
        v[0] = a + sum;
        v[1] = m1 + jd;
        v[2] = m1 - jd;
        return;
      }
      default:
        generic(v, r, sign);
        return;
    }
  }

  void generic(C* v, int r, int sign) const {
    C tmp[64];
    std::vector<C> big;
    C* o = tmp;
    if (r > 64) {
      big.resize(r);
      o = big.data();
    }
    const std::vector<C>& w = dft_[r];
    for (int q = 0; q < r; ++q) {
      C acc(0, 0);
      for (int t = 0; t < r; ++t) {
        C f = w[(t * q) % r];
        acc += v[t] * (sign < 0 ? f : std::conj(f));
      }
      o[q] = acc;
    }
    for (int q = 0; q < r; ++q) v[q] = o[q];
  }

  int n_;
  std::vector<int> radix_;
  std::vector<C> tw_;
  std::vector<std::vector<C>> dft_;  // dft_[r][q] = exp(-2*pi*i*q/r) for generic radices
};

// 3D real <-> half-complex transform on an x-fastest volume (qx, qy, qz).
// The half spectrum is laid out [z][y][kx], kx in [0, qx/2], qx even.
template <class T>
class RealFFT3D {
 public:
  using C = std::complex<T>;
  RealFFT3D(int qx, int qy, int qz, int threads)
      : qx_(qx), qy_(qy), qz_(qz), hx_(qx / 2 + 1), threads_(threads), fx_(qx), fy_(qy), fz_(qz) {
    if (qx % 2) throw std::invalid_argument("real FFT length along x must be even");
  }

  int64_t spectrum_size() const { return (int64_t)hx_ * qy_ * qz_; }
  int hx() const { return hx_; }

  void forward(const T* real, C* spec) const {
    // x: complex transform of each real line, keep kx in [0, qx/2]
    parallel_for((int64_t)qy_ * qz_, threads_, [&](int64_t b, int64_t e) {
      std::vector<C> line(qx_), scr(qx_);
      for (int64_t l = b; l < e; ++l) {
        const T* src = real + l * qx_;
        for (int x = 0; x < qx_; ++x) line[x] = C(src[x], 0);
        fx_.run(line.data(), scr.data(), -1);
        std::copy(line.begin(), line.begin() + hx_, spec + l * hx_);
      }
    });
    strided(spec, -1);
  }

  // Inverse (unnormalised): real = sum over frequencies.
  void inverse(C* spec, T* real) const {
    strided(spec, +1);
    parallel_for((int64_t)qy_ * qz_, threads_, [&](int64_t b, int64_t e) {
      std::vector<C> line(qx_), scr(qx_);
      for (int64_t l = b; l < e; ++l) {
        const C* src = spec + l * hx_;
        for (int k = 0; k < hx_; ++k) line[k] = src[k];
        for (int k = hx_; k < qx_; ++k) line[k] = std::conj(src[qx_ - k]);
        fx_.run(line.data(), scr.data(), +1);
        T* dst = real + l * qx_;
        for (int x = 0; x < qx_; ++x) dst[x] = line[x].real();
      }
    });
  }

 private:
  // y then z transforms for the forward sign, z then y for the inverse.
  void strided(C* spec, int sign) const {
    auto ypass = [&] {
      parallel_for((int64_t)qz_ * hx_, threads_, [&](int64_t b, int64_t e) {
        std::vector<C> line(qy_), scr(qy_);
        for (int64_t l = b; l < e; ++l) {
          int64_t z = l / hx_, kx = l % hx_;
          C* base = spec + z * (int64_t)qy_ * hx_ + kx;
          for (int y = 0; y < qy_; ++y) line[y] = base[(int64_t)y * hx_];
          fy_.run(line.data(), scr.data(), sign);
          for (int y = 0; y < qy_; ++y) base[(int64_t)y * hx_] = line[y];
        }
      });
    };
    auto zpass = [&] {
      const int64_t plane = (int64_t)qy_ * hx_;
      parallel_for(plane, threads_, [&](int64_t b, int64_t e) {
        std::vector<C> line(qz_), scr(qz_);
        for (int64_t l = b; l < e; ++l) {
          C* base = spec + l;
          for (int z = 0; z < qz_; ++z) line[z] = base[z * plane];
          fz_.run(line.data(), scr.data(), sign);
          for (int z = 0; z < qz_; ++z) base[z * plane] = line[z];
        }
      });
    };
    if (sign < 0) {
      ypass();
      zpass();
    } else {
      zpass();
      ypass();
    }
  }

  int qx_, qy_, qz_, hx_, threads_;
  FFT1D<T> fx_, fy_, fz_;
};

}  // namespace orc
