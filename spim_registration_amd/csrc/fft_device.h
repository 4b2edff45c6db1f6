// LDS-resident mixed-radix FFT for CDNA4 (gfx950), 64-wide wavefronts.
//
// A workgroup holds `nlines` complex lines of length n in LDS, element-major
// with the lines fastest: element (line, i) lives at i * lp + line, lp >= nlines
// (the padding spreads power-of-two strides over the LDS banks).  Each pass of
// the Stockham decimation-in-time FFT reads its r inputs into registers, a
// barrier, then writes the r outputs to their autosorted positions; no second
// buffer is needed.  Twiddles come from a table exp(-2*pi*i*m/(n*tw_step)) that
// the caller staged in LDS; the inverse direction uses their conjugates.
#pragma once

#include <hip/hip_runtime.h>

namespace mvd {

constexpr int kMaxPasses = 16;
constexpr int kThreads = 256;       // threads per workgroup for every FFT kernel
constexpr int kLdsCap = 4096;       // max n * nlines complex values per workgroup

struct FftPlan {
  int n;                 // transform length
  int npass;
  int radix[kMaxPasses];
  int ns[kMaxPasses];    // product of the radices of the earlier passes
  unsigned magic[kMaxPasses];  // ceil(2^32 / ns) for ns > 1: j / ns = umulhi(j, magic) for j * ns < 2^32
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by dir * i  (dir = -1 forward, +1 inverse)
__device__ __forceinline__ float2 mul_di(float2 a, float dir) { return make_float2(-dir * a.y, dir * a.x); }

template <int R>
__device__ __forceinline__ void butterfly(float2* v, float dir);

template <>
__device__ __forceinline__ void butterfly<2>(float2* v, float) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void butterfly<3>(float2* v, float dir) {
  const float c1 = -0.5f, s1 = 0.866025403784438646763723f;
  float2 a = v[0], sum = cadd(v[1], v[2]), dif = csub(v[1], v[2]);
  float2 m1 = make_float2(a.x + c1 * sum.x, a.y + c1 * sum.y);
  float2 jd = mul_di(cscale(dif, s1), dir);
  v[0] = cadd(a, sum);
  v[1] = cadd(m1, jd);
  v[2] = csub(m1, jd);
}

template <>
__device__ __forceinline__ void butterfly<4>(float2* v, float dir) {
  float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  float2 a2 = cadd(v[1], v[3]), a3 = mul_di(csub(v[1], v[3]), dir);
  v[0] = cadd(a0, a2);
  v[2] = csub(a0, a2);
  v[1] = cadd(a1, a3);
  v[3] = csub(a1, a3);
}

template <>
__device__ __forceinline__ void butterfly<5>(float2* v, float dir) {
  const float c1 = 0.309016994374947424102293f;   // cos(2pi/5)
  const float c2 = -0.809016994374947424102293f;  // cos(4pi/5)
  const float s1 = 0.951056516295153572116439f;   // sin(2pi/5)
  const float s2 = 0.587785252292473129168706f;   // sin(4pi/5)
  float2 a = v[0];
  float2 p14 = cadd(v[1], v[4]), m14 = csub(v[1], v[4]);
  float2 p23 = cadd(v[2], v[3]), m23 = csub(v[2], v[3]);
  float2 r1 = make_float2(a.x + c1 * p14.x + c2 * p23.x, a.y + c1 * p14.y + c2 * p23.y);
  float2 r2 = make_float2(a.x + c2 * p14.x + c1 * p23.x, a.y + c2 * p14.y + c1 * p23.y);
  float2 i1 = mul_di(make_float2(s1 * m14.x + s2 * m23.x, s1 * m14.y + s2 * m23.y), dir);
  float2 i2 = mul_di(make_float2(s2 * m14.x - s1 * m23.x, s2 * m14.y - s1 * m23.y), dir);
  v[0] = make_float2(a.x + p14.x + p23.x, a.y + p14.y + p23.y);
  v[1] = cadd(r1, i1);
  v[4] = csub(r1, i1);
  v[2] = cadd(r2, i2);
  v[3] = csub(r2, i2);
}

template <>
__device__ __forceinline__ void butterfly<8>(float2* v, float dir) {
  // even/odd split into two radix-4 DFTs, then combine with w8^q
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  butterfly<4>(e, dir);
  butterfly<4>(o, dir);
  const float h = 0.707106781186547524400844f;
  // w8^1 = h(1 + i dir), w8^2 = i dir, w8^3 = h(-1 + i dir)
  float2 t1 = make_float2(h * (o[1].x - dir * o[1].y), h * (o[1].y + dir * o[1].x));
  float2 t2 = mul_di(o[2], dir);
  float2 t3 = make_float2(h * (-o[3].x - dir * o[3].y), h * (-o[3].y + dir * o[3].x));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], t1);
  v[5] = csub(e[1], t1);
  v[2] = cadd(e[2], t2);
  v[6] = csub(e[2], t2);
  v[3] = cadd(e[3], t3);
  v[7] = csub(e[3], t3);
}

// One Stockham pass of radix R over all lines held in LDS.
template <int R>
__device__ __forceinline__ void fft_pass(float2* buf, int lp, int nlines, int n, int ns, unsigned magic,
                                         const float2* tw, int tw_step, float dir) {
  constexpr int kMaxB = (kLdsCap / R + kThreads - 1) / kThreads;
  const int m = n / R;
  const int total = m * nlines;
  const int tstride = (n / (ns * R)) * tw_step;
  // index arithmetic without integer division: nlines is a power of two and
  // j / ns comes from a multiply-high by the plan's reciprocal (j < 4096,
  // ns <= n <= 1024, so j * ns < 2^32 and the quotient is exact)
  const int lshift = __builtin_ctz((unsigned)nlines);
  auto split = [&](int idx, int& line, int& j, int& k, int& q) {
    line = idx & (nlines - 1);
    j = idx >> lshift;
    q = ns == 1 ? j : (int)__umulhi((unsigned)j, magic);
    k = j - q * ns;
  };
  float2 v[kMaxB][R];
#pragma unroll
  for (int b = 0; b < kMaxB; ++b) {
    const int idx = threadIdx.x + b * kThreads;
    if (idx < total) {
      int line, j, k, q;
      split(idx, line, j, k, q);
#pragma unroll
      for (int t = 0; t < R; ++t) {
        float2 x = buf[(j + t * m) * lp + line];
        if (t > 0 && k > 0) {
          float2 w = tw[t * k * tstride];
          w.y *= -dir;  // table holds the forward (dir = -1) twiddles
          x = cmul(x, w);
        }
        v[b][t] = x;
      }
      butterfly<R>(v[b], dir);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < kMaxB; ++b) {
    const int idx = threadIdx.x + b * kThreads;
    if (idx < total) {
      int line, j, k, q;
      split(idx, line, j, k, q);
      const int base = q * ns * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) buf[(base + t * ns) * lp + line] = v[b][t];
    }
  }
  __syncthreads();
}

// Full transform of the lines in LDS (nlines a power of two).  Caller
// synchronises before the call; the function returns after a barrier.
__device__ __forceinline__ void lds_fft(float2* buf, int lp, int nlines, const FftPlan& plan, const float2* tw,
                                        int tw_step, float dir) {
  for (int p = 0; p < plan.npass; ++p) {
    const int r = plan.radix[p], ns = plan.ns[p];
    const unsigned mg = plan.magic[p];
    switch (r) {
      case 2: fft_pass<2>(buf, lp, nlines, plan.n, ns, mg, tw, tw_step, dir); break;
      case 3: fft_pass<3>(buf, lp, nlines, plan.n, ns, mg, tw, tw_step, dir); break;
      case 4: fft_pass<4>(buf, lp, nlines, plan.n, ns, mg, tw, tw_step, dir); break;
      case 5: fft_pass<5>(buf, lp, nlines, plan.n, ns, mg, tw, tw_step, dir); break;
      default: fft_pass<8>(buf, lp, nlines, plan.n, ns, mg, tw, tw_step, dir); break;
    }
  }
}

}  // namespace mvd
