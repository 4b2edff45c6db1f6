// Register-resident FFT building blocks shared by the fast per-axis passes
// and the fused x-y plane passes.
//
// Every line is transformed by T = N / E cooperating threads that each keep
// E complex values in VGPRs.  A length is split into two or three radices
// (256 = 16 x 16, 128 = 16 x 8, 512 = 8 x 8 x 8, 64 = 8 x 8, 288 = 24 x 12,
// 144 = 12 x 12, 192 = 24 x 8, 384 = 8 x 8 x 6); butterflies run in registers and the lines meet in LDS only
// between radix passes -- one LDS exchange for a 256-point line instead of one
// per radix-2/4 stage.  All sizes are compile-time, so index arithmetic is
// shifts, masks and multiplications by constants.
//
// Stockham decimation in time: in a pass of radix R with Ns = product of the
// earlier radices, butterfly j reads x[j + t N/R], multiplies by
// exp(dir 2 pi i t (j mod Ns) / (Ns R)) and writes y[(j div Ns) Ns R + (j mod Ns)
// + t Ns].  A thread owns butterflies j = jt + T b.  After the last pass the
// thread holds X[jt + T b + t N/R_last], which is exactly the input order of a
// transform with the radices reversed -- the fused z pass goes forward,
// multiplies by the kernel spectrum and back without touching LDS in between.
#pragma once

#include "mvd_internal.h"

namespace mvd {

// Returns v unchanged but opaque to the optimiser: values derived from it are
// recomputed where they are used instead of being hoisted and kept live across
// phases or loop iterations (register pressure).
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <>
__device__ __forceinline__ void butterfly<16>(float2* v, float dir) {
  // n = 4 n2 + n1, k = k1 + 4 k2
  const float c1 = 0.923879532511286756128183f;  // cos(pi/8)
  const float s1 = 0.382683432365089771728460f;  // sin(pi/8)
  const float h = 0.707106781186547524400844f;
  float2 a[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    float2 t[4] = {v[n1], v[4 + n1], v[8 + n1], v[12 + n1]};
    butterfly<4>(t, dir);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) a[n1][k1] = t[k1];
  }
  // W16^m = cos(2 pi m / 16) + i dir sin(2 pi m / 16) for m = n1 k1
  const float cs[10] = {1.f, c1, h, s1, 0.f, -s1, -h, -c1, -1.f, -c1};
  const float sn[10] = {0.f, s1, h, c1, 1.f, c1, h, s1, 0.f, -s1};
#pragma unroll
  for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) {
      const int m = n1 * k1;
      a[n1][k1] = cmul(a[n1][k1], make_float2(cs[m], dir * sn[m]));
    }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    butterfly<4>(t, dir);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

// DFT of size N = P Q on v[0, N): n = n1 + P n2, k = k1 + Q k2; inner DFTs of
// size Q over n2, twiddles W_N^(n1 k1), outer DFTs of size P over n1.
// cs[m], sn[m]: cos and sin of 2 pi m / N.
template <int P, int Q>
__device__ __forceinline__ void butterfly_pq(float2* v, float dir, const float* cs, const float* sn) {
  constexpr int N = P * Q;
  float2 a[P][Q];
#pragma unroll
  for (int n1 = 0; n1 < P; ++n1) {
    float2 t[Q];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) t[n2] = v[n1 + P * n2];
    butterfly<Q>(t, dir);
#pragma unroll
    for (int k1 = 0; k1 < Q; ++k1) a[n1][k1] = t[k1];
  }
#pragma unroll
  for (int n1 = 1; n1 < P; ++n1)
#pragma unroll
    for (int k1 = 1; k1 < Q; ++k1) {
      const int m = (n1 * k1) % N;
      a[n1][k1] = cmul(a[n1][k1], make_float2(cs[m], dir * sn[m]));
    }
#pragma unroll
  for (int k1 = 0; k1 < Q; ++k1) {
    float2 t[P];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) t[n1] = a[n1][k1];
    butterfly<P>(t, dir);
#pragma unroll
    for (int k2 = 0; k2 < P; ++k2) v[k1 + Q * k2] = t[k2];
  }
}

template <>
__device__ __forceinline__ void butterfly<6>(float2* v, float dir) {
  const float r3 = 0.866025403784438646763723f;  // sqrt(3) / 2
  const float cs[6] = {1.f, 0.5f, -0.5f, -1.f, -0.5f, 0.5f};
  const float sn[6] = {0.f, r3, r3, 0.f, -r3, -r3};
  butterfly_pq<2, 3>(v, dir, cs, sn);
}

template <>
__device__ __forceinline__ void butterfly<12>(float2* v, float dir) {
  const float r3 = 0.866025403784438646763723f;  // sqrt(3) / 2
  const float cs[12] = {1.f, r3, 0.5f, 0.f, -0.5f, -r3, -1.f, -r3, -0.5f, 0.f, 0.5f, r3};
  const float sn[12] = {0.f, 0.5f, r3, 1.f, r3, 0.5f, 0.f, -0.5f, -r3, -1.f, -r3, -0.5f};
  butterfly_pq<4, 3>(v, dir, cs, sn);
}

template <>
__device__ __forceinline__ void butterfly<24>(float2* v, float dir) {
  const float c1 = 0.965925826289068286749743f;  // cos(pi / 12)
  const float s1 = 0.258819045102520762348899f;  // sin(pi / 12)
  const float r3 = 0.866025403784438646763723f;
  const float h = 0.707106781186547524400844f;
  const float cs[24] = {1.f, c1, r3, h, 0.5f, s1, 0.f, -s1, -0.5f, -h, -r3, -c1,
                        -1.f, -c1, -r3, -h, -0.5f, -s1, 0.f, s1, 0.5f, h, r3, c1};
  const float sn[24] = {0.f, s1, 0.5f, h, r3, c1, 1.f, c1, r3, h, 0.5f, s1,
                        0.f, -s1, -0.5f, -h, -r3, -c1, -1.f, -c1, -r3, -h, -0.5f, -s1};
  butterfly_pq<8, 3>(v, dir, cs, sn);
}

// E values per thread, radices R0 x R1 (x R2), NT threads per workgroup (one
// workgroup holds NT / (N / E) lines).
template <int N>
struct Spec;
template <>
struct Spec<64> {
  static constexpr int E = 8, R0 = 8, R1 = 8, R2 = 1, NT = kThreads;
};
template <>
struct Spec<128> {
  static constexpr int E = 16, R0 = 16, R1 = 8, R2 = 1, NT = kThreads;
};
template <>
struct Spec<144> {
  static constexpr int E = 12, R0 = 12, R1 = 12, R2 = 1, NT = 192;
};
template <>
struct Spec<192> {
  static constexpr int E = 24, R0 = 24, R1 = 8, R2 = 1, NT = kThreads;
};
template <>
struct Spec<256> {
  static constexpr int E = 16, R0 = 16, R1 = 16, R2 = 1, NT = kThreads;
};
template <>
struct Spec<288> {
  static constexpr int E = 24, R0 = 24, R1 = 12, R2 = 1, NT = 192;
};
template <>
struct Spec<384> {
  static constexpr int E = 24, R0 = 8, R1 = 8, R2 = 6, NT = kThreads;
};
template <>
struct Spec<512> {
  static constexpr int E = 16, R0 = 8, R1 = 8, R2 = 8, NT = kThreads;
};

// Slot of element i of line `line` in the line-interleaved exchange buffer of
// the y and z kernels (L lines per workgroup, line_slots<N, L>() entries).
// LDS serves a 64-bit access 32 lanes at a time from 64 four-byte banks; a
// half wave holds the L lines of 32 / L thread groups jt, and the groups must
// land in disjoint banks.  With rows of L + 1 slots the first pass's stores
// (elements R0 jt + t) are disjoint, but the loads (elements jt + T b +
// t N / R) of two groups share a bank pair (256 points; more at 512), so every
// such access takes an extra cycle.  For 256 and 512 points the rows are L
// slots wide and offset by i & 16 (256) or i & 24 (512) slots, which puts the
// groups of a half wave in disjoint banks in every exchange of both transform
// orders, plus 32 slots per 32 elements so that rows do not overlap.  The
// 288-point exchange measures no bank conflicts with rows of L + 1 or L slots;
// rows of L slots save the LDS that fits a fourth workgroup on a CU.  Other
// lengths keep rows of L + 1 slots.
template <int N, int L>
__device__ __forceinline__ int line_slot(int i, int line) {
  if constexpr (N == 256 && L == 16) return (i << 4) + line + (i & 16) + (i & ~31);
  else if constexpr (N == 512 && L == 8) return (i << 3) + line + (i & 24) + (i & ~31);
  else if constexpr (N == 288) return i * L + line;
  else return i * (L + 1) + line;
}

template <int N, int L>
constexpr int line_slots() {
  return N == 288 ? N * L : N * (L + 1);
}

// Butterflies of one pass on the registers (inputs already in pass order).
template <int N, int E, int R, int NS>
__device__ __forceinline__ void bfly_pass(float2 (&v)[E], int jt, const float2* tw, int tw_step, float dir) {
  constexpr int T = N / E;
  constexpr int TS = N / (NS * R);
#pragma unroll
  for (int b = 0; b < E / R; ++b) {
    if constexpr (NS > 1) {
      const int k = (jt + T * b) % NS;
#pragma unroll
      for (int t = 1; t < R; ++t) {
        float2 w = tw[t * k * TS * tw_step];
        w.y *= -dir;
        v[b * R + t] = cmul(v[b * R + t], w);
      }
    }
    butterfly<R>(v + b * R, dir);
  }
}

template <int N, int E, int R, int NS, class Addr>
__device__ __forceinline__ void store_pass(const float2 (&v)[E], float2* lds, int jt, Addr addr) {
  constexpr int T = N / E;
#pragma unroll
  for (int b = 0; b < E / R; ++b) {
    const int j = jt + T * b;
    const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
    for (int t = 0; t < R; ++t) lds[addr(base + t * NS)] = v[b * R + t];
  }
}

template <int N, int E, int R, class Addr>
__device__ __forceinline__ void load_pass(float2 (&v)[E], const float2* lds, int jt, Addr addr) {
  constexpr int T = N / E;
#pragma unroll
  for (int b = 0; b < E / R; ++b)
#pragma unroll
    for (int t = 0; t < R; ++t) v[b * R + t] = lds[addr(jt + T * b + t * (N / R))];
}

// Element index held in v[b * R + t] when the pass of radix R reads or the
// last pass of radix R writes.
template <int N, int E, int R>
__device__ __forceinline__ int elem_index(int jt, int b, int t) {
  return jt + (N / E) * b + t * (N / R);
}

// Full transform.  v: inputs in the order of the first radix A; on return the
// outputs in the order of the last radix.  Starts and ends with a barrier, so
// LDS may be in use by the caller before and after.
template <int N, int E, int A, int B, int C, class Addr>
__device__ __forceinline__ void fft_regs(float2 (&v)[E], float2* lds, int jt, const float2* tw, int tw_step,
                                         float dir, Addr addr) {
  bfly_pass<N, E, A, 1>(v, jt, tw, tw_step, dir);
  __syncthreads();
  store_pass<N, E, A, 1>(v, lds, jt, addr);
  __syncthreads();
  load_pass<N, E, B>(v, lds, jt, addr);
  bfly_pass<N, E, B, A>(v, jt, tw, tw_step, dir);
  if constexpr (C > 1) {
    __syncthreads();
    store_pass<N, E, B, A>(v, lds, jt, addr);
    __syncthreads();
    load_pass<N, E, C>(v, lds, jt, addr);
    bfly_pass<N, E, C, A * B>(v, jt, tw, tw_step, dir);
  }
  __syncthreads();
}

template <int N, class Addr>
__device__ __forceinline__ void fft_forward_order(float2 (&v)[Spec<N>::E], float2* lds, int jt, const float2* tw,
                                                  int tw_step, float dir, Addr addr) {
  using S = Spec<N>;
  fft_regs<N, S::E, S::R0, S::R1, S::R2>(v, lds, jt, tw, tw_step, dir, addr);
}

// Radices reversed: input order = output order of fft_forward_order.
template <int N, class Addr>
__device__ __forceinline__ void fft_reverse_order(float2 (&v)[Spec<N>::E], float2* lds, int jt, const float2* tw,
                                                  int tw_step, float dir, Addr addr) {
  using S = Spec<N>;
  if constexpr (S::R2 > 1)
    fft_regs<N, S::E, S::R2, S::R1, S::R0>(v, lds, jt, tw, tw_step, dir, addr);
  else
    fft_regs<N, S::E, S::R1, S::R0, 1>(v, lds, jt, tw, tw_step, dir, addr);
}

template <int N>
constexpr int last_radix() {
  return Spec<N>::R2 > 1 ? Spec<N>::R2 : Spec<N>::R1;
}

}  // namespace mvd
