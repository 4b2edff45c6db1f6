// Fused x + y passes for FFT boxes whose planes are 256 x 256 (q.x = q.y = 256).
//
// A 1024-thread workgroup owns one z plane.  Forward: the plane is gathered in
// two halves of 128 rows; each half is transformed along x in LDS (real to
// complex: a 128-point complex FFT of the packed even/odd samples, 8 threads
// per row, then the unpack step) and every thread keeps the values of its two
// y columns in registers (16 rows each).  The 129 columns then go through
// 256-point FFTs along y, 64 columns per round, 16 threads per column.  The DC
// and Nyquist columns are real after the x transform, so they share one
// complex column (DC + i Nyquist) and are separated after the y transform; two
// rounds cover all 129 columns.  Inverse: the same steps backwards -- y
// transforms of the spectrum columns (DC and Nyquist packed again), rows
// written back to LDS half by half, complex to real along x, fused epilogue.
//
// The plane never leaves the chip between the x and the y transform, which
// saves one read and one write of the spectrum per direction compared with
// separate x and y passes.  LDS: 256 x 65 complex for the y rounds, overlaid
// with 128 x 129 complex for an x half, plus a 256-entry DC/Nyquist buffer and
// the twiddle table -- about 137 KB, one workgroup per CU.
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "fft_regs.h"
#include "mvd_internal.h"
#include "pointwise_math.h"

namespace mvd {

namespace {

constexpr int kXYThreads = 1024;
constexpr int kXYM = 128;        // complex samples of a packed row (256 reals)
constexpr int kXYN = 256;        // rows of a plane, length of the y transforms
constexpr int kXYMP = kXYM + 1;  // LDS row pitch of an x half (complex)
constexpr int kXYLP = 65;        // LDS pitch of a y round: 64 columns + 1
constexpr int kXYHalf = 128;     // rows per x half
constexpr int kXYLds = kXYN * kXYLP;
static_assert(kXYHalf * kXYMP <= kXYLds, "x half must fit the y-round buffer");

using SX = Spec<kXYM>;  // 16 x 8: 16 values per thread, 8 threads per row
using SY = Spec<kXYN>;  // 16 x 16: 16 values per thread, 16 threads per column
static_assert(SX::E == 16 && SY::E == 16 && SX::R2 == 1 && SY::R2 == 1, "fused x-y radix layout");
static_assert(kXYThreads / (kXYM / SX::E) == kXYHalf, "one x row per 8 threads");
static_assert(kXYThreads / (kXYN / SY::E) == 64, "64 y columns per round");

// Full transform of 16 values per thread; outputs in the order of the last
// radix.  Barriers before the LDS exchange and after it.
template <int N, class Addr>
__device__ __forceinline__ void xy_fft(float2 (&v)[16], float2* lds, int jt, const float2* tw, int tw_step,
                                       float dir, Addr addr) {
  using S = Spec<N>;
  bfly_pass<N, 16, S::R0, 1>(v, jt, tw, tw_step, dir);
  __syncthreads();
  store_pass<N, 16, S::R0, 1>(v, lds, jt, addr);
  __syncthreads();
  load_pass<N, 16, S::R1>(v, lds, jt, addr);
  bfly_pass<N, 16, S::R1, S::R0>(v, jt, tw, tw_step, dir);
  __syncthreads();
}

// Two 256-point column transforms per thread (columns a and b, same row set,
// so the same twiddles): first radix-16 pass on both, LDS exchange of a then
// of b, twiddles loaded once, second pass on both.  Barriers before each
// exchange and at the end.
template <class Addr>
__device__ __forceinline__ void xy_fft_pair(float2 (&a)[16], float2 (&b)[16], float2* lds, int jt, const float2* tw,
                                            float dir, Addr addr) {
  static_assert(SY::R0 == 16 && SY::R1 == 16, "256 = 16 x 16");
  butterfly<16>(a, dir);
  __builtin_amdgcn_sched_barrier(0);
  butterfly<16>(b, dir);
  __syncthreads();
  store_pass<kXYN, 16, 16, 1>(a, lds, jt, addr);
  __syncthreads();
  load_pass<kXYN, 16, 16>(a, lds, jt, addr);
  __syncthreads();
  store_pass<kXYN, 16, 16, 1>(b, lds, jt, addr);
  __syncthreads();
  load_pass<kXYN, 16, 16>(b, lds, jt, addr);
  // second pass: butterfly j = jt reads x[jt + 16 t] with twiddle W^(t jt)
#pragma unroll
  for (int t = 1; t < 16; ++t) {
    float2 w = tw[t * jt];
    w.y *= -dir;
    a[t] = cmul(a[t], w);
    b[t] = cmul(b[t], w);
    // one twiddle live at a time: the plane already fills the register file
    __builtin_amdgcn_sched_barrier(0);
  }
  butterfly<16>(a, dir);
  __builtin_amdgcn_sched_barrier(0);  // keep the two butterflies apart (register pressure)
  butterfly<16>(b, dir);
  __syncthreads();
}

// 128-point transform of the rows of an x half, in place in LDS.
__device__ __forceinline__ void xy_rows_fft(float2* lds, const float2* tw, float dir) {
  static_assert(SX::R0 == 16 && SX::R1 == 8, "128 = 16 x 8");
  const int lane = opaque(threadIdx.x);
  const int line = lane & (kXYHalf - 1), jt = lane >> 7;
  auto addr = [line](int i) { return line * kXYMP + i; };
  float2 v[16];
  load_pass<kXYM, 16, 16>(v, lds, jt, addr);
  butterfly<16>(v, dir);
  __syncthreads();
  store_pass<kXYM, 16, 16, 1>(v, lds, jt, addr);
  __syncthreads();
  load_pass<kXYM, 16, 8>(v, lds, jt, addr);
  // radix-8 pass, butterflies k = jt and k = jt + 8 with twiddles W128^(t k).
  // Those of k + 8 follow from those of k by the constant W128^(8 t) =
  // exp(-i pi t / 8), so only seven twiddles are live at a time.
  constexpr float kC[8] = {1.f, 0.923879532511286756f, 0.707106781186547524f, 0.382683432365089772f,
                           0.f, -0.382683432365089772f, -0.707106781186547524f, -0.923879532511286756f};
  constexpr float kS[8] = {0.f, 0.382683432365089772f, 0.707106781186547524f, 0.923879532511286756f,
                           1.f, 0.923879532511286756f, 0.707106781186547524f, 0.382683432365089772f};
  float2 w[8];
#pragma unroll
  for (int t = 1; t < 8; ++t) w[t] = tw[2 * t * jt];  // exp(-2 pi i t jt / 128)
#pragma unroll
  for (int t = 1; t < 8; ++t) v[t] = cmul(v[t], make_float2(w[t].x, -dir * w[t].y));
  butterfly<8>(v, dir);
#pragma unroll
  for (int t = 1; t < 8; ++t) {
    const float2 w8 = cmul(w[t], make_float2(kC[t], -kS[t]));
    v[8 + t] = cmul(v[8 + t], make_float2(w8.x, -dir * w8.y));
  }
  butterfly<8>(v + 8, dir);
  __syncthreads();
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int t = 0; t < 8; ++t) lds[addr(elem_index<kXYM, 16, 8>(jt, b, t))] = v[b * 8 + t];
}

}  // namespace

struct XYFwdParams {
  ConvGeom g;
  const float* src;
  float2* spec;
  const float2* tw;  // 256 entries exp(-2 pi i m / 256)
  int pitch;
};

// Where the forward pass stores a plane's spectrum: the same plane can be
// needed by several block windows (see PlaneTask), so up to kMaxPlaneDests
// copies are written.
struct PlaneOut {
  int n;
  const PlaneTask* task;     // destinations of a task (n of them), or
  const void* items;         // (the launch's XYFwdParams)
  float2* single;            // the one destination without a task
  __device__ float2* base(int d, int pitch) const;  // start of the plane in destination d
};

template <int MODE>
__device__ __forceinline__ void xy_forward_body(const XYFwdParams& P, int bx, const PlaneOut& po) {
  __shared__ float2 lds[kXYLds];
  __shared__ float2 dcn[kXYN];
  __shared__ float2 tw[kXYN];
  const int tid = threadIdx.x;
  const ConvGeom& g = P.g;
  const int z = bx;
  if (tid < kXYN) tw[tid] = gmem(P.tw)[tid];
  float* fbuf = reinterpret_cast<float*>(lds);
  // y columns of this thread: cl and 64 + cl, rows jy + 16 t
  const int cl = tid & 63, jy = tid >> 6;
  float2 c0[16], c1[16];
  const bool zlive = z < g.win.z;
  int gz = g.off.z + z;
  bool z_in = gz >= 0 && gz < g.src.z;
  if (MODE == kGatherMirror && !z_in) {
    gz = mirror_single(gz, g.src.z);
    z_in = true;
  }
  const int x = tid & 255, rsub = tid >> 8;
  int gx = g.off.x + x;
  const bool x_in = gx >= 0 && gx < g.src.x;
  if (MODE == kGatherMirror && !x_in) gx = mirror_single(gx, g.src.x);
  const bool x_live = x < g.win.x;
  // gather of one half: 128 rows of 256 real samples, lanes along x
  auto gather = [&](int h, float (&vals)[32]) {
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      const int y = h * kXYHalf + rsub + 4 * i;
      float val = 0.f;
      if (zlive && x_live && y < g.win.y) {
        int gy = g.off.y + y;
        const bool y_in = gy >= 0 && gy < g.src.y;
        if (MODE == kGatherMirror) {
          // one reflection covers every window whose halo is shorter than the
          // image; the general (periodic) form handles the rest
          if (!y_in) {
            const int r = gy < 0 ? -gy : 2 * g.src.y - 2 - gy;
            gy = (r >= 0 && r < g.src.y) ? r : mirror_single(gy, g.src.y);
          }
          val = gmem(P.src)[gx + g.src.x * (gy + g.src.y * gz)];
        } else if (z_in && y_in && x_in) {
          val = gmem(P.src)[gx + g.src.x * (gy + g.src.y * gz)];
        } else {
          val = MODE == kGatherOne ? 1.f : 0.f;
        }
      }
      vals[i] = val;
    }
  };
  float vals[32];
  gather(0, vals);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    __syncthreads();  // the previous half's column reads are done
#pragma unroll
    for (int i = 0; i < 32; ++i) fbuf[(rsub + 4 * i) * 2 * kXYMP + x] = vals[i];
    // the second half's loads stay in flight while the first half is transformed
    if (h == 0) gather(1, vals);
    __syncthreads();
    xy_rows_fft(lds, tw, -1.f);
    __syncthreads();
    // unpack Z = FFT(even + i odd) into X[k] and X[M - k], in place
    for (int idx = opaque(threadIdx.x); idx < kXYHalf * (kXYM / 2 + 1); idx += kXYThreads) {
      const int l = idx / (kXYM / 2 + 1), k = idx % (kXYM / 2 + 1);
      float2* row = lds + l * kXYMP;
      const float2 a = row[k], b = row[(kXYM - k) & (kXYM - 1)];
      const float2 E = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
      const float2 O = make_float2(0.5f * (a.y + b.y), 0.5f * (b.x - a.x));
      const float2 T = cmul(tw[k], O);
      row[k] = cadd(E, T);
      if (kXYM - k != k) row[kXYM - k] = make_float2(E.x - T.x, T.y - E.y);
    }
    __syncthreads();
    // this half's rows of my columns; column 0 carries DC + i Nyquist
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const float2* row = lds + (jy + 16 * t) * kXYMP;
      c0[t + 8 * h] = cl == 0 ? make_float2(row[0].x, row[kXYM].x) : row[cl];
      c1[t + 8 * h] = row[64 + cl];
    }
  }
  // y transforms, 64 columns per round
  auto yaddr = [cl](int i) { return i * kXYLP + cl; };
  xy_fft<kXYN>(c0, lds, jy, tw, 1, -1.f, yaddr);
  // Column 0 carries Z = F0 + i F128, F0 and F128 the transforms of the real
  // DC and Nyquist columns: F0(k) = (Z(k) + conj Z(-k)) / 2 and
  // F128(k) = (Z(k) - conj Z(-k)) / 2i.  Z goes through LDS so that the DC
  // value is stored with columns 1..63 and the Nyquist value with columns
  // 64..127 and the row padding (zeros): every 128-byte line of a row is
  // written by one store instruction (a partly written line reaches HBM as a
  // read-modify-write).
  if (cl == 0) {
#pragma unroll
    for (int t = 0; t < 16; ++t) dcn[elem_index<kXYN, 16, SY::R1>(jy, 0, t)] = c0[t];
  }
  __syncthreads();
  // the DC values of this wave's 16 rows, one per lane 0..15, handed to lane 0
  // row by row
  float2 sdc = make_float2(0.f, 0.f);
  if (cl < 16) {
    const int kk = elem_index<kXYN, 16, SY::R1>(jy, 0, cl);
    sdc = cscale(cadd(dcn[kk], cconj(dcn[(kXYN - kk) & (kXYN - 1)])), 0.5f);
  }
  // (lane indices from opaque() in every destination pass: offsets hoisted
  // out of the destination loop would hold 32 more registers)
  for (int d = 0; d < po.n; ++d) {
    const auto out = gmem(po.base(d, P.pitch));
    const int lane = opaque(threadIdx.x), cl_ = lane & 63, jy_ = lane >> 6;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int ky = elem_index<kXYN, 16, SY::R1>(jy_, 0, t);
      const float2 d0 = make_float2(__shfl(sdc.x, t), __shfl(sdc.y, t));
      stream_store(&out[ky * P.pitch + cl_], cl_ == 0 ? d0 : c0[t]);
    }
  }
  xy_fft<kXYN>(c1, lds, jy, tw, 1, -1.f, yaddr);
  for (int d = 0; d < po.n; ++d) {
    const auto out = gmem(po.base(d, P.pitch));
    const int lane = opaque(threadIdx.x), cl_ = lane & 63, jy_ = lane >> 6;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int ky = elem_index<kXYN, 16, SY::R1>(jy_, 0, t);
      stream_store(&out[ky * P.pitch + 64 + cl_], c1[t]);
    }
  }
  // the Nyquist column and the padding of this wave's 16 rows: 16 lanes per
  // row, four rows per store instruction (pitch - 128 = 16 columns)
  float2 nq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int t = 4 * q + (cl >> 4), c = cl & 15;
    const int ky = elem_index<kXYN, 16, SY::R1>(jy, 0, t);
    nq[q] = make_float2(0.f, 0.f);
    if (c == 0) {
      const float2 d = cscale(csub(dcn[ky], cconj(dcn[(kXYN - ky) & (kXYN - 1)])), 0.5f);
      nq[q] = make_float2(d.y, -d.x);
    }
  }
  for (int d = 0; d < po.n; ++d) {
    const auto out = gmem(po.base(d, P.pitch));
    const int lane = opaque(threadIdx.x), cl_ = lane & 63, jy_ = lane >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = 4 * q + (cl_ >> 4), c = cl_ & 15;
      const int ky = elem_index<kXYN, 16, SY::R1>(jy_, 0, t);
      stream_store(&out[ky * P.pitch + kXYM + c], nq[q]);
    }
  }
}

struct XYInvParams {
  ConvGeom g;
  const float2* spec;
  EpilogueArgs e;
  const float2* tw;
  int pitch;
};

template <int EPI>
__device__ __forceinline__ void xy_inverse_body(const XYInvParams& P, int bx) {
  __shared__ float2 lds[kXYLds];
  __shared__ float2 nyq[kXYN];
  __shared__ float2 tw[kXYN];
  __shared__ double red_sum[kXYThreads / 64];
  __shared__ float red_max[kXYThreads / 64];
  const int tid = threadIdx.x;
  const ConvGeom& g = P.g;
  const EpilogueArgs& e = P.e;
  const double inv_lambda = rl_inv_lambda(e.lambda);
  const int zo = bx;                 // output plane inside the effective box
  const int z = g.eff_local.z + zo;  // plane of the FFT box
  if (tid < kXYN) tw[tid] = gmem(P.tw)[tid];
  const int cl = tid & 63, jy = tid >> 6;
  const auto in = gmem(P.spec) + (int64_t)z * kXYN * P.pitch;
  auto yaddr = [cl](int i) { return i * kXYLP + cl; };
  float2 c0[16], c1[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const auto row = in + elem_index<kXYN, 16, SY::R0>(jy, 0, t) * P.pitch;
    c0[t] = stream_load(&row[cl]);
    c1[t] = stream_load(&row[64 + cl]);
  }
  if (tid < kXYN) nyq[tid] = stream_load(&in[tid * P.pitch + kXYM]);
  __syncthreads();
  if (cl == 0) {
#pragma unroll
    for (int t = 0; t < 16; ++t) {  // column 0 carries DC + i Nyquist
      const float2 b = nyq[elem_index<kXYN, 16, SY::R0>(jy, 0, t)];
      c0[t] = make_float2(c0[t].x - b.y, c0[t].y + b.x);
    }
  }
  // y transforms of both column sets.  After the two LDS exchanges the y tile
  // is free, so the last radix pass of each set writes its rows of the first
  // half straight into the x tile: at most 48 plane values per thread are live.
  auto put_rows = [&](const float2 (&c)[16], int col, int h) {
    const int lane = opaque(threadIdx.x);
    const int cl = lane & 63, jy = lane >> 6;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      float2* row = lds + (jy + 16 * t) * kXYMP;
      const float2 a = c[t + 8 * h];
      if (col == 0 && cl == 0) {  // DC + i Nyquist
        row[0] = make_float2(a.x, 0.f);
        row[kXYM] = make_float2(a.y, 0.f);
      } else {
        row[col + cl] = a;
      }
    }
  };
  auto last_pass = [&](float2 (&c)[16]) {
#pragma unroll
    for (int t = 1; t < 16; ++t) {
      float2 w = tw[t * jy];  // butterfly jy of the second radix-16 pass: W256^(t jy)
      w.y = -w.y;             // inverse direction
      c[t] = cmul(c[t], w);
      __builtin_amdgcn_sched_barrier(0);  // one twiddle live at a time
    }
    butterfly<16>(c, 1.f);
  };
  butterfly<16>(c0, 1.f);
  __builtin_amdgcn_sched_barrier(0);
  butterfly<16>(c1, 1.f);
  __syncthreads();
  store_pass<kXYN, 16, 16, 1>(c0, lds, jy, yaddr);
  __syncthreads();
  load_pass<kXYN, 16, 16>(c0, lds, jy, yaddr);
  __syncthreads();
  store_pass<kXYN, 16, 16, 1>(c1, lds, jy, yaddr);
  __syncthreads();
  load_pass<kXYN, 16, 16>(c1, lds, jy, yaddr);
  __syncthreads();
  last_pass(c0);
  put_rows(c0, 0, 0);
  last_pass(c1);
  put_rows(c1, 64, 0);
  double my_sum = 0.0;
  float my_max = -1.f;
  const int ylo = g.eff_local.y, yhi = g.eff_local.y + g.eff_size.y;
  const int nx = g.eff_size.x;
  const float* fbuf = reinterpret_cast<const float*>(lds);
  // images hold fewer than 2^31 voxels (checked by the host): 32-bit offsets
  auto out_index = [&](int x, int y, int zo_) -> int {
    return (g.eff_offset.x + x) + g.dst.x * ((g.eff_offset.y + y - ylo) + g.dst.y * (g.eff_offset.z + zo_));
  };
  // A thread's epilogue voxels in half h are x = lane & 255 and the 32 rows
  // rsub + 4 j, j < 32.  Their inputs are loaded in groups of R rows
  // (in0: the view or the estimate, in1: the weight), as early as registers
  // allow: the waves' loads are then in flight during the x transform, not
  // waited for one small batch after another.
  // The views and weights are read once per iteration: streaming loads keep
  // them from displacing the estimate and ratio lines that the forward gathers
  // of neighbouring blocks (overlapping windows) read again.
  auto load_rows = [&](int h, int j0, auto& in0, auto& in1) {
    constexpr int R = sizeof(in0) / sizeof(float);
    const int lane = opaque(threadIdx.x);
    const int x = lane & 255, rsub = lane >> 8;
    const bool x_live = x < nx;
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int y = h * kXYHalf + rsub + 4 * (j0 + i);
      in0[i] = 0.f;
      in1[i] = 1.f;
      if (x_live && y >= ylo && y < yhi) {
        const int gi = out_index(x, y, zo);
        if (EPI == kEpiQuotient || EPI == kEpiResidual) {
          in0[i] = __builtin_nontemporal_load(&gmem(e.img)[gi]);
          if (EPI == kEpiResidual && e.weight_mode == kWeightArray) in1[i] = __builtin_nontemporal_load(&gmem(e.weight)[gi]);
        } else if (EPI == kEpiUpdate) {
          in0[i] = gmem(e.dst)[gi];
          in1[i] = e.weight_mode != kWeightNone ? __builtin_nontemporal_load(&gmem(e.weight)[gi]) : e.wconst;
        }
      }
    }
  };
  auto run_rows = [&](int h, int j0, const auto& in0, const auto& in1) {
    constexpr int R = sizeof(in0) / sizeof(float);
    const int lane = opaque(threadIdx.x);
    const int x = lane & 255, rsub = lane >> 8;
    const bool x_live = x < nx;
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int yl = rsub + 4 * (j0 + i);
      const int y = h * kXYHalf + yl;
      if (!x_live || y < ylo || y >= yhi) continue;
      const int gi = out_index(x, y, zo);
      const float val = fbuf[yl * 2 * kXYMP + g.eff_local.x + x];
      if (EPI == kEpiStore) {
        gmem(e.dst)[gi] = val;
      } else if (EPI == kEpiQuotient) {
        gmem(e.dst)[gi] = rl_quotient(val, in0[i]);
      } else if (EPI == kEpiResidual) {
        if (val > 0.f) {
          my_sum += nll_term(in1[i], val, in0[i]);
          gmem(e.dst)[gi] = nll_residual(in1[i], val, in0[i]);
        } else {
          gmem(e.dst)[gi] = 0.f;
        }
      } else {
        float w = in1[i];
        if (e.weight_mode == kWeightVirtual) {
          const double r = (double)in1[i] / (double)gmem(e.wsum)[gi];
          w = (float)java_min(1.0, r * e.osem);
        }
        const float last = in0[i];
        const float next = rl_next_value(last, val, w, e.lambda, inv_lambda);
        gmem(e.dst)[gi] = next;
        const float change = fabsf(next - last);
        my_sum += (double)change;
        my_max = java_maxf(my_max, change);
      }
    }
  };
  // pack X[k], X[M - k] into Z[k] = E + i O, then the 128-point row transforms
  auto rows_c2r = [&]() {
    for (int idx = opaque(threadIdx.x); idx < kXYHalf * (kXYM / 2 + 1); idx += kXYThreads) {
      const int l = idx / (kXYM / 2 + 1), k = idx % (kXYM / 2 + 1);
      float2* row = lds + l * kXYMP;
      const float2 a = row[k], b = row[kXYM - k];
      const float2 A = make_float2(a.x + b.x, a.y - b.y);
      const float2 B = make_float2(a.x - b.x, a.y + b.y);
      const float2 C = cmul(B, cconj(tw[k]));
      if (k != 0 && kXYM - k != k) row[kXYM - k] = make_float2(A.x + C.y, C.x - A.y);
      row[k] = make_float2(A.x - C.y, A.y + C.x);
    }
    __syncthreads();
    xy_rows_fft(lds, tw, 1.f);
    __syncthreads();
  };
  // The 32 rows of a half in chunks of C: the first chunk's loads are in
  // flight during the half's transform, and each later chunk is loaded while
  // the one before it is computed and stored.  The first half still holds the
  // second half's plane values in registers, the second half holds none;
  // epilogues with two inputs per voxel take chunks half as long.
  constexpr int kInputs = (EPI == kEpiUpdate || EPI == kEpiResidual) ? 2 : 1;
  auto half_epilogue = [&](int h, auto chunk) {
    constexpr int C = decltype(chunk)::value;
    float cur0[C], cur1[C];
    load_rows(h, 0, cur0, cur1);
    __syncthreads();
    rows_c2r();
#pragma unroll
    for (int j0 = 0; j0 < 32; j0 += C) {
      float nxt0[C], nxt1[C];
      if (j0 + C < 32) load_rows(h, j0 + C, nxt0, nxt1);
      run_rows(h, j0, cur0, cur1);
#pragma unroll
      for (int i = 0; i < C; ++i) {
        cur0[i] = nxt0[i];
        cur1[i] = nxt1[i];
      }
    }
  };
  // (chunk sizes: the largest that compile without register spills)
  constexpr int kChunk0 = EPI == kEpiResidual ? 2 : kInputs == 2 ? 4 : 8;
  constexpr int kChunk1 = EPI == kEpiResidual ? 16 : kInputs == 2 ? 16 : 32;
  half_epilogue(0, std::integral_constant<int, kChunk0>{});
  __syncthreads();  // the first half's readers of lds are done
  put_rows(c0, 0, 1);
  put_rows(c1, 64, 1);
  half_epilogue(1, std::integral_constant<int, kChunk1>{});
  if (EPI == kEpiUpdate || EPI == kEpiResidual) {
    for (int o = 32; o > 0; o >>= 1) {
      my_sum += __shfl_xor(my_sum, o);
      my_max = java_maxf(my_max, __shfl_xor(my_max, o));
    }
    const int wid = tid / 64;
    if ((tid & 63) == 0) {
      red_sum[wid] = my_sum;
      red_max[wid] = my_max;
    }
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      float m = -1.f;
      for (int i = 0; i < kXYThreads / 64; ++i) {
        s += red_sum[i];
        m = java_maxf(m, red_max[i]);
      }
      gmem(e.part_sum)[bx] = s;
      gmem(e.part_max)[bx] = m;
    }
  }
}

__device__ float2* PlaneOut::base(int d, int pitch) const {
  if (!task) return single;
  const XYFwdParams* it = static_cast<const XYFwdParams*>(items);
  return it[task->dest[d][0]].spec + (int64_t)task->dest[d][1] * kXYN * pitch;
}

template <int MODE>
__global__ __launch_bounds__(kXYThreads) void k_xy_forward(XYFwdParams P) {
  PlaneOut po{};
  po.n = 1;
  po.single = P.spec + (int64_t)blockIdx.x * kXYN * P.pitch;
  xy_forward_body<MODE>(P, blockIdx.x, po);
}
// tasks (optional): per workgroup the plane it transforms and the spectra it
// stores the result into, in the order the host chose; without tasks,
// workgroup -> (item, plane) through the prefix array and one store each.
template <int MODE>
__global__ __launch_bounds__(kXYThreads) void k_xy_forward_batch(const XYFwdParams* items, const int* prefix,
                                                                const int* gx, int nitems, const PlaneTask* tasks) {
  PlaneOut po{};
  if (tasks) {
    const PlaneTask* T = tasks + blockIdx.x;
    const auto item_params = items[T->item];  // private copy: the kernel's stores cannot alias it
    po.n = T->ndest;
    po.task = T;
    po.items = items;
    xy_forward_body<MODE>(item_params, T->plane, po);
  } else {
    int item, bx, by;
    batch_locate(prefix, gx, nitems, item, bx, by);
    const auto item_params = items[item];
    po.n = 1;
    po.single = item_params.spec + (int64_t)bx * kXYN * item_params.pitch;
    xy_forward_body<MODE>(item_params, bx, po);
  }
}
template <int EPI>
__global__ __launch_bounds__(kXYThreads) void k_xy_inverse(XYInvParams P) {
  xy_inverse_body<EPI>(P, blockIdx.x);
}
template <int EPI>
__global__ __launch_bounds__(kXYThreads) void k_xy_inverse_batch(const XYInvParams* items, const int* prefix,
                                                                const int* gx, int nitems) {
  int item, bx, by;
  batch_locate(prefix, gx, nitems, item, bx, by);
  const auto item_params = items[item];  // private copy: the kernel's stores cannot alias it
  xy_inverse_body<EPI>(item_params, bx);
}

// ---------------------------------------------------------------------------
// host side

namespace {
// MVD_DISABLE_FUSED_XY=1 keeps the separate x and y passes (for comparisons).
const bool g_xy_disabled = [] {
  const char* v = std::getenv("MVD_DISABLE_FUSED_XY");
  return v && v[0] == '1';
}();

bool xy_gather_ok(int gather) { return gather == kGatherMirror || gather == kGatherOne || gather == kGatherZero; }
}  // namespace

// (the forward's padding stores assume rows of 144 = 128 + 16 complex values)
bool xy_fused(const ShapePlan& sp) {
  return !g_xy_disabled && sp.q.x == kXYN && sp.q.y == kXYN && sp.pitch == kXYM + 16;
}

hipError_t launch_xy_forward(const ShapePlan& sp, const ConvGeom& g, const float* src, int gather, float2* spec,
                             hipStream_t stream) {
  if (!xy_fused(sp) || !xy_gather_ok(gather) || g.q.x != kXYN || g.q.y != kXYN) return hipErrorInvalidValue;
  XYFwdParams P{g, src, spec, sp.tw_x, sp.pitch};
  (void)hipGetLastError();
  const dim3 grid(g.q.z), block(kXYThreads);
  switch (gather) {
    case kGatherMirror: hipLaunchKernelGGL(k_xy_forward<kGatherMirror>, grid, block, 0, stream, P); break;
    case kGatherOne: hipLaunchKernelGGL(k_xy_forward<kGatherOne>, grid, block, 0, stream, P); break;
    default: hipLaunchKernelGGL(k_xy_forward<kGatherZero>, grid, block, 0, stream, P); break;
  }
  return hipGetLastError();
}

hipError_t launch_xy_inverse(const ShapePlan& sp, const ConvGeom& g, const float2* spec, const EpilogueArgs& e,
                             hipStream_t stream) {
  if (!xy_fused(sp) || g.q.x != kXYN || g.q.y != kXYN) return hipErrorInvalidValue;
  if (g.eff_size.z <= 0) return hipSuccess;
  XYInvParams P{g, spec, e, sp.tw_x, sp.pitch};
  (void)hipGetLastError();
  const dim3 grid(g.eff_size.z), block(kXYThreads);
  switch (e.mode) {
    case kEpiStore: hipLaunchKernelGGL(k_xy_inverse<kEpiStore>, grid, block, 0, stream, P); break;
    case kEpiQuotient: hipLaunchKernelGGL(k_xy_inverse<kEpiQuotient>, grid, block, 0, stream, P); break;
    case kEpiUpdate: hipLaunchKernelGGL(k_xy_inverse<kEpiUpdate>, grid, block, 0, stream, P); break;
    case kEpiResidual: hipLaunchKernelGGL(k_xy_inverse<kEpiResidual>, grid, block, 0, stream, P); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int build_xy_params(const PassDesc& d, void* out, int* gx, int* gy, int* cls) {
  const ShapePlan& sp = *d.sp;
  *gy = 1;
  if (d.kind == kPassXForward) {
    static_assert(sizeof(XYFwdParams) <= kMaxParamBytes, "parameter block");
    XYFwdParams P{d.g, d.src, d.spec, sp.tw_x, sp.pitch};
    *gx = d.g.q.z;
    *cls = d.gather;
    std::memcpy(out, &P, sizeof(P));
    return sizeof(P);
  }
  static_assert(sizeof(XYInvParams) <= kMaxParamBytes, "parameter block");
  XYInvParams P{d.g, d.spec, d.e, sp.tw_x, sp.pitch};
  *gx = d.g.eff_size.z;
  *cls = d.e.mode;
  std::memcpy(out, &P, sizeof(P));
  return sizeof(P);
}

hipError_t launch_batch_xy(int kind, int cls, const void* items, const int* prefix, const int* gx, int nitems,
                           int nwg, hipStream_t stream, const PlaneTask* tasks) {
  if (nwg <= 0) return hipSuccess;
  (void)hipGetLastError();
  const dim3 grid(nwg), block(kXYThreads);
  if (kind == kPassXForward) {
    const XYFwdParams* it = static_cast<const XYFwdParams*>(items);
    switch (cls) {
      case kGatherMirror:
        hipLaunchKernelGGL(k_xy_forward_batch<kGatherMirror>, grid, block, 0, stream, it, prefix, gx, nitems, tasks);
        break;
      case kGatherOne:
        hipLaunchKernelGGL(k_xy_forward_batch<kGatherOne>, grid, block, 0, stream, it, prefix, gx, nitems, tasks);
        break;
      case kGatherZero:
        hipLaunchKernelGGL(k_xy_forward_batch<kGatherZero>, grid, block, 0, stream, it, prefix, gx, nitems, tasks);
        break;
      default: return hipErrorInvalidValue;
    }
  } else if (kind == kPassXInverse) {
    const XYInvParams* it = static_cast<const XYInvParams*>(items);
    switch (cls) {
      case kEpiStore:
        hipLaunchKernelGGL(k_xy_inverse_batch<kEpiStore>, grid, block, 0, stream, it, prefix, gx, nitems);
        break;
      case kEpiQuotient:
        hipLaunchKernelGGL(k_xy_inverse_batch<kEpiQuotient>, grid, block, 0, stream, it, prefix, gx, nitems);
        break;
      case kEpiUpdate:
        hipLaunchKernelGGL(k_xy_inverse_batch<kEpiUpdate>, grid, block, 0, stream, it, prefix, gx, nitems);
        break;
      case kEpiResidual:
        hipLaunchKernelGGL(k_xy_inverse_batch<kEpiResidual>, grid, block, 0, stream, it, prefix, gx, nitems);
        break;
      default: return hipErrorInvalidValue;
    }
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mvd
