// Register-resident FFT passes for lengths 64, 128, 144, 192, 256, 288, 384 and 512:
// the y pass, the fused z pass (forward, kernel product, inverse) and the
// real-to-complex / complex-to-real x passes with their gather and epilogue.
// The butterfly and LDS exchange building blocks live in fft_regs.h.
#include <cstdlib>
#include <cstring>

#include "fft_regs.h"
#include "mvd_internal.h"
#include "pointwise_math.h"

namespace mvd {

// Workgroup order of the y and z line kernels.  A workgroup holding L lines
// of 8 bytes covers L * 8 bytes of every row it touches; at 64 bytes (the
// 512-point lengths) two workgroups share each 128-byte line.  Those kernels,
// and the batched z kernel at every length, run their workgroups in the
// order of xcd_chunk: the two halves of a line are read by one XCD (one HBM
// fetch instead of two) and every XCD's accesses spread over its L2 channels.

// lines per workgroup of the kernels of length N
template <int N>
static constexpr int lines_per_wg() {
  return Spec<N>::NT / (N / Spec<N>::E);
}

template <int L, bool kAlways>
constexpr bool chunked_order() {
  return kAlways || L * sizeof(float2) < 128;
}

// (column group, row) of a 2-D grid, column groups fastest
template <int L, bool kAlways = false>
__device__ __forceinline__ void line_group(int& bx, int& by) {
  if constexpr (chunked_order<L, kAlways>()) {
    const int w = xcd_chunk(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    bx = w % gridDim.x;
    by = w / gridDim.x;
  } else {
    bx = blockIdx.x;
    by = blockIdx.y;
  }
}

// logical workgroup of a 1-D (batched) grid
template <int L>
__device__ __forceinline__ int line_wg() {
  if constexpr (chunked_order<L, false>()) return xcd_chunk(blockIdx.x, gridDim.x);
  return blockIdx.x;
}

// Column groups of a y pass grid: with 64-byte groups they cover the row
// pitch, so that the padding lanes complete the row's last 128-byte line with
// zeros.  (The z passes cover hx: an extra group per row costs them more than
// the partial line.)
static int y_line_groups(const ShapePlan& sp, int L) {
  return L * (int)sizeof(float2) < 128 ? sp.pitch / L : (sp.hx + L - 1) / L;
}

static int z_line_groups(const ShapePlan& sp, int L) { return (sp.hx + L - 1) / L; }

// ---------------------------------------------------------------------------
// y pass: complex FFT along y of L consecutive kx columns for one z.

struct YFastParams {
  float2* spec;
  const float2* tw;
  int hx, pitch, zlo, ylo, yhi;
  float dir;
  // input rows outside [iylo, iyhi) of planes outside [izlo, izhi) read as zero
  int iylo = 0, iyhi = 1 << 30, izlo = 0, izhi = 1 << 30;
};

template <int N>
__device__ __forceinline__ void y_fast_body(const YFastParams& P, int bx, int by) {
  using S = Spec<N>;
  constexpr int E = S::E, T = N / E, NT = S::NT, L = NT / T;
  constexpr bool kStream = L * sizeof(float2) % 128 == 0;  // whole lines per workgroup
  __shared__ float2 tw[N];
  __shared__ float2 lds[line_slots<N, L>()];
  const int line = threadIdx.x % L, jt = threadIdx.x / L;
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = gmem(P.tw)[i];
  // lanes of the row padding [hx, pitch) transform zeros and store them, so
  // that every 128-byte line of a row is written whole
  const int col = bx * L + line;
  const bool valid = col < P.hx, store_col = col < P.pitch;
  const int z = P.zlo + by;
  const auto base = gmem(P.spec) + (int64_t)z * N * P.pitch + col;
  const bool plane_in = z >= P.izlo && z < P.izhi;
  float2 v[E];
#pragma unroll
  for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
    for (int t = 0; t < S::R0; ++t) {
      const int y = elem_index<N, E, S::R0>(jt, b, t);
      const bool in = valid && plane_in && y >= P.iylo && y < P.iyhi;
      v[b * S::R0 + t] = in ? maybe_stream_load<kStream>(&base[y * P.pitch]) : make_float2(0.f, 0.f);
    }
  auto addr = [line](int i) { return line_slot<N, L>(i, line); };
  fft_forward_order<N>(v, lds, jt, tw, 1, P.dir, addr);
  constexpr int RL = last_radix<N>();
  if (store_col) {
#pragma unroll
    for (int b = 0; b < E / RL; ++b)
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int y = elem_index<N, E, RL>(jt, b, t);
        if (y >= P.ylo && y < P.yhi) maybe_stream_store<kStream>(&base[y * P.pitch], v[b * RL + t]);
      }
  }
}

template <int N>
__global__ __launch_bounds__(Spec<N>::NT) void k_y_fast(YFastParams P) {
  int bx, by;
  line_group<lines_per_wg<N>()>(bx, by);
  y_fast_body<N>(P, bx, by);
}

// tasks (optional): one (item, plane) per wpt workgroups -- only the planes
// listed are transformed (the others are identical to a listed plane and read
// from it by the z pass); wpt = the item's column groups.
template <int N>
__global__ __launch_bounds__(Spec<N>::NT) void k_y_fast_batch(const YFastParams* items, const int* prefix, const int* gx,
                                                   int nitems, const int2* tasks, int wpt) {
  int item, bx, by;
  const int wg = line_wg<lines_per_wg<N>()>();
  if (tasks) {
    const int2 t = tasks[wg / wpt];
    item = t.x;
    bx = wg % wpt;
    by = t.y - items[item].zlo;
  } else {
    batch_locate(prefix, gx, nitems, item, bx, by, wg);
  }
  const auto item_params = items[item];  // private copy: the kernel's stores cannot alias it
  y_fast_body<N>(item_params, bx, by);
}

// ---------------------------------------------------------------------------
// z pass: forward FFT along z, product with the kernel spectrum, inverse FFT,
// write the z rows [zlo, zhi).  Without a kernel spectrum: forward + scale.

struct ZFastParams {
  float2* spec;
  const float2* kspec;
  const float2* tw;
  float scale, dir;
  int hx, pitch, qy, zlo, zhi;
  // input planes through a table (lines kernel only): plane z of the item is
  // read from plane ztab[z] of the arena of spectra, which may hold it for
  // another item (identical window planes); outputs stay in spec
  const unsigned short* ztab;
  const float2* arena;
  unsigned arena_bytes;
  // with kspec (single launches only): the forward spectrum before the
  // product, all rows (nullable)
  float2* aux = nullptr;
  // with kspec (single launches only): the lines' spectrum already
  // transformed along z, in place of spec (nullable)
  const float2* pre = nullptr;
};

template <int N>
__device__ __forceinline__ void z_fast_body(const ZFastParams& P, int bx, int by) {
  using S = Spec<N>;
  constexpr int E = S::E, T = N / E, NT = S::NT, L = NT / T;
  constexpr bool kStream = L * sizeof(float2) % 128 == 0;  // whole lines per workgroup
  __shared__ float2 tw[N];
  __shared__ float2 lds[line_slots<N, L>()];
  const int line = threadIdx.x % L, jt = threadIdx.x / L;
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = gmem(P.tw)[i];
  const int col = bx * L + line;
  const bool valid = col < P.hx;
  // spectra hold fewer than 2^31 elements (lengths up to 1024): 32-bit offsets
  const int plane = P.qy * P.pitch;
  const int off = by * P.pitch + col;
  const bool pre = P.kspec && P.pre;
  constexpr int RL = last_radix<N>();
  float2 v[E];
  if (pre) {
    // already transformed: the lines in the order the forward transform leaves them
#pragma unroll
    for (int b = 0; b < E / RL; ++b)
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int z = elem_index<N, E, RL>(jt, b, t);
        v[b * RL + t] = make_float2(0.f, 0.f);
        if (valid) v[b * RL + t] = gmem(P.pre)[z * plane + off];
      }
  } else {
#pragma unroll
    for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
      for (int t = 0; t < S::R0; ++t) {
        const int z = elem_index<N, E, S::R0>(jt, b, t);
        v[b * S::R0 + t] = make_float2(0.f, 0.f);
        if (valid) v[b * S::R0 + t] = maybe_stream_load<kStream>(&gmem(P.spec)[z * plane + off]);
      }
  }
  auto addr = [line](int i) { return line_slot<N, L>(i, line); };
  // the kernel spectrum is loaded with the line, in the order the forward
  // transform leaves its outputs, so its latency hides behind the transform
  float2 kv[E];
  if (P.kspec) {
#pragma unroll
    for (int b = 0; b < E / RL; ++b)
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int z = elem_index<N, E, RL>(jt, b, t);
        kv[b * RL + t] = make_float2(0.f, 0.f);
        if (valid) kv[b * RL + t] = gmem(P.kspec)[z * plane + off];
      }
  }
  if (!pre) fft_forward_order<N>(v, lds, jt, tw, 1, P.kspec ? -1.f : P.dir, addr);
  if (P.kspec) {
    if (P.aux && valid) {
#pragma unroll
      for (int b = 0; b < E / RL; ++b)
#pragma unroll
        for (int t = 0; t < RL; ++t) gmem(P.aux)[elem_index<N, E, RL>(jt, b, t) * plane + off] = v[b * RL + t];
    }
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = cmul(v[i], kv[i]);
    fft_reverse_order<N>(v, lds, jt, tw, 1, 1.f, addr);
    if (valid) {
#pragma unroll
      for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
        for (int t = 0; t < S::R0; ++t) {
          const int z = elem_index<N, E, S::R0>(jt, b, t);
          if (z >= P.zlo && z < P.zhi) maybe_stream_store<kStream>(&gmem(P.spec)[z * plane + off], v[b * S::R0 + t]);
        }
    }
  } else if (valid) {
#pragma unroll
    for (int b = 0; b < E / RL; ++b)
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int z = elem_index<N, E, RL>(jt, b, t);
        maybe_stream_store<kStream>(&gmem(P.spec)[z * plane + off], cscale(v[b * RL + t], P.scale));
      }
  }
}

// Many blocks convolved with one kernel spectrum (the blocks of one view): a
// workgroup owns one set of lines (bx, by) and runs it for every item, so the
// kernel spectrum of its lines is read once instead of once per block.  The
// lines of the next item are loaded while the current one is transformed (the
// items' spectra are distinct buffers).  All items share hx, qy, pitch, the
// twiddles and kspec, and every spectrum spans fewer than 2^31 bytes (the host
// checks): memory is accessed through buffer resources with 32-bit offsets,
// and stores outside [zlo, zhi) are dropped by the range check rather than
// branched around, so the wait for the prefetch does not depend on how many
// stores a lane made.
//
// TAB: the items' input planes come through their plane tables (the
// launch's tables, at most ztab_entries(N) entries, are copied to LDS first);
// every table entry addresses a plane that no item of the launch overwrites,
// so the items can run in any order.  TAB = 1 addresses the arena of spectra
// with 32-bit buffer offsets (fewer than 2^31 bytes), TAB = 2 with 64-bit
// addresses.
template <int N, int TAB>
__global__ __launch_bounds__(Spec<N>::NT, N == 512 || N == 288 || N == 192 || N == 384 ? 2 : 3) void k_z_fast_lines(const ZFastParams* __restrict__ items, int nitems) {
  using S = Spec<N>;
  constexpr int E = S::E, T = N / E, NT = S::NT, L = NT / T;
  constexpr bool kStream = L * sizeof(float2) % 128 == 0;  // whole lines per workgroup
  constexpr int RL = last_radix<N>();
  __shared__ float2 tw[N];
  __shared__ float2 lds[line_slots<N, L>()];
  constexpr bool TABLE = TAB != 0;
  __shared__ unsigned short ztab[TABLE ? ztab_entries(N) : 1];
  const ZFastParams P0 = items[0];
  const int line = threadIdx.x % L, jt = threadIdx.x / L;
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = gmem(P0.tw)[i];
  if constexpr (TABLE) {
    for (int it = 0; it < nitems; ++it) {
      const unsigned short* src = items[it].ztab;
      for (int i = threadIdx.x; i < N; i += NT) ztab[it * N + i] = gmem(src)[i];
    }
    __syncthreads();
  }
  int bx, by;
  line_group<L, true>(bx, by);
  const int col = bx * L + line;
  const uint32_t plane = (uint32_t)(P0.qy * P0.pitch) * 8u;  // bytes
  const uint32_t bytes = plane * (uint32_t)N;
  // Lanes of the padding columns [hx, pitch) transform their own (unused)
  // column and store it, so that every 128-byte line of a row is written whole
  // (a partly written line reaches HBM as a read-modify-write).  Lanes past the
  // pitch read column hx - 1 and store nothing.  A line's transform never
  // mixes with other lines.
  const uint32_t row = (uint32_t)(by * P0.pitch) * 8u;  // uniform: the SGPR offset
  const bool store_col = col < P0.pitch;
  const uint32_t off = (uint32_t)(store_col ? col : P0.hx - 1) * 8u;
  auto addr = [line](int i) { return line_slot<N, L>(i, line); };
  auto load_line = [&](int it, const float2* spec, float2(&v)[E]) {
    if constexpr (TAB == 2) {
      // 64-bit: arena + plane index * plane + row + column
      const auto base = gmem(P0.arena) + (size_t)(by * P0.pitch) + (off >> 3);
      const size_t plane_elems = (size_t)P0.qy * P0.pitch;
#pragma unroll
      for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
        for (int t = 0; t < S::R0; ++t) {
          const int z = elem_index<N, E, S::R0>(jt, b, t);
          const auto ptr = base + (size_t)ztab[it * N + z] * plane_elems;
          if constexpr (kStream) v[b * S::R0 + t] = stream_load(ptr);
          else v[b * S::R0 + t] = *ptr;
        }
    } else {
      const BufRes r = TABLE ? buf_res(P0.arena, P0.arena_bytes) : buf_res(spec, bytes);
#pragma unroll
      for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
        for (int t = 0; t < S::R0; ++t) {
          const uint32_t z = elem_index<N, E, S::R0>(jt, b, t);
          const uint32_t zp = TABLE ? (uint32_t)ztab[it * N + z] : z;
          v[b * S::R0 + t] = kStream ? buf_load2_nt(r, zp * plane + off, row) : buf_load2(r, zp * plane + off, row);
        }
    }
  };
  float2 nxt[E];
  load_line(0, P0.spec, nxt);
  float2 kv[E];
  {
    const BufRes r = buf_res(P0.kspec, bytes);
#pragma unroll
    for (int b = 0; b < E / RL; ++b)
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const uint32_t z = elem_index<N, E, RL>(jt, b, t);
        kv[b * RL + t] = buf_load2(r, z * plane + off, row);
      }
  }
  // wait for the kernel spectrum here, ahead of the loop and its prefetches
#pragma unroll
  for (int i = 0; i < E; ++i) asm volatile("" : "+v"(kv[i].x), "+v"(kv[i].y));
#pragma unroll 1
  for (int it = 0; it < nitems; ++it) {
    // twiddles re-read from LDS in every item: hoisting them out of the loop
    // would hold ~60 more registers and cost occupancy
    const float2* twl = tw + opaque(0);
    const BufRes out = buf_res(items[it].spec, bytes);
    const int zlo = items[it].zlo, zhi = items[it].zhi;
    float2 v[E];
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = nxt[i];
    if (it + 1 < nitems) load_line(it + 1, items[it + 1].spec, nxt);
    fft_forward_order<N>(v, lds, jt, twl, 1, -1.f, addr);
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = cmul(v[i], kv[i]);
    fft_reverse_order<N>(v, lds, jt, twl, 1, 1.f, addr);
#pragma unroll
    for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
      for (int t = 0; t < S::R0; ++t) {
        const int z = elem_index<N, E, S::R0>(jt, b, t);
        const bool keep = store_col && z >= zlo && z < zhi;
        const uint32_t o = keep ? (uint32_t)z * plane + off : kNoLane;
        if constexpr (kStream) buf_store2_nt(out, o, v[b * S::R0 + t], row);
        else buf_store2(out, o, v[b * S::R0 + t], row);
      }
  }
}

hipError_t launch_z_lines(int n, const void* items, int nitems, int gx, int gy, hipStream_t stream, int table) {
  if (nitems <= 0 || gx <= 0 || gy <= 0) return hipSuccess;
  if (table < 0 || table > 2 || (table && nitems * n > ztab_entries(n))) return hipErrorInvalidValue;
  (void)hipGetLastError();
  const ZFastParams* it = static_cast<const ZFastParams*>(items);
  const dim3 grid(gx, gy);
#define MVD_Z_LINES(NN)                                                                                         \
  if (table == 2) hipLaunchKernelGGL((k_z_fast_lines<NN, 2>), grid, dim3(Spec<NN>::NT), 0, stream, it, nitems);  \
  else if (table == 1) hipLaunchKernelGGL((k_z_fast_lines<NN, 1>), grid, dim3(Spec<NN>::NT), 0, stream, it, nitems); \
  else hipLaunchKernelGGL((k_z_fast_lines<NN, 0>), grid, dim3(Spec<NN>::NT), 0, stream, it, nitems);
  switch (n) {
    case 64: MVD_Z_LINES(64) break;
    case 128: MVD_Z_LINES(128) break;
    case 144: MVD_Z_LINES(144) break;
    case 256: MVD_Z_LINES(256) break;
    case 192: MVD_Z_LINES(192) break;
    case 288: MVD_Z_LINES(288) break;
    case 384:  // plane tables only (see the launch grouping)
      if (table == 2) hipLaunchKernelGGL((k_z_fast_lines<384, 2>), grid, dim3(Spec<384>::NT), 0, stream, it, nitems);
      else if (table == 1) hipLaunchKernelGGL((k_z_fast_lines<384, 1>), grid, dim3(Spec<384>::NT), 0, stream, it, nitems);
      else return hipErrorInvalidValue;
      break;
    case 512: MVD_Z_LINES(512) break;
    default: return hipErrorInvalidValue;
  }
#undef MVD_Z_LINES
  return hipGetLastError();
}

template <int N>
__global__ __launch_bounds__(Spec<N>::NT) void k_z_fast(ZFastParams P) {
  int bx, by;
  line_group<lines_per_wg<N>()>(bx, by);
  z_fast_body<N>(P, bx, by);
}

// z pass of the likelihood gradient on registers (see k_z_grad in
// conv_kernels.hip for the generic lengths): forward transform of the
// residual line R^; the correlation conj(R^) W^ times the lag-shift phase is
// accumulated into corr (mode 0) or transformed back and rows [0, kz) stored
// (mode 1); the adjoint R^ conj(P^) is transformed back and rows [zlo, zhi)
// stored in place.
struct ZGradFastParams {
  float2* spec;
  const float2* what;
  const float2* khat;
  float2* corr;
  const float2* tw;
  const float2* twx;
  const float2* twy;
  int mode, first;
  Int3 shift, q;
  int hx, pitch, zlo, zhi, kz;
  const double* rsum;  // R^(0) from the residual's double sum (nullable)
  float2* adj;         // the adjoint rows' destination (nullable: spec, in place)
  int adj_add;         // add the adjoint rows to adj instead of storing them
};

// At 288 points the kernel holds the residual line, the correlation line and
// their loads in about 220 VGPRs, which allows two waves per SIMD; asking for
// three (168 VGPRs, about 30 of them spilled to scratch) made the gradient of
// four 256^3 views (288-point window) 5 % faster -- the line loads' latency is
// what bounds it.
template <int N>
__global__ __launch_bounds__(Spec<N>::NT) __attribute__((amdgpu_waves_per_eu(N == 288 ? 3 : 1)))
void k_z_fast_grad(ZGradFastParams P) {
  using S = Spec<N>;
  constexpr int E = S::E, T = N / E, NT = S::NT, L = NT / T;
  constexpr int RL = last_radix<N>();
  __shared__ float2 tw[N];
  __shared__ float2 lds[line_slots<N, L>()];
  const int line = threadIdx.x % L, jt = threadIdx.x / L;
  for (int i = threadIdx.x; i < N; i += NT) tw[i] = gmem(P.tw)[i];
  int bx, ky;
  line_group<L>(bx, ky);
  const int col = bx * L + line;
  const bool valid = col < P.hx;
  // spectra hold fewer than 2^31 elements (lengths up to 1024): 32-bit offsets
  const int plane = P.q.y * P.pitch;
  const int off = ky * P.pitch + col;
  float2 v[E];
#pragma unroll
  for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
    for (int t = 0; t < S::R0; ++t) {
      const int z = elem_index<N, E, S::R0>(jt, b, t);
      v[b * S::R0 + t] = make_float2(0.f, 0.f);
      if (valid) v[b * S::R0 + t] = gmem(P.spec)[z * plane + off];
    }
  auto addr = [line](int i) { return line_slot<N, L>(i, line); };
  fft_forward_order<N>(v, lds, jt, tw, 1, -1.f, addr);
  // R^(0) (kz = 0: v[0] of thread jt = 0 of line (kx, ky) = (0, 0)) from the double sum
  if (P.rsum && ky == 0 && col == 0 && jt == 0) v[0] = make_float2((float)*gmem(P.rsum), 0.f);
  if (P.corr) {
    // lag-shift phase exp(-2 pi i (kx sx / qx + ky sy / qy + kz sz / qz)): the
    // (kx, ky) part is one factor per line.  Index times shift stays below
    // 1024^2: 32-bit remainders (a 64-bit one per element cost the 288-point
    // gradient 5 %)
    const int tx = (int)(((uint32_t)(valid ? col : 0) * (uint32_t)P.shift.x) % (uint32_t)P.q.x);
    const int ty = (int)(((uint32_t)ky * (uint32_t)P.shift.y) % (uint32_t)P.q.y);
    const float2 pl = cmul(gmem(P.twx)[tx], gmem(P.twy)[ty]);
    float2 c[E];
#pragma unroll
    for (int b = 0; b < E / RL; ++b)
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int z = elem_index<N, E, RL>(jt, b, t);
        float2 w = make_float2(0.f, 0.f);
        if (valid) w = gmem(P.what)[z * plane + off];
        const float2 ph = cmul(pl, tw[((uint32_t)z * (uint32_t)P.shift.z) % (uint32_t)N]);
        c[b * RL + t] = cmul(cmul(cconj(v[b * RL + t]), w), ph);
      }
    if (P.mode == 0) {
      if (valid) {
#pragma unroll
        for (int b = 0; b < E / RL; ++b)
#pragma unroll
          for (int t = 0; t < RL; ++t) {
            const int o = elem_index<N, E, RL>(jt, b, t) * plane + off;
            float2 x = c[b * RL + t];
            if (!P.first) x = cadd(x, gmem(P.corr)[o]);
            gmem(P.corr)[o] = x;
          }
      }
    } else {
      fft_reverse_order<N>(c, lds, jt, tw, 1, 1.f, addr);
      if (valid) {
#pragma unroll
        for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
          for (int t = 0; t < S::R0; ++t) {
            const int z = elem_index<N, E, S::R0>(jt, b, t);
            if (z < P.kz) gmem(P.corr)[z * plane + off] = c[b * S::R0 + t];
          }
      }
    }
  }
  if (P.zhi > P.zlo) {
#pragma unroll
    for (int b = 0; b < E / RL; ++b)
#pragma unroll
      for (int t = 0; t < RL; ++t) {
        const int z = elem_index<N, E, RL>(jt, b, t);
        float2 p = make_float2(0.f, 0.f);
        if (valid) p = gmem(P.khat)[z * plane + off];
        v[b * RL + t] = cmul(v[b * RL + t], cconj(p));
      }
    fft_reverse_order<N>(v, lds, jt, tw, 1, 1.f, addr);
    if (valid) {
      float2* out = P.adj ? P.adj : P.spec;
#pragma unroll
      for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
        for (int t = 0; t < S::R0; ++t) {
          const int z = elem_index<N, E, S::R0>(jt, b, t);
          if (z >= P.zlo && z < P.zhi) {
            const int o = z * plane + off;
            gmem(out)[o] = P.adj_add ? cadd(gmem(out)[o], v[b * S::R0 + t]) : v[b * S::R0 + t];
          }
        }
    }
  }
}

hipError_t launch_z_fast_grad(const ShapePlan& sp, float2* spec, const float2* what, const float2* khat, float2* corr,
                              int mode, int first, Int3 shift, int zlo, int zhi, int kz, const double* rsum,
                              float2* adj, bool adj_add, hipStream_t stream) {
  ZGradFastParams P{spec, what, khat, corr, sp.tw_z, sp.tw_x, sp.tw_y, mode, first, shift, sp.q,
                    sp.hx, sp.pitch, zlo, zhi, kz, rsum, adj, adj_add ? 1 : 0};
  (void)hipGetLastError();
#define MVD_ZG(N)                                                                               \
  case N: {                                                                                     \
    constexpr int L = Spec<N>::NT / (N / Spec<N>::E);                                           \
    dim3 grid(z_line_groups(sp, L), sp.q.y);                                                    \
    hipLaunchKernelGGL(k_z_fast_grad<N>, grid, dim3(Spec<N>::NT), 0, stream, P);                \
    break;                                                                                      \
  }
  switch (sp.q.z) {
    MVD_ZG(64)
    MVD_ZG(128)
    MVD_ZG(144)
    MVD_ZG(256)
    MVD_ZG(192)
    MVD_ZG(288)
    MVD_ZG(384)
    MVD_ZG(512)
    default: return hipErrorInvalidValue;
  }
#undef MVD_ZG
  return hipGetLastError();
}

template <int N>
__global__ __launch_bounds__(Spec<N>::NT) void k_z_fast_batch(const ZFastParams* items, const int* prefix, const int* gx,
                                                   int nitems) {
  int item, bx, by;
  batch_locate(prefix, gx, nitems, item, bx, by, line_wg<lines_per_wg<N>()>());
  const auto item_params = items[item];  // private copy: the kernel's stores cannot alias it
  z_fast_body<N>(item_params, bx, by);
}

// ---------------------------------------------------------------------------
// x passes: rows of 2M real samples <-> M + 1 complex values.  Rows are staged
// row-major in LDS (coalesced global traffic); the FFT reads them with the
// rows fastest across lanes.

struct XFwdFastParams {
  ConvGeom g;
  const float* src;
  float2* spec;
  const float2* tw;  // 2M entries
  int pitch;
};

// Real-to-complex transform of the workgroup's L rows of 2M real samples held
// in LDS (row l at buf + l (M + 1), samples x-fastest, as M complex pairs) and
// the store of X[0, M] with the row padding [M + 1, M + 16) as zeros, so that
// the row's last 128-byte line is written by one store (a partly written line
// reaches HBM as a read-modify-write).  row_out(l): the row's output in global
// memory, or null for rows that are not stored.  The caller has synchronised
// after filling the rows.
template <int M, class RowOut>
__device__ __forceinline__ void r2c_rows_store(float2* buf, const float2* tw, RowOut row_out) {
  using S = Spec<M>;
  constexpr int E = S::E, T = M / E, NT = S::NT, L = NT / T, MP = M + 1;
  const int line = threadIdx.x % L, jt = threadIdx.x / L;
  auto addr = [line](int i) { return line * MP + i; };
  float2 v[E];
#pragma unroll
  for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
    for (int t = 0; t < S::R0; ++t) v[b * S::R0 + t] = buf[addr(elem_index<M, E, S::R0>(jt, b, t))];
  fft_forward_order<M>(v, buf, jt, tw, 2, -1.f, addr);
  constexpr int RL = last_radix<M>();
#pragma unroll
  for (int b = 0; b < E / RL; ++b)
#pragma unroll
    for (int t = 0; t < RL; ++t) buf[addr(elem_index<M, E, RL>(jt, b, t))] = v[b * RL + t];
  __syncthreads();
  // unpack Z = FFT(even + i odd) into X[k] for the pair (k, M - k); X[M] (the
  // pair of k = 0) goes out with the padding
  constexpr int K2 = M / 2 + 1;
  for (int idx = threadIdx.x; idx < L * K2; idx += NT) {
    const int l = idx / K2, k = idx % K2;
    const auto out = row_out(l);
    if (!out) continue;
    const float2 a = buf[l * MP + k], b = buf[l * MP + (M - k) % M];
    {
      const float2 bc = cconj(b);
      const float2 E2 = cscale(cadd(a, bc), 0.5f), D = cscale(csub(a, bc), 0.5f);
      stream_store(&out[k], cadd(E2, cmul(tw[k], make_float2(D.y, -D.x))));
    }
    if (M - k != k && k != 0) {
      const float2 ac = cconj(a);
      const float2 E2 = cscale(cadd(b, ac), 0.5f), D = cscale(csub(b, ac), 0.5f);
      stream_store(&out[M - k], cadd(E2, cmul(tw[M - k], make_float2(D.y, -D.x))));
    }
  }
  for (int idx = threadIdx.x; idx < L * 16; idx += NT) {
    const int l = idx >> 4, c = idx & 15;
    const auto out = row_out(l);
    if (!out) continue;
    float2 x = make_float2(0.f, 0.f);
    if (c == 0) {
      const float2 a = buf[l * MP];
      const float2 ac = cconj(a);
      const float2 E2 = cscale(cadd(a, ac), 0.5f), D = cscale(csub(a, ac), 0.5f);
      x = cadd(E2, cmul(tw[M], make_float2(D.y, -D.x)));
    }
    stream_store(&out[M + c], x);
  }
}

// One row of the window: rows are (y, z) of the FFT box, x-fastest.  Loops
// run over whole rows so that the (y, z) bookkeeping is done once per row.
template <int M, int MODE>
__device__ __forceinline__ void x_forward_fast_body(const XFwdFastParams& P, int bx, int by) {
  static_assert(M % 16 == 0, "the row pitch is M + 16");
  using S = Spec<M>;
  constexpr int E = S::E, T = M / E, NT = S::NT, L = NT / T, MP = M + 1;
  constexpr int W = 2 * M;                                   // real samples per row
  constexpr int RPI = W >= NT ? 1 : NT / W;                  // rows per loop trip
  constexpr int EPT = W >= NT ? (W + NT - 1) / NT : 1;       // samples per thread per row
  static_assert(W >= NT || NT % W == 0, "rows per trip");
  __shared__ float2 tw[2 * M];
  __shared__ float2 buf[L * MP];
  float* fbuf = reinterpret_cast<float*>(buf);
  const ConvGeom& g = P.g;
  for (int i = threadIdx.x; i < 2 * M; i += NT) tw[i] = gmem(P.tw)[i];
  const int nrows = g.q.y * g.q.z;
  const int row0 = bx * L;
  const int x0 = RPI > 1 ? threadIdx.x % W : threadIdx.x;
  const int lsub = RPI > 1 ? threadIdx.x / W : 0;
  constexpr int NR = L / RPI;  // rows handled by each thread
  // the last trip over a row is partial when NT does not divide W
  auto x_in_row = [&](int e) { return W % NT == 0 || x0 + e * NT < W; };
  // all loads of the thread first (independent, many in flight), then LDS stores
  float vals[NR][EPT];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int l = r * RPI + lsub;
    const int row = row0 + l;
    const int y = row % g.q.y, z = row / g.q.y;
    const bool live = row < nrows && y < g.win.y && z < g.win.z;
    decltype(gmem(P.src)) srow = nullptr;
    bool row_inside = false;  // the whole (y, z) row is inside the source
    if (live) {
      if (MODE == kGatherKernelWrap) {
        int jy = y + g.src.y / 2, jz = z + g.src.z / 2;
        if (jy >= g.q.y) jy -= g.q.y;
        if (jz >= g.q.z) jz -= g.q.z;
        row_inside = jy < g.src.y && jz < g.src.z;
        srow = gmem(P.src) + (int64_t)g.src.x * (jy + (int64_t)g.src.y * jz);
      } else {
        int gy = g.off.y + y, gz = g.off.z + z;
        row_inside = gy >= 0 && gy < g.src.y && gz >= 0 && gz < g.src.z;
        if (MODE == kGatherMirror && !row_inside) {
          gy = mirror_single(gy, g.src.y);
          gz = mirror_single(gz, g.src.z);
          row_inside = true;  // the mirrored row is a source row
        }
        srow = gmem(P.src) + (int64_t)g.src.x * (gy + (int64_t)g.src.y * gz);
      }
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int x = x0 + e * NT;
      float val = 0.f;
      if (live && x < g.win.x) {
        if (MODE == kGatherKernelWrap) {
          int jx = x + g.src.x / 2;
          if (jx >= g.q.x) jx -= g.q.x;
          if (row_inside && jx < g.src.x) val = srow[jx];
        } else {
          int gx = g.off.x + x;
          const bool in_x = gx >= 0 && gx < g.src.x;
          if (MODE == kGatherMirror) {
            if (!in_x) gx = mirror_single(gx, g.src.x);
            val = srow[gx];
          } else if (row_inside && in_x) {
            val = srow[gx];
          } else {
            val = MODE == kGatherOne ? 1.f : 0.f;
          }
        }
      }
      vals[r][e] = val;
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int e = 0; e < EPT; ++e)
      if (x_in_row(e)) fbuf[(r * RPI + lsub) * 2 * MP + x0 + e * NT] = vals[r][e];
  __syncthreads();
  r2c_rows_store<M>(buf, tw, [&](int l) {
    const int row = row0 + l;
    return row < nrows ? gmem(P.spec) + (int64_t)row * P.pitch : nullptr;
  });
}

template <int M, int MODE>
__global__ __launch_bounds__(Spec<M>::NT) void k_x_forward_fast(XFwdFastParams P) {
  x_forward_fast_body<M, MODE>(P, blockIdx.x, blockIdx.y);
}

// tasks (optional): one (item, plane) per wpt = q.y / L workgroups, each
// workgroup L rows of the plane; only the planes listed are transformed.
template <int M, int MODE>
__global__ __launch_bounds__(Spec<M>::NT) void k_x_forward_fast_batch(const XFwdFastParams* items, const int* prefix, const int* gx,
                                                   int nitems, const int2* tasks, int wpt) {
  int item, bx, by;
  if (tasks) {
    const int2 t = tasks[blockIdx.x / wpt];
    item = t.x;
    bx = t.y * wpt + blockIdx.x % wpt;
    by = 0;
  } else {
    batch_locate(prefix, gx, nitems, item, bx, by);
  }
  const auto item_params = items[item];  // private copy: the kernel's stores cannot alias it
  x_forward_fast_body<M, MODE>(item_params, bx, by);
}

struct XInvFastParams {
  ConvGeom g;
  const float2* spec;
  EpilogueArgs e;
  const float2* tw;  // 2M entries
  int pitch;
};

template <int M, int EPI>
__device__ __forceinline__ void x_inverse_fast_body(const XInvFastParams& P, int bx, int by) {
  using S = Spec<M>;
  constexpr int E = S::E, T = M / E, NT = S::NT, L = NT / T, MP = M + 1;
  constexpr int W = 2 * M;
  // rows per trip and values per thread when loading M complex per row: a
  // row spans whole trips (NT | M), or a trip holds whole rows (M | NT), or
  // one row per trip with the lanes past M idle
  constexpr int RPI_C = M >= NT || NT % M ? 1 : NT / M;
  constexpr int EPT_C = M >= NT ? (M + NT - 1) / NT : 1;
  constexpr int RPI = W >= NT ? 1 : NT / W;     // rows per trip over 2M reals
  constexpr int EPT = W >= NT ? (W + NT - 1) / NT : 1;
  static_assert(W >= NT || NT % W == 0, "rows per trip");
  __shared__ float2 tw[2 * M];
  __shared__ float2 buf[L * MP];
  __shared__ double red_sum[NT / 64];
  __shared__ float red_max[NT / 64];
  __shared__ double red_rsum[NT / 64];
  const ConvGeom& g = P.g;
  for (int i = threadIdx.x; i < 2 * M; i += NT) tw[i] = gmem(P.tw)[i];
  const int ny = g.eff_size.y;
  const int nrows = ny * g.eff_size.z;
  const int row0 = bx * L;
  {
    const int k0 = RPI_C > 1 ? threadIdx.x % M : threadIdx.x;
    const int lsub = RPI_C > 1 ? threadIdx.x / M : 0;
    constexpr int NR = L / RPI_C;
    float2 vals[NR][EPT_C];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int row = row0 + r * RPI_C + lsub;
      decltype(gmem(P.spec)) srow = nullptr;
      if (row < nrows) {
        const int y = g.eff_local.y + row % ny, z = g.eff_local.z + row / ny;
        srow = gmem(P.spec) + ((int64_t)z * g.q.y + y) * P.pitch;
      }
#pragma unroll
      for (int e = 0; e < EPT_C; ++e) {
        vals[r][e] = make_float2(0.f, 0.f);
        if (srow && k0 + e * NT < M) vals[r][e] = stream_load(&srow[k0 + e * NT]);
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int e = 0; e < EPT_C; ++e)
        if (k0 + e * NT < M) buf[(r * RPI_C + lsub) * MP + k0 + e * NT] = vals[r][e];
    // the Nyquist column k = M
    for (int l = threadIdx.x; l < L; l += NT) {
      const int row = row0 + l;
      float2 val = make_float2(0.f, 0.f);
      if (row < nrows) {
        const int y = g.eff_local.y + row % ny, z = g.eff_local.z + row / ny;
        val = stream_load(&gmem(P.spec)[((int64_t)z * g.q.y + y) * P.pitch + M]);
      }
      buf[l * MP + M] = val;
    }
  }
  __syncthreads();
  constexpr int K2 = M / 2 + 1;
  for (int idx = threadIdx.x; idx < L * K2; idx += NT) {
    const int l = idx / K2, k = idx % K2;
    const float2 a = buf[l * MP + k], b = buf[l * MP + M - k];
    const float2 za = cadd(cadd(a, cconj(b)), mul_di(cmul(csub(a, cconj(b)), cconj(tw[k])), 1.f));
    if (k != 0 && M - k != k) {
      const float2 zb = cadd(cadd(b, cconj(a)), mul_di(cmul(csub(b, cconj(a)), cconj(tw[M - k])), 1.f));
      buf[l * MP + M - k] = zb;
    }
    buf[l * MP + k] = za;
  }
  __syncthreads();
  const EpilogueArgs& e = P.e;
  const int nx = g.eff_size.x;
  const int x0 = RPI > 1 ? threadIdx.x % W : threadIdx.x;
  const int lsub = RPI > 1 ? threadIdx.x / W : 0;
  constexpr int NR = L / RPI;
  // The epilogue runs over the thread's rows in chunks of CR rows (about 8
  // voxels per thread): a chunk's destination indices and inputs (the view or
  // the estimate, and the weight) are loaded while the chunk before it is
  // computed, and the first chunk's loads are issued before the transform.
  constexpr int CR = EPT >= 8 ? 1 : (8 / EPT < NR ? 8 / EPT : NR);
  static_assert(NR % CR == 0, "epilogue chunks");
  constexpr int NCH = NR / CR;
  struct Chunk {
    int gi[CR][EPT];  // images hold fewer than 2^31 voxels (checked by the host)
    float in0[CR][EPT], in1[CR][EPT];
  };
  auto load_chunk = [&](int c, Chunk& ch) {
#pragma unroll
    for (int rr = 0; rr < CR; ++rr) {
      const int row = row0 + (c * CR + rr) * RPI + lsub;
      const bool live_row = row < nrows;
      const int yo = live_row ? row % ny : 0, zo = live_row ? row / ny : 0;
      const int base = g.eff_offset.x + g.dst.x * ((g.eff_offset.y + yo) + g.dst.y * (g.eff_offset.z + zo));
#pragma unroll
      for (int ee = 0; ee < EPT; ++ee) {
        const int xo = x0 + ee * NT;
        const bool live = live_row && xo < nx;
        const int gi = live ? base + xo : -1;
        ch.gi[rr][ee] = gi;
        ch.in0[rr][ee] = 0.f;
        ch.in1[rr][ee] = 1.f;
        if (live) {
          if (EPI == kEpiQuotient || EPI == kEpiResidual) {
            ch.in0[rr][ee] = gmem(e.img)[gi];
            if (EPI == kEpiResidual && e.weight_mode == kWeightArray) ch.in1[rr][ee] = gmem(e.weight)[gi];
          } else if (EPI == kEpiUpdate) {
            ch.in0[rr][ee] = gmem(e.dst)[gi];
            ch.in1[rr][ee] = e.weight_mode != kWeightNone ? gmem(e.weight)[gi] : e.wconst;
          }
        }
      }
    }
  };
  Chunk cur;
  load_chunk(0, cur);
  const int line = threadIdx.x % L, jt = threadIdx.x / L;
  auto addr = [line](int i) { return line * MP + i; };
  float2 v[E];
#pragma unroll
  for (int b = 0; b < E / S::R0; ++b)
#pragma unroll
    for (int t = 0; t < S::R0; ++t) v[b * S::R0 + t] = buf[addr(elem_index<M, E, S::R0>(jt, b, t))];
  fft_forward_order<M>(v, buf, jt, tw, 2, 1.f, addr);
  constexpr int RL = last_radix<M>();
#pragma unroll
  for (int b = 0; b < E / RL; ++b)
#pragma unroll
    for (int t = 0; t < RL; ++t) buf[addr(elem_index<M, E, RL>(jt, b, t))] = v[b * RL + t];
  __syncthreads();
  float* fbuf = reinterpret_cast<float*>(buf);
  const double inv_lambda = rl_inv_lambda(e.lambda);
  double my_sum = 0.0, my_rsum = 0.0;
  float my_max = -1.f;
  // residual rows transformed here (rspec): r goes back into the row in LDS
  const bool r_rows = EPI == kEpiResidual && e.rspec != nullptr;
  auto run_chunk = [&](int c, const Chunk& ch) {
#pragma unroll
    for (int rr = 0; rr < CR; ++rr) {
      const int l = (c * CR + rr) * RPI + lsub;
#pragma unroll
      for (int ee = 0; ee < EPT; ++ee) {
        const int gi = ch.gi[rr][ee];
        if (gi < 0) continue;
        const int xo = x0 + ee * NT;
        const float val = fbuf[l * 2 * MP + g.eff_local.x + xo];
        if (EPI == kEpiStore) {
          gmem(e.dst)[gi] = val;
        } else if (EPI == kEpiQuotient) {
          gmem(e.dst)[gi] = rl_quotient(val, ch.in0[rr][ee]);
        } else if (EPI == kEpiResidual) {
          const float w = ch.in1[rr][ee], y = ch.in0[rr][ee];
          float rv = 0.f;
          if (val > 0.f) {
            my_sum += nll_term(w, val, y);
            rv = nll_residual(w, val, y);
          }
          if (r_rows) {
            fbuf[l * 2 * MP + g.eff_local.x + xo] = rv;  // the value this thread read
            my_rsum += (double)rv;
          } else {
            gmem(e.dst)[gi] = rv;
          }
        } else {
          float w = ch.in1[rr][ee];
          if (e.weight_mode == kWeightVirtual) {
            const double rw = (double)ch.in1[rr][ee] / (double)gmem(e.wsum)[gi];
            w = (float)java_min(1.0, rw * e.osem);
          }
          const float last = ch.in0[rr][ee];
          const float next = rl_next_value(last, val, w, e.lambda, inv_lambda);
          gmem(e.dst)[gi] = next;
          const float change = fabsf(next - last);
          my_sum += (double)change;
          my_max = java_maxf(my_max, change);
        }
      }
    }
  };
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    Chunk nxt;
    if (c + 1 < NCH) load_chunk(c + 1, nxt);
    run_chunk(c, cur);
    if (c + 1 < NCH) cur = nxt;
  }
  if (EPI == kEpiUpdate || EPI == kEpiResidual) {
    for (int o = 32; o > 0; o >>= 1) {
      my_sum += __shfl_xor(my_sum, o);
      my_max = java_maxf(my_max, __shfl_xor(my_max, o));
      if (EPI == kEpiResidual) my_rsum += __shfl_xor(my_rsum, o);
    }
    const int wid = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
      red_sum[wid] = my_sum;
      red_max[wid] = my_max;
      if (EPI == kEpiResidual) red_rsum[wid] = my_rsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0, rs = 0.0;
      float m = -1.f;
      for (int i = 0; i < NT / 64; ++i) {
        s += red_sum[i];
        m = java_maxf(m, red_max[i]);
        if (EPI == kEpiResidual) rs += red_rsum[i];
      }
      gmem(e.part_sum)[bx] = s;
      gmem(e.part_max)[bx] = m;
      if (r_rows) gmem(e.part_rsum)[bx] = rs;
    }
  }
  if (r_rows) {
    // window row = the effective row placed at eff_local, zeros elsewhere;
    // then the x transform of the rows into the residual's window spectrum
    for (int idx = threadIdx.x; idx < L * W; idx += NT) {
      const int l = idx / W, x = idx % W;
      if (x < g.eff_local.x || x >= g.eff_local.x + nx) fbuf[l * 2 * MP + x] = 0.f;
    }
    __syncthreads();
    r2c_rows_store<M>(buf, tw, [&](int l) {
      const int row = row0 + l;
      return row < nrows ? gmem(e.rspec) + ((int64_t)(g.eff_local.z + row / ny) * g.q.y + g.eff_local.y + row % ny) *
                                               e.rpitch
                         : nullptr;
    });
  }
}

template <int M, int EPI>
__global__ __launch_bounds__(Spec<M>::NT) void k_x_inverse_fast(XInvFastParams P) {
  x_inverse_fast_body<M, EPI>(P, blockIdx.x, blockIdx.y);
}

template <int M, int EPI>
__global__ __launch_bounds__(Spec<M>::NT) void k_x_inverse_fast_batch(const XInvFastParams* items, const int* prefix, const int* gx,
                                                   int nitems) {
  int item, bx, by;
  batch_locate(prefix, gx, nitems, item, bx, by);
  const auto item_params = items[item];  // private copy: the kernel's stores cannot alias it
  x_inverse_fast_body<M, EPI>(item_params, bx, by);
}

// ---------------------------------------------------------------------------
// dispatch

// MVD_FAST_SMOOTH=0 keeps the 144-, 192-, 288- and 384-point lengths on the
// generic LDS kernels (comparisons).
static const bool g_fast_smooth = [] {
  const char* v = std::getenv("MVD_FAST_SMOOTH");
  return !(v && v[0] == '0');
}();

bool fast_size(int n) {
  return n == 64 || n == 128 || n == 256 || n == 512 ||
         (g_fast_smooth && (n == 144 || n == 192 || n == 288 || n == 384));
}

int x_fast_rows_per_wg(int M) {
  switch (M) {
    case 64: return lines_per_wg<64>();
    case 128: return lines_per_wg<128>();
    case 144: return lines_per_wg<144>();
    case 256: return lines_per_wg<256>();
    case 192: return lines_per_wg<192>();
    case 288: return lines_per_wg<288>();
    case 384: return lines_per_wg<384>();
    case 512: return lines_per_wg<512>();
  }
  return 0;
}

// threads per workgroup of the kernels of length n
static int fast_threads(int n) {
  switch (n) {
    case 144: return Spec<144>::NT;
    case 288: return Spec<288>::NT;
  }
  return kThreads;
}

bool pass_is_fast(const PassDesc& d) {
  switch (d.kind) {
    case kPassXForward:
    case kPassXInverse: return fast_size(d.g.q.x / 2);
    case kPassYForward:
    case kPassYInverse: return fast_size(d.sp->q.y);
    case kPassZConv: return fast_size(d.sp->q.z);
  }
  return false;
}

template <int N>
static int lines_of() {
  return lines_per_wg<N>();
}

static int fast_lines(int n) {
  switch (n) {
    case 64: return lines_of<64>();
    case 128: return lines_of<128>();
    case 144: return lines_of<144>();
    case 256: return lines_of<256>();
    case 192: return lines_of<192>();
    case 288: return lines_of<288>();
    case 384: return lines_of<384>();
    case 512: return lines_of<512>();
  }
  return 1;
}

int build_fast_params(const PassDesc& d, void* out, int* gx, int* gy, int* cls) {
  const ShapePlan& sp = *d.sp;
  const ConvGeom& g = d.g;
  switch (d.kind) {
    case kPassXForward: {
      XFwdFastParams P{g, d.src, d.spec, sp.tw_x, sp.pitch};
      const int M = g.q.x / 2, L = fast_lines(M);
      *gx = (g.q.y * g.q.z + L - 1) / L;
      *gy = 1;
      *cls = M * 16 + d.gather;
      std::memcpy(out, &P, sizeof(P));
      return sizeof(P);
    }
    case kPassYForward:
    case kPassYInverse: {
      YFastParams P{d.spec, sp.tw_y, sp.hx, sp.pitch, d.zlo, d.ylo, d.yhi, d.dir};
      *gx = y_line_groups(sp, fast_lines(sp.q.y));
      *gy = d.zhi - d.zlo;
      *cls = sp.q.y;
      std::memcpy(out, &P, sizeof(P));
      return sizeof(P);
    }
    case kPassZConv: {
      ZFastParams P{d.spec, d.kspec, sp.tw_z, d.scale, d.dir, sp.hx, sp.pitch, sp.q.y, d.zlo, d.zhi,
                    d.ztab, d.arena, d.arena_bytes};
      *gx = z_line_groups(sp, fast_lines(sp.q.z));
      *gy = sp.q.y;
      *cls = sp.q.z;
      std::memcpy(out, &P, sizeof(P));
      return sizeof(P);
    }
    case kPassXInverse: {
      XInvFastParams P{g, d.spec, d.e, sp.tw_x, sp.pitch};
      const int M = g.q.x / 2;
      *gx = x_inverse_workgroups(sp, g);
      *gy = 1;
      *cls = M * 16 + d.e.mode;
      std::memcpy(out, &P, sizeof(P));
      return sizeof(P);
    }
  }
  return -1;
}

hipError_t launch_batch_fast(int kind, int cls, const void* items, const int* prefix, const int* gx, int nitems,
                             int nwg, hipStream_t stream, const int2* tasks, int wpt) {
  if (tasks && (wpt <= 0 || (kind != kPassXForward && kind != kPassYForward))) return hipErrorInvalidValue;
  if (nwg <= 0) return hipSuccess;
  (void)hipGetLastError();
  const dim3 grid(nwg);
  // the class encodes the length: y and z passes cls = N, x passes cls = 16 M + mode
  const int n = kind == kPassXForward || kind == kPassXInverse ? cls / 16 : cls;
  const dim3 block(fast_threads(n));
  switch (kind) {
    case kPassYForward:
    case kPassYInverse: {
      const YFastParams* it = static_cast<const YFastParams*>(items);
      switch (cls) {
        case 64: hipLaunchKernelGGL(k_y_fast_batch<64>, grid, block, 0, stream, it, prefix, gx, nitems, tasks, wpt); break;
        case 128: hipLaunchKernelGGL(k_y_fast_batch<128>, grid, block, 0, stream, it, prefix, gx, nitems, tasks, wpt); break;
        case 144: hipLaunchKernelGGL(k_y_fast_batch<144>, grid, block, 0, stream, it, prefix, gx, nitems, tasks, wpt); break;
        case 256: hipLaunchKernelGGL(k_y_fast_batch<256>, grid, block, 0, stream, it, prefix, gx, nitems, tasks, wpt); break;
        case 192: hipLaunchKernelGGL(k_y_fast_batch<192>, grid, block, 0, stream, it, prefix, gx, nitems, tasks, wpt); break;
        case 288: hipLaunchKernelGGL(k_y_fast_batch<288>, grid, block, 0, stream, it, prefix, gx, nitems, tasks, wpt); break;
        case 384: hipLaunchKernelGGL(k_y_fast_batch<384>, grid, block, 0, stream, it, prefix, gx, nitems, tasks, wpt); break;
        case 512: hipLaunchKernelGGL(k_y_fast_batch<512>, grid, block, 0, stream, it, prefix, gx, nitems, tasks, wpt); break;
        default: return hipErrorInvalidValue;
      }
      break;
    }
    case kPassZConv: {
      const ZFastParams* it = static_cast<const ZFastParams*>(items);
      switch (cls) {
        case 64: hipLaunchKernelGGL(k_z_fast_batch<64>, grid, block, 0, stream, it, prefix, gx, nitems); break;
        case 128: hipLaunchKernelGGL(k_z_fast_batch<128>, grid, block, 0, stream, it, prefix, gx, nitems); break;
        case 144: hipLaunchKernelGGL(k_z_fast_batch<144>, grid, block, 0, stream, it, prefix, gx, nitems); break;
        case 256: hipLaunchKernelGGL(k_z_fast_batch<256>, grid, block, 0, stream, it, prefix, gx, nitems); break;
        case 192: hipLaunchKernelGGL(k_z_fast_batch<192>, grid, block, 0, stream, it, prefix, gx, nitems); break;
        case 288: hipLaunchKernelGGL(k_z_fast_batch<288>, grid, block, 0, stream, it, prefix, gx, nitems); break;
        case 384: hipLaunchKernelGGL(k_z_fast_batch<384>, grid, block, 0, stream, it, prefix, gx, nitems); break;
        case 512: hipLaunchKernelGGL(k_z_fast_batch<512>, grid, block, 0, stream, it, prefix, gx, nitems); break;
        default: return hipErrorInvalidValue;
      }
      break;
    }
    case kPassXForward: {
      const XFwdFastParams* it = static_cast<const XFwdFastParams*>(items);
#define MVD_BXF(M, MODE)                                                                                   \
  case M * 16 + MODE:                                                                                      \
    hipLaunchKernelGGL((k_x_forward_fast_batch<M, MODE>), grid, block, 0, stream, it, prefix, gx, nitems, tasks, wpt); \
    break;
#define MVD_BXF4(M) MVD_BXF(M, kGatherMirror) MVD_BXF(M, kGatherOne) MVD_BXF(M, kGatherZero) MVD_BXF(M, kGatherKernelWrap)
      switch (cls) {
        MVD_BXF4(64)
        MVD_BXF4(128)
        MVD_BXF4(144)
        MVD_BXF4(256)
        MVD_BXF4(192)
        MVD_BXF4(288)
        MVD_BXF4(384)
        MVD_BXF4(512)
        default: return hipErrorInvalidValue;
      }
#undef MVD_BXF4
#undef MVD_BXF
      break;
    }
    case kPassXInverse: {
      const XInvFastParams* it = static_cast<const XInvFastParams*>(items);
#define MVD_BXI(M, EPI)                                                                                    \
  case M * 16 + EPI:                                                                                       \
    hipLaunchKernelGGL((k_x_inverse_fast_batch<M, EPI>), grid, block, 0, stream, it, prefix, gx, nitems);  \
    break;
#define MVD_BXI4(M) MVD_BXI(M, kEpiStore) MVD_BXI(M, kEpiQuotient) MVD_BXI(M, kEpiUpdate) MVD_BXI(M, kEpiResidual)
      switch (cls) {
        MVD_BXI4(64)
        MVD_BXI4(128)
        MVD_BXI4(144)
        MVD_BXI4(256)
        MVD_BXI4(192)
        MVD_BXI4(288)
        MVD_BXI4(384)
        MVD_BXI4(512)
        default: return hipErrorInvalidValue;
      }
#undef MVD_BXI4
#undef MVD_BXI
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_y_fast_window(const ShapePlan& sp, float2* spec, int zlo, int zhi, int ylo, int yhi, float dir,
                                Int3 in_lo, Int3 in_hi, hipStream_t stream) {
  YFastParams P{spec, sp.tw_y, sp.hx, sp.pitch, zlo, ylo, yhi, dir, in_lo.y, in_hi.y, in_lo.z, in_hi.z};
  const int n = sp.q.y;
  (void)hipGetLastError();
#define MVD_Y(N)                                                                                \
  case N: {                                                                                     \
    dim3 grid(y_line_groups(sp, lines_per_wg<N>()), zhi - zlo);                                 \
    hipLaunchKernelGGL(k_y_fast<N>, grid, dim3(Spec<N>::NT), 0, stream, P);                     \
    break;                                                                                      \
  }
  switch (n) {
    MVD_Y(64)
    MVD_Y(128)
    MVD_Y(144)
    MVD_Y(256)
    MVD_Y(192)
    MVD_Y(288)
    MVD_Y(384)
    MVD_Y(512)
    default: return hipErrorInvalidValue;
  }
#undef MVD_Y
  return hipGetLastError();
}

hipError_t launch_y_fast(const ShapePlan& sp, float2* spec, int zlo, int zhi, int ylo, int yhi, float dir,
                         hipStream_t stream) {
  return launch_y_fast_window(sp, spec, zlo, zhi, ylo, yhi, dir, {0, 0, 0}, {1 << 30, 1 << 30, 1 << 30}, stream);
}

hipError_t launch_z_fast(const ShapePlan& sp, float2* spec, const float2* kspec, float scale, int zlo, int zhi,
                         hipStream_t stream, float dir, float2* aux, const float2* pre) {
  ZFastParams P{spec, kspec, sp.tw_z, scale, dir, sp.hx, sp.pitch, sp.q.y, zlo, zhi, nullptr, nullptr, 0, aux, pre};
  (void)hipGetLastError();
#define MVD_Z(N)                                                                                \
  case N: {                                                                                     \
    dim3 grid(z_line_groups(sp, lines_per_wg<N>()), sp.q.y);                                    \
    hipLaunchKernelGGL(k_z_fast<N>, grid, dim3(Spec<N>::NT), 0, stream, P);                     \
    break;                                                                                      \
  }
  switch (sp.q.z) {
    MVD_Z(64)
    MVD_Z(128)
    MVD_Z(144)
    MVD_Z(256)
    MVD_Z(192)
    MVD_Z(288)
    MVD_Z(384)
    MVD_Z(512)
    default: return hipErrorInvalidValue;
  }
#undef MVD_Z
  return hipGetLastError();
}

hipError_t launch_x_forward_fast(const ShapePlan& sp, const ConvGeom& g, const float* src, int gather, float2* spec,
                                 hipStream_t stream) {
  XFwdFastParams P{g, src, spec, sp.tw_x, sp.pitch};
  const int M = g.q.x / 2;
  const int nrows = g.q.y * g.q.z;
  (void)hipGetLastError();
#define MVD_XF_MODE(N, MODE)                                                                    \
  case MODE:                                                                                    \
    hipLaunchKernelGGL((k_x_forward_fast<N, MODE>), grid, dim3(Spec<N>::NT), 0, stream, P);     \
    break;
#define MVD_XF(N)                                                                               \
  case N: {                                                                                     \
    dim3 grid((nrows + lines_per_wg<N>() - 1) / lines_per_wg<N>());                             \
    switch (gather) {                                                                           \
      MVD_XF_MODE(N, kGatherMirror)                                                             \
      MVD_XF_MODE(N, kGatherOne)                                                                \
      MVD_XF_MODE(N, kGatherZero)                                                               \
      MVD_XF_MODE(N, kGatherKernelWrap)                                                         \
      default: return hipErrorInvalidValue;                                                     \
    }                                                                                           \
    break;                                                                                      \
  }
  switch (M) {
    MVD_XF(64)
    MVD_XF(128)
    MVD_XF(144)
    MVD_XF(256)
    MVD_XF(192)
    MVD_XF(288)
    MVD_XF(384)
    MVD_XF(512)
    default: return hipErrorInvalidValue;
  }
#undef MVD_XF
#undef MVD_XF_MODE
  return hipGetLastError();
}

hipError_t launch_x_inverse_fast(const ShapePlan& sp, const ConvGeom& g, const float2* spec, const EpilogueArgs& e,
                                 hipStream_t stream, int nwg) {
  XInvFastParams P{g, spec, e, sp.tw_x, sp.pitch};
  const int M = g.q.x / 2;
  (void)hipGetLastError();
#define MVD_XI_EPI(N, EPI)                                                                      \
  case EPI:                                                                                     \
    hipLaunchKernelGGL((k_x_inverse_fast<N, EPI>), dim3(nwg), dim3(Spec<N>::NT), 0, stream, P); \
    break;
#define MVD_XI(N)                                                                               \
  case N:                                                                                       \
    switch (e.mode) {                                                                           \
      MVD_XI_EPI(N, kEpiStore)                                                                  \
      MVD_XI_EPI(N, kEpiQuotient)                                                               \
      MVD_XI_EPI(N, kEpiUpdate)                                                                 \
      MVD_XI_EPI(N, kEpiResidual)                                                               \
      default: return hipErrorInvalidValue;                                                     \
    }                                                                                           \
    break;
  switch (M) {
    MVD_XI(64)
    MVD_XI(128)
    MVD_XI(144)
    MVD_XI(256)
    MVD_XI(192)
    MVD_XI(288)
    MVD_XI(384)
    MVD_XI(512)
    default: return hipErrorInvalidValue;
  }
#undef MVD_XI
#undef MVD_XI_EPI
  return hipGetLastError();
}

}  // namespace mvd
