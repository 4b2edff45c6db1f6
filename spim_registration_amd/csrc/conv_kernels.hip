// FFT convolution passes of one block (window) of a volume, for gfx950.
//
// A 3D real convolution with circular wrap over the padded size q is done in
// four launches over a half spectrum S laid out [z][y][kx], kx in [0, q.x/2]:
//   x_forward : gather the window (with the out-of-bounds rule of the
//               reference), real-to-complex transform along x (a length q.x/2
//               complex FFT of the packed even/odd samples plus an unpack
//               step), write S rows.
//   y         : complex FFT along y of every (z, kx) column.
//   z         : forward FFT along z, product with the kernel spectrum,
//               inverse FFT along z -- one launch, the line never leaves LDS.
//               Only the z rows the output box needs are written back.
//   y (inv)   : inverse FFT along y, only for the needed z, and only the
//               needed y rows are written.
//   x_inverse : complex-to-real along x for the output rows, then the fused
//               epilogue (store, quotient or multiplicative update).
// Kernel spectra are pre-scaled by 1/(q.x q.y q.z), so every transform here
// is unnormalised.
#include <cstring>

#include "mvd_internal.h"
#include "pointwise_math.h"

namespace mvd {

struct XFwdParams {
  ConvGeom g;
  const float* src;
  int gather;
  float2* spec;
  FftPlan plan;
  const float2* tw;
  int hx, pitch;
  int lines;
};

__device__ __forceinline__ float gather_value(const XFwdParams& P, int x, int y, int z) {
  const ConvGeom& g = P.g;
  if (P.gather == kGatherKernelWrap) {
    // padded position (x, y, z) holds k[(p + k/2) mod q] when that index lies in the kernel
    int jx = x + g.src.x / 2, jy = y + g.src.y / 2, jz = z + g.src.z / 2;
    if (jx >= g.q.x) jx -= g.q.x;
    if (jy >= g.q.y) jy -= g.q.y;
    if (jz >= g.q.z) jz -= g.q.z;
    if (jx < g.src.x && jy < g.src.y && jz < g.src.z)
      return gmem(P.src)[jx + (int64_t)g.src.x * (jy + (int64_t)g.src.y * jz)];
    return 0.f;
  }
  const int gx = g.off.x + x, gy = g.off.y + y, gz = g.off.z + z;
  const bool inside = gx >= 0 && gx < g.src.x && gy >= 0 && gy < g.src.y && gz >= 0 && gz < g.src.z;
  if (inside) return gmem(P.src)[gx + (int64_t)g.src.x * (gy + (int64_t)g.src.y * gz)];
  if (P.gather == kGatherOne) return 1.f;
  if (P.gather == kGatherZero) return 0.f;
  const int mx = mirror_single(gx, g.src.x), my = mirror_single(gy, g.src.y), mz = mirror_single(gz, g.src.z);
  return gmem(P.src)[mx + (int64_t)g.src.x * (my + (int64_t)g.src.y * mz)];
}

__device__ __forceinline__ void x_forward_body(const XFwdParams& P, int bx, int by) {
  extern __shared__ float2 smem[];
  const int qx = P.g.q.x, M = qx / 2;
  const int B = P.lines, lp = B + 1;
  float2* tw = smem;
  float2* buf = smem + qx;
  for (int i = threadIdx.x; i < qx; i += kThreads) tw[i] = gmem(P.tw)[i];
  const int nrows = P.g.q.y * P.g.q.z;
  const int row0 = bx * B;
  float* fbuf = reinterpret_cast<float*>(buf);
  for (int idx = threadIdx.x; idx < B * qx; idx += kThreads) {
    const int l = idx / qx, x = idx - l * qx;
    const int row = row0 + l;
    float v = 0.f;
    if (row < nrows) {
      const int y = row % P.g.q.y, z = row / P.g.q.y;
      if (x < P.g.win.x && y < P.g.win.y && z < P.g.win.z) v = gather_value(P, x, y, z);
    }
    fbuf[2 * ((x >> 1) * lp + l) + (x & 1)] = v;
  }
  __syncthreads();
  lds_fft(buf, lp, B, P.plan, tw, 2, -1.f);
  // unpack Z = FFT(even + i odd) into X[k], k in [0, M]; thread handles the pair (k, M - k)
  const int K2 = M / 2 + 1;
  for (int idx = threadIdx.x; idx < B * K2; idx += kThreads) {
    const int l = idx / K2, k = idx - l * K2;
    const int row = row0 + l;
    if (row >= nrows) continue;
    const int km = (M - k) % M;
    const float2 a = buf[k * lp + l], b = buf[km * lp + l];
    const auto out = gmem(P.spec) + (int64_t)row * P.pitch;
    {
      const float2 bc = cconj(b);
      const float2 E = cscale(cadd(a, bc), 0.5f), D = cscale(csub(a, bc), 0.5f);
      const float2 O = make_float2(D.y, -D.x);
      out[k] = cadd(E, cmul(tw[k], O));
    }
    if (M - k != k) {
      const float2 ac = cconj(a);
      const float2 E = cscale(cadd(b, ac), 0.5f), D = cscale(csub(b, ac), 0.5f);
      const float2 O = make_float2(D.y, -D.x);
      out[M - k] = cadd(E, cmul(tw[M - k], O));
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_x_forward(XFwdParams P) { x_forward_body(P, blockIdx.x, blockIdx.y); }

__global__ __launch_bounds__(kThreads) void k_x_forward_batch(const XFwdParams* items, const int* prefix, const int* gx,
                                                   int nitems) {
  int item, bx, by;
  batch_locate(prefix, gx, nitems, item, bx, by);
  x_forward_body(items[item], bx, by);
}

struct YParams {
  float2* spec;
  int hx, pitch, qy, zlo, ylo, yhi, lines;
  float dir;
  FftPlan plan;
  const float2* tw;
};

__device__ __forceinline__ void y_body(const YParams& P, int bx, int by) {
  extern __shared__ float2 smem[];
  const int B = P.lines, lp = B + 1, qy = P.qy;
  float2* tw = smem;
  float2* buf = smem + qy;
  for (int i = threadIdx.x; i < qy; i += kThreads) tw[i] = gmem(P.tw)[i];
  const int c0 = bx * B;
  const int z = P.zlo + by;
  const auto base = gmem(P.spec) + (int64_t)z * qy * P.pitch + c0;
  for (int idx = threadIdx.x; idx < qy * B; idx += kThreads) {
    const int y = idx / B, b = idx - y * B;
    buf[y * lp + b] = (c0 + b < P.hx) ? base[(int64_t)y * P.pitch + b] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft(buf, lp, B, P.plan, tw, 1, P.dir);
  const int nrow = P.yhi - P.ylo;
  for (int idx = threadIdx.x; idx < nrow * B; idx += kThreads) {
    const int y = P.ylo + idx / B, b = idx % B;
    if (c0 + b < P.hx) base[(int64_t)y * P.pitch + b] = buf[y * lp + b];
  }
}

__global__ __launch_bounds__(kThreads) void k_y(YParams P) { y_body(P, blockIdx.x, blockIdx.y); }

__global__ __launch_bounds__(kThreads) void k_y_batch(const YParams* items, const int* prefix, const int* gx,
                                                   int nitems) {
  int item, bx, by;
  batch_locate(prefix, gx, nitems, item, bx, by);
  y_body(items[item], bx, by);
}

struct ZParams {
  float2* spec;
  const float2* kspec;
  float2* aux;  // with kspec: the forward spectrum before the product, all rows (nullable)
  const float2* pre;  // with kspec: the lines' spectrum already transformed along z (nullable)
  float scale, dir;
  int hx, pitch, qy, qz, zlo, zhi, lines;
  FftPlan plan;
  const float2* tw;
};

__device__ __forceinline__ void z_body(const ZParams& P, int bx, int by) {
  extern __shared__ float2 smem[];
  const int B = P.lines, lp = B + 1, qz = P.qz;
  float2* tw = smem;
  float2* buf = smem + qz;
  for (int i = threadIdx.x; i < qz; i += kThreads) tw[i] = gmem(P.tw)[i];
  const int c0 = bx * B;
  const int y = by;
  const int64_t plane = (int64_t)P.qy * P.pitch;
  const int64_t col = (int64_t)y * P.pitch + c0;
  const float2* in = P.kspec && P.pre ? P.pre : P.spec;
  for (int idx = threadIdx.x; idx < qz * B; idx += kThreads) {
    const int z = idx / B, b = idx - z * B;
    buf[z * lp + b] = make_float2(0.f, 0.f);
    if (c0 + b < P.hx) buf[z * lp + b] = gmem(in)[z * plane + col + b];
  }
  __syncthreads();
  if (!(P.kspec && P.pre)) lds_fft(buf, lp, B, P.plan, tw, 1, P.kspec ? -1.f : P.dir);
  if (P.kspec) {
    for (int idx = threadIdx.x; idx < qz * B; idx += kThreads) {
      const int z = idx / B, b = idx - z * B;
      if (c0 + b < P.hx) {
        const float2 v = buf[z * lp + b];
        if (P.aux) gmem(P.aux)[z * plane + col + b] = v;
        buf[z * lp + b] = cmul(v, gmem(P.kspec)[z * plane + col + b]);
      }
    }
    __syncthreads();
    lds_fft(buf, lp, B, P.plan, tw, 1, 1.f);
    const int nrow = P.zhi - P.zlo;
    for (int idx = threadIdx.x; idx < nrow * B; idx += kThreads) {
      const int z = P.zlo + idx / B, b = idx % B;
      if (c0 + b < P.hx) gmem(P.spec)[z * plane + col + b] = buf[z * lp + b];
    }
  } else {
    for (int idx = threadIdx.x; idx < qz * B; idx += kThreads) {
      const int z = idx / B, b = idx - z * B;
      if (c0 + b < P.hx) gmem(P.spec)[z * plane + col + b] = cscale(buf[z * lp + b], P.scale);
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_z(ZParams P) { z_body(P, blockIdx.x, blockIdx.y); }

// z pass of the likelihood gradient, after the x and y transforms of a block's
// residual R.  Per line: the forward transform R^ (z), then
//   adjoint     A = R^ conj(P^)  -> inverse z -> rows [zlo, zhi) of spec (in
//               place; skipped when zhi <= zlo);
//   correlation C = conj(R^) W^ e(-2 pi i <k, s> / q), the lag shift s moving
//               the lags the kernel needs, c - j for j in the kernel box, to
//               [0, kd): accumulated into corr over the blocks (mode 0), or
//               inverse-transformed along z with rows [0, kz) stored (mode 1);
//               skipped when corr is null.
// LDS holds the two line sets one after the other.
struct ZGradParams {
  float2* spec;
  const float2* what;
  const float2* khat;
  float2* corr;
  int mode, first;
  Int3 shift, q;
  int hx, pitch, zlo, zhi, kz, lines;
  FftPlan plan;
  const float2* tw;
  const float2* twx;  // exp(-2 pi i m / q.x), m < q.x
  const float2* twy;  // exp(-2 pi i m / q.y), m < q.y
  const double* rsum;  // R^(0) from the residual's double sum (nullable)
  float2* adj;         // the adjoint rows' destination (nullable: spec, in place)
  int adj_add;         // add the adjoint rows to adj instead of storing them
};

__global__ __launch_bounds__(kThreads) void k_z_grad(ZGradParams P) {
  extern __shared__ float2 smem[];
  const int B = P.lines, lp = B + 1, qz = P.q.z;
  float2* tw = smem;
  float2* buf = smem + qz;
  float2* cbuf = buf + qz * lp;
  for (int i = threadIdx.x; i < qz; i += kThreads) tw[i] = gmem(P.tw)[i];
  const int c0 = blockIdx.x * B;
  const int y = blockIdx.y;
  const int64_t plane = (int64_t)P.q.y * P.pitch;
  const int64_t col = (int64_t)y * P.pitch + c0;
  for (int idx = threadIdx.x; idx < qz * B; idx += kThreads) {
    const int z = idx / B, b = idx - z * B;
    buf[z * lp + b] = make_float2(0.f, 0.f);
    if (c0 + b < P.hx) buf[z * lp + b] = gmem(P.spec)[z * plane + col + b];
  }
  __syncthreads();
  lds_fft(buf, lp, B, P.plan, tw, 1, -1.f);
  const bool adj = P.zhi > P.zlo, corr = P.corr != nullptr;
  // the workgroup of line (kx, ky) = (0, 0) takes R^(0) from the double sum
  const bool dc = P.rsum && c0 == 0 && y == 0;
  const float dcv = dc ? (float)*gmem(P.rsum) : 0.f;
  // per-line part of the lag-shift phase (kx = c0 + b, ky = y); the products
  // of an index and a shift stay below 1024^2, so 32-bit remainders
  const int ty = (int)(((uint32_t)y * (uint32_t)P.shift.y) % (uint32_t)P.q.y);
  const float2 py = gmem(P.twy)[ty];
  for (int idx = threadIdx.x; idx < qz * B; idx += kThreads) {
    const int z = idx / B, b = idx - z * B;
    const int kx = c0 + b;
    float2 a = make_float2(0.f, 0.f), c = make_float2(0.f, 0.f);
    if (kx < P.hx) {
      const int64_t o = z * plane + col + b;
      const float2 r = dc && z == 0 && kx == 0 ? make_float2(dcv, 0.f) : buf[z * lp + b];
      if (adj) a = cmul(r, cconj(gmem(P.khat)[o]));
      if (corr) {
        const int tx = (int)(((uint32_t)kx * (uint32_t)P.shift.x) % (uint32_t)P.q.x);
        const int tz = (int)(((uint32_t)z * (uint32_t)P.shift.z) % (uint32_t)qz);
        const float2 ph = cmul(cmul(gmem(P.twx)[tx], py), tw[tz]);
        c = cmul(cmul(cconj(r), gmem(P.what)[o]), ph);
        if (P.mode == 0) {
          if (!P.first) c = cadd(c, gmem(P.corr)[o]);
          gmem(P.corr)[o] = c;
        }
      }
    }
    buf[z * lp + b] = a;
    cbuf[z * lp + b] = c;
  }
  __syncthreads();
  if (adj) lds_fft(buf, lp, B, P.plan, tw, 1, 1.f);
  if (corr && P.mode == 1) lds_fft(cbuf, lp, B, P.plan, tw, 1, 1.f);
  float2* adj_out = P.adj ? P.adj : P.spec;
  for (int idx = threadIdx.x; idx < (P.zhi - P.zlo) * B; idx += kThreads) {
    const int z = P.zlo + idx / B, b = idx % B;
    if (c0 + b < P.hx) {
      const int64_t o = z * plane + col + b;
      const float2 x = buf[z * lp + b];
      gmem(adj_out)[o] = P.adj_add ? cadd(gmem(adj_out)[o], x) : x;
    }
  }
  if (corr && P.mode == 1) {
    for (int idx = threadIdx.x; idx < P.kz * B; idx += kThreads) {
      const int z = idx / B, b = idx % B;
      if (c0 + b < P.hx) gmem(P.corr)[z * plane + col + b] = cbuf[z * lp + b];
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_z_batch(const ZParams* items, const int* prefix, const int* gx,
                                                   int nitems) {
  int item, bx, by;
  batch_locate(prefix, gx, nitems, item, bx, by);
  z_body(items[item], bx, by);
}

struct XInvParams {
  ConvGeom g;
  const float2* spec;
  EpilogueArgs e;
  FftPlan plan;
  const float2* tw;
  int hx, pitch, lines;
};

__device__ __forceinline__ void x_inverse_body(const XInvParams& P, int bx, int by) {
  extern __shared__ float2 smem[];
  __shared__ double red_sum[kThreads / 64];
  __shared__ float red_max[kThreads / 64];
  const ConvGeom& g = P.g;
  const int qx = g.q.x, M = qx / 2;
  const int B = P.lines, lp = B + 1;
  float2* tw = smem;
  float2* buf = smem + qx;
  for (int i = threadIdx.x; i < qx; i += kThreads) tw[i] = gmem(P.tw)[i];
  const int ny = g.eff_size.y;
  const int nrows = ny * g.eff_size.z;
  const int row0 = bx * B;
  const int H = M + 1;
  for (int idx = threadIdx.x; idx < B * H; idx += kThreads) {
    const int l = idx / H, k = idx - l * H;
    const int row = row0 + l;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) {
      const int y = g.eff_local.y + row % ny, z = g.eff_local.z + row / ny;
      v = gmem(P.spec)[((int64_t)z * g.q.y + y) * P.pitch + k];
    }
    buf[k * lp + l] = v;
  }
  __syncthreads();
  // pack X[k], X[M-k] into Z[k] = E + i O (even/odd inverse split)
  const int K2 = M / 2 + 1;
  for (int idx = threadIdx.x; idx < B * K2; idx += kThreads) {
    const int l = idx / K2, k = idx - l * K2;
    const float2 a = buf[k * lp + l], b = buf[(M - k) * lp + l];
    const float2 za = cadd(cadd(a, cconj(b)), mul_di(cmul(csub(a, cconj(b)), cconj(tw[k])), 1.f));
    if (k != 0 && M - k != k) {
      const float2 zb = cadd(cadd(b, cconj(a)), mul_di(cmul(csub(b, cconj(a)), cconj(tw[M - k])), 1.f));
      buf[(M - k) * lp + l] = zb;
    }
    buf[k * lp + l] = za;
  }
  __syncthreads();
  lds_fft(buf, lp, B, P.plan, tw, 2, 1.f);
  const float* fbuf = reinterpret_cast<const float*>(buf);
  const EpilogueArgs& e = P.e;
  const double inv_lambda = rl_inv_lambda(e.lambda);
  double my_sum = 0.0;
  float my_max = -1.f;
  const int nx = g.eff_size.x;
  for (int idx = threadIdx.x; idx < B * nx; idx += kThreads) {
    const int l = idx / nx, xo = idx - l * nx;
    const int row = row0 + l;
    if (row >= nrows) continue;
    const int x = g.eff_local.x + xo;
    const float v = fbuf[2 * ((x >> 1) * lp + l) + (x & 1)];
    const int yo = row % ny, zo = row / ny;
    const int64_t gi = (int64_t)(g.eff_offset.x + xo) +
                       (int64_t)g.dst.x * ((int64_t)(g.eff_offset.y + yo) + (int64_t)g.dst.y * (g.eff_offset.z + zo));
    if (e.mode == kEpiStore) {
      gmem(e.dst)[gi] = v;
    } else if (e.mode == kEpiResidual) {
      const float w = e.weight_mode == kWeightArray ? gmem(e.weight)[gi] : 1.f;
      const float y = gmem(e.img)[gi];
      if (v > 0.f) {
        my_sum += nll_term(w, v, y);
        gmem(e.dst)[gi] = nll_residual(w, v, y);
      } else {
        gmem(e.dst)[gi] = 0.f;
      }
    } else if (e.mode == kEpiQuotient) {
      gmem(e.dst)[gi] = rl_quotient(v, gmem(e.img)[gi]);
    } else {
      float w = e.wconst;
      if (e.weight_mode == kWeightArray) {
        w = gmem(e.weight)[gi];
      } else if (e.weight_mode == kWeightVirtual) {
        const double r = (double)gmem(e.weight)[gi] / (double)gmem(e.wsum)[gi];
        w = (float)java_min(1.0, r * e.osem);
      }
      const float last = gmem(e.dst)[gi];
      const float next = rl_next_value(last, v, w, e.lambda, inv_lambda);
      gmem(e.dst)[gi] = next;
      const float change = fabsf(next - last);
      my_sum += (double)change;
      my_max = java_maxf(my_max, change);
    }
  }
  if (e.mode == kEpiUpdate || e.mode == kEpiResidual) {
    for (int o = 32; o > 0; o >>= 1) {
      my_sum += __shfl_xor(my_sum, o);
      my_max = java_maxf(my_max, __shfl_xor(my_max, o));
    }
    const int wid = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
      red_sum[wid] = my_sum;
      red_max[wid] = my_max;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      float m = -1.f;
      for (int i = 0; i < kThreads / 64; ++i) {
        s += red_sum[i];
        m = java_maxf(m, red_max[i]);
      }
      gmem(e.part_sum)[bx] = s;
      gmem(e.part_max)[bx] = m;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_x_inverse(XInvParams P) { x_inverse_body(P, blockIdx.x, blockIdx.y); }

__global__ __launch_bounds__(kThreads) void k_x_inverse_batch(const XInvParams* items, const int* prefix, const int* gx,
                                                   int nitems) {
  int item, bx, by;
  batch_locate(prefix, gx, nitems, item, bx, by);
  x_inverse_body(items[item], bx, by);
}

// ---------------------------------------------------------------------------
// launchers

hipError_t launch_x_forward(const ShapePlan& sp, const ConvGeom& g, const float* src, int gather, float2* spec,
                            hipStream_t stream) {
  if (fast_size(g.q.x / 2)) return launch_x_forward_fast(sp, g, src, gather, spec, stream);
  XFwdParams P;
  P.g = g;
  P.src = src;
  P.gather = gather;
  P.spec = spec;
  P.plan = sp.px;
  P.tw = sp.tw_x;
  P.hx = sp.hx;
  P.pitch = sp.pitch;
  P.lines = sp.bx;
  const int nrows = g.q.y * g.q.z;
  const int M = g.q.x / 2;
  dim3 grid((nrows + P.lines - 1) / P.lines);
  size_t lds = sizeof(float2) * (g.q.x + (size_t)M * (P.lines + 1));
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_x_forward, grid, dim3(kThreads), lds, stream, P);
  return hipGetLastError();
}

hipError_t launch_y(const ShapePlan& sp, float2* spec, int zlo, int zhi, int ylo, int yhi, float dir,
                    hipStream_t stream) {
  if (zhi <= zlo) return hipSuccess;
  if (fast_size(sp.q.y)) return launch_y_fast(sp, spec, zlo, zhi, ylo, yhi, dir, stream);
  YParams P;
  P.spec = spec;
  P.hx = sp.hx;
  P.pitch = sp.pitch;
  P.qy = sp.q.y;
  P.zlo = zlo;
  P.ylo = ylo;
  P.yhi = yhi;
  P.lines = sp.by;
  P.dir = dir;
  P.plan = sp.py;
  P.tw = sp.tw_y;
  dim3 grid((sp.hx + P.lines - 1) / P.lines, zhi - zlo);
  size_t lds = sizeof(float2) * (sp.q.y + (size_t)sp.q.y * (P.lines + 1));
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_y, grid, dim3(kThreads), lds, stream, P);
  return hipGetLastError();
}

hipError_t launch_z(const ShapePlan& sp, float2* spec, const float2* kspec, float scale, int zlo, int zhi,
                    hipStream_t stream, float dir, float2* aux, const float2* pre) {
  if (fast_size(sp.q.z)) return launch_z_fast(sp, spec, kspec, scale, zlo, zhi, stream, dir, aux, pre);
  ZParams P;
  P.aux = aux;
  P.pre = pre;
  P.dir = dir;
  P.spec = spec;
  P.kspec = kspec;
  P.scale = scale;
  P.hx = sp.hx;
  P.pitch = sp.pitch;
  P.qy = sp.q.y;
  P.qz = sp.q.z;
  P.zlo = zlo;
  P.zhi = zhi;
  P.lines = sp.bz;
  P.plan = sp.pz;
  P.tw = sp.tw_z;
  dim3 grid((sp.hx + P.lines - 1) / P.lines, sp.q.y);
  size_t lds = sizeof(float2) * (sp.q.z + (size_t)sp.q.z * (P.lines + 1));
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_z, grid, dim3(kThreads), lds, stream, P);
  return hipGetLastError();
}

bool z_grad_supported(const ShapePlan& sp) {
  // generic lengths: two line sets of sp.bz lines in LDS
  return fast_size(sp.q.z) || sizeof(float2) * (sp.q.z + 2 * (size_t)sp.q.z * (sp.bz + 1)) <= 64 * 1024;
}

hipError_t launch_z_grad(const ShapePlan& sp, float2* spec, const float2* what, const float2* khat, float2* corr,
                         int mode, int first, Int3 shift, int zlo, int zhi, int kz, const double* rsum,
                         float2* adj, bool adj_add, hipStream_t stream) {
  if (!z_grad_supported(sp) || zlo < 0 || zhi > sp.q.z || kz < 0 || kz > sp.q.z) return hipErrorInvalidValue;
  if (fast_size(sp.q.z))
    return launch_z_fast_grad(sp, spec, what, khat, corr, mode, first, shift, zlo, zhi, kz, rsum, adj, adj_add, stream);
  ZGradParams P{};
  P.rsum = rsum;
  P.adj = adj;
  P.adj_add = adj_add ? 1 : 0;
  P.spec = spec;
  P.what = what;
  P.khat = khat;
  P.corr = corr;
  P.mode = mode;
  P.first = first;
  P.shift = shift;
  P.q = sp.q;
  P.hx = sp.hx;
  P.pitch = sp.pitch;
  P.zlo = zlo;
  P.zhi = zhi;
  P.kz = kz;
  P.lines = sp.bz;
  P.plan = sp.pz;
  P.tw = sp.tw_z;
  P.twx = sp.tw_x;
  P.twy = sp.tw_y;
  dim3 grid((sp.hx + P.lines - 1) / P.lines, sp.q.y);
  size_t lds = sizeof(float2) * (sp.q.z + 2 * (size_t)sp.q.z * (P.lines + 1));
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_z_grad, grid, dim3(kThreads), lds, stream, P);
  return hipGetLastError();
}

int x_inverse_workgroups(const ShapePlan& sp, const ConvGeom& g) {
  const int nrows = g.eff_size.y * g.eff_size.z;
  const int rows = fast_size(g.q.x / 2) ? x_fast_rows_per_wg(g.q.x / 2) : sp.bx;
  return (nrows + rows - 1) / rows;
}

hipError_t launch_x_inverse(const ShapePlan& sp, const ConvGeom& g, const float2* spec, const EpilogueArgs& e,
                            hipStream_t stream, int* nwg) {
  XInvParams P;
  P.g = g;
  P.spec = spec;
  P.e = e;
  P.plan = sp.px;
  P.tw = sp.tw_x;
  P.hx = sp.hx;
  P.pitch = sp.pitch;
  P.lines = sp.bx;
  const int M = g.q.x / 2;
  const int n = x_inverse_workgroups(sp, g);
  if (nwg) *nwg = n;
  if (n == 0) return hipSuccess;
  if (fast_size(M)) return launch_x_inverse_fast(sp, g, spec, e, stream, n);
  size_t lds = sizeof(float2) * (g.q.x + (size_t)(M + 1) * (P.lines + 1));
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_x_inverse, dim3(n), dim3(kThreads), lds, stream, P);
  return hipGetLastError();
}

int conv_inverse_workgroups(const ShapePlan& sp, const ConvGeom& g) {
  return xy_fused(sp) ? g.eff_size.z : x_inverse_workgroups(sp, g);
}

hipError_t launch_xy_forward_passes(const ShapePlan& sp, const ConvGeom& g, const float* src, int gather,
                                    float2* spec, hipStream_t stream) {
  if (xy_fused(sp)) return launch_xy_forward(sp, g, src, gather, spec, stream);
  hipError_t e = launch_x_forward(sp, g, src, gather, spec, stream);
  if (e != hipSuccess) return e;
  return launch_y(sp, spec, 0, sp.q.z, 0, sp.q.y, -1.f, stream);
}

hipError_t launch_yx_inverse_passes(const ShapePlan& sp, const ConvGeom& g, float2* spec, const EpilogueArgs& e,
                                    hipStream_t stream, int* nwg) {
  if (xy_fused(sp)) {
    if (nwg) *nwg = conv_inverse_workgroups(sp, g);
    return launch_xy_inverse(sp, g, spec, e, stream);
  }
  const int zlo = g.eff_local.z, zhi = g.eff_local.z + g.eff_size.z;
  const int ylo = g.eff_local.y, yhi = g.eff_local.y + g.eff_size.y;
  hipError_t err = launch_y(sp, spec, zlo, zhi, ylo, yhi, 1.f, stream);
  if (err != hipSuccess) return err;
  return launch_x_inverse(sp, g, spec, e, stream, nwg);
}

// ---------------------------------------------------------------------------
// batched generic passes

int build_generic_params(const PassDesc& d, void* out, int* gx, int* gy, size_t* lds) {
  const ShapePlan& sp = *d.sp;
  const ConvGeom& g = d.g;
  switch (d.kind) {
    case kPassXForward: {
      XFwdParams P{};
      P.g = g;
      P.src = d.src;
      P.gather = d.gather;
      P.spec = d.spec;
      P.plan = sp.px;
      P.tw = sp.tw_x;
      P.hx = sp.hx;
      P.pitch = sp.pitch;
      P.lines = sp.bx;
      *gx = (g.q.y * g.q.z + P.lines - 1) / P.lines;
      *gy = 1;
      *lds = sizeof(float2) * (g.q.x + (size_t)(g.q.x / 2) * (P.lines + 1));
      std::memcpy(out, &P, sizeof(P));
      return sizeof(P);
    }
    case kPassYForward:
    case kPassYInverse: {
      YParams P{};
      P.spec = d.spec;
      P.hx = sp.hx;
      P.pitch = sp.pitch;
      P.qy = sp.q.y;
      P.zlo = d.zlo;
      P.ylo = d.ylo;
      P.yhi = d.yhi;
      P.lines = sp.by;
      P.dir = d.dir;
      P.plan = sp.py;
      P.tw = sp.tw_y;
      *gx = (sp.hx + P.lines - 1) / P.lines;
      *gy = d.zhi - d.zlo;
      *lds = sizeof(float2) * (sp.q.y + (size_t)sp.q.y * (P.lines + 1));
      std::memcpy(out, &P, sizeof(P));
      return sizeof(P);
    }
    case kPassZConv: {
      ZParams P{};
      P.aux = nullptr;
      P.pre = nullptr;
      P.spec = d.spec;
      P.kspec = d.kspec;
      P.scale = d.scale;
      P.dir = d.dir;
      P.hx = sp.hx;
      P.pitch = sp.pitch;
      P.qy = sp.q.y;
      P.qz = sp.q.z;
      P.zlo = d.zlo;
      P.zhi = d.zhi;
      P.lines = sp.bz;
      P.plan = sp.pz;
      P.tw = sp.tw_z;
      *gx = (sp.hx + P.lines - 1) / P.lines;
      *gy = sp.q.y;
      *lds = sizeof(float2) * (sp.q.z + (size_t)sp.q.z * (P.lines + 1));
      std::memcpy(out, &P, sizeof(P));
      return sizeof(P);
    }
    case kPassXInverse: {
      XInvParams P{};
      P.g = g;
      P.spec = d.spec;
      P.e = d.e;
      P.plan = sp.px;
      P.tw = sp.tw_x;
      P.hx = sp.hx;
      P.pitch = sp.pitch;
      P.lines = sp.bx;
      *gx = x_inverse_workgroups(sp, g);
      *gy = 1;
      *lds = sizeof(float2) * (g.q.x + (size_t)(g.q.x / 2 + 1) * (P.lines + 1));
      std::memcpy(out, &P, sizeof(P));
      return sizeof(P);
    }
  }
  return -1;
}

hipError_t launch_batch_generic(int kind, int epi, const void* items, const int* prefix, const int* gx, int nitems,
                                int nwg, size_t lds, hipStream_t stream) {
  (void)epi;
  if (nwg <= 0) return hipSuccess;
  (void)hipGetLastError();
  switch (kind) {
    case kPassXForward:
      hipLaunchKernelGGL(k_x_forward_batch, dim3(nwg), dim3(kThreads), lds, stream,
                         static_cast<const XFwdParams*>(items), prefix, gx, nitems);
      break;
    case kPassYForward:
    case kPassYInverse:
      hipLaunchKernelGGL(k_y_batch, dim3(nwg), dim3(kThreads), lds, stream, static_cast<const YParams*>(items),
                         prefix, gx, nitems);
      break;
    case kPassZConv:
      hipLaunchKernelGGL(k_z_batch, dim3(nwg), dim3(kThreads), lds, stream, static_cast<const ZParams*>(items),
                         prefix, gx, nitems);
      break;
    case kPassXInverse:
      hipLaunchKernelGGL(k_x_inverse_batch, dim3(nwg), dim3(kThreads), lds, stream,
                         static_cast<const XInvParams*>(items), prefix, gx, nitems);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mvd
