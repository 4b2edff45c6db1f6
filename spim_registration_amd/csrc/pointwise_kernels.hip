// Pointwise and reduction kernels of the deconvolution session: first
// iteration (initial estimate), final mask, statistics reduction, weight
// normalisation.  All are single HBM passes with grid-stride loops; every
// reduction writes one partial per workgroup so results do not depend on
// scheduling.
#include "mvd_internal.h"
#include "pointwise_math.h"

namespace mvd {

constexpr int kPwThreads = 256;

template <class T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kPwThreads / 64; ++i) s += red[i];
  return s;
}

// FirstIteration.call (deconv/FirstIteration.java:84-192): per voxel the
// number c of views with a value > 0 and their mean (double); counts[i] = c.
// Partials: sum of the means and the number of voxels with c > 0.
__global__ __launch_bounds__(kPwThreads) void k_first_iteration(const float* views, int nviews, int64_t n,
                                                               float* counts, double* part_sum,
                                                               unsigned long long* part_cnt) {
  __shared__ double rs[kPwThreads / 64];
  __shared__ unsigned long long rc[kPwThreads / 64];
  double s = 0.0;
  unsigned long long c = 0;
  for (int64_t i = blockIdx.x * (int64_t)kPwThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPwThreads) {
    double sum = 0.0;
    int cnt = 0;
    for (int v = 0; v < nviews; ++v) {
      const double x = (double)views[(int64_t)v * n + i];
      if (x > 0) {
        sum += x;
        ++cnt;
      }
    }
    if (cnt > 0) {
      s += sum / cnt;
      ++c;
    }
    if (counts) counts[i] = (float)cnt;
  }
  const double bs = block_sum(s, rs);
  const unsigned long long bc = block_sum(c, rc);
  if (threadIdx.x == 0) {
    part_sum[blockIdx.x] = bs;
    part_cnt[blockIdx.x] = bc;
  }
}

hipError_t launch_first_iteration(const float* views, int nviews, int64_t n, float* counts, double* part_sum,
                                  unsigned long long* part_cnt, int nparts, hipStream_t stream) {
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_first_iteration, dim3(nparts), dim3(kPwThreads), 0, stream, views, nviews, n, counts,
                     part_sum, part_cnt);
  return hipGetLastError();
}

// Sum of a float array in double, one partial per workgroup.
__global__ __launch_bounds__(kPwThreads) void k_sum(const float* a, int64_t n, double* part_sum) {
  __shared__ double rs[kPwThreads / 64];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kPwThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPwThreads)
    s += (double)a[i];
  const double bs = block_sum(s, rs);
  if (threadIdx.x == 0) part_sum[blockIdx.x] = bs;
}

hipError_t launch_sum(const float* a, int64_t n, double* part_sum, int nparts, hipStream_t stream) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_sum, dim3(nparts), dim3(kPwThreads), 0, stream, a, n, part_sum);
  return hipGetLastError();
}

__global__ __launch_bounds__(kPwThreads) void k_fill(float* dst, float v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kPwThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPwThreads)
    dst[i] = v;
}

hipError_t launch_fill(float* dst, float v, int64_t n, hipStream_t stream) {
  const int64_t blocks = std::min<int64_t>((n + kPwThreads - 1) / kPwThreads, 8192);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(kPwThreads), 0, stream, dst, v, n);
  return hipGetLastError();
}

// MVDeconvolution.java:201-208: psi = 0 where no view has data.
__global__ __launch_bounds__(kPwThreads) void k_mask(const float* psi, const float* counts, float* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kPwThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPwThreads)
    out[i] = counts[i] == 0.f ? 0.f : psi[i];
}

hipError_t launch_mask(const float* psi, const float* counts, float* out, int64_t n, hipStream_t stream) {
  const int64_t blocks = std::min<int64_t>((n + kPwThreads - 1) / kPwThreads, 8192);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_mask, dim3((unsigned)blocks), dim3(kPwThreads), 0, stream, psi, counts, out, n);
  return hipGetLastError();
}

// Sum of the per-workgroup change sums (fixed order per thread, then a fixed
// tree) and max of the maxima -> out[0], out[1].
__global__ __launch_bounds__(kPwThreads) void k_reduce_stats(const double* part_sum, const float* part_max, int n,
                                                            double* out) {
  __shared__ double rs[kPwThreads];
  __shared__ float rm[kPwThreads];
  double s = 0.0;
  float m = -1.f;
  for (int i = threadIdx.x; i < n; i += kPwThreads) {
    s += part_sum[i];
    m = java_maxf(m, part_max[i]);
  }
  rs[threadIdx.x] = s;
  rm[threadIdx.x] = m;
  __syncthreads();
  for (int o = kPwThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      rs[threadIdx.x] += rs[threadIdx.x + o];
      rm[threadIdx.x] = java_maxf(rm[threadIdx.x], rm[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = rs[0];
    out[1] = (double)rm[0];
  }
}

hipError_t launch_reduce_stats(const double* part_sum, const float* part_max, int n, double* out,
                               hipStream_t stream) {
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_reduce_stats, dim3(1), dim3(kPwThreads), 0, stream, part_sum, part_max, n, out);
  return hipGetLastError();
}

// One workgroup per item: the statistics of one (stack, view) of a batched
// iteration, reduced in the same order as k_reduce_stats.
__global__ __launch_bounds__(kPwThreads) void k_reduce_stats_batch(const ReduceItem* items, int nitems) {
  __shared__ double rs[kPwThreads];
  __shared__ float rm[kPwThreads];
  const ReduceItem it = items[blockIdx.x];
  double s = 0.0;
  float m = -1.f;
  for (int i = threadIdx.x; i < it.n; i += kPwThreads) {
    s += it.part_sum[i];
    m = java_maxf(m, it.part_max[i]);
  }
  rs[threadIdx.x] = s;
  rm[threadIdx.x] = m;
  __syncthreads();
  for (int o = kPwThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      rs[threadIdx.x] += rs[threadIdx.x + o];
      rm[threadIdx.x] = java_maxf(rm[threadIdx.x], rm[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    it.out[0] = rs[0];
    it.out[1] = (double)rm[0];
  }
}

hipError_t launch_reduce_stats_batch(const ReduceItem* items, int nitems, hipStream_t stream) {
  if (nitems <= 0) return hipSuccess;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_reduce_stats_batch, dim3(nitems), dim3(kPwThreads), 0, stream, items, nitems);
  return hipGetLastError();
}

// WeightNormalizer (deconv/WeightNormalizer.java:138-253).  Per voxel the
// double sum of the view weights and the number of views with weight > 0.
// Precomputed weights (ApplyDirectly, :138-185): w_v = (float)(w_v / sum),
// with no guard against sum == 0.  Virtual weights (ComputeSumImage,
// :187-253): wsum = sum > 1 ? (float)sum : 1, weights unchanged.
// weights of view v start at weights + v * stride.  Partials: minimum count and
// the total count.
__global__ __launch_bounds__(kPwThreads) void k_weight_normalize(float* weights, int nviews, int64_t n,
                                                                int64_t stride, float* wsum,
                                                                int virtual_mode, int* part_min,
                                                                unsigned long long* part_cnt) {
  __shared__ unsigned long long rc[kPwThreads / 64];
  __shared__ int rmin[kPwThreads / 64];
  int mn = nviews;
  unsigned long long cnt_all = 0;
  for (int64_t i = blockIdx.x * (int64_t)kPwThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPwThreads) {
    double sum = 0.0;
    int cnt = 0;
    for (int v = 0; v < nviews; ++v) {
      const float w = weights[(int64_t)v * stride + i];
      sum += w;
      if (w > 0) ++cnt;
    }
    cnt_all += cnt;
    mn = min(mn, cnt);
    if (virtual_mode) {
      wsum[i] = sum > 1 ? (float)sum : 1.f;
    } else {
      for (int v = 0; v < nviews; ++v) {
        float* p = weights + (int64_t)v * stride + i;
        *p = (float)((double)*p / sum);
      }
    }
  }
  const unsigned long long bc = block_sum(cnt_all, rc);
  for (int o = 32; o > 0; o >>= 1) mn = min(mn, __shfl_xor(mn, o));
  if ((threadIdx.x & 63) == 0) rmin[threadIdx.x / 64] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = nviews;
    for (int i = 0; i < kPwThreads / 64; ++i) m = min(m, rmin[i]);
    part_min[blockIdx.x] = m;
    part_cnt[blockIdx.x] = bc;
  }
}

hipError_t launch_weight_normalize(float* weights, int nviews, int64_t n, int64_t stride, float* wsum,
                                   int virtual_mode, int* part_min, unsigned long long* part_cnt, int nparts,
                                   hipStream_t stream) {
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_weight_normalize, dim3(nparts), dim3(kPwThreads), 0, stream, weights, nviews, n, stride, wsum,
                     virtual_mode, part_min, part_cnt);
  return hipGetLastError();
}

// ProcessForDeconvolution.adjustForOSEM (deconv/ProcessForDeconvolution.java:372-389):
// w = Math.min(1, w * (float)osem) in float.  A weight that normalisation left
// NaN (a voxel no view weights, 0/0) stays NaN, as Java's Math.min keeps it.
__global__ __launch_bounds__(kPwThreads) void k_osem_scale(float* w, int64_t n, float osem) {
#pragma clang fp contract(off)
  for (int64_t i = blockIdx.x * (int64_t)kPwThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPwThreads)
    w[i] = java_minf(1.f, w[i] * osem);
}

hipError_t launch_osem_scale(float* weights, int64_t n, float osem, hipStream_t stream) {
  const int64_t blocks = std::min<int64_t>((n + kPwThreads - 1) / kPwThreads, 8192);
  if (blocks == 0) return hipSuccess;
  (void)hipGetLastError();  // launch errors only, not a stale one
  hipLaunchKernelGGL(k_osem_scale, dim3((unsigned)blocks), dim3(kPwThreads), 0, stream, weights, n, osem);
  return hipGetLastError();
}

}  // namespace mvd
