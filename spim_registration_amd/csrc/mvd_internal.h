// Internal declarations shared by the HIP translation units of libmvdecon_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "fft_device.h"

namespace mvd {

// How the R2C pass fills the FFT window from its source volume.
enum Gather : int {
  kGatherMirror = 0,      // mirror-single out-of-bounds (psi, MVDeconFFTThreads.java:44,78)
  kGatherOne = 1,         // constant 1 out of bounds (ratio, MVDeconFFT.java:515)
  kGatherZero = 2,        // zero out of bounds (compound kernels, MVDeconFFT.java:222-228)
  kGatherKernelWrap = 3,  // kernel wrapped so that its centre k/2 sits at index 0
};

// What the C2R pass does with the convolution result.
enum Epilogue : int {
  kEpiStore = 0,     // dst = result
  kEpiQuotient = 1,  // dst = img > 0 ? img / result : 1   (MVDeconvolution.java:494-546)
  kEpiUpdate = 2,    // psi update + statistics          (MVDeconvolution.java:603-724)
  kEpiResidual = 3,  // dst = w (1 - img / result), FP64 partials of w (result - img ln result)
};

// Weight source used by the update epilogue.
enum WeightMode : int {
  kWeightNone = 0,     // w = 1 (ProcessForDeconvolution.java:223-224)
  kWeightArray = 1,    // w = weight[i] (precomputed, already normalised)
  kWeightVirtual = 2,  // w = min(1, weight[i] / sum[i] * osem) in double (NormalizingRandomAccess.java:56-66)
};

struct Int3 {
  int x, y, z;
};

// Geometry of one FFT convolution job: a window of the source volume is
// convolved with circular wrap over the padded size q, and the sub-box
// [eff_local, eff_local + eff_size) of the window is written to the
// destination at eff_offset.
struct ConvGeom {
  Int3 src;         // source volume size
  Int3 off;         // window origin in source coordinates (may be negative)
  Int3 win;         // window size (data); [win, q) is zero padding
  Int3 q;           // FFT size (q.x even)
  Int3 eff_local;   // first output voxel inside the window
  Int3 eff_size;    // output box size
  Int3 eff_offset;  // output box origin in destination coordinates
  Int3 dst;         // destination volume size
};

// Device-resident plan of one FFT size: twiddle tables and radix plans.
struct ShapePlan {
  Int3 q;
  int hx;                    // q.x / 2 + 1 (half-spectrum row length)
  int pitch;                 // row stride of the spectrum in complex values (hx rounded up to 16)
  FftPlan px, py, pz;        // plans for lengths q.x/2, q.y, q.z
  float2* tw_x = nullptr;    // q.x entries: exp(-2 pi i m / q.x)
  float2* tw_y = nullptr;    // q.y entries
  float2* tw_z = nullptr;    // q.z entries
  int bx, by, bz;            // lines per workgroup for the x, y, z passes
  int64_t spectrum_elems() const { return (int64_t)pitch * q.y * q.z; }
};

// The dispatcher hands a launch's workgroups to the 8 XCDs in turn (workgroup
// w to XCD w mod 8) and each XCD has its own L2.  xcd_chunk maps the running
// workgroup w to the logical one it works on so that each XCD takes one
// contiguous eighth of the logical range, in order: logical neighbours run on
// one XCD at about the same time and share its L2 (lines that two workgroups
// each read half of are fetched once), and an XCD's accesses spread over all
// of its L2 channels instead of the few that a fixed stride of 8 workgroups
// selects.  The tail past the last whole round keeps its order.  A
// permutation of [0, count) for every count.
__device__ __forceinline__ int xcd_chunk(int w, int count) {
  const int per = count >> 3;
  if (w >= per * 8) return w;
  return (w & 7) * per + (w >> 3);
}

// Batched launches: workgroup wg (blockIdx.x, or its place in a reordered
// launch) belongs to the item i with prefix[i] <= wg < prefix[i + 1]; within
// the item it is workgroup (local % gx[i], local / gx[i]) of that item's own
// grid.
__device__ __forceinline__ void batch_locate(const int* prefix, const int* gx, int nitems, int& item, int& bx,
                                             int& by, int wg = blockIdx.x) {
  int lo = 0, hi = nitems;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= wg) lo = mid; else hi = mid;
  }
  const int local = wg - prefix[lo];
  item = lo;
  bx = local % gx[lo];
  by = local / gx[lo];
}

// Global-memory view of a pointer held in a parameter block: accesses through
// it compile to global_* instead of flat_* instructions.  Flat accesses also
// count against the LDS wait counter, so every LDS barrier would wait for
// them and no global load could stay in flight across a barrier.
// (The host pass of the single-source compile only parses device code; it
// keeps the generic pointer.)
template <class T>
__device__ __forceinline__ auto gmem(T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (__attribute__((address_space(1))) T*)p;
#else
  return p;
#endif
}

// Element at a 32-bit byte offset from a wave-uniform base: the compiler can
// keep the base in SGPRs and use the SGPR-base + 32-bit VGPR-offset global
// form, one VGPR and one add per address instead of 64-bit address pairs.
// The buffer behind `base` must span fewer than 2^32 bytes.
template <class T>
__device__ __forceinline__ auto gbyte(T* base, uint32_t byte_off) {
#if defined(__HIP_DEVICE_COMPILE__)
  using C = __attribute__((address_space(1))) char;
  using R = __attribute__((address_space(1))) T;
  return (R*)((C*)base + byte_off);
#else
  return (T*)((char*)base + byte_off);
#endif
}

// Buffer-resource access to a float2 array of `bytes` < 2^31 bytes: the base
// stays in SGPRs, each lane gives a 32-bit byte offset, and the hardware range
// check of raw buffers turns an access at offset >= bytes into a dropped store
// (a load returns zero).  kNoLane (2^31) is such an offset: masked lanes take
// it instead of a branch around the store, so every lane issues the same
// memory instructions.  Resource word 3 is the gfx9-family value for raw
// 32-bit-format buffers.
constexpr uint32_t kNoLane = 0x80000000u;
#if defined(__HIP_DEVICE_COMPILE__)
typedef __amdgpu_buffer_rsrc_t BufRes;
__device__ __forceinline__ BufRes buf_res(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
// soff: a wave-uniform byte offset (held in an SGPR) added to the lane's.
__device__ __forceinline__ float2 buf_load2(BufRes r, uint32_t byte_off, uint32_t soff = 0) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, soff, 0));
}
__device__ __forceinline__ void buf_store2(BufRes r, uint32_t byte_off, float2 v, uint32_t soff = 0) {
  typedef unsigned int U2 __attribute__((vector_size(8)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, byte_off, soff, 0);
}
// Streaming forms (cache policy nt): spectrum data that is read or written
// once per pass and is far larger than the caches.
constexpr int kCachePolicyNT = 2;
__device__ __forceinline__ float2 buf_load2_nt(BufRes r, uint32_t byte_off, uint32_t soff = 0) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, soff, kCachePolicyNT));
}
__device__ __forceinline__ void buf_store2_nt(BufRes r, uint32_t byte_off, float2 v, uint32_t soff = 0) {
  typedef unsigned int U2 __attribute__((vector_size(8)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, byte_off, soff, kCachePolicyNT);
}
#else
struct BufRes {
  char* p;
};
__device__ __forceinline__ BufRes buf_res(const void* base, uint32_t) { return BufRes{(char*)base}; }
__device__ __forceinline__ float2 buf_load2(BufRes r, uint32_t o, uint32_t so = 0) {
  return *(const float2*)(r.p + o + so);
}
__device__ __forceinline__ void buf_store2(BufRes r, uint32_t o, float2 v, uint32_t so = 0) {
  *(float2*)(r.p + o + so) = v;
}
__device__ __forceinline__ float2 buf_load2_nt(BufRes r, uint32_t o, uint32_t so = 0) { return buf_load2(r, o, so); }
__device__ __forceinline__ void buf_store2_nt(BufRes r, uint32_t o, float2 v, uint32_t so = 0) {
  buf_store2(r, o, v, so);
}
#endif

// Streaming (nontemporal) access to spectrum data: every pass reads and writes
// each spectrum value once and a view's block spectra are far larger than the
// caches, so they are kept from displacing data that is reused.
typedef float StreamF2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 stream_load(const float2* p) {
  const StreamF2 v = __builtin_nontemporal_load(reinterpret_cast<const StreamF2*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void stream_store(float2* p, float2 v) {
  __builtin_nontemporal_store(StreamF2{v.x, v.y}, reinterpret_cast<StreamF2*>(p));
}
// Column passes stream only when a workgroup's segment of a row covers whole
// 128-byte lines: nontemporal stores of half lines from two workgroups are
// not merged in L2 and reach HBM as read-modify-writes (measured: the
// 512-point z pass, 8 columns per workgroup, 4.2 -> 6.4 ms per launch).
template <bool ON>
__device__ __forceinline__ float2 maybe_stream_load(const float2* p) {
  if constexpr (ON) return stream_load(p);
  else return *p;
}
template <bool ON>
__device__ __forceinline__ void maybe_stream_store(float2* p, float2 v) {
  if constexpr (ON) stream_store(p, v);
  else *p = v;
}

FftPlan make_plan(int n);
inline int spectrum_pitch(int hx) { return (hx + 15) / 16 * 16; }
bool fft_size_supported(int n);
int next_fft_size(int n, bool even);

struct EpilogueArgs {
  int mode;              // Epilogue
  float* dst;            // store/quotient target, or psi for the update
  const float* img;      // observed view for the quotient
  const float* weight;   // weights for the update (WeightMode)
  const float* wsum;     // sum image for virtual weights
  int weight_mode;
  double osem;
  double lambda;
  double* part_sum;      // per-workgroup statistics of the update
  float* part_max;
  float wconst = 1.f;    // the weight of every voxel under kWeightNone
  // kEpiResidual, register x inverse only: instead of storing the residual,
  // transform its rows (the block's effective box placed in its window, zeros
  // elsewhere) along x into rspec (window rows, pitch rpitch) and write each
  // workgroup's residual sum (double) to part_rsum
  float2* rspec = nullptr;
  int rpitch = 0;
  double* part_rsum = nullptr;
};

// Launchers (all asynchronous on `stream`).
hipError_t launch_x_forward(const ShapePlan& sp, const ConvGeom& g, const float* src, int gather, float2* spec,
                            hipStream_t stream);
hipError_t launch_y(const ShapePlan& sp, float2* spec, int zlo, int zhi, int ylo, int yhi, float dir,
                    hipStream_t stream);
// kspec == nullptr: one z transform in direction `dir`, then multiply by `scale`
// (kernel spectra, gradient passes).  aux (with kspec): also store the forward
// spectrum before the product, all rows.  pre (with kspec): the lines are
// read from pre, already transformed along z (no forward transform); the
// output rows still go to spec.
hipError_t launch_z(const ShapePlan& sp, float2* spec, const float2* kspec, float scale, int zlo, int zhi,
                    hipStream_t stream, float dir = -1.f, float2* aux = nullptr, const float2* pre = nullptr);
// Gradient z pass (z_grad_supported): forward R^ of the
// y-transformed residual in spec; adjoint R^ conj(P^) inverted into rows
// [zlo, zhi) of spec; correlation conj(R^) W^ with the lag shift `shift`
// accumulated into corr (mode 0, first: overwrite) or inverted along z into
// rows [0, kz) of corr (mode 1); adj (nullable): the adjoint rows go there
// instead (added to what it holds when adj_add: the views' adjoints are
// linear and share the window, so their sum is transformed once).  rsum
// (nullable): the residual's sum in
// double, which replaces the zero-frequency value R^(0) of the transform (an
// FP32 transform of a residual of mixed signs loses that value's low bits,
// and the correlation multiplies it by the window's sum).
bool z_grad_supported(const ShapePlan& sp);
hipError_t launch_z_grad(const ShapePlan& sp, float2* spec, const float2* what, const float2* khat, float2* corr,
                         int mode, int first, Int3 shift, int zlo, int zhi, int kz, const double* rsum,
                         float2* adj, bool adj_add, hipStream_t stream);
// Returns the number of workgroups (for the update statistics) in *nwg.
hipError_t launch_x_inverse(const ShapePlan& sp, const ConvGeom& g, const float2* spec, const EpilogueArgs& e,
                            hipStream_t stream, int* nwg);
int x_inverse_workgroups(const ShapePlan& sp, const ConvGeom& g);

// The x and y passes of a convolution window, fused per plane when q.x = q.y =
// 256 (fft_xy.hip) and separate otherwise.  Forward: x then y over the whole
// box (gather rules mirror / one / zero).  Inverse: y over rows and planes of
// the effective box, then x with the epilogue; *nwg = number of statistics
// partials written (conv_inverse_workgroups).
hipError_t launch_xy_forward_passes(const ShapePlan& sp, const ConvGeom& g, const float* src, int gather,
                                    float2* spec, hipStream_t stream);
hipError_t launch_yx_inverse_passes(const ShapePlan& sp, const ConvGeom& g, float2* spec, const EpilogueArgs& e,
                                    hipStream_t stream, int* nwg);
int conv_inverse_workgroups(const ShapePlan& sp, const ConvGeom& g);

hipError_t launch_reduce_stats(const double* part_sum, const float* part_max, int n, double* out,
                               hipStream_t stream);
hipError_t launch_first_iteration(const float* views, int nviews, int64_t n, float* counts, double* part_sum,
                                  unsigned long long* part_cnt, int nparts, hipStream_t stream);
hipError_t launch_sum(const float* a, int64_t n, double* part_sum, int nparts, hipStream_t stream);
// Copies between pageable host arrays and device memory through pinned
// staging buffers, several host threads and overlapped DMA (staging.hip);
// they return when the copy is complete.
hipError_t staged_h2d(void* dst, const void* src, size_t bytes, hipStream_t stream);
hipError_t staged_d2h(void* dst, const void* src, size_t bytes, hipStream_t stream);
// Device to host that leaves `dst` untouched unless the whole copy succeeded.
hipError_t staged_d2h_all_or_nothing(void* dst, const void* src, size_t bytes, hipStream_t stream);
// Output path (export.hip): min/max with Math.min/max semantics, the 16-bit
// range widening and the TIFF stack writer (img in device or host memory).
hipError_t launch_min_max(const float* img, int64_t n, float* minmax, hipStream_t stream);
void u16_range(const float* minmax, double* out);
int write_tiff_stack(const float* img, bool on_device, int nx, int ny, int nz, int bits, double min, double max,
                     const double* calib, bool big_endian, int force_bigtiff, const char* path, hipStream_t stream);
hipError_t launch_fill(float* dst, float v, int64_t n, hipStream_t stream);
hipError_t launch_mask(const float* psi, const float* counts, float* out, int64_t n, hipStream_t stream);
hipError_t launch_weight_normalize(float* weights, int nviews, int64_t n, int64_t stride, float* wsum,
                                   int virtual_mode, int* part_min, unsigned long long* part_cnt, int nparts,
                                   hipStream_t stream);
hipError_t launch_osem_scale(float* weights, int64_t n, float osem, hipStream_t stream);

// Register-resident kernels for power-of-two lengths (fft_fast.hip).
bool fast_size(int n);
int x_fast_rows_per_wg(int M);
hipError_t launch_y_fast(const ShapePlan& sp, float2* spec, int zlo, int zhi, int ylo, int yhi, float dir,
                         hipStream_t stream);
// y pass whose input holds data only in rows [in_lo.y, in_hi.y) of planes
// [in_lo.z, in_hi.z); every other input row is read as zero (register kernels).
hipError_t launch_y_fast_window(const ShapePlan& sp, float2* spec, int zlo, int zhi, int ylo, int yhi, float dir,
                                Int3 in_lo, Int3 in_hi, hipStream_t stream);
hipError_t launch_z_fast(const ShapePlan& sp, float2* spec, const float2* kspec, float scale, int zlo, int zhi,
                         hipStream_t stream, float dir, float2* aux = nullptr, const float2* pre = nullptr);
hipError_t launch_z_fast_grad(const ShapePlan& sp, float2* spec, const float2* what, const float2* khat, float2* corr,
                              int mode, int first, Int3 shift, int zlo, int zhi, int kz, const double* rsum,
                              float2* adj, bool adj_add, hipStream_t stream);
hipError_t launch_x_forward_fast(const ShapePlan& sp, const ConvGeom& g, const float* src, int gather, float2* spec,
                                 hipStream_t stream);
hipError_t launch_x_inverse_fast(const ShapePlan& sp, const ConvGeom& g, const float2* spec, const EpilogueArgs& e,
                                 hipStream_t stream, int nwg);

// Fused x-y plane kernels (fft_xy.hip).
bool xy_fused(const ShapePlan& sp);
hipError_t launch_xy_forward(const ShapePlan& sp, const ConvGeom& g, const float* src, int gather, float2* spec,
                             hipStream_t stream);
hipError_t launch_xy_inverse(const ShapePlan& sp, const ConvGeom& g, const float2* spec, const EpilogueArgs& e,
                             hipStream_t stream);

// Batched passes (batch.hip): one launch runs the same pass for many items
// (blocks of many stacks), each item with its own geometry and buffers.
enum PassKind : int { kPassXForward = 0, kPassYForward, kPassZConv, kPassYInverse, kPassXInverse };
struct PassDesc {
  int kind;
  const ShapePlan* sp;
  ConvGeom g;
  const float* src;     // x forward source
  int gather;           // x forward out-of-bounds rule
  float2* spec;         // block spectrum
  const float2* kspec;  // z pass kernel spectrum
  EpilogueArgs e;       // x inverse epilogue
  int zlo, zhi, ylo, yhi;
  float dir, scale;
  int xy;  // x passes of a fused x-y item (no separate y pass)
  // z pass in the lines kernel only: input planes through a table of plane
  // indices into the arena of spectra (nullptr: plane z of spec)
  const unsigned short* ztab = nullptr;
  const float2* arena = nullptr;
  unsigned arena_bytes = 0;
};
constexpr int kMaxParamBytes = 640;
bool pass_is_fast(const PassDesc& d);
// Fill the kernel parameter block of one item; returns its size in bytes and
// the item's own grid (gx, gy); *lds = dynamic LDS bytes (generic) and
// *cls = template selector (fast).
int build_generic_params(const PassDesc& d, void* out, int* gx, int* gy, size_t* lds);
int build_fast_params(const PassDesc& d, void* out, int* gx, int* gy, int* cls);
hipError_t launch_batch_generic(int kind, int epi, const void* items, const int* prefix, const int* gx, int nitems,
                                int nwg, size_t lds, hipStream_t stream);
// tasks (x and y forward only, optional): one (item, plane) per wpt
// workgroups; only the listed planes of the items are transformed.
hipError_t launch_batch_fast(int kind, int cls, const void* items, const int* prefix, const int* gx, int nitems,
                             int nwg, hipStream_t stream, const int2* tasks = nullptr, int wpt = 0);
// z passes of many blocks sharing one kernel spectrum and line grid (gx, gy);
// table: the items' input planes come through their plane tables, at most
// ztab_entries(n) entries for the launch (1: the arena spans fewer than 2^31
// bytes, 32-bit offsets; 2: 64-bit addresses).  The table lives in LDS beside
// the lines: 512-point lines take up to 18 blocks of 512 planes, the shorter
// ones 6144 entries (which keeps three 256-point workgroups per CU).
constexpr int ztab_entries(int n) { return n >= 512 ? 9216 : 6144; }
hipError_t launch_z_lines(int n, const void* items, int nitems, int gx, int gy, hipStream_t stream, int table);
// One workgroup of a fused forward launch with a task list: transform plane
// `plane` of item `item` and store the spectrum into plane dest[d][1] of the
// spectrum of item dest[d][0], d < ndest (identical planes of several block
// windows are transformed once).
constexpr int kMaxPlaneDests = 8;
struct PlaneTask {
  int item, plane, ndest, pad;
  int dest[kMaxPlaneDests][2];
};
// Fused x-y items: kind kPassXForward (cls = gather) or kPassXInverse (cls = epilogue).
// tasks (forward only, optional): one PlaneTask per workgroup.
int build_xy_params(const PassDesc& d, void* out, int* gx, int* gy, int* cls);
hipError_t launch_batch_xy(int kind, int cls, const void* items, const int* prefix, const int* gx, int nitems,
                           int nwg, hipStream_t stream, const PlaneTask* tasks = nullptr);
// Task list of a fused forward launch (see session.hip).
std::vector<PlaneTask> forward_plane_tasks(const std::vector<const PassDesc*>& items, bool dedup);
struct ReduceItem {
  const double* part_sum;
  const float* part_max;
  int n;
  double* out;
};
hipError_t launch_reduce_stats_batch(const ReduceItem* items, int nitems, hipStream_t stream);

// Gradient helpers (gradient.hip).
// out = op(a) * op(b) over n spectrum values, op = conjugate when the flag is set.
hipError_t launch_spec_product(float2* out, const float2* a, const float2* b, int conj_a, int conj_b, int64_t n,
                               hipStream_t stream);
// out = conj(a) b when first, else out += conj(a) b.
hipError_t launch_spec_corr_acc(float2* out, const float2* a, const float2* b, int first, int64_t n,
                                hipStream_t stream);
// grad[j] = scale * box[kd - 1 - j] per axis: the kernel-sized box of a
// correlation whose lags were shifted by (kd - 1) / 2 (launch_z_grad), so that
// lag kd / 2 - j lands on kd - 1 - j for odd and even kd.
hipError_t launch_extract_psf_grad_shifted(const float* box, Int3 kd, float scale, float* grad, hipStream_t stream);
// grad[j] = scale * box[(c - j) mod q] for the kernel box kd (x-fastest).
hipError_t launch_extract_psf_grad(const float* box, Int3 q, Int3 kd, float scale, float* grad, int accumulate,
                                   hipStream_t stream);
hipError_t launch_crop(float* dst, const float* src, Int3 n, Int3 origin, Int3 size, hipStream_t stream);
// acc[mirror(lo + u)] += win[u] over the window positions u in [0, wd) (the
// adjoint of the mirror-single window gather; atomics for the mirrored shell).
hipError_t launch_fold_window(const float* win, Int3 wd, Int3 lo, Int3 n, float* acc, hipStream_t stream);
// The same fold as a gather over the image voxels (no atomics), for windows
// that extend past each image face by less than n - 1 (fold_gather_ok): the
// fold is added to acc (add) or written to it.
bool fold_gather_ok(Int3 wd, Int3 lo, Int3 n);
// The window's direct box is the whole image.
bool fold_gather_covers(Int3 wd, Int3 lo, Int3 n);
hipError_t launch_fold_gather(const float* win, Int3 wd, Int3 lo, Int3 n, float* acc, bool add, hipStream_t stream);
// *out = sum of the partials part[0, n) (one workgroup, a fixed order).
hipError_t launch_reduce_sum(const double* part, int n, double* out, hipStream_t stream);
// *out = sum of a[0, n) in double, through nparts partials in part.
hipError_t launch_sum_to(const float* a, int64_t n, double* part, int nparts, double* out, hipStream_t stream);
// spec[0] = (float) *v (the zero-frequency value of a spectrum).
hipError_t launch_set_dc(float2* spec, const double* v, hipStream_t stream);

// Resampling of one view into the bounding box with its blending weight
// (transform.hip).  inv: row-packed 3x4 inverse affine map (bounding-box
// coordinates -> raw view); off: bounding-box min as float; out_img / out_w
// nullable; lut: the blending table in device memory (blend_lut_device).
constexpr int kBlendLutSize = 1001;
struct TransformArgs {
  const float* raw;
  int rsize[3];
  double inv[12];
  float off[3];
  int nx, ny, nz;
  float border[3], range[3];
  float* out_img;
  float* out_w;
  const double* lut;
};
const double* blend_lut_device(int device);  // per device, held until the process ends
const char* check_transform_args(const TransformArgs& a);  // nullptr when valid
hipError_t launch_transform_view(const TransformArgs& a, hipStream_t stream);

// Plan cache and kernel-spectrum helpers (session.cpp).
void set_error(const std::string& msg);

}  // namespace mvd
