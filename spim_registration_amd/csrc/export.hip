// Output path of the deconvolved image.
//
//   * FusionHelper.minMax (process/fusion/FusionHelper.java:108-172): min and
//     max of a float image with java.lang.Math.min/max semantics (NaN wins,
//     -0.0 < +0.0), started at +/-Float.MAX_VALUE;
//   * the 16-bit conversion of the HDF5 export
//     (export/ExportSpimData2HDF5.java:227-262): the min/max widening of
//     updateAndGetMinMax and imglib2's RealUnsignedShortConverter (imglib2
//     6.1.0, not in the reference tree): u = min(65535, (int)(max(0, (a - min)
//     / (max - min) * 65535) + 0.5)) in double;
//   * the TIFF stack writer in the role of Save3dTIFF.exportImage
//     (export/Save3dTIFF.java:79-133, ImageJ FileSaver.saveAsTiffStack):
//     one uncompressed page per z slice, an ImageJ description with the
//     display range and calibration, big-endian by default like ImageJ.
//
// The conversion and the byte-order packing run on the GPU in chunks; each
// chunk is copied to pinned host memory and written while the next one is
// converted.  Every pass is one read and one write of the image in HBM.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mvd_internal.h"
#include "mvdecon.h"

#pragma clang fp contract(off)

namespace mvd {
namespace {

constexpr int kExThreads = 256;
constexpr int kMinMaxParts = 1024;
constexpr int kMinMaxUnroll = 2;  // 16-byte loads in flight per thread
constexpr int kStreamGrid = 65536;  // grid cap of the streaming passes
constexpr int64_t kChunkBytes = 64ll << 20;  // output bytes per pipeline stage

// Order-preserving integer key of a float: signed integer order of the keys
// is -NaN < -inf < ... < -0.0 < +0.0 < ... < +inf < +NaN, so integer min/max
// gives Math.min/max's order of signed zeros, and a NaN anywhere leaves a key
// beyond an infinity in the min or the max.
__host__ __device__ inline int32_t order_key(uint32_t bits) {
  return (int32_t)(bits ^ ((uint32_t)((int32_t)bits >> 31) & 0x7fffffffu));
}

__host__ __device__ inline uint32_t key_bits(int32_t key) {
  return (uint32_t)key ^ ((uint32_t)(key >> 31) & 0x7fffffffu);
}

// Per-workgroup min and max keys; the host combines the partials (integer
// min/max: the result does not depend on the order).
__global__ __launch_bounds__(kExThreads) void k_min_max(const float* __restrict__ img, int64_t n,
                                                        int2* __restrict__ part) {
  __shared__ int2 red[kExThreads / 64];
  int32_t kmn = INT32_MAX, kmx = INT32_MIN;
  const int64_t stride = (int64_t)gridDim.x * kExThreads;
  auto take = [&](float v) {
    const int32_t k = order_key(__float_as_uint(v));
    kmn = min(kmn, k);
    kmx = max(kmx, k);
  };
  const bool vec = ((uintptr_t)img & 15) == 0;
  int64_t done = 0;
  if (vec) {
    const int64_t n4 = n >> 2;
    const float4* p = reinterpret_cast<const float4*>(img);
    // kMinMaxUnroll loads in flight per thread
    for (int64_t i0 = blockIdx.x * (int64_t)kExThreads + threadIdx.x; i0 < n4; i0 += stride * kMinMaxUnroll) {
      float4 v[kMinMaxUnroll];
#pragma unroll
      for (int j = 0; j < kMinMaxUnroll; ++j) {
        const int64_t i = i0 + j * stride;
        if (i < n4) v[j] = p[i];
      }
#pragma unroll
      for (int j = 0; j < kMinMaxUnroll; ++j) {
        if (i0 + j * stride >= n4) break;
        take(v[j].x);
        take(v[j].y);
        take(v[j].z);
        take(v[j].w);
      }
    }
    done = n4 << 2;
  }
  for (int64_t i = done + blockIdx.x * (int64_t)kExThreads + threadIdx.x; i < n; i += stride) take(img[i]);
  for (int o = 32; o > 0; o >>= 1) {
    kmn = min(kmn, __shfl_xor(kmn, o));
    kmx = max(kmx, __shfl_xor(kmx, o));
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = make_int2(kmn, kmx);
  __syncthreads();
  if (threadIdx.x == 0) {
    int2 r = red[0];
    for (int w = 1; w < kExThreads / 64; ++w) {
      r.x = min(r.x, red[w].x);
      r.y = max(r.y, red[w].y);
    }
    part[blockIdx.x] = r;
  }
}

// RealUnsignedShortConverter.convert on one value (double arithmetic).
__host__ __device__ inline uint16_t to_u16(float a, double min, double scale) {
  double x = ((double)a - min) / scale * 65535.0;
  x = (x != x) ? x : (x > 0.0 ? x : 0.0);  // Math.max(0, x): NaN stays NaN
  if (x != x) return 0;                     // (int) NaN = 0
  if (x >= 65535.0) return 65535;           // Math.min(65535, ...), (int) saturates
  return (uint16_t)(int)(x + 0.5);
}

__device__ __forceinline__ uint16_t bswap16(uint16_t v) { return (uint16_t)((v >> 8) | (v << 8)); }

// Four values per thread and iteration when both pointers allow it (a
// large grid keeps enough loads in flight).
template <bool BE>
__global__ __launch_bounds__(kExThreads) void k_to_u16(const float* __restrict__ img, int64_t n, double min,
                                                       double scale, uint16_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * kExThreads;
  const bool vec = ((uintptr_t)img & 15) == 0 && ((uintptr_t)out & 7) == 0;
  int64_t done = 0;
  if (vec) {
    const int64_t n4 = n >> 2;
    const float4* p = reinterpret_cast<const float4*>(img);
    uint2* q = reinterpret_cast<uint2*>(out);
    for (int64_t i = blockIdx.x * (int64_t)kExThreads + threadIdx.x; i < n4; i += stride) {
      const float4 v = p[i];
      uint16_t u[4] = {to_u16(v.x, min, scale), to_u16(v.y, min, scale), to_u16(v.z, min, scale),
                       to_u16(v.w, min, scale)};
      if (BE)
        for (int k = 0; k < 4; ++k) u[k] = bswap16(u[k]);
      q[i] = make_uint2((uint32_t)u[0] | ((uint32_t)u[1] << 16), (uint32_t)u[2] | ((uint32_t)u[3] << 16));
    }
    done = n4 << 2;
  }
  for (int64_t i = done + blockIdx.x * (int64_t)kExThreads + threadIdx.x; i < n; i += stride) {
    const uint16_t u = to_u16(img[i], min, scale);
    out[i] = BE ? bswap16(u) : u;
  }
}

// Big-endian float32: the bytes of every value reversed.
__global__ __launch_bounds__(kExThreads) void k_swap32(const uint32_t* __restrict__ in, int64_t n,
                                                       uint32_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * kExThreads;
  const bool vec = ((uintptr_t)in & 15) == 0 && ((uintptr_t)out & 15) == 0;
  int64_t done = 0;
  if (vec) {
    const int64_t n4 = n >> 2;
    const uint4* p = reinterpret_cast<const uint4*>(in);
    uint4* q = reinterpret_cast<uint4*>(out);
    for (int64_t i = blockIdx.x * (int64_t)kExThreads + threadIdx.x; i < n4; i += stride) {
      const uint4 v = p[i];
      q[i] = make_uint4(__builtin_bswap32(v.x), __builtin_bswap32(v.y), __builtin_bswap32(v.z),
                        __builtin_bswap32(v.w));
    }
    done = n4 << 2;
  }
  for (int64_t i = done + blockIdx.x * (int64_t)kExThreads + threadIdx.x; i < n; i += stride)
    out[i] = __builtin_bswap32(in[i]);
}

int grid_for(int64_t work, int cap = kStreamGrid) {
  const int64_t g = (work + kExThreads - 1) / kExThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

struct DeviceScope {
  int prev = 0;
  bool ok = false;
  explicit DeviceScope(int device) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return;
    (void)hipGetDevice(&prev);
    ok = hipSetDevice(device) == hipSuccess;
  }
  ~DeviceScope() {
    if (ok) (void)hipSetDevice(prev);
  }
};

int fail(const std::string& m, int code) {
  set_error(m);
  return code;
}

// ---------------------------------------------------------------- TIFF layout

struct TiffWriter {
  bool big_endian = true;
  bool bigtiff = false;
  std::vector<uint8_t> buf;

  void u16(uint16_t v) {
    if (big_endian) {
      buf.push_back(v >> 8);
      buf.push_back(v & 255);
    } else {
      buf.push_back(v & 255);
      buf.push_back(v >> 8);
    }
  }
  void u32(uint32_t v) {
    if (big_endian) {
      u16(v >> 16);
      u16(v & 0xffff);
    } else {
      u16(v & 0xffff);
      u16(v >> 16);
    }
  }
  void u64(uint64_t v) {
    if (big_endian) {
      u32(v >> 32);
      u32(v & 0xffffffffu);
    } else {
      u32(v & 0xffffffffu);
      u32(v >> 32);
    }
  }
  void off(uint64_t v) { bigtiff ? u64(v) : u32((uint32_t)v); }
};

enum TiffType : uint16_t { kAscii = 2, kShort = 3, kLong = 4, kRational = 5, kLong8 = 16 };

struct Entry {
  uint16_t tag, type;
  uint64_t count, value;  // value: inline number or offset
};

// Java's Double.toString for the values written here is close to the
// shortest representation that reads back exactly.
std::string num(double v) {
  if (std::isnan(v)) return "NaN";
  if (v == std::floor(v) && std::fabs(v) < 1e7) {
    char b[64];
    snprintf(b, sizeof(b), "%.1f", v);
    return b;
  }
  for (int p = 1; p <= 17; ++p) {
    char b[64];
    snprintf(b, sizeof(b), "%.*g", p, v);
    if (strtod(b, nullptr) == v) return b;
  }
  return "0";
}

// ImageJ description of a z stack (the fields FileSaver writes for one
// channel and one frame).
std::string imagej_description(int nz, double dmin, double dmax, const double* calib) {
  std::string d = "ImageJ=1.54f\nimages=" + std::to_string(nz) + "\nslices=" + std::to_string(nz) + "\n";
  if (calib) {
    d += "unit=pixel\nspacing=" + num(calib[3]) + "\n";
  }
  d += "loop=false\nmin=" + num(dmin) + "\nmax=" + num(dmax) + "\n";
  if (calib) {
    if (calib[0] != 0) d += "xorigin=" + num(calib[0]) + "\n";
    if (calib[1] != 0) d += "yorigin=" + num(calib[1]) + "\n";
    if (calib[2] != 0) d += "zorigin=" + num(calib[2]) + "\n";
  }
  return d;
}

// Header, every IFD and the description in front of the pixel data; page z
// is the single strip at data_offset + z * page_bytes.
std::vector<uint8_t> tiff_prefix(int nx, int ny, int nz, int bits, bool big_endian, bool bigtiff,
                                 const std::string& desc, const double* calib, uint64_t* data_offset) {
  TiffWriter w;
  w.big_endian = big_endian;
  w.bigtiff = bigtiff;
  const uint64_t page = (uint64_t)nx * ny * (bits / 8);
  const int nent0 = calib ? 15 : 12;  // first IFD: description and resolution
  const int nent = calib ? 14 : 11;
  const uint64_t esz = bigtiff ? 20 : 12, csz = bigtiff ? 8 : 2, osz = bigtiff ? 8 : 4;
  const uint64_t ifd0 = bigtiff ? 16 : 8;
  const uint64_t ifd0_size = csz + nent0 * esz + osz;
  const uint64_t ifd_size = csz + nent * esz + osz;
  const uint64_t desc_off = ifd0 + ifd0_size;
  const uint64_t desc_len = desc.size() + 1;
  const uint64_t res_off = desc_off + desc_len + (desc_len & 1);
  // BigTIFF keeps 8-byte values (a rational) in the entry itself
  const uint64_t ifds_off = res_off + (calib && !bigtiff ? 16 : 0);
  const uint64_t data = ifds_off + (uint64_t)(nz - 1) * ifd_size;
  *data_offset = data;
  // resolution: 1 / pixel size as a rational (1e6 / (pixel size * 1e6))
  uint32_t rnum = 1, rden = 1;
  if (calib) {
    const double ps = calib[3];
    rnum = 1000000u;
    rden = (uint32_t)std::max(1.0, std::min(4294967295.0, std::floor(ps * 1e6 + 0.5)));
  }
  if (bigtiff) {
    w.u16(big_endian ? 0x4d4d : 0x4949);
    w.u16(43);
    w.u16(8);
    w.u16(0);
    w.u64(ifd0);
  } else {
    w.u16(big_endian ? 0x4d4d : 0x4949);
    w.u16(42);
    w.u32((uint32_t)ifd0);
  }
  auto ifd = [&](int z, uint64_t next) {
    std::vector<Entry> e;
    e.push_back({254, kLong, 1, 0});
    e.push_back({256, kLong, 1, (uint64_t)nx});
    e.push_back({257, kLong, 1, (uint64_t)ny});
    e.push_back({258, kShort, 1, (uint64_t)bits});
    e.push_back({259, kShort, 1, 1});
    e.push_back({262, kShort, 1, 1});
    if (z == 0) e.push_back({270, kAscii, desc_len, desc_off});
    e.push_back({273, (uint16_t)(bigtiff ? kLong8 : kLong), 1, data + (uint64_t)z * page});
    e.push_back({277, kShort, 1, 1});
    e.push_back({278, kLong, 1, (uint64_t)ny});
    e.push_back({279, (uint16_t)(bigtiff ? kLong8 : kLong), 1, page});
    if (calib) {
      e.push_back({282, kRational, 1, res_off});
      e.push_back({283, kRational, 1, res_off + 8});
      e.push_back({296, kShort, 1, 1});
    }
    e.push_back({339, kShort, 1, (uint64_t)(bits == 32 ? 3 : 1)});
    bigtiff ? w.u64(e.size()) : w.u16((uint16_t)e.size());
    for (const Entry& x : e) {
      w.u16(x.tag);
      w.u16(x.type);
      bigtiff ? w.u64(x.count) : w.u32((uint32_t)x.count);
      // values are left-justified in the value field
      const bool inline_val = x.type == kShort || x.type == kLong || x.type == kLong8;
      if (inline_val && x.type == kShort) {
        w.u16((uint16_t)x.value);
        w.u16(0);
        if (bigtiff) w.u32(0);
      } else if (inline_val && x.type == kLong) {
        w.u32((uint32_t)x.value);
        if (bigtiff) w.u32(0);
      } else if (x.type == kRational && bigtiff) {
        w.u32(rnum);
        w.u32(rden);
      } else {
        w.off(x.value);
      }
    }
    w.off(next);
  };
  ifd(0, nz > 1 ? ifds_off : 0);
  for (char c : desc) w.buf.push_back((uint8_t)c);
  w.buf.push_back(0);
  if (desc_len & 1) w.buf.push_back(0);
  if (calib && !bigtiff) {
    w.u32(rnum);
    w.u32(rden);
    w.u32(rnum);
    w.u32(rden);
  }
  for (int z = 1; z < nz; ++z) ifd(z, z + 1 < nz ? ifds_off + (uint64_t)z * ifd_size : 0);
  if (w.buf.size() != data) {
    set_error("internal TIFF layout error");
    return {};
  }
  return std::move(w.buf);
}

bool write_all(int fd, const void* p, size_t bytes) {
  const char* c = static_cast<const char*>(p);
  while (bytes > 0) {
    const ssize_t r = ::write(fd, c, std::min<size_t>(bytes, 1u << 30));
    if (r <= 0) return false;
    c += r;
    bytes -= (size_t)r;
  }
  return true;
}

}  // namespace

hipError_t launch_min_max(const float* img, int64_t n, float* minmax, hipStream_t stream) {
  const int parts = grid_for((n + 3) / 4, kMinMaxParts);
  int2* d_part = nullptr;
  hipError_t e = hipMalloc(&d_part, sizeof(int2) * parts);
  if (e != hipSuccess) return e;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_min_max, dim3(parts), dim3(kExThreads), 0, stream, img, n, d_part);
  e = hipGetLastError();
  std::vector<int2> h(parts);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_part, sizeof(int2) * parts, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(d_part);
  if (e != hipSuccess) return e;
  int32_t kmn = INT32_MAX, kmx = INT32_MIN;
  for (const int2& p : h) {
    kmn = std::min(kmn, p.x);
    kmx = std::max(kmx, p.y);
  }
  const float fmax = 3.40282347e+38f;
  const int32_t kinf = order_key(0x7f800000u), kninf = order_key(0xff800000u);
  if (n > 0 && (kmn < kninf || kmx > kinf)) {
    // Math.min/max: a NaN anywhere makes both NaN
    minmax[0] = minmax[1] = std::nanf("");
    return hipSuccess;
  }
  uint32_t b;
  float mn = fmax, mx = -fmax;  // start values +/-Float.MAX_VALUE
  if (n > 0) {
    b = key_bits(kmn);
    float v;
    memcpy(&v, &b, 4);
    if (!(v > fmax)) mn = v;  // +inf everywhere leaves Float.MAX_VALUE
    b = key_bits(kmx);
    memcpy(&v, &b, 4);
    if (!(v < -fmax)) mx = v;
  }
  minmax[0] = mn;
  minmax[1] = mx;
  return hipSuccess;
}

void u16_range(const float* minmax, double* out) {
  double mn = minmax[0], mx = minmax[1];
  // the widening of max reads the widened min (ExportSpimData2HDF5.java:235-236)
  const double t = mn - ((mn + mx) / 2.0) * 0.1;
  mn = (t != t) ? t : (t > 0.0 ? t : 0.0);  // Math.max(0, t)
  mx = mx + ((mn + mx) / 2.0) * 0.1;
  out[0] = mn;
  out[1] = mx;
}

hipError_t launch_to_u16(const float* img, int64_t n, double min, double max, bool big_endian, uint16_t* out,
                         hipStream_t stream) {
  const double scale = max - min;
  (void)hipGetLastError();
  if (big_endian)
    hipLaunchKernelGGL(k_to_u16<true>, dim3(grid_for((n + 3) / 4)), dim3(kExThreads), 0, stream, img, n, min, scale,
                       out);
  else
    hipLaunchKernelGGL(k_to_u16<false>, dim3(grid_for((n + 3) / 4)), dim3(kExThreads), 0, stream, img, n, min,
                       scale, out);
  return hipGetLastError();
}

// The whole export: range, conversion, pipeline to the file.
int write_tiff_stack(const float* img, bool on_device, int nx, int ny, int nz, int bits, double min, double max,
                     const double* calib, bool big_endian, int force_bigtiff, const char* path, hipStream_t stream) {
  if (!img || !path) return fail("NULL argument", -1);
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail("invalid image size", -1);
  if (bits != 16 && bits != 32) return fail("bits must be 16 or 32", -1);
  if (calib && !(std::isfinite(calib[0]) && std::isfinite(calib[1]) && std::isfinite(calib[2]) && calib[3] > 0 &&
                 std::isfinite(calib[3])))
    return fail("invalid calibration", -1);
  const int64_t n = (int64_t)nx * ny * nz;
  const int esz = bits / 8;
  // range: NaN -> computed (Save3dTIFF.java:86-89; widened for 16 bit,
  // ExportSpimData2HDF5.java:229-237)
  double lo = min, hi = max;
  if (std::isnan(min) || std::isnan(max)) {
    float mm[2];
    hipError_t e = hipSuccess;
    if (on_device) {
      e = launch_min_max(img, n, mm, stream);
    } else {
      float* d = nullptr;
      e = hipMalloc(&d, sizeof(float) * n);
      if (e == hipSuccess) e = hipMemcpy(d, img, sizeof(float) * n, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = launch_min_max(d, n, mm, stream);
      if (d) (void)hipFree(d);
    }
    if (e != hipSuccess) return fail(std::string("min/max failed: ") + hipGetErrorString(e), -2);
    if (bits == 16) {
      double r[2];
      u16_range(mm, r);
      lo = r[0];
      hi = r[1];
    } else {
      lo = mm[0];
      hi = mm[1];
    }
  }
  const uint64_t data_bytes = (uint64_t)n * esz;
  uint64_t data_off = 0;
  const std::string desc =
      imagej_description(nz, bits == 16 ? 0.0 : lo, bits == 16 ? 65535.0 : hi, calib);
  bool bigtiff = force_bigtiff != 0;
  std::vector<uint8_t> prefix = tiff_prefix(nx, ny, nz, bits, big_endian, bigtiff, desc, calib, &data_off);
  if (prefix.empty()) return -2;
  if (!bigtiff && data_off + data_bytes > 0xffffffffull) {
    bigtiff = true;  // classic TIFF offsets are 32 bits
    prefix = tiff_prefix(nx, ny, nz, bits, big_endian, true, desc, calib, &data_off);
    if (prefix.empty()) return -2;
  }
  const int fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) return fail(std::string("cannot open ") + path, -3);
  int rc = 0;
  if (!write_all(fd, prefix.data(), prefix.size())) rc = fail("write failed", -3);
  // chunks of whole groups of 4 values, two stages in flight
  const int64_t chunk = std::max<int64_t>(4, (kChunkBytes / esz) & ~int64_t(3));
  const bool need_kernel = bits == 16 || big_endian;
  if (rc == 0 && !on_device && !need_kernel) {
    // little-endian float32 from host memory: the bytes as they are
    if (!write_all(fd, img, (size_t)data_bytes)) rc = fail("write failed", -3);
    if (::close(fd) != 0 && rc == 0) rc = fail("close failed", -3);
    if (rc != 0) (void)::unlink(path);
    return rc;
  }
  void* d_in = nullptr;
  void* d_out[2] = {nullptr, nullptr};
  void* h_out[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  const int64_t cap = std::min<int64_t>(chunk, n);
  hipError_t e = hipSuccess;
  if (rc == 0) {
    for (int b = 0; b < 2 && e == hipSuccess; ++b) {
      if (need_kernel) e = hipMalloc(&d_out[b], (size_t)cap * esz);
      if (e == hipSuccess) e = hipHostMalloc(&h_out[b], (size_t)cap * esz, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[b], hipEventDisableTiming);
    }
    if (e == hipSuccess && !on_device) e = hipMalloc(&d_in, (size_t)cap * sizeof(float) * 2);
    if (e != hipSuccess) rc = fail(std::string("export buffers: ") + hipGetErrorString(e), -2);
  }
  int64_t pending = -1, pending_len = 0;
  for (int64_t c0 = 0, k = 0; rc == 0 && c0 < n; c0 += chunk, ++k) {
    const int b = (int)(k & 1);
    const int64_t len = std::min<int64_t>(chunk, n - c0);
    const float* src = img + c0;
    if (!on_device) {
      float* dst = static_cast<float*>(d_in) + (size_t)b * cap;
      e = hipMemcpyAsync(dst, src, sizeof(float) * len, hipMemcpyHostToDevice, stream);
      src = dst;
    }
    if (e == hipSuccess && bits == 16)
      e = launch_to_u16(src, len, lo, hi, big_endian, static_cast<uint16_t*>(d_out[b]), stream);
    if (e == hipSuccess && bits == 32 && big_endian) {
      (void)hipGetLastError();
      hipLaunchKernelGGL(k_swap32, dim3(grid_for((len + 3) / 4)), dim3(kExThreads), 0, stream,
                         reinterpret_cast<const uint32_t*>(src), len, static_cast<uint32_t*>(d_out[b]));
      e = hipGetLastError();
    }
    const void* from = need_kernel ? d_out[b] : static_cast<const void*>(src);
    if (e == hipSuccess) e = hipMemcpyAsync(h_out[b], from, (size_t)len * esz, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipEventRecord(ev[b], stream);
    // write the previous chunk while this one is converted and copied
    if (e == hipSuccess && pending >= 0) {
      const int pb = (int)(pending & 1);
      e = hipEventSynchronize(ev[pb]);
      if (e == hipSuccess && !write_all(fd, h_out[pb], (size_t)pending_len * esz)) rc = fail("write failed", -3);
    }
    if (e != hipSuccess && rc == 0) rc = fail(std::string("export: ") + hipGetErrorString(e), -2);
    pending = k;
    pending_len = len;
  }
  if (rc == 0 && pending >= 0) {
    const int pb = (int)(pending & 1);
    e = hipEventSynchronize(ev[pb]);
    if (e != hipSuccess) rc = fail(std::string("export: ") + hipGetErrorString(e), -2);
    else if (!write_all(fd, h_out[pb], (size_t)pending_len * esz)) rc = fail("write failed", -3);
  }
  (void)hipStreamSynchronize(stream);
  for (int b = 0; b < 2; ++b) {
    if (d_out[b]) (void)hipFree(d_out[b]);
    if (h_out[b]) (void)hipHostFree(h_out[b]);
    if (ev[b]) (void)hipEventDestroy(ev[b]);
  }
  if (d_in) (void)hipFree(d_in);
  if (::close(fd) != 0 && rc == 0) rc = fail("close failed", -3);
  if (rc != 0) (void)::unlink(path);  // no partial file is left behind
  return rc;
}

}  // namespace mvd

using namespace mvd;

extern "C" {

int mvd_min_max_device(int device, const float* img, long long n, float* minmax, void* stream) {
  if (!img || !minmax || n < 0) return fail("invalid argument", -1);
  DeviceScope ds(device);
  if (!ds.ok) return fail("invalid device", -1);
  hipError_t e = launch_min_max(img, n, minmax, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(std::string("min/max failed: ") + hipGetErrorString(e), -2);
  return 0;
}

int mvd_min_max(int device, const float* img, long long n, float* minmax) {
  if (!img || !minmax || n < 0) return fail("invalid argument", -1);
  DeviceScope ds(device);
  if (!ds.ok) return fail("invalid device", -1);
  float* d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(float) * std::max<long long>(n, 1));
  if (e == hipSuccess) e = hipMemcpy(d, img, sizeof(float) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_min_max(d, n, minmax, nullptr);
  if (d) (void)hipFree(d);
  if (e != hipSuccess) return fail(std::string("min/max failed: ") + hipGetErrorString(e), -2);
  return 0;
}

int mvd_u16_range(const float* minmax, double* out) {
  if (!minmax || !out) return fail("NULL argument", -1);
  u16_range(minmax, out);
  return 0;
}

int mvd_convert_u16_device(int device, const float* img, long long n, double min, double max, int big_endian,
                           unsigned short* out, void* stream) {
  if (!img || !out || n < 0) return fail("invalid argument", -1);
  DeviceScope ds(device);
  if (!ds.ok) return fail("invalid device", -1);
  if (n == 0) return 0;
  hipError_t e = launch_to_u16(img, n, min, max, big_endian != 0, out, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(std::string("conversion failed: ") + hipGetErrorString(e), -2);
  return 0;
}

int mvd_write_tiff(int device, const float* img, int img_on_device, int nx, int ny, int nz, int bits, double min,
                   double max, const double* calibration, int big_endian, int bigtiff, const char* path) {
  DeviceScope ds(device);
  if (!ds.ok) return fail("invalid device", -1);
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return fail("stream creation failed", -2);
  const int rc = write_tiff_stack(img, img_on_device != 0, nx, ny, nz, bits, min, max, calibration,
                                  big_endian != 0, bigtiff, path, st);
  (void)hipStreamDestroy(st);
  return rc;
}

}  // extern "C"
