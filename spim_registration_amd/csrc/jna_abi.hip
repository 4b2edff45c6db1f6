// The JNA ABI of the reference's native FFT-convolution library
// (spim/process/cuda/CUDAStandardFunctions.java:26-45,
//  spim/process/cuda/CUDAFourierConvolution.java:25-32), implemented on HIP.
//
// convolution3DfftCUDAInPlace is called by one Java thread per selected device
// (deconv/MVDeconFFT.java:447-468), so every device has its own context: a
// mutex, a stream, device buffers and a cache of kernel spectra keyed by the
// kernel's bytes and the FFT size (the reference passes the same kernel for
// every block of a view).  Nothing here may abort the JVM: errors are printed
// to stderr and the block is left unchanged.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mvdecon.h"
#include "mvd_internal.h"

namespace mvd {
namespace {

uint64_t fnv1a(const void* data, size_t bytes, uint64_t h = 1469598103934665603ull) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < bytes; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

struct KernelSpectrum {
  uint64_t hash;
  Int3 kd, q;
  float2* spec = nullptr;
  size_t elems = 0;
};

struct DevicePlan {
  Int3 q;
  ShapePlan sp;
};

struct DeviceContext {
  std::mutex mu;
  int device = -1;
  hipStream_t stream = nullptr;
  float* block = nullptr;
  size_t block_cap = 0;
  float* kern = nullptr;
  size_t kern_cap = 0;
  float2* spec = nullptr;
  size_t spec_cap = 0;
  std::list<DevicePlan> plans;
  std::list<KernelSpectrum> kcache;  // most recent first
};

std::mutex g_ctx_mu;
std::vector<std::unique_ptr<DeviceContext>> g_ctx;

DeviceContext* context(int dev) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if ((int)g_ctx.size() <= dev) g_ctx.resize(dev + 1);
  if (!g_ctx[dev]) {
    g_ctx[dev] = std::make_unique<DeviceContext>();
    g_ctx[dev]->device = dev;
  }
  return g_ctx[dev].get();
}

// exp(-2 pi i k / n), k < n, computed in long double and rounded to float.
void fill_twiddles(float2* t, int n) {
  const long double two_pi = 6.283185307179586476925286766559L;
  for (int k = 0; k < n; ++k) {
    const long double a = -two_pi * (long double)k / (long double)n;
    t[k] = make_float2((float)cosl(a), (float)sinl(a));
  }
}

std::vector<float2> twiddle_table(int n) {
  std::vector<float2> t(n);
  fill_twiddles(t.data(), n);
  return t;
}

int lines(int n) {
  int b = 16;
  while (b > 1 && n * b > kLdsCap) b >>= 1;
  return b;
}

const ShapePlan* plan_for(DeviceContext* c, Int3 q) {
  for (auto& p : c->plans)
    if (p.q.x == q.x && p.q.y == q.y && p.q.z == q.z) return &p.sp;
  if (q.x % 2 || !fft_size_supported(q.x / 2) || !fft_size_supported(q.y) || !fft_size_supported(q.z)) return nullptr;
  DevicePlan dp;
  dp.q = q;
  ShapePlan& sp = dp.sp;
  sp.q = q;
  sp.hx = q.x / 2 + 1;
  sp.pitch = spectrum_pitch(sp.hx);
  sp.px = make_plan(q.x / 2);
  sp.py = make_plan(q.y);
  sp.pz = make_plan(q.z);
  sp.bx = lines(q.x / 2 + 1);
  sp.by = lines(q.y);
  sp.bz = lines(q.z);
  float2** dst[3] = {&sp.tw_x, &sp.tw_y, &sp.tw_z};
  const int len[3] = {q.x, q.y, q.z};
  for (int i = 0; i < 3; ++i) {
    std::vector<float2> t = twiddle_table(len[i]);
    if (hipMalloc(dst[i], sizeof(float2) * len[i]) != hipSuccess) return nullptr;
    if (hipMemcpy(*dst[i], t.data(), sizeof(float2) * len[i], hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  }
  c->plans.push_back(dp);
  return &c->plans.back().sp;
}

bool grow(void** p, size_t* cap, size_t bytes) {
  if (bytes <= *cap) return true;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, bytes) != hipSuccess) return false;
  *cap = bytes;
  return true;
}

#define JNA_FAIL(msg)                                                    \
  do {                                                                   \
    std::fprintf(stderr, "libmvdecon_hip: %s\n", std::string(msg).c_str()); \
    set_error(msg);                                                      \
    return false;                                                        \
  } while (0)

#define JNA_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) JNA_FAIL(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Convolve the block `im` (dims x-fastest) with `kernel`; result into `out`.
bool convolve_block(const float* im, Int3 bd, const float* kernel, Int3 kd, int dev, float* out) {
  if (!im || !kernel || !out) JNA_FAIL("NULL array");
  if (bd.x <= 0 || bd.y <= 0 || bd.z <= 0 || kd.x <= 0 || kd.y <= 0 || kd.z <= 0) JNA_FAIL("invalid dimensions");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || dev < 0 || dev >= ndev) {
    (void)hipGetLastError();
    JNA_FAIL("invalid device id");
  }
  DeviceContext* c = context(dev);
  std::lock_guard<std::mutex> lk(c->mu);
  int prev = -1;
  (void)hipGetDevice(&prev);
  struct Restore {
    int d;
    ~Restore() {
      if (d >= 0) (void)hipSetDevice(d);
    }
  } restore{prev};
  JNA_HIP(hipSetDevice(dev));
  if (!c->stream) JNA_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const Int3 q = {next_fft_size(bd.x, true), next_fft_size(bd.y, false), next_fft_size(bd.z, false)};
  if (kd.x > q.x || kd.y > q.y || kd.z > q.z) JNA_FAIL("kernel larger than the block");
  const ShapePlan* sp = plan_for(c, q);
  if (!sp) JNA_FAIL("unsupported block size");
  const size_t nb = (size_t)bd.x * bd.y * bd.z, nk = (size_t)kd.x * kd.y * kd.z;
  if (!grow((void**)&c->block, &c->block_cap, sizeof(float) * nb)) JNA_FAIL("out of device memory (block)");
  if (!grow((void**)&c->spec, &c->spec_cap, sizeof(float2) * sp->spectrum_elems()))
    JNA_FAIL("out of device memory (spectrum)");
  // kernel spectrum, cached by content
  const uint64_t h = fnv1a(kernel, sizeof(float) * nk, fnv1a(&kd, sizeof(kd)));
  KernelSpectrum* ks = nullptr;
  for (auto it = c->kcache.begin(); it != c->kcache.end(); ++it)
    if (it->hash == h && it->q.x == q.x && it->q.y == q.y && it->q.z == q.z && it->kd.x == kd.x &&
        it->kd.y == kd.y && it->kd.z == kd.z) {
      c->kcache.splice(c->kcache.begin(), c->kcache, it);
      ks = &c->kcache.front();
      break;
    }
  if (!ks) {
    while (c->kcache.size() >= 16) {
      (void)hipFree(c->kcache.back().spec);
      c->kcache.pop_back();
    }
    KernelSpectrum k;
    k.hash = h;
    k.kd = kd;
    k.q = q;
    k.elems = sp->spectrum_elems();
    JNA_HIP(hipMalloc(&k.spec, sizeof(float2) * k.elems));
    struct Guard {  // frees the spectrum unless it made it into the cache
      float2** p;
      ~Guard() {
        if (*p) (void)hipFree(*p);
      }
    } guard{&k.spec};
    if (!grow((void**)&c->kern, &c->kern_cap, sizeof(float) * nk)) JNA_FAIL("out of device memory (kernel)");
    JNA_HIP(hipMemcpyAsync(c->kern, kernel, sizeof(float) * nk, hipMemcpyHostToDevice, c->stream));
    ConvGeom kg{};
    kg.src = kd;
    kg.win = q;
    kg.q = q;
    JNA_HIP(launch_x_forward(*sp, kg, c->kern, kGatherKernelWrap, k.spec, c->stream));
    JNA_HIP(launch_y(*sp, k.spec, 0, q.z, 0, q.y, -1.f, c->stream));
    const float scale = (float)(1.0L / ((long double)q.x * q.y * q.z));
    JNA_HIP(launch_z(*sp, k.spec, nullptr, scale, 0, q.z, c->stream));
    JNA_HIP(hipStreamSynchronize(c->stream));
    c->kcache.push_front(k);
    k.spec = nullptr;
    ks = &c->kcache.front();
  }
  JNA_HIP(staged_h2d(c->block, im, sizeof(float) * nb, c->stream));  // Java array: pageable
  ConvGeom g{};
  g.src = bd;
  g.off = {0, 0, 0};
  g.win = bd;
  g.q = q;
  g.eff_local = {0, 0, 0};
  g.eff_size = bd;
  g.eff_offset = {0, 0, 0};
  g.dst = bd;
  EpilogueArgs e{};
  e.mode = kEpiStore;
  e.dst = c->block;
  JNA_HIP(launch_xy_forward_passes(*sp, g, c->block, kGatherZero, c->spec, c->stream));
  JNA_HIP(launch_z(*sp, c->spec, ks->spec, 1.f, 0, bd.z, c->stream));
  JNA_HIP(launch_yx_inverse_passes(*sp, g, c->spec, e, c->stream, nullptr));
  // in place (out == im): im is overwritten only once the whole result has
  // reached host memory, so a failed copy leaves the caller's block as it was
  if (out == im)
    JNA_HIP(staged_d2h_all_or_nothing(out, c->block, sizeof(float) * nb, c->stream));
  else
    JNA_HIP(staged_d2h(out, c->block, sizeof(float) * nb, c->stream));
  JNA_HIP(hipStreamSynchronize(c->stream));
  return true;
}

bool dims_ok(const int* d) { return d && d[0] > 0 && d[1] > 0 && d[2] > 0; }

// ---------------------------------------------------------------------------
// The no-allocation variant: FFT lengths up to kNaMaxLen per axis; twiddles,
// block spectrum and kernel spectrum live in the file-scope array `scratch`.
constexpr int kNaMaxLen = 128;
constexpr int64_t kNaTwElems = 4 * kNaMaxLen;  // tw_x | tw_y | tw_z | unused (keeps 4 KB alignment)
constexpr int64_t kNaSpecElems = (int64_t)((kNaMaxLen / 2 + 1 + 15) / 16 * 16) * kNaMaxLen * kNaMaxLen;
constexpr int64_t kNaElems = kNaTwElems + 2 * kNaSpecElems;

// Twiddle tables handed over by value in the kernel arguments (3 KB), so that
// filling them needs no host-to-device copy.
struct NaTwiddles {
  float2 t[3 * kNaMaxLen];
};
static_assert(sizeof(NaTwiddles) <= 3200, "kernel argument size");

__global__ void k_na_twiddles(float2* dst, NaTwiddles tw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * kNaMaxLen) dst[i] = tw.t[i];
}

}  // namespace
}  // namespace mvd

// Scratch of mvd_convolution3d_inplace_noalloc, one copy per device, held
// until the process ends (about 20 MiB).  Host code never uses this name as
// an address: it takes each device's address from hipGetSymbolAddress.
static __device__ __align__(256) float2 scratch[mvd::kNaElems];

namespace mvd {
namespace {

std::mutex g_na_mu;
float2* g_na_ptr[64] = {};

// Device address of `scratch` on the current device `dev`, looked up once.
hipError_t na_scratch(int dev, float2** out) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_na_ptr[dev] = static_cast<float2*>(p);
  }
  *out = g_na_ptr[dev];
  return hipSuccess;
}

}  // namespace
}  // namespace mvd

using namespace mvd;

extern "C" {

// Failed queries must not leave an error behind for later calls on the thread.
static void clear_error() { (void)hipGetLastError(); }

int getNumDevicesCUDA(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    clear_error();
    return -1;
  }
  return n;
}

void getNameDeviceCUDA(int devCUDA, char* name) {
  if (!name) return;
  name[0] = 0;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, devCUDA) != hipSuccess) {
    clear_error();
    return;
  }
  std::strncpy(name, p.name, 255);
  name[255] = 0;
  if (name[0] == 0) {
    // some drivers report no marketing name; fall back to the ISA name
    std::snprintf(name, 256, "AMD GPU %s", p.gcnArchName);
  }
}

long long getMemDeviceCUDA(int devCUDA) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, devCUDA) != hipSuccess) {
    clear_error();
    return -1;
  }
  return (long long)p.totalGlobalMem;
}

long long getFreeMemDeviceCUDA(int devCUDA) {
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (hipSetDevice(devCUDA) != hipSuccess) {
    clear_error();
    return -1;
  }
  size_t fr = 0, tot = 0;
  hipError_t e = hipMemGetInfo(&fr, &tot);
  if (prev >= 0) (void)hipSetDevice(prev);
  if (e != hipSuccess) clear_error();
  return e == hipSuccess ? (long long)fr : -1;
}

int getCUDAcomputeCapabilityMajorVersion(int devCUDA) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, devCUDA) != hipSuccess) {
    clear_error();
    return -1;
  }
  return p.major;
}

int getCUDAcomputeCapabilityMinorVersion(int devCUDA) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, devCUDA) != hipSuccess) {
    clear_error();
    return -1;
  }
  return p.minor;
}

void convolution3DfftCUDAInPlace(float* im, int* imDim, float* kernel, int* kernelDim, int devCUDA) {
  if (!dims_ok(imDim) || !dims_ok(kernelDim)) {
    std::fprintf(stderr, "libmvdecon_hip: invalid dimensions\n");
    return;
  }
  // dims arrive reversed: {nz, ny, nx} (MVDeconFFTThreads.getCUDACoordinates, :157-165)
  const Int3 bd = {imDim[2], imDim[1], imDim[0]}, kd = {kernelDim[2], kernelDim[1], kernelDim[0]};
  (void)convolve_block(im, bd, kernel, kd, devCUDA, im);
}

float* convolution3DfftCUDA(float* im, int* imDim, float* kernel, int* kernelDim, int devCUDA) {
  if (!dims_ok(imDim) || !dims_ok(kernelDim)) return nullptr;
  const Int3 bd = {imDim[2], imDim[1], imDim[0]}, kd = {kernelDim[2], kernelDim[1], kernelDim[0]};
  const size_t nb = (size_t)bd.x * bd.y * bd.z;
  float* out = static_cast<float*>(std::malloc(sizeof(float) * nb));
  if (!out) return nullptr;
  if (!convolve_block(im, bd, kernel, kd, devCUDA, out)) {
    std::free(out);
    return nullptr;
  }
  return out;
}

void mvd_free(void* p) { std::free(p); }

int mvd_convolution3d_inplace_noalloc(float* im, const int* imDim, const float* kernel, const int* kernelDim,
                                      int devCUDA, void* stream) {
  if (!im || !kernel) return set_error("NULL array"), -1;
  if (!dims_ok(imDim) || !dims_ok(kernelDim)) return set_error("invalid dimensions"), -1;
  const Int3 bd = {imDim[2], imDim[1], imDim[0]}, kd = {kernelDim[2], kernelDim[1], kernelDim[0]};
  const Int3 q = {next_fft_size(bd.x, true), next_fft_size(bd.y, false), next_fft_size(bd.z, false)};
  if (q.x > kNaMaxLen || q.y > kNaMaxLen || q.z > kNaMaxLen)
    return set_error("block too large for the no-allocation entry: FFT lengths above 128 per axis"), -1;
  if (kd.x > q.x || kd.y > q.y || kd.z > q.z) return set_error("kernel larger than the block"), -1;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || devCUDA < 0 || devCUDA >= ndev) {
    (void)hipGetLastError();
    return set_error("invalid device id"), -1;
  }
  int prev = -1;
  (void)hipGetDevice(&prev);
  struct Restore {
    int d;
    ~Restore() {
      if (d >= 0) (void)hipSetDevice(d);
    }
  } restore{prev};
  hipError_t e = hipSetDevice(devCUDA);
  float2* base = nullptr;
  if (e == hipSuccess) e = na_scratch(devCUDA, &base);
  if (e != hipSuccess) return set_error(std::string("no-allocation scratch: ") + hipGetErrorString(e)), -2;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  ShapePlan sp{};
  sp.q = q;
  sp.hx = q.x / 2 + 1;
  sp.pitch = spectrum_pitch(sp.hx);
  sp.px = make_plan(q.x / 2);
  sp.py = make_plan(q.y);
  sp.pz = make_plan(q.z);
  sp.bx = lines(q.x / 2 + 1);
  sp.by = lines(q.y);
  sp.bz = lines(q.z);
  sp.tw_x = base;
  sp.tw_y = base + kNaMaxLen;
  sp.tw_z = base + 2 * kNaMaxLen;
  float2* spec = base + kNaTwElems;
  float2* kspec = spec + kNaSpecElems;
  if (sp.spectrum_elems() > kNaSpecElems) return set_error("internal: no-allocation spectrum size"), -2;
  NaTwiddles tw{};
  fill_twiddles(tw.t, q.x);
  fill_twiddles(tw.t + kNaMaxLen, q.y);
  fill_twiddles(tw.t + 2 * kNaMaxLen, q.z);
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_na_twiddles, dim3((3 * kNaMaxLen + 255) / 256), dim3(256), 0, st, base, tw);
  e = hipGetLastError();
  // kernel spectrum (wrapped kernel, scaled by 1 / (qx qy qz)), then the block
  ConvGeom kg{};
  kg.src = kd;
  kg.win = q;
  kg.q = q;
  if (e == hipSuccess) e = launch_x_forward(sp, kg, kernel, kGatherKernelWrap, kspec, st);
  if (e == hipSuccess) e = launch_y(sp, kspec, 0, q.z, 0, q.y, -1.f, st);
  const float scale = (float)(1.0L / ((long double)q.x * q.y * q.z));
  if (e == hipSuccess) e = launch_z(sp, kspec, nullptr, scale, 0, q.z, st);
  ConvGeom g{};
  g.src = bd;
  g.win = bd;
  g.q = q;
  g.eff_size = bd;
  g.dst = bd;
  EpilogueArgs ep{};
  ep.mode = kEpiStore;
  ep.dst = im;
  if (e == hipSuccess) e = launch_xy_forward_passes(sp, g, im, kGatherZero, spec, st);
  if (e == hipSuccess) e = launch_z(sp, spec, kspec, 1.f, 0, bd.z, st);
  if (e == hipSuccess) e = launch_yx_inverse_passes(sp, g, spec, ep, st, nullptr);
  if (e != hipSuccess) return set_error(std::string("no-allocation convolution: ") + hipGetErrorString(e)), -2;
  return 0;
}

}  // extern "C"
