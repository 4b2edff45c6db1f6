// Deconvolution session: state resident in HBM and the host-side driver of
// the MVDeconvolution iteration (reference: deconv/MVDeconvolution.java,
// deconv/MVDeconFFT.java, deconv/MVDeconInput.java, cuda/Block*.java).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mvdecon.h"
#include "mvd_internal.h"

namespace mvd {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

#define MVD_CHECK(expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
      return -2;                                                                         \
    }                                                                                    \
  } while (0)

#define MVD_REQUIRE(cond, msg) \
  do {                         \
    if (!(cond)) {             \
      set_error(msg);          \
      return -1;               \
    }                          \
  } while (0)

// ---------------------------------------------------------------------------
// FFT plans

static bool is_5smooth(int n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

// Lengths up to 1024 keep every pass within 64 KB of LDS per workgroup.
bool fft_size_supported(int n) { return n >= 1 && n <= 1024 && is_5smooth(n); }

int next_fft_size(int n, bool even) {
  int m = std::max(n, even ? 2 : 1);
  for (;; ++m)
    if (is_5smooth(m) && (!even || m % 2 == 0)) return m;
}

FftPlan make_plan(int n) {
  FftPlan p{};
  p.n = n;
  int m = n, ns = 1;
  auto push = [&](int r) {
    p.radix[p.npass] = r;
    p.ns[p.npass] = ns;
    p.magic[p.npass] = ns > 1 ? (unsigned)((((uint64_t)1 << 32) + ns - 1) / ns) : 0u;
    ns *= r;
    ++p.npass;
    m /= r;
  };
  while (m % 8 == 0) push(8);
  while (m % 4 == 0) push(4);
  while (m % 2 == 0) push(2);
  while (m % 3 == 0) push(3);
  while (m % 5 == 0) push(5);
  return p;
}

static int lines_for(int n, int want) {
  int b = want;
  while (b > 1 && n * b > kLdsCap) b >>= 1;
  return b;
}

static std::vector<float2> twiddles(int n) {
  std::vector<float2> t(n);
  const long double two_pi = 6.283185307179586476925286766559L;
  for (int k = 0; k < n; ++k) {
    const long double a = -two_pi * (long double)k / (long double)n;
    t[k] = make_float2((float)cosl(a), (float)sinl(a));
  }
  return t;
}

struct DeviceBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DeviceBuf() = default;
  DeviceBuf(const DeviceBuf&) = delete;
  DeviceBuf& operator=(const DeviceBuf&) = delete;
  DeviceBuf(DeviceBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DeviceBuf& operator=(DeviceBuf&& o) noexcept {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
    return *this;
  }
  ~DeviceBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

class PlanCache {
 public:
  ~PlanCache() {
    for (auto& kv : plans_) {
      (void)hipFree(kv.second->tw_x);
      (void)hipFree(kv.second->tw_y);
      (void)hipFree(kv.second->tw_z);
    }
  }
  // Returns nullptr (and sets the error) for unsupported sizes.
  const ShapePlan* get(Int3 q) {
    auto key = std::make_tuple(q.x, q.y, q.z);
    auto it = plans_.find(key);
    if (it != plans_.end()) return it->second.get();
    if (q.x % 2 || !fft_size_supported(q.x / 2) || !fft_size_supported(q.y) || !fft_size_supported(q.z)) {
      set_error("unsupported FFT size " + std::to_string(q.x) + "x" + std::to_string(q.y) + "x" +
                std::to_string(q.z));
      return nullptr;
    }
    auto sp = std::make_unique<ShapePlan>();
    sp->q = q;
    sp->hx = q.x / 2 + 1;
    sp->pitch = spectrum_pitch(sp->hx);
    sp->px = make_plan(q.x / 2);
    sp->py = make_plan(q.y);
    sp->pz = make_plan(q.z);
    sp->bx = lines_for(q.x / 2 + 1, 16);
    sp->by = lines_for(q.y, 16);
    sp->bz = lines_for(q.z, 16);
    auto up = [](int n, float2** dst) -> bool {
      std::vector<float2> t = twiddles(n);
      if (hipMalloc(dst, sizeof(float2) * n) != hipSuccess) return false;
      return hipMemcpy(*dst, t.data(), sizeof(float2) * n, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(q.x, &sp->tw_x) || !up(q.y, &sp->tw_y) || !up(q.z, &sp->tw_z)) {
      set_error("twiddle table allocation failed");
      return nullptr;
    }
    const ShapePlan* r = sp.get();
    plans_[key] = std::move(sp);
    return r;
  }

 private:
  std::map<std::tuple<int, int, int>, std::unique_ptr<ShapePlan>> plans_;
};

// ---------------------------------------------------------------------------
// Timing of kernel classes with HIP events on the session stream.

enum TimingClass {
  kTXForward = 0,
  kTYForward,
  kTZConv,
  kTYInverse,
  kTXInvQuotient,
  kTXInvUpdate,
  kTXInvStore,
  kTZForward,
  kTReduce,
  kTPointwise,
  kTXYForward,
  kTXYInvQuotient,
  kTXYInvUpdate,
  kTXYInvStore,
  kTTransform,
  kTNumClasses
};
static const char* kTimingNames[kTNumClasses] = {"x_forward",         "y_forward",       "z_conv",
                                                 "y_inverse",         "x_inverse_quotient", "x_inverse_update",
                                                 "x_inverse_store",   "z_forward",       "reduce_stats",
                                                 "pointwise",         "xy_forward",      "xy_inverse_quotient",
                                                 "xy_inverse_update", "xy_inverse_store", "transform_view"};

class Timer {
 public:
  ~Timer() { release(); }
  void enable(bool on) {
    on_ = on;
    std::fill(ms_, ms_ + kTNumClasses, 0.0);
    std::fill(cnt_, cnt_ + kTNumClasses, 0);
  }
  bool on() const { return on_; }
  void begin(hipStream_t s) {
    if (!on_) return;
    if (used_ == pool_.size()) {
      hipEvent_t a, b;
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      pool_.push_back({a, b, 0});
    }
    (void)hipEventRecord(pool_[used_].start, s);
  }
  void end(hipStream_t s, int cls) {
    if (!on_) return;
    (void)hipEventRecord(pool_[used_].stop, s);
    pool_[used_].cls = cls;
    ++used_;
  }
  // Call after the stream is synchronised.
  void collect() {
    for (size_t i = 0; i < used_; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, pool_[i].start, pool_[i].stop) == hipSuccess) {
        ms_[pool_[i].cls] += ms;
        cnt_[pool_[i].cls] += 1;
      }
    }
    used_ = 0;
  }
  double ms(int c) const { return ms_[c]; }
  long long count(int c) const { return cnt_[c]; }

 private:
  struct Pair {
    hipEvent_t start, stop;
    int cls;
  };
  void release() {
    for (auto& p : pool_) {
      (void)hipEventDestroy(p.start);
      (void)hipEventDestroy(p.stop);
    }
    pool_.clear();
  }
  bool on_ = false;
  std::vector<Pair> pool_;
  size_t used_ = 0;
  double ms_[kTNumClasses] = {};
  long long cnt_[kTNumClasses] = {};
};

// ---------------------------------------------------------------------------
// Block geometry: BlockGeneratorFixedSizePrecise.divideIntoBlocks
// (cuda/BlockGeneratorFixedSizePrecise.java:46-122).  Effective size bs-k+1,
// effective local offset k/2, blocks x-fastest, last block clipped.

static bool divide_into_blocks(Int3 img, Int3 ker, Int3 bs, Int3 q, std::vector<ConvGeom>* out) {
  const int is[3] = {img.x, img.y, img.z}, ks[3] = {ker.x, ker.y, ker.z}, b[3] = {bs.x, bs.y, bs.z};
  int eff[3], nb[3];
  for (int d = 0; d < 3; ++d) {
    eff[d] = b[d] - ks[d] + 1;
    if (eff[d] <= 0) return false;
    nb[d] = is[d] / eff[d] + (is[d] % eff[d] != 0);
  }
  out->clear();
  for (int k = 0; k < nb[2]; ++k)
    for (int j = 0; j < nb[1]; ++j)
      for (int i = 0; i < nb[0]; ++i) {
        const int cur[3] = {i, j, k};
        int off[3], esz[3], eoff[3], eloc[3];
        for (int d = 0; d < 3; ++d) {
          esz[d] = eff[d];
          eoff[d] = cur[d] * eff[d];
          off[d] = eoff[d] - ks[d] / 2;
          eloc[d] = ks[d] / 2;
          if (eoff[d] + esz[d] > is[d]) esz[d] = is[d] - eoff[d];
        }
        ConvGeom g;
        g.src = img;
        g.off = {off[0], off[1], off[2]};
        g.win = bs;
        g.q = q;
        g.eff_local = {eloc[0], eloc[1], eloc[2]};
        g.eff_size = {esz[0], esz[1], esz[2]};
        g.eff_offset = {eoff[0], eoff[1], eoff[2]};
        g.dst = img;
        out->push_back(g);
      }
  return true;
}

static Int3 fft_shape_for(Int3 win) {
  return {next_fft_size(win.x, true), next_fft_size(win.y, false), next_fft_size(win.z, false)};
}

// ---------------------------------------------------------------------------
// Host-side kernel preparation helpers (tiny arrays).

// FusionHelper.divideIntoPortions (spim/process/fusion/FusionHelper.java:257-280).
static std::vector<std::pair<int64_t, int64_t>> portions(int64_t n, int k) {
  std::vector<std::pair<int64_t, int64_t>> p;
  const int64_t chunk = n / k, mod = n % k;
  for (int i = 0; i < k; ++i) p.push_back({i * chunk, i == k - 1 ? chunk + mod : chunk});
  return p;
}

// Pairwise summation of doubles in input order (binary counter).
class PairSum {
 public:
  void add(double a) {
    size_t i = 0;
    double s = a;
    while (i < full_.size() && full_[i]) {
      full_[i] = false;
      s += v_[i];
      v_[i] = 0.0;
      ++i;
    }
    if (i == full_.size()) {
      full_.push_back(false);
      v_.push_back(0.0);
    }
    full_[i] = true;
    v_[i] = s;
  }
  double sum() const {
    double s = 0;
    for (double x : v_) s += x;
    return s;
  }

 private:
  std::vector<bool> full_;
  std::vector<double> v_;
};

// AdjustInput.normImg (deconv/AdjustInput.java:50-100): k /= S where S sums
// 2*threads portions and counts the first portion twice.
static void norm_img(std::vector<float>& k, int threads) {
  auto ps = portions((int64_t)k.size(), 2 * threads);
  std::vector<double> part;
  for (auto& p : ps) {
    PairSum s;
    for (int64_t j = 0; j < p.second; ++j) s.add((double)k[p.first + j]);
    part.push_back(s.sum());
  }
  PairSum total;
  total.add(part[0]);
  for (double v : part) total.add(v);
  const double sum = total.sum();
  for (auto& v : k) v = (float)((double)v / sum);
}

// Mirror.mirror along each axis (deconv/Mirror.java:52-129; MVDeconFFT.computeInvertedKernel,
// deconv/MVDeconFFT.java:336-344).  Positions p <= n/2 are swapped with n-1-p, so
// for even n the middle pair is swapped twice and stays.
static std::vector<float> inverted(const std::vector<float>& k, Int3 d) {
  std::vector<float> r = k;
  const int dims[3] = {d.x, d.y, d.z};
  for (int axis = 0; axis < 3; ++axis) {
    const int len = dims[axis], half = len / 2, mx = len - 1;
    for (int z = 0; z < d.z; ++z)
      for (int y = 0; y < d.y; ++y)
        for (int x = 0; x < d.x; ++x) {
          int p[3] = {x, y, z};
          if (p[axis] > half) continue;
          p[axis] = mx - p[axis];
          std::swap(r[x + (size_t)d.x * (y + (size_t)d.y * z)], r[p[0] + (size_t)d.x * (p[1] + (size_t)d.y * p[2])]);
        }
  }
  return r;
}

}  // namespace mvd

namespace mvd {

// One launch of a pass over a group of items (blocks of one or many stacks)
// that share the kernel class; items, prefix and per-item grids live on the device.
struct PassLaunch {
  int kind = 0, cls = 0, epi = 0, nitems = 0, nwg = 0;
  int path = 0;    // 0 generic kernels, 1 register-resident kernels, 2 fused x-y kernels
  int timing = 0;  // TimingClass of the launch
  bool lines = false;  // z pass with one kernel spectrum for all items: grid (lx, ly) of lines
  int lx = 0, ly = 0;
  bool reduce = false;
  size_t lds = 0;
  DeviceBuf items, prefix, gx;
  DeviceBuf sched;  // fused forward: one PlaneTask per workgroup
  bool has_sched = false;
  int table = 0;  // lines z pass reading its input planes through tables (1: 32-bit, 2: 64-bit offsets)
  DeviceBuf ptasks;  // separate x and y forward: (item, plane) per wpt workgroups
  int wpt = 0;
};

// MVD_XCD_SCHEDULE=0 keeps the item-major order of the forward plane
// workgroups and one transform per plane; MVD_PLANE_DEDUP=0 keeps the order
// but transforms every plane of every window.
static const bool g_xcd_schedule = [] {
  const char* v = std::getenv("MVD_XCD_SCHEDULE");
  return !(v && v[0] == '0');
}();
static const bool g_plane_dedup = [] {
  const char* v = std::getenv("MVD_PLANE_DEDUP");
  return !(v && v[0] == '0');
}();

// MVD_GRAD_FUSED=0: the gradient's spectrum products and z transforms as
// separate launches; MVD_GRAD_KEEP_W=0: the window spectrum recomputed in the
// second pass also for one block (comparisons).
static const bool g_grad_fused = [] {
  const char* v = std::getenv("MVD_GRAD_FUSED");
  return !(v && v[0] == '0');
}();
static const bool g_grad_keep_w = [] {
  const char* v = std::getenv("MVD_GRAD_KEEP_W");
  return !(v && v[0] == '0');
}();

static int mirror_single_host(int p, int n) {
  if (p >= 0 && p < n) return p;
  if (n == 1) return 0;
  const int period = 2 * n - 2;
  p %= period;
  if (p < 0) p += period;
  return p < n ? p : period - p;
}

// Task list of one fused forward launch (one plane of one block window per
// task).
//
// Identical planes.  A window plane's content is fixed by the gather rule,
// the source, the window's x-y placement and the source plane it reads: under
// the mirror rule the source plane is the reflection of the window's z; under
// the constant rules a plane outside the source in z is the constant within
// the window (whatever its x-y placement), and planes past the window's depth
// are zero.  Overlapping block windows, and windows that run past an image
// face, therefore hold many identical planes -- at 512 x 512 x 256 with 256^3
// blocks and a 51-deep kernel half of each forward launch's 4608 planes.  Each
// distinct plane is transformed once and stored into every spectrum that holds
// it (up to kMaxPlaneDests per task, more copies take more tasks); the
// transform is deterministic, so every copy is bit for bit what its own
// transform would give.
//
// Order.  Workgroup w runs on XCD w mod 8, and each XCD has its own L2.  The
// planes that read source plane Z (neighbouring windows overlap, reflected
// windows read planes twice) are sorted by (Z, window y, window x) and the
// sorted list is cut into 8 contiguous parts, one per XCD, so the workgroups
// that share a source plane run on one XCD at about the same time.  Planes
// that read nothing are dealt out evenly at the end of the parts.
namespace {

enum { kPlaneRead = 0, kPlaneOutside = 1, kPlanePadding = 2 };
// key of a window plane's content: (kind, gather, q, win.x, win.y, src, off.x, off.y, source z)
using PlaneKey = std::tuple<int, int, int, int, int, int, int, uintptr_t, int, int, int>;

// Content key of plane p of a forward item; *z = source plane read (-1: none).
PlaneKey plane_key(const PassDesc& d, int p, int* z) {
  const ConvGeom& g = d.g;
  int kind = kPlaneRead;
  *z = -1;
  if (p >= g.win.z) {
    kind = kPlanePadding;
  } else {
    const int gz = g.off.z + p;
    if (d.gather == kGatherMirror) *z = mirror_single_host(gz, g.src.z);
    else if (gz >= 0 && gz < g.src.z) *z = gz;
    else kind = kPlaneOutside;
  }
  if (kind == kPlaneRead)
    return PlaneKey{kind, d.gather, g.q.x, g.q.y, g.q.z, g.win.x, g.win.y, reinterpret_cast<uintptr_t>(d.src),
                    g.off.x, g.off.y, *z};
  return PlaneKey{kind, kind == kPlanePadding ? -1 : d.gather, g.q.x, g.q.y, g.q.z, g.win.x, g.win.y, 0, 0, 0, 0};
}

// Distinct planes of a forward launch: per content key, the (item, plane)
// occurrences in item and plane order, and the source plane and window
// placement used to order the tasks.
struct PlaneGroup {
  int z, y, x;
  std::vector<std::pair<int, int>> occ;
};
std::vector<PlaneGroup> plane_groups(const std::vector<const PassDesc*>& items, bool dedup) {
  std::vector<PlaneGroup> groups;
  std::map<PlaneKey, int> index;
  for (int i = 0; i < (int)items.size(); ++i) {
    const PassDesc& d = *items[i];
    for (int p = 0; p < d.g.q.z; ++p) {
      int z;
      const PlaneKey k = plane_key(d, p, &z);
      auto it = dedup ? index.find(k) : index.end();
      if (it == index.end()) {
        if (dedup) index.emplace(k, (int)groups.size());
        groups.push_back({z, d.g.off.y, d.g.off.x, {{i, p}}});
      } else {
        groups[it->second].occ.push_back({i, p});
      }
    }
  }
  return groups;
}

struct OrderedTask {
  int z, y, x;
  PlaneTask t;
};
// Tasks storing the plane of group g into occurrences [a, a + n) of dest.
void push_tasks(const PlaneGroup& g, const std::vector<std::pair<int, int>>& dest, std::vector<OrderedTask>* out) {
  for (size_t a = 0; a < dest.size(); a += kMaxPlaneDests) {
    PlaneTask t{};
    t.item = dest[a].first;
    t.plane = dest[a].second;
    t.ndest = (int)std::min<size_t>(kMaxPlaneDests, dest.size() - a);
    for (int d = 0; d < kMaxPlaneDests; ++d) {
      const auto& e = dest[a + std::min(d, t.ndest - 1)];  // unused entries repeat a valid one
      t.dest[d][0] = e.first;
      t.dest[d][1] = e.second;
    }
    out->push_back({g.z, g.y, g.x, t});
  }
}

// XCD-aware order (see forward_plane_tasks).
std::vector<PlaneTask> order_tasks(std::vector<OrderedTask> tasks) {
  constexpr int kXcds = 8;
  std::vector<OrderedTask> reading, idle;
  for (auto& t : tasks) (t.z >= 0 ? reading : idle).push_back(t);
  std::stable_sort(reading.begin(), reading.end(), [](const OrderedTask& a, const OrderedTask& b) {
    return a.z != b.z ? a.z < b.z : a.y != b.y ? a.y < b.y : a.x < b.x;
  });
  std::vector<std::vector<OrderedTask>> part(kXcds);
  const size_t nr = reading.size();
  for (int x = 0; x < kXcds; ++x)
    part[x].assign(reading.begin() + nr * x / kXcds, reading.begin() + nr * (x + 1) / kXcds);
  for (size_t k = 0; k < idle.size(); ++k) part[k % kXcds].push_back(idle[k]);
  std::vector<PlaneTask> out;
  out.reserve(tasks.size());
  for (size_t k = 0;; ++k) {
    bool any = false;
    for (int x = 0; x < kXcds; ++x)
      if (k < part[x].size()) {
        out.push_back(part[x][k].t);
        any = true;
      }
    if (!any) break;
  }
  return out;
}

}  // namespace

std::vector<PlaneTask> forward_plane_tasks(const std::vector<const PassDesc*>& items, bool dedup) {
  std::vector<OrderedTask> tasks;
  for (const PlaneGroup& g : plane_groups(items, dedup)) push_tasks(g, g.occ, &tasks);
  return order_tasks(std::move(tasks));
}

// Forward tasks and z-pass plane tables of one convolution of a view whose
// block spectra lie in one arena (block-batched iteration).  A distinct plane
// is transformed once when one of its occurrences lies outside its own
// block's output planes [zlo, zhi): that copy is never overwritten by a z
// pass, so every other occurrence reads it through its table and the z passes
// can run in any order.  A plane whose occurrences are all output planes is
// stored into each of them (as forward_plane_tasks does).  ztab: per z item
// (in the order of zp) q.z plane indices into the arena.
bool plan_conv_planes(const std::vector<PassDesc>& fwd, const std::vector<PassDesc>& zp, const float2* arena,
                      int64_t elems, std::vector<PlaneTask>* tasks, std::vector<unsigned short>* ztab) {
  if (fwd.empty() || fwd.size() != zp.size() || !arena || elems <= 0) return false;
  const int qz = fwd[0].g.q.z;
  std::vector<const PassDesc*> items;
  std::vector<int> slot;
  for (size_t i = 0; i < fwd.size(); ++i) {
    if (fwd[i].g.q.z != qz || fwd[i].spec != zp[i].spec) return false;
    const int64_t off = fwd[i].spec - arena;
    if (off < 0 || off % elems) return false;
    slot.push_back((int)(off / elems));
    items.push_back(&fwd[i]);
  }
  int max_slot = 0;
  for (int sl : slot) max_slot = std::max(max_slot, sl);
  if ((int64_t)(max_slot + 1) * qz > 65535 || (int64_t)fwd.size() * qz > ztab_entries(qz)) return false;
  ztab->assign(fwd.size() * qz, 0);
  std::vector<OrderedTask> out;
  for (const PlaneGroup& g : plane_groups(items, g_plane_dedup)) {
    const std::pair<int, int>* keep = nullptr;
    for (const auto& o : g.occ)
      if (o.second < zp[o.first].zlo || o.second >= zp[o.first].zhi) {
        keep = &o;
        break;
      }
    if (keep) {
      push_tasks(g, {*keep}, &out);
      for (const auto& o : g.occ) (*ztab)[o.first * qz + o.second] = (unsigned short)(slot[keep->first] * qz + keep->second);
    } else {
      push_tasks(g, g.occ, &out);
      for (const auto& o : g.occ) (*ztab)[o.first * qz + o.second] = (unsigned short)(slot[o.first] * qz + o.second);
    }
  }
  *tasks = order_tasks(std::move(out));
  return true;
}

// The same for the separate x and y passes (windows wider than the fused
// plane kernels take): the x and y forward transform only the planes listed
// in tasks, (item, plane) -- one per distinct plane that has an occurrence
// outside its own block's output planes (read there through the tables by
// every occurrence), and every occurrence of the others.
bool plan_sep_planes(const std::vector<PassDesc>& fwd, const std::vector<PassDesc>& zp, const float2* arena,
                     int64_t elems, std::vector<int2>* tasks, std::vector<unsigned short>* ztab) {
  if (fwd.empty() || fwd.size() != zp.size() || !arena || elems <= 0) return false;
  const int qz = fwd[0].g.q.z;
  std::vector<const PassDesc*> items;
  std::vector<int> slot;
  for (size_t i = 0; i < fwd.size(); ++i) {
    if (fwd[i].g.q.z != qz || fwd[i].spec != zp[i].spec) return false;
    const int64_t off = fwd[i].spec - arena;
    if (off < 0 || off % elems) return false;
    slot.push_back((int)(off / elems));
    items.push_back(&fwd[i]);
  }
  int max_slot = 0;
  for (int sl : slot) max_slot = std::max(max_slot, sl);
  if ((int64_t)(max_slot + 1) * qz > 65535 || (int64_t)fwd.size() * qz > ztab_entries(qz)) return false;
  ztab->assign(fwd.size() * qz, 0);
  tasks->clear();
  for (const PlaneGroup& g : plane_groups(items, g_plane_dedup)) {
    const std::pair<int, int>* keep = nullptr;
    for (const auto& o : g.occ)
      if (o.second < zp[o.first].zlo || o.second >= zp[o.first].zhi) {
        keep = &o;
        break;
      }
    if (keep) {
      tasks->push_back(make_int2(keep->first, keep->second));
      for (const auto& o : g.occ) (*ztab)[o.first * qz + o.second] = (unsigned short)(slot[keep->first] * qz + keep->second);
    } else {
      for (const auto& o : g.occ) {
        tasks->push_back(make_int2(o.first, o.second));
        (*ztab)[o.first * qz + o.second] = (unsigned short)(slot[o.first] * qz + o.second);
      }
    }
  }
  // the planes of one item and position next to each other: neighbouring
  // workgroups read neighbouring source rows
  std::stable_sort(tasks->begin(), tasks->end(), [](int2 a, int2 b) { return a.y != b.y ? a.y < b.y : a.x < b.x; });
  return true;
}
}  // namespace mvd

using namespace mvd;

struct mvd_session {
  int device = 0;
  hipStream_t stream = nullptr;
  Int3 img{0, 0, 0};
  int64_t n = 0;
  int nviews = 0;

  struct View {
    std::vector<float> psf_raw;  // as given
    std::vector<float> k1, k2;   // kernel1 (normalised) and kernel2
    Int3 kd{0, 0, 0};
    bool has_weight = false;
    const ShapePlan* plan = nullptr;
    std::vector<ConvGeom> blocks;
    DeviceBuf k1hat, k2hat;
    int nwg_update = 0;
  };
  std::vector<View> views;

  DeviceBuf d_views, d_weights, d_wsum, d_psi, d_ratio, d_spec, d_part_sum, d_part_max, d_stats, d_tmp;
  DeviceBuf d_fi_sum, d_fi_cnt;
  // likelihood gradient workspace
  DeviceBuf g_spec_a, g_spec_b, g_spec_c, g_spec_d, g_box, g_khat, g_psi, g_part_sum, g_part_max;
  DeviceBuf g_adj;  // the views' adjoint spectra summed (views sharing one window)
  DeviceBuf g_out;  // per call: the views' NLL sums (double) and PSF gradients (float), copied out at once
  DeviceBuf g_rsum;  // partials and the double sum of one block's residual
  DeviceBuf g_crop, g_gpsi;
  bool g_gpsi_ready = false;
  PlanCache plans;
  Timer timer;
  // The blocks of a view convolved together: one launch per pass and kernel
  // class for all blocks (one spectrum buffer per block).  Built at the first
  // iteration after prepare / new PSFs / new weights; off when the extra
  // spectra do not fit (the per-block loop then reuses d_spec).
  struct BlockBatch {
    bool built = false, enabled = true;
    DeviceBuf arena;                 // the block spectra, one after another
    std::vector<DeviceBuf> ztab;     // z-pass plane tables, per view and convolution
    std::vector<std::vector<PassLaunch>> view;  // per view: convolve1 passes, then convolve2 passes
  } bb;

  int fft_policy = MVD_FFT_SMOOTH;
  int psf_type = MVD_PSF_EFFICIENT_BAYESIAN;
  double lambda = 0.0;
  int norm_threads = 8;
  bool use_blocks = false;
  Int3 block{0, 0, 0};
  int weight_mode = -1;  // -1: decided at prepare
  double osem = 1.0;
  float wconst = 1.f;  // weight of every voxel without weight images
  bool prepared = false, psi_ready = false;

  ~mvd_session() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int timing_class(int kind, int path, int epi) {
  switch (kind) {
    case kPassXForward: return path == 2 ? kTXYForward : kTXForward;
    case kPassYForward: return kTYForward;
    case kPassZConv: return kTZConv;
    case kPassYInverse: return kTYInverse;
    default:
      if (path == 2) return epi == kEpiQuotient ? kTXYInvQuotient : epi == kEpiUpdate ? kTXYInvUpdate : kTXYInvStore;
      return epi == kEpiQuotient ? kTXInvQuotient : epi == kEpiUpdate ? kTXInvUpdate : kTXInvStore;
  }
}

// Groups the items of one pass by kernel class -- (path, template selector,
// epilogue) -- and appends one launch per group.
// fwd_tasks: the task list of the fused forward group when the caller planned
// it (plan_conv_planes); otherwise forward_plane_tasks makes one.
int build_pass_launches(const std::vector<PassDesc>& descs, std::vector<PassLaunch>* out,
                        const std::vector<PlaneTask>* fwd_tasks = nullptr,
                        const std::vector<int2>* plane_tasks = nullptr, bool wide = false) {
  // z items (register kernels) that share a kernel spectrum and shape with at
  // least one other item form their own group: one "lines" launch reads the
  // kernel spectrum once for all of them (the blocks of one view of one stack)
  // (the lines kernel addresses each spectrum with 32-bit buffer offsets:
  // spectra of 2^31 bytes or more take the per-item kernel)
  std::map<std::pair<const float2*, const ShapePlan*>, int> kshare;
  for (auto& d : descs)
    if (d.kind == kPassZConv && d.kspec && pass_is_fast(d) && d.sp->spectrum_elems() * 8 < (int64_t(1) << 31))
      ++kshare[{d.kspec, d.sp}];
  std::map<std::tuple<int, int, int, uintptr_t>, std::vector<const PassDesc*>> groups;
  for (auto& d : descs) {
    int cls = -1, path = 0;
    unsigned char tmp[kMaxParamBytes];
    int gx, gy;
    if (d.xy) {
      path = 2;
      build_xy_params(d, tmp, &gx, &gy, &cls);
    } else if (pass_is_fast(d)) {
      path = 1;
      build_fast_params(d, tmp, &gx, &gy, &cls);
    }
    // (the 384-point lines kernel exists with plane tables only: without them
    // it needs more than 256 VGPRs; those items take the batched z kernel)
    uintptr_t lines = 0;
    if (path == 1 && d.kind == kPassZConv && d.kspec && kshare[{d.kspec, d.sp}] > 1 &&
        (d.ztab != nullptr || d.sp->q.z != 384))
      lines = reinterpret_cast<uintptr_t>(d.kspec);
    groups[{path, cls, d.kind == kPassXInverse ? d.e.mode : 0, lines}].push_back(&d);
  }
  for (auto& kv : groups) {
    PassLaunch L;
    L.kind = kv.second.front()->kind;
    L.path = std::get<0>(kv.first);
    L.cls = std::get<1>(kv.first);
    L.epi = std::get<2>(kv.first);
    L.timing = timing_class(L.kind, L.path, L.epi);
    std::vector<unsigned char> blob;
    std::vector<int> prefix(1, 0), gxs;
    for (const PassDesc* d : kv.second) {
      unsigned char tmp[kMaxParamBytes];
      int gx = 0, gy = 0, cls = 0;
      size_t lds = 0;
      const int n = L.path == 2   ? build_xy_params(*d, tmp, &gx, &gy, &cls)
                    : L.path == 1 ? build_fast_params(*d, tmp, &gx, &gy, &cls)
                                  : build_generic_params(*d, tmp, &gx, &gy, &lds);
      MVD_REQUIRE(n > 0 && n <= kMaxParamBytes, "internal: parameter block size");
      blob.insert(blob.end(), tmp, tmp + n);
      gxs.push_back(std::max(gx, 1));
      prefix.push_back(prefix.back() + gx * gy);
      L.lds = std::max(L.lds, lds);
    }
    L.nitems = (int)kv.second.size();
    L.nwg = prefix.back();
    if (std::get<3>(kv.first)) {
      unsigned char tmp[kMaxParamBytes];
      int cls = 0;
      build_fast_params(*kv.second.front(), tmp, &L.lx, &L.ly, &cls);
      L.lines = true;
      int with_table = 0;
      for (const PassDesc* d : kv.second) with_table += d->ztab != nullptr;
      MVD_REQUIRE(with_table == 0 || with_table == L.nitems, "internal: z items with and without plane tables");
      L.table = with_table > 0 ? (wide ? 2 : 1) : 0;
      // the lines kernel loads an item's lines while it stores the previous
      // item's: the spectra must be distinct buffers
      std::set<const float2*> specs;
      for (const PassDesc* d : kv.second)
        MVD_REQUIRE(specs.insert(d->spec).second, "internal: z items of one launch share a spectrum");
    }
    MVD_CHECK(L.items.ensure(blob.size()));
    MVD_CHECK(hipMemcpy(L.items.p, blob.data(), blob.size(), hipMemcpyHostToDevice));
    MVD_CHECK(L.prefix.ensure(sizeof(int) * prefix.size()));
    MVD_CHECK(hipMemcpy(L.prefix.p, prefix.data(), sizeof(int) * prefix.size(), hipMemcpyHostToDevice));
    MVD_CHECK(L.gx.ensure(sizeof(int) * gxs.size()));
    MVD_CHECK(hipMemcpy(L.gx.p, gxs.data(), sizeof(int) * gxs.size(), hipMemcpyHostToDevice));
    if (L.path == 2 && L.kind == kPassXForward && g_xcd_schedule) {
      const std::vector<PlaneTask> tasks = fwd_tasks ? *fwd_tasks : forward_plane_tasks(kv.second, g_plane_dedup);
      size_t stores = 0;
      for (const PlaneTask& t : tasks) stores += t.ndest;
      // every plane is stored once, unless the caller planned plane tables
      // (then shared planes are stored only where the tables read them)
      MVD_REQUIRE(!tasks.empty() && (fwd_tasks ? (int)stores <= L.nwg : (int)stores == L.nwg),
                  "internal: forward task list");
      MVD_CHECK(L.sched.ensure(sizeof(PlaneTask) * tasks.size()));
      MVD_CHECK(hipMemcpy(L.sched.p, tasks.data(), sizeof(PlaneTask) * tasks.size(), hipMemcpyHostToDevice));
      L.nwg = (int)tasks.size();
      L.has_sched = true;
    }
    if (plane_tasks && L.path == 1 && (L.kind == kPassXForward || L.kind == kPassYForward)) {
      // every item of the pass in this group, in the order of descs
      MVD_REQUIRE(kv.second.size() == descs.size(), "internal: plane tasks over several launch groups");
      const PassDesc& d0 = *kv.second.front();
      if (L.kind == kPassXForward) {
        const int rows = x_fast_rows_per_wg(d0.g.q.x / 2);
        MVD_REQUIRE(rows > 0 && d0.g.q.y % rows == 0, "internal: rows per plane");
        L.wpt = d0.g.q.y / rows;
      } else {
        L.wpt = gxs.front();
      }
      for (const PassDesc* d : kv.second) MVD_REQUIRE(d->g.q.y == d0.g.q.y && d->zlo == 0, "internal: plane task items");
      MVD_CHECK(L.ptasks.ensure(sizeof(int2) * plane_tasks->size()));
      MVD_CHECK(hipMemcpy(L.ptasks.p, plane_tasks->data(), sizeof(int2) * plane_tasks->size(), hipMemcpyHostToDevice));
      L.nwg = (int)plane_tasks->size() * L.wpt;
    }
    out->push_back(std::move(L));
  }
  return 0;
}

hipError_t enqueue_pass_launch(const PassLaunch& L, hipStream_t st) {
  if (L.reduce) return launch_reduce_stats_batch(L.items.as<ReduceItem>(), L.nitems, st);
  if (L.lines) return launch_z_lines(L.cls, L.items.p, L.nitems, L.lx, L.ly, st, L.table);
  if (L.path == 2)
    return launch_batch_xy(L.kind, L.cls, L.items.p, L.prefix.as<int>(), L.gx.as<int>(), L.nitems, L.nwg, st,
                           L.has_sched ? L.sched.as<PlaneTask>() : nullptr);
  if (L.path == 1)
    return launch_batch_fast(L.kind, L.cls, L.items.p, L.prefix.as<int>(), L.gx.as<int>(), L.nitems, L.nwg, st,
                             L.wpt ? L.ptasks.as<int2>() : nullptr, L.wpt);
  return launch_batch_generic(L.kind, L.epi, L.items.p, L.prefix.as<int>(), L.gx.as<int>(), L.nitems, L.nwg, L.lds,
                              st);
}

// The pass items of one convolution of view v over all its blocks: convolve1
// (conv 0: psi -> ratio, quotient epilogue) or convolve2 (conv 1: ratio -> psi,
// update epilogue with statistics partials), block spectra in spec[block].
void conv_pass_descs(mvd_session* s, int v, int conv, const std::vector<float2*>& spec,
                     std::vector<PassDesc> (&pass)[5]) {
  auto& view = s->views[v];
  const int wm = s->weight_mode == MVD_WEIGHTS_NONE ? kWeightNone
                 : s->weight_mode == MVD_WEIGHTS_VIRTUAL ? kWeightVirtual
                                                         : kWeightArray;
  int off = 0;
  for (size_t bi = 0; bi < view.blocks.size(); ++bi) {
    const ConvGeom& g = view.blocks[bi];
    PassDesc d{};
    d.sp = view.plan;
    d.g = g;
    d.spec = spec[bi];
    d.src = conv == 0 ? s->d_psi.as<float>() : s->d_ratio.as<float>();
    d.gather = conv == 0 ? kGatherMirror : kGatherOne;
    d.kspec = conv == 0 ? view.k1hat.as<float2>() : view.k2hat.as<float2>();
    d.scale = 1.f;
    const int fused = xy_fused(*view.plan) ? 1 : 0;
    PassDesc x = d;
    x.kind = kPassXForward;
    x.xy = fused;
    pass[0].push_back(x);
    PassDesc yf = d;
    yf.kind = kPassYForward;
    yf.zlo = 0;
    yf.zhi = view.plan->q.z;
    yf.ylo = 0;
    yf.yhi = view.plan->q.y;
    yf.dir = -1.f;
    if (!fused) pass[1].push_back(yf);
    PassDesc z = d;
    z.kind = kPassZConv;
    z.zlo = g.eff_local.z;
    z.zhi = g.eff_local.z + g.eff_size.z;
    z.dir = -1.f;
    pass[2].push_back(z);
    PassDesc yi = d;
    yi.kind = kPassYInverse;
    yi.zlo = g.eff_local.z;
    yi.zhi = g.eff_local.z + g.eff_size.z;
    yi.ylo = g.eff_local.y;
    yi.yhi = g.eff_local.y + g.eff_size.y;
    yi.dir = 1.f;
    if (!fused) pass[3].push_back(yi);
    PassDesc xi = d;
    xi.kind = kPassXInverse;
    xi.xy = fused;
    if (conv == 0) {
      xi.e.mode = kEpiQuotient;
      xi.e.dst = s->d_ratio.as<float>();
      xi.e.img = s->d_views.as<float>() + v * s->n;
    } else {
      xi.e.mode = kEpiUpdate;
      xi.e.dst = s->d_psi.as<float>();
      xi.e.weight_mode = wm;
      xi.e.weight = wm == kWeightNone ? nullptr : s->d_weights.as<float>() + v * s->n;
      xi.e.wsum = s->d_wsum.as<float>();
      xi.e.osem = s->osem;
      xi.e.wconst = s->wconst;
      xi.e.lambda = s->lambda;
      xi.e.part_sum = s->d_part_sum.as<double>() + off;
      xi.e.part_max = s->d_part_max.as<float>() + off;
      off += conv_inverse_workgroups(*view.plan, g);
    }
    pass[4].push_back(xi);
  }
}

// Launch lists of the block-batched iteration (one spectrum buffer per block).
// Off for single-block views and when the extra spectra would take more than
// half of the free device memory.
int build_block_batch(mvd_session* s) {
  auto& bb = s->bb;
  bb.view.clear();
  bb.ztab.clear();
  size_t nb = 0, elems = 0;
  for (auto& v : s->views) {
    nb = std::max(nb, v.blocks.size());
    elems = std::max<size_t>(elems, v.plan->spectrum_elems());
  }
  size_t free_b = 0, total_b = 0;
  MVD_CHECK(hipMemGetInfo(&free_b, &total_b));
  const size_t need = nb * elems * sizeof(float2);
  bb.built = true;
  bb.enabled = nb > 1 && need <= bb.arena.bytes + free_b / 2;
  if (!bb.enabled) return 0;
  MVD_CHECK(bb.arena.ensure(need));
  MVD_CHECK(hipMemsetAsync(bb.arena.p, 0, need, s->stream));  // finite row padding
  std::vector<float2*> spec;
  for (size_t b = 0; b < nb; ++b) spec.push_back(bb.arena.as<float2>() + b * elems);
  // plane tables: the fused planes' lines kernel addresses the arena with
  // 32-bit offsets; the separate passes' one with 64-bit ones past 2^31 bytes
  const bool wide = need >= (size_t(1) << 31);
  const bool tables = g_xcd_schedule && !wide;
  bb.view.resize(s->nviews);
  for (int v = 0; v < s->nviews; ++v)
    for (int conv = 0; conv < 2; ++conv) {
      std::vector<PassDesc> pass[5];
      conv_pass_descs(s, v, conv, spec, pass);
      std::vector<PlaneTask> tasks;
      std::vector<int2> ptasks;
      std::vector<unsigned short> tab;
      const bool planned = tables && pass[0].size() == pass[2].size() && !pass[0].empty() && pass[0][0].xy &&
                           pass_is_fast(pass[2][0]) &&
                           plan_conv_planes(pass[0], pass[2], bb.arena.as<float2>(), (int64_t)elems, &tasks, &tab);
      const bool sep = !planned && g_xcd_schedule && pass[0].size() == pass[2].size() && !pass[0].empty() &&
                       !pass[0][0].xy && pass[1].size() == pass[0].size() && pass_is_fast(pass[0][0]) &&
                       pass_is_fast(pass[1][0]) && pass_is_fast(pass[2][0]) &&
                       pass[0][0].g.q.y % x_fast_rows_per_wg(pass[0][0].g.q.x / 2) == 0 &&
                       plan_sep_planes(pass[0], pass[2], bb.arena.as<float2>(), (int64_t)elems, &ptasks, &tab);
      if (planned || sep) {
        bb.ztab.emplace_back();
        DeviceBuf& dt = bb.ztab.back();
        MVD_CHECK(dt.ensure(sizeof(unsigned short) * tab.size()));
        MVD_CHECK(hipMemcpy(dt.p, tab.data(), sizeof(unsigned short) * tab.size(), hipMemcpyHostToDevice));
        const int qz = pass[2][0].g.q.z;
        for (size_t i = 0; i < pass[2].size(); ++i) {
          pass[2][i].ztab = dt.as<unsigned short>() + i * qz;
          pass[2][i].arena = bb.arena.as<float2>();
          pass[2][i].arena_bytes = wide ? 0u : (unsigned)need;
        }
      }
      for (int p = 0; p < 5; ++p) {
        int rc = build_pass_launches(pass[p], &bb.view[v], p == 0 && planned ? &tasks : nullptr,
                                     sep && p < 2 ? &ptasks : nullptr, wide);
        if (rc) return rc;
      }
    }
  return 0;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// One FFT convolution of a window: forward passes, kernel product, inverse
// passes and the fused epilogue, all on s->stream.
int run_conv(mvd_session* s, const ShapePlan& sp, const ConvGeom& g, const float* src, int gather,
             const float2* khat, const EpilogueArgs& e, float2* spec) {
  hipStream_t st = s->stream;
  Timer& T = s->timer;
  const int zlo = g.eff_local.z, zhi = g.eff_local.z + g.eff_size.z;
  const int ylo = g.eff_local.y, yhi = g.eff_local.y + g.eff_size.y;
  if (xy_fused(sp)) {
    T.begin(st);
    MVD_CHECK(launch_xy_forward(sp, g, src, gather, spec, st));
    T.end(st, kTXYForward);
    T.begin(st);
    MVD_CHECK(launch_z(sp, spec, khat, 1.f, zlo, zhi, st));
    T.end(st, kTZConv);
    T.begin(st);
    MVD_CHECK(launch_xy_inverse(sp, g, spec, e, st));
    T.end(st, e.mode == kEpiQuotient ? kTXYInvQuotient : e.mode == kEpiUpdate ? kTXYInvUpdate : kTXYInvStore);
    return 0;
  }
  T.begin(st);
  MVD_CHECK(launch_x_forward(sp, g, src, gather, spec, st));
  T.end(st, kTXForward);
  T.begin(st);
  MVD_CHECK(launch_y(sp, spec, 0, sp.q.z, 0, sp.q.y, -1.f, st));
  T.end(st, kTYForward);
  T.begin(st);
  MVD_CHECK(launch_z(sp, spec, khat, 1.f, zlo, zhi, st));
  T.end(st, kTZConv);
  T.begin(st);
  MVD_CHECK(launch_y(sp, spec, zlo, zhi, ylo, yhi, 1.f, st));
  T.end(st, kTYInverse);
  T.begin(st);
  int nwg = 0;
  MVD_CHECK(launch_x_inverse(sp, g, spec, e, st, &nwg));
  T.end(st, e.mode == kEpiQuotient ? kTXInvQuotient : e.mode == kEpiUpdate ? kTXInvUpdate : kTXInvStore);
  return 0;
}

// Spectrum of a kernel wrapped into the FFT box of sp, scaled by 1/(qx qy qz).
int kernel_spectrum(mvd_session* s, const ShapePlan& sp, const std::vector<float>& k, Int3 kd, DeviceBuf* out) {
  MVD_CHECK(out->ensure(sizeof(float2) * sp.spectrum_elems()));
  // the generic x pass leaves the row padding [hx, pitch) unwritten; zeros
  // there keep the (discarded) padding lanes of the z pass finite.  The
  // register x pass stores the padding as zeros with every row.
  if (!fast_size(sp.q.x / 2))
    MVD_CHECK(hipMemsetAsync(out->p, 0, sizeof(float2) * sp.spectrum_elems(), s->stream));
  DeviceBuf dk;
  MVD_CHECK(dk.ensure(sizeof(float) * k.size()));
  MVD_CHECK(hipMemcpyAsync(dk.p, k.data(), sizeof(float) * k.size(), hipMemcpyHostToDevice, s->stream));
  ConvGeom g{};
  g.src = kd;
  g.off = {0, 0, 0};
  g.win = sp.q;
  g.q = sp.q;
  Timer& T = s->timer;
  T.begin(s->stream);
  MVD_CHECK(launch_x_forward(sp, g, dk.as<float>(), kGatherKernelWrap, out->as<float2>(), s->stream));
  T.end(s->stream, kTXForward);
  T.begin(s->stream);
  MVD_CHECK(launch_y(sp, out->as<float2>(), 0, sp.q.z, 0, sp.q.y, -1.f, s->stream));
  T.end(s->stream, kTYForward);
  const long double n = (long double)sp.q.x * sp.q.y * sp.q.z;
  const float scale = (float)(1.0L / n);
  T.begin(s->stream);
  MVD_CHECK(launch_z(sp, out->as<float2>(), nullptr, scale, 0, sp.q.z, s->stream));
  T.end(s->stream, kTZForward);
  // The zero-frequency value is the kernel's sum over N: set from the sum in
  // long double, rounded once.  Its FP32 transform can be an ulp or two off,
  // which scales every convolution with the kernel by as much; the likelihood
  // gradient in the PSF responds to that scale about a thousand times more
  // strongly than to the PSF values themselves near a fitted estimate.
  long double sum = 0.0L;
  for (float v : k) sum += v;
  const float2 dc = make_float2((float)(sum / n), 0.f);
  MVD_CHECK(hipMemcpyAsync(out->p, &dc, sizeof(dc), hipMemcpyHostToDevice, s->stream));
  MVD_CHECK(hipStreamSynchronize(s->stream));
  return 0;
}

// "Same"-size convolution with zero out-of-bounds, as the compound-kernel
// construction calls FFTConvolution (deconv/MVDeconFFT.java:222-248, 285-295).
int same_conv_zero(mvd_session* s, const std::vector<float>& in, Int3 din, const std::vector<float>& ker, Int3 dk,
                   std::vector<float>* out) {
  ConvGeom g{};
  g.src = din;
  g.off = {-dk.x / 2, -dk.y / 2, -dk.z / 2};
  g.win = {din.x + dk.x - 1, din.y + dk.y - 1, din.z + dk.z - 1};
  g.q = fft_shape_for(g.win);
  g.eff_local = {dk.x / 2, dk.y / 2, dk.z / 2};
  g.eff_size = din;
  g.eff_offset = {0, 0, 0};
  g.dst = din;
  const ShapePlan* sp = s->plans.get(g.q);
  if (!sp) return -1;
  DeviceBuf khat, spec, din_buf, dout;
  int rc = kernel_spectrum(s, *sp, ker, dk, &khat);
  if (rc) return rc;
  MVD_CHECK(spec.ensure(sizeof(float2) * sp->spectrum_elems()));
  MVD_CHECK(din_buf.ensure(sizeof(float) * in.size()));
  MVD_CHECK(dout.ensure(sizeof(float) * in.size()));
  MVD_CHECK(hipMemcpyAsync(din_buf.p, in.data(), sizeof(float) * in.size(), hipMemcpyHostToDevice, s->stream));
  EpilogueArgs e{};
  e.mode = kEpiStore;
  e.dst = dout.as<float>();
  rc = run_conv(s, *sp, g, din_buf.as<float>(), kGatherZero, khat.as<float2>(), e, spec.as<float2>());
  if (rc) return rc;
  out->resize(in.size());
  MVD_CHECK(hipMemcpyAsync(out->data(), dout.p, sizeof(float) * in.size(), hipMemcpyDeviceToHost, s->stream));
  MVD_CHECK(hipStreamSynchronize(s->stream));
  return 0;
}

// MVDeconFFT.init for one view (deconv/MVDeconFFT.java:183-324).  With
// normalise false, kernel1 is taken as already normalised.
int init_view(mvd_session* s, int v, bool normalise = true) {
  auto& V = s->views;
  mvd_session::View& view = V[v];
  if (normalise) norm_img(view.k1, s->norm_threads);
  const int nv = s->nviews;
  const int64_t kn = (int64_t)view.k1.size();
  if (nv == 1 || s->psf_type == MVD_PSF_INDEPENDENT) {
    view.k2 = inverted(view.k1, view.kd);
  } else if (s->psf_type == MVD_PSF_EFFICIENT_BAYESIAN) {
    std::vector<float> tmp = inverted(view.k1, view.kd);
    for (int w = 0; w < nv; ++w) {
      if (w == v) continue;
      std::vector<float> a, b;
      int rc = same_conv_zero(s, inverted(view.k1, view.kd), view.kd, V[w].k1, V[w].kd, &a);
      if (rc) return rc;
      rc = same_conv_zero(s, a, view.kd, inverted(V[w].k1, V[w].kd), V[w].kd, &b);
      if (rc) return rc;
      for (int64_t i = 0; i < kn; ++i) tmp[i] = b[i] * tmp[i];
    }
    norm_img(tmp, s->norm_threads);
    view.k2 = tmp;
  } else if (s->psf_type == MVD_PSF_OPTIMIZATION_I) {
    std::vector<float> tmp = view.k1;
    for (int w = 0; w < nv; ++w) {
      if (w == v) continue;
      std::vector<float> a;
      int rc = same_conv_zero(s, view.k1, view.kd, inverted(V[w].k1, V[w].kd), V[w].kd, &a);
      if (rc) return rc;
      for (int64_t i = 0; i < kn; ++i) tmp[i] = a[i] * tmp[i];
    }
    norm_img(tmp, s->norm_threads);
    view.k2 = inverted(tmp, view.kd);
  } else {
    // OPTIMIZATION_II: float power by repeated product (MVDeconFFT.java:326-354)
    std::vector<float> e = view.k1;
    for (auto& f : e) {
      float r = f;
      for (int i = 1; i < nv; ++i) r *= f;
      f = r;
    }
    norm_img(e, s->norm_threads);
    view.k2 = inverted(e, view.kd);
  }
  return 0;
}

// FFT shape of a block window.  MVD_FFT_SMOOTH: smallest 5-smooth size per
// axis (even along x); MVD_FFT_POW2: smallest power of two (up to 1024, else
// 5-smooth); MVD_FFT_AUTOTUNE: per axis the smallest 5-smooth size and the
// smallest size the register-resident kernels take are timed on the GPU (one
// full block convolution each, best of three) and the fastest shape is kept,
// cached per window shape for the process.
static int pow2_at_least(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

// Smallest length >= w that the register-resident kernels take (along x: a
// row of 2M samples with M such a length); the power of two up to 1024 when
// none is, else the 5-smooth length sm.
static int fast_at_least(int w, bool x_axis, int sm) {
  static const int lengths[] = {64, 128, 144, 192, 256, 288, 384, 512};
  for (int m : lengths) {
    const int q = x_axis ? 2 * m : m;
    if (q >= w && fast_size(m)) return q;
  }
  const int p = pow2_at_least(w);
  return p <= 1024 ? p : sm;
}

std::mutex g_tune_mu;
struct TunedShape {
  Int3 q;
  float ms;  // one block convolution
};
std::map<std::tuple<int, int, int>, TunedShape> g_tune_cache;  // window -> FFT shape
struct TunedBlocks {
  Int3 bs, q;
};
std::map<std::tuple<int, int, int, int, int, int>, TunedBlocks> g_block_cache;  // (image, kernel) -> blocks

int time_shape(mvd_session* s, Int3 win, Int3 q, float* ms_out) {
  const ShapePlan* sp = s->plans.get(q);
  if (!sp) return -1;
  DeviceBuf spec, kspec, src;
  MVD_CHECK(spec.ensure(sizeof(float2) * sp->spectrum_elems()));
  MVD_CHECK(kspec.ensure(sizeof(float2) * sp->spectrum_elems()));
  MVD_CHECK(src.ensure(sizeof(float) * (size_t)win.x * win.y * win.z));
  MVD_CHECK(hipMemsetAsync(kspec.p, 0, sizeof(float2) * sp->spectrum_elems(), s->stream));
  MVD_CHECK(hipMemsetAsync(src.p, 0, sizeof(float) * (size_t)win.x * win.y * win.z, s->stream));
  ConvGeom g{};
  g.src = win;
  g.off = {0, 0, 0};
  g.win = win;
  g.q = q;
  g.eff_local = {0, 0, 0};
  g.eff_size = win;
  g.eff_offset = {0, 0, 0};
  g.dst = win;
  EpilogueArgs e{};
  e.mode = kEpiStore;
  e.dst = src.as<float>();
  hipEvent_t a, b;
  MVD_CHECK(hipEventCreate(&a));
  MVD_CHECK(hipEventCreate(&b));
  float best = 1e30f;
  const bool timing = s->timer.on();
  s->timer.enable(false);
  for (int rep = 0; rep < 3; ++rep) {
    MVD_CHECK(hipEventRecord(a, s->stream));
    int rc = run_conv(s, *sp, g, src.as<float>(), kGatherZero, kspec.as<float2>(), e, spec.as<float2>());
    if (rc) return rc;
    MVD_CHECK(hipEventRecord(b, s->stream));
    MVD_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    MVD_CHECK(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms);
  }
  s->timer.enable(timing);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms_out = best;
  return 0;
}

// Fastest FFT shape for a window among the smallest 5-smooth and the smallest
// register-kernel lengths per axis (all 8 combinations), timed on the GPU and
// cached.
int tune_fft_shape(mvd_session* s, Int3 win, Int3* q, float* ms_out) {
  const Int3 smooth = fft_shape_for(win);
  const Int3 reg = {fast_at_least(win.x, true, smooth.x), fast_at_least(win.y, false, smooth.y),
                    fast_at_least(win.z, false, smooth.z)};
  const auto key = std::make_tuple(win.x, win.y, win.z);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      *q = it->second.q;
      if (ms_out) *ms_out = it->second.ms;
      return 0;
    }
  }
  Int3 best = smooth;
  float best_ms = 1e30f;
  for (int c = 0; c < 8; ++c) {
    const Int3 cand = {(c & 1) ? reg.x : smooth.x, (c & 2) ? reg.y : smooth.y, (c & 4) ? reg.z : smooth.z};
    if (c && ((c & 1) && reg.x == smooth.x)) continue;
    if (c && ((c & 2) && reg.y == smooth.y)) continue;
    if (c && ((c & 4) && reg.z == smooth.z)) continue;
    float ms = 0.f;
    int rc = time_shape(s, win, cand, &ms);
    if (rc) return rc;
    if (ms < best_ms) {
      best_ms = ms;
      best = cand;
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = {best, best_ms};
  }
  *q = best;
  if (ms_out) *ms_out = best_ms;
  return 0;
}

int choose_fft_shape(mvd_session* s, Int3 win, Int3* q) {
  const Int3 smooth = fft_shape_for(win);
  if (s->fft_policy == MVD_FFT_SMOOTH) {
    *q = smooth;
    return 0;
  }
  if (s->fft_policy == MVD_FFT_POW2) {
    auto p2 = [](int w, int sm) { const int p = pow2_at_least(w); return p <= 1024 ? p : sm; };
    *q = {p2(win.x, smooth.x), p2(win.y, smooth.y), p2(win.z, smooth.z)};
    return 0;
  }
  return tune_fft_shape(s, win, q, nullptr);
}

// Autotuned block shape for a view the caller did not split (one window of
// img + K - 1 in the reference): the whole window with its best FFT shape, or
// blocks of 256 x 256 x {128, 256} whose planes take the fused x-y kernels,
// whichever convolves the image fastest (blocks x time per block).  Blocked
// and whole-image convolutions are the same operator (each block window
// carries the kernel halo and the mirror-single extension), so this changes
// rounding only.
int choose_blocks(mvd_session* s, Int3 kd, Int3* bs_out, Int3* q_out) {
  const auto key = std::make_tuple(s->img.x, s->img.y, s->img.z, kd.x, kd.y, kd.z);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_block_cache.find(key);
    if (it != g_block_cache.end()) {
      *bs_out = it->second.bs;
      *q_out = it->second.q;
      return 0;
    }
  }
  const Int3 whole = {s->img.x + kd.x - 1, s->img.y + kd.y - 1, s->img.z + kd.z - 1};
  Int3 best_bs = whole, best_q;
  float t_whole = 0.f;
  int rc = tune_fft_shape(s, whole, &best_q, &t_whole);
  if (rc) return rc;
  double best_ms = t_whole;
  for (int bz : {128, 256}) {
    const Int3 b = {256, 256, bz};
    std::vector<ConvGeom> blocks;
    if (!divide_into_blocks(s->img, kd, b, b, &blocks)) continue;    // kernel larger than the block
    float t = 0.f;
    rc = time_shape(s, b, b, &t);
    if (rc) return rc;
    if ((double)t * blocks.size() < best_ms) {
      best_ms = (double)t * blocks.size();
      best_bs = b;
      best_q = b;
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_block_cache[key] = {best_bs, best_q};
  }
  *bs_out = best_bs;
  *q_out = best_q;
  return 0;
}

int build_geometry(mvd_session* s, int v) {
  mvd_session::View& view = s->views[v];
  Int3 bs = s->use_blocks ? s->block
                          : Int3{s->img.x + view.kd.x - 1, s->img.y + view.kd.y - 1, s->img.z + view.kd.z - 1};
  Int3 q;
  int rc0 = !s->use_blocks && s->fft_policy == MVD_FFT_AUTOTUNE ? choose_blocks(s, view.kd, &bs, &q)
                                                                : choose_fft_shape(s, bs, &q);
  if (rc0) return rc0;
  view.plan = s->plans.get(q);
  if (!view.plan) return -1;
  MVD_REQUIRE(divide_into_blocks(s->img, view.kd, bs, q, &view.blocks),
              "block size smaller than the kernel (BlockGeneratorFixedSizePrecise)");
  view.nwg_update = 0;
  for (auto& g : view.blocks) view.nwg_update += conv_inverse_workgroups(*view.plan, g);
  return 0;
}

int view_spectra(mvd_session* s, int v) {
  mvd_session::View& view = s->views[v];
  int rc = kernel_spectrum(s, *view.plan, view.k1, view.kd, &view.k1hat);
  if (rc) return rc;
  return kernel_spectrum(s, *view.plan, view.k2, view.kd, &view.k2hat);
}

int ensure_workspace(mvd_session* s) {
  size_t spec = 0;
  int parts = 1;
  for (auto& view : s->views) {
    spec = std::max<size_t>(spec, (size_t)view.plan->spectrum_elems());
    parts = std::max(parts, view.nwg_update);
  }
  MVD_CHECK(s->d_spec.ensure(sizeof(float2) * spec));
  MVD_CHECK(s->d_part_sum.ensure(sizeof(double) * parts));
  MVD_CHECK(s->d_part_max.ensure(sizeof(float) * parts));
  return 0;
}

constexpr int kReduceParts = 2048;

// Sum of the per-voxel means and number of voxels with data (FirstIteration).
int first_iteration_sums(mvd_session* s, float* counts, double* sum, unsigned long long* cnt) {
  MVD_CHECK(s->d_fi_sum.ensure(sizeof(double) * kReduceParts));
  MVD_CHECK(s->d_fi_cnt.ensure(sizeof(unsigned long long) * kReduceParts));
  const int parts = (int)std::min<int64_t>(kReduceParts, std::max<int64_t>(1, (s->n + 255) / 256));
  MVD_CHECK(launch_first_iteration(s->d_views.as<float>(), s->nviews, s->n, counts,
                                   s->d_fi_sum.as<double>(), s->d_fi_cnt.as<unsigned long long>(), parts,
                                   s->stream));
  std::vector<double> ps(parts);
  std::vector<unsigned long long> pc(parts);
  MVD_CHECK(hipMemcpyAsync(ps.data(), s->d_fi_sum.p, sizeof(double) * parts, hipMemcpyDeviceToHost, s->stream));
  MVD_CHECK(hipMemcpyAsync(pc.data(), s->d_fi_cnt.p, sizeof(unsigned long long) * parts, hipMemcpyDeviceToHost,
                           s->stream));
  MVD_CHECK(hipStreamSynchronize(s->stream));
  double t = 0;
  unsigned long long c = 0;
  for (int i = 0; i < parts; ++i) {
    t += ps[i];
    c += pc[i];
  }
  *sum = t;
  *cnt = c;
  return 0;
}

}  // namespace

extern "C" {

const char* mvd_last_error(void) { return g_last_error.c_str(); }

int mvd_create(int device, int nx, int ny, int nz, int nviews, mvd_session** out) {
  MVD_REQUIRE(out != nullptr, "out is NULL");
  *out = nullptr;
  MVD_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nviews > 0, "invalid image size or view count");
  int ndev = 0;
  MVD_CHECK(hipGetDeviceCount(&ndev));
  MVD_REQUIRE(device >= 0 && device < ndev, "invalid device id");
  DeviceGuard guard(device);
  auto s = std::make_unique<mvd_session>();
  s->device = device;
  s->img = {nx, ny, nz};
  s->n = (int64_t)nx * ny * nz;
  MVD_REQUIRE(s->n < (int64_t(1) << 31), "images of 2^31 voxels or more are not supported");
  s->nviews = nviews;
  s->views.reserve(nviews);
  MVD_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  MVD_CHECK(s->d_views.ensure(sizeof(float) * s->n * nviews));
  MVD_CHECK(s->d_psi.ensure(sizeof(float) * s->n));
  MVD_CHECK(s->d_ratio.ensure(sizeof(float) * s->n));
  *out = s.release();
  return 0;
}

int mvd_destroy(mvd_session* s) {
  if (!s) return 0;
  DeviceGuard guard(s->device);
  (void)hipStreamSynchronize(s->stream);
  delete s;
  return 0;
}

static int check_view_args(mvd_session* s, const float* psf, int kx, int ky, int kz) {
  MVD_REQUIRE(s && psf, "NULL argument");
  MVD_REQUIRE((int)s->views.size() < s->nviews, "all views already added");
  MVD_REQUIRE(kx > 0 && ky > 0 && kz > 0 && (kx & 1) && (ky & 1) && (kz & 1),
              "kernel sizes must be odd (BlockGeneratorFixedSizePrecise)");
  MVD_REQUIRE(kx <= s->img.x + 4096 && ky <= s->img.y + 4096, "kernel too large");
  return 0;
}

static void register_view(mvd_session* s, const float* psf, int kx, int ky, int kz, bool has_weight) {
  s->views.emplace_back();
  auto& view = s->views.back();
  view.kd = {kx, ky, kz};
  view.psf_raw.assign(psf, psf + (size_t)kx * ky * kz);
  view.has_weight = has_weight;
  s->prepared = false;
}

int mvd_add_view(mvd_session* s, const float* img, const float* weight, const float* psf, int kx, int ky, int kz) {
  if (int rc = check_view_args(s, psf, kx, ky, kz)) return rc;
  MVD_REQUIRE(img, "NULL argument");
  DeviceGuard guard(s->device);
  const int v = (int)s->views.size();
  if (weight) {
    MVD_CHECK(s->d_weights.ensure(sizeof(float) * s->n * s->nviews));
    MVD_CHECK(staged_h2d(s->d_weights.as<float>() + v * s->n, weight, sizeof(float) * s->n, s->stream));
  }
  MVD_CHECK(staged_h2d(s->d_views.as<float>() + v * s->n, img, sizeof(float) * s->n, s->stream));
  register_view(s, psf, kx, ky, kz, weight != nullptr);
  return 0;
}

int mvd_add_view_transformed(mvd_session* s, const float* raw, int rx, int ry, int rz, const double* inverse_affine,
                             const long long* offset, const float* border, const float* range, const float* psf,
                             int kx, int ky, int kz) {
  if (int rc = check_view_args(s, psf, kx, ky, kz)) return rc;
  MVD_REQUIRE(raw && inverse_affine && offset, "NULL argument");
  MVD_REQUIRE((border == nullptr) == (range == nullptr), "weights need both border and range");
  DeviceGuard guard(s->device);
  const int v = (int)s->views.size();
  TransformArgs a{};
  a.rsize[0] = rx;
  a.rsize[1] = ry;
  a.rsize[2] = rz;
  for (int i = 0; i < 12; ++i) a.inv[i] = inverse_affine[i];
  for (int d = 0; d < 3; ++d) a.off[d] = (float)offset[d];
  a.nx = s->img.x;
  a.ny = s->img.y;
  a.nz = s->img.z;
  a.out_img = s->d_views.as<float>() + v * s->n;
  if (border) {
    MVD_CHECK(s->d_weights.ensure(sizeof(float) * s->n * s->nviews));
    for (int d = 0; d < 3; ++d) {
      a.border[d] = border[d];
      a.range[d] = range[d];
    }
    a.out_w = s->d_weights.as<float>() + v * s->n;
    a.lut = blend_lut_device(s->device);
    MVD_REQUIRE(a.lut, "blending table allocation failed");
  }
  a.raw = raw;  // checked for presence; replaced by the device copy below
  if (const char* msg = check_transform_args(a)) MVD_REQUIRE(false, msg);
  DeviceBuf d_raw;
  MVD_CHECK(d_raw.ensure(sizeof(float) * (size_t)rx * ry * rz));
  MVD_CHECK(staged_h2d(d_raw.p, raw, sizeof(float) * (size_t)rx * ry * rz, s->stream));
  a.raw = d_raw.as<float>();
  s->timer.begin(s->stream);
  MVD_CHECK(launch_transform_view(a, s->stream));
  s->timer.end(s->stream, kTTransform);
  MVD_CHECK(hipStreamSynchronize(s->stream));
  s->timer.collect();
  register_view(s, psf, kx, ky, kz, border != nullptr);
  return 0;
}

int mvd_set_weights(mvd_session* s, int mode, int osem_index, double osem_value, int norm_threads,
                    double* min_avg_out) {
  MVD_REQUIRE(s, "NULL session");
  MVD_REQUIRE((int)s->views.size() == s->nviews, "add all views first");
  MVD_REQUIRE(mode >= MVD_WEIGHTS_NONE && mode <= MVD_WEIGHTS_VIRTUAL, "invalid weight mode");
  MVD_REQUIRE(norm_threads >= 1, "norm_threads must be >= 1");
  s->bb.built = false;
  bool all_w = true;
  for (auto& v : s->views) all_w = all_w && v.has_weight;
  MVD_REQUIRE(mode == MVD_WEIGHTS_NONE || all_w, "this weight mode needs a weight image for every view");
  DeviceGuard guard(s->device);
  // Without weights there is no WeightNormalizer and the overlap counts keep
  // their initial 0 (ProcessForDeconvolution.java:89-90,331-351).
  double mn = mode == MVD_WEIGHTS_NONE ? 0 : 1, avg = mode == MVD_WEIGHTS_NONE ? 0 : 1;
  if (mode == MVD_WEIGHTS_PRECOMPUTED || mode == MVD_WEIGHTS_VIRTUAL) {
    // WeightNormalizer.process (deconv/WeightNormalizer.java:73-130): per portion
    // of 2*threads, min count and average count; the overall average is the mean
    // of the portion averages.
    if (mode == MVD_WEIGHTS_VIRTUAL) MVD_CHECK(s->d_wsum.ensure(sizeof(float) * s->n));
    DeviceBuf pmin, pcnt;
    MVD_CHECK(pmin.ensure(sizeof(int) * kReduceParts));
    MVD_CHECK(pcnt.ensure(sizeof(unsigned long long) * kReduceParts));
    int gmin = s->nviews;
    double avg_sum = 0;
    auto ps = portions(s->n, 2 * norm_threads);
    for (auto& p : ps) {
      if (p.second == 0) {
        avg_sum += std::nan("");
        continue;
      }
      const int parts = (int)std::min<int64_t>(kReduceParts, std::max<int64_t>(1, (p.second + 255) / 256));
      MVD_CHECK(launch_weight_normalize(s->d_weights.as<float>() + p.first, s->nviews, p.second, s->n,
                                        mode == MVD_WEIGHTS_VIRTUAL ? s->d_wsum.as<float>() + p.first : nullptr,
                                        mode == MVD_WEIGHTS_VIRTUAL, pmin.as<int>(),
                                        pcnt.as<unsigned long long>(), parts, s->stream));
      std::vector<int> hm(parts);
      std::vector<unsigned long long> hc(parts);
      MVD_CHECK(hipMemcpyAsync(hm.data(), pmin.p, sizeof(int) * parts, hipMemcpyDeviceToHost, s->stream));
      MVD_CHECK(hipMemcpyAsync(hc.data(), pcnt.p, sizeof(unsigned long long) * parts, hipMemcpyDeviceToHost,
                               s->stream));
      MVD_CHECK(hipStreamSynchronize(s->stream));
      unsigned long long c = 0;
      for (int i = 0; i < parts; ++i) {
        gmin = std::min(gmin, hm[i]);
        c += hc[i];
      }
      avg_sum += (double)c / (double)p.second;
    }
    mn = std::max(1, gmin);
    avg = std::max(1.0, avg_sum / ps.size());
  }
  double osem = 1.0;
  if (osem_index == 1)
    osem = mn;
  else if (osem_index == 2)
    osem = avg;
  else if (osem_index == 3)
    osem = osem_value;
  // ProcessForDeconvolution.adjustForOSEM (deconv/ProcessForDeconvolution.java:372-405)
  if (osem != 1.0 && (mode == MVD_WEIGHTS_PRECOMPUTED || mode == MVD_WEIGHTS_AS_GIVEN))
    MVD_CHECK(launch_osem_scale(s->d_weights.as<float>(), s->n * s->nviews, (float)osem, s->stream));
  MVD_CHECK(hipStreamSynchronize(s->stream));
  // NO_WEIGHTS: the constant weight image becomes Math.min(1, 1 * (float)osem)
  // (ProcessForDeconvolution.java:385-394); osem is never NaN here.
  s->wconst = 1.f;
  if (osem != 1.0 && mode == MVD_WEIGHTS_NONE) s->wconst = std::min(1.f, 1.f * (float)osem);
  s->osem = osem;
  s->weight_mode = mode;
  if (min_avg_out) {
    min_avg_out[0] = mn;
    min_avg_out[1] = avg;
  }
  return 0;
}

int mvd_prepare(mvd_session* s, int psf_type, double lambda, const int* block_xyz, int norm_threads) {
  MVD_REQUIRE(s, "NULL session");
  MVD_REQUIRE((int)s->views.size() == s->nviews, "add all views first");
  MVD_REQUIRE(psf_type >= 0 && psf_type <= 3, "invalid PSF type");
  MVD_REQUIRE(norm_threads >= 1, "norm_threads must be >= 1");
  DeviceGuard guard(s->device);
  s->bb.built = false;
  s->psf_type = psf_type;
  s->lambda = lambda;
  s->norm_threads = norm_threads;
  s->use_blocks = block_xyz != nullptr;
  if (block_xyz) {
    MVD_REQUIRE(block_xyz[0] > 0 && block_xyz[1] > 0 && block_xyz[2] > 0, "invalid block size");
    s->block = {block_xyz[0], block_xyz[1], block_xyz[2]};
  }
  if (s->weight_mode < 0) {
    bool all_w = true;
    for (auto& v : s->views) all_w = all_w && v.has_weight;
    s->weight_mode = all_w ? MVD_WEIGHTS_AS_GIVEN : MVD_WEIGHTS_NONE;
  }
  for (auto& v : s->views) v.k1 = v.psf_raw;
  // MVDeconInput.init: views in order
  for (int v = 0; v < s->nviews; ++v) {
    int rc = init_view(s, v);
    if (rc) return rc;
  }
  for (int v = 0; v < s->nviews; ++v) {
    int rc = build_geometry(s, v);
    if (rc) return rc;
    rc = view_spectra(s, v);
    if (rc) return rc;
  }
  int rc = ensure_workspace(s);
  if (rc) return rc;
  s->prepared = true;
  return 0;
}

int mvd_set_fft_policy(mvd_session* s, int policy) {
  MVD_REQUIRE(s, "NULL session");
  MVD_REQUIRE(policy >= MVD_FFT_SMOOTH && policy <= MVD_FFT_AUTOTUNE, "invalid FFT shape policy");
  s->fft_policy = policy;
  s->prepared = false;
  return 0;
}

int mvd_set_psf(mvd_session* s, int view, const float* psf) {
  MVD_REQUIRE(s && psf, "NULL argument");
  MVD_REQUIRE(s->prepared, "mvd_prepare first");
  MVD_REQUIRE(view >= 0 && view < s->nviews, "invalid view");
  DeviceGuard guard(s->device);
  auto& v = s->views[view];
  v.psf_raw.assign(psf, psf + v.k1.size());
  v.k1 = v.psf_raw;
  s->bb.built = false;
  int rc = init_view(s, view);
  if (rc) return rc;
  return view_spectra(s, view);
}

int mvd_set_psfs(mvd_session* s, const float* const* psfs) {
  MVD_REQUIRE(s && psfs, "NULL argument");
  MVD_REQUIRE(s->prepared, "mvd_prepare first");
  for (int v = 0; v < s->nviews; ++v) MVD_REQUIRE(psfs[v], "NULL PSF");
  DeviceGuard guard(s->device);
  // every kernel1 first: the compound kernels of each view read the others'
  for (int v = 0; v < s->nviews; ++v) {
    auto& view = s->views[v];
    view.psf_raw.assign(psfs[v], psfs[v] + view.k1.size());
    view.k1 = view.psf_raw;
    norm_img(view.k1, s->norm_threads);
  }
  s->bb.built = false;
  for (int v = 0; v < s->nviews; ++v) {
    int rc = init_view(s, v, false);
    if (rc) return rc;
    rc = view_spectra(s, v);
    if (rc) return rc;
  }
  return 0;
}

int mvd_psi_sum(mvd_session* s, double* out) {
  MVD_REQUIRE(s && out, "NULL argument");
  MVD_REQUIRE(s->psi_ready, "mvd_init_psi first");
  DeviceGuard guard(s->device);
  MVD_CHECK(s->d_fi_sum.ensure(sizeof(double) * kReduceParts));
  const int parts = (int)std::min<int64_t>(kReduceParts, std::max<int64_t>(1, (s->n + 255) / 256));
  MVD_CHECK(launch_sum(s->d_psi.as<float>(), s->n, s->d_fi_sum.as<double>(), parts, s->stream));
  std::vector<double> ps(parts);
  MVD_CHECK(hipMemcpyAsync(ps.data(), s->d_fi_sum.p, sizeof(double) * parts, hipMemcpyDeviceToHost, s->stream));
  MVD_CHECK(hipStreamSynchronize(s->stream));
  double t = 0;
  for (double v : ps) t += v;
  *out = t;
  return 0;
}

int mvd_init_psi(mvd_session* s, const float* psi0, double* avg_out) {
  MVD_REQUIRE(s, "NULL session");
  MVD_REQUIRE((int)s->views.size() == s->nviews, "add all views first");
  DeviceGuard guard(s->device);
  if (psi0) {
    MVD_CHECK(staged_h2d(s->d_psi.p, psi0, sizeof(float) * s->n, s->stream));
    MVD_CHECK(hipStreamSynchronize(s->stream));
    if (avg_out) *avg_out = std::nan("");
  } else {
    double sum = 0;
    unsigned long long cnt = 0;
    int rc = first_iteration_sums(s, nullptr, &sum, &cnt);
    if (rc) return rc;
    double avg = sum / (double)cnt;
    if (std::isnan(avg)) avg = 0.5;  // MVDeconvolution.java:138-142
    MVD_CHECK(launch_fill(s->d_psi.as<float>(), (float)avg, s->n, s->stream));
    MVD_CHECK(hipStreamSynchronize(s->stream));
    if (avg_out) *avg_out = avg;
  }
  s->psi_ready = true;
  return 0;
}

int mvd_iterate(mvd_session* s, int n, double* stats) {
  MVD_REQUIRE(s, "NULL session");
  MVD_REQUIRE(s->prepared, "mvd_prepare first");
  MVD_REQUIRE(s->psi_ready, "mvd_init_psi first");
  MVD_REQUIRE(n >= 0, "negative iteration count");
  if (n == 0) return 0;
  DeviceGuard guard(s->device);
  MVD_CHECK(s->d_stats.ensure(sizeof(double) * 2 * s->nviews * (size_t)n));
  float* psi = s->d_psi.as<float>();
  float* ratio = s->d_ratio.as<float>();
  float2* spec = s->d_spec.as<float2>();
  const int wm = s->weight_mode == MVD_WEIGHTS_NONE ? kWeightNone
                 : s->weight_mode == MVD_WEIGHTS_VIRTUAL ? kWeightVirtual
                                                         : kWeightArray;
  if (!s->bb.built) {
    int rc = build_block_batch(s);
    if (rc) return rc;
  }
  // blocks of a view together: per pass one launch for all blocks
  for (int it = 0; s->bb.enabled && it < n; ++it) {
    for (int v = 0; v < s->nviews; ++v) {
      for (auto& L : s->bb.view[v]) {
        s->timer.begin(s->stream);
        MVD_CHECK(enqueue_pass_launch(L, s->stream));
        s->timer.end(s->stream, L.timing);
      }
      s->timer.begin(s->stream);
      MVD_CHECK(launch_reduce_stats(s->d_part_sum.as<double>(), s->d_part_max.as<float>(), s->views[v].nwg_update,
                                    s->d_stats.as<double>() + 2 * ((size_t)it * s->nviews + v), s->stream));
      s->timer.end(s->stream, kTReduce);
    }
  }
  // block after block through one spectrum buffer
  for (int it = 0; !s->bb.enabled && it < n; ++it) {
    for (int v = 0; v < s->nviews; ++v) {
      auto& view = s->views[v];
      const ShapePlan& sp = *view.plan;
      // convolve1 + quotient: psi -> ratio (MVDeconvolution.java:383-406)
      EpilogueArgs q{};
      q.mode = kEpiQuotient;
      q.dst = ratio;
      q.img = s->d_views.as<float>() + v * s->n;
      for (auto& g : view.blocks) {
        int rc = run_conv(s, sp, g, psi, kGatherMirror, view.k1hat.as<float2>(), q, spec);
        if (rc) return rc;
      }
      // convolve2 + update: ratio -> psi (MVDeconvolution.java:418-457)
      EpilogueArgs u{};
      u.mode = kEpiUpdate;
      u.dst = psi;
      u.weight_mode = wm;
      u.weight = wm == kWeightNone ? nullptr : s->d_weights.as<float>() + v * s->n;
      u.wsum = s->d_wsum.as<float>();
      u.osem = s->osem;
      u.wconst = s->wconst;
      u.lambda = s->lambda;
      int off = 0;
      for (auto& g : view.blocks) {
        u.part_sum = s->d_part_sum.as<double>() + off;
        u.part_max = s->d_part_max.as<float>() + off;
        int rc = run_conv(s, sp, g, ratio, kGatherOne, view.k2hat.as<float2>(), u, spec);
        if (rc) return rc;
        off += conv_inverse_workgroups(sp, g);
      }
      s->timer.begin(s->stream);
      MVD_CHECK(launch_reduce_stats(s->d_part_sum.as<double>(), s->d_part_max.as<float>(), off,
                                    s->d_stats.as<double>() + 2 * ((size_t)it * s->nviews + v), s->stream));
      s->timer.end(s->stream, kTReduce);
    }
  }
  if (stats)
    MVD_CHECK(hipMemcpyAsync(stats, s->d_stats.p, sizeof(double) * 2 * s->nviews * (size_t)n,
                             hipMemcpyDeviceToHost, s->stream));
  MVD_CHECK(hipStreamSynchronize(s->stream));
  s->timer.collect();
  return 0;
}

int mvd_get_psi(mvd_session* s, float* out, int masked) {
  MVD_REQUIRE(s && out, "NULL argument");
  MVD_REQUIRE(s->psi_ready, "mvd_init_psi first");
  DeviceGuard guard(s->device);
  if (!masked) {
    MVD_CHECK(staged_d2h(out, s->d_psi.p, sizeof(float) * s->n, s->stream));
    return 0;
  }
  // the counts of views with data into the ratio buffer (scratch between
  // iterations), then psi where a view has data, in place, and one copy
  double sum;
  unsigned long long cnt;
  int rc = first_iteration_sums(s, s->d_ratio.as<float>(), &sum, &cnt);
  if (rc) return rc;
  MVD_CHECK(launch_mask(s->d_psi.as<float>(), s->d_ratio.as<float>(), s->d_ratio.as<float>(), s->n, s->stream));
  MVD_CHECK(staged_d2h(out, s->d_ratio.p, sizeof(float) * s->n, s->stream));
  return 0;
}

int mvd_export_psi(mvd_session* s, const char* path, int bits, double min, double max, const double* calibration,
                   int big_endian, int bigtiff) {
  MVD_REQUIRE(s && path, "NULL argument");
  MVD_REQUIRE(s->psi_ready, "mvd_init_psi first");
  DeviceGuard guard(s->device);
  // the masked estimate (getPsi) into the ratio buffer: view counts first,
  // then psi where a view has data, in place
  double sum;
  unsigned long long cnt;
  int rc = first_iteration_sums(s, s->d_ratio.as<float>(), &sum, &cnt);
  if (rc) return rc;
  MVD_CHECK(launch_mask(s->d_psi.as<float>(), s->d_ratio.as<float>(), s->d_ratio.as<float>(), s->n, s->stream));
  return write_tiff_stack(s->d_ratio.as<float>(), true, s->img.x, s->img.y, s->img.z, bits, min, max, calibration,
                          big_endian != 0, bigtiff, path, s->stream);
}

int mvd_nll_grad_device(mvd_session* s, const float* psi, const float* const* psfs, double* nll, int want_grad_psi,
                        float* const* grad_psf) {
  MVD_REQUIRE(s, "NULL session");
  MVD_REQUIRE((int)s->views.size() == s->nviews, "add all views first");
  MVD_REQUIRE(psi || s->psi_ready, "no estimate: pass psi or call mvd_init_psi");
  DeviceGuard guard(s->device);
  int wmode = s->weight_mode;
  if (wmode < 0) {
    bool all_w = true;
    for (auto& v : s->views) all_w = all_w && v.has_weight;
    wmode = all_w ? MVD_WEIGHTS_AS_GIVEN : MVD_WEIGHTS_NONE;
  }
  MVD_REQUIRE(wmode != MVD_WEIGHTS_VIRTUAL, "virtual weights are not supported by mvd_nll_grad");
  const Int3 n = s->img;
  hipStream_t st = s->stream;
  s->g_gpsi_ready = false;
  const float* psrc = s->d_psi.as<float>();
  if (psi) {
    MVD_CHECK(s->g_psi.ensure(sizeof(float) * s->n));
    MVD_CHECK(staged_h2d(s->g_psi.p, psi, sizeof(float) * s->n, st));
    psrc = s->g_psi.as<float>();
  }
  // dL/dpsi is accumulated in g_gpsi.  fold_once: every view has one block
  // window covering the image, folded by the gather kernel, and the first
  // view writes the gradient instead of adding to it (no clearing pass)
  // sum_adj: besides, all views share that window and its FFT shape, so their
  // adjoint spectra (linear in the residual) are summed along the views and
  // transformed back and folded once
  bool fold_once = want_grad_psi != 0, sum_adj = fold_once && g_grad_fused && g_grad_keep_w;
  const ShapePlan* first_plan = nullptr;
  Int3 first_wd{0, 0, 0}, first_lo{0, 0, 0};
  for (int v = 0; fold_once && v < s->nviews; ++v) {
    const auto& view = s->views[v];
    Int3 wd, lo;
    const ShapePlan* pl = nullptr;
    if (s->prepared && view.plan) {
      if (view.blocks.size() != 1) {
        fold_once = false;
        break;
      }
      wd = view.blocks[0].win;
      lo = view.blocks[0].off;
      pl = view.plan;
    } else {
      wd = {n.x + view.kd.x - 1, n.y + view.kd.y - 1, n.z + view.kd.z - 1};
      lo = {-view.kd.x / 2, -view.kd.y / 2, -view.kd.z / 2};
      pl = s->plans.get(fft_shape_for(wd));
    }
    fold_once = fold_gather_ok(wd, lo, n) && fold_gather_covers(wd, lo, n);
    if (v == 0) {
      first_plan = pl;
      first_wd = wd;
      first_lo = lo;
    }
    sum_adj = sum_adj && pl && pl == first_plan && z_grad_supported(*pl) && wd.x == first_wd.x &&
              wd.y == first_wd.y && wd.z == first_wd.z && lo.x == first_lo.x && lo.y == first_lo.y &&
              lo.z == first_lo.z;
  }
  sum_adj = sum_adj && fold_once;
  if (sum_adj) MVD_CHECK(s->g_adj.ensure(sizeof(float2) * first_plan->spectrum_elems()));
  ConvGeom adj_g{};  // the shared window (sum_adj)
  float* gpsi = nullptr;
  if (want_grad_psi) {
    MVD_CHECK(s->g_gpsi.ensure(sizeof(float) * s->n));
    gpsi = s->g_gpsi.as<float>();
    if (!fold_once) MVD_CHECK(launch_fill(gpsi, 0.f, s->n, st));
  }
  auto fold = [&](int v, const ConvGeom& g) -> hipError_t {
    if (fold_once) {
      if (!fold_gather_ok(g.win, g.off, n) || !fold_gather_covers(g.win, g.off, n)) return hipErrorInvalidValue;
      return launch_fold_gather(s->g_box.as<float>(), g.win, g.off, n, gpsi, v > 0, st);
    }
    if (fold_gather_ok(g.win, g.off, n)) return launch_fold_gather(s->g_box.as<float>(), g.win, g.off, n, gpsi, true, st);
    return launch_fold_window(s->g_box.as<float>(), g.win, g.off, n, gpsi, st);
  };
  float* resid = s->d_ratio.as<float>();
  // the views' NLL sums and PSF gradients in one device buffer, copied to the
  // host in one transfer after the last view (a copy to pageable memory waits
  // for the stream)
  std::vector<int64_t> kofs(s->nviews + 1, 0);
  for (int v = 0; v < s->nviews; ++v) {
    const Int3 kd = s->views[v].kd;
    kofs[v + 1] = kofs[v] + (grad_psf && grad_psf[v] ? (int64_t)kd.x * kd.y * kd.z : 0);
  }
  const size_t out_bytes = sizeof(double) * 2 * s->nviews + sizeof(float) * kofs[s->nviews];
  MVD_CHECK(s->g_out.ensure(out_bytes));
  double* nll_sums = s->g_out.as<double>();
  float* kgrads = reinterpret_cast<float*>(nll_sums + 2 * s->nviews);
  // W^ = FFT(E_mirror(psi) on a block window), kept in g_spec_a across the
  // views that share the window (one block per view, same FFT shape): views of
  // one image with kernels of one size convolve the same window
  const ShapePlan* w_plan = nullptr;
  ConvGeom w_geom{};
  auto same_window = [](const ConvGeom& a, const ConvGeom& b) {
    auto eq = [](Int3 u, Int3 w) { return u.x == w.x && u.y == w.y && u.z == w.z; };
    return eq(a.src, b.src) && eq(a.off, b.off) && eq(a.win, b.win) && eq(a.q, b.q);
  };
  for (int v = 0; v < s->nviews; ++v) {
    auto& view = s->views[v];
    const Int3 k = view.kd;
    std::vector<float> pk = view.psf_raw;
    if (psfs && psfs[v]) pk.assign(psfs[v], psfs[v] + pk.size());
    // the deconvolution's own blocks once prepared, else the single window img + K - 1
    std::vector<ConvGeom> blocks;
    const ShapePlan* spp = nullptr;
    if (s->prepared && view.plan) {
      blocks = view.blocks;
      spp = view.plan;
    } else {
      const Int3 win = {n.x + k.x - 1, n.y + k.y - 1, n.z + k.z - 1};
      const Int3 q = fft_shape_for(win);
      spp = s->plans.get(q);
      MVD_REQUIRE(spp, "no FFT plan for the gradient window (lengths up to 1024): prepare with blocks first");
      MVD_REQUIRE(divide_into_blocks(n, k, win, q, &blocks), "internal: gradient window");
    }
    const ShapePlan& sp = *spp;
    const Int3 q = sp.q;
    const int64_t hs = sp.spectrum_elems();
    int64_t crop = 0, box = (int64_t)q.x * q.y * q.z;
    int nparts = 0;
    for (auto& g : blocks) {
      crop = std::max<int64_t>(crop, (int64_t)g.eff_size.x * g.eff_size.y * g.eff_size.z);
      nparts += conv_inverse_workgroups(sp, g);
    }
    MVD_CHECK(s->g_spec_a.ensure(sizeof(float2) * hs));
    MVD_CHECK(s->g_spec_b.ensure(sizeof(float2) * hs));
    MVD_CHECK(s->g_spec_c.ensure(sizeof(float2) * hs));
    // z_grad: the residual's z pass forms both gradients' spectra in one
    // launch; keep_w: one block, whose window spectrum W^ is kept from pass 1
    const bool z_grad = g_grad_fused && z_grad_supported(sp);
    const bool keep_w = z_grad && g_grad_keep_w && blocks.size() == 1;
    // fuse_r (one block, register x and y passes): the pass-1 x inverse turns
    // the residual rows into their x transform in the window spectrum Cs and
    // sums them; the residual image is not written
    const bool fuse_r = keep_w && fast_size(q.x / 2) && fast_size(q.y) && !xy_fused(sp);
    // the summed correlation spectrum D: its own buffer unless one block's
    // correlation goes through B (keep_w, or the separate launches with one block)
    const bool own_d = z_grad ? !keep_w : blocks.size() > 1;
    if (grad_psf && grad_psf[v] && own_d) MVD_CHECK(s->g_spec_d.ensure(sizeof(float2) * hs));
    MVD_CHECK(s->g_box.ensure(sizeof(float) * box));
    MVD_CHECK(s->g_crop.ensure(sizeof(float) * crop));
    MVD_CHECK(s->g_part_sum.ensure(sizeof(double) * nparts));
    MVD_CHECK(s->g_part_max.ensure(sizeof(float) * nparts));
    // residual sums: partials (of k_sum, or of the fused x inverse) and the sum
    constexpr int kRsumParts = 1024;
    const int rparts = std::max(kRsumParts, nparts);
    MVD_CHECK(s->g_rsum.ensure(sizeof(double) * (rparts + 1)));
    double* rsum = s->g_rsum.as<double>() + rparts;
    // the kernel spectrum: the prepared view's own when the PSF is its kernel1
    const float2* Kh = nullptr;
    if (s->prepared && view.plan == spp && view.k1hat.p && pk == view.k1) {
      Kh = view.k1hat.as<float2>();
    } else {
      int rc = kernel_spectrum(s, sp, pk, k, &s->g_khat);
      if (rc) return rc;
      Kh = s->g_khat.as<float2>();
    }
    float2* A = s->g_spec_a.as<float2>();
    float2* B = s->g_spec_b.as<float2>();
    float2* Cs = s->g_spec_c.as<float2>();
    // pass 1, every block: b = E_mirror(psi) * P on the effective region, with the
    // residual r = w (1 - y / b) and the NLL partials w (b - y ln b) in the epilogue
    int off = 0;
    for (auto& g : blocks) {
      EpilogueArgs e{};
      e.mode = kEpiResidual;
      e.dst = resid;
      e.img = s->d_views.as<float>() + v * s->n;
      e.weight_mode = wmode == MVD_WEIGHTS_NONE ? kWeightNone : kWeightArray;
      e.weight = wmode == MVD_WEIGHTS_NONE ? nullptr : s->d_weights.as<float>() + v * s->n;
      e.part_sum = s->g_part_sum.as<double>() + off;
      e.part_max = s->g_part_max.as<float>() + off;
      if (keep_w) {
        // W^ in A (transformed once for the views sharing the window), the
        // convolution W^ P^ transformed back along z into B
        if (!(w_plan == spp && same_window(w_geom, g))) {
          MVD_CHECK(launch_xy_forward_passes(sp, g, psrc, kGatherMirror, A, st));
          MVD_CHECK(launch_z(sp, A, nullptr, 1.f, 0, q.z, st, -1.f));
          w_plan = spp;
          w_geom = g;
        }
        const int zlo = g.eff_local.z, zhi = g.eff_local.z + g.eff_size.z;
        MVD_CHECK(launch_z(sp, B, Kh, 1.f, zlo, zhi, st, -1.f, nullptr, A));
        if (fuse_r) {
          e.rspec = s->g_spec_c.as<float2>();
          e.rpitch = sp.pitch;
          e.part_rsum = s->g_rsum.as<double>();
        }
        int nwg = 0;
        MVD_CHECK(launch_yx_inverse_passes(sp, g, B, e, st, &nwg));
        if (fuse_r) {
          MVD_REQUIRE(nwg <= rparts, "internal: residual partials");
          MVD_CHECK(launch_reduce_sum(s->g_rsum.as<double>(), nwg, rsum, st));
        }
      } else {
        w_plan = nullptr;  // A is overwritten
        int rc = run_conv(s, sp, g, psrc, kGatherMirror, Kh, e, A);
        if (rc) return rc;
      }
      off += conv_inverse_workgroups(sp, g);
    }
    MVD_CHECK(launch_reduce_stats(s->g_part_sum.as<double>(), s->g_part_max.as<float>(), off,
                                  nll_sums + 2 * v, st));
    const bool want_psf = grad_psf && grad_psf[v];
    if (!want_grad_psi && !want_psf) continue;
    float* kgrad = kgrads + kofs[v];
    const float scale = (float)(1.0L / ((long double)q.x * q.y * q.z));
    // pass 2, every block, with the window spectrum recomputed instead of kept
    // (checkpointing: one block's spectra in memory at a time)
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
      const ConvGeom& g = blocks[bi];
      if (!keep_w) {
        // W^ = FFT(E_mirror(psi) on the block window)
        MVD_CHECK(launch_xy_forward_passes(sp, g, psrc, kGatherMirror, A, st));
        MVD_CHECK(launch_z(sp, A, nullptr, 1.f, 0, q.z, st, -1.f));
        w_plan = nullptr;
      }
      if (fuse_r) {
        // the residual rows are x-transformed already; rows outside the
        // effective box are read as zero by the y pass
        const Int3 lo = g.eff_local;
        const Int3 hi = {lo.x + g.eff_size.x, lo.y + g.eff_size.y, lo.z + g.eff_size.z};
        MVD_CHECK(launch_y_fast_window(sp, Cs, 0, q.z, 0, q.y, -1.f, lo, hi, st));
      } else {
        // R^ = FFT(residual of this block's effective region, zero elsewhere in
        // the window); the residual image itself when that region is the image
        const bool whole = g.eff_offset.x == 0 && g.eff_offset.y == 0 && g.eff_offset.z == 0 &&
                           g.eff_size.x == n.x && g.eff_size.y == n.y && g.eff_size.z == n.z;
        const float* rsrc = resid;
        if (!whole) {
          MVD_CHECK(launch_crop(s->g_crop.as<float>(), resid, n, g.eff_offset, g.eff_size, st));
          rsrc = s->g_crop.as<float>();
        }
        ConvGeom gc = g;
        gc.src = g.eff_size;
        gc.off = {-g.eff_local.x, -g.eff_local.y, -g.eff_local.z};
        MVD_CHECK(launch_xy_forward_passes(sp, gc, rsrc, kGatherZero, Cs, st));
        // R^(0) = the residual's sum, in double
        MVD_CHECK(launch_sum_to(rsrc, (int64_t)g.eff_size.x * g.eff_size.y * g.eff_size.z, s->g_rsum.as<double>(),
                                kRsumParts, rsum, st));
      }
      // the adjoint's output box: the block window
      ConvGeom gw{};
      gw.src = g.win;
      gw.win = g.win;
      gw.q = q;
      gw.eff_size = g.win;
      gw.dst = g.win;
      const bool last = bi + 1 == blocks.size();
      if (z_grad) {
        // dL/dP(j) = sum over blocks of corr(R, W)(c - j): the lag-shifted
        // correlation spectra are summed over the blocks (mode 0) and inverted
        // once after the last, or, for one block, inverted along z in the same
        // pass (mode 1); only the kernel-sized corner [0, kd) is transformed
        // further.  The adjoint R^ conj(P^) is inverted along z into Cs.
        float2* D = own_d ? s->g_spec_d.as<float2>() : B;
        MVD_REQUIRE(!want_psf || !own_d || s->g_spec_d.bytes >= sizeof(float2) * hs, "internal: correlation spectrum");
        const Int3 shift = {(k.x - 1) / 2, (k.y - 1) / 2, (k.z - 1) / 2};
        const bool adj_sum = sum_adj && keep_w;
        MVD_REQUIRE(!sum_adj || keep_w, "internal: summed adjoint without a kept window spectrum");
        MVD_CHECK(launch_z_grad(sp, Cs, A, Kh, want_psf ? D : nullptr, keep_w ? 1 : 0, bi == 0, shift, 0,
                                want_grad_psi ? g.win.z : 0, keep_w && want_psf ? k.z : 0, rsum,
                                adj_sum ? s->g_adj.as<float2>() : nullptr, adj_sum && v > 0, st));
        if (want_psf && last) {
          if (!keep_w) MVD_CHECK(launch_z(sp, D, nullptr, 1.f, 0, q.z, st, 1.f));
          ConvGeom gk{};
          gk.src = q;
          gk.win = q;
          gk.q = q;
          gk.eff_size = k;
          gk.dst = k;
          EpilogueArgs es{};
          es.mode = kEpiStore;
          es.dst = s->g_box.as<float>();
          MVD_CHECK(launch_yx_inverse_passes(sp, gk, D, es, st, nullptr));
          MVD_CHECK(launch_extract_psf_grad_shifted(s->g_box.as<float>(), k, scale, kgrad, st));
        }
        if (want_grad_psi && adj_sum) {
          adj_g = gw;
          adj_g.off = g.off;
        } else if (want_grad_psi) {
          EpilogueArgs es{};
          es.mode = kEpiStore;
          es.dst = s->g_box.as<float>();
          MVD_CHECK(launch_yx_inverse_passes(sp, gw, Cs, es, st, nullptr));
          MVD_CHECK(fold(v, g));
        }
        continue;
      }
      MVD_CHECK(launch_z(sp, Cs, nullptr, 1.f, 0, q.z, st, -1.f));
      MVD_CHECK(launch_set_dc(Cs, rsum, st));
      if (want_psf) {
        // dL/dP(j) = sum over blocks of corr(R, W)(c - j) = IFFT(sum conj(R^) W^)(c - j) / (qx qy qz):
        // the products are summed here and inverted once after the last block
        float2* D = own_d ? s->g_spec_d.as<float2>() : B;
        MVD_REQUIRE(!own_d || s->g_spec_d.bytes >= sizeof(float2) * hs, "internal: correlation spectrum");
        MVD_CHECK(launch_spec_corr_acc(D, Cs, A, bi == 0, hs, st));
        if (last) {
          MVD_CHECK(launch_z(sp, D, nullptr, 1.f, 0, q.z, st, 1.f));
          ConvGeom gb{};
          gb.src = q;
          gb.win = q;
          gb.q = q;
          gb.eff_size = q;
          gb.dst = q;
          EpilogueArgs es{};
          es.mode = kEpiStore;
          es.dst = s->g_box.as<float>();
          MVD_CHECK(launch_yx_inverse_passes(sp, gb, D, es, st, nullptr));
          MVD_CHECK(launch_extract_psf_grad(s->g_box.as<float>(), q, k, scale, kgrad, 0, st));
        }
      }
      if (want_grad_psi) {
        // adjoint on the block window: IFFT(R^ conj(P^)), folded back through the mirror extension
        MVD_CHECK(launch_spec_product(B, Cs, Kh, 0, 1, hs, st));
        MVD_CHECK(launch_z(sp, B, nullptr, 1.f, 0, q.z, st, 1.f));
        EpilogueArgs es{};
        es.mode = kEpiStore;
        es.dst = s->g_box.as<float>();
        MVD_CHECK(launch_yx_inverse_passes(sp, gw, B, es, st, nullptr));
        MVD_CHECK(fold(v, g));
      }
    }
  }
  if (sum_adj) {
    // the summed adjoint: y and x back over the window, then one fold
    EpilogueArgs es{};
    es.mode = kEpiStore;
    es.dst = s->g_box.as<float>();
    MVD_CHECK(launch_yx_inverse_passes(*first_plan, adj_g, s->g_adj.as<float2>(), es, st, nullptr));
    MVD_CHECK(fold(0, adj_g));
  }
  std::vector<double> host((out_bytes + sizeof(double) - 1) / sizeof(double));
  MVD_CHECK(hipMemcpyAsync(host.data(), s->g_out.p, out_bytes, hipMemcpyDeviceToHost, st));
  MVD_CHECK(hipStreamSynchronize(st));
  const float* hk = reinterpret_cast<const float*>(host.data() + 2 * s->nviews);
  for (int v = 0; v < s->nviews; ++v)
    if (kofs[v + 1] > kofs[v]) std::memcpy(grad_psf[v], hk + kofs[v], sizeof(float) * (kofs[v + 1] - kofs[v]));
  double total = 0.0;
  for (int v = 0; v < s->nviews; ++v) total += host[2 * v];
  if (nll) *nll = total;
  s->g_gpsi_ready = want_grad_psi != 0;
  return 0;
}

int mvd_get_grad_psi(mvd_session* s, float* out) {
  MVD_REQUIRE(s && out, "NULL argument");
  MVD_REQUIRE(s->g_gpsi_ready, "no estimate gradient: call mvd_nll_grad_device with want_grad_psi first");
  DeviceGuard guard(s->device);
  MVD_CHECK(staged_d2h(out, s->g_gpsi.p, sizeof(float) * s->n, s->stream));
  MVD_CHECK(hipStreamSynchronize(s->stream));
  return 0;
}

int mvd_nll_grad(mvd_session* s, const float* psi, const float* const* psfs, double* nll, float* grad_psi,
                 float* const* grad_psf) {
  int rc = mvd_nll_grad_device(s, psi, psfs, nll, grad_psi != nullptr, grad_psf);
  if (rc || !grad_psi) return rc;
  return mvd_get_grad_psi(s, grad_psi);
}

int mvd_get_kernels(mvd_session* s, int view, float* kernel1, float* kernel2) {
  MVD_REQUIRE(s, "NULL session");
  MVD_REQUIRE(s->prepared, "mvd_prepare first");
  MVD_REQUIRE(view >= 0 && view < s->nviews, "invalid view");
  auto& v = s->views[view];
  if (kernel1) std::copy(v.k1.begin(), v.k1.end(), kernel1);
  if (kernel2) std::copy(v.k2.begin(), v.k2.end(), kernel2);
  return 0;
}

int mvd_get_geometry(mvd_session* s, int view, int* out) {
  MVD_REQUIRE(s && out, "NULL argument");
  MVD_REQUIRE(s->prepared, "mvd_prepare first");
  MVD_REQUIRE(view >= 0 && view < s->nviews, "invalid view");
  auto& v = s->views[view];
  out[0] = (int)v.blocks.size();
  out[1] = v.plan->q.x;
  out[2] = v.plan->q.y;
  out[3] = v.plan->q.z;
  return 0;
}

int mvd_set_timing(mvd_session* s, int on) {
  MVD_REQUIRE(s, "NULL session");
  s->timer.enable(on != 0);
  return 0;
}

int mvd_get_timing(mvd_session* s, double* total_ms, long long* launches, int cap) {
  MVD_REQUIRE(s, "NULL session");
  for (int i = 0; i < kTNumClasses && i < cap; ++i) {
    if (total_ms) total_ms[i] = s->timer.ms(i);
    if (launches) launches[i] = s->timer.count(i);
  }
  return kTNumClasses;
}

const char* mvd_timing_name(int cls) {
  if (cls < 0 || cls >= kTNumClasses) return "";
  return kTimingNames[cls];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched iteration over many sessions (stacks): in view slot v every stack
// with more than v views runs convolve1 + quotient, then convolve2 + update,
// for all of its blocks; each pass is one launch per kernel class for all
// (stack, block) items, and a whole iteration can be replayed as a hipGraph.

struct mvd_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<mvd_session*> stacks;
  std::vector<std::vector<DeviceBuf>> spec;  // [stack][block]
  std::vector<PassLaunch> launches;
  DeviceBuf stats_iter, stats_hist;
  std::vector<int> stats_off;  // per stack, in doubles
  int total_stats = 0;
  bool prepared = false, use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~mvd_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int enqueue_batch_iteration(mvd_batch* b) {
  for (auto& L : b->launches) MVD_CHECK(enqueue_pass_launch(L, b->stream));
  return 0;
}

}  // namespace

extern "C" {

int mvd_batch_create(int device, mvd_batch** out) {
  MVD_REQUIRE(out, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  MVD_CHECK(hipGetDeviceCount(&ndev));
  MVD_REQUIRE(device >= 0 && device < ndev, "invalid device id");
  DeviceGuard guard(device);
  auto b = std::make_unique<mvd_batch>();
  b->device = device;
  MVD_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  *out = b.release();
  return 0;
}

int mvd_batch_destroy(mvd_batch* b) {
  if (!b) return 0;
  DeviceGuard guard(b->device);
  (void)hipStreamSynchronize(b->stream);
  delete b;
  return 0;
}

int mvd_batch_add(mvd_batch* b, mvd_session* s) {
  MVD_REQUIRE(b && s, "NULL argument");
  MVD_REQUIRE(s->device == b->device, "session is on another device");
  MVD_REQUIRE(s->prepared, "mvd_prepare the session first");
  for (auto* x : b->stacks) MVD_REQUIRE(x != s, "session already in the batch");
  b->stacks.push_back(s);
  b->prepared = false;
  return 0;
}

int mvd_batch_prepare(mvd_batch* b, int use_graph) {
  MVD_REQUIRE(b, "NULL batch");
  MVD_REQUIRE(!b->stacks.empty(), "no sessions in the batch");
  DeviceGuard guard(b->device);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->launches.clear();
  b->spec.clear();
  b->stats_off.clear();
  int maxv = 0, tot = 0;
  for (auto* s : b->stacks) {
    MVD_REQUIRE(s->prepared, "a session is not prepared");
    b->stats_off.push_back(tot);
    tot += 2 * s->nviews;
    maxv = std::max(maxv, s->nviews);
    size_t nb = 0;
    for (auto& v : s->views) nb = std::max(nb, v.blocks.size());
    b->spec.emplace_back();
    for (size_t i = 0; i < nb; ++i) {
      b->spec.back().emplace_back();
      size_t elems = 0;
      for (auto& v : s->views)
        if (i < v.blocks.size()) elems = std::max<size_t>(elems, v.plan->spectrum_elems());
      MVD_CHECK(b->spec.back().back().ensure(sizeof(float2) * elems));
      // finite row padding (before any capture)
      MVD_CHECK(hipMemsetAsync(b->spec.back().back().p, 0, sizeof(float2) * elems, b->stream));
    }
  }
  b->total_stats = tot;
  MVD_CHECK(b->stats_iter.ensure(sizeof(double) * tot));
  for (int v = 0; v < maxv; ++v) {
    for (int conv = 0; conv < 2; ++conv) {
      std::vector<PassDesc> pass[5];
      for (size_t si = 0; si < b->stacks.size(); ++si) {
        mvd_session* s = b->stacks[si];
        if (v >= s->nviews) continue;
        std::vector<float2*> spec;
        for (auto& buf : b->spec[si]) spec.push_back(buf.as<float2>());
        conv_pass_descs(s, v, conv, spec, pass);
      }
      for (int p = 0; p < 5; ++p) {
        int rc = build_pass_launches(pass[p], &b->launches);
        if (rc) return rc;
      }
    }
    // statistics of view slot v
    std::vector<ReduceItem> red;
    for (size_t si = 0; si < b->stacks.size(); ++si) {
      mvd_session* s = b->stacks[si];
      if (v >= s->nviews) continue;
      int n = 0;
      for (auto& g : s->views[v].blocks) n += conv_inverse_workgroups(*s->views[v].plan, g);
      red.push_back({s->d_part_sum.as<double>(), s->d_part_max.as<float>(), n,
                     b->stats_iter.as<double>() + b->stats_off[si] + 2 * v});
    }
    PassLaunch L;
    L.reduce = true;
    L.timing = kTReduce;
    L.nitems = (int)red.size();
    MVD_CHECK(L.items.ensure(sizeof(ReduceItem) * red.size()));
    MVD_CHECK(hipMemcpy(L.items.p, red.data(), sizeof(ReduceItem) * red.size(), hipMemcpyHostToDevice));
    b->launches.push_back(std::move(L));
  }
  b->use_graph = use_graph != 0;
  if (b->use_graph) {
    MVD_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_batch_iteration(b);
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    MVD_CHECK(ec);
    b->graph = g;
    MVD_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  b->prepared = true;
  return 0;
}

int mvd_batch_iterate(mvd_batch* b, int n, double* stats) {
  MVD_REQUIRE(b, "NULL batch");
  MVD_REQUIRE(b->prepared, "mvd_batch_prepare first");
  MVD_REQUIRE(n >= 0, "negative iteration count");
  for (auto* s : b->stacks) MVD_REQUIRE(s->psi_ready, "mvd_init_psi every session first");
  if (n == 0) return 0;
  DeviceGuard guard(b->device);
  MVD_CHECK(b->stats_hist.ensure(sizeof(double) * (size_t)b->total_stats * n));
  for (int it = 0; it < n; ++it) {
    if (b->use_graph) {
      MVD_CHECK(hipGraphLaunch(b->exec, b->stream));
    } else {
      int rc = enqueue_batch_iteration(b);
      if (rc) return rc;
    }
    MVD_CHECK(hipMemcpyAsync(b->stats_hist.as<double>() + (size_t)it * b->total_stats, b->stats_iter.p,
                             sizeof(double) * b->total_stats, hipMemcpyDeviceToDevice, b->stream));
  }
  if (stats)
    MVD_CHECK(hipMemcpyAsync(stats, b->stats_hist.p, sizeof(double) * (size_t)b->total_stats * n,
                             hipMemcpyDeviceToHost, b->stream));
  MVD_CHECK(hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"
