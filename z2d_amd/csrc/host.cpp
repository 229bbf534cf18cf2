// host.cpp - host runtime behind include/z2d_gpu.h.
//
// Mirrors the reference's painter entry points (src/painter.zig:66-344) and
// SurfaceCompositor.run (src/compositor.zig:302-440): argument checks on the
// host in the reference's order, curve flattening with the reference's
// Spline algorithm, then one device pipeline per batch (render.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/z2d_gpu.h"
#include "png.h"
#include "render.h"
#include "stroke.h"

using namespace z2d;

// launchers (render.hip)
hipError_t z2d_launch_edge_build(const JobView &J, hipStream_t st);
hipError_t z2d_launch_rowprog(const JobView &J, const uint32_t *list, uint32_t n,
                              z2d_edge *sedges, int32_t *sx, const uint32_t *soff,
                              hipStream_t st);
hipError_t z2d_launch_binning(const JobView &J, hipStream_t st);
hipError_t z2d_launch_render(const JobView &J, uint32_t tile_base, uint32_t tile_count,
                             uint32_t *cursor, uint32_t *cursor_fast, uint32_t n_blocks,
                             uint32_t n_fast_blocks, hipStream_t st);
hipError_t z2d_launch_composite_quad(const z2d::QuadJob &job, uint64_t n_quads, hipStream_t st);
hipError_t z2d_launch_flatten(const z2d_path_node *raw, const z2d::FlatJob *jobs, uint32_t n_jobs,
                              z2d::FNode *out, z2d::FlatResult *res, uint32_t *rel,
                              const uint2 *curves, uint32_t n_curves, z2d::FlatTask *tasks,
                              uint32_t *toff, z2d::FlatCurve *cv, bool emit, hipStream_t st,
                              const z2d::FillSet *fill, bool any_serial);
hipError_t z2d_launch_composite(const z2d::CompJob *dev_job, const z2d::CompJob &host_job,
                                bool stream, hipStream_t st);
hipError_t z2d_launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st);
hipError_t z2d_launch_fill_set(const FillSet &F, hipStream_t st);
hipError_t z2d_launch_multi_copy(const z2d::CopyCmd *cmds, uint32_t n, hipStream_t st);
hipError_t z2d_launch_stream_copy(const void *src, void *dst, uint64_t bytes, hipStream_t st);
hipError_t z2d_launch_stream_read2(const void *a, const void *b, void *dst, uint64_t bytes,
                                   hipStream_t st);
hipError_t z2d_launch_grad_setup(DevGradient *grads, uint32_t n_grads, hipStream_t st);
hipError_t z2d_launch_png(const z2d::PngJob &J, hipStream_t st);  // png.hip

static const uint16_t kBlueNoise[4096] = {
#include "blue_noise_64x64.inc"
};

// ---------------------------------------------------------------------------
// Fixed device scratch of z2d_gpu_composite_noalloc: the operation list
// (CompJob), its paint / gradient / stop tables and the blue-noise matrix of
// one call.  Each device holds its own copy for the life of the process; the
// host reaches it only through the address hipGetSymbolAddress returns on
// that device (the array's name in host code is a stand-in, not a device
// address).
// ---------------------------------------------------------------------------
constexpr size_t kCompScratchBytes = (size_t)1 << 20;
static __device__ __align__(256) uint8_t z2d_comp_scratch[kCompScratchBytes];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint8_t *g_comp_scratch[kMaxDevices];  // per device, resolved once

// The scratch's device address on `device` (current), resolved at the first
// context created there.  Null when the lookup failed.
static uint8_t *comp_scratch_for(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_comp_scratch[device]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(z2d_comp_scratch)) == hipSuccess)
      g_comp_scratch[device] = static_cast<uint8_t *>(ptr);
    else
      (void)hipGetLastError();
  }
  return g_comp_scratch[device];
}

// ---------------------------------------------------------------------------
// device memory helpers
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t n) { return (n + kAlign - 1) & ~(kAlign - 1); }

// Device and pinned host buffers that only grow.  A regrowth frees the old
// allocation (hipFree waits for the device), so past the first size they
// grow by half again at least: a workload whose sizes creep up (a server's
// batches) reallocates a logarithmic number of times, a fixed one once.
inline size_t grown_size(size_t n, size_t cap) {
  return std::max(std::max(n, (size_t)1 << 20), cap ? cap + cap / 2 : (size_t)0);
}

// The buffer owns its allocation: freed by release() or the destructor, so
// a holder that is destroyed (a batch, a server, a per-call batch on the
// stack) never leaks it.  Moves transfer ownership; copies are not allowed.
// A growth that fails at the grown size is retried at the exact size (the
// old allocation is already gone, so near the memory limit the exact size
// may still fit).
template <class Alloc>
struct OwnedBuf {
  uint8_t *p = nullptr;
  size_t cap = 0;
  OwnedBuf() = default;
  OwnedBuf(const OwnedBuf &) = delete;
  OwnedBuf &operator=(const OwnedBuf &) = delete;
  OwnedBuf(OwnedBuf &&o) noexcept : p(o.p), cap(o.cap) {
    o.p = nullptr;
    o.cap = 0;
  }
  OwnedBuf &operator=(OwnedBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      cap = o.cap;
      o.p = nullptr;
      o.cap = 0;
    }
    return *this;
  }
  ~OwnedBuf() { release(); }
  // exact: allocate n itself (small fixed-size buffers), not a grown size
  hipError_t ensure(size_t n, bool exact = false) {
    if (n <= cap) return hipSuccess;
    const size_t want = exact ? n : grown_size(n, cap);
    release();
    hipError_t e = Alloc::alloc(&p, want);
    if (e == hipSuccess) {
      cap = want;
      return hipSuccess;
    }
    p = nullptr;
    if (want == n) return e;
    (void)hipGetLastError();  // the failed attempt's error is superseded
    e = Alloc::alloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    cap = n;
    return hipSuccess;
  }
  void release() {
    if (p) Alloc::free(p);
    p = nullptr;
    cap = 0;
  }
};
struct DeviceAlloc {
  static hipError_t alloc(uint8_t **p, size_t n) { return hipMalloc(p, n); }
  static void free(uint8_t *p) { (void)hipFree(p); }
};
struct PinnedAlloc {
  static hipError_t alloc(uint8_t **p, size_t n) {
    return hipHostMalloc(reinterpret_cast<void **>(p), n, hipHostMallocDefault);
  }
  static void free(uint8_t *p) { (void)hipHostFree(p); }
};
using DevBuf = OwnedBuf<DeviceAlloc>;
using HostBuf = OwnedBuf<PinnedAlloc>;

// Sub-allocation plan inside one arena.
struct Layout {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = off;
    off += align_up(bytes ? bytes : 1);
    return o;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// context / surfaces
// ---------------------------------------------------------------------------
struct z2d_gpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint16_t *blue_noise = nullptr;
  std::string err;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  DevBuf comp_arena;
  HostBuf comp_stage;
  int n_cus = 256;
  float last_comp_ms = 0.0f;  // device time of the last composite kernel
  DevBuf png_work;            // scanline stream, segment staging, offsets
  DevBuf png_out;             // the encoded file
  HostBuf png_meta;
  hipEvent_t png_ev[2] = {nullptr, nullptr};
  float last_png_ms = 0.0f;   // device time of the last PNG encode
  uint8_t *comp_scratch = nullptr;  // this device's fixed composite scratch
};

struct z2d_gpu_surface {
  z2d_gpu_ctx *ctx = nullptr;
  uint32_t format = 0;
  int32_t w = 0, h = 0;
  uint8_t *data = nullptr;
  uint64_t nbytes = 0;
  bool owned = true;
  const uint8_t *clip = nullptr;  // alpha8 clip applied after a batch's draws
};

static int32_t fail(z2d_gpu_ctx *ctx, hipError_t e, const char *what) {
  if (ctx) {
    ctx->err = std::string(what) + ": " + hipGetErrorString(e);
  }
  return Z2D_ERR_DEVICE;
}
#define HIP_OK(ctx, expr)                                 \
  do {                                                    \
    hipError_t _e = (expr);                               \
    if (_e != hipSuccess) return fail((ctx), _e, #expr);  \
  } while (0)

static inline int fbits(uint32_t f) { return fmt_bits(f); }
static uint64_t surface_bytes(uint32_t fmt, int32_t w, int32_t h) {
  uint64_t n = (uint64_t)w * (uint64_t)h;
  int b = fbits(fmt);
  return b == 32 ? n * 4 : (n * (uint64_t)b + 7) / 8;
}
// Device allocations are padded to whole 32-bit words (+4) so packed-alpha
// atomics on the last word stay inside the allocation.
static uint64_t padded_bytes(uint64_t n) { return ((n + 3) & ~3ull) + 4; }

// ---------------------------------------------------------------------------
// patterns and gradients (host side of DevPaint / DevGradient)
// ---------------------------------------------------------------------------
namespace {

inline float clampf_(float v, float lo, float hi) { return fmaxf(lo, fminf(v, hi)); }
inline float zmodf_h(float x, float y) {
  float a = fmodf(x, y);
  return x < 0.0f ? fmodf(a + y, y) : a;
}

// Color.init clamping (src/color.zig:81-92, :216-223, :416-423).
DevStop stop_from(const z2d_stop &s) {
  DevStop o;
  o.offset = clampf_(s.offset, 0.0f, 1.0f);
  o.space = s.color.space;
  if (s.color.space == Z2D_COLOR_HSL) {
    float h = s.color.v[0];
    o.v[0] = (h < 0.0f || h > 360.0f) ? zmodf_h(h, 360.0f) : h;
    for (int i = 1; i < 4; i++) o.v[i] = clampf_(s.color.v[i], 0.0f, 1.0f);
  } else {
    for (int i = 0; i < 4; i++) o.v[i] = clampf_(s.color.v[i], 0.0f, 1.0f);
  }
  return o;
}

// Host copy of zpowf (same algorithm as the device's).
float zpowf_h(float x, float y) {
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (std::isnan(x) || std::isnan(y)) return NAN;
  if (y == 1.0f) return x;
  if (x == 0.0f) {
    float ip;
    bool odd = std::modf(y, &ip) == 0.0f && std::fmod(std::fabs(ip), 2.0f) == 1.0f;
    if (y < 0.0f) return odd ? std::copysign(INFINITY, x) : INFINITY;
    return odd ? x : 0.0f;
  }
  if (std::isinf(y) || std::isinf(x)) return std::pow(x, y);
  if (y == 0.5f) return std::sqrt(x);
  if (y == -0.5f) return 1.0f / std::sqrt(x);
  float yi;
  float yf = std::modf(std::fabs(y), &yi);
  if (yf != 0.0f && x < 0.0f) return NAN;
  float a1 = 1.0f;
  int ae = 0;
  if (yf != 0.0f) {
    if (yf > 0.5f) {
      yf -= 1.0f;
      yi += 1.0f;
    }
    a1 = expf(yf * logf(x));
  }
  int xe;
  float x1 = frexpf(x, &xe);
  int i = (int)yi;
  while (i != 0) {
    if (xe < -(1 << 9) || (1 << 9) < xe) {
      ae += xe;
      break;
    }
    if (i & 1) {
      a1 *= x1;
      ae += xe;
    }
    x1 *= x1;
    xe <<= 1;
    if (x1 < 0.5f) {
      x1 += x1;
      xe -= 1;
    }
    i >>= 1;
  }
  if (y < 0.0f) {
    a1 = 1.0f / a1;
    ae = -ae;
  }
  return scalbnf(a1, ae);
}

void hsl_to_rgb_h(const float *h, float *o) {
  float hue = fmodf(h[0], 360.0f);
  if (hue < 0.0f) hue += 360.0f;
  const float ns[3] = {0.0f, 8.0f, 4.0f};
  for (int i = 0; i < 3; i++) {
    float k = fmodf(ns[i] + hue / 30.0f, 12.0f);
    float a = h[1] * fminf(h[2], 1.0f - h[2]);
    o[i] = h[2] - a * fmaxf(-1.0f, fminf(fminf(k - 3.0f, 9.0f - k), 1.0f));
  }
  o[3] = h[3];
}

// LinearRGB.fromColor (src/color.zig:226-248) of an initialized color.
void color_linear_h(const DevStop &s, float *o) {
  if (s.space == Z2D_COLOR_HSL) {
    hsl_to_rgb_h(s.v, o);
    return;
  }
  if (s.space == Z2D_COLOR_LINEAR_RGB) {
    for (int i = 0; i < 4; i++) o[i] = s.v[i];
    return;
  }
  for (int i = 0; i < 3; i++) o[i] = zpowf_h(zpowf_h(s.v[i], 2.2f), 1.0f);
  o[3] = s.v[3];
}

I4 pixel_rgba(const z2d_pixel &p) {
  I4 o{0, 0, 0, 0};
  switch (p.format) {
    case Z2D_FORMAT_RGBA:
    case Z2D_FORMAT_ARGB: o = I4{p.r, p.g, p.b, p.a}; break;
    case Z2D_FORMAT_RGB:
    case Z2D_FORMAT_XRGB: o = I4{p.r, p.g, p.b, 255}; break;
    case Z2D_FORMAT_ALPHA8: o.a = p.a; break;
    case Z2D_FORMAT_ALPHA4: o.a = (uint32_t)p.a * 17u; break;
    case Z2D_FORMAT_ALPHA2: o.a = (uint32_t)p.a * 85u; break;
    default: o.a = (uint32_t)p.a * 255u; break;
  }
  return o;
}
bool pixel_opaque(const z2d_pixel &p) {
  switch (p.format) {
    case Z2D_FORMAT_XRGB:
    case Z2D_FORMAT_RGB: return true;
    case Z2D_FORMAT_ARGB:
    case Z2D_FORMAT_RGBA:
    case Z2D_FORMAT_ALPHA8: return p.a == 255;
    case Z2D_FORMAT_ALPHA4: return p.a == 15;
    case Z2D_FORMAT_ALPHA2: return p.a == 3;
    default: return p.a == 1;
  }
}
// RGBA.canDemultiply (src/pixel.zig:504-517).
bool can_demultiply(const z2d_pixel &p) {
  if (p.a == 0) return true;
  return (uint32_t)p.r * 255u / p.a <= 255u && (uint32_t)p.g * 255u / p.a <= 255u &&
         (uint32_t)p.b * 255u / p.a <= 255u;
}

// A gradient's device record (its stop pointers left for the caller) and
// its stops in Stop.List order, appended to `stops`.
static void build_gradient(const z2d_gradient_desc *g, DevGradient &d,
                           std::vector<DevStop> &stops) {
  std::memset(&d, 0, sizeof(d));
  d.type = g->type;
  d.method = g->method;
  d.polar = g->polar;
  d.n_stops = g->n_stops;
  for (int i = 0; i < 6; i++) {
    d.geom[i] = g->geom[i];
    d.m[i] = g->matrix[i];
  }
  d.identity = (g->matrix[0] == 1.0 && g->matrix[1] == 0.0 && g->matrix[2] == 0.0 &&
                g->matrix[3] == 1.0 && g->matrix[4] == 0.0 && g->matrix[5] == 0.0);
  if (g->type == Z2D_GRADIENT_RADIAL) {
    // Radial.initBuffer (src/gradient.zig:440-470)
    d.r1 = std::max(0.0, g->geom[2]);
    d.r2 = std::max(0.0, g->geom[5]);
    d.cdx = g->geom[3] - g->geom[0];
    d.cdy = g->geom[4] - g->geom[1];
    d.dr = d.r2 - d.r1;
    d.min_dr = -d.r1;
    double a = 0.0;
    a += d.cdx * d.cdx;
    a += d.cdy * d.cdy;
    a += d.dr * (-d.dr);
    d.a = a;
    d.inv_a = a != 0.0 ? 1.0 / a : 0.0;
  } else if (g->type == Z2D_GRADIENT_LINEAR) {
    // Linear.getOffset's per-gradient terms (src/gradient.zig)
    d.lin_dx = g->geom[2] - g->geom[0];
    d.lin_dy = g->geom[3] - g->geom[1];
    double dist = 0.0;
    dist += d.lin_dx * d.lin_dx;
    dist += d.lin_dy * d.lin_dy;
    d.lin_degenerate = dist == 0.0;
    d.lin_inv = dist != 0.0 ? 1.0 / dist : 0.0;
  } else if (g->type == Z2D_GRADIENT_CONIC) {
    // Conic.initBuffer: @mod(angle, pi*2)
    const double two_pi = 3.14159265358979323846 * 2.0;
    double a = std::fmod(g->geom[2], two_pi);
    d.conic_angle = g->geom[2] < 0.0 ? std::fmod(a + two_pi, two_pi) : a;
  }
  // Stop.List.addAssumeCapacity: clamp, Color.init, stable sort by offset.
  uint32_t base = (uint32_t)stops.size();
  for (uint32_t i = 0; i < g->n_stops; i++) {
    DevStop s = stop_from(g->stops[i]);
    size_t j = stops.size();
    stops.push_back(s);
    while (j > base && stops[j - 1].offset > s.offset) {
      stops[j] = stops[j - 1];
      j--;
    }
    stops[j] = s;
  }
}

// Gradients resident in HBM across batches (a server's requests): each
// distinct gradient is packed, uploaded and set up (k_grad_setup) once, on
// the table's own stream, and later draws of an equal gradient (same type,
// geometry, matrix, method and stops, compared byte for byte) use the
// resident record.  Entries are never replaced: once the table is full, or
// for gradients of more than kMaxStops stops, draws use the batch's own
// tables.  A batch that uses the table makes its stream wait for the
// table's inserts before its kernels (`ev`, recorded after every insert).
struct GradCache {
  static constexpr uint32_t kCap = 4096, kMaxStops = 16;
  std::mutex mu;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  DevBuf dev;     // records, then stops, offsets and per-stop colors
  HostBuf stage;  // one staging slot per entry (records, stops, offsets)
  size_t o_stops = 0, o_offs = 0, o_pre = 0, stage_entry = 0;
  std::unordered_multimap<uint64_t, uint32_t> index;
  std::vector<std::vector<uint8_t>> keys;
  uint32_t n = 0;
  uint64_t hits = 0, misses = 0, full = 0;
  bool ok = false;

  hipError_t init() {
    o_stops = kCap * sizeof(DevGradient);
    o_offs = o_stops + (size_t)kCap * kMaxStops * sizeof(DevStop);
    o_pre = (o_offs + (size_t)kCap * kMaxStops * sizeof(float) + 255) & ~(size_t)255;
    const size_t bytes = o_pre + (size_t)kCap * kMaxStops * sizeof(float4);
    stage_entry = (sizeof(DevGradient) + kMaxStops * (sizeof(DevStop) + sizeof(float)) + 63) &
                  ~(size_t)63;
    hipError_t e = dev.ensure(bytes, true);
    if (e == hipSuccess) e = stage.ensure(stage_entry * kCap, true);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    ok = e == hipSuccess;
    return e;
  }
  void destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (ev) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
    ev = nullptr;
    stream = nullptr;
    dev.release();
    stage.release();
    ok = false;
  }
  static void key_of(const z2d_gradient_desc *g, std::vector<uint8_t> &k) {
    k.clear();
    auto put = [&](const void *p, size_t len) {
      const uint8_t *b = static_cast<const uint8_t *>(p);
      k.insert(k.end(), b, b + len);
    };
    put(&g->type, sizeof(g->type));
    put(&g->method, sizeof(g->method));
    put(&g->polar, sizeof(g->polar));
    put(&g->n_stops, sizeof(g->n_stops));
    put(g->geom, sizeof(g->geom));
    put(g->matrix, sizeof(g->matrix));
    for (uint32_t i = 0; i < g->n_stops; i++) {
      put(&g->stops[i].offset, sizeof(float));
      put(&g->stops[i].color.space, sizeof(uint32_t));
      put(g->stops[i].color.v, sizeof(g->stops[i].color.v));
    }
  }
  // The resident record of g, inserted if new; null when g does not fit
  // (the caller packs it into the batch).  `used` is set when a record is
  // returned.
  const DevGradient *get(const z2d_gradient_desc *g, bool &used) {
    if (!ok || g->n_stops > kMaxStops) return nullptr;
    thread_local std::vector<uint8_t> k;  // the key, built without a new allocation
    key_of(g, k);
    uint64_t h = 1469598103934665603ull;
    for (uint8_t c : k) h = (h ^ c) * 1099511628211ull;
    std::lock_guard<std::mutex> lk(mu);
    auto range = index.equal_range(h);
    for (auto it = range.first; it != range.second; ++it)
      if (keys[it->second] == k) {
        hits++;
        used = true;
        return reinterpret_cast<const DevGradient *>(dev.p) + it->second;
      }
    if (n >= kCap) {
      full++;
      return nullptr;
    }
    const uint32_t i = n;
    DevGradient d;
    std::vector<DevStop> st;
    build_gradient(g, d, st);
    DevGradient *dg = reinterpret_cast<DevGradient *>(dev.p) + i;
    DevStop *ds = reinterpret_cast<DevStop *>(dev.p + o_stops) + (size_t)i * kMaxStops;
    float *doff = reinterpret_cast<float *>(dev.p + o_offs) + (size_t)i * kMaxStops;
    float4 *dpre = reinterpret_cast<float4 *>(dev.p + o_pre) + (size_t)i * kMaxStops;
    d.stops = ds;
    d.offs = doff;
    d.pre = dpre;
    uint8_t *slot = stage.p + (size_t)i * stage_entry;
    std::memcpy(slot, &d, sizeof(d));
    DevStop *hs = reinterpret_cast<DevStop *>(slot + sizeof(DevGradient));
    float *hoff = reinterpret_cast<float *>(hs + kMaxStops);
    for (size_t j = 0; j < st.size(); j++) {
      hs[j] = st[j];
      hoff[j] = st[j].offset;
    }
    // the slot is written once and never reused, so the copies may still be
    // in flight when this returns
    hipError_t e = hipMemcpyAsync(dg, slot, sizeof(DevGradient), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && !st.empty())
      e = hipMemcpyAsync(ds, hs, st.size() * sizeof(DevStop), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && !st.empty())
      e = hipMemcpyAsync(doff, hoff, st.size() * sizeof(float), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = z2d_launch_grad_setup(dg, 1, stream);
    if (e == hipSuccess) e = hipEventRecord(ev, stream);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      ok = false;  // no further inserts; draws use their batches' tables
      return nullptr;
    }
    n++;
    misses++;
    keys.push_back(k);
    index.emplace(h, i);
    used = true;
    return dg;
  }
  // The batch stream waits for every insert so far.
  hipError_t wait_on(hipStream_t st) {
    std::lock_guard<std::mutex> lk(mu);
    return hipStreamWaitEvent(st, ev, 0);
  }
};

// Accumulates paints, gradients and stops for one upload.
struct PaintTable {
  std::vector<DevPaint> paints;
  std::vector<DevGradient> grads;
  std::vector<DevStop> stops;
  std::vector<uint32_t> grad_stop_off;  // per gradient
  std::vector<int32_t> paint_grad;      // per paint: gradient index or -1
  std::vector<const DevGradient *> paint_gdev;  // per paint: a resident gradient or null
  GradCache *cache = nullptr;           // resident gradients (a server's batches)
  bool cache_used = false;

  uint32_t add_gradient(const z2d_gradient_desc *g) {
    DevGradient d;
    build_gradient(g, d, stops);
    grads.push_back(d);
    grad_stop_off.push_back((uint32_t)(stops.size() - g->n_stops));
    return (uint32_t)grads.size() - 1;
  }

  uint32_t add(const z2d_pattern_desc *p) {
    DevPaint d;
    std::memset(&d, 0, sizeof(d));
    d.kind = p->kind;
    int32_t gi = -1;
    const DevGradient *gdev = nullptr;
    if (p->kind == Z2D_PATTERN_OPAQUE) {
      d.px_format = p->pixel.format;
      d.px = pixel_rgba(p->pixel);
      d.px_opaque = pixel_opaque(p->pixel);
    } else if (p->kind == Z2D_PATTERN_GRADIENT) {
      gdev = cache ? cache->get(p->gradient, cache_used) : nullptr;
      if (!gdev) gi = (int32_t)add_gradient(p->gradient);
    } else {
      const z2d_dither_desc &dd = p->dither;
      d.dither_type = dd.type;
      d.dither_source = dd.source;
      d.dither_scale = dd.scale;
      if (dd.source == Z2D_DITHER_SRC_GRADIENT) {
        gdev = cache ? cache->get(dd.gradient, cache_used) : nullptr;
        if (!gdev) gi = (int32_t)add_gradient(dd.gradient);
      } else if (dd.source == Z2D_DITHER_SRC_PIXEL) {
        // LinearRGB.decodeRGBA: u16 demultiply, then /255
        I4 v = pixel_rgba(dd.pixel);
        uint32_t r = v.r, g = v.g, b = v.b;
        if (v.a == 0) {
          r = g = b = 0;
        } else {
          r = (r * 255u / v.a) & 0xFFFF;
          g = (g * 255u / v.a) & 0xFFFF;
          b = (b * 255u / v.a) & 0xFFFF;
        }
        d.dither_color = F4{(float)r / 255.0f, (float)g / 255.0f, (float)b / 255.0f,
                            (float)v.a / 255.0f};
      } else {
        z2d_stop tmp;
        tmp.offset = 0.0f;
        tmp.color = dd.color;
        DevStop s = stop_from(tmp);
        float c[4];
        color_linear_h(s, c);
        d.dither_color = F4{c[0], c[1], c[2], c[3]};
      }
    }
    paints.push_back(d);
    paint_grad.push_back(gi);
    paint_gdev.push_back(gdev);
    return (uint32_t)paints.size() - 1;
  }

  void clear() {
    paints.clear();
    grads.clear();
    stops.clear();
    grad_stop_off.clear();
    paint_grad.clear();
    paint_gdev.clear();
    cache_used = false;
  }

  // Writes the three tables at their device addresses.
  // Layout: paints, gradients, stops, dense stop offsets, per-stop colors
  // (the last filled on the device by k_grad_setup).
  struct Offsets {
    size_t paints, grads, stops, offs, pre;
  };
  Offsets take(Layout &L) const {
    Offsets o;
    o.paints = L.take(paints.size() * sizeof(DevPaint));
    o.grads = L.take(grads.size() * sizeof(DevGradient));
    o.stops = L.take(stops.size() * sizeof(DevStop));
    o.offs = L.take(stops.size() * sizeof(float));
    o.pre = L.take(stops.size() * sizeof(float4));
    return o;
  }

  void write(uint8_t *host_base, uint8_t *dev_base, const Offsets &o) {
    DevPaint *hp = reinterpret_cast<DevPaint *>(host_base + o.paints);
    DevGradient *hg = reinterpret_cast<DevGradient *>(host_base + o.grads);
    DevStop *hs = reinterpret_cast<DevStop *>(host_base + o.stops);
    float *hoff = reinterpret_cast<float *>(host_base + o.offs);
    DevGradient *dg = reinterpret_cast<DevGradient *>(dev_base + o.grads);
    DevStop *ds = reinterpret_cast<DevStop *>(dev_base + o.stops);
    float *doff = reinterpret_cast<float *>(dev_base + o.offs);
    float4 *dpre = reinterpret_cast<float4 *>(dev_base + o.pre);
    for (size_t i = 0; i < grads.size(); i++) {
      hg[i] = grads[i];
      hg[i].stops = ds + grad_stop_off[i];
      hg[i].offs = doff + grad_stop_off[i];
      hg[i].pre = dpre + grad_stop_off[i];
    }
    if (!stops.empty()) std::memcpy(hs, stops.data(), stops.size() * sizeof(DevStop));
    for (size_t i = 0; i < stops.size(); i++) hoff[i] = stops[i].offset;
    for (size_t i = 0; i < paints.size(); i++) {
      hp[i] = paints[i];
      hp[i].gradient = paint_grad[i] >= 0 ? dg + paint_grad[i] : paint_gdev[i];
    }
  }
};

// ---------------------------------------------------------------------------
// Flattening: fill_plotter's node state machine with the reference's Spline
// subdivision (src/internal/tess/fill_plotter.zig:21-118,
// src/internal/tess/Spline.zig:37-165).  Curves become the exact line_to
// sequence the spline would plot; the device repeats the state machine to
// build edges.  Also reports whether every contour returns to its start.
// ---------------------------------------------------------------------------
struct Pt {
  double x, y;
};
inline bool peq(Pt a, Pt b) { return a.x == b.x && a.y == b.y; }

// PathNode.isClosedNodeSet (src/internal/path_nodes.zig:23-36).
bool is_closed_node_set(const z2d_path_node *nodes, uint64_t n) {
  if (n == 0) return false;
  bool closed = false;
  for (uint64_t i = 0; i < n; i++) {
    switch (nodes[i].tag) {
      case Z2D_NODE_MOVE_TO:
        if (!closed && i != 0) return closed;
        break;
      case Z2D_NODE_CLOSE_PATH: closed = true; break;
      default: closed = false; break;
    }
  }
  return closed;
}

int32_t check_pattern(const z2d_pattern_desc *p) {
  if (!p) return Z2D_ERR_INVALID_ARGUMENT;
  if (p->kind == Z2D_PATTERN_OPAQUE &&
      (p->pixel.format == Z2D_FORMAT_ARGB || p->pixel.format == Z2D_FORMAT_RGBA) &&
      !can_demultiply(p->pixel))
    return Z2D_ERR_PIXEL_SOURCE_NOT_PREMULTIPLIED;
  if (p->kind == Z2D_PATTERN_GRADIENT && !p->gradient) return Z2D_ERR_INVALID_ARGUMENT;
  if (p->kind == Z2D_PATTERN_DITHER && p->dither.source == Z2D_DITHER_SRC_GRADIENT &&
      !p->dither.gradient)
    return Z2D_ERR_INVALID_ARGUMENT;
  return Z2D_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
struct HostDraw {
  DevDraw d;
  uint32_t surface_slot;  // index into batch surfaces
  uint32_t seq;           // insertion order
  double box[4];          // xmin, xmax, ymin, ymax of the input points (pixels)
  bool bounded;           // bounded operator
  bool whole;             // unbounded operator or non-finite box: every tile
  double edge_scale;      // polygon draws: the scale of their edges' space
  uint32_t raw_off, raw_cnt;  // fill draws: their path nodes in raw_nodes
  double tol_sq;              // fill draws: squared flattening tolerance
  // set when the batch is prepared, for its band height
  uint32_t band_lo, band_hi;  // estimated band range (inclusive)
  uint32_t tcol_lo, tcol_hi;  // estimated tile-column range (inclusive)
  uint64_t edge_pairs;        // upper bound of (edge, band) pairs
  // polygon draws whose edges stay with their owner until the batch is
  // prepared (a server request's stroke polygon): read from there, not
  // copied into the batch
  const z2d_edge *ext_edges;
  // edge_pairs computed by the owner for band height pairs_bh (0: none)
  uint64_t pairs_hint;
  uint32_t pairs_bh;
};

// Bands [floor(lo / band_h), floor(hi / band_h)] that an edge between pixel
// rows a and b can meet, capped at the surface's band count.
static uint64_t band_span(double a, double b, uint32_t nbands, uint32_t band_h) {
  const double lo = std::min(a, b), hi = std::max(a, b);
  if (!std::isfinite(lo) || !std::isfinite(hi)) return nbands;
  const double s = std::floor(hi / band_h) - std::floor(lo / band_h) + 1.0;
  return (uint64_t)std::min(s, (double)nbands);
}

struct z2d_gpu_batch {
  z2d_gpu_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;  // null: the context stream
  std::vector<z2d_gpu_surface *> surfaces;
  std::vector<z2d_path_node> raw_nodes;  // fill draws' path nodes
  std::vector<z2d_edge> in_edges;
  uint64_t n_ext_edges = 0;  // edges of the draws that borrow their edges (HostDraw::ext_edges)
  uint64_t n_nodes = 0;  // flattened nodes and edge slots of the prepared batch
  DevBuf flat_buf;       // device flattening inputs and results
  HostBuf flat_stage;
  uint32_t n_flat_jobs = 0;  // draws flattened on the device at every render
  uint32_t n_flat_serial = 0;  // of which by the per-draw walk
  const FlatJob *d_flat_jobs = nullptr;
  FNode *d_flat_out = nullptr;
  uint32_t *d_flat_rel = nullptr;
  const uint2 *d_flat_curves = nullptr;
  uint32_t n_flat_curves = 0;
  FlatTask *d_flat_tasks = nullptr;
  uint32_t *d_flat_toff = nullptr;
  std::vector<HostDraw> draws;
  PaintTable paints;
  bool prepared = false;
  // device state of the prepared batch
  DevBuf arena;
  HostBuf stage;
  JobView view{};
  uint32_t n_rowprog = 0;
  uint32_t *d_rowprog_list = nullptr;
  uint32_t *d_rowprog_soff = nullptr;
  z2d_edge *d_scratch_edges = nullptr;
  int32_t *d_scratch_x = nullptr;
  std::vector<uint32_t> surf_tile_off;  // per surface slot, + total
  float last_pipeline_ms = 0.0f, last_render_ms = 0.0f;
  uint32_t count_stats = 0;
  uint32_t band_h = TILE_H;  // rows per band (bin height), a multiple of ROWS_PER_THREAD
  uint32_t longest_first = 0;  // strip queue order of the single-launch render
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  bool render_timed = false;  // the last render recorded ev[1] before the tile render
  // The counters of the last render, copied to pinned memory at its end
  // (read after the stream completes: no extra synchronizing copy).
  HostBuf ctr_host;
  // isolate_failures: a draw whose curve exceeds the subdivision limit is
  // dropped (renders nothing) and reported in `failed` as (insertion index,
  // status) instead of failing the whole batch (the server's requests fail
  // one by one).
  bool isolate_failures = false;
  double t_count_wait = 0;  // seconds the last prepare waited for its count pass
  // async_prepare: prepare returns with its uploads still in flight (the
  // caller keeps the batch's staging buffers untouched until the stream
  // passes them: a server slot until its done event)
  bool async_prepare = false;
  std::vector<std::pair<uint32_t, int32_t>> failed;
  // hipGraph replay of the render pipeline (z2d_gpu_batch_use_graph)
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  uint32_t graph_mode = 0, graph_stats = 0;

  uint32_t slot_of(z2d_gpu_surface *s) {
    for (uint32_t i = 0; i < surfaces.size(); i++)
      if (surfaces[i] == s) return i;
    surfaces.push_back(s);
    return (uint32_t)surfaces.size() - 1;
  }
};

static inline hipStream_t bstream(const z2d_gpu_batch *b) {
  return b->stream ? b->stream : b->ctx->stream;
}

static int32_t batch_add_common(z2d_gpu_batch *b, z2d_gpu_surface *s,
                                const z2d_pattern_desc *pattern, uint32_t aa, uint32_t rule,
                                uint32_t op, uint32_t precision, HostDraw &hd,
                                const double box[4]) {
  (void)pattern;
  hd.surface_slot = b->slot_of(s);
  hd.seq = (uint32_t)b->draws.size();
  DevDraw &d = hd.d;
  // Alpha1 targets force AA off (src/painter.zig:88-91).
  uint32_t eff_aa = (s->format == Z2D_FORMAT_ALPHA1) ? Z2D_AA_NONE : aa;
  d.mode = eff_aa == Z2D_AA_NONE ? MODE_DIRECT : eff_aa == Z2D_AA_SUPERSAMPLE_4X ? MODE_SSAA
                                                                             : MODE_MSAA;
  d.rule = rule;
  d.op = op;
  d.precision = op_requires_float(op) ? Z2D_PRECISION_FLOAT : precision;
  d.paint = b->paints.add(pattern);
  const bool bounded = op_is_bounded(op);
  if (d.mode == MODE_DIRECT && !bounded) d.flags |= DRAW_ROWPROG;
  if (d.flags & DRAW_UNBALANCED) d.flags |= DRAW_ROWPROG;
  // The points' box gives the band and tile-column estimates when the batch
  // is prepared; the composite region lies within a pixel of it.
  const bool finite = std::isfinite(box[0]) && std::isfinite(box[1]) &&
                      std::isfinite(box[2]) && std::isfinite(box[3]);
  for (int i = 0; i < 4; i++) hd.box[i] = box[i];
  hd.bounded = bounded;
  hd.whole = !bounded || !finite;
  b->draws.push_back(hd);
  b->prepared = false;
  return Z2D_OK;
}

// painter.fill's checks after the pattern's (src/painter.zig:73-82, in order)
// and the node walk's own errors, which depend on the tags alone: a line or
// curve before the first move is InvalidState (fill_plotter.zig:52-71), an
// unknown tag an invalid argument.  Flattening happens on the device.
static int32_t fill_checks(const z2d_path_node *nodes, uint64_t n, const z2d_fill_opts *o) {
  if (n == 0) return Z2D_OK;
  if (!nodes) return Z2D_ERR_INVALID_ARGUMENT;
  if (!is_closed_node_set(nodes, n)) return Z2D_ERR_PATH_NOT_CLOSED;
  if (o->op >= Z2D_OP_COUNT || o->fill_rule > 1 || o->anti_aliasing_mode > 3)
    return Z2D_ERR_INVALID_ARGUMENT;
  // A non-finite coordinate on a node the walk uses makes the reference
  // panic (Polygon.inBox, src/internal/tess/Polygon.zig:146-178) or recurse
  // without end (Spline.decompose); here it is an invalid argument of this
  // call alone.
  auto finite = [](const z2d_path_node &nd, int k) {
    for (int i = 0; i < k; i++)
      if (!std::isfinite(nd.v[i])) return false;
    return true;
  };
  bool moved = false;
  for (uint64_t i = 0; i < n; i++) {
    switch (nodes[i].tag) {
      case Z2D_NODE_MOVE_TO:
        if (i == n - 1) return Z2D_OK;  // a final move ends the walk
        if (!finite(nodes[i], 2)) return Z2D_ERR_INVALID_ARGUMENT;
        moved = true;
        break;
      case Z2D_NODE_LINE_TO:
        if (!moved) return Z2D_ERR_INVALID_STATE;
        if (!finite(nodes[i], 2)) return Z2D_ERR_INVALID_ARGUMENT;
        break;
      case Z2D_NODE_CURVE_TO:
        if (!moved) return Z2D_ERR_INVALID_STATE;
        if (!finite(nodes[i], 6)) return Z2D_ERR_INVALID_ARGUMENT;
        break;
      case Z2D_NODE_CLOSE_PATH: break;
      default: return Z2D_ERR_INVALID_ARGUMENT;
    }
  }
  return Z2D_OK;
}

// Appends a checked fill: its path nodes go to the device flattener when the
// batch is prepared.
static int32_t batch_add_fill(z2d_gpu_batch *b, z2d_gpu_surface *s,
                              const z2d_pattern_desc *pattern, const z2d_path_node *nodes,
                              uint64_t n, const z2d_fill_opts *o) {
  if (n == 0) return Z2D_OK;
  if (b->raw_nodes.size() + n > 0xFFFFFFF0ull) return Z2D_ERR_OUT_OF_MEMORY;
  HostDraw hd;
  std::memset(&hd, 0, sizeof(hd));
  hd.raw_off = (uint32_t)b->raw_nodes.size();
  hd.raw_cnt = (uint32_t)n;
  const double tol = std::max(o->tolerance, 0.001);
  hd.tol_sq = tol * tol;
  b->raw_nodes.insert(b->raw_nodes.end(), nodes, nodes + n);
  const double box[4] = {NAN, NAN, NAN, NAN};  // from the device flattener
  return batch_add_common(b, s, pattern, o->anti_aliasing_mode, o->fill_rule, o->op,
                          o->precision, hd, box);
}

static int32_t batch_fill(z2d_gpu_batch *b, z2d_gpu_surface *s, const z2d_pattern_desc *pattern,
                          const z2d_path_node *nodes, uint64_t n, const z2d_fill_opts *o) {
  if (!b || !s || !o) return Z2D_ERR_INVALID_ARGUMENT;
  int32_t rc = check_pattern(pattern);
  if (rc) return rc;
  rc = fill_checks(nodes, n, o);
  if (rc) return rc;
  return batch_add_fill(b, s, pattern, nodes, n, o);
}

static int32_t batch_fill_polygon(z2d_gpu_batch *b, z2d_gpu_surface *s,
                                  const z2d_pattern_desc *pattern, const z2d_edge *edges,
                                  uint64_t n_edges, const double ext[4], uint32_t aa,
                                  uint32_t rule, uint32_t op, uint32_t precision,
                                  bool balanced, bool borrow = false, uint64_t pairs_hint = 0,
                                  uint32_t pairs_bh = 0) {
  if (!b || !s || !ext) return Z2D_ERR_INVALID_ARGUMENT;
  int32_t rc = check_pattern(pattern);
  if (rc) return rc;
  if (op >= Z2D_OP_COUNT || rule > 1 || aa > 3) return Z2D_ERR_INVALID_ARGUMENT;
  HostDraw hd;
  std::memset(&hd, 0, sizeof(hd));
  // one edge slot per edge (offsets assigned when the batch is prepared)
  if (n_edges > 0xFFFFFFF0ull) return Z2D_ERR_OUT_OF_MEMORY;
  hd.d.node_cnt = (uint32_t)n_edges;
  hd.d.input_edges = 1;
  if (borrow) {
    hd.ext_edges = edges;  // in_edge_off is given at prepare
    b->n_ext_edges += n_edges;
    hd.pairs_hint = pairs_hint;
    hd.pairs_bh = pairs_bh;
  } else {
    hd.d.in_edge_off = (uint32_t)b->in_edges.size();
    b->in_edges.insert(b->in_edges.end(), edges, edges + n_edges);
  }
  for (int i = 0; i < 4; i++) hd.d.in_ext[i] = ext[i];
  if (!balanced) hd.d.flags |= DRAW_UNBALANCED;
  // extents are in scaled space
  const double sc = (aa == Z2D_AA_NONE || s->format == Z2D_FORMAT_ALPHA1) ? 1.0 : 4.0;
  const double box[4] = {n_edges ? ext[2] / sc : NAN, n_edges ? ext[3] / sc : NAN,
                         n_edges ? ext[0] / sc : NAN, n_edges ? ext[1] / sc : NAN};
  hd.edge_scale = sc;
  return batch_add_common(b, s, pattern, aa, rule, op, precision, hd, box);
}

// Transformation.inverse succeeds (src/Transformation.zig:103-160).
static bool ctm_invertible(const double m[6]) {
  if (m[1] == 0 && m[2] == 0) return m[0] != 0 && m[3] != 0;
  return m[0] * m[3] - m[1] * m[2] != 0;
}

// painter.stroke's option adjustments (src/painter.zig:236-304) and the
// stroke plotter: the polygon in the AA mode's scaled space.
static int32_t stroke_polygon(const z2d_path_node *nodes, uint64_t n, const z2d_stroke_opts *o,
                              uint32_t format, uint32_t &aa, StrokePolygon &poly) {
  aa = format == Z2D_FORMAT_ALPHA1 ? (uint32_t)Z2D_AA_NONE : o->anti_aliasing_mode;
  const double minimum_line_width = 0.00390625;
  const bool wide = o->line_width >= 2;
  StrokeParams p;
  p.cap_mode = wide ? o->line_cap_mode : (uint32_t)Z2D_CAP_BUTT;
  p.join_mode = wide ? o->line_join_mode : (uint32_t)Z2D_JOIN_MITER;
  p.miter_limit = wide ? o->miter_limit : 10.0;
  p.thickness = o->line_width >= minimum_line_width ? o->line_width : minimum_line_width;
  p.tolerance = std::max(o->tolerance, 0.001);
  p.scale = aa == Z2D_AA_NONE ? 1.0 : 4.0;
  for (int i = 0; i < 6; i++) p.ctm[i] = o->transformation[i];
  p.dashes = o->dashes;
  p.n_dashes = o->n_dashes;
  p.dash_offset = o->dash_offset;
  return stroke_tessellate(nodes, n, p, poly);
}

// painter.stroke (src/painter.zig:214-344): host tessellation by the
// stroke plotter, then the polygon is filled non-zero.  Hairline strokes
// (a separate rasterizer in the reference) are not part of this backend.
// The host geometry of one stroke: painter.stroke's checks after the
// pattern's (src/painter.zig:221-236, in order), then the stroke polygon and
// the effective AA mode.  `empty`: zero nodes.
struct StrokeDraw {
  StrokePolygon poly;
  uint32_t aa = 0;
  bool empty = false;
  // (edge, band) bound of the polygon for band height pairs_bh (0: not
  // computed), as batch_prepare would compute it
  uint64_t pairs = 0;
  uint32_t pairs_bh = 0;
};

static int32_t stroke_draw(const z2d_path_node *nodes, uint64_t n, const z2d_stroke_opts *o,
                           uint32_t format, StrokeDraw &d) {
  if (!ctm_invertible(o->transformation)) return Z2D_ERR_INVALID_MATRIX;
  d.empty = n == 0;
  if (n == 0) return Z2D_OK;
  if (!nodes || o->op >= Z2D_OP_COUNT || o->anti_aliasing_mode > 3 || o->line_cap_mode > 2 ||
      o->line_join_mode > 2 || (o->n_dashes && !o->dashes))
    return Z2D_ERR_INVALID_ARGUMENT;
  if (o->hairline) return Z2D_ERR_UNSUPPORTED;
  return stroke_polygon(nodes, n, o, format, d.aa, d.poly);
}

// borrow: the batch reads d's edges at prepare instead of copying them (d
// must outlive that).
static int32_t batch_add_stroke(z2d_gpu_batch *b, z2d_gpu_surface *s,
                                const z2d_pattern_desc *pattern, const StrokeDraw &d,
                                const z2d_stroke_opts *o, bool borrow = false) {
  if (d.empty) return Z2D_OK;
  return batch_fill_polygon(b, s, pattern, d.poly.edges.data(), d.poly.edges.size(),
                            d.poly.ext, d.aa, Z2D_FILL_NON_ZERO, o->op, o->precision, true,
                            borrow, d.pairs, d.pairs_bh);
}

// The (edge, band) bound batch_prepare computes for a polygon draw on a
// surface of height h with band height bh.
static uint64_t polygon_pairs(const StrokeDraw &d, uint32_t format, int32_t h, uint32_t bh) {
  const double sc = (d.aa == Z2D_AA_NONE || format == Z2D_FORMAT_ALPHA1) ? 1.0 : 4.0;
  const uint32_t nb = (uint32_t)((h + bh - 1) / bh);
  uint64_t pairs = 0;
  for (const z2d_edge &e : d.poly.edges) pairs += band_span(e.y0 / sc, e.y1 / sc, nb, bh);
  return pairs;
}

static int32_t batch_stroke(z2d_gpu_batch *b, z2d_gpu_surface *s, const z2d_pattern_desc *pattern,
                            const z2d_path_node *nodes, uint64_t n, const z2d_stroke_opts *o) {
  if (!b || !s || !o) return Z2D_ERR_INVALID_ARGUMENT;
  int32_t rc = check_pattern(pattern);
  if (rc) return rc;
  StrokeDraw d;
  rc = stroke_draw(nodes, n, o, s->format, d);
  if (rc) return rc;
  return batch_add_stroke(b, s, pattern, d, o);
}

static void batch_drop_graph(z2d_gpu_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
}

static int32_t batch_prepare(z2d_gpu_batch *b) {
  z2d_gpu_ctx *ctx = b->ctx;
  if (b->prepared) return Z2D_OK;
  b->t_count_wait = 0;
  HIP_OK(ctx, hipSetDevice(ctx->device));
  batch_drop_graph(b);  // its kernel arguments point into the arena
  // resident gradients this batch uses: their uploads and setup first
  if (b->paints.cache && b->paints.cache_used) HIP_OK(ctx, b->paints.cache->wait_on(bstream(b)));
  const uint32_t ns = (uint32_t)b->surfaces.size();
  // Draws grouped by surface, insertion order kept within a surface.
  std::stable_sort(b->draws.begin(), b->draws.end(), [](const HostDraw &a, const HostDraw &c) {
    return a.surface_slot < c.surface_slot;
  });
  std::vector<uint32_t> surf_draw_off(ns + 1, 0), surf_band_off(ns + 1, 0);
  b->surf_tile_off.assign(ns + 1, 0);
  const uint32_t band_h = b->band_h;
  const uint32_t spt = band_h / ROWS_PER_THREAD;  // strips per tile
  const hipStream_t st = bstream(b);

  // ---- the fill draws' flattening, count pass: node counts, balance and
  // line-after-close flags, boxes and (edge, band) bounds come back here
  std::vector<FlatJob> jobs;
  std::vector<uint32_t> job_draw;
  std::vector<uint2> curve_list;  // (job, raw node) of every curve, in draw and node order
  for (uint32_t i = 0; i < (uint32_t)b->draws.size(); i++) {
    const HostDraw &h = b->draws[i];
    if (h.d.input_edges) continue;
    const z2d_gpu_surface *s = b->surfaces[h.surface_slot];
    FlatJob fj;
    std::memset(&fj, 0, sizeof(fj));
    fj.raw_off = h.raw_off;
    fj.raw_cnt = h.raw_cnt;
    fj.nbands = (uint32_t)((s->h + band_h - 1) / band_h);
    fj.band_h = band_h;
    fj.tol_sq = h.tol_sq;
    fj.curve_base = (uint32_t)curve_list.size();
    for (uint32_t k = 0; k < h.raw_cnt; k++)
      if (b->raw_nodes[h.raw_off + k].tag == Z2D_NODE_CURVE_TO)
        curve_list.push_back(make_uint2((uint32_t)jobs.size(), h.raw_off + k));
    jobs.push_back(fj);
    job_draw.push_back(i);
  }
  size_t fo_jobs = 0, fo_res = 0, fo_rel = 0, fo_curves = 0, fo_tasks = 0, fo_toff = 0,
         fo_cv = 0;
  const uint32_t n_curves = (uint32_t)curve_list.size();
  if (!jobs.empty()) {
    Layout FL;
    FL.take(b->raw_nodes.size() * sizeof(z2d_path_node));
    fo_curves = FL.take(curve_list.size() * sizeof(uint2));
    fo_jobs = FL.take(jobs.size() * sizeof(FlatJob));
    fo_res = FL.take(jobs.size() * sizeof(FlatResult));
    fo_rel = FL.take(b->raw_nodes.size() * sizeof(uint32_t));  // per raw node output offsets
    fo_tasks = FL.take((size_t)n_curves * FLAT_TASKS * sizeof(FlatTask));
    fo_toff = FL.take((size_t)n_curves * FLAT_TASKS * sizeof(uint32_t));
    fo_cv = FL.take((size_t)n_curves * sizeof(FlatCurve));
    hipError_t e = b->flat_buf.ensure(FL.off);
    if (e == hipSuccess) e = b->flat_stage.ensure(FL.off);
    if (e != hipSuccess) return fail(ctx, e, "flatten buffers");
    uint8_t *FH = b->flat_stage.p, *FD = b->flat_buf.p;
    std::memcpy(FH, b->raw_nodes.data(), b->raw_nodes.size() * sizeof(z2d_path_node));
    std::memcpy(FH + fo_jobs, jobs.data(), jobs.size() * sizeof(FlatJob));
    if (n_curves) std::memcpy(FH + fo_curves, curve_list.data(), curve_list.size() * sizeof(uint2));
    HIP_OK(ctx, hipMemcpyAsync(FD, FH, fo_res, hipMemcpyHostToDevice, st));
    HIP_OK(ctx, z2d_launch_flatten(reinterpret_cast<const z2d_path_node *>(FD),
                                   reinterpret_cast<const FlatJob *>(FD + fo_jobs),
                                   (uint32_t)jobs.size(), nullptr,
                                   reinterpret_cast<FlatResult *>(FD + fo_res),
                                   reinterpret_cast<uint32_t *>(FD + fo_rel),
                                   reinterpret_cast<const uint2 *>(FD + fo_curves), n_curves,
                                   reinterpret_cast<FlatTask *>(FD + fo_tasks),
                                   reinterpret_cast<uint32_t *>(FD + fo_toff),
                                   reinterpret_cast<FlatCurve *>(FD + fo_cv), false, st,
                                   nullptr, true));
    HIP_OK(ctx, hipMemcpyAsync(FH + fo_res, FD + fo_res, jobs.size() * sizeof(FlatResult),
                               hipMemcpyDeviceToHost, st));
    const auto cw0 = std::chrono::steady_clock::now();
    HIP_OK(ctx, hipStreamSynchronize(st));
    b->t_count_wait =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - cw0).count();
    const FlatResult *res = reinterpret_cast<const FlatResult *>(FH + fo_res);
    b->failed.clear();
    for (size_t k = 0; k < jobs.size(); k++) {
      HostDraw &h = b->draws[job_draw[k]];
      const FlatResult &r = res[k];
      if (r.flags & FLAT_RUNAWAY) {
        // A curve needed more than the subdivision visit limit: its control
        // points lie ~1e10 px apart (the reference would allocate millions
        // of edges) or are infinite (the reference recurses without end).
        bool fin = true;
        for (uint32_t q = 0; q < h.raw_cnt && fin; q++)
          for (int c = 0; c < 6; c++) fin &= std::isfinite(b->raw_nodes[h.raw_off + q].v[c]) ||
                                                b->raw_nodes[h.raw_off + q].tag != Z2D_NODE_CURVE_TO;
        const int32_t code = fin ? Z2D_ERR_UNSUPPORTED : Z2D_ERR_INVALID_ARGUMENT;
        if (!b->isolate_failures) return code;
        b->failed.emplace_back(h.seq, code);
        h.d.node_cnt = 0;  // the draw renders nothing: its emit job is an
        h.edge_pairs = 0;  // empty walk of the per-draw flattener
        jobs[k].raw_cnt = 0;
        jobs[k].serial = 1;
        continue;
      }
      h.d.node_cnt = r.n_nodes;
      h.edge_pairs = r.edge_pairs;
      if (r.flags & FLAT_UNBALANCED) h.d.flags |= DRAW_UNBALANCED | DRAW_ROWPROG;
      if (r.flags & FLAT_SERIAL) h.d.flags |= DRAW_SERIAL_PLOT;
      jobs[k].serial = (r.flags & FLAT_NONSIMPLE) ? 1u : 0u;
      for (int c = 0; c < 4; c++) h.box[c] = r.box[c];
      h.whole = !h.bounded || !(std::isfinite(r.box[0]) && std::isfinite(r.box[1]) &&
                                std::isfinite(r.box[2]) && std::isfinite(r.box[3]));
    }
  }
  // borrowed polygon edges follow the copied ones in the input-edge array
  {
    uint64_t ext_off = b->in_edges.size();
    for (HostDraw &h : b->draws)
      if (h.ext_edges) {
        h.d.in_edge_off = (uint32_t)std::min<uint64_t>(ext_off, 0xFFFFFFFFull);
        ext_off += h.d.node_cnt;
      }
    if (ext_off > 0xFFFFFFF0ull) return Z2D_ERR_OUT_OF_MEMORY;
  }
  // node ranges: flattened nodes of fills, one edge slot per edge of polygons
  uint64_t total_nodes = 0;
  for (HostDraw &h : b->draws) {
    h.d.node_off = (uint32_t)std::min<uint64_t>(total_nodes, 0xFFFFFFFFull);
    total_nodes += h.d.node_cnt;
  }
  b->n_nodes = total_nodes;
  for (const HostDraw &h : b->draws) surf_draw_off[h.surface_slot + 1]++;
  for (uint32_t i = 0; i < ns; i++) {
    surf_draw_off[i + 1] += surf_draw_off[i];
    const z2d_gpu_surface *s = b->surfaces[i];
    uint32_t bands = (uint32_t)((s->h + band_h - 1) / band_h);
    uint32_t tcols = (uint32_t)((s->w + TILE_W - 1) / TILE_W);
    surf_band_off[i + 1] = surf_band_off[i] + bands;
    b->surf_tile_off[i + 1] = b->surf_tile_off[i] + bands * tcols;
  }
  const uint32_t n_bands = surf_band_off[ns];
  const uint32_t n_tiles = b->surf_tile_off[ns];
  const uint32_t nd = (uint32_t)b->draws.size();
  // band pass slots: each band of a surface with its draws in groups of 64
  uint64_t n_slots = 0;
  for (uint32_t i = 0; i < ns; i++) {
    const uint64_t groups = std::max<uint64_t>(1, (surf_draw_off[i + 1] - surf_draw_off[i] + 63) / 64);
    n_slots += groups * (surf_band_off[i + 1] - surf_band_off[i]);
  }
  if (n_slots > 0xFFFFFFF0ull) return Z2D_ERR_OUT_OF_MEMORY;
  uint64_t max_entries = 0, max_rows = 0, scratch = 0, max_entry_edges = 0;
  uint64_t max_strip_entries = 0;
  std::vector<uint32_t> rowprog_list, rowprog_soff;
  std::vector<uint2> egroups;
  for (uint32_t i = 0; i < nd; i++) {
    HostDraw &h = b->draws[i];
    h.d.surface = h.surface_slot;
    {
      // band / tile-column estimates and the (edge, band) bound for band_h
      const z2d_gpu_surface *s = b->surfaces[h.surface_slot];
      const uint32_t bh = b->band_h;
      const uint32_t nb = (uint32_t)((s->h + bh - 1) / bh);
      const uint32_t nc = (uint32_t)((s->w + TILE_W - 1) / TILE_W);
      if (h.whole) {
        h.band_lo = 0;
        h.band_hi = nb - 1;
        h.tcol_lo = 0;
        h.tcol_hi = nc - 1;
      } else {
        auto clamped = [](double v, int32_t n) {
          return std::max(0.0, std::min(v, (double)(n - 1)));
        };
        h.band_lo = (uint32_t)clamped(std::floor(h.box[2]) - 2.0, s->h) / bh;
        h.band_hi = (uint32_t)clamped(std::ceil(h.box[3]) + 2.0, s->h) / bh;
        h.tcol_lo = (uint32_t)clamped(std::floor(h.box[0]) - 2.0, s->w) / TILE_W;
        h.tcol_hi = (uint32_t)clamped(std::ceil(h.box[1]) + 2.0, s->w) / TILE_W;
      }
      if (h.d.input_edges && h.pairs_bh == bh) {  // computed by the edges' owner
        h.edge_pairs = h.pairs_hint;
      } else if (h.d.input_edges) {  // (fills: from the flattener's count pass)
        h.edge_pairs = 0;
        const z2d_edge *ed = h.ext_edges ? h.ext_edges : b->in_edges.data() + h.d.in_edge_off;
        for (uint32_t k = 0; k < h.d.node_cnt; k++)
          h.edge_pairs += band_span(ed[k].y0 / h.edge_scale, ed[k].y1 / h.edge_scale, nb, bh);
      }
    }
    h.d.band_lo = h.band_lo;
    h.d.band_hi = h.band_hi;
    h.d.entry_base = (uint32_t)std::min<uint64_t>(max_entries, 0xFFFFFFFFull);
    for (uint32_t k = 0; k < h.d.node_cnt; k += EDGE_GROUP) egroups.push_back(make_uint2(i, k));
    max_entries += h.band_hi - h.band_lo + 1;
    max_strip_entries += (uint64_t)(h.band_hi - h.band_lo + 1) * spt *
                         (h.tcol_hi - h.tcol_lo + 1);
    max_entry_edges += h.edge_pairs;
    if (h.d.flags & DRAW_ROWPROG) {
      const z2d_gpu_surface *s = b->surfaces[h.surface_slot];
      h.d.row_off = (uint32_t)max_rows;
      max_rows += (uint64_t)(h.d.mode == MODE_DIRECT ? 1 : 4) * (uint64_t)s->h + 4;
      rowprog_list.push_back(i);
      rowprog_soff.push_back((uint32_t)scratch);
      scratch += h.d.node_cnt + 1;
    }
  }
  if (max_entries > 0xFFFFFFF0ull || total_nodes > 0xFFFFFFF0ull ||
      max_entry_edges > 0xFFFFFFF0ull || max_strip_entries > 0xFFFFFFF0ull ||
      (uint64_t)n_tiles * spt > 0xFFFFFFF0ull || egroups.size() > 0xFFFFFFF0ull ||
      2 * max_entries + 1 > 0xFFFFFFF0ull)
    return Z2D_ERR_OUT_OF_MEMORY;
  b->n_rowprog = (uint32_t)rowprog_list.size();

  // ---- arena layout
  Layout L;
  const size_t o_inedges = L.take((b->in_edges.size() + b->n_ext_edges) * sizeof(z2d_edge));
  const size_t o_draws = L.take(nd * sizeof(DevDraw));
  const size_t o_surfs = L.take(ns * sizeof(DevSurface));
  const PaintTable::Offsets o_pt = b->paints.take(L);
  const size_t o_sdo = L.take((ns + 1) * sizeof(uint32_t));
  const size_t o_sbo = L.take((ns + 1) * sizeof(uint32_t));
  const size_t o_tiles = L.take((size_t)n_tiles * sizeof(TileRef));
  const size_t o_slotinfo = L.take((size_t)n_slots * sizeof(uint2));
  const size_t o_egroups = L.take(egroups.size() * sizeof(uint2));
  const size_t o_bso = L.take((size_t)n_bands * sizeof(uint32_t));
  const size_t o_rpl = L.take(rowprog_list.size() * sizeof(uint32_t));
  const size_t o_rps = L.take(rowprog_soff.size() * sizeof(uint32_t));
  const size_t o_bn = L.take(sizeof(kBlueNoise));
  const size_t input_bytes = L.off;
  const size_t o_nodes = L.take(total_nodes * sizeof(FNode));  // written by k_flatten
  const size_t o_edges = L.take(total_nodes * sizeof(z2d_edge));
  const size_t o_bandcnt = L.take(((size_t)n_bands + 1) * sizeof(uint32_t));
  const size_t o_slotcnt = L.take(((size_t)n_slots + 1) * sizeof(uint32_t));
  const uint64_t scan_n = std::max<uint64_t>(
      std::max<uint64_t>(n_slots, max_entries), (uint64_t)n_tiles * spt);
  // per scan: a ticket and a status word per workgroup (at most
  // scan_n / SCAN_BLOCK + 1 workgroups)
  const uint32_t scan_words = (uint32_t)(scan_n / SCAN_BLOCK + 2);
  const size_t o_scantmp = L.take((size_t)3 * scan_words * sizeof(uint64_t));
  const size_t o_drows = L.take((size_t)nd * sizeof(int2));
  const size_t o_entries = L.take(max_entries * sizeof(BandEntry));
  const size_t o_blist = L.take(max_entries * sizeof(uint4));
  const size_t o_ecnt = L.take((2 * max_entries + 1) * sizeof(uint32_t));  // + cursors
  const size_t o_exr = L.take(max_entries * 2 * sizeof(int32_t));
  const size_t o_eidx = L.take(max_entry_edges * sizeof(REdge));
  const size_t o_stripcnt = L.take(((size_t)n_tiles * spt + 1) * sizeof(uint32_t));
  const size_t o_order = L.take(b->longest_first ? (size_t)n_tiles * spt * sizeof(uint32_t) : 0);
  const size_t o_ohist = L.take(b->longest_first ? (ORDER_BUCKETS + 1) * sizeof(uint32_t) : 0);
  const size_t o_sentries = L.take(max_strip_entries * sizeof(uint32_t));
  const size_t o_sslow = L.take((size_t)n_tiles * spt * sizeof(uint32_t));
  const size_t o_fast = L.take((size_t)nd * sizeof(FastDraw));
  const size_t o_counters = L.take(16 * sizeof(uint32_t));
  // render cursors: the general renderer's, then the fast one's (one per
  // launch: the single launch uses slot ns, per-surface launches slot s)
  const size_t o_cursors = L.take(2 * ((size_t)ns + 1) * sizeof(uint32_t));
  const size_t o_rows = L.take(max_rows * sizeof(RowRec));
  const size_t o_sedges = L.take(scratch * sizeof(z2d_edge));
  const size_t o_sx = L.take(scratch * sizeof(int32_t));
  const size_t total = L.off;
  hipError_t e = b->arena.ensure(total);
  if (e != hipSuccess) return fail(ctx, e, "batch arena");
  e = b->stage.ensure(input_bytes);
  if (e != hipSuccess) return fail(ctx, e, "batch staging");
  if (!b->ctr_host.p) {
    e = b->ctr_host.ensure(16 * sizeof(uint32_t), true);
    if (e != hipSuccess) return fail(ctx, e, "batch counters");
    std::memset(b->ctr_host.p, 0, 16 * sizeof(uint32_t));
  }
  uint8_t *H = b->stage.p;
  uint8_t *D = b->arena.p;

  if (!b->in_edges.empty())
    std::memcpy(H + o_inedges, b->in_edges.data(), b->in_edges.size() * sizeof(z2d_edge));
  for (const HostDraw &h : b->draws)
    if (h.ext_edges && h.d.node_cnt)
      std::memcpy(H + o_inedges + (size_t)h.d.in_edge_off * sizeof(z2d_edge), h.ext_edges,
                  (size_t)h.d.node_cnt * sizeof(z2d_edge));
  DevDraw *hdraws = reinterpret_cast<DevDraw *>(H + o_draws);
  for (uint32_t i = 0; i < nd; i++) hdraws[i] = b->draws[i].d;
  DevSurface *hs = reinterpret_cast<DevSurface *>(H + o_surfs);
  for (uint32_t i = 0; i < ns; i++) {
    const z2d_gpu_surface *s = b->surfaces[i];
    hs[i].data = s->data;
    hs[i].w = s->w;
    hs[i].h = s->h;
    hs[i].fmt = s->format;
    hs[i].pad = 0;
    hs[i].nbytes = s->nbytes;
    hs[i].clip = s->clip;
  }
  b->paints.write(H, D, o_pt);
  std::memcpy(H + o_sdo, surf_draw_off.data(), (ns + 1) * sizeof(uint32_t));
  std::memcpy(H + o_sbo, surf_band_off.data(), (ns + 1) * sizeof(uint32_t));
  TileRef *ht = reinterpret_cast<TileRef *>(H + o_tiles);
  for (uint32_t i = 0; i < ns; i++) {
    const z2d_gpu_surface *s = b->surfaces[i];
    const uint32_t bands = (uint32_t)((s->h + band_h - 1) / band_h);
    const uint32_t tcols = (uint32_t)((s->w + TILE_W - 1) / TILE_W);
    uint32_t k = b->surf_tile_off[i];
    for (uint32_t ty = 0; ty < bands; ty++)
      for (uint32_t tx = 0; tx < tcols; tx++) ht[k++] = TileRef{i, (uint16_t)tx, (uint16_t)ty};
  }
  if (!egroups.empty())
    std::memcpy(H + o_egroups, egroups.data(), egroups.size() * sizeof(uint2));
  uint2 *hsi = reinterpret_cast<uint2 *>(H + o_slotinfo);
  uint32_t *hbso = reinterpret_cast<uint32_t *>(H + o_bso);
  {
    uint32_t k = 0;
    for (uint32_t i = 0; i < ns; i++) {
      const uint32_t groups =
          std::max<uint32_t>(1, (surf_draw_off[i + 1] - surf_draw_off[i] + 63) / 64);
      for (uint32_t bd = surf_band_off[i]; bd < surf_band_off[i + 1]; bd++) {
        hbso[bd] = k;
        for (uint32_t g = 0; g < groups; g++) hsi[k++] = make_uint2(bd, i);
      }
    }
  }
  if (!rowprog_list.empty()) {
    std::memcpy(H + o_rpl, rowprog_list.data(), rowprog_list.size() * sizeof(uint32_t));
    std::memcpy(H + o_rps, rowprog_soff.data(), rowprog_soff.size() * sizeof(uint32_t));
  }
  std::memcpy(H + o_bn, kBlueNoise, sizeof(kBlueNoise));
  HIP_OK(ctx, hipMemcpyAsync(D, H, input_bytes, hipMemcpyHostToDevice, bstream(b)));
  HIP_OK(ctx, z2d_launch_grad_setup(reinterpret_cast<DevGradient *>(D + o_pt.grads),
                                    (uint32_t)b->paints.grads.size(), bstream(b)));

  JobView &J = b->view;
  J.nodes = reinterpret_cast<const FNode *>(D + o_nodes);
  b->n_flat_jobs = 0;
  b->n_flat_serial = 0;
  if (!jobs.empty()) {
    // the emit pass runs at the start of every render (batch_enqueue): the
    // flattened nodes at their draws' offsets, from the resident raw nodes
    FlatJob *hj = reinterpret_cast<FlatJob *>(b->flat_stage.p + fo_jobs);
    for (size_t k = 0; k < jobs.size(); k++) {
      hj[k].node_off = b->draws[job_draw[k]].d.node_off;
      hj[k].serial = jobs[k].serial;
      hj[k].raw_cnt = jobs[k].raw_cnt;
      b->n_flat_serial += jobs[k].serial ? 1u : 0u;
    }
    HIP_OK(ctx, hipMemcpyAsync(b->flat_buf.p + fo_jobs, hj, jobs.size() * sizeof(FlatJob),
                               hipMemcpyHostToDevice, st));
    b->n_flat_jobs = (uint32_t)jobs.size();
    b->d_flat_jobs = reinterpret_cast<const FlatJob *>(b->flat_buf.p + fo_jobs);
    b->d_flat_out = reinterpret_cast<FNode *>(D + o_nodes);
    b->d_flat_rel = reinterpret_cast<uint32_t *>(b->flat_buf.p + fo_rel);
    b->d_flat_curves = reinterpret_cast<const uint2 *>(b->flat_buf.p + fo_curves);
    b->n_flat_curves = n_curves;
    b->d_flat_tasks = reinterpret_cast<FlatTask *>(b->flat_buf.p + fo_tasks);
    b->d_flat_toff = reinterpret_cast<uint32_t *>(b->flat_buf.p + fo_toff);
  }
  J.in_edges = reinterpret_cast<const z2d_edge *>(D + o_inedges);
  J.edges = reinterpret_cast<z2d_edge *>(D + o_edges);
  J.draws = reinterpret_cast<DevDraw *>(D + o_draws);
  J.draw_rows = reinterpret_cast<int2 *>(D + o_drows);
  J.n_draws = nd;
  J.surfaces = reinterpret_cast<DevSurface *>(D + o_surfs);
  J.n_surfaces = ns;
  J.paints = reinterpret_cast<const DevPaint *>(D + o_pt.paints);
  J.blue_noise = reinterpret_cast<const uint16_t *>(D + o_bn);
  J.surf_draw_off = reinterpret_cast<const uint32_t *>(D + o_sdo);
  J.surf_band_off = reinterpret_cast<const uint32_t *>(D + o_sbo);
  J.n_bands = n_bands;
  J.band_cnt = reinterpret_cast<uint32_t *>(D + o_bandcnt);
  J.slot_info = reinterpret_cast<const uint2 *>(D + o_slotinfo);
  J.band_slot_off = reinterpret_cast<const uint32_t *>(D + o_bso);
  J.n_slots = (uint32_t)n_slots;
  J.slot_cnt = reinterpret_cast<uint32_t *>(D + o_slotcnt);
  J.scan_tmp = reinterpret_cast<uint64_t *>(D + o_scantmp);
  J.scan_words = scan_words;
  J.entries = reinterpret_cast<BandEntry *>(D + o_entries);
  J.max_entries = (uint32_t)max_entries;
  J.entry_ecnt = reinterpret_cast<uint32_t *>(D + o_ecnt);
  J.entry_cur = J.entry_ecnt + max_entries + 1;
  J.entry_xlo = reinterpret_cast<int32_t *>(D + o_exr);
  J.entry_xhi = J.entry_xlo + max_entries;
  J.band_list = reinterpret_cast<uint4 *>(D + o_blist);
  J.egroups = reinterpret_cast<const uint2 *>(D + o_egroups);
  J.n_egroups = (uint32_t)egroups.size();
  J.entry_edges = reinterpret_cast<REdge *>(D + o_eidx);
  J.max_entry_edges = (uint32_t)max_entry_edges;
  J.band_h = band_h;
  J.strips_per_tile = spt;
  J.strip_order = b->longest_first ? reinterpret_cast<uint32_t *>(D + o_order) : nullptr;
  J.order_hist = b->longest_first ? reinterpret_cast<uint32_t *>(D + o_ohist) : nullptr;
  J.strip_cnt = reinterpret_cast<uint32_t *>(D + o_stripcnt);
  J.strip_entries = reinterpret_cast<uint32_t *>(D + o_sentries);
  J.max_strip_entries = (uint32_t)max_strip_entries;
  J.strip_slow = reinterpret_cast<uint32_t *>(D + o_sslow);
  J.fast = reinterpret_cast<FastDraw *>(D + o_fast);
  J.counters = reinterpret_cast<uint32_t *>(D + o_counters);
  J.count_stats = b->count_stats;
  J.cursors = reinterpret_cast<uint32_t *>(D + o_cursors);
  J.rows = reinterpret_cast<RowRec *>(D + o_rows);
  J.max_rows = (uint32_t)max_rows;
  J.tiles = reinterpret_cast<const TileRef *>(D + o_tiles);
  J.n_tiles = n_tiles;
  b->d_rowprog_list = reinterpret_cast<uint32_t *>(D + o_rpl);
  b->d_rowprog_soff = reinterpret_cast<uint32_t *>(D + o_rps);
  b->d_scratch_edges = reinterpret_cast<z2d_edge *>(D + o_sedges);
  b->d_scratch_x = reinterpret_cast<int32_t *>(D + o_sx);
  if (!b->async_prepare) HIP_OK(ctx, hipStreamSynchronize(bstream(b)));
  b->prepared = true;
  return Z2D_OK;
}

// The render pipeline of a prepared batch on its stream: counters and
// cursors cleared by fill kernels (no memset nodes, so the sequence can be
// captured into a graph), edge build, row programs, binning, tile render.
// mark_render records ev[1] between the binning and the render.
static hipError_t batch_enqueue(z2d_gpu_batch *b, uint32_t mode, bool mark_render) {
  z2d_gpu_ctx *ctx = b->ctx;
  JobView &J = b->view;
  const uint32_t ns = J.n_surfaces;
  const hipStream_t st = bstream(b);
  // every per-render reset: counters, render cursors, the entries' edge
  // counts and cursors, their column ranges (empty until the edge binning
  // widens them) and the scans' look-back words; done by the flattening's
  // per-node kernel when there is one, else in a launch of their own
  FillSet F{};
  auto add = [&](uint32_t *p, uint64_t n, uint32_t v) {
    if (n == 0) return;
    F.p[F.count] = p;
    F.n[F.count] = n;
    F.v[F.count] = v;
    F.count++;
  };
  add(J.counters, 16, 0u);
  add(J.cursors, 2 * ((uint64_t)ns + 1), 0u);
  if (J.n_bands) {
    add(J.entry_ecnt, 2ull * J.max_entries + 1, 0u);
    add(reinterpret_cast<uint32_t *>(J.entry_xlo), J.max_entries, 0x7FFFFFFFu);
    add(reinterpret_cast<uint32_t *>(J.entry_xhi), J.max_entries, 0x80000000u);
    add(reinterpret_cast<uint32_t *>(J.scan_tmp), 6ull * J.scan_words, 0u);
  }
  hipError_t e = hipSuccess;
  if (b->n_flat_jobs)
    e = z2d_launch_flatten(reinterpret_cast<const z2d_path_node *>(b->flat_buf.p),
                           b->d_flat_jobs, b->n_flat_jobs, b->d_flat_out, nullptr,
                           b->d_flat_rel, b->d_flat_curves, b->n_flat_curves, b->d_flat_tasks,
                           b->d_flat_toff, nullptr, true, st, &F, b->n_flat_serial != 0);
  else
    e = z2d_launch_fill_set(F, st);
  if (e == hipSuccess) e = z2d_launch_edge_build(J, st);
  if (e == hipSuccess)
    e = z2d_launch_rowprog(J, b->d_rowprog_list, b->n_rowprog, b->d_scratch_edges,
                           b->d_scratch_x, b->d_rowprog_soff, st);
  if (e == hipSuccess) e = z2d_launch_binning(J, st);
  if (e == hipSuccess && mark_render) e = hipEventRecord(b->ev[1], st);
  // persistent workers: up to 8 four-wave workgroups of the general renderer
  // and 16 one-wave workgroups of the fast one per CU (a launch with more
  // strips than workers drains a shared queue)
  const uint32_t max_blocks = (uint32_t)ctx->n_cus * 8;
  const uint32_t max_fast = (uint32_t)ctx->n_cus * 16;
  uint32_t *cur_fast = J.cursors + (ns + 1);
  if (mode == 0) {
    const uint32_t items = J.n_tiles * J.strips_per_tile;
    if (e == hipSuccess)
      e = z2d_launch_render(J, 0, J.n_tiles, J.cursors + ns, cur_fast + ns,
                            std::min(J.n_tiles, max_blocks), std::min(items, max_fast), st);
  } else {
    for (uint32_t s = 0; s < ns && e == hipSuccess; s++) {
      const uint32_t t0 = b->surf_tile_off[s], tn = b->surf_tile_off[s + 1] - t0;
      e = z2d_launch_render(J, t0, tn, J.cursors + s, cur_fast + s, std::min(tn, max_blocks),
                            std::min(tn * J.strips_per_tile, max_fast), st);
    }
  }
  // the counters (work counts, bin-overflow flag) to pinned memory, read
  // once the stream has completed
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->ctr_host.p, J.counters, 16 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                       st);
  return e;
}

static int32_t batch_render(z2d_gpu_batch *b, uint32_t mode) {
  z2d_gpu_ctx *ctx = b->ctx;
  if (!b->prepared) {
    int32_t rc = batch_prepare(b);
    if (rc) return rc;
  }
  bool any_clip = false;
  for (const z2d_gpu_surface *s : b->surfaces) any_clip |= s->clip != nullptr;
  if (b->draws.empty() && !any_clip) return Z2D_OK;
  HIP_OK(ctx, hipSetDevice(ctx->device));
  const hipStream_t st = bstream(b);
  for (int i = 0; i < 3; i++)
    if (!b->ev[i]) HIP_OK(ctx, hipEventCreate(&b->ev[i]));
  if (!b->use_graph) {
    HIP_OK(ctx, hipEventRecord(b->ev[0], st));
    HIP_OK(ctx, batch_enqueue(b, mode, true));
    HIP_OK(ctx, hipEventRecord(b->ev[2], st));
    b->render_timed = true;
    return Z2D_OK;
  }
  // graph replay: captured once per prepared batch, launch mode and counter
  // setting (all of them are kernel arguments)
  if (b->gexec && (b->graph_mode != mode || b->graph_stats != b->view.count_stats))
    batch_drop_graph(b);
  if (!b->gexec) {
    HIP_OK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = batch_enqueue(b, mode, false);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      b->gexec = nullptr;
      return fail(ctx, e, "batch graph capture");
    }
    b->graph_mode = mode;
    b->graph_stats = b->view.count_stats;
  }
  HIP_OK(ctx, hipEventRecord(b->ev[0], st));
  HIP_OK(ctx, hipGraphLaunch(b->gexec, st));
  HIP_OK(ctx, hipEventRecord(b->ev[2], st));
  b->render_timed = false;
  return Z2D_OK;
}

static int32_t batch_wait(z2d_gpu_batch *b) {
  z2d_gpu_ctx *ctx = b->ctx;
  HIP_OK(ctx, hipStreamSynchronize(bstream(b)));
  // A bin outside the host's capacity bounds drops work: the pixels would be
  // wrong, so the render is an error (never expected: the bounds are upper
  // bounds of what the kernels write).
  if (b->ctr_host.p && reinterpret_cast<const uint32_t *>(b->ctr_host.p)[2]) {
    ctx->err = "render: a bin exceeded its capacity bound, or a scan's look-back timed out";
    return Z2D_ERR_DEVICE;
  }
  if (b->ev[0] && b->ev[2]) {
    float a = 0.0f, r = 0.0f;
    if (hipEventElapsedTime(&a, b->ev[0], b->ev[2]) == hipSuccess) b->last_pipeline_ms = a;
    if (!b->render_timed)
      b->last_render_ms = -1.0f;  // a graph replay has no mark before the render
    else if (hipEventElapsedTime(&r, b->ev[1], b->ev[2]) == hipSuccess)
      b->last_render_ms = r;
  }
  return Z2D_OK;
}

static void batch_clear(z2d_gpu_batch *b) {
  b->surfaces.clear();
  b->raw_nodes.clear();
  b->n_nodes = 0;
  b->in_edges.clear();
  b->n_ext_edges = 0;
  b->draws.clear();
  b->paints.clear();
  b->prepared = false;
}

// ---------------------------------------------------------------------------
// surfaces from descriptors
// ---------------------------------------------------------------------------
static int32_t surface_alloc(z2d_gpu_ctx *ctx, uint32_t fmt, int32_t w, int32_t h,
                             z2d_gpu_surface **out) {
  if (w < 1) return Z2D_ERR_INVALID_WIDTH;
  if (h < 1) return Z2D_ERR_INVALID_HEIGHT;
  if (fmt > Z2D_FORMAT_ALPHA1) return Z2D_ERR_INVALID_ARGUMENT;
  auto *s = new (std::nothrow) z2d_gpu_surface();
  if (!s) return Z2D_ERR_OUT_OF_MEMORY;
  s->ctx = ctx;
  s->format = fmt;
  s->w = w;
  s->h = h;
  s->nbytes = surface_bytes(fmt, w, h);
  HIP_OK(ctx, hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(&s->data, padded_bytes(s->nbytes));
  if (e != hipSuccess) {
    delete s;
    return fail(ctx, e, "surface alloc");
  }
  e = hipMemsetAsync(s->data, 0, padded_bytes(s->nbytes), ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(s->data);
    delete s;
    return fail(ctx, e, "surface clear");
  }
  *out = s;
  return Z2D_OK;
}

// A descriptor as a surface: device descriptors are wrapped, host ones are
// uploaded into a temporary device surface (copied back by the caller).
struct SurfRef {
  z2d_gpu_surface *s = nullptr;
  bool temp = false;
  z2d_surface_desc *desc = nullptr;
};

static int32_t surface_from_desc(z2d_gpu_ctx *ctx, z2d_surface_desc *d, SurfRef &r) {
  if (!d || !d->data) return Z2D_ERR_INVALID_ARGUMENT;
  if (d->width < 1) return Z2D_ERR_INVALID_WIDTH;
  if (d->height < 1) return Z2D_ERR_INVALID_HEIGHT;
  if (d->format > Z2D_FORMAT_ALPHA1) return Z2D_ERR_INVALID_ARGUMENT;
  uint64_t need = surface_bytes(d->format, d->width, d->height);
  if (d->len < need) return Z2D_ERR_INVALID_ARGUMENT;
  r.desc = d;
  if (d->reserved == Z2D_SURFACE_DEVICE) {
    auto *s = new (std::nothrow) z2d_gpu_surface();
    if (!s) return Z2D_ERR_OUT_OF_MEMORY;
    s->ctx = ctx;
    s->format = d->format;
    s->w = d->width;
    s->h = d->height;
    s->nbytes = need;
    s->data = d->data;
    s->owned = false;
    r.s = s;
    r.temp = false;
    return Z2D_OK;
  }
  int32_t rc = surface_alloc(ctx, d->format, d->width, d->height, &r.s);
  if (rc) return rc;
  r.temp = true;
  HIP_OK(ctx, hipMemcpyAsync(r.s->data, d->data, need, hipMemcpyHostToDevice, ctx->stream));
  return Z2D_OK;
}

static void surface_free(z2d_gpu_surface *s) {
  if (!s) return;
  if (s->owned && s->data) (void)hipFree(s->data);
  delete s;
}

static int32_t surface_release(z2d_gpu_ctx *ctx, SurfRef &r, bool copy_back) {
  int32_t rc = Z2D_OK;
  if (r.s && r.temp && copy_back) {
    hipError_t e =
        hipMemcpyAsync(r.desc->data, r.s->data, r.s->nbytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = fail(ctx, e, "surface download");
  }
  if (r.s) {
    if (r.temp) (void)hipStreamSynchronize(ctx->stream);
    surface_free(r.s);
  }
  r.s = nullptr;
  return rc;
}

// ---------------------------------------------------------------------------
// compositor
// ---------------------------------------------------------------------------
// A source of the quad compositor: a pixel, or a surface read at the
// destination's pixel index (same width, at least as tall, aligned for the
// vector loads) in a 4-byte format or alpha8.
static bool quad_src(const DevCompParam &p, const CompJob &job, QuadSrc &o) {
  o.data = nullptr;
  o.fmt = 0;
  o.px = I4{0, 0, 0, 0};
  if (p.kind == Z2D_PARAM_PIXEL) {
    o.kind = QUAD_PIXEL;
    o.px = p.px;
    return true;
  }
  if (p.kind != Z2D_PARAM_SURFACE || p.base != 0 || p.surf.w != job.dst.w ||
      p.surf.h < job.dst.h)
    return false;
  const uintptr_t a = reinterpret_cast<uintptr_t>(p.surf.data);
  if (fbits(p.surf.fmt) == 32 && (a & 15) == 0)
    o.kind = QUAD_SURF32;
  else if (p.surf.fmt == Z2D_FORMAT_ALPHA8 && (a & 3) == 0)
    o.kind = QUAD_ALPHA8;
  else
    return false;
  o.data = p.surf.data;
  o.fmt = p.surf.fmt;
  return true;
}

// The batches the quad compositor takes (on top of the streaming conditions):
// one operator onto the destination with a surface or pixel source, or
// {dst_in dst=src src=mask}, {op} (src and mask surfaces or pixels).
static bool quad_job(const CompJob &job, QuadJob &q) {
  if ((reinterpret_cast<uintptr_t>(job.dst.data) & 15) ||
      (reinterpret_cast<uintptr_t>(job.out.data) & 15))
    return false;
  q.dst = reinterpret_cast<const uint32_t *>(job.dst.data);
  q.out = reinterpret_cast<uint32_t *>(job.out.data);
  q.dst_fmt = job.dst.fmt;
  q.out_fmt = job.out.fmt;
  q.precision = job.precision;
  if (job.n_ops == 1) {
    const DevCompOp &o = job.ops[0];
    if (o.dst.kind != Z2D_PARAM_NONE || !quad_src(o.src, job, q.src)) return false;
    q.op = o.op;
    q.masked = 0;
    q.mask = q.src;
    return true;
  }
  if (job.n_ops == 2) {
    const DevCompOp &m = job.ops[0], &o = job.ops[1];
    if (m.op != Z2D_OP_DST_IN || o.dst.kind != Z2D_PARAM_NONE || o.src.kind != Z2D_PARAM_NONE)
      return false;
    if (!quad_src(m.dst, job, q.src) || !quad_src(m.src, job, q.mask)) return false;
    q.op = o.op;
    q.masked = 1;
    return true;
  }
  return false;
}

// noalloc: the variant that allocates no device or pinned memory: every
// surface must be a device descriptor, the job lives in the device's fixed
// scratch (a job larger than it is refused) and is staged from pageable
// host memory, and no timing events are recorded.
static int32_t composite_impl(z2d_gpu_ctx *ctx, z2d_surface_desc *dst_d, int32_t dst_x,
                              int32_t dst_y, const z2d_comp_op *ops, uint64_t n_ops,
                              uint32_t precision, z2d_surface_desc *out_d, bool noalloc = false) {
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  if (!dst_d || (n_ops && !ops) || n_ops > 16) return Z2D_ERR_INVALID_ARGUMENT;
  if (n_ops == 0) return Z2D_OK;
  if (noalloc) {
    auto on_device = [](const z2d_surface_desc *d) {
      return d == nullptr || d->reserved == Z2D_SURFACE_DEVICE;
    };
    bool ok = on_device(dst_d) && on_device(out_d);
    for (uint64_t i = 0; i < n_ops && ok; i++) {
      if (ops[i].dst.kind == Z2D_PARAM_SURFACE) ok = on_device(ops[i].dst.surface);
      if (ok && ops[i].src.kind == Z2D_PARAM_SURFACE) ok = on_device(ops[i].src.surface);
    }
    if (!ok) {
      ctx->err = "composite_noalloc: every surface must be a device descriptor";
      return Z2D_ERR_INVALID_ARGUMENT;
    }
    if (!ctx->comp_scratch) {
      ctx->err = "composite_noalloc: the device scratch address is unavailable";
      return Z2D_ERR_DEVICE;
    }
  }
  if (dst_x >= dst_d->width || dst_y >= dst_d->height) return Z2D_OK;
  for (uint64_t i = 0; i < n_ops; i++)
    if (ops[i].op >= Z2D_OP_COUNT) return Z2D_ERR_INVALID_ARGUMENT;
  // precision: float if any operator requires it
  for (uint64_t i = 0; i < n_ops; i++)
    if (op_requires_float(ops[i].op)) precision = Z2D_PRECISION_FLOAT;
  int32_t sw, sh;
  switch (ops[0].src.kind) {
    case Z2D_PARAM_NONE: return Z2D_OK;
    case Z2D_PARAM_SURFACE:
      if (!ops[0].src.surface) return Z2D_ERR_INVALID_ARGUMENT;
      sw = ops[0].src.surface->width;
      sh = ops[0].src.surface->height;
      break;
    default:
      if (dst_x != 0 || dst_y != 0) return Z2D_OK;
      sw = dst_d->width;
      sh = dst_d->height;
      break;
  }
  const int32_t ssx = dst_x < 0 ? -dst_x : 0, ssy = dst_y < 0 ? -dst_y : 0;
  const int32_t width = (sw + dst_x > dst_d->width) ? dst_d->width - dst_x : sw;
  const int32_t height = (sh + dst_y > dst_d->height) ? dst_d->height - dst_y : sh;
  if (ssx >= width || ssy >= height) return Z2D_OK;

  HIP_OK(ctx, hipSetDevice(ctx->device));
  SurfRef dst, out;
  int32_t rc = surface_from_desc(ctx, dst_d, dst);
  if (rc) return rc;
  if (out_d) {
    if (out_d->format != dst_d->format || out_d->width != dst_d->width ||
        out_d->height != dst_d->height) {
      surface_release(ctx, dst, false);
      return Z2D_ERR_INVALID_ARGUMENT;
    }
    rc = surface_from_desc(ctx, out_d, out);
    if (rc) {
      surface_release(ctx, dst, false);
      return rc;
    }
  }
  std::vector<SurfRef> srcs;
  PaintTable pt;
  CompJob job;
  std::memset(&job, 0, sizeof(job));
  auto devsurf = [](const z2d_gpu_surface *s) {
    DevSurface d;
    d.data = s->data;
    d.w = s->w;
    d.h = s->h;
    d.fmt = s->format;
    d.pad = 0;
    d.nbytes = s->nbytes;
    d.clip = nullptr;
    return d;
  };
  job.dst = devsurf(dst.s);
  job.out = out_d ? devsurf(out.s) : job.dst;
  job.x0 = ssx + dst_x;
  job.y0 = ssy + dst_y;
  job.width = width - ssx;
  job.height = (height > ssy ? height : ssy) - ssy;
  job.src_row0 = ssy;
  job.n_ops = (uint32_t)n_ops;
  job.precision = precision;
  std::vector<int32_t> paint_idx(2 * n_ops, -1);
  for (uint64_t k = 0; k < n_ops && rc == Z2D_OK; k++) {
    job.ops[k].op = ops[k].op;
    for (int side = 0; side < 2; side++) {
      const z2d_comp_param &p = side == 0 ? ops[k].dst : ops[k].src;
      DevCompParam &q = side == 0 ? job.ops[k].dst : job.ops[k].src;
      q.kind = p.kind;
      switch (p.kind) {
        case Z2D_PARAM_NONE: break;
        case Z2D_PARAM_PIXEL: q.px = pixel_rgba(p.pixel); break;
        case Z2D_PARAM_SURFACE: {
          SurfRef r;
          rc = surface_from_desc(ctx, const_cast<z2d_surface_desc *>(p.surface), r);
          if (rc) break;
          srcs.push_back(r);
          q.surf = devsurf(r.s);
          // getStride(x, y): dst-side params at the destination start,
          // src-side params at (src_start_x, src_y).
          int32_t bx = side == 0 ? job.x0 : ssx;
          int32_t by = side == 0 ? job.y0 : ssy;
          if (bx < 0 || by < 0 || bx >= r.s->w || by >= r.s->h) {
            rc = Z2D_ERR_INVALID_ARGUMENT;  // the reference reads an empty stride
            break;
          }
          q.base = (int64_t)r.s->w * by + bx;
          // rows of the source stride advance by the source's own width
        } break;
        case Z2D_PARAM_GRADIENT: {
          z2d_pattern_desc tmp;
          std::memset(&tmp, 0, sizeof(tmp));
          tmp.kind = Z2D_PATTERN_GRADIENT;
          tmp.gradient = p.gradient;
          if (!p.gradient) {
            rc = Z2D_ERR_INVALID_ARGUMENT;
            break;
          }
          paint_idx[2 * k + side] = (int32_t)pt.add(&tmp);
        } break;
        case Z2D_PARAM_DITHER: {
          z2d_pattern_desc tmp;
          std::memset(&tmp, 0, sizeof(tmp));
          tmp.kind = Z2D_PATTERN_DITHER;
          tmp.dither = p.dither;
          paint_idx[2 * k + side] = (int32_t)pt.add(&tmp);
        } break;
        default: rc = Z2D_ERR_INVALID_ARGUMENT; break;
      }
      if (rc) break;
    }
  }
  if (rc == Z2D_OK) {
    Layout L;
    const size_t o_job = L.take(sizeof(CompJob));
    const PaintTable::Offsets o_pt = pt.take(L);
    const size_t o_bn = L.take(sizeof(kBlueNoise));
    std::vector<uint8_t> host_stage;  // noalloc: pageable staging
    hipError_t e = hipSuccess;
    uint8_t *H = nullptr, *D = nullptr;
    if (noalloc) {
      if (L.off > kCompScratchBytes) {
        ctx->err = "composite_noalloc: the job needs " + std::to_string(L.off) +
                   " bytes of device scratch, more than its " +
                   std::to_string(kCompScratchBytes);
        rc = Z2D_ERR_BUDGET_EXCEEDED;
      } else {
        host_stage.assign(L.off, 0);
        H = host_stage.data();
        D = ctx->comp_scratch;
      }
    } else {
      e = ctx->comp_arena.ensure(L.off);
      if (e == hipSuccess) e = ctx->comp_stage.ensure(L.off);
      if (e != hipSuccess) rc = fail(ctx, e, "composite arena");
      H = ctx->comp_stage.p;
      D = ctx->comp_arena.p;
    }
    if (rc == Z2D_OK) {
      pt.write(H, D, o_pt);
      const DevPaint *dp = reinterpret_cast<const DevPaint *>(D + o_pt.paints);
      for (uint64_t k = 0; k < n_ops; k++) {
        if (paint_idx[2 * k] >= 0) job.ops[k].dst.paint = dp + paint_idx[2 * k];
        if (paint_idx[2 * k + 1] >= 0) job.ops[k].src.paint = dp + paint_idx[2 * k + 1];
      }
      job.blue_noise = reinterpret_cast<const uint16_t *>(D + o_bn);
      std::memcpy(H + o_job, &job, sizeof(job));
      std::memcpy(H + o_bn, kBlueNoise, sizeof(kBlueNoise));
      e = hipMemcpyAsync(D, H, L.off, hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess)
        e = z2d_launch_grad_setup(reinterpret_cast<DevGradient *>(D + o_pt.grads),
                                  (uint32_t)pt.grads.size(), ctx->stream);
      if (e == hipSuccess) {
        // Streaming path: whole 4-byte-format surfaces, aligned rows, every
        // surface source covering the destination at the origin.
        bool stream = fbits(job.dst.fmt) == 32 && fbits(job.out.fmt) == 32 && job.x0 == 0 &&
                      job.y0 == 0 && job.width == job.dst.w && job.height == job.dst.h &&
                      (job.dst.w % 4) == 0;
        for (uint64_t k = 0; k < n_ops && stream; k++)
          for (int side = 0; side < 2; side++) {
            const DevCompParam &q = side == 0 ? job.ops[k].dst : job.ops[k].src;
            if (q.kind == Z2D_PARAM_SURFACE && (q.base != 0 || q.surf.w != job.dst.w))
              stream = false;
          }
        if (!noalloc) {
          if (!ctx->ev[0]) e = hipEventCreate(&ctx->ev[0]);
          if (e == hipSuccess && !ctx->ev[1]) e = hipEventCreate(&ctx->ev[1]);
          if (e == hipSuccess) e = hipEventRecord(ctx->ev[0], ctx->stream);
        }
        QuadJob qj;
        if (e != hipSuccess) {
        } else if (stream && quad_job(job, qj))
          e = z2d_launch_composite_quad(qj, (uint64_t)job.width * (uint64_t)job.height / 4,
                                        ctx->stream);
        else
          e = z2d_launch_composite(reinterpret_cast<const CompJob *>(D + o_job), job, stream,
                                   ctx->stream);
      }
      if (e == hipSuccess && !noalloc) e = hipEventRecord(ctx->ev[1], ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      float ms = 0.0f;
      if (e == hipSuccess && !noalloc &&
          hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess)
        ctx->last_comp_ms = ms;
      if (e != hipSuccess) rc = fail(ctx, e, "composite launch");
    }
  }
  for (SurfRef &r : srcs) surface_release(ctx, r, false);
  int32_t rc2 = Z2D_OK;
  if (out_d) {
    rc2 = surface_release(ctx, out, rc == Z2D_OK);
    surface_release(ctx, dst, false);
  } else {
    rc2 = surface_release(ctx, dst, rc == Z2D_OK);
  }
  return rc ? rc : rc2;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int32_t z2d_gpu_create(int32_t device, z2d_gpu_ctx **out) {
  if (!out) return Z2D_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return Z2D_ERR_DEVICE;
  if (device < 0 || device >= count) return Z2D_ERR_INVALID_ARGUMENT;
  auto *ctx = new (std::nothrow) z2d_gpu_ctx();
  if (!ctx) return Z2D_ERR_OUT_OF_MEMORY;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return Z2D_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->n_cus = prop.multiProcessorCount;
  ctx->comp_scratch = comp_scratch_for(device);  // device is current
  *out = ctx;
  return Z2D_OK;
}

void z2d_gpu_destroy(z2d_gpu_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->comp_arena.release();
  ctx->comp_stage.release();
  ctx->png_work.release();
  ctx->png_out.release();
  ctx->png_meta.release();
  for (hipEvent_t &ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (hipEvent_t &ev : ctx->png_ev)
    if (ev) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *z2d_gpu_last_error(z2d_gpu_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

int32_t z2d_gpu_sync(z2d_gpu_ctx *ctx) {
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
  return Z2D_OK;
}

int32_t z2d_gpu_surface_create(z2d_gpu_ctx *ctx, uint32_t format, int32_t width, int32_t height,
                               z2d_gpu_surface **out) {
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  if (!out) return Z2D_ERR_INVALID_ARGUMENT;
  return surface_alloc(ctx, format, width, height, out);
}

int32_t z2d_gpu_surface_upload(z2d_gpu_surface *s, const void *host, uint64_t len) {
  if (!s || !host || len < s->nbytes) return Z2D_ERR_INVALID_ARGUMENT;
  z2d_gpu_ctx *ctx = s->ctx;
  HIP_OK(ctx, hipMemcpyAsync(s->data, host, s->nbytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
  return Z2D_OK;
}

int32_t z2d_gpu_surface_download(z2d_gpu_surface *s, void *host, uint64_t len) {
  if (!s || !host || len < s->nbytes) return Z2D_ERR_INVALID_ARGUMENT;
  z2d_gpu_ctx *ctx = s->ctx;
  HIP_OK(ctx, hipMemcpyAsync(host, s->data, s->nbytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
  return Z2D_OK;
}

int32_t z2d_gpu_surface_desc(z2d_gpu_surface *s, z2d_surface_desc *out) {
  if (!s || !out) return Z2D_ERR_INVALID_ARGUMENT;
  out->format = s->format;
  out->width = s->w;
  out->height = s->h;
  out->reserved = Z2D_SURFACE_DEVICE;
  out->data = s->data;
  out->len = s->nbytes;
  return Z2D_OK;
}

void z2d_gpu_surface_destroy(z2d_gpu_surface *s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->ctx->stream);
  surface_free(s);
}

int32_t z2d_gpu_batch_create(z2d_gpu_ctx *ctx, z2d_gpu_batch **out) {
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  if (!out) return Z2D_ERR_INVALID_ARGUMENT;
  auto *b = new (std::nothrow) z2d_gpu_batch();
  if (!b) return Z2D_ERR_OUT_OF_MEMORY;
  b->ctx = ctx;
  *out = b;
  return Z2D_OK;
}

int32_t z2d_gpu_batch_fill(z2d_gpu_batch *b, z2d_gpu_surface *surface,
                           const z2d_pattern_desc *pattern, const z2d_path_node *nodes,
                           uint64_t n_nodes, const z2d_fill_opts *opts) {
  return batch_fill(b, surface, pattern, nodes, n_nodes, opts);
}

int32_t z2d_gpu_batch_fill_polygon(z2d_gpu_batch *b, z2d_gpu_surface *surface,
                                   const z2d_pattern_desc *pattern, const z2d_edge *edges,
                                   uint64_t n_edges, const double extents[4],
                                   uint32_t aa, uint32_t rule, uint32_t op,
                                   uint32_t precision) {
  if (n_edges && !edges) return Z2D_ERR_INVALID_ARGUMENT;
  return batch_fill_polygon(b, surface, pattern, edges, n_edges, extents, aa, rule, op,
                            precision, false);
}

int32_t z2d_gpu_batch_prepare(z2d_gpu_batch *b) {
  if (!b) return Z2D_ERR_INVALID_ARGUMENT;
  return batch_prepare(b);
}

int32_t z2d_gpu_batch_render(z2d_gpu_batch *b, uint32_t mode) {
  if (!b || mode > 1) return Z2D_ERR_INVALID_ARGUMENT;
  return batch_render(b, mode);
}

int32_t z2d_gpu_batch_wait(z2d_gpu_batch *b) {
  if (!b) return Z2D_ERR_INVALID_ARGUMENT;
  return batch_wait(b);
}

int32_t z2d_gpu_composite_timing(z2d_gpu_ctx *ctx, double *kernel_ms) {
  if (!ctx || !kernel_ms) return Z2D_ERR_INVALID_ARGUMENT;
  *kernel_ms = ctx->last_comp_ms;
  return Z2D_OK;
}

int32_t z2d_gpu_batch_set_queue_order(z2d_gpu_batch *b, uint32_t longest_first) {
  if (!b || longest_first > 1) return Z2D_ERR_INVALID_ARGUMENT;
  if (longest_first != b->longest_first) {
    b->longest_first = longest_first;
    b->prepared = false;
  }
  return Z2D_OK;
}

int32_t z2d_gpu_batch_set_band_height(z2d_gpu_batch *b, uint32_t rows) {
  if (!b || rows < (uint32_t)ROWS_PER_THREAD || rows > (uint32_t)MAX_BAND_H ||
      rows % ROWS_PER_THREAD)
    return Z2D_ERR_INVALID_ARGUMENT;
  if (rows != b->band_h) {
    b->band_h = rows;
    b->prepared = false;
  }
  return Z2D_OK;
}

// The autotuner: renders the batch once per candidate (band height 8, 16, 32
// with the natural queue order, then each with the longest-first order; a
// warm-up and `reps` timed pipelines each), keeps the fastest and prepares
// the batch for it.  The surfaces are saved before and restored after every
// candidate, so the batch's targets end as they started.
int32_t z2d_gpu_batch_autotune(z2d_gpu_batch *b, uint32_t mode, uint32_t reps,
                               uint32_t *band_h, uint32_t *longest_first, double ms_out[6]) {
  if (!b || reps == 0 || mode > 1) return Z2D_ERR_INVALID_ARGUMENT;
  z2d_gpu_ctx *ctx = b->ctx;
  HIP_OK(ctx, hipSetDevice(ctx->device));
  const uint32_t cand[3] = {8, 16, 32};
  size_t total = 0;
  for (const z2d_gpu_surface *s : b->surfaces) total += (s->nbytes + 255) & ~(size_t)255;
  DevBuf saved;
  hipError_t e = saved.ensure(total ? total : 1);
  if (e != hipSuccess) return fail(ctx, e, "autotune save buffer");
  const hipStream_t st = bstream(b);
  auto copy_all = [&](bool save) {
    size_t off = 0;
    hipError_t r = hipSuccess;
    for (const z2d_gpu_surface *s : b->surfaces) {
      if (r == hipSuccess)
        r = save ? hipMemcpyAsync(saved.p + off, s->data, s->nbytes, hipMemcpyDeviceToDevice, st)
                 : hipMemcpyAsync(s->data, saved.p + off, s->nbytes, hipMemcpyDeviceToDevice, st);
      off += (s->nbytes + 255) & ~(size_t)255;
    }
    return r;
  };
  e = copy_all(true);
  double best = 1e30;
  uint32_t best_h = b->band_h, best_o = b->longest_first;
  int32_t rc = Z2D_OK;
  for (int c = 0; c < 6 && e == hipSuccess && rc == Z2D_OK; c++) {
    b->band_h = cand[c % 3];
    b->longest_first = c >= 3 ? 1u : 0u;
    b->prepared = false;
    double sum = 0.0;
    for (uint32_t r = 0; r <= reps && rc == Z2D_OK; r++) {
      rc = batch_render(b, mode);
      if (rc == Z2D_OK) rc = batch_wait(b);
      if (r > 0) sum += b->last_pipeline_ms;
    }
    if (rc == Z2D_OK) e = copy_all(false);
    const double ms = sum / reps;
    if (ms_out) ms_out[c] = ms;
    if (rc == Z2D_OK && ms < best) {
      best = ms;
      best_h = cand[c % 3];
      best_o = b->longest_first;
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  saved.release();
  if (e != hipSuccess) return fail(ctx, e, "autotune");
  if (rc) return rc;
  b->band_h = best_h;
  b->longest_first = best_o;
  b->prepared = false;
  if (band_h) *band_h = best_h;
  if (longest_first) *longest_first = best_o;
  return batch_prepare(b);
}

int32_t z2d_gpu_batch_use_graph(z2d_gpu_batch *b, uint32_t enable) {
  if (!b) return Z2D_ERR_INVALID_ARGUMENT;
  b->use_graph = enable != 0;
  if (!b->use_graph) batch_drop_graph(b);
  return Z2D_OK;
}

int32_t z2d_gpu_batch_timing(z2d_gpu_batch *b, double *pipeline_ms, double *render_ms) {
  if (!b) return Z2D_ERR_INVALID_ARGUMENT;
  if (pipeline_ms) *pipeline_ms = b->last_pipeline_ms;
  if (render_ms) *render_ms = b->last_render_ms;
  return Z2D_OK;
}

int32_t z2d_gpu_batch_stats(z2d_gpu_batch *b, uint64_t out[4]) {
  if (!b || !out) return Z2D_ERR_INVALID_ARGUMENT;
  out[0] = b->draws.size();
  out[1] = b->surfaces.size();
  out[2] = b->n_nodes;
  out[3] = b->surf_tile_off.empty() ? 0 : b->surf_tile_off.back();
  return Z2D_OK;
}

int32_t z2d_gpu_batch_counters(z2d_gpu_batch *b, uint32_t enable, uint64_t out[8]) {
  if (!b) return Z2D_ERR_INVALID_ARGUMENT;
  b->count_stats = enable ? 1u : 0u;
  b->view.count_stats = b->count_stats;
  if (out) {
    for (int i = 0; i < 8; i++) out[i] = 0;
    if (b->prepared && b->ctr_host.p) {
      HIP_OK(b->ctx, hipStreamSynchronize(bstream(b)));
      const uint32_t *c = reinterpret_cast<const uint32_t *>(b->ctr_host.p);
      out[0] = (uint64_t)c[4] | ((uint64_t)c[5] << 32);
      out[1] = c[6];
      out[2] = c[7];
      out[3] = c[2];
      out[4] = c[8];
      out[5] = b->view.max_entry_edges;
    }
  }
  return Z2D_OK;
}

int32_t z2d_gpu_batch_reset(z2d_gpu_batch *b) {
  if (!b) return Z2D_ERR_INVALID_ARGUMENT;
  (void)hipStreamSynchronize(bstream(b));
  batch_clear(b);
  return Z2D_OK;
}

void z2d_gpu_batch_destroy(z2d_gpu_batch *b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_drop_graph(b);
  b->arena.release();
  b->stage.release();
  for (int i = 0; i < 3; i++)
    if (b->ev[i]) (void)hipEventDestroy(b->ev[i]);
  delete b;
}

// One painter call on a descriptor through a single-draw batch.
static int32_t single_draw(z2d_gpu_ctx *ctx, z2d_surface_desc *surface,
                           int32_t (*add)(z2d_gpu_batch *, z2d_gpu_surface *, const void *),
                           const void *arg) {
  SurfRef ref;
  int32_t rc = surface_from_desc(ctx, surface, ref);
  if (rc) return rc;
  z2d_gpu_batch b;
  b.ctx = ctx;
  rc = add(&b, ref.s, arg);
  if (rc == Z2D_OK && !b.draws.empty()) {
    rc = batch_prepare(&b);
    if (rc == Z2D_OK) rc = batch_render(&b, 0);
    if (rc == Z2D_OK) rc = batch_wait(&b);
  }
  b.arena.release();
  b.stage.release();
  for (int i = 0; i < 3; i++)
    if (b.ev[i]) (void)hipEventDestroy(b.ev[i]);
  int32_t rc2 = surface_release(ctx, ref, rc == Z2D_OK);
  return rc ? rc : rc2;
}

struct FillArgs {
  const z2d_pattern_desc *pattern;
  const z2d_path_node *nodes;
  uint64_t n;
  const z2d_fill_opts *opts;
};
struct PolyArgs {
  const z2d_pattern_desc *pattern;
  const z2d_edge *edges;
  uint64_t n;
  const double *ext;
  uint32_t aa, rule, op, precision;
};

int32_t z2d_gpu_fill(z2d_gpu_ctx *ctx, z2d_surface_desc *surface, const z2d_pattern_desc *pattern,
                     const z2d_path_node *nodes, uint64_t n_nodes, const z2d_fill_opts *opts) {
  if (!opts || !surface) return Z2D_ERR_INVALID_ARGUMENT;
  // Host-side checks first, in the reference's order, so they also work
  // without a device.
  int32_t rc = check_pattern(pattern);
  if (rc) return rc;
  if (n_nodes == 0) return Z2D_OK;
  rc = fill_checks(nodes, n_nodes, opts);  // closed set, then the node walk's errors
  if (rc) return rc;
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  FillArgs a{pattern, nodes, n_nodes, opts};
  return single_draw(
      ctx, surface,
      [](z2d_gpu_batch *b, z2d_gpu_surface *s, const void *p) -> int32_t {
        const FillArgs &f = *static_cast<const FillArgs *>(p);
        return batch_fill(b, s, f.pattern, f.nodes, f.n, f.opts);
      },
      &a);
}

int32_t z2d_gpu_fill_polygon(z2d_gpu_ctx *ctx, z2d_surface_desc *surface,
                             const z2d_pattern_desc *pattern, const z2d_edge *edges,
                             uint64_t n_edges, const double extents[4], uint32_t aa,
                             uint32_t rule, uint32_t op, uint32_t precision) {
  if (!surface || !extents || (n_edges && !edges)) return Z2D_ERR_INVALID_ARGUMENT;
  int32_t rc = check_pattern(pattern);
  if (rc) return rc;
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  PolyArgs a{pattern, edges, n_edges, extents, aa, rule, op, precision};
  return single_draw(
      ctx, surface,
      [](z2d_gpu_batch *b, z2d_gpu_surface *s, const void *p) -> int32_t {
        const PolyArgs &f = *static_cast<const PolyArgs *>(p);
        return batch_fill_polygon(b, s, f.pattern, f.edges, f.n, f.ext, f.aa, f.rule, f.op,
                                  f.precision, false);
      },
      &a);
}

int32_t z2d_gpu_composite(z2d_gpu_ctx *ctx, z2d_surface_desc *dst, int32_t dst_x, int32_t dst_y,
                          const z2d_comp_op *ops, uint64_t n_ops, uint32_t precision,
                          z2d_surface_desc *out) {
  return composite_impl(ctx, dst, dst_x, dst_y, ops, n_ops, precision, out);
}

int32_t z2d_gpu_composite_noalloc(z2d_gpu_ctx *ctx, z2d_surface_desc *dst, int32_t dst_x,
                                  int32_t dst_y, const z2d_comp_op *ops, uint64_t n_ops,
                                  uint32_t precision, z2d_surface_desc *out) {
  return composite_impl(ctx, dst, dst_x, dst_y, ops, n_ops, precision, out, true);
}

uint64_t z2d_gpu_composite_noalloc_max_bytes(void) { return kCompScratchBytes; }

// ---------------------------------------------------------------------------
// PNG export (src/export_png.zig:56-360)
// ---------------------------------------------------------------------------
struct PngShape {
  uint64_t row_bytes = 0, raw_len = 0, n_seg = 0, max_zlib = 0, max_idat = 0, bound = 0;
  uint32_t head_len = 0, stream_bpp = 1;
};

static int32_t png_shape(uint32_t format, int32_t w, int32_t h, uint32_t profile, PngShape &o) {
  if (w < 1) return Z2D_ERR_INVALID_WIDTH;
  if (h < 1) return Z2D_ERR_INVALID_HEIGHT;
  if (format > Z2D_FORMAT_ALPHA1 || profile > Z2D_PNG_PROFILE_SRGB) return Z2D_ERR_INVALID_ARGUMENT;
  const uint64_t W = (uint64_t)w;
  switch (format) {
    case Z2D_FORMAT_XRGB:
    case Z2D_FORMAT_RGB:
      o.row_bytes = W * 3;
      o.stream_bpp = 3;
      break;
    case Z2D_FORMAT_ARGB:
    case Z2D_FORMAT_RGBA:
      o.row_bytes = W * 4;
      o.stream_bpp = 4;
      break;
    case Z2D_FORMAT_ALPHA8: o.row_bytes = W; break;
    case Z2D_FORMAT_ALPHA4: o.row_bytes = (W * 4 + 7) / 8; break;
    case Z2D_FORMAT_ALPHA2: o.row_bytes = (W * 2 + 7) / 8; break;
    default: o.row_bytes = (W + 7) / 8; break;
  }
  o.raw_len = (uint64_t)h * (o.row_bytes + 1);
  o.n_seg = (o.raw_len + PNG_SEG - 1) / PNG_SEG;
  if (o.n_seg > 0x7FFFFFFFull) return Z2D_ERR_INVALID_ARGUMENT;
  o.head_len = 8 + 25 + (profile != Z2D_PNG_PROFILE_NONE ? 16 : 0);
  o.max_zlib = 2 + o.raw_len + 5 * o.n_seg + 4;  // every segment stored
  o.max_idat = (o.max_zlib + PNG_IDAT - 1) / PNG_IDAT;
  o.bound = o.head_len + o.max_zlib + 12 * o.max_idat + 12;
  return Z2D_OK;
}

// LZ77 candidate distances for a scanline stream with pixels of `bpp` bytes
// and rows of `row` bytes (filter byte included): the previous pixel, the
// pixel above and its neighbours, then longer pixel and row periods (dither
// tiles repeat every 8 and 64 pixels).
static uint32_t png_candidates(uint64_t bpp, uint64_t row, uint32_t *out) {
  const uint64_t want[PNG_MAX_CAND] = {bpp, row, row - bpp, row + bpp, 2 * bpp, 2 * row,
                                       8 * bpp, 64 * bpp};
  uint32_t n = 0;
  for (uint64_t d : want) {
    if (d < 1 || d > 32768) continue;
    bool dup = false;
    for (uint32_t i = 0; i < n; i++) dup |= out[i] == d;
    if (!dup) out[n++] = (uint32_t)d;
  }
  return n;
}

// Enqueues the encode of a device surface on `st`: work buffers from `work`
// (grown if needed), the file at `out` (capacity `out_cap` >= the bound) and
// meta[0..3] (file bytes, zlib bytes, Adler-32, overflow flag) at `meta`.
static hipError_t png_enqueue(const z2d_gpu_surface &s, uint32_t profile, const PngShape &sh,
                              DevBuf &work, uint8_t *out, uint64_t out_cap, uint64_t *meta,
                              hipStream_t st) {
  Layout L;
  const size_t o_raw = L.take(sh.raw_len + PNG_RAW_PAD);
  const size_t o_seg = L.take(sh.n_seg * PNG_SEG_CAP);
  const size_t o_len = L.take(sh.n_seg * 4);
  const size_t o_sum = L.take(sh.n_seg * 8);
  const size_t o_off = L.take((sh.n_seg + 1) * 8);
  hipError_t e = work.ensure(L.off);
  if (e != hipSuccess) return e;
  uint8_t *D = work.p;
  PngJob J{};
  J.surface = s.data;
  J.format = s.format;
  J.profile = profile;
  J.width = s.w;
  J.height = s.h;
  J.row_bytes = sh.row_bytes;
  J.raw_len = sh.raw_len;
  J.raw = D + o_raw;
  J.n_seg = (uint32_t)sh.n_seg;
  J.n_cand = png_candidates(sh.stream_bpp, sh.row_bytes + 1, J.cand);
  J.seg = D + o_seg;
  J.seg_len = reinterpret_cast<uint32_t *>(D + o_len);
  J.seg_sum = reinterpret_cast<uint32_t *>(D + o_sum);
  J.seg_off = reinterpret_cast<uint64_t *>(D + o_off);
  J.meta = meta;
  J.out = out;
  J.out_cap = out_cap;
  J.head_len = sh.head_len;
  J.max_idat = (uint32_t)sh.max_idat;
  return z2d_launch_png(J, st);
}

int32_t z2d_gpu_png_bound(uint32_t format, int32_t width, int32_t height, uint32_t profile,
                          uint64_t *bound) {
  if (!bound) return Z2D_ERR_INVALID_ARGUMENT;
  PngShape sh;
  int32_t rc = png_shape(format, width, height, profile, sh);
  if (rc) return rc;
  *bound = sh.bound;
  return Z2D_OK;
}

int32_t z2d_gpu_png_encode(z2d_gpu_ctx *ctx, z2d_surface_desc *surface, uint32_t profile,
                           void *out, uint64_t out_cap, uint64_t *out_len) {
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  if (!surface || !out_len || (!out && out_cap)) return Z2D_ERR_INVALID_ARGUMENT;
  PngShape sh;
  int32_t rc = png_shape(surface->format, surface->width, surface->height, profile, sh);
  if (rc) return rc;
  HIP_OK(ctx, hipSetDevice(ctx->device));
  SurfRef ref;
  rc = surface_from_desc(ctx, surface, ref);
  if (rc) return rc;
  hipError_t e = ctx->png_out.ensure(sh.bound + 256);
  if (e == hipSuccess) e = ctx->png_meta.ensure(32);
  for (int i = 0; i < 2 && e == hipSuccess; i++)
    if (!ctx->png_ev[i]) e = hipEventCreate(&ctx->png_ev[i]);
  if (e != hipSuccess) {
    surface_release(ctx, ref, false);
    return fail(ctx, e, "png buffers");
  }
  uint64_t *dmeta = reinterpret_cast<uint64_t *>(ctx->png_out.p);  // meta, then the file
  uint8_t *dfile = ctx->png_out.p + 256;
  e = hipEventRecord(ctx->png_ev[0], ctx->stream);
  if (e == hipSuccess)
    e = png_enqueue(*ref.s, profile, sh, ctx->png_work, dfile, sh.bound, dmeta, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ctx->png_ev[1], ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ctx->png_meta.p, dmeta, 32, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    surface_release(ctx, ref, false);
    return fail(ctx, e, "png encode");
  }
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, ctx->png_ev[0], ctx->png_ev[1]) == hipSuccess) ctx->last_png_ms = ms;
  const uint64_t *meta = reinterpret_cast<const uint64_t *>(ctx->png_meta.p);
  const uint64_t file_len = meta[0];
  *out_len = file_len;
  if (meta[3] || file_len > sh.bound) {  // cannot happen: the buffer holds the stored bound
    surface_release(ctx, ref, false);
    ctx->err = "png encode: output exceeds its bound";
    return Z2D_ERR_DEVICE;
  }
  if (file_len > out_cap) {
    surface_release(ctx, ref, false);
    ctx->err = "png encode: out_cap smaller than the encoded size (returned in out_len)";
    return Z2D_ERR_INVALID_ARGUMENT;
  }
  e = hipMemcpyAsync(out, dfile, file_len, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  rc = surface_release(ctx, ref, false);
  if (e != hipSuccess) return fail(ctx, e, "png download");
  return rc;
}

int32_t z2d_gpu_png_timing(z2d_gpu_ctx *ctx, double *kernel_ms) {
  if (!ctx || !kernel_ms) return Z2D_ERR_INVALID_ARGUMENT;
  *kernel_ms = ctx->last_png_ms;
  return Z2D_OK;
}

// The reference stream over `bytes`-sized buffers: a copy, or (read2) two
// reads and one write; GB/s counts every byte read and written.  The passes
// rotate over three independent buffer sets, so a pass does not find its
// inputs in the last-level cache (as the compositor sweep rotates its
// inputs).
static int32_t stream_rate(z2d_gpu_ctx *ctx, uint64_t bytes, uint32_t iters, bool read2,
                           double *gbps) {
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  if (bytes < 16 || iters == 0) return Z2D_ERR_INVALID_ARGUMENT;
  bytes &= ~15ull;
  HIP_OK(ctx, hipSetDevice(ctx->device));
  constexpr int kSets = 3;
  DevBuf a[kSets], b[kSets], c[kSets];
  hipError_t e = hipSuccess;
  for (int k = 0; k < kSets && e == hipSuccess; k++) {
    e = a[k].ensure(bytes, true);
    if (e == hipSuccess) e = b[k].ensure(bytes, true);
    if (e == hipSuccess && read2) e = c[k].ensure(bytes, true);
  }
  if (e != hipSuccess) return fail(ctx, e, "stream alloc");
  auto once = [&](uint32_t i) {
    const int k = (int)(i % kSets);
    return read2 ? z2d_launch_stream_read2(a[k].p, c[k].p, b[k].p, bytes, ctx->stream)
                 : z2d_launch_stream_copy(a[k].p, b[k].p, bytes, ctx->stream);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  for (int k = 0; k < kSets && e == hipSuccess; k++) {
    e = hipMemsetAsync(a[k].p, 1, bytes, ctx->stream);
    if (e == hipSuccess && read2) e = hipMemsetAsync(c[k].p, 2, bytes, ctx->stream);
  }
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (uint32_t i = 0; i < kSets && e == hipSuccess; i++) e = once(i);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
  for (uint32_t i = 0; i < iters && e == hipSuccess; i++) e = once(i);
  if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess && gbps)
    *gbps = (read2 ? 3.0 : 2.0) * (double)bytes * iters / (ms * 1e-3) / 1e9;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(ctx, e, "stream");
  return Z2D_OK;
}

int32_t z2d_gpu_stream_copy(z2d_gpu_ctx *ctx, uint64_t bytes, uint32_t iters, double *gbps) {
  return stream_rate(ctx, bytes, iters, false, gbps);
}

int32_t z2d_gpu_stream_read2(z2d_gpu_ctx *ctx, uint64_t bytes, uint32_t iters, double *gbps) {
  return stream_rate(ctx, bytes, iters, true, gbps);
}

struct StrokeArgs {
  const z2d_pattern_desc *pattern;
  const z2d_path_node *nodes;
  uint64_t n;
  const z2d_stroke_opts *opts;
};

int32_t z2d_gpu_stroke(z2d_gpu_ctx *ctx, z2d_surface_desc *surface,
                       const z2d_pattern_desc *pattern, const z2d_path_node *nodes,
                       uint64_t n_nodes, const z2d_stroke_opts *opts) {
  if (!opts || !surface) return Z2D_ERR_INVALID_ARGUMENT;
  // painter.stroke's checks, in order (src/painter.zig:221-234)
  int32_t rc = check_pattern(pattern);
  if (rc) return rc;
  if (!ctm_invertible(opts->transformation)) return Z2D_ERR_INVALID_MATRIX;
  if (n_nodes == 0) return Z2D_OK;
  if (!nodes) return Z2D_ERR_INVALID_ARGUMENT;
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  StrokeArgs a{pattern, nodes, n_nodes, opts};
  return single_draw(
      ctx, surface,
      [](z2d_gpu_batch *b, z2d_gpu_surface *s, const void *p) -> int32_t {
        const StrokeArgs &f = *static_cast<const StrokeArgs *>(p);
        return batch_stroke(b, s, f.pattern, f.nodes, f.n, f.opts);
      },
      &a);
}

int32_t z2d_stroke_polygon(const z2d_path_node *nodes, uint64_t n_nodes,
                           const z2d_stroke_opts *opts, uint32_t surface_format, z2d_edge *edges,
                           uint64_t capacity, uint64_t *n_edges, double extents[4],
                           uint32_t *aa_mode) {
  if (!opts || !n_edges || (n_nodes && !nodes) || surface_format > Z2D_FORMAT_ALPHA1)
    return Z2D_ERR_INVALID_ARGUMENT;
  if (!ctm_invertible(opts->transformation)) return Z2D_ERR_INVALID_MATRIX;
  if (opts->hairline) return Z2D_ERR_UNSUPPORTED;
  if (opts->n_dashes && !opts->dashes) return Z2D_ERR_INVALID_ARGUMENT;
  uint32_t aa = 0;
  StrokePolygon poly;
  int32_t rc = n_nodes ? stroke_polygon(nodes, n_nodes, opts, surface_format, aa, poly) : 0;
  if (rc) return rc;
  *n_edges = poly.edges.size();
  if (edges)
    std::memcpy(edges, poly.edges.data(),
                std::min<uint64_t>(capacity, poly.edges.size()) * sizeof(z2d_edge));
  if (extents)
    for (int i = 0; i < 4; i++) extents[i] = poly.ext[i];
  if (aa_mode) *aa_mode = aa;
  return Z2D_OK;
}

int32_t z2d_gpu_batch_stroke(z2d_gpu_batch *b, z2d_gpu_surface *surface,
                             const z2d_pattern_desc *pattern, const z2d_path_node *nodes,
                             uint64_t n_nodes, const z2d_stroke_opts *opts) {
  return batch_stroke(b, surface, pattern, nodes, n_nodes, opts);
}

int32_t z2d_gpu_batch_draw_n(z2d_gpu_batch *b, z2d_gpu_surface *const *surfaces,
                             const z2d_draw_item *items, uint64_t n, uint32_t threads,
                             int32_t *status) {
  if (!b || (n && (!surfaces || !items))) return Z2D_ERR_INVALID_ARGUMENT;
  // the strokes' checks and polygons, in parallel (pure host work)
  std::vector<StrokeDraw> sd(n);
  std::vector<int32_t> src(n, Z2D_OK);
  std::vector<uint64_t> todo;
  for (uint64_t i = 0; i < n; i++)
    if (items[i].kind == Z2D_DRAW_STROKE) todo.push_back(i);
  auto one = [&](uint64_t i) {
    const z2d_draw_item &it = items[i];
    if (!surfaces[i] || !it.stroke) {
      src[i] = Z2D_ERR_INVALID_ARGUMENT;
      return;
    }
    int32_t rc = check_pattern(it.pattern);
    if (rc == Z2D_OK) rc = stroke_draw(it.nodes, it.n_nodes, it.stroke, surfaces[i]->format, sd[i]);
    src[i] = rc;
  };
  unsigned nt = threads;
  if (nt == 0) {
    const unsigned hc = std::thread::hardware_concurrency();
    nt = hc ? std::min(hc, 16u) : 4u;
  }
  nt = (unsigned)std::min<uint64_t>(nt, todo.size());
  if (nt <= 1) {
    for (uint64_t i : todo) one(i);
  } else {
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++)
      pool.emplace_back([&] {
        for (uint64_t k; (k = next.fetch_add(1)) < todo.size();) one(todo[k]);
      });
    for (std::thread &t : pool) t.join();
  }
  // the draws appended in order
  int32_t first = Z2D_OK;
  for (uint64_t i = 0; i < n; i++) {
    const z2d_draw_item &it = items[i];
    int32_t rc;
    if (it.kind == Z2D_DRAW_STROKE)
      rc = src[i] != Z2D_OK ? src[i]
                            : batch_add_stroke(b, surfaces[i], it.pattern, sd[i], it.stroke);
    else if (it.kind == Z2D_DRAW_FILL)
      rc = batch_fill(b, surfaces[i], it.pattern, it.nodes, it.n_nodes, &it.fill);
    else
      rc = Z2D_ERR_INVALID_ARGUMENT;
    if (status) status[i] = rc;
    if (rc != Z2D_OK && first == Z2D_OK) first = rc;
  }
  return first;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Serving: a request loop with continuous batching under an HBM budget.
// Requests (one surface job each) are copied at submit, admitted FIFO while
// their surface + clip mask fit the budget, and rendered in batches on two
// streams: while one batch runs on the device the next is admitted, built
// and launched on the other, so work from requests that arrive while others
// are in flight overlaps on the device.  Surfaces live in one pre-allocated
// HBM region managed by a first-fit allocator.
// ---------------------------------------------------------------------------
namespace {

struct HbmPool {
  uint8_t *base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_blocks;  // offset -> size
  hipError_t init(size_t n) {
    size = (n + 255) & ~(size_t)255;
    hipError_t e = hipMalloc(&base, size);
    if (e == hipSuccess) free_blocks[0] = size;
    return e;
  }
  uint8_t *alloc(size_t n) {
    n = (n + 255) & ~(size_t)255;
    for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
      if (it->second >= n) {
        size_t off = it->first, rem = it->second - n;
        free_blocks.erase(it);
        if (rem) free_blocks[off + n] = rem;
        return base + off;
      }
    }
    return nullptr;
  }
  void release(uint8_t *p, size_t n) {
    n = (n + 255) & ~(size_t)255;
    size_t off = (size_t)(p - base);
    auto it = free_blocks.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_blocks.end() && off + it->second == nx->first) {
      it->second += nx->second;
      free_blocks.erase(nx);
    }
    if (it != free_blocks.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_blocks.erase(it);
      }
    }
  }
  void destroy() {
    if (base) (void)hipFree(base);
    base = nullptr;
  }
};

// One draw of a request, copied at submit: a fill (opts) or a stroke (sopts
// with its own copy of the dash array).
struct SrvFill {
  uint32_t kind = Z2D_DRAW_FILL;
  z2d_pattern_desc pat;
  z2d_gradient_desc grad;
  std::vector<z2d_stop> stops;
  std::vector<z2d_path_node> nodes;
  z2d_fill_opts opts;
  z2d_stroke_opts sopts;
  std::vector<double> dashes;
  // host work, done by the server's preparation threads: a fill's checks
  // (its flattening runs on the device), a stroke's polygon
  int32_t geo_rc = Z2D_OK;
  StrokeDraw stroke;
};

struct SrvRequest {
  uint64_t id = 0;
  uint32_t format = 0;
  int32_t w = 0, h = 0;
  z2d_pixel clear{};
  std::vector<uint8_t> mask;
  std::vector<std::unique_ptr<SrvFill>> fills;
  void *out = nullptr;
  uint64_t out_len = 0;
  uint32_t encode = 0;  // 0 raw bytes, else 1 + PNG profile
  uint64_t *out_written = nullptr;
  PngShape png;
  size_t surf_bytes = 0, mask_bytes = 0;
  z2d_gpu_surface surf;
  uint8_t *mask_dev = nullptr;
  int32_t status = Z2D_OK;
  bool done = false;
  double t_queued = 0;  // when it became ready for admission (steady clock, s)
  double t_submit = 0, t_prep0 = 0, t_taken = 0;  // submitted, prep started, taken into a batch
  double t_done = 0;  // completed
};

struct SrvSlot {
  z2d_gpu_batch batch;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<SrvRequest *> reqs;
  // inputs: the batch's copy commands (surface clears and clip-mask scatters
  // before the render, result gathers after it) and its clip masks, one
  // upload; results: gathered on the device, one read-back
  HostBuf in_stage;
  DevBuf in_dev;
  DevBuf result_dev;
  HostBuf result_stage;
  std::vector<size_t> stage_off;
  std::vector<CopyCmd> pre_cmds, post_cmds;
  DevBuf png_work, png_files;  // PNG encodes of the batch: work, then meta + file
  std::vector<size_t> png_off;
  // free -> running (launched) -> copying (results on their way to the
  // callers) -> free; changed under the server's mutex
  enum State { FREE, RUNNING, COPYING } state = FREE;
  uint32_t pending = 0;  // results still being copied to callers (under the server's mutex)
};

// A finished request's result on its way from the pinned stage to the
// caller's buffer (done by the server's copy threads).
struct SrvCopy {
  SrvRequest *r;
  SrvSlot *slot;
  const uint8_t *src;
  uint64_t n;
};

}  // namespace

struct z2d_gpu_server {
  z2d_gpu_ctx *ctx = nullptr;
  HbmPool pool;
  GradCache gcache;  // gradients resident across requests
  // batches whose device work has finished, counted by a host function
  // enqueued behind each batch (wakes the workers instead of polling)
  uint64_t gpu_completions = 0;
  uint64_t budget = 0;
  uint32_t max_batch = 64;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<SrvRequest *> queue;
  std::unordered_map<uint64_t, SrvRequest *> reqs;
  std::deque<uint64_t> completed;  // completion order, for wait_any
  uint64_t next_id = 1;
  uint64_t used = 0, peak = 0;
  uint64_t n_submitted = 0, n_completed = 0, n_rejected = 0, n_batches = 0, pixel_draws = 0;
  uint32_t n_running = 0;  // batches launched and not yet finalized (under the mutex)
  // Continuous batching: a worker with a free slot launches at once while
  // fewer than kBusyBatches batches are on the device (a light load is
  // answered immediately); beyond that the device is busy anyway, so it
  // waits for a full batch, for the device to drain below that count, or
  // until the oldest queued request has waited kMaxHoldS: a saturated
  // server runs few large batches.
  static constexpr uint32_t kBusyBatches = 2;
  static constexpr double kMaxHoldS = 4e-3;
  bool stop = false;
  // worker threads, each building and launching batches on its own two
  // slots (own streams), so batch preparation runs in parallel
  static constexpr unsigned kWorkers = 4, kSlotsPerWorker = 2;
  std::vector<std::unique_ptr<SrvSlot>> slots;
  std::vector<std::thread> workers;
  std::string err;
  // preparation threads: the host geometry of submitted requests (flattened
  // fills, stroke polygons) is built in parallel before admission
  std::vector<std::thread> preppers;
  std::deque<SrvRequest *> prep_q;
  std::condition_variable cv_prep;
  bool stop_prep = false;
  // copy threads: results from the pinned stage into callers' buffers, in
  // parallel, off the worker thread
  std::vector<std::thread> copiers;
  std::deque<SrvCopy> copy_q;
  std::condition_variable cv_copy;
  bool stop_copy = false;

  // Marks a request complete (mutex held).
  void complete_locked(SrvRequest *r) {
    const double t = now_s();
    q_prep += r->t_prep0 - r->t_submit;
    q_admit += r->t_taken - r->t_queued;
    q_batch += t - r->t_taken;
    q_n++;
    r->t_done = t;
    r->done = true;
    completed.push_back(r->id);
    n_completed++;
  }
  // A slot whose results have all reached their callers is free (mutex held).
  void slot_copied_locked(SrvSlot &sl) {
    if (--sl.pending == 0) {
      sl.reqs.clear();
      sl.state = SrvSlot::FREE;
      cv_work.notify_all();
    }
  }

  void copy_loop() {
    for (;;) {
      SrvCopy c;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_copy.wait(lk, [&] { return stop_copy || !copy_q.empty(); });
        if (copy_q.empty()) return;  // stopping and drained
        c = copy_q.front();
        copy_q.pop_front();
      }
      const double c0 = now_s();
      if (c.n) std::memcpy(c.r->out, c.src, c.n);
      const double c1 = now_s();
      std::lock_guard<std::mutex> lk(mu);
      t_copythreads += c1 - c0;
      complete_locked(c.r);
      slot_copied_locked(*c.slot);
      cv_done.notify_all();
    }
  }
  // worker-thread phase times in seconds, summed over the workers (under the
  // mutex; reported at destroy when the environment sets Z2D_SERVER_TRACE):
  // batch build, the count pass wait, the rest of prepare, render +
  // result-copy enqueue, finalize, idle
  double t_build = 0, t_count = 0, t_prep = 0, t_launch = 0, t_fin = 0, t_idle = 0;
  double t_prepthreads = 0, t_copythreads = 0;  // busy time of the prep and copy threads
  double t_strokes = 0;  // of which stroke tessellation
  uint64_t n_strokes = 0, n_stroke_edges = 0;
  // summed per request: waiting for a prep thread, waiting for admission
  // once prepared, and from admission to completion
  double q_prep = 0, q_admit = 0, q_batch = 0;
  uint64_t q_n = 0;
  // seconds of CLOCK_MONOTONIC (the clock completion times are reported on)
  static double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
  }

  void prep_loop() {
    for (;;) {
      SrvRequest *r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_prep.wait(lk, [&] { return stop_prep || !prep_q.empty(); });
        if (prep_q.empty()) return;  // stopping and drained
        r = prep_q.front();
        prep_q.pop_front();
      }
      const double p0 = now_s();
      r->t_prep0 = p0;
      double ts = 0;
      uint64_t ns = 0, ne = 0;
      for (auto &f : r->fills) {
        if (f->kind == Z2D_DRAW_STROKE) {
          const double s0 = now_s();
          f->geo_rc = stroke_draw(f->nodes.data(), f->nodes.size(), &f->sopts, r->format,
                                  f->stroke);
          ts += now_s() - s0;
          ns++;
          ne += f->stroke.poly.edges.size();
          // server batches keep the default band height
          if (f->geo_rc == Z2D_OK && !f->stroke.empty) {
            f->stroke.pairs = polygon_pairs(f->stroke, r->format, r->surf.h, TILE_H);
            f->stroke.pairs_bh = TILE_H;
          }
        } else {
          f->geo_rc = fill_checks(f->nodes.data(), f->nodes.size(), &f->opts);
        }
      }
      std::lock_guard<std::mutex> lk(mu);
      r->t_queued = now_s();
      t_prepthreads += r->t_queued - p0;
      t_strokes += ts;
      n_strokes += ns;
      n_stroke_edges += ne;
      queue.push_back(r);
      cv_work.notify_all();
    }
  }

  void finalize(SrvSlot &sl) {
    const double f0 = now_s();
    finalize_impl(sl);
    const double f1 = now_s();
    std::lock_guard<std::mutex> lk(mu);
    t_fin += f1 - f0;
  }
  void finalize_impl(SrvSlot &sl) {
    // the batch's counters reached pinned memory before its done event
    uint64_t pix = 0;
    bool overflow = false;
    if (sl.batch.prepared && sl.batch.ctr_host.p) {
      const uint32_t *c = reinterpret_cast<const uint32_t *>(sl.batch.ctr_host.p);
      pix = (uint64_t)c[4] | ((uint64_t)c[5] << 32);
      overflow = c[2] != 0;
    }
    std::vector<SrvCopy> copies;
    copies.reserve(sl.reqs.size());
    for (size_t i = 0; i < sl.reqs.size(); i++) {
      SrvRequest *r = sl.reqs[i];
      if (overflow && r->status == Z2D_OK) r->status = Z2D_ERR_DEVICE;  // bin overflow
      const uint8_t *res = sl.result_stage.p + sl.stage_off[i];
      SrvCopy c{r, &sl, nullptr, 0};
      if (r->encode && r->status == Z2D_OK) {
        const uint64_t *meta = reinterpret_cast<const uint64_t *>(res);
        const uint64_t flen = meta[0];
        if (meta[3] || flen > r->png.bound) {
          r->status = Z2D_ERR_DEVICE;
        } else {
          if (r->out_written) *r->out_written = flen;
          if (r->out) {
            if (flen <= r->out_len) {
              c.src = res + 256;
              c.n = flen;
            } else {
              r->status = Z2D_ERR_INVALID_ARGUMENT;
            }
          }
        }
      } else if (r->status == Z2D_OK) {
        const uint64_t n = r->out ? std::min<uint64_t>(r->out_len, r->surf.nbytes) : 0;
        if (r->out_written) *r->out_written = n;
        c.src = res;
        c.n = n;
      }
      copies.push_back(c);
    }
    std::lock_guard<std::mutex> lk(mu);
    for (SrvRequest *r : sl.reqs) {  // the device is done with their memory
      pool.release(r->surf.data, r->surf_bytes);
      if (r->mask_dev) pool.release(r->mask_dev, r->mask_bytes);
      used -= r->surf_bytes + r->mask_bytes;
    }
    pixel_draws += pix;
    n_running--;
    sl.state = SrvSlot::COPYING;
    sl.pending = 1;  // released below, after the copies are queued
    for (const SrvCopy &c : copies) {
      if (c.n) {
        sl.pending++;
        copy_q.push_back(c);
      } else {
        complete_locked(c.r);
      }
    }
    slot_copied_locked(sl);
    cv_copy.notify_all();
    cv_done.notify_all();
    cv_work.notify_all();
  }

  // Builds and launches one batch on a free slot (requests already admitted
  // and their device memory reserved): one upload of copy commands and clip
  // masks, one launch for every surface clear and mask scatter, the
  // batch's pipeline, PNG encodes, one launch gathering every result and one
  // read-back into the pinned stage.
  void launch(SrvSlot &sl, std::vector<SrvRequest *> &take) {
    const double l0 = now_s();
    hipStream_t st = sl.stream;
    batch_clear(&sl.batch);
    sl.batch.ctx = ctx;
    sl.batch.stream = st;
    sl.batch.count_stats = 1;
    sl.batch.isolate_failures = true;
    sl.batch.async_prepare = true;  // this slot's buffers wait for its done event
    sl.batch.paints.cache = &gcache;
    std::vector<uint32_t> seq0(take.size() + 1, 0);  // each request's draws: [seq0[i], seq0[i+1])
    size_t stage = 0, files = 0, masks = 0;
    sl.stage_off.clear();
    sl.png_off.clear();
    sl.pre_cmds.clear();
    sl.post_cmds.clear();
    std::vector<size_t> mask_off(take.size(), 0);
    for (size_t ri = 0; ri < take.size(); ri++) {
      SrvRequest *r = take[ri];
      sl.stage_off.push_back(stage);
      sl.png_off.push_back(files);
      const size_t res = r->encode ? 256 + r->png.bound : r->surf.nbytes;
      stage += (res + 255) & ~(size_t)255;
      if (r->encode) files += (256 + r->png.bound + 255) & ~(size_t)255;
      if (r->mask_dev) {
        mask_off[ri] = masks;
        masks += (r->mask.size() + 255) & ~(size_t)255;
      }
    }
    const size_t n_cmd_max = 3 * take.size();
    const size_t o_masks = (n_cmd_max * sizeof(CopyCmd) + 255) & ~(size_t)255;
    const size_t in_bytes = o_masks + masks;
    hipError_t e = sl.result_stage.ensure(stage);
    if (e == hipSuccess) e = sl.result_dev.ensure(stage);
    if (e == hipSuccess) e = sl.in_stage.ensure(in_bytes);
    if (e == hipSuccess) e = sl.in_dev.ensure(in_bytes);
    if (e == hipSuccess && files) e = sl.png_files.ensure(files);
    for (size_t ri = 0; ri < take.size(); ri++) {
      SrvRequest *r = take[ri];
      if (e != hipSuccess) break;
      seq0[ri] = (uint32_t)sl.batch.draws.size();
      // Surface.initPixel: every pixel = T.fromPixel(clear)
      I4 c = pixel_rgba(r->clear);
      uint32_t raw = 0;
      switch (r->format) {
        case Z2D_FORMAT_RGBA: raw = c.r | (c.g << 8) | (c.b << 16) | (c.a << 24); break;
        case Z2D_FORMAT_ARGB: raw = c.b | (c.g << 8) | (c.r << 16) | (c.a << 24); break;
        case Z2D_FORMAT_RGB: raw = c.r | (c.g << 8) | (c.b << 16); break;
        case Z2D_FORMAT_XRGB: raw = c.b | (c.g << 8) | (c.r << 16); break;
        default: {
          int bits = fbits(r->format);
          uint32_t v = (bits == 8) ? c.a : (c.a & 0xFF) >> (8 - bits);
          uint32_t byte = 0;
          if (v)
            for (int sh = 0; sh <= 8 - bits; sh += bits) byte |= v << sh;
          raw = byte | (byte << 8) | (byte << 16) | (byte << 24);
        } break;
      }
      sl.pre_cmds.push_back(CopyCmd{nullptr, r->surf.data, padded_bytes(r->surf.nbytes), raw, 0});
      if (r->mask_dev) {
        std::memcpy(sl.in_stage.p + o_masks + mask_off[ri], r->mask.data(), r->mask.size());
        sl.pre_cmds.push_back(CopyCmd{sl.in_dev.p + o_masks + mask_off[ri], r->mask_dev,
                                      r->mask.size(), 0, 0});
      }
      r->surf.clip = r->mask_dev;
      // a clipped surface takes part in the render even with no fills
      if (r->mask_dev) sl.batch.slot_of(&r->surf);
      for (auto &f : r->fills) {
        int32_t rc = f->geo_rc;
        if (rc == Z2D_OK)
          rc = f->kind == Z2D_DRAW_STROKE
                   ? batch_add_stroke(&sl.batch, &r->surf, &f->pat, f->stroke, &f->sopts, true)
                   : batch_add_fill(&sl.batch, &r->surf, &f->pat, f->nodes.data(),
                                    f->nodes.size(), &f->opts);
        if (rc != Z2D_OK && r->status == Z2D_OK) r->status = rc;
      }
      seq0[ri + 1] = (uint32_t)sl.batch.draws.size();
      // the result: the surface, or the PNG region (meta, then the file at +256)
      const uint8_t *src = r->encode ? sl.png_files.p + sl.png_off[ri] : r->surf.data;
      const uint64_t n = r->encode ? 256 + r->png.bound : r->surf.nbytes;
      sl.post_cmds.push_back(CopyCmd{src, sl.result_dev.p + sl.stage_off[ri], n, 0, 0});
    }
    const size_t n_pre = sl.pre_cmds.size(), n_post = sl.post_cmds.size();
    if (e == hipSuccess) {
      std::memcpy(sl.in_stage.p, sl.pre_cmds.data(), n_pre * sizeof(CopyCmd));
      std::memcpy(sl.in_stage.p + n_pre * sizeof(CopyCmd), sl.post_cmds.data(),
                  n_post * sizeof(CopyCmd));
      e = hipMemcpyAsync(sl.in_dev.p, sl.in_stage.p, in_bytes, hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess)
      e = z2d_launch_multi_copy(reinterpret_cast<const CopyCmd *>(sl.in_dev.p), (uint32_t)n_pre,
                                st);
    const double l1 = now_s();
    int32_t rc = e == hipSuccess ? batch_prepare(&sl.batch) : Z2D_ERR_DEVICE;
    const double l2 = now_s();
    // draws the flattener refused fail their own request only
    for (const auto &fd : sl.batch.failed) {
      const size_t ri =
          (size_t)(std::upper_bound(seq0.begin(), seq0.end(), fd.first) - seq0.begin()) - 1;
      if (ri < take.size() && take[ri]->status == Z2D_OK) take[ri]->status = fd.second;
    }
    if (rc == Z2D_OK) rc = batch_render(&sl.batch, 0);
    for (size_t i = 0; i < take.size() && rc == Z2D_OK; i++) {
      SrvRequest *r = take[i];
      if (!r->encode) continue;
      uint8_t *region = sl.png_files.p + sl.png_off[i];
      if (png_enqueue(r->surf, r->encode - 1, r->png, sl.png_work, region + 256, r->png.bound,
                      reinterpret_cast<uint64_t *>(region), st) != hipSuccess)
        rc = Z2D_ERR_DEVICE;
    }
    if (rc == Z2D_OK &&
        z2d_launch_multi_copy(reinterpret_cast<const CopyCmd *>(sl.in_dev.p) + n_pre,
                              (uint32_t)n_post, st) != hipSuccess)
      rc = Z2D_ERR_DEVICE;
    if (rc == Z2D_OK && stage &&
        hipMemcpyAsync(sl.result_stage.p, sl.result_dev.p, stage, hipMemcpyDeviceToHost, st) !=
            hipSuccess)
      rc = Z2D_ERR_DEVICE;
    if (rc != Z2D_OK)
      for (SrvRequest *r : take)
        if (r->status == Z2D_OK) r->status = rc;
    (void)hipEventRecord(sl.done, st);
    // a wake-up for the workers once the stream passes this point (they
    // still check the done event; without the call they poll)
    (void)hipLaunchHostFunc(st, on_batch_done, this);
    sl.reqs = take;
    const double l3 = now_s();
    std::lock_guard<std::mutex> lk(mu);
    sl.state = SrvSlot::RUNNING;
    n_running++;
    n_batches++;
    t_build += l1 - l0;
    t_count += sl.batch.t_count_wait;
    t_prep += (l2 - l1) - sl.batch.t_count_wait;
    t_launch += l3 - l2;
  }

  static void on_batch_done(void *arg) {
    auto *s = static_cast<z2d_gpu_server *>(arg);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->gpu_completions++;
    }
    s->cv_work.notify_all();
  }

  void run(unsigned w) {
    (void)hipSetDevice(ctx->device);
    SrvSlot *mine[kSlotsPerWorker];
    for (unsigned k = 0; k < kSlotsPerWorker; k++) mine[k] = slots[w * kSlotsPerWorker + k].get();
    for (;;) {
      bool progressed = false;
      for (SrvSlot *slp : mine) {
        SrvSlot &sl = *slp;
        bool running;
        {
          std::lock_guard<std::mutex> lk(mu);
          running = sl.state == SrvSlot::RUNNING;
        }
        if (running && hipEventQuery(sl.done) == hipSuccess) {
          finalize(sl);
          progressed = true;
        }
      }
      SrvSlot *free_slot = nullptr;
      std::vector<SrvRequest *> take;
      {
        std::unique_lock<std::mutex> lk(mu);
        bool all_free = true;
        for (SrvSlot *slp : mine) {
          if (slp->state == SrvSlot::FREE) {
            if (!free_slot) free_slot = slp;
          } else {
            all_free = false;
          }
        }
        if (stop && queue.empty() && all_free) return;
        bool hold = false;  // wait for the batch to fill
        if (free_slot && !queue.empty() && !stop && n_running >= kBusyBatches &&
            queue.size() < max_batch && now_s() - queue.front()->t_queued < kMaxHoldS)
          hold = true;
        if (free_slot && !hold) {
          while (!queue.empty() && take.size() < max_batch) {
            SrvRequest *r = queue.front();
            size_t need = r->surf_bytes + r->mask_bytes;
            if (used + need > budget) break;
            uint8_t *sp = pool.alloc(r->surf_bytes);
            if (!sp) break;
            uint8_t *mp = nullptr;
            if (r->mask_bytes) {
              mp = pool.alloc(r->mask_bytes);
              if (!mp) {
                pool.release(sp, r->surf_bytes);
                break;
              }
            }
            r->surf.data = sp;
            r->mask_dev = mp;
            used += need;
            peak = std::max(peak, used);
            queue.pop_front();
            r->t_taken = now_s();
            take.push_back(r);
          }
        }
        if (take.empty() && !progressed) {
          const double w0 = now_s();
          if (!all_free || hold) {
            const uint64_t seen = gpu_completions;
            const size_t q0 = queue.size();
            cv_work.wait_for(lk, std::chrono::microseconds(hold ? 250 : 2000), [&] {
              return stop || gpu_completions != seen || (free_slot && queue.size() != q0);
            });
          } else
            cv_work.wait_for(lk, std::chrono::milliseconds(5),
                             [&] { return stop || !queue.empty(); });
          t_idle += now_s() - w0;
        }
      }
      if (!take.empty()) launch(*free_slot, take);
    }
  }
};

static void copy_pattern(SrvFill &f, const z2d_pattern_desc *p) {
  f.pat = *p;
  const z2d_gradient_desc *g = nullptr;
  if (p->kind == Z2D_PATTERN_GRADIENT) g = p->gradient;
  if (p->kind == Z2D_PATTERN_DITHER && p->dither.source == Z2D_DITHER_SRC_GRADIENT)
    g = p->dither.gradient;
  if (g) {
    f.grad = *g;
    f.stops.assign(g->stops, g->stops + g->n_stops);
    f.grad.stops = f.stops.data();
    if (p->kind == Z2D_PATTERN_GRADIENT)
      f.pat.gradient = &f.grad;
    else
      f.pat.dither.gradient = &f.grad;
  }
}

extern "C" {

int32_t z2d_gpu_server_create(z2d_gpu_ctx *ctx, uint64_t hbm_budget, uint32_t max_batch,
                              z2d_gpu_server **out) {
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  if (!out || hbm_budget < 4096) return Z2D_ERR_INVALID_ARGUMENT;
  auto *s = new (std::nothrow) z2d_gpu_server();
  if (!s) return Z2D_ERR_OUT_OF_MEMORY;
  s->ctx = ctx;
  s->budget = hbm_budget;
  s->max_batch = max_batch ? max_batch : 64;
  HIP_OK(ctx, hipSetDevice(ctx->device));
  hipError_t e = s->pool.init(hbm_budget + 256 * 1024);
  if (e == hipSuccess && s->gcache.init() != hipSuccess) {
    (void)hipGetLastError();  // the table is optional: batches then pack their own
    s->gcache.destroy();
  }
  for (unsigned i = 0; i < z2d_gpu_server::kWorkers * z2d_gpu_server::kSlotsPerWorker; i++)
    s->slots.push_back(std::make_unique<SrvSlot>());
  for (auto &sl : s->slots) {
    if (e != hipSuccess) break;
    e = hipStreamCreate(&sl->stream);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl->done, hipEventDisableTiming);
    sl->batch.ctx = ctx;
  }
  if (e != hipSuccess) {
    for (auto &sl : s->slots) {
      if (sl->done) (void)hipEventDestroy(sl->done);
      if (sl->stream) (void)hipStreamDestroy(sl->stream);
    }
    s->gcache.destroy();
    s->pool.destroy();
    delete s;
    return fail(ctx, e, "server create");
  }
  for (unsigned w = 0; w < z2d_gpu_server::kWorkers; w++)
    s->workers.emplace_back([s, w] { s->run(w); });
  // preparation threads: three quarters of the hardware threads less one
  // (2 to 12; the workers, copy threads and clients take the rest), or
  // Z2D_SERVER_PREP_THREADS (1 to 64)
  const unsigned hc = std::thread::hardware_concurrency();
  unsigned np = std::max(2u, std::min(12u, hc >= 4 ? 3 * hc / 4 - 1 : 2u));
  if (const char *env = std::getenv("Z2D_SERVER_PREP_THREADS")) {
    const long v = std::strtol(env, nullptr, 10);
    if (v >= 1 && v <= 64) np = (unsigned)v;
  }
  for (unsigned i = 0; i < np; i++) s->preppers.emplace_back([s] { s->prep_loop(); });
  for (unsigned i = 0; i < 4; i++) s->copiers.emplace_back([s] { s->copy_loop(); });
  *out = s;
  return Z2D_OK;
}

}  // extern "C"

// A request's copy for the server (z2d_gpu_server_submit's checks), not yet
// queued.
static int32_t server_request(z2d_gpu_server *s, const z2d_request_desc *rq, void *out,
                              uint64_t out_len, std::unique_ptr<SrvRequest> &res) {
  if (!rq) return Z2D_ERR_INVALID_ARGUMENT;
  if (rq->width < 1) return Z2D_ERR_INVALID_WIDTH;
  if (rq->height < 1) return Z2D_ERR_INVALID_HEIGHT;
  if (rq->format > Z2D_FORMAT_ALPHA1 || (rq->n_fills && !rq->fills) ||
      (rq->n_draws && !rq->draws))
    return Z2D_ERR_INVALID_ARGUMENT;
  auto r = std::make_unique<SrvRequest>();
  r->format = rq->format;
  r->w = rq->width;
  r->h = rq->height;
  r->clear = rq->clear;
  r->out = out;
  r->out_len = out_len;
  r->encode = rq->encode;
  r->out_written = rq->out_written;
  if (rq->encode > 1 + Z2D_PNG_PROFILE_SRGB) return Z2D_ERR_INVALID_ARGUMENT;
  if (rq->encode) {
    int32_t prc = png_shape(rq->format, rq->width, rq->height, rq->encode - 1, r->png);
    if (prc) return prc;
  }
  r->surf.ctx = s->ctx;
  r->surf.format = rq->format;
  r->surf.w = rq->width;
  r->surf.h = rq->height;
  r->surf.nbytes = surface_bytes(rq->format, rq->width, rq->height);
  r->surf.owned = false;
  r->surf_bytes = padded_bytes(r->surf.nbytes);
  if (!rq->encode && out && out_len < r->surf.nbytes) return Z2D_ERR_INVALID_ARGUMENT;
  if (rq->clip_mask) {
    size_t n = (size_t)rq->width * (size_t)rq->height;
    r->mask.assign(rq->clip_mask, rq->clip_mask + n);
    r->mask_bytes = n;
  }
  for (uint64_t i = 0; i < rq->n_fills; i++) {
    const z2d_fill_item &it = rq->fills[i];
    int32_t rc = check_pattern(it.pattern);
    if (rc) return rc;
    auto f = std::make_unique<SrvFill>();
    copy_pattern(*f, it.pattern);
    f->nodes.assign(it.nodes, it.nodes + it.n_nodes);
    f->opts = it.opts;
    r->fills.push_back(std::move(f));
  }
  for (uint64_t i = 0; i < rq->n_draws; i++) {
    const z2d_draw_item &it = rq->draws[i];
    if (it.kind > Z2D_DRAW_STROKE || (it.kind == Z2D_DRAW_STROKE && !it.stroke) ||
        (it.n_nodes && !it.nodes))
      return Z2D_ERR_INVALID_ARGUMENT;
    int32_t rc = check_pattern(it.pattern);
    if (rc) return rc;
    auto f = std::make_unique<SrvFill>();
    copy_pattern(*f, it.pattern);
    f->kind = it.kind;
    f->nodes.assign(it.nodes, it.nodes + it.n_nodes);
    f->opts = it.fill;
    if (it.kind == Z2D_DRAW_STROKE) {
      f->sopts = *it.stroke;
      if (f->sopts.n_dashes && !f->sopts.dashes) return Z2D_ERR_INVALID_ARGUMENT;
      f->dashes.assign(f->sopts.dashes, f->sopts.dashes + f->sopts.n_dashes);
      f->sopts.dashes = f->dashes.data();
    }
    r->fills.push_back(std::move(f));
  }
  res = std::move(r);
  return Z2D_OK;
}

// Queues a built request for preparation (mutex held).
static int32_t server_enqueue_locked(z2d_gpu_server *s, std::unique_ptr<SrvRequest> &r,
                                     double t, uint64_t *id) {
  if (r->surf_bytes + r->mask_bytes > s->budget) {
    s->n_rejected++;
    return Z2D_ERR_BUDGET_EXCEEDED;
  }
  r->id = s->next_id++;
  *id = r->id;
  SrvRequest *raw = r.release();
  raw->t_submit = t;
  s->reqs[raw->id] = raw;
  s->prep_q.push_back(raw);
  s->n_submitted++;
  return Z2D_OK;
}

extern "C" {

int32_t z2d_gpu_server_submit(z2d_gpu_server *s, const z2d_request_desc *rq, void *out,
                              uint64_t out_len, uint64_t *id) {
  if (!s || !rq || !id) return Z2D_ERR_INVALID_ARGUMENT;
  std::unique_ptr<SrvRequest> r;
  int32_t rc = server_request(s, rq, out, out_len, r);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(s->mu);
  rc = server_enqueue_locked(s, r, z2d_gpu_server::now_s(), id);
  if (rc == Z2D_OK) s->cv_prep.notify_one();
  return rc;
}

int32_t z2d_gpu_server_submit_n(z2d_gpu_server *s, const z2d_request_desc *const *reqs,
                                void *const *outs, const uint64_t *out_lens, uint64_t n,
                                uint64_t *ids, int32_t *status) {
  if (!s || (n && (!reqs || !ids))) return Z2D_ERR_INVALID_ARGUMENT;
  std::vector<std::unique_ptr<SrvRequest>> built(n);
  std::vector<int32_t> rcs(n, Z2D_OK);
  for (uint64_t i = 0; i < n; i++)
    rcs[i] = server_request(s, reqs[i], outs ? outs[i] : nullptr, out_lens ? out_lens[i] : 0,
                            built[i]);
  int32_t first = Z2D_OK;
  uint64_t queued = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    const double t = z2d_gpu_server::now_s();
    for (uint64_t i = 0; i < n; i++) {
      ids[i] = 0;
      if (rcs[i] == Z2D_OK) rcs[i] = server_enqueue_locked(s, built[i], t, &ids[i]);
      if (rcs[i] == Z2D_OK) queued++;
      if (status) status[i] = rcs[i];
      if (rcs[i] != Z2D_OK && first == Z2D_OK) first = rcs[i];
    }
  }
  if (queued == 1) s->cv_prep.notify_one();
  else if (queued > 1) s->cv_prep.notify_all();
  return first;
}


// Takes a completed request out of the server (mutex held); `gone` receives
// it, so that its memory is released after the mutex is.
static int32_t take_done(z2d_gpu_server *s, uint64_t id, int32_t *status,
                         std::unique_ptr<SrvRequest> &gone) {
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return Z2D_ERR_INVALID_ARGUMENT;
  if (!it->second->done) return 0;
  if (status) *status = it->second->status;
  gone.reset(it->second);
  s->reqs.erase(it);
  auto c = std::find(s->completed.begin(), s->completed.end(), id);
  if (c != s->completed.end()) s->completed.erase(c);
  return 1;
}

int32_t z2d_gpu_server_poll(z2d_gpu_server *s, uint64_t id, int32_t *status) {
  if (!s) return Z2D_ERR_INVALID_ARGUMENT;
  std::unique_ptr<SrvRequest> gone;  // released after the mutex
  std::lock_guard<std::mutex> lk(s->mu);
  return take_done(s, id, status, gone);
}

int32_t z2d_gpu_server_wait(z2d_gpu_server *s, uint64_t id, int32_t *status, double timeout_ms) {
  if (!s) return Z2D_ERR_INVALID_ARGUMENT;
  std::unique_ptr<SrvRequest> gone;  // released after the mutex
  std::unique_lock<std::mutex> lk(s->mu);
  auto deadline = std::chrono::steady_clock::now() +
                  std::chrono::microseconds((int64_t)(timeout_ms * 1000.0));
  for (;;) {
    int32_t rc = take_done(s, id, status, gone);
    if (rc != 0) return rc;
    if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout)
      return take_done(s, id, status, gone);
  }
}

int32_t z2d_gpu_server_wait_any(z2d_gpu_server *s, uint64_t *id, int32_t *status,
                                double timeout_ms) {
  if (!s || !id) return Z2D_ERR_INVALID_ARGUMENT;
  std::unique_ptr<SrvRequest> gone;  // released after the mutex
  std::unique_lock<std::mutex> lk(s->mu);
  auto deadline = std::chrono::steady_clock::now() +
                  std::chrono::microseconds((int64_t)(timeout_ms * 1000.0));
  for (;;) {
    if (!s->completed.empty()) {
      uint64_t got = s->completed.front();
      *id = got;
      return take_done(s, got, status, gone);
    }
    if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && s->completed.empty())
      return 0;
  }
}

int32_t z2d_gpu_server_wait_some(z2d_gpu_server *s, uint64_t *ids, int32_t *status,
                                 double *t_done, uint32_t max, double timeout_ms) {
  if (!s || !ids || max == 0) return Z2D_ERR_INVALID_ARGUMENT;
  std::vector<std::unique_ptr<SrvRequest>> gone;  // released after the mutex
  std::unique_lock<std::mutex> lk(s->mu);
  auto deadline = std::chrono::steady_clock::now() +
                  std::chrono::microseconds((int64_t)(timeout_ms * 1000.0));
  for (;;) {
    uint32_t k = 0;
    while (k < max && !s->completed.empty()) {
      const uint64_t got = s->completed.front();
      s->completed.pop_front();
      auto it = s->reqs.find(got);
      if (it == s->reqs.end()) continue;
      ids[k] = got;
      if (status) status[k] = it->second->status;
      if (t_done) t_done[k] = it->second->t_done;
      gone.emplace_back(it->second);
      s->reqs.erase(it);
      k++;
    }
    if (k) return (int32_t)k;
    if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && s->completed.empty())
      return 0;
  }
}

int32_t z2d_gpu_server_stats(z2d_gpu_server *s, uint64_t out[8]) {
  if (!s || !out) return Z2D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  out[0] = s->n_submitted;
  out[1] = s->n_completed;
  out[2] = s->n_rejected;
  out[3] = s->n_batches;
  out[4] = s->pixel_draws;
  out[5] = s->used;
  out[6] = s->peak;
  out[7] = s->queue.size() + s->prep_q.size();
  return Z2D_OK;
}

void z2d_gpu_server_destroy(z2d_gpu_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop_prep = true;
    s->cv_prep.notify_all();
  }
  for (std::thread &t : s->preppers)
    if (t.joinable()) t.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  for (std::thread &t : s->workers)
    if (t.joinable()) t.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);  // the worker left every slot free: no copies remain
    s->stop_copy = true;
    s->cv_copy.notify_all();
  }
  for (std::thread &t : s->copiers)
    if (t.joinable()) t.join();
  if (std::getenv("Z2D_SERVER_TRACE"))
    std::fprintf(stderr,
                 "[z2d server] batches %llu: build %.3f s, count-pass wait %.3f s, prepare %.3f "
                 "s, launch %.3f s, finalize %.3f s, idle %.3f s; prep threads %.3f s (%zu), "
                 "copy threads %.3f s (%zu)\n",
                 (unsigned long long)s->n_batches, s->t_build, s->t_count, s->t_prep,
                 s->t_launch, s->t_fin, s->t_idle, s->t_prepthreads, s->preppers.size(),
                 s->t_copythreads, s->copiers.size());
  if (std::getenv("Z2D_SERVER_TRACE"))
    std::fprintf(stderr,
                 "[z2d server] per request (ms): prep queue %.3f, admission queue %.3f, "
                 "batch %.3f; %.1f requests per batch\n",
                 s->q_n ? s->q_prep / s->q_n * 1e3 : 0.0, s->q_n ? s->q_admit / s->q_n * 1e3 : 0.0,
                 s->q_n ? s->q_batch / s->q_n * 1e3 : 0.0,
                 s->n_batches ? (double)s->q_n / (double)s->n_batches : 0.0);
  if (std::getenv("Z2D_SERVER_TRACE"))
    std::fprintf(stderr,
                 "[z2d server] strokes %llu: tessellation %.3f s (%.1f us each, %.0f edges "
                 "each)\n",
                 (unsigned long long)s->n_strokes, s->t_strokes,
                 s->n_strokes ? s->t_strokes / (double)s->n_strokes * 1e6 : 0.0,
                 s->n_strokes ? (double)s->n_stroke_edges / (double)s->n_strokes : 0.0);
  if (std::getenv("Z2D_SERVER_TRACE"))
    std::fprintf(stderr, "[z2d server] resident gradients: %u, hits %llu, misses %llu, full %llu\n",
                 s->gcache.n, (unsigned long long)s->gcache.hits,
                 (unsigned long long)s->gcache.misses, (unsigned long long)s->gcache.full);
  for (auto &slp : s->slots) {
    SrvSlot &sl = *slp;
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    sl.batch.arena.release();
    sl.batch.stage.release();
    for (int i = 0; i < 3; i++)
      if (sl.batch.ev[i]) (void)hipEventDestroy(sl.batch.ev[i]);
    sl.result_stage.release();
    sl.result_dev.release();
    sl.in_stage.release();
    sl.in_dev.release();
    sl.png_work.release();
    sl.png_files.release();
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  for (auto &kv : s->reqs) delete kv.second;
  s->gcache.destroy();
  s->pool.destroy();
  delete s;
}

}  // extern "C"
