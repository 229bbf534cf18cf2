// Host harness for the per-pixel composite of the device renderer
// (composite_cov in z2d_amd/csrc/composite_px.h), built for the CPU with
// Z2D_HOST_PIXEL under the address and undefined-behaviour sanitizers and
// called through a non-inlined wrapper (the "outlined" form).  Every
// operator, precision and format, coverage counts 0..16 (multisample) and
// 0..1 (direct), solid paints with and without full alpha, against the
// oracle: compositeOpaque / compositeOpacity / compositeStride
// (src/internal/raster/shared.zig:9-117, src/surface.zig:557-581) stated as
// SurfaceCompositor.run operation lists on a one-pixel surface
// (zo_composite, src/compositor.zig:302-440).  Test infrastructure only.
#define Z2D_HOST_PIXEL 1
#include "../../z2d_amd/csrc/composite_px.h"

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../oracle/z2d_oracle.h"

using namespace z2d;

__attribute__((noinline)) uint32_t composite_outlined(uint32_t raw, uint32_t fmt, int k, int kfull,
                                                      const DevDraw &D, const DevPaint &P,
                                                      const StopView &sv, const uint16_t *bn,
                                                      int x, int y) {
  return composite_cov(raw, fmt, k, kfull, D, P, sv, bn, x, y);
}

static uint64_t rng_state = 0x5EEDC0DEu;
static uint64_t next_u64() {  // SplitMix64
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static uint32_t rnd(uint32_t n) { return (uint32_t)(next_u64() % n); }

static int fmt_bits_of(uint32_t f) {
  switch (f) {
    case Z2D_FORMAT_ALPHA8: return 8;
    case Z2D_FORMAT_ALPHA4: return 4;
    case Z2D_FORMAT_ALPHA2: return 2;
    case Z2D_FORMAT_ALPHA1: return 1;
    default: return 32;
  }
}

// A raw value valid for the format (premultiplied where the format has
// color and alpha).
static uint32_t random_raw(uint32_t fmt) {
  const int bits = fmt_bits_of(fmt);
  if (bits < 32) return rnd(1u << bits);
  const uint32_t a = rnd(4) == 0 ? 255u : rnd(256);
  uint32_t c[3];
  for (int i = 0; i < 3; i++) c[i] = a ? rnd(a + 1) : 0;
  switch (fmt) {
    case Z2D_FORMAT_RGBA: return c[0] | (c[1] << 8) | (c[2] << 16) | (a << 24);
    case Z2D_FORMAT_ARGB: return c[2] | (c[1] << 8) | (c[0] << 16) | (a << 24);
    default: return rnd(256) | (rnd(256) << 8) | (rnd(256) << 16);  // RGB, XRGB
  }
}

// The oracle's result of `ops` on a one-pixel surface holding raw.
static uint32_t oracle_run(uint32_t fmt, uint32_t raw, const z2d_comp_op *ops, uint64_t n,
                           uint32_t precision) {
  uint8_t buf[8] = {0};
  const int bits = fmt_bits_of(fmt);
  if (bits == 32) std::memcpy(buf, &raw, 4); else buf[0] = (uint8_t)raw;
  z2d_surface_desc s{};
  s.format = fmt;
  s.width = 1;
  s.height = 1;
  s.data = buf;
  s.len = bits == 32 ? 4 : 1;
  zo_composite(&s, 0, 0, ops, n, precision);
  uint32_t out = 0;
  if (bits == 32) std::memcpy(&out, buf, 4); else out = buf[0] & ((1u << bits) - 1u);
  return out;
}

// The fast renderer's RGBA8 forms against the per-channel ones they stand
// for, and the other arithmetic rewrites against the forms they replace.
static uint64_t check_rewrites(uint64_t &cases) {
  uint64_t bad = 0;
  auto unpack8 = [](uint32_t w) { return I4{w & 255u, (w >> 8) & 255u, (w >> 16) & 255u, w >> 24}; };
  // src_over_rgba8: every (destination channel, source channel, source
  // alpha), the channel in every byte position, then random words
  for (uint32_t dc = 0; dc < 256; dc++)
    for (uint32_t sc = 0; sc < 256; sc++)
      for (uint32_t sa = 0; sa < 256; sa++) {
        const uint32_t ds[2] = {dc | (dc << 8) | (dc << 16) | ((255u - dc) << 24),
                                dc | (dc << 8) | (dc << 16) | (dc << 24)};
        const uint32_t sw = sc | (sc << 8) | (sc << 16) | (sa << 24);
        for (uint32_t d : ds) {
          cases++;
          if (src_over_rgba8(d, sw) != composite_src_over(d, Z2D_FORMAT_RGBA, unpack8(sw))) bad++;
        }
      }
  for (int i = 0; i < (1 << 22); i++) {
    const uint64_t r = next_u64();
    const uint32_t d = (uint32_t)r, sw = (uint32_t)(r >> 32);
    const int k = (int)(next_u64() % 17u);
    cases += 2;
    if (src_over_rgba8(d, sw) != composite_src_over(d, Z2D_FORMAT_RGBA, unpack8(sw))) bad++;
    const I4 m = mask_coverage(unpack8(sw), k);
    if (mask_coverage_rgba8(sw, k) != (m.r | (m.g << 8) | (m.b << 16) | (m.a << 24))) bad++;
  }
  // f2u8 (truncf of v + (0.5 - 2^-25), clamped) against roundf
  // with the branches it replaces: every 61st float bit pattern and the
  // neighbourhoods of the halves 0.5 .. 65535.5
  auto f2u8_ref = [](float v) -> uint32_t {
    const float r = std::roundf(v);
    if (!(r >= 0.0f)) return 0u;
    if (r > 65535.0f) return 65535u;
    return (uint32_t)r;
  };
  for (uint64_t u = 0; u <= 0xFFFFFFFFull; u += 61) {
    uint32_t w = (uint32_t)u;
    float v;
    std::memcpy(&v, &w, 4);
    cases++;
    if (f2u8(v) != f2u8_ref(v)) bad++;
  }
  for (uint32_t h = 0; h < 65536; h++) {
    float v = (float)h + 0.5f;
    for (int t = 0; t < 3; t++) {
      const float vs[2] = {v, -v};
      for (float x : vs) {
        cases++;
        if (f2u8(x) != f2u8_ref(x)) bad++;
      }
      v = std::nextafterf(v, t == 0 ? 0.0f : 1e9f);
      if (t == 1) v = std::nextafterf(v, 1e9f);
    }
  }
  // round_i32 (truncation of v + copysign(0.5 - 2^-54, v)) against
  // (int32_t)round(v): integers, halves and quarters at every scale from
  // 2^-60 to 2^30 with their float neighbours, both signs
  for (int e = -60; e <= 30; e++)
    for (int64_t k = 0; k < 1024; k++) {
      const double base = std::ldexp((double)k, e);
      for (double h : {base, base + 0.5, base - 0.5, base + 0.25, base + 0.75}) {
        for (int dir = 0; dir < 2; dir++) {
          double v = h;
          for (int st = 0; st < 4; st++) {
            if (std::fabs(v) < 2.0e9) {
              cases += 2;
              if (round_i32(v) != (int32_t)std::round(v)) bad++;
              if (round_i32(-v) != (int32_t)std::round(-v)) bad++;
            }
            v = std::nextafter(v, dir ? 1e300 : -1e300);
          }
        }
      }
    }
  // overlay / hard light on 8-bit lanes against the signed per-channel forms
  for (int i = 0; i < (1 << 24); i++) {
    const uint64_t r = next_u64();
    const uint32_t sc = r & 255u, dc = (r >> 8) & 255u, sa = (r >> 16) & 255u, da = (r >> 24) & 255u;
    const I4 d{dc, dc, dc, da}, sv{sc, sc, sc, sa};
    cases += 2;
    if (op_int(Z2D_OP_OVERLAY, d, sv).r != overlay_ch((int32_t)sc, (int32_t)dc, (int32_t)sa, (int32_t)da)) bad++;
    if (op_int(Z2D_OP_HARD_LIGHT, d, sv).r != hardlight_ch((int32_t)sc, (int32_t)dc, (int32_t)sa, (int32_t)da)) bad++;
  }
  // stops_find8 against the reference's binary search (stops_search): the
  // interpolated color (both ends' colors at t) must be bit-identical; stop
  // lists of 1..8 sorted offsets in [0, 1] drawn from a small set (ties,
  // repeated stops, 0 and 1), padded with +inf, and offsets at every stop,
  // between stops, below 0, above 1, and at their float neighbours
  for (int i = 0; i < (1 << 18); i++) {
    const uint32_t n = 1u + (uint32_t)(next_u64() % 8u);
    float st[8], col[8];
    const float grid[9] = {0.0f, 0.125f, 0.25f, 0.3f, 0.5f, 0.7f, 0.75f, 0.9f, 1.0f};
    for (uint32_t k = 0; k < 8; k++) {
      st[k] = k < n ? ((next_u64() & 3u) ? grid[next_u64() % 9u] : (float)(next_u64() % 1000u) / 999.0f)
                    : __builtin_inff();
      col[k] = (float)(next_u64() % 100000u) / 99999.0f;
    }
    std::sort(st, st + n);
    float offs[6] = {st[next_u64() % n], (float)(next_u64() % 1000u) / 999.0f, -0.25f, 1.5f, 0.0f, 1.0f};
    for (float o0 : offs) {
      const float ov[3] = {o0, std::nextafterf(o0, -1e9f), std::nextafterf(o0, 1e9f)};
      for (float o : ov) {
        int a0, a1, b0, b1;
        float ta, tb;
        bool xa, xb;
        stops_find8(st, n, o, a0, a1, ta, xa);
        stops_search(st, n, o, b0, b1, tb, xb);
        cases++;
        const float ca = xa ? 0.0f : col[a0] + (col[a1] - col[a0]) * ta;
        const float cb = xb ? 0.0f : col[b0] + (col[b1] - col[b0]) * tb;
        if (xa != xb || __builtin_bit_cast(uint32_t, ca) != __builtin_bit_cast(uint32_t, cb)) bad++;
      }
    }
  }
  return bad;
}

int main(int argc, char **argv) {
  const int trials = argc > 1 ? std::atoi(argv[1]) : 24;
  uint64_t cases = 0, bad = 0;
  {
    uint64_t rc = 0;
    const uint64_t rb = check_rewrites(rc);
    std::printf("rewrites: cases %llu mismatches %llu\n", (unsigned long long)rc, (unsigned long long)rb);
    cases += rc;
    bad += rb;
  }
  for (uint32_t fmt = 0; fmt < 8; fmt++)
    for (uint32_t op = 0; op < 28; op++)
      for (uint32_t prec = 0; prec < 2; prec++)
        for (int t = 0; t < trials; t++) {
          const uint32_t raw = random_raw(fmt);
          z2d_pixel paint{Z2D_FORMAT_RGBA, 0, 0, 0, 0};
          paint.a = (uint8_t)(rnd(3) == 0 ? 255u : rnd(256));
          paint.r = (uint8_t)(paint.a ? rnd(paint.a + 1u) : 0u);
          paint.g = (uint8_t)(paint.a ? rnd(paint.a + 1u) : 0u);
          paint.b = (uint8_t)(paint.a ? rnd(paint.a + 1u) : 0u);
          DevPaint P;
          std::memset(&P, 0, sizeof(P));
          P.kind = Z2D_PATTERN_OPAQUE;
          P.px_format = Z2D_FORMAT_RGBA;
          P.px = I4{paint.r, paint.g, paint.b, paint.a};
          P.px_opaque = paint.a == 255;
          // the renderer's flags and precision upgrade (painter.fill's
          // fillReducesToSource, Operator.requiresFloat)
          const bool fast = op == Z2D_OP_SRC || (op == Z2D_OP_SRC_OVER && paint.a == 255);
          const uint32_t precision = op_requires_float(op) ? Z2D_PRECISION_FLOAT : prec;
          for (int mode = 0; mode < 2; mode++) {
            DevDraw D;
            std::memset(&D, 0, sizeof(D));
            D.mode = mode == 0 ? MODE_MSAA : MODE_DIRECT;
            D.op = op;
            D.precision = precision;
            D.flags = fast ? DRAW_FAST_SRC : 0u;
            D.mask_bits = 8;
            const int kfull = mode == 0 ? 16 : 1;
            const StopView sv{nullptr, nullptr};
            for (int k = 0; k <= kfull; k++) {
              const uint32_t got = composite_outlined(raw, fmt, k, kfull, D, P, sv, nullptr, 0, 0);
              uint32_t want;
              z2d_comp_op ops[2];
              std::memset(ops, 0, sizeof(ops));
              if (k == 0) {
                want = raw;  // no span covers the pixel
              } else if (k >= kfull) {
                // compositeOpaque: paintStride for the fast path, else one
                // operation with the paint as source
                ops[0].op = fast ? (uint32_t)Z2D_OP_SRC : op;
                ops[0].src.kind = Z2D_PARAM_PIXEL;
                ops[0].src.pixel = paint;
                want = oracle_run(fmt, raw, ops, 1,
                                  fast ? (uint32_t)Z2D_PRECISION_INTEGER : precision);
              } else {
                // compositeOpacity at the coverage's opacity: the paint
                // masked by it (dst_in), then the operator; compositeStride
                // (the fast path) is the same pair in integer precision
                const uint8_t opacity = (uint8_t)(k * 16 - 1);
                ops[0].op = Z2D_OP_DST_IN;
                ops[0].dst.kind = Z2D_PARAM_PIXEL;
                ops[0].dst.pixel = paint;
                ops[0].src.kind = Z2D_PARAM_PIXEL;
                ops[0].src.pixel = z2d_pixel{Z2D_FORMAT_ALPHA8, 0, 0, 0, opacity};
                ops[1].op = op;
                want = oracle_run(fmt, raw, ops, 2,
                                  fast ? (uint32_t)Z2D_PRECISION_INTEGER : precision);
              }
              if (op == Z2D_OP_CLEAR && k > 0) want = 0;  // clearStride
              cases++;
              if (got != want) {
                if (bad < 12)
                  std::printf("mismatch fmt %u op %u prec %u mode %d k %d raw %08x paint "
                              "%02x%02x%02x%02x: got %08x want %08x\n",
                              fmt, op, precision, mode, k, raw, paint.r, paint.g, paint.b,
                              paint.a, got, want);
                bad++;
              }
            }
          }
        }
  std::printf("cases %llu mismatches %llu\n", (unsigned long long)cases, (unsigned long long)bad);
  return bad != 0;
}
