/* SPDX-License-Identifier: MPL-2.0
 * Restates code of z2d (Copyright (c) 2024-2026 Chris Marchesi, MPL-2.0),
 * parts of which are derived from the Cairo project (Copyright (c) 2002
 * University of Southern California; cairo-spline.c).  The Mozilla Public
 * License v2.0 applies to this file: https://mozilla.org/MPL/2.0/.
 */
/*
 * z2d_oracle.c - ORACLE (test infrastructure only).
 *
 * A scalar C restatement of the reference's painter hot path, written from
 * the Zig sources and used to check the HIP backend.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fno-fast-math).
 * Zig does not contract floating-point expressions, so neither may we.
 *
 * Conventions shared with the reference:
 *   - @round  == C round()  (half away from zero)
 *   - @rem    == fmod
 *   - @mod    == x < 0 ? fmod(fmod(x, y) + y, y) : fmod(x, y)  (Zig's
 *                lowering of float @mod)
 *   - @min/@max on floats return the non-NaN operand (fmin/fmax)
 *   - integer compositor arithmetic is u16 with truncating division
 *
 * Pinning status of each part is recorded in the header comment of the
 * test that checks it; a part with no independent check is marked
 * "parity unpinned" there.
 */
#include "z2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------
 * small helpers
 * ---------------------------------------------------------------------- */

static inline double zmin_d(double a, double b) { return fmin(a, b); }
static inline double zmax_d(double a, double b) { return fmax(a, b); }
static inline float zmin_f(float a, float b) { return fminf(a, b); }
static inline float zmax_f(float a, float b) { return fmaxf(a, b); }
static inline int32_t clamp_i32(int32_t v, int32_t lo, int32_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}
/* std.math.clamp == @max(lo, @min(v, hi)) */
static inline float clamp_f(float v, float lo, float hi) {
  return zmax_f(lo, zmin_f(v, hi));
}
static inline double clamp_d(double v, double lo, double hi) {
  return zmax_d(lo, zmin_d(v, hi));
}
/* Zig float @mod lowering (see file header). */
static inline float zmod_f(float x, float y) {
  float a = fmodf(x, y);
  if (x < 0.0f) return fmodf(a + y, y);
  return a;
}
/* std.math.atan / std.math.atan2 for f64 as Zig 0.16's standard library
 * computes them (lib/std/math/atan.zig, atan2.zig; a third-party dependency
 * of the reference, not under /root/reference).  Both restate the published
 * fdlibm algorithm (Sun Microsystems, e_atan2.c / s_atan.c, as carried by
 * musl): atan reduces |x| to one of five ranges around 0, 0.5, 1, 1.5 and
 * infinity, evaluates an odd polynomial split into two Horner chains, and
 * adds atan of the range's centre in two parts; atan2 handles signs, zeros
 * and infinities and takes atan(|y/x|).  Used by Conic.getOffset
 * (src/gradient.zig:731-742). */
static uint32_t hi_word(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return (uint32_t)(u >> 32);
}
static uint32_t lo_word(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return (uint32_t)u;
}
static double zo_atan(double x) {
  static const double atanhi[4] = {4.63647609000806093515e-01, 7.85398163397448278999e-01,
                                   9.82793723247329054082e-01, 1.57079632679489655800e+00};
  static const double atanlo[4] = {2.26987774529616870924e-17, 3.06161699786838301793e-17,
                                   1.39033110312309984516e-17, 6.12323399573676603587e-17};
  static const double aT[11] = {3.33333333333329318027e-01,  -1.99999999998764832476e-01,
                                1.42857142725034663711e-01,  -1.11111104054623557880e-01,
                                9.09088713343650656196e-02,  -7.69187620504482999495e-02,
                                6.66107313738753120669e-02,  -5.83357013379057348645e-02,
                                4.97687799461593236017e-02,  -3.65315727442169155270e-02,
                                1.62858201153657823623e-02};
  uint32_t ix = hi_word(x);
  const uint32_t sign = ix >> 31;
  ix &= 0x7FFFFFFFu;
  if (ix >= 0x44100000u) { /* |x| >= 2^66 */
    if (isnan(x)) return x;
    const double z = atanhi[3] + 0x1.0p-120;
    return sign ? -z : z;
  }
  int id;
  if (ix < 0x3FDC0000u) { /* |x| < 0.4375 */
    if (ix < 0x3E400000u) return x; /* |x| < 2^-27 */
    id = -1;
  } else {
    x = fabs(x);
    if (ix < 0x3FF30000u) { /* |x| < 1.1875 */
      if (ix < 0x3FE60000u) { /* 7/16 <= |x| < 11/16 */
        id = 0;
        x = (2.0 * x - 1.0) / (2.0 + x);
      } else { /* 11/16 <= |x| < 19/16 */
        id = 1;
        x = (x - 1.0) / (x + 1.0);
      }
    } else if (ix < 0x40038000u) { /* |x| < 2.4375 */
      id = 2;
      x = (x - 1.5) / (1.0 + 1.5 * x);
    } else { /* 2.4375 <= |x| < 2^66 */
      id = 3;
      x = -1.0 / x;
    }
  }
  const double z = x * x;
  const double w = z * z;
  const double s1 =
      z * (aT[0] + w * (aT[2] + w * (aT[4] + w * (aT[6] + w * (aT[8] + w * aT[10])))));
  const double s2 = w * (aT[1] + w * (aT[3] + w * (aT[5] + w * (aT[7] + w * aT[9]))));
  if (id < 0) return x - x * (s1 + s2);
  const double zz = atanhi[id] - ((x * (s1 + s2) - atanlo[id]) - x);
  return sign ? -zz : zz;
}
double zo_atan2(double y, double x) {
  const double pi = 3.1415926535897931160E+00, pi_lo = 1.2246467991473531772E-16;
  if (isnan(x) || isnan(y)) return x + y;
  uint32_t ix = hi_word(x), iy = hi_word(y);
  const uint32_t lx = lo_word(x), ly = lo_word(y);
  if (((ix - 0x3FF00000u) | lx) == 0) return zo_atan(y); /* x = 1.0 */
  const uint32_t m = ((iy >> 31) & 1u) | ((ix >> 30) & 2u); /* 2 sign(x) + sign(y) */
  ix &= 0x7FFFFFFFu;
  iy &= 0x7FFFFFFFu;
  if ((iy | ly) == 0) { /* y = 0 */
    if (m < 2) return y;
    return m == 2 ? pi : -pi;
  }
  if ((ix | lx) == 0) return (m & 1u) ? -pi / 2 : pi / 2; /* x = 0 */
  if (ix == 0x7FF00000u) { /* x infinite */
    if (iy == 0x7FF00000u) {
      static const double q[4] = {3.1415926535897931160E+00 / 4, -3.1415926535897931160E+00 / 4,
                                  3 * 3.1415926535897931160E+00 / 4,
                                  -3 * 3.1415926535897931160E+00 / 4};
      return q[m];
    }
    static const double r[4] = {0.0, -0.0, 3.1415926535897931160E+00, -3.1415926535897931160E+00};
    return r[m];
  }
  if (ix + (64u << 20) < iy || iy == 0x7FF00000u) /* |y/x| > 2^64 */
    return (m & 1u) ? -pi / 2 : pi / 2;
  double z;
  if ((m & 2u) && iy + (64u << 20) < ix) /* |y/x| < 2^-64, x < 0 */
    z = 0.0;
  else
    z = zo_atan(fabs(y / x));
  switch (m) {
    case 0: return z;
    case 1: return -z;
    case 2: return pi - (z - pi_lo);
    default: return (z - pi_lo) - pi;
  }
}

static inline double zmod_d(double x, double y) {
  double a = fmod(x, y);
  if (x < 0.0) return fmod(a + y, y);
  return a;
}

/* Saturating conversion of a rounded f32 color to an 8-bit channel.  The
 * reference's @intFromFloat is only defined for in-range values; for those
 * this is the identity. */
static inline uint16_t f2u(float v) {
  float r = roundf(v);
  if (!(r >= 0.0f)) return 0;
  if (r > 65535.0f) return 65535;
  return (uint16_t)r;
}

/* ------------------------------------------------------------------------
 * std.math.pow for f32 (Zig stdlib, pinned version 0.16.0; the algorithm is
 * Go's math.Pow).  Used for sRGB gamma, src/color.zig:336-354 and
 * src/internal/color_vector.zig:210-241.
 * ---------------------------------------------------------------------- */
static int is_odd_integer_f(float y) {
  float ip;
  float fp = modff(y, &ip);
  if (fp != 0.0f) return 0;
  return fmodf(fabsf(ip), 2.0f) == 1.0f;
}

float zo_powf(float x, float y) {
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (isnan(x) || isnan(y)) return NAN;
  if (y == 1.0f) return x;
  if (x == 0.0f) {
    if (y < 0.0f) return is_odd_integer_f(y) ? copysignf(INFINITY, x) : INFINITY;
    return is_odd_integer_f(y) ? x : 0.0f;
  }
  if (isinf(y)) {
    if (x == -1.0f) return 1.0f;
    if ((fabsf(x) < 1.0f) == (y > 0.0f)) return 0.0f;
    return INFINITY;
  }
  if (isinf(x)) {
    if (x < 0.0f) return zo_powf(1.0f / x, -y);
    return y < 0.0f ? 0.0f : INFINITY;
  }
  if (y == 0.5f) return sqrtf(x);
  if (y == -0.5f) return 1.0f / sqrtf(x);

  float yi;
  float yf = modff(fabsf(y), &yi);
  if (yf != 0.0f && x < 0.0f) return NAN;
  if (yi >= 2147483648.0f) {
    if (x == -1.0f) return 1.0f;
    if ((fabsf(x) < 1.0f) == (y > 0.0f)) return 0.0f;
    return INFINITY;
  }
  float a1 = 1.0f;
  int32_t ae = 0;
  if (yf != 0.0f) {
    if (yf > 0.5f) {
      yf -= 1.0f;
      yi += 1.0f;
    }
    a1 = expf(yf * logf(x));
  }
  int xe;
  float x1 = frexpf(x, &xe);
  int32_t i = (int32_t)yi;
  while (i != 0) {
    const int overflow_shift = 8 + 1; /* floatExponentBits(f32) + 1 */
    if (xe < -(1 << overflow_shift) || (1 << overflow_shift) < xe) {
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

/* ------------------------------------------------------------------------
 * pixel formats and surface memory (src/pixel.zig, src/surface.zig)
 * ---------------------------------------------------------------------- */

static inline int fmt_bits(uint32_t f) {
  switch (f) {
    case Z2D_FORMAT_ALPHA8: return 8;
    case Z2D_FORMAT_ALPHA4: return 4;
    case Z2D_FORMAT_ALPHA2: return 2;
    case Z2D_FORMAT_ALPHA1: return 1;
    default: return 32;
  }
}
static inline int fmt_has_color(uint32_t f) { return f <= Z2D_FORMAT_RGBA; }
static inline int fmt_packed(uint32_t f) { return f >= Z2D_FORMAT_ALPHA4; }

/* Alpha8.scaleIntoIntAlpha from an n-bit value (src/pixel.zig:582-630):
 * bit repetition when scaling up. */
static inline uint16_t alpha_up8(uint32_t v, int bits) {
  switch (bits) {
    case 1: return (uint16_t)(v * 255u);
    case 2: return (uint16_t)((v << 6) | (v << 4) | (v << 2) | v);
    case 4: return (uint16_t)((v << 4) + v);
    default: return (uint16_t)v;
  }
}
/* Alpha{4,2,1}.scaleIntoIntAlpha from u8: shift right. */
static inline uint32_t alpha_down(uint32_t v8, int bits) {
  if (bits >= 8) return v8 & 0xFF;
  return (v8 & 0xFF) >> (8 - bits);
}

/* mem.readPackedInt/writePackedInt little-endian at bit index i*bits. */
static inline uint32_t packed_get(const uint8_t *buf, uint64_t idx, int bits) {
  uint64_t bit = idx * (uint64_t)bits;
  return (buf[bit >> 3] >> (bit & 7)) & ((1u << bits) - 1u);
}
static inline void packed_set(uint8_t *buf, uint64_t idx, int bits, uint32_t v) {
  uint64_t bit = idx * (uint64_t)bits;
  uint8_t mask = (uint8_t)(((1u << bits) - 1u) << (bit & 7));
  buf[bit >> 3] = (uint8_t)((buf[bit >> 3] & ~mask) | ((v << (bit & 7)) & mask));
}

typedef struct {
  uint16_t r, g, b, a;
} rgba16;

/* RGBA16Vec.fromStride / RGBA16.fromPixelT for one pixel
 * (src/compositor.zig:716-747, :859-940, :664-683). */
static rgba16 load_px(const z2d_surface_desc *s, uint64_t idx) {
  rgba16 o = {0, 0, 0, 0};
  const uint8_t *p;
  switch (s->format) {
    case Z2D_FORMAT_RGBA:
      p = s->data + idx * 4;
      o.r = p[0]; o.g = p[1]; o.b = p[2]; o.a = p[3];
      break;
    case Z2D_FORMAT_ARGB:
      p = s->data + idx * 4;
      o.b = p[0]; o.g = p[1]; o.r = p[2]; o.a = p[3];
      break;
    case Z2D_FORMAT_RGB:
      p = s->data + idx * 4;
      o.r = p[0]; o.g = p[1]; o.b = p[2]; o.a = 255;
      break;
    case Z2D_FORMAT_XRGB:
      p = s->data + idx * 4;
      o.b = p[0]; o.g = p[1]; o.r = p[2]; o.a = 255;
      break;
    case Z2D_FORMAT_ALPHA8:
      o.a = s->data[idx];
      break;
    default: {
      int bits = fmt_bits(s->format);
      o.a = alpha_up8(packed_get(s->data, idx, bits), bits);
    } break;
  }
  return o;
}

/* RGBA16Vec.toStride for one pixel (src/compositor.zig:806-837, :957-1037):
 * rgb/xrgb pad byte is written as zero; packed alpha is shifted down. */
static void store_px(z2d_surface_desc *s, uint64_t idx, rgba16 v) {
  uint8_t *p;
  switch (s->format) {
    case Z2D_FORMAT_RGBA:
      p = s->data + idx * 4;
      p[0] = (uint8_t)v.r; p[1] = (uint8_t)v.g; p[2] = (uint8_t)v.b; p[3] = (uint8_t)v.a;
      break;
    case Z2D_FORMAT_ARGB:
      p = s->data + idx * 4;
      p[0] = (uint8_t)v.b; p[1] = (uint8_t)v.g; p[2] = (uint8_t)v.r; p[3] = (uint8_t)v.a;
      break;
    case Z2D_FORMAT_RGB:
      p = s->data + idx * 4;
      p[0] = (uint8_t)v.r; p[1] = (uint8_t)v.g; p[2] = (uint8_t)v.b; p[3] = 0;
      break;
    case Z2D_FORMAT_XRGB:
      p = s->data + idx * 4;
      p[0] = (uint8_t)v.b; p[1] = (uint8_t)v.g; p[2] = (uint8_t)v.r; p[3] = 0;
      break;
    case Z2D_FORMAT_ALPHA8:
      s->data[idx] = (uint8_t)v.a;
      break;
    default: {
      int bits = fmt_bits(s->format);
      packed_set(s->data, idx, bits, alpha_down(v.a, bits));
    } break;
  }
}

/* pixel.RGBA.fromPixel (src/pixel.zig:399-432) as RGBA16. */
static rgba16 pixel_to_rgba(z2d_pixel px) {
  rgba16 o = {0, 0, 0, 0};
  switch (px.format) {
    case Z2D_FORMAT_RGBA:
    case Z2D_FORMAT_ARGB:
      o.r = px.r; o.g = px.g; o.b = px.b; o.a = px.a;
      break;
    case Z2D_FORMAT_RGB:
    case Z2D_FORMAT_XRGB:
      o.r = px.r; o.g = px.g; o.b = px.b; o.a = 255;
      break;
    default:
      o.a = alpha_up8(px.a, fmt_bits(px.format));
      break;
  }
  return o;
}

/* T.fromPixel(px) written by paintStride (src/surface.zig:484-493,
 * :778-820): same bytes as running `src` through the vector store. */
static void paint_px(z2d_surface_desc *s, uint64_t idx, z2d_pixel px) {
  store_px(s, idx, pixel_to_rgba(px));
}

static uint64_t surface_bytes(uint32_t format, int32_t w, int32_t h) {
  uint64_t n = (uint64_t)w * (uint64_t)h;
  int bits = fmt_bits(format);
  if (bits == 32) return n * 4;
  return (n * (uint64_t)bits + 7) / 8;
}

/* Surface.paintStride / clearStride (src/surface.zig:303-318, :484-493,
 * :778-820): out-of-range start is a no-op, len wraps across rows. */
static void paint_stride(z2d_surface_desc *s, int32_t x, int32_t y, uint64_t len,
                         z2d_pixel px) {
  if (x < 0 || y < 0 || x >= s->width || y >= s->height) return;
  uint64_t start = (uint64_t)s->width * (uint64_t)y + (uint64_t)x;
  for (uint64_t i = start; i < start + len; i++) paint_px(s, i, px);
}
static void clear_stride(z2d_surface_desc *s, int32_t x, int32_t y, uint64_t len) {
  z2d_pixel z = {Z2D_FORMAT_RGBA, 0, 0, 0, 0};
  paint_stride(s, x, y, len, z);
}

/* RGBA.canDemultiply (src/pixel.zig:504-517). */
static int can_demultiply(z2d_pixel px) {
  if (px.a == 0) return 1;
  if ((uint32_t)px.r * 255u / px.a > 255u) return 0;
  if ((uint32_t)px.g * 255u / px.a > 255u) return 0;
  if ((uint32_t)px.b * 255u / px.a > 255u) return 0;
  return 1;
}

/* Pixel.isOpaque (src/pixel.zig:131-140). */
static int pixel_is_opaque(z2d_pixel px) {
  switch (px.format) {
    case Z2D_FORMAT_XRGB:
    case Z2D_FORMAT_RGB: return 1;
    case Z2D_FORMAT_ARGB:
    case Z2D_FORMAT_RGBA:
    case Z2D_FORMAT_ALPHA8: return px.a == 255;
    case Z2D_FORMAT_ALPHA4: return px.a == 15;
    case Z2D_FORMAT_ALPHA2: return px.a == 3;
    default: return px.a == 1;
  }
}

/* ------------------------------------------------------------------------
 * Integer operators (IntegerOps, src/compositor.zig:1158-1568).
 * u16 arithmetic; mul(a,b) = (a*b)/255 truncated; widened i32 for
 * overlay/hard_light.
 * ---------------------------------------------------------------------- */
#define U16(x) ((uint16_t)((x) & 0xFFFF))
static inline uint16_t imul(uint16_t a, uint16_t b) { return U16(((uint32_t)a * b) / 255u); }
static inline uint16_t iinv(uint16_t a) { return U16(255u - a); }
static inline uint16_t iinvmul(uint16_t a, uint16_t b) { return imul(a, iinv(b)); }
static inline int32_t wmul(int32_t a, int32_t b) { return (a * b) / 255; }
static inline int32_t winvmul(int32_t a, int32_t b) { return wmul(a, 255 - b); }
static inline int32_t wrinvmul(int32_t a, int32_t b) { return wmul(a, 255 + b); }

static uint16_t overlay_ch(uint16_t sca, uint16_t dca, uint16_t sa, uint16_t da) {
  if (U16(dca * 2u) <= da) {
    int32_t r = wmul((int32_t)sca * 2, dca) + winvmul(sca, da) + winvmul(dca, sa);
    return U16(r);
  }
  int32_t r = wrinvmul(sca, da) + wrinvmul(dca, sa) - wmul((int32_t)dca * 2, sca) -
              wmul(da, sa);
  return U16(r);
}
static uint16_t hardlight_ch(uint16_t sca, uint16_t dca, uint16_t sa, uint16_t da) {
  if (U16(sca * 2u) <= sa) {
    int32_t r = wmul((int32_t)sca * 2, dca) + winvmul(sca, da) + winvmul(dca, sa);
    return U16(r);
  }
  int32_t r = wrinvmul(sca, da) + wrinvmul(dca, sa) - wmul(sa, da) -
              wmul((int32_t)sca * 2, dca);
  return U16(r);
}
static inline uint16_t umin16(uint16_t a, uint16_t b) { return a < b ? a : b; }
static inline uint16_t umax16(uint16_t a, uint16_t b) { return a > b ? a : b; }

static rgba16 int_op(uint32_t op, rgba16 d, rgba16 s) {
  rgba16 o = {0, 0, 0, 0};
  uint16_t sa = s.a, da = d.a;
  uint16_t uni_a = U16(sa + da - imul(sa, da));
#define CH3(expr_r, expr_g, expr_b) \
  do {                              \
    o.r = U16(expr_r);              \
    o.g = U16(expr_g);              \
    o.b = U16(expr_b);              \
  } while (0)
  switch (op) {
    case Z2D_OP_SRC: return s;
    case Z2D_OP_DST: return d;
    case Z2D_OP_SRC_OVER:
      CH3(s.r + iinvmul(d.r, sa), s.g + iinvmul(d.g, sa), s.b + iinvmul(d.b, sa));
      o.a = uni_a;
      return o;
    case Z2D_OP_DST_OVER:
      CH3(d.r + iinvmul(s.r, da), d.g + iinvmul(s.g, da), d.b + iinvmul(s.b, da));
      o.a = uni_a;
      return o;
    case Z2D_OP_SRC_IN:
      CH3(imul(s.r, da), imul(s.g, da), imul(s.b, da));
      o.a = imul(sa, da);
      return o;
    case Z2D_OP_DST_IN:
      CH3(imul(d.r, sa), imul(d.g, sa), imul(d.b, sa));
      o.a = imul(da, sa);
      return o;
    case Z2D_OP_SRC_OUT:
      CH3(iinvmul(s.r, da), iinvmul(s.g, da), iinvmul(s.b, da));
      o.a = iinvmul(sa, da);
      return o;
    case Z2D_OP_DST_OUT:
      CH3(iinvmul(d.r, sa), iinvmul(d.g, sa), iinvmul(d.b, sa));
      o.a = iinvmul(da, sa);
      return o;
    case Z2D_OP_SRC_ATOP:
      CH3(imul(s.r, da) + iinvmul(d.r, sa), imul(s.g, da) + iinvmul(d.g, sa),
          imul(s.b, da) + iinvmul(d.b, sa));
      o.a = da;
      return o;
    case Z2D_OP_DST_ATOP:
      CH3(imul(d.r, sa) + iinvmul(s.r, da), imul(d.g, sa) + iinvmul(s.g, da),
          imul(d.b, sa) + iinvmul(s.b, da));
      o.a = sa;
      return o;
    case Z2D_OP_XOR:
      CH3(iinvmul(s.r, da) + iinvmul(d.r, sa), iinvmul(s.g, da) + iinvmul(d.g, sa),
          iinvmul(s.b, da) + iinvmul(d.b, sa));
      o.a = U16(iinvmul(sa, da) + iinvmul(da, sa));
      return o;
    case Z2D_OP_PLUS:
      o.r = umin16(255, U16(s.r + d.r));
      o.g = umin16(255, U16(s.g + d.g));
      o.b = umin16(255, U16(s.b + d.b));
      o.a = umin16(255, U16(sa + da));
      return o;
    case Z2D_OP_MULTIPLY:
      CH3(imul(s.r, d.r) + iinvmul(s.r, da) + iinvmul(d.r, sa),
          imul(s.g, d.g) + iinvmul(s.g, da) + iinvmul(d.g, sa),
          imul(s.b, d.b) + iinvmul(s.b, da) + iinvmul(d.b, sa));
      o.a = uni_a;
      return o;
    case Z2D_OP_SCREEN:
      CH3(s.r + d.r - imul(s.r, d.r), s.g + d.g - imul(s.g, d.g),
          s.b + d.b - imul(s.b, d.b));
      o.a = uni_a;
      return o;
    case Z2D_OP_OVERLAY:
      o.r = overlay_ch(s.r, d.r, sa, da);
      o.g = overlay_ch(s.g, d.g, sa, da);
      o.b = overlay_ch(s.b, d.b, sa, da);
      o.a = uni_a;
      return o;
    case Z2D_OP_DARKEN:
      CH3(umin16(imul(s.r, da), imul(d.r, sa)) + iinvmul(s.r, da) + iinvmul(d.r, sa),
          umin16(imul(s.g, da), imul(d.g, sa)) + iinvmul(s.g, da) + iinvmul(d.g, sa),
          umin16(imul(s.b, da), imul(d.b, sa)) + iinvmul(s.b, da) + iinvmul(d.b, sa));
      o.a = uni_a;
      return o;
    case Z2D_OP_LIGHTEN:
      CH3(umax16(imul(s.r, da), imul(d.r, sa)) + iinvmul(s.r, da) + iinvmul(d.r, sa),
          umax16(imul(s.g, da), imul(d.g, sa)) + iinvmul(s.g, da) + iinvmul(d.g, sa),
          umax16(imul(s.b, da), imul(d.b, sa)) + iinvmul(s.b, da) + iinvmul(d.b, sa));
      o.a = uni_a;
      return o;
    case Z2D_OP_HARD_LIGHT:
      o.r = hardlight_ch(s.r, d.r, sa, da);
      o.g = hardlight_ch(s.g, d.g, sa, da);
      o.b = hardlight_ch(s.b, d.b, sa, da);
      o.a = uni_a;
      return o;
    case Z2D_OP_DIFFERENCE:
      CH3(s.r + d.r - 2u * umin16(imul(s.r, da), imul(d.r, sa)),
          s.g + d.g - 2u * umin16(imul(s.g, da), imul(d.g, sa)),
          s.b + d.b - 2u * umin16(imul(s.b, da), imul(d.b, sa)));
      o.a = uni_a;
      return o;
    case Z2D_OP_EXCLUSION:
      CH3(U16(imul(s.r, da) + imul(d.r, sa) - 2u * imul(s.r, d.r)) + iinvmul(s.r, da) +
              iinvmul(d.r, sa),
          U16(imul(s.g, da) + imul(d.g, sa) - 2u * imul(s.g, d.g)) + iinvmul(s.g, da) +
              iinvmul(d.g, sa),
          U16(imul(s.b, da) + imul(d.b, sa) - 2u * imul(s.b, d.b)) + iinvmul(s.b, da) +
              iinvmul(d.b, sa));
      o.a = uni_a;
      return o;
    /* clear, and every float-only operator, map to clear in integer mode
     * (src/compositor.zig:1179-1188). */
    default: return o;
  }
#undef CH3
}

/* ------------------------------------------------------------------------
 * Float operators (FloatOps, src/compositor.zig:1571-2440), with the
 * semantics of the vectorized path the stride compositor always takes.
 * ---------------------------------------------------------------------- */
typedef struct {
  float r, g, b, a;
} rgbaf;

static float dodge_ch(float sca, float dca, float sa, float da) {
  if (sca == sa && dca == 0.0f) return sca * (1.0f - da);
  if (sca == sa) return sa * da + sca * (1.0f - da) + dca * (1.0f - sa);
  return sa * da * zmin_f(1.0f, dca / da * sa / (sa - sca)) + sca * (1.0f - da) +
         dca * (1.0f - sa);
}
/* p0 is called as p0(sca, dca, sa) against a (sca, dca, da) signature, so
 * the first branch tests dca == sa (src/compositor.zig:1933-1950). */
static float burn_ch(float sca, float dca, float sa, float da) {
  if (sca == 0.0f && dca == sa) return sa * da + dca * (1.0f - sa);
  if (sca == 0.0f) return dca * (1.0f - sa);
  return sa * da * (1.0f - zmin_f(1.0f, (1.0f - dca / da) * sa / sca)) +
         sca * (1.0f - da) + dca * (1.0f - sa);
}
static float softlight_ch(float sca, float dca, float sa, float da) {
  if (da == 0.0f) return sca;
  if (2.0f * sca <= sa) {
    float m = dca / da;
    return dca * (sa + (2.0f * sca - sa) * (1.0f - m)) + sca * (1.0f - da) +
           dca * (1.0f - sa);
  }
  if (2.0f * sca > sa && 4.0f * dca <= da) {
    float m = dca / da;
    return dca * sa +
           da * (2.0f * sca - sa) * (4.0f * m * (4.0f * m + 1.0f) * (m - 1.0f) + 7.0f * m) +
           sca * (1.0f - da) + dca * (1.0f - sa);
  }
  float m = dca / da;
  return da * (2.0f * sca - sa) * (sqrtf(m) - m) + sca - sca * da + dca;
}

typedef struct {
  float r, g, b;
} ns3;
static inline float ns_max3(ns3 c) { return zmax_f(zmax_f(c.r, c.g), c.b); }
static inline float ns_min3(ns3 c) { return zmin_f(zmin_f(c.r, c.g), c.b); }
static inline float ns_sat(ns3 c) { return ns_max3(c) - ns_min3(c); }
static inline float ns_lum(ns3 c) { return c.r * 0.3f + c.g * 0.59f + c.b * 0.11f; }
/* clipColor, vector form: the x > a branch overrides the n < 0 branch and
 * both use the unclipped channel (src/compositor.zig:2286-2318). */
static float clip_ch(float c, float l, float n, float x, float a) {
  float r = c;
  if (n < 0.0f) {
    float t = l - n;
    r = (t == 0.0f) ? 0.0f : (l + ((c - l) * l) / t);
  }
  if (x > a) {
    float t = x - l;
    r = (t == 0.0f) ? 0.0f : (l + ((c - l) * (a - l)) / t);
  }
  return r;
}
static ns3 ns_clip(ns3 c, float a) {
  float l = ns_lum(c), n = ns_min3(c), x = ns_max3(c);
  ns3 o = {clip_ch(c.r, l, n, x, a), clip_ch(c.g, l, n, x, a), clip_ch(c.b, l, n, x, a)};
  return o;
}
static ns3 ns_setlum(ns3 c, float l) {
  float d = l - ns_lum(c);
  ns3 o = {c.r + d, c.g + d, c.b + d};
  return o;
}
static ns3 ns_setsat(ns3 c, float s) {
  float n = ns_min3(c), x = ns_max3(c), d = x - n;
  ns3 o;
  o.r = (d == 0.0f) ? 0.0f : (c.r - n) * s / d;
  o.g = (d == 0.0f) ? 0.0f : (c.g - n) * s / d;
  o.b = (d == 0.0f) ? 0.0f : (c.b - n) * s / d;
  return o;
}
static rgbaf ns_to_rgba(ns3 c, rgbaf d, rgbaf s) {
  rgbaf o;
  o.r = s.r * (1.0f - d.a) + d.r * (1.0f - s.a) + c.r;
  o.g = s.g * (1.0f - d.a) + d.g * (1.0f - s.a) + c.g;
  o.b = s.b * (1.0f - d.a) + d.b * (1.0f - s.a) + c.b;
  o.a = s.a + d.a - s.a * d.a;
  return o;
}

static rgbaf float_op(uint32_t op, rgbaf d, rgbaf s) {
  rgbaf o = {0, 0, 0, 0};
  float uni_a = s.a + d.a - s.a * d.a;
  switch (op) {
    case Z2D_OP_CLEAR: return o;
    case Z2D_OP_SRC: return s;
    case Z2D_OP_DST: return d;
    case Z2D_OP_SRC_OVER:
      o.r = s.r + d.r * (1.0f - s.a);
      o.g = s.g + d.g * (1.0f - s.a);
      o.b = s.b + d.b * (1.0f - s.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_DST_OVER:
      o.r = d.r + s.r * (1.0f - d.a);
      o.g = d.g + s.g * (1.0f - d.a);
      o.b = d.b + s.b * (1.0f - d.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_SRC_IN:
      o.r = s.r * d.a; o.g = s.g * d.a; o.b = s.b * d.a; o.a = s.a * d.a;
      return o;
    case Z2D_OP_DST_IN:
      o.r = d.r * s.a; o.g = d.g * s.a; o.b = d.b * s.a; o.a = s.a * d.a;
      return o;
    case Z2D_OP_SRC_OUT:
      o.r = s.r * (1.0f - d.a); o.g = s.g * (1.0f - d.a); o.b = s.b * (1.0f - d.a);
      o.a = s.a * (1.0f - d.a);
      return o;
    case Z2D_OP_DST_OUT:
      o.r = d.r * (1.0f - s.a); o.g = d.g * (1.0f - s.a); o.b = d.b * (1.0f - s.a);
      o.a = d.a * (1.0f - s.a);
      return o;
    case Z2D_OP_SRC_ATOP:
      o.r = s.r * d.a + d.r * (1.0f - s.a);
      o.g = s.g * d.a + d.g * (1.0f - s.a);
      o.b = s.b * d.a + d.b * (1.0f - s.a);
      o.a = d.a;
      return o;
    case Z2D_OP_DST_ATOP:
      o.r = d.r * s.a + s.r * (1.0f - d.a);
      o.g = d.g * s.a + s.g * (1.0f - d.a);
      o.b = d.b * s.a + s.b * (1.0f - d.a);
      o.a = s.a;
      return o;
    case Z2D_OP_XOR:
      o.r = s.r * (1.0f - d.a) + d.r * (1.0f - s.a);
      o.g = s.g * (1.0f - d.a) + d.g * (1.0f - s.a);
      o.b = s.b * (1.0f - d.a) + d.b * (1.0f - s.a);
      o.a = s.a + d.a - 2.0f * s.a * d.a;
      return o;
    case Z2D_OP_PLUS:
      o.r = zmin_f(1.0f, s.r + d.r);
      o.g = zmin_f(1.0f, s.g + d.g);
      o.b = zmin_f(1.0f, s.b + d.b);
      o.a = zmin_f(1.0f, s.a + d.a);
      return o;
    case Z2D_OP_MULTIPLY:
      o.r = s.r * d.r + s.r * (1.0f - d.a) + d.r * (1.0f - s.a);
      o.g = s.g * d.g + s.g * (1.0f - d.a) + d.g * (1.0f - s.a);
      o.b = s.b * d.b + s.b * (1.0f - d.a) + d.b * (1.0f - s.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_SCREEN:
      o.r = s.r + d.r - s.r * d.r;
      o.g = s.g + d.g - s.g * d.g;
      o.b = s.b + d.b - s.b * d.b;
      o.a = uni_a;
      return o;
    case Z2D_OP_OVERLAY: {
#define OVL(sc, dc)                                                               \
  ((2.0f * (dc) <= d.a) ? (2.0f * (sc) * (dc) + (sc) * (1.0f - d.a) + (dc) * (1.0f - s.a)) \
                        : ((sc) * (1.0f + d.a) + (dc) * (1.0f + s.a) - 2.0f * (dc) * (sc) - \
                           d.a * s.a))
      o.r = OVL(s.r, d.r);
      o.g = OVL(s.g, d.g);
      o.b = OVL(s.b, d.b);
#undef OVL
      o.a = uni_a;
      return o;
    }
    case Z2D_OP_DARKEN:
      o.r = zmin_f(s.r * d.a, d.r * s.a) + s.r * (1.0f - d.a) + d.r * (1.0f - s.a);
      o.g = zmin_f(s.g * d.a, d.g * s.a) + s.g * (1.0f - d.a) + d.g * (1.0f - s.a);
      o.b = zmin_f(s.b * d.a, d.b * s.a) + s.b * (1.0f - d.a) + d.b * (1.0f - s.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_LIGHTEN:
      o.r = zmax_f(s.r * d.a, d.r * s.a) + s.r * (1.0f - d.a) + d.r * (1.0f - s.a);
      o.g = zmax_f(s.g * d.a, d.g * s.a) + s.g * (1.0f - d.a) + d.g * (1.0f - s.a);
      o.b = zmax_f(s.b * d.a, d.b * s.a) + s.b * (1.0f - d.a) + d.b * (1.0f - s.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_COLOR_DODGE:
      o.r = dodge_ch(s.r, d.r, s.a, d.a);
      o.g = dodge_ch(s.g, d.g, s.a, d.a);
      o.b = dodge_ch(s.b, d.b, s.a, d.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_COLOR_BURN:
      o.r = burn_ch(s.r, d.r, s.a, d.a);
      o.g = burn_ch(s.g, d.g, s.a, d.a);
      o.b = burn_ch(s.b, d.b, s.a, d.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_HARD_LIGHT: {
#define HL(sc, dc)                                                                \
  ((2.0f * (sc) <= s.a) ? (2.0f * (sc) * (dc) + (sc) * (1.0f - d.a) + (dc) * (1.0f - s.a)) \
                        : ((sc) * (1.0f + d.a) + (dc) * (1.0f + s.a) - s.a * d.a -     \
                           2.0f * (sc) * (dc)))
      o.r = HL(s.r, d.r);
      o.g = HL(s.g, d.g);
      o.b = HL(s.b, d.b);
#undef HL
      o.a = uni_a;
      return o;
    }
    case Z2D_OP_SOFT_LIGHT:
      o.r = softlight_ch(s.r, d.r, s.a, d.a);
      o.g = softlight_ch(s.g, d.g, s.a, d.a);
      o.b = softlight_ch(s.b, d.b, s.a, d.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_DIFFERENCE:
      o.r = s.r + d.r - 2.0f * zmin_f(s.r * d.a, d.r * s.a);
      o.g = s.g + d.g - 2.0f * zmin_f(s.g * d.a, d.g * s.a);
      o.b = s.b + d.b - 2.0f * zmin_f(s.b * d.a, d.b * s.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_EXCLUSION:
      o.r = (s.r * d.a + d.r * s.a - 2.0f * s.r * d.r) + s.r * (1.0f - d.a) + d.r * (1.0f - s.a);
      o.g = (s.g * d.a + d.g * s.a - 2.0f * s.g * d.g) + s.g * (1.0f - d.a) + d.g * (1.0f - s.a);
      o.b = (s.b * d.a + d.b * s.a - 2.0f * s.b * d.b) + s.b * (1.0f - d.a) + d.b * (1.0f - s.a);
      o.a = uni_a;
      return o;
    case Z2D_OP_HUE: {
      ns3 cr = {s.r * s.a, s.g * s.a, s.b * s.a};
      ns3 cd = {d.r, d.g, d.b};
      cr = ns_setsat(cr, ns_sat(cd) * s.a);
      cr = ns_setlum(cr, ns_lum(cd) * s.a);
      cr = ns_clip(cr, s.a * d.a);
      return ns_to_rgba(cr, d, s);
    }
    case Z2D_OP_SATURATION: {
      ns3 cr = {d.r * s.a, d.g * s.a, d.b * s.a};
      ns3 cs = {s.r, s.g, s.b};
      ns3 cd = {d.r, d.g, d.b};
      cr = ns_setsat(cr, ns_sat(cs) * d.a);
      cr = ns_setlum(cr, ns_lum(cd) * s.a);
      cr = ns_clip(cr, s.a * d.a);
      return ns_to_rgba(cr, d, s);
    }
    case Z2D_OP_COLOR: {
      ns3 cr = {s.r * d.a, s.g * d.a, s.b * d.a};
      ns3 cd = {d.r, d.g, d.b};
      cr = ns_setlum(cr, ns_lum(cd) * s.a);
      cr = ns_clip(cr, s.a * d.a);
      return ns_to_rgba(cr, d, s);
    }
    case Z2D_OP_LUMINOSITY: {
      ns3 cr = {d.r * s.a, d.g * s.a, d.b * s.a};
      ns3 cs = {s.r, s.g, s.b};
      cr = ns_setlum(cr, ns_lum(cs) * d.a);
      cr = ns_clip(cr, s.a * d.a);
      return ns_to_rgba(cr, d, s);
    }
    default: return o;
  }
}

static inline int op_requires_float(uint32_t op) {
  switch (op) {
    case Z2D_OP_COLOR_DODGE:
    case Z2D_OP_COLOR_BURN:
    case Z2D_OP_SOFT_LIGHT:
    case Z2D_OP_HUE:
    case Z2D_OP_SATURATION:
    case Z2D_OP_COLOR:
    case Z2D_OP_LUMINOSITY: return 1;
    default: return 0;
  }
}
static inline int op_is_bounded(uint32_t op) {
  return !(op == Z2D_OP_SRC_IN || op == Z2D_OP_DST_IN || op == Z2D_OP_SRC_OUT ||
           op == Z2D_OP_DST_ATOP);
}

/* decodeRGBAVecRaw / encodeRGBAVecRaw (src/internal/color_vector.zig:266-284). */
static inline rgbaf decode_raw(rgba16 v) {
  rgbaf o = {(float)v.r / 255.0f, (float)v.g / 255.0f, (float)v.b / 255.0f,
             (float)v.a / 255.0f};
  return o;
}
static inline rgba16 encode_raw(rgbaf v) {
  rgba16 o = {f2u(255.0f * v.r), f2u(255.0f * v.g), f2u(255.0f * v.b), f2u(255.0f * v.a)};
  return o;
}
static inline rgba16 clamp_store(rgba16 v) {
  /* @intCast(u8) on store: in range for valid premultiplied data. */
  rgba16 o = {(uint16_t)(v.r & 0xFF), (uint16_t)(v.g & 0xFF), (uint16_t)(v.b & 0xFF),
              (uint16_t)(v.a & 0xFF)};
  return o;
}

/* ------------------------------------------------------------------------
 * colors (src/color.zig:201-572)
 * ---------------------------------------------------------------------- */
typedef struct {
  uint32_t space;
  float v[4];
} zcolor;

static const float kGammaSRGB = 2.2f;

/* Color.init clamping (src/color.zig:81-92, :216-223, :416-423). */
static zcolor color_init(z2d_color c) {
  zcolor o;
  o.space = c.space;
  if (c.space == Z2D_COLOR_HSL) {
    float h = c.v[0];
    o.v[0] = (h < 0.0f || h > 360.0f) ? zmod_f(h, 360.0f) : h;
    o.v[1] = clamp_f(c.v[1], 0.0f, 1.0f);
    o.v[2] = clamp_f(c.v[2], 0.0f, 1.0f);
    o.v[3] = clamp_f(c.v[3], 0.0f, 1.0f);
  } else {
    for (int i = 0; i < 4; i++) o.v[i] = clamp_f(c.v[i], 0.0f, 1.0f);
  }
  return o;
}

/* HSL.toRGB (src/color.zig:470-488). */
static void hsl_to_rgb(const float hsla[4], float out[4]) {
  float hue = fmodf(hsla[0], 360.0f);
  if (hue < 0.0f) hue += 360.0f;
  const float ns[3] = {0.0f, 8.0f, 4.0f};
  for (int i = 0; i < 3; i++) {
    float k = fmodf(ns[i] + hue / 30.0f, 12.0f);
    float a = hsla[1] * zmin_f(hsla[2], 1.0f - hsla[2]);
    out[i] = hsla[2] - a * zmax_f(-1.0f, zmin_f(zmin_f(k - 3.0f, 9.0f - k), 1.0f));
  }
  out[3] = hsla[3];
}

/* HSL.fromRGB (src/color.zig:435-467). */
static void hsl_from_rgb(const float rgba[4], float out[4]) {
  float mx = zmax_f(rgba[0], zmax_f(rgba[1], rgba[2]));
  float mn = zmin_f(rgba[0], zmin_f(rgba[1], rgba[2]));
  float range = mx - mn;
  float light = (mn + mx) / 2.0f;
  float sat = (light == 0.0f || light == 1.0f) ? 0.0f
                                               : (mx - light) / zmin_f(light, 1.0f - light);
  float hue = 0.0f;
  if (range != 0.0f) {
    if (mx == rgba[0]) {
      hue = 60.0f * zmod_f((rgba[1] - rgba[2]) / range, 6.0f);
    } else if (mx == rgba[1]) {
      hue = 60.0f * ((rgba[2] - rgba[0]) / range + 2.0f);
    } else if (mx == rgba[2]) {
      hue = 60.0f * ((rgba[0] - rgba[1]) / range + 4.0f);
    }
  }
  if (sat < 0.0f) {
    hue += 180.0f;
    if (hue >= 360.0f) hue -= 360.0f;
    sat = fabsf(sat);
  }
  out[0] = hue;
  out[1] = sat;
  out[2] = light;
  out[3] = rgba[3];
}

/* RGB(profile).fromColor (src/color.zig:226-248): returns linear (srgb=0) or
 * gamma-encoded sRGB (srgb=1) values, straight alpha. */
static void color_to_rgb(zcolor c, int srgb, float out[4]) {
  if (c.space == Z2D_COLOR_HSL) {
    hsl_to_rgb(c.v, out);
    if (srgb) {
      float ig = 1.0f / kGammaSRGB;
      for (int i = 0; i < 3; i++) out[i] = zo_powf(out[i], ig);
    }
    return;
  }
  int is_srgb = c.space == Z2D_COLOR_SRGB;
  if (is_srgb == srgb) {
    memcpy(out, c.v, sizeof(float) * 4);
    return;
  }
  /* removeGamma of the source profile, then applyGamma of the target. */
  float g_from = is_srgb ? kGammaSRGB : 1.0f;
  float g_to = srgb ? kGammaSRGB : 1.0f;
  float ig_to = 1.0f / g_to;
  for (int i = 0; i < 3; i++) out[i] = zo_powf(zo_powf(c.v[i], g_from), ig_to);
  out[3] = c.v[3];
}

/* HSL.fromColor (src/color.zig:426-431). */
static void color_to_hsl(zcolor c, float out[4]) {
  if (c.space == Z2D_COLOR_HSL) {
    memcpy(out, c.v, sizeof(float) * 4);
    return;
  }
  float lin[4];
  color_to_rgb(c, 0, lin);
  hsl_from_rgb(lin, out);
}

static inline float lerp_f(float a, float b, float t) { return a + (b - a) * t; }

/* Gradient interpolation for one pixel.  encode=1: interpolateEncodeVec
 * (src/internal/color_vector.zig:59-83) -> pre-multiplied RGBA16.
 * encode=0: interpolateVec (:29-56) -> linear straight-alpha f32. */
static void interp_rgb_lerp(const float a[4], const float b[4], float t, float out[4]) {
  float am[4] = {a[0] * a[3], a[1] * a[3], a[2] * a[3], a[3]};
  float bm[4] = {b[0] * b[3], b[1] * b[3], b[2] * b[3], b[3]};
  for (int i = 0; i < 4; i++) out[i] = lerp_f(am[i], bm[i], t);
}
static void demul4(float v[4]) {
  if (v[3] == 0.0f) {
    v[0] = v[1] = v[2] = 0.0f;
  } else {
    v[0] = v[0] / v[3];
    v[1] = v[1] / v[3];
    v[2] = v[2] / v[3];
  }
}
/* LinearRGB.encodeRGBAVec: round(255*x) then u16 premultiply. */
static rgba16 encode_premul(const float v[4]) {
  rgba16 o = {f2u(255.0f * v[0]), f2u(255.0f * v[1]), f2u(255.0f * v[2]), f2u(255.0f * v[3])};
  o.r = U16((uint32_t)o.r * o.a / 255u);
  o.g = U16((uint32_t)o.g * o.a / 255u);
  o.b = U16((uint32_t)o.b * o.a / 255u);
  return o;
}

/* HSL.interpolateVec (src/internal/color_vector.zig:375-427). */
static void hsl_interp(const float a[4], const float b[4], float t, uint32_t polar,
                       float out[4]) {
  float am[4] = {a[0], a[1] * a[3], a[2] * a[3], a[3]};
  float bm[4] = {b[0], b[1] * b[3], b[2] * b[3], b[3]};
  float ah = am[0], bh = bm[0];
  switch (polar) {
    case Z2D_POLAR_SHORTER: {
      int gt = (bh - ah) > 180.0f;
      int lt = (bh - ah) < -180.0f;
      if (gt) ah = ah + 360.0f;
      if (lt) bh = bh + 360.0f;
    } break;
    case Z2D_POLAR_LONGER: {
      float delta = bh - ah;
      int c1 = (0.0f < delta) && (delta < 180.0f);
      int c2 = (-180.0f < delta) && (delta <= 0.0f);
      if (c1) ah = ah + 360.0f;
      if (c2) bh = bh + 360.0f;
    } break;
    case Z2D_POLAR_INCREASING:
      if (bh < ah) bh = bh + 360.0f;
      break;
    default:
      if (ah < bh) ah = ah + 360.0f;
      break;
  }
  /* lerpPolar(a_mul.h, a_mul_h, b_mul_h, t) = a.h + (b_h' - a_h') * t */
  float h = am[0] + (bh - ah) * t;
  out[0] = zmod_f(h, 360.0f);
  out[1] = lerp_f(am[1], bm[1], t);
  out[2] = lerp_f(am[2], bm[2], t);
  out[3] = lerp_f(am[3], bm[3], t);
  if (out[3] == 0.0f) {
    out[1] = 0.0f;
    out[2] = 0.0f;
  } else {
    out[1] = out[1] / out[3];
    out[2] = out[2] / out[3];
  }
}

/* HSL.toRGBVec (src/internal/color_vector.zig:319-345) == scalar toRGB. */

static void gradient_interp(uint32_t method, uint32_t polar, zcolor c0, zcolor c1, float t,
                            int encode, rgba16 *out_i, float out_f[4]) {
  float a[4], b[4], v[4];
  switch (method) {
    case Z2D_INTERP_LINEAR_RGB:
      color_to_rgb(c0, 0, a);
      color_to_rgb(c1, 0, b);
      interp_rgb_lerp(a, b, t, v);
      if (encode) {
        rgba16 o = {f2u(255.0f * v[0]), f2u(255.0f * v[1]), f2u(255.0f * v[2]),
                    f2u(255.0f * v[3])};
        *out_i = o;
      } else {
        demul4(v);
        memcpy(out_f, v, sizeof(v));
      }
      return;
    case Z2D_INTERP_SRGB:
      color_to_rgb(c0, 1, a);
      color_to_rgb(c1, 1, b);
      interp_rgb_lerp(a, b, t, v);
      demul4(v);
      for (int i = 0; i < 3; i++) v[i] = zo_powf(v[i], kGammaSRGB);
      if (encode) {
        *out_i = encode_premul(v);
      } else {
        memcpy(out_f, v, sizeof(v));
      }
      return;
    default: {
      color_to_hsl(c0, a);
      color_to_hsl(c1, b);
      float h[4];
      hsl_interp(a, b, t, polar, h);
      hsl_to_rgb(h, v);
      if (encode) {
        *out_i = encode_premul(v);
      } else {
        memcpy(out_f, v, sizeof(v));
      }
      return;
    }
  }
}

z2d_pixel zo_interpolate_encode(uint32_t method, uint32_t polar, const z2d_color *a,
                                const z2d_color *b, float t) {
  rgba16 v = {0, 0, 0, 0};
  float f[4];
  gradient_interp(method, polar, color_init(*a), color_init(*b), t, 1, &v, f);
  z2d_pixel o = {Z2D_FORMAT_RGBA, (uint8_t)v.r, (uint8_t)v.g, (uint8_t)v.b, (uint8_t)v.a};
  return o;
}

/* ------------------------------------------------------------------------
 * gradients (src/gradient.zig)
 * ---------------------------------------------------------------------- */
typedef struct {
  float offset;
  zcolor color;
} zstop;

static int is_identity(const double m[6]) {
  return m[0] == 1.0 && m[1] == 0.0 && m[2] == 0.0 && m[3] == 1.0 && m[4] == 0.0 &&
         m[5] == 0.0;
}

/* Transformation.userToDevice (src/Transformation.zig:194-206). */
static void user_to_device(const double m[6], double *x, double *y) {
  double ix = *x, iy = *y;
  double ox = m[0] * ix + m[1] * iy;
  double oy = m[2] * ix + m[3] * iy;
  *x = ox + m[4];
  *y = oy + m[5];
}

static inline double dot2(double a0, double a1, double b0, double b1) {
  double r = 0.0;
  r += a0 * b0;
  r += a1 * b1;
  return r;
}
static inline double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
  double r = 0.0;
  r += a0 * b0;
  r += a1 * b1;
  r += a2 * b2;
  return r;
}

/* Linear/Radial/Conic.getOffset (src/gradient.zig:349-375, :605-652,
 * :731-742); radial pre-calculations from Radial.initBuffer (:440-470). */
float zo_gradient_offset(const z2d_gradient_desc *g, int32_t x, int32_t y) {
  double px = (double)x + 0.5, py = (double)y + 0.5;
  switch (g->type) {
    case Z2D_GRADIENT_LINEAR: {
      if (!is_identity(g->matrix)) user_to_device(g->matrix, &px, &py);
      double dx = g->geom[2] - g->geom[0];
      double dy = g->geom[3] - g->geom[1];
      double dist = dot2(dx, dy, dx, dy);
      if (dist == 0.0) return -1.0f;
      double inv = 1.0 / dist;
      double spx = px - g->geom[0], spy = py - g->geom[1];
      return (float)clamp_d(dot2(dx, dy, spx, spy) * inv, 0.0, 1.0);
    }
    case Z2D_GRADIENT_RADIAL: {
      double r1 = zmax_d(0.0, g->geom[2]);
      double r2 = zmax_d(0.0, g->geom[5]);
      if (r1 == 0.0 && r2 == 0.0) return -1.0f;
      double cdx = g->geom[3] - g->geom[0];
      double cdy = g->geom[4] - g->geom[1];
      double dr = r2 - r1;
      double min_dr = -r1;
      double a = dot3(cdx, cdy, dr, cdx, cdy, -dr);
      double inv_a = (a != 0.0) ? 1.0 / a : 0.0;
      if (!is_identity(g->matrix)) user_to_device(g->matrix, &px, &py);
      double pdx = px - g->geom[0], pdy = py - g->geom[1];
      double b = dot3(pdx, pdy, r1, cdx, cdy, dr);
      double c = dot3(pdx, pdy, -r1, pdx, pdy, r1);
      if (a == 0.0) {
        if (b == 0.0) return -1.0f;
        double t = 0.5 * c / b;
        if (t * dr >= min_dr) return (float)clamp_d(t, 0.0, 1.0);
        return -1.0f;
      }
      double discr = dot2(b, a, b, -c);
      if (discr >= 0.0) {
        double sq = sqrt(discr);
        double t0 = (b + sq) * inv_a;
        double t1 = (b - sq) * inv_a;
        if (t0 * dr >= min_dr) return (float)clamp_d(t0, 0.0, 1.0);
        if (t1 * dr >= min_dr) return (float)clamp_d(t1, 0.0, 1.0);
      }
      return -1.0f;
    }
    default: {
      const double two_pi = 3.14159265358979323846 * 2.0;
      double angle0 = zmod_d(g->geom[2], two_pi);
      if (!is_identity(g->matrix)) user_to_device(g->matrix, &px, &py);
      double dx = px - g->geom[0], dy = py - g->geom[1];
      double ang = zmod_d(zo_atan2(dy, dx) - angle0, two_pi);
      return (float)(ang / two_pi);
    }
  }
}

/* Stop.List.addAssumeCapacity: clamp + stable sort by offset
 * (src/gradient.zig:818-830). */
static int build_stops(const z2d_gradient_desc *g, zstop *out) {
  int n = (int)g->n_stops;
  for (int i = 0; i < n; i++) {
    zstop s;
    s.offset = clamp_f(g->stops[i].offset, 0.0f, 1.0f);
    s.color = color_init(g->stops[i].color);
    /* insertion keeps equal offsets in insertion order */
    int j = i;
    while (j > 0 && out[j - 1].offset > s.offset) {
      out[j] = out[j - 1];
      j--;
    }
    out[j] = s;
  }
  return n;
}

/* Stop.List.search (src/gradient.zig:842-905). */
static void stops_search(const zstop *st, int n, float offset, zcolor *c0, zcolor *c1,
                         float *t) {
  if (offset < 0.0f || n == 0) {
    zcolor z = {Z2D_COLOR_LINEAR_RGB, {0, 0, 0, 0}};
    *c0 = z;
    *c1 = z;
    *t = 0.0f;
    return;
  }
  float o = zmin_f(offset, 1.0f);
  size_t left = 0, right = (size_t)n, mid = 0;
  while (left < right) {
    mid = left + (right - left) / 2;
    if (o >= st[mid].offset && (mid == (size_t)n - 1 || o <= st[mid + 1].offset)) break;
    if (o < st[mid].offset) {
      right = mid;
      continue;
    }
    if (o > st[mid].offset) {
      left = mid + 1;
      continue;
    }
  }
  if (mid == (size_t)n - 1) {
    *c0 = st[mid].color;
    *c1 = st[mid].color;
    *t = o - st[mid].offset;
    return;
  }
  if (mid == 0 && o < st[mid].offset) {
    *c0 = st[mid].color;
    *c1 = st[mid].color;
    *t = o / st[mid].offset;
    return;
  }
  float start = st[mid].offset, end = st[mid + 1].offset;
  float rl = end - start;
  *c0 = st[mid].color;
  *c1 = st[mid + 1].color;
  *t = (rl != 0.0f) ? (o - start) / rl : 0.0f;
}

#define ZO_MAX_STOPS 256

/* Stop.List.search over a gradient's stops as added (src/gradient.zig:797-905):
 * the two colors (after Color.init) and the offset between them. */
void zo_stop_search(const z2d_gradient_desc *g, float offset, z2d_color *c0, z2d_color *c1,
                    float *t) {
  zstop st[ZO_MAX_STOPS];
  int n = g->n_stops > ZO_MAX_STOPS ? 0 : build_stops(g, st);
  zcolor a, b;
  stops_search(st, n, offset, &a, &b, t);
  c0->space = a.space;
  c1->space = b.space;
  memcpy(c0->v, a.v, sizeof(a.v));
  memcpy(c1->v, b.v, sizeof(b.v));
}
static void gradient_sample(const z2d_gradient_desc *g, int32_t x, int32_t y, int encode,
                            rgba16 *oi, float of[4]) {
  zstop st[ZO_MAX_STOPS];
  int n = build_stops(g, st);
  zcolor c0, c1;
  float t;
  stops_search(st, n, zo_gradient_offset(g, x, y), &c0, &c1, &t);
  gradient_interp(g->method, g->polar, c0, c1, t, encode, oi, of);
}

z2d_pixel zo_gradient_pixel(const z2d_gradient_desc *g, int32_t x, int32_t y) {
  rgba16 v;
  gradient_sample(g, x, y, 1, &v, NULL);
  z2d_pixel p = {Z2D_FORMAT_RGBA, (uint8_t)v.r, (uint8_t)v.g, (uint8_t)v.b, (uint8_t)v.a};
  return p;
}

/* ------------------------------------------------------------------------
 * dither (src/Dither.zig, src/internal/color_vector.zig:90-186, :482-522)
 * ---------------------------------------------------------------------- */
#include "blue_noise_table.h"

static float bayer8(int32_t x, int32_t y) {
  int32_t yy = y ^ x;
  uint32_t m = (uint32_t)(((yy & 1) << 5) | ((x & 1) << 4) | ((yy & 2) << 2) |
                          ((x & 2) << 1) | ((yy & 4) >> 1) | ((x & 4) >> 2));
  return (float)m * (2.0f / 128.0f) - (63.0f / 128.0f);
}
static float bluenoise64(int32_t x, int32_t y) {
  int32_t mx = ((x % 64) + 64) % 64, my = ((y % 64) + 64) % 64;
  uint32_t idx = (uint32_t)((mx << 6) | my);
  return (float)zo_blue_noise_64x64[idx] * (2.0f / 8192.0f) - (4095.0f / 8192.0f);
}

/* LinearRGB.decodeRGBA (src/color.zig:262-271): u16 demultiply then /255. */
static void decode_demul(rgba16 v, float out[4]) {
  uint16_t r = v.r, g = v.g, b = v.b;
  if (v.a == 0) {
    r = g = b = 0;
  } else {
    r = U16((uint32_t)v.r * 255u / v.a);
    g = U16((uint32_t)v.g * 255u / v.a);
    b = U16((uint32_t)v.b * 255u / v.a);
  }
  out[0] = (float)r / 255.0f;
  out[1] = (float)g / 255.0f;
  out[2] = (float)b / 255.0f;
  out[3] = (float)v.a / 255.0f;
}

/* fromDitherVec for one lane: linear straight-alpha f32 output. */
static void dither_sample(const z2d_dither_desc *d, int32_t x, int32_t y, float out[4]) {
  float c[4];
  switch (d->source) {
    case Z2D_DITHER_SRC_PIXEL:
      decode_demul(pixel_to_rgba(d->pixel), c);
      break;
    case Z2D_DITHER_SRC_COLOR:
      color_to_rgb(color_init(d->color), 0, c);
      break;
    default: {
      rgba16 unused;
      gradient_sample(d->gradient, x, y, 0, &unused, c);
    } break;
  }
  float m;
  switch (d->type) {
    case Z2D_DITHER_BAYER: m = bayer8(x, y); break;
    case Z2D_DITHER_BLUE_NOISE: m = bluenoise64(x, y); break;
    default: memcpy(out, c, sizeof(c)); return;
  }
  float scale = 1.0f / (float)((1 << d->scale) - 1);
  for (int i = 0; i < 4; i++) out[i] = clamp_f(c[i] + m * scale, 0.0f, 1.0f);
}

z2d_pixel zo_dither_pixel(const z2d_dither_desc *d, int32_t x, int32_t y) {
  float v[4];
  dither_sample(d, x, y, v);
  rgba16 e = encode_premul(v);
  z2d_pixel p = {Z2D_FORMAT_RGBA, (uint8_t)e.r, (uint8_t)e.g, (uint8_t)e.b, (uint8_t)e.a};
  return p;
}

/* ------------------------------------------------------------------------
 * compositor (src/compositor.zig:221-659)
 * ---------------------------------------------------------------------- */

/* A resolved per-row operation parameter. */
typedef struct {
  uint32_t kind;
  z2d_pixel pixel;
  const z2d_gradient_desc *gradient;
  const z2d_dither_desc *dither;
  const z2d_surface_desc *surface;
  int64_t base_idx; /* stride: surface pixel index of stride start */
  int32_t x, y;     /* gradient/dither: stride start co-ordinates */
} rparam;

typedef struct {
  uint32_t op;
  rparam dst, src;
} rop;

/* One lane of the stride compositor (StrideCompositor.runPrecision,
 * src/compositor.zig:547-621). */
static void eval_param_int(const rparam *p, int64_t j, rgba16 *out) {
  switch (p->kind) {
    case Z2D_PARAM_PIXEL: *out = pixel_to_rgba(p->pixel); break;
    case Z2D_PARAM_SURFACE: *out = load_px(p->surface, (uint64_t)(p->base_idx + j)); break;
    case Z2D_PARAM_GRADIENT: {
      rgba16 v;
      gradient_sample(p->gradient, p->x + (int32_t)j, p->y, 1, &v, NULL);
      *out = v;
    } break;
    case Z2D_PARAM_DITHER: {
      float f[4];
      dither_sample(p->dither, p->x + (int32_t)j, p->y, f);
      *out = encode_premul(f);
    } break;
    default: break;
  }
}
static void eval_param_float(const rparam *p, int64_t j, rgbaf *out) {
  switch (p->kind) {
    case Z2D_PARAM_PIXEL: *out = decode_raw(pixel_to_rgba(p->pixel)); break;
    case Z2D_PARAM_SURFACE:
      *out = decode_raw(load_px(p->surface, (uint64_t)(p->base_idx + j)));
      break;
    case Z2D_PARAM_GRADIENT: {
      float f[4];
      gradient_sample(p->gradient, p->x + (int32_t)j, p->y, 0, NULL, f);
      rgbaf v = {f[0], f[1], f[2], f[3]};
      *out = v;
    } break;
    case Z2D_PARAM_DITHER: {
      float f[4];
      dither_sample(p->dither, p->x + (int32_t)j, p->y, f);
      rgbaf v = {f[0], f[1], f[2], f[3]};
      *out = v;
    } break;
    default: break;
  }
}

static void stride_run(z2d_surface_desc *dst, int64_t dst_idx, int64_t len, const rop *ops,
                       int n_ops, uint32_t precision) {
  for (int64_t j = 0; j < len; j++) {
    if (precision == Z2D_PRECISION_INTEGER) {
      rgba16 cur = {0, 0, 0, 0}, s = {0, 0, 0, 0}, d = {0, 0, 0, 0};
      for (int k = 0; k < n_ops; k++) {
        if (ops[k].src.kind == Z2D_PARAM_NONE)
          s = cur;
        else
          eval_param_int(&ops[k].src, j, &s);
        if (ops[k].dst.kind == Z2D_PARAM_NONE)
          d = load_px(dst, (uint64_t)(dst_idx + j));
        else
          eval_param_int(&ops[k].dst, j, &d);
        cur = int_op(ops[k].op, d, s);
      }
      store_px(dst, (uint64_t)(dst_idx + j), clamp_store(cur));
    } else {
      rgbaf cur = {0, 0, 0, 0}, s = {0, 0, 0, 0}, d = {0, 0, 0, 0};
      for (int k = 0; k < n_ops; k++) {
        if (ops[k].src.kind == Z2D_PARAM_NONE)
          s = cur;
        else
          eval_param_float(&ops[k].src, j, &s);
        if (ops[k].dst.kind == Z2D_PARAM_NONE)
          d = decode_raw(load_px(dst, (uint64_t)(dst_idx + j)));
        else
          eval_param_float(&ops[k].dst, j, &d);
        cur = float_op(ops[k].op, d, s);
      }
      store_px(dst, (uint64_t)(dst_idx + j), clamp_store(encode_raw(cur)));
    }
  }
}

static rparam resolve_param(const z2d_comp_param *p, int32_t sx, int32_t sy, int32_t gx,
                            int32_t gy) {
  rparam r;
  memset(&r, 0, sizeof(r));
  r.kind = p->kind;
  r.pixel = p->pixel;
  r.gradient = p->gradient;
  r.dither = &p->dither;
  r.surface = p->surface;
  r.x = gx;
  r.y = gy;
  if (p->kind == Z2D_PARAM_SURFACE) {
    const z2d_surface_desc *s = p->surface;
    if (sx < 0 || sy < 0 || sx >= s->width || sy >= s->height) {
      /* getStride on out-of-range co-ordinates yields an empty stride;
       * reading it is the reference's undefined behaviour. */
      r.base_idx = 0;
    } else {
      r.base_idx = (int64_t)s->width * sy + sx;
    }
  }
  return r;
}

#define ZO_MAX_OPS 16
void zo_composite(z2d_surface_desc *dst, int32_t dst_x, int32_t dst_y,
                  const z2d_comp_op *ops, uint64_t n_ops, uint32_t precision) {
  if (n_ops == 0 || n_ops > ZO_MAX_OPS) return;
  if (dst_x >= dst->width || dst_y >= dst->height) return;
  for (uint64_t i = 0; i < n_ops; i++)
    if (op_requires_float(ops[i].op)) precision = Z2D_PRECISION_FLOAT;
  int32_t sw, sh;
  switch (ops[0].src.kind) {
    case Z2D_PARAM_NONE: return;
    case Z2D_PARAM_SURFACE:
      sw = ops[0].src.surface->width;
      sh = ops[0].src.surface->height;
      break;
    default:
      if (dst_x != 0 || dst_y != 0) return;
      sw = dst->width;
      sh = dst->height;
      break;
  }
  int32_t src_start_x = dst_x < 0 ? -dst_x : 0;
  int32_t src_start_y = dst_y < 0 ? -dst_y : 0;
  int32_t width = (sw + dst_x > dst->width) ? dst->width - dst_x : sw;
  int32_t height = (sh + dst_y > dst->height) ? dst->height - dst_y : sh;
  if (src_start_x >= width || src_start_y >= height) return;
  int32_t scan_hi = height > src_start_y ? height : src_start_y;
  int32_t dst_start_x = src_start_x + dst_x;
  int64_t swidth = width - src_start_x;
  if (swidth < 0) swidth = 0;
  rop rops[ZO_MAX_OPS];
  for (int32_t src_y = src_start_y; src_y < scan_hi; src_y++) {
    int32_t dst_start_y = src_y + dst_y;
    for (uint64_t k = 0; k < n_ops; k++) {
      rops[k].op = ops[k].op;
      rops[k].dst = resolve_param(&ops[k].dst, dst_start_x, dst_start_y, dst_start_x,
                                  dst_start_y);
      rops[k].src =
          resolve_param(&ops[k].src, src_start_x, src_y, dst_start_x, dst_start_y);
    }
    if (dst_start_x < 0 || dst_start_y < 0 || dst_start_x >= dst->width ||
        dst_start_y >= dst->height)
      continue;
    int64_t didx = (int64_t)dst->width * dst_start_y + dst_start_x;
    stride_run(dst, didx, swidth, rops, (int)n_ops, precision);
  }
}

z2d_pixel zo_run_pixel(uint32_t precision, z2d_pixel dst, z2d_pixel src, uint32_t op) {
  uint8_t buf[4] = {0, 0, 0, 0};
  z2d_surface_desc s = {dst.format, 1, 1, 0, buf, 4};
  paint_px(&s, 0, dst);
  /* fromPixel for both sides; dst re-encoded into its own format. */
  rgba16 di = pixel_to_rgba(dst), si = pixel_to_rgba(src);
  rgba16 r;
  if (precision == Z2D_PRECISION_INTEGER)
    r = int_op(op, di, si);
  else
    r = encode_raw(float_op(op, decode_raw(di), decode_raw(si)));
  store_px(&s, 0, clamp_store(r));
  z2d_pixel out = {dst.format, 0, 0, 0, 0};
  rgba16 back = load_px(&s, 0);
  out.r = (uint8_t)back.r;
  out.g = (uint8_t)back.g;
  out.b = (uint8_t)back.b;
  if (fmt_packed(dst.format)) {
    out.a = (uint8_t)packed_get(buf, 0, fmt_bits(dst.format));
  } else {
    out.a = (uint8_t)back.a;
  }
  if (dst.format == Z2D_FORMAT_RGB || dst.format == Z2D_FORMAT_XRGB) out.a = 0;
  return out;
}

/* Alpha.scaleIntoIntAlpha / shlr (src/pixel.zig:582-630): a `from`-bit
 * alpha value in a `to`-bit field.  A 1-bit source is multiplied to the
 * target's maximum; scaling down shifts right; scaling up repeats the bits
 * ((v << d) + v for a difference of 2 or 4, four copies for 6). */
static uint32_t shlr_bits(uint32_t v, int from, int to) {
  const uint32_t max_to = (1u << to) - 1u;
  if (from == 1) return v * max_to;
  if (v == 0 || from == to) return v;
  if (from > to) return v >> (from - to);
  const int d = to - from;
  if (d == 6) return (v << 6) | (v << 4) | (v << 2) | v;
  return (v << d) + v;
}

static int pixel_alpha_bits(uint32_t f) {
  return (f == Z2D_FORMAT_ALPHA4) ? 4 : (f == Z2D_FORMAT_ALPHA2) ? 2 : (f == Z2D_FORMAT_ALPHA1) ? 1 : 8;
}

z2d_pixel zo_pixel_convert(uint32_t to, z2d_pixel p) {
  z2d_pixel o = {to, 0, 0, 0, 0};
  if (to == p.format) return p;
  if (fmt_has_color(to)) {  /* RGB-family fromPixel (src/pixel.zig:398-432) */
    const int is_rgba = (to == Z2D_FORMAT_ARGB || to == Z2D_FORMAT_RGBA);
    if (fmt_has_color(p.format)) {
      o.r = p.r;
      o.g = p.g;
      o.b = p.b;
      if (is_rgba)
        o.a = (p.format == Z2D_FORMAT_ARGB || p.format == Z2D_FORMAT_RGBA) ? p.a : 255;
    } else if (is_rgba) {
      o.a = (uint8_t)shlr_bits(p.a, pixel_alpha_bits(p.format), 8);
    }
    return o;
  }
  /* Alpha-family fromPixel (src/pixel.zig:569-577): RGB is opaque. */
  const int tb = pixel_alpha_bits(to);
  if (p.format == Z2D_FORMAT_RGB || p.format == Z2D_FORMAT_XRGB)
    o.a = (uint8_t)((1u << tb) - 1u);
  else
    o.a = (uint8_t)shlr_bits(p.a, pixel_alpha_bits(p.format), tb);
  return o;
}

/* RGBA.multiply / ARGB.multiply (src/pixel.zig:483-491). */
z2d_pixel zo_pixel_multiply(z2d_pixel p) {
  z2d_pixel o = p;
  o.r = (uint8_t)((uint32_t)p.r * p.a / 255u);
  o.g = (uint8_t)((uint32_t)p.g * p.a / 255u);
  o.b = (uint8_t)((uint32_t)p.b * p.a / 255u);
  return o;
}

/* Pixel.fromColor (src/pixel.zig:119-121):
 * LinearRGB.fromColor(Color.init(c)).encodeRGBA() (src/color.zig:227-249,
 * :272-289): linear values, round(255 x), then RGBA.multiply. */
z2d_pixel zo_pixel_from_color(const z2d_color *c) {
  float lin[4];
  color_to_rgb(color_init(*c), 0, lin);
  rgba16 v = encode_premul(lin);
  z2d_pixel o = {Z2D_FORMAT_RGBA, (uint8_t)v.r, (uint8_t)v.g, (uint8_t)v.b, (uint8_t)v.a};
  return o;
}

void zo_sample_pattern_int(const z2d_pattern_desc *p, int32_t x, int32_t y,
                           uint16_t out[4]) {
  rparam r;
  memset(&r, 0, sizeof(r));
  r.x = x;
  r.y = y;
  switch (p->kind) {
    case Z2D_PATTERN_OPAQUE: r.kind = Z2D_PARAM_PIXEL; r.pixel = p->pixel; break;
    case Z2D_PATTERN_GRADIENT: r.kind = Z2D_PARAM_GRADIENT; r.gradient = p->gradient; break;
    default: r.kind = Z2D_PARAM_DITHER; r.dither = &p->dither; break;
  }
  rgba16 v = {0, 0, 0, 0};
  eval_param_int(&r, 0, &v);
  out[0] = v.r; out[1] = v.g; out[2] = v.b; out[3] = v.a;
}

/* Pattern as a stride-compositor parameter starting at (x, y). */
static rparam pattern_param(const z2d_pattern_desc *p, int32_t x, int32_t y) {
  rparam r;
  memset(&r, 0, sizeof(r));
  r.x = x;
  r.y = y;
  switch (p->kind) {
    case Z2D_PATTERN_OPAQUE: r.kind = Z2D_PARAM_PIXEL; r.pixel = p->pixel; break;
    case Z2D_PATTERN_GRADIENT: r.kind = Z2D_PARAM_GRADIENT; r.gradient = p->gradient; break;
    default: r.kind = Z2D_PARAM_DITHER; r.dither = &p->dither; break;
  }
  return r;
}

/* fillReducesToSource (src/internal/raster/shared.zig:111-117). */
static int fill_reduces_to_source(uint32_t op, z2d_pixel px) {
  if (op == Z2D_OP_SRC) return 1;
  if (op == Z2D_OP_SRC_OVER) return pixel_is_opaque(px);
  return 0;
}

/* compositeOpaque (src/internal/raster/shared.zig:9-46). */
static void composite_opaque(uint32_t op, z2d_surface_desc *s, const z2d_pattern_desc *p,
                             int32_t x, int32_t y, uint64_t len, uint32_t precision) {
  if (op == Z2D_OP_CLEAR) {
    clear_stride(s, x, y, len);
  } else if (p->kind == Z2D_PATTERN_OPAQUE && fill_reduces_to_source(op, p->pixel)) {
    paint_stride(s, x, y, len, p->pixel);
  } else {
    if (x < 0 || y < 0 || x >= s->width || y >= s->height || len == 0) return;
    rop o;
    memset(&o, 0, sizeof(o));
    o.op = op;
    o.src = pattern_param(p, x, y);
    o.dst.kind = Z2D_PARAM_NONE;
    stride_run(s, (int64_t)s->width * y + x, (int64_t)len, &o, 1, precision);
  }
}

/* compositeOpacity (src/internal/raster/shared.zig:50-104); the fast path
 * is Surface.compositeStride (src/surface.zig:557-581, :852-876), always
 * integer. */
static void composite_opacity(uint32_t op, z2d_surface_desc *s, const z2d_pattern_desc *p,
                              int32_t x, int32_t y, uint64_t len, uint32_t precision,
                              uint8_t opacity) {
  if (op == Z2D_OP_CLEAR) {
    clear_stride(s, x, y, len);
  } else if (p->kind == Z2D_PATTERN_OPAQUE && fill_reduces_to_source(op, p->pixel)) {
    if (x < 0 || y < 0 || x >= s->width || y >= s->height) return;
    rgba16 src = pixel_to_rgba(p->pixel);
    if (opacity < 255) {
      rgba16 m = {0, 0, 0, opacity};
      src = int_op(Z2D_OP_DST_IN, src, m);
    }
    uint64_t start = (uint64_t)s->width * (uint64_t)y + (uint64_t)x;
    for (uint64_t i = start; i < start + len; i++) {
      rgba16 d = load_px(s, i);
      store_px(s, i, clamp_store(int_op(op, d, src)));
    }
  } else {
    if (x < 0 || y < 0 || x >= s->width || y >= s->height || len == 0) return;
    rop o[2];
    memset(o, 0, sizeof(o));
    o[0].op = Z2D_OP_DST_IN;
    o[0].dst = pattern_param(p, x, y);
    o[0].src.kind = Z2D_PARAM_PIXEL;
    z2d_pixel mpx = {Z2D_FORMAT_ALPHA8, 0, 0, 0, opacity};
    o[0].src.pixel = mpx;
    o[1].op = op;
    o[1].src.kind = Z2D_PARAM_NONE;
    o[1].dst.kind = Z2D_PARAM_NONE;
    stride_run(s, (int64_t)s->width * y + x, (int64_t)len, o, 2, precision);
  }
}

/* ------------------------------------------------------------------------
 * tessellation: fill_plotter + Spline (src/internal/tess/)
 * ---------------------------------------------------------------------- */
typedef struct {
  double x, y;
} zpt;

typedef struct {
  z2d_edge *edges;
  uint64_t n, cap;
  int overflow;
  double scale;
  double top, bottom, left, right;
} zpoly;

/* Polygon.addEdge (src/internal/tess/Polygon.zig:61-109). */
static void poly_add_edge(zpoly *p, zpt p0, zpt p1) {
  zpt a = {p0.x * p->scale, p0.y * p->scale};
  zpt b = {p1.x * p->scale, p1.y * p->scale};
  z2d_edge e;
  if (a.y < b.y) {
    e.y0 = a.y; e.y1 = b.y; e.x_start = a.x; e.x_inc = (b.x - a.x) / (b.y - a.y);
  } else if (a.y > b.y) {
    e.y0 = a.y; e.y1 = b.y; e.x_start = b.x; e.x_inc = (a.x - b.x) / (a.y - b.y);
  } else {
    return;
  }
  double top = e.y0 < e.y1 ? e.y0 : e.y1;
  double bottom = e.y0 < e.y1 ? e.y1 : e.y0;
  double left = a.x < b.x ? a.x : b.x;
  double right = a.x < b.x ? b.x : a.x;
  if (p->n == 0) {
    p->top = top; p->bottom = bottom; p->left = left; p->right = right;
  } else {
    if (top < p->top) p->top = top;
    if (bottom > p->bottom) p->bottom = bottom;
    if (left < p->left) p->left = left;
    if (right > p->right) p->right = right;
  }
  if (p->n >= p->cap) {
    p->overflow = 1;
    p->n++;
    return;
  }
  p->edges[p->n++] = e;
}

/* PointBuffer(1, 3) (src/internal/tess/point_buffer.zig:14-58). */
typedef struct {
  zpt items[3];
  int len;
} ptbuf;
static void pb_add(ptbuf *b, zpt p) {
  if (b->len < 3) {
    b->items[b->len++] = p;
  } else {
    b->items[1] = b->items[2];
    b->items[2] = p;
  }
}
static inline int pt_eq(zpt a, zpt b) { return a.x == b.x && a.y == b.y; }

/* SplinePlotterCtx.line_to (src/internal/tess/fill_plotter.zig:104-118). */
static int spline_line_to(zpoly *poly, ptbuf *pts, zpt p) {
  if (pts->len == 0) return Z2D_ERR_INVALID_STATE;
  zpt last = pts->items[pts->len - 1];
  if (!pt_eq(last, p)) {
    poly_add_edge(poly, last, p);
    pb_add(pts, p);
  }
  return 0;
}

typedef struct {
  zpt a, b, c, d;
} knots;

static inline double dotsq(double x, double y) { return x * x + y * y; }
static inline zpt lerp_half(zpt a, zpt b) {
  zpt r = {a.x + ((b.x - a.x) / 2.0), a.y + ((b.y - a.y) / 2.0)};
  return r;
}
/* Knots.errorSq (src/internal/tess/Spline.zig:83-123). */
static double knots_error_sq(const knots *k) {
  double bx = k->b.x - k->a.x, by = k->b.y - k->a.y;
  double cx = k->c.x - k->a.x, cy = k->c.y - k->a.y;
  if (k->a.x != k->d.x || k->a.y != k->d.y) {
    double dx = k->d.x - k->a.x, dy = k->d.y - k->a.y;
    double dd = dotsq(dx, dy);
    double bd = bx * dx + by * dy;
    if (bd >= dd) {
      bx -= dx;
      by -= dy;
    } else {
      bx -= bd / dd * dx;
      by -= bd / dd * dy;
    }
    double cd = cx * dx + cy * dy;
    if (cd >= dd) {
      cx -= dx;
      cy -= dy;
    } else {
      cx -= cd / dd * dx;
      cy -= cd / dd * dy;
    }
  }
  double be = dotsq(bx, by), ce = dotsq(cx, cy);
  return be > ce ? be : ce;
}
/* Knots.deCasteljau (src/internal/tess/Spline.zig:128-152). */
static knots knots_split(knots *k) {
  zpt ab = lerp_half(k->a, k->b);
  zpt bc = lerp_half(k->b, k->c);
  zpt cd = lerp_half(k->c, k->d);
  zpt abbc = lerp_half(ab, bc);
  zpt bccd = lerp_half(bc, cd);
  zpt fin = lerp_half(abbc, bccd);
  knots r = {fin, bccd, cd, k->d};
  k->b = ab;
  k->c = abbc;
  k->d = fin;
  return r;
}
/* Spline.decomposeInto (src/internal/tess/Spline.zig:56-71). */
static int spline_decompose_into(zpoly *poly, ptbuf *pts, knots *s1, zpt start, double tol,
                                 int depth) {
  if (knots_error_sq(s1) < tol || depth > 64) {
    if (!pt_eq(s1->a, start)) return spline_line_to(poly, pts, s1->a);
    return 0;
  }
  knots s2 = knots_split(s1);
  int rc = spline_decompose_into(poly, pts, s1, start, tol, depth + 1);
  if (rc) return rc;
  return spline_decompose_into(poly, pts, &s2, start, tol, depth + 1);
}
/* Spline.decompose (src/internal/tess/Spline.zig:37-53). */
static int spline_decompose(zpoly *poly, ptbuf *pts, zpt a, zpt b, zpt c, zpt d, double tol) {
  if (pt_eq(a, b) && pt_eq(c, d)) return spline_line_to(poly, pts, d);
  knots s1 = {a, b, c, d};
  int rc = spline_decompose_into(poly, pts, &s1, a, tol * tol, 0);
  if (rc) return rc;
  return spline_line_to(poly, pts, d);
}

/* fill_plotter.plot (src/internal/tess/fill_plotter.zig:21-97). */
static int plot_fill(const z2d_path_node *nodes, uint64_t n, double scale, double tol,
                     zpoly *poly) {
  ptbuf pts;
  pts.len = 0;
  poly->scale = scale;
  for (uint64_t i = 0; i < n; i++) {
    const z2d_path_node *nd = &nodes[i];
    zpt p0 = {nd->v[0], nd->v[1]};
    switch (nd->tag) {
      case Z2D_NODE_MOVE_TO:
        if (i == n - 1) return 0;
        pts.len = 0;
        pb_add(&pts, p0);
        break;
      case Z2D_NODE_LINE_TO:
        if (pts.len == 0) return Z2D_ERR_INVALID_STATE;
        if (!pt_eq(pts.items[pts.len - 1], p0)) {
          poly_add_edge(poly, pts.items[pts.len - 1], p0);
          pb_add(&pts, p0);
        }
        break;
      case Z2D_NODE_CURVE_TO: {
        if (pts.len == 0) return Z2D_ERR_INVALID_STATE;
        zpt p1 = {nd->v[0], nd->v[1]}, p2 = {nd->v[2], nd->v[3]}, p3 = {nd->v[4], nd->v[5]};
        int rc = spline_decompose(poly, &pts, pts.items[pts.len - 1], p1, p2, p3, tol);
        if (rc) return rc;
      } break;
      case Z2D_NODE_CLOSE_PATH:
        if (pts.len >= 3) {
          zpt last = pts.items[pts.len - 1], first = pts.items[0];
          if (pt_eq(last, first)) continue;
          poly_add_edge(poly, last, first);
          pb_add(&pts, first);
        }
        break;
      default: return Z2D_ERR_INVALID_ARGUMENT;
    }
  }
  return 0;
}

int64_t zo_plot_fill(const z2d_path_node *nodes, uint64_t n_nodes, double scale,
                     double tolerance, z2d_edge *edges_out, uint64_t cap,
                     double extents_out[4]) {
  zpoly poly;
  memset(&poly, 0, sizeof(poly));
  poly.edges = edges_out;
  poly.cap = cap;
  int rc = plot_fill(nodes, n_nodes, scale, tolerance, &poly);
  if (rc) return -(int64_t)rc;
  extents_out[0] = poly.top;
  extents_out[1] = poly.bottom;
  extents_out[2] = poly.left;
  extents_out[3] = poly.right;
  return (int64_t)poly.n;
}

/* PathNode.isClosedNodeSet (src/internal/path_nodes.zig:23-36). */
static int is_closed_node_set(const z2d_path_node *nodes, uint64_t n) {
  if (n == 0) return 0;
  int closed = 0;
  for (uint64_t i = 0; i < n; i++) {
    switch (nodes[i].tag) {
      case Z2D_NODE_MOVE_TO:
        if (!closed && i != 0) return closed;
        break;
      case Z2D_NODE_CLOSE_PATH: closed = 1; break;
      default: closed = 0; break;
    }
  }
  return closed;
}

int zo_is_closed_node_set(const z2d_path_node *nodes, uint64_t n) {
  return is_closed_node_set(nodes, n);
}

/* ------------------------------------------------------------------------
 * Polygon.inBox and WorkingEdgeSet (src/internal/tess/Polygon.zig:142-354)
 * ---------------------------------------------------------------------- */
int zo_in_box(const double ext[4], double scale, int32_t w, int32_t h) {
  double top = ext[0], bottom = ext[1], left = ext[2], right = ext[3];
  if (!isfinite(left) || !isfinite(top) || !isfinite(right) || !isfinite(bottom)) return 0;
  if (!isfinite(scale) || scale < 1.0) return 0;
  if (w < 1 || h < 1) return 0;
  if (right < 0.0 || bottom < 0.0) return 0;
  int32_t sx = (int32_t)floor(left / scale);
  int32_t sy = (int32_t)floor(top / scale);
  int32_t ex = (int32_t)ceil(right / scale);
  int32_t ey = (int32_t)ceil(bottom / scale);
  int32_t pw = ex - sx, ph = ey - sy;
  if (pw < 0 || ph < 0) return 0;
  if (pw == 0 || ph == 0) return 0;
  if (sx + pw < 0 || sy + ph < 0) return 0;
  if (sx >= w || sy >= h) return 0;
  return 1;
}

static inline double e_top(const z2d_edge *e) { return e->y0 < e->y1 ? e->y0 : e->y1; }
static inline double e_bottom(const z2d_edge *e) { return e->y0 < e->y1 ? e->y1 : e->y0; }
static inline int e_dir(const z2d_edge *e) { return e->y0 < e->y1 ? -1 : 1; }

typedef struct {
  z2d_edge *all; /* the polygon's edge array, permuted in place */
  uint64_t n_all;
  uint64_t n_active;
  int32_t *x;
  int32_t *bp;
  uint64_t n_bp;
} wes;

static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* breakpoints (src/internal/tess/Polygon.zig:232-273): sorted unique
 * round(top)/round(bottom). */
static void wes_init(wes *w, z2d_edge *edges, uint64_t n) {
  w->all = edges;
  w->n_all = n;
  w->n_active = 0;
  w->x = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
  w->bp = (int32_t *)malloc(sizeof(int32_t) * (2 * n + 1));
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; i++) {
    w->bp[k++] = (int32_t)round(e_top(&edges[i]));
    w->bp[k++] = (int32_t)round(e_bottom(&edges[i]));
  }
  qsort(w->bp, k, sizeof(int32_t), cmp_i32);
  uint64_t u = 0;
  for (uint64_t i = 0; i < k; i++)
    if (u == 0 || w->bp[u - 1] != w->bp[i]) w->bp[u++] = w->bp[i];
  w->n_bp = u;
}
static void wes_free(wes *w) {
  free(w->x);
  free(w->bp);
}
/* rescan (src/internal/tess/Polygon.zig:275-300). */
static void wes_rescan(wes *w, int32_t line_y) {
  if (w->n_all == 0) return;
  double mid = (double)line_y + 0.5;
  uint64_t to = 0;
  for (uint64_t from = 0; from < w->n_all; from++) {
    if (e_top(&w->all[from]) < mid && e_bottom(&w->all[from]) >= mid) {
      if (from != to) {
        z2d_edge t = w->all[to];
        w->all[to] = w->all[from];
        w->all[from] = t;
      }
      to++;
    }
  }
  w->n_active = to;
}
/* inc (src/internal/tess/Polygon.zig:302-307): f64, no FMA. */
static void wes_inc(wes *w, int32_t y) {
  double ym = (double)y + 0.5;
  for (uint64_t i = 0; i < w->n_active; i++) {
    const z2d_edge *e = &w->all[i];
    w->x[i] = (int32_t)round(e->x_start + (e->x_inc * (ym - e_top(e))));
  }
}
/* sort (src/internal/tess/Polygon.zig:309-324): the reference's pdqsort
 * is an insertion sort for up to 24 elements; insertion sort is used here
 * for every length (see the oracle's pinning notes). */
static void wes_sort(wes *w) {
  for (uint64_t i = 1; i < w->n_active; i++) {
    uint64_t j = i;
    while (j > 0 && w->x[j] < w->x[j - 1]) {
      int32_t tx = w->x[j];
      w->x[j] = w->x[j - 1];
      w->x[j - 1] = tx;
      z2d_edge te = w->all[j];
      w->all[j] = w->all[j - 1];
      w->all[j - 1] = te;
      j--;
    }
  }
}
/* filter (src/internal/tess/Polygon.zig:326-353). */
static uint64_t wes_filter(wes *w, uint32_t rule) {
  if (rule == Z2D_FILL_EVEN_ODD) return w->n_active;
  int32_t wn = 0;
  uint64_t to = 0;
  for (uint64_t from = 0; from < w->n_active; from++) {
    w->x[to] = w->x[from];
    if (wn == 0) {
      wn += e_dir(&w->all[from]);
      to++;
    } else {
      wn += e_dir(&w->all[from]);
      if (wn == 0) to++;
    }
  }
  return to;
}
/* initial breakpoint index (e.g. src/internal/raster/multisample.zig:121-130) */
static int wes_start_idx(const wes *w, int32_t start, uint64_t *idx) {
  for (uint64_t i = 0; i < w->n_bp; i++) {
    if (w->bp[i] >= start) {
      *idx = i > 0 ? i - 1 : 0;
      return 1;
    }
  }
  return 0;
}
static void wes_step(wes *w, uint64_t *bpi, int32_t y) {
  if (y >= w->bp[*bpi]) {
    wes_rescan(w, y);
    if (*bpi < w->n_bp - 1) (*bpi)++;
  }
}

/* ------------------------------------------------------------------------
 * rasterizers
 * ---------------------------------------------------------------------- */

/* direct.run (src/internal/raster/direct.zig:14-127). */
static void raster_direct(z2d_surface_desc *s, const z2d_pattern_desc *p, zpoly *poly,
                          uint32_t rule, uint32_t op, uint32_t precision) {
  int32_t W = s->width, H = s->height;
  uint32_t prec = op_requires_float(op) ? Z2D_PRECISION_FLOAT : precision;
  int bounded = op_is_bounded(op);
  double ext[4] = {poly->top, poly->bottom, poly->left, poly->right};
  if (!zo_in_box(ext, 1.0, W, H)) return;
  int32_t psy = bounded ? (int32_t)floor(poly->top) : 0;
  int32_t pey = bounded ? (int32_t)ceil(poly->bottom) : H - 1;
  int32_t sy = clamp_i32(psy, 0, H - 1);
  int32_t ey = clamp_i32(pey, sy, H - 1);
  wes w;
  wes_init(&w, poly->edges, poly->n);
  uint64_t bpi;
  if (!wes_start_idx(&w, sy, &bpi)) {
    wes_free(&w);
    return;
  }
  for (int32_t y = (sy > 0 ? sy : 0); y <= (ey > 0 ? ey : 0); y++) {
    wes_step(&w, &bpi, y);
    wes_inc(&w, y);
    wes_sort(&w);
    uint64_t nf = wes_filter(&w, rule);
    if (!bounded && nf == 0) {
      clear_stride(s, 0, y, (uint64_t)W);
      continue;
    }
    for (uint64_t k = 0; k < nf / 2; k++) {
      int32_t st = w.x[2 * k] > 0 ? w.x[2 * k] : 0;
      if (st >= W) break;
      int32_t en = clamp_i32(w.x[2 * k + 1], st, W);
      int32_t fl = en - st, ecl = W - en;
      if (!bounded && st > 0) clear_stride(s, 0, y, (uint64_t)st);
      if (fl > 0) composite_opaque(op, s, p, st, y, (uint64_t)fl, prec);
      if (!bounded && ecl > 0) clear_stride(s, en, y, (uint64_t)ecl);
    }
  }
  wes_free(&w);
}

/* multisample.addSpan (src/internal/raster/multisample.zig:239-279) on a
 * dense per-pixel coverage row. */
static void msaa_add_span(uint8_t *cov, uint32_t x, uint32_t len) {
  if (len == 0) return;
  const uint32_t scale = 4;
  uint32_t sx = x / scale;
  uint32_t so = x - sx * scale;
  if (so == 0 && len >= scale) {
    uint32_t front = len / scale;
    for (uint32_t i = 0; i < front; i++) cov[sx + i] += scale;
    uint32_t endc = len - front * scale;
    if (endc > scale) endc = scale;
    if (endc > 0) cov[sx + front] += (uint8_t)endc;
  } else {
    uint32_t sc = scale - so;
    if (len < sc) sc = len;
    if (sc > scale) sc = scale;
    cov[sx] += (uint8_t)sc;
    uint32_t after = len - sc;
    uint32_t mid = after / scale;
    for (uint32_t i = 0; i < mid; i++) cov[sx + 1 + i] += scale;
    uint32_t endc = after - mid * scale;
    if (endc > scale) endc = scale;
    if (endc > 0) cov[sx + 1 + mid] += (uint8_t)endc;
  }
}

void zo_msaa_add_span(uint8_t *cov, uint32_t x, uint32_t len) { msaa_add_span(cov, x, len); }

/* multisample.run (src/internal/raster/multisample.zig:22-229). */
static void raster_msaa(z2d_surface_desc *s, const z2d_pattern_desc *p, zpoly *poly,
                        uint32_t rule, uint32_t op, uint32_t precision) {
  const int32_t scale = 4;
  int32_t W = s->width, H = s->height;
  uint32_t prec = op_requires_float(op) ? Z2D_PRECISION_FLOAT : precision;
  double ext[4] = {poly->top, poly->bottom, poly->left, poly->right};
  if (!zo_in_box(ext, scale, W, H)) return;
  int32_t sy = clamp_i32((int32_t)floor(poly->top / scale), 0, H - 1);
  int32_t ey = clamp_i32((int32_t)ceil(poly->bottom / scale), sy, H - 1);
  int32_t sx = clamp_i32((int32_t)floor(poly->left / scale), 0, W - 1);
  int32_t ex = clamp_i32((int32_t)ceil(poly->right / scale), sx, W);
  int32_t dw = ex - sx;
  if (dw < 1) return;
  int32_t sxs = sx * scale, dws = dw * scale;
  if (!op_is_bounded(op)) {
    for (int32_t y = 0; y < (sy > 0 ? sy : 0); y++) clear_stride(s, 0, y, (uint64_t)W);
    for (int32_t y = (ey + 1 > 0 ? ey + 1 : 0); y < (W > 0 ? W : 0); y++)
      clear_stride(s, 0, y, (uint64_t)W);
    for (int32_t y = (sy > 0 ? sy : 0); y <= (ey > 0 ? ey : 0); y++) {
      if (sx > 0) clear_stride(s, 0, y, (uint64_t)sx);
      if (ex < W) clear_stride(s, 0, y, (uint64_t)(W - ex));
    }
  }
  wes w;
  wes_init(&w, poly->edges, poly->n);
  uint64_t bpi;
  if (!wes_start_idx(&w, sy, &bpi)) {
    wes_free(&w);
    return;
  }
  uint8_t *cov = (uint8_t *)calloc((size_t)dw + 1, 1);
  for (int32_t y = (sy > 0 ? sy : 0); y <= (ey > 0 ? ey : 0); y++) {
    memset(cov, 0, (size_t)dw + 1);
    for (int32_t o = 0; o < 4; o++) {
      int32_t ys = y * scale + o;
      wes_step(&w, &bpi, ys);
      wes_inc(&w, ys);
      wes_sort(&w);
      uint64_t nf = wes_filter(&w, rule);
      int32_t xmin = 0;
      for (uint64_t k = 0; k < nf / 2; k++) {
        int32_t a = w.x[2 * k] - sxs;
        int32_t st = a > xmin ? a : xmin;
        if (st >= dws) break;
        int32_t en = clamp_i32(w.x[2 * k + 1] - sxs, st, dws);
        int32_t fl = en - st;
        if (fl > 0) msaa_add_span(cov, (uint32_t)(st > 0 ? st : 0), (uint32_t)fl);
        xmin = en;
      }
    }
    for (int32_t i = 0; i < dw; i++) {
      uint8_t c = cov[i] > 16 ? 16 : cov[i];
      int32_t x = sx + i;
      if (c == 0) continue;
      if (c == 16) {
        composite_opaque(op, s, p, x, y, 1, prec);
      } else {
        int32_t a = c * 16 - 1;
        composite_opacity(op, s, p, x, y, 1, prec, (uint8_t)clamp_i32(a, 0, 255));
      }
    }
  }
  free(cov);
  wes_free(&w);
}

/* supersample.run (src/internal/raster/supersample.zig:17-185) with
 * Surface.downsample (src/surface.zig:447-469, :687-709). */
static void raster_ssaa(z2d_surface_desc *s, const z2d_pattern_desc *p, zpoly *poly,
                        uint32_t rule, uint32_t op, uint32_t precision) {
  const int32_t sc = 4;
  int32_t W = s->width, H = s->height;
  double ext[4] = {poly->top, poly->bottom, poly->left, poly->right};
  if (!zo_in_box(ext, sc, W, H)) return;
  int bounded = op_is_bounded(op);
  int32_t x0 = bounded ? (int32_t)floor(poly->left / sc) : 0;
  int32_t y0 = bounded ? (int32_t)floor(poly->top / sc) : 0;
  int32_t x1 = bounded ? (int32_t)ceil(poly->right / sc) : W;
  int32_t y1 = bounded ? (int32_t)ceil(poly->bottom / sc) : H;
  int32_t tws = W * sc, ths = H * sc;
  int32_t bx0 = clamp_i32(x0 * sc, 0, tws - 1);
  int32_t by0 = clamp_i32(y0 * sc, 0, ths - 1);
  int32_t bx1 = clamp_i32(x1 * sc, bx0, tws - 1);
  int32_t by1 = clamp_i32(y1 * sc, by0, ths - 1);
  int32_t mw = bx1 + 1 - bx0, mh = by1 + 1 - by0;
  if (mw < 1 || mh < 1) return;
  uint32_t mfmt;
  switch (s->format) {
    case Z2D_FORMAT_ALPHA4:
    case Z2D_FORMAT_ALPHA2:
    case Z2D_FORMAT_ALPHA1: mfmt = s->format; break;
    default: mfmt = Z2D_FORMAT_ALPHA8; break;
  }
  int mbits = fmt_bits(mfmt);
  uint32_t opaque_v = (mbits == 8) ? 255u : ((1u << mbits) - 1u);
  /* mask samples kept unpacked (0 or opaque_v) */
  uint8_t *mask = (uint8_t *)calloc((size_t)mw * (size_t)mh, 1);
  wes w;
  wes_init(&w, poly->edges, poly->n);
  uint64_t bpi;
  if (!wes_start_idx(&w, by0, &bpi)) {
    wes_free(&w);
    free(mask);
    return;
  }
  for (int32_t y = 0; y < mh; y++) {
    int32_t dy = y + by0;
    wes_step(&w, &bpi, dy);
    wes_inc(&w, dy);
    wes_sort(&w);
    uint64_t nf = wes_filter(&w, rule);
    for (uint64_t k = 0; k < nf / 2; k++) {
      int32_t a = w.x[2 * k] - bx0;
      int32_t st = a > 0 ? a : 0;
      if (st >= mw) break;
      int32_t en = clamp_i32(w.x[2 * k + 1] - bx0, st, mw);
      /* paintStride on the mask: row-wrapping linear index */
      uint64_t base = (uint64_t)mw * (uint64_t)y;
      for (int32_t i = st; i < en; i++) mask[base + (uint64_t)i] = (uint8_t)opaque_v;
    }
  }
  wes_free(&w);
  /* downsample: integer average of 16 samples at the mask's depth */
  int32_t dw = mw / sc, dh = mh / sc;
  uint64_t dbytes = surface_bytes(mfmt, dw, dh);
  uint8_t *dmask = (uint8_t *)calloc((size_t)dbytes + 1, 1);
  z2d_surface_desc md = {mfmt, dw, dh, 0, dmask, dbytes};
  for (int32_t y = 0; y < dh; y++) {
    for (int32_t x = 0; x < dw; x++) {
      uint32_t sum = 0;
      for (int i = 0; i < sc; i++)
        for (int j = 0; j < sc; j++)
          sum += mask[(uint64_t)(y * sc + i) * (uint64_t)mw + (uint64_t)(x * sc + j)];
      uint32_t avg = sum / 16u;
      uint64_t idx = (uint64_t)y * (uint64_t)dw + (uint64_t)x;
      if (mbits == 8)
        dmask[idx] = (uint8_t)avg;
      else
        packed_set(dmask, idx, mbits, avg);
    }
  }
  free(mask);
  int32_t cx = x0 > 0 ? x0 : 0, cy = y0 > 0 ? y0 : 0;
  z2d_comp_op ops[2];
  memset(ops, 0, sizeof(ops));
  ops[0].op = Z2D_OP_DST_IN;
  switch (p->kind) {
    case Z2D_PATTERN_OPAQUE: ops[0].dst.kind = Z2D_PARAM_PIXEL; ops[0].dst.pixel = p->pixel; break;
    case Z2D_PATTERN_GRADIENT:
      ops[0].dst.kind = Z2D_PARAM_GRADIENT;
      ops[0].dst.gradient = p->gradient;
      break;
    default: ops[0].dst.kind = Z2D_PARAM_DITHER; ops[0].dst.dither = p->dither; break;
  }
  ops[0].src.kind = Z2D_PARAM_SURFACE;
  ops[0].src.surface = &md;
  ops[1].op = op;
  zo_composite(s, cx, cy, ops, 2, precision);
  free(dmask);
}

static int32_t raster_dispatch(z2d_surface_desc *s, const z2d_pattern_desc *p, zpoly *poly,
                               uint32_t aa, uint32_t rule, uint32_t op, uint32_t precision) {
  switch (aa) {
    case Z2D_AA_NONE: raster_direct(s, p, poly, rule, op, precision); break;
    case Z2D_AA_SUPERSAMPLE_4X: raster_ssaa(s, p, poly, rule, op, precision); break;
    default: raster_msaa(s, p, poly, rule, op, precision); break;
  }
  return Z2D_OK;
}

static int check_pattern(const z2d_pattern_desc *p) {
  if (p->kind == Z2D_PATTERN_OPAQUE &&
      (p->pixel.format == Z2D_FORMAT_ARGB || p->pixel.format == Z2D_FORMAT_RGBA) &&
      !can_demultiply(p->pixel))
    return Z2D_ERR_PIXEL_SOURCE_NOT_PREMULTIPLIED;
  return 0;
}

/* painter.fill (src/painter.zig:66-143). */
int32_t zo_fill(z2d_surface_desc *s, const z2d_pattern_desc *p, const z2d_path_node *nodes,
                uint64_t n, const z2d_fill_opts *o) {
  int rc = check_pattern(p);
  if (rc) return rc;
  if (n == 0) return Z2D_OK;
  if (!is_closed_node_set(nodes, n)) return Z2D_ERR_PATH_NOT_CLOSED;
  uint32_t aa = (s->format == Z2D_FORMAT_ALPHA1) ? Z2D_AA_NONE : o->anti_aliasing_mode;
  double scale = (aa == Z2D_AA_NONE) ? 1.0 : 4.0;
  double tol = o->tolerance > 0.001 ? o->tolerance : 0.001;
  uint64_t cap = 1024;
  for (;;) {
    zpoly poly;
    memset(&poly, 0, sizeof(poly));
    poly.cap = cap;
    poly.edges = (z2d_edge *)malloc(sizeof(z2d_edge) * cap);
    if (!poly.edges) return Z2D_ERR_OUT_OF_MEMORY;
    rc = plot_fill(nodes, n, scale, tol, &poly);
    if (rc) {
      free(poly.edges);
      return rc;
    }
    if (poly.overflow) {
      cap = poly.n + 16;
      free(poly.edges);
      continue;
    }
    raster_dispatch(s, p, &poly, aa, o->fill_rule, o->op, o->precision);
    free(poly.edges);
    return Z2D_OK;
  }
}

int32_t zo_fill_polygon(z2d_surface_desc *s, const z2d_pattern_desc *p, const z2d_edge *edges,
                        uint64_t n_edges, const double extents[4], uint32_t aa, uint32_t rule,
                        uint32_t op, uint32_t precision) {
  int rc = check_pattern(p);
  if (rc) return rc;
  zpoly poly;
  memset(&poly, 0, sizeof(poly));
  poly.edges = (z2d_edge *)malloc(sizeof(z2d_edge) * (n_edges ? n_edges : 1));
  memcpy(poly.edges, edges, sizeof(z2d_edge) * n_edges);
  poly.n = poly.cap = n_edges;
  poly.top = extents[0];
  poly.bottom = extents[1];
  poly.left = extents[2];
  poly.right = extents[3];
  poly.scale = (aa == Z2D_AA_NONE) ? 1.0 : 4.0;
  if (s->format == Z2D_FORMAT_ALPHA1) aa = Z2D_AA_NONE;
  raster_dispatch(s, p, &poly, aa, rule, op, precision);
  free(poly.edges);
  return Z2D_OK;
}

uint64_t zo_count_coverage(const z2d_edge *edges, uint64_t n_edges, const double ext[4],
                           uint32_t rule, int32_t W, int32_t H) {
  /* MSAA coverage count over a scratch alpha8 surface with src: every pixel
   * whose coverage is nonzero becomes nonzero. */
  uint64_t bytes = (uint64_t)W * (uint64_t)H;
  uint8_t *buf = (uint8_t *)calloc((size_t)bytes, 1);
  z2d_surface_desc s = {Z2D_FORMAT_ALPHA8, W, H, 0, buf, bytes};
  z2d_pattern_desc p;
  memset(&p, 0, sizeof(p));
  p.kind = Z2D_PATTERN_OPAQUE;
  z2d_pixel px = {Z2D_FORMAT_ALPHA8, 0, 0, 0, 255};
  p.pixel = px;
  zo_fill_polygon(&s, &p, edges, n_edges, ext, Z2D_AA_MULTISAMPLE_4X, rule, Z2D_OP_SRC,
                  Z2D_PRECISION_INTEGER);
  uint64_t c = 0;
  for (uint64_t i = 0; i < bytes; i++) c += buf[i] != 0;
  free(buf);
  return c;
}

/* ---------------------------------------------------------------------------
 * PNG scanline stream (src/export_png.zig:153-305, encodeRGBAVec :313-360).
 * Rows are walked in chunks of the reference's default vector_length (16,
 * build.zig:141-145); with 16 a multiple of 8/bits the packed formats come
 * out as continuous MSB-first rows padded to a byte at the row end.
 * ------------------------------------------------------------------------- */
#define ZO_PNG_VL 16

/* encodeRGBAVec for one pixel: demultiply (pixel_vector.RGBA16.demultiply,
 * src/internal/pixel_vector.zig:27-49) when the format has alpha, then the
 * sRGB profile's decodeRGBAVecRaw -> applyGammaVec -> encodeRGBAVecRaw
 * (src/internal/color_vector.zig:210-224, :266-284). */
static void png_encode_px(uint32_t c[4], int use_alpha, uint32_t profile) {
  if (use_alpha) {
    for (int k = 0; k < 3; k++) {
      uint32_t den = c[3] > 1 ? c[3] : 1;
      c[k] = c[3] == 0 ? 0 : (uint16_t)(c[k] * 255u) / den;
    }
  }
  if (profile == Z2D_PNG_PROFILE_SRGB) {
    const float ig = 1.0f / kGammaSRGB;
    for (int k = 0; k < 4; k++) {
      float v = (float)c[k] / 255.0f;
      if (k < 3) v = zo_powf(v, ig);
      c[k] = (uint32_t)roundf(255.0f * v);
    }
  }
}

uint64_t zo_png_scanlines(const z2d_surface_desc *s, uint32_t profile, uint8_t *out) {
  const int32_t w = s->width, h = s->height;
  int bits = 32, chans = 0;
  switch (s->format) {
    case Z2D_FORMAT_XRGB: case Z2D_FORMAT_RGB: chans = 3; break;
    case Z2D_FORMAT_ARGB: case Z2D_FORMAT_RGBA: chans = 4; break;
    case Z2D_FORMAT_ALPHA8: bits = 8; break;
    case Z2D_FORMAT_ALPHA4: bits = 4; break;
    case Z2D_FORMAT_ALPHA2: bits = 2; break;
    default: bits = 1; break;
  }
  const uint64_t row_bytes = chans ? (uint64_t)w * chans : ((uint64_t)w * bits + 7) / 8;
  const uint64_t total = (uint64_t)h * (1 + row_bytes);
  if (!out) return total;
  uint8_t *o = out;
  for (int32_t y = 0; y < h; y++) {
    *o++ = 0; /* filter type 0 */
    for (int32_t x = 0; x < w; x += ZO_PNG_VL) {
      const int32_t n = w - x < ZO_PNG_VL ? w - x : ZO_PNG_VL;
      if (chans) {
        for (int32_t i = 0; i < n; i++) {
          const uint8_t *p = s->data + ((uint64_t)y * w + x + i) * 4;
          uint32_t c[4];
          if (s->format == Z2D_FORMAT_RGBA || s->format == Z2D_FORMAT_RGB) {
            c[0] = p[0]; c[1] = p[1]; c[2] = p[2]; c[3] = p[3];
          } else { /* argb / xrgb bytes are b, g, r, a */
            c[0] = p[2]; c[1] = p[1]; c[2] = p[0]; c[3] = p[3];
          }
          if (chans == 3) c[3] = 255;
          png_encode_px(c, chans == 4, profile);
          for (int k = 0; k < chans; k++) *o++ = (uint8_t)c[k];
        }
      } else if (bits == 8) {
        memcpy(o, s->data + (uint64_t)y * w + x, (size_t)n);
        o += n;
      } else {
        const int scale = 8 / bits;
        const uint64_t nb = ((uint64_t)n * bits + 7) / 8;
        memset(o, 0, nb);
        for (int32_t i = 0; i < n; i++) {
          uint32_t v = packed_get(s->data, (uint64_t)y * w + x + i, bits);
          int sh = 8 - bits - (i % scale) * bits;
          o[i / scale] |= (uint8_t)(v << sh);
        }
        o += nb;
      }
    }
  }
  return total;
}
