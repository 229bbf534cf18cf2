// dev_pixel.h - device-side pixel formats, the 28 compositing operators and
// the gradient/dither paint sources for gfx950.
//
// Semantics follow the reference's vectorized stride compositor, which every
// fill and composite in the reference goes through:
//   formats     src/pixel.zig:47-680, src/compositor.zig:716-1037
//   integer ops src/compositor.zig:1158-1568 (u16 lanes, /255 truncating)
//   float ops   src/compositor.zig:1571-2440 (f32, vector @select semantics)
//   gradients   src/gradient.zig:349-905, src/internal/color_vector.zig:29-459
//   dither      src/internal/color_vector.zig:90-186, :482-522
// All arithmetic must be built with -ffp-contract=off: the reference never
// fuses a*b+c, and bit-exact parity depends on the same rounding sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/z2d_types.h"

// Z2D_HOST_PIXEL: the pixel functions are compiled for the host as well (a
// CPU harness runs them there, under the address and undefined-behaviour
// sanitizers, against the oracle).
#ifdef Z2D_HOST_PIXEL
#define Z2D_PX __host__ __device__
#else
#define Z2D_PX __device__
#endif

namespace z2d {

// 24-bit unsigned multiply (low 32 bits of the product of the low 24 bits)
// and the two 32-bit halves of a double, on either side.
Z2D_PX __forceinline__ uint32_t mul24(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul24(a, b);
#else
  return (a & 0xFFFFFFu) * (b & 0xFFFFFFu);
#endif
}
template <typename T>
Z2D_PX __forceinline__ bool z_isnan(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return isnan(x);
#else
  return __builtin_isnan(x);
#endif
}
template <typename T>
Z2D_PX __forceinline__ bool z_isinf(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return isinf(x);
#else
  return __builtin_isinf(x);
#endif
}
Z2D_PX __forceinline__ uint32_t dhi(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint32_t)__double2hiint(x);
#else
  uint64_t u;
  __builtin_memcpy(&u, &x, 8);
  return (uint32_t)(u >> 32);
#endif
}
Z2D_PX __forceinline__ uint32_t dlo(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint32_t)__double2loint(x);
#else
  uint64_t u;
  __builtin_memcpy(&u, &x, 8);
  return (uint32_t)u;
#endif
}

// ---------------------------------------------------------------------------
// small math helpers
// ---------------------------------------------------------------------------
Z2D_PX __forceinline__ float fminz(float a, float b) { return fminf(a, b); }
Z2D_PX __forceinline__ float fmaxz(float a, float b) { return fmaxf(a, b); }
Z2D_PX __forceinline__ float clampz(float v, float lo, float hi) {
  return fmaxz(lo, fminz(v, hi));
}
// Float @mod as Zig lowers it: x<0 ? fmod(fmod(x,y)+y, y) : fmod(x,y).
Z2D_PX __forceinline__ float zmodf(float x, float y) {
  float a = fmodf(x, y);
  return x < 0.0f ? fmodf(a + y, y) : a;
}
// fmod(x, y) for y > 0: |x| < y gives x and y <= |x| < 2y gives x - y with
// x's sign, exact by Sterbenz's lemma and equal to fmod's exact remainder;
// other arguments go to fmod.
Z2D_PX __forceinline__ double fmod_pos(double x, double y) {
  const double ax = fabs(x);
  if (ax < y) return x;
  if (ax < 2.0 * y) return x - copysign(y, x);
  return fmod(x, y);
}
Z2D_PX __forceinline__ double zmodd(double x, double y) {
  if (y > 0.0) {
    const double a = fmod_pos(x, y);
    return x < 0.0 ? fmod_pos(a + y, y) : a;
  }
  double a = fmod(x, y);
  return x < 0.0 ? fmod(a + y, y) : a;
}
// round(255*v) to an 8-bit channel; saturates outside the range the
// reference defines (its @intFromFloat is undefined there).
// roundf(v) (half away from zero), then 0 below zero or for NaN and 65535
// above it, without branches: trunc(v + (0.5 - 2^-25)) is roundf for every
// v >= 0 (checked over all 2^32 floats with the sign copied onto the
// addend), and for v < 0 it is at most 0 as roundf(v) is, so the clamp
// gives 0 either way; the clamp takes NaN to 0.
Z2D_PX __forceinline__ uint32_t f2u8(float v) {
  const float r = truncf(v + 0x1.fffffep-2f);
  return (uint32_t)fminf(fmaxf(r, 0.0f), 65535.0f);
}

// (int32_t)round(v) (half away from zero) as one add and the truncating
// conversion: trunc(v + copysign(0.5 - 2^-54, v)) is round(v) for every
// double (v + 0.5 - 2^-54 reaches the next integer exactly when v's
// fraction is at least one half); out-of-range values convert as the
// rounded ones do.
Z2D_PX __forceinline__ int32_t round_i32(double v) {
  return (int32_t)(v + copysign(0x1.fffffffffffffp-2, v));
}

// std.math.pow(f32) as the reference's stdlib computes it (the Go-derived
// algorithm: x^frac via exp(frac*log x), x^int by frexp squaring).
Z2D_PX inline float zpowf(float x, float y) {
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (z_isnan(x) || z_isnan(y)) return __builtin_nanf("");
  if (y == 1.0f) return x;
  if (x == 0.0f) {
    float ip;
    bool odd = modff(y, &ip) == 0.0f && fmodf(fabsf(ip), 2.0f) == 1.0f;
    if (y < 0.0f) return odd ? copysignf(__builtin_inff(), x) : __builtin_inff();
    return odd ? x : 0.0f;
  }
  if (z_isinf(y) || z_isinf(x)) return powf(x, y);
  if (y == 0.5f) return sqrtf(x);
  if (y == -0.5f) return 1.0f / sqrtf(x);
  float yi;
  float yf = modff(fabsf(y), &yi);
  if (yf != 0.0f && x < 0.0f) return __builtin_nanf("");
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

// ---------------------------------------------------------------------------
// pixels
// ---------------------------------------------------------------------------
struct I4 {  // RGBA16 lane values (u16 semantics, held in u32)
  uint32_t r, g, b, a;
};
struct F4 {
  float r, g, b, a;
};

__host__ __device__ __forceinline__ int fmt_bits(uint32_t f) {
  return f == Z2D_FORMAT_ALPHA8 ? 8 : f == Z2D_FORMAT_ALPHA4 ? 4 : f == Z2D_FORMAT_ALPHA2 ? 2
       : f == Z2D_FORMAT_ALPHA1 ? 1 : 32;
}

// Alpha8.scaleIntoIntAlpha from n bits (bit repetition).
Z2D_PX __forceinline__ uint32_t alpha_up8(uint32_t v, int bits) {
  return bits == 1 ? v * 255u : bits == 2 ? v * 85u : bits == 4 ? v * 17u : v;
}
// Alpha{4,2,1}.scaleIntoIntAlpha from u8 (shift right).
Z2D_PX __forceinline__ uint32_t alpha_down(uint32_t a8, int bits) {
  return (a8 & 0xFFu) >> (8 - bits);
}

// Raw pixel word (as stored: 4-byte formats as their u32, alpha8 as its byte,
// packed alpha as its native value) -> RGBA16 lanes.
Z2D_PX __forceinline__ I4 unpack(uint32_t raw, uint32_t fmt) {
  I4 o;
  switch (fmt) {
    case Z2D_FORMAT_RGBA:
      o.r = raw & 0xFF; o.g = (raw >> 8) & 0xFF; o.b = (raw >> 16) & 0xFF; o.a = raw >> 24;
      break;
    case Z2D_FORMAT_ARGB:
      o.b = raw & 0xFF; o.g = (raw >> 8) & 0xFF; o.r = (raw >> 16) & 0xFF; o.a = raw >> 24;
      break;
    case Z2D_FORMAT_RGB:
      o.r = raw & 0xFF; o.g = (raw >> 8) & 0xFF; o.b = (raw >> 16) & 0xFF; o.a = 255;
      break;
    case Z2D_FORMAT_XRGB:
      o.b = raw & 0xFF; o.g = (raw >> 8) & 0xFF; o.r = (raw >> 16) & 0xFF; o.a = 255;
      break;
    case Z2D_FORMAT_ALPHA8:
      o.r = o.g = o.b = 0; o.a = raw & 0xFF;
      break;
    default:
      o.r = o.g = o.b = 0; o.a = alpha_up8(raw, fmt_bits(fmt));
      break;
  }
  return o;
}
// RGBA16 lanes -> raw word for `fmt` (u8 truncation of each lane; rgb/xrgb
// pad byte zero; packed alpha shifted down).
Z2D_PX __forceinline__ uint32_t pack(I4 v, uint32_t fmt) {
  uint32_t r = v.r & 0xFF, g = v.g & 0xFF, b = v.b & 0xFF, a = v.a & 0xFF;
  switch (fmt) {
    case Z2D_FORMAT_RGBA: return r | (g << 8) | (b << 16) | (a << 24);
    case Z2D_FORMAT_ARGB: return b | (g << 8) | (r << 16) | (a << 24);
    case Z2D_FORMAT_RGB: return r | (g << 8) | (b << 16);
    case Z2D_FORMAT_XRGB: return b | (g << 8) | (r << 16);
    case Z2D_FORMAT_ALPHA8: return a;
    default: return alpha_down(a, fmt_bits(fmt));
  }
}

// ---------------------------------------------------------------------------
// integer operators (u16 lanes)
// ---------------------------------------------------------------------------
Z2D_PX __forceinline__ uint32_t u16(uint32_t x) { return x & 0xFFFFu; }
Z2D_PX __forceinline__ uint32_t imul(uint32_t a, uint32_t b) { return u16((a * b) / 255u); }
// imul for channels a, b <= 255: (x * 0x8081) >> 23 equals x / 255 for every
// x < 65536, so two 24-bit multiplies and a shift replace the 32-bit
// multiply-high division.
Z2D_PX __forceinline__ uint32_t imul8(uint32_t a, uint32_t b) {
  return mul24(mul24(a, b), 0x8081u) >> 23;
}
Z2D_PX __forceinline__ uint32_t iinvmul(uint32_t a, uint32_t b) {
  return imul(a, u16(255u - b));
}

Z2D_PX __forceinline__ uint32_t overlay_ch(int32_t s, int32_t d, int32_t sa, int32_t da) {
  if ((int32_t)u16((uint32_t)d * 2u) <= da)
    return u16((uint32_t)((2 * s * d) / 255 + (s * (255 - da)) / 255 + (d * (255 - sa)) / 255));
  return u16((uint32_t)((s * (255 + da)) / 255 + (d * (255 + sa)) / 255 - (2 * d * s) / 255 -
                        (da * sa) / 255));
}
Z2D_PX __forceinline__ uint32_t hardlight_ch(int32_t s, int32_t d, int32_t sa, int32_t da) {
  if ((int32_t)u16((uint32_t)s * 2u) <= sa)
    return u16((uint32_t)((2 * s * d) / 255 + (s * (255 - da)) / 255 + (d * (255 - sa)) / 255));
  return u16((uint32_t)((s * (255 + da)) / 255 + (d * (255 + sa)) / 255 - (sa * da) / 255 -
                        (2 * s * d) / 255));
}
// overlay / hard light on 8-bit lanes (every lane <= 255, the case of
// unpacked pixels and paints).  Both of the reference's branches are sums
// of quotients of the same kind:
//   p:  (2sd)/255 + (s(255-da))/255 + (d(255-sa))/255
//   !p: (s(255+da))/255 + (d(255+sa))/255 - (2sd)/255 - (da sa)/255
// so the branch only selects the multipliers: three divisions per channel
// instead of seven, and (da sa)/255 once per pixel.  With lanes <= 255 every
// dividend is non-negative and below 2^18, where the unsigned quotient is
// the signed one; the sum wraps to 16 bits as the signed form's does.
Z2D_PX __forceinline__ uint32_t blend8_ch(bool p, uint32_t s, uint32_t d, uint32_t sa,
                                              uint32_t da, uint32_t q_alpha) {
  const uint32_t t = mul24(2u * s, d) / 255u;
  const uint32_t q1 = mul24(s, p ? 255u - da : 255u + da) / 255u;
  const uint32_t q2 = mul24(d, p ? 255u - sa : 255u + sa) / 255u;
  return u16(p ? t + q1 + q2 : q1 + q2 - t - q_alpha);
}
Z2D_PX __forceinline__ bool lanes8(I4 d, I4 s) {
  return (d.r | d.g | d.b | d.a | s.r | s.g | s.b | s.a) <= 255u;
}

Z2D_PX inline I4 op_int(uint32_t op, I4 d, I4 s) {
  I4 o = {0, 0, 0, 0};
  const uint32_t sa = s.a, da = d.a;
  const uint32_t ua = u16(sa + da - imul(sa, da));
  switch (op) {
    case Z2D_OP_SRC: return s;
    case Z2D_OP_DST: return d;
    case Z2D_OP_SRC_OVER:
      o.r = u16(s.r + iinvmul(d.r, sa)); o.g = u16(s.g + iinvmul(d.g, sa));
      o.b = u16(s.b + iinvmul(d.b, sa)); o.a = ua;
      return o;
    case Z2D_OP_DST_OVER:
      o.r = u16(d.r + iinvmul(s.r, da)); o.g = u16(d.g + iinvmul(s.g, da));
      o.b = u16(d.b + iinvmul(s.b, da)); o.a = ua;
      return o;
    case Z2D_OP_SRC_IN:
      o.r = imul(s.r, da); o.g = imul(s.g, da); o.b = imul(s.b, da); o.a = imul(sa, da);
      return o;
    case Z2D_OP_DST_IN:
      o.r = imul(d.r, sa); o.g = imul(d.g, sa); o.b = imul(d.b, sa); o.a = imul(da, sa);
      return o;
    case Z2D_OP_SRC_OUT:
      o.r = iinvmul(s.r, da); o.g = iinvmul(s.g, da); o.b = iinvmul(s.b, da);
      o.a = iinvmul(sa, da);
      return o;
    case Z2D_OP_DST_OUT:
      o.r = iinvmul(d.r, sa); o.g = iinvmul(d.g, sa); o.b = iinvmul(d.b, sa);
      o.a = iinvmul(da, sa);
      return o;
    case Z2D_OP_SRC_ATOP:
      o.r = u16(imul(s.r, da) + iinvmul(d.r, sa)); o.g = u16(imul(s.g, da) + iinvmul(d.g, sa));
      o.b = u16(imul(s.b, da) + iinvmul(d.b, sa)); o.a = da;
      return o;
    case Z2D_OP_DST_ATOP:
      o.r = u16(imul(d.r, sa) + iinvmul(s.r, da)); o.g = u16(imul(d.g, sa) + iinvmul(s.g, da));
      o.b = u16(imul(d.b, sa) + iinvmul(s.b, da)); o.a = sa;
      return o;
    case Z2D_OP_XOR:
      o.r = u16(iinvmul(s.r, da) + iinvmul(d.r, sa));
      o.g = u16(iinvmul(s.g, da) + iinvmul(d.g, sa));
      o.b = u16(iinvmul(s.b, da) + iinvmul(d.b, sa));
      o.a = u16(iinvmul(sa, da) + iinvmul(da, sa));
      return o;
    case Z2D_OP_PLUS:
      o.r = min(255u, u16(s.r + d.r)); o.g = min(255u, u16(s.g + d.g));
      o.b = min(255u, u16(s.b + d.b)); o.a = min(255u, u16(sa + da));
      return o;
    case Z2D_OP_MULTIPLY:
      o.r = u16(imul(s.r, d.r) + iinvmul(s.r, da) + iinvmul(d.r, sa));
      o.g = u16(imul(s.g, d.g) + iinvmul(s.g, da) + iinvmul(d.g, sa));
      o.b = u16(imul(s.b, d.b) + iinvmul(s.b, da) + iinvmul(d.b, sa));
      o.a = ua;
      return o;
    case Z2D_OP_SCREEN:
      o.r = u16(s.r + d.r - imul(s.r, d.r)); o.g = u16(s.g + d.g - imul(s.g, d.g));
      o.b = u16(s.b + d.b - imul(s.b, d.b)); o.a = ua;
      return o;
    case Z2D_OP_OVERLAY:
      if (lanes8(d, s)) {
        const uint32_t qa = mul24(da, sa) / 255u;
        o.r = blend8_ch(2u * d.r <= da, s.r, d.r, sa, da, qa);
        o.g = blend8_ch(2u * d.g <= da, s.g, d.g, sa, da, qa);
        o.b = blend8_ch(2u * d.b <= da, s.b, d.b, sa, da, qa);
      } else {
        o.r = overlay_ch(s.r, d.r, sa, da); o.g = overlay_ch(s.g, d.g, sa, da);
        o.b = overlay_ch(s.b, d.b, sa, da);
      }
      o.a = ua;
      return o;
    case Z2D_OP_DARKEN:
      o.r = u16(min(imul(s.r, da), imul(d.r, sa)) + iinvmul(s.r, da) + iinvmul(d.r, sa));
      o.g = u16(min(imul(s.g, da), imul(d.g, sa)) + iinvmul(s.g, da) + iinvmul(d.g, sa));
      o.b = u16(min(imul(s.b, da), imul(d.b, sa)) + iinvmul(s.b, da) + iinvmul(d.b, sa));
      o.a = ua;
      return o;
    case Z2D_OP_LIGHTEN:
      o.r = u16(max(imul(s.r, da), imul(d.r, sa)) + iinvmul(s.r, da) + iinvmul(d.r, sa));
      o.g = u16(max(imul(s.g, da), imul(d.g, sa)) + iinvmul(s.g, da) + iinvmul(d.g, sa));
      o.b = u16(max(imul(s.b, da), imul(d.b, sa)) + iinvmul(s.b, da) + iinvmul(d.b, sa));
      o.a = ua;
      return o;
    case Z2D_OP_HARD_LIGHT:
      if (lanes8(d, s)) {
        const uint32_t qa = mul24(da, sa) / 255u;
        o.r = blend8_ch(2u * s.r <= sa, s.r, d.r, sa, da, qa);
        o.g = blend8_ch(2u * s.g <= sa, s.g, d.g, sa, da, qa);
        o.b = blend8_ch(2u * s.b <= sa, s.b, d.b, sa, da, qa);
      } else {
        o.r = hardlight_ch(s.r, d.r, sa, da); o.g = hardlight_ch(s.g, d.g, sa, da);
        o.b = hardlight_ch(s.b, d.b, sa, da);
      }
      o.a = ua;
      return o;
    case Z2D_OP_DIFFERENCE:
      o.r = u16(s.r + d.r - 2u * min(imul(s.r, da), imul(d.r, sa)));
      o.g = u16(s.g + d.g - 2u * min(imul(s.g, da), imul(d.g, sa)));
      o.b = u16(s.b + d.b - 2u * min(imul(s.b, da), imul(d.b, sa)));
      o.a = ua;
      return o;
    case Z2D_OP_EXCLUSION:
      o.r = u16(u16(imul(s.r, da) + imul(d.r, sa) - 2u * imul(s.r, d.r)) + iinvmul(s.r, da) +
                iinvmul(d.r, sa));
      o.g = u16(u16(imul(s.g, da) + imul(d.g, sa) - 2u * imul(s.g, d.g)) + iinvmul(s.g, da) +
                iinvmul(d.g, sa));
      o.b = u16(u16(imul(s.b, da) + imul(d.b, sa) - 2u * imul(s.b, d.b)) + iinvmul(s.b, da) +
                iinvmul(d.b, sa));
      o.a = ua;
      return o;
    default:  // clear; float-only operators read as clear in integer mode
      return o;
  }
}

// ---------------------------------------------------------------------------
// float operators
// ---------------------------------------------------------------------------
Z2D_PX __forceinline__ float dodge_ch(float s, float d, float sa, float da) {
  if (s == sa && d == 0.0f) return s * (1.0f - da);
  if (s == sa) return sa * da + s * (1.0f - da) + d * (1.0f - sa);
  return sa * da * fminz(1.0f, d / da * sa / (sa - s)) + s * (1.0f - da) + d * (1.0f - sa);
}
// The reference's p0 receives (sca, dca, sa): its first branch tests
// sca == 0 && dca == sa.
Z2D_PX __forceinline__ float burn_ch(float s, float d, float sa, float da) {
  if (s == 0.0f && d == sa) return sa * da + d * (1.0f - sa);
  if (s == 0.0f) return d * (1.0f - sa);
  return sa * da * (1.0f - fminz(1.0f, (1.0f - d / da) * sa / s)) + s * (1.0f - da) +
         d * (1.0f - sa);
}
Z2D_PX __forceinline__ float soft_ch(float s, float d, float sa, float da) {
  if (da == 0.0f) return s;
  float m = d / da;
  if (2.0f * s <= sa)
    return d * (sa + (2.0f * s - sa) * (1.0f - m)) + s * (1.0f - da) + d * (1.0f - sa);
  if (4.0f * d <= da)
    return d * sa + da * (2.0f * s - sa) * (4.0f * m * (4.0f * m + 1.0f) * (m - 1.0f) + 7.0f * m) +
           s * (1.0f - da) + d * (1.0f - sa);
  return da * (2.0f * s - sa) * (sqrtf(m) - m) + s - s * da + d;
}

struct F3 {
  float r, g, b;
};
Z2D_PX __forceinline__ float mx3(F3 c) { return fmaxz(fmaxz(c.r, c.g), c.b); }
Z2D_PX __forceinline__ float mn3(F3 c) { return fminz(fminz(c.r, c.g), c.b); }
Z2D_PX __forceinline__ float lum3(F3 c) { return c.r * 0.3f + c.g * 0.59f + c.b * 0.11f; }
// clipColor with the reference's vector semantics: the x>a branch wins over
// the n<0 branch and both read the unclipped channel.
Z2D_PX __forceinline__ float clip1(float c, float l, float n, float x, float a) {
  float r = c;
  if (n < 0.0f) {
    float t = l - n;
    r = t == 0.0f ? 0.0f : l + ((c - l) * l) / t;
  }
  if (x > a) {
    float t = x - l;
    r = t == 0.0f ? 0.0f : l + ((c - l) * (a - l)) / t;
  }
  return r;
}
Z2D_PX __forceinline__ F3 clip3(F3 c, float a) {
  float l = lum3(c), n = mn3(c), x = mx3(c);
  return F3{clip1(c.r, l, n, x, a), clip1(c.g, l, n, x, a), clip1(c.b, l, n, x, a)};
}
Z2D_PX __forceinline__ F3 setlum3(F3 c, float l) {
  float d = l - lum3(c);
  return F3{c.r + d, c.g + d, c.b + d};
}
Z2D_PX __forceinline__ F3 setsat3(F3 c, float s) {
  float n = mn3(c), x = mx3(c), d = x - n;
  return F3{d == 0.0f ? 0.0f : (c.r - n) * s / d, d == 0.0f ? 0.0f : (c.g - n) * s / d,
            d == 0.0f ? 0.0f : (c.b - n) * s / d};
}
Z2D_PX __forceinline__ F4 ns_out(F3 c, F4 d, F4 s) {
  return F4{s.r * (1.0f - d.a) + d.r * (1.0f - s.a) + c.r,
            s.g * (1.0f - d.a) + d.g * (1.0f - s.a) + c.g,
            s.b * (1.0f - d.a) + d.b * (1.0f - s.a) + c.b, s.a + d.a - s.a * d.a};
}

Z2D_PX inline F4 op_float(uint32_t op, F4 d, F4 s) {
  const float ua = s.a + d.a - s.a * d.a;
  switch (op) {
    case Z2D_OP_CLEAR: return F4{0, 0, 0, 0};
    case Z2D_OP_SRC: return s;
    case Z2D_OP_DST: return d;
    case Z2D_OP_SRC_OVER:
      return F4{s.r + d.r * (1.0f - s.a), s.g + d.g * (1.0f - s.a), s.b + d.b * (1.0f - s.a), ua};
    case Z2D_OP_DST_OVER:
      return F4{d.r + s.r * (1.0f - d.a), d.g + s.g * (1.0f - d.a), d.b + s.b * (1.0f - d.a), ua};
    case Z2D_OP_SRC_IN: return F4{s.r * d.a, s.g * d.a, s.b * d.a, s.a * d.a};
    case Z2D_OP_DST_IN: return F4{d.r * s.a, d.g * s.a, d.b * s.a, s.a * d.a};
    case Z2D_OP_SRC_OUT:
      return F4{s.r * (1.0f - d.a), s.g * (1.0f - d.a), s.b * (1.0f - d.a), s.a * (1.0f - d.a)};
    case Z2D_OP_DST_OUT:
      return F4{d.r * (1.0f - s.a), d.g * (1.0f - s.a), d.b * (1.0f - s.a), d.a * (1.0f - s.a)};
    case Z2D_OP_SRC_ATOP:
      return F4{s.r * d.a + d.r * (1.0f - s.a), s.g * d.a + d.g * (1.0f - s.a),
                s.b * d.a + d.b * (1.0f - s.a), d.a};
    case Z2D_OP_DST_ATOP:
      return F4{d.r * s.a + s.r * (1.0f - d.a), d.g * s.a + s.g * (1.0f - d.a),
                d.b * s.a + s.b * (1.0f - d.a), s.a};
    case Z2D_OP_XOR:
      return F4{s.r * (1.0f - d.a) + d.r * (1.0f - s.a), s.g * (1.0f - d.a) + d.g * (1.0f - s.a),
                s.b * (1.0f - d.a) + d.b * (1.0f - s.a), s.a + d.a - 2.0f * s.a * d.a};
    case Z2D_OP_PLUS:
      return F4{fminz(1.0f, s.r + d.r), fminz(1.0f, s.g + d.g), fminz(1.0f, s.b + d.b),
                fminz(1.0f, s.a + d.a)};
    case Z2D_OP_MULTIPLY:
      return F4{s.r * d.r + s.r * (1.0f - d.a) + d.r * (1.0f - s.a),
                s.g * d.g + s.g * (1.0f - d.a) + d.g * (1.0f - s.a),
                s.b * d.b + s.b * (1.0f - d.a) + d.b * (1.0f - s.a), ua};
    case Z2D_OP_SCREEN:
      return F4{s.r + d.r - s.r * d.r, s.g + d.g - s.g * d.g, s.b + d.b - s.b * d.b, ua};
    case Z2D_OP_OVERLAY: {
#define Z2D_OVL(sc, dc)                                                                   \
  ((2.0f * (dc) <= d.a) ? (2.0f * (sc) * (dc) + (sc) * (1.0f - d.a) + (dc) * (1.0f - s.a)) \
                        : ((sc) * (1.0f + d.a) + (dc) * (1.0f + s.a) - 2.0f * (dc) * (sc) - d.a * s.a))
      return F4{Z2D_OVL(s.r, d.r), Z2D_OVL(s.g, d.g), Z2D_OVL(s.b, d.b), ua};
#undef Z2D_OVL
    }
    case Z2D_OP_DARKEN:
      return F4{fminz(s.r * d.a, d.r * s.a) + s.r * (1.0f - d.a) + d.r * (1.0f - s.a),
                fminz(s.g * d.a, d.g * s.a) + s.g * (1.0f - d.a) + d.g * (1.0f - s.a),
                fminz(s.b * d.a, d.b * s.a) + s.b * (1.0f - d.a) + d.b * (1.0f - s.a), ua};
    case Z2D_OP_LIGHTEN:
      return F4{fmaxz(s.r * d.a, d.r * s.a) + s.r * (1.0f - d.a) + d.r * (1.0f - s.a),
                fmaxz(s.g * d.a, d.g * s.a) + s.g * (1.0f - d.a) + d.g * (1.0f - s.a),
                fmaxz(s.b * d.a, d.b * s.a) + s.b * (1.0f - d.a) + d.b * (1.0f - s.a), ua};
    case Z2D_OP_COLOR_DODGE:
      return F4{dodge_ch(s.r, d.r, s.a, d.a), dodge_ch(s.g, d.g, s.a, d.a),
                dodge_ch(s.b, d.b, s.a, d.a), ua};
    case Z2D_OP_COLOR_BURN:
      return F4{burn_ch(s.r, d.r, s.a, d.a), burn_ch(s.g, d.g, s.a, d.a),
                burn_ch(s.b, d.b, s.a, d.a), ua};
    case Z2D_OP_HARD_LIGHT: {
#define Z2D_HL(sc, dc)                                                                    \
  ((2.0f * (sc) <= s.a) ? (2.0f * (sc) * (dc) + (sc) * (1.0f - d.a) + (dc) * (1.0f - s.a)) \
                        : ((sc) * (1.0f + d.a) + (dc) * (1.0f + s.a) - s.a * d.a - 2.0f * (sc) * (dc)))
      return F4{Z2D_HL(s.r, d.r), Z2D_HL(s.g, d.g), Z2D_HL(s.b, d.b), ua};
#undef Z2D_HL
    }
    case Z2D_OP_SOFT_LIGHT:
      return F4{soft_ch(s.r, d.r, s.a, d.a), soft_ch(s.g, d.g, s.a, d.a),
                soft_ch(s.b, d.b, s.a, d.a), ua};
    case Z2D_OP_DIFFERENCE:
      return F4{s.r + d.r - 2.0f * fminz(s.r * d.a, d.r * s.a),
                s.g + d.g - 2.0f * fminz(s.g * d.a, d.g * s.a),
                s.b + d.b - 2.0f * fminz(s.b * d.a, d.b * s.a), ua};
    case Z2D_OP_EXCLUSION:
      return F4{(s.r * d.a + d.r * s.a - 2.0f * s.r * d.r) + s.r * (1.0f - d.a) + d.r * (1.0f - s.a),
                (s.g * d.a + d.g * s.a - 2.0f * s.g * d.g) + s.g * (1.0f - d.a) + d.g * (1.0f - s.a),
                (s.b * d.a + d.b * s.a - 2.0f * s.b * d.b) + s.b * (1.0f - d.a) + d.b * (1.0f - s.a),
                ua};
    case Z2D_OP_HUE: {
      F3 c{s.r * s.a, s.g * s.a, s.b * s.a}, cd{d.r, d.g, d.b};
      c = setsat3(c, (mx3(cd) - mn3(cd)) * s.a);
      c = setlum3(c, lum3(cd) * s.a);
      return ns_out(clip3(c, s.a * d.a), d, s);
    }
    case Z2D_OP_SATURATION: {
      F3 c{d.r * s.a, d.g * s.a, d.b * s.a}, cs{s.r, s.g, s.b}, cd{d.r, d.g, d.b};
      c = setsat3(c, (mx3(cs) - mn3(cs)) * d.a);
      c = setlum3(c, lum3(cd) * s.a);
      return ns_out(clip3(c, s.a * d.a), d, s);
    }
    case Z2D_OP_COLOR: {
      F3 c{s.r * d.a, s.g * d.a, s.b * d.a}, cd{d.r, d.g, d.b};
      c = setlum3(c, lum3(cd) * s.a);
      return ns_out(clip3(c, s.a * d.a), d, s);
    }
    case Z2D_OP_LUMINOSITY: {
      F3 c{d.r * s.a, d.g * s.a, d.b * s.a}, cs{s.r, s.g, s.b};
      c = setlum3(c, lum3(cs) * d.a);
      return ns_out(clip3(c, s.a * d.a), d, s);
    }
    default: return F4{0, 0, 0, 0};
  }
}

__host__ __device__ __forceinline__ bool op_requires_float(uint32_t op) {
  return op == Z2D_OP_COLOR_DODGE || op == Z2D_OP_COLOR_BURN || op == Z2D_OP_SOFT_LIGHT ||
         op == Z2D_OP_HUE || op == Z2D_OP_SATURATION || op == Z2D_OP_COLOR ||
         op == Z2D_OP_LUMINOSITY;
}
__host__ __device__ __forceinline__ bool op_is_bounded(uint32_t op) {
  return !(op == Z2D_OP_SRC_IN || op == Z2D_OP_DST_IN || op == Z2D_OP_SRC_OUT ||
           op == Z2D_OP_DST_ATOP);
}

Z2D_PX __forceinline__ F4 decode_raw(I4 v) {
  return F4{(float)v.r / 255.0f, (float)v.g / 255.0f, (float)v.b / 255.0f, (float)v.a / 255.0f};
}
Z2D_PX __forceinline__ I4 encode_raw(F4 v) {
  return I4{f2u8(255.0f * v.r), f2u8(255.0f * v.g), f2u8(255.0f * v.b), f2u8(255.0f * v.a)};
}

// ---------------------------------------------------------------------------
// paint sources
// ---------------------------------------------------------------------------

// Device copy of one gradient: geometry, stored transformation, method and
// its stop table (offsets clamped, stable-sorted, colors clamped by
// Color.init on the host side).
struct DevStop {
  float offset;
  uint32_t space;
  float v[4];
};
struct DevGradient {
  uint32_t type, method, polar, n_stops;
  uint32_t identity, pad0;
  double geom[6];
  double m[6];
  // radial pre-calculations (Radial.initBuffer, src/gradient.zig:440-470)
  double r1, r2, cdx, cdy, dr, min_dr, a, inv_a;
  double conic_angle;  // @mod(angle, 2*pi)
  const DevStop *stops;
  const float *offs;   // stop offsets, dense
  const float4 *pre;   // per stop: its color in the interpolation space
  // linear: the gradient vector and 1/|v|^2 (Linear.getOffset's constants)
  double lin_dx, lin_dy, lin_inv;
  uint32_t lin_degenerate, pad1;
};

// Where a draw's stops are read from (LDS staging or the gradient itself).
struct StopView {
  const float *off;
  const float4 *pre;
};

// A resolved paint (Pattern, src/pattern.zig:32-50).
struct DevPaint {
  uint32_t kind;  // Z2D_PATTERN_*
  uint32_t px_format;
  I4 px;          // RGBA.fromPixel(pixel) for opaque patterns
  uint32_t px_opaque;
  // dither
  uint32_t dither_type, dither_source, dither_scale;
  F4 dither_color;  // pixel/color sources: linear straight-alpha f32
  const DevGradient *gradient;  // gradient pattern or gradient dither source
};

Z2D_PX __forceinline__ void user_to_device(const double *m, double &x, double &y) {
  double ox = m[0] * x + m[1] * y;
  double oy = m[2] * x + m[3] * y;
  x = ox + m[4];
  y = oy + m[5];
}
// std.math.atan / std.math.atan2 for f64 as Zig's standard library computes
// them (lib/std/math/atan.zig, atan2.zig, the reference's third-party
// dependency): the published fdlibm algorithm (Sun Microsystems
// s_atan.c / e_atan2.c).  atan reduces |x| to one of five ranges around 0,
// 0.5, 1, 1.5 and infinity, evaluates an odd polynomial as two Horner
// chains and adds atan of the range's centre in two parts; atan2 handles
// signs, zeros and infinities and takes atan(|y/x|).  A few live doubles,
// so a conic gradient's sampler stays small.
Z2D_PX __forceinline__ double zatan(double x) {
  const double atanhi0 = 4.63647609000806093515e-01, atanhi1 = 7.85398163397448278999e-01,
               atanhi2 = 9.82793723247329054082e-01, atanhi3 = 1.57079632679489655800e+00;
  const double atanlo0 = 2.26987774529616870924e-17, atanlo1 = 3.06161699786838301793e-17,
               atanlo2 = 1.39033110312309984516e-17, atanlo3 = 6.12323399573676603587e-17;
  uint32_t ix = dhi(x);
  const uint32_t sign = ix >> 31;
  ix &= 0x7FFFFFFFu;
  if (ix >= 0x44100000u) {  // |x| >= 2^66
    if (z_isnan(x)) return x;
    const double z = atanhi3 + 0x1.0p-120;
    return sign ? -z : z;
  }
  int id;
  if (ix < 0x3FDC0000u) {  // |x| < 0.4375
    if (ix < 0x3E400000u) return x;  // |x| < 2^-27
    id = -1;
  } else {
    x = fabs(x);
    if (ix < 0x3FF30000u) {    // |x| < 1.1875
      if (ix < 0x3FE60000u) {  // 7/16 <= |x| < 11/16
        id = 0;
        x = (2.0 * x - 1.0) / (2.0 + x);
      } else {  // 11/16 <= |x| < 19/16
        id = 1;
        x = (x - 1.0) / (x + 1.0);
      }
    } else if (ix < 0x40038000u) {  // |x| < 2.4375
      id = 2;
      x = (x - 1.5) / (1.0 + 1.5 * x);
    } else {  // 2.4375 <= |x| < 2^66
      id = 3;
      x = -1.0 / x;
    }
  }
  const double z = x * x;
  const double w = z * z;
  const double s1 =
      z * (3.33333333333329318027e-01 +
           w * (1.42857142725034663711e-01 +
                w * (9.09088713343650656196e-02 +
                     w * (6.66107313738753120669e-02 +
                          w * (4.97687799461593236017e-02 + w * 1.62858201153657823623e-02)))));
  const double s2 =
      w * (-1.99999999998764832476e-01 +
           w * (-1.11111104054623557880e-01 +
                w * (-7.69187620504482999495e-02 +
                     w * (-5.83357013379057348645e-02 + w * -3.65315727442169155270e-02))));
  if (id < 0) return x - x * (s1 + s2);
  const double hi = id == 0 ? atanhi0 : id == 1 ? atanhi1 : id == 2 ? atanhi2 : atanhi3;
  const double lo = id == 0 ? atanlo0 : id == 1 ? atanlo1 : id == 2 ? atanlo2 : atanlo3;
  const double zz = hi - ((x * (s1 + s2) - lo) - x);
  return sign ? -zz : zz;
}
Z2D_PX __forceinline__ double zatan2(double y, double x) {
  const double pi = 3.1415926535897931160E+00, pi_lo = 1.2246467991473531772E-16;
  if (z_isnan(x) || z_isnan(y)) return x + y;
  uint32_t ix = dhi(x), iy = dhi(y);
  const uint32_t lx = dlo(x), ly = dlo(y);
  if (((ix - 0x3FF00000u) | lx) == 0) return zatan(y);  // x = 1.0
  const uint32_t m = ((iy >> 31) & 1u) | ((ix >> 30) & 2u);  // 2 sign(x) + sign(y)
  ix &= 0x7FFFFFFFu;
  iy &= 0x7FFFFFFFu;
  if ((iy | ly) == 0) {  // y = 0
    if (m < 2) return y;
    return m == 2 ? pi : -pi;
  }
  if ((ix | lx) == 0) return (m & 1u) ? -pi / 2 : pi / 2;  // x = 0
  if (ix == 0x7FF00000u) {  // x infinite
    if (iy == 0x7FF00000u)
      return m == 0 ? pi / 4 : m == 1 ? -pi / 4 : m == 2 ? 3 * pi / 4 : -3 * pi / 4;
    return m == 0 ? 0.0 : m == 1 ? -0.0 : m == 2 ? pi : -pi;
  }
  if (ix + (64u << 20) < iy || iy == 0x7FF00000u)  // |y/x| > 2^64
    return (m & 1u) ? -pi / 2 : pi / 2;
  double z;
  if ((m & 2u) && iy + (64u << 20) < ix)  // |y/x| < 2^-64, x < 0
    z = 0.0;
  else
    z = zatan(fabs(y / x));
  if (m == 0) return z;
  if (m == 1) return -z;
  if (m == 2) return pi - (z - pi_lo);
  return (z - pi_lo) - pi;
}

Z2D_PX __forceinline__ double dclamp01(double v) { return fmax(0.0, fmin(v, 1.0)); }

// Gradient.getOffset at pixel (x, y), one function per gradient type.
Z2D_PX __forceinline__ float linear_offset(const DevGradient &g, int x, int y) {
  double px = (double)x + 0.5, py = (double)y + 0.5;
  if (!g.identity) user_to_device(g.m, px, py);
  if (g.lin_degenerate) return -1.0f;
  double spx = px - g.geom[0], spy = py - g.geom[1];
  double dt = 0.0;
  dt += g.lin_dx * spx;
  dt += g.lin_dy * spy;
  return (float)dclamp01(dt * g.lin_inv);
}
Z2D_PX __forceinline__ float radial_offset(const DevGradient &g, int x, int y) {
  double px = (double)x + 0.5, py = (double)y + 0.5;
  {
    if (g.r1 == 0.0 && g.r2 == 0.0) return -1.0f;
    if (!g.identity) user_to_device(g.m, px, py);
    double pdx = px - g.geom[0], pdy = py - g.geom[1];
    double b = 0.0;
    b += pdx * g.cdx;
    b += pdy * g.cdy;
    b += g.r1 * g.dr;
    double c = 0.0;
    c += pdx * pdx;
    c += pdy * pdy;
    c += (-g.r1) * g.r1;
    if (g.a == 0.0) {
      if (b == 0.0) return -1.0f;
      double t = 0.5 * c / b;
      if (t * g.dr >= g.min_dr) return (float)dclamp01(t);
      return -1.0f;
    }
    double discr = 0.0;
    discr += b * b;
    discr += g.a * (-c);
    if (discr >= 0.0) {
      double sq = sqrt(discr);
      double t0 = (b + sq) * g.inv_a;
      double t1 = (b - sq) * g.inv_a;
      if (t0 * g.dr >= g.min_dr) return (float)dclamp01(t0);
      if (t1 * g.dr >= g.min_dr) return (float)dclamp01(t1);
    }
    return -1.0f;
  }
}
Z2D_PX __forceinline__ float conic_offset(const DevGradient &g, int x, int y) {
  double px = (double)x + 0.5, py = (double)y + 0.5;
  const double two_pi = 3.14159265358979323846 * 2.0;
  if (!g.identity) user_to_device(g.m, px, py);
  double ang = zmodd(zatan2(py - g.geom[1], px - g.geom[0]) - g.conic_angle, two_pi);
  return (float)(ang / two_pi);
}
Z2D_PX inline float gradient_offset(const DevGradient &g, int x, int y) {
  if (g.type == Z2D_GRADIENT_LINEAR) return linear_offset(g, x, y);
  if (g.type == Z2D_GRADIENT_RADIAL) return radial_offset(g, x, y);
  return conic_offset(g, x, y);
}

// Stop.List.search.
Z2D_PX inline void stops_search(const float *st, uint32_t n_stops, float offset, int &i0,
                                    int &i1, float &t, bool &transparent) {
  const int n = (int)n_stops;
  transparent = false;
  if (offset < 0.0f || n == 0) {
    transparent = true;
    i0 = i1 = 0;
    t = 0.0f;
    return;
  }
  float o = fminz(offset, 1.0f);
  int left = 0, right = n, mid = 0;
  while (left < right) {
    mid = left + (right - left) / 2;
    if (o >= st[mid] && (mid == n - 1 || o <= st[mid + 1])) break;
    if (o < st[mid]) {
      right = mid;
      continue;
    }
    if (o > st[mid]) {
      left = mid + 1;
      continue;
    }
  }
  if (mid == n - 1) {
    i0 = i1 = mid;
    t = o - st[mid];
    return;
  }
  if (mid == 0 && o < st[mid]) {
    i0 = i1 = mid;
    t = o / st[mid];
    return;
  }
  i0 = mid;
  i1 = mid + 1;
  float rl = st[mid + 1] - st[mid];
  t = rl != 0.0f ? (o - st[mid]) / rl : 0.0f;
}

// Stop.List.search for a list of at most 8 stops whose offsets every lane
// reads at the same addresses, padded with +infinity up to 8: when the
// offset equals no stop offset, the reference's binary search can only end
// on the one interval st[m] < o < st[m + 1] that holds it, or on an end stop
// (where both colors are that stop's and t does not matter), so m follows
// from the number of stops below the offset with no data-dependent loop; an
// offset equal to some stop offset (where the search path picks among equal
// stops) goes through stops_search itself.
Z2D_PX __forceinline__ void stops_find8(const float *st, uint32_t n_stops, float offset,
                                            int &i0, int &i1, float &t, bool &transparent) {
  const int n = (int)n_stops;
  transparent = offset < 0.0f || n == 0;
  const float o = fminz(offset, 1.0f);
  const float4 a = *reinterpret_cast<const float4 *>(st);
  const float4 b = *reinterpret_cast<const float4 *>(st + 4);
  // The offset is never NaN (fminf) and the stops are finite or +inf, so
  // each difference st[k] - o is +0 exactly when st[k] == o (x - x), has its
  // sign bit set exactly when st[k] < o (a difference of distinct floats is
  // never +0, and a flushed one keeps its sign), and is +inf for padding:
  // the stops below the offset are the differences' sign bits, and a tie is
  // a zero smallest magnitude.
  const float d[8] = {a.x - o, a.y - o, a.z - o, a.w - o, b.x - o, b.y - o, b.z - o, b.w - o};
  uint32_t nb = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) nb += __builtin_bit_cast(uint32_t, d[k]) >> 31;
  const int below = (int)nb;
  const float dmin = fminf(fminf(fminf(fabsf(d[0]), fabsf(d[1])), fminf(fabsf(d[2]), fabsf(d[3]))),
                           fminf(fminf(fabsf(d[4]), fabsf(d[5])), fminf(fabsf(d[6]), fabsf(d[7]))));
  const bool tie = dmin == 0.0f;
  if (transparent) {
    i0 = i1 = 0;
    t = 0.0f;
    return;
  }
  if (tie) {
    stops_search(st, n_stops, offset, i0, i1, t, transparent);
    return;
  }
  // below == 0: before the first stop; below == n: after the last; either
  // way one stop's color, which t leaves unchanged
  i0 = below > 0 ? below - 1 : 0;
  i1 = below < n ? below : n - 1;
  const float rl = st[i1] - st[i0];
  t = rl != 0.0f ? (o - st[i0]) / rl : 0.0f;
}

// HSL.toRGB (src/color.zig:470-488).
Z2D_PX inline void hsl_to_rgb(const float *h, float *o) {
  float hue = fmodf(h[0], 360.0f);
  if (hue < 0.0f) hue += 360.0f;
  const float ns[3] = {0.0f, 8.0f, 4.0f};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float k = fmodf(ns[i] + hue / 30.0f, 12.0f);
    float a = h[1] * fminz(h[2], 1.0f - h[2]);
    o[i] = h[2] - a * fmaxz(-1.0f, fminz(fminz(k - 3.0f, 9.0f - k), 1.0f));
  }
  o[3] = h[3];
}
// HSL.fromRGB (src/color.zig:435-467).
Z2D_PX inline void hsl_from_rgb(const float *c, float *o) {
  float mx = fmaxz(c[0], fmaxz(c[1], c[2])), mn = fminz(c[0], fminz(c[1], c[2]));
  float range = mx - mn, light = (mn + mx) / 2.0f;
  float sat = (light == 0.0f || light == 1.0f) ? 0.0f : (mx - light) / fminz(light, 1.0f - light);
  float hue = 0.0f;
  if (range != 0.0f) {
    if (mx == c[0])
      hue = 60.0f * zmodf((c[1] - c[2]) / range, 6.0f);
    else if (mx == c[1])
      hue = 60.0f * ((c[2] - c[0]) / range + 2.0f);
    else if (mx == c[2])
      hue = 60.0f * ((c[0] - c[1]) / range + 4.0f);
  }
  if (sat < 0.0f) {
    hue += 180.0f;
    if (hue >= 360.0f) hue -= 360.0f;
    sat = fabsf(sat);
  }
  o[0] = hue; o[1] = sat; o[2] = light; o[3] = c[3];
}
// RGB(profile).fromColor: srgb=false -> linear, true -> gamma-encoded sRGB.
Z2D_PX inline void color_to_rgb(const DevStop &s, bool srgb, float *o) {
  if (s.space == Z2D_COLOR_HSL) {
    hsl_to_rgb(s.v, o);
    if (srgb) {
      const float ig = 1.0f / 2.2f;
      for (int i = 0; i < 3; i++) o[i] = zpowf(o[i], ig);
    }
    return;
  }
  bool is_srgb = s.space == Z2D_COLOR_SRGB;
  if (is_srgb == srgb) {
    o[0] = s.v[0]; o[1] = s.v[1]; o[2] = s.v[2]; o[3] = s.v[3];
    return;
  }
  const float gf = is_srgb ? 2.2f : 1.0f, gt = srgb ? 2.2f : 1.0f;
  const float igt = 1.0f / gt;
  for (int i = 0; i < 3; i++) o[i] = zpowf(zpowf(s.v[i], gf), igt);
  o[3] = s.v[3];
}
Z2D_PX inline void color_to_hsl(const DevStop &s, float *o) {
  if (s.space == Z2D_COLOR_HSL) {
    o[0] = s.v[0]; o[1] = s.v[1]; o[2] = s.v[2]; o[3] = s.v[3];
    return;
  }
  float lin[4];
  color_to_rgb(s, false, lin);
  hsl_from_rgb(lin, o);
}

Z2D_PX __forceinline__ float lerpf_(float a, float b, float t) { return a + (b - a) * t; }

Z2D_PX __forceinline__ I4 encode_premul(const float *v) {
  I4 o{f2u8(255.0f * v[0]), f2u8(255.0f * v[1]), f2u8(255.0f * v[2]), f2u8(255.0f * v[3])};
  o.r = u16(o.r * o.a / 255u);
  o.g = u16(o.g * o.a / 255u);
  o.b = u16(o.b * o.a / 255u);
  return o;
}

// Gradient color at (x, y): encode=true -> interpolateEncodeVec (premultiplied
// RGBA16); encode=false -> interpolateVec (linear, straight alpha f32).
// A stop's color in the gradient's interpolation space, the per-stop part of
// interpolateVec / interpolateEncodeVec: premultiplied linear or sRGB RGB, or
// HSL with premultiplied saturation and lightness.  Evaluated once per stop
// (k_grad_setup); a transparent (offset < 0) sample uses all zeros, which is
// what this gives for the transparent linear stop in every space.
Z2D_PX inline float4 stop_interp_color(const DevGradient &g, const DevStop &s) {
  float a[4];
  if (g.method == Z2D_INTERP_LINEAR_RGB || g.method == Z2D_INTERP_SRGB) {
    color_to_rgb(s, g.method == Z2D_INTERP_SRGB, a);
    return make_float4(a[0] * a[3], a[1] * a[3], a[2] * a[3], a[3]);
  }
  color_to_hsl(s, a);
  return make_float4(a[0], a[1] * a[3], a[2] * a[3], a[3]);
}

Z2D_PX inline void gradient_sample(const DevGradient &g, const StopView &sv, int x, int y,
                                       bool encode, I4 &oi, F4 &of) {
  int i0, i1;
  float t;
  bool transparent;
  stops_search(sv.off, g.n_stops, gradient_offset(g, x, y), i0, i1, t, transparent);
  const float4 A = transparent ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : sv.pre[i0];
  const float4 B = transparent ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : sv.pre[i1];
  const float am[4] = {A.x, A.y, A.z, A.w};
  const float bm[4] = {B.x, B.y, B.z, B.w};
  float v[4];
  if (g.method == Z2D_INTERP_LINEAR_RGB || g.method == Z2D_INTERP_SRGB) {
    bool srgb = g.method == Z2D_INTERP_SRGB;
    for (int i = 0; i < 4; i++) v[i] = lerpf_(am[i], bm[i], t);
    if (!srgb && encode) {
      oi = I4{f2u8(255.0f * v[0]), f2u8(255.0f * v[1]), f2u8(255.0f * v[2]), f2u8(255.0f * v[3])};
      return;
    }
    if (v[3] == 0.0f) {
      v[0] = v[1] = v[2] = 0.0f;
    } else {
      v[0] = v[0] / v[3]; v[1] = v[1] / v[3]; v[2] = v[2] / v[3];
    }
    if (srgb)
      for (int i = 0; i < 3; i++) v[i] = zpowf(v[i], 2.2f);
  } else {
    float ah = am[0], bh = bm[0];
    switch (g.polar) {
      case Z2D_POLAR_SHORTER: {
        bool gt = (bh - ah) > 180.0f, lt = (bh - ah) < -180.0f;
        if (gt) ah += 360.0f;
        if (lt) bh += 360.0f;
      } break;
      case Z2D_POLAR_LONGER: {
        float dl = bh - ah;
        bool c1b = (0.0f < dl) && (dl < 180.0f), c2b = (-180.0f < dl) && (dl <= 0.0f);
        if (c1b) ah += 360.0f;
        if (c2b) bh += 360.0f;
      } break;
      case Z2D_POLAR_INCREASING:
        if (bh < ah) bh += 360.0f;
        break;
      default:
        if (ah < bh) ah += 360.0f;
        break;
    }
    float h[4];
    h[0] = zmodf(am[0] + (bh - ah) * t, 360.0f);
    h[1] = lerpf_(am[1], bm[1], t);
    h[2] = lerpf_(am[2], bm[2], t);
    h[3] = lerpf_(am[3], bm[3], t);
    if (h[3] == 0.0f) {
      h[1] = 0.0f; h[2] = 0.0f;
    } else {
      h[1] = h[1] / h[3]; h[2] = h[2] / h[3];
    }
    hsl_to_rgb(h, v);
  }
  if (encode)
    oi = encode_premul(v);
  else
    of = F4{v[0], v[1], v[2], v[3]};
}

// Dither threshold matrices (Dither.zig:120-158, color_vector.zig:482-522).
Z2D_PX __forceinline__ float bayer8(int x, int y) {
  int yy = y ^ x;
  uint32_t m = (uint32_t)(((yy & 1) << 5) | ((x & 1) << 4) | ((yy & 2) << 2) | ((x & 2) << 1) |
                          ((yy & 4) >> 1) | ((x & 4) >> 2));
  return (float)m * (2.0f / 128.0f) - (63.0f / 128.0f);
}

// fromDitherVec for one lane: linear straight-alpha f32.
// g: the paint's gradient (p.gradient or a copy of it); read only for a
// gradient source.
Z2D_PX inline F4 dither_sample(const DevPaint &p, const DevGradient *g, const StopView &sv,
                                   const uint16_t *blue_noise, int x, int y) {
  F4 c;
  if (p.dither_source == Z2D_DITHER_SRC_GRADIENT) {
    I4 unused;
    gradient_sample(*g, sv, x, y, false, unused, c);
  } else {
    c = p.dither_color;
  }
  float m;
  if (p.dither_type == Z2D_DITHER_BAYER) {
    m = bayer8(x, y);
  } else if (p.dither_type == Z2D_DITHER_BLUE_NOISE) {
    int mx = ((x % 64) + 64) % 64, my = ((y % 64) + 64) % 64;
    m = (float)blue_noise[(mx << 6) | my] * (2.0f / 8192.0f) - (4095.0f / 8192.0f);
  } else {
    return c;
  }
  float sc = 1.0f / (float)((1 << p.dither_scale) - 1);
  return F4{clampz(c.r + m * sc, 0.0f, 1.0f), clampz(c.g + m * sc, 0.0f, 1.0f),
            clampz(c.b + m * sc, 0.0f, 1.0f), clampz(c.a + m * sc, 0.0f, 1.0f)};
}

// Integer paint sample (premultiplied RGBA, 0-255 per channel); g is the
// paint's gradient (p.gradient or a copy of it) when it has one.
Z2D_PX __forceinline__ I4 paint_int_g(const DevPaint &p, const DevGradient *g,
                                          const StopView &sv, const uint16_t *bn, int x, int y) {
  if (p.kind == Z2D_PATTERN_OPAQUE) return p.px;
  if (p.kind == Z2D_PATTERN_GRADIENT) {
    I4 o;
    F4 unused;
    gradient_sample(*g, sv, x, y, true, o, unused);
    return o;
  }
  F4 f = dither_sample(p, g, sv, bn, x, y);
  float v[4] = {f.r, f.g, f.b, f.a};
  return encode_premul(v);
}

Z2D_PX __noinline__ I4 paint_int(const DevPaint &p, const StopView &sv, const uint16_t *bn, int x,
                               int y) {
  if (p.kind == Z2D_PATTERN_OPAQUE) return p.px;
  if (p.kind == Z2D_PATTERN_GRADIENT) {
    I4 o;
    F4 unused;
    gradient_sample(*p.gradient, sv, x, y, true, o, unused);
    return o;
  }
  F4 f = dither_sample(p, p.gradient, sv, bn, x, y);
  float v[4] = {f.r, f.g, f.b, f.a};
  return encode_premul(v);
}
Z2D_PX inline F4 paint_float(const DevPaint &p, const StopView &sv, const uint16_t *bn,
                                 int x, int y) {
  if (p.kind == Z2D_PATTERN_OPAQUE) return decode_raw(p.px);
  if (p.kind == Z2D_PATTERN_GRADIENT) {
    I4 unused;
    F4 o;
    gradient_sample(*p.gradient, sv, x, y, false, unused, o);
    return o;
  }
  return dither_sample(p, p.gradient, sv, bn, x, y);
}

// The stops of a paint as stored with its gradient.
Z2D_PX __forceinline__ StopView paint_stops(const DevPaint &p) {
  if (!p.gradient) return StopView{nullptr, nullptr};
  return StopView{p.gradient->offs, p.gradient->pre};
}

}  // namespace z2d
