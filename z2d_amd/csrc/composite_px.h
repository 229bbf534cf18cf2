// composite_px.h - one pixel through one draw: the coverage mask, then the
// draw's operator (src/internal/raster/shared.zig:9-117, compositeOpaque and
// compositeOpacity, for the pixel the raster loops hand it).  Compiled for
// the device by render.hip and, with Z2D_HOST_PIXEL, for the host by the CPU
// harness that runs it under the sanitizers against the oracle.
#pragma once

#include "render.h"

namespace z2d {

Z2D_PX __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// The supersample mask's alpha for k covered samples of 16 at `bits` bits
// (supersample.zig's packed alpha mask, scaled back to 8 bits).
Z2D_PX __forceinline__ uint32_t ssaa_mask(int k, int bits) {
  switch (bits) {
    case 4: return ((15u * (uint32_t)k) / 16u) * 17u;
    case 2: return ((3u * (uint32_t)k) / 16u) * 85u;
    case 1: return ((1u * (uint32_t)k) / 16u) * 255u;
    default: return (255u * (uint32_t)k) / 16u;
  }
}

// One pixel through one draw, given its sample coverage k (of kfull): the
// paint masked by the coverage (DST_IN), then the draw's operator.  The
// supersample mode masks every pixel of its box, even at zero coverage; the
// solid fast path of painter.fill stays in integer precision.
Z2D_PX __forceinline__ uint32_t composite_cov(uint32_t raw, uint32_t fmt, int k, int kfull,
                                              const DevDraw &D, const DevPaint &P,
                                              const StopView &sv, const uint16_t *bn, int x,
                                              int y) {
  uint32_t mask_a = 255;
  bool masked = true;
  if (D.mode == MODE_SSAA) {
    mask_a = ssaa_mask(k, D.mask_bits);
  } else {
    if (k == 0) return raw;
    if (D.op == Z2D_OP_CLEAR) return pack(I4{0, 0, 0, 0}, fmt);
    if (k >= kfull) {
      if (D.flags & DRAW_FAST_SRC) return pack(P.px, fmt);
      masked = false;
    } else {
      mask_a = (uint32_t)clampi(k * 16 - 1, 0, 255);
    }
  }
  const bool int_path = D.precision == Z2D_PRECISION_INTEGER ||
                        (D.mode != MODE_SSAA && (D.flags & DRAW_FAST_SRC));
  if (int_path) {
    I4 src = P.kind == Z2D_PATTERN_OPAQUE ? P.px : paint_int(P, sv, bn, x, y);
    if (masked) src = op_int(Z2D_OP_DST_IN, src, I4{0, 0, 0, mask_a});
    return pack(op_int(D.op, unpack(raw, fmt), src), fmt);
  }
  F4 src = P.kind == Z2D_PATTERN_OPAQUE ? decode_raw(P.px) : paint_float(P, sv, bn, x, y);
  if (masked) src = op_float(Z2D_OP_DST_IN, src, decode_raw(I4{0, 0, 0, mask_a}));
  return pack(encode_raw(op_float(D.op, decode_raw(unpack(raw, fmt)), src)), fmt);
}

// Channels of s (each <= 255) masked by coverage count k.
Z2D_PX __forceinline__ I4 mask_coverage(I4 s, int k) {
  const uint32_t a = (uint32_t)clampi(k * 16 - 1, 0, 255);
  return I4{imul8(s.r, a), imul8(s.g, a), imul8(s.b, a), imul8(s.a, a)};
}
// s: the paint already masked by the pixel's coverage (channels <= 255).
Z2D_PX __forceinline__ uint32_t composite_src_over(uint32_t raw, uint32_t fmt, I4 s) {
  const I4 d = unpack(raw, fmt);
  const uint32_t sa = s.a, na = 255u - sa;
  const I4 o{u16(s.r + imul8(d.r, na)), u16(s.g + imul8(d.g, na)), u16(s.b + imul8(d.b, na)),
             u16(sa + d.a - imul8(sa, d.a))};
  return pack(o, fmt);
}

// The same two steps on RGBA8 words (red in the low byte), two channels at
// a time in the 16-bit halves of a word: each product of a channel and an
// 8-bit factor is below 2^16 (a packed 16-bit multiply), floor(x / 255) is
// (x + (x >> 8) + 1) >> 8 for every x <= 65025, and no sum carries into the
// other half, so every channel is the value the per-channel form computes.
typedef unsigned short z_u16x2 __attribute__((ext_vector_type(2)));
Z2D_PX __forceinline__ uint32_t mul_x2(uint32_t x, uint32_t m) {
  const z_u16x2 a = __builtin_bit_cast(z_u16x2, x);
  const z_u16x2 b = {(unsigned short)m, (unsigned short)m};
  return __builtin_bit_cast(uint32_t, a * b);
}
Z2D_PX __forceinline__ uint32_t div255_x2(uint32_t x) {
  return ((x + ((x >> 8) & 0x00FF00FFu) + 0x00010001u) >> 8) & 0x00FF00FFu;
}
// mask_coverage on an RGBA8 word
Z2D_PX __forceinline__ uint32_t mask_coverage_rgba8(uint32_t s, int k) {
  const uint32_t a = (uint32_t)clampi(k * 16 - 1, 0, 255);
  return div255_x2(mul_x2(s & 0x00FF00FFu, a)) | (div255_x2(mul_x2((s >> 8) & 0x00FF00FFu, a)) << 8);
}
// composite_src_over on an RGBA8 surface word d and RGBA8 source s
Z2D_PX __forceinline__ uint32_t src_over_rgba8(uint32_t d, uint32_t s) {
  const uint32_t sa = s >> 24, na = 255u - sa;
  const uint32_t q_rb = div255_x2(mul_x2(d & 0x00FF00FFu, na));
  const uint32_t q_ga = div255_x2(mul_x2((d >> 8) & 0x00FF00FFu, na));
  const uint32_t o_rb = ((s & 0x00FF00FFu) + q_rb) & 0x00FF00FFu;
  const uint32_t o_g = (((s >> 8) & 0xFFu) + (q_ga & 0xFFu)) & 0xFFu;
  const uint32_t da = d >> 24;
  const uint32_t o_a = (sa + da - imul8(sa, da)) & 0xFFu;
  return o_rb | (o_g << 8) | (o_a << 24);
}

}  // namespace z2d
