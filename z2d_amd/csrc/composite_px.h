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

}  // namespace z2d
