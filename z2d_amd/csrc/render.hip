// render.hip - gfx950 kernels of the fill/stroke/composite hot path.
//
// Pipeline for one job (any number of surfaces and draws, one stream):
//   k_edge_build  flattened nodes -> Polygon edges + extents + draw setup
//                 (fill_plotter.plot, Polygon.addEdge/inBox and the raster
//                 entry logic of src/internal/raster/{direct,multisample,
//                 supersample}.zig)
//   k_rowprog     serial row emulation, only for draws whose coverage depends
//                 on the reference's sort order (unbalanced contours) or whose
//                 direct-mode unbounded clears depend on whole rows
//   k_edge_bins / scan
//                 per (draw, band) entry, the draw's edges that meet the band
//   k_band_pass / scan
//                 ordered per-band lists of entries (band_h-row bins, draw
//                 order kept)
//   k_strip_pass / scan
//                 per strip (TILE_W x ROWS_PER_THREAD pixels), the band
//                 entries whose composite region meets it, in draw order
//   k_render      persistent strip workers, one wave per strip at a time: the
//                 strip stays in registers while each draw of its list is
//                 rasterized (exact f64 intercepts accumulated as winding
//                 deltas in LDS, wave prefix scans for per-sample winding) and
//                 composited in draw order; HBM sees one read and one write
//                 per strip pixel
//   k_composite   SurfaceCompositor.run batches (src/compositor.zig:302-440)
#include <hip/hip_runtime.h>

#include "composite_px.h"
#include "render.h"

namespace z2d {

// ---------------------------------------------------------------------------
// surface memory
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t load_raw(const DevSurface &s, int64_t idx) {
  switch (s.fmt) {
    case Z2D_FORMAT_ALPHA8: return s.data[idx];
    case Z2D_FORMAT_ALPHA4:
    case Z2D_FORMAT_ALPHA2:
    case Z2D_FORMAT_ALPHA1: {
      int bits = fmt_bits(s.fmt);
      uint64_t bit = (uint64_t)idx * bits;
      return (s.data[bit >> 3] >> (bit & 7)) & ((1u << bits) - 1u);
    }
    default: return reinterpret_cast<const uint32_t *>(s.data)[idx];
  }
}

// Packed alpha rows are not byte aligned, so neighbouring tiles can share a
// byte: packed stores go through 32-bit atomics on the containing word.
__device__ __forceinline__ void store_raw(const DevSurface &s, int64_t idx, uint32_t v) {
  switch (s.fmt) {
    case Z2D_FORMAT_ALPHA8: s.data[idx] = (uint8_t)v; break;
    case Z2D_FORMAT_ALPHA4:
    case Z2D_FORMAT_ALPHA2:
    case Z2D_FORMAT_ALPHA1: {
      int bits = fmt_bits(s.fmt);
      uint64_t bit = (uint64_t)idx * bits;
      uint32_t *word = reinterpret_cast<uint32_t *>(s.data) + (bit >> 5);
      uint32_t sh = (uint32_t)(bit & 31);
      uint32_t mask = ((1u << bits) - 1u) << sh;
      atomicAnd(word, ~mask);
      atomicOr(word, (v << sh) & mask);
    } break;
    default: reinterpret_cast<uint32_t *>(s.data)[idx] = v; break;
  }
}


// Inclusive prefix sum over the 64 lanes of a wave (DPP row shifts within
// 16-lane rows, then row broadcasts across rows), one DPP add per step: a
// shift reads 0 past the row's start (bound_ctrl), a broadcast leaves the
// rows outside its row mask as they were.  Written out because the
// compiler keeps some of the builtin form's steps as a DPP move and an add.
// Each step reads the previous one's result through DPP, which needs two
// wait states after the VALU write (s_nop 1); the first covers the
// caller's write of x.  Every lane of the wave must be active.
__device__ __forceinline__ int32_t wave_incl_scan(int32_t x) {
  asm("s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_add_u32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf"
      : "+v"(x));
  return x;
}

// Minimum / maximum over the 64 lanes of a wave as a wave-uniform value.
// The DPP steps have no bound_ctrl, so a lane whose source lies outside its
// row (or outside the broadcast's row mask) is not written and keeps its
// value; lane 63 ends with the extreme of the whole wave.  Every lane of
// the wave must be active.
__device__ __forceinline__ int32_t wave_min_i32(int32_t x) {
  asm("s_nop 1\n\t"
      "v_min_i32_dpp %0, %0, %0 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_min_i32_dpp %0, %0, %0 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_min_i32_dpp %0, %0, %0 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_min_i32_dpp %0, %0, %0 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_min_i32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_min_i32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf"
      : "+v"(x));
  return __builtin_amdgcn_readlane(x, 63);
}
__device__ __forceinline__ int32_t wave_max_i32(int32_t x) {
  asm("s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf"
      : "+v"(x));
  return __builtin_amdgcn_readlane(x, 63);
}

// ---------------------------------------------------------------------------
// k_edge_build: one wave per draw.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pt_eq(double ax, double ay, double bx, double by) {
  return ax == bx && ay == by;
}

struct ExtAcc {
  uint32_t n;
  double top, bottom, left, right;
};

// Polygon.addEdge (src/internal/tess/Polygon.zig:61-109).
__device__ __forceinline__ void add_edge(z2d_edge *out, ExtAcc &acc, double scale, double x0,
                                         double y0, double x1, double y1) {
  double ax = x0 * scale, ay = y0 * scale, bx = x1 * scale, by = y1 * scale;
  z2d_edge e;
  if (ay < by) {
    e.y0 = ay; e.y1 = by; e.x_start = ax; e.x_inc = (bx - ax) / (by - ay);
  } else if (ay > by) {
    e.y0 = ay; e.y1 = by; e.x_start = bx; e.x_inc = (ax - bx) / (ay - by);
  } else {
    return;
  }
  double top = e.y0 < e.y1 ? e.y0 : e.y1, bottom = e.y0 < e.y1 ? e.y1 : e.y0;
  double left = ax < bx ? ax : bx, right = ax < bx ? bx : ax;
  if (acc.n == 0) {
    acc.top = top; acc.bottom = bottom; acc.left = left; acc.right = right;
  } else {
    if (top < acc.top) acc.top = top;
    if (bottom > acc.bottom) acc.bottom = bottom;
    if (left < acc.left) acc.left = left;
    if (right > acc.right) acc.right = right;
  }
  out[acc.n++] = e;
}

// Polygon.inBox (src/internal/tess/Polygon.zig:142-201).
__device__ bool in_box(const double *ext, double scale, int32_t w, int32_t h) {
  double top = ext[0], bottom = ext[1], left = ext[2], right = ext[3];
  if (!isfinite(left) || !isfinite(top) || !isfinite(right) || !isfinite(bottom)) return false;
  if (w < 1 || h < 1) return false;
  if (right < 0.0 || bottom < 0.0) return false;
  int32_t sx = (int32_t)floor(left / scale), sy = (int32_t)floor(top / scale);
  int32_t ex = (int32_t)ceil(right / scale), ey = (int32_t)ceil(bottom / scale);
  int32_t pw = ex - sx, ph = ey - sy;
  if (pw <= 0 || ph <= 0) return false;
  if (sx + pw < 0 || sy + ph < 0) return false;
  if (sx >= w || sy >= h) return false;
  return true;
}

// Inclusive max over the lanes of a wave up to this one, by the DPP steps
// of wave_incl_scan with a max (no bound_ctrl: a lane without a source keeps
// its value).  Every lane of the wave must be active.
__device__ __forceinline__ int32_t wave_max_scan(int32_t v) {
  asm("s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf"
      : "+v"(v));
  return v;
}
// A lane's value read by every lane (src wave-uniform).
__device__ __forceinline__ uint32_t lane_u32(uint32_t v, uint32_t src) {
  return __builtin_amdgcn_readlane(v, src);
}
__device__ __forceinline__ double lane_f64(double v, uint32_t src) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// The value of the lane below (DPP wave_shr:1); lane 0, which has none,
// gets `first`.
__device__ __forceinline__ uint32_t lane_prev_u32(uint32_t v, uint32_t first) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)first, (int)v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ double lane_prev_f64(double v, double first) {
  const uint64_t u = (uint64_t)__double_as_longlong(v), f = (uint64_t)__double_as_longlong(first);
  const uint32_t lo = lane_prev_u32((uint32_t)u, (uint32_t)f);
  const uint32_t hi = lane_prev_u32((uint32_t)(u >> 32), (uint32_t)(f >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// v from the lane `k` below within the lane's row of 16 (DPP row_shr:k);
// the first k lanes of each row get `fill`.
template <int K>
__device__ __forceinline__ double row_shr_f64(double v, double fill) {
  const uint64_t u = (uint64_t)__double_as_longlong(v), f = (uint64_t)__double_as_longlong(fill);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)f, (int)(uint32_t)u,
                                                            0x110 + K, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(
      (int)(uint32_t)(f >> 32), (int)(uint32_t)(u >> 32), 0x110 + K, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// fmin / fmax of a double over the 64 lanes as a wave-uniform value: a
// prefix within each row of 16, then the four rows' last lanes.  (fmin and
// fmax drop NaN operands in any order, so the order of the steps does not
// change the result.)  Every lane of the wave must be active.
__device__ __forceinline__ double wave_fmin_f64(double v) {
  const double inf = __builtin_huge_val();
  v = fmin(v, row_shr_f64<1>(v, inf));
  v = fmin(v, row_shr_f64<2>(v, inf));
  v = fmin(v, row_shr_f64<4>(v, inf));
  v = fmin(v, row_shr_f64<8>(v, inf));
  return fmin(fmin(lane_f64(v, 15), lane_f64(v, 31)), fmin(lane_f64(v, 47), lane_f64(v, 63)));
}
__device__ __forceinline__ double wave_fmax_f64(double v) {
  const double ninf = -__builtin_huge_val();
  v = fmax(v, row_shr_f64<1>(v, ninf));
  v = fmax(v, row_shr_f64<2>(v, ninf));
  v = fmax(v, row_shr_f64<4>(v, ninf));
  v = fmax(v, row_shr_f64<8>(v, ninf));
  return fmax(fmax(lane_f64(v, 15), lane_f64(v, 31)), fmax(lane_f64(v, 47), lane_f64(v, 63)));
}

// fill_plotter.plot (src/internal/tess/fill_plotter.zig:21-97) with the
// PointBuffer(1, 3) of src/internal/tess/point_buffer.zig, one lane per
// node.  Only the buffer's first point, last point and fill count (capped
// at 3) are ever read.  The last point before a node is the previous node's
// point when that node is a move or a line (a line to an equal point leaves
// it equal); the host sends draws with a line right after a close to the
// serial form below.  A line emits when it moves; a close emits when it
// starts a run of closes, at least two lines moved since the subpath's
// move, and the last point is not the first.  Edge slots are a wave scan of
// the emit flags; the extents are the running min/max, with the first edge
// seeding them as in Polygon.addEdge (a NaN there sticks).
__device__ void plot_parallel(const FNode *nodes, uint32_t n, double scale, z2d_edge *out,
                              ExtAcc &acc) {
  const uint32_t lane = threadIdx.x & 63;
  const double inf = __builtin_huge_val();
  double c_fx = 0.0, c_fy = 0.0, c_px = 0.0, c_py = 0.0;
  uint32_t c_ptag = 0xFFFFFFFFu;
  bool c_have = false;
  int32_t c_cnt = 0;
  uint32_t ebase = 0;
  double m_top = inf, m_bot = -inf, m_left = inf, m_right = -inf;
  double f_top = 0.0, f_bot = 0.0, f_left = 0.0, f_right = 0.0;
  for (uint32_t base = 0; base < n; base += 64) {
    const uint32_t k = base + lane;
    FNode nd{0.0, 0.0, 0xFFFFFFFFu, 0u};
    if (k < n) nd = nodes[k];
    uint32_t tag = nd.tag;
    if (tag == Z2D_NODE_MOVE_TO && k == n - 1) tag = 0xFFFFFFFFu;  // a final move is ignored
    const double px = lane_prev_f64(nd.x, c_px), py = lane_prev_f64(nd.y, c_py);
    const uint32_t ptag = lane_prev_u32(tag, c_ptag);
    const int32_t head = wave_max_scan(tag == Z2D_NODE_MOVE_TO ? (int32_t)lane : -1);
    const int32_t hsrc = head >= 0 ? head : 0;
    const double hx = __shfl(nd.x, hsrc, 64), hy = __shfl(nd.y, hsrc, 64);
    const double fx = head >= 0 ? hx : c_fx, fy = head >= 0 ? hy : c_fy;
    const bool have = head >= 0 || c_have;
    const bool line_moves = tag == Z2D_NODE_LINE_TO && have && !pt_eq(px, py, nd.x, nd.y);
    const int32_t a = line_moves ? 1 : 0;
    const int32_t ps = wave_incl_scan(a);
    const int32_t ps_head = __shfl(ps, hsrc, 64);
    const int32_t cnt_before = head >= 0 ? ps - a - ps_head : c_cnt + ps - a;
    const bool close_emits = tag == Z2D_NODE_CLOSE_PATH && have &&
                             ptag != Z2D_NODE_CLOSE_PATH && cnt_before >= 2 &&
                             !pt_eq(px, py, fx, fy);
    const double tx = line_moves ? nd.x : fx, ty = line_moves ? nd.y : fy;
    // Polygon.addEdge (src/internal/tess/Polygon.zig:61-109)
    const double ax = px * scale, ay = py * scale, bx = tx * scale, by = ty * scale;
    const bool emit = (line_moves || close_emits) && (ay < by || ay > by);
    const uint64_t m = __ballot(emit);
    double top = 0.0, bot = 0.0, left = 0.0, right = 0.0;
    if (emit) {
      z2d_edge e;
      e.y0 = ay;
      e.y1 = by;
      if (ay < by) {
        e.x_start = ax;
        e.x_inc = (bx - ax) / (by - ay);
      } else {
        e.x_start = bx;
        e.x_inc = (ax - bx) / (ay - by);
      }
      top = e.y0 < e.y1 ? e.y0 : e.y1;
      bot = e.y0 < e.y1 ? e.y1 : e.y0;
      left = ax < bx ? ax : bx;
      right = ax < bx ? bx : ax;
      out[ebase + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = e;
      m_top = fmin(m_top, top);
      m_bot = fmax(m_bot, bot);
      m_left = fmin(m_left, left);
      m_right = fmax(m_right, right);
    }
    if (ebase == 0 && m != 0) {
      const int fl = __ffsll((unsigned long long)m) - 1;
      f_top = lane_f64(top, (uint32_t)fl);
      f_bot = lane_f64(bot, (uint32_t)fl);
      f_left = lane_f64(left, (uint32_t)fl);
      f_right = lane_f64(right, (uint32_t)fl);
    }
    ebase += (uint32_t)__popcll(m);
    c_px = lane_f64(nd.x, 63);
    c_py = lane_f64(nd.y, 63);
    c_ptag = lane_u32(tag, 63);
    c_fx = lane_f64(fx, 63);
    c_fy = lane_f64(fy, 63);
    c_have = lane_u32(have ? 1u : 0u, 63) != 0;
    c_cnt = (int32_t)lane_u32((uint32_t)(cnt_before + a), 63);
  }
  m_top = wave_fmin_f64(m_top);
  m_bot = wave_fmax_f64(m_bot);
  m_left = wave_fmin_f64(m_left);
  m_right = wave_fmax_f64(m_right);
  acc.n = ebase;
  if (ebase) {
    acc.top = f_top != f_top ? f_top : m_top;
    acc.bottom = f_bot != f_bot ? f_bot : m_bot;
    acc.left = f_left != f_left ? f_left : m_left;
    acc.right = f_right != f_right ? f_right : m_right;
  } else {
    acc.top = acc.bottom = acc.left = acc.right = 0.0;
  }
}

// The fast renderer's paint class of a paint (FAST_PAINT_NONE: the general
// renderer's).
__device__ __forceinline__ uint32_t fast_paint(const DevPaint &P) {
  if (P.kind == Z2D_PATTERN_OPAQUE) return FAST_PAINT_SOLID;
  if (P.kind == Z2D_PATTERN_GRADIENT && P.gradient) {
    const DevGradient &G = *P.gradient;
    if (G.method == Z2D_INTERP_LINEAR_RGB && G.n_stops <= FAST_STOPS && G.type <= 2u)
      return FAST_PAINT_LINEAR + G.type;
  }
  return FAST_PAINT_NONE;
}

// A solid paint's RGBA8 word (FastDraw::px) masked by coverage count k
// (mask_coverage; k = 16: the paint itself), FastDraw::srctab[k].
__device__ __forceinline__ uint32_t fast_srctab_entry(const DevPaint &P, uint32_t k) {
  const uint32_t praw = P.px.r | (P.px.g << 8) | (P.px.b << 16) | (P.px.a << 24);
  const I4 p_px{praw & 0xFFu, (praw >> 8) & 0xFFu, (praw >> 16) & 0xFFu, praw >> 24};
  const I4 m = k >= 16 ? p_px : mask_coverage(p_px, (int)k);
  return m.r | (m.g << 8) | (m.b << 16) | (m.a << 24);
}

// A fast draw's block: its region and flags, and the paint's constants
// (the solid paint's table is written by the draw's lanes 0..16 before).
__device__ void write_fast_draw(FastDraw &F, const DevDraw &D, const DevPaint &P, uint32_t pf) {
  F.rx0 = D.rx0;
  F.rx1 = D.rx1;
  F.ry0 = D.ry0;
  F.ry1 = D.ry1;
  F.flags = D.flags;
  F.rule = D.rule;
  F.paint = pf;
  F.px = P.px.r | (P.px.g << 8) | (P.px.b << 16) | (P.px.a << 24);
  F.n_stops = 0;
  F.identity = 1;
  if (pf < FAST_PAINT_LINEAR) return;
  const DevGradient &G = *P.gradient;
  F.n_stops = G.n_stops;
  F.identity = G.identity;
  F.gx = G.geom[0];
  F.gy = G.geom[1];
  for (int i = 0; i < 6; i++) F.m[i] = G.m[i];
  for (int i = 0; i < 8; i++) F.p[i] = 0.0;
  if (pf == FAST_PAINT_LINEAR) {
    F.p[0] = G.lin_dx;
    F.p[1] = G.lin_dy;
    F.p[2] = G.lin_inv;
    F.p[3] = G.lin_degenerate ? 1.0 : 0.0;
  } else if (pf == FAST_PAINT_RADIAL) {
    F.p[0] = G.r1;
    F.p[1] = G.r2;
    F.p[2] = G.cdx;
    F.p[3] = G.cdy;
    F.p[4] = G.dr;
    F.p[5] = G.min_dr;
    F.p[6] = G.a;
    F.p[7] = G.inv_a;
  } else {
    F.p[0] = G.conic_angle;
  }
  for (uint32_t i = 0; i < FAST_STOPS; i++) {
    F.offs[i] = i < G.n_stops ? G.offs[i] : __builtin_inff();
    F.pre[i] = i < G.n_stops ? G.pre[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

__global__ void __launch_bounds__(256) k_edge_build(JobView J) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t d = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (d >= J.n_draws) return;
  DevDraw &D = J.draws[d];
  const DevSurface S = J.surfaces[D.surface];
  const double scale = D.mode == MODE_DIRECT ? 1.0 : 4.0;
  z2d_edge *out = J.edges + D.node_off;
  uint32_t flags = D.flags & (DRAW_UNBALANCED | DRAW_ROWPROG | DRAW_SERIAL_PLOT);
  ExtAcc acc{0, 0.0, 0.0, 0.0, 0.0};
  if (D.input_edges) {
    for (uint32_t k = lane; k < D.node_cnt; k += 64) out[k] = J.in_edges[D.in_edge_off + k];
    acc.n = D.node_cnt;
    acc.top = D.in_ext[0]; acc.bottom = D.in_ext[1];
    acc.left = D.in_ext[2]; acc.right = D.in_ext[3];
  } else if (!(flags & DRAW_SERIAL_PLOT)) {
    plot_parallel(J.nodes + D.node_off, D.node_cnt, scale, out, acc);
  }
  // a solid paint's coverage table, for the fast renderer should the draw
  // be one (lane k: entry k)
  if (J.fast && lane <= 16 && J.paints[D.paint].kind == Z2D_PATTERN_OPAQUE)
    J.fast[d].srctab[lane] = fast_srctab_entry(J.paints[D.paint], lane);
  if (lane != 0) return;
  J.draw_rows[d] = make_int2(0, 0);  // not visible unless set below
  if (!D.input_edges && (flags & DRAW_SERIAL_PLOT)) {
    // The same state machine node by node (a line may follow a close).
    double fx = 0.0, fy = 0.0, lx = 0.0, ly = 0.0;
    int len = 0;
    const FNode *nodes = J.nodes + D.node_off;
    const uint32_t n = D.node_cnt;
    for (uint32_t i = 0; i < n; i++) {
      const FNode nd = nodes[i];
      if (nd.tag == Z2D_NODE_MOVE_TO) {
        if (i == n - 1) break;
        fx = lx = nd.x;
        fy = ly = nd.y;
        len = 1;
      } else if (nd.tag == Z2D_NODE_LINE_TO) {
        if (len == 0) continue;  // InvalidState is rejected on the host
        if (!pt_eq(lx, ly, nd.x, nd.y)) {
          add_edge(out, acc, scale, lx, ly, nd.x, nd.y);
          lx = nd.x;
          ly = nd.y;
          if (len < 3) len++;
        }
      } else if (nd.tag == Z2D_NODE_CLOSE_PATH) {
        if (len >= 3 && !pt_eq(lx, ly, fx, fy)) {
          add_edge(out, acc, scale, lx, ly, fx, fy);
          lx = fx;
          ly = fy;
        }
      }
    }
  }
  D.edge_cnt = acc.n;
  if (J.count_stats && acc.n) atomicAdd(&J.counters[6], acc.n);
  D.ext[0] = acc.top; D.ext[1] = acc.bottom; D.ext[2] = acc.left; D.ext[3] = acc.right;

  // Raster entry logic: visibility, draw window and composite region.
  const int32_t W = S.w, H = S.h;
  const bool bounded = op_is_bounded(D.op);
  if (!bounded) flags |= DRAW_UNBOUNDED;
  if (acc.n > 0) flags |= DRAW_HAS_EDGES;
  const DevPaint &P = J.paints[D.paint];
  if (P.kind == Z2D_PATTERN_OPAQUE &&
      (D.op == Z2D_OP_SRC || (D.op == Z2D_OP_SRC_OVER && P.px_opaque)))
    flags |= DRAW_FAST_SRC;
  D.rx0 = D.ry0 = D.rx1 = D.ry1 = 0;
  D.sx = D.ex = D.sy = D.ey = 0;
  D.mask_bits = 8;
  if (!in_box(D.ext, scale, W, H)) {
    D.flags = flags & ~DRAW_VISIBLE;
    return;
  }
  if (D.mode == MODE_MSAA) {
    // multisample.run (src/internal/raster/multisample.zig:46-80)
    int32_t sy = clampi((int32_t)floor(acc.top / 4.0), 0, H - 1);
    int32_t ey = clampi((int32_t)ceil(acc.bottom / 4.0), sy, H - 1);
    int32_t sx = clampi((int32_t)floor(acc.left / 4.0), 0, W - 1);
    int32_t ex = clampi((int32_t)ceil(acc.right / 4.0), sx, W);
    D.sx = sx; D.ex = ex; D.sy = sy; D.ey = ey;
    if (bounded) {
      D.rx0 = sx; D.rx1 = ex; D.ry0 = sy; D.ry1 = ey + 1;
    } else {
      D.rx0 = 0; D.rx1 = W; D.ry0 = 0; D.ry1 = H;
    }
    D.row_y0 = 4 * sy;
    D.row_n = 4 * (ey - sy + 1);
  } else if (D.mode == MODE_SSAA) {
    // supersample.run (src/internal/raster/supersample.zig:48-184)
    int32_t x0 = bounded ? (int32_t)floor(acc.left / 4.0) : 0;
    int32_t y0 = bounded ? (int32_t)floor(acc.top / 4.0) : 0;
    int32_t x1 = bounded ? (int32_t)ceil(acc.right / 4.0) : W;
    int32_t y1 = bounded ? (int32_t)ceil(acc.bottom / 4.0) : H;
    int32_t bx0 = clampi(x0 * 4, 0, W * 4 - 1), by0 = clampi(y0 * 4, 0, H * 4 - 1);
    int32_t bx1 = clampi(x1 * 4, bx0, W * 4 - 1), by1 = clampi(y1 * 4, by0, H * 4 - 1);
    int32_t dmw = (bx1 + 1 - bx0) / 4, dmh = (by1 + 1 - by0) / 4;
    int32_t cx = x0 > 0 ? x0 : 0, cy = y0 > 0 ? y0 : 0;
    int32_t cw = (dmw + cx > W) ? W - cx : dmw, ch = (dmh + cy > H) ? H - cy : dmh;
    if (cw <= 0 || ch <= 0) {
      D.flags = flags & ~DRAW_VISIBLE;
      return;
    }
    D.rx0 = cx; D.rx1 = cx + cw; D.ry0 = cy; D.ry1 = cy + ch;
    D.mask_bits = (S.fmt == Z2D_FORMAT_ALPHA4) ? 4 : (S.fmt == Z2D_FORMAT_ALPHA2) ? 2
                : (S.fmt == Z2D_FORMAT_ALPHA1) ? 1 : 8;
    D.row_y0 = by0;
    D.row_n = 4 * ch;
  } else {
    // direct.run (src/internal/raster/direct.zig:38-49)
    int32_t psy = bounded ? (int32_t)floor(acc.top) : 0;
    int32_t pey = bounded ? (int32_t)ceil(acc.bottom) : H - 1;
    int32_t sy = clampi(psy, 0, H - 1), ey = clampi(pey, sy, H - 1);
    D.sy = sy; D.ey = ey;
    if (bounded) {
      int32_t lx = (int32_t)floor(acc.left), rx = (int32_t)ceil(acc.right) + 1;
      D.rx0 = clampi(lx, 0, W); D.rx1 = clampi(rx, 0, W);
      D.ry0 = sy; D.ry1 = ey + 1;
      if (D.rx1 <= D.rx0) {
        D.flags = flags & ~DRAW_VISIBLE;
        return;
      }
    } else {
      D.rx0 = 0; D.rx1 = W; D.ry0 = 0; D.ry1 = H;
      flags |= DRAW_ROWPROG;
    }
    D.row_y0 = sy;
    D.row_n = ey - sy + 1;
  }
  D.flags = flags | DRAW_VISIBLE;
  J.draw_rows[d] = make_int2(D.ry0, D.ry1);
  if (J.fast) {
    const uint32_t pf = fast_paint(P);
    if (pf != FAST_PAINT_NONE && D.mode == MODE_MSAA && D.op == Z2D_OP_SRC_OVER &&
        D.precision == Z2D_PRECISION_INTEGER &&
        !(D.flags & (DRAW_ROWPROG | DRAW_UNBALANCED | DRAW_UNBOUNDED))) {
      D.flags |= DRAW_FAST;
      write_fast_draw(J.fast[d], D, P, pf);
    }
  }
}

// ---------------------------------------------------------------------------
// k_rowprog: the reference's serial scanline loop for one draw per thread
// (WorkingEdgeSet rescan/inc/sort/filter, src/internal/tess/Polygon.zig:
// 275-353), on a private copy of the edges.  Rescanning every sub-row is
// equivalent to rescanning at breakpoints (the set only changes there and a
// rescan of an unchanged set moves nothing), so the order history matches.
// ---------------------------------------------------------------------------
__global__ void k_rowprog(JobView J, const uint32_t *draw_list, uint32_t n_list,
                          z2d_edge *scratch_edges, int32_t *scratch_x,
                          const uint32_t *scratch_off) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_list) return;
  DevDraw &D = J.draws[draw_list[i]];
  if (!(D.flags & DRAW_VISIBLE) || !(D.flags & DRAW_ROWPROG)) return;
  const DevSurface S = J.surfaces[D.surface];
  z2d_edge *E = scratch_edges + scratch_off[i];
  int32_t *X = scratch_x + scratch_off[i];
  const uint32_t n_all = D.edge_cnt;
  for (uint32_t k = 0; k < n_all; k++) E[k] = J.edges[D.node_off + k];
  RowRec *R = J.rows + D.row_off;
  const bool direct_unbounded = D.mode == MODE_DIRECT && (D.flags & DRAW_UNBOUNDED);
  const int32_t W = S.w;
  for (int32_t r = 0; r < D.row_n; r++) {
    const int32_t ys = D.row_y0 + r;
    const double mid = (double)ys + 0.5;
    // rescan
    uint32_t to = 0;
    for (uint32_t from = 0; from < n_all; from++) {
      double top = E[from].y0 < E[from].y1 ? E[from].y0 : E[from].y1;
      double bot = E[from].y0 < E[from].y1 ? E[from].y1 : E[from].y0;
      if (top < mid && bot >= mid) {
        if (from != to) {
          z2d_edge t = E[to];
          E[to] = E[from];
          E[from] = t;
        }
        to++;
      }
    }
    // inc
    for (uint32_t k = 0; k < to; k++) {
      double top = E[k].y0 < E[k].y1 ? E[k].y0 : E[k].y1;
      X[k] = round_i32(E[k].x_start + (E[k].x_inc * (mid - top)));
    }
    // sort (insertion, stable)
    for (uint32_t a = 1; a < to; a++) {
      uint32_t b = a;
      while (b > 0 && X[b] < X[b - 1]) {
        int32_t tx = X[b]; X[b] = X[b - 1]; X[b - 1] = tx;
        z2d_edge te = E[b]; E[b] = E[b - 1]; E[b - 1] = te;
        b--;
      }
    }
    // filter
    uint32_t nf = to;
    if (D.rule == Z2D_FILL_NON_ZERO) {
      int32_t wn = 0;
      nf = 0;
      for (uint32_t from = 0; from < to; from++) {
        int dir = E[from].y0 < E[from].y1 ? -1 : 1;
        X[nf] = X[from];
        if (wn == 0) {
          wn += dir;
          nf++;
        } else {
          wn += dir;
          if (wn == 0) nf++;
        }
      }
    }
    RowRec rec;
    rec.cutoff = (nf & 1u) ? X[nf - 1] : INT32_MAX;
    rec.kind = 1;
    rec.s = rec.e = 0;
    if (direct_unbounded) {
      if (nf == 0) {
        rec.kind = 0;
      } else {
        uint32_t m = 0;
        int32_t sm = 0, em = 0;
        for (uint32_t p = 0; p < nf / 2; p++) {
          int32_t st = X[2 * p] > 0 ? X[2 * p] : 0;
          if (st >= W) break;
          int32_t en = clampi(X[2 * p + 1], st, W);
          m++;
          sm = st;
          em = en;
        }
        if (m > 0) {
          rec.kind = (m == 1) ? 2 : 3;
          rec.s = sm;
          rec.e = em;
        }
      }
    }
    R[r] = rec;
  }
}

// ---------------------------------------------------------------------------
// Flattening: the fill plotter's node walk with Spline decomposition
// (src/internal/tess/Spline.zig:37-165, fill_plotter.zig:52-118) on the
// device.  A curve starts at its predecessor's end point, so its subdivision
// is independent of the rest of the draw: k_curve_tasks splits each curve's
// recursion into FLAT_TASKS subtrees (one thread each) and k_curve_combine
// turns their counts into offsets and the curve's bounds.  k_flatten_wave
// takes one wave per fill draw and one lane per raw node: wave scans give the
// output offsets and the contour state.  Draws whose point-buffer state needs
// the whole walk (FLAT_NONSIMPLE) go through k_flatten_serial, one thread per
// draw.  The count pass (EMIT=false, once per prepared batch) returns the
// number of flattened nodes, the contour-balance and line-after-close flags,
// the (edge, band) bound and the bounding box of the tracked points, and
// stores each raw node's output offset; the emit pass (every render) writes
// the nodes.  The recursion is walked depth-first with the knots of a piece
// halved from its parent when descending and recomputed from the curve by its
// path of halvings when moving to a sibling, which reproduces the recursion's
// values exactly.
// ---------------------------------------------------------------------------
struct FPt {
  double x, y;
};
__device__ __forceinline__ bool fpeq(FPt a, FPt b) { return a.x == b.x && a.y == b.y; }
__device__ __forceinline__ FPt fhalf(FPt a, FPt b) {
  return FPt{a.x + ((b.x - a.x) / 2.0), a.y + ((b.y - a.y) / 2.0)};
}
struct FKnots {
  FPt a, b, c, d;
};
// Spline.Knots.errorSq (src/internal/tess/Spline.zig:83-123)
__device__ __forceinline__ double knots_err_sq(const FKnots &k) {
  double bx = k.b.x - k.a.x, by = k.b.y - k.a.y, cx = k.c.x - k.a.x, cy = k.c.y - k.a.y;
  if (k.a.x != k.d.x || k.a.y != k.d.y) {
    const double dx = k.d.x - k.a.x, dy = k.d.y - k.a.y;
    const double dd = dx * dx + dy * dy;
    const double bd = bx * dx + by * dy;
    if (bd >= dd) {
      bx -= dx;
      by -= dy;
    } else {
      bx -= bd / dd * dx;
      by -= bd / dd * dy;
    }
    const double cd = cx * dx + cy * dy;
    if (cd >= dd) {
      cx -= dx;
      cy -= dy;
    } else {
      cx -= cd / dd * dx;
      cy -= cd / dd * dy;
    }
  }
  const double be = bx * bx + by * by, ce = cx * cx + cy * cy;
  return be > ce ? be : ce;
}
// deCasteljau at the midpoint: the first (right = false) or second half
__device__ __forceinline__ FKnots knots_half(const FKnots &k, bool right) {
  const FPt ab = fhalf(k.a, k.b), bc = fhalf(k.b, k.c), cd = fhalf(k.c, k.d);
  const FPt abbc = fhalf(ab, bc), bccd = fhalf(bc, cd), fin = fhalf(abbc, bccd);
  return right ? FKnots{fin, bccd, cd, k.d} : FKnots{k.a, ab, abbc, fin};
}

constexpr uint32_t FLAT_MAX_VISITS = 1u << 20;  // per curve

__device__ __forceinline__ double flat_band_span(double a, double b, uint32_t nbands,
                                                 uint32_t band_h) {
  const double lo = fmin(a, b), hi = fmax(a, b);
  if (!isfinite(lo) || !isfinite(hi)) return (double)nbands;
  const double s = floor(hi / band_h) - floor(lo / band_h) + 1.0;
  return fmin(s, (double)nbands);
}

// Spline.decompose from start point a: depth-first over the halvings, leaves
// give their first point unless it is the curve's start, then d; pieces
// deeper than 64 halvings are leaves.  `pt` receives each line_to point.
template <class F>
__device__ __forceinline__ void spline_points(FPt a, FPt b, FPt c, FPt d, double tol_sq,
                                              uint32_t &flags, F &&pt) {
  if (fpeq(a, b) && fpeq(c, d)) {
    pt(d);
    return;
  }
  const FKnots root{a, b, c, d};
  uint32_t depth = 0;
  uint64_t path = 0;  // bit (depth - 1 - i): the half taken at level i
  uint32_t visits = 0;
  FKnots k = root;  // the visited piece: a first half from its parent when
                    // descending, else recomputed from the root by its path
  for (;;) {
    bool leaf = knots_err_sq(k) < tol_sq || depth > 64;
    bool runaway = false;
    if (++visits > FLAT_MAX_VISITS) {
      flags |= FLAT_RUNAWAY;
      runaway = true;
      leaf = true;
    }
    if (!leaf && depth == 64) {
      // both halves are leaves (depth 65): their first points
      if (!fpeq(k.a, a)) pt(k.a);
      const FKnots r = knots_half(k, true);
      if (!fpeq(r.a, a)) pt(r.a);
      leaf = true;
    } else if (leaf) {
      if (!fpeq(k.a, a)) pt(k.a);
    }
    if (leaf) {
      while (depth > 0 && (path & 1u)) {
        path >>= 1;
        depth--;
      }
      if (depth == 0 || runaway) break;
      path |= 1u;  // the second half of the same parent
      k = root;
      for (uint32_t i = 0; i < depth; i++) k = knots_half(k, (path >> (depth - 1 - i)) & 1u);
    } else {
      depth++;
      path <<= 1;
      k = knots_half(k, false);
    }
  }
  pt(d);
}

template <bool EMIT>
struct DevFlattener {
  FNode *out;
  uint32_t n = 0, flags = 0;
  uint64_t pairs = 0;
  uint32_t nbands, band_h;
  double cur_y = 0.0, init_y = 0.0;
  uint32_t last_tag = 0xFFFFFFFFu;
  FPt buf[3];
  int len = 0;
  double box[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};

  __device__ void track(FPt p) {
    box[0] = fmin(box[0], p.x);
    box[1] = fmax(box[1], p.x);
    box[2] = fmin(box[2], p.y);
    box[3] = fmax(box[3], p.y);
  }
  __device__ void emit(uint32_t tag, FPt p) {
    if (EMIT) out[n] = FNode{p.x, p.y, tag, 0u};
    n++;
    // the (edge, band) bound over the emitted stream, as the host computes it
    if (tag == Z2D_NODE_MOVE_TO) {
      cur_y = init_y = p.y;
    } else if (tag == Z2D_NODE_LINE_TO) {
      pairs += (uint64_t)flat_band_span(cur_y, p.y, nbands, band_h);
      cur_y = p.y;
      if (last_tag == Z2D_NODE_CLOSE_PATH) flags |= FLAT_SERIAL;
    } else if (tag == Z2D_NODE_CLOSE_PATH) {
      pairs += (uint64_t)flat_band_span(cur_y, init_y, nbands, band_h);
      cur_y = init_y;
    }
    last_tag = tag;
  }
  __device__ void add(FPt p) {
    if (len < 3) {
      buf[len++] = p;
    } else {
      buf[1] = buf[2];
      buf[2] = p;
    }
  }
  __device__ void line_to(FPt p) {
    emit(Z2D_NODE_LINE_TO, p);
    track(p);
    if (!fpeq(buf[len - 1], p)) add(p);
  }
  __device__ void curve(FPt b, FPt c, FPt d, double tol_sq) {
    spline_points(buf[len - 1], b, c, d, tol_sq, flags, [&](FPt p) { line_to(p); });
  }
};

// The per-draw walk, for the draws k_flatten_wave marked FLAT_NONSIMPLE
// (count pass) or the host marked serial (emit pass).
template <bool EMIT>
__global__ void k_flatten_serial(const z2d_path_node *__restrict__ raw,
                                 const FlatJob *__restrict__ jobs, uint32_t n_jobs, FNode *out,
                                 FlatResult *res) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  const FlatJob J = jobs[j];
  if (EMIT ? !J.serial : !(res[j].flags & FLAT_NONSIMPLE)) return;
  DevFlattener<EMIT> f;
  f.out = out + J.node_off;
  f.nbands = J.nbands;
  f.band_h = J.band_h;
  const z2d_path_node *nodes = raw + J.raw_off;
  const uint32_t n = J.raw_cnt;
  bool balanced = true;
  // the node walk of the host's checks (tags and first move already
  // validated): fill_plotter.plot with PointBuffer(1, 3)
  for (uint32_t i = 0; i < n; i++) {
    const z2d_path_node nd = nodes[i];
    const FPt p{nd.v[0], nd.v[1]};
    if (nd.tag == Z2D_NODE_MOVE_TO) {
      if (i == n - 1) {
        f.emit(Z2D_NODE_MOVE_TO, p);
        break;
      }
      if (f.len > 0 && !fpeq(f.buf[f.len - 1], f.buf[0])) balanced = false;
      f.len = 0;
      f.add(p);
      f.emit(Z2D_NODE_MOVE_TO, p);
      f.track(p);
    } else if (nd.tag == Z2D_NODE_LINE_TO) {
      f.line_to(p);
    } else if (nd.tag == Z2D_NODE_CURVE_TO) {
      const FPt b{nd.v[0], nd.v[1]}, c{nd.v[2], nd.v[3]}, d{nd.v[4], nd.v[5]};
      f.track(b);
      f.track(c);
      f.curve(b, c, d, J.tol_sq);
    } else {  // close
      f.emit(Z2D_NODE_CLOSE_PATH, FPt{0.0, 0.0});
      if (f.len >= 3 && !fpeq(f.buf[f.len - 1], f.buf[0])) f.add(f.buf[0]);
    }
  }
  if (f.len > 0 && !fpeq(f.buf[f.len - 1], f.buf[0])) balanced = false;
  if (!EMIT) {
    FlatResult r;
    r.n_nodes = f.n;
    r.flags = f.flags | (balanced ? 0u : FLAT_UNBALANCED) | FLAT_NONSIMPLE;
    r.edge_pairs = f.pairs;
    for (int i = 0; i < 4; i++) r.box[i] = f.box[i];
    res[j] = r;
  }
}

__device__ __forceinline__ void box_track(double *box, FPt p) {
  box[0] = fmin(box[0], p.x);
  box[1] = fmax(box[1], p.x);
  box[2] = fmin(box[2], p.y);
  box[3] = fmax(box[3], p.y);
}

// One wave per draw, one lane per raw node (64 at a time).  For the draws it
// keeps (every close is the last node or followed by a move, comes after a
// move, and ends a contour of at least 3 distinct points), the point buffer
// before a line or curve ends at its predecessor's end point, a close always
// returns the contour to its start, and a contour without a close is
// balanced exactly when it ends at its move point.  The distinct-point count
// is taken over the end points (a curve counted by its end point only), a
// lower bound of the buffer's length, so a draw kept here has at least 3.
template <bool EMIT>
__global__ void __launch_bounds__(256) k_flatten_wave(const z2d_path_node *__restrict__ raw,
                                                      const FlatJob *__restrict__ jobs,
                                                      uint32_t n_jobs, FNode *out,
                                                      FlatResult *res, uint32_t *rel,
                                                      const FlatCurve *__restrict__ cv,
                                                      FillSet fill) {
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  if (EMIT) {
    // the render's per-step resets (none of them is read by the flattening),
    // here instead of a launch of their own
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint32_t f = 0; f < fill.count; f++)
      for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < fill.n[f]; i += stride)
        fill.p[f][i] = fill.v[f];
  }
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (j >= n_jobs) return;
  const FlatJob J = jobs[j];
  if (EMIT && J.serial) return;
  const z2d_path_node *nodes = raw + J.raw_off;
  uint32_t *rl = rel + J.raw_off;
  FNode *o = out + J.node_off;
  const uint32_t n = J.raw_cnt;
  const uint64_t below = (1ull << lane) - 1ull;
  // carried from the previous 64 nodes (wave-uniform)
  uint32_t c_tag = NONE, c_S = 0, c_out = 0, c_mvS = 0, c_cur = J.curve_base;
  FPt c_pt{0.0, 0.0}, c_mv{0.0, 0.0};
  bool c_mv_ok = false;
  // per lane, reduced at the end
  uint64_t pairs = 0;
  uint32_t flags = 0;
  double box[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
  for (uint32_t base = 0; base < n; base += 64) {
    const uint32_t i = base + lane;
    const bool act = i < n;
    z2d_path_node nd;
    uint32_t tag = NONE;
    FPt p{0.0, 0.0};
    if (act) {
      nd = nodes[i];
      tag = nd.tag;
      if (tag == Z2D_NODE_MOVE_TO || tag == Z2D_NODE_LINE_TO) p = FPt{nd.v[0], nd.v[1]};
      else if (tag == Z2D_NODE_CURVE_TO) p = FPt{nd.v[4], nd.v[5]};
    }
    // the predecessor's tag and end point
    const uint32_t ptag = lane_prev_u32(tag, c_tag);
    const FPt pp{lane_prev_f64(p.x, c_pt.x), lane_prev_f64(p.y, c_pt.y)};
    const uint32_t ntag = (act && i + 1 < n) ? nodes[i + 1].tag : NONE;
    // S: running count of lines/curves whose end point differs from the
    // previous end point
    uint32_t S = ((tag == Z2D_NODE_LINE_TO || tag == Z2D_NODE_CURVE_TO) && !fpeq(p, pp)) ? 1u : 0u;
    S = (uint32_t)wave_incl_scan((int32_t)S) + c_S;
    // the contour's move: the last move at or before this node, and the last
    // one before it
    const uint64_t mv_mask = __ballot(tag == Z2D_NODE_MOVE_TO);
    const uint64_t at = mv_mask & (below | (1ull << lane)), before = mv_mask & below;
    const int src_at = at ? 63 - __builtin_clzll(at) : 0;
    const int src_bf = before ? 63 - __builtin_clzll(before) : 0;
    FPt mv{__shfl(p.x, src_at, 64), __shfl(p.y, src_at, 64)};
    uint32_t mvS = __shfl(S, src_at, 64);
    FPt mvb{__shfl(p.x, src_bf, 64), __shfl(p.y, src_bf, 64)};
    bool mv_ok = true;
    if (!at) {
      mv = c_mv;
      mvS = c_mvS;
      mv_ok = c_mv_ok;
    }
    if (!before) mvb = c_mv;
    uint32_t cnt = 0;
    if (tag == Z2D_NODE_MOVE_TO) {
      cnt = 1;
      if (i != n - 1) box_track(box, p);
      // the previous contour ends at the predecessor
      if ((ptag == Z2D_NODE_LINE_TO || ptag == Z2D_NODE_CURVE_TO) && !fpeq(pp, mvb))
        flags |= FLAT_UNBALANCED;
    } else if (tag == Z2D_NODE_LINE_TO) {
      cnt = 1;
      box_track(box, p);
      pairs += (uint64_t)flat_band_span(pp.y, p.y, J.nbands, J.band_h);
    } else if (tag == Z2D_NODE_CLOSE_PATH) {
      cnt = 1;
      pairs += (uint64_t)flat_band_span(pp.y, mv.y, J.nbands, J.band_h);
      if (!mv_ok || (ntag != NONE && ntag != Z2D_NODE_MOVE_TO) || 1u + S - mvS < 3u)
        flags |= FLAT_NONSIMPLE;
    }
    if ((tag == Z2D_NODE_LINE_TO || tag == Z2D_NODE_CURVE_TO) && i == n - 1 && !fpeq(p, mv))
      flags |= FLAT_UNBALANCED;
    // output offsets from the node counts (a curve's from k_curve_combine),
    // stored; the emit pass leaves curves to k_curve_tasks
    const uint64_t cm = __ballot(tag == Z2D_NODE_CURVE_TO);
    if (!EMIT && tag == Z2D_NODE_CURVE_TO) {
      const FlatCurve &r = cv[c_cur + (uint32_t)__popcll(cm & below)];
      cnt = r.cnt;
      flags |= r.flags;
      pairs += r.pairs;
      box[0] = fmin(box[0], r.box[0]);
      box[1] = fmax(box[1], r.box[1]);
      box[2] = fmin(box[2], r.box[2]);
      box[3] = fmax(box[3], r.box[3]);
    } else if (EMIT && act && tag != Z2D_NODE_CURVE_TO) {
      o[rl[i]] = FNode{tag == Z2D_NODE_CLOSE_PATH ? 0.0 : p.x,
                       tag == Z2D_NODE_CLOSE_PATH ? 0.0 : p.y, tag, 0u};
    }
    c_cur += (uint32_t)__popcll(cm);
    if (!EMIT) {
      const uint32_t x = (uint32_t)wave_incl_scan((int32_t)cnt);
      if (act) rl[i] = c_out + x - cnt;
      c_out += lane_u32(x, 63);
    }
    c_tag = lane_u32(tag, 63);
    c_pt = FPt{lane_f64(p.x, 63), lane_f64(p.y, 63)};
    c_S = lane_u32(S, 63);
    if (mv_mask) {
      const uint32_t last = 63u - (uint32_t)__builtin_clzll(mv_mask);
      c_mv = FPt{lane_f64(p.x, last), lane_f64(p.y, last)};
      c_mvS = lane_u32(S, last);
      c_mv_ok = true;
    }
  }
  if (!EMIT) {
    for (int off = 32; off > 0; off >>= 1) {
      pairs += __shfl_xor(pairs, off, 64);
      flags |= __shfl_xor(flags, off, 64);
      box[0] = fmin(box[0], __shfl_xor(box[0], off, 64));
      box[1] = fmax(box[1], __shfl_xor(box[1], off, 64));
      box[2] = fmin(box[2], __shfl_xor(box[2], off, 64));
      box[3] = fmax(box[3], __shfl_xor(box[3], off, 64));
    }
    if (lane == 0) {
      FlatResult r;
      r.n_nodes = c_out;
      r.flags = flags;
      r.edge_pairs = pairs;
      for (int q = 0; q < 4; q++) r.box[q] = box[q];
      res[j] = r;
    }
  }
}

// One piece of a curve's Spline walk (task t of FLAT_TASKS): the subtree at
// depth FLAT_TASK_DEPTH along t's bits, depth-first, or the leaf above that
// depth when t is the leftmost piece under it.  Concatenated in t order, the
// pieces give spline_points' leaf points (without the final d).
template <class F>
__device__ __forceinline__ void spline_task(FPt a, FPt b, FPt c, FPt d, double tol_sq, uint32_t t,
                                            uint32_t &flags, F &&pt) {
  const FKnots root{a, b, c, d};
  FKnots k = root;
  for (uint32_t l = 0; l < FLAT_TASK_DEPTH; l++) {
    if (knots_err_sq(k) < tol_sq) {
      if ((t & ((1u << (FLAT_TASK_DEPTH - l)) - 1u)) == 0u && !fpeq(k.a, a)) pt(k.a);
      return;
    }
    k = knots_half(k, (t >> (FLAT_TASK_DEPTH - 1 - l)) & 1u);
  }
  uint32_t depth = FLAT_TASK_DEPTH;
  uint64_t path = t;
  uint32_t visits = 0;
  // the visited piece's parent (set once the walk is below the task's
  // root): a first half's sibling is the parent's second half, the same
  // value the recomputation from the root gives
  FKnots par = k;
  for (;;) {
    bool leaf = knots_err_sq(k) < tol_sq || depth > 64;
    bool runaway = false;
    if (++visits > FLAT_MAX_VISITS) {
      flags |= FLAT_RUNAWAY;
      runaway = true;
      leaf = true;
    }
    if (!leaf && depth == 64) {
      if (!fpeq(k.a, a)) pt(k.a);
      const FKnots r = knots_half(k, true);
      if (!fpeq(r.a, a)) pt(r.a);
      leaf = true;
    } else if (leaf) {
      if (!fpeq(k.a, a)) pt(k.a);
    }
    if (leaf) {
      if (depth > FLAT_TASK_DEPTH && !(path & 1u)) {
        if (runaway) break;
        path |= 1u;  // the second half of the same parent
        k = knots_half(par, true);
        continue;
      }
      while (depth > FLAT_TASK_DEPTH && (path & 1u)) {
        path >>= 1;
        depth--;
      }
      if (depth == FLAT_TASK_DEPTH || runaway) break;
      path |= 1u;
      k = root;
      for (uint32_t i = 0; i < depth; i++) {
        if (i + 1 == depth) par = k;
        k = knots_half(k, (path >> (depth - 1 - i)) & 1u);
      }
    } else {
      depth++;
      path <<= 1;
      par = k;
      k = knots_half(k, false);
    }
  }
}

__device__ __forceinline__ void curve_points(const z2d_path_node *raw, uint32_t i, FPt &a, FPt &b,
                                             FPt &c, FPt &d) {
  // a curve of a per-node draw follows a move, line or curve
  const z2d_path_node nd = raw[i], pv = raw[i - 1];
  a = pv.tag == Z2D_NODE_CURVE_TO ? FPt{pv.v[4], pv.v[5]} : FPt{pv.v[0], pv.v[1]};
  b = FPt{nd.v[0], nd.v[1]};
  c = FPt{nd.v[2], nd.v[3]};
  d = FPt{nd.v[4], nd.v[5]};
}

// One thread per curve piece.  Count pass: the piece's point count, its
// first and last y, the (edge, band) bound between its points and its box.
// Emit pass: its points at the offset k_curve_combine gave it; the last
// piece adds the curve's end point.
template <bool EMIT>
__global__ void __launch_bounds__(256) k_curve_tasks(const z2d_path_node *__restrict__ raw,
                                                     const FlatJob *__restrict__ jobs,
                                                     const uint2 *__restrict__ curves,
                                                     uint32_t n_curves, FlatTask *tasks,
                                                     const uint32_t *__restrict__ toff,
                                                     const uint32_t *__restrict__ rel,
                                                     FNode *out) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t ci = g / FLAT_TASKS, t = g % FLAT_TASKS;
  if (ci >= n_curves) return;
  const uint2 e = curves[ci];
  const FlatJob &J = jobs[e.x];
  if (EMIT && J.serial) return;
  FPt a, b, c, d;
  curve_points(raw, e.y, a, b, c, d);
  const bool degenerate = fpeq(a, b) && fpeq(c, d);  // the curve is one line to d
  uint32_t flags = 0;
  if (EMIT) {
    FNode *o = out + J.node_off + rel[e.y] + toff[g];
    uint32_t k = 0;
    if (!degenerate)
      spline_task(a, b, c, d, J.tol_sq, t, flags,
                  [&](FPt q) { o[k++] = FNode{q.x, q.y, Z2D_NODE_LINE_TO, 0u}; });
    if (t == FLAT_TASKS - 1) o[k] = FNode{d.x, d.y, Z2D_NODE_LINE_TO, 0u};
  } else {
    FlatTask r;
    r.cnt = 0;
    r.pairs = 0;
    r.fy = r.ly = 0.0;
    r.box[0] = r.box[2] = INFINITY;
    r.box[1] = r.box[3] = -INFINITY;
    if (!degenerate)
      spline_task(a, b, c, d, J.tol_sq, t, flags, [&](FPt q) {
        if (r.cnt) r.pairs += (uint64_t)flat_band_span(r.ly, q.y, J.nbands, J.band_h);
        else r.fy = q.y;
        r.ly = q.y;
        r.cnt++;
        box_track(r.box, q);
      });
    r.flags = flags;
    tasks[g] = r;
  }
}

// Count pass, one thread per curve: piece offsets, and the curve's point
// count (pieces + its end point), (edge, band) bound from its start point
// along all its points, box (control points and points) and flags.
__global__ void __launch_bounds__(256) k_curve_combine(const z2d_path_node *__restrict__ raw,
                                                       const FlatJob *__restrict__ jobs,
                                                       const uint2 *__restrict__ curves,
                                                       uint32_t n_curves,
                                                       const FlatTask *__restrict__ tasks,
                                                       uint32_t *toff, FlatCurve *cv) {
  const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= n_curves) return;
  const uint2 e = curves[ci];
  const FlatJob &J = jobs[e.x];
  FPt a, b, c, d;
  curve_points(raw, e.y, a, b, c, d);
  FlatCurve r;
  r.cnt = 0;
  r.flags = 0;
  r.pairs = 0;
  r.box[0] = r.box[2] = INFINITY;
  r.box[1] = r.box[3] = -INFINITY;
  box_track(r.box, b);
  box_track(r.box, c);
  double cy = a.y;
  for (uint32_t t = 0; t < FLAT_TASKS; t++) {
    const FlatTask &T = tasks[ci * FLAT_TASKS + t];
    toff[ci * FLAT_TASKS + t] = r.cnt;
    r.flags |= T.flags;
    if (!T.cnt) continue;
    r.pairs += (uint64_t)flat_band_span(cy, T.fy, J.nbands, J.band_h) + T.pairs;
    cy = T.ly;
    r.cnt += T.cnt;
    r.box[0] = fmin(r.box[0], T.box[0]);
    r.box[1] = fmax(r.box[1], T.box[1]);
    r.box[2] = fmin(r.box[2], T.box[2]);
    r.box[3] = fmax(r.box[3], T.box[3]);
  }
  r.pairs += (uint64_t)flat_band_span(cy, d.y, J.nbands, J.band_h);
  box_track(r.box, d);
  r.cnt += 1;
  cv[ci] = r;
}

// ---------------------------------------------------------------------------
// Binning: ordered draw lists per band (band_h rows of one surface).
// ---------------------------------------------------------------------------
// One wave per slot: a band and 64 consecutive draws of its surface (every
// band has at least one slot).  FILL=false counts the draws whose composite
// rows meet the band; FILL=true writes, after the scan of the slot counts,
// the band list (the draws' entry indices, in draw order), the band offsets
// band_cnt and each listed entry's record (its edge list from k_edge_bins).
template <bool FILL>
__global__ void k_band_pass(JobView J) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t slot = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (slot >= J.n_slots) return;
  const uint2 si = J.slot_info[slot];  // band, surface
  const uint32_t b = si.x, s = si.y;
  const uint32_t c = slot - J.band_slot_off[b];
  const int32_t y0 = (int32_t)(b - J.surf_band_off[s]) * (int32_t)J.band_h,
                y1 = y0 + (int32_t)J.band_h;
  const uint32_t d = J.surf_draw_off[s] + c * 64 + lane, d1 = J.surf_draw_off[s + 1];
  const uint32_t lb = b - J.surf_band_off[s];  // band within the surface
  bool hit = false;
  uint32_t idx = 0;
  if (d < d1) {
    const int2 r = J.draw_rows[d];
    hit = r.x < y1 && r.y > y0;
    if (hit) {
      // the draw's entry for this band (entries are owned by draws)
      const DevDraw &D = J.draws[d];
      const uint32_t blo = D.band_lo, bhi = D.band_hi;
      if (lb < blo || lb > bhi) {
        atomicOr(&J.counters[2], 1u);  // outside the host's band estimate
        hit = false;
      } else {
        idx = D.entry_base + (lb - blo);
      }
    }
  }
  const uint64_t m = __ballot(hit);
  if (!FILL) {
    if (lane == 0) J.slot_cnt[slot] = (uint32_t)__popcll(m);
    return;
  }
  const uint32_t base = J.slot_cnt[slot];
  if (hit) {
    const uint32_t pos = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (pos < J.max_entries) {
      // only indices below the edge capacity were stored
      const uint32_t off = J.entry_ecnt[idx], cnt = J.entry_ecnt[idx + 1] - off;
      BandEntry e;
      e.draw = d;
      e.edge_off = off;
      e.edge_cnt = off >= J.max_entry_edges ? 0u : min(cnt, J.max_entry_edges - off);
      e.band = b;
      // A tile wholly right of every crossing of the entry's edges, or wholly
      // left of them, sees winding 0 (contours return to their start), so no
      // coverage; with a bounded operator outside SSAA (which composites its
      // whole mask box) nothing changes there.  Sample columns to pixels,
      // rounded outward.
      e.px_lo = INT32_MIN;
      e.px_hi = INT32_MAX;
      const DevDraw &Dd = J.draws[d];
      if (Dd.mode != MODE_SSAA &&
          !(Dd.flags & (DRAW_UNBOUNDED | DRAW_UNBALANCED | DRAW_ROWPROG | DRAW_SERIAL_PLOT))) {
        const int32_t sc = Dd.mode == MODE_DIRECT ? 1 : 4;
        const int32_t xl = J.entry_xlo[idx], xh = J.entry_xhi[idx];
        if (xl > xh) {  // no edges in the band
          e.px_lo = INT32_MAX;
          e.px_hi = INT32_MIN;
        } else if (xl != INT32_MIN && xh != INT32_MAX) {
          e.px_lo = (int32_t)floor((double)xl / sc);
          e.px_hi = (int32_t)floor((double)xh / sc);
        }
      }
      J.entries[idx] = e;
      // the strip pass's tests for this entry: a tile at column tx0 (>= 0)
      // takes it when it is not wholly beside the entry's crossings (px_lo
      // >= tx0 + TILE_W + 1 or px_hi <= tx0 - 2) and meets the composite
      // region (rx0 < tx0 + TILE_W, tx0 < rx1); a strip when its rows meet
      // [ry0, ry1)
      const int64_t lo_x = max(max((int64_t)Dd.rx0 - TILE_W, (int64_t)e.px_lo - TILE_W - 1),
                               (int64_t)-1);
      const int64_t hi_x = min(min((int64_t)Dd.rx1, (int64_t)e.px_hi + 2), (int64_t)INT32_MAX);
      uint32_t smask = 0;
      for (uint32_t st = 0; st < J.strips_per_tile; st++) {
        const int32_t sy0 = y0 + (int32_t)st * ROWS_PER_THREAD;
        if (Dd.ry1 > sy0 && Dd.ry0 < sy0 + ROWS_PER_THREAD) smask |= 1u << st;
      }
      const bool slow = !(Dd.flags & DRAW_FAST) || e.edge_cnt > PACKED_EDGE_LIMIT;
      J.band_list[pos] = make_uint4(idx, (uint32_t)(int32_t)lo_x,
                                    (uint32_t)(int32_t)max(hi_x, lo_x), smask | (slow ? 256u : 0u));
    } else {
      atomicOr(&J.counters[2], 1u);
    }
  }
  if (lane == 0) {
    if (c == 0) J.band_cnt[b] = base;
    if (slot == 0) J.band_cnt[J.n_bands] = J.slot_cnt[J.n_slots];
  }
}

// Exclusive scan of n u32 values in place; a[n] receives the total.  One
// workgroup of 1024 lanes walks the array in chunks of 16 values per lane.
__global__ void __launch_bounds__(1024) k_scan(uint32_t *a, uint32_t n) {
  constexpr uint32_t PER = 16;
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t chunk_total;
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += 1024 * PER) {
    const uint32_t i0 = base + t * PER;
    uint32_t v[PER];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
      v[k] = i0 + k < n ? a[i0 + k] : 0u;
      sum += v[k];
    }
    // wave inclusive scan of the lane sums
    const uint32_t x = (uint32_t)wave_incl_scan((int32_t)sum);
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (t < 16) {
      const uint32_t s0 = wsum[t];
      uint32_t z = s0;
      for (int off = 1; off < 16; off <<= 1) {
        const uint32_t y = __shfl_up(z, off, 16);
        if (t >= (uint32_t)off) z += y;
      }
      wsum[t] = z - s0;  // exclusive wave offsets
      if (t == 15) chunk_total = z;
    }
    __syncthreads();
    uint32_t run = carry + wsum[w] + x - sum;
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
      if (i0 + k < n) a[i0 + k] = run;
      run += v[k];
    }
    carry += chunk_total;
    __syncthreads();
  }
  if (t == 0) a[n] = carry;
}

// Exclusive scan of a[0..n) in place with a[n] = total, in one pass
// (decoupled look-back): SCAN_BLOCK values per workgroup, the workgroups
// numbered by a ticket taken when they start, so each waits only on
// workgroups already running.  A workgroup publishes its aggregate, then its
// inclusive prefix, in its status word (flag << 32 | value); its first wave
// sums its predecessors' words, 64 at a time, back to the nearest inclusive
// prefix.  stat[0] is the ticket and stat[1 + b] the status of workgroup b,
// all zero at launch.  A wait that outlasts SCAN_SPIN_LIMIT polls (never
// expected) sets bit 1 of *err and goes on, so every workgroup ends.
constexpr uint64_t SCAN_AGG = 1ull << 32, SCAN_INCL = 2ull << 32;
constexpr uint32_t SCAN_SPIN_LIMIT = 1u << 20;

__global__ void __launch_bounds__(256) k_scan_lb(uint32_t *a, uint32_t n, uint64_t *stat,
                                                 uint32_t *err) {
  __shared__ uint32_t ws[4];
  __shared__ uint32_t s_bid, s_off;
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) s_bid = (uint32_t)atomicAdd(reinterpret_cast<unsigned long long *>(stat), 1ull);
  __syncthreads();
  const uint32_t bid = s_bid;
  if (bid >= gridDim.x) return;  // (uniform) the ticket starts at zero
  const uint32_t i0 = bid * SCAN_BLOCK + t * 16;
  // a lane's 16 values: four 16-byte loads when all lie below n (a is
  // 16-byte aligned and i0 a multiple of 16)
  const bool whole = i0 + 16 <= n;
  uint32_t v[16];
  if (whole) {
#pragma unroll
    for (uint32_t q = 0; q < 4; q++) {
      const uint4 u = reinterpret_cast<const uint4 *>(a + i0)[q];
      v[4 * q] = u.x;
      v[4 * q + 1] = u.y;
      v[4 * q + 2] = u.z;
      v[4 * q + 3] = u.w;
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) v[k] = i0 + k < n ? a[i0 + k] : 0u;
  }
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < 16; k++) sum += v[k];
  const uint32_t x = (uint32_t)wave_incl_scan((int32_t)sum);
  if (lane == 63) ws[w] = x;
  __syncthreads();
  if (w == 0) {
    const uint32_t total = ws[0] + ws[1] + ws[2] + ws[3];
    uint64_t *const st = stat + 1;
    uint32_t off = 0;
    if (lane == 0)
      __hip_atomic_store(&st[bid], (bid == 0 ? SCAN_INCL : SCAN_AGG) | total, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (bid != 0) {
      uint32_t spins = 0;
      for (int32_t base = (int32_t)bid - 1;; base -= 64) {
        const int32_t j = base - (int32_t)lane;  // lane 0 the nearest predecessor
        uint64_t s = SCAN_INCL;                  // before workgroup 0: an inclusive zero
        for (;;) {
          if (j >= 0) s = __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (__ballot((s >> 32) == 0) == 0) break;
          if (++spins > SCAN_SPIN_LIMIT) {
            if (lane == 0) atomicOr(err, 2u);
            s = SCAN_INCL | (s & 0xFFFFFFFFull);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        const uint64_t pm = __ballot((s >> 32) == 2);
        const uint32_t first = pm ? (uint32_t)__builtin_ctzll(pm) : 63u;
        const uint32_t part = (uint32_t)wave_incl_scan(lane <= first ? (int32_t)(uint32_t)s : 0);
        off += (uint32_t)__builtin_amdgcn_readlane((int32_t)part, 63);
        if (pm) break;
      }
      if (lane == 0)
        __hip_atomic_store(&st[bid], SCAN_INCL | (off + total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) s_off = off;
  }
  __syncthreads();
  uint32_t run = s_off + x - sum;
  for (uint32_t j = 0; j < w; j++) run += ws[j];
  uint32_t o[16];
#pragma unroll
  for (uint32_t k = 0; k < 16; k++) {
    o[k] = run;
    run += v[k];
  }
  if (whole) {
#pragma unroll
    for (uint32_t q = 0; q < 4; q++)
      reinterpret_cast<uint4 *>(a + i0)[q] =
          make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
  } else {
#pragma unroll
    for (uint32_t k = 0; k < 16; k++)
      if (i0 + k < n) a[i0 + k] = o[k];
  }
  if (bid == gridDim.x - 1 && t == 255) a[n] = run;
}

// Exclusive scan of a[0..n) in place with a[n] = total; stat holds
// n / SCAN_BLOCK + 2 zeroed words.
static void scan_u32(uint32_t *a, uint32_t n, uint64_t *stat, uint32_t *err, hipStream_t st) {
  const uint32_t nb = n == 0 ? 1u : (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
  hipLaunchKernelGGL(k_scan_lb, dim3(nb), dim3(256), 0, st, a, n, stat, err);
}

// ---------------------------------------------------------------------------
// k_strip_pass: per strip (ROWS_PER_THREAD rows of one tile), the entries of
// its band list whose composite region meets the strip, in draw order.  One
// wave per tile walks the band's entries; FILL=false counts per strip,
// FILL=true writes after the scan.
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ void k_strip_pass(JobView J) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (t >= J.n_tiles) return;
  const TileRef T = J.tiles[t];
  const uint32_t b = J.surf_band_off[T.surface] + T.ty;
  const uint32_t e0 = min(J.band_cnt[b], J.max_entries);
  const uint32_t e1 = min(J.band_cnt[b + 1], J.max_entries);
  const uint32_t spt = J.strips_per_tile;
  const int32_t tx0 = (int32_t)T.tx * TILE_W;
  uint32_t cnt[MAX_STRIPS_PER_TILE], base[MAX_STRIPS_PER_TILE];
  uint32_t slow_strips = 0;  // strips whose list holds a general-renderer entry
#pragma unroll
  for (int s = 0; s < MAX_STRIPS_PER_TILE; s++) {
    cnt[s] = 0;
    base[s] = (FILL && (uint32_t)s < spt) ? J.strip_cnt[t * spt + s] : 0u;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t e = e0; e < e1; e += 64) {
    uint32_t mask = 0, idx = 0;
    bool slow = false;  // the entry needs the general renderer
    if (e + lane < e1) {
      const uint4 it = J.band_list[e + lane];  // see k_band_pass
      idx = it.x;
      slow = (it.w & 256u) != 0;
      if (tx0 > (int32_t)it.y && tx0 < (int32_t)it.z) mask = it.w & 255u;
    }
#pragma unroll
    for (int s = 0; s < MAX_STRIPS_PER_TILE; s++) {
      if ((uint32_t)s >= spt) break;  // uniform
      const bool hit = (mask >> s) & 1u;
      const uint64_t m = __ballot(hit);
      if (FILL && __ballot(hit && slow)) slow_strips |= 1u << s;
      if (FILL && hit) {
        const uint32_t pos = base[s] + cnt[s] + (uint32_t)__popcll(m & below);
        if (pos < J.max_strip_entries)
          J.strip_entries[pos] = idx;
        else
          atomicOr(&J.counters[2], 1u);
      }
      cnt[s] += (uint32_t)__popcll(m);
    }
  }
  if (!FILL && lane == 0) {
#pragma unroll
    for (int s = 0; s < MAX_STRIPS_PER_TILE; s++)
      if ((uint32_t)s < spt) J.strip_cnt[t * spt + s] = cnt[s];
  }
  if (FILL && lane == 0) {
    uint32_t n_slow = 0;
#pragma unroll
    for (int s = 0; s < MAX_STRIPS_PER_TILE; s++)
      if ((uint32_t)s < spt) {
        const bool sl = (slow_strips >> s) & 1u;
        J.strip_slow[t * spt + s] = sl ? 1u : 0u;
        n_slow += sl ? 1u : 0u;
      }
    if (n_slow) atomicAdd(&J.counters[9], n_slow);
  }
}

// ---------------------------------------------------------------------------
// Longest-first strip order: a counting sort of the strips by list length
// (descending), so that the render's queue hands out the long strips first
// and the short ones fill the tail.  Order within a length class is free.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t order_bucket(const JobView &J, uint32_t i) {
  const uint32_t n = J.strip_cnt[i + 1] - J.strip_cnt[i];
  return ORDER_BUCKETS - 1u - min(n, ORDER_BUCKETS - 1u);
}
__global__ void k_order_hist(JobView J, uint32_t n_strips) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_strips) atomicAdd(&J.order_hist[order_bucket(J, i)], 1u);
}
// after the scan of order_hist: each strip takes the next slot of its class
__global__ void k_order_scatter(JobView J, uint32_t n_strips) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_strips) return;
  const uint32_t pos = atomicAdd(&J.order_hist[order_bucket(J, i)], 1u);
  if (pos < n_strips) J.strip_order[pos] = i;
}

// ---------------------------------------------------------------------------
// Sample rows of an edge: the smallest y with top < y + 0.5 and the largest
// with y + 0.5 <= bottom (the active-edge test of Polygon.zig:284-285).  The
// floor estimates are corrected with those exact comparisons; values are
// clamped to [-2, 2^29] (rows outside every surface behave alike there).
// ---------------------------------------------------------------------------
constexpr double ROW_CLAMP = 536870912.0;  // 2^29
__device__ __forceinline__ int32_t row_lo(double top) {
  const double f = floor(top - 0.5) + 1.0;
  if (!(f > -2.0)) return -2;
  if (f > ROW_CLAMP) return (int32_t)ROW_CLAMP;
  int32_t y = (int32_t)f;
  if ((double)(y - 1) + 0.5 > top)
    y--;
  else if (!((double)y + 0.5 > top))
    y++;
  return y;
}
__device__ __forceinline__ int32_t row_hi(double bot) {
  const double f = floor(bot - 0.5);
  if (!(f > -2.0)) return -2;
  if (f > ROW_CLAMP) return (int32_t)ROW_CLAMP;
  int32_t y = (int32_t)f;
  if ((double)(y + 1) + 0.5 <= bot)
    y++;
  else if (!((double)y + 0.5 <= bot))
    y--;
  return y;
}
__device__ __forceinline__ REdge make_redge(const z2d_edge &e) {
  const bool down = e.y0 < e.y1;
  const double top = down ? e.y0 : e.y1, bot = down ? e.y1 : e.y0;
  REdge r;
  r.top = top;
  r.x_start = e.x_start;
  r.x_inc = e.x_inc;
  r.ylo = row_lo(top);
  r.yhi2 = row_hi(bot) * 2 + (down ? 1 : 0);
  return r;
}
// The crossing of an edge with sample row y: x_start + x_inc * ((y + 0.5) -
// top), rounded to the nearest integer, ties away from zero (the
// reference's intercept, src/internal/tess/Polygon.zig:302-307; no fused
// multiply-add).
__device__ __forceinline__ double redge_x(double top, double x_start, double x_inc, int32_t y) {
  return x_start + (x_inc * (((double)y + 0.5) - top));
}

// ---------------------------------------------------------------------------
// k_edge_bins: per (draw, band) entry, copies of the draw's edges whose
// y-range meets the band's sample rows (top < last mid && bottom >= first mid,
// the active-edge test of src/internal/raster/multisample.zig and direct.zig
// taken over the band); strips then walk only those edges.  One wave per
// group of EDGE_GROUP edge slots of a draw: each chunk of 64 edges is tested
// against the bands it can meet.  FILL=false counts per entry (atomics on
// the zeroed entry_ecnt); FILL=true, after the scan, appends to the entry's
// list through the entry_cur cursors (the order inside a list is free: the
// winding deltas of the edges are summed).
// ---------------------------------------------------------------------------
template <bool FILL>
__global__ void k_edge_bins(JobView J) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (g >= J.n_egroups) return;
  const uint2 it = J.egroups[g];  // draw, first edge slot
  const DevDraw &D = J.draws[it.x];
  if (D.mode == MODE_DIRECT && (D.flags & DRAW_UNBOUNDED)) return;  // row records instead
  const int2 rows = J.draw_rows[it.x];
  if (rows.x >= rows.y) return;  // no entries
  const uint32_t ne = D.edge_cnt;
  if (it.y >= ne) return;
  const uint32_t kend = min(it.y + EDGE_GROUP, ne);
  const int SC = D.mode == MODE_DIRECT ? 1 : 4;
  const int32_t bh = (int32_t)J.band_h;
  const int32_t rb_lo = rows.x / bh, rb_hi = (rows.y - 1) / bh;  // entry bands
  const uint32_t blo = D.band_lo, bhi = D.band_hi, ebase = D.entry_base;
  const double band_rows = (double)(bh * SC);
  const z2d_edge *edges = J.edges + D.node_off;
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t k0 = it.y; k0 < kend; k0 += 64) {
    const bool valid = k0 + lane < kend;
    z2d_edge ed{0.0, 0.0, 0.0, 0.0};
    double top = 0.0, bot = 0.0;
    int32_t lo = INT32_MAX, hi = INT32_MIN;  // candidate bands, widened by one
    int32_t xlo = INT32_MAX, xhi = INT32_MIN;  // the edge's sample columns (fill pass)
    if (valid) {
      ed = edges[k0 + lane];
      top = ed.y0 < ed.y1 ? ed.y0 : ed.y1;
      bot = ed.y0 < ed.y1 ? ed.y1 : ed.y0;
      if (FILL) {
        const double xb = ed.x_start + ed.x_inc * (bot - top);
        const double a = fmin(ed.x_start, xb), c = fmax(ed.x_start, xb);
        if (a >= -2.0e9 && c <= 2.0e9) {
          xlo = (int32_t)floor(a);
          xhi = (int32_t)floor(c);
        } else {  // huge or non-finite: the whole row
          xlo = INT32_MIN;
          xhi = INT32_MAX;
        }
      }
      const double flo = fmax(floor(top / band_rows) - 1.0, (double)rb_lo);
      const double fhi = fmin(floor(bot / band_rows) + 1.0, (double)rb_hi);
      if (flo <= fhi) {  // (a NaN edge fails the exact test below)
        lo = (int32_t)flo;
        hi = (int32_t)fhi;
      }
    }
    const REdge red = make_redge(ed);  // the copy every band of the edge receives
    const int32_t ulo = wave_min_i32(lo), uhi = wave_max_i32(hi);
    // bands in windows of 64: lane j holds the count (then the list
    // position) of band w + j, so each window takes one vector atomic
    for (int32_t w = ulo; w <= uhi; w += 64) {
      const int32_t wend = min(uhi, w + 63);
      uint32_t cnt = 0;
      for (int32_t bnd = w; bnd <= wend; bnd++) {
        const int32_t y0 = bnd * bh;
        const double ys_lo = (double)(y0 * SC) + 0.5;
        const double ys_hi = (double)((y0 + bh) * SC - 1) + 0.5;
        const bool hit = lo <= bnd && bnd <= hi && top < ys_hi && bot >= ys_lo;
        const uint32_t n = (uint32_t)__popcll(__ballot(hit));
        if ((int32_t)lane == bnd - w) cnt = n;
      }
      const int32_t bj = w + (int32_t)lane;  // this lane's band
      const bool has = cnt != 0 && bj <= wend && (uint32_t)bj >= blo && (uint32_t)bj <= bhi;
      const uint32_t idx = has ? ebase + ((uint32_t)bj - blo) : 0u;
      if (!FILL) {
        if (has) atomicAdd(&J.entry_ecnt[idx], cnt);
        continue;
      }
      const uint32_t at = has ? J.entry_ecnt[idx] + atomicAdd(&J.entry_cur[idx], cnt) : 0u;
      for (int32_t bnd = w; bnd <= wend; bnd++) {
        if ((uint32_t)bnd < blo || (uint32_t)bnd > bhi) continue;
        const int32_t y0 = bnd * bh;
        const double ys_lo = (double)(y0 * SC) + 0.5;
        const double ys_hi = (double)((y0 + bh) * SC - 1) + 0.5;
        const bool hit = lo <= bnd && bnd <= hi && top < ys_hi && bot >= ys_lo;
        const uint64_t m = __ballot(hit);
        if (m == 0) continue;
        const uint32_t base = __builtin_amdgcn_readlane(at, bnd - w);
        // the entry's column range, widened by this wave's edges in the band
        const int32_t rlo = wave_min_i32(hit ? xlo : INT32_MAX);
        const int32_t rhi = wave_max_i32(hit ? xhi : INT32_MIN);
        if (lane == 0) {
          const uint32_t eidx = ebase + ((uint32_t)bnd - blo);
          atomicMin(&J.entry_xlo[eidx], rlo);
          atomicMax(&J.entry_xhi[eidx], rhi);
        }
        if (hit) {
          const uint32_t pos = base + (uint32_t)__popcll(m & below);
          if (pos < J.max_entry_edges)
            J.entry_edges[pos] = red;
          else
            atomicOr(&J.counters[2], 1u);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// k_render: tile renderer.
// ---------------------------------------------------------------------------
// Gradients with up to this many stops are read from LDS while a draw is
// composited.
constexpr uint32_t STAGED_STOPS = 64;

// LDS of one strip worker (one wave).
struct StripShared {
  // Winding deltas per sample row and sample column of the strip, two int16
  // columns per word (column 2i in the low half).  Zero between entries:
  // the coverage phase clears what it reads.
  uint32_t acc[STRIP_SUBROWS][ACC_WORDS];
  int32_t clo[64];     // first active sample row in the strip
  uint32_t coff[64];   // exclusive offsets of active-row counts
  float soff[STAGED_STOPS];   // the draw's gradient stops, staged
  float4 spre[STAGED_STOPS];
  union {
    struct {                               // coverage (phase A)
      REdge cedge[64];                     // staged band edges
      uint8_t own[64 * STRIP_SUBROWS];     // owning edge of each work item
    } a;
    struct {                               // composite
      uint32_t ppaint[ROWS_PER_THREAD][64];  // paint samples (RGBA8)
      uint8_t plist[ROWS_PER_THREAD * 64];   // covered (row, lane) pairs
    } c;
  } u;
  uint32_t srctab[17];  // solid paint masked by each coverage count (RGBA8)
  DevPaint paint;      // the draw's paint and gradient, staged
  DevGradient grad;
};

// Loads a value every lane of the wave reads at the same address through the
// scalar data cache (constant address space: read-only for the kernel).
template <typename T>
__device__ __forceinline__ T load_uniform(const T *p) {
#if __HIP_DEVICE_COMPILE__
  return *((const __attribute__((address_space(4))) T *)p);
#else
  return *p;  // host pass only parses device code
#endif
}

// Orders this wave's LDS accesses before and after the call (lanes of one
// wave exchanging data through LDS).
// Lane `src`'s value of v (every lane of the wave active).
__device__ __forceinline__ uint32_t lane_read_u32(uint32_t v, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}
__device__ __forceinline__ double lane_read_f64(double v, uint32_t src) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = lane_read_u32((uint32_t)u, src), hi = lane_read_u32((uint32_t)(u >> 32), src);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Splits a word of two int16 deltas accumulated with 32-bit adds (the low
// half's borrows land in the high half and are undone here).
__device__ __forceinline__ void unpack_pair(uint32_t w, int32_t &lo, int32_t &hi) {
  lo = (int32_t)(int16_t)(w & 0xFFFFu);
  hi = (int32_t)(w - (uint32_t)lo) >> 16;
}



// composite_cov for the common case, kept inline: src_over in integer
// precision without a supersample mask, given the paint sample s (the same
// steps: DST_IN of the paint by the coverage alpha, then SRC_OVER).
// Paint samples of a linear or radial (TYPE) gradient pattern interpolated in linear RGB
// (Gradient.getOffset, Stop.List.search, interpolateEncodeVec of
// src/gradient.zig) for the pixels listed in sh.u.c.plist; G is wave-uniform
// and its stops are staged in sh.soff / sh.spre.
template <uint32_t TYPE>
__device__ __forceinline__ void dense_gradient(const DevGradient &G, StripShared &sh,
                                               uint32_t tot_p, uint32_t lane, int32_t tx0,
                                               int32_t ty0) {
#pragma unroll 1
  for (uint32_t it = lane; it < tot_p; it += 64) {
    const uint32_t v = sh.u.c.plist[it];
    const uint32_t pl = v & 63u, pi = v >> 6;
    const int32_t x = tx0 + (int32_t)pl, y = ty0 + (int32_t)pi;
    float off;
    if constexpr (TYPE == Z2D_GRADIENT_LINEAR)
      off = linear_offset(G, x, y);
    else
      off = radial_offset(G, x, y);
    int i0, i1;
    float t;
    bool transparent;
    stops_search(sh.soff, G.n_stops, off, i0, i1, t, transparent);
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 A = transparent ? z : sh.spre[i0];
    const float4 B = transparent ? z : sh.spre[i1];
    const uint32_t r = f2u8(255.0f * lerpf_(A.x, B.x, t));
    const uint32_t g = f2u8(255.0f * lerpf_(A.y, B.y, t));
    const uint32_t b = f2u8(255.0f * lerpf_(A.z, B.z, t));
    const uint32_t a = f2u8(255.0f * lerpf_(A.w, B.w, t));
    sh.u.c.ppaint[pi][pl] = r | (g << 8) | (b << 16) | (a << 24);
  }
}

// The rows of the src_over fast composite.  full_raw: the packed source that
// full coverage stores when the draw takes the solid fast path.
__device__ __forceinline__ void src_over_rows(uint32_t (&pix)[ROWS_PER_THREAD],
                                              const int (&kk)[ROWS_PER_THREAD], uint32_t need,
                                              bool solid, bool fast_full, int kfull,
                                              uint32_t full_raw, const StripShared &sh,
                                              uint32_t lane, uint32_t fmt, uint32_t &n_cov) {
#pragma unroll
  for (int i = 0; i < ROWS_PER_THREAD; i++) {
    if (!((need >> i) & 1u)) continue;
    n_cov++;
    const int k = kk[i];
    if (fast_full && k >= kfull) {
      pix[i] = full_raw;
      continue;
    }
    const uint32_t pv = solid ? sh.srctab[k] : sh.u.c.ppaint[i][lane];
    I4 src{pv & 0xFFu, (pv >> 8) & 0xFFu, (pv >> 16) & 0xFFu, pv >> 24};
    if (!solid && k < kfull) src = mask_coverage(src, k);
    pix[i] = composite_src_over(pix[i], fmt, src);
  }
}

__device__ __forceinline__ bool msaa_preclear(const DevDraw &D, int W, int x, int y) {
  if (y < D.sy) return true;
  if (y > D.ey) return y < W;
  if (D.sx > 0 && x < D.sx) return true;
  if (D.ex < W && x < W - D.ex) return true;
  return false;
}

__device__ __forceinline__ bool wind_in(int32_t w, bool eo) { return eo ? (w & 1) != 0 : w != 0; }

// Covered samples (of 4 x 4) per pixel of the lane's column for the strip's
// rows: each sample row's winding is the prefix sum of the deltas up to the
// sample column (lane-local over the pixel's 4 columns, a wave scan across
// lanes, two sample rows packed into one scan).  Reads and clears the deltas.
// CUT: the row records' cutoff columns apply (unbalanced contours).
template <bool EO, bool CUT>
__device__ __forceinline__ void coverage_msaa(uint32_t (&acc)[STRIP_SUBROWS][ACC_WORDS],
                                              uint32_t lane, int32_t xs_base, int32_t ys_base,
                                              const RowRec *R, int32_t row_y0, int32_t row_n,
                                              int *kk) {
  const int c_lo = (int)lane * 4;
#pragma unroll
  for (int i = 0; i < ROWS_PER_THREAD; i++) {
    int k = 0;
#pragma unroll
    for (int pr = 0; pr < 2; pr++) {
      int32_t d[2][4], tot[2];
#pragma unroll
      for (int q = 0; q < 2; q++) {
        uint2 *slot = reinterpret_cast<uint2 *>(&acc[i * 4 + pr * 2 + q][lane * 2]);
        const uint2 v = *slot;
        *slot = make_uint2(0u, 0u);
        unpack_pair(v.x, d[q][0], d[q][1]);
        unpack_pair(v.y, d[q][2], d[q][3]);
        tot[q] = d[q][0] + d[q][1] + d[q][2] + d[q][3];
      }
      const uint32_t pk = (uint32_t)tot[0] + ((uint32_t)tot[1] << 16);
      const uint32_t exc = (uint32_t)wave_incl_scan((int32_t)pk) - pk;
      int32_t base[2];
      unpack_pair(exc, base[0], base[1]);
#pragma unroll
      for (int q = 0; q < 2; q++) {
        int32_t cutoff = INT32_MAX;
        if (CUT) {
          const int32_t rr = ys_base + i * 4 + pr * 2 + q - row_y0;
          if (rr >= 0 && rr < row_n) cutoff = load_uniform(&R[rr]).cutoff;
        }
        int32_t w = base[q];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          w += d[q][j];
          const bool in = EO ? (w & 1) != 0 : w != 0;
          if (CUT)
            k += (in && xs_base + c_lo + j < cutoff) ? 1 : 0;
          else
            k += in ? 1 : 0;
        }
      }
    }
    kk[i] = k;
  }
}

// Row-paired winding deltas (the fast strip worker's layout): word [p][c]
// holds sample column c of sample rows 2p (low half) and 2p + 1 (high half)
// as lo + hi * 65536 in 32-bit arithmetic (the LDS atomics carry between the
// halves).  Every winding and partial sum of a draw whose band entry has at
// most PACKED_EDGE_LIMIT edges lies in [-32767, 32767], so the 32-bit sums
// are exact, and U = W + 0x8000 for a sum W = lo + hi * 65536 holds hi in its
// high half (mod 2^16) and lo + 0x8000 in its low half with no borrow:
// lo is odd iff bit 0 of U is set, hi iff bit 16; U ^ 0x8000 holds lo and hi
// as 16-bit halves, each zero iff that winding is.
constexpr int FAST_ACC_ROWS = STRIP_SUBROWS / 2, FAST_ACC_COLS = TILE_W * 4;

// coverage_msaa over the row-paired deltas: per pair of sample rows one
// 16-byte read of the lane's four columns, one wave scan, and per column an
// add and a packed test of both rows, counted by popcount.
// 1 in each 16-bit half of v that is nonzero, 0 in each that is zero (a
// packed unsigned min with 1; written out because the compiler otherwise
// expands the packed min into per-half compares).
__device__ __forceinline__ uint32_t pk_nonzero_u16(uint32_t v) {
  uint32_t r;
  asm("v_pk_min_u16 %0, %1, %2" : "=v"(r) : "v"(v), "s"(0x00010001u));
  return r;
}

template <bool EO>
__device__ __forceinline__ void coverage_msaa_rp(uint32_t (&acc)[FAST_ACC_ROWS][FAST_ACC_COLS],
                                                 uint32_t lane, int *kk) {
#pragma unroll
  for (int i = 0; i < ROWS_PER_THREAD; i++) {
    uint32_t k = 0;
#pragma unroll
    for (int pr = 0; pr < 2; pr++) {
      uint4 *slot = reinterpret_cast<uint4 *>(&acc[i * 2 + pr][lane * 4]);
      const uint4 p = *slot;
      *slot = make_uint4(0u, 0u, 0u, 0u);
      const uint32_t tot = p.x + p.y + p.z + p.w;
      const uint32_t exc = (uint32_t)wave_incl_scan((int32_t)tot) - tot;
      const uint32_t col[4] = {p.x, p.y, p.z, p.w};
      uint32_t u = exc + 0x8000u;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        u += col[j];
        if (EO)
          k += (uint32_t)__popc(u & 0x00010001u);
        else
          k += (uint32_t)__popc(pk_nonzero_u16(u ^ 0x8000u));
      }
    }
    kk[i] = (int)k;
  }
}

// Persistent strip workers: each wave takes (tile, strip) items from the
// queue and renders every draw of the tile's band over its ROWS_PER_THREAD x
// TILE_W pixels, which stay in registers from the first draw to the last.
// ordered: take the strips in J.strip_order (a launch over all tiles).
__global__ void __launch_bounds__(RENDER_THREADS, 3) k_render(JobView J, uint32_t tile_base,
                                                           uint32_t tile_count,
                                                           uint32_t *cursor, uint32_t ordered) {
  __shared__ StripShared shw[RENDER_THREADS / 64];
  const uint32_t lane = threadIdx.x & 63;
  StripShared &sh = shw[threadIdx.x >> 6];
  const uint32_t spt = J.strips_per_tile;
  const uint32_t n_items = tile_count * spt;
  if (load_uniform(&J.counters[9]) == 0) return;  // every strip is k_render_fast's

  // The accumulator starts at zero and every entry leaves it at zero.
  for (uint32_t i = lane; i < (uint32_t)(STRIP_SUBROWS * ACC_WORDS); i += 64)
    (&sh.acc[0][0])[i] = 0u;
  wave_lds_sync();

  uint32_t stat_pix = 0;  // work counters (count_stats), per lane
  uint32_t stat_entries = 0, stat_unpacked = 0;
  for (;;) {
    uint32_t item = 0;
    if (lane == 0) item = atomicAdd(cursor, 1u);
    item = __builtin_amdgcn_readfirstlane(item);
    if (item >= n_items) {
      if (J.count_stats) {
        // one set of atomics per wave (per-entry atomics on one address
        // serialize in L2)
        unsigned long long wpix = stat_pix;
        for (int off = 32; off > 0; off >>= 1) wpix += __shfl_xor(wpix, off, 64);
        if (lane == 0) {
          if (wpix) atomicAdd(reinterpret_cast<unsigned long long *>(&J.counters[4]), wpix);
          if (stat_entries) atomicAdd(&J.counters[7], stat_entries);
          if (stat_unpacked) atomicAdd(&J.counters[8], stat_unpacked);
        }
      }
      break;
    }
    const uint32_t sidx = ordered ? load_uniform(&J.strip_order[item]) : tile_base * spt + item;
    if (!load_uniform(&J.strip_slow[sidx])) continue;  // k_render_fast's strip
    const TileRef T = load_uniform(&J.tiles[sidx / spt]);
    const uint32_t strip = sidx % spt;
    const DevSurface S = load_uniform(&J.surfaces[T.surface]);
    const int32_t W = S.w, H = S.h;
    const int32_t tx0 = (int32_t)T.tx * TILE_W;
    const int32_t ty0 = (int32_t)T.ty * (int32_t)J.band_h + (int32_t)strip * ROWS_PER_THREAD;
    if (ty0 >= H) continue;
    const uint32_t e0 = min(load_uniform(&J.strip_cnt[sidx]), J.max_strip_entries);
    const uint32_t e1 = min(load_uniform(&J.strip_cnt[sidx + 1]), J.max_strip_entries);
    if (e0 == e1 && !S.clip) continue;
    const int32_t px = tx0 + (int32_t)lane;
    const bool col_ok = px < W;

    uint32_t pix[ROWS_PER_THREAD];
#pragma unroll
    for (int i = 0; i < ROWS_PER_THREAD; i++) {
      const int32_t py = ty0 + i;
      pix[i] = (col_ok && py < H) ? load_raw(S, (int64_t)py * W + px) : 0u;
    }
    uint32_t orig[ROWS_PER_THREAD];
#pragma unroll
    for (int i = 0; i < ROWS_PER_THREAD; i++) orig[i] = pix[i];

    for (uint32_t e = e0; e < e1; e++) {
      const BandEntry E = load_uniform(&J.entries[load_uniform(&J.strip_entries[e])]);
      const DevDraw D = load_uniform(&J.draws[E.draw]);
      // uniform: the draw's composite region misses this strip
      if (D.rx1 <= tx0 || D.rx0 >= tx0 + TILE_W) continue;
      if (D.ry1 <= ty0 || D.ry0 >= ty0 + ROWS_PER_THREAD) continue;

      const DevPaint &P = J.paints[D.paint];
      const uint32_t p_kind = load_uniform(&P.kind);
      const I4 p_px = load_uniform(&P.px);
      const DevGradient *const p_grad = load_uniform(&P.gradient);
      // a linear or radial gradient pattern interpolated in linear RGB with
      // its stops staged: sampled by dense_gradient, without the paint in LDS
      // (conic gradients go through paint_int: atan2 inlined here would
      // spill)
      uint32_t g_type = Z2D_GRADIENT_CONIC, g_method = 0, g_nstops = 0;
      if (p_grad) {
        g_type = load_uniform(&p_grad->type);
        g_method = load_uniform(&p_grad->method);
        g_nstops = load_uniform(&p_grad->n_stops);
      }
      const bool grad_dense = p_kind == Z2D_PATTERN_GRADIENT && g_type != Z2D_GRADIENT_CONIC &&
                              g_method == Z2D_INTERP_LINEAR_RGB && g_nstops <= STAGED_STOPS;
      StopView sv{nullptr, nullptr};
      if (p_kind != Z2D_PATTERN_OPAQUE && !grad_dense) {
        // the paint and its gradient in LDS for the samplers
        wave_lds_sync();
        constexpr uint32_t PW = sizeof(DevPaint) / 4, GW = sizeof(DevGradient) / 4;
        static_assert(PW <= 64 && GW <= 128, "staging sizes");
        if (lane < PW) reinterpret_cast<uint32_t *>(&sh.paint)[lane] =
            reinterpret_cast<const uint32_t *>(&J.paints[D.paint])[lane];
        if (p_grad) {
          const uint32_t *gs = reinterpret_cast<const uint32_t *>(p_grad);
          uint32_t *gd = reinterpret_cast<uint32_t *>(&sh.grad);
          if (lane < GW) gd[lane] = gs[lane];
          if (lane + 64 < GW) gd[lane + 64] = gs[lane + 64];
          if (lane == 0) sh.paint.gradient = &sh.grad;
        }
        wave_lds_sync();
      }
      if (p_grad) {
        const DevGradient G = load_uniform(p_grad);
        if (G.n_stops <= STAGED_STOPS) {
          wave_lds_sync();
          if (lane < G.n_stops) {
            sh.soff[lane] = G.offs[lane];
            sh.spre[lane] = G.pre[lane];
          }
          wave_lds_sync();
          sv = StopView{sh.soff, sh.spre};
        } else {
          sv = StopView{G.offs, G.pre};
        }
      }
      const int SC = D.mode == MODE_DIRECT ? 1 : 4;
      const int SR = ROWS_PER_THREAD * SC;  // sample rows in this strip
      const int32_t ys_base = ty0 * SC;
      const int32_t xs_base = tx0 * SC;
      const int32_t cols = TILE_W * SC;
      const bool direct_unb = D.mode == MODE_DIRECT && (D.flags & DRAW_UNBOUNDED);

      if (J.count_stats) stat_entries++;
      int kk[ROWS_PER_THREAD];
      if (direct_unb) {
        // Row records give the whole result of the reference's quirky
        // per-pair clears (src/internal/raster/direct.zig:88-124): cleared
        // pixels become transparent, the rest composite at full coverage
        // onto themselves or onto a cleared pixel.
        const RowRec *R = J.rows + D.row_off;
#pragma unroll
        for (int i = 0; i < ROWS_PER_THREAD; i++) {
          const int32_t py = ty0 + i;
          kk[i] = 0;
          if (!col_ok || py >= H) continue;
          const RowRec rr = load_uniform(&R[py - D.row_y0]);
          if (rr.kind == 1) continue;
          if (rr.kind == 0 || px < rr.s || px >= rr.e) {
            pix[i] = pack(I4{0, 0, 0, 0}, S.fmt);
            continue;
          }
          if (rr.kind != 2) pix[i] = pack(I4{0, 0, 0, 0}, S.fmt);
          kk[i] = 1;
        }
      } else {
        const REdge *edges = J.entry_edges + E.edge_off;
        const uint32_t ne = E.edge_off >= J.max_entry_edges
                                ? 0u
                                : min(E.edge_cnt, J.max_entry_edges - E.edge_off);
        const bool eo = D.rule == Z2D_FILL_EVEN_ODD;
        const bool packed = ne <= PACKED_EDGE_LIMIT;  // uniform

        // ---- phase A: winding deltas of this draw's band edges.  Each
        // crossing adds its direction at its sample column (crossings left of
        // the strip at column 0, right of it nowhere); work items are the
        // (edge, active sample row) pairs, compacted by a wave scan.
        uint32_t inner = 0;  // a crossing at a column past 0 (per lane)
        if (packed) {
          for (uint32_t c0 = 0; c0 < ne; c0 += 64) {
            const uint32_t k = c0 + lane;
            int32_t n_sr = 0, lo = 0;
            REdge ed{0.0, 0.0, 0.0, 0, 0};
            if (k < ne) {
              ed = edges[k];
              const int32_t a = max(ed.ylo, ys_base), b = min(ed.yhi2 >> 1, ys_base + SR - 1);
              if (a <= b) {
                n_sr = b - a + 1;
                lo = a - ys_base;
                // every crossing right of the strip: no effect here (the
                // crossings are monotone in the row, so the extremes are at
                // the first and last active rows)
                const double xa = redge_x(ed.top, ed.x_start, ed.x_inc, a);
                const double xb = redge_x(ed.top, ed.x_start, ed.x_inc, b);
                if (fmin(xa, xb) > (double)(xs_base + cols) + 1.0) n_sr = 0;
              }
            }
            const int32_t incl = wave_incl_scan(n_sr);
            const uint32_t total = (uint32_t)__builtin_amdgcn_readlane(incl, 63);
            if (total == 0) continue;  // uniform
            sh.u.a.cedge[lane] = ed;
            sh.clo[lane] = lo;
            const uint32_t my_off = (uint32_t)(incl - n_sr);
            sh.coff[lane] = my_off;
            // each edge marks its own items (at most SR per edge)
            for (int t = 0; t < n_sr; t++) sh.u.a.own[my_off + (uint32_t)t] = (uint8_t)lane;
            wave_lds_sync();
            for (uint32_t it = lane; it < total; it += 64) {
              const uint32_t lj = sh.u.a.own[it];
              const REdge ce = sh.u.a.cedge[lj];
              const int sr = sh.clo[lj] + (int)(it - sh.coff[lj]);
              const int32_t x = round_i32(redge_x(ce.top, ce.x_start, ce.x_inc, ys_base + sr));
              const int32_t xr = x - xs_base;
              if (xr >= cols) continue;
              const int32_t col = xr < 0 ? 0 : xr;
              inner |= (uint32_t)col;
              const int32_t dv = eo ? 1 : ((ce.yhi2 & 1) ? -1 : 1);
              atomicAdd(&sh.acc[sr][col >> 1], (col & 1) ? ((uint32_t)dv << 16) : (uint32_t)dv);
            }
            wave_lds_sync();
          }
        }

        // ---- phase B: coverage per pixel (samples covered of SC x SC), then
        // the composite in draw order.
        const RowRec *R = (D.flags & DRAW_ROWPROG) ? J.rows + D.row_off : nullptr;
        const int c_lo = (int)lane * SC;
        // every delta at column 0: each sample row's winding is the same
        // across the strip, so a pixel row's coverage is 4 samples per
        // sample row inside (no prefix sums)
        const bool uniform = packed && SC == 4 && !R && __ballot(inner != 0) == 0;
        if (uniform) {
          uint32_t w0 = 0;  // lane < 16: sample row `lane`'s column-0 word
          if (lane < (uint32_t)SR) {
            w0 = sh.acc[lane][0];
            sh.acc[lane][0] = 0u;
          }
          int32_t lo, hi;
          unpack_pair(w0, lo, hi);
          const uint64_t in_rows = __ballot(lane < (uint32_t)SR && wind_in(lo, eo));
#pragma unroll
          for (int i = 0; i < ROWS_PER_THREAD; i++)
            kk[i] = 4 * __popcll((in_rows >> (4 * i)) & 0xFull);
        } else if (packed && SC == 4) {
          if (eo) {
            if (R) coverage_msaa<true, true>(sh.acc, lane, xs_base, ys_base, R, D.row_y0, D.row_n, kk);
            else coverage_msaa<true, false>(sh.acc, lane, xs_base, ys_base, R, 0, 0, kk);
          } else {
            if (R) coverage_msaa<false, true>(sh.acc, lane, xs_base, ys_base, R, D.row_y0, D.row_n, kk);
            else coverage_msaa<false, false>(sh.acc, lane, xs_base, ys_base, R, 0, 0, kk);
          }
        } else if (packed) {
          // one sample per pixel: rows paired into one scan
  #pragma unroll
          for (int pr = 0; pr < ROWS_PER_THREAD / 2; pr++) {
            int32_t d[2];
  #pragma unroll
            for (int q = 0; q < 2; q++) {
              uint32_t *slot = &sh.acc[pr * 2 + q][lane >> 1];
              const uint32_t v = *slot;
              if ((lane & 1) == 0) *slot = 0u;
              int32_t lo, hi;
              unpack_pair(v, lo, hi);
              d[q] = (lane & 1) ? hi : lo;
            }
            const uint32_t pk = (uint32_t)d[0] + ((uint32_t)d[1] << 16);
            const uint32_t inc = (uint32_t)wave_incl_scan((int32_t)pk);
            int32_t w[2];
            unpack_pair(inc, w[0], w[1]);
  #pragma unroll
            for (int q = 0; q < 2; q++) {
              int32_t cutoff = INT32_MAX;
              if (R) {
                const int32_t rr = ys_base + pr * 2 + q - D.row_y0;
                if (rr >= 0 && rr < D.row_n) cutoff = load_uniform(&R[rr]).cutoff;
              }
              kk[pr * 2 + q] = (wind_in(w[q], eo) && xs_base + c_lo < cutoff) ? 1 : 0;
            }
          }
        } else {
          // more band edges than the packed deltas can hold: each lane
          // evaluates the crossings of its own sample rows
          if (J.count_stats) stat_unpacked++;
  #pragma unroll
          for (int i = 0; i < ROWS_PER_THREAD; i++) {
            int k = 0;
            for (int o = 0; o < SC; o++) {
              const int32_t ys = ys_base + i * SC + o;
              const double mid = (double)ys + 0.5;
              int32_t w[4] = {0, 0, 0, 0};
              for (uint32_t q = 0; q < ne; q++) {
                const REdge ed = edges[q];
                if (ys < ed.ylo || ys > (ed.yhi2 >> 1)) continue;
                const int32_t x = round_i32(ed.x_start + (ed.x_inc * (mid - ed.top)));
                const int32_t dv = eo ? 1 : ((ed.yhi2 & 1) ? -1 : 1);
                const int32_t xr = x - xs_base;
  #pragma unroll
                for (int j = 0; j < 4; j++)
                  if (j < SC && xr <= c_lo + j) w[j] += dv;
              }
              int32_t cutoff = INT32_MAX;
              if (R) {
                const int32_t rr = ys - D.row_y0;
                if (rr >= 0 && rr < D.row_n) cutoff = load_uniform(&R[rr]).cutoff;
              }
  #pragma unroll
              for (int j = 0; j < 4; j++)
                if (j < SC && wind_in(w[j], eo) && xs_base + c_lo + j < cutoff) k++;
            }
            kk[i] = k;
          }
        }
        wave_lds_sync();
      }

      uint32_t n_cov = 0;
      if (D.op == Z2D_OP_SRC_OVER && D.precision == Z2D_PRECISION_INTEGER &&
          D.mode != MODE_SSAA) {
        uint32_t need = 0;  // covered pixels inside the composite region
#pragma unroll
        for (int i = 0; i < ROWS_PER_THREAD; i++) {
          const int32_t py = ty0 + i;
          if (col_ok && py < H && px >= D.rx0 && px < D.rx1 && py >= D.ry0 && py < D.ry1 &&
              kk[i] > 0)
            need |= 1u << i;
        }
        const bool solid = p_kind == Z2D_PATTERN_OPAQUE;
        const int kfull = SC * SC;
        if (__ballot(need != 0) == 0) continue;  // nothing covered in this strip (uniform)
        if (solid) {
          // the 17 possible masked sources of a solid paint
          if (lane <= 16) {
            const I4 m = (int)lane >= kfull ? p_px : mask_coverage(p_px, (int)lane);
            sh.srctab[lane] = m.r | (m.g << 8) | (m.b << 16) | (m.a << 24);
          }
          wave_lds_sync();
        } else {
          // paint samples of the covered pixels, evaluated densely: the
          // (row, lane) pairs are listed with a wave scan and taken 64 at a
          // time by one evaluation site; results go to LDS by pixel
          const int32_t np = __popc(need);
          const int32_t incl_p = wave_incl_scan(np);
          const uint32_t tot_p = (uint32_t)__builtin_amdgcn_readlane(incl_p, 63);
          uint32_t o = (uint32_t)(incl_p - np);
#pragma unroll
          for (int i = 0; i < ROWS_PER_THREAD; i++)
            if ((need >> i) & 1u) sh.u.c.plist[o++] = (uint8_t)((i << 6) | lane);
          wave_lds_sync();
          if (grad_dense) {
            const DevGradient G = load_uniform(p_grad);
            if (G.type == Z2D_GRADIENT_LINEAR)
              dense_gradient<Z2D_GRADIENT_LINEAR>(G, sh, tot_p, lane, tx0, ty0);
            else
              dense_gradient<Z2D_GRADIENT_RADIAL>(G, sh, tot_p, lane, tx0, ty0);
          } else {
#pragma unroll 1
            for (uint32_t it = lane; it < tot_p; it += 64) {
              const uint32_t v = sh.u.c.plist[it];
              const uint32_t pl = v & 63u, pi = v >> 6;
              const I4 c = paint_int(sh.paint, sv, J.blue_noise, tx0 + (int32_t)pl,
                                     ty0 + (int32_t)pi);
              sh.u.c.ppaint[pi][pl] = c.r | (c.g << 8) | (c.b << 16) | (c.a << 24);
            }
          }
          wave_lds_sync();
        }
        const bool fast_full = (D.flags & DRAW_FAST_SRC) != 0;
        // the solid fast path stores the source where coverage is full
        const uint32_t full_raw = fast_full ? pack(p_px, S.fmt) : 0u;
        src_over_rows(pix, kk, need, solid, fast_full, kfull, full_raw, sh, lane, S.fmt, n_cov);
      } else {
        // every other operator / precision / mode: one composite instance,
        // rows in a rolled loop over a copy of the strip's pixels
        uint32_t tp[ROWS_PER_THREAD];
#pragma unroll
        for (int i = 0; i < ROWS_PER_THREAD; i++) tp[i] = pix[i];
#pragma unroll 1
        for (int i = 0; i < ROWS_PER_THREAD; i++) {
          const int32_t py = ty0 + i;
          if (!col_ok || py >= H) continue;
          const bool in_region = px >= D.rx0 && px < D.rx1 && py >= D.ry0 && py < D.ry1;
          if (!in_region) continue;
          uint32_t v = tp[i];
          if (D.mode == MODE_MSAA && (D.flags & DRAW_UNBOUNDED) && msaa_preclear(D, W, px, py))
            v = pack(I4{0, 0, 0, 0}, S.fmt);
          const int k = kk[i];
          n_cov += k > 0;
          tp[i] = composite_cov(v, S.fmt, k, SC * SC, D, P, sv, J.blue_noise, px, py);
        }
#pragma unroll
        for (int i = 0; i < ROWS_PER_THREAD; i++) pix[i] = tp[i];
      }
      if (J.count_stats) stat_pix += n_cov;  // per lane
    }

    if (S.clip) {
      // Clip stage: Surface.composite(dst, mask, .dst_in) after the draws.
#pragma unroll
      for (int i = 0; i < ROWS_PER_THREAD; i++) {
        const int32_t py = ty0 + i;
        if (!col_ok || py >= H) continue;
        const I4 m{0, 0, 0, S.clip[(int64_t)py * W + px]};
        pix[i] = pack(op_int(Z2D_OP_DST_IN, unpack(pix[i], S.fmt), m), S.fmt);
      }
    }
#pragma unroll
    for (int i = 0; i < ROWS_PER_THREAD; i++) {
      const int32_t py = ty0 + i;
      if (col_ok && py < H && pix[i] != orig[i]) store_raw(S, (int64_t)py * W + px, pix[i]);
    }
  }
}

// ---------------------------------------------------------------------------
// k_render_fast: the strip worker for the common draw (DRAW_FAST: multisample,
// src_over in integer precision, a solid paint or a linear-RGB gradient of
// at most FAST_STOPS stops, an edge list the packed deltas hold), over the
// strips whose every entry is such a draw (k_render skips them).  The
// coverage and composite arithmetic is k_render's; the difference is how the
// work reaches the wave: one wave per workgroup, the entries of a strip
// gathered 64 at a time into registers, and each draw's constants (its
// FastDraw block, with the gradient's stops inline) loaded one entry ahead
// of their use, so the dependent chain strip list -> entry -> draw -> paint
// -> gradient no longer stalls every entry.
// ---------------------------------------------------------------------------
struct FastShared {
  // winding deltas, row-paired (coverage_msaa_rp; zero between entries)
  uint32_t acc[FAST_ACC_ROWS][FAST_ACC_COLS];
  union {
    struct {                             // coverage (phase A)
      // owning lane of each work item (0x80: delta -1); the owner's edge
      // terms are read from its registers (ds_bpermute), not staged here,
      // which keeps the block within 10 KiB: 16 workgroups per CU
      uint8_t own[64 * STRIP_SUBROWS];
    } a;
    struct {                             // composite
      uint32_t ppaint[ROWS_PER_THREAD][64];  // paint samples (RGBA8)
      uint8_t plist[ROWS_PER_THREAD * 64];   // covered (row, lane) pairs
    } c;
  } u;
  FastDraw blk;  // the current entry's draw constants
  // winding of the edges wholly left of the strip, as a difference array
  // over its sample rows (zero between entries)
  int32_t ldiff[STRIP_SUBROWS + 1];
};

// Gradient.getOffset (src/gradient.zig:349-375, :605-652, :731-742) on a
// fast draw's block: the arithmetic of linear_offset / radial_offset /
// conic_offset, term for term.
__device__ __forceinline__ float fast_linear_offset(const FastDraw &g, int x, int y) {
  double px = (double)x + 0.5, py = (double)y + 0.5;
  if (!g.identity) user_to_device(g.m, px, py);
  if (g.p[3] != 0.0) return -1.0f;
  double spx = px - g.gx, spy = py - g.gy;
  double dt = 0.0;
  dt += g.p[0] * spx;
  dt += g.p[1] * spy;
  return (float)dclamp01(dt * g.p[2]);
}
__device__ __forceinline__ float fast_radial_offset(const FastDraw &g, int x, int y) {
  double px = (double)x + 0.5, py = (double)y + 0.5;
  const double r1 = g.p[0], r2 = g.p[1], cdx = g.p[2], cdy = g.p[3], dr = g.p[4],
               min_dr = g.p[5], a = g.p[6], inv_a = g.p[7];
  if (r1 == 0.0 && r2 == 0.0) return -1.0f;
  if (!g.identity) user_to_device(g.m, px, py);
  double pdx = px - g.gx, pdy = py - g.gy;
  double b = 0.0;
  b += pdx * cdx;
  b += pdy * cdy;
  b += r1 * dr;
  double c = 0.0;
  c += pdx * pdx;
  c += pdy * pdy;
  c += (-r1) * r1;
  if (a == 0.0) {
    if (b == 0.0) return -1.0f;
    double t = 0.5 * c / b;
    if (t * dr >= min_dr) return (float)dclamp01(t);
    return -1.0f;
  }
  double discr = 0.0;
  discr += b * b;
  discr += a * (-c);
  if (discr >= 0.0) {
    double sq = sqrt(discr);
    double t0 = (b + sq) * inv_a;
    double t1 = (b - sq) * inv_a;
    if (t0 * dr >= min_dr) return (float)dclamp01(t0);
    if (t1 * dr >= min_dr) return (float)dclamp01(t1);
  }
  return -1.0f;
}
__device__ __forceinline__ float fast_conic_offset(const FastDraw &g, int x, int y) {
  double px = (double)x + 0.5, py = (double)y + 0.5;
  const double two_pi = 3.14159265358979323846 * 2.0;
  if (!g.identity) user_to_device(g.m, px, py);
  double ang = zmodd(zatan2(py - g.gy, px - g.gx) - g.p[0], two_pi);
  return (float)(ang / two_pi);
}

// Paint samples of the covered pixels listed in plist (as dense_gradient).
template <uint32_t PF>
__device__ __forceinline__ void fast_gradient(const FastDraw &B, FastShared &sh, uint32_t tot_p,
                                              uint32_t lane, int32_t tx0, int32_t ty0) {
  const uint32_t n_stops = min(__builtin_amdgcn_readfirstlane(B.n_stops), FAST_STOPS);
#pragma unroll 1
  for (uint32_t it = lane; it < tot_p; it += 64) {
    // the gradient's constants are read from the block in LDS on each pass
    // instead of being held in registers across the loop
    asm volatile("" ::: "memory");
    const uint32_t v = sh.u.c.plist[it];
    const uint32_t pl = v & 63u, pi = v >> 6;
    const int32_t x = tx0 + (int32_t)pl, y = ty0 + (int32_t)pi;
    float off;
    if constexpr (PF == FAST_PAINT_LINEAR)
      off = fast_linear_offset(B, x, y);
    else if constexpr (PF == FAST_PAINT_RADIAL)
      off = fast_radial_offset(B, x, y);
    else
      off = fast_conic_offset(B, x, y);
    int i0, i1;
    float t;
    bool transparent;
    stops_find8(B.offs, n_stops, off, i0, i1, t, transparent);
    // both stop colors read whole (i0, i1 are valid indices either way)
    float4 A = B.pre[i0], C2 = B.pre[i1];
    if (transparent) A = C2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t r = f2u8(255.0f * lerpf_(A.x, C2.x, t));
    const uint32_t g = f2u8(255.0f * lerpf_(A.y, C2.y, t));
    const uint32_t b = f2u8(255.0f * lerpf_(A.z, C2.z, t));
    const uint32_t a = f2u8(255.0f * lerpf_(A.w, C2.w, t));
    sh.u.c.ppaint[pi][pl] = r | (g << 8) | (b << 16) | (a << 24);
  }
}

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int32_t rfl(int32_t v) {
  return (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
}

// At most 128 VGPRs (four waves per SIMD; LDS allows fourteen per CU).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_render_fast(JobView J, uint32_t tile_base,
                                                     uint32_t tile_count, uint32_t *cursor,
                                                     uint32_t ordered) {
  __shared__ FastShared sh;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t spt = J.strips_per_tile;
  const uint32_t n_items = tile_count * spt;
  for (uint32_t i = lane; i < (uint32_t)(FAST_ACC_ROWS * FAST_ACC_COLS); i += 64)
    (&sh.acc[0][0])[i] = 0u;
  if (lane <= (uint32_t)STRIP_SUBROWS) sh.ldiff[lane] = 0;
  wave_lds_sync();

  uint32_t stat_pix = 0, stat_entries = 0;  // work counters (count_stats)
  for (;;) {
    uint32_t item = 0;
    if (lane == 0) item = atomicAdd(cursor, 1u);
    item = __builtin_amdgcn_readfirstlane(item);
    if (item >= n_items) {
      if (J.count_stats) {
        unsigned long long wpix = stat_pix;
        for (int off = 32; off > 0; off >>= 1) wpix += __shfl_xor(wpix, off, 64);
        if (lane == 0) {
          if (wpix) atomicAdd(reinterpret_cast<unsigned long long *>(&J.counters[4]), wpix);
          if (stat_entries) atomicAdd(&J.counters[7], stat_entries);
        }
      }
      break;
    }
    const uint32_t sidx = ordered ? load_uniform(&J.strip_order[item]) : tile_base * spt + item;
    if (load_uniform(&J.strip_slow[sidx])) continue;  // k_render's strip
    const TileRef T = load_uniform(&J.tiles[sidx / spt]);
    const uint32_t strip = sidx % spt;
    const DevSurface S = load_uniform(&J.surfaces[T.surface]);
    const int32_t W = S.w, H = S.h;
    const int32_t tx0 = (int32_t)T.tx * TILE_W;
    const int32_t ty0 = (int32_t)T.ty * (int32_t)J.band_h + (int32_t)strip * ROWS_PER_THREAD;
    if (ty0 >= H) continue;
    const uint32_t e0 = min(load_uniform(&J.strip_cnt[sidx]), J.max_strip_entries);
    const uint32_t e1 = min(load_uniform(&J.strip_cnt[sidx + 1]), J.max_strip_entries);
    if (e0 == e1 && !S.clip) continue;
    const int32_t px = tx0 + (int32_t)lane;
    const bool col_ok = px < W;
    uint32_t pix[ROWS_PER_THREAD];
    uint32_t dirty = 0;  // rows whose pixel a draw composited on or the clip changed
    uint32_t vrows = 0;  // rows whose pixel of this lane's column is on the surface
#pragma unroll
    for (int i = 0; i < ROWS_PER_THREAD; i++) {
      const int32_t py = ty0 + i;
      const bool on = col_ok && py < H;
      pix[i] = on ? load_raw(S, (int64_t)py * W + px) : 0u;
      vrows |= (on ? 1u : 0u) << i;
    }
    const int32_t ys_base = ty0 * 4, xs_base = tx0 * 4;
    constexpr int32_t COLS = TILE_W * 4, SR = ROWS_PER_THREAD * 4;

    for (uint32_t c = e0; c < e1; c += 64) {
      // this chunk's entries, one per lane
      const uint32_t nc = min(64u, e1 - c);
      uint32_t g_draw = 0, g_off = 0, g_cnt = 0;
      if (lane < nc) {
        const BandEntry E = J.entries[J.strip_entries[c + lane]];
        g_draw = E.draw;
        g_off = E.edge_off;
        g_cnt = E.edge_off >= J.max_entry_edges ? 0u
                                                : min(E.edge_cnt, J.max_entry_edges - E.edge_off);
      }
      uint4 nb = make_uint4(0u, 0u, 0u, 0u);  // the next entry's block, 16 B per lane
      {
        const uint32_t d0 = __builtin_amdgcn_readlane(g_draw, 0);
        if (lane < FAST_DRAW_CHUNKS) nb = reinterpret_cast<const uint4 *>(J.fast + d0)[lane];
      }
      for (uint32_t j = 0; j < nc; j++) {
        // entry j's block to LDS, entry j+1's in flight
        wave_lds_sync();
        if (lane < FAST_DRAW_CHUNKS) reinterpret_cast<uint4 *>(&sh.blk)[lane] = nb;
        wave_lds_sync();
        if (j + 1 < nc) {
          const uint32_t dn = __builtin_amdgcn_readlane(g_draw, j + 1);
          if (lane < FAST_DRAW_CHUNKS) nb = reinterpret_cast<const uint4 *>(J.fast + dn)[lane];
        }
        const FastDraw &B = sh.blk;
        const int32_t rx0 = rfl(B.rx0), rx1 = rfl(B.rx1), ry0 = rfl(B.ry0), ry1 = rfl(B.ry1);
        // uniform: the draw's composite region misses this strip
        if (rx1 <= tx0 || rx0 >= tx0 + TILE_W) continue;
        if (ry1 <= ty0 || ry0 >= ty0 + ROWS_PER_THREAD) continue;
        const uint32_t pf = rfl(B.paint), bflags = rfl(B.flags);
        const bool eo = rfl(B.rule) == Z2D_FILL_EVEN_ODD;
        const uint32_t ne = __builtin_amdgcn_readlane(g_cnt, j);
        const REdge *edges = J.entry_edges + __builtin_amdgcn_readlane(g_off, j);
        if (J.count_stats) stat_entries++;

        // ---- phase A: winding deltas of the band edges (as k_render; an
        // edge wholly left of the strip adds its direction at column 0 of
        // each of its rows through two entries of a row difference array)
        uint32_t inner = 0;
        bool any_left = false;
        for (uint32_t c0 = 0; c0 < ne; c0 += 64) {
          const uint32_t k = c0 + lane;
          int32_t n_sr = 0, lo = 0;
          REdge ed{0.0, 0.0, 0.0, 0, 0};
          if (k < ne) {
            ed = edges[k];
            const int32_t a = max(ed.ylo, ys_base), b = min(ed.yhi2 >> 1, ys_base + SR - 1);
            if (a <= b) {
              n_sr = b - a + 1;
              lo = a - ys_base;
              const double xa = redge_x(ed.top, ed.x_start, ed.x_inc, a);
              const double xb = redge_x(ed.top, ed.x_start, ed.x_inc, b);
              if (fmin(xa, xb) > (double)(xs_base + COLS) + 1.0) {
                n_sr = 0;  // every crossing right of the strip: none here
              } else if (fmax(xa, xb) < (double)xs_base - 1.0) {
                // every crossing left of the strip (rounded below its first
                // sample column): column 0 of rows [lo, lo + n_sr)
                const int32_t dv = (!eo && (ed.yhi2 & 1)) ? -1 : 1;
                atomicAdd(&sh.ldiff[lo], dv);
                atomicAdd(&sh.ldiff[lo + n_sr], -dv);
                any_left = true;
                n_sr = 0;
              }
            }
          }
          const int32_t incl = wave_incl_scan(n_sr);
          const uint32_t total = (uint32_t)__builtin_amdgcn_readlane(incl, 63);
          if (total == 0) continue;  // uniform
          const uint32_t my_off = (uint32_t)(incl - n_sr);
          const uint32_t lo_off = ((uint32_t)lo << 16) | my_off;
          const uint8_t tag = (uint8_t)(lane | ((!eo && (ed.yhi2 & 1)) ? 0x80u : 0u));
          for (int t = 0; t < n_sr; t++) sh.u.a.own[my_off + (uint32_t)t] = tag;
          wave_lds_sync();
          // every lane takes part in each round's register reads (the owner
          // of an item may be a lane with no item of its own this round)
          for (uint32_t it0 = 0; it0 < total; it0 += 64) {
            const uint32_t it = it0 + lane;
            const bool live = it < total;
            const uint32_t v = live ? (uint32_t)sh.u.a.own[it] : 0u;
            const uint32_t lj = v & 63u;
            const uint32_t q = lane_read_u32(lo_off, lj);
            const double etop = lane_read_f64(ed.top, lj), exs = lane_read_f64(ed.x_start, lj),
                         exi = lane_read_f64(ed.x_inc, lj);
            if (!live) continue;
            const int sr = (int)(q >> 16) + (int)(it - (q & 0xFFFFu));
            const int32_t x = round_i32(redge_x(etop, exs, exi, ys_base + sr));
            const int32_t xr = x - xs_base;
            if (xr >= COLS) continue;
            const int32_t col = xr < 0 ? 0 : xr;
            inner |= (uint32_t)col;
            const uint32_t dv = (v & 0x80u) ? 0xFFFFFFFFu : 1u;
            atomicAdd(&sh.acc[sr >> 1][col], (sr & 1) ? (dv << 16) : dv);
          }
          wave_lds_sync();
        }

        if (__ballot(any_left)) {
          // the left edges' winding per row: prefix of the difference array
          wave_lds_sync();
          int32_t dl = 0;
          if (lane <= (uint32_t)SR) {
            dl = lane < (uint32_t)SR ? sh.ldiff[lane] : 0;
            sh.ldiff[lane] = 0;
          }
          const int32_t wl = wave_incl_scan(dl);
          if (lane < (uint32_t)SR && wl != 0)
            atomicAdd(&sh.acc[lane >> 1][0], (lane & 1) ? ((uint32_t)wl << 16) : (uint32_t)wl);
          wave_lds_sync();
        }

        // ---- phase B: covered samples per pixel
        int kk[ROWS_PER_THREAD];
        if (__ballot(inner != 0) == 0) {
          // every delta at column 0: uniform winding across each sample row
          // (lane = sample row; the even lane clears the pair's word)
          uint32_t w0 = 0;
          if (lane < (uint32_t)SR) w0 = sh.acc[lane >> 1][0];
          wave_lds_sync();
          if (lane < (uint32_t)SR && !(lane & 1)) sh.acc[lane >> 1][0] = 0u;
          int32_t wl, wh;
          unpack_pair(w0, wl, wh);
          const int32_t w_row = (lane & 1) ? wh : wl;
          const uint64_t in_rows = __ballot(lane < (uint32_t)SR && wind_in(w_row, eo));
#pragma unroll
          for (int i = 0; i < ROWS_PER_THREAD; i++)
            kk[i] = 4 * __popcll((in_rows >> (4 * i)) & 0xFull);
        } else if (eo) {
          coverage_msaa_rp<true>(sh.acc, lane, kk);
        } else {
          coverage_msaa_rp<false>(sh.acc, lane, kk);
        }
        wave_lds_sync();

        // ---- composite: src_over of the masked paint
        // covered pixels inside the composite region: rows [r_lo, r_hi) of
        // the strip lie in [ry0, ry1) (uniform; the region meets the strip)
        const int32_t r_lo = max(ry0, ty0) - ty0, r_hi = min(ry1, ty0 + ROWS_PER_THREAD) - ty0;
        const uint32_t rmask = (1u << r_hi) - (1u << r_lo);
        uint32_t kbits = 0;
#pragma unroll
        for (int i = 0; i < ROWS_PER_THREAD; i++) kbits |= (kk[i] > 0 ? 1u : 0u) << i;
        const uint32_t need = (px >= rx0 && px < rx1) ? (vrows & rmask & kbits) : 0u;
        if (__ballot(need != 0) == 0) continue;  // nothing covered in this strip (uniform)
        const uint32_t praw = rfl(B.px);
        const I4 p_px{praw & 0xFFu, (praw >> 8) & 0xFFu, (praw >> 16) & 0xFFu, praw >> 24};
        const bool solid = pf == FAST_PAINT_SOLID;
        if (!solid) {
          const int32_t np = __popc(need);
          const int32_t incl_p = wave_incl_scan(np);
          const uint32_t tot_p = (uint32_t)__builtin_amdgcn_readlane(incl_p, 63);
          uint32_t o = (uint32_t)(incl_p - np);
#pragma unroll
          for (int i = 0; i < ROWS_PER_THREAD; i++)
            if ((need >> i) & 1u) sh.u.c.plist[o++] = (uint8_t)((i << 6) | lane);
          wave_lds_sync();
          if (pf == FAST_PAINT_LINEAR)
            fast_gradient<FAST_PAINT_LINEAR>(B, sh, tot_p, lane, tx0, ty0);
          else if (pf == FAST_PAINT_RADIAL)
            fast_gradient<FAST_PAINT_RADIAL>(B, sh, tot_p, lane, tx0, ty0);
          else
            fast_gradient<FAST_PAINT_CONIC>(B, sh, tot_p, lane, tx0, ty0);
          wave_lds_sync();
        }
        const bool fast_full = (bflags & DRAW_FAST_SRC) != 0;
        const uint32_t full_raw = fast_full ? pack(p_px, S.fmt) : 0u;
        const bool rgba8 = S.fmt == Z2D_FORMAT_RGBA;  // uniform
        // the pixels composited on are stored (a composite that leaves a
        // pixel as it was stores the same value)
        stat_pix += (uint32_t)__popc(need);
        dirty |= need;
#pragma unroll
        for (int i = 0; i < ROWS_PER_THREAD; i++) {
          if (!((need >> i) & 1u)) continue;
          const int k = kk[i];
          if (fast_full && k >= 16) {
            pix[i] = full_raw;
          } else {
            uint32_t pv = solid ? B.srctab[k] : sh.u.c.ppaint[i][lane];
            if (rgba8) {
              if (!solid && k < 16) pv = mask_coverage_rgba8(pv, k);
              pix[i] = src_over_rgba8(pix[i], pv);
            } else {
              I4 src{pv & 0xFFu, (pv >> 8) & 0xFFu, (pv >> 16) & 0xFFu, pv >> 24};
              if (!solid && k < 16) src = mask_coverage(src, k);
              pix[i] = composite_src_over(pix[i], S.fmt, src);
            }
          }
        }
      }
    }

    if (S.clip) {
      // Clip stage: Surface.composite(dst, mask, .dst_in) after the draws.
#pragma unroll
      for (int i = 0; i < ROWS_PER_THREAD; i++) {
        const int32_t py = ty0 + i;
        if (!col_ok || py >= H) continue;
        const I4 m{0, 0, 0, S.clip[(int64_t)py * W + px]};
        const uint32_t before = pix[i];
        pix[i] = pack(op_int(Z2D_OP_DST_IN, unpack(pix[i], S.fmt), m), S.fmt);
        dirty |= (pix[i] != before ? 1u : 0u) << i;
      }
    }
#pragma unroll
    for (int i = 0; i < ROWS_PER_THREAD; i++) {
      const int32_t py = ty0 + i;
      if (col_ok && py < H && ((dirty >> i) & 1u)) store_raw(S, (int64_t)py * W + px, pix[i]);
    }
  }
}

// ---------------------------------------------------------------------------
// k_grad_setup: per-stop interpolation colors of every gradient of a job.
// ---------------------------------------------------------------------------
__global__ void k_grad_setup(DevGradient *grads, uint32_t n_grads) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= n_grads) return;
  const DevGradient &g = grads[gi];
  float4 *pre = const_cast<float4 *>(g.pre);
  for (uint32_t i = 0; i < g.n_stops; i++) pre[i] = stop_interp_color(g, g.stops[i]);
}

// ---------------------------------------------------------------------------
// k_composite: SurfaceCompositor.run batches.
// ---------------------------------------------------------------------------
__device__ __forceinline__ I4 comp_param_int(const DevCompParam &p, const CompJob &J, int64_t row_off,
                                             int32_t j, int x, int y) {
  switch (p.kind) {
    case Z2D_PARAM_PIXEL: return p.px;
    case Z2D_PARAM_SURFACE: {
      int64_t idx = p.base + row_off + j;
      if (idx < 0 || idx >= (int64_t)p.surf.w * p.surf.h) return I4{0, 0, 0, 0};
      return unpack(load_raw(p.surf, idx), p.surf.fmt);
    }
    default: return paint_int(*p.paint, paint_stops(*p.paint), J.blue_noise, x, y);
  }
}
__device__ __forceinline__ F4 comp_param_float(const DevCompParam &p, const CompJob &J,
                                               int64_t row_off, int32_t j, int x, int y) {
  switch (p.kind) {
    case Z2D_PARAM_PIXEL: return decode_raw(p.px);
    case Z2D_PARAM_SURFACE: {
      int64_t idx = p.base + row_off + j;
      if (idx < 0 || idx >= (int64_t)p.surf.w * p.surf.h) return F4{0, 0, 0, 0};
      return decode_raw(unpack(load_raw(p.surf, idx), p.surf.fmt));
    }
    default: return paint_float(*p.paint, paint_stops(*p.paint), J.blue_noise, x, y);
  }
}

__device__ __forceinline__ uint32_t comp_pixel(const CompJob &J, int32_t r, int32_t j,
                                               uint32_t dst_raw) {
  const int x = J.x0 + j, y = J.y0 + r;
  if (J.precision == Z2D_PRECISION_INTEGER) {
    I4 cur{0, 0, 0, 0};
    for (uint32_t k = 0; k < J.n_ops; k++) {
      const DevCompOp &o = J.ops[k];
      I4 s = o.src.kind == Z2D_PARAM_NONE
                 ? cur
                 : comp_param_int(o.src, J, (int64_t)r * o.src.surf.w, j, x, y);
      I4 d = o.dst.kind == Z2D_PARAM_NONE
                 ? unpack(dst_raw, J.dst.fmt)
                 : comp_param_int(o.dst, J, (int64_t)r * o.dst.surf.w, j, x, y);
      cur = op_int(o.op, d, s);
    }
    return pack(cur, J.out.fmt);
  }
  F4 cur{0, 0, 0, 0};
  for (uint32_t k = 0; k < J.n_ops; k++) {
    const DevCompOp &o = J.ops[k];
    F4 s = o.src.kind == Z2D_PARAM_NONE
               ? cur
               : comp_param_float(o.src, J, (int64_t)r * o.src.surf.w, j, x, y);
    F4 d = o.dst.kind == Z2D_PARAM_NONE
               ? decode_raw(unpack(dst_raw, J.dst.fmt))
               : comp_param_float(o.dst, J, (int64_t)r * o.dst.surf.w, j, x, y);
    cur = op_float(o.op, d, s);
  }
  return pack(encode_raw(cur), J.out.fmt);
}

// Generic path: one pixel per lane, rows on blockIdx.y (strided).
__global__ void k_composite(const CompJob *__restrict__ Jp) {
  const CompJob &J = *Jp;
  for (int32_t r = blockIdx.y; r < J.height; r += gridDim.y) {
    for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < J.width;
         j += gridDim.x * blockDim.x) {
      const int64_t didx = (int64_t)(J.y0 + r) * J.dst.w + (J.x0 + j);
      const uint32_t v = comp_pixel(J, r, j, load_raw(J.dst, didx));
      if (J.out.fmt >= Z2D_FORMAT_ALPHA4 && J.out.fmt <= Z2D_FORMAT_ALPHA1)
        store_raw(J.out, didx, v);
      else if (J.out.fmt == Z2D_FORMAT_ALPHA8)
        J.out.data[didx] = (uint8_t)v;
      else
        reinterpret_cast<uint32_t *>(J.out.data)[didx] = v;
    }
  }
}

// Streaming path for whole-surface 4-byte-format batches whose surface
// sources start at the same pixel as the destination: 4 pixels per lane via
// 16-byte loads/stores, grid-stride over the flat pixel range.
__global__ void __launch_bounds__(256) k_composite_stream(const CompJob *__restrict__ Jp,
                                                          uint64_t n_quads) {
  const CompJob &J = *Jp;
  const uint32_t *dst = reinterpret_cast<const uint32_t *>(J.dst.data);
  uint32_t *out = reinterpret_cast<uint32_t *>(J.out.data);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_quads; q += stride) {
    const uint64_t p0 = q * 4;
    const uint4 dv = *reinterpret_cast<const uint4 *>(dst + p0);
    uint32_t dvals[4] = {dv.x, dv.y, dv.z, dv.w};
    uint32_t res[4];
    const int32_t W = J.dst.w;
    const int32_t r = (int32_t)(p0 / (uint64_t)W);
    const int32_t j = (int32_t)(p0 - (uint64_t)r * W);
#pragma unroll
    for (int l = 0; l < 4; l++) {
      // the flat index maps to (r, j + l) with j + l < W because W % 4 == 0
      res[l] = comp_pixel(J, r, j + l, dvals[l]);
    }
    *reinterpret_cast<uint4 *>(out + p0) = make_uint4(res[0], res[1], res[2], res[3]);
  }
}

// ---------------------------------------------------------------------------
// k_composite_quad: the streaming compositor specialised per operator.  The
// operator is a template parameter (one instance per operator and precision,
// chosen once per launch), 4-byte formats unpack through byte shifts held in
// scalar registers, and every source is one 16-byte (or 4-byte for alpha8)
// load per four pixels.  Per pixel it computes what comp_pixel computes for
// the same batch.
// ---------------------------------------------------------------------------
struct Fmt32 {
  uint32_t sr, sb, has_a;  // red and blue byte shifts; alpha byte present
};
__device__ __forceinline__ Fmt32 fmt32(uint32_t fmt) {
  const bool bgr = fmt == Z2D_FORMAT_ARGB || fmt == Z2D_FORMAT_XRGB;
  return Fmt32{bgr ? 16u : 0u, bgr ? 0u : 16u,
               (fmt == Z2D_FORMAT_RGBA || fmt == Z2D_FORMAT_ARGB) ? 1u : 0u};
}
// unpack / pack of the four 4-byte formats
__device__ __forceinline__ I4 unpack32(uint32_t raw, Fmt32 f) {
  return I4{(raw >> f.sr) & 0xFFu, (raw >> 8) & 0xFFu, (raw >> f.sb) & 0xFFu,
            f.has_a ? raw >> 24 : 255u};
}
__device__ __forceinline__ uint32_t pack32(I4 v, Fmt32 f) {
  const uint32_t a = f.has_a ? (v.a & 0xFFu) << 24 : 0u;
  return ((v.r & 0xFFu) << f.sr) | ((v.g & 0xFFu) << 8) | ((v.b & 0xFFu) << f.sb) | a;
}
__device__ __forceinline__ void quad_load(const QuadSrc &s, Fmt32 f, uint64_t q, I4 (&o)[4]) {
  if (s.kind == QUAD_SURF32) {
    const uint4 v = reinterpret_cast<const uint4 *>(s.data)[q];
    o[0] = unpack32(v.x, f);
    o[1] = unpack32(v.y, f);
    o[2] = unpack32(v.z, f);
    o[3] = unpack32(v.w, f);
  } else if (s.kind == QUAD_ALPHA8) {
    const uint32_t v = reinterpret_cast<const uint32_t *>(s.data)[q];
#pragma unroll
    for (int l = 0; l < 4; l++) o[l] = I4{0u, 0u, 0u, (v >> (8 * l)) & 0xFFu};
  } else {
#pragma unroll
    for (int l = 0; l < 4; l++) o[l] = s.px;
  }
}

template <uint32_t OP, bool FLOAT>
__device__ __attribute__((always_inline, flatten)) void quad_run(const QuadJob &J,
                                                                  uint64_t n_quads) {
  const Fmt32 fd = fmt32(J.dst_fmt), fo = fmt32(J.out_fmt), fs = fmt32(J.src.fmt),
              fm = fmt32(J.mask.fmt);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_quads; q += stride) {
    const uint4 dv = reinterpret_cast<const uint4 *>(J.dst)[q];
    const uint32_t draw[4] = {dv.x, dv.y, dv.z, dv.w};
    I4 s[4];
    quad_load(J.src, fs, q, s);
    I4 m[4];
    if (J.masked) quad_load(J.mask, fm, q, m);
    uint32_t res[4];
#pragma unroll
    for (int l = 0; l < 4; l++) {
      const I4 d = unpack32(draw[l], fd);
      if (FLOAT) {
        F4 sv = decode_raw(s[l]);
        if (J.masked) sv = op_float(Z2D_OP_DST_IN, sv, decode_raw(m[l]));
        res[l] = pack32(encode_raw(op_float(OP, decode_raw(d), sv)), fo);
      } else {
        I4 sv = s[l];
        if (J.masked) sv = op_int(Z2D_OP_DST_IN, sv, m[l]);
        res[l] = pack32(op_int(OP, d, sv), fo);
      }
    }
    reinterpret_cast<uint4 *>(J.out)[q] = make_uint4(res[0], res[1], res[2], res[3]);
  }
}

__global__ void __launch_bounds__(256) k_composite_quad(QuadJob J, uint64_t n_quads) {
  const bool flt = J.precision != Z2D_PRECISION_INTEGER;
  switch (J.op) {
#define Z2D_QUAD_CASE(n)            \
  case n:                          \
    if (flt)                       \
      quad_run<n, true>(J, n_quads); \
    else                           \
      quad_run<n, false>(J, n_quads); \
    break;
    Z2D_QUAD_CASE(0) Z2D_QUAD_CASE(1) Z2D_QUAD_CASE(2) Z2D_QUAD_CASE(3) Z2D_QUAD_CASE(4)
    Z2D_QUAD_CASE(5) Z2D_QUAD_CASE(6) Z2D_QUAD_CASE(7) Z2D_QUAD_CASE(8) Z2D_QUAD_CASE(9)
    Z2D_QUAD_CASE(10) Z2D_QUAD_CASE(11) Z2D_QUAD_CASE(12) Z2D_QUAD_CASE(13) Z2D_QUAD_CASE(14)
    Z2D_QUAD_CASE(15) Z2D_QUAD_CASE(16) Z2D_QUAD_CASE(17) Z2D_QUAD_CASE(18) Z2D_QUAD_CASE(19)
    Z2D_QUAD_CASE(20) Z2D_QUAD_CASE(21) Z2D_QUAD_CASE(22) Z2D_QUAD_CASE(23) Z2D_QUAD_CASE(24)
    Z2D_QUAD_CASE(25) Z2D_QUAD_CASE(26) Z2D_QUAD_CASE(27)
#undef Z2D_QUAD_CASE
    default: break;
  }
}

// ---------------------------------------------------------------------------
// launchers (called from host.cpp)
// ---------------------------------------------------------------------------
}  // namespace z2d

using namespace z2d;

// Reference streams for the achievable HBM rate: a copy (1 read : 1 write)
// and, the compositor's mix, two reads and one write (dst = a ^ b).  The
// same shape as the whole-surface compositor's loop (16 B per buffer per
// lane per iteration, at most 8192 blocks of 256 threads striding over the
// buffers), so the two are compared on equal terms; also the known-size
// kernel that calibrates the HBM byte counters.
template <bool READ2>
__global__ void __launch_bounds__(256) k_stream(const uint4 *__restrict__ a,
                                                const uint4 *__restrict__ b,
                                                uint4 *__restrict__ dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint4 v = a[i];
    if (READ2) {
      const uint4 w = b[i];
      v.x ^= w.x; v.y ^= w.y; v.z ^= w.z; v.w ^= w.w;
    }
    dst[i] = v;
  }
}

static uint32_t stream_blocks(uint64_t n) {
  const uint64_t blocks = (n + 255) / 256;
  return (uint32_t)(blocks > 8192 ? 8192 : blocks);
}

hipError_t z2d_launch_stream_copy(const void *src, void *dst, uint64_t bytes, hipStream_t st) {
  const uint64_t n = bytes / 16;
  hipLaunchKernelGGL(k_stream<false>, dim3(stream_blocks(n)), dim3(256), 0, st,
                     reinterpret_cast<const uint4 *>(src), nullptr, reinterpret_cast<uint4 *>(dst),
                     n);
  return hipGetLastError();
}

hipError_t z2d_launch_stream_read2(const void *a, const void *b, void *dst, uint64_t bytes,
                                   hipStream_t st) {
  const uint64_t n = bytes / 16;
  hipLaunchKernelGGL(k_stream<true>, dim3(stream_blocks(n)), dim3(256), 0, st,
                     reinterpret_cast<const uint4 *>(a), reinterpret_cast<const uint4 *>(b),
                     reinterpret_cast<uint4 *>(dst), n);
  return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Command y of the list, strided over the x blocks: 16-byte moves where the
// addresses and the size allow, bytes otherwise; fills write whole words.
__global__ void __launch_bounds__(256) k_multi_copy(const CopyCmd *__restrict__ cmds) {
  const CopyCmd c = cmds[blockIdx.y];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (!c.src) {
    uint32_t *d = reinterpret_cast<uint32_t *>(c.dst);
    const uint64_t n = c.bytes / 4;
    for (uint64_t i = t0; i < n; i += stride) d[i] = c.value;
    return;
  }
  const bool v16 = ((reinterpret_cast<uintptr_t>(c.src) | reinterpret_cast<uintptr_t>(c.dst)) &
                    15u) == 0;
  uint64_t done = 0;
  if (v16) {
    const uint64_t n = c.bytes / 16;
    const uint4 *s4 = reinterpret_cast<const uint4 *>(c.src);
    uint4 *d4 = reinterpret_cast<uint4 *>(c.dst);
    for (uint64_t i = t0; i < n; i += stride) d4[i] = s4[i];
    done = n * 16;
  }
  for (uint64_t i = done + t0; i < c.bytes; i += stride) c.dst[i] = c.src[i];
}

hipError_t z2d_launch_multi_copy(const CopyCmd *cmds, uint32_t n, hipStream_t st) {
  for (uint32_t c0 = 0; c0 < n; c0 += 65535u) {  // grid y is at most 65535
    hipLaunchKernelGGL(k_multi_copy, dim3(64, min(65535u, n - c0)), dim3(256), 0, st, cmds + c0);
  }
  return hipGetLastError();
}

hipError_t z2d_launch_grad_setup(DevGradient *grads, uint32_t n_grads, hipStream_t st) {
  if (n_grads == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grad_setup, dim3((n_grads + 63) / 64), dim3(64), 0, st, grads, n_grads);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_fill_set(FillSet F) {
  const uint32_t f = blockIdx.y;
  uint32_t *const p = F.p[f];
  const uint64_t n = F.n[f];
  const uint32_t v = F.v[f];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t z2d_launch_fill_set(const FillSet &F, hipStream_t st) {
  if (F.count == 0) return hipSuccess;
  uint64_t most = 0;
  for (uint32_t f = 0; f < F.count; f++) most = F.n[f] > most ? F.n[f] : most;
  const uint64_t blocks = (most + 255) / 256;
  hipLaunchKernelGGL(k_fill_set, dim3((uint32_t)(blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks)), F.count),
                     dim3(256), 0, st, F);
  return hipGetLastError();
}

hipError_t z2d_launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t z2d_launch_edge_build(const JobView &J, hipStream_t st) {
  if (J.n_draws == 0) return hipSuccess;
  dim3 grid((J.n_draws + 3) / 4);  // one wave per draw
  hipLaunchKernelGGL(k_edge_build, grid, dim3(256), 0, st, J);
  return hipGetLastError();
}

hipError_t z2d_launch_rowprog(const JobView &J, const uint32_t *list, uint32_t n,
                              z2d_edge *sedges, int32_t *sx, const uint32_t *soff,
                              hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rowprog, dim3((n + 63) / 64), dim3(64), 0, st, J, list, n, sedges, sx,
                     soff);
  return hipGetLastError();
}

hipError_t z2d_launch_binning(const JobView &J, hipStream_t st) {
  if (J.n_bands == 0) return hipSuccess;
  // entry edge lists: counts, offsets, then the copies (entry_ecnt,
  // entry_cur and the entry column ranges were reset by the pipeline's
  // first fill, z2d_launch_fill_set)
  dim3 ggrid((J.n_egroups + 3) / 4);
  if (J.n_egroups) hipLaunchKernelGGL(k_edge_bins<false>, ggrid, dim3(256), 0, st, J);
  scan_u32(J.entry_ecnt, J.max_entries, J.scan_tmp, J.counters + 2, st);
  if (J.n_egroups) hipLaunchKernelGGL(k_edge_bins<true>, ggrid, dim3(256), 0, st, J);
  // band lists and entry records
  dim3 grid((J.n_slots + 3) / 4);
  hipLaunchKernelGGL(k_band_pass<false>, grid, dim3(256), 0, st, J);
  scan_u32(J.slot_cnt, J.n_slots, J.scan_tmp + J.scan_words, J.counters + 2, st);
  hipLaunchKernelGGL(k_band_pass<true>, grid, dim3(256), 0, st, J);
  if (J.n_tiles) {
    dim3 tgrid((J.n_tiles + 3) / 4);
    hipLaunchKernelGGL(k_strip_pass<false>, tgrid, dim3(256), 0, st, J);
    scan_u32(J.strip_cnt, J.n_tiles * J.strips_per_tile, J.scan_tmp + 2 * J.scan_words,
             J.counters + 2, st);
    hipLaunchKernelGGL(k_strip_pass<true>, tgrid, dim3(256), 0, st, J);
    if (J.strip_order) {
      const uint32_t ns = J.n_tiles * J.strips_per_tile;
      const dim3 og((ns + 255) / 256);
      hipLaunchKernelGGL(k_fill_u32, dim3((ORDER_BUCKETS + 1 + 255) / 256), dim3(256), 0, st,
                         J.order_hist, (uint64_t)ORDER_BUCKETS + 1, 0u);
      hipLaunchKernelGGL(k_order_hist, og, dim3(256), 0, st, J, ns);
      hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, J.order_hist, ORDER_BUCKETS);
      hipLaunchKernelGGL(k_order_scatter, og, dim3(256), 0, st, J, ns);
    }
  }
  return hipGetLastError();
}

// Count pass (emit = false): the curve pieces, their combination per curve,
// the per-node pass for every draw, then the per-draw walk for the draws it
// marked FLAT_NONSIMPLE.  Emit pass: each draw by the form its
// FlatJob.serial names (the per-draw walk launched only when any_serial),
// curves by pieces, and the fills of *fill (when given) done by the per-node
// kernel's workgroups before their draws.
hipError_t z2d_launch_flatten(const z2d_path_node *raw, const FlatJob *jobs, uint32_t n_jobs,
                              FNode *out, FlatResult *res, uint32_t *rel, const uint2 *curves,
                              uint32_t n_curves, FlatTask *tasks, uint32_t *toff, FlatCurve *cv,
                              bool emit, hipStream_t st, const FillSet *fill,
                              bool any_serial) {
  if (n_jobs == 0) return hipSuccess;
  const FillSet no_fill{};
  const dim3 wgrid((n_jobs + 3) / 4), sgrid((n_jobs + 63) / 64);
  const dim3 tgrid((uint32_t)(((uint64_t)n_curves * FLAT_TASKS + 255) / 256)),
      cgrid((n_curves + 255) / 256);
  if (emit) {
    hipLaunchKernelGGL(k_flatten_wave<true>, wgrid, dim3(256), 0, st, raw, jobs, n_jobs, out, res,
                       rel, cv, fill ? *fill : no_fill);
    if (n_curves)
      hipLaunchKernelGGL(k_curve_tasks<true>, tgrid, dim3(256), 0, st, raw, jobs, curves, n_curves,
                         tasks, toff, rel, out);
    if (any_serial)
      hipLaunchKernelGGL(k_flatten_serial<true>, sgrid, dim3(64), 0, st, raw, jobs, n_jobs, out,
                         res);
  } else {
    if (n_curves) {
      hipLaunchKernelGGL(k_curve_tasks<false>, tgrid, dim3(256), 0, st, raw, jobs, curves,
                         n_curves, tasks, toff, rel, out);
      hipLaunchKernelGGL(k_curve_combine, cgrid, dim3(256), 0, st, raw, jobs, curves, n_curves,
                         tasks, toff, cv);
    }
    hipLaunchKernelGGL(k_flatten_wave<false>, wgrid, dim3(256), 0, st, raw, jobs, n_jobs, out, res,
                       rel, cv, no_fill);
    hipLaunchKernelGGL(k_flatten_serial<false>, sgrid, dim3(64), 0, st, raw, jobs, n_jobs, out,
                       res);
  }
  return hipGetLastError();
}

// The tile render of tiles [tile_base, tile_base + tile_count): k_render_fast
// (one-wave workgroups) over the strips of fast draws, then k_render
// (four-wave workgroups) over the others; each takes strips from its own
// cursor and leaves the other's.
hipError_t z2d_launch_render(const JobView &J, uint32_t tile_base, uint32_t tile_count,
                             uint32_t *cursor, uint32_t *cursor_fast, uint32_t n_blocks,
                             uint32_t n_fast_blocks, hipStream_t st) {
  if (tile_count == 0) return hipSuccess;
  const uint32_t ordered = (J.strip_order && tile_base == 0 && tile_count == J.n_tiles) ? 1u : 0u;
  hipLaunchKernelGGL(k_render_fast, dim3(n_fast_blocks), dim3(64), 0, st, J, tile_base,
                     tile_count, cursor_fast, ordered);
  hipLaunchKernelGGL(k_render, dim3(n_blocks), dim3(RENDER_THREADS), 0, st, J, tile_base,
                     tile_count, cursor, ordered);
  return hipGetLastError();
}

hipError_t z2d_launch_composite_quad(const QuadJob &job, uint64_t n_quads, hipStream_t st) {
  if (n_quads == 0) return hipSuccess;
  uint64_t blocks = (n_quads + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_composite_quad, dim3((uint32_t)blocks), dim3(256), 0, st, job, n_quads);
  return hipGetLastError();
}

hipError_t z2d_launch_composite(const CompJob *dev_job, const CompJob &host_job, bool stream,
                                hipStream_t st) {
  if (host_job.width <= 0 || host_job.height <= 0) return hipSuccess;
  if (stream) {
    uint64_t quads = (uint64_t)host_job.width * (uint64_t)host_job.height / 4;
    uint64_t blocks = (quads + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_composite_stream, dim3((uint32_t)blocks), dim3(256), 0, st, dev_job,
                       quads);
  } else {
    uint32_t bx = (uint32_t)((host_job.width + 255) / 256);
    uint32_t by = (uint32_t)(host_job.height < 65535 ? host_job.height : 65535);
    hipLaunchKernelGGL(k_composite, dim3(bx, by), dim3(256), 0, st, dev_job);
  }
  return hipGetLastError();
}
