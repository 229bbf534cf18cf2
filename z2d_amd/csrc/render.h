// render.h - device job layout shared by the host runtime (host.cpp) and the
// kernels (render.hip).  A job is a set of surfaces and an ordered list of
// draws (fills / stroke polygons) on them; one pipeline launch renders all of
// them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_pixel.h"

namespace z2d {

// Tile shape of the renderer: TILE_W pixels x band_h rows per band entry
// (JobView::band_h, TILE_H unless tuned; a multiple of ROWS_PER_THREAD up to
// MAX_BAND_H); workgroups of 256 lanes (4 wave64s) hold four independent
// strip workers.
constexpr int TILE_W = 64;
constexpr int TILE_H = 16;
constexpr int MAX_BAND_H = 32;
constexpr int RENDER_THREADS = 256;
// Each wave renders one strip of ROWS_PER_THREAD rows of a tile on its own
// (no workgroup barriers): lane = column, one register per strip row.
constexpr int ROWS_PER_THREAD = 4;
constexpr int MAX_STRIPS_PER_TILE = MAX_BAND_H / ROWS_PER_THREAD;
constexpr int STRIP_SUBROWS = ROWS_PER_THREAD * 4;
constexpr int ACC_WORDS = TILE_W * 2;  // 4 sample columns per pixel, two int16 per word
// Winding deltas are accumulated as int16 pairs in LDS; an entry whose band
// holds more edges than this cannot overflow them and takes the exact
// per-thread path instead.
constexpr uint32_t PACKED_EDGE_LIMIT = 32767;
// Length classes of the longest-first strip order (lists longer than this
// share the first class).
constexpr uint32_t ORDER_BUCKETS = 1024;
// Edge slots of one draw per k_edge_bins work item.
constexpr uint32_t EDGE_GROUP = 64;

// Flattened path node (move_to / line_to / close_path only; curves are
// flattened on the host by the reference's Spline algorithm first).
struct FNode {
  double x, y;
  uint32_t tag, pad;
};

// Bézier flattening on the device (k_flatten): one job per fill draw.
struct FlatJob {
  uint32_t raw_off, raw_cnt;  // the draw's path nodes
  uint32_t node_off;          // where its flattened nodes go (emit pass)
  uint32_t nbands, band_h;    // for the (edge, band) bound (count pass)
  uint32_t serial;            // flattened by the per-draw walk (see FLAT_NONSIMPLE)
  double tol_sq;              // squared tolerance (>= 0.001^2)
  uint32_t curve_base;        // index of the draw's first curve in the curve list
  uint32_t pad;
};
// Curves are subdivided as FLAT_TASKS independent pieces: the subtrees of the
// Spline recursion at depth FLAT_TASK_DEPTH, in order (a leaf above that
// depth belongs to the leftmost piece under it).
constexpr uint32_t FLAT_TASK_DEPTH = 4;
constexpr uint32_t FLAT_TASKS = 1u << FLAT_TASK_DEPTH;

// Values per workgroup of the binning's exclusive scans (256 lanes x 16).
constexpr uint32_t SCAN_BLOCK = 4096;
struct FlatTask {  // count pass, per curve piece
  uint32_t cnt, flags;
  uint64_t pairs;   // (edge, band) bound between the piece's own points
  double fy, ly;    // its first and last point's y
  double box[4];
};
struct FlatCurve {  // count pass, per curve
  uint32_t cnt, flags;
  uint64_t pairs;
  double box[4];
};
enum : uint32_t {
  FLAT_UNBALANCED = 1u << 0,  // a contour does not end where it starts
  FLAT_SERIAL = 1u << 1,      // a line follows a close
  FLAT_RUNAWAY = 1u << 2,     // a curve exceeded the subdivision visit limit
  // the draw's point-buffer state is not a function of each node and its
  // predecessor (a close followed by anything but a move, a close before
  // the first move, or a contour with fewer than 3 distinct points before
  // its close): it is flattened by the per-draw walk instead of per node
  FLAT_NONSIMPLE = 1u << 3,
};
struct FlatResult {
  uint32_t n_nodes, flags;
  uint64_t edge_pairs;  // upper bound of (edge, band) pairs of the flattened draw
  double box[4];        // xmin, xmax, ymin, ymax of the tracked points
};

struct DevSurface {
  uint8_t *data;
  int32_t w, h;
  uint32_t fmt, pad;
  uint64_t nbytes;
  const uint8_t *clip;  // optional alpha8 clip mask (w*h): dst_in after all draws
};

enum DrawFlags : uint32_t {
  DRAW_VISIBLE = 1u << 0,     // passed Polygon.inBox and has a non-empty region
  DRAW_UNBOUNDED = 1u << 1,   // operator is unbounded (src_in/dst_in/src_out/dst_atop)
  DRAW_FAST_SRC = 1u << 2,    // solid pattern and op reduces to source
  DRAW_UNBALANCED = 1u << 3,  // some contour does not return to its start
  DRAW_ROWPROG = 1u << 4,     // needs the serial row emulation
  DRAW_HAS_EDGES = 1u << 5,
  DRAW_SERIAL_PLOT = 1u << 6,  // a line follows a close: plot node by node
  DRAW_FAST = 1u << 7,         // rendered by k_render_fast (see FastDraw)
};

enum DrawMode : uint32_t { MODE_DIRECT = 0, MODE_MSAA = 1, MODE_SSAA = 2 };

struct DevDraw {
  // inputs
  uint32_t node_off, node_cnt;  // flattened nodes; edge slots use the same range
  uint32_t surface;
  uint32_t mode;       // DrawMode (alpha1 targets already forced to direct)
  uint32_t rule, op;
  uint32_t precision;  // already upgraded to float if op requires it
  uint32_t paint;
  uint32_t input_edges;  // nonzero: edges given directly (stroke polygons)
  uint32_t in_edge_off;  // their offset in JobView::in_edges
  double in_ext[4];      // extents when edges are given directly
  // the host's band estimate [band_lo, band_hi] (bands of the surface) and
  // the draw's entries: entry_base + (band - band_lo)
  uint32_t band_lo, band_hi, entry_base;
  // written by k_edge_build
  uint32_t edge_cnt;
  uint32_t flags;
  double ext[4];  // top, bottom, left, right (scaled space)
  int32_t rx0, ry0, rx1, ry1;  // composite region [rx0,rx1) x [ry0,ry1)
  int32_t sx, ex, sy, ey;      // MSAA window (rows sy..ey inclusive)
  int32_t mask_bits;           // SSAA mask depth
  int32_t row_y0, row_n;       // sub-rows covered by the row records
  uint32_t row_off;            // offset into the row-record array
};

// Per sub-row record of the serial emulation (rowprog): unbalanced cutoff
// and, for direct-mode unbounded draws, the surviving span.
struct RowRec {
  int32_t cutoff;  // crossings >= cutoff are in a dropped trailing span
  int32_t kind;    // direct unbounded: 0 clear row, 1 untouched, 2 span on dst, 3 span on zero
  int32_t s, e;    // surviving span [s, e)
};

// An edge as the tile renderer reads it from a band entry's list: the
// reference's intercept terms (top, x_start, x_inc, Polygon.Edge,
// src/internal/tess/Polygon.zig:18-47) and the sample rows y it is active on,
// top < y + 0.5 <= bottom (the active-edge test, Polygon.zig:284-285), as
// [ylo, yhi] computed once by the binning pass.  yhi2 = yhi * 2 + (y0 < y1):
// the low bit is the edge's orientation (its winding direction).
struct REdge {
  double top, x_start, x_inc;
  int32_t ylo, yhi2;
};

struct BandEntry {
  uint32_t draw;
  uint32_t edge_off;  // into the entry-edge array
  uint32_t edge_cnt;  // edges of the draw whose y-range meets the band
  uint32_t band;
  // pixel columns of its edges' crossings: a tile wholly beside them has no
  // coverage ([INT32_MIN, INT32_MAX] for draws where that may still change
  // pixels)
  int32_t px_lo, px_hi;
};

// Per-draw constants of k_render_fast, the renderer of the common draw
// (multisample, src_over in integer precision, a solid paint or a linear /
// radial / conic gradient interpolated in linear RGB with at most FAST_STOPS
// stops): one fixed-size block per draw, written by k_edge_build, fetched
// by the strip worker one entry ahead of its use.
constexpr uint32_t FAST_STOPS = 8;
enum : uint32_t {
  FAST_PAINT_NONE = 0,
  FAST_PAINT_SOLID = 1,
  FAST_PAINT_LINEAR = 2,  // + gradient type
  FAST_PAINT_RADIAL = 3,
  FAST_PAINT_CONIC = 4,
};
struct alignas(16) FastDraw {
  int32_t rx0, rx1, ry0, ry1;  // composite region
  uint32_t flags, rule, paint, px;  // DRAW_* flags, fill rule, FAST_PAINT_*, solid RGBA8
  uint32_t n_stops, identity, pad0, pad1;
  uint32_t pad2[4];
  // gradient: geom[0..1], the inverse matrix, the type's constants
  // (linear: dx, dy, 1/|d|^2, degenerate; radial: r1, r2, cdx, cdy, dr,
  // min_dr, a, 1/a; conic: the start angle mod 2 pi)
  double gx, gy;
  double m[6];
  double p[8];
  float offs[FAST_STOPS];   // stop offsets
  float4 pre[FAST_STOPS];   // stop colors in the interpolation space
  // solid paint: the RGBA8 source masked by each coverage count 0..16
  // (mask_coverage; entry 16 is the paint itself)
  uint32_t srctab[20];
};
constexpr uint32_t FAST_DRAW_CHUNKS = sizeof(FastDraw) / 16;  // 16-byte lanes
static_assert(sizeof(FastDraw) % 16 == 0 && FAST_DRAW_CHUNKS <= 64, "FastDraw layout");

struct TileRef {
  uint32_t surface;
  uint16_t tx, ty;
};

struct JobView {
  const FNode *nodes;
  const z2d_edge *in_edges;  // edges of polygon draws
  z2d_edge *edges;  // edge slots, indexed like nodes
  DevDraw *draws;
  int2 *draw_rows;  // per draw: composite rows [y0, y1), empty when not visible
  uint32_t n_draws;
  DevSurface *surfaces;
  uint32_t n_surfaces;
  const DevPaint *paints;
  const uint16_t *blue_noise;
  // surface -> draw range, band range
  const uint32_t *surf_draw_off;  // n_surfaces + 1
  const uint32_t *surf_band_off;  // n_surfaces + 1
  uint32_t n_bands;
  // binning outputs
  uint32_t *band_cnt;    // n_bands + 1: offsets of the band lists in band_list
  // per band-list position: x = entry index; the tiles of the band whose
  // column tx0 satisfies y < tx0 < z (int32) take it; w = the strips it
  // meets (bits 0..7) | needs the general renderer (bit 8).  Capacity
  // max_entries.
  uint4 *band_list;
  // band pass slots: a band and 64 consecutive draws of its surface
  const uint2 *slot_info;         // per slot: band, surface
  const uint32_t *band_slot_off;  // n_bands: the band's first slot
  uint32_t n_slots;
  uint32_t *slot_cnt;             // n_slots + 1 (scanned in place)
  // the scans' look-back words: three regions of scan_words u64 (a ticket,
  // then one status word per workgroup), zero at the start of each render
  uint64_t *scan_tmp;
  uint32_t scan_words;
  BandEntry *entries;    // per draw and estimated band, capacity max_entries
  uint32_t max_entries;
  uint32_t *entry_ecnt;      // max_entries + 1: per-entry edge counts, scanned in place
  uint32_t *entry_cur;       // max_entries: fill cursors (follows entry_ecnt)
  int32_t *entry_xlo;        // max_entries: sample-column range of the entry's edges
  int32_t *entry_xhi;
  const uint2 *egroups;      // edge-binning work: draw, first edge slot
  uint32_t n_egroups;
  REdge *entry_edges;        // the edges of each entry's list, capacity max_entry_edges
  uint32_t max_entry_edges;
  uint32_t band_h;           // rows per band
  uint32_t strips_per_tile;  // band_h / ROWS_PER_THREAD
  // longest-first queue order of the single-launch render (null: natural
  // order): strips by decreasing list length, from a counting sort over
  // ORDER_BUCKETS lengths
  uint32_t *strip_order;     // n_tiles * strips_per_tile
  uint32_t *order_hist;      // ORDER_BUCKETS + 1
  // per strip (tile * strips_per_tile + strip): the band entries whose
  // composite region meets the strip, in draw order
  uint32_t *strip_cnt;       // n_tiles * strips_per_tile + 1 (scanned in place)
  uint32_t *strip_entries;   // band entry indices, capacity max_strip_entries
  uint32_t max_strip_entries;
  // per strip: nonzero when an entry of its list is not a DRAW_FAST draw
  // with a packed edge list (the general k_render takes it, k_render_fast
  // the others)
  uint32_t *strip_slow;
  FastDraw *fast;            // per draw: k_render_fast's constants
  // [2] binning overflow flag; [4..5] pixel-draws (u64); [6] edges;
  // [7] tile-draw entries rasterized; [8] entries on the unpacked path;
  // [9] strips for the general renderer
  uint32_t *counters;
  uint32_t count_stats;  // nonzero: accumulate pixel-draw / edge counts
  uint32_t *cursors;     // tile work cursors, one per render launch
  RowRec *rows;
  uint32_t max_rows;
  const TileRef *tiles;
  uint32_t n_tiles;
};

// SurfaceCompositor.run batch on device (src/compositor.zig:302-440).
struct DevCompParam {
  uint32_t kind;  // Z2D_PARAM_*
  uint32_t pad;
  I4 px;
  const DevPaint *paint;  // gradient / dither sources
  DevSurface surf;        // surface sources
  int64_t base;           // pixel index of the stride start in row 0
};
struct DevCompOp {
  uint32_t op, pad;
  DevCompParam dst, src;
};
struct CompJob {
  DevSurface dst;
  DevSurface out;  // equal to dst for in-place runs
  int32_t x0, y0;  // first destination pixel of the region
  int32_t width, height;
  int32_t src_row0;  // source row of region row 0 (src_start_y)
  uint32_t n_ops;
  uint32_t precision;
  DevCompOp ops[16];
  const uint16_t *blue_noise;
};

// One source of the quad compositor: a 4-byte-format surface, an alpha8
// surface or a constant pixel, read at the destination's own pixel index.
struct QuadSrc {
  const uint8_t *data;  // surface kinds
  uint32_t kind;        // QUAD_PIXEL / QUAD_SURF32 / QUAD_ALPHA8
  uint32_t fmt;         // surface format (QUAD_SURF32)
  I4 px;                // QUAD_PIXEL
};
enum : uint32_t { QUAD_PIXEL = 0, QUAD_SURF32 = 1, QUAD_ALPHA8 = 2 };

// Whole-surface SurfaceCompositor batches of one operator, or of
// {dst_in dst=src src=mask}, {op}: out = op(dst, src) per pixel, four pixels
// per lane.
struct QuadJob {
  const uint32_t *dst;
  uint32_t *out;
  uint32_t dst_fmt, out_fmt;
  uint32_t op;       // the operator applied to the destination
  uint32_t masked;   // nonzero: src = dst_in(src, mask) first
  uint32_t precision;
  QuadSrc src, mask;
};

// Up to six word fills in one launch (the zeroing at the start of a render
// pipeline): fill f writes v[f] to n[f] words at p[f].
struct FillSet {
  uint32_t *p[6];
  uint64_t n[6];
  uint32_t v[6];
  uint32_t count;
};

// One command of k_multi_copy: copy `bytes` from src to dst, or (src null)
// fill bytes / 4 words at dst with `value`.  Many such commands go in one
// launch (a server batch's surface clears, clip-mask scatters and result
// gathers).
struct CopyCmd {
  const uint8_t *src;
  uint8_t *dst;
  uint64_t bytes;
  uint32_t value, pad;
};

}  // namespace z2d
