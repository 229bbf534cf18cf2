/*
 * z2d_gpu.h - C ABI of the MI355X backend for z2d's painter hot path.
 *
 * The reference's boundary is painter.fill / painter.stroke
 * (src/painter.zig:66-143, :214-344) plus the compositor entry points
 * SurfaceCompositor.run (src/compositor.zig:302-440) and Surface.composite
 * (src/surface.zig:225-241).  A Zig shim keeping those signatures marshals
 * its tagged unions into the POD descriptors of z2d_types.h and calls the
 * functions below (z2d_amd/zig/painter_gpu.zig is that binding).
 *
 * All functions return a z2d_status code (0 = ok).  Argument checks run on
 * the host before any device work, in the reference's order.
 */
#ifndef Z2D_GPU_H
#define Z2D_GPU_H

#include "z2d_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct z2d_gpu_ctx z2d_gpu_ctx;
typedef struct z2d_gpu_surface z2d_gpu_surface;
typedef struct z2d_gpu_batch z2d_gpu_batch;

/* A z2d_surface_desc whose `reserved` field holds this value refers to
 * device memory (from z2d_gpu_surface_desc) instead of a host buffer. */
#define Z2D_SURFACE_DEVICE 1u

/* ---- context -------------------------------------------------------------
 * One context per host thread; it owns a HIP stream and a device arena.
 * No reference counterpart (the reference allocates per call through its
 * `alloc` parameter, src/painter.zig:67). */
int32_t z2d_gpu_create(int32_t device, z2d_gpu_ctx **out);
void z2d_gpu_destroy(z2d_gpu_ctx *ctx);
const char *z2d_gpu_last_error(z2d_gpu_ctx *ctx);

/* ---- painter ---------------------------------------------------------------
 * painter.fill (src/painter.zig:66-143).  Synchronous: on return the caller's
 * surface buffer holds the result (host buffers are uploaded and read back;
 * device descriptors are rendered in place).  Errors:
 * PIXEL_SOURCE_NOT_PREMULTIPLIED, PATH_NOT_CLOSED, INVALID_STATE; zero nodes
 * is a no-op, as in the reference. */
int32_t z2d_gpu_fill(z2d_gpu_ctx *ctx, z2d_surface_desc *surface,
                     const z2d_pattern_desc *pattern, const z2d_path_node *nodes,
                     uint64_t n_nodes, const z2d_fill_opts *opts);

/* painter.stroke (src/painter.zig:214-344): stroke tessellation (the
 * reference's stroke_plotter / dashed_plotter) runs on the host and feeds the
 * same rasterizer with the non-zero rule.  Errors in the reference's order:
 * PIXEL_SOURCE_NOT_PREMULTIPLIED, INVALID_MATRIX (the transformation is not
 * invertible); zero nodes is a no-op; INVALID_STATE.  Hairline strokes
 * return UNSUPPORTED. */
int32_t z2d_gpu_stroke(z2d_gpu_ctx *ctx, z2d_surface_desc *surface,
                       const z2d_pattern_desc *pattern, const z2d_path_node *nodes,
                       uint64_t n_nodes, const z2d_stroke_opts *opts);

/* stroke_plotter.plot (src/internal/tess/stroke_plotter.zig:40-76) with
 * painter.stroke's option adjustments (src/painter.zig:236-304) for a
 * surface of `surface_format`: the stroke's polygon in the AA mode's scaled
 * space (edges, extents top/bottom/left/right, and the effective AA mode).
 * Host only, no device work.  Writes at most `capacity` edges; *n_edges is
 * the full count. */
int32_t z2d_stroke_polygon(const z2d_path_node *nodes, uint64_t n_nodes,
                           const z2d_stroke_opts *opts, uint32_t surface_format,
                           z2d_edge *edges, uint64_t capacity, uint64_t *n_edges,
                           double extents[4], uint32_t *aa_mode);

/* The rasterizer step of painter.fill/stroke on an existing edge list
 * (direct.run / supersample.run / multisample.run,
 * src/internal/raster/{direct,supersample,multisample}.zig), with extents as
 * Polygon.extent_{top,bottom,left,right}; edges are in the scaled space of
 * the AA mode (x1 for none, x4 otherwise). */
int32_t z2d_gpu_fill_polygon(z2d_gpu_ctx *ctx, z2d_surface_desc *surface,
                             const z2d_pattern_desc *pattern, const z2d_edge *edges,
                             uint64_t n_edges, const double extents[4],
                             uint32_t anti_aliasing_mode, uint32_t fill_rule,
                             uint32_t op, uint32_t precision);

/* ---- compositor ----------------------------------------------------------
 * SurfaceCompositor.run (src/compositor.zig:302-440): `n_ops` operations
 * applied at (dst_x, dst_y).  Surface parameters may be host or device
 * descriptors.  If `out` is non-NULL the result goes to `out` (same format
 * and size as dst) and dst is left untouched; NULL = in place, as the
 * reference. */
int32_t z2d_gpu_composite(z2d_gpu_ctx *ctx, z2d_surface_desc *dst, int32_t dst_x,
                          int32_t dst_y, const z2d_comp_op *ops, uint64_t n_ops,
                          uint32_t precision, z2d_surface_desc *out);

/* z2d_gpu_composite with no allocation: no counterpart in the reference.  For
 * callers that cannot allocate device or pinned memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks and results as z2d_gpu_composite, but every
 * surface (dst, out, surface parameters) must be a device descriptor
 * (INVALID_ARGUMENT otherwise), and the operation list with its paint,
 * gradient and stop tables is staged in a fixed device scratch of
 * z2d_gpu_composite_noalloc_max_bytes() bytes (1 MiB) that each device
 * holds for the life of the process; a job that needs more returns
 * BUDGET_EXCEEDED.  Calls on one device share that scratch: calls that
 * overlap on one device (from several threads or contexts) race and must
 * be serialized by the caller.  Synchronous, like z2d_gpu_composite; it
 * records no device timing (z2d_gpu_composite_timing is not updated). */
int32_t z2d_gpu_composite_noalloc(z2d_gpu_ctx *ctx, z2d_surface_desc *dst, int32_t dst_x,
                                  int32_t dst_y, const z2d_comp_op *ops, uint64_t n_ops,
                                  uint32_t precision, z2d_surface_desc *out);
uint64_t z2d_gpu_composite_noalloc_max_bytes(void);

/* ---- text ---------------------------------------------------------------------
 * Font.loadBuffer (src/Font.zig:28-45): validates the font (magic, table
 * checksums, required tables, cmap subtable, loca format) and keeps a copy of
 * the bytes.  A font may be shared between threads (its glyph cache is
 * locked). */
typedef struct z2d_font z2d_font;
int32_t z2d_font_load(const void *data, uint64_t len, z2d_font **out);
void z2d_font_destroy(z2d_font *font);
/* Font.Meta (src/Font.zig:173-260): units per em, advance width max, number
 * of long hmetrics, flags (bit 0 lsb_is_at_x_zero, bit 1 long loca, bit 2
 * full-range cmap). */
int32_t z2d_font_info(z2d_font *font, uint32_t out[4]);
/* Glyph.init (src/internal/Glyph.zig): glyph index, advance, lsb and glyf
 * offset (-1 = no outline) for a codepoint. */
int32_t z2d_font_glyph(z2d_font *font, uint32_t codepoint, int32_t out[4]);
/* The path text.show builds (src/text.zig:73-196): glyph outlines laid out
 * with advances and kerning, in device space.  Host only; *n_out is the node
 * count (INVALID_ARGUMENT when cap is smaller). */
int32_t z2d_text_path(z2d_font *font, const char *utf8, uint64_t len, double x, double y,
                      const z2d_text_opts *opts, z2d_path_node *out, uint64_t cap,
                      uint64_t *n_out);
/* text.show (src/text.zig:73-196): that path filled with painter.fill's
 * checks and raster (z2d_gpu_fill). */
int32_t z2d_gpu_text_show(z2d_gpu_ctx *ctx, z2d_surface_desc *surface,
                          const z2d_pattern_desc *pattern, z2d_font *font, const char *utf8,
                          uint64_t len, double x, double y, const z2d_text_opts *opts);
/* text.show into a batch: one fill draw per call, rendered with the batch. */
int32_t z2d_gpu_batch_text(z2d_gpu_batch *b, z2d_gpu_surface *surface,
                           const z2d_pattern_desc *pattern, z2d_font *font, const char *utf8,
                           uint64_t len, double x, double y, const z2d_text_opts *opts);

/* ---- PNG export ------------------------------------------------------------
 * writeToPNGFile (src/export_png.zig:56-84), into memory (the caller writes
 * the file).  Same IHDR, gAMA and IEND chunks (:100-145, :422-446) and the
 * same scanline stream (:153-360: filter byte 0, demultiplied RGBA,
 * MSB-first packed alpha, optional sRGB re-encode); the zlib stream is this
 * library's own device deflate, so a decoder yields the reference's bytes
 * while the file is not byte-identical.  `surface` is a host or device
 * descriptor; `profile` a z2d_png_profile.  The reference panics on an
 * empty surface; here INVALID_WIDTH / INVALID_HEIGHT.  An `out_cap` below
 * the encoded size gives INVALID_ARGUMENT with *out_len = the size needed;
 * z2d_gpu_png_bound returns a size that always suffices. */
int32_t z2d_gpu_png_bound(uint32_t format, int32_t width, int32_t height, uint32_t profile,
                          uint64_t *bound);
int32_t z2d_gpu_png_encode(z2d_gpu_ctx *ctx, z2d_surface_desc *surface, uint32_t profile,
                           void *out, uint64_t out_cap, uint64_t *out_len);
/* Device time of the last z2d_gpu_png_encode's kernels, in ms (no reference
 * counterpart). */
int32_t z2d_gpu_png_timing(z2d_gpu_ctx *ctx, double *kernel_ms);

/* ---- device-resident surfaces ----------------------------------------------
 * Surface.init / initBuffer (src/surface.zig:97-184) for surfaces that stay
 * in HBM across calls. */
int32_t z2d_gpu_surface_create(z2d_gpu_ctx *ctx, uint32_t format, int32_t width,
                               int32_t height, z2d_gpu_surface **out);
int32_t z2d_gpu_surface_upload(z2d_gpu_surface *s, const void *host, uint64_t len);
int32_t z2d_gpu_surface_download(z2d_gpu_surface *s, void *host, uint64_t len);
/* Fills a descriptor (reserved = Z2D_SURFACE_DEVICE) for use with the
 * painter and compositor calls above. */
int32_t z2d_gpu_surface_desc(z2d_gpu_surface *s, z2d_surface_desc *out);
void z2d_gpu_surface_destroy(z2d_gpu_surface *s);

/* ---- batches ----------------------------------------------------------------
 * Many fills on many resident surfaces rendered by one pipeline launch
 * (edge build, binning, one persistent tile-render launch).  Draws on one
 * surface are applied in the order they were added, exactly like
 * successive painter.fill calls.  No reference counterpart: it batches the
 * reference's per-call painter.fill. */
int32_t z2d_gpu_batch_create(z2d_gpu_ctx *ctx, z2d_gpu_batch **out);
int32_t z2d_gpu_batch_fill(z2d_gpu_batch *b, z2d_gpu_surface *surface,
                           const z2d_pattern_desc *pattern, const z2d_path_node *nodes,
                           uint64_t n_nodes, const z2d_fill_opts *opts);
int32_t z2d_gpu_batch_stroke(z2d_gpu_batch *b, z2d_gpu_surface *surface,
                             const z2d_pattern_desc *pattern, const z2d_path_node *nodes,
                             uint64_t n_nodes, const z2d_stroke_opts *opts);
int32_t z2d_gpu_batch_fill_polygon(z2d_gpu_batch *b, z2d_gpu_surface *surface,
                                   const z2d_pattern_desc *pattern, const z2d_edge *edges,
                                   uint64_t n_edges, const double extents[4],
                                   uint32_t anti_aliasing_mode, uint32_t fill_rule,
                                   uint32_t op, uint32_t precision);
/* Uploads the batch's inputs to HBM (no rendering). */
int32_t z2d_gpu_batch_prepare(z2d_gpu_batch *b);
/* Renders the prepared batch.  mode 0: one persistent launch for all
 * surfaces; mode 1: one launch per surface.  Asynchronous. */
int32_t z2d_gpu_batch_render(z2d_gpu_batch *b, uint32_t mode);
int32_t z2d_gpu_batch_wait(z2d_gpu_batch *b);
/* Device time of the last z2d_gpu_composite call's kernel, in milliseconds
 * (HIP events on the context stream around the launch).  A measurement hook
 * with no counterpart in the reference. */
int32_t z2d_gpu_composite_timing(z2d_gpu_ctx *ctx, double *kernel_ms);
/* Rows per band (the bin height of the edge and draw lists): a multiple of 4
 * from 4 to 32, 16 by default.  The output does not depend on it.  No
 * counterpart in the reference. */
int32_t z2d_gpu_batch_set_band_height(z2d_gpu_batch *b, uint32_t rows);
/* Queue order of the single-launch render: 0 natural (tile by tile), 1
 * longest strip lists first (a counting sort on the device after binning).
 * The output does not depend on it.  No counterpart in the reference. */
int32_t z2d_gpu_batch_set_queue_order(z2d_gpu_batch *b, uint32_t longest_first);
/* Autotuner: renders the batch with band heights 8, 16 and 32, each in the
 * natural and the longest-first queue order (one warm-up and `reps` timed
 * pipelines per candidate, launch mode `mode`), keeps the fastest (*band_h,
 * *longest_first, the candidates' mean pipeline ms in ms_out: natural 8, 16,
 * 32 then longest-first 8, 16, 32; any may be null) and prepares the batch
 * for it.  The batch's surfaces are restored after every candidate, so they
 * end as they started.  No counterpart in the reference. */
int32_t z2d_gpu_batch_autotune(z2d_gpu_batch *b, uint32_t mode, uint32_t reps,
                               uint32_t *band_h, uint32_t *longest_first, double ms_out[6]);
/* enable != 0: each render of this batch replays one hipGraph holding the
 * whole pipeline (captured at the first render after the batch is prepared,
 * and again when the launch mode or the counter setting changes).  No
 * counterpart in the reference; results are those of the eager render. */
int32_t z2d_gpu_batch_use_graph(z2d_gpu_batch *b, uint32_t enable);
/* Device time of the last render: whole pipeline and the tile-render
 * kernel(s) alone, in milliseconds (HIP events on the context stream);
 * render_ms is -1 after a graph replay. */
int32_t z2d_gpu_batch_timing(z2d_gpu_batch *b, double *pipeline_ms, double *render_ms);
/* Number of draws / surfaces / flattened nodes in the batch. */
int32_t z2d_gpu_batch_stats(z2d_gpu_batch *b, uint64_t out[4]);
/* Work counters of the last render (enable before rendering): out[0] =
 * pixel-draws (pixels with nonzero coverage summed over draws), out[1] =
 * edges built, out[2] = tile-draw pairs rasterized, out[3] = nonzero if a
 * bin overflowed, out[4] = tile-draw pairs on the unpacked (very many edges)
 * path, out[5] = capacity of the per-band edge lists.  No reference
 * counterpart: the benchmark's unit of work. */
int32_t z2d_gpu_batch_counters(z2d_gpu_batch *b, uint32_t enable, uint64_t out[8]);
int32_t z2d_gpu_batch_reset(z2d_gpu_batch *b);
void z2d_gpu_batch_destroy(z2d_gpu_batch *b);

/* ---- serving ------------------------------------------------------------------
 * A request loop with continuous batching: each request is one surface job
 * (a cleared surface, fills applied in order, then an optional alpha8 clip
 * mask applied as Surface.composite(mask, .dst_in)).  Requests are copied at
 * submit, admitted while their device memory fits `hbm_budget`, and rendered
 * in batches on two streams so requests arriving while others run are
 * batched and overlapped on the device.  No reference counterpart (the
 * reference renders one call at a time). */
typedef struct z2d_gpu_server z2d_gpu_server;

typedef struct z2d_fill_item {
  const z2d_pattern_desc *pattern;
  const z2d_path_node *nodes;
  uint64_t n_nodes;
  z2d_fill_opts opts;
} z2d_fill_item;

/* A fill or a stroke of a request (painter.fill / painter.stroke). */
enum { Z2D_DRAW_FILL = 0, Z2D_DRAW_STROKE = 1 };
typedef struct z2d_draw_item {
  uint32_t kind;                   /* Z2D_DRAW_FILL or Z2D_DRAW_STROKE */
  const z2d_pattern_desc *pattern;
  const z2d_path_node *nodes;
  uint64_t n_nodes;
  z2d_fill_opts fill;              /* kind == Z2D_DRAW_FILL */
  const z2d_stroke_opts *stroke;   /* kind == Z2D_DRAW_STROKE (its dashes are copied) */
} z2d_draw_item;

/* Batch form of many painter.fill / painter.stroke calls (src/painter.zig:
 * 66-143, :214-344), item i on surfaces[i], in order: the same draws and the
 * same per-item status (status[i], optional) as z2d_gpu_batch_fill /
 * z2d_gpu_batch_stroke called once per item, a failed item left out as a
 * failed call is.  The strokes' host tessellation runs on `threads` host
 * threads (0: one per hardware thread, at most 16) before the draws are
 * appended.  Returns the first item failure, else Z2D_OK.  No reference
 * counterpart: a caller's loop of painter calls, made parallel. */
int32_t z2d_gpu_batch_draw_n(z2d_gpu_batch *b, z2d_gpu_surface *const *surfaces,
                             const z2d_draw_item *items, uint64_t n, uint32_t threads,
                             int32_t *status);

typedef struct z2d_request_desc {
  uint32_t format;
  int32_t width;
  int32_t height;
  z2d_pixel clear;            /* initial contents (Surface.initPixel) */
  const uint8_t *clip_mask;   /* optional alpha8 width*height mask */
  const z2d_fill_item *fills;
  uint64_t n_fills;
  const z2d_draw_item *draws; /* fills and strokes, drawn in order after `fills` */
  uint64_t n_draws;
  /* the result in `out`: 0 = the surface's bytes; 1 + z2d_png_profile = a
   * PNG file (writeToPNGFile, src/export_png.zig:56-84), encoded on the
   * device after the render.  A file larger than out_len completes with
   * INVALID_ARGUMENT. */
  uint32_t encode;
  uint32_t reserved;
  uint64_t *out_written;      /* optional: bytes written to `out` (or needed) */
} z2d_request_desc;

int32_t z2d_gpu_server_create(z2d_gpu_ctx *ctx, uint64_t hbm_budget, uint32_t max_batch,
                              z2d_gpu_server **out);
/* `out` (optional, >= the surface's byte size) receives the result and must
 * stay valid until the request completes.  BUDGET_EXCEEDED if the request
 * can never fit. */
int32_t z2d_gpu_server_submit(z2d_gpu_server *s, const z2d_request_desc *req, void *out,
                              uint64_t out_len, uint64_t *id);
/* Many requests submitted in one call (a front end handing over the
 * requests that arrived together): request i is reqs[i] with outs[i] and
 * out_lens[i] (outs and out_lens optional), checked and queued as
 * z2d_gpu_server_submit would, in order and under one lock.  ids[i]
 * receives its id (0 if refused) and status[i] (optional) its submit
 * status.  Returns the first refusal, else Z2D_OK. */
int32_t z2d_gpu_server_submit_n(z2d_gpu_server *s, const z2d_request_desc *const *reqs,
                                void *const *outs, const uint64_t *out_lens, uint64_t n,
                                uint64_t *ids, int32_t *status);
/* 1 = done (status filled, request released), 0 = pending. */
int32_t z2d_gpu_server_poll(z2d_gpu_server *s, uint64_t id, int32_t *status);
int32_t z2d_gpu_server_wait(z2d_gpu_server *s, uint64_t id, int32_t *status, double timeout_ms);
/* Next completed request in completion order (1), or 0 on timeout. */
int32_t z2d_gpu_server_wait_any(z2d_gpu_server *s, uint64_t *id, int32_t *status,
                                double timeout_ms);
/* Up to `max` completed requests in completion order, waiting up to
 * timeout_ms for the first: ids, statuses (optional) and completion times
 * (optional, seconds of CLOCK_MONOTONIC, stamped when the result reached
 * the caller's buffer).  Returns their count, 0 on timeout. */
int32_t z2d_gpu_server_wait_some(z2d_gpu_server *s, uint64_t *ids, int32_t *status,
                                 double *t_done, uint32_t max, double timeout_ms);
/* submitted, completed, rejected, batches, pixel-draws, bytes in use, peak
 * bytes, queued. */
int32_t z2d_gpu_server_stats(z2d_gpu_server *s, uint64_t out[8]);
void z2d_gpu_server_destroy(z2d_gpu_server *s);

/* Measurement utility with no reference counterpart: copies `bytes` between
 * two device buffers `iters` times with a plain stream-copy kernel, the
 * passes rotating over three buffer sets (so no pass reads what the last-
 * level cache kept), and reports read+write GB/s (the achievable-bandwidth
 * reference). */
int32_t z2d_gpu_stream_copy(z2d_gpu_ctx *ctx, uint64_t bytes, uint32_t iters, double *gbps);

/* Measurement utility (no reference counterpart): the compositor's access
 * mix as a reference stream, two `bytes` buffers read and one written
 * (dst = a ^ b), `iters` times over three rotating buffer sets; reports
 * read+write GB/s. */
int32_t z2d_gpu_stream_read2(z2d_gpu_ctx *ctx, uint64_t bytes, uint32_t iters, double *gbps);

/* Synchronizes the context stream. */
int32_t z2d_gpu_sync(z2d_gpu_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* Z2D_GPU_H */
