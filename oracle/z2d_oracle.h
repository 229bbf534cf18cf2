/*
 * z2d_oracle.h - CPU restatement of the reference's fill/stroke/composite
 * hot path, used ONLY as the checker: by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg.  Nothing in the product (z2d_amd/,
 * include/) links, loads or calls it.
 *
 * Every function here follows a reference file:line, cited beside its
 * definition in z2d_oracle.c.  The descriptors are the same POD types that
 * cross the product's C ABI (include/z2d_types.h), so a test can hand one
 * set of buffers to both.
 */
#ifndef Z2D_ORACLE_H
#define Z2D_ORACLE_H

#include "../include/z2d_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* painter.fill (src/painter.zig:66-143). Mutates surface->data in place. */
int32_t zo_fill(z2d_surface_desc *surface, const z2d_pattern_desc *pattern,
                const z2d_path_node *nodes, uint64_t n_nodes,
                const z2d_fill_opts *opts);

/* Rasterize an already-built edge list (the step after fill_plotter /
 * stroke_plotter in src/painter.zig:99-142, :308-343).  `scale` must match
 * the AA mode (1 for none, 4 otherwise); extents as Polygon.extent_*. */
int32_t zo_fill_polygon(z2d_surface_desc *surface,
                        const z2d_pattern_desc *pattern, const z2d_edge *edges,
                        uint64_t n_edges, const double extents[4],
                        uint32_t aa_mode, uint32_t fill_rule, uint32_t op,
                        uint32_t precision);

/* fill_plotter.plot (src/internal/tess/fill_plotter.zig:21-97): writes at
 * most `cap` edges, returns the number plotted (or -status on error);
 * extents_out = {top, bottom, left, right}. */
int64_t zo_plot_fill(const z2d_path_node *nodes, uint64_t n_nodes, double scale,
                     double tolerance, z2d_edge *edges_out, uint64_t cap,
                     double extents_out[4]);

/* SurfaceCompositor.run (src/compositor.zig:302-440). */
void zo_composite(z2d_surface_desc *dst, int32_t dst_x, int32_t dst_y,
                  const z2d_comp_op *ops, uint64_t n_ops, uint32_t precision);

/* compositor.runPixel (src/compositor.zig:626-643): one operator on two
 * pixels; result in dst's format. */
z2d_pixel zo_run_pixel(uint32_t precision, z2d_pixel dst, z2d_pixel src,
                       uint32_t op);

/* T.fromPixel (src/pixel.zig:398-432 for the 32-bit formats, :569-630 for
 * the alpha formats): `px` converted to format `to`. */
z2d_pixel zo_pixel_convert(uint32_t to, z2d_pixel px);

/* RGBA.multiply (src/pixel.zig:483-491). */
z2d_pixel zo_pixel_multiply(z2d_pixel px);

/* Pixel.fromColor (src/pixel.zig:119-121): an RGBA pixel. */
z2d_pixel zo_pixel_from_color(const z2d_color *c);

/* PathNode.isClosedNodeSet (src/internal/path_nodes.zig:23-36). */
int zo_is_closed_node_set(const z2d_path_node *nodes, uint64_t n);

/* Gradient.getOffset (src/gradient.zig:349-375, :605-652, :731-742). */
float zo_gradient_offset(const z2d_gradient_desc *g, int32_t x, int32_t y);

/* Stop.List.search (src/gradient.zig:832-905) over the gradient's stops as
 * added (addAssumeCapacity: clamped, stable by offset): the colors on both
 * sides of `offset` and the relative offset between them. */
void zo_stop_search(const z2d_gradient_desc *g, float offset, z2d_color *c0, z2d_color *c1,
                    float *t);

/* InterpolationMethod.interpolateEncodeVec for one lane
 * (src/internal/color_vector.zig:59-83, :191-459): the two colors (after
 * Color.init) interpolated at t in the method's space, encoded to a
 * pre-multiplied RGBA pixel. */
z2d_pixel zo_interpolate_encode(uint32_t method, uint32_t polar, const z2d_color *a,
                                const z2d_color *b, float t);

/* multisample.addSpan (src/internal/raster/multisample.zig:239-279) into a
 * dense per-pixel coverage array (`cov` indexed by pixel, values 0..16):
 * x and len in sample columns. */
void zo_msaa_add_span(uint8_t *cov, uint32_t x, uint32_t len);

/* Gradient.getPixel (src/gradient.zig:333-341): pre-multiplied RGBA. */
z2d_pixel zo_gradient_pixel(const z2d_gradient_desc *g, int32_t x, int32_t y);

/* Dither.getPixel (src/Dither.zig:73-98). */
z2d_pixel zo_dither_pixel(const z2d_dither_desc *d, int32_t x, int32_t y);

/* The compositor's vectorized paint-source sampling for one pixel:
 * RGBA16Vec.fromGradient / fromDither (src/compositor.zig:749-804) in
 * integer precision, out = {r,g,b,a} 0..255 pre-multiplied. */
void zo_sample_pattern_int(const z2d_pattern_desc *p, int32_t x, int32_t y,
                           uint16_t out[4]);

/* std.math.pow(f32) as used for gamma (src/color.zig:336-354). */
float zo_powf(float x, float y);

/* writePNGIDATStream's scanline stream (src/export_png.zig:153-305) with
 * encodeRGBAVec (:313-360): for each row a filter byte 0, then rgb/xrgb as
 * r,g,b; argb/rgba demultiplied r,g,b,a; alpha8 as is; alpha4/2/1 packed
 * MSB-first with each row starting on a byte.  The bytes the reference hands
 * its zlib compressor.  Returns the stream length; `out` may be NULL. */
uint64_t zo_png_scanlines(const z2d_surface_desc *s, uint32_t profile, uint8_t *out);

/* std.math.atan2 (f64) of Zig's standard library: the fdlibm algorithm
 * (Conic.getOffset, src/gradient.zig:731-742). */
double zo_atan2(double y, double x);

/* Polygon.inBox (src/internal/tess/Polygon.zig:142-201). */
int zo_in_box(const double extents[4], double scale, int32_t w, int32_t h);

/* Counts pixel-draws of the MSAA path for an edge list (pixels composited
 * with nonzero coverage); used by the benchmark's unit of work. */
uint64_t zo_count_coverage(const z2d_edge *edges, uint64_t n_edges,
                           const double extents[4], uint32_t fill_rule,
                           int32_t width, int32_t height);

#ifdef __cplusplus
}
#endif

#endif
