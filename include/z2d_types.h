/*
 * z2d_types.h - plain-old-data descriptors crossing the z2d GPU backend's
 * C ABI.  No Zig, HIP or C++ types appear here; every enum keeps the
 * numbering of the reference's Zig declaration order so a Zig shim can
 * @intFromEnum its values straight across.
 *
 * Reference enums mirrored:
 *   pixel.Format          src/pixel.zig:47-56
 *   compositor.Operator   src/compositor.zig:46-158
 *   compositor.Precision  src/compositor.zig:214-217
 *   options.FillRule      src/options.zig:23-37
 *   options.JoinMode      src/options.zig:40-55
 *   options.CapMode       src/options.zig:58-70
 *   options.AntiAliasMode src/options.zig:73-91
 *   PathNodeTag           src/internal/path_nodes.zig:9-14
 *   GradientType          src/gradient.zig:25-29
 *   InterpolationMethod   src/color.zig:104-131
 *   Dither.Type/Source    src/Dither.zig:29-52
 *   PatternType           src/pattern.zig:14-18
 */
#ifndef Z2D_TYPES_H
#define Z2D_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums ------------------------------------------------------------ */

enum z2d_format {
  Z2D_FORMAT_ARGB = 0,
  Z2D_FORMAT_XRGB = 1,
  Z2D_FORMAT_RGB = 2,
  Z2D_FORMAT_RGBA = 3,
  Z2D_FORMAT_ALPHA8 = 4,
  Z2D_FORMAT_ALPHA4 = 5,
  Z2D_FORMAT_ALPHA2 = 6,
  Z2D_FORMAT_ALPHA1 = 7,
};

/* WriteToPNGFileOptions.color_profile (src/export_png.zig:33-41): NONE is the
 * reference's null (no gAMA chunk, no re-encode); LINEAR writes gAMA 100000;
 * SRGB writes gAMA 45454 and re-encodes the color channels with pow(c, 1/2.2)
 * (src/export_png.zig:125-145, :313-360). */
enum z2d_png_profile {
  Z2D_PNG_PROFILE_NONE = 0,
  Z2D_PNG_PROFILE_LINEAR = 1,
  Z2D_PNG_PROFILE_SRGB = 2,
};

enum z2d_operator {
  Z2D_OP_CLEAR = 0,
  Z2D_OP_SRC,
  Z2D_OP_DST,
  Z2D_OP_SRC_OVER,
  Z2D_OP_DST_OVER,
  Z2D_OP_SRC_IN,
  Z2D_OP_DST_IN,
  Z2D_OP_SRC_OUT,
  Z2D_OP_DST_OUT,
  Z2D_OP_SRC_ATOP,
  Z2D_OP_DST_ATOP,
  Z2D_OP_XOR,
  Z2D_OP_PLUS,
  Z2D_OP_MULTIPLY,
  Z2D_OP_SCREEN,
  Z2D_OP_OVERLAY,
  Z2D_OP_DARKEN,
  Z2D_OP_LIGHTEN,
  Z2D_OP_COLOR_DODGE,
  Z2D_OP_COLOR_BURN,
  Z2D_OP_HARD_LIGHT,
  Z2D_OP_SOFT_LIGHT,
  Z2D_OP_DIFFERENCE,
  Z2D_OP_EXCLUSION,
  Z2D_OP_HUE,
  Z2D_OP_SATURATION,
  Z2D_OP_COLOR,
  Z2D_OP_LUMINOSITY,
  Z2D_OP_COUNT
};

enum z2d_precision { Z2D_PRECISION_INTEGER = 0, Z2D_PRECISION_FLOAT = 1 };

enum z2d_fill_rule { Z2D_FILL_NON_ZERO = 0, Z2D_FILL_EVEN_ODD = 1 };

enum z2d_join_mode { Z2D_JOIN_MITER = 0, Z2D_JOIN_ROUND = 1, Z2D_JOIN_BEVEL = 2 };

enum z2d_cap_mode { Z2D_CAP_BUTT = 0, Z2D_CAP_ROUND = 1, Z2D_CAP_SQUARE = 2 };

enum z2d_aa_mode {
  Z2D_AA_NONE = 0,
  Z2D_AA_DEFAULT = 1,
  Z2D_AA_MULTISAMPLE_4X = 2,
  Z2D_AA_SUPERSAMPLE_4X = 3,
};

enum z2d_node_tag {
  Z2D_NODE_MOVE_TO = 0,
  Z2D_NODE_LINE_TO = 1,
  Z2D_NODE_CURVE_TO = 2,
  Z2D_NODE_CLOSE_PATH = 3,
};

enum z2d_gradient_type {
  Z2D_GRADIENT_LINEAR = 0,
  Z2D_GRADIENT_RADIAL = 1,
  Z2D_GRADIENT_CONIC = 2,
};

enum z2d_interp_method {
  Z2D_INTERP_LINEAR_RGB = 0,
  Z2D_INTERP_SRGB = 1,
  Z2D_INTERP_HSL = 2,
};

enum z2d_polar_method {
  Z2D_POLAR_SHORTER = 0,
  Z2D_POLAR_LONGER = 1,
  Z2D_POLAR_INCREASING = 2,
  Z2D_POLAR_DECREASING = 3,
};

/* Color spaces of a stored color (Color union, src/color.zig:41-45). */
enum z2d_color_space {
  Z2D_COLOR_LINEAR_RGB = 0,
  Z2D_COLOR_SRGB = 1,
  Z2D_COLOR_HSL = 2,
};

enum z2d_dither_type {
  Z2D_DITHER_NONE = 0,
  Z2D_DITHER_BAYER = 1,
  Z2D_DITHER_BLUE_NOISE = 2,
};

enum z2d_dither_source {
  Z2D_DITHER_SRC_PIXEL = 0,
  Z2D_DITHER_SRC_COLOR = 1,
  Z2D_DITHER_SRC_GRADIENT = 2,
};

enum z2d_pattern_kind {
  Z2D_PATTERN_OPAQUE = 0,
  Z2D_PATTERN_GRADIENT = 1,
  Z2D_PATTERN_DITHER = 2,
};

/* Compositor operation parameter kinds
 * (SurfaceCompositor.Operation.Param, src/compositor.zig:249-294). */
enum z2d_param_kind {
  Z2D_PARAM_NONE = 0,
  Z2D_PARAM_DITHER = 1,
  Z2D_PARAM_GRADIENT = 2,
  Z2D_PARAM_PIXEL = 3,
  Z2D_PARAM_SURFACE = 4,
};

/* ---- descriptors -------------------------------------------------------- */

/* A surface: the caller's pixel buffer.  Layout per src/surface.zig:585-598
 * and src/pixel.zig: argb/xrgb/rgb/rgba 4 B/px, alpha8 1 B/px, alpha4/2/1
 * packed LSB-first with no row padding (ceil(w*h*bits/8) bytes). */
typedef struct z2d_surface_desc {
  uint32_t format;
  int32_t width;
  int32_t height;
  uint32_t reserved;
  uint8_t *data;
  uint64_t len;
} z2d_surface_desc;

/* A single pixel of any format (pixel.Pixel, src/pixel.zig:101-141).  For
 * alpha formats only `a` is used, at the format's native bit depth. */
typedef struct z2d_pixel {
  uint32_t format;
  uint8_t r, g, b, a;
} z2d_pixel;

/* A color (color.Color, src/color.zig:41-45): v = {r,g,b,a} for the RGB
 * spaces, {h,s,l,a} for HSL.  Color.init clamping is applied by the
 * consumer, so raw InitArgs values are accepted. */
typedef struct z2d_color {
  uint32_t space;
  float v[4];
} z2d_color;

/* A gradient stop as passed to addStop (src/gradient.zig:797-830); the
 * consumer clamps offsets and stable-sorts by offset like the reference. */
typedef struct z2d_stop {
  float offset;
  z2d_color color;
} z2d_stop;

/* A gradient (gradient.Gradient, src/gradient.zig:42-46).
 * geom: linear  {x0, y0, x1, y1}
 *       radial  {inner_x, inner_y, inner_radius, outer_x, outer_y, outer_radius}
 *       conic   {x, y, angle}
 * matrix: the gradient's stored `transformation` field {ax, by, cx, dy, tx,
 * ty}, i.e. the inverse of the matrix given to setTransformation. */
typedef struct z2d_gradient_desc {
  uint32_t type;
  uint32_t method;
  uint32_t polar;
  uint32_t n_stops;
  double geom[6];
  double matrix[6];
  const z2d_stop *stops;
} z2d_gradient_desc;

/* Dither (src/Dither.zig:54-69). */
typedef struct z2d_dither_desc {
  uint32_t type;
  uint32_t source;
  z2d_pixel pixel;
  z2d_color color;
  uint32_t scale;
  const z2d_gradient_desc *gradient;
} z2d_dither_desc;

/* Pattern (src/pattern.zig:32-50). */
typedef struct z2d_pattern_desc {
  uint32_t kind;
  z2d_pixel pixel;
  const z2d_gradient_desc *gradient;
  z2d_dither_desc dither;
} z2d_pattern_desc;

/* Path node (PathNode, src/internal/path_nodes.zig:16-58):
 * move_to/line_to use v[0..1]; curve_to uses p1=v[0..1], p2=v[2..3],
 * p3=v[4..5]; close_path uses nothing. */
typedef struct z2d_path_node {
  uint32_t tag;
  uint32_t reserved;
  double v[6];
} z2d_path_node;

/* Polygon edge (Polygon.Edge, src/internal/tess/Polygon.zig:18-47). */
typedef struct z2d_edge {
  double y0, y1, x_start, x_inc;
} z2d_edge;

/* painter.FillOptions (src/painter.zig:28-48). */
typedef struct z2d_fill_opts {
  uint32_t anti_aliasing_mode;
  uint32_t fill_rule;
  uint32_t op;
  uint32_t precision;
  double tolerance;
} z2d_fill_opts;

/* text.ShowTextOptions (src/text.zig:32-46): the path's transformation
 * {ax, by, cx, dy, tx, ty} (Transformation.zig field order), the font size in
 * user units (em size), and the fill options. */
typedef struct z2d_text_opts {
  double transformation[6];
  double size;
  z2d_fill_opts fill;
} z2d_text_opts;

/* painter.StrokeOptions (src/painter.zig:145-198). */
typedef struct z2d_stroke_opts {
  uint32_t anti_aliasing_mode;
  uint32_t line_cap_mode;
  uint32_t line_join_mode;
  uint32_t op;
  uint32_t precision;
  uint32_t hairline;
  const double *dashes;
  uint64_t n_dashes;
  double dash_offset;
  double line_width;
  double miter_limit;
  double tolerance;
  double transformation[6];
} z2d_stroke_opts;

/* Compositor operation parameter (src/compositor.zig:249-294). */
typedef struct z2d_comp_param {
  uint32_t kind;
  z2d_pixel pixel;
  const z2d_gradient_desc *gradient;
  z2d_dither_desc dither;
  const z2d_surface_desc *surface;
} z2d_comp_param;

/* SurfaceCompositor.Operation (src/compositor.zig:233-295). */
typedef struct z2d_comp_op {
  uint32_t op;
  z2d_comp_param dst;
  z2d_comp_param src;
} z2d_comp_op;

/* Status codes.  Positive values mirror the reference's error sets in order
 * (painter.FillError / StrokeError, src/painter.zig:54-63, :204-209;
 * Surface.Error src/surface.zig:85-91; InternalError
 * src/internal/InternalError.zig:8-10; Transformation.Error). Negative values
 * are backend conditions with no reference counterpart. */
enum z2d_status {
  Z2D_OK = 0,
  Z2D_ERR_PATH_NOT_CLOSED = 1,
  Z2D_ERR_PIXEL_SOURCE_NOT_PREMULTIPLIED = 2,
  Z2D_ERR_INVALID_WIDTH = 3,
  Z2D_ERR_INVALID_HEIGHT = 4,
  Z2D_ERR_INVALID_STATE = 5,
  Z2D_ERR_OUT_OF_MEMORY = 6,
  Z2D_ERR_INVALID_MATRIX = 7,
  Z2D_ERR_INVALID_ARGUMENT = -1,
  Z2D_ERR_DEVICE = -2,
  Z2D_ERR_BUDGET_EXCEEDED = -3,
  Z2D_ERR_NO_CONTEXT = -4,
  Z2D_ERR_UNSUPPORTED = -5,
  /* text.ShowTextError's UTF-8 errors (src/text.zig:48-67) */
  Z2D_ERR_INVALID_TEXT = -6,
  /* Font.LoadBufferError, Font.FileError and Glyph.Outline.InitError
   * (src/Font.zig, src/internal/Glyph.zig): not a font, checksum mismatch,
   * missing table, no usable cmap, bad loca format, a read past the end, a
   * malformed glyph */
  Z2D_ERR_INVALID_FONT = -7,
};

#ifdef __cplusplus
}
#endif

#endif /* Z2D_TYPES_H */
