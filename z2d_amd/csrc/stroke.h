// Host stroke tessellation: path nodes -> the polygon (edge list + extents)
// that painter.stroke fills with the non-zero rule.
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/z2d_types.h"

namespace z2d {

// stroke_plotter.PlotterOptions (src/internal/tess/stroke_plotter.zig:28-38),
// after painter.stroke's thin-line adjustments (src/painter.zig:283-304).
struct StrokeParams {
  uint32_t cap_mode;   // Z2D_CAP_*
  uint32_t join_mode;  // Z2D_JOIN_*
  double miter_limit;
  double thickness;
  double tolerance;
  double scale;        // AA scale applied to every contour point
  double ctm[6];       // ax, by, cx, dy, tx, ty
  const double *dashes;
  uint64_t n_dashes;
  double dash_offset;
};

// Polygon with edges in scaled space and its extents (top, bottom, left,
// right), as Polygon.addEdge builds them.
struct StrokePolygon {
  std::vector<z2d_edge> edges;
  double ext[4] = {0.0, 0.0, 0.0, 0.0};
};

// stroke_plotter.plot / dashed_plotter.plot.  Returns Z2D_OK or
// Z2D_ERR_INVALID_STATE (a line or curve without a current point).
int32_t stroke_tessellate(const z2d_path_node *nodes, uint64_t n, const StrokeParams &p,
                          StrokePolygon &out);

}  // namespace z2d
