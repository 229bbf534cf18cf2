// SPDX-License-Identifier: MPL-2.0
// Restates code of z2d (Copyright (c) 2024-2026 Chris Marchesi, MPL-2.0),
// parts of which are derived from the Cairo project (Copyright (c) 2002
// University of Southern California; cairo-path-stroke-polygon.c,
// cairo-pen.c, cairo-slope.c, cairo-spline.c, cairo-arc.c).  The Mozilla
// Public License v2.0 applies to this file: https://mozilla.org/MPL/2.0/.
//
// Stroke tessellation on the host, restating the reference's stroker step
// for step (the polygon is then filled non-zero by the device rasterizer):
//   stroke_plotter.zig  (plot, join, plotSingle, plotOpenJoined,
//                        plotClosedJoined; src/internal/tess/stroke_plotter.zig)
//   dashed_plotter.zig  (dash stepping and the initial-segment bookkeeping)
//   Face.zig, Slope.zig, Pen.zig, Dasher.zig, Spline.zig, point_buffer.zig,
//   Polygon.zig (Contour, addEdgesFromContour), arc.zig
//   (transformed_circle_major_axis), Transformation.zig.
// Arithmetic follows the reference's expressions and order (compiled without
// floating-point contraction); cos/sin/acos/hypot/sqrt are the C library's.
#include "stroke.h"

#include <cfloat>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>
#include <memory>

namespace z2d {
namespace {

struct P {
  double x, y;
};
inline bool peq(P a, P b) { return a.x == b.x && a.y == b.y; }
// the same bits (signed zeros kept apart)
inline bool pbits(P a, P b) { return std::memcmp(&a, &b, sizeof(P)) == 0; }

// Transformation (src/Transformation.zig): [ax by tx; cx dy ty].
struct Xf {
  double ax = 1, by = 0, cx = 0, dy = 1, tx = 0, ty = 0;
  bool identity() const {
    return ax == 1 && by == 0 && cx == 0 && dy == 1 && tx == 0 && ty == 0;
  }
  double det() const { return ax * dy - by * cx; }
  // Transformation.inverse (:103-160); false = InvalidMatrix.
  bool inverse(Xf &o) const {
    if (by == 0 && cx == 0) {
      if (ax == 0 || dy == 0) return false;
      if (ax != 1 || dy != 1) {
        o.ax = 1 / ax;
        o.by = 0;
        o.cx = 0;
        o.dy = 1 / dy;
        o.tx = -tx / ax;
        o.ty = -ty / dy;
      } else {
        o.ax = 1;
        o.by = 0;
        o.cx = 0;
        o.dy = 1;
        o.tx = -tx;
        o.ty = -ty;
      }
      return true;
    }
    const double d = det();
    if (d == 0) return false;
    const double s = 1 / d;
    o.ax = dy * s;
    o.by = -by * s;
    o.cx = -cx * s;
    o.dy = ax * s;
    o.tx = (by * ty - dy * tx) * s;
    o.ty = (cx * tx - ax * ty) * s;
    return true;
  }
  void u2d_dist(double &x, double &y) const {
    const double ix = x, iy = y;
    x = ax * ix + by * iy;
    y = cx * ix + dy * iy;
  }
  void d2u_dist(double &x, double &y) const {
    Xf inv;
    inverse(inv);  // validated by the caller
    inv.u2d_dist(x, y);
  }
};

inline int fsign(double v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }

// Slope (src/internal/tess/Slope.zig).
struct Slope {
  double dx = 0, dy = 0;
  static Slope init(P a, P b) { return Slope{b.x - a.x, b.y - a.y}; }
  int compare(Slope b) const {
    const double eps = DBL_EPSILON;
    const double bdy = std::fabs(b.dy - dy) > eps ? b.dy : dy;
    const double bdx = std::fabs(b.dx - dx) > eps ? b.dx : dx;
    const int cmp = fsign(dy * bdx - bdy * dx);
    if (cmp != 0) return cmp;
    if (dx == 0 && dy == 0 && bdx == 0 && bdy == 0) return 0;
    if (dx == 0 && dy == 0) return 1;
    if (bdx == 0 && bdy == 0) return -1;
    if (fsign(dx) != fsign(bdx) || fsign(dy) != fsign(bdy))
      return (dx > 0 || (dx == 0 && dy > 0)) ? -1 : 1;
    return 0;
  }
  double normalize() {
    double rdx, rdy, mag;
    if (dx == 0.0) {
      rdx = 0.0;
      if (dy > 0.0) {
        mag = dy;
        rdy = 1.0;
      } else {
        mag = -dy;
        rdy = -1.0;
      }
    } else if (dy == 0.0) {
      rdy = 0.0;
      if (dx > 0.0) {
        mag = dx;
        rdx = 1.0;
      } else {
        mag = -dx;
        rdx = -1.0;
      }
    } else {
      mag = std::hypot(dx, dy);
      rdx = dx / mag;
      rdy = dy / mag;
    }
    dx = rdx;
    dy = rdy;
    return mag;
  }
};

bool miter_within_limit(Slope in, Slope out, double miter_limit) {
  in.normalize();
  out.normalize();
  const double dot = in.dx * out.dx + in.dy * out.dy;
  return 2 <= miter_limit * miter_limit * (1 + dot);
}

// arc.transformed_circle_major_axis (src/internal/arc.zig:92-261).
double circle_major_axis(const Xf &m, double radius) {
  const double a = m.ax, b = m.by, c = m.cx, d = m.dy;
  const double eps = 0.00390625;
  const double det = m.ax * m.dy - m.by * m.cx;
  if (std::fabs(det * det - 1.0) < eps) {
    if (std::fabs(m.by) < eps && std::fabs(m.cx) < eps) return radius;
    if (std::fabs(m.ax) < eps && std::fabs(m.dy) < eps) return radius;
  }
  const double i = a * a + b * b;
  const double j = c * c + d * d;
  const double f = 0.5 * (i + j);
  const double g = 0.5 * (i - j);
  const double h = a * c + b * d;
  return radius * std::sqrt(f + std::hypot(g, h));
}

// Pen (src/internal/tess/Pen.zig).
struct PenVertex {
  P point;
  Slope slope_cw, slope_ccw;
};
struct Pen {
  std::vector<PenVertex> v;
  Pen(double thickness, double tolerance, const Xf &ctm) {
    const double radius = thickness / 2;
    int n;
    const double major = circle_major_axis(ctm, radius);
    if (tolerance >= major * 4) {
      n = 1;
    } else if (tolerance >= major) {
      n = 4;
    } else {
      const double delta = std::acos(1 - tolerance / major);
      if (delta == 0) {
        n = 4;
      } else {
        const int k = (int)std::ceil(2 * M_PI / delta);
        n = k < 4 ? 4 : (k % 2 != 0 ? k + 1 : k);
      }
    }
    const bool reflect = ctm.det() < 0;
    v.reserve((size_t)n);
    for (int i = 0; i < n; i++) {
      double t = 2 * M_PI * (double)i / (double)n;
      if (reflect) t = -t;
      double dx = radius * std::cos(t), dy = radius * std::sin(t);
      ctm.u2d_dist(dx, dy);
      v.push_back(PenVertex{P{dx, dy}, Slope{}, Slope{}});
    }
    for (int i = 0; i < n; i++) {
      const int next = i >= n - 1 ? 0 : i + 1;
      const int prev = i == 0 ? n - 1 : i - 1;
      v[i].slope_cw = Slope::init(v[prev].point, v[i].point);
      v[i].slope_ccw = Slope::init(v[i].point, v[next].point);
    }
  }
  // Pen.vertexIteratorFor: the vertex run [start, end) walked forward
  // (clockwise) or backward.
  void range(Slope from, Slope to, bool clockwise, int &start, int &end) const {
    const int len = (int)v.size();
    int i;
    if (clockwise) {
      int low = 0, high = len;
      i = (low + high) >> 1;
      while (high - low > 1) {
        const bool lt = v[i].slope_cw.compare(from) < 0;
        low = lt ? i : low;
        high = lt ? high : i;
        i = (low + high) >> 1;
      }
      if (v[i].slope_cw.compare(from) < 0) {
        i += 1;
        if (i == len) i = 0;
      }
      start = i;
      if (to.compare(v[i].slope_ccw) >= 0) {
        low = i;
        high = i + len;
        i = (low + high) >> 1;
        while (high - low > 1) {
          const int j = i >= len ? i - len : i;
          const bool gt = v[j].slope_cw.compare(to) > 0;
          high = gt ? i : high;
          low = gt ? low : i;
          i = (low + high) >> 1;
        }
        if (i >= len) i -= len;
      }
      end = i;
    } else {
      int low = 0, high = len;
      i = (low + high) >> 1;
      while (high - low > 1) {
        const bool lt = from.compare(v[i].slope_ccw) < 0;
        low = lt ? i : low;
        high = lt ? high : i;
        i = (low + high) >> 1;
      }
      if (from.compare(v[i].slope_ccw) < 0) {
        i += 1;
        if (i == len) i = 0;
      }
      start = i;
      if (v[i].slope_cw.compare(to) <= 0) {
        low = i;
        high = i + len;
        i = (low + high) >> 1;
        while (high - low > 1) {
          const int j = i >= len ? i - len : i;
          const bool gt = to.compare(v[j].slope_ccw) > 0;
          high = gt ? i : high;
          low = gt ? low : i;
          i = (low + high) >> 1;
        }
        if (i >= len) i -= len;
      }
      end = i;
    }
    if (start < 0) start = 0;
    if (end < 0) end = 0;
  }
  template <typename F>
  void walk(Slope from, Slope to, bool clockwise, F &&f) const {
    int idx, end;
    range(from, to, clockwise, idx, end);
    const int len = (int)v.size();
    while (idx != end) {
      f(v[idx]);
      if (clockwise) {
        idx += 1;
        if (idx == len) idx = 0;
      } else {
        if (idx == 0) idx = len;
        idx -= 1;
      }
    }
  }
};

// Polygon.Contour: a point list plotted at the end or at the front; points
// are scaled as they are plotted.  The sequence is front (stored reversed)
// followed by back, so both ends grow in amortized constant time without a
// node per point.  The reference's insertions before the first corner
// (a run of points placed, in order, ahead of what the list held) are
// collected by the caller and prepended at once.
struct Contour {
  std::vector<P> front, back;
  double scale = 1.0;
  void plot(P p) { back.push_back(P{p.x * scale, p.y * scale}); }
  void plot_reverse(P p) { front.push_back(P{p.x * scale, p.y * scale}); }
  // points already scaled, in order, ahead of the current sequence
  void prepend_scaled(const std::vector<P> &pts) {
    for (size_t i = pts.size(); i-- > 0;) front.push_back(pts[i]);
  }
  P scaled(P p) const { return P{p.x * scale, p.y * scale}; }
  bool empty() const { return front.empty() && back.empty(); }
  template <class F>
  void each(F &&f) const {
    for (size_t i = front.size(); i-- > 0;) f(front[i]);
    for (const P &p : back) f(p);
  }
  // appends o's sequence and empties o
  void concat(Contour &o) {
    back.reserve(back.size() + o.front.size() + o.back.size());
    for (size_t i = o.front.size(); i-- > 0;) back.push_back(o.front[i]);
    back.insert(back.end(), o.back.begin(), o.back.end());
    o.reset();
  }
  // takes o's sequence as this one's (this one's points dropped)
  void take(Contour &o) {
    front.swap(o.front);
    back.swap(o.back);
    o.reset();
  }
  void reset() {
    front.clear();
    back.clear();
  }
};

// Polygon.addEdge / addEdgesFromContour on already scaled points.
struct Poly {
  StrokePolygon *out;
  // The two orientations differ only in which end is the lower one, so the
  // ends are selected and one expression serves both (the same operands as
  // the reference's two branches); the extents take the same comparisons as
  // selects rather than branches.  w: the next edge's slot; x: the extents
  // (valid once the polygon has an edge).  Returns the slot after the
  // edge's (w itself for a horizontal edge, which is not added).
  static z2d_edge *add_edge(z2d_edge *w, double (&x)[4], bool first, P a, P b) {
    const bool down = a.y < b.y;
    if (!down && !(a.y > b.y)) return w;  // horizontal (or NaN)
    const P lo = down ? a : b, hi = down ? b : a;
    w->y0 = a.y;
    w->y1 = b.y;
    w->x_start = lo.x;
    w->x_inc = (hi.x - lo.x) / (hi.y - lo.y);
    const double top = lo.y, bottom = hi.y;
    const double left = a.x < b.x ? a.x : b.x, right = a.x < b.x ? b.x : a.x;
    if (first) {
      x[0] = top;
      x[1] = bottom;
      x[2] = left;
      x[3] = right;
    } else {
      x[0] = top < x[0] ? top : x[0];
      x[1] = bottom > x[1] ? bottom : x[1];
      x[2] = left < x[2] ? left : x[2];
      x[3] = right > x[3] ? right : x[3];
    }
    return w + 1;
  }
  void add_contour(const Contour &c) {
    const size_t nf = c.front.size(), nb = c.back.size();
    if (nf + nb == 0) return;
    std::vector<z2d_edge> &E = out->edges;
    const size_t n0 = E.size();
    if (E.capacity() < n0 + nf + nb) E.reserve(std::max(n0 + nf + nb, 2 * E.capacity()));
    E.resize(n0 + nf + nb);  // at most one edge per point
    z2d_edge *const w0 = E.data() + n0;
    z2d_edge *w = w0;
    double x[4] = {out->ext[0], out->ext[1], out->ext[2], out->ext[3]};
    // the sequence: front (stored reversed), then back
    const P first = nf ? c.front[nf - 1] : c.back[0];
    P last = first;
    for (size_t i = nf; i-- > 0;) {
      if (i == nf - 1) continue;  // the first point
      w = add_edge(w, x, n0 == 0 && w == w0, last, c.front[i]);
      last = c.front[i];
    }
    for (size_t i = nf ? 0 : 1; i < nb; i++) {
      w = add_edge(w, x, n0 == 0 && w == w0, last, c.back[i]);
      last = c.back[i];
    }
    w = add_edge(w, x, n0 == 0 && w == w0, last, first);
    E.resize(n0 + (size_t)(w - w0));
    for (int k = 0; k < 4; k++) out->ext[k] = x[k];
  }
};

// PointBuffer(2, 5): keeps the first two points and the last three.
struct PtBuf {
  P items[5];
  int len = 0;
  void add(P p) {
    if (len < 5) {
      items[len++] = p;
    } else {
      for (int i = 2; i < 4; i++) items[i] = items[i + 1];
      items[4] = p;
    }
  }
  void reset() { len = 0; }
  P head(int n) const { return items[n]; }
  P tail(int n) const { return items[len - n]; }
  P last() const { return items[len - 1]; }
  P first() const { return items[0]; }
};

struct Opts {
  uint32_t cap, join;
  double miter_limit, thickness, tolerance, scale;
  Xf ctm;
};

// Face (src/internal/tess/Face.zig).
struct Face {
  P p0, p1;
  double width;
  Slope dev_slope, user_slope;
  double half_width;
  P p0_cw, p0_ccw, p1_cw, p1_ccw;
  const Xf *ctm = nullptr;  // the stroke's transformation (outlives every face)

  static Face make(P p0, P p1, Slope dev, double thickness, const Xf &ctm) {
    Face f;
    make_into(f, p0, p1, dev, thickness, ctm);
    return f;
  }
  static void make_into(Face &f, P p0, P p1, Slope dev, double thickness, const Xf &ctm) {
    const double half = thickness / 2;
    double ox, oy;
    Slope user = dev;
    if (!ctm.identity()) {
      double dx = dev.dx, dy = dev.dy;
      ctm.d2u_dist(dx, dy);
      user = Slope{dx, dy};
      user.normalize();
      if (ctm.det() >= 0) {
        ox = -user.dy * half;
        oy = user.dx * half;
      } else {
        ox = user.dy * half;
        oy = -user.dx * half;
      }
      ctm.u2d_dist(ox, oy);
    } else {
      ox = -dev.dy * half;
      oy = dev.dx * half;
    }
    const double cwx = ox, cwy = oy, ccwx = -cwx, ccwy = -cwy;
    f.p0 = p0;
    f.p1 = p1;
    f.width = thickness;
    f.dev_slope = dev;
    f.user_slope = user;
    f.half_width = half;
    f.p0_cw = P{p0.x + cwx, p0.y + cwy};
    f.p0_ccw = P{p0.x + ccwx, p0.y + ccwy};
    f.p1_cw = P{p1.x + cwx, p1.y + cwy};
    f.p1_ccw = P{p1.x + ccwx, p1.y + ccwy};
    f.ctm = &ctm;
  }
  static Face init(P p0, P p1, double thickness, const Xf &ctm) {
    Slope s = Slope::init(p0, p1);
    s.normalize();
    return make(p0, p1, s, thickness, ctm);
  }
  static void init_into(Face &f, P p0, P p1, double thickness, const Xf &ctm) {
    Slope s = Slope::init(p0, p1);
    s.normalize();
    make_into(f, p0, p1, s, thickness, ctm);
  }
  static Face init_single(P pt, Slope dev, double thickness, const Xf &ctm) {
    return make(pt, pt, dev, thickness, ctm);
  }

  P intersect(const Face &out, bool clockwise) const {
    const P ip = clockwise ? p1_ccw : p1_cw;
    const P op = clockwise ? out.p0_ccw : out.p0_cw;
    Slope is = dev_slope, os = out.dev_slope;
    is.normalize();
    os.normalize();
    const double ry = ((op.x - ip.x) * is.dy * os.dy - op.y * os.dx * is.dy +
                       ip.y * is.dx * os.dy) /
                      (is.dx * os.dy - os.dx * is.dy);
    const double rx = std::fabs(is.dy) >= std::fabs(os.dy)
                          ? (ry - ip.y) * is.dx / is.dy + ip.x
                          : (ry - op.y) * os.dx / os.dy + op.x;
    return P{rx, ry};
  }

  template <typename Sink>
  void cap(Sink &&line_to, uint32_t mode, bool clockwise, const Pen *pen) const {
    if (mode == Z2D_CAP_SQUARE) {
      double ox = user_slope.dx * half_width, oy = user_slope.dy * half_width;
      ctm->u2d_dist(ox, oy);
      if (clockwise) {
        line_to(p1_ccw);
        line_to(P{p1_ccw.x + ox, p1_ccw.y + oy});
        line_to(P{p1_cw.x + ox, p1_cw.y + oy});
        line_to(p1_cw);
      } else {
        line_to(p1_cw);
        line_to(P{p1_cw.x + ox, p1_cw.y + oy});
        line_to(P{p1_ccw.x + ox, p1_ccw.y + oy});
        line_to(p1_ccw);
      }
    } else if (mode == Z2D_CAP_ROUND) {
      line_to(clockwise ? p1_ccw : p1_cw);
      pen->walk(dev_slope, Slope{-dev_slope.dx, -dev_slope.dy}, clockwise,
                [&](const PenVertex &pv) { line_to(P{p1.x + pv.point.x, p1.y + pv.point.y}); });
      line_to(clockwise ? p1_cw : p1_ccw);
    } else {
      if (clockwise) {
        line_to(p1_ccw);
        line_to(p1_cw);
      } else {
        line_to(p1_cw);
        line_to(p1_ccw);
      }
    }
  }
  template <typename Sink>
  void cap_p0(Sink &&line_to, uint32_t mode, bool clockwise, const Pen *pen) const {
    Face::init(p1, p0, width, *ctm).cap(line_to, mode, clockwise, pen);
  }
  template <typename Sink>
  void cap_p1(Sink &&line_to, uint32_t mode, bool clockwise, const Pen *pen) const {
    cap(line_to, mode, clockwise, pen);
  }
};

// The state the stroke helpers work on: the plotter's own, or a saved
// initial dash segment's.
struct State {
  const Opts *o = nullptr;
  const Pen *pen = nullptr;
  PtBuf points;
  int cw = -1;  // clockwise_: -1 unset, 0 false, 1 true
  Poly result{};
  Contour outer, inner;
  Slope current_slope;
  // join's two faces; `last` is the slot holding the face the last join
  // built for its outgoing segment, which is the next join's incoming one
  // when segments chain (Face.init is a pure function of its points, the
  // thickness and the transformation)
  Face slot[2];
  int last = -1;
  // scratch for the points plotted ahead of a contour (empty between uses)
  std::vector<P> ahead;

  void reset_contours() {
    outer.reset();
    inner.reset();
    cw = -1;
  }
};

// stroke_plotter.join.  before: the outer points go, in order, ahead of
// the outer contour's current points instead of after them.
void join(State &s, uint32_t mode, P p0, P p1, P p2, bool before) {
  if (peq(p0, p1) || peq(p1, p2)) {
    if (s.cw < 0) s.cw = 0;
    return;
  }
  const Opts &o = *s.o;
  int in_slot = s.last;
  if (in_slot < 0 || !pbits(s.slot[in_slot].p0, p0) || !pbits(s.slot[in_slot].p1, p1)) {
    in_slot = 0;
    Face::init_into(s.slot[0], p0, p1, o.thickness, o.ctm);
  }
  const int out_slot = 1 - in_slot;
  Face::init_into(s.slot[out_slot], p1, p2, o.thickness, o.ctm);
  s.last = out_slot;
  const Face &in = s.slot[in_slot];
  const Face &out = s.slot[out_slot];
  const bool join_cw = in.dev_slope.compare(out.dev_slope) < 0;
  const bool poly_cw = s.cw >= 0 ? s.cw != 0 : join_cw;
  const bool switched = join_cw != poly_cw;
  std::vector<P> &ahead = s.ahead;  // outer points placed ahead (before == true)
  auto plot_outer = [&](P p) {
    if (before) ahead.push_back(s.outer.scaled(p)); else s.outer.plot(p);
  };
  struct Flush {  // prepends the collected points however join returns
    Contour &c;
    std::vector<P> &v;
    ~Flush() {
      if (!v.empty()) c.prepend_scaled(v);
      v.clear();
    }
  } flush{s.outer, ahead};
  auto plot_inner = [&](P p) { s.inner.plot_reverse(p); };
  auto outer_j = [&](P p) { if (switched) plot_inner(p); else plot_outer(p); };
  auto inner_j = [&](P p) { if (switched) plot_outer(p); else plot_inner(p); };

  if (in.dev_slope.compare(out.dev_slope) == 0) {
    outer_j(join_cw ? in.p1_ccw : in.p1_cw);
    inner_j(join_cw ? in.p1_cw : in.p1_ccw);
    if (s.cw < 0) s.cw = poly_cw ? 1 : 0;
    return;
  }
  if (mode == Z2D_JOIN_ROUND) {
    outer_j(join_cw ? in.p1_ccw : in.p1_cw);
    s.pen->walk(in.dev_slope, out.dev_slope, join_cw,
                [&](const PenVertex &pv) { outer_j(P{p1.x + pv.point.x, p1.y + pv.point.y}); });
    outer_j(join_cw ? out.p0_ccw : out.p0_cw);
  } else {
    if (mode == Z2D_JOIN_MITER && miter_within_limit(in.dev_slope, out.dev_slope, o.miter_limit)) {
      outer_j(in.intersect(out, join_cw));
    } else {
      outer_j(join_cw ? in.p1_ccw : in.p1_cw);
      outer_j(join_cw ? out.p0_ccw : out.p0_cw);
    }
  }
  inner_j(join_cw ? in.p1_cw : in.p1_ccw);
  inner_j(p1);
  inner_j(join_cw ? out.p0_cw : out.p0_ccw);
  if (s.cw < 0) s.cw = poly_cw ? 1 : 0;
}

void plot_single(State &s, P start, P end) {
  const Opts &o = *s.o;
  const Face f = Face::init(start, end, o.thickness, o.ctm);
  auto sink = [&](P p) { s.outer.plot(p); };
  f.cap_p0(sink, o.cap, true, s.pen);
  f.cap_p1(sink, o.cap, true, s.pen);
  s.result.add_contour(s.outer);
  s.outer.reset();
  s.cw = -1;
}

void plot_open_joined(State &s, P start0, P end0, P start1, P end1) {
  const Opts &o = *s.o;
  const Face fs = Face::init(start0, end0, o.thickness, o.ctm);
  const Face fe = Face::init(start1, end1, o.thickness, o.ctm);
  const bool clockwise = s.cw >= 0 ? s.cw != 0 : true;
  // the start cap goes ahead of the first corner, in order
  std::vector<P> &ahead = s.ahead;
  fs.cap_p0([&](P p) { ahead.push_back(s.outer.scaled(p)); }, o.cap, clockwise, s.pen);
  s.outer.prepend_scaled(ahead);
  ahead.clear();
  fe.cap_p1([&](P p) { s.outer.plot(p); }, o.cap, clockwise, s.pen);
  s.outer.concat(s.inner);
  s.result.add_contour(s.outer);
  s.reset_contours();
}

void plot_closed_joined(State &s, P initial0, P initial1, P p1, P p2) {
  const uint32_t mode = s.o->join;
  if (!peq(p2, initial0)) {
    join(s, mode, p1, p2, initial0, false);
    join(s, mode, p2, initial0, initial1, false);
  } else {
    join(s, mode, p1, initial0, initial1, false);
  }
  s.result.add_contour(s.outer);
  s.result.add_contour(s.inner);
  s.reset_contours();
}

// Spline.decompose (src/internal/tess/Spline.zig:37-63) with a line_to sink.
struct Knots {
  P a, b, c, d;
};
inline P lerp_half(P a, P b) { return P{a.x + ((b.x - a.x) / 2), a.y + ((b.y - a.y) / 2)}; }
double knots_err_sq(const Knots &k) {
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
template <typename F>
void decompose_into(Knots &s1, P start, double tol_sq, int depth, F &line_to) {
  if (knots_err_sq(s1) < tol_sq || depth > 64) {
    if (!peq(s1.a, start)) line_to(s1.a);
    return;
  }
  const P ab = lerp_half(s1.a, s1.b), bc = lerp_half(s1.b, s1.c), cd = lerp_half(s1.c, s1.d);
  const P abbc = lerp_half(ab, bc), bccd = lerp_half(bc, cd), fin = lerp_half(abbc, bccd);
  Knots s2{fin, bccd, cd, s1.d};
  s1.b = ab;
  s1.c = abbc;
  s1.d = fin;
  decompose_into(s1, start, tol_sq, depth + 1, line_to);
  decompose_into(s2, start, tol_sq, depth + 1, line_to);
}
template <typename F>
void spline(P a, P b, P c, P d, double tolerance, F &&line_to) {
  if (peq(a, b) && peq(c, d)) {
    line_to(d);
    return;
  }
  Knots s1{a, b, c, d};
  decompose_into(s1, a, tolerance * tolerance, 0, line_to);
  line_to(d);
}

// ---------------------------------------------------------------------------
// stroke_plotter.Plotter (src/internal/tess/stroke_plotter.zig:78-246).
// ---------------------------------------------------------------------------
struct Plotter {
  Opts o;
  std::unique_ptr<Pen> pen;
  State s;

  explicit Plotter(const Opts &opts, StrokePolygon *out) : o(opts) {
    if (o.join == Z2D_JOIN_ROUND || o.cap == Z2D_CAP_ROUND)
      pen.reset(new Pen(o.thickness, o.tolerance, o.ctm));
    s.o = &o;
    s.pen = pen.get();
    s.result.out = out;
    s.outer.scale = s.inner.scale = o.scale;
  }

  int32_t line_to(uint32_t join_mode, P p) {
    if (s.points.len == 0) return Z2D_ERR_INVALID_STATE;
    if (peq(p, s.points.last())) return Z2D_OK;  // degenerate
    s.points.add(p);
    if (s.points.len > 2)
      join(s, join_mode, s.points.tail(3), s.points.tail(2), s.points.tail(1), false);
    return Z2D_OK;
  }

  void plot_dotted(P point) {
    if (o.cap != Z2D_CAP_ROUND) return;
    for (const PenVertex &v : pen->v) s.outer.plot(P{point.x + v.point.x, point.y + v.point.y});
    s.result.add_contour(s.outer);
    s.outer.reset();
    s.cw = -1;
  }

  void close_path() {
    switch (s.points.len) {
      case 0: break;
      case 1: plot_dotted(s.points.first()); break;
      case 2: plot_single(s, s.points.head(0), s.points.head(1)); break;
      default:
        plot_closed_joined(s, s.points.head(0), s.points.head(1), s.points.tail(2),
                           s.points.tail(1));
    }
    s.points.reset();
  }

  void finish() {
    if (s.points.len == 2) {
      plot_single(s, s.points.head(0), s.points.head(1));
    } else if (s.points.len > 2) {
      plot_open_joined(s, s.points.head(0), s.points.head(1), s.points.tail(2),
                       s.points.tail(1));
    }
  }

  int32_t run(const z2d_path_node *nodes, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) {
      const z2d_path_node &nd = nodes[i];
      const P p{nd.v[0], nd.v[1]};
      int32_t rc = Z2D_OK;
      switch (nd.tag) {
        case Z2D_NODE_MOVE_TO:
          if (s.points.len > 0) finish();
          s.points.reset();
          s.points.add(p);
          break;
        case Z2D_NODE_LINE_TO: rc = line_to(o.join, p); break;
        case Z2D_NODE_CURVE_TO: {
          if (s.points.len == 0) return Z2D_ERR_INVALID_STATE;
          if (!pen) {
            pen.reset(new Pen(o.thickness, o.tolerance, o.ctm));
            s.pen = pen.get();
          }
          spline(s.points.last(), p, P{nd.v[2], nd.v[3]}, P{nd.v[4], nd.v[5]}, o.tolerance,
                 [&](P q) { line_to(Z2D_JOIN_ROUND, q); });
        } break;
        case Z2D_NODE_CLOSE_PATH: close_path(); break;
        default: return Z2D_ERR_INVALID_ARGUMENT;
      }
      if (rc) return rc;
    }
    finish();
    return Z2D_OK;
  }
};

// Dasher (src/internal/tess/Dasher.zig).
struct Dasher {
  const double *dashes;
  uint64_t n;
  double offset;
  uint64_t idx = 0;
  bool on = true;
  double remain = 0;

  static bool valid(const double *d, uint64_t n) {
    bool ok = false;
    for (uint64_t i = 0; i < n; i++) {
      if (d[i] < 0) return false;
      if (d[i] > 0) ok = true;
    }
    return ok;
  }
  Dasher(const double *d, uint64_t cnt, double off) : dashes(d), n(cnt), offset(off) { reset(); }
  void apply_offset() {
    remain -= offset;
    while (remain < 0 || remain > dashes[idx]) {
      if (remain < 0) {
        remain += dashes[idx];
        idx = idx >= n - 1 ? 0 : idx + 1;
      } else {
        remain -= dashes[idx];
        idx = idx == 0 ? n - 1 : idx - 1;
      }
      on = !on;
    }
  }
  void reset() {
    idx = 0;
    on = true;
    remain = dashes[0];
    apply_offset();
  }
  bool step(double len) {
    bool stepped = false;
    remain -= len;
    if (remain <= 0) {
      stepped = true;
      on = !on;
      idx += 1;
      if (idx >= n) idx = 0;
      remain = dashes[idx];
    }
    return stepped;
  }
};

// ---------------------------------------------------------------------------
// dashed_plotter.Plotter (src/internal/tess/dashed_plotter.zig:70-621).
// ---------------------------------------------------------------------------
struct DashedPlotter {
  enum Init { NONE, OFF, ON };
  Opts o;
  std::unique_ptr<Pen> pen;
  State s;
  Dasher dasher;
  Init init_kind = NONE;
  P init_off{0, 0};
  State init;  // the saved initial segment (init_kind == ON)

  DashedPlotter(const Opts &opts, const double *dashes, uint64_t nd, double dash_offset,
                StrokePolygon *out)
      : o(opts), dasher(dashes, nd, dash_offset) {
    if (o.join == Z2D_JOIN_ROUND || o.cap == Z2D_CAP_ROUND)
      pen.reset(new Pen(o.thickness, o.tolerance, o.ctm));
    s.o = &o;
    s.pen = pen.get();
    s.result.out = out;
    s.outer.scale = s.inner.scale = o.scale;
    init.result.out = out;
    init.outer.scale = init.inner.scale = o.scale;
  }

  int32_t line_to(uint32_t join_mode, P p) {
    if (s.points.len == 0) return Z2D_ERR_INVALID_STATE;
    const P cur = s.points.last();
    if (peq(p, cur)) return Z2D_OK;
    const P first = cur;
    Slope slope = Slope::init(first, p);
    s.current_slope = slope;
    s.current_slope.normalize();
    o.ctm.d2u_dist(slope.dx, slope.dy);
    const double total = slope.normalize();
    double remaining = total;
    double step = std::fmin(dasher.remain, remaining);
    while (remaining > 0) {
      remaining -= step;
      double xo = slope.dx * (total - remaining), yo = slope.dy * (total - remaining);
      o.ctm.u2d_dist(xo, yo);
      const P dp{first.x + xo, first.y + yo};
      if (!peq(dp, s.points.last())) s.points.add(dp);
      if (dasher.on && s.points.len > 2)
        join(s, join_mode, s.points.tail(3), s.points.tail(2), s.points.tail(1), false);
      if (dasher.step(step)) next_segment(dp);
      step = std::fmin(dasher.remain, remaining);
    }
    return Z2D_OK;
  }

  void plot_dotted(P point, Slope slope) {
    if (o.cap == Z2D_CAP_ROUND) {
      for (const PenVertex &v : pen->v) s.outer.plot(P{point.x + v.point.x, point.y + v.point.y});
      s.result.add_contour(s.outer);
    } else if (o.cap == Z2D_CAP_SQUARE) {
      const Face f = Face::init_single(point, slope, o.thickness, o.ctm);
      double ox = f.user_slope.dx * f.half_width, oy = f.user_slope.dy * f.half_width;
      o.ctm.u2d_dist(ox, oy);
      s.outer.plot(P{f.p1_cw.x - ox, f.p1_cw.y - oy});
      s.outer.plot(P{f.p1_cw.x + ox, f.p1_cw.y + oy});
      s.outer.plot(P{f.p1_ccw.x + ox, f.p1_ccw.y + oy});
      s.outer.plot(P{f.p1_ccw.x - ox, f.p1_ccw.y - oy});
      s.result.add_contour(s.outer);
    }
    s.outer.reset();
    s.cw = -1;
  }

  void plot_segment() {
    switch (s.points.len) {
      case 0: break;
      case 1: plot_dotted(s.points.first(), s.current_slope); break;
      case 2: plot_single(s, s.points.head(0), s.points.head(1)); break;
      default:
        plot_open_joined(s, s.points.head(0), s.points.head(1), s.points.tail(2),
                         s.points.tail(1));
    }
  }

  void save_initial() {
    if (!dasher.on) {
      init_kind = ON;
      init.o = &o;
      init.pen = s.pen;
      init.points = s.points;
      init.cw = s.cw;
      init.outer.take(s.outer);
      init.inner.take(s.inner);
      init.current_slope = s.current_slope;
    } else {
      init_kind = OFF;
      init_off = s.points.first();
    }
    s.outer.reset();
    s.inner.reset();
    s.cw = -1;
  }

  void next_segment(P p) {
    if (init_kind == NONE)
      save_initial();
    else if (!dasher.on)
      plot_segment();
    s.points.reset();
    s.points.add(p);
  }

  int32_t finish_initial_dotted() {
    if (init_kind != ON) return Z2D_ERR_INVALID_STATE;
    plot_dotted(init.points.first(), init.current_slope);
    init_kind = NONE;
    return Z2D_OK;
  }

  int32_t finish_initial(P last_point, P second_to_last) {
    if (init_kind != ON) return Z2D_ERR_INVALID_STATE;
    if (init.points.len < 2) return Z2D_ERR_INVALID_STATE;
    plot_open_joined(init, last_point, second_to_last, init.points.tail(2), init.points.tail(1));
    init_kind = NONE;
    return Z2D_OK;
  }

  int32_t join_and_cap_initial() {
    if (init_kind != ON) return Z2D_ERR_INVALID_STATE;
    if (init.points.len < 2) return Z2D_ERR_INVALID_STATE;
    if (s.points.len > 2) {
      join(s, o.join, s.points.tail(2), init.points.head(0), init.points.head(1), false);
      s.outer.concat(init.outer);
      init.inner.concat(s.inner);
      init.outer.take(s.outer);  // initial outer = combined
      plot_open_joined(init, s.points.head(0), s.points.head(1), init.points.tail(2),
                       init.points.tail(1));
    } else {
      join(init, o.join, s.points.tail(2), init.points.head(0), init.points.head(1), true);
      plot_open_joined(init, s.points.first(), init.points.first(), init.points.tail(2),
                       init.points.tail(1));
    }
    init_kind = NONE;
    s.outer.reset();
    s.inner.reset();
    s.cw = -1;
    return Z2D_OK;
  }

  int32_t close_path() {
    if (s.points.len <= 0) return Z2D_ERR_INVALID_STATE;
    P target;
    if (init_kind == ON) {
      if (init.points.len == 0) return Z2D_ERR_INVALID_STATE;
      target = init.points.first();
    } else if (init_kind == OFF) {
      target = init_off;
    } else {
      target = s.points.first();
    }
    int32_t rc = line_to(o.join, target);
    if (rc) return rc;
    if (init_kind == ON) {
      if (dasher.on && s.points.len > 1) {
        if (init.points.len == 1) {
          plot_open_joined(s, s.points.head(0), s.points.head(1), s.points.tail(2),
                           s.points.tail(1));
          init_kind = NONE;
        } else {
          rc = join_and_cap_initial();
        }
      } else {
        switch (init.points.len) {
          case 0: return Z2D_ERR_INVALID_STATE;
          case 1: rc = finish_initial_dotted(); break;
          default: rc = finish_initial(init.points.head(0), init.points.head(1));
        }
      }
    } else if (init_kind == OFF) {
      init_kind = NONE;
    } else {
      switch (s.points.len) {
        case 1: plot_dotted(s.points.first(), s.current_slope); break;
        case 2: plot_single(s, s.points.head(0), s.points.head(1)); break;
        default:
          join(s, o.join, s.points.tail(2), s.points.head(0), s.points.head(1), false);
          s.result.add_contour(s.outer);
          s.result.add_contour(s.inner);
          s.reset_contours();
      }
    }
    s.points.reset();
    return rc;
  }

  int32_t finish() {
    int32_t rc = Z2D_OK;
    if (init_kind == ON) {
      switch (init.points.len) {
        case 0: return Z2D_ERR_INVALID_STATE;
        case 1: rc = finish_initial_dotted(); break;
        default: rc = finish_initial(init.points.head(0), init.points.head(1));
      }
    } else if (init_kind == OFF) {
      init_kind = NONE;
    }
    if (rc) return rc;
    if (dasher.on) plot_segment();
    return Z2D_OK;
  }

  int32_t run(const z2d_path_node *nodes, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) {
      const z2d_path_node &nd = nodes[i];
      const P p{nd.v[0], nd.v[1]};
      int32_t rc = Z2D_OK;
      switch (nd.tag) {
        case Z2D_NODE_MOVE_TO:
          rc = finish();
          dasher.reset();
          s.points.reset();
          s.points.add(p);
          break;
        case Z2D_NODE_LINE_TO: rc = line_to(o.join, p); break;
        case Z2D_NODE_CURVE_TO: {
          if (s.points.len == 0) return Z2D_ERR_INVALID_STATE;
          if (!pen) {
            pen.reset(new Pen(o.thickness, o.tolerance, o.ctm));
            s.pen = pen.get();
          }
          int32_t lrc = Z2D_OK;
          spline(s.points.last(), p, P{nd.v[2], nd.v[3]}, P{nd.v[4], nd.v[5]}, o.tolerance,
                 [&](P q) {
                   if (lrc == Z2D_OK) lrc = line_to(Z2D_JOIN_ROUND, q);
                 });
          rc = lrc;
        } break;
        case Z2D_NODE_CLOSE_PATH: rc = close_path(); break;
        default: return Z2D_ERR_INVALID_ARGUMENT;
      }
      if (rc) return rc;
    }
    return finish();
  }
};

}  // namespace

int32_t stroke_tessellate(const z2d_path_node *nodes, uint64_t n, const StrokeParams &p,
                          StrokePolygon &out) {
  Opts o;
  o.cap = p.cap_mode;
  o.join = p.join_mode;
  o.miter_limit = p.miter_limit;
  o.thickness = p.thickness;
  o.tolerance = p.tolerance;
  o.scale = p.scale;
  o.ctm.ax = p.ctm[0];
  o.ctm.by = p.ctm[1];
  o.ctm.cx = p.ctm[2];
  o.ctm.dy = p.ctm[3];
  o.ctm.tx = p.ctm[4];
  o.ctm.ty = p.ctm[5];
  out.edges.clear();
  out.ext[0] = out.ext[1] = out.ext[2] = out.ext[3] = 0.0;
  if (p.dashes && p.n_dashes && Dasher::valid(p.dashes, p.n_dashes)) {
    DashedPlotter dp(o, p.dashes, p.n_dashes, p.dash_offset, &out);
    return dp.run(nodes, n);
  }
  Plotter pl(o, &out);
  return pl.run(nodes, n);
}

}  // namespace z2d
