// SPDX-License-Identifier: MPL-2.0
// Restates code of z2d (Copyright (c) 2024-2026 Chris Marchesi, MPL-2.0).
// The Mozilla Public License v2.0 applies to this file:
// https://mozilla.org/MPL/2.0/.
//
// text.cpp - text for the device fill path: TrueType outlines (src/Font.zig,
// src/internal/Glyph.zig) laid out by text.show (src/text.zig:73-196) into
// one path, whose nodes go to the fill entry points (one fill per call, as
// the reference does; batches take many of them in one launch).
//
// Host code: parsing and layout are byte reads and a few f64 operations per
// point; the device renders the result.  Every read is bounds-checked (the
// reference's fixed reader returns EndOfStream), and the arithmetic follows
// the reference's order (f64, no contraction) so the nodes are the
// reference's bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "../../include/z2d_gpu.h"

namespace {

// ---------------------------------------------------------------------------
// Transformation (src/Transformation.zig)
// ---------------------------------------------------------------------------
struct Xf {
  double ax = 1, by = 0, cx = 0, dy = 1, tx = 0, ty = 0;
};
Xf xf_mul(const Xf &a, const Xf &b) {
  Xf r;
  r.ax = a.ax * b.ax + a.by * b.cx;
  r.by = a.ax * b.by + a.by * b.dy;
  r.cx = a.cx * b.ax + a.dy * b.cx;
  r.dy = a.cx * b.by + a.dy * b.dy;
  r.ty = a.cx * b.tx + a.dy * b.ty + a.ty;
  r.tx = a.ax * b.tx + a.by * b.ty + a.tx;
  return r;
}
Xf xf_translate(const Xf &a, double tx, double ty) {
  Xf b;
  b.tx = tx;
  b.ty = ty;
  return xf_mul(a, b);
}
Xf xf_scale(const Xf &a, double sx, double sy) {
  Xf b;
  b.ax = sx;
  b.dy = sy;
  return xf_mul(a, b);
}
bool xf_inverse(const Xf &a, Xf &o) {
  if (a.by == 0 && a.cx == 0) {
    if (a.ax == 0 || a.dy == 0) return false;
    if (a.ax != 1 || a.dy != 1) {
      o = Xf{1 / a.ax, 0, 0, 1 / a.dy, -a.tx / a.ax, -a.ty / a.dy};
    } else {
      o = Xf{1, 0, 0, 1, -a.tx, -a.ty};
    }
    return true;
  }
  const double det = a.ax * a.dy - a.by * a.cx;
  if (det == 0) return false;
  const Xf adj{a.dy, -a.by, -a.cx, a.ax, a.by * a.ty - a.dy * a.tx, a.cx * a.tx - a.ax * a.ty};
  const double s = 1 / det;
  o = Xf{adj.ax * s, adj.by * s, adj.cx * s, adj.dy * s, adj.tx * s, adj.ty * s};
  return true;
}
void xf_apply(const Xf &a, double &x, double &y) {
  const double ix = x, iy = y;
  x = a.ax * ix + a.by * iy;
  y = a.cx * ix + a.dy * iy;
  x += a.tx;
  y += a.ty;
}

// ---------------------------------------------------------------------------
// Path (src/Path.zig:116-262, :447-477): points clamped to i24 and
// transformed on entry; a move equal to a preceding move is dropped; a close
// appends the close and a move back to the initial point.
// ---------------------------------------------------------------------------
double clamp_i24(double v) { return std::fmax(-8388608.0, std::fmin(v, 8388607.0)); }

struct PathB {
  std::vector<z2d_path_node> nodes;
  Xf t;
  bool has_cur = false;
  double cur_x = 0, cur_y = 0, init_x = 0, init_y = 0;

  static z2d_path_node node(uint32_t tag, double x, double y) {
    z2d_path_node n;
    std::memset(&n, 0, sizeof(n));
    n.tag = tag;
    n.v[0] = x;
    n.v[1] = y;
    return n;
  }
  void move_to(double x, double y) {
    x = clamp_i24(x);
    y = clamp_i24(y);
    xf_apply(t, x, y);
    if (!nodes.empty() && nodes.back().tag == Z2D_NODE_MOVE_TO && nodes.back().v[0] == x &&
        nodes.back().v[1] == y)
      return;
    nodes.push_back(node(Z2D_NODE_MOVE_TO, x, y));
    init_x = cur_x = x;
    init_y = cur_y = y;
    has_cur = true;
  }
  void line_to(double x, double y) {
    if (!has_cur) return move_to(x, y);
    x = clamp_i24(x);
    y = clamp_i24(y);
    xf_apply(t, x, y);
    nodes.push_back(node(Z2D_NODE_LINE_TO, x, y));
    cur_x = x;
    cur_y = y;
  }
  bool curve_to(double x1, double y1, double x2, double y2, double x3, double y3) {
    if (!has_cur) return false;  // NoCurrentPoint
    double p[6] = {clamp_i24(x1), clamp_i24(y1), clamp_i24(x2),
                   clamp_i24(y2), clamp_i24(x3), clamp_i24(y3)};
    for (int k = 0; k < 6; k += 2) xf_apply(t, p[k], p[k + 1]);
    z2d_path_node n;
    std::memset(&n, 0, sizeof(n));
    n.tag = Z2D_NODE_CURVE_TO;
    for (int k = 0; k < 6; k++) n.v[k] = p[k];
    nodes.push_back(n);
    cur_x = p[4];
    cur_y = p[5];
    return true;
  }
  void close() {
    if (!has_cur) return;
    nodes.push_back(node(Z2D_NODE_CLOSE_PATH, 0, 0));
    nodes.push_back(node(Z2D_NODE_MOVE_TO, init_x, init_y));
  }
};

// ---------------------------------------------------------------------------
// Big-endian reads over the font bytes
// ---------------------------------------------------------------------------
struct Rd {
  const uint8_t *p = nullptr;
  uint64_t n = 0;
  bool ok = true;
  uint32_t byte(uint64_t off) {
    if (off >= n) {
      ok = false;
      return 0;
    }
    return p[off];
  }
  uint32_t u16(uint64_t off) { return (byte(off) << 8) | byte(off + 1); }
  int32_t i16(uint64_t off) { return (int16_t)(uint16_t)u16(off); }
  uint32_t u32(uint64_t off) { return (u16(off) << 16) | u16(off + 2); }
};

}  // namespace

// ---------------------------------------------------------------------------
// Font (src/Font.zig:28-45 loadBuffer, :84-171 Directory, :173-260 Meta)
// ---------------------------------------------------------------------------
struct GlyphRec {  // Glyph (src/internal/Glyph.zig): index, metrics, outline offset
  uint32_t index = 0;
  uint32_t advance = 0;
  int32_t lsb = 0;
  bool has_outline = false;
  uint32_t offset = 0;
};
struct OutlineRec {  // Glyph.Outline: nodes in glyph space (y flipped), x_min
  std::vector<z2d_path_node> nodes;
  int32_t x_min = 0;
};

struct z2d_font {
  std::vector<uint8_t> data;
  uint32_t cmap = 0, glyf = 0, head = 0, hhea = 0, hmtx = 0, loca = 0, kern = 0, gpos = 0;
  bool cmap_full = false;
  uint32_t cmap_sub = 0;
  bool lsb_is_at_x_zero = false;
  bool loca_long = false;
  uint32_t advance_width_max = 0, number_of_hmetrics = 0, units_per_em = 0;
  std::mutex mu;  // the glyph and outline caches
  std::unordered_map<uint32_t, GlyphRec> glyphs;
  std::unordered_map<uint32_t, OutlineRec> outlines;
  Rd rd() const {
    Rd r;
    r.p = data.data();
    r.n = data.size();
    return r;
  }
};

namespace {

int32_t font_parse(z2d_font &f) {
  Rd r = f.rd();
  const uint8_t *d = f.data.data();
  if (f.data.size() < 4) return Z2D_ERR_INVALID_FONT;  // EndOfStream
  const bool magic = (std::memcmp(d, "1\0\0\0", 4) == 0) || (std::memcmp(d, "OTTO", 4) == 0) ||
                     (std::memcmp(d, "\0\1\0\0", 4) == 0) || (std::memcmp(d, "true", 4) == 0);
  if (!magic) return Z2D_ERR_INVALID_FONT;  // InvalidFormat
  const uint32_t num_tables = r.u16(4);
  struct T {
    const char *tag;
    uint32_t *dst;
  } tables[] = {{"cmap", &f.cmap}, {"glyf", &f.glyf}, {"head", &f.head}, {"hhea", &f.hhea},
                {"hmtx", &f.hmtx}, {"loca", &f.loca}, {"kern", &f.kern}, {"GPOS", &f.gpos}};
  for (uint32_t i = 0; i < num_tables && r.ok; i++) {
    const uint64_t e = 12 + 16ull * i;
    uint8_t tag[4];
    for (int k = 0; k < 4; k++) tag[k] = (uint8_t)r.byte(e + k);
    for (T &t : tables) {
      if (std::memcmp(tag, t.tag, 4) != 0) continue;
      const uint32_t checksum = r.u32(e + 4), offset = r.u32(e + 8), len = r.u32(e + 12);
      uint32_t actual = 0;
      const bool is_head = std::memcmp(t.tag, "head", 4) == 0;
      for (uint64_t j = 0; j < ((uint64_t)len + 3) / 4; j++) {
        const uint32_t w = r.u32((uint64_t)offset + 4 * j);
        if (!(is_head && j == 2)) actual += w;
      }
      if (!r.ok) return Z2D_ERR_INVALID_FONT;
      if (checksum != actual) return Z2D_ERR_INVALID_FONT;  // ChecksumMismatch
      *t.dst = offset;
    }
  }
  if (!r.ok) return Z2D_ERR_INVALID_FONT;
  if (!f.cmap || !f.glyf || !f.head || !f.hhea || !f.hmtx || !f.loca)
    return Z2D_ERR_INVALID_FONT;  // MissingRequiredTable
  // cmap subtable: Unicode (0) encodings 3 / 4 or Windows (3) encodings 1 / 10
  uint32_t bmp = 0, full = 0;
  const uint32_t n_sub = r.u16((uint64_t)f.cmap + 2);
  for (uint32_t i = 0; i < n_sub && r.ok; i++) {
    const uint64_t e = (uint64_t)f.cmap + 4 + 8ull * i;
    const uint32_t platform = r.u16(e), encoding = r.u16(e + 2);
    const uint32_t off = r.u32(e + 4) + f.cmap;
    if (platform == 0) {
      if (encoding == 3) bmp = off;
      else if (encoding == 4) full = off;
    } else if (platform == 3) {
      if (encoding == 1) bmp = off;
      else if (encoding == 10) full = off;
    }
  }
  if (!r.ok) return Z2D_ERR_INVALID_FONT;
  if (full) {
    f.cmap_full = true;
    f.cmap_sub = full;
  } else if (bmp) {
    f.cmap_sub = bmp;
  } else {
    return Z2D_ERR_INVALID_FONT;  // NoSuitableCmapSubtable
  }
  // head flags & (2 >> 1): the flag word's bit 0, as the reference reads it
  f.lsb_is_at_x_zero = (r.u16((uint64_t)f.head + 14) & 1u) != 0;
  const uint32_t itl = r.u16((uint64_t)f.head + 50);
  f.units_per_em = r.u16((uint64_t)f.head + 18);
  f.advance_width_max = r.u16((uint64_t)f.hhea + 10);
  f.number_of_hmetrics = r.u16((uint64_t)f.hhea + 34);
  if (!r.ok) return Z2D_ERR_INVALID_FONT;
  if (itl > 1) return Z2D_ERR_INVALID_FONT;  // InvalidIndexToLocFormat
  f.loca_long = itl == 1;
  return Z2D_OK;
}

// Glyph.findGlyphIndexBMP (src/internal/Glyph.zig): cmap format 4 by the
// reference's binary search over the end codes.
uint32_t glyph_index_bmp(z2d_font &f, Rd &r, uint32_t cp) {
  if (cp > 0xffff) return 0;
  const uint64_t o = f.cmap_sub;
  const uint32_t seg = r.u16(o + 6) >> 1;
  uint32_t search_range = r.u16(o + 8) >> 1;
  uint32_t entry_selector = r.u16(o + 10);
  const uint32_t range_shift = r.u16(o + 12) >> 1;
  const uint64_t eco = o + 14;
  uint64_t search = eco;
  if (cp >= r.u16(search + range_shift * 2)) search += range_shift * 2;
  search -= 2;
  while (entry_selector > 0 && r.ok) {
    search_range >>= 1;
    const uint32_t end = r.u16(search + search_range * 2);
    if (cp > end) search += search_range * 2;
    entry_selector--;
  }
  search += 2;
  const uint64_t item = (search - eco) >> 1;
  const uint32_t start = r.u16(eco + seg * 2ull + 2 + 2 * item);
  const uint32_t last = r.u16(eco + 2 * item);
  if (cp < start || cp > last) return 0;
  const uint32_t offset = r.u16(eco + seg * 6ull + 2 + 2 * item);
  if (offset == 0) {
    const int32_t delta = r.i16(eco + seg * 4ull + 2 + 2 * item);
    return (uint32_t)((int32_t)cp + delta);
  }
  return r.u16(offset + (uint64_t)(cp - start) * 2 + eco + seg * 6ull + 2 + 2 * item);
}

// Glyph.findGlyphIndexFull: cmap format 12 groups by binary search.
uint32_t glyph_index_full(z2d_font &f, Rd &r, uint32_t cp) {
  const uint64_t o = f.cmap_sub;
  const uint32_t n = r.u32(o + 12);
  int64_t lo = 0, hi = (int64_t)n;
  while (lo < hi && r.ok) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    const uint64_t m = o + 16 + (uint64_t)mid * 12;
    const uint32_t start = r.u32(m), end = r.u32(m + 4);
    if (cp < start) hi = mid;
    else if (cp > end) lo = mid + 1;
    else return r.u32(m + 8) + cp - start;
  }
  return 0;
}

// Glyph.byIndex: metrics from hmtx and the outline range from loca.
bool glyph_by_index(z2d_font &f, Rd &r, uint32_t index, GlyphRec &g) {
  g.index = index;
  const uint64_t nh = f.number_of_hmetrics;
  if (index < nh) {
    g.advance = r.u16(f.hmtx + (uint64_t)index * 4);
    g.lsb = r.i16(f.hmtx + (uint64_t)index * 4 + 2);
  } else {
    g.advance = r.u16(f.hmtx + (nh - 1) * 4);
    g.lsb = r.i16(f.hmtx + nh * 4 + ((uint64_t)index - nh) * 2);
  }
  uint32_t off[2];
  for (int i = 0; i < 2; i++) {
    if (f.loca_long) off[i] = f.glyf + r.u32(f.loca + ((uint64_t)index + i) * 4);
    else off[i] = f.glyf + r.u16(f.loca + ((uint64_t)index + i) * 2) * 2;
  }
  g.has_outline = off[0] != off[1];
  g.offset = off[0];
  return r.ok;
}

bool glyph_init(z2d_font &f, uint32_t cp, GlyphRec &g) {
  auto it = f.glyphs.find(cp);
  if (it != f.glyphs.end()) {
    g = it->second;
    return true;
  }
  Rd r = f.rd();
  const uint32_t index = f.cmap_full ? glyph_index_full(f, r, cp) : glyph_index_bmp(f, r, cp);
  if (!r.ok || !glyph_by_index(f, r, index, g)) return false;
  f.glyphs[cp] = g;
  return true;
}

// ---- kerning (Glyph.getKernAdvance: GPOS pair adjustment, else kern) -----
int32_t coverage_index(Rd &r, uint64_t t, uint32_t glyph) {
  const uint32_t format = r.u16(t);
  if (format == 1) {
    const int32_t count = (int32_t)r.u16(t + 2);
    int32_t l = 0, h = count - 1;
    while (l <= h && r.ok) {
      const int32_t m = (l + h) >> 1;
      const uint32_t straw = r.u16(t + 4 + 2ull * (uint32_t)m);
      if ((int32_t)glyph < (int32_t)straw) h = m - 1;
      else if ((int32_t)glyph > (int32_t)straw) l = m + 1;
      else return m;
    }
  } else if (format == 2) {
    const int32_t count = (int32_t)r.u16(t + 2);
    int32_t l = 0, h = count - 1;
    while (l <= h && r.ok) {
      const int32_t m = (l + h) >> 1;
      const uint64_t rec = t + 4 + 6ull * (uint32_t)m;
      const uint32_t s = r.u16(rec), e = r.u16(rec + 2);
      if ((int32_t)glyph < (int32_t)s) h = m - 1;
      else if ((int32_t)glyph > (int32_t)e) l = m + 1;
      else return (int32_t)r.u16(rec + 4) + (int32_t)glyph - (int32_t)s;
    }
  } else {
    return -1;
  }
  return -1;
}

int32_t glyph_class(Rd &r, uint64_t t, uint32_t glyph) {
  const uint32_t format = r.u16(t);
  if (format == 1) {
    const uint32_t start = r.u16(t + 2), count = r.u16(t + 4);
    if (glyph >= start && glyph < start + count) return (int32_t)r.u16(t + 6 + 2ull * (glyph - start));
  } else if (format == 2) {
    const int32_t count = (int32_t)r.u16(t + 2);
    int32_t l = 0, h = count - 1;
    while (l <= h && r.ok) {
      const int32_t m = (l + h) >> 1;
      const uint64_t rec = t + 4 + 6ull * (uint32_t)m;
      const uint32_t s = r.u16(rec), e = r.u16(rec + 2);
      if (glyph < s) h = m - 1;
      else if (glyph > e) l = m + 1;
      else return (int32_t)r.u16(rec + 4);
    }
  } else {
    return -1;
  }
  return -1;
}

int32_t kern_gpos(z2d_font &f, Rd &r, uint32_t cur, uint32_t next) {
  const uint64_t G = f.gpos;
  if (r.u16(G) != 1 || r.u16(G + 2) != 0) return 0;
  const uint64_t lookup_list = G + r.u16(G + 8);
  const uint32_t lookup_count = r.u16(lookup_list);
  for (uint32_t i = 0; i < lookup_count && r.ok; i++) {
    const uint64_t lt = lookup_list + r.u16(lookup_list + 2 + 2ull * i);
    const uint32_t type = r.u16(lt);
    const uint32_t n_sub = r.u16(lt + 4);
    if (type != 2 && type != 9) continue;
    for (uint32_t s = 0; s < n_sub && r.ok; s++) {
      const uint32_t so = r.u16(lt + 6 + 2ull * s);
      uint64_t table;
      if (type == 2) {
        table = lt + so;
      } else {
        if (r.u16(lt + so + 2) != 2) break;  // extension of another lookup type
        table = lt + so + r.u32(lt + so + 4);
      }
      const uint32_t pos_format = r.u16(table);
      const uint32_t cov_off = r.u16(table + 2);
      const int32_t ci = coverage_index(r, table + cov_off, cur);
      if (ci == -1) continue;
      if (pos_format == 1) {
        const uint32_t vf1 = r.u16(table + 4), vf2 = r.u16(table + 6);
        if (!(vf1 == 4 && vf2 == 0)) return 0;
        const uint32_t pair_set_count = r.u16(table + 8);
        const uint64_t pv = table + r.u16(table + 10 + 2ull * (uint32_t)ci);
        const int32_t pair_count = (int32_t)r.u16(pv);
        if (ci >= (int32_t)pair_set_count) return 0;
        int32_t l = 0, h = pair_count - 1;
        while (l <= h && r.ok) {
          const int32_t m = (l + h) >> 1;
          const uint64_t rec = pv + 2 + 4ull * (uint32_t)m;
          const uint32_t straw = r.u16(rec);
          if (next < straw) h = m - 1;
          else if (next > straw) l = m + 1;
          else return r.i16(rec + 2);
        }
      } else if (pos_format == 2) {
        const uint32_t vf1 = r.u16(table + 4), vf2 = r.u16(table + 6);
        if (!(vf1 == 4 && vf2 == 0)) return 0;
        const uint32_t cd1 = r.u16(table + 8), cd2 = r.u16(table + 10);
        const int32_t c1 = glyph_class(r, table + cd1, cur);
        const int32_t c2 = glyph_class(r, table + cd2, next);
        const int32_t n1 = (int32_t)r.u16(table + 12), n2 = (int32_t)r.u16(table + 14);
        if (c1 < 0 || c1 >= n1) return 0;
        if (c2 < 0 || c2 >= n2) return 0;
        const uint64_t c2rec = table + 16 + 2ull * ((uint64_t)c1 * (uint64_t)n2);
        return r.i16(c2rec + 2ull * (uint32_t)c2);
      } else {
        return 0;
      }
    }
  }
  return 0;
}

int32_t kern_table(z2d_font &f, Rd &r, uint32_t cur, uint32_t next) {
  const uint64_t K = f.kern;
  const uint32_t num_tables = r.u16(K + 2);
  const uint32_t format = r.u16(K + 8);
  if (num_tables < 1) return 0;
  if (format != 1) return 0;
  int32_t l = 0, h = (int32_t)r.u16(K + 10) - 1;
  const uint32_t needle = cur << 16 | next;
  while (l <= h && r.ok) {
    const int32_t m = (l + h) >> 1;
    const uint32_t straw = r.u32(K + 18 + 6ull * (uint32_t)m);
    if (needle < straw) h = m - 1;
    else if (needle > straw) l = m + 1;
    else return r.i16(K + 18 + 6ull * (uint32_t)m + 4);
  }
  return 0;
}

bool kern_advance(z2d_font &f, uint32_t cur, uint32_t next, int32_t &out) {
  Rd r = f.rd();
  out = f.gpos ? kern_gpos(f, r, cur, next) : f.kern ? kern_table(f, r, cur, next) : 0;
  return r.ok;
}

// ---- outlines (Glyph.Outline.init / runOutline) ---------------------------
double f2dot14(uint32_t x) {
  const int32_t hi = (int32_t)(x >> 14) >= 2 ? (int32_t)(x >> 14) - 4 : (int32_t)(x >> 14);
  return (double)hi + (double)(x & 0x3FFF) / 16384.0;
}

bool quad_to(PathB &p, int32_t ctl_x, int32_t ctl_y, int32_t to_x, int32_t to_y) {
  if (!p.has_cur) return false;  // MalformedGlyph
  double x0 = p.cur_x, y0 = p.cur_y;
  Xf inv;
  if (!xf_inverse(p.t, inv)) return false;
  xf_apply(inv, x0, y0);
  const double x3 = to_x, y3 = to_y;
  const double x1 = x0 + 2.0 / 3.0 * ((double)ctl_x - x0);
  const double y1 = y0 + 2.0 / 3.0 * ((double)ctl_y - y0);
  const double x2 = x3 + 2.0 / 3.0 * ((double)ctl_x - x3);
  const double y2 = y3 + 2.0 / 3.0 * ((double)ctl_y - y3);
  return p.curve_to(x1, y1, x2, y2, x3, y3);
}

bool run_outline(z2d_font &f, PathB &p, uint32_t glyf_off, int depth) {
  if (depth > 32) return false;  // a composite that refers to itself
  Rd r = f.rd();
  const int32_t n_contours = r.i16(glyf_off);
  uint64_t pos = (uint64_t)glyf_off + 10;
  if (n_contours < 0) {
    for (;;) {
      const uint32_t flags = r.u16(pos);
      const uint32_t index = r.u16(pos + 2);
      pos += 4;
      int32_t xo, yo;
      if (flags & 0x0001) {  // args are words
        xo = r.i16(pos);
        yo = r.i16(pos + 2);
        pos += 4;
      } else {
        xo = (int8_t)r.byte(pos);
        yo = (int8_t)r.byte(pos + 1);
        pos += 2;
      }
      if (!r.ok) return false;
      const Xf saved = p.t;
      const bool scaled_offset = (flags & 0x0800) != 0;
      if (!scaled_offset) p.t = xf_translate(p.t, xo, yo);
      if (flags & 0x0008) {  // a scale
        const double s = f2dot14(r.u16(pos));
        pos += 2;
        p.t = xf_scale(p.t, s, s);
      } else if (flags & 0x0040) {  // an x and y scale
        const double sx = f2dot14(r.u16(pos)), sy = f2dot14(r.u16(pos + 2));
        pos += 4;
        p.t = xf_scale(p.t, sx, sy);
      } else if (flags & 0x0080) {  // a two by two
        Xf m;
        m.ax = f2dot14(r.u16(pos));
        m.cx = f2dot14(r.u16(pos + 2));
        m.by = f2dot14(r.u16(pos + 4));
        m.dy = f2dot14(r.u16(pos + 6));
        pos += 8;
        p.t = xf_mul(p.t, m);
      }
      if (scaled_offset) p.t = xf_translate(p.t, xo, yo);
      if (!r.ok) return false;
      GlyphRec g;
      if (!glyph_by_index(f, r, index, g)) return false;
      if (g.has_outline && !run_outline(f, p, g.offset, depth + 1)) return false;
      p.t = saved;
      if (!(flags & 0x0020)) return true;  // no more components
    }
  }
  std::vector<uint8_t> is_end;
  std::vector<uint32_t> ends;
  uint32_t len = 0;
  for (int32_t c = 0; c < n_contours; c++) {
    const uint32_t e = r.u16(pos);
    pos += 2;
    ends.push_back(e);
    len = (e + 1) & 0xFFFFu;
  }
  is_end.assign(len + 1, 0);
  for (uint32_t e : ends)
    if (e < len) is_end[e] = 1;
  pos += 2 + (uint64_t)r.u16(pos);  // instructions
  std::vector<uint8_t> flags;
  flags.reserve(len);
  for (uint32_t i = 0; i < len && r.ok;) {
    const uint8_t fl = (uint8_t)r.byte(pos++);
    flags.push_back(fl);
    if (fl & 0x08) {  // repeat
      const uint32_t cnt = r.byte(pos++);
      for (uint32_t k = 0; k < cnt; k++) {
        flags.push_back(fl);
        i++;
      }
    }
    i++;
  }
  if (!r.ok || flags.size() < len) return false;
  std::vector<int32_t> px(len), py(len);
  int16_t cur = 0;
  for (uint32_t i = 0; i < len; i++) {
    const uint8_t fl = flags[i];
    if (fl & 0x02) {
      if (fl & 0x10) cur = (int16_t)(cur + (int32_t)r.byte(pos++));
      else cur = (int16_t)(cur - (int32_t)r.byte(pos++));
    } else if (!(fl & 0x10)) {
      cur = (int16_t)(cur + r.i16(pos));
      pos += 2;
    }
    px[i] = cur;
  }
  cur = 0;
  for (uint32_t i = 0; i < len; i++) {
    const uint8_t fl = flags[i];
    if (fl & 0x04) {
      if (fl & 0x20) cur = (int16_t)(cur + (int32_t)r.byte(pos++));
      else cur = (int16_t)(cur - (int32_t)r.byte(pos++));
    } else if (!(fl & 0x20)) {
      cur = (int16_t)(cur + r.i16(pos));
      pos += 2;
    }
    py[i] = cur;
  }
  if (!r.ok) return false;
  enum { MOVE, ON, OFF } state = MOVE;
  uint32_t j = 0;
  for (uint32_t i = 0; i < len; i++) {
    const bool on = flags[i] & 0x01;
    if (state == MOVE) {
      if (!on || is_end[i]) continue;  // skips the contour-end handling too
      p.move_to(px[i], py[i]);
      j = i;
      state = ON;
    } else if (state == ON) {
      if (!on) state = OFF;
      else p.line_to(px[i], py[i]);
    } else {
      if (on) {
        if (!quad_to(p, px[i - 1], py[i - 1], px[i], py[i])) return false;
        state = ON;
      } else {
        const int32_t ex = (px[i] + px[i - 1]) >> 1, ey = (py[i] + py[i - 1]) >> 1;
        if (!quad_to(p, px[i - 1], py[i - 1], ex, ey)) return false;
      }
    }
    if (is_end[i]) {
      if (state == OFF && !quad_to(p, px[i], py[i], px[j], py[j])) return false;
      p.close();
      state = MOVE;
    }
  }
  return true;
}

bool outline_get(z2d_font &f, uint32_t cp, const GlyphRec &g, const OutlineRec *&o) {
  auto it = f.outlines.find(cp);
  if (it != f.outlines.end()) {
    o = &it->second;
    return true;
  }
  Rd r = f.rd();
  OutlineRec rec;
  rec.x_min = r.i16(g.offset + 2);
  if (!r.ok) return false;
  PathB p;
  p.t = xf_translate(xf_scale(Xf{}, 1.0, -1.0), 0.0, (double)f.units_per_em * -1.0);
  if (!run_outline(f, p, g.offset, 0)) return false;
  rec.nodes = std::move(p.nodes);
  o = &(f.outlines[cp] = std::move(rec));
  return true;
}

// std.unicode decoding as text.show uses it; false on any invalid sequence.
bool utf8_decode(const uint8_t *s, uint64_t n, std::vector<uint32_t> &out) {
  uint64_t i = 0;
  while (i < n) {
    const uint8_t b = s[i];
    uint32_t len;
    if (b < 0x80) len = 1;
    else if ((b & 0xE0) == 0xC0) len = 2;
    else if ((b & 0xF0) == 0xE0) len = 3;
    else if ((b & 0xF8) == 0xF0) len = 4;
    else return false;  // Utf8InvalidStartByte
    if (i + len > n) return false;  // InvalidSequence
    uint32_t cp;
    if (len == 1) {
      cp = b;
    } else {
      cp = b & (0x7Fu >> len);
      for (uint32_t k = 1; k < len; k++) {
        if ((s[i + k] & 0xC0) != 0x80) return false;  // Utf8ExpectedContinuation
        cp = (cp << 6) | (s[i + k] & 0x3Fu);
      }
      if (len == 2 && cp < 0x80) return false;  // Utf8OverlongEncoding
      if (len == 3 && cp < 0x800) return false;
      if (len == 3 && cp >= 0xD800 && cp <= 0xDFFF) return false;  // Utf8EncodesSurrogateHalf
      if (len == 4 && cp < 0x10000) return false;
      if (len == 4 && cp > 0x10FFFF) return false;  // Utf8CodepointTooLarge
    }
    out.push_back(cp);
    i += len;
  }
  return true;
}

}  // namespace

// text.show's path (src/text.zig:73-196): glyphs laid out along the
// baseline with their advances and kerning, each glyph's outline appended
// under translate(x + advance + pp1, y) . scale(size / units_per_em).
static int32_t text_path(z2d_font *f, const char *utf8, uint64_t len, double x, double y,
                         const z2d_text_opts *o, std::vector<z2d_path_node> &nodes) {
  if (!f || !o || (!utf8 && len)) return Z2D_ERR_INVALID_ARGUMENT;
  std::vector<uint32_t> cps;
  if (!utf8_decode(reinterpret_cast<const uint8_t *>(utf8), len, cps)) return Z2D_ERR_INVALID_TEXT;
  std::lock_guard<std::mutex> lock(f->mu);
  PathB path;
  const Xf base{o->transformation[0], o->transformation[1], o->transformation[2],
                o->transformation[3], o->transformation[4], o->transformation[5]};
  path.t = base;
  const double scale = o->size / (double)f->units_per_em;
  double advance = 0.0;
  for (size_t idx = 0; idx < cps.size(); idx++) {
    GlyphRec cur;
    if (!glyph_init(*f, cps[idx], cur)) return Z2D_ERR_INVALID_FONT;
    GlyphRec next;
    const bool has_next = idx + 1 < cps.size();
    if (has_next && !glyph_init(*f, cps[idx + 1], next)) return Z2D_ERR_INVALID_FONT;
    if (cur.has_outline) {
      const Xf saved = path.t;
      const OutlineRec *ol = nullptr;
      if (!outline_get(*f, cps[idx], cur, ol)) return Z2D_ERR_INVALID_FONT;
      const double pp1 = f->lsb_is_at_x_zero ? 0.0 : (double)(ol->x_min - cur.lsb);
      path.t = xf_scale(xf_translate(path.t, x + advance + pp1, y), scale, scale);
      for (const z2d_path_node &nd : ol->nodes) {
        switch (nd.tag) {
          case Z2D_NODE_MOVE_TO: path.move_to(nd.v[0], nd.v[1]); break;
          case Z2D_NODE_LINE_TO: path.line_to(nd.v[0], nd.v[1]); break;
          case Z2D_NODE_CURVE_TO:
            if (!path.curve_to(nd.v[0], nd.v[1], nd.v[2], nd.v[3], nd.v[4], nd.v[5]))
              return Z2D_ERR_INVALID_FONT;
            break;
          default: path.close(); break;
        }
      }
      path.t = saved;
    }
    int32_t kern = 0;
    if (has_next && !kern_advance(*f, cur.index, next.index, kern)) return Z2D_ERR_INVALID_FONT;
    advance += ((double)(cur.advance > 0 ? cur.advance : f->advance_width_max) + (double)kern) *
               scale;
  }
  nodes = std::move(path.nodes);
  return Z2D_OK;
}

extern "C" {

int32_t z2d_font_load(const void *data, uint64_t len, z2d_font **out) {
  if (!out || (!data && len)) return Z2D_ERR_INVALID_ARGUMENT;
  auto *f = new (std::nothrow) z2d_font();
  if (!f) return Z2D_ERR_OUT_OF_MEMORY;
  f->data.assign(static_cast<const uint8_t *>(data), static_cast<const uint8_t *>(data) + len);
  const int32_t rc = font_parse(*f);
  if (rc) {
    delete f;
    return rc;
  }
  *out = f;
  return Z2D_OK;
}

void z2d_font_destroy(z2d_font *f) { delete f; }

int32_t z2d_font_info(z2d_font *f, uint32_t out[4]) {
  if (!f || !out) return Z2D_ERR_INVALID_ARGUMENT;
  out[0] = f->units_per_em;
  out[1] = f->advance_width_max;
  out[2] = f->number_of_hmetrics;
  out[3] = (f->lsb_is_at_x_zero ? 1u : 0u) | (f->loca_long ? 2u : 0u) | (f->cmap_full ? 4u : 0u);
  return Z2D_OK;
}

int32_t z2d_font_glyph(z2d_font *f, uint32_t codepoint, int32_t out[4]) {
  if (!f || !out) return Z2D_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(f->mu);
  GlyphRec g;
  if (!glyph_init(*f, codepoint, g)) return Z2D_ERR_INVALID_FONT;
  out[0] = (int32_t)g.index;
  out[1] = (int32_t)g.advance;
  out[2] = g.lsb;
  out[3] = g.has_outline ? (int32_t)g.offset : -1;
  return Z2D_OK;
}

int32_t z2d_text_path(z2d_font *font, const char *utf8, uint64_t len, double x, double y,
                      const z2d_text_opts *opts, z2d_path_node *out, uint64_t cap,
                      uint64_t *n_out) {
  if (!n_out || (!out && cap)) return Z2D_ERR_INVALID_ARGUMENT;
  std::vector<z2d_path_node> nodes;
  const int32_t rc = text_path(font, utf8, len, x, y, opts, nodes);
  if (rc) return rc;
  *n_out = nodes.size();
  if (nodes.size() > cap) return Z2D_ERR_INVALID_ARGUMENT;
  if (!nodes.empty()) std::memcpy(out, nodes.data(), nodes.size() * sizeof(z2d_path_node));
  return Z2D_OK;
}

int32_t z2d_gpu_text_show(z2d_gpu_ctx *ctx, z2d_surface_desc *surface,
                          const z2d_pattern_desc *pattern, z2d_font *font, const char *utf8,
                          uint64_t len, double x, double y, const z2d_text_opts *opts) {
  if (!ctx) return Z2D_ERR_NO_CONTEXT;
  std::vector<z2d_path_node> nodes;
  const int32_t rc = text_path(font, utf8, len, x, y, opts, nodes);
  if (rc) return rc;
  return z2d_gpu_fill(ctx, surface, pattern, nodes.data(), nodes.size(), &opts->fill);
}

int32_t z2d_gpu_batch_text(z2d_gpu_batch *b, z2d_gpu_surface *surface,
                           const z2d_pattern_desc *pattern, z2d_font *font, const char *utf8,
                           uint64_t len, double x, double y, const z2d_text_opts *opts) {
  if (!b) return Z2D_ERR_INVALID_ARGUMENT;
  std::vector<z2d_path_node> nodes;
  const int32_t rc = text_path(font, utf8, len, x, y, opts, nodes);
  if (rc) return rc;
  return z2d_gpu_batch_fill(b, surface, pattern, nodes.data(), nodes.size(), &opts->fill);
}

}  // extern "C"
