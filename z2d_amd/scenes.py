"""Synthetic workloads of the benchmark configurations (SplitMix64-seeded,
deterministic).  Each generator returns plain descriptors (surfaces as numpy
buffers, node arrays, patterns, options) that can be handed to the HIP
library or to any other implementation of the same C ABI.

  star_512         512x512 RGBA8 white, one 5-point star, even-odd, MSAA
  compositor_4k    4096x4096 RGBA8 premultiplied src/dst (+ alpha8 mask)
  scene_4k         3840x2160 RGBA8 white, 10,000 random closed paths
                   (70% lines / 30% cubics, 3-64 segments, log-uniform
                   4-512 px boxes, 60% solid / 40% gradient paints,
                   50/50 non-zero / even-odd, src_over, MSAA)
"""
import ctypes as C
import math

import numpy as np

from . import abi

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (state += 0x9E3779B97F4A7C15, then the two xor-shift
    multiplies).  Outputs are produced in vectorised blocks; the sequence is
    the scalar one."""

    BLOCK = 4096

    def __init__(self, seed):
        self.s = seed & MASK64
        self._buf = []
        self._pos = 0

    def _refill(self):
        self._buf = splitmix_stream(self.s, self.BLOCK).tolist()
        self._pos = 0
        self.s = (self.s + self.BLOCK * 0x9E3779B97F4A7C15) & MASK64

    def next(self):
        if self._pos >= len(self._buf):
            self._refill()
        v = self._buf[self._pos]
        self._pos += 1
        return v

    def uniform(self, a=0.0, b=1.0):
        u = (self.next() >> 11) * (1.0 / (1 << 53))
        return a + (b - a) * u

    def randint(self, a, b):
        u = (self.next() >> 11) * (1.0 / (1 << 53))
        return a + int(math.floor(u * (b - a + 1)))

    def u8_array(self, n):
        """n uniform bytes (one u64 draw per 8 bytes, little-endian)."""
        words = np.array([self.next() for _ in range((n + 7) // 8)], dtype=np.uint64)
        return words.view(np.uint8)[:n].copy()


def splitmix_stream(seed, n):
    """Vectorised SplitMix64: n consecutive outputs as uint64."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * gamma
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def rgba_from_clamped(r, g, b, a):
    """pixel.RGBA.fromClamped (src/pixel.zig:376-396): round(255*c) then
    RGBA.multiply."""
    def q(v):
        v = min(max(v, 0.0), 1.0)
        return int(math.floor(255.0 * v + 0.5))
    rr, gg, bb, aa = q(r), q(g), q(b), q(a)
    return abi.rgba(rr * aa // 255, gg * aa // 255, bb * aa // 255, aa)


def star_512():
    """The 512x512 star: 10 vertices at -90 + 36k degrees, radius 240/96."""
    w = h = 512
    buf = np.full(w * h * 4, 255, dtype=np.uint8)
    pts = []
    for k in range(10):
        th = math.radians(-90.0 + 36.0 * k)
        r = 240.0 if k % 2 == 0 else 96.0
        pts.append((256.0 + r * math.cos(th), 256.0 + r * math.sin(th)))
    nodes = [(abi.MOVE_TO, *pts[0])] + [(abi.LINE_TO, *p) for p in pts[1:]]
    nodes += [(abi.CLOSE_PATH,), (abi.MOVE_TO, *pts[0])]
    pattern = abi.solid_pattern(abi.rgba(200, 30, 30, 255))
    opts = abi.fill_opts(abi.AA_DEFAULT, abi.EVEN_ODD, "src_over")
    return dict(format=abi.RGBA, width=w, height=h, buf=buf,
                draws=[(pattern, abi.nodes_array(nodes), opts)], keep=[])


def premultiplied_rgba(seed, n):
    """n RGBA8 pixels: a = U{0..255}, r,g,b = U{0..a} (premultiplied-valid)."""
    z = splitmix_stream(seed, 4 * n).reshape(n, 4)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    a = np.floor(u[:, 0] * 256).astype(np.int64)
    out = np.empty((n, 4), dtype=np.uint8)
    out[:, 3] = a
    for c in range(3):
        out[:, c] = np.floor(u[:, c + 1] * (a + 1)).astype(np.int64)
    return out.reshape(-1)


def compositor_inputs(w=4096, h=4096, seed=0xC2):
    n = w * h
    src = premultiplied_rgba(seed, n)
    dst = premultiplied_rgba(seed + 0x1000, n)
    z = splitmix_stream(seed + 0x2000, n)
    mask = (z >> np.uint64(56)).astype(np.uint8)
    return src, dst, mask


def _gradient_for(rng, x0, y0, s):
    gtype = rng.randint(0, 2)
    if gtype == abi.GRADIENT_LINEAR:
        geom = [rng.uniform(x0, x0 + s), rng.uniform(y0, y0 + s),
                rng.uniform(x0, x0 + s), rng.uniform(y0, y0 + s)]
    elif gtype == abi.GRADIENT_RADIAL:
        geom = [rng.uniform(x0, x0 + s), rng.uniform(y0, y0 + s), rng.uniform(0.0, s / 4),
                rng.uniform(x0, x0 + s), rng.uniform(y0, y0 + s), rng.uniform(s / 4, s)]
    else:
        geom = [rng.uniform(x0, x0 + s), rng.uniform(y0, y0 + s),
                rng.uniform(0.0, 2 * math.pi)]
    nst = rng.randint(2, 6)
    offs = sorted(rng.uniform() for _ in range(nst))
    stops = []
    for o in offs:
        col = abi.color(abi.COLOR_LINEAR_RGB, rng.uniform(), rng.uniform(), rng.uniform(),
                        rng.uniform(0.25, 1.0))
        stops.append((o, col))
    return abi.Gradient(gtype, geom, stops, abi.INTERP_LINEAR_RGB)


def random_paths(rng, n_paths, width, height, max_side=512.0, min_side=4.0, aa=abi.AA_DEFAULT,
                 gradients=True, op="src_over", curves=True):
    """The path distribution of the 4K scene: returns (draws, keep)."""
    draws, keep = [], []
    lo, hi = math.log(min_side), math.log(max(min_side, max_side))
    for _ in range(n_paths):
        nseg = rng.randint(3, 64)
        s = math.exp(rng.uniform(lo, hi))
        x0 = rng.uniform(0.0, max(0.0, width - s))
        y0 = rng.uniform(0.0, max(0.0, height - s))
        start = (rng.uniform(x0, x0 + s), rng.uniform(y0, y0 + s))
        nodes = [(abi.MOVE_TO, *start)]
        for _ in range(nseg):
            if not curves or rng.uniform() < 0.7:
                nodes.append((abi.LINE_TO, rng.uniform(x0, x0 + s), rng.uniform(y0, y0 + s)))
            else:
                c = [rng.uniform(x0, x0 + s) if i % 2 == 0 else rng.uniform(y0, y0 + s)
                     for i in range(6)]
                nodes.append((abi.CURVE_TO, *c))
        nodes.append((abi.CLOSE_PATH,))
        nodes.append((abi.MOVE_TO, *start))
        rule = abi.NON_ZERO if rng.uniform() < 0.5 else abi.EVEN_ODD
        if gradients and rng.uniform() >= 0.6:
            g = _gradient_for(rng, x0, y0, s)
            keep.append(g)
            pattern = g.pattern()
        else:
            r, gg, b = rng.uniform(), rng.uniform(), rng.uniform()
            a = rng.randint(64, 255)
            pattern = abi.solid_pattern(rgba_from_clamped(r, gg, b, a / 255.0))
        opts = abi.fill_opts(aa, rule, op)
        draws.append((pattern, abi.nodes_array(nodes), opts))
    return draws, keep


def stroke_opts(rng, aa=abi.AA_DEFAULT, op="src_over"):
    """Stroke options of the batched form: width U(0.5, 8), join and cap
    uniform over the three modes, 25% dashed with 2-6 dash lengths U(1, 20)
    and offset 0.  Returns (opts, keep)."""
    o = abi.StrokeOpts()
    o.anti_aliasing_mode = aa
    o.line_width = rng.uniform(0.5, 8.0)
    o.line_join_mode = rng.randint(0, 2)
    o.line_cap_mode = rng.randint(0, 2)
    o.op = abi.OP[op]
    o.precision = abi.PRECISION_INTEGER
    o.miter_limit = 10.0
    o.tolerance = 0.1
    o.transformation[:] = abi.IDENTITY
    keep = []
    if rng.uniform() < 0.25:
        d = (C.c_double * rng.randint(2, 6))(*[])
        for i in range(len(d)):
            d[i] = rng.uniform(1.0, 20.0)
        o.dashes = C.cast(d, C.POINTER(C.c_double))
        o.n_dashes = len(d)
        keep.append(d)
    o.dash_offset = 0.0
    return o, keep


def random_jobs(rng, n_paths, width, height, max_side=512.0, stroke_share=0.3):
    """The batched form's per-surface path list: each path from the 4K
    scene's distribution, stroked with probability `stroke_share`.  Returns
    ([(kind, pattern, nodes, opts)], keep)."""
    jobs, keep = [], []
    for _ in range(n_paths):
        draws, k = random_paths(rng, 1, width, height, max_side=max_side)
        keep += k
        pat, nodes, fopts = draws[0]
        if rng.uniform() < stroke_share:
            sopts, k2 = stroke_opts(rng)
            keep += k2
            jobs.append(("stroke", pat, nodes, sopts))
        else:
            jobs.append(("fill", pat, nodes, fopts))
    return jobs, keep


def uneven_surfaces(n_surfaces=4096, seed=0xC4):
    """The batched form's surfaces, in generation order: side log-uniform in
    16-1024 px, 80% RGBA8 / 20% alpha8 (cleared to 0), 1-50 paths each of the
    4K scene's distribution with boxes up to min(512, side), 30% stroked.
    Returns [(fmt, side, jobs, keep)]."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n_surfaces):
        side = int(round(np.exp(rng.uniform(np.log(16), np.log(1024)))))
        fmt = abi.RGBA if rng.uniform() < 0.8 else abi.ALPHA8
        jobs, keep = random_jobs(rng, rng.randint(1, 50), side, side,
                                 max_side=min(512.0, side))
        out.append((fmt, side, jobs, keep))
    return out


def served_requests(n, seed=0xC5):
    """The served form's requests: one surface job of the batched form's
    distribution each, half of them with a uniform alpha8 clip mask.
    Returns [(fmt, side, jobs, keep, clip or None)]."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n):
        side = int(round(np.exp(rng.uniform(np.log(16), np.log(1024)))))
        fmt = abi.RGBA if rng.uniform() < 0.8 else abi.ALPHA8
        jobs, keep = random_jobs(rng, rng.randint(1, 50), side, side,
                                 max_side=min(512.0, side))
        clip = None
        if rng.uniform() < 0.5:
            clip = splitmix_stream(rng.next(), (side * side + 7) // 8).view(np.uint8)
            clip = clip[: side * side].copy()
        out.append((fmt, side, jobs, keep, clip))
    return out


def scene_4k(n_paths=10000, width=3840, height=2160, seed=0xC3):
    rng = SplitMix64(seed)
    buf = np.full(width * height * 4, 255, dtype=np.uint8)
    draws, keep = random_paths(rng, n_paths, width, height)
    return dict(format=abi.RGBA, width=width, height=height, buf=buf, draws=draws, keep=keep)
