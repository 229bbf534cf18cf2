"""The oracle's restatement of Zig's std.math.atan2 (f64; the fdlibm
algorithm of lib/std/math/atan.zig and atan2.zig, used by Conic.getOffset,
src/gradient.zig:731-742) against the C library's atan2, which is accurate
to within one unit in the last place: fdlibm's error bound is below one ulp
too, so the two agree to one ulp everywhere, and a wrong coefficient, range
boundary or quadrant rule would be off by far more.  The special values
(zeros, infinities, x = 1, huge and tiny ratios) follow the algorithm's own
rules exactly."""
import ctypes as C
import math

import numpy as np

from oracle_backend import lib


def _atan2(y, x):
    f = lib().zo_atan2
    f.argtypes = [C.c_double, C.c_double]
    f.restype = C.c_double
    return f(y, x)


def _ulps(a, b):
    ia = np.array([a], np.float64).view(np.int64)[0]
    ib = np.array([b], np.float64).view(np.int64)[0]
    if (ia < 0) != (ib < 0):
        return abs(ia) + abs(ib) if a != b else 0
    return abs(int(ia) - int(ib))


def test_atan2_within_one_ulp_of_libm():
    rng = np.random.default_rng(0xA7A2)
    # ratios across every reduction range, both signs of both arguments
    ys = np.concatenate([rng.uniform(-1, 1, 40000) * 10.0 ** rng.uniform(-8, 8, 40000),
                         rng.uniform(-2000, 2000, 20000)])
    xs = np.concatenate([rng.uniform(-1, 1, 40000) * 10.0 ** rng.uniform(-8, 8, 40000),
                         rng.uniform(-2000, 2000, 20000)])
    worst = 0
    for y, x in zip(ys.tolist(), xs.tolist()):
        worst = max(worst, _ulps(_atan2(y, x), math.atan2(y, x)))
    assert worst <= 1, worst


def test_atan2_reduction_range_boundaries():
    # |y/x| right at and around 7/16, 11/16, 19/16, 39/16 and 2^66
    for r in (7 / 16, 11 / 16, 19 / 16, 39 / 16, 2.0 ** 66, 2.0 ** -27):
        for k in range(-3, 4):
            t = float(np.nextafter(r, np.inf if k > 0 else -np.inf)) if k else r
            for _ in range(abs(k) - 1 if k else 0):
                t = float(np.nextafter(t, np.inf if k > 0 else -np.inf))
            for sy, sx in ((1, 1), (-1, 1), (1, -1), (-1, -1)):
                assert _ulps(_atan2(sy * t, sx * 1.0), math.atan2(sy * t, sx * 1.0)) <= 1
                assert _ulps(_atan2(sy * t * 3.0, sx * 3.0), math.atan2(sy * t * 3.0, sx * 3.0)) <= 1


def test_atan2_special_values():
    pi = math.pi
    inf = math.inf
    assert _atan2(0.0, 5.0) == 0.0 and math.copysign(1, _atan2(-0.0, 5.0)) < 0
    assert _atan2(0.0, -5.0) == pi and _atan2(-0.0, -5.0) == -pi
    assert _atan2(3.0, 0.0) == pi / 2 and _atan2(-3.0, 0.0) == -pi / 2
    assert _atan2(inf, inf) == pi / 4 and _atan2(-inf, -inf) == -3 * pi / 4
    assert _atan2(1.0, inf) == 0.0 and _atan2(1.0, -inf) == pi
    assert _atan2(inf, 1.0) == pi / 2
    assert _atan2(1e300, 1e-300) == pi / 2          # |y/x| > 2^64
    assert _atan2(1e-300, -1e300) == pi             # |y/x| < 2^-64, x < 0: pi - (0 - pi_lo)
    assert _atan2(0.75, 1.0) == _atan2(1.5, 2.0) or _ulps(_atan2(0.75, 1.0), _atan2(1.5, 2.0)) <= 1
    assert math.isnan(_atan2(math.nan, 1.0)) and math.isnan(_atan2(1.0, math.nan))
    # the reference's own Conic.getOffset expectations (src/gradient.zig:1702-1748)
    two_pi = 2 * pi
    def conic(cx, cy, angle, x, y, sx=1.0, sy=1.0):
        px, py = (x + 0.5) / sx, (y + 0.5) / sy
        a = math.fmod(_atan2(py - cy, px - cx) - angle, two_pi)
        a = a + two_pi if a < 0 else a
        return float(np.float32(a / two_pi))
    assert conic(49.5, 49, 0, 49, 99) == 0.25
    assert conic(49.5, 49, pi / 4, 49, 99) == 0.125
    assert conic(49, 49.5, pi / 4, 99, 49) == 0.875
    assert conic(49.5, 49, 0, 98, 396, 2.0, 4.0) == float(np.float32(0.25079378))
