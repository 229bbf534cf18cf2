"""The device edge build (one lane per node) against the oracle's serial
fill plotter (src/internal/tess/fill_plotter.zig:21-97) on node lists that
exercise its state: repeated points, closes with fewer than three points,
runs of closes, a line right after a close (no move), a final move, and
subpaths crossing the 64-node chunk boundary; with curves, the device
flattening's per-node form and its per-draw fallback (a curve after a close,
a close before the first move, short closed contours)."""
import math

import numpy as np
import pytest

from z2d_amd import abi, scenes
from gpu_util import diff_report, gpu_backend, run_draws
from oracle_backend import OracleBackend

pytestmark = pytest.mark.gpu

M, L, Z, C = abi.MOVE_TO, abi.LINE_TO, abi.CLOSE_PATH, abi.CURVE_TO


def _cases():
    a, b, c, d, e = (5.2, 4.1), (40.7, 9.3), (30.1, 38.8), (12.4, 30.2), (50.5, 45.5)
    yield "line_after_close", [(M, *a), (L, *b), (L, *c), (Z,), (L, *d), (L, *e), (Z,)]
    yield "repeats", [(M, *a), (L, *a), (L, *b), (L, *b), (L, *c), (L, *c), (Z,), (Z,), (Z,),
                      (M, *a)]
    yield "short_close", [(M, *a), (L, *b), (Z,), (M, *c), (L, *d), (L, *e), (Z,), (M, *c)]
    yield "close_at_first", [(M, *a), (L, *b), (L, *c), (L, *a), (Z,), (M, *a)]
    yield "final_move", [(M, *a), (L, *b), (L, *c), (Z,), (M, *a), (M, *d)]
    yield "horizontal", [(M, 3.0, 10.0), (L, 50.0, 10.0), (L, 50.0, 40.0), (L, 3.0, 10.0),
                         (Z,), (M, 3.0, 10.0)]
    # curves: after a close (no move), as the only segment of a closed
    # contour, degenerate (a == b and c == d), right before a close, and
    # returning to the contour's start
    yield "curve_after_close", [(M, *a), (L, *b), (L, *c), (Z,), (C, *d, *e, 20.0, 50.0),
                                (L, *b), (Z,)]
    yield "curve_only_contour", [(M, *a), (C, *b, *c, *d), (Z,), (M, *e), (C, 40.0, 60.0,
                                 10.0, 60.0, *e), (Z,)]
    yield "degenerate_curve", [(M, *a), (C, *a, *c, *c), (L, *d), (Z,), (M, *b),
                               (C, *b, *b, *b), (L, *e), (L, *c), (Z,)]
    yield "curve_before_close", [(M, *a), (L, *b), (C, *c, *d, *e), (Z,)]
    yield "curve_to_start", [(M, *a), (L, *b), (C, *c, *d, *a), (Z,), (M, *c), (C, *d, *e, *c),
                             (Z,)]
    yield "close_first", [(Z,), (M, *a), (C, *b, *c, *d), (L, *e), (Z,)]
    yield "double_close_curve", [(M, *a), (C, *b, *c, *d), (L, *e), (Z,), (Z,), (C, *b, *c, *e),
                                 (Z,)]
    rng = scenes.SplitMix64(0xC0FE)
    nodes = []
    for k in range(150):  # mixed lines and curves across several 64-node chunks
        pt = (rng.uniform(2, 62), rng.uniform(2, 62))
        if k % 37 == 0:
            if k:
                nodes.append((Z,))
            nodes.append((M, *pt))
        elif rng.uniform() < 0.4:
            nodes.append((C, rng.uniform(2, 62), rng.uniform(2, 62), rng.uniform(2, 62),
                          rng.uniform(2, 62), *pt))
        else:
            nodes.append((L, *pt))
        if k % 37 == 30:
            nodes.append((Z,))  # a close inside the contour, lines/curves after it
    nodes.append((Z,))
    yield "mixed_curves_chunks", nodes
    # long subpaths across chunk boundaries, with repeats and closes near them
    rng = scenes.SplitMix64(0x9107)
    for n_sub, n_pts in ((1, 200), (5, 63), (7, 64), (3, 130)):
        nodes = []
        for _ in range(n_sub):
            cx, cy = rng.uniform(10, 54), rng.uniform(10, 54)
            start = None
            for k in range(n_pts):
                ang = 2 * math.pi * k / n_pts
                r = rng.uniform(3, 30)
                p = (cx + r * math.cos(ang), cy + r * math.sin(ang))
                if start is None:
                    start = p
                    nodes.append((M, *p))
                else:
                    nodes.append((L, *p))
                    if rng.uniform() < 0.1:
                        nodes.append((L, *p))  # repeated point
            nodes.append((Z,))
            if rng.uniform() < 0.5:
                nodes.append((Z,))
            nodes.append((M, *start))
        yield f"chunks_{n_sub}x{n_pts}", nodes


CASES = list(_cases())


@pytest.mark.parametrize("name,nodes", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("aa", [abi.AA_NONE, abi.AA_DEFAULT, abi.AA_SUPERSAMPLE_4X])
@pytest.mark.parametrize("rule", [abi.NON_ZERO, abi.EVEN_ODD])
def test_plotter_cases(name, nodes, aa, rule):
    arr = abi.nodes_array(nodes)
    pat = abi.solid_pattern(abi.rgba(200, 30, 30, 255))
    draws = [(pat, arr, abi.fill_opts(aa, rule))]
    init = np.full(64 * 64 * 4, 255, np.uint8)
    want = run_draws(OracleBackend(), abi.RGBA, 64, 64, init, draws)
    got = run_draws(gpu_backend(), abi.RGBA, 64, 64, init, draws)
    msg = diff_report(got.data(), want.data())
    assert not msg, f"{name}: {msg}"
