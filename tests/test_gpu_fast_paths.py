"""The fast tile renderer's draws on the GPU against the oracle, bit-exact:
multisample src_over fills in integer precision (painter.fill,
src/painter.zig:66-143; multisample.run, src/internal/raster/
multisample.zig) with solid paints and linear / radial / conic gradients
interpolated in linear RGB with up to 8 stops, including repeated stop
offsets (hard stops) and offsets that pixel centers hit exactly, where the
reference's binary stop search (Stop.List.search, src/gradient.zig:797-905)
decides between equal stops; mixed with draws the general renderer takes
(other operators, 9 stops, sRGB interpolation, dither) on the same
surfaces, one draw per call and all draws in one batch."""
import numpy as np
import pytest

from z2d_amd import abi
from z2d_amd.scenes import SplitMix64, rgba_from_clamped
from gpu_util import diff_report, gpu_backend, run_draws
from oracle_backend import OracleBackend

pytestmark = pytest.mark.gpu


def _stops(rng, n):
    """n stops whose offsets are snapped to sixteenths (pixel centers of the
    tests' gradients land on them) with repeats (hard stops)."""
    offs = sorted(rng.randint(0, 16) / 16.0 for _ in range(n))
    if n >= 3 and rng.uniform() < 0.5:
        offs[n // 2] = offs[n // 2 - 1]  # a hard stop
    return [(o, abi.color(abi.COLOR_LINEAR_RGB, rng.uniform(), rng.uniform(), rng.uniform(),
                          rng.uniform(0.2, 1.0))) for o in offs]


def _paint(rng, w, h, keep, general_ok):
    kind = rng.randint(0, 9)
    if kind <= 3:
        return abi.solid_pattern(rgba_from_clamped(rng.uniform(), rng.uniform(), rng.uniform(),
                                                   rng.uniform(0.25, 1.0)))
    gtype = rng.randint(0, 2)
    if gtype == abi.GRADIENT_LINEAR and rng.uniform() < 0.5:
        # pixel centers (x + 0.5) hit offsets k/16 exactly: start at x = -0.5,
        # length 16 * whole pixels
        geom = [-0.5, 0.0, -0.5 + 16.0 * rng.randint(1, max(1, w // 16 + 1)), 0.0]
    else:
        geom = [rng.uniform(-10, w + 10), rng.uniform(-10, h + 10), rng.uniform(0, 30),
                rng.uniform(-10, w + 10), rng.uniform(-10, h + 10), rng.uniform(0, 90)]
        if gtype == abi.GRADIENT_CONIC:
            geom[2] = rng.uniform(-7, 7)
    n = rng.randint(0, 8)
    method = abi.INTERP_LINEAR_RGB
    if general_ok and kind == 9:
        n = 9 if rng.uniform() < 0.5 else n  # beyond the fast renderer's stops
        method = abi.INTERP_SRGB if n != 9 else method
    matrix = abi.IDENTITY
    if rng.uniform() < 0.25:
        matrix = (rng.uniform(0.5, 2), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5),
                  rng.uniform(0.5, 2), rng.uniform(-20, 20), rng.uniform(-20, 20))
    g = abi.Gradient(gtype, geom, _stops(rng, n), method, abi.POLAR_SHORTER, matrix)
    keep.append(g)
    return g.pattern()


def _case(seed, general_ok):
    rng = SplitMix64(0xFA57 + seed)
    fmt = rng.randint(0, 7)
    w, h = rng.randint(8, 230), rng.randint(4, 120)
    init = rng.u8_array(abi.surface_nbytes(fmt, w, h))
    if fmt in (abi.RGBA, abi.ARGB):
        from z2d_amd.scenes import premultiplied_rgba
        init = premultiplied_rgba(rng.next(), w * h).reshape(-1)
    keep, draws = [], []
    for _ in range(rng.randint(1, 12)):
        nodes = []
        for _c in range(rng.randint(1, 3)):
            cx, cy = rng.uniform(-20, w + 20), rng.uniform(-20, h + 20)
            r = rng.uniform(2, max(w, h))
            nodes.append((abi.MOVE_TO, cx + rng.uniform(-r, r), cy + rng.uniform(-r, r)))
            for _k in range(rng.randint(2, 9)):
                if rng.uniform() < 0.3:
                    nodes.append((abi.CURVE_TO, *[v for _q in range(3) for v in
                                                  (cx + rng.uniform(-r, r),
                                                   cy + rng.uniform(-r, r))]))
                else:
                    nodes.append((abi.LINE_TO, cx + rng.uniform(-r, r), cy + rng.uniform(-r, r)))
            nodes.append((abi.CLOSE_PATH,))
        pat = _paint(rng, w, h, keep, general_ok)
        op = "src_over"
        aa = [abi.AA_DEFAULT, abi.AA_MULTISAMPLE_4X][rng.randint(0, 1)]
        if general_ok and rng.uniform() < 0.15:
            op = ["multiply", "src", "dst_out", "xor"][rng.randint(0, 3)]
        if general_ok and rng.uniform() < 0.08:
            pat = abi.dither_pattern(1, 8, px=rgba_from_clamped(rng.uniform(), rng.uniform(),
                                                                  rng.uniform(), 1.0))
        draws.append((pat, abi.nodes_array(nodes),
                      abi.fill_opts(aa, rng.randint(0, 1), op, abi.PRECISION_INTEGER,
                                    rng.uniform(0.05, 1.0))))
    return fmt, w, h, np.ascontiguousarray(init, np.uint8), draws, keep


@pytest.mark.parametrize("seed", list(range(80)))
def test_fast_draws_one_per_call(seed):
    fmt, w, h, init, draws, keep = _case(seed, general_ok=False)
    want = run_draws(OracleBackend(), fmt, w, h, init, draws)
    got = run_draws(gpu_backend(), fmt, w, h, init, draws)
    msg = diff_report(got.data(), want.data())
    assert not msg, f"seed {seed} fmt {fmt} {w}x{h}: {msg}"


@pytest.mark.parametrize("seed", list(range(60)))
def test_fast_and_general_draws_in_one_batch(seed):
    """All draws of a surface in one batch, fast-renderer draws mixed with
    general-renderer ones (strips holding any general draw go to the general
    renderer as a whole)."""
    from z2d_amd.gpu import Batch, DeviceSurface
    fmt, w, h, init, draws, keep = _case(1000 + seed, general_ok=True)
    want = run_draws(OracleBackend(), fmt, w, h, init, draws)
    ctx = gpu_backend().ctx
    s = DeviceSurface(ctx, fmt, w, h)
    buf = np.zeros(s.nbytes, np.uint8)
    buf[:] = init[: s.nbytes]
    s.upload(buf)
    b = Batch(ctx)
    for pat, nodes, opts in draws:
        assert b.fill(s, pat, nodes, opts) == 0
    b.prepare()
    b.render(0)
    b.wait()
    got = np.empty(s.nbytes, np.uint8)
    s.download(got)
    b.close()
    s.close()
    msg = diff_report(got, want.data())
    assert not msg, f"seed {seed} fmt {fmt} {w}x{h}: {msg}"
