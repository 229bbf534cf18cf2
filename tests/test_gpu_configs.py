"""The benchmark configurations' own inputs, checked on the GPU: the 512x512
star (bit-exact against the oracle), a slice of the compositor sweep and a
reduced 4K scene (bit-exact)."""
import ctypes as C

import numpy as np
import pytest

from z2d_amd import abi, scenes
from gpu_util import diff_report, gpu_backend, run_draws
from oracle_backend import OracleBackend
from zscene import Surface

pytestmark = pytest.mark.gpu


def test_star_512_bit_exact():
    sc = scenes.star_512()
    want = run_draws(OracleBackend(), sc["format"], sc["width"], sc["height"], sc["buf"], sc["draws"])
    got = run_draws(gpu_backend(), sc["format"], sc["width"], sc["height"], sc["buf"], sc["draws"])
    assert not diff_report(got.data(), want.data())
    # closed form: the star is simple, so even-odd and non-zero agree and the
    # covered area is 5*240*96*sin(36deg) px^2 up to the AA boundary
    cov = np.count_nonzero(np.frombuffer(bytes(got.data()), np.uint8).reshape(-1, 4)[:, 1] != 255)
    assert abs(cov - 5 * 240 * 96 * np.sin(np.radians(36))) < 0.02 * cov


def test_scene_4k_subset_bit_exact():
    rng = scenes.SplitMix64(0xC3)
    draws, keep = scenes.random_paths(rng, 150, 640, 360, max_side=200.0)
    init = np.full(640 * 360 * 4, 255, np.uint8)
    want = run_draws(OracleBackend(), abi.RGBA, 640, 360, init, draws)
    got = run_draws(gpu_backend(), abi.RGBA, 640, 360, init, draws)
    msg = diff_report(got.data(), want.data())
    assert not msg, msg


def _batch_render(draw_sets, mode):
    """Renders [(fmt, w, h, init, draws)] as one batch; returns host buffers."""
    from z2d_amd.gpu import Batch, DeviceSurface
    ctx = gpu_backend().ctx
    b = Batch(ctx)
    surfs = []
    for fmt, w, h, init, draws in draw_sets:
        s = DeviceSurface(ctx, fmt, w, h)
        s.upload(np.ascontiguousarray(init[: s.nbytes]))
        surfs.append(s)
        for pat, nodes, opts in draws:
            assert b.fill(s, pat, nodes, opts) == 0
    b.prepare()
    b.render(mode)
    b.wait()
    outs = []
    for s in surfs:
        o = np.zeros(s.nbytes, np.uint8)
        s.download(o)
        outs.append(o)
    b.close()
    return outs


@pytest.mark.parametrize("mode", [0, 1], ids=["persistent", "per_surface"])
def test_batch_many_surfaces_bit_exact(mode):
    """Uneven surfaces in one batch give the same bytes as one painter.fill at
    a time on the oracle (draw order per surface preserved)."""
    rng = scenes.SplitMix64(0xC4)
    sets = []
    for i in range(24):
        side = int(round(np.exp(rng.uniform(np.log(16), np.log(300)))))
        fmt = abi.RGBA if rng.uniform() < 0.8 else abi.ALPHA8
        draws, keep = scenes.random_paths(rng, rng.randint(1, 12), side, side,
                                          max_side=min(200.0, side))
        init = np.zeros(abi.surface_nbytes(fmt, side, side), np.uint8)
        sets.append((fmt, side, side, init, draws, keep))
    outs = _batch_render([s[:5] for s in sets], mode)
    for (fmt, w, h, init, draws, keep), got in zip(sets, outs):
        want = run_draws(OracleBackend(), fmt, w, h, init, draws)
        msg = diff_report(got, want.data())
        assert not msg, msg


def test_batch_draw_n_equals_per_item_calls():
    """z2d_gpu_batch_draw_n (strokes tessellated on host threads) appends the
    same draws as one fill / stroke call per item: 96 uneven surfaces with
    30% strokes render to the same bytes, per-item statuses included (a
    stroke with a singular matrix fails alone, as its single call does)."""
    from z2d_amd.gpu import Batch, DeviceSurface
    ctx = gpu_backend().ctx
    work = scenes.uneven_surfaces(96, 0xC4)
    bad_opts = None
    outs = []
    for use_n in (False, True):
        b = Batch(ctx)
        surfs = [DeviceSurface(ctx, fmt, side, side) for fmt, side, _j, _k in work]
        for s in surfs:
            s.upload(np.zeros(s.nbytes, np.uint8))
        jobs = []
        for s, (fmt, side, js, keep) in zip(surfs, work):
            for kind, pat, nodes, opts in js:
                jobs.append((s, kind, pat, nodes, opts))
        # one failing stroke in the middle: a singular transformation
        k = next(i for i, j in enumerate(jobs) if j[1] == "stroke" and i > len(jobs) // 2)
        s, kind, pat, nodes, opts = jobs[k]
        if bad_opts is None:
            bad_opts = abi.StrokeOpts.from_buffer_copy(opts)
            bad_opts.transformation[:] = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
        jobs.insert(k, (s, "stroke", pat, nodes, bad_opts))
        if use_n:
            st = b.draw_n(jobs, threads=8)
        else:
            st = [(b.stroke if kind == "stroke" else b.fill)(s, pat, nodes, opts)
                  for s, kind, pat, nodes, opts in jobs]
        assert st[k] == abi.ERR_INVALID_MATRIX, st[k]
        assert all(v == 0 for i, v in enumerate(st) if i != k)
        b.prepare()
        b.render(0)
        b.wait()
        got = []
        for s in surfs:
            o = np.empty(s.nbytes, np.uint8)
            s.download(o)
            got.append(o)
            s.close()
        b.close()
        outs.append(got)
    for i, (a, c) in enumerate(zip(*outs)):
        assert np.array_equal(a, c), f"surface {i}"
