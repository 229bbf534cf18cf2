"""The batch pipeline replayed as a captured hipGraph equals the eager
pipeline: the same batch (fills and strokes on several surfaces of several
formats) rendered repeatedly onto copies of the same surfaces, with eager
GPU work and synchronisation between the replays, in both launch modes and
with the work counters toggled (which re-captures the graph)."""
import numpy as np
import pytest

from z2d_amd import abi, scenes
from gpu_util import diff_report, gpu_backend

pytestmark = pytest.mark.gpu


def _build(gpu, ctx, seed):
    rng = scenes.SplitMix64(seed)
    b = gpu.Batch(ctx)
    surfs, keep = [], []
    for i in range(6):
        fmt = [abi.RGBA, abi.ALPHA8, abi.RGB, abi.ARGB, abi.ALPHA4, abi.RGBA][i]
        w, h = rng.randint(20, 300), rng.randint(20, 300)
        init = rng.u8_array(abi.surface_nbytes(fmt, w, h))
        jobs, k = scenes.random_jobs(rng, rng.randint(1, 20), w, h, max_side=min(200.0, w, h))
        keep.append(k)
        surfs.append((fmt, w, h, init, jobs))
    return b, surfs, keep


def _surfaces(gpu, ctx, b, surfs):
    out = []
    for fmt, w, h, init, jobs in surfs:
        s = gpu.DeviceSurface(ctx, fmt, w, h)
        s.upload(init)
        for kind, pat, nodes, opts in jobs:
            rc = (b.stroke if kind == "stroke" else b.fill)(s, pat, nodes, opts)
            assert rc == 0, rc
        out.append(s)
    return out


@pytest.mark.parametrize("mode", [0, 1], ids=["persistent", "per_surface"])
def test_graph_replays_equal_eager(mode):
    from z2d_amd import gpu
    ctx = gpu_backend().ctx
    b1, spec, keep = _build(gpu, ctx, 0x6A + mode)
    b2 = gpu.Batch(ctx)
    eager = _surfaces(gpu, ctx, b1, spec)
    graph = _surfaces(gpu, ctx, b2, spec)
    b2.use_graph(True)
    try:
        for rep in range(4):
            if rep == 2:  # re-capture with the counters on, off again after
                b1.counters(True)
                b2.counters(True)
            b1.render(mode)
            b1.wait()
            b2.render(mode)
            b2.wait()
            # eager device work between replays
            ctx.check(gpu.lib().z2d_gpu_stream_copy(ctx.h, 1 << 16, 1, None), "stream_copy")
            for i, (se, sg) in enumerate(zip(eager, graph)):
                a = np.zeros(se.nbytes, np.uint8)
                g = np.zeros(sg.nbytes, np.uint8)
                se.download(a)
                sg.download(g)
                msg = diff_report(g, a)
                assert not msg, f"replay {rep} surface {i}: {msg}"
            if rep == 2:
                assert b1.counters(False)[:3] == b2.counters(False)[:3]
    finally:
        for s in eager + graph:
            s.close()
        b1.close()
        b2.close()
