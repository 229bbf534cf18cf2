"""Bin height, queue order and the autotuner: the output of a batch does not
depend on the band height (4 to 32 rows) or the strip queue order, and the
autotuner leaves the batch's surfaces as they were, picks one of its
candidates, and renders the same result after."""
import numpy as np
import pytest

from gpu_util import diff_report, gpu_backend
from test_gpu_graph import _build, _surfaces

pytestmark = pytest.mark.gpu


def _contents(surfs):
    out = []
    for s in surfs:
        a = np.zeros(s.nbytes, np.uint8)
        s.download(a)
        out.append(a)
    return out


@pytest.mark.parametrize("mode", [0, 1], ids=["persistent", "per_surface"])
def test_band_height_and_order_do_not_change_output(mode):
    from z2d_amd import gpu
    ctx = gpu_backend().ctx
    results = {}
    for rows, order in ((16, 0), (4, 0), (8, 1), (12, 0), (32, 1), (16, 1)):
        b, spec, keep = _build(gpu, ctx, 0x7B + mode)
        surfs = _surfaces(gpu, ctx, b, spec)
        try:
            b.set_band_height(rows)
            b.set_queue_order(order)
            b.render(mode)
            b.wait()
            results[(rows, order)] = _contents(surfs)
        finally:
            for s in surfs:
                s.close()
            b.close()
    for key, got in results.items():
        for i, (g, w) in enumerate(zip(got, results[(16, 0)])):
            msg = diff_report(g, w)
            assert not msg, f"band height / order {key} surface {i}: {msg}"


def test_autotune_keeps_surfaces_and_output():
    from z2d_amd import gpu
    ctx = gpu_backend().ctx
    b, spec, keep = _build(gpu, ctx, 0x7E)
    surfs = _surfaces(gpu, ctx, b, spec)
    ref_b, _, _ = _build(gpu, ctx, 0x7E)
    ref = _surfaces(gpu, ctx, ref_b, spec)
    try:
        before = _contents(surfs)
        rows, longest, ms = b.autotune(0, 1)
        assert rows in (8, 16, 32) and longest in (0, 1) and all(v > 0 for v in ms.values())
        for i, (g, w) in enumerate(zip(_contents(surfs), before)):
            assert not diff_report(g, w), f"surface {i} changed by the autotuner"
        b.render(0)
        b.wait()
        ref_b.render(0)
        ref_b.wait()
        for i, (g, w) in enumerate(zip(_contents(surfs), _contents(ref))):
            msg = diff_report(g, w)
            assert not msg, f"surface {i} after tuning: {msg}"
    finally:
        for s in surfs + ref:
            s.close()
        b.close()
        ref_b.close()
