"""Shared helpers for the GPU parity tests (which compare the HIP library
against the CPU oracle on identical inputs)."""
import numpy as np

from z2d_amd import abi
from oracle_backend import OracleBackend
from zscene import Surface

_backend = None


def gpu_backend():
    global _backend
    if _backend is None:
        from z2d_amd.gpu import Context, GpuBackend
        _backend = GpuBackend(Context())
    return _backend


def run_draws(backend, fmt, w, h, init, draws):
    """Applies a list of (pattern, nodes, opts) fills to a fresh surface."""
    s = Surface(fmt, w, h)
    s.buf[: s.nbytes] = init[: s.nbytes]
    for pat, nodes, opts in draws:
        rc = backend.fill(s, pat, nodes, opts)
        assert rc == 0, f"{backend.name} fill rc={rc}"
    return s


def diff_report(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    idx = np.flatnonzero(a != b)
    if idx.size == 0:
        return ""
    return f"{idx.size} bytes differ, first at {idx[:8].tolist()}: gpu={a[idx[:8]].tolist()} oracle={b[idx[:8]].tolist()}"
