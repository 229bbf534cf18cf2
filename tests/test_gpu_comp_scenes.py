"""GPU parity on the reference's gradient / dither / operator / gamma spec
cases (spec/061-072): the HIP library against the oracle on identical
calls.  Linear-RGB, radial, dither and integer operator cases are bit-exact;
cases whose gradients go through std.math.pow (sRGB, HSL) or atan2 (conic)
may differ by 1 LSB because the device's expf/logf/atan2 are not the host
libm's."""
import numpy as np
import pytest

from gpu_util import gpu_backend
from oracle_backend import OracleBackend
from test_oracle_comp_fixtures import cases

pytestmark = pytest.mark.gpu

EXACT = {"061_linear_gradient", "063_radial_gradient", "064_radial_source",
         "068_gradient_deband", "069_gradient_dither_context", "070_compositor_ops",
         "071_gamma_linear"}


@pytest.mark.parametrize("name,fn,aa,suffix", cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_gpu_comp_scene(name, fn, aa, suffix):
    want = fn(OracleBackend(), aa).data().astype(np.int32)
    got = fn(gpu_backend(), aa).data().astype(np.int32)
    d = np.abs(got - want)
    tol = 0 if name in EXACT else 1
    assert d.max(initial=0) <= tol, f"{name}{suffix}: max diff {d.max()}, {int((d > 0).sum())} bytes"
