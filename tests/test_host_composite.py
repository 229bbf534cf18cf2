"""The device renderer's per-pixel composite (composite_cov,
z2d_amd/csrc/composite_px.h) built for the host with the address and
undefined-behaviour sanitizers and called out of line, against the oracle's
compositeOpaque / compositeOpacity / compositeStride
(src/internal/raster/shared.zig:9-117, src/surface.zig:557-581) on a
one-pixel surface: every operator, both precisions, all eight formats,
coverage 0..16 (multisample) and 0..1 (direct).  The same harness checks the
device code's arithmetic rewrites against the forms they replace: src_over
and the coverage mask on RGBA8 words two channels at a time (every channel
triple, then random words), the branchless float-to-lane conversion
(every 61st float and the neighbourhoods of the halves), overlay and hard
light on 8-bit lanes (random lanes against the signed per-channel forms).
Runs on the CPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_composite_cov_host_sanitized(tmp_path):
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    exe = str(tmp_path / "composite_host")
    cmd = [HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O1", "-g",
           "-ffp-contract=off", "-fno-fast-math",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=undefined",
           os.path.join(ROOT, "tests", "host", "composite_host.cpp"), "-o", exe,
           "-L" + os.path.join(ROOT, "oracle"), "-loracle",
           "-Wl,-rpath," + os.path.join(ROOT, "oracle")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, "48"], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "rewrites: cases" in r.stdout and "mismatches 0" in r.stdout, r.stdout[-3000:]
    assert " mismatches 0\n" in r.stdout.split("rewrites:")[1].splitlines(True)[0], r.stdout
    assert "runtime error" not in r.stderr, r.stderr[-3000:]
