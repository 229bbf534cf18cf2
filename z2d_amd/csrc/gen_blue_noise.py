"""Generates blue_noise_64x64.inc, the 64x64 ordered-dither threshold matrix
the reference's Dither pattern indexes (Pixman's generated blue-noise matrix,
MIT licence; the reference keeps it in src/internal/blue_noise.zig).  The
values have to be identical for dithered output to match bit for bit, so
they are kept here as data.  Reads the reference checkout given on the
command line (default /root/reference)."""
import hashlib
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
text = open(os.path.join(ref, "src", "internal", "blue_noise.zig")).read()
body = text[text.index("= .{") + 4: text.rindex("};")]
vals = [int(v) for v in re.findall(r"\d+", body)]
assert len(vals) == 4096
digest = hashlib.sha256(b"".join(v.to_bytes(2, "little") for v in vals)).hexdigest()
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blue_noise_64x64.inc")
with open(out, "w") as f:
    f.write("// 64x64 blue-noise dither thresholds, indexed [(x mod 64) << 6 | (y mod 64)].\n")
    f.write("// Pixman's generated matrix (pixman/dither/blue-noise-64x64.h, MIT licence).\n")
    f.write(f"// sha256 of the little-endian u16 values: {digest}\n")
    for i in range(0, 4096, 16):
        f.write("  " + ", ".join(f"{v:4d}" for v in vals[i:i + 16]) + ",\n")
print(out, digest)
