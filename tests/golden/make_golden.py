"""Copies the reference's own acceptance fixtures (spec/files/*.png, data
produced by the reference's `zig build spec -Dupdate=true`) that cover the
fill/composite hot path into tests/golden/spec/, with a SHA-256 manifest,
and the test fonts the reference's text tests load (spec/test-fonts,
src/internal/test-fonts) into tests/golden/fonts/.  Reads /root/reference;
run once where the reference is mounted."""
import glob
import hashlib
import os
import shutil
import sys

REF = os.environ.get("Z2D_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "spec")
PREFIXES = ["001", "002", "003", "004", "005", "006", "007", "013", "016", "023", "024",
            "025", "026", "028", "031", "032", "033", "046", "047", "048", "049", "050",
            "051", "052", "053", "059", "061", "062", "063", "064", "065", "066", "067",
            "068", "069", "070", "071", "072", "075", "076", "077", "078", "079", "080",
            # strokes (host tessellation + device raster)
            "008", "009", "010", "011", "012", "014", "015", "017", "018", "019", "020",
            "021", "022", "027", "029", "030", "034", "036", "037", "038", "039", "040",
            "041", "043", "044", "045", "054", "055", "056", "057", "060", "073",
            # arcs (Path.arc restated on the test side)
            "035", "042", "058",
            # text (the library's font parser and layout)
            "074"]


def main():
    src_dir = os.path.join(REF, "spec", "files")
    if not os.path.isdir(src_dir):
        sys.exit(f"reference fixtures not found at {src_dir}")
    os.makedirs(OUT, exist_ok=True)
    lines = []
    for pre in PREFIXES:
        for f in sorted(glob.glob(os.path.join(src_dir, pre + "_*.png"))):
            dst = os.path.join(OUT, os.path.basename(f))
            shutil.copyfile(f, dst)
            h = hashlib.sha256(open(dst, "rb").read()).hexdigest()
            lines.append(f"{h}  {os.path.basename(f)}")
    with open(os.path.join(OUT, "SHA256SUMS"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    fonts = os.path.join(HERE, "fonts")
    os.makedirs(fonts, exist_ok=True)
    for rel in ("spec/test-fonts/Inter-Regular.ttf", "src/internal/test-fonts/Inter-Regular.subset.ttf",
                "src/internal/test-fonts/Inter-SemiBold.subset.ttf",
                "src/internal/test-fonts/AnnotationMono-Regular.ttf"):
        shutil.copyfile(os.path.join(REF, rel), os.path.join(fonts, os.path.basename(rel)))
    print(f"copied {len(lines)} fixtures")


if __name__ == "__main__":
    main()
