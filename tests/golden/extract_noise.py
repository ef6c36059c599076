"""Regenerate noise_r8.bin from the reference's assets/textures/noise.png.

The reference loads that PNG with stb_image as one channel (texture_manager.cuh:
161-185) and divides by 255; the image is RGBA with R == G == B, so stb's luma
conversion returns R exactly.  The library ships the 256x256 R channel as raw
bytes (row-major, 65536 bytes) so that nothing at run time needs the PNG or a
PNG decoder.  Run once, where the reference checkout is available:

    python tests/golden/extract_noise.py /path/to/reference
"""
import sys
from pathlib import Path

import numpy as np
from PIL import Image


def main(ref_root: str) -> None:
    img = Image.open(Path(ref_root) / "assets" / "textures" / "noise.png")
    arr = np.asarray(img.convert("RGBA"))
    if arr.shape[:2] != (256, 256):
        raise SystemExit(f"unexpected size {arr.shape}")
    r, g, b = arr[..., 0], arr[..., 1], arr[..., 2]
    if not (np.array_equal(r, g) and np.array_equal(r, b)):
        raise SystemExit("noise.png is not grey; stb luma would differ from R")
    out = Path(__file__).resolve().parents[2] / "master_thesis_amd" / "data" / "noise_r8.bin"
    out.write_bytes(r.astype(np.uint8).tobytes())
    print(f"wrote {out} ({r.size} bytes), mean {r.mean():.3f}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
