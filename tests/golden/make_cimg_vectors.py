"""Write tests/golden/cimg_vectors.npz from the reference's own vendored CImg code.

Builds oracle/_ref/libcimg_ref.so (oracle/cimg_ref.cpp compiled against the reference
tree's vendor/cimg/CImg.h) and records, for seeded inputs:
  * CImg<float>::normalize(0,1) of small 8-bit images (CImg.h:26786-26794), including
    a constant image, one already spanning [0,1] and a 3-channel one;
  * CImg<float>::_linear_atXY of channels 0..2 (CImg.h:13475-13491) at sample points
    in range, on texel corners, on the far edges, outside the image and at NaN;
  * the bytes of depth.png as the reference writes it, depth_buffer.normalize(0,255)
    .save("depth.png") (renderer.cpp:93; CImg's libpng writer), for depth buffers with
    the cleared value 1.0 around drawn depths, a constant buffer, a buffer holding
    -0.0 and denormals, and one whose range is a few ulps wide.
The vectors are data (inputs and outputs); they let the oracle be checked where the
reference tree is absent.  Run from the repository root:
    python tests/golden/make_cimg_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def depth_buffers():
    """Seeded depth buffers (name -> [H][W] float32) for the depth.png vectors."""
    rng = np.random.default_rng(20261203)
    out = {}
    d = np.ones((23, 31), np.float32)
    d[3:19, 4:27] = rng.uniform(0.93, 0.999, (16, 23)).astype(np.float32)
    out["depth_frame"] = d
    out["depth_const"] = np.full((5, 7), 0.25, np.float32)
    d = rng.uniform(-1.0, 1.0, (6, 9)).astype(np.float32)
    d[0, 0], d[0, 1], d[1, 0] = -0.0, np.float32(1e-40), np.float32(-3e-39)
    out["depth_signed"] = d
    base = np.float32(0.7)
    d = np.array([[np.nextafter(base, np.float32(2), dtype=np.float32) if (x + y) % 3 == 0 else base
                   for x in range(8)] for y in range(4)], np.float32)
    d[2, 5] = np.float32(np.nextafter(np.nextafter(base, np.float32(2), dtype=np.float32), np.float32(2),
                                      dtype=np.float32))
    out["depth_narrow"] = d
    return out


def depth_png_vectors(ref):
    """depth.png written by CImg (renderer.cpp:93), decoded to bytes."""
    import tempfile

    from PIL import Image
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, d in depth_buffers().items():
            H, W = d.shape
            path = os.path.join(tmp, name + ".png")
            d = np.ascontiguousarray(d)
            if ref.cimg_ref_save_depth_png(path.encode(), d.ctypes.data, W, H) != 0:
                raise SystemExit(f"CImg could not write {path}")
            out[name + "_in"] = d
            out[name + "_png"] = np.asarray(Image.open(path)).copy()
    return out


def main():
    import oracle
    oracle.build()
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libcimg_ref.so could not be built (reference tree absent?)")
    rng = np.random.default_rng(20240611)
    out = {}
    images = {
        "noise4": rng.integers(0, 256, (13, 17, 4), dtype=np.uint8),
        "noise3": rng.integers(7, 200, (9, 11, 3), dtype=np.uint8),
        "const": np.full((4, 5, 3), 77, np.uint8),
        "unit": rng.integers(0, 2, (6, 6, 4), dtype=np.uint8),
    }
    for name, img in images.items():
        H, W, C = img.shape
        planar = np.ascontiguousarray(np.transpose(img, (2, 0, 1)).astype(np.float32))
        norm = np.zeros_like(planar)
        ref.cimg_ref_normalize(planar.ctypes.data, W, H, C, norm.ctypes.data)
        out[f"{name}_texels"] = img
        out[f"{name}_normalized"] = norm
        # sample points: random, integer corners, edges, outside, NaN
        n = 400
        fx = np.concatenate([rng.uniform(-2, W + 2, n), np.arange(W, dtype=np.float64), [W - 1, W - 1.0000001, 0.5, np.nan, -0.0]])
        fy = np.concatenate([rng.uniform(-2, H + 2, n), np.full(W, H - 1.0), [H - 1, 0.25, H - 1.0000001, 0.0, np.nan]])
        fx = fx.astype(np.float32)
        fy = fy.astype(np.float32)
        vals = np.zeros((3, fx.size), np.float32)
        for c in range(3):
            ref.cimg_ref_linear_atXY(norm.ctypes.data, W, H, C, fx.ctypes.data, fy.ctypes.data, c, fx.size,
                                     vals[c].ctypes.data)
        out[f"{name}_fx"] = fx
        out[f"{name}_fy"] = fy
        out[f"{name}_samples"] = vals
    out.update(depth_png_vectors(ref))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cimg_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, sorted(out))


if __name__ == "__main__":
    main()
