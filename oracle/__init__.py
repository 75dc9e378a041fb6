"""oracle — TEST INFRASTRUCTURE ONLY.

ctypes access to the CPU restatement of the reference frame path
(``oracle/rz_oracle.cpp``, built into ``oracle/_build/liboracle.so``) and to the
reference's own vendored CImg routines (``oracle/_ref/libcimg_ref.so``, built from the
reference tree where it is present).  Only tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this package; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from rasteriser_amd.abi import Args, Material, Matrices, MeshDesc  # record layouts only

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None


class OracleCounts(C.Structure):
    _fields_ = [("front_triangles", C.c_uint64), ("bin_pairs16", C.c_uint64),
                ("fragments", C.c_uint64), ("depth_passes", C.c_uint64)]


def build(reference: str = "/root/reference") -> None:
    """make -C oracle (and the CImg harness when the reference tree exists)."""
    targets = ["all"]
    if os.path.isdir(reference):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", _HERE, f"REFERENCE={reference}"] + targets, check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        L.oracle_frame_matrices.argtypes = [C.POINTER(Args), C.POINTER(Matrices)]
        L.oracle_draw_frame.argtypes = [C.POINTER(MeshDesc), P, C.c_int32, C.POINTER(Args), P, P, P,
                                        C.POINTER(OracleCounts)]
        L.oracle_edge.restype = C.c_float
        L.oracle_edge.argtypes = [C.c_float, C.c_float, P, P]
        L.oracle_barycentric.argtypes = [C.c_float, C.c_float, P, P]
        L.oracle_signed_area.restype = C.c_float
        L.oracle_signed_area.argtypes = [P]
        L.oracle_bounding_box.argtypes = [P, C.c_uint32, C.c_uint32, P]
        L.oracle_vertex_to_raster.argtypes = [P, P, C.c_int, C.c_int, P]
        L.oracle_shade.argtypes = [P, P, P, C.c_int32, P]
        L.oracle_sample.argtypes = [C.POINTER(Material), C.c_float, C.c_float, P]
        L.oracle_normalized_texture.argtypes = [C.POINTER(Material), P]
        L.oracle_linear_atXY.restype = C.c_float
        L.oracle_linear_atXY.argtypes = [C.POINTER(Material), C.c_float, C.c_float, C.c_int]
        L.oracle_cast_u32.restype = C.c_uint32
        L.oracle_cast_u32.argtypes = [C.c_float]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's CImg routines, or None when oracle/_ref was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libcimg_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        P = C.c_void_p
        R.cimg_ref_normalize.argtypes = [P, C.c_int, C.c_int, C.c_int, P]
        R.cimg_ref_linear_atXY.argtypes = [P, C.c_int, C.c_int, C.c_int, P, P, C.c_int, C.c_int, P]
        R.cimg_ref_save_depth_png.argtypes = [C.c_char_p, P, C.c_int, C.c_int]
        _REF = R
    return _REF


def matrices(args: Args) -> dict:
    m = Matrices()
    assert lib().oracle_frame_matrices(C.byref(args), C.byref(m)) == 0
    return {k: np.frombuffer(bytes(getattr(m, k)), np.float32).reshape(4, 4).copy()
            for k in ("model", "view", "modelview", "camera", "normal")}


def draw_frame(scene, args: Args, frame=None, depth=None, lights=None):
    """Run the restated draw_frame.  Returns (frame [3,H,W] u8, depth [H,W] f32,
    ids [H,W] i32 (-1 = no depth-test pass), counts dict, lights (L,10) with trans_dir)."""
    W, H = args.image_width, args.image_height
    frame = np.zeros((3, H, W), np.uint8) if frame is None else np.ascontiguousarray(frame, np.uint8).copy()
    depth = np.ones((H, W), np.float32) if depth is None else np.ascontiguousarray(depth, np.float32).copy()
    ids = np.full((H, W), -1, np.int32)
    lights = np.array(scene.lights if lights is None else lights, np.float32, copy=True)
    desc, keep = scene.desc()
    cnt = OracleCounts()
    st = lib().oracle_draw_frame(C.byref(desc), lights.ctypes.data, lights.shape[0], C.byref(args),
                                 frame.ctypes.data, depth.ctypes.data, ids.ctypes.data, C.byref(cnt))
    del keep
    if st != 0:
        raise RuntimeError(f"oracle status {st}")
    counts = {k: int(getattr(cnt, k)) for k, _ in OracleCounts._fields_}
    return frame, depth, ids, counts, lights


def draw_frame_status(scene, args: Args) -> int:
    """Status code only (for the error-path tests)."""
    W, H = args.image_width, args.image_height
    frame = np.zeros((3, H, W), np.uint8)
    depth = np.ones((H, W), np.float32)
    lights = np.array(scene.lights, np.float32, copy=True)
    desc, keep = scene.desc()
    return lib().oracle_draw_frame(C.byref(desc), lights.ctypes.data, lights.shape[0], C.byref(args),
                                   frame.ctypes.data, depth.ctypes.data, None, None)


def _f(a, n):
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    assert a.size == n
    return a


def edge(px, py, a, b) -> float:
    a, b = _f(a, 4), _f(b, 4)
    return lib().oracle_edge(px, py, a.ctypes.data, b.ctypes.data)


def barycentric(px, py, verts) -> np.ndarray:
    v = _f(verts, 12)
    out = np.zeros(3, np.float32)
    lib().oracle_barycentric(px, py, v.ctypes.data, out.ctypes.data)
    return out


def signed_area(verts) -> float:
    v = _f(verts, 12)
    return lib().oracle_signed_area(v.ctypes.data)


def bounding_box(verts, W, H):
    v = _f(verts, 12)
    out = np.zeros(4, np.uint32)
    lib().oracle_bounding_box(v.ctypes.data, W, H, out.ctypes.data)
    return tuple(int(x) for x in out)


def vertex_to_raster(camera, p, W, H) -> np.ndarray:
    m = _f(camera, 16)
    q = _f(p, 3)
    out = np.zeros(4, np.float32)
    lib().oracle_vertex_to_raster(m.ctypes.data, q.ctypes.data, W, H, out.ctypes.data)
    return out


def shade(normal, albedo, lights) -> np.ndarray:
    n, a = _f(normal, 3), _f(albedo, 3)
    L = np.ascontiguousarray(lights, np.float32).reshape(-1, 10)
    out = np.zeros(3, np.uint32)
    lib().oracle_shade(n.ctypes.data, a.ctypes.data, L.ctypes.data, L.shape[0], out.ctypes.data)
    return out


def _material(texture, diffuse=(1.0, 1.0, 1.0)):
    m = Material()
    m.diffuse_colour[:] = list(diffuse)
    tex = None
    if texture is not None:
        tex = np.ascontiguousarray(texture, np.uint8)
        m.tex_height, m.tex_width, m.tex_channels = tex.shape
        m.texels = tex.ctypes.data
    return m, tex


def sample(texture, u, v, diffuse=(1.0, 1.0, 1.0)) -> np.ndarray:
    m, keep = _material(texture, diffuse)
    out = np.zeros(3, np.float32)
    lib().oracle_sample(C.byref(m), u, v, out.ctypes.data)
    return out


def normalized_texture(texture) -> np.ndarray:
    """Planar [C][H][W] float texels after the Material constructor's normalize(0,1)."""
    m, keep = _material(texture)
    H, W, Cn = keep.shape
    out = np.zeros((Cn, H, W), np.float32)
    lib().oracle_normalized_texture(C.byref(m), out.ctypes.data)
    return out


def linear_atXY(texture, fx, fy, c) -> float:
    m, keep = _material(texture)
    return float(lib().oracle_linear_atXY(C.byref(m), float(fx), float(fy), int(c)))


def cast_u32(f) -> int:
    return int(lib().oracle_cast_u32(float(np.float32(f))))
