"""The C ABI boundary (include/rz.h) — CPU only: librz.so loads, exports every entry
point the header declares, and the ctypes records match the C layouts the header
defines (which in turn match the reference's Triangle and Light)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import rasteriser_amd as rz
from rasteriser_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rz.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rz_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = rz.load_library()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"librz.so does not export {n}"
    assert set(names) == set(abi.EXPORTED)
    assert lib.rz_abi_version() == 2


def test_record_layouts_match_the_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "rz.h"
#define P(T) printf(#T " %zu\\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\\n", offsetof(T, f));
int main(void) {
  P(rz_triangle) P(rz_light) P(rz_material) P(rz_args) P(rz_mesh_desc) P(rz_matrices) P(rz_frame_stats)
  O(rz_triangle, material) O(rz_light, trans_dir) O(rz_material, texels) O(rz_args, z_far)
  O(rz_mesh_desc, raster_space) O(rz_frame_stats, ms_frame) O(rz_frame_stats, raster_pixels)
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    types = {"rz_triangle": abi.Triangle, "rz_light": abi.Light, "rz_material": abi.Material,
             "rz_args": abi.Args, "rz_mesh_desc": abi.MeshDesc, "rz_matrices": abi.Matrices,
             "rz_frame_stats": abi.FrameStats}
    for name, T in types.items():
        assert int(out[name]) == C.sizeof(T), name
    assert int(out["rz_triangle"]) == 40 and int(out["rz_light"]) == 40   # reference Triangle / Light
    for key, (T, f) in {"rz_triangle.material": (abi.Triangle, "material"),
                        "rz_light.trans_dir": (abi.Light, "trans_dir"),
                        "rz_material.texels": (abi.Material, "texels"),
                        "rz_args.z_far": (abi.Args, "z_far"),
                        "rz_mesh_desc.raster_space": (abi.MeshDesc, "raster_space"),
                        "rz_frame_stats.ms_frame": (abi.FrameStats, "ms_frame"),
                        "rz_frame_stats.raster_pixels": (abi.FrameStats, "raster_pixels")}.items():
        assert int(out[key]) == getattr(T, f).offset, key


def test_default_args_are_the_references():
    lib = rz.load_library()
    a = abi.Args()
    lib.rz_default_args(C.byref(a), 540, 304)
    assert (a.image_width, a.image_height, a.scale, a.flat, a.wind_clockwise) == (540, 304, 1.0, 0, 0)
    assert (a.fovy_degrees, round(a.z_near, 6), a.z_far) == (45.0, 0.1, 6.0)
    assert a.aspect_ratio == C.c_float(540 / 304).value
    b = rz.make_args(540, 304)
    assert bytes(a) == bytes(b)


def test_status_strings():
    lib = rz.load_library()
    for s in range(6):
        assert lib.rz_status_string(s)
    assert lib.rz_status_string(2) == b"index out of bounds"


def test_create_without_device_fails_cleanly():
    """Without a HIP device rz_create returns RZ_ERR_NO_DEVICE instead of crashing."""
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU is visible here")
    with pytest.raises(rz.RzError) as e:
        rz.Context(0)
    assert e.value.status == abi.RZ_ERR_NO_DEVICE


def test_null_arguments_are_rejected():
    lib = rz.load_library()
    assert lib.rz_frame_matrices(None, None) == abi.RZ_ERR_INVALID_ARG
    assert lib.rz_create(0, None) == abi.RZ_ERR_INVALID_ARG
    assert lib.rz_sync(None) == abi.RZ_ERR_INVALID_ARG
    assert lib.rz_set_tuning(None, 16, 16) == abi.RZ_ERR_INVALID_ARG
