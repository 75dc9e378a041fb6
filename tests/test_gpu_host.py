"""The C++ host layer on the GPU: the `renderer` command line (the reference's flags and
program flow over librz, via the C++ draw_frame) renders an OBJ it loads itself, and
its frame.png must equal the oracle's frame of the same scene byte for byte."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from rasteriser_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "rasteriser_amd", "_lib")
RENDERER = os.path.join(LIBDIR, "renderer")

pytestmark = pytest.mark.gpu


def write_obj(path, scene):
    """The scene's one material (Kd, and map_Kd when it has a texture, saved as PNG
    beside the OBJ), its vertices, normals and uvs (%.9g: decimals that read back
    exactly), and faces v/vt/vn (v//vn for a face without uvs)."""
    mat = scene.materials[0]
    with open(path + ".mtl", "w") as f:
        f.write("newmtl m0\nKd %.9g %.9g %.9g\n" % tuple(float(c) for c in mat.diffuse))
        if mat.texture is not None:
            tex = np.ascontiguousarray(mat.texture)
            H, W, C = tex.shape
            L = C_lib()
            assert L.rzh_save_png(os.path.join(os.path.dirname(path), "tex.png").encode(), tex.ctypes.data,
                                  W, H, C) == 0
            f.write("map_Kd tex.png\n")
    with open(path, "w") as f:
        f.write(f"mtllib {os.path.basename(path)}.mtl\n")
        for v in scene.vertices:
            f.write("v %.9g %.9g %.9g\n" % tuple(float(x) for x in v[:3]))
        for n in scene.normals:
            f.write("vn %.9g %.9g %.9g\n" % tuple(float(x) for x in n[:3]))
        for t in scene.uvs:
            f.write("vt %.9g %.9g\n" % tuple(float(x) for x in t[:2]))
        f.write("usemtl m0\n")
        for t in scene.faces:
            v, n, u = t[0:3] + 1, t[3:6] + 1, t[6:9] + 1
            if (t[6:9] >= 0).all():
                f.write(f"f {v[0]}/{u[0]}/{n[0]} {v[1]}/{u[1]}/{n[1]} {v[2]}/{u[2]}/{n[2]}\n")
            else:
                f.write(f"f {v[0]}//{n[0]} {v[1]}//{n[1]} {v[2]}//{n[2]}\n")


def C_lib():
    L = C.CDLL(os.path.join(LIBDIR, "librz_host.so"))
    L.rzh_save_png.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
    return L


def write_lights(path, lights):
    with open(path, "w") as f:
        for L in lights:
            f.write(",".join("%.9g" % float(x) for x in L[:7]) + "\n")


def load_png(path):
    L = C.CDLL(os.path.join(LIBDIR, "librz_host.so"))
    L.rzh_load_image.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p]
    W, H, Ch = C.c_int(), C.c_int(), C.c_int()
    assert L.rzh_load_image(path.encode(), C.byref(W), C.byref(H), C.byref(Ch), None) == 0
    out = np.zeros((H.value, W.value, Ch.value), np.uint8)
    assert L.rzh_load_image(path.encode(), C.byref(W), C.byref(H), C.byref(Ch), out.ctypes.data) == 0
    return out


@pytest.mark.parametrize("case", ["smooth", "flat", "textured", "textured_rgb", "spin_one_frame", "spin_frames"])
def test_renderer_cli_frame_matches_oracle(tmp_path, case):
    """UV sphere (smooth, flat), a textured icosphere whose RGBA (or RGB) texture the
    renderer loads as a PNG named by map_Kd, and the spin loop (-s: the mesh uploaded
    once, one captured frame replayed with each frame's rotation, ry advanced by the
    elapsed time) for one frame, drawn before its first rotation step, and for four
    frames: the last one at the angles the renderer prints, its frame rate drawn at
    (5, 5) like the reference's (renderer.cpp:114), the rest of it the oracle's."""
    if not os.path.exists(RENDERER):
        pytest.skip("renderer not built")
    flat = case == "flat"
    if case.startswith("textured"):
        scene, args = scenes.icosphere_scene(subdivisions=4, width=480, height=360, texture_size=128)
        if case == "textured_rgb":   # 3 channels: the normalisation spans R, G and B only
            m = scene.materials[0]
            scene.materials[0] = type(m)(m.diffuse, np.ascontiguousarray(m.texture[:, :, :3]), m.normal_map)
    else:
        scene, args = scenes.uv_sphere_scene()
    args.flat = 1 if flat else 0
    obj = str(tmp_path / "sphere.obj")
    lights = str(tmp_path / "lights.csv")
    write_obj(obj, scene)
    write_lights(lights, scene.lights)
    rot = []
    for flag, angle in zip(("--rx", "--ry", "--rz"), args.tait_bryan_angles):
        rot += [flag, "%.9g" % float(angle)]
    cmd = [RENDERER, "-o", obj, "--mats-dir", str(tmp_path) + "/", "-l", lights, "-x", str(args.image_width),
           "-y", str(args.image_height), "--fovy", "60", "--near", "0.1", "--far", "100", "--out-dir",
           str(tmp_path)] + rot + (["-f"] if flat else [])
    if case.startswith("spin"):
        cmd += ["-s", "--frames", "1" if case == "spin_one_frame" else "4"]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = load_png(str(tmp_path / "frame.png"))[:, :, :3]
    if case.startswith("spin"):
        line = [l for l in r.stdout.splitlines() if l.startswith("last frame angles:")][0]
        args.tait_bryan_angles[:] = [float(np.float32(x)) for x in line.split(":")[1].split()]
    of, od, oids, counts, _ = oracle.draw_frame(scene, args)
    want = np.moveaxis(of, 0, 2)
    assert (oids >= 0).sum() > (30_000 if case.startswith("textured") else 50_000)
    if case == "spin_frames":
        rates = [l.split(":")[1].split()[0] for l in r.stdout.splitlines() if l.startswith("frame ")]
        assert len(rates) == 4
        box = (slice(5, 21), slice(5, 5 + 12 * len(rates[2])))   # the last frame shows frame 2's rate
        inside = got[box]
        assert ((inside == 255).all(axis=2) | (inside == 0).all(axis=2)).all()
        assert (inside == 255).all(axis=2).sum() > 20 * len(rates[2])
        got, want = got.copy(), want.copy()
        got[box] = 0
        want[box] = 0
    assert np.array_equal(got, want), "frame.png differs from the oracle's frame"
