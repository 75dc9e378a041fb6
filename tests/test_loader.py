"""Host layer I/O — CPU only: the OBJ/MTL loader against the reference's own
tinyobjloader (golden vectors it wrote, and the live harness where the reference tree
is present), the lights CSV reader, the PNG codec and the command line."""
import ctypes as C
import math
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "rasteriser_amd", "_lib")
RENDERER = os.path.join(LIBDIR, "renderer")
SAMPLES = "/root/reference/sampledata"
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from obj_cases import CASES, write_case  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden", "obj_cases.npz")


def need_renderer():
    if not os.path.exists(RENDERER):
        pytest.skip("renderer not built")


def dump_mesh(tmp, obj, mats_dir, lights):
    prefix = os.path.join(tmp, "dump")
    cmd = [RENDERER, "-l", lights, "--dump-mesh", prefix]
    if obj:
        cmd += ["-o", obj, "--mats-dir", mats_dir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = {k: np.fromfile(f"{prefix}.{k}.{'i32' if k == 'faces' else 'f32'}",
                          np.int32 if k == "faces" else np.float32)
           for k in ("vertices", "normals", "uvs", "faces", "lights")}
    out["materials"] = open(prefix + ".materials.txt").read()
    return out


def write_lights(tmp, text="1,-0.5,-1,512,1,1,0.9\n"):
    p = os.path.join(tmp, "lights.csv")
    with open(p, "w") as f:
        f.write(text)
    return p


def write_ppm(path, w=2, h=2):
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h) + bytes(range(3 * w * h)))


@pytest.mark.parametrize("case", sorted(CASES))
def test_obj_cases_match_tinyobjloader(case, tmp_path):
    need_renderer()
    obj, mats = write_case(case, str(tmp_path))
    write_ppm(os.path.join(tmp_path, "red.png"))
    got = dump_mesh(str(tmp_path), obj, mats, write_lights(str(tmp_path)))
    g = np.load(GOLDEN)
    for k in ("vertices", "normals", "uvs"):
        assert np.array_equal(got[k].view(np.uint32), g[f"{case}.{k}"].view(np.uint32)), (case, k)
    assert np.array_equal(got["faces"], g[f"{case}.faces"]), case
    # materials: same count and diffuse colours (texture names resolve to loaded images)
    ref_m = [l.split() for l in str(g[f"{case}.materials"]).splitlines()]
    got_m = [l.split() for l in got["materials"].splitlines()]
    assert len(ref_m) == len(got_m)
    for r, m in zip(ref_m, got_m):
        assert [np.float32(x) for x in r[:3]] == [np.float32(x) for x in m[:3]]
        assert (r[3] != "-") == (int(m[3]) > 0)


def test_polygon_case_known_answers():
    """Known answers independent of any parser: fan order, negative indices and
    material ids in order of appearance (-1 for an unknown name)."""
    g = np.load(GOLDEN)
    f = g["polygons.faces"].reshape(-1, 10)
    assert f[:3, 0:3].tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4]]     # pentagon fan
    assert f[:3, 6:9].tolist() == [[0, 1, 2], [0, 2, 0], [0, 0, 1]]
    assert f[3, 0:3].tolist() == [0, 1, 5] and f[3, 6:9].tolist() == [0, 1, 2]   # -6 -5 -1
    assert f[:4, 9].tolist() == [0, 0, 0, 0]                            # red
    assert f[4, 9] == 1 and f[4, 3:6].tolist() == [0, 0, 0] and f[4, 6] == -1   # blue, v//vn
    assert f[5:7, 9].tolist() == [-1, -1]                               # an unknown name
    assert f[7, 0:3].tolist() == [5, 0, 1] and len(f) == 8


def test_shape_dropped_before_group_like_tinyobjloader():
    """usemtl directly before g flushes the last face group and the loader then drops
    the whole shape (tiny_obj_loader.h:1590-1600): only the later shapes survive."""
    g = np.load(GOLDEN)
    f = g["dropped.faces"].reshape(-1, 10)
    assert f[:, 0:3].tolist() == [[5, 0, 1], [2, 3, 4], [3, 4, 5]]
    assert f[:, 9].tolist() == [0, 0, 0]


def test_sample_models_match_live_tinyobjloader(tmp_path):
    need_renderer()
    exe = os.path.join(ROOT, "oracle", "_ref", "tinyobj_ref")
    if not (os.path.exists(exe) and os.path.isdir(SAMPLES)):
        pytest.skip("reference tree or its tinyobjloader harness absent")
    lights = os.path.join(SAMPLES, "threepoint.csv")
    for name, counts in (("Suzanne.obj", (507, 507, 596, 968)), ("plane.obj", (4, 1, 0, 2))):
        obj = os.path.join(SAMPLES, name)
        got = dump_mesh(str(tmp_path), obj, SAMPLES + "/", lights)
        ref = os.path.join(tmp_path, "ref")
        subprocess.run([exe, obj, SAMPLES + "/", ref], check=True, capture_output=True)
        for k in ("vertices", "normals", "uvs"):
            assert np.array_equal(got[k], np.fromfile(f"{ref}.{k}.f32", np.float32)), (name, k)
        assert np.array_equal(got["faces"], np.fromfile(ref + ".faces.i32", np.int32)), name
        nv, nn, nt, nf = counts
        assert (got["vertices"].size // 3, got["normals"].size // 3, got["uvs"].size // 2,
                got["faces"].size // 10) == (nv, nn, nt, nf)


def test_lights_csv(tmp_path):
    need_renderer()
    text = "1,-0.5,-1,512,1,1,0.9\n-1,0,-1,256,1,1,0.5\n\n1,0,1,100,0,0,1\n"
    got = dump_mesh(str(tmp_path), None, None, write_lights(str(tmp_path), text))
    L = got["lights"].reshape(-1, 10)
    assert L.shape[0] == 3
    assert L[:, :7].tolist() == [[1, -0.5, -1, 512, 1, 1, np.float32(0.9)], [-1, 0, -1, 256, 1, 1, 0.5],
                                 [1, 0, 1, 100, 0, 0, 1]]


def test_cli_reports_bad_arguments():
    need_renderer()
    r = subprocess.run([RENDERER, "-x", "64"], capture_output=True, text=True)
    assert r.returncode == 1 and "Error:" in r.stderr                  # -l is required
    r = subprocess.run([RENDERER, "-l", "x.csv", "--width", "abc"], capture_output=True, text=True)
    assert r.returncode == 1 and "Error:" in r.stderr
    r = subprocess.run([RENDERER, "-l", "x.csv", "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1


def host_lib():
    path = os.path.join(LIBDIR, "librz_host.so")
    if not os.path.exists(path):
        pytest.skip("librz_host not built")
    L = C.CDLL(path)
    L.rzh_load_image.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p]
    L.rzh_save_png.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.rzh_save_depth_png.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int]
    return L


def load(L, path):
    W, H, Ch = C.c_int(), C.c_int(), C.c_int()
    assert L.rzh_load_image(path.encode(), C.byref(W), C.byref(H), C.byref(Ch), None) == 0
    out = np.zeros((H.value, W.value, Ch.value), np.uint8)
    assert L.rzh_load_image(path.encode(), C.byref(W), C.byref(H), C.byref(Ch), out.ctypes.data) == 0
    return out


def test_png_codec_against_pillow(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    L = host_lib()
    rng = np.random.default_rng(9)
    for ch, mode in ((3, "RGB"), (4, "RGBA"), (1, "L")):
        img = rng.integers(0, 256, (37, 53, ch), dtype=np.uint8)
        # Pillow writes with adaptive filters: decoding exercises every filter type
        p = str(tmp_path / f"pil{ch}.png")
        Image.fromarray(img[:, :, 0] if ch == 1 else img, mode).save(p, optimize=True)
        assert np.array_equal(load(L, p), img)
        q = str(tmp_path / f"rz{ch}.png")
        assert L.rzh_save_png(q.encode(), img.ctypes.data, 53, 37, ch) == 0
        back = np.asarray(Image.open(q))
        assert np.array_equal(back.reshape(img.shape), img)
    if os.path.exists(os.path.join(SAMPLES, "SuzanneTex.png")):
        p = os.path.join(SAMPLES, "SuzanneTex.png")
        assert np.array_equal(load(L, p), np.asarray(Image.open(p)))


def test_depth_png_matches_cimg_vectors(tmp_path):
    """depth.png = depth_buffer.normalize(0,255).save(...) (renderer.cpp:93): the bytes the
    reference's own CImg wrote for the same depth buffers (tests/golden/cimg_vectors.npz:
    cleared 1.0 around drawn depths, a constant buffer, -0.0 and denormals, a range a few
    ulps wide)."""
    Image = pytest.importorskip("PIL.Image")
    L = host_lib()
    g = np.load(os.path.join(ROOT, "tests", "golden", "cimg_vectors.npz"))
    names = [k[:-3] for k in g.files if k.startswith("depth_") and k.endswith("_in")]
    assert len(names) == 4
    for name in names:
        d = np.ascontiguousarray(g[name + "_in"])
        H, W = d.shape
        p = str(tmp_path / f"{name}.png")
        assert L.rzh_save_depth_png(p.encode(), d.ctypes.data, W, H) == 0
        assert np.array_equal(np.asarray(Image.open(p)), g[name + "_png"]), name


def test_depth_png_matches_live_cimg(tmp_path):
    """The same against the reference's CImg writer itself, on a rendered depth buffer of
    the oracle (oracle/_ref, built where the reference tree is present)."""
    Image = pytest.importorskip("PIL.Image")
    import oracle
    from rasteriser_amd import scenes
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    L = host_lib()
    scene, args = scenes.uv_sphere_scene(width=96, height=64)
    _, d, _, _, _ = oracle.draw_frame(scene, args)
    for k, depth in enumerate([d, np.where(d < 1, d * np.float32(-3), d).astype(np.float32)]):
        depth = np.ascontiguousarray(depth)
        ours, theirs = str(tmp_path / f"ours{k}.png"), str(tmp_path / f"cimg{k}.png")
        assert L.rzh_save_depth_png(ours.encode(), depth.ctypes.data, 96, 64) == 0
        assert ref.cimg_ref_save_depth_png(theirs.encode(), depth.ctypes.data, 96, 64) == 0
        a, b = np.asarray(Image.open(ours)), np.asarray(Image.open(theirs))
        assert a.shape == (64, 96) and np.array_equal(a, b), k
