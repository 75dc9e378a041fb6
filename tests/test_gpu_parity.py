"""GPU parity of the frame path: librz (through its C ABI) against the CPU oracle on
the same seeded inputs.  Face ids, depth bits and colour bytes must match exactly."""
import numpy as np
import pytest

from rasteriser_amd import scenes
from tests.parity import check

pytestmark = pytest.mark.gpu


def test_uv_sphere_512(ctx):
    scene, args = scenes.uv_sphere_scene()
    gpu, ref = check(ctx, scene, args, label="uv sphere 512")
    assert (gpu[2] >= 0).sum() > 50000


@pytest.mark.parametrize("flat", [False, True])
def test_icosphere_1080p(ctx, flat):
    scene, args = scenes.icosphere_scene(flat=flat)
    gpu, ref = check(ctx, scene, args, label=f"icosphere 1080p flat={flat}")
    assert (gpu[2] >= 0).sum() > 300000


@pytest.mark.parametrize("tile,cap", [(8, 64), (16, 512), (32, 2048), (16, 1), (16, 4)])
def test_tuning_and_overflow(ctx, tile, cap):
    """Every tile size, and tiny bin capacities that force the overflow path."""
    scene, args = scenes.icosphere_scene(flat=False, subdivisions=4, width=640, height=360)
    gpu, ref = check(ctx, scene, args, tuning=(tile, cap), label=f"tile {tile} cap {cap}")
    if cap <= 4:
        assert gpu[3]["overflow_tiles"] > 0


def test_not_cleared_buffers(ctx):
    """draw_frame does not clear: the depth test reads the existing depth and pixels it
    does not win keep their old colour (drawing.cpp:119, renderer.cpp:107-108)."""
    scene, args = scenes.uv_sphere_scene(width=300, height=200)
    W, H = 300, 200
    rng = np.random.default_rng(7)
    frame = rng.integers(0, 256, (3, H, W), dtype=np.uint8)
    depth = rng.uniform(0.85, 1.0, (H, W)).astype(np.float32)
    depth[::7, ::5] = np.float32(2.0)
    check(ctx, scene, args, frame=frame, depth=depth, label="not cleared")


def test_wind_clockwise(ctx):
    scene, args = scenes.uv_sphere_scene(width=256, height=256)
    args.wind_clockwise = 1
    check(ctx, scene, args, label="wind clockwise")


def test_screen_soup_small(ctx):
    scene, args = scenes.screen_soup_scene(n=200_000, size=1024)
    gpu, ref = check(ctx, scene, args, label="screen soup 200k")
    assert gpu[3]["front_triangles"] > 80_000


def test_repeatable(ctx):
    """Two renders of the same frame give identical visibility (order-free atomics)."""
    scene, args = scenes.icosphere_scene(subdivisions=5, width=800, height=600)
    from tests.parity import render_gpu
    a = render_gpu(ctx, scene, args)
    b = render_gpu(ctx, scene, args)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
