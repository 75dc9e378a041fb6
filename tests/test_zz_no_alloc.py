"""rz_draw_frame_noalloc (no reference counterpart: the entry point for callers that
cannot allocate device memory) against rz_draw_frame on the same inputs, through the C
ABI: face ids, depth bits and colour bytes identical, on two calls in a row (the second
smaller, so stale contents of the shared working array would show), and the working-set
limit rejected on the host before any launch."""
import numpy as np
import pytest

import rasteriser_amd as rz
from rasteriser_amd import scenes

pytestmark = pytest.mark.gpu


def _outputs(c, args):
    W, H = args.image_width, args.image_height
    return c.empty((3, H, W), np.uint8), c.empty((H, W), np.float32), c.empty((H, W), np.int32)


def _same(a, b, label):
    (fa, da, ia), (fb, db, ib) = a, b
    assert np.array_equal(ia, ib), f"{label}: face ids differ"
    assert np.array_equal(da.view(np.uint32), db.view(np.uint32)), f"{label}: depth differs"
    assert np.array_equal(fa, fb), f"{label}: colour differs"


def _frame(c, mesh, scene, args, noalloc, flags):
    f, d, i = _outputs(c, args)
    try:
        if not flags & rz.RZ_FLAG_CLEARED:   # not-cleared buffers: the same seeded contents
            rng = np.random.default_rng(11)
            H, W = args.image_height, args.image_width
            f.upload(rng.integers(0, 256, (3, H, W), dtype=np.uint8))
            d.upload(rng.uniform(0.9, 1.0, (H, W)).astype(np.float32))
        i.upload(np.full(i.shape, 9, np.int32))
        lights = scene.lights_copy()
        if noalloc:
            c.draw_noalloc(mesh, lights, args, f.ptr, d.ptr, i.ptr, flags)
        else:
            c.draw_device(mesh, lights, args, f.ptr, d.ptr, i.ptr, flags)
        c.sync()
        return f.download(), d.download(), i.download()
    finally:
        for b in (f, d, i):
            b.close()


def test_noalloc_matches_draw_frame_twice():
    cases = [
        ("icosphere 1080p", scenes.icosphere_scene(), rz.RZ_FLAG_CLEARED),
        ("uv sphere 300x200, buffers not cleared", scenes.uv_sphere_scene(width=300, height=200), 0),
    ]
    with rz.Context(0) as c:
        for label, (scene, args), flags in cases:
            print("no-alloc case:", label, flush=True)
            mesh = c.upload(scene)
            want = _frame(c, mesh, scene, args, False, flags)
            got = _frame(c, mesh, scene, args, True, flags)
            _same(got, want, label)
            assert (got[2] >= 0).sum() > 1000
            mesh.close()


def test_noalloc_overflow_capacity_and_split_pipeline():
    """A context capacity small enough to overflow bins, and the separate-shade pipeline
    (visibility words in the array): same bits as rz_draw_frame with the same tuning."""
    scene, args = scenes.icosphere_scene(subdivisions=4, width=640, height=360)
    with rz.Context(0) as c:
        mesh = c.upload(scene)
        for fused, cap in ((True, 4), (False, 0), (False, 4)):
            print("no-alloc case: fused", fused, "cap", cap, flush=True)
            c.set_fused_shade(fused)
            c.set_tuning(16, cap)
            want = _frame(c, mesh, scene, args, False, rz.RZ_FLAG_CLEARED)
            got = _frame(c, mesh, scene, args, True, rz.RZ_FLAG_CLEARED)
            _same(got, want, f"fused {fused} cap {cap}")
        mesh.close()


def test_noalloc_rejects_oversized_frames():
    """A working set beyond the array (8192x8192: the tile counters alone exceed it) is
    refused on the host with a clear error, and the context stays usable."""
    scene, args = scenes.uv_sphere_scene(width=8192, height=8192)
    small, sargs = scenes.uv_sphere_scene(width=64, height=64)
    with rz.Context(0) as c:
        mesh = c.upload(scene)
        f, d, i = _outputs(c, args)   # real buffers, though the refusal comes before any launch
        with pytest.raises(rz.RzError) as e:
            c.draw_noalloc(mesh, scene.lights_copy(), args, f.ptr, d.ptr, i.ptr)
        assert e.value.status == rz.abi.RZ_ERR_INVALID_ARG and "working memory" in str(e.value)
        for b in (f, d, i):
            b.close()
        m2 = c.upload(small)
        _same(_frame(c, m2, small, sargs, True, rz.RZ_FLAG_CLEARED),
              _frame(c, m2, small, sargs, False, rz.RZ_FLAG_CLEARED), "after a refusal")
        mesh.close()
        m2.close()
