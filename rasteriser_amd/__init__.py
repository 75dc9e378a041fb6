"""rasteriser_amd — the rasteriser's frame path on MI355X (gfx950).

The product is ``librz.so`` (HIP kernels + the C ABI of ``include/rz.h``); this module
is a thin ctypes host over that ABI for Python callers, tests and the benchmark.  It
never computes a frame itself: when ``librz.so`` is missing it raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

from . import abi
from .abi import (RZ_FLAG_ASYNC, RZ_FLAG_CLEARED, RZ_FLAG_DEVICE_BUFFERS, RZ_FLAG_MLP_SHADE, RZ_MLP_PARAMS,
                  RZ_MLP_SAMPLE_FLOATS, Args, BatchItem, FrameStats, Matrices, make_args)
from .scene import MaterialSpec, Scene, make_faces, make_lights

__all__ = ["Context", "Mesh", "RzError", "Scene", "MaterialSpec", "make_faces", "make_lights",
           "make_args", "library_path", "load_library", "frame_matrices",
           "RZ_FLAG_CLEARED", "RZ_FLAG_DEVICE_BUFFERS", "RZ_FLAG_ASYNC", "RZ_FLAG_MLP_SHADE",
           "RZ_MLP_PARAMS", "RZ_MLP_SAMPLE_FLOATS"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def library_path() -> str:
    return os.path.join(_HERE, "_lib", "librz.so")


def load_library():
    """Load librz.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"librz.so is not built ({path}); run __graft_entry__.build()")
        _LIB = abi.declare(C.CDLL(path))
    return _LIB


class RzError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        super().__init__(f"rz status {status}: {message}")


def _check(status, ctx=None):
    if status != abi.RZ_OK:
        lib = load_library()
        msg = lib.rz_last_error(ctx).decode() if ctx else lib.rz_status_string(status).decode()
        raise RzError(status, msg)


def frame_matrices(args: Args) -> dict:
    """The five matrices librz builds for ``args`` (float32, column-major (4,4) as [col][row])."""
    m = Matrices()
    _check(load_library().rz_frame_matrices(C.byref(args), C.byref(m)))
    return {k: np.frombuffer(bytes(getattr(m, k)), np.float32).reshape(4, 4).copy()
            for k in ("model", "view", "modelview", "camera", "normal")}


class Mesh:
    """A device-resident mesh (rz_mesh_create)."""

    def __init__(self, ctx: "Context", scene: Scene):
        self.ctx = ctx
        self.num_faces = scene.num_faces
        self.raster_space = scene.raster_space
        desc, keep = scene.desc()
        h = C.c_void_p()
        _check(ctx.lib.rz_mesh_create(ctx.handle, C.byref(desc), C.byref(h)), ctx.handle)
        del keep
        self.handle = h
        ctx._children.add(self)

    def close(self):
        if self.handle:
            if self.ctx.handle:
                self.ctx.lib.rz_mesh_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FrameGraph:
    """A captured run of frames (rz_graph_create); launch() enqueues them all."""

    def __init__(self, ctx: "Context", handle, frames: int):
        self.ctx, self.handle, self.frames = ctx, handle, frames
        ctx._children.add(self)

    def launch(self):
        _check(self.ctx.lib.rz_graph_launch(self.handle), self.ctx.handle)

    def close(self):
        if self.handle:
            if self.ctx.handle:
                self.ctx.lib.rz_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpinLoop:
    """The spin loop's frame captured once (rz_spin_create); frame(angles) replays it
    with that rotation."""

    def __init__(self, ctx: "Context", handle):
        self.ctx, self.handle = ctx, handle
        ctx._children.add(self)

    def frame(self, angles):
        a = (C.c_float * 3)(*[float(x) for x in angles])
        _check(self.ctx.lib.rz_spin_frame(self.handle, a), self.ctx.handle)

    def close(self):
        if self.handle:
            if self.ctx.handle:
                self.ctx.lib.rz_spin_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation of the context's device holding an array of `shape`/`dtype`."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(ctx.lib.rz_device_alloc(ctx.handle, self.nbytes, C.byref(p)), ctx.handle)
        self.ptr = int(p.value)
        ctx._children.add(self)

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, self.dtype)
        assert host.nbytes == self.nbytes
        _check(self.ctx.lib.rz_copy_to_device(self.ctx.handle, C.c_void_p(self.ptr), host.ctypes.data,
                                              self.nbytes), self.ctx.handle)

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(self.ctx.lib.rz_copy_to_host(self.ctx.handle, out.ctypes.data, C.c_void_p(self.ptr),
                                            self.nbytes), self.ctx.handle)
        return out

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy of `other`'s bytes into this array (rz_copy_device)."""
        assert other.nbytes == self.nbytes
        _check(self.ctx.lib.rz_copy_device(self.ctx.handle, C.c_void_p(self.ptr), C.c_void_p(other.ptr),
                                           self.nbytes), self.ctx.handle)

    def close(self):
        if self.ptr:
            if self.ctx.handle:
                self.ctx.lib.rz_device_free(self.ctx.handle, C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One rz_ctx: a HIP device and its stream."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        self.handle = None
        # meshes, graphs and device arrays of this context, released before it is
        self._children = weakref.WeakSet()
        h = C.c_void_p()
        _check(self.lib.rz_create(int(device), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            for child in list(self._children):
                child.close()
            self.lib.rz_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_tuning(self, tile_size: int, bin_capacity: int):
        _check(self.lib.rz_set_tuning(self.handle, tile_size, bin_capacity), self.handle)

    def tuning(self):
        t, c = C.c_int32(), C.c_int32()
        _check(self.lib.rz_get_tuning(self.handle, C.byref(t), C.byref(c)), self.handle)
        return t.value, c.value

    def set_fused_shade(self, on: bool):
        """rz_set_fused_shade: shade in the raster pass (default) or as a separate pass."""
        _check(self.lib.rz_set_fused_shade(self.handle, int(bool(on))), self.handle)

    @property
    def fused_shade(self) -> bool:
        return bool(self.lib.rz_get_fused_shade(self.handle))

    def set_batch_persistent(self, on: bool):
        """rz_set_batch_persistent: a batch frame as one persistent launch (default) or five."""
        _check(self.lib.rz_set_batch_persistent(self.handle, int(bool(on))), self.handle)

    @property
    def batch_persistent(self) -> bool:
        return bool(self.lib.rz_get_batch_persistent(self.handle))

    def set_profiling(self, on: bool):
        _check(self.lib.rz_set_profiling(self.handle, int(bool(on))), self.handle)

    def stream(self) -> int:
        return int(self.lib.rz_stream(self.handle) or 0)

    def sync(self):
        _check(self.lib.rz_sync(self.handle), self.handle)

    def empty(self, shape, dtype) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def upload(self, scene: Scene) -> Mesh:
        return Mesh(self, scene)

    def draw(self, mesh: Mesh, lights: np.ndarray, args: Args, frame=None, depth=None,
             ids: bool = False, stats: bool = False, flags: int = 0):
        """draw_frame into host buffers.  With frame/depth None the buffers start cleared
        (frame 0, depth 1.0, renderer.cpp:85-86); otherwise they are updated in place.
        ``lights`` (L,10) float32 gets trans_dir written back.  Returns
        (frame [3,H,W] u8, depth [H,W] f32, ids [H,W] i32 or None, stats dict or None).
        ``flags`` adds e.g. RZ_FLAG_MLP_SHADE."""
        W, H = args.image_width, args.image_height
        flags = int(flags) & ~(RZ_FLAG_CLEARED | RZ_FLAG_DEVICE_BUFFERS | RZ_FLAG_ASYNC)
        if frame is None and depth is None:
            flags |= RZ_FLAG_CLEARED
            frame = np.zeros((3, H, W), np.uint8)
            depth = np.ones((H, W), np.float32)
        assert frame.dtype == np.uint8 and frame.shape == (3, H, W) and frame.flags.c_contiguous
        assert depth.dtype == np.float32 and depth.shape == (H, W) and depth.flags.c_contiguous
        lights = _lights(lights)
        idbuf = np.empty((H, W), np.int32) if ids else None
        st = FrameStats() if stats else None
        _check(self.lib.rz_draw_frame(self.handle, mesh.handle, lights.ctypes.data, lights.shape[0],
                                      C.byref(args), frame.ctypes.data, depth.ctypes.data,
                                      idbuf.ctypes.data if ids else None, flags,
                                      C.byref(st) if stats else None), self.handle)
        return frame, depth, idbuf, (st.as_dict() if stats else None)

    def draw_device(self, mesh: Mesh, lights: np.ndarray, args: Args, frame_ptr: int,
                    depth_ptr: int, ids_ptr: int = 0, flags: int = RZ_FLAG_CLEARED,
                    stats: bool = False):
        """draw_frame into device buffers (RZ_FLAG_DEVICE_BUFFERS is added)."""
        lights = _lights(lights)
        st = FrameStats() if stats else None
        _check(self.lib.rz_draw_frame(self.handle, mesh.handle, lights.ctypes.data, lights.shape[0],
                                      C.byref(args), C.c_void_p(frame_ptr), C.c_void_p(depth_ptr),
                                      C.c_void_p(ids_ptr) if ids_ptr else None,
                                      flags | RZ_FLAG_DEVICE_BUFFERS,
                                      C.byref(st) if stats else None), self.handle)
        return st.as_dict() if stats else None

    def draw_noalloc(self, mesh: Mesh, lights: np.ndarray, args: Args, frame_ptr: int, depth_ptr: int,
                     ids_ptr: int = 0, flags: int = RZ_FLAG_CLEARED):
        """rz_draw_frame_noalloc: draw_frame into device buffers with librz's fixed
        per-GPU working array (no device allocation)."""
        lights = _lights(lights)
        _check(self.lib.rz_draw_frame_noalloc(self.handle, mesh.handle, lights.ctypes.data, lights.shape[0],
                                              C.byref(args), C.c_void_p(frame_ptr), C.c_void_p(depth_ptr),
                                              C.c_void_p(ids_ptr) if ids_ptr else None,
                                              flags | RZ_FLAG_DEVICE_BUFFERS), self.handle)

    def graph(self, mesh: Mesh, lights: np.ndarray, args: Args, frame_ptr: int, depth_ptr: int,
              frames: int, ids_ptr: int = 0, flags: int = RZ_FLAG_CLEARED, ry_step: float = 0.0):
        """rz_graph_create: `frames` consecutive frames captured in one HIP graph."""
        lights = _lights(lights)
        h = C.c_void_p()
        _check(self.lib.rz_graph_create(self.handle, mesh.handle, lights.ctypes.data, lights.shape[0],
                                        C.byref(args), C.c_void_p(frame_ptr), C.c_void_p(depth_ptr),
                                        C.c_void_p(ids_ptr) if ids_ptr else None,
                                        flags | RZ_FLAG_DEVICE_BUFFERS, int(frames), float(ry_step),
                                        C.byref(h)), self.handle)
        return FrameGraph(self, h, frames)

    def spin(self, mesh: Mesh, lights: np.ndarray, args: Args, frame_ptr: int, depth_ptr: int,
             ids_ptr: int = 0, flags: int = RZ_FLAG_CLEARED) -> SpinLoop:
        """rz_spin_create over device buffers."""
        lights = _lights(lights)
        h = C.c_void_p()
        _check(self.lib.rz_spin_create(self.handle, mesh.handle, lights.ctypes.data, lights.shape[0],
                                       C.byref(args), C.c_void_p(frame_ptr), C.c_void_p(depth_ptr),
                                       C.c_void_p(ids_ptr) if ids_ptr else None,
                                       flags | RZ_FLAG_DEVICE_BUFFERS, C.byref(h)), self.handle)
        return SpinLoop(self, h)

    def draw_text(self, frame_ptr: int, width: int, height: int, x: int, y: int, text: str,
                  fg=(255, 255, 255), bg=(0, 0, 0)):
        """rz_draw_text into a planar device frame (asynchronous on the context stream)."""
        f = (C.c_uint8 * 3)(*fg)
        b = (C.c_uint8 * 3)(*bg)
        _check(self.lib.rz_draw_text(self.handle, C.c_void_p(frame_ptr), int(width), int(height), int(x), int(y),
                                     text.encode(), f, b), self.handle)

    def check_quotients(self, samples: int, seed: int = 0):
        """rz_check_quotients: (pairs compared, mismatches) of the raster pass's shared-
        reciprocal quotient against the IEEE division on random in-range operands."""
        n, bad = C.c_uint64(0), C.c_uint64(0)
        _check(self.lib.rz_check_quotients(self.handle, int(samples), int(seed) & 0xFFFFFFFF, C.byref(n),
                                           C.byref(bad)), self.handle)
        return n.value, bad.value

    def _batch_items(self, batch):
        """batch: list of (Mesh, lights (L,10) float32, Args, frame DeviceArray, depth DeviceArray,
        ids DeviceArray or None)."""
        arr = (BatchItem * len(batch))()
        keep = []
        for i, (mesh, lights, args, frame, depth, ids) in enumerate(batch):
            lights = _lights(lights)
            keep.append(lights)
            arr[i].mesh = mesh.handle.value
            arr[i].lights = lights.ctypes.data
            arr[i].num_lights = lights.shape[0]
            arr[i].args = args
            arr[i].frame = frame.ptr
            arr[i].depth = depth.ptr
            arr[i].ids = ids.ptr if ids is not None else None
        return arr, keep

    def draw_batch(self, batch, flags: int = RZ_FLAG_CLEARED, stats: bool = False):
        """rz_draw_batch over device buffers; returns the summed statistics or None."""
        arr, keep = self._batch_items(batch)
        st = FrameStats() if stats else None
        _check(self.lib.rz_draw_batch(self.handle, C.addressof(arr), len(batch), flags | RZ_FLAG_DEVICE_BUFFERS,
                                      C.byref(st) if stats else None), self.handle)
        return st.as_dict() if stats else None

    def batch_graph(self, batch, frames: int, flags: int = RZ_FLAG_CLEARED):
        arr, keep = self._batch_items(batch)
        h = C.c_void_p()
        _check(self.lib.rz_batch_graph_create(self.handle, C.addressof(arr), len(batch),
                                              flags | RZ_FLAG_DEVICE_BUFFERS, int(frames), C.byref(h)),
               self.handle)
        return FrameGraph(self, h, frames)

    def autotune(self, batch, flags: int = RZ_FLAG_CLEARED):
        """rz_autotune: returns (tile_size, bin_capacity, ms per frame); the context keeps them."""
        arr, keep = self._batch_items(batch)
        t, c, ms = C.c_int32(), C.c_int32(), C.c_float()
        _check(self.lib.rz_autotune(self.handle, C.addressof(arr), len(batch), flags | RZ_FLAG_DEVICE_BUFFERS,
                                    C.byref(t), C.byref(c), C.byref(ms)), self.handle)
        return t.value, c.value, ms.value

    def time_passes(self, mesh: Mesh, lights: np.ndarray, args: Args, frame: "DeviceArray",
                    depth: "DeviceArray", reps: int = 200, flags: int = RZ_FLAG_CLEARED) -> dict:
        """rz_time_passes: microseconds per launch of each pass (graph of `reps` launches);
        "chain" is the shading pass run with neither light loop nor network."""
        lights = _lights(lights)
        us = (C.c_float * 6)()
        _check(self.lib.rz_time_passes(self.handle, mesh.handle, lights.ctypes.data, lights.shape[0],
                                       C.byref(args), C.c_void_p(frame.ptr), C.c_void_p(depth.ptr),
                                       flags | RZ_FLAG_CLEARED | RZ_FLAG_DEVICE_BUFFERS, int(reps), us),
               self.handle)
        return dict(zip(("vertex", "bin", "raster", "overflow", "shade", "chain"), [float(u) for u in us]))

    # ---- learned shader -------------------------------------------------------------
    def mlp_init(self, seed: int = 0):
        _check(self.lib.rz_mlp_init(self.handle, int(seed)), self.handle)

    def mlp_set_weights(self, params: np.ndarray):
        params = np.ascontiguousarray(params, np.float32).reshape(-1)
        if params.size != RZ_MLP_PARAMS:
            raise ValueError(f"expected {RZ_MLP_PARAMS} parameters")
        _check(self.lib.rz_mlp_set_weights(self.handle, params.ctypes.data), self.handle)

    def mlp_get_weights(self) -> np.ndarray:
        out = np.empty(RZ_MLP_PARAMS, np.float32)
        _check(self.lib.rz_mlp_get_weights(self.handle, out.ctypes.data), self.handle)
        return out

    def mlp_collect(self, mesh: Mesh, lights: np.ndarray, args: Args, samples: DeviceArray, count: int = 0) -> int:
        """Appends the plain shader's fragments of one frame to `samples` ((cap, 15) float32 on
        the device) starting at `count`; returns the new fill level (clamped to capacity)."""
        lights = _lights(lights)
        n = C.c_int64(int(count))
        _check(self.lib.rz_mlp_collect(self.handle, mesh.handle, lights.ctypes.data, lights.shape[0],
                                       C.byref(args), C.c_void_p(samples.ptr), samples.shape[0], C.byref(n)),
               self.handle)
        return n.value

    def mlp_train(self, samples: DeviceArray, n: int, steps: int, batch: int = 4096, lr: float = 3e-3,
                  seed: int = 1) -> np.ndarray:
        """Adam steps on the first n device samples; returns the per-step mean squared error."""
        losses = np.empty(int(steps), np.float32)
        _check(self.lib.rz_mlp_train(self.handle, C.c_void_p(samples.ptr), int(n), int(steps), int(batch),
                                     float(lr), int(seed), losses.ctypes.data), self.handle)
        return losses

    def mlp_infer(self, features: np.ndarray) -> np.ndarray:
        """The network on host rows (n, 12) -> (n, 3), through device buffers."""
        x = np.ascontiguousarray(features, np.float32).reshape(-1, 12)
        n = x.shape[0]
        dx, dy = self.empty((max(n, 1), 12), np.float32), self.empty((max(n, 1), 3), np.float32)
        try:
            if n:
                dx.upload(x)
            _check(self.lib.rz_mlp_infer(self.handle, C.c_void_p(dx.ptr), n, C.c_void_p(dy.ptr)), self.handle)
            return dy.download()[:n] if n else np.zeros((0, 3), np.float32)
        finally:
            dx.close()
            dy.close()

    def draw_scene(self, scene: Scene, args: Args, frame=None, depth=None):
        """rz_draw_frame_desc: upload, draw and free in one call (draw_frame's by-value form)."""
        W, H = args.image_width, args.image_height
        flags = 0
        if frame is None and depth is None:
            flags |= RZ_FLAG_CLEARED
            frame = np.zeros((3, H, W), np.uint8)
            depth = np.ones((H, W), np.float32)
        lights = _lights(scene.lights)
        desc, keep = scene.desc()
        _check(self.lib.rz_draw_frame_desc(self.handle, C.byref(desc), lights.ctypes.data,
                                           lights.shape[0], C.byref(args), frame.ctypes.data,
                                           depth.ctypes.data, flags), self.handle)
        del keep
        return frame, depth


def _lights(lights) -> np.ndarray:
    lights = np.asarray(lights)
    if lights.dtype != np.float32 or not lights.flags.c_contiguous or lights.ndim != 2 \
            or lights.shape[1] != 10:
        raise ValueError("lights must be a C-contiguous (L, 10) float32 array")
    return lights
