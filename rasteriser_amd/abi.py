"""ctypes mirror of include/rz.h — the record types of the C ABI.

The byte layouts match the reference's own structs: ``Triangle`` is
``headers/face.h:6-13`` (3+3+3+1 int32), ``Light`` is ``headers/light.h:7-14``
(direction, intensity, colour, trans_dir).  Numpy arrays with 10 int32 per face and
10 float32 per light therefore pass straight through as ``rz_triangle*`` and
``rz_light*``.
"""
import ctypes as C

import numpy as np

RZ_OK = 0
RZ_ERR_INVALID_ARG = 1
RZ_ERR_INDEX_OOB = 2
RZ_ERR_HIP = 3
RZ_ERR_OOM = 4
RZ_ERR_NO_DEVICE = 5

RZ_FLAG_CLEARED = 1
RZ_FLAG_DEVICE_BUFFERS = 2
RZ_FLAG_ASYNC = 4
RZ_FLAG_MLP_SHADE = 8

RZ_MLP_PARAMS = 5187
RZ_MLP_SAMPLE_FLOATS = 15


class Triangle(C.Structure):
    _fields_ = [("vertices", C.c_int32 * 3), ("normals", C.c_int32 * 3),
                ("uvs", C.c_int32 * 3), ("material", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("direction", C.c_float * 3), ("intensity", C.c_float),
                ("colour", C.c_float * 3), ("trans_dir", C.c_float * 3)]


class Material(C.Structure):
    _fields_ = [("diffuse_colour", C.c_float * 3), ("tex_width", C.c_int32),
                ("tex_height", C.c_int32), ("tex_channels", C.c_int32),
                ("texels", C.c_void_p), ("normal_map_width", C.c_int32),
                ("normal_map_height", C.c_int32), ("normal_map_channels", C.c_int32),
                ("reserved", C.c_int32), ("normal_map", C.c_void_p)]


class Args(C.Structure):
    _fields_ = [("image_width", C.c_uint32), ("image_height", C.c_uint32),
                ("aspect_ratio", C.c_float), ("flat", C.c_int32),
                ("wind_clockwise", C.c_int32), ("scale", C.c_float),
                ("displacement", C.c_float * 3), ("tait_bryan_angles", C.c_float * 3),
                ("fovy_degrees", C.c_float), ("z_near", C.c_float), ("z_far", C.c_float)]


class MeshDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("normals", C.c_void_p), ("uvs", C.c_void_p),
                ("faces", C.c_void_p), ("materials", C.c_void_p),
                ("num_vertices", C.c_int32), ("num_normals", C.c_int32),
                ("num_uvs", C.c_int32), ("num_faces", C.c_int32),
                ("num_materials", C.c_int32), ("raster_space", C.c_int32)]


class BatchItem(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("lights", C.c_void_p), ("num_lights", C.c_int32),
                ("reserved", C.c_int32), ("args", Args), ("frame", C.c_void_p), ("depth", C.c_void_p),
                ("ids", C.c_void_p)]


class Matrices(C.Structure):
    _fields_ = [("model", C.c_float * 16), ("view", C.c_float * 16),
                ("modelview", C.c_float * 16), ("camera", C.c_float * 16),
                ("normal", C.c_float * 16)]


class FrameStats(C.Structure):
    _fields_ = [("triangles", C.c_uint64), ("front_triangles", C.c_uint64),
                ("bin_pairs", C.c_uint64), ("overflow_tiles", C.c_uint64),
                ("fragments", C.c_uint64), ("covered_pixels", C.c_uint64),
                ("ms_vertex", C.c_float), ("ms_bin", C.c_float), ("ms_raster", C.c_float),
                ("ms_overflow", C.c_float), ("ms_shade", C.c_float), ("ms_frame", C.c_float),
                ("raster_tiles", C.c_uint64), ("raster_pixels", C.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def make_args(width, height, *, aspect_ratio=None, flat=False, wind_clockwise=False, scale=1.0,
              displacement=(0.0, 0.0, 0.0), angles=(0.0, 0.0, 0.0), fovy=45.0, z_near=0.1,
              z_far=6.0):
    """Args with the reference defaults (arguments.cpp:20-31, geometry.cpp:30)."""
    a = Args()
    a.image_width = width
    a.image_height = height
    # (float)width / (float)height in binary32, as arguments.cpp:39 computes it
    a.aspect_ratio = (float(np.float32(width) / np.float32(height))
                      if aspect_ratio is None else aspect_ratio)
    a.flat = int(bool(flat))
    a.wind_clockwise = int(bool(wind_clockwise))
    a.scale = scale
    a.displacement[:] = list(displacement)
    a.tait_bryan_angles[:] = list(angles)
    a.fovy_degrees = fovy
    a.z_near = z_near
    a.z_far = z_far
    return a


def declare(lib):
    """Attach argtypes/restype for every entry point of include/rz.h."""
    P = C.c_void_p
    sig = {
        "rz_abi_version": (C.c_int32, []),
        "rz_status_string": (C.c_char_p, [C.c_int]),
        "rz_last_error": (C.c_char_p, [P]),
        "rz_create": (C.c_int, [C.c_int32, C.POINTER(P)]),
        "rz_destroy": (None, [P]),
        "rz_set_tuning": (C.c_int, [P, C.c_int32, C.c_int32]),
        "rz_get_tuning": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "rz_set_fused_shade": (C.c_int, [P, C.c_int32]),
        "rz_get_fused_shade": (C.c_int32, [P]),
        "rz_set_profiling": (C.c_int, [P, C.c_int32]),
        "rz_stream": (P, [P]),
        "rz_sync": (C.c_int, [P]),
        "rz_mesh_create": (C.c_int, [P, C.POINTER(MeshDesc), C.POINTER(P)]),
        "rz_mesh_destroy": (None, [P, P]),
        "rz_draw_frame": (C.c_int, [P, P, P, C.c_int32, C.POINTER(Args), P, P, P, C.c_uint32,
                                    C.POINTER(FrameStats)]),
        "rz_draw_frame_desc": (C.c_int, [P, C.POINTER(MeshDesc), P, C.c_int32, C.POINTER(Args),
                                         P, P, C.c_uint32]),
        "rz_draw_frame_noalloc": (C.c_int, [P, P, P, C.c_int32, C.POINTER(Args), P, P, P, C.c_uint32]),
        "rz_graph_create": (C.c_int, [P, P, P, C.c_int32, C.POINTER(Args), P, P, P, C.c_uint32,
                                      C.c_int32, C.c_float, C.POINTER(P)]),
        "rz_graph_launch": (C.c_int, [P]),
        "rz_graph_destroy": (None, [P]),
        "rz_spin_create": (C.c_int, [P, P, P, C.c_int32, C.POINTER(Args), P, P, P, C.c_uint32, C.POINTER(P)]),
        "rz_spin_frame": (C.c_int, [P, C.POINTER(C.c_float)]),
        "rz_spin_destroy": (None, [P]),
        "rz_draw_text": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_int32, C.c_int32, C.c_char_p,
                                   C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]),
        "rz_draw_batch": (C.c_int, [P, P, C.c_int32, C.c_uint32, C.POINTER(FrameStats)]),
        "rz_set_batch_persistent": (C.c_int, [P, C.c_int32]),
        "rz_get_batch_persistent": (C.c_int32, [P]),
        "rz_batch_graph_create": (C.c_int, [P, P, C.c_int32, C.c_uint32, C.c_int32, C.POINTER(P)]),
        "rz_autotune": (C.c_int, [P, P, C.c_int32, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_float)]),
        "rz_time_passes": (C.c_int, [P, P, P, C.c_int32, C.POINTER(Args), P, P, C.c_uint32, C.c_int32,
                                     C.POINTER(C.c_float)]),
        "rz_mlp_init": (C.c_int, [P, C.c_uint64]),
        "rz_mlp_set_weights": (C.c_int, [P, P]),
        "rz_mlp_get_weights": (C.c_int, [P, P]),
        "rz_mlp_collect": (C.c_int, [P, P, P, C.c_int32, C.POINTER(Args), P, C.c_int64,
                                     C.POINTER(C.c_int64)]),
        "rz_mlp_train": (C.c_int, [P, P, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, P]),
        "rz_mlp_infer": (C.c_int, [P, P, C.c_int64, P]),
        "rz_device_alloc": (C.c_int, [P, C.c_size_t, C.POINTER(P)]),
        "rz_device_free": (C.c_int, [P, P]),
        "rz_copy_to_device": (C.c_int, [P, P, P, C.c_size_t]),
        "rz_copy_to_host": (C.c_int, [P, P, P, C.c_size_t]),
        "rz_copy_device": (C.c_int, [P, P, P, C.c_size_t]),
        "rz_check_quotients": (C.c_int, [P, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "rz_default_args": (None, [C.POINTER(Args), C.c_uint32, C.c_uint32]),
        "rz_frame_matrices": (C.c_int, [C.POINTER(Args), C.POINTER(Matrices)]),
    }
    for name, (res, argt) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = argt
    return lib


EXPORTED = ("rz_abi_version", "rz_status_string", "rz_last_error", "rz_create", "rz_destroy",
            "rz_set_tuning", "rz_get_tuning", "rz_set_fused_shade", "rz_get_fused_shade", "rz_set_profiling", "rz_stream", "rz_sync",
            "rz_mesh_create", "rz_mesh_destroy", "rz_draw_frame", "rz_draw_frame_desc", "rz_draw_frame_noalloc",
            "rz_graph_create", "rz_graph_launch", "rz_graph_destroy", "rz_spin_create", "rz_spin_frame",
            "rz_spin_destroy", "rz_draw_text",
            "rz_draw_batch", "rz_set_batch_persistent", "rz_get_batch_persistent", "rz_batch_graph_create", "rz_autotune",
            "rz_time_passes", "rz_mlp_init", "rz_mlp_set_weights", "rz_mlp_get_weights", "rz_mlp_collect",
            "rz_mlp_train", "rz_mlp_infer",
            "rz_device_alloc", "rz_device_free", "rz_copy_to_device", "rz_copy_to_host", "rz_copy_device",
            "rz_check_quotients", "rz_default_args", "rz_frame_matrices")
