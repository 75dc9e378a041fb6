"""Host-side scene container in the reference's own mesh representation.

A ``Scene`` holds what ``load_obj`` + ``load_lights`` hand to ``draw_frame``
(fileloader.cpp:79-133): positions, normals, uvs, one 40-byte ``Triangle`` per face
(separate v/vn/vt index streams and a material index), the materials, and the
lights.  ``Scene.desc()`` views it as an ``rz_mesh_desc`` without copying.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .abi import Material, MeshDesc


@dataclass
class MaterialSpec:
    """One reference ``Material`` (headers/material.h:11-25)."""
    diffuse: tuple = (1.0, 1.0, 1.0)
    texture: np.ndarray | None = None     # (H, W, C) uint8 as the image file holds it, C in {3, 4}
    normal_map: np.ndarray | None = None  # (H, W, C) uint8 tangent-space normals, C in {3, 4}


@dataclass
class Scene:
    vertices: np.ndarray                  # (V, 3) float32 model space, or (V, 4) raster space
    normals: np.ndarray                   # (N, 3) float32
    uvs: np.ndarray                       # (T, 2) float32
    faces: np.ndarray                     # (F, 10) int32: v0 v1 v2 n0 n1 n2 t0 t1 t2 material
    materials: list = field(default_factory=lambda: [MaterialSpec()])
    lights: np.ndarray = field(default_factory=lambda: np.zeros((0, 10), np.float32))
    raster_space: bool = False

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, np.float32)
        self.normals = np.ascontiguousarray(self.normals, np.float32).reshape(-1, 3)
        self.uvs = np.ascontiguousarray(self.uvs, np.float32).reshape(-1, 2)
        self.faces = np.ascontiguousarray(self.faces, np.int32).reshape(-1, 10)
        self.lights = np.ascontiguousarray(self.lights, np.float32).reshape(-1, 10)
        want = 4 if self.raster_space else 3
        if self.vertices.ndim != 2 or self.vertices.shape[1] != want:
            raise ValueError(f"vertices must be (V, {want})")

    @property
    def num_faces(self) -> int:
        return int(self.faces.shape[0])

    def lights_copy(self) -> np.ndarray:
        return np.array(self.lights, dtype=np.float32, copy=True)

    def desc(self):
        """(MeshDesc, keepalive) — the descriptor points into this scene's arrays."""
        mats = (Material * max(1, len(self.materials)))()
        keep = [self, mats]
        for i, m in enumerate(self.materials):
            mats[i].diffuse_colour[:] = [float(c) for c in m.diffuse]
            if m.texture is not None:
                tex = np.ascontiguousarray(m.texture, np.uint8)
                if tex.ndim != 3 or tex.shape[2] not in (3, 4):
                    raise ValueError("texture must be (H, W, 3|4) uint8")
                keep.append(tex)
                mats[i].tex_height, mats[i].tex_width, mats[i].tex_channels = tex.shape
                mats[i].texels = tex.ctypes.data
            if m.normal_map is not None:
                nm = np.ascontiguousarray(m.normal_map, np.uint8)
                if nm.ndim != 3 or nm.shape[2] not in (3, 4):
                    raise ValueError("normal map must be (H, W, 3|4) uint8")
                keep.append(nm)
                mats[i].normal_map_height, mats[i].normal_map_width, mats[i].normal_map_channels = nm.shape
                mats[i].normal_map = nm.ctypes.data
        d = MeshDesc()
        d.vertices = self.vertices.ctypes.data
        d.normals = self.normals.ctypes.data
        d.uvs = self.uvs.ctypes.data
        d.faces = self.faces.ctypes.data
        d.materials = C.addressof(mats)
        d.num_vertices = self.vertices.shape[0]
        d.num_normals = self.normals.shape[0]
        d.num_uvs = self.uvs.shape[0]
        d.num_faces = self.faces.shape[0]
        d.num_materials = len(self.materials)
        d.raster_space = int(self.raster_space)
        return d, keep


def make_faces(v, n=None, t=None, material=None) -> np.ndarray:
    """Pack index streams into the (F, 10) Triangle layout; t=None marks no uvs (-1)."""
    v = np.asarray(v, np.int32).reshape(-1, 3)
    f = np.empty((v.shape[0], 10), np.int32)
    f[:, 0:3] = v
    f[:, 3:6] = v if n is None else np.asarray(n, np.int32).reshape(-1, 3)
    f[:, 6:9] = -1 if t is None else np.asarray(t, np.int32).reshape(-1, 3)
    f[:, 9] = 0 if material is None else np.asarray(material, np.int32)
    return f


def make_lights(directions, intensities, colours) -> np.ndarray:
    """(L, 10) float32 in the reference Light layout; trans_dir starts at zero."""
    d = np.asarray(directions, np.float32).reshape(-1, 3)
    out = np.zeros((d.shape[0], 10), np.float32)
    out[:, 0:3] = d
    out[:, 3] = np.asarray(intensities, np.float32).reshape(-1)
    out[:, 4:7] = np.asarray(colours, np.float32).reshape(-1, 3)
    return out
