"""Write tests/golden/obj_cases.npz from the reference's own vendored tinyobjloader.

For each OBJ case in tests/golden/obj_cases.py, runs oracle/_ref/tinyobj_ref (the
reference's tiny_obj_loader.h flattened like its load_obj) and stores the vertex,
normal, uv and face arrays it produces.  Run from the repository root:
    python tests/golden/make_obj_vectors.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from obj_cases import CASES, write_case  # noqa: E402


def run_reference(obj, mtl_dir, prefix):
    exe = os.path.join(ROOT, "oracle", "_ref", "tinyobj_ref")
    subprocess.run([exe, obj, mtl_dir, prefix], check=True, capture_output=True)
    out = {}
    for key, dt in (("vertices", np.float32), ("normals", np.float32), ("uvs", np.float32), ("faces", np.int32)):
        ext = "i32" if dt is np.int32 else "f32"
        out[key] = np.fromfile(f"{prefix}.{key}.{ext}", dt)
    out["materials"] = np.array(open(prefix + ".materials.txt").read())
    return out


def main():
    import oracle
    oracle.build()
    res = {}
    with tempfile.TemporaryDirectory() as d:
        for name in CASES:
            obj, mtl_dir = write_case(name, d)
            got = run_reference(obj, mtl_dir, os.path.join(d, name))
            for k, v in got.items():
                res[f"{name}.{k}"] = v
    np.savez_compressed(os.path.join(HERE, "obj_cases.npz"), **res)
    print("wrote", len(res), "arrays")


if __name__ == "__main__":
    main()
