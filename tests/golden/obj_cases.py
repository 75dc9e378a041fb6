"""OBJ/MTL inputs for the loader checks (data only).

Each case exercises a behaviour of the reference's loader (tinyobjloader 1.0.x): its
decimal parser on unusual spellings, polygon fans, index forms and negative indices,
material switches (including the shape dropped when usemtl directly precedes g/o),
mtllib fallbacks, map_Kd options and CRLF line endings.
"""
import os

NUMBERS = """# numbers in the spellings the loader sees
v 0.1 -0.25 3
v 1e-3 -2.5E+2 +7.125
v 0.123456789012 1234567.891 -0.000000123
v 5. 6.0e1 -7.5e-1
v .5 abc 1e
v 3.14159265358979 2.718281828459045 1.4142135623730951
v 100000000000 0.000001 0.9999999999
vt 0.3333333333 0.6666666667
vt 1.0e0 -0.0
vn 0 0 1
vn 0.577350269 0.577350269 0.577350269
f 1/1/1 2/2/2 3/1/1
f 4//2 5//1 6//2
f 7 1 2
"""

POLYGONS = """mtllib missing.mtl cases.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v -1 0.5 0
v 0.5 -1 0
vt 0 0
vt 1 0
vt 1 1
vn 0 0 1
o first
usemtl red
f 1/1/1 2/2/1 3/3/1 4/1/1 5/2/1
f -6/-3/-1 -5/-2/-1 -1/-1/-1
usemtl blue
f 2//1 3//1 4//1
usemtl nosuch
f 1 2 3 4
g second
f 6/1 1/2 2/3
"""

DROPPED = """mtllib cases.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v -1 0.5 0
v 0.5 -1 0
o first
usemtl red
f 1 2 3 4 5
usemtl blue
f 2 3 4
usemtl red
g second
f 6 1 2
usemtl blue
usemtl red
f 3 4 5
o third
f 4 5 6
"""

CRLF = "mtllib cases.mtl\r\nv 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\nvn 0 0 1\r\nusemtl blue\r\nf 1//1 2//1 3//1\r\n"

MTL = """# materials
newmtl red
Kd 0.8 0.1 0.1
map_Kd -blendu off -o 0.5 0.5 -s 2 2 1 red.png
newmtl blue
Kd 0.1 0.2 0.9e0
newmtl red
Kd 1 1 1
"""

CASES = {"numbers": (NUMBERS, None), "polygons": (POLYGONS, MTL), "dropped": (DROPPED, MTL),
         "crlf": (CRLF, MTL)}


def write_case(name, directory):
    """Write case `name` into `directory`; returns (obj path, materials dir with a slash)."""
    obj, mtl = CASES[name]
    path = os.path.join(directory, name + ".obj")
    with open(path, "w", newline="") as f:
        f.write(obj)
    if mtl is not None:
        with open(os.path.join(directory, "cases.mtl"), "w") as f:
            f.write(mtl)
    return path, directory + "/"
