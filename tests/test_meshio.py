"""Mesh files (.off / .obj) as triangle meshes: the library's reader
(gwn_read_triangle_mesh) against the oracle restatement of readOFF / readOBJ /
polygons_to_triangles, and both against the arrays the files were written from."""
import os

import numpy as np
import pytest

import oracle_meshio as OM
from libigl_amd import datagen as D
from libigl_amd import gwn

CUBE_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
CUBE_Q = [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]]


def _both(path):
    V, F = gwn.read_triangle_mesh(path)
    Vo, Fo = OM.read_triangle_mesh(path)
    assert np.array_equal(V, Vo) and np.array_equal(F, Fo), path
    return V, F


def test_cube_off_and_obj(tmp_path):
    # the reference's readOFF / readOBJ tests: the cube is 8 vertices, 12 triangles
    off = tmp_path / "cube.off"
    off.write_text("OFF\n# a comment\n\n8 6 12\n" + "".join(f"{x:g} {y:g} {z:g}\n" for x, y, z in CUBE_V) +
                   "".join("4 " + " ".join(map(str, q)) + "\n" for q in CUBE_Q))
    obj = tmp_path / "cube.OBJ"
    obj.write_text("# cube\nmtllib cube.mtl\no cube\n" + "".join(f"v {x:g} {y:g} {z:g}\n" for x, y, z in CUBE_V) +
                   "vt 0 0\nvt 1 0\nvn 0 0 1\ng side\ns off\nusemtl red\n" +
                   "".join("f " + " ".join(str(i + 1) for i in q) + "\n" for q in CUBE_Q[:3]) +
                   "usemtl blue\n" + "".join("f " + " ".join(str(i + 1) for i in q) + "\n" for q in CUBE_Q[3:]))
    for p in (off, obj):
        V, F = _both(str(p))
        assert V.shape == (8, 3) and F.shape == (12, 3)
        assert np.array_equal(V, CUBE_V)
        assert np.array_equal(F, OM.fan(CUBE_Q))


def test_obj_face_word_forms_and_negative_indices(tmp_path):
    p = tmp_path / "forms.obj"
    p.write_text("v 0 0 0 1\nv 1 0 0\nv 1 1 0\nv 0 1 0 0.5\nvt 0 0\nvt 1 1\nvn 0 0 1\n"
                 "f 1/1/1 2/2/1 3/1/1 4/2/1\n"   # i/t/n, quad
                 "f 1//1 3//1 4//1\n"            # i//n
                 "f 1/2 2/1 3/2\n"               # i/t
                 "f -4 -3 -2 -1\n"               # negative: relative to the 4 vertices so far
                 "\n# trailing comment\n")
    V, F = _both(str(p))
    assert V.shape == (4, 3) and np.array_equal(V[0], [0, 0, 0])  # 4th coordinate dropped
    assert np.array_equal(F, [[0, 1, 2], [0, 2, 3], [0, 2, 3], [0, 1, 2], [0, 1, 2], [0, 2, 3]])


def test_off_variants(tmp_path):
    p = tmp_path / "c.off"
    p.write_text("COFF\n3 1 0\n0 0 0 255 0 0 255\n1 0 0 0 255 0 255\n0 1 0 0 0 255 255\n3 0 1 2 128 128 128\n")
    V, F = _both(str(p))
    assert np.array_equal(F, [[0, 1, 2]]) and np.array_equal(V[1], [1, 0, 0])
    p = tmp_path / "n.off"
    p.write_text("NOFF\n4 2 0\n0 0 0 0 0 1\n1 0 0 0 0 1\n1 1 0 0 0 1\n0 1 0 0 0 1\n"
                 "# faces\n3 0 1 2\n3 0 2 3\n")
    V, F = _both(str(p))
    assert np.array_equal(F, [[0, 1, 2], [0, 2, 3]])


@pytest.mark.parametrize("ext", ["off", "obj"])
def test_round_trip_of_a_generated_mesh(tmp_path, ext):
    V, F = D.bumpy_sphere(3, 4, 7)
    p = tmp_path / f"m.{ext}"
    with open(p, "w") as f:
        if ext == "off":
            f.write(f"OFF\n{V.shape[0]} {F.shape[0]} 0\n")
            for v in V:
                f.write("%.17g %.17g %.17g\n" % tuple(v))
            for t in F:
                f.write("3 %d %d %d\n" % tuple(t))
        else:
            for v in V:
                f.write("v %.17g %.17g %.17g\n" % tuple(v))
            for t in F:
                f.write("f %d %d %d\n" % tuple(t + 1))
    V2, F2 = _both(str(p))
    assert np.array_equal(V2, V) and np.array_equal(F2, F)


def test_reader_errors(tmp_path):
    bad = {"x.stl": "solid x\n", "h.off": "PLY\n1 0 0\n0 0 0\n", "v.obj": "v 0 0\n",
           "f.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/1 2 3\n"}
    for name, text in bad.items():
        p = tmp_path / name
        p.write_text(text)
        with pytest.raises(gwn.GWNError):
            gwn.read_triangle_mesh(str(p))
    with pytest.raises(gwn.GWNError):
        gwn.read_triangle_mesh(str(tmp_path / "missing.obj"))
