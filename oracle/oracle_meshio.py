"""Restatement of libigl's triangle-mesh file readers, for tests.

TEST INFRASTRUCTURE ONLY (the checker of libigl_amd's readers).

  read_off       readOFF.cpp:22-170
  read_obj       readOBJ.cpp:68-276 (vertices truncated to xyz as
                 read_triangle_mesh.cpp:146-151 does)
  fan            polygons_to_triangles.cpp:16-31
  read_triangle_mesh  read_triangle_mesh.cpp:117-190 (.off / .obj)

Pinned by files the tests write from known arrays (closed form), and by the
reference's own test expectations (tests/include/igl/readOFF.cpp:4-14,
readOBJ.cpp:7-16: the cube reads as 8 vertices and 12 triangles).
"""
from __future__ import annotations

import os
import re

import numpy as np


def read_off(path):
    with open(path, "r") as f:
        lines = f.read().split("\n")
    k = 0
    header = lines[k].split()[0] if lines[k].split() else ""
    if not (header.startswith("OFF") or header.startswith("NOFF") or header.startswith("COFF")):
        raise ValueError("bad OFF header")
    k += 1
    while lines[k].startswith("#") or lines[k] == "":
        k += 1
    nv, nf = (int(x) for x in lines[k].split()[:2])
    k += 1
    V = []
    while len(V) < nv:
        toks = lines[k].split()
        k += 1
        try:
            vals = [float(t) for t in toks[:6]]
        except ValueError:
            vals = []
        if len(toks) >= 3 and len(vals) >= 3:
            V.append(vals[:3])
    polys = []
    rest = "\n".join(lines[k:])
    words = iter(re.findall(r"#[^\n]*|[^\s]+|\n", rest))
    pending_line = False
    for w in words:
        if len(polys) == nf:
            break
        if w == "\n":
            continue
        if w.startswith("#"):
            continue
        n = int(w)
        face = [int(next(x for x in words if x != "\n")) for _ in range(n)]
        # anything after the last index on its line is ignored
        for x in words:
            if x == "\n":
                break
        polys.append(face)
    return np.array(V, float).reshape(-1, 3), polys


def read_obj(path):
    V, polys = [], []
    with open(path, "r") as f:
        for line in f:
            toks = line.split()
            if not toks:
                continue
            if toks[0] == "v":
                c = []
                for t in toks[1:]:
                    try:
                        c.append(float(t))
                    except ValueError:
                        break
                if len(c) < 3:
                    raise ValueError("vertex with fewer than 3 coordinates")
                V.append(c[:3])
            elif toks[0] == "f":
                n = len(V)
                face, nt, nn = [], 0, 0
                for w in toks[1:]:
                    parts = w.split("/")
                    i = int(parts[0])
                    if len(parts) >= 2 and parts[1] != "":
                        nt += 1
                    if len(parts) == 3 and parts[2] != "":
                        nn += 1
                    face.append(i + n if i < 0 else i - 1)
                if not face or nt not in (0, len(face)) or nn not in (0, len(face)):
                    raise ValueError("invalid face format")
                polys.append(face)
    return np.array(V, float).reshape(-1, 3), polys


def fan(polys):
    F = [(p[0], p[c], p[c + 1]) for p in polys for c in range(1, len(p) - 1)]
    return np.array(F, np.int64).reshape(-1, 3)


def read_triangle_mesh(path):
    ext = os.path.splitext(path)[1].lower()
    V, polys = read_off(path) if ext == ".off" else read_obj(path)
    return V, fan(polys)
