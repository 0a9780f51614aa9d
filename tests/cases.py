"""Shared problem set-up for the parity tests (fixture -> oracle / product inputs)."""
from __future__ import annotations

import ast
import math

import numpy as np

from tests.conftest import golden

DIFF_CASES = ["sq8_p1_sinsin", "sq6j_p1_blocks_q5", "sq8_p2_sinsin", "sq6j_p2_blocks", "sq4_p2_sinsin_q8",
              "cube3_p1_smooth_q1", "cube3_p1_smooth_q2", "mms1_p1_sinsin", "mms2_p2_sinsin", "mms3_p2_blocks"]
NS_CASES = ["sq4_th_kovasznay", "sq5j_th_kovasznay", "data2_th_kovasznay_q8"]

FN_CONST, FN_SINSIN, FN_BLOCKS, FN_SMOOTH3 = 0, 1, 2, 3


def load(name):
    g = golden(name)
    g["order"], g["deg"] = int(g["meta"][0]), int(g["meta"][1])
    g["problem"] = str(g["problem"][0])
    g["extra"] = str(g["extra"][0])
    g["spec"] = ast.literal_eval(str(g["spec"][0]))
    g["n"] = int(g["n"][0])
    g["ndofs"] = int(g["ndofs"][0])
    g["nvpe"] = int(g["nvpe"][0])
    return g


def kind_of(g):
    if g["problem"].startswith("diff3d"):
        return 4
    if g["problem"].startswith("ns"):
        return 3
    return 1 if g["order"] == 1 else 2


def coeffs_of(g):
    """(diffusivity, source) as (kind, c0, c1, data) tuples."""
    p = g["problem"]
    if p == "diff2d_sinsin":
        return (FN_CONST, 1.0, 0.0, None), (FN_SINSIN, -2.0 * math.pi ** 2, 0.0, None)
    if p == "diff2d_blocks":
        return (FN_BLOCKS, float(g["extra"]), 0.0, None), (FN_CONST, -1.0, 0.0, None)
    if p == "diff3d_smooth":
        return (FN_SMOOTH3, 1.5, 0.0, None), (FN_CONST, -1.0, 0.0, None)
    raise ValueError(p)


def bedge_from(conn, edges, bconn):
    E = conn.size // 3
    key = {}
    for e in range(E):
        for k in range(3):
            a, b = int(conn[3 * e + k]), int(conn[3 * e + (k + 1) % 3])
            key[(min(a, b), max(a, b))] = abs(int(edges[3 * e + k])) - 1
    out = np.zeros(bconn.size // 2, np.int32)
    for i in range(out.size):
        a, b = int(bconn[2 * i]), int(bconn[2 * i + 1])
        out[i] = key[(min(a, b), max(a, b))]
    return out


def relerr(a, b):
    """max |a-b| / max |b| (norm-based: near-zero entries are judged against
    the largest entry)."""
    sc = np.abs(b).max() if b.size else 1.0
    return float(np.abs(a - b).max() / (sc if sc > 0 else 1.0)) if b.size else 0.0


def product_mesh(g):
    from feng_amd.api import Mesh
    from tests.conftest import GOLDEN
    spec = g["spec"]
    if spec[0] == "square":
        return Mesh.square(spec[1], spec[2], spec[3])
    if spec[0] == "cube":
        return Mesh.cube(spec[1])
    return Mesh.read_msh(GOLDEN / "meshes" / spec[1])
