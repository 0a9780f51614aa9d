"""Writes the golden fixtures of tests/golden/ from the reference itself.

Runs in the build container only (it needs /root/reference and the dumper
oracle/_ref/ref_dump built by `make -C oracle ref`).  For each case it writes
a mesh file (generated meshes through the product's .msh writer, or a mesh
file the reference's own tests use, copied under tests/golden/meshes/), runs
the reference's mesh reader, DOF numbering, CSR builder and element kernels on
it through the dumper, and stores what they produce as a compressed .npz:
vertex coordinates and connectivity as read, signed edge tags, element ->
DOF maps, the initial solution with the essential values, CSR ia/ja, the
assembled values and right-hand side.
"""
from __future__ import annotations

import struct
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
DUMPER = ROOT / "oracle" / "_ref" / "ref_dump"

# name, mesh spec, problem, order, quadrature degree, extra parameter
CASES = [
    ("sq8_p1_sinsin", ("square", 8, 0.0, 0), "diff2d_sinsin", 1, 2, None),
    ("sq6j_p1_blocks_q5", ("square", 6, 0.2, 7), "diff2d_blocks", 1, 5, "4"),
    ("sq8_p2_sinsin", ("square", 8, 0.0, 0), "diff2d_sinsin", 2, 2, None),
    ("sq6j_p2_blocks", ("square", 6, 0.2, 3), "diff2d_blocks", 2, 2, "4"),
    ("sq4_p2_sinsin_q8", ("square", 4, 0.1, 5), "diff2d_sinsin", 2, 8, None),
    ("cube3_p1_smooth_q1", ("cube", 3), "diff3d_smooth", 1, 1, None),
    ("cube3_p1_smooth_q2", ("cube", 3), "diff3d_smooth", 1, 2, None),
    ("mms1_p1_sinsin", ("file", "square1.msh"), "diff2d_sinsin", 1, 2, None),
    ("mms2_p2_sinsin", ("file", "square2.msh"), "diff2d_sinsin", 2, 2, None),
    ("mms3_p2_blocks", ("file", "square3.msh"), "diff2d_blocks", 2, 2, "4"),
    ("sq4_th_kovasznay", ("square", 4, 0.0, 0), "ns_kovasznay", 2, 5, "40"),
    ("sq5j_th_kovasznay", ("square", 5, 0.2, 1), "ns_kovasznay", 2, 5, "40"),
    ("data2_th_kovasznay_q8", ("file", "data/square2.msh"), "ns_kovasznay", 2, 8, "1"),
]


def readbin(path):
    d = {}
    b = Path(path).read_bytes()
    o = 0
    while o < len(b):
        (ln,) = struct.unpack_from("<i", b, o)
        o += 4
        name = b[o:o + ln].decode()
        o += ln
        t = chr(b[o])
        o += 1
        (c,) = struct.unpack_from("<q", b, o)
        o += 8
        dt = np.float64 if t == "d" else np.int32
        sz = 8 if t == "d" else 4
        d[name] = np.frombuffer(b, dt, c, o).copy()
        o += c * sz
    return d


def main(only=None):
    from feng_amd.api import Mesh

    if not DUMPER.exists():
        raise SystemExit(f"{DUMPER} missing: run `make -C oracle ref` first")
    with tempfile.TemporaryDirectory() as tmp:
        for name, spec, problem, order, deg, extra in CASES:
            if only and name not in only:
                continue
            if spec[0] == "file":
                msh = HERE / "meshes" / spec[1]
            else:
                msh = Path(tmp) / f"{name}.msh"
                m = Mesh.square(spec[1], spec[2], spec[3]) if spec[0] == "square" else Mesh.cube(spec[1])
                m.write_msh(msh)
            out = Path(tmp) / f"{name}.bin"
            cmd = [str(DUMPER), str(msh), problem, str(order), str(deg), str(out)]
            if extra:
                cmd.append(extra)
            subprocess.run(cmd, check=True)
            d = readbin(out)
            meta = np.array([order, deg], np.int32)
            np.savez_compressed(HERE / f"{name}.npz", meta=meta, spec=np.array([str(spec)]),
                                problem=np.array([problem]), extra=np.array([extra or ""]), **d)
            print(name, "n", int(d["n"][0]), "nnz", d["ja"].size)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
