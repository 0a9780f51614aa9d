"""Product host library (libfe_host.so) against the reference fixtures:
mesh generation / .msh round trip, edge tags, DOF numbering, coloring."""
import re
from pathlib import Path

import numpy as np
import pytest

from tests.cases import DIFF_CASES, NS_CASES, kind_of, load, product_mesh

INCLUDE = Path(__file__).resolve().parent.parent / "include"


@pytest.mark.parametrize("name", DIFF_CASES + NS_CASES)
def test_mesh_and_numbering_match_reference(name):
    from feng_amd.api import Numbering
    g = load(name)
    m = product_mesh(g)
    assert np.array_equal(m.xyz, g["xyz"])
    assert np.array_equal(m.conn, g["conn"])
    if m.dim == 2:
        assert np.array_equal(m.edges, g["edges"])
    kind = kind_of(g)
    kinds, ess = ([3, 1], [1, 1]) if kind == 3 else ([kind], [1])
    num = Numbering(m, kinds, ess)
    assert (num.nInc, num.nDofs) == (g["n"], g["ndofs"])
    for f in range(len(kinds)):
        assert np.array_equal(num.adr[f], g[f"adr{f}"])


def test_msh_round_trip(tmp_path):
    from feng_amd.api import Mesh
    for m in (Mesh.square(7, 0.2, 11), Mesh.cube(3)):
        p = tmp_path / "m.msh"
        m.write_msh(p)
        r = Mesh.read_msh(p)
        assert np.array_equal(r.xyz, m.xyz) and np.array_equal(r.conn, m.conn)
        assert np.array_equal(r.bconn, m.bconn)
        if m.dim == 2:
            assert np.array_equal(r.edges, m.edges) and np.array_equal(r.bedge, m.bedge)


@pytest.mark.parametrize("maker", [lambda M: M.square(16, 0.2, 1), lambda M: M.cube(4)])
def test_coloring_valid_and_orientation_positive(maker):
    from feng_amd.api import Mesh
    m = maker(Mesh)
    col, nc = m.coloring()
    conn = m.conn.reshape(-1, m.nvpe)
    for c in range(nc):
        v = conn[col == c].ravel()
        assert np.unique(v).size == v.size
    x = m.xyz.reshape(-1, 3)
    if m.dim == 3:
        a, b, c_, d = (x[conn[:, i]] for i in range(4))
        det = np.einsum("ij,ij->i", b - a, np.cross(c_ - a, d - a))
        assert (det > 0).all()
        assert np.isclose(det.sum() / 6.0, 1.0)
    else:
        a, b, c_ = (x[conn[:, i]] for i in range(3))
        det = (b[:, 0] - a[:, 0]) * (c_[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c_[:, 0] - a[:, 0])
        assert (det > 0).all() and np.isclose(det.sum() / 2.0, 1.0)


def test_config_sizes():
    """System sizes quoted for the configs (64^2 P1: n 3969, nnz 27,281)."""
    from feng_amd.api import Mesh, Numbering
    m = Mesh.square(64)
    num = Numbering(m, [1], [1])
    assert (m.E, m.V, num.nInc) == (8192, 4225, 3969)
    c = Mesh.cube(8)
    nc = Numbering(c, [4], [1])
    assert nc.nInc == 7 ** 3 and c.E == 6 * 8 ** 3


def test_coefficient_fields_deterministic():
    from feng_amd.api import Mesh
    m = Mesh.square(32)
    k1 = m.coeff_blocks(16, 1e-2, 1e2, 5)
    k2 = m.coeff_blocks(16, 1e-2, 1e2, 5)
    assert np.array_equal(k1, k2) and k1.min() >= 1e-2 and k1.max() <= 1e2
    assert np.unique(k1).size == 256
    c = Mesh.cube(6)
    g = c.coeff_grf(0.1, 1.5, 256, 3)
    lg = np.log(g)
    assert np.all(np.isfinite(lg)) and 0.5 < lg.std() < 3.0


def _declared(header):
    txt = (INCLUDE / header).read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fe_[a-z0-9_]+)\s*\(", txt)))


def test_libraries_export_every_declared_symbol():
    import ctypes
    from feng_amd import _native
    host = _native.host()
    for name in _declared("fe_mesh.h"):
        assert hasattr(host, name), name
    lib = ctypes.CDLL(str(_native.LIB_DIR / "libfe_hip.so"))
    for name in _declared("fe_hip.h"):
        assert hasattr(lib, name), name


@pytest.mark.parametrize("dim", [2, 3])
def test_negative_jacobian_is_refused(dim):
    """A reversed element makes the mesh unusable, as feCncGeo::computeJacobians
    (feCncGeo.cpp:346-353, :388-395, :410-412) makes the reference's space
    creation fail: the check names the element, the numbering refuses it."""
    from feng_amd.api import FeError, Mesh, Numbering
    good = Mesh.square(4) if dim == 2 else Mesh.cube(2)
    assert good.check_jacobians() == (-1, None)
    conn = good.conn.reshape(good.E, good.nvpe).copy()
    bad_e = 5
    conn[bad_e, [0, 1]] = conn[bad_e, [1, 0]]  # swapped vertices: the orientation flips
    xyz = good.xyz.reshape(-1, 3)
    m = Mesh.from_arrays(dim, xyz, conn)
    e, det = m.check_jacobians()
    assert e == bad_e and det < 0.0
    v = xyz[conn[bad_e]]
    ref = np.linalg.det((v[1:dim + 1] - v[0])[:, :dim])  # the determinant of the element map
    assert abs(det - ref) <= 1e-14
    with pytest.raises(FeError):
        Numbering(m, [1 if dim == 2 else 4], [1])
    # a degenerate (zero-area) element is refused too
    xyz0 = xyz.copy()
    xyz0[conn[0, 2]] = xyz0[conn[0, 0]]
    e0, det0 = Mesh.from_arrays(dim, xyz0, good.conn.reshape(good.E, good.nvpe)).check_jacobians()
    assert e0 == 0 and det0 == 0.0


def test_msh_reader_keeps_named_boundary_groups():
    """The reference's Poiseuille mesh (data/poiseuille0.msh: "Domain" and the
    boundary groups "NoSlip", "Inlet", "Outlet", feMeshRead.cpp:1709-2005):
    every named group is read, each boundary segment keeps its group, and an
    essential set given per group (the reference's essential spaces uInlet,
    uNoSlip of navier_stokes.cpp:213-221) marks exactly the DOFs on those
    lines; generated meshes have the one group "Bord"."""
    import numpy as np

    from feng_amd.api import TRI_P1, TRI_VP2, Mesh, Numbering
    from tests.conftest import GOLDEN
    m = Mesh.read_msh(GOLDEN / "meshes" / "data" / "poiseuille0.msh")
    names, bg = m.boundary_groups()
    assert names == ["NoSlip", "Inlet", "Outlet"]
    assert m.E > 0 and bg.size == m.nB and set(bg.tolist()) == {0, 1, 2}
    xyz = m.xyz.reshape(-1, 3)
    seg = m.bconn.reshape(-1, 2)
    for g, on_line in enumerate([lambda p: np.isclose(p[:, 1], 0) | np.isclose(p[:, 1], 1),
                                 lambda p: np.isclose(p[:, 0], 0), lambda p: np.isclose(p[:, 0], 5)]):
        pts = xyz[seg[bg == g].ravel()]
        assert on_line(pts).all(), names[g]
    num = Numbering(m, [TRI_VP2, TRI_P1], [["Inlet", "NoSlip"], []])
    d = num.dof_xyz.reshape(-1, 3)
    ess = np.arange(num.nDofs) >= num.nInc
    wall = np.isclose(d[:, 0], 0) | np.isclose(d[:, 1], 0) | np.isclose(d[:, 1], 1)
    u = num.dof_field == 0
    assert np.array_equal(ess[u], wall[u])  # velocity: essential exactly on Inlet and NoSlip
    assert not ess[num.dof_field == 1].any()  # pressure: no essential DOF
    assert (np.isclose(d[:, 0], 5) & u & ~ess).any()  # the outlet's velocity DOFs stay unknowns
    # every group of the old meaning (essential = 1) equals the union of the names
    n1 = Numbering(m, [TRI_VP2, TRI_P1], [1, 0])
    n2 = Numbering(m, [TRI_VP2, TRI_P1], [["NoSlip", "Inlet", "Outlet"], []])
    assert n1.nInc == n2.nInc and all(np.array_equal(a, b) for a, b in zip(n1.adr, n2.adr))
    assert Mesh.square(4).boundary_groups()[0] == ["Bord"]


@pytest.mark.parametrize("kind", ["P1", "P2", "P3", "VP2"])
def test_numbering_periodic_pairs_between_boundary_groups(kind):
    """fe_numbering_periodic (feNumber::applyPeriodicity, feNumber.cpp:217-307)
    on the reference's Poiseuille mesh, Inlet (x = 0) -> Outlet (x = 5): the
    4 matching vertices pair their DOFs (per component), the 3 matching edges
    their edge DOFs; every slave DOF sits at its master's position + offset,
    same component; masters sorted and unique; an offset that matches nothing
    is an error, as the reference's exit."""
    from feng_amd import api as A
    from tests.conftest import GOLDEN
    m = A.Mesh.read_msh(GOLDEN / "meshes" / "data" / "poiseuille0.msh")
    k = {"P1": A.TRI_P1, "P2": A.TRI_P2, "P3": A.TRI_P3, "VP2": A.TRI_VP2}[kind]
    num = A.Numbering(m, [k], [["NoSlip"]])
    master, slave = num.periodic_dofs(0, "Inlet", "Outlet", (5.0, 0.0))
    nv, ne = {"P1": (1, 0), "P2": (1, 1), "P3": (1, 2), "VP2": (2, 2)}[kind]
    assert master.size == 4 * nv + 3 * ne
    assert np.all(np.diff(master) > 0)
    xyz = num.dof_xyz.reshape(-1, 3)
    assert np.abs(xyz[slave] - xyz[master] - [5.0, 0.0, 0.0]).max() <= 1e-8
    assert np.abs(xyz[master, 0]).max() <= 1e-12
    assert np.array_equal(num.dof_comp[slave], num.dof_comp[master])
    with pytest.raises(A.FeError):
        num.periodic_dofs(0, "Inlet", "Outlet", (4.9, 0.0))
