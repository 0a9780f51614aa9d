"""ctypes bindings of the CPU oracle (oracle/liboracle.so) -- the checker.

Loaded only by tests, smoke() and bench.py's cpu_baseline leg.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle" / "liboracle.so"

ip = C.POINTER(C.c_int)
dp = C.POINTER(C.c_double)


class OrFn(C.Structure):
    _fields_ = [("kind", C.c_int), ("c0", C.c_double), ("c1", C.c_double), ("data", dp)]


def P(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


class Oracle:
    def __init__(self, lib):
        self.L = lib
        L = lib
        L.or_tri_edges.argtypes = [C.c_int, ip, ip]
        L.or_color_elements.argtypes = [C.c_int, C.c_int, ip, C.c_int, ip]
        L.or_number.argtypes = [C.c_int, C.c_int, C.c_int, ip, C.c_int, ip, C.c_int, C.c_int, ip, ip,
                                C.c_int, ip, ip, C.POINTER(ip), ip, ip]
        L.or_csr.restype = C.c_int64
        L.or_csr.argtypes = [C.c_int, C.c_int, ip, ip, ip, C.POINTER(ip), C.POINTER(ip), ip, ip]
        L.or_quad_tri.argtypes = [C.c_int, dp, dp, dp]
        L.or_quad_tet.argtypes = [C.c_int, dp, dp, dp, dp]
        L.or_assemble_diffusion.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip, C.c_int,
                                            OrFn, OrFn, dp, ip, dp, dp]
        L.or_element_diffusion.argtypes = [C.c_int, dp, C.c_int, OrFn, C.c_int, dp]
        L.or_element_mass.argtypes = [C.c_int, dp, C.c_int, OrFn, C.c_int, dp]
        L.or_assemble_advection.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip, C.c_int,
                                            OrFn, OrFn, OrFn, dp, dp, dp]
        L.or_assemble_nonlinear_diffusion.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip,
                                                      C.c_int, C.c_int, dp, C.c_int, dp, dp, dp, dp]
        L.or_basis.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp]
        L.or_assemble_mixed.argtypes = [C.c_int, dp, C.c_int, ip, ip, ip, C.c_int, ip, ip, C.c_int, C.c_double,
                                        C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.or_l2_error_qp.argtypes = [C.c_int, dp, C.c_int, ip, ip, dp, C.c_int, dp]
        L.or_l2_error_qp.restype = C.c_double
        L.or_assemble_mass_power.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip, C.c_int,
                                             OrFn, C.c_double, dp, dp, dp]
        L.or_assemble_vector_transient_mass.argtypes = [C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip, C.c_int,
                                                        C.c_double, C.c_double, dp, dp, dp]
        L.or_assemble_transient_mass.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip, C.c_int,
                                                 OrFn, C.c_double, dp, dp, dp]
        L.or_assemble_ns.argtypes = [C.c_int, dp, C.c_int, ip, ip, ip, C.c_int, ip, ip, C.c_int, C.c_double,
                                     dp, ip, dp, dp]
        L.or_spmv.argtypes = [C.c_int, ip, ip, dp, dp, dp]
        L.or_cg_jacobi.argtypes = [C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int, dp]
        L.or_cg_jacobi_omp.argtypes = [C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int, C.c_int, dp]
        L.or_assemble_diffusion_omp.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip, C.c_int,
                                                OrFn, OrFn, dp, ip, C.c_int, C.c_int, dp, dp]
        L.or_ic0.argtypes = [C.c_int, ip, ip, dp, dp]
        L.or_pcg_ic0.argtypes = [C.c_int, ip, ip, dp, dp, dp, dp, C.c_double, C.c_int, dp]
        L.or_ilu0.argtypes = [C.c_int, ip, ip, dp, dp]
        L.or_gmres_ilu0.argtypes = [C.c_int, ip, ip, dp, dp, dp, dp, C.c_int, C.c_double, C.c_int, dp]
        L.or_l2_error_sinsin.restype = C.c_double
        L.or_l2_error_sinsin.argtypes = [C.c_int, dp, C.c_int, ip, ip, dp, C.c_int]

    # -- helpers ------------------------------------------------------------
    def tri_edges(self, conn):
        conn = np.ascontiguousarray(conn, np.int32)
        E = conn.size // 3
        edges = np.zeros(3 * E, np.int32)
        ne = self.L.or_tri_edges(E, P(conn, C.c_int), P(edges, C.c_int))
        return edges, ne

    def color(self, conn, nvpe, V):
        conn = np.ascontiguousarray(conn, np.int32)
        E = conn.size // nvpe
        col = np.zeros(E, np.int32)
        nc = self.L.or_color_elements(E, nvpe, P(conn, C.c_int), V, P(col, C.c_int))
        return col, nc

    def number(self, V, conn, nvpe, edges, nEdges, bconn, bnv, bedge, kinds, essential):
        nloc = {1: 3, 2: 6, 3: 12, 4: 4}
        conn = np.ascontiguousarray(conn, np.int32)
        E = conn.size // nvpe
        outs = [np.zeros(E * nloc[k], np.int32) for k in kinds]
        arr = (ip * len(kinds))(*[P(o, C.c_int) for o in outs])
        k = np.array(kinds, np.int32)
        e = np.array(essential, np.int32)
        nInc, nDofs = C.c_int(), C.c_int()
        edges = np.ascontiguousarray(edges if edges is not None and len(edges) else np.zeros(1), np.int32)
        bedge = np.ascontiguousarray(bedge if bedge is not None and len(bedge) else np.zeros(1), np.int32)
        bconn = np.ascontiguousarray(bconn, np.int32)
        rc = self.L.or_number(V, E, nvpe, P(conn, C.c_int), nEdges, P(edges, C.c_int), bconn.size // bnv, bnv,
                              P(bconn, C.c_int), P(bedge, C.c_int), len(kinds), P(k, C.c_int), P(e, C.c_int),
                              arr, C.byref(nInc), C.byref(nDofs))
        assert rc == 0
        return nInc.value, nDofs.value, outs

    def csr(self, n, forms):
        """forms: list of (adrI, nI, adrJ, nJ) with E elements each."""
        nf = len(forms)
        Es = (C.c_int * nf)(*[f[0].size // f[1] for f in forms])
        nI = (C.c_int * nf)(*[f[1] for f in forms])
        nJ = (C.c_int * nf)(*[f[3] for f in forms])
        keep = [np.ascontiguousarray(f[0], np.int32) for f in forms] + [np.ascontiguousarray(f[2], np.int32) for f in forms]
        aI = (ip * nf)(*[P(keep[i], C.c_int) for i in range(nf)])
        aJ = (ip * nf)(*[P(keep[nf + i], C.c_int) for i in range(nf)])
        nnz = self.L.or_csr(n, nf, Es, nI, nJ, aI, aJ, None, None)
        ia = np.zeros(n + 1, np.int32)
        ja = np.zeros(nnz, np.int32)
        self.L.or_csr(n, nf, Es, nI, nJ, aI, aJ, P(ia, C.c_int), P(ja, C.c_int))
        return ia, ja

    def assemble_diffusion(self, kind, xyz, conn, adr, n, ia, ja, deg, coeff, src, u, order=None):
        nvpe = 4 if kind == 4 else 3
        E = conn.size // nvpe
        val = np.zeros(ja.size)
        rhs = np.zeros(n)
        keep = []

        def fn(f):
            kind_, c0, c1, data = f
            o = OrFn(kind_, c0, c1, None)
            if data is not None:
                d = np.ascontiguousarray(data, np.float64)
                keep.append(d)
                o.data = P(d, C.c_double)
            return o
        order = None if order is None else np.ascontiguousarray(order, np.int32)
        rc = self.L.or_assemble_diffusion(kind, xyz.size // 3, P(xyz, C.c_double), E, P(conn, C.c_int),
                                          P(adr, C.c_int), n, P(ia, C.c_int), P(ja, C.c_int), deg, fn(coeff),
                                          fn(src), P(u, C.c_double), P(order, C.c_int), P(val, C.c_double),
                                          P(rhs, C.c_double))
        assert rc == 0, rc
        return val, rhs

    def assemble_diffusion_omp(self, kind, xyz, conn, adr, n, ia, ja, deg, coeff, src, u, colorOf, ncolors,
                               threads):
        nvpe = 4 if kind == 4 else 3
        E = conn.size // nvpe
        val = np.zeros(ja.size)
        rhs = np.zeros(n)
        keep = []

        def fn(f):
            kind_, c0, c1, data = f
            o = OrFn(kind_, c0, c1, None)
            if data is not None:
                d = np.ascontiguousarray(data, np.float64)
                keep.append(d)
                o.data = P(d, C.c_double)
            return o
        colorOf = np.ascontiguousarray(colorOf, np.int32)
        rc = self.L.or_assemble_diffusion_omp(kind, xyz.size // 3, P(xyz, C.c_double), E, P(conn, C.c_int),
                                              P(adr, C.c_int), n, P(ia, C.c_int), P(ja, C.c_int), deg, fn(coeff),
                                              fn(src), P(u, C.c_double), P(colorOf, C.c_int), ncolors, threads,
                                              P(val, C.c_double), P(rhs, C.c_double))
        assert rc == 0, rc
        return val, rhs

    def assemble_transient_mass(self, kind, xyz, conn, adr, n, ia, ja, deg, coeff, c0, udot, val, rhs):
        """Adds the transient-mass terms into val (None: residual only) and rhs."""
        nvpe = 4 if kind == 4 else 3
        E = conn.size // nvpe
        kind_, a0, a1, data = coeff
        o = OrFn(kind_, a0, a1, None)
        if data is not None:
            d = np.ascontiguousarray(data, np.float64)
            o.data = P(d, C.c_double)
        udot = np.ascontiguousarray(udot, np.float64)
        rc = self.L.or_assemble_transient_mass(kind, xyz.size // 3, P(xyz, C.c_double), E, P(conn, C.c_int),
                                               P(adr, C.c_int), n, P(ia, C.c_int), P(ja, C.c_int), deg, o, c0,
                                               P(udot, C.c_double), None if val is None else P(val, C.c_double),
                                               P(rhs, C.c_double))
        assert rc == 0, rc

    def assemble_vector_transient_mass(self, xyz, conn, adrU, n, ia, ja, deg, m, c0, udot, val, rhs):
        udot = np.ascontiguousarray(udot, np.float64)
        rc = self.L.or_assemble_vector_transient_mass(xyz.size // 3, P(xyz, C.c_double), conn.size // 3,
                                                      P(conn, C.c_int), P(adrU, C.c_int), n, P(ia, C.c_int),
                                                      P(ja, C.c_int), deg, m, c0, P(udot, C.c_double),
                                                      None if val is None else P(val, C.c_double),
                                                      P(rhs, C.c_double))
        assert rc == 0, rc

    def assemble_mass_power(self, kind, xyz, conn, adr, n, ia, ja, deg, coeff, p, u, val, rhs):
        nvpe = 4 if kind == 4 else 3
        o = OrFn(coeff[0], coeff[1], coeff[2], None)
        u = np.ascontiguousarray(u, np.float64)
        rc = self.L.or_assemble_mass_power(kind, xyz.size // 3, P(xyz, C.c_double), conn.size // nvpe,
                                           P(conn, C.c_int), P(adr, C.c_int), n, P(ia, C.c_int), P(ja, C.c_int),
                                           deg, o, p, P(u, C.c_double), None if val is None else P(val, C.c_double),
                                           P(rhs, C.c_double))
        assert rc == 0, rc

    def assemble_advection(self, kind, xyz, conn, adr, n, ia, ja, deg, vel, u, val, rhs):
        nvpe = 4 if kind == 4 else 3
        keep = []

        def fn(f):
            o = OrFn(f[0], f[1], f[2], None)
            if f[3] is not None:
                d = np.ascontiguousarray(f[3], np.float64)
                keep.append(d)
                o.data = P(d, C.c_double)
            return o
        v = [fn(f) for f in vel] + [OrFn(0, 0.0, 0.0, None)] * (3 - len(vel))
        u = np.ascontiguousarray(u, np.float64)
        rc = self.L.or_assemble_advection(kind, xyz.size // 3, P(xyz, C.c_double), conn.size // nvpe,
                                          P(conn, C.c_int), P(adr, C.c_int), n, P(ia, C.c_int), P(ja, C.c_int),
                                          deg, v[0], v[1], v[2], P(u, C.c_double),
                                          None if val is None else P(val, C.c_double), P(rhs, C.c_double))
        assert rc == 0, rc

    def assemble_nonlinear_diffusion(self, kind, xyz, conn, adr, n, ia, ja, deg, k, dk, u, val, rhs):
        """k, dk: polynomial coefficients in u (constant term first)."""
        nvpe = 4 if kind == 4 else 3
        k = np.ascontiguousarray(k, np.float64)
        dk = np.ascontiguousarray(dk, np.float64)
        u = np.ascontiguousarray(u, np.float64)
        rc = self.L.or_assemble_nonlinear_diffusion(kind, xyz.size // 3, P(xyz, C.c_double), conn.size // nvpe,
                                                    P(conn, C.c_int), P(adr, C.c_int), n, P(ia, C.c_int),
                                                    P(ja, C.c_int), deg, k.size, P(k, C.c_double), dk.size,
                                                    P(dk, C.c_double), P(u, C.c_double),
                                                    None if val is None else P(val, C.c_double), P(rhs, C.c_double))
        assert rc == 0, rc

    def basis(self, kind, r, s, t=0.0):
        """(L, dL/dr, dL/ds) of the kind's shape functions at (r, s[, t])."""
        L, dr, ds, dt = (np.zeros(15) for _ in range(4))
        nf = self.L.or_basis(kind, r, s, t, P(L, C.c_double), P(dr, C.c_double), P(ds, C.c_double),
                             P(dt, C.c_double))
        assert nf > 0, kind
        return L[:nf], dr[:nf], ds[:nf]

    def assemble_mixed(self, kind, xyz, conn, adrI, adrJ, n, ia, ja, deg, cm, cp, p, cg, u, val, rhs):
        """Mixed mass / mass power / grad-grad of (test adrI, trial adrJ), added into val / rhs."""
        nvpe = 4 if kind == 4 else 3
        u = np.ascontiguousarray(u, np.float64)
        rc = self.L.or_assemble_mixed(kind, P(xyz, C.c_double), conn.size // nvpe, P(conn, C.c_int),
                                      P(adrI, C.c_int), P(adrJ, C.c_int), n, P(ia, C.c_int), P(ja, C.c_int), deg,
                                      cm, cp, p, cg, P(u, C.c_double),
                                      None if val is None else P(val, C.c_double), P(rhs, C.c_double))
        assert rc == 0, rc

    def l2_error_qp(self, kind, xyz, conn, adr, u, deg, ue):
        ue = np.ascontiguousarray(ue, np.float64)
        return self.L.or_l2_error_qp(kind, P(xyz, C.c_double), conn.size // 3, P(conn, C.c_int), P(adr, C.c_int),
                                     P(np.ascontiguousarray(u, np.float64), C.c_double), deg, P(ue, C.c_double))

    def element_mass(self, kind, coords, deg, coeff=(0, 1.0, 0.0, None), elem=0):
        nf = {1: 3, 2: 6, 4: 4}[kind]
        Ae = np.zeros(nf * nf)
        c = np.ascontiguousarray(coords, np.float64)
        rc = self.L.or_element_mass(kind, P(c, C.c_double), deg, OrFn(coeff[0], coeff[1], coeff[2], None), elem,
                                    P(Ae, C.c_double))
        assert rc == nf, rc
        return Ae.reshape(nf, nf)

    def cg_jacobi_omp(self, ia, ja, val, b, rtol, maxit, threads):
        x = np.zeros(b.size)
        rr = C.c_double()
        it = self.L.or_cg_jacobi_omp(b.size, P(ia, C.c_int), P(ja, C.c_int), P(val, C.c_double), P(b, C.c_double),
                                     P(x, C.c_double), rtol, maxit, threads, C.byref(rr))
        return x, it, rr.value

    def spmv(self, ia, ja, val, x):
        y = np.zeros(ia.size - 1)
        self.L.or_spmv(ia.size - 1, P(ia, C.c_int), P(ja, C.c_int), P(val, C.c_double), P(x, C.c_double),
                       P(y, C.c_double))
        return y

    def cg_jacobi(self, ia, ja, val, b, rtol, maxit):
        x = np.zeros(b.size)
        rr = C.c_double()
        it = self.L.or_cg_jacobi(b.size, P(ia, C.c_int), P(ja, C.c_int), P(val, C.c_double), P(b, C.c_double),
                                 P(x, C.c_double), rtol, maxit, C.byref(rr))
        return x, it, rr.value

    def ic0(self, ia, ja, val):
        l = np.zeros(val.size)
        rc = self.L.or_ic0(ia.size - 1, P(ia, C.c_int), P(ja, C.c_int), P(val, C.c_double), P(l, C.c_double))
        return l, rc

    def pcg_ic0(self, ia, ja, val, lval, b, rtol, maxit):
        x = np.zeros(b.size)
        rr = C.c_double()
        it = self.L.or_pcg_ic0(b.size, P(ia, C.c_int), P(ja, C.c_int), P(val, C.c_double), P(lval, C.c_double),
                               P(b, C.c_double), P(x, C.c_double), rtol, maxit, C.byref(rr))
        return x, it, rr.value

    def ilu0(self, ia, ja, val):
        lu = np.zeros(val.size)
        rc = self.L.or_ilu0(ia.size - 1, P(ia, C.c_int), P(ja, C.c_int), P(val, C.c_double), P(lu, C.c_double))
        return lu, rc

    def gmres_ilu0(self, ia, ja, val, lu, b, restart, rtol, maxit):
        x = np.zeros(b.size)
        rr = C.c_double()
        it = self.L.or_gmres_ilu0(b.size, P(ia, C.c_int), P(ja, C.c_int), P(val, C.c_double), P(lu, C.c_double),
                                  P(b, C.c_double), P(x, C.c_double), restart, rtol, maxit, C.byref(rr))
        return x, it, rr.value

    def l2_error_sinsin(self, kind, xyz, conn, adr, u, deg):
        E = conn.size // 3
        return self.L.or_l2_error_sinsin(kind, P(xyz, C.c_double), E, P(conn, C.c_int), P(adr, C.c_int),
                                         P(u, C.c_double), deg)


def load_oracle():
    if not ORACLE.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "liboracle.so"], check=True)
    return Oracle(C.CDLL(str(ORACLE)))
