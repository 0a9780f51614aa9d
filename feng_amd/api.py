"""Python layer over the C ABIs: meshes, numbering and the GPU linear system.

Method names follow the reference's feLinearSystem interface
(feLinearSystem.h:85-168): setToZero, assembleMatrices, assembleResiduals,
getRHSMaxNorm, solve, correctSolution.  Every call goes to libfe_hip.so; there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N

# element kinds (include/fe_mesh.h)
TRI_P1, TRI_P2, TRI_VP2, TET_P1, TRI_P3, TRI_P4 = 1, 2, 3, 4, 5, 6
NLOC = {TRI_P1: 3, TRI_P2: 6, TRI_VP2: 12, TET_P1: 4, TRI_P3: 10, TRI_P4: 15}

# weak forms (elementSystemType values, feSysElm.h:13-77)
SOURCE, VECTOR_SOURCE, MASS, TRANSIENT_MASS, TRANSIENT_VECTOR_MASS, DIFFUSION, VECTOR_DIFFUSION = 0, 2, 6, 13, 15, 16, 18
ADVECTION, GRAD_SOURCE, VECTOR_MASS, NONLINEAR_DIFFUSION = 20, 4, 10, 19
MIXED_MASS, MIXED_MASS_POWER, MIXED_GRADGRAD = 8, 9, 28
VECTOR_CONVECTIVE_ACCELERATION, DIV_NEWTONIAN_STRESS = 22, 25
MIXED_GRADIENT, MIXED_DIVERGENCE = 26, 31

# coefficient kinds
FN_CONST, FN_SINSIN, FN_BLOCKS, FN_SMOOTH3, FN_ELEM, FN_QP, FN_KOV_U, FN_KOV_P, FN_POLY_U = range(9)

KSP_CG, KSP_GMRES, KSP_DIRECT = 0, 1, 2
PC_NONE, PC_JACOBI, PC_IC0, PC_ILU0, PC_BJ_ILU0, PC_LEARNED, PC_AMG, PC_FIELDSPLIT = range(8)
PREC_FP64, PREC_FP32_IR = 0, 1

# norm types (the reference's normType values, feNorm.h:10-85)
(NORM_L1, NORM_L1_ERROR, NORM_L2, NORM_VECTOR_L2, NORM_L2_ERROR, NORM_VECTOR_L2_ERROR, NORM_LINF, NORM_LINF_ERROR,
 NORM_LINF_UDOT_ERROR, NORM_VECTOR_LINF_ERROR, NORM_VECTOR_LINF_UDOT_ERROR, NORM_SEMI_H1, NORM_SEMI_H1_ERROR,
 NORM_VECTOR_SEMI_H1_ERROR, NORM_H1, NORM_H1_ERROR, NORM_AREA) = (1, 2, 3, 4, 5, 7, 8, 9, 10, 11, 12, 13, 14, 16, 17,
                                                                  18, 19)
NORM_INTEGRAL, NORM_INTEGRAL_SOL_DEPENDENT_F, NORM_INTEGRAL_DT, NORM_INTEGRAL_F = 20, 21, 23, 25


class FeError(RuntimeError):
    pass


def _chk(rc, what, ctx=None):
    if rc != 0:
        msg = ""
        if ctx is not None:
            msg = N.hip().fe_hip_last_error(ctx).decode(errors="replace")
        raise FeError(f"{what} failed with status {rc}: {msg}")


class Mesh:
    """Synthetic or user mesh (libfe_host.so)."""

    def __init__(self, handle):
        self.h = handle
        L = N.host()
        v = [C.c_int() for _ in range(7)]
        _chk(L.fe_mesh_info(self.h, *[C.byref(x) for x in v]), "fe_mesh_info")
        self.dim, self.V, self.E, self.nvpe, self.nB, self.bnv, self.nEdges = [x.value for x in v]
        self.xyz = np.zeros(3 * self.V)
        self.conn = np.zeros(self.E * self.nvpe, np.int32)
        self.bconn = np.zeros(max(1, self.nB * self.bnv), np.int32)
        self.edges = np.zeros(max(1, 3 * self.E if self.dim == 2 else 1), np.int32)
        self.bedge = np.zeros(max(1, self.nB), np.int32)
        _chk(L.fe_mesh_get(self.h, N.ptr(self.xyz, C.c_double), N.ptr(self.conn, C.c_int),
                           N.ptr(self.bconn, C.c_int), N.ptr(self.edges, C.c_int),
                           N.ptr(self.bedge, C.c_int)), "fe_mesh_get")
        self.bconn = self.bconn[: self.nB * self.bnv]
        self.bedge = self.bedge[: self.nB]
        if self.dim != 2:
            self.edges = np.zeros(0, np.int32)

    @staticmethod
    def square(N_, jitter=0.0, seed=0):
        h = C.c_void_p()
        _chk(N.host().fe_mesh_square(N_, jitter, seed, C.byref(h)), "fe_mesh_square")
        return Mesh(h)

    @staticmethod
    def cube(N_):
        h = C.c_void_p()
        _chk(N.host().fe_mesh_cube(N_, C.byref(h)), "fe_mesh_cube")
        return Mesh(h)

    @staticmethod
    def read_msh(path):
        h = C.c_void_p()
        _chk(N.host().fe_mesh_read_msh(str(path).encode(), C.byref(h)), "fe_mesh_read_msh")
        return Mesh(h)

    def boundary_groups(self):
        """(names of the boundary physical groups, group of each boundary element)."""
        n = C.c_int()
        buf = C.create_string_buffer(4096)
        bg = np.zeros(max(1, self.nB), np.int32)
        _chk(N.host().fe_mesh_boundary_groups(self.h, C.byref(n), buf, 4096, N.ptr(bg, C.c_int)),
             "fe_mesh_boundary_groups")
        names = buf.value.decode().split("\n") if n.value else []
        return names, bg[:self.nB]

    @staticmethod
    def from_arrays(dim, xyz, conn, bconn=None):
        """A mesh from vertex coordinates (V x 3), element connectivity
        (E x nvpe, 0-based) and optional boundary elements."""
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        nvpe = 3 if dim == 2 else 4
        conn = np.ascontiguousarray(conn, np.int32).reshape(-1)
        bnv = 2 if dim == 2 else 3
        b = np.ascontiguousarray(bconn if bconn is not None else np.zeros(1), np.int32).reshape(-1)
        nB = 0 if bconn is None else b.size // bnv
        h = C.c_void_p()
        _chk(N.host().fe_mesh_from_arrays(dim, xyz.size // 3, N.ptr(xyz, C.c_double), conn.size // nvpe, nvpe,
                                          N.ptr(conn, C.c_int), nB, bnv, N.ptr(b, C.c_int), C.byref(h)),
             "fe_mesh_from_arrays")
        return Mesh(h)

    def check_jacobians(self):
        """(first element with a Jacobian determinant <= 0 or -1, its determinant)."""
        e, d = C.c_int(), C.c_double()
        N.host().fe_mesh_check_jacobians(self.h, C.byref(e), C.byref(d))
        return e.value, (d.value if e.value >= 0 else None)

    def write_msh(self, path):
        _chk(N.host().fe_mesh_write_msh(self.h, str(path).encode()), "fe_mesh_write_msh")

    def coloring(self):
        col = np.zeros(self.E, np.int32)
        nc = C.c_int()
        _chk(N.host().fe_color_elements(self.h, N.ptr(col, C.c_int), C.byref(nc)), "fe_color_elements")
        return col, nc.value

    def coeff_blocks(self, B, lo, hi, seed):
        k = np.zeros(self.E)
        _chk(N.host().fe_coeff_blocks(self.h, B, lo, hi, seed, N.ptr(k, C.c_double)), "fe_coeff_blocks")
        return k

    def coeff_grf(self, ell, sigma, nfeat, seed):
        k = np.zeros(self.E)
        _chk(N.host().fe_coeff_grf_cube(self.h, ell, sigma, nfeat, seed, N.ptr(k, C.c_double)),
             "fe_coeff_grf_cube")
        return k

    def __del__(self):
        try:
            if self.h:
                N.host().fe_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Numbering:
    """feMetaNumber counterpart (libfe_host.so).  essential[f]: 0 / 1 (no /
    every boundary group essential for field f), or the names of the boundary
    physical groups on which field f is essential (the reference's essential
    spaces of the field, one per group)."""

    def __init__(self, mesh: Mesh, kinds, essential):
        self.mesh = mesh
        self.kinds = list(kinds)
        k = np.array(kinds, np.int32)
        self.h = C.c_void_p()
        if any(isinstance(x, (list, tuple)) for x in essential):
            names, _ = mesh.boundary_groups()
            mask = np.zeros(len(kinds), np.uint32)
            for f, x in enumerate(essential):
                if isinstance(x, (list, tuple)):
                    for g in x:
                        mask[f] |= np.uint32(1 << names.index(g))
                elif x:
                    mask[f] = np.uint32(0xFFFFFFFF)
            _chk(N.host().fe_numbering_create_groups(mesh.h, len(kinds), N.ptr(k, C.c_int),
                                                     mask.ctypes.data_as(C.POINTER(C.c_uint)), C.byref(self.h)),
                 "fe_numbering_create_groups")
        else:
            e = np.array(essential, np.int32)
            _chk(N.host().fe_numbering_create(mesh.h, len(kinds), N.ptr(k, C.c_int), N.ptr(e, C.c_int),
                                              C.byref(self.h)), "fe_numbering_create")
        a, b = C.c_int(), C.c_int()
        N.host().fe_numbering_info(self.h, C.byref(a), C.byref(b))
        self.nInc, self.nDofs = a.value, b.value
        self.adr = []
        for f, kind in enumerate(self.kinds):
            x = np.zeros(mesh.E * NLOC[kind], np.int32)
            _chk(N.host().fe_numbering_adr(self.h, f, N.ptr(x, C.c_int)), "fe_numbering_adr")
            self.adr.append(x)
        self.dof_xyz = np.zeros(3 * self.nDofs)
        self.dof_field = np.zeros(self.nDofs, np.int32)
        self.dof_comp = np.zeros(self.nDofs, np.int32)
        N.host().fe_numbering_dof_info(self.h, N.ptr(self.dof_xyz, C.c_double),
                                       N.ptr(self.dof_field, C.c_int), N.ptr(self.dof_comp, C.c_int))

    def periodic_dofs(self, field, master_group, slave_group, offset):
        """feNumber::applyPeriodicity: (master, slave) DOF arrays of `field`
        between two boundary groups (names or indices), the slave side at
        master + offset; for LinearSystemHIP.set_periodic_dofs."""
        names, _ = self.mesh.boundary_groups()
        gi = lambda g: names.index(g) if isinstance(g, str) else int(g)
        off = np.zeros(3)
        off[:len(offset)] = offset
        cnt = C.c_int()
        fn = N.host().fe_numbering_periodic
        _chk(fn(self.mesh.h, self.h, field, gi(master_group), gi(slave_group), N.ptr(off, C.c_double), None, None, 0,
                C.byref(cnt)), "fe_numbering_periodic")
        master, slave = np.zeros(cnt.value, np.int32), np.zeros(cnt.value, np.int32)
        _chk(fn(self.mesh.h, self.h, field, gi(master_group), gi(slave_group), N.ptr(off, C.c_double),
                N.ptr(master, C.c_int), N.ptr(slave, C.c_int), cnt.value, C.byref(cnt)), "fe_numbering_periodic")
        return master, slave

    def __del__(self):
        try:
            if self.h:
                N.host().fe_numbering_destroy(self.h)
                self.h = None
        except Exception:
            pass


@dataclass
class Fn:
    kind: int = FN_CONST
    c0: float = 0.0
    c1: float = 0.0
    data: np.ndarray | None = None


@dataclass
class SolveResult:
    iterations: int
    outer_iterations: int
    converged: bool
    reason: int
    relres: float
    norm_dx_inf: float
    norm_rhs_inf: float
    norm_axb_inf: float
    ms_solve: float
    ms_setup: float = 0.0


class LinearSystemHIP:
    """GPU linear system (feLinearSystem drop-in, one MI355X)."""

    def __init__(self, device=0):
        self.L = N.hip()
        self.ctx = C.c_void_p()
        _chk(self.L.fe_hip_ctx_create(device, C.byref(self.ctx)), "fe_hip_ctx_create (is a GPU present?)")
        self.nnz = 0
        self._keep = []

    def _c(self, rc, what):
        _chk(rc, what, self.ctx)

    def upload_mesh(self, mesh: Mesh, coloring=None):
        # the C-ABI reads 3 V coordinates and E nvpe vertex ids through raw pointers
        if np.asarray(mesh.xyz).size != 3 * mesh.V or np.asarray(mesh.conn).size != mesh.E * mesh.nvpe:
            raise ValueError(f"mesh arrays do not match V = {mesh.V}, E = {mesh.E}, nvpe = {mesh.nvpe}")
        self.E = mesh.E
        self._c(self.L.fe_hip_upload_mesh(self.ctx, mesh.dim, mesh.V, N.ptr(mesh.xyz, C.c_double), mesh.E,
                                          mesh.nvpe, N.ptr(mesh.conn, C.c_int)), "fe_hip_upload_mesh")
        if coloring is not None:
            col = np.ascontiguousarray(coloring, np.int32)
            self._c(self.L.fe_hip_set_coloring(self.ctx, N.ptr(col, C.c_int)), "fe_hip_set_coloring")

    def add_field(self, kind, adr):
        adr = np.ascontiguousarray(adr, np.int32)
        f = C.c_int()
        self._c(self.L.fe_hip_add_field(self.ctx, kind, N.ptr(adr, C.c_int), C.byref(f)), "fe_hip_add_field")
        return f.value

    def set_sizes(self, nInc, nDofs):
        self.n, self.nDofs = nInc, nDofs
        self._c(self.L.fe_hip_set_sizes(self.ctx, nInc, nDofs), "fe_hip_set_sizes")

    def add_form(self, sys_type, fields, degree, coeffs):
        arr = (N.FeHipFn * max(1, len(coeffs)))()
        for i, fn in enumerate(coeffs):
            arr[i].kind, arr[i].c0, arr[i].c1 = fn.kind, fn.c0, fn.c1
            if fn.data is not None:
                d = np.ascontiguousarray(fn.data, np.float64)
                self._keep.append(d)
                arr[i].data = N.ptr(d, C.c_double)
                arr[i].count = d.size
        fl = np.array(fields, np.int32)
        out = C.c_int()
        self._c(self.L.fe_hip_add_form(self.ctx, sys_type, len(fields), N.ptr(fl, C.c_int), degree,
                                       len(coeffs), arr, C.byref(out)), "fe_hip_add_form")
        self._keep = []
        return out.value

    def build_csr(self):
        nnz = C.c_int64()
        self._c(self.L.fe_hip_build_csr(self.ctx, C.byref(nnz)), "fe_hip_build_csr")
        self.nnz = nnz.value
        return self.nnz

    def get_csr(self):
        ia = np.zeros(self.n + 1, np.int32)
        ja = np.zeros(self.nnz, np.int32)
        self._c(self.L.fe_hip_get_csr(self.ctx, N.ptr(ia, C.c_int), N.ptr(ja, C.c_int)), "fe_hip_get_csr")
        return ia, ja

    def set_solution(self, u):
        u = np.ascontiguousarray(u, np.float64)
        self._c(self.L.fe_hip_set_solution(self.ctx, N.ptr(u, C.c_double)), "fe_hip_set_solution")

    def get_solution(self):
        u = np.zeros(self.nDofs)
        self._c(self.L.fe_hip_get_solution(self.ctx, N.ptr(u, C.c_double)), "fe_hip_get_solution")
        return u

    # feLinearSystem virtuals ------------------------------------------------
    def setToZero(self, matrix=True, residual=True):
        self._c(self.L.fe_hip_set_to_zero(self.ctx, int(matrix), int(residual)), "fe_hip_set_to_zero")

    def assemble(self, what=3, precision=PREC_FP64):
        self._c(self.L.fe_hip_assemble(self.ctx, what, precision), "fe_hip_assemble")

    def assembleMatrices(self, precision=PREC_FP64):
        self.assemble(1, precision)

    def assembleResiduals(self):
        self.assemble(2)

    def getRHSMaxNorm(self):
        v = C.c_double()
        self._c(self.L.fe_hip_rhs_max_norm(self.ctx, C.byref(v)), "fe_hip_rhs_max_norm")
        return v.value

    def get_system(self):
        val = np.zeros(self.nnz)
        rhs = np.zeros(self.n)
        self._c(self.L.fe_hip_get_system(self.ctx, N.ptr(val, C.c_double), N.ptr(rhs, C.c_double)),
                "fe_hip_get_system")
        return val, rhs

    def get_system32(self):
        """The FP32 values kept by an FP32_IR assembly (fe_hip_get_system32)."""
        val = np.zeros(self.nnz, np.float32)
        self._c(self.L.fe_hip_get_system32(self.ctx, N.ptr(val, C.c_float)), "fe_hip_get_system32")
        return val

    def set_system(self, val=None, rhs=None):
        v = np.ascontiguousarray(val, np.float64) if val is not None else None
        r = np.ascontiguousarray(rhs, np.float64) if rhs is not None else None
        self._c(self.L.fe_hip_set_system(self.ctx, N.ptr(v, C.c_double) if v is not None else None,
                                         N.ptr(r, C.c_double) if r is not None else None), "fe_hip_set_system")

    def solve(self, ksp=KSP_CG, pc=PC_JACOBI, precision=PREC_FP64, rtol=1e-8, atol=1e-300,
              max_iter=10000, restart=50, inner_rtol=1e-3, check_every=32, want_du=False):
        o = N.FeHipSolverOpts()
        self.L.fe_hip_default_opts(C.byref(o))
        o.ksp, o.pc, o.precision, o.rtol, o.atol = ksp, pc, precision, rtol, atol
        o.max_iter, o.restart, o.inner_rtol, o.check_every = max_iter, restart, inner_rtol, check_every
        st = N.FeHipSolveStats()
        du = np.zeros(self.n) if want_du else None
        self._c(self.L.fe_hip_solve(self.ctx, C.byref(o), N.ptr(du, C.c_double) if du is not None else None,
                                    C.byref(st)), "fe_hip_solve")
        res = SolveResult(st.iterations, st.outer_iterations, bool(st.converged), st.reason, st.relres_fp64,
                          st.norm_dx_inf, st.norm_rhs_inf, st.norm_axb_inf, st.ms_solve,
                          st.ms_setup)
        return (res, du) if want_du else res

    # constraints (the reference's constrainEssentialComponents / applyPeriodicity)
    def set_constrained_dofs(self, dofs):
        d = np.ascontiguousarray(dofs, np.int32)
        self._c(self.L.fe_hip_set_constrained_dofs(self.ctx, d.size, N.ptr(d, C.c_int)),
                "fe_hip_set_constrained_dofs")

    def constrainEssentialComponents(self):
        self._c(self.L.fe_hip_constrain_essential_components(self.ctx), "fe_hip_constrain_essential_components")

    def set_periodic_dofs(self, master, slave):
        m = np.ascontiguousarray(master, np.int32)
        s_ = np.ascontiguousarray(slave, np.int32)
        assert m.size == s_.size
        self._c(self.L.fe_hip_set_periodic_dofs(self.ctx, m.size, N.ptr(m, C.c_int), N.ptr(s_, C.c_int)),
                "fe_hip_set_periodic_dofs")

    def applyPeriodicity(self):
        self._c(self.L.fe_hip_apply_periodicity(self.ctx), "fe_hip_apply_periodicity")

    def correctSolution(self):
        self._c(self.L.fe_hip_correct_solution(self.ctx), "fe_hip_correct_solution")

    def set_profiling(self, on=True):
        self._c(self.L.fe_hip_set_profiling(self.ctx, int(on)), "fe_hip_set_profiling")

    def kernel_stats(self, which=0):
        ms = C.c_double()
        cnt = C.c_int64()
        self._c(self.L.fe_hip_kernel_stats(self.ctx, which, C.byref(ms), C.byref(cnt)), "fe_hip_kernel_stats")
        return ms.value, cnt.value

    def pc_info(self):
        a, b = C.c_int(), C.c_int()
        self._c(self.L.fe_hip_pc_info(self.ctx, C.byref(a), C.byref(b)), "fe_hip_pc_info")
        return a.value, b.value

    def get_factor(self):
        lu = np.zeros(self.nnz)
        self._c(self.L.fe_hip_get_factor(self.ctx, N.ptr(lu, C.c_double)), "fe_hip_get_factor")
        return lu

    def set_tuning(self, key, value):
        self._c(self.L.fe_hip_set_tuning(self.ctx, key.encode(), int(value)), "fe_hip_set_tuning")

    def get_tuning(self, key):
        v = C.c_int()
        self._c(self.L.fe_hip_get_tuning(self.ctx, key.encode(), C.byref(v)), "fe_hip_get_tuning")
        return v.value

    def update_coeff(self, form, index, data):
        """Replace a tabulated coefficient (FN_ELEM / FN_QP) of form `form`, or
        the constant c0 of an analytic one (a single value)."""
        d = np.ascontiguousarray(np.atleast_1d(data), np.float64)
        self._c(self.L.fe_hip_update_coeff(self.ctx, form, index, N.ptr(d, C.c_double), d.size),
                "fe_hip_update_coeff")

    def norm(self, type_, field, degree, exact=(), noalloc=False):
        """feNorm::compute on the device: norm `type_` (NORM_*) of field
        `field` with the degree-`degree` rule; exact: the Fn callbacks the
        type takes (see fe_hip_norm).  noalloc: through fe_hip_norm_noalloc
        (no device allocation; the library's static scratch)."""
        arr = (N.FeHipFn * max(1, len(exact)))()
        keep = []
        for i, fn in enumerate(exact):
            arr[i].kind, arr[i].c0, arr[i].c1 = fn.kind, fn.c0, fn.c1
            if fn.data is not None:
                d = np.ascontiguousarray(fn.data, np.float64)
                keep.append(d)
                arr[i].data = N.ptr(d, C.c_double)
                arr[i].count = d.size
        out = C.c_double()
        fn = self.L.fe_hip_norm_noalloc if noalloc else self.L.fe_hip_norm
        self._c(fn(self.ctx, type_, field, degree, len(exact), arr, C.byref(out)),
                "fe_hip_norm_noalloc" if noalloc else "fe_hip_norm")
        return out.value

    # solution history of the BDF schemes (feSolutionContainer) ---------------
    def set_bdf(self, cn):
        """setC0 + computeBDFCoefficients: u_dot = sum_j cn[j] sol[j]."""
        d = np.ascontiguousarray(cn, np.float64)
        self._c(self.L.fe_hip_set_bdf(self.ctx, d.size, N.ptr(d, C.c_double)), "fe_hip_set_bdf")

    def rotate_solutions(self):
        """feSolutionContainer::rotate: sol[j] = sol[j - 1], j = 4..1."""
        self._c(self.L.fe_hip_rotate_solutions(self.ctx), "fe_hip_rotate_solutions")

    def set_past_solution(self, j, u):
        u = np.ascontiguousarray(u, np.float64)
        assert u.size == self.nDofs
        self._c(self.L.fe_hip_set_past_solution(self.ctx, j, N.ptr(u, C.c_double)), "fe_hip_set_past_solution")

    def get_past_solution(self, j):
        u = np.zeros(self.nDofs)
        self._c(self.L.fe_hip_get_past_solution(self.ctx, j, N.ptr(u, C.c_double)), "fe_hip_get_past_solution")
        return u

    def set_essential(self, values):
        """The essential part of the current solution (nDofs - nInc values)."""
        v = np.ascontiguousarray(values, np.float64)
        assert v.size == self.nDofs - self.n
        if v.size:
            self._c(self.L.fe_hip_set_essential(self.ctx, N.ptr(v, C.c_double)), "fe_hip_set_essential")

    # aggregation multigrid (FE_PC_AMG) ---------------------------------------
    def amg_levels(self):
        """(rows, nnz) per level of the AMG hierarchy (empty before the first AMG solve)."""
        cap = 32
        lv = C.c_int()
        rows = np.zeros(cap, np.int32)
        nnz = np.zeros(cap, np.int64)
        self._c(self.L.fe_hip_amg_levels(self.ctx, C.byref(lv), N.ptr(rows, C.c_int), N.ptr(nnz, C.c_int64), cap),
                "fe_hip_amg_levels")
        k = min(lv.value, cap)
        return rows[:k].tolist(), nnz[:k].tolist()

    def amg_apply(self, r):
        """One V-cycle z = M^-1 r with the current values (FP64 vectors)."""
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        self._c(self.L.fe_hip_amg_apply(self.ctx, N.ptr(r, C.c_double), N.ptr(z, C.c_double)), "fe_hip_amg_apply")
        return z

    # learned preconditioner (FE_PC_LEARNED) ---------------------------------
    def mpnn_opts(self, **kw):
        o = N.FeHipMpnnOpts()
        self.L.fe_hip_mpnn_default_opts(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def mpnn_init(self, seed):
        self._c(self.L.fe_hip_mpnn_init(self.ctx, seed), "fe_hip_mpnn_init")

    def mpnn_set_params(self, theta):
        t = np.ascontiguousarray(theta, np.float32)
        self._c(self.L.fe_hip_mpnn_set_params(self.ctx, N.ptr(t, C.c_float), t.size), "fe_hip_mpnn_set_params")

    def mpnn_get_params(self):
        t = np.zeros(self.L.fe_hip_mpnn_param_count(), np.float32)
        self._c(self.L.fe_hip_mpnn_get_params(self.ctx, N.ptr(t, C.c_float), t.size), "fe_hip_mpnn_get_params")
        return t

    def mpnn_loss_grad(self, want_grad=True):
        loss = C.c_double()
        g = np.zeros(self.L.fe_hip_mpnn_param_count(), np.float32) if want_grad else None
        self._c(self.L.fe_hip_mpnn_loss_grad(self.ctx, C.byref(loss), N.ptr(g, C.c_float) if g is not None else None),
                "fe_hip_mpnn_loss_grad")
        return loss.value, g

    def mpnn_train_step(self, **kw):
        o = self.mpnn_opts(**kw)
        loss = C.c_double()
        self._c(self.L.fe_hip_mpnn_train_step(self.ctx, C.byref(o), C.byref(loss)), "fe_hip_mpnn_train_step")
        return loss.value

    def mpnn_tensor(self, name):
        sizes = {"X": self.n * 16, "p": self.n * 8, "q": self.n * 8, "heads": self.n * 5, "B": self.n,
                 "d": self.n, "g": self.nnz, "w0": self.nnz, "w": self.nnz}
        cnt = sizes.get(name, self.n * 32)
        out = np.zeros(cnt, np.float32)
        self._c(self.L.fe_hip_mpnn_get_tensor(self.ctx, name.encode(), N.ptr(out, C.c_float), cnt),
                "fe_hip_mpnn_get_tensor")
        return out

    def close(self):
        if self.ctx:
            self.L.fe_hip_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def essential_values(num: Numbering, fns_by_field):
    """Initial solution vector: unknowns 0, essential DOFs from the boundary
    callbacks evaluated at the DOF locations (feSolution.cpp:246-377).
    fns_by_field[f] is a callable (x, y, z, comp) -> value, or None for 0."""
    u = np.zeros(num.nDofs)
    ess = np.arange(num.nInc, num.nDofs)
    xyz = num.dof_xyz.reshape(-1, 3)
    for f, fn in enumerate(fns_by_field):
        if fn is None:
            continue
        sel = ess[num.dof_field[ess] == f]
        if sel.size:
            u[sel] = fn(xyz[sel, 0], xyz[sel, 1], xyz[sel, 2], num.dof_comp[sel])
    return u
