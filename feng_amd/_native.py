"""ctypes bindings of the two C ABIs (include/fe_mesh.h, include/fe_hip.h).

The libraries are loaded from feng_amd/lib/ (built in-tree by
feng_amd/build.py).  A missing library raises immediately: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)
c_void_pp = C.POINTER(C.c_void_p)


class FeHipFn(C.Structure):
    _fields_ = [("kind", C.c_int), ("c0", C.c_double), ("c1", C.c_double),
                ("data", c_double_p), ("count", C.c_int64)]


class FeHipSolverOpts(C.Structure):
    _fields_ = [("ksp", C.c_int), ("pc", C.c_int), ("precision", C.c_int), ("restart", C.c_int),
                ("max_iter", C.c_int), ("rtol", C.c_double), ("atol", C.c_double),
                ("dtol", C.c_double), ("inner_rtol", C.c_double), ("check_every", C.c_int)]


class FeHipSolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_iterations", C.c_int), ("converged", C.c_int),
                ("reason", C.c_int), ("relres_fp64", C.c_double), ("norm_dx_inf", C.c_double),
                ("norm_rhs_inf", C.c_double), ("norm_axb_inf", C.c_double),
                ("ms_solve", C.c_double), ("ms_setup", C.c_double)]


class FeHipMpnnOpts(C.Structure):
    _fields_ = [("lr", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double)]


class FeHipBatchStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("relres_fp64", C.c_double)]


# name -> (restype, argtypes)
HOST_API = {
    "fe_mesh_square": (C.c_int, [C.c_int, C.c_double, C.c_uint64, c_void_pp]),
    "fe_mesh_cube": (C.c_int, [C.c_int, c_void_pp]),
    "fe_mesh_from_arrays": (C.c_int, [C.c_int, C.c_int, c_double_p, C.c_int, C.c_int, c_int_p,
                                      C.c_int, C.c_int, c_int_p, c_void_pp]),
    "fe_mesh_write_msh": (C.c_int, [C.c_void_p, C.c_char_p]),
    "fe_mesh_read_msh": (C.c_int, [C.c_char_p, c_void_pp]),
    "fe_mesh_destroy": (C.c_int, [C.c_void_p]),
    "fe_mesh_check_jacobians": (C.c_int, [C.c_void_p, c_int_p, c_double_p]),
    "fe_mesh_info": (C.c_int, [C.c_void_p] + [c_int_p] * 7),
    "fe_mesh_get": (C.c_int, [C.c_void_p, c_double_p, c_int_p, c_int_p, c_int_p, c_int_p]),
    "fe_numbering_create": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, c_void_pp]),
    "fe_numbering_create_groups": (C.c_int, [C.c_void_p, C.c_int, c_int_p, C.POINTER(C.c_uint), c_void_pp]),
    "fe_mesh_boundary_groups": (C.c_int, [C.c_void_p, c_int_p, C.c_char_p, C.c_int, c_int_p]),
    "fe_numbering_periodic": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                        c_int_p, c_int_p, C.c_int, c_int_p]),
    "fe_numbering_destroy": (C.c_int, [C.c_void_p]),
    "fe_numbering_info": (C.c_int, [C.c_void_p, c_int_p, c_int_p]),
    "fe_numbering_adr": (C.c_int, [C.c_void_p, C.c_int, c_int_p]),
    "fe_numbering_dof_info": (C.c_int, [C.c_void_p, c_double_p, c_int_p, c_int_p]),
    "fe_color_elements": (C.c_int, [C.c_void_p, c_int_p, c_int_p]),
    "fe_coeff_blocks": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_uint64,
                                  c_double_p]),
    "fe_coeff_grf_cube": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_uint64,
                                    c_double_p]),
}

HIP_API = {
    "fe_hip_ctx_create": (C.c_int, [C.c_int, c_void_pp]),
    "fe_hip_ctx_destroy": (C.c_int, [C.c_void_p]),
    "fe_hip_last_error": (C.c_char_p, [C.c_void_p]),
    "fe_hip_upload_mesh": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p, C.c_int, C.c_int,
                                     c_int_p]),
    "fe_hip_set_coloring": (C.c_int, [C.c_void_p, c_int_p]),
    "fe_hip_add_field": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p]),
    "fe_hip_set_sizes": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "fe_hip_add_form": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_int_p, C.c_int, C.c_int,
                                  C.POINTER(FeHipFn), c_int_p]),
    "fe_hip_build_csr": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "fe_hip_get_csr": (C.c_int, [C.c_void_p, c_int_p, c_int_p]),
    "fe_hip_set_solution": (C.c_int, [C.c_void_p, c_double_p]),
    "fe_hip_get_solution": (C.c_int, [C.c_void_p, c_double_p]),
    "fe_hip_set_to_zero": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "fe_hip_assemble": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "fe_hip_rhs_max_norm": (C.c_int, [C.c_void_p, c_double_p]),
    "fe_hip_get_system": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "fe_hip_set_system": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "fe_hip_default_opts": (None, [C.POINTER(FeHipSolverOpts)]),
    "fe_hip_solve": (C.c_int, [C.c_void_p, C.POINTER(FeHipSolverOpts), c_double_p,
                               C.POINTER(FeHipSolveStats)]),
    "fe_hip_correct_solution": (C.c_int, [C.c_void_p]),
    "fe_hip_kernel_stats": (C.c_int, [C.c_void_p, C.c_int, c_double_p, C.POINTER(C.c_int64)]),
    "fe_hip_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "fe_hip_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "fe_hip_pc_info": (C.c_int, [C.c_void_p, c_int_p, c_int_p]),
    "fe_hip_get_factor": (C.c_int, [C.c_void_p, c_double_p]),
    "fe_hip_get_tuning": (C.c_int, [C.c_void_p, C.c_char_p, c_int_p]),
    "fe_hip_solve_batched": (C.c_int, [C.c_void_p, C.c_int, c_int_p, C.POINTER(FeHipSolverOpts),
                                       C.POINTER(FeHipBatchStats), c_double_p]),
    "fe_hip_autotune_batched": (C.c_int, [C.c_void_p, C.c_int, c_int_p, C.c_int, c_int_p, c_int_p, c_double_p]),
    "fe_hip_autotune_assembly": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, c_int_p, c_double_p]),
    "fe_hip_update_coeff": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p, C.c_int64]),
    "fe_hip_set_bdf": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "fe_hip_rotate_solutions": (C.c_int, [C.c_void_p]),
    "fe_hip_set_past_solution": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "fe_hip_get_past_solution": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "fe_hip_set_essential": (C.c_int, [C.c_void_p, c_double_p]),
    "fe_hip_get_system32": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "fe_hip_norm": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, c_double_p]),
    "fe_hip_norm_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, c_double_p]),
    "fe_hip_set_constrained_dofs": (C.c_int, [C.c_void_p, C.c_int, c_int_p]),
    "fe_hip_constrain_essential_components": (C.c_int, [C.c_void_p]),
    "fe_hip_set_periodic_dofs": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p]),
    "fe_hip_apply_periodicity": (C.c_int, [C.c_void_p]),
    "fe_hip_amg_levels": (C.c_int, [C.c_void_p, c_int_p, c_int_p, C.POINTER(C.c_int64), C.c_int]),
    "fe_hip_amg_apply": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "fe_hip_mpnn_param_count": (C.c_int, []),
    "fe_hip_mpnn_default_opts": (None, [C.POINTER(FeHipMpnnOpts)]),
    "fe_hip_mpnn_init": (C.c_int, [C.c_void_p, C.c_uint64]),
    "fe_hip_mpnn_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int]),
    "fe_hip_mpnn_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int]),
    "fe_hip_mpnn_loss_grad": (C.c_int, [C.c_void_p, c_double_p, C.POINTER(C.c_float)]),
    "fe_hip_mpnn_train_step": (C.c_int, [C.c_void_p, C.POINTER(FeHipMpnnOpts), c_double_p]),
    "fe_hip_mpnn_get_tensor": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_float), C.c_int64]),
}

_host = None
_hip = None


def _bind(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def host():
    """libfe_host.so (meshes, numbering, coloring)."""
    global _host
    if _host is None:
        path = LIB_DIR / "libfe_host.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run `python -m feng_amd.build` first")
        _host = _bind(C.CDLL(str(path), mode=C.RTLD_GLOBAL), HOST_API)
    return _host


def hip():
    """libfe_hip.so (device path).  Raises when it is not built."""
    global _hip
    if _hip is None:
        path = LIB_DIR / "libfe_hip.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run `python -m feng_amd.build` first")
        _hip = _bind(C.CDLL(str(path)), HIP_API)
    return _hip


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
