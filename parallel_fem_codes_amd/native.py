"""ctypes binding of libfemgpu.so (the C ABI declared in include/femgpu.h).

The library is built in-tree (``parallel_fem_codes_amd/libfemgpu.so``) by ``build()`` or
``make -C parallel_fem_codes_amd/csrc``. There is no fallback: if the library is missing,
``lib()`` raises, and every failing call raises ``FemGpuError`` with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(PKG, "libfemgpu.so")
CSRC = os.path.join(PKG, "csrc")

OK, ERR_ARG, ERR_HIP, ERR_ZERO_PIVOT, ERR_NOMEM, ERR_STATE, ERR_TIMEOUT = 0, -1, -2, -3, -4, -5, -6


class FemGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"femgpu error {code}: {msg}")
        self.code = code


class TransportProblem(C.Structure):
    """femgpu_transport_problem."""
    _fields_ = [("k", C.c_double), ("gamma", C.c_double), ("f0", C.c_double),
                ("bx", C.c_double * 3), ("by", C.c_double * 3), ("shock", C.c_int), ("pad", C.c_int)]


class System(C.Structure):
    """femgpu_system."""
    _fields_ = [("csr", C.c_void_p), ("ebe", C.c_void_p), ("ilu", C.c_void_p), ("bdiag", C.c_void_p)]


class SolveInfo(C.Structure):
    """femgpu_solve_info."""
    _fields_ = [("iterations", C.c_int), ("cycles", C.c_int), ("inner_iterations", C.c_int),
                ("status", C.c_int), ("rho", C.c_double), ("normb", C.c_double),
                ("true_rel_residual", C.c_double), ("prec_rel_residual", C.c_double)]


class EbeInfo(C.Structure):
    """femgpu_ebe_info."""
    _fields_ = [("groups", C.c_int), ("max_incidences", C.c_int), ("column_bytes", C.c_int), ("pad", C.c_int),
                ("bytes_f64", C.c_longlong), ("bytes_f32", C.c_longlong)]


class EulerProblem(C.Structure):
    """femgpu_euler_problem."""
    _fields_ = [("gamma", C.c_double), ("Mach", C.c_double), ("alpha", C.c_double), ("dt", C.c_double),
                ("tolerance", C.c_double), ("dimensionless", C.c_int), ("shock", C.c_int)]


class SweepInfo(C.Structure):
    """femgpu_sweep_info."""
    _fields_ = [(f, C.c_int) for f in ("kernel", "depth_lower", "depth_upper", "steps_lower", "steps_upper",
                                        "ring", "entries_per_row", "chunks_lower", "rows_per_chunk", "lds_limit",
                                        "factor_kernel", "jacobi_column_bytes", "sell_column_bytes", "long_slots",
                                        "factor_compact")]


class MixedOpts(C.Structure):
    """femgpu_mixed_opts."""
    _fields_ = [("inner_kmax", C.c_int), ("inner_cycles", C.c_int), ("inner_tol", C.c_double),
                ("max_refinements", C.c_int), ("sweep_mode", C.c_int), ("factor_sweeps", C.c_int),
                ("orthogonalization", C.c_int), ("graphs", C.c_int), ("pad", C.c_int)]


class TuneInfo(C.Structure):
    """femgpu_tune_info."""
    _fields_ = [("lanes", C.c_int), ("pad", C.c_int), ("us", C.c_double * 5), ("us_sell", C.c_double)]


class BatchInfo(C.Structure):
    """femgpu_batch_info."""
    _fields_ = [("refinements", C.c_int), ("inner_iterations", C.c_int), ("status", C.c_int), ("pad", C.c_int),
                ("normF", C.c_double), ("true_rel_residual", C.c_double)]


class ServerOpts(C.Structure):
    """femgpu_server_opts."""
    _fields_ = [("hbm_budget", C.c_size_t), ("executors", C.c_int), ("prep_threads", C.c_int),
                ("small_rows", C.c_int), ("batch_rows", C.c_int), ("max_batch", C.c_int), ("kmax", C.c_int),
                ("small_solver", C.c_int), ("wg_threads", C.c_int), ("tol", C.c_double), ("mixed", MixedOpts)]


class RequestInfo(C.Structure):
    """femgpu_request_info."""
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("batch_id", C.c_int), ("batch_size", C.c_int),
                ("iterations", C.c_int), ("refinements", C.c_int), ("status", C.c_int), ("wg_threads", C.c_int),
                ("true_rel_residual", C.c_double), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_end", C.c_double), ("prep_s", C.c_double), ("setup_s", C.c_double), ("solve_s", C.c_double)]


class DriverOpts(C.Structure):
    """femgpu_driver_opts."""
    _fields_ = [("solver", C.c_int), ("kmax", C.c_int), ("lmax", C.c_int), ("tol", C.c_double),
                ("mixed", MixedOpts), ("nonlinear_tol", C.c_double), ("nonlinear_maxiter", C.c_int),
                ("stop_by_norm", C.c_int)]


class DriverInfo(C.Structure):
    """femgpu_driver_info."""
    _fields_ = [("iterations", C.c_int), ("linear_iterations", C.c_int), ("status", C.c_int), ("pad", C.c_int),
                ("normdiff", C.c_double), ("normu", C.c_double), ("norm_a", C.c_double), ("norm_Da", C.c_double),
                ("iters", C.c_int * 64)]


class MeshFileInfo(C.Structure):
    """femgpu_mesh_file_info."""
    _fields_ = [("nnodes", C.c_int), ("nel", C.c_int), ("ndof", C.c_int), ("format", C.c_int)]


_LIB = None
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
vp = C.c_void_p
I, D = C.c_int, C.c_double

# name: (restype, argtypes)
_SIGS = {
    "femgpu_last_error": (C.c_char_p, []),
    "femgpu_version": (C.c_char_p, []),
    "femgpu_number_unknowns": (I, [I, I, i32p, i32p]),
    "femgpu_fill_lm": (I, [I, I, I, i32p, i32p, i32p, i32p]),
    "femgpu_csr_nnz": (I, [I, I, I, i32p]),
    "femgpu_csr_init": (I, [I, I, I, i32p, I, I, I, i32p, i32p, i32p, i32p, i32p, i32p]),
    "femgpu_symrcm": (I, [I, i32p, i32p, i32p]),
    "femgpu_fiedler": (I, [I, i32p, i32p, I, f64p]),
    "femgpu_spectral": (I, [I, i32p, i32p, i32p]),
    "femgpu_ebe_permutation": (I, [I, I, i32p, i32p, i32p, I, i32p, i32p]),
    "femgpu_ebe_colour": (I, [I, I, I, i32p, i32p]),
    "femgpu_ilu_levels": (I, [I, i32p, i32p, I, i32p]),
    "femgpu_create": (I, [I, vp, C.POINTER(vp)]),
    "femgpu_destroy": (I, [vp]),
    "femgpu_sync": (I, [vp]),
    "femgpu_csr_create": (I, [vp, I, I, i32p, i32p, C.POINTER(vp)]),
    "femgpu_csr_create_ex": (I, [vp, I, I, i32p, i32p, I, vp]),
    "femgpu_csr_rebuild": (I, [vp, I, I, i32p, i32p]),
    "femgpu_csr_destroy": (I, [vp]),
    "femgpu_csr_set_values": (I, [vp, vp, I]),
    "femgpu_csr_values_ptr": (I, [vp, C.POINTER(vp)]),
    "femgpu_csr_values_changed": (I, [vp]),
    "femgpu_csr_sweep_info": (I, [vp, C.POINTER(SweepInfo)]),
    "femgpu_csr_set_sweep_kernel": (I, [vp, I]),
    "femgpu_csr_set_factor_kernel": (I, [vp, I]),
    "femgpu_csrmv": (I, [vp, vp, vp]),
    "femgpu_csr_autotune": (I, [vp, C.POINTER(TuneInfo)]),
    "femgpu_csr_set_lanes": (I, [vp, I]),
    "femgpu_csrmv_f32": (I, [vp, vp, vp]),
    "femgpu_ilu0_factor": (I, [vp]),
    "femgpu_ilu0_factor_host": (I, [vp]),
    "femgpu_ilu0_apply": (I, [vp, vp, vp]),
    "femgpu_ilu0_factor_parallel": (I, [vp, I]),
    "femgpu_ilu0_set_apply": (I, [vp, I]),
    "femgpu_ilu0_apply_f32": (I, [vp, vp, vp]),
    "femgpu_ilu0_get": (I, [vp, f64p, f64p]),
    "femgpu_ilu0_factor_parallel_f32": (I, [vp, I]),
    "femgpu_ilu0_factor_to_f32": (I, [vp]),
    "femgpu_ilu0_apply_jacobi_f32": (I, [vp, I, vp, vp]),
    "femgpu_ilu0_jacobi_iteration_f32": (I, [vp, I, vp, vp, vp]),
    "femgpu_ilu0_get_f32": (I, [vp, f32p, f32p]),
    "femgpu_ebe_create": (I, [vp, I, I, I, i32p, C.POINTER(vp)]),
    "femgpu_ebe_destroy": (I, [vp]),
    "femgpu_ebe_set_values": (I, [vp, vp, I]),
    "femgpu_ebe_get_info": (I, [vp, vp]),
    "femgpu_ebemv": (I, [vp, vp, vp]),
    "femgpu_ebemv_f32": (I, [vp, vp, vp]),
    "femgpu_ddot": (I, [vp, I, vp, vp, C.POINTER(D)]),
    "femgpu_ddot_noalloc": (I, [vp, I, vp, vp, vp]),
    "femgpu_daxpy": (I, [vp, I, D, vp, vp]),
    "femgpu_dscal": (I, [vp, I, D, vp]),
    "femgpu_dcopy": (I, [vp, I, vp, vp]),
    "femgpu_dzero": (I, [vp, I, vp]),
    "femgpu_pgmres": (I, [vp, C.POINTER(System), I, I, D, vp, vp, C.POINTER(SolveInfo)]),
    "femgpu_gmres_mixed": (I, [vp, C.POINTER(System), C.POINTER(MixedOpts), D, vp, vp,
                              C.POINTER(SolveInfo)]),
    "femgpu_gmres_mixed_batched": (I, [vp, C.POINTER(System), I, i32p, C.POINTER(MixedOpts), D, vp, vp, vp]),
    "femgpu_gmres_mixed_wg": (I, [vp, C.POINTER(System), I, i32p, C.POINTER(MixedOpts), D, I, vp, vp, vp]),
    "femgpu_pcg": (I, [vp, C.POINTER(System), I, D, vp, vp, C.POINTER(SolveInfo)]),
    "femgpu_blockdiag_create": (I, [vp, I, I, I, I, i32p, i32p, i32p, vp]),
    "femgpu_blockdiag_destroy": (I, [vp]),
    "femgpu_blockdiag_setup": (I, [vp, vp]),
    "femgpu_blockdiag_apply": (I, [vp, vp, vp]),
    "femgpu_block_scaling": (I, [vp, vp, vp]),
    "femgpu_structured_mesh": (I, [I, I, D, D, D, D, I, f64p, f64p, i32p]),
    "femgpu_structured_marks": (I, [C.c_char_p, I, f64p, f64p, i32p, vp]),
    "femgpu_mesh_file_read": (I, [C.c_char_p, I, I, C.POINTER(MeshFileInfo), vp, vp, vp, vp, vp, vp]),
    "femgpu_delaunay": (I, [I, vp, vp, I, C.POINTER(I), vp, C.POINTER(I)]),
    "femgpu_mesh_file_write": (I, [C.c_char_p, I, I, I, f64p, f64p, i32p, i32p]),
    "femgpu_server_create": (I, [I, C.POINTER(ServerOpts), vp]),
    "femgpu_server_submit": (I, [vp, I, I, f64p, f64p, i32p, f64p, i32p, C.POINTER(TransportProblem),
                                 C.POINTER(C.c_longlong)]),
    "femgpu_server_wait": (I, [vp, C.c_longlong, D, vp, C.POINTER(RequestInfo)]),
    "femgpu_server_clock": (D, [vp]),
    "femgpu_server_batch_tuning": (I, [vp, C.POINTER(I)]),
    "femgpu_server_wg_tuning": (I, [vp, C.POINTER(I)]),
    "femgpu_server_destroy": (I, [vp]),
    "femgpu_st_process": (I, [vp, vp, C.POINTER(TransportProblem), C.POINTER(DriverOpts), vp,
                              C.POINTER(DriverInfo)]),
    "femgpu_ns_process": (I, [vp, vp, D, vp, vp, C.POINTER(DriverOpts), vp, C.POINTER(DriverInfo)]),
    "femgpu_tt_predictor_old_step": (I, [vp, vp, C.POINTER(TransportProblem), D, D, C.POINTER(DriverOpts), vp, vp,
                                         C.POINTER(DriverInfo)]),
    "femgpu_mesh_create": (I, [vp, I, I, I, f64p, f64p, i32p, i32p, f64p, i32p, vp, vp, I,
                              C.POINTER(vp)]),
    "femgpu_mesh_destroy": (I, [vp]),
    "femgpu_mesh_rebuild": (I, [vp, I, I, I, f64p, f64p, i32p, i32p, f64p, i32p, vp, vp, I]),
    "femgpu_ns_mesh_create": (I, [vp, I, I, I, f64p, f64p, i32p, i32p, f64p, i32p, i32p, vp, I,
                                  C.POINTER(vp)]),
    "femgpu_assemble_ns": (I, [vp, D, vp, vp, vp, vp, vp, vp]),
    "femgpu_euler_mesh_create": (I, [vp, I, I, I, f64p, f64p, i32p, i32p, f64p, i32p, i32p, vp, I,
                                     C.POINTER(vp)]),
    "femgpu_assemble_euler": (I, [vp, C.POINTER(EulerProblem), vp, vp, vp, vp, vp]),
    "femgpu_eu_predictor_old_step": (I, [vp, vp, vp, C.POINTER(EulerProblem), C.POINTER(DriverOpts), vp, vp,
                                         C.POINTER(DriverInfo)]),
    "femgpu_assemble_transport": (I, [vp, C.POINTER(TransportProblem), vp, vp, vp, vp]),
    "femgpu_assemble_transport_transient": (I, [vp, C.POINTER(TransportProblem), D, D, vp, vp,
                                                vp, vp, vp]),
}


def declared_symbols():
    """Every entry point of include/femgpu.h, parsed from the header itself."""
    import re
    hdr = open(os.path.join(os.path.dirname(PKG), "include", "femgpu.h")).read()
    return sorted(set(re.findall(r"\b(femgpu_[a-z0-9_]+)\s*\(", hdr)))


def build(force: bool = False):
    """Compile libfemgpu.so for gfx950 in-tree with hipcc (cross-compiles without a GPU)."""
    args = ["make", "-C", CSRC, "-j8"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return SO_PATH


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(SO_PATH):
            raise FemGpuError(ERR_STATE, f"{SO_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(SO_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc < 0:
        raise FemGpuError(rc, lib().femgpu_last_error().decode())
    return rc
