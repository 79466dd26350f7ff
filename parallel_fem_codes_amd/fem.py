"""Host-side mirror of the reference hot-path interface, over the libfemgpu C ABI.

Names follow the reference (``Fill_LM``, ``csr_Initialization``, ``Permutation_of_LM``,
``pgmres``, ``ILUp_precond_setup`` ...) with the argument meaning of the one-rank case. Device
vectors are torch tensors on the GPU (PyTorch is used only to own device memory and streams);
host arrays are numpy. Integer preprocessing runs in the library's C++ host code; every numerical
operation runs in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as N


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("device vectors must be contiguous GPU tensors")
    return t.data_ptr()


# ------------------------------------------------------------------------------ mesh generation

def structured_mesh(mx, my, box=(0.0, 0.0, 1.0, 1.0), right=True):
    """mx x my squares cut on one diagonal, numbered like MG/Structured_Mesh/meshGenerator.c
    (nodes row by row, squares row by row, two triangles each; "right" cut from the lower-left to
    the upper-right corner, "left" cut the other diagonal). Returns x, y, ien (int32, nel*3)."""
    x0, y0, x1, y1 = box
    nn = (mx + 1) * (my + 1)
    x, y, ien = np.empty(nn), np.empty(nn), np.empty(6 * mx * my, np.int32)
    N.check(N.lib().femgpu_structured_mesh(int(mx), int(my), float(x0), float(y0), float(x1), float(y1),
                                           1 if right else 0, x, y, ien))
    return x, y, ien


def structured_marks(problem, x, y):
    """setmark_according_boundary of the generator: (NDOF, mark per node, Type per dof)."""
    x, y = np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64)
    mark = np.empty(len(x), np.int32)
    ndof = N.check(N.lib().femgpu_structured_marks(problem.encode(), len(x), x, y, mark, None))
    t = np.empty(len(x) * ndof, np.int32)
    N.check(N.lib().femgpu_structured_marks(problem.encode(), len(x), x, y, mark, t.ctypes.data_as(C.c_void_p)))
    return ndof, mark, t


def delaunay(x, y):
    """Delaunay triangulation of the points (x, y) by the library (Bowyer-Watson): (ntri, 3) int32
    counter-clockwise triangles and the number of repeated points skipped."""
    x, y = _f64(x), _f64(y)
    n = len(x)
    cap = max(2 * n, 8)
    tri = np.empty(3 * cap, np.int32)
    nt, dup = C.c_int(), C.c_int()
    N.check(N.lib().femgpu_delaunay(n, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), cap,
                                    C.byref(nt), tri.ctypes.data_as(C.c_void_p), C.byref(dup)))
    return tri[:3 * nt.value].reshape(-1, 3).copy(), dup.value


def read_mesh(path, ndof=1, format=0):
    """A mesh file: format 0 = the generator's <PROBLEM>_<nn>_<ne>.dat, 1 = one rank's RANK_* file.
    Returns dict(x, y, type, invP, ien, etype)."""
    info = N.MeshFileInfo()
    p = str(path).encode()
    N.check(N.lib().femgpu_mesh_file_read(p, format, ndof, C.byref(info), None, None, None, None, None, None))
    x, y = np.empty(info.nnodes), np.empty(info.nnodes)
    t, inv = np.empty(info.nnodes * ndof, np.int32), np.empty(info.nnodes * ndof, np.int32)
    ien, et = np.empty(3 * info.nel, np.int32), np.empty(info.nel, np.int32)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    N.check(N.lib().femgpu_mesh_file_read(p, format, ndof, C.byref(info), vp(x), vp(y), vp(t), vp(inv), vp(ien),
                                          vp(et)))
    return dict(x=x, y=y, type=t, invP=inv, ien=ien, etype=et)


def write_mesh(path, x, y, type_, ien, ndof=1):
    """Writes the generator's format (meshGenerator.c:193-211)."""
    x, y = np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64)
    N.check(N.lib().femgpu_mesh_file_write(str(path).encode(), len(x), len(ien) // 3, ndof, x, y, _i32(type_),
                                           _i32(ien)))


def boundary_types(x, y, box=(0.0, 0.0, 1.0, 1.0), tol=1e-15):
    """Type per node: 0 (Dirichlet) on the boundary of the box, 1 inside (mark 15 / mark 0)."""
    x0, y0, x1, y1 = box
    b = (np.abs(x - x1) <= tol) | (np.abs(x - x0) <= tol) | (np.abs(y - y1) <= tol) | (np.abs(y - y0) <= tol)
    return np.where(b, 0, 1).astype(np.int32)


# ------------------------------------------------------------------------------ integer work

def number_unknowns(type_, ndof=1):
    t = _i32(type_)
    inv = np.empty_like(t)
    neq = N.check(N.lib().femgpu_number_unknowns(len(t) // ndof, ndof, t, inv))
    return inv, neq


def Fill_LM(ien, type_, invP_id, NEQ, ndof=1):
    nel = len(ien) // 3
    lm = np.empty(nel * 3 * ndof, np.int32)
    N.check(N.lib().femgpu_fill_lm(nel, ndof, NEQ, _i32(ien), _i32(type_), _i32(invP_id), lm))
    return lm


@dataclass
class CsrStructure:
    IA: np.ndarray
    JA: np.ndarray
    scheme: np.ndarray
    perm: np.ndarray
    lm: np.ndarray
    invP_id: np.ndarray
    nnz: int


def csr_Initialization(lm, s, neq, reordering="SYMRCM", permute_lm=False, type_=None, invP_id=None):
    """csr_Initialization for one rank: sorted CSR from lm, slot table, reordering and remaps."""
    code = {"NOT": 0, "SYMRCM": 1, "SPECTRAL": 2}[reordering.upper()]
    L = N.lib()
    lm = _i32(lm).copy()
    nel = len(lm) // s
    nnz = N.check(L.femgpu_csr_nnz(nel, s, neq, lm))
    IA = np.empty(neq + 1, np.int32)
    JA = np.empty(max(nnz, 1), np.int32)
    scheme = np.empty(max(nel * s * s, 1), np.int32)
    perm = np.empty(max(neq, 1), np.int32)
    t = _i32(type_) if type_ is not None else np.zeros(1, np.int32)
    inv = _i32(invP_id).copy() if invP_id is not None else np.zeros(1, np.int32)
    nt = len(t) if type_ is not None else 0
    r = N.check(L.femgpu_csr_init(nel, s, neq, lm, code, int(permute_lm), nt, t, inv, IA, JA, scheme, perm))
    assert r == nnz
    return CsrStructure(IA, JA[:nnz], scheme[:nel * s * s], perm[:neq], lm,
                        inv if invP_id is not None else None, nnz)


def symrcm(IA, JA):
    n = len(IA) - 1
    p = np.empty(max(n, 1), np.int32)
    N.check(N.lib().femgpu_symrcm(n, _i32(IA), _i32(JA), p))
    return p[:n]


def spectral(IA, JA):
    """REORDERING_SPECTRAL: perm[new] = old by ascending Fiedler-vector entry."""
    n = len(IA) - 1
    p = np.empty(max(n, 1), np.int32)
    N.check(N.lib().femgpu_spectral(n, _i32(IA), _i32(JA), p))
    return p[:n]


def fiedler(IA, JA, refine_iters=60):
    """The Fiedler vector of the pattern's graph Laplacian (zero mean, unit norm)."""
    n = len(IA) - 1
    f = np.empty(max(n, 1))
    N.check(N.lib().femgpu_fiedler(n, _i32(IA), _i32(JA), int(refine_iters), f))
    return f[:n]


def Permutation_of_LM(ien, type_, invP_id, NEQ, reordering="SYMRCM"):
    code = {"NOT": 0, "SYMRCM": 1, "SPECTRAL": 2}[reordering.upper()]
    Perm, invPerm = np.empty(NEQ + 1, np.int32), np.empty(NEQ + 1, np.int32)
    N.check(N.lib().femgpu_ebe_permutation(len(ien) // 3, NEQ, _i32(ien), _i32(type_), _i32(invP_id),
                                           code, Perm, invPerm))
    return Perm, invPerm


def ebe_colour(lm, s, NEQ):
    lm = _i32(lm)
    col = np.empty(max(len(lm) // s, 1), np.int32)
    nc = N.check(N.lib().femgpu_ebe_colour(len(lm) // s, s, NEQ, lm, col))
    return col[:len(lm) // s], nc


def ilu_levels(IA, JA, lower=True):
    n = len(IA) - 1
    lev = np.empty(max(n, 1), np.int32)
    depth = N.check(N.lib().femgpu_ilu_levels(n, _i32(IA), _i32(JA), int(lower), lev))
    return lev[:n], depth


# ------------------------------------------------------------------------------ device objects

class Device:
    """A libfemgpu context on one GPU, issuing work on the given HIP stream (torch's current
    stream by default, so torch tensors produced before a call are ready for it)."""

    def __init__(self, device=0, stream=None):
        import torch
        torch.cuda.init()
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        h = C.c_void_p()
        N.check(N.lib().femgpu_create(device, C.c_void_p(stream), C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            N.lib().femgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        N.check(N.lib().femgpu_sync(self.h))

    # BLAS (00_CommonFiles/BLAS_Operations) on device vectors
    def ddot(self, x, y):
        r = C.c_double()
        N.check(N.lib().femgpu_ddot(self.h, x.numel(), _ptr(x), _ptr(y), C.byref(r)))
        return r.value

    def daxpy(self, a, x, y):
        N.check(N.lib().femgpu_daxpy(self.h, x.numel(), float(a), _ptr(x), _ptr(y)))

    def dscal(self, a, x):
        N.check(N.lib().femgpu_dscal(self.h, x.numel(), float(a), _ptr(x)))

    def dcopy(self, x, y):
        N.check(N.lib().femgpu_dcopy(self.h, x.numel(), _ptr(x), _ptr(y)))

    def dzero(self, x):
        N.check(N.lib().femgpu_dzero(self.h, x.numel(), _ptr(x)))


class CsrMatrix:
    """Device CSR operator (csrmv) with its ILU(0) preconditioner (ILUp_precond[_setup])."""

    def __init__(self, dev: Device, IA, JA):
        self.dev, self.n, self.nnz = dev, len(IA) - 1, len(JA)
        h = C.c_void_p()
        N.check(N.lib().femgpu_csr_create(dev.h, self.n, self.nnz, _i32(IA), _i32(JA), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().femgpu_csr_destroy(self.h)
            self.h = None

    def set_values(self, AA):
        if hasattr(AA, "is_cuda"):
            N.check(N.lib().femgpu_csr_set_values(self.h, C.c_void_p(_ptr(AA)), 1))
        else:
            a = _f64(AA)
            N.check(N.lib().femgpu_csr_set_values(self.h, a.ctypes.data_as(C.c_void_p), 0))
            self.dev.sync()

    def values_ptr(self):
        p = C.c_void_p()
        N.check(N.lib().femgpu_csr_values_ptr(self.h, C.byref(p)))
        return p.value

    def values_changed(self):
        N.check(N.lib().femgpu_csr_values_changed(self.h))

    def rebuild(self, IA, JA):
        """Re-plans the matrix for a new structure, reusing its device buffers where large enough."""
        self.n, self.nnz = len(IA) - 1, len(JA)
        N.check(N.lib().femgpu_csr_rebuild(self.h, self.n, self.nnz, _i32(IA), _i32(JA)))

    def autotune(self):
        """Times the FP64 product kernels (CSR at 1, 2, 4, 8 and 16 lanes per row; sliced ELL, lanes
        0, where the structure allows it) and keeps the fastest; returns (lanes chosen, {lanes:
        microseconds per launch})."""
        t = N.TuneInfo()
        N.check(N.lib().femgpu_csr_autotune(self.h, C.byref(t)))
        us = {L: t.us[k] for k, L in enumerate((1, 2, 4, 8, 16))}
        if t.us_sell > 0:
            us[0] = t.us_sell
        return t.lanes, us

    def set_lanes(self, lanes):
        """Lanes per row of the FP64 CSR kernels (1, 2, 4, 8, 16) or 0 for FP64 sliced ELL."""
        N.check(N.lib().femgpu_csr_set_lanes(self.h, int(lanes)))

    def sweep_info(self):
        si = N.SweepInfo()
        N.check(N.lib().femgpu_csr_sweep_info(self.h, C.byref(si)))
        return {f: getattr(si, f) for f, _ in N.SweepInfo._fields_}

    def set_factor_kernel(self, kind):
        """"stream" (single-workgroup streamed factorisation, the default where its plan exists) or
        "lists" (level-synchronous update lists)."""
        N.check(N.lib().femgpu_csr_set_factor_kernel(self.h, {"stream": 0, "lists": 1}[kind]))

    def set_sweep_kernel(self, kind):
        """"stream" (single-workgroup streamed sweep, the default where its plan exists) or "packed"."""
        N.check(N.lib().femgpu_csr_set_sweep_kernel(self.h, {"stream": 0, "packed": 1}[kind]))

    def csrmv(self, P, Q):
        fn = N.lib().femgpu_csrmv_f32 if P.dtype.itemsize == 4 else N.lib().femgpu_csrmv
        N.check(fn(self.h, _ptr(P), _ptr(Q)))

    def ilu0_factor(self, on_host=False):
        N.check((N.lib().femgpu_ilu0_factor_host if on_host else N.lib().femgpu_ilu0_factor)(self.h))

    def ilu0_apply(self, p, z):
        fn = N.lib().femgpu_ilu0_apply_f32 if p.dtype.itemsize == 4 else N.lib().femgpu_ilu0_apply
        N.check(fn(self.h, _ptr(p), _ptr(z)))

    def ilu0_factors(self):
        LU, Dinv = np.empty(max(self.nnz, 1)), np.empty(max(self.n, 1))
        N.check(N.lib().femgpu_ilu0_get(self.h, LU, Dinv))
        return LU[:self.nnz], Dinv[:self.n]

    # FP32 preconditioner of the mixed solver
    def ilu0_factor_parallel_f32(self, sweeps):
        N.check(N.lib().femgpu_ilu0_factor_parallel_f32(self.h, int(sweeps)))

    def ilu0_factor_to_f32(self):
        N.check(N.lib().femgpu_ilu0_factor_to_f32(self.h))

    def ilu0_jacobi_iteration_f32(self, upper, rhs, out, inp=None):
        """One Jacobi iteration on one triangle of the FP32 factors (out must not alias rhs or inp)."""
        N.check(N.lib().femgpu_ilu0_jacobi_iteration_f32(self.h, 1 if upper else 0, _ptr(rhs), _ptr(inp), _ptr(out)))

    def ilu0_apply_jacobi_f32(self, iterations, p, z):
        N.check(N.lib().femgpu_ilu0_apply_jacobi_f32(self.h, int(iterations), _ptr(p), _ptr(z)))

    def ilu0_factors_f32(self):
        LU, Dinv = np.empty(max(self.nnz, 1), np.float32), np.empty(max(self.n, 1), np.float32)
        N.check(N.lib().femgpu_ilu0_get_f32(self.h, LU, Dinv))
        return LU[:self.nnz], Dinv[:self.n]

    def ilu0_factor_parallel(self, sweeps):
        """FP64 ILU(0) factors by parallel (fixed-point) sweeps, for ilu0_set_apply(k > 0)."""
        N.check(N.lib().femgpu_ilu0_factor_parallel(self.h, int(sweeps)))

    def ilu0_set_apply(self, jacobi_iterations):
        """FP64 preconditioner apply: 0 exact sweeps (ILUp_precond), k > 0 Jacobi iterations on the
        parallel factors."""
        N.check(N.lib().femgpu_ilu0_set_apply(self.h, int(jacobi_iterations)))

    def ILUp_precond_setup(self, F):
        """ILUp_precond_setup with p = 0: factor the current values, then F := M^-1 F in place."""
        self.ilu0_factor()
        self.ilu0_apply(F, F)


class EbeMatrix:
    """Device element-by-element operator (ebemv / ebemvNDOF3 / ebemvNDOF4)."""

    def __init__(self, dev: Device, lm, s, NEQ):
        self.dev, self.s, self.NEQ = dev, s, NEQ
        self.nel = len(lm) // s
        h = C.c_void_p()
        N.check(N.lib().femgpu_ebe_create(dev.h, self.nel, s, NEQ, _i32(lm), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().femgpu_ebe_destroy(self.h)
            self.h = None

    @property
    def info(self):
        """Layout of the gathered product: groups, max incidences per node, column bytes and the
        bytes one FP64 / FP32 product moves."""
        i = N.EbeInfo()
        N.check(N.lib().femgpu_ebe_get_info(self.h, C.byref(i)))
        return i

    def set_values(self, A):
        if hasattr(A, "is_cuda"):
            N.check(N.lib().femgpu_ebe_set_values(self.h, C.c_void_p(_ptr(A)), 1))
        else:
            a = _f64(A)
            N.check(N.lib().femgpu_ebe_set_values(self.h, a.ctypes.data_as(C.c_void_p), 0))

    def ebemv(self, P, Q):
        """Q = A P; FP64 tensors use the FP64 operator, FP32 tensors its FP32 copy."""
        if P.dtype == Q.dtype and str(P.dtype) == "torch.float32":
            N.check(N.lib().femgpu_ebemv_f32(self.h, _ptr(P), _ptr(Q)))
        else:
            N.check(N.lib().femgpu_ebemv(self.h, _ptr(P), _ptr(Q)))


def make_problem(k=1e-4, gamma=0.0, f0=0.0, bx=(0.5, 0.0, -1.0), by=(-0.5, 1.0, 0.0), shock=0):
    """Problem callbacks of the transport drivers as data (Condutivity, Reaction, Font, Velocity;
    CAU shock capture on/off). The default velocity is the rotating field (-y+0.5, x-0.5)."""
    p = N.TransportProblem()
    p.k, p.gamma, p.f0 = k, gamma, f0
    p.bx[:] = list(bx)
    p.by[:] = list(by)
    p.shock = shock
    return p


class Mesh:
    """Device-resident mesh + gather maps for Build_K_F_SUPG / Build_M_K_R_SUPG."""

    def __init__(self, dev: Device, x, y, type_, invP_id, ubc, ien, neq, lm=None, scheme=None, nnz=0):
        self.dev, self.neq, self.nel = dev, neq, len(ien) // 3
        h = C.c_void_p()
        lm_a = _i32(lm) if lm is not None else None
        sc = _i32(scheme) if scheme is not None else None
        N.check(N.lib().femgpu_mesh_create(
            dev.h, len(x), self.nel, neq, _f64(x), _f64(y), _i32(type_), _i32(invP_id), _f64(ubc), _i32(ien),
            lm_a.ctypes.data_as(C.c_void_p) if lm_a is not None else None,
            sc.ctypes.data_as(C.c_void_p) if sc is not None else None, int(nnz), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().femgpu_mesh_destroy(self.h)
            self.h = None

    def Build_K_F_SUPG(self, pb, u, F, Ke=None, AA_ptr=None):
        N.check(N.lib().femgpu_assemble_transport(self.h, C.byref(pb), _ptr(u), _ptr(Ke), _ptr(F),
                                                  C.c_void_p(AA_ptr) if AA_ptr else None))

    def Build_M_K_R_SUPG(self, pb, alpha, dt, u, du, R, Ae=None, AA_ptr=None):
        N.check(N.lib().femgpu_assemble_transport_transient(self.h, C.byref(pb), float(alpha), float(dt),
                                                            _ptr(u), _ptr(du), _ptr(Ae), _ptr(R),
                                                            C.c_void_p(AA_ptr) if AA_ptr else None))


class NsMesh:
    """Device-resident Navier-Stokes mesh (u, v, p per node) for Build_K_F_SUPG_PSPG."""

    def __init__(self, dev: Device, x, y, type3, invP3, presc3, ien, lm, neq, scheme=None, nnz=0):
        self.dev, self.neq, self.nel = dev, neq, len(ien) // 3
        h = C.c_void_p()
        sc = _i32(scheme) if scheme is not None else None
        N.check(N.lib().femgpu_ns_mesh_create(
            dev.h, len(x), self.nel, neq, _f64(x), _f64(y), _i32(type3), _i32(invP3), _f64(presc3), _i32(ien),
            _i32(lm), sc.ctypes.data_as(C.c_void_p) if sc is not None else None, int(nnz), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().femgpu_mesh_destroy(self.h)
            self.h = None

    def Build_K_F_SUPG_PSPG(self, Re, u, F, fx=None, fy=None, Ke=None, AA_ptr=None):
        """Picard tangent and residual; F has NEQ+1 entries (F[NEQ] = 0)."""
        N.check(N.lib().femgpu_assemble_ns(self.h, float(Re), _ptr(u), _ptr(fx), _ptr(fy), _ptr(Ke), _ptr(F),
                                           C.c_void_p(AA_ptr) if AA_ptr else None))


def make_euler_problem(gamma=1.4, Mach=2.0, alpha=0.5, dt=1e-2, tolerance=1e-8, dimensionless=1, shock=1):
    p = N.EulerProblem()
    p.gamma, p.Mach, p.alpha, p.dt, p.tolerance = gamma, Mach, alpha, dt, tolerance
    p.dimensionless, p.shock = dimensionless, shock
    return p


class EulerMesh(NsMesh):
    """Device-resident Euler mesh (rho, rho u, rho v, rho E per node) for Build_M_K_F_SUPG."""

    def __init__(self, dev: Device, x, y, type4, invP4, presc4, ien, lm, neq, scheme=None, nnz=0):
        self.dev, self.neq, self.nel = dev, neq, len(ien) // 3
        h = C.c_void_p()
        sc = _i32(scheme) if scheme is not None else None
        N.check(N.lib().femgpu_euler_mesh_create(
            dev.h, len(x), self.nel, neq, _f64(x), _f64(y), _i32(type4), _i32(invP4), _f64(presc4), _i32(ien),
            _i32(lm), sc.ctypes.data_as(C.c_void_p) if sc is not None else None, int(nnz), C.byref(h)))
        self.h = h

    def Build_M_K_F_SUPG(self, pb, u, du, R, Ae=None, AA_ptr=None):
        """Ae = Me + alpha*dt*Ke and R = -(Me dU + Ke U); R has NEQ+1 entries (R[NEQ] = 0)."""
        N.check(N.lib().femgpu_assemble_euler(self.h, C.byref(pb), _ptr(u), _ptr(du), _ptr(Ae), _ptr(R),
                                              C.c_void_p(AA_ptr) if AA_ptr else None))


class BlockDiag:
    """Nodal block-diagonal preconditioner (BlockDiag_precond_EBE_setup / BlockDiag_precond) and
    block scaling (Block_scaling_EBE) for element matrices with ndof unknowns per node."""

    def __init__(self, dev: Device, nnodes, ien, ndof, type_, invP, NEQ):
        self.dev, self.ndof, self.NEQ = dev, ndof, NEQ
        h = C.c_void_p()
        N.check(N.lib().femgpu_blockdiag_create(dev.h, int(nnodes), len(ien) // 3, int(ndof), int(NEQ), _i32(ien),
                                                _i32(type_), _i32(invP), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            N.lib().femgpu_blockdiag_destroy(self.h)
            self.h = None

    def setup(self, Ae, F=None):
        """Blocks from the element matrices Ae (device, element order); F := D^-1 F when given."""
        N.check(N.lib().femgpu_blockdiag_setup(self.h, _ptr(Ae)))
        if F is not None:
            self.apply(F, F)

    def apply(self, p, z):
        N.check(N.lib().femgpu_blockdiag_apply(self.h, _ptr(p), _ptr(z)))

    def scale(self, Ae, F=None):
        """Block_scaling_EBE: Ae and F scaled in place."""
        N.check(N.lib().femgpu_block_scaling(self.h, _ptr(Ae), _ptr(F)))


def _system(op, prec):
    S = N.System()
    if isinstance(op, CsrMatrix):
        S.csr = op.h
    else:
        S.ebe = op.h
    if isinstance(prec, BlockDiag):
        S.bdiag = prec.h
    else:
        S.ilu = prec.h if prec is not None else None
    return S


def pgmres(dev: Device, op, B, X, KrylovBasisVectorsQuantity=30, LinearMaxIter=1000,
           SolverTolerance=1e-8, precond=None):
    """pgmres (pgmres.c:5-200) on device vectors: B already preconditioned, X overwritten.
    precond: a factored CsrMatrix (ILU(0)) or a BlockDiag, applied on the left.
    Returns the SolveInfo (iterations = Arnoldi steps, as added to Parameters->iterations)."""
    info = N.SolveInfo()
    S = _system(op, precond)
    N.check(N.lib().femgpu_pgmres(dev.h, C.byref(S), int(KrylovBasisVectorsQuantity), int(LinearMaxIter),
                                  float(SolverTolerance), _ptr(B), _ptr(X), C.byref(info)))
    return info


def mixed_opts(inner_kmax=30, inner_tol=1e-4, max_refinements=50, sweep_mode=0, factor_sweeps=0,
               orthogonalization="MGS", graphs=False):
    """femgpu_mixed_opts from keyword values (one inner cycle per refinement)."""
    o = N.MixedOpts()
    o.inner_kmax, o.inner_cycles, o.inner_tol = inner_kmax, 1, inner_tol
    o.max_refinements, o.sweep_mode, o.factor_sweeps = max_refinements, sweep_mode, factor_sweeps
    o.orthogonalization = {"MGS": 0, "CGS": 1}[orthogonalization]
    o.graphs = 1 if graphs else 0
    return o


def gmres_mixed(dev: Device, A: CsrMatrix, F, X, tol=1e-8, inner_kmax=30, inner_tol=1e-4,
                max_refinements=50, sweep_mode=0, factor_sweeps=0, orthogonalization="MGS", graphs=False):
    """FP32 ILU(0)-GMRES inside FP64 iterative refinement; F is the raw right-hand side.

    sweep_mode 0 applies the exact FP32 sweeps of A's FP64 factorisation (A.ilu0_factor() first);
    sweep_mode k > 0 applies k Jacobi iterations per triangle. factor_sweeps k > 0 builds the FP32
    factors by k parallel ILU(0) sweeps instead, with no FP64 factorisation. orthogonalization
    "MGS" or "CGS" selects the inner Arnoldi's Gram-Schmidt variant; graphs=True replays each inner
    Arnoldi step from a captured hipGraph."""
    info = N.SolveInfo()
    o = mixed_opts(inner_kmax, inner_tol, max_refinements, sweep_mode, factor_sweeps, orthogonalization, graphs)
    S = _system(A, A)
    N.check(N.lib().femgpu_gmres_mixed(dev.h, C.byref(S), C.byref(o), float(tol), _ptr(F), _ptr(X),
                                       C.byref(info)))
    return info


def concat_csr(structures):
    """Block-diagonal concatenation of CSR patterns [(IA, JA), ...]: (IA, JA, row_off, nnz_off)."""
    row_off, nnz_off = [0], [0]
    IAs, JAs = [], []
    for IA, JA in structures:
        IA, JA = np.asarray(IA), np.asarray(JA)
        IAs.append(IA[:-1] + nnz_off[-1])
        JAs.append(JA + row_off[-1])
        row_off.append(row_off[-1] + len(IA) - 1)
        nnz_off.append(nnz_off[-1] + int(IA[-1]))
    IA = np.concatenate(IAs + [np.array([nnz_off[-1]])]).astype(np.int32)
    JA = np.concatenate(JAs).astype(np.int32)
    return IA, JA, np.array(row_off, np.int32), np.array(nnz_off, np.int64)


def gmres_mixed_batched(dev: Device, A: CsrMatrix, row_off, F, X, tol=1e-8, inner_kmax=30, inner_tol=1e-4,
                        max_refinements=50, sweep_mode=2, factor_sweeps=2, orthogonalization="MGS"):
    """Mixed precision over the independent systems of a block-diagonal A (rows row_off[s] ..
    row_off[s+1]), each verified and masked on its own; returns one BatchInfo per system."""
    ro = _i32(row_off)
    K = len(ro) - 1
    infos = (N.BatchInfo * K)()
    o = mixed_opts(inner_kmax, inner_tol, max_refinements, sweep_mode, factor_sweeps, orthogonalization)
    S = _system(A, A)
    N.check(N.lib().femgpu_gmres_mixed_batched(dev.h, C.byref(S), K, ro, C.byref(o), float(tol), _ptr(F), _ptr(X),
                                               infos))
    return list(infos)


def gmres_mixed_wg(dev: Device, A: CsrMatrix, row_off, F, X, tol=1e-8, inner_kmax=30, inner_tol=1e-4,
                   max_refinements=50, sweep_mode=2, factor_sweeps=2, threads=0):
    """The batched mixed-precision solve as one persistent launch, a workgroup per system (its own
    Krylov space: the answer does not depend on the batch); returns one BatchInfo per system."""
    ro = _i32(row_off)
    K = len(ro) - 1
    infos = (N.BatchInfo * K)()
    o = mixed_opts(inner_kmax, inner_tol, max_refinements, sweep_mode, factor_sweeps)
    S = _system(A, A)
    N.check(N.lib().femgpu_gmres_mixed_wg(dev.h, C.byref(S), K, ro, C.byref(o), float(tol), int(threads), _ptr(F),
                                          _ptr(X), infos))
    return list(infos)


def pcg(dev: Device, A: CsrMatrix, B, X, itmax=1000, tol=1e-8, precond: CsrMatrix | None = None):
    info = N.SolveInfo()
    S = _system(A, precond)
    N.check(N.lib().femgpu_pcg(dev.h, C.byref(S), int(itmax), float(tol), _ptr(B), _ptr(X), C.byref(info)))
    return info


# ----------------------------------------------------------------------------- drivers

def driver_opts(solver="fp64", kmax=30, lmax=1000, tol=1e-8, nonlinear_tol=1e-8, nonlinear_maxiter=50,
                stop_by_norm=True, inner_kmax=30, inner_tol=1e-4, max_refinements=50, sweep_mode=2,
                factor_sweeps=2, orthogonalization="MGS"):
    """Options of the nonlinear/time drivers: solver "fp64" = ILUp_precond_setup + pgmres (the
    reference's pair), "mixed" = gmres_mixed with the given FP32 preconditioner."""
    o = N.DriverOpts()
    o.solver = {"fp64": 0, "mixed": 1}[solver]
    o.kmax, o.lmax, o.tol = kmax, lmax, tol
    o.mixed = mixed_opts(inner_kmax, inner_tol, max_refinements, sweep_mode, factor_sweeps, orthogonalization)
    o.nonlinear_tol, o.nonlinear_maxiter, o.stop_by_norm = nonlinear_tol, nonlinear_maxiter, 1 if stop_by_norm else 0
    return o


def Process(M: "Mesh", A: CsrMatrix, pb, u, opts):
    """Steady transport fixed point (ST Process.c:29-86) on the device; u is updated in place."""
    info = N.DriverInfo()
    N.check(N.lib().femgpu_st_process(M.h, A.h, C.byref(pb), C.byref(opts), _ptr(u), C.byref(info)))
    return info


def Process_NS(M: "NsMesh", A: CsrMatrix, Re, u, opts, fx=None, fy=None):
    """Navier-Stokes Picard loop (NS Process.c:27-52) on the device; u is updated in place."""
    info = N.DriverInfo()
    N.check(N.lib().femgpu_ns_process(M.h, A.h, float(Re), _ptr(fx), _ptr(fy), C.byref(opts), _ptr(u),
                                      C.byref(info)))
    return info


def Predictor_Old_step(M: "Mesh", A: CsrMatrix, pb, dt, alpha, u, a, opts):
    """One predictor-multicorrector step (Predictor_Old.c:53-123); u and a are updated in place."""
    info = N.DriverInfo()
    N.check(N.lib().femgpu_tt_predictor_old_step(M.h, A.h, C.byref(pb), float(dt), float(alpha), C.byref(opts),
                                                 _ptr(u), _ptr(a), C.byref(info)))
    return info


def Predictor_Old_step_Euler(M: "EulerMesh", E: EbeMatrix, B: BlockDiag, pb, u, a, opts):
    """One Euler predictor-multicorrector step (EU Process.c:4-21 -> Predictor_Old.c:53-123) on the
    device with the EBE scheme and the BlockDiag preconditioner; u and a are updated in place.
    pb: make_euler_problem(...) carrying alpha and dt."""
    info = N.DriverInfo()
    N.check(N.lib().femgpu_eu_predictor_old_step(M.h, E.h, B.h, C.byref(pb), C.byref(opts), _ptr(u), _ptr(a),
                                                 C.byref(info)))
    return info


OBLIQUO_STATE = (1.0, 0.98481, -0.17365, 0.94643)   # rho, rho u, rho v, rho E (OBLIQUO_InitialSolution.c)


def obliquo_problem(mx, my, length=2.0, right=True, cfl=5.0, gamma=1.4):
    """The reference's OBLIQUO problem (03_Euler_Equations/2D/OBLIQUO): supersonic flow at Mach 2
    entering at -10 degrees onto the wall y = 0 (a 10-degree compression in the wall's frame), on a
    structured mesh of mx x my squares over [0, length] x [0, 1]. Marks of the generator's
    setmark_according_boundary (x = 0 and y = 1: all four dofs prescribed; y = 0: rho v
    prescribed), prescribed values of OBLIQUO_{rho,v1,v2,e}presc.c (rho v = 0 on the wall past the
    leading edge) and the initial state of OBLIQUO_InitialSolution.c at the unknowns. dt follows
    from the CFL number on the mesh spacing h = sqrt(2 A) and the freestream c + |v| (the reference
    takes dt as an input). Returns a dict of host arrays and the Euler problem parameters."""
    x, y, ien = structured_mesh(mx, my, box=(0.0, 0.0, float(length), 1.0), right=right)
    ndof, mark, t = structured_marks("OBLIQUO", x, y)
    assert ndof == 4
    invP, neq = number_unknowns(t, 4)
    lm = Fill_LM(ien, t, invP, neq, 4)
    nn = len(x)
    presc = np.tile(np.array(OBLIQUO_STATE), (nn, 1))
    presc[(np.abs(y) <= 1e-15) & (np.abs(x) > 0), 2] = 0.0
    presc = presc.ravel()
    u0 = np.zeros(neq)
    st = np.tile(np.array(OBLIQUO_STATE), nn)
    unk = t != 0
    u0[invP[unk]] = st[unk]
    rho, m1, m2, rE = OBLIQUO_STATE
    vmag = np.hypot(m1, m2) / rho
    c = np.sqrt(gamma * (gamma - 1.0) * (rE / rho - 0.5 * vmag * vmag))
    h = np.sqrt(float(length) / mx * 1.0 / my)
    dt = cfl * h / (c + vmag)
    return dict(x=x, y=y, ien=ien, type4=t, invP=invP, neq=neq, lm=lm, presc=presc, u0=u0, dt=dt,
                gamma=gamma, mach=vmag / c)


def cavity_problem(m, reorder="SYMRCM"):
    """The reference's CAVITY problem (04_SS_NavierStokes_Equations/2D/CAVITY): lid-driven cavity on a
    structured mesh of m x m squares over the unit square, (m+1)^2 nodes. Marks of the generator
    (walls: u and v prescribed; the centre node: p prescribed), prescribed values of
    CAVITY_{v1,v2,p}presc.c (u = 1 on the lid y = 1, else 0; p = 0), no body force
    (CAVITY_f{1,2}ext.c). Numbering, lm and the CSR structure (s = 9) as the steady NS driver builds
    them. Returns a dict of host arrays."""
    x, y, ien = structured_mesh(m, m)
    ndof, mark, t = structured_marks("CAVITY", x, y)
    assert ndof == 3
    invP0, neq = number_unknowns(t, 3)
    lm0 = Fill_LM(ien, t, invP0, neq, 3)
    c = csr_Initialization(lm0, 9, neq, reorder, permute_lm=True, type_=t, invP_id=invP0)
    presc = np.zeros((len(x), 3))
    presc[np.abs(y - 1.0) <= 1e-15, 0] = 1.0
    return dict(x=x, y=y, ien=ien, type3=t, invP=c.invP_id, presc=presc.ravel(), neq=neq, lm=c.lm, c=c,
                fx=np.zeros(len(x)), fy=np.zeros(len(x)))
