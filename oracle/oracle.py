"""ctypes front end of the CPU FP64 oracle (``oracle/fem_oracle.c``).

TEST INFRASTRUCTURE: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker. The HIP product library never imports,
links or falls back to it.

Functions mirror the reference hot path (file:line citations live in ``fem_oracle.h``); the
helpers at the bottom set up the steady/transient transport problems of the benchmark
configurations (unit square, rotating field, all-Dirichlet boundary).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class TransportProblem(C.Structure):
    """Mirror of ``orc_transport_problem``."""

    _fields_ = [("k", C.c_double), ("gamma", C.c_double), ("f0", C.c_double),
                ("bx", C.c_double * 3), ("by", C.c_double * 3),
                ("shock", C.c_int), ("pad", C.c_int)]


class System(C.Structure):
    """Mirror of ``orc_system``."""

    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("nel", C.c_int), ("s", C.c_int),
                ("IA", C.c_void_p), ("JA", C.c_void_p), ("AA", C.c_void_p),
                ("lm", C.c_void_p), ("Perm", C.c_void_p), ("invPerm", C.c_void_p),
                ("A", C.c_void_p), ("precond", C.c_int),
                ("pIA", C.c_void_p), ("pJA", C.c_void_p), ("LU", C.c_void_p),
                ("Dinv", C.c_void_p), ("bd_nnodes", C.c_int), ("bd_ndof", C.c_int),
                ("bd_id", C.c_void_p), ("bd_inv", C.c_void_p)]


def build(force: bool = False) -> str:
    """Compile ``liboracle.so`` with the committed Makefile (gcc, strict IEEE)."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        L.orc_mesh_structured.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_int, f64p, f64p, i32p]
        L.orc_number_unknowns.argtypes = [C.c_int, C.c_int, i32p, i32p]
        L.orc_number_unknowns.restype = C.c_int
        L.orc_fill_lm.argtypes = [C.c_int, C.c_int, C.c_int, i32p, i32p, i32p, i32p]
        L.orc_symrcm.argtypes = [C.c_int, i32p, i32p, i32p]
        L.orc_csr_permute.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, i32p]
        L.orc_csr_nnz.argtypes = [C.c_int, C.c_int, C.c_int, i32p]
        L.orc_csr_nnz.restype = C.c_int
        L.orc_csr_init.argtypes = [C.c_int, C.c_int, C.c_int, i32p, C.c_int, C.c_int, C.c_int,
                                   i32p, i32p, i32p, i32p, i32p, i32p]
        L.orc_csr_init.restype = C.c_int
        L.orc_ebe_permutation.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, C.c_int, i32p, i32p]
        L.orc_ebe_permutation.restype = C.c_int
        L.orc_build_transport_supg.argtypes = [C.c_int, C.c_int, C.c_int, f64p, f64p, i32p, i32p,
                                               i32p, f64p, f64p, C.POINTER(TransportProblem),
                                               f64p, f64p]
        L.orc_build_transport_transient.argtypes = [C.c_int, C.c_int, C.c_int, f64p, f64p, i32p,
                                                    i32p, i32p, i32p, f64p, f64p, f64p,
                                                    C.POINTER(TransportProblem), C.c_double,
                                                    C.c_double, f64p, f64p]
        L.orc_csr_assemble.argtypes = [C.c_int, C.c_int, C.c_int, i32p, f64p, f64p]
        L.orc_euler_nodal_state.argtypes = [C.c_int, i32p, i32p, f64p, f64p, f64p, f64p, f64p]
        L.orc_build_euler_supg.argtypes = [C.c_int, C.c_int, C.c_int, f64p, f64p, i32p, i32p, f64p, f64p,
                                           C.POINTER(EulerParams), f64p, f64p]
        L.orc_ns_nodal_state.argtypes = [C.c_int, i32p, i32p, f64p, f64p, f64p]
        L.orc_build_ns_supg_pspg.argtypes = [C.c_int, C.c_int, C.c_int, f64p, f64p, i32p, i32p, f64p,
                                             f64p, f64p, C.c_double, f64p, f64p]
        L.orc_ebemv.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, f64p, f64p, f64p]
        L.orc_ebemv_block.argtypes = [C.c_int, C.c_int, C.c_int, i32p, f64p, f64p, f64p]
        L.orc_csrmv.argtypes = [C.c_int, i32p, i32p, f64p, f64p, f64p]
        L.orc_ddot.argtypes = [C.c_int, f64p, f64p]
        L.orc_ddot.restype = C.c_double
        L.orc_ilu0_factor.argtypes = [C.c_int, i32p, i32p, f64p, f64p, f64p]
        L.orc_ilu0_factor.restype = C.c_int
        L.orc_ilu0_apply.argtypes = [C.c_int, i32p, i32p, f64p, f64p, f64p, f64p]
        L.orc_blockdiag_setup.argtypes = [C.c_int, C.c_int, C.c_int, i32p, f64p, f64p]
        L.orc_blockdiag_apply.argtypes = [C.c_int, C.c_int, C.c_int, i32p, f64p, f64p, f64p]
        L.orc_pgmres.argtypes = [C.POINTER(System), C.c_int, C.c_int, C.c_double, f64p, f64p,
                                 C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_pgmres.restype = C.c_int
        L.orc_pcg.argtypes = [C.POINTER(System), C.c_int, C.c_double, f64p, f64p]
        L.orc_pcg.restype = C.c_int
        _LIB = L
    return _LIB


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------------------- thin wrappers

def mesh_structured(M, N, box=(0.0, 0.0, 1.0, 1.0), right=True):
    nn, ne = (M + 1) * (N + 1), 2 * M * N
    x, y, ien = np.empty(nn), np.empty(nn), np.empty(3 * ne, np.int32)
    lib().orc_mesh_structured(M, N, *box, 1 if right else 0, x, y, ien)
    return x, y, ien


def number_unknowns(type_, ndof=1):
    t = _i(type_)
    inv = np.empty_like(t)
    neq = lib().orc_number_unknowns(len(t) // ndof, ndof, t, inv)
    return inv, neq


def fill_lm(ien, type_, invP, NEQ, ndof=1):
    nel = len(ien) // 3
    lm = np.empty(nel * 3 * ndof, np.int32)
    lib().orc_fill_lm(nel, ndof, NEQ, _i(ien), _i(type_), _i(invP), lm)
    return lm


def symrcm(IA, JA):
    n = len(IA) - 1
    p = np.empty(n, np.int32)
    lib().orc_symrcm(n, _i(IA), _i(JA), p)
    return p


def csr_init(lm, s, neq, reorder="SYMRCM", permute_lm=False, type_=None, invP=None):
    """csr_Initialization for one rank. Returns dict(IA, JA, scheme, perm, lm, invP, nnz)."""
    code = {"NOT": 0, "SYMRCM": 1}[reorder.upper()]
    lm = _i(lm).copy()
    nel = len(lm) // s
    nnz = lib().orc_csr_nnz(nel, s, neq, lm)
    IA, JA = np.empty(neq + 1, np.int32), np.empty(max(nnz, 1), np.int32)
    scheme, perm = np.empty(nel * s * s, np.int32), np.empty(max(neq, 1), np.int32)
    if type_ is None:
        t = np.zeros(1, np.int32)
        inv = np.zeros(1, np.int32)
        nt = 0
    else:
        t, inv, nt = _i(type_), _i(invP).copy(), len(type_)
    r = lib().orc_csr_init(nel, s, neq, lm, code, 1 if permute_lm else 0, nt, t, inv, IA, JA,
                           scheme, perm)
    assert r == nnz
    return dict(IA=IA, JA=JA[:nnz], scheme=scheme, perm=perm[:neq], lm=lm,
                invP=inv if type_ is not None else None, nnz=nnz)


def ebe_permutation(ien, type_, invP, NEQ, reorder="SYMRCM"):
    code = {"NOT": 0, "SYMRCM": 1}[reorder.upper()]
    Perm, invPerm = np.empty(NEQ + 1, np.int32), np.empty(NEQ + 1, np.int32)
    lib().orc_ebe_permutation(len(ien) // 3, NEQ, _i(ien), _i(type_), _i(invP), code, Perm, invPerm)
    return Perm, invPerm


def make_problem(k=1e-4, gamma=0.0, f0=0.0, bx=(0.5, 0.0, -1.0), by=(-0.5, 1.0, 0.0), shock=0):
    p = TransportProblem()
    p.k, p.gamma, p.f0 = k, gamma, f0
    p.bx[:] = list(bx)
    p.by[:] = list(by)
    p.shock = shock
    return p


def build_transport_supg(x, y, type_, invP, ien, ubc, u, pb, neq):
    nel = len(ien) // 3
    Ke, F = np.empty(9 * nel), np.empty(max(neq, 1))
    lib().orc_build_transport_supg(len(x), nel, neq, _f(x), _f(y), _i(type_), _i(invP), _i(ien),
                                   _f(ubc), _f(u), C.byref(pb), Ke, F)
    return Ke, F[:neq]


def build_transport_transient(x, y, type_, invP, ien, lm, ubc, u, du, pb, alpha, dt, NEQ):
    nel = len(ien) // 3
    Ae, R = np.empty(9 * nel), np.empty(NEQ + 1)
    lib().orc_build_transport_transient(len(x), nel, NEQ, _f(x), _f(y), _i(type_), _i(invP),
                                        _i(ien), _i(lm), _f(ubc), _f(u), _f(du), C.byref(pb),
                                        alpha, dt, Ae, R)
    return Ae, R


class EulerParams(C.Structure):
    """orc_euler_params."""
    _fields_ = [("gamma", C.c_double), ("Mach", C.c_double), ("alpha", C.c_double), ("dt", C.c_double),
                ("tolerance", C.c_double), ("dimensionless", C.c_int), ("shock", C.c_int)]


def euler_params(gamma=1.4, Mach=2.0, alpha=0.5, dt=1e-2, tolerance=1e-8, dimensionless=1, shock=1):
    p = EulerParams()
    p.gamma, p.Mach, p.alpha, p.dt, p.tolerance = gamma, Mach, alpha, dt, tolerance
    p.dimensionless, p.shock = dimensionless, shock
    return p


def euler_nodal_state(type4, invP4, u, du, presc4):
    """eval_U_dU: nodal conservative state and increment."""
    U4, dU4 = np.empty(len(type4)), np.empty(len(type4))
    lib().orc_euler_nodal_state(len(type4) // 4, _i(type4), _i(invP4), _f(np.append(u, 0.0)),
                                _f(np.append(du, 0.0)), _f(presc4), U4, dU4)
    return U4, dU4


def build_euler_supg(x, y, ien, lm, U4, dU4, params, NEQ):
    """Build_M_K_F_SUPG: Ae[nel*144], R[NEQ+1]."""
    nel = len(ien) // 3
    Ae, R = np.empty(144 * nel), np.empty(NEQ + 1)
    lib().orc_build_euler_supg(len(x), nel, NEQ, _f(x), _f(y), _i(ien), _i(lm), _f(U4), _f(dU4), C.byref(params),
                               Ae, R)
    return Ae, R


def ns_nodal_state(type3, invP3, u, presc3):
    """eval_U: nodal (u, v, p) from the unknowns and the prescribed values."""
    U3 = np.empty(len(type3))
    lib().orc_ns_nodal_state(len(type3) // 3, _i(type3), _i(invP3), _f(np.append(u, 0.0)), _f(presc3), U3)
    return U3


def build_ns_supg_pspg(x, y, ien, lm, U3, fx, fy, Re, NEQ):
    """Build_K_F_SUPG_PSPG (Picard tangent): Ke[nel*81], F[NEQ+1]."""
    nel = len(ien) // 3
    Ke, F = np.empty(81 * nel), np.empty(NEQ + 1)
    lib().orc_build_ns_supg_pspg(len(x), nel, NEQ, _f(x), _f(y), _i(ien), _i(lm), _f(U3), _f(fx), _f(fy),
                                 float(Re), Ke, F)
    return Ke, F


def csr_assemble(scheme, Ke, nnz, s=3):
    AA = np.empty(nnz + 1)
    lib().orc_csr_assemble(len(scheme) // (s * s), s, nnz, _i(scheme), _f(Ke), AA)
    return AA


def csrmv(IA, JA, AA, P):
    n = len(IA) - 1
    Q = np.empty(n)
    lib().orc_csrmv(n, _i(IA), _i(JA), _f(AA), _f(P), Q)
    return Q


def ebemv(lm, Perm, invPerm, A, P, NEQ):
    Q = np.zeros(NEQ + 1)
    Pp = np.zeros(NEQ + 1)
    Pp[:NEQ] = P[:NEQ]
    lib().orc_ebemv(len(lm) // 3, NEQ, _i(lm), _i(Perm), _i(invPerm), _f(A), Pp, Q)
    return Q[:NEQ]


def ebemv_block(lm, A, P, NEQ, s):
    Q = np.zeros(NEQ + 1)
    Pp = np.zeros(NEQ + 1)
    Pp[:NEQ] = P[:NEQ]
    lib().orc_ebemv_block(len(lm) // s, s, NEQ, _i(lm), _f(A), Pp, Q)
    return Q[:NEQ]


def ddot(x, y):
    return lib().orc_ddot(len(x), _f(x), _f(y))


def ilu0_factor(IA, JA, AA):
    n = len(IA) - 1
    LU, Dinv = np.empty(len(JA) + 1), np.empty(max(n, 1))
    rc = lib().orc_ilu0_factor(n, _i(IA), _i(JA), _f(AA), LU, Dinv)
    return LU, Dinv[:n], rc


def ilu0_apply(IA, JA, LU, Dinv, p):
    n = len(IA) - 1
    z = np.empty(max(n, 1))
    lib().orc_ilu0_apply(n, _i(IA), _i(JA), _f(LU), _f(Dinv), _f(p), z)
    return z[:n]


@dataclass
class SystemArrays:
    """Keeps the numpy buffers alive while an ``orc_system`` points at them."""
    kind: int
    n: int
    arrays: dict = field(default_factory=dict)


def blockdiag_ids(type_, invP, ndof, NEQ):
    """Per node and dof: the unknown, or NEQ for a prescribed dof (BlockDiag_precond_EBE_setup.c:28-52)."""
    return np.where(np.asarray(type_) == 0, NEQ, np.asarray(invP)).astype(np.int32)


def blockdiag_setup(ien, A, nnodes, ndof):
    """Inverse nodal diagonal blocks of the element matrices A (nel*s*s, element order)."""
    inv = np.empty(nnodes * ndof * ndof)
    lib().orc_blockdiag_setup(nnodes, len(ien) // 3, ndof, _i(ien), _f(A), inv)
    return inv


def blockdiag_apply(ids, inv, p, NEQ, ndof):
    z = np.array(p[:NEQ], dtype=np.float64, copy=True)
    lib().orc_blockdiag_apply(len(ids) // ndof, ndof, NEQ, _i(ids), _f(inv), z, z)
    return z


def make_system(kind, n, *, IA=None, JA=None, AA=None, lm=None, Perm=None, invPerm=None, A=None,
                nel=0, s=3, ilu=None, bdiag=None):
    keep = {}

    def ptr(name, a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dt)
        keep[name] = a
        return a.ctypes.data

    S = System()
    S.kind, S.n, S.nel, S.s = kind, n, nel, s
    S.IA, S.JA, S.AA = ptr("IA", IA, np.int32), ptr("JA", JA, np.int32), ptr("AA", AA, np.float64)
    S.lm, S.Perm = ptr("lm", lm, np.int32), ptr("Perm", Perm, np.int32)
    S.invPerm, S.A = ptr("invPerm", invPerm, np.int32), ptr("A", A, np.float64)
    if ilu is not None:
        pIA, pJA, LU, Dinv = ilu
        S.precond = 1
        S.pIA, S.pJA = ptr("pIA", pIA, np.int32), ptr("pJA", pJA, np.int32)
        S.LU, S.Dinv = ptr("LU", LU, np.float64), ptr("Dinv", Dinv, np.float64)
    elif bdiag is not None:
        ids, inv, ndof = bdiag
        S.precond = 2
        S.bd_nnodes, S.bd_ndof = len(ids) // ndof, ndof
        S.bd_id, S.bd_inv = ptr("bd_id", ids, np.int32), ptr("bd_inv", inv, np.float64)
    else:
        S.precond = 0
    S._keep = keep
    return S


def pgmres(S, B, kmax=30, lmax=1000, tol=1e-8):
    n = S.n
    Bp = np.zeros(n + 1)
    Bp[:n] = B[:n]
    X = np.zeros(n + 1)
    rho, cyc = C.c_double(), C.c_int()
    it = lib().orc_pgmres(C.byref(S), kmax, lmax, tol, Bp, X, C.byref(rho), C.byref(cyc))
    return X[:n], it, rho.value, cyc.value


def pcg(S, B, itmax=1000, tol=1e-8):
    n = S.n
    X = np.zeros(n + 1)
    it = lib().orc_pcg(C.byref(S), itmax, tol, _f(np.append(B[:n], 0.0)), X)
    return X[:n], it


# ----------------------------------------------------------------------------- problem setup

def unit_square_types(x, y, tol=1e-15):
    """Mark 15 (all dofs Dirichlet) on the boundary of [0,1]^2, mark 0 inside
    (MG/Structured_Mesh/meshGenerator.c:315-332 for CARTOLA/TESTE)."""
    b = (np.abs(x - 1.0) <= tol) | (np.abs(x) <= tol) | (np.abs(y - 1.0) <= tol) | (np.abs(y) <= tol)
    return np.where(b, 0, 1).astype(np.int32)


def inflow_ubc(x, y):
    """Dirichlet data of the benchmark transport problems: u = 1 on x = 0, 0.25 < y < 0.5."""
    return np.where((np.abs(x) <= 1e-15) & (y > 0.25) & (y < 0.5), 1.0, 0.0)


def steady_transport_setup(m, scheme="CSR", reorder="SYMRCM", right=True, k=1e-4, shock=0,
                           ubc_fn=inflow_ubc, bx=(0.5, 0.0, -1.0), by=(-0.5, 1.0, 0.0)):
    """Mesh of m x m squares, numbering, lm and the storage structure for the steady transport
    driver (01_SS_Transport_Equation Preprocess.c:299-357 on one rank)."""
    x, y, ien = mesh_structured(m, m, right=right)
    t = unit_square_types(x, y)
    invP, neq = number_unknowns(t)
    lm = fill_lm(ien, t, invP, neq)
    out = dict(x=x, y=y, ien=ien, type=t, neq=neq, ubc=ubc_fn(x, y),
               pb=make_problem(k=k, shock=shock, bx=bx, by=by))
    if scheme.upper() == "CSR":
        c = csr_init(lm, 3, neq, reorder=reorder, type_=t, invP=invP)
        out.update(invP=c["invP"], lm=lm, IA=c["IA"], JA=c["JA"], scheme=c["scheme"],
                   perm=c["perm"], nnz=c["nnz"])
    else:
        Perm, invPerm = ebe_permutation(ien, t, invP, neq, reorder=reorder)
        out.update(invP=invP, lm=lm, Perm=Perm, invPerm=invPerm)
    return out
