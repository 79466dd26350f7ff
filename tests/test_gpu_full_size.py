"""Parity at the workloads' full sizes, against the CPU oracle on the same inputs (the oracle is C,
so these stay within seconds of host time):
  * the 1001 x 1001 transient transport system (n = 998 001, 2 000 000 elements): element matrices,
    residual and CSR values bit-identical to the oracle, from both the fused and the element-array
    assembly; the exact ILU(0) factors bit-identical to the oracle's; the mixed-precision solve meets the 1e-8 FP64 bar in the oracle's own operator and
    lies within 1e-6 of the oracle's FP64 ILU(0) + pgmres solution; the EBE product of the element
    matrices equals the CSR product to 1e-13;
  * the 513 x 513 Navier-Stokes pattern (n = 785 410, rows of up to 11 strict entries per triangle,
    some results read beyond the streamed sweep's ring): ILU(0) factors and both sweeps bit-identical
    to the oracle;
  * the 801 x 401 Euler operator (640 000 12x12 elements): element matrices, residual, BlockDiag and
    the FP64 EBE product bit-identical to the oracle."""
import numpy as np
import pytest

from parallel_fem_codes_amd import fem, native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(scope="module")
def dev(torch_cuda):
    native.lib()   # fail loudly if the HIP library is not built
    d = fem.Device(0)
    yield d
    d.close()


def T(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a).copy()).to("cuda")


def H(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def test_transient_transport_1001(dev, torch_cuda, oracle):
    torch = torch_cuda
    m = 1000
    x, y, ien = oracle.mesh_structured(m, m)
    t = oracle.unit_square_types(x, y)
    invP0, neq = oracle.number_unknowns(t)
    lm0 = oracle.fill_lm(ien, t, invP0, neq)
    c = oracle.csr_init(lm0, 3, neq, reorder="SYMRCM", permute_lm=True, type_=t, invP=invP0)
    ubc = oracle.inflow_ubc(x, y)
    u0 = np.exp(-((x - 0.3) ** 2 + (y - 0.5) ** 2) / (2 * 0.05 ** 2))
    u = np.zeros(neq)
    u[c["invP"][t == 1]] = u0[t == 1]
    du = np.random.default_rng(12).standard_normal(neq) * 1e-3
    pbo = oracle.make_problem(k=1e-4)
    alpha, dt = 0.5, 1e-3
    Ae, R = oracle.build_transport_transient(x, y, t, c["invP"], ien, c["lm"], ubc, u, du, pbo, alpha, dt, neq)
    AAo = oracle.csr_assemble(c["scheme"], Ae, c["nnz"])
    print(f"1001x1001: n {neq}, nnz {c['nnz']}", flush=True)
    M = fem.Mesh(dev, x, y, t, c["invP"], ubc, ien, neq, lm=c["lm"], scheme=c["scheme"], nnz=c["nnz"])
    pb = fem.make_problem(k=1e-4)
    ug, dug = T(torch, u), T(torch, du)
    # element-array path: element matrices, residual and the CSR gather
    Aeg = torch.empty(len(Ae), dtype=torch.float64, device="cuda")
    Rg = torch.empty(neq + 1, dtype=torch.float64, device="cuda")
    AAg = torch.full((c["nnz"] + 1,), float("nan"), dtype=torch.float64, device="cuda")
    M.Build_M_K_R_SUPG(pb, alpha, dt, ug, dug, Rg, Ae=Aeg, AA_ptr=AAg.data_ptr())
    dev.sync()
    assert np.array_equal(H(Aeg), Ae)
    assert np.array_equal(H(Rg), R)
    assert np.array_equal(H(AAg), AAo)
    # fused path (CSR values and residual only)
    A = fem.CsrMatrix(dev, c["IA"], c["JA"])
    Rf = torch.full((neq + 1,), float("nan"), dtype=torch.float64, device="cuda")
    M.Build_M_K_R_SUPG(pb, alpha, dt, ug, dug, Rf, AA_ptr=A.values_ptr())
    A.values_changed()
    dev.sync()
    assert np.array_equal(H(Rf), R)
    # EBE product of the element matrices == CSR product (same unknown numbering)
    E = fem.EbeMatrix(dev, c["lm"], 3, neq)
    E.set_values(Aeg)
    p = T(torch, np.random.default_rng(3).standard_normal(neq))
    qe, qc = torch.empty_like(p), torch.empty_like(p)
    E.ebemv(p, qe)
    A.csrmv(p, qc)
    assert rel(H(qe), H(qc)) < 1e-13
    # mixed solve of A X = R, verified in the oracle's operator, against the oracle's FP64 solve
    X = torch.zeros(neq, dtype=torch.float64, device="cuda")
    info = fem.gmres_mixed(dev, A, Rf[:neq].clone(), X, tol=1e-8, inner_kmax=30, inner_tol=1e-4, sweep_mode=2,
                           factor_sweeps=1)
    Xh = H(X)
    res = np.linalg.norm(R[:neq] - oracle.csrmv(c["IA"], c["JA"], AAo, Xh)) / np.linalg.norm(R[:neq])
    LU, Dinv, rc = oracle.ilu0_factor(c["IA"], c["JA"], AAo)
    assert rc == 0
    # the exact ILU(0) on the device (compact fixed-slot rows on the structured meshes' pair
    # pattern): factors bit-identical to the oracle's
    assert A.sweep_info()["factor_compact"] == 2
    A.ilu0_factor()
    LUg, Dig = A.ilu0_factors()
    assert np.array_equal(LUg, LU[:-1]) and np.array_equal(Dig, Dinv)
    B = oracle.ilu0_apply(c["IA"], c["JA"], LU, Dinv, R[:neq])
    S = oracle.make_system(0, neq, IA=c["IA"], JA=c["JA"], AA=AAo, ilu=(c["IA"], c["JA"], LU, Dinv))
    Xo = oracle.pgmres(S, B, kmax=30, lmax=1000, tol=1e-12)[0][:neq]
    print(f"mixed: status {info.status}, {info.iterations} inner steps, residual in the oracle's operator "
          f"{res:.2e}, vs the oracle's solve {rel(Xh, Xo):.2e}", flush=True)
    assert info.status == 0 and res <= 1e-8
    assert rel(Xh, Xo) < 1e-6


def test_navier_stokes_pattern_513_ilu(dev, torch_cuda, oracle):
    torch = torch_cuda
    c = fem.cavity_problem(512)["c"]
    IA, JA = c.IA.astype(np.int32), c.JA.astype(np.int32)
    n = len(IA) - 1
    rng = np.random.default_rng(21)
    AA = rng.uniform(-1, 1, len(JA))
    rows = np.repeat(np.arange(n), np.diff(IA))
    AA[JA == rows] = 2.0 + np.bincount(rows, weights=np.abs(AA), minlength=n)
    AA = np.append(AA, 0.0)
    A = fem.CsrMatrix(dev, IA, JA)
    info = A.sweep_info()
    print(f"513x513 cavity pattern: n {n}, nnz {len(JA)}, {info}", flush=True)
    assert info["kernel"] == 0 and info["entries_per_row"] == 16 and info["long_slots"] > 0
    A.set_values(AA[:-1])
    A.ilu0_factor()
    LUo, Dio, rc = oracle.ilu0_factor(IA, JA, AA)
    assert rc == 0
    LU, Di = A.ilu0_factors()
    assert np.array_equal(LU, LUo[:-1]) and np.array_equal(Di, Dio)
    rhs = np.random.default_rng(22).standard_normal(n)
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    A.ilu0_apply(T(torch, rhs), z)
    assert np.array_equal(H(z), oracle.ilu0_apply(IA, JA, LUo, Dio, rhs))


def test_euler_obliquo_801x401(dev, torch_cuda, oracle):
    """The Euler workload's operator at full size (801 x 401 nodes, 640 000 elements, 4 unknowns per
    node) on a perturbed OBLIQUO state: the 12x12 element matrices and residual bit-identical to the
    oracle, the nodal BlockDiag preconditioner (blocks summed in element order, cofactor inverses)
    and the FP64 EBE product bit-identical to the oracle's, the FP32 product within FP32 rounding."""
    torch = torch_cuda
    P = fem.obliquo_problem(800, 400)
    neq, nn = P["neq"], len(P["x"])
    rng = np.random.default_rng(31)
    u = P["u0"] * (1.0 + 1e-2 * rng.uniform(-1, 1, neq))
    du = 1e-3 * rng.standard_normal(neq)
    prm = dict(gamma=1.4, Mach=2.0, alpha=0.5, dt=P["dt"], tolerance=1e-8, dimensionless=0, shock=1)
    U4, dU4 = oracle.euler_nodal_state(P["type4"], P["invP"], u, du, P["presc"])
    Ae, R = oracle.build_euler_supg(P["x"], P["y"], P["ien"], P["lm"], U4, dU4, oracle.euler_params(**prm), neq)
    print(f"801x401 OBLIQUO: neq {neq}, {len(P['ien']) // 3} elements", flush=True)
    M = fem.EulerMesh(dev, P["x"], P["y"], P["type4"], P["invP"], P["presc"], P["ien"], P["lm"], neq)
    Aeg = torch.empty(len(Ae), dtype=torch.float64, device="cuda")
    Rg = torch.empty(neq + 1, dtype=torch.float64, device="cuda")
    M.Build_M_K_F_SUPG(fem.make_euler_problem(**prm), T(torch, u), T(torch, du), Rg, Ae=Aeg)
    dev.sync()
    assert np.array_equal(H(Rg), R)
    assert np.array_equal(H(Aeg), Ae)
    # BlockDiag: setup and apply against the oracle's
    B = fem.BlockDiag(dev, nn, P["ien"], 4, P["type4"], P["invP"], neq)
    B.setup(Aeg)
    inv = oracle.blockdiag_setup(P["ien"], Ae, nn, 4)
    ids = oracle.blockdiag_ids(P["type4"], P["invP"], 4, neq)
    p = rng.standard_normal(neq)
    z = torch.empty(neq, dtype=torch.float64, device="cuda")
    B.apply(T(torch, p), z)
    assert np.array_equal(H(z), oracle.blockdiag_apply(ids, inv, p, neq, 4))
    # EBE products
    E = fem.EbeMatrix(dev, P["lm"], 12, neq)
    E.set_values(Aeg)
    del Aeg
    q = torch.empty(neq, dtype=torch.float64, device="cuda")
    E.ebemv(T(torch, p), q)
    ref = oracle.ebemv_block(P["lm"], Ae, p, neq, 12)
    assert np.array_equal(H(q), ref)
    q32 = torch.empty(neq, dtype=torch.float32, device="cuda")
    E.ebemv(T(torch, p).float(), q32)
    assert rel(H(q32).astype(np.float64), ref) < 1e-5
