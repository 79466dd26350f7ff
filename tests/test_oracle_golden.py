"""Pin the CPU oracle against fixtures written by the reference's own code (tests/golden).

Both sides are strict-IEEE builds (no fast-math, no FMA contraction) that follow the same
operation order, so everything -- integer structures, element matrices, ILU factors, GMRES
iteration counts and solutions, the CAU fixed-point sequence -- must match bit for bit.
"""
import numpy as np
import pytest

import golden_cases as G

CASES = G.names()


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


def setup(O, d):
    h = d["header"]
    nn, nel, ebe, reorder = int(h[0]), int(h[1]), int(h[2]), int(h[3])
    invP, neq = O.number_unknowns(d["type"])
    lm = O.fill_lm(d["ien"], d["type"], invP, neq)
    pa = G.problem_args(d)
    pb = O.make_problem(k=pa["k"], bx=pa["bx"], by=pa["by"], shock=pa["shock"])
    return nn, nel, ebe, reorder, invP, neq, lm, pb


@pytest.mark.parametrize("name", CASES)
def test_mesh_matches_reference_generator(oracle, name):
    d = G.load(name)
    m = d["meta"]
    x, y, ien = oracle.mesh_structured(m["M"], m["N"], right=(m["cut"] == "right"))
    assert np.array_equal(ien, d["ien"])
    # the generator prints coordinates with %.14lf
    assert np.max(np.abs(x - d["x"])) < 1e-14 and np.max(np.abs(y - d["y"])) < 1e-14
    assert np.array_equal(oracle.unit_square_types(d["x"], d["y"]), d["type"])


@pytest.mark.parametrize("name", CASES)
def test_structure_bit_exact(oracle, name):
    d = G.load(name)
    nn, nel, ebe, reorder, invP, neq, lm, pb = setup(oracle, d)
    assert np.array_equal(lm, d["ref_lm"])
    ro = "SYMRCM" if reorder else "NOT"
    if ebe:
        Perm, invPerm = oracle.ebe_permutation(d["ien"], d["type"], invP, neq, reorder=ro)
        assert np.array_equal(Perm, d["ref_Perm"])
        assert np.array_equal(invPerm, d["ref_invPerm"])
        assert np.array_equal(invP, d["ref_invP_id"])
    else:
        c = oracle.csr_init(lm, 3, neq, reorder=ro, type_=d["type"], invP=invP)
        assert np.array_equal(c["IA"], d["ref_IA"])
        assert np.array_equal(c["JA"], d["ref_JA"])
        assert np.array_equal(c["perm"], d["ref_perm"])
        assert np.array_equal(c["scheme"], d["ref_scheme"])
        assert np.array_equal(c["invP"], d["ref_invP_id"])


def _first_system(oracle, d):
    nn, nel, ebe, reorder, invP, neq, lm, pb = setup(oracle, d)
    ro = "SYMRCM" if reorder else "NOT"
    out = dict(neq=neq, ebe=ebe, lm=lm, pb=pb)
    if ebe:
        Perm, invPerm = oracle.ebe_permutation(d["ien"], d["type"], invP, neq, reorder=ro)
        out.update(invP=invP, Perm=Perm, invPerm=invPerm)
    else:
        c = oracle.csr_init(lm, 3, neq, reorder=ro, type_=d["type"], invP=invP)
        out.update(invP=c["invP"], IA=c["IA"], JA=c["JA"], scheme=c["scheme"], nnz=c["nnz"])
    Ke, F = oracle.build_transport_supg(d["x"], d["y"], d["type"], out["invP"], d["ien"], d["ubc"],
                                        np.zeros(neq), pb, neq)
    out.update(Ke=Ke, F=F)
    if not ebe:
        out["AA"] = oracle.csr_assemble(out["scheme"], Ke, out["nnz"])
    return out


@pytest.mark.parametrize("name", CASES)
def test_assembly_and_matvec(oracle, name):
    d = G.load(name)
    s = _first_system(oracle, d)
    assert same(s["F"], d["ref_F"])
    pv = np.zeros(s["neq"])
    t1 = d["type"] == 1
    pv[s["invP"][t1]] = d["probe"][t1]
    if s["ebe"]:
        assert same(s["Ke"], d["ref_A"])
        q = oracle.ebemv(s["lm"], s["Perm"], s["invPerm"], s["Ke"], pv, s["neq"])
    else:
        assert same(s["AA"], d["ref_AA"])
        q = oracle.csrmv(s["IA"], s["JA"], s["AA"], pv)
    assert same(q, d["ref_mv"])


@pytest.mark.parametrize("name", [n for n in CASES if "_csr_" in n])
def test_ilu0_factors_and_apply(oracle, name):
    d = G.load(name)
    s = _first_system(oracle, d)
    LU, Dinv, rc = oracle.ilu0_factor(s["IA"], s["JA"], s["AA"])
    assert rc == 0
    IA, JA = s["IA"], s["JA"]
    row = np.repeat(np.arange(s["neq"]), np.diff(IA))
    lower, upper = JA < row, JA > row
    assert same(LU[:-1][lower], d["ref_ILU_L"])
    assert same(LU[:-1][upper], d["ref_ILU_U"])
    assert same(Dinv, d["ref_ILU_D"])
    assert same(oracle.ilu0_apply(IA, JA, LU, Dinv, s["F"]), d["ref_MinvF"])


def _solve_sequence(oracle, d):
    """Process.c:29-86 (steady fixed point) driven through the oracle primitives."""
    h = d["header"]
    kmax, lmax, nl = int(h[5]), int(h[6]), int(h[7])
    nltol = float(d["params"][10])
    s = _first_system(oracle, d)
    neq = s["neq"]
    u = np.zeros(neq)
    iters, X1 = [], None
    normdiff, normu = nltol + 1, 1.0 / nltol
    while normdiff > nltol * normu and len(iters) < nl:
        Ke, F = oracle.build_transport_supg(d["x"], d["y"], d["type"], s["invP"], d["ien"], d["ubc"],
                                            u, s["pb"], neq)
        if s["ebe"]:
            S = oracle.make_system(1, neq, lm=s["lm"], Perm=s["Perm"], invPerm=s["invPerm"], A=Ke,
                                   nel=len(d["ien"]) // 3)
            B = F
        else:
            AA = oracle.csr_assemble(s["scheme"], Ke, s["nnz"])
            LU, Dinv, rc = oracle.ilu0_factor(s["IA"], s["JA"], AA)
            B = oracle.ilu0_apply(s["IA"], s["JA"], LU, Dinv, F)
            S = oracle.make_system(0, neq, IA=s["IA"], JA=s["JA"], AA=AA,
                                   ilu=(s["IA"], s["JA"], LU, Dinv))
        uold = u.copy()
        u, it, rho, cyc = oracle.pgmres(S, B, kmax=kmax, lmax=lmax, tol=1e-8)
        iters.append(it)
        if X1 is None:
            X1 = u.copy()
        normdiff = np.sqrt(oracle.ddot(uold - u, uold - u))
        normu = np.sqrt(oracle.ddot(u, u))
    return iters, X1, u


@pytest.mark.parametrize("name", CASES)
def test_gmres_iterations_and_solutions(oracle, name):
    d = G.load(name)
    iters, X1, u = _solve_sequence(oracle, d)
    assert iters == d["ref_iters"].tolist()
    assert same(X1, d["ref_X1"])
    assert same(u, d["ref_u"])


# ----------------------------------------------------------------------------- Navier-Stokes (NDOF = 3)

NS_CASES = G.names("ns_")


def ns_setup(oracle, d):
    """Numbering, lm and (CSR) structure of a Navier-Stokes fixture, as the reference builds them:
    per-dof running numbering, Fill_LM with NDOF = 3, csr_Initialization with s = 9 that also
    renumbers lm and invP_id (04_SS_NavierStokes_Equations/.../csr_Initialization.c:330-345)."""
    h = d["header"]
    ebe, reorder = int(h[2]), int(h[3])
    type3 = d["type3"]
    invP, neq = oracle.number_unknowns(type3, 3)
    lm = oracle.fill_lm(d["ien"], type3, invP, neq, 3)
    assert np.array_equal(lm, d["ref_lm"])      # Fill_LM output, before any renumbering
    c = None
    if not ebe:
        c = oracle.csr_init(lm, 9, neq, reorder="SYMRCM" if reorder else "NOT", permute_lm=True,
                            type_=type3, invP=invP)
        lm, invP = c["lm"], c["invP"]
    u = np.zeros(neq)
    unk = type3 != 0
    u[invP[unk]] = d["state"][unk]
    return ebe, neq, lm, invP, u, c


@pytest.mark.parametrize("name", NS_CASES)
def test_ns_structure_and_assembly_bit_exact(oracle, name):
    d = G.load(name)
    ebe, neq, lm, invP, u, c = ns_setup(oracle, d)
    assert np.array_equal(invP, d["ref_invP_id"])
    if c is not None:
        for k in ("IA", "JA", "perm", "scheme"):
            assert np.array_equal(c[k], d["ref_" + k]), k
    U3 = oracle.ns_nodal_state(d["type3"], invP, u, d["presc"])
    f = d["fsrc"].reshape(-1, 2)
    Ke, F = oracle.build_ns_supg_pspg(d["x"], d["y"], d["ien"], lm, U3, f[:, 0].copy(), f[:, 1].copy(),
                                      float(d["params"][0]), neq)
    assert np.array_equal(F, d["ref_F"])
    if ebe:
        assert np.array_equal(Ke, d["ref_A"])
    else:
        assert np.array_equal(oracle.csr_assemble(c["scheme"], Ke, c["nnz"], s=9), d["ref_AA"])


# ----------------------------------------------------------------------------- Euler (NDOF = 4)

EU_CASES = G.names("eu_")


def eu_setup(oracle, d):
    """Numbering, lm, (CSR) structure, state and increment of an Euler fixture, as the reference
    builds them (Fill_LM with NDOF = 4; csr_Initialization with s = 12 renumbering lm and invP_id)."""
    h = d["header"]
    ebe, reorder = int(h[2]), int(h[3])
    type4 = d["type4"]
    invP, neq = oracle.number_unknowns(type4, 4)
    lm = oracle.fill_lm(d["ien"], type4, invP, neq, 4)
    assert np.array_equal(lm, d["ref_lm"])
    c = None
    if not ebe:
        c = oracle.csr_init(lm, 12, neq, reorder="SYMRCM" if reorder else "NOT", permute_lm=True,
                            type_=type4, invP=invP)
        lm, invP = c["lm"], c["invP"]
    u, du = np.zeros(neq), np.zeros(neq)
    unk = type4 != 0
    u[invP[unk]] = d["state"][unk]
    du[invP[unk]] = d["incr"][unk]
    pr = d["params"]
    params = oracle.euler_params(gamma=pr[0], Mach=pr[1], alpha=pr[2], dt=pr[3], tolerance=pr[4],
                                 dimensionless=int(h[4]), shock=int(h[5]))
    return ebe, neq, lm, invP, u, du, c, params


@pytest.mark.parametrize("name", EU_CASES)
def test_euler_structure_and_assembly_bit_exact(oracle, name):
    d = G.load(name)
    ebe, neq, lm, invP, u, du, c, params = eu_setup(oracle, d)
    assert np.array_equal(invP, d["ref_invP_id"])
    if c is not None:
        for k in ("IA", "JA", "perm", "scheme"):
            assert np.array_equal(c[k], d["ref_" + k]), k
    U4, dU4 = oracle.euler_nodal_state(d["type4"], invP, u, du, d["presc"])
    Ae, R = oracle.build_euler_supg(d["x"], d["y"], d["ien"], lm, U4, dU4, params, neq)
    assert np.array_equal(R, d["ref_R"])
    if ebe:
        assert np.array_equal(Ae, d["ref_A"])
    else:
        assert np.array_equal(oracle.csr_assemble(c["scheme"], Ae, c["nnz"], s=12), d["ref_AA"])


def test_euler_blockdiag_gmres_matches_reference(oracle):
    """The Euler linear solve of the EBE scheme as the reference runs it (BlockDiag_precond_EBE_setup,
    then pgmres with ebemvNDOF4 and BlockDiag_precond): the inverse nodal blocks and D^-1 R are the
    reference's bit for bit (its 4x4 cofactor inverse included), the Arnoldi step count is equal and
    the solution agrees to 1e-10 relative L2."""
    d = G.load("eu_ebe_dimless_cau_r8")
    ebe, neq, lm, invP, u, du, c, params = eu_setup(oracle, d)
    assert ebe
    nn = len(d["x"])
    inv = oracle.blockdiag_setup(d["ien"], d["ref_A"], nn, 4)
    assert np.array_equal(inv, d["ref_invBlockDiag"])
    ids = oracle.blockdiag_ids(d["type4"], invP, 4, neq)
    B = oracle.blockdiag_apply(ids, inv, d["ref_R"], neq, 4)
    assert np.array_equal(B, d["ref_MinvR"])
    kmax, lmax, tol = int(d["header"][6]), int(d["header"][7]), float(d["params"][5])
    S = oracle.make_system(2, neq, lm=lm, A=d["ref_A"], nel=len(d["ien"]) // 3, s=12, bdiag=(ids, inv, 4))
    X, it, rho, cyc = oracle.pgmres(S, B, kmax=kmax, lmax=lmax, tol=tol)
    assert it == int(d["ref_iters"][0])
    assert np.linalg.norm(X - d["ref_X"]) <= 1e-10 * np.linalg.norm(d["ref_X"])
