"""The persistent batched mixed-precision solver (one workgroup per system) on the serving
workload's steady transport systems: every system reaches the FP64-verified 1e-8 residual (checked
again on the host in FP64), a system's answer is bit-identical alone and in a batch (its Krylov space
is its own), it agrees with a direct sparse FP64 solve within what the residual allows, and the
three workgroup sizes all converge."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from parallel_fem_codes_amd import fem, native as N, serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(scope="module")
def dev(torch_cuda):
    N.lib()
    d = fem.Device(0)
    yield d
    d.close()


def assemble(dev, torch, r):
    """The request's system with the library's assembly (Build_K_F_SUPG at u = 0): host IA, JA, AA, F."""
    t = r["type"]
    invP, neq = fem.number_unknowns(t)
    lm = fem.Fill_LM(r["ien"], t, invP, neq)
    c = fem.csr_Initialization(lm, 3, neq, "SYMRCM", type_=t, invP_id=invP)
    M = fem.Mesh(dev, r["x"], r["y"], t, c.invP_id, r["ubc"], r["ien"], neq, scheme=c.scheme, nnz=c.nnz)
    A = fem.CsrMatrix(dev, c.IA, c.JA)
    u0 = torch.zeros(neq, dtype=torch.float64, device="cuda")
    F = torch.empty(neq, dtype=torch.float64, device="cuda")
    M.Build_K_F_SUPG(r["pb"], u0, F, AA_ptr=A.values_ptr())
    AA = torch.empty(c.nnz, dtype=torch.float64, device="cuda")
    N.check(N.lib().femgpu_dcopy(dev.h, c.nnz, C.c_void_p(A.values_ptr()), C.c_void_p(AA.data_ptr())))
    dev.sync()
    return dict(IA=np.asarray(c.IA, np.int32), JA=np.asarray(c.JA, np.int32), AA=AA.cpu().numpy(),
                F=F.cpu().numpy(), n=neq)


@pytest.fixture(scope="module")
def systems(dev, torch_cuda):
    reqs = serving.make_requests(6, 31, 300, 6000, rate=1e4)
    return [assemble(dev, torch_cuda, r) for r in reqs]


def batch_solve(dev, torch, syss, threads=256, solver="wg"):
    IA, JA, row_off, nnz_off = fem.concat_csr([(s["IA"], s["JA"]) for s in syss])
    A = fem.CsrMatrix(dev, IA, JA)
    A.set_values(np.concatenate([s["AA"] for s in syss]))
    A.values_changed()
    F = torch.from_numpy(np.concatenate([s["F"] for s in syss])).cuda()
    X = torch.empty_like(F)
    if solver == "wg":
        infos = fem.gmres_mixed_wg(dev, A, row_off, F, X, tol=1e-8, inner_kmax=30, threads=threads)
    else:
        infos = fem.gmres_mixed_batched(dev, A, row_off, F, X, tol=1e-8, inner_kmax=30)
    x = X.cpu().numpy()
    return [x[row_off[k]:row_off[k + 1]] for k in range(len(syss))], infos


def csr(s):
    return sp.csr_matrix((s["AA"], s["JA"], s["IA"]), shape=(s["n"], s["n"]))


def host_rel_residual(s, x):
    return float(np.linalg.norm(s["F"] - csr(s) @ x) / np.linalg.norm(s["F"]))


def test_wg_converges_and_is_verified_in_fp64(dev, torch_cuda, systems):
    xs, infos = batch_solve(dev, torch_cuda, systems)
    for s, x, inf in zip(systems, xs, infos):
        res = host_rel_residual(s, x)
        print(f"n={s['n']}: refinements {inf.refinements} inner steps {inf.inner_iterations} "
              f"device residual {inf.true_rel_residual:.2e} host residual {res:.2e}")
        assert inf.status == 0 and inf.true_rel_residual <= 1e-8
        assert res <= 1.01e-8
        assert abs(res - inf.true_rel_residual) <= 1e-3 * res + 1e-15
        xd = spla.spsolve(csr(s).tocsc(), s["F"])
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-5


def test_wg_alone_equals_in_batch_bitwise(dev, torch_cuda, systems):
    """At one workgroup size (the size sets the order of the workgroup's sums)."""
    mix, _ = batch_solve(dev, torch_cuda, systems)
    for k, s in enumerate(systems):
        alone, info = batch_solve(dev, torch_cuda, [s])
        assert info[0].status == 0
        assert np.array_equal(alone[0], mix[k]), k
    # a different batch order and neighbours: still the same bits
    order = [3, 0, 5, 1, 4, 2]
    shuffled, _ = batch_solve(dev, torch_cuda, [systems[k] for k in order])
    for pos, k in enumerate(order):
        assert np.array_equal(shuffled[pos], mix[k]), k


@pytest.mark.parametrize("threads", [0, 256, 512, 1024])
def test_wg_workgroup_sizes(dev, torch_cuda, systems, threads):
    xs, infos = batch_solve(dev, torch_cuda, systems, threads=threads)
    assert all(i.status == 0 for i in infos)
    assert max(host_rel_residual(s, x) for s, x in zip(systems, xs)) <= 1.01e-8


def test_wg_matches_shared_krylov_batched_solver(dev, torch_cuda, systems):
    a, ia = batch_solve(dev, torch_cuda, systems)
    b, ib = batch_solve(dev, torch_cuda, systems, solver="batched")
    for s, x, y in zip(systems, a, b):
        assert np.linalg.norm(x - y) / np.linalg.norm(y) < 1e-5
    assert all(i.status == 0 for i in ib)


def test_wg_rejects_unsupported_options(dev, torch_cuda, systems):
    s = systems[0]
    A = fem.CsrMatrix(dev, s["IA"], s["JA"])
    A.set_values(s["AA"])
    F = torch_cuda.from_numpy(s["F"]).cuda()
    X = torch_cuda.empty_like(F)
    with pytest.raises(Exception):
        fem.gmres_mixed_wg(dev, A, [0, s["n"]], F, X, factor_sweeps=0)
    with pytest.raises(Exception):
        fem.gmres_mixed_wg(dev, A, [0, s["n"]], F, X, inner_kmax=65)
    with pytest.raises(Exception):
        fem.gmres_mixed_wg(dev, A, [0, s["n"]], F, X, threads=384)
