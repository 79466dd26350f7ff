"""The serving loop (continuous batching, mixed precision) on the GPU: every request converges to
the FP64-verified 1e-8 residual, its result is the same alone and in the mix and matches the
plain FP64 path (ILU(0) + pgmres) within the tolerance the 1e-8 residual allows, and small
requests are actually packed into batches."""
import numpy as np
import pytest

from parallel_fem_codes_amd import fem, native, serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def reqs():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    native.lib()
    small = serving.make_requests(10, 7, 200, 2000, rate=1e4)
    large = serving.make_requests(2, 8, 15000, 20000, rate=1e4)
    r = small[:5] + large[:1] + small[5:] + large[1:]
    for k, q in enumerate(r):
        q["arrival"] = 1e-4 * k
    return r


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("small_solver", ["persistent", "shared"])
def test_server_results_alone_in_the_mix_and_fp64(reqs, small_solver):
    """persistent: every system has its own Krylov space, so alone and in the mix agree bit for bit
    (at one workgroup size: the size sets the order of the workgroup's sums); shared: the batch shares
    one Krylov space, so they agree within what the 1e-8 residual allows."""
    opts = serving.server_opts(executors=2, prep_threads=4, small_rows=5000, batch_rows=100000, max_batch=4,
                               small_solver=small_solver, wg_threads=512)
    mix, infos, wall, batch, wg = serving.run_server(reqs, opts)
    assert batch == 4
    assert max(i.batch_size for i in infos) > 1                    # small requests were packed
    small = [i.n <= 5000 for i in infos]
    if small_solver == "persistent":
        assert all(i.wg_threads == (512 if s else 0) for i, s in zip(infos, small))
    else:
        assert all(i.wg_threads == 0 for i in infos)
    assert all(i.status == 0 for i in infos)
    assert max(i.true_rel_residual for i in infos) <= 1e-8
    assert all(i.t_end >= i.t_start >= i.t_submit for i in infos)
    alone_opts = serving.server_opts(executors=1, prep_threads=2, small_rows=5000, batch_rows=100000, max_batch=1,
                                     small_solver=small_solver, wg_threads=512)
    alone, ainfos, _, _, _ = serving.run_server(reqs, alone_opts)
    assert all(i.batch_size == 1 and i.status == 0 for i in ainfos)
    dev = fem.Device(0)
    worst_alone = worst_fp64 = 0.0
    for k, r in enumerate(reqs):
        u64, info = serving.solve_fp64(dev, r)
        assert info.status == 0
        worst_alone = max(worst_alone, rel(mix[k], alone[k]))
        worst_fp64 = max(worst_fp64, rel(mix[k], u64))
        if small_solver == "persistent":
            assert np.array_equal(mix[k], alone[k]), k
        assert rel(mix[k], alone[k]) < 1e-6, k
        assert rel(mix[k], u64) < 1e-5, k
        # Dirichlet values are carried through exactly
        assert np.array_equal(mix[k][r["type"] == 0], r["ubc"][r["type"] == 0])
    print(f"worst relative difference: alone vs mix {worst_alone:.2e}, mix vs FP64 {worst_fp64:.2e}")
    dev.close()


def test_server_matches_oracle(reqs, oracle):
    """The served answers against the CPU oracle's own steady pipeline on the same requests (the
    reference's numbering, lm, SYMRCM CSR and SUPG assembly, restated): the oracle's operator and load
    applied to each served solution leave a relative residual within the serving bar (1e-8, plus
    the difference of summation order), and the oracle's FP64 ILU(0) + pgmres solution agrees within
    what that residual allows on these convection-dominated systems."""
    sub = [r for r in reqs if len(r["x"]) <= 2500][:4]
    opts = serving.server_opts(executors=2, prep_threads=2, small_rows=5000, batch_rows=100000, max_batch=2)
    sols, infos, _, _, _ = serving.run_server(sub, opts)
    for r, u, info in zip(sub, sols, infos):
        t = r["type"]
        pb = oracle.make_problem(k=r["pb"].k, bx=tuple(r["pb"].bx), by=tuple(r["pb"].by))
        invP0, neq = oracle.number_unknowns(t)
        lm = oracle.fill_lm(r["ien"], t, invP0, neq)
        c = oracle.csr_init(lm, 3, neq, reorder="SYMRCM", type_=t, invP=invP0)
        Ke, F = oracle.build_transport_supg(r["x"], r["y"], t, c["invP"], r["ien"], r["ubc"], np.zeros(neq), pb, neq)
        AA = oracle.csr_assemble(c["scheme"], Ke, c["nnz"])
        U = np.zeros(neq)
        U[c["invP"][t == 1]] = u[t == 1]
        res = np.linalg.norm(F[:neq] - oracle.csrmv(c["IA"], c["JA"], AA, U)) / np.linalg.norm(F[:neq])
        LU, Dinv, rc = oracle.ilu0_factor(c["IA"], c["JA"], AA)
        assert rc == 0
        B = oracle.ilu0_apply(c["IA"], c["JA"], LU, Dinv, F[:neq])
        S = oracle.make_system(0, neq, IA=c["IA"], JA=c["JA"], AA=AA, ilu=(c["IA"], c["JA"], LU, Dinv))
        Uo = oracle.pgmres(S, B, kmax=30, lmax=2000, tol=1e-12)[0][:neq]
        print(f"request {r['index']} ({len(r['x'])} points, k {r['pb'].k:.2e}): residual in the oracle's "
              f"system {res:.2e} (served {info.true_rel_residual:.2e}), vs the oracle's solve {rel(U, Uo):.2e}")
        assert info.status == 0 and res <= 2e-8
        assert rel(U, Uo) < 1e-5
        assert np.array_equal(u[t == 0], r["ubc"][t == 0])


def test_server_autotunes_batch_size(reqs):
    many = serving.make_requests(48, 11, 200, 1500, rate=1e4)
    opts = serving.server_opts(executors=1, prep_threads=4, small_rows=5000, batch_rows=200000, max_batch=0,
                               small_solver="auto")
    sols, infos, wall, batch, wg = serving.run_server(many, opts)
    assert all(i.status == 0 for i in infos)
    assert batch in (0, 4, 8, 16, 32, 64)
    assert wg in (0, 256, 512, 1024)
    assert all(i.wg_threads in (256, 512, 1024) for i in infos)
