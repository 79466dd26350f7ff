"""The reference-side binding, executed: the reference's own steady-transport code (Fill_LM,
Permutation_of_LM, csr_Initialization with SYMRCM, Build_K_F_SUPG with F_assembly and
CAU_ShockCapture, the nonlinear loop of Process.c) built by oracle/ref.mk with the shim
parallel_fem_codes_amd/compat/femgpu_compat.c and libfemgpu.so in place of its ebemv, csrmv,
ILUp_precond(_setup), NO_precond_setup and pgmres objects, on the reference fixtures' inputs.

Bars: the preconditioned right-hand side M^-1 F and the operator probe A p bit-identical to the
reference's (the device ILU(0) and products keep its operation order; values cross the ABI
unchanged), GMRES steps within +-2 per nonlinear iteration and the final state within 1e-9 relative
L2 of the reference's (the bar of the device fixed-point driver's test: the GMRES reductions are
summed in another order, and the shock-capturing fixed point carries that difference: 4e-10 on
the st_ebe_not_cau_l8 fixture). With FEMGPU_SOLVER=mixed the solves are the mixed-precision solver's, verified
in FP64 to the same tolerance (parallel ILU(0) with Jacobi sweeps, or with FEMGPU_MIXED_PREC=exact
the FP32 copy of the exact ILU(0)): the state within 1e-6 of the reference's. The harness binary exists
only where the reference tree was present at build time; without it the test is skipped."""
import os
import struct
import subprocess

import numpy as np
import pytest

import golden_cases as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "st_harness_femgpu")


def read_records(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            h = f.read(32)
            if len(h) < 32:
                break
            name = h.split(b"\0", 1)[0].decode()
            kind, = struct.unpack("<i", f.read(4))
            cnt, = struct.unpack("<q", f.read(8))
            dt = np.float64 if kind else np.int32
            out[name] = np.frombuffer(f.read(cnt * np.dtype(dt).itemsize), dtype=dt).copy()
    return out


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.skipif(not os.path.exists(EXE), reason="shim harness not built (oracle/ref.mk needs the reference tree)")
@pytest.mark.parametrize("name,solver", [("st_csr_rcm_lin_r10", "fp64"), ("st_csr_rcm_cau_l10", "fp64"),
                                         ("st_csr_not_lin_r16x9", "fp64"), ("st_ebe_rcm_lin_r12", "fp64"),
                                         ("st_ebe_not_cau_l8", "fp64"), ("st_csr_not_lin_r16x9", "mixed"),
                                         ("st_csr_rcm_lin_l40", "mixed-exact"), ("st_csr_rcm_cau_r24", "mixed-exact")])
def test_reference_driver_through_the_shim(tmp_path, name, solver):
    d = G.load(name)
    inp, outp = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(inp, "wb") as f:
        for k in ("header", "params", "x", "y", "type", "ien", "ubc", "probe"):
            f.write(np.ascontiguousarray(d[k]).tobytes())
    env = dict(os.environ)
    env.pop("FEMGPU_SOLVER", None)
    env.pop("FEMGPU_MIXED_PREC", None)
    if solver.startswith("mixed"):
        env["FEMGPU_SOLVER"] = "mixed"
    if solver == "mixed-exact":
        env["FEMGPU_MIXED_PREC"] = "exact"
    print(f"shim harness: {name} ({solver})", flush=True)
    r = subprocess.run([EXE, inp, outp], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rec = read_records(outp)
    ebe = int(d["header"][2]) == 1
    # the operator probe through csrmv / ebemv of the shim
    if ebe:
        assert np.array_equal(rec["mv"], d["ref_mv"])
    else:
        assert rel(rec["mv"], d["ref_mv"]) < 1e-13
    if solver == "fp64":
        if not ebe:
            assert np.array_equal(rec["MinvF"], d["ref_MinvF"])
        its, ref_its = rec["iters"], d["ref_iters"]
        assert len(its) == len(ref_its), (its, ref_its)
        assert np.all(np.abs(its - ref_its) <= 2), (its, ref_its)
        assert rel(rec["u"], d["ref_u"]) < 1e-9
    else:
        assert rel(rec["u"], d["ref_u"]) < 1e-6, rel(rec["u"], d["ref_u"])
