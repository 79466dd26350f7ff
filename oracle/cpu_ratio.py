"""Time the reference's own CPU hot path against the oracle restatement on the same system.

Test infrastructure, run only in a container that holds the reference tree: it needs
oracle/_ref/st_harness_ofast (the reference's unmodified ILUp_precond_setup, ILUp_precond, pgmres,
csrmv and BLAS sources built by oracle/ref.mk with the reference Makefile's -Ofast -march=native)
and the oracle (oracle/liboracle.so). Both solve the steady SUPG transport system of the
benchmark (rotating field, k = 1e-4, u = 1 on the inflow segment, SYMRCM-ordered CSR, ILU(0),
GMRES(30), tol 1e-8) on m x m squares, one rank, one thread. The phases compared are the
preconditioner set-up (ILU(0) factorisation and F := M^-1 F) and pgmres; the harness's assembly
time is not compared (its Dirichlet-value lookup is a linear search written for the fixtures).

    python oracle/cpu_ratio.py [--m 200 400] [--reps 3] [--out profiles/cpu_ratio.json]

Writes a JSON summary: per size, the best-of-reps seconds of each phase on both sides, the
iteration counts (equal when the two agree), and oracle / reference time ratios.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
MPIEXEC = "/opt/conda/bin/mpiexec"

import oracle as O  # noqa: E402
from make_golden import read_records  # noqa: E402


def reference_run(x, y, t, ien, ubc, workdir, exe):
    hd = np.zeros(16, np.int32)
    hd[:8] = [len(x), len(ien) // 3, 0, 1, 0, 30, 1000, 1]
    pr = np.zeros(16)
    pr[:11] = [1e-4, 0.0, 0.0, 0.5, 0.0, -1.0, -0.5, 1.0, 0.0, 1e-8, 1e-6]
    inp, outp = os.path.join(workdir, "in.bin"), os.path.join(workdir, "out.bin")
    with open(inp, "wb") as f:
        for a in (hd, pr, x, y, t, ien, ubc, np.zeros(len(x))):
            f.write(np.ascontiguousarray(a).tobytes())
    subprocess.run([MPIEXEC, "-n", "1", exe, inp, outp], check=True, stdout=subprocess.DEVNULL, timeout=3600)
    rec = read_records(outp)
    return dict(precond_setup=float(rec["timers"][1]), pgmres=float(rec["timers"][2]),
                iterations=int(rec["iters"][0]), u=rec["u"])


def oracle_run(x, y, t, ien, ubc):
    invP0, neq = O.number_unknowns(t)
    lm = O.fill_lm(ien, t, invP0, neq)
    c = O.csr_init(lm, 3, neq, reorder="SYMRCM", type_=t, invP=invP0)
    Ke, F = O.build_transport_supg(x, y, t, c["invP"], ien, ubc, np.zeros(neq), O.make_problem(k=1e-4), neq)
    AA = O.csr_assemble(c["scheme"], Ke, c["nnz"])
    t0 = time.perf_counter()
    LU, Dinv, rc = O.ilu0_factor(c["IA"], c["JA"], AA)
    B = O.ilu0_apply(c["IA"], c["JA"], LU, Dinv, F)
    t1 = time.perf_counter()
    S = O.make_system(0, neq, IA=c["IA"], JA=c["JA"], AA=AA, ilu=(c["IA"], c["JA"], LU, Dinv))
    X, it, rho, cyc = O.pgmres(S, B, kmax=30, lmax=1000, tol=1e-8)
    t2 = time.perf_counter()
    return dict(precond_setup=t1 - t0, pgmres=t2 - t1, iterations=int(it), u=X, n=neq, nnz=int(c["nnz"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--m", type=int, nargs="+", default=[200, 400])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(ROOT, "profiles", "cpu_ratio.json"))
    a = ap.parse_args()
    exe = os.path.join(HERE, "_ref", "st_harness_ofast")
    if not os.path.exists(exe):
        sys.exit(f"{exe} missing: make -f oracle/ref.mk REF=/root/reference")
    O.build()
    cpu = "unknown"
    with open("/proc/cpuinfo") as fh:
        for line in fh:
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    out = {"cpu_model": cpu, "nproc": os.cpu_count(), "threads_used": 1, "platform": platform.platform(),
           "reference_build": "gcc -Ofast -march=native (the reference Makefile's flags), mpiexec -n 1",
           "oracle_build": "gcc -O2 -ffp-contract=off -fno-fast-math", "sizes": []}
    for m in a.m:
        x, y, ien = O.mesh_structured(m, m, right=True)
        t = O.unit_square_types(x, y)
        ubc = O.inflow_ubc(x, y)
        best_r, best_o = None, None
        with tempfile.TemporaryDirectory() as wd:
            for _ in range(a.reps):
                r = reference_run(x, y, t, ien, ubc, wd, exe)
                o = oracle_run(x, y, t, ien, ubc)
                if best_r is None or r["precond_setup"] + r["pgmres"] < best_r["precond_setup"] + best_r["pgmres"]:
                    best_r = r
                if best_o is None or o["precond_setup"] + o["pgmres"] < best_o["precond_setup"] + best_o["pgmres"]:
                    best_o = o
        diff = float(np.linalg.norm(best_o["u"] - best_r["u"]) / np.linalg.norm(best_r["u"]))
        row = {"m": m, "n": best_o["n"], "nnz": best_o["nnz"],
               "reference_s": {k: round(best_r[k], 5) for k in ("precond_setup", "pgmres")},
               "oracle_s": {k: round(best_o[k], 5) for k in ("precond_setup", "pgmres")},
               "iterations": {"reference": best_r["iterations"], "oracle": best_o["iterations"]},
               "solution_rel_diff": diff,
               "ratio_oracle_over_reference": {
                   k: round(best_o[k] / best_r[k], 4) for k in ("precond_setup", "pgmres")}}
        tot_r = best_r["precond_setup"] + best_r["pgmres"]
        tot_o = best_o["precond_setup"] + best_o["pgmres"]
        row["ratio_oracle_over_reference"]["total"] = round(tot_o / tot_r, 4)
        out["sizes"].append(row)
        print(json.dumps(row), flush=True)
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
