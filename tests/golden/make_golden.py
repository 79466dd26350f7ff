"""Write the golden fixtures of tests/golden from the reference's own code.

Runs only in a container that holds the reference tree (default /root/reference): it builds
oracle/_ref/ with oracle/ref.mk (the reference's structured mesh generator and its steady
transport hot-path sources linked with oracle/ref_harness_st.c), runs them on small cases and
stores inputs and outputs as compressed .npz files plus manifest.json (SHA-256 per file).

    python tests/golden/make_golden.py [--ref /root/reference]

The fixtures are data only: meshes, CSR structures, permutations, element matrices, ILU
factors, GMRES iteration counts and solutions produced by the reference code.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import shutil
import struct
import subprocess
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFBIN = os.path.join(ROOT, "oracle", "_ref")
MPIEXEC = "/opt/conda/bin/mpiexec"

ROT = dict(bx=(0.5, 0.0, -1.0), by=(-0.5, 1.0, 0.0))   # (-y+0.5, x-0.5)


def uniform(theta):
    return dict(bx=(float(np.cos(theta)), 0.0, 0.0), by=(float(np.sin(theta)), 0.0, 0.0))


CASES = [
    dict(name="st_csr_rcm_lin_r10", M=10, N=10, cut="right", ebe=0, reorder=1, shock=0, k=1e-4,
         vel=ROT, bc="inflow", kmax=30, lmax=400, nl=1),
    dict(name="st_csr_rcm_cau_l10", M=10, N=10, cut="left", ebe=0, reorder=1, shock=1, k=1e-4,
         vel=ROT, bc="inflow", kmax=30, lmax=400, nl=5),
    dict(name="st_csr_not_lin_r16x9", M=16, N=9, cut="right", ebe=0, reorder=0, shock=0, k=1e-2,
         vel=uniform(0.7), bc="inflow", kmax=10, lmax=400, nl=1),
    dict(name="st_ebe_rcm_lin_r12", M=12, N=12, cut="right", ebe=1, reorder=1, shock=0, k=1e-2,
         vel=ROT, bc="inflow", kmax=30, lmax=60, nl=1),
    dict(name="st_ebe_not_cau_l8", M=8, N=8, cut="left", ebe=1, reorder=0, shock=1, k=1e-3,
         vel=ROT, bc="inflow", kmax=20, lmax=60, nl=3),
    dict(name="st_csr_rcm_teste_r8", M=8, N=8, cut="right", ebe=0, reorder=1, shock=0, k=1.0,
         vel=dict(bx=(0.0, 0.0, 0.0), by=(0.0, 0.0, 0.0)), bc="ten", kmax=30, lmax=100, nl=1),
    dict(name="st_csr_rcm_cau_r24", M=24, N=24, cut="right", ebe=0, reorder=1, shock=1, k=1e-4,
         vel=ROT, bc="inflow", kmax=30, lmax=400, nl=3),
    dict(name="st_csr_rcm_lin_l40", M=40, N=40, cut="left", ebe=0, reorder=1, shock=0, k=1e-4,
         vel=ROT, bc="inflow", kmax=30, lmax=400, nl=1),
]


def run_mesh_generator(M, N, cut, workdir):
    """The reference generator on the unit square (problem TESTE marks: boundary = mark 15)."""
    txt = os.path.join(workdir, "mesh.txt")
    with open(txt, "w") as f:
        f.write(f"{cut}\t:inclination\n{M}\t:M\n{N}\t:N\n0\t:Xcria\n0\t:Ycria\n4\t:nodeCont\n")
        f.write("0.0\t0.0\t:p1\n1.0\t0.0\t:p2\n1.0\t1.0\t:p3\n0.0\t1.0\t:p4\n")
    subprocess.run([os.path.join(REFBIN, "meshGenerator"), txt, "TESTE"], cwd=workdir, check=True,
                   stdout=subprocess.DEVNULL)
    nn, ne = (M + 1) * (N + 1), 2 * M * N
    tok = open(os.path.join(workdir, f"TESTE_{nn}_{ne}.dat")).read().split()
    assert int(tok[0]) == nn
    node = np.array(tok[1:1 + 3 * nn]).reshape(nn, 3)
    x, y, t = node[:, 0].astype(float), node[:, 1].astype(float), node[:, 2].astype(np.int32)
    rest = tok[1 + 3 * nn:]
    assert int(rest[0]) == ne
    ien = np.array(rest[1:1 + 3 * ne], dtype=np.int32)
    return x, y, t, ien


NS_CASES = [
    dict(name="ns_ebe_r8", M=8, N=8, cut="right", ebe=1, reorder=0, Re=1000.0),
    dict(name="ns_csr_rcm_l10", M=10, N=10, cut="left", ebe=0, reorder=1, Re=100.0),
    dict(name="ns_csr_not_r6x5", M=6, N=5, cut="right", ebe=0, reorder=0, Re=1.0),
]


EU_CASES = [
    dict(name="eu_ebe_dimless_cau_r8", M=8, N=8, cut="right", ebe=1, reorder=0, dimless=1, shock=1, kmax=50,
         lmax=40, tol=1e-10),
    dict(name="eu_csr_rcm_dim_cau_l8", M=8, N=8, cut="left", ebe=0, reorder=1, dimless=0, shock=1),
    dict(name="eu_csr_not_dimless_none_r6x5", M=6, N=5, cut="right", ebe=0, reorder=0, dimless=1, shock=0),
]


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


def run_case(c, workdir):
    x, y, t, ien = run_mesh_generator(c["M"], c["N"], c["cut"], workdir)
    if c["bc"] == "ten":
        ubc = np.full(len(x), 10.0)
    else:
        ubc = np.where((np.abs(x) <= 1e-15) & (y > 0.25) & (y < 0.5), 1.0, 0.0)
    rng = np.random.default_rng(zlib.crc32(c["name"].encode()))
    probe = rng.standard_normal(len(x))
    hd = np.zeros(16, np.int32)
    hd[:8] = [len(x), len(ien) // 3, c["ebe"], c["reorder"], c["shock"], c["kmax"], c["lmax"], c["nl"]]
    pr = np.zeros(16)
    pr[:11] = [c["k"], 0.0, 0.0, *c["vel"]["bx"], *c["vel"]["by"], 1e-8, 1e-6]
    inp, outp = os.path.join(workdir, "in.bin"), os.path.join(workdir, "out.bin")
    with open(inp, "wb") as f:
        for a in (hd, pr, x, y, t, ien, ubc, probe):
            f.write(np.ascontiguousarray(a).tobytes())
    subprocess.run([MPIEXEC, "-n", "1", os.path.join(REFBIN, "st_harness"), inp, outp], check=True,
                   stdout=subprocess.DEVNULL, timeout=600)
    rec = read_records(outp)
    data = dict(x=x, y=y, type=t, ien=ien, ubc=ubc, probe=probe, header=hd, params=pr)
    data.update({"ref_" + k: v for k, v in rec.items()})
    return data


def run_ns_case(c, workdir):
    """Navier-Stokes assembly on a cavity-like problem: velocity prescribed on the boundary (lid
    u = 1 at y = 1), pressure pinned at the first interior node, random smooth sources and state."""
    x, y, t1, ien = run_mesh_generator(c["M"], c["N"], c["cut"], workdir)
    nn = len(x)
    bnd = t1 == 0
    type3 = np.ones((nn, 3), np.int32)
    type3[bnd, 0] = 0
    type3[bnd, 1] = 0
    type3[np.flatnonzero(~bnd)[0], 2] = 0
    presc = np.zeros((nn, 3))
    presc[bnd & (np.abs(y - 1.0) <= 1e-15), 0] = 1.0
    rng = np.random.default_rng(zlib.crc32(c["name"].encode()))
    fsrc = np.stack([np.sin(3 * x + y), np.cos(2 * y - x)], 1) * 0.1
    state = rng.uniform(-1, 1, (nn, 3))
    hd = np.zeros(16, np.int32)
    hd[:4] = [nn, len(ien) // 3, c["ebe"], c["reorder"]]
    pr = np.zeros(16)
    pr[0] = c["Re"]
    inp, outp = os.path.join(workdir, "in.bin"), os.path.join(workdir, "out.bin")
    with open(inp, "wb") as f:
        for a in (hd, pr, x, y, type3, ien, presc, fsrc, state):
            f.write(np.ascontiguousarray(a).tobytes())
    subprocess.run([MPIEXEC, "-n", "1", os.path.join(REFBIN, "ns_harness"), inp, outp], check=True,
                   stdout=subprocess.DEVNULL, timeout=600)
    rec = read_records(outp)
    data = dict(x=x, y=y, type3=type3.ravel(), ien=ien, presc=presc.ravel(), fsrc=fsrc.ravel(),
                state=state.ravel(), header=hd, params=pr)
    data.update({"ref_" + k: v for k, v in rec.items()})
    return data


def run_eu_case(c, workdir):
    """Euler assembly on a supersonic state: conservative variables of rho ~ 1, velocity ~ (2, 0.2)
    and p = 1/1.4 perturbed per node, all four dofs prescribed on the inflow boundary x = 0."""
    x, y, t1, ien = run_mesh_generator(c["M"], c["N"], c["cut"], workdir)
    nn = len(x)
    gamma = 1.4
    rng = np.random.default_rng(zlib.crc32(c["name"].encode()))
    rho = 1.0 + 0.2 * rng.uniform(-1, 1, nn)
    vx = 2.0 + 0.3 * rng.uniform(-1, 1, nn)
    vy = 0.2 + 0.3 * rng.uniform(-1, 1, nn)
    pres = (1.0 / gamma) * (1.0 + 0.3 * rng.uniform(-1, 1, nn))
    Et = pres / ((gamma - 1.0) * rho) + 0.5 * (vx * vx + vy * vy)
    state = np.stack([rho, rho * vx, rho * vy, rho * Et], 1)
    incr = 1e-2 * rng.standard_normal((nn, 4))
    inflow = np.abs(x) <= 1e-15
    type4 = np.ones((nn, 4), np.int32)
    type4[inflow] = 0
    presc = np.where(inflow[:, None], state, 0.0)
    hd = np.zeros(16, np.int32)
    hd[:8] = [nn, len(ien) // 3, c["ebe"], c["reorder"], c["dimless"], c["shock"], c.get("kmax", 0), c.get("lmax", 0)]
    pr = np.zeros(16)
    pr[:6] = [gamma, 2.0, 0.5, 1e-2, 1e-8, c.get("tol", 0.0)]
    inp, outp = os.path.join(workdir, "in.bin"), os.path.join(workdir, "out.bin")
    with open(inp, "wb") as f:
        for a in (hd, pr, x, y, type4, ien, presc, state, incr):
            f.write(np.ascontiguousarray(a).tobytes())
    subprocess.run([MPIEXEC, "-n", "1", os.path.join(REFBIN, "eu_harness"), inp, outp], check=True,
                   stdout=subprocess.DEVNULL, timeout=600)
    rec = read_records(outp)
    data = dict(x=x, y=y, type4=type4.ravel(), ien=ien, presc=presc.ravel(), state=state.ravel(),
                incr=incr.ravel(), header=hd, params=pr)
    data.update({"ref_" + k: v for k, v in rec.items()})
    return data


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", nargs="*", default=None, help="regenerate only these fixtures (keep the others)")
    a = ap.parse_args()
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref.mk"), f"REF={a.ref}"],
                   check=True)
    mpath = os.path.join(HERE, "manifest.json")
    manifest = json.load(open(mpath)) if a.only and os.path.exists(mpath) else {}
    keep = (lambda c: c["name"] in a.only) if a.only else (lambda c: True)
    work = tempfile.mkdtemp()
    try:
        for c in filter(keep, CASES):
            d = run_case(c, work)
            fn = os.path.join(HERE, c["name"] + ".npz")
            np.savez_compressed(fn, **d)
            meta = {k: v for k, v in c.items() if k != "name"}
            meta["sha256"] = hashlib.sha256(open(fn, "rb").read()).hexdigest()
            meta["iters"] = d["ref_iters"].tolist()
            manifest[c["name"]] = meta
            print(c["name"], "iters", meta["iters"], "bytes", os.path.getsize(fn))
        for c in filter(keep, EU_CASES):
            d = run_eu_case(c, work)
            fn = os.path.join(HERE, c["name"] + ".npz")
            np.savez_compressed(fn, **d)
            meta = {k: v for k, v in c.items() if k != "name"}
            meta["sha256"] = hashlib.sha256(open(fn, "rb").read()).hexdigest()
            manifest[c["name"]] = meta
            print(c["name"], "bytes", os.path.getsize(fn))
        for c in filter(keep, NS_CASES):
            d = run_ns_case(c, work)
            fn = os.path.join(HERE, c["name"] + ".npz")
            np.savez_compressed(fn, **d)
            meta = {k: v for k, v in c.items() if k != "name"}
            meta["sha256"] = hashlib.sha256(open(fn, "rb").read()).hexdigest()
            manifest[c["name"]] = meta
            print(c["name"], "bytes", os.path.getsize(fn))
    finally:
        shutil.rmtree(work)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
