"""Writes tests/golden/meshes/: mesh files printed by the reference's own structured generator
(Mesh_Generator/Structured_Mesh/meshGenerator.c, compiled in place into oracle/_ref/ by
oracle/ref.mk), for the mesh-file tests. Runs only where that build exists.

    python tests/golden/make_meshes.py
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = os.path.join(HERE, "..", "..", "oracle", "_ref", "meshGenerator")

# (problem, cut, M, N, boundary corner points)
CASES = [
    ("TESTE", "right", 6, 4, [(0, 0), (1, 0), (1, 1), (0, 1)]),
    ("CAVITY", "left", 4, 4, [(0, 0), (1, 0), (1, 1), (0, 1)]),
    ("OBLIQUO", "right", 3, 5, [(0, 0), (1, 0), (1, 1), (0, 1)]),
    ("CONE", "left", 5, 5, [(0, 0), (10, 0), (10, 10), (0, 10)]),
    ("PUDIM", "right", 4, 4, [(-1, -1), (1, -1), (1, 1), (-1, 1)]),
    ("SOD", "right", 6, 2, [(0, 0), (1, 0), (1, 0.2), (0, 0.2)]),
]


def main():
    if not os.path.exists(GEN):
        sys.exit("oracle/_ref/meshGenerator is not built (make -f oracle/ref.mk)")
    out = os.path.join(HERE, "meshes")
    os.makedirs(out, exist_ok=True)
    manifest = []
    for problem, cut, M, N, pts in CASES:
        with tempfile.TemporaryDirectory() as td:
            inp = os.path.join(td, f"{problem}.txt")
            with open(inp, "w") as f:
                f.write(f"{cut}\t:inclination\n{M}\t:M\n{N}\t:N\n1\t:Xcria\n1\t:Ycria\n{len(pts)}\t:nodeCont\n")
                for k, (px, py) in enumerate(pts):
                    f.write(f"{px}\t{py}\t:p{k}\n")
            subprocess.run([GEN, inp, problem], cwd=td, check=True, stdout=subprocess.DEVNULL)
            nn, ne = (M + 1) * (N + 1), 2 * M * N
            name = f"{problem}_{nn}_{ne}.dat"
            shutil.copy(os.path.join(td, name), os.path.join(out, name))
        xs, ys = [p[0] for p in pts], [p[1] for p in pts]
        manifest.append(dict(file=name, problem=problem, cut=cut, M=M, N=N,
                             box=[min(xs), min(ys), max(xs), max(ys)]))
    with open(os.path.join(out, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
