"""Loader for the reference-generated fixtures in tests/golden (see make_golden.py there)."""
import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(prefix="st_"):
    """Fixture names: steady transport ("st_", the default) or Navier-Stokes ("ns_")."""
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    meta = json.load(open(os.path.join(GOLDEN, "manifest.json")))[name]
    d["meta"] = meta
    return d


def problem_args(d):
    """(k, bx, by, shock) of a fixture, from the parameter block written for the harness."""
    pr = d["params"]
    return dict(k=float(pr[0]), bx=tuple(pr[3:6]), by=tuple(pr[6:9]), shock=int(d["header"][4]))
