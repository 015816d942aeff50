"""Writes tests/golden/eigs_<cfg>.json: lowest eigenvalues of the synthetic
configurations, computed with the oracle's operate_H! restatement (C) and
ARPACK (scipy.sparse.linalg.eigsh, tol 1e-13) — an eigen-solver independent of
both the oracle's TRL and the GPU path.  Run: python tests/golden/make_golden.py [mg24 cr48 ...]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "nucleartoolkit.jl_amd"))

import numpy as np  # noqa: E402
from scipy.sparse.linalg import LinearOperator, eigsh  # noqa: E402

import sm_oracle as orc  # noqa: E402
from smgpu import synth  # noqa: E402

CFG = {
    "mg24": ("sd", 4, 4, 1, (-4.0, 0.0), 1.5, 5),
    "cr48": ("pf", 4, 4, 2, (-8.0, 0.0), 1.0, 10),
    "mn50": ("pf", 5, 5, 3, (-8.0, 0.0), 1.0, 5),
}


def main(names):
    for name in names:
        space, zv, nv, soff, spe, sig, nev = CFG[name]
        path = os.path.join("/tmp", f"golden_{name}.snt")
        snt, nuc = synth.config_snt(path, space, zv, nv, synth.BASE_SEED + soff, spe, sig)
        o = orc.OracleHamiltonian(snt, nuc)
        n = [0]

        def mv(x):
            n[0] += 1
            return o.operate_H(np.ascontiguousarray(x, dtype=np.float64).ravel())

        op = LinearOperator((o.mdim, o.mdim), matvec=mv, dtype=np.float64)
        t0 = time.time()
        ev = np.sort(eigsh(op, k=nev, which="SA", tol=1e-13, ncv=max(2 * nev + 1, 40),
                           v0=np.ones(o.mdim) / np.sqrt(o.mdim))[0])
        out = {"config": name, "nucleus": nuc, "space": space, "zv": zv, "nv": nv,
               "seed": synth.BASE_SEED + soff, "mdim": o.mdim, "eigenvalues_MeV": [float(x) for x in ev],
               "method": "oracle operate_H! (C restatement) + ARPACK eigsh tol 1e-13", "matvecs": n[0],
               "seconds": round(time.time() - t0, 1)}
        with open(os.path.join(HERE, f"eigs_{name}.json"), "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out))


if __name__ == "__main__":
    main(sys.argv[1:] or list(CFG))
