"""Writes tests/golden/cg_iteration_fixtures.json: the oracle's CG iteration
counts on the CG workload (N=32768, d=8, b=128, l=1, sf2=1, sn2=1e-2,
SplitMix64 seed 2, x0 = 0, eps = 1e-10), computed once in the build
container because the oracle needs minutes per run at that size.

Each entry is or_cg_gpu_order, the oracle's restatement of the reference's
share-1.0 GPU CG path (CG.cpp:38-163 with matrixVectorBlock_GPU and the
scalarProduct / sumFinalScalarProduct trees, work-group sizes 128 / 512), once
with every product and sum rounded separately (fma = 0) and once with the
multiply-adds contracted (fma = 1). The matrix is the oracle's own assembly
(or_synth_gp + or_rbf); the tests rebuild it the same way on the host and
upload it, so the GPU sees the identical system.

Usage: python tests/golden/make_cg_iteration_fixtures.py [--quick]
(--quick writes nothing and runs N=4096 only, to check the script).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from hs_testlib import Oracle  # noqa: E402

WORKLOAD = {"N": 32768, "d": 8, "b": 128, "ell": 1.0, "sf2": 1.0, "sn2": 1e-2, "seed": 2, "eps": 1e-10}


def run(o, w):
    N, d, b = w["N"], w["d"], w["b"]
    X, y = o.synth(w["seed"], N, d)
    A = o.rbf(X, N, d, b, w["ell"], w["sf2"], w["sn2"])
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    out = {}
    for fma in (0, 1):
        t = time.time()
        x, it, dn = o.cg_gpu_order(A, rhs, N, b, w["eps"], fma=fma)
        out["fma%d" % fma] = {"iterations": int(it), "delta_final": dn, "x_sum": float(np.sum(x)),
                              "x_norm": float(np.linalg.norm(x)), "seconds": round(time.time() - t, 1)}
        print("fma=%d: %d iterations in %.0f s" % (fma, it, time.time() - t), flush=True)
    return out


def main():
    o = Oracle()
    if "--quick" in sys.argv:
        print(run(o, dict(WORKLOAD, N=4096)))
        return
    res = run(o, WORKLOAD)
    doc = {"generator": "tests/golden/make_cg_iteration_fixtures.py", "oracle": "or_cg_gpu_order",
           "workload": WORKLOAD, "results": res}
    with open(os.path.join(HERE, "cg_iteration_fixtures.json"), "w") as fh:
        json.dump(doc, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
