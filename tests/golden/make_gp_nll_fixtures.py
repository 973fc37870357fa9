"""Writes gp_nll_fixtures.json: the GP negative log marginal likelihood, its
gradient with respect to (log ell, log sf2, log sn2) and checksums of dNLL/dK,
computed in FP64 numpy from the closed forms with an explicit K^{-1}:

    NLL   = 1/2 y^T K^{-1} y + sum log diag(chol(K)) + N/2 log(2 pi)
    dNLL/dtheta = 1/2 sum_ij (K^{-1} - alpha alpha^T)_ij (dK/dtheta)_ij

dK/dlog ell = Kf * r^2 / ell^2, dK/dlog sf2 = Kf, dK/dlog sn2 = sn2 I, with
Kf the noise-free RBF part. dNLL/dK is reported in lower storage (an
off-diagonal entry counts (i,j) and (j,i): 2 Kbar_ij; diagonal Kbar_ii).
Inputs come from the SplitMix64 generator of the workloads (seeded), so only
the seeds and sizes are stored.

The gradient workload itself (seed 5, N = 16384, d = 8, b = 128, l = 1,
sf2 = 1, sn2 = 1e-2) is computed by `--large` the same way but lean in memory
(r^2 in row slabs, K^{-1} from LAPACK's Cholesky inverse dpotri), and also
with numpy's LU inverse; the two gradients' relative difference is stored as
the reference's own uncertainty ("ref_spread"). It needs about 20 GB and a few
minutes, and updates only that entry of the file.

Run from the repository root: python tests/golden/make_gp_nll_fixtures.py [--large]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
from hs_testlib import synth_gp_numpy  # noqa: E402

# (seed, N, d, matrix_bsz, ell, sf2, sn2)
CASES = [
    (11, 256, 3, 64, 0.7, 1.3, 0.05),
    (12, 1000, 8, 128, 1.0, 1.0, 1e-2),
    (13, 1536, 16, 256, 2.0, 1.0, 1e-3),
    (5, 2048, 8, 128, 1.0, 1.0, 1e-2),
]


def closed_form(seed, N, d, ell, sf2, sn2):
    X, y = synth_gp_numpy(seed, N, d)
    r2 = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    kf = sf2 * np.exp(-0.5 / ell ** 2 * r2)
    K = kf + sn2 * np.eye(N)
    C = np.linalg.cholesky(K)
    Kinv = np.linalg.inv(K)
    alpha = Kinv @ y
    nll = 0.5 * y @ alpha + np.log(np.diag(C)).sum() + 0.5 * N * np.log(2 * np.pi)
    W = Kinv - np.outer(alpha, alpha)
    grad = [0.5 * np.sum(W * kf * r2 / ell ** 2), 0.5 * np.sum(W * kf), 0.5 * np.trace(W) * sn2]
    kbar = 0.5 * W
    low = np.tril(2.0 * kbar) - np.diag(np.diag(kbar))
    return {
        "nll": float(nll),
        "grad": [float(g) for g in grad],
        "kbar_sum": float(low.sum()),
        "kbar_abs_sum": float(np.abs(low).sum()),
        "kbar_diag_sum": float(np.trace(low)),
        "cond": float(np.linalg.cond(K)),
    }


LARGE = (5, 16384, 8, 128, 1.0, 1.0, 1e-2)


def closed_form_large(seed, N, d, ell, sf2, sn2):
    from scipy.linalg import lapack

    X, y = synth_gp_numpy(seed, N, d)
    r2 = np.empty((N, N))
    for i0 in range(0, N, 256):
        r2[i0:i0 + 256] = ((X[i0:i0 + 256, None, :] - X[None, :, :]) ** 2).sum(-1)
    kf = sf2 * np.exp(-0.5 / ell ** 2 * r2)
    K = kf.copy()
    K[np.diag_indices(N)] += sn2
    C, info = lapack.dpotrf(K, lower=1)
    assert info == 0
    logdet_half = np.log(np.diag(C)).sum()
    Kinv, info = lapack.dpotri(C, lower=1)
    assert info == 0
    del C
    Kinv = np.tril(Kinv) + np.tril(Kinv, -1).T

    def grads(Ki):
        alpha = Ki @ y
        nll = 0.5 * y @ alpha + logdet_half + 0.5 * N * np.log(2 * np.pi)
        W = Ki - np.outer(alpha, alpha)
        Wk = W * kf
        g = [0.5 * np.sum(Wk * r2) / ell ** 2, 0.5 * np.sum(Wk), 0.5 * np.trace(W) * sn2]
        del Wk
        low = np.tril(W) - 0.5 * np.diag(np.diag(W))
        return nll, g, low

    nll, g, low = grads(Kinv)
    del Kinv
    nll2, g2, _ = grads(np.linalg.inv(K))
    spread = max(abs(a - c) for a, c in zip(g, g2)) / max(abs(a) for a in g)
    return {
        "nll": float(nll),
        "grad": [float(v) for v in g],
        "kbar_sum": float(low.sum()),
        "kbar_abs_sum": float(np.abs(low).sum()),
        "kbar_diag_sum": float(np.trace(low)),
        "cond": None,
        "ref_spread": float(spread),
        "nll_lu": float(nll2),
        "grad_lu": [float(v) for v in g2],
    }


def main_large():
    path = os.path.join(HERE, "gp_nll_fixtures.json")
    with open(path) as f:
        out = json.load(f)
    seed, N, d, b, ell, sf2, sn2 = LARGE
    rec = {"seed": seed, "N": N, "d": d, "b": b, "ell": ell, "sf2": sf2, "sn2": sn2}
    rec.update(closed_form_large(seed, N, d, ell, sf2, sn2))
    print(rec)
    out = [r for r in out if r["N"] != N] + [rec]
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


def main():
    if "--large" in sys.argv:
        return main_large()
    out = []
    for seed, N, d, b, ell, sf2, sn2 in CASES:
        rec = {"seed": seed, "N": N, "d": d, "b": b, "ell": ell, "sf2": sf2, "sn2": sn2}
        rec.update(closed_form(seed, N, d, ell, sf2, sn2))
        out.append(rec)
        print(rec)
    with open(os.path.join(HERE, "gp_nll_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
