"""The committed GP NLL fixtures against an independent recomputation: the
smallest case is re-derived with scipy's Cholesky solve and a central
difference of the NLL itself (h = 1e-5 in log space), so a wrong closed-form
term in the generator would show here."""
import json
import math
import os

import numpy as np
import scipy.linalg as sla

from hs_testlib import synth_gp_numpy

HERE = os.path.dirname(os.path.abspath(__file__))


def nll(X, y, ell, sf2, sn2):
    r2 = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    K = sf2 * np.exp(-0.5 / ell ** 2 * r2) + sn2 * np.eye(len(y))
    c, low = sla.cho_factor(K, lower=True)
    a = sla.cho_solve((c, low), y)
    return 0.5 * y @ a + np.log(np.diag(c)).sum() + 0.5 * len(y) * math.log(2 * math.pi)


def test_fixtures_against_fd():
    fx = json.load(open(os.path.join(HERE, "golden", "gp_nll_fixtures.json")))[0]
    X, y = synth_gp_numpy(fx["seed"], fx["N"], fx["d"])
    th = [math.log(fx["ell"]), math.log(fx["sf2"]), math.log(fx["sn2"])]
    f = lambda t: nll(X, y, *(math.exp(v) for v in t))  # noqa: E731
    assert abs(f(th) - fx["nll"]) <= 1e-10 * abs(fx["nll"])
    h = 1e-5
    for i in range(3):
        tp, tm = list(th), list(th)
        tp[i] += h
        tm[i] -= h
        fd = (f(tp) - f(tm)) / (2 * h)
        assert abs(fd - fx["grad"][i]) <= 1e-6 * max(1.0, abs(fx["grad"][i])), (i, fd, fx["grad"][i])
