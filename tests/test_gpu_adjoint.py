"""GP NLL gradient by the hand-written reverse pass (hs_gp_nll_grad_f64):
against the FP64 closed form with an explicit K^{-1} (committed fixtures,
tests/golden/make_gp_nll_fixtures.py), against dense numpy dNLL/dK, against
central finite differences (h = 1e-5 in log space) and an FP64 CPU reference
gradient at the 16384-point workload, and through torch.autograd
(gradcheck)."""
import json
import math
import os

import numpy as np
import pytest

from hs_testlib import packed_to_dense, synth_gp_numpy

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "gp_nll_fixtures.json")))


def _inputs(seed, N, d):
    import torch
    X, y = synth_gp_numpy(seed, N, d)
    return (torch.tensor(X, dtype=torch.float64, device="cuda"),
            torch.tensor(y, dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("fx", FIXTURES, ids=lambda f: "N%d_b%d_d%d" % (f["N"], f["b"], f["d"]))
def test_nll_grad_matches_closed_form(hs, dev, fx):
    import gp_autograd as ga
    print("fixture", {k: fx[k] for k in ("seed", "N", "d", "b", "ell", "sf2", "sn2")}, flush=True)
    X, y = _inputs(fx["seed"], fx["N"], fx["d"])
    theta = [math.log(fx["ell"]), math.log(fx["sf2"]), math.log(fx["sn2"])]
    nll, g, Kb = ga.nll_grad(theta, X, y, fx["b"], want_kbar=True)
    print("  nll %.15g (ref %.15g)  grad %s (ref %s)" % (nll, fx["nll"], g, fx["grad"]), flush=True)
    gref = np.array(fx["grad"])
    if fx["cond"] is None:
        # the gradient workload (N = 16384): the reference's own spread (two
        # independent inverses) is ~1e-14, so the bar is the 1e-9 relative
        # agreement of an FP64 reverse pass with an FP64 CPU reference
        rel = np.max(np.abs(np.array(g) - gref)) / np.max(np.abs(gref))
        print("  max relative gradient difference %.3e (reference spread %.1e)" % (rel, fx["ref_spread"]))
        assert rel <= 1e-9
        assert abs(nll - fx["nll"]) <= 1e-10 * abs(fx["nll"])
        tol = 1e-10
    else:
        # FP64 agreement is limited by cond(K) * eps on K^{-1}
        tol = max(1e-11, 4e-16 * fx["cond"])
        assert abs(nll - fx["nll"]) <= tol * max(1.0, abs(fx["nll"]))
        # the gradient is a sum of N^2 terms of size up to |Kbar|: its error
        # scales with sum |Kbar|, not with the (cancelled) result
        assert np.max(np.abs(np.array(g) - gref)) <= 1e-12 * fx["kbar_abs_sum"] + 1e-9 * np.max(np.abs(gref))
    low = packed_to_dense(Kb.cpu().numpy(), fx["N"], fx["b"], lower_only=True)
    assert abs(low.sum() - fx["kbar_sum"]) <= tol * fx["kbar_abs_sum"]
    assert abs(np.abs(low).sum() - fx["kbar_abs_sum"]) <= tol * fx["kbar_abs_sum"]
    assert abs(np.trace(low) - fx["kbar_diag_sum"]) <= tol * fx["kbar_abs_sum"]


@pytest.mark.parametrize("N,b,d", [(300, 64, 3), (520, 128, 5)])
def test_kbar_dense(hs, dev, N, b, d):
    """dNLL/dK entry by entry, padding included (N % b != 0)."""
    import gp_autograd as ga
    print("kbar dense N=%d b=%d d=%d" % (N, b, d), flush=True)
    ell, sf2, sn2 = 0.8, 1.1, 0.03
    Xn, yn = synth_gp_numpy(21, N, d)
    X, y = _inputs(21, N, d)
    _, _, Kb = ga.nll_grad([math.log(ell), math.log(sf2), math.log(sn2)], X, y, b, want_kbar=True)
    r2 = ((Xn[:, None, :] - Xn[None, :, :]) ** 2).sum(-1)
    K = sf2 * np.exp(-0.5 / ell ** 2 * r2) + sn2 * np.eye(N)
    Kinv = np.linalg.inv(K)
    a = Kinv @ yn
    kbar = 0.5 * (Kinv - np.outer(a, a))
    expect = np.tril(2 * kbar) - np.diag(np.diag(kbar))
    got = packed_to_dense(Kb.cpu().numpy(), N, b, lower_only=True)
    err = np.max(np.abs(got - expect))
    print("  max |Kbar - ref| = %.3e (scale %.3e, cond %.2e)" % (err, np.abs(expect).max(), np.linalg.cond(K)))
    assert err <= 1e-10 * np.abs(expect).max() * max(1.0, np.linalg.cond(K) / 1e3)
    # padding entries of the packed adjoint are exactly zero
    n = -(-N // b)
    full = Kb.cpu().numpy()
    dense_pad = packed_to_dense(full, n * b, b, lower_only=True)
    assert np.all(dense_pad[N:, :] == 0.0) and np.all(dense_pad[:, N:] == 0.0)


def test_fd_at_16384(hs, dev):
    """Central differences, h = 1e-5 in log space, on the 16384-point
    workload (d = 8, matrix_bsz = 128, ell = 1, sf2 = 1, sn2 = 1e-2)."""
    import gp_autograd as ga
    seed, N, d, b = 5, 16384, 8, 128
    print("fd N=%d" % N, flush=True)
    X, y = _inputs(seed, N, d)
    theta = [0.0, 0.0, math.log(1e-2)]
    nll, g, _ = ga.nll_grad(theta, X, y, b)
    h = 1e-5
    fd = []
    for i in range(3):
        tp = list(theta)
        tm = list(theta)
        tp[i] += h
        tm[i] -= h
        fd.append((ga.nll_only(tp, X, y, b) - ga.nll_only(tm, X, y, b)) / (2 * h))
    g, fd = np.array(g), np.array(fd)
    rel = np.max(np.abs(g - fd)) / np.max(np.abs(g))
    print("  nll %.12g grad %s fd %s max rel diff %.2e" % (nll, g, fd, rel))
    assert rel <= 1e-6


def test_torch_autograd(hs, dev):
    import torch

    import gp_autograd as ga
    fx = FIXTURES[0]
    X, y = _inputs(fx["seed"], fx["N"], fx["d"])
    theta = torch.tensor([math.log(fx["ell"]), math.log(fx["sf2"]), math.log(fx["sn2"])], dtype=torch.float64,
                         requires_grad=True)
    loss = ga.GPNLL.apply(theta, X, y, fx["b"])
    loss.backward()
    assert np.allclose(theta.grad.numpy(), fx["grad"], rtol=1e-9, atol=1e-9)
    # gradcheck differentiates numerically through the same forward
    t2 = theta.detach().clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: ga.GPNLL.apply(t, X, y, fx["b"]), (t2,), eps=1e-5, atol=1e-5,
                                    rtol=1e-6)


def test_not_positive_definite(hs, dev):
    import gp_autograd as ga
    X, y = _inputs(3, 256, 2)
    # duplicate rows and no noise make K singular
    X[1] = X[0]
    with pytest.raises(RuntimeError):
        ga.nll_grad([0.0, 0.0, math.log(1e-300)], X, y, 64)
