"""Pins of the oracle's restatement of the reference's share-1.0 GPU CG path
(or_dot_tree, or_symv_ordered, or_cg_gpu_order; CPU only).

The dot product's order is fixed by scalarProduct / sumFinalScalarProduct
(VectorOperations.cpp:84-185): a work-group of wg items covers 2 wg
elements, item l adds x[l] y[l] then x[l + wg] y[l + wg], then the halving
tree cache[l] += cache[l + s] for s = wg/2 .. 1; the partials go through the
same scheme with 512 items. The hand-derived cases below are built on
2^53, where adding 1 rounds (ties to even) and adding an even number is
exact, so each expected value follows from the order alone; a sequential
sum, an exactly rounded sum or another work-group size gives a different
number (asserted too).
"""
import math

import numpy as np
import pytest

from backends import OracleBackend
from golden_cases import case_id, check_case, kernel_cases, run_case
from hs_testlib import dense_to_packed, dptr, rbf_dense_numpy, synth_gp_numpy

TWO53 = 2.0 ** 53


def _ones_with(first, length, extra=None):
    v = np.ones(length)
    v[0] = first
    for i, val in (extra or {}).items():
        v[i] = val
    return v


@pytest.mark.parametrize("fma", [0, 1])
def test_dot_tree_pair_then_halving_order(oracle, fma):
    # one work-group of 128 items over 256 elements, y = 1: item 0 adds
    # x[0] = 2^53 and x[128] = 1 -> 2^53 + 1 rounds to 2^53 (ties to even);
    # the 127 other items hold 2; the tree adds those to 2^53 in even steps,
    # which are exact: 2^53 + 254.
    x = _ones_with(TWO53, 256)
    y = np.ones(256)
    got = oracle.dot_tree(x, y, 128, fma=fma)
    assert got == TWO53 + 254
    # sequential summation loses every 1: 2^53; exact rounding: 2^53 + 256
    assert math.fsum(x) == TWO53 + 256
    seq = 0.0
    for v in x:
        seq += v
    assert seq == TWO53 and got not in (seq, math.fsum(x))


def test_dot_tree_work_group_size_is_visible(oracle):
    # x[128] = 3. With wg = 128 it pairs with x[0] = 2^53: 2^53 + 3 is a tie
    # between 2^53 + 2 and 2^53 + 4 and rounds to the even 2^53 + 4; the
    # other 127 items hold 2: 2^53 + 4 + 254 = 2^53 + 258.
    x = _ones_with(TWO53, 256, {128: 3.0})
    y = np.ones(256)
    assert oracle.dot_tree(x, y, 128, wg=128) == TWO53 + 258
    # wg = 64: two groups of 128 elements. Group 0 pairs x[0] with x[64] = 1
    # (2^53 + 1 rounds to 2^53) and totals 2^53 + 2 * 63; group 1 pairs
    # x[128] = 3 with x[192] = 1 and totals 4 + 2 * 63 = 130; the final tree
    # adds them exactly: 2^53 + 256.
    assert oracle.dot_tree(x, y, 128, wg=64) == TWO53 + 256


def test_dot_tree_final_pass_over_partials(oracle):
    # two work-groups of 128 items (512 elements): group 0 gives 2^53 + 254
    # as above, group 1 gives 256; the 512-item final tree adds them exactly.
    x = _ones_with(TWO53, 512)
    y = np.ones(512)
    assert oracle.dot_tree(x, y, 128) == TWO53 + 510


def test_dot_tree_fma_contracts_the_second_pair_term(oracle):
    # item 0: c = x0 y0 = -(1 + 2^-51), exact; then c += x128 y128 with
    # x128 y128 = (1 + 2^-52)^2 = 1 + 2^-51 + 2^-104. Unfused, the product
    # rounds to 1 + 2^-51 first and the sum is 0; fused, the sum is the exact
    # 2^-104. Every other term is zero.
    e = 2.0 ** -52
    x = np.zeros(256)
    y = np.zeros(256)
    x[0], y[0] = -(1.0 + 2 * e), 1.0
    x[128] = y[128] = 1.0 + e
    assert oracle.dot_tree(x, y, 128, fma=0) == 0.0
    assert oracle.dot_tree(x, y, 128, fma=1) == 2.0 ** -104


@pytest.mark.parametrize("n", [1, 7, 64, 300])
def test_dot_tree_exact_on_integers(oracle, n):
    # every partial sum of small integers is exact: any order gives the sum
    rng = np.random.default_rng(n)
    b = 128
    x = rng.integers(-1000, 1000, n * b).astype(np.float64)
    y = rng.integers(-1000, 1000, n * b).astype(np.float64)
    exact = int(np.dot(x.astype(np.int64), y.astype(np.int64)))
    assert oracle.dot_tree(x, y, b) == exact
    assert oracle.dot_tree(x, y, b, fma=1) == exact


def test_dot_tree_rejects_too_many_partials(oracle):
    # 2 * 512 partials of 256 elements at most: one more element-pair group
    # makes the reference throw (VectorOperations.cpp:145-147)
    b = 128
    ok = np.ones(1024 * 256)
    assert oracle.dot_tree(ok, ok, b) == 1024 * 256
    bad = np.ones(1024 * 256 + 2 * b)
    assert math.isnan(oracle.dot_tree(bad, bad, b))


@pytest.mark.parametrize("case", [c for c in kernel_cases()
                                  if any(f[0] == "VectorOperations::scalarProduct" for f in c["calls"])],
                         ids=case_id)
@pytest.mark.parametrize("fma", [0, 1])
def test_dot_tree_reference_goldens(case, oracle, fma):
    final, env = run_case(case, OracleBackend(oracle, fma=fma))
    assert not check_case(case, final, env)


@pytest.mark.parametrize("N,b", [(200, 16), (20, 6), (300, 64), (257, 32)])
def test_symv_ordered_is_or_symv_bit_for_bit(oracle, N, b):
    """fma = 0: the block-row loop nest performs each row's sums in
    matrixVectorBlock_GPU's order, which or_symv (pinned by the reference's
    mat-vec goldens) performs row by row."""
    rng = np.random.default_rng(N + b)
    M = rng.standard_normal((N, N))
    K = M + M.T
    A = dense_to_packed(K, b)
    n = -(-N // b)
    x = np.zeros(n * b)
    x[:N] = rng.standard_normal(N)
    assert np.array_equal(oracle.symv_ordered(A, x, b, n, fma=0), oracle.symv(A, x, b, n))
    # fma = 1 stays within the rounding bound of the same product
    yf = oracle.symv_ordered(A, x, b, n, fma=1)
    ref = np.zeros(n * b)
    ref[:N] = K @ x[:N]
    bound = 2 * N * np.finfo(float).eps * (np.abs(K) @ np.abs(x[:N]))
    assert np.all(np.abs(yf[:N] - ref[:N]) <= bound)
    assert np.all(yf[N:] == 0.0)


def _composed_cg(oracle, A, rhs, N, b, eps, wg=128, wgf=512):
    """The reference's share-1.0 step sequence (CG.cpp:38-163) composed from
    the oracle's pinned primitives (or_symv, or_dot_tree, or_sub,
    or_scale_add), fma = 0."""
    n = -(-N // b)
    x = np.zeros(n * b)
    r = np.zeros(n * b)
    q = np.zeros(n * b)
    oracle.or_symv(dptr(A), dptr(x), dptr(r), b, 0, 0, n, n, n, 1)
    oracle.or_sub(dptr(rhs), dptr(r), dptr(r), b, 0, n)
    d = r.copy()
    dn = oracle.or_dot_tree(dptr(r), dptr(r), b, 0, n, wg, wgf, 0)
    d0 = dn
    it = 0
    while it < 100000 and dn > eps * eps * d0:
        oracle.or_symv(dptr(A), dptr(d), dptr(q), b, 0, 0, n, n, n, 1)
        alpha = dn / oracle.or_dot_tree(dptr(d), dptr(q), b, 0, n, wg, wgf, 0)
        oracle.or_scale_add(dptr(x), alpha, dptr(d), dptr(x), b, 0, n)
        if it % 50 == 0:
            oracle.or_symv(dptr(A), dptr(x), dptr(r), b, 0, 0, n, n, n, 1)
            oracle.or_sub(dptr(rhs), dptr(r), dptr(r), b, 0, n)
        else:
            oracle.or_scale_add(dptr(r), -1.0 * alpha, dptr(q), dptr(r), b, 0, n)
        do = dn
        dn = oracle.or_dot_tree(dptr(r), dptr(r), b, 0, n, wg, wgf, 0)
        oracle.or_scale_add(dptr(r), dn / do, dptr(d), dptr(d), b, 0, n)
        it += 1
    return x, it, dn


@pytest.mark.parametrize("N,b,sn2", [(256, 64, 1e-2), (300, 32, 1e-1), (130, 16, 1e-3)])
def test_cg_gpu_order_is_the_composed_step_sequence(oracle, N, b, sn2):
    X, y = synth_gp_numpy(7, N, 4)
    A = dense_to_packed(rbf_dense_numpy(X, 1.0, 1.0, sn2), b)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    x, it, dn = oracle.cg_gpu_order(A, rhs, N, b, 1e-10, fma=0)
    xc, itc, dnc = _composed_cg(oracle, A, rhs, N, b, 1e-10)
    assert it == itc and dn == dnc and np.array_equal(x, xc)


@pytest.mark.parametrize("fma", [0, 1])
@pytest.mark.parametrize("N,b,eps,sn2,slack", [(64, 16, 1e-10, 1.0, 1), (256, 64, 1e-10, 1.0, 1),
                                                (256, 64, 1e-10, 1e-1, 5)])
def test_cg_gpu_order_matches_textbook_and_solve(oracle, fma, N, b, eps, sn2, slack):
    from test_oracle_pins import _textbook_cg

    X, y = synth_gp_numpy(11, N, 4)
    K = rbf_dense_numpy(X, 1.0, 1.0, sn2)
    A = dense_to_packed(K, b)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    x, it, _ = oracle.cg_gpu_order(A, rhs, N, b, eps, fma=fma)
    _, itt = _textbook_cg(K, y, eps)
    assert abs(it - itt) <= slack
    assert np.linalg.norm(K @ x[:N] - y) <= 2 * eps * np.linalg.norm(y)
    ref = np.linalg.solve(K, y)
    assert np.allclose(x[:N], ref, rtol=0, atol=1e-6 * np.abs(ref).max())
    assert np.all(x[N:] == 0.0)


def test_cg_gpu_order_zero_rhs_and_imax(oracle):
    N, b = 128, 32
    X, y = synth_gp_numpy(3, N, 4)
    A = dense_to_packed(rbf_dense_numpy(X, 1.0, 1.0, 1e-2), b)
    x, it, _ = oracle.cg_gpu_order(A, np.zeros(N), N, b, 1e-10)
    assert it == 0 and np.all(x == 0.0)
    x, it, _ = oracle.cg_gpu_order(A, y.copy(), N, b, 1e-14, imax=5)
    assert it == 5
