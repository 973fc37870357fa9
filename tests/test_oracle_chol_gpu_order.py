"""Pins of the oracle's restatement of the reference's share-1.0 GPU blocked
Cholesky (or_potrf_gpu_order, or_trsm_gpu_order, or_syrk_gpu_order,
or_gemm_gpu_order, or_cholesky_gpu_order; CPU only).

The GPU kernels apply each element's terms in the same sequence as the CPU
kernels that the reference goldens pin, except the TRSM, which multiplies by
a rounded reciprocal of the diagonal (triangularSolve_optimizedGPU,
MatrixMatrixOperations.cpp:50-107) where the CPU kernel divides. fma = 1
contracts each `v - x y` / `s + x y` statement. The expected values below
come from the goldens, from the pinned CPU restatement (bit for bit at
fma = 0), and from exactly rounded arithmetic (fractions.Fraction, whose
float conversion rounds to nearest even) on inputs where the rounding of one
statement decides the result.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from backends import OracleGpuOrderBackend
from golden_cases import case_id, check_case, factor_cases, run_case
from hs_testlib import dense_to_packed, dptr, nblocks, packed_to_dense, rbf_dense_numpy, synth_gp_numpy


def _round(q):
    """Nearest double of an exact rational (ties to even)."""
    return float(q)


FACTOR_CASES = factor_cases()


def test_factor_golden_cases_present():
    names = {c[0] for t in FACTOR_CASES for c in t["calls"]}
    assert "MatrixMatrixOperations::triangularSolve_optimizedGPU" in names
    assert "MatrixOperations::cholesky_optimizedGPU" in names
    assert len(FACTOR_CASES) >= 40


@pytest.mark.parametrize("fma", [0, 1])
@pytest.mark.parametrize("case", FACTOR_CASES, ids=case_id)
def test_gpu_order_kernels_match_reference_goldens(oracle, case, fma):
    final, env = run_case(case, OracleGpuOrderBackend(oracle, fma))
    fails = check_case(case, final, env)
    assert not fails, fails


def _spd_packed(N, b, seed, d=4, ell=1.0, sn2=0.1):
    X, _ = synth_gp_numpy(seed, N, d)
    return dense_to_packed(rbf_dense_numpy(X, ell, 1.0, sn2), b), X


@pytest.mark.parametrize("N,b", [(20, 4), (20, 6), (37, 8), (64, 16)])
def test_fma0_potrf_syrk_gemm_are_the_cpu_order(oracle, N, b):
    """At fma = 0 the GPU POTRF, SYRK and GEMM orders are the CPU kernels'
    orders (or_potrf, or_syrk, or_gemm): equal bit for bit after each op of
    a whole right-looking sweep in which both use the same TRSM."""
    A, _ = _spd_packed(N, b, 7)
    n = nblocks(N, b)
    P, Q = A.copy(), A.copy()
    for k in range(n):
        bid = int(oracle.or_block_id(k, k, n))
        oracle.or_potrf(dptr(P), N, b, bid, k)
        oracle.or_potrf_gpu_order(dptr(Q), N, b, bid, k, 0)
        assert np.array_equal(P, Q), ("potrf", k)
        cnt = n - k - 1
        if cnt > 0:
            oracle.or_trsm_gpu_order(dptr(P), b, bid, k, k + 1, cnt, 0)
            oracle.or_trsm_gpu_order(dptr(Q), b, bid, k, k + 1, cnt, 0)
            oracle.or_syrk(dptr(P), b, bid, k, k + 1, cnt, n)
            oracle.or_syrk_gpu_order(dptr(Q), b, bid, k, k + 1, cnt, n, 0)
            assert np.array_equal(P, Q), ("syrk", k)
        if cnt > 1:
            oracle.or_gemm(dptr(P), b, bid, k, k + 2, cnt - 1, n)
            oracle.or_gemm_gpu_order(dptr(Q), b, bid, k, k + 2, cnt - 1, n, 0)
            assert np.array_equal(P, Q), ("gemm", k)


def _find_recip_case():
    """A value x and a diagonal l for which x * fl(1/l) != fl(x / l)."""
    for l in (3.0, 7.0, 10.0, 11.0):
        r = 1.0 / l
        for x in np.linspace(1.0, 2.0, 2001):
            if x * r != x / l:
                return float(x), l
    raise AssertionError("no such pair")


def test_trsm_multiplies_by_the_rounded_reciprocal(oracle):
    """b = 1: the stored value is x * fl(1/l) (the GPU kernel), which on this
    input is not fl(x / l) (the CPU kernel); or_trsm divides."""
    x, l = _find_recip_case()
    A = np.array([l * l, x])  # n = 2 blocks of 1: L_00 = l after POTRF
    for fma in (0, 1):
        G = A.copy()
        oracle.or_potrf_gpu_order(dptr(G), 2, 1, 0, 0, fma)
        assert G[0] == l
        oracle.or_trsm_gpu_order(dptr(G), 1, 0, 0, 1, 1, fma)
        assert G[1] == x * (1.0 / l) and G[1] != x / l
    C = A.copy()
    oracle.or_potrf(dptr(C), 2, 1, 0, 0)
    oracle.or_trsm(dptr(C), 1, 0, 0, 1, 1)
    assert C[1] == x / l


def _fma_case(x, y, v, sign=-1):
    """Inputs where fl(v + sign x y) differs from fl(v + sign fl(x y))."""
    exact = _round(Fraction(v) + sign * Fraction(x) * Fraction(y))
    split = v + sign * (x * y)
    return exact, split


def _search(pred):
    rng = np.random.default_rng(5)
    for _ in range(100000):
        x, y, v = rng.uniform(0.5, 1.5, 3)
        out = pred(float(x), float(y), float(v))
        if out is not None:
            return out
    raise AssertionError("no input found")


def test_syrk_statement_contraction(oracle):
    """b = 1 SYRK: D - c c, one statement; fma = 1 rounds once, fma = 0
    rounds the product first."""
    def pick(c, _, v):
        exact, split = _fma_case(c, c, v + 2.0)
        return (c, v + 2.0, exact, split) if exact != split else None

    c, v, exact, split = _search(pick)
    for fma, want in ((1, exact), (0, split)):
        A = np.array([1.0, c, v])  # blocks (0,0), (1,0), (1,1), b = 1
        oracle.or_syrk_gpu_order(dptr(A), 1, 0, 0, 1, 1, 2, fma)
        assert A[2] == want, fma


def test_gemm_statement_contraction(oracle):
    """b = 2 GEMM on block (2,1) of n = 3: s = b0 c0 + b1 c1 accumulated from
    zero, then a - s. fma = 1: s = fma(b1, c1, fl(b0 c0)); fma = 0: s =
    fl(fl(b0 c0) + fl(b1 c1))."""
    def pick(x, y, v):
        b0, c0, b1, c1 = x, y, v, x + y
        s0 = b0 * c0
        s_f = _round(Fraction(s0) + Fraction(b1) * Fraction(c1))
        s_n = s0 + b1 * c1
        if s_f == s_n:
            return None
        a = 3.0
        return b0, c0, b1, c1, a, a - s_f, a - s_n

    b0, c0, b1, c1, a, want_f, want_n = _search(pick)
    if want_f == want_n:
        pytest.skip("subtraction hides the difference")
    b, n = 2, 3
    for fma, want in ((1, want_f), (0, want_n)):
        A = np.zeros(6 * b * b)
        bid = lambda I, J: J * n - J * (J - 1) // 2 + (I - J)  # noqa: E731
        # L_(2,0) row 0 = (b0, b1), L_(1,0) row 0 = (c0, c1), A_(2,1)[0,0] = a
        A[bid(2, 0) * 4 + 0] = b0
        A[bid(2, 0) * 4 + 1] = b1
        A[bid(1, 0) * 4 + 0] = c0
        A[bid(1, 0) * 4 + 1] = c1
        A[bid(2, 1) * 4 + 0] = a
        oracle.or_gemm_gpu_order(dptr(A), b, 0, 0, 2, 1, n, fma)
        assert A[bid(2, 1) * 4] == want, fma


def test_potrf_and_trsm_statement_contraction(oracle):
    """b = 2 POTRF: l10 = a10 / sqrt(a00), l11 = sqrt(a11 - l10 l10) with one
    rounding (fma = 1) or two (fma = 0). TRSM of a 1 x 2 row (r0, r1):
    x0 = r0 / ... = r0 fl(1/l00), v1 = r1 - l10 x0 (one or two roundings),
    x1 = v1 fl(1/l11)."""
    def pick(x, y, v):
        a00, a10, a11 = 4.0, x, v + 3.0
        l10 = a10 / math.sqrt(a00)
        d_f = _round(Fraction(a11) - Fraction(l10) * Fraction(l10))
        d_n = a11 - l10 * l10
        return (a00, a10, a11, l10, math.sqrt(d_f), math.sqrt(d_n)) if math.sqrt(d_f) != math.sqrt(d_n) else None

    a00, a10, a11, l10, l11_f, l11_n = _search(pick)
    for fma, l11 in ((1, l11_f), (0, l11_n)):
        A = np.array([a00, 0.0, a10, a11])
        oracle.or_potrf_gpu_order(dptr(A), 2, 2, 0, 0, fma)
        assert A[0] == 2.0 and A[2] == l10 and A[3] == l11 and A[1] == 0.0, fma

    def pick_t(x, y, v):
        l00, l10, l11 = 2.0, x, v + 1.0
        r0, r1 = y + 1.0, x + v
        x0 = r0 * (1.0 / l00)
        v_f = _round(Fraction(r1) - Fraction(l10) * Fraction(x0))
        v_n = r1 - l10 * x0
        if v_f * (1.0 / l11) == v_n * (1.0 / l11):
            return None
        return l00, l10, l11, r0, r1, x0, v_f * (1.0 / l11), v_n * (1.0 / l11)

    l00, l10, l11, r0, r1, x0, x1_f, x1_n = _search(pick_t)
    for fma, x1 in ((1, x1_f), (0, x1_n)):
        # n = 2 blocks of 2: L_00 = [[l00, 0], [l10, l11]], block (1,0) rows (r0, r1), (0, 0)
        A = np.array([l00, 0.0, l10, l11, r0, r1, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0])
        oracle.or_trsm_gpu_order(dptr(A), 2, 0, 0, 1, 1, fma)
        assert A[4] == x0 and A[5] == x1, fma


@pytest.mark.parametrize("fma", [0, 1])
@pytest.mark.parametrize("N,b", [(100, 16), (257, 32), (512, 64)])
def test_whole_factor_matches_numpy(oracle, N, b, fma):
    A, X = _spd_packed(N, b, 11, sn2=0.05)
    L = oracle.cholesky_gpu_order(A, N, b, fma)
    Ld = packed_to_dense(L, N, b, lower_only=True)
    K = rbf_dense_numpy(X, 1.0, 1.0, 0.05)
    ref = np.linalg.cholesky(K)
    rel = np.abs(Ld - ref).max() / np.abs(ref).max()
    assert rel < 1e-12, rel
    # and it differs from the CPU order only in rounding
    C = oracle.cholesky(A, N, b)
    Cd = packed_to_dense(C, N, b, lower_only=True)
    assert np.abs(Cd - Ld).max() / np.abs(ref).max() < 1e-13


@pytest.mark.parametrize("N,b", [(20, 4), (37, 8), (257, 32)])
def test_tri_solve_gpu_order_fma0_is_or_tri_solve(oracle, N, b):
    A, _ = _spd_packed(N, b, 3)
    L = oracle.cholesky_gpu_order(A, N, b, 0)
    n = nblocks(N, b)
    rhs = np.zeros(n * b)
    rhs[:N] = np.random.default_rng(N).standard_normal(N)
    assert np.array_equal(oracle.tri_solve_gpu_order(L, rhs, N, b, 0), oracle.tri_solve(L, rhs, N, b))


def test_tri_solve_statement_contraction(oracle):
    """One block of b = 2: forward x1 = (r1 - l10 x0) / l11 with x0 = r0 / l00,
    backward x0' = (x0 - l10 x1') / l00 with x1' = x1 / l11; each `x - l y`
    rounded once (fma = 1) or twice (fma = 0)."""
    l00, l11 = 2.0, 3.0

    def sub(v, x, y, fma):
        return _round(Fraction(v) - Fraction(x) * Fraction(y)) if fma else v - x * y

    def expect(r0, r1, l10, fma):
        x0 = r0 / l00
        x1 = sub(r1, l10, x0, fma) / l11
        x1 = x1 / l11
        x0 = sub(x0, l10, x1, fma) / l00
        return x0, x1

    def pick(x, y, v):
        r0, r1, l10 = x + 1.0, y, v
        f, nf = expect(r0, r1, l10, 1), expect(r0, r1, l10, 0)
        return (r0, r1, l10, f, nf) if f != nf else None

    r0, r1, l10, want_f, want_n = _search(pick)
    L = np.array([l00, 0.0, l10, l11])
    for fma, want in ((1, want_f), (0, want_n)):
        got = oracle.tri_solve_gpu_order(L, np.array([r0, r1]), 2, 2, fma)
        assert tuple(got) == want, fma


def test_col_update_statement_contraction(oracle):
    """Two blocks of b = 2 with identity diagonal blocks, so that only the
    column updates round: forward x1 -= L10 x0 with s = l0 x0 + l1 x1
    accumulated from zero (fma(l1, x1, fl(l0 x0)) at fma = 1), then backward
    x0 -= L10^T x1."""
    def col(xv, rows, fma):
        out = []
        for (a0, a1), v in zip(rows, xv[1]):
            s = a0 * xv[0][0]
            s = _round(Fraction(s) + Fraction(a1) * Fraction(xv[0][1])) if fma else s + a1 * xv[0][1]
            out.append(v - s)
        return out

    def expect(x, M, fma):
        # forward: block 1 -= M x_block0 (rows of M); backward: block 0 -= M^T x_block1
        y1 = col([x[:2], x[2:]], [M[0], M[1]], fma)
        y0 = col([y1, x[:2]], [(M[0][0], M[1][0]), (M[0][1], M[1][1])], fma)
        return tuple(y0) + tuple(y1)

    def pick(a, c, v):
        M = [(a, c), (v, a + c)]
        x = [1.0 + a, 2.0 - c, 0.5 + v, 1.0]
        f, nf = expect(x, M, 1), expect(x, M, 0)
        return (M, x, f, nf) if f != nf else None

    M, x, want_f, want_n = _search(pick)
    b = 2
    L = np.zeros(3 * b * b)
    L[0], L[3] = 1.0, 1.0          # block (0,0) = I
    L[8], L[11] = 1.0, 1.0         # block (1,1) = I
    L[4:8] = [M[0][0], M[0][1], M[1][0], M[1][1]]  # block (1,0)
    for fma, want in ((1, want_f), (0, want_n)):
        got = oracle.tri_solve_gpu_order(L, np.array(x), 4, b, fma)
        assert tuple(got) == want, fma


@pytest.mark.parametrize("fma", [0, 1])
def test_tri_solve_gpu_order_matches_scipy(oracle, fma):
    from scipy.linalg import cho_solve

    N, b = 300, 32
    A, X = _spd_packed(N, b, 9, sn2=0.05)
    L = oracle.cholesky_gpu_order(A, N, b, fma)
    n = nblocks(N, b)
    rhs = np.zeros(n * b)
    rhs[:N] = np.random.default_rng(1).standard_normal(N)
    x = oracle.tri_solve_gpu_order(L, rhs, N, b, fma)
    Ld = packed_to_dense(L, N, b, lower_only=True)
    ref = cho_solve((Ld, True), rhs[:N])
    assert np.abs(x[:N] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(x[N:] == 0.0)
