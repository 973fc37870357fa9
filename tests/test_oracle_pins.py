"""Pins of the oracle's whole-path functions, none of which the reference's
own tests cover: checked against textbook routines (numpy/scipy), closed
forms and an independent textbook CG, on small inputs (CPU only)."""
import numpy as np
import pytest
import scipy.linalg as sla

from hs_testlib import (dense_to_packed, dptr, packed_len, packed_to_dense, rbf_dense_numpy, splitmix64_uniforms,
                        synth_gp_numpy)


def _spd(N, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((N, N))
    return M @ M.T / N + np.eye(N)


@pytest.mark.parametrize("N,b", [(64, 16), (70, 16), (128, 32), (20, 6), (96, 64)])
def test_cholesky_matches_numpy(oracle, N, b):
    K = _spd(N, N + b)
    A = dense_to_packed(K, b)
    L = packed_to_dense(oracle.cholesky(A, N, b), N, b, lower_only=True)
    Lref = np.linalg.cholesky(K)
    assert np.allclose(L, Lref, rtol=0, atol=1e-12)
    assert np.allclose(L @ L.T, K, rtol=0, atol=1e-12)


@pytest.mark.parametrize("N,b", [(64, 16), (70, 16), (20, 6)])
def test_tri_solve_matches_cho_solve(oracle, N, b):
    K = _spd(N, 7 * N + b)
    rhs = np.random.default_rng(3).standard_normal(N)
    L = oracle.cholesky(dense_to_packed(K, b), N, b)
    n = -(-N // b)
    x = np.zeros(n * b)
    x[:N] = rhs
    x = oracle.tri_solve(L, x, N, b)
    ref = sla.cho_solve(sla.cho_factor(K, lower=True), rhs)
    assert np.allclose(x[:N], ref, rtol=1e-12, atol=1e-12)
    assert np.all(x[N:] == 0.0)


def _textbook_cg(K, rhs, eps):
    """Plain CG with x0 = 0, true residual every 50 iterations (CG.cpp:95-101),
    delta ratio stop; written independently of the oracle."""
    x = np.zeros_like(rhs)
    r = rhs.copy()
    d = r.copy()
    dn = r @ r
    d0 = dn
    it = 0
    while it < 100000 and dn > eps * eps * d0:
        q = K @ d
        a = dn / (d @ q)
        x = x + a * d
        r = rhs - K @ x if it % 50 == 0 else r - a * q
        do, dn = dn, r @ r
        d = r + (dn / do) * d
        it += 1
    return x, it


# CG iteration counts are a rounding-sensitive quantity: on an RBF system with
# cond ~ 2e4 the oracle alone moves by +-5 iterations when only the chunking of
# its dot product changes. The count is pinned exactly on well-conditioned
# systems (sn2 = 1, cond ~ 1e2) and to 5% on the others.
@pytest.mark.parametrize("N,b,eps,sn2,slack", [(64, 16, 1e-10, 1.0, 1), (256, 64, 1e-10, 1.0, 1),
                                                (100, 32, 1e-8, 1e-1, 3), (256, 64, 1e-10, 1e-1, 5)])
def test_cg_matches_textbook_and_solve(oracle, N, b, eps, sn2, slack):
    X, y = synth_gp_numpy(11, N, 4)
    K = rbf_dense_numpy(X, 1.0, 1.0, sn2)
    A = dense_to_packed(K, b)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    x, it, delta = oracle.cg(A, rhs, N, b, eps)
    xt, itt = _textbook_cg(K, y, eps)
    assert abs(it - itt) <= slack
    ref = np.linalg.solve(K, y)
    assert np.linalg.norm(K @ x[:N] - y) <= 2 * eps * np.linalg.norm(y)
    assert np.allclose(x[:N], ref, rtol=0, atol=1e-6 * np.abs(ref).max())


def test_cg_zero_rhs_takes_no_iterations(oracle):
    N, b = 32, 16
    A = dense_to_packed(_spd(N, 1), b)
    x, it, delta = oracle.cg(A, np.zeros(N), N, b, 1e-10)
    assert it == 0 and delta == 0.0 and np.all(x == 0)


@pytest.mark.parametrize("N,d,b,ell,sf2,sn2", [(50, 3, 16, 0.7, 1.3, 1e-2), (64, 8, 64, 1.0, 1.0, 1e-2),
                                                 (33, 16, 8, 2.0, 1.0, 1e-3)])
def test_rbf_matches_numpy(oracle, N, d, b, ell, sf2, sn2):
    X, _ = synth_gp_numpy(5, N, d)
    A = oracle.rbf(X, N, d, b, ell, sf2, sn2)
    K = packed_to_dense(A, N, b)
    assert np.allclose(K, rbf_dense_numpy(X, ell, sf2, sn2), rtol=1e-14, atol=1e-15)
    # padding stays zero
    n = -(-N // b)
    full = packed_to_dense(A, n * b, b)
    assert np.all(full[N:, :] == 0) and np.all(full[:, N:] == 0)


def test_rbf_window_mode(oracle):
    N, b = 40, 16
    t = np.sin(np.arange(N) * 0.37) + 0.1 * np.cos(np.arange(N) * 1.3)
    A = np.zeros(packed_len(N, b))
    oracle.or_rbf_window(dptr(t), N, b, dptr(A))
    tin = np.concatenate([np.zeros(7), t])
    X = np.stack([tin[i:i + 8] for i in range(N)])
    K = rbf_dense_numpy(X, 1.0, 1.0, 0.01)
    assert np.allclose(packed_to_dense(A, N, b), K, rtol=1e-14, atol=1e-15)


def test_synthetic_generator_matches_numpy(oracle):
    for seed, N, d in [(1, 100, 8), (3, 37, 16)]:
        X, y = oracle.synth(seed, N, d)
        Xn, yn = synth_gp_numpy(seed, N, d)
        assert np.array_equal(X, Xn)
        assert np.allclose(y, yn, rtol=1e-15, atol=1e-15)
    u = splitmix64_uniforms(0, 3)
    # published SplitMix64 outputs for seed 0 (first three values)
    z = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert np.array_equal(u, np.array([(v >> 11) * 2.0 ** -53 for v in z]))


def test_packed_layout_ids(oracle):
    n = 7
    ids = [[oracle.or_block_id(I, J, n) for J in range(I + 1)] for I in range(n)]
    flat = sorted(v for row in ids for v in row)
    assert flat == list(range(n * (n + 1) // 2))
    # column-major enumeration: block (J,J) starts column J
    for J in range(n):
        assert ids[J][J] == sum(n - c for c in range(J))


# ---- GP prediction (K* assembly, window features, the GP mat-vec) ----------

def _cross_numpy(Xtr, Xte, ell, sf2):
    d2 = ((Xte[:, None, :] - Xtr[None, :, :]) ** 2).sum(-1)
    return sf2 * np.exp(-0.5 / ell ** 2 * d2)


@pytest.mark.parametrize("N,M,d,ell,sf2", [(40, 7, 3, 0.8, 1.7), (64, 33, 8, 1.0, 1.0), (5, 1, 16, 2.0, 0.5),
                                           (1, 4, 1, 0.3, 2.0)])
def test_rbf_cross_matches_numpy(oracle, N, M, d, ell, sf2):
    rng = np.random.default_rng(N * 100 + M)
    Xtr, Xte = rng.uniform(-1, 1, (N, d)), rng.uniform(-1, 1, (M, d))
    K = oracle.rbf_cross(Xtr, Xte, ell, sf2)
    assert K.shape == (M, N)
    assert np.allclose(K, _cross_numpy(Xtr, Xte, ell, sf2), rtol=1e-14, atol=1e-300)


def test_rbf_cross_equal_inputs_give_sf2(oracle):
    """A test input equal to training input i gives K*[j, i] = sf2 exactly
    (distance 0, exp(0) = 1); a reversed layout or a dropped transpose moves
    the ones off the diagonal."""
    rng = np.random.default_rng(5)
    Xtr = rng.uniform(-1, 1, (20, 8))
    K = oracle.rbf_cross(Xtr, Xtr[[3, 11, 19]], 1.3, 2.5)
    for j, i in enumerate([3, 11, 19]):
        assert K[j, i] == 2.5
        assert np.all(np.delete(K[j], i) < 2.5)


def test_window_features_match_reference_construction(oracle):
    """readInputVector puts the series after 7 zeros and the features of
    point i are entries i..i+7 of that padded vector."""
    t = np.arange(1.0, 13.0)
    X = oracle.window(t)
    tin = np.concatenate([np.zeros(7), t])
    assert np.array_equal(X, np.stack([tin[i:i + 8] for i in range(len(t))]))
    assert np.array_equal(X[0], [0, 0, 0, 0, 0, 0, 0, 1.0])
    assert np.array_equal(X[9], np.arange(3.0, 11.0))


@pytest.mark.parametrize("m,n", [(1, 1), (7, 300), (33, 64)])
def test_gemv_gp_matches_numpy(oracle, m, n):
    rng = np.random.default_rng(m + n)
    A, x = rng.standard_normal((m, n)), rng.standard_normal(n)
    assert np.allclose(oracle.gemv_gp(A, x), A @ x, rtol=1e-13, atol=1e-13)


def test_prediction_at_training_inputs_is_y_minus_noise_alpha(oracle):
    """K*(X, X) = K - sn2 I, so the predictive mean at the training inputs is
    K* alpha = y - sn2 alpha for alpha = K^{-1} y: ties the K* assembly, the
    mat-vec and the solve together through a closed form."""
    N, d, ell, sf2, sn2 = 120, 4, 0.9, 1.2, 0.05
    rng = np.random.default_rng(11)
    X = rng.uniform(-1, 1, (N, d))
    y = np.sin(X.sum(1))
    alpha = np.linalg.solve(rbf_dense_numpy(X, ell, sf2, sn2), y)
    mean = oracle.gemv_gp(oracle.rbf_cross(X, X, ell, sf2), alpha)
    assert np.allclose(mean, y - sn2 * alpha, rtol=1e-9, atol=1e-9)
