"""GPU path against the oracle at the block sizes the solvers run with
(64 / 128 / 256, MFMA and streaming kernels), on seeded RBF systems small
enough for the oracle to finish in seconds."""
import ctypes as C

import numpy as np
import pytest

from hs_testlib import dense_to_packed, dptr, packed_len, packed_to_dense, synth_gp_numpy

pytestmark = pytest.mark.gpu


def rbf_system(oracle, N, d, b, ell=1.0, sf2=1.0, sn2=1e-2, seed=1):
    X, y = synth_gp_numpy(seed, N, d)
    A = oracle.rbf(X, N, d, b, ell, sf2, sn2)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    return X, A, rhs


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("N,b", [(1000, 64), (2048, 128), (2560, 256), (300, 20), (4096, 128)])
def test_symv_packed_matches_oracle(hs, dev, oracle, N, b):
    say("symv_packed N=%d b=%d" % (N, b))
    _, A, _ = rbf_system(oracle, N, 8, b)
    n = -(-N // b)
    x = np.zeros(n * b)
    x[:N] = np.random.default_rng(N).standard_normal(N)
    ref = oracle.symv(A, x, b, n)
    dA, dx, dy = dev.upload(A), dev.upload(x), dev.zeros(n * b)
    hs.call("hs_symv_packed_f64", dev.h, dA.p, dx.p, dy.p, N, b)
    y = dev.download(dy, n * b)
    scale = np.abs(A).max() * np.abs(x).sum()
    assert np.max(np.abs(y - ref)) <= 1e-13 * scale
    assert np.all(y[N:] == 0.0)
    # bit-reproducible run to run
    hs.call("hs_symv_packed_f64", dev.h, dA.p, dx.p, dy.p, N, b)
    assert np.array_equal(dev.download(dy, n * b), y)


@pytest.mark.parametrize("b,k,n", [(64, 0, 6), (128, 0, 5), (256, 0, 4), (128, 2, 6), (64, 3, 7)])
def test_trailing_update_mfma_matches_oracle(hs, dev, oracle, b, k, n):
    say("trailing mfma b=%d k=%d n=%d" % (b, k, n))
    N = n * b
    rng = np.random.default_rng(b + k)
    A = rng.standard_normal(packed_len(N, b))
    ref = A.copy()
    bid = n * k - k * (k - 1) // 2
    oracle.or_syrk(dptr(ref), b, bid, k, k + 1, n - k - 1, n)
    if n - k - 2 > 0:
        oracle.or_gemm(dptr(ref), b, bid, k, k + 2, n - k - 2, n)
    dA = dev.upload(A)
    hs.call("hs_syrk_gemm_trailing_f64", dev.h, dA.p, b, k, n)
    got = dev.download(dA, len(A))
    assert np.max(np.abs(got - ref)) <= 1e-12 * b
    # separate SYRK / GEMM entries (reference argument conventions)
    dB = dev.upload(A)
    hs.call("hs_syrk_diag_f64", dev.h, dB.p, b, bid, k, k + 1, n - k - 1, n)
    if n - k - 2 > 0:
        hs.call("hs_gemm_trailing_f64", dev.h, dB.p, b, bid, k, k + 2, n - k - 2, n)
    assert np.max(np.abs(dev.download(dB, len(A)) - ref)) <= 1e-12 * b


@pytest.mark.parametrize("b", [64, 128, 256])
def test_potrf_and_trsm_match_oracle(hs, dev, oracle, b):
    say("potrf/trsm b=%d" % b)
    N = 3 * b
    _, A, _ = rbf_system(oracle, N, 8, b, sn2=1e-1, seed=b)
    ref = A.copy()
    oracle.or_potrf(dptr(ref), N, b, 0, 0)
    oracle.or_trsm(dptr(ref), b, 0, 0, 1, 2)
    dA = dev.upload(A)
    hs.call("hs_reset_info", dev.h)
    hs.call("hs_potrf_diag_f64", dev.h, dA.p, N, b, 0, 0)
    hs.call("hs_trsm_panel_f64", dev.h, dA.p, b, 0, 0, 1, 2)
    got = dev.download(dA, len(A))
    assert dev.info() == 0
    assert np.max(np.abs(got - ref)) <= 1e-11


@pytest.mark.parametrize("N,b", [(1024, 64), (1000, 64), (2048, 128), (1536, 256), (500, 20),
                                 # two-panel look-ahead edge cases: 4, 5 and 7 block columns, padded or not
                                 (512, 128), (1280, 256), (1100, 256), (448, 64), (330, 64),
                                 # 1024-row super-block steps: padded last super-block, remainder blocks
                                 (2000, 128), (2300, 256), (3000, 64),
                                 # the Cholesky + solve workload on the CPU reference path, exactly:
                                 # N = 4096, b = 128, d = 8, seed 1, l = 1, sf2 = 1, sn2 = 1e-2
                                 (4096, 128)])
def test_cholesky_and_solve_match_oracle(hs, dev, oracle, N, b):
    """sn2 = 1e-2, d = 8 (cond(K) up to ~1e5): the factor within the
    factorisations' rounding, the solution within 1e-12 relative of the
    oracle's (the reference algorithm on the CPU)."""
    say("cholesky+solve N=%d b=%d" % (N, b))
    X, A, rhs = rbf_system(oracle, N, 8, b)
    Lref = oracle.cholesky(A, N, b)
    xref = oracle.tri_solve(Lref, rhs, N, b)
    dA, dx = dev.upload(A), dev.upload(rhs)
    info = hs.call("hs_cholesky_f64", dev.h, dA.p, N, b)
    assert info == 0
    L = dev.download(dA, len(A))
    hs.call("hs_tri_solve_f64", dev.h, dA.p, dx.p, N, b)
    x = dev.download(dx, len(rhs))
    # factor: two backward-stable FP64 factorisations with different
    # summation orders differ by O(cond * u)
    assert np.max(np.abs(L - Lref)) <= 1e-10
    K = packed_to_dense(A, N, b)
    rel_res = np.linalg.norm(K @ x[:N] - rhs[:N]) / np.linalg.norm(rhs[:N])
    rel_diff = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    say("  rel residual %.3e  rel diff vs oracle %.3e" % (rel_res, rel_diff))
    assert rel_res <= 1e-12
    assert rel_diff <= 1e-12
    assert np.all(x[N:] == 0.0)


@pytest.mark.parametrize("N", [4096, 6144])
def test_long_lengthscale_construction_matches_oracle(hs, dev, oracle, N):
    """The N = 65536 workload's construction (d = 16, b = 256, l = 2,
    sn2 = 1e-3, seed 3) at sizes the oracle finishes in seconds. Its
    condition number bound kappa <= ||K||_inf / sn2 (the RBF part is positive
    semi-definite, so lambda_min >= sn2) decides the bar: 1e-12 relative to
    the oracle's solution where kappa u <= 1e-13; otherwise the normwise
    backward error within 4x the oracle's and the difference within kappa u."""
    b = 256
    say("long-lengthscale construction N=%d" % N)
    X, y = synth_gp_numpy(3, N, 16)
    A = oracle.rbf(X, N, 16, b, 2.0, 1.0, 1e-3)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    xref = oracle.tri_solve(oracle.cholesky(A, N, b), rhs, N, b)
    dA, dx = dev.upload(A), dev.upload(rhs)
    assert hs.call("hs_cholesky_f64", dev.h, dA.p, N, b) == 0
    hs.call("hs_tri_solve_f64", dev.h, dA.p, dx.p, N, b)
    x = dev.download(dx, len(rhs))
    K = packed_to_dense(A, N, b)
    knorm = np.max(np.sum(np.abs(K), axis=1))
    kappa = knorm / 1e-3
    u = np.finfo(float).eps / 2

    def omega(v):
        r = rhs[:N] - K @ v[:N]
        return np.max(np.abs(r)) / (knorm * np.max(np.abs(v[:N])) + np.max(np.abs(rhs[:N])))

    rel = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    w, wref = omega(x), omega(xref)
    say("  kappa bound %.3e (kappa u %.2e); rel diff vs oracle %.3e; backward error GPU %.3e oracle %.3e"
        % (kappa, kappa * u, rel, w, wref))
    if kappa * u <= 1e-13:
        assert rel <= 1e-12
    else:
        assert w <= max(4 * wref, 4 * u)
        assert rel <= kappa * u
    assert np.all(x[N:] == 0.0)


@pytest.mark.parametrize("N,b,sn2", [(4096, 128, 4.0), (2048, 256, 16.0), (3000, 64, 8.0)])
def test_solve_within_1e12_when_well_conditioned(hs, dev, oracle, N, b, sn2):
    """The FP64 solution agrees with the oracle's (the reference algorithm on
    the CPU) to 1e-12 relative where cond(K) * u leaves room for it."""
    say("cholesky+solve 1e-12 N=%d b=%d sn2=%g" % (N, b, sn2))
    X, A, rhs = rbf_system(oracle, N, 8, b, sn2=sn2)
    xref = oracle.tri_solve(oracle.cholesky(A, N, b), rhs, N, b)
    dA, dx = dev.upload(A), dev.upload(rhs)
    assert hs.call("hs_cholesky_f64", dev.h, dA.p, N, b) == 0
    hs.call("hs_tri_solve_f64", dev.h, dA.p, dx.p, N, b)
    x = dev.download(dx, len(rhs))
    rel = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    say("  rel diff vs oracle %.3e" % rel)
    assert rel <= 1e-12


@pytest.mark.parametrize("N,b", [(40960, 256), (36864, 128), (34816, 64)])
def test_solve_recovers_known_solution_at_scale(hs, dev, N, b):
    """Sizes past the oracle's reach, where the backward super-steps near the
    bottom use their wide-grid form: y = K x_true (SYMV on the GPU), factor,
    solve, and x must match x_true to cond(K) * u with a small backward
    error."""
    say("known-solution solve N=%d b=%d" % (N, b))
    d, ell, sf2, sn2 = 8, 1.0, 1.0, 1e-1
    n = -(-N // b)
    rng = np.random.default_rng(11)
    X = rng.uniform(-1.0, 1.0, (N, d))
    xt = np.zeros(n * b)
    xt[:N] = rng.standard_normal(N)
    dX, dxt, dy = dev.upload(X.reshape(-1)), dev.upload(xt), dev.zeros(n * b)
    dA = dev.alloc(packed_len(N, b) * 8)
    hs.call("hs_rbf_assemble_f64", dev.h, dX.p, N, d, b, ell, sf2, sn2, dA.p)
    hs.call("hs_symv_packed_f64", dev.h, dA.p, dxt.p, dy.p, N, b)
    y = dev.download(dy, n * b)
    assert hs.call("hs_cholesky_f64", dev.h, dA.p, N, b) == 0
    hs.call("hs_tri_solve_f64", dev.h, dA.p, dy.p, N, b)
    x = dev.download(dy, n * b)
    # residual against a fresh K
    hs.call("hs_rbf_assemble_f64", dev.h, dX.p, N, d, b, ell, sf2, sn2, dA.p)
    dx, dr = dev.upload(x), dev.zeros(n * b)
    hs.call("hs_symv_packed_f64", dev.h, dA.p, dx.p, dr.p, N, b)
    r = y - dev.download(dr, n * b)
    ones = np.zeros(n * b)
    ones[:N] = 1.0
    hs.call("hs_symv_packed_f64", dev.h, dA.p, dev.upload(ones).p, dr.p, N, b)
    knorm = float(np.max(dev.download(dr, n * b)))
    dA.free()
    bwd = np.max(np.abs(r)) / (knorm * np.max(np.abs(x)) + np.max(np.abs(y)))
    rel = np.linalg.norm(x - xt) / np.linalg.norm(xt)
    say("  backward error %.3e  rel error vs x_true %.3e" % (bwd, rel))
    # cond(K) <= 1 + N sf2 / sn2 < 5e5 here
    assert bwd <= 1e-15
    assert rel <= 1e-8
    assert np.all(x[N:] == 0.0)


@pytest.mark.parametrize("N,b", [(4096, 128), (4096, 256)])
def test_super_block_solve_backward_error_when_ill_conditioned(hs, dev, oracle, N, b):
    """The super-block steps multiply by explicit inverses of 1024-row
    diagonal blocks. At a noise variance of 1e-7 (cond(K) ~1e9-1e10) their
    normwise backward error must stay at the level of the oracle's
    substitution (the reference algorithm on the CPU)."""
    say("ill-conditioned solve N=%d b=%d" % (N, b))
    X, A, rhs = rbf_system(oracle, N, 8, b, sn2=1e-7)
    Lref = oracle.cholesky(A, N, b)
    xref = oracle.tri_solve(Lref, rhs, N, b)
    dA, dx = dev.upload(A), dev.upload(rhs)
    assert hs.call("hs_cholesky_f64", dev.h, dA.p, N, b) == 0
    hs.call("hs_tri_solve_f64", dev.h, dA.p, dx.p, N, b)
    x = dev.download(dx, len(rhs))
    K = packed_to_dense(A, N, b)
    knorm = np.max(np.sum(np.abs(K), axis=1))

    def omega(v):
        r = rhs[:N] - K @ v[:N]
        return np.max(np.abs(r)) / (knorm * np.max(np.abs(v[:N])) + np.max(np.abs(rhs[:N])))

    w, wref = omega(x), omega(xref)
    say("  backward error GPU %.3e, oracle %.3e" % (w, wref))
    assert w <= max(10 * wref, 1e-15)


@pytest.mark.parametrize("N,b", [(3000, 64), (2300, 256), (1000, 128)])
def test_single_direction_solves(hs, dev, oracle, N, b):
    """hs_trsv_packed_f64: trans = 0 solves L z = y, trans = 1 solves
    L^T x = z (against scipy on the dense oracle factor), and the two in a
    row give hs_tri_solve_f64's result bit for bit (the same kernels)."""
    from scipy.linalg import solve_triangular

    say("single-direction solves N=%d b=%d" % (N, b))
    X, A, rhs = rbf_system(oracle, N, 8, b, sn2=1.0)
    dA = dev.upload(A)
    assert hs.call("hs_cholesky_f64", dev.h, dA.p, N, b) == 0
    L = dev.download(dA, len(A))
    Ld = packed_to_dense(L, N, b, lower_only=True)
    zref = solve_triangular(Ld, rhs[:N], lower=True)
    xref = solve_triangular(Ld.T, zref, lower=False)
    dz = dev.upload(rhs)
    assert hs.call("hs_trsv_packed_f64", dev.h, dA.p, dz.p, N, b, 0) == 0
    z = dev.download(dz, len(rhs))
    assert hs.call("hs_trsv_packed_f64", dev.h, dA.p, dz.p, N, b, 1) == 0
    x = dev.download(dz, len(rhs))
    dx = dev.upload(rhs)
    hs.call("hs_tri_solve_f64", dev.h, dA.p, dx.p, N, b)
    x2 = dev.download(dx, len(rhs))
    ez = np.linalg.norm(z[:N] - zref) / np.linalg.norm(zref)
    ex = np.linalg.norm(x[:N] - xref) / np.linalg.norm(xref)
    say("  forward %.3e  backward %.3e" % (ez, ex))
    assert ez <= 1e-12 and ex <= 1e-12
    assert np.array_equal(x, x2)


@pytest.mark.parametrize("count", [16384 * 3, 16384 * 2 + 1000, 777])
def test_read_sum_probe_reads_every_element(hs, dev, count):
    """The streaming-read probe behind the bench's read-ceiling figure sums
    every element once (full chunks on the vector path, a partial last chunk
    element by element)."""
    x = np.random.default_rng(count).standard_normal(count)
    dx = dev.upload(x)
    g = -(-count // 16384)
    dp = dev.zeros(g)
    assert hs.call("hs_read_sum_f64", dev.h, dx.p, count, dp.p) == 0
    part = dev.download(dp, g)
    ref = [x[i * 16384:(i + 1) * 16384].sum() for i in range(g)]
    assert np.allclose(part, ref, rtol=1e-12, atol=1e-9)


def test_cholesky_reports_first_bad_pivot(hs, dev, oracle):
    N, b = 256, 64
    K = np.eye(N)
    K[130, 130] = -1.0
    A = dense_to_packed(K, b)
    dA = dev.upload(A)
    info = hs.call("hs_cholesky_f64", dev.h, dA.p, N, b)
    assert info == 131


@pytest.mark.parametrize("N,b,eps", [(2048, 128, 1e-10), (3000, 64, 1e-10), (4096, 256, 1e-8)])
def test_cg_matches_oracle(hs, dev, oracle, N, b, eps):
    say("cg N=%d b=%d eps=%g" % (N, b, eps))
    X, A, rhs = rbf_system(oracle, N, 8, b)
    xref, itref, _ = oracle.cg(A, rhs, N, b, eps)
    dA, drhs, dx = dev.upload(A), dev.upload(rhs), dev.zeros(len(rhs))
    it, df, d0 = C.c_int64(), C.c_double(), C.c_double()
    hs.call("hs_cg_f64", dev.h, dA.p, drhs.p, dx.p, N, b, eps, 100000, C.byref(it), C.byref(df), C.byref(d0))
    x = dev.download(dx, len(rhs))
    say("  iterations gpu %d oracle %d" % (it.value, itref))
    assert abs(it.value - itref) <= max(1, int(0.05 * itref))
    K = packed_to_dense(A, N, b)
    assert np.linalg.norm(K @ x[:N] - rhs[:N]) <= 3 * eps * np.linalg.norm(rhs[:N])
    assert df.value <= eps * eps * d0.value


@pytest.mark.parametrize("N,b,sn2", [(4096, 128, 4.0), (2500, 64, 1.0), (8192, 256, 16.0)])
def test_cg_iterations_within_one(hs, dev, oracle, N, b, sn2):
    """Where the convergence history is not dominated by rounding (moderate
    condition number), the iteration count matches the oracle's to +-1."""
    say("cg iterations N=%d b=%d sn2=%g" % (N, b, sn2))
    _, A, rhs = rbf_system(oracle, N, 8, b, sn2=sn2)
    xref, itref, _ = oracle.cg(A, rhs, N, b, 1e-10)
    dA, drhs, dx = dev.upload(A), dev.upload(rhs), dev.zeros(len(rhs))
    it = C.c_int64()
    hs.call("hs_cg_f64", dev.h, dA.p, drhs.p, dx.p, N, b, 1e-10, 100000, C.byref(it), None, None)
    say("  iterations gpu %d oracle %d" % (it.value, itref))
    assert abs(it.value - itref) <= 1
    x = dev.download(dx, len(rhs))
    assert np.linalg.norm(x - xref) <= 1e-9 * np.linalg.norm(xref)


def test_cg_imax_and_zero_rhs(hs, dev, oracle):
    N, b = 512, 64
    _, A, rhs = rbf_system(oracle, N, 8, b)
    dA, drhs, dx = dev.upload(A), dev.upload(rhs), dev.zeros(len(rhs))
    it = C.c_int64()
    hs.call("hs_cg_f64", dev.h, dA.p, drhs.p, dx.p, N, b, 1e-14, 7, C.byref(it), None, None)
    assert it.value == 7
    xref, itref, _ = oracle.cg(A, rhs, N, b, 1e-14, imax=7)
    assert np.allclose(dev.download(dx, len(rhs)), xref, rtol=1e-9, atol=1e-12)
    dz = dev.zeros(len(rhs))
    hs.call("hs_cg_f64", dev.h, dA.p, dz.p, dx.p, N, b, 1e-10, 100, C.byref(it), None, None)
    assert it.value == 0 and np.all(dev.download(dx, len(rhs)) == 0.0)


@pytest.mark.parametrize("N,d,b,ell,sn2", [(1000, 8, 64, 1.0, 1e-2), (2048, 16, 256, 2.0, 1e-3),
                                          (777, 8, 128, 1.0, 1e-2), (300, 5, 20, 0.8, 1e-2)])
def test_rbf_assembly_matches_oracle(hs, dev, oracle, N, d, b, ell, sn2):
    say("rbf N=%d d=%d b=%d" % (N, d, b))
    X, _ = synth_gp_numpy(3, N, d)
    ref = oracle.rbf(X, N, d, b, ell, 1.0, sn2)
    dX, dA = dev.upload(X.reshape(-1)), dev.zeros(packed_len(N, b))
    hs.call("hs_rbf_assemble_f64", dev.h, dX.p, N, d, b, ell, 1.0, sn2, dA.p)
    got = dev.download(dA, packed_len(N, b))
    assert np.max(np.abs(got - ref)) <= 4e-16 * np.maximum(1.0, np.abs(ref)).max() * 4


def test_rbf_window_matches_oracle(hs, dev, oracle):
    N, b = 1000, 64
    t = np.sin(np.arange(N) * 0.01) + 0.05 * np.random.default_rng(0).standard_normal(N)
    ref = np.zeros(packed_len(N, b))
    oracle.or_rbf_window(dptr(t), N, b, dptr(ref))
    dt, dA = dev.upload(t), dev.zeros(packed_len(N, b))
    hs.call("hs_rbf_window_f64", dev.h, dt.p, N, b, dA.p)
    assert np.max(np.abs(dev.download(dA, len(ref)) - ref)) <= 2e-15
