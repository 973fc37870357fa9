"""The reference-order blocked Cholesky (hs_cholesky_ref_order_f64 and its
four kernels) against the oracle's restatement of the reference's share-1.0
GPU path (or_*_gpu_order): bit for bit, with and without contraction, on the
reference's golden cases and on the workloads' RBF constructions."""
import time

import numpy as np
import pytest

from backends import HipRefOrderBackend, OracleGpuOrderBackend
from golden_cases import case_id, check_case, factor_cases, run_case
from hs_testlib import dense_to_packed, packed_to_dense, synth_gp_numpy

pytestmark = pytest.mark.gpu



def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("fma", [0, 1])
@pytest.mark.parametrize("case", factor_cases(), ids=case_id)
def test_ref_order_kernels_match_goldens_and_oracle(hs, dev, oracle, case, fma):
    final, env = run_case(case, HipRefOrderBackend(hs, dev, fma))
    fails = check_case(case, final, env)
    assert not fails, fails
    want, _ = run_case(case, OracleGpuOrderBackend(oracle, fma))
    for name, arr in want.items():
        assert np.array_equal(final[name], arr), name


# (N, b, d, l, sn2, seed): the workloads' constructions at sizes the oracle
# finishes in seconds
SYSTEMS = [
    (300, 20, 8, 1.0, 1e-2, 1),      # b not a multiple of the 64-wide tiles, padding
    (1000, 64, 8, 1.0, 1e-2, 1),
    (2000, 128, 8, 1.0, 1e-2, 1),
    (2300, 256, 8, 1.0, 1e-2, 1),
    (4096, 128, 8, 1.0, 1e-2, 1),    # the Cholesky + solve workload on the CPU path, exactly
    (4096, 256, 16, 2.0, 1e-3, 3),   # the N = 65536 workload's construction
]


@pytest.mark.parametrize("fma", [0, 1])
@pytest.mark.parametrize("N,b,d,ell,sn2,seed", SYSTEMS)
def test_ref_order_factor_is_bit_identical(hs, dev, oracle, N, b, d, ell, sn2, seed, fma):
    say("reference-order factor N=%d b=%d d=%d l=%g sn2=%g fma=%d" % (N, b, d, ell, sn2, fma))
    X, y = synth_gp_numpy(seed, N, d)
    A = oracle.rbf(X, N, d, b, ell, 1.0, sn2)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    t0 = time.time()
    Lref = oracle.cholesky_gpu_order(A, N, b, fma)
    xref = oracle.tri_solve_gpu_order(Lref, rhs, N, b, fma)
    t1 = time.time()
    dA, dx = dev.upload(A), dev.upload(rhs)
    info = hs.call("hs_cholesky_ref_order_f64", dev.h, dA.p, N, b, fma)
    hs.call("hs_tri_solve_ref_order_f64", dev.h, dA.p, dx.p, N, b, fma)
    dev.sync()
    t2 = time.time()
    assert info == 0
    L = dev.download(dA, len(A))
    x = dev.download(dx, len(rhs))
    ndiff = int(np.count_nonzero(L != Lref))
    # the fast factor and solve on the same input, for the record
    dF, dxf = dev.upload(A), dev.upload(rhs)
    assert hs.call("hs_cholesky_f64", dev.h, dF.p, N, b) == 0
    hs.call("hs_tri_solve_f64", dev.h, dF.p, dxf.p, N, b)
    F = dev.download(dF, len(A))
    xf = dev.download(dxf, len(rhs))
    say("  oracle %.2f s, GPU %.2f s; differing factor entries %d of %d, solution entries %d; fast path: "
        "factor max diff %.3e (relative to max |L|), solution rel diff %.3e"
        % (t1 - t0, t2 - t1, ndiff, len(A), int(np.count_nonzero(x != xref)),
           np.abs(F - Lref).max() / np.abs(Lref).max(), np.linalg.norm(xf - xref) / np.linalg.norm(xref)))
    assert ndiff == 0
    assert np.array_equal(x, xref)


def test_ref_order_differs_from_cpu_order_only_in_rounding(hs, dev, oracle):
    """The GPU order is not the CPU order (the TRSM's reciprocal), but the two
    factors agree to rounding."""
    N, b = 1000, 64
    X, _ = synth_gp_numpy(1, N, 8)
    A = oracle.rbf(X, N, 8, b, 1.0, 1.0, 1e-2)
    dA = dev.upload(A)
    assert hs.call("hs_cholesky_ref_order_f64", dev.h, dA.p, N, b, 0) == 0
    L = dev.download(dA, len(A))
    C = oracle.cholesky(A, N, b)
    assert not np.array_equal(L, C)
    assert np.abs(L - C).max() <= 1e-12 * np.abs(C).max()
    Ld = packed_to_dense(L, N, b, lower_only=True)
    K = packed_to_dense(A, N, b)
    assert np.abs(Ld @ Ld.T - K).max() <= 1e-13 * np.abs(K).max()


def test_ref_order_reports_first_bad_pivot(hs, dev, oracle):
    N, b = 300, 64
    X, _ = synth_gp_numpy(1, N, 8)
    A = oracle.rbf(X, N, 8, b, 1.0, 1.0, 1e-2)
    K = packed_to_dense(A, N, b)
    K[200, 200] = -1.0
    B = dense_to_packed(K, b)
    dB = dev.upload(B)
    info = hs.call("hs_cholesky_ref_order_f64", dev.h, dB.p, N, b, 1)
    dF = dev.upload(B)
    info_fast = hs.call("hs_cholesky_f64", dev.h, dF.p, N, b)
    assert info == 201 and info_fast == 201


def test_ref_order_argument_errors(hs, dev):
    lib = hs.load()
    d = dev.zeros(64 * 64)
    assert lib.hs_cholesky_ref_order_f64(dev.h, d.p, 64, 64, 2) == hs.HS_ERR_UNSUPPORTED
    assert lib.hs_cholesky_ref_order_f64(dev.h, None, 64, 64, 0) == -2
    assert lib.hs_cholesky_ref_order_f64(dev.h, d.p, 0, 64, 0) == -4
    assert lib.hs_potrf_ref_order_f64(dev.h, d.p, 64, 2048, 0, 0, 0) == hs.HS_ERR_UNSUPPORTED
    # SYRK over block rows past the matrix (n = 3, rows 1 .. 5)
    assert lib.hs_syrk_ref_order_f64(dev.h, d.p, 64, 0, 0, 1, 5, 3, 0) == -4


def test_ref_order_at_the_full_workload_size(hs, dev):
    """The N = 65536 workload (d = 16, b = 256, l = 2, sn2 = 1e-3, seed 3), past
    the oracle's reach: the reference-order factor and solves against the fast
    path through size-independent properties. Both solutions leave a relative
    residual at the FP64 path's level, and they differ by no more than the
    conditioning allows (kappa <= ||K||_inf / sn2)."""
    import ctypes as C

    N, d, b, ell, sn2 = 65536, 16, 256, 2.0, 1e-3
    X, y = synth_gp_numpy(3, N, d)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    packed = n * (n + 1) // 2 * b * b
    dX = dev.upload(X.reshape(-1))
    dK, dR, dF = dev.zeros(packed), dev.zeros(packed), dev.zeros(packed)
    for buf in (dK, dR, dF):
        hs.call("hs_rbf_assemble_f64", dev.h, dX.p, N, d, b, ell, 1.0, sn2, buf.p)
    dev.sync()
    t0 = time.time()
    assert hs.call("hs_cholesky_ref_order_f64", dev.h, dR.p, N, b, 1) == 0
    dxr = dev.upload(rhs)
    hs.call("hs_tri_solve_ref_order_f64", dev.h, dR.p, dxr.p, N, b, 1)
    dev.sync()
    t1 = time.time()
    assert hs.call("hs_cholesky_f64", dev.h, dF.p, N, b) == 0
    dxf = dev.upload(rhs)
    hs.call("hs_tri_solve_f64", dev.h, dF.p, dxf.p, N, b)
    dev.sync()
    t2 = time.time()
    xr, xf = dev.download(dxr, n * b), dev.download(dxf, n * b)

    def rel_residual(x):
        dx, dy = dev.upload(x), dev.zeros(n * b)
        hs.call("hs_symv_packed_f64", dev.h, dK.p, dx.p, dy.p, N, b)
        return np.linalg.norm(dev.download(dy, n * b)[:N] - y) / np.linalg.norm(y)

    rr, rf = rel_residual(xr), rel_residual(xf)
    diff = np.linalg.norm(xr - xf) / np.linalg.norm(xr)
    say("N=65536 reference order: %.1f s; fast: %.2f s; relative residual %.3e / %.3e; solutions differ by %.3e"
        % (t1 - t0, t2 - t1, rr, rf, diff))
    assert rr <= 1e-10 and rf <= 1e-10
    # kappa <= ||K||_inf / sn2 <= N sf2 / sn2: two backward-stable solves differ by O(kappa u)
    assert diff <= (N / sn2) * np.finfo(float).eps
    assert np.all(xr[N:] == 0.0) and np.all(xf[N:] == 0.0)
