"""Mixed-precision factor (FP32 storage; FP32-MFMA, BF16 hi+lo or scaled FP16
hi+lo trailing updates) and FP64 iterative refinement with the PCG fallback, against the
FP64 oracle: the refined solution must reach the FP64 backward error target
(1e-12) and agree with the oracle's FP64 solution to within what the
conditioning allows."""
import ctypes as C

import numpy as np
import pytest

from hs_testlib import dense_to_packed, packed_len, packed_to_dense, synth_gp_numpy

pytestmark = pytest.mark.gpu


def system(oracle, N, d, b, ell, sn2, seed):
    X, y = synth_gp_numpy(seed, N, d)
    A = oracle.rbf(X, N, d, b, ell, 1.0, sn2)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    return A, rhs


@pytest.mark.parametrize("N,b,shift", [(5, 5, 0.0), (1000, 64, 0.5), (777, 37, 1e-3), (4096, 256, 0.0)])
def test_to_f32_packed_rounds_each_element(hs, dev, N, b, shift):
    """The FP64 -> FP32 copy of the packed matrix (vector path for aligned
    buffers and a length divisible by 4, element path for the rest,
    e.g. 25 elements at N = b = 5) rounds every element and adds the shift
    on the diagonal entries below N only."""
    n = -(-N // b)
    rng = np.random.default_rng(N + b)
    A = rng.standard_normal(packed_len(N, b))
    dA, dL = dev.upload(A), dev.alloc(packed_len(N, b) * 4)
    assert hs.call("hs_to_f32_packed", dev.h, dA.p, dL.p, N, b, shift) == 0
    L = dev.download(dL, packed_len(N, b), dtype=np.float32)
    ref = A.copy()
    for j in range(n):
        base = (j * n - j * (j - 1) // 2) * b * b
        for i in range(min(b, N - j * b)):
            ref[base + i * b + i] += shift
    assert np.array_equal(L, ref.astype(np.float32))


# variants: 0 FP32, 1 BF16 hi+lo, 2 scaled FP16 hi+lo
VARIANTS = [0, 1, 2]
# ||L L^T - K|| / ||K|| bars: FP32 storage alone gives ~6e-8; the bf16 hi+lo
# products add ~2^-17 relative, the fp16 ones (22-bit operands) ~2^-23
RECON_BAR = {0: 2e-6, 1: 2e-5, 2: 2e-6}
FACTOR_BAR = {0: 1e-4, 1: 1e-3, 2: 1e-4}


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("N,b", [(1024, 128), (1536, 256), (1280, 256), (600, 128), (4096, 256), (3000, 256),
                                 (4096, 128)])
def test_mixed_factor_reconstructs(hs, dev, oracle, variant, N, b):
    print("mixed factor variant=%d N=%d b=%d" % (variant, N, b), flush=True)
    A, _ = system(oracle, N, 8, b, 1.0, 1e-1, 3)
    dA, dL = dev.upload(A), dev.alloc(packed_len(N, b) * 4)
    info = hs.call("hs_cholesky_mixed", dev.h, dA.p, dL.p, N, b, variant, 0.0)
    assert info == 0
    L = dev.download(dL, packed_len(N, b), dtype=np.float32).astype(np.float64)
    Ld = packed_to_dense(L, N, b, lower_only=True)
    K = packed_to_dense(A, N, b)
    err = np.linalg.norm(Ld @ Ld.T - K) / np.linalg.norm(K)
    print("  ||LL^T - K|| / ||K|| = %.2e" % err)
    assert err < RECON_BAR[variant]
    # the factor agrees with the FP64 oracle's to low-precision accuracy
    Lref = packed_to_dense(oracle.cholesky(A, N, b), N, b, lower_only=True)
    assert np.max(np.abs(Ld - Lref)) < FACTOR_BAR[variant] * np.max(np.abs(Lref))


@pytest.mark.parametrize("c", [2.0 ** -30 * 3.0, 1e-7, 1.0, 1e5, 2.0 ** 40 * 5.0])
@pytest.mark.parametrize("N,b", [(1536, 256), (1280, 128)])
def test_f16_split_scale_follows_the_matrix(hs, dev, oracle, c, N, b):
    """The fp16 split scales the panels by a power of two from the largest
    diagonal entry, so the factor of c K is sqrt(c) times the factor of K to
    the same relative accuracy whatever c is: well outside the fp16 range
    (65504 down to 6e-8) on both sides, where an unscaled split would
    overflow or lose the lo parts to underflow."""
    print("f16 split scale c=%.3e N=%d b=%d" % (c, N, b), flush=True)
    A, _ = system(oracle, N, 8, b, 1.0, 1e-1, 5)
    A = A * c
    dA, dL = dev.upload(A), dev.alloc(packed_len(N, b) * 4)
    assert hs.call("hs_cholesky_mixed", dev.h, dA.p, dL.p, N, b, 2, 0.0) == 0
    L = dev.download(dL, packed_len(N, b), dtype=np.float32).astype(np.float64)
    Ld = packed_to_dense(L, N, b, lower_only=True)
    K = packed_to_dense(A, N, b)
    err = np.linalg.norm(Ld @ Ld.T - K) / np.linalg.norm(K)
    print("  ||LL^T - K|| / ||K|| = %.2e" % err, flush=True)
    assert np.all(np.isfinite(Ld))
    assert err < RECON_BAR[2]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("N,b,ell,sn2", [(2048, 128, 1.0, 1e-2), (2048, 256, 2.0, 1e-3)])
def test_refine_reaches_fp64_backward_error(hs, dev, oracle, variant, N, b, ell, sn2):
    print("refine variant=%d N=%d b=%d" % (variant, N, b), flush=True)
    A, rhs = system(oracle, N, 16, b, ell, sn2, 7)
    dA, dL, dy, dx = dev.upload(A), dev.alloc(packed_len(N, b) * 4), dev.upload(rhs), dev.zeros(len(rhs))
    assert hs.call("hs_cholesky_mixed", dev.h, dA.p, dL.p, N, b, variant, 0.0) == 0
    rep = hs.RefineReport()
    hs.call("hs_refine_solve", dev.h, dA.p, dL.p, dy.p, dx.p, N, b, 1e-12, 30, 200, C.byref(rep))
    x = dev.download(dx, len(rhs))
    xref = oracle.tri_solve(oracle.cholesky(A, N, b), rhs, N, b)
    K = packed_to_dense(A, N, b)
    r = rhs[:N] - K @ x[:N]
    omega = np.abs(r).max() / (np.abs(K).sum(1).max() * np.abs(x).max() + np.abs(rhs).max())
    fwd = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    print("  report", rep.as_dict(), "host omega %.2e forward diff %.2e" % (omega, fwd))
    assert rep.converged == 1 and rep.backward_error <= 1e-12
    assert omega <= 2e-12
    assert fwd <= 1e-8
    assert np.all(x[N:] == 0.0)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("N,b,ell,sn2", [(2048, 256, 2.0, 1e-3), (3000, 128, 1.0, 1e-2)])
def test_refine_to_fp64_relative_residual(hs, dev, oracle, variant, N, b, ell, sn2):
    """Equal residual: refinement run to the relative residual the FP64
    factor + solves reach on the same system (and the 1e-12 backward error),
    checked on the host against the dense K."""
    print("refine to equal residual variant=%d N=%d b=%d" % (variant, N, b), flush=True)
    A, rhs = system(oracle, N, 16, b, ell, sn2, 9)
    K = packed_to_dense(A, N, b)
    # the FP64 path on the GPU: factor + solves, its relative residual
    dA, dx64 = dev.upload(A), dev.upload(rhs)
    assert hs.call("hs_cholesky_f64", dev.h, dA.p, N, b) == 0
    hs.call("hs_tri_solve_f64", dev.h, dA.p, dx64.p, N, b)
    x64 = dev.download(dx64, len(rhs))
    rr64 = np.linalg.norm(rhs[:N] - K @ x64[:N]) / np.linalg.norm(rhs[:N])
    dK, dL, dy, dx = dev.upload(A), dev.alloc(packed_len(N, b) * 4), dev.upload(rhs), dev.zeros(len(rhs))
    assert hs.call("hs_cholesky_mixed", dev.h, dK.p, dL.p, N, b, variant, 0.0) == 0
    rep = hs.RefineReport()
    hs.call("hs_refine_solve_to", dev.h, dK.p, dL.p, dy.p, dx.p, N, b, 1e-12, rr64, 60, 600, C.byref(rep))
    x = dev.download(dx, len(rhs))
    rr = np.linalg.norm(rhs[:N] - K @ x[:N]) / np.linalg.norm(rhs[:N])
    fwd = np.linalg.norm(x - x64) / np.linalg.norm(x64)
    print("  FP64 relative residual %.3e; mixed %.3e (device %.3e); report %s; forward diff vs FP64 %.2e"
          % (rr64, rr, rep.relative_residual, rep.as_dict(), fwd), flush=True)
    assert rep.converged == 1
    assert rep.relative_residual <= rr64
    # the host recomputation (different summation order) within rounding of it
    assert rr <= 2 * rr64
    assert np.all(x[N:] == 0.0)


def test_pcg_fallback_and_shift(hs, dev, oracle):
    """A shifted factor makes refinement contract slowly; the solve must switch
    to PCG and still reach the target. max_ir = 0 goes to PCG at once."""
    N, b = 2048, 128
    A, rhs = system(oracle, N, 16, b, 2.0, 1e-3, 11)
    K = packed_to_dense(A, N, b)
    knorm = np.abs(K).sum(1).max()
    dA, dL, dy, dx = dev.upload(A), dev.alloc(packed_len(N, b) * 4), dev.upload(rhs), dev.zeros(len(rhs))
    for shift, max_ir in ((1e-3 * knorm, 30), (0.0, 0)):
        assert hs.call("hs_cholesky_mixed", dev.h, dA.p, dL.p, N, b, 0, shift) == 0
        rep = hs.RefineReport()
        hs.call("hs_refine_solve", dev.h, dA.p, dL.p, dy.p, dx.p, N, b, 1e-12, max_ir, 300, C.byref(rep))
        print("  shift %.2e max_ir %d:" % (shift, max_ir), rep.as_dict())
        assert rep.used_pcg == 1 and rep.pcg_iterations > 0 and rep.converged == 1
        x = dev.download(dx, len(rhs))
        r = rhs[:N] - K @ x[:N]
        omega = np.abs(r).max() / (knorm * np.abs(x).max() + np.abs(rhs).max())
        assert omega <= 2e-12
        # the backward error target bounds the forward error by cond(K) * omega
        xref = np.linalg.solve(K, rhs[:N])
        cond = np.linalg.cond(K)
        fwd = np.linalg.norm(x[:N] - xref) / np.linalg.norm(xref)
        print("  cond %.2e forward diff %.2e" % (cond, fwd))
        assert fwd <= 10 * cond * 1e-12


@pytest.mark.parametrize("variant", [1, 2])
def test_mixed_reports_bad_pivot(hs, dev, variant):
    N, b = 512, 128
    K = np.eye(N)
    K[300, 300] = -2.0
    A = dense_to_packed(K, b)
    dA, dL = dev.upload(A), dev.alloc(packed_len(N, b) * 4)
    assert hs.call("hs_cholesky_mixed", dev.h, dA.p, dL.p, N, b, variant, 0.0) == 301


@pytest.mark.parametrize("entry,tol,scale", [("hs_syrk_gemm_trailing_f32", 2e-6, 1.0),
                                             ("hs_syrk_gemm_trailing_bf16x3", 2e-5, 1.0),
                                             ("hs_syrk_gemm_trailing_f16x3", 2e-6, 1.0),
                                             ("hs_syrk_gemm_trailing_f16x3", 2e-6, 3.0e6),
                                             ("hs_syrk_gemm_trailing_f16x3", 2e-6, 1.7e-9)])
@pytest.mark.parametrize("N,b,k", [(1024, 128, 2), (1536, 256, 0), (1280, 256, 3)])
def test_trailing_step_entries(hs, dev, entry, tol, scale, N, b, k):
    """One fused trailing-update step on FP32 storage: A_IJ -= L_Ik L_Jk^T for
    k < J <= I < n (lower triangle of the diagonal blocks), against the same
    update in FP64 on the stored FP32 values; every other entry untouched."""
    n = -(-N // b)
    rng = np.random.default_rng(N + b + k)
    # scale: panel entries far outside the fp16 range, which the fp16 entry
    # scales back into it from the panel's largest |entry|; the rest of the
    # matrix at scale^2, the size of the products it receives
    A = scale * scale * rng.uniform(-1, 1, packed_len(N, b))
    p0 = (k * n - k * (k - 1) // 2) * b * b
    A[p0:p0 + (n - k) * b * b] /= scale
    A = A.astype(np.float32)
    dA = dev.upload(A)
    assert hs.call(entry, dev.h, dA.p, b, k, n) == 0
    got = dev.download(dA, len(A), dtype=np.float32).astype(np.float64)
    ref = A.astype(np.float64)

    def blk(I, J):
        s = (J * n - J * (J - 1) // 2 + (I - J)) * b * b
        return slice(s, s + b * b)

    for J in range(k + 1, n):
        LJ = ref[blk(J, k)].reshape(b, b)
        for I in range(J, n):
            LI = ref[blk(I, k)].reshape(b, b)
            upd = ref[blk(I, J)].reshape(b, b) - LI @ LJ.T
            if I == J:
                upd = np.where(np.tril(np.ones((b, b))) > 0, upd, ref[blk(I, J)].reshape(b, b))
            ref[blk(I, J)] = upd.reshape(-1)
    # b products of panel entries <= scale added to entries <= scale^2: the
    # error bar scales with b scale^2
    assert np.max(np.abs(got - ref)) <= tol * b * scale * scale
    changed = np.zeros(len(A), dtype=bool)
    for J in range(k + 1, n):
        for I in range(J, n):
            changed[blk(I, J)] = True
    assert np.array_equal(got[~changed], A.astype(np.float64)[~changed])
