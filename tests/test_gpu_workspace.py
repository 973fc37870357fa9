"""The handle workspace contract: sized once at set-up by
hs_workspace_reserve, never grown by a solver entry; an entry that needs more
than was reserved fails with HS_ERR_WORKSPACE instead of allocating; a
prepared factor gives the same solutions as the one-shot solve."""
import ctypes as C

import numpy as np
import pytest

from hs_testlib import synth_gp_numpy

pytestmark = pytest.mark.gpu


def system(oracle, N, d, b, sn2=1e-2, seed=1):
    X, y = synth_gp_numpy(seed, N, d)
    A = oracle.rbf(X, N, d, b, 1.0, 1.0, sn2)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    return X, A, rhs


@pytest.mark.parametrize("N,b", [(3000, 128), (2048, 256)])
def test_workspace_unchanged_across_solver_calls(hs, oracle, N, b):
    print("workspace contract N=%d b=%d" % (N, b), flush=True)
    X, A, rhs = system(oracle, N, 8, b)
    n = -(-N // b)
    q = C.c_size_t()
    hs.call("hs_workspace_query", N, b, 0, C.byref(q))
    with hs.Device(0, reserve=[(N, b)]) as dev:
        w0 = dev.workspace_bytes()
        assert w0 == q.value
        dA, drhs, dx = dev.upload(A), dev.upload(rhs), dev.zeros(n * b)
        it = C.c_int64()
        hs.call("hs_cg_f64", dev.h, dA.p, drhs.p, dx.p, N, b, 1e-10, 100000, C.byref(it), None, None)
        assert dev.workspace_bytes() == w0
        hs.call("hs_cg_ref_order_f64", dev.h, dA.p, drhs.p, dx.p, N, b, 1e-10, 100000, 128, 512, 1, C.byref(it),
                None, None, None)
        assert dev.workspace_bytes() == w0
        dr = dev.zeros(1)
        hs.call("hs_dot_f64", dev.h, drhs.p, drhs.p, b, 0, n, dr.p)
        hs.call("hs_dot_ref_order_f64", dev.h, drhs.p, drhs.p, b, 0, n, 128, 512, 0, dr.p)
        dy = dev.zeros(n * b)
        hs.call("hs_symv_packed_f64", dev.h, dA.p, drhs.p, dy.p, N, b)
        # FP32 factor + refinement on a copy of K, then the FP64 factor + solves
        dL = dev.alloc(len(A) * 4)
        assert hs.call("hs_cholesky_mixed", dev.h, dA.p, dL.p, N, b, hs.HS_MIXED_BF16X3, 0.0) == 0
        rep = hs.RefineReport()
        hs.call("hs_refine_solve", dev.h, dA.p, dL.p, drhs.p, dx.p, N, b, 1e-12, 30, 300, C.byref(rep))
        assert rep.converged == 1
        assert dev.workspace_bytes() == w0
        assert hs.call("hs_cholesky_f64", dev.h, dA.p, N, b) == 0
        dxs = dev.upload(rhs)
        hs.call("hs_tri_solve_f64", dev.h, dA.p, dxs.p, N, b)
        hs.call("hs_trsv_packed_f64", dev.h, dA.p, dxs.p, N, b, 0)
        assert dev.workspace_bytes() == w0


def test_entry_beyond_reservation_fails_without_allocating(hs, oracle):
    # reserved for one block (N = 128), called at N = 8192: the mat-vec
    # partials alone need more than every entry's need at N = 128
    N, b = 8192, 128
    small, big = C.c_size_t(), C.c_size_t()
    hs.call("hs_workspace_query", 128, b, 0, C.byref(small))
    hs.call("hs_workspace_query", N, b, 0, C.byref(big))
    assert big.value > 4 * small.value
    X, A, rhs = system(oracle, N, 8, b)
    with hs.Device(0, reserve=[(128, b)]) as dev:
        w0 = dev.workspace_bytes()
        dA, drhs, dx = dev.upload(A), dev.upload(rhs), dev.zeros(len(rhs))
        it = C.c_int64()
        rc = hs.load().hs_cg_f64(dev.h, dA.p, drhs.p, dx.p, N, b, 1e-10, 100, C.byref(it), None, None)
        assert rc == hs.HS_ERR_WORKSPACE
        assert "hs_workspace_reserve" in hs.last_error()
        # the factorisation's event pool was sized for n = 1 block as well
        rc = hs.load().hs_cholesky_f64(dev.h, dA.p, N, b)
        assert rc == hs.HS_ERR_WORKSPACE
        assert dev.workspace_bytes() == w0
        dev.reserve(N, b)
        assert dev.workspace_bytes() >= w0
        assert hs.call("hs_cg_f64", dev.h, dA.p, drhs.p, dx.p, N, b, 1e-10, 100000, C.byref(it), None, None) == 0


@pytest.mark.parametrize("N,b", [(3000, 64), (2300, 256), (4096, 128), (500, 20)])
def test_prepared_factor_solves_match_one_shot_solve(hs, oracle, N, b):
    print("prepared factor N=%d b=%d" % (N, b), flush=True)
    X, A, rhs = system(oracle, N, 8, b)
    n = -(-N // b)
    with hs.Device(0, reserve=[(N, b)]) as dev:
        dA = dev.upload(A)
        assert hs.call("hs_cholesky_f64", dev.h, dA.p, N, b) == 0
        q = C.c_size_t()
        hs.call("hs_tri_prepare_query", N, b, C.byref(q))
        dP = dev.alloc(max(8, q.value))
        hs.call("hs_tri_prepare_f64", dev.h, dA.p, N, b, dP.p)
        w0 = dev.workspace_bytes()
        rng = np.random.default_rng(b)
        for rep in range(3):
            v = np.zeros(n * b)
            v[:N] = rhs[:N] if rep == 0 else rng.standard_normal(N)
            d1, d2 = dev.upload(v), dev.upload(v)
            hs.call("hs_tri_solve_f64", dev.h, dA.p, d1.p, N, b)
            hs.call("hs_tri_solve_prepared_f64", dev.h, dA.p, dP.p, d2.p, N, b)
            assert np.array_equal(dev.download(d1, n * b), dev.download(d2, n * b))
        assert dev.workspace_bytes() == w0
