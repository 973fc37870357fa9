"""The allocation-free CG entry (hs_cg_f64_noalloc) through the C ABI: its
scratch is the library's fixed device array, so it runs on a handle with no
workspace reserved, and it returns what hs_cg_f64 (the same kernels on the
reserved workspace) returns, bit for bit, on two calls in a row, the second
on a smaller system. Sizes beyond the fixed scratch are refused with an
error and nothing allocated."""
import ctypes as C

import numpy as np
import pytest

from hs_testlib import synth_gp_numpy

pytestmark = pytest.mark.gpu


def _system(oracle, N, b, seed):
    X, y = synth_gp_numpy(seed, N, 8)
    A = oracle.rbf(X, N, 8, b, 1.0, 1.0, 1e-2)
    n = -(-N // b)
    rhs = np.zeros(n * b)
    rhs[:N] = y
    return A, rhs


def _cg(hs, dev, name, dA, drhs, N, b):
    n = -(-N // b)
    dx = dev.zeros(n * b)
    it, df, d0 = C.c_int64(), C.c_double(), C.c_double()
    rc = getattr(hs.load(), name)(dev.h, dA.p, drhs.p, dx.p, N, b, 1e-10, 100000, C.byref(it), C.byref(df),
                                  C.byref(d0))
    assert rc == 0, "%s: %d %s" % (name, rc, hs.last_error())
    return dev.download(dx, n * b), it.value, df.value, d0.value


def test_noalloc_cg_matches_reserved_cg_on_two_calls(hs, oracle):
    lim = C.c_size_t()
    hs.call("hs_cg_noalloc_limit", C.byref(lim))
    print("no-alloc CG scratch limit %d bytes" % lim.value, flush=True)
    with hs.Device(0) as bare:  # no workspace reserved at all
        assert bare.workspace_bytes() == 0
        for N, b, seed in ((8192, 128, 3), (3000, 64, 4)):  # the second call on a smaller system
            print("no-alloc CG N=%d b=%d" % (N, b), flush=True)
            A, rhs = _system(oracle, N, b, seed)
            with hs.Device(0, reserve=[(N, b)]) as ref:
                rA, rrhs = ref.upload(A), ref.upload(rhs)
                want = _cg(hs, ref, "hs_cg_f64", rA, rrhs, N, b)
            dA, drhs = bare.upload(A), bare.upload(rhs)
            got = _cg(hs, bare, "hs_cg_f64_noalloc", dA, drhs, N, b)
            print("  iterations %d (reserved path %d)" % (got[1], want[1]), flush=True)
            assert got[1] == want[1] and got[2] == want[2] and got[3] == want[3]
            assert np.array_equal(got[0], want[0])
            assert bare.workspace_bytes() == 0


def test_noalloc_cg_refuses_sizes_beyond_its_scratch(hs):
    with hs.Device(0) as bare:
        N, b = 40000, 128  # 2 n(n+1)/2 b doubles of partials alone: ~100 MB
        n = -(-N // b)
        dA, dv = bare.alloc(8), bare.zeros(n * b)
        it = C.c_int64()
        rc = hs.load().hs_cg_f64_noalloc(bare.h, dA.p, dv.p, dv.p, N, b, 1e-10, 10, C.byref(it), None, None)
        assert rc == hs.HS_ERR_UNSUPPORTED
        assert "beyond its fixed" in hs.last_error()
        assert bare.workspace_bytes() == 0
