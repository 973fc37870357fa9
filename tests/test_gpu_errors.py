"""Argument validation of the main C-ABI entries on a live handle: empty
systems, unsupported block sizes and null pointers come back as the
documented negative codes (argument index), no device work is done, and the
handle stays usable afterwards."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_entry_argument_codes(hs, dev):
    lib = hs.load()
    buf = dev.zeros(4096)
    p = buf.p
    it = C.c_int64()
    cases = [
        ("hs_cholesky_f64 N=0", lib.hs_cholesky_f64(dev.h, p, 0, 64), -3),
        ("hs_cholesky_f64 b=0", lib.hs_cholesky_f64(dev.h, p, 64, 0), -4),
        ("hs_cholesky_f64 A=NULL", lib.hs_cholesky_f64(dev.h, None, 64, 64), -2),
        ("hs_tri_solve_f64 N=0", lib.hs_tri_solve_f64(dev.h, p, p, 0, 64), -4),
        ("hs_tri_solve_f64 x=NULL", lib.hs_tri_solve_f64(dev.h, p, None, 64, 64), -3),
        ("hs_trsv_packed_f64 trans=2", lib.hs_trsv_packed_f64(dev.h, p, p, 64, 64, 2), -6),
        ("hs_symv_packed_f64 N=0", lib.hs_symv_packed_f64(dev.h, p, p, p, 0, 64), -5),
        ("hs_symv_packed_f64 y=NULL", lib.hs_symv_packed_f64(dev.h, p, p, None, 64, 64), -4),
        ("hs_cg_f64 N=0", lib.hs_cg_f64(dev.h, p, p, p, 0, 64, 1e-10, 10, C.byref(it), None, None), -5),
        ("hs_cg_f64 b=96", lib.hs_cg_f64(dev.h, p, p, p, 64, 96, 1e-10, 10, C.byref(it), None, None), -6),
        ("hs_cg_f64 eps<0", lib.hs_cg_f64(dev.h, p, p, p, 64, 64, -1.0, 10, C.byref(it), None, None), -7),
        ("hs_cg_f64 imax<0", lib.hs_cg_f64(dev.h, p, p, p, 64, 64, 1e-10, -1, C.byref(it), None, None), -8),
        ("hs_rbf_assemble_f64 d=0", lib.hs_rbf_assemble_f64(dev.h, p, 64, 0, 64, 1.0, 1.0, 0.0, p), -4),
        ("hs_rbf_assemble_f64 ell=0", lib.hs_rbf_assemble_f64(dev.h, p, 64, 8, 64, 0.0, 1.0, 0.0, p), -6),
        ("hs_cholesky_mixed b=64", lib.hs_cholesky_mixed(dev.h, p, p, 64, 64, hs.HS_MIXED_FP32, 0.0), -5),
        ("hs_cholesky_mixed variant", lib.hs_cholesky_mixed(dev.h, p, p, 128, 128, 7, 0.0), -6),
        ("hs_gp_predict_f64 M=0", lib.hs_gp_predict_f64(dev.h, p, 8, p, 0, 8, 1.0, 1.0, p, p), -5),
    ]
    bad = [(name, got, want) for name, got, want in cases if got != want]
    assert not bad, bad
    assert lib.hs_last_error_string() is not None
    # the handle still works: a tiny factorisation of the identity
    N, b = 64, 64
    I = np.eye(N).reshape(-1)
    dI = dev.upload(I)
    assert lib.hs_cholesky_f64(dev.h, dI.p, N, b) == 0
    assert np.allclose(dev.download(dI, N * N), I)


def test_refine_argument_codes(hs, dev):
    lib = hs.load()
    buf = dev.zeros(4096)
    p = buf.p
    rep = hs.RefineReport()
    assert lib.hs_refine_solve(dev.h, p, p, p, p, 0, 128, 1e-12, 5, 5, C.byref(rep)) == -6
    assert lib.hs_refine_solve(dev.h, p, p, p, p, 128, 96, 1e-12, 5, 5, C.byref(rep)) == -7
    assert lib.hs_refine_solve(dev.h, p, p, p, p, 128, 128, 0.0, 5, 5, C.byref(rep)) == -8
    assert lib.hs_refine_solve(dev.h, p, p, p, p, 128, 128, 1e-12, -1, 5, C.byref(rep)) == -9
    # the relative-residual form: at least one target, none negative
    assert lib.hs_refine_solve_to(dev.h, p, p, p, p, 128, 128, 0.0, 0.0, 5, 5, C.byref(rep)) == -8
    assert lib.hs_refine_solve_to(dev.h, p, p, p, p, 128, 128, 1e-12, -1.0, 5, 5, C.byref(rep)) == -8
