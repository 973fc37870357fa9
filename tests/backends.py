"""Backends the golden cases run on: the CPU oracle and the HIP library
(through the C ABI, device memory, one handle)."""
import ctypes as C

import numpy as np

from hs_testlib import Oracle, dptr


class OracleBackend:
    def __init__(self, oracle=None, fma=0):
        self.o = oracle or Oracle()
        self.fma = fma

    def begin(self, arrays):
        self.a = {k: np.array(v, dtype=np.float64) for k, v in arrays.items()}

    def end(self):
        return self.a

    def potrf(self, A, N, b, bid, brow):
        self.o.or_potrf(dptr(self.a[A]), N, b, bid, brow)

    def trsm(self, A, b, bid, brow, bs, bc):
        self.o.or_trsm(dptr(self.a[A]), b, bid, brow, bs, bc)

    def syrk(self, A, b, bid, brow, bs, bc, n):
        self.o.or_syrk(dptr(self.a[A]), b, bid, brow, bs, bc, n)

    def gemm(self, A, b, bid, brow, bs, bc, n):
        self.o.or_gemm(dptr(self.a[A]), b, bid, brow, bs, bc, n)

    def symv(self, A, x, y, b, bsi, bsj, bci, bcj, n, reset):
        self.o.or_symv(dptr(self.a[A]), dptr(self.a[x]), dptr(self.a[y]), b, bsi, bsj, bci, bcj, n, reset)

    def trsv(self, A, x, N, b, brow, bid, tr):
        self.o.or_trsv_block(dptr(self.a[A]), dptr(self.a[x]), N, b, brow, bid, tr)

    def colupd(self, A, x, b, bs, bc, brow, bid, n, tr):
        self.o.or_col_update(dptr(self.a[A]), dptr(self.a[x]), b, bs, bc, brow, bid, n, tr)

    def vec(self, op, x, y, alpha, r, b, bs, bc):
        X, R = self.a[x], self.a[r]
        if op == "scale":
            self.o.or_scale(dptr(X), alpha, dptr(R), b, bs, bc)
        elif op == "add":
            self.o.or_add(dptr(X), dptr(self.a[y]), dptr(R), b, bs, bc)
        elif op == "sub":
            self.o.or_sub(dptr(X), dptr(self.a[y]), dptr(R), b, bs, bc)
        else:
            self.o.or_scale_add(dptr(X), alpha, dptr(self.a[y]), dptr(R), b, bs, bc)

    def dot(self, x, y, res, b, bs, bc):
        self.a[res][0] = self.o.or_dot(dptr(self.a[x]), dptr(self.a[y]), b, bs, bc, self.o.or_num_threads())

    def dot_tree(self, x, y, res, b, bs, bc, wg):
        self.a[res][0] = self.o.or_dot_tree(dptr(self.a[x]), dptr(self.a[y]), b, bs, bc, wg, 512, self.fma)


class OracleGpuOrderBackend(OracleBackend):
    """The oracle with the factor kernels in the order of the reference's
    share-1.0 GPU kernels (or_*_gpu_order)."""

    def potrf(self, A, N, b, bid, brow):
        self.o.or_potrf_gpu_order(dptr(self.a[A]), N, b, bid, brow, self.fma)

    def trsm(self, A, b, bid, brow, bs, bc):
        self.o.or_trsm_gpu_order(dptr(self.a[A]), b, bid, brow, bs, bc, self.fma)

    def syrk(self, A, b, bid, brow, bs, bc, n):
        self.o.or_syrk_gpu_order(dptr(self.a[A]), b, bid, brow, bs, bc, n, self.fma)

    def gemm(self, A, b, bid, brow, bs, bc, n):
        self.o.or_gemm_gpu_order(dptr(self.a[A]), b, bid, brow, bs, bc, n, self.fma)


class HipBackend:
    """Runs every op through libhs_amd.so on the GPU."""

    def __init__(self, hs, dev, fma=0):
        self.hs, self.dev = hs, dev
        self.fma = fma

    def begin(self, arrays):
        self.host = {k: np.array(v, dtype=np.float64) for k, v in arrays.items()}
        self.buf = {k: self.dev.upload(v) for k, v in self.host.items()}

    def end(self):
        self.dev.sync()
        out = {k: self.dev.download(self.buf[k], len(v)) for k, v in self.host.items()}
        for bf in self.buf.values():
            bf.free()
        return out

    def _c(self, name, *args):
        self.hs.call(name, self.dev.h, *args)

    def potrf(self, A, N, b, bid, brow):
        self._c("hs_potrf_diag_f64", self.buf[A].p, N, b, bid, brow)

    def trsm(self, A, b, bid, brow, bs, bc):
        self._c("hs_trsm_panel_f64", self.buf[A].p, b, bid, brow, bs, bc)

    def syrk(self, A, b, bid, brow, bs, bc, n):
        self._c("hs_syrk_diag_f64", self.buf[A].p, b, bid, brow, bs, bc, n)

    def gemm(self, A, b, bid, brow, bs, bc, n):
        self._c("hs_gemm_trailing_f64", self.buf[A].p, b, bid, brow, bs, bc, n)

    def symv(self, A, x, y, b, bsi, bsj, bci, bcj, n, reset):
        self._c("hs_symv_block_f64", self.buf[A].p, self.buf[x].p, self.buf[y].p, b, bsi, bsj, bci, bcj, n, reset)

    def trsv(self, A, x, N, b, brow, bid, tr):
        self._c("hs_trsv_block_f64", self.buf[A].p, self.buf[x].p, N, b, brow, bid, tr)

    def colupd(self, A, x, b, bs, bc, brow, bid, n, tr):
        self._c("hs_col_update_f64", self.buf[A].p, self.buf[x].p, b, bs, bc, brow, bid, n, tr)

    def vec(self, op, x, y, alpha, r, b, bs, bc):
        X, R = self.buf[x].p, self.buf[r].p
        if op == "scale":
            self._c("hs_vec_scale_f64", X, C.c_double(alpha), R, b, bs, bc)
        elif op == "add":
            self._c("hs_vec_add_f64", X, self.buf[y].p, R, b, bs, bc)
        elif op == "sub":
            self._c("hs_vec_sub_f64", X, self.buf[y].p, R, b, bs, bc)
        else:
            self._c("hs_vec_scale_add_f64", X, C.c_double(alpha), self.buf[y].p, R, b, bs, bc)

    def dot(self, x, y, res, b, bs, bc):
        self._c("hs_dot_f64", self.buf[x].p, self.buf[y].p, b, bs, bc, self.buf[res].p)

    def dot_tree(self, x, y, res, b, bs, bc, wg):
        self._c("hs_dot_ref_order_f64", self.buf[x].p, self.buf[y].p, b, bs, bc, wg, 512, self.fma,
                self.buf[res].p)


class HipRefOrderBackend(HipBackend):
    """HipBackend with the factor kernels of the reference-order Cholesky
    (hs_*_ref_order_f64)."""

    def potrf(self, A, N, b, bid, brow):
        self._c("hs_potrf_ref_order_f64", self.buf[A].p, N, b, bid, brow, self.fma)

    def trsm(self, A, b, bid, brow, bs, bc):
        self._c("hs_trsm_ref_order_f64", self.buf[A].p, b, bid, brow, bs, bc, self.fma)

    def syrk(self, A, b, bid, brow, bs, bc, n):
        self._c("hs_syrk_ref_order_f64", self.buf[A].p, b, bid, brow, bs, bc, n, self.fma)

    def gemm(self, A, b, bid, brow, bs, bc, n):
        self._c("hs_gemm_ref_order_f64", self.buf[A].p, b, bid, brow, bs, bc, n, self.fma)
