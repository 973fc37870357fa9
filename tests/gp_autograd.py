"""torch.autograd.Function over the C ABI's GP likelihood, for the checks.

GPNLL.apply(theta, X, y, b) returns the negative log marginal likelihood of
the RBF Gaussian process with theta = (log ell, log sf2, log sn2); backward
returns the gradient computed by hs_gp_nll_grad_f64 (reverse pass through the
blocked Cholesky on the device). X (N x d) and y (N) are float64 CUDA tensors;
theta is a float64 tensor of 3 on any device. This harness is test code: the
product entry point is the C function.
"""
import ctypes as C
import math

import torch

from hs_testlib import load_hs

hs = load_hs()

_dev = None


def device(N, b):
    """The harness's handle, its workspace reserved for (N, b) before the
    call (a set-up step; a no-op once reserved)."""
    global _dev
    if _dev is None:
        _dev = hs.Device(0)
    _dev.reserve(N, b)
    return _dev


def _buffers(N, b):
    n = -(-N // b)
    cnt = n * (n + 1) // 2 * b * b
    K = torch.empty(cnt, dtype=torch.float64, device="cuda")
    Kb = torch.empty(cnt, dtype=torch.float64, device="cuda")
    return K, Kb


def nll_only(theta, X, y, b):
    N, d = X.shape
    K, _ = _buffers(N, b)
    out = C.c_double()
    ell, sf2, sn2 = (math.exp(float(t)) for t in theta)
    torch.cuda.synchronize()
    rc = hs.call("hs_gp_nll_f64", device(N, b).h, C.c_void_p(X.data_ptr()), N, d, b, ell, sf2, sn2,
                 C.c_void_p(y.data_ptr()), C.c_void_p(K.data_ptr()), C.byref(out))
    if rc:
        raise RuntimeError("kernel matrix not positive definite at row %d" % rc)
    return out.value


def nll_grad(theta, X, y, b, want_kbar=False):
    N, d = X.shape
    K, Kb = _buffers(N, b)
    out = C.c_double()
    g = (C.c_double * 3)()
    ell, sf2, sn2 = (math.exp(float(t)) for t in theta)
    torch.cuda.synchronize()
    rc = hs.call("hs_gp_nll_grad_f64", device(N, b).h, C.c_void_p(X.data_ptr()), N, d, b, ell, sf2, sn2,
                 C.c_void_p(y.data_ptr()), C.c_void_p(K.data_ptr()), C.c_void_p(Kb.data_ptr()), C.byref(out), g)
    if rc:
        raise RuntimeError("kernel matrix not positive definite at row %d" % rc)
    return out.value, [g[0], g[1], g[2]], (Kb if want_kbar else None)


class GPNLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, X, y, b):
        nll, grad, _ = nll_grad(theta.detach().cpu().tolist(), X, y, b)
        ctx.save_for_backward(torch.tensor(grad, dtype=torch.float64, device=theta.device))
        return torch.tensor(nll, dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None
