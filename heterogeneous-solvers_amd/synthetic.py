"""Synthetic GP regression inputs of the measured workloads.

SplitMix64 stream from `seed`; uniforms u = (z >> 11) * 2^-53; X = 2u - 1
drawn first (row-major N x d), then N standard normals by Box-Muller from
two uniforms each, sqrt(-2 ln(1 - u1)) cos(2 pi u2); targets
y_i = sin(sum_k x_ik) + 0.1 xi_i.
"""
import numpy as np

# name -> (seed, N, d, matrix_bsz, lengthscale, signal variance, noise variance)
WORKLOADS = {
    "cholesky_small": (1, 4096, 8, 128, 1.0, 1.0, 1e-2),
    "cg": (2, 32768, 8, 128, 1.0, 1.0, 1e-2),
    "cholesky": (3, 65536, 16, 256, 2.0, 1.0, 1e-3),
    "adjoint": (5, 16384, 8, 128, 1.0, 1.0, 1e-2),
}


def splitmix64_uniforms(seed, count):
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def gp_inputs(seed, N, d):
    u = splitmix64_uniforms(seed, N * d + 2 * N)
    X = (2.0 * u[:N * d] - 1.0).reshape(N, d)
    u1 = u[N * d::2][:N]
    u2 = u[N * d + 1::2][:N]
    xi = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)
    y = np.sin(X.sum(axis=1)) + 0.1 * xi
    return X, y
