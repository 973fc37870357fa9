"""Shared test helpers: the oracle (CPU restatement, oracle/liboracle.so),
the reference's text fixture format, dense <-> packed conversion, and the
synthetic GP workload generator written independently in numpy."""
import ctypes as C
import importlib.util
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_P = C.c_void_p
_I = C.c_int
_L = C.c_longlong
_D = C.c_double


def load_hs():
    """Import heterogeneous-solvers_amd/hs_amd.py (the directory name is not a
    Python identifier, so it is loaded by path)."""
    if "hs_amd" in sys.modules:
        return sys.modules["hs_amd"]
    path = os.path.join(ROOT, "heterogeneous-solvers_amd", "hs_amd.py")
    spec = importlib.util.spec_from_file_location("hs_amd", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    sys.modules["hs_amd"] = mod
    return mod


def dptr(a):
    return a.ctypes.data_as(_P)


class Oracle:
    """ctypes view of oracle/liboracle.so (CPU FP64 restatement)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(path)
        sig = {
            "or_block_id": ([_I, _I, _I], _L),
            "or_packed_elems": ([_L, _I], _L),
            "or_potrf": ([_P, _L, _I, _I, _I], None),
            "or_trsm": ([_P, _I, _I, _I, _I, _I], None),
            "or_syrk": ([_P, _I, _I, _I, _I, _I, _I], None),
            "or_gemm": ([_P, _I, _I, _I, _I, _I, _I], None),
            "or_symv": ([_P, _P, _P, _I, _I, _I, _I, _I, _I, _I], None),
            "or_trsv_block": ([_P, _P, _L, _I, _I, _I, _I], None),
            "or_col_update": ([_P, _P, _I, _I, _I, _I, _I, _I, _I], None),
            "or_dot": ([_P, _P, _I, _I, _I, _I], _D),
            "or_scale": ([_P, _D, _P, _I, _I, _I], None),
            "or_add": ([_P, _P, _P, _I, _I, _I], None),
            "or_sub": ([_P, _P, _P, _I, _I, _I], None),
            "or_scale_add": ([_P, _D, _P, _P, _I, _I, _I], None),
            "or_cholesky": ([_P, _L, _I], None),
            "or_tri_solve": ([_P, _P, _L, _I], None),
            "or_cg": ([_P, _P, _P, _L, _I, _D, _L, _I, C.POINTER(_D)], _L),
            "or_rbf_window": ([_P, _L, _I, _P], None),
            "or_rbf": ([_P, _L, _I, _I, _D, _D, _D, _P], None),
            "or_synth_gp": ([C.c_uint64, _L, _I, _P, _P], None),
            "or_window_features": ([_P, _L, _P], None),
            "or_rbf_cross": ([_P, _L, _P, _L, _I, _D, _D, _P], None),
            "or_gemv_gp": ([_P, _P, _P, _L, _L], None),
            "or_num_threads": ([], _I),
            "or_dot_tree": ([_P, _P, _I, _I, _I, _I, _I, _I], _D),
            "or_symv_ordered": ([_P, _P, _P, _I, _I, _I], None),
            "or_cg_gpu_order": ([_P, _P, _P, _L, _I, _D, _L, _I, _I, _I, C.POINTER(_D)], _L),
            "or_potrf_gpu_order": ([_P, _L, _I, _I, _I, _I], None),
            "or_trsm_gpu_order": ([_P, _I, _I, _I, _I, _I, _I], None),
            "or_syrk_gpu_order": ([_P, _I, _I, _I, _I, _I, _I, _I], None),
            "or_gemm_gpu_order": ([_P, _I, _I, _I, _I, _I, _I, _I], None),
            "or_cholesky_gpu_order": ([_P, _L, _I, _I], None),
            "or_tri_solve_gpu_order": ([_P, _P, _L, _I, _I], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        self.L = L

    def __getattr__(self, name):
        return getattr(self.L, name)

    # convenience wrappers on numpy arrays
    def cholesky(self, A, N, b):
        A = A.copy()
        self.L.or_cholesky(dptr(A), N, b)
        return A

    def cholesky_gpu_order(self, A, N, b, fma=0):
        """Blocked Cholesky in the order of the reference's share-1.0 GPU path."""
        A = A.copy()
        self.L.or_cholesky_gpu_order(dptr(A), N, b, fma)
        return A

    def tri_solve_gpu_order(self, L_, x, N, b, fma=0):
        x = x.copy()
        self.L.or_tri_solve_gpu_order(dptr(L_), dptr(x), N, b, fma)
        return x

    def tri_solve(self, L_, x, N, b):
        x = x.copy()
        self.L.or_tri_solve(dptr(L_), dptr(x), N, b)
        return x

    def cg(self, A, rhs, N, b, eps, imax=100000, chunks=8):
        x = np.zeros_like(rhs)
        delta = _D()
        it = self.L.or_cg(dptr(A), dptr(rhs), dptr(x), N, b, eps, imax, chunks, C.byref(delta))
        return x, it, delta.value

    def cg_gpu_order(self, A, rhs, N, b, eps, imax=100000, fma=0, wg=128, wgf=512):
        """CG in the order of the reference's share-1.0 GPU path."""
        x = np.zeros_like(rhs)
        delta = _D()
        it = self.L.or_cg_gpu_order(dptr(A), dptr(rhs), dptr(x), N, b, eps, imax, wg, wgf, fma, C.byref(delta))
        return x, it, delta.value

    def dot_tree(self, x, y, b, fma=0, wg=128, wgf=512, block_start=0, block_count=None):
        n = len(x) // b
        return self.L.or_dot_tree(dptr(x), dptr(y), b, block_start, n if block_count is None else block_count,
                                  wg, wgf, fma)

    def symv_ordered(self, A, x, b, n, fma=0):
        y = np.zeros(n * b)
        self.L.or_symv_ordered(dptr(A), dptr(x), dptr(y), b, n, fma)
        return y

    def symv(self, A, x, b, n):
        y = np.zeros(n * b)
        self.L.or_symv(dptr(A), dptr(x), dptr(y), b, 0, 0, n, n, n, 1)
        return y

    def rbf(self, X, N, d, b, ell, sf2, sn2):
        A = np.zeros(packed_len(N, b))
        self.L.or_rbf(dptr(np.ascontiguousarray(X)), N, d, b, ell, sf2, sn2, dptr(A))
        return A

    def rbf_cross(self, Xtr, Xte, ell, sf2):
        Xtr, Xte = np.ascontiguousarray(Xtr, dtype=np.float64), np.ascontiguousarray(Xte, dtype=np.float64)
        N, d = Xtr.shape
        M = Xte.shape[0]
        K = np.empty(M * N)
        self.L.or_rbf_cross(dptr(Xtr), N, dptr(Xte), M, d, ell, sf2, dptr(K))
        return K.reshape(M, N)

    def gemv_gp(self, A, x):
        A = np.ascontiguousarray(A, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        m, n = A.shape
        y = np.empty(m)
        self.L.or_gemv_gp(dptr(A), dptr(x), dptr(y), n, m)
        return y

    def window(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        X = np.empty(len(t) * 8)
        self.L.or_window_features(dptr(t), len(t), dptr(X))
        return X.reshape(len(t), 8)

    def synth(self, seed, N, d):
        X = np.empty(N * d)
        y = np.empty(N)
        self.L.or_synth_gp(seed, N, d, dptr(X), dptr(y))
        return X.reshape(N, d), y


def nblocks(N, b):
    return -(-N // b)


def packed_len(N, b):
    n = nblocks(N, b)
    return n * (n + 1) // 2 * b * b


def block_id(I, J, n):
    return J * n - J * (J - 1) // 2 + (I - J)


def dense_to_packed(M, b):
    """Lower block triangle of a dense symmetric matrix into the packed layout
    (diagonal blocks full), zero padded."""
    N = M.shape[0]
    n = nblocks(N, b)
    P = np.zeros(n * b)
    Mp = np.zeros((n * b, n * b))
    Mp[:N, :N] = M
    out = np.zeros(packed_len(N, b))
    for J in range(n):
        for I in range(J, n):
            blk = Mp[I * b:(I + 1) * b, J * b:(J + 1) * b]
            s = block_id(I, J, n) * b * b
            out[s:s + b * b] = blk.reshape(-1)
    del P
    return out


def packed_to_dense(A, N, b, lower_only=False):
    """Dense matrix from the packed layout; lower_only takes the lower
    triangle of each diagonal block (as after a factorisation)."""
    n = nblocks(N, b)
    M = np.zeros((n * b, n * b))
    for J in range(n):
        for I in range(J, n):
            s = block_id(I, J, n) * b * b
            blk = A[s:s + b * b].reshape(b, b)
            if I == J and lower_only:
                blk = np.tril(blk)
            M[I * b:(I + 1) * b, J * b:(J + 1) * b] = blk
    if not lower_only:
        low = np.tril(M, -1)
        M = np.tril(M) + low.T
        # diagonal blocks were stored full; keep them as stored
        for J in range(n):
            s = block_id(J, J, n) * b * b
            M[J * b:(J + 1) * b, J * b:(J + 1) * b] = A[s:s + b * b].reshape(b, b)
    return M[:N, :N]


# ---- the reference's text formats (MatrixParser.cpp:11-167) ---------------

def parse_row_values(s):
    """getRowValuesFromString (MatrixParser.cpp:74-105): ';'-separated values;
    a value that is not a number raises ValueError."""
    out = []
    for tok in s.split(";"):
        t = tok.strip()
        try:
            out.append(float(t))
        except ValueError:
            raise ValueError("Error while parsing string '%s'" % tok)
    return out


def parse_symmetric_matrix_text(path):
    """`# N` header then N lower-triangle rows; returns the dense matrix."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    if not lines or not lines[0].startswith("#"):
        raise ValueError("Invalid matrix format. First line has to be '# <N>!'")
    N = int(lines[0][2:])
    M = np.zeros((N, N))
    for i, row in enumerate(lines[1:1 + N]):
        vals = parse_row_values(row)
        assert len(vals) == i + 1
        M[i, :i + 1] = vals
        M[:i + 1, i] = vals
    return M


def parse_rhs_text(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    if not lines or not lines[0].startswith("#"):
        raise ValueError("Invalid right-hand side format. First line has to be '# <N>!'")
    N = int(lines[0][2:])
    return np.array(parse_row_values(lines[1])[:N])


def write_matrix_text(path, K):
    """The reference's matrix format (MatrixParser.cpp:11-72): '# N', then
    row i's entries 0..i separated by ';', 17 significant digits (exact)."""
    N = K.shape[0]
    with open(path, "w") as fh:
        fh.write("# %d\n" % N)
        for i in range(N):
            fh.write(";".join("%.17g" % v for v in K[i, :i + 1]) + "\n")


def write_rhs_text(path, y):
    with open(path, "w") as fh:
        fh.write("# %d\n" % len(y))
        fh.write(";".join("%.17g" % v for v in y) + "\n")


def fixture_matrix():
    return parse_symmetric_matrix_text(os.path.join(GOLDEN_DIR, "data", "testMatrixSymmetric20x20.txt"))


def fixture_vector():
    return parse_rhs_text(os.path.join(GOLDEN_DIR, "data", "testVector_20.txt"))


# ---- synthetic GP workload, independent numpy SplitMix64 -----------------

_M64 = (1 << 64) - 1


def splitmix64_uniforms(seed, count):
    """count uniforms u = (z >> 11) * 2^-53 of the SplitMix64 stream (numpy
    uint64 arithmetic wraps mod 2^64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def synth_gp_numpy(seed, N, d):
    u = splitmix64_uniforms(seed, N * d + 2 * N)
    X = (2.0 * u[:N * d] - 1.0).reshape(N, d)
    u1 = u[N * d::2][:N]
    u2 = u[N * d + 1::2][:N]
    xi = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)
    y = np.sin(X.sum(axis=1)) + 0.1 * xi
    return X, y


def rbf_dense_numpy(X, ell, sf2, sn2):
    d2 = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    return sf2 * np.exp(-0.5 / ell ** 2 * d2) + sn2 * np.eye(X.shape[0])
