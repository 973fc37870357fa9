"""Replays the reference's GoogleTest cases (tests/golden/reference_goldens.json)
against a backend: the CPU oracle or the HIP library through the C ABI.

Each case holds the block size, the kernel calls with their arguments, how
the input vectors are filled and the expected values with the reference's
tolerance (EXPECT_NEAR ..., 1e-12). The call names map onto one backend op
each; the reference's plain / _GPU / _optimizedGPU / _CPU variants of one op
all map onto the same op, as they are pinned by the same golden arrays.
"""
import json
import math
import os
import re

import numpy as np

from hs_testlib import GOLDEN_DIR, dense_to_packed, fixture_matrix, fixture_vector, nblocks, parse_row_values

with open(os.path.join(GOLDEN_DIR, "reference_goldens.json")) as _fh:
    GOLDENS = json.load(_fh)

OPS = {
    "MatrixOperations::cholesky": "potrf",
    "MatrixOperations::cholesky_GPU": "potrf",
    "MatrixOperations::cholesky_optimizedGPU": "potrf",
    "MatrixMatrixOperations::triangularSolve": "trsm",
    "MatrixMatrixOperations::triangularSolve_optimizedGPU": "trsm",
    "MatrixMatrixOperations::triangularSolve_optimizedCPU": "trsm",
    "MatrixMatrixOperations::symmetricMatrixMatrixDiagonal": "syrk",
    "MatrixMatrixOperations::symmetricMatrixMatrixDiagonal_optimizedGPU": "syrk",
    "MatrixMatrixOperations::symmetricMatrixMatrixDiagonal_optimizedCPU": "syrk",
    "MatrixMatrixOperations::matrixMatrixStep": "gemm",
    "MatrixMatrixOperations::matrixMatrixStep_optimizedGPU": "gemm",
    "MatrixMatrixOperations::matrixMatrixStep_optimizedGPU2": "gemm",
    "MatrixMatrixOperations::matrixMatrixStep_optimizedCPU2": "gemm",
    "MatrixVectorOperations::matrixVectorBlock": "symv",
    "MatrixVectorOperations::matrixVectorBlock_GPU": "symv",
    "MatrixVectorOperations::triangularSolveBlockVector": "trsv",
    "MatrixVectorOperations::matrixVectorColumnUpdate": "colupd",
    "VectorOperations::scaleVectorBlock": "scale",
    "VectorOperations::addVectorBlock": "add",
    "VectorOperations::subVectorBlock": "sub",
    "VectorOperations::scaleAndAddVectorBlock": "scale_add",
    "VectorOperations::scalarProduct": "dot_tree",
    "VectorOperations::scalarProduct_CPU": "dot",
    "VectorOperations::sumFinalScalarProduct": None,
    "VectorOperations::sumFinalScalarProduct_CPU": None,
}


def kernel_cases():
    """Cases that run at least one kernel (the parser cases are separate)."""
    return [t for t in GOLDENS["tests"] if t["calls"]]


def factor_cases():
    """Cases whose every call is a blocked-Cholesky kernel (POTRF, TRSM,
    SYRK, GEMM)."""
    return [t for t in kernel_cases() if all(OPS.get(c[0]) in ("potrf", "trsm", "syrk", "gemm") for c in t["calls"])]


def parser_cases():
    return [t for t in GOLDENS["tests"] if not t["calls"]]


def case_id(t):
    return "%s.%s" % (t["suite"], t["name"])


def _arrays(t):
    arr = dict(GOLDENS["suites"].get(t["suite"], {}))
    arr.update(t["arrays"])
    return arr


class Env:
    """Evaluates the small integer expressions of a case."""

    def __init__(self, b, N, sizes):
        self.b, self.N = b, N
        self.n = nblocks(N, b)
        self.sizes = sizes

    def eval(self, expr):
        e = expr.strip()
        if e in ("true", "false"):
            return 1 if e == "true" else 0
        e = e.replace("A.blockCountXY", str(self.n)).replace("b.blockCountX", str(self.n))
        e = e.replace("b.rightHandSideData.size()", str(self.n * self.b))
        e = e.replace("A.matrixData.size()", str(self.sizes.get("A", 0)))
        for name, sz in self.sizes.items():
            e = e.replace(name + ".size()", str(sz))
        if not re.fullmatch(r"[\d\s+\-*.eE]+", e):
            raise ValueError("cannot evaluate " + expr)
        v = eval(e, {"__builtins__": {}})
        return v


def _buffer_name(arg):
    m = re.fullmatch(r"(\w+)\.(?:matrixData|rightHandSideData)\.data\(\)", arg)
    if m:
        return "A" if m.group(1) in ("A", "matrix") else m.group(1)
    m = re.fullmatch(r"(\w+)\.data\(\)", arg)
    if m:
        return m.group(1)
    return None


def setup_case(t):
    """Host arrays and parameters of a case before any call runs."""
    cfg = t["config"]
    b = cfg.get("matrixBlockSize")
    N = 20  # both fixtures are 20 x 20 / 20 long
    env = Env(b, N, {})
    arrays = {}
    for var, kind, _ in t["parse"]:
        if kind == "parseSymmetricMatrix":
            arrays["A"] = dense_to_packed(fixture_matrix(), b)
        else:
            v = np.zeros(env.n * b)
            v[:N] = fixture_vector()
            arrays[var] = v
    env.sizes = {k: len(v) for k, v in arrays.items()}
    for var, spec in t["vectors"].items():
        size = env.eval(spec["size"]) if "size" in spec else env.n * b
        v = np.full(int(size), spec.get("fill", 0.0))
        env.sizes[var] = len(v)
        for lo, hi, kind in spec.get("init", []):
            hi = int(env.eval(hi))
            for i in range(int(lo), hi):
                v[i] = 1.0 if kind == "one" else math.sqrt(i)
        arrays[var] = v
    return b, N, env, arrays


def run_case(t, backend):
    """Runs the calls of case t on `backend`, returns {name: final array}.
    scalarProduct (the GPU kernel, VectorOperations.cpp:84-139) runs as the
    work-group tree of the case's conf::workGroupSizeVector; the final
    reduction's work-group size is the configuration default (512)."""
    b, N, env, arrays = setup_case(t)
    wg = t["config"].get("workGroupSizeVector", 128)
    backend.begin(arrays)
    for fname, args in t["calls"]:
        op = OPS[fname]
        if op is None:
            continue
        args = [a for a in args if a != "queue"]
        bufs = [_buffer_name(a) for a in args]
        nums = [None if bufs[i] else (a if a == "workGroupCount" else env.eval(a)) for i, a in enumerate(args)]
        if op == "potrf":
            backend.potrf(bufs[0], N, b, int(nums[1]), int(nums[2]))
        elif op == "trsm":
            backend.trsm(bufs[0], b, *[int(x) for x in nums[1:5]])
        elif op in ("syrk", "gemm"):
            getattr(backend, op)(bufs[0], b, *[int(x) for x in nums[1:6]])
        elif op == "symv":
            reset = int(nums[8]) if len(args) > 8 else 1
            backend.symv(bufs[0], bufs[1], bufs[2], b, *[int(x) for x in nums[3:8]], reset)
        elif op == "trsv":
            backend.trsv(bufs[0], bufs[1], N, b, int(nums[2]), int(nums[3]), int(nums[4]))
        elif op == "colupd":
            backend.colupd(bufs[0], bufs[1], b, *[int(x) for x in nums[2:8]])
        elif op == "scale":
            backend.vec("scale", bufs[0], None, float(nums[1]), bufs[2], b, int(nums[3]), int(nums[4]))
        elif op in ("add", "sub"):
            backend.vec(op, bufs[0], bufs[1], 0.0, bufs[2], b, int(nums[3]), int(nums[4]))
        elif op == "scale_add":
            backend.vec("scale_add", bufs[0], bufs[2], float(nums[1]), bufs[3], b, int(nums[4]), int(nums[5]))
        elif op == "dot":
            backend.dot(bufs[0], bufs[1], bufs[2], b, int(nums[3]), int(nums[4]))
        elif op == "dot_tree":
            backend.dot_tree(bufs[0], bufs[1], bufs[2], b, int(nums[3]), int(nums[4]), wg)
    return backend.end(), env


def check_case(t, final, env):
    """Applies the case's expectations; returns a list of failure strings."""
    arrs = _arrays(t)
    fails = []
    for chk in t["checks"]:
        if "size_equal" in chk:
            a, c = [("A" if x in ("A.matrixData", "matrix.matrixData") else x) for x in chk["size_equal"]]
            la = len(final[a]) if a in final else len(arrs[a])
            lc = len(final[c]) if c in final else len(arrs[c])
            if la != lc:
                fails.append("size %s=%d vs %s=%d" % (a, la, c, lc))
            continue
        target = chk["target"].split(".")[0]
        target = "A" if target in ("A", "matrix") else target
        got = final[target]
        off = int(env.eval(chk["offset"]))
        tol = chk["tol"]
        if "expected_scalar" in chk:
            exp = chk["expected_scalar"]
            if exp == "resultValue":
                ka = t["known_answer"]
                exp = 0.0
                for i in range(ka["to"]):
                    if i >= ka["from"]:
                        exp += math.sqrt(i) * math.sqrt(i)
            vals = [(off, exp)]
        else:
            ref = arrs[chk["expected"]]
            cnt_expr = chk["count"]
            if cnt_expr == chk["expected"] + ".size()":
                cnt = len(ref)
            else:
                cnt = int(env.eval(cnt_expr.replace(chk["target"] + ".size()", str(len(got)))))
            vals = [(off + i, ref[i]) for i in range(cnt)]
        for idx, exp in vals:
            g = got[idx]
            if chk.get("double_eq"):
                ok = g == exp or abs(g - exp) <= 4 * np.spacing(abs(exp))
            else:
                ok = abs(g - exp) <= tol
            if not ok:
                fails.append("%s[%d]=%r expected %r (tol %g)" % (target, idx, g, exp, tol))
                if len(fails) > 8:
                    return fails
    return fails


def run_parser_case(t):
    """The parser / layout cases of testSymmetricMatrix.cpp, on the test
    helpers' parser (host code)."""
    fails = []
    if "row_string" in t:
        vals = parse_row_values(t["row_string"])
        ref = t["arrays"]["result"]
        if len(vals) != len(ref) or any(v != r for v, r in zip(vals, ref)):
            fails.append("row parse mismatch")
        return fails
    for chk in t["checks"]:
        if "throws" in chk:
            try:
                parse_row_values(chk["row_string"])
                fails.append("no exception for %r" % chk["row_string"])
            except ValueError:
                pass
            return fails
    b = t["config"]["matrixBlockSize"]
    A = dense_to_packed(fixture_matrix(), b)
    env = Env(b, 20, {"A": len(A)})
    for chk in t["checks"]:
        if "attr" in chk:
            got = {"blockSize": b, "blockCountXY": nblocks(20, b), "matrixData.size()": len(A)}[chk["attr"]]
            if got != chk["value"]:
                fails.append("%s=%d expected %d" % (chk["attr"], got, chk["value"]))
        elif "expected_scalar" in chk:
            idx = int(env.eval(chk["offset"]))
            if A[idx] != chk["expected_scalar"]:
                fails.append("A[%d]=%r expected %r" % (idx, A[idx], chk["expected_scalar"]))
    return fails
