"""Race analysis of the two-stream look-ahead schedule of the blocked
Cholesky (hs_lookahead_schedule in heterogeneous-solvers_amd/csrc/hs_factor.h):
the launch sequence of both streams with their event waits, as the library's
schedule code issues it (recorded on the host by hs_lookahead_dry_run),
replayed with vector clocks. Every pair of launches on different streams that is not ordered by
the events must touch disjoint blocks (no write shared with a read or a
write), and executing the launches in an order the events allow must give
numpy's Cholesky. Covers the panel counts the FP64 (4) and mixed (6) factors
use and the clamping to n / 2, and the mixed factor's panel planes: after a
step's panel chain S2 copies the step's panels (rows below them) into one of
two plane sets, which that step's P-panel column update and rest update read
in place of the panels."""
import itertools

import numpy as np
import pytest

def pairs_for(mode, k, n, np_, extra=None):
    if mode == 4:
        rs = k + 2 * np_
        return [(I, J) for J in range(rs, n) for I in range(J, n)]
    if mode == 5:
        rs = k + 1
        return [(I, rs) for I in range(rs, n)]
    if mode == 6:
        rs, nc = extra
        out = []
        for c in range(min(nc, n - rs)):
            out += [(I, rs + c) for I in range(rs + c, n)]
        return out
    raise ValueError

def schedule(n, P, planes=False):
    """The launch sequence of both streams with their event records and
    waits, as the library's own schedule code issues it
    (hs_lookahead_schedule in hs_factor.h, recorded on the host by
    hs_lookahead_dry_run: no device is touched). Returned as
    (stream, kind, args) launches and ('rec' | 'wait', stream, event)."""
    import ctypes as C

    from hs_testlib import load_hs

    lib = load_hs().load()
    cnt = C.c_int64()
    assert lib.hs_lookahead_dry_run(n, P, int(planes), None, 0, C.byref(cnt)) == 0
    buf = (C.c_int64 * (8 * cnt.value))()
    assert lib.hs_lookahead_dry_run(n, P, int(planes), buf, cnt.value, C.byref(cnt)) == 0
    raw = list(buf)

    def event(e):
        return ('P', e) if e <= n else ('U', e - (n + 1))

    ops = []
    for i in range(cnt.value):
        kind, s_, a0, a1, a2, a3, a4, a5 = raw[8 * i:8 * i + 8]
        if kind == 0:
            ops.append(('rec', s_, event(a0)))
        elif kind == 1:
            ops.append(('wait', s_, event(a0)))
        elif kind == 2:
            ops.append((s_, 'potrf', (a0,)))
        elif kind == 3:
            ops.append((s_, 'trsm', (a0, a1, a2)))
        elif kind == 4:
            k, mode, rs, rc, np_, reads = a0, a1, a2, a3, a4, a5
            # the planes flag the code computed must be the analysis's rule
            assert bool(reads) == shared(mode, np_, P, planes), (k, mode, np_)
            if mode == 6:
                ops.append((s_, 'trail', (k, 6, np_, (rs, rc))))
            else:
                assert rs == (k + 1 if mode == 5 else k + 2 * np_)
                ops.append((s_, 'trail', (k, mode, np_)))
        elif kind == 5:
            ops.append((s_, 'prep', (a0,)))
        else:
            raise AssertionError("unknown op kind %d" % kind)
    return ops

def plane_set(k, P):
    return (k // P) & 1


def shared(mode, np_, P, planes):
    return planes and np_ == P and (mode == 4 or (mode == 6 and P > 1))


def rw_sets(op, n, P, planes):
    kind, a = op[1], op[2]
    if kind == 'potrf': return {(a[0], a[0])}, {(a[0], a[0])}
    if kind == 'prep':
        k0 = a[0]
        R = {(I, k0 + p) for I in range(k0 + P, n) for p in range(P)}
        return R, {('W', plane_set(k0, P), I, p) for I in range(k0 + P, n) for p in range(P)}
    if kind == 'trsm':
        k, st, c = a
        return {(k, k)} | {(I, k) for I in range(st, st + c)}, {(I, k) for I in range(st, st + c)}
    k, mode, np_ = a[0], a[1], a[2]
    ps = pairs_for(mode, k, n, np_, a[3] if len(a) > 3 else None)
    R = set(ps); W = set(ps)
    sh = shared(mode, np_, P, planes)
    for (I, J) in ps:
        for p in range(np_):
            if sh:
                R |= {('W', plane_set(k, P), I, p), ('W', plane_set(k, P), J, p)}
            else:
                R |= {(I, k + p), (J, k + p)}
    return R, W

def check(n, P, b=3, planes=False, drop=None):
    ops = schedule(n, P, planes)
    if drop is not None:
        ops = [op for op in ops if op != drop]
    # vector clocks: per stream op index; event -> clock snapshot
    clocks = {1: {1: 0, 2: 0}, 2: {1: 0, 2: 0}}
    ev = {}
    real = []  # (stream, idx, clock_at_start, op)
    cnt = {1: 0, 2: 0}
    for op in ops:
        if op[0] == 'rec':
            ev[op[2]] = dict(clocks[op[1]])
        elif op[0] == 'wait':
            e = ev[op[2]]
            for s_ in (1, 2): clocks[op[1]][s_] = max(clocks[op[1]][s_], e[s_])
        else:
            s_ = op[0]; cnt[s_] += 1
            real.append((s_, cnt[s_], dict(clocks[s_]), op))
            clocks[s_][s_] = cnt[s_]
    # a happens-before b iff b's start clock for a's stream >= a's index
    def hb(a, b): return b[2][a[0]] >= a[1]
    races = 0
    for a, c in itertools.combinations(real, 2):
        if a[0] == c[0]: continue
        if hb(a, c) or hb(c, a): continue
        Ra, Wa = rw_sets(a[3], n, P, planes); Rc, Wc = rw_sets(c[3], n, P, planes)
        if (Wa & (Rc | Wc)) or (Wc & Ra):
            races += 1
    # numeric: execute in program order of a topological merge (S2 first when possible)
    rng = np.random.default_rng(n * 10 + P)
    N = n * b
    M = rng.standard_normal((N, N)); K = M @ M.T + N * np.eye(N)
    A = np.tril(K.copy())
    Wp = {}  # plane copies of panel blocks: (set, block row, panel) -> b x b
    blk = lambda I, J: (slice(I * b, I * b + b), slice(J * b, J * b + b))
    q = {1: [r for r in real if r[0] == 1], 2: [r for r in real if r[0] == 2]}
    pos = {1: 0, 2: 0}; order = []
    def ready(c):
        other = 1 if c[0] == 2 else 2
        return pos[other] >= c[2][other]   # all ops of the other stream it waits on are done
    while pos[1] < len(q[1]) or pos[2] < len(q[2]):
        progressed = False
        for s_ in (2, 1):
            if pos[s_] < len(q[s_]) and ready(q[s_][pos[s_]]):
                order.append(q[s_][pos[s_]]); pos[s_] += 1; progressed = True; break
        assert progressed, "deadlock"
    for s_, i, _, op in order:
        kind, a = op[1], op[2]
        if kind == 'potrf':
            k = a[0]; A[blk(k, k)] = np.linalg.cholesky(np.tril(A[blk(k, k)]) + np.tril(A[blk(k, k)], -1).T)
        elif kind == 'trsm':
            k, st, c_ = a; L = A[blk(k, k)]
            for I in range(st, st + c_): A[blk(I, k)] = np.linalg.solve(L, A[blk(I, k)].T).T
        elif kind == 'prep':
            k0 = a[0]
            for I in range(k0 + P, n):
                for p in range(P):
                    Wp[(plane_set(k0, P), I, p)] = A[blk(I, k0 + p)].copy()
        else:
            k, mode, np_ = a[0], a[1], a[2]
            sh = shared(mode, np_, P, planes)
            for (I, J) in pairs_for(mode, k, n, np_, a[3] if len(a) > 3 else None):
                for p in range(np_):
                    if sh:
                        A[blk(I, J)] -= Wp[(plane_set(k, P), I, p)] @ Wp[(plane_set(k, P), J, p)].T
                    else:
                        A[blk(I, J)] -= A[blk(I, k + p)] @ A[blk(J, k + p)].T
                if I == J: A[blk(I, J)] = np.tril(A[blk(I, J)])
    err = np.abs(np.tril(A) - np.linalg.cholesky(K)).max()
    return races, err


@pytest.mark.parametrize("planes", [False, True])
@pytest.mark.parametrize("P", [1, 2, 3, 4, 5, 6, 7, 8])
def test_lookahead_schedule_is_race_free_and_exact(P, planes):
    for n in range(1, 20):
        p = max(1, min(P, n // 2))
        races, err = check(n, p, planes=planes)
        assert races == 0, (n, p)
        assert err < 1e-8, (n, p, err)


@pytest.mark.parametrize("planes", [False, True])
def test_analysis_catches_a_missing_wait(planes):
    """A planted mistake, on the CPU: without S2's wait for S1's update of the
    step before (event U), the next column update races with it."""
    n, P = 12, 2
    assert check(n, P, planes=planes)[0] == 0
    races, _ = check(n, P, planes=planes, drop=('wait', 2, ('U', 0)))
    assert races > 0


def test_dry_run_matches_the_factors_panel_counts():
    """The rest-update launches of the recorded schedule at the library's
    default depth (panels = 0) are the ones the bench's roofline counts
    (hs_cholesky_schedule_f64): 62 at N = 65536, b = 256 (four panels per
    step)."""
    import ctypes as C

    from hs_testlib import load_hs

    hs = load_hs()
    for n, b in ((256, 256), (32, 128), (7, 64)):
        ops = schedule(n, 0)
        rest = [op for op in ops if op[1:2] == ('trail',) and op[2][1] == 4]
        cnt, fl = C.c_int64(), C.c_double()
        hs.call("hs_cholesky_schedule_f64", n * b, b, C.byref(cnt), C.byref(fl))
        assert len(rest) == cnt.value
        if n == 256:
            assert cnt.value == 62
