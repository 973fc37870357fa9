#!/usr/bin/env python3
"""Extract the reference's golden test vectors into JSON fixtures.

Run in the build container only (it reads the reference checkout):

    python tests/golden/extract_reference_goldens.py /root/reference

It reads the reference's GoogleTest files (tests/test*.cpp of
TimThuering/Heterogeneous-Solvers) and writes, for every test case, the data
that case holds: block/work-group sizes, the sequence of kernel calls with
their integer arguments, how the input vectors are filled, and the expected
values with the comparison tolerance. The numeric arrays are copied as
numbers; no source text is kept. The two data files the tests read
(tests/testData/*.txt) are copied verbatim to tests/golden/data/.

The parser is deliberately strict: a statement it does not understand makes
it stop with an error, so a silent mis-read cannot produce a fixture.
"""
import json
import os
import re
import shutil
import sys

FILES = [
    "testSymmetricMatrix.cpp",
    "testMatrixOperations.cpp",
    "testMatrixMatrixOperations.cpp",
    "testMatrixVectorBlock.cpp",
    "testVectorOperations.cpp",
]

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", lambda m: "\n" * m.group(0).count("\n"), text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def match_brace(text, open_idx):
    depth = 0
    for i in range(open_idx, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return i
    raise ValueError("unbalanced braces")


def parse_arrays(body):
    """All `std::vector<conf::fp_type|double> NAME = { ... };` literals."""
    out = {}
    for m in re.finditer(r"std::vector<(?:conf::fp_type|double)>\s+(\w+)\s*=\s*\{", body):
        end = match_brace(body, m.end() - 1)
        vals = [float(v) for v in re.findall(NUM, body[m.end():end])]
        out[m.group(1)] = vals
    return out


def line_of(text, idx):
    return text.count("\n", 0, idx) + 1


def split_args(s):
    args, depth, cur = [], 0, ""
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            args.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        args.append(cur.strip())
    return [re.sub(r"\s+", " ", a) for a in args]


KNOWN_STMT = [
    r"queue queue\(cpu_selector_v\)", r"queue queue\(gpu_selector_v\)",
    r"sycl::queue queue\(sycl::cpu_selector_v\)", r"queue\.wait\(\)",
    r"const usm_allocator<[^>]*> allocator\{queue\}",
    r"std::vector<conf::fp_type, usm_allocator<[^>]*>> (\w+)\(allocator\)",
    r"std::vector<conf::fp_type> vector", r"conf::fp_type resultValue = 0",
]


def parse_test(body):
    rec = {"config": {}, "parse": [], "vectors": {}, "calls": [], "checks": [], "known_answer": None}
    arrays = parse_arrays(body)
    # remove array literals before statement scanning
    scan = body
    for m in reversed(list(re.finditer(r"std::vector<(?:conf::fp_type|double)>\s+(\w+)\s*=\s*\{", scan))):
        end = match_brace(scan, m.end() - 1)
        semi = scan.index(";", end)
        scan = scan[:m.start()] + scan[semi + 1:]
    # for loops (init or check)
    loops = []
    for m in re.finditer(r"for \((?:size_t|unsigned int) i = (\d+); i < ([^;]+); (?:i\+\+|\+\+i)\)\s*\{", scan):
        end = match_brace(scan, m.end() - 1)
        loops.append((m.start(), end, m.group(1), m.group(2).strip(), scan[m.end():end]))
    covered = []
    for (s, e, lo, hi, inner) in loops:
        covered.append((s, e))
        inner_s = inner.strip()
        mm = re.fullmatch(r"EXPECT_NEAR\((\w+(?:\.\w+)?)\[([^\]]+)\], (\w+)\[i\], (" + NUM + r")\);", inner_s)
        if mm:
            target, idx, ref, tol = mm.groups()
            offset = "0"
            im = re.fullmatch(r"(.+?)\s*\+\s*i", idx)
            if idx.strip() == "i":
                offset = "0"
            elif im:
                offset = im.group(1).strip()
            else:
                raise ValueError("check index " + idx)
            rec["checks"].append({"target": target, "offset": offset, "expected": ref, "count": hi, "tol": float(tol)})
            continue
        mm = re.fullmatch(r"EXPECT_DOUBLE_EQ\((\w+)\[i\], values\[i\]\);", inner_s)
        if mm:
            rec["checks"].append({"row_values_equal": mm.group(1)})
            continue
        mm = re.fullmatch(r"(\w+)\[i\] = (1|std::sqrt\(i\));", inner_s)
        if mm:
            rec["vectors"].setdefault(mm.group(1), {}).setdefault("init", []).append(
                [int(lo), hi, "one" if mm.group(2) == "1" else "sqrt_i"])
            continue
        mm = re.fullmatch(r"(?:if \(i >= (\d+)\) \{\s*)?resultValue \+= std::sqrt\(i\) \* std::sqrt\(i\);\s*\}?\s*vector\[i\] = std::sqrt\(i\);",
                          inner_s, flags=re.S)
        if mm:
            rec["vectors"].setdefault("vector", {}).setdefault("init", []).append([int(lo), hi, "sqrt_i"])
            rec["known_answer"] = {"kind": "sum_sqrt_i_squared", "from": int(mm.group(1) or 0), "to": int(hi)}
            continue
        raise ValueError("unknown loop body: " + inner_s[:120])
    rest = scan
    for (s, e) in sorted(covered, reverse=True):
        rest = rest[:s] + rest[e + 1:]
    # string literals may hold ';': lift them out before splitting statements
    strings = re.findall(r'"([^"]*)"', rest)
    for k, lit in enumerate(strings):
        rest = rest.replace('"%s"' % lit, '"@STR%d@"' % k, 1)

    def unstr(t):
        return re.sub(r"@STR(\d+)@", lambda q: strings[int(q.group(1))], t)
    for stmt in [t.strip() for t in rest.split(";")]:
        stmt = re.sub(r"\s+", " ", stmt)
        if not stmt:
            continue
        m = re.fullmatch(r"conf::(\w+) = (\d+)", stmt)
        if m:
            rec["config"][m.group(1)] = int(m.group(2))
            continue
        m = re.fullmatch(r"(?:SymmetricMatrix|RightHandSide) (\w+) = MatrixParser::(parseSymmetricMatrix|parseRightHandSide)\((\w+), queue\)", stmt)
        if m:
            rec["parse"].append([m.group(1), m.group(2), m.group(3)])
            continue
        m = re.fullmatch(r"(\w+)\.resize\(([^,()]+(?:\(\))?(?:\.size\(\))?)(?:, (" + NUM + r"))?\)", stmt)
        if m:
            v = rec["vectors"].setdefault(m.group(1), {})
            v["size"] = m.group(2).strip()
            v["fill"] = float(m.group(3)) if m.group(3) else 0.0
            continue
        m = re.fullmatch(r"(?:unsigned int workGroupCount = )?(\w+::\w+)\((.*)\)", stmt)
        if m and not stmt.startswith("EXPECT"):
            rec["calls"].append([m.group(1), split_args(m.group(2))])
            continue
        m = re.fullmatch(r"EXPECT_NEAR\((\w+)\[(\d+)\], (" + NUM + r"|resultValue), (" + NUM + r")\)", stmt)
        if m:
            exp = m.group(3)
            rec["checks"].append({"target": m.group(1), "offset": m.group(2), "expected_scalar": exp if exp == "resultValue" else float(exp),
                                  "count": "1", "tol": float(m.group(4))})
            continue
        m = re.fullmatch(r"EXPECT_EQ\(([\w.]+)\.size\(\), ([\w.]+)\.size\(\)\)", stmt)
        if m:
            rec["checks"].append({"size_equal": [m.group(1), m.group(2)]})
            continue
        m = re.fullmatch(r"EXPECT_EQ\((\w+)\.(blockSize|blockCountXY|matrixData\.size\(\)), ([\d *]+)\)", stmt)
        if m:
            rec["checks"].append({"attr": m.group(2), "value": eval(m.group(3), {"__builtins__": {}})})
            continue
        m = re.fullmatch(r"EXPECT_DOUBLE_EQ\((\w+)\.matrixData\[([\d *+-]+)\], (" + NUM + r")\)", stmt)
        if m:
            rec["checks"].append({"target": "A.matrixData", "offset": m.group(2), "expected_scalar": float(m.group(3)),
                                  "count": "1", "tol": 0.0, "double_eq": True})
            continue
        m = re.fullmatch(r"EXPECT_THROW\(MatrixParser::getRowValuesFromString\(\"(.*)\"\), std::invalid_argument\)", stmt)
        if m:
            rec["checks"].append({"throws": "invalid_argument", "row_string": unstr(m.group(1))})
            continue
        m = re.fullmatch(r"std::vector<conf::fp_type> values = MatrixParser::getRowValuesFromString\( ?\"(.*)\"\)", stmt)
        if m:
            rec["row_string"] = unstr(m.group(1))
            continue
        if any(re.fullmatch(p, stmt) for p in KNOWN_STMT):
            continue
        if stmt in ("}",):
            continue
        raise ValueError("unknown statement: " + stmt[:160])
    return rec, arrays


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    out = {"source": "TimThuering/Heterogeneous-Solvers tests/ (MIT licence)", "suites": {}, "tests": []}
    for fname in FILES:
        text = strip_comments(open(os.path.join(ref, "tests", fname)).read())
        for m in re.finditer(r"class (\w+) : public ::testing::Test\s*\{", text):
            end = match_brace(text, m.end() - 1)
            out["suites"].setdefault(m.group(1), {}).update(parse_arrays(text[m.end():end]))
        for m in re.finditer(r"TEST(?:_F)?\((\w+), (\w+)\)\s*\{", text):
            end = match_brace(text, m.end() - 1)
            rec, arrays = parse_test(text[m.end():end])
            rec.update({"suite": m.group(1), "name": m.group(2), "source": "tests/%s:%d" % (fname, line_of(text, m.start())),
                        "arrays": arrays})
            out["tests"].append(rec)
    os.makedirs(os.path.join(here, "data"), exist_ok=True)
    for f in ("testMatrixSymmetric20x20.txt", "testVector_20.txt"):
        shutil.copyfile(os.path.join(ref, "tests", "testData", f), os.path.join(here, "data", f))
    with open(os.path.join(here, "reference_goldens.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("tests:", len(out["tests"]), "suites:", {k: len(v) for k, v in out["suites"].items()})


if __name__ == "__main__":
    main()
