"""The code examples of INTEGRATION.md compile against this build's headers
(CPU only, g++ -fsyntax-only): the ported main.cpp solver calls, the --gpr
branch and the C ABI example. Each snippet is placed in a function whose
preamble declares the names the prose leaves to the caller."""
import os
import re
import subprocess

import pytest

from hs_testlib import ROOT

DOC = os.path.join(ROOT, "INTEGRATION.md")


def blocks(lang):
    text = open(DOC).read()
    return re.findall(r"```%s\n(.*?)```" % lang, text, flags=re.S)


def compile_ok(src, tmp_path, suffix, compiler):
    f = tmp_path / ("snippet" + suffix)
    f.write_text(src)
    r = subprocess.run([compiler, "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-std=" +
                        ("c++17" if suffix == ".cpp" else "c11"), str(f)], capture_output=True, text=True)
    return r.returncode, r.stderr


CPP_PREAMBLE = """#include <memory>
#include <string>
void ported_main() {
    std::string path_gp_input = "in.txt", path_gp_output = "out.txt", path_gp_test = "test.txt";
"""


def test_cpp_snippets_compile(tmp_path):
    cpp = blocks("cpp")
    assert len(cpp) >= 3, "expected the ported main, the --gpr branch and the builder overload"
    main_snippet, gpr_snippet = cpp[0], cpp[1]
    assert "#include <hs/hs.hpp>" in main_snippet
    body = main_snippet.replace("#include <hs/hs.hpp>", "")
    src = "#include <hs/hs.hpp>\n" + CPP_PREAMBLE + body + gpr_snippet + "\n}\n"
    rc, err = compile_ok(src, tmp_path, ".cpp", "g++")
    assert rc == 0, err


C_PREAMBLE = """#include <stddef.h>
#include <stdint.h>
void example(int64_t N, int d, int b, int64_t M, double ell, double sf2, double sn2, double* X_host,
             double* K2, double* x, double* y2, float* L32, double* Xtest, double* alpha, double* Kbar) {
"""


def test_c_abi_example_compiles(tmp_path):
    c = [blk for blk in blocks("c") if "hs_abi.h" in blk]
    assert c, "no C ABI example found"
    body = c[0].replace('#include "hs_abi.h"', "")
    body = body.replace("/* ... hs_memcpy_htod(h, X, X_host, bytes) ... */", "(void)X_host;")
    src = '#include "hs_abi.h"\n' + C_PREAMBLE + body + "\n}\n"
    rc, err = compile_ok(src, tmp_path, ".c", "gcc")
    assert rc == 0, err


@pytest.mark.parametrize("name", ["hs_workspace_reserve", "hs_tri_prepare_f64", "hs_cg_ref_order_f64",
                                  "hs_refine_solve_to", "hs_cg_f64_noalloc"])
def test_new_entries_documented(name):
    assert name in open(DOC).read()
