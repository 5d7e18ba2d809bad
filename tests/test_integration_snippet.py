"""The linear-solver seam as INTEGRATION.md section 1 shows it to a
maintainer: its solveWithBackend branch, taken verbatim from the document,
is compiled against the reference's own Matrix.h (src/common/include,
Matrix.h:57-89) and linked with the library and the reference's solveCG
(the reference build in oracle/_ref).  No compute call: this runs without a
GPU, and only where the reference tree is present."""
import os
import re
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"
REF_INC = os.path.join(REFERENCE, "src", "common", "include")
LIB_DIR = os.path.join(REPO, "chuanflow_amd", "_lib")
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libcfref.so")


def section_one_snippet():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    sec = text[text.index("## 1."):text.index("## 2.")]
    m = re.search(r"```cpp\n(.*?)```", sec, re.S)
    assert m, "INTEGRATION.md section 1 has no C++ block"
    return m.group(1)


def translation_unit(snippet):
    """The snippet as a maintainer places it: the enum in Matrix.h (here a
    namespace of its own, since the reference header keeps its enum), the
    include at file scope, the branch at the top of solveWithBackend."""
    includes = [l for l in snippet.splitlines() if l.startswith("#include")]
    enum = next(l for l in snippet.splitlines() if l.startswith("enum class LinearBackend"))
    start = snippet.index("if (backend == LinearBackend::Hip)")
    branch = snippet[start:]
    return "\n".join(["#include <cf/common/Error.h>", "#include <cf/common/Matrix.h>"] + includes + [
        "namespace seam {",
        "using namespace cf;",
        enum,
        "SolverReport solveWithBackend(const SparseMatrix& A, const Field<double>& b, Field<double>& x, double tol,",
        "                              int maxIter, LinearBackend backend)",
        "{",
        branch,
        "    return solveCG(A, b, x, tol, maxIter);",
        "}",
        "} // namespace seam",
        "int main() { return seam::solveWithBackend(cf::SparseMatrix{}, {}, *new cf::Field<double>(), 1e-8, 0,",
        "                                           seam::LinearBackend::Native).converged ? 0 : 1; }",
    ]) + "\n"


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_INC, "cf", "common", "Matrix.h")),
                    reason="reference tree not present")
@pytest.mark.skipif(not os.path.isfile(REF_LIB), reason="reference build (oracle/_ref) not built")
def test_section_one_branch_compiles_and_links():
    src = translation_unit(section_one_snippet())
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "seam.cpp")
        open(cpp, "w").write(src)
        cmd = ["g++", "-std=c++20", "-Wall", "-Werror", "-Wno-unused-variable", "-I" + REF_INC,
               "-I" + os.path.join(REPO, "include"), cpp, "-o", os.path.join(d, "seam"),
               "-L" + LIB_DIR, "-lchuanflow_amd", REF_LIB, "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
               "-Wl,-rpath," + LIB_DIR]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, src + "\n" + r.stderr
