"""C-ABI checks without a GPU: the library loads and exports every symbol the
public header declares; errors come back as status codes with a message."""
import os
import re

import pytest

import helmholtz_amd as ha
from helmholtz_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "helmholtz_amd.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hh_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    assert set(_declared()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ha.lib()
    for name in _declared():
        assert hasattr(L, name), name


def test_version_and_error_string():
    L = ha.lib()
    assert L.hh_version() == 1
    assert isinstance(L.hh_last_error(), bytes)


def test_argument_errors_are_reported_without_touching_the_gpu():
    L = ha.lib()
    import ctypes as C
    out = C.c_void_p()
    nn = (C.c_int64 * 3)(10, 10, 1)
    h = (C.c_double * 3)(1.0, 1.0, 1.0)
    # dim = 4 is rejected before any device call
    rc = L.hh_op_create(4, nn, h, None, None, 1.0, 0.0, 0, 0, 2, C.byref(out))
    assert rc == _lib.HH_ERR_ARG
    rc = L.hh_op_create(2, nn, h, C.c_void_p(1), C.c_void_p(1), 1.0, 0.0, 0, 0, 3, C.byref(out))
    assert rc == _lib.HH_ERR_ARG
    assert b"bc_order" in L.hh_last_error()
    # getSommerfeldBC: a bad dimension, bc order or location is refused before any device call
    m1 = C.c_void_p(1)
    assert L.hh_sommerfeld_bc(4, nn, h, m1, 1.0, 0, 2, m1, 0) == _lib.HH_ERR_ARG
    assert L.hh_sommerfeld_bc(2, nn, h, m1, 1.0, 0, 3, m1, 0) == _lib.HH_ERR_ARG
    assert b"bc_order" in L.hh_last_error()
    assert L.hh_sommerfeld_bc(2, nn, h, m1, 1.0, 0, 2, m1, 7) == _lib.HH_ERR_ARG
    # getABL with a face mask: entries must be 0 or 1, and the mask must be given
    pad = (C.c_int64 * 3)(2, 2, 2)
    nn3 = (C.c_int64 * 3)(9, 8, 7)
    bad = (C.c_int32 * 6)(1, 1, 2, 1, 1, 1)
    assert L.hh_abl_faces(3, nn3, 0, pad, 1.0, bad, m1, 0) == _lib.HH_ERR_ARG
    assert b"faces" in L.hh_last_error()
    assert L.hh_abl_faces(3, nn3, 0, pad, 1.0, None, m1, 0) == _lib.HH_ERR_ARG
    with pytest.raises(ValueError):
        ha.getABL([9, 8, 7], False, [2, 2, 2], 1.0, [[1, 1], [1, 1]])


def test_solver_rejects_unsupported_settings():
    mg = ha.getMGparam(complex, int, 3, 1, 30, 1e-6, "GS", 0.75, 2, 2, "K", "NoMUMPS", 0.5, 0.0)
    M = ha.getRegularMesh([0, 1, 0, 1], [8, 8])
    import numpy as np
    hp = ha.HelmholtzParam(M, np.zeros(81), np.ones(81), 1.0, False, True)
    s = ha.getShiftedLaplacianMultigridSolver(hp, mg, 0.5, "GMRES", 5)
    with pytest.raises(NotImplementedError):
        ha.solveLinearSystem(None, np.ones(81, complex), s)
    # Jac-GMRES takes at most 8 GMRES steps per smoothing call
    mg9 = ha.getMGparam(complex, int, 3, 1, 30, 1e-6, "Jac-GMRES", 0.75, 9, 2, "K", "NoMUMPS", 0.5, 0.0)
    with pytest.raises(ValueError):
        ha.solveLinearSystem(None, np.ones(81, complex), ha.getShiftedLaplacianMultigridSolver(hp, mg9, 0.5, "GMRES", 5))
    # the FGMRES basis is stored in complex128 or complex64 only
    mgb = ha.getMGparam(complex, int, 3, 1, 30, 1e-6, "Jac", 0.75, 2, 2, "V", "NoMUMPS", 0.5, 0.0)
    mgb.basis = "bf16"
    with pytest.raises(ValueError):
        ha.solveLinearSystem(None, np.ones(81, complex), ha.getShiftedLaplacianMultigridSolver(hp, mgb, 0.5, "GMRES", 5))
    assert ha.getMGparam(3, 1, 30, 1e-6, "Jac", 0.8, 2, 2, "V", "NoMUMPS").basis == "c128"
    s2 = ha.getShiftedLaplacianMultigridSolver(hp, ha.getMGparam(3, 1, 30, 1e-6, "Jac", 0.8, 2, 2, "V", "NoMUMPS"),
                                               0.5)
    assert s2.Krylov == "BiCGSTAB" and s2.inner == 5 and list(s2.shift) == [0.5] * 3


def test_getmgparam_both_signatures():
    a = ha.getMGparam(complex, int, 2, 2, 30, 1e-6, "Jac", 0.75, 2, 2, "W", "NoMUMPS", 0.5, 0.0)
    b = ha.getMGparam(5, 4, 50, 1e-5, "Jac", 0.8, 2, 2, "K", "NoMUMPS", 0.5, 0.0, "FullWeighting")
    assert (a.levels, a.numCores, a.maxOuterIter, a.cycleType) == (2, 2, 30, "W")
    assert (b.levels, b.relaxParam, b.extra) == (5, 0.8, (0.5, 0.0, "FullWeighting"))


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of the header's structs (hh_mg_params,
    hh_solve_stats, hh_problem, hh_problem_result) have the C compiler's
    size and field offsets: compiled from include/helmholtz_amd.h with gcc."""
    import ctypes as C
    import shutil
    import subprocess
    from helmholtz_amd import _lib as L
    from helmholtz_amd import stream
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"hh_mg_params": L.MGParams, "hh_solve_stats": L.SolveStats, "hh_problem": stream.Problem,
               "hh_problem_result": stream.ProblemResult}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "helmholtz_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            s, f, v = line.split()
            got[(s, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def _c_fields(header: str, name: str):
    """(field, julia type) of a C struct in the header, in order."""
    body = re.search(r"typedef struct " + name + r" \{(.*?)\} " + name + ";", header, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    cmap = {"int32_t": "Int32", "int64_t": "Int64", "double": "Float64"}
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(int32_t|int64_t|double)\s*(\*?)\s*(\w+)(\[(\d+)\])?$", decl)
        assert m, decl
        t = cmap[m.group(1)]
        if m.group(2):
            t = f"Ptr{{{t}}}"
        if m.group(5):
            t = f"NTuple{{{m.group(5)},{t}}}"
        out.append((m.group(3), t))
    return out


def _julia_fields(src: str, name: str):
    body = re.search(r"struct " + name + r"\n(.*?)\nend", src, re.S).group(1)
    return [(f.strip(), t.strip()) for f, t in re.findall(r"(\w+)::([\w{},]+)", body)]


def test_julia_shim_structs_match_the_header():
    """The Julia ccall shim's HHMGParams / HHSolveStats have the header's
    fields, in order, with the matching Julia types."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    header = open(os.path.join(root, "include", "helmholtz_amd.h")).read()
    jl = open(os.path.join(root, "helmholtz.jl_amd", "julia", "HelmholtzAMD.jl")).read()
    assert _julia_fields(jl, "HHMGParams") == _c_fields(header, "hh_mg_params")
    assert _julia_fields(jl, "HHSolveStats") == _c_fields(header, "hh_solve_stats")


def test_no_cpu_fallback_without_the_library(tmp_path):
    """Without libhelmholtz_amd.so every compute entry of the package raises:
    there is no host fallback (run in a fresh interpreter so that the
    library this process loaded stays untouched)."""
    import subprocess
    import sys
    import textwrap

    script = textwrap.dedent(f"""
        import sys
        sys.path.insert(0, {os.path.join(ROOT, "helmholtz.jl_amd")!r})
        import numpy as np
        from helmholtz_amd import _lib
        _lib.LIB_PATH = {str(tmp_path / "absent.so")!r}
        import helmholtz_amd as ha
        M = ha.getRegularMesh([0.0, 1.0, 0.0, 1.0], [8, 8])
        m = np.ones(81)
        calls = [
            lambda: ha.getABL([9, 9], False, [2, 2], 1.0),
            lambda: ha.getSommerfeldBC(M, m, 1.0, False),
            lambda: ha.GetHelmholtzOperator(M, m, 1.0, np.zeros(81), False, True) @ np.ones(81, complex),
        ]
        for c in calls:
            try:
                c()
            except ha.HelmholtzError as e:
                assert "not built" in str(e), e
            else:
                raise SystemExit("a compute call ran without the library")
        print("all raised")
    """)
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all raised" in r.stdout


def test_sommerfeld_bc_host_checks():
    """getSommerfeldBC takes a real omega (the reference's omega::Float64) and
    one slowness value per node; both are checked before the library runs."""
    import numpy as np

    M = ha.getRegularMesh([0.0, 1.0, 0.0, 1.0], [8, 8])
    with pytest.raises(TypeError):
        ha.getSommerfeldBC(M, np.ones(81), complex(1.0, 0.5), False)
    with pytest.raises(ValueError):
        ha.getSommerfeldBC(M, np.ones(80), 1.0, False)


def test_helmholtz_fun_refuses_sparse_matrices():
    """getHelmholtzFun composes matrix-free operators only: the reference's
    CSR arguments are refused (there is no host SpMV path)."""
    import numpy as np
    import scipy.sparse as sp

    with pytest.raises(TypeError):
        ha.getHelmholtzFun(sp.eye(4, format="csc"), sp.eye(4, format="csc"), np.zeros(4, complex), 1)


def test_loc2cs_methods():
    """Both loc2cs methods and loc2cs3D (src/getPointSource.jl:82-100) give the
    1-based column-major index, which numpy's ravel_multi_index confirms."""
    import numpy as np

    n = [7, 5, 4]
    for sub in ([3, 2], [7, 5], [1, 1]):
        want = int(np.ravel_multi_index([v - 1 for v in sub], n[:2], order="F")) + 1
        assert ha.loc2cs(n[:2], sub) == want == ha.loc2cs(sub[0], sub[1], n[:2])
    for sub in ([3, 2, 4], [7, 5, 1], [1, 1, 1]):
        want = int(np.ravel_multi_index([v - 1 for v in sub], n, order="F")) + 1
        assert ha.loc2cs(n, sub) == want == ha.loc2cs3D(sub[0], sub[1], sub[2], n)



def _split_top(s: str):
    """Split at commas outside (), {} and [] (for Julia type tuples and C
    parameter lists)."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _balanced(s: str, start: int) -> str:
    """The text inside the parenthesis that opens at s[start]."""
    depth = 0
    for i in range(start, len(s)):
        depth += s[i] == "("
        depth -= s[i] == ")"
        if depth == 0:
            return s[start + 1:i]
    raise ValueError("unbalanced")


def test_julia_ccalls_match_the_header():
    """Every ccall of the Julia shim names a function the header declares and
    passes as many argument types as the C declaration has parameters."""
    header = open(HEADER).read()
    decl = {}
    for mt in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hh_\w+)\s*\(", header, re.M):
        params = _split_top(_balanced(header, mt.end() - 1))
        decl[mt.group(1)] = 0 if params in ([], ["void"]) else len(params)
    jl = open(os.path.join(ROOT, "helmholtz.jl_amd", "julia", "HelmholtzAMD.jl")).read()
    calls = list(re.finditer(r"ccall\(\(:(hh_\w+), LIB\)", jl))
    assert calls
    for mc in calls:
        name = mc.group(1)
        assert name in decl, name
        args = _split_top(_balanced(jl, mc.start() + len("ccall")))
        types = args[2].strip()
        assert types.startswith("(") and types.endswith(")"), (name, types)
        n = len(_split_top(types[1:-1]))
        assert n == decl[name], (name, n, decl[name])


def test_ctypes_argtypes_match_the_header():
    """The ctypes binding declares, for every header function it sets
    argtypes for, as many argument types as the C declaration has
    parameters."""
    header = open(HEADER).read()
    L = ha.lib()
    checked = 0
    for mt in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hh_\w+)\s*\(", header, re.M):
        params = _split_top(_balanced(header, mt.end() - 1))
        n = 0 if params in ([], ["void"]) else len(params)
        at = getattr(L, mt.group(1)).argtypes
        if at is not None:
            assert len(at) == n, (mt.group(1), len(at), n)
            checked += 1
    assert checked >= 30


def test_julia_shim_is_a_jinv_plugin():
    """The shim plugs into jInv as the reference's solver does
    (src/Helmholtz.jl:5-8, src/ShiftedLaplacianMultigridSolver.jl:4,17,32,104):
    the solver type subtypes jInv.LinearSolvers.AbstractSolver, and
    solveLinearSystem / copySolver / clear! are imported from jInv before the
    shim defines its methods, so they extend jInv's generic functions (a
    module-local function of the same name would not be reached by a jInv
    caller's dispatch)."""
    jl = open(os.path.join(ROOT, "helmholtz.jl_amd", "julia", "HelmholtzAMD.jl")).read()
    code = "\n".join(l.split("#", 1)[0] for l in jl.splitlines())
    assert re.search(r"^mutable struct ShiftedLaplacianMultigridSolverAMD <: AbstractSolver$", code, re.M)
    imports = {
        "AbstractSolver": r"^import jInv\.LinearSolvers\.AbstractSolver$",
        "solveLinearSystem": r"^import jInv\.LinearSolvers\.solveLinearSystem$",
        "copySolver": r"^import jInv\.LinearSolvers\.copySolver$",
        "clear!": r"^import jInv\.Utils\.clear!$",
    }
    pos = {}
    for name, pat in imports.items():
        m = re.search(pat, code, re.M)
        assert m, f"missing import of jInv's {name}"
        pos[name] = m.start()
    defs = {
        "solveLinearSystem": r"^function solveLinearSystem\(ShiftedHT, B, param::ShiftedLaplacianMultigridSolverAMD,"
                             r" doTranspose::Int64=0\)",
        "clear!": r"^function clear!\(s::ShiftedLaplacianMultigridSolverAMD\)",
        "copySolver": r"^copySolver\(s::ShiftedLaplacianMultigridSolverAMD\) =",
    }
    for name, pat in defs.items():
        m = re.search(pat, code, re.M)
        assert m, f"no method of {name} for the AMD solver"
        assert pos[name] < m.start(), f"{name} is defined before it is imported from jInv"
    # ShiftedHT is read when the hierarchy is built (checked, not ignored)
    assert re.search(r"ensure_handle!\(param, ShiftedHT\)", code)
    assert re.search(r"s\.op = matching_operator\(s, ShiftedHT\)", code)
    assert "does not match helmParam's shifted operator" in code
