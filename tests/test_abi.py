"""The C-ABI library loads and exports every entry point include/*.h declares (no GPU needed)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mpdg_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mpdg_heat_op_vmult" in names and "mpdg_linear_solve" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    import meltpooldg_amd as M
    if not os.path.exists(M.LIB_PATH):
        pytest.fail("libmpdg_heat.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(M.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds exactly the declared entry points
    assert set(M.exported_symbols()) <= set(declared_functions())


@pytest.mark.parametrize("lang,compiler", [("c", "gcc"), ("c++", "g++")])
def test_header_compiles_as_c_and_cxx(lang, compiler):
    """include/mpdg_heat.h is a self-contained C header (and valid C++ with its extern "C" guard)."""
    import subprocess
    h = os.path.join(ROOT, "include", "mpdg_heat.h")
    std = "-std=c99" if lang == "c" else "-std=c++17"
    r = subprocess.run([compiler, "-fsyntax-only", "-Wall", "-Wextra", "-Werror", std, "-x", lang, h],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_python_mirror_fails_loudly_without_gpu():
    import meltpooldg_amd as M
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError):
        M.HeatDiffuseMultiPhaseOperator(2, (2, 2, 2))


def test_python_struct_mirrors_match_the_header(tmp_path):
    """Every ctypes mirror of a parameter struct has the header's size and field offsets (a C
    program built with gcc from include/mpdg_heat.h prints them)."""
    import subprocess
    import meltpooldg_amd as M
    mirrors = {"mpdg_heat_desc": M.HeatDesc, "mpdg_gmg_params": M.GMGParams, "mpdg_solver_params": M.SolverParams,
               "mpdg_heat_material": M.Material, "mpdg_heat_operation_params": M.OperationParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mpdg_heat.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    for cname, py in mirrors.items():
        assert int(got["%s size" % cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got["%s %s" % (cname, f)]) == getattr(py, f).offset, (cname, f)
