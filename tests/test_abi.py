"""The C ABI boundary: the library loads and exports every declared symbol;
the header compiles as C and C++; host-side error mapping. No GPU calls."""
import ctypes
import os
import re
import subprocess

import pytest

import nosh_amd as nosh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nosh_amd.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(nosh_\w+)\(", src, re.M)))


def test_header_declares_the_api():
    names = declared()
    for must in ("nosh_mesh_create", "nosh_keo_apply", "nosh_jac_rebuild", "nosh_jac_apply",
                 "nosh_prec_rebuild", "nosh_prec_apply", "nosh_nls_eval_f",
                 "nosh_nls_eval_dfdp", "nosh_newton_solve", "nosh_minres_solve"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(nosh.library_path), "build first: python -m nosh_amd.build"
    lib = ctypes.CDLL(nosh.library_path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("lang,compiler", [("c", "gcc"), ("c++", "g++")])
def test_header_compiles(lang, compiler, tmp_path):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "nosh_amd.h"\nint main(void){nosh_newton_options o;'
                   '(void)o; return NOSH_OK;}\n')
    subprocess.run([compiler, "-std=c11" if lang == "c" else "-std=c++17", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"), "-fsyntax-only", str(src)], check=True)


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of nosh_newton_options / nosh_newton_stats have the C sizes."""
    src = tmp_path / "s.c"
    src.write_text('#include <stdio.h>\n#include "nosh_amd.h"\nint main(void){printf("%zu %zu",'
                   'sizeof(nosh_newton_options), sizeof(nosh_newton_stats));return 0;}\n')
    exe = tmp_path / "s"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    a, b = map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                   check=True).stdout.split())
    assert a == ctypes.sizeof(nosh.NewtonOptions)
    assert b == ctypes.sizeof(nosh.NewtonStats)


def test_every_struct_mirror_matches_header(tmp_path):
    """Each ctypes mirror has the size and field offsets of its C struct."""
    pairs = [("nosh_newton_options", nosh.NewtonOptions), ("nosh_newton_stats", nosh.NewtonStats),
             ("nosh_batch_instance_stats", nosh.BatchInstanceStats),
             ("nosh_continuation_options", nosh.ContinuationOptions),
             ("nosh_continuation_step", nosh.ContinuationStep),
             ("nosh_surrogate_options", nosh.SurrogateOptions),
             ("nosh_surrogate_stats", nosh.SurrogateStats),
             ("nosh_batch_continuation_step", nosh.BatchContinuationStep),
             ("nosh_eigen_options", nosh.EigenOptions), ("nosh_eigen_stats", nosh.EigenStats),
             ("nosh_arclength_options", nosh.ArclengthOptions),
             ("nosh_arclength_step", nosh.ArclengthStep)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "nosh_amd.h"', 'int main(void){']
    for cname, cls in pairs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    vals = iter(map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                        check=True).stdout.split()))
    for cname, cls in pairs:
        assert next(vals) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert next(vals) == getattr(cls, f).offset, (cname, f)


def test_params_marshalling_sorted():
    p = nosh._Params({"mu": 0.5, "g": 2.0})
    assert p.count == 2
    assert [p.names[i] for i in range(2)] == [b"g", b"mu"]
    assert list(p.values) == [2.0, 0.5]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(nosh, "_lib", None)
    monkeypatch.setattr(nosh, "library_path", str(tmp_path / "absent.so"))
    with pytest.raises(nosh.LibraryNotBuilt):
        nosh.Mesh([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 2]])


def test_error_codes_map_to_reference_exceptions(monkeypatch):
    nosh._load()
    with pytest.raises(nosh.LogicError):
        nosh._check(2)
    with pytest.raises(KeyError):
        nosh._check(3)
    with pytest.raises(ValueError):
        nosh._check(1)
    with pytest.raises(nosh.NoshError):
        nosh._check(4)


def test_no_gpu_is_an_error_not_a_crash():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(nosh.NoshError):
        nosh.Mesh([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 2]])
