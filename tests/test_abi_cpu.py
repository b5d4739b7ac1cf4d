"""CPU-side checks of the C-ABI boundary (no GPU needed, no compute calls).

  * librlgpu.so loads and exports every function include/rlgpu.h declares;
  * the Python binding knows a signature for every declared function;
  * status codes follow the reference's ErrorKind order (src/error.rs:21-34);
  * struct layouts used across the boundary match the header's field order.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

from rulinalg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rlgpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rlg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["rlg_gemm_f64", "rlg_lu_decompose_f64", "rlg_lu_solve_f64", "rlg_chol_decompose_f64",
                 "rlg_chol_solve_f64", "rlg_solve_mixed_f64_dev", "rlg_batch_solve_f64_dev",
                 "rlg_server_submit"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in declared_functions():
        assert n in exported, n  # unmangled: extern "C"


def test_binding_covers_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_status_codes_follow_errorkind_order():
    text = open(HEADER).read()
    codes = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"(RLG_E_[A-Z_]+)\s*=\s*(\d+)", text))
    # ErrorKind: InvalidArg, DecompFailure, AlgebraFailure, DivByZero,
    # ScalarConversionFailure, InvalidPermutation (src/error.rs:21-34)
    assert [codes[k] for k in ["RLG_E_INVALID_ARG", "RLG_E_DECOMP_FAILURE", "RLG_E_ALGEBRA_FAILURE",
                               "RLG_E_DIV_BY_ZERO", "RLG_E_SCALAR_CONVERSION",
                               "RLG_E_INVALID_PERMUTATION"]] == [1, 2, 3, 4, 5, 6]
    import rulinalg_amd as rl
    assert [k.value for k in rl.ErrorKind] == [1, 2, 3, 4, 5, 6]
    assert [k.name for k in rl.ErrorKind] == ["InvalidArg", "DecompFailure", "AlgebraFailure",
                                              "DivByZero", "ScalarConversionFailure",
                                              "InvalidPermutation"]


@pytest.mark.parametrize("struct,cname", [(_lib.MixedOpts, "rlg_mixed_opts"),
                                          (_lib.MixedReport, "rlg_mixed_report"),
                                          (_lib.BatchItem, "rlg_batch_item"),
                                          (_lib.ServerOpts, "rlg_server_opts"),
                                          (_lib.Request, "rlg_request"),
                                          (_lib.Result, "rlg_result")])
def test_struct_layouts(struct, cname, tmp_path):
    """Compile a probe against the header and compare sizes and offsets."""
    fields = [f for f, _ in struct._fields_]
    src = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){",
           f'printf("%zu\\n", sizeof({cname}));']
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "probe.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", str(c), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals[0] == C.sizeof(struct)
    for f, off in zip(fields, vals[1:]):
        assert getattr(struct, f).offset == off, f


def test_no_oracle_in_product():
    """The product library must not link or load the oracle."""
    out = subprocess.run(["nm", "-D", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "orc_" not in out
    ldd = subprocess.run(["ldd", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "liboracle" not in ldd
    for root, _, files in os.walk(os.path.join(ROOT, "rulinalg_amd")):
        for fn in files:
            if fn.endswith((".py", ".hip", ".hpp", ".cpp")):
                text = open(os.path.join(root, fn)).read()
                assert "import oracle" not in text and "from oracle" not in text, fn
