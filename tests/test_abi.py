"""The C ABI boundary (no GPU needed): the HIP library and the oracle load
and export every entry point include/ibmdem.h declares; the ctypes structs
match the header's layout; the product refuses to run without its HIP
library."""
import ctypes as C
import os
import re

import pytest

import ibmdem
import oracle
from ibmdem import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ibmdem.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(ibmdem_\w+)\s*\(", text, re.M)))


def test_header_and_wrapper_agree():
    assert set(declared()) == set(abi.EXPORTED)


def test_gpu_library_exports_every_symbol():
    if not os.path.exists(ibmdem.GPU_LIB):
        pytest.skip("HIP library not built")
    lib = C.CDLL(ibmdem.GPU_LIB)
    for name in declared():
        assert hasattr(lib, name), name


def test_oracle_exports_every_symbol():
    lib = C.CDLL(oracle.build())
    for name in declared():
        assert hasattr(lib, name), name


def test_struct_layouts():
    # sizes from the header: 3+6 ints, 3+3 doubles, ... checked by offsets
    assert abi.Case.dx.offset == 40
    assert abi.Case.reserved.offset + C.sizeof(C.c_int32 * 6) <= C.sizeof(abi.Case)
    assert C.sizeof(abi.Markers) == 8 * C.sizeof(C.c_void_p)
    assert C.sizeof(abi.Bodies) == 18 * C.sizeof(C.c_void_p)
    assert C.sizeof(abi.SolveInfo) == 24


def test_case_layout_matches_compiled_header(tmp_path):
    """Compile a probe against include/ibmdem.h and compare every offset."""
    import subprocess
    fields = [f[0] for f in abi.Case._fields_]
    src = ["#include <stddef.h>", "#include <stdio.h>", f'#include "{HEADER}"', "int main(void){"]
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof(ibmdem_case, {f}));')
    src.append('printf("%zu\\n", sizeof(ibmdem_case)); return 0; }')
    c = tmp_path / "probe.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", str(c), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[:-1] == [getattr(abi.Case, f).offset for f in fields]
    assert out[-1] == C.sizeof(abi.Case)


def test_product_fails_loudly_without_hip_library(monkeypatch):
    monkeypatch.setattr(ibmdem, "GPU_LIB", "/nonexistent/libibmdem.so")
    monkeypatch.setattr(ibmdem, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        ibmdem.gpu_library()


def test_oracle_rejects_bad_cases():
    from ibmdem import cases
    c = cases.build_case(8, 8, 8, [0, 1, 0, 0, 0, 0])   # periodic on one side only
    with pytest.raises(abi.IbmdemError):
        oracle.simulation(c)
    c = cases.build_case(1, 8, 8, [0] * 6)
    with pytest.raises(abi.IbmdemError):
        oracle.simulation(c)
