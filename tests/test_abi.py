"""The C ABI: libmcpm.so loads and exports every entry point that
include/mcpm.h declares (CPU only: no compute call needs a GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

import mcpm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mcpm_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = mcpm.load_library()
    names = _declared("mcpm.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", mcpm.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (mcpm_\w+)", out))
    assert set(names) <= exported


def test_status_strings_and_host_contact_value():
    lib = mcpm.load_library()
    assert lib.mcpm_status_string(0) == b"ok"
    assert lib.mcpm_status_string(4) == b"particle outside the domain"
    p = mcpm.Params(n_cells=10, length=10.0, capacity=10)
    # Y at n sigma^3 = 0.2 (Main_Functions.cpp:1911): 1.320591
    assert abs(mcpm.contact_y(p, 0.2) - 1.3205908) < 1e-7
    p.y_model = mcpm.Y_CARNAHAN_STARLING
    assert abs(mcpm.contact_y(p, 0.2) - 1.3205854) < 1e-6


def test_params_struct_layout_matches_header(tmp_path):
    """Every field offset of the ctypes mirror equals the C compiler's
    offsetof on include/mcpm.h, and so does the struct size."""
    fields = [f[0] for f in mcpm.CParams._fields_]
    src = tmp_path / "probe.c"
    body = "".join(f'  printf("%s %zu\\n", "{f}", offsetof(mcpm_params, {f}));\n' for f in fields)
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"mcpm.h\"\n"
                   "int main(void) {\n" + body +
                   '  printf("sizeof %zu\\n", sizeof(mcpm_params));\n  return 0;\n}\n')
    exe = tmp_path / "probe"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = dict(line.split() for line in out.strip().splitlines())
    for f in fields:
        assert getattr(mcpm.CParams, f).offset == int(got[f]), f
    assert ctypes.sizeof(mcpm.CParams) == int(got["sizeof"])


def test_probe_rejects_bad_arguments_before_touching_the_device():
    lib = mcpm.load_library()
    r = mcpm.CAccessRates()
    assert lib.mcpm_probe_random_access(1 << 20, 24, ctypes.byref(r)) == 1
    assert lib.mcpm_probe_random_access(100, 16, ctypes.byref(r)) == 1
    assert lib.mcpm_probe_random_access(1 << 20, 16, None) == 1
