"""The reference's C++ call surface (include/mcpm_dropin.h, libmcpm_dropin.so),
on the CPU: the structs have the reference's layout, the library exports
the reference's functions under their C++ signatures, and without a GPU a
call fails with a catchable error instead of ending the process."""
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "monte-carlo-particle-method_amd", "libmcpm_dropin.so")
DRIVER = os.path.join(ROOT, "tests", "_bin", "dropin_driver")

# Member offsets of the reference's structs under g++ 11 on x86-64, as a
# probe of src/cpp_headers.h reports them (the fields the step reads and
# writes), and the struct sizes.
OFFSETS = {
    "CELLS": {"id": 76, "n_n": 96, "volume": 360, "indices_inside": 368, "num_inside": 376,
              "alpha": 384, "U_space": 552, "T": 576, "n": 584, "PIJ": 616, "M4": 664,
              "M3": 760, "DM2": 840, "c": 856, "gamma": 904, "omega_max": 1936,
              "Lambda": 1944, "F2": 1968},
    "GAS": {"model": 616, "output_name": 648},
    "BOX": {"Len": 416, "delta_dim": 440, "N": 536, "file_name": 752, "step": 800},
}
SIZES = {"GAS": 872, "BOX": 816, "CELLS": 1984}

SIGNATURES = [
    "cell_update(double*, double*, double*, double*, double*, GAS*, CELLS*, BOX*, int*, "
    "double**, double*, int*)",
    "Collision_ESMC_new3(double*, double*, double*, double*, double*, GAS*, BOX*, CELLS*, int*)",
    "Collision_ESMC_new4(double*, double*, double*, double*, double*, GAS*, BOX*, CELLS*, int*)",
    "Collision_DSMC_VHS(double*, double*, double*, double*, double*, GAS*, BOX*, CELLS*, "
    "Collision*)",
    "random_generator(double*, double*, double*, double*, double*, double*, CELLS*, BOX*, GAS*, "
    "int*)",
    "Velocity_FP_parallel(double*, double*, double*, GAS*, CELLS*, BOX*, int*, double*, double*, "
    "double*, double*, double*, double*)",
    "c_gamma_calculation(double*, double*, double*, GAS*, CELLS*, BOX*)",
    "alpha_calculation(double*, double*, double*, GAS*, CELLS*, BOX*)",
    "DFP_extra_streaming(double*, double*, double*, double*, double*, GAS*, BOX*, CELLS*, int*, "
    "double)",
    "poisson_solver(double*, double*, double*, double*, GAS*, BOX*, CELLS*, int*)",
    "vlasov_taking_the_integral(double*, GAS*, BOX*, CELLS*, int*, double*, double*, double*)",
    "compute_SP_2D(GAS*, BOX*, CELLS*, double*, double*, double*, double*, int*)",
    "compute_Vlasov_2D(GAS*, BOX*, CELLS*, double*, double*, double*, double*, int*)",
    "stream_particles(double*, double*, double*, double*, double*, double*, double*, double*, "
    "double*, GAS*, BOX*, CELLS*, int*)",
    "apply_BC(double*, double*, double*, double*, double*, double*, double*, double*, double*, "
    "GAS*, BOX*, CELLS*, int*, int*, double*, double*, double*, double*, double*, double*, "
    "double*, double*, double*, double*)",
    "increase_collision_rate(double, GAS*)",
]


def test_reference_struct_layout(tmp_path):
    lines = []
    for st, fields in OFFSETS.items():
        lines.append(f'  printf("{st} sizeof %zu\\n", sizeof({st}));')
        for f in fields:
            lines.append(f'  printf("{st} {f} %zu\\n", offsetof({st}, {f}));')
    src = tmp_path / "probe.cpp"
    src.write_text("#include <cstddef>\n#include <cstdio>\n#include \"mcpm_dropin.h\"\n"
                   "#pragma GCC diagnostic ignored \"-Winvalid-offsetof\"\n"
                   "int main() {\n" + "\n".join(lines) + "\n  return 0;\n}\n")
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.strip().splitlines():
        st, f, v = line.split()
        got[(st, f)] = int(v)
    for st, fields in OFFSETS.items():
        assert got[(st, "sizeof")] == SIZES[st], st
        for f, off in fields.items():
            assert got[(st, f)] == off, (st, f)


def test_dropin_exports_the_reference_signatures():
    assert os.path.exists(DROPIN)
    out = subprocess.run(["nm", "-DC", "--defined-only", DROPIN], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (.+)$", out, flags=re.M))
    missing = [s for s in SIGNATURES if s not in exported]
    assert not missing, missing
    # and under the reference's mangled names (the ones its objects reference)
    raw = subprocess.run(["nm", "-D", "--defined-only", DROPIN], capture_output=True, text=True,
                         check=True).stdout
    assert "_Z19Collision_ESMC_new3PdS_S_S_S_P3GASP3BOXP5CELLSPi" in raw
    assert "_Z11cell_updatePdS_S_S_S_P3GASP5CELLSP3BOXPiPS_S_S6_" in raw


def test_driver_reports_a_missing_device_as_an_error(tmp_path):
    """No GPU here: the first call throws mcpm_dropin_error, the driver
    catches it and exits with its own status; nothing aborts."""
    import torch
    if torch.cuda.is_available():
        return  # a GPU machine: the GPU test runs the driver
    fin = tmp_path / "in.bin"
    n, nc = 10, 2
    with open(fin, "wb") as f:
        np.array([n], np.int64).tofile(f)
        np.array([nc, 2], np.int32).tofile(f)
        np.array([2.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 0.01, 1.0, 1.0, 0.0, 0.0, 0.1, 1.0,
                  0.5]).tofile(f)
        np.linspace(0.1, 1.9, n).tofile(f)
        np.zeros(3 * n).tofile(f)
        np.ones(nc).tofile(f)
    r = subprocess.run([DRIVER, str(fin), str(tmp_path / "out.bin"), "1", "3"],
                       capture_output=True, text=True)
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert "drop-in error 2" in r.stderr
