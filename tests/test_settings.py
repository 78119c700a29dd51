"""The settings-file reader against the reference's own input file
(tests/golden/injection.in, the reference's src/injection.in) and small
files written here, checked by known answers of readSettingsFile's formulas
(myIO.cpp:1756-3016) and by the consistency relations they imply.  CPU only:
reading a file needs no GPU."""
import math
import os

import pytest

import mcpm

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PI = math.pi


def _write(tmp_path, text, name="case.in"):
    f = tmp_path / name
    f.write_text(text)
    return str(f)


def test_injection_settings_known_answers():
    s = mcpm.read_settings(os.path.join(GOLDEN, "injection.in"))
    p = mcpm.Params.from_c(s.params)
    sigma = 3.405e-10                     # myIO.cpp:2732
    kb, m = 1.38064852e-23, 6.6335214e-26
    nc, nh = 0.1 / sigma ** 3, 1.0 / sigma ** 3
    L = 40.0 * sigma
    N_bg = 500 * 40 * 40
    Fn = nc * L * L * 1.0 / N_bg
    R = 20.0 * sigma / 2.0
    Nh = math.floor((nh - nc) * PI * (2 * R) ** 2 / 4.0 / Fn)
    assert s.problem == 12 and p.dims == 2
    assert (p.n_cells_x, p.n_cells) == (40, 40)
    assert p.boundary == mcpm.BC_SPECULAR and p.boundary_x == mcpm.BC_SPECULAR
    assert p.sigma == sigma and p.eqdist == sigma
    assert p.length_x == pytest.approx(L, rel=1e-15) and p.length == pytest.approx(L, rel=1e-15)
    assert p.Fn == pytest.approx(Fn, rel=1e-14)
    assert s.n_drop == Nh and s.n_particles == N_bg + Nh
    # consistency: the background carries nc, the disc carries (nh - nc) pi R^2
    assert p.Fn * N_bg / (L * L) == pytest.approx(nc, rel=1e-13)
    assert s.n_drop * p.Fn == pytest.approx((nh - nc) * PI * R * R, rel=1e-6)
    assert p.cell_volume == pytest.approx((L / 40) ** 2 * 1.0, rel=1e-14)
    assert p.model == mcpm.MODEL_ESMC and p.long_range == mcpm.LR_DIRECT
    assert s.model == b"ESMC" and s.output == b"injection_ESMC_4.15"
    # b, mu(T0 = 273) from sigma (myIO.cpp:2828-2829), phi = -4.15 eps, eps = 119.8 kB
    assert p.b == pytest.approx(2 * PI * sigma ** 3 / 3, rel=1e-15)
    mu = 5.0 / (16.0 * sigma * sigma) * math.sqrt(m * kb * 273.0 / PI)
    assert p.mu == pytest.approx(mu, rel=1e-14) and p.T0 == 273.0
    assert p.phi == pytest.approx(-4.15 * 119.8 * kb, rel=1e-14)
    assert (p.ast, p.bst) == (-1.64835851e-28, 6.91304716e+09)
    # dt from the mean free path (myIO.cpp:2840-2861): frac_mft min(lambda, dy) / v_wall
    lam = 4.0 * 1.4 * (5 - 1.0) * (7 - 1.0) / (5.0 * 2.4 * 3.4)
    lam = lam * math.sqrt(m / (2 * PI * kb * 100.0)) * mu / (m * nc)
    assert s.lambda_ == pytest.approx(lam, rel=1e-13)
    dt = 0.01 * min(lam, L / 40) / max(math.sqrt(kb * 100.0 / m), 0.0)
    assert p.delta_t == pytest.approx(dt, rel=1e-13)
    assert s.T_init == 100.0 and list(s.U0) == [0.0, 0.0, 0.0]
    assert list(s.direction) == [1, 1, 0]
    assert (s.every, s.after, s.reset_every, s.num_steps) == (10, 9, 1, 100000000)


def test_couette_slab_settings(tmp_path):
    text = """# heat conduction between diffuse walls, reduced units
problem 2
num_cells 1 200 1
box_len 5000.0 20.0 1.0
N 20000
n 0.2
sigma 1.0
mass 1.0
kb 1.0
T 1.5
T_wall_1 1.0
T_wall_2 2.0
U_wall_1 0.0
U_wall_2 0.0
U0 0 0 0
direction 0 1 0
dt 0.01
model ESMC
long_range 0
"""
    s = mcpm.read_settings(_write(tmp_path, text))
    p = mcpm.Params.from_c(s.params)
    assert p.dims == 1 and p.n_cells == 200 and p.length == 20.0
    assert p.boundary == mcpm.BC_DIFFUSE
    assert p.wall_lo == 1.0 + 1e-15 and p.wall_hi == 20.0 - 1.0 - 1e-15
    assert (p.T_wall_lo, p.T_wall_hi) == (1.0, 2.0)
    assert p.Fn == pytest.approx(0.2 * 5000.0 * 20.0 / 20000)
    assert p.cell_volume == pytest.approx(5000.0 * 0.1 * 1.0)
    assert p.delta_t == 0.01 and p.sigma == 1.0
    assert p.crref == pytest.approx(math.sqrt(1.5))    # sigma = 1: crref = sqrt(kB T / m)
    assert p.eqdist == pytest.approx(2 ** (1 / 6))
    # initialization's temperature: max(T, (T_w1 + T_w2) / 2)
    assert s.T_init == 1.5
    assert s.n_particles == 20000 and s.n_drop == 0


def test_defaults_comments_and_trailing_text(tmp_path):
    text = """#comment lines are skipped
problem 1 ##closed box
num_cells 1 10 1
N 100
model none
"""
    s = mcpm.read_settings(_write(tmp_path, text))
    p = mcpm.Params.from_c(s.params)
    assert p.boundary == mcpm.BC_SPECULAR and p.model == mcpm.MODEL_NONE
    # defaults (myIO.cpp:2137-2186): kB, T = 1 / kB, n = 1e-3, m = 1, Len = 1, U0 = 1
    assert p.kb == 1.38064852e-23
    assert s.T == pytest.approx(1.0 / 1.38064852e-23)
    assert s.n == 1e-3 and p.mass == 1.0 and p.length == 1.0
    assert list(s.U0) == [1.0, 1.0, 1.0]
    assert (s.num_steps, s.every, s.after) == (10, 1, 1)
    assert s.output == b"data.txt"


@pytest.mark.parametrize("bad", ["frobnicate 3\n", " #indented comment is a keyword\n",
                                 "problem 3\n", "model SPH\n", "long_range 7\n"])
def test_rejects_what_the_reference_or_this_path_cannot_run(tmp_path, bad):
    text = "problem 2\nnum_cells 1 10 1\nN 100\n" + bad
    with pytest.raises(mcpm.McpmError):
        mcpm.read_settings(_write(tmp_path, text))


def test_integer_fields_read_like_operator_extraction(tmp_path):
    # "N 1e4" reads 1 into an integer field, as iss >> (*gas).N does
    text = "problem 2\nnum_cells 1 10 1\nN 1e4\n"
    s = mcpm.read_settings(_write(tmp_path, text))
    assert s.n_particles == 1


def test_settings_struct_layout_matches_header(tmp_path):
    import ctypes
    import subprocess
    fields = [f[0] for f in mcpm.CSettings._fields_]
    cname = {"lambda_": "lambda"}
    src = tmp_path / "probe.c"
    body = "".join(f'  printf("%s %zu\\n", "{f}", offsetof(mcpm_settings, {cname.get(f, f)}));\n'
                   for f in fields)
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"mcpm.h\"\n"
                   "int main(void) {\n" + body +
                   '  printf("sizeof %zu\\n", sizeof(mcpm_settings));\n  return 0;\n}\n')
    exe = tmp_path / "probe"
    inc = os.path.join(os.path.dirname(GOLDEN), "..", "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = dict(line.split() for line in out.strip().splitlines())
    for f in fields:
        assert getattr(mcpm.CSettings, f).offset == int(got[f]), f
    assert ctypes.sizeof(mcpm.CSettings) == int(got["sizeof"])
