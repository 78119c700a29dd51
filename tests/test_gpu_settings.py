"""GPU side of the settings path: the device initializer against the
oracle's restatement (bit for bit), the PostProc file against the moments
and samples it is written from, and the settings-driven time loop."""
import os

import numpy as np
import pytest

import mcpm
import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SLAB = """problem 2
num_cells 1 100 1
box_len 4000.0 10.0 1.0
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
every 2
after 1
output slab_case
"""


def _small_injection(tmp_path):
    """The reference's injection.in on a 20 x 20 plane with 20 particles per
    cell and a 8-sigma drop."""
    text = open(os.path.join(GOLDEN, "injection.in")).read().splitlines()
    repl = {"N": "N 20", "num_cells": "num_cells 20 20 1", "box_len": "box_len 20.0 20.0 1.0",
            "drop_dia": "drop_dia 8.0", "every": "every 2", "after": "after 1"}
    out = []
    for line in text:
        k = line.split()[0] if line.strip() and not line.startswith("#") else None
        out.append(repl.get(k, line))
    f = tmp_path / "inj.in"
    f.write_text("\n".join(out) + "\n")
    return str(f)


def _slab(tmp_path):
    f = tmp_path / "slab.in"
    f.write_text(SLAB)
    return str(f)


def _pair(s, precision):
    p = mcpm.Params.from_c(s.params)
    p.precision = precision
    s.params = p.to_c()
    print(f"settings problem={s.problem} N={s.n_particles} prec={precision}", flush=True)
    return mcpm.Simulation(p), orc.OracleSim(p)


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_initialize_injection_matches_oracle(tmp_path, precision):
    s = mcpm.read_settings(_small_injection(tmp_path))
    assert s.n_drop > 0
    g, o = _pair(s, precision)
    g.initialize(s)
    o.initialize(s)
    for a, b in zip(g.download_xy(), o.download_xy()):
        np.testing.assert_array_equal(a, b)
    cg, co = g.cells(), o.cells()
    np.testing.assert_array_equal(cg["count"], co["count"])
    # the disc: the centre cells hold the liquid density
    cnt = cg["count"].reshape(20, 20)
    assert cnt[9:11, 9:11].mean() > 5 * cnt[0:3, 0:3].mean()


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_initialize_slab_matches_oracle(tmp_path, precision):
    s = mcpm.read_settings(_slab(tmp_path))
    g, o = _pair(s, precision)
    g.initialize(s)
    o.initialize(s)
    for a, b in zip(g.download(), o.download()):
        np.testing.assert_array_equal(a, b)
    y, u1, u2, u3 = g.download()
    assert abs(u2.std() - np.sqrt(1.5)) < 0.03 and abs(u1.mean()) < 0.03


def test_postproc_file_matches_moments_and_samples(tmp_path):
    s = mcpm.read_settings(_slab(tmp_path))
    p = mcpm.Params.from_c(s.params)
    g = mcpm.Simulation(p)
    g.initialize(s)
    g.init_omega_max(s.n)
    g.measure_pressure(True)
    for step in range(6):
        g.sample()
        g.step(step)
    out = tmp_path / "post"
    out.mkdir()
    g.write_postproc(s, 6, str(out))
    f = out / "slab_case_noatt_6.dat"
    lines = f.read_text().splitlines()
    header = lines[0].split()
    assert header[0] == "#variables=" and len(header) == 122 and header[-1] == "p_rep"
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[1:] if ln.strip()])
    assert rows.shape == (100, 121)
    col = {name: k for k, name in enumerate(header[1:])}
    m = g.moments()
    smp = g.samples()
    cells = g.cells()
    np.testing.assert_allclose(rows[:, col["x2"]], (np.arange(100) + 0.5) * 0.1, rtol=1e-5)
    np.testing.assert_array_equal(rows[:, col["N"]], cells["count"])
    np.testing.assert_allclose(rows[:, col["T"]], m[:, 29], rtol=1e-5)
    np.testing.assert_allclose(rows[:, col["n"]], cells["n"], rtol=1e-5)
    np.testing.assert_allclose(rows[:, col["sum_weight"]], smp[:, 0], rtol=1e-5)
    np.testing.assert_allclose(rows[:, col["sum_MMMM[5]"]], smp[:, 25], rtol=1e-5)
    np.testing.assert_allclose(rows[:, col["tavg"]], 6 * p.delta_t, rtol=1e-5)
    pr = g.pressure()
    np.testing.assert_allclose(rows[:, col["p_m"]], pr["p_m"], rtol=1e-5)
    np.testing.assert_allclose(rows[:, col["p_rep"]], pr["p_rep"], rtol=1e-5)
    np.testing.assert_allclose(rows[:, col["ptot"]], pr["p_tot"], rtol=1e-5)


def test_settings_driven_loop_on_the_plane(tmp_path):
    import run_settings
    path = _small_injection(tmp_path)
    out = tmp_path / "PostProc"
    r = run_settings.run(path, steps=5, out=str(out), log=lambda *a: None)
    assert r["written"] == [0, 2, 4]
    names = sorted(os.listdir(out))
    assert names == ["injection_ESMC_4.15_direct_%d.dat" % k for k in (0, 2, 4)]
    rows = [ln for ln in (out / names[-1]).read_text().splitlines()[1:] if ln.strip()]
    assert len(rows) == 400
    N = sum(int(ln.split()[84]) for ln in rows)
    assert N == r["particles"]
