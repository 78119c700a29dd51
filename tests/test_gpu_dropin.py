"""The reference's time loop through the drop-in C++ symbols on the GPU.

tests/_bin/dropin_driver builds GAS / BOX / CELLS and the particle arrays as
the reference's own code does and runs cell_update, Collision_ESMC_new3 (or
_new4, or for FP_dense random_generator, Velocity_FP_parallel and
DFP_extra_streaming), apply_BC, store_old_positions, stream_particles, apply_BC for a few
steps of the heat-conduction slab (200 cells, 100 particles per cell, FP64,
diffuse walls): exactly what the reference's main.cpp calls.  The state it
leaves (per-cell counts, omega_max and n, the particles listed through
cells[c].indices_inside, index[]) must equal the oracle's run of the same
steps bit for bit.  Also: particle identities through mcpm_track_ids."""
import os
import subprocess

import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "_bin", "dropin_driver")


def _write_input(path, p, arrs, omega):
    y, u1, u2, u3 = arrs
    with open(path, "wb") as f:
        np.array([y.size], np.int64).tofile(f)
        np.array([p.n_cells, 2], np.int32).tofile(f)  # problem 2: couette_flow (diffuse walls)
        np.array([p.length, p.sigma, p.mass, p.kb, p.b, p.Fn, p.cell_volume, p.delta_t,
                  p.T_wall_lo, p.T_wall_hi, p.U_wall_lo, p.U_wall_hi, p.mu, p.T0, p.visp],
                 np.float64).tofile(f)
        for a in arrs:
            np.ascontiguousarray(a, np.float64).tofile(f)
        np.ascontiguousarray(omega, np.float64).tofile(f)


def _read_output(path, n, nc):
    raw = open(path, "rb").read()
    o = 0
    cnt = np.frombuffer(raw, np.int32, nc, o); o += 4 * nc
    om = np.frombuffer(raw, np.float64, nc, o); o += 8 * nc
    nd = np.frombuffer(raw, np.float64, nc, o); o += 8 * nc
    parts = []
    for _ in range(4):
        parts.append(np.frombuffer(raw, np.float64, n, o)); o += 8 * n
    index = np.frombuffer(raw, np.int32, n, o)
    return cnt, om, nd, parts, index


@pytest.mark.parametrize("op", [3, 4, 5])
def test_reference_loop_through_dropin_symbols_matches_oracle(op, tmp_path):
    steps = 3
    p, arrs, meta = W.heat_conduction(precision=mcpm.FP64, seed=401, walls_at_sigma=True)
    p.model = {3: mcpm.MODEL_ESMC, 4: mcpm.MODEL_ESMC_NEW4, 5: mcpm.MODEL_FP_DENSE}[op]
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    o.init_omega_max(meta["n_gas"])
    omega = o.cells()["omega_max"].copy()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, p, arrs, omega)
    print(f"dropin driver op={op} N={arrs[0].size} steps={steps}", flush=True)
    env = dict(os.environ, MCPM_SEED=str(p.seed))  # the drop-in's Philox key
    r = subprocess.run([DRIVER, str(fin), str(fout), str(steps), str(op)], capture_output=True,
                       text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    for s in range(steps):
        o.step(s)
    o.cell_update()
    cnt, om, nd, parts, index = _read_output(fout, arrs[0].size, p.n_cells)
    co = o.cells()
    np.testing.assert_array_equal(cnt, co["count"])
    np.testing.assert_array_equal(om, co["omega_max"])
    np.testing.assert_array_equal(nd, co["n"])
    for k, (a, b) in enumerate(zip(parts, o.download())):
        np.testing.assert_array_equal(a, b, err_msg=f"component {k}")
    # index[] names each caller particle's cell, consistently with the lists
    assert np.bincount(index, minlength=p.n_cells).tolist() == list(cnt)


def test_track_ids_follow_each_particle_through_the_sort():
    """Free flight through a periodic slab: with identities the caller's
    own order comes back, and each particle moved by exactly u dt."""
    p, arrs, meta = W.interface(N=40000, n_cells=300, length=60.0, precision=mcpm.FP64, seed=402)
    p.model = mcpm.MODEL_NONE
    p.long_range = mcpm.LR_NONE
    g = mcpm.Simulation(p)
    g.upload(*arrs)
    g.track_ids(True)
    g.cell_update()
    y0, u1, u2, u3 = [a.copy() for a in arrs]
    ids = g.ids()
    assert sorted(ids.tolist()) == list(range(y0.size))
    y, v1, v2, v3 = g.download()
    np.testing.assert_array_equal(y, y0[ids])
    np.testing.assert_array_equal(v2, u2[ids])
    g.step(0)
    y, v1, v2, v3 = g.download()
    ids = g.ids()
    back = np.empty_like(y)
    back[ids] = y
    expect = np.mod(y0 + u2 * p.delta_t, p.length)
    d = np.abs(back - expect)
    d = np.minimum(d, p.length - d)
    assert d.max() < 1e-12
    np.testing.assert_array_equal(v2[np.argsort(ids)], u2)
