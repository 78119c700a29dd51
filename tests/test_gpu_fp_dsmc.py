"""GPU parity of the Fokker-Planck and plain-DSMC models against the oracle.

FP: bit-exact.  The per-cell moments, the noise mean and rms and the energy
rescale are sums in the same fixed workgroup order on both sides, the
closures are the same Gauss-Jordan sequence, so closures (phi included),
velocities and multi-step states agree bit for bit, in FP64 and FP32.  The
cell sizes cover the HIP kernel's register noise cache (up to 256 particles
per cached slot, several slots) and its redraw fallback beyond 3072
particles.  DSMC: bit-exact pairs, decisions and velocities.
"""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


def _pair(p, arrs):
    g = mcpm.Simulation(p)
    o = orc.OracleSim(p)
    for s in (g, o):
        s.upload(*arrs)
        s.cell_update()
    return g, o


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _fp_case(model, precision, ppc, n_cells, seed):
    p, arrs, meta = W.heat_conduction(n_cells=n_cells, ppc=ppc, seed=seed, precision=precision)
    p.model = model
    return p, arrs


def _check_fp_pair(g, o, steps):
    for step in steps:
        g.fp_closures(step)
        o.fp_closures(step)
        np.testing.assert_array_equal(g.closures(), o.closures())
        g.velocity_fp(step)
        o.velocity_fp(step)
        for k, (a, b) in enumerate(zip(g.download(), o.download())):
            np.testing.assert_array_equal(a, b, err_msg=f"component {k} step {step}")


@pytest.mark.parametrize("model", [mcpm.MODEL_FP_DENSE, mcpm.MODEL_FP_IDEAL])
@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_fp_closures_and_velocity_match_oracle(model, precision):
    p, arrs = _fp_case(model, precision, 100, 200, 51)
    g, o = _pair(p, arrs)
    _check_fp_pair(g, o, (0, 2))


@pytest.mark.parametrize("model", [mcpm.MODEL_FP_DENSE, mcpm.MODEL_FP_LINEAR])
@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
@pytest.mark.parametrize("ppc", [300, 1700, 3500])
def test_fp_large_cells_match_oracle(model, precision, ppc):
    """Cells of 300 and 1700 particles use cache slots 0..1 and 0..6 of the
    noise kept in registers; 3500 per cell (above 12 x 256) redraws the
    noise in every pass."""
    p, arrs = _fp_case(model, precision, ppc, 12, 55 + ppc)
    g, o = _pair(p, arrs)
    assert g.cells()["count"].max() > (3072 if ppc > 3072 else 256)
    _check_fp_pair(g, o, (0, 1))


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_fp_dense_full_steps_match_oracle(precision):
    p, arrs, meta = W.heat_conduction(seed=52, precision=precision)
    p.model = mcpm.MODEL_FP_DENSE
    g, o = _pair(p, arrs)
    for s in range(6):
        g.step(s)
        o.step(s)
        cg, co = g.cells(), o.cells()
        np.testing.assert_array_equal(cg["count"], co["count"])
        for k, (a, b) in enumerate(zip(g.download(), o.download())):
            np.testing.assert_array_equal(a, b, err_msg=f"component {k} step {s}")
    np.testing.assert_array_equal(g.moments(), o.moments())


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_dsmc_vhs_bit_identical(precision):
    p, arrs, meta = W.interface(N=60000, n_cells=200, length=40.0, n_liq=0.2, n_vap=0.05,
                                T=1.0, dt=0.05, precision=precision, seed=53)
    p.long_range = mcpm.LR_NONE
    p.model = mcpm.MODEL_DSMC_VHS
    p.visp = 0.7  # VHS exponents 0.6 and 0.4 go through the deterministic log/exp
    g, o = _pair(p, arrs)
    o.init_omega_max(meta["n_gas"], times_y=False)
    g.set_omega_max(o.cells()["omega_max"])
    g.debug_collision_log(True)
    for step in range(3):
        sg = g.collide_dsmc_vhs(step)
        so = o.collide_dsmc_vhs(step)
        assert (sg["candidates"], sg["accepted"]) == (so["candidates"], so["accepted"])
        pg, ig, jg, ag = g.collision_log(so["valid"] + 8)
        po, io, jo, ao = o.collision_log(so["valid"] + 8)
        order = np.argsort(pg, kind="stable")
        np.testing.assert_array_equal(pg[order], po)
        np.testing.assert_array_equal(ig[order], io)
        np.testing.assert_array_equal(jg[order], jo)
        np.testing.assert_array_equal(ag[order], ao)
        for a, b in zip(g.download(), o.download()):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(g.cells()["omega_max"], o.cells()["omega_max"])


def test_dsmc_full_steps_with_walls_bit_identical():
    p, arrs, meta = W.heat_conduction(seed=54)
    p.model = mcpm.MODEL_DSMC_VHS
    g, o = _pair(p, arrs)
    o.init_omega_max(meta["n_gas"], times_y=False)
    g.set_omega_max(o.cells()["omega_max"])
    for s in range(4):
        assert g.step(s)["accepted"] == o.step(s)["accepted"]
    for a, b in zip(g.download(), o.download()):
        np.testing.assert_array_equal(a, b)
