"""GPU parity of the evaporation boundary (BC_evaporation,
Main_Functions.cpp:5444-5617, boundary MCPM_BC_REINJECT) and of the inverted
problem's thermostat (Main_Functions.cpp:1760-1830) against the oracle."""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_reinjection_steps_bit_exact(precision):
    p, arrs, meta = W.heat_conduction(n_cells=80, ppc=60, length=10.0, precision=precision,
                                      dt_frac=1.0, seed=7)
    p.boundary = mcpm.BC_REINJECT
    p.T_wall_lo = 1.2
    print(f"reinjection precision={precision}", flush=True)
    g = mcpm.Simulation(p)
    o = orc.OracleSim(p)
    for s in (g, o):
        s.upload(*arrs)
        s.cell_update()
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    for step in range(4):
        sg = g.step(step)
        so = o.step(step)
        keys = ("candidates", "valid", "accepted")
        assert {k: sg[k] for k in keys} == {k: so[k] for k in keys}, (step, sg, so)
    assert g.escaped() == o.escaped() > 0
    for a, b in zip(g.download(), o.download()):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_thermostat_matches_oracle(precision):
    p, (y, u1, u2, u3), _ = W.heat_conduction(n_cells=64, ppc=80, length=10.0,
                                              precision=precision, seed=9)
    u1 = u1 + 0.3 * np.sin(y)
    print(f"thermostat precision={precision}", flush=True)
    g = mcpm.Simulation(p)
    o = orc.OracleSim(p)
    for s in (g, o):
        s.upload(y, u1, u2, u3)
        s.cell_update()
        s.thermostat(1.7, 0.5, 2.2, 2.5)
    tol = 1e-12 if precision == mcpm.FP64 else 2e-6
    for a, b in zip(g.download(), o.download()):
        assert np.max(np.abs(a - b)) <= tol * max(np.max(np.abs(b)), 1.0)
