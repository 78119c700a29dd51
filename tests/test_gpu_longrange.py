"""GPU parity of the direct (long_range 1) and density-expansion
(long_range 2) mean-field forces and of the kernel density estimate.

Sums over quadrature points and particles are reduced in one fixed
workgroup order on both sides, so forces, the density estimate and
multi-step states are compared bit for bit."""
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


def _close(a, b, tol=1e-12):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("boundary", ["walls", "periodic"])
def test_direct_vlasov_matches_oracle(boundary):
    if boundary == "walls":
        p, arrs, _ = W.heat_conduction(seed=61)
    else:
        p, arrs, _ = W.interface(N=80000, n_cells=500, length=100.0, seed=61)
    p.long_range = mcpm.LR_DIRECT
    p.phi = -4.0
    g, o = _pair(p, arrs)
    g.vlasov_integral()
    o.vlasov_integral()
    assert np.abs(o.cells()["F2"]).max() > 0.0
    np.testing.assert_array_equal(g.cells()["F2"], o.cells()["F2"])


@pytest.mark.parametrize("boundary", ["walls", "periodic"])
@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_kde_and_density_expansion_match_oracle(boundary, precision):
    if boundary == "walls":
        p, arrs, _ = W.heat_conduction(seed=62, precision=precision)
    else:
        p, arrs, _ = W.interface(N=80000, n_cells=500, length=100.0, seed=62, precision=precision)
    p.long_range = mcpm.LR_DENSITY_EXPANSION
    p.phi = -1.0
    g, o = _pair(p, arrs)
    for a, b in zip(g.kde(), o.kde()):
        np.testing.assert_array_equal(a, b)
    g.density_expansion()
    o.density_expansion()
    np.testing.assert_array_equal(g.cells()["F2"], o.cells()["F2"])


@pytest.mark.parametrize("lr", [mcpm.LR_DIRECT, mcpm.LR_DENSITY_EXPANSION])
def test_full_steps_with_direct_and_expansion_forces(lr):
    p, arrs, meta = W.interface(N=50000, n_cells=256, length=100.0, seed=63, precision=mcpm.FP64)
    p.long_range = lr
    # the density-expansion force scales with Fn times particles per length:
    # a small phi keeps the kick at a few per cent of the thermal speed
    p.phi = -1.0 if lr == mcpm.LR_DIRECT else -1e-4
    g, o = _pair(p, arrs)
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    for s in range(5):
        sg, so = g.step(s), o.step(s)
        for k in ("candidates", "valid", "accepted"):
            assert sg[k] == so[k], (s, k)
        cg, co = g.cells(), o.cells()
        np.testing.assert_array_equal(cg["count"], co["count"])
        np.testing.assert_array_equal(cg["F2"], co["F2"])
        for k, (a, b) in enumerate(zip(g.download(), o.download())):
            np.testing.assert_array_equal(a, b, err_msg=f"component {k} step {s}")
