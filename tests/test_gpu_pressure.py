"""GPU parity of the pressure observables: measure_pressure's face sums p_m,
accumulated on the device by the push, and PostProc's ptot / p_rep columns.

The runs below are bit-identical to the oracle particle by particle (no mean
field), so every face sees the same crossings; the sums themselves are added
in another order (device atomics), compared within 1e-12 relative.  ptot and
p_rep come from the moments' T (FP64 sums in another order): 1e-12."""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


def _pair(p, arrs):
    print(f"model={p.model} bc={p.boundary} prec={p.precision} N={arrs[0].size}", flush=True)
    g = mcpm.Simulation(p)
    o = orc.OracleSim(p)
    for s in (g, o):
        s.upload(*arrs)
        s.cell_update()
    return g, o


def _close(a, b, tol):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_pressure_with_diffuse_walls_and_esmc(precision):
    p, arrs, meta = W.heat_conduction(seed=71, n_cells=100, ppc=80, precision=precision)
    g, o = _pair(p, arrs)
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    for s in (g, o):
        s.measure_pressure(True)
    for step in range(4):
        assert g.step(step)["accepted"] == o.step(step)["accepted"]
    for a, b in zip(g.download(), o.download()):
        np.testing.assert_array_equal(a, b)
    pg, po = g.pressure(), o.pressure()
    assert po["p_m"][2:-2].min() > 0
    for k in ("p_m", "p_rep", "p_tot"):
        assert _close(pg[k], po[k], 1e-12), k
    # reset clears the face sums on both
    g.reset_samples()
    assert not g.pressure()["p_m"].any()


def test_pressure_on_a_periodic_slab_counts_the_unwrapped_flight():
    p, arrs, meta = W.interface(N=100_000, n_cells=64, length=32.0, n_liq=0.1, n_vap=0.1,
                                T=1.0, dt=0.2, precision=mcpm.FP64, seed=72)
    p.long_range = mcpm.LR_NONE
    p.model = mcpm.MODEL_NONE
    g, o = _pair(p, arrs)
    for s in (g, o):
        s.measure_pressure(True)
    for step in range(3):
        g.stream_bc(step, False)
        o.stream_bc(step, False)
    pg, po = g.pressure(), o.pressure()
    assert _close(pg["p_m"], po["p_m"], 1e-12)
    # the faces next to the wrap see as many crossings as the others
    assert po["p_m"][-1] > 0.5 * po["p_m"].mean()


def test_pressure_off_by_default_and_rejected_on_the_plane():
    p, arrs, meta = W.heat_conduction(seed=73, n_cells=20, ppc=20)
    g, _ = _pair(p, arrs)
    g.stream_bc(0, False)
    assert not g.pressure()["p_m"].any()
    p2, arrs2, _ = W.plane(nx=6, ny=5, Lx=6.0, Ly=5.0, ppc=10, drops=(), seed=74,
                           long_range=mcpm.LR_NONE)
    g2 = mcpm.Simulation(p2)
    with pytest.raises(mcpm.McpmError):
        g2.measure_pressure(True)
