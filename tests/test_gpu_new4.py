"""Collision_ESMC_new4 on the GPU against the oracle, bit for bit: the
candidate log (rank in new4's cell order, i, j, accept), velocities and
omega_max, walled (clamped partner cell) and periodic slabs, FP64 and FP32,
and whole steps with the operator as the context's model."""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


def _pair(p, arrs):
    g, o = mcpm.Simulation(p), orc.OracleSim(p)
    for s in (g, o):
        s.upload(*arrs)
        s.cell_update()
    return g, o


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
@pytest.mark.parametrize("boundary", ["walls", "periodic"])
def test_new4_collisions_bit_identical(precision, boundary):
    if boundary == "walls":
        p, arrs, meta = W.heat_conduction(precision=precision, seed=601)
    else:
        p, arrs, meta = W.interface(N=40000, n_cells=256, length=100.0, precision=precision,
                                    seed=601)
    print(f"new4 {boundary} prec={precision}", flush=True)
    g, o = _pair(p, arrs)
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    g.debug_collision_log(True)
    for step in range(3):
        sg, so = g.collide_esmc_new4(step), o.collide_esmc_new4(step)
        for k in ("candidates", "valid", "accepted"):
            assert sg[k] == so[k], (step, k)
        cap = so["valid"] + 16
        pg, ig, jg, ag = g.collision_log(cap)
        po, io, jo, ao = o.collision_log(cap)
        order = np.argsort(pg, kind="stable")
        np.testing.assert_array_equal(pg[order], po)
        np.testing.assert_array_equal(ig[order], io)
        np.testing.assert_array_equal(jg[order], jo)
        np.testing.assert_array_equal(ag[order], ao)
        for a, b in zip(g.download(), o.download()):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(g.cells()["omega_max"], o.cells()["omega_max"])


def test_new4_whole_steps():
    p, arrs, meta = W.heat_conduction(precision=mcpm.FP64, seed=602)
    p.model = mcpm.MODEL_ESMC_NEW4
    g, o = _pair(p, arrs)
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    for s in range(4):
        sg, so = g.step(s), o.step(s)
        assert (sg["candidates"], sg["valid"], sg["accepted"]) == \
               (so["candidates"], so["valid"], so["accepted"])
        assert sg["rounds"] >= 1
    for a, b in zip(g.download(), o.download()):
        np.testing.assert_array_equal(a, b)
