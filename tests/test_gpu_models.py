"""GPU parity of the other velocity models: Collision_DSMC_HS,
Collision_CBA_VHS and Collision_CBA_HS on the slab and the plane, and
Velocity_FP_linear on the slab.

DSMC_HS and CBA_VHS run on the collision rounds: pairs, decisions,
velocities, displaced positions and crm are compared bit for bit with the
oracle.  CBA_HS runs one thread per cell: counts, velocities and positions bit
for bit (its Vr_max starts from the cell temperature of the moments, whose
FP64 sums both sides take in one fixed workgroup order).  FP_linear: bit for
bit, as the other Fokker-Planck models."""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


def _slab(model, precision, seed, boundary=mcpm.BC_PERIODIC):
    p, arrs, meta = W.interface(N=60000, n_cells=200, length=40.0, n_liq=0.3, n_vap=0.05,
                                T=1.0, dt=0.05, precision=precision, seed=seed)
    p.long_range = mcpm.LR_NONE
    p.model = model
    p.boundary = boundary
    return p, arrs, meta


def _pair(p, arrs, xy=False):
    print(f"model={p.model} dims={p.dims} prec={p.precision} N={arrs[0].size}", flush=True)
    g = mcpm.Simulation(p)
    o = orc.OracleSim(p)
    for s in (g, o):
        (s.upload_xy if xy else s.upload)(*arrs)
        s.cell_update()
    return g, o


def _same_state(g, o, xy=False):
    a = g.download_xy() if xy else g.download()
    b = o.download_xy() if xy else o.download()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def _collide(sim, model, step):
    name = {mcpm.MODEL_DSMC_HS: "collide_dsmc_hs", mcpm.MODEL_CBA_VHS: "collide_cba_vhs",
            mcpm.MODEL_CBA_HS: "collide_cba_hs"}[model]
    return getattr(sim, name)(step)


@pytest.mark.parametrize("model", [mcpm.MODEL_DSMC_HS, mcpm.MODEL_CBA_VHS])
@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_round_models_bit_identical_on_the_slab(model, precision):
    p, arrs, meta = _slab(model, precision, seed=91)
    g, o = _pair(p, arrs)
    crm = np.full(p.n_cells, 7.0)
    g.set_omega_max(crm)
    o.set_omega_max(crm)
    g.debug_collision_log(True)
    for step in range(3):
        sg, so = _collide(g, model, step), _collide(o, model, step)
        assert (sg["candidates"], sg["valid"], sg["accepted"]) == \
            (so["candidates"], so["valid"], so["accepted"])
        assert so["accepted"] > 100
        pg, ig, jg, ag = g.collision_log(so["valid"] + 8)
        po, io, jo, ao = o.collision_log(so["valid"] + 8)
        order = np.argsort(pg, kind="stable")
        for a, b in ((pg, po), (ig, io), (jg, jo), (ag, ao)):
            np.testing.assert_array_equal(a[order], b)
        _same_state(g, o)
        np.testing.assert_array_equal(g.cells()["omega_max"], o.cells()["omega_max"])


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_cba_hs_bit_identical_on_the_slab(precision):
    p, arrs, meta = _slab(mcpm.MODEL_CBA_HS, precision, seed=92)
    g, o = _pair(p, arrs)
    for step in range(3):
        sg, so = g.collide_cba_hs(step), o.collide_cba_hs(step)
        assert (sg["candidates"], sg["accepted"]) == (so["candidates"], so["accepted"])
        assert so["accepted"] > 100
        _same_state(g, o)


@pytest.mark.parametrize("model", [mcpm.MODEL_DSMC_HS, mcpm.MODEL_CBA_VHS, mcpm.MODEL_CBA_HS])
@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_models_bit_identical_on_the_plane(model, boundary):
    p, arrs, meta = W.plane(nx=16, ny=16, Lx=16.0, Ly=16.0, ppc=30, drops=((8.0, 8.0, 4.0),),
                            boundary=boundary, seed=93, long_range=mcpm.LR_NONE, model=model,
                            dt=0.05)
    g, o = _pair(p, arrs, xy=True)
    crm = np.full(p.n_cells * p.n_cells_x, 7.0)
    g.set_omega_max(crm)
    o.set_omega_max(crm)
    for step in range(2):
        sg, so = _collide(g, model, step), _collide(o, model, step)
        assert (sg["candidates"], sg["accepted"]) == (so["candidates"], so["accepted"])
        assert so["accepted"] > 50
        _same_state(g, o, xy=True)


@pytest.mark.parametrize("model", [mcpm.MODEL_DSMC_HS, mcpm.MODEL_CBA_VHS, mcpm.MODEL_CBA_HS])
def test_models_full_steps_with_walls(model):
    """Whole steps (collisions, free flight, diffuse walls, re-sort): the
    displaced particles of CBA cross cell faces and walls like any other."""
    p, arrs, meta = W.heat_conduction(seed=94, n_cells=100, ppc=60)
    p.model = model
    g, o = _pair(p, arrs)
    crm = np.full(p.n_cells, 7.0)
    g.set_omega_max(crm)
    o.set_omega_max(crm)
    for s in range(4):
        assert g.step(s)["accepted"] == o.step(s)["accepted"]
    _same_state(g, o)


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_fp_linear_matches_oracle(precision):
    p, arrs, meta = W.heat_conduction(seed=95, precision=precision)
    p.model = mcpm.MODEL_FP_LINEAR
    g, o = _pair(p, arrs)
    for step in (0, 1):
        g.fp_closures(step)
        o.fp_closures(step)
        np.testing.assert_array_equal(g.closures()[:, 27], o.closures()[:, 27])
        g.velocity_fp(step)
        o.velocity_fp(step)
        for a, b in zip(g.download(), o.download()):
            np.testing.assert_array_equal(a, b)
    for s in range(2, 6):
        g.step(s)
        o.step(s)
        _same_state(g, o)
    np.testing.assert_array_equal(g.moments(), o.moments())
