"""GPU parity past one LDS histogram: more than 16384 cells.

The sort then runs two stable passes over two digits of the cell key (low,
then high) and rebuilds the cell bounds from the sorted keys; the plane's
FEM solve runs over a grid of workgroups once the plane has more than 129^2
nodes.  Order, counts, deposit, collisions and whole steps against the
oracle bit for bit (its stable counting sort is the single-pass definition);
the FEM field within 1e-10 of the oracle's banded LU, and on a uniform
periodic lattice against the closed form."""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


def _cells_equal(g, o):
    cg, co = g.cells(), o.cells()
    for k in ("count", "start", "n"):
        np.testing.assert_array_equal(cg[k], co[k])


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_slab_with_20000_cells(precision):
    p, arrs, meta = W.heat_conduction(n_cells=20000, ppc=10, seed=101, precision=precision)
    print(f"slab nc={p.n_cells} N={arrs[0].size} prec={precision}", flush=True)
    g, o = mcpm.Simulation(p), orc.OracleSim(p)
    for s in (g, o):
        s.upload(*arrs)
        s.cell_update()
    _cells_equal(g, o)
    for a, b in zip(g.download(), o.download()):
        np.testing.assert_array_equal(a, b)
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    g.debug_collision_log(True)
    for step in range(3):
        sg, so = g.step(step), o.step(step)
        assert (sg["candidates"], sg["accepted"]) == (so["candidates"], so["accepted"])
        _cells_equal(g, o)
        for a, b in zip(g.download(), o.download()):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_plane_160x160_sort_deposit_and_esmc(boundary):
    p, arrs, meta = W.plane(nx=160, ny=160, Lx=160.0, Ly=160.0, ppc=12,
                            drops=((60.0, 80.0, 20.0), (120.0, 40.0, 12.0)), boundary=boundary,
                            seed=102, long_range=mcpm.LR_NONE, precision=mcpm.FP32)
    print(f"plane 160x160 N={arrs[0].size} bc={boundary}", flush=True)
    g, o = mcpm.Simulation(p), orc.OracleSim(p)
    for s in (g, o):
        s.upload_xy(*arrs)
        s.cell_update()
    _cells_equal(g, o)
    np.testing.assert_array_equal(g.nodes()[0], o.nodes()[0])
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    for step in range(2):
        sg, so = g.step(step), o.step(step)
        assert (sg["candidates"], sg["accepted"]) == (so["candidates"], so["accepted"])
        _cells_equal(g, o)
        for a, b in zip(g.download_xy(), o.download_xy()):
            np.testing.assert_array_equal(a, b)
        # the deposit of the moved particles (from the sort keys' quarters)
        np.testing.assert_array_equal(g.nodes()[0], o.nodes()[0])


def test_plane_160x160_screened_poisson_walls_grid_cg():
    p, arrs, _ = W.plane(nx=160, ny=160, Lx=160.0, Ly=160.0, ppc=6,
                         drops=((80.0, 80.0, 30.0),), boundary=mcpm.BC_SPECULAR, seed=103,
                         long_range=mcpm.LR_SCREENED_POISSON)
    print(f"plane 160x160 SP N={arrs[0].size}", flush=True)
    g, o = mcpm.Simulation(p), orc.OracleSim(p)
    for s in (g, o):
        s.upload_xy(*arrs)
        s.cell_update()
    g.poisson_solver()
    o.poisson_solver()
    _, pg = g.nodes()
    _, po = o.nodes()
    assert np.abs(pg - po).max() <= 1e-10 * np.abs(po).max()


def test_plane_192x192_periodic_lattice_closed_form():
    p, arrs, _ = W.plane(nx=192, ny=192, Lx=192.0, Ly=192.0, lattice=True,
                         boundary=mcpm.BC_PERIODIC, long_range=mcpm.LR_SCREENED_POISSON)
    print(f"plane 192x192 lattice N={arrs[0].size}", flush=True)
    g = mcpm.Simulation(p)
    g.upload_xy(*arrs)
    g.cell_update()
    g.poisson_solver()
    _, pot = g.nodes()
    lam = W.YUKAWA_BST
    n = 4 * p.Fn / p.cell_volume
    assert np.abs(pot / (n * (1.0 + lam) * np.exp(-lam) / lam ** 2) - 1.0).max() < 1e-12
