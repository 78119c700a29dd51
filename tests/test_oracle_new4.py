"""The oracle's Collision_ESMC_new4 (collisions.cpp:722-837), pinned by what
the operator's text fixes: its order (cells from the last to the first,
each cell's candidates in turn, so the log's ranks rise), its candidate
count (cells of one particle draw none), its clamped partner cell (no
candidate loses its partner at a wall), exact momentum and energy
conservation, and the Enskog collision rate it shares with new3."""
import numpy as np

import mcpm
import oracle as orc
import workloads as W


def _sim(seed, **kw):
    p, arrs, meta = W.heat_conduction(seed=seed, **kw)
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    o.init_omega_max(meta["n_gas"])
    return p, o, meta


def test_new4_order_counts_and_conservation():
    p, o, meta = _sim(501)
    c0 = o.cells()
    _, u1, u2, u3 = o.download()
    P0 = np.array([u1.sum(), u2.sum(), u3.sum()])
    E0 = (u1 * u1 + u2 * u2 + u3 * u3).sum()
    st = o.collide_esmc_new4(0)
    pr, i, j, a = o.collision_log(st["valid"] + 8)
    assert len(pr) == st["valid"]
    # the log is in processing order: ranks rise, the i cells never rise
    assert np.all(np.diff(pr.astype(np.int64)) > 0)
    cell_i = np.searchsorted(c0["start"], i, side="right") - 1
    assert np.all(np.diff(cell_i) <= 0)
    # walls: the clamp gives every candidate a partner cell, so only i == j
    # (or an empty partner cell) drops one
    assert st["candidates"] - st["valid"] < 0.002 * st["candidates"]
    _, u1, u2, u3 = o.download()
    P1 = np.array([u1.sum(), u2.sum(), u3.sum()])
    assert np.abs(P1 - P0).max() < 1e-9
    assert abs((u1 * u1 + u2 * u2 + u3 * u3).sum() - E0) < 1e-9 * E0
    assert st["accepted"] > 0


def test_new4_draws_nothing_in_single_particle_cells():
    p, arrs, meta = W.heat_conduction(seed=502, n_cells=50, ppc=4)
    y = arrs[0].copy()
    y[np.floor(y / p.dy) == 7] = 7.5 * p.dy  # pile cell 7's particles into cell 7 ...
    keep = np.ones(y.size, bool)
    idx = np.where(np.floor(y / p.dy) == 7)[0]
    keep[idx[1:]] = False  # ... and leave one
    arrs = tuple(a[keep] for a in (y,) + tuple(arrs[1:]))
    p.capacity = arrs[0].size
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    om = np.full(p.n_cells, 3.0)
    o.set_omega_max(om)
    st = o.collide_esmc_new4(0)
    pr, i, j, a = o.collision_log(st["valid"] + 8)
    c = o.cells()
    assert c["count"][7] == 1
    cell_i = np.searchsorted(c["start"], i, side="right") - 1
    assert 7 not in set(cell_i.tolist())


def test_new4_and_new3_share_the_enskog_rate():
    acc3, acc4 = [], []
    for s in range(3):
        p, o, meta = _sim(510 + s)
        acc4.append(o.collide_esmc_new4(0)["accepted"])
        p, o, meta = _sim(510 + s)
        acc3.append(o.collide_esmc(0)["accepted"])
    r = np.mean(acc4) / np.mean(acc3)
    assert abs(r - 1.0) < 0.05
