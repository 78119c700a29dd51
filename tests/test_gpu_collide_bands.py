"""GPU parity of the collision schedule's priority bands and launch knobs.

Candidates without a partner (the point r_i + sigma dir beyond a wall,
collisions.cpp:668-671, or an empty partner cell) stay in the band lists as
empty slots; every band's admission must skip them.  A walled slab makes
them common (about one candidate in ten).  Pairs, decisions and velocities
must not depend on the number of bands, the round grid or the operations per
thread of a round: compared bit for bit with the oracle."""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bands,grid,unroll", [(1, 8, 4), (3, 64, 2), (12, 2048, 2),
                                               (16, 1024, 4), (12, 1024, 0)])
def test_esmc_bands_with_partnerless_candidates(bands, grid, unroll):
    print(f"bands={bands} grid={grid} unroll={unroll}", flush=True)
    p, arrs, meta = W.heat_conduction(n_cells=100, ppc=30, seed=71)
    g = mcpm.Simulation(p)
    o = orc.OracleSim(p)
    for s in (g, o):
        s.upload(*arrs)
        s.cell_update()
    g.set_tuning(mcpm.TUNE_BANDS, bands)
    g.set_tuning(mcpm.TUNE_ROUND_GRID, grid)
    g.set_tuning(mcpm.TUNE_ROUND_UNROLL, unroll)
    assert g.get_tuning(mcpm.TUNE_BANDS) == bands
    assert g.get_tuning(mcpm.TUNE_ROUND_UNROLL) == unroll
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    g.debug_collision_log(True)
    for step in range(3):
        sg, so = g.collide_esmc(step), o.collide_esmc(step)
        assert so["candidates"] > so["valid"]  # partnerless candidates present
        for k in ("candidates", "valid", "accepted"):
            assert sg[k] == so[k]
        lg = g.collision_log(sg["valid"] + 16)
        lo = o.collision_log(so["valid"] + 16)
        order = np.argsort(lg[0], kind="stable")
        for a, b in zip(lg, lo):
            assert np.array_equal(a[order], b)
        for a, b in zip(g.download(), o.download()):
            assert np.array_equal(a, b)
        assert np.array_equal(g.cells()["omega_max"], o.cells()["omega_max"])


def test_tuning_rejects_bad_values():
    p, arrs, meta = W.heat_conduction(n_cells=20, ppc=10, seed=72)
    g = mcpm.Simulation(p)
    for knob, bad in ((mcpm.TUNE_BANDS, 0), (mcpm.TUNE_BANDS, 17), (mcpm.TUNE_ROUND_GRID, 12),
                      (mcpm.TUNE_SORT_CHUNK, 1000), (mcpm.TUNE_ROUND_UNROLL, 3), (99, 1)):
        with pytest.raises(mcpm.McpmError):
            g.set_tuning(knob, bad)
