"""Batched replicas against unbatched runs: every replica of a batched
context evolves bit for bit as the same system in a context of its own
(positions, velocities, cell counts, omega_max, field), whatever the other
replicas do.  Replicas differ in cells, particles, densities, temperatures
and seeds, as in the replica workload."""
import numpy as np
import pytest

import mcpm
import workloads as W

pytestmark = pytest.mark.gpu


def _run(common, reps, arrs, metas, steps):
    print(f"batch R={len(reps)} N={sum(a[0].size for a in arrs)} bc={common.boundary} "
          f"lr={common.long_range} prec={common.precision}", flush=True)
    b = mcpm.Simulation.batch(common, reps)
    counts = [a[0].size for a in arrs]
    b.upload_batch(counts, *[np.concatenate([a[q] for a in arrs]) for q in range(4)])
    b.cell_update()
    b.init_omega_max_batch([m["n_gas"] for m in metas])
    singles = []
    for d, a, m in zip(reps, arrs, metas):
        s = mcpm.Simulation(d["params"])
        s.upload(*a)
        s.cell_update()
        s.init_omega_max(m["n_gas"])
        singles.append(s)
    for step in range(steps):
        sb = b.step(step)
        acc = sum(s.step(step)["accepted"] for s in singles)
        assert sb["accepted"] == acc
    got = b.download()
    cb = b.cells()
    off = np.cumsum([0] + counts)
    coff = np.cumsum([0] + [d["n_cells"] for d in reps])
    for r, s in enumerate(singles):
        for x, y in zip(got, s.download()):
            np.testing.assert_array_equal(x[off[r]:off[r + 1]], y)
        cs = s.cells()
        for k in ("count", "omega_max", "n"):
            np.testing.assert_array_equal(cb[k][coff[r]:coff[r + 1]], cs[k])
        if common.long_range == mcpm.LR_SCREENED_POISSON:
            np.testing.assert_array_equal(cb["F2"][coff[r]:coff[r + 1]], cs["F2"])


@pytest.mark.parametrize("precision", [mcpm.FP32, mcpm.FP64])
def test_batched_interfaces_match_unbatched_runs(precision):
    common, reps, arrs, metas = W.replicas(n_rep=6, seed0=5, N_lo=4000, N_hi=30000,
                                           cells_lo=64, cells_hi=300, length=40.0,
                                           precision=precision)
    _run(common, reps, arrs, metas, steps=3)


def test_batched_walled_slabs_without_field():
    common, reps, arrs, metas = W.replicas(n_rep=5, seed0=40, N_lo=3000, N_hi=20000,
                                           cells_lo=50, cells_hi=200, length=30.0,
                                           precision=mcpm.FP64, boundary=mcpm.BC_SPECULAR,
                                           long_range=mcpm.LR_NONE)
    _run(common, reps, arrs, metas, steps=3)


def test_batch_over_16384_cells_uses_the_two_digit_sort():
    common, reps, arrs, metas = W.replicas(n_rep=12, seed0=90, N_lo=5000, N_hi=20000,
                                           cells_lo=1500, cells_hi=2000, length=400.0)
    assert sum(d["n_cells"] for d in reps) > 16384
    _run(common, reps, arrs, metas, steps=2)


def test_batch_rejects_what_it_does_not_run():
    common, reps, arrs, metas = W.replicas(n_rep=2, seed0=1, N_lo=1000, N_hi=2000,
                                           cells_lo=20, cells_hi=30, length=20.0)
    b = mcpm.Simulation.batch(common, reps)
    with pytest.raises(mcpm.McpmError):
        b.upload(*arrs[0])
    common.model = mcpm.MODEL_FP_DENSE
    with pytest.raises(mcpm.McpmError):
        mcpm.Simulation.batch(common, reps)
