"""The windowed sort (per-chunk histograms over a few cells) against the
full-histogram sort and the oracle: the same stable order bit for bit, on a
walled slab, a periodic slab (windows wrap), batched replicas (chunks never
span two replicas; also past 65536 cells, where a wavefront scans several
cells) and when a particle leaves its window (the sort is redone by the full
histogram)."""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


def _runs(g):
    return g.get_tuning(mcpm.TUNE_SORT_WINDOW_RUNS), g.get_tuning(mcpm.TUNE_SORT_WINDOW_REDONE)


@pytest.mark.parametrize("case", ["walls", "periodic"])
def test_windowed_sort_matches_oracle_and_full_histogram(case):
    if case == "walls":
        p, arrs, meta = W.heat_conduction(precision=mcpm.FP32, seed=701)
    else:
        p, arrs, meta = W.interface(N=200000, n_cells=512, length=100.0, precision=mcpm.FP32,
                                    seed=701)
    sims = []
    for window in (1, 0):
        g = mcpm.Simulation(p)
        g.set_tuning(mcpm.TUNE_SORT_WINDOW, window)
        g.upload(*arrs)
        g.cell_update()
        g.init_omega_max(meta["n_gas"])
        sims.append(g)
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    o.init_omega_max(meta["n_gas"])
    for s in range(4):
        st = [g.step(s) for g in sims]
        so = o.step(s)
        assert st[0]["accepted"] == st[1]["accepted"] == so["accepted"]
    for a, b, c in zip(sims[0].download(), sims[1].download(), o.download()):
        np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(b, c)
    for k in ("count", "start", "n"):
        np.testing.assert_array_equal(sims[0].cells()[k], o.cells()[k])
    runs, redone = _runs(sims[0])
    assert runs >= 4 and redone == 0, (runs, redone)
    assert _runs(sims[1])[0] == 0


def test_a_particle_leaving_its_window_redoes_the_sort():
    p, arrs, meta = W.heat_conduction(precision=mcpm.FP64, seed=702)
    y, u1, u2, u3 = [a.copy() for a in arrs]
    k = int(np.argmin(np.abs(y - 10.0)))
    u2[k] = 60.0  # from mid-slab, 60 x dt = 2.6 sigma = 26 cells in one step
    g, o = mcpm.Simulation(p), orc.OracleSim(p)
    for s in (g, o):
        s.upload(y, u1, u2, u3)
        s.cell_update()
    g.stream_bc(0, with_kick=False)
    o.stream_bc(0, with_kick=False)
    g.cell_update()
    o.cell_update()
    runs, redone = _runs(g)
    assert redone >= 1
    for a, b in zip(g.download(), o.download()):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(g.cells()["count"], o.cells()["count"])


def test_windowed_sort_on_batched_replicas():
    common, reps, arrs, metas = W.replicas(n_rep=8, seed0=30, N_lo=20000, N_hi=80000,
                                           cells_lo=128, cells_hi=400, length=60.0)
    bs = []
    for window in (1, 0):
        b = mcpm.Simulation.batch(common, reps)
        b.set_tuning(mcpm.TUNE_SORT_WINDOW, window)
        counts = [a[0].size for a in arrs]
        b.upload_batch(counts, *[np.concatenate([a[q] for a in arrs]) for q in range(4)])
        b.cell_update()
        b.init_omega_max_batch([m["n_gas"] for m in metas])
        bs.append(b)
    for s in range(3):
        assert bs[0].step(s)["accepted"] == bs[1].step(s)["accepted"]
    for a, c in zip(bs[0].download(), bs[1].download()):
        np.testing.assert_array_equal(a, c)
    for k in ("count", "start", "n", "F2"):
        np.testing.assert_array_equal(bs[0].cells()[k], bs[1].cells()[k])
    runs, redone = _runs(bs[0])
    assert runs >= 3 and redone == 0, (runs, redone)


def test_windowed_sort_on_more_than_65536_cells():
    """192 replicas, about 73k cells: the per-cell scan of the windowed sort
    then takes several consecutive cells per wavefront (each cell's chunk
    bounds searched from the previous cell's, across replica boundaries too);
    the same order and cells as the full-histogram sort, bit for bit."""
    common, reps, arrs, metas = W.replicas(n_rep=192, seed0=40, N_lo=7000, N_hi=9000,
                                           cells_lo=360, cells_hi=400, length=60.0)
    assert sum(r["n_cells"] for r in reps) > 65536
    counts = [a[0].size for a in arrs]
    print(f"windowed sort R=192 N={sum(counts)} cells={sum(r['n_cells'] for r in reps)}",
          flush=True)
    bs = []
    for window in (1, 0):
        b = mcpm.Simulation.batch(common, reps)
        b.set_tuning(mcpm.TUNE_SORT_WINDOW, window)
        b.upload_batch(counts, *[np.concatenate([a[q] for a in arrs]) for q in range(4)])
        b.cell_update()
        b.init_omega_max_batch([m["n_gas"] for m in metas])
        bs.append(b)
    for s in range(3):
        assert bs[0].step(s)["accepted"] == bs[1].step(s)["accepted"]
    for a, c in zip(bs[0].download(), bs[1].download()):
        np.testing.assert_array_equal(a, c)
    for k in ("count", "start", "n", "F2"):
        np.testing.assert_array_equal(bs[0].cells()[k], bs[1].cells()[k])
    runs, redone = _runs(bs[0])
    assert runs >= 3 and redone == 0, (runs, redone)
