"""Graph-replayed steps (mcpm_graph_capture / mcpm_graph_steps) against the
same steps called one by one (mcpm_step): positions, velocities, cell counts,
densities, omega_max and the field bit for bit, and the same collision
counts.  Covered: the interface slab in FP32 and FP64 with eager steps
between replays, batched replicas, a capacity too small for a step (the run
recaptures and repeats the step), and particles that leave their sort
windows (the host finishes that step's sort and the run goes on), also on
more than 16384 cells."""
import numpy as np
import pytest

import mcpm
import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pair(p, arrs, n_gas):
    sims = []
    for _ in range(2):
        s = mcpm.Simulation(p)
        s.upload(*arrs)
        s.cell_update()
        s.init_omega_max(n_gas)
        sims.append(s)
    return sims


def _same_state(a, b):
    for k, (x, y) in enumerate(zip(a.download(), b.download())):
        np.testing.assert_array_equal(x, y, err_msg=f"particle array {k}")
    ca, cb = a.cells(), b.cells()
    for k in ("count", "start", "n", "omega_max", "F2"):
        np.testing.assert_array_equal(ca[k], cb[k], err_msg=k)


def _sum(stats, keys=("candidates", "valid", "accepted")):
    return {k: sum(s[k] for s in stats) for k in keys}


@pytest.mark.parametrize("precision,eager_unroll,graph_unroll",
                         [(mcpm.FP32, 4, 2), (mcpm.FP64, 2, 4)])
def test_graph_steps_equal_eager_steps(precision, eager_unroll, graph_unroll):
    """Also across the operations per thread of a round (2 or 4)."""
    p, arrs, meta = W.interface(N=300_000, n_cells=512, length=100.0, precision=precision,
                                seed=71)
    _say(f"graph interface N=300000 nc=512 prec={precision} unroll {eager_unroll}/{graph_unroll}")
    e, g = _pair(p, arrs, meta["n_gas"])
    e.set_tuning(mcpm.TUNE_ROUND_UNROLL, eager_unroll)
    g.set_tuning(mcpm.TUNE_ROUND_UNROLL, graph_unroll)
    eager = [e.step(s) for s in range(9)]
    g.graph_capture()
    parts = [g.graph_steps(0, 3)]
    parts.append(g.step(3))          # an eager step between replays
    parts.append(g.graph_steps(4, 2))
    parts.append(g.graph_steps(6, 3))
    assert g.graph_steps(9, 0)["candidates"] == 0  # no step, no change
    assert _sum(parts) == _sum(eager)
    _same_state(e, g)


def test_graph_steps_of_batched_replicas():
    common, reps, arrs, metas = W.replicas(n_rep=6, N_lo=20_000, N_hi=60_000, cells_lo=128,
                                           cells_hi=300)
    counts = [a[0].size for a in arrs]
    _say(f"graph replicas R=6 N={sum(counts)}")
    sims = []
    for _ in range(2):
        b = mcpm.Simulation.batch(common, reps)
        b.upload_batch(counts, *[np.concatenate([a[q] for a in arrs]) for q in range(4)])
        b.cell_update()
        b.init_omega_max_batch([m["n_gas"] for m in metas])
        sims.append(b)
    e, g = sims
    eager = [e.step(s) for s in range(5)]
    g.graph_capture()
    got = g.graph_steps(0, 5)
    assert _sum([got]) == _sum(eager)
    _same_state(e, g)


def test_graph_capacity_too_small_recaptures():
    p, arrs, meta = W.interface(N=100_000, n_cells=256, length=100.0, seed=72)
    _say("graph capacity N=100000 nc=256")
    e, g = _pair(p, arrs, meta["n_gas"])
    eager = [e.step(s) for s in range(4)]
    g.graph_capture(op_capacity=4096)
    got = g.graph_steps(0, 4)
    assert _sum([got]) == _sum(eager)
    _same_state(e, g)


def test_graph_particles_leaving_their_windows():
    p, arrs, meta = W.interface(N=100_000, n_cells=256, length=100.0, seed=73)
    y, u1, u2, u3 = (a.copy() for a in arrs)
    # a few particles cross many cells per step (20 cells at dt = 0.02)
    u2[::5000] = 400.0
    _say("graph window overflow N=100000 nc=256")
    e, g = _pair(p, (y, u1, u2, u3), meta["n_gas"])
    eager = [e.step(s) for s in range(6)]
    g.graph_capture()
    got = g.graph_steps(0, 6)
    assert _sum([got]) == _sum(eager)
    _same_state(e, g)


def test_graph_capture_rejects_what_it_does_not_run():
    p, arrs, meta = W.heat_conduction(n_cells=50, ppc=40, seed=74)
    p.model = mcpm.MODEL_DSMC_VHS
    s = mcpm.Simulation(p)
    s.upload(*arrs)
    s.cell_update()
    with pytest.raises(mcpm.McpmError):
        s.graph_capture()
    q, a2, m2 = W.interface(N=20_000, n_cells=64, length=20.0, seed=75)
    t = mcpm.Simulation(q)
    t.upload(*a2)
    t.cell_update()
    with pytest.raises(mcpm.McpmError):
        t.graph_steps(0, 1)  # nothing captured


def test_graph_steps_at_the_interface_size():
    """20M particles, 4096 cells (the benchmark's interface): three steps one
    call at a time and three graph replays from the same state give the same
    cells (counts, starts, densities, omega_max, field) and the same FP64
    energy and momentum sums, bit for bit."""
    p, arrs, meta = W.interface()
    _say("graph interface N=20000000 nc=4096")
    sims = []
    for _ in range(2):
        s = mcpm.Simulation(p)
        s.upload(*arrs)
        s.cell_update()
        s.init_omega_max(meta["n_gas"])
        s.step(0)
        sims.append(s)
    del arrs
    e, g = sims
    eager = [e.step(s) for s in (1, 2, 3)]
    g.graph_capture()
    got = g.graph_steps(1, 3)
    assert _sum([got]) == _sum(eager)
    ce, cg = e.cells(), g.cells()
    for k in ("count", "start", "n", "omega_max", "F2"):
        np.testing.assert_array_equal(ce[k], cg[k], err_msg=k)
    np.testing.assert_array_equal(e.energy_momentum(), g.energy_momentum())


@pytest.mark.parametrize("precision", [mcpm.FP32, mcpm.FP64])
def test_graph_steps_between_diffuse_walls_without_field(precision):
    """The dense slab's geometry (diffuse walls at T = 1.0 and 1.5, 0.005
    sigma cells, no mean field) at 40 particles per cell: wall re-emission
    inside a captured step (the general push), the same steps bit for bit."""
    p, arrs, meta = W.dense_slab(n_density=0.3, ppc=40, seed=76)
    p.precision = precision
    _say(f"graph dense slab nc={p.n_cells} N={arrs[0].size} prec={precision}")
    e, g = _pair(p, arrs, meta["n_gas"])
    eager = [e.step(s) for s in range(6)]
    g.graph_capture()
    got = g.graph_steps(0, 6)
    assert _sum([got]) == _sum(eager)
    _same_state(e, g)


def test_graph_steps_of_fp64_replicas_without_field():
    common, reps, arrs, metas = W.replicas(n_rep=4, N_lo=20_000, N_hi=40_000, cells_lo=128,
                                           cells_hi=200, precision=mcpm.FP64,
                                           long_range=mcpm.LR_NONE)
    counts = [a[0].size for a in arrs]
    _say(f"graph replicas FP64 R=4 N={sum(counts)}")
    sims = []
    for _ in range(2):
        b = mcpm.Simulation.batch(common, reps)
        b.upload_batch(counts, *[np.concatenate([a[q] for a in arrs]) for q in range(4)])
        b.cell_update()
        b.init_omega_max_batch([m["n_gas"] for m in metas])
        sims.append(b)
    e, g = sims
    eager = [e.step(s) for s in range(4)]
    g.graph_capture()
    got = g.graph_steps(0, 4)
    assert _sum([got]) == _sum(eager)
    _same_state(e, g)


def test_graph_particles_leaving_their_windows_on_many_cells():
    """64 replicas, more than 16384 cells in all: the full-histogram sort that
    finishes a step whose particles left their windows is then the two-digit
    sort, which writes no sorted y array, and the next replayed step draws
    its candidates from the y array refilled after it."""
    common, reps, arrs, metas = W.replicas(n_rep=64, N_lo=6000, N_hi=10000, cells_lo=260,
                                           cells_hi=300)
    assert sum(r["n_cells"] for r in reps) > 16384
    arrs = [[a.copy() for a in t] for t in arrs]
    for t in arrs[::8]:
        t[2][::1000] = 1500.0  # about 20 cells per step at dt = 0.02
    counts = [a[0].size for a in arrs]
    _say(f"graph replicas window overflow R=64 N={sum(counts)}")
    sims = []
    for _ in range(2):
        b = mcpm.Simulation.batch(common, reps)
        b.upload_batch(counts, *[np.concatenate([a[q] for a in arrs]) for q in range(4)])
        b.cell_update()
        b.init_omega_max_batch([m["n_gas"] for m in metas])
        sims.append(b)
    e, g = sims
    eager = [e.step(s) for s in range(6)]
    g.graph_capture()
    redone = g.get_tuning(mcpm.TUNE_SORT_WINDOW_REDONE)
    got = g.graph_steps(0, 6)
    assert g.get_tuning(mcpm.TUNE_SORT_WINDOW_REDONE) > redone  # the case happened
    assert _sum([got]) == _sum(eager)
    _same_state(e, g)
