"""Variable particle counts on the device: mcpm_addremove
(inverted_addremove_particles, Main_Functions.cpp:1206-1450) against the
oracle, bit for bit.

* One call on random particle weights: cells over 256 particles (several
  tiles of the compaction's ranks), splits, removals and untouched cells,
  FP64 and FP32.
* Five steps of a diffuse-wall slab whose middle third has cells of weight
  R: particles crossing into the light cells split, those crossing into the
  heavy ones are thinned; each step, then the add/remove, on both sides.
* The errors: no cell weights, identity tracking, capacity (the state is
  left as it was).
"""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu

R = 3


def _say(msg):
    print(msg, flush=True)


def _pair(p, arrs):
    g, o = mcpm.Simulation(p), orc.OracleSim(p)
    for s in (g, o):
        s.upload(*arrs)
        s.cell_update()
    return g, o


def _same(g, o, tag):
    cg, co = g.cells(), o.cells()
    for k in ("count", "start", "n"):
        np.testing.assert_array_equal(cg[k], co[k], err_msg=f"{k} {tag}")
    for k, (a, b) in enumerate(zip(g.download(), o.download())):
        np.testing.assert_array_equal(a, b, err_msg=f"component {k} {tag}")
    np.testing.assert_array_equal(g.particle_ratio(), o.particle_ratio(), err_msg=tag)


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_one_call_on_random_weights(precision):
    _say(f"addremove random weights prec={precision}")
    p, arrs, meta = W.heat_conduction(n_cells=24, ppc=700, seed=301, precision=precision)
    N = len(arrs[0])
    p.capacity = R * N
    g, o = _pair(p, arrs)
    rng = np.random.default_rng(7)
    cr = rng.choice([1, R], size=p.n_cells).astype(np.int32)
    cr[5] = 1  # a light cell with only heavy particles below
    for s in (g, o):
        s.set_cell_ratio(cr)
    _same(g, o, "weights from cells")
    w = rng.choice([1, R], size=N).astype(np.int32)
    st = o.cells()
    w[st["start"][5]:st["start"][5] + st["count"][5]] = R
    for s in (g, o):
        s.set_particle_ratio(w)
    sg, so = g.addremove(R, 11), o.addremove(R, 11)
    assert sg == so
    assert so["split"] > 0 and so["deleted"] > 0
    _same(g, o, "after")


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_inverted_zone_steps(precision):
    _say(f"addremove steps prec={precision}")
    p, arrs, meta = W.heat_conduction(n_cells=60, ppc=80, seed=302, precision=precision)
    p.capacity = 3 * len(arrs[0])
    g, o = _pair(p, arrs)
    cr = np.ones(p.n_cells, np.int32)
    cr[p.n_cells // 3: 2 * p.n_cells // 3] = R
    for s in (g, o):
        s.set_cell_ratio(cr)
        s.init_omega_max(meta["n_gas"])
    _same(g, o, "start")
    moved = 0
    for step in range(5):
        sg, so = g.step(step), o.step(step)
        for k in ("candidates", "valid", "accepted"):
            assert sg[k] == so[k], (step, k)
        ag, ao = g.addremove(R, step), o.addremove(R, step)
        assert ag == ao, step
        moved += ao["split"] + ao["deleted"]
        _same(g, o, f"step {step}")
    assert moved > 0


def test_errors_leave_the_state():
    _say("addremove errors")
    p, arrs, meta = W.heat_conduction(n_cells=20, ppc=50, seed=303)
    N = len(arrs[0])
    g = mcpm.Simulation(p)
    g.upload(*arrs)
    g.cell_update()
    with pytest.raises(mcpm.McpmError):
        g.addremove(R, 0)  # no cell weights
    g.set_cell_ratio(np.ones(p.n_cells, np.int32))
    g.set_particle_ratio(np.full(N, R, np.int32))  # every particle would split: N R > capacity
    before = g.download()
    with pytest.raises(mcpm.McpmError) as e:
        g.addremove(R, 0)
    assert e.value.code == 3  # MCPM_ERR_CAPACITY
    for a, b in zip(before, g.download()):
        np.testing.assert_array_equal(a, b)
    g.track_ids(True)
    with pytest.raises(mcpm.McpmError):
        g.addremove(R, 0)
