"""GPU parity on the shapes of the five benchmark configurations.

* Heat-conduction slab (200 cells x 100 particles, FP64) with both contact
  values Y(n): the reference's polynomial and Carnahan-Starling
  (Main_Functions.cpp:1906-1915); whole steps bit for bit, and the
  omega_max initialisation (main.cpp:250-269) computed on each side.
* Dense slab (4000 cells, partner cells up to 200 cells away): the
  configuration's own check, 1000 particles per cell (4M), pairs and counts
  bit for bit for two steps; and at 15 particles per cell, ESMC and FP_dense
  over several steps.
* Interface slab: covered at full size in test_gpu_parity.
* 256 batched replicas (about 53M particles): every replica bit for bit
  its own unbatched run for two steps; and two small replicas of a batch
  against the oracle.
* 512 x 512 droplet plane: at three particles per cell (about 1M) the sort,
  the deposit and ESMC steps against the oracle bit for bit; at 64M
  particles the properties that do not depend on size (every particle in
  exactly one cell, cells in key order, deposit total, momentum and energy
  of a collision step).
"""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _slab_pair(p, arrs):
    g, o = mcpm.Simulation(p), orc.OracleSim(p)
    for s in (g, o):
        s.upload(*arrs)
        s.cell_update()
    return g, o


def _same_slab(g, o, tag=""):
    cg, co = g.cells(), o.cells()
    for k in ("count", "start", "n"):
        np.testing.assert_array_equal(cg[k], co[k], err_msg=f"{k} {tag}")
    for k, (a, b) in enumerate(zip(g.download(), o.download())):
        np.testing.assert_array_equal(a, b, err_msg=f"component {k} {tag}")


def _same_log(g, o, so):
    cap = so["valid"] + 16
    pg, ig, jg, ag = g.collision_log(cap)
    po, io, jo, ao = o.collision_log(cap)
    order = np.argsort(pg, kind="stable")
    np.testing.assert_array_equal(pg[order], po)
    np.testing.assert_array_equal(ig[order], io)
    np.testing.assert_array_equal(jg[order], jo)
    np.testing.assert_array_equal(ag[order], ao)


@pytest.mark.parametrize("y_model", [mcpm.Y_REFERENCE, mcpm.Y_CARNAHAN_STARLING])
@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_heat_conduction_steps_and_omega_init(y_model, precision):
    p, arrs, meta = W.heat_conduction(precision=precision, seed=201)
    p.y_model = y_model
    _say(f"heat y_model={y_model} prec={precision}")
    g, o = _slab_pair(p, arrs)
    g.init_omega_max(meta["n_gas"])
    o.init_omega_max(meta["n_gas"])
    om = o.cells()["omega_max"]
    np.testing.assert_array_equal(g.cells()["omega_max"], om)
    # omega_max = 4 pi sigma^2 10 sqrt(kT/m) n dt Y(n_gas) around the mean
    y = orc.contact_y(p, meta["n_gas"])
    ref = 4.0 * np.pi * 10.0 * np.sqrt(meta["T0"]) * meta["n_gas"] * p.delta_t * y
    assert abs(np.median(om) / ref - 1.0) < 0.05
    g.debug_collision_log(True)
    for s in range(5):
        sg, so = g.step(s), o.step(s)
        for k in ("candidates", "valid", "accepted"):
            assert sg[k] == so[k], (s, k)
        _same_slab(g, o, f"step {s}")
    np.testing.assert_array_equal(g.cells()["omega_max"], o.cells()["omega_max"])
    np.testing.assert_array_equal(g.moments(), o.moments())


def test_carnahan_starling_changes_the_collision_count():
    """The Y(n) switch reaches the device: at n sigma^3 = 0.2 the two forms
    differ by 5e-6 relative, at 0.45 by 1.5 %; the candidate counts follow."""
    got = {}
    for ym in (mcpm.Y_REFERENCE, mcpm.Y_CARNAHAN_STARLING):
        p, arrs, meta = W.heat_conduction(precision=mcpm.FP64, seed=202, n_density=0.45,
                                          n_cells=100, ppc=200)
        p.y_model = ym
        g = mcpm.Simulation(p)
        g.upload(*arrs)
        g.cell_update()
        g.init_omega_max(meta["n_gas"])
        got[ym] = g.collide_esmc(0)["candidates"]
    yr = W.contact_y_ref(0.45)
    eta = 0.45 * 2.0 * np.pi / 3.0 / 4.0
    ycs = (1.0 - eta / 2.0) / (1.0 - eta) ** 3
    ratio = got[mcpm.Y_CARNAHAN_STARLING] / got[mcpm.Y_REFERENCE]
    # omega_max carries Y(n_gas), and the candidate count 0.5 omega_max N_c
    assert ycs / yr > 1.015
    assert abs(ratio / (ycs / yr) - 1.0) < 0.003


def test_dense_slab_4M_two_steps_bit_exact():
    p, arrs, meta = W.dense_slab(n_density=0.3, ppc=1000, seed=203)
    _say(f"dense slab nc={p.n_cells} N={arrs[0].size}")
    g, o = _slab_pair(p, arrs)
    _same_slab(g, o, "sort")
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    g.debug_collision_log(True)
    for s in range(2):
        sg, so = g.step(s), o.step(s)
        for k in ("candidates", "valid", "accepted"):
            assert sg[k] == so[k], (s, k)
        _same_log(g, o, so)
        _same_slab(g, o, f"step {s}")


@pytest.mark.parametrize("model", [mcpm.MODEL_ESMC, mcpm.MODEL_FP_DENSE])
@pytest.mark.parametrize("n_density", [0.1, 0.3])
def test_dense_slab_geometry_few_particles_per_cell(model, n_density):
    p, arrs, meta = W.dense_slab(n_density=n_density, ppc=15, seed=204)
    p.model = model
    _say(f"dense slab ppc=15 model={model} n={n_density}")
    g, o = _slab_pair(p, arrs)
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    for s in range(4):
        sg, so = g.step(s), o.step(s)
        for k in ("candidates", "valid", "accepted"):
            assert sg[k] == so[k], (s, k)
        _same_slab(g, o, f"step {s}")


def _batch(common, reps, arrs, metas):
    b = mcpm.Simulation.batch(common, reps)
    counts = [a[0].size for a in arrs]
    b.upload_batch(counts, *[np.concatenate([a[q] for a in arrs]) for q in range(4)])
    b.cell_update()
    b.init_omega_max_batch([m["n_gas"] for m in metas])
    return b, counts


def test_replicas_256_each_equal_to_its_unbatched_run():
    common, reps, arrs, metas = W.replicas(n_rep=256, seed0=0)
    total = sum(a[0].size for a in arrs)
    _say(f"replicas R=256 N={total}")
    assert 4.0e7 < total < 6.5e7
    b, counts = _batch(common, reps, arrs, metas)
    steps = 2
    acc_b = [b.step(s)["accepted"] for s in range(steps)]
    got = b.download()
    cb = b.cells()
    b.close()
    off = np.cumsum([0] + counts)
    coff = np.cumsum([0] + [d["n_cells"] for d in reps])
    acc_s = np.zeros(steps, dtype=np.int64)
    for r, (d, a, m) in enumerate(zip(reps, arrs, metas)):
        s = mcpm.Simulation(d["params"])
        s.upload(*a)
        s.cell_update()
        s.init_omega_max(m["n_gas"])
        for t in range(steps):
            acc_s[t] += s.step(t)["accepted"]
        for x, y in zip(got, s.download()):
            np.testing.assert_array_equal(x[off[r]:off[r + 1]], y, err_msg=f"replica {r}")
        cs = s.cells()
        for k in ("count", "omega_max", "n", "F2"):
            np.testing.assert_array_equal(cb[k][coff[r]:coff[r + 1]], cs[k],
                                          err_msg=f"replica {r} {k}")
        s.close()
    assert list(acc_s) == acc_b


@pytest.mark.parametrize("precision", [mcpm.FP32, mcpm.FP64])
def test_small_replicas_of_a_batch_match_the_oracle(precision):
    common, reps, arrs, metas = W.replicas(n_rep=2, seed0=7, N_lo=5000, N_hi=20000,
                                           cells_lo=100, cells_hi=300, length=40.0,
                                           precision=precision)
    b, counts = _batch(common, reps, arrs, metas)
    oracles = []
    for d, a, m in zip(reps, arrs, metas):
        o = orc.OracleSim(d["params"])
        o.upload(*a)
        o.cell_update()
        o.init_omega_max(m["n_gas"])
        oracles.append(o)
    off = np.cumsum([0] + counts)
    coff = np.cumsum([0] + [d["n_cells"] for d in reps])
    for s in range(4):
        sb = b.step(s)
        assert sb["accepted"] == sum(o.step(s)["accepted"] for o in oracles)
        got = b.download()
        cb = b.cells()
        for r, o in enumerate(oracles):
            for x, y in zip(got, o.download()):
                np.testing.assert_array_equal(x[off[r]:off[r + 1]], y)
            co = o.cells()
            for k in ("count", "omega_max", "F2"):
                np.testing.assert_array_equal(cb[k][coff[r]:coff[r + 1]], co[k])


def test_droplet_plane_512_sort_deposit_esmc_vs_oracle():
    p, arrs, meta = W.droplets(N=800_000, long_range=mcpm.LR_NONE, precision=mcpm.FP32, seed=205)
    _say(f"droplets 512x512 N={arrs[0].size}")
    g, o = mcpm.Simulation(p), orc.OracleSim(p)
    for s in (g, o):
        s.upload_xy(*arrs)
        s.cell_update()
    for k in ("count", "start", "n"):
        np.testing.assert_array_equal(g.cells()[k], o.cells()[k])
    np.testing.assert_array_equal(g.nodes()[0], o.nodes()[0])
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    for s in range(2):
        sg, so = g.step(s), o.step(s)
        for k in ("candidates", "valid", "accepted"):
            assert sg[k] == so[k], (s, k)
        for k in ("count", "start"):
            np.testing.assert_array_equal(g.cells()[k], o.cells()[k])
        for a, b in zip(g.download_xy(), o.download_xy()):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(g.nodes()[0], o.nodes()[0])


def test_droplet_plane_64M_properties():
    p, arrs, meta = W.droplets(precision=mcpm.FP32, seed=206)
    N = arrs[0].size
    _say(f"droplets 512x512 N={N}")
    assert 6.0e7 < N < 7.0e7
    g = mcpm.Simulation(p)
    g.upload_xy(*arrs)
    del arrs
    g.cell_update()
    c = g.cells()
    assert int(c["count"].sum()) == N
    np.testing.assert_array_equal(c["start"][1:], np.cumsum(c["count"])[:-1].astype(c["start"].dtype))
    x, y, u1, u2, u3 = g.download_xy()
    kx = np.minimum(np.floor(x / p.dx), p.n_cells_x - 1).astype(np.int64)
    ky = np.minimum(np.floor(y / p.dy), p.n_cells - 1).astype(np.int64)
    assert np.all(np.diff(ky * p.n_cells_x + kx) >= 0)
    # each particle credits Fn / V to one node copy in each of 4 cells; on
    # a periodic plane the node values sum to 4 N Fn / V over the copies,
    # i.e. N Fn / V over the unique nodes
    nd, _ = g.nodes()
    uniq = nd[:-1, :-1]
    assert abs(uniq.sum() * p.cell_volume / (N * p.Fn) - 1.0) < 1e-9
    g.init_omega_max(meta["n_gas"])
    P0 = np.array([u1.sum(dtype=np.float64), u2.sum(dtype=np.float64), u3.sum(dtype=np.float64)])
    E0 = float((u1.astype(np.float64) ** 2 + u2.astype(np.float64) ** 2 + u3.astype(np.float64) ** 2).sum())
    st = g.collide_esmc(0)
    assert st["accepted"] > 0.01 * st["valid"]
    x2, y2, u1, u2, u3 = g.download_xy()
    np.testing.assert_array_equal(x2, x)
    np.testing.assert_array_equal(y2, y)
    P1 = np.array([u1.sum(dtype=np.float64), u2.sum(dtype=np.float64), u3.sum(dtype=np.float64)])
    E1 = float((u1.astype(np.float64) ** 2 + u2.astype(np.float64) ** 2 + u3.astype(np.float64) ** 2).sum())
    # FP32 storage rounds each of the 6 components an accepted pair writes
    assert np.abs(P1 - P0).max() < 1e-6 * np.sqrt(st["accepted"]) * 10
    assert abs(E1 - E0) < 1e-6 * E0
