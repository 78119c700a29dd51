"""The oracle's inverted_addremove_particles (Main_Functions.cpp:1206-1450),
pinned by what the operation must conserve and by its published rules:

* splitting a particle of weight R in a cell of weight 1 into R particles
  keeps the total weight exactly, keeps every original particle as it was,
  puts the copies at their parent's position, and perturbs each copy's
  velocity by h z with z the Philox normals of its own draw block and h
  Silverman's rule 1.06 sig n^(-1/5) (n = n_rem + R, :1343), sig the weighted
  spread about the reference's mean (the last kept particle's velocity times
  its weight over the cell's weight, :1302-1313);
* removing a light particle from a heavy cell happens with probability
  1 - 1/R (:1262-1268): the deleted fraction and the expected total weight
  are checked statistically, the particles kept are exactly those whose
  draw is not below 1 - 1/R;
* a cell's density carries its weight, n = n_ratio Fn N / V (:2588).
"""
import math

import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

R = 4


def _normals(p, j, step):
    w = orc.philox((p, j, step & 0xFFFFFFFF, (mcpm_tag() << 24) ^ ((step >> 32) << 8) ^ 1),
                   (SEED & 0xFFFFFFFF, SEED >> 32))
    z = []
    for a, b in ((w[0], w[1]), (w[2], w[3])):
        u1 = (float(a) + 0.5) * 2.0 ** -32
        u2 = (float(b) + 0.5) * 2.0 ** -32
        rad = math.sqrt(-2.0 * orc.det_log(u1))
        s, c = orc.det_sincos2pi(u2)
        z += [rad * c, rad * s]
    return z[:3]


def _deletion_draw(p, step):
    w = orc.philox((p, 0, step & 0xFFFFFFFF, (mcpm_tag() << 24) ^ ((step >> 32) << 8) ^ 0),
                   (SEED & 0xFFFFFFFF, SEED >> 32))
    return (float(w[0]) + 0.5) * 2.0 ** -32


def mcpm_tag():
    return 11  # MCPM_TAG_ADDREM


SEED = 5


def _sorted_slab(n_cells=20, ppc=60, precision=mcpm.FP64, seed=SEED):
    p, arrs, meta = W.heat_conduction(n_cells=n_cells, ppc=ppc, seed=seed, precision=precision)
    p.seed = seed
    p.capacity = R * len(arrs[0])  # room for every particle to split
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    return p, o


def test_cell_weight_enters_the_density():
    p, o = _sorted_slab()
    n0 = o.cells()["n"].copy()
    ratio = np.where(np.arange(p.n_cells) % 3 == 0, 5, 1).astype(np.int32)
    o.set_cell_ratio(ratio)
    cnt = o.cells()["count"].astype(np.float64)
    assert np.array_equal(o.cells()["n"], ratio.astype(np.float64) * p.Fn * cnt / p.cell_volume)
    assert np.allclose(o.cells()["n"], n0 * ratio, rtol=1e-15, atol=0)
    assert np.array_equal(o.particle_ratio(), np.repeat(ratio, o.cells()["count"]))
    o.set_cell_ratio(None)
    assert np.array_equal(o.cells()["n"], n0)


def test_nothing_changes_when_weights_match():
    p, o = _sorted_slab()
    o.set_cell_ratio(np.full(p.n_cells, 1, np.int32))
    before = o.download()
    st = o.addremove(R, 3)
    assert st == {"before": len(before[0]), "split": 0, "created": 0, "deleted": 0,
                  "after": len(before[0])}
    for a, b in zip(before, o.download()):
        assert np.array_equal(a, b)


def test_split_keeps_weight_and_places_copies():
    p, o = _sorted_slab()
    o.set_cell_ratio(np.ones(p.n_cells, np.int32))
    N = len(o.download()[0])
    rng = np.random.default_rng(3)
    w = np.where(rng.random(N) < 0.2, R, 1).astype(np.int32)
    o.set_particle_ratio(w)
    cells0 = o.cells()
    y0, u10, u20, u30 = o.download()
    U0 = np.stack([u10, u20, u30], axis=1)
    step = 9
    st = o.addremove(R, step)
    nh = int((w == R).sum())
    assert st["split"] == nh and st["created"] == (R - 1) * nh and st["deleted"] == 0
    assert st["after"] == N + (R - 1) * nh == int(w.sum())  # total weight kept exactly
    cells1 = o.cells()
    y1, u11, u21, u31 = o.download()
    U1 = np.stack([u11, u21, u31], axis=1)
    assert np.all(o.particle_ratio() == 1)
    for c in range(p.n_cells):
        s0, n0 = cells0["start"][c], cells0["count"][c]
        s1, n1 = cells1["start"][c], cells1["count"][c]
        heavy = np.nonzero(w[s0:s0 + n0] == R)[0]
        assert n1 == n0 + (R - 1) * len(heavy)
        # the originals first, untouched and in order
        assert np.array_equal(y1[s1:s1 + n0], y0[s0:s0 + n0])
        assert np.array_equal(U1[s1:s1 + n0], U0[s0:s0 + n0])
        if len(heavy) == 0:
            continue
        # Silverman's rule on the reference's mean and weighted spread
        Wc = float(w[s0:s0 + n0].sum())
        last = s0 + n0 - 1
        m = U0[last] * w[last] / Wc
        sig = math.sqrt(float((((U0[s0:s0 + n0] - m) ** 2).sum(axis=1) * w[s0:s0 + n0]).sum()) / Wc)
        n_rem = n0 - len(heavy)
        h = 1.06 * sig * (n_rem + R) ** (-1.0 / 5.0)
        d = s1 + n0
        for q, u in enumerate(heavy):
            parent = s0 + u
            for j in range(R - 1):
                assert y1[d] == y0[parent]
                z = _normals(parent, j, step)
                assert np.allclose(U1[d] - U0[parent], h * np.array(z), rtol=1e-9, atol=1e-12), \
                    (c, q, j)
                d += 1


def test_removal_rate_and_expected_weight():
    p, o = _sorted_slab(n_cells=40, ppc=400)
    N = len(o.download()[0])
    o.set_cell_ratio(np.full(p.n_cells, R, np.int32))
    o.set_particle_ratio(np.ones(N, np.int32))
    y0, u10, _, _ = o.download()
    step = 2**33 + 17  # a step with high bits: they enter the draw
    st = o.addremove(R, step)
    pdel = 1.0 - 1.0 / R
    frac = st["deleted"] / N
    assert abs(frac - pdel) < 4.0 * math.sqrt(pdel * (1 - pdel) / N)
    # the weight survives in expectation: kept x R against N
    assert abs(st["after"] * R - N) < 4.0 * R * math.sqrt(pdel * (1 - pdel) * N)
    keep = np.array([_deletion_draw(k, step) >= pdel for k in range(N)])
    y1, u11, _, _ = o.download()
    assert np.array_equal(y1, y0[keep]) and np.array_equal(u11, u10[keep])
    assert np.all(o.particle_ratio() == R)


def test_fp32_copies_are_rounded_like_records():
    p, o = _sorted_slab(precision=mcpm.FP32)
    o.set_cell_ratio(np.ones(p.n_cells, np.int32))
    N = len(o.download()[0])
    o.set_particle_ratio(np.full(N, 2, np.int32))
    o.addremove(2, 1)
    _, u1, u2, u3 = o.download()
    for a in (u1, u2, u3):
        assert np.array_equal(a, a.astype(np.float32).astype(np.float64))


def test_state_and_capacity_errors():
    p, o = _sorted_slab()
    with pytest.raises(mcpm.McpmError):
        o.addremove(R, 0)  # no cell weights
    p2, arrs, _ = W.heat_conduction(n_cells=20, ppc=60, seed=SEED)
    N = len(arrs[0])
    p2.capacity = N + 10
    o2 = orc.OracleSim(p2)
    o2.upload(*arrs)
    o2.cell_update()
    o2.set_cell_ratio(np.ones(p2.n_cells, np.int32))
    o2.set_particle_ratio(np.full(N, R, np.int32))
    before = o2.download()
    with pytest.raises(mcpm.McpmError):
        o2.addremove(R, 0)
    for a, b in zip(before, o2.download()):
        assert np.array_equal(a, b)
