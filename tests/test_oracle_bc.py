"""What pins the oracle's evaporation boundary (BC_evaporation,
Main_Functions.cpp:5444-5617) and the inverted problem's thermostat
(Main_Functions.cpp:1760-1830): closed forms of free flight and of the
rescale, and the distribution of the re-inserted particles."""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W


def _reinject(precision=mcpm.FP64, n_cells=50, ppc=200, T=1.3, dt_frac=4.0, seed=3):
    p, arrs, meta = W.heat_conduction(n_cells=n_cells, ppc=ppc, length=10.0, precision=precision,
                                      dt_frac=dt_frac, seed=seed)
    p.boundary = mcpm.BC_REINJECT
    p.model = mcpm.MODEL_NONE
    p.T_wall_lo = T
    return p, arrs


def _split(before, after, dt, p, f32):
    """The step re-sorts the particles: match the expected free flights of
    the particles that stay inside against the particles after the step (as
    rows of (y, u1, u2, u3)); the rest of `after` are the re-inserted ones."""
    y0, v1, v2, v3 = before
    flight = y0 + v2 * dt
    out = (flight < p.wall_lo) | (flight > p.wall_hi)
    if f32:
        flight = flight.astype(np.float32).astype(np.float64)
    exp = np.stack([flight, v1, v2, v3], axis=1)[~out]
    got = np.stack(after, axis=1)
    e = set(map(tuple, exp))
    inside = np.array([tuple(r) in e for r in got])
    return out, exp, got, inside


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_reinjection_is_free_flight_for_those_inside(precision):
    """One free-flight step: a particle whose flight stays in [eps, L - eps]
    ends at y + v dt; exactly the others are re-inserted, into [L/4, 3L/4]."""
    p, (y, u1, u2, u3) = _reinject(precision)
    o = orc.OracleSim(p)
    o.upload(y, u1, u2, u3)
    o.cell_update()
    before = o.download()
    o.stream_bc(0, False)
    after = o.download()
    out, exp, got, inside = _split(before, after, p.delta_t, p, precision == mcpm.FP32)
    assert out.sum() == o.escaped() > 100
    assert inside.sum() == len(exp)
    rein = got[~inside]
    assert len(rein) == out.sum()
    assert np.all(rein[:, 0] >= 0.25 * p.length) and np.all(rein[:, 0] <= 0.75 * p.length)


def test_reinserted_particles_are_maxwellian_and_uniform():
    p, (y, u1, u2, u3) = _reinject(ppc=2000, dt_frac=20.0, T=1.7)
    o = orc.OracleSim(p)
    o.upload(y, u1, u2, u3)
    o.cell_update()
    before = o.download()
    o.stream_bc(0, False)
    out, exp, got, inside = _split(before, o.download(), p.delta_t, p, False)
    rein = got[~inside]
    n = len(rein)
    assert n > 20000
    for k in (1, 2, 3):
        v = rein[:, k]
        assert abs(v.mean()) < 4.0 * np.sqrt(1.7 / n)
        assert abs(v.var() / 1.7 - 1.0) < 4.0 * np.sqrt(2.0 / n)
    # positions uniform on [L/4, 3L/4]: mean L/2, variance (L/2)^2 / 12
    L = p.length
    q = rein[:, 0]
    assert abs(q.mean() - 0.5 * L) < 4.0 * (0.5 * L) / np.sqrt(12.0 * n)
    assert abs(q.var() / ((0.5 * L) ** 2 / 12.0) - 1.0) < 0.05


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_thermostat_sets_zone_temperatures_and_keeps_means(precision):
    p, (y, u1, u2, u3), _ = W.heat_conduction(n_cells=40, ppc=100, length=10.0, precision=precision)
    u1 = u1 + 0.3 * np.sin(y)  # a velocity profile, so that cell means are not zero
    o = orc.OracleSim(p)
    o.upload(y, u1, u2, u3)
    o.cell_update()
    before = o.moments().reshape(-1, mcpm.NMOM)
    r0 = o.download()
    zone_lo, zone_hi, Tc, Th = 2.0, 1.5, 0.6, 2.4
    o.thermostat(zone_lo, Tc, zone_hi, Th)
    after = o.moments().reshape(-1, mcpm.NMOM)
    r1 = o.download()
    centre = (np.arange(p.n_cells) + 0.5) * p.length / p.n_cells
    lo, hi = centre < zone_lo, centre > p.length - zone_hi
    tol = 1e-12 if precision == mcpm.FP64 else 2e-6
    assert np.allclose(after[lo, 29], Tc, rtol=tol * 10)
    assert np.allclose(after[hi, 29], Th, rtol=tol * 10)
    assert np.allclose(after[:, :3], before[:, :3], atol=tol * 10)
    mid = ~(lo | hi)
    np.testing.assert_array_equal(after[mid], before[mid])
    # particles of untouched cells keep their velocities bit for bit
    cells = np.minimum((r0[0] / (p.length / p.n_cells)).astype(int), p.n_cells - 1)
    keep = mid[cells]
    for k in (1, 2, 3):
        np.testing.assert_array_equal(r1[k][keep], r0[k][keep])
