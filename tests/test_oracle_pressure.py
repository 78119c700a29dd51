"""The oracle's pressure observables: measure_pressure's face sums p_m
(Main_Functions.cpp:13-34) and PostProc's ptot / p_rep columns
(myIO.cpp:855, 890).  CPU only.

What pins what:
* p_m, kinetic closed form: in free flight through a uniform periodic slab,
  the particles crossing a face in dt carry sum m |v_y| = n_sim A dt m <v_y^2>,
  so after S steps p_m / (S dt A) = n kB T (A the cross-section, Fn n_sim = n)
  -- a factor of two, a velocity other than |U2| or a missing face fails;
* p_m, one step between specular walls against a numpy count of the cells
  between each particle's old and new cell (the reference's index range, not
  the path), with the post-boundary |U2|;
* p_rep / ptot: the ideal-gas limit n kB T at low density, the Enskog factor
  1 + n b Y(n) and the mean-field term -2 b n^2 phi separately.
"""
import math

import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W


def _gas(n, T, N, n_cells, length, dt, boundary, seed, model=mcpm.MODEL_NONE):
    p, arrs, meta = W.interface(N=N, n_cells=n_cells, length=length, n_liq=n, n_vap=n, T=T,
                                dt=dt, precision=mcpm.FP64, seed=seed)
    p.long_range = mcpm.LR_NONE
    p.model = model
    p.boundary = boundary
    if boundary == mcpm.BC_SPECULAR:
        p.wall_lo, p.wall_hi = 1e-12, length - 1e-12
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    return p, o, meta


def test_face_flux_is_n_kT_in_free_flight():
    n, T, dt, S = 0.1, 1.3, 0.1, 20
    p, o, meta = _gas(n, T, 200_000, 20, 20.0, dt, mcpm.BC_PERIODIC, seed=61)
    o.measure_pressure(True)
    for s in range(S):
        o.stream_bc(s, False)
    pm = o.pressure()["p_m"]
    area = p.cell_volume / p.dy
    est = pm / (S * dt * area)
    nkT = meta["n_gas"] * p.kb * T
    assert est.mean() == pytest.approx(nkT, rel=0.02)
    assert np.abs(est / nkT - 1.0).max() < 0.12
    # off: no further sums; reset clears them
    o.measure_pressure(False)
    o.stream_bc(S, False)
    np.testing.assert_array_equal(o.pressure()["p_m"], pm)
    o.reset_samples()
    assert not o.pressure()["p_m"].any()


def test_face_sums_between_walls_match_numpy_count():
    p, o, meta = _gas(0.1, 1.0, 5_000, 10, 10.0, 0.7, mcpm.BC_SPECULAR, seed=62)
    y0, a0, b0, c0 = o.download()
    o.measure_pressure(True)
    o.stream_bc(0, False)
    y1, a1, b1, c1 = o.download()
    # specular walls flip u2 only: (u1, u3) identifies each particle
    k0 = np.lexsort((c0, a0))
    k1 = np.lexsort((c1, a1))
    assert np.array_equal(a0[k0], a1[k1]) and np.array_equal(c0[k0], c1[k1])
    yo, yn, u2 = y0[k0], y1[k1], b1[k1]
    assert (np.abs(u2) == np.abs(b0[k0])).all()
    expect = np.zeros(p.n_cells)
    no = np.floor(yo / p.dy).astype(int)
    nn = np.floor(yn / p.dy).astype(int)
    for a, b, v in zip(no, nn, u2):
        for c in range(min(a, b), max(a, b)):
            expect[c] += p.Fn * p.mass * abs(v)
    pm = o.pressure()["p_m"]
    assert expect.sum() > 0
    np.testing.assert_allclose(pm, expect, rtol=1e-12, atol=0)
    # reflected particles count the cells between old and new positions,
    # not the faces of their path: the wall cells see fewer sums
    assert pm[0] < pm[4] and pm[-2] < pm[4]


def test_ptot_and_prep_closed_forms():
    n, T = 0.3, 1.2
    p, o, meta = _gas(n, T, 100_000, 10, 10.0, 0.01, mcpm.BC_PERIODIC, seed=63)
    p_phi = p.phi
    pr = o.pressure()
    m = o.moments()
    nd = o.cells()["n"]
    Tc = m[:, 29]
    Y = np.array([orc.contact_y(p, x) for x in nd])
    np.testing.assert_allclose(pr["p_rep"], nd * p.kb * Tc * (1.0 + nd * p.b * Y), rtol=1e-14)
    np.testing.assert_allclose(pr["p_tot"] - pr["p_rep"], -2.0 * p.b * nd * nd * p_phi,
                               rtol=1e-12)
    # the Enskog factor at n sigma^3 = 0.3 is about 1.95; the cell kinetic
    # temperature is T within sampling noise
    assert (pr["p_rep"] / (nd * p.kb * T)).mean() == pytest.approx(
        1.0 + n * p.b * orc.contact_y(p, n), rel=0.01)
    # dilute: the ideal-gas pressure
    p2, o2, meta2 = _gas(1e-4, T, 100_000, 10, 10.0, 0.01, mcpm.BC_PERIODIC, seed=64)
    pr2 = o2.pressure()
    assert (pr2["p_rep"] / (o2.cells()["n"] * T)).mean() == pytest.approx(1.0, rel=0.01)


def test_pressure_is_slab_only():
    p, arrs, _ = W.plane(nx=6, ny=5, Lx=6.0, Ly=5.0, ppc=10, drops=(), seed=65,
                         long_range=mcpm.LR_NONE)
    o = orc.OracleSim(p)
    with pytest.raises(mcpm.McpmError):
        o.measure_pressure(True)
