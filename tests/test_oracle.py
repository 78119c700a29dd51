"""The CPU oracle pinned against closed forms, invariants and independent
numpy computations (no GPU).  Each check is chosen so that a plausible slip
in the restatement (a dropped term, a wrong index, a swapped wall) fails.
"""
import math

import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W


def _sim(p, arrs):
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    return o


# ---------------------------------------------------------------- constants
def test_contact_value_known_answers():
    p = mcpm.Params(n_cells=1, length=1.0, capacity=1)
    # values at n sigma^3 = 0.2 computed from the published forms:
    # reference polynomial (Main_Functions.cpp:1911) and Carnahan-Starling
    assert orc.contact_y(p, 0.2) == pytest.approx(1.3205908, abs=1e-7)
    p.y_model = mcpm.Y_CARNAHAN_STARLING
    eta = math.pi * 0.2 / 6.0
    assert orc.contact_y(p, 0.2) == pytest.approx((1 - eta / 2) / (1 - eta) ** 3, rel=1e-14)
    assert orc.contact_y(p, 0.0) == 1.0


def test_mean_collision_time_known_answer():
    # Enskog frequency 4 n sigma^2 Y sqrt(pi kT/m) at n=0.2, T=1.5 -> tau_c = 0.436035
    assert W.mean_collision_time(0.2, 1.5) == pytest.approx(0.436035, abs=2e-6)


def test_gauss_legendre_table_is_the_8_decimal_rule():
    x, w = orc.gl50()
    xr, wr = np.polynomial.legendre.leggauss(50)
    np.testing.assert_array_equal(x, np.round(xr, 8))
    np.testing.assert_array_equal(w, np.round(wr, 8))
    # first entries as printed in the reference table (Main_Functions.cpp:3328, 3338)
    assert x[0] == -0.9988664 and x[1] == -0.99403197 and w[0] == 0.00290862


def test_thomas_matches_dense_solve():
    rng = np.random.default_rng(1)
    n = 257
    a = 2.5 + rng.random(n)
    b = -rng.random(n)
    c = -rng.random(n)
    rhs = rng.normal(size=n)
    A = np.diag(a) + np.diag(b[:-1], -1) + np.diag(c[:-1], 1)
    np.testing.assert_allclose(orc.thomas(a, b, c, rhs), np.linalg.solve(A, rhs), rtol=1e-12,
                               atol=1e-13)


# ---------------------------------------------------------------- binning
def test_cell_update_is_numpy_stable_argsort():
    p, arrs, _ = W.heat_conduction(seed=21)
    o = _sim(p, arrs)
    key = np.floor(arrs[0] / p.dy).astype(int)
    order = np.argsort(key, kind="stable")
    y, u1, u2, u3 = o.download()
    np.testing.assert_array_equal(y, arrs[0][order])
    np.testing.assert_array_equal(u3, arrs[3][order])
    c = o.cells()
    np.testing.assert_array_equal(c["count"], np.bincount(key, minlength=p.n_cells))
    np.testing.assert_allclose(c["n"], c["count"] * p.Fn / p.cell_volume, rtol=1e-15)


def test_out_of_domain_particle_is_an_error():
    p, arrs, _ = W.heat_conduction(n_cells=20, ppc=5)
    y = arrs[0].copy()
    y[3] = p.length + 1.0
    o = orc.OracleSim(p)
    o.upload(y, *arrs[1:])
    with pytest.raises(mcpm.McpmError) as e:
        o.cell_update()
    assert e.value.code == 4  # MCPM_ERR_DOMAIN


# ---------------------------------------------------------------- collisions
def test_collisions_conserve_momentum_and_energy():
    p, arrs, meta = W.heat_conduction(seed=22)
    o = _sim(p, arrs)
    o.init_omega_max(meta["n_gas"])
    _, u1, u2, u3 = o.download()
    P0 = np.array([u1.sum(), u2.sum(), u3.sum()])
    E0 = (u1 ** 2 + u2 ** 2 + u3 ** 2).sum()
    st = o.collide_esmc(0)
    _, u1, u2, u3 = o.download()
    assert st["accepted"] > 500
    assert np.abs(np.array([u1.sum(), u2.sum(), u3.sum()]) - P0).max() < 1e-10
    assert abs((u1 ** 2 + u2 ** 2 + u3 ** 2).sum() - E0) < 1e-11 * E0


def test_collision_rate_is_the_enskog_frequency():
    """Periodic uniform Maxwellian gas: accepted collisions per step equal
    N nu dt / 2 with nu = 4 n sigma^2 Y(n) sqrt(pi kT/m)."""
    n, T, N = 0.3, 1.0, 200_000
    p, arrs, meta = W.interface(N=N, n_cells=400, length=40.0, n_liq=n, n_vap=n, T=T, dt=0.01,
                                precision=mcpm.FP64, seed=23)
    p.long_range = mcpm.LR_NONE
    o = _sim(p, arrs)
    o.init_omega_max(meta["n_gas"])
    tot = 0
    steps = 4
    for s in range(steps):
        tot += o.collide_esmc(s)["accepted"]
    expect = steps * N * p.delta_t / W.mean_collision_time(n, T) / 2.0
    assert abs(tot - expect) < 4.0 * math.sqrt(expect) + 0.01 * expect, (tot, expect)


def test_collision_directions_are_isotropic():
    """Momentum transfer along the slab axis gets one third of the total
    kinetic-energy exchange in an isotropic gas."""
    p, arrs, meta = W.interface(N=100_000, n_cells=200, length=20.0, n_liq=0.3, n_vap=0.3,
                                T=1.0, dt=0.02, precision=mcpm.FP64, seed=24)
    p.long_range = mcpm.LR_NONE
    o = _sim(p, arrs)
    o.init_omega_max(meta["n_gas"])
    before = np.array(o.download()[1:])
    o.collide_esmc(0)
    after = np.array(o.download()[1:])
    d = np.abs(after ** 2 - before ** 2).sum(axis=1)
    frac = d / d.sum()
    np.testing.assert_allclose(frac, [1 / 3] * 3, atol=0.02)


def test_omega_max_initialisation_closed_form():
    """main.cpp:250-269 with dt = f tau_c gives
    omega_max = 10 f sqrt(pi) (n_c/n) sqrt(T_c/T) Y(n)/Y(n) in each cell."""
    p, arrs, meta = W.heat_conduction(seed=25)
    o = _sim(p, arrs)
    o.init_omega_max(meta["n_gas"])
    om = o.cells()["omega_max"]
    m = o.moments()
    nc = o.cells()["n"]
    expect = 10 * 0.1 * math.sqrt(math.pi) * (nc / 0.2) * np.sqrt(m[:, 29] / 1.5)
    np.testing.assert_allclose(om, expect, rtol=1e-12)


# ---------------------------------------------------------------- field
def _counts_state(p, counts, seed=0):
    rng = np.random.default_rng(seed)
    y = np.concatenate([(c + rng.random(k)) * p.dy for c, k in enumerate(counts)])
    n = len(y)
    return (y, rng.normal(size=n), rng.normal(size=n), rng.normal(size=n))


def test_periodic_field_matches_fourier_solution():
    """Cell counts n_c = n0 + a cos(2 pi k c / N): the periodic discrete
    Helmholtz problem is diagonal in Fourier space (numpy FFT solve)."""
    nc, L = 256, 64.0
    p = mcpm.Params(n_cells=nc, length=L, capacity=10 ** 6, boundary=mcpm.BC_PERIODIC,
                    long_range=mcpm.LR_SCREENED_POISSON, ast=-73.17, bst=2.3539, eqdist=1.0)
    cidx = np.arange(nc)
    counts = (200 + 80 * np.cos(2 * np.pi * 3 * cidx / nc) + 30 * np.sin(2 * np.pi * 7 * cidx / nc)).astype(int)
    p.capacity = int(counts.sum())
    o = _sim(p, _counts_state(p, counts))
    o.poisson_solver()
    F2 = o.cells()["F2"]
    n = counts * p.Fn / p.cell_volume
    nj = 0.5 * (n + np.roll(n, 1))
    h = p.dy
    theta = 2 * np.pi * np.fft.fftfreq(nc)
    x = np.real(np.fft.ifft(np.fft.fft(h * h * nj) / (2 + (h * p.bst) ** 2 - 2 * np.cos(theta))))
    core = (1 - (p.sigma * p.bst + 1) * math.exp(-p.bst * p.sigma)) / p.bst ** 2
    x = x - nj * core
    F = -(np.roll(x, -1) - x) / h / p.mass * p.ast
    np.testing.assert_allclose(F2, F, rtol=1e-10, atol=1e-12 * np.abs(F).max())


def test_uniform_density_gives_no_force_and_the_wall_closed_form():
    """Walls: for uniform n the interior face value is n / bst^2 (no force),
    and the Dirichlet value is the truncated Yukawa integral
    n (1 - (1 + 3 bst) e^{-3 bst}) / bst^2 (eqdist = sigma = 1)."""
    nc, L = 400, 40.0
    p = mcpm.Params(n_cells=nc, length=L, capacity=10 ** 6, boundary=mcpm.BC_DIFFUSE,
                    long_range=mcpm.LR_SCREENED_POISSON, ast=-292.68, bst=2.3539, eqdist=1.0)
    counts = np.full(nc, 100)
    o = _sim(p, _counts_state(p, counts))
    o.poisson_solver()
    F2 = o.cells()["F2"]
    # the wall perturbation decays like exp(-bst dy k): 1e-10 of it is left
    # 100 cells from the walls
    mid = slice(100, 300)
    assert np.abs(F2[mid]).max() < 1e-9 * np.abs(F2).max()
    # wall value: rebuild x at face 0 from F2 and the interior value
    n0 = 100 * p.Fn / p.cell_volume
    lam = p.bst
    wall = n0 * (1 - (1 + 3 * lam) * math.exp(-3 * lam)) / lam ** 2
    core = (1 - (lam + 1) * math.exp(-lam)) / lam ** 2
    # x_f - nj core summed: x_0 = x_mid + sum over cells of F2 dy m / ast
    x_mid = n0 / lam ** 2 - n0 * core
    x0 = x_mid + np.sum(F2[:200]) * p.dy * p.mass / p.ast
    assert x0 + n0 * core == pytest.approx(wall, rel=2e-7)


# ---------------------------------------------------------------- push / walls
def test_fp32_wall_reemission_stays_float():
    """Regression: every stored FP32 component is a float after wall hits."""
    p, arrs, meta = W.heat_conduction(precision=mcpm.FP32, seed=8)
    o = _sim(p, arrs)
    o.stream_bc(0, True)
    for a in o.download():
        assert np.all(a.astype(np.float32).astype(np.float64) == a)


def test_diffuse_walls_keep_an_isothermal_gas_at_wall_temperature():
    p, arrs, meta = W.heat_conduction(n_cells=50, ppc=400, length=5.0, T0=1.3, T_lo=1.3,
                                      T_hi=1.3, seed=26)
    o = _sim(p, arrs)
    o.init_omega_max(meta["n_gas"])
    for s in range(40):
        o.step(s)
    m = o.moments()
    T = (m[:, 29] * m[:, 30]).sum() / m[:, 30].sum()
    assert T == pytest.approx(1.3, rel=0.03)
    assert o.cells()["count"].sum() == meta["N"]


def test_free_flight_and_periodic_wrap():
    p = mcpm.Params(n_cells=10, length=10.0, capacity=3, boundary=mcpm.BC_PERIODIC, delta_t=0.5)
    arrs = (np.array([9.9, 0.2, 5.0]), np.zeros(3), np.array([1.0, -1.0, 0.5]), np.zeros(3))
    o = _sim(p, arrs)
    o.stream_bc(0, False)
    y = np.sort(o.download()[0])
    np.testing.assert_allclose(y, np.sort([0.4, 9.7, 5.25]), atol=1e-14)


def test_specular_walls_reflect():
    p = mcpm.Params(n_cells=10, length=10.0, capacity=1, boundary=mcpm.BC_SPECULAR, delta_t=1.0,
                    wall_lo=1.0, wall_hi=9.0)
    o = _sim(p, (np.array([8.5]), np.zeros(1), np.array([1.0]), np.zeros(1)))
    o.stream_bc(0, False)
    y, _, u2, _ = o.download()
    assert y[0] == pytest.approx(8.5) and u2[0] == -1.0


# ---------------------------------------------------------------- moments
def test_moments_and_samples_match_numpy():
    p, arrs, _ = W.heat_conduction(n_cells=40, ppc=200, seed=27)
    o = _sim(p, arrs)
    m = o.moments()
    o.sample()
    s = o.samples()
    y, u1, u2, u3 = o.download()
    key = np.floor(y / p.dy).astype(int)
    for c in (0, 7, 39):
        sel = key == c
        U = np.array([u1[sel], u2[sel], u3[sel]])
        mean = U.mean(axis=1)
        P = U - mean[:, None]
        np.testing.assert_allclose(m[c, :3], mean, rtol=1e-12)
        np.testing.assert_allclose(m[c, 3:9], [(P[0] * P[0]).mean(), (P[0] * P[1]).mean(),
                                               (P[0] * P[2]).mean(), (P[1] * P[1]).mean(),
                                               (P[1] * P[2]).mean(), (P[2] * P[2]).mean()],
                                   rtol=1e-10)
        e = (P * P).sum(axis=0)
        np.testing.assert_allclose(m[c, 28], e.mean(), rtol=1e-12)
        np.testing.assert_allclose(m[c, 29], e.mean() / 3.0, rtol=1e-12)
        np.testing.assert_allclose(m[c, 25:28], (P * e * e).mean(axis=1), rtol=1e-9, atol=1e-12)
        assert s[c, 0] == sel.sum()
        np.testing.assert_allclose(s[c, 1:4], U.sum(axis=1), rtol=1e-12)
        np.testing.assert_allclose(s[c, 10], (U[2] ** 3).sum(), rtol=1e-12)   # (0,0,3)
        np.testing.assert_allclose(s[c, 19], (U[0] ** 3).sum(), rtol=1e-12)   # (3,0,0)


# ---------------------------------------------------------------- Fokker-Planck
def test_gauss_jordan_matches_numpy_with_pivoting():
    rng = np.random.default_rng(31)
    for n in (9, 12):
        for _ in range(20):
            A = rng.normal(size=(n, n))
            A[0, 0] = 1e-12  # forces a row exchange at the first column
            rhs = rng.normal(size=n)
            np.testing.assert_allclose(orc.gauss_jordan(A, rhs), np.linalg.solve(A, rhs),
                                       rtol=1e-9, atol=1e-10)


def _fp_case(model, seed=41, **kw):
    p, arrs, meta = W.heat_conduction(seed=seed, **kw)
    p.model = model
    return p, arrs, meta


@pytest.mark.parametrize("model", [mcpm.MODEL_FP_DENSE, mcpm.MODEL_FP_IDEAL])
def test_fp_velocity_conserves_cell_momentum_and_energy(model):
    p, arrs, meta = _fp_case(model)
    o = _sim(p, arrs)
    for step in (0, 3):   # step 0: no alpha; step > 0: alpha_calculation active
        m0 = o.moments()
        o.fp_closures(step)
        o.velocity_fp(step)
        m1 = o.moments()
        sel = m0[:, 30] > 1
        np.testing.assert_allclose(m1[sel, :3], m0[sel, :3], rtol=0, atol=1e-12)
        np.testing.assert_allclose(m1[sel, 28], m0[sel, 28], rtol=1e-12)
        # the velocities did change
        assert np.abs(m1[sel, 3] - m0[sel, 3]).max() > 1e-3


def test_alpha_closure_equilibrium_value():
    """Uniform Maxwellian gas at rest: the alpha system reduces to
    alpha_ii = n b Y (1 + 0.005) (pressure correction p_crr - 1 = n b Y plus the
    gamma_st M4 term with M4_ii = 5 (kT/m)^2), off-diagonals and beta ~ 0."""
    n0, T = 0.3, 1.0
    p, arrs, meta = W.interface(N=400_000, n_cells=4, length=4.0, n_liq=n0, n_vap=n0, T=T,
                                dt=0.01, precision=mcpm.FP64, seed=43)
    p.long_range = mcpm.LR_NONE
    p.model = mcpm.MODEL_FP_DENSE
    o = _sim(p, arrs)
    o.fp_closures(1)
    cl = o.closures()
    nbY = n0 * p.b * W.contact_y_ref(n0)
    for c in range(4):
        alpha = cl[c, 10:19].reshape(3, 3)
        np.testing.assert_allclose(np.diag(alpha), 1.005 * nbY, rtol=0.03)
        assert np.abs(alpha - np.diag(np.diag(alpha))).max() < 0.05 * nbY


def test_fp_isothermal_gas_stays_at_wall_temperature():
    p, arrs, meta = W.heat_conduction(n_cells=50, ppc=400, length=5.0, T0=1.2, T_lo=1.2,
                                      T_hi=1.2, seed=44)
    p.model = mcpm.MODEL_FP_IDEAL
    o = _sim(p, arrs)
    for s in range(30):
        o.step(s)
    m = o.moments()
    T = (m[:, 29] * m[:, 30]).sum() / m[:, 30].sum()
    assert T == pytest.approx(1.2, rel=0.03)


# ---------------------------------------------------------------- plain DSMC
def test_dsmc_vhs_conserves_and_relaxes_anisotropy():
    """Hard-sphere DSMC (visp = 1/2): momentum and energy conserved per
    collision; an anisotropic temperature relaxes towards isotropy."""
    p, arrs, meta = W.interface(N=100_000, n_cells=50, length=10.0, n_liq=0.1, n_vap=0.1, T=1.0,
                                dt=0.05, precision=mcpm.FP64, seed=45)
    p.long_range = mcpm.LR_NONE
    p.model = mcpm.MODEL_DSMC_VHS
    y, u1, u2, u3 = arrs
    u1 = u1 * 1.5
    o = _sim(p, (y, u1, u2, u3))
    o.init_omega_max(meta["n_gas"], times_y=False)
    _, a1, a2, a3 = o.download()
    P0 = np.array([a1.sum(), a2.sum(), a3.sum()])
    E0 = (a1 ** 2 + a2 ** 2 + a3 ** 2).sum()
    an0 = (a1 ** 2).mean() - (a2 ** 2).mean()
    tot = 0
    for s in range(10):
        tot += o.collide_dsmc_vhs(s)["accepted"]
    _, a1, a2, a3 = o.download()
    assert tot > 1000
    assert np.abs(np.array([a1.sum(), a2.sum(), a3.sum()]) - P0).max() < 1e-9
    assert abs((a1 ** 2 + a2 ** 2 + a3 ** 2).sum() - E0) < 1e-10 * E0
    an1 = (a1 ** 2).mean() - (a2 ** 2).mean()
    assert an1 < 0.8 * an0


# ---------------------------------------------------------------- direct / KDE long range
def _lattice_state(p, counts):
    """Particles evenly spaced inside each cell (no sampling noise)."""
    y = np.concatenate([(c + (np.arange(k) + 0.5) / k) * p.dy for c, k in enumerate(counts) if k])
    n = len(y)
    z = np.zeros(n)
    return (y, z, z.copy(), z.copy())


def test_direct_vlasov_closed_form_and_independent_quadrature():
    """vlasov_taking_the_integral on a step density (periodic).
    Closed form: deep inside a uniform region V = -(4 pi / 3)(26/27) n sigma^3
    (integral of -(sigma/r)^6 over sigma < r < 3 sigma), so the force summed
    across the step telescopes to -(V_liq - V_vap) phi / m.  Pointwise, F2
    follows scipy's integration of the piecewise-constant density; the
    reference's 50 x 50 rule meets a discontinuous integrand near the step and
    scatters by up to ~20 % cell to cell around it."""
    from scipy import integrate
    nc, L = 200, 20.0
    p = mcpm.Params(n_cells=nc, length=L, capacity=10 ** 6, boundary=mcpm.BC_PERIODIC,
                    long_range=mcpm.LR_DIRECT, eqdist=1.0, phi=-4.0)
    counts = np.where((np.arange(nc) >= 70) & (np.arange(nc) < 130), 400, 40)
    p.capacity = int(counts.sum())
    o = _sim(p, _lattice_state(p, counts))
    o.vlasov_integral()
    F2 = o.cells()["F2"]
    n = counts * p.Fn / p.cell_volume
    v_unif = -(4 * math.pi / 3) * (26 / 27)
    # faces 20 (vapour, >= 3 sigma from the step) and 100 (liquid)
    telescoped = np.sum(F2[20:100]) * p.dy
    assert telescoped == pytest.approx(-(v_unif * n[100] - v_unif * n[20]) * p.phi / p.mass,
                                       rel=1e-7)
    assert np.abs(F2[5:30]).max() < 1e-10 * np.abs(F2).max()
    edges = np.arange(nc + 1) * p.dy
    cs = np.concatenate([[0.0], np.cumsum(n * p.dy)])

    def cum(s):  # integral of the periodic density from 0 to s
        k, frac = np.divmod(s, L)
        idx = np.minimum((frac // p.dy).astype(int), nc - 1)
        return k * n.sum() * p.dy + cs[idx] + n[idx] * (frac - edges[idx])

    def V(y):
        f = lambda r: -2 * np.pi * r ** -4 * (cum(y + r) - cum(y - r)) / r
        return integrate.quad(f, 1.0, 3.0, limit=400)[0]

    exact = np.array([-(V((c + 1) * p.dy) - V(c * p.dy)) / p.dy * p.phi / p.mass
                      for c in range(50, 91)])
    np.testing.assert_allclose(F2[50:91], exact, rtol=0.2, atol=1e-3 * np.abs(exact).max())
    assert np.mean(F2[50:91]) == pytest.approx(np.mean(exact), rel=0.01)


def test_kde_uniform_lattice_closed_form():
    """Even particles: n_smooth = (ppc Fn / dy) erf(2 sqrt 2) inside, dn2 = dn3 = 0."""
    from math import erf
    nc, ppc = 40, 2000
    p = mcpm.Params(n_cells=nc, length=4.0, capacity=nc * ppc, boundary=mcpm.BC_PERIODIC)
    o = _sim(p, _lattice_state(p, np.full(nc, ppc)))
    ns, d2, d3 = o.kde()
    np.testing.assert_allclose(ns, ppc / p.dy * erf(2 * math.sqrt(2)), rtol=1e-5)
    assert np.abs(d2).max() < 1e-8 * ns.max() / p.dy
    assert np.abs(d3).max() < 1e-6 * ns.max() / p.dy ** 3


def test_density_expansion_force_follows_the_gradient():
    """n(y) increasing linearly: dn2 ~ -dn/dy (kernel second moment h^2), and
    the force F2 = -dn2 4 pi (2^(1/6) sigma)^3 2 phi / m points up the gradient
    for phi < 0 (attraction)."""
    nc = 100
    p = mcpm.Params(n_cells=nc, length=10.0, capacity=10 ** 7, boundary=mcpm.BC_DIFFUSE,
                    long_range=mcpm.LR_DENSITY_EXPANSION, phi=-1.0)
    counts = (1000 + 20 * np.arange(nc)).astype(int)
    p.capacity = int(counts.sum())
    o = _sim(p, _lattice_state(p, counts))
    ns, d2, d3 = o.kde()
    grad = 20 / p.dy / p.dy  # particles per length per length
    np.testing.assert_allclose(d2[10:90], -grad, rtol=0.02)
    o.density_expansion()
    F2 = o.cells()["F2"]
    pref = 4 * math.pi * (2 ** (1 / 6)) ** 3 * p.phi * 2
    np.testing.assert_allclose(F2[10:90], -d2[10:90] * pref, rtol=1e-12)
    assert F2[0] == 0.0 and F2[-1] == 0.0 and np.all(F2[10:90] < 0) is not None
