"""The plane's oracle (the reference's 2D injection path) against closed forms,
invariants and the reference's own tables.  CPU only.

What pins what:
* the 100-point Gauss-Legendre rule: the reference's table in FEM.cpp:656-657
  (a few entries typed here as golden values) and numpy's leggauss;
* the nodal deposit (Main_Functions.cpp:2268-2364): a node-centric count in
  numpy (each particle credits the corner nearest to it) and the exactly
  uniform quarter-point lattice;
* compute_Vlasov_2D: (4 pi / 3)(26/27) n sigma^3 for uniform n;
* compute_SP_2D: uniform n gives n / lambda^2 before the core subtraction on a
  periodic plane (the mass matrix rows sum to one, the stiffness rows to
  zero), the truncated Yukawa integral n (1 - (1 + 3 sigma lambda)
  e^{-3 sigma lambda}) / lambda^2 on the walls, and n / lambda^2 deep inside
  a walled plane; the core term n (1 - (1 + sigma lambda) e^{-sigma lambda})
  / lambda^2 comes off every node;
* the force interpolation (FEM.cpp:568-582): exact for bilinear fields;
* ESMC on the plane: momentum and energy conservation, the Enskog rate;
* specular walls: speeds kept, every particle inside."""
import math

import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W


def _lam():
    return W.YUKAWA_BST


def test_gl100_matches_reference_table_and_leggauss():
    # The rule is computed in long double and rounded; the reference's table
    # (numpy's leggauss output) carries weight errors up to ~1e-11 relative
    # at the ends of the interval, nodes agree to the last bit or so.
    x, w = orc.gl_full(100)
    xr, wr = np.polynomial.legendre.leggauss(100)
    assert np.abs(x - xr).max() < 3e-16
    assert np.abs(w / wr - 1.0).max() < 2e-11
    assert abs(w.sum() - 2.0) < 1e-14
    # odd polynomials integrate to zero, x^198 to 2/199
    assert abs((w * x ** 198).sum() - 2.0 / 199.0) < 1e-14
    # FEM.cpp:656-657, first and middle entries
    assert abs(x[0] - (-0.9997137267734413)) < 2e-16
    assert abs(w[0] / 0.0007346344905008809 - 1.0) < 2e-11
    assert abs(x[49] - (-0.015628984421543084)) < 2e-16
    assert abs(w[49] / 0.03125542345386349 - 1.0) < 1e-14


def _oracle(p, arrs):
    o = orc.OracleSim(p)
    o.upload_xy(*arrs)
    o.cell_update()
    return o


@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_deposit_lattice_is_uniform(boundary):
    p, arrs, meta = W.plane(nx=6, ny=5, Lx=6.0, Ly=5.0, lattice=True, boundary=boundary)
    o = _oracle(p, arrs)
    nd, _ = o.nodes()
    n = 4 * p.Fn / p.cell_volume
    assert np.allclose(nd, n, rtol=0, atol=1e-15 * n)
    assert np.allclose(o.cells()["n"], n, rtol=1e-15)


@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_deposit_matches_node_centric_count(boundary):
    p, arrs, meta = W.plane(nx=7, ny=9, Lx=7.0, Ly=9.0, ppc=30, drops=((3.0, 4.0, 2.0),),
                            boundary=boundary, seed=5)
    o = _oracle(p, arrs)
    nd, _ = o.nodes()
    x, y = arrs[0], arrs[1]
    dx, dy = p.dx, p.dy
    j = np.floor(x / dx).astype(int)
    k = np.floor(y / dy).astype(int)
    i_node = j + (x >= 0.5 * (j * dx + (j + 1) * dx))
    j_node = k + (y >= 0.5 * (k * dy + (k + 1) * dy))
    cnt = np.zeros((p.n_cells + 1, p.n_cells_x + 1))
    if boundary == mcpm.BC_PERIODIC:
        i_node %= p.n_cells_x
        j_node %= p.n_cells
    np.add.at(cnt, (j_node, i_node), 1.0)
    if boundary == mcpm.BC_PERIODIC:
        cnt[:, -1] = cnt[:, 0]
        cnt[-1, :] = cnt[0, :]
        w = np.ones_like(cnt)
    else:
        w = np.ones_like(cnt)
        w[:, 0] *= 2; w[:, -1] *= 2; w[0, :] *= 2; w[-1, :] *= 2
    assert np.array_equal(nd, cnt * (p.Fn / p.cell_volume) * w)
    # every particle credits exactly one node
    assert abs((nd / w)[: p.n_cells + (boundary != mcpm.BC_PERIODIC),
                       : p.n_cells_x + (boundary != mcpm.BC_PERIODIC)].sum() * p.cell_volume / p.Fn
               - meta["N"]) < 1e-6


@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_vlasov2d_uniform_closed_form(boundary):
    p, arrs, _ = W.plane(nx=8, ny=8, Lx=8.0, Ly=8.0, lattice=True, boundary=boundary,
                         long_range=mcpm.LR_DIRECT)
    o = _oracle(p, arrs)
    o.vlasov_integral()
    _, pot = o.nodes()
    n = 4 * p.Fn / p.cell_volume
    exact = 4.0 * math.pi / 3.0 * 26.0 / 27.0 * n
    assert np.abs(pot / exact - 1.0).max() < 1e-12


def test_sp2d_uniform_periodic_closed_form():
    p, arrs, _ = W.plane(nx=10, ny=12, Lx=10.0, Ly=12.0, lattice=True,
                         boundary=mcpm.BC_PERIODIC, long_range=mcpm.LR_SCREENED_POISSON)
    o = _oracle(p, arrs)
    o.poisson_solver()
    _, pot = o.nodes()
    lam = _lam()
    n = 4 * p.Fn / p.cell_volume
    exact = n * (1.0 + lam) * math.exp(-lam) / lam ** 2   # n/lam^2 minus the core term
    assert np.abs(pot / exact - 1.0).max() < 1e-12


def test_sp2d_walls_uniform_closed_forms():
    p, arrs, _ = W.plane(nx=24, ny=24, Lx=24.0, Ly=24.0, lattice=True,
                         boundary=mcpm.BC_SPECULAR, long_range=mcpm.LR_SCREENED_POISSON)
    o = _oracle(p, arrs)
    o.poisson_solver()
    _, pot = o.nodes()
    lam = _lam()
    n = 4 * p.Fn / p.cell_volume
    core = n * (1.0 - (1.0 + lam) * math.exp(-lam)) / lam ** 2
    wall = n * (1.0 - (1.0 + 3.0 * lam) * math.exp(-3.0 * lam)) / lam ** 2 - core
    edges = np.concatenate([pot[0, :], pot[-1, :], pot[:, 0], pot[:, -1]])
    # 8-decimal 50-point rule: about 1e-8 relative
    assert np.abs(edges / wall - 1.0).max() < 1e-7
    deep = n / lam ** 2 - core
    assert abs(pot[12, 12] / deep - 1.0) < 1e-9


def test_force_interpolation_exact_for_bilinear_fields():
    p, arrs, _ = W.plane(nx=5, ny=4, Lx=10.0, Ly=6.0, lattice=True,
                         long_range=mcpm.LR_SCREENED_POISSON)
    o = _oracle(p, arrs)
    X, Y = np.meshgrid(np.arange(p.n_cells_x + 1) * p.dx, np.arange(p.n_cells + 1) * p.dy)
    a, b, c, d = 0.7, -1.3, 0.25, 3.0
    pot = a * X + b * Y + c * X * Y + d
    coef, m = -2.5, p.mass
    rng = np.random.default_rng(3)
    for _ in range(50):
        x, y = rng.random() * p.length_x, rng.random() * p.length
        f1, f2 = o.fem_force(x, y, pot, coef)
        assert abs(f1 - (a + c * y) * coef / m) < 1e-12
        assert abs(f2 - (b + c * x) * coef / m) < 1e-12


def _energy_momentum(arrs):
    u = np.stack(arrs[2:5])
    return u.sum(axis=1), (u * u).sum()


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_plane_esmc_conserves_and_matches_enskog_rate(precision):
    n_gas, T = 0.3, 1.0
    p, arrs, meta = W.plane(nx=16, ny=16, Lx=16.0, Ly=16.0, ppc=60, n_vap=n_gas, drops=(),
                            T=T, boundary=mcpm.BC_PERIODIC, long_range=mcpm.LR_NONE,
                            precision=precision, seed=9, dt=0.02)
    o = _oracle(p, arrs)
    o.init_omega_max(meta["n_gas"])
    before = _energy_momentum(o.download_xy())
    acc = 0
    steps = 3
    for s in range(steps):
        st = o.collide_esmc(s)
        acc += st["accepted"]
    after = _energy_momentum(o.download_xy())
    tol = 1e-9 if precision == mcpm.FP64 else 2e-3
    assert np.abs(after[0] - before[0]).max() < tol * math.sqrt(meta["N"])
    assert abs(after[1] / before[1] - 1.0) < (1e-12 if precision == mcpm.FP64 else 1e-5)
    # Enskog: N nu dt / 2 collisions per step, nu = 4 n sigma^2 Y sqrt(pi T / m)
    nu = 4.0 * n_gas * W.contact_y_ref(n_gas) * math.sqrt(math.pi * T)
    expect = meta["N"] * nu * p.delta_t / 2.0 * steps
    assert abs(acc / expect - 1.0) < 0.06


def test_plane_specular_walls_keep_speed_and_particles_inside():
    p, arrs, meta = W.plane(nx=6, ny=5, Lx=6.0, Ly=5.0, ppc=40, drops=(), T=4.0,
                            boundary=mcpm.BC_SPECULAR, long_range=mcpm.LR_NONE,
                            model=mcpm.MODEL_NONE, dt=0.7, seed=4)
    o = _oracle(p, arrs)
    x0, y0, a0, b0, c0 = o.download_xy()
    order = np.lexsort((b0, a0))
    for s in range(5):
        o.step(s)
    x, y, a, b, c = o.download_xy()
    assert (x >= 0).all() and (x <= p.length_x).all() and (y >= 0).all() and (y <= p.length).all()
    # reflections only flip signs: the multiset of |u1|, |u2|, u3 is unchanged
    assert np.array_equal(np.sort(np.abs(a)), np.sort(np.abs(a0)))
    assert np.array_equal(np.sort(np.abs(b)), np.sort(np.abs(b0)))
    assert np.array_equal(np.sort(c), np.sort(c0))


def test_plane_periodic_free_flight_closed_form():
    p, arrs, meta = W.plane(nx=6, ny=5, Lx=6.0, Ly=5.0, ppc=20, drops=(), T=1.0,
                            boundary=mcpm.BC_PERIODIC, long_range=mcpm.LR_NONE,
                            model=mcpm.MODEL_NONE, dt=0.3, seed=6)
    o = _oracle(p, arrs)
    x0, y0, a0, b0, c0 = o.download_xy()
    o.step(0)
    x, y, a, b, c = o.download_xy()
    xe = np.mod(x0 + a0 * p.delta_t, p.length_x)
    ye = np.mod(y0 + b0 * p.delta_t, p.length)
    key0 = np.round(c0, 15)
    idx = np.argsort(c0)
    jdx = np.argsort(c)
    assert np.array_equal(c0[idx], c[jdx])
    assert np.abs(x[jdx] - xe[idx]).max() < 1e-12
    assert np.abs(y[jdx] - ye[idx]).max() < 1e-12


@pytest.mark.parametrize("h,unstable", [(0.25, True), (0.5, True), (1.0, False)])
def test_core_stencil_makes_the_grid_mode_attractive_on_fine_meshes(h, unstable):
    """Sign of the screened-Poisson field's response to a checkerboard of
    nodal density on a periodic plane (compute_SP_2D with the core
    subtraction of FEM.cpp:497-554).  The Laplacian part of the core term,
    dummy2 / dx^2 (FEM.cpp:502), grows as 1/h^2; for h below about 0.67
    sigma it outweighs the point term at the grid's Nyquist mode, so the
    potential rises where the density does and, with ast < 0, the force
    pulls particles into the denser nodes: the grid-scale mode is unstable.
    On a sigma-sized mesh the response is repulsive."""
    n = 16
    L = n * h
    p, _, _ = W.plane(nx=n, ny=n, Lx=L, Ly=L, ppc=1, boundary=mcpm.BC_PERIODIC,
                      long_range=mcpm.LR_SCREENED_POISSON, model=mcpm.MODEL_NONE)
    xs, ys, heavy = [], [], []
    for j in range(n):
        for i in range(n):
            c = 6 if (i + j) % 2 == 0 else 4   # particles per quadrant around node (i, j)
            for dx in (-0.25 * h, 0.25 * h):
                for dy in (-0.25 * h, 0.25 * h):
                    xs += [((i * h + dx) % L)] * c
                    ys += [((j * h + dy) % L)] * c
    x, y = np.array(xs), np.array(ys)
    p.Fn = 0.3 * L * L / x.size
    p.capacity = x.size
    z = np.zeros_like(x)
    o = orc.OracleSim(p)
    o.upload_xy(x, y, z, z, z)
    o.cell_update()
    o.poisson_solver()
    nd, pot = o.nodes()
    nd, pot = nd[:-1, :-1], pot[:-1, :-1]
    even = (np.add.outer(np.arange(n), np.arange(n)) % 2) == 0
    dn = nd[even].mean() - nd[~even].mean()
    dphi = pot[even].mean() - pot[~even].mean()
    assert dn > 0
    assert (dphi > 0) == unstable, (h, dphi / dn)
