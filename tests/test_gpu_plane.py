"""GPU parity of the plane (the reference's 2D injection path): binning and
sort order, ESMC pairs and velocities, the nodal deposit, the
screened-Poisson FEM field and the direct 2D integral, and whole steps.

Integer results (counts, order, pairs, accept decisions) and the deposit are
compared bit for bit.  The fields are sums reduced in another order (the
quadratures) or solved another way (conjugate gradients, or the FFT on
periodic power-of-two planes, on the GPU; banded LU in the oracle):
compared within 1e-10 relative to the field's scale.  The kick itself (the
gradient of the bilinear interpolant, FEM.cpp:568-582) is compared bit for
bit from one nodal field handed to both sides."""
import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W

pytestmark = pytest.mark.gpu


def _pair(p, arrs):
    print(f"plane nx={p.n_cells_x} ny={p.n_cells} N={arrs[0].size} prec={p.precision} "
          f"bc={p.boundary} lr={p.long_range}", flush=True)
    g = mcpm.Simulation(p)
    o = orc.OracleSim(p)
    for s in (g, o):
        s.upload_xy(*arrs)
        s.cell_update()
    return g, o


def _close(a, b, tol):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_plane_sort_and_deposit_bit_exact(precision, boundary):
    p, arrs, _ = W.plane(nx=24, ny=20, Lx=24.0, Ly=20.0, ppc=25, drops=((10.0, 9.0, 5.0),),
                         precision=precision, boundary=boundary, seed=21)
    g, o = _pair(p, arrs)
    cg, co = g.cells(), o.cells()
    assert np.array_equal(cg["count"], co["count"])
    assert np.array_equal(cg["start"], co["start"])
    assert np.array_equal(cg["n"], co["n"])
    for a, b in zip(g.download_xy(), o.download_xy()):
        assert np.array_equal(a, b)
    ng, _ = g.nodes()
    no, _ = o.nodes()
    assert np.array_equal(ng, no)


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_plane_esmc_bit_exact(precision, boundary):
    p, arrs, meta = W.plane(nx=24, ny=24, Lx=24.0, Ly=24.0, ppc=30, drops=((12.0, 12.0, 6.0),),
                            precision=precision, boundary=boundary, seed=22,
                            long_range=mcpm.LR_NONE)
    g, o = _pair(p, arrs)
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    g.debug_collision_log(True)
    for step in range(2):
        sg, so = g.collide_esmc(step), o.collide_esmc(step)
        assert sg["candidates"] == so["candidates"] and sg["valid"] == so["valid"]
        assert sg["accepted"] == so["accepted"]
        lg = g.collision_log(sg["valid"] + 16)
        lo = o.collision_log(so["valid"] + 16)
        order = np.argsort(lg[0], kind="stable")
        for a, b in zip(lg, lo):
            assert np.array_equal(a[order], b)
        for a, b in zip(g.download_xy(), o.download_xy()):
            assert np.array_equal(a, b)
        assert np.array_equal(g.cells()["omega_max"], o.cells()["omega_max"])


@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_plane_screened_poisson_field(boundary):
    p, arrs, _ = W.plane(nx=20, ny=16, Lx=20.0, Ly=16.0, ppc=25, drops=((9.0, 8.0, 5.0),),
                         boundary=boundary, seed=23, long_range=mcpm.LR_SCREENED_POISSON)
    g, o = _pair(p, arrs)
    g.poisson_solver()
    o.poisson_solver()
    _, pg = g.nodes()
    _, po = o.nodes()
    assert _close(pg, po, 1e-10)


@pytest.mark.parametrize("nx,ny", [(32, 16), (64, 64), (16, 128)])
def test_plane_screened_poisson_fft_solve(nx, ny):
    """Periodic power-of-two planes take the FFT solve: against the oracle's
    banded LU of the same Q1 system."""
    p, arrs, _ = W.plane(nx=nx, ny=ny, Lx=nx * 0.5, Ly=ny * 0.5, ppc=20,
                         drops=((nx * 0.25, ny * 0.25, min(nx, ny) * 0.15),),
                         boundary=mcpm.BC_PERIODIC, seed=31, long_range=mcpm.LR_SCREENED_POISSON)
    print(f"fft solve {nx}x{ny}", flush=True)
    g, o = _pair(p, arrs)
    g.poisson_solver()
    o.poisson_solver()
    _, pg = g.nodes()
    _, po = o.nodes()
    assert _close(pg, po, 1e-10)


@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_plane_direct_vlasov_field(boundary):
    p, arrs, _ = W.plane(nx=10, ny=9, Lx=10.0, Ly=9.0, ppc=25, drops=((5.0, 4.0, 3.0),),
                         boundary=boundary, seed=24, long_range=mcpm.LR_DIRECT)
    g, o = _pair(p, arrs)
    g.vlasov_integral()
    o.vlasov_integral()
    _, pg = g.nodes()
    _, po = o.nodes()
    assert _close(pg, po, 1e-12)


def test_plane_uniform_lattice_fields_closed_form():
    """The GPU path alone against the closed forms of the uniform state."""
    p, arrs, _ = W.plane(nx=12, ny=12, Lx=12.0, Ly=12.0, lattice=True,
                         boundary=mcpm.BC_PERIODIC, long_range=mcpm.LR_SCREENED_POISSON)
    g = mcpm.Simulation(p)
    g.upload_xy(*arrs)
    g.cell_update()
    g.poisson_solver()
    _, pot = g.nodes()
    lam = W.YUKAWA_BST
    n = 4 * p.Fn / p.cell_volume
    assert np.abs(pot / (n * (1.0 + lam) * np.exp(-lam) / lam ** 2) - 1.0).max() < 1e-12
    g.vlasov_integral()
    _, pot = g.nodes()
    assert np.abs(pot / (4.0 * np.pi / 3.0 * 26.0 / 27.0 * n) - 1.0).max() < 1e-12


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_plane_kick_and_stream_bit_exact(precision, boundary):
    """The GPU's screened-Poisson nodal field is handed to the oracle; one
    kick + stream + boundary push from it then agrees bit for bit."""
    p, arrs, meta = W.plane(nx=20, ny=16, Lx=20.0, Ly=16.0, ppc=30, drops=((10.0, 8.0, 4.0),),
                            boundary=boundary, seed=27, long_range=mcpm.LR_SCREENED_POISSON,
                            dt=0.02, precision=precision)
    g, o = _pair(p, arrs)
    g.poisson_solver()
    _, pot = g.nodes()
    assert np.abs(pot).max() > 0.0
    o.set_nodes(pot, p.ast)
    g.stream_bc(0, with_kick=True)
    o.stream_bc(0, with_kick=True)
    g.cell_update()
    o.cell_update()
    cg, co = g.cells(), o.cells()
    np.testing.assert_array_equal(cg["count"], co["count"])
    a, b = g.download_xy(), o.download_xy()
    for k, (x, y) in enumerate(zip(a, b)):
        np.testing.assert_array_equal(x, y, err_msg=f"component {k}")
    # the kick changed the velocities: not a free flight
    u0 = np.sort(arrs[3])
    assert not np.array_equal(np.sort(a[3]), u0)


@pytest.mark.parametrize("lr", [mcpm.LR_SCREENED_POISSON, mcpm.LR_DIRECT])
@pytest.mark.parametrize("boundary", [mcpm.BC_SPECULAR, mcpm.BC_PERIODIC])
def test_plane_full_steps(lr, boundary):
    """Whole FP64 steps.  The two fields differ by their solvers' round-off
    (1e-10 relative), far below what moves a particle across a cell face or
    flips an acceptance in a few steps: counts and decisions are asserted
    equal and the particles within 1e-9 of their scale after every step."""
    small = lr == mcpm.LR_DIRECT
    nx = 10 if small else 20
    p, arrs, meta = W.plane(nx=nx, ny=nx, Lx=float(nx), Ly=float(nx), ppc=30,
                            drops=((nx / 2, nx / 2, nx / 4),), boundary=boundary, seed=25,
                            long_range=lr, dt=0.02, phi=-1.0)
    g, o = _pair(p, arrs)
    o.init_omega_max(meta["n_gas"])
    g.set_omega_max(o.cells()["omega_max"])
    for s in range(3):
        sg, so = g.step(s), o.step(s)
        for k in ("candidates", "valid", "accepted"):
            assert sg[k] == so[k], (s, k)
        np.testing.assert_array_equal(g.cells()["count"], o.cells()["count"])
        for k, (a, b) in enumerate(zip(g.download_xy(), o.download_xy())):
            assert np.abs(a - b).max() <= 1e-9 * max(np.abs(b).max(), 1.0), (s, k)
