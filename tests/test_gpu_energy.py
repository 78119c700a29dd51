"""The FP64 energy check on the device (mcpm_energy_momentum): equal to the
sums over the downloaded particles, and conserved by the Enskog collisions
(exact exchanges, round-off only, in FP64)."""
import numpy as np
import pytest

import mcpm
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [mcpm.FP64, mcpm.FP32])
def test_energy_equals_host_sums(precision):
    p, arrs, meta = W.heat_conduction(n_cells=100, ppc=200, precision=precision, seed=2)
    print(f"energy precision={precision}", flush=True)
    g = mcpm.Simulation(p)
    g.upload(*arrs)
    g.cell_update()
    e = g.energy_momentum()
    y, u1, u2, u3 = g.download()
    mf = p.mass * p.Fn
    ref = np.array([0.5 * mf * np.sum(u1 * u1 + u2 * u2 + u3 * u3), mf * u1.sum(), mf * u2.sum(),
                    mf * u3.sum()])
    assert abs(e[0] - ref[0]) <= 1e-12 * ref[0]
    assert np.all(np.abs(e[1:] - ref[1:]) <= 1e-9 * np.sqrt(ref[0] * mf * y.size))


def test_esmc_conserves_energy_and_momentum_fp64():
    p, arrs, meta = W.heat_conduction(n_cells=100, ppc=200, seed=4)
    g = mcpm.Simulation(p)
    g.upload(*arrs)
    g.cell_update()
    g.init_omega_max(meta["n_gas"])
    e0 = g.energy_momentum()
    st = g.collide_esmc(0)
    assert st["accepted"] > 100
    e1 = g.energy_momentum()
    assert abs(e1[0] - e0[0]) <= 1e-12 * e0[0]
    assert np.all(np.abs(e1[1:] - e0[1:]) <= 1e-10 * np.sqrt(e0[0] * p.mass * p.Fn * 20000))


def test_energy_on_the_plane():
    p, arrs, meta = W.droplets(nx=64, ny=64, Lx=16.0, Ly=16.0, N=64 * 64 * 40, n_drops=3,
                               r_lo=2.0, r_hi=3.5, seed=5)
    g = mcpm.Simulation(p)
    g.upload_xy(*arrs)
    g.cell_update()
    x, y, u1, u2, u3 = g.download_xy()
    e = g.energy_momentum()
    ref = 0.5 * p.mass * p.Fn * np.sum(u1 * u1 + u2 * u2 + u3 * u3)
    assert abs(e[0] - ref) <= 1e-12 * ref
