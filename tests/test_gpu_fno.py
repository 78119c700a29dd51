"""The field surrogate on the GPU (csrc/mcpm_fno.hip) against its CPU oracle
(oracle/fno_oracle.py): forward pass, loss and full gradient on the same
parameters and inputs; then training on snapshots that the exact GPU solver
writes, the surrogate field inside a plane context, and what the trained
network gains over the field without it."""
import numpy as np
import pytest

import fno_oracle as F
import mcpm
import workloads as W

pytestmark = pytest.mark.gpu


def _inputs(n0, n1, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.arange(n1), np.arange(n0), indexing="ij")
    u = np.sin(2 * np.pi * xx / n0 * 3) * np.cos(2 * np.pi * yy / n1 * 2) + 0.3 * rng.normal(size=(n1, n0))
    t = 0.5 * np.cos(2 * np.pi * xx / n0) + 0.2 * rng.normal(size=(n1, n0))
    return u.astype(np.float32), t.astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n0,n1,m", [(64, 64, 16), (128, 32, 8)])
def test_forward_matches_oracle(n0, n1, m):
    p = F.init_params(m, seed=n0 + m)
    u, _ = _inputs(n0, n1, 1)
    print(f"fno forward n0={n0} n1={n1} m={m}", flush=True)
    net = mcpm.FieldSurrogate(n0, n1, m, p)
    y = net.forward(u)
    yo, _ = F.forward(p, u.astype(np.float64), m)
    assert _rel(y, yo) < 2e-3, _rel(y, yo)
    # no BF16 in the oracle: the roundings account for the rest
    ye, _ = F.forward(p, u.astype(np.float64), m, exact=True)
    assert _rel(y, ye) < 5e-2


@pytest.mark.parametrize("n0,n1,m", [(64, 64, 16), (64, 32, 8)])
def test_loss_and_gradient_match_oracle(n0, n1, m):
    p = F.init_params(m, seed=3)
    off = F.layout(m)
    rng = np.random.default_rng(9)
    for l in range(F.LAYERS):  # spectral weights large enough to matter
        n = 2 * m * m * F.C * F.C * 2
        p[off[f"R{l}"]:off[f"R{l}"] + n] = rng.uniform(-1, 1, n).astype(np.float32) / F.C
    u, t = _inputs(n0, n1, 2)
    w0 = 0.1
    print(f"fno loss/grad n0={n0} n1={n1} m={m}", flush=True)
    net = mcpm.FieldSurrogate(n0, n1, m, p)
    L, g, y = net.loss_grad(u, t, w0)
    Lo, go, yo = F.loss_grad(p, u.astype(np.float64), t.astype(np.float64), m, w0)
    assert abs(L - Lo) < 2e-3 * Lo
    groups = [("wl", F.C), ("bl", F.C), ("wq", F.C), ("bq", 1)]
    for l in range(F.LAYERS):
        groups += [(f"W{l}", F.C * F.C), (f"b{l}", F.C), (f"R{l}", 2 * m * m * F.C * F.C * 2)]
    for name, n in groups:
        a, b = g[off[name]:off[name] + n], go[off[name]:off[name] + n]
        assert _rel(a, b) < 2e-2, (name, _rel(a, b))


def _force(pot, h):
    """Gradient of the Q1 interpolant at cell centres (periodic nodes)."""
    gx = 0.5 * ((pot[:-1, 1:] - pot[:-1, :-1]) + (pot[1:, 1:] - pot[1:, :-1])) / h
    gy = 0.5 * ((pot[1:, :-1] - pot[:-1, :-1]) + (pot[1:, 1:] - pot[:-1, 1:])) / h
    return np.stack([gx, gy])


def _droplet_sim(n=64, seed=3):
    p, arrs, meta = W.droplets(nx=n, ny=n, Lx=n / 4.0, Ly=n / 4.0, N=n * n * 60, n_drops=3,
                               r_lo=2.0, r_hi=3.5, seed=seed)
    sim = mcpm.Simulation(p)
    sim.upload_xy(*arrs)
    sim.cell_update()
    sim.init_omega_max(meta["n_gas"])
    return p, sim


def test_training_on_solver_snapshots_and_the_surrogate_field():
    n, m, S, hold = 64, 16, 40, 8
    p, sim = _droplet_sim(n)
    print("fno training: collecting snapshots", flush=True)
    sim.poisson_solver()
    nd, pot = sim.nodes()
    net = mcpm.FieldSurrogate(n, n, m, F.init_params(m, seed=11))
    net.set_norm(nd.mean(), 1.0 / nd.std(), pot.mean(), 1.0 / pot.std())
    net.reserve_snapshots(S)
    for s in range(S):
        sim.step(s)
        net.add_snapshot(sim, 0.05)
    assert net.snapshot_count() == S

    # force error of the network's field against the exact one, measured
    # through the context on the last snapshot's positions
    def ctx_force_error():
        st = net.field(sim, accept_residual=-1.0)
        _, pnet = sim.nodes()
        sim.poisson_solver()
        _, pex = sim.nodes()
        h = p.dx
        fe, fn = _force(pex, h), _force(pnet, h)
        return np.linalg.norm(fn - fe) / np.linalg.norm(fe), st

    e0, st0 = ctx_force_error()
    assert st0["used_network"] and st0["residual"] > 0
    losses = net.train(600, 1e-3, seed=5, n_snap=S - hold)
    assert np.all(np.isfinite(losses))
    print("loss first/last", losses[:5].mean(), losses[-20:].mean(), flush=True)
    assert losses[-20:].mean() < 0.5 * losses[:5].mean()
    e1, st1 = ctx_force_error()
    # the best field without the spectral path: a pointwise gain of the density
    _, pex = sim.nodes()
    nd, _ = sim.nodes()
    f_ex = _force(pex, p.dx)
    f_n = _force(nd, p.dx)
    a = float(np.sum(f_ex * f_n) / np.sum(f_n * f_n))
    e_local = np.linalg.norm(a * f_n - f_ex) / np.linalg.norm(f_ex)
    print(f"force error: untrained {e0:.3f} trained {e1:.3f} best pointwise {e_local:.3f}; "
          f"residual {st0['residual']:.3e} -> {st1['residual']:.3e}", flush=True)
    assert e1 < 0.5 * e0
    assert e1 < e_local
    assert st1["residual"] < st0["residual"]


def test_surrogate_field_exact_mode_equals_the_solver_and_drives_the_step():
    n, m = 64, 8
    p, sim = _droplet_sim(n, seed=4)
    sim.poisson_solver()
    nd, pex = sim.nodes()
    net = mcpm.FieldSurrogate(n, n, m, F.init_params(m, seed=2))
    net.set_norm(nd.mean(), 1.0 / nd.std(), pex.mean(), 1.0 / pex.std())
    st = net.field(sim, exact=True)
    assert not st["used_network"]
    _, pf = sim.nodes()
    assert np.max(np.abs(pf - pex)) < 1e-9 * np.max(np.abs(pex))
    st = net.field(sim, accept_residual=1e9)  # any residual accepted
    assert st["used_network"] and st["cg_iterations"] == 0
    # attached to the step: exact every second step, else the network's field
    sim.set_field_surrogate(net, accept_residual=-1.0, every=2)
    sim.step(0)
    assert not sim.field_surrogate_stats()["used_network"]
    sim.step(1)
    assert sim.field_surrogate_stats()["used_network"]
    sim.set_field_surrogate(None)
    sim.step(2)
    x, y, *_ = sim.download_xy()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(y))
