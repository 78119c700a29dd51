"""What pins the CPU oracle of the field surrogate (oracle/fno_oracle.py):
its backward pass against central finite differences, its spectral layer
against numpy's FFT, its BF16 rounding against torch's, its loss gradient,
and the closed-form FEM transfer function against the 9-point stencils of
the Q1 system (FEM.cpp:73-110)."""
import numpy as np
import pytest

import fno_oracle as F


def _small(seed=0, n0=16, n1=8, m=4):
    rng = np.random.default_rng(seed)
    p = F.init_params(m, seed).astype(np.float64)
    # larger spectral weights than the initial ones, so that the spectral
    # path matters in the finite differences
    off = F.layout(m)
    for l in range(F.LAYERS):
        n = 2 * m * m * F.C * F.C * 2
        p[off[f"R{l}"]:off[f"R{l}"] + n] = rng.uniform(-1, 1, n) / F.C
    u = rng.normal(size=(n1, n0))
    t = rng.normal(size=(n1, n0))
    return p, u, t, m


def test_backward_matches_central_differences():
    p, u, t, m = _small()
    w0 = 0.3
    L, g, _ = F.loss_grad(p, u, t, m, w0, exact=True)
    off = F.layout(m)
    rng = np.random.default_rng(1)
    idx = [off["wl"] + 3, off["bl"] + 7, off["wq"] + 5, off["bq"]]
    for l in range(F.LAYERS):
        idx += [off[f"W{l}"] + int(rng.integers(F.C * F.C)), off[f"b{l}"] + int(rng.integers(F.C))]
        n = 2 * m * m * F.C * F.C * 2
        idx += [off[f"R{l}"] + int(rng.integers(n)) for _ in range(3)]
        # the kx = 0 and the negative-ky rows, real and imaginary parts
        idx += [off[f"R{l}"] + ((((2 * m - 1) * m + 0) * F.C + 2) * F.C + 9) * 2 + 1]
    for k in idx:
        h = 1e-4 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (F.loss_grad(pp, u, t, m, w0, exact=True)[0] - F.loss_grad(pm, u, t, m, w0, exact=True)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-5 * abs(fd), (k, fd, g[k])


def test_loss_gradient_in_y_matches_differences():
    rng = np.random.default_rng(3)
    y = rng.normal(size=(8, 16))
    t = rng.normal(size=(8, 16))
    L, dy = F.loss_and_dy(y, t, 0.2)
    for (j, i) in [(0, 0), (3, 15), (7, 4)]:
        e = np.zeros_like(y)
        e[j, i] = 1e-6
        fd = (F.loss_and_dy(y + e, t, 0.2)[0] - F.loss_and_dy(y - e, t, 0.2)[0]) / 2e-6
        assert abs(fd - dy[j, i]) < 1e-7


def test_spectral_layer_is_the_band_limited_fft_filter():
    """With W = 0, b = 0, R = identity, lift w = 1: z0 = the real band-limited
    part of u with kx in [0, m), ky in [0, m) or [N1 - m, N1), the kx = 0
    column symmetrised (numpy's ifft2 of the Hermitian-completed mask)."""
    n0, n1, m = 32, 16, 4
    off = F.layout(m)
    p = np.zeros(off["n"])
    p[off["wl"]:off["wl"] + F.C] = 1.0
    for l in range(F.LAYERS):
        r = np.zeros((2 * m, m, F.C, F.C, 2))
        r[:, :, np.arange(F.C), np.arange(F.C), 0] = 1.0
        p[off[f"R{l}"]:off[f"R{l}"] + r.size] = r.ravel()
    u = np.random.default_rng(5).normal(size=(n1, n0))
    _, cache = F.forward(p, u, m, exact=True)
    z0 = cache["z"][0][:, :, 11]
    w = np.zeros((n1, n0))
    ky = np.r_[np.arange(m), np.arange(n1 - m, n1)]
    for kx in range(m):
        w[ky, kx] = 1.0 if kx == 0 else 2.0
    wm = np.roll(np.flip(w, (0, 1)), (1, 1), (0, 1))  # w(-k)
    ref = np.real(np.fft.ifft2(np.fft.fft2(u) * (w + wm) / 2))
    assert np.allclose(z0, ref, atol=1e-12)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(7).normal(size=4096) * np.exp(np.random.default_rng(8).normal(size=4096) * 5)
    x = np.r_[x, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -9, -1 - 2.0 ** -8, 0.0]
    ref = torch.tensor(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(F.bf16(x), ref)
    assert F.bf16(np.array([1 + 2.0 ** -8]))[0] == 1.0          # tie to even
    assert F.bf16(np.array([1 + 3 * 2.0 ** -8]))[0] == 1 + 2.0 ** -6  # tie to even, upwards


def test_fem_transfer_solves_the_q1_stencils():
    """phi = ifft(T fft(n)) satisfies (K + lambda^2 M) phi = M n with the
    assembled Q1 stencils of square cells (FEM.cpp:73-110)."""
    n1, n0, h, lam = 24, 32, 0.25, 2.353893
    n = np.random.default_rng(2).uniform(0.02, 0.55, size=(n1, n0))
    phi = F.fem_solve_fft(n, h, lam)

    def apply9(v, c, e, d):
        s = c * v
        for dy, dx in ((0, 1), (0, -1), (1, 0), (-1, 0)):
            s = s + e * np.roll(v, (dy, dx), (0, 1))
        for dy, dx in ((1, 1), (1, -1), (-1, 1), (-1, -1)):
            s = s + d * np.roll(v, (dy, dx), (0, 1))
        return s
    V = h * h
    A = apply9(phi, 8 / 3 + 4 * lam ** 2 * V / 9, -1 / 3 + lam ** 2 * V / 9, -1 / 3 + lam ** 2 * V / 36)
    b = apply9(n, 4 * V / 9, V / 9, V / 36)
    assert np.max(np.abs(A - b)) < 1e-12 * np.max(np.abs(b))
    # uniform density: phi = n / lambda^2
    assert np.allclose(F.fem_solve_fft(np.full((8, 8), 0.3), h, lam), 0.3 / lam ** 2)


def test_layout_matches_the_library():
    mcpm = pytest.importorskip("mcpm")
    for m in (8, 16):
        assert mcpm.FieldSurrogate.param_count(m) == F.layout(m)["n"]


def test_package_initial_parameters_match_the_oracle_layout():
    mcpm = pytest.importorskip("mcpm")
    for m, seed in ((8, 1), (16, 4)):
        assert np.array_equal(mcpm.FieldSurrogate.init_params(m, seed), F.init_params(m, seed))


def _nodes_of(x, y, n0, n1, h, w):
    """Periodic node densities from particle quarter counts (each node sums
    the four cell quarters that touch it), with the seam row and column."""
    qx = np.floor(x / (h / 2)).astype(np.int64) % (2 * n0)
    qy = np.floor(y / (h / 2)).astype(np.int64) % (2 * n1)
    q = np.bincount(qy * (2 * n0) + qx, minlength=4 * n0 * n1).reshape(2 * n1, 2 * n0)
    q = q + np.roll(q, 1, axis=1)
    q = q + np.roll(q, 1, axis=0)
    n = q[0::2, 0::2] * w
    out = np.empty((n1 + 1, n0 + 1))
    out[:-1, :-1] = n
    out[-1, :-1] = n[0]
    out[:-1, -1] = n[:, 0]
    out[-1, -1] = n[0, 0]
    return out


def test_spectral_force_floor_closed_cases():
    """mcpm.spectral_force_floor: no error left when every mode is kept or
    the density has only retained modes; with no mode kept, the rest is the
    force of the FEM solution alone (the oracle's FFT solve)."""
    mcpm = pytest.importorskip("mcpm")
    n1 = n0 = 32
    h, lam, sig = 0.25, 2.353893, 1.0
    rng = np.random.default_rng(5)
    n = rng.uniform(0.02, 0.55, size=(n1, n0))
    seam = mcpm._periodic_nodes(n)
    assert mcpm.spectral_force_floor(seam, h, lam, sig, n0)["floor"] < 1e-13
    jj, ii = np.mgrid[0:n1, 0:n0]
    smooth = 0.3 + 0.1 * np.cos(2 * np.pi * 3 * ii / n0) * np.sin(2 * np.pi * 2 * jj / n1)
    r = mcpm.spectral_force_floor(mcpm._periodic_nodes(smooth), h, lam, sig, 4)
    assert r["floor"] < 1e-12 and r["modes_only"] < 1e-12
    exact = mcpm.plane_force(mcpm._periodic_nodes(F.fem_solve_fft(n, h, lam) - mcpm.plane_core(n, h, lam, sig)), h)
    fem = mcpm.plane_force(mcpm._periodic_nodes(F.fem_solve_fft(n, h, lam)), h)
    r0 = mcpm.spectral_force_floor(seam, h, lam, sig, 0)
    assert abs(r0["modes_only"] - np.linalg.norm(fem) / np.linalg.norm(exact)) < 1e-12


def test_sixteen_modes_cannot_reach_1e3_force_error_at_244_particles_per_cell():
    """The droplet plane's density at 0.25 sigma cells and about 244
    particles per cell: its particle noise puts force at wave numbers no
    16-mode operator keeps, so the best 16-mode operator (exact on its
    modes, the best pointwise gain above) misses the FEM force by far more
    than 1e-3; the floor falls as modes are added."""
    mcpm = pytest.importorskip("mcpm")
    W = pytest.importorskip("workloads")
    p, arrs, meta = W.droplets(nx=128, ny=128, Lx=32.0, Ly=32.0, N=4_000_000, n_drops=3,
                               r_lo=3.0, r_hi=6.0)
    h = p.length / p.n_cells
    nodes = _nodes_of(arrs[0], arrs[1], p.n_cells_x, p.n_cells, h, p.Fn / (h * h))
    fl = [mcpm.spectral_force_floor(nodes, h, p.bst, p.sigma, m)["floor"] for m in (16, 32, 64)]
    assert fl[0] > 1e-2, fl
    assert fl[0] > fl[1] > fl[2], fl
