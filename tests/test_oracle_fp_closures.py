"""The oracle's Fokker-Planck closures pinned by what they are built to do.

c_gamma_calculation (Fokker_Planck.cpp:590-804) and alpha_calculation
(:200-587) assemble 9x9 and 12x12 systems coefficient by coefficient.  Here
the solutions are checked against the conditions those systems encode,
evaluated directly on each cell's particles with numpy (sample averages of
the peculiar velocities, never the assembled matrices), on strongly
non-equilibrium cells (anisotropic, skewed, heat-carrying):

* velocity drift A(v) = C v + gamma (|v|^2 - <|v|^2>) + Lambda (v |v|^2 - Q):
  - it adds nothing to the pressure tensor: <A_i v_j + A_j v_i> = 0;
  - it sets the heat-flux production <A_i |v|^2 + 2 v_i A.v> to
    Y (3 - 2 Pr) Q_i / tau with Pr = 2/3 (the linear Langevin part removes
    3 Y Q_i / tau, so Q relaxes at 2 Pr Y / tau);
  - Lambda = -det(P - <|v|^2>/3 I) / (tau <|v|^2>^4) (cell_update,
    Main_Functions.cpp:2718-2728), tau = 2 mu(T) / (n kB T);
* position drift alpha v + beta (|v|^2 - <|v|^2>) of FP_dense (after step 0):
  - momentum-flux rows <v_j (alpha_i . v)> + Q_i beta_i equal the dense-gas
    corrections (mu_crr - 1 + coeff_mu) dev_ij + delta_ij ((p_crr - 1) tr/3
    - w div U / (m n)) - gamma_st M4_ij, with mu_crr = 1 + 2 nbY/5,
    p_crr = 1 + nbY, w = (4/9) sqrt(pi) Y n^2 sigma^4 sqrt(m kB T), the
    1.002 coefficients and div U from the neighbouring cells' mean velocity
    (one-sided at the walls), as the reference writes them (its beta_i
    column carries Q_i in every row of block i);
  - energy rows <(|v|^2 - <|v|^2>)(alpha_i . v)> + (<|v|^4> - <|v|^2>^2)
    beta_i equal (k_crr - 1 + coeff_k) Q_i - gamma_st (M5_i - <|v|^2> Q_i).
A wrong matrix coefficient (a 2 for a 4, a dropped Q, a swapped index)
breaks one of these identities."""
import math

import numpy as np
import pytest

import mcpm
import oracle as orc
import workloads as W


def _cells_with_designed_velocities(model, seed, ppc=600, n_cells=6):
    p, arrs, meta = W.heat_conduction(n_cells=n_cells, ppc=ppc, seed=seed, n_density=0.3)
    p.model = model
    y = arrs[0]
    rng = np.random.default_rng(seed)
    k = np.minimum(np.floor(y / p.dy).astype(int), n_cells - 1)
    u = np.empty((y.size, 3))
    for c in range(n_cells):
        sel = k == c
        m = int(sel.sum())
        z = rng.standard_normal((m, 3))
        L = np.eye(3) + 0.4 * rng.standard_normal((3, 3))      # anisotropic pressure
        skew = 0.3 * rng.standard_normal(3)                      # heat flux
        v = z @ L.T + skew * (z ** 2 - 1.0) + 0.2 * rng.standard_normal(3) + 0.05 * c
        u[sel] = v
    return p, (y, u[:, 0].copy(), u[:, 1].copy(), u[:, 2].copy())


def _per_cell(o, p):
    y, u1, u2, u3 = o.download()
    cells = o.cells()
    U = np.stack([u1, u2, u3], axis=1)
    out = []
    for c in range(p.n_cells):
        s, n = cells["start"][c], cells["count"][c]
        out.append(U[s:s + n])
    return out, cells


def _moments(u):
    U = u.mean(axis=0)
    v = u - U
    e = (v * v).sum(axis=1)
    P = v.T @ v / len(v)
    Q = (v * e[:, None]).mean(axis=0)
    DM2 = e.mean()
    DM4 = (e * e).mean()
    M4 = (v.T * e) @ v / len(v)
    M5 = (v * (e * e)[:, None]).mean(axis=0)
    return dict(U=U, v=v, e=e, P=P, Q=Q, DM2=DM2, DM4=DM4, M4=M4, M5=M5)


def _sym(c6):
    return np.array([[c6[0], c6[1], c6[2]], [c6[1], c6[3], c6[4]], [c6[2], c6[4], c6[5]]])


@pytest.mark.parametrize("model", [mcpm.MODEL_FP_DENSE, mcpm.MODEL_FP_IDEAL])
def test_velocity_drift_keeps_pressure_and_sets_heat_flux_production(model):
    p, arrs = _cells_with_designed_velocities(model, seed=801)
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    o.fp_closures(0)
    cl = o.closures()
    parts, cells = _per_cell(o, p)
    Pr = 2.0 / 3.0
    for c, u in enumerate(parts):
        m = _moments(u)
        v, e = m["v"], m["e"]
        n = cells["n"][c]
        T = p.mass * m["DM2"] / (3.0 * p.kb)
        mu = p.mu * math.sqrt(T / p.T0)  # visp = 1/2
        tau = 2.0 * mu / (n * p.kb * T)
        Lam = -np.linalg.det(m["P"] - m["DM2"] / 3.0 * np.eye(3)) / (tau * m["DM2"] ** 4)
        assert abs(cl[c, 0] - Lam) <= 1e-10 * abs(Lam), c
        assert abs(cl[c, 27] - tau) <= 1e-12 * tau
        Y = orc.contact_y(p, n) if model == mcpm.MODEL_FP_DENSE else 1.0
        C, g, L = _sym(cl[c, 1:7]), cl[c, 7:10], cl[c, 0]
        A = v @ C.T + np.outer(e - m["DM2"], g) + L * (v * e[:, None] - m["Q"])
        S = (A.T @ v + v.T @ A) / len(v)
        mag = np.mean(np.abs(A)) * np.mean(np.abs(v))
        assert mag > 1e-3 and np.abs(S).max() <= 1e-9 * mag, (c, S)
        H = (A * e[:, None]).mean(axis=0) + 2.0 * (v * (A * v).sum(axis=1)[:, None]).mean(axis=0)
        target = Y * (3.0 - 2.0 * Pr) / tau * m["Q"]
        assert np.abs(target).max() > 0.01 * mag
        assert np.abs(H - target).max() <= 1e-9 * np.abs(target).max(), (c, H, target)


def test_position_drift_matches_the_dense_gas_corrections():
    p, arrs = _cells_with_designed_velocities(mcpm.MODEL_FP_DENSE, seed=802)
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    o.fp_closures(1)  # alpha and beta after step 0
    cl = o.closures()
    parts, cells = _per_cell(o, p)
    mom = [_moments(u) for u in parts]
    nc = p.n_cells
    for c, m in enumerate(mom):
        v, e = m["v"], m["e"]
        n = cells["n"][c]
        T = p.mass * m["DM2"] / (3.0 * p.kb)
        Y = orc.contact_y(p, n)
        nbY = n * p.b * Y
        p_crr, mu_crr, k_crr = 1.0 + nbY, 1.0 + 2.0 * nbY / 5.0, 1.0 + 3.0 * nbY / 5.0
        w = (4.0 / 9.0) * math.sqrt(math.pi) * Y * n ** 2 * p.sigma ** 4 * math.sqrt(p.mass * p.kb * T)
        coeff_mu = 3.0 * 1.002 / 5.0 * nbY ** 2 / (1.0 + 2.0 * nbY / 5.0)
        coeff_k = 1.002 * nbY ** 2 / (1.0 + 3.0 * nbY / 5.0) / 5.0
        gst = -0.001 * (n * p.b * Y / (p.kb * T / p.mass))
        if c == 0:
            divU = (mom[1]["U"][1] - m["U"][1]) / p.dy
        elif c == nc - 1:
            divU = (m["U"][1] - mom[c - 1]["U"][1]) / p.dy
        else:
            divU = (mom[c + 1]["U"][1] - mom[c - 1]["U"][1]) / p.dy / 2.0
        assert abs(cl[c, 26] - divU) <= 1e-10 * max(abs(divU), 1e-3)
        assert abs(cl[c, 22] - gst) <= 1e-12 * abs(gst)
        P, Q, tr = m["P"], m["Q"], np.trace(m["P"])
        dev = P - tr / 3.0 * np.eye(3)
        alpha = cl[c, 10:19].reshape(3, 3)
        beta = cl[c, 19:22]
        Av = v @ alpha.T  # (alpha_i . v)
        Lmom = (Av.T @ v) / len(v) + np.outer(beta * Q, np.ones(3))  # [i, j]
        rhs = (mu_crr - 1.0 + coeff_mu) * dev + np.eye(3) * ((p_crr - 1.0) * tr / 3.0
                                                                - w * divU / (p.mass * n))
        rhs = rhs - gst * m["M4"]
        scale = np.abs(rhs).max()
        assert scale > 1e-3
        assert np.abs(Lmom - rhs).max() <= 1e-9 * scale, (c, Lmom, rhs)
        Le = ((e - m["DM2"])[:, None] * Av).mean(axis=0) + (m["DM4"] - m["DM2"] ** 2) * beta
        rhs_e = (k_crr - 1.0 + coeff_k) * Q - gst * m["M5"] + gst * m["DM2"] * Q
        assert np.abs(Le - rhs_e).max() <= 1e-9 * np.abs(rhs_e).max(), (c, Le, rhs_e)
        # phi = alpha . Mp (Main_Functions.cpp:2734-2746) with Mp the mean
        # peculiar velocity, zero up to rounding
        vs = math.sqrt(m["DM2"])
        assert np.abs(cl[c, 23:26]).max() <= 1e-12 * np.abs(alpha).max() * vs


def test_position_drift_is_zero_before_the_first_step():
    p, arrs = _cells_with_designed_velocities(mcpm.MODEL_FP_DENSE, seed=803)
    o = orc.OracleSim(p)
    o.upload(*arrs)
    o.cell_update()
    o.fp_closures(0)
    cl = o.closures()
    assert np.all(cl[:, 10:22] == 0.0)  # alpha_calculation runs for step > 0 only
