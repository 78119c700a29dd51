"""Synthetic workloads of the slab step: parameters plus initial particles.

Three slab set-ups, sized by arguments so that the same code builds the full
problem and the small versions the parity tests run:

* heat_conduction: Enskog DSMC Fourier heat conduction between diffuse walls
  (20 sigma slab, 200 cells, 100 particles per cell, n sigma^3 = 0.2, T0 = 1.5,
  walls 1.0 / 2.0, dt = 0.1 mean collision time, FP64).
* dense_slab: the same geometry refined (4000 cells, 2500 particles per cell,
  10M particles) at n sigma^3 = 0.1 or 0.3, walls 1.0 / 1.5, dt = 0.05 mean
  collision time, FP32 velocities.
* interface: liquid-vapour planar interface in a periodic 400 sigma box with
  4096 cells and 20M particles; n sigma^3 = 0.55 in the central third, 0.02
  elsewhere, T = 0.7, attractive tail -eps (sigma/r)^6 through the
  screened-Poisson mean field, dt = 0.02, FP32.
* plane: the 2D problems (the reference's `injection`, myIO.cpp:2729-2761):
  liquid discs in vapour on an Lx x Ly plane of unit depth, specular walls
  or periodic axes, direct or screened-Poisson attraction.  `lattice=True`
  places four particles per cell at the quarter points instead, an exactly
  uniform density for closed-form checks.

Units are reduced: sigma = m = kB = eps = 1.  Initial positions are uniform
in each density region and velocities Maxwellian (numpy Philox generator),
the reference's `initialization` (Main_Functions.cpp:4657-4801) in spirit.
"""
from __future__ import annotations

import math

import numpy as np

from mcpm import (BC_DIFFUSE, BC_PERIODIC, BC_SPECULAR, FP32, FP64, LR_DIRECT, LR_NONE,
                  LR_SCREENED_POISSON, MODEL_ESMC, Params, Y_REFERENCE)

# Yukawa fit of the reference (myIO.cpp:2891-2892) in reduced units:
# bst = 6.91304716e9 1/m * 3.405e-10 m, ast = -1.64835851e-28 J m / (eps sigma)
# with eps = 119.8 kB, sigma = 3.405e-10 m.  It fits -4 eps (sigma/r)^6; the
# interface problem's tail -eps (sigma/r)^6 scales ast by 1/4.
YUKAWA_BST = 6.91304716e9 * 3.405e-10
YUKAWA_AST_4EPS = -1.64835851e-28 / (119.8 * 1.380649e-23 * 3.405e-10)


def contact_y_ref(n: float, sigma: float = 1.0) -> float:
    b = 2.0 * math.pi * sigma ** 3 / 3.0
    nb = n * b
    return 1.0 + 0.625 * n * b + 0.2869 * nb * nb + 0.115 * nb * nb * nb


def mean_collision_time(n: float, T: float, sigma: float = 1.0, m: float = 1.0) -> float:
    """1 / nu with the Enskog frequency nu = 4 n sigma^2 Y(n) sqrt(pi kB T / m)."""
    return 1.0 / (4.0 * n * sigma ** 2 * contact_y_ref(n, sigma) * math.sqrt(math.pi * T / m))


def maxwellian(rng, n, T, m=1.0, dtype=np.float64):
    s = math.sqrt(T / m)
    return (rng.standard_normal(n) * s, rng.standard_normal(n) * s, rng.standard_normal(n) * s)


def heat_conduction(n_cells=200, ppc=100, length=20.0, n_density=0.2, T0=1.5, T_lo=1.0,
                    T_hi=2.0, dt_frac=0.1, precision=FP64, seed=1, walls_at_sigma=False):
    N = n_cells * ppc
    area = N / (n_density * length)
    dy = length / n_cells
    eps = 1e-15
    lo, hi = (1.0 + eps, length - 1.0 - eps) if walls_at_sigma else (eps, length - eps)
    p = Params(n_cells=n_cells, length=length, capacity=N, precision=precision,
               boundary=BC_DIFFUSE, y_model=Y_REFERENCE, long_range=LR_NONE, seed=seed,
               wall_lo=lo, wall_hi=hi, T_wall_lo=T_lo, T_wall_hi=T_hi,
               cell_volume=dy * area, Fn=1.0,
               delta_t=dt_frac * mean_collision_time(n_density, T0))
    rng = np.random.Generator(np.random.Philox(seed))
    y = lo + (hi - lo) * rng.random(N)
    u1, u2, u3 = maxwellian(rng, N, T0)
    return p, (y, u1, u2, u3), {"n_gas": n_density, "T0": T0, "N": N, "area": area}


def dense_slab(n_density=0.3, n_cells=4000, ppc=2500, length=20.0, T0=1.25, T_lo=1.0,
               T_hi=1.5, dt_frac=0.05, precision=FP32, seed=1):
    p, arrs, meta = heat_conduction(n_cells=n_cells, ppc=ppc, length=length, n_density=n_density,
                                    T0=T0, T_lo=T_lo, T_hi=T_hi, dt_frac=dt_frac,
                                    precision=precision, seed=seed)
    return p, arrs, meta


def interface(N=20_000_000, n_cells=4096, length=400.0, n_liq=0.55, n_vap=0.02, T=0.7,
              dt=0.02, precision=FP32, seed=1, capacity=None):
    L3 = length / 3.0
    area = N / (n_liq * L3 + n_vap * 2.0 * L3)
    dy = length / n_cells
    n_l = int(round(n_liq * L3 * area))
    n_v = N - n_l
    rng = np.random.Generator(np.random.Philox(seed))
    yl = L3 + L3 * rng.random(n_l)
    yv = rng.random(n_v) * (2.0 * L3)
    yv = np.where(yv < L3, yv, yv + L3)
    y = np.concatenate([yl, yv])
    u1, u2, u3 = maxwellian(rng, N, T)
    p = Params(n_cells=n_cells, length=length, capacity=capacity or N, precision=precision,
               boundary=BC_PERIODIC, y_model=Y_REFERENCE, long_range=LR_SCREENED_POISSON,
               seed=seed, cell_volume=dy * area, Fn=1.0, delta_t=dt,
               ast=YUKAWA_AST_4EPS / 4.0, bst=YUKAWA_BST, eqdist=1.0)
    n_gas = N / (length * area)
    return p, (y, u1, u2, u3), {"n_gas": n_gas, "T0": T, "N": N, "area": area}


def replicas(n_rep=256, seed0=0, N_lo=50_000, N_hi=500_000, nliq_lo=0.45, nliq_hi=0.6,
             T_lo=0.6, T_hi=0.9, cells_lo=512, cells_hi=2048, n_vap=0.02, length=400.0,
             dt=0.02, precision=FP32, boundary=BC_PERIODIC, long_range=LR_SCREENED_POISSON):
    """Independent interface slabs for one batched context: replica r has
    seed seed0 + r, N log-uniform in [N_lo, N_hi], liquid density uniform in
    [nliq_lo, nliq_hi], T uniform in [T_lo, T_hi], cells uniform in
    {cells_lo..cells_hi}.  Returns (common params, replica dicts, per-replica
    arrays, per-replica meta)."""
    rng = np.random.Generator(np.random.Philox(10_000 + seed0))
    reps, arrs, metas = [], [], []
    common = None
    for r in range(n_rep):
        N = int(round(math.exp(rng.uniform(math.log(N_lo), math.log(N_hi)))))
        n_liq = rng.uniform(nliq_lo, nliq_hi)
        T = rng.uniform(T_lo, T_hi)
        nc = int(rng.integers(cells_lo, cells_hi + 1))
        p, a, meta = interface(N=N, n_cells=nc, length=length, n_liq=n_liq, n_vap=n_vap, T=T,
                               dt=dt, precision=precision, seed=seed0 + r)
        p.boundary = boundary
        p.long_range = long_range
        if common is None:
            common = p
        reps.append({"n_cells": nc, "capacity": N, "seed": seed0 + r,
                     "cell_volume": p.cell_volume, "Fn": p.Fn, "params": p})
        arrs.append(a)
        metas.append(meta)
    return common, reps, arrs, metas


def plane(nx=40, ny=40, Lx=40.0, Ly=40.0, ppc=20, n_vap=0.1, n_liq=0.55, drops=((20.0, 20.0, 10.0),),
          T=0.835, dt=0.01, precision=FP64, seed=1, boundary=BC_SPECULAR, long_range=LR_DIRECT,
          model=MODEL_ESMC, lattice=False, capacity=None, phi=-4.0):
    """Droplets of density n_liq in vapour n_vap; `drops` lists (x, y, radius)."""
    rng = np.random.Generator(np.random.Philox(seed))
    dx, dy = Lx / nx, Ly / ny
    if lattice:
        j, k = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
        xs, ys = [], []
        for fx in (0.25, 0.75):
            for fy in (0.25, 0.75):
                xs.append(((j + fx) * dx).ravel())
                ys.append(((k + fy) * dy).ravel())
        x = np.concatenate(xs)
        y = np.concatenate(ys)
        n_bg = x.size
        Fn = n_vap * Lx * Ly * 1.0 / n_bg
    else:
        n_bg = ppc * nx * ny
        Fn = n_vap * Lx * Ly * 1.0 / n_bg           # myIO.cpp:2745
        x = rng.random(n_bg) * Lx
        y = rng.random(n_bg) * Ly
        for (cx, cy, r) in drops:
            nd = int(math.floor((n_liq - n_vap) * math.pi * r * r / Fn))   # myIO.cpp:2748
            rad = r * np.sqrt(rng.random(nd))
            ang = 2.0 * math.pi * rng.random(nd)
            xd = np.clip(cx + rad * np.cos(ang), 0.0, np.nextafter(Lx, 0))
            yd = np.clip(cy + rad * np.sin(ang), 0.0, np.nextafter(Ly, 0))
            x = np.concatenate([x, xd])
            y = np.concatenate([y, yd])
    N = x.size
    u1, u2, u3 = maxwellian(rng, N, T)
    p = Params(n_cells=ny, length=Ly, capacity=capacity or N, precision=precision,
               boundary=boundary, y_model=Y_REFERENCE, long_range=long_range, model=model,
               seed=seed, Fn=Fn, delta_t=dt, ast=YUKAWA_AST_4EPS / 4.0, bst=YUKAWA_BST,
               eqdist=1.0, phi=phi, dims=2, n_cells_x=nx, length_x=Lx, boundary_x=boundary)
    n_gas = N * Fn / (Lx * Ly)
    return p, (x, y, u1, u2, u3), {"n_gas": n_gas, "T0": T, "N": N, "Fn": Fn}


def droplets(nx=512, ny=512, Lx=128.0, Ly=128.0, N=64_000_000, n_drops=20, r_lo=5.0, r_hi=15.0,
             n_liq=0.55, n_vap=0.02, T=0.75, dt=0.01, precision=FP32, seed=1,
             long_range=LR_SCREENED_POISSON, capacity=None):
    """Periodic plane with random liquid discs: n_drops discs, radii uniform in
    [r_lo, r_hi], at n_liq in vapour n_vap, about N particles in all.  Discs
    are placed one by one at random centres that keep a gap of at least one
    sigma to every disc placed before (periodic distance), so no region holds
    the liquid density twice."""
    rng = np.random.Generator(np.random.Philox(seed + 1000))
    drops = []
    tries = 0
    while len(drops) < n_drops:
        tries += 1
        if tries > 100_000:
            raise ValueError("cannot place the discs without overlap")
        cx, cy, r = rng.random() * Lx, rng.random() * Ly, r_lo + (r_hi - r_lo) * rng.random()
        ok = True
        for (qx, qy, qr) in drops:
            dx = abs(cx - qx)
            dy = abs(cy - qy)
            dx, dy = min(dx, Lx - dx), min(dy, Ly - dy)
            if math.hypot(dx, dy) < r + qr + 1.0:
                ok = False
                break
        if ok:
            drops.append((cx, cy, r))
    # liquid fraction of the particles from the disc areas
    a_liq = sum(math.pi * r * r for (_, _, r) in drops)
    mass = n_vap * Lx * Ly + (n_liq - n_vap) * a_liq
    ppc = max(1, int(round(N * (n_vap * Lx * Ly / mass) / (nx * ny))))
    p, arrs, meta = plane(nx=nx, ny=ny, Lx=Lx, Ly=Ly, ppc=ppc, n_vap=n_vap, n_liq=n_liq,
                          drops=[], T=T, dt=dt, precision=precision, seed=seed,
                          boundary=BC_PERIODIC, long_range=long_range, capacity=capacity)
    x, y, u1, u2, u3 = arrs
    rng2 = np.random.Generator(np.random.Philox(seed + 2000))
    xs, ys = [x], [y]
    for (cx, cy, r) in drops:
        nd = int(math.floor((n_liq - n_vap) * math.pi * r * r / p.Fn))
        rad = r * np.sqrt(rng2.random(nd))
        ang = 2.0 * math.pi * rng2.random(nd)
        xs.append(np.mod(cx + rad * np.cos(ang), Lx))
        ys.append(np.mod(cy + rad * np.sin(ang), Ly))
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    x = np.where(x < Lx, x, 0.0)
    y = np.where(y < Ly, y, 0.0)
    Nt = x.size
    v1, v2, v3 = maxwellian(rng2, Nt, T)
    if capacity is None or capacity < Nt:
        p.capacity = Nt
    meta.update({"N": Nt, "n_gas": Nt * p.Fn / (Lx * Ly), "drops": drops})
    return p, (x, y, v1, v2, v3), meta
