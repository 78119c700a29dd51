"""The random-number and elementary-function contract shared by the HIP
kernels and the oracle (include/mcpm_rng.h), checked on the CPU.

* Philox4x32-10 against the known-answer vectors published with the
  Random123 library (Salmon et al., SC'11, kat_vectors: philox4x32 R=10).
* The deterministic log / exp / sincos(2 pi u) against numpy to a few ulp.
* The candidate priority is a bijection (no two candidates share one).
"""
import math

import numpy as np
import pytest

import oracle as orc


@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, expect):
    out = orc.philox(ctr, key)
    assert [int(v) for v in out] == list(expect)


def test_det_log_matches_numpy():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.random(20000), rng.random(2000) * 1e-9 + 1e-300,
                         [2.0 ** -32, 0.5, 1.0, 1.0 - 2 ** -53, 1e-300, 123.456]])
    xs = xs[xs > 0]
    got = np.array([orc.det_log(x) for x in xs])
    ref = np.log(xs)
    tol = 4 * np.finfo(float).eps * np.maximum(1.0, np.abs(ref))
    assert np.all(np.abs(got - ref) <= tol)
    assert orc.det_log(1.0) == 0.0


def test_det_exp_matches_numpy():
    rng = np.random.default_rng(4)
    xs = np.concatenate([rng.uniform(-30, 30, 20000), [0.0, -1e-10, 1e-10, -700.0, 700.0]])
    got = np.array([orc.det_exp(x) for x in xs])
    ref = np.exp(xs)
    assert np.all(np.abs(got - ref) <= 4 * np.finfo(float).eps * ref)
    assert orc.det_exp(0.0) == 1.0


def test_det_sincos_matches_numpy():
    rng = np.random.default_rng(5)
    us = np.concatenate([rng.random(20000), [0.0, 0.125, 0.25, 0.375, 0.5, 0.75, 1 - 2 ** -40]])
    s = np.empty_like(us)
    c = np.empty_like(us)
    for k, u in enumerate(us):
        s[k], c[k] = orc.det_sincos2pi(u)
    # reference in x87 extended precision: 2*pi*u rounded to double would
    # itself be off by up to 1.4e-15 near u = 1
    two_pi = 2 * np.longdouble("3.14159265358979323846264338327950288")
    arg = two_pi * us.astype(np.longdouble)
    assert np.all(np.abs(s - np.sin(arg)) <= 2.5e-16)
    assert np.all(np.abs(c - np.cos(arg)) <= 2.5e-16)
    assert orc.det_sincos2pi(0.0) == (0.0, 1.0)


def test_priority_is_a_bijection():
    key = 0x9E3779B9
    vals = np.array([orc.priority(k, key) for k in range(200000)], dtype=np.uint64)
    assert len(np.unique(vals)) == len(vals)
    # and looks random: about half the values above 2^31
    frac = np.mean(vals >= 2 ** 31)
    assert abs(frac - 0.5) < 0.01


def test_philox_uniformity():
    words = np.concatenate([orc.philox((k, 7, 0, 0x02000000), (1, 0)) for k in range(5000)])
    u = (words.astype(np.float64) + 0.5) * 2.0 ** -32
    assert 0.0 < u.min() and u.max() < 1.0
    # chi-square on 16 bins
    h, _ = np.histogram(u, bins=16, range=(0, 1))
    e = len(u) / 16
    chi2 = ((h - e) ** 2 / e).sum()
    assert chi2 < 40.0  # 15 dof, p ~ 5e-4
