"""Pins of the oracle's composed operations and diagnostics (oracle/pn_oracle.c): the ring
averages, and the parts that only a whole sweep or step exercises, each checked against
what the mathematics fixes -- invariants, conservation, exact solutions, brute force --
chosen so that a plausible mistake in the restatement fails one of them.
"""
from __future__ import annotations

import math

import numpy as np

import oracle
from penguin_2_0_amd import configs

TWO_PI = 2.0 * math.pi


def nonuniform_phi(o: oracle.OracleDisk, seed: int = 3) -> np.ndarray:
    """Give every ring of o the same stretched azimuthal grid (sum of dphi = 2 pi) and a
    random roll; returns dphi."""
    J = o.grid.jmax
    rng = np.random.default_rng(seed)
    w = 1.0 + 0.4 * np.sin(np.arange(J) * TWO_PI / J * 3.0) + 0.2 * rng.random(J)
    dy = w / w.sum() * TWO_PI
    y = np.concatenate([[0.0], np.cumsum(dy)[:-1]])
    for R in (o.rings,):
        R["dy"][:] = dy
        R["y"][:] = y
        R["yc"][:] = y + 0.5 * dy
        R["yvol"][:] = dy
        R["rot_j"][:] = rng.integers(0, J, len(R))
    return dy


def test_ring_means_of_axisymmetric_fields_are_the_ring_values():
    """A field constant along every ring averages to that constant on any azimuthal grid and
    any roll (a wrong divisor -- the column count instead of the sum of dphi -- fails here)."""
    w = configs.small(24, 60)
    o = oracle.OracleDisk(w.grid, w.params, 1, 0.0)
    nonuniform_phi(o)
    rng = np.random.default_rng(5)
    for f in ("r", "p", "u", "v", "w"):
        c = rng.uniform(-2.0, 3.0, len(o.rings))
        o.rings[f][:] = c[:, None]
        got = o.ring_means(f)
        np.testing.assert_allclose(got, c, rtol=4e-16 * w.grid.jmax, atol=0)


def test_ring_means_brute_force_in_the_lab_frame():
    """Against the dphi-weighted average of the lab-frame field (data column j sits at lab
    column j + rot_j), on a stretched grid where a mix-up of data and lab columns changes the
    weights."""
    w = configs.small(24, 60)
    o = oracle.OracleDisk(w.grid, w.params, 1, 0.0)
    dy = nonuniform_phi(o)
    rng = np.random.default_rng(11)
    for k, f in enumerate(("r", "p", "u", "v", "w")):
        o.rings[f][:] = rng.normal(size=o.rings[f].shape) + k
        lab = np.stack([np.roll(o.rings[f][i], int(o.rings["rot_j"][i])) for i in range(len(o.rings))])
        want = (lab * dy).sum(axis=1) / dy.sum()
        got = o.ring_means(f)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)
        # the data-frame weighting (no roll) is a different number on this grid
        wrong = (o.rings[f] * dy).sum(axis=1) / dy.sum()
        assert np.abs(wrong - want).max() > 1e-3


def test_ring_means_of_a_pure_harmonic_vanish():
    """On a uniform full ring, cos(m phi) for 1 <= m < jmax averages to zero (the discrete
    orthogonality of the harmonics), and a constant plus it to the constant."""
    w = configs.small(8, 64)
    o = oracle.OracleDisk(w.grid, w.params, 1, 0.0)
    J = w.grid.jmax
    yc = o.rings["yc"][0]
    o.rings["rot_j"][:] = np.arange(len(o.rings)) * 7 % J
    for m in (1, 5, 31):
        o.rings["v"][:] = 0.25 + np.cos(m * yc)[None, :]
        np.testing.assert_allclose(o.ring_means("v"), 0.25, rtol=0, atol=1e-15)
