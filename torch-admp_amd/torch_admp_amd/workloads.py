"""Seeded synthetic inputs for the benchmark configurations (host-side numpy).

ionic_box      uniform +-1 ions with zero net charge in a cubic box
spc_water      rigid 3-site SPC-like water (O -0.8476 e, H +0.4238 e, OH 1.0 A,
               HOH 109.47 deg) with random orientations at 1 g/cm^3
"""

from __future__ import annotations

import math

import numpy as np

WATER_VOLUME_A3 = 18.01528 / 6.02214076e23 / 1.0 * 1e24  # A^3 per molecule at 1 g/cm^3


def ionic_box(n: int, length: float, seed: int = 0):
    """n positions uniform in [0, L)^3 and a shuffled half +1 / half -1 charge set."""
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * length
    q = np.r_[np.ones(n // 2), -np.ones(n - n // 2)]
    if n % 2:
        q[-1] = 0.0
    rng.shuffle(q)
    box = np.eye(3) * length
    return pos, q, box


def _random_rotations(rng, m):
    """m uniform random rotation matrices from unit quaternions."""
    u = rng.normal(size=(m, 4))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w, x, y, z = u.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def spc_water(n_mol: int = 8000, seed: int = 0, jitter: float = 0.3):
    """n_mol rigid waters on a jittered cubic lattice at liquid density.

    Returns positions (3 n_mol, 3) ordered O, H, H per molecule, charges and
    the cubic box; the box edge is (n_mol * 29.915 A^3)^(1/3) (62.09 A for 8000).
    """
    rng = np.random.default_rng(seed)
    length = (n_mol * WATER_VOLUME_A3) ** (1.0 / 3.0)
    side = int(math.ceil(n_mol ** (1.0 / 3.0) - 1e-9))
    a = length / side
    grid = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    grid = grid[rng.permutation(len(grid))[:n_mol]]
    centres = (grid + 0.5 + jitter * (rng.random((n_mol, 3)) - 0.5)) * a
    half = math.radians(109.47) / 2
    local = np.array([[0.0, 0.0, 0.0],
                      [math.sin(half), 0.0, math.cos(half)],
                      [-math.sin(half), 0.0, math.cos(half)]])  # OH = 1.0 A
    rot = _random_rotations(rng, n_mol)
    pos = centres[:, None, :] + np.einsum("mij,kj->mki", rot, local)
    q = np.tile([-0.8476, 0.4238, 0.4238], n_mol)
    return pos.reshape(-1, 3), q, np.eye(3) * length
