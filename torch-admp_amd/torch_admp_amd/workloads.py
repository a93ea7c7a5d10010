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


def qeq_species_box(n: int = 3000, length: float = 30.0, seed: int = 0):
    """n atoms of 3 species (equal counts) uniform in a cubic box, per-species
    chi ~ N(5, 1) eV, J ~ U(6, 10) eV and Gaussian width ~ U(0.8, 1.2) A.
    Returns positions, box, chi, hardness (= J / 2, the coefficient of q^2 in
    the QEq site energy), eta."""
    rng = np.random.default_rng(seed)
    species = np.repeat(np.arange(3), -(-n // 3))[:n]
    rng.shuffle(species)
    chi_s = rng.normal(5.0, 1.0, 3)
    j_s = rng.uniform(6.0, 10.0, 3)
    eta_s = rng.uniform(0.8, 1.2, 3)
    pos = rng.random((n, 3)) * length
    return pos, np.eye(3) * length, chi_s[species], j_s[species] / 2.0, eta_s[species]


def electrode_cell(seed: int = 0, n_ions: int = 2000, lx: float = 40.0, ly: float = 40.0,
                   lz: float = 120.0, a: float = 4.078, min_sep: float = 2.0):
    """Two fcc(111) gold-like slabs of 4 layers x 200 atoms (a 10 x 10 grid of
    the 2-atom rectangular cell of the (111) plane, strained to lx x ly) facing
    each other across a gap filled with n_ions random +-1 ions (at least
    min_sep apart); the cell is lz tall, about twice the system height, for the
    slab correction.  Returns positions (1600 electrode atoms first: left slab
    then right slab), charges (0 on the electrodes) and the box."""
    rng = np.random.default_rng(seed)
    d = a / math.sqrt(2.0)               # in-plane nearest-neighbour distance
    dz = a / math.sqrt(3.0)              # (111) layer spacing
    cx, cy = 10, 10
    sx, sy = lx / (cx * d), ly / (cy * d * math.sqrt(3.0))
    base = np.array([[0.0, 0.0], [0.5 * d, 0.5 * d * math.sqrt(3.0)]])
    cells = np.stack(np.meshgrid(np.arange(cx), np.arange(cy), indexing="ij"), -1).reshape(-1, 2)
    plane = (cells[:, None, :] * np.array([d, d * math.sqrt(3.0)]) + base[None]).reshape(-1, 2)
    shift = [np.array([0.0, 0.0]), np.array([0.5 * d, d / (2.0 * math.sqrt(3.0))]),
             np.array([d, d / math.sqrt(3.0)])]  # A B C stacking

    def slab(z0, direction):
        layers = []
        for k in range(4):
            xy = (plane + shift[k % 3]) * np.array([sx, sy])
            xy = np.mod(xy, [lx, ly])
            z = np.full((len(xy), 1), z0 + direction * k * dz)
            layers.append(np.hstack([xy, z]))
        return np.concatenate(layers)

    height = lz / 2.0
    left = slab(1.0, 1.0)
    right = slab(height - 1.0, -1.0)
    z_lo, z_hi = 1.0 + 3 * dz + 3.0, height - 1.0 - 3 * dz - 3.0
    ions = np.empty((0, 3))
    while len(ions) < n_ions:
        cand = rng.random((4 * n_ions, 3)) * np.array([lx, ly, z_hi - z_lo]) + np.array([0, 0, z_lo])
        for p in cand:
            if len(ions) == n_ions:
                break
            if len(ions):
                dd = ions - p
                dd[:, :2] -= np.round(dd[:, :2] / [lx, ly]) * [lx, ly]
                if (dd * dd).sum(1).min() < min_sep * min_sep:
                    continue
            ions = np.vstack([ions, p])
    q_ions = np.r_[np.ones(n_ions // 2), -np.ones(n_ions - n_ions // 2)]
    rng.shuffle(q_ions)
    pos = np.concatenate([left, right, ions])
    q = np.r_[np.zeros(len(left) + len(right)), q_ions]
    return pos, q, np.diag([lx, ly, lz])
