"""Direct Ewald summation in numpy (test oracle for the mesh restatement).

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

This is textbook Ewald (no mesh, no B-splines): it pins
``pme_ref.reciprocal_energy_frame`` (src/torch_admp/pme.py:351-446) up to the
SPME interpolation error, and the split into real + reciprocal + self +
non-neutral terms through the kappa-independence of the total.

    E_real  = eps/2 sum_{i,j,n}' q_i q_j erfc(kappa r)/r
    E_recip = eps (2 pi / V) sum_{m != 0} exp(-k^2 / 4 kappa^2) / k^2 |sum_j q_j e^{2 pi i m.s_j}|^2
    E_self  = -eps kappa / sqrt(pi) sum q^2
    E_nn    = -eps pi Q^2 / (2 V kappa^2)

with k = 2 pi A^T m ("columns", the physical reciprocal vectors) or
k = 2 pi m A ("reference", the rows convention of src/torch_admp/recip.py:304).
"""

from __future__ import annotations

import itertools
import math

import numpy as np
from scipy.special import erfc

DIELECTRIC = 14.399645468667817


def ewald_reciprocal(pos, box, q, kappa, mmax, convention="columns"):
    pos = np.asarray(pos, float)
    box = np.asarray(box, float)
    q = np.asarray(q, float)
    A = np.linalg.inv(box)
    V = abs(np.linalg.det(box))
    s = pos @ A
    rng = [np.arange(-m, m + 1) for m in mmax]
    M = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 3)
    M = M[np.any(M != 0, axis=1)].astype(float)
    k = 2 * np.pi * (M @ A.T if convention == "columns" else M @ A)
    k2 = (k ** 2).sum(1)
    phase = 2 * np.pi * (M @ s.T)
    S = (q[None, :] * np.exp(1j * phase)).sum(1)
    return DIELECTRIC * (2 * np.pi / V * np.exp(-k2 / (4 * kappa ** 2)) / k2 * np.abs(S) ** 2).sum()


def ewald_real(pos, box, q, kappa, rmax):
    """Real-space Ewald sum over all images with r < rmax (both orderings, halved)."""
    pos = np.asarray(pos, float)
    box = np.asarray(box, float)
    q = np.asarray(q, float)
    V = abs(np.linalg.det(box))
    heights = [V / np.linalg.norm(np.cross(box[a], box[b])) for a, b in ((1, 2), (2, 0), (0, 1))]
    nmax = [int(math.ceil(rmax / h)) + 1 for h in heights]
    e = 0.0
    for n in itertools.product(*[range(-m, m + 1) for m in nmax]):
        sh = np.array(n, float) @ box
        d = np.linalg.norm(pos[None, :, :] + sh - pos[:, None, :], axis=-1)
        mask = (d < rmax) & (d > 1e-12)
        qq = q[:, None] * q[None, :]
        e += 0.5 * (qq[mask] * erfc(kappa * d[mask]) / d[mask]).sum()
    return DIELECTRIC * e


def ewald_total(pos, box, q, kappa, rmax, mmax, convention="columns"):
    q = np.asarray(q, float)
    V = abs(np.linalg.det(np.asarray(box, float)))
    e_self = -DIELECTRIC * kappa / math.sqrt(math.pi) * (q ** 2).sum()
    e_nn = -DIELECTRIC * math.pi / (2 * V * kappa ** 2) * q.sum() ** 2
    return (ewald_real(pos, box, q, kappa, rmax)
            + ewald_reciprocal(pos, box, q, kappa, mmax, convention) + e_self + e_nn)
