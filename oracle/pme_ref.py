"""Float64 CPU restatement of the reference PME Coulomb path (test oracle).

TEST INFRASTRUCTURE ONLY -- imported by tests/, smoke() and bench.py's
cpu_baseline leg, never by the product package.

Every function follows one reference location (paths relative to the
reference root, torch-admp 1.1.5):

=====================  ==========================================  ===========================
function               reference                                   pinned by (tests/)
=====================  ==========================================  ===========================
bspline6               src/torch_admp/recip.py:102-143             closed forms: knot values
                                                                   1,26,66,26,1 /120, partition
                                                                   of unity, symmetry, Cox-de
                                                                   Boor recursion
spread_charges         src/torch_admp/recip.py:13-64,67-99,        brute-force per-atom
                       146-252                                     stencil sum; total charge
kpts_integer           src/torch_admp/recip.py:255-285             numpy.fft.fftfreq ordering
kpts                   src/torch_admp/recip.py:288-308             closed form (rows of A)
theta_k                src/torch_admp/pme.py:410-432               |b(k)|^-1 of Essmann 1995
reciprocal_energy      src/torch_admp/pme.py:351-446,661-683       direct Ewald k-sum
                                                                   (oracle/ewald_ref.py),
                                                                   kappa-independence
real_energy            src/torch_admp/pme.py:312-349               LAMMPS coul/long real part
                                                                   via total, brute force
self_energy            src/torch_admp/pme.py:448-461               closed form, Ewald total
non_neutral_energy     src/torch_admp/pme.py:463-493               Ewald total, charged system
slab_corr_energy       src/torch_admp/pme.py:495-527               LAMMPS slab fixture
obc_energy             src/torch_admp/pme.py:529-561               LAMMPS coul/cut fixture
kappa_from_ethresh     src/torch_admp/pme.py:113-115               closed form
kmesh_rule             src/torch_admp/pme.py:383-392               closed form
setup_ewald_parameters src/torch_admp/pme.py:564-658               reference test values
=====================  ==========================================  ===========================

Forces and box gradients come from torch autograd exactly as in the reference
(src/torch_admp/utils.py:84-112); they are pinned by central finite differences.
"""

from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np
import torch

# 1/(4 pi eps0) in eV*Angstrom/e^2 from scipy CODATA, as src/torch_admp/utils.py:268-278
DIELECTRIC = 14.399645468667817
SQRT_PI = math.sqrt(math.pi)
PME_ORDER = 6  # src/torch_admp/pme.py:167-169 (order 6 only)


def kappa_from_ethresh(rcut: float, ethresh: float) -> float:
    """kappa = sqrt(-ln(2 ethresh)) / rcut  (src/torch_admp/pme.py:113-115)."""
    return math.sqrt(-math.log(2.0 * ethresh)) / rcut


def kmesh_rule(box: torch.Tensor, kappa: float, ethresh: float,
               spacing: Optional[Sequence[float]] = None) -> list:
    """Mesh size rule of src/torch_admp/pme.py:383-392 (uses the box diagonal)."""
    diag = torch.diagonal(box.detach().reshape(3, 3)).double()
    if spacing is not None:
        sp = torch.as_tensor(spacing, dtype=torch.float64)
        k = torch.ceil(diag / sp)
    else:
        k = torch.ceil(2 * kappa * diag / (3.0 * ethresh ** (1.0 / 5.0)))
    return [int(v) for v in k.tolist()]


def bspline6(u: torch.Tensor) -> torch.Tensor:
    """Order-6 cardinal B-spline by pieces (src/torch_admp/recip.py:102-143).

    Same piece boundaries as the reference: the first piece requires u >= 0,
    the later ones only test the upper bound.
    """
    p0 = u ** 5 / 120
    p1 = p0 - (u - 1) ** 5 / 20
    p2 = p1 + (u - 2) ** 5 / 8
    p3 = p2 - (u - 3) ** 5 / 6
    p4 = u ** 5 / 24 - u ** 4 + 19 * u ** 3 / 2 - 89 * u ** 2 / 2 + 409 * u / 4 - 1829 / 20
    p5 = -(u ** 5) / 120 + u ** 4 / 4 - 3 * u ** 3 + 18 * u ** 2 - 54 * u + 324 / 5
    zero = torch.zeros_like(u)
    out = torch.where(u < 6.0, p5, zero)
    out = torch.where(u < 5.0, p4, out)
    out = torch.where(u < 4.0, p3, out)
    out = torch.where(u < 3.0, p2, out)
    out = torch.where(u < 2.0, p1, out)
    out = torch.where((u >= 0.0) & (u < 1.0), p0, out)
    return out


def stencil_shifts(dtype=torch.long) -> torch.Tensor:
    """216 integer stencil offsets in the reference order (src/torch_admp/pme.py:173-183).

    The reference builds them with meshgrid(ij) -> stack(x=s1, y=s0, z=s2)
    -> transpose(0, 3) -> reshape; the resulting row order only affects the
    summation order of the scatter, not its value.
    """
    r = torch.arange(-PME_ORDER // 2, PME_ORDER // 2)
    s0, s1, s2 = torch.meshgrid(r, r, r, indexing="ij")
    return torch.stack((s1, s0, s2)).transpose(0, 3).reshape(1, PME_ORDER ** 3, 3).to(dtype)


def spread_charges(pos: torch.Tensor, box_inv: torch.Tensor, q: torch.Tensor,
                   kmesh: Sequence[int]) -> torch.Tensor:
    """Order-6 charge spreading onto a K1 x K2 x K3 mesh (src/torch_admp/recip.py:216-252).

    R = K * (r @ box_inv); m0 = ceil(R); u0 = m0 - R + 3;
    Q[(m0 + s) mod K] += q * prod_d M6(u0_d + s_d) for the 216 offsets s in {-3..2}^3.
    """
    K = torch.tensor(list(kmesh), dtype=torch.long)
    R = (pos @ box_inv) * K.to(pos.dtype)            # recip.py:33, 61
    m0 = torch.ceil(R).to(torch.long)                # recip.py:62
    u0 = (m0 - R) + PME_ORDER / 2                    # recip.py:63
    sh = stencil_shifts()
    u = u0[:, None, :] + sh.to(pos.dtype)            # recip.py:94
    theta = torch.prod(bspline6(u), dim=-1)          # recip.py:97-98
    w = q.reshape(-1, 1) * theta                     # recip.py:168-169
    idx = torch.remainder(m0[:, None, :] + sh, K)    # recip.py:190-192 (fmod of shifted value)
    flat = (idx[..., 0] * K[1] + idx[..., 1]) * K[2] + idx[..., 2]
    mesh = torch.zeros(int(K.prod()), dtype=pos.dtype)
    mesh = mesh.index_add(0, flat.reshape(-1), w.reshape(-1))
    return mesh.reshape(int(K[0]), int(K[1]), int(K[2]))


def kpts_integer(kmesh: Sequence[int]) -> torch.Tensor:
    """Integer k points in FFT order, Nyquist negative (src/torch_admp/recip.py:255-285)."""
    axes = []
    for k in kmesh:
        k = int(k)
        a = torch.arange(-(k - 1) // 2, (k + 1) // 2)
        axes.append(torch.roll(a, shifts=[-(k - 1) // 2]))
    g = torch.meshgrid(*axes, indexing="ij")
    return torch.stack([x.reshape(-1) for x in g], dim=1)


def kpts(box_inv: torch.Tensor, kint: torch.Tensor, convention: str = "reference") -> torch.Tensor:
    """k vectors and k^2 (src/torch_admp/recip.py:288-308).

    ``reference``: k = 2 pi m @ box_inv  (rows of box_inv, recip.py:304).
    ``columns``:   k = 2 pi box_inv @ m  (the reciprocal vectors proper).
    They agree for orthorhombic and symmetric boxes.
    """
    m = kint.to(box_inv.dtype)
    if convention == "reference":
        k = 2 * math.pi * (m @ box_inv)
    elif convention == "columns":
        k = 2 * math.pi * (m @ box_inv.transpose(0, 1))
    else:
        raise ValueError(f"unknown reciprocal convention {convention!r}")
    return torch.cat((k, (k ** 2).sum(1, keepdim=True)), dim=1).transpose(0, 1)


def theta_k(kint: torch.Tensor, kmesh: Sequence[int], dtype=torch.float64) -> torch.Tensor:
    """B-spline structure-factor modulus prod_d sum_{m=-2..2} M6(m+3) cos(2 pi m k_d/K_d)
    (src/torch_admp/pme.py:412-432)."""
    m = torch.linspace(-PME_ORDER // 2 + 1, PME_ORDER // 2 - 1, PME_ORDER - 1,
                       dtype=dtype).reshape(-1, 1, 1)
    Kf = torch.tensor(list(kmesh), dtype=dtype).reshape(1, 1, 3)
    terms = bspline6(m + PME_ORDER / 2) * torch.cos(2 * math.pi * m * kint[None].to(dtype) / Kf)
    return torch.prod(terms.sum(0), dim=1)


def reciprocal_energy_frame(pos, box, q, kappa, kmesh, convention="reference"):
    """One frame of the SPME reciprocal energy (src/torch_admp/pme.py:378-443, 661-683).

    E = eps * sum_{k != 0} (2 pi / V) exp(-k^2/4 kappa^2)/k^2 * |FFT(Q)(k)|^2 / theta(k)^2.
    Differentiable in pos, box and q through torch autograd.
    """
    box_inv = torch.linalg.inv(box)
    volume = torch.det(box)
    Q = spread_charges(pos, box_inv, q, kmesh)
    kint = kpts_integer(kmesh)
    kv = kpts(box_inv, kint, convention)
    th = theta_k(kint, kmesh, pos.dtype)
    S = torch.fft.fftn(Q).reshape(-1)
    ksq = kv[3, 1:]
    coeff = 2 * math.pi / volume / ksq * torch.exp(-ksq / 4 / kappa ** 2)
    E = coeff * ((S[1:].real ** 2 + S[1:].imag ** 2) / th[1:] ** 2)
    return E.sum() * DIELECTRIC


def safe_inverse(x: torch.Tensor, threshold: float = 1e-8) -> torch.Tensor:
    """1/x where |x| > threshold else 0 (src/torch_admp/utils.py:66-81)."""
    return torch.where(x.abs() > threshold, 1 / x, torch.zeros_like(x))


def real_energy(q, pairs, ds, bs, kappa):
    """eps * sum_P erfc(kappa d) qi qj safe_inv(d, 1e-4) bs (src/torch_admp/pme.py:312-349).

    q: (nf, N); pairs: (nf, P, 2); ds, bs: (nf, P).
    """
    qi = torch.gather(q, 1, pairs[:, :, 0])
    qj = torch.gather(q, 1, pairs[:, :, 1])
    e = torch.erfc(kappa * ds) * qi * qj * safe_inverse(ds, 1e-4) * bs
    return e.sum(-1) * DIELECTRIC


def self_energy(q, kappa):
    """-eps kappa/sqrt(pi) sum q^2 (src/torch_admp/pme.py:448-461)."""
    return -(kappa / SQRT_PI * q ** 2).sum(-1) * DIELECTRIC


def non_neutral_energy(q, box, kappa):
    """-eps pi Q^2 / (2 V kappa^2) (src/torch_admp/pme.py:463-493)."""
    vol = torch.det(box)
    return -math.pi / (2 * vol * kappa ** 2) * DIELECTRIC * q.sum(-1) ** 2


def slab_corr_energy(q, pos, box, axis=2):
    """eps 2pi/V (M^2 - Q sum q z^2 - Q^2 L^2/12) (src/torch_admp/pme.py:495-527)."""
    vol = torch.det(box)
    pre = 2 * math.pi / vol * DIELECTRIC
    z = pos[:, :, axis]
    M = (q * z).sum(-1)
    Q = q.sum(-1)
    L = torch.norm(box[:, axis], dim=-1)
    return pre * (M ** 2 - Q * (q * z ** 2).sum(-1) - Q ** 2 * L ** 2 / 12)


def obc_energy(q, pairs, ds, bs):
    """Open-boundary Coulomb eps sum qi qj safe_inv(d, 1e-8) bs (src/torch_admp/pme.py:529-561)."""
    qi = torch.gather(q, 1, pairs[:, :, 0])
    qj = torch.gather(q, 1, pairs[:, :, 1])
    return (qi * qj * DIELECTRIC * safe_inverse(ds) * bs).sum(-1)


class CoulombOracle:
    """Stateless restatement of CoulombForceModule._forward_impl (src/torch_admp/pme.py:210-310).

    Inputs are already standardised to (nf, ...) shapes (src/torch_admp/base_force.py:175-339).
    Stores the energy components as attributes like the reference.
    """

    def __init__(self, rcut, ethresh=1e-5, kspace=True, rspace=True, slab_corr=False,
                 slab_axis=2, kappa=None, kmesh=None, spacing=None, convention="reference"):
        self.rcut = rcut
        self.ethresh = ethresh
        self.kspace = kspace
        self.rspace = rspace
        self.slab_corr = slab_corr
        self.slab_axis = slab_axis
        if kappa is None:
            kappa = kappa_from_ethresh(rcut, ethresh) if kspace else 0.0
        self.kappa = kappa
        if isinstance(kmesh, int):
            kmesh = [kmesh] * 3
        self.kmesh = kmesh
        if isinstance(spacing, float):
            spacing = [spacing] * 3
        self.spacing = spacing
        self.convention = convention

    def mesh_for(self, box_frame):
        if self.kmesh is not None:
            return list(self.kmesh)
        return kmesh_rule(box_frame, self.kappa, self.ethresh, self.spacing)

    def energy(self, pos, box, pairs, ds, bs, q):
        nf = pos.shape[0]
        zero = torch.zeros(nf, dtype=pos.dtype)
        self.real = self.recip = self.self_e = self.non_neutral = self.slab = zero
        if box is None:
            return obc_energy(q, pairs, ds, bs)
        if self.rspace:
            self.real = real_energy(q, pairs, ds, bs, self.kappa)
        if self.kspace:
            self.recip = torch.stack([
                reciprocal_energy_frame(pos[i], box[i], q[i], self.kappa,
                                        self.mesh_for(box[i]), self.convention)
                for i in range(nf)])
            self.self_e = self_energy(q, self.kappa)
            self.non_neutral = non_neutral_energy(q, box, self.kappa)
        if self.slab_corr:
            self.slab = slab_corr_energy(q, pos, box, self.slab_axis)
        return self.real + self.recip + self.self_e + self.non_neutral + self.slab


def energy_forces(pos, box, q, pairs, ds_fn, bs, oracle: CoulombOracle):
    """Energy, forces (-dE/dr) and dE/dbox for one frame via autograd, as the
    reference user does with calc_grads (src/torch_admp/utils.py:84-112).

    ``ds_fn(pos, box)`` must recompute the pair distances differentiably.
    """
    pos = pos.detach().clone().requires_grad_(True)
    box_t = None if box is None else box.detach().clone().requires_grad_(True)
    ds = ds_fn(pos, box_t)
    e = oracle.energy(pos[None], None if box_t is None else box_t[None], pairs[None],
                      ds[None], bs[None].to(pos.dtype), q[None])
    inputs = [pos] + ([box_t] if box_t is not None else [])
    grads = torch.autograd.grad(e.sum(), inputs)
    f = -grads[0]
    gbox = grads[1] if box_t is not None else None
    return e.detach()[0], f.detach(), None if gbox is None else gbox.detach()


def setup_ewald_parameters(rcut, box=None, threshold=1e-5, spacing=None, method="openmm"):
    """Host-side Ewald parameter rule (src/torch_admp/pme.py:564-658)."""
    from scipy import special

    if rcut <= 0.0:
        raise ValueError(f"rcut must be positive, got {rcut}")
    if box is None:
        return 0.1, 1, 1, 1
    box = np.asarray(box.detach().cpu() if isinstance(box, torch.Tensor) else box, dtype=float)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert np.inner(box[a], box[b]) == 0.0, "Only orthogonal box is supported currently."
    if method == "openmm":
        kappa = np.sqrt(-np.log(2 * threshold)) / rcut
        ks = [int(np.ceil(2 * kappa * box[i, i] / (3.0 * threshold ** 0.2))) for i in range(3)]
    elif method == "gromacs":
        assert spacing is not None, "Spacing must be provided for gromacs method."
        kappa, i = 5.0, 0
        while special.erfc(kappa * rcut) > threshold:
            i += 1
            kappa *= 2
        lo, hi = 0.0, kappa
        for _ in range(i + 60):
            kappa = (lo + hi) / 2
            if special.erfc(kappa * rcut) > threshold:
                lo = kappa
            else:
                hi = kappa
        ks = [int(np.ceil(box[i, i] / spacing)) for i in range(3)]
    else:
        raise ValueError(f"Invalid method: {method}.Valid methods: 'openmm', 'gromacs'")
    return kappa, ks[0], ks[1], ks[2]
