"""Charge equilibration (QEq) with the reference call surface
(src/torch_admp/qeq.py:33-1670), solved by the HIP library.

The QEq energy is quadratic in the charges,
    E(q) = chi.q + sum hardness q^2 + E_coul(q) + E_damp(q),
so its Hessian is the fixed linear operator
    H = 2 diag(hardness) + [self terms] + J_pair + J_recip + [non-neutral, slab].
admp_qeq_matvec applies H without forming it (real-space and damping pair
coefficients are computed once per geometry; the reciprocal part is spread ->
r2c FFT -> influence function -> c2r FFT -> gather), and admp_qeq_solve runs a
projected, Jacobi-preconditioned conjugate-gradient minimisation on
{A q = b} with every scalar kept on the device.

Solver correspondence: ``solve_pgrad(method="lbfgs" | "quadratic")`` and
``pgrad_optimize`` converge to the same constrained minimum the reference's
LBFGS and Polak-Ribiere solvers reach, with the same stopping metric
|P grad E| / N < eps (src/torch_admp/qeq.py:714-765); iteration counts differ.
``solve_matrix_inversion`` / ``matinv_optimize`` build H column by column with
the matrix-free product and solve the KKT system densely.
"""

import ctypes
import math
from typing import Dict, List, Optional, Union

import torch

from . import _lib
from .base_force import BaseForceModule
from . import _ops  # noqa: F401  (registers torch.ops.admp)
from .pme import CoulombForceModule
from .utils import (calc_grads, calc_pgrads, safe_inverse, vector_projection,
                    vector_projection_coeff_matrix)

QEQ_KSPACE, QEQ_RSPACE, QEQ_DAMPING, QEQ_SLAB = 1, 2, 4, 8


class _QeqProblem(ctypes.Structure):
    """Mirror of admp_qeq_problem (include/admp_amd.h)."""

    _fields_ = [
        ("pos", ctypes.c_void_p), ("n", ctypes.c_int64), ("box", ctypes.c_void_p),
        ("cutoff", ctypes.c_double), ("kmesh", ctypes.c_int32 * 3),
        ("precision", ctypes.c_int32), ("convention", ctypes.c_int32),
        ("kappa", ctypes.c_double), ("dielectric", ctypes.c_double),
        ("flags", ctypes.c_int32), ("slab_axis", ctypes.c_int32),
        ("pairs", ctypes.c_void_p), ("ds", ctypes.c_void_p), ("bs", ctypes.c_void_p),
        ("n_pairs", ctypes.c_int64), ("eta", ctypes.c_void_p), ("hdiag", ctypes.c_void_p),
        ("chi", ctypes.c_void_p), ("precond", ctypes.c_void_p), ("cons", ctypes.c_void_p),
        ("cons_vals", ctypes.c_void_p), ("n_cons", ctypes.c_int32),
    ]


class SiteForceModule(BaseForceModule):
    """sum(chi q + hardness q^2) (src/torch_admp/qeq.py:132-209)."""

    def _forward_impl(self, positions: torch.Tensor, box: Optional[torch.Tensor],
                      pairs: torch.Tensor, ds: torch.Tensor, buffer_scales: torch.Tensor,
                      params: Dict[str, torch.Tensor]) -> torch.Tensor:
        nf, na = positions.size(0), positions.size(1)
        ec = self.const_lib.energy_coeff
        chi = params["chi"].reshape(nf, na) * ec
        hardness = params["hardness"].reshape(nf, na) * ec
        q = params["charge"].reshape(nf, na)
        return torch.sum(chi * q + hardness * q ** 2, dim=-1)


class GaussianDampingForceModule(BaseForceModule):
    """Gaussian short-range damping (src/torch_admp/qeq.py:33-129): the pair
    sum runs in the HIP pair kernel, the per-atom term is elementwise."""

    def _forward_impl(self, positions: torch.Tensor, box: Optional[torch.Tensor],
                      pairs: torch.Tensor, ds: torch.Tensor, buffer_scales: torch.Tensor,
                      params: Dict[str, torch.Tensor]) -> torch.Tensor:
        nf, na = positions.size(0), positions.size(1)
        q = params["charge"].reshape(nf, na)
        eta = params["eta"].reshape(nf, na) * self.const_lib.length_coeff
        eps = self.const_lib.dielectric
        # pair mode 2: Gaussian-damped pairs
        e_pair = torch.ops.admp.pair_energy(q, ds, pairs, buffer_scales, eta, 2, 0.0, eps)
        pre_self = safe_inverse(eta, threshold=1e-4) / (2 * self.const_lib.sqrt_pi)
        return e_pair + torch.sum(pre_self * q * q, dim=-1) * eps


class QEqForceModule(BaseForceModule):
    """Charge equilibration (src/torch_admp/qeq.py:212-765).

    forward and func_energy script (torch.jit.script) and save; the solvers
    and the Hessian are Python methods around the library's QEq entry points."""

    sel: Optional[List[int]]

    def __init__(self, rcut: float, ethresh: float = 1e-5, kspace: bool = True,
                 rspace: bool = True, slab_corr: bool = False, slab_axis: int = 2,
                 max_iter: int = 20, ls_eps: float = 1e-4, eps: float = 1e-4,
                 units_dict: Optional[Dict] = None, damping: bool = True,
                 sel: Optional[List[int]] = None, kappa: Optional[float] = None,
                 spacing: Union[List[float], float, None] = None,
                 kmesh: Union[List[int], int, None] = None, precision: str = "float64",
                 convention: str = "reference") -> None:
        BaseForceModule.__init__(self, units_dict)
        models: Dict[str, BaseForceModule] = {
            "site": SiteForceModule(units_dict=units_dict),
            "coulomb": CoulombForceModule(rcut=rcut, ethresh=ethresh, kspace=kspace,
                                          rspace=rspace, slab_corr=slab_corr,
                                          slab_axis=slab_axis, units_dict=units_dict,
                                          kappa=kappa, spacing=spacing, kmesh=kmesh,
                                          precision=precision, convention=convention),
        }
        if damping:
            models["damping"] = GaussianDampingForceModule(units_dict=units_dict)
        self.submodels = torch.nn.ModuleDict(models)
        self.rcut = rcut
        self.max_iter = max_iter
        self.ls_eps = ls_eps
        self.eps = eps
        self.converge_iter: int = -1
        self.residual: float = float("nan")
        self.sel = sel
        self.slab_axis = slab_axis
        self.slab_corr = slab_corr
        self.damping = damping
        self.cg_max_iter = max(25 * max_iter, 200)

    def get_rcut(self) -> float:
        return self.rcut

    def get_sel(self) -> Optional[List[int]]:
        return self.sel

    def _forward_impl(self, positions: torch.Tensor, box: Optional[torch.Tensor],
                      pairs: torch.Tensor, ds: torch.Tensor, buffer_scales: torch.Tensor,
                      params: Dict[str, torch.Tensor]) -> torch.Tensor:
        """Sum of the submodels (src/torch_admp/qeq.py:351-406)."""
        energy = torch.zeros(positions.size(0), device=positions.device, dtype=positions.dtype)
        for model in self.submodels.values():
            energy = energy + model._forward_impl(positions, box, pairs, ds, buffer_scales,
                                                  params)
        return energy

    @torch.jit.export
    def func_energy(self, charges: torch.Tensor, positions: torch.Tensor,
                    box: Optional[torch.Tensor], chi: torch.Tensor, hardness: torch.Tensor,
                    eta: torch.Tensor, pairs: torch.Tensor, ds: torch.Tensor,
                    buffer_scales: torch.Tensor) -> torch.Tensor:
        """Total energy at the given charges (src/torch_admp/qeq.py:408-458)."""
        params = {"charge": charges, "chi": chi, "hardness": hardness, "eta": eta}
        energy = torch.zeros(1, device=positions.device, dtype=positions.dtype)
        for model in self.submodels.values():
            energy = energy + model(positions, box, pairs, ds, buffer_scales, params)
        return energy

    # -- the HIP problem description ------------------------------------------
    def _problem(self, positions, box, chi, hardness, eta, pairs, ds, buffer_scales,
                 constraint_matrix=None, constraint_vals=None):
        _lib.require_gpu(positions)
        coul: CoulombForceModule = self.submodels["coulomb"]  # type: ignore[assignment]
        lc = self.const_lib.length_coeff
        ec = self.const_lib.energy_coeff
        eps = self.const_lib.dielectric
        dev = positions.device
        f64 = torch.float64
        pos = (positions.detach().reshape(-1, 3).to(f64) * lc).contiguous()
        n = pos.shape[0]
        prs = pairs.detach().reshape(-1, 2).to(torch.int64).contiguous()
        dsv = (ds.detach().reshape(-1).to(f64) * lc).contiguous()
        bsv = buffer_scales.detach().reshape(-1).to(f64).contiguous()
        eta_i = (eta.detach().reshape(-1).to(f64) * lc).contiguous()
        hard = hardness.detach().reshape(-1).to(f64) * ec
        flags = 0
        if coul.rspace_flag:
            flags |= QEQ_RSPACE
        if self.damping:
            flags |= QEQ_DAMPING
        box_int = None
        if box is not None:
            box_int = box.detach().reshape(3, 3).to(f64) * lc
            if coul.kspace_flag:
                flags |= QEQ_KSPACE
            if coul.slab_corr_flag:
                flags |= QEQ_SLAB
        hdiag = 2 * hard
        if flags & QEQ_KSPACE:
            hdiag = hdiag - 2 * eps * coul.kappa / math.sqrt(math.pi)
        precond = 2 * hard
        if self.damping:
            dself = eps * safe_inverse(eta_i, 1e-4) / math.sqrt(math.pi)
            hdiag = hdiag + dself
            precond = precond + dself
        precond = precond if bool((precond > 0).all()) else None
        keep = [pos, prs, dsv, bsv, eta_i, hdiag.contiguous()]
        pr = _QeqProblem()
        pr.pos = pos.data_ptr()
        pr.n = n
        boxh = _lib.box_array(box_int)
        keep.append(boxh)
        pr.box = ctypes.cast(boxh, ctypes.c_void_p) if boxh is not None else None
        pr.cutoff = float(self.rcut * lc)
        km = coul.mesh_for_box(box_int) if box_int is not None and (flags & QEQ_KSPACE) else [1, 1, 1]
        for d in range(3):
            pr.kmesh[d] = int(km[d])
        pr.precision = _lib.PRECISION[coul.precision]
        pr.convention = _lib.KCONV[coul.convention]
        pr.kappa = float(coul.kappa)
        pr.dielectric = eps
        pr.flags = flags
        pr.slab_axis = self.slab_axis
        pr.pairs = prs.data_ptr()
        pr.ds = dsv.data_ptr()
        pr.bs = bsv.data_ptr()
        pr.n_pairs = prs.shape[0]
        pr.eta = eta_i.data_ptr()
        pr.hdiag = keep[5].data_ptr()
        if chi is not None:
            chi_i = (chi.detach().reshape(-1).to(f64) * ec).contiguous()
            keep.append(chi_i)
            pr.chi = chi_i.data_ptr()
        if precond is not None:
            pc = precond.contiguous()
            keep.append(pc)
            pr.precond = pc.data_ptr()
        if constraint_matrix is not None:
            A = constraint_matrix.detach().to(f64).reshape(-1, n).contiguous()
            b = constraint_vals.detach().to(f64).reshape(-1).contiguous()
            keep += [A, b]
            pr.cons = A.data_ptr()
            pr.cons_vals = b.data_ptr()
            pr.n_cons = A.shape[0]
        if box_int is not None and (flags & QEQ_KSPACE):
            coul._kmesh = torch.tensor(km, dtype=torch.long)
            coul.__dict__.pop("_kmesh_list", None)
        lib = _lib.load()
        nbytes = ctypes.c_size_t()
        _lib.check(lib.admp_qeq_workspace_bytes(ctypes.byref(pr), ctypes.byref(nbytes)),
                   "admp_qeq_workspace_bytes")
        ws = _lib.workspace(nbytes.value, dev)
        return pr, keep, ws, nbytes.value

    def hessian_matvec(self, v, positions, box, chi, hardness, eta, pairs, ds, buffer_scales):
        """H v in the user's energy unit per charge^2 (no graph)."""
        pr, keep, ws, nb = self._problem(positions, box, None, hardness, eta, pairs, ds,
                                         buffer_scales)
        lib = _lib.load()
        vv = v.detach().reshape(-1).to(torch.float64).contiguous()
        out = torch.empty_like(vv)
        _lib.check(lib.admp_qeq_matvec(ctypes.byref(pr), _lib.ptr(vv), _lib.ptr(out),
                                       _lib.ptr(ws), nb, _lib.stream_of(vv)), "admp_qeq_matvec")
        return out / self.const_lib.energy_coeff

    def calc_hessian(self, positions, box, chi, hardness, eta, pairs, ds, buffer_scales):
        """Dense d2E/dq2 (src/torch_admp/qeq.py:460-515), built column by column
        with the matrix-free product instead of N double-backward passes."""
        pr, keep, ws, nb = self._problem(positions, box, None, hardness, eta, pairs, ds,
                                         buffer_scales)
        lib = _lib.load()
        n = pr.n
        cols = torch.empty(n, n, dtype=torch.float64, device=positions.device)
        e = torch.zeros(n, dtype=torch.float64, device=positions.device)
        stream = _lib.stream_of(e)
        for j in range(n):
            e.zero_()
            e[j] = 1.0
            _lib.check(lib.admp_qeq_matvec(ctypes.byref(pr), _lib.ptr(e), _lib.ptr(cols[j]),
                                           _lib.ptr(ws), nb, stream), "admp_qeq_matvec")
        hess = 0.5 * (cols + cols.T) / self.const_lib.energy_coeff
        return hess.to(positions.dtype)

    def solve_matrix_inversion(self, positions, box, chi, hardness, eta, pairs, ds,
                               buffer_scales, constraint_matrix, constraint_vals,
                               check_hessian: bool = False):
        """Dense KKT solve (src/torch_admp/qeq.py:517-607).
        Returns (energy, q_opt, diag(H), fermi)."""
        hessian = self.calc_hessian(positions, box, chi, hardness, eta, pairs, ds, buffer_scales)
        n = hessian.shape[0]
        if constraint_matrix is None:
            coeff = hessian
            vector = -chi
        else:
            nc = constraint_matrix.shape[0]
            coeff = torch.cat([torch.cat([hessian, constraint_matrix.T], 1),
                               torch.cat([constraint_matrix,
                                          torch.zeros(nc, nc, device=positions.device,
                                                      dtype=hessian.dtype)], 1)], 0)
            vector = torch.cat([-chi, constraint_vals])
        if check_hessian:
            print(torch.all(torch.diag(hessian) > 0.0))
        q_full = torch.linalg.solve(coeff, vector.reshape(-1, 1).to(coeff.dtype)).reshape(-1)
        q_opt = q_full[:n].detach().clone().requires_grad_(True)
        energy = self.func_energy(q_opt, positions, box, chi, hardness, eta, pairs, ds,
                                  buffer_scales)
        fermi = torch.mean(chi + hessian @ q_full[:n])
        return energy, q_full[:n], torch.diag(hessian), fermi

    def solve_pgrad(self, q0, positions, box, chi, hardness, eta, pairs, ds, buffer_scales,
                    constraint_matrix, constraint_vals, coeff_matrix=None,
                    reinit_q: bool = False, method: str = "lbfgs"):
        """Projected-gradient solve (src/torch_admp/qeq.py:609-712) by the HIP
        projected CG; returns (energy, q_opt) with q_opt detached, energy
        differentiable in positions at fixed charges.

        Afterwards ``self.converge_iter`` holds the number of projected-CG
        iterations the solve took (tens, one Hessian product each), not the
        reference's count of LBFGS / line-search outer steps (at most
        ``max_iter``); ``self.residual`` holds the final |P grad E| / n."""
        if method not in ("lbfgs", "quadratic"):
            raise ValueError(f"Method {method} is not supported.")
        n = positions.reshape(-1, 3).shape[0]
        if q0 is None:
            q0 = torch.rand(n, device=positions.device, dtype=positions.dtype)
        assert q0.shape[0] == n
        q, iters, resid = self._pcg(q0, positions, box, chi, hardness, eta, pairs, ds,
                                    buffer_scales, constraint_matrix, constraint_vals, self.eps)
        self.converge_iter = iters
        self.residual = resid
        q_opt = q.to(positions.dtype).detach().requires_grad_(True)
        energy = self.func_energy(q_opt, positions, box, chi, hardness, eta, pairs, ds,
                                  buffer_scales)
        return energy, q_opt

    def solve_implicit(self, positions, box, chi, hardness, eta, pairs, ds, buffer_scales,
                       constraint_matrix=None, constraint_vals=None, tol: float = 1e-10):
        """Equilibrated charges q*(positions, box, chi, hardness, eta, ds, b) as a
        differentiable function of its inputs.

        Forward: the projected CG solve to |P grad E| / N < tol.  Backward (the
        implicit-function theorem on the KKT system [H A^T; A 0][q; mu] = [-chi; b]):
        one adjoint solve H u + A^T nu = g, A u = 0 for the upstream gradient g,
        then dL/dchi = -u, dL/dhardness = -2 u q*, dL/db = nu and, for the
        geometry (positions, box, ds, eta), dL/dtheta = -d/dtheta (u^T J q*) with
        u^T J q* = (E_c(u + q*) - E_c(u - q*)) / 2 over the Coulomb and damping
        terms, whose gradients come from the library's backward kernels.  The
        reference differentiates through its dense solve instead
        (src/torch_admp/qeq.py:517-607).
        """
        n = positions.reshape(-1, 3).shape[0]
        if constraint_matrix is not None:
            constraint_matrix = constraint_matrix.detach().to(torch.float64).reshape(-1, n)
            if constraint_vals is None:
                constraint_vals = torch.zeros(constraint_matrix.shape[0], dtype=torch.float64,
                                              device=positions.device)
        else:
            constraint_vals = None
        return _QEqImplicit.apply(self, positions, box, chi, hardness, eta, ds, constraint_vals,
                                  pairs, buffer_scales, constraint_matrix, float(tol))

    def _coulomb_damping_energy(self, charges, positions, box, eta, pairs, ds, buffer_scales):
        """E_c(q): the Coulomb and damping terms only (quadratic in q)."""
        params = {"charge": charges, "eta": eta}
        energy = torch.zeros(1, device=positions.device, dtype=positions.dtype)
        for name, model in self.submodels.items():
            if name != "site":
                energy = energy + model(positions, box, pairs, ds, buffer_scales, params)
        return energy.sum()

    def _pcg(self, q0, positions, box, chi, hardness, eta, pairs, ds, buffer_scales,
             constraint_matrix, constraint_vals, tol, max_iter=None, check_every=4):
        pr, keep, ws, nb = self._problem(positions, box, chi, hardness, eta, pairs, ds,
                                         buffer_scales, constraint_matrix, constraint_vals)
        lib = _lib.load()
        if q0 is None:
            q0 = torch.zeros(pr.n, dtype=torch.float64, device=positions.device)
        start = q0.detach().reshape(-1).to(torch.float64).contiguous()
        q = torch.empty_like(start)
        iters = ctypes.c_int32()
        resid = ctypes.c_double()
        # the stopping metric is |P grad E| / N in the user's energy unit
        tol_int = float(tol) * self.const_lib.energy_coeff
        _lib.check(lib.admp_qeq_solve(ctypes.byref(pr), _lib.ptr(start), tol_int,
                                      int(max_iter or self.cg_max_iter), int(check_every),
                                      _lib.ptr(q), ctypes.byref(iters), ctypes.byref(resid),
                                      _lib.ptr(ws), nb, _lib.stream_of(start)), "admp_qeq_solve")
        return q, int(iters.value), float(resid.value) / self.const_lib.energy_coeff

    def optimality(self, charges, positions, box, chi, hardness, eta, pairs, ds, buffer_scales,
                   constraint_matrix, coeff_matrix) -> torch.Tensor:
        """|P grad E| / N (src/torch_admp/qeq.py:714-765)."""
        energy = self.func_energy(charges, positions, box, chi, hardness, eta, pairs, ds,
                                  buffer_scales)
        pgrads = calc_pgrads(energy, charges, constraint_matrix, coeff_matrix)
        return torch.norm(pgrads) / charges.shape[0]


def pgrad_optimize(module: QEqForceModule, q0, positions, box, chi, hardness, eta, pairs, ds,
                   buffer_scales, constraint_matrix, constraint_vals, coeff_matrix=None,
                   reinit_q: bool = False, method: str = "lbfgs", **kwargs):
    """Module-level projected-gradient solve (src/torch_admp/qeq.py:1117-1270)."""
    return module.solve_pgrad(q0, positions, box, chi, hardness, eta, pairs, ds, buffer_scales,
                              constraint_matrix, constraint_vals, coeff_matrix, reinit_q, method)


def matinv_optimize(module: QEqForceModule, positions, box, chi, hardness, eta, pairs, ds,
                    buffer_scales, constraint_matrix, constraint_vals, **kwargs):
    """Module-level dense solve (src/torch_admp/qeq.py:1530-1617): returns
    (energy, q_opt), q_opt a detached leaf that requires grad and the energy
    evaluated at it."""
    _, q_full, _, _ = module.solve_matrix_inversion(positions, box, chi, hardness, eta, pairs,
                                                    ds, buffer_scales, constraint_matrix,
                                                    constraint_vals)
    q_opt = q_full.detach().clone().requires_grad_(True)
    energy = module.func_energy(q_opt, positions, box, chi, hardness, eta, pairs, ds,
                                buffer_scales)
    return energy, q_opt


def calc_hessian(func_energy, positions, box, chi, hardness, eta, pairs, ds, buffer_scales):
    """Module-level Hessian (src/torch_admp/qeq.py:1620-1670); ``func_energy``
    must be the bound ``func_energy`` of a QEqForceModule."""
    module = getattr(func_energy, "__self__", None)
    if not isinstance(module, QEqForceModule):
        raise TypeError("calc_hessian expects QEqForceModule.func_energy")
    return module.calc_hessian(positions, box, chi, hardness, eta, pairs, ds, buffer_scales)


__all__ = ["SiteForceModule", "GaussianDampingForceModule", "QEqForceModule", "pgrad_optimize",
           "matinv_optimize", "calc_hessian", "vector_projection", "vector_projection_coeff_matrix",
           "calc_grads"]


class _QEqImplicit(torch.autograd.Function):
    """q* of the QEq solve with the implicit-function backward (see
    QEqForceModule.solve_implicit)."""

    @staticmethod
    def forward(ctx, module, positions, box, chi, hardness, eta, ds, constraint_vals, pairs,
                buffer_scales, constraint_matrix, tol):
        q, iters, resid = module._pcg(None, positions, box, chi, hardness, eta, pairs, ds,
                                      buffer_scales, constraint_matrix, constraint_vals, tol)
        if iters < 0:
            raise RuntimeError(f"QEq solve did not converge (|P grad E| / N = {resid:.3e})")
        module.converge_iter = iters
        module.residual = resid
        q = q.to(positions.dtype)
        ctx.module = module
        ctx.pairs = pairs
        ctx.buffer_scales = buffer_scales
        ctx.constraint_matrix = constraint_matrix
        ctx.tol = tol
        ctx.has_box = box is not None
        ctx.has_cons = constraint_vals is not None
        ctx.save_for_backward(positions, box if box is not None else torch.empty(0), chi,
                              hardness, eta, ds, q)
        return q

    @staticmethod
    def backward(ctx, g):
        module = ctx.module
        positions, box, chi, hardness, eta, ds, q = ctx.saved_tensors
        box = box if ctx.has_box else None
        A = ctx.constraint_matrix
        n = q.shape[0]
        g = g.detach().reshape(-1).to(torch.float64)
        zeros_b = None if A is None else torch.zeros(A.shape[0], dtype=torch.float64,
                                                     device=g.device)
        # adjoint solve: min 1/2 u^T H u - g^T u on {A u = 0}
        scale = float(g.abs().max()) if g.numel() else 0.0
        u = torch.zeros(n, dtype=torch.float64, device=g.device)
        if scale > 0.0:
            u, iters, resid = module._pcg(None, positions, box, -g, hardness, eta, ctx.pairs, ds,
                                          ctx.buffer_scales, A, zeros_b,
                                          max(ctx.tol * scale, 1e-300))
            if iters < 0:
                raise RuntimeError("QEq adjoint solve did not converge "
                                   f"(|P (H u - g)| / N = {resid:.3e})")
        grads = [None] * 12
        uq = u.to(q.dtype)
        if ctx.needs_input_grad[3]:
            grads[3] = (-uq).reshape(chi.shape)
        if ctx.needs_input_grad[4]:
            grads[4] = (-2.0 * uq * q).reshape(hardness.shape)
        if ctx.needs_input_grad[7] and A is not None:
            # nu = (A A^T)^-1 A (g - H u)
            hu = module.hessian_matvec(u, positions, box, None, hardness, eta, ctx.pairs, ds,
                                       ctx.buffer_scales).to(torch.float64)
            r = g - hu
            nu = torch.linalg.solve(A @ A.T, A @ r)
            grads[7] = nu.to(q.dtype)
        geo = [(1, positions), (2, box), (5, eta), (6, ds)]
        want = [(i, t) for i, t in geo if t is not None and ctx.needs_input_grad[i]]
        if want:
            with torch.enable_grad():
                inputs = {i: t.detach().requires_grad_(True) for i, t in want}
                pos_ = inputs.get(1, positions.detach())
                box_ = inputs.get(2, None if box is None else box.detach())
                eta_ = inputs.get(5, eta.detach())
                ds_ = inputs.get(6, ds.detach())
                e_p = module._coulomb_damping_energy(uq + q, pos_, box_, eta_, ctx.pairs, ds_,
                                                     ctx.buffer_scales)
                e_m = module._coulomb_damping_energy(uq - q, pos_, box_, eta_, ctx.pairs, ds_,
                                                     ctx.buffer_scales)
                mixed = 0.5 * (e_p - e_m)
                out = torch.autograd.grad(mixed, [inputs[i] for i, _ in want], allow_unused=True)
            for (i, t), gr in zip(want, out):
                grads[i] = None if gr is None else -gr
        return tuple(grads)


def solve_batch(positions: List[torch.Tensor], boxes: List[torch.Tensor],
                chi: List[torch.Tensor], hardness: List[torch.Tensor], eta: List[torch.Tensor],
                rcut: Union[float, List[float]], ethresh: Union[float, List[float]] = 1e-5,
                total_charge: Optional[List[float]] = None,
                tol: Union[float, List[float]] = 1e-8, max_iter: int = 1000, check_every: int = 4,
                damping: bool = True, kappa: Optional[List[float]] = None,
                kmesh: Optional[List[List[int]]] = None, precision: str = "float64",
                convention: str = "reference", q0: Optional[List[torch.Tensor]] = None):
    """Charge equilibration of several periodic systems at once.

    System s has positions (n_s, 3), box (3, 3) rows, chi, hardness (the q^2
    coefficient) and Gaussian widths eta (n_s,) on the device, in the
    reference's default units (Angstrom, eV).  Each is solved for the minimum
    of QEqForceModule's energy (Ewald real and reciprocal space with
    CoulombForceModule's kappa and mesh rule for ``ethresh``, Gaussian damping
    when ``damping``) on sum q = total_charge[s], to |P grad E| / n_s < tol.
    ``rcut``, ``ethresh`` and ``tol`` are one value or one per system.
    All systems advance together through admp_qeq_solve_batch: one Hessian
    product for the whole batch per iteration, one CG step per system, and
    each system stops when it converges.  Returns (charges per system,
    iterations per system (-1 if not converged), residuals per system).

    The reference solves one frame per call (src/torch_admp/qeq.py:609-765);
    each system's fixed point is the one QEqForceModule.solve_pgrad reaches.
    """
    from .nblist import TorchNeighborList
    from .utils import DIELECTRIC

    nsys = len(positions)
    if nsys == 0:
        return [], [], []
    _lib.require_gpu(*positions)
    dev = positions[0].device
    f64 = torch.float64
    eps = DIELECTRIC
    per = lambda v: [float(x) for x in v] if isinstance(v, (list, tuple)) else [float(v)] * nsys  # noqa: E731
    rcs, eths, tols = per(rcut), per(ethresh), per(tol)
    n_atoms, pos_l, pairs_l, ds_l, bs_l, hd_l, pc_l, chi_l, eta_l, q0_l = ([] for _ in range(10))
    boxes_h, kap_h, km_h = [], [], []
    off = 0
    for s in range(nsys):
        pos = positions[s].detach().reshape(-1, 3).to(f64).contiguous()
        box = boxes[s].detach().reshape(3, 3).to(f64)
        n = pos.shape[0]
        # CoulombForceModule's rule (src/torch_admp/pme.py:113-115,383-392)
        kap = (float(kappa[s]) if kappa is not None
               else math.sqrt(-math.log(2.0 * eths[s])) / rcs[s])
        km = ([int(k) for k in kmesh[s]] if kmesh is not None else
              [int(math.ceil(2.0 * kap * float(box[d, d]) / (3.0 * eths[s] ** 0.2)))
               for d in range(3)])
        nbl = TorchNeighborList(rcs[s])
        with torch.no_grad():
            prs = nbl(pos, box.to(dev)).to(torch.int64).reshape(-1, 2)
            ds = nbl.get_ds().reshape(-1).to(f64)
            bs = nbl.get_buffer_scales().reshape(-1).to(f64)
        hard = hardness[s].detach().reshape(-1).to(f64)
        et = eta[s].detach().reshape(-1).to(f64)
        hdiag = 2 * hard - 2 * eps * kap / math.sqrt(math.pi)
        pc = 2 * hard
        if damping:
            dself = eps * safe_inverse(et, 1e-4) / math.sqrt(math.pi)
            hdiag = hdiag + dself
            pc = pc + dself
        n_atoms.append(n)
        pos_l.append(pos)
        pairs_l.append(prs + off)
        ds_l.append(ds)
        bs_l.append(bs)
        hd_l.append(hdiag)
        pc_l.append(pc)
        chi_l.append(chi[s].detach().reshape(-1).to(f64))
        eta_l.append(et)
        q0_l.append(torch.zeros(n, dtype=f64, device=dev) if q0 is None
                    else q0[s].detach().reshape(-1).to(f64))
        boxes_h.append(box.cpu().numpy().reshape(-1))
        kap_h.append(kap)
        km_h.append(km)
        off += n
    cat = lambda xs: torch.cat(xs).contiguous()  # noqa: E731
    pos_a, prs_a, ds_a, bs_a = cat(pos_l), cat(pairs_l), cat(ds_l), cat(bs_l)
    hd_a, chi_a, eta_a, q0_a = cat(hd_l), cat(chi_l), cat(eta_l), cat(q0_l)
    pc_a = cat(pc_l)
    precond = pc_a if bool((pc_a > 0).all()) else None
    import numpy as np

    n_arr = np.asarray(n_atoms, dtype=np.int64)
    box_arr = np.ascontiguousarray(np.stack(boxes_h), dtype=np.float64)
    rc_arr = np.asarray(rcs, dtype=np.float64)
    tol_arr = np.asarray(tols, dtype=np.float64)
    kap_arr = np.asarray(kap_h, dtype=np.float64)
    km_arr = np.ascontiguousarray(np.asarray(km_h, dtype=np.int32))
    tq = np.asarray([0.0] * nsys if total_charge is None else total_charge, dtype=np.float64)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    lib = _lib.load()
    prec = _lib.PRECISION[precision]
    nbytes = ctypes.c_size_t()
    _lib.check(lib.admp_qeq_solve_batch_workspace_bytes(nsys, p(n_arr), p(box_arr), p(rc_arr),
                                                        p(km_arr), prec, prs_a.shape[0],
                                                        ctypes.byref(nbytes)),
               "admp_qeq_solve_batch_workspace_bytes")
    ws = _lib.workspace(nbytes.value, dev)
    q = torch.empty_like(q0_a)
    iters = np.zeros(nsys, dtype=np.int32)
    resid = np.zeros(nsys, dtype=np.float64)
    _lib.check(lib.admp_qeq_solve_batch(
        nsys, p(n_arr), p(box_arr), p(rc_arr), p(kap_arr), p(km_arr), prec,
        _lib.KCONV[convention], eps, QEQ_DAMPING if damping else 0, _lib.ptr(pos_a),
        _lib.ptr(prs_a), _lib.ptr(ds_a), _lib.ptr(bs_a), prs_a.shape[0], _lib.ptr(eta_a),
        _lib.ptr(hd_a), _lib.ptr(chi_a), _lib.ptr(precond), p(tq), _lib.ptr(q0_a), p(tol_arr),
        int(max_iter), int(check_every), _lib.ptr(q), p(iters), p(resid), _lib.ptr(ws),
        nbytes.value, _lib.stream_of(pos_a)), "admp_qeq_solve_batch")
    return (list(torch.split(q, n_atoms)), [int(i) for i in iters], [float(r) for r in resid])
