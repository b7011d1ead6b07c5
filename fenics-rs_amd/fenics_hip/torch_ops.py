"""Differentiable compliance of a linear-elastic body with per-cell Young's
moduli, as a torch.autograd.Function over the libfenics_hip C ABI.

    J(E) = f^T u,   K(E) u = f,   K(E) = sum_c E_c K0_c

(K0_c: the elasticity.rs:133-136 element matrix at E = 1).  The adjoint
equation K lambda = dJ/du = f gives lambda = u because K is symmetric, so

    dJ/dE_c = -lambda_c^T K0_c u_c = -u_c^T K0_c u_c,

evaluated on the GPU by fh_cell_bilinear_k0_dev (one thread per cell).  The
forward solve is the same GPU Jacobi-PCG as fh_solve (FP64 or mixed
precision).  Tensors stay on the device; PyTorch provides the memory and the
stream synchronisation only.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import (FH_OK, FP64, ElasticMaterial, FenicsError, Mesh, SolveStats, VectorDirichletBC,
               last_error, lib)


def _ptr(t: torch.Tensor):
    assert t.is_cuda and t.is_contiguous() and t.dtype == torch.float64
    return C.c_void_p(t.data_ptr())


def _chk(code):
    if code != FH_OK:
        raise FenicsError(code, last_error())


class ElasticityProblem:
    """Cantilever-style problem: mesh, Poisson ratio, body force, Dirichlet dofs."""

    def __init__(self, mesh: Mesh, nu: float, body_force, bc: VectorDirichletBC, tol=1e-12,
                 max_iter=100000, precision=FP64):
        ElasticMaterial(1.0, nu)  # validates nu (elasticity.rs:23-26)
        self.mesh, self.nu, self.tol, self.max_iter, self.precision = mesh, nu, tol, max_iter, precision
        self.f3 = np.ascontiguousarray(body_force, dtype=np.float64)
        pd = list(bc.prescribed_dofs)
        self.dofs = np.ascontiguousarray([d for d, _ in pd], dtype=np.uint64)
        self.vals = np.ascontiguousarray([v for _, v in pd], dtype=np.float64)
        self.n = 3 * mesh.nv
        self.last_stats = None

    def solve(self, E_cell: torch.Tensor):
        """u(E) and the condensed load b (both device tensors)."""
        L, m = lib(), self.mesh
        dev = E_cell.device
        E_cell = E_cell.detach().to(torch.float64).contiguous()
        torch.cuda.synchronize(dev)
        K = C.c_void_p()
        _chk(L.fh_assemble_elasticity_dev(m.h, 1.0, self.nu, _ptr(E_cell), C.byref(K)))
        try:
            b = torch.empty(self.n, dtype=torch.float64, device=dev)
            bc = torch.empty_like(b)
            u = torch.empty_like(b)
            _chk(L.fh_assemble_body_force_dev(m.h, self.f3.ctypes.data, _ptr(b)))
            Kc = C.c_void_p()
            _chk(L.fh_apply_dirichlet_dev(K, _ptr(b), len(self.dofs), self.dofs.ctypes.data,
                                          self.vals.ctypes.data, C.byref(Kc), _ptr(bc)))
            try:
                st = SolveStats()
                _chk(L.fh_solve_dev(Kc, _ptr(bc), None, self.tol, self.max_iter, self.precision,
                                    _ptr(u), C.byref(st)))
                self.last_stats = st
            finally:
                L.fh_matrix_destroy(Kc)
        finally:
            L.fh_matrix_destroy(K)
        return u, bc

    def cell_gradient(self, w: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
        """-w_c^T K0_c u_c for every cell (device)."""
        out = torch.empty(self.mesh.nc, dtype=torch.float64, device=u.device)
        torch.cuda.synchronize(u.device)
        _chk(lib().fh_cell_bilinear_k0_dev(self.mesh.h, self.nu, _ptr(w.contiguous()),
                                           _ptr(u.contiguous()), _ptr(out)))
        return out

    def compliance(self, E_cell: torch.Tensor) -> torch.Tensor:
        return Compliance.apply(E_cell, self)


class Compliance(torch.autograd.Function):
    """J = f^T u(E) with the adjoint-state gradient dJ/dE_c = -u_c^T K0_c u_c."""

    @staticmethod
    def forward(ctx, E_cell: torch.Tensor, problem: ElasticityProblem):
        u, b = problem.solve(E_cell)
        ctx.problem = problem
        ctx.save_for_backward(u)
        return torch.dot(b, u)

    @staticmethod
    def backward(ctx, grad_J):
        (u,) = ctx.saved_tensors
        # adjoint: K lambda = dJ/du = b  ->  lambda = u (K symmetric)
        g = ctx.problem.cell_gradient(u, u)
        return grad_J * g, None
