"""Adjoint shape and wave-number gradients of the Dirichlet problem.

The reference solves the interior Dirichlet problem for the P0 Neumann trace
t with P1 Dirichlet data g (BEMManager.run, BEM_manager.py:990-1011):

    V(x) t = (1/2 M(x) + K(x)) g            (x = vertex positions)

It has no derivatives.  ``NeumannTrace`` is that solve as a
``torch.autograd.Function`` of (vertices, k, g).  Its backward never forms
dV/dx or dK/dx: one transposed LU solve gives the adjoint
mu = V^-T conj(dL/dt), and ``bem_shape_gradient`` re-evaluates every
quadrature point pair of every panel pair to contract

    dL/dx = -Re(mu^T dV t) + Re(mu^T dK g) + 1/2 Re(mu^T dM g)
    dL/dk = -Re(mu^T dV/dk t) + Re(mu^T dK/dk g)
    dL/dg = conj((1/2 M + K)^T mu)

on the device (dM through ``bem_area_gradient``).  Gradients that flow into
g (its dependence on x or k) are left to torch's autograd of the caller's g.
Complex tensors follow torch's convention: the gradient of a real loss L
with respect to z is dL/dRe z + i dL/dIm z.
"""
from __future__ import annotations

import torch

from . import _lib
from .bem_functions.utils.types import KernelType


class DirichletProblem:
    """Mesh connectivity, kernel and quadrature of a NeumannTrace solve, and
    the device context it runs in."""

    def __init__(self, panels, kernel: int = KernelType.HELMHOLTZ, gauss_qr: int = 5, flags: int = 0,
                 sqrt_n: float = 1.0, device: int = 0, context: _lib.Context | None = None):
        self.device = torch.device("cuda", device)
        self.panels = torch.as_tensor(panels, dtype=torch.int32).to(self.device).contiguous()
        self.kernel = int(kernel)
        self.gauss_qr = int(gauss_qr)
        self.flags = int(flags)
        self.sqrt_n = float(sqrt_n)
        self.ctx = context if context is not None else _lib.Context(device)

    def set_vertices(self, vertices: torch.Tensor) -> None:
        """Uploads the mesh at these vertex positions (every call: the
        forward and backward of one solve may be separated by other solves
        on the same context)."""
        v = vertices.detach().to(torch.float64).contiguous()
        self.ctx.set_mesh(v, self.panels, self.gauss_qr)

    def desc(self, op: int, k: float, deg_trial: int):
        return _lib.Context.op_desc(op, self.kernel, k, self.sqrt_n, 0, deg_trial, _lib.PREC_FP64, _lib.C128,
                                    self.flags)


def panel_areas(vertices: torch.Tensor, panels: torch.Tensor) -> torch.Tensor:
    """Differentiable panel areas 1/2 |(x2-x1) x (x3-x1)| (mesh_manager.py:114-124)."""
    p = panels.long()
    a, b, c = vertices[p[:, 0]], vertices[p[:, 1]], vertices[p[:, 2]]
    return 0.5 * torch.linalg.vector_norm(torch.linalg.cross(b - a, c - a), dim=1)


class NeumannTrace(torch.autograd.Function):
    """t = V^-1 (1/2 M + K) g on the device, differentiable in vertices, k and g."""

    @staticmethod
    def forward(ctx, vertices: torch.Tensor, k: torch.Tensor, g: torch.Tensor, problem: DirichletProblem):
        prob = problem
        prob.set_vertices(vertices)
        kk = float(k)
        C = prob.ctx
        nf = int(prob.panels.shape[0])
        gc = g.detach().to(torch.complex128).contiguous()
        A = torch.zeros(nf, nf, dtype=torch.complex128, device=prob.device)
        b = torch.zeros(nf, dtype=torch.complex128, device=prob.device)
        C.assemble_dirichlet(prob.desc(_lib.OP_V, kk, 0), A, gc, b, alpha=1.0, alpha_k=1.0, alpha_m=0.5,
                             conj_g=False)
        ipiv = torch.empty(nf, dtype=torch.int32, device=prob.device)
        C.lu_factor(A, ipiv)
        t = b.clone()
        C.lu_solve(A, ipiv, t)
        ctx.prob = prob
        ctx.k = kk
        ctx.real_out = not g.is_complex() and prob.kernel == KernelType.LAPLACE
        ctx.g_real = not g.is_complex()
        ctx.save_for_backward(vertices.detach().to(torch.float64).contiguous(), gc, t, A, ipiv)
        return t.real.clone() if ctx.real_out else t

    @staticmethod
    def backward(ctx, gt: torch.Tensor):
        vertices, gc, t, LU, ipiv = ctx.saved_tensors
        prob: DirichletProblem = ctx.prob
        prob.set_vertices(vertices)
        C = prob.ctx
        # adjoint: V^T mu = conj(dL/dt)
        mu = gt.to(torch.complex128).conj().contiguous().clone()
        C.lu_solve_transpose(LU, ipiv, mu)
        nv = int(vertices.shape[0])
        grad_x = torch.zeros(nv, 3, dtype=torch.float64, device=prob.device)
        grad_k = torch.zeros(1, dtype=torch.float64, device=prob.device)
        C.shape_gradient(prob.desc(_lib.OP_V, ctx.k, 0), mu, t, grad_x, grad_k, alpha=-1.0)
        # the K contraction also returns K^T mu when g needs a gradient
        bt = torch.zeros(nv, dtype=torch.complex128, device=prob.device) if ctx.needs_input_grad[2] else None
        C.shape_gradient(prob.desc(_lib.OP_K, ctx.k, 1), mu, gc, grad_x, grad_k, alpha=1.0, kt=bt)
        # 1/2 dM g: M_iv = A_i / 3 for the three vertices v of panel i
        gbar = gc[prob.panels.long()].mean(dim=1)
        C.area_gradient((0.5 * (mu * gbar).real).contiguous(), grad_x)
        grad_g = None
        if bt is not None:
            # (1/2 M + K)^T mu: K^T mu from the contraction, P1 x P0 mass
            C.mass_apply(1, 0, mu, bt, alpha=0.5, conj_x=False)
            grad_g = bt.conj()
            if ctx.g_real:
                grad_g = grad_g.real.contiguous()
        gk = grad_k.reshape(()) if ctx.needs_input_grad[1] else None
        return grad_x, gk, grad_g, None


def neumann_trace(vertices: torch.Tensor, k: torch.Tensor, g: torch.Tensor, problem: DirichletProblem):
    return NeumannTrace.apply(vertices, k, g, problem)


def point_source(vertices: torch.Tensor, k: torch.Tensor, x0, kernel: int) -> torch.Tensor:
    """Dirichlet data of a source at x0 outside the surface: exp(i k r)/r
    (Helmholtz) or 1/r (Laplace), r = |x - x0|, differentiable in x and k."""
    r = torch.linalg.vector_norm(vertices - torch.as_tensor(x0, dtype=vertices.dtype, device=vertices.device),
                                 dim=1)
    if kernel == KernelType.LAPLACE:
        return 1.0 / r
    return torch.exp(1j * (k * r)) / r


def misfit(t: torch.Tensor, t_target: torch.Tensor, vertices: torch.Tensor, panels: torch.Tensor) -> torch.Tensor:
    """J = ||t - t_target||^2_M with the P0 mass (panel areas)."""
    return (panel_areas(vertices, panels) * (t - t_target).abs() ** 2).sum()


# ---------------------------------------------------------------------------
# Transmission problem: derivative with respect to the interior wave number
# ---------------------------------------------------------------------------
def transmission_dA_contract(bm, mu: torch.Tensor, x: torch.Tensor) -> complex:
    """mu^T (dA/dk_i) x for the assembled 2 x 2 Calderon system of a
    transmission BEMManager, k_i = k sqrt(n_i) the interior wave number.  The
    interior blocks enter A as -W at (D,D), -K^T at (D,N), +K at (N,D) and -V
    at (N,N) (BEMManager.matA_add_M with multiplier +1, BEM_manager.py:608-628),
    so with the k-derivative products of the library (BEM_FLAG_DK)

        mu^T dA x = -mu_D.dW x_D - x_N.dK mu_D + mu_N.dK x_D - mu_N.dV x_N

    (plain bilinear dots; the transposed block is taken as x_N . dK mu_D, so
    no transposed product is needed).  With P0 Neumann unknowns on every panel
    and P1 Dirichlet unknowns on every vertex the four forms come from one
    fused pass (bem_vkw_dk_forms); otherwise from four matrix-free products."""
    if not bm._is_transmission:
        raise RuntimeError("transmission_dA_contract needs a transmission BEMManager")
    C = bm.ctx
    ki = bm._k * bm._sqrt_ni
    qN, qD = bm.shape_func_degree_Neumann, bm.shape_func_degree_Dirichlet
    D0, N0 = bm.get_Dirichlet_offset_j(), bm.get_Neumann_offset_j()
    nD, nN = bm.M_Dirichlet, bm.N_Neumann
    mu = mu.to(torch.complex128).contiguous()
    x = x.to(torch.complex128).contiguous()
    xD, xN = x[D0:D0 + nD].contiguous(), x[N0:N0 + nN].contiguous()
    mD, mN = mu[D0:D0 + nD].contiguous(), mu[N0:N0 + nN].contiguous()
    mm = bm._core_manager._mesh_manager
    if bm.fuse_layers and qN == 0 and qD == 1 and nN == mm.get_num_of_panels() and nD == mm.get_num_of_vertices():
        # the four forms from one evaluation per unordered point pair
        d = _lib.Context.op_desc(_lib.OP_V, _lib.KernelType_TRANSMISSION, ki, 1.0, 0, 0, _lib.PREC_FP64, _lib.C128,
                                 bm._quadrature_flags)
        c = C.vkw_dk_forms(d, mN, xN, mD, xD)
        return complex((-c[3] - c[2] + c[1] - c[0]).item())
    flags = bm._quadrature_flags | _lib.FLAG_DK

    def dapply(op, dt, dr, v, n_out):
        y = torch.zeros(n_out, dtype=torch.complex128, device=x.device)
        d = _lib.Context.op_desc(op, _lib.KernelType_TRANSMISSION, ki, 1.0, dt, dr, _lib.PREC_FP64, _lib.C128, flags)
        C.apply(d, v, y, alpha=1.0, conj_x=False)
        return y

    s = -(mD * dapply(_lib.OP_W, qD, qD, xD, nD)).sum()
    s = s - (xN * dapply(_lib.OP_K, qN, qD, mD, nN)).sum()
    s = s + (mN * dapply(_lib.OP_K, qN, qD, xD, nN)).sum()
    s = s - (mN * dapply(_lib.OP_V, qN, qN, xN, nN)).sum()
    return complex(s.item())


def transmission_k_gradient(bm, grad_x: torch.Tensor) -> float:
    """dJ/dk_i of a real objective J(x) of the transmission solution x =
    A(k_i)^-1 b (b, the incident-field data, held fixed), given
    grad_x = dJ/dRe x + i dJ/dIm x: one transposed solve with the factors of
    BEMManager.solve, A^T mu = conj(grad_x), then -Re(mu^T dA/dk_i x)."""
    mu = grad_x.conj().to(bm._LU.dtype).contiguous().clone()
    bm.ctx.lu_solve_transpose(bm._LU, bm._ipiv, mu)
    return -transmission_dA_contract(bm, mu, bm.raw_solved).real
