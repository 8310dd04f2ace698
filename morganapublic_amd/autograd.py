"""torch.autograd harness around the solve and its adjoint.

u = A(k)^{-1} f on the interior system, with A(k) the stiffness matrix of a
per-element coefficient k.  backward() runs the adjoint solve A lam = dJ/du
(A is symmetric) and the adjoint element kernel dJ/dk_e = -lam_e^T Khat_e u_e,
both in libmgfe.so; torch only holds the device buffers.

The context holds one interior matrix at a time.  forward() records the
assembly sequence number of the matrix it solved with and keeps its k;
backward() re-assembles A(k) from that k when the context has assembled
anything since (another forward pass, for instance), so the adjoint solve
always uses the forward pass's own operator.
"""
from __future__ import annotations

import torch

from . import _lib as L


class DiffusionSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k: torch.Tensor, f: torch.Tensor, lib_ctx: L.Context, mode: int, rule: int,
                tol: float):
        if k.device.type != "cuda" or f.device.type != "cuda":
            raise ValueError("k and f must live on the GPU")
        k = k.detach().to(torch.float64).contiguous()
        f = f.detach().to(torch.float64).contiguous()
        torch.cuda.synchronize()
        _assemble(lib_ctx, k, mode, rule)
        u = torch.zeros_like(f)
        info = lib_ctx.pcg_device(f.data_ptr(), u.data_ptr(), mode=mode, tol=tol)
        lib_ctx.synchronize()
        ctx.lib_ctx, ctx.mode, ctx.rule, ctx.tol = lib_ctx, mode, rule, tol
        ctx.info = info
        ctx.seq = _seqs(lib_ctx, mode)
        ctx.save_for_backward(u, k)
        return u

    @staticmethod
    def backward(ctx, grad_u: torch.Tensor):
        u, k = ctx.saved_tensors
        g = grad_u.detach().to(torch.float64).contiguous()
        lam = torch.zeros_like(g)
        torch.cuda.synchronize()
        lib = ctx.lib_ctx
        if _seqs(lib, ctx.mode) != ctx.seq:
            # the context was re-assembled after this forward pass: restore A(k)
            _assemble(lib, k, ctx.mode, ctx.rule)
        lib.pcg_device(g.data_ptr(), lam.data_ptr(), mode=ctx.mode, tol=ctx.tol)
        dk = torch.zeros(lib.nT, dtype=torch.float64, device=u.device)
        lib.adjoint_grad_k_device(L.SYS_INTERIOR, u.data_ptr(), lam.data_ptr(), dk.data_ptr(),
                                  rule=ctx.rule)
        lib.synchronize()
        return dk, lam, None, None, None, None


def _assemble(lib_ctx: L.Context, k: torch.Tensor, mode: int, rule: int) -> None:
    """A(k) on the interior system in FP64 (and FP32 for the mixed solvers)."""
    lib_ctx.field_set_device(L.FIELD_STIFF, k.data_ptr())
    lib_ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, stiff_coeff=L.COEFF_P0, stiff_rule=rule)
    if mode != L.SOLVE_F64:
        lib_ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, stiff_coeff=L.COEFF_P0,
                         stiff_rule=rule, prec=L.PREC_F32)


def _seqs(lib_ctx: L.Context, mode: int) -> tuple:
    """Assembly sequence numbers of the interior values a solve in `mode` reads."""
    s64 = lib_ctx.assembly_seq(L.SYS_INTERIOR, L.PREC_F64)
    return (s64,) if mode == L.SOLVE_F64 else (s64, lib_ctx.assembly_seq(L.SYS_INTERIOR, L.PREC_F32))


def diffusion_solve(k, f, lib_ctx, mode=L.SOLVE_F64, rule=4, tol=1e-10):
    return DiffusionSolve.apply(k, f, lib_ctx, mode, rule, tol)
