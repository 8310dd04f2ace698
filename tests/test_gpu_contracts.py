"""C-ABI contracts of include/mgfe.h that the parity tests do not reach:
status codes without an info struct, no device memory left behind by failed
host-copy calls, the FP32 copy of the mixed solvers tied to the operator it
was assembled from, and the assembly sequence numbers."""
import ctypes as C

import numpy as np
import pytest

from morganapublic_amd import _lib as L
from morganapublic_amd import workloads as W

pytestmark = pytest.mark.gpu


def _case(ctx, n=8, order=1):
    p = W.config_problem(5 if order == 1 else 2, n=n)
    m = p.mesh
    ctx.mesh_set(m.xyz, m.tets, m.btri, m.bgeo)
    ctx.dofmap_build(order)
    ctx.dirichlet_set(list(p.dirichlet_geo))
    ctx.pattern_build(L.SYS_INTERIOR)
    fixed = ctx.dirichlet_mask()
    return p, np.ascontiguousarray(p.f[fixed == 0])


def _device_used():
    torch = pytest.importorskip("torch")
    free, total = torch.cuda.mem_get_info()
    return total - free


@pytest.mark.parametrize("mode", [L.SOLVE_F64, L.SOLVE_F32_REFINE, L.SOLVE_F32_RELIABLE])
def test_pcg_device_without_info_reports_nonconvergence(ctx, mode):
    # belosBlockCG::solve returns true whatever happens (belosBlockCG.hpp:75);
    # the C ABI returns MG_ERR_NOCONV also when the caller passes no info
    torch = pytest.importorskip("torch")
    p, b = _case(ctx)
    ctx.field_set(L.FIELD_STIFF, p.k)
    kw = dict(stiff_coeff=L.COEFF_P0, stiff_rule=1)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, **kw)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, prec=L.PREC_F32, **kw)
    db = torch.tensor(b, device="cuda")
    dx = torch.zeros_like(db)
    prm = L.SolverParams(mode, 1e-10, 0.0, 3, 1, 32, 0)
    rc = ctx._L.mg_pcg_solve_device(ctx.h, C.byref(prm), C.c_void_p(db.data_ptr()),
                                    C.c_void_p(dx.data_ptr()), None)
    assert rc == L.MG_ERR_NOCONV
    assert b"iteration limit" in ctx._L.mg_last_error(ctx.h)
    # converged solves still return MG_OK without an info struct
    prm = L.SolverParams(mode, 1e-10, 0.0, 100000, 30, 32, 0)
    rc = ctx._L.mg_pcg_solve_device(ctx.h, C.byref(prm), C.c_void_p(db.data_ptr()),
                                    C.c_void_p(dx.data_ptr()), None)
    assert rc == L.MG_OK
    rc = ctx._L.mg_pcg_solve(ctx.h, C.byref(L.SolverParams(mode, 1e-10, 0.0, 3, 1, 32, 0)),
                             b.ctypes.data_as(C.c_void_p), np.zeros_like(b).ctypes.data_as(C.c_void_p),
                             None)
    assert rc == L.MG_ERR_NOCONV


def test_failed_host_calls_release_their_device_memory(ctx):
    # a solve on a system whose FP64 values were never assembled fails before
    # or after allocating its temporaries; repeated calls must not grow the
    # device footprint
    _, b = _case(ctx, n=12)
    x = np.zeros_like(b)
    for _ in range(3):
        with pytest.raises(L.MgfeError) as ei:
            ctx.pcg(b)
        assert ei.value.code == L.MG_ERR_STATE
    ctx.synchronize()
    before = _device_used()
    for _ in range(50):
        with pytest.raises(L.MgfeError):
            ctx.pcg(b)
        with pytest.raises(L.MgfeError):
            ctx.spmv(L.SYS_INTERIOR, x)
        with pytest.raises(L.MgfeError):
            ctx.adjoint_grad_k(L.SYS_INTERIOR, x, x, rule=7)
    ctx.synchronize()
    assert _device_used() <= before


@pytest.mark.parametrize("mode", [L.SOLVE_F32_REFINE, L.SOLVE_F32_RELIABLE])
def test_fp32_copy_of_another_operator_is_not_used(ctx, mode):
    # FP64 values of operator A (k), then FP32 values of operator B (a mass
    # matrix): the mixed solvers must iterate on A rounded to FP32, and reach
    # A's FP64 solution
    p, b = _case(ctx, n=10)
    ctx.field_set(L.FIELD_STIFF, p.k)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, stiff_coeff=L.COEFF_P0, stiff_rule=1)
    x64, _ = ctx.pcg(b, mode=L.SOLVE_F64, tol=1e-12)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_MASS, mass_rule=4, prec=L.PREC_F32)
    x, info = ctx.pcg(b, mode=mode, tol=1e-10, max_iters=20000)
    assert info["status"] == L.MG_OK and info["relres"] <= 1e-10, info
    assert np.linalg.norm(x - x64) <= 1e-8 * np.linalg.norm(x64)
    # the same operator in both precisions is reused as assembled
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, stiff_coeff=L.COEFF_P0, stiff_rule=1,
                 prec=L.PREC_F32)
    s32 = ctx.assembly_seq(L.SYS_INTERIOR, L.PREC_F32)
    x, info = ctx.pcg(b, mode=mode, tol=1e-10)
    assert info["relres"] <= 1e-10
    assert ctx.assembly_seq(L.SYS_INTERIOR, L.PREC_F32) == s32


def test_assembly_sequence_numbers(ctx):
    p, _ = _case(ctx, n=4)
    assert ctx.assembly_seq(L.SYS_INTERIOR, L.PREC_F64) == 0
    assert ctx.assembly_seq(L.SYS_INTERIOR, L.PREC_F32) == 0
    ctx.field_set(L.FIELD_STIFF, p.k)
    kw = dict(stiff_coeff=L.COEFF_P0, stiff_rule=1)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, **kw)
    s1 = ctx.assembly_seq(L.SYS_INTERIOR, L.PREC_F64)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, prec=L.PREC_F32, **kw)
    s2 = ctx.assembly_seq(L.SYS_INTERIOR, L.PREC_F32)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, **kw)
    s3 = ctx.assembly_seq(L.SYS_INTERIOR, L.PREC_F64)
    assert 0 < s1 < s2 < s3
    with pytest.raises(L.MgfeError):
        ctx.assembly_seq(7, L.PREC_F64)
