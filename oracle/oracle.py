"""ctypes view of liboracle.so — TEST INFRASTRUCTURE ONLY.

The FP64 CPU restatement of the reference path (oracle.c).  Imported by the
tests, by __graft_entry__.smoke() and by the CPU-baseline leg of bench.py,
always as the checker / baseline, never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OP_STIFF = 1
OP_MASS = 2

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_I64 = C.c_int64
_I = C.c_int


def _declare(L):
    L.or_quad_rule.argtypes = [_I, _P, _P]
    L.or_quad_rule.restype = _I
    L.or_basis.argtypes = [_I, _P, _P, _P, _P, _P]
    L.or_global_grad.argtypes = [_I, _P, _P, _P, _P, _P]
    L.or_tet_jacobian.argtypes = [_P, _P]
    L.or_det3.argtypes = [_P]
    L.or_det3.restype = C.c_double
    L.or_inv3.argtypes = [_P, _P]
    L.or_element_matrix.argtypes = [_I, _I, _P, C.c_double, C.c_double, _I, _I, _P]
    L.or_element_matrix.restype = _I
    L.or_edges.argtypes = [_I64, _I64, _P, _P, _P]
    L.or_edges.restype = _I64
    L.or_elem_dofs.argtypes = [_I, _I64, _I64, _P, _P, _P]
    L.or_assemble.argtypes = [_I, _I, _I64, _P, _I64, _P, _P, _P, _I, _I,
                              C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_P), C.POINTER(_P),
                              C.POINTER(_P)]
    L.or_assemble.restype = _I
    L.or_load_vector.argtypes = [_I, _I64, _P, _I64, _P, _I, _P, _I, _P]
    L.or_load_vector.restype = _I
    L.or_restrict.argtypes = [_I64, _P, _P, _P, _P, C.POINTER(_I64), C.POINTER(_I64),
                              C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]
    L.or_restrict.restype = _I
    L.or_free.argtypes = [_P]
    L.or_spmv.argtypes = [_I64, _P, _P, _P, _P, _P]
    L.or_pcg.argtypes = [_I64, _P, _P, _P, _P, _P, C.c_double, _I, C.POINTER(_I),
                         C.POINTER(C.c_double)]
    L.or_pcg.restype = _I
    L.or_dirichlet_mask.argtypes = [_I, _I64, _I64, _P, _P, _I, _P, _I64, _P, _P]
    L.or_dual_fv.argtypes = [_I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P]
    L.or_bernoulli.argtypes = [C.c_double]
    L.or_bernoulli.restype = C.c_double
    L.or_dd_loss.argtypes = [_I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                             _P]
    L.or_dd_loss.restype = C.c_double
    L.or_dd_grad.argtypes = [_I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                             _P]
    L.or_adjoint_grad_k.argtypes = [_I, _I64, _P, _I64, _P, _P, _I, _P, _P, _P]
    L.or_accumulate.argtypes = [_I, _I, _P, _P, _P, _I64, _I64, _P, _P, _I, _I, _P, _P, _P]
    L.or_accumulate.restype = _I64
    L.or_pcg_fixed.argtypes = [_I64, _P, _P, _P, _P, _P, _I, _I, C.POINTER(C.c_double)]
    L.or_pcg_fixed.restype = C.c_double


def _p(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def quad_rule(npts: int):
    pts = np.zeros((15, 3))
    w = np.zeros(15)
    nq = lib().or_quad_rule(npts, _p(pts), _p(w))
    if nq == 0:
        raise ValueError("unknown rule %d" % npts)
    return pts[:nq].copy(), w[:nq].copy()


def basis(order: int, y):
    n = 4 if order == 1 else 10
    y = _c(y, np.float64)
    v, gx, gy, gz = (np.zeros(n) for _ in range(4))
    lib().or_basis(order, _p(y), _p(v), _p(gx), _p(gy), _p(gz))
    return v, np.stack([gx, gy, gz])


def global_grad(order: int, P, y):
    n = 4 if order == 1 else 10
    P, y = _c(P, np.float64), _c(y, np.float64)
    gx, gy, gz = (np.zeros(n) for _ in range(3))
    lib().or_global_grad(order, _p(P), _p(y), _p(gx), _p(gy), _p(gz))
    return np.stack([gx, gy, gz])


def jacobian(P):
    P = _c(P, np.float64)
    J = np.zeros(9)
    lib().or_tet_jacobian(_p(P), _p(J))
    return J.reshape(3, 3)


def element_matrix(ops: int, order: int, P, c_stiff=1.0, c_mass=1.0, rule_stiff=4, rule_mass=4):
    """Column-major element matrix returned as A[row=test, col=field]."""
    n = 4 if order == 1 else 10
    P = _c(P, np.float64)
    mat = np.zeros(n * n)
    if lib().or_element_matrix(ops, order, _p(P), c_stiff, c_mass, rule_stiff, rule_mass, _p(mat)):
        raise ValueError("bad rule")
    return mat.reshape(n, n).T.copy()


def edges(mesh):
    tets = _c(mesh.tets, np.int32)
    ee = np.zeros(6 * mesh.nT, dtype=np.int32)
    en = np.zeros(2 * 6 * mesh.nT, dtype=np.int32)
    E = lib().or_edges(mesh.nV, mesh.nT, _p(tets), _p(ee), _p(en))
    return ee.reshape(-1, 6), en[:2 * E].reshape(-1, 2).copy()


def elem_dofs(order: int, mesh, elem_edge=None):
    n = 4 if order == 1 else 10
    tets = _c(mesh.tets, np.int32)
    if order == 2 and elem_edge is None:
        elem_edge, _ = edges(mesh)
    ee = _c(elem_edge, np.int32) if elem_edge is not None else None
    d = np.zeros(n * mesh.nT, dtype=np.int32)
    lib().or_elem_dofs(order, mesh.nV, mesh.nT, _p(tets), _p(ee), _p(d))
    return d.reshape(-1, n)


def _take_csr(nrows, nnz, rp, ci, va):
    n, m = nrows.value, nnz.value
    rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int32)), shape=(n + 1,)).copy()
    colidx = np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_int32)), shape=(max(m, 1),))[:m].copy()
    vals = np.ctypeslib.as_array(C.cast(va, C.POINTER(C.c_double)), shape=(max(m, 1),))[:m].copy()
    for p in (rp, ci, va):
        lib().or_free(p)
    return rowptr, colidx, vals


def assemble(order: int, mesh, ops: int, c_stiff=None, c_mass=None, rule_stiff=4, rule_mass=4):
    """Full CSR (rowptr, colidx, vals) as Epetra's GlobalAssemble leaves it."""
    xyz, tets = _c(mesh.xyz, np.float64), _c(mesh.tets, np.int32)
    cs = _c(c_stiff, np.float64) if c_stiff is not None else None
    cm = _c(c_mass, np.float64) if c_mass is not None else None
    nrows, nnz = _I64(), _I64()
    rp, ci, va = _P(), _P(), _P()
    rc = lib().or_assemble(order, ops, mesh.nV, _p(xyz), mesh.nT, _p(tets), _p(cs), _p(cm),
                           rule_stiff, rule_mass, C.byref(nrows), C.byref(nnz), C.byref(rp),
                           C.byref(ci), C.byref(va))
    if rc:
        raise ValueError("or_assemble failed")
    return _take_csr(nrows, nnz, rp, ci, va)


LOAD_CONST, LOAD_P0, LOAD_NODAL = 0, 1, 2


def load_vector(order: int, mesh, kind: int, f, rule=4):
    """F_j = int f phi_j on the full dof numbering (buildEpetraVector over fnEL3dA)."""
    xyz, tets = _c(mesh.xyz, np.float64), _c(mesh.tets, np.int32)
    ndofs = mesh.nV if order == 1 else mesh.nV + edges(mesh)[1].shape[0]
    f = _c(np.atleast_1d(np.asarray(f, np.float64)), np.float64)
    F = np.zeros(ndofs)
    if lib().or_load_vector(order, mesh.nV, _p(xyz), mesh.nT, _p(tets), kind, _p(f), rule, _p(F)):
        raise ValueError("bad rule")
    return F


def restrict(rowptr, colidx, vals, fixed):
    rowptr, colidx = _c(rowptr, np.int32), _c(colidx, np.int32)
    vals, fixed = _c(vals, np.float64), _c(fixed, np.uint8)
    nrows, nnz = _I64(), _I64()
    rp, ci, va = _P(), _P(), _P()
    lib().or_restrict(rowptr.size - 1, _p(rowptr), _p(colidx), _p(vals), _p(fixed),
                      C.byref(nrows), C.byref(nnz), C.byref(rp), C.byref(ci), C.byref(va))
    return _take_csr(nrows, nnz, rp, ci, va)


def spmv(rowptr, colidx, vals, x):
    rowptr, colidx = _c(rowptr, np.int32), _c(colidx, np.int32)
    vals, x = _c(vals, np.float64), _c(x, np.float64)
    y = np.zeros(rowptr.size - 1)
    lib().or_spmv(rowptr.size - 1, _p(rowptr), _p(colidx), _p(vals), _p(x), _p(y))
    return y


def pcg(rowptr, colidx, vals, b, tol=1e-10, maxit=100000, x0=None):
    rowptr, colidx = _c(rowptr, np.int32), _c(colidx, np.int32)
    vals, b = _c(vals, np.float64), _c(b, np.float64)
    x = np.zeros_like(b) if x0 is None else _c(x0, np.float64).copy()
    it, rel = _I(), C.c_double()
    st = lib().or_pcg(rowptr.size - 1, _p(rowptr), _p(colidx), _p(vals), _p(b), _p(x), tol,
                      maxit, C.byref(it), C.byref(rel))
    return x, {"status": st, "iters": it.value, "relres": rel.value}


def dirichlet_mask(order: int, mesh, geo_list, edge_nodes=None):
    btri, bgeo = _c(mesh.btri, np.int32), _c(mesh.bgeo, np.int32)
    gl = _c(geo_list, np.int32)
    nE = 0
    en = None
    if order == 2:
        if edge_nodes is None:
            _, edge_nodes = edges(mesh)
        en = _c(edge_nodes, np.int32)
        nE = en.shape[0]
    N = mesh.nV + nE
    fixed = np.zeros(N, dtype=np.uint8)
    lib().or_dirichlet_mask(order, mesh.nV, btri.shape[0], _p(btri), _p(bgeo), gl.size, _p(gl),
                            nE, _p(en), _p(fixed))
    return fixed


def dual_fv(mesh, elem_edge=None, edge_nodes=None):
    if elem_edge is None:
        elem_edge, edge_nodes = edges(mesh)
    xyz, tets = _c(mesh.xyz, np.float64), _c(mesh.tets, np.int32)
    ee, en = _c(elem_edge, np.int32), _c(edge_nodes, np.int32)
    nE = en.shape[0]
    surf, nrm = np.zeros(nE), np.zeros(3 * nE)
    cvol, sig = np.zeros(mesh.nV), np.zeros(6 * mesh.nT)
    lib().or_dual_fv(mesh.nV, _p(xyz), mesh.nT, _p(tets), nE, _p(ee), _p(en), _p(surf), _p(nrm),
                     _p(cvol), _p(sig))
    return {"edge_surf": surf, "edge_n": nrm.reshape(-1, 3), "cell_volume": cvol,
            "sigma_el": sig.reshape(-1, 6)}


def bernoulli(x: float) -> float:
    return lib().or_bernoulli(float(x))


def dd_loss(mesh, dual, eps, rho, mu, phi, dens, fixed, elem_edge, edge_nodes):
    rp, rn = np.zeros(mesh.nV), np.zeros(mesh.nV)
    arrs = dict(xyz=_c(mesh.xyz, np.float64), tets=_c(mesh.tets, np.int32),
                ee=_c(elem_edge, np.int32), en=_c(edge_nodes, np.int32),
                sig=_c(dual["sigma_el"], np.float64), eps=_c(eps, np.float64),
                rho=_c(rho, np.float64), mu=_c(mu, np.float64), phi=_c(phi, np.float64),
                dens=_c(dens, np.float64), fixed=_c(fixed, np.uint8))
    loss = lib().or_dd_loss(mesh.nV, _p(arrs["xyz"]), mesh.nT, _p(arrs["tets"]),
                            arrs["en"].shape[0], _p(arrs["ee"]), _p(arrs["en"]), _p(arrs["sig"]),
                            _p(arrs["eps"]), _p(arrs["rho"]), _p(arrs["mu"]), _p(arrs["phi"]),
                            _p(arrs["dens"]), _p(arrs["fixed"]), _p(rp), _p(rn))
    return loss, rp, rn


def dd_grad(mesh, dual, eps, rho, mu, phi, dens, fixed, elem_edge, edge_nodes):
    ge, gm = np.zeros(mesh.nT), np.zeros(mesh.nT)
    arrs = dict(xyz=_c(mesh.xyz, np.float64), tets=_c(mesh.tets, np.int32),
                ee=_c(elem_edge, np.int32), en=_c(edge_nodes, np.int32),
                sig=_c(dual["sigma_el"], np.float64), eps=_c(eps, np.float64),
                rho=_c(rho, np.float64), mu=_c(mu, np.float64), phi=_c(phi, np.float64),
                dens=_c(dens, np.float64), fixed=_c(fixed, np.uint8))
    lib().or_dd_grad(mesh.nV, _p(arrs["xyz"]), mesh.nT, _p(arrs["tets"]), arrs["en"].shape[0],
                     _p(arrs["ee"]), _p(arrs["en"]), _p(arrs["sig"]), _p(arrs["eps"]),
                     _p(arrs["rho"]), _p(arrs["mu"]), _p(arrs["phi"]), _p(arrs["dens"]),
                     _p(arrs["fixed"]), _p(ge), _p(gm))
    return ge, gm


def adjoint_grad_k(order: int, mesh, dofs, u, lam, rule=4):
    xyz, tets = _c(mesh.xyz, np.float64), _c(mesh.tets, np.int32)
    d = _c(dofs, np.int32)
    u, lam = _c(u, np.float64), _c(lam, np.float64)
    g = np.zeros(mesh.nT)
    lib().or_adjoint_grad_k(order, mesh.nV, _p(xyz), mesh.nT, _p(tets), _p(d), rule, _p(u),
                            _p(lam), _p(g))
    return g


def accumulate(order: int, mesh, ops: int, sdofs, rowptr, colidx, c_stiff=None, c_mass=None,
               rule_stiff=4, rule_mass=4, threads=8):
    """Element matrices of the whole mesh added into a given CSR pattern
    (or_accumulate), on `threads` host threads that each take a contiguous
    block of elements into their own value array; the blocks are summed in
    block order.  sdofs: (nT, n) system index per element basis, -1 where
    the dof is not in the system.  Returns (vals, number of element pairs
    missing from the pattern)."""
    import concurrent.futures as cf
    xyz, tets = _c(mesh.xyz, np.float64), _c(mesh.tets, np.int32)
    sd = _c(sdofs, np.int32)
    rp, ci = _c(rowptr, np.int32), _c(colidx, np.int32)
    cs = _c(c_stiff, np.float64) if c_stiff is not None else None
    cm = _c(c_mass, np.float64) if c_mass is not None else None
    nT = mesh.nT
    threads = max(1, min(threads, nT))
    cuts = [nT * t // threads for t in range(threads + 1)]
    L = lib()

    def run(t):
        v = np.zeros(ci.size)
        miss = L.or_accumulate(ops, order, _p(xyz), _p(tets), _p(sd), cuts[t], cuts[t + 1],
                               _p(cs), _p(cm), rule_stiff, rule_mass, _p(rp), _p(ci), _p(v))
        if miss < 0:
            raise ValueError("bad rule")
        return v, miss

    with cf.ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(run, range(threads)))
    vals = parts[0][0]
    for v, _ in parts[1:]:
        vals += v
    return vals, sum(m for _, m in parts)


def pcg_fixed(rowptr, colidx, vals, b, iters, threads=1):
    """`iters` Jacobi-PCG iterations from x = 0 on `threads` host threads
    (or_pcg_fixed); returns (x, seconds per iteration, explicit relres)."""
    rp, ci, va = _c(rowptr, np.int32), _c(colidx, np.int32), _c(vals, np.float64)
    b = _c(b, np.float64)
    x = np.zeros_like(b)
    rel = C.c_double()
    t = lib().or_pcg_fixed(b.size, _p(rp), _p(ci), _p(va), _p(b), _p(x), iters, threads,
                           C.byref(rel))
    return x, t, rel.value
