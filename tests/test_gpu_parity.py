"""GPU parity: libmgfe.so (through its C ABI) against the FP64 oracle.

Integer/index results (edges, dofs, Dirichlet set, CSR pattern, slot map) must
be bit-exact.  FP64 matrix entries: |a - a_oracle| <= 1e-12 * max|a_oracle| of
the same row.  FP32 entries: <= 2e-6 * max|row|.  Solutions: explicit FP64
relative residual <= 1e-10 and agreement with a direct solve.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from morganapublic_amd import _lib as L
from morganapublic_amd import workloads as W

pytestmark = pytest.mark.gpu

ALL_FACES = [1, 2, 3, 4, 5, 6]


def setup(ctx, mesh, order, dirichlet=ALL_FACES):
    ctx.mesh_set(mesh.xyz, mesh.tets, mesh.btri, mesh.bgeo)
    ctx.dofmap_build(order)
    ctx.pattern_build(L.SYS_FULL)
    if dirichlet is not None:
        ctx.dirichlet_set(dirichlet)
        ctx.pattern_build(L.SYS_INTERIOR)


def row_scale(rp, va):
    """max |a| of each entry's row; a row that is zero throughout (P2 vertex
    rows under the 1-point rule: the vertex basis gradients vanish at the
    centroid) is measured against the largest entry of the matrix."""
    s = np.zeros(va.size)
    for r in range(rp.size - 1):
        if rp[r + 1] > rp[r]:
            s[rp[r]:rp[r + 1]] = np.abs(va[rp[r]:rp[r + 1]]).max()
    s[s == 0] = np.abs(va).max()
    return s


@pytest.mark.parametrize("order,n", [(1, 2), (1, 5), (2, 2), (2, 4)])
def test_numbering_bit_exact(ctx, order, n):
    m = W.kuhn_cube(n, jitter=0.2, seed=n)
    setup(ctx, m, order)
    ee, en = O.edges(m)
    gee, gen = ctx.edges()
    np.testing.assert_array_equal(gen, en)
    np.testing.assert_array_equal(gee, ee)
    np.testing.assert_array_equal(ctx.elem_dofs(), O.elem_dofs(order, m, ee))
    np.testing.assert_array_equal(ctx.dirichlet_mask(), O.dirichlet_mask(order, m, ALL_FACES, en))


@pytest.mark.parametrize("order,n", [(1, 3), (1, 6), (2, 2), (2, 4)])
def test_pattern_bit_exact(ctx, order, n):
    m = W.kuhn_cube(n, jitter=0.2, seed=1)
    setup(ctx, m, order)
    rp, ci, va = O.assemble(order, m, O.OP_STIFF, rule_stiff=4)
    grp, gci = ctx.pattern(L.SYS_FULL)
    np.testing.assert_array_equal(grp, rp)
    np.testing.assert_array_equal(gci, ci)
    cnt = W.kuhn_counts(n)
    assert gci.size == cnt["nnz_p%d" % order]
    fixed = O.dirichlet_mask(order, m, ALL_FACES)
    rpi, cii, _ = O.restrict(rp, ci, va, fixed)
    grpi, gcii = ctx.pattern(L.SYS_INTERIOR)
    np.testing.assert_array_equal(grpi, rpi)
    np.testing.assert_array_equal(gcii, cii)
    assert gcii.size == cnt["nnz_p%d_int" % order]


@pytest.mark.parametrize("order", [1, 2])
def test_slot_map_consistent(ctx, order):
    m = W.kuhn_cube(3, jitter=0.2, seed=2)
    setup(ctx, m, order)
    rp, ci = ctx.pattern(L.SYS_FULL)
    slot = ctx.slot_map(L.SYS_FULL)  # [e, i (field/column), j (test/row)]
    d = ctx.elem_dofs()
    nl = d.shape[1]
    row_of = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    cols = d[:, :, None].repeat(nl, axis=2)   # column dof of basis i
    rows = d[:, None, :].repeat(nl, axis=1)   # row dof of basis j
    assert slot.min() >= 0
    np.testing.assert_array_equal(ci[slot], cols)
    np.testing.assert_array_equal(row_of[slot], rows)


CASES = [
    # order, terms, stiff rule, mass rule, coefficient on stiffness
    (1, O.OP_STIFF | O.OP_MASS, 1, 4, False),  # config 1 operator: K + M, k = 1
    (1, O.OP_STIFF, 4, 4, True),
    (1, O.OP_MASS, 4, 15, False),
    (2, O.OP_STIFF, 4, 4, True),               # config 2 operator
    (2, O.OP_STIFF, 11, 4, True),
    (2, O.OP_STIFF | O.OP_MASS, 15, 15, True),
    (2, O.OP_MASS, 4, 11, False),
    # rules not exact on the integrand: the quadrature-table form of the rows
    (2, O.OP_STIFF, 1, 4, True),
    (2, O.OP_MASS, 4, 4, False),
    (2, O.OP_STIFF | O.OP_MASS, 1, 4, True),
    (1, O.OP_MASS, 4, 1, False),
]


@pytest.mark.parametrize("order,terms,rs,rm,coef", CASES)
@pytest.mark.parametrize("sysid", [L.SYS_FULL, L.SYS_INTERIOR])
def test_assembly_fp64_matches_oracle(ctx, order, terms, rs, rm, coef, sysid):
    m = W.kuhn_cube(4, jitter=0.2, seed=7)
    k = np.exp(W.stream_normal(0, W.STREAM_K, m.nT)) if coef else None
    setup(ctx, m, order)
    if coef:
        ctx.field_set(L.FIELD_STIFF, k)
    ctx.assemble(sysid, terms, stiff_coeff=L.COEFF_P0 if coef else L.COEFF_CONST,
                 stiff_rule=rs, mass_rule=rm)
    got = ctx.values(sysid)
    rp, ci, va = O.assemble(order, m, terms, c_stiff=k, rule_stiff=rs, rule_mass=rm)
    if sysid == L.SYS_INTERIOR:
        rp, ci, va = O.restrict(rp, ci, va, O.dirichlet_mask(order, m, ALL_FACES))
    err = np.abs(got - va) / row_scale(rp, va)
    assert err.max() <= 1e-12, err.max()


@pytest.mark.parametrize("order,rs", [(1, 1), (2, 11), (2, 1)])
def test_assembly_fp32(ctx, order, rs):
    m = W.kuhn_cube(4, jitter=0.2, seed=7)
    k = np.exp(W.stream_normal(0, W.STREAM_K, m.nT))
    setup(ctx, m, order)
    ctx.field_set(L.FIELD_STIFF, k)
    ctx.assemble(L.SYS_FULL, L.TERM_STIFF, stiff_coeff=L.COEFF_P0, stiff_rule=rs,
                 prec=L.PREC_F32)
    got = ctx.values(L.SYS_FULL, prec=L.PREC_F32).astype(np.float64)
    rp, ci, va = O.assemble(order, m, O.OP_STIFF, c_stiff=k, rule_stiff=rs)
    err = np.abs(got - va) / row_scale(rp, va)
    assert err.max() <= 2e-6, err.max()


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("prec", [L.PREC_F64, L.PREC_F32])
@pytest.mark.parametrize("sysid", [L.SYS_FULL, L.SYS_INTERIOR])
@pytest.mark.parametrize("fields", [True, False])
def test_assembly_coefficients(ctx, order, prec, sysid, fields):
    """K + M with P0 fields on both terms, or with non-unit constants."""
    m = W.kuhn_cube(4, jitter=0.2, seed=8)
    if fields:
        ks = np.exp(W.stream_normal(0, W.STREAM_K, m.nT))
        km = 0.5 + W.stream_uniform(0, W.STREAM_RHO, m.nT)
        kw = dict(stiff_coeff=L.COEFF_P0, mass_coeff=L.COEFF_P0)
    else:
        ks, km = np.full(m.nT, 2.5), np.full(m.nT, 0.75)
        kw = dict(stiff_const=2.5, mass_const=0.75)
    setup(ctx, m, order)
    if fields:
        ctx.field_set(L.FIELD_STIFF, ks)
        ctx.field_set(L.FIELD_MASS, km)
    ctx.assemble(sysid, L.TERM_STIFF | L.TERM_MASS, stiff_rule=4, mass_rule=15, prec=prec, **kw)
    got = ctx.values(sysid, prec=prec).astype(np.float64)
    rp, ci, va = O.assemble(order, m, O.OP_STIFF | O.OP_MASS, c_stiff=ks, c_mass=km,
                            rule_stiff=4, rule_mass=15)
    if sysid == L.SYS_INTERIOR:
        rp, ci, va = O.restrict(rp, ci, va, O.dirichlet_mask(order, m, ALL_FACES))
    err = np.abs(got - va) / row_scale(rp, va)
    assert err.max() <= (1e-12 if prec == L.PREC_F64 else 2e-6), err.max()


def test_assembly_deterministic(ctx):
    m = W.kuhn_cube(8, jitter=0.2, seed=3)
    setup(ctx, m, 2)
    ctx.field_set(L.FIELD_STIFF, np.exp(W.stream_normal(0, W.STREAM_K, m.nT)))
    runs = []
    for _ in range(3):
        ctx.assemble(L.SYS_FULL, L.TERM_STIFF, stiff_coeff=L.COEFF_P0, stiff_rule=11)
        runs.append(ctx.values(L.SYS_FULL))
    assert all(np.array_equal(runs[0], r) for r in runs[1:])


# (vertex lanes per row, edge rows per workgroup, edge lanes per row, 16-byte
# P2 records instead of the 12-byte form)
SHAPES = [("1", "64", "1", "0"), ("2", "128", "2", "0"), ("4", "256", "1", "0"),
          ("4", "256", "2", "0"), ("4", "256", "2", "1"), ("1", "64", "1", "1")]


def _row_scaled_diff(ctx, sysid, a, b):
    rp, _ = ctx.pattern(sysid)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    scale = np.zeros(rp.size - 1)
    np.maximum.at(scale, rows, np.abs(a.astype(np.float64)))
    scale[scale == 0] = 1.0
    return float((np.abs(a.astype(np.float64) - b) / scale[rows]).max())


@pytest.mark.parametrize("order,rs", [(1, 1), (2, 11), (2, 15)])
@pytest.mark.parametrize("prec", [L.PREC_F64, L.PREC_F32])
@pytest.mark.parametrize("sysid", [L.SYS_FULL, L.SYS_INTERIOR])
def test_assembly_launch_shapes_bit_identical(ctx, monkeypatch, order, rs, prec, sysid):
    # lanes per row, rows per workgroup and the record form change the
    # schedule and the bytes read, never the order in which a row receives
    # its incidences; FP64 P2 rows with vertex patches (12-byte form off)
    # form the element factors in the row's relabelled frame, the 16-byte
    # form reads them from the factor kernel: the two agree to rounding
    m = W.kuhn_cube(7, jitter=0.2, seed=5)
    k = np.exp(W.stream_normal(0, W.STREAM_K, m.nT))
    out = {}
    for vl, er, el, r16 in SHAPES:
        monkeypatch.setenv("MGFE_VERTEX_LPR", vl)
        monkeypatch.setenv("MGFE_VERTEX_LPR32", vl)
        monkeypatch.setenv("MGFE_EDGE_ROWS", er)
        monkeypatch.setenv("MGFE_EDGE_LPR", el)
        monkeypatch.setenv("MGFE_REC16", r16)
        setup(ctx, m, order)
        ctx.field_set(L.FIELD_STIFF, k)
        ctx.assemble(sysid, L.TERM_STIFF | L.TERM_MASS, stiff_coeff=L.COEFF_P0,
                     stiff_rule=rs, mass_rule=15, prec=prec)
        out.setdefault(r16, []).append(ctx.values(sysid, prec=prec))
    for group in out.values():
        assert all(np.array_equal(group[0], o) for o in group[1:])
    if order == 1 or prec == L.PREC_F32:
        assert np.array_equal(out["0"][0], out["1"][0])
    else:
        assert _row_scaled_diff(ctx, sysid, out["0"][0], out["1"][0]) <= 1e-14


@pytest.mark.parametrize("terms,rs,rm", [("stiff", 11, 15), ("both", 11, 15), ("both", 1, 4)])
@pytest.mark.parametrize("prec", [L.PREC_F64, L.PREC_F32])
@pytest.mark.parametrize("sysid", [L.SYS_FULL, L.SYS_INTERIOR])
def test_edge_record_forms_bit_identical(ctx, monkeypatch, terms, rs, rm, prec, sysid):
    # P2 edge-dof rows read 12-byte records that carry their patch words
    # (5-bit positions) or 12-byte records plus separate 4-byte patch words
    # (7-bit positions), and every patched row reads its row data from one
    # packed word or from the separate row arrays (MGFE_PROW=0): the same
    # positions, patch indices and ranges, so the same bits, for every
    # lanes-per-row shape; the oracle pins the default form
    m = W.kuhn_cube(6, jitter=0.2, seed=23)
    k = np.exp(W.stream_normal(0, W.STREAM_K, m.nT))
    t = L.TERM_STIFF | (L.TERM_MASS if terms == "both" else 0)
    out = {}
    for e5, er in (("1", "1"), ("1", "0"), ("0", "1")):
        for lpr in ("2", "1"):
            monkeypatch.setenv("MGFE_EDGE5", e5)
            monkeypatch.setenv("MGFE_PROW", er)
            monkeypatch.setenv("MGFE_EDGE_LPR", lpr)
            setup(ctx, m, 2)
            ctx.field_set(L.FIELD_STIFF, k)
            ctx.assemble(sysid, t, stiff_coeff=L.COEFF_P0, stiff_rule=rs, mass_rule=rm, prec=prec)
            out[(e5, er, lpr)] = ctx.values(sysid, prec=prec)
    ref = out[("1", "1", "2")]
    for key, v in out.items():
        assert np.array_equal(ref, v), key
    if prec == L.PREC_F64 and terms == "stiff":
        rp, ci = ctx.pattern(sysid)
        orp, oci, ova = O.assemble(2, m, O.OP_STIFF, c_stiff=k, rule_stiff=4)
        if sysid == L.SYS_INTERIOR:
            orp, oci, ova = O.restrict(orp, oci, ova, O.dirichlet_mask(2, m, list(ALL_FACES)))
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
        assert np.abs(ref - ova).max() <= 1e-12 * np.abs(ova).max()


@pytest.mark.parametrize("order,n", [(1, 9), (1, 14), (2, 5), (2, 7)])
@pytest.mark.parametrize("terms,rs,rm", [("stiff", 4, 15), ("both", 1, 4)])
@pytest.mark.parametrize("prec", [L.PREC_F64, L.PREC_F32])
@pytest.mark.parametrize("sysid", [L.SYS_FULL, L.SYS_INTERIOR])
def test_ell_records_bit_identical(ctx, monkeypatch, order, n, terms, rs, rm, prec, sysid):
    # patched row blocks read their incidence records from an ELL copy laid
    # out for their lanes per row and their patch vertex ids from a padded
    # copy (MGFE_ELL=2: P1 and P2 systems, 1: P1 systems, 0: neither; other
    # lanes per row: the CSR order), P1 records in the 8-byte form at four
    # lanes per row (MGFE_P1Q=0: the 12-byte form): the same records and
    # vertices in the same order, so the same bits
    m = W.kuhn_cube(n, jitter=0.2, seed=37)
    k = np.exp(W.stream_normal(0, W.STREAM_K, m.nT))
    t = L.TERM_STIFF | (L.TERM_MASS if terms == "both" else 0)
    if order == 2 and rs == 4:
        rs = 11
    out = []
    # (MGFE_P1Q_SPAN=2: element offsets past 3 do not fit the 8-byte form,
    # so the copy falls back to 12-byte records after trying it; "rows": the
    # separate row arrays instead of the packed row words)
    for ell, lpr, q in (("0", None, "1"), ("2", None, "1"), ("1", None, "1"), ("1", None, "0"),
                        ("2", "1", "1"), ("2", "2", "0"), ("1", None, "span"),
                        ("1", None, "rows")):
        monkeypatch.setenv("MGFE_ELL", ell)
        monkeypatch.setenv("MGFE_PROW", "0" if q == "rows" else "1")
        monkeypatch.setenv("MGFE_P1Q", "0" if q == "0" else "1")
        if q == "span":
            monkeypatch.setenv("MGFE_P1Q_SPAN", "2")
        else:
            monkeypatch.delenv("MGFE_P1Q_SPAN", raising=False)
        for var in ("MGFE_VERTEX_LPR", "MGFE_VERTEX_LPR32", "MGFE_EDGE_LPR"):
            if lpr is None:
                monkeypatch.delenv(var, raising=False)
            else:
                monkeypatch.setenv(var, lpr)
        setup(ctx, m, order)
        ctx.field_set(L.FIELD_STIFF, k)
        ctx.assemble(sysid, t, stiff_coeff=L.COEFF_P0, stiff_rule=rs, mass_rule=rm, prec=prec)
        out.append(ctx.values(sysid, prec=prec))
    for o in out[1:]:
        assert np.array_equal(out[0], o)
    if prec == L.PREC_F64 and terms == "stiff":
        orp, oci, ova = O.assemble(order, m, O.OP_STIFF, c_stiff=k, rule_stiff=4)
        if sysid == L.SYS_INTERIOR:
            orp, oci, ova = O.restrict(orp, oci, ova, O.dirichlet_mask(order, m, ALL_FACES))
        assert (np.abs(out[1] - ova) / row_scale(orp, ova)).max() <= 1e-12


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("prec", [L.PREC_F64, L.PREC_F32])
@pytest.mark.parametrize("sysid", [L.SYS_FULL, L.SYS_INTERIOR])
def test_patch_overflow_bit_identical(ctx, monkeypatch, order, prec, sysid):
    # a P2 edge-row launch stages at most a fixed number of each block's
    # patch vertices in LDS and reads the coordinates of the others from
    # memory: the same coordinates, so the same bits (vertex-row launches
    # stage every patch in full and ignore the limit)
    m = W.kuhn_cube(9 if order == 1 else 5, jitter=0.2, seed=29)
    k = np.exp(W.stream_normal(0, W.STREAM_K, m.nT))
    out = []
    for cap in (None, "40", "1"):
        if cap is None:
            monkeypatch.delenv("MGFE_PATCH_STAGE", raising=False)
        else:
            monkeypatch.setenv("MGFE_PATCH_STAGE", cap)
        setup(ctx, m, order)
        ctx.field_set(L.FIELD_STIFF, k)
        ctx.assemble(sysid, L.TERM_STIFF | L.TERM_MASS, stiff_coeff=L.COEFF_P0,
                     stiff_rule=4 if order == 2 else 1, mass_rule=15, prec=prec)
        out.append(ctx.values(sysid, prec=prec))
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


# record forms: (MGFE_PATCH, MGFE_PATCH_MAX): forms without patches;
# patched records without patches; patches everywhere; patches on some row
# blocks only
PATCHES = [("0", None), ("2", None), ("1", None), ("1", "90"), ("1", "160"), ("1", "260")]


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("terms,rs,rm", [("stiff", 1, 4), ("both", 1, 4), ("both", 4, 1)])
@pytest.mark.parametrize("prec", [L.PREC_F64, L.PREC_F32])
@pytest.mark.parametrize("sysid", [L.SYS_FULL, L.SYS_INTERIOR])
def test_vertex_patches(ctx, monkeypatch, order, terms, rs, rm, prec, sysid):
    # a row block's element vertices staged in LDS (vertex patch) instead of
    # gathered per incidence.  P1 rows form the factors in the relabelled
    # frame with or without a patch: every form gives the same bits.  FP64
    # P2 rows with a patch form them in the relabelled frame, P2 rows without
    # one (and FP32 P2 rows) read the factor kernel's: equal bits between
    # the forms without patches, agreement to rounding across, and the same
    # bits for every launch shape of a form
    m = W.kuhn_cube(13 if order == 1 else 6, jitter=0.2, seed=21)
    k = np.exp(W.stream_normal(0, W.STREAM_K, m.nT))
    t = L.TERM_STIFF | (L.TERM_MASS if terms == "both" else 0)
    if order == 2:  # exact rules, or both inexact (table form)
        rs, rm = (11, 15) if rs == 1 else (1, 4)
    out = []
    for mode, umax in PATCHES:
        monkeypatch.setenv("MGFE_PATCH", mode)
        if umax is None:
            monkeypatch.delenv("MGFE_PATCH_MAX", raising=False)
        else:
            monkeypatch.setenv("MGFE_PATCH_MAX", umax)
        setup(ctx, m, order)
        ctx.field_set(L.FIELD_STIFF, k)
        ctx.assemble(sysid, t, stiff_coeff=L.COEFF_P0, stiff_rule=rs, mass_rule=rm, prec=prec)
        out.append(ctx.values(sysid, prec=prec))
        if mode == "1" and umax is None:
            for vl in ("1", "2"):
                monkeypatch.setenv("MGFE_VERTEX_LPR", vl)
                monkeypatch.setenv("MGFE_VERTEX_LPR32", vl)
                monkeypatch.setenv("MGFE_EDGE_LPR", vl)
                setup(ctx, m, order)
                ctx.field_set(L.FIELD_STIFF, k)
                ctx.assemble(sysid, t, stiff_coeff=L.COEFF_P0, stiff_rule=rs, mass_rule=rm,
                             prec=prec)
                assert np.array_equal(out[-1], ctx.values(sysid, prec=prec))
            for var in ("MGFE_VERTEX_LPR", "MGFE_VERTEX_LPR32", "MGFE_EDGE_LPR"):
                monkeypatch.delenv(var, raising=False)
    assert np.array_equal(out[0], out[1])
    for o in out[2:]:
        if order == 1 or prec == L.PREC_F32:
            assert np.array_equal(out[0], o)
        else:
            assert _row_scaled_diff(ctx, sysid, out[0], o) <= 1e-14


@pytest.mark.parametrize("order", [1, 2])
def test_spmv_matches_oracle(ctx, order):
    m = W.kuhn_cube(5, jitter=0.2, seed=9)
    setup(ctx, m, order)
    ctx.assemble(L.SYS_FULL, L.TERM_STIFF | L.TERM_MASS, stiff_rule=4, mass_rule=15)
    rp, ci = ctx.pattern(L.SYS_FULL)
    va = ctx.values(L.SYS_FULL)
    x = 2 * W.stream_uniform(0, 3, rp.size - 1) - 1
    y = ctx.spmv(L.SYS_FULL, x)
    yr = O.spmv(rp, ci, va, x)
    assert np.abs(y - yr).max() <= 1e-13 * np.abs(yr).max()
    y32 = None
    ctx.assemble(L.SYS_FULL, L.TERM_STIFF | L.TERM_MASS, stiff_rule=4, mass_rule=15,
                 prec=L.PREC_F32)
    y32 = ctx.spmv(L.SYS_FULL, x, prec=L.PREC_F32)
    assert np.abs(y32 - yr).max() <= 1e-5 * np.abs(yr).max()


def _solve_case(ctx, cfg, n):
    p = W.config_problem(cfg, n=n)
    setup(ctx, p.mesh, p.order, list(p.dirichlet_geo))
    if p.k is not None:
        ctx.field_set(L.FIELD_STIFF, p.k)
    terms = L.TERM_STIFF | (L.TERM_MASS if p.mass else 0)
    coeff = L.COEFF_P0 if p.k is not None else L.COEFF_CONST
    ctx.assemble(L.SYS_INTERIOR, terms, stiff_coeff=coeff, stiff_rule=1 if p.order == 1 else 11,
                 mass_rule=4)
    fixed = ctx.dirichlet_mask()
    b = p.f[fixed == 0]
    return p, b


@pytest.mark.parametrize("cfg,n", [(2, 12), (5, 24)])
@pytest.mark.parametrize("rows", ["64", "128"])
def test_spmv_column_forms_bit_identical(ctx, monkeypatch, cfg, n, rows):
    # the 16-bit clustered column stream decodes to the int32 columns (at
    # these sizes P2 vertex rows span several clusters: vertex and edge dofs),
    # so SpMV and every PCG iterate agree bit for bit; both row-block shapes
    # (64 rows x 4 lanes, 128 rows x 2 lanes) against the oracle
    monkeypatch.setenv("MGFE_SPMV_ROWS", rows)
    out = []
    for col32 in ("0", "1"):
        monkeypatch.setenv("MGFE_SPMV_COL32", col32)
        p, b = _solve_case(ctx, cfg, n)
        ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, stiff_coeff=L.COEFF_P0,
                     stiff_rule=1 if p.order == 1 else 11, prec=L.PREC_F32)
        x = 2 * W.stream_uniform(0, 3, b.size) - 1
        y = ctx.spmv(L.SYS_INTERIOR, x)
        y32 = ctx.spmv(L.SYS_INTERIOR, x, prec=L.PREC_F32)
        x64, i64 = ctx.pcg(b, mode=L.SOLVE_F64, tol=1e-10)
        xr, ir = ctx.pcg(b, mode=L.SOLVE_F32_RELIABLE, tol=1e-10)
        assert i64["relres"] <= 1e-10 and ir["relres"] <= 1e-10
        out.append((y, y32, x64, xr, i64["iters"], ir["iters"]))
    rp, ci = ctx.pattern(L.SYS_INTERIOR)
    va = ctx.values(L.SYS_INTERIOR)
    assert np.abs(out[0][0] - O.spmv(rp, ci, va, x)).max() <= 1e-13 * np.abs(out[0][0]).max()
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("cfg,n", [(1, 20), (2, 6), (5, 10)])
@pytest.mark.parametrize("mode", [L.SOLVE_F64, L.SOLVE_F32_REFINE, L.SOLVE_F32_RELIABLE])
def test_pcg_solution(ctx, cfg, n, mode):
    p, b = _solve_case(ctx, cfg, n)
    x, info = ctx.pcg(b, mode=mode, tol=1e-10)
    assert info["status"] == 0 and info["relres"] <= 1e-10, info
    rp, ci = ctx.pattern(L.SYS_INTERIOR)
    va = ctx.values(L.SYS_INTERIOR)
    N = rp.size - 1
    A = sp.csr_matrix((va, ci, rp), shape=(N, N))
    rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert rel <= 1.5e-10
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7
    if mode != L.SOLVE_F64:
        assert info["outer"] >= 1


def test_pcg_matches_oracle_pcg(ctx):
    p, b = _solve_case(ctx, 1, 12)
    x, info = ctx.pcg(b, mode=L.SOLVE_F64, tol=1e-10)
    rp, ci = ctx.pattern(L.SYS_INTERIOR)
    va = ctx.values(L.SYS_INTERIOR)
    xo, io = O.pcg(rp, ci, va, b, tol=1e-10)
    assert io["status"] == 0
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    # same algorithm, same stopping test: iteration counts agree closely
    assert abs(info["iters"] - io["iters"]) <= 2


def test_pcg_reports_nonconvergence(ctx):
    p, b = _solve_case(ctx, 1, 10)
    x, info = ctx.pcg(b, mode=L.SOLVE_F64, tol=1e-10, max_iters=3, raise_on_fail=False)
    assert info["status"] == L.MG_ERR_NOCONV and info["relres"] > 1e-10
    with pytest.raises(L.MgfeError):
        ctx.pcg(b, mode=L.SOLVE_F64, tol=1e-10, max_iters=3)


# rules exact on the stiffness integrand take the closed form (P1: all; P2:
# 4, 11, 15 points), the P2 1-point rule the quadrature-table form
@pytest.mark.parametrize("order,rule", [(1, 4), (1, 1), (2, 4), (2, 11), (2, 1)])
@pytest.mark.parametrize("sysid", [L.SYS_FULL, L.SYS_INTERIOR])
def test_adjoint_kernel_matches_oracle(ctx, order, rule, sysid):
    m = W.kuhn_cube(4, jitter=0.2, seed=11)
    setup(ctx, m, order)
    fixed = ctx.dirichlet_mask()
    free = np.ones(fixed.size, bool) if sysid == L.SYS_FULL else fixed == 0
    nf = int(free.sum())
    u = W.stream_uniform(0, 7, nf) - 0.5
    lam = W.stream_uniform(1, 7, nf) - 0.5
    g = ctx.adjoint_grad_k(sysid, u, lam, rule=rule)
    uf = np.zeros(fixed.size)
    lf = np.zeros(fixed.size)
    uf[free] = u
    lf[free] = lam
    go = O.adjoint_grad_k(order, m, ctx.elem_dofs(), uf, lf, rule=rule)
    assert np.abs(g - go).max() <= 1e-12 * np.abs(go).max()


def test_adjoint_gradient_through_solve_fd(ctx):
    # J(k) = ||u(k) - u_obs||^2 with A(k) u = f; dJ/dk = -lam^T Khat u, A lam = 2(u - u_obs)
    p = W.config_problem(5, n=5)
    setup(ctx, p.mesh, 1, ALL_FACES)
    fixed = ctx.dirichlet_mask()
    b = p.f[fixed == 0]
    uobs = (2 * W.stream_uniform(0, W.STREAM_UOBS, fixed.size) - 1)[fixed == 0] * 1e-2

    def J(k):
        ctx.field_set(L.FIELD_STIFF, k)
        ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, stiff_coeff=L.COEFF_P0, stiff_rule=1)
        u, _ = ctx.pcg(b, tol=1e-13)
        return float(np.sum((u - uobs) ** 2)), u

    k = p.k.copy()
    J0, u = J(k)
    lam, _ = ctx.pcg(2 * (u - uobs), tol=1e-13)
    g = ctx.adjoint_grad_k(L.SYS_INTERIOR, u, lam, rule=1)
    rng = np.random.default_rng(0)
    for e in rng.choice(p.mesh.nT, 6, replace=False):
        h = 1e-6 * k[e]
        kp, km = k.copy(), k.copy()
        kp[e] += h
        km[e] -= h
        fd = (J(kp)[0] - J(km)[0]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-5 * np.abs(g).max(), (e, fd, g[e])


def test_dual_fv_matches_oracle(ctx):
    m = W.kuhn_cube(4, jitter=0.2, seed=5)
    setup(ctx, m, 1, None)
    ctx.dual_fv_build()
    d = ctx.dual_fv()
    ee, en = O.edges(m)
    o = O.dual_fv(m, ee, en)
    for key in ("edge_surf", "cell_volume", "sigma_el", "edge_n"):
        a, b = d[key], o[key]
        assert np.abs(a - b).max() <= 1e-14 * max(1.0, np.abs(b).max()), key


def _dd(n=4):
    m = W.kuhn_cube(n, jitter=0.2, seed=0)
    eps = 1 + 4 * W.stream_uniform(0, W.STREAM_K, m.nT)
    rho = 2 * W.stream_uniform(0, W.STREAM_RHO, m.nT) - 1
    mu = np.exp(0.5 * W.stream_normal(0, W.STREAM_MU, m.nT))
    dens = np.exp(W.stream_normal(0, W.STREAM_DENSITY, m.nV))
    phi = m.xyz[:, 0] + 0.1 * (2 * W.stream_uniform(0, W.STREAM_PHI, m.nV) - 1)
    fixed = O.dirichlet_mask(1, m, [1, 2])
    phi[fixed == 1] = m.xyz[fixed == 1, 0]
    return m, eps, rho, mu, phi, dens, fixed


def test_dd_residual_and_vjp_match_oracle(ctx):
    m, eps, rho, mu, phi, dens, fixed = _dd()
    setup(ctx, m, 1, None)
    ctx.dual_fv_build()
    loss, rp, rn = ctx.dd_residual(eps, rho, mu, phi, dens, fixed)
    ee, en = O.edges(m)
    d = O.dual_fv(m, ee, en)
    lo, rpo, rno = O.dd_loss(m, d, eps, rho, mu, phi, dens, fixed, ee, en)
    assert abs(loss - lo) <= 1e-12 * lo
    assert np.abs(rp - rpo).max() <= 1e-13 * np.abs(rpo).max()
    assert np.abs(rn - rno).max() <= 1e-13 * np.abs(rno).max()
    ge, gm = ctx.dd_vjp(eps, rho, mu, phi, dens, fixed)
    geo, gmo = O.dd_grad(m, d, eps, rho, mu, phi, dens, fixed, ee, en)
    assert np.abs(ge - geo).max() <= 1e-12 * np.abs(geo).max()
    assert np.abs(gm - gmo).max() <= 1e-12 * np.abs(gmo).max()


def test_dd_vjp_central_differences(ctx):
    m, eps, rho, mu, phi, dens, fixed = _dd()
    setup(ctx, m, 1, None)
    ctx.dual_fv_build()
    ge, gm = ctx.dd_vjp(eps, rho, mu, phi, dens, fixed)
    rng = np.random.default_rng(1)
    h = 1e-6
    for e in rng.choice(m.nT, 16, replace=False):
        for which, g in (("eps", ge), ("mu", gm)):
            a = {"eps": eps.copy(), "mu": mu.copy()}
            a[which][e] += h
            lp = ctx.dd_residual(a["eps"], rho, a["mu"], phi, dens, fixed)[0]
            a[which][e] -= 2 * h
            lm = ctx.dd_residual(a["eps"], rho, a["mu"], phi, dens, fixed)[0]
            assert abs((lp - lm) / (2 * h) - g[e]) <= 1e-6 * max(1.0, abs(g[e]))


def test_bernoulli_device(ctx):
    x = np.array([0.0, 1e-12, -1e-6, 5e-3, 0.0099999, 0.01, -0.01, 0.5, -3.0, 40.0, -40.0, 800.0])
    y = ctx.bernoulli(x)
    ref = np.array([O.bernoulli(v) for v in x])
    np.testing.assert_allclose(y, ref, rtol=2e-16, atol=0)
    assert y[0] == 1.0


def test_mesh_rejects_inverted_tet(ctx):
    m = W.kuhn_cube(2)
    t = m.tets.copy()
    t[3, [2, 3]] = t[3, [3, 2]]
    with pytest.raises(L.MgfeError) as ei:
        ctx.mesh_set(m.xyz, t, m.btri, m.bgeo)
    assert ei.value.code == L.MG_ERR_MESH
    bad = m.tets.copy()
    bad[0, 0] = m.nV + 5
    with pytest.raises(L.MgfeError):
        ctx.mesh_set(m.xyz, bad, m.btri, m.bgeo)


def test_state_errors(ctx):
    m = W.kuhn_cube(2)
    ctx.mesh_set(m.xyz, m.tets, m.btri, m.bgeo)
    with pytest.raises(L.MgfeError) as ei:
        ctx.pattern_build(L.SYS_FULL)
    assert ei.value.code == L.MG_ERR_STATE
    ctx.dofmap_build(2)
    with pytest.raises(L.MgfeError):
        ctx.pattern_build(L.SYS_INTERIOR)  # no Dirichlet set yet
    ctx.pattern_build(L.SYS_FULL)
    with pytest.raises(L.MgfeError):
        ctx.assemble(L.SYS_FULL, L.TERM_STIFF, stiff_coeff=L.COEFF_P0)  # no field
    with pytest.raises(L.MgfeError):
        ctx.assemble(L.SYS_FULL, L.TERM_STIFF, stiff_rule=7)


def test_copy_device(ctx):
    import torch
    src = torch.arange(1 << 20, dtype=torch.float64, device="cuda")
    dst = torch.zeros_like(src)
    ctx.copy_device(dst.data_ptr(), src.data_ptr(), src.numel() * 8)
    ctx.synchronize()
    assert torch.equal(src, dst)
    with pytest.raises(L.MgfeError):
        ctx.copy_device(dst.data_ptr(), src.data_ptr(), 12)


def test_dd_device_entry_points_match_host(ctx):
    import torch
    m, eps, rho, mu, phi, dens, fixed = _dd()
    setup(ctx, m, 1, None)
    ctx.dual_fv_build()
    loss, rp, rn = ctx.dd_residual(eps, rho, mu, phi, dens, fixed)
    ge, gm = ctx.dd_vjp(eps, rho, mu, phi, dens, fixed)
    d = {k: torch.tensor(v, device="cuda") for k, v in
         dict(eps=eps, rho=rho, mu=mu, phi=phi, dens=dens).items()}
    dfix = torch.tensor(fixed.astype(np.uint8), device="cuda")
    out = {k: torch.zeros(n, dtype=torch.float64, device="cuda") for k, n in
           dict(rp=m.nV, rn=m.nV, ge=m.nT, gm=m.nT, loss=1).items()}
    args = [d[k].data_ptr() for k in ("eps", "rho", "mu", "phi", "dens")] + [dfix.data_ptr()]
    ctx.dd_residual_device(*args, out["rp"].data_ptr(), out["rn"].data_ptr(), out["loss"].data_ptr())
    ctx.synchronize()
    assert out["loss"].item() == loss
    assert np.array_equal(out["rp"].cpu().numpy(), rp) and np.array_equal(out["rn"].cpu().numpy(), rn)
    ctx.dd_vjp_device(*args, out["rp"].data_ptr(), out["rn"].data_ptr(), out["ge"].data_ptr(),
                      out["gm"].data_ptr())
    ctx.synchronize()
    assert np.array_equal(out["ge"].cpu().numpy(), ge) and np.array_equal(out["gm"].cpu().numpy(), gm)


@pytest.mark.parametrize("mode", [L.SOLVE_F32_REFINE, L.SOLVE_F32_RELIABLE])
def test_mixed_solve_follows_fp64_reassembly(ctx, mode):
    # FP64 + FP32 assembled with k1, then FP64 alone with k2: the mixed solver
    # must use k2 (its FP32 copy is refreshed from the newer FP64 values)
    p, b = _solve_case(ctx, 5, 6)
    kw = dict(stiff_coeff=L.COEFF_P0, stiff_rule=1)
    ctx.field_set(L.FIELD_STIFF, p.k)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, **kw)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, prec=L.PREC_F32, **kw)
    ctx.pcg(b, mode=mode, tol=1e-10)
    ctx.field_set(L.FIELD_STIFF, 3.0 * p.k + 1.0)
    ctx.assemble(L.SYS_INTERIOR, L.TERM_STIFF, **kw)
    x64, _ = ctx.pcg(b, mode=L.SOLVE_F64, tol=1e-12)
    x, info = ctx.pcg(b, mode=mode, tol=1e-10)
    assert info["relres"] <= 1e-10
    assert np.linalg.norm(x - x64) <= 1e-8 * np.linalg.norm(x64)


@pytest.mark.parametrize("mode", [L.SOLVE_F64, L.SOLVE_F32_RELIABLE])
def test_pcg_device_unaligned_vectors(ctx, mode):
    # caller buffers 8 bytes off a 16-byte boundary take the scalar path of
    # the elementwise kernels; both paths reach the same solution
    torch = pytest.importorskip("torch")
    p, b = _solve_case(ctx, 2, 8)
    n = b.size
    out = []
    for off in (0, 1):
        db = torch.zeros(n + 2, dtype=torch.float64, device="cuda")
        dx = torch.zeros(n + 2, dtype=torch.float64, device="cuda")
        db[off:off + n] = torch.from_numpy(b).cuda()
        info = ctx.pcg_device(db.data_ptr() + 8 * off, dx.data_ptr() + 8 * off, mode=mode,
                              tol=1e-10)
        assert info["status"] == L.MG_OK and info["relres"] <= 1e-10, info
        out.append(dx[off:off + n].cpu().numpy())
    rp, ci = ctx.pattern(L.SYS_INTERIOR)
    A = sp.csr_matrix((ctx.values(L.SYS_INTERIOR), ci, rp), shape=(n, n))
    for x in out:
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1.5e-10
    assert np.linalg.norm(out[0] - out[1]) <= 1e-8 * np.linalg.norm(out[0])

