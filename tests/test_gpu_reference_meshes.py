"""Parity on the reference's own mesh fixtures (tests/morganaMeshes/*.msh of
the morganaPublic tree, copied under tests/golden/meshes), through the mirror
of the reference interfaces (fe.py) and through the raw C ABI."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from morganapublic_amd import _lib as L
from morganapublic_amd import fe
from morganapublic_amd import workloads as W

pytestmark = pytest.mark.gpu
MESHES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "meshes", "*.msh")))


@pytest.mark.parametrize("path", MESHES, ids=[os.path.basename(p) for p in MESHES])
@pytest.mark.parametrize("order,rule", [(1, 1), (2, 4)])
def test_reference_mesh_assembly(ctx, path, order, rule):
    m = W.read_gmsh22(path)
    ctx.mesh_set(m.xyz, m.tets, m.btri, m.bgeo)
    ctx.dofmap_build(order)
    ctx.pattern_build(L.SYS_FULL)
    k = 1.0 + np.arange(m.nT)  # distinct per-element coefficients
    ctx.field_set(L.FIELD_STIFF, k)
    ctx.assemble(L.SYS_FULL, L.TERM_STIFF | L.TERM_MASS, stiff_coeff=L.COEFF_P0, stiff_rule=rule,
                 mass_rule=15)
    rp, ci = ctx.pattern(L.SYS_FULL)
    va = ctx.values(L.SYS_FULL)
    orp, oci, ova = O.assemble(order, m, O.OP_STIFF | O.OP_MASS, c_stiff=k, rule_stiff=rule,
                               rule_mass=15)
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(ci, oci)
    assert np.abs(va - ova).max() <= 1e-12 * np.abs(ova).max()
    geos = sorted(set(m.bgeo.tolist()))[:1]
    ctx.dirichlet_set(geos)
    ee, en = O.edges(m)
    np.testing.assert_array_equal(ctx.dirichlet_mask(), O.dirichlet_mask(order, m, geos, en))


def test_traits_matrix_builder_mirror():
    # the steps of tests/morganaFiniteElements/traitsMatrixBuilderTest5.cpp:
    # mignon3dB.msh, P1 field/test, P0 coefficient "1", opEL3dA, INT_PRECISION 1
    path = os.path.join(os.path.dirname(__file__), "golden", "meshes", "mignon3dB.msh")
    with L.Context(0) as ctx:
        mesh = fe.Mesh3d.from_gmsh(path)
        space = fe.FeSpace(ctx, mesh, 1)
        op = fe.opEL3dA()
        op.setCoeff(np.ones(mesh.getNumElements()))
        A = fe.MatrixBuilder(space, op, stiff_points=fe.MatrixBuilder.PRECISION_TO_POINTS[1])
        M = A.buildEpetraCrs()
        orp, oci, ova = O.assemble(1, mesh.mesh, O.OP_STIFF, rule_stiff=1)
        np.testing.assert_array_equal(M.rowptr, orp)
        np.testing.assert_array_equal(M.colidx, oci)
        assert np.abs(M.values - ova).max() <= 1e-14
        assert M.NumGlobalRows() == mesh.getNumNodes()


def test_solver_factory_mirror():
    p = W.config_problem(1, n=10)
    with L.Context(0) as ctx:
        space = fe.FeSpace(ctx, fe.Mesh3d(p.mesh), 1)
        space.setDirichlet(p.dirichlet_geo)
        A = fe.MatrixBuilder(space, fe.opEL3dF(1.0) + fe.opEL3dG(), stiff_points=1,
                             mass_points=4, interior=True)
        A.assemble()
        b = p.f[space.fixedMask() == 0]
        x = np.zeros_like(b)
        prec = fe.factoryPreconditioner.create(
            {"precClass": "precIfpack", "precType": "point relaxation",
             "relaxation: type": "Jacobi", "overlap": 0})
        prec.initialize()
        prec.compute()
        solver = fe.factoryLinearSolver.create(
            {"Solver": "belosCG", "Block Size": 1, "Maximum Iterations": 1000,
             "Convergence Tolerance": 1e-10, "Verbosity": 0})
        solver.setProblem(A, x, b)
        solver.setLeftPrec(prec)
        assert solver.solve()
        assert solver.info["relres"] <= 1e-10
        short = fe.factoryLinearSolver.create({"Solver": "belosCG", "Maximum Iterations": 2,
                                               "Convergence Tolerance": 1e-10})
        short.setProblem(A, x, b)
        short.setLeftPrec(prec)
        assert not short.solve()
    with pytest.raises(ValueError):
        fe.factoryLinearSolver.create({"Solver": "belosGmres"})
    with pytest.raises(ValueError):
        fe.factoryPreconditioner.create({"precClass": "precIfpack", "precType": "ILU"})


def test_traits_vector_builder_mirror():
    # fnEL3dA over a P1 space of a reference mesh: constant, P0 and nodal f
    path = os.path.join(os.path.dirname(__file__), "golden", "meshes", "mignon3dB.msh")
    with L.Context(0) as ctx:
        mesh = fe.Mesh3d.from_gmsh(path)
        space = fe.FeSpace(ctx, mesh, 1)
        m = mesh.mesh
        for f, kind in ((2.5, O.LOAD_CONST), (1.0 + np.arange(m.nT), O.LOAD_P0),
                        (np.sin(m.xyz[:, 0]) + m.xyz[:, 2], O.LOAD_NODAL)):
            F = fe.VectorBuilder(space, fe.fnEL3dA(f), points=4).buildEpetraVector()
            ref = O.load_vector(1, m, kind, f, 4)
            assert np.abs(F - ref).max() <= 1e-13 * np.abs(ref).max()


def test_mixed_precision_solver_mirror():
    p = W.config_problem(1, n=10)
    with L.Context(0) as ctx:
        space = fe.FeSpace(ctx, fe.Mesh3d(p.mesh), 1)
        space.setDirichlet(p.dirichlet_geo)
        A = fe.MatrixBuilder(space, fe.opEL3dF(1.0) + fe.opEL3dG(), stiff_points=1,
                             mass_points=4, interior=True)
        A.assemble()
        b = p.f[space.fixedMask() == 0]
        prec = fe.factoryPreconditioner.create({"precClass": "precIfpack"})
        prec.initialize()
        prec.compute()
        for precision in ("mixed", "float+refine"):
            x = np.zeros_like(b)
            solver = fe.factoryLinearSolver.create(
                {"Solver": "belosCG", "Maximum Iterations": 1000, "Convergence Tolerance": 1e-10,
                 "Precision": precision})
            solver.setProblem(A, x, b)
            solver.setLeftPrec(prec)
            assert solver.solve()
            assert solver.info["relres"] <= 1e-10 and solver.info["outer"] >= 1
