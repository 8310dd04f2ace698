"""Host-side mirror of the reference's assembly and solve interfaces.

Names, argument meaning and error behaviour follow Morgana's classes so that a
caller of the reference reads the same steps here; the work is done by
libmgfe.so through `Context`.

  reference                                         here
  meshInit3d::gmMesh_to_stdB (meshInit3d.hpp:487)   Mesh3d.from_gmsh / from_arrays
  dofMapStatic3d (dofMapStatic3d.hpp:423-712)       FeSpace (order 1 or 2)
  opEL3dA / opEL3dF / opEL3dG / opEL3dL             opEL3dA / opEL3dF / opEL3dG / opEL3dL
  epetraMatrixManip::sum (.cpp:510-527)             operator + operator
  matrixBuilder<localMatrixEL3d<...>>::buildEpetraCrs
        (matrixBuilder.hpp:22-60,
         traitsMatrixBuilder.hpp:366-441)           MatrixBuilder(...).buildEpetraCrs()
  fnEL3dA + traitsVectorBuilder::buildEpetraVector
        (traitsVectorBuilder.hpp:217-274)           VectorBuilder(space, fnEL3dA(f)).buildEpetraVector()
  factoryPreconditioner + preconditionerIfpack
        (factoryPreconditioner.hpp:53-87,
         preconditionerIfpack.cpp:26-63)            factoryPreconditioner.create(params)
  factoryLinearSolver + belosBlockCG
        (factoryLinearSolver.hpp:51-91,
         belosBlockCG.hpp:46-83)                    factoryLinearSolver.create(params)

Differences kept on purpose: Dirichlet data is imposed by elimination (the
interior system) rather than by Lagrange multipliers, and solve() returns
whether the explicit FP64 residual met the tolerance (belosBlockCG::solve
returns true even without convergence).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import _lib as L
from . import workloads as W


class Mesh3d:
    """Tetrahedral mesh with gmsh loader semantics (node gid = file order)."""

    def __init__(self, mesh: W.TetMesh):
        self.mesh = mesh

    @classmethod
    def from_gmsh(cls, path: str) -> "Mesh3d":
        return cls(W.read_gmsh22(path))

    @classmethod
    def from_arrays(cls, xyz, tets, btri=None, bgeo=None) -> "Mesh3d":
        btri = np.zeros((0, 3), np.int32) if btri is None else np.asarray(btri, np.int32)
        bgeo = np.zeros(0, np.int32) if bgeo is None else np.asarray(bgeo, np.int32)
        return cls(W.TetMesh(n=0, xyz=np.asarray(xyz, np.float64),
                             tets=np.asarray(tets, np.int32), btri=btri, bgeo=bgeo))

    def getNumElements(self) -> int:
        return self.mesh.nT

    def getNumNodes(self) -> int:
        return self.mesh.nV


class FeSpace:
    """fePr3d<order> field on a mesh with its dof map (dofMapStatic3d)."""

    def __init__(self, ctx: L.Context, mesh: Mesh3d, order: int):
        if order not in (1, 2):
            raise ValueError("fePr3d order must be 1 or 2")
        self.ctx, self.mesh, self.order = ctx, mesh, order
        m = mesh.mesh
        ctx.mesh_set(m.xyz, m.tets, m.btri, m.bgeo)
        ctx.dofmap_build(order)
        self._dirichlet = None

    def getNumDofsG(self) -> int:
        return self.ctx.ndofs

    def mapListG(self):
        """0-based global list ids of every element's local basis (feIndicesLG)."""
        return self.ctx.elem_dofs()

    def setDirichlet(self, geo_ids) -> int:
        """Fixes every dof on boundary triangles with these geoIds (bcCardFE)."""
        self._dirichlet = tuple(int(g) for g in geo_ids)
        n = self.ctx.dirichlet_set(list(self._dirichlet))
        self.ctx.pattern_build(L.SYS_INTERIOR)
        return n

    def fixedMask(self):
        return self.ctx.dirichlet_mask()


@dataclasses.dataclass
class _Term:
    kind: int  # L.TERM_STIFF or L.TERM_MASS
    coeff: object  # float or per-element array


class _Operator:
    def __init__(self, terms):
        self.terms = list(terms)

    def __add__(self, other: "_Operator") -> "_Operator":
        kinds = [t.kind for t in self.terms + other.terms]
        if len(set(kinds)) != len(kinds):
            raise ValueError("sum of two operators of the same kind is not supported")
        return _Operator(self.terms + other.terms)


class opEL3dA(_Operator):
    """c grad(u).grad(v), c a P0 field (opEL3dA.hpp:129-165)."""

    def __init__(self, coeff=None):
        super().__init__([_Term(L.TERM_STIFF, coeff)])

    def setCoeff(self, coeff):
        self.terms[0].coeff = np.asarray(coeff, np.float64)


class opEL3dF(_Operator):
    """c grad(u).grad(v), c constant (opEL3dF.hpp:102)."""

    def __init__(self, c: float = 1.0):
        super().__init__([_Term(L.TERM_STIFF, float(c))])


class opEL3dG(_Operator):
    """u v (opEL3dG.hpp:83-108)."""

    def __init__(self):
        super().__init__([_Term(L.TERM_MASS, 1.0)])


class opEL3dL(_Operator):
    """c u v, c a P0 field (opEL3dL.hpp:129)."""

    def __init__(self, coeff=None):
        super().__init__([_Term(L.TERM_MASS, coeff)])

    def setCoeff(self, coeff):
        self.terms[0].coeff = np.asarray(coeff, np.float64)


@dataclasses.dataclass
class CrsMatrix:
    """The host view of an assembled matrix (Epetra row map, 0-based gids)."""
    rowptr: np.ndarray
    colidx: np.ndarray
    values: np.ndarray

    def NumGlobalRows(self) -> int:
        return self.rowptr.size - 1

    def NumGlobalNonzeros(self) -> int:
        return self.colidx.size

    def to_scipy(self):
        import scipy.sparse as sp
        n = self.NumGlobalRows()
        return sp.csr_matrix((self.values, self.colidx, self.rowptr), shape=(n, n))

    def _write(self, name: str, market: bool):
        n = self.NumGlobalRows()
        rows = np.repeat(np.arange(1, n + 1), np.diff(self.rowptr))
        with open(name, "w") as f:
            if market:
                f.write("%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n"
                        % (n, n, self.NumGlobalNonzeros()))
            for r, c, v in zip(rows, self.colidx + 1, self.values):
                f.write("%d %d %s\n" % (r, c, repr(float(v))))

    def printToFile(self, name: str):
        """MatrixMarket coordinate file (real general, 1-based, structural zeros
        kept), as tpetraBlockMatrixAssemble::printToFile
        (tpetraBlockMatrixAssemble.cpp:470-475); values read back exactly."""
        self._write(name, True)

    def printToMatlabFile(self, name: str):
        """'row col value' lines, 1-based, as epetraBlockMatrixAssemble::printToFile
        (epetraBlockMatrixAssemble.cpp:412-416, EpetraExt::RowMatrixToMatlabFile)."""
        self._write(name, False)


class MatrixBuilder:
    """matrixBuilder over localMatrixEL3d<FIELD,TEST,...,INT_PRECISION>.

    int_points: quadrature points of the stiffness / mass terms (1, 4, 11, 15);
    the reference's INT_PRECISION 1, 2, 3 are 1, 4, 15 points.
    """

    PRECISION_TO_POINTS = {1: 1, 2: 4, 3: 15}

    def __init__(self, space: FeSpace, op: _Operator, stiff_points: int = 4,
                 mass_points: int = 4, interior: bool = False, fp32: bool = False):
        self.space, self.op = space, op
        self.stiff_points, self.mass_points = stiff_points, mass_points
        self.sys = L.SYS_INTERIOR if interior else L.SYS_FULL
        self.prec = L.PREC_F32 if fp32 else L.PREC_F64
        if self.sys == L.SYS_FULL and L.SYS_FULL not in space.ctx._nnz:
            space.ctx.pattern_build(L.SYS_FULL)

    def assemble(self):
        """Fills the device-resident matrix (no host copy)."""
        ctx = self.space.ctx
        kw = dict(stiff_rule=self.stiff_points, mass_rule=self.mass_points, prec=self.prec)
        terms = 0
        for t in self.op.terms:
            terms |= t.kind
            if t.kind == L.TERM_STIFF:
                key, field = ("stiff", L.FIELD_STIFF)
            else:
                key, field = ("mass", L.FIELD_MASS)
            if np.ndim(t.coeff) == 0:
                kw[key + "_coeff"] = L.COEFF_CONST
                kw[key + "_const"] = 1.0 if t.coeff is None else float(t.coeff)
            else:
                ctx.field_set(field, t.coeff)
                kw[key + "_coeff"] = L.COEFF_P0
        ctx.assemble(self.sys, terms, **kw)

    def buildEpetraCrs(self) -> CrsMatrix:
        self.assemble()
        rp, ci = self.space.ctx.pattern(self.sys)
        return CrsMatrix(rp, ci, self.space.ctx.values(self.sys, self.prec))


class fnEL3dA:
    """Functional f v (fnEL3dA.hpp): f a constant, a P0 field (one value per
    element) or a field of the space's element type (one value per dof)."""

    def __init__(self, coeff=1.0):
        self.coeff = coeff

    def setCoeff(self, coeff):
        self.coeff = coeff


class VectorBuilder:
    """vectorBuilder over localVectorEL3d: buildEpetraVector
    (traitsVectorBuilder.hpp:217-274) with `points` quadrature points."""

    def __init__(self, space: FeSpace, fn: fnEL3dA, points: int = 4, interior: bool = False):
        self.space, self.fn, self.points = space, fn, points
        self.sys = L.SYS_INTERIOR if interior else L.SYS_FULL
        if self.sys not in space.ctx._nnz:
            space.ctx.pattern_build(self.sys)

    def buildEpetraVector(self) -> np.ndarray:
        f = np.asarray(self.fn.coeff, np.float64)
        ctx = self.space.ctx
        if f.ndim == 0 or f.size == 1:
            kind = L.LOAD_CONST
        elif f.size == ctx.nT and f.size != ctx.ndofs:
            kind = L.LOAD_P0
        elif f.size == ctx.ndofs:
            kind = L.LOAD_NODAL
        else:
            raise ValueError("coefficient must be a scalar, one value per element or per dof")
        return ctx.assemble_load(self.sys, kind, f, self.points)


class JacobiPreconditioner:
    """Ifpack point Jacobi (the only preconditioner this path provides)."""

    def __init__(self, params: dict):
        self.params = dict(params)

    def initialize(self):
        return True

    def compute(self):
        return True


class factoryPreconditioner:
    @staticmethod
    def create(params: dict) -> JacobiPreconditioner:
        cls = str(params.get("precClass", "precIfpack"))
        ptype = str(params.get("precType", "point relaxation"))
        relax = str(params.get("relaxation: type", "Jacobi"))
        if cls not in ("precIfpack", "1") or ptype != "point relaxation" or relax != "Jacobi":
            raise ValueError("only Ifpack point-relaxation Jacobi is provided "
                             "(precClass=%s precType=%s relaxation: type=%s)" % (cls, ptype, relax))
        return JacobiPreconditioner(params)


class BelosCG:
    """Block CG with block size 1, preconditioned by point Jacobi, on the
    interior (Dirichlet-eliminated) system assembled on the device."""

    def __init__(self, params: dict):
        if int(params.get("Block Size", 1)) != 1:
            raise ValueError("Block Size must be 1")
        self.maxit = int(params.get("Maximum Iterations", 100000))
        self.tol = float(params.get("Convergence Tolerance", 1e-10))
        prec = str(params.get("Precision", "double")).lower()
        # "float+refine": FP32 inner solves restarted by FP64 refinement;
        # "mixed" / "float+reliable": one FP32 recursion with FP64 residual
        # replacement (keeps the search direction, fewer iterations)
        self.mode = {"float+refine": L.SOLVE_F32_REFINE, "mixed": L.SOLVE_F32_RELIABLE,
                     "float+reliable": L.SOLVE_F32_RELIABLE}.get(prec, L.SOLVE_F64)
        it = params.get("Inner Tolerance")
        self.inner_tol = None if it is None else float(it)
        self.verbosity = int(params.get("Verbosity", 0))
        self.A = self.x = self.b = None
        self.prec = None
        self.info = None

    def setProblem(self, A: MatrixBuilder, X: np.ndarray, B: np.ndarray):
        if A.sys != L.SYS_INTERIOR:
            raise ValueError("CG needs the interior (SPD) system: MatrixBuilder(interior=True)")
        self.A, self.x, self.b = A, X, np.asarray(B, np.float64)

    def setLeftPrec(self, prec: JacobiPreconditioner):
        self.prec = prec

    setRightPrec = setLeftPrec

    def solve(self) -> bool:
        if self.A is None:
            raise RuntimeError("setProblem first")
        if self.prec is None:
            raise RuntimeError("this path always preconditions with point Jacobi: set it")
        ctx = self.A.space.ctx
        x, info = ctx.pcg(self.b, mode=self.mode, tol=self.tol, inner_tol=self.inner_tol,
                          max_iters=self.maxit, raise_on_fail=False)
        self.x[:] = x
        self.info = info
        if self.verbosity:
            print("CG: %d iterations, %d refinements, ||b-Ax||/||b|| = %.3e"
                  % (info["iters"], info["outer"], info["relres"]))
        return info["status"] == L.MG_OK

    def getNumIters(self) -> int:
        return 0 if self.info is None else self.info["iters"]


class factoryLinearSolver:
    @staticmethod
    def create(params: dict) -> BelosCG:
        solver = str(params.get("Solver", "belosCG"))
        if solver not in ("belosCG", "2"):
            raise ValueError("only belosCG is provided on this path (got %s)" % solver)
        return BelosCG(params)
