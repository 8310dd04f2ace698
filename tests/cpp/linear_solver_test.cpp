// The solve steps of the reference's factoryLinearSolver / belosBlockCG with
// an Ifpack point-Jacobi preconditioner, through mgfe.hpp: K + M on a P1
// space with Dirichlet conditions on every boundary geoId, load f = 1, CG to
// 1e-10.  Writes the interior solution for the parity test.
//   linear_solver_test <mesh.msh> <precision: double|mixed|float+refine> <out.bin>
#include <cstdio>
#include <set>

#include "mgfe.hpp"

int main(int argc, char **argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s mesh.msh precision out.bin\n", argv[0]);
    return 2;
  }
  try {
    mgfe::Context world;
    mgfe::meshInit3d init;
    init.gmMesh_to_stdB(argv[1]);
    const mgfe::mesh3d &grid3d = init.getGrid3d();
    mgfe::feSpace3d space(world, grid3d, 1);
    std::set<int32_t> geo(grid3d.geoIds.begin(), grid3d.geoIds.end());
    space.setDirichlet(std::vector<int32_t>(geo.begin(), geo.end()));

    mgfe::opEL3dF stiff(1.0);
    mgfe::opEL3dG mass;
    mgfe::matrixBuilder A(space, stiff + mass, 1, 4, /*interior=*/true);
    A.assemble();
    mgfe::fnEL3dA fn;
    fn.setCoeff(1.0);
    mgfe::vectorBuilder vb(space, fn, 4, /*interior=*/true);
    std::vector<double> b, x;
    vb.buildEpetraVector(b);

    mgfe::ParameterList precList;
    precList.set("precClass", "precIfpack").set("precType", "point relaxation");
    precList.set("relaxation: type", "Jacobi").set("overlap", 0);
    auto prec = mgfe::factoryPreconditioner::create(precList);
    prec->initialize();
    prec->compute();

    mgfe::ParameterList solverList;
    solverList.set("Solver", "belosCG").set("Block Size", 1).set("Maximum Iterations", 2000);
    solverList.set("Convergence Tolerance", 1e-10).set("Precision", argv[2]);
    auto solver = mgfe::factoryLinearSolver::create(solverList);
    solver->setProblem(A, x, b);
    solver->setLeftPrec(prec);
    const bool ok = solver->solve();

    std::FILE *out = std::fopen(argv[3], "wb");
    if (!out) return 1;
    const int64_t n = (int64_t)x.size();
    const int64_t iters = solver->getNumIters();
    const double relres = solver->info().relres;
    std::fwrite(&n, sizeof(int64_t), 1, out);
    std::fwrite(&iters, sizeof(int64_t), 1, out);
    std::fwrite(&relres, sizeof(double), 1, out);
    std::fwrite(b.data(), sizeof(double), b.size(), out);
    std::fwrite(x.data(), sizeof(double), x.size(), out);
    std::fclose(out);
    std::printf("converged %d iterations %lld relres %.3e\n", ok ? 1 : 0, (long long)iters, relres);
    return ok ? 0 : 3;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
