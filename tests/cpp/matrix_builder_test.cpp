// The steps of the reference's traitsMatrixBuilderTest5.cpp (mesh loaded by
// meshInit3d::gmMesh_to_stdB, P0 coefficient, opEL3dA, matrixBuilder) and of
// a fnEL3dA load vector, through the C++ host interface mgfe.hpp.  Writes the
// CSR matrix and the load vector to a binary file for the parity test.
//   matrix_builder_test <mesh.msh> <order> <intPoints> <out.bin>
#include <cstdio>
#include <cstdlib>

#include "mgfe.hpp"

int main(int argc, char **argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s mesh.msh order intPoints out.bin\n", argv[0]);
    return 2;
  }
  try {
    mgfe::Context world;
    mgfe::meshInit3d init;
    init.gmMesh_to_stdB(argv[1]);
    const mgfe::mesh3d &grid3d = init.getGrid3d();
    const int order = std::atoi(argv[2]), points = std::atoi(argv[3]);

    mgfe::feSpace3d space(world, grid3d, order);
    mgfe::P0Field coeff;  // distinct per-element values: 1, 2, 3, ...
    for (int64_t e = 0; e < grid3d.getNumElements(); ++e) coeff.values.push_back(1.0 + e);
    mgfe::opEL3dA op;
    op.setCoeff(coeff);
    mgfe::opEL3dG mass;
    mgfe::matrixBuilder assembler(space, op + mass, points, 15);
    mgfe::CrsMatrix matrix;
    assembler.buildEpetraCrs(matrix);

    mgfe::fnEL3dA fn;
    fn.setCoeff(coeff);
    mgfe::vectorBuilder vb(space, fn, points);
    std::vector<double> F;
    vb.buildEpetraVector(F);

    std::FILE *out = std::fopen(argv[4], "wb");
    if (!out) return 1;
    const int64_t head[3] = {matrix.NumGlobalRows(), matrix.NumGlobalNonzeros(),
                             (int64_t)F.size()};
    std::fwrite(head, sizeof(int64_t), 3, out);
    std::fwrite(matrix.rowptr.data(), sizeof(int32_t), matrix.rowptr.size(), out);
    std::fwrite(matrix.colidx.data(), sizeof(int32_t), matrix.colidx.size(), out);
    std::fwrite(matrix.values.data(), sizeof(double), matrix.values.size(), out);
    std::fwrite(F.data(), sizeof(double), F.size(), out);
    std::fclose(out);
    matrix.printToFile(std::string(argv[4]) + ".mtx");
    matrix.printToMatlabFile(std::string(argv[4]) + ".m");
    std::printf("rows %lld nnz %lld\n", (long long)head[0], (long long)head[1]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
