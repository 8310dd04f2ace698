/*
 * mgfe.hpp — C++ host interface over the C ABI of libmgfe.so (mgfe.h),
 * mirroring the names and call sequence of Morgana's assembly and solve
 * classes so that C++ host code reads the same steps:
 *
 *   reference (morganaPublic tree)                        here (namespace mgfe)
 *   meshInit3d::gmMesh_to_stdB (meshInit3d.hpp:487-690)   meshInit3d::gmMesh_to_stdB
 *   feStaticField3d<fePr3d<0>> coefficient                P0Field
 *   dofMapStatic3d (dofMapStatic3d.hpp:423-712)           feSpace3d (order 1 or 2)
 *   opEL3dA / opEL3dF / opEL3dG / opEL3dL                 opEL3dA / opEL3dF / opEL3dG / opEL3dL
 *   epetraMatrixManip::sum (.cpp:510-527)                 operator+ of two operators
 *   traitsMatrixBuilder::buildEpetraCrs (:366-441)        matrixBuilder::buildEpetraCrs
 *   fnEL3dA + traitsVectorBuilder::buildEpetraVector
 *       (traitsVectorBuilder.hpp:217-274)                 vectorBuilder::buildEpetraVector
 *   Teuchos::ParameterList                                ParameterList
 *   factoryPreconditioner::create (.hpp:53-87)            factoryPreconditioner::create
 *   factoryLinearSolver::create (.hpp:51-91)              factoryLinearSolver::create
 *   virtualLinearSolver::setProblem/setLeftPrec/solve     virtualLinearSolver (same methods)
 *   tpetraBlockMatrixAssemble::printToFile (.cpp:470-475) CrsMatrix::printToFile (MatrixMarket)
 *   epetraBlockMatrixAssemble::printToFile (.cpp:412-416) CrsMatrix::printToMatlabFile
 *
 * Plain std::vector containers stand in for Epetra's, and failures throw
 * std::runtime_error with mg_last_error's text instead of asserting.  One
 * Context owns one GPU stream; everything built from it shares that stream.
 */
#ifndef MGFE_HPP
#define MGFE_HPP

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mgfe.h"

namespace mgfe {

// ---- context --------------------------------------------------------------
class Context {
 public:
  explicit Context(int device = 0) {
    if (mg_ctx_create(device, &ctx_) != MG_OK) throw std::runtime_error("mgfe: no usable GPU");
  }
  ~Context() { mg_ctx_destroy(ctx_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  mg_ctx *get() const { return ctx_; }
  void check(int rc) const {
    if (rc != MG_OK) throw std::runtime_error(std::string("mgfe: ") + mg_last_error(ctx_));
  }

 private:
  mg_ctx *ctx_ = nullptr;
};

// ---- mesh -----------------------------------------------------------------
struct mesh3d {
  std::vector<double> xyz;      // 3 per node
  std::vector<int32_t> tets;    // 4 per element, 1-based node gids
  std::vector<int32_t> btri;    // 3 per boundary triangle, 1-based
  std::vector<int32_t> geoIds;  // one per boundary triangle
  int64_t getNumNodes() const { return (int64_t)xyz.size() / 3; }
  int64_t getNumElements() const { return (int64_t)tets.size() / 4; }
};

// gmsh 2.x ASCII reader with the loader's semantics (loadMesh.hpp:2065-2350):
// node gid = file order, tetrahedra (type 4) in file order, boundary triangles
// (type 2) with their geometric tag (second tag) as geoId.
class meshInit3d {
 public:
  void gmMesh_to_stdB(const std::string &file) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("mgfe: cannot open " + file);
    std::string tok;
    while (in >> tok && tok != "$Nodes") {
    }
    int64_t nv = 0;
    in >> nv;
    grid.xyz.resize(3 * nv);
    for (int64_t i = 0; i < nv; ++i) {
      int64_t id;
      in >> id >> grid.xyz[3 * i] >> grid.xyz[3 * i + 1] >> grid.xyz[3 * i + 2];
    }
    while (in >> tok && tok != "$Elements") {
    }
    int64_t ne = 0;
    in >> ne;
    std::string line;
    std::getline(in, line);
    for (int64_t i = 0; i < ne; ++i) {
      std::getline(in, line);
      std::istringstream ls(line);
      int64_t id, type, ntag;
      ls >> id >> type >> ntag;
      std::vector<int64_t> tags(ntag);
      for (auto &t : tags) ls >> t;
      if (type == 4) {
        for (int k = 0; k < 4; ++k) {
          int64_t v;
          ls >> v;
          grid.tets.push_back((int32_t)v);
        }
      } else if (type == 2) {
        for (int k = 0; k < 3; ++k) {
          int64_t v;
          ls >> v;
          grid.btri.push_back((int32_t)v);
        }
        grid.geoIds.push_back((int32_t)(ntag > 1 ? tags[1] : tags[0]));
      }
    }
  }
  const mesh3d &getGrid3d() const { return grid; }
  mesh3d grid;
};

// ---- fields and spaces ----------------------------------------------------
// P0 coefficient: one value per element (feStaticField3d<fePr3d<0>>)
struct P0Field {
  std::vector<double> values;
};

// fePr3d<order> on a mesh with its dof map: vertex dofs (gid = node gid), then
// edge dofs in lexicographic order of their sorted node pairs
class feSpace3d {
 public:
  feSpace3d(Context &ctx, const mesh3d &grid, int order) : ctx_(ctx), order_(order) {
    if (order != 1 && order != 2) throw std::runtime_error("mgfe: fePr3d order must be 1 or 2");
    ctx.check(mg_mesh_set(ctx.get(), grid.getNumNodes(), grid.xyz.data(), grid.getNumElements(),
                          grid.tets.data(), (int64_t)grid.geoIds.size(),
                          grid.btri.empty() ? nullptr : grid.btri.data(),
                          grid.geoIds.empty() ? nullptr : grid.geoIds.data()));
    int64_t ne = 0;
    ctx.check(mg_dofmap_build(ctx.get(), order, &nDofs_, &ne));
    nElements_ = grid.getNumElements();
  }
  int getOrder() const { return order_; }
  int64_t getNumDofsG() const { return nDofs_; }
  int64_t getNumElements() const { return nElements_; }
  // 0-based global ids of every element's local basis (feIndicesLG)
  std::vector<int32_t> mapListG() const {
    std::vector<int32_t> d(nElements_ * (order_ == 1 ? 4 : 10));
    ctx_.check(mg_elem_dofs_get(ctx_.get(), d.data()));
    return d;
  }
  // Dirichlet dofs on the boundary triangles of these geoIds; the interior
  // (free) dofs are renumbered in increasing order
  int64_t setDirichlet(const std::vector<int32_t> &geoIds) {
    ctx_.check(mg_dirichlet_set(ctx_.get(), (int)geoIds.size(), geoIds.data(), &nFree_));
    return nFree_;
  }
  std::vector<uint8_t> fixedMask() const {
    std::vector<uint8_t> m(nDofs_);
    ctx_.check(mg_dirichlet_mask_get(ctx_.get(), m.data()));
    return m;
  }
  int64_t getNumFreeDofs() const { return nFree_; }
  Context &context() const { return ctx_; }

 private:
  Context &ctx_;
  int order_;
  int64_t nDofs_ = 0, nElements_ = 0, nFree_ = 0;
};

// ---- operators --------------------------------------------------------------
struct opTerm {
  int kind;                   // MG_TERM_STIFF or MG_TERM_MASS
  double constant = 1.0;      // used without a P0 coefficient
  const P0Field *coeff = nullptr;
};

class opEL3d {
 public:
  std::vector<opTerm> terms;
};

// c grad(u).grad(v), c a P0 field (opEL3dA.hpp:129-165)
class opEL3dA : public opEL3d {
 public:
  opEL3dA() { terms.push_back({MG_TERM_STIFF}); }
  void setCoeff(const P0Field &c) { terms[0].coeff = &c; }
};
// c grad(u).grad(v), c constant (opEL3dF.hpp:102)
class opEL3dF : public opEL3d {
 public:
  explicit opEL3dF(double c = 1.0) { terms.push_back({MG_TERM_STIFF, c}); }
};
// u v (opEL3dG.hpp:83-108)
class opEL3dG : public opEL3d {
 public:
  opEL3dG() { terms.push_back({MG_TERM_MASS}); }
};
// c u v, c a P0 field (opEL3dL.hpp:129)
class opEL3dL : public opEL3d {
 public:
  opEL3dL() { terms.push_back({MG_TERM_MASS}); }
  void setCoeff(const P0Field &c) { terms[0].coeff = &c; }
};

// sum of a stiffness and a mass operator, assembled in one pass
inline opEL3d operator+(const opEL3d &a, const opEL3d &b) {
  opEL3d s;
  s.terms = a.terms;
  for (const opTerm &t : b.terms) {
    for (const opTerm &u : s.terms)
      if (u.kind == t.kind) throw std::runtime_error("mgfe: two operators of the same kind");
    s.terms.push_back(t);
  }
  return s;
}

// ---- matrices ---------------------------------------------------------------
struct CrsMatrix {
  std::vector<int32_t> rowptr, colidx;  // 0-based gids (Epetra row map)
  std::vector<double> values;
  int64_t NumGlobalRows() const { return (int64_t)rowptr.size() - 1; }
  int64_t NumGlobalNonzeros() const { return (int64_t)colidx.size(); }

  // MatrixMarket coordinate file (real general, 1-based, every stored entry
  // including structural zeros), as Tpetra's MatrixMarket writer that
  // tpetraBlockMatrixAssemble::printToFile calls; values with 17 significant
  // digits, so they read back exactly
  void printToFile(const std::string &name) const {
    write(name, true);
  }
  // "row col value" lines, 1-based (Matlab's spconvert form), as
  // EpetraExt::RowMatrixToMatlabFile that epetraBlockMatrixAssemble::printToFile
  // calls
  void printToMatlabFile(const std::string &name) const { write(name, false); }

 private:
  void write(const std::string &name, bool market) const {
    std::FILE *f = std::fopen(name.c_str(), "w");
    if (!f) throw std::runtime_error("CrsMatrix: cannot open " + name);
    const int64_t n = NumGlobalRows();
    if (market)
      std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%lld %lld %lld\n",
                   (long long)n, (long long)n, (long long)NumGlobalNonzeros());
    for (int64_t r = 0; r < n; ++r)
      for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k)
        std::fprintf(f, "%lld %d %.17g\n", (long long)(r + 1), colidx[k] + 1, values[k]);
    if (std::fclose(f) != 0) throw std::runtime_error("CrsMatrix: cannot write " + name);
  }
};

// matrixBuilder over localMatrixEL3d: intPoints = quadrature points of the
// stiffness and mass terms (1, 4, 11, 15; the reference's INT_PRECISION 1, 2,
// 3 are 1, 4, 15 points).  interior: the Dirichlet-eliminated system.
class matrixBuilder {
 public:
  matrixBuilder(feSpace3d &space, const opEL3d &op, int stiffPoints = 4, int massPoints = 4,
                bool interior = false)
      : space_(space), op_(op), stiff_(stiffPoints), mass_(massPoints),
        sys_(interior ? MG_SYS_INTERIOR : MG_SYS_FULL) {
    int64_t nr = 0, nz = 0;
    space.context().check(mg_pattern_build(space.context().get(), sys_, &nr, &nz));
    nRows_ = nr;
    nnz_ = nz;
  }
  // assembles the device-resident matrix (no host copy)
  void assemble(int prec = MG_PREC_F64) {
    Context &ctx = space_.context();
    mg_operator o{0, MG_COEFF_CONST, 1.0, stiff_, MG_COEFF_CONST, 1.0, mass_};
    for (const opTerm &t : op_.terms) {
      o.terms |= t.kind;
      const int field = t.kind == MG_TERM_STIFF ? MG_FIELD_STIFF : MG_FIELD_MASS;
      int coeffKind = MG_COEFF_CONST;
      if (t.coeff) {
        ctx.check(mg_field_set(ctx.get(), field, t.coeff->values.data()));
        coeffKind = MG_COEFF_P0;
      }
      if (t.kind == MG_TERM_STIFF) {
        o.stiff_coeff = coeffKind;
        o.stiff_const = t.constant;
      } else {
        o.mass_coeff = coeffKind;
        o.mass_const = t.constant;
      }
    }
    ctx.check(mg_assemble(ctx.get(), sys_, &o, prec));
  }
  void buildEpetraCrs(CrsMatrix &A) {
    assemble();
    Context &ctx = space_.context();
    A.rowptr.resize(nRows_ + 1);
    A.colidx.resize(nnz_);
    A.values.resize(nnz_);
    ctx.check(mg_pattern_get(ctx.get(), sys_, A.rowptr.data(), A.colidx.data()));
    ctx.check(mg_values_get(ctx.get(), sys_, MG_PREC_F64, A.values.data()));
  }
  int system() const { return sys_; }
  feSpace3d &space() const { return space_; }

 private:
  feSpace3d &space_;
  opEL3d op_;
  int stiff_, mass_, sys_;
  int64_t nRows_ = 0, nnz_ = 0;
};

// ---- load vectors -------------------------------------------------------------
// functional f v (fnEL3dA.hpp): f constant, per element or per dof
class fnEL3dA {
 public:
  void setCoeff(double c) {
    kind = MG_LOAD_CONST;
    values.assign(1, c);
  }
  void setCoeff(const P0Field &c) {
    kind = MG_LOAD_P0;
    values = c.values;
  }
  void setCoeffNodal(const std::vector<double> &f) {
    kind = MG_LOAD_NODAL;
    values = f;
  }
  int kind = MG_LOAD_CONST;
  std::vector<double> values{1.0};
};

class vectorBuilder {
 public:
  vectorBuilder(feSpace3d &space, const fnEL3dA &fn, int points = 4, bool interior = false)
      : space_(space), fn_(fn), points_(points),
        sys_(interior ? MG_SYS_INTERIOR : MG_SYS_FULL) {
    int64_t nr = 0, nz = 0;
    space.context().check(mg_pattern_build(space.context().get(), sys_, &nr, &nz));
    nRows_ = nr;
  }
  void buildEpetraVector(std::vector<double> &F) {
    F.assign(nRows_, 0.0);
    Context &ctx = space_.context();
    ctx.check(mg_assemble_load(ctx.get(), sys_, fn_.kind, fn_.values.data(),
                               (int64_t)fn_.values.size(), points_, F.data()));
  }

 private:
  feSpace3d &space_;
  fnEL3dA fn_;
  int points_, sys_;
  int64_t nRows_ = 0;
};

// ---- parameters, preconditioner, solver -------------------------------------
// Teuchos::ParameterList stand-in: string-keyed values read back typed
class ParameterList {
 public:
  template <typename V>
  ParameterList &set(const std::string &k, const V &v) {
    std::ostringstream os;
    os.precision(17);
    os << v;
    values_[k] = os.str();
    return *this;
  }
  ParameterList &set(const std::string &k, const char *v) { return set(k, std::string(v)); }
  template <typename V>
  V get(const std::string &k, const V &dflt) const {
    auto it = values_.find(k);
    if (it == values_.end()) return dflt;
    std::istringstream is(it->second);
    V v{};
    is >> v;
    return v;
  }
  std::string get(const std::string &k, const char *dflt) const {
    auto it = values_.find(k);
    return it == values_.end() ? std::string(dflt) : it->second;
  }
  bool isParameter(const std::string &k) const { return values_.count(k) > 0; }

 private:
  std::map<std::string, std::string> values_;
};

// Ifpack point Jacobi, the only preconditioner of this path
class virtualPreconditioner {
 public:
  void initialize() { initialized_ = true; }
  void compute() { computed_ = true; }
  bool ready() const { return initialized_ && computed_; }

 private:
  bool initialized_ = false, computed_ = false;
};

class factoryPreconditioner {
 public:
  static std::shared_ptr<virtualPreconditioner> create(const ParameterList &p) {
    const std::string cls = p.get("precClass", "precIfpack");
    const std::string type = p.get("precType", "point relaxation");
    const std::string relax = p.get("relaxation: type", "Jacobi");
    if ((cls != "precIfpack" && cls != "1") || type != "point relaxation" || relax != "Jacobi")
      throw std::runtime_error("mgfe: only Ifpack point-relaxation Jacobi is provided");
    return std::make_shared<virtualPreconditioner>();
  }
};

// Block CG, block size 1, preconditioned by point Jacobi on the interior
// system.  solve() returns whether the explicit FP64 residual ||b - A x|| /
// ||b|| met "Convergence Tolerance" (belosBlockCG::solve returns true either
// way).  "Precision": "double" (default), "mixed" (FP32 recursion with FP64
// residual replacement) or "float+refine" (FP32 inner solves restarted by FP64
// refinement).
class virtualLinearSolver {
 public:
  explicit virtualLinearSolver(const ParameterList &p) {
    if (p.get<int>("Block Size", 1) != 1) throw std::runtime_error("mgfe: Block Size must be 1");
    maxit_ = p.get<int>("Maximum Iterations", 100000);
    tol_ = p.get<double>("Convergence Tolerance", 1e-10);
    const std::string prec = p.get("Precision", "double");
    mode_ = prec == "mixed" || prec == "float+reliable" ? MG_SOLVE_F32_RELIABLE
            : prec == "float+refine"                     ? MG_SOLVE_F32_REFINE
                                                         : MG_SOLVE_F64;
    innerTol_ = p.get<double>("Inner Tolerance", 0.0);
  }
  void setProblem(matrixBuilder &A, std::vector<double> &x, const std::vector<double> &b) {
    if (A.system() != MG_SYS_INTERIOR)
      throw std::runtime_error("mgfe: CG needs the interior (SPD) system");
    A_ = &A;
    x_ = &x;
    b_ = &b;
  }
  void setLeftPrec(const std::shared_ptr<virtualPreconditioner> &p) { prec_ = p; }
  void setRightPrec(const std::shared_ptr<virtualPreconditioner> &p) { prec_ = p; }
  bool solve() {
    if (!A_) throw std::runtime_error("mgfe: setProblem first");
    if (!prec_ || !prec_->ready())
      throw std::runtime_error("mgfe: set an initialized and computed point-Jacobi preconditioner");
    if (mode_ != MG_SOLVE_F64) A_->assemble(MG_PREC_F32);
    Context &ctx = A_->space().context();
    x_->resize(b_->size());
    mg_solver_params prm{mode_, tol_, innerTol_, maxit_, 30, 32, 1};
    const int rc = mg_pcg_solve(ctx.get(), &prm, b_->data(), x_->data(), &info_);
    if (rc != MG_OK && rc != MG_ERR_NOCONV) ctx.check(rc);
    return info_.status == MG_OK;
  }
  int getNumIters() const { return info_.iters; }
  const mg_solve_info &info() const { return info_; }

 private:
  matrixBuilder *A_ = nullptr;
  std::vector<double> *x_ = nullptr;
  const std::vector<double> *b_ = nullptr;
  std::shared_ptr<virtualPreconditioner> prec_;
  int maxit_ = 100000, mode_ = MG_SOLVE_F64;
  double tol_ = 1e-10, innerTol_ = 0.0;
  mg_solve_info info_{};
};

class factoryLinearSolver {
 public:
  static std::shared_ptr<virtualLinearSolver> create(const ParameterList &p) {
    const std::string s = p.get("Solver", "belosCG");
    if (s != "belosCG" && s != "2")
      throw std::runtime_error("mgfe: only belosCG (block CG, point Jacobi) is provided");
    return std::make_shared<virtualLinearSolver>(p);
  }
};

}  // namespace mgfe

#endif  // MGFE_HPP
