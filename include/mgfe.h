/*
 * mgfe.h — C ABI of libmgfe.so, the MI355X (gfx950) implementation of
 * Morgana's finite-element assembly and Jacobi-PCG path.
 *
 * Conventions
 *   - every call returns an int status: MG_OK (0) or a negative MG_ERR_*;
 *     mg_last_error(ctx) returns the text of the last failure.  No C++
 *     exception crosses this boundary.
 *   - host pointers are caller-owned and only read/written during the call;
 *     device buffers belong to the context.  Functions suffixed _device take
 *     device pointers (memory on the context's GPU) and run on the context
 *     stream without host synchronisation unless stated.
 *   - one HIP stream per context; a context is not thread-safe.
 *   - indices: mesh vertex ids in tets/btri are 1-based (gmsh / Morgana gids);
 *     every other index (dofs, CSR rows/columns, edges) is 0-based, i.e. the
 *     Epetra gid = Morgana gid - 1 (localMatrixEL3d.hpp:191-199, base 0).
 *   - values are double unless the name or a prec argument says FP32.
 *
 * Reference interfaces replaced (paths relative to the morganaPublic tree):
 *   meshInit3d::gmMesh_to_stdB + meshDoctor3d::checkJacobian  -> mg_mesh_set
 *   connect3d::buildConnectivity (edges) + dofMapStatic3d::startup
 *                                                             -> mg_dofmap_build
 *   traitsMatrixBuilder::buildEpetraCrs (pattern)             -> mg_pattern_build
 *   traitsMatrixBuilder::buildEpetraCrs (values), opEL3dA/F/G/L,
 *   integratorOEL3d_STD                                       -> mg_assemble
 *   Epetra_CrsMatrix::Apply                                    -> mg_spmv
 *   factoryLinearSolver(belosCG) + preconditionerIfpack(Jacobi) -> mg_pcg_solve
 *   dual3d<linearTetra>::buildFiniteVolumeData                -> mg_dual_fv_build
 */
#ifndef MGFE_H
#define MGFE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define MG_OK 0
#define MG_ERR_ARG -1       /* bad argument (null pointer, out-of-range value) */
#define MG_ERR_HIP -2       /* a HIP runtime call failed                       */
#define MG_ERR_STATE -3     /* call made before its prerequisite               */
#define MG_ERR_MESH -4      /* a tetrahedron with det(J) <= 0 or a bad id      */
#define MG_ERR_NOCONV -5    /* solver reached its iteration limit              */
#define MG_ERR_BREAKDOWN -6 /* p.Ap <= 0 or NaN inside CG                      */
#define MG_ERR_SIZE -7      /* problem exceeds an index-width limit            */

/* ---- systems, operators, precisions ------------------------------------- */
#define MG_SYS_FULL 0     /* all dofs: the matrix buildEpetraCrs produces    */
#define MG_SYS_INTERIOR 1 /* Dirichlet dofs eliminated (SPD system for PCG)  */

#define MG_TERM_STIFF 1 /* c grad(u).grad(v)  (opEL3dA, opEL3dF)             */
#define MG_TERM_MASS 2  /* c u v              (opEL3dG, opEL3dL)             */

#define MG_COEFF_CONST 0 /* scalar coefficient                              */
#define MG_COEFF_P0 1    /* per-element (P0) field set by mg_field_set      */

#define MG_FIELD_STIFF 0 /* P0 coefficient of the stiffness term            */
#define MG_FIELD_MASS 1  /* P0 coefficient of the mass term                 */

#define MG_PREC_F64 0
#define MG_PREC_F32 1 /* FP32 element contraction and FP32 values         */

#define MG_SOLVE_F64 0        /* plain FP64 Jacobi-PCG                        */
#define MG_SOLVE_F32_REFINE 1 /* FP32 inner Jacobi-PCG, FP64 refinement      */
/* FP32 Jacobi-PCG whose residual is replaced by the FP64 residual b - A x
 * (x accumulated in FP64) each time the recursive residual has dropped by
 * inner_tol; the search direction is kept across replacements. */
#define MG_SOLVE_F32_RELIABLE 2

typedef struct mg_ctx mg_ctx;

typedef struct {
  int terms;          /* MG_TERM_STIFF | MG_TERM_MASS                         */
  int stiff_coeff;    /* MG_COEFF_CONST or MG_COEFF_P0                        */
  double stiff_const; /* used when stiff_coeff == MG_COEFF_CONST              */
  int stiff_rule;     /* quadrature points: 1, 4, 11 or 15                    */
  int mass_coeff;
  double mass_const;
  int mass_rule;
} mg_operator;

typedef struct {
  int mode;          /* MG_SOLVE_F64, MG_SOLVE_F32_REFINE or MG_SOLVE_F32_RELIABLE */
  double tol;        /* ||b - A x||_2 / ||b||_2 target (FP64 residual)          */
  double inner_tol;  /* F32_REFINE: relative tolerance of each FP32 inner solve;
                        F32_RELIABLE: residual drop between FP64 replacements */
  int max_iters;     /* bound on the total number of (inner) CG iterations     */
  int max_outer;     /* bound on refinement steps (F32_REFINE)                 */
  int check_every;   /* iterations queued between host convergence checks     */
  int use_x0;        /* nonzero: x holds the initial guess                     */
} mg_solver_params;

typedef struct {
  int status;         /* MG_OK, MG_ERR_NOCONV or MG_ERR_BREAKDOWN              */
  int iters;          /* CG iterations (sum over inner solves)                 */
  int outer;          /* refinement steps (0 for MG_SOLVE_F64)                 */
  double relres;      /* final ||b - A x||_2 / ||b||_2 computed in FP64         */
  double solve_ms;    /* device time of the solve (HIP events)                 */
} mg_solve_info;

/* ---- context ------------------------------------------------------------- */
int mg_ctx_create(int device, mg_ctx **out);
int mg_ctx_destroy(mg_ctx *ctx);
const char *mg_last_error(const mg_ctx *ctx);
/* the context's HIP stream (hipStream_t) for callers that enqueue their own work */
void *mg_ctx_stream(mg_ctx *ctx);
int mg_ctx_synchronize(mg_ctx *ctx);
/* Measurement utility, no counterpart in the reference: a plain device copy of
 * nbytes (a multiple of 16, 16-byte aligned pointers) on the context's stream,
 * one read and one write of every byte.  Calibrates hardware byte counters and
 * gives the attainable HBM copy bandwidth next to the kernels' rooflines. */
int mg_copy_device(mg_ctx *ctx, void *dst, const void *src, int64_t nbytes);

/* ---- mesh ----------------------------------------------------------------
 * replaces meshInit3d::gmMesh_to_stdB (meshInit3d.hpp:487-690) fed by the gmsh
 * 2.1 loader (loadMesh.hpp:2065-2350) and the det > 0 assertion of
 * meshDoctor3d::checkJacobian (meshDoctor3d.hpp:195-227): fails with
 * MG_ERR_MESH if a tetrahedron has det(J) <= 0 or a vertex id is out of range.
 * xyz[3 nV]; tets[4 nT] 1-based; btri[3 nB] 1-based boundary triangles with
 * their geoIds bgeo[nB] (the second gmsh tag).                              */
int mg_mesh_set(mg_ctx *ctx, int64_t nV, const double *xyz, int64_t nT, const int32_t *tets,
                int64_t nB, const int32_t *btri, const int32_t *bgeo);

/* P0 field (one value per element) on the device: feStaticField3d<fePr3d<0>>
 * setDofL (feStaticField3d.hpp) followed by operator setCoeff.             */
int mg_field_set(mg_ctx *ctx, int which, const double *values);
int mg_field_set_device(mg_ctx *ctx, int which, const double *values_device);

/* ---- dof map ---------------------------------------------------------------
 * replaces connect3d::buildConnectivity edge numbering (connect3d.hpp:417-491,
 * lexicographic global edge gids of pVectGlobalManip.hpp:417-578) and
 * dofMapStatic3d::startup (dofMapStatic3d.hpp:471-556): vertex dof = node gid,
 * P2 edge dof = numVertices + edge gid.                                     */
int mg_dofmap_build(mg_ctx *ctx, int order, int64_t *n_dofs, int64_t *n_edges);
/* feStaticInterface3d::feIndicesLG for every element: dofs[nT * nloc]      */
int mg_elem_dofs_get(mg_ctx *ctx, int32_t *dofs);
/* edges as sorted 0-based vertex pairs, edge_nodes[2 nE]                   */
int mg_edges_get(mg_ctx *ctx, int32_t *edge_nodes, int32_t *elem_edge);

/* Dirichlet dofs: every dof on a boundary triangle whose geoId is listed
 * (boundary-card selection of bcCardFE / boundCard).  Defines the numbering
 * of MG_SYS_INTERIOR (free dofs in increasing order).                      */
int mg_dirichlet_set(mg_ctx *ctx, int n_geo, const int32_t *geo_ids, int64_t *n_free);
int mg_dirichlet_mask_get(mg_ctx *ctx, uint8_t *fixed);

/* ---- sparsity pattern ------------------------------------------------------
 * replaces the structure part of traitsMatrixBuilder::buildEpetraCrs
 * (traitsMatrixBuilder.hpp:366-441) + Epetra GlobalAssemble: every (row,col)
 * pair of dofs sharing a tetrahedron, columns sorted, structural zeros kept.
 * Also builds the row-incidence records the numeric assembly uses (the
 * element -> CSR slot map is built on the way and released).  The pattern
 * lives until mg_mesh_set,
 * mg_dofmap_build or (interior system) mg_dirichlet_set; calling this again
 * before that returns its sizes and keeps the assembled values.            */
int mg_pattern_build(mg_ctx *ctx, int sys, int64_t *n_rows, int64_t *nnz);
int mg_pattern_get(mg_ctx *ctx, int sys, int32_t *rowptr, int32_t *colidx);
/* slot[e*nloc*nloc + i*nloc + j] = CSR index of (row dof j, column dof i) of
 * element e, in the column-major order of the reference's local matrix
 * (opColMajor, traitsMatrixBuilder.hpp:57-62); -1 for eliminated pairs.
 * Recomputed on the device for each call (the assembly does not use it, so
 * it is not kept resident: 4 nloc^2 bytes per element).                    */
int mg_slot_map_get(mg_ctx *ctx, int sys, int32_t *slot);

/* ---- numeric assembly ------------------------------------------------------
 * replaces the value part of traitsMatrixBuilder::buildEpetraCrs with the
 * local matrices of integratorOEL3d_STD::integration (integratorOEL3d_STD.hpp:
 * 142-170) over opEL3dA/opEL3dF/opEL3dG/opEL3dL (opEL3dA.hpp:129-165,
 * opEL3dG.hpp:83-108).  Values stay on the device; mg_values_get copies them.
 * prec = MG_PREC_F32 fills the FP32 copy (geometry in FP64, contraction and
 * accumulation in FP32).                                                    */
int mg_assemble(mg_ctx *ctx, int sys, const mg_operator *op, int prec);
/* Allocation-free variant of mg_assemble.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and the same values, bit for bit; the per-element factor records
 * go to a static 128 MiB device array (one per GPU, held until the process
 * ends) instead of context memory, and nothing is allocated or freed.
 * MG_ERR_SIZE when the records need more than the array.  Records are
 * needed by FP32 P2 assemblies (24 bytes per element, 32 with a mass term:
 * up to 5.59 M / 4.19 M elements) and by FP64 P2 assemblies of a pattern
 * with row blocks that have no vertex patch (48 / 64 bytes: 2.79 M /
 * 2.09 M); P1 and fully patched FP64 P2 assemblies take none.  The pattern must have been
 * built by mg_pattern_build (which allocates).  Every context on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * host threads) race.                                                      */
int mg_assemble_noalloc(mg_ctx *ctx, int sys, const mg_operator *op, int prec);
int mg_values_get(mg_ctx *ctx, int sys, int prec, void *vals);
/* device pointer of the nnz assembled values (double* or float*), in CSR
 * order, owned by the context until the pattern is rebuilt or destroyed     */
int mg_values_device(mg_ctx *ctx, int sys, int prec, void **vals_device);
/* assembly sequence number of the values a system currently holds in `prec`
 * (increasing over the context's assemblies; 0: none).  No counterpart in the
 * reference (an Epetra matrix is a new object per buildEpetraCrs call): lets
 * a caller that keeps a matrix across calls check that it has not been
 * re-assembled in between, as the adjoint solve of a gradient must.         */
int mg_assembly_seq(mg_ctx *ctx, int sys, int prec, uint64_t *seq);

/* ---- load vector and Dirichlet lifting -------------------------------------
 * replaces traitsVectorBuilder::buildEpetraVector (traitsVectorBuilder.hpp:217-274)
 * over fnEL3dA (f v) integrated by integratorFEL3d_STD
 * (integratorFEL3d_STD.hpp:99-127): F_j = sum_e int_e f phi_j with `rule`
 * points, on the rows of `sys` (full or interior numbering).  f: MG_LOAD_CONST
 * (nf = 1), MG_LOAD_P0 (nf = nT, one value per element) or MG_LOAD_NODAL
 * (nf = ndofs, a field of the same element type on the full numbering).     */
#define MG_LOAD_CONST 0
#define MG_LOAD_P0 1
#define MG_LOAD_NODAL 2
int mg_assemble_load(mg_ctx *ctx, int sys, int kind, const double *f, int64_t nf, int rule,
                     double *F);
int mg_assemble_load_device(mg_ctx *ctx, int sys, int kind, const double *f, int rule,
                            double *F);
/* Non-homogeneous Dirichlet data g (full numbering, read at the fixed dofs):
 * b (interior numbering) -= A_full[free, fixed] g, with the full FP64 matrix
 * assembled.  The reference imposes such data through its Lagrange
 * multiplier system (operatorLA3d.hpp:21-90); elimination moves it to the
 * right-hand side instead. */
int mg_dirichlet_lift(mg_ctx *ctx, const double *g, double *b);
int mg_dirichlet_lift_device(mg_ctx *ctx, const double *g, double *b);

/* ---- linear algebra ----------------------------------------------------------
 * Epetra_CrsMatrix::Apply: y = A x on the assembled system.                */
int mg_spmv(mg_ctx *ctx, int sys, int prec, const double *x, double *y);
int mg_spmv_device(mg_ctx *ctx, int sys, int prec, const void *x_device, void *y_device);

/* factoryLinearSolver{"Solver"=belosCG,"Maximum Iterations","Convergence
 * Tolerance"} + setProblem/solve (factoryLinearSolver.hpp:51-91,
 * belosBlockCG.hpp:46-83) preconditioned by Ifpack point Jacobi
 * (preconditionerIfpack.cpp:26-63), on MG_SYS_INTERIOR.  Unlike
 * belosBlockCG::solve (which returns true regardless), the status reports
 * convergence of the explicit FP64 residual ||b - A x|| / ||b||.            */
int mg_pcg_solve(mg_ctx *ctx, const mg_solver_params *p, const double *b, double *x,
                 mg_solve_info *info);
int mg_pcg_solve_device(mg_ctx *ctx, const mg_solver_params *p, const double *b_device,
                        double *x_device, mg_solve_info *info);

/* ---- gradients (not in the reference) --------------------------------------
 * dJ/dk_e = - lam_e^T Khat_e u_e for the stiffness term with P0 coefficient
 * (Khat_e = the element stiffness at unit coefficient); u and lam on the
 * numbering of sys (zero on eliminated dofs).                              */
int mg_adjoint_grad_k(mg_ctx *ctx, int sys, int rule, const double *u, const double *lam,
                      double *dJdk);
int mg_adjoint_grad_k_device(mg_ctx *ctx, int sys, int rule, const double *u_device,
                             const double *lam_device, double *dJdk_device);

/* ---- dual mesh and Scharfetter-Gummel --------------------------------------
 * dual3d<linearTetra>::buildFiniteVolumeData (dual3d.hpp:606-705, 709-1025):
 * edge_surf[nE], edge_n[3 nE] (unit, oriented low->high vertex id),
 * cell_volume[nV]; sigma_el[6 nT] is the per-element share of edge_surf.   */
int mg_dual_fv_build(mg_ctx *ctx);
int mg_dual_fv_get(mg_ctx *ctx, double *edge_surf, double *edge_n, double *cell_volume,
                   double *sigma_el);

/* Drift-diffusion residual at the state (phi, dens) on the P1 vertex dofs:
 *  R_phi = K(eps) phi - b(rho), R_n = sum of SG edge fluxes, zero at fixed
 *  vertices; loss = sum R_phi^2 + R_n^2.  Per-element eps, rho, mu.         */
int mg_dd_residual(mg_ctx *ctx, const double *eps, const double *rho, const double *mu,
                   const double *phi, const double *dens, const uint8_t *fixed, double *r_phi,
                   double *r_n, double *loss);
/* d loss / d eps and d loss / d mu (FP64 VJP kernels)                      */
int mg_dd_vjp(mg_ctx *ctx, const double *eps, const double *rho, const double *mu,
              const double *phi, const double *dens, const uint8_t *fixed, double *g_eps,
              double *g_mu);
/* The same on device buffers (the context's GPU), queued on the context
 * stream without host synchronisation.  r_phi, r_n (nV) receive the
 * residuals; loss, when not null, is a device double receiving the same
 * value mg_dd_residual returns. */
int mg_dd_residual_device(mg_ctx *ctx, const double *eps, const double *rho, const double *mu,
                          const double *phi, const double *dens, const uint8_t *fixed,
                          double *r_phi, double *r_n, double *loss);
int mg_dd_vjp_device(mg_ctx *ctx, const double *eps, const double *rho, const double *mu,
                     const double *phi, const double *dens, const uint8_t *fixed, double *r_phi,
                     double *r_n, double *g_eps, double *g_mu);
/* Bernoulli function B(x) = x / (e^x - 1) in FP64 on the device (n values)  */
int mg_bernoulli(mg_ctx *ctx, int64_t n, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif /* MGFE_H */
