/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C, FP64, single-threaded restatement of Morgana's finite-element
 * assembly and linear-solve path (feof81/morganaPublic).  It exists to CHECK
 * the HIP library (libmgfe.so) and to time the reference algorithm on the
 * host CPU.  Only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  The product never links, calls or falls back to it.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Pinning (what checks this code independently of its own formulas) is listed
 * per function; functions without an independent pin say "parity unpinned".
 *
 * Conventions: vertex ids in `tets` are 1-based (gmsh/Morgana gids), every
 * other index array is 0-based (Epetra gid = Morgana gid - 1).
 */
#ifndef MG_ORACLE_H
#define MG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* operator kinds (one element matrix may combine a stiffness and a mass part) */
#define OR_OP_STIFF 1 /* opEL3dA / opEL3dF : c * grad(u).grad(v) */
#define OR_OP_MASS 2  /* opEL3dG / opEL3dL : c * u v             */

int or_quad_rule(int npts, double *pts, double *w);
void or_basis(int order, const double *y, double *val, double *gx, double *gy, double *gz);
void or_tet_jacobian(const double *P, double *J);
double or_det3(const double *J);
void or_inv3(const double *J, double *Ji);
void or_global_grad(int order, const double *P, const double *y, double *gx, double *gy,
                    double *gz);
int or_element_matrix(int ops, int order, const double *P, double c_stiff, double c_mass,
                      int rule_stiff, int rule_mass, double *mat);

int64_t or_edges(int64_t nV, int64_t nT, const int32_t *tets, int32_t *elem_edge,
                 int32_t *edge_nodes);
void or_elem_dofs(int order, int64_t nV, int64_t nT, const int32_t *tets,
                  const int32_t *elem_edge, int32_t *dofs);

int or_assemble(int order, int ops, int64_t nV, const double *xyz, int64_t nT,
                const int32_t *tets, const double *c_stiff, const double *c_mass,
                int rule_stiff, int rule_mass, int64_t *nrows, int64_t *nnz,
                int32_t **rowptr, int32_t **colidx, double **vals);
int64_t or_accumulate(int ops, int order, const double *xyz, const int32_t *tets,
                      const int32_t *sdofs, int64_t e0, int64_t e1, const double *c_stiff,
                      const double *c_mass, int rule_stiff, int rule_mass,
                      const int32_t *rowptr, const int32_t *colidx, double *vals);
int or_load_vector(int order, int64_t nV, const double *xyz, int64_t nT, const int32_t *tets,
                   int kind, const double *f, int rule, double *F);
int or_restrict(int64_t nrows, const int32_t *rowptr, const int32_t *colidx,
                const double *vals, const uint8_t *fixed, int64_t *nint, int64_t *nnz_int,
                int32_t **rowptr_int, int32_t **colidx_int, double **vals_int);
void or_free(void *p);

void or_spmv(int64_t n, const int32_t *rowptr, const int32_t *colidx, const double *vals,
             const double *x, double *y);
double or_pcg_fixed(int64_t n, const int32_t *rowptr, const int32_t *colidx, const double *vals,
                    const double *b, double *x, int iters, int threads, double *relres);
int or_pcg(int64_t n, const int32_t *rowptr, const int32_t *colidx, const double *vals,
           const double *b, double *x, double tol, int maxit, int *iters, double *relres);

void or_dirichlet_mask(int order, int64_t nV, int64_t nB, const int32_t *btri,
                       const int32_t *bgeo, int ngeo, const int32_t *geo_list, int64_t nE,
                       const int32_t *edge_nodes, uint8_t *fixed);

void or_dual_fv(int64_t nV, const double *xyz, int64_t nT, const int32_t *tets, int64_t nE,
                const int32_t *elem_edge, const int32_t *edge_nodes, double *edge_surf,
                double *edge_n, double *cell_volume, double *sigma_el);
double or_bernoulli(double x);
double or_dd_loss(int64_t nV, const double *xyz, int64_t nT, const int32_t *tets, int64_t nE,
                  const int32_t *elem_edge, const int32_t *edge_nodes,
                  const double *sigma_el, const double *eps, const double *rho,
                  const double *mu, const double *phi, const double *dens,
                  const uint8_t *fixed, double *r_phi, double *r_n);
void or_dd_grad(int64_t nV, const double *xyz, int64_t nT, const int32_t *tets, int64_t nE,
                const int32_t *elem_edge, const int32_t *edge_nodes, const double *sigma_el,
                const double *eps, const double *rho, const double *mu, const double *phi,
                const double *dens, const uint8_t *fixed, double *g_eps, double *g_mu);

void or_adjoint_grad_k(int order, int64_t nV, const double *xyz, int64_t nT,
                       const int32_t *tets, const int32_t *dofs, int rule,
                       const double *u, const double *lam, double *dJdk);

#ifdef __cplusplus
}
#endif
#endif
