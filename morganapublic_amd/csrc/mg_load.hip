// Load vector (right-hand side) assembly and Dirichlet lifting.
//
// F_j = sum_e sum_q w_q det(J_e) f(e, Y_q) phi_j(Y_q): the functional fnEL3dA
// (f * v) integrated by integratorFEL3d_STD and summed by
// traitsVectorBuilder::buildEpetraVector (traitsVectorBuilder.hpp:217-274).
// As for the matrix, one lane owns one row (dof) and walks that row's
// incidence records (element e, local basis l) in their fixed order, so the
// sum is deterministic without atomics.  The quadrature is contracted once on
// the reference element: t1[l] = sum_q w_q phi_l(Y_q) (constant or P0 f) and
// tm[l][m] = sum_q w_q phi_l(Y_q) phi_m(Y_q) (f a field of the same element
// type, f(Y) = sum_m f_m phi_m(Y)); the element contributes
// det(J_e) (f_e t1[l]) or det(J_e) sum_m tm[l][m] f[dof(e, m)].
//
// Lifting of non-homogeneous Dirichlet data: b_int -= A[free, fixed] g, from
// the full FP64 matrix, each free row summing its fixed columns in CSR order.
#include "mg_internal.h"

namespace mg {

struct LoadTables {
  double t1[10];
  double tm[100];
};

static LoadTables load_tables(int order, int rule) {
  const int n = order == 1 ? 4 : 10;
  LoadTables t{};
  double pts[45], w[15], v[10];
  const int nq = quad_rule(rule, pts, w);
  for (int q = 0; q < nq; ++q) {
    ref_basis(order, &pts[3 * q], v, nullptr);
    for (int i = 0; i < n; ++i) {
      t.t1[i] += w[q] * v[i];
      for (int j = 0; j < n; ++j) t.tm[i * n + j] += w[q] * v[i] * v[j];
    }
  }
  return t;
}

__device__ __forceinline__ double tet_det(const double *__restrict__ xyz, int4 t) {
  const int id[4] = {t.x, t.y, t.z, t.w};
  double d[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) d[k][r] = xyz[3 * (size_t)id[k + 1] + r] - xyz[3 * (size_t)id[0] + r];
  // det of J = [d0 d1 d2] (columns P_k - P_0)
  return d[0][0] * (d[1][1] * d[2][2] - d[2][1] * d[1][2]) -
         d[1][0] * (d[0][1] * d[2][2] - d[2][1] * d[0][2]) +
         d[2][0] * (d[0][1] * d[1][2] - d[1][1] * d[0][2]);
}

// kind: MG_LOAD_CONST (f[0]), MG_LOAD_P0 (f[e]), MG_LOAD_NODAL (f on the
// full dof numbering); rec_stride: int32 words per incidence record
template <int NL>
__global__ __launch_bounds__(256) void k_load_rows(
    int64_t nrows, const int32_t *__restrict__ inc_ptr, const int32_t *__restrict__ rec,
    int rec_stride, const int32_t *__restrict__ tets, const double *__restrict__ xyz,
    const int32_t *__restrict__ dofs, int kind, const double *__restrict__ f, LoadTables tab,
    double *__restrict__ F) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
       r += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int i = inc_ptr[r]; i < inc_ptr[r + 1]; ++i) {
      const int32_t el = rec[(size_t)i * rec_stride];
      const int e = el >> 4, l = el & 15;
      const double det = tet_det(xyz, reinterpret_cast<const int4 *>(tets)[e]);
      double c;
      if (kind == MG_LOAD_NODAL) {
        c = 0.0;
#pragma unroll
        for (int m = 0; m < NL; ++m) c += tab.tm[l * NL + m] * f[dofs[(size_t)e * NL + m]];
      } else {
        c = (kind == MG_LOAD_P0 ? f[e] : f[0]) * tab.t1[l];
      }
      s += det * c;
    }
    F[r] = s;
  }
}

__global__ __launch_bounds__(256) void k_lift(int64_t n, const int32_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ colidx,
                                              const double *__restrict__ vals,
                                              const int32_t *__restrict__ sysmap,
                                              const uint8_t *__restrict__ fixed,
                                              const double *__restrict__ g,
                                              double *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t di = sysmap[i];
    if (di < 0) continue;
    double s = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int j = colidx[k];
      if (fixed[j]) s += vals[k] * g[j];
    }
    b[di] -= s;
  }
}

static int64_t load_len(Ctx *c, int kind) {
  return kind == MG_LOAD_CONST ? 1 : kind == MG_LOAD_P0 ? c->nT : c->ndofs;
}

static void load_launch(Ctx *c, int sys, int kind, const double *f, int rule, double *F) {
  System &S = c->sys[sys];
  const LoadTables tab = load_tables(c->order, rule);
  const int rs = S.rec_words;
  const int blocks = (int)std::min<int64_t>(blocks_for(S.nrows, 256), 4096);
  if (S.nrows == 0) return;
  if (c->nloc == 10)
    k_load_rows<10><<<blocks, 256, 0, c->stream>>>(S.nrows, S.inc_ptr, (const int32_t *)S.inc_rec,
                                                   rs, c->tets, c->xyz, c->dofs, kind, f, tab, F);
  else
    k_load_rows<4><<<blocks, 256, 0, c->stream>>>(S.nrows, S.inc_ptr, (const int32_t *)S.inc_rec,
                                                  rs, c->tets, c->xyz, c->dofs, kind, f, tab, F);
  MG_HIP(hipGetLastError());
}

static void lift_launch(Ctx *c, const double *g, double *b) {
  System &S = c->sys[MG_SYS_FULL];
  const int blocks = (int)std::min<int64_t>(blocks_for(S.nrows, 256), 4096);
  k_lift<<<blocks, 256, 0, c->stream>>>(S.nrows, S.rowptr, S.colidx, S.vals64, c->sysmap,
                                        c->fixed, g, b);
  MG_HIP(hipGetLastError());
}

static void load_checks(Ctx *c, int sys, int kind, int rule) {
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "load: bad system");
  MG_CHECK(kind == MG_LOAD_CONST || kind == MG_LOAD_P0 || kind == MG_LOAD_NODAL, MG_ERR_ARG,
           "load: bad coefficient kind");
  MG_CHECK(rule == 1 || rule == 4 || rule == 11 || rule == 15, MG_ERR_ARG,
           "load: rule must be 1,4,11,15");
  MG_CHECK(c->sys[sys].built, MG_ERR_STATE, "load: pattern not built");
}

}  // namespace mg

using namespace mg;

#define API_BEGIN(c)           \
  if (!(c)) return MG_ERR_ARG; \
  try {                        \
    MG_HIP(hipSetDevice((c)->device));
#define API_END(c)                  \
  }                                 \
  catch (const mg::Error &e) {      \
    (c)->err = e.msg;               \
    return e.code;                  \
  }                                 \
  catch (const std::exception &e) { \
    (c)->err = e.what();            \
    return MG_ERR_HIP;              \
  }                                 \
  return MG_OK;

extern "C" {

int mg_assemble_load(mg_ctx *c, int sys, int kind, const double *f, int64_t nf, int rule,
                     double *F) {
  API_BEGIN(c)
  load_checks(c, sys, kind, rule);
  MG_CHECK(f && F, MG_ERR_ARG, "load: null pointer");
  MG_CHECK(nf == load_len(c, kind), MG_ERR_ARG, "load: f has the wrong length");
  const int64_t n = c->sys[sys].nrows;
  DevTmp<double> df(c, nf), dF(c, n);
  MG_HIP(hipMemcpyAsync(df, f, sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
  load_launch(c, sys, kind, df, rule, dF);
  MG_HIP(hipMemcpyAsync(F, dF, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_assemble_load_device(mg_ctx *c, int sys, int kind, const double *f, int rule,
                            double *F) {
  API_BEGIN(c)
  load_checks(c, sys, kind, rule);
  MG_CHECK(f && F, MG_ERR_ARG, "load: null pointer");
  load_launch(c, sys, kind, f, rule, F);
  API_END(c)
}

int mg_dirichlet_lift(mg_ctx *c, const double *g, double *b) {
  API_BEGIN(c)
  MG_CHECK(g && b, MG_ERR_ARG, "lift: null pointer");
  MG_CHECK(c->dirichlet_ok, MG_ERR_STATE, "lift: Dirichlet dofs not set");
  MG_CHECK(c->sys[MG_SYS_FULL].built && c->sys[MG_SYS_FULL].has64, MG_ERR_STATE,
           "lift: full FP64 matrix not assembled");
  DevTmp<double> dg(c, c->ndofs), db(c, c->nfree);
  MG_HIP(hipMemcpyAsync(dg, g, sizeof(double) * c->ndofs, hipMemcpyHostToDevice, c->stream));
  MG_HIP(hipMemcpyAsync(db, b, sizeof(double) * c->nfree, hipMemcpyHostToDevice, c->stream));
  lift_launch(c, dg, db);
  MG_HIP(hipMemcpyAsync(b, db, sizeof(double) * c->nfree, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_dirichlet_lift_device(mg_ctx *c, const double *g, double *b) {
  API_BEGIN(c)
  MG_CHECK(g && b, MG_ERR_ARG, "lift: null pointer");
  MG_CHECK(c->dirichlet_ok, MG_ERR_STATE, "lift: Dirichlet dofs not set");
  MG_CHECK(c->sys[MG_SYS_FULL].built && c->sys[MG_SYS_FULL].has64, MG_ERR_STATE,
           "lift: full FP64 matrix not assembled");
  lift_launch(c, g, b);
  API_END(c)
}

}  // extern "C"
