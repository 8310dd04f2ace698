// Numeric assembly of P1/P2 tetrahedral operators into a fixed CSR pattern,
// and the adjoint element kernel for d(loss)/d(per-element coefficient).
//
// Two kernels per assembly:
//  1. k_geom (one thread per element): vertex coordinates -> J, det J (signed,
//     as geoMapInterface::getGradientDet), J^{-1} by cofactors (tensor3d::
//     computeInverse), and the 7 per-element factors the local matrix needs:
//     c det (J^{-1} J^{-T})_k for the 6 symmetric pairs k, and c_mass det.
//  2. k_asm_batched (one wavefront per schedule block of consecutive CSR rows,
//     the block's entries staged in LDS): the block's (element, local basis l)
//     incidences come in precomputed batches of <= 64 that share l and hit 64
//     distinct rows.  Each lane forms row l of its element's matrix as
//     sum_k factor_k tab[k][l][m] (tab = the quadrature of integratorOEL3d_STD
//     contracted once on the reference element, in constant memory, read with
//     scalar loads because l is uniform in the batch) and adds it into its
//     row's LDS entries.  Lanes of a batch never collide and batches run in a
//     fixed order, so the sums are deterministic without atomics.  The block's
//     contiguous CSR value range is then written with coalesced stores.
#include <cstring>

#include "mg_internal.h"

namespace mg {

__constant__ double c_tab64[7 * 100];
__constant__ float c_tab32[7 * 100];

struct TabKey {
  int order = -1, stiff_rule = -1, mass_rule = -1, prec = -1;
  bool operator==(const TabKey &o) const {
    return order == o.order && stiff_rule == o.stiff_rule && mass_rule == o.mass_rule &&
           prec == o.prec;
  }
};

// per device: the table currently resident in constant memory
static TabKey g_tab_key[64];

static void upload_tables(Ctx *c, int order, int stiff_rule, int mass_rule, int prec) {
  TabKey k{order, stiff_rule, mass_rule, prec};
  if (c->device < 64 && g_tab_key[c->device] == k) return;
  double tab[700];
  build_tables(order, stiff_rule, mass_rule, tab);
  const int n = order == 1 ? 4 : 10;
  if (prec == MG_PREC_F64) {
    MG_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tab64), tab, sizeof(double) * 7 * n * n, 0,
                                  hipMemcpyHostToDevice, c->stream));
  } else {
    float t32[700];
    for (int i = 0; i < 7 * n * n; ++i) t32[i] = (float)tab[i];
    MG_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tab32), t32, sizeof(float) * 7 * n * n, 0,
                                  hipMemcpyHostToDevice, c->stream));
  }
  MG_HIP(hipStreamSynchronize(c->stream));
  if (c->device < 64) g_tab_key[c->device] = k;
}

template <typename T>
__device__ __forceinline__ T tab(int i);
template <>
__device__ __forceinline__ double tab<double>(int i) {
  return c_tab64[i];
}
template <>
__device__ __forceinline__ float tab<float>(int i) {
  return c_tab32[i];
}

struct GeomOut {
  double det;
  double A[6];  // (J^{-1} J^{-T}) for (00,11,22,01,02,12)
};

__device__ __forceinline__ GeomOut tet_geometry(const double *__restrict__ xyz, const int32_t *id) {
  double T[3][3];
  const double x0 = xyz[3 * id[0]], y0 = xyz[3 * id[0] + 1], z0 = xyz[3 * id[0] + 2];
  for (int cc = 0; cc < 3; ++cc) {
    T[0][cc] = xyz[3 * id[cc + 1]] - x0;
    T[1][cc] = xyz[3 * id[cc + 1] + 1] - y0;
    T[2][cc] = xyz[3 * id[cc + 1] + 2] - z0;
  }
  const double d = T[0][0] * T[1][1] * T[2][2] - T[0][0] * T[1][2] * T[2][1] -
                   T[1][0] * T[0][1] * T[2][2] + T[1][0] * T[0][2] * T[2][1] +
                   T[2][0] * T[0][1] * T[1][2] - T[2][0] * T[0][2] * T[1][1];
  const double r = 1.0 / d;
  double I[3][3];
  I[0][0] = (T[1][1] * T[2][2] - T[1][2] * T[2][1]) * r;
  I[0][1] = -(T[0][1] * T[2][2] - T[0][2] * T[2][1]) * r;
  I[0][2] = (T[0][1] * T[1][2] - T[0][2] * T[1][1]) * r;
  I[1][0] = -(T[1][0] * T[2][2] - T[1][2] * T[2][0]) * r;
  I[1][1] = (T[0][0] * T[2][2] - T[0][2] * T[2][0]) * r;
  I[1][2] = -(T[0][0] * T[1][2] - T[0][2] * T[1][0]) * r;
  I[2][0] = (T[1][0] * T[2][1] - T[1][1] * T[2][0]) * r;
  I[2][1] = -(T[0][0] * T[2][1] - T[0][1] * T[2][0]) * r;
  I[2][2] = (T[0][0] * T[1][1] - T[0][1] * T[1][0]) * r;
  GeomOut g;
  g.det = d;
  const int KA[6] = {0, 1, 2, 0, 0, 1}, KB[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int k = 0; k < 6; ++k)
    g.A[k] = I[KA[k]][0] * I[KB[k]][0] + I[KA[k]][1] * I[KB[k]][1] + I[KA[k]][2] * I[KB[k]][2];
  return g;
}

template <typename T>
__global__ void k_geom(int64_t nT, const int32_t *__restrict__ tets, const double *__restrict__ xyz,
                       double ks_const, const double *__restrict__ ks_field, double km_const,
                       const double *__restrict__ km_field, T *__restrict__ geom) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nT) return;
  int32_t id[4];
  const int4 t4 = *reinterpret_cast<const int4 *>(tets + 4 * e);
  id[0] = t4.x;
  id[1] = t4.y;
  id[2] = t4.z;
  id[3] = t4.w;
  const GeomOut g = tet_geometry(xyz, id);
  const double ks = (ks_field ? ks_field[e] : ks_const) * g.det;
  const double km = (km_field ? km_field[e] : km_const) * g.det;
  T out[8];
#pragma unroll
  for (int k = 0; k < 6; ++k) out[k] = (T)(ks * g.A[k]);
  out[6] = (T)km;
  out[7] = (T)0;
  T *dst = geom + 8 * e;
#pragma unroll
  for (int k = 0; k < 8; ++k) dst[k] = out[k];
}

template <typename T>
struct Geom8 {
  T v[8];
};

template <typename T>
__device__ __forceinline__ Geom8<T> load_geom(const T *__restrict__ geom, int32_t el) {
  Geom8<T> g;
  const T *src = geom + 8 * (size_t)(el >> 4);
#pragma unroll
  for (int k = 0; k < 8; ++k) g.v[k] = src[k];
  return g;
}

// One wavefront per schedule block.  Two-deep software pipeline: while batch b
// is contracted and added, the geometry of batch b+1 and the records of batch
// b+2 are in flight.
template <int NL, typename T, bool MASS, typename Rec>
__global__ __launch_bounds__(64) void k_asm_batched(const AsmBlock *__restrict__ blocks,
                                                    const AsmBatch *__restrict__ batches,
                                                    const Rec *__restrict__ rec,
                                                    const int32_t *__restrict__ rowptr,
                                                    const T *__restrict__ geom,
                                                    T *__restrict__ vals, int slots_cap) {
  extern __shared__ __align__(16) char lds_raw[];
  T *acc = reinterpret_cast<T *>(lds_raw);
  int32_t *lrp = reinterpret_cast<int32_t *>(acc + slots_cap + 1);
  const int lane = threadIdx.x;
  const AsmBlock B = blocks[blockIdx.x];
  const int s0 = rowptr[B.row_begin];
  const int ns = rowptr[B.row_end] - s0;
  for (int s = lane; s < ns; s += 64) acc[s] = (T)0;
  for (int i = lane; i <= B.row_end - B.row_begin; i += 64) lrp[i] = rowptr[B.row_begin + i] - s0;
  __syncthreads();
  const int dummy = slots_cap;  // target of eliminated columns, never written out

  int b = B.batch_begin;
  const int bend = B.batch_end;
  AsmBatch m0{0, 0};
  if (b < bend) m0 = batches[b];
  Rec r0{}, r1{};
  int c0 = m0.meta >> 8, c1 = 0;
  if (lane < c0) r0 = rec[m0.offset + lane];
  Geom8<T> g0{};
  if (lane < c0) g0 = load_geom(geom, r0.el);
  AsmBatch m1{0, 0};
  if (b + 1 < bend) {
    m1 = batches[b + 1];
    c1 = m1.meta >> 8;
    if (lane < c1) r1 = rec[m1.offset + lane];
  }
  for (; b < bend; ++b) {
    // prefetch: geometry of b+1, records of b+2
    Geom8<T> g1{};
    if (lane < c1) g1 = load_geom(geom, r1.el);
    Rec r2{};
    int c2 = 0;
    if (b + 2 < bend) {
      const AsmBatch m2 = batches[b + 2];
      c2 = m2.meta >> 8;
      if (lane < c2) r2 = rec[m2.offset + lane];
    }
    const int l = __builtin_amdgcn_readfirstlane(m0.meta & 0xFF);
    if (lane < c0) {
      const int base = lrp[r0.row];
      T v[NL];
#pragma unroll
      for (int m = 0; m < NL; ++m) {
        T s = g0.v[0] * tab<T>((0 * NL + l) * NL + m);
#pragma unroll
        for (int k = 1; k < 6; ++k) s += g0.v[k] * tab<T>((k * NL + l) * NL + m);
        if (MASS) s += g0.v[6] * tab<T>((6 * NL + l) * NL + m);
        v[m] = s;
      }
      int idx[NL];
#pragma unroll
      for (int m = 0; m < NL; ++m) idx[m] = r0.pos[m] == 0xFF ? dummy : base + r0.pos[m];
      T a[NL];
#pragma unroll
      for (int m = 0; m < NL; ++m) a[m] = acc[idx[m]];
#pragma unroll
      for (int m = 0; m < NL; ++m) acc[idx[m]] = a[m] + v[m];
    }
    // rotate the pipeline
    m0 = m1;
    c0 = c1;
    r0 = r1;
    g0 = g1;
    c1 = c2;
    r1 = r2;
    if (b + 2 < bend) m1 = batches[b + 2];
  }
  __syncthreads();
  for (int s = lane; s < ns; s += 64) vals[s0 + s] = acc[s];
}

template <int NL, typename T, typename Rec>
static void launch_rows(Ctx *c, System &S, bool mass, const T *geom, T *vals) {
  const size_t lds = sizeof(T) * (size_t)(S.asm_slots + 1) + sizeof(int32_t) * (S.asm_rows + 1);
  const int nblk = (int)S.n_blocks;
  if (mass)
    k_asm_batched<NL, T, true, Rec><<<nblk, 64, lds, c->stream>>>(
        S.blocks, S.batches, (const Rec *)S.brec, S.rowptr, geom, vals, S.asm_slots);
  else
    k_asm_batched<NL, T, false, Rec><<<nblk, 64, lds, c->stream>>>(
        S.blocks, S.batches, (const Rec *)S.brec, S.rowptr, geom, vals, S.asm_slots);
  MG_HIP(hipGetLastError());
}

static bool rule_ok(int r) { return r == 1 || r == 4 || r == 11 || r == 15; }

template <typename T>
static void assemble_t(Ctx *c, int sys, const mg_operator *op, int prec) {
  System &S = c->sys[sys];
  const bool stiff = op->terms & MG_TERM_STIFF, mass = op->terms & MG_TERM_MASS;
  upload_tables(c, c->order, stiff ? op->stiff_rule : 0, mass ? op->mass_rule : 0, prec);
  const size_t gb = sizeof(T) * 8 * (size_t)c->nT;
  if (c->geom_bytes < gb) {
    dfree(c, c->geom);
    c->geom = dalloc<char>(c, gb);
    c->geom_bytes = gb;
  }
  T *geom = (T *)c->geom;
  const double ksc = stiff ? op->stiff_const : 0.0, kmc = mass ? op->mass_const : 0.0;
  const double *ksf = (stiff && op->stiff_coeff == MG_COEFF_P0) ? c->field[MG_FIELD_STIFF] : nullptr;
  const double *kmf = (mass && op->mass_coeff == MG_COEFF_P0) ? c->field[MG_FIELD_MASS] : nullptr;
  k_geom<T><<<blocks_for(c->nT, 256), 256, 0, c->stream>>>(c->nT, c->tets, c->xyz, ksc, ksf, kmc,
                                                         kmf, geom);
  MG_HIP(hipGetLastError());
  T *vals = prec == MG_PREC_F64 ? (T *)S.vals64 : (T *)S.vals32;
  if (c->nloc == 10)
    launch_rows<10, T, BRecP2>(c, S, mass, geom, vals);
  else
    launch_rows<4, T, BRecP1>(c, S, mass, geom, vals);
}

// ---- adjoint element kernel ---------------------------------------------------
// dJ/dk_e = - lam_e^T Khat_e u_e, Khat_e = sum_k det A_k tab[k] (unit coefficient)
template <int NL>
__global__ void k_adjoint_k(int64_t nT, const int32_t *__restrict__ tets,
                            const double *__restrict__ xyz, const int32_t *__restrict__ dofs,
                            const int32_t *__restrict__ map, const double *__restrict__ u,
                            const double *__restrict__ lam, double *__restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nT) return;
  int32_t id[4];
  for (int v = 0; v < 4; ++v) id[v] = tets[4 * e + v];
  const GeomOut g = tet_geometry(xyz, id);
  double ue[NL], le[NL];
#pragma unroll
  for (int m = 0; m < NL; ++m) {
    int32_t d = dofs[NL * e + m];
    if (map) d = map[d];
    ue[m] = d >= 0 ? u[d] : 0.0;
    le[m] = d >= 0 ? lam[d] : 0.0;
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double t = 0.0;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      double w = 0.0;
#pragma unroll
      for (int m = 0; m < NL; ++m) w += c_tab64[(k * NL + l) * NL + m] * ue[m];
      t += le[l] * w;
    }
    s += g.det * g.A[k] * t;
  }
  out[e] = -s;
}

static void adjoint_launch(Ctx *c, int sys, int rule, const double *u, const double *lam,
                           double *out) {
  upload_tables(c, c->order, rule, 0, MG_PREC_F64);
  const int32_t *map = sys == MG_SYS_FULL ? nullptr : c->sysmap;
  if (c->nloc == 10)
    k_adjoint_k<10><<<blocks_for(c->nT, 128), 128, 0, c->stream>>>(c->nT, c->tets, c->xyz, c->dofs,
                                                                  map, u, lam, out);
  else
    k_adjoint_k<4><<<blocks_for(c->nT, 256), 256, 0, c->stream>>>(c->nT, c->tets, c->xyz, c->dofs,
                                                                 map, u, lam, out);
  MG_HIP(hipGetLastError());
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

int mg_assemble(mg_ctx *c, int sys, const mg_operator *op, int prec) {
  API_BEGIN(c)
  MG_CHECK(op, MG_ERR_ARG, "assemble: null operator");
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "assemble: bad system");
  MG_CHECK(prec == MG_PREC_F64 || prec == MG_PREC_F32, MG_ERR_ARG, "assemble: bad precision");
  System &S = c->sys[sys];
  MG_CHECK(S.built, MG_ERR_STATE, "assemble: pattern not built");
  MG_CHECK(op->terms & (MG_TERM_STIFF | MG_TERM_MASS), MG_ERR_ARG, "assemble: no term");
  if (op->terms & MG_TERM_STIFF) {
    MG_CHECK(rule_ok(op->stiff_rule), MG_ERR_ARG, "assemble: stiffness rule must be 1,4,11,15");
    MG_CHECK(op->stiff_coeff == MG_COEFF_CONST || c->field[MG_FIELD_STIFF], MG_ERR_STATE,
             "assemble: stiffness P0 field not set");
  }
  if (op->terms & MG_TERM_MASS) {
    MG_CHECK(rule_ok(op->mass_rule), MG_ERR_ARG, "assemble: mass rule must be 1,4,11,15");
    MG_CHECK(op->mass_coeff == MG_COEFF_CONST || c->field[MG_FIELD_MASS], MG_ERR_STATE,
             "assemble: mass P0 field not set");
  }
  if (prec == MG_PREC_F64) {
    assemble_t<double>(c, sys, op, prec);
    S.has64 = true;
    S.dinv64_ok = false;
  } else {
    assemble_t<float>(c, sys, op, prec);
    S.has32 = true;
    S.dinv32_ok = false;
  }
  API_END(c)
}

int mg_adjoint_grad_k_device(mg_ctx *c, int sys, int rule, const double *u, const double *lam,
                             double *dJdk) {
  API_BEGIN(c)
  MG_CHECK(u && lam && dJdk, MG_ERR_ARG, "adjoint: null pointer");
  MG_CHECK(rule_ok(rule), MG_ERR_ARG, "adjoint: rule must be 1,4,11,15");
  MG_CHECK(c->dofs, MG_ERR_STATE, "adjoint: dof map not built");
  MG_CHECK(sys == MG_SYS_FULL || c->dirichlet_ok, MG_ERR_STATE, "adjoint: no interior numbering");
  adjoint_launch(c, sys, rule, u, lam, dJdk);
  API_END(c)
}

int mg_adjoint_grad_k(mg_ctx *c, int sys, int rule, const double *u, const double *lam,
                      double *dJdk) {
  API_BEGIN(c)
  MG_CHECK(u && lam && dJdk, MG_ERR_ARG, "adjoint: null pointer");
  MG_CHECK(rule_ok(rule), MG_ERR_ARG, "adjoint: rule must be 1,4,11,15");
  MG_CHECK(c->dofs, MG_ERR_STATE, "adjoint: dof map not built");
  MG_CHECK(sys == MG_SYS_FULL || c->dirichlet_ok, MG_ERR_STATE, "adjoint: no interior numbering");
  const int64_t n = sys == MG_SYS_FULL ? c->ndofs : c->nfree;
  double *du = dalloc<double>(c, n), *dl = dalloc<double>(c, n), *dg = dalloc<double>(c, c->nT);
  MG_HIP(hipMemcpyAsync(du, u, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  MG_HIP(hipMemcpyAsync(dl, lam, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  adjoint_launch(c, sys, rule, du, dl, dg);
  MG_HIP(hipMemcpyAsync(dJdk, dg, sizeof(double) * c->nT, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  dfree(c, du);
  dfree(c, dl);
  dfree(c, dg);
  API_END(c)
}

}  // extern "C"
