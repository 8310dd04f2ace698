// Context, mesh upload/validation, P0 fields and the host-side reference
// tables (quadrature + reference basis) of libmgfe.so.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>

#include "mg_internal.h"

namespace mg {

void dfree(Ctx *c, void *p) {
  if (!p) return;
  for (auto &q : c->owned)
    if (q == p) {
      (void)hipFree(p);
      q = nullptr;
      return;
    }
}

void free_system(Ctx *c, System &S) {
  for (void *p : {(void *)S.rowptr, (void *)S.colidx, (void *)S.diagpos,
                  (void *)S.inc_ptr, S.inc_rec, (void *)S.vals64, (void *)S.vals32,
                  (void *)S.dinv64, (void *)S.dinv32, (void *)S.sdesc, (void *)S.cofs,
                  (void *)S.long_rows, (void *)S.patch_ptr, (void *)S.patch_vid,
                  (void *)S.row_lv, (void *)S.inc_lv, S.ell_rec, (void *)S.vid_pad,
                  (void *)S.ell_lv, (void *)S.ell_ebase, (void *)S.prow})
    dfree(c, p);
  S = System();
}

void *scratch(Ctx *c, size_t bytes) {
  if (bytes > c->tmp_bytes) {
    dfree(c, c->tmp);
    c->tmp = dalloc<char>(c, bytes);
    c->tmp_bytes = bytes;
  }
  return c->tmp;
}

// ---- quadrature on the unit tetrahedron, weights summing to 1/6 ------------
// 1, 4 and 15 points: intStaticCard<linearTetra,STANDARD,1|2|3>
// (intStaticCards.cpp:705-829, constants as printed there).  11 points: the
// degree-4 Keast rule (centroid weight -74/5625).
int quad_rule(int npts, double *pts, double *w) {
  int q = 0;
  auto put = [&](double x, double y, double z, double ww) {
    pts[3 * q] = x;
    pts[3 * q + 1] = y;
    pts[3 * q + 2] = z;
    w[q++] = ww;
  };
  switch (npts) {
    case 1:
      put(0.25, 0.25, 0.25, 1.0 / 6.0);
      break;
    case 4: {
      const double a = (5.0 - std::sqrt(5.0)) / 20.0, b = (5.0 + 3 * std::sqrt(5.0)) / 20.0;
      put(a, a, a, 1.0 / 24.0);
      put(a, a, b, 1.0 / 24.0);
      put(a, b, a, 1.0 / 24.0);
      put(b, a, a, 1.0 / 24.0);
    } break;
    case 11: {
      const double w0 = -74.0 / 5625.0, w1 = 343.0 / 45000.0, w2 = 56.0 / 2250.0;
      const double c1 = 1.0 / 14.0, d1 = 11.0 / 14.0;
      const double r = std::sqrt(5.0 / 14.0);
      const double a = (1.0 + r) / 4.0, b = (1.0 - r) / 4.0;
      put(0.25, 0.25, 0.25, w0);
      put(c1, c1, c1, w1);
      put(d1, c1, c1, w1);
      put(c1, d1, c1, w1);
      put(c1, c1, d1, w1);
      put(a, b, b, w2);
      put(b, a, b, w2);
      put(b, b, a, w2);
      put(a, a, b, w2);
      put(a, b, a, w2);
      put(b, a, a, w2);
    } break;
    case 15: {
      const double r5 = 0.25;
      const double s5[2] = {0.09197107805272303, 0.3197936278296299};
      const double t5[2] = {0.7240867658418310, 0.04061911651111023};
      const double u5 = 0.05635083268962915, v5 = 0.4436491673103708;
      const double A5 = 0.01975308641975309;
      const double B5[2] = {0.01198951396316977, 0.01151136787104540};
      const double C5 = 0.008818342151675485;
      put(r5, r5, r5, A5);
      for (int k = 0; k < 2; ++k) {
        put(s5[k], s5[k], s5[k], B5[k]);
        put(t5[k], s5[k], s5[k], B5[k]);
        put(s5[k], t5[k], s5[k], B5[k]);
        put(s5[k], s5[k], t5[k], B5[k]);
      }
      put(u5, u5, v5, C5);
      put(u5, v5, u5, C5);
      put(v5, u5, u5, C5);
      put(v5, v5, u5, C5);
      put(v5, u5, v5, C5);
      put(u5, v5, v5, C5);
    } break;
    default:
      return 0;
  }
  return q;
}

// Reference basis written with barycentric coordinates lambda = (1-x-y-z, x, y, z):
// P1: lambda_i.  P2 (order of fePr3d<2>, fePr3d.hpp:553-592): vertices
// lambda_i (2 lambda_i - 1), then edges 12, 23, 31, 14, 24, 34 (geoShapes.cpp
// edgeToPoint) as 4 lambda_a lambda_b.
const int kEdgeV[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

void relabel_tables(int perm[10][4], int sigma[10][10]) {
  int p[4] = {0, 1, 2, 3};
  for (int l = 0; l < 10; ++l) {
    // first even permutation (lexicographic order) with the required property
    int q[4] = {0, 1, 2, 3};
    bool found = false;
    do {
      int inv = 0;
      for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) inv += q[i] > q[j];
      if (inv & 1) continue;
      const bool ok = l < 4 ? q[0] == l
                            : ((q[0] == kEdgeV[l - 4][0] && q[1] == kEdgeV[l - 4][1]) ||
                               (q[0] == kEdgeV[l - 4][1] && q[1] == kEdgeV[l - 4][0]));
      if (ok) {
        for (int i = 0; i < 4; ++i) p[i] = q[i];
        found = true;
        break;
      }
    } while (std::next_permutation(q, q + 4));
    (void)found;
    for (int i = 0; i < 4; ++i) perm[l][i] = p[i];
    for (int m = 0; m < 4; ++m) sigma[l][m] = p[m];
    for (int j = 0; j < 6; ++j) {
      const int a = p[kEdgeV[j][0]], b = p[kEdgeV[j][1]];
      int old = -1;
      for (int k = 0; k < 6; ++k)
        if ((kEdgeV[k][0] == a && kEdgeV[k][1] == b) || (kEdgeV[k][0] == b && kEdgeV[k][1] == a))
          old = k;
      sigma[l][4 + j] = 4 + old;
    }
  }
}

void ref_basis(int order, const double *y, double *val, double *grad) {
  const double L[4] = {1.0 - y[0] - y[1] - y[2], y[0], y[1], y[2]};
  const double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const int n = order == 1 ? 4 : 10;
  if (order == 1) {
    for (int i = 0; i < 4; ++i) {
      if (val) val[i] = L[i];
      if (grad)
        for (int d = 0; d < 3; ++d) grad[d * n + i] = dL[i][d];
    }
    return;
  }
  for (int i = 0; i < 4; ++i) {
    if (val) val[i] = L[i] * (2.0 * L[i] - 1.0);
    if (grad)
      for (int d = 0; d < 3; ++d) grad[d * n + i] = (4.0 * L[i] - 1.0) * dL[i][d];
  }
  for (int j = 0; j < 6; ++j) {
    const int a = kEdgeV[j][0], b = kEdgeV[j][1];
    if (val) val[4 + j] = 4.0 * L[a] * L[b];
    if (grad)
      for (int d = 0; d < 3; ++d) grad[d * n + 4 + j] = 4.0 * (L[a] * dL[b][d] + L[b] * dL[a][d]);
  }
}

// Element tensors on the reference tetrahedron.  For an affine tet with
// J = [P2-P1 P3-P1 P4-P1] and A = J^{-1} J^{-T}, the quadrature of
// integratorOEL3d_STD (sum_q w_q det c grad(phi_i).grad(phi_j)) equals
//   sum_k (c det A_k) tab[k][i][j],
//   tab[k][i][j] = sum_q w_q g_i^a(q) g_j^b(q) (+ the (b,a) term when a != b),
// k over (a,b) = (0,0),(1,1),(2,2),(0,1),(0,2),(1,2); and the mass quadrature
// equals (c det) tab[6][i][j] with tab[6][i][j] = sum_q w_q phi_i(q) phi_j(q).
void build_tables(int order, int stiff_rule, int mass_rule, double *tab) {
  const int n = order == 1 ? 4 : 10;
  const int KA[6] = {0, 1, 2, 0, 0, 1}, KB[6] = {0, 1, 2, 1, 2, 2};
  std::memset(tab, 0, sizeof(double) * 7 * n * n);
  double pts[45], w[15], g[30], v[10];
  int nq = stiff_rule ? quad_rule(stiff_rule, pts, w) : 0;
  for (int q = 0; q < nq; ++q) {
    ref_basis(order, &pts[3 * q], nullptr, g);
    for (int k = 0; k < 6; ++k)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          double t = g[KA[k] * n + i] * g[KB[k] * n + j];
          if (KA[k] != KB[k]) t += g[KB[k] * n + i] * g[KA[k] * n + j];
          tab[(k * n + i) * n + j] += w[q] * t;
        }
  }
  nq = mass_rule ? quad_rule(mass_rule, pts, w) : 0;
  for (int q = 0; q < nq; ++q) {
    ref_basis(order, &pts[3 * q], v, nullptr);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) tab[(6 * n + i) * n + j] += w[q] * v[i] * v[j];
  }
}

// ---- mesh validation --------------------------------------------------------
// Converts the 1-based ids to 0-based and counts tets with an out-of-range id or
// det(J) <= 0 (meshDoctor3d::checkJacobian, meshDoctor3d.hpp:195-227).
__global__ void k_mesh_check(int64_t nT, int64_t nV, const int32_t *__restrict__ tets1,
                             const double *__restrict__ xyz, int32_t *__restrict__ tets0,
                             int *__restrict__ bad) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nT) return;
  int id[4];
  bool ok = true;
  for (int v = 0; v < 4; ++v) {
    id[v] = tets1[4 * e + v] - 1;
    ok &= (id[v] >= 0 && id[v] < nV);
    tets0[4 * e + v] = id[v];
  }
  if (ok) {
    double J[3][3];
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) J[r][c] = xyz[3 * id[c + 1] + r] - xyz[3 * id[0] + r];
    double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                 J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    ok = det > 0.0;
  }
  if (!ok) atomicAdd(bad, 1);
}

__global__ void k_btri_check(int64_t nB, int64_t nV, const int32_t *__restrict__ b1,
                             int32_t *__restrict__ b0, int *__restrict__ bad) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nB) return;
  int v = b1[t] - 1;
  b0[t] = v;
  if (v < 0 || v >= nV) atomicAdd(bad, 1);
}

// grid-stride 16-byte copy (one read, one write per byte)
__global__ __launch_bounds__(256) void k_copy16(int64_t n, const int4 *__restrict__ src,
                                                int4 *__restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

static void reset_derived(Ctx *c) {
  for (int s = 0; s < 2; ++s) {
    free_system(c, c->sys[s]);
  }
  for (void *p : {(void *)c->edge_keys, (void *)c->edge_nodes, (void *)c->elem_edge,
                  (void *)c->dofs, (void *)c->fixed, (void *)c->sysmap, (void *)c->edge_surf,
                  (void *)c->edge_n, (void *)c->cell_volume, (void *)c->sigma_el,
                  (void *)c->sigma_len, (void *)c->vinc_ptr, (void *)c->vinc,
                  (void *)c->vinc_pos, (void *)c->dd_contrib})
    dfree(c, p);
  c->edge_keys = nullptr;
  c->edge_nodes = c->elem_edge = c->dofs = c->sysmap = nullptr;
  c->fixed = nullptr;
  c->edge_surf = c->edge_n = c->cell_volume = c->sigma_el = c->sigma_len = nullptr;
  c->vinc_ptr = c->vinc = nullptr;
  c->dd_contrib = nullptr;
  c->vinc_pos = nullptr;
  c->edges_ok = c->dirichlet_ok = c->dual_ok = false;
  c->order = c->nloc = 0;
  c->nE = c->ndofs = c->nfree = 0;
}

}  // namespace mg

using namespace mg;

#define API_BEGIN(c)             \
  if (!(c)) return MG_ERR_ARG;   \
  try {                          \
    MG_HIP(hipSetDevice((c)->device));
#define API_END(c)                   \
  }                                  \
  catch (const mg::Error &e) {       \
    (c)->err = e.msg;                \
    return e.code;                   \
  }                                  \
  catch (const std::exception &e) {  \
    (c)->err = e.what();             \
    return MG_ERR_HIP;               \
  }                                  \
  return MG_OK;

extern "C" {

int mg_ctx_create(int device, mg_ctx **out) {
  if (!out) return MG_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MG_ERR_HIP;
  mg_ctx *c = new (std::nothrow) mg_ctx();
  if (!c) return MG_ERR_HIP;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return MG_ERR_HIP;
  }
  try {
    noalloc_init(c);
  } catch (const mg::Error &) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MG_ERR_HIP;
  }
  *out = c;
  return MG_OK;
}

int mg_ctx_destroy(mg_ctx *c) {
  if (!c) return MG_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (void *p : c->owned)
    if (p) (void)hipFree(p);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return MG_OK;
}

const char *mg_last_error(const mg_ctx *c) { return c ? c->err.c_str() : "null context"; }

void *mg_ctx_stream(mg_ctx *c) { return c ? (void *)c->stream : nullptr; }

int mg_ctx_synchronize(mg_ctx *c) {
  API_BEGIN(c)
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_copy_device(mg_ctx *c, void *dst, const void *src, int64_t nbytes) {
  API_BEGIN(c)
  MG_CHECK(dst && src && nbytes >= 0 && nbytes % 16 == 0 && (uintptr_t)dst % 16 == 0 &&
               (uintptr_t)src % 16 == 0,
           MG_ERR_ARG, "copy: 16-byte aligned pointers and a multiple of 16 bytes required");
  const int64_t n = nbytes / 16;
  if (n > 0) {
    const int blocks = (int)std::min<int64_t>(blocks_for(n, 256), 256 * 32);
    k_copy16<<<blocks, 256, 0, c->stream>>>(n, (const int4 *)src, (int4 *)dst);
    MG_HIP(hipGetLastError());
  }
  API_END(c)
}

int mg_mesh_set(mg_ctx *c, int64_t nV, const double *xyz, int64_t nT, const int32_t *tets,
                int64_t nB, const int32_t *btri, const int32_t *bgeo) {
  API_BEGIN(c)
  MG_CHECK(nV > 0 && nT > 0 && xyz && tets, MG_ERR_ARG, "mesh: empty or null arrays");
  MG_CHECK(nB >= 0 && (nB == 0 || (btri && bgeo)), MG_ERR_ARG, "mesh: boundary arrays");
  MG_CHECK(nV < (1ll << 31) && nT < (1ll << 27), MG_ERR_SIZE,
           "mesh: nV must be < 2^31 and nT < 2^27 (packed element ids)");
  reset_derived(c);
  for (void *p : {(void *)c->xyz, (void *)c->tets, (void *)c->btri, (void *)c->bgeo,
                  (void *)c->field[0], (void *)c->field[1]})
    dfree(c, p);
  c->field[0] = c->field[1] = nullptr;
  c->nV = nV;
  c->nT = nT;
  c->nB = nB;
  c->xyz = dalloc<double>(c, 3 * nV);
  c->tets = dalloc<int32_t>(c, 4 * nT);
  c->btri = dalloc<int32_t>(c, 3 * nB);
  c->bgeo = dalloc<int32_t>(c, nB);
  int32_t *t1 = (int32_t *)scratch(c, sizeof(int32_t) * (4 * nT + 3 * nB) + 64);
  int32_t *b1 = t1 + 4 * nT;
  int *bad = dalloc<int>(c, 2);
  MG_HIP(hipMemsetAsync(bad, 0, 2 * sizeof(int), c->stream));
  MG_HIP(hipMemcpyAsync(c->xyz, xyz, sizeof(double) * 3 * nV, hipMemcpyHostToDevice, c->stream));
  MG_HIP(hipMemcpyAsync(t1, tets, sizeof(int32_t) * 4 * nT, hipMemcpyHostToDevice, c->stream));
  if (nB) {
    MG_HIP(hipMemcpyAsync(b1, btri, sizeof(int32_t) * 3 * nB, hipMemcpyHostToDevice, c->stream));
    MG_HIP(hipMemcpyAsync(c->bgeo, bgeo, sizeof(int32_t) * nB, hipMemcpyHostToDevice, c->stream));
  }
  k_mesh_check<<<blocks_for(nT, 256), 256, 0, c->stream>>>(nT, nV, t1, c->xyz, c->tets, bad);
  if (nB)
    k_btri_check<<<blocks_for(3 * nB, 256), 256, 0, c->stream>>>(nB, nV, b1, c->btri, bad + 1);
  MG_HIP(hipGetLastError());
  int hb[2];
  MG_HIP(hipMemcpyAsync(hb, bad, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  dfree(c, bad);
  if (hb[0] || hb[1]) {
    c->nT = 0;
    throw Error{MG_ERR_MESH, "mesh: " + std::to_string(hb[0]) +
                                 " tetrahedra with det(J) <= 0 or a bad vertex id, " +
                                 std::to_string(hb[1]) + " bad boundary ids"};
  }
  API_END(c)
}

int mg_field_set(mg_ctx *c, int which, const double *values) {
  API_BEGIN(c)
  MG_CHECK(which == MG_FIELD_STIFF || which == MG_FIELD_MASS, MG_ERR_ARG, "field: bad slot");
  MG_CHECK(values, MG_ERR_ARG, "field: null values");
  MG_CHECK(c->nT > 0, MG_ERR_STATE, "field: no mesh");
  if (!c->field[which]) c->field[which] = dalloc<double>(c, c->nT);
  MG_HIP(hipMemcpyAsync(c->field[which], values, sizeof(double) * c->nT, hipMemcpyHostToDevice,
                        c->stream));
  ++c->field_ver[which];
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_field_set_device(mg_ctx *c, int which, const double *values) {
  API_BEGIN(c)
  MG_CHECK(which == MG_FIELD_STIFF || which == MG_FIELD_MASS, MG_ERR_ARG, "field: bad slot");
  MG_CHECK(values, MG_ERR_ARG, "field: null values");
  MG_CHECK(c->nT > 0, MG_ERR_STATE, "field: no mesh");
  if (!c->field[which]) c->field[which] = dalloc<double>(c, c->nT);
  MG_HIP(hipMemcpyAsync(c->field[which], values, sizeof(double) * c->nT,
                        hipMemcpyDeviceToDevice, c->stream));
  ++c->field_ver[which];
  API_END(c)
}

int mg_dofmap_build(mg_ctx *c, int order, int64_t *n_dofs, int64_t *n_edges) {
  API_BEGIN(c)
  MG_CHECK(order == 1 || order == 2, MG_ERR_ARG, "dofmap: order must be 1 or 2");
  MG_CHECK(c->nT > 0, MG_ERR_STATE, "dofmap: no mesh");
  for (int s = 0; s < 2; ++s) {
    free_system(c, c->sys[s]);
  }
  dfree(c, c->fixed);
  dfree(c, c->sysmap);
  c->fixed = nullptr;
  c->sysmap = nullptr;
  c->dirichlet_ok = false;
  c->order = order;
  c->nloc = order == 1 ? 4 : 10;
  build_edges(c);
  build_dofs(c);
  if (n_dofs) *n_dofs = c->ndofs;
  if (n_edges) *n_edges = c->nE;
  API_END(c)
}

int mg_elem_dofs_get(mg_ctx *c, int32_t *dofs) {
  API_BEGIN(c)
  MG_CHECK(dofs, MG_ERR_ARG, "null output");
  MG_CHECK(c->dofs, MG_ERR_STATE, "dof map not built");
  MG_HIP(hipMemcpyAsync(dofs, c->dofs, sizeof(int32_t) * c->nloc * c->nT, hipMemcpyDeviceToHost,
                        c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_edges_get(mg_ctx *c, int32_t *edge_nodes, int32_t *elem_edge) {
  API_BEGIN(c)
  MG_CHECK(c->edges_ok, MG_ERR_STATE, "edges not built");
  if (edge_nodes)
    MG_HIP(hipMemcpyAsync(edge_nodes, c->edge_nodes, sizeof(int32_t) * 2 * c->nE,
                          hipMemcpyDeviceToHost, c->stream));
  if (elem_edge)
    MG_HIP(hipMemcpyAsync(elem_edge, c->elem_edge, sizeof(int32_t) * 6 * c->nT,
                          hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_dirichlet_set(mg_ctx *c, int n_geo, const int32_t *geo_ids, int64_t *n_free) {
  API_BEGIN(c)
  MG_CHECK(n_geo >= 0 && n_geo <= 64 && (n_geo == 0 || geo_ids), MG_ERR_ARG,
           "dirichlet: 0..64 geoIds");
  MG_CHECK(c->dofs, MG_ERR_STATE, "dirichlet: dof map not built");
  build_dirichlet(c, n_geo, geo_ids);
  if (n_free) *n_free = c->nfree;
  API_END(c)
}

int mg_dirichlet_mask_get(mg_ctx *c, uint8_t *fixed) {
  API_BEGIN(c)
  MG_CHECK(fixed, MG_ERR_ARG, "null output");
  MG_CHECK(c->dirichlet_ok, MG_ERR_STATE, "dirichlet not set");
  MG_HIP(hipMemcpyAsync(fixed, c->fixed, c->ndofs, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_pattern_build(mg_ctx *c, int sys, int64_t *n_rows, int64_t *nnz) {
  API_BEGIN(c)
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "pattern: bad system");
  MG_CHECK(c->dofs, MG_ERR_STATE, "pattern: dof map not built");
  MG_CHECK(sys == MG_SYS_FULL || c->dirichlet_ok, MG_ERR_STATE,
           "pattern: interior system needs mg_dirichlet_set");
  // a pattern stays valid until the mesh, the dof map or the Dirichlet set
  // change (those calls release it); rebuilding it would drop its values
  if (!c->sys[sys].built) build_pattern(c, sys);
  if (n_rows) *n_rows = c->sys[sys].nrows;
  if (nnz) *nnz = c->sys[sys].nnz;
  API_END(c)
}

int mg_pattern_get(mg_ctx *c, int sys, int32_t *rowptr, int32_t *colidx) {
  API_BEGIN(c)
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "bad system");
  System &S = c->sys[sys];
  MG_CHECK(S.built, MG_ERR_STATE, "pattern not built");
  if (rowptr)
    MG_HIP(hipMemcpyAsync(rowptr, S.rowptr, sizeof(int32_t) * (S.nrows + 1), hipMemcpyDeviceToHost,
                          c->stream));
  if (colidx)
    MG_HIP(hipMemcpyAsync(colidx, S.colidx, sizeof(int32_t) * S.nnz, hipMemcpyDeviceToHost,
                          c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_slot_map_get(mg_ctx *c, int sys, int32_t *slot) {
  API_BEGIN(c)
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "bad system");
  System &S = c->sys[sys];
  MG_CHECK(S.built, MG_ERR_STATE, "pattern not built");
  MG_CHECK(slot, MG_ERR_ARG, "null output");
  const size_t ncon = (size_t)c->nT * c->nloc * c->nloc;
  DevTmp<int32_t> ds(c, ncon);
  slot_map(c, sys, ds);
  MG_HIP(hipMemcpyAsync(slot, ds, sizeof(int32_t) * ncon, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_values_get(mg_ctx *c, int sys, int prec, void *vals) {
  API_BEGIN(c)
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "bad system");
  MG_CHECK(vals, MG_ERR_ARG, "null output");
  System &S = c->sys[sys];
  if (prec == MG_PREC_F64) {
    MG_CHECK(S.has64, MG_ERR_STATE, "no FP64 values assembled");
    MG_HIP(hipMemcpyAsync(vals, S.vals64, sizeof(double) * S.nnz, hipMemcpyDeviceToHost,
                          c->stream));
  } else {
    MG_CHECK(S.has32, MG_ERR_STATE, "no FP32 values assembled");
    MG_HIP(hipMemcpyAsync(vals, S.vals32, sizeof(float) * S.nnz, hipMemcpyDeviceToHost,
                          c->stream));
  }
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_assembly_seq(mg_ctx *c, int sys, int prec, uint64_t *seq) {
  API_BEGIN(c)
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "bad system");
  MG_CHECK(prec == MG_PREC_F64 || prec == MG_PREC_F32, MG_ERR_ARG, "bad precision");
  MG_CHECK(seq, MG_ERR_ARG, "null output");
  const System &S = c->sys[sys];
  *seq = prec == MG_PREC_F64 ? (S.has64 ? S.seq64 : 0) : (S.has32 ? S.seq32 : 0);
  API_END(c)
}

int mg_values_device(mg_ctx *c, int sys, int prec, void **vals) {
  API_BEGIN(c)
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "bad system");
  MG_CHECK(vals, MG_ERR_ARG, "null output");
  System &S = c->sys[sys];
  MG_CHECK(prec == MG_PREC_F64 ? S.has64 : S.has32, MG_ERR_STATE, "values not assembled");
  *vals = prec == MG_PREC_F64 ? (void *)S.vals64 : (void *)S.vals32;
  API_END(c)
}

}  // extern "C"
