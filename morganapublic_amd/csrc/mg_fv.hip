// Barycentric dual-mesh data and the Scharfetter-Gummel drift-diffusion
// residual with its FP64 vector-Jacobian products.
//
// Dual data (dual3d<linearTetra>::buildFiniteVolumeData, dual3d.hpp:606-1025):
// per tet, the 12 interface triangles (edge midpoint, face barycentre, element
// barycentre) have normals N = (P2-P1) x (P3-P1); they are summed per edge with
// their sign aligned to the edge direction, edgeSurf = |edgeN| / 2, and each
// vertex receives the 6 sub-cell volumes of every tet around it.  Sums over
// elements are gathers over sorted incidence lists (deterministic).
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "mg_internal.h"

namespace mg {

__constant__ int c_ev[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
// the same edge table for compile-time indexing (local edge j = (kEA[j], kEB[j]))
constexpr int kEA[6] = {0, 1, 2, 0, 1, 2}, kEB[6] = {1, 2, 0, 3, 3, 3};

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// per element: aligned interface normal sum for each local edge, and the
// sub-cell volume each local vertex receives
__global__ __launch_bounds__(128) void k_dual_elem(int64_t nT, const int32_t *__restrict__ tets,
                            const double *__restrict__ xyz, double *__restrict__ ne,
                            double *__restrict__ subvol) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nT) return;
  int32_t tv[4];
  double P[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    tv[v] = tets[4 * e + v];
#pragma unroll
    for (int d = 0; d < 3; ++d) P[v][d] = xyz[3 * tv[v] + d];
  }
  double BA[3], sv[4] = {0, 0, 0, 0};
#pragma unroll
  for (int d = 0; d < 3; ++d) BA[d] = (P[0][d] + P[1][d] + P[2][d] + P[3][d]) / 4.0;
  // edge by edge, with its two faces (the faces of FV holding both ends);
  // face centroids and edge midpoints are formed where used, so only the
  // coordinates, the barycentre and the sub-volumes stay live (no scratch)
  constexpr int FV[4][3] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {0, 3, 2}};
  constexpr int EF[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 3}, {1, 2}, {2, 3}};
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double Em[3], N[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 3; ++d) Em[d] = (P[kEA[j]][d] + P[kEB[j]][d]) / 2.0;
    // edge vector from the lower to the higher global vertex (a select, not a
    // swap of indices: P stays indexed by constants)
    const bool fwd = tv[kEA[j]] <= tv[kEB[j]];
    double t[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double u = P[kEB[j]][d] - P[kEA[j]][d];
      t[d] = fwd ? u : -u;
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int f = EF[j][q];
      double F[3], a[3], b[3], n[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        F[d] = (P[FV[f][0]][d] + P[FV[f][1]][d] + P[FV[f][2]][d]) / 3.0;
        a[d] = F[d] - Em[d];
        b[d] = BA[d] - Em[d];
      }
      cross3(a, b, n);
      const double sgn = (n[0] * t[0] + n[1] * t[1] + n[2] * t[2]) >= 0.0 ? 1.0 : -1.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) N[d] += sgn * n[d];
      // the two sub-volumes (edge midpoint, face centroid, barycentre, end vertex)
      double m[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        a[d] = Em[d] - BA[d];
        b[d] = F[d] - BA[d];
      }
      cross3(a, b, m);
#pragma unroll
      for (int end = 0; end < 2; ++end) {
        const int v = end == 0 ? kEA[j] : kEB[j];
        sv[v] += fabs(m[0] * (P[v][0] - BA[0]) + m[1] * (P[v][1] - BA[1]) +
                      m[2] * (P[v][2] - BA[2])) / 6.0;
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) ne[18 * e + 3 * j + d] = N[d];
  }
  for (int v = 0; v < 4; ++v) subvol[4 * e + v] = sv[v];
}

__global__ void k_edge_gather(int64_t nE, const int32_t *__restrict__ ptr,
                              const int32_t *__restrict__ inc, const double *__restrict__ ne,
                              double *__restrict__ surf, double *__restrict__ nrm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nE) return;
  double s[3] = {0, 0, 0};
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const int32_t ej = inc[k];
    for (int d = 0; d < 3; ++d) s[d] += ne[3 * (int64_t)ej + d];
  }
  const double nn = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  surf[i] = 0.5 * nn;
  for (int d = 0; d < 3; ++d) nrm[3 * i + d] = s[d] / nn;
}

__global__ void k_sigma(int64_t nT, const int32_t *__restrict__ elem_edge,
                        const double *__restrict__ ne, const double *__restrict__ nrm,
                        double *__restrict__ sigma) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * nT) return;
  const double *nh = nrm + 3 * (int64_t)elem_edge[i];
  sigma[i] = 0.5 * (ne[3 * i] * nh[0] + ne[3 * i + 1] * nh[1] + ne[3 * i + 2] * nh[2]);
}

// sigma_len[6e + j] = sigma_el[6e + j] / |edge j of e|: the only form in
// which the SG fluxes use the dual areas, so the residual kernels need no
// edge lengths (no square roots or divisions per edge)
__global__ void k_sigma_len(int64_t nT, const int32_t *__restrict__ tets,
                            const double *__restrict__ xyz, const double *__restrict__ sigma,
                            double *__restrict__ sigma_len) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nT) return;
  const int4 t = reinterpret_cast<const int4 *>(tets)[e];
  const int32_t id[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double s2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double tt = xyz[3 * (size_t)id[kEA[j]] + d] - xyz[3 * (size_t)id[kEB[j]] + d];
      s2 += tt * tt;
    }
    sigma_len[6 * e + j] = sigma[6 * e + j] / sqrt(s2);
  }
}

__global__ void k_vertex_gather(int64_t nV, const int32_t *__restrict__ ptr,
                                const int32_t *__restrict__ inc, const double *__restrict__ subvol,
                                double *__restrict__ cvol) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  double s = 0.0;
  for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
    const int32_t el = inc[k];  // e << 2 | l
    s += subvol[4 * (int64_t)(el >> 2) + (el & 3)];
  }
  cvol[v] = s;
}

__global__ void k_keys_i32(int64_t n, const int32_t *__restrict__ key_src, int shift,
                           int32_t *__restrict__ keys, int32_t *__restrict__ vals, int mode) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = key_src[i];
  // mode 0: value = i; mode 1: value = (i / 4) << 2 | (i % 4)
  vals[i] = mode == 0 ? (int32_t)i : (int32_t)(((i >> 2) << shift) | (i & 3));
}

__global__ void k_count_i32(int64_t n, const int32_t *__restrict__ keys, int32_t *__restrict__ cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[keys[i]], 1);
}

// CSR of the items grouped by integer key (stable: original order inside a key)
static void group_by(Ctx *c, const int32_t *keys_src, int64_t n, int64_t nkeys, int mode,
                     int32_t **ptr_out, int32_t **inc_out) {
  int32_t *k0 = dalloc<int32_t>(c, n), *k1 = dalloc<int32_t>(c, n);
  int32_t *v0 = dalloc<int32_t>(c, n), *v1 = dalloc<int32_t>(c, n);
  k_keys_i32<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, keys_src, 2, k0, v0, mode);
  int bits = 1;
  while ((1ll << bits) <= nkeys) ++bits;
  size_t tb = 0;
  MG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, (int)n, 0, bits,
                                            c->stream));
  void *t = scratch(c, tb);
  MG_HIP(hipcub::DeviceRadixSort::SortPairs(t, tb, k0, k1, v0, v1, (int)n, 0, bits, c->stream));
  int32_t *cnt = dalloc<int32_t>(c, nkeys + 1);
  MG_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nkeys + 1), c->stream));
  k_count_i32<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, k1, cnt);
  int32_t *ptr = dalloc<int32_t>(c, nkeys + 1);
  tb = 0;
  MG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, ptr, (int)(nkeys + 1), c->stream));
  t = scratch(c, tb);
  MG_HIP(hipcub::DeviceScan::ExclusiveSum(t, tb, cnt, ptr, (int)(nkeys + 1), c->stream));
  MG_HIP(hipGetLastError());
  MG_HIP(hipStreamSynchronize(c->stream));
  dfree(c, k0);
  dfree(c, k1);
  dfree(c, v0);
  dfree(c, cnt);
  *ptr_out = ptr;
  *inc_out = v1;
}

__global__ void k_inverse(int64_t n, const int32_t *__restrict__ perm, int32_t *__restrict__ inv) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) inv[perm[k]] = (int32_t)k;
}

static void build_vertex_incidence(Ctx *c) {
  if (c->vinc_ptr) return;
  group_by(c, c->tets, 4 * c->nT, c->nV, 1, &c->vinc_ptr, &c->vinc);
  c->vinc_pos = dalloc<int32_t>(c, 4 * c->nT);
  k_inverse<<<blocks_for(4 * c->nT, 256), 256, 0, c->stream>>>(4 * c->nT, c->vinc, c->vinc_pos);
  MG_HIP(hipGetLastError());
}

void build_dual(Ctx *c) {
  MG_CHECK(c->edges_ok, MG_ERR_STATE, "dual: dof map (edges) not built");
  const int64_t nT = c->nT, nE = c->nE, nV = c->nV;
  double *ne = dalloc<double>(c, 18 * nT), *sv = dalloc<double>(c, 4 * nT);
  k_dual_elem<<<blocks_for(nT, 128), 128, 0, c->stream>>>(nT, c->tets, c->xyz, ne, sv);
  MG_HIP(hipGetLastError());
  int32_t *eptr = nullptr, *einc = nullptr;
  group_by(c, c->elem_edge, 6 * nT, nE, 0, &eptr, &einc);
  build_vertex_incidence(c);
  for (double *p : {c->edge_surf, c->edge_n, c->cell_volume, c->sigma_el, c->sigma_len,
                    c->dd_contrib})
    dfree(c, p);
  c->edge_surf = dalloc<double>(c, nE);
  c->edge_n = dalloc<double>(c, 3 * nE);
  c->cell_volume = dalloc<double>(c, nV);
  c->sigma_el = dalloc<double>(c, 6 * nT);
  c->sigma_len = dalloc<double>(c, 6 * nT);
  c->dd_contrib = dalloc<double>(c, 8 * nT);
  k_edge_gather<<<blocks_for(nE, 256), 256, 0, c->stream>>>(nE, eptr, einc, ne, c->edge_surf,
                                                          c->edge_n);
  k_sigma<<<blocks_for(6 * nT, 256), 256, 0, c->stream>>>(nT, c->elem_edge, ne, c->edge_n,
                                                        c->sigma_el);
  k_sigma_len<<<blocks_for(nT, 256), 256, 0, c->stream>>>(nT, c->tets, c->xyz, c->sigma_el,
                                                        c->sigma_len);
  k_vertex_gather<<<blocks_for(nV, 256), 256, 0, c->stream>>>(nV, c->vinc_ptr, c->vinc, sv,
                                                            c->cell_volume);
  MG_HIP(hipGetLastError());
  MG_HIP(hipStreamSynchronize(c->stream));
  for (void *p : {(void *)ne, (void *)sv, (void *)eptr, (void *)einc}) dfree(c, p);
  c->dual_ok = true;
}

// ---- Scharfetter-Gummel -------------------------------------------------------------
// B(x) = x / (e^x - 1) in FP64; Taylor series 1 - x/2 + x^2/12 - x^4/720 + x^6/30240
// for |x| < 1e-2 (cancellation-free near 0, B(0) = 1).
__device__ __forceinline__ double bernoulli(double x) {
  const double ax = fabs(x);
  if (ax < 1e-2) {
    const double x2 = x * x;
    return 1.0 - 0.5 * x + x2 / 12.0 * (1.0 - x2 / 60.0 * (1.0 - x2 / 42.0));
  }
  if (x > 700.0) return x * exp(-x);
  return x / expm1(x);
}

__global__ void k_bernoulli(int64_t n, const double *__restrict__ x, double *__restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = bernoulli(x[i]);
}

// An element's vertex coordinates, and the signed det of J with the
// barycentric gradients G (row q = grad lambda_q), from the coordinates
// (one division).
__device__ __forceinline__ double p1_geometry(const double P[4][3], double G[4][3]) {
  double T[3][3];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
#pragma unroll
    for (int r = 0; r < 3; ++r) T[r][cc] = P[cc + 1][r] - P[0][r];
  const double d = T[0][0] * T[1][1] * T[2][2] - T[0][0] * T[1][2] * T[2][1] -
                   T[1][0] * T[0][1] * T[2][2] + T[1][0] * T[0][2] * T[2][1] +
                   T[2][0] * T[0][1] * T[1][2] - T[2][0] * T[0][2] * T[1][1];
  const double rd = 1.0 / d;
  G[1][0] = (T[1][1] * T[2][2] - T[1][2] * T[2][1]) * rd;
  G[1][1] = -(T[0][1] * T[2][2] - T[0][2] * T[2][1]) * rd;
  G[1][2] = (T[0][1] * T[1][2] - T[0][2] * T[1][1]) * rd;
  G[2][0] = -(T[1][0] * T[2][2] - T[1][2] * T[2][0]) * rd;
  G[2][1] = (T[0][0] * T[2][2] - T[0][2] * T[2][0]) * rd;
  G[2][2] = -(T[0][0] * T[1][2] - T[0][2] * T[1][0]) * rd;
  G[3][0] = (T[1][0] * T[2][1] - T[1][1] * T[2][0]) * rd;
  G[3][1] = -(T[0][0] * T[2][1] - T[0][1] * T[2][0]) * rd;
  G[3][2] = (T[0][0] * T[1][1] - T[0][1] * T[1][0]) * rd;
#pragma unroll
  for (int k = 0; k < 3; ++k) G[0][k] = -(G[1][k] + G[2][k] + G[3][k]);
  return d;
}

// SG flux of local edge (a, b) (a, b compile-time) with coefficient coef
// (mu sigma / L: the length is in sigma_len), oriented from a to b:
// coef (n_b B(phi_b - phi_a) - n_a B(phi_a - phi_b)).
// One Bernoulli evaluation per edge: with t = |x|, B(-t) = t + B(t) (both
// terms positive, no cancellation).
__device__ __forceinline__ double sg_edge_flux(const double ph[4], const double dn[4], int a,
                                               int b, double coef) {
  const double dab = ph[b] - ph[a];
  const double t = fabs(dab);
  const double bt = bernoulli(t), bmt = t + bt;  // B(t), B(-t)
  const double bp = dab >= 0.0 ? bt : bmt;       // B(dab)
  const double bm = dab >= 0.0 ? bmt : bt;       // B(-dab)
  return coef * (dn[b] * bp - dn[a] * bm);
}

// The residual in two passes.  k_dd_elem: one lane per element computes its
// geometry, grad(phi_h) and the SG fluxes of its 6 edges once, and writes, for
// each local vertex l, the Poisson term eps vol grad(lambda_l).grad(phi_h) -
// rho vol / 4 and the SG term sum over the edges j at l (in j order) of
// +-F_j (F_j oriented from the edge's first to its second local vertex), as
// one double2 at the position of the incidence (e, l) in the vertex
// incidence list (vinc_pos), so the scattered accesses are writes.
// k_dd_gather: one lane per vertex sums its incidences' terms, a contiguous
// range, in incidence order (deterministic, no atomics), and the workgroup
// writes its partial of the loss sum (r_phi^2 + r_n^2).
__global__ __launch_bounds__(128) void k_dd_elem(
    int64_t nT, const int32_t *__restrict__ tets, const double *__restrict__ xyz,
    const double *__restrict__ sigma_len, const double *__restrict__ eps,
    const double *__restrict__ rho, const double *__restrict__ mu, const double *__restrict__ phi,
    const double *__restrict__ dens, const int32_t *__restrict__ pos,
    double2 *__restrict__ contrib) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nT) return;
  const int4 ps = reinterpret_cast<const int4 *>(pos)[e];
  const int4 t = reinterpret_cast<const int4 *>(tets)[e];
  const int32_t id[4] = {t.x, t.y, t.z, t.w};
  // every per-element and per-vertex operand is loaded once, up front
  double P[4][3], ph[4], dn[4], sg[6];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int d = 0; d < 3; ++d) P[q][d] = xyz[3 * (size_t)id[q] + d];
    ph[q] = phi[id[q]];
    dn[q] = dens[id[q]];
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) sg[j] = sigma_len[6 * e + j];
  const double ee = eps[e], re = rho[e], me = mu[e];
  // Poisson terms first, so the geometry is dead during the SG part (100
  // VGPRs, 4 waves per SIMD, against 148 with the geometry kept live)
  double cp[4];
  {
    double G[4][3];
    const double vol = p1_geometry(P, G) / 6.0;
    double gp[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < 3; ++k) gp[k] += ph[q] * G[q][k];
#pragma unroll
    for (int l = 0; l < 4; ++l)
      cp[l] = ee * vol * (G[l][0] * gp[0] + G[l][1] * gp[1] + G[l][2] * gp[2]) - re * vol / 4.0;
  }
  // each local vertex receives +-F_j of its edges in j order
  double cn[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int a = kEA[j], b = kEB[j];
    const double F = sg_edge_flux(ph, dn, a, b, me * sg[j]);
    cn[a] += F;
    cn[b] -= F;
  }
  contrib[ps.x] = make_double2(cp[0], cn[0]);
  contrib[ps.y] = make_double2(cp[1], cn[1]);
  contrib[ps.z] = make_double2(cp[2], cn[2]);
  contrib[ps.w] = make_double2(cp[3], cn[3]);
}

__global__ __launch_bounds__(256) void k_dd_gather(int64_t nV, const int32_t *__restrict__ vptr,
                                                   const double2 *__restrict__ contrib,
                                                   const uint8_t *__restrict__ fixed,
                                                   double *__restrict__ rphi,
                                                   double *__restrict__ rn,
                                                   double *__restrict__ partial) {
  __shared__ double sh[4];
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (v < nV) {
    constexpr int U = 8;  // loads in flight per lane (addresses known up front)
    const int k0 = vptr[v], k1 = vptr[v + 1];
    double sp = 0.0, sn = 0.0;
    for (int k = k0; k < k1; k += U) {
      double2 c[U];
#pragma unroll
      for (int u = 0; u < U; ++u) c[u] = contrib[k + u < k1 ? k + u : k0];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (k + u < k1) {
          sp += c[u].x;
          sn += c[u].y;
        }
    }
    const bool fx = fixed[v] != 0;
    sp = fx ? 0.0 : sp;
    sn = fx ? 0.0 : sn;
    rphi[v] = sp;
    rn[v] = sn;
    sq = sp * sp + sn * sn;
  }
  if (partial) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_down(sq, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  }
}

__global__ __launch_bounds__(128) void k_dd_vjp(
    int64_t nT, const int32_t *__restrict__ tets, const double *__restrict__ xyz,
    const double *__restrict__ sigma_len, const double *__restrict__ phi,
    const double *__restrict__ dens, const double *__restrict__ rphi,
    const double *__restrict__ rn, double *__restrict__ g_eps, double *__restrict__ g_mu) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nT) return;
  const int4 t = reinterpret_cast<const int4 *>(tets)[e];
  const int32_t id[4] = {t.x, t.y, t.z, t.w};
  double P[4][3], ph[4], dn[4], rp[4], rv[4], sg[6];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int d = 0; d < 3; ++d) P[q][d] = xyz[3 * (size_t)id[q] + d];
    ph[q] = phi[id[q]];
    dn[q] = dens[id[q]];
    rp[q] = rphi[id[q]];
    rv[q] = rn[id[q]];
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) sg[j] = sigma_len[6 * e + j];
  double ge = 0.0;
  {
    double G[4][3];
    const double vol = p1_geometry(P, G) / 6.0;
    double gp[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < 3; ++k) gp[k] += ph[q] * G[q][k];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      ge += 2.0 * rp[q] * vol * (G[q][0] * gp[0] + G[q][1] * gp[1] + G[q][2] * gp[2]);
  }
  double gm = 0.0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int a = kEA[j], b = kEB[j];
    gm += 2.0 * (rv[a] - rv[b]) * sg_edge_flux(ph, dn, a, b, sg[j]);
  }
  g_eps[e] = ge;
  g_mu[e] = gm;
}

// total of the n block partials in a fixed order (one workgroup of 256: lane t
// sums partials t, t + 256, ... in order, then a fixed shuffle tree), the same
// for every entry point
__global__ __launch_bounds__(256) void k_sum_fixed(int n, const double *__restrict__ partial,
                                                   double *__restrict__ out) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

static int dd_gather_blocks(const Ctx *c) { return blocks_for(c->nV, 256); }

struct DDBuffers {
  double *eps, *rho, *mu, *phi, *dens, *rphi, *rn, *ge, *gm, *partial;
  uint8_t *fixed;
};

// device copies of the host-side inputs (and outputs) of mg_dd_residual /
// mg_dd_vjp, released with the scope
struct DDHost : DDBuffers {
  Ctx *c;
  explicit DDHost(Ctx *c_) : DDBuffers{}, c(c_) {}
  ~DDHost() {
    for (void *p : {(void *)eps, (void *)rho, (void *)mu, (void *)ge, (void *)gm, (void *)phi,
                    (void *)dens, (void *)rphi, (void *)rn, (void *)partial, (void *)fixed})
      dfree(c, p);
  }
  DDHost(const DDHost &) = delete;
  DDHost &operator=(const DDHost &) = delete;
};

static void dd_upload(Ctx *c, DDHost &B, const double *eps, const double *rho, const double *mu,
                      const double *phi, const double *dens, const uint8_t *fixed) {
  const int64_t nT = c->nT, nV = c->nV;
  B.eps = dalloc<double>(c, nT);
  B.rho = dalloc<double>(c, nT);
  B.mu = dalloc<double>(c, nT);
  B.ge = dalloc<double>(c, nT);
  B.gm = dalloc<double>(c, nT);
  B.phi = dalloc<double>(c, nV);
  B.dens = dalloc<double>(c, nV);
  B.rphi = dalloc<double>(c, nV);
  B.rn = dalloc<double>(c, nV);
  B.partial = dalloc<double>(c, dd_gather_blocks(c) + 1);
  B.fixed = dalloc<uint8_t>(c, nV);
  MG_HIP(hipMemcpyAsync(B.eps, eps, sizeof(double) * nT, hipMemcpyHostToDevice, c->stream));
  MG_HIP(hipMemcpyAsync(B.rho, rho, sizeof(double) * nT, hipMemcpyHostToDevice, c->stream));
  MG_HIP(hipMemcpyAsync(B.mu, mu, sizeof(double) * nT, hipMemcpyHostToDevice, c->stream));
  MG_HIP(hipMemcpyAsync(B.phi, phi, sizeof(double) * nV, hipMemcpyHostToDevice, c->stream));
  MG_HIP(hipMemcpyAsync(B.dens, dens, sizeof(double) * nV, hipMemcpyHostToDevice, c->stream));
  MG_HIP(hipMemcpyAsync(B.fixed, fixed, nV, hipMemcpyHostToDevice, c->stream));
}

// residual; with loss != nullptr also the loss sum(r_phi^2 + r_n^2), whose
// block partials go to B.partial (dd_gather_blocks entries)
static void dd_residual_launch(Ctx *c, DDBuffers &B, double *loss = nullptr) {
  double2 *contrib = reinterpret_cast<double2 *>(c->dd_contrib);
  k_dd_elem<<<blocks_for(c->nT, 128), 128, 0, c->stream>>>(c->nT, c->tets, c->xyz, c->sigma_len,
                                                          B.eps, B.rho, B.mu, B.phi, B.dens,
                                                          c->vinc_pos, contrib);
  const int gb = dd_gather_blocks(c);
  k_dd_gather<<<gb, 256, 0, c->stream>>>(c->nV, c->vinc_ptr, contrib, B.fixed, B.rphi, B.rn,
                                         loss ? B.partial : nullptr);
  if (loss) k_sum_fixed<<<1, 256, 0, c->stream>>>(gb, B.partial, loss);
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

int mg_dual_fv_build(mg_ctx *c) {
  API_BEGIN(c)
  build_dual(c);
  API_END(c)
}

int mg_dual_fv_get(mg_ctx *c, double *edge_surf, double *edge_n, double *cell_volume,
                   double *sigma_el) {
  API_BEGIN(c)
  MG_CHECK(c->dual_ok, MG_ERR_STATE, "dual data not built");
  if (edge_surf)
    MG_HIP(hipMemcpyAsync(edge_surf, c->edge_surf, sizeof(double) * c->nE, hipMemcpyDeviceToHost,
                          c->stream));
  if (edge_n)
    MG_HIP(hipMemcpyAsync(edge_n, c->edge_n, sizeof(double) * 3 * c->nE, hipMemcpyDeviceToHost,
                          c->stream));
  if (cell_volume)
    MG_HIP(hipMemcpyAsync(cell_volume, c->cell_volume, sizeof(double) * c->nV,
                          hipMemcpyDeviceToHost, c->stream));
  if (sigma_el)
    MG_HIP(hipMemcpyAsync(sigma_el, c->sigma_el, sizeof(double) * 6 * c->nT,
                          hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_dd_residual(mg_ctx *c, const double *eps, const double *rho, const double *mu,
                   const double *phi, const double *dens, const uint8_t *fixed, double *r_phi,
                   double *r_n, double *loss) {
  API_BEGIN(c)
  MG_CHECK(eps && rho && mu && phi && dens && fixed, MG_ERR_ARG, "dd: null input");
  MG_CHECK(c->dual_ok, MG_ERR_STATE, "dd: dual data not built");
  DDHost B(c);
  dd_upload(c, B, eps, rho, mu, phi, dens, fixed);
  double *dloss = B.partial + dd_gather_blocks(c);
  dd_residual_launch(c, B, loss ? dloss : nullptr);
  if (r_phi)
    MG_HIP(hipMemcpyAsync(r_phi, B.rphi, sizeof(double) * c->nV, hipMemcpyDeviceToHost, c->stream));
  if (r_n)
    MG_HIP(hipMemcpyAsync(r_n, B.rn, sizeof(double) * c->nV, hipMemcpyDeviceToHost, c->stream));
  if (loss)
    MG_HIP(hipMemcpyAsync(loss, dloss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_dd_vjp(mg_ctx *c, const double *eps, const double *rho, const double *mu,
              const double *phi, const double *dens, const uint8_t *fixed, double *g_eps,
              double *g_mu) {
  API_BEGIN(c)
  MG_CHECK(eps && rho && mu && phi && dens && fixed && g_eps && g_mu, MG_ERR_ARG,
           "dd: null argument");
  MG_CHECK(c->dual_ok, MG_ERR_STATE, "dd: dual data not built");
  DDHost B(c);
  dd_upload(c, B, eps, rho, mu, phi, dens, fixed);
  dd_residual_launch(c, B);
  k_dd_vjp<<<blocks_for(c->nT, 128), 128, 0, c->stream>>>(c->nT, c->tets, c->xyz, c->sigma_len,
                                                          B.phi, B.dens, B.rphi, B.rn, B.ge, B.gm);
  MG_HIP(hipGetLastError());
  MG_HIP(hipMemcpyAsync(g_eps, B.ge, sizeof(double) * c->nT, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipMemcpyAsync(g_mu, B.gm, sizeof(double) * c->nT, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mg_dd_residual_device(mg_ctx *c, const double *eps, const double *rho, const double *mu,
                          const double *phi, const double *dens, const uint8_t *fixed,
                          double *r_phi, double *r_n, double *loss) {
  API_BEGIN(c)
  MG_CHECK(eps && rho && mu && phi && dens && fixed && r_phi && r_n, MG_ERR_ARG,
           "dd: null argument");
  MG_CHECK(c->dual_ok, MG_ERR_STATE, "dd: dual data not built");
  DDBuffers B{const_cast<double *>(eps), const_cast<double *>(rho), const_cast<double *>(mu),
              const_cast<double *>(phi), const_cast<double *>(dens), r_phi, r_n, nullptr, nullptr,
              nullptr, const_cast<uint8_t *>(fixed)};
  if (loss) B.partial = (double *)scratch(c, sizeof(double) * dd_gather_blocks(c));
  dd_residual_launch(c, B, loss);
  API_END(c)
}

int mg_dd_vjp_device(mg_ctx *c, const double *eps, const double *rho, const double *mu,
                     const double *phi, const double *dens, const uint8_t *fixed, double *r_phi,
                     double *r_n, double *g_eps, double *g_mu) {
  API_BEGIN(c)
  MG_CHECK(eps && rho && mu && phi && dens && fixed && r_phi && r_n && g_eps && g_mu,
           MG_ERR_ARG, "dd: null argument");
  MG_CHECK(c->dual_ok, MG_ERR_STATE, "dd: dual data not built");
  DDBuffers B{const_cast<double *>(eps), const_cast<double *>(rho), const_cast<double *>(mu),
              const_cast<double *>(phi), const_cast<double *>(dens), r_phi, r_n, g_eps, g_mu,
              nullptr, const_cast<uint8_t *>(fixed)};
  dd_residual_launch(c, B);
  k_dd_vjp<<<blocks_for(c->nT, 128), 128, 0, c->stream>>>(c->nT, c->tets, c->xyz, c->sigma_len,
                                                          B.phi, B.dens, B.rphi, B.rn, B.ge, B.gm);
  MG_HIP(hipGetLastError());
  API_END(c)
}

int mg_bernoulli(mg_ctx *c, int64_t n, const double *x, double *y) {
  API_BEGIN(c)
  MG_CHECK(n >= 0 && (n == 0 || (x && y)), MG_ERR_ARG, "bernoulli: bad arguments");
  if (n == 0) return MG_OK;
  DevTmp<double> dx(c, n), dy(c, n);
  MG_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  k_bernoulli<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, dx, dy);
  MG_HIP(hipGetLastError());
  MG_HIP(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

}  // extern "C"
