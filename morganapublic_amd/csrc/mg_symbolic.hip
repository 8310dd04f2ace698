// Symbolic phase on the GPU: edge numbering, dof map, Dirichlet numbering,
// CSR pattern, element -> CSR slot map, row-incidence records (12- or
// 16-byte P2 form), SpMV row-block descriptors and the 16-bit column form.
//
// Everything is sort-based (rocPRIM radix sort through hipCUB) so it scales to
// the 10^8 (row, column) pairs of a 64^3 P2 mesh without host work, and every
// result is deterministic (stable sorts, integer scans, no float atomics).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstring>
#include <type_traits>

#include "mg_internal.h"

namespace mg {

__constant__ int c_edge_v[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
__constant__ int c_sigma[10][10];  // relabelled basis -> original basis, per row basis l

static int bitlen(uint64_t x) {
  int b = 0;
  while (x) {
    ++b;
    x >>= 1;
  }
  return b;
}

template <typename K, typename V>
static void sort_pairs(Ctx *c, const K *kin, K *kout, const V *vin, V *vout, int64_t n,
                       int end_bit) {
  MG_CHECK(n < (1ll << 31), MG_ERR_SIZE, "sort: more than 2^31 items");
  size_t tb = 0;
  MG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, (int)n, 0, end_bit,
                                            c->stream));
  void *t = scratch(c, tb);
  MG_HIP(hipcub::DeviceRadixSort::SortPairs(t, tb, kin, kout, vin, vout, (int)n, 0, end_bit,
                                            c->stream));
}

static void inclusive_sum(Ctx *c, const int32_t *in, int32_t *out, int64_t n) {
  size_t tb = 0;
  MG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out, (int)n, c->stream));
  void *t = scratch(c, tb);
  MG_HIP(hipcub::DeviceScan::InclusiveSum(t, tb, in, out, (int)n, c->stream));
}

static void exclusive_sum(Ctx *c, const int32_t *in, int32_t *out, int64_t n) {
  size_t tb = 0;
  MG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, (int)n, c->stream));
  void *t = scratch(c, tb);
  MG_HIP(hipcub::DeviceScan::ExclusiveSum(t, tb, in, out, (int)n, c->stream));
}

template <typename T>
static T read_back(Ctx *c, const T *d) {
  T h;
  MG_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  return h;
}

// ---- edges --------------------------------------------------------------------
__global__ void k_edge_keys(int64_t nT, const int32_t *__restrict__ tets, uint64_t *__restrict__ keys,
                            int32_t *__restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * nT) return;
  int64_t e = i / 6;
  int j = (int)(i - 6 * e);
  uint32_t a = (uint32_t)tets[4 * e + c_edge_v[j][0]], b = (uint32_t)tets[4 * e + c_edge_v[j][1]];
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  keys[i] = ((uint64_t)lo << 32) | hi;
  vals[i] = (int32_t)i;
}

__global__ void k_heads_u64(int64_t n, const uint64_t *__restrict__ keys, uint64_t invalid,
                            int32_t *__restrict__ head) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = keys[i];
  head[i] = (k != invalid) && (i == 0 || keys[i - 1] != k);
}

__global__ void k_edge_scatter(int64_t n, const uint64_t *__restrict__ keys,
                               const int32_t *__restrict__ vals, const int32_t *__restrict__ id,
                               const int32_t *__restrict__ head, int32_t *__restrict__ elem_edge,
                               uint64_t *__restrict__ edge_keys, int32_t *__restrict__ edge_nodes) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t eid = id[i] - 1;
  elem_edge[vals[i]] = eid;
  if (head[i]) {
    uint64_t k = keys[i];
    edge_keys[eid] = k;
    edge_nodes[2 * eid] = (int32_t)(k >> 32);
    edge_nodes[2 * eid + 1] = (int32_t)(k & 0xffffffffu);
  }
}

// connect3d::buildConnectivity (connect3d.hpp:417-491) + the single-rank
// pMapItem global numbering (pVectGlobalManip.hpp:417-578): edge gid = rank
// of the sorted (node, node) pair in lexicographic order.
void build_edges(Ctx *c) {
  const int64_t n = 6 * c->nT;
  uint64_t *k0 = dalloc<uint64_t>(c, n), *k1 = dalloc<uint64_t>(c, n);
  int32_t *v0 = dalloc<int32_t>(c, n), *v1 = dalloc<int32_t>(c, n);
  int32_t *head = dalloc<int32_t>(c, n), *id = dalloc<int32_t>(c, n);
  k_edge_keys<<<blocks_for(n, 256), 256, 0, c->stream>>>(c->nT, c->tets, k0, v0);
  MG_HIP(hipGetLastError());
  sort_pairs(c, k0, k1, v0, v1, n, 32 + bitlen((uint64_t)c->nV));
  k_heads_u64<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, k1, ~0ull, head);
  inclusive_sum(c, head, id, n);
  c->nE = read_back(c, id + n - 1);
  c->edge_keys = dalloc<uint64_t>(c, c->nE);
  c->edge_nodes = dalloc<int32_t>(c, 2 * c->nE);
  c->elem_edge = dalloc<int32_t>(c, n);
  k_edge_scatter<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, k1, v1, id, head, c->elem_edge,
                                                          c->edge_keys, c->edge_nodes);
  MG_HIP(hipGetLastError());
  MG_HIP(hipStreamSynchronize(c->stream));
  for (void *p : {(void *)k0, (void *)k1, (void *)v0, (void *)v1, (void *)head, (void *)id})
    dfree(c, p);
  c->edges_ok = true;
}

// ---- dofs -----------------------------------------------------------------------
// dofMapStatic3d::startup (dofMapStatic3d.hpp:471-527): vertex dofs first, then
// one dof per edge at numVertices + edge gid.
__global__ void k_dofs(int64_t nT, int nloc, int64_t nV, const int32_t *__restrict__ tets,
                       const int32_t *__restrict__ elem_edge, int32_t *__restrict__ dofs) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nT) return;
  for (int l = 0; l < 4; ++l) dofs[nloc * e + l] = tets[4 * e + l];
  if (nloc == 10)
    for (int j = 0; j < 6; ++j) dofs[10 * e + 4 + j] = (int32_t)(nV + elem_edge[6 * e + j]);
}

void build_dofs(Ctx *c) {
  c->ndofs = c->order == 1 ? c->nV : c->nV + c->nE;
  MG_CHECK(c->ndofs < (1ll << 31), MG_ERR_SIZE, "dofs: more than 2^31");
  c->dofs = dalloc<int32_t>(c, c->nloc * c->nT);
  k_dofs<<<blocks_for(c->nT, 256), 256, 0, c->stream>>>(c->nT, c->nloc, c->nV, c->tets,
                                                        c->elem_edge, c->dofs);
  MG_HIP(hipGetLastError());
  MG_HIP(hipStreamSynchronize(c->stream));
}

// ---- Dirichlet ------------------------------------------------------------------
__device__ int64_t find_edge(const uint64_t *__restrict__ keys, int64_t nE, uint64_t k) {
  int64_t lo = 0, hi = nE - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (nE > 0 && keys[lo] == k) ? lo : -1;
}

__global__ void k_dirichlet(int64_t nB, const int32_t *__restrict__ btri,
                            const int32_t *__restrict__ bgeo, int ngeo,
                            const int32_t *__restrict__ geo, int order, int64_t nV, int64_t nE,
                            const uint64_t *__restrict__ edge_keys, uint8_t *__restrict__ fixed,
                            int *__restrict__ bad) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nB) return;
  bool sel = false;
  for (int g = 0; g < ngeo; ++g) sel |= bgeo[t] == geo[g];
  if (!sel) return;
  int v[3] = {btri[3 * t], btri[3 * t + 1], btri[3 * t + 2]};
  for (int k = 0; k < 3; ++k) fixed[v[k]] = 1;
  if (order == 2)
    for (int k = 0; k < 3; ++k) {
      uint32_t a = (uint32_t)v[k], b = (uint32_t)v[(k + 1) % 3];
      uint64_t key = ((uint64_t)(a < b ? a : b) << 32) | (a < b ? b : a);
      int64_t id = find_edge(edge_keys, nE, key);
      if (id < 0)
        atomicAdd(bad, 1);
      else
        fixed[nV + id] = 1;
    }
}

__global__ void k_free_flags(int64_t n, const uint8_t *__restrict__ fixed, int32_t *__restrict__ f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = fixed[i] ? 0 : 1;
}

__global__ void k_sysmap(int64_t n, const uint8_t *__restrict__ fixed,
                         const int32_t *__restrict__ scan, int32_t *__restrict__ map) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) map[i] = fixed[i] ? -1 : scan[i];
}

void build_dirichlet(Ctx *c, int n_geo, const int32_t *geo_ids) {
  const int64_t N = c->ndofs;
  if (!c->fixed) c->fixed = dalloc<uint8_t>(c, N);
  if (!c->sysmap) c->sysmap = dalloc<int32_t>(c, N);
  int32_t *dgeo = dalloc<int32_t>(c, 64);
  int *bad = dalloc<int>(c, 1);
  MG_HIP(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
  MG_HIP(hipMemsetAsync(c->fixed, 0, N, c->stream));
  if (n_geo)
    MG_HIP(hipMemcpyAsync(dgeo, geo_ids, sizeof(int32_t) * n_geo, hipMemcpyHostToDevice,
                          c->stream));
  if (c->nB && n_geo)
    k_dirichlet<<<blocks_for(c->nB, 256), 256, 0, c->stream>>>(c->nB, c->btri, c->bgeo, n_geo,
                                                               dgeo, c->order, c->nV, c->nE,
                                                               c->edge_keys, c->fixed, bad);
  int32_t *f = dalloc<int32_t>(c, N + 1), *s = dalloc<int32_t>(c, N + 1);
  k_free_flags<<<blocks_for(N, 256), 256, 0, c->stream>>>(N, c->fixed, f);
  MG_HIP(hipMemsetAsync(f + N, 0, sizeof(int32_t), c->stream));
  exclusive_sum(c, f, s, N + 1);
  k_sysmap<<<blocks_for(N, 256), 256, 0, c->stream>>>(N, c->fixed, s, c->sysmap);
  MG_HIP(hipGetLastError());
  c->nfree = read_back(c, s + N);
  c->nfree_vertex = read_back(c, s + c->nV);
  int hb = read_back(c, bad);
  for (void *p : {(void *)dgeo, (void *)bad, (void *)f, (void *)s}) dfree(c, p);
  MG_CHECK(hb == 0, MG_ERR_MESH, "dirichlet: a boundary triangle edge is not a mesh edge");
  // a new Dirichlet set invalidates the interior system
  free_system(c, c->sys[MG_SYS_INTERIOR]);
  c->dirichlet_ok = true;
}

// ---- pattern ----------------------------------------------------------------------
// Contribution c = e*nl*nl + i*nl + j is the entry (row = test basis j, column =
// field basis i) of element e's column-major local matrix, the order in which
// traitsMatrixBuilder::buildEpetraCrs hands it to InsertGlobalValues(COLUMN_MAJOR).
__global__ void k_pattern_keys(int64_t nT, int nl, const int32_t *__restrict__ dofs,
                               const int32_t *__restrict__ map, uint64_t nrows, uint64_t invalid,
                               uint64_t *__restrict__ keys, int32_t *__restrict__ vals) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = nT * nl * nl;
  if (c >= n) return;
  int64_t e = c / (nl * nl);
  int rem = (int)(c - e * nl * nl);
  int i = rem / nl, j = rem - i * nl;
  int32_t row = dofs[e * nl + j], col = dofs[e * nl + i];
  if (map) {
    row = map[row];
    col = map[col];
  }
  keys[c] = (row < 0 || col < 0) ? invalid : (uint64_t)row * nrows + (uint64_t)col;
  vals[c] = (int32_t)c;
}

__global__ void k_pattern_scatter(int64_t n, const uint64_t *__restrict__ keys,
                                  const int32_t *__restrict__ vals, const int32_t *__restrict__ id,
                                  const int32_t *__restrict__ head, uint64_t nrows,
                                  uint64_t invalid, int32_t *__restrict__ slot,
                                  int32_t *__restrict__ colidx, int32_t *__restrict__ rowcnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = keys[i];
  if (k == invalid) {
    slot[vals[i]] = -1;
    return;
  }
  int32_t s = id[i] - 1;
  slot[vals[i]] = s;
  if (head[i]) {
    colidx[s] = (int32_t)(k % nrows);
    atomicAdd(&rowcnt[k / nrows], 1);
  }
}

__global__ void k_diagpos(int64_t nrows, const int32_t *__restrict__ rowptr,
                          const int32_t *__restrict__ colidx, int32_t *__restrict__ diag) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int lo = rowptr[r], hi = rowptr[r + 1] - 1, d = -1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int cc = colidx[mid];
    if (cc == r) {
      d = mid;
      break;
    }
    if (cc < r)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  diag[r] = d;
}

__global__ void k_inc_keys(int64_t nT, int nl, const int32_t *__restrict__ dofs,
                           const int32_t *__restrict__ map, uint64_t invalid,
                           uint64_t *__restrict__ keys, int32_t *__restrict__ vals) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nT * nl) return;
  int64_t e = p / nl;
  int l = (int)(p - e * nl);
  int32_t row = dofs[p];
  if (map) row = map[row];
  keys[p] = row < 0 ? invalid : (uint64_t)row * nl + l;
  vals[p] = (int32_t)((e << 4) | l);
}

__global__ void k_inc_count(int64_t n, const uint64_t *__restrict__ keys, uint64_t invalid, int nl,
                            int32_t *__restrict__ cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = keys[i];
  if (k != invalid) atomicAdd(&cnt[k / (uint64_t)nl], 1);
}

// 16-bit column form of one SpMV row block (one workgroup per block, the
// block's <= kSpmvNnz entries spread over 256 lanes).  Clusters are formed
// greedily: the smallest column not yet placed opens cluster c at base_c, and
// every unplaced column in [base_c, base_c + kColSpan) joins it as
// c << 13 | (col - base_c).  The columns of a block of consecutive mesh rows
// lie in a few narrow bands (the vertex and edge dofs of neighbouring mesh
// planes), so a few clusters cover them; a block that needs more than
// kColClusters sets *fail and the system keeps int32 columns.
__global__ __launch_bounds__(256) void k_col_compress(BlkDesc *__restrict__ desc,
                                                      const int32_t *__restrict__ colidx,
                                                      uint16_t *__restrict__ ofs, int ncols,
                                                      int *__restrict__ fail) {
  constexpr int PER = kSpmvNnz / 256;
  __shared__ int red[4];
  const int b = blockIdx.x;
  const int s0 = desc[b].s0, ns = desc[b].ns;
  int col[PER];
  bool left[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int k = threadIdx.x + u * 256;
    left[u] = k < ns;
    col[u] = left[u] ? colidx[s0 + k] : 0;
  }
  for (int cl = 0; cl < kColClusters; ++cl) {
    int m = INT32_MAX;
#pragma unroll
    for (int u = 0; u < PER; ++u)
      if (left[u]) m = min(m, col[u]);
    for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    m = min(min(red[0], red[1]), min(red[2], red[3]));
    __syncthreads();
    // bases stay <= max(ncols - kColSpan, 0), so any 16-bit entry of the
    // array decodes inside [0, ncols) against any block's table (the SpMV
    // streams read up to 3 entries past a block's end without a clamp); the
    // window [base, base + kColSpan) still holds m
    if (m != INT32_MAX) m = min(m, max(ncols - kColSpan, 0));
    if (threadIdx.x == 0) desc[b].base[cl] = m == INT32_MAX ? 0 : m;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      if (left[u] && col[u] - m < kColSpan) {
        ofs[s0 + threadIdx.x + u * 256] = (uint16_t)((cl << 13) | (col[u] - m));
        left[u] = false;
      }
    }
  }
  bool any = false;
#pragma unroll
  for (int u = 0; u < PER; ++u) any |= left[u];
  if (any) *fail = 1;
}

// Records with uint8 (RecP1, RecP2) or uint16 (RecP1w, RecP2w) positions; the
// largest value of the position type marks an eliminated column.
template <int NL, typename Rec>
__global__ void k_inc_records(int64_t n_inc, const uint64_t *__restrict__ keys,
                              const int32_t *__restrict__ vals, const int32_t *__restrict__ slot,
                              const int32_t *__restrict__ rowptr, Rec *__restrict__ rec,
                              int *__restrict__ bad) {
  using P = typename std::remove_reference<decltype(rec->pos[0])>::type;
  constexpr int kElim = (int)(P)~0u;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inc) return;
  int32_t v = vals[i];
  int64_t e = v >> 4;
  int l = v & 15;
  int64_t row = (int64_t)(keys[i] / NL);
  int32_t base = rowptr[row];
  Rec r;
  r.el = v;
  for (int m = 0; m < NL; ++m) {
    // column of relabelled basis m = original basis c_sigma[l][m]; row basis l
    int32_t s = slot[e * NL * NL + c_sigma[l][m] * NL + l];
    int p = s < 0 ? kElim : s - base;
    if (s >= 0 && (p < 0 || p >= kElim)) {
      atomicAdd(bad, 1);
      p = kElim;
    }
    r.pos[m] = (P)p;
  }
  rec[i] = r;
}

// Compact P2 records: the 9 positions other than the row's own column, 7 bits
// each; the own column (relabelled basis 0 or 4) must be the row's diagonal.
__global__ void k_inc_records_c(int64_t n_inc, const uint64_t *__restrict__ keys,
                                const int32_t *__restrict__ vals,
                                const int32_t *__restrict__ slot,
                                const int32_t *__restrict__ rowptr,
                                const int32_t *__restrict__ diagpos, RecP2c *__restrict__ rec,
                                int *__restrict__ bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inc) return;
  const int32_t v = vals[i];
  const int64_t e = v >> 4;
  const int l = v & 15;
  const int lrow = l < 4 ? 0 : 4;
  const int64_t row = (int64_t)(keys[i] / 10);
  const int32_t base = rowptr[row];
  uint64_t w = 0;
  int k = 0;
  for (int m = 0; m < 10; ++m) {
    const int32_t s = slot[e * 100 + c_sigma[l][m] * 10 + l];
    if (m == lrow) {
      if (s != diagpos[row]) atomicAdd(bad, 1);
      continue;
    }
    int p = s < 0 ? 0x7F : s - base;
    if (s >= 0 && (p < 0 || p >= 0x7F)) {
      atomicAdd(bad, 1);
      p = 0x7F;
    }
    w |= (uint64_t)p << (7 * k);
    ++k;
  }
  rec[i] = RecP2c{v, (uint32_t)w, (uint32_t)(w >> 32)};
}

// Assembly launches: the vertex-dof rows (relabelled local basis 0) and the
// edge-dof rows (basis 4) of a P2 system go to separate launches; each launch
// uses 64-row blocks and the LDS size of its largest block.
static int env_int(const char *name, int dflt) {
  const char *v = getenv(name);
  return v ? atoi(v) : dflt;
}

// Patched P1 records: the 8-bit positions of relabelled bases 1-3; the own
// column (basis 0) must be the row's diagonal.  k_patch adds the patch indices.
__global__ void k_inc_records_p1p(int64_t n_inc, const uint64_t *__restrict__ keys,
                                  const int32_t *__restrict__ vals,
                                  const int32_t *__restrict__ slot,
                                  const int32_t *__restrict__ rowptr,
                                  const int32_t *__restrict__ diagpos, RecP1p *__restrict__ rec,
                                  int *__restrict__ bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inc) return;
  const int32_t v = vals[i];
  const int64_t e = v >> 4;
  const int l = v & 15;
  const int64_t row = (int64_t)(keys[i] / 4);
  const int32_t base = rowptr[row];
  uint32_t w = 0;
  for (int m = 0; m < 4; ++m) {
    const int32_t s = slot[e * 16 + c_sigma[l][m] * 4 + l];
    if (m == 0) {
      if (s != diagpos[row]) atomicAdd(bad, 1);
      continue;
    }
    int p = s < 0 ? 0xFF : s - base;
    if (s >= 0 && (p < 0 || p >= 0xFF)) {
      atomicAdd(bad, 1);
      p = 0xFF;
    }
    w |= (uint32_t)p << (8 * (m - 1));
  }
  rec[i] = RecP1p{v, w, 0u};
}

// Vertex patch of patch block b: the vertices of the elements of the block's
// incidences, sorted in LDS (bitonic) and made distinct.  Blocks b < nvb hold
// vertex-dof rows from row 0, the others edge-dof rows from row nvrows.
// FILL false: cnt[b] = the number of distinct vertices (-1: more than
// kPatchKeys references).  FILL true, for blocks given a patch
// (pptr[b + 1] > pptr[b]): the vertices go to vid[pptr[b] ..]; a lane per row
// sets row_lv[r] and the patch indices of each of the row's incidences (in
// its P1 record, in inc_lv for P2): of relabelled vertices 1-3 (vertex rows)
// or 2-3 and the orientation bit (P2 edge rows, whose endpoints are
// relabelled vertices 0 and 1 of the row's first record).
__device__ __forceinline__ int patch_find(const int32_t *u, int n, int32_t v) {
  int lo = 0, hi = n;  // first index with u[index] >= v
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (u[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void patch_block_rows(int b, int64_t nvrows, int nvb, int64_t nrows,
                                                 int64_t &r0, int64_t &r1) {
  if (b < nvb) {
    r0 = (int64_t)b * kPatchRows;
    r1 = min<int64_t>(r0 + kPatchRows, nvrows);
  } else {
    r0 = nvrows + (int64_t)(b - nvb) * kPatchRows;
    r1 = min<int64_t>(r0 + kPatchRows, nrows);
  }
}

template <bool FILL, typename Rec>
__global__ __launch_bounds__(256) void k_patch(int64_t nvrows, int nvb, int64_t nrows,
                                               const int32_t *__restrict__ inc_ptr,
                                               Rec *__restrict__ rec,
                                               const int32_t *__restrict__ tets,
                                               const int32_t *__restrict__ pptr,
                                               int32_t *__restrict__ out,
                                               uint32_t *__restrict__ row_lv,
                                               uint32_t *__restrict__ inc_lv) {
  __shared__ int32_t key[kPatchKeys];
  __shared__ int32_t uniq[kPatchMax];
  __shared__ int32_t part[257];
  const int b = blockIdx.x, t = threadIdx.x;
  int64_t r0, r1;
  patch_block_rows(b, nvrows, nvb, nrows, r0, r1);
  const int i0 = inc_ptr[r0], i1 = inc_ptr[r1];
  const int n4 = 4 * (i1 - i0);
  if (FILL) {
    if (pptr[b + 1] == pptr[b]) return;
  } else if (n4 > kPatchKeys) {
    if (t == 0) out[b] = -1;
    return;
  }
  int P = 1;
  while (P < n4) P <<= 1;
  for (int j = t; j < P; j += 256)
    key[j] = j < n4 ? tets[4 * (size_t)(rec[i0 + (j >> 2)].el >> 4) + (j & 3)] : INT_MAX;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < P; i += 256) {
        const int o = i ^ j;
        if (o > i) {
          const int32_t a = key[i], c = key[o];
          if ((a > c) == ((i & k) == 0)) {
            key[i] = c;
            key[o] = a;
          }
        }
      }
      __syncthreads();
    }
  // distinct values: thread t counts the first occurrences in its run of
  // consecutive keys, an exclusive scan of the counts gives their indices
  const int per = (P + 255) / 256;
  const int a0 = min(t * per, n4), a1 = min(a0 + per, n4);
  int nu = 0;
  for (int i = a0; i < a1; ++i) nu += (i == 0 || key[i] != key[i - 1]);
  part[t + 1] = nu;
  __syncthreads();
  if (t == 0) {
    part[0] = 0;
    for (int k = 1; k <= 256; ++k) part[k] += part[k - 1];
  }
  __syncthreads();
  const int U = part[256];
  if (!FILL) {
    if (t == 0) out[b] = U;
    return;
  }
  int u = part[t];
  for (int i = a0; i < a1; ++i)
    if (i == 0 || key[i] != key[i - 1]) {
      uniq[u] = key[i];
      out[pptr[b] + u] = key[i];
      ++u;
    }
  __syncthreads();
  auto find = [&](int64_t e, int l, int m) {
    return (uint32_t)patch_find(uniq, U, tets[4 * e + c_sigma[l][m]]);
  };
  for (int64_t r = r0 + t; r < r1; r += 256) {
    const int j0 = inc_ptr[r], j1 = inc_ptr[r + 1];
    uint32_t own = 0;
    int32_t va = -1;  // the row's first endpoint (edge rows)
    if (j0 < j1) {
      const int32_t el = rec[j0].el;
      const int64_t e = el >> 4;
      const int l = el & 15;
      own = find(e, l, 0);
      if (l >= 4) {
        own |= find(e, l, 1) << 16;
        va = tets[4 * e + c_sigma[l][0]];
      }
    }
    row_lv[r] = own;
    for (int j = j0; j < j1; ++j) {
      Rec q = rec[j];
      const int64_t e = q.el >> 4;
      const int l = q.el & 15;
      if constexpr (std::is_same<Rec, RecP1p>::value) {
        const uint32_t v0 = find(e, l, 1), v1 = find(e, l, 2), v2 = find(e, l, 3);
        q.lo = (q.lo & 0xFFFFFFu) | (v0 << 24);
        q.hi = (v0 >> 8) | (v1 << 4) | (v2 << 16);
        rec[j] = q;
      } else if (l < 4) {
        inc_lv[j] = find(e, l, 1) | (find(e, l, 2) << 10) | (find(e, l, 3) << 20);
      } else {
        const uint32_t sw = tets[4 * e + c_sigma[l][0]] == va ? 0u : 1u;
        inc_lv[j] = find(e, l, 2) | (find(e, l, 3) << 10) | (sw << 30);
      }
    }
  }
}

// RecP2e edge-row records from the RecP2c records [i0, n_inc) and their patch
// words: 7-bit positions re-packed in 5 bits (0x7F -> 31, eliminated), the
// 10-bit patch indices of relabelled vertices 2 and 3 in 9 bits, and the
// orientation bit.  bad counts fields that do not fit.
__global__ void k_edge5(int64_t i0, int64_t n_inc, RecP2c *__restrict__ rec,
                        const uint32_t *__restrict__ inc_lv, int *__restrict__ bad) {
  const int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inc) return;
  const RecP2c q = rec[i];
  const uint64_t w7 = (uint64_t)q.lo | ((uint64_t)q.hi << 32);
  uint64_t w = 0;
  int nbad = 0;
  for (int k = 0; k < 9; ++k) {
    const uint32_t p = (uint32_t)(w7 >> (7 * k)) & 0x7Fu;
    nbad += (p != 0x7Fu && p >= 31u);
    w |= (uint64_t)(p == 0x7Fu ? 31u : (p & 31u)) << (5 * k);
  }
  const uint32_t lv = inc_lv[i];
  const uint32_t a = lv & 0x3FFu, b = (lv >> 10) & 0x3FFu, sw = (lv >> 30) & 1u;
  nbad += (a > 511u) + (b > 511u);
  w |= ((uint64_t)(a & 511u) << 45) | ((uint64_t)(b & 511u) << 54) | ((uint64_t)sw << 63);
  if (nbad) atomicAdd(bad, nbad);
  reinterpret_cast<RecP2e *>(rec)[i] = RecP2e{q.el, (uint32_t)w, (uint32_t)(w >> 32)};
}

// Packed row words of a patched system (System::prow); bad counts rows
// whose fields do not fit.
__global__ void k_prow(int64_t nvrows, int64_t nrows, const int32_t *__restrict__ rowptr,
                       const int32_t *__restrict__ inc_ptr, const int32_t *__restrict__ diagpos,
                       const uint32_t *__restrict__ row_lv, uint64_t *__restrict__ prow,
                       int *__restrict__ bad) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int64_t r0 = r < nvrows ? (r / kPatchRows) * kPatchRows
                                : nvrows + ((r - nvrows) / kPatchRows) * kPatchRows;
  const int vo = rowptr[r] - rowptr[r0];
  const int io = inc_ptr[r] - inc_ptr[r0];
  const int cnt = inc_ptr[r + 1] - inc_ptr[r];
  const int self = diagpos[r] - rowptr[r];
  const uint32_t a = row_lv[r] & 0xFFFFu, b = row_lv[r] >> 16;
  const bool ok = vo >= 0 && vo <= 0xFFFF && self >= 0 && self <= 255 && io >= 0 && io <= 4095 &&
                  cnt >= 0 && cnt <= 255 && a <= 1023u && b <= 1023u;
  prow[r] = (uint64_t)(vo & 0xFFFF) | ((uint64_t)(self & 255) << 16) |
            ((uint64_t)(io & 4095) << 24) | ((uint64_t)(cnt & 255) << 36) |
            ((uint64_t)(a & 1023u) << 44) | ((uint64_t)(b & 1023u) << 54);
  if (!ok) atomicAdd(bad, 1);
}

// RecP1q form of a patched P1 record, its element taken relative to eb (the
// lowest element of its block) and accepted below 2^span (span <= 22, the
// field's width); nbad counts the fields that do not fit
__device__ __forceinline__ RecP1q p1q_pack(const RecP1p &q, int eb, int span, int &nbad) {
  uint64_t w = 0;
  for (int m = 0; m < 3; ++m) {
    const uint32_t p = (q.lo >> (8 * m)) & 0xFFu;
    if (p != 0xFFu && p >= 31u) ++nbad;
    w |= (uint64_t)(p == 0xFFu ? 31u : p & 31u) << (5 * m);
  }
  const uint32_t lv[3] = {(q.lo >> 24) | ((q.hi & 0xFu) << 8), (q.hi >> 4) & 0xFFFu,
                          (q.hi >> 16) & 0xFFFu};
  for (int o = 0; o < 3; ++o) {
    if (lv[o] > 511u) ++nbad;
    w |= (uint64_t)(lv[o] & 511u) << (15 + 9 * o);
  }
  const int64_t off = (int64_t)(q.el >> 4) - eb;
  if (off < 0 || off >= (1 << span)) ++nbad;
  w |= (uint64_t)(off & ((1 << 22) - 1)) << 42;
  return RecP1q{(uint32_t)w, (uint32_t)(w >> 32)};
}

// ELL copy of the incidence records (and P2 patch words) of every patch
// block: one workgroup per block fills its slots (step k, lane t) from the
// CSR-ordered records (layout: System::ell_rec), as they are (Out = Rec) or
// in the RecP1q form (Out = RecP1q: ebase[b] = the block's lowest element,
// slots of rows past the block's last or without incidences hold a zero
// record, bad counts records that do not fit).
template <typename Rec, typename Out>
__global__ __launch_bounds__(256) void k_ell_fill(int64_t nvrows, int nvb, int64_t nrows,
                                                  const int32_t *__restrict__ inc_ptr,
                                                  const Rec *__restrict__ rec,
                                                  const uint32_t *__restrict__ inc_lv,
                                                  int64_t base_v, int steps_v, int lpr_v,
                                                  int64_t base_e, int steps_e, int lpr_e,
                                                  Out *__restrict__ ell,
                                                  uint32_t *__restrict__ ell_lv,
                                                  int32_t *__restrict__ ebase, int span,
                                                  int *__restrict__ bad) {
  constexpr bool kPack = !std::is_same<Rec, Out>::value;
  const int b = blockIdx.x;
  int64_t r0, r1;
  patch_block_rows(b, nvrows, nvb, nrows, r0, r1);
  const bool vx = b < nvb;
  const int lpr = vx ? lpr_v : lpr_e, steps = vx ? steps_v : steps_e;
  const int nt = kPatchRows * lpr;
  const int64_t e0 = (vx ? base_v : base_e) + (int64_t)(vx ? b : b - nvb) * steps * nt;
  int eb = 0, nbad = 0;
  if constexpr (kPack) {
    __shared__ int emin;
    if (threadIdx.x == 0) emin = INT_MAX;
    __syncthreads();
    int m = INT_MAX;
    const int ia = r0 < r1 ? inc_ptr[r0] : 0, ib = r0 < r1 ? inc_ptr[r1] : 0;
    for (int j = ia + (int)threadIdx.x; j < ib; j += blockDim.x) m = min(m, rec[j].el >> 4);
    atomicMin(&emin, m);
    __syncthreads();
    eb = emin == INT_MAX ? 0 : emin;
    if (threadIdx.x == 0) ebase[b] = eb;
  }
  for (int s = threadIdx.x; s < steps * nt; s += blockDim.x) {
    const int k = s / nt, t = s - k * nt;
    const int64_t r = r0 + t / lpr;
    const int p = t % lpr + k * lpr;
    int64_t src = -1;
    if (r < r1) {
      const int i0 = inc_ptr[r], n = inc_ptr[r + 1] - i0;
      if (n > 0) src = i0 + (p < n ? p : n - 1);
    }
    if constexpr (kPack)
      ell[e0 + s] = src < 0 ? Out{0u, 0u} : p1q_pack(rec[src], eb, span, nbad);
    else
      ell[e0 + s] = rec[src < 0 ? 0 : src];
    if (ell_lv) ell_lv[e0 + s] = inc_lv[src < 0 ? 0 : src];
  }
  if (nbad) atomicAdd(bad, nbad);
}

// Vertex patches of a RecP1p / RecP2c system (vertex rows [0, nvrows), P2
// edge rows after them); returns the patch size of every patch block (0: no
// patch).  A block gets one when its distinct vertices fit the index form
// and its staged entries and coordinates fit the 64 KiB an assembly block
// stages (MGFE_PATCH=2: patched records, no patches; MGFE_PATCH_MAX: a lower
// vertex limit, which mixes blocks with and without patches on small meshes).
template <typename Rec>
static std::vector<int32_t> build_patches(Ctx *c, System &S, const std::vector<int32_t> &rp,
                                          int64_t nvrows) {
  const int nvb = (int)((nvrows + kPatchRows - 1) / kPatchRows);
  const int64_t nb = nvb + (S.nrows - nvrows + kPatchRows - 1) / kPatchRows;
  std::vector<int32_t> hc(nb, 0), pp(nb + 1, 0);
  const bool on = env_int("MGFE_PATCH", 1) == 1;
  const int lim = std::is_same<Rec, RecP1p>::value ? kPatchMax : kPatchMaxP2;
  const int umax = std::min(lim, std::max(1, env_int("MGFE_PATCH_MAX", lim)));
  if (on && nb > 0) {
    int32_t *cnt = dalloc<int32_t>(c, nb);
    k_patch<false, Rec><<<(unsigned)nb, 256, 0, c->stream>>>(nvrows, nvb, S.nrows, S.inc_ptr,
                                                             (Rec *)S.inc_rec, c->tets, nullptr,
                                                             cnt, nullptr, nullptr);
    MG_HIP(hipGetLastError());
    MG_HIP(hipMemcpyAsync(hc.data(), cnt, sizeof(int32_t) * nb, hipMemcpyDeviceToHost,
                          c->stream));
    MG_HIP(hipStreamSynchronize(c->stream));
    dfree(c, cnt);
  }
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t r0 = b < nvb ? b * kPatchRows : nvrows + (b - nvb) * kPatchRows;
    const int64_t r1 = std::min<int64_t>(r0 + kPatchRows, b < nvb ? nvrows : S.nrows);
    const int ent = rp[r1] - rp[r0];
    const int u = hc[b];
    const bool ok = u > 0 && u <= umax && patch_lds(ent, u) <= 64 * 1024;
    hc[b] = ok ? u : 0;
    pp[b + 1] = pp[b] + hc[b];
  }
  S.patch_ptr = dalloc<int32_t>(c, nb + 1);
  S.patch_vid = dalloc<int32_t>(c, std::max<int64_t>(pp[nb], 1));
  S.row_lv = dalloc<uint32_t>(c, S.nrows);
  MG_HIP(hipMemcpyAsync(S.patch_ptr, pp.data(), sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice,
                        c->stream));
  MG_HIP(hipMemsetAsync(S.row_lv, 0, sizeof(uint32_t) * S.nrows, c->stream));
  if (pp[nb] > 0) {
    if (!std::is_same<Rec, RecP1p>::value) {
      S.inc_lv = dalloc<uint32_t>(c, S.n_inc);
      MG_HIP(hipMemsetAsync(S.inc_lv, 0, sizeof(uint32_t) * S.n_inc, c->stream));
    }
    k_patch<true, Rec><<<(unsigned)nb, 256, 0, c->stream>>>(nvrows, nvb, S.nrows, S.inc_ptr,
                                                            (Rec *)S.inc_rec, c->tets, S.patch_ptr,
                                                            S.patch_vid, S.row_lv, S.inc_lv);
    MG_HIP(hipGetLastError());
  }
  MG_HIP(hipStreamSynchronize(c->stream));
  return hc;
}


// Padded patch vertex ids (System::vid_pad): one workgroup per patch block.
__global__ __launch_bounds__(256) void k_vid_pad(int nvb, int nb, const int32_t *__restrict__ pptr,
                                                 const int32_t *__restrict__ vid,
                                                 int64_t base_v, int stride_v, int64_t base_e,
                                                 int stride_e, int32_t *__restrict__ out) {
  const int b = blockIdx.x;
  if (b >= nb) return;
  const bool vx = b < nvb;
  const int stride = vx ? stride_v : stride_e;
  const int64_t o = (vx ? base_v : base_e) + (int64_t)(vx ? b : b - nvb) * stride;
  const int p0 = pptr[b], np = pptr[b + 1] - p0;
  for (int p = threadIdx.x; p < stride; p += blockDim.x)
    out[o + p] = np > 0 ? vid[p0 + (p < np ? p : 0)] : 0;
}

// The ELL copy of a patched system's records (System::ell_*), laid out for
// the lanes per row of its FP64 vertex-row and edge-row segments; P1 records
// in the 8-byte RecP1q form where they fit (MGFE_P1Q=0: RecP1p).
static void build_ell(Ctx *c, System &S, int64_t nvrows, const std::vector<int32_t> &psz) {
  const size_t nb = psz.size();
  int lpr[2] = {0, 0};
  for (const RowSegment &sg : S.segs) {
    if (sg.gacc) continue;  // accumulated in place: not patched
    int &l = lpr[sg.lrow == 0 ? 0 : 1];
    if (l == 0) l = sg.lpr;
    MG_CHECK(l == sg.lpr, MG_ERR_STATE, "ell: segments of one kind with different lanes per row");
  }
  std::vector<int32_t> ip(S.nrows + 1);
  MG_HIP(hipMemcpyAsync(ip.data(), S.inc_ptr, sizeof(int32_t) * (S.nrows + 1),
                        hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  const int nvb = (int)((nvrows + kPatchRows - 1) / kPatchRows);
  int steps[2] = {1, 1};
  for (size_t b = 0; b < nb; ++b) {
    const int g = (int)b < nvb ? 0 : 1;
    const int l = std::max(1, lpr[g]);
    const int64_t r0 = g == 0 ? (int64_t)b * kPatchRows : nvrows + (int64_t)(b - nvb) * kPatchRows;
    const int64_t r1 = std::min<int64_t>(r0 + kPatchRows, g == 0 ? nvrows : S.nrows);
    for (int64_t r = r0; r < r1; ++r) steps[g] = std::max(steps[g], (ip[r + 1] - ip[r] + l - 1) / l);
  }
  const int64_t nbv = std::min<int64_t>(nvb, (int64_t)nb), nbe = (int64_t)nb - nbv;
  const int64_t sv = nbv * steps[0] * kPatchRows * std::max(1, lpr[0]);
  const int64_t se = nbe * steps[1] * kPatchRows * std::max(1, lpr[1]);
  MG_CHECK(sv + se < (1ll << 31), MG_ERR_SIZE, "ell: more than 2^31 slots");
  const int64_t ns = std::max<int64_t>(sv + se, 1);
  // 8-byte P1 records: rows of <= 31 entries, patches of <= 512 vertices
  // (the element offsets are checked as the copy is filled, against 2^22 or
  // the smaller MGFE_P1Q_SPAN bits, which tests use to force the fallback)
  int max_patch = 0;
  for (int32_t n : psz) max_patch = std::max(max_patch, n);
  if (S.rec_kind == kRecP1p && S.max_row_nnz <= 31 && max_patch <= 512 &&
      env_int("MGFE_P1Q", 1) != 0 && nb > 0) {
    S.ell_rec = dalloc<RecP1q>(c, ns);
    S.ell_ebase = dalloc<int32_t>(c, nb);
    int *bad = dalloc<int>(c, 1);
    MG_HIP(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    k_ell_fill<RecP1p, RecP1q><<<(unsigned)nb, 256, 0, c->stream>>>(
        nvrows, nvb, S.nrows, S.inc_ptr, (const RecP1p *)S.inc_rec, nullptr, 0, steps[0],
        std::max(1, lpr[0]), sv, steps[1], std::max(1, lpr[1]), (RecP1q *)S.ell_rec, nullptr,
        S.ell_ebase, std::min(22, std::max(0, env_int("MGFE_P1Q_SPAN", 22))), bad);
    MG_HIP(hipGetLastError());
    const int nbad = read_back(c, bad);
    dfree(c, bad);
    S.ell_q = nbad == 0;
    if (!S.ell_q) {
      dfree(c, S.ell_rec);
      dfree(c, S.ell_ebase);
      S.ell_rec = nullptr;
      S.ell_ebase = nullptr;
    }
  }
  if (!S.ell_q) {
    S.ell_rec = dalloc<char>(c, (size_t)S.rec_words * 4 * ns);
    if (S.inc_lv) S.ell_lv = dalloc<uint32_t>(c, ns);
    if (nb > 0) {
      if (S.rec_kind == kRecP1p)
        k_ell_fill<RecP1p, RecP1p><<<(unsigned)nb, 256, 0, c->stream>>>(
            nvrows, nvb, S.nrows, S.inc_ptr, (const RecP1p *)S.inc_rec, S.inc_lv, 0, steps[0],
            std::max(1, lpr[0]), sv, steps[1], std::max(1, lpr[1]), (RecP1p *)S.ell_rec,
            S.ell_lv, nullptr, 0, nullptr);
      else
        k_ell_fill<RecP2c, RecP2c><<<(unsigned)nb, 256, 0, c->stream>>>(
            nvrows, nvb, S.nrows, S.inc_ptr, (const RecP2c *)S.inc_rec, S.inc_lv, 0, steps[0],
            std::max(1, lpr[0]), sv, steps[1], std::max(1, lpr[1]), (RecP2c *)S.ell_rec,
            S.ell_lv, nullptr, 0, nullptr);
      MG_HIP(hipGetLastError());
    }
  }
  MG_HIP(hipStreamSynchronize(c->stream));
  S.ell_base[0] = 0;
  S.ell_base[1] = sv;
  for (int g = 0; g < 2; ++g) {
    S.ell_steps[g] = steps[g];
    S.ell_lpr[g] = lpr[g];
  }
  // padded patch vertex ids: per region, the largest number of vertices a
  // launch of it stages
  int stride[2] = {1, 1};
  for (const RowSegment &sg : S.segs)
    if (sg.patch > 0) stride[sg.lrow == 0 ? 0 : 1] = std::max(stride[sg.lrow == 0 ? 0 : 1], sg.patch);
  const int64_t pv = nbv * stride[0], pe = nbe * stride[1];
  S.vid_pad = dalloc<int32_t>(c, std::max<int64_t>(pv + pe, 1));
  if (nb > 0) {
    k_vid_pad<<<(unsigned)nb, 256, 0, c->stream>>>(nvb, (int)nb, S.patch_ptr, S.patch_vid, 0,
                                                   stride[0], pv, stride[1], S.vid_pad);
    MG_HIP(hipGetLastError());
  }
  MG_HIP(hipStreamSynchronize(c->stream));
  S.vpad_base[0] = 0;
  S.vpad_base[1] = pv;
  S.vpad_stride[0] = stride[0];
  S.vpad_stride[1] = stride[1];
}

// LDS budget of a staged row block: its entries plus the 256-slot region
// that takes eliminated columns, in FP64, within 64 KiB
static bool block_fits(int entries) { return (entries + 256) * 8 <= 64 * 1024; }

// psz: the vertex patch size of every patch block (patched records: vertex
// rows from row 0, then P2 edge rows from nvrows, in kPatchRows-row blocks),
// nullptr otherwise
static void make_segments(System &S, const std::vector<int32_t> &rp, int64_t nvrows, int nl,
                          const std::vector<int32_t> *psz) {
  S.segs.clear();
  auto push = [&](int64_t a, int64_t b, int lrow, int rows, int lpr, int lpr32, bool gacc) {
    if (b <= a) return;
    int mx = 0;
    for (int64_t r = a; r < b; r += rows) mx = std::max(mx, rp[std::min<int64_t>(r + rows, b)] - rp[r]);
    S.segs.push_back(RowSegment{(int32_t)a, (int32_t)b, gacc ? 0 : mx, lrow, rows,
                                gacc ? 1 : lpr, gacc ? 1 : lpr32, gacc ? 1 : 0, 0, 0});
  };
  auto add = [&](int64_t a, int64_t b, int lrow, int rows, int lpr, int lpr32) {
    if (b <= a) return;
    // halve the rows per block until every block fits the LDS budget (64
    // rows at the least); 64-row blocks that still do not fit go to
    // segments that accumulate in global memory instead
    for (;;) {
      int mx = 0;
      for (int64_t r = a; r < b; r += rows)
        mx = std::max(mx, rp[std::min<int64_t>(r + rows, b)] - rp[r]);
      if (block_fits(mx) || rows == 64) break;
      rows /= 2;
    }
    int64_t run = a;
    bool run_g = !block_fits(rp[std::min<int64_t>(a + rows, b)] - rp[a]);
    for (int64_t r = a; r < b; r += rows) {
      const bool g = !block_fits(rp[std::min<int64_t>(r + rows, b)] - rp[r]);
      if (g != run_g) {
        push(run, r, lrow, rows, lpr, lpr32, run_g);
        run = r;
        run_g = g;
      }
    }
    push(run, b, lrow, rows, lpr, lpr32, run_g);
  };
  auto pick = [](int v, int a, int b, int c) { return v == a || v == b || v == c; };
  // launch shapes (MGFE_VERTEX_LPR 1/2/4, MGFE_EDGE_ROWS 64/128/256,
  // MGFE_EDGE_LPR 1/2) only change speed, never the result; defaults from a
  // sweep at 64^3 P2 (edge rows: 128 x 2 lanes 0.283-0.288 ms per FP64
  // assembly, 256 x 2 0.292-0.305, 64 x 2 0.292; 4 lanes per edge row 0.39+;
  // with whole-record loads, one lane per row: 128 x 1 0.284-0.287, 256 x 1
  // 0.278-0.280, 64 x 1 0.292-0.294 against 0.273-0.277 for 128 x 2)
  // FP32 P2 vertex rows: 2 lanes per row (64 x 2: 166.9 us per 64^3 P2 FP32
  // assembly, 64 x 4: 171.4-172.3, 64 x 1: 189; FP64 64 x 2 is 3% slower);
  // FP32 P1 vertex rows: 4 (128^3 P1: 404 us against 473 with 2)
  const int vl = env_int("MGFE_VERTEX_LPR", 4),
            vl32 = env_int("MGFE_VERTEX_LPR32", nl == 10 ? 2 : 4),
            er = env_int("MGFE_EDGE_ROWS", 128), el = env_int("MGFE_EDGE_LPR", 2);
  const int vlpr = pick(vl, 1, 2, 4) ? vl : 4;
  const int vlpr32 = pick(vl32, 1, 2, 4) ? vl32 : (nl == 10 ? 2 : 4);
  const int erows = pick(er, 64, 128, 256) ? er : 128;
  const int elpr = (el == 1 || el == 2) ? el : 2;
  if (psz) {
    // kPatchRows-row blocks, segments split where a block's kind changes
    // (staged: 0, accumulated in place: 1, staged with a patch: 2) and
    // where a patched segment's largest entries and patch would together
    // pass the 64 KiB LDS budget
    auto region = [&](int64_t a, int64_t b, int lrow, int pb0, int lpr, int lpr32) {
      auto kind = [&](int64_t r) {
        if (!block_fits(rp[std::min<int64_t>(r + kPatchRows, b)] - rp[r])) return 1;
        return (*psz)[pb0 + (r - a) / kPatchRows] > 0 ? 2 : 0;
      };
      int64_t run = a;
      int rk = -1, me = 0, mp = 0;
      auto flush = [&](int64_t x) {
        if (x <= run) return;
        push(run, x, lrow, kPatchRows, lpr, lpr32, rk == 1);
        if (rk == 2) {
          // a P2 edge-row launch stages the patches of all but the largest
          // 0.5% of its blocks in full (the others read their remaining
          // vertices from memory), so one outlier does not size every
          // block's LDS; MGFE_PATCH_STAGE > 0: stage at most that many
          // vertices.  Vertex-row launches stage every patch in full.
          const int q0 = pb0 + (int)((run - a) / kPatchRows);
          const int q1 = pb0 + (int)((x - a + kPatchRows - 1) / kPatchRows);
          int slots = mp;
          if (lrow == 4) {
            std::vector<int32_t> u(psz->begin() + q0, psz->begin() + q1);
            std::sort(u.begin(), u.end());
            const size_t k = std::min(u.size() - 1, (size_t)(0.995 * (double)u.size()));
            slots = std::max(1, u[k]);
            const int cap = env_int("MGFE_PATCH_STAGE", 0);
            if (cap > 0) slots = std::min(slots, cap);
          }
          S.segs.back().patch = slots;
          S.segs.back().pb0 = q0;
        }
        run = x;
      };
      for (int64_t r = a; r < b; r += kPatchRows) {
        const int k = kind(r);
        const int ent = rp[std::min<int64_t>(r + kPatchRows, b)] - rp[r];
        const int u = (*psz)[pb0 + (r - a) / kPatchRows];
        if (k != rk || (k == 2 && patch_lds(std::max(me, ent), std::max(mp, u)) > 64 * 1024)) {
          flush(r);
          rk = k;
          me = mp = 0;
        }
        me = std::max(me, ent);
        mp = std::max(mp, u);
      }
      flush(b);
    };
    const int nvb = (int)((nvrows + kPatchRows - 1) / kPatchRows);
    region(0, nvrows, 0, 0, vlpr, vlpr32);
    region(nvrows, S.nrows, 4, nvb, elpr, elpr);
    return;
  }
  add(0, nvrows, 0, 64, vlpr, vlpr32);
  add(nvrows, S.nrows, 4, erows, elpr, elpr);
}

// The (row, column) pairs of every element contribution sorted and made
// unique: slot[c] = CSR position of contribution c (-1: eliminated); with
// cols_out, a new array of the column of every position (8 zeroed entries of
// padding after it: the SpMV streams read a lane's 4 entries at once), and
// with cnt the entries per row.  Returns nnz.
static int64_t pattern_pass(Ctx *c, int sys, int32_t *slot, int32_t **cols_out, int32_t *cnt) {
  const int nl = c->nloc;
  const int32_t *map = sys == MG_SYS_FULL ? nullptr : c->sysmap;
  const int64_t nrows = sys == MG_SYS_FULL ? c->ndofs : c->nfree;
  const int64_t ncon = c->nT * nl * nl;
  const int eb = bitlen((uint64_t)nrows * (uint64_t)nrows) + 1;
  const uint64_t invalid = (eb >= 64) ? ~0ull : ((1ull << eb) - 1);
  uint64_t *k0 = dalloc<uint64_t>(c, ncon), *k1 = dalloc<uint64_t>(c, ncon);
  int32_t *v0 = dalloc<int32_t>(c, ncon), *v1 = dalloc<int32_t>(c, ncon);
  k_pattern_keys<<<blocks_for(ncon, 256), 256, 0, c->stream>>>(c->nT, nl, c->dofs, map,
                                                               (uint64_t)nrows, invalid, k0, v0);
  MG_HIP(hipGetLastError());
  sort_pairs(c, k0, k1, v0, v1, ncon, eb);
  dfree(c, k0);
  dfree(c, v0);
  int32_t *head = dalloc<int32_t>(c, ncon), *id = dalloc<int32_t>(c, ncon);
  k_heads_u64<<<blocks_for(ncon, 256), 256, 0, c->stream>>>(ncon, k1, invalid, head);
  inclusive_sum(c, head, id, ncon);
  const int64_t nnz = read_back(c, id + ncon - 1);
  int32_t *cols = dalloc<int32_t>(c, nnz + 8);
  MG_HIP(hipMemsetAsync(cols + nnz, 0, 8 * sizeof(int32_t), c->stream));
  int32_t *tcnt = cnt ? nullptr : dalloc<int32_t>(c, nrows + 1);
  if (!cnt) MG_HIP(hipMemsetAsync(tcnt, 0, sizeof(int32_t) * (nrows + 1), c->stream));
  k_pattern_scatter<<<blocks_for(ncon, 256), 256, 0, c->stream>>>(
      ncon, k1, v1, id, head, (uint64_t)nrows, invalid, slot, cols, cnt ? cnt : tcnt);
  MG_HIP(hipGetLastError());
  MG_HIP(hipStreamSynchronize(c->stream));
  for (void *p : {(void *)k1, (void *)v1, (void *)head, (void *)id, (void *)tcnt}) dfree(c, p);
  if (cols_out)
    *cols_out = cols;
  else
    dfree(c, cols);
  return nnz;
}

// The element -> CSR position map of a built system (nT nloc^2 entries,
// contribution c = (e nloc + i) nloc + j for field basis i, test basis j).
// Recomputed on request: the assembly does not read it, so it is not kept
// on the device between calls.
void slot_map(Ctx *c, int sys, int32_t *slot) {
  const int64_t nnz = pattern_pass(c, sys, slot, nullptr, nullptr);
  MG_CHECK(nnz == c->sys[sys].nnz, MG_ERR_STATE, "slot map: pattern changed");
}

void build_pattern(Ctx *c, int sys) {
  System &S = c->sys[sys];
  free_system(c, S);
  const int nl = c->nloc;
  const int32_t *map = sys == MG_SYS_FULL ? nullptr : c->sysmap;
  const int64_t nrows = sys == MG_SYS_FULL ? c->ndofs : c->nfree;
  S.nrows = nrows;
  MG_CHECK(nrows > 0, MG_ERR_STATE, "pattern: no rows");
  const int64_t ncon = c->nT * nl * nl;
  MG_CHECK(ncon < (1ll << 31), MG_ERR_SIZE, "pattern: more than 2^31 element contributions");
  // the slot map lives only while the incidence records are built
  int32_t *slot = dalloc<int32_t>(c, ncon);
  int32_t *cnt = dalloc<int32_t>(c, nrows + 1);
  MG_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nrows + 1), c->stream));
  S.nnz = pattern_pass(c, sys, slot, &S.colidx, cnt);
  S.rowptr = dalloc<int32_t>(c, nrows + 1);
  exclusive_sum(c, cnt, S.rowptr, nrows + 1);
  S.diagpos = dalloc<int32_t>(c, nrows);
  k_diagpos<<<blocks_for(nrows, 256), 256, 0, c->stream>>>(nrows, S.rowptr, S.colidx, S.diagpos);
  MG_HIP(hipGetLastError());

  // row incidence: (element, local basis) pairs grouped by row, ordered by
  // (local basis, element) inside a row
  const int64_t npair = c->nT * nl;
  const int eb2 = bitlen((uint64_t)nrows * nl) + 1;
  const uint64_t inv2 = (1ull << eb2) - 1;
  uint64_t *k0 = dalloc<uint64_t>(c, npair), *k1 = dalloc<uint64_t>(c, npair);
  int32_t *v0 = dalloc<int32_t>(c, npair), *v1 = dalloc<int32_t>(c, npair);
  k_inc_keys<<<blocks_for(npair, 256), 256, 0, c->stream>>>(c->nT, nl, c->dofs, map, inv2, k0, v0);
  MG_HIP(hipGetLastError());
  sort_pairs(c, k0, k1, v0, v1, npair, eb2);
  MG_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nrows + 1), c->stream));
  k_inc_count<<<blocks_for(npair, 256), 256, 0, c->stream>>>(npair, k1, inv2, nl, cnt);
  S.inc_ptr = dalloc<int32_t>(c, nrows + 1);
  exclusive_sum(c, cnt, S.inc_ptr, nrows + 1);
  S.n_inc = read_back(c, S.inc_ptr + nrows);
  int *bad = dalloc<int>(c, 1);
  MG_HIP(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
  {
    int perm[10][4], sigma[10][10];
    relabel_tables(perm, sigma);
    MG_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_sigma), sigma, sizeof(sigma), 0,
                                  hipMemcpyHostToDevice, c->stream));
    MG_HIP(hipStreamSynchronize(c->stream));
  }
  std::vector<int32_t> rp(nrows + 1);
  MG_HIP(hipMemcpyAsync(rp.data(), S.rowptr, sizeof(int32_t) * (nrows + 1), hipMemcpyDeviceToHost,
                        c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  int maxnnz = 0;
  for (int64_t r = 0; r < nrows; ++r) maxnnz = std::max(maxnnz, rp[r + 1] - rp[r]);
  S.max_row_nnz = maxnnz;
  // record form by the longest row: P2 12-byte records with 7-bit positions
  // up to 127 entries (MGFE_REC16=1 keeps the 16-byte form), uint8
  // positions up to 254, uint16 positions beyond (up to 65534)
  MG_CHECK(maxnnz <= 65534, MG_ERR_SIZE, "pattern: a row has more than 65534 entries");
  const bool compact = nl == 10 && maxnnz <= 127 && env_int("MGFE_REC16", 0) == 0;
  // patched records: P1 rows up to 254 entries, P2 rows up to 127
  // (MGFE_PATCH=0: the forms without patches)
  const bool patched = (nl == 4 ? maxnnz <= 254 : compact) && env_int("MGFE_PATCH", 1) != 0;
  const bool wide = maxnnz > 254;
  const int gi = blocks_for(S.n_inc, 256);
  if (compact) {
    S.inc_rec = dalloc<RecP2c>(c, S.n_inc);
    S.rec_words = 3;
    S.rec_kind = kRecP2c;
    k_inc_records_c<<<gi, 256, 0, c->stream>>>(S.n_inc, k1, v1, slot, S.rowptr, S.diagpos,
                                               (RecP2c *)S.inc_rec, bad);
  } else if (nl == 10 && !wide) {
    S.inc_rec = dalloc<RecP2>(c, S.n_inc);
    S.rec_words = 4;
    S.rec_kind = kRecP2;
    k_inc_records<10, RecP2><<<gi, 256, 0, c->stream>>>(S.n_inc, k1, v1, slot, S.rowptr,
                                                        (RecP2 *)S.inc_rec, bad);
  } else if (nl == 10) {
    S.inc_rec = dalloc<RecP2w>(c, S.n_inc);
    S.rec_words = 6;
    S.rec_kind = kRecP2w;
    k_inc_records<10, RecP2w><<<gi, 256, 0, c->stream>>>(S.n_inc, k1, v1, slot, S.rowptr,
                                                         (RecP2w *)S.inc_rec, bad);
  } else if (patched) {
    S.inc_rec = dalloc<RecP1p>(c, S.n_inc);
    S.rec_words = 3;
    S.rec_kind = kRecP1p;
    k_inc_records_p1p<<<gi, 256, 0, c->stream>>>(S.n_inc, k1, v1, slot, S.rowptr, S.diagpos,
                                                 (RecP1p *)S.inc_rec, bad);
  } else if (!wide) {
    S.inc_rec = dalloc<RecP1>(c, S.n_inc);
    S.rec_words = 2;
    S.rec_kind = kRecP1;
    k_inc_records<4, RecP1><<<gi, 256, 0, c->stream>>>(S.n_inc, k1, v1, slot, S.rowptr,
                                                       (RecP1 *)S.inc_rec, bad);
  } else {
    S.inc_rec = dalloc<RecP1w>(c, S.n_inc);
    S.rec_words = 3;
    S.rec_kind = kRecP1w;
    k_inc_records<4, RecP1w><<<gi, 256, 0, c->stream>>>(S.n_inc, k1, v1, slot, S.rowptr,
                                                        (RecP1w *)S.inc_rec, bad);
  }
  MG_HIP(hipGetLastError());
  int hb = read_back(c, bad);
  for (void *p : {(void *)k0, (void *)k1, (void *)v0, (void *)v1, (void *)cnt, (void *)bad,
                  (void *)slot})
    dfree(c, p);
  MG_CHECK(hb == 0, MG_ERR_STATE, "pattern: a record position does not fit its form");
  const int64_t nvrows =
      nl == 4 ? nrows : (sys == MG_SYS_FULL ? c->nV : c->nfree_vertex);
  S.nvrows = nvrows;
  if (S.rec_kind == kRecP1p || (S.rec_kind == kRecP2c && patched)) {
    const std::vector<int32_t> psz = S.rec_kind == kRecP1p
                                         ? build_patches<RecP1p>(c, S, rp, nvrows)
                                         : build_patches<RecP2c>(c, S, rp, nvrows);
    make_segments(S, rp, nvrows, nl, &psz);
    // P2 edge-dof rows of <= 31 entries whose patches have <= 511 vertices:
    // their records take the RecP2e form, patch words included (12 bytes per
    // incidence instead of 16; MGFE_EDGE5=0 keeps the separate patch words)
    if (S.rec_kind == kRecP2c && S.inc_lv && nvrows < nrows && env_int("MGFE_EDGE5", 1) != 0) {
      int max_edge = 0, max_patch = 0;
      for (int64_t r = nvrows; r < nrows; ++r) max_edge = std::max(max_edge, rp[r + 1] - rp[r]);
      const size_t nvb = (size_t)((nvrows + kPatchRows - 1) / kPatchRows);
      for (size_t b = nvb; b < psz.size(); ++b) max_patch = std::max(max_patch, psz[b]);
      if (max_edge <= 31 && max_patch <= 511) {
        int32_t i0 = 0;
        MG_HIP(hipMemcpyAsync(&i0, S.inc_ptr + nvrows, sizeof(int32_t), hipMemcpyDeviceToHost,
                              c->stream));
        MG_HIP(hipStreamSynchronize(c->stream));
        int *bad5 = dalloc<int>(c, 1);
        MG_HIP(hipMemsetAsync(bad5, 0, sizeof(int), c->stream));
        if (S.n_inc > i0)
          k_edge5<<<blocks_for(S.n_inc - i0, 256), 256, 0, c->stream>>>(
              i0, S.n_inc, (RecP2c *)S.inc_rec, S.inc_lv, bad5);
        MG_HIP(hipGetLastError());
        const int nb5 = read_back(c, bad5);
        dfree(c, bad5);
        MG_CHECK(nb5 == 0, MG_ERR_STATE, "pattern: an edge-row record does not fit the 5-bit form");
        S.edge5 = true;
        S.edge_inc0 = i0;
      }
    }
    // packed row words (MGFE_PROW=0: the separate row arrays)
    if (env_int("MGFE_PROW", 1) != 0 && nrows > 0) {
      S.prow = dalloc<uint64_t>(c, nrows);
      int *bad = dalloc<int>(c, 1);
      MG_HIP(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
      k_prow<<<blocks_for(nrows, 256), 256, 0, c->stream>>>(nvrows, nrows, S.rowptr, S.inc_ptr,
                                                            S.diagpos, S.row_lv, S.prow, bad);
      MG_HIP(hipGetLastError());
      const int nbad = read_back(c, bad);
      dfree(c, bad);
      if (nbad) {
        dfree(c, S.prow);
        S.prow = nullptr;
      }
    }
    // ELL records and padded patch ids: P1 systems (128^3 P1 FP64 assembly
    // 310 -> 300 us with 12-byte ELL records, -> 289 us with the 8-byte
    // form); P2 rows measured no faster with them (MGFE_ELL=2: all patched
    // systems, 0: none)
    const int ell = env_int("MGFE_ELL", 1);
    if (ell == 2 || (ell == 1 && nl == 4)) build_ell(c, S, nvrows, psz);
  } else {
    make_segments(S, rp, nvrows, nl, nullptr);
  }
  {
    // SpMV row blocks
    std::vector<int32_t> sb;
    int64_t r = 0;
    sb.push_back(0);
    // very short rows (mean <= kSpmvNnz / 128 entries): 128-row blocks summed
    // by 2 lanes per row, so a block still carries ~kSpmvNnz entries
    // (MGFE_SPMV_ROWS=64|128 overrides)
    const int64_t mean = nrows > 0 ? S.nnz / nrows : 0;
    const int want = env_int("MGFE_SPMV_ROWS", mean <= kSpmvNnz / (2 * kSpmvRows) ? 128 : 64);
    const int rows_cap = want == 128 ? 2 * kSpmvRows : kSpmvRows;
    S.srl = rows_cap == kSpmvRows ? 4 : 2;
    // blocks of consecutive rows; a row longer than a block stands alone in
    // the long-row list, and the blocks around it end before it and start
    // after it
    std::vector<int32_t> lr;
    std::vector<std::pair<int32_t, int32_t>> br;
    while (r < nrows) {
      if (rp[r + 1] - rp[r] > kSpmvNnz) {
        lr.push_back((int32_t)r);
        ++r;
        continue;
      }
      int64_t r1 = r + 1;
      while (r1 < nrows && r1 - r < rows_cap && rp[r1 + 1] - rp[r] <= kSpmvNnz) ++r1;
      br.push_back({(int32_t)r, (int32_t)r1});
      r = r1;
    }
    S.n_sblk = (int32_t)br.size();
    std::vector<BlkDesc> d(S.n_sblk);
    for (int32_t b = 0; b < S.n_sblk; ++b) {
      d[b] = BlkDesc{};
      d[b].r0 = br[b].first;
      d[b].r1 = br[b].second;
      d[b].s0 = rp[br[b].first];
      d[b].ns = rp[br[b].second] - rp[br[b].first];
    }
    S.n_long = (int32_t)lr.size();
    if (S.n_long) {
      S.long_rows = dalloc<int32_t>(c, lr.size());
      MG_HIP(hipMemcpyAsync(S.long_rows, lr.data(), sizeof(int32_t) * lr.size(),
                            hipMemcpyHostToDevice, c->stream));
    }
    S.sdesc = dalloc<BlkDesc>(c, d.size());
    MG_HIP(hipMemcpyAsync(S.sdesc, d.data(), sizeof(BlkDesc) * d.size(), hipMemcpyHostToDevice,
                          c->stream));
    MG_HIP(hipStreamSynchronize(c->stream));
  }
  // 16-bit columns for the SpMV streams (MGFE_SPMV_COL32=1 keeps int32)
  if (S.n_sblk > 0 && env_int("MGFE_SPMV_COL32", 0) == 0) {
    S.cofs = dalloc<uint16_t>(c, S.nnz + 8);
    // zeroed: the entries of long rows (in no block) are never written, and
    // the padding after the array is read past a block's end
    MG_HIP(hipMemsetAsync(S.cofs, 0, (S.nnz + 8) * sizeof(uint16_t), c->stream));
    int *fail = dalloc<int>(c, 1);
    MG_HIP(hipMemsetAsync(fail, 0, sizeof(int), c->stream));
    k_col_compress<<<S.n_sblk, 256, 0, c->stream>>>(S.sdesc, S.colidx, S.cofs, (int)S.nrows,
                                                    fail);
    MG_HIP(hipGetLastError());
    S.c16 = read_back(c, fail) == 0;
    dfree(c, fail);
    if (!S.c16) {
      dfree(c, S.cofs);
      S.cofs = nullptr;
    }
  }
  S.vals64 = dalloc<double>(c, S.nnz + 8);
  MG_HIP(hipMemsetAsync(S.vals64 + S.nnz, 0, 8 * sizeof(double), c->stream));
  S.vals32 = dalloc<float>(c, S.nnz + 8);
  MG_HIP(hipMemsetAsync(S.vals32 + S.nnz, 0, 8 * sizeof(float), c->stream));
  S.dinv64 = dalloc<double>(c, nrows);
  S.dinv32 = dalloc<float>(c, nrows);
  S.built = true;
}

}  // namespace mg
