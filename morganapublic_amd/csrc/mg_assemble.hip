// Numeric assembly of P1/P2 tetrahedral operators into a fixed CSR pattern,
// and the adjoint element kernel for d(loss)/d(per-element coefficient).
//
// A P2 assembly is three launches, a P1 assembly one.
// k_elem_factors: one lane per element computes, in FP64 from the vertex
// coordinates, the 6 edge factors S_ab = c vol grad(lambda_a).grad(lambda_b)
// (and c_mass det with a mass term), stored per element (P1 rows form them
// in place instead, from the row block's vertex patch in LDS where it has
// one).  k_asm_rows, once over the vertex-dof rows and
// once over the P2 edge-dof rows: a workgroup owns consecutive CSR rows and
// stages their value range in LDS; LPR lanes per row walk the row's
// (element, local basis l) incidence records, whose element is relabelled by
// an even vertex permutation so that the row's dof is local basis 0 (vertex)
// or 4 (edge).  The quadrature rules are symmetric, so the relabelled element
// matrix is the same matrix, and its row 0 / row 4 is a fixed linear map of
// the edge factors: the closed form of the barycentric integrals when the
// rule is exact on the integrand (row_values), otherwise the quadrature of
// integratorOEL3d_STD contracted once on the reference element (constant
// memory, compile-time indices).  The lanes of a row add their element rows
// in ordered phases, so every entry sums its contributions in incidence order
// (deterministic, no global atomics), and the workgroup writes its contiguous
// value range with coalesced nontemporal stores.
#include <cstring>
#include <mutex>
#include <type_traits>

#include "mg_internal.h"

// Static scratch of the allocation-free assembly (mg_assemble_noalloc): the
// per-element factor records live here instead of in context memory.  Every
// GPU the library runs on holds its own copy for the life of the process.
// Host code never uses this name as an address: the device address comes
// from hipGetSymbolAddress (noalloc_init) and is what the kernels receive.
static constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace mg {

// Called by mg_ctx_create with the context's device current: looks up that
// device's address of the scratch array once and keeps it per device.
void noalloc_init(Ctx *c) {
  static std::mutex mu;
  static void *addr[64] = {};
  std::lock_guard<std::mutex> lock(mu);
  if (c->device < 64 && addr[c->device]) {
    c->noalloc_scratch = addr[c->device];
    return;
  }
  void *p = nullptr;
  MG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
  MG_CHECK(p != nullptr, MG_ERR_HIP, "noalloc: no device address for the scratch array");
  if (c->device < 64) addr[c->device] = p;
  c->noalloc_scratch = p;
}

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

// (x, y) of a vertex as one 16-byte load at 8-byte alignment
typedef double dbl2_a8 __attribute__((ext_vector_type(2), aligned(8)));

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

// ---- per-element factors ------------------------------------------------------
// With n_j = det J grad(lambda_j) (n_1 = e2 x e3, n_2 = e3 x e1, n_3 = e1 x e2,
// n_0 = -(n_1 + n_2 + n_3), e_k = P_k - P_0; det J signed as
// geoMapInterface::getGradientDet, positive on a valid mesh), the record of an
// element holds, for its 6 edges (a, b) in kEdgeV order,
//   S_ab = k / (6 det) n_a . n_b  =  k vol grad(lambda_a) . grad(lambda_b)
// and, with a mass term, c_mass det.  Since sum_b grad(lambda_b) = 0 the
// diagonal follows as S_aa = -sum_{b != a} S_ab.  Every stiffness and mass row
// of the element is a fixed linear map of these 6 (7) numbers, so the row
// kernels read one small record per incidence instead of 4 vertex coordinates.
// Layout: records of 6 (stiffness) or 8 (with mass: 7 used) values; FP32
// stiffness records are 24 bytes (measured -9% for a 128^3 P1 FP32 assembly
// against 32-byte records).
template <typename T, bool MASS>
constexpr int geo_stride() {
  return MASS ? 8 : 6;
}

// The 6 edge factors (kEdgeV order) and c_mass det of one element from its
// vertex coordinates: the arithmetic of k_elem_factors, shared with the P1
// row kernels that form the factors in place.
template <typename T, bool MASS>
__device__ __forceinline__ void edge_factors(const double P[4][3], double ks, double km,
                                             T out[8]) {
  // explicit fma only: every kernel that forms factors (the factor kernel,
  // P1 rows, patched P2 rows, each launch shape) must round them alike,
  // whatever the compiler would fuse in its context
#pragma clang fp contract(off)
  double d[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) d[k][r] = P[k + 1][r] - P[0][r];
  // a b - c e with one rounding of the product c e
  auto x2 = [](double a, double b, double c, double e) { return fma(a, b, -(c * e)); };
  double n[4][3];
  n[1][0] = x2(d[1][1], d[2][2], d[1][2], d[2][1]);
  n[1][1] = x2(d[1][2], d[2][0], d[1][0], d[2][2]);
  n[1][2] = x2(d[1][0], d[2][1], d[1][1], d[2][0]);
  n[2][0] = x2(d[2][1], d[0][2], d[2][2], d[0][1]);
  n[2][1] = x2(d[2][2], d[0][0], d[2][0], d[0][2]);
  n[2][2] = x2(d[2][0], d[0][1], d[2][1], d[0][0]);
  n[3][0] = x2(d[0][1], d[1][2], d[0][2], d[1][1]);
  n[3][1] = x2(d[0][2], d[1][0], d[0][0], d[1][2]);
  n[3][2] = x2(d[0][0], d[1][1], d[0][1], d[1][0]);
#pragma unroll
  for (int r = 0; r < 3; ++r) n[0][r] = -(n[1][r] + n[2][r] + n[3][r]);
  const double det = fma(d[0][0], n[1][0], fma(d[0][1], n[1][1], d[0][2] * n[1][2]));
  const double sc = ks / (6.0 * det);
  constexpr int EA[6] = {0, 1, 2, 0, 1, 2}, EB[6] = {1, 2, 0, 3, 3, 3};  // kEdgeV
#pragma unroll
  for (int j = 0; j < 6; ++j)
    out[j] = (T)(sc * fma(n[EA[j]][0], n[EB[j]][0],
                          fma(n[EA[j]][1], n[EB[j]][1], n[EA[j]][2] * n[EB[j]][2])));
  out[6] = MASS ? (T)(km * det) : (T)0;
  out[7] = (T)0;
}

template <typename T, bool MASS>
__global__ __launch_bounds__(256) void k_elem_factors(
    int64_t nT, const int32_t *__restrict__ tets, const double *__restrict__ xyz,
    double ks_const, const double *__restrict__ ks_field, double km_const,
    const double *__restrict__ km_field, T *__restrict__ geo) {
  // lanes past the last element repeat it (their records are not stored), so
  // that every lane reaches the barrier below
  const int64_t e0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t ec = min<int64_t>(e0 + threadIdx.x, nT - 1);
  const int4 t = reinterpret_cast<const int4 *>(tets)[ec];
  const int id[4] = {t.x, t.y, t.z, t.w};
  // a vertex in two loads: (x, y), then z
  double P[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const dbl2_a8 q = *reinterpret_cast<const dbl2_a8 *>(xyz + 3 * (size_t)id[v]);
    P[v][0] = q.x;
    P[v][1] = q.y;
    P[v][2] = xyz[3 * (size_t)id[v] + 2];
  }
  const double ks = ks_field ? ks_field[ec] : ks_const;
  const double km = MASS ? (km_field ? km_field[ec] : km_const) : 0.0;
  constexpr int N = geo_stride<T, MASS>();
  T out[8];
  edge_factors<T, MASS>(P, ks, km, out);
  // the block's records are staged in LDS and leave as one contiguous run of
  // 16-byte stores (per-lane records 48 bytes apart: measured -25% for this
  // kernel, -2% for the FP64 assembly)
  __shared__ T stage[256 * 8];
#pragma unroll
  for (int k = 0; k < N; ++k) stage[threadIdx.x * N + k] = out[k];
  __syncthreads();
  const int nb = (int)min<int64_t>(256, nT - e0);
  if constexpr (sizeof(T) == 8) {
    double2 *dst = reinterpret_cast<double2 *>(geo + (size_t)e0 * N);
    const double2 *src = reinterpret_cast<const double2 *>(stage);
    for (int i = threadIdx.x; i < nb * N / 2; i += 256) dst[i] = src[i];
  } else if constexpr (N == 6) {
    float2 *dst = reinterpret_cast<float2 *>(geo + (size_t)e0 * N);
    const float2 *src = reinterpret_cast<const float2 *>(stage);
    for (int i = threadIdx.x; i < nb * 3; i += 256) dst[i] = src[i];
  } else {
    float4 *dst = reinterpret_cast<float4 *>(geo + (size_t)e0 * N);
    const float4 *src = reinterpret_cast<const float4 *>(stage);
    for (int i = threadIdx.x; i < nb * N / 4; i += 256) dst[i] = src[i];
  }
}

// What the P1 row kernels need to form element factors in place.
// The vertex patches of a patched segment: per patch block, its vertices
// (ptr, vid), each row's own vertex or endpoints (row_lv), the P2 patch words
// (inc_lv; P1 records hold their patch indices); slots is the number of
// vertices whose coordinates the segment stages per block.
struct PatchSrc {
  const int32_t *ptr;
  const int32_t *vid;
  const uint32_t *row_lv;
  const uint32_t *inc_lv;
  int slots;
  int pb0;  // patch block of the segment's first row
  // the ELL copy of the records (ell != 0: laid out for this launch's lanes
  // per row): a lane's step-k record of the segment's block q at
  // ell_rec[ell_base + (q * ell_steps + k) * NT + lane]
  const void *ell_rec;
  const uint32_t *ell_lv;
  int64_t ell_base;
  int ell_steps;
  int ell;
  // padded patch vertex ids (vid_pad non-null): the segment's block q stages
  // vid_pad[vpad_base + q * vpad_stride + p], p < slots
  const int32_t *vid_pad;
  int64_t vpad_base;
  int vpad_stride;
  // RecP1q records: the lowest element of each patch block
  const int32_t *ebase;
  // packed row words (System::prow), or null
  const uint64_t *prow;
};

struct ElemSrc {
  const int32_t *tets;
  const double *xyz;
  const double *ksf, *kmf;
  double ksc, kmc;
};

// The element of an incidence is relabelled by the even vertex permutation
// that moves the row's basis to local index 0 (vertex) or 4 (edge (0,1)); the
// relabelled edge j is original edge sigma_l(j).  EdgePack.w[l] holds
// sigma_l(j) in bits 3j..3j+2 (kernel arguments; selected with ALU ops, so no
// memory access depends on l).
// EdgePack.v4 holds, for P1 rows, the original vertex of relabelled vertex m
// of row basis l in bits 2 (4 m + l) .. 2 (4 m + l) + 1.
struct EdgePack {
  uint32_t w[10];
  uint32_t v4;
};

static EdgePack pack_edges() {
  int perm[10][4], sigma[10][10];
  relabel_tables(perm, sigma);
  EdgePack p{};
  for (int l = 0; l < 10; ++l)
    for (int j = 0; j < 6; ++j) p.w[l] |= (uint32_t)(sigma[l][4 + j] - 4) << (3 * j);
  for (int m = 0; m < 4; ++m)
    for (int l = 0; l < 4; ++l) p.v4 |= (uint32_t)sigma[l][m] << (2 * (4 * m + l));
  return p;
}

// value k of six held in registers (by value: a select over loads from a
// local array can be folded back into a runtime index, which puts the array
// in scratch memory)
// lane S of each group of four lanes (quad_perm DPP), for every lane of the group
template <int S>
__device__ __forceinline__ double quad_bcast(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, S * 0x55, 0xF, 0xF, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), S * 0x55, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <typename T>
__device__ __forceinline__ T pick6v(T s0, T s1, T s2, T s3, T s4, T s5, uint32_t k) {
  const T a = (k & 1u) ? s1 : s0;
  const T b = (k & 1u) ? s3 : s2;
  const T c = (k & 1u) ? s5 : s4;
  const T ab = (k & 2u) ? b : a;
  return (k & 4u) ? c : ab;
}

template <typename T>
__device__ __forceinline__ T pick6(const T *s, uint32_t k) {
  const T a = (k & 1u) ? s[1] : s[0];
  const T b = (k & 1u) ? s[3] : s[2];
  const T c = (k & 1u) ? s[5] : s[4];
  const T ab = (k & 2u) ? b : a;
  return (k & 4u) ? c : ab;
}

// An incidence record as the row kernels read it (whole dwords): el, the
// position of relabelled column m (pos), and the code of an eliminated column.
template <typename Rec, int LROW>
struct RecView;
template <int LROW>
struct RecView<RecP2, LROW> {
  static constexpr int kElim = 0xFF;
  int32_t el;
  uint32_t w[3];
  __device__ __forceinline__ int pos(int m) const { return (w[m >> 2] >> (8 * (m & 3))) & 0xFF; }
  __device__ __forceinline__ static RecView load(const RecP2 *__restrict__ rec, int j, int) {
    const int4 q = reinterpret_cast<const int4 *>(rec)[j];  // el, pos[10], pad[2]
    return RecView{q.x, {(uint32_t)q.y, (uint32_t)q.z, (uint32_t)q.w}};
  }
};
template <int LROW>
struct RecView<RecP1, LROW> {
  static constexpr int kElim = 0xFF;
  int32_t el;
  uint32_t w;
  __device__ __forceinline__ int pos(int m) const { return (w >> (8 * m)) & 0xFF; }
  __device__ __forceinline__ static RecView load(const RecP1 *__restrict__ rec, int j, int) {
    const int2 q = reinterpret_cast<const int2 *>(rec)[j];  // el, pos[4]
    return RecView{q.x, (uint32_t)q.y};
  }
};
// wide records: uint16 positions
template <int LROW>
struct RecView<RecP2w, LROW> {
  static constexpr int kElim = 0xFFFF;
  int32_t el;
  uint32_t w[5];
  __device__ __forceinline__ int pos(int m) const { return (w[m >> 1] >> (16 * (m & 1))) & 0xFFFF; }
  __device__ __forceinline__ static RecView load(const RecP2w *__restrict__ rec, int j, int) {
    const int2 *q = reinterpret_cast<const int2 *>(rec) + 3 * (size_t)j;  // 24 bytes, 8-aligned
    const int2 q0 = q[0], q1 = q[1], q2 = q[2];
    return RecView{q0.x, {(uint32_t)q0.y, (uint32_t)q1.x, (uint32_t)q1.y, (uint32_t)q2.x,
                          (uint32_t)q2.y}};
  }
};
template <int LROW>
struct RecView<RecP1w, LROW> {
  static constexpr int kElim = 0xFFFF;
  int32_t el;
  uint32_t w[2];
  __device__ __forceinline__ int pos(int m) const { return (w[m >> 1] >> (16 * (m & 1))) & 0xFFFF; }
  __device__ __forceinline__ static RecView load(const RecP1w *__restrict__ rec, int j, int) {
    const RecP1w q = rec[j];  // one dwordx3 load
    return RecView{q.el, {(uint32_t)q.pos[0] | ((uint32_t)q.pos[1] << 16),
                          (uint32_t)q.pos[2] | ((uint32_t)q.pos[3] << 16)}};
  }
};
// patched P1: own column (basis 0) = the diagonal, `self`; lv(o) = patch index
// of relabelled vertex o + 1
template <int LROW>
struct RecView<RecP1p, LROW> {
  static constexpr int kElim = 0xFF;
  int32_t el;
  uint32_t lo, hi;
  int self;
  __device__ __forceinline__ int pos(int m) const {
    return m == 0 ? self : (int)((lo >> (8 * (m - 1))) & 0xFFu);
  }
  __device__ __forceinline__ int lv(int o) const {
    return (int)(o == 0 ? ((lo >> 24) | ((hi & 0xFu) << 8)) : ((hi >> (o == 1 ? 4 : 16)) & 0xFFFu));
  }
  __device__ __forceinline__ static RecView load(const RecP1p *__restrict__ rec, int j, int self) {
    const RecP1p q = rec[j];  // one dwordx3 load
    RecView r;
    r.el = q.el;
    r.lo = q.lo;
    r.hi = q.hi;
    r.self = self;
    return r;
  }
};
// 8-byte patched P1 (ELL copies only): el = the element's offset from its
// block's lowest, times 16; own column (basis 0) = the diagonal, `self`
template <int LROW>
struct RecView<RecP1q, LROW> {
  static constexpr int kElim = 31;
  int32_t el;
  uint32_t lo, hi;
  int self;
  __device__ __forceinline__ int pos(int m) const {
    return m == 0 ? self : (int)((lo >> (5 * (m - 1))) & 31u);
  }
  __device__ __forceinline__ int lv(int o) const {
    return (int)(o == 0 ? (lo >> 15) & 511u : o == 1 ? ((lo >> 24) | (hi << 8)) & 511u : (hi >> 1) & 511u);
  }
  // one dwordx2 load, nontemporal: every slot of the ELL copy is read once,
  // and the L2 is left to the coordinates and coefficients that neighbouring
  // blocks share (128^3 P1 assembly: FP64 282 -> 269 us, FP32 255 -> 244)
  __device__ __forceinline__ static RecView load(const RecP1q *__restrict__ rec, int j, int self) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 q = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(rec) + j);
    RecView r;
    r.el = (int32_t)(q.y >> 10) << 4;
    r.lo = q.x;
    r.hi = q.y;
    r.self = self;
    return r;
  }
};
// compact P2: the row's own column (basis LROW) is its diagonal, `self`
template <int LROW>
struct RecP2cView {
  static constexpr int kElim = 0x7F;
  int32_t el;
  uint32_t lo, hi;
  int self;
  __device__ __forceinline__ int pos(int m) const {
    if (m == LROW) return self;
    const int k = m < LROW ? m : m - 1;  // packed index, compile-time
    const int b = 7 * k;
    const uint32_t v = b >= 32 ? (hi >> (b - 32)) : (b > 25 ? ((lo >> b) | (hi << (32 - b))) : (lo >> b));
    return (int)(v & 0x7Fu);
  }
};
// patched P2 edge rows: 5-bit positions, the row's own column (basis LROW)
// is its diagonal, the patch word is part of the record
template <int LROW>
struct RecView<RecP2e, LROW> {
  static constexpr int kElim = 31;
  int32_t el;
  uint32_t lo, hi;
  int self;
  __device__ __forceinline__ int pos(int m) const {
    if (m == LROW) return self;
    const int k = m < LROW ? m : m - 1;  // packed index, compile-time
    const int b = 5 * k;
    const uint32_t v = b >= 32 ? (hi >> (b - 32)) : (b > 27 ? ((lo >> b) | (hi << (32 - b))) : (lo >> b));
    return (int)(v & 31u);
  }
  // in the layout of a separate patch word: the 10-bit fields of relabelled
  // vertices 2 and 3, the orientation in bit 30
  __device__ __forceinline__ uint32_t patch_word() const {
    return ((hi >> 13) & 511u) | (((hi >> 22) & 511u) << 10) | ((hi >> 31) << 30);
  }
  __device__ __forceinline__ static RecView load(const RecP2e *__restrict__ rec, int j, int self) {
    const RecP2e q = rec[j];  // one dwordx3 load
    RecView r;
    r.el = q.el;
    r.lo = q.lo;
    r.hi = q.hi;
    r.self = self;
    return r;
  }
};
template <int LROW>
struct RecView<RecP2c, LROW> : RecP2cView<LROW> {
  __device__ __forceinline__ static RecView load(const RecP2c *__restrict__ rec, int j,
                                                 int self) {
    const RecP2c q = rec[j];  // one dwordx3 load
    RecView r;
    r.el = q.el;
    r.lo = q.lo;
    r.hi = q.hi;
    r.self = self;
    return r;
  }
};

// ---- closed form of the stiffness row for quadrature rules exact on it ------
// With D_pq = k vol grad(lambda_p) . grad(lambda_q) of the relabelled element,
// the stiffness entry is Q_lm(D), where Q integrates the products of the basis
// gradients with  int lambda_p = vol/4  and  int lambda_p lambda_q = vol (1 + delta_pq) / 20:
//   P1:                  Q_ij = D_ij
//   P2 vertex i, vertex j: D_ij ((4/5)(1 + delta_ij) - 1)
//   P2 vertex i, edge ab:  4 (D_ib E(i,a) + D_ia E(i,b)),  E(i,a) = (1 + delta_ia)/5 - 1/4
//   P2 edge ab, edge cd:   16 (D_bd F(a,c) + D_bc F(a,d) + D_ad F(b,c) + D_ac F(b,d)),
//                          F(p,q) = (1 + delta_pq)/20
// This is what any quadrature rule exact on the integrand (P1: all rules;
// P2: 4, 11, 15 points) returns, up to rounding.  Otherwise the row is
//   sum_k g_k tab[k][LROW][m],  g_(ab) = c det (J^{-1} J^{-T})_ab = 6 D_(a+1)(b+1),
// with the quadrature of integratorOEL3d_STD contracted once on the reference
// element (tab, constant memory, compile-time indices).
__host__ __device__ constexpr int edge_a(int j) {
  return j == 0 ? 0 : j == 1 ? 1 : j == 2 ? 2 : j == 3 ? 0 : j == 4 ? 1 : 2;
}
__host__ __device__ constexpr int edge_b(int j) {
  return j == 0 ? 1 : j == 1 ? 2 : j == 2 ? 0 : j == 3 ? 3 : j == 4 ? 3 : 3;
}
__host__ __device__ constexpr double dlt(int p, int q) { return p == q ? 1.0 : 0.0; }
__host__ __device__ constexpr double coefE(int i, int a) { return (1.0 + dlt(i, a)) / 5.0 - 0.25; }
__host__ __device__ constexpr double coefF(int p, int q) { return (1.0 + dlt(p, q)) / 20.0; }

// Relabelled edges a row needs: vertex rows (0,1), (2,0), (0,3); edge rows
// all but (2,3); the quadrature-table form all six.
template <int LROW, bool EXACT>
__host__ __device__ constexpr bool edge_needed(int j) {
  return !EXACT || (LROW == 0 ? (j == 0 || j == 2 || j == 3) : j != 5);
}

// Row LROW (0: vertex, 4: edge (0,1)) of the relabelled element matrix from
// the relabelled edge values e (kEdgeV order: e0 = D01, e1 = D12, e2 = D02,
// e3 = D03, e4 = D13, e5 = D23) and md = c_mass det.
template <int NL, int LROW, typename T, bool MASS, bool EXACT>
__device__ __forceinline__ void row_values(const T e[6], T md, T v[NL]) {
  // every kernel shape must round alike: explicit fma only, no contraction
  // chosen per instantiation
#pragma clang fp contract(off)
  const T D[4][4] = {{-(e[0] + e[2] + e[3]), e[0], e[2], e[3]},
                     {e[0], -(e[0] + e[1] + e[4]), e[1], e[4]},
                     {e[2], e[1], -(e[1] + e[2] + e[5]), e[5]},
                     {e[3], e[4], e[5], -(e[3] + e[4] + e[5])}};
  if (!EXACT) {
    const T gk[6] = {(T)6 * D[1][1], (T)6 * D[2][2], (T)6 * D[3][3],
                     (T)6 * e[1],    (T)6 * e[4],    (T)6 * e[5]};
#pragma unroll
    for (int m = 0; m < NL; ++m) {
      T sum = gk[0] * tab<T>((0 * NL + LROW) * NL + m);
#pragma unroll
      for (int k = 1; k < 6; ++k) sum = fma(gk[k], tab<T>((k * NL + LROW) * NL + m), sum);
      if (MASS) sum = fma(md, tab<T>((6 * NL + LROW) * NL + m), sum);
      v[m] = sum;
    }
    return;
  }
  if (NL == 4) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      v[j] = MASS ? fma(md, tab<T>((6 * NL + LROW) * NL + j), D[0][j]) : D[0][j];
    return;
  }
#pragma unroll
  for (int m = 0; m < NL; ++m) {
    T q;
    if (LROW == 0) {
      if (m < 4) {
        q = D[0][m] * (T)(0.8 * (1.0 + dlt(0, m)) - 1.0);
      } else {
        const int a = edge_a(m - 4), b = edge_b(m - 4);
        q = fma(D[0][b], (T)(4.0 * coefE(0, a)), D[0][a] * (T)(4.0 * coefE(0, b)));
      }
    } else {
      if (m < 4) {
        q = fma(D[1][m], (T)(4.0 * coefE(m, 0)), D[0][m] * (T)(4.0 * coefE(m, 1)));
      } else {
        const int cc = edge_a(m - 4), d = edge_b(m - 4);
        // row edge (a,b) = (0,1): D_bd F(a,c) + D_bc F(a,d) + D_ad F(b,c) + D_ac F(b,d)
        q = fma(D[1][d], (T)(16.0 * coefF(0, cc)),
                fma(D[1][cc], (T)(16.0 * coefF(0, d)),
                    fma(D[0][d], (T)(16.0 * coefF(1, cc)), D[0][cc] * (T)(16.0 * coefF(1, d)))));
      }
    }
    if (MASS) q = fma(md, tab<T>((6 * NL + LROW) * NL + m), q);
    v[m] = q;
  }
}

// ROWS rows per workgroup, LPR lanes per row, the block's CSR entries staged
// in LDS (plus, with ELIM, a slot at index `dummy` that takes the
// contributions to eliminated columns).  The lanes of a row take its
// incidences in turn (lane sub takes sub, sub + LPR, ...) and add their rows
// of the element matrices in LPR ordered phases, so every entry sums its
// contributions in incidence order, independent of ROWS and LPR.  An
// incidence reads its record, then its element's edge values: an FP64
// vertex row gathers the 3 it needs by address; edge rows and FP32 rows load
// the whole element record in 16-byte loads and select in registers.
// An FP64 vertex row's incidences (about 24) are software-pipelined: while
// incidence j is added, the edge values of j+1 and the record of j+3 are in
// flight; edge rows (about 5 incidences) and FP32 rows load theirs in chunks
// instead.
//
// GACC (blocks whose entries exceed the LDS budget; LPR = 1): each lane
// zeroes and accumulates its own row in the values array itself, in the same
// incidence order, so the result is the staged form's bit for bit.
template <int NL, int LROW, typename T, bool MASS, bool EXACT, bool ELIM, int ROWS, int LPR,
          typename Rec, bool GACC = false, bool PATCH = false>
// patched P1 rows: at most 80 VGPRs (FP64), so 6 waves per SIMD fit (5:
// +1.4%, 7: +20% from spills at 128^3); FP32 P1 rows at 5 waves per SIMD
// (-5% against 6).  P2 blocks stage more entries, and their LDS allows about
// 3 waves per SIMD for vertex rows, 4 for edge rows
__global__ __launch_bounds__(ROWS *LPR, !PATCH ? 1 : NL == 4 ? (sizeof(T) == 4 ? 5 : 6) : LROW == 0 ? 3 : 4) void k_asm_rows(
    int32_t row_begin, int32_t row_end, int32_t dummy, EdgePack ep,
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ inc_ptr,
    const int32_t *__restrict__ diagpos, const Rec *__restrict__ rec, const T *__restrict__ geo,
    ElemSrc es, PatchSrc pc, T *__restrict__ vals) {
  static_assert(!GACC || LPR == 1, "global accumulation takes one lane per row");
  static_assert(!PATCH || (ROWS == kPatchRows && !GACC),
                "vertex patches: kPatchRows-row staged blocks");
  extern __shared__ __align__(16) char lds_raw[];
  constexpr int NT = ROWS * LPR;
  constexpr int GS = geo_stride<T, MASS>();
  const int lane = threadIdx.x;
  const int sub = lane % LPR;
  const int r0 = row_begin + xcd_chunk(blockIdx.x, gridDim.x) * ROWS;
  const int r1 = min(r0 + ROWS, row_end);
  const int s0 = rowptr[r0];
  const int ns = rowptr[r1] - s0;
  T *acc = GACC ? vals + s0 : reinterpret_cast<T *>(lds_raw);
  // Edge rows and FP32 rows load their incidences in chunks (below) and
  // issue the lane's row pointers before the LDS is cleared, needing them
  // together after it (one round trip; otherwise the row offset is loaded
  // only after the loop guard has waited for the incidence range).
  // Measured at 64^3 P2: FP64 vertex rows are fastest pipelined.
  // P1 rows form their element factors in place (no factor kernel, no factor
  // array: measured 706 -> 593 us per 128^3 P1 FP64 assembly with vertex
  // gathers, before patches); P2 rows read the factor records (in place:
  // P2 FP64 the same, P2 FP32 +13%).
  // P1 rows (both precisions) form the factors in place, in the relabelled
  // element's own vertex order (row vertex first), so the row's edge values
  // come out of the arithmetic directly, without selecting them; patched rows
  // read the vertex coordinates from the block's patch in LDS, so that a
  // record and a coefficient are the only per-incidence global loads.
  constexpr bool kInl = NL == 4;
  constexpr bool kChunked = kInl || PATCH || LROW == 4 || sizeof(T) == 4;
  const int r = r0 + lane / LPR;
  const int rr = kChunked ? min(r, r1 - 1) : r;
  // compact records leave out the row's own column: its position is the
  // row's diagonal
  constexpr bool kSelf = std::is_same<Rec, RecP2c>::value || std::is_same<Rec, RecP1p>::value ||
                        std::is_same<Rec, RecP2e>::value || std::is_same<Rec, RecP1q>::value;
  // RecP1q records exist only as ELL copies, their elements relative to the
  // block's lowest (eb0)
  constexpr bool kQ = std::is_same<Rec, RecP1q>::value;
  static_assert(!kQ || PATCH, "RecP1q records: patched blocks");
  const bool ell = kQ || pc.ell;
  int eb0 = 0;
  // records that carry their P2 patch words
  constexpr bool kRecLv = std::is_same<Rec, RecP2e>::value;
  using RV = RecView<Rec, LROW>;
  int base = 0, iend = 0, i = 0, self = 0;
  uint32_t lvr = 0;
  // rows of a patched system, in its kPatchRows-row blocks, with packed row
  // words: one 8-byte load per lane (128^3 P1 FP64 assembly -2%, 64^3 P2
  // -1%); also the FP32 P2 launches, which read factor records instead of
  // the patches but stage the same blocks
  const bool prow = kChunked && ROWS == kPatchRows && pc.prow != nullptr;
  if (kChunked) {
    if (prow) {
      const int inc0 = inc_ptr[r0];
      const uint64_t w = pc.prow[rr];
      const int i0 = inc0 + (int)((w >> 24) & 4095u);
      base = (int)(w & 0xFFFFu);
      self = (int)((w >> 16) & 255u);
      iend = i0 + (int)((w >> 36) & 255u);
      i = i0 + sub;
      lvr = (uint32_t)((w >> 44) & 1023u) | ((uint32_t)((w >> 54) & 1023u) << 16);
    } else {
      base = rowptr[rr] - s0;
      iend = inc_ptr[rr + 1];
      i = inc_ptr[rr] + sub;
      if (kSelf) self = diagpos[rr] - s0 - base;  // position of the diagonal in the row
    }
  }
  // vertex patch: coordinates as three arrays of pc.slots doubles after the
  // staged entries; npm = the patch's last index (record indices are clamped
  // to it: a row's records always point into its own block's patch)
  double *PX = nullptr;
  int npm = 0, p0 = 0;
  // coordinates of patch vertex lv (staged, or past the staged ones): the
  // LDS read is unconditional (clamped index) and only the memory read sits
  // in a branch, a form the compiler keeps apart
  // (vertex-row launches stage every patch in full: their patches are alike)
  constexpr bool kOvf = NL == 10 && LROW == 4;
  auto patch_xyz = [&](int lv) {
    const int ls = kOvf ? min(lv, pc.slots - 1) : lv;
    double3 q;
    q.x = PX[ls];
    q.y = PX[pc.slots + ls];
    q.z = PX[2 * pc.slots + ls];
    if (kOvf && lv >= pc.slots) {
      // two 16-byte loads, (x, y) and (y, z): unlike the scalar LDS reads,
      // so the compiler cannot merge the two branches into a flat load
      // (which would wait for every load in flight)
      const size_t v = (size_t)pc.vid[p0 + lv];
      const dbl2_a8 xy = *reinterpret_cast<const dbl2_a8 *>(es.xyz + 3 * v);
      const dbl2_a8 yz = *reinterpret_cast<const dbl2_a8 *>(es.xyz + 3 * v + 1);
      q.x = xy.x;
      q.y = xy.y;
      q.z = yz.y;
    }
    return q;
  };
  // Patched rows: the lane's first KP incidences (records, then their
  // coefficients) are requested before the patch is staged, so that their
  // latency overlaps the staging and the barrier; a block then waits for
  // about three memory round trips before it computes.  Vertex rows carry
  // about 24 incidences, P2 edge rows about 5.
  constexpr int KP = !PATCH ? 1 : LROW == 0 ? 6 : 3;
  RV Rp[KP];
  uint32_t lvw[KP];  // P2 patch words
  double ksp[KP], kmp[KP];
  int lastp = 0;
  // ELL records: the block's first slot and last step
  int64_t eb = 0;
  int emax = 0;
  // record source of the lane's incidence step `st` (CSR position j when
  // the records are read in CSR order)
  auto rec_src = [&](int st, int j) -> int64_t {
    return ell ? eb + (int64_t)min(st, emax) * NT + lane : (int64_t)j;
  };
  if constexpr (PATCH) {
    lastp = iend > i - sub ? iend - 1 : 0;
    const int pb = pc.pb0 + (r0 - row_begin) / kPatchRows;
    const Rec *rsrc = rec;
    const uint32_t *lsrc = pc.inc_lv;
    if (ell) {
      eb = pc.ell_base + (int64_t)((r0 - row_begin) / kPatchRows) * pc.ell_steps * NT;
      emax = pc.ell_steps - 1;
      rsrc = reinterpret_cast<const Rec *>(pc.ell_rec);
      lsrc = pc.ell_lv;
    }
    if constexpr (kQ) eb0 = pc.ebase[pb];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int j = i + k * LPR;
      const int64_t src = rec_src(k, j < iend ? j : lastp);
      Rp[k] = RV::load(rsrc, src, self);
      if constexpr (kRecLv)
        lvw[k] = Rp[k].patch_word();
      else if constexpr (NL == 10)
        lvw[k] = lsrc[src];
    }
    PX = reinterpret_cast<double *>(lds_raw +
                                    ((((size_t)dummy + (ELIM ? 2 : 0)) * sizeof(T) + 7) & ~(size_t)7));
    p0 = pc.ptr[pb];
    const int np = pc.ptr[pb + 1] - p0;
    npm = max(np - 1, 0);
    // the first pc.slots vertices are staged; a P2 edge-row block with a
    // larger patch (a rare outlier: the launch's LDS is sized for its
    // typical blocks) reads the coordinates of the others from memory
    const int ns_p = kOvf ? min(np, pc.slots) : np;
    if (!prow) lvr = pc.row_lv[rr];
    // up to three vertices per lane with their loads in flight together;
    // from the padded ids (all pc.slots of them, whatever the block's patch
    // size, so the loads do not wait for it) or the block's own
    constexpr int KV = 3;
    int vid[KV];
    const int *vsrc = pc.vid + p0;
    int nst = ns_p;
    if (pc.vid_pad) {
      vsrc = pc.vid_pad + pc.vpad_base + (int64_t)((r0 - row_begin) / kPatchRows) * pc.vpad_stride;
      nst = pc.slots;
    }
    // P2 blocks read their patch ids nontemporally (read once; 64^3 P2 FP64
    // assembly -4% to +1% on different boxes; P1 blocks: the same or slower)
#pragma unroll
    for (int u = 0; u < KV; ++u)
      vid[u] = lane + u * NT >= nst ? 0
               : NL == 10       ? __builtin_nontemporal_load(vsrc + lane + u * NT)
                                : vsrc[lane + u * NT];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int el = (Rp[k].el >> 4) + eb0;
      ksp[k] = es.ksf ? es.ksf[el] : es.ksc;
      kmp[k] = MASS ? (es.kmf ? es.kmf[el] : es.kmc) : 0.0;
    }
    auto put = [&](int p, int v) {
      const dbl2_a8 q = *reinterpret_cast<const dbl2_a8 *>(es.xyz + 3 * (size_t)v);
      const double z = es.xyz[3 * (size_t)v + 2];
      PX[p] = q.x;
      PX[pc.slots + p] = q.y;
      PX[2 * pc.slots + p] = z;
    };
#pragma unroll
    for (int u = 0; u < KV; ++u)
      if (lane + u * NT < nst) put(lane + u * NT, vid[u]);
    for (int p = lane + KV * NT; p < nst; p += NT) put(p, vsrc[p]);
  }
  if constexpr (GACC) {
    // each lane clears its own row (it is the only one to touch it)
    if (r < r1)
      for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) vals[k] = (T)0;
  } else {
    for (int s = lane; s < ns; s += NT) acc[s] = (T)0;
  }
  __syncthreads();
  if (kChunked) asm volatile("" ::"v"(base), "v"(iend), "v"(i), "v"(self));
  // Patched FP64 P2 vertex rows at 4 lanes per row (one row per group of
  // four lanes): the row's own diagonal entry, which every incidence adds
  // to, sums in registers instead of LDS.  Every lane of the row adds the
  // four lanes' terms of each step in lane order (DPP broadcasts), the order
  // in which the LDS phases add them, so the sum has the same bits; lane 0
  // writes it to the staged row at the end: a tenth fewer LDS add
  // instructions, which bound these rows (64^3 P2 FP64 assembly -1.1% to
  // -2.6%).  P1 rows (a quarter fewer adds) measured the same time with the
  // extra registers, P2 edge rows (read-add-write) +7%.
  constexpr bool kRegDiag = PATCH && NL == 10 && LROW == 0 && LPR == 4 && sizeof(T) == 8;
  T dsum = (T)0;
  // the row's vertex, or its two endpoints (patched rows)
  double own[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  if constexpr (PATCH) {
#pragma unroll
    for (int h = 0; h < (LROW == 0 ? 1 : 2); ++h) {
      const int lv = min((int)((lvr >> (16 * h)) & 0xFFFFu), npm);
      const double3 q = patch_xyz(lv);
      own[h][0] = q.x;
      own[h][1] = q.y;
      own[h][2] = q.z;
    }
  }
  if (r < r1) {
    if (!kChunked) {
      base = rowptr[r] - s0;
      iend = inc_ptr[r + 1];
      i = inc_ptr[r] + sub;
      if (kSelf) self = diagpos[r] - s0 - base;
    }
    // Pipelined form: three slots rotate; unrolled by three so every load
    // lands in the slot that consumes it and no loaded register is copied.
    // In both forms, indices past the row's last incidence are clamped (loads
    // stay in bounds) and their LDS update is masked off (measured -1.5%
    // against sending it to the dummy region, where the padding lanes of
    // all rows hit the same slots).
    const int last = iend > i - sub ? iend - 1 : 0;  // a row without incidences reads record 0
    auto at = [&](int j) { return j < iend ? j : last; };
    struct Slot {
      RV r;
      T e[6];
      T md;
    };
    // relabelled edge j of incidence basis l is original edge (w >> 3j) & 7
    auto edge_word = [&](int l) {
      if (LROW == 0) {
        const uint32_t a = (l & 1) ? ep.w[1] : ep.w[0], b = (l & 1) ? ep.w[3] : ep.w[2];
        return (l & 2) ? b : a;
      }
      return pick6(ep.w + 4, (uint32_t)(l - 4));
    };
    auto load_vals = [&](Slot &A) {
      const uint32_t w = edge_word(A.r.el & 15);
      const T *g = geo + (size_t)(A.r.el >> 4) * GS;
      // FP32 records and FP64 edge rows: the whole record in 16-byte loads,
      // the needed values selected in registers.  A gather is paced by the
      // cache lines it touches (one per lane and load), not by its bytes:
      // measured -11% FP32 edge rows (2 loads instead of 5), -7% FP32
      // vertex rows (2 instead of 3), -3% FP64 edge rows (3 instead of 5);
      // FP64 vertex rows (3 either way) are slower with the wider registers.
      constexpr bool kWide = sizeof(T) == 4 || LROW == 4;
      if constexpr (kWide) {
        T f[8];
        if constexpr (sizeof(T) == 8) {
#pragma unroll
          for (int k = 0; k < GS / 2; ++k) {
            const double2 q = reinterpret_cast<const double2 *>(g)[k];
            f[2 * k] = q.x;
            f[2 * k + 1] = q.y;
          }
        } else if constexpr (GS == 6) {
          // 24-byte records: 8-byte aligned
          typedef float flt4_a8 __attribute__((ext_vector_type(4), aligned(8)));
          const flt4_a8 q0 = *reinterpret_cast<const flt4_a8 *>(g);
          const float2 q1 = reinterpret_cast<const float2 *>(g)[2];
          f[0] = q0.x, f[1] = q0.y, f[2] = q0.z, f[3] = q0.w;
          f[4] = q1.x, f[5] = q1.y, f[6] = f[7] = (T)0;
        } else {
          const float4 q0 = reinterpret_cast<const float4 *>(g)[0];
          const float4 q1 = reinterpret_cast<const float4 *>(g)[1];
          f[0] = q0.x, f[1] = q0.y, f[2] = q0.z, f[3] = q0.w;
          f[4] = q1.x, f[5] = q1.y, f[6] = q1.z, f[7] = q1.w;
        }
#pragma unroll
        for (int j = 0; j < 6; ++j)
          A.e[j] = edge_needed<LROW, EXACT>(j)
                       ? pick6v(f[0], f[1], f[2], f[3], f[4], f[5], (w >> (3 * j)) & 7u)
                       : (T)0;
        A.md = MASS ? f[6] : (T)0;
      } else {
#pragma unroll
        for (int j = 0; j < 6; ++j)
          A.e[j] = edge_needed<LROW, EXACT>(j) ? g[(w >> (3 * j)) & 7u] : (T)0;
        A.md = MASS ? g[6] : (T)0;
      }
    };
    // add incidence j's row of the element matrix (slot A) to the row's entries
    auto add = [&](const Slot &A, int j) {
      T v[NL];
      row_values<NL, LROW, T, MASS, EXACT>(A.e, A.md, v);
      const bool valid = j < iend;
      if constexpr (kRegDiag) {
        // lane s of the row holds incidence j - sub + s (padding past iend)
        const int j0 = j - sub;
        const double t0 = quad_bcast<0>(v[LROW]), t1 = quad_bcast<1>(v[LROW]);
        const double t2 = quad_bcast<2>(v[LROW]), t3 = quad_bcast<3>(v[LROW]);
        dsum = j0 < iend ? dsum + t0 : dsum;
        dsum = j0 + 1 < iend ? dsum + t1 : dsum;
        dsum = j0 + 2 < iend ? dsum + t2 : dsum;
        dsum = j0 + 3 < iend ? dsum + t3 : dsum;
      }
      if constexpr (GACC) {
        if (valid) {
#pragma unroll
          for (int m = 0; m < NL; ++m) {
            const int pm = A.r.pos(m);
            if (!ELIM || pm != RV::kElim) acc[base + pm] += v[m];
          }
        }
        return;
      }
      const int bsel = valid ? base : dummy;
      int idx[NL];
#pragma unroll
      for (int m = 0; m < NL; ++m) {
        const int pm = A.r.pos(m);
        idx[m] = bsel + pm;
        if (ELIM) idx[m] = pm == RV::kElim ? dummy : idx[m];
      }
      // FP64 vertex rows: one LDS atomic add (ds_add_f64, no return) per
      // entry; otherwise read, add, write (measured: FP64 edge rows -19%,
      // ds_add_f32 several times slower than the plain form).  In a phase
      // every active lane owns a different row, so the plain form has no
      // conflicts (the dummy region excepted, whose values are discarded).
      // Padding steps (j past the row's last incidence) add nothing.
      // The lanes of a row are in one wavefront, whose LDS operations
      // execute in issue order; the compiler barrier keeps the phases in
      // that order.
#pragma unroll
      for (int ph = 0; ph < LPR; ++ph) {
        if ((LPR == 1 || sub == ph) && valid) {
          if constexpr (sizeof(T) == 8 && LROW == 0) {
#pragma unroll
            for (int m = 0; m < NL; ++m)
              if (!kRegDiag || m != LROW) unsafeAtomicAdd(&acc[idx[m]], v[m]);
          } else {
            T a[NL];
#pragma unroll
            for (int m = 0; m < NL; ++m) a[m] = acc[idx[m]];
#pragma unroll
            for (int m = 0; m < NL; ++m) acc[idx[m]] = a[m] + v[m];
          }
        }
        if (LPR > 1) asm volatile("" ::: "memory");
      }
    };
    if constexpr (PATCH) {
      // the first KP incidences are in Rp/ksp/kmp already; each incidence's
      // relabelled vertices 1-3 come from the patch, vertex 0 is the row's
      bool first = true;
      const Rec *rsrc = ell ? reinterpret_cast<const Rec *>(pc.ell_rec) : rec;
      const uint32_t *lsrc = ell ? pc.ell_lv : pc.inc_lv;
      for (int c0 = 0; i < iend; i += KP * LPR, c0 += KP) {
        if (!first) {
#pragma unroll
          for (int k = 0; k < KP; ++k) {
            const int64_t src = rec_src(c0 + k, at(i + k * LPR));
            Rp[k] = RV::load(rsrc, src, self);
            if constexpr (kRecLv)
              lvw[k] = Rp[k].patch_word();
            else if constexpr (NL == 10)
              lvw[k] = lsrc[src];
          }
#pragma unroll
          for (int k = 0; k < KP; ++k) {
            const int el = (Rp[k].el >> 4) + eb0;
            ksp[k] = es.ksf ? es.ksf[el] : es.ksc;
            kmp[k] = MASS ? (es.kmf ? es.kmf[el] : es.kmc) : 0.0;
          }
        }
        first = false;
        // relabelled vertex 0 (and 1, edge rows) from the row's, the rest
        // from the patch: Q = the patch vertices' coordinates of incidence k
        constexpr int M0 = LROW == 0 ? 1 : 2;
        auto coords = [&](int k, double (&Q)[4 - M0][3]) {
#pragma unroll
          for (int m = M0; m < 4; ++m) {
            int lv;
            if constexpr (NL == 4)
              lv = Rp[k].lv(m - M0);
            else
              lv = (int)((lvw[k] >> (10 * (m - M0))) & 0x3FFu);
            lv = min(lv, npm);
            const double3 q = patch_xyz(lv);
            Q[m - M0][0] = q.x;
            Q[m - M0][1] = q.y;
            Q[m - M0][2] = q.z;
          }
        };
        // P2 rows read incidence k+1's coordinates before incidence k's LDS
        // adds, which would otherwise stand between them in the LDS queue
        // (the compiler cannot move loads from the patch above stores into
        // the staged entries: both are LDS): 64^3 P2 FP64 assembly -2.7%.
        // P1 rows spill with the extra registers at their 6 waves per SIMD.
        constexpr bool kPipe = NL == 10;
        double Qn[4 - M0][3];
        if constexpr (kPipe) coords(0, Qn);
#pragma unroll
        for (int k = 0; k < KP; ++k) {
          double Qc[4 - M0][3];
          if constexpr (kPipe) {
#pragma unroll
            for (int m = 0; m < 4 - M0; ++m)
#pragma unroll
              for (int cc = 0; cc < 3; ++cc) Qc[m][cc] = Qn[m][cc];
            if (k + 1 < KP) coords(k + 1, Qn);
          }
          // edge-row lanes past their row's last incidence skip it (whole
          // waves past it branch over); vertex rows run it masked, as a
          // branch costs registers there
          if (LROW == 0 || i + k * LPR < iend) {
            if constexpr (!kPipe) coords(k, Qc);
            double P[4][3];
            if constexpr (LROW == 0) {
#pragma unroll
              for (int cc = 0; cc < 3; ++cc) P[0][cc] = own[0][cc];
            } else {
              const bool sw = (lvw[k] >> 30) & 1u;
#pragma unroll
              for (int cc = 0; cc < 3; ++cc) {
                P[0][cc] = sw ? own[1][cc] : own[0][cc];
                P[1][cc] = sw ? own[0][cc] : own[1][cc];
              }
            }
#pragma unroll
            for (int m = M0; m < 4; ++m)
#pragma unroll
              for (int cc = 0; cc < 3; ++cc) P[m][cc] = Qc[m - M0][cc];
            T f[8];
            edge_factors<T, MASS>(P, ksp[k], kmp[k], f);
            Slot A;
            A.r = Rp[k];
#pragma unroll
            for (int j = 0; j < 6; ++j) A.e[j] = edge_needed<LROW, EXACT>(j) ? f[j] : (T)0;
            A.md = MASS ? f[6] : (T)0;
            add(A, i + k * LPR);
          }
        }
      }
    } else if constexpr (kInl) {
      // P1 rows without a patch.  A lane takes 3 incidences at once (2: 626
      // us, 4: 764 us at 128^3): their records, then vertex ids and
      // coefficients, then the vertex coordinates in relabelled order, and
      // forms each element's factors as the patched rows do (the same
      // values, bit for bit).
      constexpr int KC = 3;
      const int4 *__restrict__ tets4 = reinterpret_cast<const int4 *>(es.tets);
      for (; i < iend; i += KC * LPR) {
        RV R[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) R[k] = RV::load(rec, at(i + k * LPR), self);
        int4 t[KC];
        double ks[KC], km[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          const int el = R[k].el >> 4;
          t[k] = tets4[el];
          ks[k] = es.ksf ? es.ksf[el] : es.ksc;
          km[k] = MASS ? (es.kmf ? es.kmf[el] : es.kmc) : 0.0;
        }
        // P[k][m]: relabelled vertex m of incidence k (m = 0: the row's)
        double P[KC][4][3];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          const int l = R[k].el & 15;
          const int id[4] = {t[k].x, t[k].y, t[k].z, t[k].w};
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            const uint32_t o = (ep.v4 >> (2 * (4 * m + l))) & 3u;
            const int a = (o & 1u) ? id[1] : id[0], b = (o & 1u) ? id[3] : id[2];
            const int v = (o & 2u) ? b : a;
            const dbl2_a8 q = *reinterpret_cast<const dbl2_a8 *>(es.xyz + 3 * (size_t)v);
            P[k][m][0] = q.x;
            P[k][m][1] = q.y;
            P[k][m][2] = es.xyz[3 * (size_t)v + 2];
          }
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          // the factors of the relabelled element are its row's edge values
          T f[8];
          edge_factors<T, MASS>(P[k], ks[k], km[k], f);
          Slot A;
          A.r = R[k];
#pragma unroll
          for (int j = 0; j < 6; ++j) A.e[j] = edge_needed<LROW, EXACT>(j) ? f[j] : (T)0;
          A.md = MASS ? f[6] : (T)0;
          add(A, i + k * LPR);
        }
      }
    } else if constexpr (kChunked) {
      // A lane takes up to KC incidences at once, records first, then all
      // their edge values, so a P2 edge row (about 5 incidences) or a vertex
      // row (about 24, 6 per lane) needs two load round trips, with nothing
      // carried in flight from one chunk to the next.
      constexpr int KC = (LROW == 4 && LPR > 1) ? 3 : 6;
      for (; i < iend; i += KC * LPR) {
        Slot S[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) S[k].r = RV::load(rec, at(i + k * LPR), self);
#pragma unroll
        for (int k = 0; k < KC; ++k) load_vals(S[k]);
#pragma unroll
        for (int k = 0; k < KC; ++k) add(S[k], i + k * LPR);
      }
    } else {
      Slot S0, S1, S2;
      auto step = [&](Slot &A, Slot &B, int j) {
        load_vals(B);
        add(A, j);
        A.r = RV::load(rec, at(j + 3 * LPR), self);
      };
      S0.r = RV::load(rec, at(i), self);
      S1.r = RV::load(rec, at(i + LPR), self);
      S2.r = RV::load(rec, at(i + 2 * LPR), self);
      load_vals(S0);
      for (; i < iend; i += 3 * LPR) {
        step(S0, S1, i);
        step(S1, S2, i + LPR);
        step(S2, S0, i + 2 * LPR);
      }
    }
  }
  if constexpr (GACC) return;
  if constexpr (kRegDiag)
    if (r < r1 && sub == 0) acc[base + self] = dsum;
  __syncthreads();
  // values are written once and not read again by this assembly: streaming
  // (nontemporal) stores keep them from evicting element records from L2
  for (int s = lane; s < ns; s += NT) __builtin_nontemporal_store(acc[s], vals + s0 + s);
}

// rules exact on the term's integrand (affine element, P0 coefficient)
static bool rule_exact(int order, int term, int rule) {
  if (term == MG_TERM_STIFF) return order == 1 ? rule >= 1 : rule >= 4;
  return order == 1 ? rule >= 4 : rule >= 11;  // mass: degree 2 / degree 4
}

template <int V>
using IC = std::integral_constant<int, V>;
template <bool V>
using BC = std::integral_constant<bool, V>;
template <typename R>
struct RecTag {
  using type = R;
};

// Whether an assembly of S needs the factor kernel: P1 rows and patched FP64
// P2 rows form the factors in place.
static bool needs_factors(const System &S, int nl, size_t tsize) {
  if (nl == 4) return false;
  if (tsize == 4) return true;
  for (const RowSegment &sg : S.segs)
    if (!sg.patch) return true;
  return false;
}

// Whether a segment's rows use its vertex patch: P1 rows in both precisions,
// P2 rows in FP64 (FP32 P2 rows read the factor records: measured 154 us per
// 64^3 P2 FP32 assembly against 173 us in place, where the edge rows repeat
// the element geometry six times per element)
static bool use_patch(const RowSegment &sg, int nl, size_t tsize) {
  return sg.patch > 0 && (nl == 4 || tsize == 8);
}

// geo_buf: where the element factor records go (nullptr: the context's own
// array, grown as needed)
template <int NL, typename T, typename Rec>
static void launch_rows(Ctx *c, System &S, const mg_operator *op, T *vals, bool elim,
                        T *geo_buf) {
  const bool stiff = op->terms & MG_TERM_STIFF, mass = op->terms & MG_TERM_MASS;
  const double ksc = stiff ? op->stiff_const : 0.0, kmc = mass ? op->mass_const : 0.0;
  const double *ksf = (stiff && op->stiff_coeff == MG_COEFF_P0) ? c->field[MG_FIELD_STIFF] : nullptr;
  const double *kmf = (mass && op->mass_coeff == MG_COEFF_P0) ? c->field[MG_FIELD_MASS] : nullptr;
  const bool exact = (!stiff || rule_exact(c->order, MG_TERM_STIFF, op->stiff_rule)) &&
                     (!mass || rule_exact(c->order, MG_TERM_MASS, op->mass_rule));
  static const EdgePack ep = pack_edges();
  const ElemSrc es{c->tets, c->xyz, ksf, kmf, ksc, kmc};
  // no factor kernel when every segment forms the factors in place (as kInl
  // in k_asm_rows)
  const bool inl = !needs_factors(S, NL, sizeof(T));
  // element records
  const int stride = mass ? geo_stride<T, true>() : geo_stride<T, false>();
  const size_t gbytes = sizeof(T) * stride * (size_t)c->nT;
  if (!inl && !geo_buf && c->geom_bytes < gbytes) {
    dfree(c, c->geom);
    c->geom = dalloc<char>(c, gbytes);
    c->geom_bytes = gbytes;
  }
  T *geo = geo_buf ? geo_buf : (T *)c->geom;
  if (c->nT > 0 && !inl) {
    if (mass)
      k_elem_factors<T, true><<<blocks_for(c->nT, 256), 256, 0, c->stream>>>(
          c->nT, c->tets, c->xyz, ksc, ksf, kmc, kmf, geo);
    else
      k_elem_factors<T, false><<<blocks_for(c->nT, 256), 256, 0, c->stream>>>(
          c->nT, c->tets, c->xyz, ksc, ksf, kmc, kmf, geo);
    MG_HIP(hipGetLastError());
  }
  // one segment's launch with the records read as R
  auto seg_launch = [&](const RowSegment &sg, auto rec_tag) {
    using R = typename decltype(rec_tag)::type;
    // (RecP1q records are read from the ELL copy alone)
    const R *rec = std::is_same<R, RecP1q>::value ? nullptr : (const R *)S.inc_rec;
    const int nblk = (sg.row_end - sg.row_begin + sg.rows - 1) / sg.rows;
    // staged entries, two slots that take the contributions to eliminated
    // columns (interior systems), then (at 8-byte alignment) the coordinates
    // of the first sg.patch vertices of each block's patch
    const bool patched = use_patch(sg, NL, sizeof(T));
    size_t lds = sg.gacc ? 0 : sizeof(T) * (size_t)(sg.lds_slots + (elim ? 2 : 0));
    if (patched) lds = ((lds + 7) & ~(size_t)7) + 24 * (size_t)sg.patch;
    // the segment's first block in its ELL region
    const int g = sg.lrow == 0 ? 0 : 1;
    const int nvb = (int)((S.nvrows + kPatchRows - 1) / kPatchRows);
    const int64_t q0 = g == 0 ? sg.pb0 : sg.pb0 - nvb;
    PatchSrc pc{S.patch_ptr, S.patch_vid, S.row_lv, S.inc_lv, sg.patch, sg.pb0, S.ell_rec,
                S.ell_lv,
                S.ell_base[g] + q0 * S.ell_steps[g] * kPatchRows * std::max(1, S.ell_lpr[g]),
                S.ell_steps[g], 0,
                S.vid_pad && sg.patch <= S.vpad_stride[g] ? S.vid_pad : nullptr,
                S.vpad_base[g] + q0 * S.vpad_stride[g], S.vpad_stride[g], S.ell_ebase,
                S.prow};
    auto go = [&](auto lr, auto ms, auto ex, auto rw, auto lp, auto ga, auto pt) {
      // the grid and the LDS size come from the segment's block rows
      MG_CHECK(decltype(rw)::value == sg.rows, MG_ERR_STATE,
               "assembly: a launch shape differs from its segment's row blocks");
      // patched launches read the ELL copy of the records when it is laid
      // out for their lanes per row and holds their record form
      pc.ell = decltype(pt)::value && S.ell_rec &&
               S.ell_lpr[decltype(lr)::value == 0 ? 0 : 1] == decltype(lp)::value &&
               S.ell_q == std::is_same<R, RecP1q>::value;
      MG_CHECK((pc.ell || !std::is_same<R, RecP1q>::value), MG_ERR_STATE,
               "assembly: 8-byte P1 records without their ELL copy");
      if (elim)
        k_asm_rows<NL, decltype(lr)::value, T, decltype(ms)::value, decltype(ex)::value, true,
                   decltype(rw)::value, decltype(lp)::value, R, decltype(ga)::value,
                   decltype(pt)::value>
            <<<nblk, decltype(rw)::value * decltype(lp)::value, lds, c->stream>>>(
                sg.row_begin, sg.row_end, sg.lds_slots, ep, S.rowptr, S.inc_ptr, S.diagpos, rec, geo,
                es, pc, vals);
      else
        k_asm_rows<NL, decltype(lr)::value, T, decltype(ms)::value, decltype(ex)::value, false,
                   decltype(rw)::value, decltype(lp)::value, R, decltype(ga)::value,
                   decltype(pt)::value>
            <<<nblk, decltype(rw)::value * decltype(lp)::value, lds, c->stream>>>(
                sg.row_begin, sg.row_end, sg.lds_slots, ep, S.rowptr, S.inc_ptr, S.diagpos, rec, geo,
                es, pc, vals);
    };
    // vertex-dof rows: 64 rows per workgroup, 1/2/4 lanes per row; P2 edge-dof
    // rows: 64/128/256 rows per workgroup, 1/2 lanes per row; blocks past the
    // LDS budget: 64 rows, one lane per row, accumulated in place
    auto shape = [&](auto lr, auto ms, auto ex) {
      const BC<false> st{};
      if (sg.gacc) {
        go(lr, ms, ex, IC<64>{}, IC<1>{}, BC<true>{}, st);
        return;
      }
      if constexpr (decltype(lr)::value == 0) {
        const int lpr = sizeof(T) == 4 ? sg.lpr32 : sg.lpr;
        if constexpr (std::is_same<R, RecP1p>::value ||
                      (std::is_same<R, RecP2c>::value && sizeof(T) == 8)) {
          if (patched) {
            const BC<true> pt{};
            if (lpr == 4)
              go(lr, ms, ex, IC<kPatchRows>{}, IC<4>{}, st, pt);
            else if (lpr == 2)
              go(lr, ms, ex, IC<kPatchRows>{}, IC<2>{}, st, pt);
            else
              go(lr, ms, ex, IC<kPatchRows>{}, IC<1>{}, st, pt);
            return;
          }
        }
        if (lpr == 4)
          go(lr, ms, ex, IC<64>{}, IC<4>{}, st, st);
        else if (lpr == 2)
          go(lr, ms, ex, IC<64>{}, IC<2>{}, st, st);
        else
          go(lr, ms, ex, IC<64>{}, IC<1>{}, st, st);
      } else {
        if constexpr ((std::is_same<R, RecP2c>::value || std::is_same<R, RecP2e>::value) &&
                      sizeof(T) == 8) {
          if (patched) {
            const BC<true> pt{};
            sg.lpr == 2 ? go(lr, ms, ex, IC<kPatchRows>{}, IC<2>{}, st, pt)
                        : go(lr, ms, ex, IC<kPatchRows>{}, IC<1>{}, st, pt);
            return;
          }
        }
        if constexpr (std::is_same<R, RecP2e>::value) {
          // RecP2e rows lie in kPatchRows-row patch blocks
          sg.lpr == 2 ? go(lr, ms, ex, IC<kPatchRows>{}, IC<2>{}, st, st)
                      : go(lr, ms, ex, IC<kPatchRows>{}, IC<1>{}, st, st);
          return;
        }
        if (sg.rows == 256)
          sg.lpr == 2 ? go(lr, ms, ex, IC<256>{}, IC<2>{}, st, st)
                      : go(lr, ms, ex, IC<256>{}, IC<1>{}, st, st);
        else if (sg.rows == 128)
          sg.lpr == 2 ? go(lr, ms, ex, IC<128>{}, IC<2>{}, st, st)
                      : go(lr, ms, ex, IC<128>{}, IC<1>{}, st, st);
        else
          sg.lpr == 2 ? go(lr, ms, ex, IC<64>{}, IC<2>{}, st, st)
                      : go(lr, ms, ex, IC<64>{}, IC<1>{}, st, st);
      }
    };
    auto terms = [&](auto lr) {
      auto sh = [&](auto ms, auto ex) {
        // RecP1q: patched P1 blocks at 4 lanes per row (checked by the caller)
        if constexpr (std::is_same<R, RecP1q>::value)
          go(lr, ms, ex, IC<kPatchRows>{}, IC<4>{}, BC<false>{}, BC<true>{});
        else
          shape(lr, ms, ex);
      };
      if (exact)
        mass ? sh(BC<true>{}, BC<true>{}) : sh(BC<false>{}, BC<true>{});
      else
        mass ? sh(BC<true>{}, BC<false>{}) : sh(BC<false>{}, BC<false>{});
    };
    if constexpr (std::is_same<R, RecP2e>::value) {
      terms(IC<4>{});  // edge-dof rows only
    } else {
      if (sg.lrow == 0)
        terms(IC<0>{});
      else if constexpr (NL == 10)
        terms(IC<4>{});
    }
    MG_HIP(hipGetLastError());
  };
  for (const RowSegment &sg : S.segs) {
    // patched P1 blocks at 4 lanes per row read the 8-byte records of the
    // ELL copy when it has them
    if constexpr (std::is_same<Rec, RecP1p>::value) {
      if (S.ell_q && !sg.gacc && use_patch(sg, NL, sizeof(T)) && S.ell_lpr[0] == 4 &&
          (sizeof(T) == 4 ? sg.lpr32 : sg.lpr) == 4) {
        seg_launch(sg, RecTag<RecP1q>{});
        continue;
      }
    }
    // P2 edge-dof rows of a system with RecP2e edge records
    if constexpr (std::is_same<Rec, RecP2c>::value) {
      if (sg.lrow == 4 && S.edge5) {
        seg_launch(sg, RecTag<RecP2e>{});
        continue;
      }
    }
    seg_launch(sg, RecTag<Rec>{});
  }
}

static bool rule_ok(int r) { return r == 1 || r == 4 || r == 11 || r == 15; }

template <typename T>
static void assemble_t(Ctx *c, int sys, const mg_operator *op, int prec, T *geo_buf) {
  System &S = c->sys[sys];
  const bool stiff = op->terms & MG_TERM_STIFF, mass = op->terms & MG_TERM_MASS;
  upload_tables(c, c->order, stiff ? op->stiff_rule : 0, mass ? op->mass_rule : 0, prec);
  T *vals = prec == MG_PREC_F64 ? (T *)S.vals64 : (T *)S.vals32;
  const bool elim = sys == MG_SYS_INTERIOR;
  switch (S.rec_kind) {
    case kRecP2c: launch_rows<10, T, RecP2c>(c, S, op, vals, elim, geo_buf); break;
    case kRecP2: launch_rows<10, T, RecP2>(c, S, op, vals, elim, geo_buf); break;
    case kRecP2w: launch_rows<10, T, RecP2w>(c, S, op, vals, elim, geo_buf); break;
    case kRecP1: launch_rows<4, T, RecP1>(c, S, op, vals, elim, geo_buf); break;
    case kRecP1w: launch_rows<4, T, RecP1w>(c, S, op, vals, elim, geo_buf); break;
    case kRecP1p: launch_rows<4, T, RecP1p>(c, S, op, vals, elim, geo_buf); break;
    default: MG_CHECK(false, MG_ERR_STATE, "assemble: no incidence records");
  }
}

// bytes of the element factor records an assembly writes (0: the rows form
// them in place)
static size_t factor_bytes(const Ctx *c, int sys, const mg_operator *op, int prec) {
  const bool mass = op->terms & MG_TERM_MASS;
  if (!needs_factors(c->sys[sys], c->nloc, prec == MG_PREC_F64 ? 8 : 4)) return 0;
  const size_t per = prec == MG_PREC_F64 ? 8 * (mass ? 8 : 6) : 4 * (mass ? 8 : 6);
  return per * (size_t)c->nT;
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

// The same for rules exact on the stiffness integrand (P1: all rules; P2: 4,
// 11, 15 points), in closed form.  With G_ij = grad(lambda_i).grad(lambda_j)
// (barycentric coordinates) and vol = det / 6:
//   P1: lam^T Khat u = vol sum_ij G_ij lam_i u_j;
//   P2: grad u_h = sum_i c_i grad(lambda_i) with c_i = sum_p alpha_ip lambda_p,
//       alpha_ii = 3 u_i, alpha_ip = 4 u_(ip) - u_i (u_(ip): the edge dof),
//       and int lambda_p lambda_q = vol (1 + delta_pq) / 20, so
//       lam^T Khat u = vol / 20 sum_ij G_ij (s^lam_i s^u_j + sum_p alpha^lam_ip alpha^u_jp),
//       s_i = sum_p alpha_ip = 4 sum_(p != i) u_(ip).
// About 150 FP64 operations per P2 element instead of the table form's 660
// multiply-adds with their constant-memory table.
template <int NL>
__global__ __launch_bounds__(256) void k_adjoint_exact(int64_t nT, const int32_t *__restrict__ tets,
                                                       const double *__restrict__ xyz,
                                                       const int32_t *__restrict__ dofs,
                                                       const int32_t *__restrict__ map,
                                                       const double *__restrict__ u,
                                                       const double *__restrict__ lam,
                                                       double *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nT) return;
  const int4 t = reinterpret_cast<const int4 *>(tets)[e];
  const int32_t id[4] = {t.x, t.y, t.z, t.w};
  int32_t d[NL];
  if constexpr (NL == 10) {
    // 40 bytes per element: five 8-byte loads
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      const int2 q = reinterpret_cast<const int2 *>(dofs + 10 * e)[m];
      d[2 * m] = q.x;
      d[2 * m + 1] = q.y;
    }
  } else {
    const int4 q = reinterpret_cast<const int4 *>(dofs)[e];
    d[0] = q.x, d[1] = q.y, d[2] = q.z, d[3] = q.w;
  }
  if (map) {
#pragma unroll
    for (int m = 0; m < NL; ++m) d[m] = map[d[m]];
  }
  double ue[NL], le[NL];
#pragma unroll
  for (int m = 0; m < NL; ++m) {
    ue[m] = d[m] >= 0 ? u[d[m]] : 0.0;
    le[m] = d[m] >= 0 ? lam[d[m]] : 0.0;
  }
  const GeomOut g = tet_geometry(xyz, id);
  // G over the barycentric gradients: lambda_1..3 from the metric, lambda_0
  // = 1 - the others
  double G[4][4];
  const int KA[6] = {0, 1, 2, 0, 0, 1}, KB[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int k = 0; k < 6; ++k) G[KA[k] + 1][KB[k] + 1] = G[KB[k] + 1][KA[k] + 1] = g.A[k];
#pragma unroll
  for (int j = 1; j < 4; ++j) G[0][j] = G[j][0] = -(G[1][j] + G[2][j] + G[3][j]);
  G[0][0] = -(G[0][1] + G[0][2] + G[0][3]);
  double s = 0.0;
  if constexpr (NL == 4) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) s += G[i][j] * le[i] * ue[j];
    s *= g.det / 6.0;
  } else {
    // local edges (0,1) (1,2) (2,0) (0,3) (1,3) (2,3) are dofs 4-9
    constexpr int EI[4][4] = {{-1, 4, 6, 7}, {4, -1, 5, 8}, {6, 5, -1, 9}, {7, 8, 9, -1}};
    double au[4][4], al[4][4], su[4], sl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      su[i] = 0.0;
      sl[i] = 0.0;
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        if (p == i) {
          au[i][p] = 3.0 * ue[i];
          al[i][p] = 3.0 * le[i];
        } else {
          au[i][p] = 4.0 * ue[EI[i][p]] - ue[i];
          al[i][p] = 4.0 * le[EI[i][p]] - le[i];
          su[i] += ue[EI[i][p]];
          sl[i] += le[EI[i][p]];
        }
      }
      su[i] *= 4.0;
      sl[i] *= 4.0;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        double mij = sl[i] * su[j];
#pragma unroll
        for (int p = 0; p < 4; ++p) mij += al[i][p] * au[j][p];
        s += G[i][j] * mij;
      }
    s *= g.det / 120.0;
  }
  out[e] = -s;
}

static void adjoint_launch(Ctx *c, int sys, int rule, const double *u, const double *lam,
                           double *out) {
  const int32_t *map = sys == MG_SYS_FULL ? nullptr : c->sysmap;
  if (rule_exact(c->order, MG_TERM_STIFF, rule)) {
    if (c->nloc == 10)
      k_adjoint_exact<10><<<blocks_for(c->nT, 256), 256, 0, c->stream>>>(
          c->nT, c->tets, c->xyz, c->dofs, map, u, lam, out);
    else
      k_adjoint_exact<4><<<blocks_for(c->nT, 256), 256, 0, c->stream>>>(
          c->nT, c->tets, c->xyz, c->dofs, map, u, lam, out);
    MG_HIP(hipGetLastError());
    return;
  }
  upload_tables(c, c->order, rule, 0, MG_PREC_F64);
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

namespace mg {

// checks shared by mg_assemble and mg_assemble_noalloc; returns the operator key
static OpKey assemble_checks(Ctx *c, int sys, const mg_operator *op, int prec) {
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
  (void)S;
  OpKey key;
  key.terms = op->terms;
  if (op->terms & MG_TERM_STIFF) {
    key.stiff_coeff = op->stiff_coeff;
    key.stiff_rule = op->stiff_rule;
    if (op->stiff_coeff == MG_COEFF_P0)
      key.stiff_field = c->field_ver[MG_FIELD_STIFF];
    else
      key.stiff_const = op->stiff_const;
  }
  if (op->terms & MG_TERM_MASS) {
    key.mass_coeff = op->mass_coeff;
    key.mass_rule = op->mass_rule;
    if (op->mass_coeff == MG_COEFF_P0)
      key.mass_field = c->field_ver[MG_FIELD_MASS];
    else
      key.mass_const = op->mass_const;
  }
  return key;
}

static void assemble_run(Ctx *c, int sys, const mg_operator *op, int prec, const OpKey &key,
                         void *geo_buf) {
  System &S = c->sys[sys];
  if (prec == MG_PREC_F64) {
    assemble_t<double>(c, sys, op, prec, (double *)geo_buf);
    S.has64 = true;
    S.dinv64_ok = false;
    S.seq64 = ++c->asm_seq;
    S.key64 = key;
  } else {
    assemble_t<float>(c, sys, op, prec, (float *)geo_buf);
    S.has32 = true;
    S.dinv32_ok = false;
    S.seq32 = ++c->asm_seq;
    S.key32 = key;
  }
}

}  // namespace mg

extern "C" {

int mg_assemble(mg_ctx *c, int sys, const mg_operator *op, int prec) {
  API_BEGIN(c)
  const OpKey key = assemble_checks(c, sys, op, prec);
  assemble_run(c, sys, op, prec, key, nullptr);
  API_END(c)
}

int mg_assemble_noalloc(mg_ctx *c, int sys, const mg_operator *op, int prec) {
  API_BEGIN(c)
  const OpKey key = assemble_checks(c, sys, op, prec);
  const size_t need = factor_bytes(c, sys, op, prec);
  MG_CHECK(need <= kNoAllocScratchBytes, MG_ERR_SIZE,
           "assemble_noalloc: the element factor records take " + std::to_string(need) +
               " bytes, more than the " + std::to_string(kNoAllocScratchBytes) +
               "-byte static scratch (use mg_assemble)");
  MG_CHECK(c->noalloc_scratch, MG_ERR_STATE, "assemble_noalloc: no scratch address");
  assemble_run(c, sys, op, prec, key, c->noalloc_scratch);
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
  DevTmp<double> du(c, n), dl(c, n), dg(c, c->nT);
  MG_HIP(hipMemcpyAsync(du, u, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  MG_HIP(hipMemcpyAsync(dl, lam, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  adjoint_launch(c, sys, rule, du, dl, dg);
  MG_HIP(hipMemcpyAsync(dJdk, dg, sizeof(double) * c->nT, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  API_END(c)
}

}  // extern "C"
