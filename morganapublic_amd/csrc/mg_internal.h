// Internal definitions shared by the translation units of libmgfe.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mgfe.h"

namespace mg {

// Incidence record: element e sees the row dof as its local basis l.  The
// element is relabelled by the even vertex permutation perm[l] (relabel_tables)
// that moves that basis to local index 0 (vertex) or 4 (edge); pos[m'] is the
// position, inside the row, of the column dof of relabelled basis m' (0xFF when
// the column is eliminated).
struct alignas(16) RecP2 {
  int32_t el;  // e << 4 | l
  uint8_t pos[10];
  uint8_t pad[2];
};
struct alignas(8) RecP1 {
  int32_t el;
  uint8_t pos[4];
};
// Compact P2 record (12 bytes, used when every row has <= 127 entries): the
// row's own column (relabelled basis 0 for a vertex row, 4 for an edge row)
// is the row's diagonal and is not stored; the other 9 positions take 7 bits
// each, in increasing relabelled-basis order (0x7F: eliminated column).
struct alignas(4) RecP2c {
  int32_t el;
  uint32_t lo, hi;  // 63 bits of positions
};

// Wide records for rows past 254 entries: uint16 positions (0xFFFF:
// eliminated column), 24 bytes (P2) / 12 bytes (P1).
struct alignas(8) RecP2w {
  int32_t el;
  uint16_t pos[10];
};
struct alignas(4) RecP1w {
  int32_t el;
  uint16_t pos[4];
};
// Patched P1 record (12 bytes, rows of <= 254 entries): the row's own column
// (relabelled basis 0) is its diagonal; bits 0-23 of (lo, hi) hold the 8-bit
// positions of relabelled bases 1-3 (0xFF: eliminated column), bits 24-59
// the 12-bit indices, in the row block's vertex patch, of relabelled
// vertices 1-3 (all zero for a row block without a patch).
struct alignas(4) RecP1p {
  int32_t el;
  uint32_t lo, hi;
};
// Patched P2 edge-row record (12 bytes, a RecP2c system whose edge-dof rows
// all have <= 31 entries and whose edge-row patches have <= 511 vertices):
// el as in RecP2c; (lo, hi) hold the 5-bit positions of the 9 columns other
// than the row's own (bits 0-44, increasing relabelled basis, 31: eliminated
// column), the 9-bit patch indices of relabelled vertices 2 and 3 (bits
// 45-53, 54-62) and, in bit 63, whether relabelled vertex 0 is the row's
// second endpoint.  It replaces a RecP2c record plus its 4-byte patch word.
struct alignas(4) RecP2e {
  int32_t el;
  uint32_t lo, hi;
};
// Patched P1 record of the ELL copy (8 bytes; systems whose rows have <= 31
// entries, whose patches have <= 512 vertices and whose patch blocks each
// reach elements within 2^22 of the block's lowest): bits 0-14 the 5-bit
// positions of relabelled bases 1-3 (31: eliminated column), bits 15-41 the
// 9-bit patch indices of relabelled vertices 1-3, bits 42-63 the element
// minus the block's lowest element (System::ell_ebase).
struct alignas(8) RecP1q {
  uint32_t lo, hi;
};
enum RecKind { kRecP1 = 1, kRecP1w, kRecP2, kRecP2c, kRecP2w, kRecP1p };

// Vertex patches of row blocks: the sorted distinct vertices of the
// elements incident to kPatchRows consecutive rows (vertex-dof and P2
// edge-dof rows in separate blocks), whose coordinates a row block stages in
// LDS.  A block gets a patch when its vertex references fit the LDS sort
// (kPatchKeys), it has at most kPatchMax distinct vertices and its entries
// and coordinates fit the assembly's LDS budget.
constexpr int kPatchRows = 64;
constexpr int kPatchMax = 2048;    // P1: 12-bit indices in the records
constexpr int kPatchMaxP2 = 1023;  // P2: 10-bit indices in the patch words
constexpr int kPatchKeys = 8192;
// LDS bytes of a patched FP64 row block: its entries and the 256-slot
// region of eliminated columns, then the patch coordinates (3 doubles each)
inline size_t patch_lds(int64_t entries, int64_t verts) {
  return (size_t)(entries + 256) * 8 + 24 * (size_t)verts;
}

// Rows of one assembly launch: blocks of `rows` rows, each staging its CSR
// entries in LDS; lrow is the local basis every incidence of these rows is relabelled to
// (0 for vertex-dof rows, 4 for P2 edge-dof rows).
struct RowSegment {
  int32_t row_begin, row_end, lds_slots, lrow;
  int32_t rows, lpr;  // rows per workgroup, lanes per row
  int32_t lpr32;      // lanes per row of FP32 assemblies
  // 1: the block's entries exceed the LDS budget and are accumulated in the
  // values array itself, one lane per row (no staging)
  int32_t gacc;
  // > 0: every block has a vertex patch of at most this many vertices, staged
  // in LDS after the entries; pb0 = the patch block of the first row
  int32_t patch, pb0;
};

constexpr int kSpmvRows = 64;  // rows per SpMV block (128 for short rows)
// entries per SpMV block: 1024 (4 per lane) measured 4.7% faster than 2048
// (8 per lane) at 64^3 P2 (the smaller register set runs 6 waves per SIMD
// instead of 3), and 4% faster per FP32 iteration of the mixed PCG (154
// against 160 us: 8 waves instead of 5); 512 slower again
constexpr int kSpmvNnz = 1024;
// 16-bit column form of the SpMV streams: entry = cluster << 13 | offset
// against kColClusters bases per row block, each cluster 8192 columns wide
constexpr int kColClusters = 8;
constexpr int kColSpan = 1 << 13;

// SpMV row block: rows [r0, r1), CSR entries [s0, s0 + ns), and the column
// cluster bases of the 16-bit form; 64 bytes, read with one scalar load
struct alignas(64) BlkDesc {
  int32_t r0, r1, s0, ns;
  int32_t base[kColClusters];
  int32_t pad[4];
};

struct OpKey {
  int terms = 0, stiff_coeff = 0, stiff_rule = 0, mass_coeff = 0, mass_rule = 0;
  double stiff_const = 0.0, mass_const = 0.0;
  uint64_t stiff_field = 0, mass_field = 0;  // field versions (0: constant coefficient)
  bool operator==(const OpKey &o) const {
    return terms == o.terms && stiff_coeff == o.stiff_coeff && stiff_rule == o.stiff_rule &&
           mass_coeff == o.mass_coeff && mass_rule == o.mass_rule &&
           stiff_const == o.stiff_const && mass_const == o.mass_const &&
           stiff_field == o.stiff_field && mass_field == o.mass_field;
  }
};

struct System {
  bool built = false;
  int64_t nrows = 0, nnz = 0, n_inc = 0;
  int64_t nvrows = 0;  // vertex-dof rows (rows [0, nvrows)); P1: all rows
  int max_row_nnz = 0;
  int32_t *rowptr = nullptr;   // nrows + 1
  int32_t *colidx = nullptr;   // nnz
  int32_t *diagpos = nullptr;  // nrows, CSR index of the diagonal entry
  int32_t *inc_ptr = nullptr;  // nrows + 1 (row -> incidences)
  void *inc_rec = nullptr;     // n_inc records (RecP1, RecP1w, RecP1p, RecP2, RecP2c (+ RecP2e), RecP2w), sorted by (row, l, e)
  int rec_words = 0;           // int32 words per record: 2 (RecP1), 3 (RecP1w, RecP1p, RecP2c), 4 (RecP2), 6 (RecP2w)
  int rec_kind = 0;            // RecKind
  std::vector<RowSegment> segs;
  // vertex patches (RecP1p systems, and RecP2c systems with inc_lv): patch
  // block b (kPatchRows rows: vertex-dof rows from row 0, then edge-dof rows
  // from the first edge row) has vertices patch_vid[patch_ptr[b] ..
  // patch_ptr[b + 1]) (none: equal pointers); row_lv[r] holds the patch
  // index of row r's vertex (vertex rows) or endpoints (edge rows: low and
  // high 16 bits).  P2 patch word of incidence i, inc_lv[i]: the 10-bit patch
  // indices of relabelled vertices 1-3 (vertex rows) or 2-3 (edge rows, with
  // bit 30 set when relabelled vertex 0 is the row's second endpoint); P1
  // records carry theirs.
  int32_t *patch_ptr = nullptr;
  int32_t *patch_vid = nullptr;
  uint32_t *row_lv = nullptr;
  uint32_t *inc_lv = nullptr;
  // the records of the P2 edge-dof rows (incidences from edge_inc0 on) are
  // in the RecP2e form, which carries their patch words
  bool edge5 = false;
  int64_t edge_inc0 = 0;
  // one packed word per row of a patched system: the row's value offset in
  // its kPatchRows-row patch block (bits 0-15), its diagonal's position in
  // the row (16-23), its first incidence relative to the block's first
  // (24-35), its incidence count (36-43) and the patch indices of its vertex
  // or endpoints (44-53, 54-63): one load in place of rowptr, inc_ptr (2),
  // diagpos and row_lv (null: a row did not fit, the arrays are read)
  uint64_t *prow = nullptr;
  // ELL copy of the patch blocks' incidence records, one region for the
  // vertex-dof blocks (0) and one for the P2 edge-dof blocks (1), each with
  // ell_steps[g] steps per block: slot (step k, lane t) of the region's block
  // q is ell_rec[ell_base[g] + (q * ell_steps[g] + k) * kPatchRows * lpr + t],
  // lane t = (row in block) * lpr + sub holding the row's incidence
  // sub + k * lpr (past a row's last incidence: a copy of its last record; a
  // row without any: record 0), so that a block's records are addressed
  // without loading anything first; ell_lv: the P2 patch words in the same
  // order; ell_lpr[g]: the lanes per row it was laid out for (0: none)
  void *ell_rec = nullptr;
  uint32_t *ell_lv = nullptr;
  int64_t ell_base[2] = {0, 0};
  int ell_steps[2] = {0, 0};
  int ell_lpr[2] = {0, 0};
  // P1 systems: the ELL copy holds RecP1q records (8 bytes) instead of
  // RecP1p, with each patch block's lowest element in ell_ebase[b]
  bool ell_q = false;
  int32_t *ell_ebase = nullptr;
  // the first vpad_stride[g] patch vertex ids of region g's block q at
  // vid_pad[vpad_base[g] + q * vpad_stride[g] + p] (past a block's patch: its
  // first vertex), so that a block's patch is addressed without loading its
  // extent first; built together with the ELL copy
  int32_t *vid_pad = nullptr;
  int64_t vpad_base[2] = {0, 0};
  int vpad_stride[2] = {0, 0};
  // SpMV row blocks: each holds <= kSpmvNnz entries and <= kSpmvRows rows
  BlkDesc *sdesc = nullptr;
  int32_t n_sblk = 0;
  // rows longer than one block (kSpmvNnz entries) are in no block: a
  // workgroup per row takes them (k_long_rows)
  int32_t *long_rows = nullptr;
  int32_t n_long = 0;
  int srl = 4;  // lanes per row summing a block: 4 (64-row blocks) or 2 (128-row blocks)
  // the columns in 16-bit form (c16: every block fits kColClusters clusters,
  // bases in sdesc), read by the SpMV streams instead of colidx
  uint16_t *cofs = nullptr;  // nnz
  bool c16 = false;
  double *vals64 = nullptr;
  float *vals32 = nullptr;
  bool has64 = false, has32 = false;
  // assembly sequence numbers of the two value arrays: the mixed solvers
  // refresh the FP32 copy from FP64 values assembled after it
  uint64_t seq64 = 0, seq32 = 0;
  // what each value array holds (operator, rules, constants and the versions
  // of the P0 fields it read): the FP32 array stands in for the FP64 one in
  // the mixed solvers only when both hold the same operator
  OpKey key64, key32;
  double *dinv64 = nullptr;
  float *dinv32 = nullptr;
  bool dinv64_ok = false, dinv32_ok = false;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void *> owned;  // every device allocation, freed by destroy

  // mesh
  int64_t nV = 0, nT = 0, nB = 0;
  double *xyz = nullptr;   // 3 nV
  int32_t *tets = nullptr; // 4 nT, 0-based
  int32_t *btri = nullptr; // 3 nB, 0-based
  int32_t *bgeo = nullptr; // nB
  double *field[2] = {nullptr, nullptr};
  uint64_t field_ver[2] = {0, 0};  // bumped by every mg_field_set / mg_field_set_device

  // dof map
  int order = 0, nloc = 0;
  int64_t nE = 0, ndofs = 0;
  uint64_t *edge_keys = nullptr;  // nE sorted (lo << 32 | hi)
  int32_t *edge_nodes = nullptr;  // 2 nE
  int32_t *elem_edge = nullptr;   // 6 nT
  int32_t *dofs = nullptr;        // nloc nT
  bool edges_ok = false;

  // Dirichlet
  uint8_t *fixed = nullptr;   // ndofs
  int32_t *sysmap = nullptr;  // ndofs -> interior index or -1
  int64_t nfree = 0, nfree_vertex = 0;
  bool dirichlet_ok = false;

  System sys[2];
  uint64_t asm_seq = 0;  // numbers the assemblies of this context

  // per-element geometry/coefficient records for the current assembly
  void *geom = nullptr;
  size_t geom_bytes = 0;

  // solver workspace
  void *solve_ws = nullptr;
  size_t solve_ws_bytes = 0;

  // dual mesh
  bool dual_ok = false;
  double *edge_surf = nullptr, *edge_n = nullptr, *cell_volume = nullptr, *sigma_el = nullptr;
  double *sigma_len = nullptr;  // 6 nT: sigma_el over the edge's length (the SG coefficient's geometry)
  int32_t *vinc_ptr = nullptr;  // vertex -> (element, local vertex) incidence
  int32_t *vinc = nullptr;
  int32_t *vinc_pos = nullptr;   // 4 nT: incidence (4e + l) -> its index in vinc
  double *dd_contrib = nullptr;  // 8 nT: drift-diffusion residual terms in vinc order

  // scratch for library primitives (sorts, scans)
  void *tmp = nullptr;
  size_t tmp_bytes = 0;

  // this device's address of the static scratch array of the
  // allocation-free assembly (mg_assemble_noalloc), set by mg_ctx_create
  void *noalloc_scratch = nullptr;
};

// ---- error helpers ----------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};

#define MG_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::mg::Error{MG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define MG_CHECK(cond, code, msg)                  \
  do {                                             \
    if (!(cond)) throw ::mg::Error{(code), (msg)}; \
  } while (0)

template <typename T>
T *dalloc(Ctx *c, size_t n) {
  void *p = nullptr;
  if (n == 0) n = 1;
  MG_HIP(hipMalloc(&p, n * sizeof(T)));
  c->owned.push_back(p);
  return static_cast<T *>(p);
}
void dfree(Ctx *c, void *p);

// A device temporary of a host-copy entry point: released when it goes out of
// scope, also when the call ends early through MG_CHECK / MG_HIP.
template <typename T>
struct DevTmp {
  Ctx *c;
  T *p;
  DevTmp(Ctx *c_, size_t n) : c(c_), p(dalloc<T>(c_, n)) {}
  ~DevTmp() { dfree(c, p); }
  DevTmp(const DevTmp &) = delete;
  DevTmp &operator=(const DevTmp &) = delete;
  operator T *() const { return p; }
};

// A pair of HIP events destroyed with the scope.
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  EventPair() {
    if (hipEventCreate(&e0) != hipSuccess) e0 = nullptr;
    if (hipEventCreate(&e1) != hipSuccess) e1 = nullptr;
  }
  ~EventPair() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  EventPair(const EventPair &) = delete;
  EventPair &operator=(const EventPair &) = delete;
};
void free_system(Ctx *c, System &S);  // releases a system's buffers and resets it
void *scratch(Ctx *c, size_t bytes);  // context-owned reusable temporary

inline int blocks_for(int64_t n, int threads) { return (int)((n + threads - 1) / threads); }

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X (each
// with its own L2).  xcd_chunk maps the hardware block index to a logical
// chunk index so that XCD k works on the k-th contiguous eighth of the chunks:
// consecutive rows (and the vector entries / mesh data they touch) then stay
// in one L2.  A bijection on [0, nblocks).
__device__ __forceinline__ int xcd_chunk(int b, int nblocks) {
  const int q = nblocks >> 3;
  return b < (q << 3) ? (b & 7) * q + (b >> 3) : b;
}

// ---- cross-unit entry points ------------------------------------------------
void build_edges(Ctx *c);
void build_dofs(Ctx *c);
void build_dirichlet(Ctx *c, int n_geo, const int32_t *geo_ids);
void build_pattern(Ctx *c, int sys);
void noalloc_init(Ctx *c);  // device address of the static scratch array (device current)
void slot_map(Ctx *c, int sys, int32_t *slot);  // element -> CSR position, computed on request
void ensure_dinv(Ctx *c, int sys, int prec);
void spmv_launch(Ctx *c, int sys, int prec, const void *x, void *y);

// Even vertex permutations perm[l] (new local vertex -> old) moving local basis
// l to 0 (l < 4) or local edge l-4 to edge 0 = (0,1); sigma[l][m'] = old local
// basis of relabelled basis m'.  The quadrature rules are invariant under
// vertex permutations, so the relabelled element matrix is the same matrix.
void relabel_tables(int perm[10][4], int sigma[10][10]);
extern const int kEdgeV[6][2];

// host-side reference tables (restating the reference basis and quadrature)
int quad_rule(int npts, double *pts, double *w);
void ref_basis(int order, const double *y, double *val, double *grad /* [3][n] */);
// stiffness table: tab[k][l][m], k over (00,11,22,01,02,12); mass table tab[6][l][m]
void build_tables(int order, int stiff_rule, int mass_rule, double *tab /* 7*100 */);

}  // namespace mg

struct mg_ctx : public mg::Ctx {};
