// CSR SpMV and Jacobi-preconditioned CG: FP64 plain; FP32 inner solves
// restarted by FP64 iterative refinement; or one FP32 recursion with FP64
// residual replacement (pcg_reliable).
//
// SpMV works on row blocks (<= kSpmvRows rows, <= kSpmvNnz entries, fixed at
// pattern time): the block's (column, value) pairs are streamed with
// coalesced loads (columns in a 16-bit clustered form when the pattern
// allows it), the products staged in LDS, and each row summed by 4 lanes in a
// fixed order.  One CG iteration is three kernels, grid-stride over a
// fixed grid:
//   k_spmv_pq_stream : q = A p and the per-workgroup partials of p.q
//   k_update         : alpha, x += alpha p, r -= alpha q, partials of r.D^{-1}r, r.r
//   k_pupdate        : convergence test, beta, p = D^{-1} r + beta p
// Every workgroup of a kernel first sums the partials the previous kernel
// wrote, in the same fixed order, so all workgroups hold bit-identical
// scalars without communicating inside a kernel (a device-scope fence per
// workgroup would write back the L2 of its XCD each time).  The host reads the
// state every `check_every` iterations; once converged, queued kernels return
// at entry.  No reduction uses float atomics.
#include <cmath>
#include <cstdint>
#include <initializer_list>
#include <type_traits>

#include "mg_internal.h"

namespace mg {

static constexpr int kThreads = 256;
static constexpr int kMaxBlocks = 2048;  // SpMV grid cap = size of every partial array
// vector-kernel grid cap: every workgroup of k_update / k_pupdate re-totals the
// previous kernel's partials, so fewer, fuller workgroups read less of them.
// One workgroup per CU of an MI355X (64^3 P2 interior, against 1024: FP64
// PCG 179.4 -> 177.0 ms, FP32 with FP64 residual replacement 143.5 -> 140.1
// ms; 128 or 384: slower)
static constexpr int kMaxVecBlocks = 256;

// Host-visible solver state; written by block 0 of a kernel or by a one-block
// finalize kernel, read by later kernels (kernel boundaries order the writes).
struct PcgState {
  double tol2, rr, rz, pq, bb;
  int maxit, iters, done, status;
};

// Per-workgroup partial sums of the reductions (kMaxBlocks each).
struct Partials {
  double *pq, *rz0, *rz1, *rr, *tmp;
};

// Every thread of the block gets the same total of one value per thread
// (fixed shuffle tree + fixed order over the waves).
__device__ __forceinline__ double block_sum_all(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  __syncthreads();
  if (ln == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  __syncthreads();
  return t;
}

// Total of the n partials a previous kernel wrote.  Every workgroup computes it
// in the same order, so all of them obtain the same bits without any
// cross-workgroup communication inside the kernel.
__device__ __forceinline__ double total(const double *__restrict__ partial, int n, double *sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += partial[i];
  return block_sum_all(v, sh);
}

// Dot product of one CSR row with x by LPR lanes; each round issues U
// independent (column, value) loads per lane before the U gathers of x, so a
// lane keeps U loads in flight.
template <typename T, int LPR>
__device__ __forceinline__ T row_dot(int row, int lane, const int32_t *__restrict__ rowptr,
                                     const int32_t *__restrict__ colidx, const T *__restrict__ vals,
                                     const T *__restrict__ x) {
  constexpr int U = 4;
  const int k0 = rowptr[row], k1 = rowptr[row + 1];
  T s = (T)0;
  for (int kb = k0 + lane; kb < k1; kb += LPR * U) {
    int col[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kb + u * LPR;
      const bool ok = k < k1;
      col[u] = ok ? colidx[k] : 0;
      v[u] = ok ? vals[k] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (kb + u * LPR < k1) s += v[u] * x[col[u]];
  }
#pragma unroll
  for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
  return s;
}

// CSR-stream product of row blocks: a block's entries are streamed by all
// 256 threads (coalesced) into LDS as vals * x[col], then RL lanes per row
// (4, or 2 for 128-row blocks) sum their row in a fixed order.  Three blocks
// are in flight per workgroup (stream_blocks), so the stream never waits on
// loads issued in the same step.
//
// Columns come from colidx (int32) or, when every row block fits
// kColClusters clusters (C16), from the 16-bit form: entry = cluster << 13 |
// offset, column = (block descriptor).base[cluster] + offset, which halves
// the index bytes of the stream.  Both forms give the same columns, hence
// bit-identical products.  A block's rows, entry range and cluster bases come
// in one 64-byte descriptor (scalar loads, no dependent rowptr loads); the
// bases go to a small LDS table per pipeline slot, which the gather stage
// reads one step later (the finish stage's barriers lie in between).
struct Cols {
  const int32_t *idx;   // int32 columns
  const uint16_t *ofs;  // 16-bit form
};

template <typename T>
struct StreamBlk {
  static constexpr int PER = kSpmvNnz / kThreads;
  int r0, r1, s0, ns;
  int ra, re;  // this lane's row: its entries [ra, re) relative to s0
  int col[PER];  // column, or the raw 16-bit entry until gather_blk decodes it
  int *tab;      // this slot's cluster-base table in LDS (16-bit form)
  T v[PER];
  T g[PER];  // x at col
  T xr;      // the row operand of the epilogue at this lane's row
};

// branch-free: every lane issues its (column, value) loads for entries
// PER*lane .. PER*lane + PER - 1; lanes past the block's end re-read entry s0,
// and entries past it go to the spare LDS slot
template <typename T, bool C16, int RL>
__device__ __forceinline__ void load_blk(StreamBlk<T> &B, int b, const BlkDesc *__restrict__ desc,
                                         const int32_t *__restrict__ rowptr, const Cols cs,
                                         const T *__restrict__ vals) {
  const BlkDesc &D = desc[__builtin_amdgcn_readfirstlane(b)];  // uniform: scalar loads
  B.r0 = D.r0;
  B.r1 = D.r1;
  B.s0 = D.s0;
  B.ns = D.ns;
  if constexpr (C16) {
    // issued with the loads above (one wait), not after them
    int base[kColClusters];
#pragma unroll
    for (int i = 0; i < kColClusters; ++i) base[i] = D.base[i];
    __builtin_amdgcn_sched_barrier(0);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int i = 0; i < kColClusters; ++i) B.tab[i] = base[i];
    }
  }
  const int row = min(B.r0 + (int)(threadIdx.x / RL), B.r1);  // == r1: empty range
  B.ra = rowptr[row] - B.s0;
  B.re = rowptr[min(row + 1, B.r1)] - B.s0;
  // a lane's PER entries are consecutive: the columns in one load, the
  // values in one or two 16-byte loads (the arrays carry 8 entries of
  // padding, so the last block's over-read stays in bounds).  Measured
  // against PER loads strided by the workgroup size: -2.6% per FP64 PCG
  // iteration, -1% per FP32 one.
  static_assert(StreamBlk<T>::PER == 4, "four entries per lane");
  const int k0 = threadIdx.x * 4;
  const int kb = B.s0 + (k0 < B.ns ? k0 : 0);
  // Entries past the block's end (up to 3) belong to the next block or to the
  // zeroed padding; their products go to the spare LDS slot, and their
  // gather indices stay inside x without a clamp: int32 entries are columns
  // of the matrix (or 0), and the pattern build keeps every cluster base
  // <= max(n - kColSpan, 0), so any 16-bit entry decodes inside [0, n)
  // against any block's table.
  if constexpr (C16) {
    typedef uint32_t u3a4 __attribute__((ext_vector_type(3), aligned(4)));
    const u3a4 w = *reinterpret_cast<const u3a4 *>(cs.ofs + (kb & ~1));
    const bool odd = kb & 1;
    B.col[0] = (int)(odd ? (w.x >> 16) : (w.x & 0xFFFFu));
    B.col[1] = (int)(odd ? (w.y & 0xFFFFu) : (w.x >> 16));
    B.col[2] = (int)(odd ? (w.y >> 16) : (w.y & 0xFFFFu));
    B.col[3] = (int)(odd ? (w.z & 0xFFFFu) : (w.y >> 16));
  } else {
    typedef int i4a4 __attribute__((ext_vector_type(4), aligned(4)));
    const i4a4 q = *reinterpret_cast<const i4a4 *>(cs.idx + kb);
    B.col[0] = q.x, B.col[1] = q.y, B.col[2] = q.z, B.col[3] = q.w;
  }
  if constexpr (sizeof(T) == 8) {
    typedef double d2a8 __attribute__((ext_vector_type(2), aligned(8)));
    const d2a8 q0 = *reinterpret_cast<const d2a8 *>(vals + kb);
    const d2a8 q1 = *reinterpret_cast<const d2a8 *>(vals + kb + 2);
    B.v[0] = q0.x, B.v[1] = q0.y, B.v[2] = q1.x, B.v[3] = q1.y;
  } else {
    typedef float f4a4 __attribute__((ext_vector_type(4), aligned(4)));
    const f4a4 q = *reinterpret_cast<const f4a4 *>(vals + kb);
    B.v[0] = q.x, B.v[1] = q.y, B.v[2] = q.z, B.v[3] = q.w;
  }
}

// Gathers x at the block's columns (issued, not waited for).
template <typename T, bool C16, int RL>
__device__ __forceinline__ void gather_blk(StreamBlk<T> &B, const T *__restrict__ x,
                                           const T *__restrict__ rowvec) {
#pragma unroll
  for (int u = 0; u < StreamBlk<T>::PER; ++u) {
    const int e = B.col[u];
    const int cidx = C16 ? B.tab[(uint32_t)e >> 13] + (e & (kColSpan - 1)) : e;
    B.g[u] = x[cidx];
  }
  B.xr = rowvec[min(B.r0 + (int)(threadIdx.x / RL), B.r1 - 1)];
}

// Stages the products in LDS and sums each row with RL lanes; returns the row
// value in lane sub == 0 of each RL-lane group (row index in *row_out, -1 when
// the group has no row).
template <typename T, int RL>
__device__ __forceinline__ T finish_blk(const StreamBlk<T> &B, T *prod, int *row_out) {
  if (B.ns > 0) {
#pragma unroll
    for (int u = 0; u < StreamBlk<T>::PER; ++u) {
      const int k = (int)threadIdx.x * StreamBlk<T>::PER + u;
      prod[k < B.ns ? k : kSpmvNnz] = B.v[u] * B.g[u];
    }
  }
  __syncthreads();
  const int row = B.r0 + (int)(threadIdx.x / RL), sub = threadIdx.x % RL;
  T sum = (T)0;
  for (int k = B.ra + sub; k < B.re; k += RL) sum += prod[k];
#pragma unroll
  for (int o = 1; o < RL; o <<= 1) sum += __shfl_xor(sum, o, RL);
  __syncthreads();
  *row_out = (row < B.r1) ? row : -1;
  return sum;
}

// Grid-stride loop over the row blocks, software-pipelined over three
// blocks: while block b is summed, the gathers of block b + G and the
// (column, value) loads of block b + 2G are in flight.  Each step waits only
// for loads issued one step earlier (the vector-memory counter completes in
// issue order), and every buffer is written by loads only, never copied.
// epi(row, value, rowvec[row]) runs in lane sub == 0 of each row's group.
// A block holds kThreads / RL rows (RL = 4: 64 rows; RL = 2: 128 rows, for
// patterns whose rows are short).
template <typename T, bool C16, int RL, typename Epi>
__device__ __forceinline__ void stream_blocks(int nb, const BlkDesc *__restrict__ desc,
                                              const int32_t *__restrict__ rowptr, const Cols cs,
                                              const T *__restrict__ vals,
                                              const T *__restrict__ x,
                                              const T *__restrict__ rowvec, T *prod,
                                              int (*tab)[kColClusters], Epi epi) {
  const int G = gridDim.x;
  int b = blockIdx.x;
  if (b >= nb) return;
  StreamBlk<T> K0, K1, K2;
  K0.tab = tab[0];
  K1.tab = tab[1];
  K2.tab = tab[2];
  load_blk<T, C16, RL>(K0, b, desc, rowptr, cs, vals);
  load_blk<T, C16, RL>(K1, min(b + G, nb - 1), desc, rowptr, cs, vals);
  if (C16) __syncthreads();  // K0's cluster bases
  gather_blk<T, C16, RL>(K0, x, rowvec);
  // X: gathered, summed now; Y: loaded, gathered now; Z: loaded now
  auto step = [&](StreamBlk<T> &X, StreamBlk<T> &Y, StreamBlk<T> &Z) {
    gather_blk<T, C16, RL>(Y, x, rowvec);
    load_blk<T, C16, RL>(Z, min(b + 2 * G, nb - 1), desc, rowptr, cs, vals);
    int row;
    const T s = finish_blk<T, RL>(X, prod, &row);
    if (threadIdx.x % RL == 0 && row >= 0) epi(row, s, X.xr);
    b += G;
    return b < nb;
  };
  while (step(K0, K1, K2) && step(K1, K2, K0) && step(K2, K0, K1)) {
  }
}

template <typename T, bool C16, int RL>
__global__ __launch_bounds__(kThreads) void k_spmv_stream(
    int nb, const BlkDesc *__restrict__ desc, const int32_t *__restrict__ rowptr,
    const int32_t *__restrict__ colidx, const uint16_t *__restrict__ cofs, const T *__restrict__ vals, const T *__restrict__ x,
    T *__restrict__ y) {
  __shared__ T prod[kSpmvNnz + 1];
  __shared__ int tab[3][kColClusters];
  stream_blocks<T, C16, RL>(nb, desc, rowptr, Cols{colidx, cofs}, vals, x, x, prod, tab,
                        [&](int row, T s, T) { y[row] = s; });
}

// q = A p; partial p.q (row blocks)
template <typename T, bool C16, int RL>
__global__ __launch_bounds__(kThreads) void k_spmv_pq_stream(
    int nb, const BlkDesc *__restrict__ desc, const int32_t *__restrict__ rowptr,
    const int32_t *__restrict__ colidx, const uint16_t *__restrict__ cofs, const T *__restrict__ vals, const T *__restrict__ p,
    T *__restrict__ q, double *__restrict__ pq_part, const PcgState *__restrict__ st) {
  __shared__ T prod[kSpmvNnz + 1];
  __shared__ double sh[kThreads / 64];
  if (st->done) return;
  double acc = 0.0;
  __shared__ int tab[3][kColClusters];
  stream_blocks<T, C16, RL>(nb, desc, rowptr, Cols{colidx, cofs}, vals, p, p, prod, tab,
                        [&](int row, T s, T pr) {
    q[row] = s;
    acc += (double)(s * pr);
  });
  const double bsum = block_sum_all(acc, sh);
  if (threadIdx.x == 0) pq_part[blockIdx.x] = bsum;
}

// FP64 residual r = b - A x; partial r.r (row blocks)
template <bool C16, int RL>
__global__ __launch_bounds__(kThreads) void k_resid64_stream(
    int nb, const BlkDesc *__restrict__ desc, const int32_t *__restrict__ rowptr,
    const int32_t *__restrict__ colidx, const uint16_t *__restrict__ cofs, const double *__restrict__ vals,
    const double *__restrict__ x, const double *__restrict__ b, double *__restrict__ r,
    double *__restrict__ part) {
  __shared__ double prod[kSpmvNnz + 1];
  __shared__ double sh[kThreads / 64];
  double rr = 0.0;
  __shared__ int tab[3][kColClusters];
  stream_blocks<double, C16, RL>(nb, desc, rowptr, Cols{colidx, cofs}, vals, x, b, prod, tab,
                             [&](int row, double ax, double br) {
    const double ri = br - ax;
    r[row] = ri;
    rr += ri * ri;
  });
  const double s = block_sum_all(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Rows longer than an SpMV block belong to no block: one workgroup takes one
// such row at a time and its lanes sum the row's entries lane-strided, then
// in a fixed tree over lanes and waves (deterministic).  MODE 0: y = A x;
// 1: q = A p with the p.q partials; 2: r = b - A x with the r.r partials.
// Partials go to part[blockIdx.x] (the caller passes the slots after the
// stream kernel's).
template <typename T>
__device__ __forceinline__ T block_sum_t(T v, T *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  __syncthreads();
  if (ln == 0) sh[w] = v;
  __syncthreads();
  T t = (T)0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  __syncthreads();
  return t;
}

template <typename T, int MODE>
__global__ __launch_bounds__(kThreads) void k_long_rows(
    int nl, const int32_t *__restrict__ long_rows, const int32_t *__restrict__ rowptr,
    const int32_t *__restrict__ colidx, const T *__restrict__ vals, const T *__restrict__ x,
    const T *__restrict__ b, T *__restrict__ y, double *__restrict__ part,
    const PcgState *__restrict__ st) {
  __shared__ T shv[kThreads / 64];
  __shared__ double sh[kThreads / 64];
  if (MODE == 1 && st->done) return;
  double acc = 0.0;
  for (int i = blockIdx.x; i < nl; i += gridDim.x) {
    const int row = long_rows[i];
    T s = (T)0;
    for (int k = rowptr[row] + (int)threadIdx.x; k < rowptr[row + 1]; k += kThreads)
      s += vals[k] * x[colidx[k]];
    s = block_sum_t(s, shv);
    if (threadIdx.x == 0) {
      if (MODE == 2) {
        const double ri = (double)b[row] - (double)s;
        y[row] = (T)ri;
        acc += ri * ri;
      } else {
        y[row] = s;
        if (MODE == 1) acc += (double)(s * x[row]);
      }
    }
  }
  if (MODE != 0) {
    const double t = block_sum_all(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// The elementwise passes read 16-byte vectors (VEC: every array 16-byte
// aligned), 8 entries per array per thread and round, all loads of a round
// issued before any use; the n % (16 / sizeof(T)) last entries go to the
// scalar path of block 0.
template <typename T>
struct V16;
template <>
struct V16<float> {
  using type = float4;
};
template <>
struct V16<double> {
  using type = double2;
};

// alpha = rz / pq; x += alpha p; r -= alpha q; partials r.D^{-1}r and r.r
template <typename T, bool VEC>
__global__ __launch_bounds__(kThreads) void k_update(int n, T *__restrict__ x, T *__restrict__ r,
                                                     const T *__restrict__ p,
                                                     const T *__restrict__ q,
                                                     const T *__restrict__ dinv,
                                                     const double *__restrict__ pq_part, int npq,
                                                     const double *__restrict__ rz_prev, int nprev,
                                                     double *__restrict__ rz_out,
                                                     double *__restrict__ rr_out,
                                                     PcgState *__restrict__ st) {
  __shared__ double sh[kThreads / 64];
  if (st->done) return;
  const double pq = total(pq_part, npq, sh);
  const double rz = total(rz_prev, nprev, sh);
  if (!(pq > 0.0) || !isfinite(pq)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->pq = pq;
      st->status = MG_ERR_BREAKDOWN;
      st->done = 1;
    }
    return;
  }
  const T alpha = (T)(rz / pq);
  double a1 = 0.0, a2 = 0.0;
  auto one = [&](int i) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    a1 += (double)(ri * (dinv[i] * ri));
    a2 += (double)(ri * ri);
  };
  if constexpr (VEC) {
    using VT = typename V16<T>::type;
    constexpr int VW = 16 / sizeof(T), U = 8 / VW;
    const int nv = n / VW;
    VT *xv = reinterpret_cast<VT *>(x), *rv = reinterpret_cast<VT *>(r);
    const VT *pv = reinterpret_cast<const VT *>(p), *qv = reinterpret_cast<const VT *>(q),
             *dv = reinterpret_cast<const VT *>(dinv);
    for (int j0 = blockIdx.x * kThreads * U + threadIdx.x; j0 < nv; j0 += gridDim.x * kThreads * U) {
      VT X[U], R[U], P[U], Q[U], D[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = min(j0 + u * kThreads, nv - 1);
        X[u] = xv[j];
        R[u] = rv[j];
        P[u] = pv[j];
        Q[u] = qv[j];
        D[u] = dv[j];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (j0 + u * kThreads >= nv) continue;
        T *xe = reinterpret_cast<T *>(&X[u]), *re = reinterpret_cast<T *>(&R[u]);
        const T *pe = reinterpret_cast<const T *>(&P[u]), *qe = reinterpret_cast<const T *>(&Q[u]),
                *de = reinterpret_cast<const T *>(&D[u]);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
          xe[k] += alpha * pe[k];
          const T ri = re[k] - alpha * qe[k];
          re[k] = ri;
          a1 += (double)(ri * (de[k] * ri));
          a2 += (double)(ri * ri);
        }
        xv[j0 + u * kThreads] = X[u];
        rv[j0 + u * kThreads] = R[u];
      }
    }
    if (blockIdx.x == 0 && (int)threadIdx.x < n - nv * VW) one(nv * VW + threadIdx.x);
  } else {
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) one(i);
  }
  const double b1 = block_sum_all(a1, sh);
  const double b2 = block_sum_all(a2, sh);
  if (threadIdx.x == 0) {
    rz_out[blockIdx.x] = b1;
    rr_out[blockIdx.x] = b2;
  }
}

// convergence test on r.r; beta = rz_new / rz; p = D^{-1} r + beta p
template <typename T, bool VEC>
__global__ __launch_bounds__(kThreads) void k_pupdate(int n, const T *__restrict__ r,
                                                      T *__restrict__ p,
                                                      const T *__restrict__ dinv,
                                                      const double *__restrict__ rz_new_part,
                                                      const double *__restrict__ rz_part,
                                                      const double *__restrict__ rr_part, int np,
                                                      int iter, PcgState *__restrict__ st) {
  __shared__ double sh[kThreads / 64];
  if (st->done) return;
  const double rz_new = total(rz_new_part, np, sh);
  const double rz = total(rz_part, np, sh);
  const double rr = total(rr_part, np, sh);
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (first) {
    st->iters = iter + 1;
    st->rr = rr;
    st->rz = rz_new;
  }
  if (!isfinite(rr)) {
    if (first) {
      st->status = MG_ERR_BREAKDOWN;
      st->done = 1;
    }
    return;
  }
  if (rr <= st->tol2) {
    if (first) {
      st->status = MG_OK;
      st->done = 1;
    }
    return;
  }
  if (iter + 1 >= st->maxit) {
    if (first) {
      st->status = MG_ERR_NOCONV;
      st->done = 1;
    }
    return;
  }
  const T beta = (T)(rz_new / rz);
  if constexpr (VEC) {
    using VT = typename V16<T>::type;
    constexpr int VW = 16 / sizeof(T), U = 8 / VW;
    const int nv = n / VW;
    VT *pv = reinterpret_cast<VT *>(p);
    const VT *rv = reinterpret_cast<const VT *>(r), *dv = reinterpret_cast<const VT *>(dinv);
    for (int j0 = blockIdx.x * kThreads * U + threadIdx.x; j0 < nv; j0 += gridDim.x * kThreads * U) {
      VT P[U], R[U], D[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = min(j0 + u * kThreads, nv - 1);
        P[u] = pv[j];
        R[u] = rv[j];
        D[u] = dv[j];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (j0 + u * kThreads >= nv) continue;
        T *pe = reinterpret_cast<T *>(&P[u]);
        const T *re = reinterpret_cast<const T *>(&R[u]), *de = reinterpret_cast<const T *>(&D[u]);
#pragma unroll
        for (int k = 0; k < VW; ++k) pe[k] = de[k] * re[k] + beta * pe[k];
        pv[j0 + u * kThreads] = P[u];
      }
    }
    if (blockIdx.x == 0 && (int)threadIdx.x < n - nv * VW) {
      const int i = nv * VW + threadIdx.x;
      p[i] = dinv[i] * r[i] + beta * p[i];
    }
  } else {
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
      p[i] = dinv[i] * r[i] + beta * p[i];
  }
}

// r = b - A x (or r = b when x is null), p = D^{-1} r; partials r.p, r.r
template <typename T, int LPR>
__global__ __launch_bounds__(kThreads) void k_start(int n, const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colidx,
                                                    const T *__restrict__ vals,
                                                    const T *__restrict__ x,
                                                    const T *__restrict__ b, T *__restrict__ r,
                                                    T *__restrict__ p, const T *__restrict__ dinv,
                                                    double *__restrict__ rz_out,
                                                    double *__restrict__ rr_out) {
  __shared__ double sh[kThreads / 64];
  const int lane = threadIdx.x % LPR;
  const int stride = gridDim.x * (kThreads / LPR);
  double a1 = 0.0, a2 = 0.0;
  for (int row = blockIdx.x * (kThreads / LPR) + threadIdx.x / LPR; row < n; row += stride) {
    T ax = (T)0;
    if (x) ax = row_dot<T, LPR>(row, lane, rowptr, colidx, vals, x);
    if (lane == 0) {
      const T ri = b[row] - ax;
      const T zi = dinv[row] * ri;
      r[row] = ri;
      p[row] = zi;
      a1 += (double)(ri * zi);
      a2 += (double)(ri * ri);
    }
  }
  const double b1 = block_sum_all(a1, sh);
  const double b2 = block_sum_all(a2, sh);
  if (threadIdx.x == 0) {
    rz_out[blockIdx.x] = b1;
    rr_out[blockIdx.x] = b2;
  }
}

// one workgroup: st->rr (and st->done) from the partials of k_start
__global__ void k_start_finalize(const double *__restrict__ rz_part,
                                 const double *__restrict__ rr_part, int np,
                                 PcgState *__restrict__ st) {
  __shared__ double sh[kThreads / 64];
  const double rz = total(rz_part, np, sh);
  const double rr = total(rr_part, np, sh);
  if (threadIdx.x == 0) {
    st->rz = rz;
    st->rr = rr;
    st->iters = 0;
    st->status = MG_OK;
    st->done = (rr <= st->tol2) ? 1 : 0;
  }
}

__global__ void k_dot_self(int n, const double *__restrict__ v, double *__restrict__ part) {
  __shared__ double sh[kThreads / 64];
  double a = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    a += v[i] * v[i];
  const double s = block_sum_all(a, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// one workgroup: folds the first n_from partials into slot 0 and zeroes the
// slots up to n_to, so a reader summing n_to partials obtains the same total
__global__ void k_start_partials_to(double *__restrict__ part, int n_from, int n_to) {
  __shared__ double sh[kThreads / 64];
  const double t = total(part, n_from, sh);
  __syncthreads();
  for (int i = threadIdx.x; i < (n_from > n_to ? n_from : n_to); i += blockDim.x)
    part[i] = i == 0 ? t : 0.0;
}

// one workgroup: *out = total of np partials
__global__ void k_finalize_sum(const double *__restrict__ part, int np, double *__restrict__ out) {
  __shared__ double sh[kThreads / 64];
  const double t = total(part, np, sh);
  if (threadIdx.x == 0) *out = t;
}

template <typename T>
__global__ void k_dinv(int n, const int32_t *__restrict__ diag, const T *__restrict__ vals,
                       T *__restrict__ dinv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dinv[i] = diag[i] >= 0 ? (T)1 / vals[diag[i]] : (T)0;
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, double s, float *__restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)(a[i] * s);
}

__global__ void k_axpy_f32(int64_t n, double a, const float *__restrict__ d, double *__restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += a * (double)d[i];
}

// Residual replacement of the FP32 recursion: partials of r.D^{-1}r and r.r
// of the replaced residual r.
__global__ __launch_bounds__(kThreads) void k_replace_partials(int n, const float *__restrict__ r,
                                                               const float *__restrict__ dinv,
                                                               double *__restrict__ rz_out,
                                                               double *__restrict__ rr_out) {
  __shared__ double sh[kThreads / 64];
  double a1 = 0.0, a2 = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const float ri = r[i];
    a1 += (double)(ri * (dinv[i] * ri));
    a2 += (double)(ri * ri);
  }
  const double b1 = block_sum_all(a1, sh);
  const double b2 = block_sum_all(a2, sh);
  if (threadIdx.x == 0) {
    rz_out[blockIdx.x] = b1;
    rr_out[blockIdx.x] = b2;
  }
}

// The direction update the recursion skipped when it stopped, made with the
// replaced residual r (new units, r = r64 / s_new) and p still in the old
// units (s_old): in new units p' = pscale p with pscale = s_old / s_new and
// r.z_old' = pscale^2 r.z_old, so p = D^{-1} r + beta p' with
// beta p' = rz_repl / (rz_old pscale) p.
__global__ __launch_bounds__(kThreads) void k_replace_pupdate(int n, const float *__restrict__ r,
                                                              float *__restrict__ p,
                                                              const float *__restrict__ dinv,
                                                              const double *__restrict__ rz_repl,
                                                              const double *__restrict__ rz_old,
                                                              int np, double pscale) {
  __shared__ double sh[kThreads / 64];
  const double a = total(rz_repl, np, sh);
  const double o = total(rz_old, np, sh);
  const float beta = (float)(a / (o * pscale));
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    p[i] = dinv[i] * r[i] + beta * p[i];
}

// one workgroup: the recursion restarts its stopping test from the replaced
// residual (tol2 = delta^2 r.r), keeping the iteration count
__global__ void k_replace_state(const double *__restrict__ rz_part,
                                const double *__restrict__ rr_part, int np, double delta2,
                                PcgState *__restrict__ st) {
  __shared__ double sh[kThreads / 64];
  const double rz = total(rz_part, np, sh);
  const double rr = total(rr_part, np, sh);
  if (threadIdx.x == 0) {
    st->rz = rz;
    st->rr = rr;
    st->tol2 = delta2 * rr;
    st->status = MG_OK;
    st->done = 0;
  }
}

// ---------------------------------------------------------------------------
static int lanes_per_row(const System &S) {
  const double avg = (double)S.nnz / (double)S.nrows;
  return avg > 20.0 ? 8 : 4;
}

static int grid_rows(int64_t n, int lpr) {
  int64_t b = (n * lpr + kThreads - 1) / kThreads;
  return (int)(b < kMaxVecBlocks ? (b < 1 ? 1 : b) : kMaxVecBlocks);
}

static int grid_vec(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  return (int)(b < kMaxVecBlocks ? (b < 1 ? 1 : b) : kMaxVecBlocks);
}

void ensure_dinv(Ctx *c, int sys, int prec) {
  System &S = c->sys[sys];
  if (prec == MG_PREC_F64) {
    if (S.dinv64_ok) return;
    MG_CHECK(S.has64, MG_ERR_STATE, "no FP64 values");
    k_dinv<double><<<blocks_for(S.nrows, 256), 256, 0, c->stream>>>((int)S.nrows, S.diagpos,
                                                                    S.vals64, S.dinv64);
    S.dinv64_ok = true;
  } else {
    if (S.dinv32_ok) return;
    MG_CHECK(S.has32, MG_ERR_STATE, "no FP32 values");
    k_dinv<float><<<blocks_for(S.nrows, 256), 256, 0, c->stream>>>((int)S.nrows, S.diagpos,
                                                                   S.vals32, S.dinv32);
    S.dinv32_ok = true;
  }
  MG_HIP(hipGetLastError());
}

// workgroups of k_long_rows (0: no long rows)
static int grid_long(const System &S) { return S.n_long < 64 ? S.n_long : 64; }

// workgroups of the stream kernels; their partials and the long-row kernel's
// (npart in all) fit the kMaxBlocks-entry partial arrays
static int grid_blocks(const System &S) {
  const int cap = kMaxBlocks - grid_long(S);
  return S.n_sblk < cap ? (S.n_sblk < 1 ? 1 : S.n_sblk) : cap;
}
static int npart(const System &S) { return grid_blocks(S) + grid_long(S); }

// Calls f(C16, RL) with the system's column form and lanes per row as
// compile-time constants.
template <typename F>
static void stream_variant(const System &S, F f) {
  using T16 = std::integral_constant<bool, true>;
  using T32 = std::integral_constant<bool, false>;
  using L2 = std::integral_constant<int, 2>;
  using L4 = std::integral_constant<int, 4>;
  if (S.c16)
    S.srl == 2 ? f(T16{}, L2{}) : f(T16{}, L4{});
  else
    S.srl == 2 ? f(T32{}, L2{}) : f(T32{}, L4{});
}

template <typename T>
static void spmv_t(Ctx *c, System &S, const T *vals, const T *x, T *y) {
  stream_variant(S, [&](auto cf, auto rl) {
    k_spmv_stream<T, decltype(cf)::value, decltype(rl)::value>
        <<<grid_blocks(S), kThreads, 0, c->stream>>>(S.n_sblk, S.sdesc, S.rowptr, S.colidx, S.cofs,
                                                     vals, x, y);
  });
  if (S.n_long)
    k_long_rows<T, 0><<<grid_long(S), kThreads, 0, c->stream>>>(
        S.n_long, S.long_rows, S.rowptr, S.colidx, vals, x, nullptr, y, nullptr, nullptr);
  MG_HIP(hipGetLastError());
}

// q = A p with the p.q partials (npart(S) of them), in the system's column form
template <typename T>
static void spmv_pq(Ctx *c, const System &S, int gb, const T *vals, const T *p, T *q,
                    double *pq_part, const PcgState *st) {
  stream_variant(S, [&](auto cf, auto rl) {
    k_spmv_pq_stream<T, decltype(cf)::value, decltype(rl)::value><<<gb, kThreads, 0, c->stream>>>(
        S.n_sblk, S.sdesc, S.rowptr, S.colidx, S.cofs, vals, p, q, pq_part, st);
  });
  if (S.n_long)
    k_long_rows<T, 1><<<grid_long(S), kThreads, 0, c->stream>>>(
        S.n_long, S.long_rows, S.rowptr, S.colidx, vals, p, nullptr, q, pq_part + gb, st);
}

void spmv_launch(Ctx *c, int sys, int prec, const void *x, void *y) {
  System &S = c->sys[sys];
  if (prec == MG_PREC_F64)
    spmv_t<double>(c, S, S.vals64, (const double *)x, (double *)y);
  else
    spmv_t<float>(c, S, S.vals32, (const float *)x, (float *)y);
}

// convert FP64 values to FP32 when the FP32 copy was not assembled directly
__global__ void k_vals_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

static bool aligned16(std::initializer_list<const void *> ps) {
  for (const void *q : ps)
    if (reinterpret_cast<uintptr_t>(q) & 15) return false;
  return true;
}

// k_update / k_pupdate with 16-byte vector access when every array allows it
template <typename T>
static void vec_update(Ctx *c, int grid, int n, T *x, T *r, const T *p, const T *q, const T *dinv,
                       const double *pq_part, int npq, const double *rz_prev, int nprev,
                       double *rz_out, double *rr_out, PcgState *st) {
  if (aligned16({x, r, p, q, dinv}))
    k_update<T, true><<<grid, kThreads, 0, c->stream>>>(n, x, r, p, q, dinv, pq_part, npq, rz_prev,
                                                        nprev, rz_out, rr_out, st);
  else
    k_update<T, false><<<grid, kThreads, 0, c->stream>>>(n, x, r, p, q, dinv, pq_part, npq,
                                                         rz_prev, nprev, rz_out, rr_out, st);
}

template <typename T>
static void vec_pupdate(Ctx *c, int grid, int n, const T *r, T *p, const T *dinv,
                        const double *rz_new_part, const double *rz_part, const double *rr_part,
                        int np, int iter, PcgState *st) {
  if (aligned16({r, p, dinv}))
    k_pupdate<T, true><<<grid, kThreads, 0, c->stream>>>(n, r, p, dinv, rz_new_part, rz_part,
                                                         rr_part, np, iter, st);
  else
    k_pupdate<T, false><<<grid, kThreads, 0, c->stream>>>(n, r, p, dinv, rz_new_part, rz_part,
                                                          rr_part, np, iter, st);
}

struct Workspace {
  double *x64, *r64;
  float *b32, *x32, *r32, *p32, *q32;
  double *p64, *q64, *rr64;
  Partials part;
  PcgState *st;
};

static Workspace workspace(Ctx *c, int64_t n) {
  size_t need = sizeof(double) * 5 * n + sizeof(float) * 5 * n + sizeof(double) * 5 * kMaxBlocks +
                sizeof(PcgState) + 20 * 256;
  if (c->solve_ws_bytes < need) {
    dfree(c, c->solve_ws);
    c->solve_ws = dalloc<char>(c, need);
    c->solve_ws_bytes = need;
  }
  char *p = (char *)c->solve_ws;
  auto take = [&](size_t bytes) {
    char *q = p;
    p += (bytes + 255) & ~(size_t)255;
    return q;
  };
  Workspace w;
  w.x64 = (double *)take(sizeof(double) * n);
  w.r64 = (double *)take(sizeof(double) * n);
  w.p64 = (double *)take(sizeof(double) * n);
  w.q64 = (double *)take(sizeof(double) * n);
  w.rr64 = (double *)take(sizeof(double) * n);
  w.b32 = (float *)take(sizeof(float) * n);
  w.x32 = (float *)take(sizeof(float) * n);
  w.r32 = (float *)take(sizeof(float) * n);
  w.p32 = (float *)take(sizeof(float) * n);
  w.q32 = (float *)take(sizeof(float) * n);
  w.part.pq = (double *)take(sizeof(double) * kMaxBlocks);
  w.part.rz0 = (double *)take(sizeof(double) * kMaxBlocks);
  w.part.rz1 = (double *)take(sizeof(double) * kMaxBlocks);
  w.part.rr = (double *)take(sizeof(double) * kMaxBlocks);
  w.part.tmp = (double *)take(sizeof(double) * kMaxBlocks);
  w.st = (PcgState *)take(sizeof(PcgState));
  return w;
}

static PcgState read_state(Ctx *c, const PcgState *d) {
  PcgState h;
  MG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  return h;
}

__global__ void k_init_state(PcgState *st, double tol2, int maxit) {
  if (threadIdx.x == 0) {
    PcgState h{};
    h.tol2 = tol2;
    h.maxit = maxit;
    *st = h;
  }
}

static void init_state(Ctx *c, PcgState *d, double tol2, int maxit) {
  k_init_state<<<1, 64, 0, c->stream>>>(d, tol2, maxit);
  MG_HIP(hipGetLastError());
}

// Runs Jacobi-PCG on A (type T) from x (x_is_zero: x == 0), stopping when the
// recursive residual satisfies r.r <= tol2 or after maxit iterations.
template <typename T, int LPR>
static PcgState pcg_core(Ctx *c, System &S, const T *vals, const T *dinv, const T *b, T *x, T *r,
                         T *p, T *q, bool x_is_zero, double tol2, int maxit, int check_every,
                         Workspace &w) {
  const int n = (int)S.nrows;
  init_state(c, w.st, tol2, maxit);
  const int gr = grid_rows(n, LPR), gv = grid_vec(n), gb = grid_blocks(S);
  if (x_is_zero) MG_HIP(hipMemsetAsync(x, 0, sizeof(T) * n, c->stream));
  Partials &P = w.part;
  // k_start writes the r.z partials where iteration 0 reads its "previous" r.z
  k_start<T, LPR><<<gr, kThreads, 0, c->stream>>>(n, S.rowptr, S.colidx, vals,
                                                  x_is_zero ? nullptr : x, b, r, p, dinv, P.rz1,
                                                  P.tmp);
  // the vector kernels must see one partial per workgroup of the same grid
  k_start_finalize<<<1, kThreads, 0, c->stream>>>(P.rz1, P.tmp, gr, w.st);
  MG_HIP(hipGetLastError());
  PcgState h = read_state(c, w.st);
  if (h.done || maxit <= 0) return h;
  if (gr != gv) {
    // re-express the start partials on the vector grid: one k_dot-like pass is
    // not needed because k_update only reads the total; store it in slot 0
    k_start_partials_to<<<1, kThreads, 0, c->stream>>>(P.rz1, gr, gv);
  }
  const int chunk = check_every > 0 ? check_every : 32;
  int it = 0;
  while (true) {
    for (int k = 0; k < chunk; ++k, ++it) {
      double *rz_prev = (it & 1) ? P.rz0 : P.rz1;
      double *rz_new = (it & 1) ? P.rz1 : P.rz0;
      spmv_pq<T>(c, S, gb, vals, p, q, P.pq, w.st);
      vec_update<T>(c, gv, n, x, r, p, q, dinv, P.pq, npart(S), rz_prev, gv,
                                                  rz_new, P.rr, w.st);
      vec_pupdate<T>(c, gv, n, r, p, dinv, rz_new, rz_prev, P.rr, gv, it,
                                                   w.st);
    }
    MG_HIP(hipGetLastError());
    h = read_state(c, w.st);
    if (h.done) return h;
  }
}

template <typename T>
static PcgState pcg_dispatch(Ctx *c, System &S, const T *vals, const T *dinv, const T *b, T *x,
                             T *r, T *p, T *q, bool x_is_zero, double tol2, int maxit,
                             int check_every, Workspace &w) {
  if (lanes_per_row(S) == 8)
    return pcg_core<T, 8>(c, S, vals, dinv, b, x, r, p, q, x_is_zero, tol2, maxit, check_every, w);
  return pcg_core<T, 4>(c, S, vals, dinv, b, x, r, p, q, x_is_zero, tol2, maxit, check_every, w);
}

static double true_relres(Ctx *c, System &S, const double *b, const double *x, Workspace &w,
                          double bnorm);

// The mixed solvers' FP32 matrix: the FP32 assembly when it is at least as
// recent as the FP64 one (the two assemblies of one operator), otherwise the
// FP64 values rounded, so an FP64 re-assembly never leaves a stale FP32 copy.
static void refresh_f32(Ctx *c, System &S) {
  if (S.has32 && S.seq32 >= S.seq64 && S.key32 == S.key64) return;
  k_vals_to_f32<<<blocks_for(S.nnz, 256), 256, 0, c->stream>>>(S.nnz, S.vals64, S.vals32);
  MG_HIP(hipGetLastError());
  S.has32 = true;
  S.seq32 = S.seq64;
  S.key32 = S.key64;
  S.dinv32_ok = false;
}

// FP32 Jacobi-PCG whose residual is replaced by the FP64 one whenever the
// recursive residual has dropped by delta since the last replacement: the
// iterate is accumulated in FP64 (x += s x32), r = b - A x is recomputed with
// the FP64 matrix, and the recursion continues with the same search
// direction (rescaled), so no Krylov information is thrown away as it is by
// restarted refinement.  A replacement that does not reduce the true residual
// restarts the direction (p = D^{-1} r).  Returns the number of FP32
// iterations; *relres / *status / *outer as in pcg_solve.
template <int LPR>
static int pcg_reliable(Ctx *c, System &S, const double *b, double *x, double bnorm, double tol,
                        double delta, int maxit, int check_every, Workspace &w, double *relres,
                        int *status, int *outer) {
  const int n = (int)S.nrows;
  const int gr = grid_rows(n, LPR), gv = grid_vec(n), gb = grid_blocks(S);
  Partials &P = w.part;
  const int chunk = check_every > 0 ? check_every : 32;
  *outer = 0;
  *status = MG_ERR_NOCONV;
  double rel = true_relres(c, S, b, x, w, bnorm);  // w.rr64 = b - A x
  *relres = rel;
  if (rel <= tol) {
    *status = MG_OK;
    return 0;
  }
  double scale = rel * bnorm;  // r32 = r64 / scale, x += scale * x32
  auto start_direction = [&](int it) {
    // r32 = r64 / scale, p = D^{-1} r32; r.z partials where iteration `it` reads them
    k_to_f32<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, w.rr64, 1.0 / scale, w.b32);
    double *rz_prev = (it & 1) ? P.rz0 : P.rz1;
    k_start<float, LPR><<<gr, kThreads, 0, c->stream>>>(n, S.rowptr, S.colidx, S.vals32, nullptr,
                                                        w.b32, w.r32, w.p32, S.dinv32, rz_prev,
                                                        P.tmp);
    k_start_finalize<<<1, kThreads, 0, c->stream>>>(rz_prev, P.tmp, gr, w.st);
    if (gr != gv) k_start_partials_to<<<1, kThreads, 0, c->stream>>>(rz_prev, gr, gv);
    k_replace_state<<<1, kThreads, 0, c->stream>>>(rz_prev, P.tmp, gr, delta * delta, w.st);
    MG_HIP(hipGetLastError());
  };
  init_state(c, w.st, 0.0, maxit);
  MG_HIP(hipMemsetAsync(w.x32, 0, sizeof(float) * n, c->stream));
  start_direction(0);
  int it = 0;
  while (true) {
    for (int k = 0; k < chunk; ++k, ++it) {
      double *rz_prev = (it & 1) ? P.rz0 : P.rz1;
      double *rz_new = (it & 1) ? P.rz1 : P.rz0;
      spmv_pq<float>(c, S, gb, S.vals32, w.p32, w.q32, P.pq, w.st);
      vec_update<float>(c, gv, n, w.x32, w.r32, w.p32, w.q32, S.dinv32,
                                                      P.pq, npart(S), rz_prev, gv, rz_new, P.rr, w.st);
      vec_pupdate<float>(c, gv, n, w.r32, w.p32, S.dinv32, rz_new, rz_prev,
                                                       P.rr, gv, it, w.st);
    }
    MG_HIP(hipGetLastError());
    const PcgState h = read_state(c, w.st);
    if (!h.done) continue;
    it = h.iters;  // iterations actually performed (the rest of the chunk was skipped)
    // fold the FP32 increment into x and replace the residual
    k_axpy_f32<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, scale, w.x32, x);
    MG_HIP(hipMemsetAsync(w.x32, 0, sizeof(float) * n, c->stream));
    MG_HIP(hipGetLastError());
    const double prev = rel;
    rel = true_relres(c, S, b, x, w, bnorm);
    *relres = rel;
    ++*outer;
    if (rel <= tol) {
      *status = MG_OK;
      return it;
    }
    if (h.status == MG_ERR_BREAKDOWN && !(rel < prev)) {
      *status = MG_ERR_BREAKDOWN;
      return it;
    }
    if (it >= maxit) return it;
    const double ns = rel * bnorm;
    if (h.status == MG_OK && rel < 0.5 * prev) {
      // keep the direction.  The stopped iteration (it - 1) read r.z from
      // rz_old and wrote the recursive r.z where iteration `it` reads it;
      // that slot now takes the replaced residual's partials.
      k_to_f32<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, w.rr64, 1.0 / ns, w.r32);
      double *rz_next = (it & 1) ? P.rz0 : P.rz1;
      double *rz_old = (it & 1) ? P.rz1 : P.rz0;
      k_replace_partials<<<gv, kThreads, 0, c->stream>>>(n, w.r32, S.dinv32, rz_next, P.rr);
      k_replace_pupdate<<<gv, kThreads, 0, c->stream>>>(n, w.r32, w.p32, S.dinv32, rz_next,
                                                        rz_old, gv, scale / ns);
      k_replace_state<<<1, kThreads, 0, c->stream>>>(rz_next, P.rr, gv, delta * delta, w.st);
      MG_HIP(hipGetLastError());
      scale = ns;
    } else {
      // stagnation or breakdown of the recursion: restart the direction
      scale = ns;
      start_direction(it);
    }
  }
}

static double true_relres(Ctx *c, System &S, const double *b, const double *x, Workspace &w,
                          double bnorm) {
  const int n = (int)S.nrows;
  (void)n;
  const int g = grid_blocks(S);
  stream_variant(S, [&](auto cf, auto rl) {
    k_resid64_stream<decltype(cf)::value, decltype(rl)::value><<<g, kThreads, 0, c->stream>>>(
        S.n_sblk, S.sdesc, S.rowptr, S.colidx, S.cofs, S.vals64, x, b, w.rr64, w.part.tmp);
  });
  if (S.n_long)
    k_long_rows<double, 2><<<grid_long(S), kThreads, 0, c->stream>>>(
        S.n_long, S.long_rows, S.rowptr, S.colidx, S.vals64, x, b, w.rr64, w.part.tmp + g,
        nullptr);
  k_finalize_sum<<<1, kThreads, 0, c->stream>>>(w.part.tmp, npart(S), &w.st->rr);
  MG_HIP(hipGetLastError());
  PcgState h = read_state(c, w.st);
  return std::sqrt(h.rr) / bnorm;
}

static void pcg_solve(Ctx *c, const mg_solver_params *prm, const double *b, double *x,
                      mg_solve_info *info) {
  System &S = c->sys[MG_SYS_INTERIOR];
  MG_CHECK(S.built && S.has64, MG_ERR_STATE, "solve: interior FP64 matrix not assembled");
  const int64_t n = S.nrows;
  Workspace w = workspace(c, n);
  EventPair ev;
  MG_CHECK(ev.e0 && ev.e1, MG_ERR_HIP, "solve: hipEventCreate failed");
  MG_HIP(hipEventRecord(ev.e0, c->stream));
  ensure_dinv(c, MG_SYS_INTERIOR, MG_PREC_F64);
  // ||b||
  {
    init_state(c, w.st, 0.0, 0);
    k_dot_self<<<grid_vec(n), kThreads, 0, c->stream>>>((int)n, b, w.part.tmp);
    k_finalize_sum<<<1, kThreads, 0, c->stream>>>(w.part.tmp, grid_vec(n), &w.st->bb);
    MG_HIP(hipGetLastError());
  }
  double bnorm = std::sqrt(read_state(c, w.st).bb);
  if (bnorm == 0.0) bnorm = 1.0;
  const double tol = prm->tol > 0 ? prm->tol : 1e-10;
  const int maxit = prm->max_iters > 0 ? prm->max_iters : 100000;
  int iters = 0, outer = 0, status = MG_ERR_NOCONV;
  double relres = 0.0;
  if (!prm->use_x0) MG_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, c->stream));

  if (prm->mode == MG_SOLVE_F64) {
    bool zero = !prm->use_x0;
    for (int restart = 0; restart < 8; ++restart) {
      const double tol2 = tol * tol * bnorm * bnorm;
      PcgState h = pcg_dispatch<double>(c, S, S.vals64, S.dinv64, b, x, w.r64, w.p64, w.q64,
                                        zero, tol2, maxit - iters, prm->check_every, w);
      iters += h.iters;
      zero = false;
      relres = true_relres(c, S, b, x, w, bnorm);
      if (h.status == MG_ERR_BREAKDOWN && relres > tol) {
        status = MG_ERR_BREAKDOWN;
        break;
      }
      if (relres <= tol) {
        status = MG_OK;
        break;
      }
      if (iters >= maxit) break;
    }
  } else if (prm->mode == MG_SOLVE_F32_RELIABLE) {
    refresh_f32(c, S);
    ensure_dinv(c, MG_SYS_INTERIOR, MG_PREC_F32);
    const double delta = prm->inner_tol > 0 ? prm->inner_tol : 1e-2;
    if (lanes_per_row(S) == 8)
      iters = pcg_reliable<8>(c, S, b, x, bnorm, tol, delta, maxit, prm->check_every, w, &relres,
                              &status, &outer);
    else
      iters = pcg_reliable<4>(c, S, b, x, bnorm, tol, delta, maxit, prm->check_every, w, &relres,
                              &status, &outer);
  } else {
    refresh_f32(c, S);
    ensure_dinv(c, MG_SYS_INTERIOR, MG_PREC_F32);
    const double itol = prm->inner_tol > 0 ? prm->inner_tol : 1e-5;
    const int max_outer = prm->max_outer > 0 ? prm->max_outer : 30;
    for (outer = 0; outer <= max_outer; ++outer) {
      relres = true_relres(c, S, b, x, w, bnorm);  // w.rr64 = b - A x
      if (relres <= tol) {
        status = MG_OK;
        break;
      }
      if (outer == max_outer || iters >= maxit) break;
      const double rn = relres * bnorm;
      // the inner solve never goes below inner_tol, nor further than this step
      // needs to reach tol (with a factor 1/4 of margin)
      const double need = 0.25 * tol / relres;
      const double step_tol = need > itol ? (need < 0.5 ? need : 0.5) : itol;
      k_to_f32<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, w.rr64, 1.0 / rn, w.b32);
      MG_HIP(hipGetLastError());
      // inner: A32 d = r32 / ||r||, ||r32|| = 1
      PcgState h = pcg_dispatch<float>(c, S, S.vals32, S.dinv32, w.b32, w.x32, w.r32, w.p32,
                                       w.q32, true, step_tol * step_tol, maxit - iters,
                                       prm->check_every, w);
      iters += h.iters;
      if (h.status == MG_ERR_BREAKDOWN) {
        status = MG_ERR_BREAKDOWN;
        break;
      }
      k_axpy_f32<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, rn, w.x32, x);
      MG_HIP(hipGetLastError());
    }
  }
  MG_HIP(hipEventRecord(ev.e1, c->stream));
  MG_HIP(hipEventSynchronize(ev.e1));
  float ms = 0.f;
  MG_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
  if (info) {
    info->status = status;
    info->iters = iters;
    info->outer = outer;
    info->relres = relres;
    info->solve_ms = ms;
  }
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

int mg_spmv_device(mg_ctx *c, int sys, int prec, const void *x, void *y) {
  API_BEGIN(c)
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "spmv: bad system");
  MG_CHECK(x && y, MG_ERR_ARG, "spmv: null vector");
  System &S = c->sys[sys];
  MG_CHECK(S.built && (prec == MG_PREC_F64 ? S.has64 : S.has32), MG_ERR_STATE,
           "spmv: matrix not assembled");
  spmv_launch(c, sys, prec, x, y);
  API_END(c)
}

int mg_spmv(mg_ctx *c, int sys, int prec, const double *x, double *y) {
  API_BEGIN(c)
  MG_CHECK(sys == MG_SYS_FULL || sys == MG_SYS_INTERIOR, MG_ERR_ARG, "spmv: bad system");
  MG_CHECK(x && y, MG_ERR_ARG, "spmv: null vector");
  System &S = c->sys[sys];
  MG_CHECK(S.built && (prec == MG_PREC_F64 ? S.has64 : S.has32), MG_ERR_STATE,
           "spmv: matrix not assembled");
  const int64_t n = S.nrows;
  if (prec == MG_PREC_F64) {
    DevTmp<double> dx(c, n), dy(c, n);
    MG_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    spmv_launch(c, sys, prec, dx, dy);
    MG_HIP(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    MG_HIP(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> hx(n), hy(n);
    for (int64_t i = 0; i < n; ++i) hx[i] = (float)x[i];
    DevTmp<float> dx(c, n), dy(c, n);
    MG_HIP(hipMemcpyAsync(dx, hx.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    spmv_launch(c, sys, prec, dx, dy);
    MG_HIP(hipMemcpyAsync(hy.data(), dy, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    MG_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) y[i] = hy[i];
  }
  API_END(c)
}

int mg_pcg_solve_device(mg_ctx *c, const mg_solver_params *p, const double *b, double *x,
                        mg_solve_info *info) {
  API_BEGIN(c)
  MG_CHECK(p && b && x, MG_ERR_ARG, "solve: null argument");
  MG_CHECK(p->mode == MG_SOLVE_F64 || p->mode == MG_SOLVE_F32_REFINE ||
               p->mode == MG_SOLVE_F32_RELIABLE,
           MG_ERR_ARG, "solve: bad mode");
  // the status is returned whether or not the caller asked for the info
  mg_solve_info loc{};
  pcg_solve(c, p, b, x, &loc);
  if (info) *info = loc;
  if (loc.status != MG_OK) {
    c->err = loc.status == MG_ERR_NOCONV ? "solve: iteration limit reached"
                                         : "solve: CG breakdown (p.Ap <= 0 or NaN)";
    return loc.status;
  }
  API_END(c)
}

int mg_pcg_solve(mg_ctx *c, const mg_solver_params *p, const double *b, double *x,
                 mg_solve_info *info) {
  API_BEGIN(c)
  MG_CHECK(p && b && x, MG_ERR_ARG, "solve: null argument");
  MG_CHECK(p->mode == MG_SOLVE_F64 || p->mode == MG_SOLVE_F32_REFINE ||
               p->mode == MG_SOLVE_F32_RELIABLE,
           MG_ERR_ARG, "solve: bad mode");
  System &S = c->sys[MG_SYS_INTERIOR];
  MG_CHECK(S.built && S.has64, MG_ERR_STATE, "solve: interior FP64 matrix not assembled");
  const int64_t n = S.nrows;
  DevTmp<double> db(c, n), dx(c, n);
  MG_HIP(hipMemcpyAsync(db, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (p->use_x0)
    MG_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  mg_solve_info loc{};
  pcg_solve(c, p, db, dx, &loc);
  MG_HIP(hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  if (info) *info = loc;
  if (loc.status != MG_OK) {
    c->err = loc.status == MG_ERR_NOCONV ? "solve: iteration limit reached"
                                         : "solve: CG breakdown (p.Ap <= 0 or NaN)";
    return loc.status;
  }
  API_END(c)
}

}  // extern "C"
