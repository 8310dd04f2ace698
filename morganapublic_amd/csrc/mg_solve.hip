// CSR SpMV and Jacobi-preconditioned CG (FP64 plain, or FP32 inner solves
// inside an FP64 iterative refinement).
//
// One CG iteration is three kernels, all grid-stride over a fixed grid:
//   k_spmv_pq : q = A p and the block partials of p.q
//   k_update  : x += alpha p, r -= alpha q, partials of r.D^{-1}r and r.r
//   k_pupdate : p = D^{-1} r + beta p
// Scalars (alpha, beta, r.z, convergence flag) live in device memory: the last
// workgroup to finish a reduction (ticket counter) sums the per-workgroup
// partials in a fixed order and computes the next scalar, so the host only
// reads the state every `check_every` iterations and no reduction uses float
// atomics.  Once converged, the queued kernels return at entry.
#include <cmath>

#include "mg_internal.h"

namespace mg {

static constexpr int kThreads = 256;
static constexpr int kMaxBlocks = 2048;

struct PcgState {
  double rz, alpha, beta, rr, tol2, pq, bb;
  int iters, maxit, done, status;
  unsigned ticket[4];
};

__device__ __forceinline__ double ld_coherent(const double *p) {
  unsigned long long v = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)v);
}

// deterministic block sum of one double per thread (fixed tree)
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  __syncthreads();
  if (ln == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  return t;  // valid in thread 0
}

// Publishes this block's partial(s); returns true in the block that finished last.
__device__ __forceinline__ bool last_block(unsigned *ticket, int *sh_flag) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = atomicAdd(ticket, 1u);
    *sh_flag = (t == gridDim.x - 1);
  }
  __syncthreads();
  const bool last = *sh_flag;
  if (last) __threadfence();
  return last;
}

// sum of partial[0 .. n) in a fixed order by the whole block
__device__ __forceinline__ double reduce_partials(const double *partial, int n, double *sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += ld_coherent(partial + i);
  return block_sum(v, sh);
}

template <typename T, int LPR>
__device__ __forceinline__ T row_dot(int row, int lane, const int32_t *__restrict__ rowptr,
                                     const int32_t *__restrict__ colidx, const T *__restrict__ vals,
                                     const T *__restrict__ x) {
  const int k0 = rowptr[row], k1 = rowptr[row + 1];
  T s = (T)0;
  for (int k = k0 + lane; k < k1; k += LPR) s += vals[k] * x[colidx[k]];
#pragma unroll
  for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
  return s;
}

template <typename T, int LPR>
__global__ __launch_bounds__(kThreads) void k_spmv(int n, const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ colidx,
                                                   const T *__restrict__ vals,
                                                   const T *__restrict__ x, T *__restrict__ y) {
  const int lane = threadIdx.x % LPR;
  const int stride = gridDim.x * (kThreads / LPR);
  for (int row = blockIdx.x * (kThreads / LPR) + threadIdx.x / LPR; row < n; row += stride) {
    const T s = row_dot<T, LPR>(row, lane, rowptr, colidx, vals, x);
    if (lane == 0) y[row] = s;
  }
}

template <typename T, int LPR>
__global__ __launch_bounds__(kThreads) void k_spmv_pq(int n, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ colidx,
                                                      const T *__restrict__ vals,
                                                      const T *__restrict__ p, T *__restrict__ q,
                                                      double *__restrict__ partial,
                                                      PcgState *__restrict__ st) {
  __shared__ double sh[kThreads / 64];
  __shared__ int flag;
  if (st->done) return;
  const int lane = threadIdx.x % LPR;
  const int stride = gridDim.x * (kThreads / LPR);
  double acc = 0.0;
  for (int row = blockIdx.x * (kThreads / LPR) + threadIdx.x / LPR; row < n; row += stride) {
    const T s = row_dot<T, LPR>(row, lane, rowptr, colidx, vals, p);
    if (lane == 0) {
      q[row] = s;
      acc += (double)(s * p[row]);
    }
  }
  const double b = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
  if (last_block(&st->ticket[0], &flag)) {
    const double pq = reduce_partials(partial, gridDim.x, sh);
    if (threadIdx.x == 0) {
      st->ticket[0] = 0;
      st->pq = pq;
      if (!(pq > 0.0) || !isfinite(pq)) {
        st->status = MG_ERR_BREAKDOWN;
        st->done = 1;
      } else {
        st->alpha = st->rz / pq;
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_update(int n, T *__restrict__ x, T *__restrict__ r,
                                                     const T *__restrict__ p,
                                                     const T *__restrict__ q,
                                                     const T *__restrict__ dinv,
                                                     double *__restrict__ partial,
                                                     PcgState *__restrict__ st) {
  __shared__ double sh[kThreads / 64];
  __shared__ int flag;
  if (st->done) return;
  const T alpha = (T)st->alpha;
  double rz = 0.0, rr = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    rz += (double)(ri * (dinv[i] * ri));
    rr += (double)(ri * ri);
  }
  const double b1 = block_sum(rz, sh);
  const double b2 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = b1;
    partial[2 * blockIdx.x + 1] = b2;
  }
  if (last_block(&st->ticket[1], &flag)) {
    double v1 = 0.0, v2 = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
      v1 += ld_coherent(partial + 2 * i);
      v2 += ld_coherent(partial + 2 * i + 1);
    }
    const double s1 = block_sum(v1, sh);
    const double s2 = block_sum(v2, sh);
    if (threadIdx.x == 0) {
      st->ticket[1] = 0;
      st->iters += 1;
      st->rr = s2;
      if (!isfinite(s2)) {
        st->status = MG_ERR_BREAKDOWN;
        st->done = 1;
      } else if (s2 <= st->tol2) {
        st->status = MG_OK;
        st->done = 1;
      } else if (st->iters >= st->maxit) {
        st->status = MG_ERR_NOCONV;
        st->done = 1;
      }
      st->beta = s1 / st->rz;
      st->rz = s1;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_pupdate(int n, const T *__restrict__ r,
                                                      T *__restrict__ p,
                                                      const T *__restrict__ dinv,
                                                      const PcgState *__restrict__ st) {
  if (st->done) return;
  const T beta = (T)st->beta;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    p[i] = dinv[i] * r[i] + beta * p[i];
}

// r = b - A x (or r = b when x is null), p = D^{-1} r; st->rz = r.p, st->rr = r.r
template <typename T, int LPR>
__global__ __launch_bounds__(kThreads) void k_start(int n, const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colidx,
                                                    const T *__restrict__ vals,
                                                    const T *__restrict__ x,
                                                    const T *__restrict__ b, T *__restrict__ r,
                                                    T *__restrict__ p, const T *__restrict__ dinv,
                                                    double *__restrict__ partial,
                                                    PcgState *__restrict__ st) {
  __shared__ double sh[kThreads / 64];
  __shared__ int flag;
  const int lane = threadIdx.x % LPR;
  const int stride = gridDim.x * (kThreads / LPR);
  double rz = 0.0, rr = 0.0;
  for (int row = blockIdx.x * (kThreads / LPR) + threadIdx.x / LPR; row < n; row += stride) {
    T ax = (T)0;
    if (x) ax = row_dot<T, LPR>(row, lane, rowptr, colidx, vals, x);
    if (lane == 0) {
      const T ri = b[row] - ax;
      const T zi = dinv[row] * ri;
      r[row] = ri;
      p[row] = zi;
      rz += (double)(ri * zi);
      rr += (double)(ri * ri);
    }
  }
  const double b1 = block_sum(rz, sh);
  const double b2 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = b1;
    partial[2 * blockIdx.x + 1] = b2;
  }
  if (last_block(&st->ticket[2], &flag)) {
    double v1 = 0.0, v2 = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
      v1 += ld_coherent(partial + 2 * i);
      v2 += ld_coherent(partial + 2 * i + 1);
    }
    const double s1 = block_sum(v1, sh);
    const double s2 = block_sum(v2, sh);
    if (threadIdx.x == 0) {
      st->ticket[2] = 0;
      st->rz = s1;
      st->rr = s2;
      st->done = (s2 <= st->tol2) ? 1 : 0;
      st->status = MG_OK;
    }
  }
}

// FP64 residual r = b - A x with st->rr = r.r (refinement and final check)
template <int LPR>
__global__ __launch_bounds__(kThreads) void k_resid64(int n, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ colidx,
                                                      const double *__restrict__ vals,
                                                      const double *__restrict__ x,
                                                      const double *__restrict__ b,
                                                      double *__restrict__ r,
                                                      double *__restrict__ partial,
                                                      PcgState *__restrict__ st) {
  __shared__ double sh[kThreads / 64];
  __shared__ int flag;
  const int lane = threadIdx.x % LPR;
  const int stride = gridDim.x * (kThreads / LPR);
  double rr = 0.0;
  for (int row = blockIdx.x * (kThreads / LPR) + threadIdx.x / LPR; row < n; row += stride) {
    const double ax = row_dot<double, LPR>(row, lane, rowptr, colidx, vals, x);
    if (lane == 0) {
      const double ri = b[row] - ax;
      r[row] = ri;
      rr += ri * ri;
    }
  }
  const double s = block_sum(rr, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
  if (last_block(&st->ticket[3], &flag)) {
    const double t = reduce_partials(partial, gridDim.x, sh);
    if (threadIdx.x == 0) {
      st->ticket[3] = 0;
      st->rr = t;
    }
  }
}

__global__ void k_dot_self(int n, const double *__restrict__ v, double *__restrict__ partial,
                           PcgState *__restrict__ st) {
  __shared__ double sh[kThreads / 64];
  __shared__ int flag;
  double a = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    a += v[i] * v[i];
  const double s = block_sum(a, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
  if (last_block(&st->ticket[3], &flag)) {
    const double t = reduce_partials(partial, gridDim.x, sh);
    if (threadIdx.x == 0) {
      st->ticket[3] = 0;
      st->bb = t;
    }
  }
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

// ---------------------------------------------------------------------------
static int lanes_per_row(const System &S) {
  const double avg = (double)S.nnz / (double)S.nrows;
  return avg > 20.0 ? 8 : 4;
}

static int grid_rows(int64_t n, int lpr) {
  int64_t b = (n * lpr + kThreads - 1) / kThreads;
  return (int)(b < kMaxBlocks ? (b < 1 ? 1 : b) : kMaxBlocks);
}

static int grid_vec(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  return (int)(b < kMaxBlocks ? (b < 1 ? 1 : b) : kMaxBlocks);
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

template <typename T>
static void spmv_t(Ctx *c, System &S, const T *vals, const T *x, T *y) {
  const int n = (int)S.nrows;
  if (lanes_per_row(S) == 8)
    k_spmv<T, 8><<<grid_rows(n, 8), kThreads, 0, c->stream>>>(n, S.rowptr, S.colidx, vals, x, y);
  else
    k_spmv<T, 4><<<grid_rows(n, 4), kThreads, 0, c->stream>>>(n, S.rowptr, S.colidx, vals, x, y);
  MG_HIP(hipGetLastError());
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

struct Workspace {
  double *x64, *r64;
  float *b32, *x32, *r32, *p32, *q32;
  double *p64, *q64, *rr64;
  double *partial;
  PcgState *st;
};

static Workspace workspace(Ctx *c, int64_t n) {
  size_t need = sizeof(double) * 5 * n + sizeof(float) * 5 * n + sizeof(double) * 2 * kMaxBlocks +
                sizeof(PcgState) + 16 * 256;
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
  w.partial = (double *)take(sizeof(double) * 2 * kMaxBlocks);
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
  const int gr = grid_rows(n, LPR), gv = grid_vec(n);
  if (x_is_zero) MG_HIP(hipMemsetAsync(x, 0, sizeof(T) * n, c->stream));
  k_start<T, LPR><<<gr, kThreads, 0, c->stream>>>(n, S.rowptr, S.colidx, vals,
                                                  x_is_zero ? nullptr : x, b, r, p, dinv,
                                                  w.partial, w.st);
  MG_HIP(hipGetLastError());
  PcgState h = read_state(c, w.st);
  if (h.done || maxit <= 0) return h;
  const int chunk = check_every > 0 ? check_every : 32;
  while (true) {
    for (int k = 0; k < chunk; ++k) {
      k_spmv_pq<T, LPR><<<gr, kThreads, 0, c->stream>>>(n, S.rowptr, S.colidx, vals, p, q,
                                                        w.partial, w.st);
      k_update<T><<<gv, kThreads, 0, c->stream>>>(n, x, r, p, q, dinv, w.partial, w.st);
      k_pupdate<T><<<gv, kThreads, 0, c->stream>>>(n, r, p, dinv, w.st);
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
                          double bnorm) {
  const int n = (int)S.nrows;
  if (lanes_per_row(S) == 8)
    k_resid64<8><<<grid_rows(n, 8), kThreads, 0, c->stream>>>(n, S.rowptr, S.colidx, S.vals64, x,
                                                              b, w.rr64, w.partial, w.st);
  else
    k_resid64<4><<<grid_rows(n, 4), kThreads, 0, c->stream>>>(n, S.rowptr, S.colidx, S.vals64, x,
                                                              b, w.rr64, w.partial, w.st);
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
  hipEvent_t e0, e1;
  MG_HIP(hipEventCreate(&e0));
  MG_HIP(hipEventCreate(&e1));
  MG_HIP(hipEventRecord(e0, c->stream));
  ensure_dinv(c, MG_SYS_INTERIOR, MG_PREC_F64);
  // ||b||
  {
    init_state(c, w.st, 0.0, 0);
    k_dot_self<<<grid_vec(n), kThreads, 0, c->stream>>>((int)n, b, w.partial, w.st);
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
  } else {
    if (!S.has32) {
      k_vals_to_f32<<<blocks_for(S.nnz, 256), 256, 0, c->stream>>>(S.nnz, S.vals64, S.vals32);
      MG_HIP(hipGetLastError());
      S.has32 = true;
      S.dinv32_ok = false;
    }
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
      k_to_f32<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, w.rr64, 1.0 / rn, w.b32);
      MG_HIP(hipGetLastError());
      // inner: A32 d = r32 / ||r||, ||r32|| = 1
      PcgState h = pcg_dispatch<float>(c, S, S.vals32, S.dinv32, w.b32, w.x32, w.r32, w.p32,
                                       w.q32, true, itol * itol, maxit - iters, prm->check_every,
                                       w);
      iters += h.iters;
      if (h.status == MG_ERR_BREAKDOWN) {
        status = MG_ERR_BREAKDOWN;
        break;
      }
      k_axpy_f32<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, rn, w.x32, x);
      MG_HIP(hipGetLastError());
    }
  }
  MG_HIP(hipEventRecord(e1, c->stream));
  MG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MG_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
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
    double *dx = dalloc<double>(c, n), *dy = dalloc<double>(c, n);
    MG_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    spmv_launch(c, sys, prec, dx, dy);
    MG_HIP(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    MG_HIP(hipStreamSynchronize(c->stream));
    dfree(c, dx);
    dfree(c, dy);
  } else {
    std::vector<float> hx(n), hy(n);
    for (int64_t i = 0; i < n; ++i) hx[i] = (float)x[i];
    float *dx = dalloc<float>(c, n), *dy = dalloc<float>(c, n);
    MG_HIP(hipMemcpyAsync(dx, hx.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    spmv_launch(c, sys, prec, dx, dy);
    MG_HIP(hipMemcpyAsync(hy.data(), dy, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    MG_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) y[i] = hy[i];
    dfree(c, dx);
    dfree(c, dy);
  }
  API_END(c)
}

int mg_pcg_solve_device(mg_ctx *c, const mg_solver_params *p, const double *b, double *x,
                        mg_solve_info *info) {
  API_BEGIN(c)
  MG_CHECK(p && b && x, MG_ERR_ARG, "solve: null argument");
  MG_CHECK(p->mode == MG_SOLVE_F64 || p->mode == MG_SOLVE_F32_REFINE, MG_ERR_ARG,
           "solve: bad mode");
  pcg_solve(c, p, b, x, info);
  if (info && info->status != MG_OK) {
    c->err = info->status == MG_ERR_NOCONV ? "solve: iteration limit reached"
                                           : "solve: CG breakdown (p.Ap <= 0 or NaN)";
    return info->status;
  }
  API_END(c)
}

int mg_pcg_solve(mg_ctx *c, const mg_solver_params *p, const double *b, double *x,
                 mg_solve_info *info) {
  API_BEGIN(c)
  MG_CHECK(p && b && x, MG_ERR_ARG, "solve: null argument");
  MG_CHECK(p->mode == MG_SOLVE_F64 || p->mode == MG_SOLVE_F32_REFINE, MG_ERR_ARG,
           "solve: bad mode");
  System &S = c->sys[MG_SYS_INTERIOR];
  MG_CHECK(S.built, MG_ERR_STATE, "solve: interior pattern not built");
  const int64_t n = S.nrows;
  double *db = dalloc<double>(c, n), *dx = dalloc<double>(c, n);
  MG_HIP(hipMemcpyAsync(db, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (p->use_x0)
    MG_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  mg_solve_info loc{};
  pcg_solve(c, p, db, dx, &loc);
  MG_HIP(hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  MG_HIP(hipStreamSynchronize(c->stream));
  dfree(c, db);
  dfree(c, dx);
  if (info) *info = loc;
  if (loc.status != MG_OK) {
    c->err = loc.status == MG_ERR_NOCONV ? "solve: iteration limit reached"
                                         : "solve: CG breakdown (p.Ap <= 0 or NaN)";
    return loc.status;
  }
  API_END(c)
}

}  // extern "C"
