// Sparse matrix-vector products and the preconditioned conjugate-gradient
// solver for gfx950: all-FP64, or mixed precision (FP32 matrix values and
// FP32 Krylov vectors, with the solution and the true residual in FP64).
//
// Vectors inside the solver are interleaved [n][NRP] (NRP = right-hand sides
// padded to a power of two), so one pass over the matrix serves every
// right-hand side and the gathers of x read NRP consecutive values.
//
// Reductions are deterministic and need no atomics or device-scope fences:
// each block of a reducing kernel writes its partial sums, and the kernel
// that consumes the result (the next one on the stream, which sees every
// write of the previous kernel) adds the partials in block order; every block
// of the consumer forms the same sum.
//
// Mixed precision uses reliable updates: the FP32 CG recurrence keeps its
// search direction across corrections; whenever its recursive residual has
// dropped by a factor delta, the FP32 correction is added to the FP64
// solution, the residual is replaced by the true FP64 residual b - K x, and
// the direction is re-orthogonalised against it. Iteration counts therefore
// stay close to those of FP64 CG, while every iteration streams FP32 values.
#include "tfm_common.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <type_traits>
#include <vector>

namespace tfm {

constexpr int BS = 256;   // threads per block of the solver kernels
constexpr int RB = 1024;  // max blocks of a reducing kernel (= partials per reduction)
constexpr int MAXR = 16;  // right-hand sides per solve
constexpr int NSCAL = 4;  // per-iteration scalars: rz, pq, rr, spare
enum { S_RZ = 0, S_PQ = 1, S_RR = 2 };

// ---------------------------------------------------------- reductions
// Block partials: thread t contributes v[k] to right-hand side t % NRP; the
// block's sums go to part[blockIdx.x][k][r].
template <int NV, int NRP>
__device__ __forceinline__ void write_block_partials(const double (&v)[NV], double *sh,
                                                     double *part) {
  const int tid = threadIdx.x;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) sh[k * BS + tid] = v[k];
  __syncthreads();
  for (int s = BS / 2; s >= NRP; s >>= 1) {
    if (tid < s)
#pragma unroll
      for (int k = 0; k < NV; ++k) sh[k * BS + tid] += sh[k * BS + tid + s];
    __syncthreads();
  }
  if (tid < NV * NRP) {
    const int k = tid / NRP, r = tid % NRP;
    part[(size_t)blockIdx.x * NV * NRP + tid] = sh[k * BS + r];
  }
}

// Sums the NVT partial slots written by nblk blocks (in block order) into
// out[0..NVT) in LDS; identical in every block that calls it.
template <int NVT>
__device__ __forceinline__ void sum_partials(const double *part, int nblk, double *sh,
                                             double *out) {
  constexpr int CH = BS / NVT;
  const int tid = threadIdx.x;
  double acc = 0.0;
  if (tid < CH * NVT) {
    const int slot = tid % NVT, c = tid / NVT;
    for (int b = c; b < nblk; b += CH) acc += part[(size_t)b * NVT + slot];
  }
  __syncthreads();
  sh[tid] = acc;
  __syncthreads();
  if (tid < NVT) {
    double t = 0.0;
    for (int c = 0; c < CH; ++c) t += sh[c * NVT + tid];
    out[tid] = t;
  }
  __syncthreads();
}

// ------------------------------------------------------- node-block SpMV
// A team of TL lanes per node; lane l owns the neighbour blocks p = l, l+TL,
// ...: it loads adj[p], the ND x NRP entries of x at that neighbour, and the
// ND x ND block of values (ND contiguous values per row), so each lane does
// ND*ND*NRP FMAs per one index load. Team sums are butterfly-reduced.
// Optionally writes per-block partials of dot = sum_i y_i x_i per rhs.
template <typename T, int ND, int TL, int NRP>
__global__ void __launch_bounds__(BS) k_spmv(tfm_mesh_t m, const T *__restrict__ val,
                                             const T *__restrict__ x, T *__restrict__ y,
                                             const int *alldone, double *part_pq) {
  if (alldone && *alldone) return;
  constexpr int TPB = BS / TL;
  __shared__ double sh[BS > TPB * NRP ? BS : TPB * NRP];
  const int tid = threadIdx.x, lane = tid % TL, team = tid / TL;
  double dot[NRP];
#pragma unroll
  for (int r = 0; r < NRP; ++r) dot[r] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * TPB;
  for (int64_t n = (int64_t)blockIdx.x * TPB + team; n < m.n_nod; n += stride) {
    const int a0 = m.adj_ptr[n];
    const int deg = m.adj_ptr[n + 1] - a0;
    const T *vn = val + m.val_ptr[n];
    T acc[ND][NRP];
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int r = 0; r < NRP; ++r) acc[a][r] = T(0);
    if (deg > 0) {
      const int L = ND * deg;
      for (int p = lane; p < deg; p += TL) {
        const int64_t col = (int64_t)m.adj[a0 + p] * ND;
        T xv[ND][NRP];
#pragma unroll
        for (int b = 0; b < ND; ++b)
#pragma unroll
          for (int r = 0; r < NRP; ++r) xv[b][r] = x[(col + b) * NRP + r];
#pragma unroll
        for (int a = 0; a < ND; ++a)
#pragma unroll
          for (int b = 0; b < ND; ++b) {
            const T v = vn[a * L + p * ND + b];
#pragma unroll
            for (int r = 0; r < NRP; ++r) acc[a][r] += v * xv[b][r];
          }
      }
    } else if (lane < ND) {  // orphan node: only its diagonal is stored
#pragma unroll
      for (int a = 0; a < ND; ++a)
        if (a == lane)
#pragma unroll
          for (int r = 0; r < NRP; ++r) acc[a][r] = vn[a] * x[(n * ND + a) * NRP + r];
    }
#pragma unroll
    for (int off = 1; off < TL; off <<= 1)
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int r = 0; r < NRP; ++r) acc[a][r] += __shfl_xor(acc[a][r], off, TL);
    if (lane == 0) {
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int r = 0; r < NRP; ++r) {
          const int64_t o = (n * ND + a) * NRP + r;
          y[o] = acc[a][r];
          if (part_pq) dot[r] += (double)acc[a][r] * (double)x[o];
        }
    }
  }
  if (!part_pq) return;
  // team leaders hold the dots; regroup so thread t carries rhs t % NRP
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int r = 0; r < NRP; ++r) sh[team * NRP + r] = dot[r];
  __syncthreads();
  double v[1] = {tid < TPB * NRP ? sh[tid] : 0.0};
  write_block_partials<1, NRP>(v, sh, part_pq);
}

// ------------------------------------------------------- general CSR SpMV
template <int TL>
__global__ void __launch_bounds__(BS) k_spmv_csr(int64_t n, const int64_t *rowptr,
                                                 const int32_t *col, const double *val,
                                                 const double *x, double *y, const int *alldone,
                                                 double *part_pq) {
  if (alldone && *alldone) return;
  constexpr int TPB = BS / TL;
  __shared__ double sh[BS];
  const int tid = threadIdx.x, lane = tid % TL, team = tid / TL;
  double dot = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + team; i < n; i += (int64_t)gridDim.x * TPB) {
    double acc = 0.0;
    for (int64_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += TL) acc += val[k] * x[col[k]];
#pragma unroll
    for (int off = 1; off < TL; off <<= 1) acc += __shfl_xor(acc, off, TL);
    if (lane == 0) {
      y[i] = acc;
      dot += acc * x[i];
    }
  }
  if (!part_pq) return;
  double v[1] = {dot};
  write_block_partials<1, 1>(v, sh, part_pq);
}

// ------------------------------------------------------- preconditioner
// Inverse of the nodal diagonal blocks (block Jacobi) or of the diagonal
// (Jacobi) in the inner precision; a zero pivot keeps the identity.
template <typename T, int ND>
__global__ void k_precond(int64_t n_nod, const double *dblk, int pc, T *minv) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nod) return;
  constexpr int ND2 = ND * ND;
  double A[ND][ND], R[ND][ND];
  for (int a = 0; a < ND; ++a)
    for (int b = 0; b < ND; ++b) {
      A[a][b] = dblk[n * ND2 + a * ND + b];
      R[a][b] = a == b ? 1.0 : 0.0;
    }
  if (pc == TFM_PC_JACOBI || ND == 1) {
    for (int a = 0; a < ND; ++a)
      for (int b = 0; b < ND; ++b) R[a][b] = (a == b && A[a][a] != 0.0) ? 1.0 / A[a][a] : (a == b);
  } else {
    // Gauss-Jordan with partial pivoting on the small SPD block
    for (int c = 0; c < ND; ++c) {
      int piv = c;
      for (int r = c + 1; r < ND; ++r)
        if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
      if (A[piv][c] == 0.0) continue;
      if (piv != c)
        for (int k = 0; k < ND; ++k) {
          double t = A[c][k]; A[c][k] = A[piv][k]; A[piv][k] = t;
          t = R[c][k]; R[c][k] = R[piv][k]; R[piv][k] = t;
        }
      const double inv = 1.0 / A[c][c];
      for (int k = 0; k < ND; ++k) { A[c][k] *= inv; R[c][k] *= inv; }
      for (int r = 0; r < ND; ++r)
        if (r != c) {
          const double f = A[r][c];
          for (int k = 0; k < ND; ++k) { A[r][k] -= f * A[c][k]; R[r][k] -= f * R[c][k]; }
        }
    }
  }
  for (int a = 0; a < ND; ++a)
    for (int b = 0; b < ND; ++b) minv[n * ND2 + a * ND + b] = (T)R[a][b];
}

template <typename T, int ND, int NRP>
__device__ __forceinline__ void apply_pc(const T *minv, const T *r, int64_t n, int rr,
                                         T (&z)[ND], T (&rv)[ND]) {
#pragma unroll
  for (int a = 0; a < ND; ++a) rv[a] = r[(n * ND + a) * NRP + rr];
#pragma unroll
  for (int a = 0; a < ND; ++a) {
    T acc = T(0);
#pragma unroll
    for (int b = 0; b < ND; ++b) acc += minv[n * ND * ND + a * ND + b] * rv[b];
    z[a] = acc;
  }
}

// ------------------------------------------------------------ PCG steps
struct PcgState {
  double *scal;      // [(max_iter + 2) * NRP * NSCAL]
  double *tol2;      // [NRP] stop the current segment when rr <= tol2 (< 0: never live)
  int *done;         // [NRP] iteration at which rhs r stopped, -1 while running
  int *alldone;      // 1 once every rhs stopped
  double *part_pq;   // [RB * NRP]        written by the SpMV
  double *part_upd;  // [RB * 2 * NRP]    written by the update
  double *part_aux;  // [RB * 3 * NRP]    start / reliable-update reductions
};

// z = M r; p = z; partials of (r.z, r.r) for a (re)start from the residual in r.
template <typename T, int ND, int NRP>
__global__ void __launch_bounds__(BS) k_pcg_start(int64_t n_nod, const T *r, T *z, T *p,
                                                  const T *minv, PcgState st) {
  __shared__ double sh[2 * BS];
  const int tid = threadIdx.x;
  double acc[2] = {0.0, 0.0};
  const int64_t total = n_nod * NRP;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int64_t n = t / NRP;
    const int rr = (int)(t % NRP);
    T rv[ND], zv[ND];
    apply_pc<T, ND, NRP>(minv, r, n, rr, zv, rv);
#pragma unroll
    for (int a = 0; a < ND; ++a) {
      z[(n * ND + a) * NRP + rr] = zv[a];
      p[(n * ND + a) * NRP + rr] = zv[a];
      acc[0] += (double)rv[a] * (double)zv[a];
      acc[1] += (double)rv[a] * (double)rv[a];
    }
  }
  write_block_partials<2, NRP>(acc, sh, st.part_aux);
}

// One block: rz and rr of the (re)start at iteration `it`; live flags.
template <int NRP>
__global__ void __launch_bounds__(BS) k_pcg_start_finish(int it, int nblk, PcgState st) {
  __shared__ double sh[BS];
  __shared__ double tot[2 * NRP];
  sum_partials<2 * NRP>(st.part_aux, nblk, sh, tot);
  const int tid = threadIdx.x;
  __shared__ int live[NRP];
  if (tid < NRP) {
    double *sc = st.scal + ((size_t)it * NRP + tid) * NSCAL;
    sc[S_RZ] = tot[tid];
    sc[S_RR] = tot[NRP + tid];
    const bool lv = st.tol2[tid] >= 0.0 && tot[NRP + tid] > st.tol2[tid];
    st.done[tid] = lv ? -1 : it;
    live[tid] = lv;
  }
  __syncthreads();
  if (tid == 0) {
    int any = 0;
    for (int r = 0; r < NRP; ++r) any |= live[r];
    *st.alldone = !any;
  }
}

// alpha = rz/pq; x += alpha p; r -= alpha q; z = M r; partials of (r.z, r.r).
template <typename T, int ND, int NRP>
__global__ void __launch_bounds__(BS) k_pcg_update(int64_t n_nod, int it, T *x, T *r, T *z,
                                                   const T *p, const T *q, const T *minv,
                                                   PcgState st, int n_pq) {
  if (*st.alldone) return;
  __shared__ double sh[2 * BS];
  __shared__ double pq[NRP];
  __shared__ double alpha[NRP];
  sum_partials<NRP>(st.part_pq, n_pq, sh, pq);
  const int tid = threadIdx.x;
  if (tid < NRP) {
    double *sc = st.scal + ((size_t)it * NRP + tid) * NSCAL;
    alpha[tid] = (st.done[tid] < 0 && pq[tid] != 0.0) ? sc[S_RZ] / pq[tid] : 0.0;
    if (blockIdx.x == 0) sc[S_PQ] = pq[tid];
  }
  __syncthreads();
  double acc[2] = {0.0, 0.0};
  const int64_t total = n_nod * NRP;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int64_t n = t / NRP;
    const int rr = (int)(t % NRP);
    const T al = (T)alpha[rr];
#pragma unroll
    for (int a = 0; a < ND; ++a) {
      const int64_t o = (n * ND + a) * NRP + rr;
      x[o] += al * p[o];
      r[o] -= al * q[o];
    }
    T rv[ND], zv[ND];
    apply_pc<T, ND, NRP>(minv, r, n, rr, zv, rv);
#pragma unroll
    for (int a = 0; a < ND; ++a) {
      z[(n * ND + a) * NRP + rr] = zv[a];
      acc[0] += (double)rv[a] * (double)zv[a];
      acc[1] += (double)rv[a] * (double)rv[a];
    }
  }
  write_block_partials<2, NRP>(acc, sh, st.part_upd);
}

// beta = rz_{it+1}/rz_it; stop test on rr_{it+1}; p = z + beta p. Block 0
// records the scalars and the stop flags for the host and the next kernels.
template <typename T, int ND, int NRP>
__global__ void __launch_bounds__(BS) k_pcg_dir(int64_t n_nod, int it, const T *z, T *p,
                                                PcgState st, int n_upd) {
  if (*st.alldone) return;
  __shared__ double sh[BS];
  __shared__ double tot[2 * NRP];
  __shared__ double beta[NRP];
  __shared__ int live[NRP];
  sum_partials<2 * NRP>(st.part_upd, n_upd, sh, tot);
  const int tid = threadIdx.x;
  if (tid < NRP) {
    const double rz1 = tot[tid], rr1 = tot[NRP + tid];
    const double rz0 = st.scal[((size_t)it * NRP + tid) * NSCAL + S_RZ];
    const bool was = st.done[tid] >= 0;
    const bool now = !was && (rr1 <= st.tol2[tid] || !(rr1 == rr1));
    live[tid] = !was && !now;
    beta[tid] = rz0 != 0.0 ? rz1 / rz0 : 0.0;
    if (blockIdx.x == 0) {
      double *sc = st.scal + ((size_t)(it + 1) * NRP + tid) * NSCAL;
      sc[S_RZ] = rz1;
      sc[S_RR] = rr1;
      if (now) st.done[tid] = it + 1;
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid == 0) {
    int any = 0;
    for (int r = 0; r < NRP; ++r) any |= live[r];
    *st.alldone = !any;
  }
  const int64_t total = n_nod * ND * NRP;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int rr = (int)(t % NRP);
    if (live[rr]) p[t] = z[t] + (T)beta[rr] * p[t];
  }
}

// Reliable update, part 1: z = M r for the replaced residuals; partials of
// (r.z, r.p, r.r).
template <typename T, int ND, int NRP>
__global__ void __launch_bounds__(BS) k_reliable_dots(int64_t n_nod, const T *r, T *z, const T *p,
                                                      const T *minv, PcgState st) {
  __shared__ double sh[3 * BS];
  const int tid = threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  const int64_t total = n_nod * NRP;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int64_t n = t / NRP;
    const int rr = (int)(t % NRP);
    T rv[ND], zv[ND];
    apply_pc<T, ND, NRP>(minv, r, n, rr, zv, rv);
#pragma unroll
    for (int a = 0; a < ND; ++a) {
      const int64_t o = (n * ND + a) * NRP + rr;
      z[o] = zv[a];
      acc[0] += (double)rv[a] * (double)zv[a];
      acc[1] += (double)rv[a] * (double)p[o];
      acc[2] += (double)rv[a] * (double)rv[a];
    }
  }
  write_block_partials<3, NRP>(acc, sh, st.part_aux);
}

// Reliable update, part 2 (one block): rz of the iteration at which each
// continuing rhs stopped, the denominator of its next beta (0: restart).
// A rhs in `restart` restarts from the new residual (beta = 0, p = z).
template <int NRP>
__global__ void k_reliable_prep(int mask, int restart, int seg_start, PcgState st, double *rzold) {
  const int r = threadIdx.x;
  if (r >= NRP) return;
  const int d = st.done[r];
  double v = 0.0;
  if (((mask >> r) & 1) && !((restart >> r) & 1) && d > seg_start)
    v = st.scal[((size_t)(d - 1) * NRP + r) * NSCAL + S_RZ];
  rzold[r] = v;
}

// Reliable update, part 3: for the continuing rhs (bit set in `mask`) the kept
// direction is made orthogonal to the new residual, p' = p - (r.p / r.z) z,
// and the interrupted direction update is completed, p = z + beta p' with
// beta = (r.z) / rz_old; the scalar recurrence restarts at iteration `it`.
template <typename T, int ND, int NRP>
__global__ void __launch_bounds__(BS) k_reliable_apply(int64_t n_nod, int it, int mask,
                                                       const T *z, T *p, PcgState st,
                                                       const double *rzold, int nblk) {
  __shared__ double sh[BS];
  __shared__ double tot[3 * NRP];
  __shared__ double cz[NRP], cp[NRP];
  __shared__ int live[NRP];
  sum_partials<3 * NRP>(st.part_aux, nblk, sh, tot);
  const int tid = threadIdx.x;
  if (tid < NRP) {
    const bool lv = (mask >> tid) & 1;
    const double rz = tot[tid], rp = tot[NRP + tid];
    double beta = (lv && rzold[tid] != 0.0) ? rz / rzold[tid] : 0.0;
    double c = (lv && rz != 0.0) ? rp / rz : 0.0;
    if (!(beta == beta) || !(c == c)) beta = c = 0.0;
    live[tid] = lv;
    cz[tid] = 1.0 - beta * c;  // p = cz z + cp p
    cp[tid] = beta;
    if (blockIdx.x == 0) {
      double *sc = st.scal + ((size_t)it * NRP + tid) * NSCAL;
      if (lv) {
        sc[S_RZ] = rz;
        sc[S_RR] = tot[2 * NRP + tid];
      }
      st.done[tid] = lv ? -1 : (st.done[tid] >= 0 ? st.done[tid] : it);
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid == 0) *st.alldone = mask == 0;
  const int64_t total = n_nod * ND * NRP;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int rr = (int)(t % NRP);
    if (live[rr]) p[t] = (T)cz[rr] * z[t] + (T)cp[rr] * p[t];
  }
}

// ------------------------------------------------------- layout helpers
// dst[i*NRP + r] = zero ? 0 : (T) src[i]
template <typename T, int NRP>
__global__ void k_set_column(int64_t n, int r, const double *src, T *dst, int zero) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i * NRP + r] = zero ? T(0) : (T)src[i];
}

// xo[i] += (double) xi[i*NRP + r]; xi[i*NRP + r] = 0
template <typename T, int NRP>
__global__ void k_take_column(int64_t n, int r, T *xi, double *xo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    xo[i] += (double)xi[i * NRP + r];
    xi[i * NRP + r] = T(0);
  }
}

template <typename T>
__global__ void k_fill_zero(int64_t n, T *dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = T(0);
}

// Matrix-free Dirichlet residual: res = b - (con ? x : f) where f = K x~ and
// x~ = x with constrained entries zeroed (the rows/cols of K that
// FEM.assemble_matrix replaces by the identity, base.py:413-419).
__global__ void k_mask(int64_t n, const uint8_t *con, const double *x, double *xm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xm[i] = con[i] ? 0.0 : x[i];
}

__global__ void __launch_bounds__(BS) k_residual(int64_t n, const uint8_t *con, const double *b,
                                                 const double *x, const double *f, double *res,
                                                 double *part) {
  __shared__ double sh[2 * BS];
  double acc[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    const double v = b[i] - (con && con[i] ? x[i] : f[i]);
    res[i] = v;
    acc[0] += v * v;
    acc[1] += b[i] * b[i];
  }
  write_block_partials<2, 1>(acc, sh, part);
}

__global__ void __launch_bounds__(BS) k_sum2(const double *part, int nblk, double *out) {
  __shared__ double sh[BS];
  __shared__ double tot[2];
  sum_partials<2>(part, nblk, sh, tot);
  if (threadIdx.x < 2) out[threadIdx.x] = tot[threadIdx.x];
}

// ------------------------------------------------------- launch helpers
static unsigned grid_for(int64_t work, int per_block) {
  const int64_t g = (work + per_block - 1) / per_block;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, RB));
}

// Team width for the node-block SpMV from the mean node degree.
static int team_width(const tfm_mesh_t &m) {
  const int64_t n_adj = m.n_nod > 0 ? (int64_t)m.nnz / std::max(1, m.ndof * m.ndof) : 0;
  const double deg = (double)n_adj / std::max<int64_t>(1, m.n_nod);
  return deg > 16.0 ? 32 : deg > 8.0 ? 16 : 8;
}

// Launches the node-block SpMV; returns the grid size (number of partials).
template <typename T, int ND, int NRP>
static int spmv_launch(hipStream_t s, const tfm_mesh_t &m, const T *val, const T *x, T *y,
                       const int *alldone, double *part_pq, unsigned *grid_out) {
  const int tl = team_width(m);
  const unsigned g = grid_for(m.n_nod, BS / tl);
  switch (tl) {
    case 8: hipLaunchKernelGGL((k_spmv<T, ND, 8, NRP>), dim3(g), dim3(BS), 0, s, m, val, x, y, alldone, part_pq); break;
    case 16: hipLaunchKernelGGL((k_spmv<T, ND, 16, NRP>), dim3(g), dim3(BS), 0, s, m, val, x, y, alldone, part_pq); break;
    default: hipLaunchKernelGGL((k_spmv<T, ND, 32, NRP>), dim3(g), dim3(BS), 0, s, m, val, x, y, alldone, part_pq); break;
  }
  TFM_CHECK_LAUNCH();
  if (grid_out) *grid_out = g;
  return TFM_OK;
}

#define TFM_NRP_SWITCH(nrp, ...)                                                               \
  switch (nrp) {                                                                               \
    case 1: { constexpr int NRP = 1; __VA_ARGS__; break; }                                     \
    case 2: { constexpr int NRP = 2; __VA_ARGS__; break; }                                     \
    case 4: { constexpr int NRP = 4; __VA_ARGS__; break; }                                     \
    case 8: { constexpr int NRP = 8; __VA_ARGS__; break; }                                     \
    case 16: { constexpr int NRP = 16; __VA_ARGS__; break; }                                   \
    default: TFM_REQUIRE(false, "number of right-hand sides must be a power of two <= 16");    \
  }

#define TFM_ND_SWITCH(nd, ...)                                                                 \
  switch (nd) {                                                                                \
    case 1: { constexpr int ND = 1; __VA_ARGS__; break; }                                      \
    case 2: { constexpr int ND = 2; __VA_ARGS__; break; }                                      \
    case 3: { constexpr int ND = 3; __VA_ARGS__; break; }                                      \
    default: TFM_REQUIRE(false, "ndof must be 1, 2 or 3");                                     \
  }

int launch_spmv(hipStream_t s, const tfm_mesh_t &m, int prec, const void *vals, const void *x,
                void *y, int nrhs) {
  if (m.n_nod == 0) return TFM_OK;
  int rc = TFM_OK;
  TFM_ND_SWITCH(m.ndof, TFM_NRP_SWITCH(nrhs, {
    if (prec == 0)
      rc = spmv_launch<double, ND, NRP>(s, m, (const double *)vals, (const double *)x, (double *)y,
                                        nullptr, nullptr, nullptr);
    else
      rc = spmv_launch<float, ND, NRP>(s, m, (const float *)vals, (const float *)x, (float *)y,
                                       nullptr, nullptr, nullptr);
  }));
  return rc;
}

}  // namespace tfm

// ====================================================================
// PCG handle and driver
// ====================================================================
using namespace tfm;

struct tfm_pcg {
  int64_t n = 0;  // vector length (DOFs)
  int nrhs = 0;   // capacity
  int nrp = 1;    // capacity padded to a power of two
  int max_iter = 0;
  int device = 0;
  // inner vectors, interleaved [n][nrp]; sized for doubles, used as float in mixed mode
  double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double *minv = nullptr;                                  // sized n*3 doubles
  double *xm = nullptr, *r64 = nullptr, *f64 = nullptr;    // FP64 work vectors [n]
  double *work = nullptr;                                  // internal-force scratch
  int64_t work_len = 0;
  double *scal = nullptr, *tol2 = nullptr;
  double *part_pq = nullptr, *part_upd = nullptr, *part_aux = nullptr, *part_res = nullptr;
  double *res2 = nullptr, *rzold = nullptr;
  int *done = nullptr, *alldone = nullptr;
  int *h_done = nullptr;                                   // pinned host mirrors
  double *h_res = nullptr;
  // optional event timing of the inner SpMV launches
  bool timing = false;
  double spmv_ms = 0.0;
  int64_t spmv_count = 0;
  hipEvent_t ev[64] = {};
};

static int pow2_cap(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

static int pcg_alloc(tfm_pcg *h) {
  const int64_t n = h->n;
  const int nrp = h->nrp;
  const size_t vbytes = sizeof(double) * (size_t)n * nrp;
  for (double **v : {&h->x, &h->r, &h->z, &h->p, &h->q}) TFM_CHECK_HIP(hipMalloc(v, vbytes));
  TFM_CHECK_HIP(hipMalloc(&h->minv, sizeof(double) * (size_t)n * 3));
  for (double **v : {&h->xm, &h->r64, &h->f64}) TFM_CHECK_HIP(hipMalloc(v, sizeof(double) * (size_t)n));
  TFM_CHECK_HIP(hipMalloc(&h->scal, sizeof(double) * (size_t)(h->max_iter + 2) * nrp * NSCAL));
  TFM_CHECK_HIP(hipMalloc(&h->tol2, sizeof(double) * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->part_pq, sizeof(double) * RB * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->part_upd, sizeof(double) * RB * 2 * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->part_aux, sizeof(double) * RB * 3 * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->part_res, sizeof(double) * RB * 2));
  TFM_CHECK_HIP(hipMalloc(&h->res2, sizeof(double) * 2));
  TFM_CHECK_HIP(hipMalloc(&h->rzold, sizeof(double) * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->done, sizeof(int) * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->alldone, sizeof(int)));
  TFM_CHECK_HIP(hipHostMalloc(&h->h_done, sizeof(int) * (nrp + 1), hipHostMallocDefault));
  TFM_CHECK_HIP(hipHostMalloc(&h->h_res, sizeof(double) * 4, hipHostMallocDefault));
  for (auto &e : h->ev) TFM_CHECK_HIP(hipEventCreate(&e));
  return TFM_OK;
}

static void pcg_free(tfm_pcg *h) {
  for (double *ptr : {h->x, h->r, h->z, h->p, h->q, h->minv, h->xm, h->r64, h->f64, h->work,
                      h->scal, h->tol2, h->part_pq, h->part_upd, h->part_aux, h->part_res,
                      h->res2, h->rzold})
    if (ptr) (void)hipFree(ptr);
  if (h->done) (void)hipFree(h->done);
  if (h->alldone) (void)hipFree(h->alldone);
  if (h->h_done) (void)hipHostFree(h->h_done);
  if (h->h_res) (void)hipHostFree(h->h_res);
  for (auto &e : h->ev)
    if (e) (void)hipEventDestroy(e);
}

extern "C" int tfm_pcg_create(tfm_pcg_t **out, int64_t n, int32_t nrhs, int32_t max_iter) {
  TFM_REQUIRE(out != nullptr, "null handle pointer");
  TFM_REQUIRE(n > 0, "system size must be positive");
  TFM_REQUIRE(nrhs >= 1 && nrhs <= MAXR, "nrhs must be in [1, 16]");
  TFM_REQUIRE(max_iter >= 1, "max_iter must be positive");
  auto *h = new tfm_pcg();
  h->n = n;
  h->nrhs = nrhs;
  h->nrp = pow2_cap(nrhs);
  h->max_iter = max_iter;
  (void)hipGetDevice(&h->device);
  const int rc = pcg_alloc(h);
  if (rc != TFM_OK) {
    pcg_free(h);
    delete h;
    return rc;
  }
  *out = h;
  return TFM_OK;
}

extern "C" int tfm_pcg_destroy(tfm_pcg_t *h) {
  if (!h) return TFM_OK;
  pcg_free(h);
  delete h;
  return TFM_OK;
}

extern "C" int tfm_pcg_timing(tfm_pcg_t *h, int32_t enable, double *spmv_ms,
                              int64_t *spmv_launches) {
  TFM_REQUIRE(h != nullptr, "null solver handle");
  if (spmv_ms) *spmv_ms = h->spmv_ms;
  if (spmv_launches) *spmv_launches = h->spmv_count;
  h->spmv_ms = 0.0;
  h->spmv_count = 0;
  h->timing = enable != 0;
  return TFM_OK;
}

namespace {

// Operator callbacks of one solve: the inner SpMV (y = A x in the inner type,
// with optional dot partials; returns the grid size through *g) and the FP64
// true residual of rhs r (res = b - A x; returns ||res||^2 and ||b||^2).
template <typename T> struct Ops {
  std::function<int(const T *, T *, const int *, double *, unsigned *)> spmv;
  std::function<int(const double *, const double *, double *, double *, double *)> residual;
};

int sum_residual(tfm_pcg *h, hipStream_t s, unsigned nblk, double *rr, double *bb) {
  hipLaunchKernelGGL(k_sum2, dim3(1), dim3(BS), 0, s, (const double *)h->part_res, (int)nblk, h->res2);
  TFM_CHECK_LAUNCH();
  TFM_CHECK_HIP(hipMemcpyAsync(h->h_res, h->res2, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipStreamSynchronize(s));
  *rr = h->h_res[0];
  *bb = h->h_res[1];
  return TFM_OK;
}

// The PCG driver (see the file comment). fp64 selects the target of a segment:
// the final tolerance directly (FP64), or a reduction by `delta` before the
// next reliable update (mixed).
template <typename T, int ND, int NRP>
int run_pcg(tfm_pcg *h, hipStream_t s, int64_t n_nod, Ops<T> &ops, const double *b, double *x,
            int nrhs, double rtol, double delta, int32_t *iters, double *relres, int32_t *outer) {
  const bool fp64 = std::is_same<T, double>::value;
  const int64_t n = n_nod * ND;
  T *xv = (T *)h->x, *rv = (T *)h->r, *zv = (T *)h->z, *pv = (T *)h->p, *qv = (T *)h->q;
  const T *minv = (const T *)h->minv;
  PcgState st{h->scal, h->tol2, h->done, h->alldone, h->part_pq, h->part_upd, h->part_aux};
  const unsigned g1 = (unsigned)((n + 255) / 256);
  const unsigned gv = grid_for(n_nod * NRP, BS);
  std::vector<double> htol(NRP, -1.0), bnorm(NRP, 0.0), rprev(NRP, 0.0), rnow(NRP, 0.0);
  std::vector<int> fin(NRP, 1), stall(NRP, 0);
  for (int r = 0; r < nrhs; ++r) iters[r] = 0;

  // residuals of the initial guesses, staged as the inner residual columns
  auto stage = [&](int r, bool first) -> int {
    double rr2, bb2;
    int rc = ops.residual(b + (int64_t)r * n, x + (int64_t)r * n, h->r64, &rr2, &bb2);
    if (rc != TFM_OK) return rc;
    if (!std::isfinite(rr2)) {
      set_error("non-finite residual in the PCG solve");
      return TFM_ERR_NOCONV;
    }
    if (first) bnorm[r] = std::sqrt(bb2);
    const double rn = std::sqrt(rr2);
    rprev[r] = first ? rn : rnow[r];
    rnow[r] = rn;
    relres[r] = bnorm[r] > 0 ? rn / bnorm[r] : rn;
    fin[r] = rn <= rtol * bnorm[r];
    const double target = (fp64 ? 0.7 : 0.5) * rtol * bnorm[r];
    const double seg = fp64 ? target : std::max(delta * rn, target);
    htol[r] = fin[r] ? -1.0 : seg * seg;
    hipLaunchKernelGGL((k_set_column<T, NRP>), dim3(g1), dim3(256), 0, s, n, r,
                       (const double *)h->r64, rv, fin[r] ? 1 : 0);
    TFM_CHECK_LAUNCH();
    return TFM_OK;
  };
  hipLaunchKernelGGL((k_fill_zero<T>), dim3((unsigned)((n * NRP + 255) / 256)), dim3(256), 0, s,
                     n * NRP, xv);
  hipLaunchKernelGGL((k_fill_zero<T>), dim3((unsigned)((n * NRP + 255) / 256)), dim3(256), 0, s,
                     n * NRP, rv);
  TFM_CHECK_LAUNCH();
  for (int r = 0; r < nrhs; ++r) {
    const int rc = stage(r, true);
    if (rc != TFM_OK) return rc;
  }
  TFM_CHECK_HIP(hipMemcpyAsync(h->tol2, htol.data(), sizeof(double) * NRP, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL((k_pcg_start<T, ND, NRP>), dim3(gv), dim3(BS), 0, s, n_nod, (const T *)rv, zv,
                     pv, minv, st);
  TFM_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_pcg_start_finish<NRP>), dim3(1), dim3(BS), 0, s, 0, (int)gv, st);
  TFM_CHECK_LAUNCH();

  const int CHK = 16;
  int it = 0, n_outer = 0;
  bool any_live = false;
  for (int r = 0; r < nrhs; ++r) any_live |= !fin[r];
  int seg_start = 0;
  while (any_live) {
    // run the current segment in batches of CHK iterations
    bool seg_done = false;
    while (!seg_done && it < h->max_iter) {
      const int start = it;
      const int stop = std::min(h->max_iter, it + CHK);
      for (; it < stop; ++it) {
        if (h->timing) TFM_CHECK_HIP(hipEventRecord(h->ev[2 * (it - start)], s));
        unsigned npq = 0;
        int rc = ops.spmv(pv, qv, h->alldone, h->part_pq, &npq);
        if (rc != TFM_OK) return rc;
        if (h->timing) TFM_CHECK_HIP(hipEventRecord(h->ev[2 * (it - start) + 1], s));
        hipLaunchKernelGGL((k_pcg_update<T, ND, NRP>), dim3(gv), dim3(BS), 0, s, n_nod, it, xv, rv,
                           zv, (const T *)pv, (const T *)qv, minv, st, (int)npq);
        TFM_CHECK_LAUNCH();
        hipLaunchKernelGGL((k_pcg_dir<T, ND, NRP>), dim3(gv), dim3(BS), 0, s, n_nod, it,
                           (const T *)zv, pv, st, (int)gv);
        TFM_CHECK_LAUNCH();
      }
      TFM_CHECK_HIP(hipMemcpyAsync(h->h_done, h->done, sizeof(int) * NRP, hipMemcpyDeviceToHost, s));
      TFM_CHECK_HIP(hipStreamSynchronize(s));
      seg_done = true;
      int end = 0;
      for (int r = 0; r < NRP; ++r) {
        seg_done &= h->h_done[r] >= 0;
        end = std::max(end, h->h_done[r]);
      }
      if (h->timing) {  // launches after the last rhs stopped were empty
        const int real_end = seg_done ? std::min(it, end) : it;
        for (int i = start; i < real_end; ++i) {
          float ms = 0.0f;
          TFM_CHECK_HIP(hipEventElapsedTime(&ms, h->ev[2 * (i - start)], h->ev[2 * (i - start) + 1]));
          h->spmv_ms += ms;
          h->spmv_count += 1;
        }
      }
    }
    for (int r = 0; r < nrhs; ++r)
      if (!fin[r]) iters[r] += std::max(0, (h->h_done[r] >= 0 ? h->h_done[r] : it) - seg_start);
    // fold the corrections into x and replace the residuals by true ones;
    // FP64 (whose corrections only absorb the drift of the recursive residual)
    // and a mixed solve whose true residual stopped falling restart from the
    // new residual, other mixed solves continue their recurrence
    int mask = 0, restart = 0;
    any_live = false;
    for (int r = 0; r < nrhs; ++r) {
      if (fin[r]) continue;
      hipLaunchKernelGGL((k_take_column<T, NRP>), dim3(g1), dim3(256), 0, s, n, r, xv,
                         x + (int64_t)r * n);
      TFM_CHECK_LAUNCH();
      const int rc = stage(r, false);
      if (rc != TFM_OK) return rc;
      if (!fin[r]) {
        mask |= 1 << r;
        const bool stuck = rnow[r] > 0.9 * rprev[r];
        stall[r] = stuck ? stall[r] + 1 : 0;
        if (fp64 || stuck) restart |= 1 << r;
        any_live = true;
      }
    }
    // a residual that no correction improves has reached the attainable
    // accuracy of the arithmetic: give up instead of cycling
    for (int r = 0; r < nrhs; ++r)
      if (!fin[r] && stall[r] >= 8) {
        set_error("PCG stagnated above the tolerance (attainable accuracy reached)");
        if (outer) *outer = n_outer;
        return TFM_ERR_NOCONV;
      }
    ++n_outer;
    if (!any_live || it >= h->max_iter) break;
    TFM_CHECK_HIP(hipMemcpyAsync(h->tol2, htol.data(), sizeof(double) * NRP, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_reliable_dots<T, ND, NRP>), dim3(gv), dim3(BS), 0, s, n_nod,
                       (const T *)rv, zv, (const T *)pv, minv, st);
    TFM_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_reliable_prep<NRP>), dim3(1), dim3(64), 0, s, mask, restart, seg_start,
                       st, h->rzold);
    TFM_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_reliable_apply<T, ND, NRP>), dim3(gv), dim3(BS), 0, s, n_nod, it, mask,
                       (const T *)zv, pv, st, (const double *)h->rzold, (int)gv);
    TFM_CHECK_LAUNCH();
    seg_start = it;
  }
  if (outer) *outer = n_outer;
  for (int r = 0; r < nrhs; ++r)
    if (!fin[r]) {
      set_error("PCG did not reach the tolerance within max_iter iterations");
      return TFM_ERR_NOCONV;
    }
  return TFM_OK;
}

template <typename T>
int build_precond(hipStream_t s, int64_t n_nod, int nd, const double *dblk, int pc, T *minv) {
  const unsigned g = (unsigned)((n_nod + 255) / 256);
  TFM_ND_SWITCH(nd, hipLaunchKernelGGL((k_precond<T, ND>), dim3(g), dim3(256), 0, s, n_nod, dblk,
                                       pc, minv));
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

// FP64 true residual of the finite-element system: with the FP64 values when
// present, otherwise matrix-free through the element kernels.
int fem_residual(tfm_pcg *h, hipStream_t s, const tfm_system_t &sys, const double *b,
                 const double *x, double *res, double *rr, double *bb) {
  const tfm_mesh_t &m = *sys.mesh;
  const int64_t n = m.n_nod * m.ndof;
  const unsigned g1 = (unsigned)((n + 255) / 256);
  const uint8_t *con = nullptr;
  if (sys.val64) {
    int rc = launch_spmv(s, m, 0, sys.val64, x, h->f64, 1);
    if (rc != TFM_OK) return rc;
  } else {
    TFM_REQUIRE(sys.mat != nullptr, "matrix-free residual needs the material");
    const int64_t wl = m.n_elem * m.npe * m.ndof;
    if (h->work_len < wl) {
      if (h->work) (void)hipFree(h->work);
      h->work = nullptr;
      TFM_CHECK_HIP(hipMalloc(&h->work, sizeof(double) * std::max<int64_t>(wl, 1)));
      h->work_len = wl;
    }
    hipLaunchKernelGGL(k_mask, dim3(g1), dim3(256), 0, s, n, sys.con, x, h->xm);
    TFM_CHECK_LAUNCH();
    int rc = launch_internal_force(s, m, *sys.mat, h->xm, nullptr, h->f64, nullptr, nullptr,
                                   nullptr, nullptr, h->work);
    if (rc != TFM_OK) return rc;
    con = sys.con;
  }
  const unsigned gr = grid_for(n, BS);
  hipLaunchKernelGGL(k_residual, dim3(gr), dim3(BS), 0, s, n, con, b, x, (const double *)h->f64,
                     res, h->part_res);
  TFM_CHECK_LAUNCH();
  return sum_residual(h, s, gr, rr, bb);
}

}  // namespace

extern "C" int tfm_pcg_solve(tfm_pcg_t *h, tfm_stream_t stream, const tfm_system_t *sys,
                             int32_t prec, int32_t pc, const double *b, double *x, int32_t nrhs,
                             double rtol, double inner_rtol, int32_t *iters, double *relres,
                             int32_t *outer) {
  TFM_REQUIRE(h && sys && sys->mesh && b && x && iters && relres, "null argument");
  const tfm_mesh_t &m = *sys->mesh;
  const int64_t n = m.n_nod * m.ndof;
  TFM_REQUIRE(n == h->n, "system size differs from the solver handle");
  TFM_REQUIRE(nrhs >= 1 && nrhs <= h->nrhs, "nrhs exceeds the solver handle");
  TFM_REQUIRE(pc == TFM_PC_JACOBI || pc == TFM_PC_BLOCK_JACOBI, "unknown preconditioner");
  TFM_REQUIRE(prec == TFM_PREC_FP64 || prec == TFM_PREC_MIXED, "unknown precision mode");
  TFM_REQUIRE(sys->con && sys->diag_blocks, "system needs con and diag_blocks");
  TFM_REQUIRE(prec == TFM_PREC_FP64 ? sys->val64 != nullptr : sys->val32 != nullptr,
              "matrix values for the requested precision are missing");
  TFM_REQUIRE(rtol > 0.0, "rtol must be positive");
  TFM_REQUIRE(inner_rtol > 0.0 && inner_rtol < 1.0, "inner_rtol must be in (0, 1)");
  if (outer) *outer = 0;
  hipStream_t s = (hipStream_t)stream;
  const int nrp = pow2_cap(nrhs);
  int rc = TFM_OK;
  auto residual = [&](const double *bb, const double *xx, double *res, double *rr, double *b2) -> int {
    return fem_residual(h, s, *sys, bb, xx, res, rr, b2);
  };
  TFM_ND_SWITCH(m.ndof, TFM_NRP_SWITCH(nrp, {
    if (prec == TFM_PREC_FP64) {
      rc = build_precond<double>(s, m.n_nod, ND, sys->diag_blocks, pc, (double *)h->minv);
      if (rc != TFM_OK) return rc;
      Ops<double> ops;
      ops.spmv = [&](const double *xi, double *yo, const int *ad, double *part, unsigned *g) -> int {
        return spmv_launch<double, ND, NRP>(s, m, sys->val64, xi, yo, ad, part, g);
      };
      ops.residual = residual;
      rc = run_pcg<double, ND, NRP>(h, s, m.n_nod, ops, b, x, nrhs, rtol, inner_rtol, iters,
                                    relres, outer);
    } else {
      rc = build_precond<float>(s, m.n_nod, ND, sys->diag_blocks, pc, (float *)h->minv);
      if (rc != TFM_OK) return rc;
      Ops<float> ops;
      ops.spmv = [&](const float *xi, float *yo, const int *ad, double *part, unsigned *g) -> int {
        return spmv_launch<float, ND, NRP>(s, m, sys->val32, xi, yo, ad, part, g);
      };
      ops.residual = residual;
      rc = run_pcg<float, ND, NRP>(h, s, m.n_nod, ops, b, x, nrhs, rtol, inner_rtol, iters,
                                   relres, outer);
    }
  }));
  return rc;
}

// ===================================================================
// General CSR matrices (no mesh): the operator and Jacobi-PCG behind
// torchfem.sparse for user-supplied sparse matrices.
// ===================================================================
namespace {

__global__ void k_csr_diag(int64_t n, const int64_t *rowptr, const int32_t *col,
                           const double *val, double *d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    if (col[k] == i) v += val[k];
  d[i] = v;
}

int csr_spmv(hipStream_t s, int64_t n, const int64_t *rowptr, const int32_t *col,
             const double *val, const double *x, double *y, const int *alldone, double *part,
             unsigned *grid_out) {
  const unsigned g = grid_for(n, BS / 8);
  hipLaunchKernelGGL((k_spmv_csr<8>), dim3(g), dim3(BS), 0, s, n, rowptr, col, val, x, y,
                     alldone, part);
  TFM_CHECK_LAUNCH();
  if (grid_out) *grid_out = g;
  return TFM_OK;
}

}  // namespace

extern "C" int tfm_csr_spmv(tfm_stream_t s, int64_t n, const int64_t *rowptr, const int32_t *col,
                            const double *val, const double *x, double *y) {
  TFM_REQUIRE(n >= 0, "negative size");
  if (n == 0) return TFM_OK;
  TFM_REQUIRE(rowptr && col && val && x && y, "null argument");
  return csr_spmv((hipStream_t)s, n, rowptr, col, val, x, y, nullptr, nullptr, nullptr);
}

extern "C" int tfm_csr_pcg_solve(tfm_pcg_t *h, tfm_stream_t stream, int64_t n,
                                 const int64_t *rowptr, const int32_t *col, const double *val,
                                 const double *b, double *x, double rtol, int32_t *iters,
                                 double *relres) {
  TFM_REQUIRE(h && rowptr && col && val && b && x && iters && relres, "null argument");
  TFM_REQUIRE(n == h->n, "system size differs from the solver handle");
  TFM_REQUIRE(rtol > 0.0, "rtol must be positive");
  hipStream_t s = (hipStream_t)stream;
  const unsigned g1 = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_csr_diag, dim3(g1), dim3(256), 0, s, n, rowptr, col, val, h->f64);
  TFM_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_precond<double, 1>), dim3(g1), dim3(256), 0, s, n, (const double *)h->f64,
                     (int)TFM_PC_JACOBI, h->minv);
  TFM_CHECK_LAUNCH();
  Ops<double> ops;
  ops.spmv = [&](const double *xi, double *yo, const int *ad, double *part, unsigned *g) -> int {
    return csr_spmv(s, n, rowptr, col, val, xi, yo, ad, part, g);
  };
  ops.residual = [&](const double *bb, const double *xx, double *res, double *rr, double *b2) -> int {
    int rc = csr_spmv(s, n, rowptr, col, val, xx, h->f64, nullptr, nullptr, nullptr);
    if (rc != TFM_OK) return rc;
    const unsigned gr = grid_for(n, BS);
    hipLaunchKernelGGL(k_residual, dim3(gr), dim3(BS), 0, s, n, (const uint8_t *)nullptr, bb, xx,
                       (const double *)h->f64, res, h->part_res);
    TFM_CHECK_LAUNCH();
    return sum_residual(h, s, gr, rr, b2);
  };
  int32_t outer = 0;
  return run_pcg<double, 1, 1>(h, s, n, ops, b, x, 1, rtol, 0.5, iters, relres, &outer);
}
