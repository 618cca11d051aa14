// Sparse matrix-vector products and the preconditioned conjugate-gradient
// solver (FP64, and FP32-inner with FP64 iterative refinement) for gfx950.
//
// Vectors inside the solver are interleaved [n][nrp] (nrp = right-hand sides
// padded to a power of two), so one pass over the matrix serves every
// right-hand side and the gathers of x read nrp consecutive values.
// Reductions are deterministic: every block writes its partial sums and the
// last block to finish adds them in a fixed order (an integer ticket counter
// decides which block is last; no float atomics anywhere).
#include "tfm_common.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace tfm {

constexpr int BS = 256;          // threads per block of the solver kernels
constexpr int RED_BLOCKS = 1024; // fixed grid of the reducing kernels
constexpr int MAXR = 16;         // right-hand sides per solve
constexpr int NSCAL = 4;         // per-iteration scalars: rz, pq, rr, spare
enum { S_RZ = 0, S_PQ = 1, S_RR = 2 };

// ---------------------------------------------------------- reductions
// Called by every thread of the block after it wrote its contribution for
// value v < nv to red[v * BS + tid]; rows of red with stride `period` (a power
// of two dividing BS) are summed per slot. The last block of the grid adds the
// block partials in block order and hands the totals to `finish`.
template <class Fin>
__device__ void grid_reduce(double *red, int nv, int period, double *partials,
                            unsigned *counter, Fin finish) {
  const int tid = threadIdx.x;
  __syncthreads();
  for (int s = BS / 2; s >= period; s >>= 1) {
    if (tid < s)
      for (int v = 0; v < nv; ++v) red[v * BS + tid] += red[v * BS + tid + s];
    __syncthreads();
  }
  const int nt = nv * period;
  if (tid < nt) {
    const int v = tid / period, r = tid - v * period;
    partials[(size_t)blockIdx.x * nt + tid] = red[v * BS + r];
  }
  __threadfence();
  __syncthreads();
  __shared__ unsigned ticket;
  if (tid == 0) ticket = atomicAdd(counter, 1u);
  __syncthreads();
  if (ticket != gridDim.x - 1) return;
  __threadfence();
  // last block: thread t sums value slot (t % nt) over blocks b = t / nt (mod chunks)
  const int chunks = BS / nt;
  double acc = 0.0;
  if (tid < chunks * nt) {
    const int slot = tid % nt, c = tid / nt;
    for (int b = c; b < (int)gridDim.x; b += chunks)
      acc += __hip_atomic_load(&partials[(size_t)b * nt + slot], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
  }
  red[tid] = acc;
  __syncthreads();
  if (tid < nt) {
    double tot = 0.0;
    for (int c = 0; c < chunks; ++c) tot += red[c * nt + tid];
    finish(tid / period, tid % period, tot);
  }
  if (tid == 0) *counter = 0u;
}

// ------------------------------------------------------- node-block SpMV
// One team of TL lanes per node (grid-stride); the team walks the node's
// ND rows (ND*ND*deg contiguous values) and reduces its ND x nrp sums.
// Optionally forms dot = sum_i y_i x_i per right-hand side.
template <typename T, int ND, int TL, int NR>
__global__ void __launch_bounds__(BS) k_spmv_nodes(tfm_mesh_t m, const T *__restrict__ val,
                                                   const T *__restrict__ x, T *__restrict__ y,
                                                   int nrp, const int *alldone,
                                                   double *scal_pq, double *partials,
                                                   unsigned *counter) {
  if (alldone && *alldone) return;
  constexpr int TPB = BS / TL;
  __shared__ double red[MAXR * 64 > BS ? MAXR * 64 : BS];
  const int tid = threadIdx.x, lane = tid % TL, team = tid / TL;
  double dot[NR];
  for (int r = 0; r < NR; ++r) dot[r] = 0.0;
  const int64_t nteams = (int64_t)gridDim.x * TPB;
  for (int64_t n = (int64_t)blockIdx.x * TPB + team; n < m.n_nod; n += nteams) {
    const int a0 = m.adj_ptr[n];
    const int deg = m.adj_ptr[n + 1] - a0;
    const int64_t vb = m.val_ptr[n];
    T acc[ND][NR];
    for (int a = 0; a < ND; ++a)
      for (int r = 0; r < NR; ++r) acc[a][r] = T(0);
    if (deg > 0) {
      const int L = ND * deg;
      for (int q = lane; q < ND * L; q += TL) {
        int a = 0;
        if constexpr (ND == 2) a = q >= L;
        if constexpr (ND == 3) a = (q >= L) + (q >= 2 * L);
        const int rem = q - a * L;
        const int p = rem / ND, b = rem - p * ND;
        const T v = val[vb + q];
        const int64_t xc = ((int64_t)m.adj[a0 + p] * ND + b) * nrp;
#pragma unroll
        for (int aa = 0; aa < ND; ++aa)
          if (aa == a)
#pragma unroll
            for (int r = 0; r < NR; ++r)
              if (r < nrp) acc[aa][r] += v * x[xc + r];
      }
    } else if (lane < ND) {  // orphan node: diagonal only
#pragma unroll
      for (int aa = 0; aa < ND; ++aa)
        if (aa == lane) {
          const T v = val[vb + lane];
#pragma unroll
          for (int r = 0; r < NR; ++r)
            if (r < nrp) acc[aa][r] = v * x[(n * ND + lane) * nrp + r];
        }
    }
#pragma unroll
    for (int off = 1; off < TL; off <<= 1)
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[a][r] += __shfl_xor(acc[a][r], off, TL);
    if (lane == 0) {
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int r = 0; r < NR; ++r)
          if (r < nrp) {
            const int64_t o = (n * ND + a) * nrp + r;
            y[o] = acc[a][r];
            if (scal_pq) dot[r] += (double)acc[a][r] * (double)x[o];
          }
    }
  }
  if (!scal_pq) return;
  // per-rhs block sums: team leaders store their dots, threads r < nrp add them
  __syncthreads();
  if (lane == 0)
    for (int r = 0; r < NR; ++r) red[r * TPB + team] = dot[r];
  __syncthreads();
  double bs = 0.0;
  if (tid < nrp)
    for (int t = 0; t < TPB; ++t) bs += red[tid * TPB + t];
  __syncthreads();
  for (int t = tid; t < BS; t += BS) red[t] = 0.0;
  __syncthreads();
  if (tid < nrp) red[tid] = bs;
  // reduce "rows" of period nrp: only entries < nrp are nonzero
  grid_reduce(red, 1, nrp, partials, counter,
              [&](int, int r, double tot) { scal_pq[r * NSCAL] = tot; });
}

// ------------------------------------------------------- general CSR SpMV
template <int TL>
__global__ void __launch_bounds__(BS) k_spmv_csr(int64_t n, const int64_t *rowptr,
                                                 const int32_t *col, const double *val,
                                                 const double *x, double *y, const int *alldone,
                                                 double *scal_pq, double *partials,
                                                 unsigned *counter) {
  if (alldone && *alldone) return;
  constexpr int TPB = BS / TL;
  __shared__ double red[2 * BS];
  const int tid = threadIdx.x, lane = tid % TL, team = tid / TL;
  double dot = 0.0;
  const int64_t nteams = (int64_t)gridDim.x * TPB;
  for (int64_t i = (int64_t)blockIdx.x * TPB + team; i < n; i += nteams) {
    double acc = 0.0;
    for (int64_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += TL) acc += val[k] * x[col[k]];
#pragma unroll
    for (int off = 1; off < TL; off <<= 1) acc += __shfl_xor(acc, off, TL);
    if (lane == 0) {
      y[i] = acc;
      dot += acc * x[i];
    }
  }
  if (!scal_pq) return;
  red[tid] = dot;
  grid_reduce(red, 1, 1, partials, counter, [&](int, int, double tot) { scal_pq[0] = tot; });
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

// z_n = Minv_n r_n for node n, right-hand side r.
template <typename T, int ND>
__device__ __forceinline__ void apply_pc(const T *minv, const T *r, int64_t n, int rr, int nrp,
                                         T (&z)[ND], T (&rv)[ND]) {
  for (int a = 0; a < ND; ++a) rv[a] = r[(n * ND + a) * nrp + rr];
  for (int a = 0; a < ND; ++a) {
    T acc = T(0);
    for (int b = 0; b < ND; ++b) acc += minv[n * ND * ND + a * ND + b] * rv[b];
    z[a] = acc;
  }
}

// ------------------------------------------------------------ PCG steps
struct PcgState {
  double *scal;   // [(max_iter + 2) * nrp * NSCAL]
  double *bb;     // [nrp] ||b||^2 of the current (inner) right-hand sides
  double *tol2;   // [nrp] squared absolute tolerance per rhs
  int *done;      // [nrp] iteration at which rhs r converged, -1 while running
  int *alldone;   // 1 once every rhs converged
  double *partials;
  unsigned *counter;
};

// r = b - q (q = A x, or NULL for x = 0); z = M r; p = z; rz, rr, bb.
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_pcg_init(int64_t n_nod, int nrp, int nrhs, const T *b,
                                                 const T *q, T *r, T *p, const T *minv,
                                                 PcgState st) {
  __shared__ double red[3 * BS];
  const int tid = threadIdx.x;
  double s_rz = 0.0, s_rr = 0.0, s_bb = 0.0;
  const int64_t total = n_nod * nrp;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int64_t n = t / nrp;
    const int rr = (int)(t - n * nrp);
    T rv[ND], z[ND];
    for (int a = 0; a < ND; ++a) {
      const int64_t o = (n * ND + a) * nrp + rr;
      const T bv = b[o];
      r[o] = q ? bv - q[o] : bv;
      s_bb += (double)bv * (double)bv;
    }
    apply_pc<T, ND>(minv, r, n, rr, nrp, z, rv);
    for (int a = 0; a < ND; ++a) {
      p[(n * ND + a) * nrp + rr] = z[a];
      s_rz += (double)rv[a] * (double)z[a];
      s_rr += (double)rv[a] * (double)rv[a];
    }
  }
  red[tid] = s_rz;
  red[BS + tid] = s_rr;
  red[2 * BS + tid] = s_bb;
  grid_reduce(red, 3, nrp, st.partials, st.counter, [&](int v, int rr, double tot) {
    if (v == 0) st.scal[rr * NSCAL + S_RZ] = tot;
    if (v == 2) st.bb[rr] = tot;
    if (v == 1) {
      st.scal[rr * NSCAL + S_RR] = tot;
      const bool pad = rr >= nrhs;
      st.done[rr] = (pad || tot <= st.tol2[rr]) ? 0 : -1;
    }
  });
}

// alpha = rz/pq; x += alpha p; r -= alpha q; z = M r; next rz and rr; test.
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_pcg_update(int64_t n_nod, int nrp, int it, T *x, T *r,
                                                   T *z, const T *p, const T *q, const T *minv,
                                                   PcgState st) {
  if (*st.alldone) return;
  __shared__ double red[2 * BS];
  __shared__ double alpha[MAXR];
  const int tid = threadIdx.x;
  if (tid < nrp) {
    const double *sc = st.scal + ((size_t)it * nrp + tid) * NSCAL;
    alpha[tid] = (st.done[tid] < 0 && sc[S_PQ] != 0.0) ? sc[S_RZ] / sc[S_PQ] : 0.0;
  }
  __syncthreads();
  double s_rz = 0.0, s_rr = 0.0;
  const int64_t total = n_nod * nrp;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int64_t n = t / nrp;
    const int rr = (int)(t - n * nrp);
    const T al = (T)alpha[rr];
    for (int a = 0; a < ND; ++a) {
      const int64_t o = (n * ND + a) * nrp + rr;
      x[o] += al * p[o];
      r[o] -= al * q[o];
    }
    T rv[ND], zv[ND];
    apply_pc<T, ND>(minv, r, n, rr, nrp, zv, rv);
    for (int a = 0; a < ND; ++a) {
      z[(n * ND + a) * nrp + rr] = zv[a];
      s_rz += (double)rv[a] * (double)zv[a];
      s_rr += (double)rv[a] * (double)rv[a];
    }
  }
  red[tid] = s_rz;
  red[BS + tid] = s_rr;
  grid_reduce(red, 2, nrp, st.partials, st.counter, [&](int v, int rr, double tot) {
    double *sc = st.scal + ((size_t)(it + 1) * nrp + rr) * NSCAL;
    if (v == 0) sc[S_RZ] = tot;
    if (v == 1) {
      sc[S_RR] = tot;
      if (st.done[rr] < 0 && (tot <= st.tol2[rr] || !(tot == tot))) st.done[rr] = it + 1;
    }
  });
}

// beta = rz_{it+1}/rz_it; p = z + beta p
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_pcg_dir(int64_t n_nod, int nrp, int it, const T *z, T *p,
                                                PcgState st) {
  if (*st.alldone) return;
  __shared__ double beta[MAXR];
  __shared__ int live[MAXR];
  const int tid = threadIdx.x;
  if (tid < nrp) {
    const double *s0 = st.scal + ((size_t)it * nrp + tid) * NSCAL;
    const double *s1 = st.scal + ((size_t)(it + 1) * nrp + tid) * NSCAL;
    beta[tid] = s0[S_RZ] != 0.0 ? s1[S_RZ] / s0[S_RZ] : 0.0;
    live[tid] = st.done[tid] < 0;
  }
  __syncthreads();
  const int64_t total = n_nod * ND * nrp;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int rr = (int)(t % nrp);
    if (live[rr]) p[t] = z[t] + (T)beta[rr] * p[t];
  }
}

// Sets *alldone when every rhs has converged (single thread).
__global__ void k_check_done(int nrp, const int *done, int *alldone) {
  int all = 1;
  for (int r = 0; r < nrp; ++r) all &= done[r] >= 0;
  *alldone = all;
}

// ------------------------------------------------------- layout helpers
// dst[i*nrp + r] = (To) src[r*n + i] for r < nrhs, 0 for padding; optionally masked
template <typename To, typename Ti>
__global__ void k_interleave(int64_t n, int nrhs, int nrp, const Ti *src, To *dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nrp) return;
  const int64_t i = t / nrp;
  const int r = (int)(t - i * nrp);
  dst[t] = r < nrhs ? (To)src[(int64_t)r * n + i] : To(0);
}

// dst[r*n + i] (+)= (double) src[i*nrp + r]
template <typename Ti>
__global__ void k_deinterleave(int64_t n, int nrhs, int nrp, const Ti *src, double *dst,
                               int accumulate) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nrhs) return;
  const int r = (int)(t / n);
  const int64_t i = t - (int64_t)r * n;
  const double v = (double)src[i * nrp + r];
  dst[t] = accumulate ? dst[t] + v : v;
}

// Column r of an interleaved FP32 right-hand side from an FP64 vector:
// dst[i*nrp + r] = zero ? 0 : (float) src[i].
__global__ void k_set_column(int64_t n, int nrp, int r, const double *src, float *dst, int zero) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i * nrp + r] = zero ? 0.0f : (float)src[i];
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
                                                 double *out /*[2]: rr, bb*/, double *partials,
                                                 unsigned *counter) {
  __shared__ double red[2 * BS];
  const int tid = threadIdx.x;
  double rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BS + tid; i < n; i += (int64_t)gridDim.x * BS) {
    const double v = b[i] - (con && con[i] ? x[i] : f[i]);
    res[i] = v;
    rr += v * v;
    bb += b[i] * b[i];
  }
  red[tid] = rr;
  red[BS + tid] = bb;
  grid_reduce(red, 2, 1, partials, counter, [&](int v, int, double tot) { out[v] = tot; });
}

// ------------------------------------------------------- launch helpers
static unsigned grid_for(int64_t work, int per_block) {
  int64_t g = (work + per_block - 1) / per_block;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, RED_BLOCKS));
}

// Team width for the node-block SpMV from the values per node.
static int team_width(const tfm_mesh_t &m) {
  const double per_node = (double)m.nnz / std::max<int64_t>(1, m.n_nod);
  return per_node > 400 ? 64 : per_node > 100 ? 16 : per_node > 24 ? 8 : 4;
}

template <typename T, int ND, int NR>
static int spmv_nr(hipStream_t s, const tfm_mesh_t &m, const T *val, const T *x, T *y, int nrp,
                   const int *alldone, double *scal_pq, double *partials, unsigned *counter) {
  const int tl = team_width(m);
  const unsigned g = grid_for(m.n_nod, BS / tl);
#define TFM_SPMV_TL(TLV)                                                                        \
  hipLaunchKernelGGL((k_spmv_nodes<T, ND, TLV, NR>), dim3(g), dim3(BS), 0, s, m, val, x, y, nrp, \
                     alldone, scal_pq, partials, counter)
  switch (tl) {
    case 4: TFM_SPMV_TL(4); break;
    case 8: TFM_SPMV_TL(8); break;
    case 16: TFM_SPMV_TL(16); break;
    default: TFM_SPMV_TL(64); break;
  }
#undef TFM_SPMV_TL
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

template <typename T, int ND>
static int spmv_nd(hipStream_t s, const tfm_mesh_t &m, const T *val, const T *x, T *y, int nrp,
                   const int *alldone, double *scal_pq, double *partials, unsigned *counter) {
  if (nrp == 1) return spmv_nr<T, ND, 1>(s, m, val, x, y, nrp, alldone, scal_pq, partials, counter);
  if (nrp <= 4) return spmv_nr<T, ND, 4>(s, m, val, x, y, nrp, alldone, scal_pq, partials, counter);
  return spmv_nr<T, ND, 16>(s, m, val, x, y, nrp, alldone, scal_pq, partials, counter);
}

template <typename T>
static int spmv_any(hipStream_t s, const tfm_mesh_t &m, const T *val, const T *x, T *y, int nrp,
                    const int *alldone, double *scal_pq, double *partials, unsigned *counter) {
  TFM_REQUIRE(nrp >= 1 && nrp <= MAXR, "number of right-hand sides out of range");
  switch (m.ndof) {
    case 1: return spmv_nd<T, 1>(s, m, val, x, y, nrp, alldone, scal_pq, partials, counter);
    case 2: return spmv_nd<T, 2>(s, m, val, x, y, nrp, alldone, scal_pq, partials, counter);
    case 3: return spmv_nd<T, 3>(s, m, val, x, y, nrp, alldone, scal_pq, partials, counter);
    default: TFM_REQUIRE(false, "ndof must be 1, 2 or 3");
  }
}

int launch_spmv(hipStream_t s, const tfm_mesh_t &m, int prec, const void *vals, const void *x,
                void *y, int nrhs) {
  if (m.n_nod == 0) return TFM_OK;
  if (prec == 0)
    return spmv_any<double>(s, m, (const double *)vals, (const double *)x, (double *)y, nrhs,
                            nullptr, nullptr, nullptr, nullptr);
  return spmv_any<float>(s, m, (const float *)vals, (const float *)x, (float *)y, nrhs, nullptr,
                         nullptr, nullptr, nullptr);
}

}  // namespace tfm

// ====================================================================
// PCG handle and driver
// ====================================================================
using namespace tfm;

struct tfm_pcg {
  int64_t n = 0;       // vector length (DOFs)
  int nrhs = 0;        // capacity
  int max_iter = 0;
  int device = 0;
  // inner vectors (sized for doubles, used as float in mixed mode), interleaved
  double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *b = nullptr;
  double *minv = nullptr;            // preconditioner, sized n*3 doubles
  double *xm = nullptr, *r64 = nullptr, *f64 = nullptr;   // FP64 outer vectors [n]
  double *work = nullptr;            // internal-force scratch
  int64_t work_len = 0;
  double *scal = nullptr, *bb = nullptr, *tol2 = nullptr, *partials = nullptr, *res2 = nullptr;
  int *done = nullptr, *alldone = nullptr;
  unsigned *counter = nullptr;
  int *h_done = nullptr;             // pinned host mirrors
  double *h_res = nullptr;
};

static int pow2_cap(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

static int pcg_alloc(tfm_pcg *h) {
  const int64_t n = h->n;
  const int nrp = pow2_cap(h->nrhs);
  const size_t vbytes = sizeof(double) * (size_t)n * nrp;
  TFM_CHECK_HIP(hipMalloc(&h->x, vbytes));
  TFM_CHECK_HIP(hipMalloc(&h->r, vbytes));
  TFM_CHECK_HIP(hipMalloc(&h->z, vbytes));
  TFM_CHECK_HIP(hipMalloc(&h->p, vbytes));
  TFM_CHECK_HIP(hipMalloc(&h->q, vbytes));
  TFM_CHECK_HIP(hipMalloc(&h->b, vbytes));
  TFM_CHECK_HIP(hipMalloc(&h->minv, sizeof(double) * (size_t)n * 3));
  TFM_CHECK_HIP(hipMalloc(&h->xm, sizeof(double) * (size_t)n));
  TFM_CHECK_HIP(hipMalloc(&h->r64, sizeof(double) * (size_t)n));
  TFM_CHECK_HIP(hipMalloc(&h->f64, sizeof(double) * (size_t)n));
  TFM_CHECK_HIP(hipMalloc(&h->scal, sizeof(double) * (size_t)(h->max_iter + 2) * nrp * NSCAL));
  TFM_CHECK_HIP(hipMalloc(&h->bb, sizeof(double) * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->tol2, sizeof(double) * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->res2, sizeof(double) * 2));
  TFM_CHECK_HIP(hipMalloc(&h->partials, sizeof(double) * RED_BLOCKS * 4 * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->done, sizeof(int) * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->alldone, sizeof(int)));
  TFM_CHECK_HIP(hipMalloc(&h->counter, sizeof(unsigned)));
  TFM_CHECK_HIP(hipMemset(h->counter, 0, sizeof(unsigned)));
  TFM_CHECK_HIP(hipHostMalloc(&h->h_done, sizeof(int) * (nrp + 1), hipHostMallocDefault));
  TFM_CHECK_HIP(hipHostMalloc(&h->h_res, sizeof(double) * 4, hipHostMallocDefault));
  return TFM_OK;
}

static void pcg_free(tfm_pcg *h) {
  for (double *ptr : {h->x, h->r, h->z, h->p, h->q, h->b, h->minv, h->xm, h->r64, h->f64,
                      h->work, h->scal, h->bb, h->tol2, h->res2, h->partials})
    if (ptr) (void)hipFree(ptr);
  if (h->done) (void)hipFree(h->done);
  if (h->alldone) (void)hipFree(h->alldone);
  if (h->counter) (void)hipFree(h->counter);
  if (h->h_done) (void)hipHostFree(h->h_done);
  if (h->h_res) (void)hipHostFree(h->h_res);
}

extern "C" int tfm_pcg_create(tfm_pcg_t **out, int64_t n, int32_t nrhs, int32_t max_iter) {
  TFM_REQUIRE(out != nullptr, "null handle pointer");
  TFM_REQUIRE(n > 0, "system size must be positive");
  TFM_REQUIRE(nrhs >= 1 && nrhs <= MAXR, "nrhs must be in [1, 16]");
  TFM_REQUIRE(max_iter >= 1, "max_iter must be positive");
  auto *h = new tfm_pcg();
  h->n = n;
  h->nrhs = nrhs;
  h->max_iter = max_iter;
  (void)hipGetDevice(&h->device);
  int rc = pcg_alloc(h);
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

namespace {

int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Host view of a running inner solve.
struct Inner {
  tfm_pcg *h;
  hipStream_t s;
  const tfm_mesh_t *m;
  int nrhs, nrp;
  PcgState st;
};

// Runs PCG iterations on the interleaved vectors of the handle. `spmv(x, y,
// pq)` computes y = A x and, when pq is non-null, the dots p.q per rhs.
template <typename T, int ND, class SpMV>
int inner_pcg(Inner &in, SpMV spmv, const T *minv, int max_iter, int *iters_out,
              bool x_is_zero) {
  tfm_pcg *h = in.h;
  hipStream_t s = in.s;
  const int64_t n_nod = in.m ? in.m->n_nod : h->n;
  const int nrp = in.nrp;
  T *x = (T *)h->x, *r = (T *)h->r, *z = (T *)h->z, *p = (T *)h->p, *q = (T *)h->q;
  const T *b = (const T *)h->b;
  const unsigned gv = grid_for(n_nod * nrp, BS);
  if (!x_is_zero) {
    int rc = spmv(x, q, nullptr, nullptr);
    if (rc != TFM_OK) return rc;
  }
  TFM_CHECK_HIP(hipMemsetAsync(h->alldone, 0, sizeof(int), s));
  hipLaunchKernelGGL((k_pcg_init<T, ND>), dim3(gv), dim3(BS), 0, s, n_nod, nrp, in.nrhs, b,
                     x_is_zero ? (const T *)nullptr : (const T *)q, r, p, minv, in.st);
  TFM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_check_done, dim3(1), dim3(1), 0, s, nrp, (const int *)h->done, h->alldone);
  TFM_CHECK_LAUNCH();
  const int CHK = 16;
  int it = 0;
  bool finished = false;
  while (it < max_iter && !finished) {
    const int stop = std::min(max_iter, it + CHK);
    for (; it < stop; ++it) {
      double *pq = in.st.scal + (size_t)it * nrp * NSCAL + S_PQ;
      int rc = spmv(p, q, h->alldone, pq);
      if (rc != TFM_OK) return rc;
      hipLaunchKernelGGL((k_pcg_update<T, ND>), dim3(gv), dim3(BS), 0, s, n_nod, nrp, it, x, r,
                         z, (const T *)p, (const T *)q, minv, in.st);
      TFM_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_check_done, dim3(1), dim3(1), 0, s, nrp, (const int *)h->done,
                         h->alldone);
      TFM_CHECK_LAUNCH();
      hipLaunchKernelGGL((k_pcg_dir<T, ND>), dim3(gv), dim3(BS), 0, s, n_nod, nrp, it,
                         (const T *)z, p, in.st);
      TFM_CHECK_LAUNCH();
    }
    TFM_CHECK_HIP(hipMemcpyAsync(h->h_done, h->done, sizeof(int) * nrp, hipMemcpyDeviceToHost, s));
    TFM_CHECK_HIP(hipStreamSynchronize(s));
    finished = true;
    for (int r2 = 0; r2 < nrp; ++r2) finished &= h->h_done[r2] >= 0;
  }
  if (it == 0) {  // converged before the first iteration
    TFM_CHECK_HIP(hipMemcpyAsync(h->h_done, h->done, sizeof(int) * nrp, hipMemcpyDeviceToHost, s));
    TFM_CHECK_HIP(hipStreamSynchronize(s));
    finished = true;
    for (int r2 = 0; r2 < nrp; ++r2) finished &= h->h_done[r2] >= 0;
  }
  for (int r2 = 0; r2 < in.nrhs; ++r2) iters_out[r2] = h->h_done[r2] >= 0 ? h->h_done[r2] : it;
  return finished ? TFM_OK : TFM_ERR_NOCONV;
}

template <typename T>
int inner_dispatch(Inner &in, const T *val, const T *minv, int max_iter, int *iters, bool x0) {
  tfm_pcg *h = in.h;
  auto spmv = [&](const T *xin, T *yout, const int *alldone, double *pq) {
    return spmv_any<T>(in.s, *in.m, val, xin, yout, in.nrp, alldone, pq, h->partials, h->counter);
  };
  switch (in.m->ndof) {
    case 1: return inner_pcg<T, 1>(in, spmv, minv, max_iter, iters, x0);
    case 2: return inner_pcg<T, 2>(in, spmv, minv, max_iter, iters, x0);
    case 3: return inner_pcg<T, 3>(in, spmv, minv, max_iter, iters, x0);
    default: TFM_REQUIRE(false, "ndof must be 1, 2 or 3");
  }
}

template <typename T>
int build_precond(hipStream_t s, const tfm_mesh_t &m, const double *dblk, int pc, T *minv) {
  const unsigned g = (unsigned)((m.n_nod + 255) / 256);
  switch (m.ndof) {
    case 1: hipLaunchKernelGGL((k_precond<T, 1>), dim3(g), dim3(256), 0, s, m.n_nod, dblk, pc, minv); break;
    case 2: hipLaunchKernelGGL((k_precond<T, 2>), dim3(g), dim3(256), 0, s, m.n_nod, dblk, pc, minv); break;
    case 3: hipLaunchKernelGGL((k_precond<T, 3>), dim3(g), dim3(256), 0, s, m.n_nod, dblk, pc, minv); break;
    default: TFM_REQUIRE(false, "ndof must be 1, 2 or 3");
  }
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

// FP64 residual res = b - K_bc x for one rhs; returns ||res||^2 and ||b||^2.
int true_residual(tfm_pcg *h, hipStream_t s, const tfm_system_t &sys, const double *b,
                  const double *x, double *res, double *rr, double *bbv) {
  const tfm_mesh_t &m = *sys.mesh;
  const int64_t n = m.n_nod * m.ndof;
  const unsigned g1 = (unsigned)((n + 255) / 256);
  if (sys.val64) {
    int rc = spmv_any<double>(s, m, sys.val64, x, h->f64, 1, nullptr, nullptr, nullptr, nullptr);
    if (rc != TFM_OK) return rc;
    hipLaunchKernelGGL(k_residual, dim3(grid_for(n, BS)), dim3(BS), 0, s, n,
                       (const uint8_t *)nullptr, b, x, (const double *)h->f64, res, h->res2,
                       h->partials, h->counter);
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
    hipLaunchKernelGGL(k_residual, dim3(grid_for(n, BS)), dim3(BS), 0, s, n, sys.con, b, x,
                       (const double *)h->f64, res, h->res2, h->partials, h->counter);
  }
  TFM_CHECK_LAUNCH();
  TFM_CHECK_HIP(hipMemcpyAsync(h->h_res, h->res2, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipStreamSynchronize(s));
  *rr = h->h_res[0];
  *bbv = h->h_res[1];
  return TFM_OK;
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
  hipStream_t s = (hipStream_t)stream;
  const int nrp = pow2_at_least(nrhs);
  const unsigned gi = (unsigned)((n * nrp + 255) / 256);
  Inner in{h, s, &m, nrhs, nrp,
           PcgState{h->scal, h->bb, h->tol2, h->done, h->alldone, h->partials, h->counter}};
  std::vector<double> htol(nrp, 0.0);
  for (int r = 0; r < nrhs; ++r) iters[r] = 0;
  if (outer) *outer = 0;

  if (prec == TFM_PREC_FP64) {
    // all-FP64 PCG on the assembled values, stopping on ||r|| <= rtol ||b||
    std::vector<double> bb(nrhs);
    for (int r = 0; r < nrhs; ++r) {
      double rr2, bb2;
      int rc = true_residual(h, s, *sys, b + (int64_t)r * n, x + (int64_t)r * n, h->r64, &rr2, &bb2);
      if (rc != TFM_OK) return rc;
      bb[r] = bb2;
      htol[r] = rtol * rtol * bb2;
    }
    TFM_CHECK_HIP(hipMemcpyAsync(h->tol2, htol.data(), sizeof(double) * nrp, hipMemcpyHostToDevice, s));
    int rc = build_precond<double>(s, m, sys->diag_blocks, pc, h->minv);
    if (rc != TFM_OK) return rc;
    hipLaunchKernelGGL((k_interleave<double, double>), dim3(gi), dim3(256), 0, s, n, nrhs, nrp, b, h->b);
    hipLaunchKernelGGL((k_interleave<double, double>), dim3(gi), dim3(256), 0, s, n, nrhs, nrp, x, h->x);
    TFM_CHECK_LAUNCH();
    std::vector<int> it(nrp, 0);
    const int rc_in = inner_dispatch<double>(in, sys->val64, h->minv, h->max_iter, it.data(), false);
    if (rc_in != TFM_OK && rc_in != TFM_ERR_NOCONV) return rc_in;
    const unsigned go = (unsigned)((n * nrhs + 255) / 256);
    hipLaunchKernelGGL((k_deinterleave<double>), dim3(go), dim3(256), 0, s, n, nrhs, nrp, (const double *)h->x, x, 0);
    TFM_CHECK_LAUNCH();
    for (int r = 0; r < nrhs; ++r) {
      double rr2, bb2;
      rc = true_residual(h, s, *sys, b + (int64_t)r * n, x + (int64_t)r * n, h->r64, &rr2, &bb2);
      if (rc != TFM_OK) return rc;
      relres[r] = bb2 > 0 ? std::sqrt(rr2 / bb2) : std::sqrt(rr2);
      iters[r] = it[r];
    }
    if (outer) *outer = 1;
    if (rc_in == TFM_ERR_NOCONV) {
      set_error("PCG did not converge within max_iter iterations");
      return TFM_ERR_NOCONV;
    }
    return TFM_OK;
  }

  // mixed precision: FP32 PCG on the correction, FP64 matrix-free residual
  int rc = build_precond<float>(s, m, sys->diag_blocks, pc, (float *)h->minv);
  if (rc != TFM_OK) return rc;
  std::vector<double> bnorm2(nrhs), rnorm2(nrhs);
  int total_budget = h->max_iter;
  const int max_outer = 30;
  int k = 0;
  for (;; ++k) {
    // residuals of every rhs in FP64 (into h->r64 per rhs, then staged as FP32)
    bool all_ok = true;
    for (int r = 0; r < nrhs; ++r) {
      double rr2, bb2;
      rc = true_residual(h, s, *sys, b + (int64_t)r * n, x + (int64_t)r * n, h->r64, &rr2, &bb2);
      if (rc != TFM_OK) return rc;
      bnorm2[r] = bb2;
      rnorm2[r] = rr2;
      relres[r] = bb2 > 0 ? std::sqrt(rr2 / bb2) : std::sqrt(rr2);
      const bool ok = rr2 <= rtol * rtol * bb2;
      all_ok &= ok;
      // stage this rhs' residual as column r of the inner right-hand side
      // (a zero column for an rhs that is already converged)
      hipLaunchKernelGGL(k_set_column, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, nrp,
                         r, (const double *)h->r64, (float *)h->b, ok ? 1 : 0);
      TFM_CHECK_LAUNCH();
      // inner tolerance: inner_rtol, but no tighter than what the outer test needs
      const double need = ok ? 1.0 : std::max(inner_rtol, 0.3 * rtol * std::sqrt(bb2 / std::max(rr2, 1e-300)));
      htol[r] = ok ? -1.0 : need * need * rr2;
    }
    if (all_ok || k >= max_outer || total_budget <= 0) break;
    for (int r = nrhs; r < nrp; ++r) {
      hipLaunchKernelGGL(k_set_column, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, nrp,
                         r, (const double *)h->r64, (float *)h->b, 1);
      TFM_CHECK_LAUNCH();
      htol[r] = -1.0;
    }
    TFM_CHECK_HIP(hipMemcpyAsync(h->tol2, htol.data(), sizeof(double) * nrp, hipMemcpyHostToDevice, s));
    TFM_CHECK_HIP(hipMemsetAsync(h->x, 0, sizeof(float) * (size_t)n * nrp, s));
    std::vector<int> it(nrp, 0);
    const int rc_in = inner_dispatch<float>(in, sys->val32, (const float *)h->minv, total_budget,
                                            it.data(), true);
    if (rc_in != TFM_OK && rc_in != TFM_ERR_NOCONV) return rc_in;
    int used = 0;
    for (int r = 0; r < nrhs; ++r) {
      iters[r] += it[r];
      used = std::max(used, it[r]);
    }
    total_budget -= std::max(used, 1);
    const unsigned go = (unsigned)((n * nrhs + 255) / 256);
    hipLaunchKernelGGL((k_deinterleave<float>), dim3(go), dim3(256), 0, s, n, nrhs, nrp,
                       (const float *)h->x, x, 1);
    TFM_CHECK_LAUNCH();
  }
  if (outer) *outer = k;
  bool ok = true;
  for (int r = 0; r < nrhs; ++r) ok &= relres[r] <= rtol;
  if (!ok) {
    set_error("mixed-precision PCG did not reach the tolerance within max_iter iterations");
    return TFM_ERR_NOCONV;
  }
  return TFM_OK;
}

// ===================================================================
// General CSR matrices (no mesh): the operator and Jacobi-PCG behind
// torchfem.sparse for user-supplied sparse matrices.
// ===================================================================
namespace {

__global__ void k_csr_diag_block(int64_t n, const int64_t *rowptr, const int32_t *col,
                                 const double *val, double *d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    if (col[k] == i) v += val[k];
  d[i] = v;
}

int csr_spmv(hipStream_t s, int64_t n, const int64_t *rowptr, const int32_t *col,
             const double *val, const double *x, double *y, const int *alldone, double *pq,
             double *partials, unsigned *counter) {
  const unsigned g = grid_for(n, BS / 8);
  hipLaunchKernelGGL((k_spmv_csr<8>), dim3(g), dim3(BS), 0, s, n, rowptr, col, val, x, y,
                     alldone, pq, partials, counter);
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

}  // namespace

extern "C" int tfm_csr_spmv(tfm_stream_t s, int64_t n, const int64_t *rowptr, const int32_t *col,
                            const double *val, const double *x, double *y) {
  TFM_REQUIRE(n >= 0, "negative size");
  if (n == 0) return TFM_OK;
  TFM_REQUIRE(rowptr && col && val && x && y, "null argument");
  return csr_spmv((hipStream_t)s, n, rowptr, col, val, x, y, nullptr, nullptr, nullptr, nullptr);
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
  hipLaunchKernelGGL(k_csr_diag_block, dim3(g1), dim3(256), 0, s, n, rowptr, col, val, h->f64);
  TFM_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_precond<double, 1>), dim3(g1), dim3(256), 0, s, n, (const double *)h->f64,
                     (int)TFM_PC_JACOBI, h->minv);
  TFM_CHECK_LAUNCH();
  // ||b||^2 for the absolute tolerance
  int rc = csr_spmv(s, n, rowptr, col, val, x, h->f64, nullptr, nullptr, nullptr, nullptr);
  if (rc != TFM_OK) return rc;
  hipLaunchKernelGGL(k_residual, dim3(grid_for(n, BS)), dim3(BS), 0, s, n,
                     (const uint8_t *)nullptr, b, x, (const double *)h->f64, h->r64, h->res2,
                     h->partials, h->counter);
  TFM_CHECK_LAUNCH();
  TFM_CHECK_HIP(hipMemcpyAsync(h->h_res, h->res2, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipStreamSynchronize(s));
  const double bb2 = h->h_res[1];
  double tol2 = rtol * rtol * bb2;
  TFM_CHECK_HIP(hipMemcpyAsync(h->tol2, &tol2, sizeof(double), hipMemcpyHostToDevice, s));
  TFM_CHECK_HIP(hipMemcpyAsync(h->b, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  TFM_CHECK_HIP(hipMemcpyAsync(h->x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  Inner in{h, s, nullptr, 1, 1,
           PcgState{h->scal, h->bb, h->tol2, h->done, h->alldone, h->partials, h->counter}};
  auto spmv = [&](const double *xin, double *yout, const int *alldone, double *pq) {
    return csr_spmv(s, n, rowptr, col, val, xin, yout, alldone, pq, h->partials, h->counter);
  };
  int it = 0;
  const int rc_in = inner_pcg<double, 1>(in, spmv, (const double *)h->minv, h->max_iter, &it, false);
  if (rc_in != TFM_OK && rc_in != TFM_ERR_NOCONV) return rc_in;
  TFM_CHECK_HIP(hipMemcpyAsync(x, h->x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  rc = csr_spmv(s, n, rowptr, col, val, x, h->f64, nullptr, nullptr, nullptr, nullptr);
  if (rc != TFM_OK) return rc;
  hipLaunchKernelGGL(k_residual, dim3(grid_for(n, BS)), dim3(BS), 0, s, n,
                     (const uint8_t *)nullptr, b, x, (const double *)h->f64, h->r64, h->res2,
                     h->partials, h->counter);
  TFM_CHECK_LAUNCH();
  TFM_CHECK_HIP(hipMemcpyAsync(h->h_res, h->res2, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipStreamSynchronize(s));
  *iters = it;
  *relres = bb2 > 0 ? std::sqrt(h->h_res[0] / bb2) : std::sqrt(h->h_res[0]);
  if (rc_in == TFM_ERR_NOCONV) {
    set_error("CG did not converge within max_iter iterations");
    return TFM_ERR_NOCONV;
  }
  return TFM_OK;
}
