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
#include <cstdlib>
#include <functional>
#include <type_traits>
#include <vector>

namespace tfm {

constexpr int BS = 256;   // threads per block of the solver kernels
constexpr int BSV = 256;  // threads per block of the PCG vector update (kernel B)
constexpr int RBA = 2048; // max blocks of the SpMV kernels (= eta partials per rhs)
constexpr int RBV = 512;  // max blocks of the vector kernels whose partials every block of A sums
constexpr int RBW = 4096; // max blocks of the vector kernels when a scalar kernel sums them
constexpr int MAXR = 16;  // right-hand sides per solve
constexpr int CGB = 1;    // partial sets of the coarse correction (after the update's)
constexpr int NSCAL = 4;  // per-iteration scalars: rz, pq, rr, spare
constexpr int PCG_CHK_MAX = 64;  // most iterations between two host checks
enum { S_RZ = 0, S_PQ = 1, S_RR = 2 };

// ---------------------------------------------------------- reductions
// Block partials: thread t contributes v[k] to right-hand side t % NRP; the
// block's sums go to part[blockIdx.x][k][r]. Each wave first folds its lanes
// of equal t % NRP with a butterfly, then one LDS stage adds the waves (a
// fixed order, so the result is reproducible).
template <int NV, int NRP, int BSZ = BS>
__device__ __forceinline__ void write_block_partials(const double (&v)[NV], double *sh,
                                                     double *part) {
  constexpr int NW = BSZ / 64;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  double w[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) w[k] = v[k];
#pragma unroll
  for (int off = NRP; off < 64; off <<= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) w[k] += __shfl_xor(w[k], off, 64);
  __syncthreads();  // sh may still be read by the caller
  if (lane < NRP)
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[(wave * NV + k) * NRP + lane] = w[k];
  __syncthreads();
  if (tid < NV * NRP) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < NW; ++q) t += sh[q * NV * NRP + tid];
    part[(size_t)blockIdx.x * NV * NRP + tid] = t;
  }
}

// Sums the NVT partial slots written by nblk blocks into out[0..NVT) in LDS,
// in a fixed order, so that every block that calls it gets the same bits.
// Thread t adds the partials b = t/NVT + k*CH of slot t % NVT with four
// loads in flight; a wave butterfly and one LDS stage over the waves follow
// (one thread adding a 256-long LDS column serially costs microseconds).
template <int NVT, int BSZ = BS>
__device__ __forceinline__ void sum_partials(const double *part, int nblk, double *sh,
                                             double *out) {
  static_assert(64 % NVT == 0 && BSZ % 64 == 0, "slots must tile a wave");
  constexpr int CH = BSZ / NVT, NW = BSZ / 64;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int slot = tid % NVT;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int b = tid / NVT;
  for (; b + 3 * CH < nblk; b += 4 * CH) {
    a0 += part[(size_t)b * NVT + slot];
    a1 += part[(size_t)(b + CH) * NVT + slot];
    a2 += part[(size_t)(b + 2 * CH) * NVT + slot];
    a3 += part[(size_t)(b + 3 * CH) * NVT + slot];
  }
  for (; b < nblk; b += CH) a0 += part[(size_t)b * NVT + slot];
  double acc = (a0 + a1) + (a2 + a3);
#pragma unroll
  for (int off = NVT; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
  __syncthreads();  // sh may still be read by the caller
  if (lane < NVT) sh[wave * NVT + lane] = acc;
  __syncthreads();
  if (tid < NVT) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < NW; ++q) t += sh[q * NVT + tid];
    out[tid] = t;
  }
  __syncthreads();
}

// ------------------------------------------------------- node-block SpMV
// K contiguous values loaded as one unit (one global_load_dwordxN per lane
// where the width allows; the compiler otherwise splits rows unevenly).
template <typename T, int K> struct Vec {
  T v[K];
};

template <typename T, int K>
__device__ __forceinline__ Vec<T, K> ldv(const T *p) {
  Vec<T, K> r;
  __builtin_memcpy(&r, p, sizeof(r));
  return r;
}

template <typename T, int K>
__device__ __forceinline__ void stv(T *p, const Vec<T, K> &v) {
  __builtin_memcpy(p, &v, sizeof(v));
}

// Neighbour blocks per lane whose loads are issued before any FMA: two for
// 16-lane FP64 teams (a 27-block Hexa1 row in one pass), one for 32-lane
// teams and for the FP32 products summed in FP64 (whose wider products
// would otherwise cost the kernel two waves per SIMD).
template <typename T, int TL, int NRP> constexpr int hoisted_blocks() {
  return (TL < 32 && NRP <= 2 && sizeof(T) == 8) ? 2 : 1;
}

// FP32 values for one or two right-hand sides are stored in difference
// form: the diagonal block of node i holds the row-sum block
// R_i = sum_j K_ij (summed in FP64, then rounded) instead of K_ii, and the
// product is
//   (K x)_i = sum_{j != i} K_ij (x_j - x_i) + R_i x_i,
// an identity for any matrix. Rounding K_ij to FP32 then perturbs K x by
// terms of size eps |K| |x_j - x_i| instead of eps |K| |x|: for a smooth x,
// where K x is small by cancellation, the FP32 operator stays close to the
// FP64 one, and the FP32 CG between two FP64 refinements loses far fewer
// iterations to the mismatch. These products are also summed in FP64 (a
// product of two FP32 values is exact in FP64; no cost on a bandwidth-bound
// kernel). The many-rhs kernels keep the plain rounded values and FP32 sums:
// their registers hold ND*NRP sums already, and the L2-bound gathers, not
// the iteration count, limit them.
template <typename T, int NRP> constexpr bool diff_form() { return sizeof(T) == 4 && NRP <= 2; }
template <typename T, int NRP>
using acc_t = typename std::conditional<(sizeof(T) == 4 && NRP <= 2), double, T>::type;

// A team of TL lanes per node; lane l owns the neighbour blocks p = l, l+TL,
// ...: it loads adj[p], the ND rows of ND values of the block, and the
// ND x NRP entries of x at that neighbour (contiguous in the interleaved
// layout), so each lane does ND*ND*NRP FMAs per index load. The first KB
// blocks of every lane are loaded before any arithmetic, so a lane keeps
// KB*(ND+1) loads in flight. The butterfly leaves the team sums in every
// lane; output element e = a*NRP + r is handled by lane e % TL.
// Where node n's row data start: its first adjacency slot, its degree and
// its first value. Loaded ahead of the product by the PCG kernel.
struct NodePtr {
  int a0, deg;
  int64_t v0;
};

__device__ __forceinline__ NodePtr node_ptr(const tfm_mesh_t &m, int64_t n) {
  const int a0 = m.adj_ptr[n];
  return NodePtr{a0, m.adj_ptr[n + 1] - a0, m.val_ptr[n]};
}

template <typename T, int ND, int TL, int NRP, bool D16, bool BFLY = true>
__device__ __forceinline__ void node_product_at(const tfm_mesh_t &m, const T *__restrict__ val,
                                                const T *__restrict__ x, int64_t n,
                                                const NodePtr &np, int lane,
                                                acc_t<T, NRP> (&acc)[ND][NRP]) {
  using A = acc_t<T, NRP>;
  // neighbour k of the node: from the 16-bit offsets (D16) or the int32 indices
  auto nbr = [&](int k) -> int64_t { return D16 ? n + (int64_t)m.adj16[k] : (int64_t)m.adj[k]; };
  constexpr int KB = hoisted_blocks<T, TL, NRP>();
  // products and sums in A: a product of two FP32 values is exact in FP64,
  // so the row sum carries no rounding beyond that of the stored operands
  A sum[ND][NRP];
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int r = 0; r < NRP; ++r) sum[a][r] = A(0);
  const int a0 = np.a0;
  const int deg = np.deg;
  const T *vn = val + np.v0;
  // difference form: the node's own entries, subtracted from every neighbour's
  Vec<T, diff_form<T, NRP>() ? ND * NRP : 1> xs{};
  if constexpr (diff_form<T, NRP>())
    if (deg > 0) xs = ldv<T, ND * NRP>(x + n * ND * NRP);
  auto fma_block = [&](const Vec<T, ND> (&rows)[ND], const Vec<T, ND * NRP> &xv, bool self) {
    T d[ND * NRP];
#pragma unroll
    for (int q = 0; q < ND * NRP; ++q) {
      if constexpr (diff_form<T, NRP>()) d[q] = self ? xv.v[q] : xv.v[q] - xs.v[q];
      else d[q] = xv.v[q];
    }
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int b = 0; b < ND; ++b)
#pragma unroll
        for (int r = 0; r < NRP; ++r) sum[a][r] += (A)rows[a].v[b] * (A)d[b * NRP + r];
  };
  if (deg > 0) {
    const int L = ND * deg;
    int col[KB];
    Vec<T, ND> rows[KB][ND];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int p = lane + k * TL;
      col[k] = -1;
      if (p < deg) {
        col[k] = (int)nbr(a0 + p);
#pragma unroll
        for (int a = 0; a < ND; ++a) rows[k][a] = ldv<T, ND>(vn + a * L + p * ND);
      }
    }
#pragma unroll
    for (int k = 0; k < KB; ++k)
      if (col[k] >= 0)
        fma_block(rows[k], ldv<T, ND * NRP>(x + (int64_t)col[k] * ND * NRP), col[k] == n);
    for (int p = lane + KB * TL; p < deg; p += TL) {  // rows wider than the hoisted blocks
      Vec<T, ND> rw[ND];
#pragma unroll
      for (int a = 0; a < ND; ++a) rw[a] = ldv<T, ND>(vn + a * L + p * ND);
      const int64_t c = nbr(a0 + p);
      fma_block(rw, ldv<T, ND * NRP>(x + c * ND * NRP), c == n);
    }
  } else if (lane < ND) {  // orphan node: only its diagonal is stored
#pragma unroll
    for (int a = 0; a < ND; ++a)
      if (a == lane)
#pragma unroll
        for (int r = 0; r < NRP; ++r) sum[a][r] = (A)vn[a] * (A)x[(n * ND + a) * NRP + r];
  }
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int r = 0; r < NRP; ++r) acc[a][r] = sum[a][r];
  if constexpr (BFLY) {
#pragma unroll
    for (int off = 1; off < TL; off <<= 1)
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int r = 0; r < NRP; ++r) acc[a][r] += __shfl_xor(acc[a][r], off, TL);
  }
}

// FP32 with four or more rhs: the team's ND*NRP sums per lane are reduced
// through LDS instead of a butterfly (48 values over 5 butterfly steps cost
// 240 lane exchanges per node for 16 rhs). Each lane writes its sums as one
// row of a TL x (ND*NRP) table (padded to an odd row length, so neither
// the row writes nor the column reads conflict), then lane l adds column
// e = l + k*TL over the TL rows in order. outv[k] is entry e of the node's
// outputs. The team is inside one wave, whose LDS accesses complete in
// order, so only the compiler has to be kept from reordering them.
template <typename T, int TL, int NRP> constexpr bool lds_reduce() {
  return sizeof(T) == 4 && NRP >= 4 && TL < 64;
}

template <typename T, int ND, int TL, int NRP, bool D16>
__device__ __forceinline__ void node_product_lds(const tfm_mesh_t &m, const T *__restrict__ val,
                                                 const T *__restrict__ x, int64_t n,
                                                 const NodePtr &np, int lane, int team,
                                                 T (&outv)[(ND * NRP + TL - 1) / TL]) {
  constexpr int NE = ND * NRP, KE = (NE + TL - 1) / TL, PAD = NE + 1;
  __shared__ T red[BS / TL][TL][PAD];
  acc_t<T, NRP> acc[ND][NRP];
  node_product_at<T, ND, TL, NRP, D16, false>(m, val, x, n, np, lane, acc);
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int r = 0; r < NRP; ++r) red[team][lane][a * NRP + r] = acc[a][r];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int k = 0; k < KE; ++k) {
    const int e = lane + k * TL;
    T t = T(0);
    if (e < NE)
      for (int q = 0; q < TL; ++q) t += red[team][q][e];
    outv[k] = t;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, int ND, int TL, int NRP, bool D16>
__device__ __forceinline__ void node_product(const tfm_mesh_t &m, const T *__restrict__ val,
                                             const T *__restrict__ x, int64_t n, int lane,
                                             acc_t<T, NRP> (&acc)[ND][NRP]) {
  node_product_at<T, ND, TL, NRP, D16>(m, val, x, n, node_ptr(m, n), lane, acc);
}

// Several right-hand sides (NRP >= 4): one 64-lane wave per node. Lane
// l = q*NRP + 4*g + t (t < 4) works on rhs group g (rhs 4g..4g+3) and on the
// neighbour blocks p = s, s + S, ... of slice s = 4*q + t (S = 256/NRP
// slices). The lanes of one slice load the same block rows (one transaction
// serves them) and the four rhs of their group at the neighbour in one
// 16-byte load (contiguous in the interleaved layout). A butterfly over the
// slices (lane offsets 1, 2 and multiples of NRP) leaves the ND x 4 sums of
// group g in every lane of the group, so lane l holds entry e = l of the
// node's ND*NRP outputs (rhs l % NRP). ND*4 accumulators per lane instead of
// ND*NRP keep the kernel's occupancy.
template <int TL, int NRP> constexpr bool rhs_lanes() { return TL == 64 && NRP >= 4; }

template <typename T, int ND, int NRP, bool D16>
__device__ __forceinline__ void node_product_mr(const tfm_mesh_t &m, const T *__restrict__ val,
                                                const T *__restrict__ x, int64_t n,
                                                const NodePtr &np, int lane, T (&acc)[ND][4]) {
  constexpr int S = 256 / NRP;
  const int j = lane % NRP, g = j / 4;
  const int sl = (lane / NRP) * 4 + (j % 4);
  auto nbr = [&](int k) -> int64_t { return D16 ? n + (int64_t)m.adj16[k] : (int64_t)m.adj[k]; };
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[a][t] = T(0);
  const T *vn = val + np.v0;
  auto block = [&](int p, int64_t c) {
    Vec<T, ND> rows[ND];
    Vec<T, 4> xv[ND];
#pragma unroll
    for (int a = 0; a < ND; ++a) rows[a] = ldv<T, ND>(vn + a * ND * np.deg + p * ND);
#pragma unroll
    for (int b = 0; b < ND; ++b) xv[b] = ldv<T, 4>(x + (c * ND + b) * NRP + 4 * g);
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int b = 0; b < ND; ++b)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[a][t] += rows[a].v[b] * xv[b].v[t];
  };
  if (np.deg > 0) {
    int p = sl;
    for (; p + S < np.deg; p += 2 * S) {  // two blocks in flight per lane
      const int64_t c0 = nbr(np.a0 + p), c1 = nbr(np.a0 + p + S);
      block(p, c0);
      block(p + S, c1);
    }
    if (p < np.deg) block(p, nbr(np.a0 + p));
  } else if (sl == 0) {  // orphan node: only its diagonal is stored
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[a][t] = vn[a] * x[(n * ND + a) * NRP + 4 * g + t];
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    if (off >= 4 && off < NRP) continue;  // these offsets change the rhs group
#pragma unroll
    for (int a = 0; a < ND; ++a)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[a][t] += __shfl_xor(acc[a][t], off, 64);
  }
}

// Entry e = a*NRP + r of a node's outputs from the group sums of the lane
// that holds rhs r (a select chain, no indexed registers).
template <typename T, int ND, int NRP>
__device__ __forceinline__ T pick_entry(const T (&acc)[ND][4], int e) {
  const int a = e / NRP, t = e % 4;
  T w = acc[0][0];
#pragma unroll
  for (int aa = 0; aa < ND; ++aa)
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
      if (aa == a && tt == t) w = acc[aa][tt];
  return w;
}


// y = K x over the node blocks (every right-hand side of the interleaved x).
template <typename T, int ND, int TL, int NRP, bool D16>
__global__ void __launch_bounds__(BS) k_spmv(tfm_mesh_t m, const T *__restrict__ val,
                                             const T *__restrict__ x, T *__restrict__ y) {
  constexpr int TPB = BS / TL;
  const int lane = threadIdx.x % TL, team = threadIdx.x / TL;
  for (int64_t n = (int64_t)blockIdx.x * TPB + team; n < m.n_nod; n += (int64_t)gridDim.x * TPB) {
    if constexpr (rhs_lanes<TL, NRP>()) {
      T am[ND][4];
      node_product_mr<T, ND, NRP, D16>(m, val, x, n, node_ptr(m, n), lane, am);
      if (lane < ND * NRP) y[n * ND * NRP + lane] = pick_entry<T, ND, NRP>(am, lane);
    } else if constexpr (lds_reduce<T, TL, NRP>()) {
      constexpr int KE = (ND * NRP + TL - 1) / TL;
      T outv[KE];
      node_product_lds<T, ND, TL, NRP, D16>(m, val, x, n, node_ptr(m, n), lane, team, outv);
#pragma unroll
      for (int k = 0; k < KE; ++k)
        if (lane + k * TL < ND * NRP) y[n * ND * NRP + lane + k * TL] = outv[k];
    } else {
      acc_t<T, NRP> acc[ND][NRP];
      node_product<T, ND, TL, NRP, D16>(m, val, x, n, lane, acc);
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int r = 0; r < NRP; ++r)
          if ((a * NRP + r) % TL == lane) y[(n * ND + a) * NRP + r] = (T)acc[a][r];
    }
  }
}

// ------------------------------------------------------- general CSR SpMV
template <int TL>
__device__ __forceinline__ double csr_row(const int64_t *rowptr, const int32_t *col,
                                          const double *val, const double *x, int64_t i,
                                          int lane) {
  double acc = 0.0;
  for (int64_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += TL) acc += val[k] * x[col[k]];
#pragma unroll
  for (int off = 1; off < TL; off <<= 1) acc += __shfl_xor(acc, off, TL);
  return acc;
}

template <int TL>
__global__ void __launch_bounds__(BS) k_spmv_csr(int64_t n, const int64_t *rowptr,
                                                 const int32_t *col, const double *val,
                                                 const double *x, double *y) {
  constexpr int TPB = BS / TL;
  const int lane = threadIdx.x % TL, team = threadIdx.x / TL;
  for (int64_t i = (int64_t)blockIdx.x * TPB + team; i < n; i += (int64_t)gridDim.x * TPB) {
    const double acc = csr_row<TL>(rowptr, col, val, x, i, lane);
    if (lane == 0) y[i] = acc;
  }
}

// ------------------------------------------------------- preconditioner
// Inverse of the nodal diagonal blocks (block Jacobi) or of the diagonal
// (Jacobi) in the inner precision; a zero pivot keeps the identity. The
// inverse is symmetric and stored packed: the upper triangle, row by row,
// ND*(ND+1)/2 values per node (pc_index).
template <int ND> constexpr int pc_size() { return ND * (ND + 1) / 2; }

template <int ND>
__host__ __device__ constexpr int pc_index(int a, int b) {
  return a <= b ? a * ND - a * (a - 1) / 2 + (b - a) : b * ND - b * (b - 1) / 2 + (a - b);
}

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
  // the two triangles of the inverse agree up to rounding; their mean keeps
  // the preconditioner exactly symmetric
  for (int a = 0; a < ND; ++a)
    for (int b = a; b < ND; ++b)
      minv[n * pc_size<ND>() + pc_index<ND>(a, b)] = (T)(0.5 * (R[a][b] + R[b][a]));
}

template <typename T, int ND, int NRP>
__device__ __forceinline__ void apply_pc(const T *minv, const T *r, int64_t n, int rr,
                                         T (&z)[ND], T (&rv)[ND]) {
#pragma unroll
  for (int a = 0; a < ND; ++a) rv[a] = r[(n * ND + a) * NRP + rr];
  const Vec<T, pc_size<ND>()> mi = ldv<T, pc_size<ND>()>(minv + n * pc_size<ND>());
#pragma unroll
  for (int a = 0; a < ND; ++a) {
    T acc = T(0);
#pragma unroll
    for (int b = 0; b < ND; ++b) acc += mi.v[pc_index<ND>(a, b)] * rv[b];
    z[a] = acc;
  }
}

// ------------------------------------------------------------ PCG steps
// One iteration is two kernels, the form of PCG that applies the operator to
// the preconditioned residual u = M r instead of to the search direction:
//   A (spmv_dir): gamma = r.u and rr = r.r from the partials of B; stop test;
//                 beta = gamma / gamma_prev; w = K u; p = u + beta p;
//                 s = w + beta s (so that s = K p); partials of eta = p.s
//   B (update):   alpha = gamma / eta; x += alpha p; r -= alpha s; u = M r;
//                 partials of (r.u, r.r)
// In exact arithmetic these are the iterates of textbook PCG; eta is a dot
// product of the stored vectors (not a recurrence), and every dot is FP64.
// Per-rhs coefficients prepared by a one-block scalar kernel for a wide-grid
// kernel (beta for A, alpha for B), read once per block instead of every
// block summing partial sets; `any` is 0 once no rhs is running.
struct PcgCoef {
  double c[MAXR];
  int live[MAXR];
  int any;
  int pad;
};

struct PcgState {
  double *scal;      // [(max_iter + 2) * NRP * NSCAL]: gamma, eta, rr per iteration
  double *tol2;      // [NRP] stop the current segment when rr <= tol2 (< 0: rhs not running)
  double *rzold;     // [NRP] gamma that divides the first beta of a segment (0: restart)
  int *done;         // [NRP] iteration at which rhs r stopped, -1 while running
  int *alldone;      // 1 once every rhs stopped
  double *part_pq;   // [pq_cap * NRP]   eta partials (A)
  double *part_upd;  // [RBW * 2 * NRP]  (gamma, rr) partials (B and the segment start)
  double *part_aux;  // [RBW * NRP]      r.p partials (segment start)
  int pq_cap;        // blocks of A whose partials part_pq can hold
  PcgCoef *coef_a;   // beta of A (scalar-kernel mode)
  PcgCoef *coef_b;   // alpha of B (scalar-kernel mode)
  int xcd;           // grid-stride kernels give each XCD one contiguous share (block_span)
  int *segf;         // [NRP] iteration at which rhs r's current segment started
};

// Items [first, end) with step `stride` of a block of a grid-stride kernel,
// per_block items per round. Workgroups go to the 8 XCDs round-robin (block
// b on XCD b % 8); with `xcd` the blocks of one XCD share one contiguous
// eighth of the items, so what an XCD gathers and streams stays in its L2.
// The eighths are whole multiples of per_block, so item t is still handled
// by thread t % per_block (the rhs slot of the partial sums).
struct Span {
  int64_t first, end, stride;
};

__device__ __forceinline__ Span block_span(int64_t total, int per_block, int xcd) {
  if (xcd && gridDim.x % 8 == 0) {
    const int64_t c = ((total + 7) / 8 + per_block - 1) / per_block * per_block;
    const int64_t f = (int64_t)(blockIdx.x % 8) * c;
    return Span{f + (int64_t)(blockIdx.x / 8) * per_block, f + c < total ? f + c : total,
                (int64_t)(gridDim.x / 8) * per_block};
  }
  return Span{(int64_t)blockIdx.x * per_block, total, (int64_t)gridDim.x * per_block};
}

// Host view of where every rhs stopped, passed by value to the segment start.
struct SegInfo {
  int d[MAXR];
};

// Prologue of A: the scalars of iteration `it`, identical in every block (the
// partials are summed in block order); block 0 records them and the stop
// flags. Returns nonzero while some rhs is running. `ended` (the alldone flag
// read at kernel entry, so that its load overlaps the partial loads) makes it
// return 0 before recording anything.
template <int NRP>
__device__ __forceinline__ int dir_prologue(int it, const double *part_in, int n_in,
                                            const PcgState &st, double *sh, double *tot,
                                            double *beta, int *live, int ended = 0) {
  sum_partials<2 * NRP>(part_in, n_in, sh, tot);
  if (ended) return 0;
  const int tid = threadIdx.x;
  if (tid < NRP) {
    const double g = tot[tid], rr = tot[NRP + tid];
    // a concurrent write of done[] by block 0 only turns -1 into `it`, and
    // only for a rhs that stops now: every block sees it as not running
    const bool was = st.done[tid] >= 0;
    const bool now = !was && (rr <= st.tol2[tid] || !(rr == rr));
    const bool lv = !was && !now;
    const double den = it == st.segf[tid] ? st.rzold[tid]
                                       : st.scal[((size_t)(it - 1) * NRP + tid) * NSCAL + S_RZ];
    live[tid] = lv;
    beta[tid] = (lv && den != 0.0) ? g / den : 0.0;
    if (blockIdx.x == 0) {
      double *sc = st.scal + ((size_t)it * NRP + tid) * NSCAL;
      sc[S_RZ] = g;
      sc[S_RR] = rr;
      if (now) st.done[tid] = it;
    }
  }
  __syncthreads();
  int any = 0;
#pragma unroll
  for (int r = 0; r < NRP; ++r) any |= live[r];
  if (blockIdx.x == 0 && tid == 0) *st.alldone = !any;
  return any;
}

// A over the node blocks. TL is a multiple of NRP, so the lane that updates
// element e = a*NRP + r (lane e % TL) belongs to rhs tid % NRP, the slot of
// its eta partial. CO: beta and the running flags come from coef_a (prepared
// by k_pcg_scal_dir) instead of the partial sets at part_in.
template <typename T, int ND, int TL, int NRP, bool CO, bool D16>
__device__ __forceinline__ void spmv_dir_body(const tfm_mesh_t &m, const T *__restrict__ val,
                                              const T *__restrict__ u, T *__restrict__ p,
                                              T *__restrict__ s, int it,
                                              const double *part_in, int n_in,
                                              const PcgState &st) {
  static_assert(TL % NRP == 0, "team width must be a multiple of the rhs count");
  __shared__ double sh[BS];
  __shared__ double tot[2 * NRP], beta[NRP];
  __shared__ int live[NRP];
  constexpr int TPB = BS / TL;
  const int lane = threadIdx.x % TL, team = threadIdx.x / TL;
  const Span sp_ = block_span(m.n_nod, TPB, CO ? 0 : st.xcd);
  const int64_t stride = sp_.stride, n_end = sp_.end;
  int64_t n = sp_.first + team;
  // grid-stride mode: row pointers run one node ahead of the product, so
  // each node's loads start from pointers already in registers (the first
  // ones load during the prologue). The CO mode, mostly one node per team
  // on the full grid, loads them in place.
  NodePtr np{0, 0, 0};
  if constexpr (!CO) {
    if (n < n_end) np = node_ptr(m, n);
  }
  if constexpr (CO) {
    const PcgCoef &cf = *st.coef_a;
    if (!cf.any) return;
    if (threadIdx.x < NRP) {
      beta[threadIdx.x] = cf.c[threadIdx.x];
      live[threadIdx.x] = cf.live[threadIdx.x];
    }
    __syncthreads();
  } else {
    const int ended = *st.alldone;
    if (!dir_prologue<NRP>(it, part_in, n_in, st, sh, tot, beta, live, ended)) return;
  }
  // lane l updates the elements e = l, l + TL, ... of the node's ND*NRP
  // (offset e from the node's first entry); they all belong to rhs l % NRP
  constexpr int NE = ND * NRP, KE = (NE + TL - 1) / TL;
  const T b = (T)beta[lane % NRP];
  const bool lv = live[lane % NRP] != 0;
  double dot = 0.0;
  for (; n < n_end; n += stride) {
    NodePtr cur;
    if constexpr (!CO) {
      cur = np;
      if (n + stride < n_end) np = node_ptr(m, n + stride);
    } else {
      cur = node_ptr(m, n);
    }
    const int64_t nb = n * NE;
    T up[KE], pp[KE], sp[KE];  // issued with the SpMV loads, used after it
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int e = lane + k * TL;
      if (lv && e < NE) {
        up[k] = u[nb + e];
        pp[k] = p[nb + e];
        sp[k] = s[nb + e];
      }
    }
    acc_t<T, NRP> acc[ND][NRP];
    T am[ND][4], outv[KE];
    if constexpr (rhs_lanes<TL, NRP>())
      node_product_mr<T, ND, NRP, D16>(m, val, u, n, cur, lane, am);
    else if constexpr (lds_reduce<T, TL, NRP>())
      node_product_lds<T, ND, TL, NRP, D16>(m, val, u, n, cur, lane, team, outv);
    else
      node_product_at<T, ND, TL, NRP, D16>(m, val, u, n, cur, lane, acc);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const int e = lane + k * TL;
      if (lv && e < NE) {
        T w = T(0);
        if constexpr (rhs_lanes<TL, NRP>()) {
          w = pick_entry<T, ND, NRP>(am, e);  // e = lane: rhs lane % NRP
        } else if constexpr (lds_reduce<T, TL, NRP>()) {
          w = outv[k];
        } else {
#pragma unroll
          for (int a = 0; a < ND; ++a)
#pragma unroll
            for (int r = 0; r < NRP; ++r)
              if (a * NRP + r == e) w = (T)acc[a][r];
        }
        const T pv = up[k] + b * pp[k];
        const T sv = w + b * sp[k];
        p[nb + e] = pv;
        s[nb + e] = sv;
        dot += (double)pv * (double)sv;
      }
    }
  }
  double v[1] = {dot};
  write_block_partials<1, NRP>(v, sh, st.part_pq);
}

template <typename T, int ND, int TL, int NRP, bool CO, bool D16>
__global__ void __launch_bounds__(BS) k_pcg_spmv_dir(tfm_mesh_t m, const T *__restrict__ val,
                                                     const T *__restrict__ u, T *__restrict__ p,
                                                     T *__restrict__ s, int it,
                                                     const double *part_in, int n_in,
                                                     PcgState st) {
  spmv_dir_body<T, ND, TL, NRP, CO, D16>(m, val, u, p, s, it, part_in, n_in, st);
}

// The same kernel held to 64 VGPRs (8 waves per SIMD), so that a grid of
// 2048 blocks of 4 waves is resident in one round on 256 CUs.
template <typename T, int ND, int TL, int NRP, bool CO, bool D16>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_pcg_spmv_dir_w8(tfm_mesh_t m, const T *__restrict__ val, const T *__restrict__ u,
                  T *__restrict__ p, T *__restrict__ s, int it,
                  const double *part_in, int n_in, PcgState st) {
  spmv_dir_body<T, ND, TL, NRP, CO, D16>(m, val, u, p, s, it, part_in, n_in, st);
}

// Held to 80 VGPRs (6 waves per SIMD), for kernels that spill at 64.
template <typename T, int ND, int TL, int NRP, bool CO, bool D16>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(6, 6)))
k_pcg_spmv_dir_w6(tfm_mesh_t m, const T *__restrict__ val, const T *__restrict__ u,
                  T *__restrict__ p, T *__restrict__ s, int it,
                  const double *part_in, int n_in, PcgState st) {
  spmv_dir_body<T, ND, TL, NRP, CO, D16>(m, val, u, p, s, it, part_in, n_in, st);
}

// A over a general CSR matrix (one rhs).
template <int TL>
__global__ void __launch_bounds__(BS) k_pcg_csr_dir(int64_t n, const int64_t *rowptr,
                                                    const int32_t *col, const double *val,
                                                    const double *u, double *p, double *s, int it,
                                                    const double *part_in, int n_in,
                                                    PcgState st) {
  if (*st.alldone) return;
  __shared__ double sh[BS];
  __shared__ double tot[2], beta[1];
  __shared__ int live[1];
  if (!dir_prologue<1>(it, part_in, n_in, st, sh, tot, beta, live)) return;
  constexpr int TPB = BS / TL;
  const int lane = threadIdx.x % TL, team = threadIdx.x / TL;
  double dot = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + team; i < n; i += (int64_t)gridDim.x * TPB) {
    const double w = csr_row<TL>(rowptr, col, val, u, i, lane);
    if (lane == 0) {
      const double pv = u[i] + beta[0] * p[i];
      const double sv = w + beta[0] * s[i];
      p[i] = pv;
      s[i] = sv;
      dot += pv * sv;
    }
  }
  double v[1] = {dot};
  write_block_partials<1, 1>(v, sh, st.part_pq);
}

// B: alpha = gamma / eta; x += alpha p; r -= alpha s; u = M r; partials of
// (r.u, r.r) for the next A. pq_in holds n_pq partial sets of eta.
template <typename T, int ND, int NRP, bool CO>
__global__ void __launch_bounds__(BSV) k_pcg_update(int64_t n_nod, int it, T *__restrict__ x,
                                                   T *__restrict__ r, T *__restrict__ u,
                                                   const T *__restrict__ p,
                                                   const T *__restrict__ s,
                                                   const T *__restrict__ minv, PcgState st,
                                                   const double *pq_in, int n_pq) {
  __shared__ double sh[2 * BSV];
  __shared__ double pq[NRP];
  __shared__ double alpha[NRP];
  const int tid = threadIdx.x;
  if constexpr (CO) {  // alpha prepared by k_pcg_scal_alpha
    const PcgCoef &cf = *st.coef_b;
    if (!cf.any) return;
    if (tid < NRP) alpha[tid] = cf.c[tid];
  } else {
    const int ended = *st.alldone;  // tested after the sums, so the loads overlap
    sum_partials<NRP, BSV>(pq_in, n_pq, sh, pq);
    if (ended) return;
    if (tid < NRP) {
      double *sc = st.scal + ((size_t)it * NRP + tid) * NSCAL;
      alpha[tid] = (st.done[tid] < 0 && pq[tid] != 0.0) ? sc[S_RZ] / pq[tid] : 0.0;
      if (blockIdx.x == 0) sc[S_PQ] = pq[tid];
    }
  }
  __syncthreads();
  double acc[2] = {0.0, 0.0};
  const Span sp_ = block_span(n_nod * NRP, BSV, CO ? 0 : st.xcd);
  for (int64_t t = sp_.first + tid; t < sp_.end; t += sp_.stride) {
    const int64_t n = t / NRP;
    const int rr = (int)(t % NRP);
    const T al = (T)alpha[rr];
    const Vec<T, pc_size<ND>()> mi = ldv<T, pc_size<ND>()>(minv + n * pc_size<ND>());
    T rv[ND], zv[ND];
    if constexpr (NRP == 1) {  // the node's ND entries are contiguous
      Vec<T, ND> xv = ldv<T, ND>(x + n * ND), rw = ldv<T, ND>(r + n * ND);
      const Vec<T, ND> pv = ldv<T, ND>(p + n * ND), sv = ldv<T, ND>(s + n * ND);
#pragma unroll
      for (int a = 0; a < ND; ++a) {
        xv.v[a] += al * pv.v[a];
        rw.v[a] -= al * sv.v[a];
        rv[a] = rw.v[a];
      }
      stv<T, ND>(x + n * ND, xv);
      stv<T, ND>(r + n * ND, rw);
    } else {
#pragma unroll
      for (int a = 0; a < ND; ++a) {
        const int64_t o = (n * ND + a) * NRP + rr;
        x[o] += al * p[o];
        rv[a] = r[o] - al * s[o];
        r[o] = rv[a];
      }
    }
#pragma unroll
    for (int a = 0; a < ND; ++a) {
      T z = T(0);
#pragma unroll
      for (int b2 = 0; b2 < ND; ++b2) z += mi.v[pc_index<ND>(a, b2)] * rv[b2];
      zv[a] = z;
      acc[0] += (double)rv[a] * (double)z;
      acc[1] += (double)rv[a] * (double)rv[a];
    }
    if constexpr (NRP == 1) {
      Vec<T, ND> zo;
#pragma unroll
      for (int a = 0; a < ND; ++a) zo.v[a] = zv[a];
      stv<T, ND>(u + n * ND, zo);
    } else {
#pragma unroll
      for (int a = 0; a < ND; ++a) u[(n * ND + a) * NRP + rr] = zv[a];
    }
  }
  write_block_partials<2, NRP, BSV>(acc, sh, st.part_upd);
}

// Scalar kernels (one block) for wide grids. k_pcg_scal_dir prepares A of
// iteration `it`: the same scalars, stop test and records as the prologue of
// A (dir_prologue, run here by a single block), published in coef_a.
template <int NRP>
__global__ void __launch_bounds__(BS) k_pcg_scal_dir(int it, const double *part_in,
                                                     int n_in, PcgState st) {
  __shared__ double sh[BS];
  __shared__ double tot[2 * NRP], beta[NRP];
  __shared__ int live[NRP];
  const int any = dir_prologue<NRP>(it, part_in, n_in, st, sh, tot, beta, live);
  const int tid = threadIdx.x;
  if (tid < NRP) {
    st.coef_a->c[tid] = beta[tid];
    st.coef_a->live[tid] = live[tid];
  }
  if (tid == 0) st.coef_a->any = any;
}

// k_pcg_scal_alpha prepares B of iteration `it`: eta from the nblk partial
// sets of A (summed in block order by 1024 threads, four loads in flight per
// thread), alpha = gamma / eta, published in coef_b.
template <int NRP>
__global__ void __launch_bounds__(1024) k_pcg_scal_alpha(int it, const double *part, int nblk,
                                                         PcgState st) {
  __shared__ double sh[1024];
  constexpr int CH = 1024 / NRP;
  const int tid = threadIdx.x;
  const int slot = tid % NRP, c = tid / NRP;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int b = c;
  for (; b + 3 * CH < nblk; b += 4 * CH) {
    a0 += part[(size_t)b * NRP + slot];
    a1 += part[(size_t)(b + CH) * NRP + slot];
    a2 += part[(size_t)(b + 2 * CH) * NRP + slot];
    a3 += part[(size_t)(b + 3 * CH) * NRP + slot];
  }
  for (; b < nblk; b += CH) a0 += part[(size_t)b * NRP + slot];
  sh[tid] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  for (int w = CH / 2; w >= 1; w >>= 1) {
    if (tid < w * NRP) sh[tid] += sh[tid + w * NRP];
    __syncthreads();
  }
  if (tid < NRP) {
    double *sc = st.scal + ((size_t)it * NRP + tid) * NSCAL;
    const double pq = sh[tid];
    const bool lv = st.done[tid] < 0;
    sc[S_PQ] = pq;
    st.coef_b->c[tid] = (lv && pq != 0.0) ? sc[S_RZ] / pq : 0.0;
    st.coef_b->live[tid] = lv;
  }
  if (tid == 0) st.coef_b->any = !*st.alldone;
}

// Segment start, part 1: u = M r for the (replaced) residuals; partials of
// (r.u, r.r) for the next A and of r.p for the continuation.
template <typename T, int ND, int NRP>
__global__ void __launch_bounds__(BS) k_pcg_seg_dots(int64_t n_nod, const T *r, T *u, const T *p,
                                                     const T *minv, PcgState st) {
  __shared__ double sh[2 * BS];
  const int tid = threadIdx.x;
  double acc[2] = {0.0, 0.0};
  double rp = 0.0;
  const int64_t total = n_nod * NRP;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int64_t n = t / NRP;
    const int rr = (int)(t % NRP);
    T rv[ND], zv[ND];
    apply_pc<T, ND, NRP>(minv, r, n, rr, zv, rv);
#pragma unroll
    for (int a = 0; a < ND; ++a) {
      const int64_t o = (n * ND + a) * NRP + rr;
      u[o] = zv[a];
      acc[0] += (double)rv[a] * (double)zv[a];
      acc[1] += (double)rv[a] * (double)rv[a];
      rp += (double)rv[a] * (double)p[o];
    }
  }
  write_block_partials<2, NRP>(acc, sh, st.part_upd);
  double v[1] = {rp};
  write_block_partials<1, NRP>(v, sh, st.part_aux);
}

// Segment start, part 2. A rhs whose bit is set in `mask` starts its next
// segment at iteration `it`. If its bit is also set in `cont`, it keeps its
// search direction: p' = p - (r.p / r.u) u is orthogonal to the new residual,
// and the first beta of the segment divides by gamma of the iteration at
// which it stopped (seg.d - 1); the caller then forms s = K p'. Otherwise it
// restarts from p = s = 0 (beta = 0). A rhs outside `mask` has either
// finished or is still inside its own segment, and is left as is; `running`
// holds every rhs that continues to iterate (mask included).
template <typename T, int ND, int NRP>
__global__ void __launch_bounds__(BS) k_pcg_seg_apply(int64_t n_nod, int it, int mask, int cont,
                                                      int running, SegInfo seg, const T *u, T *p,
                                                      T *s, PcgState st, int nblk, int nblk_aux) {
  __shared__ double sh[BS];
  __shared__ double tot[2 * NRP], rp[NRP], cz[NRP];
  __shared__ int mode[NRP];
  sum_partials<2 * NRP>(st.part_upd, nblk, sh, tot);
  sum_partials<NRP>(st.part_aux, nblk_aux, sh, rp);
  const int tid = threadIdx.x;
  if (tid < NRP) {
    const bool run = (mask >> tid) & 1;
    const int d = seg.d[tid];
    bool keep = run && ((cont >> tid) & 1) && d >= 1;
    const double g = tot[tid];
    const double c = keep && g != 0.0 ? rp[tid] / g : 0.0;
    const double den = keep ? st.scal[((size_t)(d - 1) * NRP + tid) * NSCAL + S_RZ] : 0.0;
    if (!(c == c) || !(den == den) || den == 0.0 || g == 0.0) keep = false;
    mode[tid] = keep ? 1 : (run ? 2 : 0);
    cz[tid] = keep ? c : 0.0;
    if (blockIdx.x == 0) {
      if (run) {
        st.rzold[tid] = keep ? den : 0.0;
        st.done[tid] = -1;
        st.segf[tid] = it;
      } else if (!((running >> tid) & 1)) {  // finished (or an unused slot): stopped
        st.done[tid] = d >= 0 ? d : it;
        st.segf[tid] = it;
      }
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid == 0) *st.alldone = running == 0;
  const int64_t total = n_nod * ND * NRP;
  for (int64_t t = (int64_t)blockIdx.x * BS + tid; t < total; t += (int64_t)gridDim.x * BS) {
    const int rr = (int)(t % NRP);
    const int md = mode[rr];
    if (md == 1) {
      p[t] -= (T)cz[rr] * u[t];
    } else if (md == 2) {
      p[t] = T(0);
      s[t] = T(0);
    }
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

// Refinement residual res = b - K x with the row sums of K x carried in
// double-double (an exact product by FMA, an exact sum by two-sum, per
// term): the residual then has no FP64 rounding floor of the size
// eps * |K| |x|, which near a tight tolerance would stall the refinement.
// One team of TL lanes per node; partials of (res.res, b.b) per block.
__device__ __forceinline__ void dd_add(double &hi, double &lo, double a, double b2) {
  // (hi, lo) += (a, b2)
  const double s = hi + a;
  const double v = s - hi;
  const double e = (hi - (s - v)) + (a - v) + lo + b2;
  hi = s + e;
  lo = e - (hi - s);
}

template <int ND, int TL>
__global__ void __launch_bounds__(BS) k_residual_dd(tfm_mesh_t m, const double *__restrict__ val,
                                                    const double *__restrict__ b,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ res, double *part) {
  __shared__ double sh[2 * BS];
  constexpr int TPB = BS / TL;
  const int lane = threadIdx.x % TL, team = threadIdx.x / TL;
  double acc2[2] = {0.0, 0.0};
  for (int64_t n = (int64_t)blockIdx.x * TPB + team; n < m.n_nod; n += (int64_t)gridDim.x * TPB) {
    double hi[ND], lo[ND];
#pragma unroll
    for (int a = 0; a < ND; ++a) hi[a] = lo[a] = 0.0;
    const int a0 = m.adj_ptr[n];
    const int deg = m.adj_ptr[n + 1] - a0;
    const double *vn = val + m.val_ptr[n];
    auto block = [&](const Vec<double, ND> (&rows)[ND], const Vec<double, ND> &xv) {
#pragma unroll
      for (int a = 0; a < ND; ++a)
#pragma unroll
        for (int c = 0; c < ND; ++c) {
          const double ph = rows[a].v[c] * xv.v[c];
          dd_add(hi[a], lo[a], ph, fma(rows[a].v[c], xv.v[c], -ph));
        }
    };
    if (deg > 0) {
      const int L = ND * deg;
      for (int p = lane; p < deg; p += TL) {
        Vec<double, ND> rows[ND];
#pragma unroll
        for (int a = 0; a < ND; ++a) rows[a] = ldv<double, ND>(vn + a * L + p * ND);
        block(rows, ldv<double, ND>(x + (int64_t)m.adj[a0 + p] * ND));
      }
    } else if (lane < ND) {
#pragma unroll
      for (int a = 0; a < ND; ++a)
        if (a == lane) {
          const double ph = vn[a] * x[n * ND + a];
          dd_add(hi[a], lo[a], ph, fma(vn[a], x[n * ND + a], -ph));
        }
    }
#pragma unroll
    for (int off = 1; off < TL; off <<= 1)
#pragma unroll
      for (int a = 0; a < ND; ++a) {
        const double oh = __shfl_xor(hi[a], off, TL), ol = __shfl_xor(lo[a], off, TL);
        dd_add(hi[a], lo[a], oh, ol);
      }
#pragma unroll
    for (int a = 0; a < ND; ++a)
      if (a == lane) {
        const int64_t i = n * ND + a;
        double rh = b[i], rl = 0.0;
        dd_add(rh, rl, -hi[a], -lo[a]);
        const double r = rh + rl;
        res[i] = r;
        acc2[0] += r * r;
        acc2[1] += b[i] * b[i];
      }
  }
  write_block_partials<2, 1>(acc2, sh, part);
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

__global__ void __launch_bounds__(1024) k_sum2(const double *part, int nblk, double *out) {
  // nblk partial pairs (block order), 512 pairs of accumulators
  __shared__ double sh[1024];
  const int tid = threadIdx.x, slot = tid % 2, c = tid / 2;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int b = c;
  for (; b + 3 * 512 < nblk; b += 4 * 512) {
    a0 += part[(size_t)b * 2 + slot];
    a1 += part[(size_t)(b + 512) * 2 + slot];
    a2 += part[(size_t)(b + 1024) * 2 + slot];
    a3 += part[(size_t)(b + 1536) * 2 + slot];
  }
  for (; b < nblk; b += 512) a0 += part[(size_t)b * 2 + slot];
  sh[tid] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  for (int w = 256; w >= 1; w >>= 1) {
    if (tid < 2 * w) sh[tid] += sh[tid + 2 * w];
    __syncthreads();
  }
  if (tid < 2) out[tid] = sh[tid];
}

// ------------------------------------------------------- launch helpers
static unsigned grid_for(int64_t work, int per_block, int cap) {
  const int64_t g = (work + per_block - 1) / per_block;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

// g rounded up to a multiple of the number of CUs of the current device.
static unsigned round_to_cus(unsigned g) {
  static const int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 1;
    return std::max(1, n);
  }();
  return (unsigned)(((int64_t)g + cus - 1) / cus * cus);
}

// Whether to launch the 64-VGPR variant of a kernel: when it needs no scratch
// (TFM_SPMV_W8=0/1 in the environment forces the choice).
static bool pick_w8(const void *kernel_w8) {
  static const int forced = env_int("TFM_SPMV_W8", -1);
  if (forced >= 0) return forced != 0;
  hipFuncAttributes attr{};
  if (hipFuncGetAttributes(&attr, kernel_w8) != hipSuccess) return false;
  return attr.localSizeBytes == 0;
}

// Blocks of BS threads of `kernel` resident at once on the current device
// (occupancy per CU times CUs): the grid of a grid-stride kernel that should
// run in a single round.
static int resident_grid(const void *kernel) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, BS, 0) != hipSuccess)
    return RBA;
  return std::max(1, per_cu) * std::max(1, cus);
}

// Launch shape of the node-block SpMV kernels for value type T: the team
// width, the grid cap, and whether one-block scalar kernels prepare beta and
// alpha (then A and B read one coefficient record per block instead of
// summing partial sets, and A's grid is not limited by its consumers).
// Measured on MI355X with the Hexa1 64^3 matrix: FP32 streams fastest with
// 16-lane teams (two hoisted blocks per lane) on 2048 blocks, whose partials
// every block of B can sum directly; FP64 needs the full grid (one node per
// team, 5.8 TB/s vs 4.7 TB/s at 2048 blocks) and the scalar kernels.
// Narrower rows (2-D, heat) use 16-lane teams. TFM_SPMV_GRID / TFM_SPMV_SCAL
// in the environment override the grid cap and the mode (tuning aid).
struct SpmvShape {
  int tl;
  int cap;
  bool scal;  // one-block scalar kernels prepare the coefficients of A and B
};

template <typename T>
static SpmvShape spmv_shape(const tfm_mesh_t &m, int nrp) {
  const int64_t n_adj = m.n_nod > 0 ? (int64_t)m.nnz / std::max(1, m.ndof * m.ndof) : 0;
  const double deg = (double)n_adj / std::max<int64_t>(1, m.n_nod);
  SpmvShape sh;
  const bool fp64 = sizeof(T) == 8;
  sh.tl = (fp64 && deg > 16.0) || deg > 32.0 ? 32 : deg > 16.0 ? 16 : 8;
  // FP64 scalar fields with short rows: 4-lane teams (two hoisted blocks per
  // lane) ran the thermal Tetra1 80^3 step 11% faster than 8-lane teams;
  // in FP32 they were 11% slower
  if (fp64 && m.ndof == 1 && deg <= 16.0) sh.tl = 4;
  static const int tl_env = env_int("TFM_SPMV_TL", 0);
  if (tl_env == 4 || tl_env == 8 || tl_env == 16 || tl_env == 32) sh.tl = tl_env;
  sh.scal = fp64;
  // FP64 with one or two rhs: the full grid (one node per team); with more,
  // a capped grid-stride grid, so that the one-block scalar kernel sums few
  // partial sets (35k blocks x 16 rhs of Hexa2 32^3 took it 84 us). FP32
  // with 16 rhs on Hexa2 32^3: 1024 blocks ran the step in 1,015 ms against
  // 1,065 ms with 512
  sh.cap = fp64 ? (nrp >= 4 ? RBA / 2 : (1 << 30)) : (nrp <= 2 ? RBA : RBA / 2);
  static const int g_env = env_int("TFM_SPMV_GRID", 0);
  static const int f_env = env_int("TFM_SPMV_SCAL", -1);
  if (g_env > 0) sh.cap = g_env;
  if (f_env >= 0) sh.scal = f_env != 0;
  if (!sh.scal) sh.cap = std::min(sh.cap, RBA);  // every block of B sums A's partial sets
  return sh;
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

// Team width TL: 4 (FP64 scalar fields with short rows), 8 (rows of at most
// 16 neighbour blocks on average), 16 or 32 lanes per node, or a whole wave
// with the rhs
// across lanes (node_product_mr) for FP64 with four or more right-hand sides,
// where a lane cannot hold ND*NRP doubles without losing occupancy. (In FP32
// the lane-per-block product measured faster for 16 rhs: 274 against 307 us
// on Hexa2 32^3.) T and NRP must be in scope.
#define TFM_WIDE_TEAM (sizeof(T) == 8 && NRP >= 4)
#define TFM_TEAM_SWITCH(tl, ...)                                                               \
  switch (TFM_WIDE_TEAM ? 64 : (tl)) {                                                         \
    case 4: {                                                                                  \
      constexpr int TL = TFM_WIDE_TEAM ? 64 : (NRP > 8 ? 16 : NRP > 4 ? 8 : 4);                \
      __VA_ARGS__;                                                                             \
      break;                                                                                   \
    }                                                                                          \
    case 8: { constexpr int TL = TFM_WIDE_TEAM ? 64 : (NRP > 8 ? 16 : 8); __VA_ARGS__; break; } \
    case 16: { constexpr int TL = TFM_WIDE_TEAM ? 64 : 16; __VA_ARGS__; break; }               \
    default: { constexpr int TL = TFM_WIDE_TEAM ? 64 : 32; __VA_ARGS__; break; }               \
  }

template <typename T, int ND, int NRP>
static int spmv_launch(hipStream_t s, const tfm_mesh_t &m, const T *val, const T *x, T *y) {
  const SpmvShape sh = spmv_shape<T>(m, NRP);
  // the plain FP32 product with 16 rhs ran faster on 512 blocks (178 us)
  // than on the fused product's 1024 (204 us), on Hexa2 32^3
  const int cap = sizeof(T) == 4 && NRP > 8 ? std::min(sh.cap, RBA / 4) : sh.cap;
  TFM_TEAM_SWITCH(sh.tl, {
    const unsigned g = grid_for(m.n_nod, BS / TL, cap);
    if (m.adj16)
      hipLaunchKernelGGL((k_spmv<T, ND, TL, NRP, true>), dim3(g), dim3(BS), 0, s, m, val, x, y);
    else
      hipLaunchKernelGGL((k_spmv<T, ND, TL, NRP, false>), dim3(g), dim3(BS), 0, s, m, val, x, y);
  });
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

template <typename T, int ND, int TL, int NRP, bool D16>
static int spmv_dir_launch_d(hipStream_t s, const SpmvShape &sh, const tfm_mesh_t &m,
                             const T *val, const T *u, T *p, T *sv, int it,
                             const double *part_in, int n_in, const PcgState &st, unsigned *g) {
  if (sh.scal) {
    *g = grid_for(m.n_nod, BS / TL, sh.cap);
    TFM_REQUIRE((int64_t)*g <= st.pq_cap, "SpMV grid exceeds the partial-sum buffer");
    hipLaunchKernelGGL((k_pcg_spmv_dir<T, ND, TL, NRP, true, D16>), dim3(*g), dim3(BS), 0, s, m,
                       val, u, p, sv, it, part_in, n_in, st);
  } else {
    // grid-stride over the nodes with every block resident in one round;
    // the 8-wave variant unless holding it to 64 VGPRs spills registers,
    // then the 6-wave one unless 80 VGPRs spill
    static const auto kern =
        pick_w8((const void *)k_pcg_spmv_dir_w8<T, ND, TL, NRP, false, D16>)
            ? k_pcg_spmv_dir_w8<T, ND, TL, NRP, false, D16>
        : pick_w8((const void *)k_pcg_spmv_dir_w6<T, ND, TL, NRP, false, D16>)
            ? k_pcg_spmv_dir_w6<T, ND, TL, NRP, false, D16>
            : k_pcg_spmv_dir<T, ND, TL, NRP, false, D16>;
    static const int resident = resident_grid((const void *)kern);
    *g = grid_for(m.n_nod, BS / TL, std::min(sh.cap, resident));
    TFM_REQUIRE((int64_t)*g <= st.pq_cap, "SpMV grid exceeds the partial-sum buffer");
    hipLaunchKernelGGL(kern, dim3(*g), dim3(BS), 0, s, m, val, u, p, sv, it, part_in,
                       n_in, st);
  }
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

template <typename T, int ND, int NRP>
static int spmv_dir_launch(hipStream_t s, const tfm_mesh_t &m, const T *val, const T *u, T *p,
                           T *sv, int it, const double *part_in, int n_in,
                           const PcgState &st, unsigned *grid_out) {
  const SpmvShape sh = spmv_shape<T>(m, NRP);
  int rc = TFM_OK;
  TFM_TEAM_SWITCH(sh.tl, {
    rc = m.adj16 ? spmv_dir_launch_d<T, ND, TL, NRP, true>(s, sh, m, val, u, p, sv, it, part_in,
                                                           n_in, st, grid_out)
                 : spmv_dir_launch_d<T, ND, TL, NRP, false>(s, sh, m, val, u, p, sv, it, part_in,
                                                            n_in, st, grid_out);
  });
  return rc;
}

int launch_spmv(hipStream_t s, const tfm_mesh_t &m, int prec, const void *vals, const void *x,
                void *y, int nrhs) {
  if (m.n_nod == 0) return TFM_OK;
  int rc = TFM_OK;
  TFM_ND_SWITCH(m.ndof, TFM_NRP_SWITCH(nrhs, {
    if (prec == 0)
      rc = spmv_launch<double, ND, NRP>(s, m, (const double *)vals, (const double *)x, (double *)y);
    else
      rc = spmv_launch<float, ND, NRP>(s, m, (const float *)vals, (const float *)x, (float *)y);
  }));
  return rc;
}

// ====================================================================
// Two-level preconditioner: an additive coarse correction on node
// aggregates, M^{-1} r = B^{-1} r + P (P^T K P)^{-1} P^T r. The columns of P
// are the rigid-body modes of each aggregate's nodes about its centroid,
// zero on constrained DOFs.
// ====================================================================
template <int ND> constexpr int coarse_modes() { return ND == 3 ? 6 : ND == 2 ? 3 : 1; }

// Mode values of node i in aggregate I: M[a][k] for DOF a and mode k.
template <int ND>
__device__ __forceinline__ void node_modes(const tfm_mesh_t &m, const double *cen,
                                           const uint8_t *con, int64_t i, int I,
                                           double (&M)[ND][coarse_modes<ND>()]) {
  constexpr int NM = coarse_modes<ND>();
  if constexpr (ND == 1) {
    M[0][0] = con[i] ? 0.0 : 1.0;
  } else {
    double d[ND];
    for (int a = 0; a < ND; ++a) d[a] = m.nodes[i * ND + a] - cen[(int64_t)I * ND + a];
    for (int a = 0; a < ND; ++a)
      for (int k = 0; k < NM; ++k) M[a][k] = (k == a) ? 1.0 : 0.0;
    if constexpr (ND == 3) {  // rotations e_k x d about the centroid
      M[1][3] = -d[2]; M[2][3] = d[1];
      M[0][4] = d[2];  M[2][4] = -d[0];
      M[0][5] = -d[1]; M[1][5] = d[0];
    } else {
      M[0][2] = -d[1]; M[1][2] = d[0];
    }
    for (int a = 0; a < ND; ++a)
      if (con[i * ND + a])
        for (int k = 0; k < NM; ++k) M[a][k] = 0.0;
  }
}

struct CoarseView {
  int n_agg, n_c;
  const int32_t *agg, *agg_ptr, *agg_nodes, *slot_agg, *pair_off, *pair_node, *pair_pos;
  const double *cen;
  double *rc, *ec;  // [n_c * NRP]
};

// Ac[(I nm + k), (J nm + l)] for the aggregate pair (I, slot): one wave per
// (I, s); lane t sums M_i^T K_ij M_j over the node blocks (i, p) = t, t + 64,
// ... of the pair's list, then the wave adds the lane sums by a butterfly
// (a fixed order: the result is reproducible).
template <int ND>
__global__ void __launch_bounds__(64) k_coarse_operator(tfm_mesh_t m, CoarseView cv,
                                                        const double *val, const uint8_t *con,
                                                        double *Ac) {
  constexpr int NM = coarse_modes<ND>();
  const int64_t b = blockIdx.x;
  const int I = (int)(b / 27);
  const int J = cv.slot_agg[b];
  if (J < 0) return;
  const int t = threadIdx.x;
  double acc[NM][NM];
  for (int k = 0; k < NM; ++k)
    for (int l = 0; l < NM; ++l) acc[k][l] = 0.0;
  for (int q = cv.pair_off[b] + t; q < cv.pair_off[b + 1]; q += 64) {
    const int64_t i = cv.pair_node[q];
    const int p = cv.pair_pos[q];
    const int a0 = m.adj_ptr[i];
    const int deg = m.adj_ptr[i + 1] - a0;
    const int64_t j = m.adj[a0 + p];
    double Mi[ND][NM], Mj[ND][NM];
    node_modes<ND>(m, cv.cen, con, i, I, Mi);
    node_modes<ND>(m, cv.cen, con, j, J, Mj);
    const double *kv = val + m.val_ptr[i] + p * ND;
    double KM[ND][NM];  // K_ij M_j
    for (int a = 0; a < ND; ++a)
      for (int l = 0; l < NM; ++l) KM[a][l] = 0.0;
    for (int a = 0; a < ND; ++a)
      for (int c = 0; c < ND; ++c) {
        const double kac = kv[(int64_t)a * ND * deg + c];
        for (int l = 0; l < NM; ++l) KM[a][l] += kac * Mj[c][l];
      }
    for (int a = 0; a < ND; ++a)
      for (int k = 0; k < NM; ++k)
        for (int l = 0; l < NM; ++l) acc[k][l] += Mi[a][k] * KM[a][l];
  }
  for (int off = 32; off >= 1; off >>= 1)
    for (int k = 0; k < NM; ++k)
      for (int l = 0; l < NM; ++l) acc[k][l] += __shfl_xor(acc[k][l], off, 64);
  if (t < NM * NM) {
    const int k = t / NM, l = t % NM;
    double v = 0.0;
    for (int kk = 0; kk < NM; ++kk)
      for (int ll = 0; ll < NM; ++ll)
        if (kk == k && ll == l) v = acc[kk][ll];
    Ac[((int64_t)I * NM + k) * cv.n_c + (int64_t)J * NM + l] = v;
  }
}

// Sums of v over the lanes of each wave that share lane % NRP (a butterfly
// in a fixed order), then over the block's waves in wave order, for the
// values of one block of NT threads; the result is valid in threads
// 0 .. NRP - 1 (rhs = thread). Every thread of the block must call it.
template <int NRP, int NT = BS>
__device__ __forceinline__ double block_sum_by_rhs(double v, double *sh) {
  for (int off = 32; off >= NRP; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  if (lane < NRP) sh[wave * NRP + lane] = v;
  __syncthreads();
  double tsum = 0.0;
  if (threadIdx.x < NRP)
    for (int w = 0; w < NT / 64; ++w) tsum += sh[w * NRP + threadIdx.x];
  __syncthreads();
  return tsum;
}

// r_c = P^T r: one block per aggregate; thread t takes rhs t % NRP of every
// (BS / NRP)-th node of the aggregate, and the block adds the thread sums by
// waves (block_sum_by_rhs). (Blocks of 1024 threads, one node per thread,
// measured slower: 11.2 against 8.9 us on the 64^3 cube.)
template <typename T, int ND, int NRP>
__global__ void __launch_bounds__(BS) k_coarse_restrict(tfm_mesh_t m, CoarseView cv,
                                                        const uint8_t *con, const T *r,
                                                        const int *alldone) {
  constexpr int NM = coarse_modes<ND>(), CH = BS / NRP;
  __shared__ double sh[BS / 64 * NRP];
  if (*alldone) return;
  const int I = blockIdx.x, t = threadIdx.x, rr = t % NRP, c = t / NRP;
  double acc[NM];
  for (int k = 0; k < NM; ++k) acc[k] = 0.0;
  for (int q = cv.agg_ptr[I] + c; q < cv.agg_ptr[I + 1]; q += CH) {
    const int64_t i = cv.agg_nodes[q];
    double M[ND][NM];
    node_modes<ND>(m, cv.cen, con, i, I, M);
    for (int a = 0; a < ND; ++a) {
      const double rv = (double)r[(i * ND + a) * NRP + rr];
      for (int k = 0; k < NM; ++k) acc[k] += M[a][k] * rv;
    }
  }
  for (int k = 0; k < NM; ++k) {
    const double v = block_sum_by_rhs<NRP>(acc[k], sh);
    if (t < NRP) cv.rc[((int64_t)I * NM + k) * NRP + t] = v;
  }
}

// e_c = Ainv r_c, one wave per row (four rows per block), columns split
// over the lanes with four loads in flight, and the lane sums added by a
// butterfly.
template <typename TI, int NRP>
__global__ void __launch_bounds__(BS) k_coarse_solve(CoarseView cv, const TI *inv,
                                                     const int *alldone) {
  if (*alldone) return;
  const int lane = threadIdx.x % 64;
  const int row = blockIdx.x * (BS / 64) + threadIdx.x / 64;
  if (row >= cv.n_c) return;
  double acc[NRP];
  for (int r = 0; r < NRP; ++r) acc[r] = 0.0;
  const TI *ar = inv + (int64_t)row * cv.n_c;
  int col = lane;
  for (; col + 192 < cv.n_c; col += 256) {
    const double a0 = (double)ar[col], a1 = (double)ar[col + 64];
    const double a2 = (double)ar[col + 128], a3 = (double)ar[col + 192];
    for (int r = 0; r < NRP; ++r)
      acc[r] += a0 * cv.rc[(int64_t)col * NRP + r] + a1 * cv.rc[(int64_t)(col + 64) * NRP + r] +
                a2 * cv.rc[(int64_t)(col + 128) * NRP + r] +
                a3 * cv.rc[(int64_t)(col + 192) * NRP + r];
  }
  for (; col < cv.n_c; col += 64) {
    const double a = (double)ar[col];
    for (int r = 0; r < NRP; ++r) acc[r] += a * cv.rc[(int64_t)col * NRP + r];
  }
  for (int off = 32; off >= 1; off >>= 1)
    for (int r = 0; r < NRP; ++r) acc[r] += __shfl_xor(acc[r], off, 64);
  if (lane < NRP) {
    double e = 0.0;
    for (int r = 0; r < NRP; ++r)
      if (r == lane) e = acc[r];
    cv.ec[(int64_t)row * NRP + lane] = e;
  }
}

// u += P e_c, one thread per (node, rhs). Block 0 also writes r_c . e_c
// per rhs as one partial set in the (gamma, rr) layout of the PCG's update
// partials (rr slot 0), summed in a fixed order (block_sum_by_rhs).
template <typename T, int ND, int NRP>
__global__ void __launch_bounds__(256) k_coarse_prolong(tfm_mesh_t m, CoarseView cv,
                                                        const uint8_t *con, T *u, double *part,
                                                        const int *alldone) {
  constexpr int NM = coarse_modes<ND>();
  static_assert(BS == 256, "k_coarse_prolong runs blocks of BS = 256 threads");
  __shared__ double sh[BS / 64 * NRP];
  if (*alldone) return;
  if (blockIdx.x == 0) {
    const int rr = threadIdx.x % NRP, c = threadIdx.x / NRP;
    double d = 0.0;
    for (int row = c; row < cv.n_c; row += BS / NRP)
      d += cv.rc[(int64_t)row * NRP + rr] * cv.ec[(int64_t)row * NRP + rr];
    const double t2 = block_sum_by_rhs<NRP>(d, sh);
    if (threadIdx.x < NRP) {
      part[threadIdx.x] = t2;
      part[NRP + threadIdx.x] = 0.0;
    }
  }
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= m.n_nod * NRP) return;
  const int64_t i = t / NRP;
  const int rr = (int)(t % NRP);
  const int I = cv.agg[i];
  double M[ND][NM];
  node_modes<ND>(m, cv.cen, con, i, I, M);
  double e[NM];
  for (int k = 0; k < NM; ++k) e[k] = cv.ec[((int64_t)I * NM + k) * NRP + rr];
  for (int a = 0; a < ND; ++a) {
    double v = 0.0;
    for (int k = 0; k < NM; ++k) v += M[a][k] * e[k];
    T *ua = u + (i * ND + a) * NRP + rr;
    *ua = (T)((double)*ua + v);
  }
}

// ====================================================================
// Dense SPD inverse by block Gauss-Jordan (the coarse inverse of the
// two-level preconditioner). With the pivot block K = [k0, k0 + nb) and
// D = A_KK^{-1}, one step maps
//   A_KK -> D,   A_iK -> -A_iK D,   A_Kj -> D A_Kj,
//   A_ij -> A_ij - A_iK D A_Kj                      (i, j outside K),
// and after a step over every block the matrix holds A^{-1}. Each pivot
// block is then a Schur complement of an SPD matrix, so no pivoting is
// needed; a pivot that is not positive sets the status. A step is three
// launches: the pivot block's inverse (one block, in LDS), the panel
// R = [-A_iK D; D] and the rank-nb update into the other buffer.
// ====================================================================
constexpr int GJ_NB = 64;  // pivot block
constexpr int GJ_T = 64;   // update tile, rows and columns
constexpr int GJ_QC = 32;  // rows of the pivot panel staged at once in the update
constexpr int GJ_PR = 16;  // rows per block of the panel kernel

// D [GJ_NB x GJ_NB] = A_KK^{-1} by scalar Gauss-Jordan, the block held in
// registers: thread t owns row t / 4, columns 16 (t % 4) .. 16 (t % 4) + 15.
// The pivot's row and column go through LDS, double-buffered, so a step has
// one barrier. A block narrower than GJ_NB is padded with the identity,
// whose inverse it keeps.
__global__ void __launch_bounds__(256) k_gj_pivot(const double *a, int64_t n, int k0, int nb,
                                                  double *D, int32_t *status) {
  __shared__ double colb[2][GJ_NB], rowb[2][GJ_NB];
  const int t = threadIdx.x;
  const int i = t / 4, j0 = (t % 4) * 16;
  double v[16];
  for (int jj = 0; jj < 16; ++jj) {
    const int j = j0 + jj;
    v[jj] = (i < nb && j < nb) ? a[(int64_t)(k0 + i) * n + k0 + j] : (i == j ? 1.0 : 0.0);
  }
  int bad = -1;
  for (int pb = 0; pb < GJ_NB; pb += 16) {
#pragma unroll
    for (int pp = 0; pp < 16; ++pp) {
      const int p = pb + pp, buf = pp & 1;
      if (j0 == pb) colb[buf][i] = v[pp];
      if (i == p)
        for (int jj = 0; jj < 16; ++jj) rowb[buf][j0 + jj] = v[jj];
      __syncthreads();
      double piv = rowb[buf][p];
      if (!(piv > 0.0)) {
        if (bad < 0) bad = p;
        piv = 1.0;
      }
      const double inv = 1.0 / piv;
      const double ci = colb[buf][i] * inv;
      for (int jj = 0; jj < 16; ++jj) {
        const int j = j0 + jj;
        const double rj = rowb[buf][j];
        double w;
        if (i == p)
          w = j == p ? inv : inv * rj;
        else if (j == p)
          w = -ci;
        else
          w = v[jj] - ci * rj;
        v[jj] = w;
      }
    }
  }
  for (int jj = 0; jj < 16; ++jj) D[i * GJ_NB + j0 + jj] = v[jj];
  if (t == 0 && bad >= 0 && *status == 0) *status = k0 + bad + 1;
}

// Rt [nb x n] = R^T with R[i][q] = D[i - k0][q] for i in K, else
// -sum_p A[i][k0 + p] D[p][q]; GJ_PR rows per block, four columns q per
// thread.
__global__ void __launch_bounds__(256) k_gj_panel(const double *a, int64_t n, int k0, int nb,
                                                  const double *D, double *Rt) {
  __shared__ double Ds[GJ_NB][GJ_NB];
  __shared__ double As[GJ_PR][GJ_NB + 1];  // As[r][p] = A[i0 + r][k0 + p]
  const int t = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * GJ_PR;
  for (int e = t; e < GJ_NB * GJ_NB; e += 256) Ds[e / GJ_NB][e % GJ_NB] = D[e];
  for (int e = t; e < GJ_PR * GJ_NB; e += 256) {
    const int r = e / GJ_NB, q = e % GJ_NB;
    const int64_t i = i0 + r;
    As[r][q] = (i < n && q < nb) ? a[i * n + k0 + q] : 0.0;
  }
  __syncthreads();
  const int r = t % GJ_PR, qg = (t / GJ_PR) * 4;
  const int64_t i = i0 + r;
  if (i >= n) return;
  const bool inK = i >= k0 && i < k0 + nb;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (!inK)
    for (int p = 0; p < GJ_NB; ++p) {
      const double x = As[r][p];
      for (int k = 0; k < 4; ++k) acc[k] += x * Ds[p][qg + k];
    }
  for (int k = 0; k < 4; ++k) {
    const int q = qg + k;
    if (q >= nb) break;
    Rt[(int64_t)q * n + i] = inK ? Ds[i - k0][q] : -acc[k];
  }
}

// dst = the step's image of src, one 64 x 64 tile per block:
//   j in K:      dst[i][j] = R[i][j - k0]
//   j outside K: dst[i][j] = (i in K ? 0 : src[i][j]) + sum_q R[i][q] src[k0 + q][j].
// Thread (ty, tx) owns rows 4 ty .. 4 ty + 3 and columns 4 tx .. 4 tx + 3;
// its accumulators start from its src values, loaded before the product so
// that their latency overlaps the staging.
__global__ void __launch_bounds__(256) k_gj_update(const double *src, double *dst, int64_t n,
                                                   int k0, int nb, const double *Rt) {
  __shared__ double Rs[GJ_QC][GJ_T];  // Rs[q][r] = R[i0 + r][q0 + q], zero beyond nb
  __shared__ double Bs[GJ_QC][GJ_T];  // Bs[q][c] = src[k0 + q0 + q][j0 + c]
  const int t = threadIdx.x, tx = t % 16, ty = t / 16;
  const int64_t i0 = (int64_t)blockIdx.y * GJ_T, j0 = (int64_t)blockIdx.x * GJ_T;
  double acc[4][4];  // starts from the src values
  for (int x = 0; x < 4; ++x) {
    const int64_t i = i0 + ty * 4 + x;
    const bool iK = i >= k0 && i < k0 + nb;
    for (int y = 0; y < 4; ++y) {
      const int64_t j = j0 + tx * 4 + y;
      const bool jK = j >= k0 && j < k0 + nb;
      acc[x][y] = (i < n && j < n && !iK && !jK) ? src[i * n + j] : 0.0;
    }
  }
  for (int q0 = 0; q0 < nb; q0 += GJ_QC) {
    const int qn = min(GJ_QC, nb - q0);
    for (int e = t; e < GJ_QC * GJ_T; e += 256) {
      const int q = e / GJ_T, c = e % GJ_T;
      const bool okq = q < qn;
      const int64_t i = i0 + c, j = j0 + c;
      Rs[q][c] = (okq && i < n) ? Rt[(int64_t)(q0 + q) * n + i] : 0.0;
      Bs[q][c] = (okq && j < n) ? src[(int64_t)(k0 + q0 + q) * n + j] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int q = 0; q < GJ_QC; ++q) {
      double rv[4], bv[4];
      for (int x = 0; x < 4; ++x) rv[x] = Rs[q][ty * 4 + x];
      for (int y = 0; y < 4; ++y) bv[y] = Bs[q][tx * 4 + y];
      for (int x = 0; x < 4; ++x)
        for (int y = 0; y < 4; ++y) acc[x][y] += rv[x] * bv[y];
    }
    __syncthreads();
  }
  for (int x = 0; x < 4; ++x) {
    const int64_t i = i0 + ty * 4 + x;
    if (i >= n) break;
    for (int y = 0; y < 4; ++y) {
      const int64_t j = j0 + tx * 4 + y;
      if (j >= n) break;
      dst[i * n + j] = (j >= k0 && j < k0 + nb) ? Rt[(j - k0) * n + i] : acc[x][y];
    }
  }
}

}  // namespace tfm

// ====================================================================
// PCG handle and driver
// ====================================================================
using namespace tfm;

struct tfm_coarse {
  tfm_mesh_t m{};  // the mesh (device arrays not owned)
  int n_agg = 0, nm = 0, n_c = 0;
  int64_t n_pairs = 0;
  int32_t *agg = nullptr, *agg_ptr = nullptr, *agg_nodes = nullptr, *slot_agg = nullptr;
  int32_t *pair_off = nullptr, *pair_node = nullptr, *pair_pos = nullptr;
  double *cen = nullptr;
  double *rc = nullptr, *ec = nullptr;  // [n_c * MAXR]
  int *never = nullptr;                 // a device flag that stays 0 (segment starts)
  const double *inv64 = nullptr;        // the caller's inverse of P^T K P
  const float *inv32 = nullptr;
  CoarseView view() const {
    return CoarseView{n_agg, n_c, agg, agg_ptr, agg_nodes, slot_agg, pair_off, pair_node,
                      pair_pos, cen, rc, ec};
  }
};

static void coarse_free(tfm_coarse *h) {
  for (void *ptr : {(void *)h->agg, (void *)h->agg_ptr, (void *)h->agg_nodes, (void *)h->slot_agg,
                    (void *)h->pair_off, (void *)h->pair_node, (void *)h->pair_pos,
                    (void *)h->cen, (void *)h->rc, (void *)h->ec, (void *)h->never})
    if (ptr) (void)hipFree(ptr);
}

struct tfm_pcg {
  tfm_coarse *coarse = nullptr;  // attached coarse space (TFM_PC_TWO_LEVEL)
  int64_t n = 0;  // vector length (DOFs)
  int nrhs = 0;   // capacity
  int nrp = 1;    // capacity padded to a power of two
  int max_iter = 0;
  int device = 0;
  // inner vectors, interleaved [n][nrp]; sized for doubles, used as float in
  // mixed mode: x (correction), r (residual), u = M r, p (direction), s = K p
  double *x = nullptr, *r = nullptr, *u = nullptr, *p = nullptr, *s = nullptr;
  double *minv = nullptr;                                  // sized n*3 doubles
  double *xm = nullptr, *r64 = nullptr, *f64 = nullptr;    // FP64 work vectors [n]
  double *work = nullptr;                                  // internal-force scratch
  int64_t work_len = 0;
  double *scal = nullptr, *tol2 = nullptr, *rzold = nullptr;
  double *part_pq = nullptr, *part_upd = nullptr, *part_aux = nullptr, *part_res = nullptr;
  PcgCoef *coef = nullptr;  // [2]: coefficients of A and B (scalar-kernel mode)
  double *res2 = nullptr;
  int *done = nullptr, *alldone = nullptr, *segf = nullptr;
  int pq_cap = 0;                                          // rows of part_pq
  int res_cap = 0;                                         // rows (pairs) of part_res
  int *h_done = nullptr;                                   // pinned host mirrors
  double *h_res = nullptr;
  double *h_scal = nullptr;                                // a batch's scalars
  // optional event timing of the SpMV launches (kernel A) of the iterations
  // the FP32 arithmetic of the last mixed solve broke down (p.Kp <= 0, or a
  // true residual that stopped falling), and the FP64 iterations that finished it
  int breakdown = 0;
  int32_t fallback_iters = 0;
  int timing = 0;  // 0: off; k: event pair around the SpMV of every k-th iteration
  double spmv_ms = 0.0;
  int64_t spmv_count = 0;
  hipEvent_t ev[2 * 64] = {};                              // two per iteration of a batch
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
  for (double **v : {&h->x, &h->r, &h->u, &h->p, &h->s}) TFM_CHECK_HIP(hipMalloc(v, vbytes));
  TFM_CHECK_HIP(hipMalloc(&h->minv, sizeof(double) * (size_t)n * 3));
  for (double **v : {&h->xm, &h->r64, &h->f64}) TFM_CHECK_HIP(hipMalloc(v, sizeof(double) * (size_t)n));
  TFM_CHECK_HIP(hipMalloc(&h->scal, sizeof(double) * (size_t)(h->max_iter + 2) * nrp * NSCAL));
  TFM_CHECK_HIP(hipMalloc(&h->tol2, sizeof(double) * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->rzold, sizeof(double) * nrp));
  // one eta partial set per block of A; its widest grid has one node per
  // 8-lane share of a block, at most n / 8 + 1 blocks
  // one partial set per block of the full-grid A: at least 4 nodes per block
  // (the 64-lane team), n / 4 covers it for one DOF per node
  h->pq_cap = (int)std::max<int64_t>(RBA, n / 4 + 1);
  TFM_CHECK_HIP(hipMalloc(&h->part_pq, sizeof(double) * (size_t)h->pq_cap * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->part_upd, sizeof(double) * (RBW + CGB) * 2 * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->part_aux, sizeof(double) * RBW * nrp));
  h->res_cap = (int)std::max<int64_t>(RBV, n / 8 + 1);  // >= n_nod / 8 node teams
  TFM_CHECK_HIP(hipMalloc(&h->part_res, sizeof(double) * (size_t)h->res_cap * 2));
  TFM_CHECK_HIP(hipMalloc(&h->coef, sizeof(PcgCoef) * 2));
  TFM_CHECK_HIP(hipMalloc(&h->res2, sizeof(double) * 2));
  TFM_CHECK_HIP(hipMalloc(&h->done, sizeof(int) * nrp));
  TFM_CHECK_HIP(hipMalloc(&h->alldone, sizeof(int)));
  TFM_CHECK_HIP(hipMalloc(&h->segf, sizeof(int) * nrp));
  TFM_CHECK_HIP(hipHostMalloc(&h->h_done, sizeof(int) * (nrp + 1), hipHostMallocDefault));
  TFM_CHECK_HIP(hipHostMalloc(&h->h_res, sizeof(double) * 4, hipHostMallocDefault));
  TFM_CHECK_HIP(hipHostMalloc(&h->h_scal, sizeof(double) * PCG_CHK_MAX * nrp * NSCAL,
                              hipHostMallocDefault));
  for (auto &e : h->ev) TFM_CHECK_HIP(hipEventCreate(&e));
  return TFM_OK;
}

static void pcg_free(tfm_pcg *h) {
  for (double *ptr : {h->x, h->r, h->u, h->p, h->s, h->minv, h->xm, h->r64, h->f64, h->work,
                      h->scal, h->tol2, h->rzold, h->part_pq, h->part_upd, h->part_aux,
                      h->part_res, h->res2})
    if (ptr) (void)hipFree(ptr);
  if (h->coef) (void)hipFree(h->coef);
  if (h->done) (void)hipFree(h->done);
  if (h->alldone) (void)hipFree(h->alldone);
  if (h->segf) (void)hipFree(h->segf);
  if (h->h_done) (void)hipHostFree(h->h_done);
  if (h->h_res) (void)hipHostFree(h->h_res);
  if (h->h_scal) (void)hipHostFree(h->h_scal);
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
  h->timing = enable > 0 ? enable : 0;
  return TFM_OK;
}

extern "C" int tfm_coarse_create(tfm_coarse_t **out, const tfm_mesh_t *m, int32_t n_agg,
                                 const int32_t *agg, const int32_t *agg_ptr,
                                 const int32_t *agg_nodes, const double *centroids,
                                 const int32_t *slot_agg, const int32_t *pair_off,
                                 int64_t n_pairs, const int32_t *pair_node,
                                 const int32_t *pair_pos) {
  TFM_REQUIRE(out && m && agg && agg_ptr && agg_nodes && centroids && slot_agg && pair_off,
              "null argument");
  TFM_REQUIRE(n_pairs == 0 || (pair_node && pair_pos), "null pair arrays");
  TFM_REQUIRE(n_agg >= 1 && n_pairs >= 0, "bad aggregate or pair count");
  TFM_REQUIRE(m->ndof == 1 || m->ndof == m->dim, "ndof must be 1 or the dimension");
  const int nm = m->ndof == 3 ? 6 : m->ndof == 2 ? 3 : 1;
  TFM_REQUIRE((int64_t)n_agg * nm <= 32768, "coarse space above 32768 unknowns");
  auto *h = new tfm_coarse();
  h->m = *m;
  h->n_agg = n_agg;
  h->nm = nm;
  h->n_c = n_agg * nm;
  h->n_pairs = n_pairs;
  auto copy = [&](auto **dst, const auto *src, size_t count) -> int {
    using E = std::remove_pointer_t<std::remove_reference_t<decltype(*dst)>>;
    TFM_CHECK_HIP(hipMalloc((void **)dst, sizeof(E) * std::max<size_t>(count, 1)));
    if (count) TFM_CHECK_HIP(hipMemcpy(*dst, src, sizeof(E) * count, hipMemcpyDeviceToDevice));
    return TFM_OK;
  };
  int rc = TFM_OK;
  if (rc == TFM_OK) rc = copy(&h->agg, agg, (size_t)m->n_nod);
  if (rc == TFM_OK) rc = copy(&h->agg_ptr, agg_ptr, (size_t)n_agg + 1);
  if (rc == TFM_OK) rc = copy(&h->agg_nodes, agg_nodes, (size_t)m->n_nod);
  if (rc == TFM_OK) rc = copy(&h->cen, centroids, (size_t)n_agg * m->dim);
  if (rc == TFM_OK) rc = copy(&h->slot_agg, slot_agg, (size_t)n_agg * 27);
  if (rc == TFM_OK) rc = copy(&h->pair_off, pair_off, (size_t)n_agg * 27 + 1);
  if (rc == TFM_OK) rc = copy(&h->pair_node, pair_node, (size_t)n_pairs);
  if (rc == TFM_OK) rc = copy(&h->pair_pos, pair_pos, (size_t)n_pairs);
  if (rc == TFM_OK && hipMalloc(&h->rc, sizeof(double) * h->n_c * MAXR) != hipSuccess) rc = TFM_ERR_HIP;
  if (rc == TFM_OK && hipMalloc(&h->ec, sizeof(double) * h->n_c * MAXR) != hipSuccess) rc = TFM_ERR_HIP;
  if (rc == TFM_OK && hipMalloc(&h->never, sizeof(int)) != hipSuccess) rc = TFM_ERR_HIP;
  if (rc == TFM_OK && hipMemset(h->never, 0, sizeof(int)) != hipSuccess) rc = TFM_ERR_HIP;
  if (rc != TFM_OK) {
    set_error("allocation of the coarse space failed");
    coarse_free(h);
    delete h;
    return rc;
  }
  *out = h;
  return TFM_OK;
}

extern "C" int tfm_coarse_destroy(tfm_coarse_t *h) {
  if (!h) return TFM_OK;
  coarse_free(h);
  delete h;
  return TFM_OK;
}

extern "C" int tfm_coarse_size(const tfm_coarse_t *h, int32_t *n_c) {
  TFM_REQUIRE(h && n_c, "null argument");
  *n_c = h->n_c;
  return TFM_OK;
}

extern "C" int tfm_coarse_operator(tfm_stream_t stream, const tfm_coarse_t *h,
                                   const tfm_system_t *sys, double *Ac) {
  TFM_REQUIRE(h && sys && sys->mesh && sys->val64 && sys->con && Ac, "null argument");
  const tfm_mesh_t &m = *sys->mesh;
  TFM_REQUIRE(m.n_nod == h->m.n_nod && m.ndof == h->m.ndof, "system and coarse space differ");
  hipStream_t s = (hipStream_t)stream;
  const int64_t nn = (int64_t)h->n_c * h->n_c;
  hipLaunchKernelGGL((k_fill_zero<double>), dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s,
                     nn, Ac);
  TFM_CHECK_LAUNCH();
  const CoarseView cv = h->view();
  switch (m.ndof) {
    case 1: hipLaunchKernelGGL(k_coarse_operator<1>, dim3((unsigned)h->n_agg * 27), dim3(64), 0, s, m, cv, sys->val64, sys->con, Ac); break;
    case 2: hipLaunchKernelGGL(k_coarse_operator<2>, dim3((unsigned)h->n_agg * 27), dim3(64), 0, s, m, cv, sys->val64, sys->con, Ac); break;
    default: hipLaunchKernelGGL(k_coarse_operator<3>, dim3((unsigned)h->n_agg * 27), dim3(64), 0, s, m, cv, sys->val64, sys->con, Ac); break;
  }
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

extern "C" int tfm_coarse_set_inverse(tfm_coarse_t *h, const double *inv64, const float *inv32) {
  TFM_REQUIRE(h && inv64, "null argument");
  h->inv64 = inv64;
  h->inv32 = inv32;
  return TFM_OK;
}

extern "C" int tfm_pcg_set_coarse(tfm_pcg_t *p, tfm_coarse_t *c) {
  TFM_REQUIRE(p != nullptr, "null solver handle");
  p->coarse = c;
  return TFM_OK;
}

extern "C" int tfm_spd_inverse_work(int32_t n, int64_t *count) {
  TFM_REQUIRE(count != nullptr, "null argument");
  TFM_REQUIRE(n >= 1 && n <= 32768, "matrix size out of range (1 .. 32768)");
  *count = (int64_t)n * n + (int64_t)GJ_NB * n + GJ_NB * GJ_NB;
  return TFM_OK;
}

extern "C" int tfm_spd_inverse(tfm_stream_t stream, int32_t n, double *a, double *work,
                               int32_t *status) {
  TFM_REQUIRE(a && work && status, "null argument");
  TFM_REQUIRE(n >= 1 && n <= 32768, "matrix size out of range (1 .. 32768)");
  hipStream_t s = (hipStream_t)stream;
  const int64_t nn = (int64_t)n * n;
  double *buf[2] = {a, work};
  double *Rt = work + nn;
  double *D = Rt + (int64_t)GJ_NB * n;
  const unsigned tiles = (unsigned)((n + GJ_T - 1) / GJ_T);
  int cur = 0;
  for (int k0 = 0; k0 < n; k0 += GJ_NB) {
    const int nb = std::min(GJ_NB, n - k0);
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, s, buf[cur], (int64_t)n, k0, nb, D,
                       status);
    hipLaunchKernelGGL(k_gj_panel, dim3((unsigned)((n + GJ_PR - 1) / GJ_PR)), dim3(256), 0, s,
                       buf[cur], (int64_t)n, k0, nb, D, Rt);
    hipLaunchKernelGGL(k_gj_update, dim3(tiles, tiles), dim3(256), 0, s, buf[cur], buf[cur ^ 1],
                       (int64_t)n, k0, nb, Rt);
    TFM_CHECK_LAUNCH();
    cur ^= 1;
  }
  if (cur == 1)
    TFM_CHECK_HIP(hipMemcpyAsync(a, work, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
  return TFM_OK;
}

extern "C" int tfm_pcg_fp64_fallback(const tfm_pcg_t *h, int32_t *fp64_iters) {
  TFM_REQUIRE(h != nullptr && fp64_iters != nullptr, "null argument");
  *fp64_iters = h->fallback_iters;
  return TFM_OK;
}

namespace {

// Operator callbacks of one solve: kernel A of an iteration (reads n_in
// partial sets of (gamma, rr) at part_in; returns its grid, the number of eta
// partial sets, through *g), the plain inner SpMV y = K x, and the FP64 true
// residual of one rhs (res = b - K x; returns ||res||^2 and ||b||^2). With
// `scal`, one-block scalar kernels prepare the coefficients of A and B.
template <typename T> struct Ops {
  bool scal = false;
  std::function<int(const T *, T *, T *, int, const double *, int, const PcgState &, unsigned *)>
      spmv_dir;
  std::function<int(const T *, T *)> spmv;
  std::function<int(const double *, const double *, double *, double *, double *)> residual;
  // two-level preconditioner: u += P Ac^{-1} P^T r and the partial sets of
  // (P^T r).(Ac^{-1} P^T r) at `part`; skipped once *flag is set (empty if unused)
  std::function<int(const T *, T *, double *, const int *)> coarse;
};

// The coarse correction's three launches: r_c = P^T r, e_c = Ac^{-1} r_c
// (with the partials of r_c . e_c), u += P e_c.
template <typename T, int ND, int NRP>
std::function<int(const T *, T *, double *, const int *)> coarse_ops(hipStream_t s,
                                                                     const tfm_mesh_t &m,
                                                                     const tfm_coarse &c,
                                                                     const uint8_t *con,
                                                                     const T *inv) {
  const CoarseView cv = c.view();
  const tfm_mesh_t mm = m;
  return [=](const T *r, T *u, double *part, const int *flag) -> int {
    hipLaunchKernelGGL((k_coarse_restrict<T, ND, NRP>), dim3((unsigned)cv.n_agg), dim3(BS), 0, s,
                       mm, cv, con, r, flag);
    TFM_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_coarse_solve<T, NRP>), dim3((unsigned)((cv.n_c + 3) / 4)), dim3(BS), 0,
                       s, cv, inv, flag);
    TFM_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_coarse_prolong<T, ND, NRP>),
                       dim3((unsigned)((mm.n_nod * NRP + 255) / 256)), dim3(256), 0, s, mm, cv, con,
                       u, part, flag);
    TFM_CHECK_LAUNCH();
    return TFM_OK;
  };
}

int sum_residual(tfm_pcg *h, hipStream_t s, unsigned nblk, double *rr, double *bb) {
  hipLaunchKernelGGL(k_sum2, dim3(1), dim3(1024), 0, s, (const double *)h->part_res, (int)nblk, h->res2);
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
            int nrhs, double rtol, double delta, double final_rtol, int32_t *iters,
            double *relres, int32_t *outer) {
  const bool fp64 = std::is_same<T, double>::value;
  const int64_t n = n_nod * ND;
  T *xv = (T *)h->x, *rv = (T *)h->r, *uv = (T *)h->u, *pv = (T *)h->p, *sv = (T *)h->s;
  const T *minv = (const T *)h->minv;
  static const int xcd = env_int("TFM_XCD_MAP", 1);
  PcgState st{h->scal,     h->tol2,     h->rzold,    h->done,     h->alldone,
              h->part_pq,  h->part_upd, h->part_aux, h->pq_cap,   h->coef,
              h->coef + 1, xcd,        h->segf};
  const unsigned g1 = (unsigned)((n + 255) / 256);
  // vector kernels: one (node, rhs) per thread when a scalar kernel sums
  // their partials (at most RBW/NRP blocks, so that it sums few partial
  // sets), else at most RBV blocks for the prologue of A
  static const int vcap = std::min(RBW, std::max(1, env_int("TFM_VEC_GRID", RBV)));
  // whole multiples of the CU count, so that no CU runs a second round alone
  const unsigned gv = std::min<unsigned>(ops.scal ? std::max(RBV, RBW / NRP) : vcap,
                                         round_to_cus(grid_for(n_nod * NRP, BSV, RBW)));
  // the coarse solve's partial sets follow the update's in part_upd
  const int n_part = (int)gv + (ops.coarse ? 1 : 0);
  double *const part_coarse = h->part_upd + (size_t)gv * 2 * NRP;
  const int *never = h->coarse ? h->coarse->never : nullptr;
  std::vector<double> htol(NRP, -1.0), bnorm(NRP, 0.0), rprev(NRP, 0.0), rnow(NRP, 0.0);
  std::vector<int> fin(NRP, 1), stall(NRP, 0);
  for (int r = 0; r < nrhs; ++r) iters[r] = 0;
  h->breakdown = 0;

  // true residual of rhs r at the current x, staged as the inner residual column
  auto stage = [&](int r, bool first) -> int {
    double rr2, bb2;
    int rc = ops.residual(b + (int64_t)r * n, x + (int64_t)r * n, h->r64, &rr2, &bb2);
    if (rc != TFM_OK) return rc;
    if (!std::isfinite(rr2)) {
      set_error("non-finite residual in the PCG solve");
      return TFM_ERR_NOCONV;
    }
    if (first) bnorm[r] = std::sqrt(bb2);
    if (first && bb2 == 0.0) {  // b = 0: x = 0 whatever the initial guess
      hipLaunchKernelGGL((k_fill_zero<double>), dim3(g1), dim3(256), 0, s, n, x + (int64_t)r * n);
      TFM_CHECK_LAUNCH();
      rr2 = 0.0;
    }
    const double rn = std::sqrt(rr2);
    rprev[r] = first ? rn : rnow[r];
    rnow[r] = rn;
    relres[r] = bnorm[r] > 0 ? rn / bnorm[r] : rn;
    fin[r] = rn <= rtol * bnorm[r];
    // the last segment stops below rtol: FP64 by a margin for the drift of
    // its recursive residual, mixed at the caller's final_rtol
    const double target = (fp64 ? 0.7 * rtol : final_rtol) * bnorm[r];
    // a mixed segment reduces the residual by delta, and the one that can
    // end the solve aims at the final target
    const double seg = fp64 || delta * rn <= rtol * bnorm[r] ? target : delta * rn;
    htol[r] = fin[r] ? -1.0 : seg * seg;
    hipLaunchKernelGGL((k_set_column<T, NRP>), dim3(g1), dim3(256), 0, s, n, r,
                       (const double *)h->r64, rv, fin[r] ? 1 : 0);
    TFM_CHECK_LAUNCH();
    return TFM_OK;
  };
  // segment start of the rhs in `mask`: u = M r, then keep (cont) or reset
  // (mask & ~cont) their directions; `running` lists every rhs still iterating
  auto seg_start_kernels = [&](int it, int mask, int cont, int running,
                               const SegInfo &seg) -> int {
    TFM_CHECK_HIP(hipMemcpyAsync(h->tol2, htol.data(), sizeof(double) * NRP, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_pcg_seg_dots<T, ND, NRP>), dim3(gv), dim3(BS), 0, s, n_nod,
                       (const T *)rv, uv, (const T *)pv, minv, st);
    TFM_CHECK_LAUNCH();
    if (ops.coarse) {
      const int rc = ops.coarse(rv, uv, part_coarse, never);
      if (rc != TFM_OK) return rc;
    }
    hipLaunchKernelGGL((k_pcg_seg_apply<T, ND, NRP>), dim3(gv), dim3(BS), 0, s, n_nod, it, mask,
                       cont, running, seg, (const T *)uv, pv, sv, st, n_part, (int)gv);
    TFM_CHECK_LAUNCH();
    if (cont) return ops.spmv(pv, sv);
    return TFM_OK;
  };
  for (T *v : {xv, rv, pv, sv}) {
    hipLaunchKernelGGL((k_fill_zero<T>), dim3((unsigned)((n * NRP + 255) / 256)), dim3(256), 0, s,
                       n * NRP, v);
    TFM_CHECK_LAUNCH();
  }
  SegInfo seg{};
  int mask = 0;
  for (int r = 0; r < nrhs; ++r) {
    const int rc = stage(r, true);
    if (rc != TFM_OK) return rc;
    if (!fin[r]) mask |= 1 << r;
  }
  for (int r = 0; r < NRP; ++r) seg.d[r] = -1;
  {
    const int rc = seg_start_kernels(0, mask, 0, mask, seg);
    if (rc != TFM_OK) return rc;
  }

  // Every rhs runs its own sequence of segments: when one reaches its
  // segment target, the host folds its correction into x and restarts it at
  // the next batch boundary while the others keep iterating, so no rhs waits
  // for the slowest. Batches between host checks are sized from the observed
  // contraction of rr so that the first running rhs to reach its target
  // stops near the batch end (few idle iterations, few host round trips).
  const int CHK = 12, CHK_MIN = 4, CHK_MAX = PCG_CHK_MAX;
  std::vector<int> segf(NRP, 0), nseg(NRP, 0);
  std::vector<double> rr_seen(NRP, -1.0);
  std::vector<int> it_seen(NRP, -1);
  int it = 0, n_outer = 0, chk = CHK;
  int running = mask;
  while (running && it < h->max_iter) {
    const int start = it;
    const int stop = std::min(h->max_iter, it + chk);
    for (; it < stop; ++it) {
      if (ops.scal) {
        hipLaunchKernelGGL((k_pcg_scal_dir<NRP>), dim3(1), dim3(BS), 0, s, it,
                           (const double *)h->part_upd, n_part, st);
        TFM_CHECK_LAUNCH();
      }
      const bool timed = h->timing && it % h->timing == 0;
      if (timed) TFM_CHECK_HIP(hipEventRecord(h->ev[2 * (it - start)], s));
      unsigned npq = 0;
      int rc = ops.spmv_dir(uv, pv, sv, it, h->part_upd, n_part, st, &npq);
      if (rc != TFM_OK) return rc;
      if (timed) TFM_CHECK_HIP(hipEventRecord(h->ev[2 * (it - start) + 1], s));
      if (ops.scal) {
        hipLaunchKernelGGL((k_pcg_scal_alpha<NRP>), dim3(1), dim3(1024), 0, s, it,
                           (const double *)h->part_pq, (int)npq, st);
        TFM_CHECK_LAUNCH();
        hipLaunchKernelGGL((k_pcg_update<T, ND, NRP, true>), dim3(gv), dim3(BSV), 0, s, n_nod, it,
                           xv, rv, uv, (const T *)pv, (const T *)sv, minv, st,
                           (const double *)h->part_pq, (int)npq);
      } else {
        hipLaunchKernelGGL((k_pcg_update<T, ND, NRP, false>), dim3(gv), dim3(BSV), 0, s, n_nod, it,
                           xv, rv, uv, (const T *)pv, (const T *)sv, minv, st,
                           (const double *)h->part_pq, (int)npq);
      }
      TFM_CHECK_LAUNCH();
      if (ops.coarse) {
        rc = ops.coarse(rv, uv, part_coarse, h->alldone);
        if (rc != TFM_OK) return rc;
      }
    }
    TFM_CHECK_HIP(hipMemcpyAsync(h->h_done, h->done, sizeof(int) * NRP, hipMemcpyDeviceToHost, s));
    TFM_CHECK_HIP(hipMemcpyAsync(h->h_scal, h->scal + (size_t)start * NRP * NSCAL,
                                 sizeof(double) * (stop - start) * NRP * NSCAL,
                                 hipMemcpyDeviceToHost, s));
    TFM_CHECK_HIP(hipStreamSynchronize(s));
    // p.Kp <= 0 in an iteration that ran: K is not positive definite
    for (int i = start; i < stop; ++i)
      for (int r = 0; r < nrhs; ++r) {
        if (!((running >> r) & 1) || (h->h_done[r] >= 0 && i >= h->h_done[r])) continue;
        const double eta = h->h_scal[((size_t)(i - start) * NRP + r) * NSCAL + S_PQ];
        if (!(eta > 0.0)) {
          set_error("the matrix is not positive definite (p.Kp <= 0 in CG)");
          for (int q = 0; q < nrhs; ++q)  // iterations of the unfinished segments
            if ((running >> q) & 1)
              iters[q] += std::max(0, (h->h_done[q] >= 0 ? std::min(i, h->h_done[q]) : i) - segf[q]);
          if (outer) *outer = *std::max_element(nseg.begin(), nseg.end());
          h->breakdown = 1;
          return TFM_ERR_NOCONV;
        }
      }
    if (h->timing) {  // launches after every running rhs stopped were empty
      int end = 0;
      bool all_stopped = true;
      for (int r = 0; r < nrhs; ++r)
        if ((running >> r) & 1) {
          all_stopped &= h->h_done[r] >= 0;
          end = std::max(end, h->h_done[r]);
        }
      const int real_end = all_stopped ? std::min(it, end) : it;
      for (int i = start; i < real_end; ++i) {
        if (i % h->timing) continue;
        float ms = 0.0f;
        TFM_CHECK_HIP(hipEventElapsedTime(&ms, h->ev[2 * (i - start)], h->ev[2 * (i - start) + 1]));
        h->spmv_ms += ms;
        h->spmv_count += 1;
      }
    }
    // the rhs that reached their segment target: fold the corrections into x
    // and replace the residuals by true ones; FP64 (whose corrections only
    // absorb the drift of the recursive residual) and a mixed solve whose true
    // residual stopped falling restart from the new residual, other mixed
    // solves keep their direction
    const double *last = h->h_scal + (size_t)(stop - 1 - start) * NRP * NSCAL;
    int cont = 0;
    mask = 0;
    for (int r = 0; r < NRP; ++r) seg.d[r] = h->h_done[r];
    for (int r = 0; r < nrhs; ++r) {
      if (!((running >> r) & 1) || h->h_done[r] < 0) continue;
      iters[r] += std::max(0, h->h_done[r] - segf[r]);
      ++nseg[r];
      hipLaunchKernelGGL((k_take_column<T, NRP>), dim3(g1), dim3(256), 0, s, n, r, xv,
                         x + (int64_t)r * n);
      TFM_CHECK_LAUNCH();
      const int rc = stage(r, false);
      if (rc != TFM_OK) return rc;
      rr_seen[r] = -1.0;
      it_seen[r] = -1;
      if (fin[r]) {
        running &= ~(1 << r);
        continue;
      }
      mask |= 1 << r;
      const bool stuck = rnow[r] > 0.9 * rprev[r];
      stall[r] = stuck ? stall[r] + 1 : 0;
      if (!fp64 && !stuck && seg.d[r] > segf[r]) cont |= 1 << r;
    }
    // a residual that no correction improves has reached the attainable
    // accuracy of the arithmetic: give up instead of cycling
    for (int r = 0; r < nrhs; ++r)
      if (!fin[r] && stall[r] >= 8) {
        set_error("PCG stagnated above the tolerance (attainable accuracy reached)");
        if (outer) *outer = *std::max_element(nseg.begin(), nseg.end());
        h->breakdown = 1;
        return TFM_ERR_NOCONV;
      }
    if (mask) {
      const int rc = seg_start_kernels(it, mask, cont, running, seg);
      if (rc != TFM_OK) return rc;
      for (int r = 0; r < nrhs; ++r)
        if ((mask >> r) & 1) segf[r] = it;
    }
    // next batch: iterations until the first running rhs reaches its segment
    // target at its observed rate (rr of iteration stop-1); a rhs that has
    // just started a segment counts CHK
    int want = CHK_MAX;
    for (int r = 0; r < nrhs; ++r) {
      if (!((running >> r) & 1)) continue;
      int k = CHK;
      if (!((mask >> r) & 1)) {
        const double rr = last[r * NSCAL + S_RR];
        if (it_seen[r] >= 0 && rr > 0.0 && rr < rr_seen[r] && htol[r] > 0.0) {
          const double rate = std::log(rr / rr_seen[r]) / (stop - 1 - it_seen[r]);
          k = (int)std::ceil(std::log(htol[r] / rr) / rate) + 2;
        }
        rr_seen[r] = rr;
        it_seen[r] = stop - 1;
      }
      want = std::min(want, std::min(std::max(k, CHK_MIN), CHK_MAX));
    }
    chk = want;
  }
  for (int r = 0; r < nrhs; ++r)
    if ((running >> r) & 1) iters[r] += std::max(0, it - segf[r]);
  n_outer = *std::max_element(nseg.begin(), nseg.end());
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
// extended: double-double row sums (tolerances near the FP64 floor of
// eps |K| |x|, below 1e-10); otherwise one FP64 SpMV.
int fem_residual(tfm_pcg *h, hipStream_t s, const tfm_system_t &sys, const double *b,
                 const double *x, double *res, double *rr, double *bb, bool extended) {
  const tfm_mesh_t &m = *sys.mesh;
  const int64_t n = m.n_nod * m.ndof;
  const unsigned g1 = (unsigned)((n + 255) / 256);
  const uint8_t *con = nullptr;
  if (sys.val64 && !extended) {
    int rc = launch_spmv(s, m, 0, sys.val64, x, h->f64, 1);
    if (rc != TFM_OK) return rc;
  } else if (sys.val64) {
    const unsigned gr = grid_for(m.n_nod, BS / 32, h->res_cap);  // one node per team
    TFM_ND_SWITCH(m.ndof, hipLaunchKernelGGL((k_residual_dd<ND, 32>), dim3(gr), dim3(BS), 0, s, m,
                                             sys.val64, b, x, res, h->part_res));
    TFM_CHECK_LAUNCH();
    return sum_residual(h, s, gr, rr, bb);
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
  const unsigned gr = grid_for(n, BS, RBV);
  hipLaunchKernelGGL(k_residual, dim3(gr), dim3(BS), 0, s, n, con, b, x, (const double *)h->f64,
                     res, h->part_res);
  TFM_CHECK_LAUNCH();
  return sum_residual(h, s, gr, rr, bb);
}

}  // namespace

extern "C" int tfm_pcg_solve(tfm_pcg_t *h, tfm_stream_t stream, const tfm_system_t *sys,
                             int32_t prec, int32_t pc, const double *b, double *x, int32_t nrhs,
                             double rtol, double inner_rtol, double final_rtol,
                             int32_t *iters, double *relres, int32_t *outer) {
  TFM_REQUIRE(h && sys && sys->mesh && b && x && iters && relres, "null argument");
  const tfm_mesh_t &m = *sys->mesh;
  const int64_t n = m.n_nod * m.ndof;
  TFM_REQUIRE(n == h->n, "system size differs from the solver handle");
  TFM_REQUIRE(nrhs >= 1 && nrhs <= h->nrhs, "nrhs exceeds the solver handle");
  TFM_REQUIRE(pc == TFM_PC_JACOBI || pc == TFM_PC_BLOCK_JACOBI || pc == TFM_PC_TWO_LEVEL,
              "unknown preconditioner");
  const tfm_coarse *cs = pc == TFM_PC_TWO_LEVEL ? h->coarse : nullptr;
  if (pc == TFM_PC_TWO_LEVEL) {
    TFM_REQUIRE(cs != nullptr, "TFM_PC_TWO_LEVEL needs a coarse space (tfm_pcg_set_coarse)");
    TFM_REQUIRE(cs->m.n_nod == m.n_nod && cs->m.ndof == m.ndof,
                "the coarse space belongs to another mesh");
    TFM_REQUIRE(cs->inv64 != nullptr && (prec == TFM_PREC_FP64 || cs->inv32 != nullptr),
                "the coarse inverse is missing (tfm_coarse_set_inverse)");
    pc = TFM_PC_BLOCK_JACOBI;
  }
  TFM_REQUIRE(prec == TFM_PREC_FP64 || prec == TFM_PREC_MIXED, "unknown precision mode");
  TFM_REQUIRE(sys->con && sys->diag_blocks, "system needs con and diag_blocks");
  TFM_REQUIRE(prec == TFM_PREC_FP64 ? sys->val64 != nullptr : sys->val32 != nullptr,
              "matrix values for the requested precision are missing");
  TFM_REQUIRE(prec == TFM_PREC_FP64 ||
                  sys->val32_form == (pow2_cap(nrhs) <= 2 ? TFM_FP32_DIFF : TFM_FP32_PLAIN),
              "FP32 values in the wrong form for this number of right-hand sides "
              "(TFM_FP32_DIFF for one or two, TFM_FP32_PLAIN for more)");
  TFM_REQUIRE(rtol > 0.0, "rtol must be positive");
  TFM_REQUIRE(inner_rtol > 0.0 && inner_rtol < 1.0, "inner_rtol must be in (0, 1)");
  TFM_REQUIRE(final_rtol > 0.0 && final_rtol <= rtol, "final_rtol must be in (0, rtol]");
  if (outer) *outer = 0;
  hipStream_t s = (hipStream_t)stream;
  const int nrp = pow2_cap(nrhs);
  int rc = TFM_OK;
  const bool extended = rtol < 1e-10;
  auto residual = [&](const double *bb, const double *xx, double *res, double *rr, double *b2) -> int {
    return fem_residual(h, s, *sys, bb, xx, res, rr, b2, extended);
  };
  h->fallback_iters = 0;
  TFM_ND_SWITCH(m.ndof, TFM_NRP_SWITCH(nrp, {
    auto run_fp64 = [&](int32_t *its, int32_t *out) -> int {
      int rc64 = build_precond<double>(s, m.n_nod, ND, sys->diag_blocks, pc, (double *)h->minv);
      if (rc64 != TFM_OK) return rc64;
      Ops<double> ops;
      ops.scal = spmv_shape<double>(m, NRP).scal;
      ops.spmv_dir = [&](const double *u, double *p, double *sv, int it,
                         const double *pin, int n_in, const PcgState &st, unsigned *g) -> int {
        return spmv_dir_launch<double, ND, NRP>(s, m, sys->val64, u, p, sv, it, pin, n_in, st,
                                                g);
      };
      ops.spmv = [&](const double *xi, double *yo) -> int {
        return spmv_launch<double, ND, NRP>(s, m, sys->val64, xi, yo);
      };
      ops.residual = residual;
      if (cs) ops.coarse = coarse_ops<double, ND, NRP>(s, m, *cs, sys->con, cs->inv64);
      return run_pcg<double, ND, NRP>(h, s, m.n_nod, ops, b, x, nrhs, rtol, inner_rtol,
                                      final_rtol, its, relres, out);
    };
    if (prec == TFM_PREC_FP64) {
      rc = run_fp64(iters, outer);
    } else {
      rc = build_precond<float>(s, m.n_nod, ND, sys->diag_blocks, pc, (float *)h->minv);
      if (rc != TFM_OK) return rc;
      Ops<float> ops;
      ops.scal = spmv_shape<float>(m, NRP).scal;
      ops.spmv_dir = [&](const float *u, float *p, float *sv, int it, const double *pin,
                         int n_in, const PcgState &st, unsigned *g) -> int {
        return spmv_dir_launch<float, ND, NRP>(s, m, sys->val32, u, p, sv, it, pin, n_in, st,
                                               g);
      };
      ops.spmv = [&](const float *xi, float *yo) -> int {
        return spmv_launch<float, ND, NRP>(s, m, sys->val32, xi, yo);
      };
      ops.residual = residual;
      if (cs) ops.coarse = coarse_ops<float, ND, NRP>(s, m, *cs, sys->con, cs->inv32);
      rc = run_pcg<float, ND, NRP>(h, s, m.n_nod, ops, b, x, nrhs, rtol, inner_rtol, final_rtol,
                                   iters, relres, outer);
      // FP32 rounding of a matrix whose element stiffnesses span many orders
      // of magnitude (void elements of a density design) can make the FP32
      // operator indefinite or cap the accuracy of its corrections: then
      // finish in FP64 from the x of the last refinement
      if (rc == TFM_ERR_NOCONV && h->breakdown && sys->val64) {
        std::vector<int32_t> its64(nrhs, 0);
        int32_t out64 = 0;
        rc = run_fp64(its64.data(), &out64);
        for (int r = 0; r < nrhs; ++r) iters[r] += its64[r];
        h->fallback_iters = *std::max_element(its64.begin(), its64.end());
        if (outer) *outer += out64;
      }
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

constexpr int CSR_TL = 8;

int csr_spmv(hipStream_t s, int64_t n, const int64_t *rowptr, const int32_t *col,
             const double *val, const double *x, double *y) {
  const unsigned g = grid_for(n, BS / CSR_TL, RBA);
  hipLaunchKernelGGL((k_spmv_csr<CSR_TL>), dim3(g), dim3(BS), 0, s, n, rowptr, col, val, x, y);
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

}  // namespace

extern "C" int tfm_csr_spmv(tfm_stream_t s, int64_t n, const int64_t *rowptr, const int32_t *col,
                            const double *val, const double *x, double *y) {
  TFM_REQUIRE(n >= 0, "negative size");
  if (n == 0) return TFM_OK;
  TFM_REQUIRE(rowptr && col && val && x && y, "null argument");
  return csr_spmv((hipStream_t)s, n, rowptr, col, val, x, y);
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
                     (int)TFM_PC_JACOBI, (double *)h->minv);
  TFM_CHECK_LAUNCH();
  Ops<double> ops;
  ops.spmv_dir = [&](const double *u, double *p, double *sv, int it, const double *pin,
                     int n_in, const PcgState &st, unsigned *g) -> int {
    *g = grid_for(n, BS / CSR_TL, RBA);
    TFM_REQUIRE((int64_t)*g <= st.pq_cap, "SpMV grid exceeds the partial-sum buffer");
    hipLaunchKernelGGL((k_pcg_csr_dir<CSR_TL>), dim3(*g), dim3(BS), 0, s, n, rowptr, col, val, u,
                       p, sv, it, pin, n_in, st);
    TFM_CHECK_LAUNCH();
    return TFM_OK;
  };
  ops.spmv = [&](const double *xi, double *yo) -> int {
    return csr_spmv(s, n, rowptr, col, val, xi, yo);
  };
  ops.residual = [&](const double *bb, const double *xx, double *res, double *rr, double *b2) -> int {
    int rc = csr_spmv(s, n, rowptr, col, val, xx, h->f64);
    if (rc != TFM_OK) return rc;
    const unsigned gr = grid_for(n, BS, RBV);
    hipLaunchKernelGGL(k_residual, dim3(gr), dim3(BS), 0, s, n, (const uint8_t *)nullptr, bb, xx,
                       (const double *)h->f64, res, h->part_res);
    TFM_CHECK_LAUNCH();
    return sum_residual(h, s, gr, rr, b2);
  };
  int32_t outer = 0;
  return run_pcg<double, 1, 1>(h, s, n, ops, b, x, 1, rtol, 0.5, 0.7 * rtol, iters, relres,
                               &outer);
}
