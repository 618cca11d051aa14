// Krylov solvers for general sparse matrices (CSR, FP64) behind
// torchfem.sparse: preconditioned MINRES for symmetric matrices and
// restarted GMRES for any square matrix, both Jacobi-preconditioned.
//
// The iteration scalars live on the device. Each iteration is a short chain
// of kernels: row-parallel kernels write per-block partial sums, and a
// one-block scalar kernel adds them in block order (deterministic, no
// atomics) and advances the recurrences (Lanczos and Givens rotations for
// MINRES, Gram-Schmidt and Givens for GMRES). The host only reads a small
// state record every few iterations to test convergence, and checks the
// true residual b - A x before it returns.
#include "tfm_common.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace tfm {
namespace {

constexpr int KB = 256;       // threads per block of the row kernels
constexpr int KTL = 8;        // lanes per CSR row in the products
constexpr int KMAXB = 1024;   // max blocks of a reducing row kernel
constexpr int KCHK = 8;       // iterations between host checks

unsigned kgrid(int64_t work, int per_block) {
  const int64_t g = (work + per_block - 1) / per_block;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, KMAXB));
}

// Block partial sums of up to NV values per thread: the block's totals go to
// part[v * gridDim.x + blockIdx.x] (wave butterfly, then one LDS stage).
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *part) {
  __shared__ double sh[NV][KB / 64];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == 0) sh[k][wave] = t;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < KB / 64; ++w) t += sh[threadIdx.x][w];
    part[threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// Sum of nb partials in block order by one block (the same bits on every call).
__device__ double sum_in_order(const double *part, int nb, double *sh) {
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) t += part[b];
  // fixed tree over the threads
  sh[threadIdx.x] = t;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Row i of A times x, by a team of KTL lanes (all lanes get the sum).
__device__ __forceinline__ double row_dot(const int64_t *rowptr, const int32_t *col,
                                          const double *val, const double *x, const double *scl,
                                          int64_t i, int lane) {
  double t = 0.0;
  for (int64_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += KTL) {
    const int64_t c = col[k];
    t += val[k] * (scl ? scl[c] * x[c] : x[c]);
  }
#pragma unroll
  for (int o = 1; o < KTL; o <<= 1) t += __shfl_xor(t, o, KTL);
  return t;
}

// inverse diagonal magnitudes 1 / |A_ii| (1 for a zero diagonal): an SPD
// Jacobi preconditioner for symmetric (possibly indefinite) matrices
__global__ void k_inv_absdiag(int64_t n, const int64_t *rowptr, const int32_t *col,
                              const double *val, double *dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    if (col[k] == i) d += val[k];
  d = fabs(d);
  dinv[i] = d > 0.0 ? 1.0 / d : 1.0;
}

// r = b - A x; z = dinv r (optional); partials of (r.z or r.r, b.b)
__global__ void __launch_bounds__(KB) k_residual(int64_t n, const int64_t *rowptr,
                                                 const int32_t *col, const double *val,
                                                 const double *b, const double *x,
                                                 const double *dinv, double *r, double *z,
                                                 double *part) {
  const int lane = threadIdx.x % KTL, team = threadIdx.x / KTL;
  double acc[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * (KB / KTL) + team; i < n;
       i += (int64_t)gridDim.x * (KB / KTL)) {
    const double t = row_dot(rowptr, col, val, x, nullptr, i, lane);
    if (lane == 0) {
      const double ri = b[i] - t;
      r[i] = ri;
      const double zi = dinv ? dinv[i] * ri : ri;
      if (z) z[i] = zi;
      acc[0] += ri * zi;
      acc[1] += b[i] * b[i];
    }
  }
  block_partials<2>(acc, part);
}

// ================================================================ MINRES
// Preconditioned MINRES (Paige and Saunders, SIAM J. Numer. Anal. 12, 1975):
// the Lanczos process on M^{-1} A in the M-inner product, the tridiagonal
// least-squares problem solved by Givens rotations, and x updated through
// the three-term recurrence of the search directions w.
struct MrState {
  double beta, oldb, alfa, dbar, epsln, oldeps, phibar, cs, sn, delta, gamma, phi, denom;
  double beta1, tol;
  int live;  // 0 once converged or broken down
  int it;    // iterations done
};

// Start: beta1 = sqrt(r.z) from the residual partials (r = b - A x, z = M r).
__global__ void __launch_bounds__(256) k_mr_start(const double *part, int nb, double rtol,
                                                  MrState *st) {
  __shared__ double sh[256];
  const double rz = sum_in_order(part, nb, sh);
  if (threadIdx.x == 0) {
    const double b1 = rz > 0.0 ? sqrt(rz) : 0.0;
    MrState s{};
    s.beta = b1;
    s.beta1 = b1;
    s.phibar = b1;
    s.cs = -1.0;
    s.tol = 0.5 * rtol * b1;  // the estimate is in the M-norm: a margin
    s.live = b1 > 0.0;
    *st = s;
  }
}

// Lanczos product: v = y / beta, q = A v - (beta / oldb) r1; partials of v.q
__global__ void __launch_bounds__(KB) k_mr_lanczos(int64_t n, const int64_t *rowptr,
                                                   const int32_t *col, const double *val,
                                                   const double *y, const double *r1, double *v,
                                                   double *q, const MrState *st, double *part) {
  if (!st->live) return;
  const double rb = 1.0 / st->beta;
  const double c1 = st->it > 0 ? st->beta / st->oldb : 0.0;
  const int lane = threadIdx.x % KTL, team = threadIdx.x / KTL;
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * (KB / KTL) + team; i < n;
       i += (int64_t)gridDim.x * (KB / KTL)) {
    const double t = row_dot(rowptr, col, val, y, nullptr, i, lane) * rb;
    if (lane == 0) {
      const double vi = y[i] * rb;
      const double qi = t - c1 * r1[i];
      v[i] = vi;
      q[i] = qi;
      acc[0] += vi * qi;
    }
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(256) k_mr_alfa(const double *part, int nb, MrState *st) {
  __shared__ double sh[256];
  if (!st->live) return;
  const double a = sum_in_order(part, nb, sh);
  if (threadIdx.x == 0) st->alfa = a;
}

// r2' = q - (alfa / beta) r2 (into the r1 buffer); z = M r2'; partials r2'.z
__global__ void __launch_bounds__(KB) k_mr_orth(int64_t n, const double *q, const double *r2,
                                                double *r2new, const double *dinv, double *z,
                                                const MrState *st, double *part) {
  if (!st->live) return;
  const double c = st->alfa / st->beta;
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x; i < n; i += (int64_t)gridDim.x * KB) {
    const double ri = q[i] - c * r2[i];
    const double zi = dinv[i] * ri;
    r2new[i] = ri;
    z[i] = zi;
    acc[0] += ri * zi;
  }
  block_partials<1>(acc, part);
}

// The new Lanczos coefficient beta and the Givens rotations.
__global__ void __launch_bounds__(256) k_mr_rotate(const double *part, int nb, MrState *st) {
  __shared__ double sh[256];
  if (!st->live) return;
  const double rz = sum_in_order(part, nb, sh);
  if (threadIdx.x != 0) return;
  MrState s = *st;
  s.oldb = s.beta;
  s.beta = rz > 0.0 ? sqrt(rz) : 0.0;
  s.oldeps = s.epsln;
  s.delta = s.cs * s.dbar + s.sn * s.alfa;
  const double gbar = s.sn * s.dbar - s.cs * s.alfa;
  s.epsln = s.sn * s.beta;
  s.dbar = -s.cs * s.beta;
  const double g = sqrt(gbar * gbar + s.beta * s.beta);
  s.gamma = g > 1e-300 ? g : 1e-300;
  s.cs = gbar / s.gamma;
  s.sn = s.beta / s.gamma;
  s.phi = s.cs * s.phibar;
  s.phibar = s.sn * s.phibar;
  s.denom = 1.0 / s.gamma;
  s.it += 1;
  // converged (estimate of the preconditioned residual) or the Krylov space
  // is exhausted (beta = 0: the last update below is the exact solution)
  *st = s;
}

// w' = (v - oldeps w1 - delta w2) / gamma (into the w1 buffer); x += phi w'
__global__ void __launch_bounds__(KB) k_mr_update(int64_t n, const double *v, double *w1,
                                                  const double *w2, double *x, MrState *st) {
  const MrState s = *st;
  if (!s.live) return;
  for (int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x; i < n; i += (int64_t)gridDim.x * KB) {
    const double wi = (v[i] - s.oldeps * w1[i] - s.delta * w2[i]) * s.denom;
    w1[i] = wi;
    x[i] += s.phi * wi;
  }
}

__global__ void k_mr_check(MrState *st, int maxit) {
  MrState &s = *st;
  if (s.live && (s.phibar <= s.tol || s.beta <= 1e-300 || s.it >= maxit)) s.live = 0;
}

// ================================================================= GMRES
// Restarted GMRES(m) with right Jacobi preconditioning, x = x0 + D^{-1} V y:
// the Arnoldi basis by classical Gram-Schmidt with one reorthogonalisation
// pass, the Hessenberg least-squares problem by Givens rotations.
struct GmState {
  double beta;   // norm of the restart residual
  double resid;  // |g_{j+1}|: the least-squares residual estimate
  double tol;
  int j;         // Arnoldi steps done in this cycle
  int live;
};

__global__ void __launch_bounds__(256) k_gm_start(const double *part, int nb, double tol, int m,
                                                  GmState *st, double *g, double *H) {
  __shared__ double sh[256];
  const double rr = sum_in_order(part, nb, sh);
  if (threadIdx.x == 0) {
    const double b = rr > 0.0 ? sqrt(rr) : 0.0;
    st->beta = b;
    st->resid = b;
    st->tol = tol;
    st->j = 0;
    st->live = b > tol;
    g[0] = b;
  }
  for (int k = threadIdx.x + 1; k <= m; k += blockDim.x) g[k] = 0.0;
  for (int k = threadIdx.x; k < (m + 1) * m; k += blockDim.x) H[k] = 0.0;
}

// v0 = r / beta
__global__ void k_gm_first(int64_t n, const double *r, double *v0, const GmState *st) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !st->live) return;
  v0[i] = r[i] / st->beta;
}

// w = A D^{-1} v_j, j = st->j
__global__ void __launch_bounds__(KB) k_gm_product(int64_t n, const int64_t *rowptr,
                                                   const int32_t *col, const double *val,
                                                   const double *V, int64_t ldv,
                                                   const double *dinv, double *w,
                                                   const GmState *st) {
  if (!st->live) return;
  const double *vj = V + st->j * ldv;
  const int lane = threadIdx.x % KTL, team = threadIdx.x / KTL;
  for (int64_t i = (int64_t)blockIdx.x * (KB / KTL) + team; i < n;
       i += (int64_t)gridDim.x * (KB / KTL)) {
    const double t = row_dot(rowptr, col, val, vj, dinv, i, lane);
    if (lane == 0) w[i] = t;
  }
}

// partials of V_k . w for k = blockIdx.y <= j (one dot per grid row)
__global__ void __launch_bounds__(KB) k_gm_dots(int64_t n, const double *V, const double *w,
                                                int64_t ldv, const GmState *st, double *part) {
  if (!st->live || (int)blockIdx.y > st->j) return;
  const double *vk = V + blockIdx.y * ldv;
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x; i < n; i += (int64_t)gridDim.x * KB)
    acc[0] += vk[i] * w[i];
  block_partials<1>(acc, part + (size_t)blockIdx.y * gridDim.x);
}

// h_kj += sum of the partials (k <= j): column j of the Hessenberg matrix
__global__ void __launch_bounds__(256) k_gm_hsum(const double *part, int nb, int m,
                                                 const GmState *st, double *H, double *hc) {
  __shared__ double sh[256];
  if (!st->live) return;
  const int j = st->j;
  for (int k = 0; k <= j; ++k) {
    const double t = sum_in_order(part + (size_t)k * nb, nb, sh);
    if (threadIdx.x == 0) {
      H[(size_t)j * (m + 1) + k] += t;
      hc[k] = t;
    }
  }
}

// w -= sum_{k <= j} hc_k V_k; partials of w.w
__global__ void __launch_bounds__(KB) k_gm_subtract(int64_t n, const double *V, int64_t ldv,
                                                    const double *hc, double *w,
                                                    const GmState *st, double *part) {
  if (!st->live) return;
  const int j = st->j;
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x; i < n; i += (int64_t)gridDim.x * KB) {
    double wi = w[i];
    for (int k = 0; k <= j; ++k) wi -= hc[k] * V[k * ldv + i];
    w[i] = wi;
    acc[0] += wi * wi;
  }
  block_partials<1>(acc, part);
}

// h_{j+1,j} = |w|; Givens rotations of column j; residual estimate.
__global__ void __launch_bounds__(256) k_gm_givens(const double *part, int nb, int m, GmState *st,
                                                   double *H, double *cs, double *sn, double *g,
                                                   double *hnext) {
  __shared__ double sh[256];
  if (!st->live) return;
  const double ww = sum_in_order(part, nb, sh);
  if (threadIdx.x != 0) return;
  const int j = st->j;
  double *h = H + (size_t)j * (m + 1);
  const double hn = ww > 0.0 ? sqrt(ww) : 0.0;
  h[j + 1] = hn;
  *hnext = hn;
  for (int k = 0; k < j; ++k) {  // earlier rotations on the new column
    const double a = h[k], b = h[k + 1];
    h[k] = cs[k] * a + sn[k] * b;
    h[k + 1] = -sn[k] * a + cs[k] * b;
  }
  const double a = h[j], b = h[j + 1];
  const double r = sqrt(a * a + b * b);
  const double c = r > 0.0 ? a / r : 1.0, s = r > 0.0 ? b / r : 0.0;
  cs[j] = c;
  sn[j] = s;
  h[j] = r;
  h[j + 1] = 0.0;
  g[j + 1] = -s * g[j];
  g[j] = c * g[j];
  st->resid = fabs(g[j + 1]);
  st->j = j + 1;
  // stop the cycle when converged, the basis is full or w vanished
  if (st->resid <= st->tol || j + 1 >= m || hn <= 1e-300 * st->beta) st->live = 0;
}

// v_{j+1} = w / h_{j+1,j} (the step just taken, j = st->j - 1)
__global__ void k_gm_next(int64_t n, const double *w, double *V, int64_t ldv, const GmState *st,
                          const double *hnext) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = st->j;  // index of the new vector
  if (*hnext <= 0.0) return;
  V[j * ldv + i] = w[i] / *hnext;
}

// y = R^{-1} g (upper triangular, k = st->j columns)
__global__ void k_gm_backsolve(const GmState *st, int m, const double *H, const double *g,
                               double *y) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int k = st->j;
  for (int i = k - 1; i >= 0; --i) {
    double t = g[i];
    for (int c = i + 1; c < k; ++c) t -= H[(size_t)c * (m + 1) + i] * y[c];
    const double d = H[(size_t)i * (m + 1) + i];
    y[i] = d != 0.0 ? t / d : 0.0;
  }
}

// x += D^{-1} sum_k y_k V_k
__global__ void __launch_bounds__(KB) k_gm_apply(int64_t n, const double *V, int64_t ldv,
                                                 const double *y, const double *dinv,
                                                 const GmState *st, double *x) {
  const int k = st->j;
  for (int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x; i < n; i += (int64_t)gridDim.x * KB) {
    double t = 0.0;
    for (int c = 0; c < k; ++c) t += y[c] * V[c * ldv + i];
    x[i] += dinv[i] * t;
  }
}

}  // namespace
}  // namespace tfm

using namespace tfm;

struct tfm_krylov {
  int64_t n = 0;
  int m = 0;  // GMRES restart length
  int device = 0;
  double *vec = nullptr;   // work vectors, (m + 1 + 6) * n
  double *dinv = nullptr;  // [n]
  double *part = nullptr;  // partial sums [(m + 1) * KMAXB]
  double *small = nullptr; // Hessenberg, rotations, g, y, hc: (m+1)*m + 5*(m+1) + 1
  void *state = nullptr;   // MrState / GmState
  void *h_state = nullptr; // pinned host mirror of the state
};

namespace {
void krylov_free(tfm_krylov *h) {
  for (double *p : {h->vec, h->dinv, h->part, h->small})
    if (p) (void)hipFree(p);
  if (h->state) (void)hipFree(h->state);
  if (h->h_state) (void)hipHostFree(h->h_state);
}

int krylov_alloc(tfm_krylov *h) {
  const size_t n = (size_t)h->n, m = (size_t)h->m;
  TFM_CHECK_HIP(hipMalloc(&h->vec, sizeof(double) * n * (m + 1 + 6)));
  TFM_CHECK_HIP(hipMalloc(&h->dinv, sizeof(double) * n));
  TFM_CHECK_HIP(hipMalloc(&h->part, sizeof(double) * (m + 2) * KMAXB));
  TFM_CHECK_HIP(hipMalloc(&h->small, sizeof(double) * ((m + 1) * m + 5 * (m + 1) + 1)));
  const size_t ss = std::max(sizeof(MrState), sizeof(GmState));
  TFM_CHECK_HIP(hipMalloc(&h->state, ss));
  TFM_CHECK_HIP(hipHostMalloc(&h->h_state, ss, hipHostMallocDefault));
  return TFM_OK;
}

// ||b - A x|| / ||b|| (and whether it is <= rtol), on the handle's stream
int true_relres(tfm_krylov *h, hipStream_t s, const int64_t *rowptr, const int32_t *col,
                const double *val, const double *b, const double *x, double *relres) {
  const int64_t n = h->n;
  double *r = h->vec + (size_t)(h->m + 1) * n;
  const unsigned g = kgrid(n, KB / KTL);
  hipLaunchKernelGGL(k_residual, dim3(g), dim3(KB), 0, s, n, rowptr, col, val, b, x,
                     (const double *)nullptr, r, (double *)nullptr, h->part);
  TFM_CHECK_LAUNCH();
  std::vector<double> hp(2 * g);
  TFM_CHECK_HIP(hipMemcpyAsync(hp.data(), h->part, sizeof(double) * 2 * g, hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipStreamSynchronize(s));
  double rr = 0.0, bb = 0.0;
  for (unsigned k = 0; k < g; ++k) {
    rr += hp[k];
    bb += hp[g + k];
  }
  *relres = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
  return TFM_OK;
}
}  // namespace

extern "C" int tfm_krylov_create(tfm_krylov_t **out, int64_t n, int32_t restart) {
  TFM_REQUIRE(out != nullptr, "null handle pointer");
  TFM_REQUIRE(n > 0, "system size must be positive");
  TFM_REQUIRE(restart >= 1 && restart <= 512, "restart length must be in [1, 512]");
  auto *h = new tfm_krylov();
  h->n = n;
  h->m = restart;
  (void)hipGetDevice(&h->device);
  const int rc = krylov_alloc(h);
  if (rc != TFM_OK) {
    krylov_free(h);
    delete h;
    return rc;
  }
  *out = h;
  return TFM_OK;
}

extern "C" int tfm_krylov_destroy(tfm_krylov_t *h) {
  if (!h) return TFM_OK;
  krylov_free(h);
  delete h;
  return TFM_OK;
}

extern "C" int tfm_csr_minres_solve(tfm_krylov_t *h, tfm_stream_t stream, int64_t n,
                                    const int64_t *rowptr, const int32_t *col,
                                    const double *val, const double *b, double *x, double rtol,
                                    int32_t max_iter, int32_t *iters, double *relres) {
  TFM_REQUIRE(h && rowptr && col && val && b && x && iters && relres, "null argument");
  TFM_REQUIRE(n == h->n, "system size differs from the solver handle");
  TFM_REQUIRE(rtol > 0.0 && max_iter >= 1, "rtol and max_iter must be positive");
  hipStream_t s = (hipStream_t)stream;
  const unsigned g1 = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_inv_absdiag, dim3(g1), dim3(256), 0, s, n, rowptr, col, val, h->dinv);
  TFM_CHECK_LAUNCH();
  // vectors: r1, r2, y, q, v, w (three buffers)
  double *vb = h->vec;
  double *r1 = vb, *r2 = vb + n, *y = vb + 2 * n, *q = vb + 3 * n, *v = vb + 4 * n;
  double *w[3] = {vb + 5 * n, vb + 6 * n, vb + 7 * n};
  TFM_REQUIRE(h->m + 1 + 6 >= 8, "MINRES needs a restart length of at least 1");
  MrState *st = (MrState *)h->state;
  MrState *hs = (MrState *)h->h_state;
  const unsigned gr = kgrid(n, KB / KTL), gv = kgrid(n, KB);
  *iters = 0;
  // restarts from the current x if the true residual misses rtol
  for (int cycle = 0; cycle < 8; ++cycle) {
    hipLaunchKernelGGL(k_residual, dim3(gr), dim3(KB), 0, s, n, rowptr, col, val, b,
                       (const double *)x, (const double *)h->dinv, r1, y, h->part);
    TFM_CHECK_LAUNCH();
    TFM_CHECK_HIP(hipMemcpyAsync(r2, r1, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    for (double *wk : w) TFM_CHECK_HIP(hipMemsetAsync(wk, 0, sizeof(double) * n, s));
    hipLaunchKernelGGL(k_mr_start, dim3(1), dim3(256), 0, s, (const double *)h->part, (int)gr,
                       rtol, st);
    TFM_CHECK_LAUNCH();
    // w roles: wa = w_{k-2}, wb = w_{k-1}; the new w goes to wa's buffer
    int ia = 0, ib = 1;
    double *ra = r1, *rb = r2;  // r1 (previous), r2 (current)
    int live = 1, done_it = 0;
    while (live) {
      for (int k = 0; k < KCHK; ++k) {
        hipLaunchKernelGGL(k_mr_lanczos, dim3(gr), dim3(KB), 0, s, n, rowptr, col, val,
                           (const double *)y, (const double *)ra, v, q, (const MrState *)st,
                           h->part);
        TFM_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_mr_alfa, dim3(1), dim3(256), 0, s, (const double *)h->part, (int)gr, st);
        TFM_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_mr_orth, dim3(gv), dim3(KB), 0, s, n, (const double *)q,
                           (const double *)rb, ra, (const double *)h->dinv, y,
                           (const MrState *)st, h->part);
        TFM_CHECK_LAUNCH();
        std::swap(ra, rb);  // the new residual vector is in the old r1 buffer
        hipLaunchKernelGGL(k_mr_rotate, dim3(1), dim3(256), 0, s, (const double *)h->part, (int)gv, st);
        TFM_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_mr_update, dim3(gv), dim3(KB), 0, s, n, (const double *)v, w[ia],
                           (const double *)w[ib], x, st);
        TFM_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_mr_check, dim3(1), dim3(1), 0, s, st, (int)max_iter);
        TFM_CHECK_LAUNCH();
        // the new w (in w[ia]) becomes w_{k-1}; the old w_{k-1} becomes w_{k-2}
        std::swap(ia, ib);
      }
      TFM_CHECK_HIP(hipMemcpyAsync(hs, st, sizeof(MrState), hipMemcpyDeviceToHost, s));
      TFM_CHECK_HIP(hipStreamSynchronize(s));
      live = hs->live;
      done_it = hs->it;
      if (!std::isfinite(hs->phibar)) {
        set_error("non-finite residual in MINRES");
        return TFM_ERR_NOCONV;
      }
    }
    *iters += done_it;
    int rc = true_relres(h, s, rowptr, col, val, b, x, relres);
    if (rc != TFM_OK) return rc;
    if (*relres <= rtol) return TFM_OK;
    if (*iters >= max_iter) break;
  }
  set_error("minres did not reach the tolerance");
  return TFM_ERR_NOCONV;
}

extern "C" int tfm_csr_gmres_solve(tfm_krylov_t *h, tfm_stream_t stream, int64_t n,
                                   const int64_t *rowptr, const int32_t *col, const double *val,
                                   const double *b, double *x, double rtol, int32_t max_iter,
                                   int32_t *iters, double *relres) {
  TFM_REQUIRE(h && rowptr && col && val && b && x && iters && relres, "null argument");
  TFM_REQUIRE(n == h->n, "system size differs from the solver handle");
  TFM_REQUIRE(rtol > 0.0 && max_iter >= 1, "rtol and max_iter must be positive");
  hipStream_t s = (hipStream_t)stream;
  const int m = h->m;
  const int64_t ldv = n;
  double *V = h->vec;                        // m + 1 basis vectors
  double *w = h->vec + (size_t)(m + 1) * n;  // work vector (also the residual)
  double *H = h->small;
  double *cs = H + (size_t)(m + 1) * m, *sn = cs + (m + 1), *g = sn + (m + 1), *y = g + (m + 1);
  double *hc = y + (m + 1), *hnext = hc + (m + 1);
  GmState *st = (GmState *)h->state;
  GmState *hs = (GmState *)h->h_state;
  const unsigned g1 = (unsigned)((n + 255) / 256);
  const unsigned gr = kgrid(n, KB / KTL), gv = kgrid(n, KB);
  hipLaunchKernelGGL(k_inv_absdiag, dim3(g1), dim3(256), 0, s, n, rowptr, col, val, h->dinv);
  TFM_CHECK_LAUNCH();
  // b.b for the tolerance
  double rel0 = 0.0;
  int rc = true_relres(h, s, rowptr, col, val, b, x, &rel0);
  if (rc != TFM_OK) return rc;
  *iters = 0;
  *relres = rel0;
  if (rel0 <= rtol) return TFM_OK;
  // ||b|| from the partials of the residual kernel (second set)
  std::vector<double> hp(2 * gr);
  TFM_CHECK_HIP(hipMemcpy(hp.data(), h->part, sizeof(double) * 2 * gr, hipMemcpyDeviceToHost));
  double bb = 0.0;
  for (unsigned k = 0; k < gr; ++k) bb += hp[gr + k];
  const double tol = rtol * std::sqrt(bb) * 0.5;  // margin below rtol for the true residual
  while (*iters < max_iter) {
    hipLaunchKernelGGL(k_residual, dim3(gr), dim3(KB), 0, s, n, rowptr, col, val, b,
                       (const double *)x, (const double *)nullptr, w, (double *)nullptr, h->part);
    TFM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gm_start, dim3(1), dim3(256), 0, s, (const double *)h->part, (int)gr,
                       tol > 0.0 ? tol : 0.0, m, st, g, H);
    TFM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gm_first, dim3(g1), dim3(256), 0, s, n, (const double *)w, V,
                       (const GmState *)st);
    TFM_CHECK_LAUNCH();
    int live = 1;
    while (live) {
      for (int k = 0; k < KCHK; ++k) {
        // the step uses v_j with j = st->j, read on the device
        hipLaunchKernelGGL(k_gm_product, dim3(gr), dim3(KB), 0, s, n, rowptr, col, val,
                           (const double *)V, ldv, (const double *)h->dinv, w,
                           (const GmState *)st);
        TFM_CHECK_LAUNCH();
        for (int pass = 0; pass < 2; ++pass) {  // Gram-Schmidt and one reorthogonalisation
          hipLaunchKernelGGL(k_gm_dots, dim3(gv, m), dim3(KB), 0, s, n, (const double *)V,
                             (const double *)w, ldv, (const GmState *)st, h->part);
          TFM_CHECK_LAUNCH();
          hipLaunchKernelGGL(k_gm_hsum, dim3(1), dim3(256), 0, s, (const double *)h->part, (int)gv,
                             m, (const GmState *)st, H, hc);
          TFM_CHECK_LAUNCH();
          hipLaunchKernelGGL(k_gm_subtract, dim3(gv), dim3(KB), 0, s, n, (const double *)V, ldv,
                             (const double *)hc, w, (const GmState *)st, h->part);
          TFM_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(256), 0, s, (const double *)h->part, (int)gv,
                           m, st, H, cs, sn, g, hnext);
        TFM_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_gm_next, dim3(g1), dim3(256), 0, s, n, (const double *)w, V, ldv,
                           (const GmState *)st, (const double *)hnext);
        TFM_CHECK_LAUNCH();
      }
      TFM_CHECK_HIP(hipMemcpyAsync(hs, st, sizeof(GmState), hipMemcpyDeviceToHost, s));
      TFM_CHECK_HIP(hipStreamSynchronize(s));
      live = hs->live;
      if (!std::isfinite(hs->resid)) {
        set_error("non-finite residual in GMRES");
        return TFM_ERR_NOCONV;
      }
    }
    hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(64), 0, s, (const GmState *)st, m,
                       (const double *)H, (const double *)g, y);
    TFM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gm_apply, dim3(gv), dim3(KB), 0, s, n, (const double *)V, ldv,
                       (const double *)y, (const double *)h->dinv, (const GmState *)st, x);
    TFM_CHECK_LAUNCH();
    *iters += hs->j;
    const double prev = *relres;
    rc = true_relres(h, s, rowptr, col, val, b, x, relres);
    if (rc != TFM_OK) return rc;
    if (*relres <= rtol) return TFM_OK;
    // a cycle that no longer reduces the true residual has reached the
    // attainable accuracy of FP64 arithmetic for this matrix
    if (hs->j == 0 || *relres > 0.99 * prev) {
      set_error("GMRES stagnated above the tolerance (attainable accuracy reached)");
      return TFM_ERR_NOCONV;
    }
  }
  set_error("GMRES did not reach the tolerance within max_iter iterations");
  return TFM_ERR_NOCONV;
}
