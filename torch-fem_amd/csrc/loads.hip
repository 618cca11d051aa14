// Distributed loads and shape-function evaluation for gfx950: boundary
// facets of a nodal selection, consistent facet (surface / line) loads, body
// loads, and the shape gradients B = J^{-1} dN/dxi at arbitrary local points.
//
// Boundary facets are found without sorting or atomics. A facet of element e
// lies on the boundary when no other element incident to its smallest node
// holds all of its nodes. Its place in the output follows the order of the
// facets' sorted node tuples (the order torch.unique gives the reference's
// facet list, base.py:452-470): facets are bucketed by their smallest node
// (a prefix sum over the nodes), and inside a bucket a facet's rank is the
// number of bucket members with a smaller tuple. Facet contributions are then
// summed at the nodes by a gather over the node's incident elements in
// ascending order, so every result is bitwise reproducible.
#include "tfm_common.h"

namespace tfm {
namespace {

// ------------------------------------------------------------ facet types
// Isoparametric facets (elements.py: Bar1 172-213, Tria1 310-360, Quad1
// 491-567, Quad2 599-695) with their integration rules.
struct FBar1 {
  static constexpr int NFN = 2, IDIM = 1, NGP = 1;
  __device__ static void gauss(int, double *xi, double &w) { xi[0] = 0.0; w = 2.0; }
  __device__ static void shape(const double *xi, double *N, double (*dN)[8]) {
    N[0] = 0.5 * (1.0 - xi[0]);
    N[1] = 0.5 * (1.0 + xi[0]);
    dN[0][0] = -0.5;
    dN[0][1] = 0.5;
  }
};

struct FTria1 {
  static constexpr int NFN = 3, IDIM = 2, NGP = 1;
  __device__ static void gauss(int, double *xi, double &w) {
    xi[0] = xi[1] = 1.0 / 3.0;
    w = 0.5;
  }
  __device__ static void shape(const double *xi, double *N, double (*dN)[8]) {
    N[0] = 1.0 - xi[0] - xi[1];
    N[1] = xi[0];
    N[2] = xi[1];
    dN[0][0] = -1.0; dN[0][1] = 1.0; dN[0][2] = 0.0;
    dN[1][0] = -1.0; dN[1][1] = 0.0; dN[1][2] = 1.0;
  }
};

// Quadrilateral facets: bilinear (4 nodes) or serendipity (8 nodes) on the
// reference square, 2x2 Gauss points g = i1 + 2*i2 with weight 1.
template <int NN> struct FQuad {
  static constexpr int NFN = NN, IDIM = 2, NGP = 4;
  __device__ static void gauss(int g, double *xi, double &w) {
    xi[0] = (g & 1) ? kGauss : -kGauss;
    xi[1] = (g & 2) ? kGauss : -kGauss;
    w = 1.0;
  }
  __device__ static void shape(const double *xi, double *N, double (*dN)[8]) {
    constexpr double c[8][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1},
                                {0, -1},  {1, 0},  {0, 1}, {-1, 0}};
    for (int I = 0; I < NN; ++I) {
      const double fa = axis_f(xi[0], c[I][0]), fb = axis_f(xi[1], c[I][1]);
      const double da = axis_df(xi[0], c[I][0]), db = axis_df(xi[1], c[I][1]);
      if (NN == 4) {
        N[I] = 0.25 * fa * fb;
        dN[0][I] = 0.25 * da * fb;
        dN[1][I] = 0.25 * fa * db;
      } else if (I < 4) {  // serendipity vertex: (1/4) fa fb (x c0 + y c1 - 1)
        const double s = xi[0] * c[I][0] + xi[1] * c[I][1] - 1.0;
        N[I] = 0.25 * fa * fb * s;
        dN[0][I] = 0.25 * (da * fb * s + fa * fb * c[I][0]);
        dN[1][I] = 0.25 * (fa * db * s + fa * fb * c[I][1]);
      } else {  // mid-side: (1/2) (1 - x^2) on the axis with c = 0
        N[I] = 0.5 * fa * fb;
        dN[0][I] = 0.5 * da * fb;
        dN[1][I] = 0.5 * fa * db;
      }
    }
  }
};

// Facet tables of the element types: local node lists in the winding whose
// normal J0 x J1 (or (J0_y, -J0_x) for edges) points out of the element
// (elements.py: Quad1 facets, Tetra1 facets, Hexa1/Hexa2 facets).
template <class E> struct Facets;
template <> struct Facets<Quad1> {
  using F = FBar1;
  static constexpr int NF = 4;
  static constexpr int t[4][8] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
};
template <> struct Facets<Tetra1> {
  using F = FTria1;
  static constexpr int NF = 4;
  static constexpr int t[4][8] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {0, 3, 2}};
};
template <> struct Facets<Hexa1> {
  using F = FQuad<4>;
  static constexpr int NF = 6;
  static constexpr int t[6][8] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                  {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
};
template <> struct Facets<Hexa2> {
  using F = FQuad<8>;
  static constexpr int NF = 6;
  static constexpr int t[6][8] = {{0, 3, 2, 1, 11, 10, 9, 8},  {4, 5, 6, 7, 12, 13, 14, 15},
                                  {0, 1, 5, 4, 8, 17, 12, 16}, {1, 2, 6, 5, 9, 18, 13, 17},
                                  {2, 3, 7, 6, 10, 19, 14, 18}, {3, 0, 4, 7, 11, 16, 15, 19}};
};

// Node ids of facet lf of element e, and their sorted tuple.
template <class E>
__device__ __forceinline__ void facet_nodes(const tfm_mesh_t &m, int64_t e, int lf, int *g) {
  using FS = Facets<E>;
  const int32_t *en = m.elements + e * E::NPE;
#pragma unroll
  for (int i = 0; i < FS::F::NFN; ++i) g[i] = en[FS::t[lf][i]];
}

template <int K> __device__ __forceinline__ void sort_small(int *v) {
#pragma unroll
  for (int i = 1; i < K; ++i)
#pragma unroll
    for (int j = i; j > 0; --j)
      if (v[j] < v[j - 1]) {
        const int t = v[j];
        v[j] = v[j - 1];
        v[j - 1] = t;
      }
}

template <int K> __device__ __forceinline__ bool tuple_less(const int *a, const int *b) {
#pragma unroll
  for (int i = 0; i < K; ++i)
    if (a[i] != b[i]) return a[i] < b[i];
  return false;
}

template <class E> __device__ __forceinline__ bool elem_has(const int32_t *en, int node) {
  bool hit = false;
#pragma unroll
  for (int j = 0; j < E::NPE; ++j) hit |= en[j] == node;
  return hit;
}

// Slot s = e*NF + lf: flag[s] = 1 if the facet's nodes are all selected and
// no other element holds all of them; pos[s] = -1 (set later for flagged ones).
template <class E>
__global__ void k_facet_flag(tfm_mesh_t m, const uint8_t *mask, int8_t *flag, int32_t *pos) {
  using FS = Facets<E>;
  constexpr int NFN = FS::F::NFN;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m.n_elem * FS::NF) return;
  const int64_t e = s / FS::NF;
  const int lf = (int)(s - e * FS::NF);
  int g[NFN];
  facet_nodes<E>(m, e, lf, g);
  pos[s] = -1;
  bool sel = true;
  int lo = g[0];
#pragma unroll
  for (int i = 0; i < NFN; ++i) {
    sel &= mask[g[i]] != 0;
    lo = min(lo, g[i]);
  }
  bool shared = false;
  if (sel) {
    for (int k = m.ne_ptr[lo]; k < m.ne_ptr[lo + 1] && !shared; ++k) {
      const int64_t o = m.ne_elem[k];
      if (o == e) continue;
      const int32_t *on = m.elements + o * E::NPE;
      bool all = true;
#pragma unroll
      for (int i = 0; i < NFN; ++i) all &= elem_has<E>(on, g[i]);
      shared = all;
    }
  }
  flag[s] = (sel && !shared) ? 1 : 0;
}

// Visits the flagged facets whose smallest node is n, over n's incident
// elements in ascending order: fn(slot, sorted tuple).
template <class E, class Fn>
__device__ __forceinline__ void for_bucket(const tfm_mesh_t &m, const int8_t *flag, int64_t n,
                                           Fn &&fn) {
  using FS = Facets<E>;
  constexpr int NFN = FS::F::NFN;
  for (int k = m.ne_ptr[n]; k < m.ne_ptr[n + 1]; ++k) {
    const int64_t e = m.ne_elem[k];
    const int loc = m.ne_loc[k];
    for (int lf = 0; lf < FS::NF; ++lf) {
      bool has = false;
#pragma unroll
      for (int i = 0; i < NFN; ++i) has |= FS::t[lf][i] == loc;
      if (!has) continue;
      int g[NFN];
      facet_nodes<E>(m, e, lf, g);
      sort_small<NFN>(g);
      if (g[0] != n) continue;  // the facet belongs to its smallest node's bucket
      const int64_t s = e * FS::NF + lf;
      if (flag[s]) fn(s, g);
    }
  }
}

template <class E>
__global__ void k_facet_count(tfm_mesh_t m, const uint8_t *mask, const int8_t *flag,
                              int32_t *cnt) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= m.n_nod) return;
  int c = 0;
  if (mask[n]) for_bucket<E>(m, flag, n, [&](int64_t, const int *) { ++c; });
  cnt[n] = c;
}

// Rank of every facet of bucket n among the bucket's tuples; writes its slot
// to list[off[n] + rank] and the position to pos[slot].
template <class E>
__global__ void k_facet_rank(tfm_mesh_t m, const uint8_t *mask, const int8_t *flag,
                             const int32_t *off, int32_t *pos, int32_t *list) {
  constexpr int NFN = Facets<E>::F::NFN;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= m.n_nod || !mask[n]) return;
  if (off[n + 1] == off[n]) return;
  for_bucket<E>(m, flag, n, [&](int64_t s, const int *key) {
    int rank = 0;
    int kk[NFN];
#pragma unroll
    for (int i = 0; i < NFN; ++i) kk[i] = key[i];
    for_bucket<E>(m, flag, n, [&](int64_t, const int *other) {
      rank += tuple_less<NFN>(other, kk) ? 1 : 0;
    });
    const int p = off[n] + rank;
    pos[s] = p;
    list[p] = (int32_t)s;
  });
}

// ---------------------------------------------------------------- scan
// Exclusive prefix sum of cnt[0..n) into off[0..n]; three passes over blocks
// of SB counts (block sums, a one-block scan of those, then the offsets).
constexpr int SB = 1024;

__global__ void __launch_bounds__(256) k_scan_blocks(const int32_t *cnt, int64_t n, int32_t *bsum) {
  __shared__ int32_t red[256];
  const int64_t b0 = (int64_t)blockIdx.x * SB;
  int32_t acc = 0;
  for (int i = threadIdx.x; i < SB; i += 256)
    if (b0 + i < n) acc += cnt[b0 + i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

// One block: bsum[0..nb) -> exclusive offsets (in place), total to bsum[nb].
__global__ void __launch_bounds__(1024) k_scan_top(int32_t *bsum, int nb) {
  __shared__ int32_t tmp[1024];
  int32_t carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const int32_t v = i < nb ? bsum[i] : 0;
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // Hillis-Steele inclusive scan
      const int32_t add = (int)threadIdx.x >= d ? tmp[threadIdx.x - d] : 0;
      __syncthreads();
      tmp[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + tmp[threadIdx.x] - v;
    const int32_t total = tmp[1023];
    __syncthreads();
    carry += total;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

// Thread t owns counts [4t, 4t+4) of the block: their sum enters an LDS scan
// over the 256 threads, then the thread writes its four offsets. In place
// (off may alias cnt): every count is read before the block writes.
__global__ void __launch_bounds__(256) k_scan_apply(const int32_t *cnt, int64_t n,
                                                    const int32_t *bsum, int32_t *off) {
  __shared__ int32_t part[256];
  const int64_t b0 = (int64_t)blockIdx.x * SB + 4 * threadIdx.x;
  int32_t v[4], sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] = b0 + i < n ? cnt[b0 + i] : 0;
    sum += v[i];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {  // inclusive Hillis-Steele scan
    const int32_t add = (int)threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  int32_t run = bsum[blockIdx.x] + part[threadIdx.x] - sum;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (b0 + i < n) off[b0 + i] = run;
    run += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) off[n] = bsum[gridDim.x];
}

// ------------------------------------------------------- facet integration
// load kinds: 0 pressure along the outward normal (scalar), 1 one vector for
// every facet, 2 one vector per facet [n_facets, K].
template <class E, int K>
__global__ void k_facet_integrate(tfm_mesh_t m, const int32_t *list, int64_t nfac,
                                  const double *load, int kind, double *fc) {
  using FS = Facets<E>;
  using F = typename FS::F;
  constexpr int NFN = F::NFN, D = E::DIM;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nfac) return;
  const int64_t s = list[f];
  const int64_t e = s / FS::NF;
  const int lf = (int)(s - e * FS::NF);
  int g[NFN];
  facet_nodes<E>(m, e, lf, g);
  double X[NFN][D];
#pragma unroll
  for (int i = 0; i < NFN; ++i)
#pragma unroll
    for (int d = 0; d < D; ++d) X[i][d] = m.nodes[(int64_t)g[i] * D + d];
  double acc[NFN][K];
#pragma unroll
  for (int i = 0; i < NFN; ++i)
#pragma unroll
    for (int c = 0; c < K; ++c) acc[i][c] = 0.0;
  for (int q = 0; q < F::NGP; ++q) {
    double xi[2] = {0.0, 0.0}, w = 0.0, N[8], dN[2][8];
    F::gauss(q, xi, w);
    F::shape(xi, N, dN);
    double J[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
    for (int a = 0; a < F::IDIM; ++a)
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < NFN; ++i) v += dN[a][i] * X[i][d];
        J[a][d] = v;
      }
    // measure sqrt(det(J J^T)) and the (unnormalised) normal
    double nrm[3] = {0.0, 0.0, 0.0}, meas;
    if (F::IDIM == 2) {
      nrm[0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      nrm[1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
      nrm[2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const double a2 = J[0][0] * J[0][0] + J[0][1] * J[0][1] + J[0][2] * J[0][2];
      const double b2 = J[1][0] * J[1][0] + J[1][1] * J[1][1] + J[1][2] * J[1][2];
      const double ab = J[0][0] * J[1][0] + J[0][1] * J[1][1] + J[0][2] * J[1][2];
      meas = sqrt(a2 * b2 - ab * ab);
    } else {
      nrm[0] = J[0][1];
      nrm[1] = -J[0][0];
      meas = sqrt(J[0][0] * J[0][0] + J[0][1] * J[0][1]);
    }
    double val[K];
    if (kind == 0) {
      const double len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      const double sc = load[0] / (len > 1e-12 ? len : 1e-12);  // normalize() clamps at 1e-12
#pragma unroll
      for (int c = 0; c < K; ++c) val[c] = sc * nrm[c];
    } else {
      const double *lv = load + (kind == 2 ? f * K : 0);
#pragma unroll
      for (int c = 0; c < K; ++c) val[c] = lv[c];
    }
    const double wm = w * meas;
#pragma unroll
    for (int i = 0; i < NFN; ++i)
#pragma unroll
      for (int c = 0; c < K; ++c) acc[i][c] += wm * N[i] * val[c];
  }
  double *out = fc + f * NFN * K;
#pragma unroll
  for (int i = 0; i < NFN; ++i)
#pragma unroll
    for (int c = 0; c < K; ++c) out[i * K + c] = acc[i][c];
}

// F[n][c] = sum over node n's incident elements (ascending), and their loaded
// facets (ascending local index), of the facet contribution at node n.
template <class E, int K>
__global__ void k_facet_gather(tfm_mesh_t m, const uint8_t *mask, const int32_t *pos,
                               const double *fc, double *F) {
  using FS = Facets<E>;
  constexpr int NFN = FS::F::NFN;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= m.n_nod) return;
  double acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = 0.0;
  if (mask[n]) {
    for (int k = m.ne_ptr[n]; k < m.ne_ptr[n + 1]; ++k) {
      const int64_t e = m.ne_elem[k];
      const int loc = m.ne_loc[k];
      for (int lf = 0; lf < FS::NF; ++lf) {
        int at = -1;
#pragma unroll
        for (int i = 0; i < NFN; ++i) at = FS::t[lf][i] == loc ? i : at;
        if (at < 0) continue;
        const int p = pos[e * FS::NF + lf];
        if (p < 0) continue;
#pragma unroll
        for (int c = 0; c < K; ++c) acc[c] += fc[((int64_t)p * NFN + at) * K + c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < K; ++c) F[n * K + c] = acc[c];
}

template <class E>
__global__ void k_facet_nodes(tfm_mesh_t m, const int32_t *list, int64_t nfac, int64_t *out) {
  using FS = Facets<E>;
  constexpr int NFN = FS::F::NFN;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nfac) return;
  const int64_t s = list[f];
  int g[NFN];
  facet_nodes<E>(m, s / FS::NF, (int)(s % FS::NF), g);
#pragma unroll
  for (int i = 0; i < NFN; ++i) out[f * NFN + i] = g[i];
}

// ----------------------------------------------------- element shape values
// N_I(xi) of the element families: tensor-product / serendipity functions from
// the nodal reference coordinates, barycentric coordinates for the tetrahedron.
template <class E> __device__ __forceinline__ void shape_values(const double *xi, double *N) {
  if constexpr (E::ET == TFM_TETRA1) {
    N[0] = 1.0 - xi[0] - xi[1] - xi[2];
    N[1] = xi[0];
    N[2] = xi[1];
    N[3] = xi[2];
  } else {
    for (int I = 0; I < E::NPE; ++I) {
      double p = 1.0, s = 0.0;
      bool vertex = true;
      for (int d = 0; d < E::DIM; ++d) {
        p *= axis_f(xi[d], E::iso[I][d]);
        s += xi[d] * E::iso[I][d];
        vertex &= E::iso[I][d] != 0.0;
      }
      if (E::ET == TFM_HEXA2)
        N[I] = vertex ? 0.125 * p * (s - 2.0) : 0.25 * p;
      else
        N[I] = p / (double)(1 << E::DIM);
    }
  }
}

// Body load per element: fe[e][I][c] = sum_gp w N_I detJ scale_e load_e[c].
template <class E, int K>
__global__ void k_body_load(tfm_mesh_t m, const double *scale, const double *load,
                            int per_elem, double *fe, int32_t *bad) {
  constexpr int D = E::DIM;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.n_elem) return;
  double X[E::NPE][D];
  for (int i = 0; i < E::NPE; ++i) {
    const int64_t nd = m.elements[e * E::NPE + i];
    for (int d = 0; d < D; ++d) X[i][d] = m.nodes[nd * D + d];
  }
  double wn[E::NPE];
  for (int i = 0; i < E::NPE; ++i) wn[i] = 0.0;
  for (int g = 0; g < E::NGP; ++g) {
    double xi[3] = {0.0, 0.0, 0.0}, w = 0.0, N[E::NPE], dN[D][E::NPE];
    E::gauss(g, xi, w);
    E::dshape(xi, dN);
    shape_values<E>(xi, N);
    double J[D][D], Ji[D][D];
    for (int i = 0; i < D; ++i)
      for (int k = 0; k < D; ++k) {
        double acc = 0.0;
        for (int I = 0; I < E::NPE; ++I) acc += dN[i][I] * X[I][k];
        J[i][k] = acc;
      }
    const double det = inv_jac<D>(J, Ji);
    if (!(det > 0.0)) *bad = 1;
    for (int i = 0; i < E::NPE; ++i) wn[i] += w * N[i] * det;
  }
  const double sc = scale ? scale[e] : 1.0;
  const double *lv = load + (per_elem ? e * K : 0);
  for (int i = 0; i < E::NPE; ++i)
    for (int c = 0; c < K; ++c) fe[(e * E::NPE + i) * K + c] = wn[i] * sc * lv[c];
}

// Consistent element mass matrices (base.py:376-396): one 64-lane block per
// element stages N_I and w detJ rho_e scale_e at every Gauss point in LDS,
// then m[(I,a),(J,b)] = delta_ab sum_g w detJ rho scale N_I N_J.
template <class E, int ND>
__global__ void __launch_bounds__(64) k_element_mass(tfm_mesh_t m, const double *rho,
                                                     const double *scale, double *mout,
                                                     int32_t *bad) {
  constexpr int NPE = E::NPE, NGP = E::NGP, D = E::DIM, W = NPE * ND;
  __shared__ double X[NPE][D];
  __shared__ double Nv[NGP][NPE];
  __shared__ double wd[NGP];
  const int64_t e = xcd_block();
  const int tid = threadIdx.x;
  for (int t = tid; t < NPE * D; t += 64) {
    const int j = t / D, d = t - j * D;
    X[j][d] = m.nodes[(int64_t)m.elements[e * NPE + j] * D + d];
  }
  __syncthreads();
  for (int g = tid; g < NGP; g += 64) {
    double xi[3] = {0.0, 0.0, 0.0}, w = 0.0, N[NPE], dN[D][NPE];
    E::gauss(g, xi, w);
    E::dshape(xi, dN);
    shape_values<E>(xi, N);
    double J[D][D], Ji[D][D];
    for (int i = 0; i < D; ++i)
      for (int k = 0; k < D; ++k) {
        double acc = 0.0;
        for (int I = 0; I < NPE; ++I) acc += dN[i][I] * X[I][k];
        J[i][k] = acc;
      }
    const double det = inv_jac<D>(J, Ji);
    if (!(det > 0.0)) *bad = 1;
    for (int I = 0; I < NPE; ++I) Nv[g][I] = N[I];
    wd[g] = w * det * rho[e] * (scale ? scale[e] : 1.0);
  }
  __syncthreads();
  for (int t = tid; t < W * W; t += 64) {
    const int r = t / W, c = t - r * W;
    const int I = r / ND, a = r - I * ND, J = c / ND, b = c - J * ND;
    double v = 0.0;
    if (a == b)
      for (int g = 0; g < NGP; ++g) v += wd[g] * Nv[g][I] * Nv[g][J];
    mout[e * W * W + t] = v;
  }
}

// B[q][e][k][I] = (J^{-1} dN/dxi)[k][I] and detJ[q][e] at local points xi[q]
// (the quantities of base.py:293-314 at arbitrary points).
template <class E>
__global__ void k_shape_gradients(tfm_mesh_t m, const double *xi, int nq, double *B,
                                  double *detJ, int32_t *bad) {
  constexpr int D = E::DIM;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nq * m.n_elem) return;
  const int q = (int)(t / m.n_elem);
  const int64_t e = t - (int64_t)q * m.n_elem;
  double x[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < D; ++d) x[d] = xi[q * D + d];
  double dN[D][E::NPE];
  E::dshape(x, dN);
  double J[D][D], Ji[D][D];
  for (int i = 0; i < D; ++i)
    for (int k = 0; k < D; ++k) J[i][k] = 0.0;
  for (int I = 0; I < E::NPE; ++I) {
    const int64_t nd = m.elements[e * E::NPE + I];
    for (int k = 0; k < D; ++k) {
      const double xk = m.nodes[nd * D + k];
      for (int i = 0; i < D; ++i) J[i][k] += dN[i][I] * xk;
    }
  }
  const double det = inv_jac<D>(J, Ji);
  if (!(det > 0.0)) *bad = 1;
  detJ[t] = det;
  double *b = B + t * D * E::NPE;
  for (int k = 0; k < D; ++k)
    for (int I = 0; I < E::NPE; ++I) {
      double acc = 0.0;
      for (int i = 0; i < D; ++i) acc += Ji[k][i] * dN[i][I];
      b[k * E::NPE + I] = acc;
    }
}

template <class Fn> int with_etype(int etype, Fn &&fn) {
  switch (etype) {
    case TFM_QUAD1: return fn(Quad1{});
    case TFM_TETRA1: return fn(Tetra1{});
    case TFM_HEXA1: return fn(Hexa1{});
    case TFM_HEXA2: return fn(Hexa2{});
    default: set_error("unknown element type"); return TFM_ERR_ARG;
  }
}

template <class Fn> int with_k(int k, Fn &&fn) {
  switch (k) {
    case 1: return fn(std::integral_constant<int, 1>{});
    case 2: return fn(std::integral_constant<int, 2>{});
    case 3: return fn(std::integral_constant<int, 3>{});
    default: set_error("loads must have 1, 2 or 3 components"); return TFM_ERR_ARG;
  }
}

unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace
}  // namespace tfm

using namespace tfm;

extern "C" int tfm_facet_count(const tfm_mesh_t *m, int32_t *nf, int32_t *nfn) {
  TFM_REQUIRE(m && nf && nfn, "null argument");
  switch (m->etype) {
    case TFM_QUAD1: *nf = 4; *nfn = 2; return TFM_OK;
    case TFM_TETRA1: *nf = 4; *nfn = 3; return TFM_OK;
    case TFM_HEXA1: *nf = 6; *nfn = 4; return TFM_OK;
    case TFM_HEXA2: *nf = 6; *nfn = 8; return TFM_OK;
    default: TFM_REQUIRE(false, "unknown element type");
  }
}

extern "C" int tfm_boundary_facets(tfm_stream_t stream, const tfm_mesh_t *m,
                                   const uint8_t *mask, int8_t *flag, int32_t *pos,
                                   int32_t *off, int32_t *bsum, int32_t *list,
                                   int64_t list_cap, int64_t *n_facets) {
  TFM_REQUIRE(m && mask && flag && pos && off && bsum && list && n_facets, "null argument");
  TFM_REQUIRE(m->n_nod >= 0 && m->n_elem >= 0, "bad mesh sizes");
  hipStream_t s = (hipStream_t)stream;
  *n_facets = 0;
  if (m->n_elem == 0 || m->n_nod == 0) return TFM_OK;
  int32_t nf = 0, nfn = 0;
  int rc = tfm_facet_count(m, &nf, &nfn);
  if (rc != TFM_OK) return rc;
  const int64_t nslots = m->n_elem * nf;
  const unsigned nb = (unsigned)((m->n_nod + SB - 1) / SB);
  rc = with_etype(m->etype, [&](auto et) -> int {
    using E = decltype(et);
    hipLaunchKernelGGL(k_facet_flag<E>, dim3(blocks_for(nslots)), dim3(256), 0, s, *m, mask, flag, pos);
    TFM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_facet_count<E>, dim3(blocks_for(m->n_nod)), dim3(256), 0, s, *m, mask,
                       (const int8_t *)flag, off);
    TFM_CHECK_LAUNCH();
    return TFM_OK;
  });
  if (rc != TFM_OK) return rc;
  // off holds the counts; scan them in place through bsum
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, s, (const int32_t *)off, m->n_nod, bsum);
  TFM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, bsum, (int)nb);
  TFM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, s, (const int32_t *)off, m->n_nod,
                     (const int32_t *)bsum, off);
  TFM_CHECK_LAUNCH();
  int32_t total = 0;
  TFM_CHECK_HIP(hipMemcpyAsync(&total, off + m->n_nod, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipStreamSynchronize(s));
  *n_facets = total;
  TFM_REQUIRE(total <= list_cap, "facet list capacity too small");
  return with_etype(m->etype, [&](auto et) -> int {
    using E = decltype(et);
    hipLaunchKernelGGL(k_facet_rank<E>, dim3(blocks_for(m->n_nod)), dim3(256), 0, s, *m, mask,
                       (const int8_t *)flag, (const int32_t *)off, pos, list);
    TFM_CHECK_LAUNCH();
    return TFM_OK;
  });
}

extern "C" int tfm_facet_nodes(tfm_stream_t stream, const tfm_mesh_t *m, const int32_t *list,
                               int64_t n_facets, int64_t *out) {
  TFM_REQUIRE(m && (n_facets == 0 || (list && out)), "null argument");
  if (n_facets == 0) return TFM_OK;
  hipStream_t s = (hipStream_t)stream;
  return with_etype(m->etype, [&](auto et) -> int {
    using E = decltype(et);
    hipLaunchKernelGGL(k_facet_nodes<E>, dim3(blocks_for(n_facets)), dim3(256), 0, s, *m, list,
                       n_facets, out);
    TFM_CHECK_LAUNCH();
    return TFM_OK;
  });
}

extern "C" int tfm_facet_load(tfm_stream_t stream, const tfm_mesh_t *m, const uint8_t *mask,
                              const int32_t *pos, const int32_t *list, int64_t n_facets,
                              const double *load, int32_t kind, int32_t k, double *fc,
                              double *F) {
  TFM_REQUIRE(m && mask && pos && load && F, "null argument");
  TFM_REQUIRE(kind >= 0 && kind <= 2, "unknown load kind");
  TFM_REQUIRE(kind != 0 || k == m->dim, "a pressure load has one component per dimension");
  TFM_REQUIRE(n_facets == 0 || (list && fc), "null facet arrays");
  hipStream_t s = (hipStream_t)stream;
  return with_etype(m->etype, [&](auto et) -> int {
    using E = decltype(et);
    return with_k(k, [&](auto kc) -> int {
      constexpr int K = decltype(kc)::value;
      if (n_facets > 0) {
        hipLaunchKernelGGL((k_facet_integrate<E, K>), dim3(blocks_for(n_facets, 64)), dim3(64), 0,
                           s, *m, list, n_facets, load, kind, fc);
        TFM_CHECK_LAUNCH();
      }
      if (m->n_nod > 0) {
        hipLaunchKernelGGL((k_facet_gather<E, K>), dim3(blocks_for(m->n_nod)), dim3(256), 0, s,
                           *m, mask, pos, (const double *)fc, F);
        TFM_CHECK_LAUNCH();
      }
      return TFM_OK;
    });
  });
}

extern "C" int tfm_body_load(tfm_stream_t stream, const tfm_mesh_t *m, const double *scale,
                             const double *load, int32_t per_element, int32_t k, double *fe,
                             double *F, int32_t *bad_jac) {
  TFM_REQUIRE(m && load && fe && F && bad_jac, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (m->n_elem == 0) {
    if (m->n_nod > 0) TFM_CHECK_HIP(hipMemsetAsync(F, 0, sizeof(double) * m->n_nod * k, s));
    return TFM_OK;
  }
  return with_etype(m->etype, [&](auto et) -> int {
    using E = decltype(et);
    return with_k(k, [&](auto kc) -> int {
      constexpr int K = decltype(kc)::value;
      hipLaunchKernelGGL((k_body_load<E, K>), dim3(blocks_for(m->n_elem, 64)), dim3(64), 0, s, *m,
                         scale, load, per_element, fe, bad_jac);
      TFM_CHECK_LAUNCH();
      tfm_mesh_t mk = *m;  // the ordered node gather with K components per node
      mk.ndof = K;
      return launch_gather(s, mk, fe, F);
    });
  });
}

extern "C" int tfm_element_mass(tfm_stream_t stream, const tfm_mesh_t *m, const double *rho,
                                const double *scale, double *mass, int32_t *bad_jac) {
  TFM_REQUIRE(m && rho && mass && bad_jac, "null argument");
  TFM_REQUIRE(m->ndof >= 1 && m->ndof <= 3, "ndof must be 1, 2 or 3");
  if (m->n_elem == 0) return TFM_OK;
  hipStream_t s = (hipStream_t)stream;
  return with_etype(m->etype, [&](auto et) -> int {
    using E = decltype(et);
    const dim3 g((unsigned)m->n_elem);
    switch (m->ndof) {
      case 1: hipLaunchKernelGGL((k_element_mass<E, 1>), g, dim3(64), 0, s, *m, rho, scale, mass, bad_jac); break;
      case 2: hipLaunchKernelGGL((k_element_mass<E, 2>), g, dim3(64), 0, s, *m, rho, scale, mass, bad_jac); break;
      default: hipLaunchKernelGGL((k_element_mass<E, 3>), g, dim3(64), 0, s, *m, rho, scale, mass, bad_jac); break;
    }
    TFM_CHECK_LAUNCH();
    return TFM_OK;
  });
}

extern "C" int tfm_shape_gradients(tfm_stream_t stream, const tfm_mesh_t *m, const double *xi,
                                   int32_t nq, double *B, double *detJ, int32_t *bad_jac) {
  TFM_REQUIRE(m && xi && B && detJ && bad_jac, "null argument");
  TFM_REQUIRE(nq >= 1, "at least one local point is needed");
  const int64_t nt = (int64_t)nq * m->n_elem;
  if (nt == 0) return TFM_OK;
  hipStream_t s = (hipStream_t)stream;
  return with_etype(m->etype, [&](auto et) -> int {
    using E = decltype(et);
    hipLaunchKernelGGL(k_shape_gradients<E>, dim3(blocks_for(nt, 128)), dim3(128), 0, s, *m, xi,
                       nq, B, detJ, bad_jac);
    TFM_CHECK_LAUNCH();
    return TFM_OK;
  });
}
