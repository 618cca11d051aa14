// Element integration, deterministic assembly, internal forces and parameter
// sensitivities for gfx950. All floating-point work is FP64; the FP32 matrix
// copy for the mixed-precision solver is rounded from the FP64 sums.
#include "tfm_common.h"

namespace tfm {

// Shape-function tables, evaluated at compile time (see tfm_common.h).
__constant__ ShapeTable<Quad1> c_quad1 = make_table<Quad1>();
__constant__ ShapeTable<Tetra1> c_tetra1 = make_table<Tetra1>();
__constant__ ShapeTable<Hexa1> c_hexa1 = make_table<Hexa1>();
__constant__ ShapeTable<Hexa2> c_hexa2 = make_table<Hexa2>();

template <> __device__ const ShapeTable<Quad1> &shape_table<Quad1>() { return c_quad1; }
template <> __device__ const ShapeTable<Tetra1> &shape_table<Tetra1>() { return c_tetra1; }
template <> __device__ const ShapeTable<Hexa1> &shape_table<Hexa1>() { return c_hexa1; }
template <> __device__ const ShapeTable<Hexa2> &shape_table<Hexa2>() { return c_hexa2; }

// Material parameter i of element e in the layout of `kind`.
template <int MAT, int DIM>
__device__ __forceinline__ double mat_param(const tfm_material_t &mat, int64_t e, int i) {
  if constexpr (MAT == TFM_MAT_ISO) return i == 0 ? mat.p0[e] : mat.p1[e];
  else return mat.p0[e * mat_nparam(MAT, DIM) + i];
}

__device__ __forceinline__ double mat_scale(const tfm_material_t &mat, int64_t e) {
  return mat.scale ? mat.scale[e] : 1.0;
}

// Jacobian of element geometry X[NPE][DIM] at Gauss point g (base.py:307-313).
template <class E>
__device__ __forceinline__ double geometry(const double (*X)[E::DIM], int g,
                                           double (&Ji)[E::DIM][E::DIM]) {
  const auto &T = shape_table<E>();
  double J[E::DIM][E::DIM];
  for (int i = 0; i < E::DIM; ++i)
    for (int k = 0; k < E::DIM; ++k) {
      double acc = 0.0;
      for (int N = 0; N < E::NPE; ++N) acc += T.dN[g][i][N] * X[N][k];
      J[i][k] = acc;
    }
  return inv_jac<E::DIM>(J, Ji);
}

// B[:, N] = J^{-1} dN/dxi[:, N]
template <class E>
__device__ __forceinline__ void bcol(const double (&Ji)[E::DIM][E::DIM], int g, int N,
                                     double *b) {
  const auto &T = shape_table<E>();
  for (int i = 0; i < E::DIM; ++i) {
    double acc = 0.0;
    for (int k = 0; k < E::DIM; ++k) acc += Ji[i][k] * T.dN[g][k][N];
    b[i] = acc;
  }
}

// ========================================================= pattern expansion
// Row pointer of DOF row (n, a) of the node-block layout.
__device__ __forceinline__ int64_t row_start(const tfm_mesh_t &m, int64_t n, int a, int deg) {
  return m.val_ptr[n] + (deg > 0 ? (int64_t)a * m.ndof * deg : a);
}

__global__ void k_pattern_rows(tfm_mesh_t m, int64_t *rows, int64_t *cols, int64_t *rowptr,
                               int32_t *diag_map) {
  const int64_t n = xcd_block();
  const int nd = m.ndof;
  const int a0 = m.adj_ptr[n];
  const int deg = m.adj_ptr[n + 1] - a0;
  int pself = 0;
  for (int p = 0; p < deg; ++p)
    if (m.adj[a0 + p] == n) pself = p;
  if (threadIdx.x < nd) {
    const int a = threadIdx.x;
    const int64_t rs = row_start(m, n, a, deg);
    if (rowptr) {
      rowptr[n * nd + a] = rs;
      if (n == m.n_nod - 1 && a == nd - 1) rowptr[m.n_nod * nd] = m.nnz;
    }
    if (diag_map) diag_map[n * nd + a] = (int32_t)(deg > 0 ? rs + (int64_t)pself * nd + a : rs);
  }
  if (!rows && !cols) return;
  if (deg == 0) {
    if (threadIdx.x < nd) {
      const int64_t q = m.val_ptr[n] + threadIdx.x;
      if (rows) rows[q] = n * nd + threadIdx.x;
      if (cols) cols[q] = n * nd + threadIdx.x;
    }
    return;
  }
  const int L = nd * deg;
  for (int q = threadIdx.x; q < nd * L; q += blockDim.x) {
    const int a = q / L, rem = q - a * L, p = rem / nd, b = rem - p * nd;
    const int64_t o = m.val_ptr[n] + q;
    if (rows) rows[o] = n * nd + a;
    if (cols) cols[o] = (int64_t)m.adj[a0 + p] * nd + b;
  }
}

// k_map[e, i*nd+a, j*nd+b] = slot of (row n_i*nd+a, col n_j*nd+b) (base.py:94-104)
__global__ void k_pattern_kmap(tfm_mesh_t m, int32_t *k_map) {
  const int64_t e = xcd_block();
  const int npe = m.npe, nd = m.ndof, W = npe * nd;
  for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
    const int r = t / W, c = t - r * W;
    const int i = r / nd, a = r - i * nd, j = c / nd, b = c - j * nd;
    const int64_t ni = m.elements[e * npe + i];
    const int deg = m.adj_ptr[ni + 1] - m.adj_ptr[ni];
    const int pos = m.epos[(e * npe + i) * npe + j];
    k_map[e * W * W + t] = (int32_t)(row_start(m, ni, a, deg) + (int64_t)pos * nd + b);
  }
}

int launch_pattern_expand(hipStream_t s, const tfm_mesh_t &m, int64_t *rows, int64_t *cols,
                          int64_t *rowptr, int32_t *k_map, int32_t *diag_map) {
  if (rows || cols || rowptr || diag_map) {
    hipLaunchKernelGGL(k_pattern_rows, dim3((unsigned)m.n_nod), dim3(256), 0, s, m, rows, cols,
                       rowptr, diag_map);
    TFM_CHECK_LAUNCH();
  }
  if (k_map && m.n_elem > 0) {
    hipLaunchKernelGGL(k_pattern_kmap, dim3((unsigned)m.n_elem), dim3(256), 0, s, m, k_map);
    TFM_CHECK_LAUNCH();
  }
  return TFM_OK;
}

// ========================================================= element matrices
// One 64-lane block per element; lanes own (I, J) node pairs and loop over the
// Gauss points with the element's J^{-1} and w*detJ staged in LDS.
template <class E, int NDOF, int MAT>
__global__ void __launch_bounds__(64) k_element_matrices(tfm_mesh_t m, tfm_material_t mat,
                                                         double *kout, int32_t *bad) {
  constexpr int NPE = E::NPE, NGP = E::NGP, D = E::DIM, NP = mat_nparam(MAT, D);
  constexpr int W = NPE * NDOF, ND2 = NDOF * NDOF;
  __shared__ double X[NPE][D];
  __shared__ double Gm[NGP][D * D + 1];
  __shared__ double MP[NP];
  const int64_t e = xcd_block();
  const int tid = threadIdx.x;
  for (int t = tid; t < NPE * D; t += 64) {
    const int j = t / D, d = t - j * D;
    X[j][d] = m.nodes[(int64_t)m.elements[e * NPE + j] * D + d];
  }
  for (int t = tid; t < NP; t += 64) MP[t] = mat_param<MAT, D>(mat, e, t);
  __syncthreads();
  const double sc = mat_scale(mat, e);
  for (int g = tid; g < NGP; g += 64) {
    double Ji[D][D];
    const double det = geometry<E>(X, g, Ji);
    if (!(det > 0.0)) *bad = 1;
    for (int i = 0; i < D; ++i)
      for (int k = 0; k < D; ++k) Gm[g][i * D + k] = Ji[i][k];
    Gm[g][D * D] = shape_table<E>().w[g] * det * sc;
  }
  __syncthreads();
  for (int t = tid; t < NPE * NPE; t += 64) {
    const int I = t / NPE, J = t - I * NPE;
    double blk[ND2];
    for (int q = 0; q < ND2; ++q) blk[q] = 0.0;
    for (int g = 0; g < NGP; ++g) {
      double Ji[D][D];
      for (int i = 0; i < D; ++i)
        for (int k = 0; k < D; ++k) Ji[i][k] = Gm[g][i * D + k];
      double bI[D], bJ[D], c[ND2];
      bcol<E>(Ji, g, I, bI);
      bcol<E>(Ji, g, J, bJ);
      coupling_block<D, NDOF, MAT>(bI, bJ, MP, Gm[g][D * D], c);
      for (int q = 0; q < ND2; ++q) blk[q] += c[q];
    }
    for (int a = 0; a < NDOF; ++a)
      for (int b = 0; b < NDOF; ++b)
        kout[e * W * W + (int64_t)(I * NDOF + a) * W + J * NDOF + b] = blk[a * NDOF + b];
  }
}

// Element matrices from a tangent given per Gauss point, C[g][e][i][J][k][L]
// (a finite-strain material's dP/dF): k = sum_g w detJ B^T C_g B times the
// optional volume scale. The element's eight (or fewer) tangents are staged
// in LDS next to its J^{-1} and w*detJ.
template <class E>
__global__ void __launch_bounds__(64) k_element_matrices_gp(tfm_mesh_t m, const double *cgp,
                                                            const double *scale, double *kout,
                                                            int32_t *bad) {
  constexpr int NPE = E::NPE, NGP = E::NGP, D = E::DIM, NC = D * D * D * D;
  constexpr int W = NPE * D, ND2 = D * D;
  __shared__ double X[NPE][D];
  __shared__ double Gm[NGP][D * D + 1];
  __shared__ double Cg[NGP][NC];
  const int64_t e = xcd_block();
  const int tid = threadIdx.x;
  for (int t = tid; t < NPE * D; t += 64) {
    const int j = t / D, d = t - j * D;
    X[j][d] = m.nodes[(int64_t)m.elements[e * NPE + j] * D + d];
  }
  for (int t = tid; t < NGP * NC; t += 64) {
    const int g = t / NC, c = t - g * NC;
    Cg[g][c] = cgp[((int64_t)g * m.n_elem + e) * NC + c];
  }
  __syncthreads();
  const double sc = scale ? scale[e] : 1.0;
  for (int g = tid; g < NGP; g += 64) {
    double Ji[D][D];
    const double det = geometry<E>(X, g, Ji);
    if (!(det > 0.0)) *bad = 1;
    for (int i = 0; i < D; ++i)
      for (int k = 0; k < D; ++k) Gm[g][i * D + k] = Ji[i][k];
    Gm[g][D * D] = shape_table<E>().w[g] * det * sc;
  }
  __syncthreads();
  for (int t = tid; t < NPE * NPE; t += 64) {
    const int I = t / NPE, J = t - I * NPE;
    double blk[ND2];
    for (int q = 0; q < ND2; ++q) blk[q] = 0.0;
    for (int g = 0; g < NGP; ++g) {
      double Ji[D][D];
      for (int i = 0; i < D; ++i)
        for (int k = 0; k < D; ++k) Ji[i][k] = Gm[g][i * D + k];
      double bI[D], bJ[D], c[ND2];
      bcol<E>(Ji, g, I, bI);
      bcol<E>(Ji, g, J, bJ);
      coupling_block<D, D, TFM_MAT_GENERAL>(bI, bJ, Cg[g], Gm[g][D * D], c);
      for (int q = 0; q < ND2; ++q) blk[q] += c[q];
    }
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b)
        kout[e * W * W + (int64_t)(I * D + a) * W + J * D + b] = blk[a * D + b];
  }
}

// ================================================================ assembly
// One 64-lane block per node n assembles the node's ndof CSR rows: the
// incident elements are visited in ascending order in chunks of KCH; each
// (element, column node) pair forms its ndof x ndof block from the staged
// geometry, and the row entries add the blocks in element order. Every CSR
// value is written exactly once, so no float atomics are needed and the
// result is bitwise reproducible. The Dirichlet epilogue of
// FEM.assemble_matrix (base.py:413-419) is applied on the way out.
template <class E> constexpr int assemble_chunk() { return E::NPE >= 16 ? 4 : 64 / E::NPE; }

template <class E, int NDOF, int MAT>
__global__ void __launch_bounds__(64) k_assemble_rows(tfm_mesh_t m, tfm_material_t mat,
                                                      const uint8_t *con, double *v64,
                                                      float *v32, double *dblk, int32_t *bad,
                                                      int diff32) {
  constexpr int NPE = E::NPE, NGP = E::NGP, D = E::DIM, NP = mat_nparam(MAT, D);
  constexpr int ND2 = NDOF * NDOF, KCH = assemble_chunk<E>();
  __shared__ double X[KCH][NPE][D];
  __shared__ double Gm[KCH][NGP][D * D + 1];
  __shared__ double MP[KCH][NP];
  __shared__ double S[KCH * NPE][ND2];
  __shared__ int Eid[KCH], Iloc[KCH];
  __shared__ int pself;
  extern __shared__ double dyn[];
  double *acc = dyn;                                              // [ND2 * maxdeg]
  signed char *slot = reinterpret_cast<signed char *>(acc + ND2 * m.max_deg);  // [KCH * maxdeg]

  const int64_t n = xcd_block();
  const int tid = threadIdx.x;
  const int a0 = m.adj_ptr[n];
  const int deg = m.adj_ptr[n + 1] - a0;
  const int64_t vb = m.val_ptr[n];
  if (deg == 0) {  // orphan node: only its diagonal is stored (base.py:89-90)
    if (tid < NDOF) {
      const int64_t r = n * NDOF + tid;
      const double v = con[r] ? 1.0 : 0.0;
      if (v64) v64[vb + tid] = v;
      if (v32) v32[vb + tid] = (float)v;
    }
    for (int t = tid; t < ND2; t += 64) {
      const int a = t / NDOF, b = t - a * NDOF;
      dblk[n * ND2 + t] = (a == b && con[n * NDOF + a]) ? 1.0 : 0.0;
    }
    return;
  }
  const int inc0 = m.ne_ptr[n];
  const int ninc = m.ne_ptr[n + 1] - inc0;
  const int L = NDOF * deg;
  for (int q = tid; q < ND2 * deg; q += 64) acc[q] = 0.0;

  for (int c0 = 0; c0 < ninc; c0 += KCH) {
    const int kc = min(KCH, ninc - c0);
    __syncthreads();
    if (tid < kc) {
      Eid[tid] = m.ne_elem[inc0 + c0 + tid];
      Iloc[tid] = m.ne_loc[inc0 + c0 + tid];
    }
    for (int t = tid; t < KCH * deg; t += 64) slot[t] = -1;
    __syncthreads();
    if (c0 == 0 && tid == 0) {
      const int64_t e0 = Eid[0];
      pself = m.epos[(e0 * NPE + Iloc[0]) * NPE + Iloc[0]];
    }
    for (int t = tid; t < kc * NPE * D; t += 64) {
      const int k = t / (NPE * D), r = t - k * NPE * D, j = r / D, d = r - j * D;
      X[k][j][d] = m.nodes[(int64_t)m.elements[(int64_t)Eid[k] * NPE + j] * D + d];
    }
    for (int t = tid; t < kc * NP; t += 64) {
      const int k = t / NP, i = t - k * NP;
      MP[k][i] = mat_param<MAT, D>(mat, Eid[k], i);
    }
    for (int t = tid; t < kc * NPE; t += 64) {
      const int k = t / NPE, j = t - k * NPE;
      const int64_t e = Eid[k];
      slot[k * deg + m.epos[(e * NPE + Iloc[k]) * NPE + j]] = (signed char)j;
    }
    __syncthreads();
    for (int t = tid; t < kc * NGP; t += 64) {
      const int k = t / NGP, g = t - k * NGP;
      double Ji[D][D];
      const double det = geometry<E>(X[k], g, Ji);
      if (!(det > 0.0)) *bad = 1;
      for (int i = 0; i < D; ++i)
        for (int kk = 0; kk < D; ++kk) Gm[k][g][i * D + kk] = Ji[i][kk];
      Gm[k][g][D * D] = shape_table<E>().w[g] * det * mat_scale(mat, Eid[k]);
    }
    __syncthreads();
    for (int t = tid; t < kc * NPE; t += 64) {
      const int k = t / NPE, j = t - k * NPE, i = Iloc[k];
      double blk[ND2];
      for (int q = 0; q < ND2; ++q) blk[q] = 0.0;
      for (int g = 0; g < NGP; ++g) {
        double Ji[D][D];
        for (int a = 0; a < D; ++a)
          for (int b = 0; b < D; ++b) Ji[a][b] = Gm[k][g][a * D + b];
        double bi[D], bj[D], c[ND2];
        bcol<E>(Ji, g, i, bi);
        bcol<E>(Ji, g, j, bj);
        coupling_block<D, NDOF, MAT>(bi, bj, MP[k], Gm[k][g][D * D], c);
        for (int q = 0; q < ND2; ++q) blk[q] += c[q];
      }
      for (int q = 0; q < ND2; ++q) S[t][q] = blk[q];
    }
    __syncthreads();
    for (int q = tid; q < ND2 * deg; q += 64) {
      const int a = q / L, rem = q - a * L, p = rem / NDOF, b = rem - p * NDOF;
      double v = acc[q];
      for (int k = 0; k < kc; ++k) {
        const int j = slot[k * deg + p];
        if (j >= 0) v += S[k * NPE + j][a * NDOF + b];
      }
      acc[q] = v;
    }
  }
  __syncthreads();
  for (int q = tid; q < ND2 * deg; q += 64) {
    const int a = q / L, rem = q - a * L, p = rem / NDOF, b = rem - p * NDOF;
    const int64_t row = n * NDOF + a;
    const int64_t col = (int64_t)m.adj[a0 + p] * NDOF + b;
    double v = acc[q];
    if (con[row] || con[col]) v = (row == col) ? 1.0 : 0.0;
    acc[q] = v;
    if (v64) v64[vb + q] = v;
    if (v32 && (p != pself || !diff32)) v32[vb + q] = (float)v;
    if (p == pself) dblk[n * ND2 + a * NDOF + b] = v;
  }
  // FP32 difference form: the diagonal block holds the row sums of the
  // final FP64 values, added in ascending slot order
  if (v32 && diff32) {
    __syncthreads();
    if (tid < ND2) {
      const int a = tid / NDOF, b = tid - a * NDOF;
      double r = 0.0;
      for (int p = 0; p < deg; ++p) r += acc[a * L + p * NDOF + b];
      v32[vb + a * L + pself * NDOF + b] = (float)r;
    }
  }
}

// Node-tiled assembly. A 64-lane block assembles the rows of NW consecutive
// nodes. A node owns NPE * IP lanes: IP of its incident elements are taken at
// once (incidence t = r IP + ip in round r), each by NPE lanes, lane j for the
// element's column node j. Per incidence, lane g < NGP evaluates the Jacobian
// at Gauss point g once and shares J^{-1} and the scaled row gradient
// w_g det J_g s_e dN_i/dx through LDS; lane j then forms the geometric moments
//   M_ij[a][c] = sum_g w_g det J_g s_e (dN_i/dx)_a (dN_j/dx)_c
// and the material block once (isotropic elasticity: lam M + G M^T + G tr(M) I;
// heat: kappa : M), the per-point form of coupling_block summed over the
// points. The blocks of a round are added into per-node row accumulators in
// LDS one incidence group after the other (ascending t, a fixed order, so the
// sums are bitwise reproducible), and the tile's rows, contiguous in the value
// arrays, leave in one coalesced pass with the Dirichlet epilogue of
// k_assemble_rows. The element data of RB rounds are loaded together: a batch
// costs three dependent global round trips.
// IP and RB per element type, from timings on the 64^3 Hexa1, 80^3 Tetra1 and
// 32^3 Hexa2 cubes (IP 1-4 and RB 1-4 swept; occupancy is bound by VGPRs and
// LDS together)
template <class E> struct NodeTile;
template <> struct NodeTile<Quad1> { static constexpr int IP = 4, RB = 1; };
template <> struct NodeTile<Tetra1> { static constexpr int IP = 2, RB = 1; };
template <> struct NodeTile<Hexa1> { static constexpr int IP = 2, RB = 2; };
template <> struct NodeTile<Hexa2> { static constexpr int IP = 1, RB = 4; };
// the general (anisotropic) stiffness has 81 parameters per element, too many
// to hold per lane: it keeps the one-node-per-block kernel
template <int MAT> constexpr bool node_tiled() { return MAT != TFM_MAT_GENERAL; }
template <class E> constexpr int tile_nodes() { return 64 / (E::NPE * NodeTile<E>::IP); }

template <class E, int NDOF, int MAT>
__global__ void __launch_bounds__(64) k_assemble_nodes(tfm_mesh_t m, tfm_material_t mat,
                                                       const uint8_t *con, double *v64,
                                                       float *v32, double *dblk, int32_t *bad,
                                                       int diff32) {
  constexpr int NPE = E::NPE, NGP = E::NGP, D = E::DIM, NP = mat_nparam(MAT, D);
  constexpr int IP = NodeTile<E>::IP, RB = NodeTile<E>::RB;
  constexpr int ND2 = NDOF * NDOF, LPN = NPE * IP, NW = tile_nodes<E>(), NG = NW * IP;
  constexpr int GW = D * D + D;  // per point: J^{-1}, then the scaled row gradient
  constexpr int CQ = 4;          // Dirichlet-bit chunk: 64 CQ columns
  static_assert(NW >= 1 && NGP <= NPE, "a tile holds whole nodes, one point per lane");
  static_assert(MAT != TFM_MAT_GENERAL, "the general stiffness keeps k_assemble_rows");
  extern __shared__ double acc[];  // [NW][ND2 * max_deg] row accumulators, then
                                   // [NW][max_deg] Dirichlet bits of the columns
  // row strides padded by two doubles: the 8-byte banks of rows that lanes
  // read at the same time then differ (no bank conflicts between groups)
  constexpr int TS = D * NPE + 2, XS = NPE * D + 2, GS = NGP * GW + 2;
  __shared__ double tbl[NGP][TS];  // dN[g][a][N] at [g][a * NPE + N]
  __shared__ double tw[NGP];
  __shared__ double Xs[NG][XS];   // [group][N * D + c]
  __shared__ double Gm[NG][GS];   // [group][g * GW + ...]
  __shared__ int64_t s_vb[NW + 1];
  __shared__ int s_a0[NW], s_deg[NW], s_self[NW], s_crow[NW];

  const int lane = threadIdx.x;
  const int w = lane / LPN, ip = (lane / NPE) % IP, j = lane % NPE, grp = w * IP + ip;
  const int64_t n0 = xcd_block() * (int64_t)NW;
  const int64_t n = n0 + w;
  const bool valid = lane < NW * LPN && n < m.n_nod;
  const int stride = ND2 * m.max_deg;
  unsigned char *cmask = reinterpret_cast<unsigned char *>(acc + NW * stride);
  int inc0 = 0, ninc = 0;
  if (valid) {
    inc0 = m.ne_ptr[n];
    ninc = m.ne_ptr[n + 1] - inc0;
  }
  if (lane < NW) {
    const int64_t nn = n0 + lane;
    const bool ok = nn < m.n_nod;
    s_a0[lane] = ok ? m.adj_ptr[nn] : 0;
    s_deg[lane] = ok ? m.adj_ptr[nn + 1] - m.adj_ptr[nn] : 0;
    s_vb[lane] = m.val_ptr[ok ? nn : m.n_nod];
    s_self[lane] = -1;
    int cr = 0;
    for (int a = 0; a < NDOF; ++a) cr |= ok && con[nn * NDOF + a] ? 1 << a : 0;
    s_crow[lane] = cr;
    if (lane == NW - 1) s_vb[NW] = m.val_ptr[nn + 1 < m.n_nod ? nn + 1 : m.n_nod];
  }
  const auto &T = shape_table<E>();
  for (int q = lane; q < NGP * D * NPE; q += 64) tbl[q / (D * NPE)][q % (D * NPE)] = (&T.dN[0][0][0])[q];
  for (int q = lane; q < NGP; q += 64) tw[q] = T.w[q];
  for (int q = lane; q < NW * stride; q += 64) acc[q] = 0.0;
  __syncthreads();

  // Dirichlet bits of every column node of the tile's rows (two dependent
  // loads each), in chunks of 64 CQ columns; the first chunk rides along with
  // the element loads of the first batch, and the epilogue reads LDS only
  const int nq = NW * m.max_deg;
  auto cm_adj = [&](int q0, int (&nbr)[CQ]) {
#pragma unroll
    for (int c = 0; c < CQ; ++c) {
      const int q = q0 + c * 64 + lane;
      nbr[c] = -1;
      if (q < nq) {
        const int ww = q / m.max_deg, p = q - ww * m.max_deg;
        if (p < s_deg[ww]) nbr[c] = m.adj[s_a0[ww] + p];
      }
    }
  };
  auto cm_con = [&](int q0, const int (&nbr)[CQ]) {
#pragma unroll
    for (int c = 0; c < CQ; ++c) {
      const int q = q0 + c * 64 + lane;
      int bits = 0;
      if (nbr[c] >= 0)
        for (int b = 0; b < NDOF; ++b) bits |= con[(int64_t)nbr[c] * NDOF + b] ? 1 << b : 0;
      if (q < nq) cmask[q] = (unsigned char)bits;
    }
  };

  const int rounds = (m.max_inc + IP - 1) / IP;
  for (int r0 = 0; r0 < rounds; r0 += RB) {
    // the batch's element data: incidences, then connectivity, slots and
    // parameters, then coordinates
    int e[RB], il[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int t = (r0 + b) * IP + ip;
      e[b] = -1;
      il[b] = 0;
      if (valid && t < ninc) {
        e[b] = m.ne_elem[inc0 + t];
        il[b] = m.ne_loc[inc0 + t];
      }
    }
    int nbr[CQ];
    if (r0 == 0) cm_adj(0, nbr);
    int64_t id[RB];
    int P[RB];  // slot of the lane's column node in the row node's adjacency
    double mp[RB][NP], se[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      id[b] = 0;
      P[b] = 0;
      se[b] = 0.0;
      for (int q = 0; q < NP; ++q) mp[b][q] = 0.0;
      if (e[b] >= 0) {
        const int64_t eb = e[b];
        id[b] = m.elements[eb * NPE + j];
        P[b] = m.epos[(eb * NPE + il[b]) * NPE + j];
        for (int q = 0; q < NP; ++q) mp[b][q] = mat_param<MAT, D>(mat, eb, q);
        se[b] = mat_scale(mat, eb);
      }
    }
    if (r0 == 0) cm_con(0, nbr);
    double X[RB][D];
#pragma unroll
    for (int b = 0; b < RB; ++b)
      for (int d = 0; d < D; ++d) X[b][d] = e[b] >= 0 ? m.nodes[id[b] * D + d] : 0.0;

#pragma unroll
    for (int b = 0; b < RB; ++b) {
      if (r0 + b >= rounds) break;
      const bool act = e[b] >= 0;
      const int i = il[b];
      if (act) {
        for (int d = 0; d < D; ++d) Xs[grp][j * D + d] = X[b][d];
        if (r0 + b == 0 && ip == 0 && j == i) s_self[w] = P[b];
      }
      __syncthreads();
      if (act && j < NGP) {  // lane g = j: the geometry of point g
        const int g = j;
        double J[D][D];
        for (int a = 0; a < D; ++a)
          for (int c = 0; c < D; ++c) J[a][c] = 0.0;
        for (int N = 0; N < NPE; ++N)
          for (int a = 0; a < D; ++a)
            for (int c = 0; c < D; ++c) J[a][c] += tbl[g][a * NPE + N] * Xs[grp][N * D + c];
        double Ji[D][D];
        const double det = inv_jac<D>(J, Ji);
        if (!(det > 0.0)) *bad = 1;
        const double wd = tw[g] * det * se[b];
        double *G = &Gm[grp][g * GW];
        for (int a = 0; a < D; ++a) {
          double v = 0.0;
          for (int c = 0; c < D; ++c) {
            G[a * D + c] = Ji[a][c];
            v += Ji[a][c] * tbl[g][c * NPE + i];
          }
          G[D * D + a] = v * wd;
        }
      }
      __syncthreads();
      double blk[ND2];
      if (act) {
        double M[D * D];
        for (int q = 0; q < D * D; ++q) M[q] = 0.0;
        for (int g = 0; g < NGP; ++g) {
          const double *G = &Gm[grp][g * GW];
          double bj[D];
          for (int a = 0; a < D; ++a) {
            double v = 0.0;
            for (int c = 0; c < D; ++c) v += G[a * D + c] * tbl[g][c * NPE + j];
            bj[a] = v;
          }
          if constexpr (MAT == TFM_MAT_HEAT_ISO) {
            for (int a = 0; a < D; ++a) M[0] += G[D * D + a] * bj[a];
          } else {
            for (int a = 0; a < D; ++a)
              for (int c = 0; c < D; ++c) M[a * D + c] += G[D * D + a] * bj[c];
          }
        }
        if constexpr (MAT == TFM_MAT_ISO) {
          double tr = 0.0;
          for (int a = 0; a < D; ++a) tr += M[a * D + a];
          for (int a = 0; a < D; ++a)
            for (int c = 0; c < D; ++c)
              blk[a * NDOF + c] = mp[b][0] * M[a * D + c] + mp[b][1] * M[c * D + a] +
                                  (a == c ? mp[b][1] * tr : 0.0);
        } else if constexpr (MAT == TFM_MAT_HEAT_ISO) {
          blk[0] = mp[b][0] * M[0];
        } else {
          double v = 0.0;
          for (int q = 0; q < D * D; ++q) v += mp[b][q] * M[q];
          blk[0] = v;
        }
      }
      // add the blocks in ascending incidence order: group ip at step ip
      for (int st = 0; st < IP; ++st) {
        if (act && ip == st) {
          double *A = acc + w * stride;
          const int L = NDOF * s_deg[w];
          for (int a = 0; a < NDOF; ++a)
            for (int c = 0; c < NDOF; ++c) A[a * L + P[b] * NDOF + c] += blk[a * NDOF + c];
        }
        __syncthreads();
      }
    }
  }
  for (int q0 = rounds > 0 ? 64 * CQ : 0; q0 < nq; q0 += 64 * CQ) {
    int nbr[CQ];
    cm_adj(q0, nbr);
    cm_con(q0, nbr);
  }
  __syncthreads();
  // the tile's rows are contiguous in the value arrays: [s_vb[0], s_vb[NW])
  const int64_t vb0 = s_vb[0];
  const int total = (int)(s_vb[NW] - vb0);
  int wq = 0, beg = 0, end = (int)(s_vb[1] - vb0);  // lane's node and its value range
  for (int q = lane; q < total; q += 64) {
    while (end <= q) {
      ++wq;
      beg = end;
      end = (int)(s_vb[wq + 1] - vb0);
    }
    const int64_t nn = n0 + wq;
    const int deg = s_deg[wq];
    const int ql = q - beg;
    if (deg == 0) {  // orphan node: only its diagonal is stored (base.py:89-90)
      const double v = (s_crow[wq] >> ql) & 1 ? 1.0 : 0.0;
      if (v64) v64[vb0 + q] = v;
      if (v32) v32[vb0 + q] = (float)v;
      continue;
    }
    const int L = NDOF * deg;
    int a = 0, rem = ql;  // ql = a L + rem
    for (int k = 1; k < NDOF; ++k)
      if (rem >= L) {
        rem -= L;
        ++a;
      }
    const int p = rem / NDOF, b = rem - p * NDOF;
    const int self = s_self[wq];
    double *A = acc + wq * stride;
    double v = A[ql];
    if (((s_crow[wq] >> a) & 1) || ((cmask[wq * m.max_deg + p] >> b) & 1))
      v = (p == self && a == b) ? 1.0 : 0.0;  // row == col
    A[ql] = v;
    if (v64) v64[vb0 + q] = v;
    if (v32 && (p != self || !diff32)) v32[vb0 + q] = (float)v;
    if (p == self) dblk[nn * ND2 + a * NDOF + b] = v;
  }
  for (int q = lane; q < NW * ND2; q += 64) {  // orphan diagonal blocks
    const int ww = q / ND2, ab = q - ww * ND2, a = ab / NDOF, b = ab - a * NDOF;
    const int64_t nn = n0 + ww;
    if (nn < m.n_nod && s_deg[ww] == 0)
      dblk[nn * ND2 + ab] = (a == b && ((s_crow[ww] >> a) & 1)) ? 1.0 : 0.0;
  }
  // FP32 difference form: the diagonal block holds the row sums of the
  // final FP64 values, added in ascending slot order
  if (v32 && diff32) {
    __syncthreads();
    for (int q = lane; q < NW * ND2; q += 64) {
      const int ww = q / ND2, ab = q - ww * ND2, a = ab / NDOF, b = ab - a * NDOF;
      const int deg = s_deg[ww];
      if (n0 + ww >= m.n_nod || deg == 0) continue;
      const int L = NDOF * deg;
      const double *A = acc + ww * stride;
      double r = 0.0;
      for (int p = 0; p < deg; ++p) r += A[a * L + p * NDOF + b];
      v32[s_vb[ww] + a * L + s_self[ww] * NDOF + b] = (float)r;
    }
  }
}

// =========================================================== element teams
// Kernels below give each element a team of LG = next_pow2(NPE) lanes, one
// lane per element node; Jacobians and gradients are summed across the team
// with butterfly shuffles, so every lane ends with the full tensors.
template <class E> constexpr int team_lanes() { return E::NPE <= 4 ? 4 : E::NPE <= 8 ? 8 : 32; }

template <int LG, int NV> __device__ __forceinline__ void team_sum(double (&v)[NV]) {
#pragma unroll
  for (int off = 1; off < LG; off <<= 1)
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] += __shfl_xor(v[i], off, LG);
}

// Per-lane geometry at Gauss point g: J^{-1}, detJ and this lane's column of B.
template <class E, int LG>
__device__ __forceinline__ double team_geometry(const double *Xj, bool act, int j, int g,
                                                double (&Ji)[E::DIM][E::DIM], double *bj) {
  constexpr int D = E::DIM;
  const auto &T = shape_table<E>();
  double Jv[D * D];
  for (int i = 0; i < D; ++i)
    for (int k = 0; k < D; ++k) Jv[i * D + k] = act ? T.dN[g][i][j] * Xj[k] : 0.0;
  team_sum<LG, D * D>(Jv);
  double J[D][D];
  for (int i = 0; i < D; ++i)
    for (int k = 0; k < D; ++k) J[i][k] = Jv[i * D + k];
  const double det = inv_jac<D>(J, Ji);
  for (int i = 0; i < D; ++i) {
    double acc = 0.0;
    if (act)
      for (int k = 0; k < D; ++k) acc += Ji[i][k] * T.dN[g][k][j];
    bj[i] = acc;
  }
  return det;
}

// Gradient H[a][i] = sum_N u[N][a] B[i][N], summed over the team (base.py:1052).
template <int LG, int D, int NDOF>
__device__ __forceinline__ void team_grad(const double *uj, const double *bj, double (&H)[NDOF * D]) {
  for (int a = 0; a < NDOF; ++a)
    for (int i = 0; i < D; ++i) H[a * D + i] = uj[a] * bj[i];
  team_sum<LG, NDOF * D>(H);
}

// Internal element forces fe[e][j][a] = sum_g w detJ s sum_i B[i][j] q[a][i]
// with q = flux_prev + flux(grad u) (base.py:1050-1083 / 1250-1268); also
// the Gauss-point means of grad u and q.
template <class E, int NDOF, int MAT>
__global__ void __launch_bounds__(256) k_internal_force(tfm_mesh_t m, tfm_material_t mat,
                                                        const double *u, const double *qprev,
                                                        double *fe, double *grad_mean,
                                                        double *flux_mean, double *grad_gp,
                                                        double *flux_gp) {
  constexpr int NPE = E::NPE, NGP = E::NGP, D = E::DIM, NP = mat_nparam(MAT, D);
  constexpr int LG = team_lanes<E>(), EPB = 256 / LG, NQ = NDOF * D;
  __shared__ double MP[EPB][NP];
  const int tid = threadIdx.x, slotE = tid / LG, j = tid % LG;
  const int64_t blk = xcd_block();
  const int64_t e = blk * EPB + slotE;
  const bool live = e < m.n_elem;
  const bool act = live && j < NPE;
  for (int t = tid; t < EPB * NP; t += 256) {
    const int k = t / NP, i = t - k * NP;
    const int64_t ee = blk * EPB + k;
    MP[k][i] = ee < m.n_elem ? mat_param<MAT, D>(mat, ee, i) : 0.0;
  }
  double Xj[D], uj[NDOF];
  int64_t node = 0;
  if (act) node = m.elements[e * NPE + j];
  for (int d = 0; d < D; ++d) Xj[d] = act ? m.nodes[node * D + d] : 0.0;
  for (int a = 0; a < NDOF; ++a) uj[a] = act ? u[node * NDOF + a] : 0.0;
  __syncthreads();
  const double sc = live ? mat_scale(mat, e) : 0.0;
  double fj[NDOF], Hs[NQ], Qs[NQ];
  for (int a = 0; a < NDOF; ++a) fj[a] = 0.0;
  for (int q = 0; q < NQ; ++q) Hs[q] = Qs[q] = 0.0;
  for (int g = 0; g < NGP; ++g) {
    double Ji[D][D], bj[D], H[NQ], Q[NQ];
    const double det = team_geometry<E, LG>(Xj, act, j, g, Ji, bj);
    team_grad<LG, D, NDOF>(uj, bj, H);
    flux_from_grad<D, NDOF, MAT>(H, MP[slotE], Q);
    if (qprev && live)
      for (int q = 0; q < NQ; ++q) Q[q] += qprev[((int64_t)g * m.n_elem + e) * NQ + q];
    const double wd = shape_table<E>().w[g] * det * sc;
    for (int a = 0; a < NDOF; ++a) {
      double acc = 0.0;
      for (int i = 0; i < D; ++i) acc += bj[i] * Q[a * D + i];
      fj[a] += wd * acc;
    }
    for (int q = 0; q < NQ; ++q) {
      Hs[q] += H[q];
      Qs[q] += Q[q];
    }
    if (live && (grad_gp || flux_gp)) {
      const int64_t o = ((int64_t)g * m.n_elem + e) * NQ;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (q % LG == j) {  // component q is written by lane q % LG
          if (grad_gp) grad_gp[o + q] = H[q];
          if (flux_gp) flux_gp[o + q] = Q[q];
        }
    }
  }
  if (act)
    for (int a = 0; a < NDOF; ++a) fe[(e * NPE + j) * NDOF + a] = fj[a];
  if (live && (grad_mean || flux_mean)) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (q % LG == j) {
        if (grad_mean) grad_mean[e * NQ + q] = Hs[q] / NGP;
        if (flux_mean) flux_mean[e * NQ + q] = Qs[q] / NGP;
      }
  }
}

// F[n*nd+a] = sum over incident elements (ascending) of fe[e][loc][a]: the
// scatter of FEM.assemble_rhs (base.py:439-445) as an ordered gather.
template <int NDOF>
__global__ void k_gather_nodes(tfm_mesh_t m, const double *fe, double *F) {
  const int64_t t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= m.n_nod * NDOF) return;
  const int64_t n = t / NDOF;
  const int a = (int)(t - n * NDOF);
  double acc = 0.0;
  int k = m.ne_ptr[n];
  const int k_end = m.ne_ptr[n + 1];
  auto at = [&](int i) -> int64_t {
    return ((int64_t)m.ne_elem[i] * m.npe + m.ne_loc[i]) * NDOF + a;
  };
  // four incident elements per round: their index and value loads are in
  // flight together, and the sum keeps the ascending element order
  for (; k + 3 < k_end; k += 4) {
    const int64_t i0 = at(k), i1 = at(k + 1), i2 = at(k + 2), i3 = at(k + 3);
    const double v0 = fe[i0], v1 = fe[i1], v2 = fe[i2], v3 = fe[i3];
    acc += v0;
    acc += v1;
    acc += v2;
    acc += v3;
  }
  for (; k < k_end; ++k) acc += fe[at(k)];
  F[t] = acc;
}

// Assembly from given element matrices kin [E, W, W]: same node-row order and
// Dirichlet epilogue as k_assemble_rows.
template <int NDOF>
__global__ void __launch_bounds__(64) k_assemble_elements(tfm_mesh_t m, const double *kin,
                                                          const uint8_t *con, double *v64,
                                                          float *v32, double *dblk, int diff32) {
  constexpr int ND2 = NDOF * NDOF;
  extern __shared__ signed char slot_dyn[];  // [max_inc * max_deg]
  const int64_t n = xcd_block();
  const int tid = threadIdx.x;
  const int npe = m.npe, W = npe * NDOF;
  const int a0 = m.adj_ptr[n];
  const int deg = m.adj_ptr[n + 1] - a0;
  const int64_t vb = m.val_ptr[n];
  if (deg == 0) {
    if (tid < NDOF) {
      const double v = con[n * NDOF + tid] ? 1.0 : 0.0;
      if (v64) v64[vb + tid] = v;
      if (v32) v32[vb + tid] = (float)v;
    }
    for (int t = tid; t < ND2; t += 64) {
      const int a = t / NDOF, b = t - a * NDOF;
      dblk[n * ND2 + t] = (a == b && con[n * NDOF + a]) ? 1.0 : 0.0;
    }
    return;
  }
  const int inc0 = m.ne_ptr[n], ninc = m.ne_ptr[n + 1] - inc0;
  for (int t = tid; t < ninc * deg; t += 64) slot_dyn[t] = -1;
  __syncthreads();
  for (int t = tid; t < ninc * npe; t += 64) {
    const int k = t / npe, j = t - k * npe;
    const int64_t e = m.ne_elem[inc0 + k];
    slot_dyn[k * deg + m.epos[(e * npe + m.ne_loc[inc0 + k]) * npe + j]] = (signed char)j;
  }
  __syncthreads();
  int pself = 0;
  for (int p = 0; p < deg; ++p)
    if (m.adj[a0 + p] == n) pself = p;
  const int L = NDOF * deg;
  // entry (a, p, b) of the node's rows with the Dirichlet epilogue applied
  auto entry = [&](int a, int p, int b) -> double {
    double v = 0.0;
    for (int k = 0; k < ninc; ++k) {
      const int j = slot_dyn[k * deg + p];
      if (j >= 0) {
        const int64_t e = m.ne_elem[inc0 + k];
        const int i = m.ne_loc[inc0 + k];
        v += kin[e * W * W + (int64_t)(i * NDOF + a) * W + j * NDOF + b];
      }
    }
    const int64_t row = n * NDOF + a;
    const int64_t col = (int64_t)m.adj[a0 + p] * NDOF + b;
    if (con[row] || con[col]) v = (row == col) ? 1.0 : 0.0;
    return v;
  };
  for (int q = tid; q < ND2 * deg; q += 64) {
    const int a = q / L, rem = q - a * L, p = rem / NDOF, b = rem - p * NDOF;
    const double v = entry(a, p, b);
    if (v64) v64[vb + q] = v;
    if (v32 && (p != pself || !diff32)) v32[vb + q] = (float)v;
    if (p == pself) dblk[n * ND2 + a * NDOF + b] = v;
  }
  // FP32 difference form: the diagonal block holds the FP64 row sums
  if (v32 && diff32 && tid < ND2) {
    const int a = tid / NDOF, b = tid - a * NDOF;
    double r = 0.0;
    for (int p = 0; p < deg; ++p) r += entry(a, p, b);
    v32[vb + a * L + pself * NDOF + b] = (float)r;
  }
}

// Sensitivities of lam^T K(p) u per element (see tfm_sensitivity in tfm.h).
template <class E, int NDOF, int MAT>
__global__ void __launch_bounds__(256) k_sensitivity(tfm_mesh_t m, tfm_material_t mat,
                                                     const double *u, const double *lam,
                                                     int nrhs, double sign, double *out) {
  constexpr int NPE = E::NPE, NGP = E::NGP, D = E::DIM, NP = mat_nparam(MAT, D);
  constexpr int LG = team_lanes<E>(), EPB = 256 / LG, NQ = NDOF * D;
  constexpr bool ISO = (MAT == TFM_MAT_ISO || MAT == TFM_MAT_HEAT_ISO);
  constexpr int NO = (MAT == TFM_MAT_ISO ? 2 : MAT == TFM_MAT_HEAT_ISO ? 1 : NP) + 1;
  constexpr int PER = (NO - 1 + LG - 1) / LG;  // general layouts: outputs per lane
  __shared__ double MP[EPB][NP];
  const int tid = threadIdx.x, slotE = tid / LG, j = tid % LG;
  const int64_t blk = xcd_block();
  const int64_t e = blk * EPB + slotE;
  const bool live = e < m.n_elem;
  const bool act = live && j < NPE;
  for (int t = tid; t < EPB * NP; t += 256) {
    const int k = t / NP, i = t - k * NP;
    const int64_t ee = blk * EPB + k;
    MP[k][i] = ee < m.n_elem ? mat_param<MAT, D>(mat, ee, i) : 0.0;
  }
  double Xj[D];
  int64_t node = 0;
  if (act) node = m.elements[e * NPE + j];
  for (int d = 0; d < D; ++d) Xj[d] = act ? m.nodes[node * D + d] : 0.0;
  __syncthreads();
  const double sc = live ? mat_scale(mat, e) : 0.0;
  const int64_t ndofs = m.n_nod * NDOF;
  for (int r = 0; r < nrhs; ++r) {
    double uj[NDOF], lj[NDOF];
    for (int a = 0; a < NDOF; ++a) {
      uj[a] = act ? u[r * ndofs + node * NDOF + a] : 0.0;
      lj[a] = act ? lam[r * ndofs + node * NDOF + a] : 0.0;
    }
    double iso_acc[2] = {0.0, 0.0}, gen_acc[PER > 0 ? PER : 1], sacc = 0.0;
    for (int o = 0; o < (PER > 0 ? PER : 1); ++o) gen_acc[o] = 0.0;
    for (int g = 0; g < NGP; ++g) {
      double Ji[D][D], bj[D], Hu[NQ], Hl[NQ], Q[NQ];
      const double det = team_geometry<E, LG>(Xj, act, j, g, Ji, bj);
      team_grad<LG, D, NDOF>(uj, bj, Hu);
      team_grad<LG, D, NDOF>(lj, bj, Hl);
      const double wd = shape_table<E>().w[g] * det;
      // d/dscale: sum_g w detJ Hl : flux(Hu)
      flux_from_grad<D, NDOF, MAT>(Hu, MP[slotE], Q);
      double en = 0.0;
      for (int q = 0; q < NQ; ++q) en += Hl[q] * Q[q];
      sacc += wd * en;
      if constexpr (MAT == TFM_MAT_ISO) {
        double tl = 0.0, tu = 0.0, dd = 0.0;
        for (int i = 0; i < D; ++i) {
          tl += Hl[i * D + i];
          tu += Hu[i * D + i];
        }
        for (int i = 0; i < D; ++i)
          for (int k = 0; k < D; ++k) dd += Hl[i * D + k] * (Hu[i * D + k] + Hu[k * D + i]);
        iso_acc[0] += wd * sc * tl * tu;
        iso_acc[1] += wd * sc * dd;
      } else if constexpr (MAT == TFM_MAT_HEAT_ISO) {
        double dd = 0.0;
        for (int i = 0; i < D; ++i) dd += Hl[i] * Hu[i];
        iso_acc[0] += wd * sc * dd;
      } else {
        // general: d/dC[i,J,k,L] = Hl[i][J] eps_u[k][L]; heat: d/dKAPPA[i,j] =
        // dT_l/dx_i dT_u/dx_j. Output c lives on lane c % LG, slot c / LG;
        // the loops are unrolled so no register array is indexed at run time.
        // mechanics pairs Hl with the strain sym(Hu): the residual depends on
        // C only through C : sym(grad u) (elasticity.py:121-123)
        double Hv[NQ];
        for (int x = 0; x < NQ; ++x) Hv[x] = Hu[x];
        if constexpr (MAT == TFM_MAT_GENERAL)
          for (int k = 0; k < D; ++k)
            for (int l = 0; l < D; ++l) Hv[k * D + l] = 0.5 * (Hu[k * D + l] + Hu[l * D + k]);
#pragma unroll
        for (int x = 0; x < NQ; ++x)
#pragma unroll
          for (int y = 0; y < NQ; ++y) {
            const int c = x * NQ + y;
            if (c % LG == j) gen_acc[c / LG] += wd * sc * Hl[x] * Hv[y];
          }
      }
    }
    if (!live) continue;
    double *o = out + ((int64_t)r * m.n_elem + e) * NO;
    if constexpr (ISO) {
      if (j == 0) {
        o[0] = sign * iso_acc[0];
        if (MAT == TFM_MAT_ISO) o[1] = sign * iso_acc[1];
        o[NO - 1] = sign * sacc;
      }
    } else {
      for (int q = 0; q < PER; ++q) {
        const int c = j + q * LG;
        if (c < NP) o[c] = sign * gen_acc[q];
      }
      if (j == 0) o[NO - 1] = sign * sacc;
    }
  }
}

// ================================================================ dispatch
#define TFM_DISPATCH(ETYPE, ...)                                                                \
  switch (ETYPE) {                                                                             \
    case TFM_QUAD1: { using ET = Quad1; __VA_ARGS__; break; }                                  \
    case TFM_TETRA1: { using ET = Tetra1; __VA_ARGS__; break; }                                \
    case TFM_HEXA1: { using ET = Hexa1; __VA_ARGS__; break; }                                  \
    case TFM_HEXA2: { using ET = Hexa2; __VA_ARGS__; break; }                                  \
    default: TFM_REQUIRE(false, "unsupported element type");                                    \
  }

static int check_mesh_mat(const tfm_mesh_t &m, const tfm_material_t &mat) {
  const bool heat = (mat.kind == TFM_MAT_HEAT_ISO || mat.kind == TFM_MAT_HEAT_GENERAL);
  TFM_REQUIRE(mat.kind >= TFM_MAT_ISO && mat.kind <= TFM_MAT_HEAT_GENERAL, "unknown material kind");
  TFM_REQUIRE(heat ? m.ndof == 1 : m.ndof == m.dim,
              "material kind does not match the DOFs per node");
  TFM_REQUIRE(mat.p0 && (mat.kind != TFM_MAT_ISO || mat.p1), "material parameters missing");
  int dim = 0, npe = 0;
  switch (m.etype) {
    case TFM_QUAD1: dim = 2; npe = 4; break;
    case TFM_TETRA1: dim = 3; npe = 4; break;
    case TFM_HEXA1: dim = 3; npe = 8; break;
    case TFM_HEXA2: dim = 3; npe = 20; break;
    default: TFM_REQUIRE(false, "unsupported element type");
  }
  TFM_REQUIRE(m.dim == dim && m.npe == npe, "mesh dim/npe do not match the element type");
  return TFM_OK;
}

// Runs BODY with compile-time ET, ND (DOFs per node) and MK (material kind).
#define TFM_DISPATCH_ALL(m, mat, ...)                                                           \
  do {                                                                                         \
    int _rc = check_mesh_mat(m, mat);                                                          \
    if (_rc != TFM_OK) return _rc;                                                             \
    const bool _heat = (mat.kind == TFM_MAT_HEAT_ISO || mat.kind == TFM_MAT_HEAT_GENERAL);    \
    TFM_DISPATCH(m.etype, {                                                                    \
      if (_heat) {                                                                             \
        constexpr int ND = 1;                                                                  \
        if (mat.kind == TFM_MAT_HEAT_ISO) { constexpr int MK = TFM_MAT_HEAT_ISO; __VA_ARGS__; } \
        else { constexpr int MK = TFM_MAT_HEAT_GENERAL; __VA_ARGS__; }                         \
      } else {                                                                                 \
        constexpr int ND = ET::DIM;                                                            \
        if (mat.kind == TFM_MAT_ISO) { constexpr int MK = TFM_MAT_ISO; __VA_ARGS__; }          \
        else { constexpr int MK = TFM_MAT_GENERAL; __VA_ARGS__; }                              \
      }                                                                                        \
    });                                                                                        \
  } while (0)

int launch_element_matrices(hipStream_t s, const tfm_mesh_t &m, const tfm_material_t &mat,
                            double *k, int32_t *bad) {
  if (m.n_elem == 0) return TFM_OK;
  TFM_DISPATCH_ALL(m, mat, {
    hipLaunchKernelGGL((k_element_matrices<ET, ND, MK>), dim3((unsigned)m.n_elem), dim3(64), 0,
                       s, m, mat, k, bad);
    TFM_CHECK_LAUNCH();
  });
  return TFM_OK;
}

int launch_assemble(hipStream_t s, const tfm_mesh_t &m, const tfm_material_t &mat,
                    const uint8_t *con, double *v64, float *v32, double *dblk, int32_t *bad,
                    bool diff32) {
  if (m.n_nod == 0) return TFM_OK;
  const int rows_only = env_int("TFM_ASSEMBLE_ROWS", 0);
  TFM_DISPATCH_ALL(m, mat, {
    constexpr int NW = tile_nodes<ET>();
    const size_t tile_shm = (sizeof(double) * ND * ND + 1) * NW * (size_t)m.max_deg;
    if (node_tiled<MK>() && !rows_only && tile_shm <= 48 * 1024) {
      constexpr int MT = node_tiled<MK>() ? MK : TFM_MAT_ISO;
      const int64_t nb = (m.n_nod + NW - 1) / NW;
      hipLaunchKernelGGL((k_assemble_nodes<ET, ND, MT>), dim3((unsigned)nb), dim3(64), tile_shm,
                         s, m, mat, con, v64, v32, dblk, bad, (int)diff32);
      TFM_CHECK_LAUNCH();
      return TFM_OK;
    }
    // one node per block: the general stiffness, node degrees whose tile would
    // not fit in LDS, or TFM_ASSEMBLE_ROWS=1
    constexpr int KCH = assemble_chunk<ET>();
    const size_t shm = sizeof(double) * ND * ND * (size_t)m.max_deg + (size_t)KCH * m.max_deg;
    TFM_REQUIRE(shm <= 48 * 1024, "node degree too large for the assembly kernel");
    hipLaunchKernelGGL((k_assemble_rows<ET, ND, MK>), dim3((unsigned)m.n_nod), dim3(64), shm, s,
                       m, mat, con, v64, v32, dblk, bad, (int)diff32);
    TFM_CHECK_LAUNCH();
  });
  return TFM_OK;
}

int launch_internal_force(hipStream_t s, const tfm_mesh_t &m, const tfm_material_t &mat,
                          const double *u, const double *flux_prev, double *f,
                          double *grad_mean, double *flux_mean, double *grad_gp,
                          double *flux_gp, double *work) {
  TFM_DISPATCH_ALL(m, mat, {
    constexpr int EPB = 256 / team_lanes<ET>();
    if (m.n_elem > 0) {
      const unsigned nb = (unsigned)((m.n_elem + EPB - 1) / EPB);
      hipLaunchKernelGGL((k_internal_force<ET, ND, MK>), dim3(nb), dim3(256), 0, s, m, mat, u,
                         flux_prev, work, grad_mean, flux_mean, grad_gp, flux_gp);
      TFM_CHECK_LAUNCH();
    }
    if (f) {
      const int64_t nt = m.n_nod * ND;
      hipLaunchKernelGGL((k_gather_nodes<ND>), dim3((unsigned)((nt + 255) / 256)), dim3(256), 0,
                         s, m, work, f);
      TFM_CHECK_LAUNCH();
    }
  });
  return TFM_OK;
}

int launch_element_matrices_gp(hipStream_t s, const tfm_mesh_t &m, const double *cgp,
                               const double *scale, double *k, int32_t *bad) {
  if (m.n_elem == 0) return TFM_OK;
  TFM_REQUIRE(m.ndof == m.dim, "a per-point tangent needs a mechanics mesh");
  TFM_DISPATCH(m.etype, {
    hipLaunchKernelGGL((k_element_matrices_gp<ET>), dim3((unsigned)m.n_elem), dim3(64), 0, s, m,
                       cgp, scale, k, bad);
    TFM_CHECK_LAUNCH();
  });
  return TFM_OK;
}

int launch_assemble_elements(hipStream_t s, const tfm_mesh_t &m, const double *k,
                             const uint8_t *con, double *v64, float *v32, double *dblk,
                             bool diff32) {
  if (m.n_nod == 0) return TFM_OK;
  const size_t shm = (size_t)m.max_inc * m.max_deg;
  TFM_REQUIRE(shm <= 48 * 1024, "node degree too large for the assembly kernel");
  switch (m.ndof) {
    case 1: hipLaunchKernelGGL(k_assemble_elements<1>, dim3((unsigned)m.n_nod), dim3(64), shm, s, m, k, con, v64, v32, dblk, (int)diff32); break;
    case 2: hipLaunchKernelGGL(k_assemble_elements<2>, dim3((unsigned)m.n_nod), dim3(64), shm, s, m, k, con, v64, v32, dblk, (int)diff32); break;
    case 3: hipLaunchKernelGGL(k_assemble_elements<3>, dim3((unsigned)m.n_nod), dim3(64), shm, s, m, k, con, v64, v32, dblk, (int)diff32); break;
    default: TFM_REQUIRE(false, "ndof must be 1, 2 or 3");
  }
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

int launch_gather(hipStream_t s, const tfm_mesh_t &m, const double *fe, double *F) {
  const int64_t nt = m.n_nod * m.ndof;
  if (nt == 0) return TFM_OK;
  const unsigned g = (unsigned)((nt + 255) / 256);
  switch (m.ndof) {
    case 1: hipLaunchKernelGGL(k_gather_nodes<1>, dim3(g), dim3(256), 0, s, m, fe, F); break;
    case 2: hipLaunchKernelGGL(k_gather_nodes<2>, dim3(g), dim3(256), 0, s, m, fe, F); break;
    case 3: hipLaunchKernelGGL(k_gather_nodes<3>, dim3(g), dim3(256), 0, s, m, fe, F); break;
    default: TFM_REQUIRE(false, "ndof must be 1, 2 or 3");
  }
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

int launch_sensitivity(hipStream_t s, const tfm_mesh_t &m, const tfm_material_t &mat,
                       const double *u, const double *lam, int nrhs, double sign, double *out) {
  if (m.n_elem == 0 || nrhs == 0) return TFM_OK;
  TFM_DISPATCH_ALL(m, mat, {
    constexpr int EPB = 256 / team_lanes<ET>();
    const unsigned nb = (unsigned)((m.n_elem + EPB - 1) / EPB);
    hipLaunchKernelGGL((k_sensitivity<ET, ND, MK>), dim3(nb), dim3(256), 0, s, m, mat, u, lam,
                       nrhs, sign, out);
    TFM_CHECK_LAUNCH();
  });
  return TFM_OK;
}

}  // namespace tfm
