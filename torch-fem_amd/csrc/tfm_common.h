// Internal declarations shared by the HIP sources of libtfm_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <string>

#include "../../include/tfm.h"

namespace tfm {

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);

#define TFM_CHECK_HIP(expr)                                                    \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::tfm::set_error(std::string("HIP error in ") + #expr + ": " +           \
                       hipGetErrorString(_e));                                 \
      return TFM_ERR_HIP;                                                      \
    }                                                                          \
  } while (0)

#define TFM_REQUIRE(cond, msg)                                                 \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::tfm::set_error(msg);                                                   \
      return TFM_ERR_ARG;                                                      \
    }                                                                          \
  } while (0)

// Check the launch just issued.
#define TFM_CHECK_LAUNCH() TFM_CHECK_HIP(hipGetLastError())

// --------------------------------------------------------------- elements
// Isoparametric element definitions. The shape functions, node order and
// Gauss points follow src/torchfem/elements.py (Quad1 491-567, Tetra1 727-790,
// Hexa1 936-1075, Hexa2 1078-1406); the tables below are evaluated at compile
// time from the closed forms, so a kernel reads dN/dxi[g][d][node] and w[g].

constexpr double kGauss = 0.57735026918962576451; // 1/sqrt(3)

// Coordinate factor of a Lagrange/serendipity shape function along one axis:
// (1 + x*c) for a vertex coordinate c = +-1, (1 - x^2) for a mid-side c = 0.
constexpr double axis_f(double x, double c) { return c == 0.0 ? 1.0 - x * x : 1.0 + x * c; }
constexpr double axis_df(double x, double c) { return c == 0.0 ? -2.0 * x : c; }

struct Quad1 {
  static constexpr int ET = TFM_QUAD1, DIM = 2, NPE = 4, NGP = 4;
  static constexpr double iso[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  // Gauss point g = i1 + 2*i2 (elements.py:559-566), weight 1
  static constexpr void gauss(int g, double *xi, double &w) {
    xi[0] = (g & 1) ? kGauss : -kGauss;
    xi[1] = (g & 2) ? kGauss : -kGauss;
    w = 1.0;
  }
  static constexpr void dshape(const double *xi, double (*dN)[NPE]) {
    for (int I = 0; I < NPE; ++I) {
      const double a = axis_f(xi[0], iso[I][0]), b = axis_f(xi[1], iso[I][1]);
      dN[0][I] = 0.25 * axis_df(xi[0], iso[I][0]) * b;
      dN[1][I] = 0.25 * a * axis_df(xi[1], iso[I][1]);
    }
  }
};

struct Tetra1 {
  static constexpr int ET = TFM_TETRA1, DIM = 3, NPE = 4, NGP = 1;
  // One point at the centroid with weight 1/6 (elements.py:782-790)
  static constexpr void gauss(int, double *xi, double &w) {
    xi[0] = xi[1] = xi[2] = 0.25;
    w = 1.0 / 6.0;
  }
  static constexpr void dshape(const double *, double (*dN)[NPE]) {
    for (int d = 0; d < 3; ++d) {
      dN[d][0] = -1.0;
      for (int I = 1; I < 4; ++I) dN[d][I] = (I - 1 == d) ? 1.0 : 0.0;
    }
  }
};

struct Hexa1 {
  static constexpr int ET = TFM_HEXA1, DIM = 3, NPE = 8, NGP = 8;
  static constexpr double iso[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1},
                                       {-1, 1, -1},  {-1, -1, 1}, {1, -1, 1},
                                       {1, 1, 1},    {-1, 1, 1}};
  // Gauss point g = i1 + 2*i2 + 4*i3 (elements.py:1000-1009), weight 1
  static constexpr void gauss(int g, double *xi, double &w) {
    xi[0] = (g & 1) ? kGauss : -kGauss;
    xi[1] = (g & 2) ? kGauss : -kGauss;
    xi[2] = (g & 4) ? kGauss : -kGauss;
    w = 1.0;
  }
  static constexpr void dshape(const double *xi, double (*dN)[NPE]) {
    for (int I = 0; I < NPE; ++I) {
      double f[3] = {}, df[3] = {};
      for (int d = 0; d < 3; ++d) {
        f[d] = axis_f(xi[d], iso[I][d]);
        df[d] = axis_df(xi[d], iso[I][d]);
      }
      dN[0][I] = 0.125 * df[0] * f[1] * f[2];
      dN[1][I] = 0.125 * f[0] * df[1] * f[2];
      dN[2][I] = 0.125 * f[0] * f[1] * df[2];
    }
  }
};

struct Hexa2 {
  static constexpr int ET = TFM_HEXA2, DIM = 3, NPE = 20, NGP = 8;
  static constexpr double iso[20][3] = {
      {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1},
      {1, -1, 1},   {1, 1, 1},   {-1, 1, 1}, {0, -1, -1}, {1, 0, -1},
      {0, 1, -1},   {-1, 0, -1}, {0, -1, 1}, {1, 0, 1},   {0, 1, 1},
      {-1, 0, 1},   {-1, -1, 0}, {1, -1, 0}, {1, 1, 0},   {-1, 1, 0}};
  // Reduced 2x2x2 Gauss rule, same order as Hexa1 (elements.py:1393-1406)
  static constexpr void gauss(int g, double *xi, double &w) { Hexa1::gauss(g, xi, w); }
  // Serendipity functions: vertices 1/8 prod(1 + x c)(sum x c - 2),
  // mid-sides 1/4 prod over the two vertex axes times (1 - x^2) on the third.
  static constexpr void dshape(const double *xi, double (*dN)[NPE]) {
    for (int I = 0; I < NPE; ++I) {
      double f[3] = {}, df[3] = {};
      for (int d = 0; d < 3; ++d) {
        f[d] = axis_f(xi[d], iso[I][d]);
        df[d] = axis_df(xi[d], iso[I][d]);
      }
      if (I < 8) {
        const double s = xi[0] * iso[I][0] + xi[1] * iso[I][1] + xi[2] * iso[I][2] - 2.0;
        const double p = f[0] * f[1] * f[2];
        dN[0][I] = 0.125 * (df[0] * f[1] * f[2] * s + p * iso[I][0]);
        dN[1][I] = 0.125 * (f[0] * df[1] * f[2] * s + p * iso[I][1]);
        dN[2][I] = 0.125 * (f[0] * f[1] * df[2] * s + p * iso[I][2]);
      } else {
        dN[0][I] = 0.25 * df[0] * f[1] * f[2];
        dN[1][I] = 0.25 * f[0] * df[1] * f[2];
        dN[2][I] = 0.25 * f[0] * f[1] * df[2];
      }
    }
  }
};

template <class E> struct ShapeTable {
  double dN[E::NGP][E::DIM][E::NPE];
  double w[E::NGP];
};

template <class E> constexpr ShapeTable<E> make_table() {
  ShapeTable<E> t{};
  for (int g = 0; g < E::NGP; ++g) {
    double xi[3] = {0.0, 0.0, 0.0};
    double w = 0.0;
    E::gauss(g, xi, w);
    E::dshape(xi, t.dN[g]);
    t.w[g] = w;
  }
  return t;
}

template <class E> __device__ const ShapeTable<E> &shape_table();

// integer tuning knob from the environment (dflt when unset or empty)
inline int env_int(const char *name, int dflt) {
  const char *v = std::getenv(name);
  return (v && *v) ? std::atoi(v) : dflt;
}

// Logical index of this block for one-block-per-item kernels: workgroups are
// dispatched to the 8 XCDs round-robin (block b on XCD b % 8), so the blocks
// of XCD x take the contiguous range [x q + min(x, r), ...) of the G = 8 q + r
// items (a bijection of [0, G)). Neighbouring items share element data,
// nodes and vector entries, which then stay in that XCD's L2.
__device__ __forceinline__ int64_t xcd_block() {
  const unsigned G = gridDim.x, b = blockIdx.x, x = b & 7u, q = G >> 3, r = G & 7u;
  return (int64_t)x * q + (x < r ? x : r) + (b >> 3);
}

// ----------------------------------------------------------- small algebra
// J[i][j] = d x_j / d xi_i (base.py:307-309); returns det and J^{-1}.
template <int D> __device__ __forceinline__ double inv_jac(const double (&J)[D][D], double (&Ji)[D][D]) {
  if constexpr (D == 2) {
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double r = 1.0 / det;
    Ji[0][0] = J[1][1] * r;
    Ji[0][1] = -J[0][1] * r;
    Ji[1][0] = -J[1][0] * r;
    Ji[1][1] = J[0][0] * r;
    return det;
  } else {
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double r = 1.0 / det;
    Ji[0][0] = c00 * r;
    Ji[1][0] = c01 * r;
    Ji[2][0] = c02 * r;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    return det;
  }
}

// Number of material parameters read per element for each layout.
__host__ __device__ constexpr int mat_nparam(int kind, int dim) {
  return kind == TFM_MAT_ISO ? 2
         : kind == TFM_MAT_GENERAL ? dim * dim * dim * dim
         : kind == TFM_MAT_HEAT_ISO ? 1
                                    : dim * dim;
}

// Block (i, k) of the stiffness coupling between column vectors bp = B[:, p]
// and bq = B[:, q]: K[(p,i),(q,k)] = sum_{J,L} bp[J] C[i,J,k,L] bq[L]
// (the einsum of base.py:1088); heat: K[p,q] = sum_ij kappa_ij bp[i] bq[j]
// (base.py:1274). mp holds the element's material parameters.
template <int DIM, int NDOF, int MAT>
__device__ __forceinline__ void coupling_block(const double *bp, const double *bq,
                                               const double *mp, double s,
                                               double *out /* [NDOF*NDOF] */) {
  if constexpr (MAT == TFM_MAT_ISO) {
    const double lam = mp[0] * s, G = mp[1] * s;
    double dot = 0.0;
    for (int J = 0; J < DIM; ++J) dot += bp[J] * bq[J];
    for (int i = 0; i < DIM; ++i)
      for (int k = 0; k < DIM; ++k)
        out[i * NDOF + k] = lam * bp[i] * bq[k] + G * bp[k] * bq[i] + (i == k ? G * dot : 0.0);
  } else if constexpr (MAT == TFM_MAT_GENERAL) {
    for (int i = 0; i < DIM; ++i)
      for (int k = 0; k < DIM; ++k) {
        double acc = 0.0;
        for (int J = 0; J < DIM; ++J)
          for (int L = 0; L < DIM; ++L)
            acc += bp[J] * mp[((i * DIM + J) * DIM + k) * DIM + L] * bq[L];
        out[i * NDOF + k] = acc * s;
      }
  } else if constexpr (MAT == TFM_MAT_HEAT_ISO) {
    double dot = 0.0;
    for (int J = 0; J < DIM; ++J) dot += bp[J] * bq[J];
    out[0] = mp[0] * s * dot;
  } else {
    double acc = 0.0;
    for (int i = 0; i < DIM; ++i)
      for (int j = 0; j < DIM; ++j) acc += mp[i * DIM + j] * bp[i] * bq[j];
    out[0] = acc * s;
  }
}

// Flux from a gradient H[a][J] (a < NDOF, J < DIM): mechanics sigma = C : sym(H)
// (elasticity.py:119-127), heat q = KAPPA . grad T (conductivity.py:91-93).
template <int DIM, int NDOF, int MAT>
__device__ __forceinline__ void flux_from_grad(const double *H, const double *mp, double *q) {
  if constexpr (MAT == TFM_MAT_ISO) {
    double tr = 0.0;
    for (int i = 0; i < DIM; ++i) tr += H[i * DIM + i];
    for (int i = 0; i < DIM; ++i)
      for (int j = 0; j < DIM; ++j)
        q[i * DIM + j] = mp[1] * (H[i * DIM + j] + H[j * DIM + i]) + (i == j ? mp[0] * tr : 0.0);
  } else if constexpr (MAT == TFM_MAT_GENERAL) {
    for (int i = 0; i < DIM; ++i)
      for (int j = 0; j < DIM; ++j) {
        double acc = 0.0;
        for (int k = 0; k < DIM; ++k)
          for (int l = 0; l < DIM; ++l)
            acc += mp[((i * DIM + j) * DIM + k) * DIM + l] * 0.5 * (H[k * DIM + l] + H[l * DIM + k]);
        q[i * DIM + j] = acc;
      }
  } else if constexpr (MAT == TFM_MAT_HEAT_ISO) {
    for (int i = 0; i < DIM; ++i) q[i] = mp[0] * H[i];
  } else {
    for (int i = 0; i < DIM; ++i) {
      double acc = 0.0;
      for (int j = 0; j < DIM; ++j) acc += mp[i * DIM + j] * H[j];
      q[i] = acc;
    }
  }
}

// ------------------------------------------------------- kernel launchers
int launch_element_matrices(hipStream_t s, const tfm_mesh_t &m, const tfm_material_t &mat,
                            double *k, int32_t *bad);
int launch_assemble(hipStream_t s, const tfm_mesh_t &m, const tfm_material_t &mat,
                    const uint8_t *con, double *v64, float *v32, double *dblk, int32_t *bad,
                    bool diff32);
int launch_internal_force(hipStream_t s, const tfm_mesh_t &m, const tfm_material_t &mat,
                          const double *u, const double *flux_prev, double *f,
                          double *grad_mean, double *flux_mean, double *grad_gp,
                          double *flux_gp, double *work);
int launch_element_matrices_gp(hipStream_t s, const tfm_mesh_t &m, const double *cgp,
                               const double *scale, double *k, int32_t *bad);
int launch_assemble_elements(hipStream_t s, const tfm_mesh_t &m, const double *k,
                             const uint8_t *con, double *v64, float *v32, double *dblk,
                             bool diff32);
int launch_gather(hipStream_t s, const tfm_mesh_t &m, const double *fe, double *F);
int launch_sensitivity(hipStream_t s, const tfm_mesh_t &m, const tfm_material_t &mat,
                       const double *u, const double *lam, int nrhs, double sign, double *out);
int launch_pattern_expand(hipStream_t s, const tfm_mesh_t &m, int64_t *rows, int64_t *cols,
                          int64_t *rowptr, int32_t *k_map, int32_t *diag_map);
int launch_spmv(hipStream_t s, const tfm_mesh_t &m, int prec, const void *vals, const void *x,
                void *y, int nrhs);

}  // namespace tfm
