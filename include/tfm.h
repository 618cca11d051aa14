/*
 * tfm.h - C ABI of libtfm_hip.so, the MI355X (gfx950) assemble-and-solve path
 * behind the torchfem Python API.
 *
 * Conventions (the same ones the reference uses for its only native binding,
 * the AmgX ctypes table in src/torchfem/amgx.py:100-201):
 *   - every entry point returns an int status, TFM_OK (0) on success;
 *     tfm_last_error() returns a message for the last failure of this thread;
 *   - opaque handles come from *_create(&h, ...) and are released by an
 *     explicit *_destroy(h) (amgx.py:204-208, 385-395);
 *   - device arrays are raw pointers owned by the caller (torch tensors on the
 *     caller side); the library allocates only inside explicit handles;
 *   - every device call takes the caller's hipStream_t and is stream ordered;
 *     the only host synchronisations are the convergence checks of the Krylov
 *     solver (the reference syncs there too, sparse.py:579-596).
 *
 * Index conventions follow the reference: local DOF order inside an element is
 * node-major / component-minor (base.py:73-76) and the global matrix pattern is
 * the sorted set of (row, col) pairs of all element couplings plus the full
 * diagonal, stored row-major (base.py:79-118).
 */
#ifndef TFM_H
#define TFM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *tfm_stream_t; /* a hipStream_t; NULL is the null stream */

enum {
  TFM_OK = 0,
  TFM_ERR_ARG = 1,     /* invalid argument (sizes, pointers, unsupported type) */
  TFM_ERR_HIP = 2,     /* a HIP runtime call failed */
  TFM_ERR_NOCONV = 3,  /* the Krylov solver did not reach its tolerance */
  TFM_ERR_JACOBIAN = 4 /* an element has a non-positive Jacobian determinant */
};

/* Element types (src/torchfem/elements.py). */
enum {
  TFM_QUAD1 = 1,  /* 4-node bilinear quad, 2x2 Gauss (elements.py:491-567)   */
  TFM_TETRA1 = 2, /* 4-node linear tetrahedron, 1 point (elements.py:727-790) */
  TFM_HEXA1 = 3,  /* 8-node trilinear hexahedron, 2x2x2 (elements.py:936-1075) */
  TFM_HEXA2 = 4   /* 20-node serendipity hexahedron, 2x2x2 (elements.py:1078-1406) */
};

/* Material parameter layouts. */
enum {
  /* Isotropic elasticity: C = lam I(x)I + G (dik djl + dil djk) with per-element
   * lam (p0) and G (p1). Plane stress passes lam' = E nu / (1 - nu^2)
   * (materials/elasticity.py:44-70, 164-182). */
  TFM_MAT_ISO = 1,
  /* General linear elastic tangent C[e, i, j, k, l] (p0), dim^4 per element. */
  TFM_MAT_GENERAL = 2,
  /* Isotropic conductivity KAPPA = kappa I with per-element kappa (p0)
   * (materials/conductivity.py:30-45). */
  TFM_MAT_HEAT_ISO = 3,
  /* General conductivity KAPPA[e, i, j] (p0), dim^2 per element. */
  TFM_MAT_HEAT_GENERAL = 4
};

/* Solver precision modes. */
enum {
  TFM_PREC_FP64 = 0,  /* all-FP64 PCG on FP64 matrix values                     */
  TFM_PREC_MIXED = 1  /* FP32 inner PCG on FP32 values + FP64 iterative refinement
                         on the matrix-free FP64 residual                        */
};

/* Forms of the FP32 values (tfm_assemble): the rounded FP64 values, or the
 * difference form whose nodal diagonal blocks hold the rounded FP64 row sums
 * R_n = sum_j K_nj; the FP32 products use the difference form for one or two
 * right-hand sides and the plain form for more. */
enum {
  TFM_FP32_PLAIN = 0,
  TFM_FP32_DIFF = 1
};

/* Preconditioners. */
enum {
  TFM_PC_JACOBI = 0,       /* inverse of diag(K) (cupy path, sparse.py:406-421) */
  TFM_PC_BLOCK_JACOBI = 1, /* inverse of the ndof x ndof nodal diagonal blocks  */
  TFM_PC_TWO_LEVEL = 2     /* block-Jacobi plus an additive coarse correction on
                              node aggregates (tfm_coarse_*; in the place of the
                              reference's AMG preconditioner, sparse.py:422-442) */
};

/* Mesh and node-block sparsity pattern, all arrays on the device.
 * The CSR values of node n's ndof rows are contiguous: entry (a, p, b) of
 * node n (row n*ndof+a, column adj[adj_ptr[n]+p]*ndof+b) sits at
 * val_ptr[n] + a*ndof*deg(n) + p*ndof + b, which is exactly the row-major
 * order of the reference's glob_idx (base.py:91, 112-117). A node that belongs
 * to no element ("orphan", deg == 0) holds only its ndof diagonal entries. */
typedef struct tfm_mesh {
  const double *nodes;       /* [n_nod, dim] */
  const int32_t *elements;   /* [n_elem, npe] */
  int64_t n_nod;
  int64_t n_elem;
  int32_t dim;               /* spatial dimension (2 or 3) */
  int32_t npe;               /* nodes per element */
  int32_t ndof;              /* DOFs per node (dim for mechanics, 1 for heat) */
  int32_t etype;             /* TFM_QUAD1 ... TFM_HEXA2 */
  int32_t max_deg;           /* max node degree (adjacent nodes incl. self) */
  int32_t max_inc;           /* max number of elements incident to a node */
  int64_t nnz;               /* number of stored matrix entries */
  const int32_t *adj_ptr;    /* [n_nod + 1] */
  const int32_t *adj;        /* [adj_ptr[n_nod]] sorted adjacent nodes, incl. self */
  const int64_t *val_ptr;    /* [n_nod + 1] first CSR value of the node's rows */
  const int32_t *ne_ptr;     /* [n_nod + 1] node -> element incidence */
  const int32_t *ne_elem;    /* [n_elem*npe] incident elements, ascending */
  const int32_t *ne_loc;     /* [n_elem*npe] local index of the node there */
  const uint16_t *epos;      /* [n_elem, npe, npe] position of node j in adj(node i) */
  const int16_t *adj16;      /* optional [adj_ptr[n_nod]] adj[k] - (row node), when every
                                such difference fits in 16 bits (NULL otherwise); the
                                SpMV kernels then read half the index bytes */
} tfm_mesh_t;

/* Per-element material parameters (device). */
typedef struct tfm_material {
  int32_t kind;              /* TFM_MAT_* */
  int32_t _pad;
  const double *p0;          /* see TFM_MAT_* */
  const double *p1;          /* see TFM_MAT_* */
  const double *scale;       /* [n_elem] volume scale (planar thickness) or NULL */
} tfm_material_t;

/* An assembled (or matrix-free) linear system K u = b with Dirichlet DOFs. */
typedef struct tfm_system {
  const tfm_mesh_t *mesh;
  const tfm_material_t *mat;
  const uint8_t *con;        /* [n_nod*ndof] 1 where the DOF is constrained */
  const double *val64;       /* FP64 CSR values or NULL */
  const float *val32;        /* FP32 CSR values in difference form (see tfm_assemble) or NULL */
  const double *diag_blocks; /* [n_nod, ndof, ndof] nodal diagonal blocks of K */
  int32_t val32_form;        /* TFM_FP32_* of val32: DIFF for a solve of one or two
                                right-hand sides, PLAIN for more */
  int32_t _pad;
} tfm_system_t;

typedef struct tfm_pattern tfm_pattern_t;
typedef struct tfm_pcg tfm_pcg_t;
typedef struct tfm_krylov tfm_krylov_t;
typedef struct tfm_coarse tfm_coarse_t;

/* ---------------------------------------------------------------- errors */
const char *tfm_last_error(void);
int tfm_version(void);

/* ---------------------------------------------- sparsity pattern (host) */
/* Replaces FEM.__init__'s pattern block (base.py:72-122): builds the node
 * adjacency, the node->element incidence and the element -> slot positions
 * from a host connectivity array. Integer work only, on the host. */
int tfm_pattern_create(tfm_pattern_t **h, const int64_t *elements, int64_t n_elem,
                       int32_t npe, int64_t n_nod, int32_t ndof);
/* Sizes: nnz, total adjacency length, max degree, max incidence. */
int tfm_pattern_info(const tfm_pattern_t *h, int64_t *nnz, int64_t *n_adj,
                     int32_t *max_deg, int32_t *max_inc);
/* Copies the arrays into caller host buffers sized from tfm_pattern_info. */
int tfm_pattern_export(const tfm_pattern_t *h, int32_t *adj_ptr, int32_t *adj,
                       int64_t *val_ptr, int32_t *ne_ptr, int32_t *ne_elem,
                       int32_t *ne_loc, uint16_t *epos);
int tfm_pattern_destroy(tfm_pattern_t *h);

/* ------------------------------------------------- sparsity pattern (device) */
/* The same pattern built on the device from device connectivity
 * elements [n_elem, npe] (int32), in two calls (FEM.__init__'s pattern block,
 * base.py:72-122, with integer atomics for counts and per-node sorts for
 * order, so every array equals tfm_pattern_export's bit for bit).
 * Count: fills ne_ptr [n_nod+1], ne_elem, ne_loc [n_elem*npe], adj_ptr
 * [n_nod+1] and val_ptr [n_nod+1] (int64); work is int32 scratch of at least
 * 3*n_nod + 16 + 4*(n_nod/1024 + 2) entries; info[0..3] (host) receives the
 * adjacency length, nnz, the max degree and the max incidence. At most 192
 * adjacent nodes per node. Synchronises the stream once. */
int tfm_pattern_device_count(tfm_stream_t s, const int32_t *elements, int64_t n_elem,
                             int32_t npe, int64_t n_nod, int32_t ndof, int32_t *ne_ptr,
                             int32_t *ne_elem, int32_t *ne_loc, int32_t *adj_ptr,
                             int64_t *val_ptr, int32_t *work, int64_t *info);
/* Fill: adj [info[0]], epos [n_elem*npe*npe] and, when adj16 is not NULL, the
 * 16-bit offsets adj[k] - (row node); *adj16_fits (host) tells whether all of
 * them fit (else adj16 must not be used). */
int tfm_pattern_device_fill(tfm_stream_t s, const int32_t *elements, int64_t n_elem,
                            int32_t npe, int64_t n_nod, const int32_t *ne_ptr,
                            const int32_t *ne_elem, const int32_t *adj_ptr, int32_t *adj,
                            uint16_t *epos, int16_t *adj16, int32_t *work,
                            int32_t *adj16_fits);

/* ------------------------------------------- pattern expansion (device) */
/* The reference's index arrays on the device: glob_idx rows/cols (int64,
 * base.py:112-117), CSR row pointer, k_map (int32 [n_elem*(npe*ndof)^2],
 * base.py:94-104) and diag_map (int32 [n_dofs], base.py:106-108).
 * Any output pointer may be NULL. */
int tfm_pattern_expand(tfm_stream_t s, const tfm_mesh_t *m, int64_t *rows,
                       int64_t *cols, int64_t *rowptr, int32_t *k_map,
                       int32_t *diag_map);

/* --------------------------------------------- element integration (device) */
/* Element matrices k[e] = sum_gp w detJ B^T C B (times scale), FP64
 * [n_elem, npe*ndof, npe*ndof]; replaces the BCB einsum of
 * Mechanics/Heat.integrate_material (base.py:1086-1090, 1270-1276) and
 * FEM.k0 (base.py:228-245). *bad_jac is set to 1 if some detJ <= 0. */
int tfm_element_matrices(tfm_stream_t s, const tfm_mesh_t *m,
                         const tfm_material_t *mat, double *k, int32_t *bad_jac);

/* Element matrices from a tangent given per Gauss point,
 * tangent [n_gp, n_elem, dim, dim, dim, dim] (a finite-strain material's
 * ddsdde = dP/dF, materials/hyperelasticity.py:83-132): the BCB sum of
 * Mechanics.integrate_material (base.py:1086-1090) per point, times
 * scale [n_elem] (NULL = 1). Mechanics meshes only. */
int tfm_element_matrices_gp(tfm_stream_t s, const tfm_mesh_t *m, const double *tangent,
                            const double *scale, double *k, int32_t *bad_jac);

/* Consistent element mass matrices mass [n_elem, npe*ndof, npe*ndof]:
 * FEM.integrate_mass (base.py:376-396) with compute_m = rho detJ
 * (solid.py:60-62) or rho thickness detJ (planar.py:94-96, scale = the
 * thickness, NULL = 1); rho [n_elem]. */
int tfm_element_mass(tfm_stream_t s, const tfm_mesh_t *m, const double *rho,
                     const double *scale, double *mass, int32_t *bad_jac);

/* Fused integration + deterministic assembly + Dirichlet epilogue: writes the
 * CSR values of K in FP64 and/or FP32 (either may be NULL) and the nodal
 * diagonal blocks. Replaces integrate_material's k plus
 * FEM.assemble_matrix (base.py:398-426) with no float atomics.
 * The FP32 values are the FP64 ones rounded (fp32_form TFM_FP32_PLAIN), or
 * with TFM_FP32_DIFF each node's diagonal block holds the rounded FP64
 * row-sum block R_n = sum_j K_nj instead (the "difference form": the FP32
 * products compute sum_{j != n} K_nj (x_j - x_n) + R_n x_n, which equals
 * K x exactly and keeps the FP32 operator close to the FP64 one on smooth
 * vectors). */
int tfm_assemble(tfm_stream_t s, const tfm_mesh_t *m, const tfm_material_t *mat,
                 const uint8_t *con, double *val64, float *val32,
                 double *diag_blocks, int32_t *bad_jac, int32_t fp32_form);

/* Same Dirichlet epilogue and outputs as tfm_assemble, from caller-supplied
 * element matrices k [n_elem, npe*ndof, npe*ndof] (FEM.assemble_matrix,
 * base.py:398-426, as an ordered gather instead of index_add_). */
int tfm_assemble_elements(tfm_stream_t s, const tfm_mesh_t *m, const double *k,
                          const uint8_t *con, double *val64, float *val32,
                          double *diag_blocks, int32_t fp32_form);

/* Internal forces and fluxes for a nodal field u (FP64):
 * element forces fe[e][j][a] = sum_gp w detJ B^T (flux_prev + flux(grad u))
 * written to work [n_elem, npe, ndof]; f[n_nod*ndof] (may be NULL) gathers them
 * in ascending element order (FEM.assemble_rhs, base.py:428-445). Gradients
 * and fluxes per Gauss point ([n_gp, n_elem, ndof, dim]) and their Gauss-point
 * means ([n_elem, ndof, dim]) are written where the pointer is not NULL.
 * flux_prev [n_gp, n_elem, ndof, dim] may be NULL (zero). Replaces the force
 * part of integrate_material (base.py:1050-1083, 1250-1268). */
int tfm_internal_force(tfm_stream_t s, const tfm_mesh_t *m,
                       const tfm_material_t *mat, const double *u,
                       const double *flux_prev, double *f, double *grad_mean,
                       double *flux_mean, double *grad_gp, double *flux_gp,
                       double *work);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tfm_internal_force without the work argument. The
 * element forces are staged in a static device buffer of
 * TFM_STATIC_WORK_DOUBLES doubles (32 MiB) that every GPU the library runs on
 * holds until the process ends, so n_elem * npe * ndof must not exceed it
 * (TFM_ERR_ARG otherwise). The call allocates nothing. All calls on one GPU
 * share that buffer: calls that overlap on one GPU (from other streams or
 * threads) race. */
#define TFM_STATIC_WORK_DOUBLES (4 * 1024 * 1024)
int tfm_internal_force_static(tfm_stream_t s, const tfm_mesh_t *m,
                              const tfm_material_t *mat, const double *u,
                              const double *flux_prev, double *f, double *grad_mean,
                              double *flux_mean, double *grad_gp, double *flux_gp);

/* F[n*ndof+a] = sum over the elements of node n (ascending) of
 * fe[e][local][a]: FEM.assemble_rhs (base.py:428-445) without float atomics. */
int tfm_gather_nodes(tfm_stream_t s, const tfm_mesh_t *m, const double *fe, double *F);

/* Per-element parameter sensitivities of lam^T K(p) u for nrhs pairs of nodal
 * fields (u, lam stored [nrhs, n_dofs]); out is [nrhs, n_elem, n_out] with
 *   TFM_MAT_ISO:          n_out = 3: d/dlam, d/dG, d/dscale
 *   TFM_MAT_GENERAL:      n_out = dim^4 + 1: d/dC[i,j,k,l], d/dscale
 *   TFM_MAT_HEAT_ISO:     n_out = 2: d/dkappa, d/dscale
 *   TFM_MAT_HEAT_GENERAL: n_out = dim^2 + 1: d/dKAPPA[i,j], d/dscale
 * scaled by `sign`. With sign = -1, lam the adjoint and u the solution this is
 * the implicit-function gradient -lam^T dR/dp of NewtonRaphsonAdjoint.backward
 * (sparse.py:641-724) for a linear material, with K_e recomputed, never stored. */
int tfm_sensitivity(tfm_stream_t s, const tfm_mesh_t *m, const tfm_material_t *mat,
                    const double *u, const double *lam, int32_t nrhs, double sign,
                    double *out);

/* ----------------------------------------------------- sparse kernels */
/* y = K x on the node-block CSR values, FP64 (prec 0) or FP32 (prec 1; the
 * values in the difference form, TFM_FP32_DIFF, products summed in FP64). */
int tfm_spmv(tfm_stream_t s, const tfm_mesh_t *m, int32_t prec, const void *vals,
             const void *x, void *y);

/* y = A x for a general CSR matrix (int64 row pointer, int32 columns), FP64;
 * the operator behind torchfem.sparse on matrices without a mesh. */
int tfm_csr_spmv(tfm_stream_t s, int64_t n, const int64_t *rowptr,
                 const int32_t *col, const double *val, const double *x, double *y);

/* ------------------------------------------------------------ PCG solver */
/* Workspace handle for up to nrhs simultaneous right-hand sides of size n
 * (like the reference's AmgXSolver handle, amgx.py:211-395). */
int tfm_pcg_create(tfm_pcg_t **h, int64_t n, int32_t nrhs, int32_t max_iter);
int tfm_pcg_destroy(tfm_pcg_t *h);

/* Solves K x = b for nrhs right-hand sides (b, x stored [nrhs, n], FP64; x holds
 * the initial guess on entry, cf. the warm start of sparse.py:103-128).
 * prec: TFM_PREC_FP64 or TFM_PREC_MIXED; pc: TFM_PC_*. Stops when
 * ||b - K x||_2 <= rtol ||b||_2 for every right-hand side (the criterion of
 * the reference's CG/MINRES with rtol = stol, sparse.py:406-421).
 * Mixed mode: each FP32 segment reduces the residual by inner_rtol before the
 * next FP64 refinement, and the segment that can end the solve aims at
 * final_rtol ||b||_2 (0 < final_rtol <= rtol; a margin below rtol for the error
 * of the solution); FP64 mode aims at 0.7 rtol.
 * A right-hand side with b = 0 gets x = 0 whatever the initial guess (as
 * scipy's cg/minres return for ||b|| = 0).
 * iters[nrhs] receives the total inner iterations, relres[nrhs] the final FP64
 * relative residual, outer[0] the refinement steps (mixed mode).
 * A mixed solve whose FP32 iterations break down (p.Kp <= 0 in FP32, or a
 * true residual that stops falling above rtol) continues in FP64 from the x of
 * its last refinement when sys->val64 is given; iters then counts both parts
 * (tfm_pcg_fp64_fallback returns the FP64 part).
 * Returns TFM_ERR_NOCONV when max_iter is exhausted, when p.Kp <= 0 in FP64
 * (K is not positive definite), or when the true residual stops falling above
 * rtol (attainable accuracy reached). */
int tfm_pcg_solve(tfm_pcg_t *h, tfm_stream_t s, const tfm_system_t *sys,
                  int32_t prec, int32_t pc, const double *b, double *x,
                  int32_t nrhs, double rtol, double inner_rtol, double final_rtol,
                  int32_t *iters, double *relres, int32_t *outer);

/* Event timing of the inner SpMV launches of tfm_pcg_solve on the solve's
 * stream: returns the accumulated milliseconds and launch count since the
 * previous call, resets them, and sets the sampling: enable = k > 0 times the
 * SpMV of every k-th iteration (an event pair per timed launch), 0 disables.
 * Measurement plumbing with no counterpart in the reference. */
int tfm_pcg_timing(tfm_pcg_t *h, int32_t enable, double *spmv_ms,
                   int64_t *spmv_launches);

/* FP64 iterations that finished the last tfm_pcg_solve after its FP32
 * iterations broke down (the maximum over its right-hand sides; 0 when the
 * FP32 iterations sufficed or the solve ran in FP64). No counterpart in the
 * reference, which solves in FP64 only. */
int tfm_pcg_fp64_fallback(const tfm_pcg_t *h, int32_t *fp64_iters);

/* ------------------------------------------- two-level preconditioner */
/* Coarse space of TFM_PC_TWO_LEVEL, in the place of the aggregation AMG the
 * reference reaches through AmgX (sparse.py:422-442, amgx.py:71-98): nodes
 * are grouped into n_agg aggregates and every aggregate carries the
 * rigid-body modes of its nodes about its centroid (3 translations and 3
 * rotations in 3-D, 2 + 1 in 2-D, the constant for a scalar field), zero on
 * constrained DOFs. The preconditioner is
 *   M^{-1} r = B^{-1} r + P (P^T K P)^{-1} P^T r
 * with B the nodal diagonal blocks and P the modes.
 * Device arrays of the caller (copied): agg [n_nod] the aggregate of each
 * node; agg_ptr [n_agg+1] / agg_nodes [n_nod] the nodes of each aggregate,
 * ascending; centroids [n_agg*dim]; slot_agg [n_agg*27] the aggregates
 * J that K couples to aggregate I, one per slot (-1: none); pair_off
 * [n_agg*27+1] / pair_node, pair_pos [n_pairs] the node blocks (i, p) (node
 * i, position p in its adjacency) of every (I, slot), in ascending (i, p).
 * Create once per mesh and aggregation; the mesh's arrays must outlive the
 * handle. */
int tfm_coarse_create(tfm_coarse_t **h, const tfm_mesh_t *m, int32_t n_agg,
                      const int32_t *agg, const int32_t *agg_ptr,
                      const int32_t *agg_nodes, const double *centroids,
                      const int32_t *slot_agg, const int32_t *pair_off,
                      int64_t n_pairs, const int32_t *pair_node,
                      const int32_t *pair_pos);
int tfm_coarse_destroy(tfm_coarse_t *h);
/* Coarse size n_c = n_agg * modes. */
int tfm_coarse_size(const tfm_coarse_t *h, int32_t *n_c);
/* Ac [n_c*n_c] = P^T K P, dense, from the system's FP64 values (sys->val64)
 * and Dirichlet mask; deterministic. */
int tfm_coarse_operator(tfm_stream_t s, const tfm_coarse_t *h, const tfm_system_t *sys,
                        double *Ac);
/* The caller's inverse of Ac (FP64, and FP32 for mixed-precision solves),
 * row-major n_c x n_c, read by every preconditioner application; the
 * buffers must stay valid while solves use them. */
int tfm_coarse_set_inverse(tfm_coarse_t *h, const double *inv64, const float *inv32);
/* Attaches the coarse space to a solver handle for TFM_PC_TWO_LEVEL (NULL
 * detaches). */
int tfm_pcg_set_coarse(tfm_pcg_t *p, tfm_coarse_t *c);

/* Inverse of a dense symmetric positive definite matrix, in place: a
 * [n*n] row-major on the device, 1 <= n <= 32768. work holds
 * tfm_spd_inverse_work(n) doubles; status (device int32, zeroed by the
 * caller) becomes k + 1 if the k-th pivot is not positive, and a then holds
 * no inverse. Block Gauss-Jordan with 64-column pivot blocks, three kernels
 * per block. No counterpart in the reference (AmgX factors its coarsest
 * level internally): the coarse inverse of TFM_PC_TWO_LEVEL. */
int tfm_spd_inverse_work(int32_t n, int64_t *count);
int tfm_spd_inverse(tfm_stream_t s, int32_t n, double *a, double *work, int32_t *status);

/* Preconditioned CG for a general SPD CSR matrix (FP64, Jacobi). Like
 * tfm_pcg_solve it fails with TFM_ERR_NOCONV as soon as an iteration finds
 * p.Ap <= 0 (the matrix is not positive definite). */
int tfm_csr_pcg_solve(tfm_pcg_t *h, tfm_stream_t s, int64_t n,
                      const int64_t *rowptr, const int32_t *col,
                      const double *val, const double *b, double *x, double rtol,
                      int32_t *iters, double *relres);

/* ---------------------------------------- Krylov solvers for general CSR */
/* Workspace handle of tfm_csr_minres_solve / tfm_csr_gmres_solve for systems
 * of size n; `restart` is the GMRES restart length (basis vectors per cycle),
 * at least 1 (the AmgXSolver-like handle convention, amgx.py:211-395). */
int tfm_krylov_create(tfm_krylov_t **h, int64_t n, int32_t restart);
int tfm_krylov_destroy(tfm_krylov_t *h);

/* Preconditioned MINRES for a symmetric (possibly indefinite) CSR matrix,
 * FP64, Jacobi preconditioner 1/|A_ii| (the "minres" method of sparse_solve,
 * sparse.py:406-421 on the GPU, 493-502 on the CPU). x holds the initial
 * guess; stops when ||b - A x||_2 <= rtol ||b||_2 (checked on the true
 * residual; the Lanczos recurrence restarts from x if its estimate stopped
 * early). TFM_ERR_NOCONV after max_iter iterations. */
int tfm_csr_minres_solve(tfm_krylov_t *h, tfm_stream_t s, int64_t n,
                         const int64_t *rowptr, const int32_t *col, const double *val,
                         const double *b, double *x, double rtol, int32_t max_iter,
                         int32_t *iters, double *relres);

/* Restarted GMRES for any square CSR matrix, FP64, right Jacobi
 * preconditioning, classical Gram-Schmidt with one reorthogonalisation pass:
 * the solve behind sparse_solve's "spsolve" method (sparse.py:403-404,
 * 490-492) and non-symmetric systems, to the true relative residual rtol. */
int tfm_csr_gmres_solve(tfm_krylov_t *h, tfm_stream_t s, int64_t n,
                        const int64_t *rowptr, const int32_t *col, const double *val,
                        const double *b, double *x, double rtol, int32_t max_iter,
                        int32_t *iters, double *relres);

/* ------------------------------------------------------- distributed loads */
/* Facets per element (*nf) and nodes per facet (*nfn) of the mesh's element
 * type (the facet tables of src/torchfem/elements.py). */
int tfm_facet_count(const tfm_mesh_t *m, int32_t *nf, int32_t *nfn);

/* FEM._boundary_facets (base.py:452-470): the facets whose nodes are all set
 * in mask [n_nod] and that belong to one element only, in the order of their
 * sorted node tuples (the reference's torch.unique order). Work arrays:
 * flag [n_elem*nf] int8, off [n_nod+1], bsum [ceil(n_nod/1024)+1].
 * Outputs: pos [n_elem*nf] (the facet's index in the list, or -1),
 * list [list_cap] (slot e*nf + local facet of each listed facet) and
 * *n_facets (host). Synchronises the stream once to read the count; fails
 * with TFM_ERR_ARG if it exceeds list_cap (pos is complete then, list not). */
int tfm_boundary_facets(tfm_stream_t s, const tfm_mesh_t *m, const uint8_t *mask,
                        int8_t *flag, int32_t *pos, int32_t *off, int32_t *bsum,
                        int32_t *list, int64_t list_cap, int64_t *n_facets);

/* Node ids [n_facets, nfn] (int64, the element's winding) of listed facets. */
int tfm_facet_nodes(tfm_stream_t s, const tfm_mesh_t *m, const int32_t *list,
                    int64_t n_facets, int64_t *out);

/* FEM._integrate_facet_load (base.py:472-494) on the listed facets: consistent
 * nodal loads F [n_nod, k] = sum_gp w N detJ t, detJ = sqrt(det(J J^T)).
 * kind 0: pressure load[0] along the unit outward normal (k = dim);
 * kind 1: one traction load[k] for every facet; kind 2: load [n_facets, k].
 * fc [n_facets, nfn, k] receives the facet contributions, which are summed
 * at each node over its incident elements in ascending order. */
int tfm_facet_load(tfm_stream_t s, const tfm_mesh_t *m, const uint8_t *mask,
                   const int32_t *pos, const int32_t *list, int64_t n_facets,
                   const double *load, int32_t kind, int32_t k, double *fc, double *F);

/* FEM.integrate_body_load (base.py:496-513): F [n_nod, k] from a load per
 * unit volume, load[k] for every element (per_element 0) or [n_elem, k],
 * times scale [n_elem] (planar thickness; NULL = 1). fe [n_elem, npe, k] is
 * work space. *bad_jac is set to 1 if some detJ <= 0. */
int tfm_body_load(tfm_stream_t s, const tfm_mesh_t *m, const double *scale,
                  const double *load, int32_t per_element, int32_t k, double *fe,
                  double *F, int32_t *bad_jac);

/* FEM.eval_shape_functions (base.py:293-314) at nq local points xi [nq, dim]:
 * B [nq, n_elem, dim, npe] = J^{-1} dN/dxi and detJ [nq, n_elem].
 * *bad_jac is set to 1 if some detJ <= 0. */
int tfm_shape_gradients(tfm_stream_t s, const tfm_mesh_t *m, const double *xi, int32_t nq,
                        double *B, double *detJ, int32_t *bad_jac);

#ifdef __cplusplus
}
#endif
#endif /* TFM_H */
