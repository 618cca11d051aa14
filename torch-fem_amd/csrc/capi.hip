// C-ABI entry points of libtfm_hip.so and the host-side sparsity-pattern
// builder (integer work only).
#include "tfm_common.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

namespace tfm {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

}  // namespace tfm

using namespace tfm;

extern "C" const char *tfm_last_error(void) { return g_last_error.c_str(); }

extern "C" int tfm_version(void) { return 1; }

// ===================================================================
// Sparsity pattern (host)
//
// The reference (base.py:72-122) forms the packed keys (row << 32 | col) of
// every element DOF pair, sorts and uniques them together with the full
// diagonal, and maps each element entry to its slot with searchsorted. Every
// element couples all DOFs of all its nodes, so the same set is the
// Kronecker product of the node adjacency (nodes sharing an element) with
// the full ndof x ndof block, plus the bare diagonal of nodes that belong to
// no element. It is built here per node: the adjacency of a node is the
// sorted union of its incident elements' nodes, which gives the columns of
// the node's rows in ascending order, i.e. the same row-major order as the
// reference's sorted keys.
// ===================================================================
struct tfm_pattern {
  int64_t n_nod = 0, n_elem = 0;
  int32_t npe = 0, ndof = 0;
  int32_t max_deg = 0, max_inc = 0;
  int64_t nnz = 0;
  std::vector<int32_t> adj_ptr, adj, ne_ptr, ne_elem, ne_loc;
  std::vector<int64_t> val_ptr;
  std::vector<uint16_t> epos;
};

namespace {

template <class F> void parallel_for(int64_t n, F f) {
  const int64_t nt = std::max<int64_t>(
      1, std::min<int64_t>((int64_t)std::thread::hardware_concurrency(), (n + 65535) / 65536));
  const int64_t nthreads = std::min<int64_t>(nt, 16);
  if (nthreads <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int64_t t = 0; t < nthreads; ++t) {
    const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo < hi) th.emplace_back([=] { f(lo, hi); });
  }
  for (auto &x : th) x.join();
}

}  // namespace

extern "C" int tfm_pattern_create(tfm_pattern_t **out, const int64_t *elements, int64_t n_elem,
                                  int32_t npe, int64_t n_nod, int32_t ndof) {
  TFM_REQUIRE(out != nullptr, "null handle pointer");
  TFM_REQUIRE(n_elem >= 0 && n_nod >= 0 && npe > 0 && ndof > 0, "invalid pattern sizes");
  TFM_REQUIRE(n_elem == 0 || elements != nullptr, "null connectivity");
  TFM_REQUIRE(n_nod * ndof < (int64_t)INT32_MAX, "too many DOFs for int32 indices");
  for (int64_t t = 0; t < n_elem * npe; ++t)
    TFM_REQUIRE(elements[t] >= 0 && elements[t] < n_nod, "element references a node out of range");

  auto *P = new tfm_pattern();
  P->n_nod = n_nod;
  P->n_elem = n_elem;
  P->npe = npe;
  P->ndof = ndof;

  // node -> element incidence, elements ascending per node
  P->ne_ptr.assign(n_nod + 1, 0);
  for (int64_t t = 0; t < n_elem * npe; ++t) P->ne_ptr[elements[t] + 1]++;
  for (int64_t i = 0; i < n_nod; ++i) P->ne_ptr[i + 1] += P->ne_ptr[i];
  P->ne_elem.resize(n_elem * npe);
  P->ne_loc.resize(n_elem * npe);
  {
    std::vector<int32_t> fill(P->ne_ptr.begin(), P->ne_ptr.end() - 1);
    for (int64_t e = 0; e < n_elem; ++e)
      for (int j = 0; j < npe; ++j) {
        const int64_t nd = elements[e * npe + j];
        const int32_t k = fill[nd]++;
        P->ne_elem[k] = (int32_t)e;
        P->ne_loc[k] = j;
      }
  }
  for (int64_t i = 0; i < n_nod; ++i)
    P->max_inc = std::max(P->max_inc, P->ne_ptr[i + 1] - P->ne_ptr[i]);

  // adjacency: sorted unique nodes of the incident elements
  std::vector<int32_t> deg(n_nod, 0);
  std::vector<std::vector<int32_t>> lists(n_nod);
  parallel_for(n_nod, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> buf;
    for (int64_t i = lo; i < hi; ++i) {
      buf.clear();
      for (int k = P->ne_ptr[i]; k < P->ne_ptr[i + 1]; ++k) {
        const int64_t e = P->ne_elem[k];
        for (int j = 0; j < npe; ++j) buf.push_back((int32_t)elements[e * npe + j]);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      lists[i] = buf;
      deg[i] = (int32_t)buf.size();
    }
  });
  P->adj_ptr.assign(n_nod + 1, 0);
  for (int64_t i = 0; i < n_nod; ++i) P->adj_ptr[i + 1] = P->adj_ptr[i] + deg[i];
  P->adj.resize(P->adj_ptr[n_nod]);
  P->val_ptr.assign(n_nod + 1, 0);
  for (int64_t i = 0; i < n_nod; ++i) {
    std::copy(lists[i].begin(), lists[i].end(), P->adj.begin() + P->adj_ptr[i]);
    P->max_deg = std::max(P->max_deg, deg[i]);
    // a node in some element owns ndof rows of ndof*deg entries; an orphan
    // node keeps only its diagonal (base.py:89-90)
    P->val_ptr[i + 1] = P->val_ptr[i] + (deg[i] > 0 ? (int64_t)ndof * ndof * deg[i] : ndof);
    std::vector<int32_t>().swap(lists[i]);
  }
  P->nnz = P->val_ptr[n_nod];
  if (P->max_deg > 65535) {
    delete P;
    set_error("node degree exceeds 65535");
    return TFM_ERR_ARG;
  }

  // epos[e][i][j]: position of node j of e in the adjacency of node i of e
  P->epos.resize(n_elem * npe * npe);
  parallel_for(n_elem, [&](int64_t lo, int64_t hi) {
    for (int64_t e = lo; e < hi; ++e)
      for (int i = 0; i < npe; ++i) {
        const int64_t ni = elements[e * npe + i];
        const int32_t *b = P->adj.data() + P->adj_ptr[ni];
        const int32_t *en = P->adj.data() + P->adj_ptr[ni + 1];
        for (int j = 0; j < npe; ++j) {
          const int32_t nj = (int32_t)elements[e * npe + j];
          P->epos[(e * npe + i) * npe + j] = (uint16_t)(std::lower_bound(b, en, nj) - b);
        }
      }
  });
  *out = P;
  return TFM_OK;
}

extern "C" int tfm_pattern_info(const tfm_pattern_t *h, int64_t *nnz, int64_t *n_adj,
                                int32_t *max_deg, int32_t *max_inc) {
  TFM_REQUIRE(h != nullptr, "null pattern handle");
  if (nnz) *nnz = h->nnz;
  if (n_adj) *n_adj = h->adj_ptr.empty() ? 0 : h->adj_ptr.back();
  if (max_deg) *max_deg = h->max_deg;
  if (max_inc) *max_inc = h->max_inc;
  return TFM_OK;
}

extern "C" int tfm_pattern_export(const tfm_pattern_t *h, int32_t *adj_ptr, int32_t *adj,
                                  int64_t *val_ptr, int32_t *ne_ptr, int32_t *ne_elem,
                                  int32_t *ne_loc, uint16_t *epos) {
  TFM_REQUIRE(h != nullptr, "null pattern handle");
  auto cp = [](auto &v, auto *dst) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  cp(h->adj_ptr, adj_ptr);
  cp(h->adj, adj);
  cp(h->val_ptr, val_ptr);
  cp(h->ne_ptr, ne_ptr);
  cp(h->ne_elem, ne_elem);
  cp(h->ne_loc, ne_loc);
  cp(h->epos, epos);
  return TFM_OK;
}

extern "C" int tfm_pattern_destroy(tfm_pattern_t *h) {
  delete h;
  return TFM_OK;
}

// ===================================================================
// Device entry points
// ===================================================================
static int check_mesh(const tfm_mesh_t *m) {
  TFM_REQUIRE(m != nullptr, "null mesh");
  TFM_REQUIRE(m->n_nod >= 0 && m->n_elem >= 0, "negative mesh sizes");
  TFM_REQUIRE(m->ndof >= 1 && m->ndof <= 3, "ndof must be 1, 2 or 3");
  TFM_REQUIRE(m->adj_ptr && m->adj && m->val_ptr && m->ne_ptr && m->ne_elem && m->ne_loc &&
                  m->epos,
              "mesh pattern arrays missing");
  TFM_REQUIRE(m->n_elem == 0 || (m->nodes && m->elements), "mesh geometry missing");
  return TFM_OK;
}

#define TFM_TRY(expr)                                                                          \
  do {                                                                                         \
    int _rc = (expr);                                                                          \
    if (_rc != TFM_OK) return _rc;                                                             \
  } while (0)

extern "C" int tfm_pattern_expand(tfm_stream_t s, const tfm_mesh_t *m, int64_t *rows,
                                  int64_t *cols, int64_t *rowptr, int32_t *k_map,
                                  int32_t *diag_map) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(m->nnz < (int64_t)INT32_MAX || (!k_map && !diag_map),
              "k_map/diag_map are int32 and need nnz < 2^31");
  return launch_pattern_expand((hipStream_t)s, *m, rows, cols, rowptr, k_map, diag_map);
}

extern "C" int tfm_element_matrices(tfm_stream_t s, const tfm_mesh_t *m,
                                    const tfm_material_t *mat, double *k, int32_t *bad_jac) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(mat && k && bad_jac, "null argument");
  return launch_element_matrices((hipStream_t)s, *m, *mat, k, bad_jac);
}

extern "C" int tfm_element_matrices_gp(tfm_stream_t s, const tfm_mesh_t *m,
                                       const double *tangent, const double *scale, double *k,
                                       int32_t *bad_jac) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(tangent && k && bad_jac, "null argument");
  return launch_element_matrices_gp((hipStream_t)s, *m, tangent, scale, k, bad_jac);
}

extern "C" int tfm_assemble(tfm_stream_t s, const tfm_mesh_t *m, const tfm_material_t *mat,
                            const uint8_t *con, double *val64, float *val32,
                            double *diag_blocks, int32_t *bad_jac, int32_t fp32_form) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(mat && con && diag_blocks && bad_jac, "null argument");
  TFM_REQUIRE(val64 || val32, "no output values requested");
  TFM_REQUIRE(fp32_form == TFM_FP32_PLAIN || fp32_form == TFM_FP32_DIFF, "unknown FP32 form");
  return launch_assemble((hipStream_t)s, *m, *mat, con, val64, val32, diag_blocks, bad_jac,
                         fp32_form == TFM_FP32_DIFF);
}

extern "C" int tfm_assemble_elements(tfm_stream_t s, const tfm_mesh_t *m, const double *k,
                                     const uint8_t *con, double *val64, float *val32,
                                     double *diag_blocks, int32_t fp32_form) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(k && con && diag_blocks, "null argument");
  TFM_REQUIRE(val64 || val32, "no output values requested");
  TFM_REQUIRE(fp32_form == TFM_FP32_PLAIN || fp32_form == TFM_FP32_DIFF, "unknown FP32 form");
  return launch_assemble_elements((hipStream_t)s, *m, k, con, val64, val32, diag_blocks,
                                  fp32_form == TFM_FP32_DIFF);
}

extern "C" int tfm_internal_force(tfm_stream_t s, const tfm_mesh_t *m,
                                  const tfm_material_t *mat, const double *u,
                                  const double *flux_prev, double *f, double *grad_mean,
                                  double *flux_mean, double *grad_gp, double *flux_gp,
                                  double *work) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(mat && u && work, "null argument");
  return launch_internal_force((hipStream_t)s, *m, *mat, u, flux_prev, f, grad_mean, flux_mean,
                               grad_gp, flux_gp, work);
}

// ------------------------------------------------- allocation-free variant
// tfm_internal_force_static stages the element forces in this file-scope
// device array instead of a caller's work buffer. Host code never passes the
// array's name to a kernel: each device's copy has its own address, taken
// once per device with hipGetSymbolAddress and kept in g_scratch_ptr.
static __device__ __align__(256) double scratch[TFM_STATIC_WORK_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch array (looked up on its first use).
int scratch_ptr(double **out) {
  int dev = 0;
  TFM_CHECK_HIP(hipGetDevice(&dev));
  TFM_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    double *p = nullptr;
    TFM_CHECK_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    TFM_REQUIRE(p != nullptr, "no device address for the static work buffer");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TFM_OK;
}
}  // namespace

extern "C" int tfm_internal_force_static(tfm_stream_t s, const tfm_mesh_t *m,
                                         const tfm_material_t *mat, const double *u,
                                         const double *flux_prev, double *f,
                                         double *grad_mean, double *flux_mean,
                                         double *grad_gp, double *flux_gp) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(mat && u, "null argument");
  TFM_REQUIRE(m->n_elem * m->npe * m->ndof <= (int64_t)TFM_STATIC_WORK_DOUBLES,
              "mesh too large for the static work buffer of tfm_internal_force_static "
              "(n_elem * npe * ndof must not exceed TFM_STATIC_WORK_DOUBLES)");
  double *work = nullptr;
  TFM_TRY(scratch_ptr(&work));
  // every element writes all of its npe * ndof entries before the gather
  // reads them, so the buffer needs no clearing between calls
  return launch_internal_force((hipStream_t)s, *m, *mat, u, flux_prev, f, grad_mean, flux_mean,
                               grad_gp, flux_gp, work);
}

extern "C" int tfm_gather_nodes(tfm_stream_t s, const tfm_mesh_t *m, const double *fe,
                                double *F) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(fe && F, "null argument");
  return launch_gather((hipStream_t)s, *m, fe, F);
}

extern "C" int tfm_sensitivity(tfm_stream_t s, const tfm_mesh_t *m, const tfm_material_t *mat,
                               const double *u, const double *lam, int32_t nrhs, double sign,
                               double *out) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(mat && u && lam && out, "null argument");
  TFM_REQUIRE(nrhs >= 1, "nrhs must be positive");
  return launch_sensitivity((hipStream_t)s, *m, *mat, u, lam, nrhs, sign, out);
}

extern "C" int tfm_spmv(tfm_stream_t s, const tfm_mesh_t *m, int32_t prec, const void *vals,
                        const void *x, void *y) {
  TFM_TRY(check_mesh(m));
  TFM_REQUIRE(vals && x && y, "null argument");
  TFM_REQUIRE(prec == 0 || prec == 1, "prec must be 0 (FP64) or 1 (FP32)");
  return launch_spmv((hipStream_t)s, *m, prec, vals, x, y, 1);
}
