// The node-block sparsity pattern built on the device (FEM.__init__'s
// pattern block, base.py:72-122): node -> element incidence, the sorted node
// adjacency, the CSR value offsets and the element -> slot positions.
//
// Integer work only. Counts use integer atomics (whose totals do not depend
// on the order), and every list is put into its final order by a per-node
// sort, so the result is the same on every run and identical to the host
// builder's: incident elements ascending (then by local index), adjacency
// ascending, and a node's rows holding deg(n) * ndof entries each (a node
// of no element keeps its ndof diagonal entries).
#include "tfm_common.h"

#include <algorithm>
#include <vector>

namespace tfm {
namespace {

constexpr int PB = 256;
constexpr int PMAXDEG = 192;  // most adjacent nodes per node (incl. itself)
constexpr int PSB = 1024;     // counts per block of the scans

unsigned pblocks(int64_t n, int bs = PB) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

// ------------------------------------------------------------------ scans
// Exclusive prefix sum of cnt[0..n) into out[0..n] (out[n] = total) in three
// passes; T is int32 or int64, the counts int32.
template <typename T>
__global__ void __launch_bounds__(256) k_pscan_blocks(const int32_t *cnt, int64_t n, T *bsum) {
  __shared__ T red[256];
  const int64_t b0 = (int64_t)blockIdx.x * PSB;
  T acc = 0;
  for (int i = threadIdx.x; i < PSB; i += 256)
    if (b0 + i < n) acc += (T)cnt[b0 + i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

template <typename T>
__global__ void __launch_bounds__(1024) k_pscan_top(T *bsum, int nb) {
  __shared__ T tmp[1024];
  T carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const T v = i < nb ? bsum[i] : 0;
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      const T add = (int)threadIdx.x >= d ? tmp[threadIdx.x - d] : 0;
      __syncthreads();
      tmp[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + tmp[threadIdx.x] - v;
    const T total = tmp[1023];
    __syncthreads();
    carry += total;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

template <typename T>
__global__ void __launch_bounds__(256) k_pscan_apply(const int32_t *cnt, int64_t n, const T *bsum,
                                                     T *out) {
  __shared__ T part[256];
  const int64_t b0 = (int64_t)blockIdx.x * PSB + 4 * threadIdx.x;
  T v[4], sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] = b0 + i < n ? (T)cnt[b0 + i] : 0;
    sum += v[i];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const T add = (int)threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  T run = bsum[blockIdx.x] + part[threadIdx.x] - sum;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (b0 + i < n) out[b0 + i] = run;
    run += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = bsum[gridDim.x];
}

template <typename T>
int pscan(hipStream_t s, const int32_t *cnt, int64_t n, T *bsum, T *out) {
  const unsigned nb = (unsigned)std::max<int64_t>(1, (n + PSB - 1) / PSB);
  hipLaunchKernelGGL(k_pscan_blocks<T>, dim3(nb), dim3(256), 0, s, cnt, n, bsum);
  TFM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_pscan_top<T>, dim3(1), dim3(1024), 0, s, bsum, (int)nb);
  TFM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_pscan_apply<T>, dim3(nb), dim3(256), 0, s, cnt, n, (const T *)bsum, out);
  TFM_CHECK_LAUNCH();
  return TFM_OK;
}

// --------------------------------------------------------------- incidence
__global__ void k_inc_count(const int32_t *el, int64_t nent, int64_t n_nod, int32_t *cnt,
                            int32_t *bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nent) return;
  const int32_t n = el[t];
  if (n < 0 || n >= n_nod) {
    *bad = 1;
    return;
  }
  atomicAdd(&cnt[n], 1);
}

__global__ void k_inc_fill(const int32_t *el, int64_t nent, int npe, const int32_t *ptr,
                           int32_t *cursor, int32_t *ne_elem, int32_t *ne_loc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nent) return;
  const int32_t n = el[t];
  const int32_t k = ptr[n] + atomicAdd(&cursor[n], 1);
  ne_elem[k] = (int32_t)(t / npe);
  ne_loc[k] = (int32_t)(t % npe);
}

// Incident elements of each node ascending (then by local index): a fixed
// order whatever order the atomics filled them in.
__global__ void k_inc_sort(int64_t n_nod, const int32_t *ptr, int32_t *ne_elem, int32_t *ne_loc,
                           int32_t *max_inc) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nod) return;
  const int32_t a = ptr[n], b = ptr[n + 1];
  for (int32_t i = a + 1; i < b; ++i) {
    const int32_t e = ne_elem[i], l = ne_loc[i];
    int32_t j = i - 1;
    while (j >= a && (ne_elem[j] > e || (ne_elem[j] == e && ne_loc[j] > l))) {
      ne_elem[j + 1] = ne_elem[j];
      ne_loc[j + 1] = ne_loc[j];
      --j;
    }
    ne_elem[j + 1] = e;
    ne_loc[j + 1] = l;
  }
  atomicMax(max_inc, b - a);
}

// ---------------------------------------------------------------- adjacency
// The sorted set of nodes that share an element with n (n included when it
// belongs to an element), built by insertion with binary search in a local
// list of at most PMAXDEG entries. Returns its size, -1 on overflow.
__device__ int node_union(const int32_t *el, int npe, const int32_t *ne_ptr,
                          const int32_t *ne_elem, int64_t n, int32_t *lst) {
  int deg = 0;
  for (int32_t k = ne_ptr[n]; k < ne_ptr[n + 1]; ++k) {
    const int32_t *en = el + (int64_t)ne_elem[k] * npe;
    for (int j = 0; j < npe; ++j) {
      const int32_t c = en[j];
      int lo = 0, hi = deg;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (lst[mid] < c) lo = mid + 1;
        else hi = mid;
      }
      if (lo < deg && lst[lo] == c) continue;
      if (deg == PMAXDEG) return -1;
      for (int q = deg; q > lo; --q) lst[q] = lst[q - 1];
      lst[lo] = c;
      ++deg;
    }
  }
  return deg;
}

// deg[n] and the value count of n's rows (deg * ndof^2, or ndof for a node
// of no element)
__global__ void k_adj_count(const int32_t *el, int npe, int64_t n_nod, const int32_t *ne_ptr,
                            const int32_t *ne_elem, int ndof, int32_t *deg, int32_t *nval,
                            int32_t *max_deg, int32_t *bad) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nod) return;
  int32_t lst[PMAXDEG];
  const int d = node_union(el, npe, ne_ptr, ne_elem, n, lst);
  if (d < 0) {
    *bad = 2;
    deg[n] = 0;
    nval[n] = ndof;
    return;
  }
  deg[n] = d;
  nval[n] = d > 0 ? d * ndof * ndof : ndof;
  atomicMax(max_deg, d);
}

// adj (and the 16-bit offsets adj - n when adj16 is given; *wide is set if
// some offset does not fit)
__global__ void k_adj_fill(const int32_t *el, int npe, int64_t n_nod, const int32_t *ne_ptr,
                           const int32_t *ne_elem, const int32_t *adj_ptr, int32_t *adj,
                           int16_t *adj16, int32_t *wide) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nod) return;
  int32_t lst[PMAXDEG];
  const int d = node_union(el, npe, ne_ptr, ne_elem, n, lst);
  const int32_t a0 = adj_ptr[n];
  for (int q = 0; q < d; ++q) {
    adj[a0 + q] = lst[q];
    if (adj16) {
      const int64_t delta = (int64_t)lst[q] - n;
      if (delta < -32768 || delta > 32767) *wide = 1;
      adj16[a0 + q] = (int16_t)delta;
    }
  }
}

// epos[e][i][j]: position of element node j in the adjacency of node i
__global__ void k_epos(const int32_t *el, int64_t n_elem, int npe, const int32_t *adj_ptr,
                       const int32_t *adj, uint16_t *epos) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_elem * npe * npe) return;
  const int64_t e = t / (npe * npe);
  const int r = (int)(t - e * npe * npe), i = r / npe, j = r - i * npe;
  const int32_t ni = el[e * npe + i], nj = el[e * npe + j];
  int lo = adj_ptr[ni], hi = adj_ptr[ni + 1];
  const int base = lo;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (adj[mid] < nj) lo = mid + 1;
    else hi = mid;
  }
  epos[t] = (uint16_t)(lo - base);
}

}  // namespace
}  // namespace tfm

using namespace tfm;

extern "C" int tfm_pattern_device_count(tfm_stream_t stream, const int32_t *elements,
                                        int64_t n_elem, int32_t npe, int64_t n_nod, int32_t ndof,
                                        int32_t *ne_ptr, int32_t *ne_elem, int32_t *ne_loc,
                                        int32_t *adj_ptr, int64_t *val_ptr, int32_t *work,
                                        int64_t *info) {
  TFM_REQUIRE(n_nod >= 0 && n_elem >= 0 && npe >= 1 && ndof >= 1, "bad pattern sizes");
  TFM_REQUIRE(ne_ptr && adj_ptr && val_ptr && work && info, "null argument");
  TFM_REQUIRE(n_elem == 0 || (elements && ne_elem && ne_loc), "null element arrays");
  hipStream_t s = (hipStream_t)stream;
  const int64_t nent = n_elem * npe;
  // work layout (int32): cnt/deg [n_nod], nval [n_nod], cursor [n_nod],
  // flags [4] (bad, max_inc, max_deg, spare), bsum [2 * nb] (int64 view)
  int32_t *cnt = work, *nval = work + n_nod, *cursor = work + 2 * n_nod, *flags = work + 3 * n_nod;
  int64_t *bsum64 = reinterpret_cast<int64_t *>(work + ((3 * n_nod + 4 + 1) & ~(int64_t)1));
  int32_t *bsum32 = reinterpret_cast<int32_t *>(bsum64);
  TFM_CHECK_HIP(hipMemsetAsync(work, 0, sizeof(int32_t) * (3 * n_nod + 4), s));
  if (nent > 0) {
    hipLaunchKernelGGL(k_inc_count, dim3(pblocks(nent)), dim3(PB), 0, s, elements, nent, n_nod,
                       cnt, flags);
    TFM_CHECK_LAUNCH();
  }
  int rc = pscan<int32_t>(s, cnt, n_nod, bsum32, ne_ptr);
  if (rc != TFM_OK) return rc;
  if (nent > 0) {
    hipLaunchKernelGGL(k_inc_fill, dim3(pblocks(nent)), dim3(PB), 0, s, elements, nent, npe,
                       (const int32_t *)ne_ptr, cursor, ne_elem, ne_loc);
    TFM_CHECK_LAUNCH();
  }
  if (n_nod > 0) {
    hipLaunchKernelGGL(k_inc_sort, dim3(pblocks(n_nod)), dim3(PB), 0, s, n_nod,
                       (const int32_t *)ne_ptr, ne_elem, ne_loc, flags + 1);
    TFM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_adj_count, dim3(pblocks(n_nod, 64)), dim3(64), 0, s, elements, (int)npe,
                       n_nod, (const int32_t *)ne_ptr, (const int32_t *)ne_elem, (int)ndof, cnt,
                       nval, flags + 2, flags);
    TFM_CHECK_LAUNCH();
  }
  rc = pscan<int32_t>(s, cnt, n_nod, bsum32, adj_ptr);
  if (rc != TFM_OK) return rc;
  rc = pscan<int64_t>(s, nval, n_nod, bsum64, val_ptr);
  if (rc != TFM_OK) return rc;
  int32_t hf[4] = {0, 0, 0, 0};
  int32_t n_adj = 0;
  int64_t nnz = 0;
  TFM_CHECK_HIP(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipMemcpyAsync(&n_adj, adj_ptr + n_nod, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipMemcpyAsync(&nnz, val_ptr + n_nod, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipStreamSynchronize(s));
  TFM_REQUIRE(hf[0] != 1, "element connectivity references a node outside [0, n_nod)");
  TFM_REQUIRE(hf[0] != 2, "a node has more than 192 adjacent nodes");
  info[0] = n_adj;
  info[1] = nnz;
  info[2] = hf[2];  // max degree
  info[3] = hf[1];  // max incidence
  return TFM_OK;
}

extern "C" int tfm_pattern_device_fill(tfm_stream_t stream, const int32_t *elements,
                                       int64_t n_elem, int32_t npe, int64_t n_nod,
                                       const int32_t *ne_ptr, const int32_t *ne_elem,
                                       const int32_t *adj_ptr, int32_t *adj, uint16_t *epos,
                                       int16_t *adj16, int32_t *work, int32_t *adj16_fits) {
  TFM_REQUIRE(ne_ptr && adj_ptr && work && adj16_fits, "null argument");
  hipStream_t s = (hipStream_t)stream;
  TFM_CHECK_HIP(hipMemsetAsync(work, 0, sizeof(int32_t), s));
  if (n_nod > 0) {
    hipLaunchKernelGGL(k_adj_fill, dim3(pblocks(n_nod, 64)), dim3(64), 0, s, elements, (int)npe,
                       n_nod, ne_ptr, ne_elem, adj_ptr, adj, adj16, work);
    TFM_CHECK_LAUNCH();
  }
  const int64_t ne = n_elem * npe * npe;
  if (ne > 0) {
    hipLaunchKernelGGL(k_epos, dim3(pblocks(ne)), dim3(PB), 0, s, elements, n_elem, (int)npe,
                       adj_ptr, (const int32_t *)adj, epos);
    TFM_CHECK_LAUNCH();
  }
  int32_t wide = 0;
  TFM_CHECK_HIP(hipMemcpyAsync(&wide, work, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  TFM_CHECK_HIP(hipStreamSynchronize(s));
  *adj16_fits = adj16 != nullptr && wide == 0;
  return TFM_OK;
}
