// Device build of the CSR structure: the sorted unique column set of every row from
// cell_dofs x cell_dofs (+ the diagonal for keep_diagonal), i.e. SparsityPatternBuilder::build
// with diagonal_pattern(sorted) (dolfin/la/SparsityPatternBuilder.cpp:95-111,245-264,
// dolfin/la/SparsityPattern.cpp:511-542) and the layout EigenMatrix::init gives it
// (dolfin/la/EigenMatrix.cpp:384-410).
//
// Every (cell, i, j) pair becomes one 64-bit key row * 2^cb + col (cb = bits of the largest
// dof); the keys are radix sorted on the device (rocPRIM, only the 2 * cb significant bits),
// duplicates dropped, and one pass over the unique keys writes the column indices and, at
// each change of row, the row pointers of the rows in between (so rows without entries get
// empty ranges). Sorted keys give sorted columns within each row, as the host build does.
// Vector-P2 dofmaps whose dofs are node-blocked (component-major, as UFC numbers them, or
// node-major after renumbering) sort the node graph's 10 x 10 pairs per cell instead of the
// 30 x 30 dof pairs and expand each node entry to its 3 x 3 block (DFA_CSR_GENERAL=1 forces
// the general build).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include <cstdlib>

#include "assembler.h"

namespace dfa {
namespace {

constexpr int kCsrThreads = 256;

__global__ __launch_bounds__(kCsrThreads) void csr_keys_kernel(int64_t nc, int32_t nd,
                                                               const int32_t* __restrict__ cd,
                                                               int cb, uint64_t* __restrict__ keys) {
  const int64_t npair = static_cast<int64_t>(nd) * nd;
  const int64_t total = nc * npair;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kCsrThreads + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * kCsrThreads) {
    const int64_t c = k / npair;
    const int32_t ij = static_cast<int32_t>(k - c * npair);
    const int32_t i = ij / nd, j = ij - i * nd;
    const uint64_t r = static_cast<uint32_t>(cd[c * nd + i]);
    const uint64_t col = static_cast<uint32_t>(cd[c * nd + j]);
    keys[k] = r << cb | col;
  }
}

__global__ __launch_bounds__(kCsrThreads) void csr_diag_keys_kernel(int64_t ndofs, int cb,
                                                                    uint64_t* __restrict__ keys) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kCsrThreads + threadIdx.x; r < ndofs;
       r += static_cast<int64_t>(gridDim.x) * kCsrThreads)
    keys[r] = static_cast<uint64_t>(r) << cb | static_cast<uint64_t>(r);
}

// row_ptr[r] = first position of row r among the unique keys, for r in [0, ndofs]
__global__ __launch_bounds__(kCsrThreads) void csr_rows_kernel(int64_t nnz, int64_t ndofs, int cb,
                                                               const uint64_t* __restrict__ u,
                                                               int32_t* __restrict__ row_ptr,
                                                               int32_t* __restrict__ col) {
  const uint64_t mask = (uint64_t{1} << cb) - 1;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kCsrThreads + threadIdx.x; k <= nnz;
       k += static_cast<int64_t>(gridDim.x) * kCsrThreads) {
    // rows (prev, cur] start at k; prev = row of key k - 1 (or -1), cur = row of key k (or
    // ndofs past the end)
    const int64_t prev = k == 0 ? -1 : static_cast<int64_t>(u[k - 1] >> cb);
    const int64_t cur = k == nnz ? ndofs : static_cast<int64_t>(u[k] >> cb);
    for (int64_t r = prev + 1; r <= cur; ++r) row_ptr[r] = static_cast<int32_t>(k);
    if (k < nnz) col[k] = static_cast<int32_t>(u[k] & mask);
  }
}

// ---- vector P2: node-blocked dofmaps ----------------------------------------------------
// When every cell's dofs are dof(i, a) = dof(0, a) + i * s for the three components i, with
// s = ndofs / 3 and dof(0, a) < s (component-major, the UFC numbering) or s = 1 and
// dof(0, a) divisible by 3 (node-major), the sparsity is the node graph's with every node
// entry expanded to a 3 x 3 block: row (i, n) holds (j, m) for every node m adjacent to n and
// every j, sorted as j * s' + m (component-major) or 3 m + j (node-major). The node graph has
// 10 x 10 pairs per cell instead of 30 x 30, so it is sorted instead.

// flag[0] |= 1: not component-major; flag[0] |= 2: not node-major
__global__ __launch_bounds__(kCsrThreads) void vp2_layout_kernel(int64_t nc, int64_t nn,
                                                                 const int32_t* __restrict__ cd,
                                                                 int32_t* __restrict__ flag) {
  int bad = 0;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kCsrThreads + threadIdx.x; k < nc * 10;
       k += static_cast<int64_t>(gridDim.x) * kCsrThreads) {
    const int64_t c = k / 10;
    const int a = static_cast<int>(k - c * 10);
    const int64_t d0 = cd[c * 30 + a], d1 = cd[c * 30 + 10 + a], d2 = cd[c * 30 + 20 + a];
    if (!(d0 >= 0 && d0 < nn && d1 == d0 + nn && d2 == d0 + 2 * nn)) bad |= 1;
    if (!(d0 >= 0 && d0 % 3 == 0 && d1 == d0 + 1 && d2 == d0 + 2)) bad |= 2;
  }
  if (bad) atomicOr(flag, bad);
}

__global__ __launch_bounds__(kCsrThreads) void vp2_node_ids_kernel(int64_t nc, bool node_major,
                                                                   const int32_t* __restrict__ cd,
                                                                   int32_t* __restrict__ nid) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kCsrThreads + threadIdx.x; k < nc * 10;
       k += static_cast<int64_t>(gridDim.x) * kCsrThreads) {
    const int64_t c = k / 10;
    const int a = static_cast<int>(k - c * 10);
    const int32_t d0 = cd[c * 30 + a];
    nid[k] = node_major ? d0 / 3 : d0;
  }
}

// dof-level row pointers from the node graph's (nn nodes, nnz_n node entries)
__global__ __launch_bounds__(kCsrThreads) void vp2_rows_kernel(int64_t nn, bool node_major,
                                                               const int32_t* __restrict__ nrp,
                                                               int32_t* __restrict__ rp) {
  const int64_t nnz_n = nrp[nn];
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kCsrThreads + threadIdx.x; r <= 3 * nn;
       r += static_cast<int64_t>(gridDim.x) * kCsrThreads) {
    int64_t v;
    if (r == 3 * nn) {
      v = 9 * nnz_n;
    } else if (node_major) {
      const int64_t n = r / 3, i = r - 3 * n;
      v = 9 * static_cast<int64_t>(nrp[n]) + i * 3 * (nrp[n + 1] - nrp[n]);
    } else {
      const int64_t i = r / nn, n = r - i * nn;
      v = i * 3 * nnz_n + 3 * static_cast<int64_t>(nrp[n]);
    }
    rp[r] = static_cast<int32_t>(v);
  }
}

// dof-level columns: node entry e = (n, m) at position e of the node CSR fills its 3 x 3 block
__global__ __launch_bounds__(kCsrThreads) void vp2_cols_kernel(int64_t nnz_n, int64_t nn, int cb,
                                                               bool node_major,
                                                               const uint64_t* __restrict__ ukeys,
                                                               const int32_t* __restrict__ nrp,
                                                               const int32_t* __restrict__ rp,
                                                               int32_t* __restrict__ col) {
  const uint64_t mask = (uint64_t{1} << cb) - 1;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * kCsrThreads + threadIdx.x; e < nnz_n;
       e += static_cast<int64_t>(gridDim.x) * kCsrThreads) {
    const int64_t n = static_cast<int64_t>(ukeys[e] >> cb), m = static_cast<int64_t>(ukeys[e] & mask);
    const int64_t q = e - nrp[n], deg = nrp[n + 1] - nrp[n];
    for (int i = 0; i < 3; ++i) {
      const int64_t r = node_major ? 3 * n + i : i * nn + n;
      for (int j = 0; j < 3; ++j) {
        const int64_t slot = node_major ? rp[r] + 3 * q + j : rp[r] + j * deg + q;
        col[slot] = static_cast<int32_t>(node_major ? 3 * m + j : j * nn + m);
      }
    }
  }
}

unsigned grid_for(int64_t n) {
  const int64_t b = (n + kCsrThreads - 1) / kCsrThreads;
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(b, 65536)));
}

struct DeviceBuffer {
  void* p = nullptr;
  ~DeviceBuffer() {
    if (p) (void)hipFree(p);
  }
};

// Sorts the keys in a (nkeys, significant bits `bits`) with b as the second buffer and drops
// duplicates; returns the unique keys (a or b) and their count, or nullptr when the scratch
// memory cannot be allocated.
uint64_t* sort_unique(uint64_t* a, uint64_t* b, int64_t nkeys, int bits, hipStream_t s,
                      DeviceBuffer& tmp, DeviceBuffer& cnt, int64_t& count) {
  rocprim::double_buffer<uint64_t> keys(a, b);
  size_t sort_bytes = 0, uniq_bytes = 0;
  hip_check(rocprim::radix_sort_keys(nullptr, sort_bytes, keys, static_cast<size_t>(nkeys), 0, bits, s),
            "radix_sort_keys (size)");
  hip_check(rocprim::unique(nullptr, uniq_bytes, a, b, static_cast<uint32_t*>(nullptr),
                            static_cast<size_t>(nkeys), rocprim::equal_to<uint64_t>(), s),
            "unique (size)");
  if (hipMalloc(&tmp.p, std::max<size_t>(std::max(sort_bytes, uniq_bytes), 16)) != hipSuccess ||
      hipMalloc(&cnt.p, sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  hip_check(rocprim::radix_sort_keys(tmp.p, sort_bytes, keys, static_cast<size_t>(nkeys), 0, bits, s),
            "radix_sort_keys");
  uint64_t* sorted = keys.current();
  uint64_t* uniq = keys.alternate();
  hip_check(rocprim::unique(tmp.p, uniq_bytes, sorted, uniq, static_cast<uint32_t*>(cnt.p),
                            static_cast<size_t>(nkeys), rocprim::equal_to<uint64_t>(), s),
            "unique");
  uint32_t n32 = 0;
  hip_check(hipMemcpyAsync(&n32, cnt.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  count = n32;
  return uniq;
}

int bits_for(int64_t n) {
  int cb = 1;
  while ((int64_t{1} << cb) < n) ++cb;
  return cb;
}

// Copies row_ptr (and col) of the finished device CSR into out; frees both on failure.
void download_csr(int64_t ndofs, int64_t nnz, int32_t* rp, int32_t* col, HostCSR& out, bool download_col,
                  hipStream_t s) {
  out.row_ptr.resize(static_cast<size_t>(ndofs) + 1);
  out.col.resize(download_col ? static_cast<size_t>(nnz) : 0);
  hipError_t e = hipMemcpyAsync(out.row_ptr.data(), rp, out.row_ptr.size() * sizeof(int32_t),
                                hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && nnz > 0 && download_col)
    e = hipMemcpyAsync(out.col.data(), col, out.col.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(rp);
    (void)hipFree(col);
    hip_check(e, "CSR download");
  }
}

// The node-blocked vector-P2 build. Returns 1 when done, 0 when the dofmap is not
// node-blocked in either layout (the general build applies), -1 when device memory is short.
int build_vp2_csr_device(int64_t nc, const int32_t* d_cell_dofs, int64_t ndofs, HostCSR& out,
                         int32_t** d_row_ptr, int32_t** d_col, hipStream_t s, bool download_col,
                         int32_t** d_node_rp, int32_t** d_node_col) {
  const int64_t nn = ndofs / 3;
  DeviceBuffer flag, nid, ka, kb, tmp, cnt, nrp, ncol;
  if (hipMalloc(&flag.p, sizeof(int32_t)) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  hip_check(hipMemsetAsync(flag.p, 0, sizeof(int32_t), s), "hipMemsetAsync");
  hipLaunchKernelGGL(vp2_layout_kernel, dim3(grid_for(nc * 10)), dim3(kCsrThreads), 0, s, nc, nn,
                     d_cell_dofs, static_cast<int32_t*>(flag.p));
  hip_check(hipGetLastError(), "vp2_layout_kernel");
  int32_t fl = 0;
  hip_check(hipMemcpyAsync(&fl, flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  if ((fl & 1) && (fl & 2)) return 0;
  const bool node_major = (fl & 1) != 0;
  const int64_t nkeys = nc * 100;
  const size_t kbytes = static_cast<size_t>(nkeys) * sizeof(uint64_t);
  if (hipMalloc(&nid.p, static_cast<size_t>(nc) * 10 * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&ka.p, kbytes) != hipSuccess || hipMalloc(&kb.p, kbytes) != hipSuccess ||
      hipMalloc(&nrp.p, static_cast<size_t>(nn + 1) * sizeof(int32_t)) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  int32_t* ids = static_cast<int32_t*>(nid.p);
  hipLaunchKernelGGL(vp2_node_ids_kernel, dim3(grid_for(nc * 10)), dim3(kCsrThreads), 0, s, nc,
                     node_major, d_cell_dofs, ids);
  const int cb = bits_for(nn);
  hipLaunchKernelGGL(csr_keys_kernel, dim3(grid_for(nkeys)), dim3(kCsrThreads), 0, s, nc, 10, ids, cb,
                     static_cast<uint64_t*>(ka.p));
  hip_check(hipGetLastError(), "vp2 node keys");
  int64_t nnz_n = 0;
  uint64_t* uniq = sort_unique(static_cast<uint64_t*>(ka.p), static_cast<uint64_t*>(kb.p), nkeys, 2 * cb, s,
                               tmp, cnt, nnz_n);
  if (!uniq) return -1;
  if (9 * nnz_n > INT32_MAX) throw std::runtime_error("nnz exceeds int32 (la_index) range");
  if (hipMalloc(&ncol.p, std::max<size_t>(static_cast<size_t>(nnz_n) * sizeof(int32_t), 16)) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  int32_t* node_rp = static_cast<int32_t*>(nrp.p);
  hipLaunchKernelGGL(csr_rows_kernel, dim3(grid_for(nnz_n + 1)), dim3(kCsrThreads), 0, s, nnz_n, nn, cb,
                     uniq, node_rp, static_cast<int32_t*>(ncol.p));
  hip_check(hipGetLastError(), "vp2 node rows");
  const int64_t nnz = 9 * nnz_n;
  int32_t* rp = nullptr;
  int32_t* col = nullptr;
  hip_check(hipMalloc(&rp, static_cast<size_t>(ndofs + 1) * sizeof(int32_t)), "hipMalloc");
  if (hipMalloc(&col, std::max<size_t>(static_cast<size_t>(nnz) * sizeof(int32_t), 16)) != hipSuccess) {
    (void)hipFree(rp);
    hip_check(hipErrorOutOfMemory, "hipMalloc");
  }
  hipLaunchKernelGGL(vp2_rows_kernel, dim3(grid_for(ndofs + 1)), dim3(kCsrThreads), 0, s, nn, node_major,
                     node_rp, rp);
  hipLaunchKernelGGL(vp2_cols_kernel, dim3(grid_for(nnz_n)), dim3(kCsrThreads), 0, s, nnz_n, nn, cb,
                     node_major, uniq, node_rp, rp, col);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipFree(rp);
    (void)hipFree(col);
    hip_check(e, "vp2_rows_kernel / vp2_cols_kernel");
  }
  download_csr(ndofs, nnz, rp, col, out, download_col, s);
  out.node_layout = node_major ? 2 : 1;
  if (d_node_rp && d_node_col) {  // the caller keeps the node graph on the device
    *d_node_rp = static_cast<int32_t*>(nrp.p);
    *d_node_col = static_cast<int32_t*>(ncol.p);
    nrp.p = nullptr;
    ncol.p = nullptr;
    *d_row_ptr = rp;
    *d_col = col;
    return 1;
  }
  // the node graph, for host code that works per node (a ninth of the column indices)
  out.node_row_ptr.resize(static_cast<size_t>(nn) + 1);
  out.node_col.resize(static_cast<size_t>(nnz_n));
  hipError_t e2 = hipMemcpyAsync(out.node_row_ptr.data(), node_rp, out.node_row_ptr.size() * sizeof(int32_t),
                                 hipMemcpyDeviceToHost, s);
  if (e2 == hipSuccess && nnz_n > 0)
    e2 = hipMemcpyAsync(out.node_col.data(), ncol.p, out.node_col.size() * sizeof(int32_t),
                        hipMemcpyDeviceToHost, s);
  if (e2 == hipSuccess) e2 = hipStreamSynchronize(s);
  if (e2 != hipSuccess) {
    (void)hipFree(rp);
    (void)hipFree(col);
    hip_check(e2, "node CSR download");
  }
  *d_row_ptr = rp;
  *d_col = col;
  return 1;
}

}  // namespace

bool build_csr_device(int64_t nc, int32_t nd, const int32_t* d_cell_dofs, int64_t ndofs,
                      bool keep_diagonal, HostCSR& out, int32_t** d_row_ptr, int32_t** d_col,
                      hipStream_t s, bool download_col, int32_t** d_node_rp, int32_t** d_node_col) {
  if (ndofs <= 0 || ndofs > INT32_MAX) return false;
  // vector P2 with a node-blocked dofmap (keep_diagonal, which adds the diagonal of dofs in
  // no cell, takes the general build)
  if (nd == 30 && nc > 0 && ndofs % 3 == 0 && !keep_diagonal && !std::getenv("DFA_CSR_GENERAL")) {
    const int r = build_vp2_csr_device(nc, d_cell_dofs, ndofs, out, d_row_ptr, d_col, s, download_col,
                                       d_node_rp, d_node_col);
    if (r > 0) return true;
    if (r < 0) return false;
  }
  const int cb = bits_for(ndofs);
  const int64_t ncell_keys = nc * nd * nd;
  const int64_t nkeys = ncell_keys + (keep_diagonal ? ndofs : 0);
  if (nkeys <= 0 || static_cast<uint64_t>(nkeys) > 0xffffffffull) return false;
  DeviceBuffer ka, kb, tmp, cnt;
  const size_t kbytes = static_cast<size_t>(nkeys) * sizeof(uint64_t);
  if (hipMalloc(&ka.p, kbytes) != hipSuccess || hipMalloc(&kb.p, kbytes) != hipSuccess) {
    (void)hipGetLastError();  // out of device memory: the caller builds on the host
    return false;
  }
  uint64_t* a = static_cast<uint64_t*>(ka.p);
  uint64_t* b = static_cast<uint64_t*>(kb.p);
  if (ncell_keys > 0)
    hipLaunchKernelGGL(csr_keys_kernel, dim3(grid_for(ncell_keys)), dim3(kCsrThreads), 0, s, nc, nd,
                       d_cell_dofs, cb, a);
  if (keep_diagonal)
    hipLaunchKernelGGL(csr_diag_keys_kernel, dim3(grid_for(ndofs)), dim3(kCsrThreads), 0, s, ndofs,
                       cb, a + ncell_keys);
  hip_check(hipGetLastError(), "csr_keys_kernel");
  int64_t nnz = 0;
  uint64_t* uniq = sort_unique(a, b, nkeys, 2 * cb, s, tmp, cnt, nnz);
  if (!uniq) return false;
  if (nnz > INT32_MAX) throw std::runtime_error("nnz exceeds int32 (la_index) range");
  int32_t* rp = nullptr;
  int32_t* col = nullptr;
  hip_check(hipMalloc(&rp, static_cast<size_t>(ndofs + 1) * sizeof(int32_t)), "hipMalloc");
  if (hipMalloc(&col, std::max<size_t>(static_cast<size_t>(nnz) * sizeof(int32_t), 16)) != hipSuccess) {
    (void)hipFree(rp);
    hip_check(hipErrorOutOfMemory, "hipMalloc");
  }
  hipLaunchKernelGGL(csr_rows_kernel, dim3(grid_for(nnz + 1)), dim3(kCsrThreads), 0, s, nnz, ndofs,
                     cb, uniq, rp, col);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipFree(rp);
    (void)hipFree(col);
    hip_check(e, "csr_rows_kernel");
  }
  download_csr(ndofs, nnz, rp, col, out, download_col, s);
  *d_row_ptr = rp;
  *d_col = col;
  return true;
}

}  // namespace dfa
