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
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

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

}  // namespace

bool build_csr_device(int64_t nc, int32_t nd, const int32_t* d_cell_dofs, int64_t ndofs,
                      bool keep_diagonal, HostCSR& out, int32_t** d_row_ptr, int32_t** d_col,
                      hipStream_t s, bool download_col) {
  if (ndofs <= 0 || ndofs > INT32_MAX) return false;
  int cb = 1;
  while ((int64_t{1} << cb) < ndofs) ++cb;
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
  rocprim::double_buffer<uint64_t> keys(a, b);
  size_t sort_bytes = 0, uniq_bytes = 0;
  hip_check(rocprim::radix_sort_keys(nullptr, sort_bytes, keys, static_cast<size_t>(nkeys), 0,
                                     2 * cb, s),
            "radix_sort_keys (size)");
  uint64_t* sorted = nullptr;
  uint64_t* uniq = nullptr;
  hip_check(rocprim::unique(nullptr, uniq_bytes, a, b, static_cast<uint32_t*>(nullptr),
                            static_cast<size_t>(nkeys), rocprim::equal_to<uint64_t>(), s),
            "unique (size)");
  if (hipMalloc(&tmp.p, std::max<size_t>(std::max(sort_bytes, uniq_bytes), 16)) != hipSuccess ||
      hipMalloc(&cnt.p, sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  hip_check(rocprim::radix_sort_keys(tmp.p, sort_bytes, keys, static_cast<size_t>(nkeys), 0, 2 * cb, s),
            "radix_sort_keys");
  sorted = keys.current();
  uniq = keys.alternate();
  hip_check(rocprim::unique(tmp.p, uniq_bytes, sorted, uniq, static_cast<uint32_t*>(cnt.p),
                            static_cast<size_t>(nkeys), rocprim::equal_to<uint64_t>(), s),
            "unique");
  uint32_t nnz32 = 0;
  hip_check(hipMemcpyAsync(&nnz32, cnt.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  const int64_t nnz = nnz32;
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
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    out.row_ptr.resize(static_cast<size_t>(ndofs) + 1);
    out.col.resize(download_col ? static_cast<size_t>(nnz) : 0);
    e = hipMemcpyAsync(out.row_ptr.data(), rp, out.row_ptr.size() * sizeof(int32_t),
                       hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && nnz > 0 && download_col)
      e = hipMemcpyAsync(out.col.data(), col, out.col.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    (void)hipFree(rp);
    (void)hipFree(col);
    hip_check(e, "csr_rows_kernel / download");
  }
  *d_row_ptr = rp;
  *d_col = col;
  return true;
}

}  // namespace dfa
