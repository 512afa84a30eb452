// Consumer side of the device CSR tensor: y = alpha A x + beta y on the assembler's structure
// (the role of EigenMatrix::mult, dolfin/la/EigenMatrix.cpp:230-255) and the diagonal (the
// Jacobi preconditioner of EigenKrylovSolver, dolfin/la/EigenKrylovSolver.cpp:52-56), so
// that an assembled matrix can be used without leaving HBM.
//
// Streamed SpMV (the CSR-stream scheme of CSR-Adaptive, Greathouse & Daga, SC'14): a
// workgroup takes one row block of at most 4096 nonzeros and 256 rows, reads its values and
// columns as one contiguous coalesced stream (eight independent loads per thread in flight),
// gathers x, stages the products in LDS and reduces each row there with a power-of-two group
// of lanes. A row longer than 4096 nonzeros is a block of its own, summed by the whole
// workgroup. Workgroups are mapped so that each XCD sweeps a contiguous range of rows, which
// keeps the x gathers of neighbouring rows in one L2. HBM-bound: 12 bytes per nonzero plus
// the row pointers and y.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dfa {
namespace {

constexpr int kLaThreads = 512;
constexpr int kNpt = kSpmvBudget / kLaThreads;  // products per thread

__global__ __launch_bounds__(kLaThreads) void spmv_stream_kernel(
    int nblocks, const int32_t* __restrict__ blocks, const int32_t* __restrict__ rp,
    const int32_t* __restrict__ col, const double* __restrict__ vals, double alpha,
    const double* __restrict__ x, double beta, double* __restrict__ y) {
  __shared__ double prod[kSpmvBudget];
  __shared__ int32_t srp[kSpmvMaxRows + 1];
  __shared__ double wsum[kLaThreads / 64];
  const int q = nblocks >> 3, rem = nblocks & 7;
  const int xcd = blockIdx.x & 7;
  const int blk = xcd * q + min(xcd, rem) + static_cast<int>(blockIdx.x >> 3);
  const int tid = threadIdx.x;
  // (first row, first nonzero) pairs: the value stream starts without waiting for row_ptr
  const int2 b0 = reinterpret_cast<const int2*>(blocks)[blk];
  const int2 b1 = reinterpret_cast<const int2*>(blocks)[blk + 1];
  const int r0 = b0.x, nr = b1.x - b0.x;
  const int base = b0.y, end = b1.y;
  if (end - base > kSpmvBudget) {  // a single long row
    double s = 0.0;
    for (int k = base + tid; k < end; k += kLaThreads) s = fma(vals[k], x[col[k]], s);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((tid & 63) == 0) wsum[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kLaThreads / 64; ++w) t += wsum[w];
      y[r0] = beta == 0.0 ? alpha * t : fma(beta, y[r0], alpha * t);
    }
    return;
  }
  for (int i = tid; i <= nr; i += kLaThreads) srp[i] = rp[r0 + i] - base;
  int c[kNpt];
  double v[kNpt];
#pragma unroll
  for (int j = 0; j < kNpt; ++j) {
    const int k = base + j * kLaThreads + tid;
    // read once: streaming loads keep L2 for the x gathers
    c[j] = k < end ? __builtin_nontemporal_load(col + k) : 0;
    v[j] = k < end ? __builtin_nontemporal_load(vals + k) : 0.0;
  }
#pragma unroll
  for (int j = 0; j < kNpt; ++j) {
    const int k = j * kLaThreads + tid;
    if (base + k < end) prod[k] = v[j] * x[c[j]];
  }
  __syncthreads();
  int lg = 0;  // lanes per row: 2^lg, the largest power of two with nr * 2^lg <= threads, <= 64
  while (lg < 6 && (nr << (lg + 1)) <= kLaThreads) ++lg;
  const int L = 1 << lg, groups = kLaThreads >> lg;
  const int ln = tid & (L - 1);
  for (int lr = tid >> lg; lr - (tid >> lg) < nr; lr += groups) {  // uniform trip count
    double s = 0.0;
    if (lr < nr)
      for (int k = srp[lr] + ln; k < srp[lr + 1]; k += L) s += prod[k];
    for (int o = L >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lr < nr && ln == 0) y[r0 + lr] = beta == 0.0 ? alpha * s : fma(beta, y[r0 + lr], alpha * s);
  }
}

__global__ __launch_bounds__(kLaThreads) void diagonal_kernel(int64_t nrows, const int32_t* __restrict__ rp,
                                                              const int32_t* __restrict__ col,
                                                              const double* __restrict__ vals,
                                                              double* __restrict__ d) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kLaThreads + threadIdx.x;
  if (r >= nrows) return;
  int lo = rp[r], hi = rp[r + 1];
  double v = 0.0;
  while (lo < hi) {  // sorted columns: binary search for the diagonal entry
    const int mid = (lo + hi) >> 1;
    const int c = col[mid];
    if (c == r) {
      v = vals[mid];
      break;
    }
    if (c < r) lo = mid + 1;
    else hi = mid;
  }
  d[r] = v;
}

// Newton-increment Dirichlet values: out = x0 - g on Dirichlet dofs, 0 elsewhere.
__global__ __launch_bounds__(kLaThreads) void dirichlet_x0_kernel(int64_t n,
                                                                  const uint8_t* __restrict__ mask,
                                                                  const double* __restrict__ g,
                                                                  const double* __restrict__ x0,
                                                                  double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kLaThreads + threadIdx.x;
  if (i >= n) return;
  out[i] = mask[i] ? x0[i] - g[i] : 0.0;
}

// Slot of column c in row r of a sorted CSR row (-1 if absent), for block entry k.
__global__ __launch_bounds__(kLaThreads) void block_slots_kernel(int32_t m, const int32_t* __restrict__ rows,
                                                                 int32_t n, const int32_t* __restrict__ cols,
                                                                 const int32_t* __restrict__ rp,
                                                                 const int32_t* __restrict__ col,
                                                                 int32_t* __restrict__ slot) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kLaThreads + threadIdx.x;
  if (k >= static_cast<int64_t>(m) * n) return;
  const int32_t r = rows[k / n], c = cols[k % n];
  int lo = rp[r], hi = rp[r + 1];
  int found = -1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int cm = col[mid];
    if (cm == c) {
      found = mid;
      break;
    }
    if (cm < c) lo = mid + 1;
    else hi = mid;
  }
  slot[k] = found;
}

// values[slot[k]] += block[k] (atomic: a block may name one entry twice).
__global__ __launch_bounds__(kLaThreads) void block_add_kernel(int64_t nk, const int32_t* __restrict__ slot,
                                                               const double* __restrict__ block,
                                                               double* __restrict__ values) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kLaThreads + threadIdx.x;
  if (k >= nk) return;
  unsafeAtomicAdd(values + slot[k], block[k]);
}

}  // namespace

hipError_t launch_block_slots(int32_t m, const int32_t* rows, int32_t n, const int32_t* cols,
                              const int32_t* row_ptr, const int32_t* col, int32_t* slot,
                              hipStream_t s) {
  const int64_t nk = static_cast<int64_t>(m) * n;
  if (nk <= 0) return hipSuccess;
  hipLaunchKernelGGL(block_slots_kernel, dim3(static_cast<unsigned>((nk + kLaThreads - 1) / kLaThreads)),
                     dim3(kLaThreads), 0, s, m, rows, n, cols, row_ptr, col, slot);
  return hipGetLastError();
}

hipError_t launch_block_add(int64_t nk, const int32_t* slot, const double* block, double* values,
                            hipStream_t s) {
  if (nk <= 0) return hipSuccess;
  hipLaunchKernelGGL(block_add_kernel, dim3(static_cast<unsigned>((nk + kLaThreads - 1) / kLaThreads)),
                     dim3(kLaThreads), 0, s, nk, slot, block, values);
  return hipGetLastError();
}

hipError_t launch_dirichlet_x0(int64_t n, const uint8_t* mask, const double* g, const double* x0,
                               double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t grid = (n + kLaThreads - 1) / kLaThreads;
  hipLaunchKernelGGL(dirichlet_x0_kernel, dim3(static_cast<unsigned>(grid)), dim3(kLaThreads), 0, s,
                     n, mask, g, x0, out);
  return hipGetLastError();
}

hipError_t launch_spmv(const CsrView& A, double alpha, const double* x, double beta, double* y,
                       hipStream_t s) {
  if (A.nrows == 0) return hipSuccess;
  if (!A.blocks || A.nblocks <= 0 || A.nblocks > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(spmv_stream_kernel, dim3(static_cast<unsigned>(A.nblocks)), dim3(kLaThreads), 0,
                     s, static_cast<int>(A.nblocks), A.blocks, A.row_ptr, A.col, A.values, alpha, x,
                     beta, y);
  return hipGetLastError();
}

hipError_t launch_diagonal(const CsrView& A, double* d, hipStream_t s) {
  if (A.nrows == 0) return hipSuccess;
  const int64_t grid = (A.nrows + kLaThreads - 1) / kLaThreads;
  hipLaunchKernelGGL(diagonal_kernel, dim3(static_cast<unsigned>(grid)), dim3(kLaThreads), 0, s,
                     A.nrows, A.row_ptr, A.col, A.values, d);
  return hipGetLastError();
}

}  // namespace dfa
