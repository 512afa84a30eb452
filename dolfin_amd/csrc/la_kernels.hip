// Consumer side of the device CSR tensor: y = alpha A x + beta y on the assembler's structure
// (the role of EigenMatrix::mult, dolfin/la/EigenMatrix.cpp:230-255) and the diagonal (the
// Jacobi preconditioner of EigenKrylovSolver, dolfin/la/EigenKrylovSolver.cpp:52-56), so
// that an assembled matrix can be used without leaving HBM.
//
// One power-of-two group of G lanes per row (G from the mean row length: 16 for P1
// tetrahedra, 64 for vector P2): the group reads the row's values and columns coalesced,
// gathers x, and reduces with cross-lane shuffles. HBM-bound: 12 bytes per nonzero plus the
// row pointers and y; x is gathered through L2.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dfa {
namespace {

constexpr int kLaThreads = 256;

template <int G>
__global__ __launch_bounds__(kLaThreads) void spmv_kernel(int64_t nrows, const int32_t* __restrict__ rp,
                                                          const int32_t* __restrict__ col,
                                                          const double* __restrict__ vals,
                                                          double alpha, const double* __restrict__ x,
                                                          double beta, double* __restrict__ y) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * kLaThreads + threadIdx.x;
  const int64_t row = gid / G;
  const int ln = threadIdx.x % G;
  double s = 0.0;
  if (row < nrows) {
    const int b = rp[row], e = rp[row + 1];
    for (int k = b + ln; k < e; k += G) s = fma(vals[k], x[col[k]], s);
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  if (row < nrows && ln == 0) y[row] = beta == 0.0 ? alpha * s : fma(beta, y[row], alpha * s);
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

template <int G>
hipError_t launch_spmv_g(const CsrView& A, double alpha, const double* x, double beta, double* y,
                         hipStream_t s) {
  const int64_t threads = A.nrows * G;
  const int64_t grid = (threads + kLaThreads - 1) / kLaThreads;
  hipLaunchKernelGGL((spmv_kernel<G>), dim3(static_cast<unsigned>(grid)), dim3(kLaThreads), 0, s,
                     A.nrows, A.row_ptr, A.col, A.values, alpha, x, beta, y);
  return hipGetLastError();
}

}  // namespace

int spmv_group(int64_t nrows, int64_t nnz) {
  const double mean = nrows > 0 ? static_cast<double>(nnz) / static_cast<double>(nrows) : 1.0;
  int g = 4;
  while (g < 64 && g < mean) g *= 2;
  return g;
}

hipError_t launch_spmv(const CsrView& A, double alpha, const double* x, double beta, double* y,
                       hipStream_t s) {
  if (A.nrows == 0) return hipSuccess;
  switch (spmv_group(A.nrows, A.nnz)) {
    case 4: return launch_spmv_g<4>(A, alpha, x, beta, y, s);
    case 8: return launch_spmv_g<8>(A, alpha, x, beta, y, s);
    case 16: return launch_spmv_g<16>(A, alpha, x, beta, y, s);
    case 32: return launch_spmv_g<32>(A, alpha, x, beta, y, s);
    default: return launch_spmv_g<64>(A, alpha, x, beta, y, s);
  }
}

hipError_t launch_diagonal(const CsrView& A, double* d, hipStream_t s) {
  if (A.nrows == 0) return hipSuccess;
  const int64_t grid = (A.nrows + kLaThreads - 1) / kLaThreads;
  hipLaunchKernelGGL(diagonal_kernel, dim3(static_cast<unsigned>(grid)), dim3(kLaThreads), 0, s,
                     A.nrows, A.row_ptr, A.col, A.values, d);
  return hipGetLastError();
}

}  // namespace dfa
