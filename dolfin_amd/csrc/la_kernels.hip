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

#include <climits>

#include "kernels.h"

namespace dfa {
namespace {

constexpr int kLaThreads = 512;
constexpr int kNpt = kSpmvBudget / kLaThreads;  // products per thread

__global__ __launch_bounds__(kLaThreads) void spmv_stream_kernel(
    int nblocks, const int32_t* __restrict__ blocks, const int32_t* __restrict__ rp,
    const int32_t* __restrict__ col, const int32_t* __restrict__ blk_col0, const uint16_t* __restrict__ col16,
    const double* __restrict__ vals, double alpha, const double* __restrict__ x, double beta,
    double* __restrict__ y) {
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
  // Every load is unconditional (indices past the block clamped to its last nonzero, whose
  // products are never stored), so that all 2 x kNpt loads of a thread are in flight at once;
  // values and columns are read once: streaming loads keep L2 for the x gathers.
  const int last = end - 1;
  const int c0 = col16 ? blk_col0[blk] : -1;
  if (end == base) {  // rows without entries: nothing to read
#pragma unroll
    for (int j = 0; j < kNpt; ++j) {
      c[j] = 0;
      v[j] = 0.0;
    }
  } else if (c0 >= 0) {  // 2-byte column offsets
    uint32_t o[kNpt];
#pragma unroll
    for (int j = 0; j < kNpt; ++j) {
      const int k = min(base + j * kLaThreads + tid, last);
      o[j] = __builtin_nontemporal_load(col16 + k);
      v[j] = __builtin_nontemporal_load(vals + k);
    }
#pragma unroll
    for (int j = 0; j < kNpt; ++j) c[j] = c0 + static_cast<int>(o[j]);
  } else {
#pragma unroll
    for (int j = 0; j < kNpt; ++j) {
      const int k = min(base + j * kLaThreads + tid, last);
      c[j] = __builtin_nontemporal_load(col + k);
      v[j] = __builtin_nontemporal_load(vals + k);
    }
  }
  double xv[kNpt];  // gathered unconditionally too (clamped entries name a valid column)
#pragma unroll
  for (int j = 0; j < kNpt; ++j) xv[j] = x[c[j]];
#pragma unroll
  for (int j = 0; j < kNpt; ++j) {
    const int k = j * kLaThreads + tid;
    if (base + k < end) prod[k] = v[j] * xv[j];
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

// Node-blocked SpMV for vector-P2 structures expanded from the node graph: with nodes n, m and
// components i, j, the dof rows (i, n) hold the 3 x 3 blocks (i, j) of every node entry
// (n, m), at row_ptr(i, n) + j d + q (component-major, row (i, n) = i nn + n, column
// j nn + m) or row_ptr(3 n + i) + 3 q + j (node-major, column 3 m + j), for the q-th of the d
// node entries of node n; the row starts follow from the node row pointers. Each node column
// is read once for its nine values (the dof-level column array, 4 bytes per value, is not
// read) and x once per node entry and component. A workgroup takes a chunk of consecutive
// nodes (at most kNodeBudget node entries): the node columns and the chunk's node pointers
// are loaded together, every entry finds its node by a binary search in LDS, gathers x and
// its nine values, stages the three row products in LDS, and each of the chunk's 3 x nodes
// rows is summed there by a power-of-two group of lanes. Workgroups are mapped so that each
// XCD sweeps a contiguous range of chunks.
constexpr int kNodeThreads = 512;
constexpr int kNodeEpt = kNodeBudget / kNodeThreads;  // node entries per thread
template <bool NODE_MAJOR>
__global__ __launch_bounds__(kNodeThreads) void spmv_node_kernel(
    int nchunks, const int32_t* __restrict__ chunks, int64_t nn, int64_t nnz_n, const int32_t* __restrict__ nrp,
    const int32_t* __restrict__ ncol, const double* __restrict__ vals, double alpha,
    const double* __restrict__ x, double beta, double* __restrict__ y) {
  __shared__ double prod[3 * kNodeBudget];
  __shared__ int32_t snp[kNodeMaxNodes + 1];
  const int q8 = nchunks >> 3, rem = nchunks & 7;
  const int xcd = blockIdx.x & 7;
  const int ch = xcd * q8 + min(xcd, rem) + static_cast<int>(blockIdx.x >> 3);
  const int tid = threadIdx.x;
  const int2 c0 = reinterpret_cast<const int2*>(chunks)[ch];
  const int2 c1 = reinterpret_cast<const int2*>(chunks)[ch + 1];
  const int n0 = c0.x, nnodes = c1.x - c0.x, E0 = c0.y, E = c1.y - c0.y;
  // value and x indices fit 32 bits: nnz and the dof count are at most INT32_MAX
  const uint32_t nn32 = static_cast<uint32_t>(nn), nnz32 = static_cast<uint32_t>(nnz_n);
  if (tid <= nnodes) snp[tid] = nrp[n0 + tid] - E0;
  // unconditional loads (entries past the chunk clamped to its last, never stored), so that a
  // thread's loads are all in flight at once
  uint32_t m[kNodeEpt];
#pragma unroll
  for (int h = 0; h < kNodeEpt; ++h) {
    const int e = min(tid + h * kNodeThreads, E - 1);
    m[h] = static_cast<uint32_t>(__builtin_nontemporal_load(ncol + E0 + e));
  }
  double xm[kNodeEpt][3];
#pragma unroll
  for (int h = 0; h < kNodeEpt; ++h)
#pragma unroll
    for (int j = 0; j < 3; ++j) xm[h][j] = NODE_MAJOR ? x[3u * m[h] + j] : x[j * nn32 + m[h]];
  __syncthreads();
  double v[kNodeEpt][9];
#pragma unroll
  for (int h = 0; h < kNodeEpt; ++h) {
    const int e = min(tid + h * kNodeThreads, E - 1);
    int a = 0, b = nnodes - 1;  // the node ln with snp[ln] <= e < snp[ln + 1]
    while (a < b) {
      const int mid = (a + b + 1) >> 1;
      if (snp[mid] <= e) a = mid;
      else b = mid - 1;
    }
    const uint32_t s0 = static_cast<uint32_t>(snp[a]);
    const uint32_t d = static_cast<uint32_t>(snp[a + 1]) - s0, q = static_cast<uint32_t>(e) - s0;
    const uint32_t ent = static_cast<uint32_t>(E0) + s0;  // first node entry of node n0 + a
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const uint32_t base = NODE_MAJOR ? 9u * ent + 3u * i * d : 3u * i * nnz32 + 3u * ent;
#pragma unroll
      for (int j = 0; j < 3; ++j)
        v[h][3 * i + j] = __builtin_nontemporal_load(vals + (NODE_MAJOR ? base + 3u * q + j : base + j * d + q));
    }
  }
#pragma unroll
  for (int h = 0; h < kNodeEpt; ++h) {
    const int e = tid + h * kNodeThreads;
    if (e < E)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        prod[i * kNodeBudget + e] =
            fma(v[h][3 * i + 2], xm[h][2], fma(v[h][3 * i + 1], xm[h][1], v[h][3 * i] * xm[h][0]));
  }
  __syncthreads();
  // rows (i, node ln), r = i * nnodes + ln; lanes per row: the largest power of two with
  // rows * 2^lg <= threads, at most 32
  const int nrows = 3 * nnodes;
  int lg = 0;
  while (lg < 5 && (nrows << (lg + 1)) <= kNodeThreads) ++lg;
  const int L = 1 << lg, groups = kNodeThreads >> lg;
  const int ln0 = tid & (L - 1);
  for (int r = tid >> lg; r - (tid >> lg) < nrows; r += groups) {  // uniform trip count
    double sum = 0.0;
    int i = 0, ln = 0;
    if (r < nrows) {
      i = r / nnodes;
      ln = r - i * nnodes;
      for (int e = snp[ln] + ln0; e < snp[ln + 1]; e += L) sum += prod[i * kNodeBudget + e];
    }
    for (int o = L >> 1; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    if (r < nrows && ln0 == 0) {
      const uint32_t n = static_cast<uint32_t>(n0 + ln);
      const uint32_t row = NODE_MAJOR ? 3u * n + i : i * nn32 + n;
      y[row] = beta == 0.0 ? alpha * sum : fma(beta, y[row], alpha * sum);
    }
  }
}

// Narrow columns of each streamed row block: the block's smallest column and the columns as
// offsets from it when they span at most 65536 (else -1: the block reads the 4-byte columns).
// A single row longer than a block keeps the 4-byte columns.
__global__ __launch_bounds__(kLaThreads) void narrow_columns_kernel(const int32_t* __restrict__ blocks,
                                                                   const int32_t* __restrict__ col,
                                                                   int32_t* __restrict__ blk_col0,
                                                                   uint16_t* __restrict__ col16) {
  __shared__ int wmin[kLaThreads / 64], wmax[kLaThreads / 64];
  const int blk = blockIdx.x, tid = threadIdx.x;
  const int base = blocks[2 * blk + 1], end = blocks[2 * blk + 3];
  int lo = INT32_MAX, hi = INT32_MIN;
  for (int k = base + tid; k < end; k += kLaThreads) {
    lo = min(lo, col[k]);
    hi = max(hi, col[k]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  if ((tid & 63) == 0) {
    wmin[tid >> 6] = lo;
    wmax[tid >> 6] = hi;
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < kLaThreads / 64; ++w) {
    lo = min(lo, wmin[w]);
    hi = max(hi, wmax[w]);
  }
  const bool narrow = end - base <= kSpmvBudget && end > base &&
                      static_cast<int64_t>(hi) - static_cast<int64_t>(lo) <= 65535;
  if (tid == 0) blk_col0[blk] = narrow ? lo : -1;
  if (narrow)
    for (int k = base + tid; k < end; k += kLaThreads) col16[k] = static_cast<uint16_t>(col[k] - lo);
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

// ---- conjugate gradients -------------------------------------------------------------------
// Per iteration: Ap = A p (SpMV above), then three vector kernels that each write one partial
// sum per workgroup. The kernel that needs a sum (alpha = r.z / p.Ap in the update, beta =
// r.z' / r.z in the direction update) reduces the previous kernel's partials itself, in
// every workgroup and in index order, so all workgroups get the same value (deterministic;
// the kernel boundary makes the partials visible, so no device-scope fence is needed, which on
// several XCDs would write back L2, and no one-workgroup kernel runs in between). Workgroup 0
// records the scalars in CgState; r.z is double-buffered by iteration parity because the
// direction kernel's workgroups read the old value while workgroup 0 stores the new one. A
// converged or broken-down state makes every later kernel return at once, so the iterate is
// frozen until the host reads the state.
constexpr int kCgThreads = 256;

__device__ __forceinline__ double cg_block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kCgThreads / 64; ++w) t += red[w];
  __syncthreads();
  return t;  // every thread
}

// the workgroup's NV sums into part[q * gridDim.x + blockIdx.x]
template <int NV>
__device__ __forceinline__ void cg_partials(const double (&v)[NV], double* __restrict__ part) {
  __shared__ double red[kCgThreads / 64];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double t = cg_block_sum(v[q], red);
    if (threadIdx.x == 0) part[static_cast<int64_t>(q) * gridDim.x + blockIdx.x] = t;
  }
}

// the sums over g workgroups' partials, in workgroup order (one workgroup; every thread)
template <int NV>
__device__ __forceinline__ void cg_totals(const double* __restrict__ part, int g, double (&tot)[NV]) {
  __shared__ double red[kCgThreads / 64];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double a = 0.0;
    for (int w = threadIdx.x; w < g; w += kCgThreads) a += part[static_cast<int64_t>(q) * g + w];
    tot[q] = cg_block_sum(a, red);
  }
}

__global__ __launch_bounds__(kCgThreads) void cg_init_finish(const double* __restrict__ part, int g, CgState* st) {
  double tot[3];
  cg_totals<3>(part, g, tot);
  if (threadIdx.x == 0) {
    st->rzs[1] = tot[0];  // iteration 1
    st->rr = tot[1];
    st->bb = tot[2];
  }
}

// the p.Ap partials, apart from the update's r.z' and r.r partials that the direction kernel
// reads while the next iteration's p.Ap is not yet written
__device__ __forceinline__ double* cg_dot_part(double* part) { return part + 2 * kCgMaxGroups; }

// The vector kernels take four elements per thread at a time (stride = the grid): all loads
// of a batch are issued before the arithmetic and the stores, then a remainder loop. J: the
// Jacobi preconditioner (z = dinv r), else z = r.
#define CG_LOOP(i, s, n)                                                              \
  const int64_t s = static_cast<int64_t>(gridDim.x) * kCgThreads;                     \
  int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x;            \
  for (; i + 3 * s < n; i += 4 * s)

// r = b - A x is in r: z = D r, p = z; partials of rz = r.z, rr = r.r, bb = b.b
template <bool J>
__global__ __launch_bounds__(kCgThreads) void cg_init_kernel(int64_t n, const double* __restrict__ b,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ dinv, double* __restrict__ p,
                                                             double* __restrict__ part) {
  double v[3] = {0.0, 0.0, 0.0};
  CG_LOOP(i, s, n) {
    double rv[4], bv[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      rv[u] = r[i + u * s];
      bv[u] = b[i + u * s];
      dv[u] = J ? dinv[i + u * s] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double zi = J ? dv[u] * rv[u] : rv[u];
      p[i + u * s] = zi;
      v[0] = fma(rv[u], zi, v[0]);
      v[1] = fma(rv[u], rv[u], v[1]);
      v[2] = fma(bv[u], bv[u], v[2]);
    }
  }
  for (; i < n; i += s) {
    const double ri = r[i], zi = J ? dinv[i] * ri : ri;
    p[i] = zi;
    v[0] = fma(ri, zi, v[0]);
    v[1] = fma(ri, ri, v[1]);
    v[2] = fma(b[i], b[i], v[2]);
  }
  cg_partials<3>(v, part);
}

// partials of p.Ap
__global__ __launch_bounds__(kCgThreads) void cg_dot_kernel(int64_t n, const double* __restrict__ p,
                                                            const double* __restrict__ ap, double* __restrict__ part,
                                                            const CgState* st) {
  if (st->done >= 0 || st->broke) return;
  double v[1] = {0.0};
  CG_LOOP(i, s, n) {
    double pv[4], av[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      pv[u] = p[i + u * s];
      av[u] = ap[i + u * s];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) v[0] = fma(pv[u], av[u], v[0]);
  }
  for (; i < n; i += s) v[0] = fma(p[i], ap[i], v[0]);
  cg_partials<1>(v, cg_dot_part(part));
}

// alpha = r.z / p.Ap (breakdown when p.Ap = 0); x += alpha p, r -= alpha Ap; partials of
// r.z' = r.(D r) and r.r
template <bool J>
__global__ __launch_bounds__(kCgThreads) void cg_update_kernel(int64_t n, int64_t it, double* __restrict__ x,
                                                               const double* __restrict__ p,
                                                               double* __restrict__ r,
                                                               const double* __restrict__ ap,
                                                               const double* __restrict__ dinv,
                                                               double* __restrict__ part, CgState* st) {
  if (st->done >= 0 || st->broke) return;
  double pap[1];
  cg_totals<1>(cg_dot_part(part), static_cast<int>(gridDim.x), pap);
  if (pap[0] == 0.0) {  // the same in every workgroup
    if (blockIdx.x == 0 && threadIdx.x == 0) st->broke = 1;
    return;
  }
  const double alpha = st->rzs[it & 1] / pap[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = alpha;
  double v[2] = {0.0, 0.0};
  CG_LOOP(i, s, n) {
    double xv[4], pv[4], rv[4], av[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xv[u] = x[i + u * s];
      pv[u] = p[i + u * s];
      rv[u] = r[i + u * s];
      av[u] = ap[i + u * s];
      dv[u] = J ? dinv[i + u * s] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      x[i + u * s] = fma(alpha, pv[u], xv[u]);
      const double ri = fma(-alpha, av[u], rv[u]);
      r[i + u * s] = ri;
      const double zi = J ? dv[u] * ri : ri;
      v[0] = fma(ri, zi, v[0]);
      v[1] = fma(ri, ri, v[1]);
    }
  }
  for (; i < n; i += s) {
    x[i] = fma(alpha, p[i], x[i]);
    const double ri = fma(-alpha, ap[i], r[i]);
    r[i] = ri;
    const double zi = J ? dinv[i] * ri : ri;
    v[0] = fma(ri, zi, v[0]);
    v[1] = fma(ri, ri, v[1]);
  }
  cg_partials<2>(v, part);
}

// beta = r.z' / r.z, converged at `it` when r.r <= tol^2; else p = D r + beta p
template <bool J>
__global__ __launch_bounds__(kCgThreads) void cg_dir_kernel(int64_t n, int64_t it, const double* __restrict__ r,
                                                            const double* __restrict__ dinv, double* __restrict__ p,
                                                            const double* __restrict__ part, CgState* st) {
  if (st->done >= 0 || st->broke) return;
  double tot[2];
  cg_totals<2>(part, static_cast<int>(gridDim.x), tot);
  const double rz = st->rzs[it & 1];
  const double beta = rz != 0.0 ? tot[0] / rz : 0.0;
  const bool converged = tot[1] <= st->tol2;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rzs[(it + 1) & 1] = tot[0];
    st->rr = tot[1];
    st->beta = beta;
    if (converged) st->done = it;
  }
  if (converged) return;  // the iterate is final; p is not needed
  CG_LOOP(i, s, n) {
    double rv[4], pv[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      rv[u] = r[i + u * s];
      pv[u] = p[i + u * s];
      dv[u] = J ? dinv[i + u * s] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) p[i + u * s] = fma(beta, pv[u], J ? dv[u] * rv[u] : rv[u]);
  }
  for (; i < n; i += s) p[i] = fma(beta, p[i], J ? dinv[i] * r[i] : r[i]);
}
#undef CG_LOOP

// dinv = 1 / d where d != 0, else 1 (the Jacobi preconditioner)
__global__ __launch_bounds__(kCgThreads) void cg_jacobi_kernel(int64_t n, double* __restrict__ d) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kCgThreads + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kCgThreads)
    d[i] = d[i] != 0.0 ? 1.0 / d[i] : 1.0;
}

// Number of cells holding each dof (the diagonal SystemAssembler leaves on a Dirichlet row:
// 1 per cell).
__global__ __launch_bounds__(kLaThreads) void dof_cell_count_kernel(int64_t nk, const int32_t* __restrict__ cd,
                                                                   int32_t* __restrict__ cnt) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kLaThreads + threadIdx.x;
  if (k < nk) atomicAdd(&cnt[cd[k]], 1);
}

// SystemAssembler's per-cell Dirichlet elimination (SystemAssembler::apply_bc,
// SystemAssembler.cpp:1051-1075) summed over the cells, applied to the assembled square system
// (A, b) without it: a Dirichlet row i becomes cnt(i) on the diagonal and 0 elsewhere with
// b_i = cnt(i) g_i (every cell holding i sets its element row to e_i and b_e,i = g_i); for a
// free row j, the entry A_ji of a Dirichlet column moves to the right-hand side,
// b_j -= A_ji g_i, and becomes 0 (the element columns are zeroed after the lifting). The
// free rows' diagonals are unchanged. A fixed grid of L-lane groups (L = 16 for short rows,
// 64 for long ones); group q owns the rows q, q + G, q + 2G, ... (G groups), so the contiguous
// Dirichlet rows of a mesh face spread over many groups. A group reads the masks of L of its
// rows at once (one per lane), then works through its Dirichlet rows among them with one entry
// per lane. A_ji (the sparsity is symmetric) comes from the transpose map tmap (the slot of
// A_ji for each A_ij) when TM, else from a binary search in row j. Free rows take several
// liftings: atomic additions; A_ji is read only when g_i lifts it (g_i != 0 and b wanted).
template <int L, bool TM>
__global__ __launch_bounds__(kLaThreads) void dirichlet_rows_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                                   const int32_t* __restrict__ col,
                                                                   const int32_t* __restrict__ tmap,
                                                                   const uint8_t* __restrict__ mask,
                                                                   const double* __restrict__ g,
                                                                   const int32_t* __restrict__ cnt,
                                                                   double* __restrict__ values,
                                                                   double* __restrict__ b) {
  const int ln = static_cast<int>(threadIdx.x & (L - 1));
  const int shift = static_cast<int>(threadIdx.x & (63 & ~(L - 1)));  // the group's bits in the ballot
  const uint64_t gmask = L == 64 ? ~uint64_t{0} : ((uint64_t{1} << L) - 1);
  const int64_t G = static_cast<int64_t>(gridDim.x) * (kLaThreads / L);
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kLaThreads / L) + (threadIdx.x / L);
  const int64_t rounds = (n + L * G - 1) / (L * G);  // the same for every group
  for (int64_t t = 0; t < rounds; ++t) {
    const int64_t mine = q + (L * t + ln) * G;
    const bool on = mine < n && mask[mine];
    uint64_t bits = (static_cast<uint64_t>(__ballot(on)) >> shift) & gmask;
    while (bits) {
      const int64_t i = q + (L * t + __ffsll(static_cast<unsigned long long>(bits)) - 1) * G;
      bits &= bits - 1;
      const double gi = g[i];
      const int32_t r1 = rp[i + 1];
      for (int32_t k = rp[i] + ln; k < r1; k += L) {
        const int32_t j = col[k];
        if (j == i) {
          values[k] = static_cast<double>(cnt[i]);
          continue;
        }
        values[k] = 0.0;
        if (mask[j]) continue;
        int32_t tk;
        if constexpr (TM) {
          tk = tmap[k];
        } else {
          int32_t lo = rp[j], hi = rp[j + 1];
          const int32_t end = hi;
          while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (col[mid] < i) lo = mid + 1;
            else hi = mid;
          }
          tk = lo < end && col[lo] == static_cast<int32_t>(i) ? lo : -1;
        }
        if (tk < 0) continue;
        if (b && gi != 0.0) unsafeAtomicAdd(b + j, -values[tk] * gi);  // A_ji read only when it lifts
        values[tk] = 0.0;
      }
      if (b && ln == 0) b[i] = static_cast<double>(cnt[i]) * gi;
    }
  }
}

// Dirichlet conditions on a vector-P2 structure expanded from the node graph (the layout of
// spmv_node_kernel: dof rows (i, n), columns (j, m), value of node entry q of node n at
// node_slot). Three passes that read only the node rows involved:
//   node_bc_bits_kernel    bits[n] = the Dirichlet components of node n; touched[n] = 0;
//   dirichlet_node_rows_kernel  each Dirichlet row (i, n) gets 0 in every entry and cnt on its
//                          diagonal, b = cnt g; every node m in such a row that has a free row
//                          is marked touched;
//   dirichlet_node_pull_kernel  each free row (i, m) of a touched node reads and zeroes its
//                          entries in Dirichlet columns and subtracts their sum of A g from b,
//                          reduced over the lanes in a fixed order (no atomics, and no search
//                          for the transposed entry since the pass runs over the free rows).
// The second and third find their nodes by a ballot over the flag bytes of 64 nodes per wave,
// the nodes of a wave strided over the whole range (Dirichlet nodes cluster on the boundary, so
// this spreads them evenly over the waves), and work on several nodes at once, one per group of
// lanes. Values are read and written with streaming hints so that the scattered lines do not
// displace the assembly's inputs from the caches (measured: without them the next assembly
// pass ran 20 % slower).
template <bool NODE_MAJOR>
__device__ __forceinline__ int64_t node_dof(int64_t nn, int64_t node, int c) {
  return NODE_MAJOR ? 3 * node + c : c * nn + node;
}
template <bool NODE_MAJOR>
__device__ __forceinline__ int64_t node_slot(int64_t nnz_n, int32_t n0, int32_t d, int i, int j, int32_t q) {
  if constexpr (NODE_MAJOR) return 9 * static_cast<int64_t>(n0) + 3 * (i * static_cast<int64_t>(d) + q) + j;
  else return 3 * (i * nnz_n + n0) + j * static_cast<int64_t>(d) + q;
}

template <bool NODE_MAJOR>
__global__ __launch_bounds__(kLaThreads) void node_bc_bits_kernel(int64_t nn, const uint8_t* __restrict__ mask,
                                                                 uint8_t* __restrict__ bits,
                                                                 uint8_t* __restrict__ touched) {
  const int64_t n = static_cast<int64_t>(blockIdx.x) * kLaThreads + threadIdx.x;
  if (n >= nn) return;
  int v = 0;
  for (int c = 0; c < 3; ++c) v |= mask[node_dof<NODE_MAJOR>(nn, n, c)] ? 1 << c : 0;
  bits[n] = static_cast<uint8_t>(v);
  touched[n] = 0;
}

// The nodes whose flag byte is nonzero: wave w of the G in the grid tests nodes w + (64 t + lane) G
// and takes the flagged ones 64 / L at a time, calling body(n, lane in the group, live) in every
// group of L lanes together (live false in a group left without a node, n is then 0), so that
// the body's shuffles stay within a group.
template <int L, class Body>
__device__ __forceinline__ void for_flagged_nodes(int64_t nn, const uint8_t* __restrict__ flag, Body&& body) {
  const int ln = static_cast<int>(threadIdx.x & 63);
  const int grp = ln / L;
  const int64_t G = static_cast<int64_t>(gridDim.x) * (kLaThreads / 64);
  const int64_t wv = static_cast<int64_t>(blockIdx.x) * (kLaThreads / 64) + (threadIdx.x >> 6);
  const int64_t rounds = (nn + 64 * G - 1) / (64 * G);  // the same for every wave
  for (int64_t t = 0; t < rounds; ++t) {
    const int64_t mine = wv + (64 * t + ln) * G;
    uint64_t on = static_cast<uint64_t>(__ballot(mine < nn && flag[mine] != 0));
    while (on) {
      int l = -1;
#pragma unroll
      for (int k = 0; k < 64 / L; ++k) {
        if (!on) break;
        const int f = __ffsll(static_cast<unsigned long long>(on)) - 1;
        on &= on - 1;
        if (k == grp) l = f;
      }
      body(l >= 0 ? wv + (64 * t + l) * G : 0, ln & (L - 1), l >= 0);
    }
  }
}

template <bool NODE_MAJOR>
__global__ __launch_bounds__(kLaThreads) void dirichlet_node_rows_kernel(
    int64_t nn, int64_t nnz_n, const int32_t* __restrict__ nrp, const int32_t* __restrict__ ncol,
    const uint8_t* __restrict__ bits, const double* __restrict__ g, const int32_t* __restrict__ cnt,
    double* __restrict__ values, double* __restrict__ b, uint8_t* __restrict__ touched) {
  constexpr int L = 32;
  for_flagged_nodes<L>(nn, bits, [&](int64_t n, int ln, bool live) {
    if (!live) return;
    const int am = bits[n];
    const int32_t n0 = nrp[n], d = nrp[n + 1] - n0;
    double dg[3];
    for (int i = 0; i < 3; ++i) dg[i] = am >> i & 1 ? static_cast<double>(cnt[node_dof<NODE_MAJOR>(nn, n, i)]) : 0.0;
    // the node entries: neighbours with a free row marked, the diagonal node entry found
    int32_t qs = -1;
    for (int32_t q = ln; q < d; q += L) {
      const int32_t m = __builtin_nontemporal_load(ncol + n0 + q);
      if (m == n) qs = q;
      if (bits[m] != 7) touched[m] = 1;
    }
    for (int o = L >> 1; o > 0; o >>= 1) qs = max(qs, __shfl_xor(qs, o));
    // a Dirichlet row (i, n) is 3 d contiguous values (its node entries' three columns each):
    // stored with consecutive lanes on consecutive values (one instruction per L values
    // instead of one per column component; measured: 120 -> 43 us without these stores)
    for (int i = 0; i < 3; ++i) {
      if (!(am >> i & 1)) continue;
      const int64_t base = node_slot<NODE_MAJOR>(nnz_n, n0, d, i, 0, 0);
      const int diag = NODE_MAJOR ? 3 * qs + i : i * d + qs;
      for (int jl = ln; jl < 3 * d; jl += L) __builtin_nontemporal_store(jl == diag ? dg[i] : 0.0, values + base + jl);
    }
    if (b && ln == 0)
      for (int i = 0; i < 3; ++i)
        if (am >> i & 1) {
          const int64_t r = node_dof<NODE_MAJOR>(nn, n, i);
          b[r] = dg[i] * g[r];
        }
  });
}

template <bool NODE_MAJOR>
__global__ __launch_bounds__(kLaThreads) void dirichlet_node_pull_kernel(
    int64_t nn, int64_t nnz_n, const int32_t* __restrict__ nrp, const int32_t* __restrict__ ncol,
    const uint8_t* __restrict__ bits, const uint8_t* __restrict__ touched, const double* __restrict__ g,
    double* __restrict__ values, double* __restrict__ b) {
  constexpr int L = 32;
  for_flagged_nodes<L>(nn, touched, [&](int64_t n, int ln, bool live) {
    const int free_rows = live ? ~bits[n] & 7 : 0;
    const int32_t n0 = live ? nrp[n] : 0, d = live ? nrp[n + 1] - n0 : 0;
    double bv[3] = {0.0, 0.0, 0.0};  // loaded before the entries, so that the update waits on nothing
    if (b && ln == 0)
      for (int i = 0; i < 3; ++i)
        if (free_rows >> i & 1) bv[i] = b[node_dof<NODE_MAJOR>(nn, n, i)];
    double s[3] = {0.0, 0.0, 0.0};
    for (int32_t q = ln; q < d; q += L) {
      const int32_t m = __builtin_nontemporal_load(ncol + n0 + q);
      const int bm = bits[m];
      if (!bm) continue;
      for (int j = 0; j < 3; ++j) {
        if (!(bm >> j & 1)) continue;
        const double gj = g[node_dof<NODE_MAJOR>(nn, m, j)];
        const bool lift = b && gj != 0.0;  // the value is read only when it lifts into b
        for (int i = 0; i < 3; ++i) {
          if (!(free_rows >> i & 1)) continue;
          double* v = values + node_slot<NODE_MAJOR>(nnz_n, n0, d, i, j, q);
          if (lift) s[i] = fma(__builtin_nontemporal_load(v), gj, s[i]);
          __builtin_nontemporal_store(0.0, v);
        }
      }
    }
    if (!b) return;
    for (int i = 0; i < 3; ++i)
      for (int o = L >> 1; o > 0; o >>= 1) s[i] += __shfl_xor(s[i], o);
    if (ln == 0)
      for (int i = 0; i < 3; ++i)
        if (free_rows >> i & 1) b[node_dof<NODE_MAJOR>(nn, n, i)] = bv[i] - s[i];
  });
}

// tmap[k] = slot of A_ji for the entry k = A_ij (-1 when the pattern has no A_ji): 16 lanes
// per row, a binary search in row j per entry.
__global__ __launch_bounds__(kLaThreads) void transpose_map_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                                  const int32_t* __restrict__ col,
                                                                  int32_t* __restrict__ tmap) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (kLaThreads / 16) + (threadIdx.x >> 4);
  if (i >= n) return;
  const int ln = static_cast<int>(threadIdx.x & 15);
  for (int32_t k = rp[i] + ln; k < rp[i + 1]; k += 16) {
    const int32_t j = col[k];
    int32_t lo = rp[j], hi = rp[j + 1];
    const int32_t end = hi;
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (col[mid] < i) lo = mid + 1;
      else hi = mid;
    }
    tmap[k] = lo < end && col[lo] == static_cast<int32_t>(i) ? lo : -1;
  }
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

int cg_grid(int64_t n) {
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((n + 4 * kCgThreads - 1) / (4 * kCgThreads),
                                                                 kCgMaxGroups)));
}

hipError_t launch_cg_init(int64_t n, const double* b, const double* r, const double* dinv, double* p,
                          double* part, CgState* st, hipStream_t s) {
  const int g = cg_grid(n);
  if (dinv)
    hipLaunchKernelGGL(cg_init_kernel<true>, dim3(g), dim3(kCgThreads), 0, s, n, b, r, dinv, p, part);
  else
    hipLaunchKernelGGL(cg_init_kernel<false>, dim3(g), dim3(kCgThreads), 0, s, n, b, r, dinv, p, part);
  hipLaunchKernelGGL(cg_init_finish, dim3(1), dim3(kCgThreads), 0, s, part, g, st);
  return hipGetLastError();
}

hipError_t launch_cg_iteration(const CsrView& A, int64_t it, double* x, double* p, double* r, double* ap,
                               const double* dinv, double* part, CgState* st, hipStream_t s) {
  const int64_t n = A.nrows;
  hipError_t e = launch_spmv(A, 1.0, p, 0.0, ap, s);
  if (e != hipSuccess) return e;
  const int g = cg_grid(n);
  hipLaunchKernelGGL(cg_dot_kernel, dim3(g), dim3(kCgThreads), 0, s, n, p, ap, part, st);
  if (dinv)
    hipLaunchKernelGGL(cg_update_kernel<true>, dim3(g), dim3(kCgThreads), 0, s, n, it, x, p, r, ap, dinv, part, st);
  else
    hipLaunchKernelGGL(cg_update_kernel<false>, dim3(g), dim3(kCgThreads), 0, s, n, it, x, p, r, ap, dinv, part, st);
  if (dinv)
    hipLaunchKernelGGL(cg_dir_kernel<true>, dim3(g), dim3(kCgThreads), 0, s, n, it, r, dinv, p, part, st);
  else
    hipLaunchKernelGGL(cg_dir_kernel<false>, dim3(g), dim3(kCgThreads), 0, s, n, it, r, dinv, p, part, st);
  return hipGetLastError();
}

hipError_t launch_cg_jacobi(int64_t n, double* d, hipStream_t s) {
  hipLaunchKernelGGL(cg_jacobi_kernel, dim3(cg_grid(n)), dim3(kCgThreads), 0, s, n, d);
  return hipGetLastError();
}

hipError_t launch_dof_cell_count(int64_t nk, const int32_t* cell_dofs, int32_t* cnt, hipStream_t s) {
  if (nk <= 0) return hipSuccess;
  hipLaunchKernelGGL(dof_cell_count_kernel, dim3(static_cast<unsigned>((nk + kLaThreads - 1) / kLaThreads)),
                     dim3(kLaThreads), 0, s, nk, cell_dofs, cnt);
  return hipGetLastError();
}

hipError_t launch_dirichlet_rows(const CsrView& A, const int32_t* tmap, const uint8_t* mask, const double* g,
                                 const int32_t* cnt, double* values, double* b, hipStream_t s) {
  if (A.nrows == 0) return hipSuccess;
  const bool long_rows = A.nnz > 32 * A.nrows;  // more than 32 entries per row on average
  const int L = long_rows ? 64 : 16;
  // about one wave per SIMD slot of the chip (1024 workgroups of 8 waves), each group looping
  // over its strided rows: 2048 and 4096 workgroups (rows in fewer rounds, but waves in
  // several batches) and 512 measured slower
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((A.nrows + L * (kLaThreads / L) - 1) /
                                                                  (L * (kLaThreads / L)), 1024));
  const dim3 gd(static_cast<unsigned>(grid)), bd(kLaThreads);
  if (long_rows) {
    if (tmap) hipLaunchKernelGGL((dirichlet_rows_kernel<64, true>), gd, bd, 0, s, A.nrows, A.row_ptr, A.col, tmap,
                                 mask, g, cnt, values, b);
    else hipLaunchKernelGGL((dirichlet_rows_kernel<64, false>), gd, bd, 0, s, A.nrows, A.row_ptr, A.col, tmap,
                            mask, g, cnt, values, b);
  } else {
    if (tmap) hipLaunchKernelGGL((dirichlet_rows_kernel<16, true>), gd, bd, 0, s, A.nrows, A.row_ptr, A.col, tmap,
                                 mask, g, cnt, values, b);
    else hipLaunchKernelGGL((dirichlet_rows_kernel<16, false>), gd, bd, 0, s, A.nrows, A.row_ptr, A.col, tmap,
                            mask, g, cnt, values, b);
  }
  return hipGetLastError();
}

hipError_t launch_dirichlet_nodes(const CsrView& A, const uint8_t* mask, const double* g, const int32_t* cnt,
                                  uint8_t* scratch, double* values, double* b, hipStream_t s) {
  if (A.nrows == 0) return hipSuccess;
  if (!A.node_layout || !A.node_rp || !A.node_col || !scratch || A.nrows % 3 || A.node_nnz * 9 != A.nnz)
    return hipErrorInvalidValue;
  const int64_t nn = A.nrows / 3;
  uint8_t* bits = scratch;
  uint8_t* touched = scratch + nn;
  const int64_t g1 = (nn + kLaThreads - 1) / kLaThreads;
  // about one wave per SIMD slot of the chip: each wave tests a strided sample of the nodes
  const int64_t g2 = std::max<int64_t>(1, std::min<int64_t>(((nn + 63) / 64 + kLaThreads / 64 - 1) / (kLaThreads / 64), 1024));
  if (g1 > INT32_MAX) return hipErrorInvalidValue;
  const dim3 bd(kLaThreads), d1(static_cast<unsigned>(g1)), d2(static_cast<unsigned>(g2));
  if (A.node_layout == 2) {
    hipLaunchKernelGGL(node_bc_bits_kernel<true>, d1, bd, 0, s, nn, mask, bits, touched);
    hipLaunchKernelGGL(dirichlet_node_rows_kernel<true>, d2, bd, 0, s, nn, A.node_nnz, A.node_rp, A.node_col, bits,
                       g, cnt, values, b, touched);
    hipLaunchKernelGGL(dirichlet_node_pull_kernel<true>, d2, bd, 0, s, nn, A.node_nnz, A.node_rp, A.node_col, bits,
                       touched, g, values, b);
  } else {
    hipLaunchKernelGGL(node_bc_bits_kernel<false>, d1, bd, 0, s, nn, mask, bits, touched);
    hipLaunchKernelGGL(dirichlet_node_rows_kernel<false>, d2, bd, 0, s, nn, A.node_nnz, A.node_rp, A.node_col, bits,
                       g, cnt, values, b, touched);
    hipLaunchKernelGGL(dirichlet_node_pull_kernel<false>, d2, bd, 0, s, nn, A.node_nnz, A.node_rp, A.node_col, bits,
                       touched, g, values, b);
  }
  return hipGetLastError();
}

hipError_t launch_transpose_map(const CsrView& A, int32_t* tmap, hipStream_t s) {
  if (A.nrows == 0) return hipSuccess;
  const int64_t grid = (A.nrows + kLaThreads / 16 - 1) / (kLaThreads / 16);
  if (grid > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(transpose_map_kernel, dim3(static_cast<unsigned>(grid)), dim3(kLaThreads), 0, s, A.nrows,
                     A.row_ptr, A.col, tmap);
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
  if (A.node_layout) {
    if (!A.node_rp || !A.node_col || !A.node_chunks || A.node_nchunks <= 0 || A.node_nchunks > INT32_MAX ||
        A.nrows % 3 || A.node_nnz * 9 != A.nnz)
      return hipErrorInvalidValue;
    const int nch = static_cast<int>(A.node_nchunks);
    if (A.node_layout == 2)
      hipLaunchKernelGGL(spmv_node_kernel<true>, dim3(static_cast<unsigned>(nch)), dim3(kNodeThreads), 0, s, nch,
                         A.node_chunks, A.nrows / 3, A.node_nnz, A.node_rp, A.node_col, A.values, alpha, x, beta,
                         y);
    else
      hipLaunchKernelGGL(spmv_node_kernel<false>, dim3(static_cast<unsigned>(nch)), dim3(kNodeThreads), 0, s, nch,
                         A.node_chunks, A.nrows / 3, A.node_nnz, A.node_rp, A.node_col, A.values, alpha, x, beta,
                         y);
    return hipGetLastError();
  }
  if (!A.blocks || A.nblocks <= 0 || A.nblocks > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(spmv_stream_kernel, dim3(static_cast<unsigned>(A.nblocks)), dim3(kLaThreads), 0,
                     s, static_cast<int>(A.nblocks), A.blocks, A.row_ptr, A.col,
                     A.col16 ? A.blk_col0 : nullptr, A.col16, A.values, alpha, x, beta, y);
  return hipGetLastError();
}

hipError_t launch_narrow_columns(const CsrView& A, int32_t* blk_col0, uint16_t* col16, hipStream_t s) {
  if (A.nblocks <= 0) return hipSuccess;
  if (!A.blocks || !A.col || A.nblocks > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(narrow_columns_kernel, dim3(static_cast<unsigned>(A.nblocks)), dim3(kLaThreads), 0, s,
                     A.blocks, A.col, blk_col0, col16);
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
