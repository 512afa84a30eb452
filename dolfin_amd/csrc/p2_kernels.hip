// Vector P2 elasticity on tetrahedra for gfx950.
//
// Form (demo/undocumented/elasticity/cpp/Elasticity.ufl with P1 mu, lmbda):
//   a(u, v) = inner(2 mu eps(u) + lmbda tr(eps(u)) I, eps(v))*dx,   L(v) = inner(f, v)*dx
// Local dof order is component-major, (i, a) -> i*10 + a, nodes a = 4 vertices then the 6
// UFC edges (2,3),(1,3),(1,2),(0,3),(0,2),(0,1) (python/test/unit/fem/test_dofmap.py:112-120).
//
// Element matrix K[(i,a),(j,b)] = delta_ij tr S^mu_ab + S^mu_ab[j][i] + S^lmbda_ab[i][j] with
// S^c_ab[p][q] = int c d_p phi_a d_q phi_b. The P2 basis gradients are linear in the
// barycentric coordinates and factor into at most two outer products,
//   grad phi_vertex_k = (4 lambda_k - 1) grad lambda_k,
//   grad phi_edge(k,l) = 4 lambda_l grad lambda_k + 4 lambda_k grad lambda_l,
// so with W^c_mr = int c lambda_m lambda_r (exact: c is P1, int lambda^alpha closed form)
//   S^c_ab = sum_{t,s} (u_t^T W^c u'_s) w_t (x) w'_s,
// at most 4 scalar bilinear forms and 4 outer products of 3-vectors per node pair. This is
// ~20 kFLOP per cell, so the kernel is bound by the 900-entry scatter, not by arithmetic
// (a 4 x 900 MFMA contraction would not shorten it).
//
// Scatter: per cell the rank of node b among the column nodes of node a is precomputed
// (100 bytes, built with the CSR); slot = row_ptr[dof(i,a)] + rank*s1 + j*s2 with
// (s1, s2) = (1, deg_a) for UFC component-major numbering or (3, 1) for node-major.
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "common.h"
#include "kernels.h"
#include "p2_device.h"

namespace dfa {

namespace {

using namespace p2dev;

template <bool WANT_A, bool WANT_B, bool BC>
__global__ __launch_bounds__(kThreads) void vp2_atomic_kernel(VP2Topo t, VP2Pass p) {
  __shared__ CellData sc[kCells];
  const int tid = threadIdx.x;
  const int64_t nchunks = (t.num_cells + kCells - 1) / kCells;
  for (int64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const int64_t cbase = chunk * kCells;
    const int ncell = static_cast<int>(min<int64_t>(kCells, t.num_cells - cbase));
    // phase 1: per-cell tables (16 threads per cell)
    {
      const int lc = tid >> 4, ln = tid & 15;
      if (lc < ncell) {
        CellData& c = sc[lc];
        const int64_t cell = cbase + lc;
        for (int k = ln; k < 30; k += 16) {
          const int32_t d = t.cell_dofs[cell * 30 + k];
          c.dof[k] = d;
          c.rbase[k] = t.row_ptr[d];
          if (k < 10) c.deg[k] = (t.row_ptr[d + 1] - t.row_ptr[d]) / 3;
          if constexpr (BC) {
            const uint8_t m = p.bc_mask[d];
            c.bcf[k] = m;
            c.bcv[k] = m ? p.bc_value[d] : 0.0;
          }
        }
        if constexpr (WANT_A)
          for (int k = ln; k < 100; k += 16) c.rank[k] = t.rank[cell * 100 + k];
        if (ln == 0) {
          double pc[4][3], mu[4], la[4];
          for (int v = 0; v < 4; ++v) {
            const int64_t vid = t.cells[cell * 4 + v];
            for (int i = 0; i < 3; ++i) pc[v][i] = p.x[vid * 3 + i];
            const int32_t kd = t.coeff_dofs[cell * 4 + v];
            mu[v] = (WANT_A || BC) ? p.mu[kd] : 0.0;
            la[v] = (WANT_A || BC) ? p.lmbda[kd] : 0.0;
          }
          vp2_cell_setup(pc, mu, la, c);
        }
      }
    }
    __syncthreads();
    if constexpr (WANT_B) {
      for (int k = tid; k < ncell * 30; k += kThreads) {
        CellData& c = sc[k / 30];
        const int e = k % 30, i = e / 10, a = e % 10;
        c.be[e] = p.f3[i] * (a < 4 ? -c.vol / 20.0 : c.vol / 5.0);
      }
      __syncthreads();
    }
    // phase 2: node pairs
    if constexpr (WANT_A || BC) for (int task = tid; task < ncell * 100; task += kThreads) {
      const int lc = task / 100, a = (task % 100) / 10, b = task % 10;
      const CellData& c = sc[lc];
      double K[3][3];
      vp2_block(c, a, b, K);
      if constexpr (BC) {
        // lifting of the original entries, then elimination (SystemAssembler::apply_bc)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const int r = i * 10 + a, col = j * 10 + b;
            if (WANT_B && c.bcf[col] && !c.bcf[r])
              unsafeAtomicAdd(const_cast<double*>(&c.be[r]), -K[i][j] * c.bcv[col]);
            if (c.bcf[r] || c.bcf[col]) K[i][j] = (r == col) ? 1.0 : 0.0;
          }
      }
      if constexpr (WANT_A) {
        const int rk = c.rank[a * 10 + b];
        const int s1 = t.component_major ? 1 : 3;
        const int s2 = t.component_major ? c.deg[a] : 1;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double* row = p.values + c.rbase[i * 10 + a] + rk * s1;
#pragma unroll
          for (int j = 0; j < 3; ++j) unsafeAtomicAdd(row + j * s2, K[i][j]);
        }
      }
    }
    __syncthreads();
    if constexpr (WANT_B) {
      for (int k = tid; k < ncell * 30; k += kThreads) {
        const CellData& c = sc[k / 30];
        const int e = k % 30;
        const double v = (BC && c.bcf[e]) ? c.bcv[e] : c.be[e];
        unsafeAtomicAdd(p.b + c.dof[e], v);
      }
    }
    __syncthreads();
  }
}

template <bool A, bool B, bool BC>
hipError_t launch_t(const VP2Topo& t, const VP2Pass& p, hipStream_t s) {
  const int64_t nchunks = (t.num_cells + kCells - 1) / kCells;
  const int64_t grid = std::min<int64_t>(nchunks, 256 * 16);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((vp2_atomic_kernel<A, B, BC>), dim3(static_cast<unsigned>(grid)), dim3(kThreads),
                     0, s, t, p);
  return hipGetLastError();
}


// without Dirichlet rows: four waves per SIMD (at most 128 VGPRs)
template <bool WANT_A, bool WANT_B, bool BC>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(BC ? 1 : 4)))
void vp2_blocked_kernel(VP2BlockTopo t, VP2Pass p, VP2Lds L) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int nb = t.num_blocks;
  const int q = nb >> 3, r = nb & 7;
  const int xcd = blockIdx.x & 7, jj = blockIdx.x >> 3;
  vp2_block_body<WANT_A, WANT_B, BC>(t, p, L, lds, xcd * q + min(xcd, r) + jj);
}

template <bool A, bool B, bool BC>
hipError_t launch_blocked_t(const VP2BlockTopo& t, const VP2Pass& p, hipStream_t s) {
  const VP2Lds L = vp2_lds(t.max_verts, t.max_owned, t.max_vals, B);
  auto kern = vp2_blocked_kernel<A, B, BC>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, L.total);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(t.num_blocks), dim3(kBT), L.total, s, t, p, L);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_vp2_atomic(const VP2Topo& t, const VP2Pass& p, hipStream_t s) {
  if (p.bc) {
    if (p.want_A && p.want_b) return launch_t<true, true, true>(t, p, s);
    if (p.want_A) return launch_t<true, false, true>(t, p, s);
    return launch_t<false, true, true>(t, p, s);
  }
  if (p.want_A && p.want_b) return launch_t<true, true, false>(t, p, s);
  if (p.want_A) return launch_t<true, false, false>(t, p, s);
  if (p.want_b) return launch_t<false, true, false>(t, p, s);
  return hipErrorInvalidValue;
}



std::string vp2_kernel_name(bool blocked, const VP2Pass& p) {
  return std::string(blocked ? "vp2_blocked_kernel" : "vp2_atomic_kernel") + "<WANT_A=" +
         (p.want_A ? "1" : "0") + ",WANT_B=" + (p.want_b ? "1" : "0") + ",BC=" + (p.bc ? "1" : "0") + ">";
}

size_t vp2_blocked_lds_bytes(const VP2BlockTopo& t, bool want_b) {
  return static_cast<size_t>(vp2_lds(t.max_verts, t.max_owned, t.max_vals, want_b).total);
}

hipError_t launch_vp2_blocked(const VP2BlockTopo& t, const VP2Pass& p, hipStream_t s) {
  if (t.num_blocks == 0) return hipSuccess;
  if (p.bc) {
    if (p.want_A && p.want_b) return launch_blocked_t<true, true, true>(t, p, s);
    if (p.want_A) return launch_blocked_t<true, false, true>(t, p, s);
    return launch_blocked_t<false, true, true>(t, p, s);
  }
  if (p.want_A && p.want_b) return launch_blocked_t<true, true, false>(t, p, s);
  if (p.want_A) return launch_blocked_t<true, false, false>(t, p, s);
  if (p.want_b) return launch_blocked_t<false, true, false>(t, p, s);
  return hipErrorInvalidValue;
}

}  // namespace dfa
