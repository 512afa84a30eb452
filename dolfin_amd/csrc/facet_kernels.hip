// Exterior-facet integrals of the P1 forms alpha*u*v*ds and g*v*ds, and the vector-P2
// traction inner(t, v)*ds.
//
// Assembler::assemble_exterior_facets (dolfin/fem/Assembler.cpp:196-298) tabulates, for every
// facet with one incident cell, a cell-sized tensor that is nonzero on the facet's vertices and
// adds it with the cell's dofs; SystemAssembler adds the same tensors to the cell tensor before
// apply_bc (dolfin/fem/SystemAssembler.cpp:421-481). The facet terms here are exact P1 facet
// integrals: on a facet F with d = gdim vertices, int_F lambda_a lambda_b = |F| (1 + delta_ab)
// / (d (d + 1)). One thread per facet with global FP64 atomics: boundary facets are few
// (196,608 on a 128^3 cube against 12.6M cells) and touch distinct rows mostly.
//
// Dirichlet rows and columns: elimination is linear in the cell tensor apart from the unit
// diagonal and b_i = g that apply_bc writes once per cell, and the cell pass of
// dfa_assemble_system writes those. The facet pass therefore drops Dirichlet rows and lifts
// Dirichlet columns into b, which together equals eliminating the merged cell tensor.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"
#include "p1_device.h"

namespace dfa {
namespace {

constexpr int kFacetThreads = 256;

template <int D, bool BC>
__global__ __launch_bounds__(kFacetThreads) void facet_kernel(FacetTopo t, FacetPass p) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kFacetThreads + threadIdx.x;
  if (k >= t.num_facets) return;
  if (p.mask && !p.mask[k]) return;
  constexpr int N = D + 1;
  const int64_t c = t.facet_cell[k];
  const int lf = t.facet_local[k];
  // facet vertex j is cell vertex j + (j >= lf)
  int32_t d[D], kd[D];
  double xv[D][D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const int a = j + (j >= lf ? 1 : 0);
    const int64_t v = t.cells[c * N + a];
    d[j] = t.cell_dofs[c * N + a];
    kd[j] = t.coeff_dofs[c * N + a];
#pragma unroll
    for (int q = 0; q < D; ++q) xv[j][q] = p.x[v * D + q];
  }
  double meas = 1.0;
  if constexpr (D == 2) {
    const double dx = xv[1][0] - xv[0][0], dy = xv[1][1] - xv[0][1];
    meas = sqrt(dx * dx + dy * dy);
  } else if constexpr (D == 3) {
    double e1[3], e2[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      e1[q] = xv[1][q] - xv[0][q];
      e2[q] = xv[2][q] - xv[0][q];
    }
    const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                 cz = e1[0] * e2[1] - e1[1] * e2[0];
    meas = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
  }
  const double w = meas / (D * (D + 1.0));
  double be[D];
  double gs = 0.0, gl[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    gl[j] = p.has_b ? p.g[kd[j]] : 0.0;
    gs += gl[j];
  }
#pragma unroll
  for (int i = 0; i < D; ++i) be[i] = p.has_b ? w * (gl[i] + gs) : 0.0;
  bool isbc[D];
  double gv[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    isbc[j] = BC && p.bc_mask[d[j]] != 0;
    gv[j] = isbc[j] ? p.bc_value[d[j]] : 0.0;
  }
  const double wa = p.alpha * w;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    if (isbc[i]) continue;  // a Dirichlet row keeps what the cell pass wrote
#pragma unroll
    for (int j = 0; j < D; ++j) {
      if (!p.has_a) break;
      const double m = (i == j) ? 2.0 * wa : wa;
      if (isbc[j]) {
        be[i] -= m * gv[j];  // lifting of a Dirichlet column
      } else if (p.values) {
        p1dev::atomic_add_global(p.values + p1dev::find_slot(t.row_ptr, t.col, d[i], d[j]), m);
      }
    }
    if (p.b && (p.has_b || BC)) p1dev::atomic_add_global(p.b + d[i], be[i]);
  }
}

// Vector-P2 traction inner(t, v)*ds, t constant: on a facet F the P2 basis integrates to 0 at
// the vertex nodes and to |F| / 3 at the three edge nodes of F (the edges of the cell that do
// not touch the opposite vertex). Only the load vector has terms, so a Dirichlet row is left
// as the cell pass wrote it and nothing is lifted.
template <bool BC>
__global__ __launch_bounds__(kFacetThreads) void vp2_traction_kernel(FacetTopo t, FacetPass p) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kFacetThreads + threadIdx.x;
  if (k >= t.num_facets) return;
  if (p.mask && !p.mask[k]) return;
  const int64_t c = t.facet_cell[k];
  const int lf = t.facet_local[k];
  double xv[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int64_t v = t.cells[c * 4 + j + (j >= lf ? 1 : 0)];
#pragma unroll
    for (int q = 0; q < 3; ++q) xv[j][q] = p.x[v * 3 + q];
  }
  double e1[3], e2[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    e1[q] = xv[1][q] - xv[0][q];
    e2[q] = xv[2][q] - xv[0][q];
  }
  const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
               cz = e1[0] * e2[1] - e1[1] * e2[0];
  const double w = 0.5 * sqrt(cx * cx + cy * cy + cz * cz) / 3.0;
  const double tr[3] = {p.g[0], p.g[1], p.g[2]};
  // local edge e joins vertices (ea[e], eb[e]): nodes 4..9 of the cell
  constexpr int ea[6] = {2, 1, 1, 0, 0, 0}, eb[6] = {3, 3, 2, 3, 2, 1};
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    if (ea[e] == lf || eb[e] == lf) continue;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int32_t d = t.cell_dofs[c * 30 + i * 10 + 4 + e];
      if (BC && p.bc_mask[d]) continue;
      p1dev::atomic_add_global(p.b + d, tr[i] * w);
    }
  }
}

template <int D>
hipError_t launch_facets_d(const FacetTopo& t, const FacetPass& p, hipStream_t s) {
  const int64_t grid = (t.num_facets + kFacetThreads - 1) / kFacetThreads;
  if (p.bc_mask)
    hipLaunchKernelGGL((facet_kernel<D, true>), dim3(static_cast<unsigned>(grid)), dim3(kFacetThreads), 0, s, t, p);
  else
    hipLaunchKernelGGL((facet_kernel<D, false>), dim3(static_cast<unsigned>(grid)), dim3(kFacetThreads), 0, s, t, p);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_facets(const FacetTopo& t, const FacetPass& p, hipStream_t s) {
  if (t.num_facets == 0) return hipSuccess;
  if (p.vp2_traction) {
    if (p.gdim != 3 || !p.g) return hipErrorInvalidValue;
    if (!p.b) return hipSuccess;
    const int64_t grid = (t.num_facets + kFacetThreads - 1) / kFacetThreads;
    if (p.bc_mask)
      hipLaunchKernelGGL((vp2_traction_kernel<true>), dim3(static_cast<unsigned>(grid)), dim3(kFacetThreads), 0, s, t, p);
    else
      hipLaunchKernelGGL((vp2_traction_kernel<false>), dim3(static_cast<unsigned>(grid)), dim3(kFacetThreads), 0, s, t, p);
    return hipGetLastError();
  }
  switch (p.gdim) {
    case 1: return launch_facets_d<1>(t, p, s);
    case 2: return launch_facets_d<2>(t, p, s);
    case 3: return launch_facets_d<3>(t, p, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace dfa
