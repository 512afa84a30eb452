// Device-side building blocks of the P1 kernels (geometry, element tensors, Dirichlet
// elimination, the blocked-scatter block body), shared by the standalone kernels and the
// batched persistent kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace dfa {
namespace p1dev {

constexpr int kThreads = 256;

__device__ __forceinline__ void atomic_add_global(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add_lds(double* p, double v) { unsafeAtomicAdd(p, v); }

// Cofactor gradients g[a] = det(J) * grad(lambda_a) and |det J| of an affine simplex.
template <int D>
__device__ __forceinline__ void p1_geometry(const double (&p)[D + 1][D], double (&g)[D + 1][D],
                                            double& absdet) {
  if constexpr (D == 3) {
    double e1[3], e2[3], e3[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      e1[i] = p[1][i] - p[0][i];
      e2[i] = p[2][i] - p[0][i];
      e3[i] = p[3][i] - p[0][i];
    }
    // rows of adj(J) for J = [e1 e2 e3]: e2 x e3, e3 x e1, e1 x e2
    g[1][0] = e2[1] * e3[2] - e2[2] * e3[1];
    g[1][1] = e2[2] * e3[0] - e2[0] * e3[2];
    g[1][2] = e2[0] * e3[1] - e2[1] * e3[0];
    g[2][0] = e3[1] * e1[2] - e3[2] * e1[1];
    g[2][1] = e3[2] * e1[0] - e3[0] * e1[2];
    g[2][2] = e3[0] * e1[1] - e3[1] * e1[0];
    g[3][0] = e1[1] * e2[2] - e1[2] * e2[1];
    g[3][1] = e1[2] * e2[0] - e1[0] * e2[2];
    g[3][2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double det = e1[0] * g[1][0] + e1[1] * g[1][1] + e1[2] * g[1][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) g[0][i] = -(g[1][i] + g[2][i] + g[3][i]);
    absdet = fabs(det);
  } else if constexpr (D == 2) {
    const double e1x = p[1][0] - p[0][0], e1y = p[1][1] - p[0][1];
    const double e2x = p[2][0] - p[0][0], e2y = p[2][1] - p[0][1];
    g[1][0] = e2y;
    g[1][1] = -e2x;
    g[2][0] = -e1y;
    g[2][1] = e1x;
    g[0][0] = -(g[1][0] + g[2][0]);
    g[0][1] = -(g[1][1] + g[2][1]);
    absdet = fabs(e1x * e2y - e1y * e2x);
  } else {
    const double e = p[1][0] - p[0][0];
    g[1][0] = 1.0;
    g[0][0] = -1.0;
    absdet = fabs(e);
  }
}

template <int D>
struct Fact {
  static constexpr double value = D == 1 ? 1.0 : (D == 2 ? 2.0 : 6.0);
};

// 1/d to full FP64 accuracy without the IEEE division sequence: hardware reciprocal
// estimate refined by two Newton steps (d is a nonzero, finite |det J|).
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// Element matrix (row-major, N x N) and vector of one P1 cell.
template <int D, int AFORM, bool WANT_B>
__device__ __forceinline__ void p1_element(const double (&p)[D + 1][D], const double (&kap)[D + 1],
                                           const double (&fv)[D + 1], double (&A)[D + 1][D + 1],
                                           double (&be)[D + 1]) {
  constexpr int N = D + 1;
  double g[N][D];
  double absdet;
  p1_geometry<D>(p, g, absdet);
  if constexpr (AFORM == DFA_FORM_P1_POISSON) {
    double ks = kap[0];
#pragma unroll
    for (int a = 1; a < N; ++a) ks += kap[a];
    constexpr double c = 1.0 / (N * Fact<D>::value);
    const double w = ks * c * recip(absdet);
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = a; b < N; ++b) {
        double s = g[a][0] * g[b][0];
#pragma unroll
        for (int i = 1; i < D; ++i) s = fma(g[a][i], g[b][i], s);
        A[a][b] = w * s;
        A[b][a] = w * s;
      }
  } else if constexpr (AFORM == DFA_FORM_P1_MASS) {
    constexpr double c = 1.0 / (Fact<D>::value * (D + 1.0) * (D + 2.0));
    const double w = absdet * c;
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = 0; b < N; ++b) A[a][b] = (a == b) ? 2.0 * w : w;
  }
  if constexpr (WANT_B) {
    double fs = fv[0];
#pragma unroll
    for (int a = 1; a < N; ++a) fs += fv[a];
    constexpr double c = 1.0 / (Fact<D>::value * (D + 1.0) * (D + 2.0));
    const double w = absdet * c;
#pragma unroll
    for (int a = 0; a < N; ++a) be[a] = w * (fv[a] + fs);
  }
}

// SystemAssembler::apply_bc for a square form (dolfin/fem/SystemAssembler.cpp:1051-1075),
// applied to the element tensors in local column order.
template <int N>
__device__ __forceinline__ void apply_bc(double (&A)[N][N], double (&be)[N], const bool (&isbc)[N],
                                         const double (&g)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (!isbc[i]) continue;
#pragma unroll
    for (int j = 0; j < N; ++j) A[i][j] = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) be[r] -= A[r][i] * g[i];
#pragma unroll
    for (int r = 0; r < N; ++r) A[r][i] = 0.0;
    be[i] = g[i];
    A[i][i] = 1.0;
  }
}

__device__ __forceinline__ int find_slot(const int32_t* __restrict__ row_ptr,
                                         const int32_t* __restrict__ col, int32_t r, int32_t c) {
  int lo = row_ptr[r], hi = row_ptr[r + 1];
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}


// ---------------------------------------------------------------------------------------
// Blocked (owner-computes) scatter: workgroup per row block.
// ---------------------------------------------------------------------------------------
struct LdsLayout {
  int node, f, bcv, bcf, own, vals, bv, rdof, rbase, diag, total;
};

__host__ __device__ inline int align16(int n) { return (n + 15) & ~15; }

// Node record in LDS, read with 128-bit loads. Small: coordinates and kappa (32 bytes for
// triangles/tets; f and the Dirichlet value in separate arrays). Wide: coordinates, kappa, f,
// Dirichlet value (48 bytes): a 12-dword stride starts consecutive records on 8 different
// 4-bank groups, while a 32-byte stride reuses only 4 of them.
__host__ __device__ constexpr int node_stride(int D, bool wide = false) {
  return wide ? (D == 1 ? 4 : 6) : (D == 1 ? 2 : 4);
}

// LDS layout of one block. Strided: owned row i occupies values [i*kRowStride, +16), ranks
// index it directly. Compact: rows packed back to back (any row length), located through a
// per-node table.
__host__ __device__ inline LdsLayout lds_layout(int D, int maxN, int maxR, int maxNnz, bool f,
                                                bool bc, bool strided, bool wide) {
  LdsLayout L;
  int o = 0;
  L.node = o; o += align16(maxN * node_stride(D, wide) * 8);
  L.f = o;    o += (f && !wide) ? align16(maxN * 8) : 0;
  L.bcv = o;  o += (bc && !wide) ? align16(maxN * 8) : 0;
  L.bcf = o;  o += bc ? align16(maxN) : 0;
  L.own = o;  o += strided ? 0 : align16(maxN * 4);
  L.vals = o; o += align16((strided ? maxR * kRowStride : maxNnz) * 8);
  L.bv = o;   o += f ? align16(maxR * 8) : 0;
  L.rdof = o; o += align16(maxR * 4);
  L.rbase = o; o += strided ? align16(2 * maxR * 4) : 0;
  L.diag = o;  o += strided ? align16(maxR) : 0;
  L.total = o;
  return L;
}

template <int D>
__device__ __forceinline__ void load_node(const double* __restrict__ rec, double (&pc)[D], double& kap) {
  if constexpr (D == 1) {
    const double2 r = *reinterpret_cast<const double2*>(rec);
    pc[0] = r.x;
    kap = r.y;
  } else {
    const double2 r0 = *reinterpret_cast<const double2*>(rec);
    const double2 r1 = *reinterpret_cast<const double2*>(rec + 2);
    pc[0] = r0.x;
    pc[1] = r0.y;
    if constexpr (D == 3) {
      pc[2] = r1.x;
      kap = r1.y;
    } else {
      kap = r1.x;
    }
  }
}

// XCD-aware block order: workgroup i runs on XCD i % 8; each XCD gets a contiguous range of
// (spatially adjacent) blocks so that shared ghost nodes hit the same L2.
__device__ __forceinline__ int xcd_block(int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
  return xcd * q + min(xcd, r) + j;
}

// Prologue shared by the assembly and adjoint kernels: gather the block's nodes into LDS.
// ZF (wide tetrahedral records without Dirichlet data): x, y, z, kappa, z, f, so that z and
// f come with one 128-bit read.
template <int D, int T, bool NEED_K, bool NEED_F, bool BC, bool OWN, bool WIDE, bool ZF = false>
__device__ __forceinline__ void gather_nodes(const P1BlockTopo& t, const double* __restrict__ x,
                                             const double* __restrict__ kappa,
                                             const double* __restrict__ f,
                                             const uint8_t* __restrict__ bcm,
                                             const double* __restrict__ bcv, int blk, double* snode,
                                             double* sf, double* sbcv, uint8_t* sbcf, int32_t* sown) {
  constexpr int NS = node_stride(D, WIDE);
  const int n0 = t.node_ptr[blk], nn = t.node_ptr[blk + 1] - n0;
  for (int i = threadIdx.x; i < nn; i += T) {
    const int32_t v = t.node_vertex[n0 + i];
    double rec[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) rec[k] = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) rec[k] = x[static_cast<int64_t>(v) * D + k];
    if constexpr (NEED_K || NEED_F) {
      const int32_t kd = t.kdof_is_vertex ? v : t.node_kdof[n0 + i];
      if constexpr (NEED_K) rec[D] = kappa[kd];
      if constexpr (NEED_F) {
        if constexpr (ZF) rec[D + 2] = f[kd];
        else if constexpr (WIDE) rec[D + 1] = f[kd];
        else sf[i] = f[kd];
      }
    }
    if constexpr (BC) {
      const int32_t dd = t.node_dof[n0 + i];
      const uint8_t m = bcm[dd];
      sbcf[i] = m;
      const double g = m ? bcv[dd] : 0.0;
      if constexpr (WIDE) rec[D + 2] = g;
      else sbcv[i] = g;
    }
    if constexpr (ZF) rec[D + 1] = rec[D - 1];
#pragma unroll
    for (int k = 0; k < NS; k += 2)
      *reinterpret_cast<double2*>(snode + i * NS + k) = make_double2(rec[k], rec[k + 1]);
    if constexpr (OWN) sown[i] = -1;
  }
}

template <int NS>
__device__ __forceinline__ void load_rec(const double* __restrict__ p, double (&rec)[NS]) {
#pragma unroll
  for (int k = 0; k < NS; k += 2) {
    const double2 r = *reinterpret_cast<const double2*>(p + k);
    rec[k] = r.x;
    rec[k + 1] = r.y;
  }
}

// Sum over each 16-lane row of the wavefront, result in every lane, with DPP moves only
// (VALU; no LDS-pipe permute): xor 1 and xor 2 within quads, then rotate the row by 4 and 8.
__device__ __forceinline__ double dpp_move(double v, int ctrl) {
  const int2 a = *reinterpret_cast<int2*>(&v);
  int2 r;
  switch (ctrl) {
    case 0xB1:
      r.x = __builtin_amdgcn_update_dpp(0, a.x, 0xB1, 0xF, 0xF, false);
      r.y = __builtin_amdgcn_update_dpp(0, a.y, 0xB1, 0xF, 0xF, false);
      break;
    case 0x4E:
      r.x = __builtin_amdgcn_update_dpp(0, a.x, 0x4E, 0xF, 0xF, false);
      r.y = __builtin_amdgcn_update_dpp(0, a.y, 0x4E, 0xF, 0xF, false);
      break;
    case 0x124:
      r.x = __builtin_amdgcn_update_dpp(0, a.x, 0x124, 0xF, 0xF, false);
      r.y = __builtin_amdgcn_update_dpp(0, a.y, 0x124, 0xF, 0xF, false);
      break;
    default:
      r.x = __builtin_amdgcn_update_dpp(0, a.x, 0x128, 0xF, 0xF, false);
      r.y = __builtin_amdgcn_update_dpp(0, a.y, 0x128, 0xF, 0xF, false);
      break;
  }
  return *reinterpret_cast<double*>(&r);
}

__device__ __forceinline__ double row16_sum(double v) {
  v += dpp_move(v, 0xB1);   // quad_perm [1,0,3,2]
  v += dpp_move(v, 0x4E);   // quad_perm [2,3,0,1]
  v += dpp_move(v, 0x124);  // row_ror 4
  v += dpp_move(v, 0x128);  // row_ror 8
  return v;
}

// One row block of the blocked scatter (the workgroup's whole job in the standalone kernel,
// one work unit of the batched persistent kernel).
template <int D, int AFORM, bool WANT_B, bool BC, bool STRIDED, int T, bool WIDE, int RR = 0>
__device__ __forceinline__ void p1_block_body(const P1BlockTopo& t, const P1Pass& p,
                                              const LdsLayout& L, unsigned char* lds, int blk) {
  constexpr int N = D + 1;
  constexpr int NS = node_stride(D, WIDE);
  constexpr bool WANT_A = AFORM >= 0;
  constexpr bool NEED_K = AFORM == DFA_FORM_P1_POISSON;
  double* snode = reinterpret_cast<double*>(lds + L.node);
  double* sf = reinterpret_cast<double*>(lds + L.f);
  double* sbcv = reinterpret_cast<double*>(lds + L.bcv);
  uint8_t* sbcf = lds + L.bcf;
  int32_t* sown = reinterpret_cast<int32_t*>(lds + L.own);
  double* svals = reinterpret_cast<double*>(lds + L.vals);
  double* sbv = reinterpret_cast<double*>(lds + L.bv);
  int32_t* srdof = reinterpret_cast<int32_t*>(lds + L.rdof);
  int32_t* srbase = reinterpret_cast<int32_t*>(lds + L.rbase);

  const int tid = threadIdx.x;
  const bool do_A = WANT_A && p.want_A;
  constexpr bool ROWSUM_DIAG = STRIDED && AFORM == DFA_FORM_P1_POISSON && !BC;
  int8_t* sdiag = reinterpret_cast<int8_t*>(lds + L.diag);
  // cell records: issued first so that their latency overlaps the node gather
  const int c0 = t.cell_ptr[blk], ncb = t.cell_ptr[blk + 1] - c0;
  const ulonglong2* __restrict__ crec = reinterpret_cast<const ulonglong2*>(t.cell_rec) + c0;
  // rows longer than 16 entries (the layout without strided rows): 8-bit ranks, two words
  static_assert(STRIDED || RR == 0, "the layout for long rows streams its cell records");
  const ulonglong2* __restrict__ wrec =
      STRIDED ? nullptr : reinterpret_cast<const ulonglong2*>(t.cell_rank_wide) + c0;
  ulonglong2 rreg[RR > 0 ? RR : 1];
  if constexpr (RR > 0) {
#pragma unroll
    for (int j = 0; j < RR; ++j)
      rreg[j] = (tid + j * T < ncb) ? crec[tid + j * T] : make_ulonglong2(0, 0);
  }

  gather_nodes<D, T, NEED_K, WANT_B, BC, !STRIDED, WIDE>(t, p.x, p.kappa, p.f, p.bc_mask,
                                                         p.bc_value, blk, snode, sf, sbcv, sbcf,
                                                         sown);
  const int r0 = t.row_ptr_b[blk], nr = t.row_ptr_b[blk + 1] - r0;
  if (do_A) {
    const int nz = STRIDED ? nr * kRowStride : t.blk_nnz[blk];
    for (int i = tid; i < (nz + 1) / 2; i += T)
      reinterpret_cast<double2*>(svals)[i] = make_double2(0.0, 0.0);
  }
  for (int i = tid; i < nr; i += T) {
    const int32_t dof = t.row_dof[r0 + i];
    srdof[i] = dof;
    if constexpr (STRIDED) {
      // CSR start of each owned row, fetched now so that the write-out issues no dependent
      // global loads (the row length follows from the next row's start or the dof after)
      if (do_A) {
        const int32_t b0 = t.csr_row_ptr[dof];
        srbase[i] = b0;
        srbase[nr + i] = t.csr_row_ptr[dof + 1] - b0;
      }
    }
    if constexpr (WANT_B) sbv[i] = 0.0;
    if constexpr (ROWSUM_DIAG) sdiag[i] = static_cast<int8_t>(t.row_diag[r0 + i]);
  }
  if constexpr (!STRIDED) {
    __syncthreads();
    for (int i = tid; i < nr; i += T) {
      const uint32_t info = t.row_info[r0 + i];
      sown[info & 0xffffu] = static_cast<int>(info >> 16) | (i << 16);
    }
  }
  __syncthreads();

  auto process = [&](uint64_t ids, uint64_t rk, ulonglong2 wr) {
    int l[N];
    double pc[N][D], kap[N], fv[N], gv[N];
#pragma unroll
    for (int a = 0; a < N; ++a) {
      l[a] = static_cast<int>((ids >> (16 * a)) & 0xffffu);
      if constexpr (WIDE && !BC && D == 3) {
        // x, y | z, kappa as two 128-bit reads, f alone as a 64-bit read
        const double* rp = snode + l[a] * NS;
        const double2 r0 = *reinterpret_cast<const double2*>(rp);
        const double2 r1 = *reinterpret_cast<const double2*>(rp + 2);
        pc[a][0] = r0.x;
        pc[a][1] = r0.y;
        pc[a][2] = r1.x;
        kap[a] = r1.y;
        fv[a] = WANT_B ? rp[4] : 0.0;
        gv[a] = 0.0;
      } else if constexpr (WIDE) {
        double rec[NS];
        load_rec<NS>(snode + l[a] * NS, rec);
#pragma unroll
        for (int q = 0; q < D; ++q) pc[a][q] = rec[q];
        kap[a] = rec[D];
        fv[a] = WANT_B ? rec[D + 1] : 0.0;
        gv[a] = BC ? rec[D + 2] : 0.0;
      } else {
        load_node<D>(snode + l[a] * NS, pc[a], kap[a]);
        fv[a] = WANT_B ? sf[l[a]] : 0.0;
        gv[a] = BC ? sbcv[l[a]] : 0.0;
      }
    }
    double A[N][N], be[N];
#pragma unroll
    for (int a = 0; a < N; ++a) be[a] = 0.0;
    p1_element<D, AFORM, WANT_B>(pc, kap, fv, A, be);
    if constexpr (BC) {
      bool isbc[N];
#pragma unroll
      for (int a = 0; a < N; ++a) isbc[a] = sbcf[l[a]] != 0;
      apply_bc<N>(A, be, isbc, gv);
    }
#pragma unroll
    for (int a = 0; a < N; ++a) {
      int off, row;
      if constexpr (STRIDED) {
        if (l[a] >= nr) continue;
        row = l[a];
        off = l[a] * kRowStride;
      } else {
        const int own = sown[l[a]];
        if (own < 0) continue;
        off = own & 0xffff;
        row = own >> 16;
      }
      if (do_A) {
        const uint32_t ra = static_cast<uint32_t>(rk >> (16 * a));
        const uint64_t wa = (a < 2) ? wr.x : wr.y;
#pragma unroll
        for (int b = 0; b < N; ++b) {
          const int slot = STRIDED ? static_cast<int>((ra >> (4 * b)) & 15u)
                                   : static_cast<int>((wa >> (8 * ((a & 1) * 4 + b))) & 0xffu);
          if constexpr (ROWSUM_DIAG) {
            if (b == a) continue;  // recovered from the row sum in the write-out
          }
          atomic_add_lds(svals + off + slot, A[a][b]);
        }
      }
      if constexpr (WANT_B) atomic_add_lds(sbv + row, be[a]);
    }
  };
  if constexpr (RR > 0) {
    // all of this thread's cell records were loaded into registers in the prologue
#pragma unroll
    for (int j = 0; j < RR; ++j)
      if (tid + j * T < ncb) process(rreg[j].x, rreg[j].y, make_ulonglong2(0, 0));
  } else {
    // streaming: records prefetched one iteration ahead
    ulonglong2 nxt = tid < ncb ? crec[tid] : make_ulonglong2(0, 0);
    ulonglong2 wnxt = (!STRIDED && tid < ncb) ? wrec[tid] : make_ulonglong2(0, 0);
    for (int k = tid; k < ncb; k += T) {
      const ulonglong2 cur = nxt, wcur = wnxt;
      if (k + T < ncb) {
        nxt = crec[k + T];
        if constexpr (!STRIDED) wnxt = wrec[k + T];
      }
      process(cur.x, cur.y, wcur);
    }
  }
  __syncthreads();

  if (do_A) {
    if constexpr (STRIDED) {
      // 16 lanes per row; consecutive rows are consecutive dofs, so a wave stores 4 adjacent
      // CSR rows.
      const int grp = tid >> 4, ln = tid & 15;
      for (int i = grp; i < nr; i += T / 16) {
        const int base = srbase[i];
        const int len = srbase[nr + i];
        if constexpr (ROWSUM_DIAG) {
          // kappa-Poisson element rows sum to zero: K_aa = -sum_{b != a} K_ab
          double v = ln < len ? svals[i * kRowStride + ln] : 0.0;
          const double rs = row16_sum(v);
          if (ln == sdiag[i]) v = -rs;
          if (ln < len) {
            if (p.add) v += p.values[base + ln];
            p.values[base + ln] = v;
          }
        } else if (ln < len) {
          double v = svals[i * kRowStride + ln];
          if (p.add) v += p.values[base + ln];
          p.values[base + ln] = v;
        }
      }
    } else {
      // runs of consecutive rows: each run is one contiguous CSR range
      const int q0 = t.run_ptr[blk], nq = t.run_ptr[blk + 1] - q0;
      const int wave = tid >> 6, lane = tid & 63;
      for (int q = wave; q < nq; q += T / 64) {
        const uint64_t run = t.runs[q0 + q];
        const int64_t gstart = static_cast<int64_t>(static_cast<uint32_t>(run));
        const int loff = static_cast<int>((run >> 32) & 0xffffu);
        const int len = static_cast<int>(run >> 48);
        for (int jj = lane; jj < len; jj += 64) {
          double v = svals[loff + jj];
          if (p.add) v += p.values[gstart + jj];
          p.values[gstart + jj] = v;
        }
      }
    }
  }
  if constexpr (WANT_B)
    for (int i = tid; i < nr; i += T) {
      const int32_t dof = srdof[i];
      p.b[dof] = p.add ? p.b[dof] + sbv[i] : sbv[i];
    }
}


// ---------------------------------------------------------------------------------------
// Edge-chain scatter for the tetrahedral kappa-Poisson operator (+ load vector) without
// Dirichlet rows, strided LDS rows, 48-byte node records. A lane walks one chain of cells
// (p, q, r_k, r_{k+1}), k < m, around the edge (p, q): it reads each of the m + 3 nodes once,
// keeps the sums for (p, q), (p, r_k), (q, r_k) and the load vector in registers until they
// are complete and adds every finished entry to each owned row at the recorded ranks. The
// diagonal is minus the row sum at write-out (the element rows sum to zero).
// ---------------------------------------------------------------------------------------

// 1/d (d > 0, finite) from the hardware estimate and one Newton step: the estimate's error
// e becomes e^2, below 1e-14 relative, far inside the 1e-12 parity bar.
__device__ __forceinline__ double recip1(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

// Byte address in LDS of rank `nib` of the owned row whose byte base is `row_base`.
__device__ __forceinline__ double* row_slot(double* svals, uint32_t row_base, uint32_t nib) {
  return reinterpret_cast<double*>(reinterpret_cast<char*>(svals) + row_base + (nib << 3));
}

// The walk of one chain; whether a node is an owned row is tested once per node (p and q once
// per chain, r_k once per cell).
template <bool WANT_B>
__device__ __forceinline__ void chain_walk(const uint32_t (&d)[kChainWords], const double* snode,
                                           double* svals, double* sbv, uint32_t nr) {
  constexpr uint32_t kRowBytes = kRowStride * 8;
  const uint32_t lp = d[0] & 0xffffu, lq = d[0] >> 16;
  const int m = static_cast<int>(d[4] >> 16);
  auto rid = [&](int k) { return (d[1 + k / 2] >> (16 * (k & 1))) & 0xffffu; };
  auto rank_byte = [&](int j) { return (d[5 + j / 4] >> (8 * (j % 4))) & 0xffu; };
  const bool op = lp < nr, oq = lq < nr;
  // the finished entry (u, v) = (v, u) with value w, into every owned row of the pair
  auto emit = [&](bool ou, uint32_t u, bool ov, uint32_t v, uint32_t byte, double w) {
    if (ou) atomic_add_lds(row_slot(svals, u * kRowBytes, byte & 15u), w);
    if (ov) atomic_add_lds(row_slot(svals, v * kRowBytes, byte >> 4), w);
  };
  auto emit_b = [&](bool ou, uint32_t u, double w) {
    if constexpr (WANT_B) {
      if (ou) atomic_add_lds(sbv + u, w);
    }
  };
  // node record: x, y | z, kappa | z, f
  auto xy = [&](uint32_t l) { return *reinterpret_cast<const double2*>(snode + l * 6); };
  auto zk = [&](uint32_t l) { return *reinterpret_cast<const double2*>(snode + l * 6 + 2); };
  auto fv = [&](uint32_t l) { return snode[l * 6 + 5]; };

  const double2 pxy = xy(lp), pzk = zk(lp), qxy = xy(lq), qzk = zk(lq);
  const double px = pxy.x, py = pxy.y, pz = pzk.x;
  const double e1x = qxy.x - px, e1y = qxy.y - py, e1z = qzk.x - pz;
  const double kpq = pzk.y + qzk.y;
  const double fpq = WANT_B ? fv(lp) + fv(lq) : 0.0;
  uint32_t cur = rid(0);
  bool oc = cur < nr;
  double e2x, e2y, e2z, kr, fr;
  {
    const double2 rxy = xy(cur), rzk = zk(cur);
    e2x = rxy.x - px;
    e2y = rxy.y - py;
    e2z = rzk.x - pz;
    kr = rzk.y;
    fr = WANT_B ? fv(cur) : 0.0;
  }
  double apq = 0.0, cp = 0.0, cq = 0.0, swf = 0.0, sw = 0.0, cb = 0.0;
#pragma unroll
  for (int k = 0; k < kChainLen; ++k) {
    if (k < m) {
      const uint32_t ln = rid(k + 1);
      const bool on = ln < nr;
      const double2 nxy = xy(ln), nzk = zk(ln);
      const double fn = WANT_B ? fv(ln) : 0.0;
      const double e3x = nxy.x - px, e3y = nxy.y - py, e3z = nzk.x - pz;
      // cofactor gradients of q, r_k, r_{k+1}: g1 = e2 x e3, g2 = e3 x e1, g3 = e1 x e2
      const double g1x = e2y * e3z - e2z * e3y, g1y = e2z * e3x - e2x * e3z, g1z = e2x * e3y - e2y * e3x;
      const double g2x = e3y * e1z - e3z * e1y, g2y = e3z * e1x - e3x * e1z, g2z = e3x * e1y - e3y * e1x;
      const double g3x = e1y * e2z - e1z * e2y, g3y = e1z * e2x - e1x * e2z, g3z = e1x * e2y - e1y * e2x;
      const double adet = fabs(e1x * g1x + e1y * g1y + e1z * g1z);
      // kappa-Poisson: K_ab = mean(kappa) / (6 |det|) g_a . g_b, g_p = -(g1 + g2 + g3)
      const double w = (kpq + kr + nzk.y) * (1.0 / 24.0) * recip1(adet);
      const double g11 = fma(g1x, g1x, fma(g1y, g1y, g1z * g1z));
      const double g22 = fma(g2x, g2x, fma(g2y, g2y, g2z * g2z));
      const double g33 = fma(g3x, g3x, fma(g3y, g3y, g3z * g3z));
      const double g12 = fma(g1x, g2x, fma(g1y, g2y, g1z * g2z));
      const double g13 = fma(g1x, g3x, fma(g1y, g3y, g1z * g3z));
      const double g23 = fma(g2x, g3x, fma(g2y, g3y, g2z * g3z));
      apq -= w * (g11 + g12 + g13);
      cp -= w * (g12 + g22 + g23);
      cq += w * g12;
      // r_k is in no later cell of the chain: its entries with p and q are complete
      emit(op, lp, oc, cur, rank_byte(3 * k), cp);
      emit(oq, lq, oc, cur, rank_byte(3 * k + 1), cq);
      emit(oc, cur, on, ln, rank_byte(3 * k + 2), w * g23);
      cp = -w * (g13 + g23 + g33);
      cq = w * g13;
      if constexpr (WANT_B) {
        // load vector: b_a += |det| / 120 (f_a + sum of the cell's f)
        const double wb = adet * (1.0 / 120.0);
        const double F = fpq + fr + fn;
        swf = fma(wb, F, swf);
        sw += wb;
        emit_b(oc, cur, fma(wb, F + fr, cb));
        cb = wb * (F + fn);
      }
      e2x = e3x;
      e2y = e3y;
      e2z = e3z;
      kr = nzk.y;
      fr = fn;
      cur = ln;
      oc = on;
    }
  }
  emit(op, lp, oc, cur, rank_byte(18), cp);
  emit(oq, lq, oc, cur, rank_byte(19), cq);
  emit(op, lp, oq, lq, rank_byte(20), apq);
  if constexpr (WANT_B) {
    emit_b(oc, cur, cb);
    emit_b(op, lp, fma(fv(lp), sw, swf));
    emit_b(oq, lq, fma(fv(lq), sw, swf));
  }
}

// DPP move with bound control (every lane of the row is active here, so no old value is
// needed): VALU only, no LDS-pipe permute.
template <int CTRL>
__device__ __forceinline__ double dpp_bc(double v) {
  const int2 a = *reinterpret_cast<int2*>(&v);
  int2 r;
  r.x = __builtin_amdgcn_mov_dpp(a.x, CTRL, 0xF, 0xF, true);
  r.y = __builtin_amdgcn_mov_dpp(a.y, CTRL, 0xF, 0xF, true);
  return *reinterpret_cast<double*>(&r);
}

// The same over 16-lane rows: the quad sums, then rotations of the row by 4 and 8 lanes.
__device__ __forceinline__ double row16_sum_bc(double v) {
  v += dpp_bc<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_bc<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_bc<0x124>(v);  // row_ror 4
  v += dpp_bc<0x128>(v);  // row_ror 8
  return v;
}

// A block header read through the constant address space: the plan is read-only for the whole
// launch, so the (uniform) load is a scalar load into SGPRs.
__device__ __forceinline__ int4 block_header(const int4* hdr, int blk) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) int4 ConstInt4;
  return ((ConstInt4*)hdr)[blk];
#else
  return hdr[blk];
#endif
}

// Per-block data of the edge-chain kernels, loaded in two rounds: node ids, owned-row words
// and chain records (addressed from the block index and its header), then the nodes'
// coordinates and coefficients; held in registers until stored to LDS.
template <int T, int RR>
struct ChainLoad {
  static constexpr int NSL = (kMaxChainBlockNodes + T - 1) / T;  // node slots per thread
  int32_t v[NSL], kd[NSL];
  int2 rw;
  int32_t dof;
  uint32_t d[RR][kChainWords];
  double2 xy[NSL], zk[NSL];
  double f[NSL];
};

// round 1
template <bool WANT_B, int T, int RR>
__device__ __forceinline__ void chain_load_ids(const P1BlockTopo& t, int blk, const int4& hd,
                                               ChainLoad<T, RR>& c) {
  constexpr int NSL = ChainLoad<T, RR>::NSL;
  const int tid = threadIdx.x;
  const int SN = t.slot_nodes, SR = t.slot_rows;
  const int64_t sn0 = static_cast<int64_t>(blk) * SN, sr0 = static_cast<int64_t>(blk) * SR;
#pragma unroll
  for (int h = 0; h < NSL; ++h) {
    const int i = tid + h * T;
    c.v[h] = 0;
    c.kd[h] = 0;
    if (i < SN) {
      c.v[h] = t.slot_vertex[sn0 + i];
      c.kd[h] = t.slot_kdof ? t.slot_kdof[sn0 + i] : c.v[h];
    }
  }
  c.rw = make_int2(0, 0);
  c.dof = 0;
  if (tid < SR) {
    c.rw = t.slot_row[sr0 + tid];
    if constexpr (WANT_B) c.dof = t.slot_rdof[sr0 + tid];
  }
  const uint4* __restrict__ crec = reinterpret_cast<const uint4*>(t.chain_rec) + 3 * static_cast<int64_t>(hd.x);
#pragma unroll
  for (int j = 0; j < RR; ++j) {
    const int i = tid + j * T;
#pragma unroll
    for (int h = 0; h < 3; ++h) {
      const uint4 w = i < hd.y ? crec[3 * i + h] : make_uint4(0, 0, 0, 0);
      c.d[j][4 * h] = w.x;
      c.d[j][4 * h + 1] = w.y;
      c.d[j][4 * h + 2] = w.z;
      c.d[j][4 * h + 3] = w.w;
    }
  }
}

// round 2 (padding slots repeat node 0: harmless reads)
template <bool WANT_B, int T, int RR>
__device__ __forceinline__ void chain_gather(const P1BlockTopo& t, const P1Pass& p, ChainLoad<T, RR>& c) {
  constexpr int NSL = ChainLoad<T, RR>::NSL;
  const int tid = threadIdx.x;
#pragma unroll
  for (int h = 0; h < NSL; ++h) {
    c.xy[h] = make_double2(0.0, 0.0);
    c.zk[h] = make_double2(0.0, 0.0);
    c.f[h] = 0.0;
    if (tid + h * T < t.slot_nodes) {
      const int64_t vx = 3 * static_cast<int64_t>(c.v[h]);
      c.xy[h] = make_double2(p.x[vx], p.x[vx + 1]);
      c.zk[h] = make_double2(p.x[vx + 2], p.kappa[c.kd[h]]);
      c.f[h] = WANT_B ? p.f[c.kd[h]] : 0.0;
    }
  }
}

// LDS images of one block: node records x, y | z, kappa | z, f, owned-row words (CSR start,
// length | diagonal << 8), row dofs, zeroed value rows and load-vector entries.
template <bool WANT_B, int T, int RR>
__device__ __forceinline__ void chain_store(const P1BlockTopo& t, const LdsLayout& L, unsigned char* lds,
                                            int nr, const ChainLoad<T, RR>& c) {
  constexpr int NSL = ChainLoad<T, RR>::NSL;
  double* snode = reinterpret_cast<double*>(lds + L.node);
  double* svals = reinterpret_cast<double*>(lds + L.vals);
  double* sbv = reinterpret_cast<double*>(lds + L.bv);
  int32_t* srdof = reinterpret_cast<int32_t*>(lds + L.rdof);
  int2* srow = reinterpret_cast<int2*>(lds + L.rbase);
  const int tid = threadIdx.x;
#pragma unroll
  for (int h = 0; h < NSL; ++h) {
    const int i = tid + h * T;
    if (i < t.slot_nodes) {
      double2* rec = reinterpret_cast<double2*>(snode + i * 6);
      rec[0] = c.xy[h];
      rec[1] = c.zk[h];
      rec[2] = make_double2(c.zk[h].x, c.f[h]);
    }
  }
  if (tid < t.slot_rows) {
    srow[tid] = c.rw;
    if constexpr (WANT_B) {
      srdof[tid] = c.dof;
      sbv[tid] = 0.0;
    }
  }
  for (int i = tid; i < (nr * kRowStride + 1) / 2; i += T)
    reinterpret_cast<double2*>(svals)[i] = make_double2(0.0, 0.0);
}

// Write-out of one block from LDS: 16 lanes per row (a row holds at most 16 entries; the walk
// never writes a slot at or past the row length, nor the diagonal, so both read as zero); the
// diagonal is minus the row sum.
template <bool WANT_B, int T>
__device__ __forceinline__ void chain_write(const P1Pass& p, const LdsLayout& L, unsigned char* lds, int nr) {
  const double* svals = reinterpret_cast<const double*>(lds + L.vals);
  const double* sbv = reinterpret_cast<const double*>(lds + L.bv);
  const int32_t* srdof = reinterpret_cast<const int32_t*>(lds + L.rdof);
  const int2* srow = reinterpret_cast<const int2*>(lds + L.rbase);
  const int tid = threadIdx.x;
  const int grp = tid >> 4, ln = tid & 15;
  for (int i = grp; i < nr; i += T / 16) {
    const int2 w = srow[i];
    const int len = w.y & 0xff, dg = w.y >> 8;
    double v = svals[i * kRowStride + ln];
    const double rs = row16_sum_bc(v);
    if (ln == dg) v = -rs;
    if (ln < len) {
      double* out = p.values + static_cast<int64_t>(w.x) + ln;
      *out = p.add ? *out + v : v;
    }
  }
  if constexpr (WANT_B)
    for (int i = tid; i < nr; i += T) {
      const int32_t di = srdof[i];
      p.b[di] = p.add ? p.b[di] + sbv[i] : sbv[i];
    }
}

// One row block by edge chains (the batched kernel's unit).
template <bool WANT_B, int T, int RR>
__device__ __forceinline__ void p1_chain_body(const P1BlockTopo& t, const P1Pass& p,
                                              const LdsLayout& L, unsigned char* lds, int blk) {
  const int4 hd = block_header(t.blk_hdr, blk);  // first chain, chains, nodes, owned rows
  ChainLoad<T, RR> c;
  chain_load_ids<WANT_B, T, RR>(t, blk, hd, c);
  chain_gather<WANT_B, T, RR>(t, p, c);
  chain_store<WANT_B, T, RR>(t, L, lds, hd.w, c);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RR; ++j)
    if (static_cast<int>(threadIdx.x) + j * T < hd.y)
      chain_walk<WANT_B>(c.d[j], reinterpret_cast<const double*>(lds + L.node),
                         reinterpret_cast<double*>(lds + L.vals), reinterpret_cast<double*>(lds + L.bv),
                         static_cast<uint32_t>(hd.w));
  __syncthreads();
  chain_write<WANT_B, T>(p, L, lds, hd.w);
}

}  // namespace p1dev
}  // namespace dfa
