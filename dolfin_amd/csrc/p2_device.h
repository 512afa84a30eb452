// Device-side building blocks of the vector-P2 elasticity kernels (exact element weights,
// node-pair contraction, the blocked-scatter block body), shared by the standalone kernels
// and the batched persistent kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.h"
#include "kernels.h"

namespace dfa {
namespace {
constexpr int kThreads = 256;
constexpr int kCells = 16;  // cells per workgroup iteration

__constant__ int kEdge[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};

// int lambda_n lambda_m lambda_r / |T| on a tetrahedron.
__device__ __forceinline__ double bary3(int n, int m, int r) {
  if (n == m && m == r) return 1.0 / 20.0;
  if (n == m || m == r || n == r) return 1.0 / 60.0;
  return 1.0 / 120.0;
}

struct CellData {
  double D[4][3];      // grad lambda
  double Wm[4][4], Wl[4][4];
  double Rm[4], Rl[4]; // row sums of W
  double Tm, Tl;       // total sums
  double vol;
  int32_t dof[30];
  int32_t rbase[30];   // row_ptr[dof]
  int32_t deg[10];     // neighbours per node (row length / 3)
  uint8_t rank[100];
  uint8_t bcf[30];
  double bcv[30];
  double be[30];
};

// Geometry and exact coefficient weights of one cell (held in LDS by the atomic kernel, in
// registers by the blocked one).
struct CellGeom {
  double D[4][3];
  double Wm[4][4], Wl[4][4];
  double Rm[4], Rl[4];
  double Tm, Tl;
  double vol;
};

template <class C>
__device__ __forceinline__ void vp2_cell_setup(const double (&p)[4][3], const double (&mu)[4],
                                               const double (&la)[4], C& c) {
  double e1[3], e2[3], e3[3];
  for (int i = 0; i < 3; ++i) {
    e1[i] = p[1][i] - p[0][i];
    e2[i] = p[2][i] - p[0][i];
    e3[i] = p[3][i] - p[0][i];
  }
  double g[4][3];
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
  const double inv = 1.0 / det;
  for (int i = 0; i < 3; ++i) g[0][i] = -(g[1][i] + g[2][i] + g[3][i]);
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 3; ++i) c.D[k][i] = g[k][i] * inv;
  c.vol = fabs(det) / 6.0;
  c.Tm = c.Tl = 0.0;
  for (int m = 0; m < 4; ++m) {
    c.Rm[m] = c.Rl[m] = 0.0;
    for (int r = 0; r < 4; ++r) {
      double wm = 0.0, wl = 0.0;
      for (int n = 0; n < 4; ++n) {
        const double w = bary3(n, m, r);
        wm += mu[n] * w;
        wl += la[n] * w;
      }
      c.Wm[m][r] = c.vol * wm;
      c.Wl[m][r] = c.vol * wl;
      c.Rm[m] += c.vol * wm;
      c.Rl[m] += c.vol * wl;
    }
    c.Tm += c.Rm[m];
    c.Tl += c.Rl[m];
  }
}

// Factor of grad phi_a: terms (kind, index, gradient index). kind 0: u = 4 e_k - 1;
// kind 1: u = 4 e_k.
struct Factor {
  int n;
  int kind[2], idx[2], grad[2];
};

__device__ __forceinline__ Factor factor_of(int a) {
  Factor f;
  if (a < 4) {
    f.n = 1;
    f.kind[0] = 0;
    f.idx[0] = a;
    f.grad[0] = a;
  } else {
    const int k = kEdge[a - 4][0], l = kEdge[a - 4][1];
    f.n = 2;
    f.kind[0] = 1; f.idx[0] = l; f.grad[0] = k;
    f.kind[1] = 1; f.idx[1] = k; f.grad[1] = l;
  }
  return f;
}

// u^T W u' for the two factor kinds.
__device__ __forceinline__ double alpha(int k1, int i1, int k2, int i2, const double (&W)[4][4],
                                        const double (&R)[4], double T) {
  if (k1 == 0 && k2 == 0) return 16.0 * W[i1][i2] - 4.0 * R[i1] - 4.0 * R[i2] + T;
  if (k1 == 0) return 4.0 * (4.0 * W[i1][i2] - R[i2]);
  if (k2 == 0) return 4.0 * (4.0 * W[i1][i2] - R[i1]);
  return 16.0 * W[i1][i2];
}

// 3x3 block K[(i,a),(j,b)] of the elasticity element matrix.
template <class C>
__device__ __forceinline__ void vp2_block(const C& c, int a, int b, double (&K)[3][3]) {
  const Factor fa = factor_of(a), fb = factor_of(b);
  double Sm[3][3] = {}, Sl[3][3] = {};
  for (int t = 0; t < fa.n; ++t)
    for (int s = 0; s < fb.n; ++s) {
      const double am = alpha(fa.kind[t], fa.idx[t], fb.kind[s], fb.idx[s], c.Wm, c.Rm, c.Tm);
      const double al = alpha(fa.kind[t], fa.idx[t], fb.kind[s], fb.idx[s], c.Wl, c.Rl, c.Tl);
      const double* wa = c.D[fa.grad[t]];
      const double* wb = c.D[fb.grad[s]];
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const double o = wa[p] * wb[q];
          Sm[p][q] += am * o;
          Sl[p][q] += al * o;
        }
    }
  const double tr = Sm[0][0] + Sm[1][1] + Sm[2][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) K[i][j] = (i == j ? tr : 0.0) + Sm[j][i] + Sl[i][j];
}

}  // namespace

namespace p2dev {

// grad phi_a = sum_t (u_t . lambda) w_t with t = 0, 1:
//   vertex a:     u_0 = 4 e_a - 1, w_0 = grad lambda_a, w_1 = 0;
//   edge (k, l):  u_0 = 4 e_l, w_0 = grad lambda_k;  u_1 = 4 e_k, w_1 = grad lambda_l.
// For a compile-time node b the scalars alpha_ts = v_t . u'_s (v_t = u_t^T W) need only
// constant indices, so nothing below indexes a register array at run time.
__device__ __forceinline__ double sel4(int i, double a0, double a1, double a2, double a3) {
  return i == 0 ? a0 : (i == 1 ? a1 : (i == 2 ? a2 : a3));
}

// ---------------------------------------------------------------------------------------
// Lean per-cell data for the blocked kernel. With c P1 (nodal values c_m, S = sum c_m) the
// exact weights have the closed form
//   W_mr = int c lambda_m lambda_r = vol (1 + delta_mr)(S + c_m + c_r) / 120,
//   R_m  = sum_r W_mr = vol (S + c_m) / 20,
// so u_t^T W needs only c at the factor's index. Vertex and edge nodes get separate code
// paths (one and two gradient factors), selected per work item.
// ---------------------------------------------------------------------------------------
struct LeanGeom {
  double D[4][3];
  double cm[4], cl[4];
  double Sm, Sl, vol;
};

__device__ __forceinline__ void lean_setup(const double (&p)[4][3], const double (&mu)[4],
                                           const double (&la)[4], LeanGeom& g) {
  double e1[3], e2[3], e3[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e1[i] = p[1][i] - p[0][i];
    e2[i] = p[2][i] - p[0][i];
    e3[i] = p[3][i] - p[0][i];
  }
  double c[4][3];
  c[1][0] = e2[1] * e3[2] - e2[2] * e3[1];
  c[1][1] = e2[2] * e3[0] - e2[0] * e3[2];
  c[1][2] = e2[0] * e3[1] - e2[1] * e3[0];
  c[2][0] = e3[1] * e1[2] - e3[2] * e1[1];
  c[2][1] = e3[2] * e1[0] - e3[0] * e1[2];
  c[2][2] = e3[0] * e1[1] - e3[1] * e1[0];
  c[3][0] = e1[1] * e2[2] - e1[2] * e2[1];
  c[3][1] = e1[2] * e2[0] - e1[0] * e2[2];
  c[3][2] = e1[0] * e2[1] - e1[1] * e2[0];
  const double det = e1[0] * c[1][0] + e1[1] * c[1][1] + e1[2] * c[1][2];
  double inv = __builtin_amdgcn_rcp(det);
  double er = fma(-det, inv, 1.0);
  inv = fma(inv, er, inv);
  er = fma(-det, inv, 1.0);
  inv = fma(inv, er, inv);
#pragma unroll
  for (int i = 0; i < 3; ++i) c[0][i] = -(c[1][i] + c[2][i] + c[3][i]);
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i) g.D[k][i] = c[k][i] * inv;
  g.vol = fabs(det) * (1.0 / 6.0);
  g.Sm = mu[0] + mu[1] + mu[2] + mu[3];
  g.Sl = la[0] + la[1] + la[2] + la[3];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    g.cm[k] = mu[k];
    g.cl[k] = la[k];
  }
}

// Node a of a work item: gradient directions w_t and, per coefficient field, the scalars
// that fix v_t = u_t^T W (W in closed form above) at any index m:
//   vertex i:    v_0[m] = s (m == i ? 2S + 10 c_i : 4 c_i - 2S - 2 c_m),   sum_m v_0[m] = s (24 c_i - 6S)
//   edge (k, l): v_0[m] = 4s (1 + [m == l]) (S + c_l + c_m), v_1[m] = 4s (1 + [m == k]) (S + c_k + c_m)
// with s = vol / 120. Only the base values c_i / c_l / c_k are kept in registers; v_t[m] is
// formed where a compile-time m needs it.
struct NodeW {
  double w[2][3];
  double bm[2], bl[2];  // c_i (vertex) or c_l, c_k (edge), per field
  int i0, i1;           // i (vertex) or l, k (edge)
};

template <bool VERT>
__device__ __forceinline__ void node_w(const LeanGeom& g, int a, NodeW& f) {
  if constexpr (VERT) {
    f.i0 = a;
    f.i1 = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) f.w[0][q] = sel4(a, g.D[0][q], g.D[1][q], g.D[2][q], g.D[3][q]);
    f.bm[0] = sel4(a, g.cm[0], g.cm[1], g.cm[2], g.cm[3]);
    f.bl[0] = sel4(a, g.cl[0], g.cl[1], g.cl[2], g.cl[3]);
  } else {
    const int e = a - 4;
    const int k = (22 >> (2 * e)) & 3, l = (1775 >> (2 * e)) & 3;
    f.i0 = l;
    f.i1 = k;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      f.w[0][q] = sel4(k, g.D[0][q], g.D[1][q], g.D[2][q], g.D[3][q]);
      f.w[1][q] = sel4(l, g.D[0][q], g.D[1][q], g.D[2][q], g.D[3][q]);
    }
    f.bm[0] = sel4(l, g.cm[0], g.cm[1], g.cm[2], g.cm[3]);
    f.bl[0] = sel4(l, g.cl[0], g.cl[1], g.cl[2], g.cl[3]);
    f.bm[1] = sel4(k, g.cm[0], g.cm[1], g.cm[2], g.cm[3]);
    f.bl[1] = sel4(k, g.cl[0], g.cl[1], g.cl[2], g.cl[3]);
  }
}

// v_t[M] (without the factor s) for field values S, c_M and the node's base value.
template <bool VERT, int M>
__device__ __forceinline__ double node_v_at(const NodeW& f, int t, double S, double cM, double base) {
  if constexpr (VERT) {
    return f.i0 == M ? 2.0 * S + 10.0 * base : 4.0 * base - 2.0 * S - 2.0 * cM;
  } else {
    const int idx = t == 0 ? f.i0 : f.i1;
    return (idx == M ? 8.0 : 4.0) * (S + base + cM);
  }
}

// Adds K[(i,a),(j,B)] (B compile-time) to LDS: row(i) + col(j) addresses the entry.
template <int B, bool VERT, class Emit>
__device__ __forceinline__ void node_block_emit(const LeanGeom& g, const NodeW& f, Emit&& emit) {
  constexpr bool VB = B < 4;
  constexpr int EK = VB ? B : (22 >> (2 * (B - 4))) & 3;
  constexpr int EL = VB ? 0 : (1775 >> (2 * (B - 4))) & 3;
  constexpr int NT = VERT ? 1 : 2;
  const double s = g.vol * (1.0 / 120.0);
  // z_t = sum_s alpha_ts grad lambda_{grd_s}, per field
  double zm[2][3], zl[2][3];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const double bm = f.bm[t], bl = f.bl[t];
    double am0, al0, am1 = 0.0, al1 = 0.0;
    if constexpr (VB) {
      // u'_0 = 4 e_B - 1: alpha = 4 v_t[B] - sum_m v_t[m]
      const double summ = VERT ? 24.0 * bm - 6.0 * g.Sm : 24.0 * (g.Sm + bm);
      const double suml = VERT ? 24.0 * bl - 6.0 * g.Sl : 24.0 * (g.Sl + bl);
      am0 = s * (4.0 * node_v_at<VERT, B>(f, t, g.Sm, g.cm[B], bm) - summ);
      al0 = s * (4.0 * node_v_at<VERT, B>(f, t, g.Sl, g.cl[B], bl) - suml);
    } else {
      // u'_0 = 4 e_EL (direction EK), u'_1 = 4 e_EK (direction EL)
      am0 = 4.0 * s * node_v_at<VERT, EL>(f, t, g.Sm, g.cm[EL], bm);
      al0 = 4.0 * s * node_v_at<VERT, EL>(f, t, g.Sl, g.cl[EL], bl);
      am1 = 4.0 * s * node_v_at<VERT, EK>(f, t, g.Sm, g.cm[EK], bm);
      al1 = 4.0 * s * node_v_at<VERT, EK>(f, t, g.Sl, g.cl[EK], bl);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      zm[t][q] = am0 * g.D[EK][q];
      zl[t][q] = al0 * g.D[EK][q];
      if constexpr (!VB) {
        zm[t][q] = fma(am1, g.D[EL][q], zm[t][q]);
        zl[t][q] = fma(al1, g.D[EL][q], zl[t][q]);
      }
    }
  }
  // K[i][j] = delta_ij sum_t w_t . zm_t + sum_t (w_t[j] zm_t[i] + w_t[i] zl_t[j])
  double tr = 0.0;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int q = 0; q < 3; ++q) tr = fma(f.w[t][q], zm[t][q], tr);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double v = i == j ? tr : 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t) v = fma(f.w[t][j], zm[t][i], fma(f.w[t][i], zl[t][j], v));
      emit(i, j, v);
    }
}

constexpr int kBT = 256;

struct VP2Lds {
  int vert, vals, info, rows, base, bacc, total;
};

__host__ __device__ inline int al16(int n) { return (n + 15) & ~15; }

__host__ __device__ inline VP2Lds vp2_lds(int maxV, int maxO, int maxVals, bool b) {
  VP2Lds L;
  int o = 0;
  L.vert = o; o += al16(maxV * 6 * 8);
  L.vals = o; o += al16(maxVals * 8);
  L.info = o; o += al16(maxO * 4);
  L.rows = o; o += al16(maxO * 3 * 4);
  L.base = o; o += al16(maxO * 3 * 4);
  L.bacc = o; o += b ? al16(maxO * 3 * 8) : 0;
  L.total = o;
  return L;
}


// One node block of the blocked vector-P2 scatter (standalone kernel or batched work unit).
template <bool WANT_A, bool WANT_B, bool BC>
__device__ __forceinline__ void vp2_block_body(const VP2BlockTopo& t, const VP2Pass& p,
                                               const VP2Lds& L, unsigned char* lds, int blk) {
  double* svert = reinterpret_cast<double*>(lds + L.vert);
  double* svals = reinterpret_cast<double*>(lds + L.vals);
  uint32_t* sinfo = reinterpret_cast<uint32_t*>(lds + L.info);
  int32_t* srows = reinterpret_cast<int32_t*>(lds + L.rows);
  int32_t* sbase = reinterpret_cast<int32_t*>(lds + L.base);
  double* sbacc = reinterpret_cast<double*>(lds + L.bacc);
  const int tid = threadIdx.x;

  const int v0 = t.vert_ptr[blk], nv = t.vert_ptr[blk + 1] - v0;
  const int o0 = t.own_ptr[blk], no = t.own_ptr[blk + 1] - o0;
  const int t0 = t.task_ptr[blk], nt = t.task_ptr[blk + 1] - t0;
  const int nvx = t.task_nvert[blk], ned = nt - nvx;
  const int pv = (nvx + 63) & ~63;
  // Work items are laid out in wavefront-aligned segments (below); item kk -> task k.
  auto item = [&](int kk, int& k, bool& vert) {
    vert = kk < pv;
    k = vert ? kk : nvx + (kk - pv);
    return vert ? kk < nvx : kk < pv + ned;
  };
  const bool need_c = WANT_A || BC;  // the body force alone needs no coefficients
  // The prologue's loads in two rounds, each issued before its first wait: the first slot of
  // every per-thread list with its index clamped into the block (a block has at least one
  // vertex and owned node; clamped entries are read, never stored), then the vertex gathers
  // and the CSR row starts. (Guarded loads, one per loop iteration, made every load wait for
  // the previous one.)
  const int iv = min(tid, nv - 1), io = min(tid, no - 1), ir = min(tid, 3 * no - 1);
  const int64_t vid = t.vert_id[v0 + iv];
  const int32_t vkd = t.vert_kdof[v0 + iv];
  const uint32_t oinfo = t.own_info[o0 + io];
  const int32_t orow = t.own_rows[3 * o0 + ir];
  {
    const double2 r0 = make_double2(p.x[vid * 3], p.x[vid * 3 + 1]);
    const double2 r1 = make_double2(p.x[vid * 3 + 2], need_c ? p.mu[vkd] : 0.0);
    const double2 r2 = make_double2(need_c ? p.lmbda[vkd] : 0.0, 0.0);
    const int32_t obase = WANT_A ? t.row_ptr[orow] : 0;
    if (tid < nv) {
      double* rec = svert + tid * 6;
      *reinterpret_cast<double2*>(rec) = r0;
      *reinterpret_cast<double2*>(rec + 2) = r1;
      *reinterpret_cast<double2*>(rec + 4) = r2;
    }
    if (tid < no) sinfo[tid] = oinfo;
    // owned rows and their CSR starts, fetched now so that the write-out issues no loads
    if (tid < 3 * no) {
      srows[tid] = orow;
      if (WANT_A) sbase[tid] = obase;
    }
  }
  for (int i = tid + kBT; i < nv; i += kBT) {  // blocks of more than kBT vertices
    const int64_t v = t.vert_id[v0 + i];
    const int32_t kd = t.vert_kdof[v0 + i];
    double* rec = svert + i * 6;
    *reinterpret_cast<double2*>(rec) = make_double2(p.x[v * 3], p.x[v * 3 + 1]);
    *reinterpret_cast<double2*>(rec + 2) = make_double2(p.x[v * 3 + 2], need_c ? p.mu[kd] : 0.0);
    *reinterpret_cast<double2*>(rec + 4) = make_double2(need_c ? p.lmbda[kd] : 0.0, 0.0);
  }
  for (int i = tid + kBT; i < no; i += kBT) sinfo[i] = t.own_info[o0 + i];
  for (int i = tid + kBT; i < 3 * no; i += kBT) {
    const int32_t row = t.own_rows[3 * o0 + i];
    srows[i] = row;
    if (WANT_A) sbase[i] = t.row_ptr[row];
  }
  if (WANT_A) {
    const int nvals = t.blk_vals[blk];
    for (int i = tid; i < (nvals + 1) / 2; i += kBT)
      reinterpret_cast<double2*>(svals)[i] = make_double2(0.0, 0.0);
  }
  if (WANT_B)
    for (int i = tid; i < 3 * no; i += kBT) sbacc[i] = 0.0;
  __syncthreads();

  // Work is laid out in wavefront-aligned segments so that every wavefront runs one code path:
  // the vertex-node items (one gradient factor), then the edge-node items; each item (a cell
  // and one of its owned nodes) is one thread over all ten node pairs b. (Splitting edge items
  // over two threads, five pairs each, made a block take five wavefront passes for four
  // wavefronts: 7 % slower on the 64^3 workload.)
  for (int kk = tid; kk < pv + ned; kk += kBT) {
    int k;
    bool vert;
    if (!item(kk, k, vert)) continue;
    // task words are read once: streaming loads keep L2 for the vertex gathers
    const uint64_t w0 = __builtin_nontemporal_load(t.task + 3 * (t0 + k));
    const uint64_t w1 = __builtin_nontemporal_load(t.task + 3 * (t0 + k) + 1);
    const uint64_t w2 = __builtin_nontemporal_load(t.task + 3 * (t0 + k) + 2);
    const int o = static_cast<int>(w1 & 0xffffu);
    const int a = static_cast<int>((w1 >> 16) & 0xffu);
    double pc[4][3], mu[4], la[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double* rec = svert + static_cast<int>((w0 >> (16 * v)) & 0xffffu) * 6;
      const double2 r0 = *reinterpret_cast<const double2*>(rec);
      const double2 r1 = *reinterpret_cast<const double2*>(rec + 2);
      const double2 r2 = *reinterpret_cast<const double2*>(rec + 4);
      pc[v][0] = r0.x;
      pc[v][1] = r0.y;
      pc[v][2] = r1.x;
      mu[v] = r1.y;
      la[v] = r2.x;
    }
    LeanGeom c;
    lean_setup(pc, mu, la, c);
    const uint32_t info = sinfo[o];
    const int off = static_cast<int>(info & 0xffffu);
    const int deg = static_cast<int>(info >> 16);
    double be[3] = {0.0, 0.0, 0.0};  // Dirichlet lifting here; the body force is added below
    // Dirichlet data of the cell (rows of node a, columns of every node b)
    uint32_t bcrow = 0, bccol = 0;  // bit i: row (i,a); bit j*10+b: column (j,b)
    double growv[3] = {0, 0, 0};
    const int32_t* cdofs = nullptr;
    if constexpr (BC) {
      cdofs = t.cell_dofs + static_cast<int64_t>(t.task_cell[t0 + k]) * 30;
      for (int e = 0; e < 30; ++e)
        if (p.bc_mask[cdofs[e]]) bccol |= 1u << e;
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (bccol >> (i * 10 + a) & 1u) {
          bcrow |= 1u << i;
          growv[i] = p.bc_value[cdofs[i * 10 + a]];
        }
    }
    if constexpr (WANT_A || BC) {
      auto run = [&](auto vconst) {
        constexpr bool VERT = decltype(vconst)::value;
        NodeW fa;
        node_w<VERT>(c, a, fa);
        auto one = [&](auto bconst) {
          constexpr int b = decltype(bconst)::value;
          int rk;
          if constexpr (b < 5) rk = static_cast<int>((w1 >> (24 + 8 * b)) & 0xffu);
          else rk = static_cast<int>((w2 >> (8 * (b - 5))) & 0xffu);
          const int s2 = t.component_major ? deg : 1;
          double* base = svals + off + (t.component_major ? rk : 3 * rk);
          node_block_emit<b, VERT>(c, fa, [&](int i, int j, double v) {
            if constexpr (BC) {
              const bool cb = (bccol >> (j * 10 + b)) & 1u;
              if (WANT_B && cb && !(bcrow >> i & 1u)) be[i] -= v * p.bc_value[cdofs[j * 10 + b]];
              if (cb || (bcrow >> i & 1u)) v = (i == j && a == b) ? 1.0 : 0.0;
            }
            if constexpr (WANT_A) unsafeAtomicAdd(base + i * 3 * deg + j * s2, v);
          });
        };
        one(std::integral_constant<int, 0>{});
        one(std::integral_constant<int, 1>{});
        one(std::integral_constant<int, 2>{});
        one(std::integral_constant<int, 3>{});
        one(std::integral_constant<int, 4>{});
        one(std::integral_constant<int, 5>{});
        one(std::integral_constant<int, 6>{});
        one(std::integral_constant<int, 7>{});
        one(std::integral_constant<int, 8>{});
        one(std::integral_constant<int, 9>{});
      };
      // the segment decides the path: uniform over the wavefront
      if (vert) run(std::integral_constant<bool, true>{});
      else run(std::integral_constant<bool, false>{});
    }
    if (WANT_B) {
      // body force f3 . phi_a: -vol/20 at vertex nodes, vol/5 at edge nodes
#pragma unroll
      for (int i = 0; i < 3; ++i) be[i] += p.f3[i] * (a < 4 ? -c.vol / 20.0 : c.vol / 5.0);
    }
    if constexpr (WANT_B) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double v = (BC && (bcrow >> i & 1u)) ? growv[i] : be[i];
        unsafeAtomicAdd(sbacc + 3 * o + i, v);
      }
    }
  }
  __syncthreads();

  // write-out, one owned node per wavefront: its three rows (3 deg values each) are contiguous
  // in LDS and are three contiguous CSR ranges. Streaming (non-temporal) stores: the values
  // are not read again by this kernel, and keeping 4.4 GB of them out of L2 measured 10 %
  // faster on the 64^3 workload.
  const int wave = tid >> 6, lane = tid & 63;
  for (int o = wave; o < no; o += kBT / 64) {
    const uint32_t info = sinfo[o];
    const int off = static_cast<int>(info & 0xffffu);
    const int len = 3 * static_cast<int>(info >> 16);  // one row
    if (WANT_A) {
      const int64_t b0 = sbase[3 * o], b1 = sbase[3 * o + 1], b2 = sbase[3 * o + 2];
      for (int j = lane; j < 3 * len; j += 64) {
        const int comp = j >= len ? (j >= 2 * len ? 2 : 1) : 0;
        const int64_t base = comp == 0 ? b0 : (comp == 1 ? b1 : b2);
        double* dst = p.values + base + (j - comp * len);
        const double v = p.add ? svals[off + j] + *dst : svals[off + j];
        __builtin_nontemporal_store(v, dst);
      }
    }
    if (WANT_B && lane < 3) {
      double* dst = p.b + srows[3 * o + lane];
      *dst = p.add ? *dst + sbacc[3 * o + lane] : sbacc[3 * o + lane];
    }
  }
}


}  // namespace p2dev
}  // namespace dfa
