// Adjoint (reverse-mode) assembly of J = u^T K(c, x) u - F(x)^T u for gfx950.
//
// K is never formed: each cell recomputes its geometry and adds its contribution to
// dJ/dc (nodal coefficient gradient) and dJ/dx (vertex coordinate gradient). With
// grad(u_h) piecewise polynomial on an affine cell,
//   d vol / d x_{l,q}        = vol * D_lq                  (D = grad lambda),
//   d grad(u_h) / d x_{l,q}  = -grad(u_h)[:, q] * D_l      (column q of grad u times D_l),
// the discrete form of the shape derivative dJ[V] = int k(-2 (grad V grad u).grad u +
// |grad u|^2 div V) - f u div V checked in python/test/unit/fem/test_coordinate_derivative.py:47-49.
// dJ/dkappa_k of the P1 Poisson energy is vol |grad u|^2 / (d+1) per cell containing k
// (DOLFIN's derivative(J, kappa) form). The sign of det J never enters: only |det J| does,
// and d|det J| = |det J| tr(J^{-1} dJ) for either orientation.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace dfa {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <int D>
__device__ __forceinline__ void simplex_grads(const double (&p)[D + 1][D], double (&G)[D + 1][D],
                                              double& vol) {
  if constexpr (D == 3) {
    double e1[3], e2[3], e3[3];
#pragma unroll
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
#pragma unroll
    for (int i = 0; i < 3; ++i) g[0][i] = -(g[1][i] + g[2][i] + g[3][i]);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i) G[k][i] = g[k][i] * inv;
    vol = fabs(det) / 6.0;
  } else if constexpr (D == 2) {
    const double e1x = p[1][0] - p[0][0], e1y = p[1][1] - p[0][1];
    const double e2x = p[2][0] - p[0][0], e2y = p[2][1] - p[0][1];
    const double det = e1x * e2y - e1y * e2x;
    const double inv = 1.0 / det;
    G[1][0] = e2y * inv;
    G[1][1] = -e2x * inv;
    G[2][0] = -e1y * inv;
    G[2][1] = e1x * inv;
    G[0][0] = -(G[1][0] + G[2][0]);
    G[0][1] = -(G[1][1] + G[2][1]);
    vol = fabs(det) / 2.0;
  } else {
    const double e = p[1][0] - p[0][0];
    G[1][0] = 1.0 / e;
    G[0][0] = -G[1][0];
    vol = fabs(e);
  }
}

template <int D>
__global__ __launch_bounds__(kThreads) void adjoint_p1_kernel(AdjointP1 a) {
  constexpr int N = D + 1;
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  double Jc = 0.0;
  if (c < a.num_cells) {
    int32_t v[N], d[N], k[N];
    double p[N][D], G[N][D], vol;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      v[i] = a.cells[c * N + i];
      d[i] = a.cell_dofs[c * N + i];
      k[i] = a.coeff_dofs[c * N + i];
#pragma unroll
      for (int q = 0; q < D; ++q) p[i][q] = a.x[static_cast<int64_t>(v[i]) * D + q];
    }
    simplex_grads<D>(p, G, vol);
    double gu[D] = {}, ks = 0.0, us = 0.0, fs = 0.0, uf = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double ui = a.u[d[i]], fi = a.f[k[i]];
#pragma unroll
      for (int q = 0; q < D; ++q) gu[q] += ui * G[i][q];
      ks += a.kappa[k[i]];
      us += ui;
      fs += fi;
      uf += ui * fi;
    }
    const double kb = ks / N;
    double g2 = 0.0;
#pragma unroll
    for (int q = 0; q < D; ++q) g2 += gu[q] * gu[q];
    const double E = kb * vol * g2;
    const double Gl = vol / ((D + 1.0) * (D + 2.0)) * (uf + us * fs);
    Jc = E - Gl;
#pragma unroll
    for (int i = 0; i < N; ++i) unsafeAtomicAdd(a.dJdk + k[i], vol * g2 / N);
#pragma unroll
    for (int l = 0; l < N; ++l) {
      double gdl = 0.0;
#pragma unroll
      for (int q = 0; q < D; ++q) gdl += gu[q] * G[l][q];
#pragma unroll
      for (int q = 0; q < D; ++q)
        unsafeAtomicAdd(a.dJdx + static_cast<int64_t>(v[l]) * D + q,
                        kb * vol * (g2 * G[l][q] - 2.0 * gu[q] * gdl) - Gl * G[l][q]);
    }
  }
  const double s = wave_sum(Jc);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(a.J, s);
}

__constant__ int kEdgeA[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};

__device__ __forceinline__ double bary3(int n, int m, int r) {
  if (n == m && m == r) return 1.0 / 20.0;
  if (n == m || m == r || n == r) return 1.0 / 60.0;
  return 1.0 / 120.0;
}

// Vector P2 elasticity energy E = int 2 mu eps:eps + lmbda (tr eps)^2 with
// grad u_h = sum_m lambda_m Gm (Gm 3x3), integrated exactly with W_mr = int c lambda_m lambda_r.
__global__ __launch_bounds__(kThreads) void adjoint_vp2_kernel(AdjointVP2 a) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  double Jc = 0.0;
  if (c < a.num_cells) {
    int32_t v[4];
    double p[4][3], D[4][3], vol;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = a.cells[c * 4 + i];
#pragma unroll
      for (int q = 0; q < 3; ++q) p[i][q] = a.x[static_cast<int64_t>(v[i]) * 3 + q];
    }
    simplex_grads<3>(p, D, vol);
    // Gm[i][p] = sum_a u_{ia} Q_a[m][p]
    double G[4][3][3] = {};
    double uload = 0.0;
    for (int i = 0; i < 3; ++i) {
      double fv = 0.0, fe = 0.0;
      for (int k = 0; k < 4; ++k) {
        const double u = a.u[a.cell_dofs[c * 30 + i * 10 + k]];
        fv += u;
        for (int m = 0; m < 4; ++m) {
          const double s = (m == k ? 3.0 : -1.0) * u;
#pragma unroll
          for (int q = 0; q < 3; ++q) G[m][i][q] += s * D[k][q];
        }
      }
      for (int e = 0; e < 6; ++e) {
        const double u = a.u[a.cell_dofs[c * 30 + i * 10 + 4 + e]];
        fe += u;
        const int k = kEdgeA[e][0], l = kEdgeA[e][1];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          G[l][i][q] += 4.0 * u * D[k][q];
          G[k][i][q] += 4.0 * u * D[l][q];
        }
      }
      uload += a.f3[i] * (-fv / 20.0 + fe / 5.0);
    }
    uload *= vol;
    double mu[4], la[4];
    int32_t kd[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      kd[n] = a.coeff_dofs[c * 4 + n];
      mu[n] = a.mu[kd[n]];
      la[n] = a.lmbda[kd[n]];
    }
    double tr[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) tr[m] = G[m][0][0] + G[m][1][1] + G[m][2][2];
    double ee[4][4], tt[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = m; r < 4; ++r) {
        double e = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int q = 0; q < 3; ++q)
            e += 0.25 * (G[m][i][q] + G[m][q][i]) * (G[r][i][q] + G[r][q][i]);
        ee[m][r] = ee[r][m] = 2.0 * e;
        tt[m][r] = tt[r][m] = tr[m] * tr[r];
      }
    double Wm[4][4], Wl[4][4];
    double E = 0.0;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double wm = 0.0, wl = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          wm += mu[n] * bary3(n, m, r);
          wl += la[n] * bary3(n, m, r);
        }
        Wm[m][r] = wm;
        Wl[m][r] = wl;
        E += vol * (wm * ee[m][r] + wl * tt[m][r]);
      }
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      double dm = 0.0, dl = 0.0;
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          dm += bary3(n, m, r) * ee[m][r];
          dl += bary3(n, m, r) * tt[m][r];
        }
      unsafeAtomicAdd(a.dJdmu + kd[n], vol * dm);
      unsafeAtomicAdd(a.dJdlmbda + kd[n], vol * dl);
    }
    // H[q][p] = sum_{m,i} Gm[i][q] Pm[i][p],  Pm = dE/dGm = 2 vol sum_r (Wm_mr 2 eps_r + Wl_mr tr_r I)
    double H[3][3] = {};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double P[3][3] = {};
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int q = 0; q < 3; ++q)
            P[i][q] += 2.0 * vol * (Wm[m][r] * (G[r][i][q] + G[r][q][i]) + (i == q ? Wl[m][r] * tr[r] : 0.0));
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
          for (int pp = 0; pp < 3; ++pp) H[q][pp] += G[m][i][q] * P[i][pp];
    }
    Jc = E - uload;
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double t = 0.0;
#pragma unroll
        for (int pp = 0; pp < 3; ++pp) t += H[q][pp] * D[l][pp];
        unsafeAtomicAdd(a.dJdx + static_cast<int64_t>(v[l]) * 3 + q, Jc * D[l][q] - t);
      }
  }
  const double s = wave_sum(Jc);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(a.J, s);
}

}  // namespace

hipError_t launch_adjoint_p1(const AdjointP1& a, hipStream_t s) {
  if (a.num_cells == 0) return hipSuccess;
  const unsigned grid = static_cast<unsigned>((a.num_cells + kThreads - 1) / kThreads);
  switch (a.gdim) {
    case 1: hipLaunchKernelGGL(adjoint_p1_kernel<1>, dim3(grid), dim3(kThreads), 0, s, a); break;
    case 2: hipLaunchKernelGGL(adjoint_p1_kernel<2>, dim3(grid), dim3(kThreads), 0, s, a); break;
    case 3: hipLaunchKernelGGL(adjoint_p1_kernel<3>, dim3(grid), dim3(kThreads), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_adjoint_vp2(const AdjointVP2& a, hipStream_t s) {
  if (a.num_cells == 0) return hipSuccess;
  const unsigned grid = static_cast<unsigned>((a.num_cells + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(adjoint_vp2_kernel, dim3(grid), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace dfa
