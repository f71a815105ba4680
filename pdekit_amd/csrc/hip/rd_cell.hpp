// Fused residual-distribution cell kernel for gfx950.
//
// One thread owns one element and performs, in registers, the work of
//   gather (GeometryCache / SolutionCache push_back_to_buffer,
//           src/interpolation/GeometryCache.hpp:329-400, SolutionCache.hpp:361-388)
//   geometry metric (GeometryMetric.hpp:366-425, 3D 487-560)
//   P+1 flux metric (FluxSpaceMetric.hpp:562-743)
//   element scheme  (PGLDA.hpp:109-311, PGN.hpp:96-290, PGB.hpp:314-498,
//                    PSI limiter LFResidualLimiter.hpp:84-110)
//   scatter res[v] += Phi_v, lambda[v] += |T| max lambda
//           (PGRDMExplicitCellWorker.hpp:546-567)
// The division by the dual volume is done once per node in the nodal pass.
//
// Algebraic forms used on the device (same mathematics as the reference):
//   * flux divergence at q: sum_f sum_xi dVf[xi][q][f] G_xi(f) with the
//     contravariant flux G_xi(f) = sum_d Jinv[d][xi] F_d(f) (affine cells);
//   * N dissipation  sum_{m!=n} K+_n S^-1 K+_m (U_n - U_m) = K+_n (U_n - S^-1 sum_m K+_m U_m),
//     S = sum_m K+_m, so one extra solve replaces the O(nd^2) pair loop;
//   * S^-1 is never formed: S x = r and S y = sum K+_m U_m are solved by
//     Gaussian elimination with partial pivoting done by predicated row swaps.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "rd_physics.hpp"

namespace pdrd {

struct CellArgs
{
  const double *__restrict__ tab;      // packed FE tables of the block
  const uint32_t *__restrict__ gconn;  // [NG][ne] (structure of arrays)
  const uint32_t *__restrict__ sconn;  // [ND][ne]
  const double *__restrict__ coords;   // [nv][D]
  const double *__restrict__ u;        // [N][M]
  double *__restrict__ acc_res;        // [N][M]
  double *__restrict__ acc_ws;         // [N]
  int *__restrict__ status;
  uint32_t first;  // first element of the launch (block order)
  uint32_t count;  // elements in the launch
  uint32_t ne;     // elements in the block (SoA stride)
  int atomic;      // 1: FP64 atomics, 0: plain read-modify-write (one colour)
  PhysParams pp;
};

template <int D, int NG, int ND, int NF, int NQ>
struct TabOffsets
{
  static constexpr int qp = 0;
  static constexpr int w = qp + NQ * D;
  static constexpr int Vg = w + NQ;
  static constexpr int dVg = Vg + NQ * NG;
  static constexpr int V = dVg + D * NQ * NG;
  static constexpr int dV = V + NQ * ND;
  static constexpr int Vf = dV + D * NQ * ND;
  static constexpr int dVf = Vf + NQ * NF;
  static constexpr int Vgf = dVf + D * NQ * NF;
  static constexpr int Vsf = Vgf + NF * NG;
  static constexpr int sol_nodes = Vsf + NF * ND;
  static constexpr int flux_nodes = sol_nodes + ND * D;
  static constexpr int size = flux_nodes + NF * D;
};

// Workgroups are dealt round-robin to the 8 XCDs; give each XCD one
// contiguous range of workgroups so neighbouring cells (which share nodes)
// meet in the same L2.
__device__ __forceinline__ uint32_t xcd_block_index()
{
  const uint32_t nb = gridDim.x, bid = blockIdx.x;
  const uint32_t x = bid & 7u, k = bid >> 3, q8 = nb >> 3, r8 = nb & 7u;
  return x * q8 + (x < r8 ? x : r8) + k;
}

__device__ __forceinline__ void scatter_add(double *p, double v, int atomic)
{
  if (atomic)
    unsafeAtomicAdd(p, v);
  else
    *p += v;
}

template <int D>
__device__ __forceinline__ double det_inv(const double J[D][D], double Ji[D][D])
{
  if constexpr (D == 2) {
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double id = 1.0 / det;
    Ji[0][0] = id * J[1][1];
    Ji[1][1] = id * J[0][0];
    Ji[0][1] = -id * J[0][1];
    Ji[1][0] = -id * J[1][0];
    return det;
  } else {
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double id = 1.0 / det;
    Ji[0][0] = c00 * id;
    Ji[1][0] = c01 * id;
    Ji[2][0] = c02 * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    return det;
  }
}

// Solve A X = B (M x M, R right-hand sides) in place; partial pivoting by
// predicated swaps so that every index stays a compile-time constant.
template <int M, int R>
__device__ __forceinline__ void solve_pp(double A[M][M], double B[M][R])
{
#pragma unroll
  for (int c = 0; c < M; ++c) {
#pragma unroll
    for (int r = c + 1; r < M; ++r) {
      const bool sw = fabs(A[r][c]) > fabs(A[c][c]);
#pragma unroll
      for (int j = c; j < M; ++j) {
        const double t = A[c][j];
        A[c][j] = sw ? A[r][j] : t;
        A[r][j] = sw ? t : A[r][j];
      }
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const double t = B[c][j];
        B[c][j] = sw ? B[r][j] : t;
        B[r][j] = sw ? t : B[r][j];
      }
    }
    const double ip = 1.0 / A[c][c];
#pragma unroll
    for (int r = c + 1; r < M; ++r) {
      const double f = A[r][c] * ip;
#pragma unroll
      for (int j = c + 1; j < M; ++j) A[r][j] -= f * A[c][j];
#pragma unroll
      for (int j = 0; j < R; ++j) B[r][j] -= f * B[c][j];
    }
  }
#pragma unroll
  for (int c = M - 1; c >= 0; --c) {
    const double ip = 1.0 / A[c][c];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      double s = B[c][j];
#pragma unroll
      for (int k = c + 1; k < M; ++k) s -= A[c][k] * B[k][j];
      B[c][j] = s * ip;
    }
  }
}

// y += coef * K+ x for the compact Euler K+ of one node.
template <int D>
__device__ __forceinline__ void apply_kplus(const Acoustic<D> &ac, double l0, double dp, double dm, const double *x,
                                            double coef, double *y)
{
  constexpr int M = D + 2;
  double sp = 0.0, sm = 0.0;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    sp += ac.lp[k] * x[k];
    sm += ac.lm[k] * x[k];
  }
  sp *= dp;
  sm *= dm;
#pragma unroll
  for (int k = 0; k < M; ++k) y[k] += coef * (l0 * x[k] + ac.rp[k] * sp + ac.rm[k] * sm);
}

template <int D, int NG, int ND, int NF, int NQ, bool AFFINE, int PH, int SC>
__global__ __launch_bounds__(256) void rd_cell_kernel(CellArgs a)
{
  constexpr int M = PhysTraits<PH>::M;
  static_assert(PhysTraits<PH>::DIM == D, "physics / element dimension mismatch");
  using O = TabOffsets<D, NG, ND, NF, NQ>;
  const uint32_t local = xcd_block_index() * blockDim.x + threadIdx.x;
  if (local >= a.count) return;
  const uint32_t e = a.first + local;
  const double *__restrict__ tab = a.tab;
  bool bad = false;

  // ---- gather ----------------------------------------------------------
  double X[NG][D];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const size_t v = a.gconn[static_cast<size_t>(g) * a.ne + e];
#pragma unroll
    for (int d = 0; d < D; ++d) X[g][d] = a.coords[v * D + d];
  }
  uint32_t nid[ND];
  double U[ND][M];
#pragma unroll
  for (int n = 0; n < ND; ++n) {
    nid[n] = a.sconn[static_cast<size_t>(n) * a.ne + e];
#pragma unroll
    for (int k = 0; k < M; ++k) U[n][k] = a.u[static_cast<size_t>(nid[n]) * M + k];
  }

  // ---- geometry --------------------------------------------------------
  constexpr int NJ = AFFINE ? 1 : NQ;
  double Jinv[NJ][D][D];
  double detJ[NJ];
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    double J[D][D];
#pragma unroll
    for (int xi = 0; xi < D; ++xi)
#pragma unroll
      for (int x = 0; x < D; ++x) {
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) s += tab[O::dVg + (xi * NQ + q) * NG + g] * X[g][x];
        J[xi][x] = s;
      }
    detJ[q] = det_inv<D>(J, Jinv[q]);
  }

  // ---- P+1 flux element: divergence of the interpolated flux at q -------
  double div[NQ][M];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int k = 0; k < M; ++k) div[q][k] = 0.0;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double s[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      double t = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) t += tab[O::Vsf + f * ND + n] * U[n][k];
      s[k] = t;
    }
    double F[D][M];
    point_flux<PH, double>(s, a.pp, F, bad);
    if constexpr (AFFINE) {
      double G[D][M];
#pragma unroll
      for (int xi = 0; xi < D; ++xi)
#pragma unroll
        for (int k = 0; k < M; ++k) {
          double t = 0.0;
#pragma unroll
          for (int d = 0; d < D; ++d) t += Jinv[0][d][xi] * F[d][k];
          G[xi][k] = t;
        }
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int k = 0; k < M; ++k) {
          double t = div[q][k];
#pragma unroll
          for (int xi = 0; xi < D; ++xi) t += tab[O::dVf + (xi * NQ + q) * NF + f] * G[xi][k];
          div[q][k] = t;
        }
    } else {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        double b[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          double t = 0.0;
#pragma unroll
          for (int xi = 0; xi < D; ++xi) t += Jinv[q][d][xi] * tab[O::dVf + (xi * NQ + q) * NF + f];
          b[d] = t;
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
          double t = div[q][k];
#pragma unroll
          for (int d = 0; d < D; ++d) t += b[d] * F[d][k];
          div[q][k] = t;
        }
      }
    }
  }

  // ---- element scheme ----------------------------------------------------
  constexpr bool NDIS = (SC != SC_LDA);
  double lda[ND][M];
  double dis[ND][M];
  double lam[ND];
  double fint[M];
#pragma unroll
  for (int n = 0; n < ND; ++n) {
    lam[n] = 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) { lda[n][k] = 0.0; dis[n][k] = 0.0; }
  }
#pragma unroll
  for (int k = 0; k < M; ++k) fint[k] = 0.0;
  double vol = 0.0;

#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int jq = AFFINE ? 0 : q;
    const double wj = tab[O::w + q] * detJ[jq];
    vol += wj;
#pragma unroll
    for (int k = 0; k < M; ++k) fint[k] += wj * div[q][k];
    // physical gradients of the solution basis at q
    double gphys[ND][D];
#pragma unroll
    for (int n = 0; n < ND; ++n)
#pragma unroll
      for (int x = 0; x < D; ++x) {
        double t = 0.0;
#pragma unroll
        for (int xi = 0; xi < D; ++xi) t += Jinv[jq][x][xi] * tab[O::dV + (xi * NQ + q) * ND + n];
        gphys[n][x] = t;
      }

    if constexpr (M == 1) {
      double vel0, vel1;
      if constexpr (PH == PH_ADVECTION2D) {
        vel0 = a.pp.p0;
        vel1 = a.pp.p1;
      } else {
        double uq = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) uq += tab[O::V + q * ND + n] * U[n][0];
        vel0 = uq;
        vel1 = 1.0;
      }
      double kp[ND];
      double S = 0.0, SU = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        const double kn = vel0 * gphys[n][0] + vel1 * gphys[n][1];
        kp[n] = fmax(0.0, kn);
        lam[n] = fmax(lam[n], fmax(1.e-6, fabs(kn)));
        S += kp[n];
        SU += kp[n] * U[n][0];
      }
      const double inv = 1.0 / S;
      const double r = div[q][0];
      const double ub = SU * inv;
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        lda[n][0] += wj * ((kp[n] * inv) * r);
        if constexpr (NDIS) dis[n][0] += wj * (kp[n] * (U[n][0] - ub));
      }
    } else {
      double uq[M];
#pragma unroll
      for (int k = 0; k < M; ++k) {
        double t = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) t += tab[O::V + q * ND + n] * U[n][k];
        uq[k] = t;
      }
      const EulerState<double, D> st = euler_state<double, D>(uq, a.pp.p0);
      bad = bad || st.bad;
      const double gm1 = a.pp.p0 - 1.0;
      // compact K+ of every node
      double l0[ND], dp[ND], dm[ND], nh[ND][D];
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        double nn = 0.0;
#pragma unroll
        for (int x = 0; x < D; ++x) nn += gphys[n][x] * gphys[n][x];
        const double norm = sqrt(nn);
        const double small = (D == 2) ? 1.e-9 : 0.0;
        const bool tiny = (D == 2) ? (norm < small) : (norm == 0.0);
        const double inorm = tiny ? 0.0 : 1.0 / norm;
        double um = 0.0;
#pragma unroll
        for (int x = 0; x < D; ++x) {
          nh[n][x] = gphys[n][x] * inorm;
          um += st.V[x] * nh[n][x];
        }
        const double e0 = norm * um, ep = norm * (um + st.a), em = norm * (um - st.a);
        lam[n] = fmax(lam[n], fmax(1.e-6, fmax(fabs(e0), fmax(fabs(ep), fabs(em)))));
        const double p0 = fmax(0.0, e0);
        l0[n] = tiny ? 0.0 : p0;
        dp[n] = tiny ? 0.0 : fmax(0.0, ep) - p0;
        dm[n] = tiny ? 0.0 : fmax(0.0, em) - p0;
      }
      // S = sum K+ ; Z = sum K+ U
      double A[M][M];
      double B[M][2];
      double sl0 = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) sl0 += l0[n];
#pragma unroll
      for (int i = 0; i < M; ++i) {
#pragma unroll
        for (int j = 0; j < M; ++j) A[i][j] = (i == j) ? sl0 : 0.0;
        B[i][0] = div[q][i];
        B[i][1] = 0.0;
      }
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        Acoustic<D> ac;
        acoustic_vectors<D>(st, gm1, nh[n], ac);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
          for (int j = 0; j < M; ++j) A[i][j] += dp[n] * ac.rp[i] * ac.lp[j] + dm[n] * ac.rm[i] * ac.lm[j];
        if constexpr (NDIS) {
          double z[M];
#pragma unroll
          for (int k = 0; k < M; ++k) z[k] = 0.0;
          apply_kplus<D>(ac, l0[n], dp[n], dm[n], U[n], 1.0, z);
#pragma unroll
          for (int k = 0; k < M; ++k) B[k][1] += z[k];
        }
      }
      solve_pp<M, 2>(A, B);
      double x[M], y[M];
#pragma unroll
      for (int k = 0; k < M; ++k) { x[k] = B[k][0]; y[k] = B[k][1]; }
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        Acoustic<D> ac;
        acoustic_vectors<D>(st, gm1, nh[n], ac);
        apply_kplus<D>(ac, l0[n], dp[n], dm[n], x, wj, lda[n]);
        if constexpr (NDIS) {
          double du[M];
#pragma unroll
          for (int k = 0; k < M; ++k) du[k] = U[n][k] - y[k];
          apply_kplus<D>(ac, l0[n], dp[n], dm[n], du, wj, dis[n]);
        }
      }
    }
  }

  // ---- blend / limit -----------------------------------------------------
  double phi[ND][M];
  if constexpr (SC == SC_LDA) {
#pragma unroll
    for (int n = 0; n < ND; ++n)
#pragma unroll
      for (int k = 0; k < M; ++k) phi[n][k] = lda[n][k];
  } else if constexpr (SC == SC_N) {
#pragma unroll
    for (int n = 0; n < ND; ++n)
#pragma unroll
      for (int k = 0; k < M; ++k) phi[n][k] = lda[n][k] + dis[n][k];
  } else if constexpr (SC == SC_B) {
    double an = 0.0;
#pragma unroll
    for (int n = 0; n < ND; ++n) an += fabs(lda[n][0] + dis[n][0]);
    const double theta = fabs(fint[0]) / (an + 1.e-6);
#pragma unroll
    for (int n = 0; n < ND; ++n)
#pragma unroll
      for (int k = 0; k < M; ++k) phi[n][k] = lda[n][k] + theta * dis[n][k];
  } else {
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double tot = fint[k];
      const bool lim = fabs(tot) > 1.e-14;
      const double itot = lim ? 1.0 / tot : 0.0;
      double xs[ND];
      double sx = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        xs[n] = fmax((lda[n][k] + dis[n][k]) * itot, 0.0) + 1.e-12;
        sx += xs[n];
      }
#pragma unroll
      for (int n = 0; n < ND; ++n) phi[n][k] = lim ? xs[n] / sx * tot : lda[n][k] + dis[n][k];
    }
  }

  double lmax = lam[0];
#pragma unroll
  for (int n = 1; n < ND; ++n) lmax = fmax(lmax, lam[n]);
  const double wave = vol * lmax;

  // ---- scatter -------------------------------------------------------------
#pragma unroll
  for (int n = 0; n < ND; ++n) {
    const size_t b = static_cast<size_t>(nid[n]);
#pragma unroll
    for (int k = 0; k < M; ++k) scatter_add(a.acc_res + b * M + k, phi[n][k], a.atomic);
    scatter_add(a.acc_ws + b, wave, a.atomic);
  }
  if (bad) atomicOr(a.status, 1);
}

}  // namespace pdrd
