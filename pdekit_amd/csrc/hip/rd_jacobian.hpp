// Element residuals and Jacobians of the implicit forms of the schemes
// (compute_adv_res_and_jacobian):
//   LDA (PGLDA.hpp:315-505): Phi_m = sum_q w K+_m S+^-1 r_q (+ the
//     shock-capturing term), J_mn = sum_q w K+_m S+^-1 K_n;
//   N (PGN.hpp:294-400): Phi_m = sum_q sum_{n != m} w K+_m S-^-1 K-_n (u_m - u_n),
//     J_mn = -w K+_m S-^-1 K-_n, J_mm = w (K+_m - K+_m S-^-1 K-_m);
//   B (PGB.hpp:503-800): the B residual lda + theta_m n_m, J = J_LDA +
//     theta_m J_stab, J_stab,mn = -w K+_m S+^-1 K+_n, J_stab,mm =
//     w (K+_m - K+_m S+^-1 K+_m);
// with S+- = sum_n K+-_n, K_n = A(u_q) . grad phi_n, theta the nodal input or
// the cell-wise theta of equation 0. One thread per cell, FP64. K+- of the
// Euler equations come from the compact form l0 x + alpha r0 + beta e_n (see
// cell_residual) applied to the unit vectors; K = K+ + K-. The flux
// divergence at the quadrature points is the P+1 flux element's, as in
// cell_residual. Element blocks are written row-major (row = node * m +
// equation) to the caller's cell slot perm[e]; the Jacobian is accumulated
// in place, the stabilisation part of B in a scratch block of the same size.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "rd_cell.hpp"

namespace pdrd {

struct JacArgs
{
  const double *__restrict__ tab;
  const uint32_t *__restrict__ gconn;  // [NG][ne]
  const uint32_t *__restrict__ sconn;  // [ND][ne]
  const uint32_t *__restrict__ perm;   // [ne] caller cell index of each block position
  const double *__restrict__ coords;
  const double *__restrict__ u;
  const double *__restrict__ nodal;    // [N] blending coefficient (B) / artificial viscosity (LDA), or null
  double *__restrict__ phi;            // [ne][ND][M]
  double *__restrict__ jac;            // [ne][ND M][ND M]
  double *__restrict__ jstab;          // [ne][ND M][ND M] scratch (B)
  double *__restrict__ wave;           // [ne]
  int *__restrict__ status;
  uint32_t ne;
  PhysParams pp;
  // FP32 output of the mixed-precision form instead of jac / jstab: the
  // blocks entry-major, jlin[(i * ND M + j) * ne + lin_pos[e]] (jstab_lin: the
  // B stabilisation scratch in the same layout); phi / wave may then be null
  float *__restrict__ jlin;
  float *__restrict__ jstab_lin;
  const uint32_t *__restrict__ lin_pos;  // [ne] patch position of each block position
};

// One cell's Jacobian block: row-major FP64 at d, or entry-major FP32 at f
// with stride `stride`.
struct JacBlock
{
  double *d;
  float *f;
  size_t stride;
  __device__ __forceinline__ void zero(int i) const
  {
    if (d) d[i] = 0.0;
    else f[i * stride] = 0.0f;
  }
  __device__ __forceinline__ void add(int i, double v) const
  {
    if (d) d[i] += v;
    else f[i * stride] += static_cast<float>(v);
  }
  __device__ __forceinline__ double get(int i) const { return d ? d[i] : static_cast<double>(f[i * stride]); }
  __device__ __forceinline__ void set(int i, double v) const
  {
    if (d) d[i] = v;
    else f[i * stride] = static_cast<float>(v);
  }
};

// Gauss-Jordan inverse with partial pivoting of a small matrix.
template <int M>
__device__ __forceinline__ void small_inverse(double (&A)[M][M], double (&X)[M][M])
{
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) X[i][j] = (i == j) ? 1.0 : 0.0;
  for (int c = 0; c < M; ++c) {
    int piv = c;
    double best = fabs(A[c][c]);
    for (int r = c + 1; r < M; ++r)
      if (fabs(A[r][c]) > best) {
        best = fabs(A[r][c]);
        piv = r;
      }
    if (piv != c)
      for (int k = 0; k < M; ++k) {
        const double t = A[c][k];
        A[c][k] = A[piv][k];
        A[piv][k] = t;
        const double s = X[c][k];
        X[c][k] = X[piv][k];
        X[piv][k] = s;
      }
    const double d = 1.0 / A[c][c];
    for (int k = 0; k < M; ++k) {
      A[c][k] *= d;
      X[c][k] *= d;
    }
    for (int r = 0; r < M; ++r) {
      if (r == c) continue;
      const double f = A[r][c];
      for (int k = 0; k < M; ++k) {
        A[r][k] -= f * A[c][k];
        X[r][k] -= f * X[c][k];
      }
    }
  }
}

template <int D, int NG, int ND, int NF, int NQ, bool AFFINE, int PH, int SC>
__global__ __launch_bounds__(64) void rd_jacobian_kernel(JacArgs a)
{
  constexpr int M = PhysTraits<PH>::M;
  constexpr int NDM = ND * M;
  using O = TabOffsets<D, NG, ND, NF, NQ>;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.ne) return;
  const double *__restrict__ tab = a.tab;
  bool bad = false;
  double X[NG][D];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const size_t v = a.gconn[static_cast<size_t>(g) * a.ne + e];
#pragma unroll
    for (int d = 0; d < D; ++d) X[g][d] = a.coords[v * D + d];
  }
  double U[ND][M], nodal[ND];
#pragma unroll
  for (int n = 0; n < ND; ++n) {
    const size_t v = a.sconn[static_cast<size_t>(n) * a.ne + e];
#pragma unroll
    for (int k = 0; k < M; ++k) U[n][k] = a.u[v * M + k];
    nodal[n] = a.nodal ? a.nodal[v] : 0.0;
  }
  const uint32_t c = a.perm[e];
  const bool lin = a.jlin != nullptr;
  const uint32_t pos = lin ? a.lin_pos[e] : 0u;
  const JacBlock jac{lin ? nullptr : a.jac + static_cast<size_t>(c) * NDM * NDM, lin ? a.jlin + pos : nullptr, a.ne};
  const JacBlock jst{lin ? nullptr : (a.jstab ? a.jstab + static_cast<size_t>(c) * NDM * NDM : nullptr),
                     lin ? (a.jstab_lin ? a.jstab_lin + pos : nullptr) : nullptr, a.ne};
  // Scalar cells of up to 6 nodes accumulate their Jacobian (and the B
  // stabilisation block) in registers and store it once; the others
  // accumulate in place in the caller's block.
  constexpr bool REG = (M == 1) && (ND <= 6);
  double Jr[REG ? NDM * NDM : 1], Jsr[(REG && SC == SC_B) ? NDM * NDM : 1];
  if constexpr (REG) {
#pragma unroll
    for (int i = 0; i < NDM * NDM; ++i) Jr[i] = 0.0;
    if constexpr (SC == SC_B)
#pragma unroll
      for (int i = 0; i < NDM * NDM; ++i) Jsr[i] = 0.0;
  } else {
    for (int i = 0; i < NDM * NDM; ++i) {
      jac.zero(i);
      if (SC == SC_B) jst.zero(i);
    }
  }
  const bool ext_theta = (SC == SC_B) && a.nodal;
  const bool visc = (SC == SC_LDA) && a.nodal;
  double tmax = 0.0;
#pragma unroll
  for (int n = 0; n < ND; ++n) tmax = fmax(tmax, nodal[n]);
  const bool skip = ext_theta && tmax < 1.e-6;

  // geometry at every quadrature point
  constexpr int NJ = AFFINE ? 1 : NQ;
  double Jinv[NJ][D][D], detJ[NJ];
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    double J[D][D];
#pragma unroll
    for (int xi = 0; xi < D; ++xi)
#pragma unroll
      for (int x = 0; x < D; ++x) {
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) s += tab[O::dVgT + (q * D + xi) * NG + g] * X[g][x];
        J[xi][x] = s;
      }
    detJ[q] = det_inv<D>(J, Jinv[q]);
  }

  // flux divergence of the P+1 flux element at the quadrature points
  double divr[NQ][M];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int k = 0; k < M; ++k) divr[q][k] = 0.0;
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
    PhysParams ppf = a.pp;
    if constexpr (PH == PH_ROTATION2D) {
      double xf[2] = {0.0, 0.0};
#pragma unroll
      for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int d = 0; d < 2; ++d) xf[d] += tab[O::Vgf + f * NG + g] * X[g][d];
      ppf.p0 = xf[1];
      ppf.p1 = -xf[0];
    }
    point_flux<PH, double>(s, ppf, F, bad);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int jq = AFFINE ? 0 : q;
#pragma unroll
      for (int k = 0; k < M; ++k) {
        double t = divr[q][k];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          double b = 0.0;
#pragma unroll
          for (int xi = 0; xi < D; ++xi) b += Jinv[jq][d][xi] * tab[O::dVfT + (f * D + xi) * NQ + q];
          t += b * F[d][k];
        }
        divr[q][k] = t;
      }
    }
  }

  double lda[ND][M], ndis[ND][M];
#pragma unroll
  for (int n = 0; n < ND; ++n)
#pragma unroll
    for (int k = 0; k < M; ++k) lda[n][k] = ndis[n][k] = 0.0;
  double fint[M];
#pragma unroll
  for (int k = 0; k < M; ++k) fint[k] = 0.0;
  double vol = 0.0, lmax = 0.0;

  double Kp[ND][M][M], Kx[ND][M][M];  // K+_n and (K_n for LDA / B, K-_n for N)
  for (int q = 0; q < NQ; ++q) {
    const int jq = AFFINE ? 0 : q;
    const double wj = tab[O::w + q] * detJ[jq];
    vol += wj;
    double uq[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      double t = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) t += tab[O::V + q * ND + n] * U[n][k];
      uq[k] = t;
    }
    // physical gradients of the basis at q
    double G[ND][D];
#pragma unroll
    for (int n = 0; n < ND; ++n)
#pragma unroll
      for (int x = 0; x < D; ++x) {
        double t = 0.0;
#pragma unroll
        for (int xi = 0; xi < D; ++xi) t += Jinv[jq][x][xi] * tab[O::dVT + (q * D + xi) * ND + n];
        G[n][x] = t;
      }
    if (visc) {
      double nuq = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) nuq += tab[O::V + q * ND + n] * nodal[n];
      if (nuq != 0.0) {
        double gu[D][M];
#pragma unroll
        for (int x = 0; x < D; ++x)
#pragma unroll
          for (int k = 0; k < M; ++k) {
            double t = 0.0;
#pragma unroll
            for (int n = 0; n < ND; ++n) t += G[n][x] * U[n][k];
            gu[x][k] = t;
          }
#pragma unroll
        for (int n = 0; n < ND; ++n)
#pragma unroll
          for (int x = 0; x < D; ++x) {
            const double cx = wj * nuq * G[n][x];
#pragma unroll
            for (int k = 0; k < M; ++k) lda[n][k] += cx * gu[x][k];
          }
      }
    }
    // K+ and K / K- of every node
    if constexpr (M == 1) {
      double ax = (PH == PH_BURGERS2D) ? uq[0] : a.pp.p0;
      double ay = (PH == PH_BURGERS2D) ? 1.0 : a.pp.p1;
      if constexpr (PH == PH_ROTATION2D) {
        double xq[2] = {0.0, 0.0};
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
          for (int d = 0; d < 2; ++d) xq[d] += tab[O::Vg + q * NG + g] * X[g][d];
        ax = xq[1];
        ay = -xq[0];
      }
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        const double k = ax * G[n][0] + ay * G[n][1];
        lmax = fmax(lmax, fabs(k));
        Kp[n][0][0] = fmax(0.0, k);
        Kx[n][0][0] = (SC == SC_N) ? fmin(0.0, k) : k;
      }
    } else {
      const EulerState<double, D> st = euler_state<double, D>(uq, a.pp.p0);
      bad = bad || st.bad;
      const double gm1 = a.pp.p0 - 1.0;
      double l0v[M], r0[M];
      l0v[0] = 0.5 * gm1 * st.v2;
      r0[0] = 1.0;
#pragma unroll
      for (int x = 0; x < D; ++x) {
        l0v[1 + x] = -gm1 * st.V[x];
        r0[1 + x] = st.V[x];
      }
      l0v[M - 1] = gm1;
      r0[M - 1] = st.H;
      const double i2a = 0.5 / st.a, i2a2 = i2a / st.a;
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        double nn = 0.0;
#pragma unroll
        for (int x = 0; x < D; ++x) nn += G[n][x] * G[n][x];
        const double norm = sqrt(nn);
        const bool tiny = (D == 2) ? (norm < 1.e-9) : (norm == 0.0);
        const double inorm = tiny ? 0.0 : 1.0 / norm;
        double nh[D], um = 0.0;
#pragma unroll
        for (int x = 0; x < D; ++x) {
          nh[x] = G[n][x] * inorm;
          um += st.V[x] * nh[x];
        }
        const double e0 = norm * um, ep = norm * (um + st.a), em = norm * (um - st.a);
        lmax = fmax(lmax, fmax(fabs(e0), fmax(fabs(ep), fabs(em))));
        // sign +1: K+, -1: K-
        for (int sgn = 0; sgn < 2; ++sgn) {
          const double p0 = sgn == 0 ? fmax(0.0, e0) : fmin(0.0, e0);
          const double dpn = tiny ? 0.0 : (sgn == 0 ? fmax(0.0, ep) : fmin(0.0, ep)) - p0;
          const double dmn = tiny ? 0.0 : (sgn == 0 ? fmax(0.0, em) : fmin(0.0, em)) - p0;
          const double l0 = tiny ? 0.0 : p0;
          const double sg = dpn + dmn, dl = dpn - dmn;
#pragma unroll
          for (int j = 0; j < M; ++j) {
            // column j: the operator applied to the unit vector e_j
            const double P = l0v[j];
            double Q = (j == 0) ? -um : ((j <= D) ? nh[j - 1] : 0.0);
            const double al = (sg * P + st.a * dl * Q) * i2a2;
            const double be = 0.5 * sg * Q + dl * P * i2a;
            double col[M];
            col[0] = ((j == 0) ? l0 : 0.0) + al;
#pragma unroll
            for (int x = 0; x < D; ++x) col[1 + x] = ((j == 1 + x) ? l0 : 0.0) + al * st.V[x] + be * nh[x];
            col[M - 1] = ((j == M - 1) ? l0 : 0.0) + al * st.H + be * um;
#pragma unroll
            for (int i = 0; i < M; ++i) {
              if (sgn == 0) {
                Kp[n][i][j] = col[i];
              } else {
                Kx[n][i][j] = (SC == SC_N) ? col[i] : Kp[n][i][j] + col[i];
              }
            }
          }
        }
      }
    }
    // S = sum K+ (LDA, B) or sum K- (N), and its inverse
    double S[M][M], Si[M][M];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
      for (int j = 0; j < M; ++j) {
        double t = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) t += (SC == SC_N) ? Kx[n][i][j] : Kp[n][i][j];
        S[i][j] = t;
      }
    small_inverse<M>(S, Si);
#pragma unroll
    for (int k = 0; k < M; ++k) fint[k] += wj * divr[q][k];
    if constexpr (REG) {
      // the scalar form of the block loops below, fully unrolled
      const double si = Si[0][0];
#pragma unroll
      for (int mm = 0; mm < ND; ++mm) {
        const double bm = Kp[mm][0][0] * si;
        if (SC != SC_N) lda[mm][0] += wj * (bm * divr[q][0]);
#pragma unroll
        for (int nn = 0; nn < ND; ++nn) {
          const double p = bm * Kx[nn][0][0];
          if (SC != SC_N) {
            Jr[mm * NDM + nn] += wj * p;
          } else if (nn != mm) {
            ndis[mm][0] += wj * (p * (U[mm][0] - U[nn][0]));
            Jr[mm * NDM + nn] += -wj * p;
          } else {
            Jr[mm * NDM + nn] += wj * (Kp[mm][0][0] - p);
          }
        }
        if (SC == SC_B && !skip)
#pragma unroll
          for (int nn = 0; nn < ND; ++nn) {
            const double p = bm * Kp[nn][0][0];
            if (nn != mm) {
              ndis[mm][0] += wj * (p * (U[mm][0] - U[nn][0]));
              Jsr[mm * NDM + nn] += -wj * p;
            } else {
              Jsr[mm * NDM + nn] += wj * (Kp[mm][0][0] - p);
            }
          }
      }
      continue;
    }
    for (int mm = 0; mm < ND; ++mm) {
      double B[M][M];  // K+_m S^-1
#pragma unroll
      for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) {
          double t = 0.0;
#pragma unroll
          for (int k = 0; k < M; ++k) t += Kp[mm][i][k] * Si[k][j];
          B[i][j] = t;
        }
      if (SC != SC_N) {
#pragma unroll
        for (int i = 0; i < M; ++i) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < M; ++j) t += B[i][j] * divr[q][j];
          lda[mm][i] += wj * t;
        }
      }
      for (int nn = 0; nn < ND; ++nn) {
        double P[M][M];
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
          for (int j = 0; j < M; ++j) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < M; ++k) t += B[i][k] * Kx[nn][k][j];
            P[i][j] = t;
          }
        const int row = (mm * M) * NDM + nn * M;
        if (SC != SC_N) {
#pragma unroll
          for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) jac.add(row + i * NDM + j, wj * P[i][j]);
        } else if (nn != mm) {
#pragma unroll
          for (int i = 0; i < M; ++i) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < M; ++j) t += P[i][j] * (U[mm][j] - U[nn][j]);
            ndis[mm][i] += wj * t;
#pragma unroll
            for (int j = 0; j < M; ++j) jac.add(row + i * NDM + j, -wj * P[i][j]);
          }
        } else {
#pragma unroll
          for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) jac.add(row + i * NDM + j, wj * (Kp[mm][i][j] - P[i][j]));
        }
      }
      if (SC == SC_B && !skip) {
        for (int nn = 0; nn < ND; ++nn) {
          double P[M][M];
#pragma unroll
          for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < M; ++j) {
              double t = 0.0;
#pragma unroll
              for (int k = 0; k < M; ++k) t += B[i][k] * Kp[nn][k][j];
              P[i][j] = t;
            }
          const int row = (mm * M) * NDM + nn * M;
          if (nn != mm) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
              double t = 0.0;
#pragma unroll
              for (int j = 0; j < M; ++j) t += P[i][j] * (U[mm][j] - U[nn][j]);
              ndis[mm][i] += wj * t;
#pragma unroll
              for (int j = 0; j < M; ++j) jst.add(row + i * NDM + j, -wj * P[i][j]);
            }
          } else {
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
              for (int j = 0; j < M; ++j) jst.add(row + i * NDM + j, wj * (Kp[mm][i][j] - P[i][j]));
          }
        }
      }
    }
  }

  // residuals, the blend of B, wave speed
  double *phi = a.phi ? a.phi + static_cast<size_t>(c) * NDM : nullptr;
  if (SC == SC_N) {
    if (phi)
#pragma unroll
      for (int n = 0; n < ND; ++n)
#pragma unroll
        for (int k = 0; k < M; ++k) phi[n * M + k] = ndis[n][k];
  } else if (SC == SC_B) {
    double th[ND];
    if (ext_theta) {
#pragma unroll
      for (int n = 0; n < ND; ++n) th[n] = nodal[n];
    } else {
      double an = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) an += fabs(lda[n][0] + ndis[n][0]);
      const double t0 = fabs(fint[0]) / (an + 1.e-6);
#pragma unroll
      for (int n = 0; n < ND; ++n) th[n] = t0;
    }
    if (phi)
#pragma unroll
      for (int n = 0; n < ND; ++n)
#pragma unroll
        for (int k = 0; k < M; ++k) phi[n * M + k] = lda[n][k] + th[n] * ndis[n][k];
    if constexpr (REG) {
      if (!skip)
#pragma unroll
        for (int i = 0; i < NDM; ++i)
#pragma unroll
          for (int j = 0; j < NDM; ++j) Jr[i * NDM + j] += th[i] * Jsr[i * NDM + j];
    } else if (!skip) {
      for (int i = 0; i < NDM; ++i) {
        const double t = th[i / M];
        for (int j = 0; j < NDM; ++j) jac.add(i * NDM + j, t * jst.get(i * NDM + j));
      }
    }
  } else {
    if (phi)
#pragma unroll
      for (int n = 0; n < ND; ++n)
#pragma unroll
        for (int k = 0; k < M; ++k) phi[n * M + k] = lda[n][k];
  }
  if constexpr (REG)
#pragma unroll
    for (int i = 0; i < NDM * NDM; ++i) jac.set(i, Jr[i]);
  if (a.wave) a.wave[c] = vol * fmax(lmax, 1.e-6);
  if (bad) atomicOr(a.status, 1);
}

// ---------------------------------------------------------------------------
// Euler cells: one wave per block row of the element Jacobian. A workgroup
// holds 64 cells and ND waves; wave m computes, for its 64 cells, the block
// row of node m, J_mn = sum_q w B_m X_n with B_m = K+_m S^-1 and
//   LDA: X_n = K_n;  N: -K-_n, plus K+_m on the diagonal block;
//   B: the LDA block and the stabilisation block -B_m K+_n (plus K+_m on the
//      diagonal), combined with theta_m once the residuals are summed.
// Every K is applied in its compact form X = l I + r0 alpha^T + e beta^T
// (alpha = A1 l0v + A2 f, beta = B1 f + B2 l0v; e = (0, n, V.n),
// f = (-V.n, n, 0) for the unit normal n of the node, r0 = (1, V, H),
// l0v = (gm1 |V|^2 / 2, -gm1 V, gm1)), so B_m X_n = l B_m + (B_m r0) alpha^T +
// (B_m e) beta^T costs about 3 M^2 FMAs and no K is held as a matrix. The
// per-cell work that is not per row (flux divergence, S and its inverse) is
// repeated by every wave, so the waves never wait on each other. The rows
// are accumulated in registers, NCH block columns per pass over the
// quadrature points, and stored once. The residuals use the same sums:
// lda_m = w K+_m S+^-1 r, and the N / B parts w K+_m (u_m - S^-1 sum_n X_n u_n)
// (= sum_{n != m} w B_m X_n (u_m - u_n), since sum_n X_n = S).
template <int D>
struct NodeQ
{
  double um, nh[D];
  double lp, sgp, dlp;  // K+: l0+, sigma+ = dp + dm, delta+ = dp - dm
  double lm, sgm, dlm;  // K-
  double kl, kn;        // K = K+ + K-: l = V.g, sigma = 0, delta = 2 |g| a (0 where g is tiny)
};

template <int D, int NG, int ND, int NF, int NQ, bool AFFINE, int PH, int SC>
__global__ __launch_bounds__(64 * ND) void rd_jacobian_rows_kernel(JacArgs a)
{
  constexpr int M = PhysTraits<PH>::M;
  static_assert(M == D + 2, "Euler cells only");
  constexpr int NDM = ND * M;
  constexpr int NACC = (SC == SC_B) ? 2 : 1;
  constexpr int CAP = 64 / (M * M * NACC);
  constexpr int NCH = CAP < 1 ? 1 : (CAP > ND ? ND : CAP);
  constexpr int NPASS = (ND + NCH - 1) / NCH;
  using O = TabOffsets<D, NG, ND, NF, NQ>;
  // the flux divergence of the workgroup's 64 cells, [NQ][M][64]
  extern __shared__ double divl[];
  const int m = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const uint32_t e0 = blockIdx.x * 64u + lane;
  // lanes past the end work on the last cell and store nothing, so that
  // every lane reaches the barriers below
  const bool active = e0 < a.ne;
  const uint32_t e = active ? e0 : a.ne - 1;
  const double *__restrict__ tab = a.tab;
  bool bad = false;
  double X[NG][D];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const size_t v = a.gconn[static_cast<size_t>(g) * a.ne + e];
#pragma unroll
    for (int d = 0; d < D; ++d) X[g][d] = a.coords[v * D + d];
  }
  double U[ND][M], nodal[ND];
#pragma unroll
  for (int n = 0; n < ND; ++n) {
    const size_t v = a.sconn[static_cast<size_t>(n) * a.ne + e];
#pragma unroll
    for (int k = 0; k < M; ++k) U[n][k] = a.u[v * M + k];
    nodal[n] = a.nodal ? a.nodal[v] : 0.0;
  }
  const uint32_t c = a.perm[e];
  const bool lin = a.jlin != nullptr;
  const uint32_t pos = lin ? a.lin_pos[e] : 0u;
  const JacBlock jac{lin ? nullptr : a.jac + static_cast<size_t>(c) * NDM * NDM, lin ? a.jlin + pos : nullptr, a.ne};
  const bool ext_theta = (SC == SC_B) && a.nodal;
  const bool visc = (SC == SC_LDA) && a.nodal;
  double tmax = 0.0;
#pragma unroll
  for (int n = 0; n < ND; ++n) tmax = fmax(tmax, nodal[n]);
  const bool skip = ext_theta && tmax < 1.e-6;
  double Um[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double t = U[0][k];
#pragma unroll
    for (int n = 1; n < ND; ++n) t = (n == m) ? U[n][k] : t;
    Um[k] = t;
  }
  // geometry at quadrature point q
  auto geometry = [&](const int q, double (&Ji)[D][D]) {
    double J[D][D];
#pragma unroll
    for (int xi = 0; xi < D; ++xi)
#pragma unroll
      for (int x = 0; x < D; ++x) {
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) s += tab[O::dVgT + (q * D + xi) * NG + g] * X[g][x];
        J[xi][x] = s;
      }
    return det_inv<D>(J, Ji);
  };
  double Ja[D][D];
  const double deta = AFFINE ? geometry(0, Ja) : 0.0;

  {
    // flux divergence of the P+1 flux element at the quadrature points: wave
    // m takes the flux nodes m, m + ND, ...; the partial sums are added into
    // LDS in wave order (deterministic)
    double divr[NQ][M];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int k = 0; k < M; ++k) divr[q][k] = 0.0;
#pragma unroll 1
    for (int f = m; f < NF; f += ND) {
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
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        double Jq[D][D];
        if constexpr (AFFINE) {
#pragma unroll
          for (int x = 0; x < D; ++x)
#pragma unroll
            for (int y = 0; y < D; ++y) Jq[x][y] = Ja[x][y];
        } else {
          geometry(q, Jq);
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
          double t = divr[q][k];
#pragma unroll
          for (int d = 0; d < D; ++d) {
            double b = 0.0;
#pragma unroll
            for (int xi = 0; xi < D; ++xi) b += Jq[d][xi] * tab[O::dVfT + (f * D + xi) * NQ + q];
            t += b * F[d][k];
          }
          divr[q][k] = t;
        }
      }
    }
#pragma unroll 1
    for (int w = 0; w < ND; ++w) {
      if (w == m) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
          for (int k = 0; k < M; ++k) {
            double *p = divl + (q * M + k) * 64 + lane;
            *p = (w == 0) ? divr[q][k] : *p + divr[q][k];
          }
      }
      __syncthreads();
    }
  }

  const double gm1 = a.pp.p0 - 1.0;
  double lda_m[M], dis_m[M], t0[ND], fint0 = 0.0, vol = 0.0, lmax = 0.0;
#pragma unroll
  for (int k = 0; k < M; ++k) lda_m[k] = dis_m[k] = 0.0;
#pragma unroll
  for (int n = 0; n < ND; ++n) t0[n] = 0.0;
  double theta = 0.0;

#pragma unroll 1
  for (int pass = 0; pass < NPASS; ++pass) {
    const int n0 = pass * NCH;
    double acc[NACC][NCH][M][M];
#pragma unroll
    for (int t = 0; t < NACC; ++t)
#pragma unroll
      for (int j = 0; j < NCH; ++j)
#pragma unroll
        for (int r = 0; r < M; ++r)
#pragma unroll
          for (int cc = 0; cc < M; ++cc) acc[t][j][r][cc] = 0.0;

#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      double Jinv[D][D];
      double detq;
      if constexpr (AFFINE) {
#pragma unroll
        for (int x = 0; x < D; ++x)
#pragma unroll
          for (int y = 0; y < D; ++y) Jinv[x][y] = Ja[x][y];
        detq = deta;
      } else {
        detq = geometry(q, Jinv);
      }
      const double wj = tab[O::w + q] * detq;
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
      const double av = st.a;
      const double i2a = 0.5 * st.inv_a, i2a2 = i2a * st.inv_a;
      double r0[M], l0v[M];
      r0[0] = 1.0;
      l0v[0] = 0.5 * gm1 * st.v2;
#pragma unroll
      for (int x = 0; x < D; ++x) {
        r0[1 + x] = st.V[x];
        l0v[1 + x] = -gm1 * st.V[x];
      }
      r0[M - 1] = st.H;
      l0v[M - 1] = gm1;
      // node parameters at q
      NodeQ<D> nq[ND];
      double G[ND][D];
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        double nn = 0.0;
#pragma unroll
        for (int x = 0; x < D; ++x) {
          double t = 0.0;
#pragma unroll
          for (int xi = 0; xi < D; ++xi) t += Jinv[x][xi] * tab[O::dVT + (q * D + xi) * ND + n];
          G[n][x] = t;
          nn += t * t;
        }
        double norm, inorm;
        sqrt_rsqrt(nn, norm, inorm);
        norm = (nn > 0.0) ? norm : 0.0;
        const bool tiny = (D == 2) ? (norm < 1.e-9) : (norm == 0.0);
        inorm = tiny ? 0.0 : inorm;
        double um = 0.0;
#pragma unroll
        for (int x = 0; x < D; ++x) {
          nq[n].nh[x] = G[n][x] * inorm;
          um += st.V[x] * nq[n].nh[x];
        }
        nq[n].um = um;
        const double e0 = norm * um, ep = norm * (um + av), em = norm * (um - av);
        lmax = fmax(lmax, fmax(fabs(e0), fmax(fabs(ep), fabs(em))));
        const double pp0 = fmax(0.0, e0), mp0 = fmin(0.0, e0);
        const double dpp = tiny ? 0.0 : fmax(0.0, ep) - pp0, dmp = tiny ? 0.0 : fmax(0.0, em) - pp0;
        const double dpm = tiny ? 0.0 : fmin(0.0, ep) - mp0, dmm = tiny ? 0.0 : fmin(0.0, em) - mp0;
        nq[n].lp = tiny ? 0.0 : pp0;
        nq[n].sgp = dpp + dmp;
        nq[n].dlp = dpp - dmp;
        nq[n].lm = tiny ? 0.0 : mp0;
        nq[n].sgm = dpm + dmm;
        nq[n].dlm = dpm - dmm;
        nq[n].kl = tiny ? 0.0 : e0;
        nq[n].kn = tiny ? 0.0 : norm;
      }
      // compact coefficients (l, A1, A2, B1, B2) of K+ (sel 0), K- (1), K (2)
      auto coef = [&](const NodeQ<D> &k, int sel, double (&cf)[5]) {
        if (sel == 2) {
          // K = (V.g) I + |g| (r0 f^T + e l0v^T)
          cf[0] = k.kl;
          cf[1] = 0.0;
          cf[2] = k.kn;
          cf[3] = 0.0;
          cf[4] = k.kn;
          return;
        }
        const double l = sel == 0 ? k.lp : k.lm;
        const double sg = sel == 0 ? k.sgp : k.sgm;
        const double dl = sel == 0 ? k.dlp : k.dlm;
        cf[0] = l;
        cf[1] = sg * i2a2;
        cf[2] = av * dl * i2a2;
        cf[3] = 0.5 * sg;
        cf[4] = dl * i2a;
      };
      auto evec = [&](const NodeQ<D> &k, double (&ev)[M], double (&fv)[M]) {
        ev[0] = 0.0;
        fv[0] = -k.um;
#pragma unroll
        for (int x = 0; x < D; ++x) ev[1 + x] = fv[1 + x] = k.nh[x];
        ev[M - 1] = k.um;
        fv[M - 1] = 0.0;
      };
      // X y
      auto apply = [&](const NodeQ<D> &k, int sel, const double (&y)[M], double (&out)[M]) {
        double cf[5], ev[M], fv[M];
        coef(k, sel, cf);
        evec(k, ev, fv);
        double P = 0.0, Q = 0.0;
#pragma unroll
        for (int j = 0; j < M; ++j) {
          P += l0v[j] * y[j];
          Q += fv[j] * y[j];
        }
        const double al = cf[1] * P + cf[2] * Q, be = cf[3] * Q + cf[4] * P;
#pragma unroll
        for (int i = 0; i < M; ++i) out[i] = cf[0] * y[i] + al * r0[i] + be * ev[i];
      };
      const int ssel = (SC == SC_N) ? 1 : 0;
      // S = sum_n X_n and its inverse
      double S[M][M], Si[M][M];
#pragma unroll
      for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) {
          S[i][j] = 0.0;
          Si[i][j] = (i == j) ? 1.0 : 0.0;
        }
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        double cf[5], ev[M], fv[M];
        coef(nq[n], ssel, cf);
        evec(nq[n], ev, fv);
#pragma unroll
        for (int j = 0; j < M; ++j) {
          const double al = cf[1] * l0v[j] + cf[2] * fv[j];
          const double be = cf[3] * fv[j] + cf[4] * l0v[j];
#pragma unroll
          for (int i = 0; i < M; ++i) S[i][j] += ((i == j) ? cf[0] : 0.0) + r0[i] * al + ev[i] * be;
        }
      }
      solve_pp<M, M>(S, Si);
      // node m's parameters
      NodeQ<D> km = nq[0];
#pragma unroll
      for (int n = 1; n < ND; ++n)
        if (n == m) km = nq[n];
      // B_m = K+_m Si = l Si + r0 (alpha^T Si) + e (beta^T Si)
      double Bm[M][M];
      {
        double cf[5], ev[M], fv[M];
        coef(km, 0, cf);
        evec(km, ev, fv);
        double pa[M], pb[M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
          double sl = 0.0, sf = 0.0;
#pragma unroll
          for (int k = 0; k < M; ++k) {
            sl += l0v[k] * Si[k][j];
            sf += fv[k] * Si[k][j];
          }
          pa[j] = cf[1] * sl + cf[2] * sf;
          pb[j] = cf[3] * sf + cf[4] * sl;
        }
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
          for (int j = 0; j < M; ++j) Bm[i][j] = cf[0] * Si[i][j] + r0[i] * pa[j] + ev[i] * pb[j];
      }
      double cB[M];  // B_m r0
#pragma unroll
      for (int i = 0; i < M; ++i) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) t += Bm[i][k] * r0[k];
        cB[i] = t;
      }
      // acc[t][j] += w B_m X_n for the pass's block columns
#pragma unroll
      for (int j = 0; j < NCH; ++j) {
        const int n = n0 + j;
        if (n < ND) {
          NodeQ<D> kn = nq[0];
#pragma unroll
          for (int n2 = 1; n2 < ND; ++n2)
            if (n2 == n) kn = nq[n2];
          double ev[M], fv[M];
          evec(kn, ev, fv);
          double dB[M];  // B_m e_n
#pragma unroll
          for (int i = 0; i < M; ++i) {
            double t = 0.0;
#pragma unroll
            for (int k = 1; k < M; ++k) t += Bm[i][k] * ev[k];
            dB[i] = t;
          }
#pragma unroll
          for (int t = 0; t < NACC; ++t) {
            // LDA: K_n; N: K-_n (negated); B: K_n (t = 0), K+_n (t = 1, negated)
            const int sel = (SC == SC_N) ? 1 : (t == 0 ? 2 : 0);
            const double sgn = (SC == SC_N || t == 1) ? -wj : wj;
            double cf[5];
            coef(kn, sel, cf);
#pragma unroll
            for (int i = 0; i < M; ++i) {
              const double ga = sgn * (cf[1] * cB[i] + cf[4] * dB[i]);
              const double de = sgn * (cf[2] * cB[i] + cf[3] * dB[i]);
              const double li = sgn * cf[0];
#pragma unroll
              for (int cc = 0; cc < M; ++cc) acc[t][j][i][cc] += li * Bm[i][cc] + ga * l0v[cc] + de * fv[cc];
            }
          }
          if ((SC == SC_N || SC == SC_B) && n == m) {
            // + w K+_m on the diagonal block (N: the Jacobian; B: its stabilisation part)
            double cf[5], evm[M], fvm[M];
            coef(km, 0, cf);
            evec(km, evm, fvm);
#pragma unroll
            for (int cc = 0; cc < M; ++cc) {
              const double al = cf[1] * l0v[cc] + cf[2] * fvm[cc];
              const double be = cf[3] * fvm[cc] + cf[4] * l0v[cc];
#pragma unroll
              for (int i = 0; i < M; ++i)
                acc[NACC - 1][j][i][cc] += wj * (((i == cc) ? cf[0] : 0.0) + r0[i] * al + evm[i] * be);
            }
          }
        }
      }
      if (pass == 0) {
        // residuals and wave speed
        double r[M];
#pragma unroll
        for (int k = 0; k < M; ++k) r[k] = divl[(q * M + k) * 64 + lane];
        vol += wj;
        fint0 += wj * r[0];
        if constexpr (SC != SC_N) {
          double x[M], y[M];
#pragma unroll
          for (int i = 0; i < M; ++i) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < M; ++k) t += Si[i][k] * r[k];
            x[i] = t;
          }
          apply(km, 0, x, y);
#pragma unroll
          for (int k = 0; k < M; ++k) lda_m[k] += wj * y[k];
          if constexpr (SC == SC_B) {
#pragma unroll
            for (int n = 0; n < ND; ++n) {
              apply(nq[n], 0, x, y);
              t0[n] += wj * y[0];
            }
          }
        }
        if constexpr (SC == SC_N || SC == SC_B) {
          // ubar = Si sum_n X_n u_n
          double su[M], ub[M], y[M];
#pragma unroll
          for (int k = 0; k < M; ++k) su[k] = 0.0;
#pragma unroll
          for (int n = 0; n < ND; ++n) {
            apply(nq[n], ssel, U[n], y);
#pragma unroll
            for (int k = 0; k < M; ++k) su[k] += y[k];
          }
#pragma unroll
          for (int i = 0; i < M; ++i) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < M; ++k) t += Si[i][k] * su[k];
            ub[i] = t;
          }
          double du[M];
#pragma unroll
          for (int k = 0; k < M; ++k) du[k] = Um[k] - ub[k];
          apply(km, 0, du, y);
#pragma unroll
          for (int k = 0; k < M; ++k) dis_m[k] += wj * y[k];
          if constexpr (SC == SC_B) {
#pragma unroll
            for (int n = 0; n < ND; ++n) {
#pragma unroll
              for (int k = 0; k < M; ++k) du[k] = U[n][k] - ub[k];
              apply(nq[n], 0, du, y);
              t0[n] += wj * y[0];
            }
          }
        }
        if (visc) {
          double nuq = 0.0;
#pragma unroll
          for (int n = 0; n < ND; ++n) nuq += tab[O::V + q * ND + n] * nodal[n];
          if (nuq != 0.0) {
            double Gm[D];
#pragma unroll
            for (int x = 0; x < D; ++x) {
              double t = G[0][x];
#pragma unroll
              for (int n = 1; n < ND; ++n) t = (n == m) ? G[n][x] : t;
              Gm[x] = t;
            }
#pragma unroll
            for (int k = 0; k < M; ++k) {
              double t = 0.0;
#pragma unroll
              for (int n = 0; n < ND; ++n) {
                double gg = 0.0;
#pragma unroll
                for (int x = 0; x < D; ++x) gg += Gm[x] * G[n][x];
                t += gg * U[n][k];
              }
              lda_m[k] += wj * nuq * t;
            }
          }
        }
      }
    }

    if (pass == 0) {
      // theta of B, the residual row of node m, the wave speed
      if constexpr (SC == SC_B) {
        if (ext_theta) {
          double t = nodal[0];
#pragma unroll
          for (int n = 1; n < ND; ++n) t = (n == m) ? nodal[n] : t;
          theta = t;
        } else {
          double an = 0.0;
#pragma unroll
          for (int n = 0; n < ND; ++n) an += fabs(t0[n]);
          theta = fabs(fint0) / (an + 1.e-6);
        }
      }
      if (active && a.phi) {
        double *phi = a.phi + static_cast<size_t>(c) * NDM + m * M;
#pragma unroll
        for (int k = 0; k < M; ++k)
          phi[k] = (SC == SC_N) ? dis_m[k] : ((SC == SC_B && !skip) ? lda_m[k] + theta * dis_m[k] : lda_m[k]);
      }
      if (active && m == 0 && a.wave) a.wave[c] = vol * fmax(lmax, 1.e-6);
    }
    // store the pass's block columns of row block m
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
      const int n = n0 + j;
      if (active && n < ND) {
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
          for (int cc = 0; cc < M; ++cc) {
            double v = acc[0][j][i][cc];
            if constexpr (SC == SC_B) v += skip ? 0.0 : theta * acc[NACC - 1][j][i][cc];
            jac.set((m * M + i) * NDM + n * M + cc, v);
          }
      }
    }
  }
  if (active && bad) atomicOr(a.status, 1);
}

}  // namespace pdrd
