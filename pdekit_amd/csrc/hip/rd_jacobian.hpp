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
// the cell-wise theta of equation 0, all in FP64. Scalar cells: one thread
// per cell (rd_jacobian_kernel; cells of up to 6 nodes accumulate in
// registers, larger ones in place in the caller's block, the stabilisation
// part of B in a scratch block of the same size). Euler cells: one lane per
// (cell, quadrature point) with K+- in their compact form
// (rd_jacobian_qp_kernel below). The flux divergence at the quadrature
// points is the P+1 flux element's, as in cell_residual. Element blocks are
// written row-major (row = node * m + equation) to the caller's cell slot
// perm[e], or entry-major in FP32 for the mixed-precision form.
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
// Euler cells, one lane per (cell, quadrature point): LPC = NQ lanes per
// cell, 64 / NQ cells per wave (the 64 mod NQ lanes left over idle), so the
// per-point work -- the K+- parameters of the nodes, S and its inverse -- is
// done once. Lane q forms its point's share of the residuals
// and of every block J_mn = w B_m X_n (B_m = K+_m S^-1; X_n = K_n for LDA,
// -K-_n plus K+_m on the diagonal for N, K_n - theta_m K+_n plus theta_m K+_m
// on the diagonal for B), all K applied in the compact form
// l I + r0 alpha^T + e beta^T. The sums over the points run through a
// wave-private LDS slab: every lane writes its values, then each lane adds
// the group's values of the entries it owns (a fixed order: deterministic).
// The lanes of a cell are in one wave, so no workgroup barrier is needed, but
// the wave waits for its outstanding LDS accesses (s_waitcnt lgkmcnt(0))
// between the writes and the reads of an exchange and before the slab is
// written again: without those waits back-to-back exchanges read values of
// the next one (measured: residuals 20 % off). The flux divergence is split
// over the lanes by flux node and summed the same way.

// K+- parameters of one node at a quadrature point (compact form)
template <int D>
struct NodeQ
{
  double um, nh[D];
  double lp, sgp, dlp;  // K+: l0+, sigma+ = dp + dm, delta+ = dp - dm
  double lm, sgm, dlm;  // K-
  double kl, kn;        // K = K+ + K-: l = V.g, sigma = 0, delta = 2 |g| a (0 where g is tiny)
};

template <int NQ>
struct QpLanes
{
  static_assert(NQ <= 64, "one cell's points must fit in a wave");
  static constexpr int value = NQ;
};

// Stride of the exchange slab rows (doubles): 64 lanes plus one, so that the
// lanes of a group, reading one column of consecutive rows, fall in different
// banks.
constexpr int QP_XS = 65;

// Waves per SIMD of the per-point kernel and rows of a wave's exchange slab.
// The 2D Euler P1 / Q1 cells run at 2 waves (a few spilled registers; measured
// 1.2x faster than 1 wave without spills, the exchanges' LDS waits of one
// wave hidden behind the other's arithmetic), with 36 slab rows so that 8
// waves fit in a CU's LDS; the others at 1 wave with 64 rows. The Jacobian
// blocks of a row block go through the slab several at a time (as many as
// fit), so that the wave waits for the LDS once per batch rather than once
// per block.
template <int M, int ND>
struct QpWaves
{
  static constexpr int value = (M == 4 && ND <= 4) ? 2 : 1;
};

template <int M, int ND>
struct QpRows
{
  static constexpr int value = QpWaves<M, ND>::value > 1 ? 36 : 64;
};

// The two halves of an exchange. Value j of a lane goes to row base + j; the
// lane of index q in its group (gl: the group's first lane in the wave) owns
// the values r LPC + q, r < R. With that order the 64 lanes of a read fall
// on distinct banks (row stride QP_XS = 1 mod 16 doubles).
template <int LPC, int R>
__device__ __forceinline__ void slab_put(double *xw, int lane, int base, const double (&v)[LPC * R])
{
#pragma unroll
  for (int j = 0; j < LPC * R; ++j) xw[(base + j) * QP_XS + lane] = v[j];
}

// out[r] = sum over the lanes l of the group of v_l[r LPC + q].
template <int LPC, int R>
__device__ __forceinline__ void slab_sum(const double *xw, int gl, int q, int base, double (&out)[R])
{
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < LPC; ++l) s += xw[(base + r * LPC + q) * QP_XS + gl + l];
    out[r] = s;
  }
}

// Waits for the wave's outstanding LDS accesses; the memory clobber keeps the
// compiler from moving LDS accesses across the wait.
__device__ __forceinline__ void lds_wait() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// One exchange: the slab's previous reads (by other lanes) complete before it
// is overwritten, and the writes before it is read.
template <int LPC, int R>
__device__ __forceinline__ void group_scatter(double *xw, int lane, int gl, int q, const double (&v)[LPC * R],
                                              double (&out)[R])
{
  lds_wait();
  slab_put<LPC, R>(xw, lane, 0, v);
  lds_wait();
  slab_sum<LPC, R>(xw, gl, q, 0, out);
}

template <int D, int NG, int ND, int NF, int NQ, bool AFFINE, int PH, int SC>
__global__ __launch_bounds__(64, (QpWaves<PhysTraits<PH>::M, ND>::value)) void rd_jacobian_qp_kernel(JacArgs a)
{
  constexpr int M = PhysTraits<PH>::M;
  constexpr int QPR = QpRows<M, ND>::value;
  static_assert(M == D + 2, "Euler cells only");
  constexpr int NDM = ND * M;
  constexpr int LPC = QpLanes<NQ>::value;
  constexpr int CPB = 64 / LPC;
  constexpr int RJ = (M * M + LPC - 1) / LPC;
  constexpr int NT = ND + 2;  // theta inputs: t0 of every node, fint0, vol
  constexpr int RT = (NT + LPC - 1) / LPC;
  constexpr int RP = (M + LPC - 1) / LPC;
  constexpr int JB = QPR / (LPC * RJ);  // Jacobian blocks per exchange
  static_assert(JB >= 1 && LPC * RT <= QPR && LPC * RP <= QPR && LPC <= QPR, "exchange slab too small");
  using O = TabOffsets<D, NG, ND, NF, NQ>;
  extern __shared__ double xl[];
  const int lane = threadIdx.x & 63;
  double *xw = xl;  // one wave per workgroup
  // the wave's last 64 mod LPC lanes (spare) and the lanes of cells past the
  // end work on the last cell and store nothing; a spare lane reads the
  // exchanges of the first group (in bounds, results unused)
  const bool spare = lane >= CPB * LPC;
  const int q = spare ? 0 : lane % LPC;
  const int gl = spare ? 0 : lane - q;
  const uint32_t e0 = spare ? a.ne : blockIdx.x * CPB + lane / LPC;
  const bool active = e0 < a.ne;
  const uint32_t e = active ? e0 : a.ne - 1;
  const bool qa = q < NQ;
  const int qq = qa ? q : NQ - 1;
  bool bad = false;
  // the cell's indices first, then the values they address (two rounds of
  // memory latency; loads interleaved with their uses, or behind a branch per
  // node, were issued one after another)
  uint32_t gv[NG], sv[ND];
#pragma unroll
  for (int g = 0; g < NG; ++g) gv[g] = a.gconn[static_cast<size_t>(g) * a.ne + e];
#pragma unroll
  for (int n = 0; n < ND; ++n) sv[n] = a.sconn[static_cast<size_t>(n) * a.ne + e];
  const uint32_t c = a.perm[e];
  const bool lin = a.jlin != nullptr;
  const uint32_t pos = lin ? a.lin_pos[e] : 0u;
  // the table values of the lane's point (vector loads, one round with the
  // cell's data rather than one per phase)
  const double *__restrict__ tab = a.tab;
  double tV[ND], tdV[D][ND], tdVg[AFFINE ? 1 : D][NG];
  const double tw = tab[O::w + qq];
#pragma unroll
  for (int n = 0; n < ND; ++n) {
    tV[n] = tab[O::V + qq * ND + n];
#pragma unroll
    for (int xi = 0; xi < D; ++xi) tdV[xi][n] = tab[O::dVT + (qq * D + xi) * ND + n];
  }
  if constexpr (!AFFINE)
#pragma unroll
    for (int xi = 0; xi < D; ++xi)
#pragma unroll
      for (int g = 0; g < NG; ++g) tdVg[xi][g] = tab[O::dVgT + (qq * D + xi) * NG + g];
  double X[NG][D];
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int d = 0; d < D; ++d) X[g][d] = a.coords[static_cast<size_t>(gv[g]) * D + d];
  double U[ND][M], nodal[ND];
#pragma unroll
  for (int n = 0; n < ND; ++n)
#pragma unroll
    for (int k = 0; k < M; ++k) U[n][k] = a.u[static_cast<size_t>(sv[n]) * M + k];
#pragma unroll
  for (int n = 0; n < ND; ++n) nodal[n] = 0.0;
  if (a.nodal)
#pragma unroll
    for (int n = 0; n < ND; ++n) nodal[n] = a.nodal[sv[n]];
  const JacBlock jac{lin ? nullptr : a.jac + static_cast<size_t>(c) * NDM * NDM, lin ? a.jlin + pos : nullptr, a.ne};
  const bool ext_theta = (SC == SC_B) && a.nodal;
  const bool visc = (SC == SC_LDA) && a.nodal;
  double tmax = 0.0;
#pragma unroll
  for (int n = 0; n < ND; ++n) tmax = fmax(tmax, nodal[n]);
  const bool skip = ext_theta && tmax < 1.e-6;
  auto geometry = [&](const int qp, double (&Ji)[D][D]) {
    double J[D][D];
#pragma unroll
    for (int xi = 0; xi < D; ++xi)
#pragma unroll
      for (int x = 0; x < D; ++x) {
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) s += tab[O::dVgT + (qp * D + xi) * NG + g] * X[g][x];
        J[xi][x] = s;
      }
    return det_inv<D>(J, Ji);
  };

  // the point's geometry
  double Jinv[D][D];
  double detq;
  if constexpr (AFFINE) {
    detq = geometry(0, Jinv);  // the same Jacobian at every point
  } else {
    double J[D][D];
#pragma unroll
    for (int xi = 0; xi < D; ++xi)
#pragma unroll
      for (int x = 0; x < D; ++x) {
        double t = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) t += tdVg[xi][g] * X[g][x];
        J[xi][x] = t;
      }
    detq = det_inv<D>(J, Jinv);
  }

  // flux divergence at the lane's point: the group's lanes evaluate the flux
  // nodes f = q + i LPC and exchange the fluxes F_f (D M values each, value
  // (d, k) of the lane's i-th node at slab row i D M + d M + k, in batches
  // of FB nodes); lane q then forms div F(q) = sum_f sum_d b_fd F_f[d] with
  // b_f = Jinv_q grad_ref psi_f(q) (the Jacobian of its own point only)
  double r[M];
  {
    constexpr int NFL = (NF + LPC - 1) / LPC;  // flux nodes per lane
    constexpr int FV = D * M;
    constexpr int FB = QPR / FV;           // flux-node slots per exchange
    static_assert(FB >= 1, "exchange slab too small");
#pragma unroll
    for (int k = 0; k < M; ++k) r[k] = 0.0;
#pragma unroll
    for (int i0 = 0; i0 < NFL; i0 += FB) {
      lds_wait();
#pragma unroll
      for (int i = i0; i < i0 + FB && i < NFL; ++i) {
        const int f = q + i * LPC;
        if (f < NF) {
          double sf[M];
#pragma unroll
          for (int k = 0; k < M; ++k) {
            double t = 0.0;
#pragma unroll
            for (int n = 0; n < ND; ++n) t += tab[O::Vsf + f * ND + n] * U[n][k];
            sf[k] = t;
          }
          double F[D][M];
          point_flux<PH, double>(sf, a.pp, F, bad);
#pragma unroll
          for (int d = 0; d < D; ++d)
#pragma unroll
            for (int k = 0; k < M; ++k) xw[((i - i0) * FV + d * M + k) * QP_XS + lane] = F[d][k];
        }
      }
      lds_wait();
      // rolled: unrolled, every read of the slab and the table was hoisted
      // ahead (registers beyond the file)
#pragma unroll 1
      for (int f = i0 * LPC; f < NF && f < (i0 + FB) * LPC; ++f) {
        const int i = f / LPC, src = gl + f % LPC;
        double b[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          double t = 0.0;
#pragma unroll
          for (int xi = 0; xi < D; ++xi) t += Jinv[d][xi] * tab[O::dVfT + (f * D + xi) * NQ + qq];
          b[d] = t;
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
          double t = r[k];
#pragma unroll
          for (int d = 0; d < D; ++d) t += b[d] * xw[((i - i0) * FV + d * M + k) * QP_XS + src];
          r[k] = t;
        }
      }
    }
  }

  // the point's state, node parameters, S and its inverse
  const double wj = qa ? tw * detq : 0.0;
  double uq[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double t = 0.0;
#pragma unroll
    for (int n = 0; n < ND; ++n) t += tV[n] * U[n][k];
    uq[k] = t;
  }
  const EulerState<double, D> st = euler_state<double, D>(uq, a.pp.p0);
  bad = bad || st.bad;
  const double gm1 = a.pp.p0 - 1.0;
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
  NodeQ<D> nq[ND];
  double G[ND][D];
  double lmax = 0.0;
#pragma unroll
  for (int n = 0; n < ND; ++n) {
    double nn = 0.0;
#pragma unroll
    for (int x = 0; x < D; ++x) {
      double t = 0.0;
#pragma unroll
      for (int xi = 0; xi < D; ++xi) t += Jinv[x][xi] * tdV[xi][n];
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
    const double e0p = norm * um, ep = norm * (um + av), em = norm * (um - av);
    lmax = fmax(lmax, fmax(fabs(e0p), fmax(fabs(ep), fabs(em))));
    const double pp0 = fmax(0.0, e0p), mp0 = fmin(0.0, e0p);
    const double dpp = tiny ? 0.0 : fmax(0.0, ep) - pp0, dmp = tiny ? 0.0 : fmax(0.0, em) - pp0;
    const double dpm = tiny ? 0.0 : fmin(0.0, ep) - mp0, dmm = tiny ? 0.0 : fmin(0.0, em) - mp0;
    nq[n].lp = tiny ? 0.0 : pp0;
    nq[n].sgp = dpp + dmp;
    nq[n].dlp = dpp - dmp;
    nq[n].lm = tiny ? 0.0 : mp0;
    nq[n].sgm = dpm + dmm;
    nq[n].dlm = dpm - dmm;
    nq[n].kl = tiny ? 0.0 : e0p;
    nq[n].kn = tiny ? 0.0 : norm;
  }
  // compact coefficients (l, A1, A2, B1, B2) of K+ (sel 0), K- (1), K (2)
  auto coef = [&](const NodeQ<D> &k, int sel, double (&cf)[5]) {
    if (sel == 2) {
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
  double Si[M][M];
  {
    double S[M][M];
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
  }

  // the point's residual shares: LDA part w K+_n S+^-1 r, N / B part
  // w K+_n (u_n - S^-1 sum X_n u_n)
  constexpr bool HL = SC != SC_N, HD = SC != SC_LDA;
  double lda[HL ? ND : 1][M], dis[HD ? ND : 1][M];
  if constexpr (HL) {
    double x[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) t += Si[i][k] * r[k];
      x[i] = t;
    }
#pragma unroll
    for (int n = 0; n < ND; ++n) {
      double y[M];
      apply(nq[n], 0, x, y);
#pragma unroll
      for (int k = 0; k < M; ++k) lda[n][k] = wj * y[k];
    }
    if (visc) {
      double nuq = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) nuq += tV[n] * nodal[n];
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
        for (int k = 0; k < M; ++k) {
          double t = 0.0;
#pragma unroll
          for (int x = 0; x < D; ++x) t += G[n][x] * gu[x][k];
          lda[n][k] += wj * nuq * t;
        }
    }
  }
  if constexpr (HD) {
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
#pragma unroll
    for (int n = 0; n < ND; ++n) {
      double du[M];
#pragma unroll
      for (int k = 0; k < M; ++k) du[k] = U[n][k] - ub[k];
      apply(nq[n], 0, du, y);
#pragma unroll
      for (int k = 0; k < M; ++k) dis[n][k] = wj * y[k];
    }
  }

  // sums over the points: theta inputs (t0 of every node, fint0, vol) to
  // every lane, the largest eigenvalue, then theta
  double tot[NT];
  {
    double v[LPC * RT], o[RT];
#pragma unroll
    for (int t = 0; t < LPC * RT; ++t) v[t] = 0.0;
    if constexpr (SC == SC_B)
#pragma unroll
      for (int n = 0; n < ND; ++n) v[n] = lda[n][0] + dis[n][0];
    v[ND] = wj * r[0];
    v[ND + 1] = wj;
    group_scatter<LPC, RT>(xw, lane, gl, q, v, o);
#pragma unroll
    for (int rr = 0; rr < RT; ++rr) xw[(rr * LPC + q) * QP_XS + gl] = o[rr];
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#pragma unroll
    for (int t = 0; t < NT; ++t) tot[t] = xw[t * QP_XS + gl];
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    xw[lane] = lmax;
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    double lm = 0.0;
#pragma unroll
    for (int l = 0; l < LPC; ++l) lm = fmax(lm, xw[gl + l]);
    lmax = lm;
  }
  double th[ND];
#pragma unroll
  for (int n = 0; n < ND; ++n) th[n] = 0.0;
  if constexpr (SC == SC_B) {
    if (ext_theta) {
#pragma unroll
      for (int n = 0; n < ND; ++n) th[n] = skip ? 0.0 : nodal[n];
    } else {
      double an = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) an += fabs(tot[n]);
      const double t0 = fabs(tot[ND]) / (an + 1.e-6);
#pragma unroll
      for (int n = 0; n < ND; ++n) th[n] = t0;
    }
  }

  // residuals: lane q stores the entries q RP ... of every node
  if (a.phi) {
    double *phi = a.phi + static_cast<size_t>(c) * NDM;
    constexpr int PB = QPR / (LPC * RP);  // nodes per exchange
#pragma unroll
    for (int n0 = 0; n0 < ND; n0 += PB) {
      lds_wait();
#pragma unroll
      for (int n = n0; n < n0 + PB && n < ND; ++n) {
        double v[LPC * RP];
#pragma unroll
        for (int t = 0; t < LPC * RP; ++t) {
          double val = 0.0;
          if (t < M) {
            if constexpr (SC == SC_LDA) val = lda[n][t];
            else if constexpr (SC == SC_N) val = dis[n][t];
            else val = lda[n][t] + th[n] * dis[n][t];
          }
          v[t] = val;
        }
        slab_put<LPC, RP>(xw, lane, (n - n0) * LPC * RP, v);
      }
      lds_wait();
#pragma unroll
      for (int n = n0; n < n0 + PB && n < ND; ++n) {
        double o[RP];
        slab_sum<LPC, RP>(xw, gl, q, (n - n0) * LPC * RP, o);
#pragma unroll
        for (int rr = 0; rr < RP; ++rr) {
          const int k = rr * LPC + q;
          if (active && k < M) phi[n * M + k] = o[rr];
        }
      }
    }
  }
  if (active && q == 0 && a.wave) a.wave[c] = tot[ND + 1] * fmax(lmax, 1.e-6);

  // Jacobian blocks, row block by row block
  // unrolled for the cells of up to 4 nodes (their parameters then stay in
  // place instead of being selected for every row block; fewer spills, except
  // for B, whose row blocks also carry theta_m K+_m)
#pragma unroll ((ND <= 4 && SC != SC_B) ? ND : 1)
  for (int m = 0; m < ND; ++m) {
    NodeQ<D> km = nq[0];
    double thm = th[0];
#pragma unroll
    for (int n = 1; n < ND; ++n)
      if (n == m) {
        km = nq[n];
        thm = th[n];
      }
    double cfm[5], evm[M], fvm[M];
    coef(km, 0, cfm);
    evec(km, evm, fvm);
    // B_m = K+_m Si = l Si + r0 (alpha^T Si) + e (beta^T Si)
    double Bm[M][M];
    {
      double pa[M], pb[M];
#pragma unroll
      for (int j = 0; j < M; ++j) {
        double sl = 0.0, sf = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) {
          sl += l0v[k] * Si[k][j];
          sf += fvm[k] * Si[k][j];
        }
        pa[j] = cfm[1] * sl + cfm[2] * sf;
        pb[j] = cfm[3] * sf + cfm[4] * sl;
      }
      // with the point's weight: w B_m
#pragma unroll
      for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) Bm[i][j] = wj * (cfm[0] * Si[i][j] + r0[i] * pa[j] + evm[i] * pb[j]);
    }
    double cB[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) t += Bm[i][k] * r0[k];
      cB[i] = t;
    }
    // JB blocks per exchange: the previous batch's reads complete before the
    // slab is overwritten, then the batch's writes before its reads
#pragma unroll
    for (int n0 = 0; n0 < ND; n0 += JB) {
    lds_wait();
#pragma unroll
    for (int n = n0; n < n0 + JB && n < ND; ++n) {
      double ev[M], fv[M], cf[5];
      evec(nq[n], ev, fv);
      if constexpr (SC == SC_LDA) {
        coef(nq[n], 2, cf);
      } else if constexpr (SC == SC_N) {
        coef(nq[n], 1, cf);
#pragma unroll
        for (int t = 0; t < 5; ++t) cf[t] = -cf[t];
      } else {
        double cp[5];
        coef(nq[n], 2, cf);
        coef(nq[n], 0, cp);
#pragma unroll
        for (int t = 0; t < 5; ++t) cf[t] -= thm * cp[t];
      }
      double dB[M];
#pragma unroll
      for (int i = 0; i < M; ++i) {
        double t = 0.0;
#pragma unroll
        for (int k = 1; k < M; ++k) t += Bm[i][k] * ev[k];
        dB[i] = t;
      }
      // the diagonal block's K+_m term: N weight 1, B weight theta_m
      const double dw = (n == m) ? ((SC == SC_N) ? 1.0 : ((SC == SC_B) ? thm : 0.0)) : 0.0;
      double v[LPC * RJ];
#pragma unroll
      for (int t = M * M; t < LPC * RJ; ++t) v[t] = 0.0;
#pragma unroll
      for (int i = 0; i < M; ++i) {
        const double ga = cf[1] * cB[i] + cf[4] * dB[i];
        const double de = cf[2] * cB[i] + cf[3] * dB[i];
#pragma unroll
        for (int j = 0; j < M; ++j) {
          double val = cf[0] * Bm[i][j] + ga * l0v[j] + de * fv[j];
          if constexpr (SC != SC_LDA) {
            const double al = cfm[1] * l0v[j] + cfm[2] * fvm[j];
            const double be = cfm[3] * fvm[j] + cfm[4] * l0v[j];
            val += wj * dw * (((i == j) ? cfm[0] : 0.0) + r0[i] * al + evm[i] * be);
          }
          v[i * M + j] = val;
        }
      }
      slab_put<LPC, RJ>(xw, lane, (n - n0) * LPC * RJ, v);
    }
    lds_wait();
#pragma unroll
    for (int n = n0; n < n0 + JB && n < ND; ++n) {
      double o[RJ];
      slab_sum<LPC, RJ>(xw, gl, q, (n - n0) * LPC * RJ, o);
#pragma unroll
      for (int rr = 0; rr < RJ; ++rr) {
        const int t = rr * LPC + q;
        if (active && t < M * M) jac.set((m * M + t / M) * NDM + n * M + t % M, o[rr]);
      }
    }
    }
  }
  if (active && bad) atomicOr(a.status, 1);
}

}  // namespace pdrd
