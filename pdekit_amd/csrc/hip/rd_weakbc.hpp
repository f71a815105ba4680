// Weak boundary conditions on boundary facets: WeakWall, WeakSubInlet,
// WeakSubOutlet, WeakFarfield, WeakSymmetry, WeakSuperInlet (Euler) and
// WeakDirichlet (any physics) (src/solver/rdm/bc/WeakWall.hpp,
// WeakSubInlet.hpp, WeakSubOutlet.hpp, WeakFarfield.hpp, WeakSymmetry.hpp,
// WeakSuperInlet.hpp, WeakDirichlet.hpp over WeakBC::apply,
// WeakBC.hpp:225-352, 786-821).
//
// One thread per facet. Per facet quadrature point q (state u_q = V u):
//   WeakWall:       dF = F(u_g) n - F(u_q) n, u_g = u_q without its normal
//                   velocity (same density and pressure);
//   WeakSuperInlet: dF = 0 (only the wave speed is added);
//   the others:     dF = K-(n) (u_g - u_q) with the ghost state u_g
//     WeakSubInlet:  from the total temperature and pressure at the local
//                    Mach number (gas constant 287.046),
//     WeakSubOutlet: u_q with the energy of the given pressure,
//     WeakFarfield, WeakDirichlet: the given state,
//     WeakSymmetry:  u_q with its normal velocity reflected;
// with K- = R min(Lambda, 0) L, for the Euler equations in the compact form
// l0 x + alpha r0 + beta e (see cell_residual), for a scalar min(a.n, 0).
// The facet residual sum_q w_q |J| V(q, n) dF (times the WeakWall relaxation
// 1 / (3 p) for p > 1) is rescaled so that its nodal parts sum to the facet
// total sum_q w_q |J| dF; each node receives it in its partial-sum row
// together with V_dual * max_q max |eigenvalue| (from 1e-9 at every point;
// WeakFarfield keeps one running maximum from 1e-8), and the nodal step then
// divides by the dual volume like the cells' residuals.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "rd_physics.hpp"

namespace pdrd {

enum {
  BC_WEAK_WALL = 0,
  BC_WEAK_SUBINLET = 1,
  BC_WEAK_SUBOUTLET = 2,
  BC_WEAK_FARFIELD = 3,
  BC_WEAK_SYMMETRY = 4,
  BC_WEAK_SUPERINLET = 5,
  BC_WEAK_DIRICHLET = 6
};

struct WeakArgs
{
  const uint32_t *__restrict__ nodes;   // [nf][NFN] solution nodes
  const uint32_t *__restrict__ rows;    // [nf][NFN] row of each (facet, node) in slot_acc
  const double *__restrict__ normal;    // [nf][D] outward unit normals
  const double *__restrict__ jac;       // [nf] facet Jacobians
  const double *__restrict__ w;         // [nq] facet rule
  const double *__restrict__ V;         // [nq][NFN]
  const double *__restrict__ u;         // [N][M] state at the start of the sweep
  const double *__restrict__ vol;       // [N] dual volumes
  double *__restrict__ slot_acc;        // [rows][M + 1]
  int *__restrict__ status;
  uint32_t nf;
  int nq, type;
  double relax;                         // WeakWall relaxation (1 for P1)
  double params[5];                     // SubInlet: T_tot, p_tot; SubOutlet: p; Farfield, Dirichlet: state
  PhysParams pp;                        // gamma (Euler) or the advection velocity
  const double *__restrict__ qstate;    // [nf][nq][M] WeakDirichlet state at every facet point, or null
  const double *__restrict__ qvel;      // [nf][nq][2] rotating advection: the velocity (y, -x) of every facet point
};

// K-(n) x for the Euler equations at state st (unit normal, un = V.n).
template <int D>
__device__ __forceinline__ void euler_kminus(const EulerState<double, D> &st, const double *nrm, double un,
                                             double gm1, const double *x, double *out)
{
  constexpr int M = D + 2;
  const double e0 = un, ep = un + st.a, em = un - st.a;
  const double m0 = fmin(0.0, e0);
  const double dpn = fmin(0.0, ep) - m0, dmn = fmin(0.0, em) - m0;
  const double sgn = dpn + dmn, dln = dpn - dmn;
  const double i2a = 0.5 / st.a, i2a2 = i2a / st.a;
  double P = 0.5 * gm1 * st.v2 * x[0];
#pragma unroll
  for (int d = 0; d < D; ++d) P -= gm1 * st.V[d] * x[1 + d];
  P += gm1 * x[M - 1];
  double Q = -un * x[0];
#pragma unroll
  for (int d = 0; d < D; ++d) Q += nrm[d] * x[1 + d];
  const double al = (sgn * P + st.a * dln * Q) * i2a2;
  const double be = 0.5 * sgn * Q + dln * P * i2a;
  out[0] = m0 * x[0] + al;
#pragma unroll
  for (int d = 0; d < D; ++d) out[1 + d] = m0 * x[1 + d] + al * st.V[d] + be * nrm[d];
  out[M - 1] = m0 * x[M - 1] + al * st.H + be * un;
}

// Correction dF at one facet point and the largest |eigenvalue| there.
// gq: the boundary state of this point (WeakDirichlet with an expression of
// the point, WeakDirichletBC.hpp:216-229), or null for the constant one.
template <int PH>
__device__ __forceinline__ double weak_point(const WeakArgs &a, const double *uq, const double *nrm, double *dF,
                                             bool &bad, const double *gq, const double *qv)
{
  constexpr int D = PhysTraits<PH>::DIM, M = PhysTraits<PH>::M;
  if constexpr (M == 1) {
    // WeakDirichlet of a scalar: dF = min(a.n, 0) (g - u); the rotating
    // advection's velocity is the point's (y, -x)
    const double ax = (PH == PH_BURGERS2D) ? uq[0] : ((PH == PH_ROTATION2D) ? qv[0] : a.pp.p0);
    const double ay = (PH == PH_BURGERS2D) ? 1.0 : ((PH == PH_ROTATION2D) ? qv[1] : a.pp.p1);
    const double an = ax * nrm[0] + ay * nrm[1];
    dF[0] = fmin(an, 0.0) * ((gq ? gq[0] : a.params[0]) - uq[0]);
    return fabs(an);
  } else {
    const double gamma = a.pp.p0, gm1 = gamma - 1.0;
    const EulerState<double, D> st = euler_state<double, D>(uq, gamma);
    bad = bad || st.bad;
    double un = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) un += st.V[d] * nrm[d];
    const double emax = fmax(fabs(un), fmax(fabs(un + st.a), fabs(un - st.a)));
    if (a.type == BC_WEAK_SUPERINLET) {
#pragma unroll
      for (int k = 0; k < M; ++k) dF[k] = 0.0;
      return emax;
    }
    double g[M];
    if (a.type == BC_WEAK_WALL) {
      double v2t = 0.0;
      g[0] = st.rho;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double vt = st.V[d] - un * nrm[d];
        v2t += vt * vt;
        g[1 + d] = vt * st.rho;
      }
      g[M - 1] = uq[M - 1] + 0.5 * st.rho * (v2t - st.v2);
      const EulerState<double, D> sg = euler_state<double, D>(g, gamma);
      double gn = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) gn += sg.V[d] * nrm[d];
      // F(s) n = (rho u_n, rho v u_n + p n, rho H u_n)
      dF[0] = g[0] * gn - uq[0] * un;
#pragma unroll
      for (int d = 0; d < D; ++d) dF[1 + d] = (g[1 + d] * gn + sg.P * nrm[d]) - (uq[1 + d] * un + st.P * nrm[d]);
      dF[M - 1] = g[0] * sg.H * gn - uq[0] * st.H * un;
      return emax;
    }
    if (a.type == BC_WEAK_SUBINLET) {
      const double R_gas = 287.046;
      const double Ma2 = st.v2 / (st.a * st.a);
      const double fac = 1.0 + 0.5 * gm1 * Ma2;
      const double Tg = a.params[0] / fac;
      const double pg = a.params[1] / pow(fac, gamma / gm1);
      g[0] = pg / (R_gas * Tg);
      g[1] = sqrt(Ma2) * sqrt(gamma * R_gas * Tg) * g[0];
#pragma unroll
      for (int d = 1; d < D; ++d) g[1 + d] = 0.0;
      double r2u2 = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) r2u2 += g[1 + d] * g[1 + d];
      g[M - 1] = pg / gm1 + 0.5 * (r2u2 / g[0]);
    } else if (a.type == BC_WEAK_SUBOUTLET) {
#pragma unroll
      for (int k = 0; k < M - 1; ++k) g[k] = uq[k];
      double r2u2 = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) r2u2 += uq[1 + d] * uq[1 + d];
      g[M - 1] = a.params[0] / gm1 + 0.5 * r2u2 / uq[0];
    } else if (a.type == BC_WEAK_SYMMETRY) {
      g[0] = uq[0];
#pragma unroll
      for (int d = 0; d < D; ++d) g[1 + d] = uq[0] * (st.V[d] - 2.0 * un * nrm[d]);
      g[M - 1] = uq[M - 1];
    } else {
#pragma unroll
      for (int k = 0; k < M; ++k) g[k] = gq ? gq[k] : a.params[k];
    }
    double x[M];
#pragma unroll
    for (int k = 0; k < M; ++k) x[k] = g[k] - uq[k];
    euler_kminus<D>(st, nrm, un, gm1, x, dF);
    return emax;
  }
}

template <int PH, int NFN>
__global__ __launch_bounds__(128) void weak_bc_kernel(WeakArgs a)
{
  constexpr int D = PhysTraits<PH>::DIM, M = PhysTraits<PH>::M;
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.nf) return;
  double U[NFN][M];
  uint32_t nid[NFN];
#pragma unroll
  for (int n = 0; n < NFN; ++n) {
    nid[n] = a.nodes[static_cast<size_t>(f) * NFN + n];
#pragma unroll
    for (int k = 0; k < M; ++k) U[n][k] = a.u[static_cast<size_t>(nid[n]) * M + k];
  }
  double nrm[D];
#pragma unroll
  for (int d = 0; d < D; ++d) nrm[d] = a.normal[static_cast<size_t>(f) * D + d];
  const double jac = a.jac[f];
  double bres[NFN][M], total[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    total[k] = 0.0;
#pragma unroll
    for (int n = 0; n < NFN; ++n) bres[n][k] = 0.0;
  }
  const double floor = (a.type == BC_WEAK_FARFIELD) ? 1.e-8 : 1.e-9;
  double coeff = 0.0;
  bool bad = false;
  for (int q = 0; q < a.nq; ++q) {
    const double wq = jac * a.w[q];
    double uq[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      double t = 0.0;
#pragma unroll
      for (int n = 0; n < NFN; ++n) t += a.V[q * NFN + n] * U[n][k];
      uq[k] = t;
    }
    double dF[M];
    const double *gq = a.qstate ? a.qstate + (static_cast<size_t>(f) * a.nq + q) * M : nullptr;
    const double *qv = (PH == PH_ROTATION2D) ? a.qvel + (static_cast<size_t>(f) * a.nq + q) * 2 : nullptr;
    const double emax = weak_point<PH>(a, uq, nrm, dF, bad, gq, qv);
    coeff = fmax(coeff, fmax(floor, emax));
#pragma unroll
    for (int k = 0; k < M; ++k) {
      total[k] += wq * dF[k];
#pragma unroll
      for (int n = 0; n < NFN; ++n) bres[n][k] += a.relax * a.V[q * NFN + n] * wq * dF[k];
    }
  }
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double sum = 0.0;
#pragma unroll
    for (int n = 0; n < NFN; ++n) sum += bres[n][k];
    if (fabs(sum) > 1.e-14) {
      const double sc = total[k] / sum;
#pragma unroll
      for (int n = 0; n < NFN; ++n) bres[n][k] *= sc;
    }
  }
#pragma unroll
  for (int n = 0; n < NFN; ++n) {
    double *row = a.slot_acc + static_cast<size_t>(a.rows[static_cast<size_t>(f) * NFN + n]) * (M + 1);
#pragma unroll
    for (int k = 0; k < M; ++k) row[k] = bres[n][k];
    row[M] = a.vol[nid[n]] * coeff;
  }
  if (bad) atomicOr(a.status, 1);
}

}  // namespace pdrd
