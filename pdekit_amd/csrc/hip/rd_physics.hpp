// Point-wise physics of the cell loop, device side.
//
// Restates compute_properties / flux / the upwind matrix K+ = R max(L,0) L^-1
// of src/physics/scalar/RotationAdvection2D.hpp:201-336 (constant velocity),
// src/physics/scalar/Burgers2D.hpp:154-275, src/physics/euler/
// Euler2DCons.hpp:159-536 and Euler3DCons.hpp:159-450.
//
// For the Euler systems K+ is applied in the form
//   K+ = l0+ I + (lp+ - l0+) r_p l_p^T + (lm+ - l0+) r_m l_m^T,
// which follows from R L = I and the multiplicity of the convective
// eigenvalue u.n: only the two acoustic eigenvector pairs (r_p, l_p),
// (r_m, l_m) enter. Entries of r and l are those of the reference's R and
// L columns/rows for the eigenvalues |g| (u.n + a) and |g| (u.n - a).
#pragma once

#include <hip/hip_runtime.h>

namespace pdrd {

// PH_ROTATION2D: advection with the velocity (y, -x) of the point (the
// kernels pass it in the PhysParams of each point).
enum { PH_ADVECTION2D = 0, PH_BURGERS2D = 1, PH_EULER2D = 2, PH_EULER3D = 3, PH_ROTATION2D = 4 };
enum { SC_LDA = 0, SC_N = 1, SC_B = 2, SC_PSI = 3 };

template <int PH> struct PhysTraits;
template <> struct PhysTraits<PH_ADVECTION2D> { static constexpr int M = 1, DIM = 2; };
template <> struct PhysTraits<PH_ROTATION2D> { static constexpr int M = 1, DIM = 2; };
template <> struct PhysTraits<PH_BURGERS2D> { static constexpr int M = 1, DIM = 2; };
template <> struct PhysTraits<PH_EULER2D> { static constexpr int M = 4, DIM = 2; };
template <> struct PhysTraits<PH_EULER3D> { static constexpr int M = 5, DIM = 3; };

struct PhysParams
{
  double p0, p1;  // advection velocity (ax, ay) -- for PH_ROTATION2D the point's -- or (gamma, R)
};

// Thermodynamic state at a point (Euler). T is the arithmetic type of the
// interpolated state.
template <typename T, int D>
struct EulerState
{
  T rho, inv_rho, V[D], v2, P, H, a;
  bool bad;
};

template <typename T, int D>
__device__ __forceinline__ EulerState<T, D> euler_state(const T *s, T gamma)
{
  EulerState<T, D> st;
  const T gm1 = gamma - T(1);
  st.rho = s[0];
  st.inv_rho = T(1) / st.rho;
  T v2 = T(0);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    st.V[d] = s[1 + d] * st.inv_rho;
    v2 += st.V[d] * st.V[d];
  }
  st.v2 = v2;
  const T rhoE = s[D + 1];
  st.P = gm1 * (rhoE - T(0.5) * st.rho * st.v2);
  st.bad = !(st.P > T(0));
  const T RT = st.P * st.inv_rho;
  st.H = rhoE * st.inv_rho + RT;
  st.a = sqrt(gamma * RT);
  return st;
}

// Physical flux F[d][k] at a state.
template <int PH, typename T>
__device__ __forceinline__ void point_flux(const T *s, const PhysParams &pp, T F[][PhysTraits<PH>::M], bool &bad)
{
  if constexpr (PH == PH_ADVECTION2D || PH == PH_ROTATION2D) {
    F[0][0] = T(pp.p0) * s[0];
    F[1][0] = T(pp.p1) * s[0];
  } else if constexpr (PH == PH_BURGERS2D) {
    F[0][0] = T(0.5) * s[0] * s[0];
    F[1][0] = s[0];
  } else {
    constexpr int D = PhysTraits<PH>::DIM;
    const EulerState<T, D> st = euler_state<T, D>(s, T(pp.p0));
    bad = bad || st.bad;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const T md = s[1 + d];
      F[d][0] = md;
#pragma unroll
      for (int k = 0; k < D; ++k) F[d][1 + k] = md * st.V[k] + ((k == d) ? st.P : T(0));
      F[d][D + 1] = md * st.H;
    }
  }
}

// ---------------------------------------------------------------------------
// Flux differences for the mixed-precision cell loop. The flux divergence
// sum_f dVf[q][f] F(s_f) equals sum_f dVf[q][f] (F(s_f) - F(ub)) because the
// derivatives of the flux basis sum to zero, and s_f = ub + sum_n Vsf[f][n]
// (U_n - ub) because its values sum to one. With ub one of the element's
// nodal states, the differences ds = s_f - ub and dF = F(ub + ds) - F(ub)
// are O(h |grad u|); they are evaluated in T (float) from closed forms free
// of cancellation, so their relative error is that of T, not of T against
// |F|.
// ---------------------------------------------------------------------------
template <int PH, typename T>
struct DeltaBase
{
  static constexpr int D = PhysTraits<PH>::DIM;
  T u;                      // scalar state (Burgers)
  T rho, v[D], m[D], E, p;  // Euler reference state
};

template <int PH, typename T>
__device__ __forceinline__ void delta_base(const double *ub, const PhysParams &pp, DeltaBase<PH, T> &b)
{
  if constexpr (PhysTraits<PH>::M == 1) {
    b.u = T(ub[0]);
  } else {
    constexpr int D = PhysTraits<PH>::DIM;
    const double ir = 1.0 / ub[0];
    double v2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double vd = ub[1 + d] * ir;
      b.v[d] = T(vd);
      b.m[d] = T(ub[1 + d]);
      v2 += vd * ub[1 + d];
    }
    b.rho = T(ub[0]);
    b.E = T(ub[D + 1]);
    b.p = T((pp.p0 - 1.0) * (ub[D + 1] - 0.5 * v2));
  }
}

template <int PH, typename T>
__device__ __forceinline__ void flux_delta(const DeltaBase<PH, T> &b, const T *ds, const PhysParams &pp,
                                           T dF[][PhysTraits<PH>::M], bool &bad)
{
  if constexpr (PH == PH_ADVECTION2D || PH == PH_ROTATION2D) {
    dF[0][0] = T(pp.p0) * ds[0];
    dF[1][0] = T(pp.p1) * ds[0];
  } else if constexpr (PH == PH_BURGERS2D) {
    dF[0][0] = ds[0] * (b.u + T(0.5) * ds[0]);
    dF[1][0] = ds[0];
  } else {
    constexpr int D = PhysTraits<PH>::DIM;
    constexpr int M = D + 2;
    const T gm1 = T(pp.p0 - 1.0);
    const T ir1 = T(1) / (b.rho + ds[0]);
    T dv[D], v1[D];
    T v1sq = T(0), sdv = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      dv[d] = (ds[1 + d] - b.v[d] * ds[0]) * ir1;
      v1[d] = b.v[d] + dv[d];
      v1sq += v1[d] * v1[d];
      sdv += (v1[d] + b.v[d]) * dv[d];
    }
    // d(rho |v|^2 / 2) = (drho |v'|^2 + rho (v' + v) . dv) / 2
    const T dp = gm1 * (ds[M - 1] - T(0.5) * (ds[0] * v1sq + b.rho * sdv));
    bad = bad || !(b.p + dp > T(0));
    const T Ep1 = b.E + b.p + ds[M - 1] + dp;  // E' + p'
#pragma unroll
    for (int d = 0; d < D; ++d) {
      dF[d][0] = ds[1 + d];
#pragma unroll
      for (int k = 0; k < D; ++k) dF[d][1 + k] = ds[1 + d] * v1[k] + b.m[d] * dv[k] + ((k == d) ? dp : T(0));
      dF[d][M - 1] = dv[d] * Ep1 + b.v[d] * (ds[M - 1] + dp);
    }
  }
}

}  // namespace pdrd
