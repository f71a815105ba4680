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
// SC_BPSI: the PSI / LDA blend, (1 - theta) LDA + theta PSI with the cell
// theta of PGB
enum { SC_LDA = 0, SC_N = 1, SC_B = 2, SC_PSI = 3, SC_BPSI = 4 };

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

// 1/x for |x| and 1/|x| in the normal range: the hardware reciprocal
// estimate (relative error below 2^-22) refined by two Newton steps
// y <- y + y (1 - x y), each of which squares the relative error, so the
// result is within an ulp or two of the IEEE quotient at a third of the
// instructions of the IEEE division sequence. x = 0 gives NaN.
__device__ __forceinline__ double recip(double x)
{
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// sqrt(x) and 1/sqrt(x) for x > 0 in the normal range, from one hardware
// reciprocal square root estimate y: Goldschmidt's coupled iteration on
// g = x y -> sqrt(x) and h = y / 2 -> 1 / (2 sqrt(x)) (r = 1/2 - g h,
// g <- g + g r, h <- h + h r; two steps take the relative error from 2^-22
// below 2^-60), then one correction of g from its residual x - g^2. Replaces
// a square root and a division. x = 0 gives NaN for both.
__device__ __forceinline__ void sqrt_rsqrt(double x, double &s, double &rs)
{
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  s = fma(fma(-g, g, x), h, g);
  rs = 2.0 * h;
}

// Thermodynamic state at a point (Euler); inv_a = 1 / a.
template <typename T, int D>
struct EulerState
{
  T rho, inv_rho, V[D], v2, P, H, a, inv_a;
  bool bad;
};

template <typename T, int D>
__device__ __forceinline__ EulerState<T, D> euler_state(const T *s, T gamma)
{
  EulerState<T, D> st;
  const T gm1 = gamma - T(1);
  st.rho = s[0];
  st.inv_rho = recip(st.rho);
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
  sqrt_rsqrt(gamma * RT, st.a, st.inv_a);
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

}  // namespace pdrd
