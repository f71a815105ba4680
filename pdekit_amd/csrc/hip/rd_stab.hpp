// Providers of the cell loop's nodal inputs, on the device:
//   * RDBlendingCoeff "LF_Blend" (src/solver/rdm/blending_coeff/
//     LFBlendingCoeffImpl.hpp:79-136): per cell the largest
//     |u_v - u_avg| / (|u_v| + |u_avg| + 1e-12) over its nodes (first
//     component), the nodal maximum of that over the node's cells, and the
//     cell value 1 - the largest nodal value of the cell;
//   * ArtificialViscosity::compute_artificial_viscosity (src/solver/art_visc/
//     ArtificialViscosity.hpp:371-540): S = ||u_h - P u_h||^2 / ||u_h||^2 with
//     P the modal expansion without its leading terms -- on simplices the L2
//     projection onto the polynomials of degree < p, on quadrilaterals the
//     tensor Legendre expansion without the modes the reference flags (mode
//     ids 0 .. 2p, see quad_zeroed_modes) --, s = log10 S,
//     eps0 = |K|^(1/dim) / p, s0 = 1 / p^4, and nu = 0, factor eps0, or
//     factor eps0 (1 + sin(pi (s - s0) / (2 kappa))) / 2 below, above or
//     inside s0 -+ kappa;
// both taken to the nodes as MeshFunctionTools::interp_from_cells_to_nodes_
// p1_max does on a conforming mesh: a vertex takes the largest value of its
// cells, any other node the P1 / Q1 interpolation of its cell's vertex values.
// One thread per cell. Maxima of non-negative doubles go through 64-bit
// unsigned atomic max on their bit patterns (the order of non-negative IEEE
// doubles), so the results do not depend on the order of the cells.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pdrd {

constexpr int STAB_MAX_ND = 10;
constexpr int STAB_MAX_NQ = 16;
constexpr int STAB_MAX_NPM = 10;

struct StabArgs
{
  const uint32_t *__restrict__ sconn;  // [nd][ne] solution nodes (structure of arrays)
  const uint32_t *__restrict__ gconn;  // [ng][ne] vertices
  const double *__restrict__ coords;   // [nv_mesh][dim]
  const double *__restrict__ u;        // [N][m]
  const double *__restrict__ p1w;      // [nd][nvert] P1 / Q1 functions of the vertices at the nodes
  const double *__restrict__ in;       // lf_cell: nodal maxima of the cell indicator
  double *__restrict__ out;            // [N] maxima / interpolated values
  // smoothness indicator: rule w[nq], nodal basis V[nq][nd], geometry
  // derivatives dVg[dim][nq][ng], orthonormal lower-degree basis psi[nq][npm]
  // (simplices) or the zeroed modal terms hm[nq][nd] (quadrilaterals)
  const double *__restrict__ w, *__restrict__ V, *__restrict__ dVg, *__restrict__ psi, *__restrict__ hm;
  uint32_t ne;
  int m, nd, nvert, ng, dim, nq, npm, p;
  double kappa, factor;
};

__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v)
{
  atomicMax(reinterpret_cast<unsigned long long *>(addr), static_cast<unsigned long long>(__double_as_longlong(v)));
}

__global__ void stab_fill_kernel(double *__restrict__ p, uint32_t n, double v)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// LF_Blend, first pass: the cell indicator, maximised into every node.
__global__ void lf_indicator_kernel(StabArgs a)
{
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.ne) return;
  double un[STAB_MAX_ND];
  uint32_t nid[STAB_MAX_ND];
  double avg = 0.0;
  for (int n = 0; n < a.nd; ++n) {
    nid[n] = a.sconn[static_cast<size_t>(n) * a.ne + e];
    un[n] = a.u[static_cast<size_t>(nid[n]) * a.m];
    avg += un[n];
  }
  avg /= a.nd;
  double ksi = 0.0;
  for (int n = 0; n < a.nd; ++n) ksi = fmax(ksi, fabs(un[n] - avg) / (fabs(un[n]) + fabs(avg) + 1.e-12));
  for (int n = 0; n < a.nd; ++n) atomic_max_nonneg(a.out + nid[n], ksi);
}

// LF_Blend, second pass: the cell value 1 - max over the cell's nodes,
// maximised into the cell's vertices.
__global__ void lf_cell_kernel(StabArgs a)
{
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.ne) return;
  double t = 0.0;
  for (int n = 0; n < a.nd; ++n) t = fmax(t, a.in[a.sconn[static_cast<size_t>(n) * a.ne + e]]);
  const double fc = 1. - t;
  for (int v = 0; v < a.nvert; ++v) atomic_max_nonneg(a.out + a.sconn[static_cast<size_t>(v) * a.ne + e], fc);
}

// Smoothness-indicator viscosity of each cell, maximised into its vertices.
__global__ void av_cell_kernel(StabArgs a)
{
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.ne) return;
  const int D = a.dim;
  double X[4][3];
  for (int g = 0; g < a.ng; ++g) {
    const size_t v = a.gconn[static_cast<size_t>(g) * a.ne + e];
    for (int d = 0; d < D; ++d) X[g][d] = a.coords[v * D + d];
  }
  double un[STAB_MAX_ND];
  for (int n = 0; n < a.nd; ++n) un[n] = a.u[static_cast<size_t>(a.sconn[static_cast<size_t>(n) * a.ne + e]) * a.m];
  double uq[STAB_MAX_NQ], wj[STAB_MAX_NQ], coef[STAB_MAX_NPM];
  double vol = 0.0;
  for (int q = 0; q < a.nq; ++q) {
    double J[3][3];
    for (int xi = 0; xi < D; ++xi)
      for (int x = 0; x < D; ++x) {
        double s = 0.0;
        for (int g = 0; g < a.ng; ++g) s += a.dVg[(static_cast<size_t>(xi) * a.nq + q) * a.ng + g] * X[g][x];
        J[xi][x] = s;
      }
    double det;
    if (D == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    } else {
      det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    }
    wj[q] = a.w[q] * det;
    vol += wj[q];
    double s = 0.0;
    for (int n = 0; n < a.nd; ++n) s += a.V[q * a.nd + n] * un[n];
    uq[q] = s;
  }
  double nom = 0.0, denom = 0.0;
  if (a.hm) {
    // quadrilaterals: u - u_{p-1} from the nodal values through the table of
    // the zeroed modal terms
    for (int q = 0; q < a.nq; ++q) {
      double d = 0.0;
      for (int n = 0; n < a.nd; ++n) d += a.hm[q * a.nd + n] * un[n];
      nom += wj[q] * d * d;
      denom += wj[q] * uq[q] * uq[q];
    }
  } else {
    for (int k = 0; k < a.npm; ++k) {
      double s = 0.0;
      for (int q = 0; q < a.nq; ++q) s += a.w[q] * a.psi[q * a.npm + k] * uq[q];
      coef[k] = s;
    }
    for (int q = 0; q < a.nq; ++q) {
      double pr = 0.0;
      for (int k = 0; k < a.npm; ++k) pr += coef[k] * a.psi[q * a.npm + k];
      nom += wj[q] * (uq[q] - pr) * (uq[q] - pr);
      denom += wj[q] * uq[q] * uq[q];
    }
  }
  const double Se = (denom <= 1.e-10) ? 0.0 : nom / denom;
  const double se = log(Se) / log(10.);
  const double h = pow(vol, 1. / D);
  const double eps0 = h / a.p;
  const double s0 = 1. / (static_cast<double>(a.p) * a.p * a.p * a.p);
  double nu;
  if (se < s0 - a.kappa)
    nu = 0.0;
  else if (se > s0 + a.kappa)
    nu = a.factor * eps0;
  else
    nu = a.factor * 0.5 * eps0 * (1. + sin(M_PI * (se - s0) / (2. * a.kappa)));
  for (int v = 0; v < a.nvert; ++v) atomic_max_nonneg(a.out + a.sconn[static_cast<size_t>(v) * a.ne + e], nu);
}

// The nodes other than vertices: P1 / Q1 interpolation of the cell's vertex
// values (every cell holding a node writes the same value).
__global__ void p1_interp_kernel(StabArgs a)
{
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.ne) return;
  double fv[4];
  for (int v = 0; v < a.nvert; ++v) fv[v] = a.out[a.sconn[static_cast<size_t>(v) * a.ne + e]];
  for (int n = a.nvert; n < a.nd; ++n) {
    double s = 0.0;
    for (int v = 0; v < a.nvert; ++v) s += a.p1w[n * a.nvert + v] * fv[v];
    a.out[a.sconn[static_cast<size_t>(n) * a.ne + e]] = s;
  }
}

}  // namespace pdrd
