// Fused residual-distribution cell kernel for gfx950.
//
// One thread owns one element and performs, in registers, the work of
//   gather (GeometryCache / SolutionCache push_back_to_buffer,
//           src/interpolation/GeometryCache.hpp:329-400, SolutionCache.hpp:361-388)
//   geometry metric (GeometryMetric.hpp:366-425, 3D 487-560)
//   P+1 flux metric (FluxSpaceMetric.hpp:562-743)
//   element scheme  (PGLDA.hpp:109-311, PGN.hpp:96-290, PGB.hpp:314-498,
//                    PSI limiter LFResidualLimiter.hpp:84-110, and the PSI /
//                    LDA blend by PGB's cell theta, PGB.hpp:437-477)
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
#include <type_traits>

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
  const double *__restrict__ nodal;  // [N] blending coefficient (B) or artificial viscosity (LDA), or null
  // when non-zero, vertex ids >= nv_check or node ids >= nn_check are not
  // followed: the cell is skipped and bit 2 of *status is set (connectivity
  // the host has not validated)
  uint32_t nv_check, nn_check;
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
  // Device-only copies appended by the host, laid out so that the values one
  // iteration of a rolled loop needs are contiguous (one scalar load each):
  static constexpr int dVfT = size;                  // [NF][D][NQ]
  static constexpr int dVT = dVfT + NF * D * NQ;     // [NQ][D][ND]
  static constexpr int dVgT = dVT + NQ * D * ND;     // [NQ][D][NG]
  static constexpr int dev_size = dVgT + NQ * D * NG;
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
    const double id = recip(det);
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
    const double id = recip(det);
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
  double ipv[M];
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
    const double ip = recip(A[c][c]);
    ipv[c] = ip;
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
    const double ip = ipv[c];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      double s = B[c][j];
#pragma unroll
      for (int k = c + 1; k < M; ++k) s -= A[c][k] * B[k][j];
      B[c][j] = s * ip;
    }
  }
}

// Nodal states as the quadrature loop reads them: from the thread's
// registers, or from the patch's LDS copy (which frees the registers of the
// state after the flux-node loop).
template <int ND, int M>
struct URegs
{
  const double (&U)[ND][M];
  __device__ __forceinline__ double operator()(int n, int k) const { return U[n][k]; }
};

template <int ND, int M>
struct ULds
{
  const double *ul;   // [local node][M]
  const int (&row)[ND];  // local node of each cell node
  __device__ __forceinline__ double operator()(int n, int k) const { return ul[row[n] * M + k]; }
};

// A nodal scalar input of the cell (external blending coefficient of the B
// scheme, artificial viscosity of the LDA scheme): none, from the patch's
// LDS rows, or from global memory by node id.
struct NodalNone
{
  __device__ __forceinline__ double operator()(int) const { return 0.0; }
};

// The nodal source field of the reaction residuals (PREC 3) from the
// patch's LDS rows: source(n, k) of the cell's node n.
template <int ND, int M>
struct SourceLds
{
  const double *s;       // [local node][M]
  const int (&row)[ND];  // local node of each cell node
  __device__ __forceinline__ double operator()(int n, int k) const { return s[row[n] * M + k]; }
};

// PREC 5 (the B scheme with a source field and an external blending
// coefficient): source(n, k) as SourceLds, theta(n) from a second LDS column.
template <int ND, int M>
struct SourceThetaLds
{
  const double *s;       // [local node][M]
  const double *th;      // [local node]
  const int (&row)[ND];
  __device__ __forceinline__ double operator()(int n, int k) const { return s[row[n] * M + k]; }
  __device__ __forceinline__ double operator()(int n) const { return th[row[n]]; }
};

template <int ND>
struct NodalLds
{
  const double *v;       // [local node]
  const int (&row)[ND];
  __device__ __forceinline__ double operator()(int n) const { return v[row[n]]; }
};

template <int ND>
struct NodalGlobal
{
  const double *__restrict__ v;  // [N]
  const uint32_t (&nid)[ND];
  __device__ __forceinline__ double operator()(int n) const { return v[nid[n]]; }
};

// The K+ parameters of one node at one quadrature point of an Euler cell.
// With the acoustic pairs written as
//   r+- = (rho / 2a) (r0 +- a e_n),  l+- = (l0v +- a f_n) / (rho a),
//   r0 = (1, V, H), e_n = (0, n, V.n), l0v = (gm1 |V|^2 / 2, -gm1 V, gm1),
//   f_n = (-V.n, n, 0),
// K+_n = l0_n I + dp_n r+ l+^T + dm_n r- l-^T becomes
//   K+_n x = l0_n x + alpha r0 + beta e_n,
//   alpha = (sg_n P + a dl_n Q) / 2a^2, beta = sg_n Q / 2 + dl_n P / 2a,
//   P = l0v . x, Q = f_n . x, sg = dp + dm, dl = dp - dm,
// and S = sum_n K+_n needs only sum l0, sum sg, w = sum dl n and
// T = sum sg n n^T. nh = g / |g| for the physical gradient g of the node's
// basis function, um = V . nh, emax the largest |eigenvalue| |g| (um, um +- a).
template <int D>
struct EulerNodeK
{
  double l0, sg, dl, um, nh[D], emax;
};

// The geometric part of a node's K+ parameters: nh = g / |g| and |g| for the
// physical gradient g of the node's basis function (tiny: |g| negligible).
template <int D>
struct EulerNodeG
{
  double nh[D], norm;
  bool tiny;
};

// Jq: the inverse Jacobian at the point, dvk: the reference gradients of the
// ND basis functions there ([D][ND]).
template <int D, int ND>
__device__ __forceinline__ EulerNodeG<D> euler_node_g(const double (&Jq)[D][D], const double *dvk, int n)
{
  EulerNodeG<D> ge;
  double g[D];
#pragma unroll
  for (int x = 0; x < D; ++x) {
    double t = 0.0;
#pragma unroll
    for (int xi = 0; xi < D; ++xi) t += Jq[x][xi] * dvk[xi * ND + n];
    g[x] = t;
  }
  double nn = 0.0;
#pragma unroll
  for (int x = 0; x < D; ++x) nn += g[x] * g[x];
  double norm, inorm;
  sqrt_rsqrt(nn, norm, inorm);
  norm = (nn > 0.0) ? norm : 0.0;  // sqrt_rsqrt(0) is NaN
  ge.tiny = (D == 2) ? (norm < 1.e-9) : (norm == 0.0);
  inorm = ge.tiny ? 0.0 : inorm;
#pragma unroll
  for (int x = 0; x < D; ++x) ge.nh[x] = g[x] * inorm;
  ge.norm = norm;
  return ge;
}

// The K+ parameters from the geometric part and the point's velocity V and
// sound speed a.
template <int D>
__device__ __forceinline__ EulerNodeK<D> euler_node_k(const EulerNodeG<D> &ge, const double (&V)[D], double a)
{
  EulerNodeK<D> k;
  double u_n = 0.0;
#pragma unroll
  for (int x = 0; x < D; ++x) {
    k.nh[x] = ge.nh[x];
    u_n += V[x] * k.nh[x];
  }
  k.um = u_n;
  const double norm = ge.norm;
  const bool tiny = ge.tiny;
  const double e0 = norm * u_n, ep = norm * (u_n + a), em = norm * (u_n - a);
  k.emax = fmax(fabs(e0), fmax(fabs(ep), fabs(em)));
  const double p0 = fmax(0.0, e0);
  const double dpn = tiny ? 0.0 : fmax(0.0, ep) - p0;
  const double dmn = tiny ? 0.0 : fmax(0.0, em) - p0;
  k.l0 = tiny ? 0.0 : p0;
  k.sg = dpn + dmn;
  k.dl = dpn - dmn;
  return k;
}

template <int D, int ND>
__device__ __forceinline__ EulerNodeK<D> euler_node_k(const double (&Jq)[D][D], const double *dvk, int n,
                                                      const double (&V)[D], double a)
{
  return euler_node_k<D>(euler_node_g<D, ND>(Jq, dvk, n), V, a);
}

// out = K+_n v at a point with velocity V, total enthalpy H, sound speed a,
// i2a = 1 / 2a, i2a2 = 1 / 2a^2 and l0v as above.
template <int D>
__device__ __forceinline__ void euler_kplus(const EulerNodeK<D> &k, const double (&l0v)[D + 2], const double (&V)[D],
                                            double H, double a, double i2a, double i2a2, const double (&v)[D + 2],
                                            double (&out)[D + 2])
{
  constexpr int M = D + 2;
  double P = 0.0;
#pragma unroll
  for (int c = 0; c < M; ++c) P += l0v[c] * v[c];
  double Q = -k.um * v[0];
#pragma unroll
  for (int x = 0; x < D; ++x) Q += k.nh[x] * v[1 + x];
  const double al = (k.sg * P + a * k.dl * Q) * i2a2;
  const double be = 0.5 * k.sg * Q + k.dl * P * i2a;
  out[0] = k.l0 * v[0] + al;
#pragma unroll
  for (int x = 0; x < D; ++x) out[1 + x] = k.l0 * v[1 + x] + al * V[x] + be * k.nh[x];
  out[M - 1] = k.l0 * v[M - 1] + al * H + be * k.um;
}

// S = sum_n K+_n from the node sums (sl0 = sum l0, ssg = sum sg, w, T) as
// S = sl0 I + ssg/2a^2 r0 l0v^T + 1/2 sum sg e f^T + 1/2a (r0 (F w)^T + (E w) l0v^T).
template <int D>
__device__ __forceinline__ void euler_kplus_sum(double sl0, double ssg, const double (&w)[D], const double (&T)[D][D],
                                                const double (&l0v)[D + 2], const double (&V)[D], double H,
                                                double i2a, double i2a2, double (&A)[D + 2][D + 2])
{
  constexpr int M = D + 2;
  double r0[M], ew[M], fw[M];
  r0[0] = 1.0;
#pragma unroll
  for (int x = 0; x < D; ++x) r0[1 + x] = V[x];
  r0[M - 1] = H;
  double vw = 0.0;
#pragma unroll
  for (int x = 0; x < D; ++x) vw += V[x] * w[x];
  ew[0] = 0.0;
  fw[0] = -vw;
#pragma unroll
  for (int x = 0; x < D; ++x) {
    ew[1 + x] = w[x];
    fw[1 + x] = w[x];
  }
  ew[M - 1] = vw;
  fw[M - 1] = 0.0;
  double TV[D], VTV = 0.0;
#pragma unroll
  for (int x = 0; x < D; ++x) {
    double t = 0.0;
#pragma unroll
    for (int y = 0; y < D; ++y) t += T[x][y] * V[y];
    TV[x] = t;
    VTV += V[x] * t;
  }
  const double c1 = ssg * i2a2;
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) {
      // 1/2 sum_n sg_n e_n f_n^T
      double et = 0.0;
      if (i >= 1 && i <= D) {
        if (j == 0) et = -TV[i - 1];
        else if (j <= D) et = T[i - 1][j - 1];
      } else if (i == M - 1) {
        if (j == 0) et = -VTV;
        else if (j <= D) et = TV[j - 1];
      }
      A[i][j] = ((i == j) ? sl0 : 0.0) + c1 * r0[i] * l0v[j] + 0.5 * et + i2a * (r0[i] * fw[j] + ew[i] * l0v[j]);
    }
}

// Residuals phi[ND][M] and wave speed of one element from its vertex
// coordinates X and nodal state U. lds_div: this thread's slice of NQ*M
// doubles of LDS (stride `stride` between entries) that carries the flux
// divergence from the flux-node loop to the quadrature loop.
//
// PREC 0: everything in FP64. PREC 2: FP64 with the
// nodal input `nodal` -- for the B scheme the external blending coefficient
// (set_blending_coeff, PGB.hpp:141-148, 467-477), for the LDA scheme the
// artificial viscosity of the shock-capturing term (PGLDA.hpp:245-256).
// PREC 3: FP64 with a nodal source field (`nodal` gives source(n, k)): the
// reaction residuals compute_adv_reaction_residuals, whose residual at a
// quadrature point is div F - s_q, s_q = V . source (PGLDA.hpp:553-741,
// PGN.hpp:470-660, PGB.hpp:825-1050).
// DLDS: the caller's LDS column at lds_div extends past the NQ M divergence
// slices to ND (M + 1) entries (the patch kernel's work region). The
// dissipative schemes of the 3D P2 cells then run the quadrature in two
// passes that carry the per-point results in that column (see TWO below).
template <int D, int NG, int ND, int NF, int NQ, bool AFFINE, int PH, int SC, int PREC, bool DLDS = false, class UA,
          class NA = NodalNone>
__device__ __forceinline__ void cell_residual(const double *__restrict__ tab, const double (&X)[NG][D],
                                              const double (&U)[ND][PhysTraits<PH>::M], const UA &Uq,
                                              const PhysParams &pp, double (&phi)[ND][PhysTraits<PH>::M],
                                              double &wave, bool &bad, double *lds_div, int stride,
                                              const NA &nodal = NA{})
{
  constexpr int M = PhysTraits<PH>::M;
  static_assert(PhysTraits<PH>::DIM == D, "physics / element dimension mismatch");
  using O = TabOffsets<D, NG, ND, NF, NQ>;
  // ---- geometry ----------------------------------------------------------
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
        for (int g = 0; g < NG; ++g) s += tab[O::dVgT + (q * D + xi) * NG + g] * X[g][x];
        J[xi][x] = s;
      }
    detJ[q] = det_inv<D>(J, Jinv[q]);
  }

  // ---- P+1 flux element: divergence of the interpolated flux at q ---------
  // The flux-node loop stays rolled: each trip reads one row of Vsf and one
  // [D][NQ] slab of dVf with scalar loads, so the table never has to be held
  // in (and spilled from) scalar registers as a whole.
  {
  double divr[NQ][M];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int k = 0; k < M; ++k) divr[q][k] = 0.0;
  // two trips at a time for the Euler cells of more than 4 nodes: their table
  // loads wait once per two trips (the scalar registers do not hold both
  // trips' rows at once otherwise; measured 5 % off that loop on P2 triangles)
#pragma unroll((M > 1 && ND > 4) ? 2 : 1)
  for (int f = 0; f < NF; ++f) {
    const double *vs = tab + O::Vsf + f * ND;
    const double *dvf = tab + O::dVfT + f * D * NQ;
    double s[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      double t = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) t += vs[n] * U[n][k];
      s[k] = t;
    }
    double F[D][M];
    PhysParams ppf = pp;
    if constexpr (PH == PH_ROTATION2D) {
      // the velocity (y, -x) at the flux node's coordinates Vgf X
      double xf[2] = {0.0, 0.0};
#pragma unroll
      for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int d = 0; d < 2; ++d) xf[d] += tab[O::Vgf + f * NG + g] * X[g][d];
      ppf.p0 = xf[1];
      ppf.p1 = -xf[0];
    }
    point_flux<PH, double>(s, ppf, F, bad);
    double G[AFFINE ? D : 1][M];
    if constexpr (AFFINE) {
#pragma unroll
      for (int xi = 0; xi < D; ++xi)
#pragma unroll
        for (int k = 0; k < M; ++k) {
          double t = 0.0;
#pragma unroll
          for (int d = 0; d < D; ++d) t += Jinv[0][d][xi] * F[d][k];
          G[xi][k] = t;
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if constexpr (AFFINE) {
#pragma unroll
        for (int k = 0; k < M; ++k) {
          double t = divr[q][k];
#pragma unroll
          for (int xi = 0; xi < D; ++xi) t += dvf[xi * NQ + q] * G[xi][k];
          divr[q][k] = t;
        }
      } else {
        double b[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          double t = 0.0;
#pragma unroll
          for (int xi = 0; xi < D; ++xi) t += Jinv[q][d][xi] * dvf[xi * NQ + q];
          b[d] = t;
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
          double t = divr[q][k];
#pragma unroll
          for (int d = 0; d < D; ++d) t += b[d] * F[d][k];
          divr[q][k] = t;
        }
      }
    }
  }
  if constexpr (PREC == 3 || PREC == 5) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int k = 0; k < M; ++k) {
        double sq = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) sq += tab[O::V + q * ND + n] * nodal(n, k);
        divr[q][k] -= sq;
      }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int k = 0; k < M; ++k) lds_div[(q * M + k) * stride] = divr[q][k];
  }

  // ---- element scheme ------------------------------------------------------
  constexpr bool NDIS = (SC != SC_LDA);
  double lda[ND][M];
  double dis[NDIS ? ND : 1][M];
  // max over nodes and quadrature points of |eigenvalue|; the floor 1e-6 of
  // every (node, point) term is applied once at the end (max(1e-6, max x) =
  // max over x of max(1e-6, x))
  double lmax = 0.0;
  double fint[M];
#pragma unroll
  for (int n = 0; n < ND; ++n) {
#pragma unroll
    for (int k = 0; k < M; ++k) lda[n][k] = 0.0;
  }
#pragma unroll
  for (int n = 0; n < (NDIS ? ND : 1); ++n)
#pragma unroll
    for (int k = 0; k < M; ++k) dis[n][k] = 0.0;
#pragma unroll
  for (int k = 0; k < M; ++k) fint[k] = 0.0;
  double vol = 0.0;

  // TWO: the dissipative schemes of the 3D P2 cells (affine, NQ 3 M <= ND (M
  // + 1)) in two passes over the quadrature points, so that neither holds the
  // K+ parameters of all nodes beside the 2 ND M residual sums:
  //   1. per point q: S_q = sum_n K+_n (one evaluation of each node's
  //      parameters, which also gives sum_n K+_n U_n), the solve for the LDA
  //      weights x_q = w_q S_q^-1 r_q and ub_q = S_q^-1 sum_n K+_n U_n; the
  //      point's (V, a, H) replace its divergence slice in the LDS column and
  //      x_q, w_q ub_q follow the slices;
  //   2. per point q again, every node: K+_n x_q and K+_n (w_q U_n - w_q ub_q)
  //      added to the node's LDA and dissipation sums (a second evaluation of
  //      the parameters).
  constexpr bool TWO = (D == 3) && NDIS && AFFINE && DLDS && (M == D + 2) && (NQ * 3 * M <= ND * (M + 1));

  // P1 simplices of the Euler physics (affine, D + 1 nodes): the basis
  // gradients, and with them the geometric part of every node's K+
  // parameters, are the same at every quadrature point; formed once per cell
  // (not for the P1 tetrahedra's LDA, which the compiler fits in 2 waves
  // without the registers they take)
  constexpr bool CG = (M > 1) && AFFINE && (ND == D + 1) && !(D == 3 && SC == SC_LDA);
  EulerNodeG<D> cg[CG ? ND : 1];
  if constexpr (CG) {
#pragma unroll
    for (int n = 0; n < ND; ++n) cg[n] = euler_node_g<D, ND>(Jinv[0], tab + O::dVT, n);
  }

  auto quad_point = [&](const int q, const double (&r)[M], const double (&Jq)[D][D], const double detq) {
    const double wj = tab[O::w + q] * detq;
    vol += wj;
#pragma unroll
    for (int k = 0; k < M; ++k) fint[k] += wj * r[k];
    const double *dvq = tab + O::dVT + q * D * ND;  // reference gradients [D][ND] at q

    if constexpr (PREC == 2 && SC == SC_LDA) {
      // shock capturing: w |J| nu_q grad(phi_n) . grad(u_k), nu interpolated
      // from the nodes like the state, added before the LDA share of q
      double nuq = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) nuq += tab[O::V + q * ND + n] * nodal(n);
      if (nuq != 0.0) {
        double gu[D][M];
#pragma unroll
        for (int x = 0; x < D; ++x)
#pragma unroll
          for (int k = 0; k < M; ++k) gu[x][k] = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) {
#pragma unroll
          for (int x = 0; x < D; ++x) {
            double gx = 0.0;
#pragma unroll
            for (int xi = 0; xi < D; ++xi) gx += Jq[x][xi] * dvq[xi * ND + n];
#pragma unroll
            for (int k = 0; k < M; ++k) gu[x][k] += gx * Uq(n, k);
          }
        }
#pragma unroll
        for (int n = 0; n < ND; ++n) {
#pragma unroll
          for (int x = 0; x < D; ++x) {
            double gx = 0.0;
#pragma unroll
            for (int xi = 0; xi < D; ++xi) gx += Jq[x][xi] * dvq[xi * ND + n];
            const double cx = wj * nuq * gx;
#pragma unroll
            for (int k = 0; k < M; ++k) lda[n][k] += cx * gu[x][k];
          }
        }
      }
    }

    if constexpr (M == 1) {
      double vel0, vel1;
      if constexpr (PH == PH_ADVECTION2D) {
        vel0 = double(pp.p0);
        vel1 = double(pp.p1);
      } else if constexpr (PH == PH_ROTATION2D) {
        // the velocity (y, -x) at the quadrature point's coordinates Vg X
        double xq[2] = {0.0, 0.0};
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
          for (int d = 0; d < 2; ++d) xq[d] += tab[O::Vg + q * NG + g] * X[g][d];
        vel0 = double(xq[1]);
        vel1 = double(-xq[0]);
      } else {
        double uq = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) uq += tab[O::V + q * ND + n] * Uq(n, 0);
        vel0 = uq;
        vel1 = 1.0;
      }
      // k_n = a . grad(phi_n) = sum_xi c_xi dphi_n/dxi with c_xi = sum_x a_x dxi/dx
      const double *dvk = dvq;
      double c[D];
#pragma unroll
      for (int xi = 0; xi < D; ++xi) c[xi] = vel0 * double(Jq[0][xi]) + vel1 * double(Jq[1][xi]);
      double kp[ND];
      double S = 0.0, SU = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        double knt = double(0);
#pragma unroll
        for (int xi = 0; xi < D; ++xi) knt += c[xi] * dvk[xi * ND + n];
        const double kn = knt;
        kp[n] = fmax(kn, 0.0);
        lmax = fmax(lmax, fabs(kn));
        S += kp[n];
        SU += kp[n] * Uq(n, 0);
      }
      const double inv = 1.0 / S;
      const double ub = SU * inv;
      const double wr = wj * inv * r[0];
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        lda[n][0] += kp[n] * wr;
        if constexpr (NDIS) dis[n][0] += wj * (kp[n] * (Uq(n, 0) - ub));
      }
    } else {
      double uq[M];
#pragma unroll
      for (int k = 0; k < M; ++k) {
        double t = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) t += tab[O::V + q * ND + n] * Uq(n, k);
        uq[k] = t;
      }
      const EulerState<double, D> st = euler_state<double, D>(uq, double(pp.p0));
      bad = bad || st.bad;
      const double gm1 = pp.p0 - 1.0;
      const double *dvk = dvq;
      // compact K+ of every node (EulerNodeK, euler_kplus)
      using NodeK = EulerNodeK<D>;
      // the K+ parameters of node n (recomputed where keeping them for all
      // nodes costs spills)
      auto node_k = [&](const int n) {
        if constexpr (CG)
          return euler_node_k<D>(cg[n], st.V, st.a);
        else
          return euler_node_k<D, ND>(Jq, dvk, n, st.V, st.a);
      };
      // 3D: recompute in the application instead of keeping ND sets
      constexpr bool KEEP = (D == 2);
      NodeK nk[KEEP ? ND : 1];
      double sl0 = 0.0, ssg = 0.0, w[D], T[D][D];
#pragma unroll
      for (int x = 0; x < D; ++x) {
        w[x] = 0.0;
#pragma unroll
        for (int y = 0; y < D; ++y) T[x][y] = 0.0;
      }
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        const NodeK k = node_k(n);
        if constexpr (KEEP) nk[n] = k;
        lmax = fmax(lmax, k.emax);
        sl0 += k.l0;
        ssg += k.sg;
#pragma unroll
        for (int x = 0; x < D; ++x) {
          w[x] += k.dl * k.nh[x];
#pragma unroll
          for (int y = 0; y < D; ++y) T[x][y] += k.sg * k.nh[x] * k.nh[y];
        }
      }
      const double a = st.a;
      const double i2a = 0.5 * st.inv_a, i2a2 = i2a * st.inv_a;
      double l0v[M];
      l0v[0] = 0.5 * gm1 * st.v2;
#pragma unroll
      for (int x = 0; x < D; ++x) l0v[1 + x] = -gm1 * st.V[x];
      l0v[M - 1] = gm1;
      auto kplus = [&](const NodeK &k, const double (&v)[M], double (&out)[M]) {
        euler_kplus<D>(k, l0v, st.V, st.H, a, i2a, i2a2, v, out);
      };
      auto node_at = [&](const int n) -> NodeK {
        if constexpr (KEEP) {
          return nk[n];
        } else if constexpr (SC == SC_LDA) {
          // an opaque copy of the index: the parameters are recomputed here,
          // not kept alive from the loop over S (which common-subexpression
          // elimination would otherwise do); measured faster for LDA (fewer
          // registers: 2 waves per SIMD on P1, no scratch on P2) and slower
          // for the dissipative schemes, whose second application per node
          // reuses the parameters
          int nn = n;
          asm volatile("" : "+s"(nn));
          return node_k(nn);
        } else {
          return node_k(n);
        }
      };
      // S as in euler_kplus_sum, written out here: the call changes the
      // register allocation of the 2D quadrilateral cells (a 16-byte spill)
      // S = sl0 I + ssg/2a^2 r0 l0v^T + 1/2 sum sg e f^T + 1/2a (r0 (F w)^T + (E w) l0v^T)
      double r0[M], ew[M], fw[M];
      r0[0] = 1.0;
#pragma unroll
      for (int x = 0; x < D; ++x) r0[1 + x] = st.V[x];
      r0[M - 1] = st.H;
      double vw = 0.0;
#pragma unroll
      for (int x = 0; x < D; ++x) vw += st.V[x] * w[x];
      ew[0] = 0.0;
      fw[0] = -vw;
#pragma unroll
      for (int x = 0; x < D; ++x) {
        ew[1 + x] = w[x];
        fw[1 + x] = w[x];
      }
      ew[M - 1] = vw;
      fw[M - 1] = 0.0;
      double TV[D], VTV = 0.0;
#pragma unroll
      for (int x = 0; x < D; ++x) {
        double t = 0.0;
#pragma unroll
        for (int y = 0; y < D; ++y) t += T[x][y] * st.V[y];
        TV[x] = t;
        VTV += st.V[x] * t;
      }
      const double c1 = ssg * i2a2;
      constexpr int R = NDIS ? 2 : 1;
      double A[M][M];
      double B[M][R];
#pragma unroll
      for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) {
          // 1/2 sum_n sg_n e_n f_n^T
          double et = 0.0;
          if (i >= 1 && i <= D) {
            if (j == 0) et = -TV[i - 1];
            else if (j <= D) et = T[i - 1][j - 1];
          } else if (i == M - 1) {
            if (j == 0) et = -VTV;
            else if (j <= D) et = TV[j - 1];
          }
          A[i][j] = ((i == j) ? sl0 : 0.0) + c1 * r0[i] * l0v[j] + 0.5 * et + i2a * (r0[i] * fw[j] + ew[i] * l0v[j]);
        }
#pragma unroll
      for (int i = 0; i < M; ++i) B[i][0] = r[i];
      if constexpr (NDIS) {
#pragma unroll
        for (int i = 0; i < M; ++i) B[i][1] = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) {
          double un[M], z[M];
#pragma unroll
          for (int k = 0; k < M; ++k) un[k] = Uq(n, k);
          kplus(node_at(n), un, z);
#pragma unroll
          for (int k = 0; k < M; ++k) B[k][1] += z[k];
        }
      }
      solve_pp<M, R>(A, B);
      double x[M];
#pragma unroll
      for (int k = 0; k < M; ++k) x[k] = B[k][0] * wj;
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        const NodeK kn = node_at(n);
        double t[M];
        kplus(kn, x, t);
#pragma unroll
        for (int k = 0; k < M; ++k) lda[n][k] += t[k];
        if constexpr (NDIS) {
          double du[M];
#pragma unroll
          for (int k = 0; k < M; ++k) du[k] = wj * (Uq(n, k) - B[k][R - 1]);
          kplus(kn, du, t);
#pragma unroll
          for (int k = 0; k < M; ++k) dis[n][k] += t[k];
        }
      }
    }
  };

  if constexpr (TWO) {
    const double gm1 = pp.p0 - 1.0;
    const double(&Jq)[D][D] = Jinv[0];
    // LDS column: [q M, q M + M) the divergence of q, then (V, a, H) of q;
    // [NQ M + 2 M q, + M) x_q, [NQ M + 2 M q + M, + M) w_q ub_q
    double *const qx = lds_div + static_cast<size_t>(NQ) * M * stride;
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      double r[M];
#pragma unroll
      for (int k = 0; k < M; ++k) r[k] = lds_div[(q * M + k) * stride];
      const double wj = tab[O::w + q] * detJ[0];
      vol += wj;
#pragma unroll
      for (int k = 0; k < M; ++k) fint[k] += wj * r[k];
      const double *dvk = tab + O::dVT + q * D * ND;
      double uq[M];
#pragma unroll
      for (int k = 0; k < M; ++k) {
        double t = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) t += tab[O::V + q * ND + n] * Uq(n, k);
        uq[k] = t;
      }
      const EulerState<double, D> st = euler_state<double, D>(uq, double(pp.p0));
      bad = bad || st.bad;
      const double i2a = 0.5 * st.inv_a, i2a2 = i2a * st.inv_a;
      double l0v[M];
      l0v[0] = 0.5 * gm1 * st.v2;
#pragma unroll
      for (int x = 0; x < D; ++x) l0v[1 + x] = -gm1 * st.V[x];
      l0v[M - 1] = gm1;
      double sl0 = 0.0, ssg = 0.0, w[D], T[D][D], B[M][2];
#pragma unroll
      for (int x = 0; x < D; ++x) {
        w[x] = 0.0;
#pragma unroll
        for (int y = 0; y < D; ++y) T[x][y] = 0.0;
      }
#pragma unroll
      for (int k = 0; k < M; ++k) {
        B[k][0] = r[k];
        B[k][1] = 0.0;
      }
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        const EulerNodeK<D> k = euler_node_k<D, ND>(Jq, dvk, n, st.V, st.a);
        lmax = fmax(lmax, k.emax);
        sl0 += k.l0;
        ssg += k.sg;
#pragma unroll
        for (int x = 0; x < D; ++x) {
          w[x] += k.dl * k.nh[x];
#pragma unroll
          for (int y = 0; y < D; ++y) T[x][y] += k.sg * k.nh[x] * k.nh[y];
        }
        double un[M], z[M];
#pragma unroll
        for (int c = 0; c < M; ++c) un[c] = Uq(n, c);
        euler_kplus<D>(k, l0v, st.V, st.H, st.a, i2a, i2a2, un, z);
#pragma unroll
        for (int c = 0; c < M; ++c) B[c][1] += z[c];
      }
      double A[M][M];
      euler_kplus_sum<D>(sl0, ssg, w, T, l0v, st.V, st.H, i2a, i2a2, A);
      solve_pp<M, 2>(A, B);
#pragma unroll
      for (int x = 0; x < D; ++x) lds_div[(q * M + x) * stride] = st.V[x];
      lds_div[(q * M + D) * stride] = st.a;
      lds_div[(q * M + D + 1) * stride] = st.H;
#pragma unroll
      for (int k = 0; k < M; ++k) {
        qx[(2 * M * q + k) * stride] = B[k][0] * wj;
        qx[(2 * M * q + M + k) * stride] = B[k][1] * wj;
      }
    }
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      double V[D], x[M], wub[M];
#pragma unroll
      for (int d = 0; d < D; ++d) V[d] = lds_div[(q * M + d) * stride];
      const double a = lds_div[(q * M + D) * stride];
      const double H = lds_div[(q * M + D + 1) * stride];
#pragma unroll
      for (int k = 0; k < M; ++k) {
        x[k] = qx[(2 * M * q + k) * stride];
        wub[k] = qx[(2 * M * q + M + k) * stride];
      }
      const double wj = tab[O::w + q] * detJ[0];
      const double ia = recip(a);
      const double i2a = 0.5 * ia, i2a2 = i2a * ia;
      double v2 = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) v2 += V[d] * V[d];
      double l0v[M];
      l0v[0] = 0.5 * gm1 * v2;
#pragma unroll
      for (int d = 0; d < D; ++d) l0v[1 + d] = -gm1 * V[d];
      l0v[M - 1] = gm1;
      const double *dvk = tab + O::dVT + q * D * ND;
#pragma unroll
      for (int n = 0; n < ND; ++n) {
        const EulerNodeK<D> k = euler_node_k<D, ND>(Jq, dvk, n, V, a);
        double t[M], du[M];
        euler_kplus<D>(k, l0v, V, H, a, i2a, i2a2, x, t);
#pragma unroll
        for (int c = 0; c < M; ++c) {
          lda[n][c] += t[c];
          du[c] = wj * Uq(n, c) - wub[c];
        }
        euler_kplus<D>(k, l0v, V, H, a, i2a, i2a2, du, t);
#pragma unroll
        for (int c = 0; c < M; ++c) dis[n][c] += t[c];
      }
    }
  } else {
#pragma unroll 1
  for (int q = 0; q < NQ; ++q) {
    double r[M];
#pragma unroll
    for (int k = 0; k < M; ++k) r[k] = lds_div[(q * M + k) * stride];
    if constexpr (AFFINE) {
      quad_point(q, r, Jinv[0], detJ[0]);
    } else {
      // recompute the Jacobian at q instead of indexing a per-q array
      double J[D][D], Ji[D][D];
#pragma unroll
      for (int xi = 0; xi < D; ++xi)
#pragma unroll
        for (int x = 0; x < D; ++x) {
          double t = 0.0;
#pragma unroll
          for (int g = 0; g < NG; ++g) t += tab[O::dVgT + (q * D + xi) * NG + g] * X[g][x];
          J[xi][x] = t;
        }
      const double dq = det_inv<D>(J, Ji);
      quad_point(q, r, Ji, dq);
    }
  }
  }

  // ---- blend / limit -------------------------------------------------------
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
  } else if constexpr (SC == SC_B && (PREC == 2 || PREC == 5)) {
    // external nodal coefficient; below 1e-6 at every node the cell keeps
    // its LDA residuals
    double th[ND];
    double tmax = 0.0;
#pragma unroll
    for (int n = 0; n < ND; ++n) {
      th[n] = nodal(n);
      tmax = fmax(tmax, th[n]);
    }
    const bool skip = tmax < 1.e-6;
#pragma unroll
    for (int n = 0; n < ND; ++n)
#pragma unroll
      for (int k = 0; k < M; ++k) phi[n][k] = skip ? lda[n][k] : lda[n][k] + th[n] * dis[n][k];
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
    // PSI: the x+ limiter of the N residuals, per equation; the PSI / LDA
    // blend then takes (1 - theta) LDA + theta PSI with PGB's cell theta
    double theta = 1.0;
    if constexpr (SC == SC_BPSI) {
      double an = 0.0;
#pragma unroll
      for (int n = 0; n < ND; ++n) an += fabs(lda[n][0] + dis[n][0]);
      theta = fabs(fint[0]) / (an + 1.e-6);
    }
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
      for (int n = 0; n < ND; ++n) {
        const double psi = lim ? xs[n] / sx * tot : lda[n][k] + dis[n][k];
        phi[n][k] = (SC == SC_BPSI) ? (1.0 - theta) * lda[n][k] + theta * psi : psi;
      }
    }
  }

  wave = vol * fmax(lmax, 1.e-6);
}

template <int D, int NG, int NQ, int M>
__device__ __forceinline__ void gather_coords(const CellArgs &a, uint32_t e, double (&X)[NG][D])
{
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const size_t v = a.gconn[static_cast<size_t>(g) * a.ne + e];
#pragma unroll
    for (int d = 0; d < D; ++d) X[g][d] = a.coords[v * D + d];
  }
}

// Element-per-thread cell loop with a global scatter: FP64 atomics or, one
// colour per launch, plain read-modify-write.
template <int D, int NG, int ND, int NF, int NQ, bool AFFINE, int PH, int SC, int PREC = 0>
__global__ __launch_bounds__(256) void rd_cell_kernel(CellArgs a)
{
  constexpr int M = PhysTraits<PH>::M;
  extern __shared__ double lds[];
  const uint32_t local = xcd_block_index() * blockDim.x + threadIdx.x;
  if (local >= a.count) return;
  const uint32_t e = a.first + local;
  bool bad = false;
  uint32_t nid[ND];
#pragma unroll
  for (int n = 0; n < ND; ++n) nid[n] = a.sconn[static_cast<size_t>(n) * a.ne + e];
  if (a.nn_check) {
    bool out = false;
#pragma unroll
    for (int g = 0; g < NG; ++g) out = out || a.gconn[static_cast<size_t>(g) * a.ne + e] >= a.nv_check;
#pragma unroll
    for (int n = 0; n < ND; ++n) out = out || nid[n] >= a.nn_check;
    if (out) {
      atomicOr(a.status, 2);
      return;
    }
  }
  double X[NG][D];
  gather_coords<D, NG, NQ, M>(a, e, X);
  double U[ND][M];
#pragma unroll
  for (int n = 0; n < ND; ++n)
#pragma unroll
    for (int k = 0; k < M; ++k) U[n][k] = a.u[static_cast<size_t>(nid[n]) * M + k];
  double phi[ND][M];
  double wave;
  if constexpr (PREC == 2)
    cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, 2>(a.tab, X, U, URegs<ND, M>{U}, a.pp, phi, wave, bad,
                                                     lds + threadIdx.x, static_cast<int>(blockDim.x),
                                                     NodalGlobal<ND>{a.nodal, nid});
  else
    cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, 0>(a.tab, X, U, URegs<ND, M>{U}, a.pp, phi, wave, bad,
                                                     lds + threadIdx.x, static_cast<int>(blockDim.x));
#pragma unroll
  for (int n = 0; n < ND; ++n) {
    const size_t b = static_cast<size_t>(nid[n]);
#pragma unroll
    for (int k = 0; k < M; ++k) scatter_add(a.acc_res + b * M + k, phi[n][k], a.atomic);
    scatter_add(a.acc_ws + b, wave, a.atomic);
  }
  if (bad) atomicOr(a.status, 1);
}

// ---------------------------------------------------------------------------
// Patch assembly: one workgroup owns one patch of up to PATCH cells (a
// spatially compact run of the Morton-ordered cells). The states of the
// patch's nodes are gathered once into LDS; every cell writes its residuals
// and wave speed to the LDS slot of each of its (cell, node) pairs, the pairs
// being grouped by node (the rank of a pair within its node comes with the
// connectivity), so each node then sums a contiguous run of slots in a fixed
// order -- no atomics and the same result on every run. Then
//   * nodes held by this patch only are finished here: residual / V, wave
//     speed, and -- in iteration mode -- dt = CFL V / sum lambda,
//     u -= dt res (after the strong Dirichlet value),
//   * nodes shared with other patches write their partial sums to their slot;
//     patch_fixup_kernel sums the slots of every shared node in a fixed order.
// Between two iterations the fix-up can be left out: the next sweep's
// patches then finish their shared nodes themselves while gathering them
// (prev_acc: the previous sweep's rows; base_in: the shared nodes' states at
// its start), each holder summing the same rows in the same order, so all
// holders agree bit for bit with each other and with the fix-up. The holder
// owning the node's first row (its lowest-numbered patch) stores the new
// state to u and, when the sweep after this one is deferred too, to
// base_out. The rows and the states alternate between two buffers, since
// other patches of a sweep still read the previous sweep's.
// ---------------------------------------------------------------------------
constexpr int LC_BITS = 10;  // local node index bits of a patch connectivity entry
constexpr int LC_MASK = (1 << LC_BITS) - 1;

struct PatchArgs
{
  const double *__restrict__ tab;
  const uint32_t *__restrict__ gconn;       // [NG][ne] global vertex ids
  const uint16_t *__restrict__ lconn;       // [ND][ne] local node | rank << LC_BITS
  const uint32_t *__restrict__ pnode_off;   // [np + 1] first slot of each patch
  const uint32_t *__restrict__ pnodes;      // [slots] global node id of each slot
  const uint32_t *__restrict__ pn_inner;    // [np] number of patch-only nodes (listed first)
  const uint16_t *__restrict__ pn_coff;     // [slots] first (cell, node) pair of the slot's node
  const double *__restrict__ coords;
  double *__restrict__ u;                   // [N][M] (updated in place in iteration mode)
  const double *__restrict__ vol;           // [N]
  const int *__restrict__ dir_index;        // [N] row in dir_vals or -1
  const double *__restrict__ dir_vals;      // [ndir][M]
  // linear sweeps of the mixed-precision form (PREC 4): the element
  // Jacobians of the block in FP32, [ND M][ND M][ne] (entry-major, patch
  // order), and the linearisation point: its state u_ref [N][M], its FP64
  // residual lin_res [N][M] and time step lin_dt [N]
  const float *__restrict__ jlin;
  const double *__restrict__ u_ref;
  const double *__restrict__ lin_res;
  const double *__restrict__ lin_dt;
  double *__restrict__ res;                 // [N][M]
  double *__restrict__ dt;                  // [N]
  double *__restrict__ ws;                  // [N]
  double *__restrict__ slot_acc;            // [shared slots][M + 1] partial sums, grouped by node
  const uint32_t *__restrict__ slot_pos;    // [slots] row of a shared slot in slot_acc
  const uint32_t *__restrict__ pn_sid;      // [slots] shared-node index of a shared slot
  const uint2 *__restrict__ pn_rows;        // [slots] rows [x, y) of a shared slot's node
  const double *__restrict__ prev_acc;      // rows of the previous sweep, not yet summed, or null
  const double *__restrict__ base_in;       // [n_shared][M] states at the start of the previous sweep
  double *__restrict__ base_out;            // [n_shared][M] states at the start of this sweep, or null
  double *__restrict__ norm_partial;        // [np][M + 1]
  const double *__restrict__ nodal;         // [N] blending coefficient (B) / artificial viscosity (LDA), PREC 2;
                                            // [N][M] source field, PREC 3 and 5
  const double *__restrict__ theta;         // [N] PREC 5: the B scheme's external blending coefficient
  int *__restrict__ status;
  uint32_t ne, np, first_patch;             // first_patch: offset of this block's patches
  int maxl;                                 // LDS rows of node state
  int update;                               // 0: assemble (res, ws), 1: iterate
  int outputs;                              // iterate: also store res, dt, ws and norm partials
  int excl;                                 // norm without Dirichlet rows
  double cfl;
  PhysParams pp;
};

// Largest (and default) number of cells per patch (= threads per
// workgroup): 256 for scalar physics on elements of up to 6 nodes, 128 for
// the Euler systems (larger per-thread state) and for P3 / Q2 elements (more
// nodes per patch), 64 for the 3D Euler equations on P2 tetrahedra (whose
// (cell, node) pair sums of 128 cells alone would fill 60 KB of LDS). A block
// may use 64 or 128 instead (PDRD_OPT_PATCH_CELLS).
template <int M, int ND>
struct PatchSize
{
  static constexpr int value = (M == 1 && ND <= 6) ? 256 : ((M == 5 && ND > 4) ? 64 : 128);
};

// Waves per SIMD the register allocation of a patch kernel must leave room
// for: 5 for the scalar affine P1 / P2 cells (at most 96 VGPRs, measured
// faster than 4 waves); 2 for the 2D Euler cells of up to 4 nodes (little or
// no scratch, measured faster than 1 wave); the others keep the compiler's
// choice (the FP64 3D Euler cells run as fast at 1 wave, where capping them
// at 2 spills a few hundred bytes per thread to scratch memory; the 2D Euler
// P2 / Q2 / P3 cells spill 250-1600 bytes per thread at 2 waves). The linear
// sweeps (PREC 4, J delta per cell) are bound by the Jacobians' bandwidth and
// keep 4 waves (2 for blocks of more than 24 unknowns).
template <int M, int ND, bool AFFINE, int SC, int PREC>
struct PatchMinWaves
{
  static constexpr int value =
      (PREC == 4) ? (ND * M <= 24 ? 4 : 2)
                  : ((M == 1 && ND <= 6 && AFFINE) ? 5 : ((M == 4 && ND <= 4) ? 2 : 1));
};

// Dynamic LDS of one patch workgroup of P cells, in doubles: node state
// [maxl][M], the pair offsets of the nodes [maxl] (u16), then a work region
// shared in turn by the flux-divergence slices, the (cell, node) pair sums
// and the norm reduction.
template <int NQ, int ND, int M>
__host__ __device__ constexpr size_t patch_work_doubles(int P)
{
  const size_t div = static_cast<size_t>(NQ) * M * P;
  const size_t pairs = static_cast<size_t>(P) * ND * (M + 1);
  const size_t red = static_cast<size_t>(P) * (M + 1);
  const size_t w = div > pairs ? div : pairs;
  return w > red ? w : red;
}

__host__ __device__ constexpr size_t patch_coff_doubles(int maxl) { return (static_cast<size_t>(maxl) + 3) / 4; }

template <int NQ, int ND, int M>
__host__ __device__ constexpr size_t patch_lds_doubles(int maxl, int P)
{
  return static_cast<size_t>(maxl) * M + patch_coff_doubles(maxl) + patch_work_doubles<NQ, ND, M>(P);
}

// With a nodal input (PREC 2) its values follow the work region, one per
// local node; with a source field (PREC 3), M per local node. A linear sweep
// (PREC 4) has no divergence slices: its work region holds the pair sums
// and the norm reduction, followed by the state differences delta of the
// patch's nodes, [maxl][M] floats.
template <int NQ, int ND, int M>
__host__ __device__ constexpr size_t patch_lin_work_doubles(int P)
{
  const size_t pairs = static_cast<size_t>(P) * ND * (M + 1);
  const size_t red = static_cast<size_t>(P) * (M + 1);
  return pairs > red ? pairs : red;
}

template <int NQ, int ND, int M, int PREC>
__host__ __device__ constexpr size_t patch_lds_doubles_prec(int maxl, int P)
{
  if (PREC == 4)
    return static_cast<size_t>(maxl) * M + patch_coff_doubles(maxl) + patch_lin_work_doubles<NQ, ND, M>(P) +
           (static_cast<size_t>(maxl) * M + 1) / 2;
  return patch_lds_doubles<NQ, ND, M>(maxl, P) +
         (PREC == 2 ? static_cast<size_t>(maxl)
                    : (PREC == 3 ? static_cast<size_t>(maxl) * M : (PREC == 5 ? static_cast<size_t>(maxl) * (M + 1) : 0)));
}

// Finishes node g from its sums acc[M + 1] (residual sum, wave-speed sum);
// u_old is the node's state at the start of the sweep.
// Sum of the partial-sum rows [r0, r1) of one shared node, in row order;
// the loads of four rows are issued before their additions.
template <int M>
__device__ __forceinline__ void sum_rows(const double *__restrict__ rows, uint32_t r0, uint32_t r1, double (&acc)[M + 1])
{
  constexpr int RB = 4;
#pragma unroll
  for (int k = 0; k <= M; ++k) acc[k] = 0.0;
  for (uint32_t s = r0; s < r1; s += RB) {
    double v[RB][M + 1];
#pragma unroll
    for (int j = 0; j < RB; ++j)
      if (s + j < r1)
#pragma unroll
        for (int k = 0; k <= M; ++k) v[j][k] = rows[static_cast<size_t>(s + j) * (M + 1) + k];
#pragma unroll
    for (int j = 0; j < RB; ++j)
      if (s + j < r1)
#pragma unroll
        for (int k = 0; k <= M; ++k) acc[k] += v[j][k];
  }
}

// The pseudo-time step of node g: residual r = sum / V, dt = CFL V / sum
// lambda, new state = (strong Dirichlet value or u_old) - dt r. In a linear
// sweep (lin_res set) the sums are those of J delta and r = lin_res + sum / V,
// dt = lin_dt: the residual linearised about u_ref with the time step there.
// Both the in-kernel finish and the deferred one use it, so they round alike.
// v = vol[g] and di = dir_index[g] come from the caller (loaded ahead).
template <int M>
__device__ __forceinline__ void node_step(const PatchArgs &a, uint32_t g, const double *acc, const double *u_old,
                                          double v, int di, double (&r)[M], double (&un)[M], double &d)
{
  const double iv = 1.0 / v;
  d = a.lin_res ? a.lin_dt[g] : a.cfl * v / acc[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const size_t j = static_cast<size_t>(g) * M + k;
    r[k] = a.lin_res ? acc[k] * iv + a.lin_res[j] : acc[k] * iv;
    const double u0 = (di >= 0) ? a.dir_vals[static_cast<size_t>(di) * M + k] : u_old[k];
    un[k] = u0 - d * r[k];
  }
}

template <int M>
__device__ __forceinline__ void finish_node(const PatchArgs &a, uint32_t g, const double *acc, const double *u_old,
                                            double v, int di, double (&sq)[M + 1])
{
  if (a.update) {
    double r[M], un[M], d;
    node_step<M>(a, g, acc, u_old, v, di, r, un, d);
    const bool count = a.outputs && !(a.excl && di >= 0);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const size_t j = static_cast<size_t>(g) * M + k;
      a.u[j] = un[k];
      if (a.outputs) a.res[j] = r[k];
      if (count) sq[k] += r[k] * r[k];
    }
    if (count) sq[M] += 1.0;
    if (a.outputs) {
      a.dt[g] = d;
      a.ws[g] = acc[M];
    }
  } else {
    const double iv = 1.0 / v;
#pragma unroll
    for (int k = 0; k < M; ++k) a.res[static_cast<size_t>(g) * M + k] = acc[k] * iv;
    a.ws[g] = acc[M];
  }
}

template <int M>
__device__ __forceinline__ void block_sum_store(double (&sq)[M + 1], double *red, double *out)
{
  // red: blockDim * (M + 1) doubles of LDS
  __syncthreads();
#pragma unroll
  for (int k = 0; k <= M; ++k) red[k * blockDim.x + threadIdx.x] = sq[k];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k <= M; ++k) red[k * blockDim.x + threadIdx.x] += red[k * blockDim.x + threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k <= M; ++k) out[k] = red[k * blockDim.x];
  }
}

// One patch (local index pl of the block described by a) on one workgroup.
template <int D, int NG, int ND, int NF, int NQ, bool AFFINE, int PH, int SC, int PREC>
__device__ __forceinline__ void patch_body(const PatchArgs &a, uint32_t pl, double *lds)
{
  constexpr int M = PhysTraits<PH>::M;
  const int PATCH = blockDim.x;     // cells per patch of this block (a multiple of 64, <= PatchSize)
  const int NPAIR = PATCH * ND;      // stride of one component in the pair sums
  double *ul = lds;
  uint16_t *coff = reinterpret_cast<uint16_t *>(lds + static_cast<size_t>(a.maxl) * M);
  double *work = lds + static_cast<size_t>(a.maxl) * M + patch_coff_doubles(a.maxl);

  const uint32_t p = a.first_patch + pl;
  const uint32_t s0 = a.pnode_off[p];
  const int nl = static_cast<int>(a.pnode_off[p + 1] - s0);
  const int tid = threadIdx.x;
  const uint32_t e0 = pl * PATCH;
  const int ecount = static_cast<int>(min(static_cast<uint32_t>(PATCH), a.ne - e0));
  const bool active = tid < ecount;

  // the cell's vertices and local nodes do not depend on LDS: issue their
  // loads before the gather
  double X[NG][D];
  uint16_t lc[ND];
  if (active) {
    const uint32_t e = e0 + tid;
    if constexpr (PREC != 4) {
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const size_t v = a.gconn[static_cast<size_t>(g) * a.ne + e];
#pragma unroll
        for (int d = 0; d < D; ++d) X[g][d] = a.coords[v * D + d];
      }
    }
#pragma unroll
    for (int n = 0; n < ND; ++n) lc[n] = a.lconn[static_cast<size_t>(n) * a.ne + e];
  }
  double *nodal_l = work + patch_work_doubles<NQ, ND, M>(PATCH);  // PREC 2 and 3 only
  float *delta_l = reinterpret_cast<float *>(work + patch_lin_work_doubles<NQ, ND, M>(PATCH));  // PREC 4 only
  // (no global store before the cell work: the table loads stay scalar)
  const int ninner = static_cast<int>(a.pn_inner[p]);
  // The kernels at 1 wave per SIMD load what the node finish needs of the
  // thread's first FP nodes (node id and slot here, volume and Dirichlet
  // index after the gather's barrier) so that it arrives during the cell
  // work: at the finish, those loads behind the node id were two rounds of
  // memory latency per node that no other wave hid (a sixth of the wave's
  // life on the P2 triangles).
  // (the Euler cells only, not the P1 tetrahedra's LDA, which the compiler
  // fits in 2 waves)
  constexpr int FP = (M > 1 && PatchMinWaves<M, ND, AFFINE, SC, PREC>::value == 1 && PREC != 4 &&
                      (ND > 4 || SC != SC_LDA))
                         ? 4
                         : 0;
  constexpr int FPS = FP > 0 ? FP : 1;
  uint32_t pg[FPS], psp[FPS];
  double pv[FPS];
  int pdi[FPS];
  if constexpr (FP > 0) {
#pragma unroll
    for (int j = 0; j < FP; ++j) {
      const int l = min(tid + j * PATCH, nl - 1);
      pg[j] = a.pnodes[s0 + l];
      psp[j] = a.slot_pos[s0 + l];
    }
  }
  for (int l = tid; l < nl; l += PATCH) {
    const uint32_t g = a.pnodes[s0 + l];
    double un[M];
    if (a.prev_acc && l >= ninner) {
      // shared node whose previous sweep is summed here
      const uint32_t i = a.pn_sid[s0 + l];
      const uint2 rr = a.pn_rows[s0 + l];
      double acc[M + 1], uo[M], r[M], d;
#pragma unroll
      for (int k = 0; k < M; ++k) uo[k] = a.base_in[static_cast<size_t>(i) * M + k];
      sum_rows<M>(a.prev_acc, rr.x, rr.y, acc);
      node_step<M>(a, g, acc, uo, a.vol[g], a.dir_index[g], r, un, d);
    } else {
#pragma unroll
      for (int k = 0; k < M; ++k) un[k] = a.u[static_cast<size_t>(g) * M + k];
    }
#pragma unroll
    for (int k = 0; k < M; ++k) ul[l * M + k] = un[k];
    if constexpr (PREC == 4)
#pragma unroll
      for (int k = 0; k < M; ++k)
        delta_l[l * M + k] = static_cast<float>(un[k] - a.u_ref[static_cast<size_t>(g) * M + k]);
    coff[l] = a.pn_coff[s0 + l];
    if constexpr (PREC == 2) nodal_l[l] = a.nodal[g];
    if constexpr (PREC == 3 || PREC == 5)
#pragma unroll
      for (int k = 0; k < M; ++k) nodal_l[l * M + k] = a.nodal[static_cast<size_t>(g) * M + k];
    if constexpr (PREC == 5) nodal_l[static_cast<size_t>(a.maxl) * M + l] = a.theta[g];
  }
  __syncthreads();
  // issued after the barrier (which waits for every outstanding load), so
  // that they complete during the cell work
  if constexpr (FP > 0) {
#pragma unroll
    for (int j = 0; j < FP; ++j) {
      pv[j] = a.vol[pg[j]];
      pdi[j] = a.dir_index[pg[j]];
    }
  }

  bool bad = false;
  double phi[ND][M];
  double wave = 0.0;
  if constexpr (PREC == 4) {
    // linear sweep: J delta with the cell's FP32 Jacobian (entry-major over
    // the block's cells, so that every load is coalesced), written straight
    // to the (cell, node) pair slots (no divergence slices share the work
    // region here); the node loop stays rolled, one block row per trip
    if (active) {
      constexpr int NDM = ND * M;
      const int NPAIR_ = blockDim.x * ND;
      float dv[NDM];
#pragma unroll
      for (int n = 0; n < ND; ++n)
#pragma unroll
        for (int k = 0; k < M; ++k) dv[n * M + k] = delta_l[(lc[n] & LC_MASK) * M + k];
      const float *__restrict__ jc = a.jlin + (e0 + tid);
      const size_t stride = a.ne;
#pragma unroll 1
      for (int n = 0; n < ND; ++n) {
        const uint16_t l = a.lconn[static_cast<size_t>(n) * a.ne + e0 + tid];
        const int slot = coff[l & LC_MASK] + (l >> LC_BITS);
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const float *__restrict__ row = jc + static_cast<size_t>((n * M + k) * NDM) * stride;
          float y = 0.0f;
#pragma unroll
          for (int j = 0; j < NDM; ++j) y += row[static_cast<size_t>(j) * stride] * dv[j];
          work[k * NPAIR_ + slot] = static_cast<double>(y);
        }
        work[M * NPAIR_ + slot] = 0.0;
      }
    }
  } else if (active) {
    double U[ND][M];
    int row[ND];
#pragma unroll
    for (int n = 0; n < ND; ++n) {
      row[n] = lc[n] & LC_MASK;
#pragma unroll
      for (int k = 0; k < M; ++k) U[n][k] = ul[row[n] * M + k];
    }
    if constexpr (PREC == 5 && M == 1)
      cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC, true>(
          a.tab, X, U, URegs<ND, M>{U}, a.pp, phi, wave, bad, work + tid, PATCH,
          SourceThetaLds<ND, M>{nodal_l, nodal_l + static_cast<size_t>(a.maxl) * M, row});
    else if constexpr (PREC == 5)
      cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC, true>(
          a.tab, X, U, ULds<ND, M>{ul, row}, a.pp, phi, wave, bad, work + tid, PATCH,
          SourceThetaLds<ND, M>{nodal_l, nodal_l + static_cast<size_t>(a.maxl) * M, row});
    else if constexpr (PREC == 3 && M == 1)
      cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC, true>(a.tab, X, U, URegs<ND, M>{U}, a.pp, phi, wave, bad, work + tid,
                                                       PATCH, SourceLds<ND, M>{nodal_l, row});
    else if constexpr (PREC == 3)
      cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC, true>(a.tab, X, U, ULds<ND, M>{ul, row}, a.pp, phi, wave, bad,
                                                       work + tid, PATCH, SourceLds<ND, M>{nodal_l, row});
    else if constexpr (PREC == 2 && M == 1)
      cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC, true>(a.tab, X, U, URegs<ND, M>{U}, a.pp, phi, wave, bad, work + tid,
                                                       PATCH, NodalLds<ND>{nodal_l, row});
    else if constexpr (PREC == 2)
      cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC, true>(a.tab, X, U, ULds<ND, M>{ul, row}, a.pp, phi, wave, bad,
                                                       work + tid, PATCH, NodalLds<ND>{nodal_l, row});
    else if constexpr (M == 1)
      cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC, true>(a.tab, X, U, URegs<ND, M>{U}, a.pp, phi, wave, bad, work + tid,
                                                       PATCH);
    else
      cell_residual<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC, true>(a.tab, X, U, ULds<ND, M>{ul, row}, a.pp, phi, wave, bad,
                                                       work + tid, PATCH);
  }
  __syncthreads();  // the work region changes from divergence slices to pair sums
  if (active && PREC != 4) {
#pragma unroll
    for (int n = 0; n < ND; ++n) {
      const int slot = coff[lc[n] & LC_MASK] + (lc[n] >> LC_BITS);
#pragma unroll
      for (int k = 0; k < M; ++k) work[k * NPAIR + slot] = phi[n][k];
      work[M * NPAIR + slot] = wave;
    }
  }
  __syncthreads();

  const int npair = ecount * ND;
  double sq[M + 1];
#pragma unroll
  for (int k = 0; k <= M; ++k) sq[k] = 0.0;
  // node l of the patch; pre: its node id, slot, volume and Dirichlet index
  // were loaded ahead (pgv, pspv, pvv, pdiv)
  auto finish_l = [&](const int l, const bool pre, const uint32_t pgv, const uint32_t pspv, const double pvv,
                      const int pdiv) {
    const int b = coff[l];
    const int e = (l + 1 < nl) ? static_cast<int>(coff[l + 1]) : npair;
    double acc[M + 1];
#pragma unroll
    for (int k = 0; k <= M; ++k) acc[k] = 0.0;
    for (int i = b; i < e; ++i)
#pragma unroll
      for (int k = 0; k <= M; ++k) acc[k] += work[k * NPAIR + i];
    if (l < ninner) {
      const uint32_t g = pre ? pgv : a.pnodes[s0 + l];
      finish_node<M>(a, g, acc, ul + l * M, pre ? pvv : a.vol[g], pre ? pdiv : a.dir_index[g], sq);
    } else {
      const uint32_t sp = pre ? pspv : a.slot_pos[s0 + l];
      double *slot = a.slot_acc + static_cast<size_t>(sp) * (M + 1);
#pragma unroll
      for (int k = 0; k <= M; ++k) slot[k] = acc[k];
      if (a.prev_acc || a.base_out) {
        // the holder of the node's first row stores its start state
        const uint32_t i = a.pn_sid[s0 + l];
        if (sp == a.pn_rows[s0 + l].x) {
#pragma unroll
          for (int k = 0; k < M; ++k) {
            const double v = ul[l * M + k];
            if (a.prev_acc) a.u[static_cast<size_t>(a.pnodes[s0 + l]) * M + k] = v;
            if (a.base_out) a.base_out[static_cast<size_t>(i) * M + k] = v;
          }
        }
      }
    }
  };
  if constexpr (FP > 0) {
#pragma unroll
    for (int j = 0; j < FP; ++j) {
      const int l = tid + j * PATCH;
      if (l < nl) finish_l(l, true, pg[j], psp[j], pv[j], pdi[j]);
    }
    for (int l = tid + FP * PATCH; l < nl; l += PATCH) finish_l(l, false, 0u, 0u, 0.0, 0);
  } else {
    (void)pg;
    (void)psp;
    (void)pv;
    (void)pdi;
    for (int l = tid; l < nl; l += PATCH) finish_l(l, false, 0u, 0u, 0.0, 0);
  }
  if (bad) atomicOr(a.status, 1);
  if (a.update && a.outputs) block_sum_store<M>(sq, work, a.norm_partial + static_cast<size_t>(p) * (M + 1));
}

template <int D, int NG, int ND, int NF, int NQ, bool AFFINE, int PH, int SC, int PREC>
__global__ __launch_bounds__((PatchSize<PhysTraits<PH>::M, ND>::value),
                             (PatchMinWaves<PhysTraits<PH>::M, ND, AFFINE, SC, PREC>::value)) void rd_patch_kernel(PatchArgs a)
{
  extern __shared__ double lds[];
  const uint32_t pl = xcd_block_index();
  if (pl >= a.np) return;
  patch_body<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC>(a, pl, lds);
}

// Patches of several blocks (of several methods) of one element type /
// physics / scheme in one launch: workgroup g runs patch g - first[i] of
// args[i], first[i] <= g < first[i + 1]. The dynamic LDS is the largest of
// the entries'.
template <int D, int NG, int ND, int NF, int NQ, bool AFFINE, int PH, int SC, int PREC>
__global__ __launch_bounds__((PatchSize<PhysTraits<PH>::M, ND>::value),
                             (PatchMinWaves<PhysTraits<PH>::M, ND, AFFINE, SC, PREC>::value)) void
rd_patch_batch_kernel(const PatchArgs *__restrict__ args, const uint32_t *__restrict__ first, int count)
{
  extern __shared__ double lds[];
  const uint32_t g = xcd_block_index();
  if (g >= first[count]) return;
  int i = 0;
  while (i + 1 < count && first[i + 1] <= g) ++i;
  patch_body<D, NG, ND, NF, NQ, AFFINE, PH, SC, PREC>(args[i], g - first[i], lds);
}

struct FixupArgs
{
  const uint32_t *__restrict__ nodes;     // [nb] shared nodes
  const uint32_t *__restrict__ slot_off;  // [nb + 1]
  const double *__restrict__ slot_acc;    // rows of node i: [slot_off[i], slot_off[i + 1])
  PatchArgs pa;                           // node outputs as in the patch kernel
  uint32_t nb;
  double *__restrict__ norm_partial;      // [gridDim.x][M + 1]
};

// Sums of the shared nodes of one method: local block lb of nblk blocks
// (grid-stride over the nodes), norm partial of the block at lb.
template <int M>
__device__ __forceinline__ void fixup_body(const FixupArgs &f, uint32_t lb, uint32_t nblk, double *red)
{
  double sq[M + 1];
#pragma unroll
  for (int k = 0; k <= M; ++k) sq[k] = 0.0;
  for (uint32_t i = lb * blockDim.x + threadIdx.x; i < f.nb; i += nblk * blockDim.x) {
    double acc[M + 1];
    sum_rows<M>(f.slot_acc, f.slot_off[i], f.slot_off[i + 1], acc);
    const uint32_t g = f.nodes[i];
    double u_old[M];
#pragma unroll
    for (int k = 0; k < M; ++k) u_old[k] = f.pa.u[static_cast<size_t>(g) * M + k];
    finish_node<M>(f.pa, g, acc, u_old, f.pa.vol[g], f.pa.dir_index[g], sq);
  }
  if (f.pa.update && f.pa.outputs) block_sum_store<M>(sq, red, f.norm_partial + static_cast<size_t>(lb) * (M + 1));
}

template <int M>
__global__ __launch_bounds__(256) void patch_fixup_kernel(FixupArgs f)
{
  __shared__ double red[256 * (M + 1)];
  fixup_body<M>(f, blockIdx.x, gridDim.x, red);
}

// Shared-node sums of several methods in one launch: block g serves
// args[i] as its block g - first[i] of first[i + 1] - first[i].
template <int M>
__global__ __launch_bounds__(256) void patch_fixup_batch_kernel(const FixupArgs *__restrict__ args,
                                                                const uint32_t *__restrict__ first, int count)
{
  __shared__ double red[256 * (M + 1)];
  const uint32_t g = blockIdx.x;
  if (g >= first[count]) return;
  int i = 0;
  while (i + 1 < count && first[i + 1] <= g) ++i;
  fixup_body<M>(args[i], g - first[i], first[i + 1] - first[i], red);
}

}  // namespace pdrd
