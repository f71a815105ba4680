// Registry of the compiled cell-kernel instantiations.
#pragma once

#include <hip/hip_runtime.h>

#include "rd_cell.hpp"
#include "rd_jacobian.hpp"

namespace pdrd {

using CellLauncher = hipError_t (*)(const CellArgs &, unsigned grid, unsigned block, hipStream_t);
using PatchLauncher = hipError_t (*)(const PatchArgs &, unsigned grid, int cells, hipStream_t);
using PatchBatchLauncher = hipError_t (*)(const PatchArgs *args, const uint32_t *first, int count, unsigned grid,
                                          int cells, size_t lds_bytes, hipStream_t);
using JacobianLauncher = hipError_t (*)(const JacArgs &, unsigned grid, hipStream_t);

struct CellKey
{
  int shape, sol_order, nq, physics, scheme;
};

// Everything the host needs to launch one instantiation.
struct KernelSet
{
  CellLauncher cell = nullptr;    // element per thread, global scatter
  PatchLauncher patch = nullptr;  // patch assembly
  PatchBatchLauncher patch_batch = nullptr;  // patch assembly of several blocks / methods
  // the same with a nodal input (LDA / B) or a source field
  PatchBatchLauncher patch_nodal_batch = nullptr;
  PatchBatchLauncher patch_src_batch = nullptr;
  PatchLauncher patch_lin = nullptr;  // linear sweep of the mixed-precision form: J delta per cell
  // with the nodal input of the scheme (LDA: artificial viscosity, B:
  // external blending coefficient); null for N and PSI, which take none
  CellLauncher cell_nodal = nullptr;
  PatchLauncher patch_nodal = nullptr;
  // FP64 patch assembly with a nodal source field (reaction residuals), and
  // for the B scheme with an external blending coefficient beside it
  PatchLauncher patch_src = nullptr;
  PatchLauncher patch_src_theta = nullptr;
  PatchBatchLauncher patch_src_theta_batch = nullptr;
  // element residuals and Jacobians of the implicit form (null for PSI)
  JacobianLauncher jacobian = nullptr;
  int patch_size = 0;             // largest (default) cells per patch workgroup
  size_t (*patch_lds_bytes)(int maxl, int cells) = nullptr;
  size_t (*patch_src_lds_bytes)(int maxl, int cells) = nullptr;
  size_t (*patch_nodal_lds_bytes)(int maxl, int cells) = nullptr;
  size_t (*patch_src_theta_lds_bytes)(int maxl, int cells) = nullptr;
  size_t (*patch_lin_lds_bytes)(int maxl, int cells) = nullptr;
  explicit operator bool() const { return cell != nullptr; }
};

// Element types built: (shape, P, dim, ng, nd, nf, quadrature points, affine)
//   triangle P1 / P4 rule (6 points), P2 / P4 rule, P3 / P6 rule (12 points),
//   quadrilateral Q1 and Q2 / P4 rule (8 points), tetrahedron P1 / P2 rule
//   (4 points), tetrahedron P2 / P2 rule and P4 rule (11 points)
#define PDRD_ELEMENT_TYPES(X)                      \
  X(0, 1, 2, 3, 3, 6, 6, true)                     \
  X(0, 2, 2, 3, 6, 10, 6, true)                    \
  X(0, 3, 2, 3, 10, 15, 12, true)                  \
  X(1, 1, 2, 4, 4, 9, 8, false)                    \
  X(1, 2, 2, 4, 9, 16, 8, false)                   \
  X(2, 1, 3, 4, 4, 10, 4, true)                    \
  X(2, 2, 3, 4, 10, 20, 4, true)                   \
  X(2, 2, 3, 4, 10, 20, 11, true)

template <int D, int NG, int ND, int NF, int NQ, bool AFF, int PH, int SC, int PREC = 0>
hipError_t launch_cell(const CellArgs &a, unsigned grid, unsigned block, hipStream_t s)
{
  constexpr int M = PhysTraits<PH>::M;
  const size_t lds = static_cast<size_t>(NQ) * M * block * sizeof(double);
  hipLaunchKernelGGL((rd_cell_kernel<D, NG, ND, NF, NQ, AFF, PH, SC, PREC>), dim3(grid), dim3(block), lds, s, a);
  return hipGetLastError();
}

template <int D, int NG, int ND, int NF, int NQ, bool AFF, int PH, int SC, int PREC>
hipError_t launch_patch(const PatchArgs &a, unsigned grid, int cells, hipStream_t s)
{
  constexpr int M = PhysTraits<PH>::M;
  const size_t lds = patch_lds_doubles_prec<NQ, ND, M, PREC>(a.maxl, cells) * sizeof(double);
  hipLaunchKernelGGL((rd_patch_kernel<D, NG, ND, NF, NQ, AFF, PH, SC, PREC>), dim3(grid), dim3(cells), lds, s, a);
  return hipGetLastError();
}

template <int D, int NG, int ND, int NF, int NQ, bool AFF, int PH, int SC, int PREC>
hipError_t launch_patch_batch(const PatchArgs *args, const uint32_t *first, int count, unsigned grid, int cells,
                              size_t lds, hipStream_t s)
{
  hipLaunchKernelGGL((rd_patch_batch_kernel<D, NG, ND, NF, NQ, AFF, PH, SC, PREC>), dim3(grid), dim3(cells), lds, s,
                     args, first, count);
  return hipGetLastError();
}

template <int D, int NG, int ND, int NF, int NQ, bool AFF, int PH, int SC>
hipError_t launch_jacobian(const JacArgs &a, unsigned grid, hipStream_t s)
{
  if constexpr (PhysTraits<PH>::M > 1) {
    // the Euler cells: LPC lanes per cell, one wave per workgroup, a slab of
    // QpRows rows (grid from the cell count: `grid` counts workgroups of 64
    // cells)
    constexpr int CPB = 64 / QpLanes<NQ>::value;
    constexpr size_t lds = static_cast<size_t>(QP_XS) * QpRows<PhysTraits<PH>::M, ND>::value * sizeof(double);
    (void)grid;
    hipLaunchKernelGGL((rd_jacobian_qp_kernel<D, NG, ND, NF, NQ, AFF, PH, SC>), dim3((a.ne + CPB - 1) / CPB), dim3(64),
                       lds, s, a);
  }
  else
    hipLaunchKernelGGL((rd_jacobian_kernel<D, NG, ND, NF, NQ, AFF, PH, SC>), dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <int NQ, int ND, int M>
size_t patch_lds_bytes(int maxl, int cells)
{
  return patch_lds_doubles<NQ, ND, M>(maxl, cells) * sizeof(double);
}

template <int NQ, int ND, int M>
size_t patch_src_lds_bytes(int maxl, int cells)
{
  return patch_lds_doubles_prec<NQ, ND, M, 3>(maxl, cells) * sizeof(double);
}

template <int NQ, int ND, int M>
size_t patch_nodal_lds_bytes(int maxl, int cells)
{
  return patch_lds_doubles_prec<NQ, ND, M, 2>(maxl, cells) * sizeof(double);
}

template <int NQ, int ND, int M>
size_t patch_src_theta_lds_bytes(int maxl, int cells)
{
  return patch_lds_doubles_prec<NQ, ND, M, 5>(maxl, cells) * sizeof(double);
}

template <int NQ, int ND, int M>
size_t patch_lin_lds_bytes(int maxl, int cells)
{
  return patch_lds_doubles_prec<NQ, ND, M, 4>(maxl, cells) * sizeof(double);
}

template <int D, int NG, int ND, int NF, int NQ, bool AFF, int PH, int SC>
KernelSet kernel_set()
{
  KernelSet k;
  k.cell = &launch_cell<D, NG, ND, NF, NQ, AFF, PH, SC>;
  k.patch = &launch_patch<D, NG, ND, NF, NQ, AFF, PH, SC, 0>;
  k.patch_batch = &launch_patch_batch<D, NG, ND, NF, NQ, AFF, PH, SC, 0>;
  // the linear sweep depends on the block size only: one instantiation per
  // element type and equation count (that of the advection physics for the
  // scalar ones)
  if constexpr ((PH == PH_ADVECTION2D || PH == PH_EULER2D || PH == PH_EULER3D) && SC == SC_LDA) {
    k.patch_lin = &launch_patch<D, NG, ND, NF, NQ, AFF, PH, SC, 4>;
    k.patch_lin_lds_bytes = &patch_lin_lds_bytes<NQ, ND, PhysTraits<PH>::M>;
  }
  if constexpr (SC == SC_LDA || SC == SC_B) {
    k.cell_nodal = &launch_cell<D, NG, ND, NF, NQ, AFF, PH, SC, 2>;
    k.patch_nodal = &launch_patch<D, NG, ND, NF, NQ, AFF, PH, SC, 2>;
    k.patch_nodal_batch = &launch_patch_batch<D, NG, ND, NF, NQ, AFF, PH, SC, 2>;
    k.patch_nodal_lds_bytes = &patch_nodal_lds_bytes<NQ, ND, PhysTraits<PH>::M>;
  }
  k.patch_src = &launch_patch<D, NG, ND, NF, NQ, AFF, PH, SC, 3>;
  k.patch_src_batch = &launch_patch_batch<D, NG, ND, NF, NQ, AFF, PH, SC, 3>;
  if constexpr (SC == SC_B) {
    k.patch_src_theta = &launch_patch<D, NG, ND, NF, NQ, AFF, PH, SC, 5>;
    k.patch_src_theta_batch = &launch_patch_batch<D, NG, ND, NF, NQ, AFF, PH, SC, 5>;
    k.patch_src_theta_lds_bytes = &patch_src_theta_lds_bytes<NQ, ND, PhysTraits<PH>::M>;
  }
  if constexpr (SC != SC_PSI && SC != SC_BPSI) k.jacobian = &launch_jacobian<D, NG, ND, NF, NQ, AFF, PH, SC>;
  k.patch_size = PatchSize<PhysTraits<PH>::M, ND>::value;
  k.patch_lds_bytes = &patch_lds_bytes<NQ, ND, PhysTraits<PH>::M>;
  k.patch_src_lds_bytes = &patch_src_lds_bytes<NQ, ND, PhysTraits<PH>::M>;
  return k;
}

// Lookup functions, one per compilation unit of instantiations.
KernelSet find_cell_scalar_adv(const CellKey &k);
KernelSet find_cell_scalar_burgers(const CellKey &k);
KernelSet find_cell_scalar_rot(const CellKey &k);
KernelSet find_cell_euler2d(const CellKey &k);
KernelSet find_cell_euler3d(const CellKey &k);

inline KernelSet find_cell(const CellKey &k)
{
  switch (k.physics) {
    case PH_ADVECTION2D: return find_cell_scalar_adv(k);
    case PH_BURGERS2D: return find_cell_scalar_burgers(k);
    case PH_ROTATION2D: return find_cell_scalar_rot(k);
    case PH_EULER2D: return find_cell_euler2d(k);
    case PH_EULER3D: return find_cell_euler3d(k);
  }
  return KernelSet{};
}

template <int SH, int P, int D, int NG, int ND, int NF, int NQ, bool AFF, int PH>
KernelSet pick_cell(const CellKey &k)
{
  if constexpr (D != PhysTraits<PH>::DIM) {
    (void)k;
    return KernelSet{};
  } else {
    if (k.shape == SH && k.sol_order == P && k.nq == NQ) {
      switch (k.scheme) {
        case SC_LDA: return kernel_set<D, NG, ND, NF, NQ, AFF, PH, SC_LDA>();
        case SC_N: return kernel_set<D, NG, ND, NF, NQ, AFF, PH, SC_N>();
        case SC_B: return kernel_set<D, NG, ND, NF, NQ, AFF, PH, SC_B>();
        case SC_PSI: return kernel_set<D, NG, ND, NF, NQ, AFF, PH, SC_PSI>();
        case SC_BPSI: return kernel_set<D, NG, ND, NF, NQ, AFF, PH, SC_BPSI>();
      }
    }
    return KernelSet{};
  }
}

}  // namespace pdrd

// Body of a lookup function for one physics (PHYS_): every element type of
// matching dimension and every scheme.
#define PDRD_CELL_CASE(SH, P, D, NG, ND, NF, NQ, AFF) \
  if (KernelSet f = pick_cell<SH, P, D, NG, ND, NF, NQ, AFF, PHYS_>(k)) return f;
