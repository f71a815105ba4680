// Registry of the compiled cell-kernel instantiations.
#pragma once

#include <hip/hip_runtime.h>

#include "rd_cell.hpp"

namespace pdrd {

using CellLauncher = hipError_t (*)(const CellArgs &, unsigned grid, unsigned block, hipStream_t);

struct CellKey
{
  int shape, sol_order, nq, physics, scheme;
};

// Element types built: (shape, P, quadrature points) with their sizes.
//   triangle P1 / P4 rule (6 points), P2 / P4 rule, P3 / P6 rule (12 points),
//   quadrilateral Q1 / P4 rule (8 points), tetrahedron P1 / P2 rule (4 points)
#define PDRD_ELEMENT_TYPES(X)                      \
  X(0, 1, 2, 3, 3, 6, 6, true)                     \
  X(0, 2, 2, 3, 6, 10, 6, true)                    \
  X(0, 3, 2, 3, 10, 15, 12, true)                  \
  X(1, 1, 2, 4, 4, 9, 8, false)                    \
  X(2, 1, 3, 4, 4, 10, 4, true)

template <int D, int NG, int ND, int NF, int NQ, bool AFF, int PH, int SC>
hipError_t launch_cell(const CellArgs &a, unsigned grid, unsigned block, hipStream_t s)
{
  hipLaunchKernelGGL((rd_cell_kernel<D, NG, ND, NF, NQ, AFF, PH, SC>), dim3(grid), dim3(block), 0, s, a);
  return hipGetLastError();
}

// Lookup functions, one per compilation unit of instantiations.
CellLauncher find_cell_scalar_adv(const CellKey &k);
CellLauncher find_cell_scalar_burgers(const CellKey &k);
CellLauncher find_cell_euler2d(const CellKey &k);
CellLauncher find_cell_euler3d(const CellKey &k);

inline CellLauncher find_cell(const CellKey &k)
{
  switch (k.physics) {
    case PH_ADVECTION2D: return find_cell_scalar_adv(k);
    case PH_BURGERS2D: return find_cell_scalar_burgers(k);
    case PH_EULER2D: return find_cell_euler2d(k);
    case PH_EULER3D: return find_cell_euler3d(k);
  }
  return nullptr;
}

}  // namespace pdrd

namespace pdrd {

template <int SH, int P, int D, int NG, int ND, int NF, int NQ, bool AFF, int PH>
CellLauncher pick_cell(const CellKey &k)
{
  if constexpr (D != PhysTraits<PH>::DIM) {
    (void)k;
    return nullptr;
  } else {
    if (k.shape == SH && k.sol_order == P && k.nq == NQ) {
      switch (k.scheme) {
        case SC_LDA: return &launch_cell<D, NG, ND, NF, NQ, AFF, PH, SC_LDA>;
        case SC_N: return &launch_cell<D, NG, ND, NF, NQ, AFF, PH, SC_N>;
        case SC_B: return &launch_cell<D, NG, ND, NF, NQ, AFF, PH, SC_B>;
        case SC_PSI: return &launch_cell<D, NG, ND, NF, NQ, AFF, PH, SC_PSI>;
      }
    }
    return nullptr;
  }
}

}  // namespace pdrd

// Body of a lookup function for one physics (PHYS_): every element type of
// matching dimension and every scheme.
#define PDRD_CELL_CASE(SH, P, D, NG, ND, NF, NQ, AFF) \
  if (CellLauncher f = pick_cell<SH, P, D, NG, ND, NF, NQ, AFF, PHYS_>(k)) return f;
