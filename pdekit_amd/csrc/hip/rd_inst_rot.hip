// Cell-kernel instantiations for one physics (compiled as its own unit so the
// instantiations build in parallel).
#include "rd_launch.hpp"

namespace pdrd {

#define PHYS_ PH_ROTATION2D
KernelSet find_cell_scalar_rot(const CellKey &k)
{
  PDRD_ELEMENT_TYPES(PDRD_CELL_CASE)
  return KernelSet{};
}
#undef PHYS_

}  // namespace pdrd
