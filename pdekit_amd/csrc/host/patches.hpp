// Patch plan of the cell loop: cells grouped into spatially compact patches
// (runs of the Morton order of cell centroids), one GPU workgroup per patch.
#pragma once

#include <cstdint>
#include <vector>

namespace pdrd {

struct PatchBlockIn
{
  uint32_t ne = 0;
  int ng = 0, nd = 0;
  const uint32_t *geo = nullptr;  // [ne][ng] vertex ids
  const uint32_t *sol = nullptr;  // [ne][nd] node ids
  int patch_size = 256;
};

struct PatchBlockOut
{
  std::vector<uint32_t> order;  // position in the block -> caller cell index
  uint32_t first_patch = 0, np = 0;
  int maxl = 0;                 // most nodes of one patch
  std::vector<uint32_t> gconn;  // [ng][ne] SoA, block order
  std::vector<uint16_t> lconn;  // [nd][ne] SoA, patch-local node indices
};

struct PatchPlan
{
  std::vector<PatchBlockOut> blocks;
  std::vector<uint32_t> pnode_off;     // [np + 1]
  std::vector<uint32_t> pnodes;        // [slots] node id of each slot
  std::vector<uint32_t> pn_inner;      // [np] nodes held by this patch only (listed first)
  std::vector<uint16_t> pcol;          // [np][pcol_stride] colour offsets, ncol last
  std::vector<uint32_t> shared_nodes;  // nodes held by several patches, ascending
  std::vector<uint32_t> shared_off;    // [nshared + 1]
  std::vector<uint32_t> shared_slots;  // slots of each shared node, ascending
  int max_colours = 0;
};

// Builds the plan for all blocks of a mesh: the Morton key of every cell
// centroid (over the bounding box of all vertices), runs of patch_size cells,
// greedy colouring of the cells inside each patch (cells are then sorted by
// colour inside the patch), and the inner / shared split of every node.
// Throws std::invalid_argument when a patch would need more than max_colours
// colours or more than 65535 local nodes.
PatchPlan build_patch_plan(int dim, uint32_t nv, const double *coords, uint32_t nn,
                           const std::vector<PatchBlockIn> &blocks, int max_colours, int pcol_stride);

}  // namespace pdrd
