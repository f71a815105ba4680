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
  // [nd][ne] SoA: patch-local node index l (low LCONN_BITS bits) and the rank
  // of this (cell, node) pair among the pairs of node l (high bits); pairs of
  // one node are ranked by cell position in the patch.
  std::vector<uint16_t> lconn;
};

constexpr int LCONN_BITS = 10;
constexpr int PATCH_MAX_NODES = 1 << LCONN_BITS;
constexpr int PATCH_MAX_RANK = 1 << (16 - LCONN_BITS);

struct PatchPlan
{
  std::vector<PatchBlockOut> blocks;
  std::vector<uint32_t> pnode_off;     // [np + 1]
  std::vector<uint32_t> pnodes;        // [slots] node id of each slot
  std::vector<uint32_t> pn_inner;      // [np] nodes held by this patch only (listed first)
  // [slots] first position of the slot's node in the patch's list of
  // (cell, node) pairs grouped by node: pairs of local node l occupy
  // [pn_coff[l], pn_coff[l + 1]) (the last node ends at cells * nd)
  std::vector<uint16_t> pn_coff;
  std::vector<uint32_t> shared_nodes;  // nodes held by several patches, ascending
  std::vector<uint32_t> shared_off;    // [nshared + 1]
  std::vector<uint32_t> shared_slots;  // slots of each shared node, ascending (then extra rows)
  int max_rank = 0;                    // most cells of one patch sharing one node
};

// Builds the plan for all blocks of a mesh: the Morton key of every cell
// centroid (over the bounding box of all vertices), runs of patch_size cells,
// the inner / shared split of every node and the node-grouped order of the
// (cell, node) pairs of each patch. Throws std::invalid_argument when a patch
// has more than PATCH_MAX_NODES nodes, a node is held by more than
// PATCH_MAX_RANK cells of one patch, or a patch has 65536 pairs or more.
// extra_rows (optional, per node): further partial sums a node receives
// besides its patches' (weak boundary facets); such nodes are treated as
// shared, and their extra rows follow the patch rows in shared_slots as
// UINT32_MAX entries.
PatchPlan build_patch_plan(int dim, uint32_t nv, const double *coords, uint32_t nn,
                           const std::vector<PatchBlockIn> &blocks,
                           const std::vector<uint32_t> *extra_rows = nullptr);

}  // namespace pdrd
