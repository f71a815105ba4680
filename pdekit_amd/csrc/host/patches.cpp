#include "patches.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace pdrd {

namespace {

// Spread the low 21 bits of x so that two zero bits separate them (3D), or
// the low 32 bits with one zero bit between them (2D).
uint64_t spread3(uint64_t x)
{
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

uint64_t spread2(uint64_t x)
{
  x &= 0xffffffffULL;
  x = (x | x << 16) & 0x0000ffff0000ffffULL;
  x = (x | x << 8) & 0x00ff00ff00ff00ffULL;
  x = (x | x << 4) & 0x0f0f0f0f0f0f0f0fULL;
  x = (x | x << 2) & 0x3333333333333333ULL;
  x = (x | x << 1) & 0x5555555555555555ULL;
  return x;
}

}  // namespace

PatchPlan build_patch_plan(int dim, uint32_t nv, const double *coords, uint32_t nn,
                           const std::vector<PatchBlockIn> &blocks, const std::vector<uint32_t> *extra_rows)
{
  auto extra = [&](uint32_t v) -> uint32_t { return extra_rows ? (*extra_rows)[v] : 0u; };
  PatchPlan plan;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (uint32_t v = 0; v < nv; ++v)
    for (int d = 0; d < dim; ++d) {
      lo[d] = std::min(lo[d], coords[static_cast<size_t>(v) * dim + d]);
      hi[d] = std::max(hi[d], coords[static_cast<size_t>(v) * dim + d]);
    }
  const double levels = (dim == 3) ? 2097151.0 : 4294967295.0;

  // 1. Morton order and patch runs of every block
  std::vector<std::vector<uint32_t>> patch_cells;  // per global patch: caller cell ids
  std::vector<int> patch_block;
  plan.blocks.resize(blocks.size());
  for (size_t b = 0; b < blocks.size(); ++b) {
    const PatchBlockIn &in = blocks[b];
    PatchBlockOut &out = plan.blocks[b];
    std::vector<uint64_t> key(in.ne);
    for (uint32_t c = 0; c < in.ne; ++c) {
      double cen[3] = {0, 0, 0};
      for (int g = 0; g < in.ng; ++g)
        for (int d = 0; d < dim; ++d) cen[d] += coords[static_cast<size_t>(in.geo[static_cast<size_t>(c) * in.ng + g]) * dim + d];
      uint64_t k = 0;
      for (int d = 0; d < dim; ++d) {
        const double t = (hi[d] > lo[d]) ? (cen[d] / in.ng - lo[d]) / (hi[d] - lo[d]) : 0.0;
        const uint64_t qd = static_cast<uint64_t>(std::min(std::max(t, 0.0), 1.0) * levels);
        k |= (dim == 3 ? spread3(qd) : spread2(qd)) << d;
      }
      key[c] = k;
    }
    std::vector<uint32_t> ord(in.ne);
    std::iota(ord.begin(), ord.end(), 0u);
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t c) { return key[a] < key[c]; });
    out.first_patch = static_cast<uint32_t>(patch_cells.size());
    for (uint32_t s = 0; s < in.ne; s += in.patch_size) {
      const uint32_t e = std::min<uint32_t>(in.ne, s + in.patch_size);
      patch_cells.emplace_back(ord.begin() + s, ord.begin() + e);
      patch_block.push_back(static_cast<int>(b));
    }
    out.np = static_cast<uint32_t>(patch_cells.size()) - out.first_patch;
  }
  const size_t np = patch_cells.size();

  // 2. nodes of every patch and how many patches hold each node
  std::vector<std::vector<uint32_t>> pn(np);
  std::vector<uint32_t> cnt(nn, 0);
  for (size_t p = 0; p < np; ++p) {
    const PatchBlockIn &in = blocks[patch_block[p]];
    std::vector<uint32_t> &L = pn[p];
    for (uint32_t c : patch_cells[p])
      for (int k = 0; k < in.nd; ++k) L.push_back(in.sol[static_cast<size_t>(c) * in.nd + k]);
    std::sort(L.begin(), L.end());
    L.erase(std::unique(L.begin(), L.end()), L.end());
    if (L.size() > static_cast<size_t>(PATCH_MAX_NODES)) throw std::invalid_argument("patch with too many nodes");
    if (patch_cells[p].size() * in.nd > 65535) throw std::invalid_argument("patch with too many (cell, node) pairs");
    for (uint32_t v : L) cnt[v]++;
  }

  // 3. inner nodes first, then shared; slots; local connectivity with ranks;
  //    node-grouped pair offsets
  plan.pnode_off.assign(np + 1, 0);
  plan.pn_inner.assign(np, 0);
  std::vector<int32_t> pos(nn, -1);
  for (size_t p = 0; p < np; ++p) {
    std::vector<uint32_t> inner, shared;
    for (uint32_t v : pn[p]) (cnt[v] == 1 && extra(v) == 0 ? inner : shared).push_back(v);
    plan.pn_inner[p] = static_cast<uint32_t>(inner.size());
    pn[p] = inner;
    pn[p].insert(pn[p].end(), shared.begin(), shared.end());
    plan.pnode_off[p + 1] = plan.pnode_off[p] + static_cast<uint32_t>(pn[p].size());
  }
  plan.pnodes.reserve(plan.pnode_off[np]);
  for (size_t p = 0; p < np; ++p) plan.pnodes.insert(plan.pnodes.end(), pn[p].begin(), pn[p].end());
  plan.pn_coff.assign(plan.pnode_off[np], 0);

  for (size_t b = 0; b < blocks.size(); ++b) {
    const PatchBlockIn &in = blocks[b];
    PatchBlockOut &out = plan.blocks[b];
    out.order.clear();
    out.order.reserve(in.ne);
    out.gconn.assign(static_cast<size_t>(in.ng) * in.ne, 0);
    out.lconn.assign(static_cast<size_t>(in.nd) * in.ne, 0);
    for (uint32_t pl = 0; pl < out.np; ++pl) {
      const size_t p = out.first_patch + pl;
      const std::vector<uint32_t> &L = pn[p];
      out.maxl = std::max(out.maxl, static_cast<int>(L.size()));
      for (size_t l = 0; l < L.size(); ++l) pos[L[l]] = static_cast<int32_t>(l);
      std::vector<int> deg(L.size(), 0);
      for (uint32_t c : patch_cells[p]) {
        const size_t e = out.order.size();
        out.order.push_back(c);
        for (int g = 0; g < in.ng; ++g) out.gconn[static_cast<size_t>(g) * in.ne + e] = in.geo[static_cast<size_t>(c) * in.ng + g];
        for (int k = 0; k < in.nd; ++k) {
          const int l = pos[in.sol[static_cast<size_t>(c) * in.nd + k]];
          const int r = deg[l]++;
          if (r >= PATCH_MAX_RANK) throw std::invalid_argument("node held by too many cells of one patch");
          plan.max_rank = std::max(plan.max_rank, r + 1);
          out.lconn[static_cast<size_t>(k) * in.ne + e] = static_cast<uint16_t>(l | (r << LCONN_BITS));
        }
      }
      uint32_t off = 0;
      for (size_t l = 0; l < L.size(); ++l) {
        plan.pn_coff[plan.pnode_off[p] + l] = static_cast<uint16_t>(off);
        off += static_cast<uint32_t>(deg[l]);
      }
      for (uint32_t v : L) pos[v] = -1;
    }
  }

  // 4. slots of every shared node, in patch order
  std::vector<uint32_t> shared_id(nn, UINT32_MAX);
  for (uint32_t v = 0; v < nn; ++v)
    if (cnt[v] > 1 || (cnt[v] == 1 && extra(v) > 0)) {
      shared_id[v] = static_cast<uint32_t>(plan.shared_nodes.size());
      plan.shared_nodes.push_back(v);
    }
  plan.shared_off.assign(plan.shared_nodes.size() + 1, 0);
  for (uint32_t v : plan.shared_nodes) plan.shared_off[shared_id[v] + 1] = cnt[v] + extra(v);
  for (size_t i = 0; i < plan.shared_nodes.size(); ++i) plan.shared_off[i + 1] += plan.shared_off[i];
  plan.shared_slots.assign(plan.shared_off.back(), UINT32_MAX);
  std::vector<uint32_t> fillp(plan.shared_off.begin(), plan.shared_off.end() - 1);
  for (size_t p = 0; p < np; ++p)
    for (uint32_t l = plan.pn_inner[p]; l < plan.pnode_off[p + 1] - plan.pnode_off[p]; ++l) {
      const uint32_t v = plan.pnodes[plan.pnode_off[p] + l];
      plan.shared_slots[fillp[shared_id[v]]++] = plan.pnode_off[p] + l;
    }
  return plan;
}

}  // namespace pdrd
