// The explicit continuous RD method on one device: the C ABI entry points
// declared in include/pdekit_rd.h that own device state, the nodal passes
// (dual-volume scaling, time step, update, norms) and HIP-graph replay of the
// pseudo-time iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/pdekit_rd.h"
#include "../host/capi_common.hpp"
#include "../host/facets.hpp"
#include "../host/fe_tables.hpp"
#include "../host/mesh.hpp"
#include "../host/patches.hpp"
#include "rd_launch.hpp"
#include "rd_stab.hpp"
#include "rd_weakbc.hpp"

using pdrd::set_error;

namespace {

constexpr int kNormBlock = 256;
constexpr int kNormMaxBlocks = 1024;

struct Block
{
  pdrd_fe_desc fe;
  uint32_t ne = 0;
  std::vector<double> tables;
  std::vector<uint32_t> geo, sol;      // host copies, [ne][ng] / [ne][nd] (caller order)
  std::vector<uint32_t> perm;          // block position -> caller cell index
  std::vector<uint32_t> colour_start;  // [ncolours + 1]
  double *d_tab = nullptr;
  uint32_t *d_gconn = nullptr;  // [ng][ne] in colour-sorted order
  uint32_t *d_sconn = nullptr;  // [nd][ne]
  pdrd::KernelSet ks;
  // patch assembly (Morton-ordered patches)
  uint32_t *d_pgconn = nullptr;  // [ng][ne] in patch order
  uint16_t *d_lconn = nullptr;   // [nd][ne] patch-local node indices
  uint32_t first_patch = 0, np = 0;
  int maxl = 0;
  int cells = 0;  // cells per patch (threads per workgroup)
  // providers of the nodal inputs: vertex functions at the nodes [nd][nvert];
  // smoothness-indicator rule w, V, dVg, psi packed
  double *d_p1w = nullptr, *d_av = nullptr;
  int av_nq = 0, av_npm = 0;
  // mixed precision: patch position of every cell (caller index -> position),
  // the element Jacobians in FP32 entry-major over the patch order
  // [(nd m)^2][ne], the B stabilisation scratch of the Jacobian kernel in the
  // same layout, the colour position -> patch position map, and the linear
  // sweep's kernel (the scalar / Euler instantiation of this element type)
  std::vector<uint32_t> ppos;
  float *d_jlin = nullptr, *d_jstab_lin = nullptr;
  uint32_t *d_lin_pos = nullptr;
  pdrd::PatchLauncher lin = nullptr;
};

// A weak boundary condition as the caller gives it, and on the device.
struct WeakSpec
{
  int type = 0;
  std::vector<uint32_t> nodes;
  double params[5] = {0, 0, 0, 0, 0};
  int quad_order = 4;
};

struct WeakDev
{
  int type = 0, nfn = 0, nq = 0;
  uint32_t nf = 0;
  double relax = 1.0, params[5] = {0, 0, 0, 0, 0};
  uint32_t *d_nodes = nullptr, *d_rows = nullptr;
  double *d_normal = nullptr, *d_jac = nullptr, *d_w = nullptr, *d_V = nullptr;
  // WeakDirichlet state at every facet point (set by the caller from an
  // expression of the point), or null; the facet points' coordinates
  double *d_qstate = nullptr;
  std::vector<double> qpoints;  // [nf][nq][dim]
  double *d_qvel = nullptr;     // rotating advection: (y, -x) at every facet point
};

}  // namespace

struct pdrd_method
{
  int dim = 2, physics = 0, scheme = 0, m = 1;
  uint32_t nv = 0, nn = 0;
  double params[4] = {0, 0, 0, 0};
  std::vector<double> coords;
  std::vector<Block> blocks;
  bool ready = false;
  int device = 0;

  // options
  int scatter_mode = 2;  // 0 atomics, 1 colours, 2 patches
  bool patch_ok = false;
  std::string patch_reason;
  int precision = 0;
  std::vector<WeakSpec> weak_specs;  // weak BCs added before pdrd_initialize
  std::vector<WeakDev> weak;
  // mixed precision (linearised defect correction). An FP64 sweep sets the
  // linearisation point: u_ref = u, its FP64 residual in d_res and time step
  // in d_dt; every `refresh` iterations it also recomputes the element
  // Jacobians (FP32, per block). The sweeps in between are linear:
  // res = R(u_ref) + sum_e J_e (u - u_ref)_e / V, dt = dt(u_ref). Its fixed
  // point is the FP64 steady state; the FP32 rounding scales with u - u_ref.
  int refresh = 50;
  int since_refresh = 1 << 30;  // iterations since the Jacobians were computed
  bool lin_valid = false;       // u_ref, d_res, d_dt and the Jacobians form a linearisation
  double *d_uref = nullptr;
  // with weak BCs: R(u_ref) without its facet part (the linear sweeps run the
  // facet kernels at the current state), and per shared node its number of
  // facet rows
  double *d_linres = nullptr;
  uint32_t *d_shared_extra = nullptr;
  // nodal inputs of the cell loop: the external blending coefficient of the
  // B scheme (set_blending_coeff) and the artificial viscosity of the LDA
  // scheme (set_artificial_viscosity), device [N] each, and a scratch [N]
  // for their providers
  double *d_blend = nullptr, *d_visc = nullptr, *d_stab_tmp = nullptr;
  bool use_blend = false, use_visc = false;
  // nodal source field [N][m] (set_vec_function(sources)): the cells then
  // compute the reaction residuals
  double *d_src = nullptr;
  bool use_src = false;
  int patch_cells = 0;  // requested cells per patch (0: each kernel's default)
  int block_threads = 256;
  int timing = 0;

  // device state
  double *d_coords = nullptr, *d_u = nullptr, *d_res = nullptr, *d_dt = nullptr, *d_ws = nullptr;
  double *d_vol = nullptr, *d_acc_res = nullptr, *d_acc_ws = nullptr, *d_partial = nullptr;
  int *d_status = nullptr;
  uint32_t *d_dir_nodes = nullptr;
  double *d_dir_vals = nullptr;
  unsigned char *d_is_dir = nullptr;
  int *d_dir_index = nullptr;  // [N] row of the node's Dirichlet values or -1
  uint32_t n_dir = 0;
  std::vector<uint32_t> dir_nodes_host;
  std::vector<double> dir_vals_host;
  // patch plan
  uint32_t np_total = 0, n_shared = 0;
  uint32_t *d_pnode_off = nullptr, *d_pnodes = nullptr, *d_pn_inner = nullptr;
  uint16_t *d_pn_coff = nullptr;
  uint32_t *d_shared_nodes = nullptr, *d_shared_off = nullptr, *d_slot_pos = nullptr, *d_pn_sid = nullptr;
  double *d_slot_acc = nullptr, *d_patch_partial = nullptr;
  // second row buffer and the shared nodes' states [n_shared][m] of the two
  // parities of a chain of sweeps whose shared-node sums are deferred
  double *d_slot_acc_b = nullptr, *d_base[2] = {nullptr, nullptr};
  uint2 *d_pn_rows = nullptr;  // [slots] rows of a shared slot's node
  int chain = -1;              // PDRD_OPT_CHAIN (-1: the default of the method's kernels)
  size_t bytes = 0;

  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double last_cell_ms = 0.0, last_node_ms = 0.0;

  // cached graph of one iteration without outputs (atomic / colour
  // assembly)
  hipGraphExec_t graph = nullptr;
  double graph_cfl = -1.0;
};

namespace {

// ---- nodal kernels -------------------------------------------------------------

// residual = acc / V, wave speed = acc_ws; the accumulators are cleared for
// the next cell loop.
__global__ void finalize_kernel(uint32_t nn, int m, double *__restrict__ acc_res, double *__restrict__ acc_ws,
                                const double *__restrict__ vol, double *__restrict__ res, double *__restrict__ ws)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const double iv = 1.0 / vol[i];
  for (int k = 0; k < m; ++k) {
    const size_t j = static_cast<size_t>(i) * m + k;
    res[j] = acc_res[j] * iv;
    acc_res[j] = 0.0;
  }
  ws[i] = acc_ws[i];
  acc_ws[i] = 0.0;
}

// Residual of the linear sweeps at the linearisation point without its
// weak-facet part (the facet kernels run in the linear sweeps too, at the
// current state): for every shared node with extra facet rows -- the last
// `extra[i]` of its rows [off[i], off[i + 1]) -- linres = res - sum / V.
// linres holds a copy of res on entry.
__global__ void lin_res_kernel(uint32_t n, int m, const uint32_t *__restrict__ nodes, const uint32_t *__restrict__ off,
                               const uint32_t *__restrict__ extra, const double *__restrict__ rows,
                               const double *__restrict__ vol, double *__restrict__ linres)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || extra[i] == 0) return;
  const uint32_t g = nodes[i];
  const double iv = 1.0 / vol[g];
  for (int k = 0; k < m; ++k) {
    double sum = 0.0;
    for (uint32_t r = off[i + 1] - extra[i]; r < off[i + 1]; ++r) sum += rows[static_cast<size_t>(r) * (m + 1) + k];
    linres[static_cast<size_t>(g) * m + k] -= sum * iv;
  }
}

__global__ void dirichlet_kernel(uint32_t n, int m, const uint32_t *__restrict__ nodes,
                                 const double *__restrict__ vals, double *__restrict__ u)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t b = static_cast<size_t>(nodes[i]) * m;
  for (int k = 0; k < m; ++k) u[b + k] = vals[static_cast<size_t>(i) * m + k];
}

__global__ void dt_kernel(uint32_t nn, double cfl, const double *__restrict__ vol, const double *__restrict__ ws,
                          double *__restrict__ dt)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  dt[i] = cfl * vol[i] / ws[i];
}

__global__ void update_kernel(uint32_t nn, int m, const double *__restrict__ dt, const double *__restrict__ res,
                              double *__restrict__ u)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  for (int k = 0; k < m; ++k) {
    const size_t j = static_cast<size_t>(i) * m + k;
    u[j] -= dt[i] * res[j];
  }
}

// Fused nodal pass of one pseudo-time iteration (after the cell loop and the
// strong Dirichlet values): res = acc / V, dt = CFL V / sum lambda,
// u -= dt res, accumulators cleared. When `outputs` is set, res / dt / ws are
// stored and per-block partial sums of res^2 (field-wise, optionally without
// Dirichlet rows) and the count of summed rows are written to `partial`.
template <int M>
__global__ __launch_bounds__(kNormBlock) void node_pass_kernel(uint32_t nn, double cfl, double *__restrict__ acc_res,
                                                               double *__restrict__ acc_ws,
                                                               const double *__restrict__ vol, double *__restrict__ u,
                                                               double *__restrict__ res, double *__restrict__ dt,
                                                               double *__restrict__ ws, int outputs, int excl,
                                                               const unsigned char *__restrict__ is_dir,
                                                               double *__restrict__ partial)
{
  double sq[M + 1];
#pragma unroll
  for (int k = 0; k <= M; ++k) sq[k] = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
    const double v = vol[i];
    const double w = acc_ws[i];
    const double iv = 1.0 / v;
    const double d = cfl * v / w;
    const bool count = outputs && !(excl && is_dir[i]);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const size_t j = static_cast<size_t>(i) * M + k;
      const double r = acc_res[j] * iv;
      acc_res[j] = 0.0;
      u[j] -= d * r;
      if (outputs) res[j] = r;
      if (count) sq[k] += r * r;
    }
    if (count) sq[M] += 1.0;
    acc_ws[i] = 0.0;
    if (outputs) {
      dt[i] = d;
      ws[i] = w;
    }
  }
  if (!outputs) return;
  __shared__ double red[kNormBlock][M + 1];
#pragma unroll
  for (int k = 0; k <= M; ++k) red[threadIdx.x][k] = sq[k];
  __syncthreads();
  for (int s = kNormBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k <= M; ++k) red[threadIdx.x][k] += red[threadIdx.x + s][k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k <= M; ++k) partial[blockIdx.x * (M + 1) + k] = red[0][k];
  }
}

// Per-block partial sums of (res * scale)^2 per field (scale = dt or 1) and
// the number of rows summed.
template <int M>
__global__ __launch_bounds__(kNormBlock) void norm_partial_kernel(uint32_t nn, const double *__restrict__ res,
                                                                  const double *__restrict__ scale, int excl,
                                                                  const unsigned char *__restrict__ is_dir,
                                                                  double *__restrict__ partial)
{
  double sq[M + 1];
#pragma unroll
  for (int k = 0; k <= M; ++k) sq[k] = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
    if (excl && is_dir[i]) continue;
    const double s = scale ? scale[i] : 1.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double r = res[static_cast<size_t>(i) * M + k] * s;
      sq[k] += r * r;
    }
    sq[M] += 1.0;
  }
  __shared__ double red[kNormBlock][M + 1];
#pragma unroll
  for (int k = 0; k <= M; ++k) red[threadIdx.x][k] = sq[k];
  __syncthreads();
  for (int s = kNormBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k <= M; ++k) red[threadIdx.x][k] += red[threadIdx.x + s][k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k <= M; ++k) partial[blockIdx.x * (M + 1) + k] = red[0][k];
  }
}

// ---- helpers ----------------------------------------------------------------------

int hip_fail(hipError_t e, const char *what)
{
  return set_error(PDRD_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_CHECK(call)                                       \
  do {                                                        \
    hipError_t _e = (call);                                   \
    if (_e != hipSuccess) return hip_fail(_e, #call);         \
  } while (0)

template <typename T>
int dev_alloc(pdrd_method *h, T **p, size_t count)
{
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
  if (e != hipSuccess) return set_error(PDRD_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  h->bytes += count * sizeof(T);
  return PDRD_OK;
}

unsigned grid_for(uint32_t n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

int field_size(const pdrd_method *h, int field, size_t *count, double **dptr)
{
  switch (field) {
    case PDRD_SOLUTION: *count = static_cast<size_t>(h->nn) * h->m; *dptr = h->d_u; return PDRD_OK;
    case PDRD_RESIDUAL: *count = static_cast<size_t>(h->nn) * h->m; *dptr = h->d_res; return PDRD_OK;
    case PDRD_TIME_STEP: *count = h->nn; *dptr = h->d_dt; return PDRD_OK;
    case PDRD_WAVE_SPEED: *count = h->nn; *dptr = h->d_ws; return PDRD_OK;
    case PDRD_DUAL_VOLUME: *count = h->nn; *dptr = h->d_vol; return PDRD_OK;
    case PDRD_BLEND_COEFF:
      if (!h->d_blend) return set_error(PDRD_EINVAL, "no blending coefficient has been set or computed");
      *count = h->nn;
      *dptr = h->d_blend;
      return PDRD_OK;
    case PDRD_ART_VISC:
      if (!h->d_visc) return set_error(PDRD_EINVAL, "no artificial viscosity has been set or computed");
      *count = h->nn;
      *dptr = h->d_visc;
      return PDRD_OK;
  }
  return set_error(PDRD_EINVAL, "unknown field");
}

// The nodal input the cells of this method take, or null: the blending
// coefficient for the B scheme, the artificial viscosity for the LDA scheme.
const double *nodal_input(const pdrd_method *h)
{
  if (h->scheme == pdrd::SC_B && h->use_blend) return h->d_blend;
  if (h->scheme == pdrd::SC_LDA && h->use_visc) return h->d_visc;
  return nullptr;
}

// With a source field: the B scheme's external blending coefficient, which
// its reaction residuals use (PGB.hpp:825-1050, the skip test at :878), or
// null. The LDA scheme's reaction residuals apply no artificial viscosity
// (PGLDA.hpp:553-741), so a viscosity set beside a source field is not used.
const double *source_theta(const pdrd_method *h)
{
  return (h->use_src && h->scheme == pdrd::SC_B && h->use_blend) ? h->d_blend : nullptr;
}

// The patch kernel and its LDS bytes for a block with this method's inputs.
pdrd::PatchLauncher input_kernel(const pdrd_method *h, const Block &b)
{
  if (h->use_src) return source_theta(h) ? b.ks.patch_src_theta : b.ks.patch_src;
  return nodal_input(h) ? b.ks.patch_nodal : b.ks.patch;
}

size_t input_lds_bytes(const pdrd_method *h, const Block &b)
{
  if (h->use_src)
    return source_theta(h) ? b.ks.patch_src_theta_lds_bytes(b.maxl, b.cells) : b.ks.patch_src_lds_bytes(b.maxl, b.cells);
  return nodal_input(h) ? b.ks.patch_nodal_lds_bytes(b.maxl, b.cells) : b.ks.patch_lds_bytes(b.maxl, b.cells);
}

// Checks that every block has the kernel of the method's inputs and that its
// LDS fits a workgroup.
int check_inputs(const pdrd_method *h)
{
  for (const Block &b : h->blocks) {
    if (!input_kernel(h, b)) return set_error(PDRD_ENOTSUP, "no kernel for this scheme with these nodal inputs");
    if (input_lds_bytes(h, b) > 65536)
      return set_error(PDRD_ENOTSUP, "the nodal inputs do not fit the patch LDS: set PDRD_OPT_PATCH_CELLS to 64");
  }
  return PDRD_OK;
}

// Cell loop of the atomic / colour modes over every block (and colour),
// accumulating into acc_res / acc_ws.
int launch_cells(pdrd_method *h)
{
  if (h->precision) return set_error(PDRD_ENOTSUP, "the mixed-precision cells run with the patch assembly only");
  if (h->use_src) return set_error(PDRD_ENOTSUP, "the source term runs with the patch assembly only");
  pdrd::CellArgs a{};
  a.nodal = nodal_input(h);
  a.coords = h->d_coords;
  a.u = h->d_u;
  a.acc_res = h->d_acc_res;
  a.acc_ws = h->d_acc_ws;
  a.status = h->d_status;
  a.pp.p0 = h->params[0];
  a.pp.p1 = h->params[1];
  for (Block &b : h->blocks) {
    a.tab = b.d_tab;
    a.gconn = b.d_gconn;
    a.sconn = b.d_sconn;
    a.ne = b.ne;
    if (h->scatter_mode == 0) {
      a.atomic = 1;
      a.first = 0;
      a.count = b.ne;
      HIP_CHECK((a.nodal ? b.ks.cell_nodal : b.ks.cell)(a, grid_for(b.ne, h->block_threads), h->block_threads,
                                                        h->stream));
    } else {
      a.atomic = 0;
      for (size_t c = 0; c + 1 < b.colour_start.size(); ++c) {
        a.first = b.colour_start[c];
        a.count = b.colour_start[c + 1] - b.colour_start[c];
        if (a.count == 0) continue;
        HIP_CHECK((a.nodal ? b.ks.cell_nodal : b.ks.cell)(a, grid_for(a.count, h->block_threads), h->block_threads,
                                                          h->stream));
      }
    }
  }
  return PDRD_OK;
}

unsigned fixup_blocks(const pdrd_method *h)
{
  return std::min<unsigned>(kNormMaxBlocks, std::max<unsigned>(1, grid_for(h->n_shared, 256)));
}

// Place of one sweep in a chain of iterations whose shared-node sums are
// deferred to the next sweep: pend = buffer holding the previous sweep's rows
// and start states (-1: none, the state in u is complete), write = buffer of
// this sweep's rows (and start states when record is set: the next sweep is
// deferred too).
struct Chain
{
  int pend = -1;
  int write = 0;
  bool record = false;
};

double *rows_buffer(const pdrd_method *h, int i) { return i ? h->d_slot_acc_b : h->d_slot_acc; }

// lin: a linear sweep of the mixed-precision form (J delta about u_ref).
pdrd::PatchArgs patch_args(const pdrd_method *h, const Block *b, int update, int outputs, int excl, double cfl,
                           Chain c = {}, bool lin = false)
{
  pdrd::PatchArgs a{};
  if (b) {
    a.tab = b->d_tab;
    a.gconn = b->d_pgconn;
    a.lconn = b->d_lconn;
    a.ne = b->ne;
    a.np = b->np;
    a.first_patch = b->first_patch;
    a.maxl = b->maxl;
  }
  a.pnode_off = h->d_pnode_off;
  a.pnodes = h->d_pnodes;
  a.pn_inner = h->d_pn_inner;
  a.pn_coff = h->d_pn_coff;
  a.coords = h->d_coords;
  a.u = h->d_u;
  a.vol = h->d_vol;
  a.dir_index = h->d_dir_index;
  a.dir_vals = h->d_dir_vals;
  if (lin) {
    a.jlin = b ? b->d_jlin : nullptr;
    a.u_ref = h->d_uref;
    a.lin_res = h->d_linres ? h->d_linres : h->d_res;
    a.lin_dt = h->d_dt;
  }
  a.res = h->d_res;
  a.dt = h->d_dt;
  a.ws = h->d_ws;
  a.slot_acc = rows_buffer(h, c.write);
  a.slot_pos = h->d_slot_pos;
  a.pn_sid = h->d_pn_sid;
  a.pn_rows = h->d_pn_rows;
  if (c.pend >= 0) {
    a.prev_acc = rows_buffer(h, c.pend);
    a.base_in = h->d_base[c.pend];
  }
  if (c.record) a.base_out = h->d_base[c.write];
  a.norm_partial = h->d_patch_partial;
  a.nodal = h->use_src ? h->d_src : nodal_input(h);
  a.theta = source_theta(h);
  a.status = h->d_status;
  a.update = update;
  a.outputs = outputs;
  a.excl = excl;
  a.cfl = cfl;
  a.pp.p0 = h->params[0];
  a.pp.p1 = h->params[1];
  return a;
}

pdrd::FixupArgs fixup_args(const pdrd_method *h, int update, int outputs, int excl, double cfl, int buffer = 0,
                           bool lin = false)
{
  pdrd::FixupArgs f{};
  f.nodes = h->d_shared_nodes;
  f.slot_off = h->d_shared_off;
  f.slot_acc = rows_buffer(h, buffer);
  f.pa = patch_args(h, nullptr, update, outputs, excl, cfl, Chain{}, lin);
  f.nb = h->n_shared;
  f.norm_partial = h->d_patch_partial + static_cast<size_t>(h->np_total) * (h->m + 1);
  return f;
}

// The patch kernels of every block (FP64, with the nodal input of the
// scheme or the source field, or the linear sweep of the mixed form).
int launch_patch_blocks(pdrd_method *h, int update, int outputs, int excl, double cfl, Chain c = {},
                        bool lin = false)
{
  for (Block &b : h->blocks) {
    const pdrd::PatchArgs a = patch_args(h, &b, update, outputs, excl, cfl, c, lin);
    const pdrd::PatchLauncher f =
        lin ? b.lin : input_kernel(h, b);
    HIP_CHECK(f(a, b.np, b.cells, h->stream));
  }
  return PDRD_OK;
}

// The weak boundary conditions' facet kernels: after the patch kernels of a
// sweep (whose facet nodes are all shared, so u holds their state at the
// sweep's start) and before the shared-node sums; rows into row buffer
// `buffer`.
int launch_weak(pdrd_method *h, hipStream_t s, int buffer = 0)
{
  for (const WeakDev &w : h->weak) {
    if (w.nf == 0) continue;
    pdrd::WeakArgs a{};
    a.nodes = w.d_nodes;
    a.rows = w.d_rows;
    a.normal = w.d_normal;
    a.jac = w.d_jac;
    a.w = w.d_w;
    a.V = w.d_V;
    a.u = h->d_u;
    a.vol = h->d_vol;
    a.slot_acc = rows_buffer(h, buffer);
    a.status = h->d_status;
    a.nf = w.nf;
    a.nq = w.nq;
    a.type = w.type;
    a.relax = w.relax;
    a.qstate = w.d_qstate;
    a.qvel = w.d_qvel;
    for (int i = 0; i < 5; ++i) a.params[i] = w.params[i];
    a.pp.p0 = h->params[0];
    a.pp.p1 = h->params[1];
    const dim3 grid((w.nf + 127) / 128);
    const int key = h->physics * 8 + w.nfn;
#define WEAK_CASE(PH, NFN)                                                                   \
  case PH * 8 + NFN:                                                                         \
    hipLaunchKernelGGL((pdrd::weak_bc_kernel<PH, NFN>), grid, dim3(128), 0, s, a);           \
    break;
    switch (key) {
      WEAK_CASE(pdrd::PH_ADVECTION2D, 2)
      WEAK_CASE(pdrd::PH_ADVECTION2D, 3)
      WEAK_CASE(pdrd::PH_BURGERS2D, 2)
      WEAK_CASE(pdrd::PH_BURGERS2D, 3)
      WEAK_CASE(pdrd::PH_ROTATION2D, 2)
      WEAK_CASE(pdrd::PH_ROTATION2D, 3)
      WEAK_CASE(pdrd::PH_EULER2D, 2)
      WEAK_CASE(pdrd::PH_EULER2D, 3)
      WEAK_CASE(pdrd::PH_EULER3D, 3)
      WEAK_CASE(pdrd::PH_EULER3D, 6)
      default: return set_error(PDRD_ENOTSUP, "weak boundary condition on this facet kind");
    }
#undef WEAK_CASE
    HIP_CHECK(hipGetLastError());
  }
  return PDRD_OK;
}

// The sums of the shared nodes (rows of row buffer `buffer`) and their nodal
// step.
int launch_fixup(pdrd_method *h, int update, int outputs, int excl, double cfl, int buffer = 0, bool lin = false)
{
  if (h->n_shared == 0) return PDRD_OK;
  const pdrd::FixupArgs f = fixup_args(h, update, outputs, excl, cfl, buffer, lin);
  const unsigned nb = fixup_blocks(h);
#define FIX_CASE(MM)                                                                                    \
  case MM:                                                                                              \
    hipLaunchKernelGGL(pdrd::patch_fixup_kernel<MM>, dim3(nb), dim3(256), 0, h->stream, f);             \
    break;
  switch (h->m) {
    FIX_CASE(1)
    FIX_CASE(4)
    FIX_CASE(5)
    default: return set_error(PDRD_ENOTSUP, "unsupported number of equations");
  }
#undef FIX_CASE
  HIP_CHECK(hipGetLastError());
  return PDRD_OK;
}

// Patch assembly of every block followed by the sums of the shared nodes.
// update = 0: residual and wave speed of every node; update = 1: one
// pseudo-time iteration (the strong Dirichlet values are applied in the
// nodal step), with res / dt / ws / norm partials when outputs is set.
int launch_patches(pdrd_method *h, int update, int outputs, int excl, double cfl)
{
  if (int rc = launch_patch_blocks(h, update, outputs, excl, cfl)) return rc;
  if (int rc = launch_weak(h, h->stream)) return rc;
  return launch_fixup(h, update, outputs, excl, cfl);
}

// One pseudo-time iteration of a chain (patch assembly): the shared nodes of
// the previous sweep are finished in this sweep's gather when c.pend >= 0;
// the shared-node sums of this sweep are summed here (close) or left to the
// next sweep.
int launch_chain_sweep(pdrd_method *h, int outputs, int excl, double cfl, Chain c, bool close, bool lin = false)
{
  if (int rc = launch_patch_blocks(h, 1, outputs, excl, cfl, c, lin)) return rc;
  if (int rc = launch_weak(h, h->stream, c.write)) return rc;
  return close ? launch_fixup(h, 1, outputs, excl, cfl, c.write, lin) : PDRD_OK;
}

// After an FP64 sweep of the mixed form with weak BCs: linres = its residual
// without the facet rows (of row buffer `buffer`).
int lin_residual(pdrd_method *h, int buffer)
{
  const size_t nm = static_cast<size_t>(h->nn) * h->m;
  HIP_CHECK(hipMemcpyAsync(h->d_linres, h->d_res, nm * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  if (h->n_shared == 0) return PDRD_OK;
  hipLaunchKernelGGL(lin_res_kernel, dim3(grid_for(h->n_shared, 256)), dim3(256), 0, h->stream, h->n_shared, h->m,
                     h->d_shared_nodes, h->d_shared_off, h->d_shared_extra, rows_buffer(h, buffer), h->d_vol,
                     h->d_linres);
  HIP_CHECK(hipGetLastError());
  return PDRD_OK;
}

int launch_dirichlet(pdrd_method *h)
{
  if (h->n_dir == 0) return PDRD_OK;
  hipLaunchKernelGGL(dirichlet_kernel, dim3(grid_for(h->n_dir, 256)), dim3(256), 0, h->stream, h->n_dir, h->m,
                     h->d_dir_nodes, h->d_dir_vals, h->d_u);
  HIP_CHECK(hipGetLastError());
  return PDRD_OK;
}

unsigned norm_blocks(uint32_t nn)
{
  return std::min<unsigned>(kNormMaxBlocks, std::max<unsigned>(1, grid_for(nn, kNormBlock)));
}

int launch_node_pass(pdrd_method *h, double cfl, int outputs, int excl)
{
  const unsigned nb = norm_blocks(h->nn);
#define NODE_CASE(MM)                                                                                        \
  case MM:                                                                                                   \
    hipLaunchKernelGGL(node_pass_kernel<MM>, dim3(nb), dim3(kNormBlock), 0, h->stream, h->nn, cfl,             \
                       h->d_acc_res, h->d_acc_ws, h->d_vol, h->d_u, h->d_res, h->d_dt, h->d_ws, outputs, excl, \
                       h->d_is_dir, h->d_partial);                                                           \
    break;
  switch (h->m) {
    NODE_CASE(1)
    NODE_CASE(4)
    NODE_CASE(5)
    default: return set_error(PDRD_ENOTSUP, "unsupported number of equations");
  }
#undef NODE_CASE
  HIP_CHECK(hipGetLastError());
  return PDRD_OK;
}

int reduce_partials(pdrd_method *h, unsigned nb, double *norm, const double *buffer = nullptr)
{
  std::vector<double> part(static_cast<size_t>(nb) * (h->m + 1));
  HIP_CHECK(hipMemcpyAsync(part.data(), buffer ? buffer : h->d_partial, part.size() * sizeof(double),
                           hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  std::vector<double> s(h->m + 1, 0.0);
  for (unsigned b = 0; b < nb; ++b)
    for (int k = 0; k <= h->m; ++k) s[k] += part[static_cast<size_t>(b) * (h->m + 1) + k];
  const double cnt = s[h->m] > 0 ? s[h->m] : 1.0;
  for (int k = 0; k < h->m; ++k) norm[k] = std::sqrt(s[k] / cnt);
  return PDRD_OK;
}

int norm_impl(pdrd_method *h, const double *scale, int excl, double *norm)
{
  const unsigned nb = norm_blocks(h->nn);
#define NORM_CASE(MM)                                                                                           \
  case MM:                                                                                                      \
    hipLaunchKernelGGL(norm_partial_kernel<MM>, dim3(nb), dim3(kNormBlock), 0, h->stream, h->nn, h->d_res, scale, \
                       excl, h->d_is_dir, h->d_partial);                                                        \
    break;
  switch (h->m) {
    NORM_CASE(1)
    NORM_CASE(4)
    NORM_CASE(5)
    default: return set_error(PDRD_ENOTSUP, "unsupported number of equations");
  }
#undef NORM_CASE
  HIP_CHECK(hipGetLastError());
  return reduce_partials(h, nb, norm);
}

int check_ready(pdrd_method *h)
{
  if (!h) return set_error(PDRD_EINVAL, "null method");
  if (!h->ready) return set_error(PDRD_EINVAL, "method not initialized");
  return PDRD_OK;
}

// The mixed-precision form's linearisation no longer matches the method's
// state or inputs: the next iteration starts with an FP64 sweep and fresh
// Jacobians.
void lin_invalidate(pdrd_method *h)
{
  h->lin_valid = false;
  h->since_refresh = 1 << 30;
}

void drop_graph(pdrd_method *h)
{
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  h->graph_cfl = -1.0;
}

// Dual volume of every node: the whole volume of each cell holding it, the
// volume integrated with the Gauss rule of the cell's order
// (RDTimeUpdate.hpp:71-194).
std::vector<double> dual_volumes(const pdrd_method *h)
{
  std::vector<double> vol(h->nn, 0.0);
  const int D = h->dim;
  for (const Block &b : h->blocks) {
    const pdrd::Shape shape = static_cast<pdrd::Shape>(b.fe.shape);
    std::vector<double> qp, w, dVg;
    pdrd::quadrature(shape, b.fe.sol_order, qp, w);
    pdrd::LagrangeBasis geo(shape, b.fe.geo_order, pdrd::PointSet::Equidist);
    geo.derivatives(qp, dVg);
    const int nq = static_cast<int>(w.size()), ng = b.fe.ng, nd = b.fe.nd;
    for (uint32_t c = 0; c < b.ne; ++c) {
      double cv = 0.0;
      for (int q = 0; q < nq; ++q) {
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int xi = 0; xi < D; ++xi)
          for (int x = 0; x < D; ++x) {
            double s = 0.0;
            for (int g = 0; g < ng; ++g)
              s += dVg[(static_cast<size_t>(xi) * nq + q) * ng + g] *
                   h->coords[static_cast<size_t>(b.geo[static_cast<size_t>(c) * ng + g]) * D + x];
            J[xi][x] = s;
          }
        const double det = (D == 2) ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                                    : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                          J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                          J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        cv += w[q] * det;
      }
      for (int n = 0; n < nd; ++n) vol[b.sol[static_cast<size_t>(c) * nd + n]] += cv;
    }
  }
  return vol;
}

void free_device(pdrd_method *h)
{
  drop_graph(h);
  for (Block &b : h->blocks) {
    (void)hipFree(b.d_tab);
    (void)hipFree(b.d_gconn);
    (void)hipFree(b.d_sconn);
    (void)hipFree(b.d_pgconn);
    (void)hipFree(b.d_lconn);
    (void)hipFree(b.d_p1w);
    (void)hipFree(b.d_av);
    b.d_p1w = nullptr;
    b.d_av = nullptr;
    b.d_tab = nullptr;
    b.d_gconn = nullptr;
    b.d_sconn = nullptr;
    b.d_pgconn = nullptr;
    b.d_lconn = nullptr;
  }
  uint32_t **u32s[] = {&h->d_pnode_off, &h->d_pnodes, &h->d_pn_inner, &h->d_shared_nodes, &h->d_shared_off,
                       &h->d_slot_pos, &h->d_pn_sid};
  for (uint32_t **p : u32s) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  for (WeakDev &w : h->weak) {
    (void)hipFree(w.d_nodes);
    (void)hipFree(w.d_rows);
    (void)hipFree(w.d_normal);
    (void)hipFree(w.d_jac);
    (void)hipFree(w.d_w);
    (void)hipFree(w.d_V);
    (void)hipFree(w.d_qstate);
    (void)hipFree(w.d_qvel);
  }
  h->weak.clear();
  (void)hipFree(h->d_pn_coff);
  (void)hipFree(h->d_slot_acc);
  (void)hipFree(h->d_patch_partial);
  for (double **q : {&h->d_slot_acc_b, &h->d_base[0], &h->d_base[1]}) {
    (void)hipFree(*q);
    *q = nullptr;
  }
  (void)hipFree(h->d_pn_rows);
  h->d_pn_rows = nullptr;
  for (double **q : {&h->d_uref, &h->d_linres}) {
    (void)hipFree(*q);
    *q = nullptr;
  }
  (void)hipFree(h->d_shared_extra);
  h->d_shared_extra = nullptr;
  h->lin_valid = false;
  for (Block &b : h->blocks) {
    for (float **q : {&b.d_jlin, &b.d_jstab_lin}) {
      (void)hipFree(*q);
      *q = nullptr;
    }
    (void)hipFree(b.d_lin_pos);
    b.d_lin_pos = nullptr;
  }
  for (double **q : {&h->d_blend, &h->d_visc, &h->d_stab_tmp, &h->d_src}) {
    (void)hipFree(*q);
    *q = nullptr;
  }
  h->use_blend = h->use_visc = h->use_src = false;
  (void)hipFree(h->d_dir_index);
  h->d_pn_coff = nullptr;
  h->d_slot_acc = nullptr;
  h->d_patch_partial = nullptr;
  h->d_dir_index = nullptr;
  double **dp[] = {&h->d_coords, &h->d_u, &h->d_res, &h->d_dt, &h->d_ws, &h->d_vol, &h->d_acc_res, &h->d_acc_ws,
                   &h->d_partial, &h->d_dir_vals};
  for (double **p : dp) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  (void)hipFree(h->d_status);
  (void)hipFree(h->d_dir_nodes);
  (void)hipFree(h->d_is_dir);
  h->d_status = nullptr;
  h->d_dir_nodes = nullptr;
  h->d_is_dir = nullptr;
  for (hipEvent_t &e : h->ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  h->stream = nullptr;
  h->own_stream = false;
  h->bytes = 0;
}

// One iteration without outputs: cells, Dirichlet values, fused nodal pass.
int enqueue_iteration(pdrd_method *h, double cfl, int outputs, int excl)
{
  if (h->scatter_mode == 2) return launch_patches(h, 1, outputs, excl, cfl);
  int rc = launch_cells(h);
  if (rc) return rc;
  rc = launch_dirichlet(h);
  if (rc) return rc;
  return launch_node_pass(h, cfl, outputs, excl);
}

// Set-up of the mixed-precision form at its first use: u_ref, the FP32
// Jacobian blocks of every block in patch order, the colour position ->
// patch position map of the Jacobian kernel, the linear sweep kernel.
int lin_setup(pdrd_method *h)
{
  if (h->d_uref) return PDRD_OK;
  if (h->scheme == PDRD_PSI || h->scheme == PDRD_BPSI)
    return set_error(PDRD_ENOTSUP, "the mixed-precision form needs element Jacobians (none for PSI)");
  for (Block &b : h->blocks) {
    const int lin_phys = (h->m == 1) ? pdrd::PH_ADVECTION2D : h->physics;
    const pdrd::KernelSet lk =
        pdrd::find_cell(pdrd::CellKey{b.fe.shape, b.fe.sol_order, b.fe.nq, lin_phys, pdrd::SC_LDA});
    if (!lk.patch_lin || !b.ks.jacobian) return set_error(PDRD_ENOTSUP, "no linear sweep for this element type");
    if (lk.patch_lin_lds_bytes(b.maxl, b.cells) > 65536)
      return set_error(PDRD_ENOTSUP, "linear sweep: patch LDS above 64 KiB");
    b.lin = lk.patch_lin;
  }
  const size_t nm = static_cast<size_t>(h->nn) * h->m;
  if (int rc = dev_alloc(h, &h->d_uref, nm)) return rc;
  if (!h->weak.empty())
    if (int rc = dev_alloc(h, &h->d_linres, nm)) return rc;
  for (Block &b : h->blocks) {
    const size_t ndm = static_cast<size_t>(b.fe.nd) * h->m;
    if (int rc = dev_alloc(h, &b.d_jlin, ndm * ndm * b.ne)) return rc;
    if (h->scheme == PDRD_B && !(h->m == 1 && b.fe.nd <= 6))
      if (int rc = dev_alloc(h, &b.d_jstab_lin, ndm * ndm * b.ne)) return rc;
    std::vector<uint32_t> pos(b.ne);
    for (uint32_t e = 0; e < b.ne; ++e) pos[e] = b.ppos[b.perm[e]];
    if (int rc = dev_alloc(h, &b.d_lin_pos, pos.size())) return rc;
    HIP_CHECK(hipMemcpy(b.d_lin_pos, pos.data(), pos.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  return PDRD_OK;
}

// The element Jacobians at the current state, FP32 into every block's d_jlin.
int lin_jacobians(pdrd_method *h)
{
  for (Block &b : h->blocks) {
    pdrd::JacArgs a{};
    a.tab = b.d_tab;
    a.gconn = b.d_gconn;
    a.sconn = b.d_sconn;
    a.perm = b.d_lin_pos;  // (the kernel's caller-slot index; unused with the FP32 output)
    a.coords = h->d_coords;
    a.u = h->d_u;
    a.nodal = nodal_input(h);
    a.status = h->d_status;
    a.ne = b.ne;
    a.pp.p0 = h->params[0];
    a.pp.p1 = h->params[1];
    a.jlin = b.d_jlin;
    a.jstab_lin = b.d_jstab_lin;
    a.lin_pos = b.d_lin_pos;
    if (b.ne) HIP_CHECK(b.ks.jacobian(a, grid_for(b.ne, 64), h->stream));
  }
  h->since_refresh = 0;
  return PDRD_OK;
}

bool mixed(const pdrd_method *h) { return h->precision == 1 && h->refresh > 1 && h->scatter_mode == 2; }

int iterate_norm(pdrd_method *h, double *norm)
{
  if (h->scatter_mode == 2)
    return reduce_partials(h, h->np_total + (h->n_shared ? fixup_blocks(h) : 0), norm, h->d_patch_partial);
  return reduce_partials(h, norm_blocks(h->nn), norm);
}

// Iterations of the patch assembly. Chained (PDRD_OPT_CHAIN), the
// shared-node sums of one sweep are done in the next sweep's gather, so a
// sweep is one launch per element type (plus the weak facet kernels); a
// sweep closes the chain (runs the fix-up) when it is the call's last or,
// in the mixed-precision form, when the next sweep is of the other kind (FP64
// or linear, whose nodal steps differ). Unchained, every sweep runs the fix-up. The sweeps are
// launched directly: replaying them from captured HIP graphs measured slower
// on MI355X (25.6 against 11.1 us per sweep of a 32 768-cell P1 mesh, the
// launches being dearer than the kernels). With timing on, events bracket the
// patch kernels (cell interval) and the facet kernels + fix-up (node
// interval) of every sweep.
// Default: chained for the scalar physics (fewer launches, little shared-node
// work per patch), not for the Euler systems, whose patches hold more shared
// nodes of more equations, so that the deferred steps cost more in their
// gather than the fix-up pass does.
bool chains(const pdrd_method *h) { return h->chain >= 0 ? h->chain == 1 : h->m == 1; }

int iterate_patches(pdrd_method *h, double cfl, int n_iter, int excl, double *norm)
{
  const bool mix = mixed(h);
  if (mix)
    if (int rc = lin_setup(h)) return rc;
  const bool chain = chains(h);
  double cell = 0.0, node = 0.0;
  int timed = 0;
  int pend = -1;
  // mixed form: iteration `it` is an FP64 sweep when there is no valid
  // linearisation, when the Jacobians are due, and at the call's last
  // iteration (whose outputs and norm are FP64); the others are linear
  auto fp64_at = [&](int it, int since) {
    return !mix || it == n_iter - 1 || since >= h->refresh || (it == 0 && !h->lin_valid);
  };
  if (!h->timing) HIP_CHECK(hipEventRecord(h->ev[3], h->stream));
  for (int it = 0; it < n_iter; ++it) {
    const bool f64 = fp64_at(it, h->since_refresh);
    const bool lin = !f64;
    if (mix && f64) {
      // new linearisation point: Jacobians when due, then u_ref = u; the
      // sweep below writes R(u_ref) and dt(u_ref) as its outputs
      if (!h->lin_valid || h->since_refresh >= h->refresh)
        if (int rc = lin_jacobians(h)) return rc;
      HIP_CHECK(hipMemcpyAsync(h->d_uref, h->d_u, static_cast<size_t>(h->nn) * h->m * sizeof(double),
                               hipMemcpyDeviceToDevice, h->stream));
    }
    const bool last = it == n_iter - 1;
    const bool next_f64 = !last && fp64_at(it + 1, h->since_refresh + 1);
    // chains run within runs of sweeps of one kind
    const bool close = !chain || last || (mix && (f64 || next_f64));
    const int out = (last || (mix && f64)) ? 1 : 0;
    Chain c;
    c.pend = pend;
    c.write = (pend == 0) ? 1 : 0;
    c.record = !close;
    if (h->timing) {
      HIP_CHECK(hipEventRecord(h->ev[0], h->stream));
      if (int rc = launch_patch_blocks(h, 1, out, excl, cfl, c, lin)) return rc;
      HIP_CHECK(hipEventRecord(h->ev[1], h->stream));
      if (int rc = launch_weak(h, h->stream, c.write)) return rc;
      if (close)
        if (int rc = launch_fixup(h, 1, out, excl, cfl, c.write, lin)) return rc;
      HIP_CHECK(hipEventRecord(h->ev[2], h->stream));
      HIP_CHECK(hipEventSynchronize(h->ev[2]));
      float a = 0.f, b = 0.f;
      HIP_CHECK(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
      HIP_CHECK(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
      if (h->timing == 1 || lin) {
        cell += a;
        node += b;
        ++timed;
      }
    } else {
      if (int rc = launch_chain_sweep(h, out, excl, cfl, c, close, lin)) return rc;
    }
    if (mix && f64) {
      h->lin_valid = true;
      if (h->d_linres)
        if (int rc = lin_residual(h, c.write)) return rc;
    }
    ++h->since_refresh;
    pend = close ? -1 : c.write;
  }
  if (h->timing) {
    h->last_cell_ms = timed ? cell / timed : 0.0;
    h->last_node_ms = timed ? node / timed : 0.0;
  } else {
    HIP_CHECK(hipEventRecord(h->ev[1], h->stream));
    HIP_CHECK(hipEventSynchronize(h->ev[1]));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, h->ev[3], h->ev[1]));
    h->last_cell_ms = ms;
    h->last_node_ms = 0.0;
  }
  if (norm) return iterate_norm(h, norm);
  return PDRD_OK;
}

}  // namespace

#define PDRD_TRY try {
#define PDRD_CATCH                                   \
  }                                                  \
  catch (const std::exception &ex)                   \
  {                                                  \
    return set_error(PDRD_EINVAL, ex.what());        \
  }

extern "C" {

int pdrd_create(int dim, uint32_t nb_vertices, const double *vertex_coords, uint32_t nb_nodes, int physics,
                int scheme, const double *params, int nb_params, pdrd_method_t *out)
{
  PDRD_TRY
  if (!out || !vertex_coords) return set_error(PDRD_EINVAL, "null argument");
  if (dim != 2 && dim != 3) return set_error(PDRD_EINVAL, "dim must be 2 or 3");
  if (physics < 0 || physics > 4) return set_error(PDRD_EINVAL, "unknown physics");
  if (scheme < 0 || scheme > 4) return set_error(PDRD_EINVAL, "unknown scheme");
  const int pdim = (physics == PDRD_EULER3D) ? 3 : 2;
  if (pdim != dim) return set_error(PDRD_EINVAL, "physics dimension does not match the mesh");
  if (nb_nodes == 0 || nb_vertices == 0) return set_error(PDRD_EINVAL, "empty mesh");
  auto *h = new pdrd_method();
  h->dim = dim;
  h->physics = physics;
  h->scheme = scheme;
  h->m = (physics == PDRD_EULER2D) ? 4 : (physics == PDRD_EULER3D) ? 5 : 1;
  h->nv = nb_vertices;
  h->nn = nb_nodes;
  if (physics == PDRD_EULER2D || physics == PDRD_EULER3D) {
    h->params[0] = 1.4;
    h->params[1] = 287.05;
  } else if (physics == PDRD_ADVECTION2D) {
    h->params[0] = 0.0;
    h->params[1] = 1.0;
  }
  for (int i = 0; i < nb_params && i < 4; ++i) h->params[i] = params[i];
  h->coords.assign(vertex_coords, vertex_coords + static_cast<size_t>(nb_vertices) * dim);
  *out = h;
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_add_cells(pdrd_method_t h, const pdrd_fe_desc *fe, const double *tables, uint32_t ne,
                   const uint32_t *geo_conn, const uint32_t *sol_conn)
{
  PDRD_TRY
  if (!h || !fe || !tables || (ne && (!geo_conn || !sol_conn))) return set_error(PDRD_EINVAL, "null argument");
  if (h->ready) return set_error(PDRD_EINVAL, "cells must be added before pdrd_initialize");
  if (fe->dim != h->dim) return set_error(PDRD_EINVAL, "element dimension does not match the mesh");
  // the descriptor must be one this library builds
  pdrd_fe_desc chk;
  int rc = pdrd_fe_describe(fe->shape, fe->geo_order, fe->sol_order, fe->quad_order, fe->point_set, &chk);
  if (rc) return rc;
  if (chk.ng != fe->ng || chk.nd != fe->nd || chk.nf != fe->nf || chk.nq != fe->nq ||
      chk.table_size != fe->table_size)
    return set_error(PDRD_EINVAL, "inconsistent element descriptor");
  if (fe->geo_order != 1) return set_error(PDRD_ENOTSUP, "only P1/Q1 geometry is built");
  pdrd::CellKey key{fe->shape, fe->sol_order, fe->nq, h->physics, h->scheme};
  const pdrd::KernelSet ks = pdrd::find_cell(key);
  if (!ks) return set_error(PDRD_ENOTSUP, "no cell kernel for this element type / physics / scheme");
  for (size_t i = 0; i < static_cast<size_t>(ne) * fe->ng; ++i)
    if (geo_conn[i] >= h->nv) return set_error(PDRD_EINVAL, "geometry connectivity out of range");
  for (size_t i = 0; i < static_cast<size_t>(ne) * fe->nd; ++i)
    if (sol_conn[i] >= h->nn) return set_error(PDRD_EINVAL, "solution connectivity out of range");
  Block b;
  b.fe = *fe;
  b.ne = ne;
  b.tables.assign(tables, tables + fe->table_size);
  b.geo.assign(geo_conn, geo_conn + static_cast<size_t>(ne) * fe->ng);
  b.sol.assign(sol_conn, sol_conn + static_cast<size_t>(ne) * fe->nd);
  b.ks = ks;
  h->blocks.push_back(std::move(b));
  return PDRD_OK;
  PDRD_CATCH
}

static_assert(pdrd::LC_BITS == pdrd::LCONN_BITS, "patch connectivity layout differs between host and device");

// The packed FE tables of a block followed by transposed copies of the
// derivative tables (dVf as [nf][dim][nq], dV as [nq][dim][nd], dVg as
// [nq][dim][ng]); the offsets match TabOffsets in rd_cell.hpp.
static std::vector<double> device_tables(const pdrd_fe_desc &fe, const std::vector<double> &t)
{
  const int D = fe.dim, NG = fe.ng, ND = fe.nd, NF = fe.nf, NQ = fe.nq;
  const size_t o_dVg = static_cast<size_t>(NQ) * D + NQ + static_cast<size_t>(NQ) * NG;
  const size_t o_V = o_dVg + static_cast<size_t>(D) * NQ * NG;
  const size_t o_dV = o_V + static_cast<size_t>(NQ) * ND;
  const size_t o_Vf = o_dV + static_cast<size_t>(D) * NQ * ND;
  const size_t o_dVf = o_Vf + static_cast<size_t>(NQ) * NF;
  std::vector<double> out(t);
  for (int f = 0; f < NF; ++f)
    for (int xi = 0; xi < D; ++xi)
      for (int q = 0; q < NQ; ++q) out.push_back(t[o_dVf + (static_cast<size_t>(xi) * NQ + q) * NF + f]);
  for (int q = 0; q < NQ; ++q)
    for (int xi = 0; xi < D; ++xi)
      for (int n = 0; n < ND; ++n) out.push_back(t[o_dV + (static_cast<size_t>(xi) * NQ + q) * ND + n]);
  for (int q = 0; q < NQ; ++q)
    for (int xi = 0; xi < D; ++xi)
      for (int g = 0; g < NG; ++g) out.push_back(t[o_dVg + (static_cast<size_t>(xi) * NQ + q) * NG + g]);
  return out;
}

int pdrd_initialize(pdrd_method_t h)
{
  PDRD_TRY
  if (!h) return set_error(PDRD_EINVAL, "null method");
  if (h->ready) return set_error(PDRD_EINVAL, "already initialized");
  if (h->blocks.empty()) return set_error(PDRD_EINVAL, "no cells");
  // every node must belong to a cell (a node without cells has no dual volume)
  std::vector<char> used(h->nn, 0);
  for (const Block &b : h->blocks)
    for (uint32_t v : b.sol) used[v] = 1;
  for (uint32_t i = 0; i < h->nn; ++i)
    if (!used[i]) return set_error(PDRD_EINVAL, "solution node not held by any cell");
  HIP_CHECK(hipGetDevice(&h->device));
  int rc;
  const size_t nm = static_cast<size_t>(h->nn) * h->m;
  if ((rc = dev_alloc(h, &h->d_coords, h->coords.size()))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_u, nm))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_res, nm))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_acc_res, nm))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_dt, h->nn))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_ws, h->nn))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_acc_ws, h->nn))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_vol, h->nn))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_partial, static_cast<size_t>(kNormMaxBlocks) * (h->m + 1)))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_status, 1))) { free_device(h); return rc; }
  if ((rc = dev_alloc(h, &h->d_is_dir, h->nn))) { free_device(h); return rc; }

  // colouring: greedy in caller order per block, then a stable sort by colour
  for (Block &b : h->blocks) {
    pdrd::SolMesh s;
    s.nb_nodes = h->nn;
    s.cell_offset.resize(b.ne + 1);
    for (uint32_t c = 0; c <= b.ne; ++c) s.cell_offset[c] = c * b.fe.nd;
    s.cell_nodes = b.sol;
    std::vector<uint32_t> cells(b.ne);
    for (uint32_t c = 0; c < b.ne; ++c) cells[c] = c;
    uint32_t ncol = 0;
    const std::vector<uint32_t> col = pdrd::greedy_colour(cells, s, &ncol);
    b.colour_start.assign(ncol + 1, 0);
    for (uint32_t c = 0; c < b.ne; ++c) b.colour_start[col[c] + 1]++;
    for (uint32_t k = 0; k < ncol; ++k) b.colour_start[k + 1] += b.colour_start[k];
    std::vector<uint32_t> fill(b.colour_start.begin(), b.colour_start.end() - 1);
    b.perm.assign(b.ne, 0);
    for (uint32_t c = 0; c < b.ne; ++c) b.perm[fill[col[c]]++] = c;
    const int ng = b.fe.ng, nd = b.fe.nd;
    std::vector<uint32_t> g(static_cast<size_t>(ng) * b.ne), so(static_cast<size_t>(nd) * b.ne);
    for (uint32_t p = 0; p < b.ne; ++p) {
      const uint32_t c = b.perm[p];
      for (int k = 0; k < ng; ++k) g[static_cast<size_t>(k) * b.ne + p] = b.geo[static_cast<size_t>(c) * ng + k];
      for (int k = 0; k < nd; ++k) so[static_cast<size_t>(k) * b.ne + p] = b.sol[static_cast<size_t>(c) * nd + k];
    }
    const std::vector<double> dtab = device_tables(b.fe, b.tables);
    if ((rc = dev_alloc(h, &b.d_tab, dtab.size()))) { free_device(h); return rc; }
    if ((rc = dev_alloc(h, &b.d_gconn, g.size()))) { free_device(h); return rc; }
    if ((rc = dev_alloc(h, &b.d_sconn, so.size()))) { free_device(h); return rc; }
    HIP_CHECK(hipMemcpy(b.d_tab, dtab.data(), dtab.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(b.d_gconn, g.data(), g.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(b.d_sconn, so.data(), so.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  // patch plan (the default assembly); any block it does not fit sends the
  // whole method to the colour-by-colour scatter
  {
    std::vector<pdrd::PatchBlockIn> pin;
    for (const Block &b : h->blocks) {
      pdrd::PatchBlockIn x;
      x.ne = b.ne;
      x.ng = b.fe.ng;
      x.nd = b.fe.nd;
      x.geo = b.geo.data();
      x.sol = b.sol.data();
      x.patch_size = h->patch_cells ? std::min(h->patch_cells, b.ks.patch_size) : b.ks.patch_size;
      pin.push_back(x);
    }
    // boundary facets of the weak BCs; their nodes get one extra partial-sum
    // row per facet
    std::vector<pdrd::FacetSet> fsets;
    std::vector<uint32_t> extra(h->nn, 0);
    if (!h->weak_specs.empty()) {
      std::vector<pdrd::FacetBlockIn> fin;
      for (const Block &b : h->blocks) {
        pdrd::FacetBlockIn x;
        x.shape = b.fe.shape;
        x.order = b.fe.sol_order;
        x.ng = b.fe.ng;
        x.nd = b.fe.nd;
        x.ne = b.ne;
        x.geo = b.geo.data();
        x.sol = b.sol.data();
        fin.push_back(x);
      }
      for (const WeakSpec &ws : h->weak_specs) {
        std::vector<char> in_set(h->nn, 0);
        for (uint32_t v : ws.nodes) in_set[v] = 1;
        fsets.push_back(pdrd::boundary_facets(h->dim, h->coords.data(), fin, in_set));
        for (uint32_t v : fsets.back().nodes) extra[v]++;
      }
    }
    // A block whose patches need more than 64 KiB of LDS at its default size
    // is planned again with half as many cells per patch (down to 64); a size
    // the caller chose is kept, and the method then falls back to colouring.
    pdrd::PatchPlan plan;
    for (;;) {
      h->patch_ok = true;
      try {
        plan = pdrd::build_patch_plan(h->dim, h->nv, h->coords.data(), h->nn, pin, fsets.empty() ? nullptr : &extra);
      } catch (const std::exception &ex) {
        h->patch_ok = false;
        h->patch_reason = ex.what();
        break;
      }
      bool shrunk = false;
      for (size_t i = 0; i < h->blocks.size(); ++i)
        // (room for one more double per node: the nodal input of the scheme)
        if (h->blocks[i].ks.patch_lds_bytes(plan.blocks[i].maxl, pin[i].patch_size) +
                static_cast<size_t>(plan.blocks[i].maxl) * sizeof(double) >
            65536) {
          if (!h->patch_cells && pin[i].patch_size > 64) {
            pin[i].patch_size /= 2;
            shrunk = true;
          } else {
            h->patch_ok = false;
            h->patch_reason = "patch LDS above 64 KiB";
          }
        }
      if (!h->patch_ok || !shrunk) break;
    }
    if (h->patch_ok) {
      for (size_t i = 0; i < h->blocks.size(); ++i) {
        Block &b = h->blocks[i];
        const pdrd::PatchBlockOut &o = plan.blocks[i];
        b.first_patch = o.first_patch;
        b.np = o.np;
        b.maxl = o.maxl;
        b.cells = pin[i].patch_size;
        b.ppos.assign(b.ne, 0);
        for (uint32_t e = 0; e < b.ne; ++e) b.ppos[o.order[e]] = e;
        if ((rc = dev_alloc(h, &b.d_pgconn, o.gconn.size()))) { free_device(h); return rc; }
        if ((rc = dev_alloc(h, &b.d_lconn, o.lconn.size()))) { free_device(h); return rc; }
        HIP_CHECK(hipMemcpy(b.d_pgconn, o.gconn.data(), o.gconn.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(b.d_lconn, o.lconn.data(), o.lconn.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
      }
      // the row of every shared slot in the node-grouped partial sums
      std::vector<uint32_t> slot_pos(plan.pnodes.size(), 0);
      for (size_t k = 0; k < plan.shared_slots.size(); ++k)
        if (plan.shared_slots[k] != UINT32_MAX) slot_pos[plan.shared_slots[k]] = static_cast<uint32_t>(k);
      // the shared-node index and the node's rows of every shared slot (0
      // for the others)
      std::vector<uint32_t> pn_sid(plan.pnodes.size(), 0);
      std::vector<uint2> pn_rows(plan.pnodes.size(), uint2{0u, 0u});
      {
        std::vector<uint32_t> sid(h->nn, 0);
        for (size_t i = 0; i < plan.shared_nodes.size(); ++i) sid[plan.shared_nodes[i]] = static_cast<uint32_t>(i);
        for (size_t p = 0; p < plan.pn_inner.size(); ++p)
          for (uint32_t k = plan.pnode_off[p] + plan.pn_inner[p]; k < plan.pnode_off[p + 1]; ++k) {
            const uint32_t i = sid[plan.pnodes[k]];
            pn_sid[k] = i;
            pn_rows[k] = uint2{plan.shared_off[i], plan.shared_off[i + 1]};
          }
      }
      h->np_total = static_cast<uint32_t>(plan.pn_inner.size());
      h->n_shared = static_cast<uint32_t>(plan.shared_nodes.size());
      auto up32 = [&](uint32_t **d, const std::vector<uint32_t> &v) -> int {
        int r = dev_alloc(h, d, v.size());
        if (r) return r;
        if (!v.empty() && hipMemcpy(*d, v.data(), v.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
          return set_error(PDRD_EHIP, "hipMemcpy of the patch plan");
        return PDRD_OK;
      };
      if ((rc = up32(&h->d_pnode_off, plan.pnode_off)) || (rc = up32(&h->d_pnodes, plan.pnodes)) ||
          (rc = up32(&h->d_pn_inner, plan.pn_inner)) || (rc = up32(&h->d_shared_nodes, plan.shared_nodes)) ||
          (rc = up32(&h->d_shared_off, plan.shared_off)) || (rc = up32(&h->d_slot_pos, slot_pos)) ||
          (rc = up32(&h->d_pn_sid, pn_sid))) {
        free_device(h);
        return rc;
      }
      if ((rc = dev_alloc(h, &h->d_pn_coff, plan.pn_coff.size()))) { free_device(h); return rc; }
      HIP_CHECK(hipMemcpy(h->d_pn_coff, plan.pn_coff.data(), plan.pn_coff.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
      if ((rc = dev_alloc(h, &h->d_slot_acc, static_cast<size_t>(plan.shared_slots.size()) * (h->m + 1))) ||
          (rc = dev_alloc(h, &h->d_slot_acc_b, static_cast<size_t>(plan.shared_slots.size()) * (h->m + 1))) ||
          (rc = dev_alloc(h, &h->d_base[0], static_cast<size_t>(h->n_shared) * h->m)) ||
          (rc = dev_alloc(h, &h->d_base[1], static_cast<size_t>(h->n_shared) * h->m)) ||
          (rc = dev_alloc(h, &h->d_pn_rows, pn_rows.size()))) {
        free_device(h);
        return rc;
      }
      HIP_CHECK(hipMemcpy(h->d_pn_rows, pn_rows.data(), pn_rows.size() * sizeof(uint2), hipMemcpyHostToDevice));
      if ((rc = dev_alloc(h, &h->d_patch_partial, static_cast<size_t>(h->np_total + kNormMaxBlocks) * (h->m + 1)))) {
        free_device(h);
        return rc;
      }
      // weak BCs: the row of every (facet, node) pair after the node's patch rows
      if (!fsets.empty()) {
        std::vector<uint32_t> sx(h->n_shared);
        for (uint32_t i = 0; i < h->n_shared; ++i) sx[i] = extra[plan.shared_nodes[i]];
        if ((rc = up32(&h->d_shared_extra, sx))) {
          free_device(h);
          return rc;
        }
        std::vector<uint32_t> next(h->nn, UINT32_MAX);
        for (uint32_t i = 0; i < h->n_shared; ++i) {
          const uint32_t v = plan.shared_nodes[i];
          next[v] = plan.shared_off[i + 1] - extra[v];
        }
        for (size_t k = 0; k < fsets.size(); ++k) {
          const pdrd::FacetSet &fs = fsets[k];
          const WeakSpec &ws = h->weak_specs[k];
          WeakDev wd;
          wd.type = ws.type;
          wd.nfn = fs.nfn;
          wd.nf = static_cast<uint32_t>(fs.size());
          for (int i = 0; i < 5; ++i) wd.params[i] = ws.params[i];
          const int forder = (fs.nfn > fs.nfv) ? 2 : 1;
          wd.relax = (ws.type == PDRD_WEAK_WALL && forder > 1) ? 1.0 / (3.0 * forder) : 1.0;
          std::vector<uint32_t> rows(fs.nodes.size());
          for (size_t j = 0; j < fs.nodes.size(); ++j) rows[j] = next[fs.nodes[j]]++;
          std::vector<double> w, V;
          if (wd.nf) pdrd::facet_tables(h->dim, fs.nfn, ws.quad_order + 1, w, V);
          wd.nq = static_cast<int>(w.size());
          auto upd = [&](double **d, const std::vector<double> &v) -> int {
            int r = dev_alloc(h, d, v.size());
            if (r) return r;
            if (!v.empty() && hipMemcpy(*d, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
              return set_error(PDRD_EHIP, "hipMemcpy of the boundary facets");
            return PDRD_OK;
          };
          // coordinates of the facet points: the vertices' P1 / linear functions
          // at the same rule (straight facets)
          {
            std::vector<double> w1, V1;
            if (wd.nf) pdrd::facet_tables(h->dim, fs.nfv, ws.quad_order + 1, w1, V1);
            wd.qpoints.assign(static_cast<size_t>(wd.nf) * wd.nq * h->dim, 0.0);
            for (uint32_t f = 0; f < wd.nf; ++f)
              for (int q = 0; q < wd.nq; ++q)
                for (int v = 0; v < fs.nfv; ++v) {
                  const uint32_t vx = fs.verts[static_cast<size_t>(f) * fs.nfv + v];
                  for (int d = 0; d < h->dim; ++d)
                    wd.qpoints[(static_cast<size_t>(f) * wd.nq + q) * h->dim + d] +=
                        V1[static_cast<size_t>(q) * fs.nfv + v] * h->coords[static_cast<size_t>(vx) * h->dim + d];
                }
          }
          h->weak.push_back(wd);
          WeakDev &wr = h->weak.back();
          if ((rc = up32(&wr.d_nodes, fs.nodes)) || (rc = up32(&wr.d_rows, rows)) ||
              (rc = upd(&wr.d_normal, fs.normal)) || (rc = upd(&wr.d_jac, fs.jac)) || (rc = upd(&wr.d_w, w)) ||
              (rc = upd(&wr.d_V, V))) {
            free_device(h);
            return rc;
          }
          if (h->physics == pdrd::PH_ROTATION2D) {
            // AroundOrigin2D: the velocity (y, -x) of every facet point
            std::vector<double> vel(static_cast<size_t>(wr.nf) * wr.nq * 2);
            for (size_t i = 0; i < vel.size() / 2; ++i) {
              vel[2 * i] = wr.qpoints[2 * i + 1];
              vel[2 * i + 1] = -wr.qpoints[2 * i];
            }
            if ((rc = upd(&wr.d_qvel, vel))) {
              free_device(h);
              return rc;
            }
          }
        }
      }
    } else {
      if (!h->weak_specs.empty()) {
        free_device(h);
        return set_error(PDRD_ENOTSUP, "weak boundary conditions need the patch assembly: " + h->patch_reason);
      }
      h->scatter_mode = 1;
    }
  }
  if ((rc = dev_alloc(h, &h->d_dir_index, h->nn))) { free_device(h); return rc; }
  {
    std::vector<int> di(h->nn, -1);
    HIP_CHECK(hipMemcpy(h->d_dir_index, di.data(), di.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  const std::vector<double> vol = dual_volumes(h);
  HIP_CHECK(hipMemcpy(h->d_coords, h->coords.data(), h->coords.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(h->d_vol, vol.data(), vol.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemset(h->d_u, 0, nm * sizeof(double)));
  HIP_CHECK(hipMemset(h->d_res, 0, nm * sizeof(double)));
  HIP_CHECK(hipMemset(h->d_acc_res, 0, nm * sizeof(double)));
  HIP_CHECK(hipMemset(h->d_dt, 0, h->nn * sizeof(double)));
  HIP_CHECK(hipMemset(h->d_ws, 0, h->nn * sizeof(double)));
  HIP_CHECK(hipMemset(h->d_acc_ws, 0, h->nn * sizeof(double)));
  HIP_CHECK(hipMemset(h->d_status, 0, sizeof(int)));
  HIP_CHECK(hipMemset(h->d_is_dir, 0, h->nn));
  if (!h->stream) {
    HIP_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->own_stream = true;
  }
  for (hipEvent_t &e : h->ev) HIP_CHECK(hipEventCreate(&e));
  HIP_CHECK(hipDeviceSynchronize());
  h->ready = true;
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_destroy(pdrd_method_t h)
{
  if (!h) return PDRD_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_device(h);
  delete h;
  return PDRD_OK;
}

int pdrd_set_option(pdrd_method_t h, int option, int value)
{
  if (!h) return set_error(PDRD_EINVAL, "null method");
  if (option == PDRD_OPT_SCATTER || option == PDRD_OPT_REFRESH || option == PDRD_OPT_CHAIN) lin_invalidate(h);
  switch (option) {
    case PDRD_OPT_SCATTER:
      if (value < 0 || value > 2) return set_error(PDRD_EINVAL, "scatter mode is 0, 1 or 2");
      if (value == 2 && h->ready && !h->patch_ok)
        return set_error(PDRD_ENOTSUP, "patch assembly unavailable: " + h->patch_reason);
      if (value != 2 && !h->weak_specs.empty())
        return set_error(PDRD_ENOTSUP, "weak boundary conditions run with the patch assembly only");
      if (value != 2 && h->use_src) return set_error(PDRD_ENOTSUP, "the source term runs with the patch assembly only");
      h->scatter_mode = value;
      break;
    case PDRD_OPT_PRECISION:
      if (value != 0 && value != 1) return set_error(PDRD_EINVAL, "precision is 0 (FP64) or 1 (mixed)");
      if (value == 1 && (h->scheme == PDRD_PSI || h->scheme == PDRD_BPSI))
        return set_error(PDRD_ENOTSUP, "the mixed-precision form needs element Jacobians (none for PSI)");
      h->precision = value;
      lin_invalidate(h);
      break;
    case PDRD_OPT_BLOCK:
      if (value < 64 || value > 256 || (value % 64) != 0)
        return set_error(PDRD_EINVAL, "block size must be a multiple of 64 in [64, 256]");
      h->block_threads = value;
      break;
    case PDRD_OPT_TIMING:
      if (value < 0 || value > 2) return set_error(PDRD_EINVAL, "timing is 0, 1 or 2");
      h->timing = value;
      break;
    case PDRD_OPT_REFRESH:
      if (value < 0) return set_error(PDRD_EINVAL, "refresh period must be >= 0");
      h->refresh = value;
      h->since_refresh = 1 << 30;
      break;
    case PDRD_OPT_CHAIN:
      if (value < -1 || value > 1) return set_error(PDRD_EINVAL, "chain is -1 (default), 0 or 1");
      h->chain = value;
      break;
    case PDRD_OPT_PATCH_CELLS:
      if (h->ready) return set_error(PDRD_EINVAL, "cells per patch are set before pdrd_initialize");
      if (value != 0 && value != 64 && value != 128 && value != 256)
        return set_error(PDRD_EINVAL, "cells per patch is 0 (default), 64, 128 or 256");
      h->patch_cells = value;
      break;
    default: return set_error(PDRD_EINVAL, "unknown option");
  }
  drop_graph(h);
  return PDRD_OK;
}

int pdrd_set_stream(pdrd_method_t h, void *stream)
{
  if (!h) return set_error(PDRD_EINVAL, "null method");
  if (h->stream) HIP_CHECK(hipStreamSynchronize(h->stream));
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  h->own_stream = false;
  h->stream = static_cast<hipStream_t>(stream);
  if (!h->stream && h->ready) {
    HIP_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->own_stream = true;
  }
  drop_graph(h);
  return PDRD_OK;
}

int pdrd_synchronize(pdrd_method_t h)
{
  if (int rc = check_ready(h)) return rc;
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return PDRD_OK;
}

int pdrd_set_field(pdrd_method_t h, int field, const double *host)
{
  if (int rc = check_ready(h)) return rc;
  lin_invalidate(h);
  if (!host) return set_error(PDRD_EINVAL, "null buffer");
  size_t n;
  double *d;
  if (int rc = field_size(h, field, &n, &d)) return rc;
  if (field == PDRD_DUAL_VOLUME) return set_error(PDRD_EINVAL, "the dual volume is computed by the method");
  HIP_CHECK(hipMemcpyAsync(d, host, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return PDRD_OK;
}

int pdrd_get_field(pdrd_method_t h, int field, double *host)
{
  if (int rc = check_ready(h)) return rc;
  if (!host) return set_error(PDRD_EINVAL, "null buffer");
  size_t n;
  double *d;
  if (int rc = field_size(h, field, &n, &d)) return rc;
  HIP_CHECK(hipMemcpyAsync(host, d, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return PDRD_OK;
}

int pdrd_field_device_ptr(pdrd_method_t h, int field, double **dptr)
{
  if (int rc = check_ready(h)) return rc;
  if (!dptr) return set_error(PDRD_EINVAL, "null output");
  size_t n;
  return field_size(h, field, &n, dptr);
}

int pdrd_add_weak_bc(pdrd_method_t h, int type, uint32_t n, const uint32_t *nodes, const double *params,
                     int nparams, int quad_order)
{
  PDRD_TRY
  if (!h || (n && !nodes)) return set_error(PDRD_EINVAL, "null argument");
  if (h->ready) return set_error(PDRD_EINVAL, "weak boundary conditions are added before pdrd_initialize");
  if (type < PDRD_WEAK_WALL || type > PDRD_WEAK_DIRICHLET) return set_error(PDRD_EINVAL, "unknown weak BC type");
  if (h->physics != pdrd::PH_EULER2D && h->physics != pdrd::PH_EULER3D && type != PDRD_WEAK_DIRICHLET)
    return set_error(PDRD_ENOTSUP, "only WeakDirichlet applies to scalar physics");
  // parameters: WeakSubInlet T_tot, p_tot; WeakSubOutlet p; WeakFarfield and
  // WeakDirichlet the boundary state; the others none
  int need = 0;
  if (type == PDRD_WEAK_SUBINLET) need = 2;
  if (type == PDRD_WEAK_SUBOUTLET) need = 1;
  if (type == PDRD_WEAK_FARFIELD || type == PDRD_WEAK_DIRICHLET) need = h->m;
  if (nparams < need || (need && !params)) return set_error(PDRD_EINVAL, "missing weak BC parameters");
  if (quad_order < 1) return set_error(PDRD_EINVAL, "quadrature order must be >= 1");
  if (h->scatter_mode != 2) return set_error(PDRD_ENOTSUP, "weak boundary conditions run with the patch assembly only");
  WeakSpec w;
  w.type = type;
  w.quad_order = quad_order;
  for (int i = 0; i < need; ++i) w.params[i] = params[i];
  for (uint32_t i = 0; i < n; ++i) {
    if (nodes[i] >= h->nn) return set_error(PDRD_EINVAL, "weak BC node out of range");
    w.nodes.push_back(nodes[i]);
  }
  h->weak_specs.push_back(std::move(w));
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_weak_bc_info(pdrd_method_t h, int k, int64_t *info3)
{
  if (!h || !info3) return set_error(PDRD_EINVAL, "null argument");
  if (!h->ready) return set_error(PDRD_EINVAL, "not initialized");
  if (k < 0 || k >= static_cast<int>(h->weak.size())) return set_error(PDRD_EINVAL, "no such weak BC");
  info3[0] = h->weak[k].nf;
  info3[1] = h->weak[k].nfn;
  info3[2] = h->weak[k].nq;
  return PDRD_OK;
}

int pdrd_weak_bc_points(pdrd_method_t h, int k, double *xq)
{
  if (!h || !xq) return set_error(PDRD_EINVAL, "null argument");
  if (!h->ready) return set_error(PDRD_EINVAL, "not initialized");
  if (k < 0 || k >= static_cast<int>(h->weak.size())) return set_error(PDRD_EINVAL, "no such weak BC");
  const std::vector<double> &p = h->weak[k].qpoints;
  std::copy(p.begin(), p.end(), xq);
  return PDRD_OK;
}

int pdrd_set_weak_bc_state(pdrd_method_t h, int k, const double *state, size_t count)
{
  if (!h) return set_error(PDRD_EINVAL, "null argument");
  if (!h->ready) return set_error(PDRD_EINVAL, "not initialized");
  lin_invalidate(h);
  if (k < 0 || k >= static_cast<int>(h->weak.size())) return set_error(PDRD_EINVAL, "no such weak BC");
  WeakDev &w = h->weak[k];
  if (w.type != PDRD_WEAK_DIRICHLET) return set_error(PDRD_ENOTSUP, "point states are for WeakDirichlet");
  if (!state) {
    (void)hipFree(w.d_qstate);
    w.d_qstate = nullptr;
    return PDRD_OK;
  }
  const size_t n = static_cast<size_t>(w.nf) * w.nq * h->m;
  if (count != n)
    return set_error(PDRD_EINVAL, "state holds " + std::to_string(count) + " values, the facet points need " +
                                      std::to_string(n) + " ([facets][points][m])");
  if (!w.d_qstate)
    if (int rc = dev_alloc(h, &w.d_qstate, n)) return rc;
  HIP_CHECK(hipMemcpyAsync(w.d_qstate, state, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return PDRD_OK;
}

}  // extern "C"

namespace {

// Per-block tables of the providers, built at their first use.
int stab_tables(pdrd_method *h, Block &b, bool smoothness)
{
  const pdrd::Shape shape = static_cast<pdrd::Shape>(b.fe.shape);
  const int p = b.fe.sol_order;
  const pdrd::PointSet ps = static_cast<pdrd::PointSet>(b.fe.point_set);
  if (b.fe.nd > pdrd::STAB_MAX_ND) return set_error(PDRD_ENOTSUP, "element too large for the providers");
  if (!b.d_p1w) {
    const std::vector<double> p1w = pdrd::vertex_functions(shape, pdrd::reference_nodes(shape, p, ps));
    if (int rc = dev_alloc(h, &b.d_p1w, p1w.size())) return rc;
    HIP_CHECK(hipMemcpy(b.d_p1w, p1w.data(), p1w.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (smoothness && !b.d_av) {
    const pdrd::ElementTables t = pdrd::make_element_tables(shape, 1, p, std::max(2 * p, 1), ps);
    int npm = 0;
    // simplices: an orthonormal lower-degree basis; quadrilaterals: the table
    // of the zeroed modal terms (npm = 0)
    const std::vector<double> psi = (shape == pdrd::Shape::Quad)
                                        ? pdrd::quad_zeroed_modes(p, t.qp, t.sol_nodes)
                                        : pdrd::lower_degree_basis(shape, p, t.qp, t.w, &npm);
    if (t.nq > pdrd::STAB_MAX_NQ || npm > pdrd::STAB_MAX_NPM)
      return set_error(PDRD_ENOTSUP, "smoothness-indicator rule too large");
    std::vector<double> pack;
    pack.insert(pack.end(), t.w.begin(), t.w.end());
    pack.insert(pack.end(), t.V.begin(), t.V.end());
    pack.insert(pack.end(), t.dVg.begin(), t.dVg.end());
    pack.insert(pack.end(), psi.begin(), psi.end());
    if (int rc = dev_alloc(h, &b.d_av, pack.size())) return rc;
    HIP_CHECK(hipMemcpy(b.d_av, pack.data(), pack.size() * sizeof(double), hipMemcpyHostToDevice));
    b.av_nq = t.nq;
    b.av_npm = npm;
  }
  return PDRD_OK;
}

pdrd::StabArgs stab_args(const pdrd_method *h, const Block &b)
{
  pdrd::StabArgs a{};
  a.sconn = b.d_sconn;
  a.gconn = b.d_gconn;
  a.coords = h->d_coords;
  a.u = h->d_u;
  a.p1w = b.d_p1w;
  a.ne = b.ne;
  a.m = h->m;
  a.nd = b.fe.nd;
  a.ng = b.fe.ng;
  a.dim = h->dim;
  a.nvert = pdrd::shape_nb_vertices(static_cast<pdrd::Shape>(b.fe.shape));
  a.p = b.fe.sol_order;
  if (b.d_av) {
    a.nq = b.av_nq;
    a.npm = b.av_npm;
    a.w = b.d_av;
    a.V = a.w + a.nq;
    a.dVg = a.V + static_cast<size_t>(a.nq) * a.nd;
    a.psi = a.dVg + static_cast<size_t>(h->dim) * a.nq * a.ng;
    if (b.fe.shape == PDRD_QUAD) a.hm = a.psi;  // the zeroed modal terms [nq][nd]
  }
  return a;
}

int stab_fill(pdrd_method *h, double *d, double v)
{
  hipLaunchKernelGGL(pdrd::stab_fill_kernel, dim3(grid_for(h->nn, 256)), dim3(256), 0, h->stream, d, h->nn, v);
  HIP_CHECK(hipGetLastError());
  return PDRD_OK;
}

// P1 / Q1 interpolation of the vertex values in `out` to the other nodes.
int stab_interp(pdrd_method *h, double *out)
{
  for (Block &b : h->blocks) {
    if (b.fe.sol_order == 1) continue;
    pdrd::StabArgs a = stab_args(h, b);
    a.out = out;
    hipLaunchKernelGGL(pdrd::p1_interp_kernel, dim3(grid_for(b.ne, 256)), dim3(256), 0, h->stream, a);
    HIP_CHECK(hipGetLastError());
  }
  return PDRD_OK;
}

}  // namespace

extern "C" {

int pdrd_set_blending_coeff(pdrd_method_t h, const double *theta)
{
  if (int rc = check_ready(h)) return rc;
  lin_invalidate(h);
  if (!theta) {
    if (h->use_blend) drop_graph(h);
    h->use_blend = false;
    return PDRD_OK;
  }
  if (!h->d_blend)
    if (int rc = dev_alloc(h, &h->d_blend, h->nn)) return rc;
  HIP_CHECK(hipMemcpyAsync(h->d_blend, theta, h->nn * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  if (!h->use_blend) drop_graph(h);
  h->use_blend = true;
  if (int rc = check_inputs(h)) {
    h->use_blend = false;
    return rc;
  }
  return PDRD_OK;
}

int pdrd_set_artificial_viscosity(pdrd_method_t h, const double *nu)
{
  if (int rc = check_ready(h)) return rc;
  lin_invalidate(h);
  if (!nu) {
    if (h->use_visc) drop_graph(h);
    h->use_visc = false;
    return PDRD_OK;
  }
  if (!h->d_visc)
    if (int rc = dev_alloc(h, &h->d_visc, h->nn)) return rc;
  HIP_CHECK(hipMemcpyAsync(h->d_visc, nu, h->nn * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  if (!h->use_visc) drop_graph(h);
  h->use_visc = true;
  return PDRD_OK;
}

int pdrd_set_sources(pdrd_method_t h, const double *sources)
{
  if (int rc = check_ready(h)) return rc;
  lin_invalidate(h);
  if (!sources) {
    h->use_src = false;
    return PDRD_OK;
  }
  if (h->scatter_mode != 2) return set_error(PDRD_ENOTSUP, "the source term runs with the patch assembly only");
  h->use_src = true;
  const int chk = check_inputs(h);
  h->use_src = false;
  if (chk) return chk;
  const size_t nm = static_cast<size_t>(h->nn) * h->m;
  if (!h->d_src)
    if (int rc = dev_alloc(h, &h->d_src, nm)) return rc;
  HIP_CHECK(hipMemcpyAsync(h->d_src, sources, nm * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  h->use_src = true;
  return PDRD_OK;
}

int pdrd_compute_blending_coeff(pdrd_method_t h, int kind)
{
  PDRD_TRY
  if (int rc = check_ready(h)) return rc;
  lin_invalidate(h);
  if (kind != PDRD_BLEND_LF) return set_error(PDRD_EINVAL, "unknown blending coefficient");
  for (Block &b : h->blocks)
    if (int rc = stab_tables(h, b, false)) return rc;
  if (!h->d_blend)
    if (int rc = dev_alloc(h, &h->d_blend, h->nn)) return rc;
  if (!h->d_stab_tmp)
    if (int rc = dev_alloc(h, &h->d_stab_tmp, h->nn)) return rc;
  if (int rc = stab_fill(h, h->d_stab_tmp, 0.0)) return rc;
  if (int rc = stab_fill(h, h->d_blend, 0.0)) return rc;
  for (Block &b : h->blocks) {
    pdrd::StabArgs a = stab_args(h, b);
    a.out = h->d_stab_tmp;
    hipLaunchKernelGGL(pdrd::lf_indicator_kernel, dim3(grid_for(b.ne, 256)), dim3(256), 0, h->stream, a);
    HIP_CHECK(hipGetLastError());
  }
  for (Block &b : h->blocks) {
    pdrd::StabArgs a = stab_args(h, b);
    a.in = h->d_stab_tmp;
    a.out = h->d_blend;
    hipLaunchKernelGGL(pdrd::lf_cell_kernel, dim3(grid_for(b.ne, 256)), dim3(256), 0, h->stream, a);
    HIP_CHECK(hipGetLastError());
  }
  if (int rc = stab_interp(h, h->d_blend)) return rc;
  if (!h->use_blend) drop_graph(h);
  h->use_blend = true;
  if (int rc = check_inputs(h)) {
    h->use_blend = false;
    return rc;
  }
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_compute_artificial_viscosity(pdrd_method_t h, double kappa, double factor)
{
  PDRD_TRY
  if (int rc = check_ready(h)) return rc;
  lin_invalidate(h);
  if (!(kappa > 0.0) || !(factor >= 0.0)) return set_error(PDRD_EINVAL, "kappa must be > 0 and factor >= 0");
  for (Block &b : h->blocks)
    if (int rc = stab_tables(h, b, true)) return rc;
  if (!h->d_visc)
    if (int rc = dev_alloc(h, &h->d_visc, h->nn)) return rc;
  if (int rc = stab_fill(h, h->d_visc, 0.0)) return rc;
  for (Block &b : h->blocks) {
    pdrd::StabArgs a = stab_args(h, b);
    a.out = h->d_visc;
    a.kappa = kappa;
    a.factor = factor;
    hipLaunchKernelGGL(pdrd::av_cell_kernel, dim3(grid_for(b.ne, 256)), dim3(256), 0, h->stream, a);
    HIP_CHECK(hipGetLastError());
  }
  if (int rc = stab_interp(h, h->d_visc)) return rc;
  if (!h->use_visc) drop_graph(h);
  h->use_visc = true;
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_element_jacobians(pdrd_method_t h, int block, double *phi, double *jac, double *wave)
{
  PDRD_TRY
  if (int rc = check_ready(h)) return rc;
  if (block < 0 || block >= static_cast<int>(h->blocks.size())) return set_error(PDRD_EINVAL, "no such block");
  Block &b = h->blocks[block];
  if (!b.ks.jacobian) return set_error(PDRD_ENOTSUP, "the PSI scheme has no implicit form");
  if (b.ne == 0) return PDRD_OK;
  const size_t ndm = static_cast<size_t>(b.fe.nd) * h->m;
  double *d_phi = nullptr, *d_jac = nullptr, *d_stab = nullptr, *d_wave = nullptr;
  uint32_t *d_perm = nullptr;
  auto release = [&]() {
    (void)hipFree(d_phi);
    (void)hipFree(d_jac);
    (void)hipFree(d_stab);
    (void)hipFree(d_wave);
    (void)hipFree(d_perm);
  };
  if (hipMalloc(&d_phi, b.ne * ndm * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_jac, b.ne * ndm * ndm * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_stab, (h->scheme == pdrd::SC_B ? b.ne * ndm * ndm : 1) * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_wave, b.ne * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_perm, b.ne * sizeof(uint32_t)) != hipSuccess) {
    release();
    return set_error(PDRD_ENOMEM, "device memory for the element Jacobians");
  }
  hipError_t e = hipMemcpyAsync(d_perm, b.perm.data(), b.ne * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream);
  pdrd::JacArgs a{};
  a.tab = b.d_tab;
  a.gconn = b.d_gconn;
  a.sconn = b.d_sconn;
  a.perm = d_perm;
  a.coords = h->d_coords;
  a.u = h->d_u;
  a.nodal = nodal_input(h);
  a.phi = d_phi;
  a.jac = d_jac;
  a.jstab = d_stab;
  a.wave = d_wave;
  a.status = h->d_status;
  a.ne = b.ne;
  a.pp.p0 = h->params[0];
  a.pp.p1 = h->params[1];
  if (e == hipSuccess) e = hipEventRecord(h->ev[0], h->stream);
  if (e == hipSuccess) e = b.ks.jacobian(a, grid_for(b.ne, 64), h->stream);
  if (e == hipSuccess) e = hipEventRecord(h->ev[1], h->stream);
  if (e == hipSuccess && phi)
    e = hipMemcpyAsync(phi, d_phi, b.ne * ndm * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess && jac)
    e = hipMemcpyAsync(jac, d_jac, b.ne * ndm * ndm * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess && wave)
    e = hipMemcpyAsync(wave, d_wave, b.ne * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
  release();
  if (e != hipSuccess) return hip_fail(e, "element Jacobians");
  h->last_cell_ms = ms;  // the Jacobian kernel alone
  h->last_node_ms = 0.0;
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_add_strong_dirichlet(pdrd_method_t h, uint32_t n, const uint32_t *nodes, const double *values)
{
  if (int rc = check_ready(h)) return rc;
  lin_invalidate(h);
  if (n == 0) return PDRD_OK;
  if (!nodes || !values) return set_error(PDRD_EINVAL, "null argument");
  for (uint32_t i = 0; i < n; ++i)
    if (nodes[i] >= h->nn) return set_error(PDRD_EINVAL, "Dirichlet node out of range");
  // a node listed again takes its latest values (boundary conditions are
  // applied in the order they were added)
  std::vector<int> row(h->nn, -1);
  for (size_t i = 0; i < h->dir_nodes_host.size(); ++i) row[h->dir_nodes_host[i]] = static_cast<int>(i);
  for (uint32_t i = 0; i < n; ++i) {
    int r = row[nodes[i]];
    if (r < 0) {
      r = static_cast<int>(h->dir_nodes_host.size());
      row[nodes[i]] = r;
      h->dir_nodes_host.push_back(nodes[i]);
      h->dir_vals_host.resize(h->dir_vals_host.size() + h->m);
    }
    for (int k = 0; k < h->m; ++k)
      h->dir_vals_host[static_cast<size_t>(r) * h->m + k] = values[static_cast<size_t>(i) * h->m + k];
  }
  (void)hipFree(h->d_dir_nodes);
  (void)hipFree(h->d_dir_vals);
  h->d_dir_nodes = nullptr;
  h->d_dir_vals = nullptr;
  int rc;
  if ((rc = dev_alloc(h, &h->d_dir_nodes, h->dir_nodes_host.size()))) return rc;
  if ((rc = dev_alloc(h, &h->d_dir_vals, h->dir_vals_host.size()))) return rc;
  HIP_CHECK(hipMemcpy(h->d_dir_nodes, h->dir_nodes_host.data(), h->dir_nodes_host.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(h->d_dir_vals, h->dir_vals_host.data(), h->dir_vals_host.size() * sizeof(double),
                      hipMemcpyHostToDevice));
  h->n_dir = static_cast<uint32_t>(h->dir_nodes_host.size());
  std::vector<unsigned char> flag(h->nn, 0);
  for (uint32_t v : h->dir_nodes_host) flag[v] = 1;
  HIP_CHECK(hipMemcpy(h->d_is_dir, flag.data(), h->nn, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(h->d_dir_index, row.data(), h->nn * sizeof(int), hipMemcpyHostToDevice));
  drop_graph(h);
  return PDRD_OK;
}

int pdrd_assemble_rhs(pdrd_method_t h)
{
  if (int rc = check_ready(h)) return rc;
  lin_invalidate(h);
  HIP_CHECK(hipEventRecord(h->ev[0], h->stream));
  if (h->scatter_mode == 2) {
    if (int rc = launch_patches(h, 0, 0, 0, 0.0)) return rc;
    HIP_CHECK(hipEventRecord(h->ev[1], h->stream));
  } else {
    if (int rc = launch_cells(h)) return rc;
    HIP_CHECK(hipEventRecord(h->ev[1], h->stream));
    hipLaunchKernelGGL(finalize_kernel, dim3(grid_for(h->nn, 256)), dim3(256), 0, h->stream, h->nn, h->m,
                       h->d_acc_res, h->d_acc_ws, h->d_vol, h->d_res, h->d_ws);
    HIP_CHECK(hipGetLastError());
  }
  if (int rc = launch_dirichlet(h)) return rc;
  HIP_CHECK(hipEventRecord(h->ev[2], h->stream));
  HIP_CHECK(hipEventSynchronize(h->ev[2]));
  float a = 0.f, b = 0.f;
  HIP_CHECK(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
  HIP_CHECK(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
  h->last_cell_ms = a;
  h->last_node_ms = b;
  return PDRD_OK;
}

int pdrd_assemble_lhs_and_rhs(pdrd_method_t h, double cfl)
{
  if (int rc = pdrd_assemble_rhs(h)) return rc;
  hipLaunchKernelGGL(dt_kernel, dim3(grid_for(h->nn, 256)), dim3(256), 0, h->stream, h->nn, cfl, h->d_vol, h->d_ws,
                     h->d_dt);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return PDRD_OK;
}

int pdrd_solve(pdrd_method_t h)
{
  if (int rc = check_ready(h)) return rc;
  lin_invalidate(h);
  hipLaunchKernelGGL(update_kernel, dim3(grid_for(h->nn, 256)), dim3(256), 0, h->stream, h->nn, h->m, h->d_dt,
                     h->d_res, h->d_u);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return PDRD_OK;
}

int pdrd_compute_residual_norm(pdrd_method_t h, double *norm)
{
  if (int rc = check_ready(h)) return rc;
  if (!norm) return set_error(PDRD_EINVAL, "null output");
  return norm_impl(h, h->d_dt, 0, norm);
}

int pdrd_residual_norm_L2(pdrd_method_t h, int exclude_dirichlet, double *norm)
{
  if (int rc = check_ready(h)) return rc;
  if (!norm) return set_error(PDRD_EINVAL, "null output");
  return norm_impl(h, nullptr, exclude_dirichlet ? 1 : 0, norm);
}

int pdrd_iterate(pdrd_method_t h, double cfl, int n_iter, int exclude_dirichlet, double *norm)
{
  if (int rc = check_ready(h)) return rc;
  if (n_iter < 1) return set_error(PDRD_EINVAL, "n_iter must be >= 1");
  const int excl = exclude_dirichlet ? 1 : 0;
  if (h->scatter_mode == 2) return iterate_patches(h, cfl, n_iter, excl, norm);
  if (h->timing) {
    // eager iterations with events around the cell loop and the nodal pass
    double cell = 0.0, node = 0.0;
    for (int it = 0; it < n_iter; ++it) {
      const int out = (it == n_iter - 1) ? 1 : 0;
      HIP_CHECK(hipEventRecord(h->ev[0], h->stream));
      if (int rc = launch_cells(h)) return rc;
      HIP_CHECK(hipEventRecord(h->ev[1], h->stream));
      if (int rc = launch_dirichlet(h)) return rc;
      if (int rc = launch_node_pass(h, cfl, out, excl)) return rc;
      HIP_CHECK(hipEventRecord(h->ev[2], h->stream));
      HIP_CHECK(hipEventSynchronize(h->ev[2]));
      float a = 0.f, b = 0.f;
      HIP_CHECK(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
      HIP_CHECK(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
      cell += a;
      node += b;
    }
    h->last_cell_ms = cell / n_iter;
    h->last_node_ms = node / n_iter;
    if (norm) return iterate_norm(h, norm);
    return PDRD_OK;
  }
  if (h->precision) return set_error(PDRD_ENOTSUP, "the mixed-precision form runs with the patch assembly only");
  HIP_CHECK(hipEventRecord(h->ev[0], h->stream));
  if (n_iter > 1) {
    if (!h->graph || h->graph_cfl != cfl) {
      drop_graph(h);
      hipGraph_t g = nullptr;
      HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_iteration(h, cfl, 0, excl);
      hipError_t e = hipStreamEndCapture(h->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
      h->graph_cfl = cfl;
    }
    for (int it = 0; it < n_iter - 1; ++it) HIP_CHECK(hipGraphLaunch(h->graph, h->stream));
  }
  if (int rc = enqueue_iteration(h, cfl, 1, excl)) return rc;
  HIP_CHECK(hipEventRecord(h->ev[1], h->stream));
  HIP_CHECK(hipEventSynchronize(h->ev[1]));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
  h->last_cell_ms = ms;
  h->last_node_ms = 0.0;
  if (norm) return iterate_norm(h, norm);
  return PDRD_OK;
}

int pdrd_check_status(pdrd_method_t h)
{
  if (int rc = check_ready(h)) return rc;
  int s = 0;
  HIP_CHECK(hipStreamSynchronize(h->stream));
  HIP_CHECK(hipMemcpy(&s, h->d_status, sizeof(int), hipMemcpyDeviceToHost));
  if (s & 1) return set_error(PDRD_ENEGPRESSURE, "non-positive pressure met in a cell");
  return PDRD_OK;
}

int pdrd_method_info(pdrd_method_t h, int64_t *info)
{
  if (!h || !info) return set_error(PDRD_EINVAL, "null argument");
  int64_t cells = 0, ncol = 0;
  for (const Block &b : h->blocks) {
    cells += b.ne;
    ncol = std::max<int64_t>(ncol, b.colour_start.empty() ? 0 : static_cast<int64_t>(b.colour_start.size()) - 1);
  }
  info[0] = h->nn;
  info[1] = h->m;
  info[2] = static_cast<int64_t>(h->blocks.size());
  info[3] = cells;
  info[4] = ncol;
  info[5] = static_cast<int64_t>(h->bytes);
  return PDRD_OK;
}

int pdrd_get_option(pdrd_method_t h, int option, int *value)
{
  if (!h || !value) return set_error(PDRD_EINVAL, "null argument");
  switch (option) {
    case PDRD_OPT_SCATTER: *value = h->scatter_mode; return PDRD_OK;
    case PDRD_OPT_PRECISION: *value = h->precision; return PDRD_OK;
    case PDRD_OPT_BLOCK: *value = h->block_threads; return PDRD_OK;
    case PDRD_OPT_TIMING: *value = h->timing; return PDRD_OK;
    case PDRD_OPT_REFRESH: *value = h->refresh; return PDRD_OK;
    case PDRD_OPT_CHAIN: *value = chains(h) ? 1 : 0; return PDRD_OK;
    case PDRD_OPT_PATCH_CELLS: {
      int c = 0;
      for (const Block &b : h->blocks) c = std::max(c, b.cells);
      *value = h->ready ? c : h->patch_cells;
      return PDRD_OK;
    }
  }
  return set_error(PDRD_EINVAL, "unknown option");
}

int pdrd_last_timing(pdrd_method_t h, double *cell_ms, double *node_ms)
{
  if (!h) return set_error(PDRD_EINVAL, "null method");
  if (cell_ms) *cell_ms = h->last_cell_ms;
  if (node_ms) *node_ms = h->last_node_ms;
  return PDRD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: the iterations of several methods in shared launches.
// ---------------------------------------------------------------------------
struct pdrd_batch
{
  int device = -1;
  hipStream_t stream = nullptr;
  pdrd::PatchArgs *d_pargs = nullptr;
  pdrd::FixupArgs *d_fargs = nullptr;
  uint32_t *d_first = nullptr;
  size_t pargs_cap = 0, fargs_cap = 0, first_cap = 0;
  int graphs = 0;  // 1: replay the iterations of a call from a captured HIP graph
};

namespace {

template <typename T>
int grow(T **p, size_t *cap, size_t need)
{
  if (need <= *cap) return PDRD_OK;
  (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  const size_t n = std::max<size_t>(need, 64);
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T));
  if (e != hipSuccess) return set_error(PDRD_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  *cap = n;
  return PDRD_OK;
}

// One launch of the batch: a kernel and its slice of the argument arrays.
struct BatchLaunch
{
  pdrd::PatchBatchLauncher patch = nullptr;  // null: shared-node sums
  int m = 0;                                 // equations (shared-node sums)
  int cells = 0;                             // threads per workgroup of a patch launch
  size_t args0 = 0, first0 = 0;              // offsets in the argument / prefix arrays
  int count = 0;
  unsigned grid = 0;
  size_t lds = 0;
};

// Kinds of batched sweeps. A method whose iterations chain (and only outside
// graph capture) takes the place in the chain of the kind; any other method
// runs a complete sweep (fix-up included) of every kind, with outputs in the
// last.
enum BatchKind
{
  BK_OPEN = 0,     // first of several sweeps
  BK_MID_W1 = 1,   // deferred sweep writing buffer 1
  BK_MID_W0 = 2,   // deferred sweep writing buffer 0
  BK_LAST_W1 = 3,  // closing sweep writing buffer 1
  BK_LAST_W0 = 4,  // closing sweep writing buffer 0
  BK_SINGLE = 5,   // the only sweep of a call
  BK_COUNT = 6
};

struct KindPlace
{
  Chain c;
  bool close = true;
  int outputs = 0;
};

KindPlace kind_place(int kind, bool chain)
{
  KindPlace k;
  const bool last = kind == BK_LAST_W1 || kind == BK_LAST_W0 || kind == BK_SINGLE;
  k.outputs = last ? 1 : 0;
  if (!chain) return k;  // complete sweep into buffer 0
  switch (kind) {
    case BK_OPEN: k.c.write = 0; k.c.record = true; k.close = false; break;
    case BK_MID_W1: k.c.pend = 0; k.c.write = 1; k.c.record = true; k.close = false; break;
    case BK_MID_W0: k.c.pend = 1; k.c.write = 0; k.c.record = true; k.close = false; break;
    case BK_LAST_W1: k.c.pend = 0; k.c.write = 1; break;
    case BK_LAST_W0: k.c.pend = 1; k.c.write = 0; break;
    default: break;
  }
  return k;
}

// One batched sweep of kind `kind`: the patch launches (arguments of the kind
// at kind * args_per_kind), the weak facet kernels of every method, the
// shared-node sums of the methods that close in this kind.
int enqueue_batch(pdrd_batch *b, const std::vector<BatchLaunch> &patches, const std::vector<BatchLaunch> &fixups,
                  size_t args_per_kind, int kind, int count, const pdrd_method_t *methods, const bool *chain)
{
  for (const BatchLaunch &l : patches) {
    if (l.grid == 0) continue;
    HIP_CHECK(l.patch(b->d_pargs + kind * args_per_kind + l.args0, b->d_first + l.first0, l.count, l.grid, l.cells,
                      l.lds, b->stream));
  }
  for (int i = 0; i < count; ++i)
    if (int rc = launch_weak(methods[i], b->stream, kind_place(kind, chain[i]).c.write)) return rc;
  for (const BatchLaunch &l : fixups) {
    if (l.grid == 0) continue;
    const pdrd::FixupArgs *fa = b->d_fargs + l.args0;
#define FIXB_CASE(MM)                                                                                            \
  case MM:                                                                                                       \
    hipLaunchKernelGGL(pdrd::patch_fixup_batch_kernel<MM>, dim3(l.grid), dim3(256), 0, b->stream, fa,             \
                       b->d_first + l.first0, l.count);                                                         \
    break;
    switch (l.m) {
      FIXB_CASE(1)
      FIXB_CASE(4)
      FIXB_CASE(5)
      default: return set_error(PDRD_ENOTSUP, "unsupported number of equations");
    }
#undef FIXB_CASE
    HIP_CHECK(hipGetLastError());
  }
  return PDRD_OK;
}

}  // namespace

extern "C" {

int pdrd_batch_create(pdrd_batch_t *out)
{
  PDRD_TRY
  if (!out) return set_error(PDRD_EINVAL, "null argument");
  pdrd_batch *b = new pdrd_batch();
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return set_error(PDRD_EHIP, std::string("pdrd_batch_create: ") + hipGetErrorString(e));
  }
  *out = b;
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_batch_destroy(pdrd_batch_t b)
{
  if (!b) return PDRD_OK;
  (void)hipStreamSynchronize(b->stream);
  (void)hipFree(b->d_pargs);
  (void)hipFree(b->d_fargs);
  (void)hipFree(b->d_first);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return PDRD_OK;
}

int pdrd_batch_set_graphs(pdrd_batch_t b, int on)
{
  if (!b) return set_error(PDRD_EINVAL, "null batch");
  if (on != 0 && on != 1) return set_error(PDRD_EINVAL, "graphs is 0 or 1");
  b->graphs = on;
  return PDRD_OK;
}

int pdrd_batch_iterate(pdrd_batch_t b, int count, const pdrd_method_t *methods, const double *cfl, int n_iter,
                       int exclude_dirichlet, double *norms)
{
  PDRD_TRY
  if (!b || !methods || !cfl || count < 1) return set_error(PDRD_EINVAL, "null or empty batch");
  if (n_iter < 1) return set_error(PDRD_EINVAL, "n_iter must be >= 1");
  const int excl = exclude_dirichlet ? 1 : 0;
  for (int i = 0; i < count; ++i) {
    pdrd_method *h = methods[i];
    if (int rc = check_ready(h)) return rc;
    if (h->device != b->device) return set_error(PDRD_EINVAL, "batch methods must live on the batch's device");
    if (h->scatter_mode != 2) return set_error(PDRD_ENOTSUP, "batched iterations need the patch assembly");
    if (mixed(h)) return set_error(PDRD_ENOTSUP, "batched iterations of mixed-precision methods are not built");
    for (int j = 0; j < i; ++j)
      if (methods[j] == h) return set_error(PDRD_EINVAL, "a method appears twice in one batch");
  }
  for (int i = 0; i < count; ++i) HIP_CHECK(hipStreamSynchronize(methods[i]->stream));

  // sweeps of methods whose iterations chain defer their shared-node sums
  // (direct launches only: the graph replays one kind of sweep)
  std::unique_ptr<bool[]> chain_v(new bool[count]);
  for (int i = 0; i < count; ++i) chain_v[i] = !b->graphs && chains(methods[i]);
  const bool *chain = chain_v.get();

  // group the blocks of all methods by kernel, in order of first appearance
  std::vector<BatchLaunch> patches;
  std::vector<std::vector<std::pair<int, size_t>>> members;  // (method, block) per launch
  for (int i = 0; i < count; ++i)
    for (size_t k = 0; k < methods[i]->blocks.size(); ++k) {
      const Block &bl = methods[i]->blocks[k];
      // the kernel of the method's inputs: plain, with a nodal input, or with a source field
      const pdrd_method *hm = methods[i];
      const pdrd::PatchBatchLauncher fn =
          hm->use_src ? (source_theta(hm) ? bl.ks.patch_src_theta_batch : bl.ks.patch_src_batch)
                      : (nodal_input(hm) ? bl.ks.patch_nodal_batch : bl.ks.patch_batch);
      if (!fn) return set_error(PDRD_ENOTSUP, "no batched kernel for this method's inputs");
      size_t g = 0;
      while (g < patches.size() && (patches[g].patch != fn || patches[g].cells != bl.cells)) ++g;
      if (g == patches.size()) {
        BatchLaunch l;
        l.patch = fn;
        l.cells = bl.cells;
        patches.push_back(l);
        members.emplace_back();
      }
      members[g].emplace_back(i, k);
    }
  // patch arguments of every kind ([kind][launch members]) and prefix sums of
  // workgroups (shared by the kinds)
  std::vector<pdrd::PatchArgs> pargs;
  std::vector<uint32_t> first;
  std::vector<pdrd::PatchArgs> pk[BK_COUNT];
  for (size_t g = 0; g < patches.size(); ++g) {
    BatchLaunch &l = patches[g];
    l.first0 = first.size();
    l.args0 = pk[0].size();
    l.count = static_cast<int>(members[g].size());
    uint32_t acc = 0;
    first.push_back(0);
    for (const auto &mk : members[g]) {
      pdrd_method *h = methods[mk.first];
      const Block &bl = h->blocks[mk.second];
      for (int k = 0; k < BK_COUNT; ++k) {
        const KindPlace kp = kind_place(k, chain[mk.first]);
        pk[k].push_back(patch_args(h, &bl, 1, kp.outputs, excl, cfl[mk.first], kp.c));
      }
      l.lds = std::max(l.lds, input_lds_bytes(h, bl));
      acc += bl.np;
      first.push_back(acc);
    }
    l.grid = acc;
  }
  const size_t ph = pk[0].size();
  for (int k = 0; k < BK_COUNT; ++k) pargs.insert(pargs.end(), pk[k].begin(), pk[k].end());
  // shared-node sums of every kind: one launch per number of equations over
  // the methods that close in that kind
  std::vector<BatchLaunch> fixups[BK_COUNT];
  std::vector<pdrd::FixupArgs> fargs;
  for (int k = 0; k < BK_COUNT; ++k)
    for (int mm : {1, 4, 5}) {
      BatchLaunch l;
      l.m = mm;
      l.first0 = first.size();
      l.args0 = fargs.size();
      uint32_t acc = 0;
      first.push_back(0);
      for (int i = 0; i < count; ++i) {
        pdrd_method *h = methods[i];
        const KindPlace kp = kind_place(k, chain[i]);
        if (h->m != mm || h->n_shared == 0 || !kp.close) continue;
        fargs.push_back(fixup_args(h, 1, kp.outputs, excl, cfl[i], kp.c.write));
        acc += fixup_blocks(h);
        first.push_back(acc);
        ++l.count;
      }
      if (l.count == 0) {
        first.pop_back();
        continue;
      }
      l.grid = acc;
      fixups[k].push_back(l);
    }
  if (int rc = grow(&b->d_pargs, &b->pargs_cap, pargs.size())) return rc;
  if (int rc = grow(&b->d_fargs, &b->fargs_cap, fargs.size())) return rc;
  if (int rc = grow(&b->d_first, &b->first_cap, first.size())) return rc;
  if (!pargs.empty())
    HIP_CHECK(hipMemcpy(b->d_pargs, pargs.data(), pargs.size() * sizeof(pdrd::PatchArgs), hipMemcpyHostToDevice));
  if (!fargs.empty())
    HIP_CHECK(hipMemcpy(b->d_fargs, fargs.data(), fargs.size() * sizeof(pdrd::FixupArgs), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(b->d_first, first.data(), first.size() * sizeof(uint32_t), hipMemcpyHostToDevice));

  auto sweep = [&](int kind) { return enqueue_batch(b, patches, fixups[kind], ph, kind, count, methods, chain); };
  if (n_iter == 1) {
    if (int rc = sweep(BK_SINGLE)) return rc;
  } else if (!b->graphs) {
    if (int rc = sweep(BK_OPEN)) return rc;
    for (int it = 1; it < n_iter - 1; ++it)
      if (int rc = sweep((it & 1) ? BK_MID_W1 : BK_MID_W0)) return rc;
    if (int rc = sweep(((n_iter - 1) & 1) ? BK_LAST_W1 : BK_LAST_W0)) return rc;
  } else {
    // no method chains here: every kind is a complete sweep, the last with outputs
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = sweep(BK_OPEN);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    for (int it = 0; it < n_iter - 1; ++it) {
      e = hipGraphLaunch(ge, b->stream);
      if (e != hipSuccess) break;
    }
    (void)hipStreamSynchronize(b->stream);
    (void)hipGraphExecDestroy(ge);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    if (int rc2 = sweep(BK_SINGLE)) return rc2;
  }
  HIP_CHECK(hipStreamSynchronize(b->stream));
  if (norms) {
    size_t off = 0;
    for (int i = 0; i < count; ++i) {
      if (int rc = iterate_norm(methods[i], norms + off)) return rc;
      off += methods[i]->m;
    }
  }
  return PDRD_OK;
  PDRD_CATCH
}

}  // extern "C"
