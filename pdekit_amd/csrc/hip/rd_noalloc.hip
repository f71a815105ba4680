// pdrd_assemble_rhs_noalloc: the cell loop of pdrd_assemble_rhs (FP64
// atomics, PDRD_OPT_SCATTER = 0) over caller-owned device arrays, for callers
// that cannot allocate device memory. Its scratch -- the device layout of the
// FE tables, the nodal accumulators and the status word -- lives in one
// file-scope device array; the library allocates nothing per call and the
// entry point calls no allocating routine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "../../../include/pdekit_rd.h"
#include "../host/capi_common.hpp"
#include "rd_launch.hpp"

using pdrd::set_error;

namespace {

// Scratch layout, in doubles: [0, kTabDoubles) FE tables in the device
// layout (TabOffsets<...>::dev_size doubles), then one 256-byte slot for
// the status word, then the accumulators: residual sums [nn][m] followed by
// wave-speed sums [nn].
constexpr size_t kScratchBytes = PDRD_NOALLOC_SCRATCH_BYTES;
constexpr size_t kTabDoubles = 8192;
constexpr size_t kStatusDoubles = 32;
constexpr size_t kAccDoubles = kScratchBytes / sizeof(double) - kTabDoubles - kStatusDoubles;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double *g_ptr[kMaxDevices] = {};

// The device address of this device's copy of `scratch` (each device holds its
// own), looked up once per device with that device current.
int scratch_ptr(int dev, double **out)
{
  if (dev < 0 || dev >= kMaxDevices) return set_error(PDRD_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(PDRD_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_ptr[dev];
  return PDRD_OK;
}

// Zeroes n doubles (a fill kernel rather than hipMemsetAsync, so the call can
// be captured into a graph).
__global__ void clear_kernel(double *__restrict__ p, size_t n)
{
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = 0.0;
}

// The device layout of the FE tables (TabOffsets in rd_cell.hpp): the packed
// tables as pdrd_fe_fill writes them, then dVf as [nf][dim][nq], dV as
// [nq][dim][nd], dVg as [nq][dim][ng]. One thread per output double.
__global__ void table_layout_kernel(const double *__restrict__ t, int D, int NG, int ND, int NF, int NQ, int size,
                                    double *__restrict__ out)
{
  const int o_dVg = NQ * D + NQ + NQ * NG;
  const int o_V = o_dVg + D * NQ * NG;
  const int o_dV = o_V + NQ * ND;
  const int o_Vf = o_dV + D * NQ * ND;
  const int o_dVf = o_Vf + NQ * NF;
  const int o_dVfT = size, o_dVT = o_dVfT + NF * D * NQ, o_dVgT = o_dVT + NQ * D * ND;
  const int dev_size = o_dVgT + NQ * D * NG;
  auto value = [&](int i) -> double {
    if (i < size) return t[i];
    if (i < o_dVT) {
      const int j = i - o_dVfT, f = j / (D * NQ), xi = (j / NQ) % D, q = j % NQ;
      return t[o_dVf + (xi * NQ + q) * NF + f];
    }
    if (i < o_dVgT) {
      const int j = i - o_dVT, q = j / (D * ND), xi = (j / ND) % D, n = j % ND;
      return t[o_dV + (xi * NQ + q) * ND + n];
    }
    const int j = i - o_dVgT, q = j / (D * NG), xi = (j / NG) % D, g = j % NG;
    return t[o_dVg + (xi * NQ + q) * NG + g];
  };
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < dev_size) out[i] = value(i);
}

// residual = acc / V, wave speed = acc_ws; the status word to the caller.
__global__ void finish_kernel(uint32_t nn, int m, const double *__restrict__ acc_res, const double *__restrict__ acc_ws,
                              const double *__restrict__ vol, double *__restrict__ res, double *__restrict__ ws,
                              const int *__restrict__ status, int *__restrict__ status_out)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (status_out && i == 0) *status_out = *status;
  if (i >= nn) return;
  const double iv = 1.0 / vol[i];
  for (int k = 0; k < m; ++k) {
    const size_t j = static_cast<size_t>(i) * m + k;
    res[j] = acc_res[j] * iv;
  }
  ws[i] = acc_ws[i];
}

unsigned blocks_for(size_t n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

}  // namespace

extern "C" {

int pdrd_assemble_rhs_noalloc(int physics, int scheme, const double *params, int nb_params, const pdrd_fe_desc *fe,
                              const double *tables, uint32_t ne, const uint32_t *geo_conn, const uint32_t *sol_conn,
                              uint32_t nb_vertices, const double *vertex_coords, uint32_t nb_nodes, const double *u,
                              const double *dual_volume, double *residual, double *wave_speed, int *status,
                              void *stream)
{
  if (!fe || !tables || !vertex_coords || !u || !dual_volume || !residual || !wave_speed ||
      (ne && (!geo_conn || !sol_conn)))
    return set_error(PDRD_EINVAL, "null argument");
  if (physics < 0 || physics > 4) return set_error(PDRD_EINVAL, "unknown physics");
  if (scheme < 0 || scheme > 4) return set_error(PDRD_EINVAL, "unknown scheme");
  if (nb_params < 0 || (nb_params > 0 && !params)) return set_error(PDRD_EINVAL, "bad parameters");
  const int dim = (physics == PDRD_EULER3D) ? 3 : 2;
  if (fe->dim != dim) return set_error(PDRD_EINVAL, "element dimension does not match the physics");
  if (nb_nodes == 0 || nb_vertices == 0) return set_error(PDRD_EINVAL, "empty mesh");
  pdrd_fe_desc chk;
  if (int rc = pdrd_fe_describe(fe->shape, fe->geo_order, fe->sol_order, fe->quad_order, fe->point_set, &chk))
    return rc;
  if (chk.ng != fe->ng || chk.nd != fe->nd || chk.nf != fe->nf || chk.nq != fe->nq ||
      chk.table_size != fe->table_size)
    return set_error(PDRD_EINVAL, "inconsistent element descriptor");
  if (fe->geo_order != 1) return set_error(PDRD_ENOTSUP, "only P1/Q1 geometry is built");
  const pdrd::KernelSet ks = pdrd::find_cell(pdrd::CellKey{fe->shape, fe->sol_order, fe->nq, physics, scheme});
  if (!ks) return set_error(PDRD_ENOTSUP, "no cell kernel for this element type / physics / scheme");

  const int D = fe->dim, NG = fe->ng, ND = fe->nd, NF = fe->nf, NQ = fe->nq;
  const size_t size = fe->table_size;
  const size_t dev_size = size + static_cast<size_t>(NF) * D * NQ + static_cast<size_t>(NQ) * D * ND +
                          static_cast<size_t>(NQ) * D * NG;
  if (dev_size > kTabDoubles) return set_error(PDRD_ENOTSUP, "FE tables exceed the scratch table region");
  const int m = (physics == PDRD_EULER2D) ? 4 : (physics == PDRD_EULER3D) ? 5 : 1;
  const size_t acc = static_cast<size_t>(nb_nodes) * (m + 1);
  if (acc > kAccDoubles)
    return set_error(PDRD_EINVAL, "nb_nodes * (m + 1) = " + std::to_string(acc) + " exceeds the scratch limit of " +
                                      std::to_string(kAccDoubles) + " doubles");

  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(PDRD_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  double *base = nullptr;
  if (int rc = scratch_ptr(dev, &base)) return rc;
  double *d_tab = base;
  int *d_status = reinterpret_cast<int *>(base + kTabDoubles);
  double *d_acc_res = base + kTabDoubles + kStatusDoubles;
  double *d_acc_ws = d_acc_res + static_cast<size_t>(nb_nodes) * m;
  const hipStream_t s = static_cast<hipStream_t>(stream);

  // the previous call's accumulators and status are still in the array
  const unsigned clear_blocks = std::min(blocks_for(acc + kStatusDoubles, 256), 2048u);
  hipLaunchKernelGGL(clear_kernel, dim3(clear_blocks), dim3(256), 0, s, base + kTabDoubles, acc + kStatusDoubles);
  hipLaunchKernelGGL(table_layout_kernel, dim3(blocks_for(dev_size, 256)), dim3(256), 0, s, tables, D, NG, ND, NF, NQ,
                     static_cast<int>(size), d_tab);
  if ((e = hipGetLastError()) != hipSuccess) return set_error(PDRD_EHIP, hipGetErrorString(e));

  if (ne) {
    pdrd::CellArgs a{};
    a.tab = d_tab;
    a.gconn = geo_conn;
    a.sconn = sol_conn;
    a.coords = vertex_coords;
    a.u = u;
    a.acc_res = d_acc_res;
    a.acc_ws = d_acc_ws;
    a.status = d_status;
    a.first = 0;
    a.count = ne;
    a.ne = ne;
    a.atomic = 1;
    // the defaults of pdrd_create
    const bool euler = physics == PDRD_EULER2D || physics == PDRD_EULER3D;
    a.pp.p0 = euler ? 1.4 : 0.0;
    a.pp.p1 = euler ? 287.05 : (physics == PDRD_ADVECTION2D ? 1.0 : 0.0);
    if (nb_params > 0) a.pp.p0 = params[0];
    if (nb_params > 1) a.pp.p1 = params[1];
    a.nodal = nullptr;
    a.nv_check = nb_vertices;
    a.nn_check = nb_nodes;
    if ((e = ks.cell(a, blocks_for(ne, 256), 256, s)) != hipSuccess)
      return set_error(PDRD_EHIP, std::string("cell kernel: ") + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(finish_kernel, dim3(blocks_for(nb_nodes, 256)), dim3(256), 0, s, nb_nodes, m, d_acc_res,
                     d_acc_ws, dual_volume, residual, wave_speed, d_status, status);
  if ((e = hipGetLastError()) != hipSuccess) return set_error(PDRD_EHIP, hipGetErrorString(e));
  return PDRD_OK;
}

}  // extern "C"
