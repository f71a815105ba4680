// The flux-node interpolation of the P2 Euler triangle cells, S = Vsf U
// (Vsf: the 10 x 6 table of the P3 flux nodes, U: a cell's 6 x 4 nodal
// states), in two forms over the same cells, each repeated R times on
// register-resident data so that the arithmetic, not memory, is timed:
//   valu: one lane per cell, 240 FP64 FMAs per repeat (the patch kernel's
//         form);
//   mfma: v_mfma_f64_16x16x4f64 with Vsf padded to 16 x 8 (two K steps) and
//         four cells' 6 x 4 states as one 16-column tile, 16 tiles per wave
//         of 64 cells.
// The VALU form writes S of every cell, the MFMA form its raw accumulators;
// the program finds the accumulator layout under which the two agree
// (≤ 1e-12) and prints one JSON line with the time and FP64 rate of each. MI355X's FP64 matrix
// peak equals its FP64 vector peak (78.6 TFLOP/s), so the matrix form can
// only win by issuing beside the vector ALU, and the padding of the 10 x 6
// table to 16 x 8 spends 53 % of its work on zeros.
//
// Build: make -C pdekit_amd/csrc probe; run: pdekit_amd/probe_mfma [cells] [repeats]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

constexpr int NF = 10, ND = 6, M = 4;

__constant__ double c_vsf[NF * ND];

#define CHECK(x)                                                                          \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      std::fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      std::exit(1);                                                                       \
    }                                                                                     \
  } while (0)

// u: [E][ND][M], s: [E][NF][M]
__global__ __launch_bounds__(256) void valu_kernel(const double *__restrict__ u, double *__restrict__ s, int E, int R)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double U[ND][M], acc[NF][M];
#pragma unroll
  for (int n = 0; n < ND; ++n)
#pragma unroll
    for (int k = 0; k < M; ++k) U[n][k] = u[(static_cast<size_t>(e) * ND + n) * M + k];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int k = 0; k < M; ++k) acc[f][k] = 0.0;
  for (int r = 0; r < R; ++r) {
    // opaque states: the repeats cannot be folded
#pragma unroll
    for (int n = 0; n < ND; ++n)
#pragma unroll
      for (int k = 0; k < M; ++k) asm volatile("" : "+v"(U[n][k]));
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k < M; ++k) {
        double t = acc[f][k];
#pragma unroll
        for (int n = 0; n < ND; ++n) t = fma(c_vsf[f * ND + n], U[n][k], t);
        acc[f][k] = t;
      }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int k = 0; k < M; ++k) s[(static_cast<size_t>(e) * NF + f) * M + k] = acc[f][k];
}

typedef double double4v __attribute__((ext_vector_type(4)));

// One wave per 64 cells (16 tiles of 4 cells). Lane l holds
//   A[i][kk] = Vsf[i][4 step + kk] at i = l % 16, kk = l / 16 (zero past the table),
//   B[kk][j] = U_cell[4 step + kk][comp] at kk = l / 16, j = l % 16 (cell = 4 tile + j / 4, comp = j % 4),
// and its four accumulator values of every tile are stored as they are,
// raw[((wave * 16 + tile) * 64 + l) * 4 + r]; the host finds which (row,
// column) of D each (lane, r) holds by comparing with the VALU result.
__global__ __launch_bounds__(64) void mfma_kernel(const double *__restrict__ u, double *__restrict__ s, int E, int R)
{
  const int l = threadIdx.x;
  const int base = blockIdx.x * 64;  // first cell of the wave
  const int i = l % 16, kk = l / 16;
  double a[2];
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    const int n = 4 * st + kk;
    a[st] = (i < NF && n < ND) ? c_vsf[i * ND + n] : 0.0;
  }
  double b[16][2];
#pragma unroll
  for (int t = 0; t < 16; ++t)
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      const int n = 4 * st + kk;
      const int cell = base + 4 * t + i / 4, comp = i % 4;
      b[t][st] = (n < ND && cell < E) ? u[(static_cast<size_t>(cell) * ND + n) * M + comp] : 0.0;
    }
  double4v acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = double4v{0.0, 0.0, 0.0, 0.0};
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
      for (int st = 0; st < 2; ++st) asm volatile("" : "+v"(b[t][st]));
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
      for (int st = 0; st < 2; ++st) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[st], b[t][st], acc[t], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 16; ++t)
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) s[((static_cast<size_t>(blockIdx.x) * 16 + t) * 64 + l) * 4 + rr] = acc[t][rr];
}

}  // namespace

int main(int argc, char **argv)
{
  const int E = argc > 1 ? std::atoi(argv[1]) : 524288;
  const int R = argc > 2 ? std::atoi(argv[2]) : 200;
  if (E <= 0 || R <= 0) return 2;
  std::vector<double> vsf(NF * ND), u(static_cast<size_t>(E) * ND * M);
  unsigned x = 12345u;
  auto rnd = [&]() {
    x = x * 1664525u + 1013904223u;
    return (x >> 8) * (1.0 / 16777216.0) - 0.5;
  };
  for (double &v : vsf) v = rnd();
  for (double &v : u) v = rnd();
  CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_vsf), vsf.data(), vsf.size() * sizeof(double)));
  double *du = nullptr, *s1 = nullptr, *s2 = nullptr;
  const int waves = (E + 63) / 64;
  const size_t nu = u.size(), ns = static_cast<size_t>(E) * NF * M, nraw = static_cast<size_t>(waves) * 16 * 64 * 4;
  CHECK(hipMalloc(&du, nu * sizeof(double)));
  CHECK(hipMalloc(&s1, ns * sizeof(double)));
  CHECK(hipMalloc(&s2, nraw * sizeof(double)));
  CHECK(hipMemcpy(du, u.data(), nu * sizeof(double), hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  auto timed = [&](bool mfma) {
    float best = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {  // the first call loads the code object
      CHECK(hipEventRecord(e0));
      if (mfma)
        hipLaunchKernelGGL(mfma_kernel, dim3((E + 63) / 64), dim3(64), 0, 0, du, s2, E, R);
      else
        hipLaunchKernelGGL(valu_kernel, dim3((E + 255) / 256), dim3(256), 0, 0, du, s1, E, R);
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(e1));
      CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0 && ms < best) best = ms;
    }
    return best;
  };
  const float tv = timed(false), tm = timed(true);
  std::vector<double> a(ns), raw(nraw);
  CHECK(hipMemcpy(a.data(), s1, ns * sizeof(double), hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(raw.data(), s2, nraw * sizeof(double), hipMemcpyDeviceToHost));
  // candidate (row, column) of the accumulator value r of lane l
  auto row_col = [](int layout, int l, int r, int &row, int &col) {
    switch (layout) {
      case 0: row = 4 * (l / 16) + r; col = l % 16; break;   // rows in blocks of four
      case 1: row = l / 16 + 4 * r; col = l % 16; break;     // rows interleaved
      case 2: row = l % 16; col = 4 * (l / 16) + r; break;   // transposed, blocks
      default: row = l % 16; col = l / 16 + 4 * r; break;    // transposed, interleaved
    }
  };
  const char *names[4] = {"row 4(l/16)+r, col l%16", "row l/16+4r, col l%16", "row l%16, col 4(l/16)+r",
                          "row l%16, col l/16+4r"};
  int layout = -1;
  double num = 0.0, den = 0.0;
  for (int cand = 0; cand < 4 && layout < 0; ++cand) {
    num = den = 0.0;
    for (int w = 0; w < waves; ++w)
      for (int t = 0; t < 16; ++t)
        for (int l = 0; l < 64; ++l)
          for (int r = 0; r < 4; ++r) {
            int row, col;
            row_col(cand, l, r, row, col);
            const int cell = w * 64 + 4 * t + col / 4, comp = col % 4;
            if (row >= NF || cell >= E) continue;
            const double ref = a[(static_cast<size_t>(cell) * NF + row) * M + comp];
            const double got = raw[((static_cast<size_t>(w) * 16 + t) * 64 + l) * 4 + r];
            num = std::fmax(num, std::fabs(ref - got));
            den = std::fmax(den, std::fabs(ref));
          }
    if (den > 0 && num / den <= 1e-12) layout = cand;
  }
  // useful FP64 work: 240 FMAs per cell and repeat; the matrix form also
  // multiplies the padding (16 x 16 x 8 per tile of 4 cells: 512 FMAs per cell
  // and repeat)
  const double useful = 2.0 * 240.0 * E * static_cast<double>(R);
  std::printf("{\"probe\": \"P2 flux-node interpolation Vsf U (10x6 by 6x4 per cell)\", \"cells\": %d, \"repeats\": %d, "
              "\"valu_ms\": %.4f, \"mfma_ms\": %.4f, \"valu_tflops\": %.2f, \"mfma_tflops_useful\": %.2f, "
              "\"mfma_tflops_issued\": %.2f, \"mfma_over_valu\": %.3f, \"max_rel_diff\": %.3e, \"d_layout\": \"%s\"}\n",
              E, R, tv, tm, useful / (tv * 1e-3) / 1e12, useful / (tm * 1e-3) / 1e12,
              useful * 512.0 / 240.0 / (tm * 1e-3) / 1e12, tv / tm, den > 0 ? num / den : 0.0,
              layout >= 0 ? names[layout] : "none matches");
  (void)hipFree(du);
  (void)hipFree(s1);
  (void)hipFree(s2);
  return layout >= 0 ? 0 : 1;
}
