// Device helpers shared by the kernel translation units: wavefront (64-lane)
// reductions and scans, cell-grid indexing, order-independent fixed-point
// scatter, and the FP32 Gaussian-overlap / spline primitives of the model.
#pragma once
#include <hip/hip_runtime.h>

#include "agb_common.h"

namespace agb {

// ---------------------------------------------------------- wave collectives
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// ------------------------------------------------------------ work queues
// Persistent blocks take work items from a queue head in memory; every block
// leaves once the head has passed the item count, so the grid always drains.
// `slot` is block-shared scratch for the broadcast.
__device__ __forceinline__ int next_item(unsigned *head, int *slot) {
  __syncthreads();                       // the previous item is finished by all waves
  if (threadIdx.x == 0) *slot = (int)atomicAdd(head, 1u);
  __syncthreads();
  return *slot;
}

// ------------------------------------------------------------------ cells
__device__ __forceinline__ void cell_coords(const GridParams &g, float x, float y, float z, int &cx, int &cy,
                                            int &cz) {
  cx = min(max((int)floorf((x - g.lo[0]) * g.inv_cs), 0), g.dim[0] - 1);
  cy = min(max((int)floorf((y - g.lo[1]) * g.inv_cs), 0), g.dim[1] - 1);
  cz = min(max((int)floorf((z - g.lo[2]) * g.inv_cs), 0), g.dim[2] - 1);
}
__device__ __forceinline__ int cell_index(const GridParams &g, float x, float y, float z) {
  int cx, cy, cz;
  cell_coords(g, x, y, z, cx, cy, cz);
  return (cz * g.dim[1] + cy) * g.dim[0] + cx;
}

// -------------------------------------------------- fixed-point scatter add
// Each contribution is rounded to 2^-32 once, so sums are exact and their
// value does not depend on the order in which waves arrive.
__device__ __forceinline__ void fix_add(long long *p, double v) {
  if (v != 0.0) atomicAdd((unsigned long long *)p, (unsigned long long)__double2ll_rn(v * kFix));
}
__device__ __forceinline__ double fix_get(long long v) { return (double)v * kFixInv; }

// ------------------------------------------------- overlap-volume switch
// quintic smoothstep of the raw overlap volume on [0.01, 0.1]
// (reference agbnp3_cpu_simd_kernels.c:2448-2471)
__device__ __forceinline__ void vol_switch(float v, float &s, float &sp, float &spp) {
  if (v > kVolMinB) { s = 1.f; sp = 0.f; spp = 0.f; return; }
  if (v < kVolMinA) { s = 0.f; sp = 0.f; spp = 0.f; return; }
  const float d = 1.0f / (kVolMinB - kVolMinA);
  const float u = (v - kVolMinA) * d, u2 = u * u;
  s = u2 * u * (10.f - 15.f * u + 6.f * u2);
  sp = d * 30.f * u2 * (1.f - 2.f * u + u2);
  spp = d * d * 60.f * u * (1.f - 3.f * u + 2.f * u2);
}

// --------------------------------------------------------- Q4 spline lookup
// node pairs (y_k, y2_k) of one table in LDS; a = d / R_j on the [0, 20)
// grid; k > 510 is the flat zero tail (agbnp3_utils.c:331-377, 266-295).
// `fixed` selects the textbook derivative sign of the y2_{k+1} term.
template <bool WANT_Q, bool WANT_DQ>
__device__ __forceinline__ void q4_lookup(const float2 *__restrict__ tab, float a, float dx, float dxinv,
                                          float dp1, float dp2, float sgn2, float &q, float &dq) {
  const float xh = a * dxinv;
  const int k = (int)xh;
  q = 0.f;
  dq = 0.f;
  if (k >= 0 && k <= kTblN - 2) {
    const float2 n0 = tab[k], n1 = tab[k + 1];
    const float A = (float)(k + 1) - xh, B = xh - (float)k;
    const float A2 = A * A, B2 = B * B;
    if (WANT_Q) q = A * n0.x + B * n1.x + ((A2 * A - A) * n0.y + (B2 * B - B) * n1.y) * dp2;
    if (WANT_DQ) dq = (n1.x - n0.x) * dxinv - ((3.f * A2 - 1.f) * n0.y + sgn2 * (3.f * B2 - 1.f) * n1.y) * dp1;
  }
}

}  // namespace agb
