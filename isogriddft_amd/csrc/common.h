// Shared helpers for the gfx950 kernels of the IsoGridDFT Kohn-Sham inner loop.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IG_THREADS 256

// Status codes returned through the C ABI.
#define IG_OK 0
#define IG_EINVAL -1
#define IG_EUNSUPPORTED -2
#define IG_EHIP -3

#define IG_CHECK_LAUNCH()                         \
  do {                                            \
    if (hipGetLastError() != hipSuccess) return IG_EHIP; \
  } while (0)

// Warp (64-lane wavefront) and block reductions in double precision.
__device__ __forceinline__ double ig_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Reduce NV values per thread over a 256-thread block; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void ig_block_sum(double (&v)[NV], double* smem /* >= 4*NV */) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = ig_wave_sum(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) smem[wave * NV + i] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = smem[i] + smem[NV + i] + smem[2 * NV + i] + smem[3 * NV + i];
  }
  __syncthreads();
}
