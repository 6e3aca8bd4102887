// Weighted tall-skinny Gram products G = X W Y^T for orbital-major blocks.
//
// X is (m, N) and Y is (n, N) with each orbital contiguous (row stride ldx / ldy),
// W = diag(w) the cell-volume metric (w == NULL: W = I).  This is the projected-matrix / overlap
// product of the reference eigensolver (ks/eigensolver.py:246-272,
// ops/reductions_jax.py:89-98, ks/eigensolver_jax.py:104-105).
//
// FP32: MFMA v_mfma_f32_32x32x2f32, a 64x64 output tile per workgroup (one 32x32
// tile per wave), split-K over the grid; each workgroup writes its FP32-accumulated
// partial as FP64 and a second kernel sums the partials in FP64 (deterministic).
// FP64: the same split-K structure with a 4x4 register tile of FP64 FMAs per thread.
#include "common.h"

typedef float floatx16 __attribute__((ext_vector_type(16)));

namespace {

constexpr int TM = 64;   // output tile rows (X)
constexpr int TN = 64;   // output tile cols (Y)
constexpr int KT = 64;   // K tile staged in LDS (FP32)
constexpr int KT64 = 32; // K tile staged in LDS (FP64)

template <typename T, int KT>
__device__ __forceinline__ void stage_tile(T* __restrict__ sx, const T* __restrict__ X, long long ldx, int m,
                                           int i0, long long k, long long kend, const T* __restrict__ w,
                                           bool weighted) {
  // 64 x KT tile, 64 consecutive threads read one row segment (coalesced).
  for (int idx = threadIdx.x; idx < TM * KT; idx += IG_THREADS) {
    const int row = idx / KT;
    const int col = idx - row * KT;
    const int i = i0 + row;
    const long long kk = k + col;
    T v = T(0);
    if (i < m && kk < kend) {
      v = X[(long long)i * ldx + kk];
      if (weighted) v *= w[kk];
    }
    sx[row * (KT + 1) + col] = v;
  }
}

__global__ __launch_bounds__(IG_THREADS) void gram_f32_mfma(const float* __restrict__ X, long long ldx, int m,
                                                            const float* __restrict__ Y, long long ldy, int n,
                                                            const float* __restrict__ w, long long N,
                                                            long long kchunk, int tiles_n,
                                                            double* __restrict__ work) {
  __shared__ float sx[TM * (KT + 1)];
  __shared__ float sy[TN * (KT + 1)];
  const int tile = blockIdx.x;
  const int kb = blockIdx.y;
  const int i0 = (tile / tiles_n) * TM;
  const int j0 = (tile % tiles_n) * TN;
  const long long kbeg = (long long)kb * kchunk;
  const long long kend = min(kbeg + kchunk, N);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (long long k = kbeg; k < kend; k += KT) {
    stage_tile<float, KT>(sx, X, ldx, m, i0, k, kend, w, w != nullptr);
    stage_tile<float, KT>(sy, Y, ldy, n, j0, k, kend, w, false);
    __syncthreads();
    const float* ax = sx + (wm * 32 + (lane & 31)) * (KT + 1) + (lane >> 5);
    const float* by = sy + (wn * 32 + (lane & 31)) * (KT + 1) + (lane >> 5);
#pragma unroll 8
    for (int kk = 0; kk < KT; kk += 2) {
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ax[kk], by[kk], acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // C[i][j]: j = lane % 32, i = 8 * (r / 4) + 4 * (lane / 32) + r % 4.
  double* out = work + (long long)kb * m * n;
  const int j = j0 + wn * 32 + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = i0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i < m && j < n) out[(long long)i * n + j] = (double)acc[r];
  }
}

__global__ __launch_bounds__(IG_THREADS) void gram_f64_simt(const double* __restrict__ X, long long ldx, int m,
                                                            const double* __restrict__ Y, long long ldy, int n,
                                                            const double* __restrict__ w, long long N,
                                                            long long kchunk, int tiles_n,
                                                            double* __restrict__ work) {
  constexpr int KT = KT64;
  __shared__ double sx[TM * (KT + 1)];
  __shared__ double sy[TN * (KT + 1)];
  const int tile = blockIdx.x;
  const int kb = blockIdx.y;
  const int i0 = (tile / tiles_n) * TM;
  const int j0 = (tile % tiles_n) * TN;
  const long long kbeg = (long long)kb * kchunk;
  const long long kend = min(kbeg + kchunk, N);
  const int tx = threadIdx.x & 15;
  const int ty = threadIdx.x >> 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
  for (long long k = kbeg; k < kend; k += KT) {
    stage_tile<double, KT>(sx, X, ldx, m, i0, k, kend, w, w != nullptr);
    stage_tile<double, KT>(sy, Y, ldy, n, j0, k, kend, w, false);
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < KT; ++kk) {
      double xa[4], yb[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) xa[a] = sx[(ty + 16 * a) * (KT + 1) + kk];
#pragma unroll
      for (int c = 0; c < 4; ++c) yb[c] = sy[(tx + 16 * c) * (KT + 1) + kk];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = fma(xa[a], yb[c], acc[a][c]);
    }
    __syncthreads();
  }
  double* out = work + (long long)kb * m * n;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = i0 + ty + 16 * a;
      const int j = j0 + tx + 16 * c;
      if (i < m && j < n) out[(long long)i * n + j] = acc[a][c];
    }
}

// G (+)= sum over partials; symmetrize optionally handled on the host.
__global__ __launch_bounds__(IG_THREADS) void gram_reduce(const double* __restrict__ work, int nparts, long long mn,
                                                          double* __restrict__ G, int accumulate) {
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < mn;
       t += (long long)gridDim.x * IG_THREADS) {
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += work[(long long)p * mn + t];
    G[t] = accumulate ? G[t] + s : s;
  }
}

struct GramPlan {
  int tiles_m, tiles_n, nkb;
  long long kchunk;
};

GramPlan plan(int m, int n, long long N) {
  GramPlan g;
  g.tiles_m = (m + TM - 1) / TM;
  g.tiles_n = (n + TN - 1) / TN;
  const int tiles = g.tiles_m * g.tiles_n;
  long long nkb = (N + 4095) / 4096;
  long long cap = 2048 / tiles;
  if (cap < 8) cap = 8;
  if (nkb > cap) nkb = cap;
  if (nkb < 1) nkb = 1;
  long long kc = (N + nkb - 1) / nkb;
  kc = ((kc + KT - 1) / KT) * KT;
  g.kchunk = kc;
  g.nkb = (int)((N + kc - 1) / kc);
  return g;
}

}  // namespace

extern "C" {

long long ig_gram_work_size(int m, int n, long long N) {
  if (m <= 0 || n <= 0 || N <= 0) return 0;
  const GramPlan g = plan(m, n, N);
  return (long long)g.nkb * m * n;
}

int ig_gram_f32(const float* X, long long ldx, int m, const float* Y, long long ldy, int n,
                const float* w, long long N, double* G, int accumulate, double* work,
                long long work_elems, hipStream_t s) {
  if (!X || !Y || !G || !work || m <= 0 || n <= 0 || N <= 0) return IG_EINVAL;
  if (ldx < N || ldy < N) return IG_EINVAL;
  const GramPlan g = plan(m, n, N);
  if (work_elems < (long long)g.nkb * m * n) return IG_EINVAL;
  hipLaunchKernelGGL(gram_f32_mfma, dim3(g.tiles_m * g.tiles_n, g.nkb), dim3(IG_THREADS), 0, s, X, ldx, m,
                     Y, ldy, n, w, N, g.kchunk, g.tiles_n, work);
  const long long mn = (long long)m * n;
  hipLaunchKernelGGL(gram_reduce, dim3((unsigned)((mn + IG_THREADS - 1) / IG_THREADS)), dim3(IG_THREADS), 0, s,
                     work, g.nkb, mn, G, accumulate);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_gram_f64(const double* X, long long ldx, int m, const double* Y, long long ldy, int n,
                const double* w, long long N, double* G, int accumulate, double* work,
                long long work_elems, hipStream_t s) {
  if (!X || !Y || !G || !work || m <= 0 || n <= 0 || N <= 0) return IG_EINVAL;
  if (ldx < N || ldy < N) return IG_EINVAL;
  const GramPlan g = plan(m, n, N);
  if (work_elems < (long long)g.nkb * m * n) return IG_EINVAL;
  hipLaunchKernelGGL(gram_f64_simt, dim3(g.tiles_m * g.tiles_n, g.nkb), dim3(IG_THREADS), 0, s, X, ldx, m,
                     Y, ldy, n, w, N, g.kchunk, g.tiles_n, work);
  const long long mn = (long long)m * n;
  hipLaunchKernelGGL(gram_reduce, dim3((unsigned)((mn + IG_THREADS - 1) / IG_THREADS)), dim3(IG_THREADS), 0, s,
                     work, g.nkb, mn, G, accumulate);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

}  // extern "C"
