// Split-K tall-skinny product on the matrix cores: C = A B^T (FP64 result).
//
// A is (m x K) and B is (n x K), row-major with unit column stride (row strides lda /
// ldb), K the grid axis (10^6 - 10^7 points), m <= 256 and n <= 128 per output block.
// These are the Gram blocks of the block iteration ([Z; R] [R; HR]^T, Z Z^T, ...: the
// reference's projected matrices, ops/reductions_jax.py:157-159, ks/eigensolver_jax.py:
// 108-146).
//
// Each workgroup takes one panel of K and accumulates the whole partial product of its
// (256 x 96) output block in MFMA accumulators (FP32 v_mfma_f32_16x16x4f32, FP64
// v_mfma_f64_16x16x4f64), so each element of A and B is read from HBM once.  The panel is
// walked in slices of one 128-byte line per row: the wavefronts load a slice of all
// 352 rows with coalesced 128-bit loads into LDS (double buffered, one barrier per slice),
// and each wavefront then reads its MFMA operands from LDS (its row tiles of A, every
// column tile of B).  The sum over k does not depend on the order in which the k values
// meet the MFMA steps, so lane l feeds the KV consecutive values at k = KV (l / TR) of
// the slice to the KV steps (whole 128-bit LDS reads).  Rows past m and columns past n
// are clamped to the last row (their results are not written); values past K are zero.
// The partials (FP32 accumulation inside a panel, FP64 across panels) go to a workspace
// [P][m][n] and a second kernel sums them in a fixed order, so the result is
// deterministic.
#include "common.h"

#include <type_traits>

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct GS;
// MFMA shape per precision: TR x TR output tiles, NA accumulators per lane; lane l feeds
// row l % TR of a tile and k group l / TR (64 / TR groups per MFMA step).
template <>
struct GS<float> {
  // 16 x 16 tiles (the 32 x 32 x 2 shape measured slower here: 6.3 against 5.7 ms for
  // 240 x 96 at N = 224^3), sixteen wavefronts of one row tile each (<= 128 VGPRs)
  static constexpr int TR = 16, NA = 4, GW = 16;
  typedef f32x4 acc_t;
  static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // row (A side) of accumulator r of lane l in its tile; the column is l % TR
  static __device__ __forceinline__ int row(int r, int l) { return 4 * (l >> 4) + r; }
};
template <>
struct GS<double> {
  static constexpr int TR = 16, NA = 4, GW = 8;  // eight wavefronts of two row tiles
  typedef f64x4 acc_t;
  static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int l) { return (l >> 4) + 4 * r; }
};

// One K slice is one 128-byte line of every row: KS = 128 / sizeof(T) values, staged in
// LDS (double buffered) with a 16-byte pad per row.  Lane l of a wavefront feeds the
// KV = KS TR / 64 consecutive values at k = KV (l / TR) of the slice to the KV MFMA steps
// (128-bit LDS reads), row l % TR of its tiles.  An output block is 256 x 96 = 16 x 6 tiles
// of 16 x 16: FP32 one row tile per wavefront (16), FP64 two (8 wavefronts).
constexpr int GROWS = 256;  // rows of A per output block
constexpr int GCOLS = 96;   // rows of B (columns of C) per output block
constexpr int GLINE = 128;  // bytes of a row per slice
constexpr int GPITCH = GLINE + 16;

template <typename T>
__global__ __launch_bounds__(GS<T>::GW * 64, 1) void gemm_nt_split(const T* __restrict__ A, long long lda, int m,
                                                            const T* __restrict__ B, long long ldb, int n,
                                                            long long K, long long kpanel, int nbm,
                                                            double* __restrict__ W) {
  using G = GS<T>;
  constexpr int TR = G::TR, NA = G::NA, GW = G::GW;
  constexpr int KS = GLINE / (int)sizeof(T), KV = KS * TR / 64, V16 = 16 / (int)sizeof(T);
  constexpr int GRT = GROWS / TR, GCT = GCOLS / TR;  // tiles of an output block
  constexpr int ROWS = GROWS + GCOLS;                // LDS rows: A then B
  constexpr int SEGS = ROWS * (GLINE / 16);          // 16-byte segments of a full slice
  constexpr int SPL = (SEGS + GW * 64 - 1) / (GW * 64);  // segments per lane
  constexpr int TM = GRT / GW;                       // row tiles per wavefront
  static_assert(GRT % GW == 0 && KV % V16 == 0, "whole tiles per wavefront, whole 128-bit reads");
  typedef T vec_t __attribute__((ext_vector_type(V16)));
  __shared__ __attribute__((aligned(16))) unsigned char s_buf[2][ROWS * GPITCH];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ob = blockIdx.y;
  const int r0 = (ob % nbm) * GROWS, c0 = (ob / nbm) * GCOLS;
  const long long kbeg = (long long)blockIdx.x * kpanel;
  const long long kend = min(kbeg + kpanel, K);
  // rows of this block, in whole tiles: only these are staged and multiplied (small
  // products, e.g. a few soft-locked rows, cost in proportion)
  const int mtile = (min(m - r0, GROWS) + TR - 1) / TR, ntile = (min(n - c0, GCOLS) + TR - 1) / TR;
  const int mrow = mtile * TR, nrow = ntile * TR;

  // staging role: segment g = tid + GW 64 e of the slice -> staged row g / 8 (A rows first,
  // then B rows), 16-byte column g % 8; rows past m / n inside the last tile are clamped to
  // the last row (their results are not written)
  const T* src[SPL];
  int dst[SPL];
  bool sok[SPL];
#pragma unroll
  for (int e = 0; e < SPL; ++e) {
    const int g = tid + GW * 64 * e;
    sok[e] = g < (mrow + nrow) * (GLINE / 16);
    const int rr = sok[e] ? g / (GLINE / 16) : 0, c = g % (GLINE / 16);
    const int r = rr < mrow ? rr : GROWS + rr - mrow;  // LDS row
    const T* base;
    if (rr < mrow) base = A + (long long)min(r0 + rr, m - 1) * lda;
    else base = B + (long long)min(c0 + rr - mrow, n - 1) * ldb;
    src[e] = base + c * V16;
    dst[e] = r * GPITCH + c * 16;
  }
  vec_t stg[SPL];
  auto fetch = [&](long long kc) {  // one slice into registers; values past kend are zero
    const bool full = kc + KS <= kend;
#pragma unroll
    for (int e = 0; e < SPL; ++e) {
      if (!sok[e]) continue;
      if (full) {
        stg[e] = *reinterpret_cast<const vec_t*>(src[e] + kc);
      } else {
        const int cc = (dst[e] % GPITCH) / (int)sizeof(T);
#pragma unroll
        for (int v = 0; v < V16; ++v) stg[e][v] = (kc + cc + v < kend) ? src[e][kc + v] : T(0);
      }
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int e = 0; e < SPL; ++e)
      if (sok[e]) *reinterpret_cast<vec_t*>(&s_buf[buf][dst[e]]) = stg[e];
  };

  typename G::acc_t acc[TM][GCT];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int u = 0; u < GCT; ++u)
#pragma unroll
      for (int r = 0; r < NA; ++r) acc[t][u][r] = T(0);

  // operand reads: row l % TR of a tile, the KV values at k = KV (l / TR) of the slice
  const int orow = (lane % TR) * GPITCH + (lane / TR) * KV * (int)sizeof(T);
  auto operand = [&](const unsigned char* sb, int row0, T (&o)[KV]) {
#pragma unroll
    for (int h = 0; h < KV / V16; ++h) {
      const vec_t x = *reinterpret_cast<const vec_t*>(sb + row0 * GPITCH + orow + 16 * h);
#pragma unroll
      for (int v = 0; v < V16; ++v) o[h * V16 + v] = x[v];
    }
  };
  auto compute = [&](int buf) {
    const unsigned char* sb = s_buf[buf];
    T bo[GCT][KV];
#pragma unroll
    for (int u = 0; u < GCT; ++u)
      if (u < ntile) operand(sb, GROWS + u * TR, bo[u]);
#pragma unroll
    for (int t = 0; t < TM; ++t) {
      if (wave + GW * t >= mtile) continue;  // wavefront-uniform
      T ao[KV];
      operand(sb, (wave + GW * t) * TR, ao);
#pragma unroll
      for (int s2 = 0; s2 < KV; ++s2)
#pragma unroll
        for (int u = 0; u < GCT; ++u)
          if (u < ntile) acc[t][u] = G::mfma(ao[s2], bo[u][s2], acc[t][u]);
    }
  };

  const long long nsl = (kend - kbeg + KS - 1) / KS;
  if (nsl > 0) {
    fetch(kbeg);
    stash(0);
  }
  __syncthreads();
  for (long long c = 0; c < nsl; ++c) {
    if (c + 1 < nsl) fetch(kbeg + KS * (c + 1));
    compute((int)(c & 1));
    if (c + 1 < nsl) stash((int)((c + 1) & 1));
    __syncthreads();
  }

  double* Wp = W + (long long)blockIdx.x * m * n;
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int u = 0; u < GCT; ++u) {
      const int j = c0 + u * TR + (lane % TR);
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        const int i = r0 + (wave + GW * t) * TR + G::row(r, lane);
        if (i < m && j < n) Wp[(long long)i * n + j] = (double)acc[t][u][r];
      }
    }
}

// C[e] = sum_p W[p][e] in the order p = 0, 1, ...
__global__ __launch_bounds__(IG_THREADS) void gemm_nt_reduce(const double* __restrict__ W, int P, long long mn,
                                                             double* __restrict__ C, long long ldc, int n) {
  const long long e = (long long)blockIdx.x * IG_THREADS + threadIdx.x;
  if (e >= mn) return;
  double s = 0.0;
  for (int p = 0; p < P; ++p) s += W[(long long)p * mn + e];
  C[(e / n) * ldc + e % n] = s;
}

// Panels: one workgroup per CU and output block, at least 64 chunks per panel.
inline int gemm_nt_panels(long long K, int nblocks) {
  const long long want = 256 / nblocks > 0 ? 256 / nblocks : 1;
  const long long most = (K + 1023) / 1024;
  return (int)max(1LL, min(want, most));
}

template <typename T>
int launch_gemm_nt(const T* A, long long lda, int m, const T* B, long long ldb, int n, long long K, double* W,
                   long long wsize, double* C, long long ldc, hipStream_t s) {
  if (!A || !B || !W || !C || m <= 0 || n <= 0 || K <= 0 || lda < K || ldb < K || ldc < n) return IG_EINVAL;
  // 128-bit loads of a row: 16-byte aligned rows
  if ((((uintptr_t)A | (uintptr_t)B) & 15) || ((lda * (long long)sizeof(T)) & 15) ||
      ((ldb * (long long)sizeof(T)) & 15))
    return IG_EUNSUPPORTED;
  const int nbm = (m + GROWS - 1) / GROWS, nbn = (n + GCOLS - 1) / GCOLS;
  const int P = gemm_nt_panels(K, nbm * nbn);
  constexpr int KS = GLINE / (int)sizeof(T);
  long long kpanel = (K + P - 1) / P;
  kpanel = (kpanel + KS - 1) / KS * KS;  // panels start on 128-byte lines
  const int Pr = (int)((K + kpanel - 1) / kpanel);
  if ((long long)Pr * m * n > wsize) return IG_EINVAL;
  hipLaunchKernelGGL((gemm_nt_split<T>), dim3(Pr, nbm * nbn), dim3(GS<T>::GW * 64), 0, s, A, lda, m, B, ldb, n, K, kpanel,
                     nbm, W);
  IG_CHECK_LAUNCH();
  const long long mn = (long long)m * n;
  hipLaunchKernelGGL(gemm_nt_reduce, dim3((unsigned)((mn + IG_THREADS - 1) / IG_THREADS)), dim3(IG_THREADS), 0, s, W,
                     Pr, mn, C, ldc, n);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

long long gemm_nt_work(int m, int n, long long K, int elem) {
  if (m <= 0 || n <= 0 || K <= 0) return -1;
  const int nbm = (m + GROWS - 1) / GROWS, nbn = (n + GCOLS - 1) / GCOLS;
  const int P = gemm_nt_panels(K, nbm * nbn);
  const int KS = GLINE / elem;
  long long kpanel = (K + P - 1) / P;
  kpanel = (kpanel + KS - 1) / KS * KS;
  return ((K + kpanel - 1) / kpanel) * (long long)m * n;
}

}  // namespace

extern "C" {

long long ig_gemm_nt_work_size(int m, int n, long long K) {
  // the larger of the two precisions' panel counts (FP64 slices hold half the values)
  return max(gemm_nt_work(m, n, K, 4), gemm_nt_work(m, n, K, 8));
}

int ig_gemm_nt_f32(const float* A, long long lda, int m, const float* B, long long ldb, int n, long long K,
                   double* work, long long work_size, double* C, long long ldc, hipStream_t stream) {
  return launch_gemm_nt<float>(A, lda, m, B, ldb, n, K, work, work_size, C, ldc, stream);
}

int ig_gemm_nt_f64(const double* A, long long lda, int m, const double* B, long long ldb, int n, long long K,
                   double* work, long long work_size, double* C, long long ldc, hipStream_t stream) {
  return launch_gemm_nt<double>(A, lda, m, B, ldb, n, K, work, work_size, C, ldc, stream);
}

}  // extern "C"
