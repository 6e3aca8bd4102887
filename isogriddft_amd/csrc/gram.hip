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

// G (+)= sum over partials.  A workgroup owns 32 consecutive entries; its 8 slices of 32
// lanes each sum every 8th partial (coalesced 256-byte rows) and the slice sums are
// combined in a fixed order (deterministic).  With tcols > 0 the partials hold the
// transpose (rows of tcols entries) and G is written transposed.
constexpr int GR_OUT = 32, GR_SL = IG_THREADS / GR_OUT;
__global__ __launch_bounds__(IG_THREADS) void gram_reduce(const double* __restrict__ work, int nparts, long long mn,
                                                          double* __restrict__ G, int accumulate, int tcols = 0,
                                                          int trows = 0) {
  __shared__ double part[GR_SL][GR_OUT];
  const int tx = threadIdx.x % GR_OUT, sl = threadIdx.x / GR_OUT;
  const long long t = (long long)blockIdx.x * GR_OUT + tx;
  double s = 0.0;
  if (t < mn)
    for (int p = sl; p < nparts; p += GR_SL) s += work[(long long)p * mn + t];
  part[sl][tx] = s;
  __syncthreads();
  if (sl == 0 && t < mn) {
    double tot = part[0][tx];
#pragma unroll
    for (int q = 1; q < GR_SL; ++q) tot += part[q][tx];
    // partial entry t = (i, j) of the transposed product -> G[j][i]
    const long long g = tcols > 0 ? (t % tcols) * trows + t / tcols : t;
    G[g] = accumulate ? G[g] + tot : tot;
  }
}

inline unsigned reduce_blocks(long long mn) { return (unsigned)((mn + GR_OUT - 1) / GR_OUT); }

// ---------------------------------------------------------------------------------
// Narrow products (m <= 256 rows of X, n <= 64 rows of Y): the whole m x n output fits
// one workgroup, so the grid splits only K.  Thread (tm, tn) = (tid / 8, tid % 8) owns
// rows tm + 32 a (a < RM) and columns tn + 8 b (b < RN) in registers (RM, RN the
// smallest powers of two covering m / 32 and n / 8, so no FMA is spent on rows or
// columns outside the product); a K tile of 16 is
// staged through LDS (K-major, FP64) with the next tile's loads in flight in registers.
// Each pass over the operands is HBM-bound; the FP64 partials per K chunk are summed by
// gram_reduce in a fixed order (deterministic).
constexpr int NW_KT = 16, NW_RM = 8, NW_MMAX = 32 * NW_RM, NW_NMAX = 64;
constexpr int NW_YP = NW_NMAX * NW_KT / 2 / IG_THREADS;  // Y pairs per thread and tile (max)

template <typename T>
struct Pair;
template <>
struct Pair<float> {
  typedef float t __attribute__((ext_vector_type(2)));
};
template <>
struct Pair<double> {
  typedef double t __attribute__((ext_vector_type(2)));
};

template <typename T, int RM, int RN, bool PAIR>
__global__ __launch_bounds__(IG_THREADS) void gram_narrow(const T* __restrict__ X, long long ldx, int m,
                                                          const T* __restrict__ Y, long long ldy, int n,
                                                          const T* __restrict__ w, long long N, long long kchunk,
                                                          double* __restrict__ work) {
  using PT = typename Pair<T>::t;
  constexpr int NW_XP = 32 * RM * NW_KT / 2 / IG_THREADS;  // X pairs per thread and tile
  __shared__ double sx[NW_KT][NW_MMAX];
  __shared__ double sy[NW_KT][NW_NMAX];
  const int tid = threadIdx.x;
  const int tm = tid >> 3, tn = tid & 7;
  const long long kbeg = (long long)blockIdx.x * kchunk;
  const long long kend = min(kbeg + kchunk, N);
  double acc[RM][RN];
#pragma unroll
  for (int a = 0; a < RM; ++a)
#pragma unroll
    for (int b = 0; b < RN; ++b) acc[a][b] = 0.0;
  // staging (PAIR): element pair q of a tile is (row q / 8, k = 2 (q % 8)); pairs never
  // straddle the chunk end because chunks and N are even (the host guarantees both).
  // Otherwise element q is (row q / 16, k = q % 16), one scalar load each.
  PT px[NW_XP], py[NW_YP];
  T qx[2 * NW_XP], qy[2 * NW_YP];
  auto fetch_scalar = [&](long long k0) {
#pragma unroll
    for (int r = 0; r < 2 * NW_XP; ++r) {
      const int q = tid + r * IG_THREADS;
      const int row = q >> 4;
      const long long k = k0 + (q & 15);
      T v = T(0);
      if (row < m && k < kend) {
        v = X[(long long)row * ldx + k];
        if (w) v *= w[k];
      }
      qx[r] = v;
    }
#pragma unroll
    for (int r = 0; r < 2 * NW_YP; ++r) {
      const int q = tid + r * IG_THREADS;
      const int row = q >> 4;
      const long long k = k0 + (q & 15);
      qy[r] = (row < n && k < kend) ? Y[(long long)row * ldy + k] : T(0);
    }
  };
  auto commit_scalar = [&]() {
#pragma unroll
    for (int r = 0; r < 2 * NW_XP; ++r) {
      const int q = tid + r * IG_THREADS;
      sx[q & 15][q >> 4] = (double)qx[r];
    }
#pragma unroll
    for (int r = 0; r < 2 * NW_YP; ++r) {
      const int q = tid + r * IG_THREADS;
      sy[q & 15][q >> 4] = (double)qy[r];
    }
  };
  auto fetch = [&](long long k0) {
    if (!PAIR) {
      fetch_scalar(k0);
      return;
    }
#pragma unroll
    for (int r = 0; r < NW_XP; ++r) {
      const int q = tid + r * IG_THREADS;
      const int row = q >> 3;
      const long long k = k0 + 2 * (q & 7);
      PT v = PT(0);
      if (row < m && k < kend) {
        v = *reinterpret_cast<const PT*>(X + (long long)row * ldx + k);
        if (w) {
          const PT ww = *reinterpret_cast<const PT*>(w + k);
          v = v * ww;
        }
      }
      px[r] = v;
    }
#pragma unroll
    for (int r = 0; r < NW_YP; ++r) {
      const int q = tid + r * IG_THREADS;
      const int row = q >> 3;
      const long long k = k0 + 2 * (q & 7);
      PT v = PT(0);
      if (row < n && k < kend) v = *reinterpret_cast<const PT*>(Y + (long long)row * ldy + k);
      py[r] = v;
    }
  };
  auto commit = [&]() {
    if (!PAIR) {
      commit_scalar();
      return;
    }
#pragma unroll
    for (int r = 0; r < NW_XP; ++r) {
      const int q = tid + r * IG_THREADS;
      const int row = q >> 3, kk = 2 * (q & 7);
      if (row < NW_MMAX) {
        sx[kk][row] = (double)px[r][0];
        sx[kk + 1][row] = (double)px[r][1];
      }
    }
#pragma unroll
    for (int r = 0; r < NW_YP; ++r) {
      const int q = tid + r * IG_THREADS;
      const int row = q >> 3, kk = 2 * (q & 7);
      if (row < NW_NMAX) {
        sy[kk][row] = (double)py[r][0];
        sy[kk + 1][row] = (double)py[r][1];
      }
    }
  };
  fetch(kbeg);
  for (long long k0 = kbeg; k0 < kend; k0 += NW_KT) {
    __syncthreads();  // the previous tile has been consumed
    commit();
    __syncthreads();
    if (k0 + NW_KT < kend) fetch(k0 + NW_KT);
#pragma unroll 4
    for (int kk = 0; kk < NW_KT; ++kk) {
      double xa[RM], yb[RN];
#pragma unroll
      for (int a = 0; a < RM; ++a) xa[a] = sx[kk][tm + 32 * a];
#pragma unroll
      for (int b = 0; b < RN; ++b) yb[b] = sy[kk][tn + 8 * b];
#pragma unroll
      for (int a = 0; a < RM; ++a)
#pragma unroll
        for (int b = 0; b < RN; ++b) acc[a][b] = fma(xa[a], yb[b], acc[a][b]);
    }
  }
  double* out = work + (long long)blockIdx.x * m * n;
#pragma unroll
  for (int a = 0; a < RM; ++a)
#pragma unroll
    for (int b = 0; b < RN; ++b) {
      const int i = tm + 32 * a, j = tn + 8 * b;
      if (i < m && j < n) out[(long long)i * n + j] = acc[a][b];
    }
}

struct NarrowPlan {
  bool use, swap, pair;
  int nchunk;
  long long kchunk;
};

NarrowPlan narrow_plan(int m, int n, long long N, long long ldx, long long ldy) {
  NarrowPlan g{false, false, false, 0, 0};
  const bool fits = n <= NW_NMAX && m <= NW_MMAX;
  const bool fits_t = m <= NW_NMAX && n <= NW_MMAX;
  if (!fits && !fits_t) return g;
  g.use = true;
  g.swap = !fits;
  // paired (2-element) loads need even N and even row strides (pointer alignment is
  // checked by the caller)
  g.pair = !(N & 1) && !(ldx & 1) && !(ldy & 1);
  long long nch = (N + 511) / 512;  // at least ~512 K values per workgroup
  if (nch > 1024) nch = 1024;
  if (nch < 1) nch = 1;
  long long kc = (N + nch - 1) / nch;
  kc = ((kc + NW_KT - 1) / NW_KT) * NW_KT;
  g.kchunk = kc;
  g.nchunk = (int)((N + kc - 1) / kc);
  return g;
}

template <typename T>
int gram_narrow_launch(const NarrowPlan& g, const T* X, long long ldx, int m, const T* Y, long long ldy, int n,
                       const T* w, long long N, double* G, int accumulate, double* work, hipStream_t s) {
  // with swap the kernel computes Y W X^T (n x m) and the reduction writes it transposed
  const T* A = g.swap ? Y : X;
  const T* B = g.swap ? X : Y;
  const long long lda = g.swap ? ldy : ldx, ldb = g.swap ? ldx : ldy;
  const int ma = g.swap ? n : m, nb = g.swap ? m : n;
  const dim3 grid(g.nchunk), block(IG_THREADS);
#define IG_NARROW(RM_, RN_)                                                                                      \
  do {                                                                                                           \
    if (g.pair)                                                                                                  \
      hipLaunchKernelGGL((gram_narrow<T, RM_, RN_, true>), grid, block, 0, s, A, lda, ma, B, ldb, nb, w, N,       \
                         g.kchunk, work);                                                                        \
    else                                                                                                         \
      hipLaunchKernelGGL((gram_narrow<T, RM_, RN_, false>), grid, block, 0, s, A, lda, ma, B, ldb, nb, w, N,      \
                         g.kchunk, work);                                                                        \
  } while (0)
#define IG_NARROW_N(RM_)   \
  do {                     \
    if (nb <= 8)           \
      IG_NARROW(RM_, 1);   \
    else if (nb <= 16)     \
      IG_NARROW(RM_, 2);   \
    else if (nb <= 32)     \
      IG_NARROW(RM_, 4);   \
    else                   \
      IG_NARROW(RM_, 8);   \
  } while (0)
  if (ma <= 32)
    IG_NARROW_N(1);
  else if (ma <= 64)
    IG_NARROW_N(2);
  else if (ma <= 128)
    IG_NARROW_N(4);
  else
    IG_NARROW_N(8);
#undef IG_NARROW_N
#undef IG_NARROW
  const long long mn = (long long)m * n;
  hipLaunchKernelGGL(gram_reduce, dim3(reduce_blocks(mn)), dim3(IG_THREADS), 0, s, work, g.nchunk, mn, G, accumulate,
                     g.swap ? m : 0, g.swap ? n : 0);
  IG_CHECK_LAUNCH();
  return IG_OK;
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
  long long need = (long long)g.nkb * m * n;
  const NarrowPlan nw = narrow_plan(m, n, N, 2, 2);
  if (nw.use && (long long)nw.nchunk * m * n > need) need = (long long)nw.nchunk * m * n;
  return need;
}

int ig_gram_f32(const float* X, long long ldx, int m, const float* Y, long long ldy, int n,
                const float* w, long long N, double* G, int accumulate, double* work,
                long long work_elems, hipStream_t s) {
  if (!X || !Y || !G || !work || m <= 0 || n <= 0 || N <= 0) return IG_EINVAL;
  if (ldx < N || ldy < N) return IG_EINVAL;
  const NarrowPlan nw = narrow_plan(m, n, N, ldx, ldy);
  if (nw.use && work_elems >= (long long)nw.nchunk * m * n) {
    NarrowPlan g = nw;
    const bool aligned = ((reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y) |
                           reinterpret_cast<uintptr_t>(w)) & (2 * sizeof(float) - 1)) == 0;
    g.pair = g.pair && aligned;
    return gram_narrow_launch<float>(g, X, ldx, m, Y, ldy, n, w, N, G, accumulate, work, s);
  }
  const GramPlan g = plan(m, n, N);
  if (work_elems < (long long)g.nkb * m * n) return IG_EINVAL;
  hipLaunchKernelGGL(gram_f32_mfma, dim3(g.tiles_m * g.tiles_n, g.nkb), dim3(IG_THREADS), 0, s, X, ldx, m,
                     Y, ldy, n, w, N, g.kchunk, g.tiles_n, work);
  const long long mn = (long long)m * n;
  hipLaunchKernelGGL(gram_reduce, dim3(reduce_blocks(mn)), dim3(IG_THREADS), 0, s, work, g.nkb, mn, G, accumulate, 0, 0);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_gram_f64(const double* X, long long ldx, int m, const double* Y, long long ldy, int n,
                const double* w, long long N, double* G, int accumulate, double* work,
                long long work_elems, hipStream_t s) {
  if (!X || !Y || !G || !work || m <= 0 || n <= 0 || N <= 0) return IG_EINVAL;
  if (ldx < N || ldy < N) return IG_EINVAL;
  const NarrowPlan nw = narrow_plan(m, n, N, ldx, ldy);
  if (nw.use && work_elems >= (long long)nw.nchunk * m * n) {
    NarrowPlan g = nw;
    const bool aligned = ((reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y) |
                           reinterpret_cast<uintptr_t>(w)) & (2 * sizeof(double) - 1)) == 0;
    g.pair = g.pair && aligned;
    return gram_narrow_launch<double>(g, X, ldx, m, Y, ldy, n, w, N, G, accumulate, work, s);
  }
  const GramPlan g = plan(m, n, N);
  if (work_elems < (long long)g.nkb * m * n) return IG_EINVAL;
  hipLaunchKernelGGL(gram_f64_simt, dim3(g.tiles_m * g.tiles_n, g.nkb), dim3(IG_THREADS), 0, s, X, ldx, m,
                     Y, ldy, n, w, N, g.kchunk, g.tiles_n, work);
  const long long mn = (long long)m * n;
  hipLaunchKernelGGL(gram_reduce, dim3(reduce_blocks(mn)), dim3(IG_THREADS), 0, s, work, g.nkb, mn, G, accumulate, 0, 0);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

}  // extern "C"
