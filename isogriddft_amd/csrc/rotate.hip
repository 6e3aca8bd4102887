// Basis rotation of the block iteration on the matrix cores: Y = Q X.
//
// Q is a small (M x K) row-major matrix, X holds K rows of N grid values (row stride ldx),
// Y holds M rows (row stride ldy).  This is the subspace update of the reference's block
// iteration (ks/eigensolver_jax.py:184-201: basis <- coefficients^T [basis; residuals];
// ks/eigensolver.py:916-932), with the long grid axis N as the GEMM's column dimension.
//
// Each wavefront owns CT column tiles of the grid and every output row: it streams its
// columns of X once, from the first row of X to the last, and accumulates all M rows of Y
// in registers (TM row tiles x CT column tiles of MFMA accumulators).  The coefficients go
// through LDS in chunks of KC columns of Q, double-buffered and shared by the four
// wavefronts of a workgroup, laid out so that every MFMA A operand is one conflict-free
// LDS read.  Y is written only after the wavefront has consumed all of X in its columns,
// and no wavefront touches another's columns, so Y may alias the first M rows of X (an
// in-place rotation needs no temporary and no column panels).
//
// FP32: v_mfma_f32_32x32x2f32 (32 x 32 tiles, K step 2).  FP64: v_mfma_f64_16x16x4f64
// (16 x 16 tiles, K step 4).  Accumulation is in the operand precision, in a fixed order
// (deterministic).
#include "common.h"

#include <type_traits>

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Rot;
template <>
struct Rot<float> {
  static constexpr int TR = 32;   // tile rows and columns
  static constexpr int KS = 2;    // K per MFMA
  static constexpr int KC = 16;   // K per LDS chunk
  static constexpr int NA = 16;   // accumulators per lane per tile
  static constexpr int TMAX = 8;  // row tiles: M <= 256
  typedef f32x16 acc_t;
  static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  // row of accumulator r of lane l inside its tile
  static __device__ __forceinline__ int row(int r, int l) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }
};
template <>
struct Rot<double> {
  static constexpr int TR = 16;
  static constexpr int KS = 4;
  static constexpr int KC = 8;
  static constexpr int NA = 4;
  static constexpr int TMAX = 16;
  typedef f64x4 acc_t;
  static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // (16 x 16 x 4 FP64: register r of lane l holds row l / 16 + 4 r, unlike the FP32 16 x 16
  // shape's 4 (l / 16) + r; both measured with one-hot operands on gfx950)
  static __device__ __forceinline__ int row(int r, int l) { return (l >> 4) + 4 * r; }
};

constexpr int RW = 4;          // wavefronts per workgroup
constexpr int AP = 65;         // LDS stride of one A operand (64 lanes + 1 pad: conflict-free staging)
constexpr int XPD = 2;         // chunks of X requested ahead of the MFMAs (latency of HBM under load)
constexpr unsigned RBAD = 0x40000000u;  // byte offset past every buffer window below (< 2^30)

// Buffer-resource loads (raw, stride 0): the hardware returns 0 at byte offsets >= the
// window size, so rows past K, coefficients past M and columns past the last row need no
// branch.
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t rmk(const void* ptr, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(ptr), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float rld(rsrc_t r, unsigned off, float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ double rld(rsrc_t r, unsigned off, double) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

template <typename T, int TM, int CT>
__global__ __launch_bounds__(RW * 64) void rotate_mfma(const T* __restrict__ Q, int ldq, const T* X, long long ldx,
                                                       T* Y, long long ldy, int M, int K, long long N) {
  using R = Rot<T>;
  constexpr int TR = R::TR, KS = R::KS, KC = R::KC, NS = KC / KS;
  constexpr unsigned S = sizeof(T);
  constexpr int QE = (TR * TM * KC + RW * 64 - 1) / (RW * 64);  // staged coefficients per thread and chunk
  constexpr bool QEXACT = (TR * TM * KC) % (RW * 64) == 0;
  constexpr int NSLOT = XPD + 1;                                 // X operand slots (chunks in registers)
  static_assert((RW * 64) % KC == 0, "whole rows of the chunk per staging pass");
  __shared__ T sq[2][TM * NS * AP];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const long long n0 = ((long long)blockIdx.x * RW + wave) * (CT * TR);
  const int kq = lane / TR;  // K offset of this lane's B operand inside a step
  const int cl = lane % TR;  // column of this lane inside a tile
  const int nch = (K + KC - 1) / KC;

  // staging role: element (i, kk) of the chunk, kk fastest over consecutive threads
  const int skk = tid % KC;
  const int si0 = tid / KC;
  constexpr int SROWS = RW * 64 / KC;
  const rsrc_t rq = rmk(Q, (unsigned)((long long)M * ldq * S));
  T qreg[QE];
  auto stage_load = [&](int c) {
    const int k = c * KC + skk;
#pragma unroll
    for (int e = 0; e < QE; ++e) {
      const int i = si0 + SROWS * e;
      qreg[e] = rld(rq, k < K ? (unsigned)(i * ldq + k) * S : RBAD, T(0));
    }
  };
  auto stage_store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < QE; ++e) {
      const int i = si0 + SROWS * e;
      const int t = i / TR, r = i % TR;
      const int s = skk / KS, q = skk % KS;
      if (QEXACT || i < TR * TM) sq[buf][(t * NS + s) * AP + r + TR * q] = qreg[e];
    }
  };

  // B operands: one window of KS rows of X per step, the lane's row kq and column n inside it
  unsigned xoff[CT];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const long long n = n0 + ct * TR + cl;
    xoff[ct] = (unsigned)(kq * ldx + n) * S;
  }
  T xb[NSLOT][CT][NS];
  auto x_load = [&](int c, int slot) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int k0 = c * KC + s * KS;
      const int rows = K - k0;
      const unsigned bytes = rows > 0 ? (unsigned)((long long)(min(rows, KS) - 1) * ldx + N) * S : 0u;
      const rsrc_t rx = rmk(X + (rows > 0 ? (long long)k0 * ldx : 0), bytes);
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) xb[slot][ct][s] = rld(rx, xoff[ct], T(0));
    }
  };

  typename R::acc_t acc[TM][CT];
#pragma unroll
  for (int t = 0; t < TM; ++t)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int r = 0; r < R::NA; ++r) acc[t][ct][r] = T(0);

  // One chunk: request X for chunk c + XPD and the coefficients of chunk c + 1, run the
  // MFMAs of chunk c, publish the staged coefficients.  The slot is a compile-time
  // constant, so the X operands stay in registers.
  auto chunk = [&](int c, auto slot) {
    constexpr int cur = decltype(slot)::value;
    constexpr int buf = cur & 1;
    if (c + XPD < nch) x_load(c + XPD, (cur + XPD) % NSLOT);
    if (c + 1 < nch) stage_load(c + 1);
    const T* sa = sq[(c & 1)];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
#pragma unroll
      for (int t = 0; t < TM; ++t) {
        const T a = sa[(t * NS + s) * AP + lane];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[t][ct] = R::mfma(a, xb[cur][ct][s], acc[t][ct]);
      }
      // keep the next step's A reads behind this step's MFMAs (they overlap the MFMA
      // pipeline) instead of hoisting a whole chunk of operands into registers
      // (reading them one step ahead instead was measured no faster in FP32 and slower in
      // FP64: 27.9 against 22.1 ms for 192 x 240 at N = 224^3)
      __builtin_amdgcn_sched_barrier(0);
    }
    (void)buf;
    // the other buffer was last read in chunk c - 1, which every wavefront has finished
    if (c + 1 < nch) stage_store((c + 1) & 1);
    __syncthreads();
  };
  stage_load(0);
#pragma unroll
  for (int d = 0; d < XPD; ++d)
    if (d < nch) x_load(d, d);
  stage_store(0);
  __syncthreads();
  for (int c = 0; c < nch; c += NSLOT) {
    chunk(c, std::integral_constant<int, 0>{});
    if (c + 1 < nch) chunk(c + 1, std::integral_constant<int, 1>{});
    if (c + 2 < nch) chunk(c + 2, std::integral_constant<int, 2>{});
  }
  static_assert(NSLOT == 3, "the chunk loop is unrolled over three X slots");

#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const long long n = n0 + ct * TR + cl;
    if (n < N) {
#pragma unroll
      for (int t = 0; t < TM; ++t)
#pragma unroll
        for (int r = 0; r < R::NA; ++r) {
          const int i = t * TR + R::row(r, lane);
          if (i < M) Y[(long long)i * ldy + n] = acc[t][ct][r];
        }
    }
  }
}

template <typename T, int TM, int CT>
int launch_rotate_t(const T* Q, int ldq, const T* X, long long ldx, T* Y, long long ldy, int M, int K, long long N,
                    hipStream_t s) {
  const long long cols = (long long)RW * CT * Rot<T>::TR;
  const long long blocks = (N + cols - 1) / cols;
  if (blocks > 0x7fffffffLL) return IG_EINVAL;
  hipLaunchKernelGGL((rotate_mfma<T, TM, CT>), dim3((unsigned)blocks), dim3(RW * 64), 0, s, Q, ldq, X, ldx, Y, ldy,
                     M, K, N);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

// ct: column tiles per wavefront (0: automatic).  Two tiles share every A operand read
// (half the LDS reads and coefficient staging per flop) but double the accumulators: they
// pay only for short Q (M <= 64 measured at N = 224^3: FP32 2.74 against 2.94 ms, FP64
// 4.78 against 5.22 ms for M = 48, K = 192); for M = 192 one tile keeps 2-3 wavefronts per
// SIMD (FP32 10.1 against 11.5 ms, FP64 22.1 against 28.9 ms for K = 240).
template <typename T>
int launch_rotate(const T* Q, int ldq, const T* X, long long ldx, T* Y, long long ldy, int M, int K, long long N,
                  int ct, hipStream_t s) {
  if (M <= 0 || K <= 0 || N <= 0 || ldq < K || ldx < N || ldy < N) return IG_EINVAL;
  if (!Q || !X || !Y) return IG_EINVAL;
  // buffer windows (the coefficients; KS rows of X) must stay below the 2^30-byte marker
  if ((long long)M * ldq * (long long)sizeof(T) >= (long long)RBAD) return IG_EUNSUPPORTED;
  if (((long long)Rot<T>::KS * ldx + 64) * (long long)sizeof(T) >= (long long)RBAD) return IG_EUNSUPPORTED;
  constexpr int TR = Rot<T>::TR;
  const int tm = (M + TR - 1) / TR;
  if (tm > Rot<T>::TMAX) return IG_EUNSUPPORTED;
  // Y may overlap X only row-aligned (same row stride, offset by whole rows), so that every
  // aliased element lies in the same column, i.e. in the same wavefront's columns
  {
    const long long x0 = (long long)(uintptr_t)X, y0 = (long long)(uintptr_t)Y;
    const long long x1 = x0 + ((long long)(K - 1) * ldx + N) * (long long)sizeof(T);
    const long long y1 = y0 + ((long long)(M - 1) * ldy + N) * (long long)sizeof(T);
    if (y0 < x1 && x0 < y1) {
      const long long row = ldx * (long long)sizeof(T);
      if (ldy != ldx || (y0 - x0) % row != 0) return IG_EINVAL;
    }
  }
  if (ct == 0) ct = tm * Rot<T>::TR <= 64 ? 2 : 1;
  if (ct != 1 && ct != 2) return IG_EINVAL;
#define IG_ROT_CASE(TM_)                                                                   \
  case TM_:                                                                                \
    return ct == 2 ? launch_rotate_t<T, TM_, 2>(Q, ldq, X, ldx, Y, ldy, M, K, N, s)        \
                   : launch_rotate_t<T, TM_, 1>(Q, ldq, X, ldx, Y, ldy, M, K, N, s);
  switch (tm) {
    IG_ROT_CASE(1)
    IG_ROT_CASE(2)
    IG_ROT_CASE(3)
    IG_ROT_CASE(4)
    IG_ROT_CASE(5)
    IG_ROT_CASE(6)
    IG_ROT_CASE(7)
    IG_ROT_CASE(8)
    default:
      break;
  }
  if constexpr (Rot<T>::TMAX > 8) {
    switch (tm) {
      IG_ROT_CASE(9)
      IG_ROT_CASE(10)
      IG_ROT_CASE(11)
      IG_ROT_CASE(12)
      IG_ROT_CASE(13)
      IG_ROT_CASE(14)
      IG_ROT_CASE(15)
      IG_ROT_CASE(16)
      default:
        break;
    }
  }
#undef IG_ROT_CASE
  return IG_EUNSUPPORTED;
}

}  // namespace

extern "C" {

int ig_rotate_f32(const float* Q, int ldq, const float* X, long long ldx, float* Y, long long ldy, int M, int K,
                  long long N, int ct, hipStream_t stream) {
  return launch_rotate<float>(Q, ldq, X, ldx, Y, ldy, M, K, N, ct, stream);
}

int ig_rotate_f64(const double* Q, int ldq, const double* X, long long ldx, double* Y, long long ldy, int M, int K,
                  long long N, int ct, hipStream_t stream) {
  return launch_rotate<double>(Q, ldq, X, ldx, Y, ldy, M, K, N, ct, stream);
}

}  // extern "C"
