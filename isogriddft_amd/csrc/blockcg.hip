// Row-wise block conjugate-gradient kernels: k independent CG iterations, one per row
// of a (k, N) block with row stride ld, under the inner product <a, b>_w = sum w a b.
// They carry the kinetic preconditioner of the block eigensolver (a few CG steps on
// (T + c) x = r per residual row); the operator apply itself is the stencil kernel.
//
//   row_dots:     out[i] = sum_n w[n] a[i][n] b[i][n]
//   cg_update:    x += alpha_i p, r -= alpha_i q, out[i] = sum_n w[n] dinv[n] r[i][n]^2
//   cg_direction: p = dinv r + beta_i p
//   cg_start:     r = isw b - q, p = dinv r, out[i] = sum_n w[n] r[i][n] p[i][n]
//                 (the first residual of x0 = dinv isw b, with q = A x0)
//   cg_finish:    y = sw (x + alpha_i p) (the last x update mapped back, x not written)
// alpha / beta are FP64 device vectors of length k (no host round trip).  Sums are FP64;
// each (row, chunk) block writes a partial and one block per row adds the partials in a
// fixed order, so results are bit-reproducible.
#include "common.h"

#include <initializer_list>

namespace {

constexpr int CG_MAX_CHUNKS = 256;

// Workgroup -> (row, column chunk) of a k-row, nch-chunk launch on a 1-D grid of
// 8 k ceil(nch / 8) workgroups.  Workgroup b runs on XCD b % 8, and each XCD takes its
// chunks (xcd, xcd + 8, ...) for every row in turn, so the per-point fields a chunk reads
// for all rows (dinv, w, isw, sw) come into that XCD's L2 once instead of once per row.
__device__ __forceinline__ bool row_chunk(int k, int nch, int& row, int& cx) {
  const int b = blockIdx.x;
  const int r = b >> 3;
  row = r % k;
  cx = (r / k) * 8 + (b & 7);
  return cx < nch;
}

// 16-byte vectors: float4 / double2.  The vector kernels run when N, ld and the chunk are
// multiples of VW and every base pointer is 16-byte aligned (the host checks), otherwise
// the scalar ones.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float t __attribute__((ext_vector_type(4)));
  static constexpr int W = 4;
};
template <>
struct Vec16<double> {
  typedef double t __attribute__((ext_vector_type(2)));
  static constexpr int W = 2;
};

template <typename T>
__device__ __forceinline__ double vsum_prod(const typename Vec16<T>::t& a, const typename Vec16<T>::t& b) {
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < Vec16<T>::W; ++e) s += (double)a[e] * (double)b[e];
  return s;
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void row_dots_vec(const T* __restrict__ a, const T* __restrict__ b,
                                                           const T* __restrict__ w, long long N, long long ld,
                                                           long long chunk, int k, int nch,
                                                           double* __restrict__ partial) {
  using VT = typename Vec16<T>::t;
  constexpr int VW = Vec16<T>::W;
  __shared__ double sm[4];
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T* ar = a + row * ld;
  const T* br = b + row * ld;
  double s = 0.0;
  for (long long n = beg + (long long)threadIdx.x * VW; n < end; n += (long long)IG_THREADS * VW) {
    const VT va = *reinterpret_cast<const VT*>(ar + n);
    const VT vb = *reinterpret_cast<const VT*>(br + n);
    if (w) {
      const VT vw = *reinterpret_cast<const VT*>(w + n);
#pragma unroll
      for (int e = 0; e < VW; ++e) s += (double)va[e] * (double)vb[e] * (double)vw[e];
    } else {
      s += vsum_prod<T>(va, vb);
    }
  }
  double v[1] = {s};
  ig_block_sum<1>(v, sm);
  if (threadIdx.x == 0) partial[(long long)row * nch + cx] = v[0];
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void cg_update_vec(T* __restrict__ x, T* __restrict__ r,
                                                            const T* __restrict__ p, const T* __restrict__ q,
                                                            const T* __restrict__ dinv, const T* __restrict__ w,
                                                            const double* __restrict__ alpha, long long N,
                                                            long long ld, long long chunk, int k, int nch,
                                                            double* __restrict__ partial) {
  using VT = typename Vec16<T>::t;
  constexpr int VW = Vec16<T>::W;
  __shared__ double sm[4];
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T a = (T)alpha[row];
  T* xr = x + row * ld;
  T* rr = r + row * ld;
  const T* pr = p + row * ld;
  const T* qr = q + row * ld;
  double s = 0.0;
  for (long long n = beg + (long long)threadIdx.x * VW; n < end; n += (long long)IG_THREADS * VW) {
    VT vx = *reinterpret_cast<const VT*>(xr + n);
    VT vr = *reinterpret_cast<const VT*>(rr + n);
    const VT vp = *reinterpret_cast<const VT*>(pr + n);
    const VT vq = *reinterpret_cast<const VT*>(qr + n);
    const VT vd = *reinterpret_cast<const VT*>(dinv + n);
    VT vw;
    if (w) vw = *reinterpret_cast<const VT*>(w + n);
#pragma unroll
    for (int e = 0; e < VW; ++e) {
      vx[e] = vx[e] + a * vp[e];
      vr[e] = vr[e] - a * vq[e];
      const double t = (double)vr[e] * (double)vr[e] * (double)vd[e];
      s += w ? t * (double)vw[e] : t;
    }
    *reinterpret_cast<VT*>(xr + n) = vx;
    *reinterpret_cast<VT*>(rr + n) = vr;
  }
  double v[1] = {s};
  ig_block_sum<1>(v, sm);
  if (threadIdx.x == 0) partial[(long long)row * nch + cx] = v[0];
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void cg_direction_vec(T* __restrict__ p, const T* __restrict__ r,
                                                               const T* __restrict__ dinv,
                                                               const double* __restrict__ beta, long long N,
                                                               long long ld, long long chunk, int k, int nch) {
  using VT = typename Vec16<T>::t;
  constexpr int VW = Vec16<T>::W;
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T b = (T)beta[row];
  T* pr = p + row * ld;
  const T* rr = r + row * ld;
  for (long long n = beg + (long long)threadIdx.x * VW; n < end; n += (long long)IG_THREADS * VW) {
    VT vp = *reinterpret_cast<const VT*>(pr + n);
    const VT vr = *reinterpret_cast<const VT*>(rr + n);
    const VT vd = *reinterpret_cast<const VT*>(dinv + n);
#pragma unroll
    for (int e = 0; e < VW; ++e) vp[e] = vd[e] * vr[e] + b * vp[e];
    *reinterpret_cast<VT*>(pr + n) = vp;
  }
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void row_dots_kernel(const T* __restrict__ a, const T* __restrict__ b,
                                                              const T* __restrict__ w, long long N, long long ld,
                                                              long long chunk, int k, int nch,
                                                              double* __restrict__ partial) {
  __shared__ double sm[4];
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T* ar = a + row * ld;
  const T* br = b + row * ld;
  double s = 0.0;
  for (long long n = beg + threadIdx.x; n < end; n += IG_THREADS) {
    const double v = (double)ar[n] * (double)br[n];
    s += w ? v * (double)w[n] : v;
  }
  double v[1] = {s};
  ig_block_sum<1>(v, sm);
  if (threadIdx.x == 0) partial[(long long)row * nch + cx] = v[0];
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void cg_update_kernel(T* __restrict__ x, T* __restrict__ r,
                                                               const T* __restrict__ p, const T* __restrict__ q,
                                                               const T* __restrict__ dinv, const T* __restrict__ w,
                                                               const double* __restrict__ alpha, long long N,
                                                               long long ld, long long chunk, int k, int nch,
                                                               double* __restrict__ partial) {
  __shared__ double sm[4];
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T a = (T)alpha[row];
  T* xr = x + row * ld;
  T* rr = r + row * ld;
  const T* pr = p + row * ld;
  const T* qr = q + row * ld;
  double s = 0.0;
  for (long long n = beg + threadIdx.x; n < end; n += IG_THREADS) {
    xr[n] = xr[n] + a * pr[n];
    const T rn = rr[n] - a * qr[n];
    rr[n] = rn;
    const double v = (double)rn * (double)rn * (double)dinv[n];
    s += w ? v * (double)w[n] : v;
  }
  double v[1] = {s};
  ig_block_sum<1>(v, sm);
  if (threadIdx.x == 0) partial[(long long)row * nch + cx] = v[0];
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void cg_direction_kernel(T* __restrict__ p, const T* __restrict__ r,
                                                                  const T* __restrict__ dinv,
                                                                  const double* __restrict__ beta, long long N,
                                                                  long long ld, long long chunk, int k, int nch) {
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T b = (T)beta[row];
  T* pr = p + row * ld;
  const T* rr = r + row * ld;
  for (long long n = beg + threadIdx.x; n < end; n += IG_THREADS) pr[n] = dinv[n] * rr[n] + b * pr[n];
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void cg_start_vec(const T* __restrict__ b, long long ldb,
                                                           const T* __restrict__ isw, const T* __restrict__ q,
                                                           const T* __restrict__ dinv, const T* __restrict__ w,
                                                           T* __restrict__ r, T* __restrict__ p, long long N,
                                                           long long ld, long long chunk, int k, int nch,
                                                           double* __restrict__ partial) {
  using VT = typename Vec16<T>::t;
  constexpr int VW = Vec16<T>::W;
  __shared__ double sm[4];
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T* br = b + row * ldb;
  const T* qr = q + row * ld;
  T* rr = r + row * ld;
  T* pr = p + row * ld;
  double s = 0.0;
  for (long long n = beg + (long long)threadIdx.x * VW; n < end; n += (long long)IG_THREADS * VW) {
    const VT vb = *reinterpret_cast<const VT*>(br + n);
    const VT vs = *reinterpret_cast<const VT*>(isw + n);
    const VT vq = *reinterpret_cast<const VT*>(qr + n);
    const VT vd = *reinterpret_cast<const VT*>(dinv + n);
    VT vw;
    if (w) vw = *reinterpret_cast<const VT*>(w + n);
    VT vr, vp;
#pragma unroll
    for (int e = 0; e < VW; ++e) {
      vr[e] = vs[e] * vb[e] - vq[e];
      vp[e] = vd[e] * vr[e];
      const double t = (double)vr[e] * (double)vp[e];
      s += w ? t * (double)vw[e] : t;
    }
    *reinterpret_cast<VT*>(rr + n) = vr;
    *reinterpret_cast<VT*>(pr + n) = vp;
  }
  double v[1] = {s};
  ig_block_sum<1>(v, sm);
  if (threadIdx.x == 0) partial[(long long)row * nch + cx] = v[0];
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void cg_start_kernel(const T* __restrict__ b, long long ldb,
                                                              const T* __restrict__ isw, const T* __restrict__ q,
                                                              const T* __restrict__ dinv, const T* __restrict__ w,
                                                              T* __restrict__ r, T* __restrict__ p, long long N,
                                                              long long ld, long long chunk, int k, int nch,
                                                              double* __restrict__ partial) {
  __shared__ double sm[4];
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T* br = b + row * ldb;
  const T* qr = q + row * ld;
  T* rr = r + row * ld;
  T* pr = p + row * ld;
  double s = 0.0;
  for (long long n = beg + threadIdx.x; n < end; n += IG_THREADS) {
    const T rn = isw[n] * br[n] - qr[n];
    const T pn = dinv[n] * rn;
    rr[n] = rn;
    pr[n] = pn;
    const double t = (double)rn * (double)pn;
    s += w ? t * (double)w[n] : t;
  }
  double v[1] = {s};
  ig_block_sum<1>(v, sm);
  if (threadIdx.x == 0) partial[(long long)row * nch + cx] = v[0];
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void cg_finish_vec(const T* __restrict__ x, const T* __restrict__ p,
                                                            const T* __restrict__ sw,
                                                            const double* __restrict__ alpha, T* __restrict__ y,
                                                            long long ldy, long long N, long long ld,
                                                            long long chunk, int k, int nch) {
  using VT = typename Vec16<T>::t;
  constexpr int VW = Vec16<T>::W;
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T a = (T)alpha[row];
  const T* xr = x + row * ld;
  const T* pr = p + row * ld;
  T* yr = y + row * ldy;
  for (long long n = beg + (long long)threadIdx.x * VW; n < end; n += (long long)IG_THREADS * VW) {
    const VT vx = *reinterpret_cast<const VT*>(xr + n);
    const VT vp = *reinterpret_cast<const VT*>(pr + n);
    const VT vs = *reinterpret_cast<const VT*>(sw + n);
    VT vy;
#pragma unroll
    for (int e = 0; e < VW; ++e) vy[e] = vs[e] * (vx[e] + a * vp[e]);
    *reinterpret_cast<VT*>(yr + n) = vy;
  }
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void cg_finish_kernel(const T* __restrict__ x, const T* __restrict__ p,
                                                               const T* __restrict__ sw,
                                                               const double* __restrict__ alpha, T* __restrict__ y,
                                                               long long ldy, long long N, long long ld,
                                                               long long chunk, int k, int nch) {
  int row, cx;
  if (!row_chunk(k, nch, row, cx)) return;
  const long long beg = (long long)cx * chunk;
  const long long end = min(beg + chunk, N);
  const T a = (T)alpha[row];
  const T* xr = x + row * ld;
  const T* pr = p + row * ld;
  T* yr = y + row * ldy;
  for (long long n = beg + threadIdx.x; n < end; n += IG_THREADS) yr[n] = sw[n] * (xr[n] + a * pr[n]);
}

__global__ void rows_finish(const double* __restrict__ partial, int nparts, int k, double* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < nparts; ++c) s += partial[(long long)i * nparts + c];
    out[i] = s;
  }
}

struct RowPlan {
  int nch;
  long long chunk;
};

// Workgroups of these kernels resident on the device at once: every one runs at 8 waves
// per SIMD, i.e. 8 workgroups of 256 lanes per CU.
long long resident_groups() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  return 8LL * cus;
}

// k rows split into chunks so that the k x nch workgroups fit in one residency round:
// all of them start together and carry equal work, so none is left to stream its chunk
// alone after the rest have finished (ceil(2048 / k) chunks overshoot one round for most
// k, e.g. 2064 workgroups for k = 48 against 2048 resident).
RowPlan row_plan(int k, long long N, int align) {
  long long nch = resident_groups() / k;
  const long long cap = (N + IG_THREADS - 1) / IG_THREADS;  // at least one element per thread
  if (nch > cap) nch = cap;
  if (nch > CG_MAX_CHUNKS) nch = CG_MAX_CHUNKS;
  if (nch >= 8) nch -= nch % 8;  // whole groups of 8 chunks: one per XCD, none idle
  if (nch < 1) nch = 1;
  RowPlan g;
  g.chunk = (N + nch - 1) / nch;
  g.chunk = (g.chunk + align - 1) / align * align;
  g.nch = (int)((N + g.chunk - 1) / g.chunk);
  return g;
}

inline unsigned cg_blocks(int k, int nch) { return (unsigned)(8LL * k * ((nch + 7) / 8)); }

inline bool al16(const void* q) { return q == nullptr || (reinterpret_cast<uintptr_t>(q) & 15u) == 0; }

template <typename T>
inline bool use_vec(long long N, long long ld, std::initializer_list<const void*> ptrs) {
  constexpr int VW = Vec16<T>::W;
  if (N % VW != 0 || ld % VW != 0) return false;
  for (const void* q : ptrs)
    if (!al16(q)) return false;
  return true;
}

template <typename T>
int row_dots(const T* a, const T* b, const T* w, int k, long long N, long long ld, double* work, double* out,
             hipStream_t s) {
  if (!a || !b || !work || !out || k <= 0 || k > 65535 || N <= 0 || ld < N) return IG_EINVAL;
  const bool vec = use_vec<T>(N, ld, {a, b, w});
  const RowPlan g = row_plan(k, N, vec ? Vec16<T>::W : 1);
  const dim3 grid(cg_blocks(k, g.nch));
  if (vec)
    hipLaunchKernelGGL(row_dots_vec<T>, grid, dim3(IG_THREADS), 0, s, a, b, w, N, ld, g.chunk,
                       k, g.nch, work);
  else
    hipLaunchKernelGGL(row_dots_kernel<T>, grid, dim3(IG_THREADS), 0, s, a, b, w, N, ld,
                       g.chunk, k, g.nch, work);
  hipLaunchKernelGGL(rows_finish, dim3((k + 255) / 256), dim3(256), 0, s, work, g.nch, k, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int cg_update(T* x, T* r, const T* p, const T* q, const T* dinv, const T* w, const double* alpha, int k, long long N,
              long long ld, double* work, double* out, hipStream_t s) {
  if (!x || !r || !p || !q || !dinv || !alpha || !work || !out || k <= 0 || k > 65535 || N <= 0 || ld < N)
    return IG_EINVAL;
  const bool vec = use_vec<T>(N, ld, {x, r, p, q, dinv, w});
  const RowPlan g = row_plan(k, N, vec ? Vec16<T>::W : 1);
  const dim3 grid(cg_blocks(k, g.nch));
  if (vec)
    hipLaunchKernelGGL(cg_update_vec<T>, grid, dim3(IG_THREADS), 0, s, x, r, p, q, dinv, w, alpha, N, ld,
                       g.chunk, k, g.nch, work);
  else
    hipLaunchKernelGGL(cg_update_kernel<T>, grid, dim3(IG_THREADS), 0, s, x, r, p, q, dinv, w, alpha, N,
                       ld, g.chunk, k, g.nch, work);
  hipLaunchKernelGGL(rows_finish, dim3((k + 255) / 256), dim3(256), 0, s, work, g.nch, k, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int cg_direction(T* p, const T* r, const T* dinv, const double* beta, int k, long long N, long long ld,
                 hipStream_t s) {
  if (!p || !r || !dinv || !beta || k <= 0 || k > 65535 || N <= 0 || ld < N) return IG_EINVAL;
  const bool vec = use_vec<T>(N, ld, {p, r, dinv});
  const RowPlan g = row_plan(k, N, vec ? Vec16<T>::W : 1);
  const dim3 grid(cg_blocks(k, g.nch));
  if (vec)
    hipLaunchKernelGGL(cg_direction_vec<T>, grid, dim3(IG_THREADS), 0, s, p, r, dinv, beta, N, ld,
                       g.chunk, k, g.nch);
  else
    hipLaunchKernelGGL(cg_direction_kernel<T>, grid, dim3(IG_THREADS), 0, s, p, r, dinv, beta, N, ld,
                       g.chunk, k, g.nch);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int cg_start(const T* b, long long ldb, const T* isw, const T* q, const T* dinv, const T* w, T* r, T* p, int k,
             long long N, long long ld, double* work, double* out, hipStream_t s) {
  if (!b || !isw || !q || !dinv || !r || !p || !work || !out || k <= 0 || k > 65535 || N <= 0 || ld < N ||
      ldb < N)
    return IG_EINVAL;
  const bool vec = use_vec<T>(N, ld, {b, isw, q, dinv, w, r, p}) && ldb % Vec16<T>::W == 0;
  const RowPlan g = row_plan(k, N, vec ? Vec16<T>::W : 1);
  const dim3 grid(cg_blocks(k, g.nch));
  if (vec)
    hipLaunchKernelGGL(cg_start_vec<T>, grid, dim3(IG_THREADS), 0, s, b, ldb, isw, q, dinv, w, r, p, N,
                       ld, g.chunk, k, g.nch, work);
  else
    hipLaunchKernelGGL(cg_start_kernel<T>, grid, dim3(IG_THREADS), 0, s, b, ldb, isw, q, dinv, w, r, p,
                       N, ld, g.chunk, k, g.nch, work);
  hipLaunchKernelGGL(rows_finish, dim3((k + 255) / 256), dim3(256), 0, s, work, g.nch, k, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int cg_finish(const T* x, const T* p, const T* sw, const double* alpha, T* y, long long ldy, int k, long long N,
              long long ld, hipStream_t s) {
  if (!x || !p || !sw || !alpha || !y || k <= 0 || k > 65535 || N <= 0 || ld < N || ldy < N) return IG_EINVAL;
  const bool vec = use_vec<T>(N, ld, {x, p, sw, y}) && ldy % Vec16<T>::W == 0;
  const RowPlan g = row_plan(k, N, vec ? Vec16<T>::W : 1);
  const dim3 grid(cg_blocks(k, g.nch));
  if (vec)
    hipLaunchKernelGGL(cg_finish_vec<T>, grid, dim3(IG_THREADS), 0, s, x, p, sw, alpha, y, ldy, N, ld,
                       g.chunk, k, g.nch);
  else
    hipLaunchKernelGGL(cg_finish_kernel<T>, grid, dim3(IG_THREADS), 0, s, x, p, sw, alpha, y, ldy, N, ld,
                       g.chunk, k, g.nch);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

}  // namespace

extern "C" {

int ig_row_dots_work_size(int k) { return k * CG_MAX_CHUNKS; }

int ig_row_dots_f64(const double* a, const double* b, const double* w, int k, long long N, long long ld,
                    double* work, double* out, hipStream_t s) {
  return row_dots<double>(a, b, w, k, N, ld, work, out, s);
}
int ig_row_dots_f32(const float* a, const float* b, const float* w, int k, long long N, long long ld,
                    double* work, double* out, hipStream_t s) {
  return row_dots<float>(a, b, w, k, N, ld, work, out, s);
}
int ig_cg_update_f64(double* x, double* r, const double* p, const double* q, const double* dinv, const double* w,
                     const double* alpha, int k, long long N, long long ld, double* work, double* out,
                     hipStream_t s) {
  return cg_update<double>(x, r, p, q, dinv, w, alpha, k, N, ld, work, out, s);
}
int ig_cg_update_f32(float* x, float* r, const float* p, const float* q, const float* dinv, const float* w,
                     const double* alpha, int k, long long N, long long ld, double* work, double* out,
                     hipStream_t s) {
  return cg_update<float>(x, r, p, q, dinv, w, alpha, k, N, ld, work, out, s);
}
int ig_cg_direction_f64(double* p, const double* r, const double* dinv, const double* beta, int k, long long N,
                        long long ld, hipStream_t s) {
  return cg_direction<double>(p, r, dinv, beta, k, N, ld, s);
}
int ig_cg_direction_f32(float* p, const float* r, const float* dinv, const double* beta, int k, long long N,
                        long long ld, hipStream_t s) {
  return cg_direction<float>(p, r, dinv, beta, k, N, ld, s);
}

int ig_cg_start_f64(const double* b, long long ldb, const double* isw, const double* q, const double* dinv,
                    const double* w, double* r, double* p, int k, long long N, long long ld, double* work,
                    double* out, hipStream_t s) {
  return cg_start<double>(b, ldb, isw, q, dinv, w, r, p, k, N, ld, work, out, s);
}
int ig_cg_start_f32(const float* b, long long ldb, const float* isw, const float* q, const float* dinv,
                    const float* w, float* r, float* p, int k, long long N, long long ld, double* work, double* out,
                    hipStream_t s) {
  return cg_start<float>(b, ldb, isw, q, dinv, w, r, p, k, N, ld, work, out, s);
}
int ig_cg_finish_f64(const double* x, const double* p, const double* sw, const double* alpha, double* y,
                     long long ldy, int k, long long N, long long ld, hipStream_t s) {
  return cg_finish<double>(x, p, sw, alpha, y, ldy, k, N, ld, s);
}
int ig_cg_finish_f32(const float* x, const float* p, const float* sw, const double* alpha, float* y, long long ldy,
                     int k, long long N, long long ld, hipStream_t s) {
  return cg_finish<float>(x, p, sw, alpha, y, ldy, k, N, ld, s);
}

}  // extern "C"
