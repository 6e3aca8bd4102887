// Reductions and fused Krylov (preconditioned BiCGSTAB) vector updates.
//
// All Krylov scalars live in a small device array `sc` (double) so an iteration
// never synchronises with the host:
//   sc[0] rho = (rhat, r)   sc[1] rho_prev   sc[2] alpha   sc[3] omega
//   sc[4] (rhat, v)         sc[5] (t, s)     sc[6] (t, t)  sc[7] (r, r)
//   sc[8] breakdown flag    sc[9] breakdown raised by the last x step (folded into
//                           sc[8] by the single-block reduction that follows it)
// The reference solves the same interior systems with SciPy BiCGSTAB
// (poisson/open_boundary.py:1017-1064, :1229-1276); iterates differ, the discrete
// solution is the same.
#include "common.h"

#include <mutex>

namespace {

constexpr int RED_BLOCKS_MAX = 1024;

// Partial sums of up to three (weighted) dot products.
template <typename T, int ND>
__global__ __launch_bounds__(IG_THREADS) void dots_partial(const T* __restrict__ a0, const T* __restrict__ b0,
                                                           const T* __restrict__ a1, const T* __restrict__ b1,
                                                           const T* __restrict__ a2, const T* __restrict__ b2,
                                                           const T* __restrict__ w, long long n,
                                                           double* __restrict__ partial) {
  __shared__ double smem[4 * 3];
  double acc[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) acc[i] = 0.0;
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < n;
       t += (long long)gridDim.x * IG_THREADS) {
    const double ww = w ? (double)w[t] : 1.0;
    acc[0] += (double)a0[t] * (double)b0[t] * ww;
    if (ND > 1) acc[1] += (double)a1[t] * (double)b1[t] * ww;
    if (ND > 2) acc[2] += (double)a2[t] * (double)b2[t] * ww;
  }
  ig_block_sum<ND>(acc, smem);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < ND; ++i) partial[blockIdx.x * ND + i] = acc[i];
  }
}

// Sum `nparts` partials of ND values; out[slot_i] (op: 0 store, 1 = shift rho).
// With fold_flag, out[9] (the x step's breakdown) is also folded into the flag out[8].
template <int ND>
__global__ __launch_bounds__(IG_THREADS) void dots_final(const double* __restrict__ partial, int nparts,
                                                         double* __restrict__ out, int o0, int o1, int o2,
                                                         int shift_rho, int fold_flag) {
  __shared__ double smem[4 * 3];
  double acc[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) acc[i] = 0.0;
  for (int t = threadIdx.x; t < nparts; t += IG_THREADS) {
#pragma unroll
    for (int i = 0; i < ND; ++i) acc[i] += partial[t * ND + i];
  }
  ig_block_sum<ND>(acc, smem);
  if (threadIdx.x == 0) {
    if (shift_rho) out[1] = out[0];
    if (fold_flag && out[9] != 0.0) out[8] = 1.0;
    out[o0] = acc[0];
    if (ND > 1) out[o1] = acc[1];
    if (ND > 2) out[o2] = acc[2];
  }
}

inline int red_blocks(long long n) {
  long long b = (n + IG_THREADS * 8 - 1) / (IG_THREADS * 8);
  if (b < 1) b = 1;
  if (b > RED_BLOCKS_MAX) b = RED_BLOCKS_MAX;
  return (int)b;
}

template <typename T>
int dots_impl(int nd, const T* a0, const T* b0, const T* a1, const T* b1, const T* a2, const T* b2,
              const T* w, long long n, double* work, double* out, int o0, int o1, int o2,
              int shift_rho, hipStream_t s) {
  if (nd < 1 || nd > 3 || n <= 0 || !work || !out || !a0 || !b0) return IG_EINVAL;
  if (nd > 1 && (!a1 || !b1)) return IG_EINVAL;
  if (nd > 2 && (!a2 || !b2)) return IG_EINVAL;
  const int nb = red_blocks(n);
  if (nd == 1) {
    hipLaunchKernelGGL((dots_partial<T, 1>), dim3(nb), dim3(IG_THREADS), 0, s, a0, b0, a1, b1, a2, b2, w, n, work);
    hipLaunchKernelGGL((dots_final<1>), dim3(1), dim3(IG_THREADS), 0, s, work, nb, out, o0, o1, o2, shift_rho,
                       0);
  } else if (nd == 2) {
    hipLaunchKernelGGL((dots_partial<T, 2>), dim3(nb), dim3(IG_THREADS), 0, s, a0, b0, a1, b1, a2, b2, w, n, work);
    hipLaunchKernelGGL((dots_final<2>), dim3(1), dim3(IG_THREADS), 0, s, work, nb, out, o0, o1, o2, shift_rho,
                       0);
  } else {
    hipLaunchKernelGGL((dots_partial<T, 3>), dim3(nb), dim3(IG_THREADS), 0, s, a0, b0, a1, b1, a2, b2, w, n, work);
    hipLaunchKernelGGL((dots_final<3>), dim3(1), dim3(IG_THREADS), 0, s, work, nb, out, o0, o1, o2, shift_rho,
                       0);
  }
  IG_CHECK_LAUNCH();
  return IG_OK;
}

// p = r + beta (p - omega v);  phat = dinv * p.
template <typename T>
__global__ __launch_bounds__(IG_THREADS) void bicg_p(long long n, const T* __restrict__ r, T* __restrict__ pv,
                                                     const T* __restrict__ v, const T* __restrict__ dinv,
                                                     T* __restrict__ phat, const double* __restrict__ sc) {
  // after a breakdown (flag set by an earlier step of the stream) every update is a no-op:
  // the iterate stays the last finite one
  if (sc[8] != 0.0) return;
  const double beta = (sc[0] / sc[1]) * (sc[2] / sc[3]);
  const double omega = sc[3];
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < n;
       t += (long long)gridDim.x * IG_THREADS) {
    const T pn = r[t] + T(beta) * (pv[t] - T(omega) * v[t]);
    pv[t] = pn;
    phat[t] = dinv ? dinv[t] * pn : pn;
  }
}

// alpha = rho / (rhat, v);  s = r - alpha v;  shat = dinv * s.
template <typename T>
__global__ __launch_bounds__(IG_THREADS) void bicg_s(long long n, const T* __restrict__ r, const T* __restrict__ v,
                                                     const T* __restrict__ dinv, T* __restrict__ sv,
                                                     T* __restrict__ shat, double* __restrict__ sc) {
  if (sc[8] != 0.0) return;
  // every block evaluates the same breakdown test, so none writes s from a non-finite alpha
  const bool broken = !(fabs(sc[4]) > 1e-300) || !(fabs(sc[0]) > 1e-300);
  const double alpha = broken ? 0.0 : sc[0] / sc[4];
  if (broken) {
    __syncthreads();  // every lane of block 0 has read sc[8] before it is set
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[8] = 1.0;
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[2] = alpha;
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < n;
       t += (long long)gridDim.x * IG_THREADS) {
    const T sn = r[t] - T(alpha) * v[t];
    sv[t] = sn;
    shat[t] = dinv ? dinv[t] * sn : sn;
  }
}

// omega = (t, s) / (t, t);  x += alpha phat + omega shat;  r = s - omega t;
// block partials of (rhat, r) and (r, r).
template <typename T>
__global__ __launch_bounds__(IG_THREADS) void bicg_x(long long n, T* __restrict__ x, const T* __restrict__ phat,
                                                     const T* __restrict__ shat, const T* __restrict__ sv,
                                                     const T* __restrict__ tv, T* __restrict__ r,
                                                     const T* __restrict__ rhat, double* __restrict__ sc,
                                                     double* __restrict__ partial) {
  __shared__ double smem[4 * 2];
  // After a breakdown of an earlier step (flag set) x and r keep their last values and the
  // partials re-measure that r, so the host reads the residual of the iterate it gets back.
  // (t, t) = 0 here (t = A shat vanishes, e.g. s = 0 when the rhs spans an invariant
  // subspace) still takes x += alpha phat, r = s, then raises the flag: the next steps,
  // whose beta would divide by omega = 0, are no-ops.  Every block sees the same scalars:
  // this kernel raises sc[9], never sc[8], because a block that starts after block 0 has
  // finished would otherwise read the new flag and skip the update the others made.
  const bool frozen = sc[8] != 0.0;
  const bool broken = frozen || !(sc[6] > 0.0);
  const double alpha = sc[2];
  const double omega = broken ? 0.0 : sc[5] / sc[6];
  double acc[2] = {0.0, 0.0};
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < n;
       t += (long long)gridDim.x * IG_THREADS) {
    T rn;
    if (frozen) {
      rn = r[t];
    } else {
      x[t] += T(alpha) * phat[t] + T(omega) * shat[t];
      rn = sv[t] - T(omega) * tv[t];
      r[t] = rn;
    }
    acc[0] += (double)rhat[t] * (double)rn;
    acc[1] += (double)rn * (double)rn;
  }
  ig_block_sum<2>(acc, smem);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (broken) sc[9] = 1.0;
    else sc[3] = omega;
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x * 2 + 0] = acc[0];
    partial[blockIdx.x * 2 + 1] = acc[1];
  }
}

template <typename T>
int bicg_step_p(long long n, const T* r, T* pv, const T* v, const T* dinv, T* phat, const double* sc,
                hipStream_t s) {
  if (n <= 0 || !r || !pv || !v || !phat || !sc) return IG_EINVAL;
  hipLaunchKernelGGL(bicg_p<T>, dim3(red_blocks(n)), dim3(IG_THREADS), 0, s, n, r, pv, v, dinv, phat, sc);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int bicg_step_s(long long n, const T* r, const T* v, const T* dinv, T* sv, T* shat, double* sc,
                hipStream_t s) {
  if (n <= 0 || !r || !v || !sv || !shat || !sc) return IG_EINVAL;
  hipLaunchKernelGGL(bicg_s<T>, dim3(red_blocks(n)), dim3(IG_THREADS), 0, s, n, r, v, dinv, sv, shat, sc);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int bicg_step_x(long long n, T* x, const T* phat, const T* shat, const T* sv, const T* tv, T* r,
                const T* rhat, double* sc, double* work, hipStream_t s) {
  if (n <= 0 || !x || !phat || !shat || !sv || !tv || !r || !rhat || !sc || !work) return IG_EINVAL;
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(bicg_x<T>, dim3(nb), dim3(IG_THREADS), 0, s, n, x, phat, shat, sv, tv, r, rhat, sc, work);
  hipLaunchKernelGGL((dots_final<2>), dim3(1), dim3(IG_THREADS), 0, s, work, nb, sc, 0, 7, 0, 1, 1);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

// ---- the entry point that allocates nothing: its partial sums go to a file-scope device
// array instead of a caller workspace.  Each GPU holds its own copy of `scratch` (24 KiB:
// at most RED_BLOCKS_MAX blocks x 3 partials, every length n fits), whose address is looked
// up once per device with that device current.  In host code `scratch` names a host-side
// stand-in, never a device address: only the looked-up pointer is passed to kernels.
constexpr int NOALLOC_SCRATCH_DOUBLES = RED_BLOCKS_MAX * 3;
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mutex;

int scratch_for_current_device(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return IG_EHIP;
  if (dev >= NOALLOC_MAX_DEVICES) return IG_EUNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return IG_EHIP;
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return IG_OK;
}

}  // namespace

extern "C" {

int ig_reduction_work_size(long long n) { return red_blocks(n) * 3; }

int ig_dots_noalloc_f64(int nd, const double* a0, const double* b0, const double* a1, const double* b1,
                        const double* a2, const double* b2, const double* w, long long n, double* out, int o0,
                        int o1, int o2, hipStream_t s) {
  if (nd < 1 || nd > 3 || n <= 0 || !out || !a0 || !b0) return IG_EINVAL;
  if ((long long)red_blocks(n) * nd > NOALLOC_SCRATCH_DOUBLES) return IG_EINVAL;  // partials must fit
  double* work = nullptr;
  const int st = scratch_for_current_device(&work);
  if (st != IG_OK) return st;
  return dots_impl<double>(nd, a0, b0, a1, b1, a2, b2, w, n, work, out, o0, o1, o2, 0, s);
}

int ig_dots_f64(int nd, const double* a0, const double* b0, const double* a1, const double* b1,
                const double* a2, const double* b2, const double* w, long long n, double* work,
                double* out, int o0, int o1, int o2, hipStream_t s) {
  return dots_impl<double>(nd, a0, b0, a1, b1, a2, b2, w, n, work, out, o0, o1, o2, 0, s);
}

int ig_dots_f32(int nd, const float* a0, const float* b0, const float* a1, const float* b1,
                const float* a2, const float* b2, const float* w, long long n, double* work,
                double* out, int o0, int o1, int o2, hipStream_t s) {
  return dots_impl<float>(nd, a0, b0, a1, b1, a2, b2, w, n, work, out, o0, o1, o2, 0, s);
}

int ig_bicgstab_p_f64(long long n, const double* r, double* p, const double* v, const double* dinv,
                      double* phat, const double* sc, hipStream_t s) {
  return bicg_step_p<double>(n, r, p, v, dinv, phat, sc, s);
}
int ig_bicgstab_p_f32(long long n, const float* r, float* p, const float* v, const float* dinv,
                      float* phat, const double* sc, hipStream_t s) {
  return bicg_step_p<float>(n, r, p, v, dinv, phat, sc, s);
}
int ig_bicgstab_s_f64(long long n, const double* r, const double* v, const double* dinv, double* sv,
                      double* shat, double* sc, hipStream_t s) {
  return bicg_step_s<double>(n, r, v, dinv, sv, shat, sc, s);
}
int ig_bicgstab_s_f32(long long n, const float* r, const float* v, const float* dinv, float* sv,
                      float* shat, double* sc, hipStream_t s) {
  return bicg_step_s<float>(n, r, v, dinv, sv, shat, sc, s);
}
int ig_bicgstab_x_f64(long long n, double* x, const double* phat, const double* shat, const double* sv,
                      const double* tv, double* r, const double* rhat, double* sc, double* work,
                      hipStream_t s) {
  return bicg_step_x<double>(n, x, phat, shat, sv, tv, r, rhat, sc, work, s);
}
int ig_bicgstab_x_f32(long long n, float* x, const float* phat, const float* shat, const float* sv,
                      const float* tv, float* r, const float* rhat, double* sc, double* work,
                      hipStream_t s) {
  return bicg_step_x<float>(n, x, phat, shat, sv, tv, r, rhat, sc, work, s);
}

}  // extern "C"
