// Discrete Green-identity terms of the monitor-grid kinetic operator
// (audit/h2_monitor_grid_kinetic_green_identity_audit.py:143-218, :295-330).
//
// For a real field f with contravariant flux F^a = c^{ab} d_b f (c = J g^{-1}):
//   K_op   = <f, T f>_W                        T f is supplied by the caller (any closure)
//   K_grad = 1/2 sum_t w_t d_a f g^{ab} d_b f  np.gradient(edge_order=2) derivatives
//   K_bdry = sum over face nodes that lie on no edge of s_face 1/2 f F^a dS, with
//            s = +1 on the minimum face and -1 on the maximum face of axis a, dS the
//            product of the other two logical spacings and F^a built with the chosen
//            closure's derivatives (zero-ghost centred or np.gradient edge_order=2)
//   N      = <f, f>_W
// The continuous identity is K_op = K_grad + K_bdry; the discrete closure mismatch is
// K_op - K_grad - K_bdry.  One pass writes per-block partials of the four sums and a
// single block adds them in a fixed order, so the result is bit-reproducible.
#include "common.h"

namespace {

constexpr int GI_MAX_BLOCKS = 1024;

// d f / d xi along one axis at index i of n (stride s); closure 0 = zero-ghost centred,
// 1 = np.gradient edge_order 2 on a uniform axis.
__device__ __forceinline__ double dfield(const double* __restrict__ f, long long t, int i, int n, long long s,
                                         double h, int closure) {
  if (i > 0 && i < n - 1) return (f[t + s] - f[t - s]) / (2.0 * h);
  if (closure == 0) return i == 0 ? (f[t + s] - 0.0) / (2.0 * h) : (0.0 - f[t - s]) / (2.0 * h);
  if (i == 0) return (-1.5 / h) * f[t] + (2.0 / h) * f[t + s] + (-0.5 / h) * f[t + 2 * s];
  return (0.5 / h) * f[t - 2 * s] + (-2.0 / h) * f[t - s] + (1.5 / h) * f[t];
}

__global__ __launch_bounds__(IG_THREADS) void green_terms_kernel(
    const double* __restrict__ f, const double* __restrict__ tf, const double* __restrict__ coef,
    const double* __restrict__ invj, const double* __restrict__ w, int nx, int ny, int nz, double hx, double hy,
    double hz, int closure, double* __restrict__ partial) {
  __shared__ double sm[4 * 4];
  const long long N = (long long)nx * ny * nz;
  const long long sx = (long long)ny * nz;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    const int k = (int)(t % nz), j = (int)((t / nz) % ny), i = (int)(t / sx);
    const double ft = f[t], wt = w[t];
    const double cxx = coef[t], cyy = coef[N + t], czz = coef[2 * N + t];
    const double cxy = coef[3 * N + t], cxz = coef[4 * N + t], cyz = coef[5 * N + t];
    const double gx = dfield(f, t, i, nx, sx, hx, 1);
    const double gy = dfield(f, t, j, ny, nz, hy, 1);
    const double gz = dfield(f, t, k, nz, 1, hz, 1);
    const double q = cxx * gx * gx + cyy * gy * gy + czz * gz * gz +
                     2.0 * (cxy * gx * gy + cxz * gx * gz + cyz * gy * gz);
    if (tf) acc[0] += wt * ft * tf[t];
    acc[1] += wt * (0.5 * q * invj[t]);
    acc[3] += wt * ft * ft;
    // Face nodes off every edge: exactly one logical index sits on its boundary.
    const bool bi = (i == 0 || i == nx - 1), bj = (j == 0 || j == ny - 1), bk = (k == 0 || k == nz - 1);
    if ((int)bi + (int)bj + (int)bk == 1) {
      const double dx = dfield(f, t, i, nx, sx, hx, closure);
      const double dy = dfield(f, t, j, ny, nz, hy, closure);
      const double dz = dfield(f, t, k, nz, 1, hz, closure);
      double flux, area, sgn;
      if (bi) {
        flux = cxx * dx + cxy * dy + cxz * dz;
        area = hy * hz;
        sgn = i == 0 ? 0.5 : -0.5;
      } else if (bj) {
        flux = cxy * dx + cyy * dy + cyz * dz;
        area = hx * hz;
        sgn = j == 0 ? 0.5 : -0.5;
      } else {
        flux = cxz * dx + cyz * dy + czz * dz;
        area = hx * hy;
        sgn = k == 0 ? 0.5 : -0.5;
      }
      acc[2] += sgn * ft * flux * area;
    }
  }
  ig_block_sum<4>(acc, sm);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) partial[(long long)blockIdx.x * 4 + v] = acc[v];
  }
}

__global__ void green_finish(const double* __restrict__ partial, int nparts, double* __restrict__ out) {
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += partial[(long long)p * 4 + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

}  // namespace

extern "C" {

int ig_green_identity_work_size(void) { return 4 * GI_MAX_BLOCKS; }

int ig_green_identity_f64(const double* f, const double* tf, const double* coef, const double* invj,
                          const double* w, int nx, int ny, int nz, double hx, double hy, double hz, int closure,
                          double* work, double* out, hipStream_t s) {
  if (!f || !coef || !invj || !w || !work || !out) return IG_EINVAL;
  if (nx < 3 || ny < 3 || nz < 3 || !(hx > 0.0) || !(hy > 0.0) || !(hz > 0.0)) return IG_EINVAL;
  if (closure != 0 && closure != 1) return IG_EINVAL;
  const long long N = (long long)nx * ny * nz;
  long long nb = (N + IG_THREADS * 4 - 1) / (IG_THREADS * 4);
  if (nb > GI_MAX_BLOCKS) nb = GI_MAX_BLOCKS;
  hipLaunchKernelGGL(green_terms_kernel, dim3((unsigned)nb), dim3(IG_THREADS), 0, s, f, tf, coef, invj, w, nx, ny,
                     nz, hx, hy, hz, closure, work);
  IG_CHECK_LAUNCH();
  hipLaunchKernelGGL(green_finish, dim3(1), dim3(64), 0, s, work, (int)nb, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

}  // extern "C"
