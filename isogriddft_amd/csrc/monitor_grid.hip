// Monitor-grid generator: atom-centred monitor field and the weighted harmonic map
// (grid/monitor_geometry.py:51-136, :268-396, :481-577), plus the metric of a mapped
// grid (:324-373).
//
// Coordinates are SoA (x | y | z, N each, C order with z fastest).  The Jacobi sweep
// evaluates the reference's expressions in the reference's order with contraction off,
// so a sweep reproduces the NumPy iterate; the sweep's largest coordinate change and the
// backtracking's smallest Jacobian are per-workgroup partials reduced in a fixed order.
#include "common.h"

#pragma clang fp contract(off)

namespace {

// One atom: position, then kinetic (weight, radius, exponent), local (weight, radius,
// exponent), projector (weight, radius, exponent) and the cap.
struct MonAtom {
  double x, y, z;
  double kw, kr, ke, lw, lr, le, pw, pr, pe, cap;
};

__global__ __launch_bounds__(IG_THREADS) void monitor_field_kernel(const double* __restrict__ xyz, long long N,
                                                                   const MonAtom* __restrict__ atoms, int natoms,
                                                                   double* __restrict__ out) {
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    double v = 1.0;
    for (int a = 0; a < natoms; ++a) {
      const MonAtom& p = atoms[a];
      const double dx = xyz[t] - p.x, dy = xyz[N + t] - p.y, dz = xyz[2 * N + t] - p.z;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      const double kin = p.kw * exp(-pow(r / p.kr, p.ke));
      const double loc = p.lw / (1.0 + pow(r / p.lr, p.le));
      const double prj = p.pw > 0.0 ? p.pw * exp(-pow(r / p.pr, p.pe)) : 0.0;
      const double raw = kin + loc + prj;
      v += p.cap * (1.0 - exp(-raw / p.cap));
    }
    out[t] = v;
  }
}

// 7-point interior average blended with weight s, then max(., 1) everywhere.
__global__ __launch_bounds__(IG_THREADS) void monitor_smooth_kernel(const double* __restrict__ m, int nx, int ny,
                                                                    int nz, double s, double* __restrict__ out) {
  const long long N = (long long)nx * ny * nz;
  const long long sx = (long long)ny * nz;
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    const int k = (int)(t % nz), j = (int)((t / nz) % ny), i = (int)(t / sx);
    double v = m[t];
    if (s > 0.0 && i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1) {
      const double avg = (m[t] + m[t - sx] + m[t + sx] + m[t - nz] + m[t + nz] + m[t - 1] + m[t + 1]) / 7.0;
      v = (1.0 - s) * m[t] + s * avg;
    }
    out[t] = s > 0.0 ? fmax(v, 1.0) : v;
  }
}

// One damped Jacobi sweep of div(M grad X) = 0 on the interior; faces are copied.
// partial[block] = max |X_new - X| over the block's interior points and components.
__global__ __launch_bounds__(IG_THREADS) void harmonic_sweep_kernel(const double* __restrict__ mon,
                                                                    const double* __restrict__ c,
                                                                    double* __restrict__ out, int nx, int ny, int nz,
                                                                    double dx, double dy, double dz, double relax,
                                                                    double* __restrict__ partial) {
  __shared__ double sm[IG_THREADS / 64];
  const long long N = (long long)nx * ny * nz;
  const long long sx = (long long)ny * nz;
  double mx = 0.0;
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    const int k = (int)(t % nz), j = (int)((t / nz) % ny), i = (int)(t / sx);
    if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1) {
      out[t] = c[t];
      out[N + t] = c[N + t];
      out[2 * N + t] = c[2 * N + t];
      continue;
    }
    const double m0 = mon[t];
    const double axm = 0.5 * (m0 + mon[t - sx]) / (dx * dx);
    const double axp = 0.5 * (m0 + mon[t + sx]) / (dx * dx);
    const double aym = 0.5 * (m0 + mon[t - nz]) / (dy * dy);
    const double ayp = 0.5 * (m0 + mon[t + nz]) / (dy * dy);
    const double azm = 0.5 * (m0 + mon[t - 1]) / (dz * dz);
    const double azp = 0.5 * (m0 + mon[t + 1]) / (dz * dz);
    const double diag = axm + axp + aym + ayp + azm + azp;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double* f = c + q * N;
      const double cand =
          (axm * f[t - sx] + axp * f[t + sx] + aym * f[t - nz] + ayp * f[t + nz] + azm * f[t - 1] + azp * f[t + 1]) /
          diag;
      const double u = (1.0 - relax) * f[t] + relax * cand;
      out[q * N + t] = u;
      mx = fmax(mx, fabs(u - f[t]));
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = sm[0];
    for (int w = 1; w < IG_THREADS / 64; ++w) v = fmax(v, sm[w]);
    partial[blockIdx.x] = v;
  }
}

// d f / d xi along one axis (stride s, n points, spacing h): centred inside, the
// second-order one-sided formulas of np.gradient(edge_order=2) at the two ends.
__device__ __forceinline__ double grad_axis(const double* f, long long t, int i, int n, long long s, double h) {
  if (i == 0) return (-1.5 / h) * f[t] + (2.0 / h) * f[t + s] + (-0.5 / h) * f[t + 2 * s];
  if (i == n - 1) return (0.5 / h) * f[t - 2 * s] + (-2.0 / h) * f[t - s] + (1.5 / h) * f[t];
  return (f[t + s] - f[t - s]) / (2.0 * h);
}

// Covariant basis B[a][i] = d x_i / d xi_a of the coordinates c at point t.
__device__ __forceinline__ void basis_at(const double* c, long long N, long long t, int i, int j, int k, int nx,
                                         int ny, int nz, double hx, double hy, double hz, double B[3][3]) {
  const long long sx = (long long)ny * nz;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    B[0][q] = grad_axis(c + q * N, t, i, nx, sx, hx);
    B[1][q] = grad_axis(c + q * N, t, j, ny, nz, hy);
    B[2][q] = grad_axis(c + q * N, t, k, nz, 1, hz);
  }
}

__device__ __forceinline__ double det3(const double B[3][3]) {
  return B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) - B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
         B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
}

// partial[block] = min det B over the block's points of c.
__global__ __launch_bounds__(IG_THREADS) void jacobian_min_kernel(const double* __restrict__ c, int nx, int ny, int nz,
                                                                  double hx, double hy, double hz,
                                                                  double* __restrict__ partial) {
  __shared__ double sm[IG_THREADS / 64];
  const long long N = (long long)nx * ny * nz;
  double mn = __builtin_huge_val();
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    const int k = (int)(t % nz), j = (int)((t / nz) % ny), i = (int)(t / ((long long)ny * nz));
    double B[3][3];
    basis_at(c, N, t, i, j, k, nx, ny, nz, hx, hy, hz, B);
    mn = fmin(mn, det3(B));
  }
  for (int off = 32; off > 0; off >>= 1) mn = fmin(mn, __shfl_down(mn, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = sm[0];
    for (int w = 1; w < IG_THREADS / 64; ++w) v = fmin(v, sm[w]);
    partial[blockIdx.x] = v;
  }
}

// J = det B, g^-1 = (B B^T)^-1 (9 entries per point, row-major), cell volume J hx hy hz.
__global__ __launch_bounds__(IG_THREADS) void mapping_metric_kernel(const double* __restrict__ c, int nx, int ny,
                                                                    int nz, double hx, double hy, double hz,
                                                                    double* __restrict__ jac, double* __restrict__ ginv,
                                                                    double* __restrict__ vol) {
  const long long N = (long long)nx * ny * nz;
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    const int k = (int)(t % nz), j = (int)((t / nz) % ny), i = (int)(t / ((long long)ny * nz));
    double B[3][3];
    basis_at(c, N, t, i, j, k, nx, ny, nz, hx, hy, hz, B);
    double g[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) g[a][b] = B[a][0] * B[b][0] + B[a][1] * B[b][1] + B[a][2] * B[b][2];
    const double J = det3(B);
    const double dg = det3(g);
    double* gi = ginv + 9 * t;
    gi[0] = (g[1][1] * g[2][2] - g[1][2] * g[2][1]) / dg;
    gi[1] = (g[0][2] * g[2][1] - g[0][1] * g[2][2]) / dg;
    gi[2] = (g[0][1] * g[1][2] - g[0][2] * g[1][1]) / dg;
    gi[3] = (g[1][2] * g[2][0] - g[1][0] * g[2][2]) / dg;
    gi[4] = (g[0][0] * g[2][2] - g[0][2] * g[2][0]) / dg;
    gi[5] = (g[0][2] * g[1][0] - g[0][0] * g[1][2]) / dg;
    gi[6] = (g[1][0] * g[2][1] - g[1][1] * g[2][0]) / dg;
    gi[7] = (g[0][1] * g[2][0] - g[0][0] * g[2][1]) / dg;
    gi[8] = (g[0][0] * g[1][1] - g[0][1] * g[1][0]) / dg;
    jac[t] = J;
    vol[t] = J * hx * hy * hz;
  }
}

__global__ void reduce_extreme_kernel(const double* __restrict__ partial, int n, int take_max,
                                      double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double v = partial[0];
  for (int i = 1; i < n; ++i) v = take_max ? fmax(v, partial[i]) : fmin(v, partial[i]);
  out[0] = v;
}

inline int blocks_for(long long n) {
  long long b = (n + IG_THREADS - 1) / IG_THREADS;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

extern "C" {

int ig_grid_work_size(void) { return 2048 + 8; }

int ig_monitor_field_f64(const double* xyz, long long N, const double* atoms, int natoms, double* out,
                         hipStream_t s) {
  if (!xyz || !out || N <= 0 || natoms < 0 || (natoms > 0 && !atoms)) return IG_EINVAL;
  hipLaunchKernelGGL(monitor_field_kernel, dim3(blocks_for(N)), dim3(IG_THREADS), 0, s, xyz, N,
                     (const MonAtom*)atoms, natoms, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_monitor_smooth_f64(const double* mon, int nx, int ny, int nz, double smoothing, double* out, hipStream_t s) {
  if (!mon || !out || nx < 1 || ny < 1 || nz < 1 || mon == out) return IG_EINVAL;
  hipLaunchKernelGGL(monitor_smooth_kernel, dim3(blocks_for((long long)nx * ny * nz)), dim3(IG_THREADS), 0, s, mon,
                     nx, ny, nz, smoothing, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_harmonic_sweep_f64(const double* mon, const double* coords, double* out, int nx, int ny, int nz, double dx,
                          double dy, double dz, double relax, double* work, double* max_change, hipStream_t s) {
  if (!mon || !coords || !out || !work || !max_change || coords == out || nx < 3 || ny < 3 || nz < 3)
    return IG_EINVAL;
  const int nb = blocks_for((long long)nx * ny * nz);
  hipLaunchKernelGGL(harmonic_sweep_kernel, dim3(nb), dim3(IG_THREADS), 0, s, mon, coords, out, nx, ny, nz, dx, dy,
                     dz, relax, work);
  hipLaunchKernelGGL(reduce_extreme_kernel, dim3(1), dim3(64), 0, s, work, nb, 1, max_change);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_mapping_min_jacobian_f64(const double* coords, int nx, int ny, int nz, double hx, double hy, double hz,
                                double* work, double* min_jac, hipStream_t s) {
  if (!coords || !work || !min_jac || nx < 3 || ny < 3 || nz < 3) return IG_EINVAL;
  const int nb = blocks_for((long long)nx * ny * nz);
  hipLaunchKernelGGL(jacobian_min_kernel, dim3(nb), dim3(IG_THREADS), 0, s, coords, nx, ny, nz, hx, hy, hz, work);
  hipLaunchKernelGGL(reduce_extreme_kernel, dim3(1), dim3(64), 0, s, work, nb, 0, min_jac);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_mapping_metric_f64(const double* coords, int nx, int ny, int nz, double hx, double hy, double hz,
                          double* jacobian, double* inverse_metric, double* cell_volumes, hipStream_t s) {
  if (!coords || !jacobian || !inverse_metric || !cell_volumes || nx < 3 || ny < 3 || nz < 3) return IG_EINVAL;
  hipLaunchKernelGGL(mapping_metric_kernel, dim3(blocks_for((long long)nx * ny * nz)), dim3(IG_THREADS), 0, s,
                     coords, nx, ny, nz, hx, hy, hz, jacobian, inverse_metric, cell_volumes);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

}  // extern "C"
