// Multigrid V-cycle kernels of the open-boundary Poisson preconditioner on monitor grids.
//
// The monitor Laplacian takes 2h-wide centred differences in flux form, so its diagonal
// (a = b) terms couple node i with i +- 2 e_a only: the grid splits into eight parity
// sub-lattices (i mod 2 in each axis) that only the small mixed (a != b) terms connect.
// The coarse spaces respect that split.  Level 1 aggregates the interior nodes of one
// parity inside a 4 x 4 x 4 block of the fine grid (fine node i belongs to parity
// p = i & 1 and aggregate S = (i - p) >> 2, per axis); deeper levels aggregate 2 x 2 x 2
// coarse points of the same parity.  Each coarse operator keeps the same-parity couplings
// of the Galerkin product P^T A P only, i.e. eight independent 7-point operators per level.
//
// Layouts: fine vectors are the grid in C order (x slowest).  A coarse vector of level
// dimensions (cx, cy, cz) is [p][Sx][Sy][Sz] with p = 4 px + 2 py + pz, m = 8 cx cy cz
// entries; its operator is coef[7][m] = {diagonal, -x, +x, -y, +y, -z, +z} couplings
// (couplings that leave the level's box are zero).  Aggregates that hold no interior node
// carry diagonal 1 and zero couplings.
//
// Kernels (T = float / double, every sum in T):
//   fine_restrict   bc[p][S] = sum over the 8 interior children i of (b - t)[i]   (t may be null)
//   fine_prolong    x[i] += ec[p(i)][S(i)] on interior nodes
//   smooth          x += dw (b - t)      (fine damped Jacobi; dw = omega / diag, 0 on faces)
//   mg7_jacobi      xo = xi + omega (b - A xi) / diag   (xi null: xo = omega b / diag)
//   mg7_restrict    bc[p][S] = sum over children 2S + d of (b - A x)  (x null: sum of b)
//   mg7_prolong     x[p][s] += ec[p][s >> 1]
// Probing (hierarchy set-up, FP64): node i of parity class cls = 4 px + 2 py + pz lies on
// sub-lattice point s = i >> 1 with colour (sx + 2 sy + 3 sz) mod 7; classes cls and
// 7 - cls share a probe (pair = min of the two), so probe q = 7 pair + colour.
//   probe_vector    e[i] = 1 on the interior nodes of probe q, else 0
//   probe_read      ent[k][i] = out[i] where probe q holds node i + o_k (o_k the k-th
//                   7-point offset in fine nodes: 0, -+2 e_x, -+2 e_y, -+2 e_z)
// The solve these kernels precondition replaces the reference's unpreconditioned
// interior BiCGSTAB (poisson/open_boundary.py:1017-1064) and the line-preconditioned CG of
// its JAX path (poisson/poisson_jax.py:411-579); the discrete solution is the same.
#include "common.h"

namespace {

__device__ __forceinline__ long long cidx(int p, int sx, int sy, int sz, int cx, int cy, int cz) {
  return (((long long)p * cx + sx) * cy + sy) * cz + sz;
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void fine_restrict(const T* __restrict__ b, const T* __restrict__ t,
                                                            int nx, int ny, int nz, T* __restrict__ bc, int cx,
                                                            int cy, int cz) {
  const long long m = 8LL * cx * cy * cz;
  for (long long q = (long long)blockIdx.x * IG_THREADS + threadIdx.x; q < m;
       q += (long long)gridDim.x * IG_THREADS) {
    const int sz = (int)(q % cz);
    const int sy = (int)((q / cz) % cy);
    const int sx = (int)((q / ((long long)cz * cy)) % cx);
    const int p = (int)(q / ((long long)cz * cy * cx));
    const int px = p >> 2, py = (p >> 1) & 1, pz = p & 1;
    T acc = T(0);
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) {
      const int ix = px + 4 * sx + 2 * dx;
      if (ix < 1 || ix > nx - 2) continue;
#pragma unroll
      for (int dy = 0; dy < 2; ++dy) {
        const int iy = py + 4 * sy + 2 * dy;
        if (iy < 1 || iy > ny - 2) continue;
#pragma unroll
        for (int dz = 0; dz < 2; ++dz) {
          const int iz = pz + 4 * sz + 2 * dz;
          if (iz < 1 || iz > nz - 2) continue;
          const long long i = ((long long)ix * ny + iy) * nz + iz;
          acc += t ? b[i] - t[i] : b[i];
        }
      }
    }
    bc[q] = acc;
  }
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void fine_prolong(T* __restrict__ x, const T* __restrict__ ec, int nx,
                                                           int ny, int nz, int cx, int cy, int cz) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * IG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * IG_THREADS) {
    const int iz = (int)(i % nz);
    const int iy = (int)((i / nz) % ny);
    const int ix = (int)(i / ((long long)nz * ny));
    if (ix < 1 || ix > nx - 2 || iy < 1 || iy > ny - 2 || iz < 1 || iz > nz - 2) continue;
    const int px = ix & 1, py = iy & 1, pz = iz & 1;
    x[i] += ec[cidx(4 * px + 2 * py + pz, (ix - px) >> 2, (iy - py) >> 2, (iz - pz) >> 2, cx, cy, cz)];
  }
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void smooth(T* __restrict__ x, const T* __restrict__ b,
                                                     const T* __restrict__ t, const T* __restrict__ dw,
                                                     long long n) {
  for (long long i = (long long)blockIdx.x * IG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * IG_THREADS)
    x[i] += dw[i] * (b[i] - t[i]);
}

// (A x)[q] of a 7-point level operator at coarse point (p, sx, sy, sz).
template <typename T>
__device__ __forceinline__ T apply7(const T* __restrict__ x, const T* __restrict__ coef, long long m, long long q,
                                    int sx, int sy, int sz, int cx, int cy, int cz) {
  const long long sX = (long long)cy * cz, sY = cz;
  T v = coef[q] * x[q];
  if (sx > 0) v += coef[m + q] * x[q - sX];
  if (sx < cx - 1) v += coef[2 * m + q] * x[q + sX];
  if (sy > 0) v += coef[3 * m + q] * x[q - sY];
  if (sy < cy - 1) v += coef[4 * m + q] * x[q + sY];
  if (sz > 0) v += coef[5 * m + q] * x[q - 1];
  if (sz < cz - 1) v += coef[6 * m + q] * x[q + 1];
  return v;
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void mg7_jacobi(T* __restrict__ xo, const T* __restrict__ xi,
                                                         const T* __restrict__ b, const T* __restrict__ coef,
                                                         T omega, int cx, int cy, int cz) {
  const long long m = 8LL * cx * cy * cz;
  for (long long q = (long long)blockIdx.x * IG_THREADS + threadIdx.x; q < m;
       q += (long long)gridDim.x * IG_THREADS) {
    if (!xi) {
      xo[q] = omega * b[q] / coef[q];
      continue;
    }
    const int sz = (int)(q % cz);
    const int sy = (int)((q / cz) % cy);
    const int sx = (int)((q / ((long long)cz * cy)) % cx);
    xo[q] = xi[q] + omega * (b[q] - apply7(xi, coef, m, q, sx, sy, sz, cx, cy, cz)) / coef[q];
  }
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void mg7_restrict(T* __restrict__ bc, const T* __restrict__ b,
                                                           const T* __restrict__ x, const T* __restrict__ coef,
                                                           int cx, int cy, int cz, int dx, int dy, int dz) {
  const long long md = 8LL * dx * dy * dz;
  const long long m = 8LL * cx * cy * cz;
  for (long long q = (long long)blockIdx.x * IG_THREADS + threadIdx.x; q < md;
       q += (long long)gridDim.x * IG_THREADS) {
    const int Sz = (int)(q % dz);
    const int Sy = (int)((q / dz) % dy);
    const int Sx = (int)((q / ((long long)dz * dy)) % dx);
    const int p = (int)(q / ((long long)dz * dy * dx));
    T acc = T(0);
    for (int ex = 0; ex < 2; ++ex) {
      const int sx = 2 * Sx + ex;
      if (sx >= cx) continue;
      for (int ey = 0; ey < 2; ++ey) {
        const int sy = 2 * Sy + ey;
        if (sy >= cy) continue;
        for (int ez = 0; ez < 2; ++ez) {
          const int sz = 2 * Sz + ez;
          if (sz >= cz) continue;
          const long long c = cidx(p, sx, sy, sz, cx, cy, cz);
          acc += x ? b[c] - apply7(x, coef, m, c, sx, sy, sz, cx, cy, cz) : b[c];
        }
      }
    }
    bc[q] = acc;
  }
}

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void mg7_prolong(T* __restrict__ x, const T* __restrict__ ec, int cx,
                                                          int cy, int cz, int dx, int dy, int dz) {
  const long long m = 8LL * cx * cy * cz;
  for (long long q = (long long)blockIdx.x * IG_THREADS + threadIdx.x; q < m;
       q += (long long)gridDim.x * IG_THREADS) {
    const int sz = (int)(q % cz);
    const int sy = (int)((q / cz) % cy);
    const int sx = (int)((q / ((long long)cz * cy)) % cx);
    const int p = (int)(q / ((long long)cz * cy * cx));
    x[q] += ec[cidx(p, sx >> 1, sy >> 1, sz >> 1, dx, dy, dz)];
  }
}

__device__ __forceinline__ bool probe_node(long long i, int nx, int ny, int nz, int& pair, int& col) {
  const int iz = (int)(i % nz);
  const int iy = (int)((i / nz) % ny);
  const int ix = (int)(i / ((long long)nz * ny));
  if (ix < 1 || ix > nx - 2 || iy < 1 || iy > ny - 2 || iz < 1 || iz > nz - 2) return false;
  const int cls = 4 * (ix & 1) + 2 * (iy & 1) + (iz & 1);
  pair = cls >= 4 ? 7 - cls : cls;
  col = ((ix >> 1) + 2 * (iy >> 1) + 3 * (iz >> 1)) % 7;
  return true;
}

__global__ __launch_bounds__(IG_THREADS) void probe_vector(double* __restrict__ e, int q, int nx, int ny, int nz) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * IG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * IG_THREADS) {
    int pair, col;
    e[i] = (probe_node(i, nx, ny, nz, pair, col) && 7 * pair + col == q) ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(IG_THREADS) void probe_read(double* __restrict__ ent, const double* __restrict__ out,
                                                         int q, int nx, int ny, int nz) {
  // colour shift of the 7-point offsets (o_x + 2 o_y + 3 o_z) / 2, kept in [0, 7)
  const int shift[7] = {0, 6, 1, 5, 2, 4, 3};
  const long long n = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * IG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * IG_THREADS) {
    int pair, col;
    if (!probe_node(i, nx, ny, nz, pair, col) || pair != q / 7) continue;
#pragma unroll
    for (int k = 0; k < 7; ++k)
      if ((col + shift[k]) % 7 == q % 7) ent[(long long)k * n + i] = out[i];
  }
}

inline unsigned grid_for(long long n) {
  long long g = (n + IG_THREADS - 1) / IG_THREADS;
  if (g > 8192) g = 8192;
  return (unsigned)(g < 1 ? 1 : g);
}

inline bool dims_ok(int a, int b, int c) { return a > 0 && b > 0 && c > 0; }
inline bool fine_level_ok(int nx, int ny, int nz, int cx, int cy, int cz) {
  return dims_ok(nx, ny, nz) && cx == (nx + 3) / 4 && cy == (ny + 3) / 4 && cz == (nz + 3) / 4;
}
inline bool next_level_ok(int cx, int cy, int cz, int dx, int dy, int dz) {
  return dims_ok(cx, cy, cz) && dx == (cx + 1) / 2 && dy == (cy + 1) / 2 && dz == (cz + 1) / 2;
}

template <typename T>
int run_fine_restrict(const T* b, const T* t, int nx, int ny, int nz, T* bc, int cx, int cy, int cz, hipStream_t s) {
  if (!b || !bc || !fine_level_ok(nx, ny, nz, cx, cy, cz)) return IG_EINVAL;
  hipLaunchKernelGGL(fine_restrict<T>, dim3(grid_for(8LL * cx * cy * cz)), dim3(IG_THREADS), 0, s, b, t, nx, ny, nz,
                     bc, cx, cy, cz);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int run_fine_prolong(T* x, const T* ec, int nx, int ny, int nz, int cx, int cy, int cz, hipStream_t s) {
  if (!x || !ec || !fine_level_ok(nx, ny, nz, cx, cy, cz)) return IG_EINVAL;
  hipLaunchKernelGGL(fine_prolong<T>, dim3(grid_for((long long)nx * ny * nz)), dim3(IG_THREADS), 0, s, x, ec, nx, ny,
                     nz, cx, cy, cz);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int run_smooth(T* x, const T* b, const T* t, const T* dw, long long n, hipStream_t s) {
  if (!x || !b || !t || !dw || n <= 0) return IG_EINVAL;
  hipLaunchKernelGGL(smooth<T>, dim3(grid_for(n)), dim3(IG_THREADS), 0, s, x, b, t, dw, n);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int run_jacobi(T* xo, const T* xi, const T* b, const T* coef, double omega, int cx, int cy, int cz, hipStream_t s) {
  if (!xo || !b || !coef || !dims_ok(cx, cy, cz) || xo == xi) return IG_EINVAL;
  hipLaunchKernelGGL(mg7_jacobi<T>, dim3(grid_for(8LL * cx * cy * cz)), dim3(IG_THREADS), 0, s, xo, xi, b, coef,
                     (T)omega, cx, cy, cz);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int run_restrict(T* bc, const T* b, const T* x, const T* coef, int cx, int cy, int cz, int dx, int dy, int dz,
                 hipStream_t s) {
  if (!bc || !b || (x && !coef) || !next_level_ok(cx, cy, cz, dx, dy, dz)) return IG_EINVAL;
  hipLaunchKernelGGL(mg7_restrict<T>, dim3(grid_for(8LL * dx * dy * dz)), dim3(IG_THREADS), 0, s, bc, b, x, coef, cx,
                     cy, cz, dx, dy, dz);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int run_prolong(T* x, const T* ec, int cx, int cy, int cz, int dx, int dy, int dz, hipStream_t s) {
  if (!x || !ec || !next_level_ok(cx, cy, cz, dx, dy, dz)) return IG_EINVAL;
  hipLaunchKernelGGL(mg7_prolong<T>, dim3(grid_for(8LL * cx * cy * cz)), dim3(IG_THREADS), 0, s, x, ec, cx, cy, cz,
                     dx, dy, dz);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int run_probe_vector(double* e, int q, int nx, int ny, int nz, hipStream_t s) {
  if (!e || q < 0 || q >= 28 || !dims_ok(nx, ny, nz)) return IG_EINVAL;
  hipLaunchKernelGGL(probe_vector, dim3(grid_for((long long)nx * ny * nz)), dim3(IG_THREADS), 0, s, e, q, nx, ny, nz);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int run_probe_read(double* ent, const double* out, int q, int nx, int ny, int nz, hipStream_t s) {
  if (!ent || !out || q < 0 || q >= 28 || !dims_ok(nx, ny, nz)) return IG_EINVAL;
  hipLaunchKernelGGL(probe_read, dim3(grid_for((long long)nx * ny * nz)), dim3(IG_THREADS), 0, s, ent, out, q, nx,
                     ny, nz);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

}  // namespace

extern "C" {

int ig_mg_probe_vector_f64(double* e, int q, int nx, int ny, int nz, hipStream_t s) {
  return run_probe_vector(e, q, nx, ny, nz, s);
}
int ig_mg_probe_read_f64(double* ent, const double* out, int q, int nx, int ny, int nz, hipStream_t s) {
  return run_probe_read(ent, out, q, nx, ny, nz, s);
}

int ig_mg_fine_restrict_f64(const double* b, const double* t, int nx, int ny, int nz, double* bc, int cx, int cy,
                            int cz, hipStream_t s) {
  return run_fine_restrict<double>(b, t, nx, ny, nz, bc, cx, cy, cz, s);
}
int ig_mg_fine_restrict_f32(const float* b, const float* t, int nx, int ny, int nz, float* bc, int cx, int cy,
                            int cz, hipStream_t s) {
  return run_fine_restrict<float>(b, t, nx, ny, nz, bc, cx, cy, cz, s);
}
int ig_mg_fine_prolong_f64(double* x, const double* ec, int nx, int ny, int nz, int cx, int cy, int cz,
                           hipStream_t s) {
  return run_fine_prolong<double>(x, ec, nx, ny, nz, cx, cy, cz, s);
}
int ig_mg_fine_prolong_f32(float* x, const float* ec, int nx, int ny, int nz, int cx, int cy, int cz,
                           hipStream_t s) {
  return run_fine_prolong<float>(x, ec, nx, ny, nz, cx, cy, cz, s);
}
int ig_mg_smooth_f64(double* x, const double* b, const double* t, const double* dw, long long n, hipStream_t s) {
  return run_smooth<double>(x, b, t, dw, n, s);
}
int ig_mg_smooth_f32(float* x, const float* b, const float* t, const float* dw, long long n, hipStream_t s) {
  return run_smooth<float>(x, b, t, dw, n, s);
}
int ig_mg7_jacobi_f64(double* xo, const double* xi, const double* b, const double* coef, double omega, int cx,
                      int cy, int cz, hipStream_t s) {
  return run_jacobi<double>(xo, xi, b, coef, omega, cx, cy, cz, s);
}
int ig_mg7_jacobi_f32(float* xo, const float* xi, const float* b, const float* coef, double omega, int cx, int cy,
                      int cz, hipStream_t s) {
  return run_jacobi<float>(xo, xi, b, coef, omega, cx, cy, cz, s);
}
int ig_mg7_restrict_f64(double* bc, const double* b, const double* x, const double* coef, int cx, int cy, int cz,
                        int dx, int dy, int dz, hipStream_t s) {
  return run_restrict<double>(bc, b, x, coef, cx, cy, cz, dx, dy, dz, s);
}
int ig_mg7_restrict_f32(float* bc, const float* b, const float* x, const float* coef, int cx, int cy, int cz, int dx,
                        int dy, int dz, hipStream_t s) {
  return run_restrict<float>(bc, b, x, coef, cx, cy, cz, dx, dy, dz, s);
}
int ig_mg7_prolong_f64(double* x, const double* ec, int cx, int cy, int cz, int dx, int dy, int dz, hipStream_t s) {
  return run_prolong<double>(x, ec, cx, cy, cz, dx, dy, dz, s);
}
int ig_mg7_prolong_f32(float* x, const float* ec, int cx, int cy, int cz, int dx, int dy, int dz, hipStream_t s) {
  return run_prolong<float>(x, ec, cx, cy, cz, dx, dy, dz, s);
}

}  // extern "C"
