// Curvilinear kinetic / Laplacian stencils for the monitor grid and the legacy grid.
//
// Monitor operator (ops/kinetic.py:212-358 of the reference):
//     Lap psi = (1/J) sum_a D_a ( sum_b c^{ab} D_b psi ),   c^{ab} = J g^{ab},
// D the centred first difference on the uniform logical axis with either a zero
// ghost ("trial_fix", CL = 0) or NumPy's edge_order=2 one-sided ends ("production",
// CL = 1).  TR = 1 gives the exact Euclidean transpose
//     Lap^T x = sum_b D_b^T ( sum_a c^{ab} D_a^T (J^-1 x) ).
// One fused pass: a workgroup owns a TY x TZ tile of the (y, z) plane and marches
// along x; a six-plane LDS ring holds the input planes with a y/z halo, fluxes of
// the current plane are staged in LDS, the x fluxes stay in registers.  The output
// is   out = scale * Lap psi + V * psi   (V optional), optionally zeroed on the
// boundary nodes (the Poisson interior restriction).
#include "common.h"

namespace {

template <typename T>
struct MonParams {
  const T* in;
  T* out;
  const T* coef;  // 6*N: cxx cyy czz cxy cxz cyz
  const T* invj;  // N
  const T* veff;  // N or null
  long long ld;   // elements between consecutive orbitals (in and out)
  int nblk, nx, ny, nz;
  T ih2x, ih2y, ih2z;  // 1/(2h)
  T scale;
  int interior_only;
  int nty, ntz, nxc, xc, units;
};

// Coefficient (times 2h) of row `row` at column offset `off` of the 1D derivative.
template <int CL>
__device__ __forceinline__ int rowcoef(int row, int off, int n) {
  if (CL == 1) {
    if (row == 0) return off == 0 ? -3 : (off == 1 ? 4 : (off == 2 ? -1 : 0));
    if (row == n - 1) return off == 0 ? 3 : (off == -1 ? -4 : (off == -2 ? 1 : 0));
  }
  return off == 1 ? 1 : (off == -1 ? -1 : 0);
}

// Forward derivative row i (un-scaled); f(d) returns the value at offset d with
// zero for ghost positions.
template <int CL, typename T, typename F>
__device__ __forceinline__ T dfwd(int i, int n, F f) {
  if (CL == 0 || (i > 0 && i < n - 1)) return f(1) - f(-1);
  if (i == 0) return T(-3) * f(0) + T(4) * f(1) - f(2);
  return T(3) * f(0) - T(4) * f(-1) + f(-2);
}

// Transposed derivative row i (un-scaled): sum_d D[i+d][i] f(d).
template <int CL, typename T, typename F>
__device__ __forceinline__ T dtr(int i, int n, F f) {
  if (CL == 0 || (i >= 3 && i <= n - 4)) return f(-1) - f(1);
  T s = T(0);
#pragma unroll
  for (int d = -2; d <= 2; ++d) {
    const int row = i + d;
    if (row < 0 || row >= n) continue;
    const int c = rowcoef<CL>(row, -d, n);
    if (c != 0) s += T(c) * f(d);
  }
  return s;
}

template <int CL, bool TR, typename T, typename F>
__device__ __forceinline__ T dop(int i, int n, F f) {
  if (TR) return dtr<CL, T>(i, n, f);
  return dfwd<CL, T>(i, n, f);
}

template <typename T, int TY, int TZ, int CL, bool TR>
__global__ __launch_bounds__(IG_THREADS) void monitor_stencil_kernel(MonParams<T> p) {
  constexpr int HF = TR ? 2 : 1;  // flux halo
  constexpr int HP = 2 * HF;      // input halo
  constexpr int PY = TY + 2 * HP;
  constexpr int PZ = TZ + 2 * HP;
  constexpr int PLANE = PY * PZ;
  constexpr int NSLOT = 6;
  constexpr int RG = IG_THREADS / TZ;  // row groups
  constexpr int RPT = TY / RG;         // rows per thread
  static_assert(RPT * RG == TY, "tile rows must divide");

  __shared__ T s_in[NSLOT * PLANE];
  __shared__ T s_fy[(TY + 2 * HF) * TZ];
  __shared__ T s_fz[TY * (TZ + 2 * HF)];

  // XCD-aware block -> (work unit, orbital): consecutive blocks on one XCD share a
  // work unit so the geometry stream is reused from that XCD's L2.
  const int b = blockIdx.x;
  const int xcd = b & 7;
  const int r = b >> 3;
  const int orb = r % p.nblk;
  const int u = (r / p.nblk) * 8 + xcd;
  if (u >= p.units) return;
  const int xci = u % p.nxc;
  const int tzi = (u / p.nxc) % p.ntz;
  const int tyi = u / (p.nxc * p.ntz);

  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const long long nyz = (long long)ny * nz;
  const int j0 = min(tyi * TY, max(ny - TY, 0));
  const int k0 = min(tzi * TZ, max(nz - TZ, 0));
  const int x0 = xci * p.xc;
  const int x1 = min(x0 + p.xc, nx);

  const T* __restrict__ in = p.in + (long long)orb * p.ld;
  T* __restrict__ out = p.out + (long long)orb * p.ld;
  const long long N = (long long)nx * nyz;
  const T* __restrict__ cxx = p.coef;
  const T* __restrict__ cyy = p.coef + N;
  const T* __restrict__ czz = p.coef + 2 * N;
  const T* __restrict__ cxy = p.coef + 3 * N;
  const T* __restrict__ cxz = p.coef + 4 * N;
  const T* __restrict__ cyz = p.coef + 5 * N;

  const int tid = threadIdx.x;
  const int tz = tid % TZ;
  const int tyb = tid / TZ;

  auto load_plane = [&](int q) {
    T* dst = s_in + ((q % NSLOT + NSLOT) % NSLOT) * PLANE;
    const bool qin = (q >= 0 && q < nx);
    for (int idx = tid; idx < PLANE; idx += IG_THREADS) {
      const int ly = idx / PZ;
      const int lz = idx - ly * PZ;
      const int j = j0 + ly - HP;
      const int k = k0 + lz - HP;
      T v = T(0);
      if (qin && j >= 0 && j < ny && k >= 0 && k < nz) {
        const long long g = (long long)q * nyz + (long long)j * nz + k;
        v = in[g];
        if (TR) v *= p.invj[g];
      }
      dst[idx] = v;
    }
  };

  // Ring slot offsets of planes s-2 .. s+2 (refreshed every step).
  int sbase[5];
  // Gradient (scaled) at global (s, j, k) / local (ly, lz) from the ring.
  auto grad = [&](int s, int j, int k, int ly, int lz, T& gx, T& gy, T& gz) {
    const T* pl = s_in + sbase[2];
    const int off = ly * PZ + lz;
    gx = p.ih2x * dop<CL, TR, T>(s, nx, [&](int d) { return s_in[sbase[2 + d] + off]; });
    gy = p.ih2y * dop<CL, TR, T>(j, ny, [&](int d) { return pl[off + d * PZ]; });
    gz = p.ih2z * dop<CL, TR, T>(k, nz, [&](int d) { return pl[off + d]; });
  };

  T fx[5][RPT];
  T dyz[3][RPT];
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) fx[a][rr] = T(0);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) dyz[a][rr] = T(0);

  const int sb = x0 - 2;
  const int se = x1 + 2;
  for (int q = sb - 2; q < sb + 2; ++q) load_plane(q);

  for (int s = sb; s < se; ++s) {
    // Phase A: prefetch plane s+2 into the slot of plane s-4.
    load_plane(s + 2);
#pragma unroll
    for (int d = 0; d < 5; ++d) sbase[d] = (((s - 2 + d) % NSLOT + NSLOT) % NSLOT) * PLANE;
    __syncthreads();

    // Phase B: fluxes of plane s.
    const bool sin_ = (s >= 0 && s < nx);
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const int ty = tyb + rr * RG;
      const int j = j0 + ty;
      const int k = k0 + tz;
      T Fx = T(0), Fy = T(0), Fz = T(0);
      if (sin_ && j < ny && k < nz) {
        T gx, gy, gz;
        grad(s, j, k, ty + HP, tz + HP, gx, gy, gz);
        const long long g = (long long)s * nyz + (long long)j * nz + k;
        const T a_xx = cxx[g], a_yy = cyy[g], a_zz = czz[g];
        const T a_xy = cxy[g], a_xz = cxz[g], a_yz = cyz[g];
        if (TR) {
          // F_b = sum_a c^{ab} G_a with G = D^T u; c symmetric so same form.
          Fx = a_xx * gx + a_xy * gy + a_xz * gz;
          Fy = a_xy * gx + a_yy * gy + a_yz * gz;
          Fz = a_xz * gx + a_yz * gy + a_zz * gz;
        } else {
          Fx = a_xx * gx + a_xy * gy + a_xz * gz;
          Fy = a_xy * gx + a_yy * gy + a_yz * gz;
          Fz = a_xz * gx + a_yz * gy + a_zz * gz;
        }
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) fx[a][rr] = fx[a + 1][rr];
      fx[4][rr] = Fx;
      s_fy[(ty + HF) * TZ + tz] = Fy;
      s_fz[ty * (TZ + 2 * HF) + tz + HF] = Fz;
    }
    // y-halo rows of F_y and z-halo columns of F_z.
    {
      constexpr int NHY = 2 * HF * TZ;
      constexpr int NHZ = 2 * HF * TY;
      for (int idx = tid; idx < NHY + NHZ; idx += IG_THREADS) {
        if (idx < NHY) {
          const int hr = idx / TZ;           // 0 .. 2HF-1
          const int cz = idx - hr * TZ;
          const int fr = hr < HF ? hr : TY + hr;  // F row index in s_fy
          const int j = j0 + fr - HF;
          const int k = k0 + cz;
          T F = T(0);
          if (sin_ && j >= 0 && j < ny && k < nz) {
            T gx, gy, gz;
            grad(s, j, k, fr - HF + HP, cz + HP, gx, gy, gz);
            const long long g = (long long)s * nyz + (long long)j * nz + k;
            F = cxy[g] * gx + cyy[g] * gy + cyz[g] * gz;
          }
          s_fy[fr * TZ + cz] = F;
        } else {
          const int i2 = idx - NHY;
          const int hc = i2 / TY;
          const int ry = i2 - hc * TY;
          const int fc = hc < HF ? hc : TZ + hc;  // F column index in s_fz
          const int j = j0 + ry;
          const int k = k0 + fc - HF;
          T F = T(0);
          if (sin_ && j < ny && k >= 0 && k < nz) {
            T gx, gy, gz;
            grad(s, j, k, ry + HP, fc - HF + HP, gx, gy, gz);
            const long long g = (long long)s * nyz + (long long)j * nz + k;
            F = cxz[g] * gx + cyz[g] * gy + czz[g] * gz;
          }
          s_fz[ry * (TZ + 2 * HF) + fc] = F;
        }
      }
    }
    __syncthreads();

    // Phase C: y/z divergence of plane s, then output plane o = s - 2.
    const int o = s - 2;
    const bool emit = (o >= x0 && o < x1);
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const int ty = tyb + rr * RG;
      const int j = j0 + ty;
      const int k = k0 + tz;
      T d = T(0);
      if (sin_ && j < ny && k < nz) {
        const int fy0 = (ty + HF) * TZ + tz;
        const int fz0 = ty * (TZ + 2 * HF) + tz + HF;
        d = p.ih2y * dop<CL, TR, T>(j, ny, [&](int dd) { return s_fy[fy0 + dd * TZ]; }) +
            p.ih2z * dop<CL, TR, T>(k, nz, [&](int dd) { return s_fz[fz0 + dd]; });
      }
      dyz[2][rr] = dyz[1][rr];
      dyz[1][rr] = dyz[0][rr];
      dyz[0][rr] = d;
      if (emit && j < ny && k < nz) {
        const T dx = p.ih2x * dop<CL, TR, T>(o, nx, [&](int dd) { return fx[2 + dd][rr]; });
        const long long g = (long long)o * nyz + (long long)j * nz + k;
        const T lap = dx + dyz[2][rr];
        T v;
        if (TR) {
          const T xin = in[g];
          v = p.scale * lap;
          if (p.veff) v += p.veff[g] * xin;
        } else {
          const T psi = s_in[(o % NSLOT) * PLANE + (ty + HP) * PZ + tz + HP];
          v = p.scale * (lap * p.invj[g]);
          if (p.veff) v += p.veff[g] * psi;
        }
        if (p.interior_only &&
            (o == 0 || o == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1))
          v = T(0);
        out[g] = v;
      }
    }
  }
}

template <typename T, int TY, int TZ, int CL, bool TR>
int launch_monitor_t(const MonParams<T>& base, hipStream_t stream) {
  MonParams<T> p = base;
  p.nty = (p.ny + TY - 1) / TY;
  p.ntz = (p.nz + TZ - 1) / TZ;
  if (p.xc <= 0 || p.xc > p.nx) p.xc = p.nx;
  p.nxc = (p.nx + p.xc - 1) / p.xc;
  p.units = p.nty * p.ntz * p.nxc;
  const long long blocks = 8LL * p.nblk * ((p.units + 7) / 8);
  if (blocks <= 0 || blocks > 0x7fffffffLL) return IG_EINVAL;
  hipLaunchKernelGGL((monitor_stencil_kernel<T, TY, TZ, CL, TR>), dim3((unsigned)blocks),
                     dim3(IG_THREADS), 0, stream, p);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int launch_monitor(const MonParams<T>& p, int closure, int transpose, int tile, hipStream_t s) {
  // tile: 0 = auto; otherwise TZ (32 or 64)
  int tz = tile;
  if (tz == 0) tz = (p.nz % 64 == 0 || p.nz < 48) ? 64 : 32;
  if (sizeof(T) == 8) {
    if (transpose) tz = 32;  // keeps the transposed FP64 tile under 64 KB of LDS
    if (tz == 64) {
      if (transpose) return closure ? launch_monitor_t<T, 8, 64, 1, true>(p, s) : launch_monitor_t<T, 8, 64, 0, true>(p, s);
      return closure ? launch_monitor_t<T, 8, 64, 1, false>(p, s) : launch_monitor_t<T, 8, 64, 0, false>(p, s);
    }
    if (transpose) return closure ? launch_monitor_t<T, 8, 32, 1, true>(p, s) : launch_monitor_t<T, 8, 32, 0, true>(p, s);
    return closure ? launch_monitor_t<T, 8, 32, 1, false>(p, s) : launch_monitor_t<T, 8, 32, 0, false>(p, s);
  }
  if (tz == 64) {
    if (transpose) return closure ? launch_monitor_t<T, 8, 64, 1, true>(p, s) : launch_monitor_t<T, 8, 64, 0, true>(p, s);
    return closure ? launch_monitor_t<T, 16, 64, 1, false>(p, s) : launch_monitor_t<T, 16, 64, 0, false>(p, s);
  }
  if (transpose) return closure ? launch_monitor_t<T, 8, 32, 1, true>(p, s) : launch_monitor_t<T, 8, 32, 0, true>(p, s);
  return closure ? launch_monitor_t<T, 16, 32, 1, false>(p, s) : launch_monitor_t<T, 16, 32, 0, false>(p, s);
}

// ------------------------------------------------------------------ legacy 7-point

template <typename T>
struct LegParams {
  const T* in;
  T* out;
  const T* axm;  // nx entries each: a-, a+ per axis
  const T* axp;
  const T* aym;
  const T* ayp;
  const T* azm;
  const T* azp;
  const T* veff;
  long long ld;
  int nblk, nx, ny, nz;
  T scale;
  int interior_only;
  int transpose;
};

template <typename T>
__global__ __launch_bounds__(IG_THREADS) void legacy_stencil_kernel(LegParams<T> p) {
  const long long nyz = (long long)p.ny * p.nz;
  const long long N = (long long)p.nx * nyz;
  const long long total = N * p.nblk;
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < total;
       t += (long long)gridDim.x * IG_THREADS) {
    const int m = (int)(t / N);
    const long long g = t - (long long)m * N;
    const int i = (int)(g / nyz);
    const int j = (int)((g / p.nz) % p.ny);
    const int k = (int)(g % p.nz);
    const T* f = p.in + (long long)m * p.ld + g;
    const T c = f[0];
    T lap = T(0);
    auto axis = [&](const T* am, const T* ap, int idx, int n, long long stride) {
      const T fm = idx > 0 ? f[-stride] : T(0);
      const T fp = idx < n - 1 ? f[stride] : T(0);
      if (!p.transpose) return ap[idx] * (fp - c) - am[idx] * (c - fm);
      const T left = idx > 0 ? ap[idx - 1] * fm : T(0);
      const T right = idx < n - 1 ? am[idx + 1] * fp : T(0);
      return left + right - (am[idx] + ap[idx]) * c;
    };
    lap += axis(p.axm, p.axp, i, p.nx, nyz);
    lap += axis(p.aym, p.ayp, j, p.ny, p.nz);
    lap += axis(p.azm, p.azp, k, p.nz, 1);
    T v = p.scale * lap;
    if (p.veff) v += p.veff[g] * c;
    if (p.interior_only &&
        (i == 0 || i == p.nx - 1 || j == 0 || j == p.ny - 1 || k == 0 || k == p.nz - 1))
      v = T(0);
    p.out[(long long)m * p.ld + g] = v;
  }
}

template <typename T>
int launch_legacy(const LegParams<T>& p, hipStream_t s) {
  const long long total = (long long)p.nx * p.ny * p.nz * p.nblk;
  long long blocks = (total + IG_THREADS - 1) / IG_THREADS;
  if (blocks > 65536) blocks = 65536;
  if (blocks <= 0) return IG_EINVAL;
  hipLaunchKernelGGL(legacy_stencil_kernel<T>, dim3((unsigned)blocks), dim3(IG_THREADS), 0, s, p);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int monitor_entry(const T* in, T* out, int nblk, long long ld, int nx, int ny, int nz,
                  double hx, double hy, double hz, const T* coef, const T* invj, const T* veff,
                  double scale, int closure, int transpose, int interior_only, int xchunk,
                  int tile, hipStream_t stream) {
  if (!in || !out || !coef || !invj || nblk <= 0 || nx < 3 || ny < 3 || nz < 3) return IG_EINVAL;
  if (ld < (long long)nx * ny * nz) return IG_EINVAL;
  if (closure < 0 || closure > 1) return IG_EINVAL;
  if (tile != 0 && tile != 32 && tile != 64) return IG_EINVAL;
  MonParams<T> p{};
  p.in = in;
  p.out = out;
  p.coef = coef;
  p.invj = invj;
  p.veff = veff;
  p.ld = ld;
  p.nblk = nblk;
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.ih2x = T(0.5 / hx);
  p.ih2y = T(0.5 / hy);
  p.ih2z = T(0.5 / hz);
  p.scale = T(scale);
  p.interior_only = interior_only;
  p.xc = xchunk;
  return launch_monitor<T>(p, closure, transpose, tile, stream);
}

template <typename T>
int legacy_entry(const T* in, T* out, int nblk, long long ld, int nx, int ny, int nz,
                 const T* coeffs, const T* veff, double scale, int transpose, int interior_only,
                 hipStream_t stream) {
  if (!in || !out || !coeffs || nblk <= 0 || nx < 2 || ny < 2 || nz < 2) return IG_EINVAL;
  if (ld < (long long)nx * ny * nz) return IG_EINVAL;
  LegParams<T> p{};
  p.in = in;
  p.out = out;
  p.axm = coeffs;
  p.axp = coeffs + nx;
  p.aym = coeffs + 2 * nx;
  p.ayp = coeffs + 2 * nx + ny;
  p.azm = coeffs + 2 * nx + 2 * ny;
  p.azp = coeffs + 2 * nx + 2 * ny + nz;
  p.veff = veff;
  p.ld = ld;
  p.nblk = nblk;
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.scale = T(scale);
  p.interior_only = interior_only;
  p.transpose = transpose;
  return launch_legacy<T>(p, stream);
}

}  // namespace

extern "C" {

int ig_monitor_stencil_f64(const double* in, double* out, int nblk, long long ld, int nx, int ny,
                           int nz, double hx, double hy, double hz, const double* coef,
                           const double* invj, const double* veff, double scale, int closure,
                           int transpose, int interior_only, int xchunk, int tile,
                           hipStream_t stream) {
  return monitor_entry<double>(in, out, nblk, ld, nx, ny, nz, hx, hy, hz, coef, invj, veff, scale,
                               closure, transpose, interior_only, xchunk, tile, stream);
}

int ig_monitor_stencil_f32(const float* in, float* out, int nblk, long long ld, int nx, int ny,
                           int nz, double hx, double hy, double hz, const float* coef,
                           const float* invj, const float* veff, double scale, int closure,
                           int transpose, int interior_only, int xchunk, int tile,
                           hipStream_t stream) {
  return monitor_entry<float>(in, out, nblk, ld, nx, ny, nz, hx, hy, hz, coef, invj, veff, scale,
                              closure, transpose, interior_only, xchunk, tile, stream);
}

int ig_legacy_stencil_f64(const double* in, double* out, int nblk, long long ld, int nx, int ny,
                          int nz, const double* coeffs, const double* veff, double scale,
                          int transpose, int interior_only, hipStream_t stream) {
  return legacy_entry<double>(in, out, nblk, ld, nx, ny, nz, coeffs, veff, scale, transpose,
                              interior_only, stream);
}

int ig_legacy_stencil_f32(const float* in, float* out, int nblk, long long ld, int nx, int ny,
                          int nz, const float* coeffs, const float* veff, double scale,
                          int transpose, int interior_only, hipStream_t stream) {
  return legacy_entry<float>(in, out, nblk, ld, nx, ny, nz, coeffs, veff, scale, transpose,
                             interior_only, stream);
}

}  // extern "C"
