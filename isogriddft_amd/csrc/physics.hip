// Pointwise physics fields and the multipole machinery of the open-boundary solve.
//
//  * local GTH potential field and its nuclear gradient  (pseudo/local.py:138-180, :303-357)
//  * spin-polarised LSDA, Slater exchange + PW92 correlation (replaces xc/lsda.py:70-129)
//  * separable GTH projectors on atom support boxes: field, projection, application
//                                                        (pseudo/nonlocal.py:99-321)
//  * raw Cartesian multipole moments to order 4, boundary values from moments,
//    and the corrected-moment quadratic sub-cell quadrature
//                                                        (poisson/open_boundary.py:223-495)
#include "common.h"
#include "gth.h"

namespace {

constexpr double PI = 3.14159265358979323846;

// Raw-moment monomials, degrees 0..4 (35 entries), ordered by degree and within a
// degree by (px, py, pz) descending lexicographically: 1; x y z; xx xy xz yy yz zz;
// xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz; xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz
// xyzz xzzz yyyy yyyz yyzz yzzz zzzz.  The host mirrors this order.
__host__ __device__ constexpr int nmono(int order) {
  return order <= 0 ? 1 : order == 1 ? 4 : order == 2 ? 10 : order == 3 ? 20 : 35;
}

// Values of all monomials of degree <= ORDER at (x, y, z), in table order.
template <int ORDER>
__device__ __forceinline__ void monomials(double x, double y, double z, double* m) {
  m[0] = 1.0;
  if (ORDER >= 1) { m[1] = x; m[2] = y; m[3] = z; }
  if (ORDER >= 2) { m[4] = x * x; m[5] = x * y; m[6] = x * z; m[7] = y * y; m[8] = y * z; m[9] = z * z; }
  if (ORDER >= 3) {
    m[10] = m[4] * x; m[11] = m[4] * y; m[12] = m[4] * z; m[13] = m[7] * x; m[14] = m[5] * z;
    m[15] = m[9] * x; m[16] = m[7] * y; m[17] = m[7] * z; m[18] = m[9] * y; m[19] = m[9] * z;
  }
  if (ORDER >= 4) {
    m[20] = m[10] * x; m[21] = m[10] * y; m[22] = m[10] * z; m[23] = m[4] * m[7]; m[24] = m[4] * m[8];
    m[25] = m[4] * m[9]; m[26] = m[16] * x; m[27] = m[13] * z; m[28] = m[15] * y; m[29] = m[19] * x;
    m[30] = m[16] * y; m[31] = m[16] * z; m[32] = m[7] * m[9]; m[33] = m[19] * y; m[34] = m[19] * z;
  }
}

inline int red_blocks(long long n) {
  long long b = (n + IG_THREADS * 4 - 1) / (IG_THREADS * 4);
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return (int)b;
}

// ------------------------------------------------------------------ local GTH potential

__global__ __launch_bounds__(IG_THREADS) void vloc_kernel(const double* __restrict__ xyz, long long N, int natoms,
                                                          const GthAtom* __restrict__ atoms, double* __restrict__ out,
                                                          int accumulate) {
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    const double x = xyz[t], y = xyz[N + t], z = xyz[2 * N + t];
    double v = accumulate ? out[t] : 0.0;
    for (int a = 0; a < natoms; ++a) {
      const double dx = x - atoms[a].x, dy = y - atoms[a].y, dz = z - atoms[a].z;
      v += gth_vloc(sqrt(dx * dx + dy * dy + dz * dz), atoms[a]);
    }
    out[t] = v;
  }
}

// partial[block][3*natoms] of dE/dR_A = -sum_p q_p V'(r) (r_p - R_A)/r,  q = rho w.
__global__ __launch_bounds__(IG_THREADS) void vloc_grad_kernel(const double* __restrict__ xyz, const double* __restrict__ q,
                                                               long long N, int natoms, const GthAtom* __restrict__ atoms,
                                                               double* __restrict__ partial) {
  __shared__ double smem[4 * 3];
  for (int a = 0; a < natoms; ++a) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
      const double dx = xyz[t] - atoms[a].x, dy = xyz[N + t] - atoms[a].y, dz = xyz[2 * N + t] - atoms[a].z;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r > 0.0) {
        const double f = -q[t] * gth_dvloc(r, atoms[a]) / r;
        acc[0] += f * dx;
        acc[1] += f * dy;
        acc[2] += f * dz;
      }
    }
    ig_block_sum<3>(acc, smem);
    if (threadIdx.x == 0) {
      partial[(long long)blockIdx.x * 3 * natoms + 3 * a + 0] = acc[0];
      partial[(long long)blockIdx.x * 3 * natoms + 3 * a + 1] = acc[1];
      partial[(long long)blockIdx.x * 3 * natoms + 3 * a + 2] = acc[2];
    }
  }
}

__global__ void sum_partials(const double* __restrict__ partial, int nparts, int nv, double* __restrict__ out) {
  for (int v = threadIdx.x; v < nv; v += blockDim.x) {
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += partial[(long long)p * nv + v];
    out[v] = s;
  }
}

// ------------------------------------------------------------------ LSDA (Slater + PW92)

struct Pw {
  double A, a1, b1, b2, b3, b4;
};
__constant__ Pw c_pw[3] = {
    {0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294},
    {0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517},
    {0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671},
};

__device__ __forceinline__ void pw_g(double rs, const Pw& p, double& g, double& dg) {
  const double srs = sqrt(rs);
  const double q0 = -2.0 * p.A * (1.0 + p.a1 * rs);
  const double q1 = 2.0 * p.A * (p.b1 * srs + p.b2 * rs + p.b3 * rs * srs + p.b4 * rs * rs);
  const double q1p = p.A * (p.b1 / srs + 2.0 * p.b2 + 3.0 * p.b3 * srs + 4.0 * p.b4 * rs);
  const double lg = log1p(1.0 / q1);
  g = q0 * lg;
  dg = -2.0 * p.A * p.a1 * lg - q0 * q1p / (q1 * q1 + q1);
}

// VWN interpolation function (Vosko, Wilk, Nusair 1980, "VWN5" fit) of x = sqrt(rs) and
// its x derivative for the parameter set (A, b, c, x0).
struct Vwn {
  double A, b, c, x0;
};
__constant__ Vwn c_vwn[3] = {
    {0.0310907, 3.72744, 12.9352, -0.10498},                        // paramagnetic
    {0.01554535, 7.06042, 18.0578, -0.32500},                       // ferromagnetic
    {-1.0 / (6.0 * PI * PI), 1.13107, 13.0045, -0.0047584},         // spin stiffness
};

__device__ __forceinline__ void vwn_g(double x, const Vwn& p, double& g, double& dg) {
  const double X = x * x + p.b * x + p.c;
  const double X0 = p.x0 * p.x0 + p.b * p.x0 + p.c;
  const double Q = sqrt(4.0 * p.c - p.b * p.b);
  const double t = 2.0 * x + p.b;
  const double at = atan(Q / t);
  const double k = p.b * p.x0 / X0;
  g = p.A * (log(x * x / X) + 2.0 * p.b / Q * at - k * (log((x - p.x0) * (x - p.x0) / X) + 2.0 * (p.b + 2.0 * p.x0) / Q * at));
  const double den = t * t + Q * Q;
  dg = p.A * (2.0 / x - t / X - 4.0 * p.b / den - k * (2.0 / (x - p.x0) - t / X - 4.0 * (p.b + 2.0 * p.x0) / den));
}

// Spin-polarised Slater exchange + correlation per point.  FUNC 0: Perdew-Wang 1992;
// FUNC 1: VWN5 with the VWN spin interpolation
//   eps_c = eps_P + alpha f(z) / f''(0) (1 - z^4) + (eps_F - eps_P) f(z) z^4.
template <int FUNC>
__global__ __launch_bounds__(IG_THREADS) void lsda_kernel(const double* __restrict__ ru, const double* __restrict__ rd,
                                                          long long N, double* __restrict__ eps,
                                                          double* __restrict__ vu, double* __restrict__ vd,
                                                          int* __restrict__ flag) {
  const double fz20 = 1.709921;
  const double fzden = 0.5198420997897464;  // 2^(4/3) - 2
  const double cx = 1.2407009817988002;     // (6/pi)^(1/3)
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    double a = ru[t], b = rd[t];
    if (!(a >= -1e-14) || !(b >= -1e-14) || !isfinite(a) || !isfinite(b)) atomicOr(flag, 1);
    a = fmax(a, 0.0);
    b = fmax(b, 0.0);
    const double rho = a + b;
    double e = 0.0, pu = 0.0, pd = 0.0;
    if (rho > 1e-30) {
      const double rs = cbrt(3.0 / (4.0 * PI * rho));
      double zeta = (a - b) / rho;
      zeta = fmin(1.0, fmax(-1.0, zeta));
      const double exv = -0.75 * cx * (a * cbrt(a) + b * cbrt(b));
      const double vxu = -cx * cbrt(a), vxd = -cx * cbrt(b);
      double g0, d0, g1, d1, g2, d2, ac, dac, fpp;
      if (FUNC == 0) {
        pw_g(rs, c_pw[0], g0, d0);
        pw_g(rs, c_pw[1], g1, d1);
        pw_g(rs, c_pw[2], g2, d2);
        ac = -g2;
        dac = -d2;
        fpp = fz20;
      } else {
        // x = sqrt(rs): d/drs = (d/dx) / (2x)
        const double x = sqrt(rs);
        vwn_g(x, c_vwn[0], g0, d0);
        vwn_g(x, c_vwn[1], g1, d1);
        vwn_g(x, c_vwn[2], ac, dac);
        const double ix = 0.5 / x;
        d0 *= ix;
        d1 *= ix;
        dac *= ix;
        fpp = 1.7099209341613657;  // 4 / (9 (2^(1/3) - 1))
      }
      const double opz = cbrt(1.0 + zeta), omz = cbrt(1.0 - zeta);
      const double fz = ((1.0 + zeta) * opz + (1.0 - zeta) * omz - 2.0) / fzden;
      const double dfz = (4.0 / 3.0) * (opz - omz) / fzden;
      const double z3 = zeta * zeta * zeta, z4 = z3 * zeta;
      const double ec = g0 + ac * fz / fpp * (1.0 - z4) + (g1 - g0) * fz * z4;
      const double decdrs = d0 + dac * fz / fpp * (1.0 - z4) + (d1 - d0) * fz * z4;
      const double decdz = ac / fpp * (dfz * (1.0 - z4) - 4.0 * z3 * fz) + (g1 - g0) * (dfz * z4 + 4.0 * z3 * fz);
      const double vcc = ec - rs / 3.0 * decdrs;
      e = exv / rho + ec;
      pu = vxu + vcc - (zeta - 1.0) * decdz;
      pd = vxd + vcc - (zeta + 1.0) * decdz;
    }
    eps[t] = e;
    vu[t] = pu;
    vd[t] = pd;
  }
}

// ------------------------------------------------------------------ nonlocal projectors

struct Box {
  int i0, j0, k0, ni, nj, nk;  // support box origin and extents (grid indices)
};

struct ProjDesc {
  int l, m, i;   // angular momentum, real-harmonic m, zero-based projector index
  double rl;     // channel radius
  double pref;   // sqrt(2) / (rl^(l+(4i+3)/2) sqrt(Gamma(l+(4i+3)/2)))
};

__device__ __forceinline__ long long box_to_grid(const Box& b, long long q, int ny, int nz) {
  const int bk = (int)(q % b.nk);
  const int bj = (int)((q / b.nk) % b.nj);
  const int bi = (int)(q / ((long long)b.nk * b.nj));
  return ((long long)(b.i0 + bi) * ny + (b.j0 + bj)) * nz + (b.k0 + bk);
}

__global__ __launch_bounds__(IG_THREADS) void proj_field_kernel(const double* __restrict__ xyz, long long N, int ny, int nz,
                                                                Box box, double ax, double ay, double az, int nproj,
                                                                const ProjDesc* __restrict__ pd, double* __restrict__ P,
                                                                int deriv) {
  // deriv = 0: P[p][q] = p(r) Y(r);  deriv = 1..3: d/dR_{x,y,z} of the projector (atom moves).
  const long long nb = (long long)box.ni * box.nj * box.nk;
  const double y1 = 0.4886025119029199;  // sqrt(3/(4 pi))
  const double y0 = 0.28209479177387814; // 1/sqrt(4 pi)
  for (long long q = (long long)blockIdx.x * IG_THREADS + threadIdx.x; q < nb; q += (long long)gridDim.x * IG_THREADS) {
    const long long g = box_to_grid(box, q, ny, nz);
    const double dx = xyz[g] - ax, dy = xyz[N + g] - ay, dz = xyz[2 * N + g] - az;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    for (int p = 0; p < nproj; ++p) {
      const ProjDesc d = pd[p];
      const int pw = d.l + 2 * d.i;
      const double gau = exp(-0.5 * r2 / (d.rl * d.rl));
      // f(r) = pref r^pw e^{-r^2/2rl^2}; for l = 1 the harmonic (y,z,x)/r combines with
      // r^pw into r^(pw-1) * component.
      double val;
      double grad[3] = {0.0, 0.0, 0.0};
      if (d.l == 0) {
        const double rp = pw == 0 ? 1.0 : pow(r, (double)pw);
        val = d.pref * rp * gau * y0;
        if (deriv) {
          // d/dr_vec of f(r) = f'(r) r_vec / r; f' = pref e (pw r^(pw-1) - r^(pw+1)/rl^2)
          const double fr = d.pref * gau * ((pw > 0 ? pw * pow(r, (double)(pw - 2)) : 0.0) - (pw == 0 ? 1.0 : pow(r, (double)pw)) / (d.rl * d.rl)) * y0;
          grad[0] = fr * dx;
          grad[1] = fr * dy;
          grad[2] = fr * dz;
        }
      } else {
        const double comp = d.m == -1 ? dy : (d.m == 0 ? dz : dx);
        const int cidx = d.m == -1 ? 1 : (d.m == 0 ? 2 : 0);
        const double rp = pw - 1 == 0 ? 1.0 : pow(r, (double)(pw - 1));  // r^(pw-1), pw >= 1
        val = (r > 0.0) ? d.pref * y1 * rp * gau * comp : 0.0;
        if (deriv && r > 0.0) {
          // h(r) = pref y1 r^(pw-1) e; value = h(r) * comp; grad = h' r_vec/r comp + h e_c
          const double h = d.pref * y1 * rp * gau;
          const double hp = d.pref * y1 * gau * ((pw - 1 > 0 ? (pw - 1) * pow(r, (double)(pw - 2)) : 0.0) - rp * r / (d.rl * d.rl));
          grad[0] = hp * dx / r * comp;
          grad[1] = hp * dy / r * comp;
          grad[2] = hp * dz / r * comp;
          grad[cidx] += h;
        }
      }
      // the atom moves: d/dR = - d/dr
      P[(long long)p * nb + q] = deriv ? -grad[deriv - 1] : val;
    }
  }
}

// partial[blk][chunk][nproj] = sum_{q in chunk} P[p][q] w(q) psi_m(q)
template <typename T>
__global__ __launch_bounds__(IG_THREADS) void proj_dot_kernel(const T* __restrict__ psi, long long ld, Box box, int ny,
                                                              int nz, const double* __restrict__ P, int nproj,
                                                              const double* __restrict__ w, double* __restrict__ partial) {
  __shared__ double smem[4 * 8];
  const int m = blockIdx.y;
  const long long nb = (long long)box.ni * box.nj * box.nk;
  const T* f = psi + (long long)m * ld;
  for (int p0 = 0; p0 < nproj; p0 += 8) {
    double acc[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) acc[p] = 0.0;
    for (long long q = (long long)blockIdx.x * IG_THREADS + threadIdx.x; q < nb; q += (long long)gridDim.x * IG_THREADS) {
      const long long g = box_to_grid(box, q, ny, nz);
      const double fw = (double)f[g] * w[g];
#pragma unroll
      for (int p = 0; p < 8; ++p)
        if (p0 + p < nproj) acc[p] += P[(long long)(p0 + p) * nb + q] * fw;
    }
    ig_block_sum<8>(acc, smem);
    if (threadIdx.x == 0) {
      for (int p = 0; p < 8 && p0 + p < nproj; ++p)
        partial[((long long)m * gridDim.x + blockIdx.x) * nproj + p0 + p] = acc[p];
    }
  }
}

// coef[m][p] = sum_chunks partial; d[m][p] = sum_p' H[p][p'] coef[m][p']
__global__ void proj_finalize(const double* __restrict__ partial, int nchunks, int nblk, int nproj,
                              const double* __restrict__ H, double* __restrict__ coef, double* __restrict__ d) {
  const int m = blockIdx.x;
  __shared__ double c[64];
  for (int p = threadIdx.x; p < nproj; p += blockDim.x) {
    double s = 0.0;
    for (int ch = 0; ch < nchunks; ++ch) s += partial[((long long)m * nchunks + ch) * nproj + p];
    c[p] = s;
    if (coef) coef[(long long)m * nproj + p] = s;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < nproj; p += blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < nproj; ++q) s += H[p * nproj + q] * c[q];
    d[(long long)m * nproj + p] = s;
  }
}

// out_m(q) += sum_p P[p][q] d[m][p]
template <typename T>
__global__ __launch_bounds__(IG_THREADS) void proj_apply_kernel(T* __restrict__ out, long long ld, Box box, int ny, int nz,
                                                                const double* __restrict__ P, int nproj,
                                                                const double* __restrict__ d) {
  const int m = blockIdx.y;
  const long long nb = (long long)box.ni * box.nj * box.nk;
  T* o = out + (long long)m * ld;
  for (long long q = (long long)blockIdx.x * IG_THREADS + threadIdx.x; q < nb; q += (long long)gridDim.x * IG_THREADS) {
    const long long g = box_to_grid(box, q, ny, nz);
    double s = 0.0;
    for (int p = 0; p < nproj; ++p) s += P[(long long)p * nb + q] * d[(long long)m * nproj + p];
    o[g] = (T)((double)o[g] + s);
  }
}

// ------------------------------------------------------------------ orbital density

// rho(t) (=|+=) sum_m occ[m] psi_m(t)^2, FP64 accumulation and output.
template <typename T>
__global__ __launch_bounds__(IG_THREADS) void density_kernel(const T* __restrict__ psi, long long ld, int nblk,
                                                             long long N, const double* __restrict__ occ,
                                                             double* __restrict__ rho, int accumulate) {
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    double s = accumulate ? rho[t] : 0.0;
    for (int m = 0; m < nblk; ++m) {
      const double v = (double)psi[(long long)m * ld + t];
      s = fma(occ[m] * v, v, s);
    }
    rho[t] = s;
  }
}

// ------------------------------------------------------------------ multipoles

template <int ORDER>
__global__ __launch_bounds__(IG_THREADS) void moments_kernel(const double* __restrict__ rho, const double* __restrict__ w,
                                                             const double* __restrict__ frac, const double* __restrict__ xyz,
                                                             long long N, double cx, double cy, double cz,
                                                             double* __restrict__ partial) {
  constexpr int NM = nmono(ORDER);
  __shared__ double smem[4 * NM];
  double acc[NM];
#pragma unroll
  for (int i = 0; i < NM; ++i) acc[i] = 0.0;
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    double q = rho[t] * w[t];
    if (frac) q *= frac[t];
    double mm[NM];
    monomials<ORDER>(xyz[t] - cx, xyz[N + t] - cy, xyz[2 * N + t] - cz, mm);
#pragma unroll
    for (int i = 0; i < NM; ++i) acc[i] += q * mm[i];
  }
  ig_block_sum<NM>(acc, smem);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NM; ++i) partial[(long long)blockIdx.x * NM + i] = acc[i];
  }
}

// Boundary potential phi = sum_l Q_l(r) / r^(2l+1) with Q_l homogeneous of degree l
// (coefficients over the monomial table); written on the six faces, zero inside.
__global__ __launch_bounds__(IG_THREADS) void boundary_kernel(const double* __restrict__ xyz, int nx, int ny, int nz,
                                                              double cx, double cy, double cz, int order,
                                                              const double* __restrict__ qcoef, double* __restrict__ out) {
  const long long N = (long long)nx * ny * nz;
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    const int k = (int)(t % nz);
    const int j = (int)((t / nz) % ny);
    const int i = (int)(t / ((long long)ny * nz));
    const bool bnd = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1);
    if (!bnd) {
      out[t] = 0.0;
      continue;
    }
    const double dx = xyz[t] - cx, dy = xyz[N + t] - cy, dz = xyz[2 * N + t] - cz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    double mm[35];
    monomials<4>(dx, dy, dz, mm);
    double phi = 0.0;
    double rpow = r;  // r^(2l+1)
    int off = 0;
    for (int l = 0; l <= order; ++l) {
      const int nl = (l + 1) * (l + 2) / 2;
      double ql = 0.0;
      for (int s = 0; s < nl; ++s) ql += qcoef[off + s] * mm[off + s];
      phi += ql / rpow;
      rpow *= r2;
      off += nl;
    }
    out[t] = phi;
  }
}

// Quadratic least-squares fit of rho on the 4x4x4 node stencil of each selected cell,
// in coordinates centred on the stencil centroid and scaled by the largest offset (the
// fitted function does not depend on that affine change of basis; centring keeps the
// 10 x 10 normal equations well conditioned).  Writes coef[cell][14]: the 10 basis
// coefficients (1, x, y, z, xx, xy, xz, yy, yz, zz of the scaled coordinates), the
// centroid and the inverse scale.
__global__ __launch_bounds__(IG_THREADS) void cell_fit_kernel(const double* __restrict__ rho, const double* __restrict__ xyz,
                                                              int nx, int ny, int nz, const int* __restrict__ cells,
                                                              int ncells, double* __restrict__ coef) {
  const long long N = (long long)nx * ny * nz;
  for (int c = blockIdx.x * IG_THREADS + threadIdx.x; c < ncells; c += gridDim.x * IG_THREADS) {
    const int ci = cells[3 * c], cj = cells[3 * c + 1], ck = cells[3 * c + 2];
    double ox = 0.0, oy = 0.0, oz = 0.0;
    for (int a = -1; a <= 2; ++a)
      for (int b = -1; b <= 2; ++b)
        for (int d = -1; d <= 2; ++d) {
          const long long g = ((long long)(ci + a) * ny + (cj + b)) * nz + (ck + d);
          ox += xyz[g];
          oy += xyz[N + g];
          oz += xyz[2 * N + g];
        }
    ox *= 1.0 / 64.0;
    oy *= 1.0 / 64.0;
    oz *= 1.0 / 64.0;
    double sc = 0.0;
    for (int a = -1; a <= 2; ++a)
      for (int b = -1; b <= 2; ++b)
        for (int d = -1; d <= 2; ++d) {
          const long long g = ((long long)(ci + a) * ny + (cj + b)) * nz + (ck + d);
          sc = fmax(sc, fmax(fabs(xyz[g] - ox), fmax(fabs(xyz[N + g] - oy), fabs(xyz[2 * N + g] - oz))));
        }
    const double is = 1.0 / sc;
    double A[55];
    double rhs[10];
    for (int e = 0; e < 55; ++e) A[e] = 0.0;
    for (int e = 0; e < 10; ++e) rhs[e] = 0.0;
    for (int a = -1; a <= 2; ++a)
      for (int b = -1; b <= 2; ++b)
        for (int d = -1; d <= 2; ++d) {
          const long long g = ((long long)(ci + a) * ny + (cj + b)) * nz + (ck + d);
          double m[10];
          monomials<2>((xyz[g] - ox) * is, (xyz[N + g] - oy) * is, (xyz[2 * N + g] - oz) * is, m);
          const double f = rho[g];
          int e = 0;
          for (int r = 0; r < 10; ++r) {
            rhs[r] += m[r] * f;
            for (int s = 0; s <= r; ++s) A[e++] += m[r] * m[s];
          }
        }
    // Cholesky A = L L^T in place (packed lower triangle, row-major).
    for (int r = 0; r < 10; ++r) {
      for (int s = 0; s <= r; ++s) {
        double v = A[r * (r + 1) / 2 + s];
        for (int t2 = 0; t2 < s; ++t2) v -= A[r * (r + 1) / 2 + t2] * A[s * (s + 1) / 2 + t2];
        if (r == s) A[r * (r + 1) / 2 + s] = sqrt(fmax(v, 1e-300));
        else A[r * (r + 1) / 2 + s] = v / A[s * (s + 1) / 2 + s];
      }
    }
    double y[10];
    for (int r = 0; r < 10; ++r) {
      double v = rhs[r];
      for (int s = 0; s < r; ++s) v -= A[r * (r + 1) / 2 + s] * y[s];
      y[r] = v / A[r * (r + 1) / 2 + r];
    }
    for (int r = 9; r >= 0; --r) {
      double v = y[r];
      for (int s = r + 1; s < 10; ++s) v -= A[s * (s + 1) / 2 + r] * y[s];
      y[r] = v / A[r * (r + 1) / 2 + r];
    }
    for (int r = 0; r < 10; ++r) coef[(long long)c * 14 + r] = y[r];
    coef[(long long)c * 14 + 10] = ox;
    coef[(long long)c * 14 + 11] = oy;
    coef[(long long)c * 14 + 12] = oz;
    coef[(long long)c * 14 + 13] = is;
  }
}

// Sums over the 2x2x2 sub-cell midpoints of the selected cells (trilinear position and
// Jacobian).  MODE 0: raw moments about c of q = rho_fit J subvol (corrected moments);
// MODE 1: [sum |d|, sum |d| x, sum |d| y, sum |d| z] and MODE 2: raw moments about c of
// d = (rho_fit - rho_trilinear) J subvol (the split correction of the reference's
// legacy_split mode, poisson/open_boundary.py:559-719).
template <int ORDER, int MODE>
__global__ __launch_bounds__(IG_THREADS) void cell_moments_kernel(const double* __restrict__ xyz, const double* __restrict__ jac,
                                                                  const double* __restrict__ rho,
                                                                  int nx, int ny, int nz, const int* __restrict__ cells,
                                                                  int ncells, const double* __restrict__ coef, double subvol,
                                                                  double cx, double cy, double cz, double* __restrict__ partial) {
  constexpr int NM = MODE == 1 ? 4 : nmono(ORDER);
  __shared__ double smem[4 * NM];
  const long long N = (long long)nx * ny * nz;
  double acc[NM];
#pragma unroll
  for (int i = 0; i < NM; ++i) acc[i] = 0.0;
  for (int c = blockIdx.x * IG_THREADS + threadIdx.x; c < ncells; c += gridDim.x * IG_THREADS) {
    const int ci = cells[3 * c], cj = cells[3 * c + 1], ck = cells[3 * c + 2];
    double X[8], Y[8], Z[8], Jc[8], Rc[8];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 2; ++d) {
          const long long g = ((long long)(ci + a) * ny + (cj + b)) * nz + (ck + d);
          const int e = a * 4 + b * 2 + d;
          X[e] = xyz[g];
          Y[e] = xyz[N + g];
          Z[e] = xyz[2 * N + g];
          Jc[e] = jac[g];
          Rc[e] = MODE == 0 ? 0.0 : rho[g];
        }
    const double* cf = coef + (long long)c * 14;
    const double ox = cf[10], oy = cf[11], oz = cf[12], is = cf[13];
    for (int s = 0; s < 8; ++s) {
      const double u = (s & 4) ? 0.75 : 0.25, v = (s & 2) ? 0.75 : 0.25, ww = (s & 1) ? 0.75 : 0.25;
      double tri[8];
      for (int e = 0; e < 8; ++e) {
        const double fu = (e & 4) ? u : 1.0 - u, fv = (e & 2) ? v : 1.0 - v, fw = (e & 1) ? ww : 1.0 - ww;
        tri[e] = fu * fv * fw;
      }
      double xs = 0.0, ys = 0.0, zs = 0.0, js = 0.0, rt = 0.0;
      for (int e = 0; e < 8; ++e) {
        xs += tri[e] * X[e];
        ys += tri[e] * Y[e];
        zs += tri[e] * Z[e];
        js += tri[e] * Jc[e];
        rt += tri[e] * Rc[e];
      }
      double m2[10];
      monomials<2>((xs - ox) * is, (ys - oy) * is, (zs - oz) * is, m2);
      double fit = 0.0;
      for (int r = 0; r < 10; ++r) fit += cf[r] * m2[r];
      const double q = (MODE == 0 ? fit : fit - rt) * js * subvol;
      if constexpr (MODE == 1) {
        const double aq = fabs(q);
        acc[0] += aq;
        acc[1] += aq * xs;
        acc[2] += aq * ys;
        acc[3] += aq * zs;
      } else {
        double mm[nmono(ORDER)];
        monomials<ORDER>(xs - cx, ys - cy, zs - cz, mm);
#pragma unroll
        for (int i = 0; i < nmono(ORDER); ++i) acc[i] += q * mm[i];
      }
    }
  }
  ig_block_sum<NM>(acc, smem);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NM; ++i) partial[(long long)blockIdx.x * NM + i] = acc[i];
  }
}

template <int ORDER>
int moments_t(const double* rho, const double* w, const double* frac, const double* xyz, long long N,
              const double* center, double* work, double* out, hipStream_t s) {
  constexpr int NM = nmono(ORDER);
  const int nb = red_blocks(N);
  hipLaunchKernelGGL(moments_kernel<ORDER>, dim3(nb), dim3(IG_THREADS), 0, s, rho, w, frac, xyz, N, center[0],
                     center[1], center[2], work);
  hipLaunchKernelGGL(sum_partials, dim3(1), dim3(64), 0, s, work, nb, NM, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <int ORDER, int MODE = 0>
int cell_moments_t(const double* xyz, const double* jac, int nx, int ny, int nz, const int* cells, int ncells,
                   const double* coef, double subvol, const double* center, double* work, double* out, hipStream_t s,
                   const double* rho = nullptr) {
  constexpr int NM = MODE == 1 ? 4 : nmono(ORDER);
  int nb = (ncells + IG_THREADS - 1) / IG_THREADS;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  const double c0 = center ? center[0] : 0.0, c1 = center ? center[1] : 0.0, c2 = center ? center[2] : 0.0;
  hipLaunchKernelGGL((cell_moments_kernel<ORDER, MODE>), dim3(nb), dim3(IG_THREADS), 0, s, xyz, jac, rho, nx, ny, nz,
                     cells, ncells, coef, subvol, c0, c1, c2, work);
  hipLaunchKernelGGL(sum_partials, dim3(1), dim3(64), 0, s, work, nb, NM, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

inline unsigned grid1d(long long n) {
  long long b = (n + IG_THREADS - 1) / IG_THREADS;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// ------------------------------------------------------------------ moment adjoints
// The multipole moments are linear functionals of rho; their adjoints map a vector q of
// moment weights back to a field: out = sum_j q_j d moment_j / d rho.

struct MomentWeights {
  double q[35];
};

// Nodal moments sum_p w_p (1 - frac_p) rho_p mono_j(x_p - c)  (frac may be NULL).
template <int ORDER>
__global__ __launch_bounds__(IG_THREADS) void moments_adjoint_kernel(const double* __restrict__ w,
                                                                     const double* __restrict__ frac,
                                                                     const double* __restrict__ xyz, long long N,
                                                                     double cx, double cy, double cz, MomentWeights q,
                                                                     double* __restrict__ out, int accumulate) {
  constexpr int NM = nmono(ORDER);
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    double mm[NM];
    monomials<ORDER>(xyz[t] - cx, xyz[N + t] - cy, xyz[2 * N + t] - cz, mm);
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) sum += q.q[i] * mm[i];
    const double v = w[t] * (frac ? 1.0 - frac[t] : 1.0) * sum;
    out[t] = accumulate ? out[t] + v : v;
  }
}

// Sub-cell quadratic-fit moments (cell_fit_kernel + cell_moments_kernel<ORDER, 0>): per
// cell, with A the normal matrix of the 4x4x4 stencil fit and a = sum_s d_s J_s subvol
// sum_j q_j mono_j(x_s - c) (d_s the fit basis at sample s), the stencil nodes receive
// d_p . A^-1 a.  Scattered with atomic adds (cells share stencil nodes).
template <int ORDER>
__global__ __launch_bounds__(IG_THREADS) void cell_moments_adjoint_kernel(
    const double* __restrict__ xyz, const double* __restrict__ jac, int nx, int ny, int nz,
    const int* __restrict__ cells, int ncells, double subvol, double cx, double cy, double cz, MomentWeights q,
    double* __restrict__ out) {
  constexpr int NM = nmono(ORDER);
  const long long N = (long long)nx * ny * nz;
  for (int c = blockIdx.x * IG_THREADS + threadIdx.x; c < ncells; c += gridDim.x * IG_THREADS) {
    const int ci = cells[3 * c], cj = cells[3 * c + 1], ck = cells[3 * c + 2];
    // centroid and scale of the stencil, as the forward fit
    double ox = 0.0, oy = 0.0, oz = 0.0;
    for (int a = -1; a <= 2; ++a)
      for (int b = -1; b <= 2; ++b)
        for (int d = -1; d <= 2; ++d) {
          const long long g = ((long long)(ci + a) * ny + (cj + b)) * nz + (ck + d);
          ox += xyz[g];
          oy += xyz[N + g];
          oz += xyz[2 * N + g];
        }
    ox *= 1.0 / 64.0;
    oy *= 1.0 / 64.0;
    oz *= 1.0 / 64.0;
    double sc = 0.0;
    for (int a = -1; a <= 2; ++a)
      for (int b = -1; b <= 2; ++b)
        for (int d = -1; d <= 2; ++d) {
          const long long g = ((long long)(ci + a) * ny + (cj + b)) * nz + (ck + d);
          sc = fmax(sc, fmax(fabs(xyz[g] - ox), fmax(fabs(xyz[N + g] - oy), fabs(xyz[2 * N + g] - oz))));
        }
    const double is = 1.0 / sc;
    double A[55];
    for (int e = 0; e < 55; ++e) A[e] = 0.0;
    for (int a = -1; a <= 2; ++a)
      for (int b = -1; b <= 2; ++b)
        for (int d = -1; d <= 2; ++d) {
          const long long g = ((long long)(ci + a) * ny + (cj + b)) * nz + (ck + d);
          double m[10];
          monomials<2>((xyz[g] - ox) * is, (xyz[N + g] - oy) * is, (xyz[2 * N + g] - oz) * is, m);
          int e = 0;
          for (int r = 0; r < 10; ++r)
            for (int t2 = 0; t2 <= r; ++t2) A[e++] += m[r] * m[t2];
        }
    for (int r = 0; r < 10; ++r) {
      for (int t2 = 0; t2 <= r; ++t2) {
        double v = A[r * (r + 1) / 2 + t2];
        for (int u = 0; u < t2; ++u) v -= A[r * (r + 1) / 2 + u] * A[t2 * (t2 + 1) / 2 + u];
        if (r == t2) A[r * (r + 1) / 2 + t2] = sqrt(fmax(v, 1e-300));
        else A[r * (r + 1) / 2 + t2] = v / A[t2 * (t2 + 1) / 2 + t2];
      }
    }
    // a = sum over the 2x2x2 sub-cell samples of d_s * J_s subvol * (q . mono(x_s - c))
    double X[8], Y[8], Z[8], Jc[8];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 2; ++d) {
          const long long g = ((long long)(ci + a) * ny + (cj + b)) * nz + (ck + d);
          const int e = a * 4 + b * 2 + d;
          X[e] = xyz[g];
          Y[e] = xyz[N + g];
          Z[e] = xyz[2 * N + g];
          Jc[e] = jac[g];
        }
    double av[10];
    for (int r = 0; r < 10; ++r) av[r] = 0.0;
    for (int smp = 0; smp < 8; ++smp) {
      const double u = (smp & 4) ? 0.75 : 0.25, v = (smp & 2) ? 0.75 : 0.25, ww = (smp & 1) ? 0.75 : 0.25;
      double xs = 0.0, ys = 0.0, zs = 0.0, js = 0.0;
      for (int e = 0; e < 8; ++e) {
        const double fu = (e & 4) ? u : 1.0 - u, fv = (e & 2) ? v : 1.0 - v, fw = (e & 1) ? ww : 1.0 - ww;
        const double tri = fu * fv * fw;
        xs += tri * X[e];
        ys += tri * Y[e];
        zs += tri * Z[e];
        js += tri * Jc[e];
      }
      double mm[NM];
      monomials<ORDER>(xs - cx, ys - cy, zs - cz, mm);
      double qs = 0.0;
#pragma unroll
      for (int i = 0; i < NM; ++i) qs += q.q[i] * mm[i];
      double m2[10];
      monomials<2>((xs - ox) * is, (ys - oy) * is, (zs - oz) * is, m2);
      const double wgt = js * subvol * qs;
      for (int r = 0; r < 10; ++r) av[r] += wgt * m2[r];
    }
    // z = A^-1 a through the Cholesky factor
    double z[10];
    for (int r = 0; r < 10; ++r) {
      double v = av[r];
      for (int t2 = 0; t2 < r; ++t2) v -= A[r * (r + 1) / 2 + t2] * z[t2];
      z[r] = v / A[r * (r + 1) / 2 + r];
    }
    for (int r = 9; r >= 0; --r) {
      double v = z[r];
      for (int t2 = r + 1; t2 < 10; ++t2) v -= A[t2 * (t2 + 1) / 2 + r] * z[t2];
      z[r] = v / A[r * (r + 1) / 2 + r];
    }
    for (int a = -1; a <= 2; ++a)
      for (int b = -1; b <= 2; ++b)
        for (int d = -1; d <= 2; ++d) {
          const long long g = ((long long)(ci + a) * ny + (cj + b)) * nz + (ck + d);
          double m[10];
          monomials<2>((xyz[g] - ox) * is, (xyz[N + g] - oy) * is, (xyz[2 * N + g] - oz) * is, m);
          double beta = 0.0;
          for (int r = 0; r < 10; ++r) beta += m[r] * z[r];
          atomicAdd(&out[g], beta);
        }
  }
}

template <int ORDER>
int moments_adjoint_t(const double* w, const double* frac, const double* xyz, long long N, const double* center,
                      const double* q, double* out, int accumulate, hipStream_t s) {
  MomentWeights mq;
  for (int i = 0; i < 35; ++i) mq.q[i] = i < nmono(ORDER) ? q[i] : 0.0;
  hipLaunchKernelGGL(moments_adjoint_kernel<ORDER>, dim3(grid1d(N)), dim3(IG_THREADS), 0, s, w, frac, xyz, N,
                     center[0], center[1], center[2], mq, out, accumulate);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <int ORDER>
int cell_moments_adjoint_t(const double* xyz, const double* jac, int nx, int ny, int nz, const int* cells,
                           int ncells, double subvol, const double* center, const double* q, double* out,
                           hipStream_t s) {
  MomentWeights mq;
  for (int i = 0; i < 35; ++i) mq.q[i] = i < nmono(ORDER) ? q[i] : 0.0;
  int nb = (ncells + IG_THREADS - 1) / IG_THREADS;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(cell_moments_adjoint_kernel<ORDER>, dim3(nb), dim3(IG_THREADS), 0, s, xyz, jac, nx, ny, nz,
                     cells, ncells, subvol, center[0], center[1], center[2], mq, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

}  // namespace

extern "C" {

int ig_physics_work_size(long long N) { return 1024 * 35 + 64; }

int ig_vloc_gth_f64(const double* xyz, long long N, int natoms, const double* atoms, double* out, int accumulate,
                    hipStream_t s) {
  if (!xyz || !atoms || !out || N <= 0 || natoms < 0) return IG_EINVAL;
  hipLaunchKernelGGL(vloc_kernel, dim3(grid1d(N)), dim3(IG_THREADS), 0, s, xyz, N, natoms,
                     (const GthAtom*)atoms, out, accumulate);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_vloc_gth_grad_f64(const double* xyz, const double* q, long long N, int natoms, const double* atoms,
                         double* work, double* out, hipStream_t s) {
  if (!xyz || !q || !atoms || !out || !work || N <= 0 || natoms <= 0 || natoms > 64) return IG_EINVAL;
  // partials: nb x 3 natoms doubles, which must fit the ig_physics_work_size buffer
  // (1024 x 35 of its doubles): 64 atoms allow 186 blocks
  int nb = red_blocks(N) > 256 ? 256 : red_blocks(N);
  const int cap = (1024 * 35) / (3 * natoms);
  if (nb > cap) nb = cap;
  hipLaunchKernelGGL(vloc_grad_kernel, dim3(nb), dim3(IG_THREADS), 0, s, xyz, q, N, natoms, (const GthAtom*)atoms,
                     work);
  hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, s, work, nb, 3 * natoms, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_lsda_pw92_f64(const double* rho_up, const double* rho_down, long long N, double* eps, double* v_up,
                     double* v_down, int* flag, hipStream_t s) {
  if (!rho_up || !rho_down || !eps || !v_up || !v_down || !flag || N <= 0) return IG_EINVAL;
  hipLaunchKernelGGL(lsda_kernel<0>, dim3(grid1d(N)), dim3(IG_THREADS), 0, s, rho_up, rho_down, N, eps, v_up, v_down,
                     flag);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_lsda_vwn_f64(const double* rho_up, const double* rho_down, long long N, double* eps, double* v_up,
                     double* v_down, int* flag, hipStream_t s) {
  if (!rho_up || !rho_down || !eps || !v_up || !v_down || !flag || N <= 0) return IG_EINVAL;
  hipLaunchKernelGGL(lsda_kernel<1>, dim3(grid1d(N)), dim3(IG_THREADS), 0, s, rho_up, rho_down, N, eps, v_up, v_down,
                     flag);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_projector_fields_f64(const double* xyz, int nx, int ny, int nz, const int* box, const double* atom_pos,
                            int nproj, const void* descs, double* P, int deriv, hipStream_t s) {
  if (!xyz || !box || !atom_pos || !descs || !P || nproj <= 0 || nproj > 64 || deriv < 0 || deriv > 3) return IG_EINVAL;
  Box b{box[0], box[1], box[2], box[3], box[4], box[5]};
  if (b.i0 < 0 || b.j0 < 0 || b.k0 < 0 || b.ni <= 0 || b.nj <= 0 || b.nk <= 0 || b.i0 + b.ni > nx ||
      b.j0 + b.nj > ny || b.k0 + b.nk > nz)
    return IG_EINVAL;
  const long long N = (long long)nx * ny * nz;
  const long long nb = (long long)b.ni * b.nj * b.nk;
  hipLaunchKernelGGL(proj_field_kernel, dim3(grid1d(nb)), dim3(IG_THREADS), 0, s, xyz, N, ny, nz, b, atom_pos[0],
                     atom_pos[1], atom_pos[2], nproj, (const ProjDesc*)descs, P, deriv);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_projector_desc_size(void) { return (int)sizeof(ProjDesc); }

static int project_common(int nx, int ny, int nz, const int* box, int nproj, int nblk, Box& b, int& nchunks) {
  if (nproj <= 0 || nproj > 64 || nblk <= 0 || nblk > 65535) return IG_EINVAL;
  b = Box{box[0], box[1], box[2], box[3], box[4], box[5]};
  if (b.i0 < 0 || b.j0 < 0 || b.k0 < 0 || b.ni <= 0 || b.nj <= 0 || b.nk <= 0 || b.i0 + b.ni > nx ||
      b.j0 + b.nj > ny || b.k0 + b.nk > nz)
    return IG_EINVAL;
  const long long nb = (long long)b.ni * b.nj * b.nk;
  long long c = (nb + IG_THREADS * 8 - 1) / (IG_THREADS * 8);
  if (c < 1) c = 1;
  if (c > 64) c = 64;
  nchunks = (int)c;
  return IG_OK;
}

int ig_nonlocal_work_size(int nblk, int nproj) { return 64 * nblk * nproj + 8; }

#define IG_PROJECT(T, SFX)                                                                                   \
  int ig_nonlocal_project_##SFX(const T* psi, long long ld, int nblk, int nx, int ny, int nz, const int* box, \
                                const double* P, int nproj, const double* w, const double* H, double* work,  \
                                double* coef, double* d, hipStream_t s) {                                    \
    if (!psi || !box || !P || !w || !H || !work || !d) return IG_EINVAL;                                     \
    Box b;                                                                                                   \
    int nchunks;                                                                                             \
    if (project_common(nx, ny, nz, box, nproj, nblk, b, nchunks) != IG_OK) return IG_EINVAL;                 \
    hipLaunchKernelGGL(proj_dot_kernel<T>, dim3(nchunks, nblk), dim3(IG_THREADS), 0, s, psi, ld, b, ny, nz,  \
                       P, nproj, w, work);                                                                   \
    hipLaunchKernelGGL(proj_finalize, dim3(nblk), dim3(64), 0, s, work, nchunks, nblk, nproj, H, coef, d);   \
    IG_CHECK_LAUNCH();                                                                                       \
    return IG_OK;                                                                                            \
  }                                                                                                          \
  int ig_nonlocal_apply_##SFX(T* out, long long ld, int nblk, int nx, int ny, int nz, const int* box,        \
                              const double* P, int nproj, const double* d, hipStream_t s) {                  \
    if (!out || !box || !P || !d) return IG_EINVAL;                                                          \
    Box b;                                                                                                   \
    int nchunks;                                                                                             \
    if (project_common(nx, ny, nz, box, nproj, nblk, b, nchunks) != IG_OK) return IG_EINVAL;                 \
    const long long nb = (long long)b.ni * b.nj * b.nk;                                                      \
    unsigned gx = (unsigned)((nb + IG_THREADS - 1) / IG_THREADS);                                            \
    if (gx > 4096) gx = 4096;                                                                                \
    hipLaunchKernelGGL(proj_apply_kernel<T>, dim3(gx, nblk), dim3(IG_THREADS), 0, s, out, ld, b, ny, nz, P,  \
                       nproj, d);                                                                            \
    IG_CHECK_LAUNCH();                                                                                       \
    return IG_OK;                                                                                            \
  }

IG_PROJECT(double, f64)
IG_PROJECT(float, f32)

#define IG_DENSITY(SFX, T)                                                                                 \
  int ig_density_##SFX(const T* psi, long long ld, int nblk, long long N, const double* occ, double* rho,   \
                       int accumulate, hipStream_t s) {                                                    \
    if (!psi || !occ || !rho || nblk <= 0 || N <= 0 || ld < N) return IG_EINVAL;                           \
    hipLaunchKernelGGL(density_kernel<T>, dim3(grid1d(N)), dim3(IG_THREADS), 0, s, psi, ld, nblk, N, occ,  \
                       rho, accumulate);                                                                   \
    IG_CHECK_LAUNCH();                                                                                     \
    return IG_OK;                                                                                          \
  }
IG_DENSITY(f64, double)
IG_DENSITY(f32, float)
#undef IG_DENSITY

int ig_multipole_moments_f64(const double* rho, const double* w, const double* frac, const double* xyz, long long N,
                             const double* center, int order, double* work, double* out, hipStream_t s) {
  if (!rho || !w || !xyz || !center || !work || !out || N <= 0) return IG_EINVAL;
  switch (order) {
    case 0: return moments_t<0>(rho, w, frac, xyz, N, center, work, out, s);
    case 1: return moments_t<1>(rho, w, frac, xyz, N, center, work, out, s);
    case 2: return moments_t<2>(rho, w, frac, xyz, N, center, work, out, s);
    case 3: return moments_t<3>(rho, w, frac, xyz, N, center, work, out, s);
    case 4: return moments_t<4>(rho, w, frac, xyz, N, center, work, out, s);
    default: return IG_EUNSUPPORTED;
  }
}

int ig_boundary_values_f64(const double* xyz, int nx, int ny, int nz, const double* center, int order,
                           const double* qcoef, double* out, hipStream_t s) {
  if (!xyz || !center || !qcoef || !out || order < 0 || order > 4) return IG_EINVAL;
  const long long N = (long long)nx * ny * nz;
  hipLaunchKernelGGL(boundary_kernel, dim3(grid1d(N)), dim3(IG_THREADS), 0, s, xyz, nx, ny, nz, center[0], center[1],
                     center[2], order, qcoef, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_cell_fit_f64(const double* rho, const double* xyz, int nx, int ny, int nz, const int* cells, int ncells,
                    double* coef, hipStream_t s) {
  if (!rho || !xyz || !cells || !coef || ncells < 0) return IG_EINVAL;
  if (ncells == 0) return IG_OK;
  hipLaunchKernelGGL(cell_fit_kernel, dim3(grid1d(ncells)), dim3(IG_THREADS), 0, s, rho, xyz, nx, ny, nz, cells,
                     ncells, coef);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_cell_moments_f64(const double* xyz, const double* jac, int nx, int ny, int nz, const int* cells, int ncells,
                        const double* coef, double subvol, const double* center, int order, double* work,
                        double* out, hipStream_t s) {
  if (!xyz || !jac || !cells || !coef || !center || !work || !out || ncells <= 0) return IG_EINVAL;
  switch (order) {
    case 0: return cell_moments_t<0>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s);
    case 1: return cell_moments_t<1>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s);
    case 2: return cell_moments_t<2>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s);
    case 3: return cell_moments_t<3>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s);
    case 4: return cell_moments_t<4>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s);
    default: return IG_EUNSUPPORTED;
  }
}

int ig_cell_split_moments_f64(const double* rho, const double* xyz, const double* jac, int nx, int ny, int nz,
                              const int* cells, int ncells, const double* coef, double subvol, const double* center,
                              int order, double* work, double* out, hipStream_t s) {
  if (!rho || !xyz || !jac || !cells || !coef || !work || !out || ncells <= 0) return IG_EINVAL;
  if (!center) return cell_moments_t<0, 1>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, nullptr, work, out, s, rho);
  switch (order) {
    case 0: return cell_moments_t<0, 2>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s, rho);
    case 1: return cell_moments_t<1, 2>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s, rho);
    case 2: return cell_moments_t<2, 2>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s, rho);
    case 3: return cell_moments_t<3, 2>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s, rho);
    case 4: return cell_moments_t<4, 2>(xyz, jac, nx, ny, nz, cells, ncells, coef, subvol, center, work, out, s, rho);
    default: return IG_EUNSUPPORTED;
  }
}

int ig_moments_adjoint_f64(const double* w, const double* frac, const double* xyz, long long N,
                           const double* center, int order, const double* q, double* out, int accumulate,
                           hipStream_t s) {
  if (!w || !xyz || !center || !q || !out || N <= 0) return IG_EINVAL;
  switch (order) {
    case 0: return moments_adjoint_t<0>(w, frac, xyz, N, center, q, out, accumulate, s);
    case 1: return moments_adjoint_t<1>(w, frac, xyz, N, center, q, out, accumulate, s);
    case 2: return moments_adjoint_t<2>(w, frac, xyz, N, center, q, out, accumulate, s);
    case 3: return moments_adjoint_t<3>(w, frac, xyz, N, center, q, out, accumulate, s);
    case 4: return moments_adjoint_t<4>(w, frac, xyz, N, center, q, out, accumulate, s);
    default: return IG_EUNSUPPORTED;
  }
}

int ig_cell_moments_adjoint_f64(const double* xyz, const double* jac, int nx, int ny, int nz, const int* cells,
                                int ncells, double subvol, const double* center, int order, const double* q,
                                double* out, hipStream_t s) {
  if (!xyz || !jac || !cells || !center || !q || !out || ncells < 0) return IG_EINVAL;
  if (ncells == 0) return IG_OK;
  switch (order) {
    case 0: return cell_moments_adjoint_t<0>(xyz, jac, nx, ny, nz, cells, ncells, subvol, center, q, out, s);
    case 1: return cell_moments_adjoint_t<1>(xyz, jac, nx, ny, nz, cells, ncells, subvol, center, q, out, s);
    case 2: return cell_moments_adjoint_t<2>(xyz, jac, nx, ny, nz, cells, ncells, subvol, center, q, out, s);
    case 3: return cell_moments_adjoint_t<3>(xyz, jac, nx, ny, nz, cells, ncells, subvol, center, q, out, s);
    case 4: return cell_moments_adjoint_t<4>(xyz, jac, nx, ny, nz, cells, ncells, subvol, center, q, out, s);
    default: return IG_EUNSUPPORTED;
  }
}

}  // extern "C"
