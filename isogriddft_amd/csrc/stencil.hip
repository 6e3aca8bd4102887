// Curvilinear kinetic / Laplacian stencils for the monitor grid and the legacy grid.
//
// Monitor operator (ops/kinetic.py:212-358 of the reference):
//     Lap psi = (1/J) sum_a D_a ( sum_b c^{ab} D_b psi ),   c^{ab} = J g^{ab},
// D the centred first difference on the uniform logical axis with either a zero
// ghost ("trial_fix", CL = 0) or NumPy's edge_order=2 one-sided ends ("production",
// CL = 1).  TR = 1 gives the exact Euclidean transpose
//     Lap^T x = sum_b D_b^T ( sum_a c^{ab} D_a^T (J^-1 x) ).
// The flux coefficients arrive pre-scaled, c^{ab} / (4 h_a h_b), so the differences D
// are applied un-scaled.  One fused pass: a workgroup owns a TY x TZ tile of the (y, z)
// plane and NB orbitals, and marches
// along x; a six-plane LDS ring holds the input planes with a y/z halo, fluxes of
// the current plane are staged in LDS, the x fluxes stay in registers.  The output
// is   out = scale * Lap psi + V * psi   (V optional), optionally zeroed on the
// boundary nodes (the Poisson interior restriction).
#include "common.h"

// No implicit FMA contraction in this file: every fused multiply-add is written
// explicitly, so the flux of a point is bit-identical whether it is evaluated as a
// tile point or as a halo point of the neighbouring (possibly overlapping) tile.
#pragma clang fp contract(off)

namespace {

template <typename T>
__device__ __forceinline__ T flux3(T c0, T g0, T c1, T g1, T c2, T g2) {
  return fma(c0, g0, fma(c1, g1, c2 * g2));
}

template <typename T>
struct MonParams {
  const T* in;
  T* out;
  const T* coef;  // 6*N: c^{ab} / (4 h_a h_b) for ab = xx yy zz xy xz yz
  const T* invj;  // N
  const T* veff;  // N or null
  long long ld;   // elements between consecutive orbitals (in and out)
  int nblk, nx, ny, nz;
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
  if (i == 0) return fma(T(-3), f(0), fma(T(4), f(1), -f(2)));
  return fma(T(3), f(0), fma(T(-4), f(-1), f(-2)));
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
    if (c != 0) s = fma(T(c), f(d), s);
  }
  return s;
}

template <int CL, bool TR, typename T, typename F>
__device__ __forceinline__ T dop(int i, int n, F f) {
  if (TR) return dtr<CL, T>(i, n, f);
  return dfwd<CL, T>(i, n, f);
}

// Buffer-resource access (MUBUF, raw, stride 0): the hardware range check returns 0
// for loads and drops stores at byte offsets >= num_records.  Every field here is
// smaller than BAD bytes, so BAD (or BAD + a valid offset) marks an inactive access:
// zero ghosts, inactive tile points and suppressed stores need no branch.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr unsigned BAD = 0x40000000u;

template <typename T>
__device__ __forceinline__ rsrc_t mkbuf(const T* ptr, long long n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(ptr), (short)0, (int)(n * (long long)sizeof(T)),
                                           0x00020000);
}
__device__ __forceinline__ float bld(rsrc_t r, unsigned off, float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ double bld(rsrc_t r, unsigned off, double) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bst(rsrc_t r, unsigned off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
}
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void bst(rsrc_t r, unsigned off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)off, 0, 0);
}

template <typename T, int TY, int TZ, int CL, bool TR, int NB, int PD>
__global__ __launch_bounds__(IG_THREADS) void monitor_stencil_kernel(MonParams<T> p) {
  constexpr int HF = TR ? 2 : 1;  // flux halo
  constexpr int HP = 2 * HF;      // input halo
  constexpr int PY = TY + 2 * HP;
  constexpr int PZ = TZ + 2 * HP;
  constexpr int PLANE = PY * PZ;
  constexpr int NSLOT = 6;             // ring of input planes; also the unroll period
  constexpr int RG = IG_THREADS / TZ;  // row groups
  constexpr int RPT = TY / RG;         // rows per thread
  constexpr int NLD = (PLANE + IG_THREADS - 1) / IG_THREADS;  // ring elements per thread
  constexpr int NHY = 2 * HF * TZ;     // y-halo flux points
  constexpr int NHZ = 2 * HF * TY;     // z-halo flux points
  constexpr int NHI = (NHY + NHZ + IG_THREADS - 1) / IG_THREADS;
  constexpr int FYS = (TY + 2 * HF) * TZ;
  constexpr int FZS = TY * (TZ + 2 * HF);
  constexpr unsigned S = sizeof(T);
  static_assert(RPT * RG == TY, "tile rows must divide");
  static_assert(PD >= 1 && NSLOT % PD == 0, "prefetch depth must divide the ring period");

  constexpr int PSTR = NLD * IG_THREADS;  // ring slot stride: every thread owns NLD entries
  __shared__ T s_in[NB][NSLOT * PSTR];
  // y fluxes [FYS], z fluxes [FZS] and one spare entry that absorbs the stores of
  // threads without a halo point, so no store is predicated.
  __shared__ T s_fl[NB][FYS + FZS + 1];

  // XCD-aware block -> (work unit, orbital group).  Workgroup b runs on XCD b % 8, so
  // each XCD takes one contiguous range of work units (z tiles fastest, then y tiles,
  // then x chunks): tiles that share halo rows and cache lines run together on one L2.
  // Within an XCD the orbital groups of a unit are consecutive, so they also share
  // that unit's geometry stream from L2.
  const int ngroups = (p.nblk + NB - 1) / NB;
  const int b = blockIdx.x;
  const int xcd = b & 7;
  const int r = b >> 3;
  const int grp = r % ngroups;
  const int upx = (p.units + 7) / 8;
  const int u = xcd * upx + r / ngroups;
  if (u >= p.units) return;
  const int tzi = u % p.ntz;
  const int tyi = (u / p.ntz) % p.nty;
  const int xci = u / (p.ntz * p.nty);

  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const int nyz = ny * nz;
  const unsigned plane_bytes = (unsigned)nyz * S;
  const int j0 = min(tyi * TY, max(ny - TY, 0));
  const int k0 = min(tzi * TZ, max(nz - TZ, 0));
  const int x0 = xci * p.xc;
  const int x1 = min(x0 + p.xc, nx);
  const long long N = (long long)nx * nyz;

  rsrc_t rin[NB], rout[NB];
  bool orb_ok[NB];
#pragma unroll
  for (int n = 0; n < NB; ++n) {
    const int orb = grp * NB + n;
    orb_ok[n] = orb < p.nblk;
    const int oc = orb_ok[n] ? orb : p.nblk - 1;
    rin[n] = mkbuf(p.in + (long long)oc * p.ld, N);
    rout[n] = mkbuf(p.out + (long long)oc * p.ld, N);
  }
  const rsrc_t rxx = mkbuf(p.coef, N);
  const rsrc_t ryy = mkbuf(p.coef + N, N);
  const rsrc_t rzz = mkbuf(p.coef + 2 * N, N);
  const rsrc_t rxy = mkbuf(p.coef + 3 * N, N);
  const rsrc_t rxz = mkbuf(p.coef + 4 * N, N);
  const rsrc_t ryz = mkbuf(p.coef + 5 * N, N);
  const rsrc_t rij = mkbuf(p.invj, N);
  const rsrc_t rv = mkbuf(p.veff ? p.veff : p.invj, p.veff ? N : 0);  // no V: every load returns 0

  const int tid = threadIdx.x;
  const int tz = tid % TZ;
  const int tyb = tid / TZ;

  // ---- per-thread constants, computed once: in-plane byte offsets (BAD outside).
  unsigned poff[NLD];
#pragma unroll
  for (int rr = 0; rr < NLD; ++rr) {
    const int idx = tid + rr * IG_THREADS;
    const int ly = idx / PZ;
    const int lz = idx - ly * PZ;
    const int j = j0 + ly - HP;
    const int k = k0 + lz - HP;
    poff[rr] = (idx < PLANE && j >= 0 && j < ny && k >= 0 && k < nz) ? (unsigned)(j * nz + k) * S : BAD;
  }
  unsigned toff[RPT];
  unsigned tbnd = 0;
#pragma unroll
  for (int rr = 0; rr < RPT; ++rr) {
    const int j = j0 + tyb + rr * RG;
    const int k = k0 + tz;
    toff[rr] = (j < ny && k < nz) ? (unsigned)(j * nz + k) * S : BAD;
    if (j == 0 || j == ny - 1 || k == 0 || k == nz - 1) tbnd |= 1u << rr;
  }
  unsigned hoff[NHI];
  int hkind[NHI];
  int hj[NHI], hk[NHI], hlds[NHI], hst[NHI];
#pragma unroll
  for (int hi = 0; hi < NHI; ++hi) {
    const int idx = tid + hi * IG_THREADS;
    int j = j0, k = k0;
    bool ok = false;
    hkind[hi] = 0;
    hlds[hi] = HP * PZ + HP;
    hst[hi] = FYS + FZS;
    if (idx < NHY) {
      const int hr = idx / TZ;
      const int cz = idx - hr * TZ;
      const int fr = hr < HF ? hr : TY + hr;
      j = j0 + fr - HF;
      k = k0 + cz;
      hkind[hi] = 1;
      hlds[hi] = (fr - HF + HP) * PZ + cz + HP;
      hst[hi] = fr * TZ + cz;
      ok = j >= 0 && j < ny && k < nz;
    } else if (idx < NHY + NHZ) {
      const int i2 = idx - NHY;
      const int hcn = i2 / TY;
      const int ry = i2 - hcn * TY;
      const int fc = hcn < HF ? hcn : TZ + hcn;
      j = j0 + ry;
      k = k0 + fc - HF;
      hkind[hi] = 2;
      hlds[hi] = (ry + HP) * PZ + fc - HF + HP;
      hst[hi] = FYS + ry * (TZ + 2 * HF) + fc;
      ok = j < ny && k >= 0 && k < nz;
    }
    hj[hi] = j;
    hk[hi] = k;
    hoff[hi] = ok ? (unsigned)(j * nz + k) * S : BAD;
  }

  // ---- software pipeline registers, PD planes deep: each set is refilled with the
  // plane PD steps ahead right after its last use.
  T pf[PD][NB][NLD];
  T gc[PD][RPT][6];
  T hc[PD][NHI][3];
  T ic[PD][RPT];
  T vc[PD][RPT];
  T xcur[PD][NB][RPT];
  const T tag = T(0);

  auto pbytes = [&](int q) { return (q >= 0 && q < nx) ? (unsigned)q * plane_bytes : BAD; };
  auto fetch_plane = [&](int q, int w) {
    const unsigned pb = pbytes(q);
#pragma unroll
    for (int rr = 0; rr < NLD; ++rr) {
      const unsigned o = pb + poff[rr];
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        pf[w][n][rr] = bld(rin[n], o, tag);
        if (TR) pf[w][n][rr] *= bld(rij, o, tag);
      }
    }
  };
  auto store_plane = [&](int slot, int w) {
#pragma unroll
    for (int rr = 0; rr < NLD; ++rr)
#pragma unroll
      for (int n = 0; n < NB; ++n)
        s_in[n][slot * PSTR + tid + rr * IG_THREADS] = pf[w][n][rr];
  };
  auto fetch_flux_geom = [&](int q, int w) {
    const unsigned pb = pbytes(q);
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const unsigned o = pb + toff[rr];
      gc[w][rr][0] = bld(rxx, o, tag); gc[w][rr][1] = bld(ryy, o, tag); gc[w][rr][2] = bld(rzz, o, tag);
      gc[w][rr][3] = bld(rxy, o, tag); gc[w][rr][4] = bld(rxz, o, tag); gc[w][rr][5] = bld(ryz, o, tag);
    }
#pragma unroll
    for (int hi = 0; hi < NHI; ++hi) {
      const unsigned o = pb + hoff[hi];
      const T a_xy = bld(rxy, o, tag), a_yy = bld(ryy, o, tag), a_yz = bld(ryz, o, tag);
      const T a_xz = bld(rxz, o, tag), a_zz = bld(rzz, o, tag);
      const bool yrow = hkind[hi] == 1;
      hc[w][hi][0] = yrow ? a_xy : a_xz;
      hc[w][hi][1] = yrow ? a_yy : a_yz;
      hc[w][hi][2] = yrow ? a_yz : a_zz;
    }
  };
  auto fetch_out_geom = [&](int o, int w) {
    const unsigned pb = pbytes(o);
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const unsigned off = pb + toff[rr];
      ic[w][rr] = bld(rij, off, tag);
      vc[w][rr] = bld(rv, off, tag);
#pragma unroll
      for (int n = 0; n < NB; ++n)
        if (TR) xcur[w][n][rr] = bld(rin[n], off, tag);
    }
  };

  // F_x and the y/z divergence of the last planes, indexed by (plane - sb) mod 6.
  T fxr[NB][NSLOT][RPT];
  T dyr[NB][NSLOT][RPT];
#pragma unroll
  for (int n = 0; n < NB; ++n)
#pragma unroll
    for (int a = 0; a < NSLOT; ++a)
#pragma unroll
      for (int rr = 0; rr < RPT; ++rr) fxr[n][a][rr] = dyr[n][a][rr] = T(0);

  const int sb = x0 - 2;
  const int se = x1 + 2;
  // LDS ring slot of plane q: (q - sb + 2) mod 6.  Register queue slots: psi plane q in
  // (q - sb - 2) mod PD, flux geometry of plane q in (q - sb) mod PD, output geometry
  // of plane q in (q - sb + 2) mod PD; 6 % PD == 0 keeps them constant when unrolled.
  for (int q = sb - 2; q < sb + 2; ++q) {
    fetch_plane(q, 0);
    store_plane(q - sb + 2, 0);
  }
#pragma unroll
  for (int d = 0; d < PD; ++d) {
    fetch_plane(sb + 2 + d, d);
    fetch_flux_geom(sb + d, d);
    fetch_out_geom(sb - 2 + d, d);
  }
  const T sc = p.scale;

  for (int s0 = sb; s0 < se; s0 += NSLOT) {
    // Unrolled by the ring period: every slot and register index below is a constant.
#pragma unroll
    for (int U = 0; U < NSLOT; ++U) {
      const int s = s0 + U;
      const int W = U % PD;  // a constant once the loop is unrolled
      // Phase A: commit plane s+2, request plane s+2+PD.
      store_plane((U + 4) % NSLOT, W);
      fetch_plane(s + 2 + PD, W);
      __syncthreads();

      // Phase B: fluxes of plane s (coefficients are zero outside the domain).
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        const T* ring = s_in[n];
        auto grad = [&](int j, int k, int off, T& gx, T& gy, T& gz) {
          const T* pl = ring + ((U + 2) % NSLOT) * PSTR;
          gx = dop<CL, TR, T>(s, nx, [&](int d) { return ring[((U + 2 + d + NSLOT) % NSLOT) * PSTR + off]; });
          gy = dop<CL, TR, T>(j, ny, [&](int d) { return pl[off + d * PZ]; });
          gz = dop<CL, TR, T>(k, nz, [&](int d) { return pl[off + d]; });
        };
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
          const int ty = tyb + rr * RG;
          T gx, gy, gz;
          grad(j0 + ty, k0 + tz, (ty + HP) * PZ + tz + HP, gx, gy, gz);
          fxr[n][U][rr] = flux3(gc[W][rr][0], gx, gc[W][rr][3], gy, gc[W][rr][4], gz);
          s_fl[n][(ty + HF) * TZ + tz] = flux3(gc[W][rr][3], gx, gc[W][rr][1], gy, gc[W][rr][5], gz);
          s_fl[n][FYS + ty * (TZ + 2 * HF) + tz + HF] = flux3(gc[W][rr][4], gx, gc[W][rr][5], gy, gc[W][rr][2], gz);
        }
#pragma unroll
        for (int hi = 0; hi < NHI; ++hi) {
          T gx, gy, gz;
          grad(hj[hi], hk[hi], hlds[hi], gx, gy, gz);
          const T F = flux3(hc[W][hi][0], gx, hc[W][hi][1], gy, hc[W][hi][2], gz);
          s_fl[n][hst[hi]] = F;
        }
      }
      fetch_flux_geom(s + PD, W);
      __syncthreads();

      // Phase C: y/z divergence of plane s, then output plane o = s - 2.
      const int o = s - 2;
      const bool emit = (o >= x0 && o < x1);
      const int oc = min(max(o, 0), nx - 1);
      const bool obnd = (o == 0 || o == nx - 1);
      const unsigned ob = emit ? (unsigned)o * plane_bytes : BAD;
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        const unsigned obn = orb_ok[n] ? ob : BAD;
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
          const int ty = tyb + rr * RG;
          const int fy0 = (ty + HF) * TZ + tz;
          const int fz0 = ty * (TZ + 2 * HF) + tz + HF;
          dyr[n][U][rr] = dop<CL, TR, T>(j0 + ty, ny, [&](int dd) { return s_fl[n][fy0 + dd * TZ]; }) +
                          dop<CL, TR, T>(k0 + tz, nz, [&](int dd) { return s_fl[n][FYS + fz0 + dd]; });
          const T dx = dop<CL, TR, T>(oc, nx, [&](int dd) { return fxr[n][(U - 2 + dd + NSLOT) % NSLOT][rr]; });
          const T lap = dx + dyr[n][(U - 2 + NSLOT) % NSLOT][rr];
          T v;
          if (TR) {
            v = fma(vc[W][rr], xcur[W][n][rr], sc * lap);
          } else {
            const T psi = s_in[n][U * PSTR + (ty + HP) * PZ + tz + HP];
            v = fma(vc[W][rr], psi, sc * (lap * ic[W][rr]));
          }
          if (p.interior_only && (obnd || ((tbnd >> rr) & 1u))) v = T(0);
          bst(rout[n], obn + toff[rr], v);
        }
      }
      fetch_out_geom(s - 2 + PD, W);
    }
  }
}

template <typename T, int TY, int TZ, int CL, bool TR, int NB = 1, int PD = 1>
int launch_monitor_t(const MonParams<T>& base, hipStream_t stream) {
  MonParams<T> p = base;
  // buffer-resource addressing: every field must stay below 1 GiB (the BAD offset)
  if ((long long)p.nx * p.ny * p.nz * (long long)sizeof(T) >= (1LL << 30)) return IG_EUNSUPPORTED;
  p.nty = (p.ny + TY - 1) / TY;
  p.ntz = (p.nz + TZ - 1) / TZ;
  if (p.xc <= 0) {
    // Auto x chunking.  Blocks are latency bound, so the launch should fill every CU's
    // resident-block capacity once, while each chunk pays ~5 extra plane steps (4 halo
    // planes + the prologue).  Model time = rounds * (xc + 5) with rounds =
    // ceil(blocks / capacity) and take the best chunk count.
    static int occ = -1;
    if (occ < 0) {
      int n = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, monitor_stencil_kernel<T, TY, TZ, CL, TR, NB, PD>,
                                                      IG_THREADS, 0) != hipSuccess || n < 1)
        n = 1;
      occ = n;
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    const long long capacity = (long long)occ * cus;
    const long long tiles = (long long)p.nty * p.ntz;
    const long long groups = (p.nblk + NB - 1) / NB;
    const int maxc = p.nx / 8 > 1 ? p.nx / 8 : 1;
    long long best = -1;
    int best_c = 1;
    for (int c = 1; c <= maxc; ++c) {
      const int xc = (p.nx + c - 1) / c;
      const long long blocks = 8LL * ((tiles * c + 7) / 8) * groups;
      const long long rounds = (blocks + capacity - 1) / capacity;
      const long long cost = rounds * (xc + 5);
      if (best < 0 || cost < best) { best = cost; best_c = c; }
    }
    p.xc = (p.nx + best_c - 1) / best_c;
  }
  if (p.xc > p.nx) p.xc = p.nx;
  p.nxc = (p.nx + p.xc - 1) / p.xc;
  p.units = p.nty * p.ntz * p.nxc;
  const long long blocks = 8LL * ((p.nblk + NB - 1) / NB) * ((p.units + 7) / 8);
  if (blocks <= 0 || blocks > 0x7fffffffLL) return IG_EINVAL;
  hipLaunchKernelGGL((monitor_stencil_kernel<T, TY, TZ, CL, TR, NB, PD>), dim3((unsigned)blocks),
                     dim3(IG_THREADS), 0, stream, p);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

#define IG_MON_CASE(TY_, TZ_, NB_, PD_)                                                            \
  if (ty == TY_ && tz == TZ_ && nb == NB_ && pd == PD_)                                            \
    return closure ? launch_monitor_t<T, TY_, TZ_, 1, false, NB_, PD_>(p, s)                       \
                   : launch_monitor_t<T, TY_, TZ_, 0, false, NB_, PD_>(p, s);

// tile: 0 = automatic; 32 / 64 = z tile width with the default rows; otherwise
// PD*100000 + NB*10000 + TZ*100 + TY (e.g. 223216: two orbitals per workgroup, 16 x 32
// tiles, loads issued two planes ahead).
template <typename T>
int launch_monitor(const MonParams<T>& p, int closure, int transpose, int tile, hipStream_t s) {
  if (transpose) return closure ? launch_monitor_t<T, 8, 32, 1, true>(p, s) : launch_monitor_t<T, 8, 32, 0, true>(p, s);
  if (tile < 0) return IG_EINVAL;
  int pd = tile / 100000;
  int nb = (tile / 10000) % 10;
  const int t2 = tile % 10000;
  int tz = t2 >= 100 ? t2 / 100 : t2;
  int ty = t2 >= 100 ? t2 % 100 : 0;
  if (tile == 0) {
    // measured defaults (96^3 and 224^3 blocks): FP32 four orbitals of 8 x 32 tiles,
    // FP64 two orbitals of 8 x 32 tiles with loads two planes ahead
    tz = 32;
    ty = 8;
    if (sizeof(T) == 8) {
      nb = p.nblk >= 2 ? 2 : 1;
      pd = nb == 2 ? 2 : 1;
    } else {
      nb = p.nblk >= 4 ? 4 : (p.nblk >= 2 ? 2 : 1);
      pd = 1;
    }
  }
  if (tz == 0) tz = 32;
  if (ty == 0) ty = tz == 64 ? 8 : 16;
  if (nb == 0) nb = (p.nblk >= 2 && tz == 32) ? 2 : 1;
  if (nb > p.nblk) nb = (p.nblk >= 2 && tz == 32) ? 2 : 1;
  if (pd == 0) pd = 1;
  if constexpr (sizeof(T) == 8) {
    IG_MON_CASE(8, 64, 1, 1)
    IG_MON_CASE(8, 32, 1, 1)
    IG_MON_CASE(8, 32, 1, 2)
    IG_MON_CASE(16, 32, 1, 1)
    IG_MON_CASE(16, 32, 1, 2)
    IG_MON_CASE(8, 32, 2, 1)
    IG_MON_CASE(8, 32, 2, 2)
    IG_MON_CASE(8, 32, 2, 3)
    IG_MON_CASE(16, 32, 2, 1)
    IG_MON_CASE(16, 32, 2, 2)
  } else {
    IG_MON_CASE(8, 64, 1, 1)
    IG_MON_CASE(8, 64, 1, 2)
    IG_MON_CASE(8, 32, 1, 1)
    IG_MON_CASE(16, 32, 1, 1)
    IG_MON_CASE(16, 32, 1, 2)
    IG_MON_CASE(16, 32, 1, 3)
    IG_MON_CASE(8, 32, 2, 1)
    IG_MON_CASE(8, 32, 2, 2)
    IG_MON_CASE(8, 32, 2, 3)
    IG_MON_CASE(16, 32, 2, 1)
    IG_MON_CASE(16, 32, 2, 2)
    IG_MON_CASE(16, 32, 2, 3)
    IG_MON_CASE(8, 32, 4, 1)
    IG_MON_CASE(8, 32, 4, 2)
    IG_MON_CASE(16, 32, 4, 1)
  }
  return IG_EUNSUPPORTED;
}
#undef IG_MON_CASE

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
                  const T* coef, const T* invj, const T* veff,
                  double scale, int closure, int transpose, int interior_only, int xchunk,
                  int tile, hipStream_t stream) {
  if (!in || !out || !coef || !invj || nblk <= 0 || nx < 3 || ny < 3 || nz < 3) return IG_EINVAL;
  if (ld < (long long)nx * ny * nz) return IG_EINVAL;
  if (closure < 0 || closure > 1) return IG_EINVAL;
  if (tile < 0 || (tile != 0 && tile != 32 && tile != 64 && tile < 100)) return IG_EINVAL;
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
                           int nz, const double* coef, const double* invj, const double* veff,
                           double scale, int closure, int transpose, int interior_only,
                           int xchunk, int tile, hipStream_t stream) {
  return monitor_entry<double>(in, out, nblk, ld, nx, ny, nz, coef, invj, veff, scale,
                               closure, transpose, interior_only, xchunk, tile, stream);
}

int ig_monitor_stencil_f32(const float* in, float* out, int nblk, long long ld, int nx, int ny,
                           int nz, const float* coef, const float* invj, const float* veff,
                           double scale, int closure, int transpose, int interior_only,
                           int xchunk, int tile, hipStream_t stream) {
  return monitor_entry<float>(in, out, nblk, ld, nx, ny, nz, coef, invj, veff, scale,
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
