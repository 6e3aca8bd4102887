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
  int order;  // block order of the full-z kernel: 0 units of one XCD consecutive, 1 y sweep,
              // 2 y sweep in bands of eight pairs
  int ug;     // block order of the 32-point-tile kernel: 0 units of one XCD consecutive,
              // g > 0 items of g consecutive units dealt round robin to the XCDs
  const T* dotw;     // full-z kernel with row dots: weights w of sum_n w[n] in[m][n] out[m][n]
  double* dpart;     // its partial sums, [nblk][units]
  long long dparts;  // doubles available at dpart
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
// Output stores with the non-temporal hint (cache policy bit 1, "nt" on gfx950): the
// operator output is not re-read by the launch, so it should not displace the geometry and
// halo data in L2 / Infinity Cache (224^3 k = 48 FP32 block: 2237 -> 2296 GB/s; 96^3
// production closure: 816 -> 828 GB/s; results identical).
constexpr int STORE_NT = 2;
__device__ __forceinline__ void bst(rsrc_t r, unsigned off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, STORE_NT);
}
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void bst(rsrc_t r, unsigned off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)off, 0, STORE_NT);
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
    // Launches of several residency rounds (many orbital groups) run faster with chunks of
    // ~56 planes than the model's long ones: 224^3, 48 FP32 orbitals, 4 per workgroup,
    // 1.90 ms at 4 chunks of 56 against 2.04 ms at 1 chunk (measured; 3 chunks 2.04 ms,
    // 48- and 64-plane chunks 2.13 / 2.21 ms).
    const long long blocks1 = 8LL * ((tiles * best_c + 7) / 8) * groups;
    if (blocks1 > capacity && p.nx >= 112) {
      const int c56 = (p.nx + 28) / 56;
      if (c56 > best_c) best_c = c56;
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

// ---------------------------------------------------------------------------------
// z-vectorised forward trial_fix stencil.  Each lane owns V = 16 / sizeof(T)
// consecutive z points (one 128-bit access), a tile row of TZ = 16 V points is 16
// lanes (one DPP row), the tile is 16 rows.  Workgroup = 5 waves: waves 0-3 compute
// the 16 x TZ interior points; wave 4 computes the halo fluxes the interior needs (F_y
// on rows -1 and 16 with lanes 0-31, F_z at z = k0-1 and k0+TZ with lanes 32-63).  Per
// x step there is one barrier: the input ring has 4 planes (s-1 .. s+2 in use, plane
// s+3 written after the barrier into the slot of s-1), the staged F_y / F_z halo are
// double buffered, z neighbours of F_z come from the adjacent lane by DPP.
// Output plane o = s - 1 is written at step s.
template <typename T>
struct ZVec;
template <>
struct ZVec<float> {
  typedef float t __attribute__((ext_vector_type(4)));
  static constexpr int V = 4;
};
template <>
struct ZVec<double> {
  typedef double t __attribute__((ext_vector_type(2)));
  static constexpr int V = 2;
};
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ typename ZVec<T>::t zld(rsrc_t r, unsigned off) {
  return __builtin_bit_cast(typename ZVec<T>::t, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}
// Output stores carry the non-temporal hint (cache policy bit 1, "nt" on gfx950): H psi is
// written once and not read again by this launch, so it should not displace the geometry
// and the psi halo rows that the neighbouring tiles and orbital groups re-read from the
// caches (c2 FP32 block: 51.7 -> 49.6 us, identical results).  Non-temporal psi loads
// were measured slower (56.6 us): the halo rows need the cache.
template <typename T>
__device__ __forceinline__ void zst(rsrc_t r, unsigned off, typename ZVec<T>::t v) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, (int)off, 0, STORE_NT);
}
// Value of the lane one to the left (CTRL = 0x111, row_shr:1) or right (0x101,
// row_shl:1) inside its 16-lane row; lanes without a source get 0 (replaced by the
// staged halo value by the caller).
template <int CTRL>
__device__ __forceinline__ float zdpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double zdpp(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffull), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}

constexpr int ZV_NS = 4;

// Tile shape per precision: FP32 rows of 8 lanes x 4 points (TZ = 32), 24 rows, 4 waves
// (3 interior + 1 halo); FP64 rows of 16 lanes x 2 points (TZ = 32), 16 rows, 5 waves.
template <typename T>
struct ZShape;
template <>
struct ZShape<float> {
  static constexpr int LZ = 8, TY = 24, NT = 256;
};
template <>
struct ZShape<double> {
  static constexpr int LZ = 16, TY = 16, NT = 320;
};

template <typename T, int NB, int WPE>
__global__ __launch_bounds__(ZShape<T>::NT) __attribute__((amdgpu_waves_per_eu(WPE)))
void monitor_zv_kernel(MonParams<T> p) {
  using VT = typename ZVec<T>::t;
  constexpr int V = ZVec<T>::V;
  constexpr int LZ = ZShape<T>::LZ, TY = ZShape<T>::TY, NT = ZShape<T>::NT, TZ = LZ * V;
  constexpr int PY = TY + 4;          // ring rows: tile rows with a 2-row halo
  // A ring plane keeps the TZ-point rows packed back to back (a 16-lane LDS pass reads two
  // consecutive 128-byte rows: conflict free) and the 2-point z halo of every row in a
  // separate array of 4 values per row (k0-2, k0-1, k0+TZ, k0+TZ+1).
  constexpr int PLANE = PY * TZ + V;  // + a spare vector absorbing stores of threads without a ring element
  constexpr int HPL = PY * 4 + 4;     // halo values per plane + a spare
  constexpr int FYR = TY + 2;         // F_y rows -1 .. TY
  constexpr unsigned S = sizeof(T);
  static_assert(TY * LZ + 2 * LZ + 2 * TY == NT, "interior, y-halo and z-halo units fill the workgroup");
  static_assert(PY * LZ <= NT && PY * 4 <= NT, "one ring vector and one ring halo scalar per thread");
  static_assert(16 % LZ == 0, "tile rows must tile the 16-lane DPP rows");

  __shared__ __attribute__((aligned(16))) T s_ring[NB][ZV_NS * PLANE];
  __shared__ T s_rh[NB][ZV_NS * HPL];
  __shared__ __attribute__((aligned(16))) T s_fy[2][NB][FYR * TZ];
  __shared__ T s_fzh[2][NB][2 * TY];

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
  const unsigned plane_bytes = (unsigned)(ny * nz) * S;
  const int j0 = min(tyi * TY, ny - TY);
  const int k0 = min(tzi * TZ, nz - TZ);
  const int x0 = xci * p.xc;
  const int x1 = min(x0 + p.xc, nx);
  const long long N = (long long)nx * ny * nz;

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
  const rsrc_t rxx = mkbuf(p.coef, N), ryy = mkbuf(p.coef + N, N), rzz = mkbuf(p.coef + 2 * N, N);
  const rsrc_t rxy = mkbuf(p.coef + 3 * N, N), rxz = mkbuf(p.coef + 4 * N, N), ryz = mkbuf(p.coef + 5 * N, N);
  const rsrc_t rij = mkbuf(p.invj, N);
  const rsrc_t rv = mkbuf(p.veff ? p.veff : p.invj, p.veff ? N : 0);

  const int tid = threadIdx.x;
  // ---- ring loading role (every thread): at most one vector and one halo scalar per plane
  unsigned ld_off = BAD;
  int ld_lds = PY * TZ;
  if (tid < PY * LZ) {
    const int lr = tid / LZ, llz = tid % LZ;
    const int lj = j0 - 2 + lr;
    ld_lds = lr * TZ + V * llz;
    if (lj >= 0 && lj < ny) ld_off = (unsigned)(lj * nz + k0 + V * llz) * S;
  }
  unsigned hs_off = BAD;
  int hs_lds = PY * 4;
  if (tid < PY * 4) {
    const int hr = tid >> 2, h = tid & 3;
    const int k = h < 2 ? k0 - 2 + h : k0 + TZ + (h - 2);
    const int j = j0 - 2 + hr;
    hs_lds = tid;
    if (j >= 0 && j < ny && k >= 0 && k < nz) hs_off = (unsigned)(j * nz + k) * S;
  }

  // ---- compute role
  const bool is_int = tid < TY * LZ;
  const bool is_yh = !is_int && tid < FYR * LZ;
  const bool is_zh = tid >= FYR * LZ;
  int row, lz, side = 0;
  if (is_int) {
    row = tid / LZ;
    lz = tid % LZ;
  } else if (is_yh) {
    const int q = tid - TY * LZ;
    row = (q / LZ) ? TY : -1;
    lz = q % LZ;
  } else {
    const int q = tid - FYR * LZ;
    row = q >> 1;
    side = q & 1;
    lz = 0;
  }
  const int gj = j0 + row;
  const bool jok = gj >= 0 && gj < ny;
  const unsigned voff = (!is_zh && jok) ? (unsigned)(gj * nz + k0 + V * lz) * S : BAD;
  const unsigned ooff = is_int ? voff : BAD;
  const int cidx = (row + 2) * TZ + V * lz;
  const int hrow = (row + 2) * 4;  // halo values of the unit's ring row
  const int zl_idx = lz == 0 ? hrow + 1 : 0;
  const int zr_idx = lz == LZ - 1 ? hrow + 2 : 0;
  const int kz = side ? k0 + TZ : k0 - 1;
  const unsigned zoff = (is_zh && jok && kz >= 0 && kz < nz) ? (unsigned)(gj * nz + kz) * S : BAD;
  // z-halo point: its value and y neighbours live in the halo array, its inner z
  // neighbour in the packed row
  const int zh_c = hrow + 1 + side;
  const int zh_out = hrow + (side ? 3 : 0);
  const int zh_in = (row + 2) * TZ + (side ? TZ - 1 : 0);
  unsigned bmask = 0;
#pragma unroll
  for (int e = 0; e < V; ++e) {
    const int k = k0 + V * lz + e;
    if (gj == 0 || gj == ny - 1 || k == 0 || k == nz - 1) bmask |= 1u << e;
  }

  auto pbytes = [&](int q) { return (q >= 0 && q < nx) ? (unsigned)q * plane_bytes : BAD; };

  VT stg[NB];
  T stgh[NB];
  auto fetch_plane = [&](int q) {
    const unsigned pb = pbytes(q);
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      stg[n] = zld<T>(rin[n], pb + ld_off);
      stgh[n] = bld(rin[n], pb + hs_off, T(0));
    }
  };
  auto store_plane = [&](int slot) {
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      *reinterpret_cast<VT*>(&s_ring[n][slot * PLANE + ld_lds]) = stg[n];
      s_rh[n][slot * HPL + hs_lds] = stgh[n];
    }
  };
  // flux coefficients of the unit, two planes in flight (set = parity of plane - sb)
  VT gxx[2], gyy[2], gzz[2], gxy[2], gxz[2], gyz[2];
  T zxz[2], zyz[2], zzz[2];
  VT oij, ov;  // 1/J and V at the next output plane
  auto fetch_geom = [&](int q, int w) {
    const unsigned pb = pbytes(q);
    gxx[w] = zld<T>(rxx, pb + voff);
    gyy[w] = zld<T>(ryy, pb + voff);
    gzz[w] = zld<T>(rzz, pb + voff);
    gxy[w] = zld<T>(rxy, pb + voff);
    gxz[w] = zld<T>(rxz, pb + voff);
    gyz[w] = zld<T>(ryz, pb + voff);
    zxz[w] = bld(rxz, pb + zoff, T(0));
    zyz[w] = bld(ryz, pb + zoff, T(0));
    zzz[w] = bld(rzz, pb + zoff, T(0));
  };
  auto fetch_out = [&](int o) {
    const unsigned pb = pbytes(o);
    oij = zld<T>(rij, pb + ooff);
    ov = zld<T>(rv, pb + ooff);
  };

  VT fx[NB][ZV_NS];   // F_x of the last planes, indexed by (plane - sb) mod 4
  VT fz[NB];          // F_z of plane s (interior)
  VT pc[NB][2];       // psi at the owned points of planes s and s - 1 (parity of s - sb)
  VT dyz[NB][2];      // y/z divergence of planes s and s - 1
#pragma unroll
  for (int n = 0; n < NB; ++n) {
#pragma unroll
    for (int a = 0; a < ZV_NS; ++a) fx[n][a] = VT(0);
    fz[n] = pc[n][0] = pc[n][1] = dyz[n][0] = dyz[n][1] = VT(0);
  }

  const int sb = x0 - 1;  // steps sb .. x1; output o = s - 1 in [x0, x1)
  // ring slot of plane q: (q - sb + 1) mod 4
  for (int q = sb - 1; q <= sb + 2; ++q) {
    fetch_plane(q);
    store_plane(q - sb + 1);
  }
  fetch_plane(sb + 3);
  fetch_geom(sb, 0);
  fetch_geom(sb + 1, 1);
  fetch_out(sb - 1);
  const T sc = p.scale;
  __syncthreads();

  for (int s0 = sb; s0 <= x1; s0 += ZV_NS) {
#pragma unroll
    for (int U = 0; U < ZV_NS; ++U) {
      const int s = s0 + U;
      if (s > x1) break;
      const int sm = U % ZV_NS, s0l = (U + 1) % ZV_NS, sp = (U + 2) % ZV_NS;  // slots of s-1, s, s+1
      const int fb = U & 1;                                                   // staging buffer / parity
      // ---------------- phase B: fluxes of plane s
      auto vgrad = [&](int n, VT& c, VT& gx, VT& gy, VT& gz) {
        const T* pl = &s_ring[n][s0l * PLANE];
        c = *reinterpret_cast<const VT*>(&pl[cidx]);
        const VT ym = *reinterpret_cast<const VT*>(&pl[cidx - TZ]);
        const VT yp = *reinterpret_cast<const VT*>(&pl[cidx + TZ]);
        const VT xm = *reinterpret_cast<const VT*>(&s_ring[n][sm * PLANE + cidx]);
        const VT xp = *reinterpret_cast<const VT*>(&s_ring[n][sp * PLANE + cidx]);
        // z neighbours: from the adjacent lane (DPP); the tile-row edge lanes read the
        // ring halo (all other lanes read one broadcast address: no bank conflicts)
        const T hl = s_rh[n][s0l * HPL + zl_idx], hr = s_rh[n][s0l * HPL + zr_idx];
        const T nl = zdpp<0x111>(c[V - 1]), nr = zdpp<0x101>(c[0]);
        const T zl = lz == 0 ? hl : nl, zr = lz == LZ - 1 ? hr : nr;
        gx = xp - xm;
        gy = yp - ym;
#pragma unroll
        for (int e = 0; e < V; ++e) gz[e] = (e + 1 < V ? c[e + 1] : zr) - (e > 0 ? c[e - 1] : zl);
      };
      if (is_int) {
#pragma unroll
        for (int n = 0; n < NB; ++n) {
          VT c, gx, gy, gz, f1, f2, f3;
          vgrad(n, c, gx, gy, gz);
#pragma unroll
          for (int e = 0; e < V; ++e) {
            f1[e] = flux3(gxx[fb][e], gx[e], gxy[fb][e], gy[e], gxz[fb][e], gz[e]);
            f2[e] = flux3(gxy[fb][e], gx[e], gyy[fb][e], gy[e], gyz[fb][e], gz[e]);
            f3[e] = flux3(gxz[fb][e], gx[e], gyz[fb][e], gy[e], gzz[fb][e], gz[e]);
          }
          *reinterpret_cast<VT*>(&s_fy[fb][n][(row + 1) * TZ + V * lz]) = f2;
          fx[n][(U + 1) % ZV_NS] = f1;
          fz[n] = f3;
          pc[n][fb] = c;
        }
      } else if (is_yh) {
#pragma unroll
        for (int n = 0; n < NB; ++n) {
          VT c, gx, gy, gz, f2;
          vgrad(n, c, gx, gy, gz);
#pragma unroll
          for (int e = 0; e < V; ++e) f2[e] = flux3(gxy[fb][e], gx[e], gyy[fb][e], gy[e], gyz[fb][e], gz[e]);
          *reinterpret_cast<VT*>(&s_fy[fb][n][(row + 1) * TZ + V * lz]) = f2;
        }
      } else {
#pragma unroll
        for (int n = 0; n < NB; ++n) {
          const T* hp = &s_rh[n][s0l * HPL];
          const T gx = s_rh[n][sp * HPL + zh_c] - s_rh[n][sm * HPL + zh_c];
          const T gy = hp[zh_c + 4] - hp[zh_c - 4];
          const T inner = s_ring[n][s0l * PLANE + zh_in];
          const T gz = side ? hp[zh_out] - inner : inner - hp[zh_out];
          s_fzh[fb][n][2 * row + side] = flux3(zxz[fb], gx, zyz[fb], gy, zzz[fb], gz);
        }
      }
      // flux coefficients of plane s + 2 into the set just used
      fetch_geom(s + 2, fb);
      __syncthreads();
      // ---------------- phase C: divergence of plane s, output plane o = s - 1
      const int o = s - 1;
      const unsigned ob = (o >= x0 && o < x1) ? (unsigned)o * plane_bytes : BAD;
      const bool obnd = (o == 0 || o == nx - 1);
      if (is_int) {
#pragma unroll
        for (int n = 0; n < NB; ++n) {
          const VT fyp = *reinterpret_cast<const VT*>(&s_fy[fb][n][(row + 2) * TZ + V * lz]);
          const VT fym = *reinterpret_cast<const VT*>(&s_fy[fb][n][row * TZ + V * lz]);
          const T hl = s_fzh[fb][n][2 * row], hr = s_fzh[fb][n][2 * row + 1];
          const T nl = zdpp<0x111>(fz[n][V - 1]);
          const T nr = zdpp<0x101>(fz[n][0]);
          const T fzl = lz == 0 ? hl : nl;
          const T fzr = lz == LZ - 1 ? hr : nr;
          VT d;
#pragma unroll
          for (int e = 0; e < V; ++e)
            d[e] = (fyp[e] - fym[e]) + ((e + 1 < V ? fz[n][e + 1] : fzr) - (e > 0 ? fz[n][e - 1] : fzl));
          dyz[n][fb] = d;
          const VT dx = fx[n][(U + 1) % ZV_NS] - fx[n][(U + 3) % ZV_NS];
          const VT dyzo = dyz[n][fb ^ 1];
          const VT po = pc[n][fb ^ 1];
          VT v;
#pragma unroll
          for (int e = 0; e < V; ++e) {
            const T lap = dx[e] + dyzo[e];
            T val = fma(ov[e], po[e], sc * (lap * oij[e]));
            if (p.interior_only && (obnd || ((bmask >> e) & 1u))) val = T(0);
            v[e] = val;
          }
          zst<T>(rout[n], (orb_ok[n] ? ob : BAD) + ooff, v);
        }
      }
      // commit plane s + 3 into the slot of plane s - 1, then request the next data
      store_plane(sm);
      fetch_plane(s + 4);
      fetch_out(s);
    }
  }
}

template <typename T, int NB, int WPE = 1>
int launch_monitor_zv_t(const MonParams<T>& base, hipStream_t stream) {
  constexpr int V = ZVec<T>::V;
  constexpr int TZ = ZShape<T>::LZ * V, TY = ZShape<T>::TY, NT = ZShape<T>::NT;
  MonParams<T> p = base;
  if ((long long)p.nx * p.ny * p.nz * (long long)sizeof(T) >= (1LL << 30)) return IG_EUNSUPPORTED;
  if (p.nz % V != 0 || p.nz < TZ || p.ny < TY || p.nx < 1 || p.ld % V != 0) return IG_EUNSUPPORTED;
  // 128-bit accesses: every operand base must be 16-byte aligned
  auto al = [](const void* q) { return q == nullptr || (reinterpret_cast<uintptr_t>(q) & 15u) == 0; };
  if (!al(p.in) || !al(p.out) || !al(p.coef) || !al(p.invj) || !al(p.veff)) return IG_EUNSUPPORTED;
  p.nty = (p.ny + TY - 1) / TY;
  p.ntz = (p.nz + TZ - 1) / TZ;
  const long long groups = (p.nblk + NB - 1) / NB;
  if (p.xc <= 0) {
    static int occ = -1;
    if (occ < 0) {
      int n = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, monitor_zv_kernel<T, NB, WPE>, NT, 0) != hipSuccess || n < 1)
        n = 1;
      occ = n;
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    const long long capacity = (long long)occ * cus;
    const long long tiles = (long long)p.nty * p.ntz;
    const int maxc = p.nx / 8 > 1 ? p.nx / 8 : 1;
    long long best = -1;
    int best_c = 1;
    for (int c = 1; c <= maxc; ++c) {
      const int xc = (p.nx + c - 1) / c;
      const long long blocks = 8LL * ((tiles * c + 7) / 8) * groups;
      const long long cost = ((blocks + capacity - 1) / capacity) * (xc + 6);
      if (best < 0 || cost < best) {
        best = cost;
        best_c = c;
      }
    }
    p.xc = (p.nx + best_c - 1) / best_c;
  }
  if (p.xc > p.nx) p.xc = p.nx;
  p.nxc = (p.nx + p.xc - 1) / p.xc;
  p.units = p.nty * p.ntz * p.nxc;
  const long long blocks = 8LL * groups * ((p.units + 7) / 8);
  if (blocks <= 0 || blocks > 0x7fffffffLL) return IG_EINVAL;
  hipLaunchKernelGGL((monitor_zv_kernel<T, NB, WPE>), dim3((unsigned)blocks), dim3(NT), 0, stream, p);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

template <typename T>
int launch_monitor_zv(const MonParams<T>& p, int nb, hipStream_t s) {
  if (nb <= 0 || (nb % 10) > p.nblk) nb = p.nblk >= 2 ? 2 : 1;
  // nb = NB + 10 * WPE (waves per SIMD the register allocation must allow; 0 = no bound)
  switch (nb) {
    case 1: return launch_monitor_zv_t<T, 1>(p, s);
    case 2: return launch_monitor_zv_t<T, 2>(p, s);
    case 31: return launch_monitor_zv_t<T, 1, 3>(p, s);
    case 41: return launch_monitor_zv_t<T, 1, 4>(p, s);
    case 32: return launch_monitor_zv_t<T, 2, 3>(p, s);
    default: return IG_EUNSUPPORTED;
  }
}


// ---------------------------------------------------------------------------------
// Register-queue forward trial_fix stencil: many orbitals per geometry fetch.
// Same tile (TY x TZ = 8 x 32, 256 lanes), same expression tree and x march as
// monitor_stencil_kernel, but the x neighbours come from registers: every lane owns one
// interior point and at most one extra point of the 12 x 36 plane (a y/z halo-flux point,
// or a point only the current plane needs), and keeps psi of its points for the planes
// s-1 .. s+2 in a 4-deep register queue (plane s+3 is requested as s-1 retires).  LDS
// holds only the current plane and the F_y / F_z staging, ~4 KB per orbital instead of
// ~14 KB, so a workgroup carries NB = 8-12 orbitals of one tile, loads the tile's geometry
// once for all of them, and two workgroups still fit on a CU.  Two barriers per plane step:
// after the plane is written (gradients and fluxes read it) and after the fluxes are staged
// (the divergence reads them); the next plane overwrites the buffers only after both.
template <typename T, int NB>
__global__ __launch_bounds__(IG_THREADS, 2) void monitor_rq_kernel(MonParams<T> p) {
  constexpr int TY = 8, TZ = 32, HP = 2, PY = TY + 2 * HP, PZ = TZ + 2 * HP, PLANE = PY * PZ;
  constexpr int FYS = (TY + 2) * TZ;      // F_y rows -1 .. TY
  constexpr int FZS = TY * (TZ + 2);      // F_z columns -1 .. TZ
  constexpr int NYH = 2 * TZ, NZH = 2 * TY;            // halo-flux points
  constexpr int NOR = 2 * TZ, NOC = 2 * TY, NCO = 4;   // outer rows, outer columns, corners
  constexpr int NEX = NYH + NZH + NOR + NOC + NCO;     // lanes with an extra point
  constexpr unsigned S = sizeof(T);
  static_assert(TY * TZ == IG_THREADS && NEX <= IG_THREADS, "one interior and at most one extra point per lane");

  // one spare entry each absorbs the writes of lanes without an extra point
  __shared__ T s_pl[NB][PLANE + 1];
  __shared__ T s_fl[NB][FYS + FZS + 1];

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
  const unsigned plane_bytes = (unsigned)(ny * nz) * S;
  const int j0 = min(tyi * TY, max(ny - TY, 0));
  const int k0 = min(tzi * TZ, max(nz - TZ, 0));
  const int x0 = xci * p.xc;
  const int x1 = min(x0 + p.xc, nx);
  const long long N = (long long)nx * ny * nz;

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
  const rsrc_t rxx = mkbuf(p.coef, N), ryy = mkbuf(p.coef + N, N), rzz = mkbuf(p.coef + 2 * N, N);
  const rsrc_t rxy = mkbuf(p.coef + 3 * N, N), rxz = mkbuf(p.coef + 4 * N, N), ryz = mkbuf(p.coef + 5 * N, N);
  const rsrc_t rij = mkbuf(p.invj, N);
  const rsrc_t rv = mkbuf(p.veff ? p.veff : p.invj, p.veff ? N : 0);

  const int tid = threadIdx.x;
  auto goff = [&](int row, int col) -> unsigned {  // in-plane byte offset, BAD outside the domain
    const int j = j0 + row, k = k0 + col;
    return (j >= 0 && j < ny && k >= 0 && k < nz) ? (unsigned)(j * nz + k) * S : BAD;
  };
  // interior point
  const int ty = tid / TZ, tz = tid % TZ;
  const int li = (ty + HP) * PZ + tz + HP;
  const unsigned ioff = goff(ty, tz);
  // extra point: kind 1 = F_y halo row, 2 = F_z halo column, 0 = plane only, -1 = none
  int erow = 0, ecol = 0, ekind = -1, est = FYS + FZS;
  if (tid < NYH) {
    const int hr = tid / TZ, c = tid % TZ;
    erow = hr ? TY : -1;
    ecol = c;
    ekind = 1;
    est = (hr ? TY + 1 : 0) * TZ + c;
  } else if (tid < NYH + NZH) {
    const int i2 = tid - NYH, side = i2 / TY, rr = i2 % TY;
    erow = rr;
    ecol = side ? TZ : -1;
    ekind = 2;
    est = FYS + rr * (TZ + 2) + (side ? TZ + 1 : 0);
  } else if (tid < NYH + NZH + NOR) {
    const int i3 = tid - NYH - NZH, hr = i3 / TZ;
    erow = hr ? TY + 1 : -2;
    ecol = i3 % TZ;
    ekind = 0;
  } else if (tid < NYH + NZH + NOR + NOC) {
    const int i4 = tid - NYH - NZH - NOR, side = i4 / TY;
    erow = i4 % TY;
    ecol = side ? TZ + 1 : -2;
    ekind = 0;
  } else if (tid < NEX) {
    const int i5 = tid - NYH - NZH - NOR - NOC;
    erow = (i5 >> 1) ? TY : -1;
    ecol = (i5 & 1) ? TZ : -1;
    ekind = 0;
  }
  const int ei = ekind >= 0 ? (erow + HP) * PZ + ecol + HP : PLANE;
  const unsigned eoff = ekind >= 0 ? goff(erow, ecol) : BAD;
  const unsigned hoff = ekind >= 1 ? eoff : BAD;  // geometry of a halo-flux point
  const bool yrow = ekind == 1;
  const bool has_flux = ekind >= 1;
  const unsigned toff = (j0 + ty < ny && k0 + tz < nz) ? ioff : BAD;

  auto pbytes = [&](int q) { return (q >= 0 && q < nx) ? (unsigned)q * plane_bytes : BAD; };
  const T tag = T(0);

  // register queues, slot of plane q: (q - sb + 1) & 3
  T qi[NB][4], qe[NB][4];
  T fx[NB][2];   // F_x at the interior point, slot (q - sb) & 1: F_x(s) replaces F_x(s - 2)
  T dyz[NB][2];  // y/z divergence, slot (q - sb) & 1
  T gc[2][6], hc[2][3], oij[2], ov[2];  // geometry sets, slot (q - sb) & 1
  auto fetch_psi = [&](int q, int slot) {
    const unsigned pb = pbytes(q);
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      qi[n][slot] = bld(rin[n], pb + ioff, tag);
      qe[n][slot] = bld(rin[n], pb + eoff, tag);
    }
  };
  auto fetch_geom = [&](int q, int w) {
    const unsigned pb = pbytes(q);
    const unsigned o = pb + ioff;
    gc[w][0] = bld(rxx, o, tag); gc[w][1] = bld(ryy, o, tag); gc[w][2] = bld(rzz, o, tag);
    gc[w][3] = bld(rxy, o, tag); gc[w][4] = bld(rxz, o, tag); gc[w][5] = bld(ryz, o, tag);
    const unsigned h = pb + hoff;
    const T a_xy = bld(rxy, h, tag), a_yy = bld(ryy, h, tag), a_yz = bld(ryz, h, tag);
    const T a_xz = bld(rxz, h, tag), a_zz = bld(rzz, h, tag);
    hc[w][0] = yrow ? a_xy : a_xz;
    hc[w][1] = yrow ? a_yy : a_yz;
    hc[w][2] = yrow ? a_yz : a_zz;
  };
  auto fetch_out = [&](int q, int w) {
    const unsigned o = pbytes(q) + ioff;
    oij[w] = bld(rij, o, tag);
    ov[w] = bld(rv, o, tag);
  };

  const int sb = x0 - 1;  // steps sb .. x1, output plane o = s - 1 in [x0, x1)
#pragma unroll
  for (int n = 0; n < NB; ++n) {
    fx[n][0] = fx[n][1] = T(0);
    dyz[n][0] = dyz[n][1] = T(0);
  }
#pragma unroll
  for (int d = 0; d < 4; ++d) fetch_psi(sb - 1 + d, d);
  fetch_geom(sb, 0);
  fetch_geom(sb + 1, 1);
  fetch_out(sb - 1, 1);
  fetch_out(sb, 0);
  const T sc = p.scale;

  for (int s0 = sb; s0 <= x1; s0 += 4) {
#pragma unroll
    for (int U = 0; U < 4; ++U) {
      const int s = s0 + U;
      if (s > x1) break;
      const int sm = U & 3, s0l = (U + 1) & 3, sp = (U + 2) & 3;  // slots of s-1, s, s+1
      const int W = U & 1;                                         // geometry / dyz set of s
      // ---- phase A: publish plane s
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        s_pl[n][li] = qi[n][s0l];
        s_pl[n][ei] = qe[n][s0l];
      }
      __syncthreads();
      // ---- phase B: gradients and fluxes of plane s
      T fxs[NB];
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        const T* pl = s_pl[n];
        {
          const T gx = qi[n][sp] - qi[n][sm];
          const T gy = pl[li + PZ] - pl[li - PZ];
          const T gz = pl[li + 1] - pl[li - 1];
          fxs[n] = flux3(gc[W][0], gx, gc[W][3], gy, gc[W][4], gz);
          s_fl[n][(ty + 1) * TZ + tz] = flux3(gc[W][3], gx, gc[W][1], gy, gc[W][5], gz);
          s_fl[n][FYS + ty * (TZ + 2) + tz + 1] = flux3(gc[W][4], gx, gc[W][5], gy, gc[W][2], gz);
        }
        if (has_flux) {
          const T gx = qe[n][sp] - qe[n][sm];
          const T gy = pl[ei + PZ] - pl[ei - PZ];
          const T gz = pl[ei + 1] - pl[ei - 1];
          s_fl[n][est] = flux3(hc[W][0], gx, hc[W][1], gy, hc[W][2], gz);
        }
      }
      fetch_geom(s + 2, W);
      __syncthreads();
      // ---- phase C: y/z divergence of plane s, output plane o = s - 1
      const int o = s - 1;
      const unsigned ob = (o >= x0 && o < x1) ? (unsigned)o * plane_bytes : BAD;
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        const int fy0 = (ty + 1) * TZ + tz;
        const int fz0 = FYS + ty * (TZ + 2) + tz + 1;
        dyz[n][W] = (s_fl[n][fy0 + TZ] - s_fl[n][fy0 - TZ]) + (s_fl[n][fz0 + 1] - s_fl[n][fz0 - 1]);
        const T dx = fxs[n] - fx[n][W];
        fx[n][W] = fxs[n];
        const T lap = dx + dyz[n][W ^ 1];
        const T v = fma(ov[W ^ 1], qi[n][sm], sc * (lap * oij[W ^ 1]));
        bst(rout[n], (orb_ok[n] ? ob : BAD) + toff, v);
      }
      fetch_out(s + 1, W ^ 1);
      fetch_psi(s + 3, sm);
    }
  }
}

// auto_pick: the launcher's own choice, declined (IG_EUNSUPPORTED) when the launch would hold
// fewer than ~2 residency rounds of workgroups, where the general / z-vectorised kernels
// fill the GPU better (96^3, 16 orbitals: 0.11-0.21 ms here against 0.052 ms).
template <typename T, int NB>
int launch_monitor_rq_t(const MonParams<T>& base, hipStream_t stream, bool auto_pick = false) {
  MonParams<T> p = base;
  if ((long long)p.nx * p.ny * p.nz * (long long)sizeof(T) >= (1LL << 30)) return IG_EUNSUPPORTED;
  if (p.interior_only) return IG_EUNSUPPORTED;
  p.nty = (p.ny + 7) / 8;
  p.ntz = (p.nz + 31) / 32;
  const long long groups = (p.nblk + NB - 1) / NB;
  if (p.xc <= 0) {
    // chunks of ~56 planes (the measured best of the 224^3 block for the kernel above)
    const int c = p.nx >= 112 ? (p.nx + 28) / 56 : 1;
    p.xc = (p.nx + c - 1) / c;
  }
  if (p.xc > p.nx) p.xc = p.nx;
  p.nxc = (p.nx + p.xc - 1) / p.xc;
  p.units = p.nty * p.ntz * p.nxc;
  const long long blocks = 8LL * groups * ((p.units + 7) / 8);
  if (blocks <= 0 || blocks > 0x7fffffffLL) return IG_EINVAL;
  if (auto_pick && blocks < 2048) return IG_EUNSUPPORTED;
  hipLaunchKernelGGL((monitor_rq_kernel<T, NB>), dim3((unsigned)blocks), dim3(IG_THREADS), 0, stream, p);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

// ---------------------------------------------------------------------------------
// z-vectorised register-queue forward trial_fix stencil.  Every lane owns one vector
// unit of V = 16 / sizeof(T) consecutive z points (128-bit global and LDS accesses) and
// at most one scalar point of the tile's z halo.  A tile row is LZ lanes (TZ = 32
// points); the 256 lanes are the PY = TY + 4 rows -2 .. TY + 1 of the plane (FP32: 8
// lanes x 32 rows, TY = 28; FP64: 16 lanes x 16 rows, TY = 12), so every lane runs the
// same instructions: interior rows produce output, rows -1 / TY only their y flux, rows
// -2 / TY + 1 only plane values.  The scalar points: lanes lz = 0 / LZ - 1 of the interior
// rows own the z-flux points k0 - 1 / k0 + TZ, of rows -1 / TY the plane values there;
// lanes 1 / LZ - 2 of the interior rows the plane values at k0 - 2 / k0 + TZ + 1.  psi of
// both points is kept for the planes s - 1 .. s + 2 in per-lane register queues (x
// neighbours from registers), LDS holds the current plane (rows with their two z-halo
// points on either side: every z and y neighbour is one LDS read) and the staged F_y /
// F_z.  Same expression tree as the general kernel: bit-identical results.  Two barriers
// per plane step (plane written; fluxes staged).
template <typename T>
struct QShape;
template <>
struct QShape<float> {
  static constexpr int LZ = 8, TY = 28;
};
template <>
struct QShape<double> {
  static constexpr int LZ = 16, TY = 12;
};

template <typename T>
__device__ __forceinline__ typename ZVec<T>::t lds_vec(const T* a, int i) {
  return *reinterpret_cast<const typename ZVec<T>::t*>(a + i);
}
template <typename T>
__device__ __forceinline__ void lds_vec_st(T* a, int i, typename ZVec<T>::t v) {
  *reinterpret_cast<typename ZVec<T>::t*>(a + i) = v;
}

template <typename T, int NB>
__global__ __launch_bounds__(IG_THREADS, 2) void monitor_zq_kernel(MonParams<T> p) {
  using VT = typename ZVec<T>::t;
  constexpr int V = ZVec<T>::V;
  constexpr int LZ = QShape<T>::LZ, TY = QShape<T>::TY, TZ = LZ * V;
  constexpr int PY = TY + 4;          // plane rows -2 .. TY + 1
  constexpr int RP = TZ + 2 * V;      // plane row: [halo vector | TZ points | halo vector]
  constexpr int PLANE = PY * RP;      // + one spare vector for lanes without a point
  constexpr int FYR = TY + 2;         // F_y rows -1 .. TY, packed
  constexpr int FYS = FYR * TZ;       // + spare vector
  constexpr int FZS = TY * RP;        // F_z rows 0 .. TY - 1 with the k0 - 1 / k0 + TZ halo, + spare
  constexpr unsigned S = sizeof(T);
  static_assert(PY * LZ == IG_THREADS, "one plane vector per lane");
  static_assert(TZ == 32, "32-point tile rows");

  __shared__ __attribute__((aligned(16))) T s_pl[NB][PLANE + V];
  __shared__ __attribute__((aligned(16))) T s_fy[NB][FYS + V];
  __shared__ __attribute__((aligned(16))) T s_fz[NB][FZS + V];

  const int ngroups = (p.nblk + NB - 1) / NB;
  const int b = blockIdx.x;
  const int xcd = b & 7;
  const int r = b >> 3;
  const int grp = r % ngroups;
  int u;
  if (p.ug <= 0) {
    const int upx = (p.units + 7) / 8;
    u = xcd * upx + r / ngroups;
  } else {
    // items of ug consecutive units (z tiles fastest) dealt round robin to the XCDs: at a
    // given moment the eight XCDs hold 8 ug consecutive units, so the z- and y-halo lines
    // at the edges of an XCD's units are fetched by two XCDs at about the same time
    const int v = r / ngroups;
    u = ((v / p.ug) * 8 + xcd) * p.ug + v % p.ug;
  }
  if (u >= p.units) return;
  const int tzi = u % p.ntz;
  const int tyi = (u / p.ntz) % p.nty;
  const int xci = u / (p.ntz * p.nty);
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const unsigned plane_bytes = (unsigned)(ny * nz) * S;
  const int j0 = max(min(tyi * TY, ny - TY), 0);
  const int k0 = max(min(tzi * TZ, nz - TZ), 0);
  const int x0 = xci * p.xc;
  const int x1 = min(x0 + p.xc, nx);
  const long long N = (long long)nx * ny * nz;

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
  const rsrc_t rxx = mkbuf(p.coef, N), ryy = mkbuf(p.coef + N, N), rzz = mkbuf(p.coef + 2 * N, N);
  const rsrc_t rxy = mkbuf(p.coef + 3 * N, N), rxz = mkbuf(p.coef + 4 * N, N), ryz = mkbuf(p.coef + 5 * N, N);
  const rsrc_t rij = mkbuf(p.invj, N);
  const rsrc_t rv = mkbuf(p.veff ? p.veff : p.invj, p.veff ? N : 0);

  const int tid = threadIdx.x;
  // ---- vector unit: row r of the plane (-2 .. TY + 1), z points k0 + V lz .. + V - 1
  const int row = tid / LZ - 2, lz = tid % LZ;
  const int gj = j0 + row;
  const bool jok = gj >= 0 && gj < ny;
  const bool kok = k0 + V * lz < nz;  // nz % V == 0: all V points in or all out
  const unsigned voff = (jok && kok) ? (unsigned)(gj * nz + k0 + V * lz) * S : BAD;
  const bool interior = row >= 0 && row < TY;
  const bool yflux = row >= -1 && row <= TY;
  const unsigned goff = yflux ? voff : BAD;      // flux coefficients (rows -1 .. TY)
  const unsigned ooff = interior ? voff : BAD;   // 1/J, V and the output
  const int pw = (row + 2) * RP + V + V * lz;    // plane position of the vector
  const int rm = max(row - 1, -2), rp = min(row + 1, TY + 1);
  const int pym = (rm + 2) * RP + V + V * lz, pyp = (rp + 2) * RP + V + V * lz;
  const int fyw = yflux ? (row + 1) * TZ + V * lz : FYS;
  const int cr = min(max(row, 0), TY - 1);       // clamped row for the divergence reads
  const int fym = cr * TZ + V * lz, fyp = (cr + 2) * TZ + V * lz;
  const int fzw = interior ? row * RP + V + V * lz : FZS;
  const int fzc = cr * RP + V + V * lz;
  // ---- scalar point of the z halo
  int ecol = 0, epos = PLANE, efz = FZS;
  bool eflux = false, ehas = false;
  if (interior || row == -1 || row == TY) {
    if (lz == 0 || lz == LZ - 1) {
      const bool left = lz == 0;
      ecol = left ? k0 - 1 : k0 + TZ;
      epos = (row + 2) * RP + (left ? V - 1 : V + TZ);
      ehas = true;
      if (interior) {
        eflux = true;
        efz = row * RP + (left ? V - 1 : V + TZ);
      }
    } else if (interior && (lz == 1 || lz == LZ - 2)) {
      const bool left = lz == 1;
      ecol = left ? k0 - 2 : k0 + TZ + 1;
      epos = (row + 2) * RP + (left ? V - 2 : V + TZ + 1);
      ehas = true;
    }
  }
  const unsigned eoff = (ehas && jok && ecol >= 0 && ecol < nz) ? (unsigned)(gj * nz + ecol) * S : BAD;
  const unsigned hoff = eflux ? eoff : BAD;
  const int ec = eflux ? epos : RP + 1;  // neighbour reads of non-flux lanes stay inside the plane

  auto pbytes = [&](int q) { return (q >= 0 && q < nx) ? (unsigned)q * plane_bytes : BAD; };
  const T tag = T(0);

  VT q[NB][4];    // psi of the vector unit, slot of plane q: (q - sb + 1) & 3
  T qe[NB][4];    // psi of the scalar point, same slots
  VT fx[NB][2];   // F_x, slot (q - sb) & 1: F_x(s) replaces F_x(s - 2)
  VT dyz[NB];     // y/z divergence of the previous plane
  VT gc[6];       // flux coefficients of the current plane
  T hc[3];        // z-flux coefficients of the scalar point
  VT oij, ov;     // 1/J and V of the output plane
  auto fetch_psi = [&](int qq, int slot) {
    const unsigned pb = pbytes(qq);
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      q[n][slot] = zld<T>(rin[n], pb + voff);
      qe[n][slot] = bld(rin[n], pb + eoff, tag);
    }
  };
  auto fetch_geom = [&](int qq) {
    const unsigned o = pbytes(qq) + goff;
    gc[0] = zld<T>(rxx, o); gc[1] = zld<T>(ryy, o); gc[2] = zld<T>(rzz, o);
    gc[3] = zld<T>(rxy, o); gc[4] = zld<T>(rxz, o); gc[5] = zld<T>(ryz, o);
    const unsigned h = pbytes(qq) + hoff;
    hc[0] = bld(rxz, h, tag); hc[1] = bld(ryz, h, tag); hc[2] = bld(rzz, h, tag);
  };
  auto fetch_out = [&](int qq) {
    const unsigned o = pbytes(qq) + ooff;
    oij = zld<T>(rij, o);
    ov = zld<T>(rv, o);
  };

  const int sb = x0 - 1;  // steps sb .. x1, output plane o = s - 1 in [x0, x1)
#pragma unroll
  for (int n = 0; n < NB; ++n) {
    fx[n][0] = fx[n][1] = VT(0);
    dyz[n] = VT(0);
  }
#pragma unroll
  for (int d = 0; d < 4; ++d) fetch_psi(sb - 1 + d, d);
  fetch_geom(sb);
  fetch_out(sb - 1);
  const T sc = p.scale;

  // Whole groups of 4 steps, no exit inside the unrolled body (a branch there makes the
  // compiler drain every outstanding load at the join); steps past x1 load zeros and
  // store nothing.
  for (int s0 = sb; s0 <= x1; s0 += 4) {
#pragma unroll
    for (int U = 0; U < 4; ++U) {
      const int s = s0 + U;
      const int sm = U & 3, s0l = (U + 1) & 3, sp = (U + 2) & 3;  // slots of s-1, s, s+1
      const int W = U & 1;
      // ---- phase A: publish plane s
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        lds_vec_st<T>(s_pl[n], pw, q[n][s0l]);
        s_pl[n][epos] = qe[n][s0l];
      }
      __syncthreads();
      // ---- phase B: gradients and fluxes of plane s
      VT fxs[NB], fzs[NB];
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        const T* pl = s_pl[n];
        const VT c = q[n][s0l];
        const VT gx = q[n][sp] - q[n][sm];
        const VT gy = lds_vec<T>(pl, pyp) - lds_vec<T>(pl, pym);
        const T zl = pl[pw - 1], zr = pl[pw + V];
        VT gz, f2;
#pragma unroll
        for (int e = 0; e < V; ++e) {
          gz[e] = (e + 1 < V ? c[e + 1] : zr) - (e > 0 ? c[e - 1] : zl);
          fxs[n][e] = flux3(gc[0][e], gx[e], gc[3][e], gy[e], gc[4][e], gz[e]);
          f2[e] = flux3(gc[3][e], gx[e], gc[1][e], gy[e], gc[5][e], gz[e]);
          fzs[n][e] = flux3(gc[4][e], gx[e], gc[5][e], gy[e], gc[2][e], gz[e]);
        }
        lds_vec_st<T>(s_fy[n], fyw, f2);
        lds_vec_st<T>(s_fz[n], fzw, fzs[n]);
        // keep F_x here: evaluated after the barrier it would hold the three gradients live
        asm volatile("" : "+v"(fxs[n]));
        {
          const T egx = qe[n][sp] - qe[n][sm];
          const T egy = pl[ec + RP] - pl[ec - RP];
          const T egz = pl[ec + 1] - pl[ec - 1];
          s_fz[n][efz] = flux3(hc[0], egx, hc[1], egy, hc[2], egz);
        }
      }
      fetch_geom(s + 1);
      __syncthreads();
      // ---- phase C: y/z divergence of plane s, output plane o = s - 1
      const int o = s - 1;
      const unsigned ob = (o >= x0 && o < x1) ? (unsigned)o * plane_bytes : BAD;
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        const VT fyp_v = lds_vec<T>(s_fy[n], fyp), fym_v = lds_vec<T>(s_fy[n], fym);
        const T fzl = s_fz[n][fzc - 1], fzr = s_fz[n][fzc + V];
        const VT fz = fzs[n];
        const VT dx = fxs[n] - fx[n][W];
        fx[n][W] = fxs[n];
        const VT po = q[n][sm];
        VT d, v;
#pragma unroll
        for (int e = 0; e < V; ++e) {
          d[e] = (fyp_v[e] - fym_v[e]) + ((e + 1 < V ? fz[e + 1] : fzr) - (e > 0 ? fz[e - 1] : fzl));
          const T lap = dx[e] + dyz[n][e];
          v[e] = fma(ov[e], po[e], sc * (lap * oij[e]));
        }
        dyz[n] = d;
        zst<T>(rout[n], (orb_ok[n] ? ob : BAD) + ooff, v);
      }
      fetch_out(s);
      fetch_psi(s + 3, sm);
    }
  }
}

template <typename T, int NB>
int launch_monitor_zq_t(const MonParams<T>& base, hipStream_t stream, bool auto_pick = false) {
  constexpr int V = ZVec<T>::V;
  constexpr int TY = QShape<T>::TY, TZ = QShape<T>::LZ * V;
  MonParams<T> p = base;
  if ((long long)p.nx * p.ny * p.nz * (long long)sizeof(T) >= (1LL << 30)) return IG_EUNSUPPORTED;
  if (p.interior_only) return IG_EUNSUPPORTED;
  // 128-bit accesses: every z row and every orbital must start on a 16-byte boundary
  if (p.nz % V != 0 || p.ld % V != 0) return IG_EUNSUPPORTED;
  if (((uintptr_t)p.in | (uintptr_t)p.out | (uintptr_t)p.coef | (uintptr_t)p.invj | (uintptr_t)p.veff) & 15)
    return IG_EUNSUPPORTED;
  p.nty = (p.ny + TY - 1) / TY;
  p.ntz = (p.nz + TZ - 1) / TZ;
  const long long groups = (p.nblk + NB - 1) / NB;
  if (p.xc <= 0) {
    // chunks of ~56 planes, rounded to 4 m + 2 planes: a chunk runs its planes plus two
    // halo steps in whole groups of 4 (224 planes: 58, 58, 58, 50)
    const int c = p.nx >= 112 ? (p.nx + 28) / 56 : 1;
    p.xc = (p.nx + c - 1) / c;
    p.xc = ((p.xc + 1) / 4) * 4 + 2;
  }
  if (p.xc > p.nx) p.xc = p.nx;
  p.nxc = (p.nx + p.xc - 1) / p.xc;
  p.units = p.nty * p.ntz * p.nxc;
  const long long per = p.ug > 0 ? 8LL * p.ug : 8LL;
  const long long blocks = 8LL * groups * ((p.units + per - 1) / per) * (p.ug > 0 ? p.ug : 1);
  if (blocks <= 0 || blocks > 0x7fffffffLL) return IG_EINVAL;
  if (auto_pick && blocks < 2048) return IG_EUNSUPPORTED;
  hipLaunchKernelGGL((monitor_zq_kernel<T, NB>), dim3((unsigned)blocks), dim3(IG_THREADS), 0, stream, p);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

// ---------------------------------------------------------------------------------
// Pipelined full-z register-queue forward trial_fix stencil, for grids whose z rows fit
// one wavefront (nz <= 64 V: FP32 256, FP64 128 points; the default for large FP32 blocks).
// A tile spans the whole z extent, so no z halo is ever fetched.  (With 32-point tiles
// every tile row is one 128-byte line and its four z-halo points pull the two
// neighbouring lines from HBM: at 224^3, k = 48 that doubled the psi read traffic, and
// the z-neighbour tile's workgroup does not reliably hit them in L2.)  Wavefront w owns
// plane row w - 2 of a TY-row tile: rows 0 .. TY - 1 produce output, rows -1 / TY only
// their y flux, rows -2 / TY + 1 only plane values (wavefront-uniform branches); lane l
// owns the V points V l .. V l + V - 1.  psi of the lane's points for the planes
// s - 1 .. s + 2 is kept in register queues (x neighbours), z neighbours of psi and of
// F_z come from the adjacent lane by lane permutes (ds_bpermute: no LDS storage, no bank
// conflicts; the zero ghosts at k = -1 / nz are the permute of lane 0 / 63 masked, or a
// lane past nz holding zeros), and LDS holds only the plane and F_y, both double
// buffered: step s computes the fluxes of plane s and the divergence and output of plane
// s - 1, publishes plane s + 1, and ends with the step's only barrier.  The work per
// step that does not depend on the orbital (6 + 2 coefficient loads per lane, the
// barrier) is shared by the NB orbitals of the workgroup, so NB = 3 is used: 12
// wavefronts at <= 168 VGPRs (3 per SIMD) and 144 KB of LDS fill a CU.  Same expression
// tree as the general kernel: bit-identical results.  Block order (MonParams::order): the
// y-halo rows of a tile are the interior rows of its y neighbours, and the halo planes of
// an x chunk are the last planes of the previous chunk; ordering the workgroups so that
// y neighbours run at the same time (across XCDs, through the memory-side cache) and an
// XCD takes the chunks of one tile pair in sequence (through its L2) serves most of both
// from cache instead of HBM.
template <typename T>
__device__ __forceinline__ T lane_from(int addr, T v);
template <>
__device__ __forceinline__ float lane_from<float>(int addr, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, v)));
}
template <>
__device__ __forceinline__ double lane_from<double>(int addr, double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)(b & 0xffffffffull));
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}

template <typename T, int NB, int GD, int TY, bool DOT = false>
__global__ __launch_bounds__((TY + 4) * 64) void monitor_zp_kernel(MonParams<T> p) {
  using VT = typename ZVec<T>::t;
  constexpr int V = ZVec<T>::V;
  constexpr int PY = TY + 4;   // plane rows -2 .. TY + 1, one per wavefront
  constexpr int NT = PY * 64;
  constexpr int RW = 64 * V;   // points per row (lanes past nz hold zeros)
  constexpr unsigned S = sizeof(T);
  static_assert(NT <= 1024, "one wavefront per plane row");
  static_assert(GD == 1 || GD == 2, "geometry one or two planes ahead");

  __shared__ __attribute__((aligned(16))) T s_pl[2][NB][PY * RW];
  __shared__ __attribute__((aligned(16))) T s_fy[2][NB][PY * RW];

  const int ngroups = (p.nblk + NB - 1) / NB;
  const int b = blockIdx.x;
  const int xcd = b & 7;
  const int r = b >> 3;
  const int grp = r % ngroups;
  int tyi, xci;
  if (p.order == 0) {
    // each XCD takes a contiguous range of units (y tiles fastest), the orbital groups of
    // a unit back to back: two y-adjacent units share their boundary rows in L2
    const int upx = (p.units + 7) / 8;
    const int u = xcd * upx + r / ngroups;
    if (u >= p.units) return;
    tyi = u % p.nty;
    xci = u / p.nty;
  } else {
    // y sweep: an item is a pair of y-adjacent tiles at one x chunk; at a given moment the
    // eight XCDs hold eight consecutive pairs of the same chunk, so the rows on either side
    // of an XCD boundary are fetched by two XCDs at about the same time (one HBM read,
    // the other served by the memory-side cache)
    const int v = r / ngroups;
    const int npairs = (p.nty + 1) >> 1;
    const int item = (v >> 1) * 8 + xcd;
    if (item >= npairs * p.nxc) return;
    int pair;
    if (p.order == 1) {
      xci = item / npairs;
      pair = item % npairs;
    } else {
      // bands of eight pairs swept chunk by chunk: XCD c keeps pair c of a full band for
      // all its chunks, so the halo planes of a chunk were just read by the same XCD
      const int full = (npairs >> 3) * 8 * p.nxc;
      if (item < full) {
        const int w = item % (8 * p.nxc);
        xci = w >> 3;
        pair = (item / (8 * p.nxc)) * 8 + (w & 7);
      } else {
        const int rp = npairs & 7;
        const int w = item - full;
        xci = w / rp;
        pair = (npairs & ~7) + w % rp;
      }
    }
    tyi = 2 * pair + (v & 1);
    if (tyi >= p.nty) return;
  }
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const unsigned plane_bytes = (unsigned)(ny * nz) * S;
  const int j0 = max(min(tyi * TY, ny - TY), 0);
  const int x0 = xci * p.xc;
  const int x1 = min(x0 + p.xc, nx);
  const long long N = (long long)nx * ny * nz;

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
  const rsrc_t rxx = mkbuf(p.coef, N), ryy = mkbuf(p.coef + N, N), rzz = mkbuf(p.coef + 2 * N, N);
  const rsrc_t rxy = mkbuf(p.coef + 3 * N, N), rxz = mkbuf(p.coef + 4 * N, N), ryz = mkbuf(p.coef + 5 * N, N);
  const rsrc_t rij = mkbuf(p.invj, N);
  const rsrc_t rv = mkbuf(p.veff ? p.veff : p.invj, p.veff ? N : 0);
  const rsrc_t rdw = mkbuf(DOT ? p.dotw : p.invj, DOT ? N : 0);

  const int tid = threadIdx.x;
  const int row = __builtin_amdgcn_readfirstlane(tid >> 6) - 2;  // uniform per wavefront
  const int ln = tid & 63;
  const int gj = j0 + row;
  const bool jok = gj >= 0 && gj < ny;
  const bool kok = V * ln < nz;  // nz % V == 0: all V points in or all out
  const unsigned voff = (jok && kok) ? (unsigned)(gj * nz + V * ln) * S : BAD;
  const bool interior = row >= 0 && row < TY;
  const bool yflux = row >= -1 && row <= TY;
  const unsigned goff = yflux ? voff : BAD;
  const unsigned ooff = interior ? voff : BAD;
  const int pw = (row + 2) * RW + V * ln;  // LDS position of the lane's vector (plane and F_y)
  // y neighbours; the outermost wavefronts (whose results are unused) read their own row
  const int pym = row > -2 ? pw - RW : pw, pyp = row < TY + 1 ? pw + RW : pw;
  const int al = ((ln + 63) & 63) * 4, ar = ((ln + 1) & 63) * 4;  // permute addresses of the z neighbours
  const bool zl_ghost = ln == 0, zr_ghost = ln == 63;  // lanes past nz hold zeros already

  auto pbytes = [&](int q) { return (q >= 0 && q < nx) ? (unsigned)q * plane_bytes : BAD; };

  // psi in registers only for the plane being published and the one in flight, slot of
  // plane q: (q - sb) & 1; a lane reads its own psi of planes s - 1 and s back from its
  // slots of the two plane buffers (plane s - 1's slot is overwritten by the lane itself
  // when it publishes plane s + 1, after that read)
  VT q[NB][2];
  VT fx[NB][2];   // F_x, slot (q - sb) & 1
  VT dz[NB];      // z part of the divergence of the previous plane
  VT gc[GD][6];   // flux coefficients, set (q - sb) % GD
  VT oij, ov;     // 1/J and V of the output plane
  VT odw;         // dot weights of the output plane (DOT)
  // (DOT) FP64 sums of w psi out per orbital over the rows this tile owns: a clamped last
  // y tile recomputes rows of the previous one, which count there only
  double dacc[NB];
  const bool dot_rows = interior && gj >= tyi * TY;
#pragma unroll
  for (int n = 0; n < NB; ++n) dacc[n] = 0.0;
  auto fetch_psi = [&](int qq, int slot) {
    const unsigned pb = pbytes(qq);
#pragma unroll
    for (int n = 0; n < NB; ++n) q[n][slot] = zld<T>(rin[n], pb + voff);
  };
  // wavefront-uniform roles: the y-halo rows need only the F_y coefficients, the outer rows
  // no geometry at all, and only interior rows read 1/J and V
  auto fetch_geom = [&](int qq, int w) {
    const unsigned o = pbytes(qq) + goff;
    if (interior) {
      gc[w][0] = zld<T>(rxx, o); gc[w][2] = zld<T>(rzz, o); gc[w][4] = zld<T>(rxz, o);
    }
    if (yflux) {
      gc[w][1] = zld<T>(ryy, o); gc[w][3] = zld<T>(rxy, o); gc[w][5] = zld<T>(ryz, o);
    }
  };
  auto fetch_out = [&](int qq) {
    if (interior) {
      const unsigned o = pbytes(qq) + ooff;
      oij = zld<T>(rij, o);
      ov = zld<T>(rv, o);
      if (DOT) odw = zld<T>(rdw, o);
    }
  };

  const int sb = x0 - 1;  // steps sb .. x1 (whole groups of 4), output plane o = s - 1 in [x0, x1)
#pragma unroll
  for (int n = 0; n < NB; ++n) {
    fx[n][0] = fx[n][1] = VT(0);
    dz[n] = VT(0);
  }
  fetch_psi(sb - 1, 1);
  fetch_psi(sb, 0);
#pragma unroll
  for (int d = 0; d < GD; ++d) fetch_geom(sb + d, d);
  fetch_out(sb - 1);
  const T sc = p.scale;
  // planes sb - 1 and sb into their buffers; buffer of plane q: (q - sb) & 1
#pragma unroll
  for (int n = 0; n < NB; ++n) {
    lds_vec_st<T>(s_pl[1][n], pw, q[n][1]);
    lds_vec_st<T>(s_pl[0][n], pw, q[n][0]);
  }
  fetch_psi(sb + 1, 1);
  fetch_psi(sb + 2, 0);
  __syncthreads();

  for (int s0 = sb; s0 <= x1; s0 += 4) {
#pragma unroll
    for (int U = 0; U < 4; ++U) {
      const int s = s0 + U;
      const int W = U & 1;  // buffer / F_x slot of plane s; psi(s + 1) in register slot W ^ 1
      const int G = GD == 2 ? W : 0;
      const int o = s - 1;
      const unsigned ob = (o >= x0 && o < x1) ? (unsigned)o * plane_bytes : BAD;
      if (interior) {
#pragma unroll
      for (int n = 0; n < NB; ++n) {
        // ---- fluxes of plane s
        const T* pl = s_pl[W][n];
        const VT c = lds_vec<T>(pl, pw);
        const VT pm = lds_vec<T>(s_pl[W ^ 1][n], pw);  // psi(s - 1)
        const VT gx = q[n][W ^ 1] - pm;
        const VT gy = lds_vec<T>(pl, pyp) - lds_vec<T>(pl, pym);
        T zl = lane_from<T>(al, c[V - 1]), zr = lane_from<T>(ar, c[0]);
        zl = zl_ghost ? T(0) : zl;
        zr = zr_ghost ? T(0) : zr;
        VT gz, fxs, f2, f3;
#pragma unroll
        for (int e = 0; e < V; ++e) {
          gz[e] = (e + 1 < V ? c[e + 1] : zr) - (e > 0 ? c[e - 1] : zl);
          fxs[e] = flux3(gc[G][0][e], gx[e], gc[G][3][e], gy[e], gc[G][4][e], gz[e]);
          f2[e] = flux3(gc[G][3][e], gx[e], gc[G][1][e], gy[e], gc[G][5][e], gz[e]);
          f3[e] = flux3(gc[G][4][e], gx[e], gc[G][5][e], gy[e], gc[G][2][e], gz[e]);
        }
        lds_vec_st<T>(s_fy[W][n], pw, f2);
        T fzl = lane_from<T>(al, f3[V - 1]), fzr = lane_from<T>(ar, f3[0]);
        fzl = zl_ghost ? T(0) : fzl;
        fzr = zr_ghost ? T(0) : fzr;
        VT dzn;
#pragma unroll
        for (int e = 0; e < V; ++e) dzn[e] = (e + 1 < V ? f3[e + 1] : fzr) - (e > 0 ? f3[e - 1] : fzl);
        // ---- divergence and output of plane s - 1 (F_y staged in the previous step)
        const T* fyo = s_fy[W ^ 1][n];
        const VT fyp_v = lds_vec<T>(fyo, pyp), fym_v = lds_vec<T>(fyo, pym);
        const VT dx = fxs - fx[n][W];
        fx[n][W] = fxs;
        const VT po = pm;
        VT v;
#pragma unroll
        for (int e = 0; e < V; ++e) {
          const T d = (fyp_v[e] - fym_v[e]) + dz[n][e];
          const T lap = dx[e] + d;
          v[e] = fma(ov[e], po[e], sc * (lap * oij[e]));
        }
        dz[n] = dzn;
        zst<T>(rout[n], (orb_ok[n] ? ob : BAD) + ooff, v);
        if (DOT && dot_rows && ob != BAD) {
          // lanes past nz and rows past ny read zero weights (range-checked loads)
#pragma unroll
          for (int e = 0; e < V; ++e) dacc[n] += (double)odw[e] * (double)po[e] * (double)v[e];
        }
        // ---- publish plane s + 1
        lds_vec_st<T>(s_pl[W ^ 1][n], pw, q[n][W ^ 1]);
      }
      } else if (yflux) {
        // y-halo rows: F_y of plane s only
#pragma unroll
        for (int n = 0; n < NB; ++n) {
          const T* pl = s_pl[W][n];
          const VT c = lds_vec<T>(pl, pw);
          const VT gx = q[n][W ^ 1] - lds_vec<T>(s_pl[W ^ 1][n], pw);
          const VT gy = lds_vec<T>(pl, pyp) - lds_vec<T>(pl, pym);
          T zl = lane_from<T>(al, c[V - 1]), zr = lane_from<T>(ar, c[0]);
          zl = zl_ghost ? T(0) : zl;
          zr = zr_ghost ? T(0) : zr;
          VT f2;
#pragma unroll
          for (int e = 0; e < V; ++e) {
            const T gz = (e + 1 < V ? c[e + 1] : zr) - (e > 0 ? c[e - 1] : zl);
            f2[e] = flux3(gc[G][3][e], gx[e], gc[G][1][e], gy[e], gc[G][5][e], gz);
          }
          lds_vec_st<T>(s_fy[W][n], pw, f2);
          lds_vec_st<T>(s_pl[W ^ 1][n], pw, q[n][W ^ 1]);
        }
      } else {
#pragma unroll
        for (int n = 0; n < NB; ++n) lds_vec_st<T>(s_pl[W ^ 1][n], pw, q[n][W ^ 1]);
      }
      fetch_geom(s + GD, G);
      fetch_out(s);
      fetch_psi(s + 3, W ^ 1);
      __syncthreads();
    }
  }
  if (DOT) {
    // wavefront sums (fixed butterfly), then the rows in order through LDS (the loop's last
    // barrier has retired every read of s_fy): one partial per orbital and (y tile, x
    // chunk), summed in a fixed order by the finishing kernel
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      double a = dot_rows ? dacc[n] : 0.0;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
      dacc[n] = a;
    }
    double* red = reinterpret_cast<double*>(&s_fy[0][0][0]);
    if (ln == 0) {
#pragma unroll
      for (int n = 0; n < NB; ++n) red[(row + 2) * NB + n] = dacc[n];
    }
    __syncthreads();
    if (tid < NB) {
      double t = 0.0;
      for (int r = 0; r < PY; ++r) t += red[r * NB + tid];
      const int orb = grp * NB + tid;
      if (orb < p.nblk) p.dpart[(long long)orb * p.units + tyi + p.nty * xci] = t;
    }
  }
}

// x chunk of the full-z kernel: ~56 planes, rounded to 4 m + 2 planes, so a chunk runs its
// planes plus two halo steps in whole groups of 4 (224 planes: 58, 58, 58, 50)
inline int zp_chunk(int nx) {
  const int c = nx >= 112 ? (nx + 28) / 56 : 1;
  int xc = (nx + c - 1) / c;
  xc = ((xc + 1) / 4) * 4 + 2;
  return xc > nx ? nx : xc;
}

// out[m] = sum_u part[m][u] over the units in order (the row dots of the full-z kernel)
__global__ void zp_dots_finish(const double* __restrict__ part, int units, int nblk, double* __restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nblk) return;
  double t = 0.0;
  for (int u = 0; u < units; ++u) t += part[(long long)m * units + u];
  out[m] = t;
}

template <typename T, int NB, int GD, int TY, bool DOT = false>
int launch_monitor_zp_t(const MonParams<T>& base, hipStream_t stream, bool auto_pick = false,
                        double* dots = nullptr) {
  constexpr int V = ZVec<T>::V;
  MonParams<T> p = base;
  if ((long long)p.nx * p.ny * p.nz * (long long)sizeof(T) >= (1LL << 30)) return IG_EUNSUPPORTED;
  if (p.interior_only) return IG_EUNSUPPORTED;
  if (p.nz > 64 * V || p.nz % V != 0 || p.ld % V != 0) return IG_EUNSUPPORTED;
  if (((uintptr_t)p.in | (uintptr_t)p.out | (uintptr_t)p.coef | (uintptr_t)p.invj | (uintptr_t)p.veff |
       (uintptr_t)(DOT ? p.dotw : nullptr)) & 15)
    return IG_EUNSUPPORTED;
  p.nty = (p.ny + TY - 1) / TY;
  p.ntz = 1;
  const long long groups = (p.nblk + NB - 1) / NB;
  if (p.xc <= 0) p.xc = zp_chunk(p.nx);
  if (p.xc > p.nx) p.xc = p.nx;
  p.nxc = (p.nx + p.xc - 1) / p.xc;
  p.units = p.nty * p.nxc;
  if (DOT && (!p.dotw || !p.dpart || !dots || (long long)p.nblk * p.units > p.dparts)) return IG_EINVAL;
  const long long items = (long long)((p.nty + 1) / 2) * p.nxc;
  const long long blocks = p.order == 0 ? 8LL * groups * ((p.units + 7) / 8) : 16LL * groups * ((items + 7) / 8);
  if (blocks <= 0 || blocks > 0x7fffffffLL) return IG_EINVAL;
  if (auto_pick && blocks < 512) return IG_EUNSUPPORTED;
  hipLaunchKernelGGL((monitor_zp_kernel<T, NB, GD, TY, DOT>), dim3((unsigned)blocks), dim3((TY + 4) * 64), 0, stream,
                     p);
  if (DOT)
    hipLaunchKernelGGL(zp_dots_finish, dim3((p.nblk + 255) / 256), dim3(256), 0, stream, p.dpart, p.units, p.nblk,
                       dots);
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
  if (tile >= 9200000) {
    if (closure) return IG_EUNSUPPORTED;
    // rows per tile x 10 + orbitals per workgroup (+ 100: coefficients two planes ahead),
    // + 1000 x block order (0 unit order, 1 y sweep, 2 banded y sweep)
    MonParams<T> q = p;
    q.order = (tile - 9200000) / 1000;
    if (q.order > 2) return IG_EUNSUPPORTED;
    switch ((tile - 9200000) % 1000) {
      case 81: return launch_monitor_zp_t<T, 1, 1, 8>(q, s);
      case 82: return launch_monitor_zp_t<T, 2, 1, 8>(q, s);
      case 83: return launch_monitor_zp_t<T, 3, 1, 8>(q, s);
      case 182: return launch_monitor_zp_t<T, 2, 2, 8>(q, s);
      case 122: return launch_monitor_zp_t<T, 2, 1, 12>(q, s);
      case 43: return launch_monitor_zp_t<T, 3, 1, 4>(q, s);
      case 64: return launch_monitor_zp_t<T, 4, 1, 6>(q, s);
      default: return IG_EUNSUPPORTED;
    }
  }
  if (tile >= 9000000) {
    if (closure) return IG_EUNSUPPORTED;
    // orbitals per workgroup, + 100 x units per interleaved item (0: contiguous ranges)
    MonParams<T> q = p;
    q.ug = (tile - 9000000) / 100;
    if (q.ug > 64) return IG_EUNSUPPORTED;
    switch ((tile - 9000000) % 100) {
      case 1: return launch_monitor_zq_t<T, 1>(q, s);
      case 2: return launch_monitor_zq_t<T, 2>(q, s);
      case 3: return launch_monitor_zq_t<T, 3>(q, s);
      case 4: return launch_monitor_zq_t<T, 4>(q, s);
      default: return IG_EUNSUPPORTED;
    }
  }
  if (tile >= 8000000) {
    if (closure) return IG_EUNSUPPORTED;
    switch (tile - 8000000) {
      case 4: return launch_monitor_rq_t<T, 4>(p, s);
      case 6: return launch_monitor_rq_t<T, 6>(p, s);
      case 8: return launch_monitor_rq_t<T, 8>(p, s);
      case 12: return launch_monitor_rq_t<T, 12>(p, s);
      default: return IG_EUNSUPPORTED;
    }
  }
  if (tile >= 7000000) return closure ? IG_EUNSUPPORTED : launch_monitor_zv<T>(p, tile - 7000000, s);
  // FP32 blocks whose geometry stream (8 fields, 32 B a point) outgrows the 256 MB
  // Infinity Cache: four orbitals per workgroup (general kernel) halve the geometry
  // re-reads and beat the z-vectorised kernel's two (224^3, k = 48: 2248 against
  // 2070 GB/s; at 96^3, where the geometry stays cached, 2368 against 2629).
  const bool fp32_big = sizeof(T) == 4 && p.nblk >= 4 && (long long)p.nx * p.ny * p.nz * 32 > (128LL << 20);
  // Large blocks: the register-queue kernel, whose workgroups carry 6 (FP32) / 4 (FP64)
  // orbitals with one geometry fetch and still fit 3 / 2 to a CU (224^3, k = 48: FP32
  // 1.69 ms against 1.90 ms for the general kernel, FP64 3.22 ms against 3.70 ms for the
  // z-vectorised one).
  // Large blocks (geometry beyond the Infinity Cache): FP32 rows that fit one wavefront
  // take the pipelined full-z kernel in the banded y-sweep order (224^3, k = 48: 1.15 ms;
  // 1.24 ms in unit order against 1.67 ms for the register-queue kernel), FP64 the
  // z-vectorised register-queue kernel with unit pairs dealt round robin to the XCDs
  // (2.74 ms; 3.03 ms in contiguous ranges against 3.22 ms); shapes these decline fall back
  // to the register-queue kernel.
  if (tile == 0 && closure == 0 && !transpose && !p.interior_only &&
      (long long)p.nx * p.ny * p.nz * (long long)sizeof(T) * 8 > (128LL << 20)) {
    int st = IG_EUNSUPPORTED;
    if (sizeof(T) == 4 && p.nblk >= 3 && p.nz >= 128) {
      MonParams<T> q = p;
      q.order = 2;  // banded y sweep: 224^3, k = 48: 1.150 ms against 1.215 ms in unit order
      st = launch_monitor_zp_t<T, 3, 1, 8>(q, s, true);
    }
    if (st == IG_EUNSUPPORTED && sizeof(T) == 8 && p.nblk >= 3) {
      MonParams<T> q = p;
      q.ug = 2;  // pairs of units dealt round robin: 224^3, k = 48: 2.744 ms against 3.037 ms
      st = launch_monitor_zq_t<T, 3>(q, s, true);
    }
    if (st == IG_EUNSUPPORTED && sizeof(T) == 4 && p.nblk >= 6) st = launch_monitor_rq_t<T, 6>(p, s, true);
    else if (st == IG_EUNSUPPORTED && sizeof(T) == 8 && p.nblk >= 4) st = launch_monitor_rq_t<T, 4>(p, s, true);
    if (st != IG_EUNSUPPORTED) return st;
  }
  if (tile == 0 && closure == 0 && !fp32_big) {
    const int st = launch_monitor_zv<T>(p, 0, s);
    if (st != IG_EUNSUPPORTED) return st;
  }
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

// H psi with the row dots out[m] = sum_n w[n] in[m][n] out[m][n] of the same pass (the
// (p, A p) products of a CG step): the full-z kernel, 3 orbitals per workgroup, banded
// y-sweep order; IG_EUNSUPPORTED where that kernel does not apply.
template <typename T>
int monitor_dots_entry(const T* in, T* out, int nblk, long long ld, int nx, int ny, int nz, const T* coef,
                       const T* invj, const T* veff, double scale, const T* w, double* work, long long work_elems,
                       double* dots, hipStream_t stream) {
  if (!in || !out || !coef || !invj || !w || !work || !dots || nblk <= 0 || nx < 3 || ny < 3 || nz < 3)
    return IG_EINVAL;
  if (ld < (long long)nx * ny * nz || work_elems <= 0) return IG_EINVAL;
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
  p.order = 2;
  p.dotw = w;
  p.dpart = work;
  p.dparts = work_elems;
  return launch_monitor_zp_t<T, 3, 1, 8, true>(p, stream, false, dots);
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

long long ig_monitor_stencil_dots_work_size(int nblk, int nx, int ny, int nz) {
  if (nblk <= 0 || nx < 3 || ny < 3 || nz < 3) return -1;
  const int xc = zp_chunk(nx);
  return (long long)nblk * ((ny + 7) / 8) * ((nx + xc - 1) / xc);
}

int ig_monitor_stencil_dots_f64(const double* in, double* out, int nblk, long long ld, int nx, int ny, int nz,
                                const double* coef, const double* invj, const double* veff, double scale,
                                const double* w, double* work, long long work_elems, double* dots,
                                hipStream_t stream) {
  return monitor_dots_entry<double>(in, out, nblk, ld, nx, ny, nz, coef, invj, veff, scale, w, work, work_elems,
                                    dots, stream);
}

int ig_monitor_stencil_dots_f32(const float* in, float* out, int nblk, long long ld, int nx, int ny, int nz,
                                const float* coef, const float* invj, const float* veff, double scale,
                                const float* w, double* work, long long work_elems, double* dots,
                                hipStream_t stream) {
  return monitor_dots_entry<float>(in, out, nblk, ld, nx, ny, nz, coef, invj, veff, scale, w, work, work_elems,
                                   dots, stream);
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
