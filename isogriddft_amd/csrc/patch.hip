// Near-core local-GTH patch correction (pseudo/local.py:379-703).
//
// Per atom the reference samples the density on a fine Cartesian patch (a 25^3
// lattice clipped to the sphere r <= R) by inverse-distance-squared interpolation
// from the k nearest main-grid nodes, integrates rho V_loc there, compares with the
// main-grid quadrature over the same sphere and embeds the difference as a frozen
// correction field.  The neighbour sets depend on the geometry only, so they are
// found once (exact brute force, ig_knn_f64) and every density evaluation is a
// gather + reduction (ig_idw_quadrature_f64).
//
//  ig_knn_f64            k nearest grid nodes of each query point, ordered by
//                        (distance, node index).  Queries x node chunks: each lane owns
//                        one query and keeps its best KC (distance^2, index) pairs in
//                        registers; node coordinates are staged through LDS and read
//                        as broadcasts.  A second pass merges the per-chunk lists.
//  ig_idw_quadrature_f64 rho at the patch points (exact-hit rule and normalised
//                        1/d^2 weights as pseudo/local.py:444-458) and sum rho V.
//  ig_patch_main_f64     sum over r <= R of rho V w and of rho w on the main grid.
//  ig_patch_field_f64    out += scale V (or += scale) on r <= R.
#include "common.h"
#include "gth.h"

#pragma clang fp contract(off)

namespace {

constexpr int KNN_TILE = 256;  // node coordinates staged per LDS pass
constexpr int KNN_NO_INDEX = 0x7fffffff;

// Strict total order of candidates: distance^2, then node index.
__device__ __forceinline__ bool knn_before(double d, int i, double e, int j) { return d < e || (d == e && i < j); }

// Insert (d, i) into the ascending list (bd, bi); the displaced last entry drops out.
template <int KC>
__device__ __forceinline__ void knn_insert(double (&bd)[KC], int (&bi)[KC], double d, int i) {
  if (!knn_before(d, i, bd[KC - 1], bi[KC - 1])) return;
#pragma unroll
  for (int j = 0; j < KC; ++j) {
    if (knn_before(d, i, bd[j], bi[j])) {
      const double td = bd[j];
      const int ti = bi[j];
      bd[j] = d;
      bi[j] = i;
      d = td;
      i = ti;
    }
  }
}

// grid (ceil(nq / 256), nchunk): best KC nodes of node range [y * chunk, (y + 1) * chunk).
template <int KC>
__global__ __launch_bounds__(IG_THREADS) void knn_partial_kernel(const double* __restrict__ xyz, long long N,
                                                                 const double* __restrict__ q, int nq, long long chunk,
                                                                 double* __restrict__ cd, int* __restrict__ ci) {
  __shared__ double sx[KNN_TILE], sy[KNN_TILE], sz[KNN_TILE];
  const int t = blockIdx.x * IG_THREADS + threadIdx.x;
  const bool active = t < nq;
  double qx = 0.0, qy = 0.0, qz = 0.0;
  if (active) {
    qx = q[t];
    qy = q[nq + t];
    qz = q[2 * (long long)nq + t];
  }
  double bd[KC];
  int bi[KC];
#pragma unroll
  for (int j = 0; j < KC; ++j) {
    bd[j] = __builtin_huge_val();
    bi[j] = KNN_NO_INDEX;
  }
  const long long lo = (long long)blockIdx.y * chunk;
  const long long hi = lo + chunk < N ? lo + chunk : N;
  for (long long base = lo; base < hi; base += KNN_TILE) {
    __syncthreads();
    const long long p = base + threadIdx.x;
    if (p < hi) {
      sx[threadIdx.x] = xyz[p];
      sy[threadIdx.x] = xyz[N + p];
      sz[threadIdx.x] = xyz[2 * N + p];
    }
    __syncthreads();
    const int n = hi - base < KNN_TILE ? (int)(hi - base) : KNN_TILE;
    if (active) {
      for (int j = 0; j < n; ++j) {
        const double dx = sx[j] - qx, dy = sy[j] - qy, dz = sz[j] - qz;
        knn_insert<KC>(bd, bi, dx * dx + dy * dy + dz * dz, (int)(base + j));
      }
    }
  }
  if (active) {
    const long long o = ((long long)blockIdx.y * nq + t) * KC;
#pragma unroll
    for (int j = 0; j < KC; ++j) {
      cd[o + j] = bd[j];
      ci[o + j] = bi[j];
    }
  }
}

// One lane per query: merge the nchunk lists, emit the first k (index, distance).
template <int KC>
__global__ __launch_bounds__(IG_THREADS) void knn_merge_kernel(const double* __restrict__ cd, const int* __restrict__ ci,
                                                               int nq, int nchunk, int k, int* __restrict__ idx,
                                                               double* __restrict__ dist) {
  const int t = blockIdx.x * IG_THREADS + threadIdx.x;
  if (t >= nq) return;
  double bd[KC];
  int bi[KC];
#pragma unroll
  for (int j = 0; j < KC; ++j) {
    bd[j] = __builtin_huge_val();
    bi[j] = KNN_NO_INDEX;
  }
  for (int c = 0; c < nchunk; ++c) {
    const long long o = ((long long)c * nq + t) * KC;
    for (int j = 0; j < KC; ++j) knn_insert<KC>(bd, bi, cd[o + j], ci[o + j]);
  }
#pragma unroll
  for (int j = 0; j < KC; ++j) {
    if (j < k) {
      idx[(long long)t * k + j] = bi[j];
      dist[(long long)t * k + j] = sqrt(bd[j]);
    }
  }
}

template <int KC>
int knn_t(const double* xyz, long long N, const double* q, int nq, int k, double* work, long long work_bytes,
          int* idx, double* dist, hipStream_t s) {
  const int qb = (nq + IG_THREADS - 1) / IG_THREADS;
  // about 2048 workgroups over the chip, chunks of at least 4 LDS tiles
  long long nchunk = (2048 + qb - 1) / qb;
  const long long min_chunk = 4LL * KNN_TILE;
  if (nchunk * min_chunk > N) nchunk = (N + min_chunk - 1) / min_chunk;
  if (nchunk < 1) nchunk = 1;
  if (nchunk > 65535) nchunk = 65535;
  const long long per = (long long)nq * KC;
  while (nchunk > 1 && nchunk * per * (long long)(sizeof(double) + sizeof(int)) > work_bytes) --nchunk;
  if (nchunk * per * (long long)(sizeof(double) + sizeof(int)) > work_bytes) return IG_EINVAL;
  const long long chunk = (N + nchunk - 1) / nchunk;
  double* cd = work;
  int* ci = (int*)(work + nchunk * per);
  hipLaunchKernelGGL(knn_partial_kernel<KC>, dim3(qb, (unsigned)nchunk), dim3(IG_THREADS), 0, s, xyz, N, q, nq, chunk,
                     cd, ci);
  hipLaunchKernelGGL(knn_merge_kernel<KC>, dim3(qb), dim3(IG_THREADS), 0, s, cd, ci, nq, (int)nchunk, k, idx, dist);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

// values[t] = rho interpolated at query t; partial[block] = sum_t values[t] vq[t].
__global__ __launch_bounds__(IG_THREADS) void idw_kernel(const double* __restrict__ field, const int* __restrict__ idx,
                                                         const double* __restrict__ dist, int nq, int k,
                                                         const double* __restrict__ vq, double* __restrict__ values,
                                                         double* __restrict__ partial) {
  __shared__ double smem[4];
  double acc[1] = {0.0};
  for (int t = blockIdx.x * IG_THREADS + threadIdx.x; t < nq; t += gridDim.x * IG_THREADS) {
    const int* id = idx + (long long)t * k;
    const double* d = dist + (long long)t * k;
    double v;
    if (d[0] <= 1.0e-12) {
      v = field[id[0]];
    } else {
      double ws = 0.0;
      for (int j = 0; j < k; ++j) {
        const double dj = d[j] > 1.0e-12 ? d[j] : 1.0e-12;
        ws += 1.0 / (dj * dj);
      }
      v = 0.0;
      for (int j = 0; j < k; ++j) {
        const double dj = d[j] > 1.0e-12 ? d[j] : 1.0e-12;
        v += field[id[j]] * ((1.0 / (dj * dj)) / ws);
      }
    }
    if (values) values[t] = v;
    acc[0] += v * vq[t];
  }
  ig_block_sum<1>(acc, smem);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

// partial[block] = {sum_{r<=R} rho V w, sum_{r<=R} rho w, count}.
__global__ __launch_bounds__(IG_THREADS) void patch_main_kernel(const double* __restrict__ xyz,
                                                                const double* __restrict__ rho,
                                                                const double* __restrict__ w, long long N, GthAtom a,
                                                                double R, double* __restrict__ partial) {
  __shared__ double smem[4 * 3];
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    const double dx = xyz[t] - a.x, dy = xyz[N + t] - a.y, dz = xyz[2 * N + t] - a.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r <= R) {
      const double rw = rho[t] * w[t];
      acc[0] += rho[t] * gth_vloc(r, a) * w[t];
      acc[1] += rw;
      acc[2] += 1.0;
    }
  }
  ig_block_sum<3>(acc, smem);
  if (threadIdx.x == 0) {
    partial[blockIdx.x * 3 + 0] = acc[0];
    partial[blockIdx.x * 3 + 1] = acc[1];
    partial[blockIdx.x * 3 + 2] = acc[2];
  }
}

__global__ __launch_bounds__(IG_THREADS) void patch_field_kernel(const double* __restrict__ xyz, long long N, GthAtom a,
                                                                 double R, int constant, double scale,
                                                                 double* __restrict__ out) {
  for (long long t = (long long)blockIdx.x * IG_THREADS + threadIdx.x; t < N; t += (long long)gridDim.x * IG_THREADS) {
    const double dx = xyz[t] - a.x, dy = xyz[N + t] - a.y, dz = xyz[2 * N + t] - a.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r <= R) out[t] += constant ? scale : scale * gth_vloc(r, a);
  }
}

__global__ void patch_sum_kernel(const double* __restrict__ partial, int nparts, int nv, double* __restrict__ out) {
  for (int v = threadIdx.x; v < nv; v += blockDim.x) {
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += partial[(long long)p * nv + v];
    out[v] = s;
  }
}

inline int blocks_for(long long n, int cap) {
  long long b = (n + IG_THREADS * 4 - 1) / (IG_THREADS * 4);
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

}  // namespace

extern "C" {

long long ig_knn_work_bytes(int nq, int k) {
  if (nq <= 0 || k <= 0 || k > 32) return -1;
  const int kc = k <= 8 ? 8 : k <= 16 ? 16 : 32;
  return 64LL * nq * kc * (long long)(sizeof(double) + sizeof(int));
}

int ig_knn_f64(const double* xyz, long long N, const double* queries, int nq, int k, void* work,
               long long work_bytes, int* idx, double* dist, hipStream_t s) {
  if (!xyz || !queries || !work || !idx || !dist || N <= 0 || nq <= 0 || k <= 0 || k > 32 || k > N ||
      N > 0x7ffffffe)
    return IG_EINVAL;
  if (k <= 8) return knn_t<8>(xyz, N, queries, nq, k, (double*)work, work_bytes, idx, dist, s);
  if (k <= 16) return knn_t<16>(xyz, N, queries, nq, k, (double*)work, work_bytes, idx, dist, s);
  return knn_t<32>(xyz, N, queries, nq, k, (double*)work, work_bytes, idx, dist, s);
}

int ig_idw_quadrature_f64(const double* field, const int* idx, const double* dist, int nq, int k, const double* vq,
                          double* values, double* work, double* out, hipStream_t s) {
  if (!field || !idx || !dist || !vq || !work || !out || nq <= 0 || k <= 0) return IG_EINVAL;
  const int nb = blocks_for(nq, 1024);
  hipLaunchKernelGGL(idw_kernel, dim3(nb), dim3(IG_THREADS), 0, s, field, idx, dist, nq, k, vq, values, work);
  hipLaunchKernelGGL(patch_sum_kernel, dim3(1), dim3(64), 0, s, work, nb, 1, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_patch_main_f64(const double* xyz, const double* rho, const double* w, long long N, const double* atom,
                      double radius, double* work, double* out, hipStream_t s) {
  if (!xyz || !rho || !w || !atom || !work || !out || N <= 0) return IG_EINVAL;
  const int nb = blocks_for(N, 1024);
  hipLaunchKernelGGL(patch_main_kernel, dim3(nb), dim3(IG_THREADS), 0, s, xyz, rho, w, N, *(const GthAtom*)atom,
                     radius, work);
  hipLaunchKernelGGL(patch_sum_kernel, dim3(1), dim3(64), 0, s, work, nb, 3, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

int ig_patch_field_f64(const double* xyz, long long N, const double* atom, double radius, int constant, double scale,
                       double* out, hipStream_t s) {
  if (!xyz || !atom || !out || N <= 0) return IG_EINVAL;
  long long b = (N + IG_THREADS - 1) / IG_THREADS;
  if (b > 16384) b = 16384;
  hipLaunchKernelGGL(patch_field_kernel, dim3((unsigned)b), dim3(IG_THREADS), 0, s, xyz, N, *(const GthAtom*)atom,
                     radius, constant, scale, out);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

}  // extern "C"
