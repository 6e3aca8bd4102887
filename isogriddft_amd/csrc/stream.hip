// Streaming calibration kernels: the HBM bandwidth this GPU delivers to a plain
// 128-bit copy and to a plain read, the practical ceiling the stencil and the
// row-wise CG kernels are compared with, and the scale of the PMC byte counters
// (a copy of known size).  Not part of the reference's algorithm; no counterpart.
#include "common.h"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Each workgroup walks a contiguous range of 16-byte vectors, UNR vectors per lane in
// flight: all UNR loads are issued before the first store.
template <int UNR>
__global__ __launch_bounds__(IG_THREADS) void stream_copy_kernel(const f32x4* __restrict__ in, f32x4* __restrict__ out,
                                                                 long long nvec, long long per_block) {
  const long long b0 = (long long)blockIdx.x * per_block;
  const long long b1 = b0 + per_block < nvec ? b0 + per_block : nvec;
  for (long long base = b0 + threadIdx.x; base < b1; base += (long long)IG_THREADS * UNR) {
    f32x4 v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const long long i = base + (long long)u * IG_THREADS;
      if (i < b1) v[u] = in[i];
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const long long i = base + (long long)u * IG_THREADS;
      if (i < b1) __builtin_nontemporal_store(v[u], &out[i]);
    }
  }
}

template <int UNR>
__global__ __launch_bounds__(IG_THREADS) void stream_read_kernel(const f32x4* __restrict__ in, long long nvec,
                                                                 long long per_block, double* __restrict__ partial) {
  __shared__ double red[4];
  const long long b0 = (long long)blockIdx.x * per_block;
  const long long b1 = b0 + per_block < nvec ? b0 + per_block : nvec;
  float acc = 0.f;
  for (long long base = b0 + threadIdx.x; base < b1; base += (long long)IG_THREADS * UNR) {
    f32x4 v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const long long i = base + (long long)u * IG_THREADS;
      v[u] = i < b1 ? in[i] : f32x4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
  }
  double s[1] = {(double)acc};
  ig_block_sum<1>(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

int stream_grid(long long nvec, int per_cu, long long* per_block) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  long long blocks = (long long)cus * per_cu;
  if (blocks > nvec) blocks = nvec > 0 ? nvec : 1;
  *per_block = (nvec + blocks - 1) / blocks;
  return (int)blocks;
}

}  // namespace

extern "C" {

// Workgroups launched by ig_stream_read_f32 (the length of its `partial` output).
int ig_stream_blocks(long long n) {
  long long pb = 0;
  return stream_grid(n / 4, 8, &pb);
}

// out[i] = in[i], i < n (n a multiple of 4; 16-byte aligned pointers), non-temporal
// stores, 8 workgroups per CU.
int ig_stream_copy_f32(const float* in, float* out, long long n, hipStream_t stream) {
  if (!in || !out || n <= 0 || (n & 3) || ((uintptr_t)in & 15) || ((uintptr_t)out & 15)) return IG_EINVAL;
  long long pb = 0;
  const int blocks = stream_grid(n / 4, 8, &pb);
  hipLaunchKernelGGL(stream_copy_kernel<4>, dim3(blocks), dim3(IG_THREADS), 0, stream,
                     reinterpret_cast<const f32x4*>(in), reinterpret_cast<f32x4*>(out), n / 4, pb);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

// partial[b] = sum of workgroup b's range of in (FP32 lanes, FP64 block sums);
// partial holds ig_stream_blocks(n) doubles.
int ig_stream_read_f32(const float* in, long long n, double* partial, hipStream_t stream) {
  if (!in || !partial || n <= 0 || (n & 3) || ((uintptr_t)in & 15)) return IG_EINVAL;
  long long pb = 0;
  const int blocks = stream_grid(n / 4, 8, &pb);
  hipLaunchKernelGGL(stream_read_kernel<4>, dim3(blocks), dim3(IG_THREADS), 0, stream,
                     reinterpret_cast<const f32x4*>(in), n / 4, pb, partial);
  IG_CHECK_LAUNCH();
  return IG_OK;
}

}  // extern "C"
