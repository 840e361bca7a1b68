// Device side of the frequency plan: the largest point-to-source distance,
// which fixes the frequency step of simus (element centres,
// src/fast_simus/simus.py:219-222) and pfield (sub-element centroids, clipped
// below at c / (2 fc) by the caller; src/fast_simus/pfield.py:271-282).
//
// One pass over the points in FP64; each thread keeps its own maximum over a
// grid-stride range of points and all sources (staged in LDS), the workgroup
// reduces it, and one atomic maximum on the bit pattern of the non-negative
// double (monotone in the value, so the result does not depend on the order)
// folds the workgroups together. The result is one double in device memory:
// the caller reads it with one synchronisation.
#include <algorithm>

#include "fs_device.h"

namespace fs {

constexpr int MD_SRC = 1024;  // sources staged in LDS per pass

__global__ void k_zero_u64(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

__global__ __launch_bounds__(256) void k_max_dist(const double* __restrict__ x, const double* __restrict__ z,
                                                  long long n, const double* __restrict__ sx,
                                                  const double* __restrict__ sz, int m,
                                                  unsigned long long* __restrict__ out) {
  __shared__ double lx[MD_SRC], lz[MD_SRC];
  __shared__ double red[256];
  double best = 0.0;
  for (int m0 = 0; m0 < m; m0 += MD_SRC) {
    const int mc = min(MD_SRC, m - m0);
    __syncthreads();
    for (int i = threadIdx.x; i < mc; i += blockDim.x) {
      lx[i] = sx[m0 + i];
      lz[i] = sz[m0 + i];
    }
    __syncthreads();
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
      const double px = x[p], pz = z[p];
      double b2 = 0.0;
      for (int i = 0; i < mc; ++i) {
        const double dx = px - lx[i], dz = pz - lz[i];
        b2 = fmax(b2, fma(dx, dx, dz * dz));
      }
      best = fmax(best, sqrt(b2));
    }
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

}  // namespace fs

using namespace fs;

extern "C" int fs_max_distance(const double* x, const double* z, long long n_points, const double* src_x,
                               const double* src_z, int n_src, double* out, void* stream) {
  if (!out || n_points < 0 || n_src <= 0 || !src_x || !src_z) return FS_ERR_INVALID;
  if (n_points > 0 && (!x || !z)) return FS_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  unsigned long long* o = (unsigned long long*)out;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, o);
  if (n_points > 0) {
    const long long blocks = std::min<long long>((n_points + 255) / 256, 4096);
    hipLaunchKernelGGL(k_max_dist, dim3((unsigned)blocks), dim3(256), 0, st, x, z, n_points, src_x, src_z, n_src, o);
  }
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}
