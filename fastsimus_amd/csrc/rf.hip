// Spectrum -> RF (reference: _irfft_and_threshold, simus.py:336-367) and its adjoint.
//
// Only the selected band [idx0, idx0 + NF) of the n_ff-bin spectrum is
// non-zero, and n_fft is often not a power of two (e.g. 4612 = 4 * 1153), so
// the inverse real DFT is evaluated directly over the band:
//
//   u[n] = (1/N) sum_{k in band, k <= N/2} w_k (Re X_k cos(2 pi k n / N) + Im X_k sin(2 pi k n / N))
//
// (irfft of conj(X): bins above N/2 are cropped, the imaginary parts of the DC
// and Nyquist bins are ignored, w_k = 1 at DC / Nyquist and 2 elsewhere).
// A 64(n) x 32(e) output tile per workgroup: the band is staged through LDS in
// chunks of 64 bins, all lanes of a wave read the same 8 elements (LDS
// broadcast), and each lane advances its own twiddle exp(i 2 pi n / N) by an
// FP32 recurrence reseeded from an exact integer (k n mod N) at every chunk.
// The -100 dB smooth gate needs max|u| per transmit: every workgroup folds its
// tile maximum into one word with an integer atomicMax on the float bits
// (order independent, so the result is deterministic).
#include <algorithm>

#include "fs_device.h"

namespace fs {

constexpr int RT = 64;  // output samples per tile (one per lane)
constexpr int RE = 32;  // elements per tile (8 per wave)
constexpr int RK = 64;  // bins per LDS chunk

__global__ void k_fill_u32(unsigned* p, int n, unsigned v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_zero_words(unsigned* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

int fill_zero(void* ptr, size_t n_words, hipStream_t st) {
  if (n_words == 0) return FS_OK;
  const int blocks = (int)std::min<size_t>((n_words + 255) / 256, 65536);
  hipLaunchKernelGGL(k_zero_words, dim3(blocks), dim3(256), 0, st, (unsigned*)ptr, n_words);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

__global__ void k_place(const cf* __restrict__ spect, int NT, int NF, int E, int nff, int idx0, cf* __restrict__ full) {
  const long n = (long)NT * nff * E;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int e = (int)(i % E);
    const long r = i / E;
    const int b = (int)(r % nff);
    const int t = (int)(r / nff);
    const int k = b - idx0;
    full[i] = (k >= 0 && k < NF) ? spect[((size_t)t * NF + k) * E + e] : cf{0.f, 0.f};
  }
}

__device__ __forceinline__ float bin_weight(int k, int N) {
  if (k > N / 2) return 0.f;
  if (k == 0 || (2 * k == N)) return 1.f;
  return 2.f;
}

__device__ __forceinline__ cf twiddle(long m, int N) {
  // exp(i 2 pi m / N) for 0 <= m < N, angle formed in FP64 then reduced once.
  const float fr = (float)((double)m / (double)N);
  float s, c;
  sincospif(2.0f * fr, &s, &c);
  return {c, s};
}

__global__ __launch_bounds__(256) void k_idft(const cf* __restrict__ spect, int NF, int E, int idx0, int N, int nkeep,
                                              float* __restrict__ rf, unsigned* __restrict__ peak) {
  __shared__ cf xs[RK][RE];
  __shared__ float wmax[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = blockIdx.x * RT + lane;
  const int e0 = blockIdx.y * RE;
  const int t = blockIdx.z;
  const cf* X = spect + (size_t)t * NF * E;
  const float invN = 1.0f / (float)N;
  const cf st = twiddle(n % N, N);
  float acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = 0.f;

  for (int i0 = 0; i0 < NF; i0 += RK) {
    __syncthreads();
    for (int q = threadIdx.x; q < RK * RE; q += 256) {
      const int kk = q / RE, ee = q % RE;
      const int i = i0 + kk, e = e0 + ee;
      cf v = {0.f, 0.f};
      if (i < NF && e < E) {
        const float w = bin_weight(idx0 + i, N) * invN;
        const cf x = X[(size_t)i * E + e];
        v = {x.x * w, x.y * w};
      }
      xs[kk][ee] = v;
    }
    __syncthreads();
    cf tw = twiddle(((long)(idx0 + i0) * n) % N, N);
    const int kmax = min(RK, NF - i0);
    for (int kk = 0; kk < kmax; ++kk) {
      const cf* row = &xs[kk][wave * 8];
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] = fmaf(row[j].x, tw.x, fmaf(row[j].y, tw.y, acc[j]));
      tw = cmul(tw, st);
    }
  }
  float m = 0.f;
  if (n < nkeep) {
    float* out = rf + ((size_t)t * nkeep + n) * E;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = e0 + wave * 8 + j;
      if (e < E) {
        out[e] = acc[j];
        m = fmaxf(m, fabsf(acc[j]));
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (lane == 0) wmax[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float b = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    atomicMax(peak + t, __float_as_uint(b));
  }
}

__device__ __forceinline__ float gate_of(float a) { return 0.5f * (1.0f + tanhf((a - 1e-5f) / 1e-6f)); }

__global__ void k_gate(const float* __restrict__ u, const unsigned* __restrict__ peak, int per_tx, int NT,
                       float* __restrict__ y) {
  const long n = (long)per_tx * NT;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int t = (int)(i / per_tx);
    const float P = __uint_as_float(peak[t]);
    const float v = u[i];
    y[i] = v * gate_of(fabsf(v) / (P + 1e-30f));
  }
}

// ---------------------------------------------------------------- adjoint of the gate
// y_i = u_i g(a_i), a_i = |u_i| / (P + eps), P = max_j |u_j| attained at index m:
//   du_i = dy_i (g_i + |u_i| g'_i / (P + eps)) + [i == m] sign(u_m) sum_j dy_j u_j g'_j (-|u_j| / (P + eps)^2)
__device__ __forceinline__ float gate_deriv(float a) {
  const float th = tanhf((a - 1e-5f) / 1e-6f);
  return 0.5f * (1.0f - th * th) / 1e-6f;
}

__global__ void k_absmax(const float* __restrict__ u, int per, unsigned* __restrict__ peak) {
  const int t = blockIdx.y;
  float m = 0.f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < per; i += gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(u[(size_t)t * per + i]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(peak + t, __float_as_uint(m));
}

// first index attaining the peak (atomicMin over candidates) and the per-block partial sums of the P term
__global__ void k_gate_adj1(const float* __restrict__ u, const float* __restrict__ dy, const unsigned* __restrict__ peak,
                            int per_tx, float* __restrict__ du, unsigned* __restrict__ argm, float* __restrict__ psum) {
  __shared__ float red[256];
  const int t = blockIdx.y;
  const float P = __uint_as_float(peak[t]);
  const float inv = 1.0f / (P + 1e-30f);
  float s = 0.f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < per_tx; i += gridDim.x * blockDim.x) {
    const size_t g = (size_t)t * per_tx + i;
    const float v = u[g], d = dy[g];
    const float a = fabsf(v) * inv;
    const float gp = gate_deriv(a);
    du[g] = d * (gate_of(a) + fabsf(v) * gp * inv);
    s += d * v * gp * (-fabsf(v) * inv * inv);
    if (fabsf(v) == P) atomicMin(argm + t, (unsigned)i);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) psum[(size_t)t * gridDim.x + blockIdx.x] = red[0];
}

__global__ void k_gate_adj2(const float* __restrict__ u, const unsigned* __restrict__ argm, const float* __restrict__ psum,
                            int nblk, int per_tx, int NT, float* __restrict__ du) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NT) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += psum[(size_t)t * nblk + b];
  const unsigned m = argm[t];
  if (m < (unsigned)per_tx) {
    const size_t g = (size_t)t * per_tx + m;
    const float v = u[g];
    du[g] += (v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f)) * s;
  }
}

// d spect[t][i][e] = (w_k / N) sum_{n < nkeep} du[t][n][e] exp(i 2 pi k n / N), k = idx0 + i
__global__ __launch_bounds__(256) void k_dft_adj(const float* __restrict__ du, int NF, int E, int idx0, int N, int nkeep,
                                                 cf* __restrict__ dspect) {
  __shared__ float us[RK][RE];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * RT + lane;  // band index
  const int e0 = blockIdx.y * RE;
  const int t = blockIdx.z;
  const int k = idx0 + i;
  const cf st = twiddle(((long)k) % N, N);
  cf acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = {0.f, 0.f};
  const float* U = du + (size_t)t * nkeep * E;
  for (int n0 = 0; n0 < nkeep; n0 += RK) {
    __syncthreads();
    for (int q = threadIdx.x; q < RK * RE; q += 256) {
      const int nn = q / RE, ee = q % RE;
      const int n = n0 + nn, e = e0 + ee;
      us[nn][ee] = (n < nkeep && e < E) ? U[(size_t)n * E + e] : 0.f;
    }
    __syncthreads();
    cf tw = twiddle(((long)k * n0) % N, N);
    const int nmax = min(RK, nkeep - n0);
    for (int nn = 0; nn < nmax; ++nn) {
      const float* row = &us[nn][wave * 8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        acc[j].x = fmaf(row[j], tw.x, acc[j].x);
        acc[j].y = fmaf(row[j], tw.y, acc[j].y);
      }
      tw = cmul(tw, st);
    }
  }
  if (i >= NF) return;
  const float w = bin_weight(k, N) / (float)N;
  cf* out = dspect + ((size_t)t * NF + i) * E;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int e = e0 + wave * 8 + j;
    if (e < E) out[e] = {acc[j].x * w, acc[j].y * w};
  }
}

// ---------------------------------------------------------------- host side
int spectrum_to_rf(const cf* spect, int NT, int NF, int E, int nff, int idx0, int N, cf* full, float* rf,
                   float* rf_u, unsigned* peak, hipStream_t st) {
  const int nkeep = (N + 1) / 2;
  hipLaunchKernelGGL(k_fill_u32, dim3((NT + 255) / 256), dim3(256), 0, st, peak, NT, 0u);
  if (full) {
    const long n = (long)NT * nff * E;
    hipLaunchKernelGGL(k_place, dim3((int)std::min<long>((n + 255) / 256, 65536)), dim3(256), 0, st, spect, NT, NF, E,
                       nff, idx0, full);
  }
  float* u = rf_u ? rf_u : rf;
  dim3 grid((nkeep + RT - 1) / RT, (E + RE - 1) / RE, NT);
  hipLaunchKernelGGL(k_idft, grid, dim3(256), 0, st, spect, NF, E, idx0, N, nkeep, u, peak);
  const long n = (long)nkeep * E * NT;
  hipLaunchKernelGGL(k_gate, dim3((int)std::min<long>((n + 255) / 256, 65536)), dim3(256), 0, st, u, peak, nkeep * E,
                     NT, rf);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

size_t rf_adjoint_ws(int NT, int nkeep, int E) {
  const size_t per = (size_t)nkeep * E;
  return ((size_t)NT * 3 * sizeof(unsigned) + 255) / 256 * 256 + (size_t)NT * 256 * sizeof(float) +
         ((per * NT * sizeof(float) + 255) / 256) * 256;
}

int rf_adjoint(const float* u, const float* dy, int NT, int NF, int E, int idx0, int N, cf* dspect, char* ws,
               hipStream_t st) {
  const int nkeep = (N + 1) / 2;
  const int per = nkeep * E;
  unsigned* peak = (unsigned*)ws;
  unsigned* argm = peak + NT;
  const size_t o1 = ((size_t)NT * 3 * sizeof(unsigned) + 255) / 256 * 256;
  float* psum = (float*)(ws + o1);
  float* du = psum + (size_t)NT * 256;
  hipLaunchKernelGGL(k_fill_u32, dim3((NT + 255) / 256), dim3(256), 0, st, peak, NT, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3((NT + 255) / 256), dim3(256), 0, st, argm, NT, 0xffffffffu);
  // recompute the peak of the pre-gate RF (max of |u| per transmit)
  {
    const int nblk = 256;
    dim3 g(nblk, NT);
    hipLaunchKernelGGL(k_absmax, g, dim3(256), 0, st, u, per, peak);
    hipLaunchKernelGGL(k_gate_adj1, g, dim3(256), 0, st, u, dy, peak, per, du, argm, psum);
    hipLaunchKernelGGL(k_gate_adj2, dim3((NT + 255) / 256), dim3(256), 0, st, u, argm, psum, nblk, per, NT, du);
  }
  dim3 grid((NF + RT - 1) / RT, (E + RE - 1) / RE, NT);
  hipLaunchKernelGGL(k_dft_adj, grid, dim3(256), 0, st, du, NF, E, idx0, N, nkeep, dspect);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

}  // namespace fs
