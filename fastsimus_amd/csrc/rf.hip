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
// For every transmit this is one real matrix product U (n x E) = W (n x 2 NF)
// Y (2 NF x E) with W[n][2k] = w cos / N, W[n][2k+1] = w sin / N and Y the real
// and imaginary parts of the band, which runs on the matrix cores in split
// BF16 (hi + lo, hi*hi + hi*lo + lo*hi, FP32 accumulation): k_idft_table
// writes W once per call in MFMA fragment order (the phase b n mod N formed
// exactly in integers), k_idft_mfma stages 8 row tiles of it per 16-bin chunk
// in LDS (double-buffered) and splits the spectrum's fragments in registers.
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

typedef __bf16 rf_bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 rf_bf16x2 __attribute__((ext_vector_type(2)));
typedef float rf_floatx4 __attribute__((ext_vector_type(4)));
typedef float rf_floatx2 __attribute__((ext_vector_type(2)));
typedef unsigned rf_u32x4 __attribute__((ext_vector_type(4)));

constexpr int IT_N = 8;  // row tiles (16 output samples each) per workgroup
constexpr int kMaxGridZ = 65535;  // transmits per launch (grid z)
constexpr int IK = 16;   // bins per K chunk (32 real products per MFMA)


__device__ __forceinline__ rf_floatx4 rf_mma(rf_u32x4 a, rf_u32x4 b, rf_floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(rf_bf16x8, a), __builtin_bit_cast(rf_bf16x8, b), c,
                                                 0, 0, 0);
}

// W fragments [nt][kc][hl][lane] (A operand of v_mfma_f32_16x16x32_bf16: lane -> row lane % 16,
// K slots 8 (lane / 16) + 0..7): row n = 16 nt + lane % 16, bins b = idx0 + 16 kc + 4 (lane / 16) + q
// (q = 0..3), slots (w_b / N) (cos, sin)(2 pi b n / N); zero for n >= nkeep, beyond the band and
// above N / 2.
__global__ void k_idft_table(int NF, int idx0, int N, int nkeep, int ntp, int nkc, rf_u32x4* __restrict__ W) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)ntp * nkc * 64) return;
  const int lane = (int)(i & 63);
  const long r = i >> 6;  // nt * nkc + kc
  const int kc = (int)(r % nkc), nt = (int)(r / nkc);
  const int n = 16 * nt + (lane & 15);
  float v[8];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = IK * kc + 4 * (lane >> 4) + q;
    const int b = idx0 + k;
    float c = 0.f, sn = 0.f;
    if (n < nkeep && k < NF && 2 * b <= N) {
      const float w = bin_weight(b, N) / (float)N;
      const cf tw = twiddle(((long)b * n) % N, N);
      c = w * tw.x;
      sn = w * tw.y;
    }
    v[2 * q] = c;
    v[2 * q + 1] = sn;
  }
  unsigned h[4], l[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) split_bf16_pair(v[2 * q], v[2 * q + 1], h[q], l[q]);
  W[(r * 2 + 0) * 64 + lane] = (rf_u32x4){h[0], h[1], h[2], h[3]};
  W[(r * 2 + 1) * 64 + lane] = (rf_u32x4){l[0], l[1], l[2], l[3]};
}

// u[t][n][e] for one transmit t (blockIdx.z), IT_N row tiles (blockIdx.x) and NW element tiles of 16
// (wave w: elements 16 (NW blockIdx.y + w) + 0..15). B operand: lane -> element lane % 16, bins
// 16 kc + 4 (lane / 16) + 0..3 as (Re, Im) pairs, matching W's slots.
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_idft_mfma(const cf* __restrict__ spect, int NF, int E, int nkeep, int nkc,
                                                       const rf_u32x4* __restrict__ W, float* __restrict__ u,
                                                       unsigned* __restrict__ peak) {
  constexpr int ENT = IT_N * 2 * 64;   // W fragments per chunk per workgroup
  constexpr int PER = ENT / (64 * NW);  // of them per thread
  __shared__ rf_u32x4 wa[2][ENT];
  __shared__ float wmax[NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nt0 = blockIdx.x * IT_N;
  const int e = (blockIdx.y * NW + wave) * 16 + (lane & 15);
  const int t = blockIdx.z;
  const bool ev = e < E;
  const cf* X = spect + (size_t)t * NF * E + (ev ? e : 0);
  const rf_u32x4* Wb = W + (size_t)nt0 * nkc * 128;
  rf_u32x4 pre[PER];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int q = threadIdx.x + j * 64 * NW;  // (row tile q / 128, hi|lo and lane q % 128)
      pre[j] = Wb[((size_t)(q >> 7) * nkc + kc) * 128 + (q & 127)];
    }
  };
  auto stash = [&](int b) {
#pragma unroll
    for (int j = 0; j < PER; ++j) wa[b][threadIdx.x + j * 64 * NW] = pre[j];
  };
  rf_floatx4 acc[IT_N];
#pragma unroll
  for (int i = 0; i < IT_N; ++i) acc[i] = (rf_floatx4){0.f, 0.f, 0.f, 0.f};
  fetch(0);
  stash(0);
  __syncthreads();
  for (int kc = 0; kc < nkc; ++kc) {
    const int b = kc & 1;
    if (kc + 1 < nkc) fetch(kc + 1);
    float y[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = IK * kc + 4 * (lane >> 4) + q;
      const cf x = (ev && k < NF) ? X[(size_t)k * E] : cf{0.f, 0.f};
      y[2 * q] = x.x;
      y[2 * q + 1] = x.y;
    }
    unsigned h[4], l[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) split_bf16_pair(y[2 * q], y[2 * q + 1], h[q], l[q]);
    const rf_u32x4 bh = {h[0], h[1], h[2], h[3]}, bl = {l[0], l[1], l[2], l[3]};
#pragma unroll
    for (int i = 0; i < IT_N; ++i) {
      const rf_u32x4 ah = wa[b][(i * 2 + 0) * 64 + lane], al = wa[b][(i * 2 + 1) * 64 + lane];
      acc[i] = rf_mma(ah, bh, acc[i]);
      acc[i] = rf_mma(ah, bl, acc[i]);
      acc[i] = rf_mma(al, bh, acc[i]);
    }
    if (kc + 1 < nkc) stash(b ^ 1);  // the other buffer: every wave finished reading it before the last barrier
    __syncthreads();
  }
  float m = 0.f;
  if (ev) {
#pragma unroll
    for (int i = 0; i < IT_N; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = 16 * (nt0 + i) + 4 * (lane >> 4) + r;
        if (n < nkeep) {
          u[((size_t)t * nkeep + n) * E + e] = acc[i][r];
          m = fmaxf(m, fabsf(acc[i][r]));
        }
      }
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (lane == 0) wmax[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = 0.f;
    for (int w = 0; w < NW; ++w) b = fmaxf(b, wmax[w]);
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
// Workspace of spectrum_to_rf: the per-transmit peak words, then the W fragments.
struct RfLayout {
  int nkeep, ntp, nkc;
  size_t w_offset, bytes;
};
static RfLayout rf_layout(int NT, int NF, int N) {
  RfLayout L;
  L.nkeep = (N + 1) / 2;
  L.ntp = (L.nkeep + 16 * IT_N - 1) / (16 * IT_N) * IT_N;
  L.nkc = (NF + IK - 1) / IK;
  L.w_offset = ((size_t)NT * sizeof(unsigned) + 255) / 256 * 256;
  L.bytes = L.w_offset + (size_t)L.ntp * L.nkc * 128 * sizeof(rf_u32x4);
  return L;
}
size_t rf_workspace(int NT, int NF, int N) { return rf_layout(NT, NF, N).bytes; }

int spectrum_to_rf(const cf* spect, int NT, int NF, int E, int nff, int idx0, int N, cf* full, float* rf,
                   float* rf_u, char* ws, hipStream_t st) {
  const RfLayout L = rf_layout(NT, NF, N);
  const int nkeep = L.nkeep;
  unsigned* peak = (unsigned*)ws;
  rf_u32x4* W = (rf_u32x4*)(ws + L.w_offset);
  hipLaunchKernelGGL(k_fill_u32, dim3((NT + 255) / 256), dim3(256), 0, st, peak, NT, 0u);
  if (full) {
    const long n = (long)NT * nff * E;
    hipLaunchKernelGGL(k_place, dim3((int)std::min<long>((n + 255) / 256, 65536)), dim3(256), 0, st, spect, NT, NF, E,
                       nff, idx0, full);
  }
  float* u = rf_u ? rf_u : rf;
  {
    const long n = (long)L.ntp * L.nkc * 64;
    hipLaunchKernelGGL(k_idft_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, NF, idx0, N, nkeep, L.ntp,
                       L.nkc, W);
  }
  // transmits on grid z, in launches of at most kMaxGridZ
  for (int t0 = 0; t0 < NT; t0 += kMaxGridZ) {
    const int nt = std::min(kMaxGridZ, NT - t0);
    const cf* sp = spect + (size_t)t0 * NF * E;
    float* ut = u + (size_t)t0 * nkeep * E;
    if (E > 64) {
      dim3 grid(L.ntp / IT_N, (E + 127) / 128, nt);
      hipLaunchKernelGGL(k_idft_mfma<8>, grid, dim3(512), 0, st, sp, NF, E, nkeep, L.nkc, W, ut, peak + t0);
    } else {
      dim3 grid(L.ntp / IT_N, (E + 63) / 64, nt);
      hipLaunchKernelGGL(k_idft_mfma<4>, grid, dim3(256), 0, st, sp, NF, E, nkeep, L.nkc, W, ut, peak + t0);
    }
  }
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
    for (int t0 = 0; t0 < NT; t0 += kMaxGridZ) {  // transmits on grid y, in launches of at most kMaxGridZ
      const int nt = std::min(kMaxGridZ, NT - t0);
      const size_t o = (size_t)t0 * per;
      dim3 g(nblk, nt);
      hipLaunchKernelGGL(k_absmax, g, dim3(256), 0, st, u + o, per, peak + t0);
      hipLaunchKernelGGL(k_gate_adj1, g, dim3(256), 0, st, u + o, dy + o, peak + t0, per, du + o, argm + t0,
                         psum + (size_t)t0 * nblk);
    }
    hipLaunchKernelGGL(k_gate_adj2, dim3((NT + 255) / 256), dim3(256), 0, st, u, argm, psum, nblk, per, NT, du);
  }
  for (int t0 = 0; t0 < NT; t0 += kMaxGridZ) {
    const int nt = std::min(kMaxGridZ, NT - t0);
    dim3 grid((NF + RT - 1) / RT, (E + RE - 1) / RE, nt);
    hipLaunchKernelGGL(k_dft_adj, grid, dim3(256), 0, st, du + (size_t)t0 * nkeep * E, NF, E, idx0, N, nkeep,
                       dspect + (size_t)t0 * NF * E);
  }
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

}  // namespace fs
