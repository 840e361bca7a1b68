// Consumers of the RF stack: IQ demodulation and delay-and-sum beamforming.
//
// The reference's multi-firing harness (tests/_picmus_phantom_shared.py:407-460) turns every
// firing's RF into IQ (PyMUST rf2iq) and beamforms it onto a pixel grid (PyMUST dasmtx), summing the
// firings. Both run here on the device, on the rf[t][n][e] layout fs_spectrum_to_rf writes.
//
// rf2iq: IQ = 2 LPF(rf(t) exp(-i 2 pi fc t)), the low-pass being a Butterworth filter run forward and
// backward (zero phase) over the signal extended at both ends by its odd reflection, each pass
// starting from the filter's steady state for the first value it sees (the algorithm of
// filtfilt). The filter comes as cascaded second-order sections designed on the host; one thread
// filters one (firing, channel) sequence, reading and writing sample rows coalesced across channels.
//
// das: for pixel p and firing t, the transmit wave arrives at tau_tx = min_e(delay_e + |p - x_e| / c)
// and the echo reaches element e at tau = tau_tx + |p - x_e| / c; the IQ sample at (tau - t0) fs is
// interpolated linearly, rotated by exp(i 2 pi fc tau) and summed over the elements inside the
// receive aperture |x_p - x_e| <= z_p / (2 F) (F the f-number), and over the firings.
#include <algorithm>

#include "fs_device.h"

namespace fs {

// ---------------------------------------------------------------- IQ demodulation
constexpr int kMaxSections = 8;
struct IirDev {
  int n_sections, padlen;
  float b[kMaxSections][3], a[kMaxSections][2];  // b0 b1 b2, a1 a2 (a0 = 1)
  float zi[kMaxSections][2];                    // steady-state states of each section per unit input
};

struct Cx {
  float x, y;
};

// One cascade step (transposed direct form II per section) on a complex sample; NS sections.
template <int NS>
__device__ __forceinline__ Cx iir_step(const IirDev& f, Cx (&s)[kMaxSections][2], Cx v) {
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const float b0 = f.b[k][0], b1 = f.b[k][1], b2 = f.b[k][2], a1 = f.a[k][0], a2 = f.a[k][1];
    const Cx y = {fmaf(b0, v.x, s[k][0].x), fmaf(b0, v.y, s[k][0].y)};
    s[k][0] = {fmaf(b1, v.x, fmaf(-a1, y.x, s[k][1].x)), fmaf(b1, v.y, fmaf(-a1, y.y, s[k][1].y))};
    s[k][1] = {fmaf(b2, v.x, -a2 * y.x), fmaf(b2, v.y, -a2 * y.y)};
    v = y;
  }
  return v;
}

// States of every section after an infinitely long constant input v.
template <int NS>
__device__ __forceinline__ void iir_init(const IirDev& f, Cx (&s)[kMaxSections][2], Cx v) {
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    s[k][0] = {f.zi[k][0] * v.x, f.zi[k][0] * v.y};
    s[k][1] = {f.zi[k][1] * v.x, f.zi[k][1] * v.y};
    // the next section sees this one's constant output: v times its DC gain
    const float g = (f.b[k][0] + f.b[k][1] + f.b[k][2]) / (1.0f + f.a[k][0] + f.a[k][1]);
    v = {v.x * g, v.y * g};
  }
}

// rf sample n of the sequence, mixed down: rf(n) exp(-i 2 pi (fc / fs) n); the phase in revolutions is
// reduced in FP64.
__device__ __forceinline__ Cx mixed(const float* __restrict__ rf, size_t stride, int n, double cyc) {
  const float v = rf[(size_t)n * stride];
  const double ph = cyc * (double)n;
  const float r = (float)(ph - floor(ph));
  return {v * __builtin_amdgcn_cosf(r), -v * __builtin_amdgcn_sinf(r)};
}

// Odd extension of the mixed signal: index i in [0, N + 2 pad), original sample i - pad.
__device__ __forceinline__ Cx extended(const float* __restrict__ rf, size_t stride, int N, int pad, int i, double cyc,
                                       Cx first, Cx last) {
  if (i < pad) {
    const Cx m = mixed(rf, stride, pad - i, cyc);
    return {2.f * first.x - m.x, 2.f * first.y - m.y};
  }
  if (i < pad + N) return mixed(rf, stride, i - pad, cyc);
  const Cx m = mixed(rf, stride, N - 2 - (i - pad - N), cyc);
  return {2.f * last.x - m.x, 2.f * last.y - m.y};
}

// rf[t][n][e] -> iq[t][n][e] (interleaved complex), or iq[t][e][n] (CM: channel-major, each sequence
// contiguous, the layout k_das gathers from fastest). tail[seq][pad]: forward output past the end.
//
// Each pass over a sequence is split into RF_K chunks run by RF_K threads at once. The cascade is a
// linear system with M = 2 NS real states per component (re and im alike): a chunk of length len
// maps its initial state S to A^len S + phi, phi being its final state from a zero start. Phase 1:
// every chunk computes phi (zero start, no output). Phase 2: one thread per sequence chains the
// chunks, S_{k+1} = A^len S_k + phi_k, from the pass's steady-state start (A^len evaluated on the
// host in FP64). Phase 3: every chunk reruns from its S_k and writes its outputs. The backward pass
// reads the forward outputs of other chunks of the same workgroup after a barrier. Inputs are read
// in blocks of RF_BLOCK samples before the block's recursion runs.
constexpr int RF_BLOCK = 16;
constexpr int RF_CH = 32;  // channels per workgroup (a wave: two chunks of 32 consecutive channels)
constexpr int RF_K = 32;   // chunks per sequence
template <int NS>
struct IirChunks {
  int lenf, lenb;                   // chunk lengths of the forward and backward passes
  float af[2 * NS][2 * NS];         // A^lenf (state index 2 k + j: state j of section k)
  float ab[2 * NS][2 * NS];         // A^lenb
};

template <int NS>
__device__ __forceinline__ void load_state(Cx (&s)[kMaxSections][2], const Cx* v) {
#pragma unroll
  for (int m = 0; m < 2 * NS; ++m) s[m / 2][m % 2] = v[m];
}
template <int NS>
__device__ __forceinline__ void store_state(const Cx (&s)[kMaxSections][2], Cx* v) {
#pragma unroll
  for (int m = 0; m < 2 * NS; ++m) v[m] = s[m / 2][m % 2];
}

// Phase 2 of a pass for one sequence: slot[k] holds phi_k on entry and S_k on return.
template <int NS>
__device__ __forceinline__ void chain_chunks(Cx (*slot)[RF_CH][2 * NS], int c, int n_chunks,
                                             const float (&A)[2 * NS][2 * NS], const Cx (&s0)[kMaxSections][2]) {
  constexpr int M = 2 * NS;
  Cx S[M];
#pragma unroll
  for (int m = 0; m < M; ++m) S[m] = s0[m / 2][m % 2];
  for (int k = 0; k < n_chunks; ++k) {
    Cx nx[M];
#pragma unroll
    for (int r = 0; r < M; ++r) {
      Cx acc = slot[k][c][r];
#pragma unroll
      for (int q = 0; q < M; ++q) acc = {fmaf(A[r][q], S[q].x, acc.x), fmaf(A[r][q], S[q].y, acc.y)};
      nx[r] = acc;
    }
#pragma unroll
    for (int m = 0; m < M; ++m) {
      slot[k][c][m] = S[m];
      S[m] = nx[m];
    }
  }
}

template <int NS, bool CM>
__global__ __launch_bounds__(RF_CH * RF_K) void k_rf2iq(const float* __restrict__ rf, int T, int N, int E, double cyc,
                                                       const IirDev f, const IirChunks<NS> ch,
                                                       float2* __restrict__ iq, float2* __restrict__ tail) {
  __shared__ Cx slot[RF_K][RF_CH][2 * NS];
  __shared__ Cx vlast[RF_CH];
  const int c = threadIdx.x % RF_CH, k = threadIdx.x / RF_CH;
  const int e = blockIdx.x * RF_CH + c;
  const int t = blockIdx.y;
  const bool valid = e < E;
  const int ee = valid ? e : E - 1;  // lanes past the last channel read a real one and write nothing
  const size_t base = (size_t)t * N * E + ee;
  const float* x = rf + base;
  float2* y = iq + (CM ? ((size_t)t * E + ee) * N : base);
  const size_t ys = CM ? 1 : (size_t)E;  // stride of the output samples
  float2* tl = tail + ((size_t)t * E + ee) * f.padlen;
  const int pad = f.padlen;
  const int L = N + 2 * pad;  // forward pass: the extended signal
  const int Lb = N + pad;     // backward pass: extended indices L - 1 down to pad
  const Cx first = mixed(x, E, 0, cyc), last = mixed(x, E, N - 1, cyc);
  Cx s[kMaxSections][2];

  // ---- forward pass
  const int f0 = k * ch.lenf, f1 = min(f0 + ch.lenf, L);
  auto forward = [&](bool emit) {
    for (int i0 = f0; i0 < f1; i0 += RF_BLOCK) {
      Cx in[RF_BLOCK];
#pragma unroll
      for (int j = 0; j < RF_BLOCK; ++j)
        in[j] = i0 + j < f1 ? extended(x, E, N, pad, i0 + j, cyc, first, last) : Cx{0.f, 0.f};
#pragma unroll
      for (int j = 0; j < RF_BLOCK; ++j) {
        const int i = i0 + j;
        if (i < f1) {
          const Cx v = iir_step<NS>(f, s, in[j]);
          if (emit && valid) {
            if (i >= pad && i < pad + N) y[(size_t)(i - pad) * ys] = make_float2(v.x, v.y);
            else if (i >= pad + N) tl[i - pad - N] = make_float2(v.x, v.y);
          }
          if (emit && i == L - 1) vlast[c] = v;
        }
      }
    }
  };
#pragma unroll
  for (int q = 0; q < NS; ++q) s[q][0] = s[q][1] = Cx{0.f, 0.f};
  forward(false);
  store_state<NS>(s, slot[k][c]);
  __syncthreads();
  if (k == 0) {
    Cx s0[kMaxSections][2];
    iir_init<NS>(f, s0, extended(x, E, N, pad, 0, cyc, first, last));
    chain_chunks<NS>(slot, c, (L + ch.lenf - 1) / ch.lenf, ch.af, s0);
  }
  __syncthreads();
  load_state<NS>(s, slot[k][c]);
  forward(true);
  __syncthreads();  // forward outputs (global) and vlast visible to the workgroup

  // ---- backward pass over j = L - 1 - i, j in [0, Lb); every index of a block is read before the
  // block writes any of them, and a chunk writes only its own indices
  const int b0 = k * ch.lenb, b1 = min(b0 + ch.lenb, Lb);
  auto backward = [&](bool emit) {
    for (int j0 = b0; j0 < b1; j0 += RF_BLOCK) {
      Cx in[RF_BLOCK];
#pragma unroll
      for (int j = 0; j < RF_BLOCK; ++j) {
        const int i = L - 1 - (j0 + j);
        float2 w = make_float2(0.f, 0.f);
        if (j0 + j < b1) w = i >= pad + N ? tl[i - pad - N] : y[(size_t)(i - pad) * ys];
        in[j] = {w.x, w.y};
      }
#pragma unroll
      for (int j = 0; j < RF_BLOCK; ++j) {
        const int i = L - 1 - (j0 + j);
        if (j0 + j < b1) {
          const Cx o = iir_step<NS>(f, s, in[j]);
          if (emit && valid && i < pad + N) y[(size_t)(i - pad) * ys] = make_float2(2.f * o.x, 2.f * o.y);
        }
      }
    }
  };
#pragma unroll
  for (int q = 0; q < NS; ++q) s[q][0] = s[q][1] = Cx{0.f, 0.f};
  backward(false);
  __syncthreads();  // every chunk's phase-1 reads of the forward outputs done before any rewrite
  store_state<NS>(s, slot[k][c]);
  __syncthreads();
  if (k == 0) {
    Cx s0[kMaxSections][2];
    iir_init<NS>(f, s0, vlast[c]);
    chain_chunks<NS>(slot, c, (Lb + ch.lenb - 1) / ch.lenb, ch.ab, s0);
  }
  __syncthreads();
  load_state<NS>(s, slot[k][c]);
  backward(true);
}

// ---------------------------------------------------------------- delay-and-sum
// img[p] (compound) or img[t][p]: pixels (px, pz); elements (ex, ez) float64; delays [t][e] float64 (NaN
// marks an element that does not transmit); iq[t][n][e], or iq[t][e][n] (CM: the two interpolation
// nodes of a term are adjacent, and the lanes of a wave, neighbouring pixels, read nearby samples of
// one sequence).
// A workgroup is DAS_PX pixels (one per lane) x DAS_GW waves; wave g takes the firing groups g,
// g + DAS_GW, ... of DAS_TG firings each, and the compound image sums the waves' partial sums in a
// fixed order through LDS (deterministic). Per group, the FP64 distance to every element is
// evaluated twice (transmit arrival, then receive), not once per firing and element, from the FP32
// reciprocal square root refined by one FP64 Newton step (relative error ~1e-14). The sample index
// (tau - t0) fs and the phase fc tau of a term are sums of a firing part (tau_tx) and an element part
// (|p - x_e| / c), each split in FP64 into an integer and a fraction (the phase modulo one
// revolution): the term adds the two fractions in FP32 (error ~1e-7 sample or revolution, against
// ~3e-4 rad at fc for a 60 us path carried in FP32). The interpolation weight and the rotation are FP32.
constexpr int DAS_TG = 8;   // firings per group
constexpr int DAS_PX = 64;  // pixels per workgroup
constexpr int DAS_GW = 8;   // waves per workgroup
constexpr double DAS_INDEX_LIMIT = 1e9;  // sample indices beyond this are out of any record

__device__ __forceinline__ double das_dist(double dx, double dz) {
  const double d2 = fma(dx, dx, dz * dz);
  if (!(d2 >= 1e-30)) return sqrt(d2);  // at (or next to) an element: the exact square root
  double y = (double)__builtin_amdgcn_rsqf((float)d2);
  y = y * fma(-0.5 * d2, y * y, 1.5);
  return d2 * y;
}

template <bool CM>
__global__ __launch_bounds__(DAS_PX * DAS_GW) void k_das(const float2* __restrict__ iq, int T, int N, int E,
                                                         const double* __restrict__ ex, const double* __restrict__ ez,
                                                         const double* __restrict__ delay,
                                                         const float* __restrict__ px, const float* __restrict__ pz,
                                                         int P, double fs, double fc, double inv_c, double t0,
                                                         float half_aperture_per_depth, int compound,
                                                         float2* __restrict__ img) {
  __shared__ float2 red[DAS_GW][DAS_PX];
  const int lane = threadIdx.x % DAS_PX;
  const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / DAS_PX));
  const int p = blockIdx.x * DAS_PX + lane;
  const bool valid = p < P;
  const int pp = valid ? p : P - 1;  // lanes past the last pixel compute a real one and write nothing
  const double x = px[pp], z = pz[pp];
  const float amax = fabsf((float)z) * half_aperture_per_depth;
  const int i_hi = N - 2;  // last left interpolation node (N >= 2)
  float sx = 0.f, sy = 0.f;
  for (int tg = g * DAS_TG; tg < T; tg += DAS_GW * DAS_TG) {
    // transmit arrival of each firing of the group: min over transmitting elements (FP64)
    double ttx[DAS_TG];
#pragma unroll
    for (int j = 0; j < DAS_TG; ++j) ttx[j] = INFINITY;
    for (int e = 0; e < E; ++e) {
      const double rc = das_dist(x - ex[e], z - ez[e]) * inv_c;
#pragma unroll
      for (int j = 0; j < DAS_TG; ++j) {
        if (tg + j < T) {
          const double d = delay[(size_t)(tg + j) * E + e];
          if (isfinite(d)) ttx[j] = fmin(ttx[j], d + rc);
        }
      }
    }
    // firing parts: sample index (tau_tx - t0) fs = ia + fa, phase fc tau_tx = integer + pa
    int ia[DAS_TG];
    float fa[DAS_TG], pa[DAS_TG];
    bool live[DAS_TG];
#pragma unroll
    for (int j = 0; j < DAS_TG; ++j) {
      const double A = (ttx[j] - t0) * fs;
      live[j] = tg + j < T && A > -DAS_INDEX_LIMIT && A < DAS_INDEX_LIMIT;  // false for an infinite tau_tx
      const double Af = live[j] ? floor(A) : 0.0;
      ia[j] = (int)Af;
      fa[j] = live[j] ? (float)(A - Af) : 0.f;
      const double ph = live[j] ? fc * ttx[j] : 0.0;
      pa[j] = (float)(ph - floor(ph));
    }
    float ax[DAS_TG], ay[DAS_TG];
#pragma unroll
    for (int j = 0; j < DAS_TG; ++j) ax[j] = ay[j] = 0.f;
    for (int e = 0; e < E; ++e) {
      const double dx = x - ex[e];
      if (fabsf((float)dx) > amax) continue;
      const double rc = das_dist(dx, z - ez[e]) * inv_c;
      const double B = rc * fs;
      if (!(B < DAS_INDEX_LIMIT)) continue;
      const double Bf = floor(B);
      const int ib = (int)Bf;
      const float fb = (float)(B - Bf);
      const double phb = fc * rc;
      const float pb = (float)(phb - floor(phb));
#pragma unroll
      for (int j = 0; j < DAS_TG; ++j) {
        if (!live[j]) continue;
        // si = (ia + ib) + (fa + fb), fractions in [0, 1)
        float fr = fa[j] + fb;
        int i = ia[j] + ib;
        if (fr >= 1.f) {
          fr -= 1.f;
          ++i;
        }
        if (i < 0 || i > N - 1 || (i == N - 1 && fr > 0.f)) continue;  // si outside [0, N - 1]
        const int i0 = min(i, i_hi);
        const float w = (float)(i - i0) + fr;
        const float2* q = CM ? iq + ((size_t)(tg + j) * E + e) * N + i0 : iq + ((size_t)(tg + j) * N + i0) * E + e;
        const float2 a = q[0];
        const float2 b = q[CM ? 1 : E];  // N >= 2: i0 + 1 <= N - 1
        const float vr = fmaf(w, b.x - a.x, a.x), vi = fmaf(w, b.y - a.y, a.y);
        const float r = __builtin_amdgcn_fractf(pa[j] + pb);
        const float c = __builtin_amdgcn_cosf(r), sn = __builtin_amdgcn_sinf(r);
        ax[j] = fmaf(vr, c, fmaf(-vi, sn, ax[j]));
        ay[j] = fmaf(vr, sn, fmaf(vi, c, ay[j]));
      }
    }
#pragma unroll
    for (int j = 0; j < DAS_TG; ++j) {
      if (tg + j < T) {
        if (compound) {
          sx += ax[j];
          sy += ay[j];
        } else if (valid) {
          img[(size_t)(tg + j) * P + p] = make_float2(ax[j], ay[j]);
        }
      }
    }
  }
  if (compound) {  // uniform over the workgroup
    red[g][lane] = make_float2(sx, sy);
    __syncthreads();
    if (g == 0 && valid) {
      float2 acc = red[0][lane];
#pragma unroll
      for (int q = 1; q < DAS_GW; ++q) acc = make_float2(acc.x + red[q][lane].x, acc.y + red[q][lane].y);
      img[p] = acc;
    }
  }
}

}  // namespace fs

using namespace fs;

extern "C" {

size_t fs_rf2iq_workspace_bytes(int n_firings, int n_channels, const fs_iir* filter) {
  if (n_firings <= 0 || n_channels <= 0 || !filter || filter->padlen < 0) return 0;
  return std::max<size_t>(256, (size_t)n_firings * n_channels * std::max(filter->padlen, 1) * sizeof(float2));
}

}  // extern "C"

namespace fs {
// A^len of the cascade (zero input): column q is the state after len steps from unit state q, in FP64.
template <int NS>
static void chunk_matrix(const double (&b)[kMaxSections][3], const double (&a)[kMaxSections][2], int len,
                         float (&out)[2 * NS][2 * NS]) {
  for (int q = 0; q < 2 * NS; ++q) {
    double s[NS][2] = {};
    s[q / 2][q % 2] = 1.0;
    for (int n = 0; n < len; ++n) {
      double v = 0.0;
      for (int k = 0; k < NS; ++k) {
        const double yk = b[k][0] * v + s[k][0];
        const double s0 = b[k][1] * v - a[k][0] * yk + s[k][1];
        const double s1 = b[k][2] * v - a[k][1] * yk;
        s[k][0] = s0;
        s[k][1] = s1;
        v = yk;
      }
    }
    for (int r = 0; r < 2 * NS; ++r) out[r][q] = (float)s[r / 2][r % 2];
  }
}

template <int NS>
static void launch_rf2iq(const double (&b)[kMaxSections][3], const double (&a)[kMaxSections][2], const IirDev& f,
                         const float* rf, int T, int N, int E, double cyc, bool cm, float2* iq, float2* tail,
                         hipStream_t st) {
  IirChunks<NS> ch;
  ch.lenf = (N + 2 * f.padlen + RF_K - 1) / RF_K;
  ch.lenb = (N + f.padlen + RF_K - 1) / RF_K;
  chunk_matrix<NS>(b, a, ch.lenf, ch.af);
  chunk_matrix<NS>(b, a, ch.lenb, ch.ab);
  const dim3 grid((E + RF_CH - 1) / RF_CH, T);
  if (cm)
    hipLaunchKernelGGL((k_rf2iq<NS, true>), grid, dim3(RF_CH * RF_K), 0, st, rf, T, N, E, cyc, f, ch, iq, tail);
  else
    hipLaunchKernelGGL((k_rf2iq<NS, false>), grid, dim3(RF_CH * RF_K), 0, st, rf, T, N, E, cyc, f, ch, iq, tail);
}
}  // namespace fs

extern "C" {

int fs_rf2iq(const float* rf, int n_firings, int n_samples, int n_channels, double fs_hz, double fc_hz,
             const fs_iir* filter, int channel_major, float* iq, void* workspace, size_t workspace_bytes,
             void* stream) {
  if (!rf || !iq || !filter || n_firings <= 0 || n_channels <= 0 || !(fs_hz > 0) || !(fc_hz >= 0))
    return FS_ERR_INVALID;
  if (filter->n_sections < 1 || filter->n_sections > kMaxSections || filter->padlen < 0) return FS_ERR_INVALID;
  if (n_samples <= filter->padlen || n_samples < 2) return FS_ERR_INVALID;  // the odd extension needs padlen + 1 samples
  if (n_firings > 65535) return FS_ERR_UNSUPPORTED;
  if (n_samples > (1 << 28)) return FS_ERR_UNSUPPORTED;
  if (!workspace || workspace_bytes < fs_rf2iq_workspace_bytes(n_firings, n_channels, filter)) return FS_ERR_WORKSPACE;
  IirDev f{};
  double b[kMaxSections][3] = {}, a[kMaxSections][2] = {};
  f.n_sections = filter->n_sections;
  f.padlen = filter->padlen;
  for (int k = 0; k < filter->n_sections; ++k) {
    const double* sos = filter->sos[k];
    if (!(sos[3] != 0.0)) return FS_ERR_INVALID;
    for (int j = 0; j < 3; ++j) b[k][j] = sos[j] / sos[3];
    a[k][0] = sos[4] / sos[3];
    a[k][1] = sos[5] / sos[3];
    for (int j = 0; j < 3; ++j) f.b[k][j] = (float)b[k][j];
    f.a[k][0] = (float)a[k][0];
    f.a[k][1] = (float)a[k][1];
    f.zi[k][0] = (float)filter->zi[k][0];
    f.zi[k][1] = (float)filter->zi[k][1];
  }
  hipStream_t st = (hipStream_t)stream;
  const double cyc = fc_hz / fs_hz;
  const bool cm = channel_major != 0;
  float2* out = (float2*)iq;
  float2* tail = (float2*)workspace;
  switch (f.n_sections) {
    case 1: launch_rf2iq<1>(b, a, f, rf, n_firings, n_samples, n_channels, cyc, cm, out, tail, st); break;
    case 2: launch_rf2iq<2>(b, a, f, rf, n_firings, n_samples, n_channels, cyc, cm, out, tail, st); break;
    case 3: launch_rf2iq<3>(b, a, f, rf, n_firings, n_samples, n_channels, cyc, cm, out, tail, st); break;
    case 4: launch_rf2iq<4>(b, a, f, rf, n_firings, n_samples, n_channels, cyc, cm, out, tail, st); break;
    case 5: launch_rf2iq<5>(b, a, f, rf, n_firings, n_samples, n_channels, cyc, cm, out, tail, st); break;
    case 6: launch_rf2iq<6>(b, a, f, rf, n_firings, n_samples, n_channels, cyc, cm, out, tail, st); break;
    case 7: launch_rf2iq<7>(b, a, f, rf, n_firings, n_samples, n_channels, cyc, cm, out, tail, st); break;
    case 8: launch_rf2iq<8>(b, a, f, rf, n_firings, n_samples, n_channels, cyc, cm, out, tail, st); break;
    default: return FS_ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

int fs_das(const float* iq, int n_firings, int n_samples, int n_channels, int channel_major, const double* elem_x,
           const double* elem_z, const double* tx_delay, const float* pixel_x, const float* pixel_z, long long n_pixels,
           double fs_hz, double fc_hz, double speed_of_sound, double t0, double f_number, int compound, float* image,
           void* stream) {
  if (!iq || !elem_x || !elem_z || !tx_delay || !pixel_x || !pixel_z || !image) return FS_ERR_INVALID;
  if (n_firings <= 0 || n_samples < 2 || n_channels <= 0 || n_pixels <= 0) return FS_ERR_INVALID;
  if (!(fs_hz > 0) || !(fc_hz >= 0) || !(speed_of_sound > 0) || !(f_number >= 0)) return FS_ERR_INVALID;
  if (n_pixels > (1LL << 31) - 256) return FS_ERR_UNSUPPORTED;
  const int P = (int)n_pixels;
  // aperture half-width per unit depth: z / (2 F); F = 0 opens the full aperture
  const float hapd = f_number > 0 ? (float)(0.5 / f_number) : INFINITY;
  const dim3 grid((P + DAS_PX - 1) / DAS_PX);
  if (channel_major)
    hipLaunchKernelGGL(k_das<true>, grid, dim3(DAS_PX * DAS_GW), 0, (hipStream_t)stream, (const float2*)iq, n_firings, n_samples,
                       n_channels, elem_x, elem_z, tx_delay, pixel_x, pixel_z, P, fs_hz, fc_hz, 1.0 / speed_of_sound,
                       t0, hapd, compound ? 1 : 0, (float2*)image);
  else
    hipLaunchKernelGGL(k_das<false>, grid, dim3(DAS_PX * DAS_GW), 0, (hipStream_t)stream, (const float2*)iq, n_firings, n_samples,
                       n_channels, elem_x, elem_z, tx_delay, pixel_x, pixel_z, P, fs_hz, fc_hz, 1.0 / speed_of_sound,
                       t0, hapd, compound ? 1 : 0, (float2*)image);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

}  // extern "C"
