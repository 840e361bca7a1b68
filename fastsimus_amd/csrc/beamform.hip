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

// One cascade step (transposed direct form II per section) on a complex sample.
__device__ __forceinline__ Cx iir_step(const IirDev& f, Cx (&s)[kMaxSections][2], Cx v) {
#pragma unroll
  for (int k = 0; k < kMaxSections; ++k) {
    if (k >= f.n_sections) break;
    const float b0 = f.b[k][0], b1 = f.b[k][1], b2 = f.b[k][2], a1 = f.a[k][0], a2 = f.a[k][1];
    const Cx y = {fmaf(b0, v.x, s[k][0].x), fmaf(b0, v.y, s[k][0].y)};
    s[k][0] = {fmaf(b1, v.x, fmaf(-a1, y.x, s[k][1].x)), fmaf(b1, v.y, fmaf(-a1, y.y, s[k][1].y))};
    s[k][1] = {fmaf(b2, v.x, -a2 * y.x), fmaf(b2, v.y, -a2 * y.y)};
    v = y;
  }
  return v;
}

// States of every section after an infinitely long constant input v.
__device__ __forceinline__ void iir_init(const IirDev& f, Cx (&s)[kMaxSections][2], Cx v) {
#pragma unroll
  for (int k = 0; k < kMaxSections; ++k) {
    if (k >= f.n_sections) break;
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

// rf[t][n][e] -> iq[t][n][e] (interleaved complex). tail[seq][pad]: forward output past the end.
__global__ __launch_bounds__(256) void k_rf2iq(const float* __restrict__ rf, int T, int N, int E, double cyc,
                                               const IirDev f, float2* __restrict__ iq, float2* __restrict__ tail) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (e >= E) return;
  const size_t base = (size_t)t * N * E + e;
  const float* x = rf + base;
  float2* y = iq + base;
  float2* tl = tail + ((size_t)t * E + e) * f.padlen;
  const int pad = f.padlen;
  const Cx first = mixed(x, E, 0, cyc), last = mixed(x, E, N - 1, cyc);
  Cx s[kMaxSections][2];
  // forward pass over the extended signal
  const Cx e0 = extended(x, E, N, pad, 0, cyc, first, last);
  iir_init(f, s, e0);
  Cx v;
  for (int i = 0; i < N + 2 * pad; ++i) {
    v = iir_step(f, s, extended(x, E, N, pad, i, cyc, first, last));
    if (i >= pad && i < pad + N) y[(size_t)(i - pad) * E] = make_float2(v.x, v.y);
    else if (i >= pad + N) tl[i - pad - N] = make_float2(v.x, v.y);
  }
  // backward pass from the end of the extension, starting in the steady state of its first value
  iir_init(f, s, v);
  for (int i = N + 2 * pad - 1; i >= pad; --i) {
    Cx u;
    if (i >= pad + N) {
      const float2 w = tl[i - pad - N];
      u = {w.x, w.y};
    } else {
      const float2 w = y[(size_t)(i - pad) * E];
      u = {w.x, w.y};
    }
    const Cx o = iir_step(f, s, u);
    if (i < pad + N) y[(size_t)(i - pad) * E] = make_float2(2.f * o.x, 2.f * o.y);
  }
}

// ---------------------------------------------------------------- delay-and-sum
// img[p] (compound) or img[t][p]: pixels (px, pz); elements (ex, ez) float64; delays [t][e] float64 (NaN
// marks an element that does not transmit); iq[t][n][e].
// Delays are FP64 (a 60 us path in FP32 would carry ~3e-4 rad of phase error at fc); the interpolation
// weight and the rotation after the phase reduction are FP32.
__global__ __launch_bounds__(256) void k_das(const float2* __restrict__ iq, int T, int N, int E,
                                             const double* __restrict__ ex, const double* __restrict__ ez,
                                             const double* __restrict__ delay, const float* __restrict__ px,
                                             const float* __restrict__ pz, int P, double fs, double fc, double inv_c,
                                             double t0, float half_aperture_per_depth, int compound,
                                             float2* __restrict__ img) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double x = px[p], z = pz[p];
  const float amax = fabsf((float)z) * half_aperture_per_depth;
  float sx = 0.f, sy = 0.f;
  for (int t = 0; t < T; ++t) {
    const double* dt = delay + (size_t)t * E;
    double ttx = INFINITY;
    for (int e = 0; e < E; ++e) {
      const double d = dt[e];
      if (!isfinite(d)) continue;
      const double dx = x - ex[e], dz = z - ez[e];
      ttx = fmin(ttx, d + sqrt(fma(dx, dx, dz * dz)) * inv_c);
    }
    float ax = 0.f, ay = 0.f;
    if (isfinite(ttx)) {
      const float2* q = iq + (size_t)t * N * E;
      for (int e = 0; e < E; ++e) {
        const double dx = x - ex[e], dz = z - ez[e];
        if (fabsf((float)dx) > amax) continue;
        const double tau = ttx + sqrt(fma(dx, dx, dz * dz)) * inv_c;
        const double si = (tau - t0) * fs;
        if (!(si >= 0.0) || si > (double)(N - 1)) continue;
        const int i0 = min((int)si, N - 2 >= 0 ? N - 2 : 0);
        const float w = (float)(si - (double)i0);
        const float2 a = q[(size_t)i0 * E + e];
        const float2 b = N > 1 ? q[(size_t)(i0 + 1) * E + e] : a;
        const float vr = fmaf(w, b.x - a.x, a.x), vi = fmaf(w, b.y - a.y, a.y);
        const double ph = fc * tau;
        const float r = (float)(ph - floor(ph));
        const float c = __builtin_amdgcn_cosf(r), s = __builtin_amdgcn_sinf(r);
        ax = fmaf(vr, c, fmaf(-vi, s, ax));
        ay = fmaf(vr, s, fmaf(vi, c, ay));
      }
    }
    if (compound) {
      sx += ax;
      sy += ay;
    } else {
      img[(size_t)t * P + p] = make_float2(ax, ay);
    }
  }
  if (compound) img[p] = make_float2(sx, sy);
}

}  // namespace fs

using namespace fs;

extern "C" {

size_t fs_rf2iq_workspace_bytes(int n_firings, int n_channels, const fs_iir* filter) {
  if (n_firings <= 0 || n_channels <= 0 || !filter || filter->padlen < 0) return 0;
  return std::max<size_t>(256, (size_t)n_firings * n_channels * std::max(filter->padlen, 1) * sizeof(float2));
}

int fs_rf2iq(const float* rf, int n_firings, int n_samples, int n_channels, double fs_hz, double fc_hz,
             const fs_iir* filter, float* iq, void* workspace, size_t workspace_bytes, void* stream) {
  if (!rf || !iq || !filter || n_firings <= 0 || n_channels <= 0 || !(fs_hz > 0) || !(fc_hz >= 0))
    return FS_ERR_INVALID;
  if (filter->n_sections < 1 || filter->n_sections > kMaxSections || filter->padlen < 0) return FS_ERR_INVALID;
  if (n_samples <= filter->padlen || n_samples < 2) return FS_ERR_INVALID;  // the odd extension needs padlen + 1 samples
  if (n_firings > 65535) return FS_ERR_UNSUPPORTED;
  if (!workspace || workspace_bytes < fs_rf2iq_workspace_bytes(n_firings, n_channels, filter)) return FS_ERR_WORKSPACE;
  IirDev f{};
  f.n_sections = filter->n_sections;
  f.padlen = filter->padlen;
  for (int k = 0; k < filter->n_sections; ++k) {
    const double* sos = filter->sos[k];
    if (!(sos[3] != 0.0)) return FS_ERR_INVALID;
    for (int j = 0; j < 3; ++j) f.b[k][j] = (float)(sos[j] / sos[3]);
    f.a[k][0] = (float)(sos[4] / sos[3]);
    f.a[k][1] = (float)(sos[5] / sos[3]);
    f.zi[k][0] = (float)filter->zi[k][0];
    f.zi[k][1] = (float)filter->zi[k][1];
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_rf2iq, dim3((n_channels + 255) / 256, n_firings), dim3(256), 0, st, rf, n_firings, n_samples,
                     n_channels, fc_hz / fs_hz, f, (float2*)iq, (float2*)workspace);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

int fs_das(const float* iq, int n_firings, int n_samples, int n_channels, const double* elem_x, const double* elem_z,
           const double* tx_delay, const float* pixel_x, const float* pixel_z, long long n_pixels, double fs_hz,
           double fc_hz, double speed_of_sound, double t0, double f_number, int compound, float* image, void* stream) {
  if (!iq || !elem_x || !elem_z || !tx_delay || !pixel_x || !pixel_z || !image) return FS_ERR_INVALID;
  if (n_firings <= 0 || n_samples <= 0 || n_channels <= 0 || n_pixels <= 0) return FS_ERR_INVALID;
  if (!(fs_hz > 0) || !(fc_hz >= 0) || !(speed_of_sound > 0) || !(f_number >= 0)) return FS_ERR_INVALID;
  if (n_pixels > (1LL << 31) - 256) return FS_ERR_UNSUPPORTED;
  const int P = (int)n_pixels;
  // aperture half-width per unit depth: z / (2 F); F = 0 opens the full aperture
  const float hapd = f_number > 0 ? (float)(0.5 / f_number) : INFINITY;
  hipLaunchKernelGGL(k_das, dim3((P + 255) / 256), dim3(256), 0, (hipStream_t)stream, (const float2*)iq, n_firings,
                     n_samples, n_channels, elem_x, elem_z, tx_delay, pixel_x, pixel_z, P, fs_hz, fc_hz,
                     1.0 / speed_of_sound, t0, hapd, compound ? 1 : 0, (float2*)image);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

}  // extern "C"
