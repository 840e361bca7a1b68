// Device-side building blocks shared by the gfx950 kernels.
//
// Green's function of one (scatterer s, element e, sub-element j) term at
// frequency f (reference math: _pfield_math.py:71-247, simus.py:263-333):
//
//   g(f) = amp * exp(-att * f) * exp(i 2 pi f t)
//   amp  = obliquity(theta) * directivity(theta) / sqrt(r_c) / n_sub
//   att  = alpha_dB / (20 / ln 10) * 1e-4 * r_c         (Np per Hz)
//   t    = r_c / c                                       (s)
//   r_c  = max(r, c / (2 fc)),  theta = asin(dx / r) - theta_e
//
// r and the phase f * t (mod 1) are evaluated in FP64; everything that only
// scales amplitude is FP32. Phases are carried in revolutions so that the
// range reduction is an exact FP64 fract().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fastsimus.h"

namespace fs {

constexpr float kPi = 3.14159265358979323846f;
constexpr double kNeperToDb = 8.6858896380650365;  // 20 / ln(10)

struct cf {
  float x, y;
};
__device__ __forceinline__ cf cmul(cf a, cf b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ void cmac(cf& acc, cf a, cf b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
}

// Scalars of one problem, copied by value into every kernel.
struct Problem {
  int S, E, NS, NF, NT;
  int fulldir, baffle;
  double f0, df, inv_c, min_dist;
  float att_coef;       // alpha_dB / (20/ln10) * 1e-4   (Np / (m Hz))
  float kappa;          // pi * fc * seg / c  (centre-frequency sinc scale)
  float kappa_per_hz;   // pi * seg / c       (per-frequency sinc scale)
  float baffle_b;
  float inv_nsub;
  double radius, apex;
  float corr;
  const float *x, *z, *rc;
  const double *sub_x, *sub_z, *ecos, *esin;
  const double *delay, *apod;
  const double *pulse_re, *pulse_im, *probe;
};

__device__ __forceinline__ bool out_of_field(const Problem& p, float xs, float zs) {
  bool out = zs < 0.0f;
  if (isfinite(p.radius)) {
    double X = xs, Z = (double)zs + p.apex;
    out = out || (X * X + Z * Z <= p.radius * p.radius);
  }
  return out;
}

// ---------------------------------------------------------------- per-term geometry
// Evaluated with the gfx950 hardware transcendentals: v_sin_f32 / v_cos_f32
// take revolutions (max abs error 1.3e-7 measured over [0, 1)), v_exp_f32
// (rel 8e-8), v_rsq_f64 refined by one Newton step (rel ~1e-15). The phase is
// reduced in FP64.
constexpr float kLog2e = 1.4426950408889634f;

__device__ __forceinline__ float sinc_fast(float a) {
  // sin(a)/a: Taylor to a^8 below |a| = 1 (error < 3e-8), hardware sine above.
  const float a2 = a * a;
  if (fabsf(a) < 1.0f) {
    return fmaf(a2, fmaf(a2, fmaf(a2, fmaf(a2, 2.7557319e-6f, -1.9841270e-4f), 8.3333333e-3f), -0.16666667f), 1.0f);
  }
  return __builtin_amdgcn_sinf(a * (0.5f / kPi)) * __builtin_amdgcn_rcpf(a);
}

struct GeoF {
  double t;    // r_c / c
  float amp;   // obliquity * directivity / sqrt(r_c) / n_sub  (directivity at fc unless fulldir)
  float att;   // Np / Hz, pre-multiplied by log2(e)
  float sint;  // sin(theta)
};

// The element-side constants of geometry_fast for (element e, sub-element j):
// centroid and element orientation. A lane that keeps its element over a loop
// of scatterers loads them once.
struct SubGeo {
  double x, z;
  float ce, se;
};
__device__ __forceinline__ SubGeo sub_geo(const Problem& p, int e, int j) {
  const int idx = e * p.NS + j;
  return {p.sub_x[idx], p.sub_z[idx], (float)p.ecos[e], (float)p.esin[e]};
}

__device__ __forceinline__ GeoF geometry_fast(const Problem& p, double xs, double zs, const SubGeo& sg) {
  const double dx = xs - sg.x;
  const double dz = zs - sg.z;
  const double r2 = fma(dx, dx, dz * dz);
  const bool zero = !(r2 > 0.0);
  double y = __builtin_amdgcn_rsq(zero ? 1.0 : r2);
  y = y * fma(-0.5 * r2, y * y, 1.5);  // Newton step on 1/sqrt
  const double r = zero ? 0.0 : r2 * y;
  const double rcl = fmax(r, p.min_dist);
  const float rinv = (float)y;
  const float sa = zero ? 1.0f : (float)dx * rinv;
  const float ca = zero ? 0.0f : fabsf((float)dz) * rinv;
  const float st = sa * sg.ce - ca * sg.se;
  const float ct = ca * sg.ce + sa * sg.se;
  float obl = (p.baffle == FS_BAFFLE_SOFT) ? ct : (p.baffle == FS_BAFFLE_RIGID) ? 1.0f : ct / (ct + p.baffle_b);
  if (!(ct > 0.0f)) obl = 1e-16f;
  const float rcf = (float)rcl;
  float amp = obl * __builtin_amdgcn_rsqf(rcf) * p.inv_nsub;
  if (!p.fulldir) amp *= sinc_fast(p.kappa * st);
  GeoF g;
  g.t = rcl * p.inv_c;
  g.amp = amp;
  g.att = p.att_coef * rcf * kLog2e;
  g.sint = st;
  return g;
}

__device__ __forceinline__ GeoF geometry_fast(const Problem& p, double xs, double zs, int e, int j) {
  return geometry_fast(p, xs, zs, sub_geo(p, e, j));
}

// amp exp(-att f) exp(i 2 pi frac(f t)) with hardware sin/cos/exp2
__device__ __forceinline__ cf green_fast(const GeoF& g, double f, float amp) {
  const double nu = f * g.t;
  const float ph = (float)(nu - floor(nu));
  const float m = amp * __builtin_amdgcn_exp2f(-g.att * (float)f);
  return {m * __builtin_amdgcn_cosf(ph), m * __builtin_amdgcn_sinf(ph)};
}

// per-step multiplier exp((-att + i 2 pi t) df) with hardware sin/cos/exp2
__device__ __forceinline__ cf green_step_fast(const GeoF& g, double df) {
  const double nu = df * g.t;
  const float ph = (float)(nu - floor(nu));
  const float m = __builtin_amdgcn_exp2f(-g.att * (float)df);
  return {m * __builtin_amdgcn_cosf(ph), m * __builtin_amdgcn_sinf(ph)};
}

// ---------------------------------------------------------------- geometry table entries
// The frequency-independent part of each (scatterer, element, sub-element) term, written once per call
// by k_geo_table (simus_mfma.hip) in rows of `ept` elements, [s][e][j]: the clipped distance r_c (FP64),
// (log2 of the amplitude, FP32 r_c) and the per-frequency step (complex64).
typedef unsigned fs_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned fs_u32x2 __attribute__((ext_vector_type(2)));

// The NSUB table entries of one (scatterer, element): one vector load each.
template <int NSUB>
struct TableTerm {
  double r[NSUB];
  float l[NSUB], rf[NSUB];
  float2 q[NSUB];
};
// Buffer resource over one table row (wave-uniform base, the row's size in
// bytes as its record count: lanes reading past it get zeros, not a fault).
constexpr int kRawBufferWord3 = 0x00020000;  // raw (untyped) buffer, 32-bit data format
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, kRawBufferWord3);
}

// The NSUB entries of table row `row` = s * ept (wave-uniform; ept elements per row) for element e,
// by buffer loads: scalar row address plus the lane's 32-bit byte offset.
template <int NSUB>
__device__ __forceinline__ TableTerm<NSUB> load_term(const double* __restrict__ rt, const float* __restrict__ la,
                                                     const float2* __restrict__ qt, size_t row, int e, int ept) {
  TableTerm<NSUB> t;
  const __amdgpu_buffer_rsrc_t rr = row_rsrc(rt + row * NSUB, ept * NSUB * 8);
  const __amdgpu_buffer_rsrc_t rl = row_rsrc(la + 2 * row * NSUB, ept * NSUB * 8);
  const __amdgpu_buffer_rsrc_t rq = row_rsrc(qt + row * NSUB, ept * NSUB * 8);
  if constexpr (NSUB == 2) {
    const fs_u32x4 r = __builtin_bit_cast(fs_u32x4, __builtin_amdgcn_raw_buffer_load_b128(rr, e * 16, 0, 0));
    const fs_u32x4 l = __builtin_bit_cast(fs_u32x4, __builtin_amdgcn_raw_buffer_load_b128(rl, e * 16, 0, 0));
    const fs_u32x4 q = __builtin_bit_cast(fs_u32x4, __builtin_amdgcn_raw_buffer_load_b128(rq, e * 16, 0, 0));
    t.r[0] = __builtin_bit_cast(double, (fs_u32x2){r.x, r.y});
    t.r[1] = __builtin_bit_cast(double, (fs_u32x2){r.z, r.w});
    t.l[0] = __uint_as_float(l.x);
    t.rf[0] = __uint_as_float(l.y);
    t.l[1] = __uint_as_float(l.z);
    t.rf[1] = __uint_as_float(l.w);
    t.q[0] = make_float2(__uint_as_float(q.x), __uint_as_float(q.y));
    t.q[1] = make_float2(__uint_as_float(q.z), __uint_as_float(q.w));
  } else {
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      const int o = (e * NSUB + j);
      t.r[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rr, o * 8, 0, 0));
      const fs_u32x2 l = __builtin_bit_cast(fs_u32x2, __builtin_amdgcn_raw_buffer_load_b64(rl, o * 8, 0, 0));
      t.l[j] = __uint_as_float(l.x);
      t.rf[j] = __uint_as_float(l.y);
      const fs_u32x2 q = __builtin_bit_cast(fs_u32x2, __builtin_amdgcn_raw_buffer_load_b64(rq, o * 8, 0, 0));
      t.q[j] = make_float2(__uint_as_float(q.x), __uint_as_float(q.y));
    }
  }
  return t;
}

// Split-bf16 form of a pair of FP32 values: hi = bf16(a) | bf16(b) << 16 (round to nearest even) and
// lo = the same of the remainders a - hi_a, b - hi_b (exact in FP32).
typedef __bf16 fs_bf16x2 __attribute__((ext_vector_type(2)));
typedef float fs_floatx2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split_bf16_pair(float a, float b, unsigned& hi, unsigned& lo) {
  const fs_bf16x2 h = __builtin_convertvector((fs_floatx2){a, b}, fs_bf16x2);
  const unsigned hb = __builtin_bit_cast(unsigned, h);
  const float ah = __uint_as_float(hb << 16), bh = __uint_as_float(hb & 0xffff0000u);
  hi = hb;
  lo = __builtin_bit_cast(unsigned, __builtin_convertvector((fs_floatx2){a - ah, b - bh}, fs_bf16x2));
}

int timing_start(hipStream_t st, int cls = FS_TIMING_SPECTRUM);
void timing_stop(int slot, hipStream_t st);

}  // namespace fs
