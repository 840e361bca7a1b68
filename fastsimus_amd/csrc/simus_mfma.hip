// Batched SIMUS through the multistatic matrix on the gfx950 matrix cores.
//
// Every transmit t shares the Green's functions G(f, e, s). Writing the
// reference algorithm (simus.py:263-333, _simus_strategies.py:23-77) per
// frequency f with transmit vector A_t(f, e') = w exp(i 2 pi f tau):
//
//   X_t(f, e) = corr probe(f) sum_s rc_s G(f,e,s) [PP(f) sum_e' G(f,e',s) A_t(f,e')]
//             = corr probe(f) PP(f) sum_e' K_f[e, e'] A_t(f, e'),
//   K_f[e, e'] = sum_{s in field} rc_s G(f, e, s) G(f, e', s)        (complex symmetric, E x E)
//
// so the scatterer sum is done once for all transmits. K_f is a rank-S
// product H^T H with H = sqrt(rc) G (see below for rc < 0), evaluated with
// v_mfma_f32_16x16x32_bf16 in split precision: every operand is hi + lo
// bf16 and hi*hi + hi*lo + lo*hi is accumulated in FP32 (on the diagonal tiles,
// which are symmetric, hi*hi + 2 hi*lo and the symmetric part of the sum).
//
// Complex -> real: the 32-deep MFMA contraction index runs over
// (real part, imaginary part) of 16 scatterers; one fragment of element e
// holds [Re H(e, s0..s3), Im H(e, s0..s3)] for the lane group's 4 scatterers.
// The sign of rc is carried by H itself: H = i sqrt|rc| G for rc < 0 (a
// quarter revolution added to the phase), so that H H^T = rc G G^T. With the
// fragment F(e) on both sides:
//   Re K = (Re H | -Im H)(e) . (Re H | Im H)(e')
//   Im K = (Im H |  Re H)(e) . (Re H | Im H)(e')
// i.e. the A operands are the fragment with its upper half negated or its
// halves swapped, and the B operand is the fragment as stored: one LDS layout
// for both sides, no transposes.
//
// Workgroup = ET waves (one thread per scatterer quad and element), one block
// of FBK frequencies and one chunk of scatterers. Per tile of 16 scatterers:
//   1. every thread evaluates G for 4 scatterers x 1 element at the FBK
//      frequencies (FP64 phase seed once per block, FP32 recurrence inside
//      it; the frequency-independent distance, amplitude and step come from
//      a table built once per call, or are evaluated per block where no table
//      is used) and writes split-bf16 fragments to LDS;
//   2. every wave runs its share of the upper-triangular 16x16 tiles of K_f
//      for its frequencies, with accumulators in registers.
// The LDS holds two tiles: tile n is multiplied while tile n + 1 is generated,
// and the two waves of a SIMD take these in opposite order so that matrix-core
// and vector work overlap. Partial K per scatterer chunk goes to HBM;
// k_msm_apply sums the chunks in a fixed order and applies every transmit
// (E x E x n_tx per frequency, FP32).
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "fs_device.h"

namespace fs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) u32x2 lds_u32x2;  // an LDS (local data share) location

constexpr int MST = 16;  // scatterers per tile (32-deep contraction: re + im)

// LDS of one workgroup: two buffers (the tile being multiplied and the tile
// being generated) of split-bf16 fragments [k][hi|lo][tile][lane].
template <int ET, int FBK>
struct FragLds {
  static constexpr int per_buf = FBK * 2 * ET * 64;  // u32x4 entries per buffer
  static constexpr size_t bytes = (size_t)2 * per_buf * 16;
  __device__ static __forceinline__ int idx(int k, int hl, int tile, int lane) {
    return ((k * 2 + hl) * ET + tile) * 64 + lane;
  }
};


// Twice each of the 8 bf16 in a fragment: one added to every exponent field (bit 7 of each 16-bit
// half). Exact for normal values; a zero becomes 2^-126 and a subnormal lo part at most 2^-125, both
// far below the FP32 resolution of any sum they enter. No carry crosses the halves: that would need
// an exponent field of 255 with the sign set (-inf or NaN), which no Green's-function value has.
__device__ __forceinline__ u32x4 twice_bf16x8(u32x4 v) {
  return v + (u32x4){0x00800080u, 0x00800080u, 0x00800080u, 0x00800080u};
}

__device__ __forceinline__ floatx4 mma(u32x4 a, u32x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

// ---------------------------------------------------------------- tile jobs
// Upper-triangular 16x16 tile pairs (a <= b) of the E x E matrix, a = row (A
// operand), b = column (B operand). WPP waves share a frequency group, each
// owning one list with only two distinct rows: 9 pairs each for 8 tiles
// (E <= 128, WPP = 4), 5 each for 4 tiles (E <= 64, WPP = 2).
template <int ET, int W>
struct JobList;
template <> struct JobList<8, 0> {
  static constexpr int n = 9;
  static constexpr int a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 7};
  static constexpr int b[9] = {0, 1, 2, 3, 4, 5, 6, 7, 7};
};
template <> struct JobList<8, 1> {
  static constexpr int n = 9;
  static constexpr int a[9] = {1, 1, 1, 1, 1, 1, 1, 6, 6};
  static constexpr int b[9] = {1, 2, 3, 4, 5, 6, 7, 6, 7};
};
template <> struct JobList<8, 2> {
  static constexpr int n = 9;
  static constexpr int a[9] = {2, 2, 2, 2, 2, 2, 5, 5, 5};
  static constexpr int b[9] = {2, 3, 4, 5, 6, 7, 5, 6, 7};
};
template <> struct JobList<8, 3> {
  static constexpr int n = 9;
  static constexpr int a[9] = {3, 3, 3, 3, 3, 4, 4, 4, 4};
  static constexpr int b[9] = {3, 4, 5, 6, 7, 4, 5, 6, 7};
};
template <> struct JobList<4, 0> {
  static constexpr int n = 5;
  static constexpr int a[5] = {0, 0, 0, 0, 3};
  static constexpr int b[5] = {0, 1, 2, 3, 3};
};
template <> struct JobList<4, 1> {
  static constexpr int n = 5;
  static constexpr int a[5] = {1, 1, 1, 2, 2};
  static constexpr int b[5] = {1, 2, 3, 2, 3};
};

template <int ET, int W>
__host__ __device__ constexpr bool uses_row(int r) {
  for (int j = 0; j < JobList<ET, W>::n; ++j)
    if (JobList<ET, W>::a[j] == r) return true;
  return false;
}
template <int ET, int W>
__host__ __device__ constexpr bool uses_col(int c) {
  for (int j = 0; j < JobList<ET, W>::n; ++j)
    if (JobList<ET, W>::b[j] == c) return true;
  return false;
}

// Workgroup shape: one thread per (scatterer quad, element) of a tile, i.e.
// ET waves; FBK frequencies per block, KPW per wave, WPP waves per group of KPW
// frequencies. SHALLOW blocks (FBK = 2, KPW = 1: half the accumulators) are
// required where the Green's functions need more registers (more than two
// sub-elements, or directivity at every frequency) and selectable elsewhere;
// the default is FBK = 4, KPW = 2.
template <int ET, bool SHALLOW>
struct MsmShape {
  static constexpr int FBK = SHALLOW ? 2 : 4;
  static constexpr int KPW = FBK / 2;           // frequencies per wave
  static constexpr int NW = ET;                 // waves per workgroup
  static constexpr int THREADS = 64 * NW;
  static constexpr int WPP = NW * KPW / FBK;    // waves sharing one frequency group
  static constexpr int WG_PER_CU = ET == 8 ? 1 : 2;
};

// Green's-function variants of k_msm: FAST = soft baffle, centre-frequency
// directivity with |kappa sin(theta)| <= 2.5 (branch-free); GENERIC = any
// baffle, centre-frequency directivity; FULLDIR = directivity at every
// frequency; TABLE = the frequency-independent part (clipped distance and
// log2 of the amplitude) read from a table that k_geo_table fills once per
// call, for any baffle with a non-negative amplitude and centre-frequency
// directivity. Distances below 1e-15 m are taken as 1e-15 m (both are clipped
// to c / (2 fc) for the amplitude and phase), which only moves the angle of a
// scatterer sitting on a sub-element centroid; GENERIC, FULLDIR and TABLE
// handle an exact zero distance as the reference does (angle pi/2).
enum MsmMode { MSM_FAST = 0, MSM_GENERIC = 1, MSM_FULLDIR = 2, MSM_TABLE = 3 };

// sin(a)/a for |a| <= 2.5: Taylor series to a^14 (truncation < 5e-9 on that range).
__device__ __forceinline__ float sinc_poly_2p5(float a) {
  const float x = a * a;
  float r = fmaf(x, 7.6471637e-13f, -1.6059044e-10f);
  r = fmaf(x, r, 2.5052108e-8f);
  r = fmaf(x, r, -2.7557319e-6f);
  r = fmaf(x, r, 1.9841270e-4f);
  r = fmaf(x, r, -8.3333333e-3f);
  r = fmaf(x, r, 0.16666667f);
  return fmaf(-x, r, 1.0f);
}

// sin(a)/a for |a| <= pi: Taylor series to a^16 (truncation < 1e-8 on that range).
__device__ __forceinline__ float sinc_poly_pi(float a) {
  const float x = a * a;
  float r = fmaf(x, -2.8114572e-15f, 7.6471637e-13f);
  r = fmaf(x, r, -1.6059044e-10f);
  r = fmaf(x, r, 2.5052108e-8f);
  r = fmaf(x, r, -2.7557319e-6f);
  r = fmaf(x, r, 1.9841270e-4f);
  r = fmaf(x, r, -8.3333333e-3f);
  r = fmaf(x, r, 0.16666667f);
  return fmaf(-x, r, 1.0f);
}

// Per-thread element data, constant over the whole launch. NSUB = 0 stands for a sub-element count
// known only at run time (p.NS, more than 4): then the centroids are read where they are used.
template <int NSUB>
struct ElemData {
  double sx[NSUB > 0 ? NSUB : 1], sz[NSUB > 0 ? NSUB : 1];
  float ce, se;
  int ee;     // element index clamped into range
  int ephys;  // the thread's element (may be a padding index >= E, below the table's padded width)
  bool valid;
};

// Elements in groups of 64. The workgroup's element v (0 .. 16 ET - 1) is element v of group ga for
// v < 64 and element v - 64 of group gb above: a work unit of a pair of groups (ga < gb) yields the
// blocks (ga, ga), (ga, gb) and (gb, gb) of the multistatic matrix. With at most 128 elements the
// pair is (0, 1) and the map is the identity.
__device__ __forceinline__ int phys_element(int v, int ga, int gb) { return v < 64 ? 64 * ga + v : 64 * gb + (v - 64); }

template <int NSUB, int ET>
__device__ __forceinline__ ElemData<NSUB> load_element(const Problem& p, int ga, int gb) {
  ElemData<NSUB> d;
  const int e = phys_element(threadIdx.x % (ET * 16), ga, gb);
  d.ephys = e;
  d.valid = e < p.E;
  const int ee = d.valid ? e : 0;
  d.ee = ee;
#pragma unroll
  for (int j = 0; j < NSUB; ++j) {
    d.sx[j] = p.sub_x[ee * NSUB + j];
    d.sz[j] = p.sub_z[ee * NSUB + j];
  }
  d.ce = (float)p.ecos[ee];
  d.se = (float)p.esin[ee];
  return d;
}

// Frequency-block constants: phase and attenuation per metre of path.
struct BlockScalars {
  double nu0;  // f_b / c      (revolutions per metre at the block's first frequency)
  float nud;   // df / c       (revolutions per metre per frequency step)
  float a0;    // att_coef log2(e) f_b
  float ad;    // att_coef log2(e) df
  float fb, df;
  bool poly_ok;
};

__device__ __forceinline__ BlockScalars block_scalars(const Problem& p, double fb_hz) {
  BlockScalars b;
  b.nu0 = fb_hz * p.inv_c;
  b.nud = (float)(p.df * p.inv_c);
  b.a0 = (float)((double)p.att_coef * (double)kLog2e * fb_hz);
  b.ad = (float)((double)p.att_coef * (double)kLog2e * p.df);
  b.fb = (float)fb_hz;
  b.df = (float)p.df;
  b.poly_ok = p.kappa <= kPi;
  return b;
}

// ---------------------------------------------------------------- scatterer prep
// Per-scatterer inputs of k_msm: (x, z, w, 0) with w = sqrt|rc| / n_sub carrying
// the sign of rc; w = 0 for rc = 0 and for scatterers out of the field.
__global__ void k_scat_prep(Problem p, float4* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= p.S) return;
  const float x = p.x[s], z = p.z[s], r = p.rc[s];
  const bool live = r != 0.f && !out_of_field(p, x, z);
  const float w = live ? copysignf(sqrtf(fabsf(r)) * p.inv_nsub, r) : 0.f;
  out[s] = make_float4(x, z, w, 0.f);
}

// Per-scatterer inputs of the plain kernels' geometry table: (x, z, 1 / n_sub, 0), the weight 0 for rc = 0
// and for scatterers out of the field (the plain kernels apply rc themselves).
__global__ void k_scat_prep_unit(Problem p, float4* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= p.S) return;
  const float x = p.x[s], z = p.z[s];
  const bool live = p.rc[s] != 0.f && !out_of_field(p, x, z);
  out[s] = make_float4(x, z, live ? p.inv_nsub : 0.f, 0.f);
}

// ---------------------------------------------------------------- geometry table
// Frequency-independent part of every (scatterer, element, sub-element) term,
// evaluated once per call (reference math: _pfield_math.py:71-247); the distance
// and the phase reduction of the step in FP64, the values stored as FP32 in FP32:
//   rt = r_c = max(r, c / (2 fc))                         (m, FP64)
//   la = (log2(obliquity(theta) directivity(theta) |w| / sqrt(r_c)), float(r_c))
//   qt = exp((-att + i 2 pi / c) r_c df)                  (the per-frequency step)
// with w = sqrt|rc| / n_sub from k_scat_prep (the sign of rc is applied in
// k_msm), directivity at the centre frequency, theta = asin(dx / r) - theta_e
// (pi/2 at r = 0). la = -inf for dead scatterers and padded elements, whose
// rt is left finite (c / (2 fc)) so that their terms are exactly zero; row
// s = S is such a dead row for every element, read by k_msm in place of the
// scatterers that pad the last tile of a unit.
// Layout [s][e][j] with EP (padded element count) elements per scatterer, so
// that a lane of k_msm reads its NSUB terms of one scatterer as one vector load.
// Requires a non-negative amplitude: sinc(kappa sin(theta)) >= 0 (kappa <= pi)
// and an obliquity >= 0 (soft, rigid, or impedance b >= 0).
template <int NSUB>
__global__ void k_geo_table(Problem p, const float4* __restrict__ scat, int EP, double* __restrict__ rt,
                            float* __restrict__ la, float2* __restrict__ qt) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)(p.S + 1) * EP) return;
  const int s = (int)(idx / EP);
  const int e = (int)(idx - (long long)s * EP);
  const float4 sd = s < p.S ? scat[s] : make_float4(0.f, 1.f, 0.f, 0.f);  // row S: dead
  const double w = fabs((double)sd.z);
#pragma unroll
  for (int j = 0; j < NSUB; ++j) {
    double rcl = p.min_dist;
    float l = -INFINITY;
    if (e < p.E && w > 0.0) {
      const double dx = (double)sd.x - p.sub_x[e * NSUB + j];
      const double dz = (double)sd.y - p.sub_z[e * NSUB + j];
      const double r = sqrt(fma(dx, dx, dz * dz));
      rcl = fmax(r, p.min_dist);
      const double ir = r > 0.0 ? 1.0 / r : 0.0;
      // the angle terms and the amplitude are stored in FP32: evaluated in FP32 (library log2f, sinf)
      const float sa = r > 0.0 ? (float)(dx * ir) : 1.f;
      const float ca = (float)(fabs(dz) * ir);
      const float ce = (float)p.ecos[e], se = (float)p.esin[e];
      const float st = sa * ce - ca * se;
      const float ct = ca * ce + sa * se;
      float obl = (p.baffle == FS_BAFFLE_SOFT) ? ct : (p.baffle == FS_BAFFLE_RIGID) ? 1.f : ct / (ct + p.baffle_b);
      if (!(ct > 0.f)) obl = 1e-16f;
      const float a = p.kappa * st;
      const float dir = fabsf(a) < 1e-2f ? fmaf(a * a, fmaf(a * a, 1.f / 120.f, -1.f / 6.f), 1.f) : sinf(a) / a;
      const float amp = obl * dir * ((float)w / sqrtf((float)rcl));
      l = amp > 0.f ? log2f(amp) : -INFINITY;
    }
    // per-frequency step: phase reduced to [-1/2, 1/2] revolution in FP64, then FP32 sincospif / exp2f
    const double rev = p.df * p.inv_c * rcl;
    float qs, qc;
    sincospif(2.f * (float)(rev - rint(rev)), &qs, &qc);
    const float qm = exp2f((float)(-(double)p.att_coef * (double)kLog2e * p.df * rcl));
    rt[idx * NSUB + j] = rcl;
    la[2 * (idx * NSUB + j)] = l;
    la[2 * (idx * NSUB + j) + 1] = (float)rcl;
    qt[idx * NSUB + j] = make_float2(qm * qc, qm * qs);
  }
}

// ---------------------------------------------------------------- G fragments
// Thread (scatterer quad g, element e): for each of the quad's 4 scatterers
// and each sub-element, the frequency-independent geometry once per block
// (FP64 distance and phase reduction), the seed g(f_b) and the step
// q = exp((-att + i 2 pi t) df), then the block's FBK frequencies by the FP32
// recurrence g <- g q. Sub-elements are summed, sqrt|rc| is folded into the
// amplitude and a quarter revolution into the phase where rc < 0 (H = i sqrt|rc| G),
// and every frequency is written as split-bf16 fragments.
template <int NSUB, int FBK, int MODE, int ET>
__device__ __forceinline__ void gen_fragments(const Problem& p, const ElemData<NSUB>& ed, const BlockScalars& bs,
                                              const float4* __restrict__ scat, int s0, int s_end, u32x4* frag) {
  using L = FragLds<ET, FBK>;
  const int ns = NSUB > 0 ? NSUB : p.NS;
  const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / (ET * 16)));  // wave-uniform
  const int e = threadIdx.x % (ET * 16);
  float re[FBK][4], im[FBK][4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int s = s0 + 4 * g + i;
    const bool valid = s < s_end;
    const float4 sd = scat[valid ? s : s_end - 1];  // in-bounds load, masked below
    const float ws = (valid && ed.valid) ? sd.z : 0.f;
    const float w = fabsf(ws);
    const double quarter = ws < 0.f ? 0.25 : 0.0;
    const double xd = (double)sd.x, zd = (double)sd.y;
#pragma unroll
    for (int j = 0; j < ns; ++j) {
      // sub-element centroids: held in registers where they fit, else re-read (cached)
      constexpr bool held = NSUB > 0 && NSUB <= 2 && MODE == MSM_FAST;
      const double sxj = held ? ed.sx[j] : p.sub_x[ed.ee * ns + j];
      const double szj = held ? ed.sz[j] : p.sub_z[ed.ee * ns + j];
      const double dx = xd - sxj;
      const double dz = zd - szj;
      double r2 = fma(dx, dx, dz * dz);
      bool zero = false;
      if constexpr (MODE != MSM_FAST) zero = !(r2 > 0.0);
      r2 = fmax(r2, 1e-30);  // keeps the FP32 estimate below finite
      // 1/r: FP32 hardware estimate, one FP64 Newton step (relative error ~1e-14)
      const float rinv = __builtin_amdgcn_rsqf((float)r2);
      double y = (double)rinv;
      y = y * fma(-0.5 * r2, y * y, 1.5);
      const double rcl = fmax(zero ? 0.0 : r2 * y, p.min_dist);
      const float sa = zero ? 1.0f : (float)dx * rinv;
      const float ca = zero ? 0.0f : fabsf((float)dz) * rinv;
      const float st = sa * ed.ce - ca * ed.se;
      const float ct = ca * ed.ce + sa * ed.se;
      float obl;
      if constexpr (MODE == MSM_FAST) {
        obl = ct;
      } else {
        obl = (p.baffle == FS_BAFFLE_SOFT)    ? ct
              : (p.baffle == FS_BAFFLE_RIGID) ? 1.0f
                                              : ct * __builtin_amdgcn_rcpf(ct + p.baffle_b);
      }
      if (!(ct > 0.0f)) obl = 1e-16f;
      const float rcf = (float)rcl;
      float amp = obl * __builtin_amdgcn_rsqf(rcf) * w;
      if constexpr (MODE == MSM_FAST) {
        amp *= sinc_poly_2p5(p.kappa * st);
      } else if constexpr (MODE == MSM_GENERIC) {
        amp *= bs.poly_ok ? sinc_poly_pi(p.kappa * st) : sinc_fast(p.kappa * st);
      }
      const double nu0 = fma(bs.nu0, rcl, quarter);
      const float ph0 = (float)(nu0 - floor(nu0));
      const float phd = __builtin_amdgcn_fractf(bs.nud * rcf);
      const float m0 = amp * __builtin_amdgcn_exp2f(-bs.a0 * rcf);
      const float md = __builtin_amdgcn_exp2f(-bs.ad * rcf);
      cf gv = {m0 * __builtin_amdgcn_cosf(ph0), m0 * __builtin_amdgcn_sinf(ph0)};
      const cf q = {md * __builtin_amdgcn_cosf(phd), md * __builtin_amdgcn_sinf(phd)};
#pragma unroll
      for (int k = 0; k < FBK; ++k) {
        if constexpr (MODE == MSM_FULLDIR) {
          const float d = sinc_fast(p.kappa_per_hz * fmaf((float)k, bs.df, bs.fb) * st);
          re[k][i] = j == 0 ? d * gv.x : fmaf(d, gv.x, re[k][i]);
          im[k][i] = j == 0 ? d * gv.y : fmaf(d, gv.y, im[k][i]);
        } else {
          re[k][i] = j == 0 ? gv.x : re[k][i] + gv.x;
          im[k][i] = j == 0 ? gv.y : im[k][i] + gv.y;
        }
        if (k + 1 < FBK) gv = cmul(gv, q);
      }
    }
  }
  const int tile = e >> 4;
  const int lane = (e & 15) + 16 * g;
#pragma unroll
  for (int k = 0; k < FBK; ++k) {
    unsigned h0, h1, h2, h3, l0, l1, l2, l3;
    split_bf16_pair(re[k][0], re[k][1], h0, l0);
    split_bf16_pair(re[k][2], re[k][3], h1, l1);
    split_bf16_pair(im[k][0], im[k][1], h2, l2);
    split_bf16_pair(im[k][2], im[k][3], h3, l3);
    frag[L::idx(k, 0, tile, lane)] = (u32x4){h0, h1, h2, h3};
    frag[L::idx(k, 1, tile, lane)] = (u32x4){l0, l1, l2, l3};
  }
}

// One thread's table entries for one tile (its 4 scatterers).
template <int NSUB>
struct TableTile {
  TableTerm<NSUB> t[4];
  int wbits[4];  // bits of w = +-sqrt|rc| / n_sub: the sign of rc (wave-uniform)
};

// ept: elements per table row (the element count padded to the element groups), e: the thread's element
template <int NSUB, int ET>
__device__ __forceinline__ void fetch_table(const float4* __restrict__ scat, const double* __restrict__ rt,
                                            const float* __restrict__ la, const float2* __restrict__ qt, int s0,
                                            int s_end, int s_dead, int ept, int e, TableTile<NSUB>& tt) {
  constexpr int EP = ET * 16;
  const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / EP));  // wave-uniform
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int s = s0 + 4 * g + i;
    const bool valid = s < s_end;
    // scatterers past the unit's end read the table's dead row (amplitude 2^-inf = 0)
    tt.wbits[i] = __float_as_int(scat[valid ? s : s_end - 1].z);
    tt.t[i] = load_term<NSUB>(rt, la, qt, (size_t)(valid ? s : s_dead) * ept, e, ept);
  }
}

// Fragments from the geometry table: per (scatterer, sub-element) only the
// block's seed is evaluated here: the phase frac(f_b r_c / c + quarter) in
// FP64 and the magnitude 2^(la - att f_b r_c log2 e); the block's FBK
// frequencies follow by the FP32 recurrence with the tabulated step q, as in
// gen_fragments.
template <int NSUB, int FBK, int ET>
__device__ __forceinline__ void gen_fragments_table(const BlockScalars& bs, const TableTile<NSUB>& tt, u32x4* frag) {
  using L = FragLds<ET, FBK>;
  constexpr int EP = ET * 16;
  const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / EP));  // wave-uniform
  const int e = threadIdx.x % EP;
  float re[FBK][4], im[FBK][4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double quarter = tt.wbits[i] < 0 ? 0.25 : 0.0;
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      const double r = tt.t[i].r[j];
      const float lj = tt.t[i].l[j];
      const float rf = tt.t[i].rf[j];
      const float ph0 = (float)__builtin_amdgcn_fract(fma(bs.nu0, r, quarter));  // v_fract_f64: x - floor(x)
      const float m0 = __builtin_amdgcn_exp2f(fmaf(-bs.a0, rf, lj));
      cf gv = {m0 * __builtin_amdgcn_cosf(ph0), m0 * __builtin_amdgcn_sinf(ph0)};
      const cf q = {tt.t[i].q[j].x, tt.t[i].q[j].y};
#pragma unroll
      for (int k = 0; k < FBK; ++k) {
        re[k][i] = j == 0 ? gv.x : re[k][i] + gv.x;
        im[k][i] = j == 0 ? gv.y : im[k][i] + gv.y;
        if (k + 1 < FBK) gv = cmul(gv, q);
      }
    }
  }
  const int tile = e >> 4;
  const int lane = (e & 15) + 16 * g;
#pragma unroll
  for (int k = 0; k < FBK; ++k) {
    unsigned h0, h1, h2, h3, l0, l1, l2, l3;
    split_bf16_pair(re[k][0], re[k][1], h0, l0);
    split_bf16_pair(re[k][2], re[k][3], h1, l1);
    split_bf16_pair(im[k][0], im[k][1], h2, l2);
    split_bf16_pair(im[k][2], im[k][3], h3, l3);
    frag[L::idx(k, 0, tile, lane)] = (u32x4){h0, h1, h2, h3};
    frag[L::idx(k, 1, tile, lane)] = (u32x4){l0, l1, l2, l3};
  }
}

// ---------------------------------------------------------------- MFMA tiles of one wave
// Accumulators of one wave: its job list at the KPW frequencies k0 .. k0 + KPW - 1 of the block.
template <int ET, int FBK, int KPW, int W>
struct WaveTiles {
  using J = JobList<ET, W>;
  static constexpr int NJ = J::n;
  floatx4 acc[KPW][NJ][2];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < KPW; ++k)
#pragma unroll
      for (int j = 0; j < NJ; ++j) acc[k][j][0] = acc[k][j][1] = (floatx4){0.f, 0.f, 0.f, 0.f};
  }

  // One 16-scatterer tile: for each of the wave's frequencies, build the A
  // operands of its two rows (Re K: upper half negated, Im K: halves swapped),
  // then stream its columns.
  __device__ __forceinline__ void step(const u32x4* frag, int k0) {
    using L = FragLds<ET, FBK>;
    const int lane = threadIdx.x & 63;
    const u32x4 nim = {0u, 0u, 0x80008000u, 0x80008000u};
    // Row operands are read from LDS in the shape they are used in: as stored
    // (then negated in place) and with halves swapped (two 8-byte volatile
    // reads, which the compiler does not merge, straight into one register
    // quad), instead of one read plus register moves. The byte offset of the
    // negated read passes through an empty asm statement so that the compiler
    // cannot prove it equal to the column reads and share one read.
    unsigned o_neg = (unsigned)lane * 16u;
    asm volatile("" : "+v"(o_neg));
    const unsigned o_zw = (unsigned)lane * 16u + 8u, o_xy = (unsigned)lane * 16u;
    const char* fbytes = (const char*)frag;
#pragma unroll
    for (int kk = 0; kk < KPW; ++kk) {
      const int k = k0 + kk;
      u32x4 arh[ET], aih[ET], arl[ET], ail[ET];
#pragma unroll
      for (int r = 0; r < ET; ++r) {
        if (!uses_row<ET, W>(r)) continue;
        const unsigned eh = (unsigned)L::idx(k, 0, r, 0) * 16u, el = (unsigned)L::idx(k, 1, r, 0) * 16u;
        arh[r] = *(const u32x4*)(fbytes + o_neg + eh) ^ nim;
        arl[r] = *(const u32x4*)(fbytes + o_neg + el) ^ nim;
        const u32x2 hzw = *(const volatile lds_u32x2*)(fbytes + o_zw + eh);
        const u32x2 hxy = *(const volatile lds_u32x2*)(fbytes + o_xy + eh);
        const u32x2 lzw = *(const volatile lds_u32x2*)(fbytes + o_zw + el);
        const u32x2 lxy = *(const volatile lds_u32x2*)(fbytes + o_xy + el);
        aih[r] = (u32x4){hzw.x, hzw.y, hxy.x, hxy.y};
        ail[r] = (u32x4){lzw.x, lzw.y, lxy.x, lxy.y};
      }
#pragma unroll
      for (int c = 0; c < ET; ++c) {
        if (!uses_col<ET, W>(c)) continue;
        const u32x4 bh = frag[L::idx(k, 0, c, lane)];
        const u32x4 bl = frag[L::idx(k, 1, c, lane)];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          if (J::b[j] != c) continue;
          const int r = J::a[j];
          floatx4 kr = acc[kk][j][0], ki = acc[kk][j][1];
          kr = mma(arh[r], bh, kr);
          ki = mma(aih[r], bh, ki);
          if (r == c) {
            // Diagonal tile: with C = hi lo^T, the split product is D + C + C^T (D = hi hi^T and the
            // tile are symmetric). Accumulating D + 2 C, whose symmetric part k_msm_fold takes, needs
            // two products instead of three.
            const u32x4 bl2 = twice_bf16x8(bl);
            kr = mma(arh[r], bl2, kr);
            ki = mma(aih[r], bl2, ki);
          } else {
            kr = mma(arh[r], bl, kr);
            ki = mma(aih[r], bl, ki);
            kr = mma(arl[r], bh, kr);
            ki = mma(ail[r], bh, ki);
          }
          acc[kk][j][0] = kr;
          acc[kk][j][1] = ki;
        }
      }
    }
  }

  // Kpart[chunk][kglob][e][e'] (complex), D layout: lane -> (row 4*(lane/16)+i, col lane%16)
  __device__ __forceinline__ void store(cf* kpart, int E_pad, int kglob0) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int kk = 0; kk < KPW; ++kk) {
      cf* Kk = kpart + (size_t)(kglob0 + kk) * E_pad * E_pad;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int col = 16 * J::b[j] + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * J::a[j] + 4 * (lane >> 4) + i;
          Kk[(size_t)row * E_pad + col] = {acc[kk][j][0][i], acc[kk][j][1][i]};
        }
      }
    }
  }
};

// One wave's part of a workgroup: tiles are generated one ahead into the other
// LDS buffer. The two waves that share a SIMD run in opposite order within an
// iteration (one multiplies while the other generates) so that matrix-core and
// vector work overlap; one barrier per tile separates the buffers' roles.
template <int NSUB, int ET, int MODE, bool SHALLOW, int W>
__device__ __forceinline__ void msm_wave(const Problem& p, const float4* __restrict__ scat,
                                         const double* __restrict__ rt, const float* __restrict__ la,
                                         const float2* __restrict__ qt, int ept, int sb, int se, int ga, int gb,
                                         int fb, int k0, bool mfma_first, char* lds, cf* kpart) {
  using SH = MsmShape<ET, SHALLOW>;
  constexpr int FBK = SH::FBK;
  using L = FragLds<ET, FBK>;
  u32x4* const frag = (u32x4*)lds;  // buffer b at frag + b * per_buf
  const ElemData<NSUB> ed = load_element<NSUB, ET>(p, ga, gb);
  const BlockScalars bs = block_scalars(p, p.f0 + (double)(fb * FBK) * p.df);
  WaveTiles<ET, FBK, SH::KPW, W> t;
  t.zero();
  int buf = 0;
  if constexpr (MODE == MSM_TABLE) {
    TableTile<NSUB> nxt;
    if (sb < se) {
      fetch_table<NSUB, ET>(scat, rt, la, qt, sb, se, p.S, ept, ed.ephys, nxt);
      gen_fragments_table<NSUB, FBK, ET>(bs, nxt, frag);
    }
    __syncthreads();
    for (int s0 = sb; s0 < se; s0 += MST) {
      const bool more = s0 + MST < se;
      const int nb = buf ^ 1;
      if (mfma_first) t.step(frag + buf * L::per_buf, k0);
      if (more) {
        if (!mfma_first) __builtin_amdgcn_s_setprio(1);
        fetch_table<NSUB, ET>(scat, rt, la, qt, s0 + MST, se, p.S, ept, ed.ephys, nxt);
        gen_fragments_table<NSUB, FBK, ET>(bs, nxt, frag + nb * L::per_buf);
        __builtin_amdgcn_s_setprio(0);
      }
      if (!mfma_first) t.step(frag + buf * L::per_buf, k0);
      __syncthreads();
      buf = nb;
    }
  } else {
    if (sb < se) gen_fragments<NSUB, FBK, MODE, ET>(p, ed, bs, scat, sb, se, frag);
    __syncthreads();
    for (int s0 = sb; s0 < se; s0 += MST) {
      const bool more = s0 + MST < se;
      const int nb = buf ^ 1;
      if (mfma_first) t.step(frag + buf * L::per_buf, k0);
      if (more) {
        // The generate-first wave of a SIMD gets the vector units first, so that its
        // multiply phase starts while the other wave still has vector work to overlap.
        if (!mfma_first) __builtin_amdgcn_s_setprio(1);
        gen_fragments<NSUB, FBK, MODE, ET>(p, ed, bs, scat, s0 + MST, se, frag + nb * L::per_buf);
        __builtin_amdgcn_s_setprio(0);
      }
      if (!mfma_first) t.step(frag + buf * L::per_buf, k0);
      __syncthreads();
      buf = nb;
    }
  }
  t.store(kpart, ET * 16, fb * FBK + k0);
}

// ---------------------------------------------------------------- work units
// A unit is a contiguous scatterer range of one phantom and one pair of element groups
// (ga | gb << 16) with its own partial K slot; an item is (unit, frequency block). Units are built
// on the host, largest first, and passed by value (so that a captured graph replays them).
struct MsmUnit {
  int s_begin, s_end, slot, pair;
};
constexpr int MSM_MAX_UNITS = 192;
// Element limit of the batched method: 16 groups of 64, i.e. 120 pairs of groups, each a work unit of a
// scatterer chunk (at most MSM_MAX_UNITS units per launch).
constexpr int MSM_MAX_ELEMENTS = 1024;
struct MsmUnits {
  int n;    // units
  int nfb;  // frequency blocks
  int pad0, pad1;
  MsmUnit u[MSM_MAX_UNITS];
};
constexpr int MSM_MAX_PHANTOMS = 192;
struct MsmPhantoms {
  int n;       // phantoms
  int groups;  // element groups of 64 (units per scatterer chunk: one, or one per pair of groups above 2 groups)
  int2 ph[MSM_MAX_PHANTOMS];  // (first K slot, scatterer chunks) of each phantom: pair q, chunk c at first + q chunks + c
};

// Pair index of element groups a < b among G groups (pairs ordered (0,1), (0,2), .., (1,2), ..)
__host__ __device__ __forceinline__ int group_pair_index(int a, int b, int G) { return a * G - a * (a + 1) / 2 + (b - a - 1); }

// Where K[r][col] of physical elements lives among the partial matrices of the pairs of element groups:
// the pair and the (virtual) row and column in its 16 ET x 16 ET matrix. Off-diagonal blocks come from
// the pair of their two groups; the diagonal block of group g from the pair (g, g + 1), the last one's
// from (G - 2, G - 1). With at most two groups there is one pair and the map is the identity.
__device__ __forceinline__ void k_location(int r, int col, int G, int& pair, int& vr, int& vc) {
  if (G <= 2) {
    pair = 0;
    vr = r;
    vc = col;
    return;
  }
  const int gr = r >> 6, gc = col >> 6;
  int a, b;
  if (gr == gc) {
    a = gr + 1 < G ? gr : G - 2;
    b = a + 1;
  } else {
    a = min(gr, gc);
    b = max(gr, gc);
  }
  vr = (gr == a ? 0 : 64) + (r & 63);
  vc = (gc == a ? 0 : 64) + (col & 63);
  pair = group_pair_index(a, b, G);
}

// Persistent: every workgroup takes items from a queue (atomic counter, reset
// by the host before the launch) until none is left; items come largest unit
// first. Every wave runs the same sequence of barriers: two per item, one
// before its tile loop and one per tile.
template <int NSUB, int ET, int MODE, bool SHALLOW, int W>
__device__ __forceinline__ void msm_persistent(const Problem& p, const float4* __restrict__ scat,
                                               const double* __restrict__ rt, const float* __restrict__ la,
                                               const float2* __restrict__ qt, const MsmUnits& units,
                                               int* __restrict__ queue, int nfpad, int ept,
                                               cf* __restrict__ kpart, int k0, bool mfma_first, char* lds,
                                               int* next_item) {
  const int n_items = units.n * units.nfb;
  for (;;) {
    if (threadIdx.x == 0) *next_item = atomicAdd(queue, 1);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(*next_item);
    __syncthreads();
    if (item >= n_items) break;  // uniform over the workgroup
    const int ui = item / units.nfb;
    const int fb = item - ui * units.nfb;
    const MsmUnit u = units.u[ui];
    cf* kp = kpart + (size_t)u.slot * nfpad * (ET * 16) * (ET * 16);
    msm_wave<NSUB, ET, MODE, SHALLOW, W>(p, scat, rt, la, qt, ept, u.s_begin, u.s_end, u.pair & 0xffff, u.pair >> 16,
                                         fb, k0, mfma_first, lds, kp);
  }
}

// ET waves per workgroup; wave w owns frequencies KPW (w / WPP) ..
// KPW (w / WPP) + KPW - 1 of each item's block and job list w % WPP. Waves w
// and w + 4 (and w, w + 1) take opposite orders of multiply and generate.
template <int NSUB, int ET, int MODE, bool SHALLOW>
__global__ __launch_bounds__(64 * ET, ET == 8 ? 1 : 2) void k_msm(Problem p, const float4* __restrict__ scat,
                                                                 const double* __restrict__ rt,
                                                                 const float* __restrict__ la,
                                                                 const float2* __restrict__ qt, const MsmUnits units,
                                                                 int* __restrict__ queue, int nfpad, int ept,
                                                                 cf* __restrict__ kpart) {
  using SH = MsmShape<ET, SHALLOW>;
  extern __shared__ __attribute__((aligned(16))) char lds_raw[];
  __shared__ int next_item;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int k0 = SH::KPW * (w / SH::WPP);
  const bool mfma_first = ((w ^ (w >> 2)) & 1) == 0;
  if constexpr (SH::WPP == 4) {
    switch (w % 4) {
      case 0: msm_persistent<NSUB, ET, MODE, SHALLOW, 0>(p, scat, rt, la, qt, units, queue, nfpad, ept, kpart, k0, mfma_first, lds_raw, &next_item); break;
      case 1: msm_persistent<NSUB, ET, MODE, SHALLOW, 1>(p, scat, rt, la, qt, units, queue, nfpad, ept, kpart, k0, mfma_first, lds_raw, &next_item); break;
      case 2: msm_persistent<NSUB, ET, MODE, SHALLOW, 2>(p, scat, rt, la, qt, units, queue, nfpad, ept, kpart, k0, mfma_first, lds_raw, &next_item); break;
      default: msm_persistent<NSUB, ET, MODE, SHALLOW, 3>(p, scat, rt, la, qt, units, queue, nfpad, ept, kpart, k0, mfma_first, lds_raw, &next_item); break;
    }
  } else {
    if (w % 2 == 0)
      msm_persistent<NSUB, ET, MODE, SHALLOW, 0>(p, scat, rt, la, qt, units, queue, nfpad, ept, kpart, k0, mfma_first, lds_raw, &next_item);
    else
      msm_persistent<NSUB, ET, MODE, SHALLOW, 1>(p, scat, rt, la, qt, units, queue, nfpad, ept, kpart, k0, mfma_first, lds_raw, &next_item);
  }
}

__global__ void k_set_int(int* __restrict__ ptr, int value) {
  if (threadIdx.x == 0) *ptr = value;
}

// ---------------------------------------------------------------- fold the scatterer chunks
// Sums the partial multistatic matrices of every (phantom, pair of element groups) over its scatterer
// chunks, in chunk order, into the first chunk's slot, mirrors each off-diagonal upper tile into
// the lower triangle and replaces each diagonal tile A (k_msm accumulates hi hi^T + 2 hi lo^T there)
// by its symmetric part (A + A^T) / 2, so that k_msm_apply reads every matrix whole, row by row, from
// one slot. One workgroup per (frequency, phantom, pair), one 16 x 16 tile at a time through LDS, so
// that both the direct and the mirrored writes are row-contiguous.
__global__ __launch_bounds__(256) void k_msm_fold(int E_pad, int nfpad, const MsmPhantoms phantoms,
                                                  cf* __restrict__ kpart) {
  __shared__ float2 tile[16][17];
  const int k = blockIdx.x, ph = blockIdx.y, pair = blockIdx.z;
  const int nch = phantoms.ph[ph].y;
  if (nch == 0) return;  // k_msm_apply reads no slot for such a phantom
  const int first = phantoms.ph[ph].x + pair * nch;
  const int ET = E_pad / 16;
  const int rr = threadIdx.x >> 4, cc = threadIdx.x & 15;
  cf* K0 = kpart + ((size_t)first * nfpad + k) * E_pad * E_pad;
  for (int a = 0; a < ET; ++a) {
    for (int b = a; b < ET; ++b) {
      const size_t o = (size_t)(16 * a + rr) * E_pad + 16 * b + cc;
      float sx = 0.f, sy = 0.f;
      for (int c = 0; c < nch; ++c) {
        const cf w = K0[(size_t)c * nfpad * E_pad * E_pad + o];
        sx += w.x;
        sy += w.y;
      }
      tile[rr][cc] = make_float2(sx, sy);
      __syncthreads();
      const float2 t = tile[cc][rr];  // element (16 b + rr, 16 a + cc) of the mirrored tile
      if (a != b) {
        K0[o] = {sx, sy};
        K0[(size_t)(16 * b + rr) * E_pad + 16 * a + cc] = {t.x, t.y};
      } else {
        K0[o] = {0.5f * (sx + t.x), 0.5f * (sy + t.y)};  // t = A^T: exactly symmetric (x + y = y + x)
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------- apply all transmits
// out[t][k][e] = corr probe_k PP_k sum_e' K_k[e][e'] A_t(k, e'), K_k whole in the first chunk's slot (k_msm_fold)
constexpr int AP_CH = 32;  // e' per LDS chunk

// Grid (frequency, phantom, block of 128 elements); phantom ph reads its folded K and writes
// out[ph][t][k][e] for the block's elements. MAXT transmits per thread and pass (2 MAXT per pass over
// the two thread parities): 40, or fewer when the problem has fewer transmits, so that the per-thread
// accumulation does no padded work.
template <int MAXT>
__global__ __launch_bounds__(256) void k_msm_apply(Problem p, int E_pad, const MsmPhantoms phantoms, int nfpad,
                                                   const cf* __restrict__ kpart, cf* __restrict__ out_all) {
  __shared__ cf Ks[128][AP_CH + 1];
  __shared__ cf As[AP_CH][2 * MAXT + 1];
  const int k = blockIdx.x;
  const int slot0 = phantoms.ph[blockIdx.y].x, nchunks = phantoms.ph[blockIdx.y].y;
  const int G = phantoms.groups;
  const int rbase = blockIdx.z * 128;
  cf* __restrict__ out = out_all + (size_t)blockIdx.y * p.NT * p.NF * p.E;
  const int tid = threadIdx.x;
  const int e = rbase + (tid & 127);
  const int tg = tid >> 7;  // transmit parity
  const double f = p.f0 + (double)k * p.df;
  for (int t0 = 0; t0 < p.NT; t0 += 2 * MAXT) {
    cf acc[MAXT];
#pragma unroll
    for (int i = 0; i < MAXT; ++i) acc[i] = {0.f, 0.f};
    for (int c0 = 0; c0 < p.E; c0 += AP_CH) {
      __syncthreads();
      for (int q = tid; q < 128 * AP_CH; q += 256) {
        const int rr = q / AP_CH, cc = q % AP_CH;
        const int r = rbase + rr, col = c0 + cc;
        cf v = {0.f, 0.f};
        if (r < p.E && col < p.E && nchunks > 0) {
          int pair, vr, vc;
          k_location(r, col, G, pair, vr, vc);
          const int first = slot0 + pair * nchunks;
          v = kpart[(((size_t)first * nfpad + k) * E_pad + vr) * E_pad + vc];
        }
        Ks[rr][cc] = v;
      }
      for (int q = tid; q < AP_CH * 2 * MAXT; q += 256) {
        const int cc = q % AP_CH, ti = q / AP_CH;
        const int t = t0 + ti, col = c0 + cc;
        cf a = {0.f, 0.f};
        if (t < p.NT && col < p.E) {
          const double tau = p.delay[(size_t)t * p.E + col];
          const double wgt = p.apod[(size_t)t * p.E + col];
          const double nu = f * tau;
          const float ph = (float)(nu - floor(nu));
          float s, c;
          sincospif(2.0f * ph, &s, &c);
          a = {(float)wgt * c, (float)wgt * s};
        }
        As[cc][ti] = a;
      }
      __syncthreads();
      if (e < p.E) {
        const int cmax = min(AP_CH, p.E - c0);
        for (int cc = 0; cc < cmax; ++cc) {
          const cf kv = Ks[e - rbase][cc];
#pragma unroll
          for (int i = 0; i < MAXT; ++i) cmac(acc[i], kv, As[cc][2 * i + tg]);
        }
      }
    }
    if (e < p.E) {
      const double q = p.probe[k];
      const float sre = (float)(p.pulse_re[k] * q * q * p.corr), sim = (float)(p.pulse_im[k] * q * q * p.corr);
#pragma unroll
      for (int i = 0; i < MAXT; ++i) {
        const int t = t0 + 2 * i + tg;
        if (t < p.NT) out[((size_t)t * p.NF + k) * p.E + e] = cmul(acc[i], cf{sre, sim});
      }
    }
  }
}

// The same product on the matrix cores, for at most 128 elements (one pair of element groups, K whole
// in the first chunk's slot): per frequency X (E_pad x transmits) = K (E_pad x E_pad) A, a complex GEMM
// in split bf16 as in k_msm (hi*hi + hi*lo + lo*hi, FP32 accumulation), with the complex parts in the
// 32-deep contraction: Re X = [Re K | -Im K] . [Re A | Im A], Im X = [Im K | Re K] . [Re A | Im A].
// Grid (frequency, phantom), 4 waves. The transmit weights of ATT 16-transmit tiles (at most) are
// built once per pass into LDS as B fragments (lane l: transmit 16 tt + l % 16, elements
// 16 c + 4 (l / 16) + i); every wave loads the K fragments of its ET / 4 row tiles straight from HBM
// (lane l: row 16 a + l % 16, the 4 consecutive columns of its k-block) and accumulates its
// ET / 4 x ATT output tiles.
template <int ET, int ATT>
__global__ __launch_bounds__(256) void k_msm_apply_mfma(Problem p, const MsmPhantoms phantoms, int nfpad,
                                                        const cf* __restrict__ kpart, cf* __restrict__ out_all) {
  constexpr int EP = ET * 16;
  constexpr int RT = ET / 4;  // row tiles per wave
  extern __shared__ __attribute__((aligned(16))) char lds_raw[];
  u32x4* bfrag = (u32x4*)lds_raw;  // [tt][c][hl][lane]
  const int k = blockIdx.x;
  const int slot0 = phantoms.ph[blockIdx.y].x, nchunks = phantoms.ph[blockIdx.y].y;
  cf* __restrict__ out = out_all + (size_t)blockIdx.y * p.NT * p.NF * p.E;
  const cf* __restrict__ K = kpart + ((size_t)slot0 * nfpad + k) * EP * EP;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const double f = p.f0 + (double)k * p.df;
  const u32x4 nim = {0u, 0u, 0x80008000u, 0x80008000u};
  const double q = p.probe[k];
  const cf scale = {(float)(p.pulse_re[k] * q * q * p.corr), (float)(p.pulse_im[k] * q * q * p.corr)};
  for (int t0 = 0; t0 < p.NT; t0 += 16 * ATT) {
    const int ntt = min(ATT, (p.NT - t0 + 15) / 16);
    // B fragments of the transmit weights A(e', t) = w exp(i 2 pi f tau), zero past NT and E
    for (int idx = tid; idx < ntt * ET * 64; idx += 256) {
      const int l = idx & 63, c = (idx >> 6) % ET, tt = (idx >> 6) / ET;
      const int t = t0 + 16 * tt + (l & 15);
      float re[4], im[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int col = 16 * c + 4 * (l >> 4) + i;
        re[i] = im[i] = 0.f;
        if (t < p.NT && col < p.E) {
          const double tau = p.delay[(size_t)t * p.E + col];
          const double wgt = p.apod[(size_t)t * p.E + col];
          const double nu = f * tau;
          const float ph = (float)(nu - floor(nu));
          float sn, cs;
          sincospif(2.0f * ph, &sn, &cs);
          re[i] = (float)wgt * cs;
          im[i] = (float)wgt * sn;
        }
      }
      unsigned h0, h1, h2, h3, l0, l1, l2, l3;
      split_bf16_pair(re[0], re[1], h0, l0);
      split_bf16_pair(re[2], re[3], h1, l1);
      split_bf16_pair(im[0], im[1], h2, l2);
      split_bf16_pair(im[2], im[3], h3, l3);
      bfrag[((tt * ET + c) * 2 + 0) * 64 + l] = (u32x4){h0, h1, h2, h3};
      bfrag[((tt * ET + c) * 2 + 1) * 64 + l] = (u32x4){l0, l1, l2, l3};
    }
    __syncthreads();
    floatx4 acc[RT][ATT][2];
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
      for (int tt = 0; tt < ATT; ++tt) acc[a][tt][0] = acc[a][tt][1] = (floatx4){0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < ET; ++c) {
      u32x4 arh[RT], arl[RT], aih[RT], ail[RT];
#pragma unroll
      for (int a = 0; a < RT; ++a) {
        const int row = 16 * (RT * w + a) + (lane & 15);
        const int col = 16 * c + 4 * (lane >> 4);
        float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0;
        if (nchunks > 0) {
          const float4* kr = (const float4*)(K + (size_t)row * EP + col);
          v0 = kr[0];  // K[row][col], K[row][col + 1]
          v1 = kr[1];  // K[row][col + 2], K[row][col + 3]
        }
        unsigned h0, h1, h2, h3, l0, l1, l2, l3;
        split_bf16_pair(v0.x, v0.z, h0, l0);  // Re of columns col, col + 1
        split_bf16_pair(v1.x, v1.z, h1, l1);  // Re of col + 2, col + 3
        split_bf16_pair(v0.y, v0.w, h2, l2);  // Im
        split_bf16_pair(v1.y, v1.w, h3, l3);
        arh[a] = (u32x4){h0, h1, h2, h3} ^ nim;
        arl[a] = (u32x4){l0, l1, l2, l3} ^ nim;
        aih[a] = (u32x4){h2, h3, h0, h1};
        ail[a] = (u32x4){l2, l3, l0, l1};
      }
#pragma unroll
      for (int tt = 0; tt < ATT; ++tt) {
        if (tt >= ntt) continue;  // uniform
        const u32x4 bh = bfrag[((tt * ET + c) * 2 + 0) * 64 + lane];
        const u32x4 bl = bfrag[((tt * ET + c) * 2 + 1) * 64 + lane];
#pragma unroll
        for (int a = 0; a < RT; ++a) {
          floatx4 xr = acc[a][tt][0], xi = acc[a][tt][1];
          xr = mma(arh[a], bh, xr);
          xi = mma(aih[a], bh, xi);
          xr = mma(arh[a], bl, xr);
          xi = mma(aih[a], bl, xi);
          xr = mma(arl[a], bh, xr);
          xi = mma(ail[a], bh, xi);
          acc[a][tt][0] = xr;
          acc[a][tt][1] = xi;
        }
      }
    }
    // D layout: lane l holds rows (elements) 4 (l / 16) + i and column (transmit) l % 16 of its tile
#pragma unroll
    for (int tt = 0; tt < ATT; ++tt) {
      if (tt >= ntt) continue;  // uniform
      const int t = t0 + 16 * tt + (lane & 15);
#pragma unroll
      for (int a = 0; a < RT; ++a) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int e = 16 * (RT * w + a) + 4 * (lane >> 4) + i;
          if (t < p.NT && e < p.E)
            out[((size_t)t * p.NF + k) * p.E + e] = cmul(cf{acc[a][tt][0][i], acc[a][tt][1][i]}, scale);
        }
      }
    }
    __syncthreads();  // the next pass rewrites the B fragments
  }
}

// ---------------------------------------------------------------- host side
struct MsmPlan {
  int ET, FBK, nfb, nfpad, threads, grid, n_items;
  int groups;  // element groups of 64
  int npair;   // units per scatterer chunk: 1, or one per pair of groups above 2 groups
  int ept;     // elements per geometry-table row: 64 groups
  bool shallow, table;
  size_t lds, kpart_bytes, scat_offset, rt_offset, la_offset, qt_offset, queue_offset, ws_bytes;
  MsmUnits units;
  MsmPhantoms phantoms;
};

constexpr int MAX_DEVICES = 64;

static int current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return 0;
  return dev;
}

// CU count of the current device, queried once per device (not a stream operation, so safe
// while a stream is being captured).
static int device_cus() {
  static int cached[MAX_DEVICES] = {0};
  const int dev = current_device();
  if (cached[dev] <= 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}

// ---------------------------------------------------------------- tuning cache
// Tile-shape choices of k_msm per problem signature (device, element tiles,
// sub-elements, Green's-function variant, frequencies, phantom count, log2 of
// the scatterer tiles), set by an autotuner; absent entries use the defaults.
static bool msm_fast_mode(const Problem& p) { return !p.fulldir && p.baffle == FS_BAFFLE_SOFT && p.kappa <= 2.5f; }

static bool shallow_required(const Problem& p) { return p.NS > 2 || p.fulldir; }

// The geometry table holds log2 of the amplitude, so it needs amplitudes >= 0:
// centre-frequency directivity with kappa <= pi and a non-negative obliquity; and a lane holds the
// table entries of all its sub-elements, so at most 4 of them.
static bool table_possible(const Problem& p) {
  return p.NS <= 4 && !p.fulldir && p.kappa <= kPi && !(p.baffle == FS_BAFFLE_IMPEDANCE && !(p.baffle_b >= 0.f));
}
constexpr size_t MSM_TABLE_AUTO_MAX = (size_t)32 << 30;  // bytes of table used without being asked

static std::array<long long, 8> tuning_key(const Problem& p, int P) {
  long long tiles = std::max(1, (p.S + MST - 1) / MST), lg = 0;
  while ((1LL << (lg + 1)) <= tiles) ++lg;
  return {(long long)current_device(), (p.E + 63) / 64, p.NS, p.fulldir ? 2 : (msm_fast_mode(p) ? 0 : 1),
          p.NF, P, lg, 0};
}

static std::mutex tuning_mutex;
static std::map<std::array<long long, 8>, fs_tuning> tuning_table;

fs_tuning msm_tuning(const Problem& p, int P) {
  std::lock_guard<std::mutex> lock(tuning_mutex);
  const auto it = tuning_table.find(tuning_key(p, P));
  if (it != tuning_table.end()) return it->second;
  fs_tuning t{};
  return t;
}

int msm_tuning_set(const Problem& p, int P, const fs_tuning* t) {
  std::lock_guard<std::mutex> lock(tuning_mutex);
  if (!t) {
    tuning_table.erase(tuning_key(p, P));
    return FS_OK;
  }
  if ((t->msm_fbk != 0 && t->msm_fbk != 2 && t->msm_fbk != 4) || t->msm_units < 0 || t->msm_units > MSM_MAX_UNITS ||
      t->msm_item_div < 0 || t->msm_item_div > 1024 || t->msm_table < 0 || t->msm_table > 2)
    return FS_ERR_INVALID;
  tuning_table[tuning_key(p, P)] = *t;
  return FS_OK;
}

// Scatterer chunks of a single phantom: the count whose last round of
// equal items is fullest (rounds / chunks smallest), at least 8 tiles each;
// every chunk is npair units (one per pair of element groups).
static int pick_chunks(int nfb, int tiles, int slots, int npair) {
  const int cmax = std::max(1, std::min(std::min(16, tiles / 8), MSM_MAX_UNITS / npair));
  double best = 1e30;
  int best_c = 1;
  for (int c = 1; c <= cmax; ++c) {
    const double rounds = (double)(((long long)nfb * c * npair + slots - 1) / slots);
    const double cost = rounds / c;
    if (cost < best * 0.995) {
      best = cost;
      best_c = c;
    }
  }
  return best_c;
}

// Units of a batch of phantoms [off[q], off[q + 1]): one phantom is split as
// pick_chunks says (or as tuned); in a batch every unit stays below
// 1/item_div (default 1/8) of a queue slot's share of the work, so that the
// largest-first queue ends evenly.
template <int ET, bool SHALLOW>
static int msm_plan_et(const Problem& p, int P, const int64_t* off, const fs_tuning& tune, MsmPlan* out) {
  using SH = MsmShape<ET, SHALLOW>;
  MsmPlan& m = *out;
  m = MsmPlan{};
  m.ET = ET;
  m.FBK = SH::FBK;
  m.shallow = SHALLOW;
  m.threads = SH::THREADS;
  m.nfb = (p.NF + m.FBK - 1) / m.FBK;
  m.nfpad = m.nfb * m.FBK;
  m.groups = (p.E + 63) / 64;
  m.npair = m.groups <= 2 ? 1 : m.groups * (m.groups - 1) / 2;
  m.ept = 64 * m.groups;
  if ((long long)P * m.npair > MSM_MAX_UNITS) return FS_ERR_UNSUPPORTED;
  std::vector<int> pair_code(m.npair, 1 << 16);  // (0, 1) with at most two groups
  if (m.groups > 2)
    for (int a = 0; a < m.groups; ++a)
      for (int b = a + 1; b < m.groups; ++b) pair_code[group_pair_index(a, b, m.groups)] = a | (b << 16);
  const int slots = device_cus() * SH::WG_PER_CU;
  std::vector<int> tiles(P), chunks(P);
  long long total = 0;
  for (int q = 0; q < P; ++q) {
    tiles[q] = (int)((off[q + 1] - off[q] + MST - 1) / MST);
    total += std::max(tiles[q], 1);
  }
  if (P == 1) {
    chunks[0] = tune.msm_units > 0 ? std::min(std::min(tune.msm_units, std::max(tiles[0], 1)), MSM_MAX_UNITS / m.npair)
                                   : pick_chunks(m.nfb, std::max(tiles[0], 1), slots, m.npair);
  } else {
    const double div = tune.msm_item_div > 0 ? (double)tune.msm_item_div : 8.0;
    double tmax = std::max(8.0, (double)total * m.nfb * m.npair / (slots * div));
    for (;;) {
      int n = 0;
      for (int q = 0; q < P; ++q) {
        chunks[q] = std::max(1, (int)std::ceil(tiles[q] / tmax));
        n += chunks[q];
      }
      if (n * m.npair <= MSM_MAX_UNITS) break;
      tmax *= 1.25;
    }
  }
  std::vector<MsmUnit> us;
  m.phantoms.n = P;
  m.phantoms.groups = m.groups;
  for (int q = 0; q < P; ++q) {
    const int per = std::max(1, (tiles[q] + chunks[q] - 1) / chunks[q]) * MST;
    m.phantoms.ph[q] = make_int2((int)us.size(), chunks[q]);
    for (int pi = 0; pi < m.npair; ++pi)
      for (int c = 0; c < chunks[q]; ++c) {
        const long long b = std::min<long long>(off[q] + (long long)c * per, off[q + 1]);
        const long long e = std::min<long long>(b + per, off[q + 1]);
        us.push_back(MsmUnit{(int)b, (int)e, (int)us.size(), pair_code[pi]});
      }
  }
  std::stable_sort(us.begin(), us.end(),
                   [](const MsmUnit& a, const MsmUnit& b) { return a.s_end - a.s_begin > b.s_end - b.s_begin; });
  m.units.n = (int)us.size();
  m.units.nfb = m.nfb;
  for (size_t i = 0; i < us.size(); ++i) m.units.u[i] = us[i];
  m.n_items = m.units.n * m.nfb;
  m.grid = std::max(1, std::min(m.n_items, slots));
  const size_t ep = (size_t)ET * 16;
  m.kpart_bytes = (size_t)m.units.n * m.nfpad * ep * ep * sizeof(cf);
  m.scat_offset = (m.kpart_bytes + 255) / 256 * 256;
  m.rt_offset = m.scat_offset + ((size_t)off[P] * sizeof(float4) + 255) / 256 * 256;
  const size_t terms = (size_t)(off[P] + 1) * m.ept * p.NS;  // + the dead row
  const size_t table_bytes = terms * (sizeof(double) + 2 * sizeof(float) + sizeof(float2));
  m.table = table_possible(p) && tune.msm_table != 2 && (tune.msm_table == 1 || table_bytes <= MSM_TABLE_AUTO_MAX);
  m.la_offset = m.rt_offset + (m.table ? (terms * sizeof(double) + 255) / 256 * 256 : 0);
  m.qt_offset = m.la_offset + (m.table ? (terms * 2 * sizeof(float) + 255) / 256 * 256 : 0);
  m.queue_offset = m.qt_offset + (m.table ? (terms * sizeof(float2) + 255) / 256 * 256 : 0);
  m.ws_bytes = m.queue_offset + 256;
  m.lds = FragLds<ET, SH::FBK>::bytes;
  return FS_OK;
}

static int msm_plan(const Problem& p, int P, const int64_t* off, MsmPlan* m) {
  if (P < 1 || P > MSM_MAX_PHANTOMS) return FS_ERR_UNSUPPORTED;
  if (off[0] != 0 || off[P] != p.S) return FS_ERR_INVALID;
  for (int q = 0; q < P; ++q)
    if (off[q + 1] < off[q]) return FS_ERR_INVALID;
  const fs_tuning tune = msm_tuning(p, P);
  const bool shallow = shallow_required(p) || tune.msm_fbk == 2;
  if (p.E <= 64)
    return shallow ? msm_plan_et<4, true>(p, P, off, tune, m) : msm_plan_et<4, false>(p, P, off, tune, m);
  return shallow ? msm_plan_et<8, true>(p, P, off, tune, m) : msm_plan_et<8, false>(p, P, off, tune, m);
}

size_t mfma_workspace_batch(const Problem& p, int P, const int64_t* off) {
  if (p.E > MSM_MAX_ELEMENTS) return 256;
  MsmPlan m;
  if (msm_plan(p, P, off, &m) != FS_OK) return 0;
  return m.ws_bytes;
}

size_t mfma_workspace(const Problem& p) {
  const int64_t off[2] = {0, p.S};
  return mfma_workspace_batch(p, 1, off);
}

template <int ET, int ATT>
static int launch_apply_mfma(const Problem& p, int P, const MsmPlan& m, const cf* kpart, cf* spect, hipStream_t st) {
  auto kern = k_msm_apply_mfma<ET, ATT>;
  constexpr int lds = ATT * ET * 2 * 64 * 16;
  static bool lds_set[MAX_DEVICES] = {false};  // dynamic LDS above 64 KiB: set once per device
  const int dev = current_device();
  if (!lds_set[dev]) {
    if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
      return FS_ERR_RESOURCE;
    lds_set[dev] = true;
  }
  hipLaunchKernelGGL(kern, dim3(p.NF, P), dim3(256), lds, st, p, m.phantoms, m.nfpad, kpart, spect);
  return FS_OK;
}

template <int NSUB, int ET, int MODE, bool SHALLOW>
static int launch_msm(const Problem& p, const MsmPlan& m, const float4* scat, const double* rt, const float* la,
                      const float2* qt, int* queue, cf* kpart, hipStream_t st) {
  auto kern = k_msm<NSUB, ET, MODE, SHALLOW>;
  static bool lds_set[MAX_DEVICES] = {false};  // dynamic LDS above 64 KiB: set once per device
  const int dev = current_device();
  if (!lds_set[dev]) {
    if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)m.lds) != hipSuccess)
      return FS_ERR_RESOURCE;
    lds_set[dev] = true;
  }
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, queue, 0);
  const int slot = timing_start(st);
  hipLaunchKernelGGL(kern, dim3(m.grid), dim3(m.threads), m.lds, st, p, scat, rt, la, qt, m.units, queue, m.nfpad,
                     m.ept, kpart);
  timing_stop(slot, st);
  return FS_OK;
}

template <int NSUB, int ET>
static int launch_msm_fd(const Problem& p, const MsmPlan& m, const float4* scat, const double* rt, const float* la,
                         const float2* qt, int* queue, cf* kpart, hipStream_t st) {
  if constexpr (NSUB == 0) {  // run-time sub-element count: no table, one frequency per wave
    if (p.fulldir) return launch_msm<0, ET, MSM_FULLDIR, true>(p, m, scat, rt, la, qt, queue, kpart, st);
    if (msm_fast_mode(p)) return launch_msm<0, ET, MSM_FAST, true>(p, m, scat, rt, la, qt, queue, kpart, st);
    return launch_msm<0, ET, MSM_GENERIC, true>(p, m, scat, rt, la, qt, queue, kpart, st);
  } else {
  if (m.table) {
    if constexpr (NSUB > 2) return launch_msm<NSUB, ET, MSM_TABLE, true>(p, m, scat, rt, la, qt, queue, kpart, st);
    else
      return m.shallow ? launch_msm<NSUB, ET, MSM_TABLE, true>(p, m, scat, rt, la, qt, queue, kpart, st)
                       : launch_msm<NSUB, ET, MSM_TABLE, false>(p, m, scat, rt, la, qt, queue, kpart, st);
  }
  if constexpr (NSUB > 2) {
    if (p.fulldir) return launch_msm<NSUB, ET, MSM_FULLDIR, true>(p, m, scat, rt, la, qt, queue, kpart, st);
    if (msm_fast_mode(p)) return launch_msm<NSUB, ET, MSM_FAST, true>(p, m, scat, rt, la, qt, queue, kpart, st);
    return launch_msm<NSUB, ET, MSM_GENERIC, true>(p, m, scat, rt, la, qt, queue, kpart, st);
  } else {
    if (p.fulldir) return launch_msm<NSUB, ET, MSM_FULLDIR, true>(p, m, scat, rt, la, qt, queue, kpart, st);
    if (msm_fast_mode(p))
      return m.shallow ? launch_msm<NSUB, ET, MSM_FAST, true>(p, m, scat, rt, la, qt, queue, kpart, st)
                       : launch_msm<NSUB, ET, MSM_FAST, false>(p, m, scat, rt, la, qt, queue, kpart, st);
    return m.shallow ? launch_msm<NSUB, ET, MSM_GENERIC, true>(p, m, scat, rt, la, qt, queue, kpart, st)
                     : launch_msm<NSUB, ET, MSM_GENERIC, false>(p, m, scat, rt, la, qt, queue, kpart, st);
  }
  }  // NSUB > 0
}

template <int NSUB>
static int launch_msm_et(const Problem& p, const MsmPlan& m, const float4* scat, double* rt, float* la, float2* qt,
                         int* queue, cf* kpart, hipStream_t st) {
  if constexpr (NSUB == 0) {
    if (m.table) return FS_ERR_UNSUPPORTED;  // table_possible excludes it
  } else if (m.table && p.S > 0) {
    const long long n = (long long)(p.S + 1) * m.ept;
    hipLaunchKernelGGL(k_geo_table<NSUB>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, scat, m.ept, rt,
                       la, qt);
  }
  return m.ET == 4 ? launch_msm_fd<NSUB, 4>(p, m, scat, rt, la, qt, queue, kpart, st)
                   : launch_msm_fd<NSUB, 8>(p, m, scat, rt, la, qt, queue, kpart, st);
}

// Spectra of P phantoms sharing one probe, transmit set and frequency plan:
// spect[ph][t][k][e]. Only kernel launches (no host synchronisation, no memset),
// so that the call can be captured in a graph.
int mfma_spectrum_batch(const Problem& p, int P, const int64_t* off, cf* spect, char* ws, size_t ws_bytes,
                        hipStream_t st) {
  if (p.E > MSM_MAX_ELEMENTS) return FS_ERR_UNSUPPORTED;
  MsmPlan m;
  int rc = msm_plan(p, P, off, &m);
  if (rc != FS_OK) return rc;
  if (ws_bytes < m.ws_bytes) return FS_ERR_WORKSPACE;
  cf* kpart = (cf*)ws;
  float4* scat = (float4*)(ws + m.scat_offset);
  int* queue = (int*)(ws + m.queue_offset);
  double* rt = m.table ? (double*)(ws + m.rt_offset) : nullptr;
  float* la = m.table ? (float*)(ws + m.la_offset) : nullptr;
  float2* qt = m.table ? (float2*)(ws + m.qt_offset) : nullptr;
  if (p.S > 0) hipLaunchKernelGGL(k_scat_prep, dim3((p.S + 255) / 256), dim3(256), 0, st, p, scat);
  switch (p.NS) {
    case 1: rc = launch_msm_et<1>(p, m, scat, rt, la, qt, queue, kpart, st); break;
    case 2: rc = launch_msm_et<2>(p, m, scat, rt, la, qt, queue, kpart, st); break;
    case 3: rc = launch_msm_et<3>(p, m, scat, rt, la, qt, queue, kpart, st); break;
    case 4: rc = launch_msm_et<4>(p, m, scat, rt, la, qt, queue, kpart, st); break;
    default: rc = launch_msm_et<0>(p, m, scat, rt, la, qt, queue, kpart, st); break;
  }
  if (rc != FS_OK) return rc;
  hipLaunchKernelGGL(k_msm_fold, dim3(p.NF, P, m.npair), dim3(256), 0, st, m.ET * 16, m.nfpad, m.phantoms, kpart);
  const dim3 ag(p.NF, P, (p.E + 127) / 128);
  if (m.groups <= 2 && p.NT > 16) {
    // matrix cores: 5 transmit tiles per pass at 128 elements, 10 at 64 (80 KB of B fragments)
    rc = m.ET == 8 ? launch_apply_mfma<8, 5>(p, P, m, kpart, spect, st) : launch_apply_mfma<4, 10>(p, P, m, kpart, spect, st);
    if (rc != FS_OK) return rc;
  } else if (p.NT <= 2)
    hipLaunchKernelGGL(k_msm_apply<1>, ag, dim3(256), 0, st, p, m.ET * 16, m.phantoms, m.nfpad, kpart, spect);
  else if (p.NT <= 16)
    hipLaunchKernelGGL(k_msm_apply<8>, ag, dim3(256), 0, st, p, m.ET * 16, m.phantoms, m.nfpad, kpart, spect);
  else
    hipLaunchKernelGGL(k_msm_apply<40>, ag, dim3(256), 0, st, p, m.ET * 16, m.phantoms, m.nfpad, kpart, spect);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

// Geometry table of the plain kernels (k_rx): rows of p.E elements (no padding), amplitude without
// sqrt|rc|, for 1 to 4 sub-elements, a non-negative amplitude and centre-frequency directivity.
bool plain_geo_table_possible(const Problem& p) { return p.NS >= 1 && table_possible(p); }

void launch_plain_geo_table(const Problem& p, float4* scat, double* rt, float* la, float2* qt, hipStream_t st) {
  if (p.S > 0) hipLaunchKernelGGL(k_scat_prep_unit, dim3((p.S + 255) / 256), dim3(256), 0, st, p, scat);
  const long long n = (long long)(p.S + 1) * p.E;
  const dim3 grid((unsigned)((n + 255) / 256));
  switch (p.NS) {
    case 1: hipLaunchKernelGGL(k_geo_table<1>, grid, dim3(256), 0, st, p, scat, p.E, rt, la, qt); break;
    case 2: hipLaunchKernelGGL(k_geo_table<2>, grid, dim3(256), 0, st, p, scat, p.E, rt, la, qt); break;
    case 3: hipLaunchKernelGGL(k_geo_table<3>, grid, dim3(256), 0, st, p, scat, p.E, rt, la, qt); break;
    default: hipLaunchKernelGGL(k_geo_table<4>, grid, dim3(256), 0, st, p, scat, p.E, rt, la, qt); break;
  }
}

int mfma_spectrum(const Problem& p, cf* spect, char* ws, size_t ws_bytes, hipStream_t st) {
  const int64_t off[2] = {0, p.S};
  return mfma_spectrum_batch(p, 1, off, spect, ws, ws_bytes, st);
}

}  // namespace fs
