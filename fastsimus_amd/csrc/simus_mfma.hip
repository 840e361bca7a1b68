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
// product H^T diag(sign rc) H with H = sqrt|rc| G, evaluated with
// v_mfma_f32_16x16x32_bf16 in split precision: every operand is hi + lo
// bf16 and hi*hi + hi*lo + lo*hi is accumulated in FP32.
//
// Complex -> real: the 32-deep MFMA contraction index runs over
// (real part, imaginary part) of 16 scatterers; one fragment of element e
// holds [Re H(e, s0..s3), Im H(e, s0..s3)] for the lane group's 4 scatterers.
// With that fragment F(e) as both operands:
//   Re K = (Re H | -Im H)(e) . sign (Re H | Im H)(e')
//   Im K = (Im H |  Re H)(e) . sign (Re H | Im H)(e')
// i.e. the A operands are the fragment with its upper half negated or its
// halves swapped, and the B operand is the fragment with the sign of rc
// applied: no transposes, one LDS layout for both sides.
//
// Workgroup = 4 waves, one block of FBK frequencies and one chunk of
// scatterers. Per tile of 16 scatterers:
//   1. every thread evaluates G for 4 scatterers x EPT elements at the FBK
//      frequencies (FP64 r and phase seed, FP32 recurrence inside the block)
//      and writes split-bf16 fragments to LDS;
//   2. every wave runs its share of the upper-triangular 16x16 tiles of K_f
//      (2x2 block split of the tile grid) with accumulators in registers.
// Partial K per scatterer chunk goes to HBM; k_msm_apply sums the chunks in a
// fixed order and applies every transmit (E x E x n_tx per frequency, FP32).
#include <algorithm>

#include "fs_device.h"

namespace fs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int MST = 16;  // scatterers per tile (32-deep contraction: re + im)

template <int ET, int FBK>
struct FragLds {
  static constexpr int per_k = 2 * 2 * ET * 64;  // (T1|T2) x (hi|lo) x tiles x lanes
  static constexpr size_t bytes = (size_t)FBK * per_k * 16;
  __device__ static __forceinline__ int idx(int k, int t, int hl, int tile, int lane) {
    return (((k * 2 + t) * 2 + hl) * ET + tile) * 64 + lane;
  }
};

__device__ __forceinline__ void split_pair(float a, float b, unsigned& hi, unsigned& lo) {
  const bf16x2 h = __builtin_convertvector((floatx2){a, b}, bf16x2);
  const unsigned hb = __builtin_bit_cast(unsigned, h);
  const float ah = __uint_as_float(hb << 16), bh = __uint_as_float(hb & 0xffff0000u);
  const bf16x2 l = __builtin_convertvector((floatx2){a - ah, b - bh}, bf16x2);
  hi = hb;
  lo = __builtin_bit_cast(unsigned, l);
}

__device__ __forceinline__ floatx4 mma(u32x4 a, u32x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

// ---------------------------------------------------------------- tile jobs
// Upper-triangular 16x16 tile pairs (a <= b) of the E x E matrix. A frequency
// pair of the block is shared by WPP waves, each owning one list: 9 pairs
// each for 8 tiles (E <= 128, WPP = 4), 5 each for 4 tiles (E <= 64,
// WPP = 2), grouped so that each wave touches few distinct rows/cols.
template <int ET, int W>
struct JobList;
template <> struct JobList<8, 0> {
  static constexpr int n = 9;
  static constexpr int a[9] = {0, 0, 0, 0, 1, 1, 1, 2, 2};
  static constexpr int b[9] = {0, 1, 2, 3, 1, 2, 3, 2, 3};
};
template <> struct JobList<8, 1> {
  static constexpr int n = 9;
  static constexpr int a[9] = {4, 4, 4, 5, 5, 5, 6, 6, 7};
  static constexpr int b[9] = {5, 6, 7, 5, 6, 7, 6, 7, 7};
};
template <> struct JobList<8, 2> {
  static constexpr int n = 9;
  static constexpr int a[9] = {0, 0, 0, 0, 1, 1, 1, 1, 4};
  static constexpr int b[9] = {4, 5, 6, 7, 4, 5, 6, 7, 4};
};
template <> struct JobList<8, 3> {
  static constexpr int n = 9;
  static constexpr int a[9] = {2, 2, 2, 2, 3, 3, 3, 3, 3};
  static constexpr int b[9] = {4, 5, 6, 7, 3, 4, 5, 6, 7};
};
template <> struct JobList<4, 0> {
  static constexpr int n = 5;
  static constexpr int a[5] = {0, 0, 0, 0, 1};
  static constexpr int b[5] = {0, 1, 2, 3, 1};
};
template <> struct JobList<4, 1> {
  static constexpr int n = 5;
  static constexpr int a[5] = {1, 1, 2, 2, 3};
  static constexpr int b[5] = {2, 3, 2, 3, 3};
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
// ET waves; FBK = 4 frequencies per block, two per wave, WPP waves per pair.
template <int ET>
struct MsmShape {
  static constexpr int FBK = 4;
  static constexpr int NW = ET;                 // waves per workgroup
  static constexpr int THREADS = 64 * NW;
  static constexpr int WPP = NW * 2 / FBK;      // waves sharing one frequency pair
  static constexpr int WG_PER_CU = ET == 8 ? 1 : 2;
};

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

// Per-thread element data, constant over the whole launch.
template <int NSUB>
struct ElemData {
  double sx[NSUB], sz[NSUB];
  float ce, se;
  bool valid;
};

template <int NSUB, int ET>
__device__ __forceinline__ ElemData<NSUB> load_element(const Problem& p) {
  ElemData<NSUB> d;
  const int e = threadIdx.x % (ET * 16);
  d.valid = e < p.E;
  const int ee = d.valid ? e : 0;
#pragma unroll
  for (int j = 0; j < NSUB; ++j) {
    d.sx[j] = p.sub_x[ee * NSUB + j];
    d.sz[j] = p.sub_z[ee * NSUB + j];
  }
  d.ce = (float)p.ecos[ee];
  d.se = (float)p.esin[ee];
  return d;
}

// ---------------------------------------------------------------- G fragments
// Thread (scatterer quad g, element e): for each of the quad's 4 scatterers
// and each sub-element, the frequency-independent geometry once per block
// (FP64 distance and phase reduction), the seed g(f_b) and the step
// q = exp((-att + i 2 pi t) df), then the block's FBK frequencies by the FP32
// recurrence g <- g q. Sub-elements are summed, sqrt|rc| is folded into the
// amplitude, and every frequency is written as split-bf16 fragments
// (T1 plain, T2 with the sign of rc applied).
template <int NSUB, int FBK, bool FULLDIR, int ET>
__device__ __forceinline__ void gen_fragments(const Problem& p, const ElemData<NSUB>& ed, bool poly_ok, int s0,
                                              int s_end, double fb_hz, u32x4* lds) {
  using L = FragLds<ET, FBK>;
  const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / (ET * 16)));  // wave-uniform
  const int e = threadIdx.x % (ET * 16);
  cf gk[4][NSUB], q[4][NSUB];
  float sth[4][NSUB];
  unsigned m01 = 0u, m23 = 0u;
  const float fbf = (float)fb_hz, dff = (float)p.df;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int s = s0 + 4 * g + i;
    const bool valid = s < s_end;
    const int sc = valid ? s : s_end - 1;  // in-bounds load, masked below
    const float x = p.x[sc], z = p.z[sc], r = p.rc[sc];
    const bool live = valid && ed.valid && r != 0.f && !out_of_field(p, x, z);
    const float w = live ? __builtin_amdgcn_sqrtf(fabsf(r)) * p.inv_nsub : 0.f;
    const unsigned sb = (live && r < 0.f) ? 1u : 0u;
    if (i < 2)
      m01 |= sb << (15 + 16 * i);
    else
      m23 |= sb << (15 + 16 * (i - 2));
    const double xd = (double)x, zd = (double)z;
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      const double dx = xd - ed.sx[j];
      const double dz = zd - ed.sz[j];
      const double r2 = fma(dx, dx, dz * dz);
      const bool zero = !(r2 > 0.0);
      double y = __builtin_amdgcn_rsq(zero ? 1.0 : r2);
      y = y * fma(-0.5 * r2, y * y, 1.5);  // Newton step on 1/sqrt
      const double rcl = fmax(zero ? 0.0 : r2 * y, p.min_dist);
      const float rinv = (float)y;
      const float sa = zero ? 1.0f : (float)dx * rinv;
      const float ca = zero ? 0.0f : fabsf((float)dz) * rinv;
      const float st = sa * ed.ce - ca * ed.se;
      const float ct = ca * ed.ce + sa * ed.se;
      float obl = (p.baffle == FS_BAFFLE_SOFT)    ? ct
                  : (p.baffle == FS_BAFFLE_RIGID) ? 1.0f
                                                  : ct * __builtin_amdgcn_rcpf(ct + p.baffle_b);
      if (!(ct > 0.0f)) obl = 1e-16f;
      const float rcf = (float)rcl;
      float amp = obl * __builtin_amdgcn_rsqf(rcf) * w;
      if constexpr (!FULLDIR) amp *= poly_ok ? sinc_poly_pi(p.kappa * st) : sinc_fast(p.kappa * st);
      sth[i][j] = st;
      const double t = rcl * p.inv_c;
      const float att = p.att_coef * rcf * kLog2e;
      const double nu0 = fb_hz * t, nud = p.df * t;
      const float ph0 = (float)(nu0 - floor(nu0)), phd = (float)(nud - floor(nud));
      const float m0 = amp * __builtin_amdgcn_exp2f(-att * fbf);
      const float md = __builtin_amdgcn_exp2f(-att * dff);
      gk[i][j] = {m0 * __builtin_amdgcn_cosf(ph0), m0 * __builtin_amdgcn_sinf(ph0)};
      q[i][j] = {md * __builtin_amdgcn_cosf(phd), md * __builtin_amdgcn_sinf(phd)};
    }
  }
  const int tile = e >> 4;
  const int lane = (e & 15) + 16 * g;
  const u32x4 msk = {m01, m23, m01, m23};
#pragma unroll
  for (int k = 0; k < FBK; ++k) {
    float re[4], im[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      re[i] = im[i] = 0.f;
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        if constexpr (FULLDIR) {
          const float d = sinc_fast(p.kappa_per_hz * (fbf + (float)k * dff) * sth[i][j]);
          re[i] = fmaf(d, gk[i][j].x, re[i]);
          im[i] = fmaf(d, gk[i][j].y, im[i]);
        } else {
          re[i] += gk[i][j].x;
          im[i] += gk[i][j].y;
        }
        if (k + 1 < FBK) gk[i][j] = cmul(gk[i][j], q[i][j]);
      }
    }
    unsigned h0, h1, h2, h3, l0, l1, l2, l3;
    split_pair(re[0], re[1], h0, l0);
    split_pair(re[2], re[3], h1, l1);
    split_pair(im[0], im[1], h2, l2);
    split_pair(im[2], im[3], h3, l3);
    const u32x4 hi = {h0, h1, h2, h3}, lo = {l0, l1, l2, l3};
    lds[L::idx(k, 0, 0, tile, lane)] = hi;
    lds[L::idx(k, 0, 1, tile, lane)] = lo;
    lds[L::idx(k, 1, 0, tile, lane)] = hi ^ msk;
    lds[L::idx(k, 1, 1, tile, lane)] = lo ^ msk;
  }
}

// ---------------------------------------------------------------- MFMA tiles of one wave
// Accumulators of one wave: its job list at the two frequencies k0, k0 + 1 of the block.
template <int ET, int FBK, int W>
struct WaveTiles {
  using J = JobList<ET, W>;
  static constexpr int NJ = J::n;
  floatx4 acc[2][NJ][2];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < NJ; ++j) acc[k][j][0] = acc[k][j][1] = (floatx4){0.f, 0.f, 0.f, 0.f};
  }

  // One 16-scatterer tile: for each of the wave's frequencies, load the sign-applied
  // column fragments this wave needs, then stream its rows.
  __device__ __forceinline__ void step(const u32x4* lds, int k0) {
    using L = FragLds<ET, FBK>;
    const int lane = threadIdx.x & 63;
    const u32x4 nim = {0u, 0u, 0x80008000u, 0x80008000u};
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const int k = k0 + kk;
      u32x4 bh[ET], bl[ET];
#pragma unroll
      for (int c = 0; c < ET; ++c) {
        if (uses_col<ET, W>(c)) {
          bh[c] = lds[L::idx(k, 1, 0, c, lane)];
          bl[c] = lds[L::idx(k, 1, 1, c, lane)];
        }
      }
#pragma unroll
      for (int r = 0; r < ET; ++r) {
        if (!uses_row<ET, W>(r)) continue;
        const u32x4 h = lds[L::idx(k, 0, 0, r, lane)];
        const u32x4 l = lds[L::idx(k, 0, 1, r, lane)];
        const u32x4 arh = h ^ nim, arl = l ^ nim;
        const u32x4 aih = {h.z, h.w, h.x, h.y}, ail = {l.z, l.w, l.x, l.y};
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          if (J::a[j] != r) continue;
          const int c = J::b[j];
          floatx4 kr = acc[kk][j][0], ki = acc[kk][j][1];
          kr = mma(arh, bh[c], kr);
          ki = mma(aih, bh[c], ki);
          kr = mma(arh, bl[c], kr);
          ki = mma(aih, bl[c], ki);
          kr = mma(arl, bh[c], kr);
          ki = mma(ail, bh[c], ki);
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
    for (int kk = 0; kk < 2; ++kk) {
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

template <int NSUB, int ET, bool FULLDIR, int W>
__device__ __forceinline__ void msm_wave(const Problem& p, int sb, int se, int fb, int k0, u32x4* lds, cf* kpart) {
  constexpr int FBK = MsmShape<ET>::FBK;
  const ElemData<NSUB> ed = load_element<NSUB, ET>(p);
  const bool poly_ok = !FULLDIR && p.kappa <= kPi;  // centre-frequency sinc argument |kappa sin| <= pi
  const double fb_hz = p.f0 + (double)(fb * FBK) * p.df;
  WaveTiles<ET, FBK, W> t;
  t.zero();
  for (int s0 = sb; s0 < se; s0 += MST) {
    gen_fragments<NSUB, FBK, FULLDIR, ET>(p, ed, poly_ok, s0, se, fb_hz, lds);
    __syncthreads();
    t.step(lds, k0);
    __syncthreads();
  }
  t.store(kpart, ET * 16, fb * FBK + k0);
}

// Grid (frequency block, scatterer chunk); ET waves per workgroup. Wave w owns
// frequencies 2 (w / WPP) and 2 (w / WPP) + 1 of the block and job list w % WPP.
template <int NSUB, int ET, bool FULLDIR>
__global__ __launch_bounds__(MsmShape<ET>::THREADS, MsmShape<ET>::WG_PER_CU) void k_msm(Problem p, int s_per_chunk,
                                                                                         int nfpad, cf* __restrict__ kpart) {
  using SH = MsmShape<ET>;
  extern __shared__ u32x4 lds_frag[];
  const int fb = blockIdx.x;
  const int chunk = blockIdx.y;
  const int sb = chunk * s_per_chunk;
  const int se = min(sb + s_per_chunk, p.S);
  cf* kp = kpart + (size_t)chunk * nfpad * (ET * 16) * (ET * 16);
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int k0 = 2 * (w / SH::WPP);
  if constexpr (SH::WPP == 4) {
    switch (w % 4) {
      case 0: msm_wave<NSUB, ET, FULLDIR, 0>(p, sb, se, fb, k0, lds_frag, kp); break;
      case 1: msm_wave<NSUB, ET, FULLDIR, 1>(p, sb, se, fb, k0, lds_frag, kp); break;
      case 2: msm_wave<NSUB, ET, FULLDIR, 2>(p, sb, se, fb, k0, lds_frag, kp); break;
      default: msm_wave<NSUB, ET, FULLDIR, 3>(p, sb, se, fb, k0, lds_frag, kp); break;
    }
  } else {
    if (w % 2 == 0)
      msm_wave<NSUB, ET, FULLDIR, 0>(p, sb, se, fb, k0, lds_frag, kp);
    else
      msm_wave<NSUB, ET, FULLDIR, 1>(p, sb, se, fb, k0, lds_frag, kp);
  }
}

// ---------------------------------------------------------------- apply all transmits
// out[t][k][e] = corr probe_k PP_k sum_e' K_k[e][e'] A_t(k, e'),  K_k = sum_c Kpart[c][k] (upper tiles, symmetric)
constexpr int AP_CH = 32;  // e' per LDS chunk

__global__ __launch_bounds__(256) void k_msm_apply(Problem p, int E_pad, int nchunks, int nfpad,
                                                   const cf* __restrict__ kpart, cf* __restrict__ out) {
  __shared__ cf Ks[128][AP_CH + 1];
  __shared__ cf As[AP_CH][2 * 40 + 1];
  const int k = blockIdx.x;
  const int tid = threadIdx.x;
  const int e = tid & 127;
  const int tg = tid >> 7;  // transmit parity
  const double f = p.f0 + (double)k * p.df;
  constexpr int MAXT = 40;  // transmits per pass per thread parity (80 per pass)
  for (int t0 = 0; t0 < p.NT; t0 += 2 * MAXT) {
    cf acc[MAXT];
#pragma unroll
    for (int i = 0; i < MAXT; ++i) acc[i] = {0.f, 0.f};
    for (int c0 = 0; c0 < p.E; c0 += AP_CH) {
      __syncthreads();
      for (int q = tid; q < 128 * AP_CH; q += 256) {
        const int r = q / AP_CH, cc = q % AP_CH;
        const int col = c0 + cc;
        cf v = {0.f, 0.f};
        if (r < p.E && col < p.E) {
          const int ra = ((r >> 4) <= (col >> 4)) ? r : col;
          const int cb = ((r >> 4) <= (col >> 4)) ? col : r;
          float sx = 0.f, sy = 0.f;
          for (int c = 0; c < nchunks; ++c) {
            const cf w = kpart[(((size_t)c * nfpad + k) * E_pad + ra) * E_pad + cb];
            sx += w.x;
            sy += w.y;
          }
          v = {sx, sy};
        }
        Ks[r][cc] = v;
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
          const cf kv = Ks[e][cc];
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

// ---------------------------------------------------------------- host side
struct MsmPlan {
  int ET, FBK, nfb, nfpad, chunks, s_per_chunk, threads;
  size_t lds, kpart_bytes;
};

static int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return 256;
  return cus;
}

// Scatterer chunks per frequency block: the count whose last round of
// workgroups is fullest (rounds / chunks smallest), at least 8 tiles per chunk.
static int pick_chunks(int nfb, int tiles, int slots) {
  const int cmax = std::max(1, std::min(16, tiles / 8));
  double best = 1e30;
  int best_c = 1;
  for (int c = 1; c <= cmax; ++c) {
    const double rounds = (double)(((long long)nfb * c + slots - 1) / slots);
    const double cost = rounds / c;
    if (cost < best * 0.995) {
      best = cost;
      best_c = c;
    }
  }
  return best_c;
}

template <int ET>
static MsmPlan msm_plan_et(const Problem& p) {
  using SH = MsmShape<ET>;
  MsmPlan m{};
  m.ET = ET;
  m.FBK = SH::FBK;
  m.threads = SH::THREADS;
  m.nfb = (p.NF + m.FBK - 1) / m.FBK;
  m.nfpad = m.nfb * m.FBK;
  const int tiles = std::max(1, (p.S + MST - 1) / MST);
  const int chunks = pick_chunks(m.nfb, tiles, device_cus() * SH::WG_PER_CU);
  m.s_per_chunk = ((tiles + chunks - 1) / chunks) * MST;
  m.chunks = std::max(1, (p.S + m.s_per_chunk - 1) / m.s_per_chunk);
  const size_t ep = (size_t)ET * 16;
  m.kpart_bytes = (size_t)m.chunks * m.nfpad * ep * ep * sizeof(cf);
  m.lds = FragLds<ET, SH::FBK>::bytes;
  return m;
}

static MsmPlan msm_plan(const Problem& p) { return p.E <= 64 ? msm_plan_et<4>(p) : msm_plan_et<8>(p); }

size_t mfma_workspace(const Problem& p) {
  if (p.E > 128 || p.NS > 4) return 256;
  return (msm_plan(p).kpart_bytes + 255) / 256 * 256;
}

template <int NSUB, int ET, bool FD>
static int launch_msm(const Problem& p, const MsmPlan& m, cf* kpart, hipStream_t st) {
  auto kern = k_msm<NSUB, ET, FD>;
  if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)m.lds) != hipSuccess)
    return FS_ERR_RESOURCE;
  const int slot = timing_start(st);
  hipLaunchKernelGGL(kern, dim3(m.nfb, m.chunks), dim3(m.threads), m.lds, st, p, m.s_per_chunk, m.nfpad, kpart);
  timing_stop(slot, st);
  return FS_OK;
}

template <int NSUB, int ET>
static int launch_msm_fd(const Problem& p, const MsmPlan& m, cf* kpart, hipStream_t st) {
  return p.fulldir ? launch_msm<NSUB, ET, true>(p, m, kpart, st) : launch_msm<NSUB, ET, false>(p, m, kpart, st);
}

template <int NSUB>
static int launch_msm_et(const Problem& p, const MsmPlan& m, cf* kpart, hipStream_t st) {
  return m.ET == 4 ? launch_msm_fd<NSUB, 4>(p, m, kpart, st) : launch_msm_fd<NSUB, 8>(p, m, kpart, st);
}

int mfma_spectrum(const Problem& p, cf* spect, char* ws, size_t ws_bytes, hipStream_t st) {
  if (p.E > 128) return FS_ERR_UNSUPPORTED;
  const MsmPlan m = msm_plan(p);
  if (ws_bytes < m.kpart_bytes) return FS_ERR_WORKSPACE;
  cf* kpart = (cf*)ws;
  int rc;
  switch (p.NS) {
    case 1: rc = launch_msm_et<1>(p, m, kpart, st); break;
    case 2: rc = launch_msm_et<2>(p, m, kpart, st); break;
    case 3: rc = launch_msm_et<3>(p, m, kpart, st); break;
    case 4: rc = launch_msm_et<4>(p, m, kpart, st); break;
    default: return FS_ERR_UNSUPPORTED;
  }
  if (rc != FS_OK) return rc;
  hipLaunchKernelGGL(k_msm_apply, dim3(p.NF), dim3(256), 0, st, p, m.ET * 16, m.chunks, m.nfpad, kpart, spect);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

}  // namespace fs
