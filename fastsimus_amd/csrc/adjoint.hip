// Adjoint of the SIMUS spectrum: gradients of a real loss L with respect to the
// reflectivities rc_s, the scatterer positions (x_s, z_s) and the speed of sound c.
//
// Forward (per transmit t, selected frequency f, element e; simus.py:263-333,
// _simus_strategies.py:23-77):
//   X_t(f,e) = c_f sum_s rc_s P_t(f,s) G(f,e,s),   c_f = corr probe(f)
//   P_t(f,s) = PP(f) sum_e G(f,e,s) A_t(f,e),       0 where s is out of field
// With the cotangent g_t(f,e) = dL/dRe X + i dL/dIm X (one value per output):
//   dL/drc_s = sum_{t,f} Re(c_f P_t(f,s) B_t(f,s)),  B_t(f,s) = sum_e G(f,e,s) conj(g_t(f,e))
//   dL/dth_s = sum_{t,f,e} Re(G_th(f,e,s) W_t(f,e,s)) rc_s,
//   W_t(f,e,s) = c_f PP(f) A_t(f,e) B_t(f,s) + c_f P_t(f,s) conj(g_t(f,e))
// for th in {x_s, z_s, c} (c summed over s). The Green's functions and their
// derivatives are recomputed here, never stored:
//   g_j(f)      = amp E_j(f),  E_j(f) = exp(f r_c (-alpha' + i 2 pi / c))
//   dg_j/dth    = E_j(f) (a_th + b_th f)
// with real a_th (obliquity, directivity and 1/sqrt(r_c) terms) and complex
// b_th (propagation terms); see coeffs_fast() below.
//
// Two implementations of the same algebra:
// vector (any element count): k_adj_tx (lane = scatterer): P_t, B_t for one
// transmit, frequency blocks of 32, stored to HBM; the rc gradient is reduced
// on the fly. k_adj_pos (lane = scatterer): regenerates E_j and the derivative
// coefficients and accumulates the x, z, c gradients against W_t.
// matrix cores (n_elem <= 128, n_sub <= 2): k_adjm, the sums over elements as
// split-BF16 MFMA products for 16 transmits at a time (see "matrix-core form").
// Per-(transmit, block, scatterer) partial sums are reduced in fixed order.
#include <algorithm>

#include "fs_device.h"

namespace fs {

constexpr int AFB = 32;

struct Dcoef {
  float ax, az, ac;  // real parts
  cf bx, bz, bc;     // per-Hz parts
  float fx, fz, fc;  // per-frequency directivity only: factors of sum_f f D'(y_f) E W (y_f = kappa_per_hz f sin(theta))
};

// sin(y) / y and its derivative, with the hardware sine / cosine in revolutions
__device__ __forceinline__ void sinc_slope(float y, float& D, float& Dp) {
  if (fabsf(y) < 1e-3f) {
    D = 1.0f - y * y * (1.0f / 6.0f);
    Dp = -y * (1.0f / 3.0f);
  } else {
    const float rev = y * (0.5f / kPi);
    const float iy = __builtin_amdgcn_rcpf(y);
    D = __builtin_amdgcn_sinf(rev) * iy;
    Dp = (__builtin_amdgcn_cosf(rev) - D) * iy;
  }
}

// amp, propagation and their derivatives for one (s, e, j) term, with the gfx950 hardware
// transcendentals (1/r from the FP32 estimate plus one FP64 Newton step, v_sin / v_cos in
// revolutions, reciprocals) and the propagation constants of GeoF (attenuation pre-multiplied
// by log2 e). With FULLDIR the directivity is a factor D(y_f) of every frequency, not part of
// amp: amp and the a / b coefficients leave it out, and f_th = amp kappa_per_hz d(sin theta)/d th
// (d(sin theta)/dc = 0, but y_f scales as 1/c: f_c = -amp kappa_per_hz sin(theta) / c) multiply
// sum_f f D'(y_f) E W.
template <bool FULLDIR = false>
__device__ __forceinline__ void coeffs_fast(const Problem& p, double xs, double zs, int e, int j, float fc_hz,
                                            GeoF& geo, Dcoef& dc) {
  const int idx = e * p.NS + j;
  const double dx = xs - p.sub_x[idx];
  const double dz = zs - p.sub_z[idx];
  double r2 = fma(dx, dx, dz * dz);
  const bool zero = !(r2 > 0.0);
  r2 = fmax(r2, 1e-30);
  const float y32 = __builtin_amdgcn_rsqf((float)r2);
  double y = (double)y32;
  y = y * fma(-0.5 * r2, y * y, 1.5);
  const double r = zero ? 0.0 : r2 * y;
  const bool clipped = r < p.min_dist;
  const double rcl = clipped ? p.min_dist : r;
  const float rinv = (float)y;
  const float sa = zero ? 1.0f : (float)dx * rinv;
  const float ca = zero ? 0.0f : fabsf((float)dz) * rinv;
  const float ce = (float)p.ecos[e], se = (float)p.esin[e];
  const float st = sa * ce - ca * se;
  const float ct = ca * ce + sa * se;
  const bool horizon = !(ct > 0.0f);
  float obl, dobl;
  if (p.baffle == FS_BAFFLE_SOFT) {
    obl = ct;
    dobl = -st;
  } else if (p.baffle == FS_BAFFLE_RIGID) {
    obl = 1.0f;
    dobl = 0.0f;
  } else {
    const float iden = __builtin_amdgcn_rcpf(ct + p.baffle_b);
    obl = ct * iden;
    dobl = -st * p.baffle_b * iden * iden;
  }
  if (horizon) {
    obl = 1e-16f;
    dobl = 0.0f;
  }
  const float yk = p.kappa * st;
  float D = 1.0f, dDdy = 0.0f;
  if constexpr (!FULLDIR) sinc_slope(yk, D, dDdy);
  const float rcf = (float)rcl;
  const float rs = __builtin_amdgcn_rsqf(rcf) * p.inv_nsub;
  const float amp = obl * D * rs;
  const float r2inv = rinv * rinv;
  const float thx = fabsf((float)dz) * r2inv;
  const float thz = -(float)dx * (dz >= 0.0 ? 1.0f : -1.0f) * r2inv;
  const float amp_th = (dobl * D + obl * dDdy * p.kappa * ct) * rs;
  const float rhox = clipped ? 0.0f : sa;
  const float rhoz = clipped ? 0.0f : (float)dz * rinv;
  const float rhoc = clipped ? 0.5f * __builtin_amdgcn_rcpf(fc_hz) : 0.0f;
  const float half = -0.5f * __builtin_amdgcn_rcpf(rcf);
  const float inv_c = (float)p.inv_c;
  const cf kp = {-p.att_coef, 2.0f * kPi * inv_c};
  dc.ax = amp_th * thx + amp * half * rhox;
  dc.az = amp_th * thz + amp * half * rhoz;
  dc.bx = {amp * kp.x * rhox, amp * kp.y * rhox};
  dc.bz = {amp * kp.x * rhoz, amp * kp.y * rhoz};
  const float dD_dc = dDdy * st * (-p.kappa * inv_c);
  dc.ac = obl * dD_dc * rs + amp * half * rhoc;
  dc.bc = {amp * kp.x * rhoc, amp * (kp.y * rhoc - 2.0f * kPi * rcf * inv_c * inv_c)};
  if constexpr (FULLDIR) {
    const float ak = amp * p.kappa_per_hz;
    dc.fx = ak * ct * thx;
    dc.fz = ak * ct * thz;
    dc.fc = -ak * st * inv_c;
  } else {
    dc.fx = dc.fz = dc.fc = 0.0f;
  }
  geo.t = rcl * p.inv_c;
  geo.amp = amp;
  geo.att = p.att_coef * rcf * kLog2e;
  geo.sint = st;
}

// TA[t][e][k] = c_k PP_k A_t(k, e),  TG[t][e][k] = c_k conj(g_t(k, e)),  c_k = corr probe_k (0 beyond NF)
__global__ void k_adj_tables(Problem p, int nfp, const cf* __restrict__ gspect, cf* __restrict__ TA,
                             cf* __restrict__ TG, cf* __restrict__ Gc) {
  const long total = (long)p.NT * p.E * nfp;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i % nfp);
    const long te = i / nfp;
    const int e = (int)(te % p.E);
    const int t = (int)(te / p.E);
    cf ta = {0.f, 0.f}, tg = {0.f, 0.f}, gc = {0.f, 0.f};
    if (k < p.NF) {
      const double f = p.f0 + (double)k * p.df;
      const double nu = f * p.delay[te];
      const float ph = (float)(nu - floor(nu));
      float s, c;
      sincospif(2.0f * ph, &s, &c);
      const float w = (float)p.apod[te];
      const double q = p.probe[k];
      const float ck = (float)(p.corr * q);
      const cf pp = {(float)(p.pulse_re[k] * q), (float)(p.pulse_im[k] * q)};
      const cf a = {w * c, w * s};
      const cf pa = cmul(pp, a);
      ta = {ck * pa.x, ck * pa.y};
      const cf g = gspect[((size_t)t * p.NF + k) * p.E + e];
      gc = {g.x, -g.y};
      tg = {ck * g.x, -ck * g.y};
    }
    TA[i] = ta;
    TG[i] = tg;
    Gc[i] = gc;
  }
}

// Sub-elements in groups of NSUB: with GROUPED the problem has p.NS of them and those past p.NS get zero
// amplitude and coefficients; without it p.NS == NSUB.
template <int NSUB, bool GROUPED>
__device__ __forceinline__ int adj_groups(const Problem& p) {
  return GROUPED ? (p.NS + NSUB - 1) / NSUB : 1;
}

// P_t (without rc) and B_t for one transmit; drc_part[fb][s] = sum_k Re(c_k P B)
template <int NSUB, bool FULLDIR, bool GROUPED = false>
__global__ __launch_bounds__(256) void k_adj_tx(Problem p, int tx, int nfp, const cf* __restrict__ A,
                                                const cf* __restrict__ Gc, const cf* __restrict__ ppc,
                                                cf* __restrict__ Pb, cf* __restrict__ Bb, float* __restrict__ drc_part) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int fb = blockIdx.y;
  const int k0 = fb * AFB;
  const bool valid = s < p.S;
  const float xs = valid ? p.x[s] : 0.f;
  const float zs = valid ? p.z[s] : 1e-2f;
  const bool out = !valid || out_of_field(p, xs, zs);
  cf P[AFB], B[AFB];
#pragma unroll
  for (int k = 0; k < AFB; ++k) P[k] = B[k] = {0.f, 0.f};
  const double fstart = p.f0 + (double)k0 * p.df;
  const cf* At = A + (size_t)tx * p.E * nfp + k0;
  const cf* Gt = Gc + (size_t)tx * p.E * nfp + k0;
  const int ngroups = adj_groups<NSUB, GROUPED>(p);
  for (int e = 0; e < p.E; ++e)
  for (int jg = 0; jg < ngroups; ++jg) {
    cf g[NSUB], q[NSUB];
    float st[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      const int jj = jg * NSUB + j;
      const bool ok = !GROUPED || jj < p.NS;
      const GeoF geo = geometry_fast(p, xs, zs, e, ok ? jj : 0);
      g[j] = green_fast(geo, fstart, ok ? geo.amp : 0.f);
      q[j] = green_step_fast(geo, p.df);
      st[j] = geo.sint;
    }
    const cf* Ae = At + (size_t)e * nfp;
    const cf* Ge = Gt + (size_t)e * nfp;
#pragma unroll
    for (int k = 0; k < AFB; ++k) {
      cf G = {0.f, 0.f};
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        if constexpr (FULLDIR) {
          const float d = sinc_fast(p.kappa_per_hz * (float)(fstart + (double)k * p.df) * st[j]);
          G.x = fmaf(g[j].x, d, G.x);
          G.y = fmaf(g[j].y, d, G.y);
        } else {
          G.x += g[j].x;
          G.y += g[j].y;
        }
        g[j] = cmul(g[j], q[j]);
      }
      cmac(P[k], G, Ae[k]);
      cmac(B[k], G, Ge[k]);
    }
  }
  if (!valid) return;
  float drc = 0.f;
  cf* Ps = Pb + ((size_t)fb * p.S + s) * AFB;
  cf* Bs = Bb + ((size_t)fb * p.S + s) * AFB;
#pragma unroll
  for (int k = 0; k < AFB; ++k) {
    // ppc[k] = (PP_k, c_k) packed: PP in .x/.y of entry 2k, c_k in entry 2k+1 .x
    const cf pp = ppc[2 * (k0 + k)];
    const float ck = ppc[2 * (k0 + k) + 1].x;
    const cf v = out ? cf{0.f, 0.f} : cmul(P[k], pp);
    Ps[k] = v;
    Bs[k] = B[k];
    const cf pb = cmul(v, B[k]);
    drc = fmaf(ck, pb.x, drc);
  }
  drc_part[(size_t)fb * p.S + s] = drc;
}

template <int NSUB, bool FULLDIR, bool GROUPED = false>
__global__ __launch_bounds__(256) void k_adj_pos(Problem p, int tx, int nfp, const cf* __restrict__ TA,
                                                 const cf* __restrict__ TG, const cf* __restrict__ Pb,
                                                 const cf* __restrict__ Bb, float* __restrict__ dxp,
                                                 float* __restrict__ dzp, float* __restrict__ dcp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int fb = blockIdx.y;
  const int k0 = fb * AFB;
  const bool valid = s < p.S;
  const float xs = valid ? p.x[s] : 0.f;
  const float zs = valid ? p.z[s] : 1e-2f;
  const float rcs = valid ? p.rc[s] : 0.f;
  const bool live = valid && !out_of_field(p, xs, zs) && rcs != 0.f;
  cf P[AFB], B[AFB];
  const cf* Ps = Pb + ((size_t)fb * p.S + (valid ? s : 0)) * AFB;
  const cf* Bs = Bb + ((size_t)fb * p.S + (valid ? s : 0)) * AFB;
#pragma unroll
  for (int k = 0; k < AFB; ++k) {
    P[k] = Ps[k];
    B[k] = Bs[k];
  }
  float gx = 0.f, gz = 0.f, gc = 0.f;
  const double fstart = p.f0 + (double)k0 * p.df;
  const float fc_hz = (float)(p.min_dist > 0 ? 0.5 / (p.min_dist * p.inv_c) : 1.0);
  const cf* TAt = TA + (size_t)tx * p.E * nfp + k0;
  const cf* TGt = TG + (size_t)tx * p.E * nfp + k0;
  const int ngroups = adj_groups<NSUB, GROUPED>(p);
  for (int e = 0; e < p.E; ++e)
  for (int jg = 0; jg < ngroups; ++jg) {
    cf E[NSUB], q[NSUB];
    float st[NSUB];
    Dcoef dc[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      const int jj = jg * NSUB + j;
      const bool ok = !GROUPED || jj < p.NS;
      GeoF geo;
      coeffs_fast<FULLDIR>(p, xs, zs, e, ok ? jj : 0, fc_hz, geo, dc[j]);
      E[j] = green_fast(geo, fstart, ok ? 1.0f : 0.0f);  // E_j carries the propagation factor only
      q[j] = green_step_fast(geo, p.df);
      st[j] = geo.sint;
    }
    const cf* TAe = TAt + (size_t)e * nfp;
    const cf* TGe = TGt + (size_t)e * nfp;
    // dg/dth = E (a + b f) is linear in f: accumulate S0 = sum_k E W and S1 = sum_k f E W over the
    // block, then Re(dg W) summed over k = a Re(S0) + Re(b S1). With the directivity at every
    // frequency, dg/dth = E (D(y_f) (a + b f) + f_th f D'(y_f)): S0 and S1 carry the factor D(y_f)
    // and S2 = sum_k f D'(y_f) E W adds f_th Re(S2).
    cf S0[NSUB], S1[NSUB], S2[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) S0[j] = S1[j] = S2[j] = {0.f, 0.f};
#pragma unroll
    for (int k = 0; k < AFB; ++k) {
      const float f = (float)(fstart + (double)k * p.df);
      cf W = {0.f, 0.f};
      cmac(W, TAe[k], B[k]);
      cmac(W, P[k], TGe[k]);
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        cf EW = cmul(E[j], W);
        if constexpr (FULLDIR) {
          float D, Dp;
          sinc_slope(p.kappa_per_hz * f * st[j], D, Dp);
          const float fdp = f * Dp;
          S2[j].x = fmaf(fdp, EW.x, S2[j].x);
          S2[j].y = fmaf(fdp, EW.y, S2[j].y);
          EW = {EW.x * D, EW.y * D};
        }
        S0[j].x += EW.x;
        S0[j].y += EW.y;
        S1[j].x = fmaf(f, EW.x, S1[j].x);
        S1[j].y = fmaf(f, EW.y, S1[j].y);
        E[j] = cmul(E[j], q[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      gx += dc[j].ax * S0[j].x + (dc[j].bx.x * S1[j].x - dc[j].bx.y * S1[j].y);
      gz += dc[j].az * S0[j].x + (dc[j].bz.x * S1[j].x - dc[j].bz.y * S1[j].y);
      gc += dc[j].ac * S0[j].x + (dc[j].bc.x * S1[j].x - dc[j].bc.y * S1[j].y);
      if constexpr (FULLDIR) {
        gx += dc[j].fx * S2[j].x;
        gz += dc[j].fz * S2[j].x;
        gc += dc[j].fc * S2[j].x;
      }
    }
  }
  if (!valid) return;
  const size_t o = (size_t)fb * p.S + s;
  dxp[o] = live ? gx * rcs : 0.f;
  dzp[o] = live ? gz * rcs : 0.f;
  dcp[o] = live ? gc * rcs : 0.f;
}

// out[s] (+)= sum_b part[b][s]  (fixed order); accumulate=0 overwrites
__global__ void k_sum_parts(const float* __restrict__ part, int nb, int S, int accumulate, float* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  float a = 0.f;
  for (int b = 0; b < nb; ++b) a += part[(size_t)b * S + s];
  out[s] = accumulate ? out[s] + a : a;
}

// d_c = sum over all (t, b, s) of the dc partials, in double, in two fixed-order stages:
// workgroup g sums the segment [g L, (g + 1) L) into seg[g], then one workgroup sums seg.
constexpr int DC_SEGMENTS = 1024;

__device__ __forceinline__ double block_sum_256(double a, double* red) {
  red[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(256) void k_sum_dc_segments(const float* __restrict__ part, long n,
                                                         double* __restrict__ seg) {
  __shared__ double red[256];
  const long len = (n + DC_SEGMENTS - 1) / DC_SEGMENTS;
  const long b = (long)blockIdx.x * len, e = min(n, b + len);
  double a = 0.0;
  for (long i = b + threadIdx.x; i < e; i += 256) a += (double)part[i];
  const double t = block_sum_256(a, red);
  if (threadIdx.x == 0) seg[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_sum_dc(const double* __restrict__ seg, double* __restrict__ out) {
  __shared__ double red[256];
  double a = 0.0;
  for (int i = threadIdx.x; i < DC_SEGMENTS; i += 256) a += seg[i];
  const double t = block_sum_256(a, red);
  if (threadIdx.x == 0) out[0] = t;
}

__global__ void k_ppc(Problem p, int nfp, cf* __restrict__ ppc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfp) return;
  if (k < p.NF) {
    const double q = p.probe[k];
    ppc[2 * k] = {(float)(p.pulse_re[k] * q), (float)(p.pulse_im[k] * q)};
    ppc[2 * k + 1] = {(float)(p.corr * q), 0.f};
  } else {
    ppc[2 * k] = {0.f, 0.f};
    ppc[2 * k + 1] = {0.f, 0.f};
  }
}

// ---------------------------------------------------------------- matrix-core form
// The transmit sums above are contractions over elements with the transmit
// index free, i.e. products of matrices. Per frequency f, with
//   R(e, m) = A_t(f, e)        (m = t' < 16),
//   R(e, m) = conj(g_t(f, e))  (m = 16 + t'),   t = 16 q + t' (transmit group q)
// and G(s, e) the Green's function summed over sub-elements:
//   Y = G R                         (16 scatterers x 32, contraction over e)
//   P_t(s) = PP Y(s, t'),  B_t(s) = Y(s, 16 + t')
//   L(s, m) = c PP Y(s, m ^ 16)     (so that W = L R^T is the W of k_adj_pos summed over t)
//   W(s, e) = sum_m L(s, m) R(e, m) = sum_t c PP A_t(e) B_t(s) + c P_t(s) conj(g_t(e))
// and sum_e G(s, e) W(s, e) = 2 sum_t c P_t(s) B_t(s), so that
//   dL/drc_s  = 1/2 Re sum_{f,e} G W = 1/2 sum_{e,j} amp_j Re S0_j   (G = sum_j amp_j E_j)
//   dL/dth_s  = rc_s sum_{e,j} a_th Re S0_j + Re(b_th S1_j),
//   S0_j = sum_f E_j W, S1_j = sum_f f E_j W      (as in k_adj_pos)
// Y and W^T are split-BF16 MFMA products (hi*hi + hi*lo + lo*hi, FP32
// accumulation) with the complex parts in the 32-deep contraction
// ([Re | Im] of 16 complex indices). For the real part one operand is negated
// in its upper half, for the imaginary part its halves are swapped: G (the A
// operand) in Y, L^T (the B operand, written in both forms by the reduction)
// in W^T. G is generated once per (scatterer, element, frequency) for all
// transmits of a group instead of once per transmit.
//
// Workgroup: 16 scatterers x nw element slices (one wave per 16 elements),
// AM_FC frequencies. Lane l of wave w holds scatterer s0 + l % 16 and elements
// 16 w + 4 (l / 16) + c, c < 4: the A-operand layout of G in Y = G R and the
// result layout of W^T = R L^T, so that G, E_j and W of one term sit in the
// same lane. Y is summed over the waves through LDS.
constexpr int AM_FC = 32;    // frequencies per chunk = the FP64 reseed interval of the recurrence (3 k + 2)
constexpr int AM_CPI = 8;    // chunks per work item (the per-item geometry and gradient coefficients are shared)
constexpr int AM_TG = 16;    // transmits per group
constexpr int AM_MAXW = 8;   // element slices per workgroup (E <= 128)
constexpr int AM_NSUB_MAX = 2;
constexpr int AM_YROW = 16;  // LDS row of Y: 16 scatterers
// Y of one slice is a block of 32 (m) x 16 (s) floats, [x][h][8]: the 4 rows s = 4 g .. 4 g + 3 of column
// m = 16 x + col (one MFMA result lane, written as one 16-byte store) sit at 4-float slot
// 8 h + ((col + 4 (g mod 2)) mod 8), h = 4 (col / 8) + g. The LDS has 32 banks of 4 bytes and serves a
// 16-byte store 8 lanes at a time and a 4-byte access 32 lanes at a time: 8 consecutive lanes (8 columns,
// one g) land on 8 distinct slots modulo 8, and each 32 lanes of the slice sum (am_reduce: 4 consecutive
// columns x 8 rows) on 32 distinct banks, for its loads of Y as for its stores of L, so none of them
// conflicts (rows of 16 floats with XOR-swizzled groups cost 8 conflict cycles per 16-byte store).
__device__ __forceinline__ int am_yidx(int m, int s) {
  const int col = m & 15, g = s >> 2;
  return (m >> 4) * 256 + (8 * ((col >> 3) * 4 + g) + ((col + 4 * (g & 1)) & 7)) * 4 + (s & 3);
}

typedef __bf16 am_bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 am_bf16x2 __attribute__((ext_vector_type(2)));
typedef float am_floatx4 __attribute__((ext_vector_type(4)));
typedef float am_floatx2 __attribute__((ext_vector_type(2)));
typedef unsigned am_u32x4 __attribute__((ext_vector_type(4)));


__device__ __forceinline__ am_floatx4 am_mma(am_u32x4 a, am_u32x4 b, am_floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(am_bf16x8, a), __builtin_bit_cast(am_bf16x8, b),
                                                 c, 0, 0, 0);
}

__device__ __forceinline__ size_t am_ridx(int k, int nw, int w, int x, int hl, int lane) {
  return ((((size_t)k * nw + w) * 2 + x) * 2 + hl) * 64 + lane;
}

// Gc[t][e][k] = conj(g_t(f_k, e)) (0 beyond NF)
__global__ void k_adjm_conj(Problem p, int nfp, const cf* __restrict__ gspect, cf* __restrict__ Gc) {
  const long total = (long)p.NT * p.E * nfp;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i % nfp);
    const long te = i / nfp;
    const int e = (int)(te % p.E);
    const int t = (int)(te / p.E);
    cf v = {0.f, 0.f};
    if (k < p.NF) {
      const cf g = gspect[((size_t)t * p.NF + k) * p.E + e];
      v = {g.x, -g.y};
    }
    Gc[i] = v;
  }
}

// Split-BF16 fragments of R for transmit group q, per frequency and element slice w:
//   R1[k][w][n][hl][lane]: B operand of Y = G R, lane (m = 16 n + l % 16; elements 16 w + 4 (l / 16) + c),
//     as [Re of the 4 | Im of the 4], times c_k PP_k (so that the product is c PP Y);
//   R2[k][w][x][hl][lane]: A operand of W^T = R L^T, lane (element 16 w + l % 16; m = 16 x + 4 (l / 16) + c),
//     interleaved: word c = (Re, Im) of entry c (the order of the contraction index inside a fragment is
//     free as long as both operands share it; the L^T operands are written in the same order);
// zero for t >= NT, e >= E and k >= NF (k < nfp).
__global__ void k_adjm_rtab(Problem p, int nfp, int q, int nw, const cf* __restrict__ A, const cf* __restrict__ Gc,
                            const cf* __restrict__ ppc, am_u32x4* __restrict__ R1, am_u32x4* __restrict__ R2) {
  const long total = (long)nfp * nw * 2 * 64;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int lane = (int)(i & 63);
    const int x = (int)((i >> 6) & 1);
    const long kw = i >> 7;
    const int w = (int)(kw % nw);
    const int k = (int)(kw / nw);
    const int g = lane >> 4, col = lane & 15;
    const cf* src = x == 0 ? A : Gc;
    float re[4], im[4];
    unsigned h0, l0, h1, l1, h2, l2, h3, l3;
    // R1: one transmit, four elements
    {
      const int t = AM_TG * q + col;
      const cf pp = ppc[2 * k];
      const float ck = ppc[2 * k + 1].x;
      const cf cpp = {ck * pp.x, ck * pp.y};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int e = 16 * w + 4 * g + c;
        cf v = {0.f, 0.f};
        if (t < p.NT && e < p.E && k < p.NF) v = cmul(cpp, src[((size_t)t * p.E + e) * nfp + k]);
        re[c] = v.x;
        im[c] = v.y;
      }
      split_bf16_pair(re[0], re[1], h0, l0);
      split_bf16_pair(re[2], re[3], h1, l1);
      split_bf16_pair(im[0], im[1], h2, l2);
      split_bf16_pair(im[2], im[3], h3, l3);
      R1[am_ridx(k, nw, w, x, 0, lane)] = (am_u32x4){h0, h1, h2, h3};
      R1[am_ridx(k, nw, w, x, 1, lane)] = (am_u32x4){l0, l1, l2, l3};
    }
    // R2: one element, four transmits
    {
      const int e = 16 * w + col;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int t = AM_TG * q + 4 * g + c;
        cf v = {0.f, 0.f};
        if (t < p.NT && e < p.E && k < p.NF) v = src[((size_t)t * p.E + e) * nfp + k];
        re[c] = v.x;
        im[c] = v.y;
      }
      split_bf16_pair(re[0], im[0], h0, l0);
      split_bf16_pair(re[1], im[1], h1, l1);
      split_bf16_pair(re[2], im[2], h2, l2);
      split_bf16_pair(re[3], im[3], h3, l3);
      R2[am_ridx(k, nw, w, x, 0, lane)] = (am_u32x4){h0, h1, h2, h3};
      R2[am_ridx(k, nw, w, x, 1, lane)] = (am_u32x4){l0, l1, l2, l3};
    }
  }
}

// Operands of Y = G R for one frequency: G = sum_j amp_j E_j (the amplitudes ride in the recurrence: Ea_j =
// amp_j E_j) split as [Re | -Im] (real part) and [Im | Re] (imaginary part), hi and lo.
struct AmOps {
  am_u32x4 a1h, a1l, a2h, a2l;
  float gr[4], gi[4];
};

template <int NSUB>
__device__ __forceinline__ void am_gen(const cf (&Ea)[4][NSUB], AmOps& o) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    o.gr[c] = Ea[c][0].x;
    o.gi[c] = Ea[c][0].y;
#pragma unroll
    for (int j = 1; j < NSUB; ++j) {
      o.gr[c] += Ea[c][j].x;
      o.gi[c] += Ea[c][j].y;
    }
  }
  const am_u32x4 nim = {0u, 0u, 0x80008000u, 0x80008000u};
  unsigned h0, l0, h1, l1, h2, l2, h3, l3;
  split_bf16_pair(o.gr[0], o.gr[1], h0, l0);
  split_bf16_pair(o.gr[2], o.gr[3], h1, l1);
  split_bf16_pair(o.gi[0], o.gi[1], h2, l2);
  split_bf16_pair(o.gi[2], o.gi[3], h3, l3);
  o.a1h = (am_u32x4){h0, h1, h2, h3} ^ nim;
  o.a1l = (am_u32x4){l0, l1, l2, l3} ^ nim;
  o.a2h = (am_u32x4){h2, h3, h0, h1};
  o.a2l = (am_u32x4){l2, l3, l0, l1};
}

// Y over one slice's 16 elements; result lane (m = 16 x + col, scatterers 4 g .. 4 g + 3)
__device__ __forceinline__ void am_mfma1(const AmOps& o, const am_u32x4 (&b1)[2][2], am_floatx4 (&yr)[2],
                                         am_floatx4 (&yi)[2]) {
  const am_floatx4 zero4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    yr[x] = am_mma(o.a1h, b1[x][0], zero4);
    yi[x] = am_mma(o.a2h, b1[x][0], zero4);
  }
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    yr[x] = am_mma(o.a1h, b1[x][1], yr[x]);
    yi[x] = am_mma(o.a2h, b1[x][1], yi[x]);
  }
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    yr[x] = am_mma(o.a1l, b1[x][0], yr[x]);
    yi[x] = am_mma(o.a2l, b1[x][0], yi[x]);
  }
}

// Y of every slice summed in a fixed order (R1 carries c PP, so the sum is c PP Y); L(s, m ^ 16) = c PP Y(s, m)
// written as the two
// split-BF16 B operands of W^T = R L^T in the interleaved order of R2: word c of a lane's fragment is
// (Re L, -Im L) of its entry c (real part) or (Im L, Re L) (imaginary part), so each entry is one
// 32-bit word per operand: the sign flip of the upper half or the swap of the halves of its split pair.
// A lane writes 4 words (hi and lo of both parts); the 64 lanes of a wave write 64 consecutive words.
// LDS: ypr / ypi [w][32 x 16 block, am_yidx]; lfrag [x][part][hl][lane][4 words].
template <int NW>
__device__ __forceinline__ void am_reduce(const float* ypr, const float* ypi, unsigned short* lfrag) {
#pragma unroll
  for (int r = 0; r < 512 / (64 * NW); ++r) {
    // entry (m, sl) of thread t: each 32 threads take 4 consecutive m and 8 of the 16 rows, so that their L
    // stores (4 words of 8 lanes) and Y loads fall on 32 distinct banks
    const int t = threadIdx.x + r * 64 * NW;
    const int m = 4 * (t >> 6) + ((t >> 3) & 3), sl = 8 * ((t >> 5) & 1) + (t & 7);
    float tr[NW], ti[NW];
#pragma unroll
    for (int v = 0; v < NW; ++v) {
      tr[v] = ypr[v * 32 * AM_YROW + am_yidx(m, sl)];
      ti[v] = ypi[v * 32 * AM_YROW + am_yidx(m, sl)];
    }
    float yx = 0.f, yy = 0.f;
#pragma unroll
    for (int v = 0; v < NW; ++v) {
      yx += tr[v];
      yy += ti[v];
    }
    const cf Lv = {yx, yy};
    const int mt = m ^ 16;
    const int xx = mt >> 4, mm = mt & 15;
    const int ln = sl + 16 * (mm >> 2), cc = mm & 3;
    unsigned hb, lb;
    split_bf16_pair(Lv.x, Lv.y, hb, lb);  // (Re, Im) in the low / high half
    unsigned* f = (unsigned*)lfrag + (xx * 4 * 64 + ln) * 4 + cc;  // [x][part 0, hl 0][ln][cc]
    f[0] = hb ^ 0x80000000u;                                       // part 0 (real), hi: (Re, -Im)
    f[64 * 4] = lb ^ 0x80000000u;                                  // part 0, lo
    f[2 * 64 * 4] = __builtin_amdgcn_alignbit(hb, hb, 16);         // part 1 (imaginary), hi: (Im, Re)
    f[3 * 64 * 4] = __builtin_amdgcn_alignbit(lb, lb, 16);         // part 1, lo
  }
}

// One work item: scatterers [16 blockIdx.x, +16), the AM_FC frequencies [AM_FC blockIdx.y, +AM_FC)
// (the tables are zero past the band, so a padded frequency adds nothing), transmit group q, NW
// element slices (NW waves, NW * 16 >= E). Writes the item's partial sums (one per scatterer) of
// dL/drc and, times rc_s, of dL/dx, dL/dz, dL/dc at [(q nch + blockIdx.y) S + s].
//
// Three-stage software pipeline over the frequencies, one workgroup barrier per stage step:
//   A(i): G(f_i) and this slice's Y(f_i) = G R on the matrix cores, stored to LDS buffer i % 2;
//   B(i - 1): Y(f_{i-1}) summed over the slices into L(f_{i-1}) (LDS buffer (i - 1) % 2);
//   C(i - 2): W^T(f_{i-2}) = R L^T on the matrix cores and the lane's terms of f_{i-2}.
// Every LDS buffer is written one step after its last reader's step, so one barrier per step
// orders all of them, and each wave's vector work overlaps its own matrix products and the LDS
// latencies of the other stages. E, G of three consecutive frequencies rotate through three
// register sets (the step sequence is unrolled by three), so nothing is copied.
template <int NSUB, int NW>
__global__ __launch_bounds__(64 * NW) void k_adjm(Problem p, int q, int nch, const am_u32x4* __restrict__ R1,
                                                  const am_u32x4* __restrict__ R2,
                                                  float* __restrict__ prc, float* __restrict__ px,
                                                  float* __restrict__ pz, float* __restrict__ pc) {
  static_assert(AM_FC % 3 == 2, "the steady steps come in groups of three");
  constexpr int nw = NW;
  constexpr int YB = NW * 32 * AM_YROW;       // floats per Y buffer
  constexpr int LB = 2 * 2 * 2 * 64 * 8;      // shorts per L buffer
  __shared__ __attribute__((aligned(16))) float ypr[2 * YB];   // [buf][w][m][s]: Re Y of each slice
  __shared__ __attribute__((aligned(16))) float ypi[2 * YB];   // Im Y
  __shared__ __attribute__((aligned(16))) unsigned short lfrag[2 * LB];  // [buf][x][part][hl][lane][8]
  __shared__ float red[NW][4][16];
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, col = lane & 15;
  int k0 = 0;                                 // first frequency of the chunk being run
  const int klast = nch * AM_FC - 1;          // last padded frequency (tables hold nch * AM_FC)
  const int s = blockIdx.x * 16 + col;
  const bool valid = s < p.S;
  const float xs = valid ? p.x[s] : 0.f;
  const float zs = valid ? p.z[s] : 1e-2f;
  const bool in_field = valid && !out_of_field(p, xs, zs);

  // amp_j E_j(f) = amp_j exp(f r_c (-alpha' + i 2 pi / c)) by FP32 recurrence from an FP64-phase seed at
  // the start of every chunk; the amplitude and the per-step factor hold for all chunks. An in-field term of
  // zero amplitude (a zero of the directivity) runs with amplitude 1e-20 instead, so that its S0 and S1 can
  // be recovered by dividing by the amplitude (its G differs from zero by 1e-20 |E|). The delay r_c / c and the
  // attenuation of each term are kept for the chunks' seeds, so that the geometry is evaluated once per item.
  cf EA[4][NSUB], EB[4][NSUB], EC[4][NSUB], Q[4][NSUB];
  float amp[4][NSUB];
  double tt[4][NSUB];
  float ta[4][NSUB];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int e = 16 * w + 4 * g + c;
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      const GeoF geo = geometry_fast(p, xs, zs, e < p.E ? e : p.E - 1, j);
      tt[c][j] = geo.t;
      ta[c][j] = geo.att;
      amp[c][j] = (in_field && e < p.E) ? (geo.amp != 0.f ? geo.amp : 1e-20f) : 0.f;
      Q[c][j] = green_step_fast(geo, p.df);
      EA[c][j] = EB[c][j] = EC[c][j] = {0.f, 0.f};
    }
  }
  // S0_j = sum_f amp_j E_j W over the item's frequencies f_0 .. f_K (in order), and V_j = sum_k of the partial
  // sums of S0_j before step k, so that sum_f f amp_j E_j W = f_K S0_j - df V_j (summation by parts)
  cf S0[4][NSUB], V[4][NSUB];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int j = 0; j < NSUB; ++j) S0[c][j] = V[c][j] = {0.f, 0.f};
  am_u32x4 b1[2][2], a2[2][2];

  // stage A(i): operands and Y of f_i from E(f_i) (matrix products), then stored to Y buffer i % 2
  auto stage_a_mma = [&](int i, const cf (&Ei)[4][NSUB], am_floatx4 (&yr)[2], am_floatx4 (&yi)[2]) {
    AmOps o;
    am_gen<NSUB>(Ei, o);
    am_mfma1(o, b1, yr, yi);
    const int kn = min(k0 + i + 1, klast);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int hl = 0; hl < 2; ++hl) b1[x][hl] = R1[am_ridx(kn, nw, w, x, hl, lane)];
  };
  auto stage_a_store = [&](int i, const am_floatx4 (&yr)[2], const am_floatx4 (&yi)[2]) {
    float* yb = ypr + (i & 1) * YB;
    float* ib = ypi + (i & 1) * YB;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      *(am_floatx4*)(yb + w * 32 * AM_YROW + am_yidx(16 * x + col, 4 * g)) = yr[x];
      *(am_floatx4*)(ib + w * 32 * AM_YROW + am_yidx(16 * x + col, 4 * g)) = yi[x];
    }
  };
  auto stage_a = [&](int i, const cf (&Ei)[4][NSUB]) {
    am_floatx4 yr[2], yi[2];
    stage_a_mma(i, Ei, yr, yi);
    stage_a_store(i, yr, yi);
  };
  // stage B(i): L of f_i from Y buffer i % 2 into L buffer i % 2
  auto stage_b = [&](int i) { am_reduce<NW>(ypr + (i & 1) * YB, ypi + (i & 1) * YB, lfrag + (i & 1) * LB); };
  // stage C(i): W^T of f_i (A = R as stored, B = the two variants of L^T; result lane = the lane's own
  // terms), then the terms of f_i
  auto stage_c_mma = [&](int i, am_floatx4& wr, am_floatx4& wi) {
    const am_u32x4* lf4 = (const am_u32x4*)(lfrag + (i & 1) * LB);
    wr = (am_floatx4){0.f, 0.f, 0.f, 0.f};
    wi = (am_floatx4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const am_u32x4 brh = lf4[(x * 4 + 0) * 64 + lane];
      const am_u32x4 brl = lf4[(x * 4 + 1) * 64 + lane];
      const am_u32x4 bih = lf4[(x * 4 + 2) * 64 + lane];
      const am_u32x4 bil = lf4[(x * 4 + 3) * 64 + lane];
      wr = am_mma(a2[x][0], brh, wr);
      wi = am_mma(a2[x][0], bih, wi);
      wr = am_mma(a2[x][0], brl, wr);
      wi = am_mma(a2[x][0], bil, wi);
      wr = am_mma(a2[x][1], brh, wr);
      wi = am_mma(a2[x][1], bih, wi);
    }
    const int kn = min(k0 + i + 1, klast);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int hl = 0; hl < 2; ++hl) a2[x][hl] = R2[am_ridx(kn, nw, w, x, hl, lane)];
  };
  auto stage_c_acc = [&](const cf (&Ei)[4][NSUB], const am_floatx4& wr, const am_floatx4& wi) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const cf Wc = {wr[c], wi[c]};
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        const cf ea = Ei[c][j];
        V[c][j].x += S0[c][j].x;
        V[c][j].y += S0[c][j].y;
        S0[c][j].x = fmaf(ea.x, Wc.x, fmaf(-ea.y, Wc.y, S0[c][j].x));
        S0[c][j].y = fmaf(ea.x, Wc.y, fmaf(ea.y, Wc.x, S0[c][j].y));
      }
    }
  };
  auto stage_c = [&](int i, const cf (&Ei)[4][NSUB]) {
    am_floatx4 wr, wi;
    stage_c_mma(i, wr, wi);
    stage_c_acc(Ei, wr, wi);
  };
  auto advance = [&](const cf (&Ei)[4][NSUB], cf (&Eo)[4][NSUB]) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int j = 0; j < NSUB; ++j) Eo[c][j] = cmul(Ei[c][j], Q[c][j]);
  };
  // steady step i (2 <= i < AM_FC): Ea = E(f_{i-2}) (then E(f_{i+1})), Ec = E(f_i). The matrix products of
  // C(i - 2) and A(i) are issued first; the vector work of B(i - 1) and of the terms of f_{i-2} runs while
  // they complete, and Y(f_i) is stored last.
  auto steady = [&](int i, cf (&Ea)[4][NSUB], cf (&Ec)[4][NSUB]) {
    am_floatx4 wr, wi, yr[2], yi[2];
    stage_c_mma(i - 2, wr, wi);
    stage_a_mma(i, Ec, yr, yi);
    stage_b(i - 1);
    stage_c_acc(Ea, wr, wi);
    stage_a_store(i, yr, yi);
    advance(Ec, Ea);
    __syncthreads();
  };

#pragma unroll 1
  for (int cc = 0; cc < AM_CPI; ++cc) {
  const int chunk = blockIdx.y * AM_CPI + cc;
  if (chunk >= nch) break;  // uniform over the workgroup
  k0 = chunk * AM_FC;
  const double fstart = p.f0 + (double)k0 * p.df;
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int j = 0; j < NSUB; ++j) EC[c][j] = green_fast(GeoF{tt[c][j], 0.f, ta[c][j], 0.f}, fstart, amp[c][j]);
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int hl = 0; hl < 2; ++hl) {
      b1[x][hl] = R1[am_ridx(k0, nw, w, x, hl, lane)];
      a2[x][hl] = R2[am_ridx(k0, nw, w, x, hl, lane)];
    }
  // step 0: E(f_0) in EC, E(f_1) -> EA
  stage_a(0, EC);
  advance(EC, EA);
  __syncthreads();
  // step 1: E(f_1) in EA, E(f_2) -> EB
  stage_a(1, EA);
  stage_b(0);
  advance(EA, EB);
  __syncthreads();
  // steps 2 .. AM_FC - 1, three at a time: (E(f_{i-2}), E(f_i)) = (EC, EB), (EA, EC), (EB, EA)
  for (int i = 2; i < AM_FC; i += 3) {
    steady(i, EC, EB);
    steady(i + 1, EA, EC);
    steady(i + 2, EB, EA);
  }
  // drain: step AM_FC (B of the last frequency, C of the one before), step AM_FC + 1 (C of the last)
  stage_b(AM_FC - 1);
  stage_c(AM_FC - 2, EC);
  __syncthreads();
  stage_c(AM_FC - 1, EA);
  __syncthreads();  // the next chunk's first steps rewrite the LDS buffers
  }  // chunks

  // derivative coefficients once per item; sums over the lane's terms, its 4 lane groups and the slices.
  // The amplitudes do not depend on frequency, so sum_f Re(G W) = sum_j Re S0_j: the rc gradient comes from
  // the same sums (S0 = 0 for padded elements and out-of-field scatterers). The position and speed-of-sound
  // terms take S0_j / amp_j and S1_j / amp_j (amp_j != 0 for every in-field term of a real element).
  float racc = 0.f;
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int j = 0; j < NSUB; ++j) racc += S0[c][j].x;
  float gx = 0.f, gz = 0.f, gcs = 0.f;
  if (in_field) {
    const float fc_hz = (float)(p.min_dist > 0 ? 0.5 / (p.min_dist * p.inv_c) : 1.0);
    const int kend = min((int)(blockIdx.y + 1) * AM_CPI, nch) * AM_FC - 1;  // the item's last frequency
    const float fend = (float)(p.f0 + (double)kend * p.df);
    const float dff = (float)p.df;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int e = 16 * w + 4 * g + c;
      if (e >= p.E) continue;
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        GeoF geo;
        Dcoef dc;
        coeffs_fast(p, xs, zs, e, j, fc_hz, geo, dc);
        const float ia = 1.0f / amp[c][j];
        const float s0 = S0[c][j].x * ia;
        const cf s1 = {(fend * S0[c][j].x - dff * V[c][j].x) * ia, (fend * S0[c][j].y - dff * V[c][j].y) * ia};
        gx += dc.ax * s0 + (dc.bx.x * s1.x - dc.bx.y * s1.y);
        gz += dc.az * s0 + (dc.bz.x * s1.x - dc.bz.y * s1.y);
        gcs += dc.ac * s0 + (dc.bc.x * s1.x - dc.bc.y * s1.y);
      }
    }
  }
  float v[4] = {0.5f * racc, gx, gz, gcs};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] += __shfl_xor(v[i], 16);
    v[i] += __shfl_xor(v[i], 32);
  }
  if (g == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) red[w][i][col] = v[i];
  }
  __syncthreads();
  if (threadIdx.x < 16 && valid) {
    float a[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < NW; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] += red[u][i][threadIdx.x];
    const float rcs = p.rc[s];
    const bool live = in_field && rcs != 0.f;
    const size_t o = ((size_t)q * gridDim.y + blockIdx.y) * p.S + s;
    prc[o] = a[0];
    px[o] = live ? a[1] * rcs : 0.f;
    pz[o] = live ? a[2] * rcs : 0.f;
    pc[o] = live ? a[3] * rcs : 0.f;
  }
}

// ---------------------------------------------------------------- host side
void launch_tx_table(const Problem& p, int nfp, cf* A, hipStream_t st);  // simus_fp32.hip

struct AdjLayout {
  int nfb, nfp;
  size_t oA, oTA, oTG, oGc, oppc, oP, oB, odrc, odx, odz, odc, oseg, total;
};

static size_t al(size_t v) { return (v + 255) & ~(size_t)255; }

static AdjLayout adj_layout(const Problem& p) {
  AdjLayout L{};
  L.nfb = (p.NF + AFB - 1) / AFB;
  L.nfp = L.nfb * AFB;
  const size_t tab = (size_t)p.NT * p.E * L.nfp * sizeof(cf);
  const size_t pb = (size_t)L.nfb * p.S * AFB * sizeof(cf);
  const size_t part = (size_t)L.nfb * p.S * sizeof(float);
  size_t o = 0;
  L.oA = o;   o = al(o + tab);
  L.oTA = o;  o = al(o + tab);
  L.oTG = o;  o = al(o + tab);
  L.oGc = o;  o = al(o + tab);
  L.oppc = o; o = al(o + 2 * (size_t)L.nfp * sizeof(cf));
  L.oP = o;   o = al(o + pb);
  L.oB = o;   o = al(o + pb);
  L.odrc = o; o = al(o + part);
  L.odx = o;  o = al(o + part);
  L.odz = o;  o = al(o + part);
  L.odc = o;  o = al(o + part * p.NT);
  L.oseg = o; o = al(o + DC_SEGMENTS * sizeof(double));
  L.total = o;
  return L;
}


template <int NSUB, bool FD, bool GR = false>
static void adj_launch(const Problem& p, const AdjLayout& L, char* ws, int t, float* d_rc, float* d_x, float* d_z,
                       hipStream_t st) {
  const dim3 grid((p.S + 255) / 256, L.nfb);
  cf* A = (cf*)(ws + L.oA);
  cf* TA = (cf*)(ws + L.oTA);
  cf* TG = (cf*)(ws + L.oTG);
  cf* Gc = (cf*)(ws + L.oGc);
  cf* ppc = (cf*)(ws + L.oppc);
  cf* Pb = (cf*)(ws + L.oP);
  cf* Bb = (cf*)(ws + L.oB);
  float* drc = (float*)(ws + L.odrc);
  float* dx = (float*)(ws + L.odx);
  float* dz = (float*)(ws + L.odz);
  float* dcp = (float*)(ws + L.odc) + (size_t)t * L.nfb * p.S;
  const int slot = timing_start(st, FS_TIMING_ADJOINT);
  hipLaunchKernelGGL((k_adj_tx<NSUB, FD, GR>), grid, dim3(256), 0, st, p, t, L.nfp, A, Gc, ppc, Pb, Bb, drc);
  hipLaunchKernelGGL((k_adj_pos<NSUB, FD, GR>), grid, dim3(256), 0, st, p, t, L.nfp, TA, TG, Pb, Bb, dx, dz, dcp);
  timing_stop(slot, st);
  const dim3 g1((p.S + 255) / 256);
  hipLaunchKernelGGL(k_sum_parts, g1, dim3(256), 0, st, drc, L.nfb, p.S, t > 0 ? 1 : 0, d_rc);
  hipLaunchKernelGGL(k_sum_parts, g1, dim3(256), 0, st, dx, L.nfb, p.S, t > 0 ? 1 : 0, d_x);
  hipLaunchKernelGGL(k_sum_parts, g1, dim3(256), 0, st, dz, L.nfb, p.S, t > 0 ? 1 : 0, d_z);
}

static int adjoint_vector(const Problem& p, const cf* dspect, float* d_rc, float* d_x, float* d_z, double* d_c,
                          char* ws, hipStream_t st) {
  const AdjLayout L = adj_layout(p);
  {
    const long n = (long)p.NT * p.E * L.nfp;
    const int blocks = (int)std::min<long>((n + 255) / 256, 65536);
    launch_tx_table(p, L.nfp, (cf*)(ws + L.oA), st);
    hipLaunchKernelGGL(k_adj_tables, dim3(blocks), dim3(256), 0, st, p, L.nfp, dspect, (cf*)(ws + L.oTA),
                       (cf*)(ws + L.oTG), (cf*)(ws + L.oGc));
    hipLaunchKernelGGL(k_ppc, dim3((L.nfp + 255) / 256), dim3(256), 0, st, p, L.nfp, (cf*)(ws + L.oppc));
  }
  for (int t = 0; t < p.NT; ++t) {
    const bool fd = p.fulldir != 0;
    switch (p.NS) {
      case 1: fd ? adj_launch<1, true>(p, L, ws, t, d_rc, d_x, d_z, st) : adj_launch<1, false>(p, L, ws, t, d_rc, d_x, d_z, st); break;
      case 2: fd ? adj_launch<2, true>(p, L, ws, t, d_rc, d_x, d_z, st) : adj_launch<2, false>(p, L, ws, t, d_rc, d_x, d_z, st); break;
      case 3: fd ? adj_launch<3, true>(p, L, ws, t, d_rc, d_x, d_z, st) : adj_launch<3, false>(p, L, ws, t, d_rc, d_x, d_z, st); break;
      case 4: fd ? adj_launch<4, true>(p, L, ws, t, d_rc, d_x, d_z, st) : adj_launch<4, false>(p, L, ws, t, d_rc, d_x, d_z, st); break;
      default: fd ? adj_launch<4, true, true>(p, L, ws, t, d_rc, d_x, d_z, st) : adj_launch<4, false, true>(p, L, ws, t, d_rc, d_x, d_z, st); break;
    }
  }
  hipLaunchKernelGGL(k_sum_dc_segments, dim3(DC_SEGMENTS), dim3(256), 0, st, (const float*)(ws + L.odc),
                     (long)p.NT * L.nfb * p.S, (double*)(ws + L.oseg));
  hipLaunchKernelGGL(k_sum_dc, dim3(1), dim3(256), 0, st, (const double*)(ws + L.oseg), d_c);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

// Matrix-core form: workspace [A | Gc | ppc | R1 | R2 | partials rc, x, z, c | dc segments].
struct AdjmLayout {
  int nfp, nw, ngrp, nch, nitem;
  size_t oA, oGc, oppc, oR1, oR2, orc, ox, oz, oc, oseg, total;
};

static AdjmLayout adjm_layout(const Problem& p) {
  AdjmLayout L{};
  L.nch = (p.NF + AM_FC - 1) / AM_FC;
  L.nfp = L.nch * AM_FC;
  L.nitem = (L.nch + AM_CPI - 1) / AM_CPI;
  L.nw = p.E <= 64 ? 4 : AM_MAXW;  // element slices (waves) per workgroup: k_adjm<NSUB, 4 or 8>
  L.ngrp = (p.NT + AM_TG - 1) / AM_TG;
  const size_t tab = (size_t)p.NT * p.E * L.nfp * sizeof(cf);
  const size_t rt = (size_t)L.nfp * L.nw * 2 * 2 * 64 * sizeof(am_u32x4);
  const size_t part = (size_t)L.ngrp * L.nitem * p.S * sizeof(float);
  size_t o = 0;
  L.oA = o;   o = al(o + tab);
  L.oGc = o;  o = al(o + tab);
  L.oppc = o; o = al(o + 2 * (size_t)L.nfp * sizeof(cf));
  L.oR1 = o;  o = al(o + rt);
  L.oR2 = o;  o = al(o + rt);
  L.orc = o;  o = al(o + part);
  L.ox = o;   o = al(o + part);
  L.oz = o;   o = al(o + part);
  L.oc = o;   o = al(o + part);
  L.oseg = o; o = al(o + DC_SEGMENTS * sizeof(double));
  L.total = o;
  return L;
}

static int adjoint_matrix(const Problem& p, const cf* dspect, float* d_rc, float* d_x, float* d_z, double* d_c,
                          char* ws, hipStream_t st) {
  const AdjmLayout L = adjm_layout(p);
  cf* A = (cf*)(ws + L.oA);
  cf* Gc = (cf*)(ws + L.oGc);
  cf* ppc = (cf*)(ws + L.oppc);
  am_u32x4* R1 = (am_u32x4*)(ws + L.oR1);
  am_u32x4* R2 = (am_u32x4*)(ws + L.oR2);
  float* prc = (float*)(ws + L.orc);
  float* px = (float*)(ws + L.ox);
  float* pz = (float*)(ws + L.oz);
  float* pc = (float*)(ws + L.oc);
  {
    const long n = (long)p.NT * p.E * L.nfp;
    const int blocks = (int)std::min<long>((n + 255) / 256, 65536);
    launch_tx_table(p, L.nfp, A, st);
    hipLaunchKernelGGL(k_adjm_conj, dim3(blocks), dim3(256), 0, st, p, L.nfp, dspect, Gc);
    hipLaunchKernelGGL(k_ppc, dim3((L.nfp + 255) / 256), dim3(256), 0, st, p, L.nfp, ppc);
  }
  const long nr = (long)L.nfp * L.nw * 2 * 64;
  const int rblocks = (int)std::min<long>((nr + 255) / 256, 65536);
  const dim3 grid((p.S + 15) / 16, L.nitem);
  for (int q = 0; q < L.ngrp; ++q) {
    hipLaunchKernelGGL(k_adjm_rtab, dim3(rblocks), dim3(256), 0, st, p, L.nfp, q, L.nw, (const cf*)A,
                       (const cf*)Gc, (const cf*)ppc, R1, R2);
    const dim3 block(64 * L.nw);
    const int slot = timing_start(st, FS_TIMING_ADJOINT);
    if (p.NS == 1 && L.nw == 4)
      hipLaunchKernelGGL((k_adjm<1, 4>), grid, block, 0, st, p, q, L.nch, R1, R2, prc, px, pz, pc);
    else if (p.NS == 1)
      hipLaunchKernelGGL((k_adjm<1, 8>), grid, block, 0, st, p, q, L.nch, R1, R2, prc, px, pz, pc);
    else if (L.nw == 4)
      hipLaunchKernelGGL((k_adjm<2, 4>), grid, block, 0, st, p, q, L.nch, R1, R2, prc, px, pz, pc);
    else
      hipLaunchKernelGGL((k_adjm<2, 8>), grid, block, 0, st, p, q, L.nch, R1, R2, prc, px, pz, pc);
    timing_stop(slot, st);
  }
  const int nb = L.ngrp * L.nitem;
  const dim3 g1((p.S + 255) / 256);
  hipLaunchKernelGGL(k_sum_parts, g1, dim3(256), 0, st, prc, nb, p.S, 0, d_rc);
  hipLaunchKernelGGL(k_sum_parts, g1, dim3(256), 0, st, px, nb, p.S, 0, d_x);
  hipLaunchKernelGGL(k_sum_parts, g1, dim3(256), 0, st, pz, nb, p.S, 0, d_z);
  hipLaunchKernelGGL(k_sum_dc_segments, dim3(DC_SEGMENTS), dim3(256), 0, st, (const float*)pc, (long)nb * p.S,
                     (double*)(ws + L.oseg));
  hipLaunchKernelGGL(k_sum_dc, dim3(1), dim3(256), 0, st, (const double*)(ws + L.oseg), d_c);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

// method (fs_adjoint_method, passed with every call): FS_ADJOINT_AUTO = matrix cores where they apply
// (n_elem <= 128, n_sub <= 2, centre-frequency directivity), else vector; FS_ADJOINT_MATRIX = matrix cores only; FS_ADJOINT_VECTOR = vector only.
static bool matrix_applies(const Problem& p) { return p.E <= 16 * AM_MAXW && p.NS <= AM_NSUB_MAX && !p.fulldir; }

static bool use_matrix(const Problem& p, int method) { return method != FS_ADJOINT_VECTOR && matrix_applies(p); }

int fill_zero(void* ptr, size_t n_words, hipStream_t st);  // rf.hip

size_t adjoint_workspace(const Problem& p, int method) {
  return use_matrix(p, method) ? adjm_layout(p).total : adj_layout(p).total;
}

int adjoint_fp32(const Problem& p, int method, const cf* dspect, float* d_rc, float* d_x, float* d_z, double* d_c,
                 char* ws, hipStream_t st) {
  if (method == FS_ADJOINT_MATRIX && !matrix_applies(p)) return FS_ERR_UNSUPPORTED;
  if (p.S == 0) return fill_zero(d_c, 2, st);  // no scatterers: the loss does not depend on c
  if (use_matrix(p, method)) return adjoint_matrix(p, dspect, d_rc, d_x, d_z, d_c, ws, st);
  return adjoint_vector(p, dspect, d_rc, d_x, d_z, d_c, ws, st);
}

}  // namespace fs
