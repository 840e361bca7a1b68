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
// k_adj_tx (lane = scatterer): P_t, B_t for one transmit, frequency blocks of
// 32, stored to HBM; the rc gradient is reduced on the fly.
// k_adj_pos (lane = scatterer): regenerates E_j and the derivative
// coefficients and accumulates the x, z, c gradients against W_t.
// Per-(transmit, block, scatterer) partial sums are reduced in fixed order.
#include <algorithm>

#include "fs_device.h"

namespace fs {

constexpr int AFB = 32;

struct Dcoef {
  float ax, az, ac;  // real parts
  cf bx, bz, bc;     // per-Hz parts
};

// amp, propagation and their derivatives for one (s, e, j) term, with the gfx950 hardware
// transcendentals (1/r from the FP32 estimate plus one FP64 Newton step, v_sin / v_cos in
// revolutions, reciprocals) and the propagation constants of GeoF (attenuation pre-multiplied
// by log2 e).
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
  float D, dDdy;
  if (fabsf(yk) < 1e-3f) {
    D = 1.0f - yk * yk / 6.0f;
    dDdy = -yk / 3.0f;
  } else {
    const float rev = yk * (0.5f / kPi);
    const float iy = __builtin_amdgcn_rcpf(yk);
    D = __builtin_amdgcn_sinf(rev) * iy;
    dDdy = (__builtin_amdgcn_cosf(rev) - D) * iy;
  }
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

// P_t (without rc) and B_t for one transmit; drc_part[fb][s] = sum_k Re(c_k P B)
template <int NSUB>
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
  for (int e = 0; e < p.E; ++e) {
    cf g[NSUB], q[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      const GeoF geo = geometry_fast(p, xs, zs, e, j);
      g[j] = green_fast(geo, fstart, geo.amp);
      q[j] = green_step_fast(geo, p.df);
    }
    const cf* Ae = At + (size_t)e * nfp;
    const cf* Ge = Gt + (size_t)e * nfp;
#pragma unroll
    for (int k = 0; k < AFB; ++k) {
      cf G = {0.f, 0.f};
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        G.x += g[j].x;
        G.y += g[j].y;
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

template <int NSUB>
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
  for (int e = 0; e < p.E; ++e) {
    cf E[NSUB], q[NSUB];
    Dcoef dc[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      GeoF geo;
      coeffs_fast(p, xs, zs, e, j, fc_hz, geo, dc[j]);
      E[j] = green_fast(geo, fstart, 1.0f);  // E_j carries the propagation factor only
      q[j] = green_step_fast(geo, p.df);
    }
    const cf* TAe = TAt + (size_t)e * nfp;
    const cf* TGe = TGt + (size_t)e * nfp;
    // dg/dth = E (a + b f) is linear in f: accumulate S0 = sum_k E W and S1 = sum_k f E W over the
    // block, then Re(dg W) summed over k = a Re(S0) + Re(b S1).
    cf S0[NSUB], S1[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) S0[j] = S1[j] = {0.f, 0.f};
#pragma unroll
    for (int k = 0; k < AFB; ++k) {
      const float f = (float)(fstart + (double)k * p.df);
      cf W = {0.f, 0.f};
      cmac(W, TAe[k], B[k]);
      cmac(W, P[k], TGe[k]);
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        const cf EW = cmul(E[j], W);
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

size_t adjoint_workspace(const Problem& p) { return adj_layout(p).total; }

template <int NSUB>
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
  hipLaunchKernelGGL(k_adj_tx<NSUB>, grid, dim3(256), 0, st, p, t, L.nfp, A, Gc, ppc, Pb, Bb, drc);
  hipLaunchKernelGGL(k_adj_pos<NSUB>, grid, dim3(256), 0, st, p, t, L.nfp, TA, TG, Pb, Bb, dx, dz, dcp);
  const dim3 g1((p.S + 255) / 256);
  hipLaunchKernelGGL(k_sum_parts, g1, dim3(256), 0, st, drc, L.nfb, p.S, t > 0 ? 1 : 0, d_rc);
  hipLaunchKernelGGL(k_sum_parts, g1, dim3(256), 0, st, dx, L.nfb, p.S, t > 0 ? 1 : 0, d_x);
  hipLaunchKernelGGL(k_sum_parts, g1, dim3(256), 0, st, dz, L.nfb, p.S, t > 0 ? 1 : 0, d_z);
}

int adjoint_fp32(const Problem& p, const cf* dspect, float* d_rc, float* d_x, float* d_z, double* d_c, char* ws,
                 hipStream_t st) {
  if (p.S == 0) return FS_ERR_INVALID;
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
    switch (p.NS) {
      case 1: adj_launch<1>(p, L, ws, t, d_rc, d_x, d_z, st); break;
      case 2: adj_launch<2>(p, L, ws, t, d_rc, d_x, d_z, st); break;
      case 3: adj_launch<3>(p, L, ws, t, d_rc, d_x, d_z, st); break;
      case 4: adj_launch<4>(p, L, ws, t, d_rc, d_x, d_z, st); break;
      default: return FS_ERR_UNSUPPORTED;
    }
  }
  hipLaunchKernelGGL(k_sum_dc_segments, dim3(DC_SEGMENTS), dim3(256), 0, st, (const float*)(ws + L.odc),
                     (long)p.NT * L.nfb * p.S, (double*)(ws + L.oseg));
  hipLaunchKernelGGL(k_sum_dc, dim3(1), dim3(256), 0, st, (const double*)(ws + L.oseg), d_c);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

}  // namespace fs
