// Plain FP32 SIMUS spectrum: one transmit at a time, TX then RX.
//
// Reference algorithm per selected frequency f (simus.py:263-333,
// _simus_strategies.py:23-77):
//   P(f, s)  = PP(f) * sum_e G(f, e, s) A(f, e),   0 where s is out of field
//   X(f, e)  = corr * probe(f) * sum_s rc_s P(f, s) G(f, e, s)
// with G the sub-element mean of the Green's function and
// A(f, e) = w_e exp(i 2 pi f tau_e).
//
// Layout / mapping (MI355X, wave64):
//  * frequencies are processed in blocks of 32; at the start of every block
//    each (s, e, j) term is reseeded from an FP64 phase f * r_c / c mod 1 and
//    advanced by an FP32 complex recurrence inside the block;
//  * k_tx: lane = scatterer, elements are walked uniformly by the wave, so the
//    transmit table A(f, e) is a wave-uniform (scalar) load; A is geometric in f,
//    so (without per-frequency directivity) it is read at each block start and
//    folded into the recurrence step of every term; a thread takes one or two
//    frequency blocks, whose 32 P(f, s) accumulators each live in VGPRs; the
//    result Q = rc PP P goes to HBM as Q[block][s][32];
//  * k_rx: lane = element, scatterers are walked uniformly, so Q(f, s) is a
//    scalar load; the 32 X(f, e) accumulators live in VGPRs; every wave writes a
//    private partial sum (no atomics) and k_reduce adds them in a fixed order,
//    so results are bitwise reproducible. Where the amplitude is non-negative
//    (centre-frequency directivity, up to 4 sub-elements) the frequency-
//    independent part of each term (clipped distance, log2 amplitude, step) is
//    read from a geometry table written once per call and shared by all
//    transmits, so a block's seed is one FP64 phase reduction, one exp2 and one
//    sine / cosine pair per term.
#include <algorithm>

#include "fs_device.h"

namespace fs {

constexpr int FB = 32;  // frequencies per reseed block

// A[t][e][k] = w exp(i 2 pi f_k tau), k < nfp (padded to a multiple of FB).
__global__ void k_tx_table(Problem p, int nfp, cf* __restrict__ A) {
  const long total = (long)p.NT * p.E * nfp;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int k = (int)(i % nfp);
    const long te = i / nfp;
    const double tau = p.delay[te];
    const double w = p.apod[te];
    const double f = p.f0 + (double)k * p.df;
    const double nu = f * tau;
    const float ph = (float)(nu - floor(nu));
    float s, c;
    sincospif(2.0f * ph, &s, &c);
    A[i] = {(float)w * c, (float)w * s};
  }
}

void launch_tx_table(const Problem& p, int nfp, cf* A, hipStream_t st) {
  const long n = (long)p.NT * p.E * nfp;
  hipLaunchKernelGGL(k_tx_table, dim3((int)std::min<long>((n + 255) / 256, 65536)), dim3(256), 0, st, p, nfp, A);
}

// PP(f_k) = pulse * probe, and probe, padded with zeros beyond NF.
__global__ void k_spectra(Problem p, int nfp, cf* __restrict__ pp, float* __restrict__ pr) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfp) return;
  if (k < p.NF) {
    const double q = p.probe[k];
    pp[k] = {(float)(p.pulse_re[k] * q), (float)(p.pulse_im[k] * q)};
    pr[k] = (float)q;
  } else {
    pp[k] = {0.f, 0.f};
    pr[k] = 0.f;
  }
}

// Sub-elements [j0, j0 + NSUB) of a group: with GROUPED the problem has p.NS sub-elements, taken
// NSUB at a time, and those past p.NS get a zero amplitude (their terms are exactly 0); without it
// p.NS == NSUB and there is one group.
template <int NSUB, bool GROUPED>
__device__ __forceinline__ int n_groups(const Problem& p) {
  return GROUPED ? (p.NS + NSUB - 1) / NSUB : 1;
}
template <bool GROUPED>
__device__ __forceinline__ GeoF group_geometry(const Problem& p, double xs, double zs, int e, int jj) {
  const bool ok = !GROUPED || jj < p.NS;
  GeoF g = geometry_fast(p, xs, zs, e, ok ? jj : 0);
  if (!ok) g.amp = 0.f;
  return g;
}

// A thread takes BPT consecutive 32-frequency blocks of the launch (lb = BPT blockIdx.y + b,
// b < BPT, lb < nb), so that the per-(scatterer, element, sub-element) geometry and frequency
// step are evaluated once for all of them; every block is still seeded from its own FP64 phase.
// MODE 0 (simus): Q[lb][s][k] = rc_s PP_k P_k(s).
// MODE 1 (pfield): pw[fb][s] = sum_{k in block} |PP_k P_k(s)|^2 (reference _freq_step_body, _pfield_strategies.py:21-50).
template <int NSUB, bool FULLDIR, int MODE, int BPT, bool GROUPED = false>
__global__ __launch_bounds__(256) void k_tx(Problem p, int tx, int nfp, int fb0, int nb, const cf* __restrict__ A,
                                            const cf* __restrict__ pp, cf* __restrict__ Q, float* __restrict__ pw) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int lb0 = blockIdx.y * BPT;
  const bool valid = s < p.S;
  const float xs = valid ? p.x[s] : 0.f;
  const float zs = valid ? p.z[s] : 1e-2f;
  const float rcs = (MODE == 0 && valid) ? p.rc[s] : 1.f;
  const bool out = !valid || out_of_field(p, xs, zs) || rcs == 0.f;

  cf P[BPT][FB];
#pragma unroll
  for (int b = 0; b < BPT; ++b)
#pragma unroll
    for (int k = 0; k < FB; ++k) P[b][k] = {0.f, 0.f};

  const cf* At = A + (size_t)tx * p.E * nfp + (size_t)(fb0 + lb0) * FB;
  const double* taus = p.delay + (size_t)tx * p.E;  // A(f, e) = w_e exp(i 2 pi f tau_e)
  const int ngroups = n_groups<NSUB, GROUPED>(p);
  for (int e = 0; e < p.E; ++e) {
  // Without per-frequency directivity the transmit weight is geometric in f as well: h_j = g_j A
  // advances by q_j exp(i 2 pi df tau_e) (phase reduced in FP64), seeded with A at the block start,
  // so the inner step is the recurrence and the sum of h_j into P, with no multiply by A.
  cf astep = {1.f, 0.f};
  if constexpr (!FULLDIR) {
    const double nu = p.df * taus[e];
    const float ph = (float)(nu - floor(nu));
    astep = {__builtin_amdgcn_cosf(ph), __builtin_amdgcn_sinf(ph)};
  }
  for (int jg = 0; jg < ngroups; ++jg) {
    GeoF geo[NSUB];
    cf q[NSUB];
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      geo[j] = group_geometry<GROUPED>(p, xs, zs, e, jg * NSUB + j);
      q[j] = green_step_fast(geo[j], p.df);
      if constexpr (!FULLDIR) q[j] = cmul(q[j], astep);
    }
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      if (lb0 + b >= nb) break;  // uniform over the launch row
      const double fstart = p.f0 + (double)((fb0 + lb0 + b) * FB) * p.df;
      cf g[NSUB];
      const cf* Ae = At + (size_t)e * nfp + b * FB;
      if constexpr (FULLDIR) {
#pragma unroll
        for (int j = 0; j < NSUB; ++j) g[j] = green_fast(geo[j], fstart, geo[j].amp);
#pragma unroll
        for (int k = 0; k < FB; ++k) {
          cf G = {0.f, 0.f};
#pragma unroll
          for (int j = 0; j < NSUB; ++j) {
            const float d = sinc_fast(p.kappa_per_hz * (float)(fstart + (double)k * p.df) * geo[j].sint);
            G.x = fmaf(g[j].x, d, G.x);
            G.y = fmaf(g[j].y, d, G.y);
            g[j] = cmul(g[j], q[j]);
          }
          cmac(P[b][k], G, Ae[k]);
        }
      } else {
        const cf a0 = Ae[0];
#pragma unroll
        for (int j = 0; j < NSUB; ++j) g[j] = cmul(green_fast(geo[j], fstart, geo[j].amp), a0);
#pragma unroll
        for (int k = 0; k < FB; ++k) {
#pragma unroll
          for (int j = 0; j < NSUB; ++j) {
            P[b][k].x += g[j].x;
            P[b][k].y += g[j].y;
            g[j] = cmul(g[j], q[j]);
          }
        }
      }
    }
  }
  }
  if (!valid) return;
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int lb = lb0 + b;
    if (lb >= nb) break;
    const int k0 = (fb0 + lb) * FB;
    if constexpr (MODE == 0) {
      cf* Qs = Q + ((size_t)lb * p.S + s) * FB;
#pragma unroll
      for (int k = 0; k < FB; ++k) {
        cf v = out ? cf{0.f, 0.f} : cmul(P[b][k], pp[k0 + k]);
        Qs[k] = {v.x * rcs, v.y * rcs};
      }
    } else {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < FB; ++k) {
        const cf v = cmul(P[b][k], pp[k0 + k]);  // pp is zero beyond NF
        acc = fmaf(v.x, v.x, fmaf(v.y, v.y, acc));
      }
      pw[(size_t)(fb0 + lb) * p.S + s] = out ? 0.f : acc;
    }
  }
}

// out[s] = sqrt(corr * sum_fb pw[fb][s])
__global__ void k_pfield_reduce(const float* __restrict__ pw, int nfb, int S, float corr, float* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  float acc = 0.f;
  for (int b = 0; b < nfb; ++b) acc += pw[(size_t)b * S + s];
  out[s] = sqrtf(acc * corr);
}

// part[(chunk * 4 + wave)][k - k_begin][e], k_begin = first frequency of the block group.
// TAB: the frequency-independent part of every term (clipped distance, log2 of the amplitude, the step)
// comes from the geometry table [s][e][j] (rows of p.E elements), so that a block's seed costs one FP64
// phase reduction, one exp2 and one sine / cosine pair per term instead of the FP64 geometry.
template <int NSUB, bool FULLDIR, bool GROUPED = false, bool TAB = false>
__global__ __launch_bounds__(256) void k_rx(Problem p, int nbuf, int fb0, int s_per_wave, const cf* __restrict__ Q,
                                            cf* __restrict__ part, const double* __restrict__ rt,
                                            const float* __restrict__ la, const float2* __restrict__ qt) {
  static_assert(!TAB || (!FULLDIR && !GROUPED), "the table holds centre-frequency directivity, up to 4 sub-elements");
  const int lane = threadIdx.x & 63;
  // wave-uniform by construction; readfirstlane tells the compiler, so that the
  // scatterer loop is a scalar loop and Q, x, z, rc are scalar loads
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int e = blockIdx.x * 64 + lane;
  const int fb = fb0 + blockIdx.y;
  const int k0 = fb * FB;
  const int chunk = blockIdx.z;
  const bool valid = e < p.E;
  const int ee = valid ? e : 0;

  cf X[FB];
#pragma unroll
  for (int k = 0; k < FB; ++k) X[k] = {0.f, 0.f};

  const double fstart = p.f0 + (double)k0 * p.df;
  const double nu0 = fstart * p.inv_c;                                     // revolutions per metre at fstart
  const float a0 = (float)((double)p.att_coef * (double)kLog2e * fstart);  // log2 attenuation per metre
  const int sb = (chunk * 4 + wave) * s_per_wave;
  const int se = min(sb + s_per_wave, p.S);
  const cf* Qb = Q + (size_t)blockIdx.y * p.S * FB;
  SubGeo sg[NSUB];
  if constexpr (!GROUPED) {
#pragma unroll
    for (int j = 0; j < NSUB; ++j) sg[j] = sub_geo(p, ee, j);
  }
  for (int s = sb; s < se; ++s) {
    const cf* Qs = Qb + (size_t)s * FB;
    if (p.rc[s] == 0.f) continue;  // wave-uniform: the scatterer contributes nothing
    const float xs = p.x[s], zs = p.z[s];
    if (out_of_field(p, xs, zs)) continue;
    // the scatterer's 32 Q values (scalar loads) are requested before the geometry,
    // whose vector work then covers their latency
    cf qv[FB];
#pragma unroll
    for (int k = 0; k < FB; ++k) qv[k] = Qs[k];
    const int ngroups = n_groups<NSUB, GROUPED>(p);
    for (int jg = 0; jg < ngroups; ++jg) {
    cf g[NSUB], q[NSUB];
    GeoF geo[NSUB];
    if constexpr (TAB) {
      const TableTerm<NSUB> tt = load_term<NSUB>(rt, la, qt, (size_t)s * p.E, ee, p.E);
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        const float ph = (float)__builtin_amdgcn_fract(nu0 * tt.r[j]);  // v_fract_f64: x - floor(x)
        const float m = __builtin_amdgcn_exp2f(fmaf(-a0, tt.rf[j], tt.l[j]));
        g[j] = {m * __builtin_amdgcn_cosf(ph), m * __builtin_amdgcn_sinf(ph)};
        q[j] = {tt.q[j].x, tt.q[j].y};
      }
    } else {
#pragma unroll
    for (int j = 0; j < NSUB; ++j) {
      if constexpr (GROUPED) {
        const int jj = jg * NSUB + j;
        geo[j] = geometry_fast(p, xs, zs, sub_geo(p, ee, jj < p.NS ? jj : 0));
        if (jj >= p.NS) geo[j].amp = 0.f;
      } else {
        geo[j] = geometry_fast(p, xs, zs, sg[j]);
      }
      g[j] = green_fast(geo[j], fstart, geo[j].amp);
      q[j] = green_step_fast(geo[j], p.df);
    }
    }
#pragma unroll
    for (int k = 0; k < FB; ++k) {
      cf G = {0.f, 0.f};
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        if constexpr (FULLDIR) {
          const float d = sinc_fast(p.kappa_per_hz * (float)(fstart + (double)k * p.df) * geo[j].sint);
          G.x = fmaf(g[j].x, d, G.x);
          G.y = fmaf(g[j].y, d, G.y);
        } else {
          G.x += g[j].x;
          G.y += g[j].y;
        }
        g[j] = cmul(g[j], q[j]);
      }
      cmac(X[k], G, qv[k]);
    }
    }  // sub-element groups
  }
  if (!valid) return;
  cf* out = part + ((size_t)(chunk * 4 + wave) * nbuf + blockIdx.y * FB) * p.E + e;
#pragma unroll
  for (int k = 0; k < FB; ++k) out[(size_t)k * p.E] = X[k];
}

// spect[tx][k][e] = corr * probe(k) * sum_w part[w][k][e]   (k < NF)
// spect[tx][k][e] = corr probe_k sum_w part[w][k - k_begin][e]. Block (64, RG): 64 consecutive
// elements of one frequency; thread row y sums parts y, y + RG, ... and the RG row sums are added
// in a fixed order, so the result does not depend on scheduling.
constexpr int RG = 4;

__global__ __launch_bounds__(64 * RG) void k_reduce(Problem p, int nbuf, int nparts, int k_begin,
                                                    const cf* __restrict__ part, const float* __restrict__ pr,
                                                    int tx, cf* __restrict__ spect) {
  __shared__ float2 rows[RG][64];
  const int e = blockIdx.x * 64 + threadIdx.x;
  const int k = k_begin + blockIdx.y;
  const bool ok = e < p.E && k < p.NF;
  float sx = 0.f, sy = 0.f;
  if (ok) {
    for (int w = threadIdx.y; w < nparts; w += RG) {
      const cf v = part[((size_t)w * nbuf + (k - k_begin)) * p.E + e];
      sx += v.x;
      sy += v.y;
    }
  }
  rows[threadIdx.y][threadIdx.x] = make_float2(sx, sy);
  __syncthreads();
  if (threadIdx.y == 0 && ok) {
    float tx_ = 0.f, ty_ = 0.f;
#pragma unroll
    for (int r = 0; r < RG; ++r) {
      tx_ += rows[r][threadIdx.x].x;
      ty_ += rows[r][threadIdx.x].y;
    }
    const float sc = p.corr * pr[k];
    spect[((size_t)tx * p.NF + k) * p.E + e] = {tx_ * sc, ty_ * sc};
  }
}

// ----------------------------------------------------------------------------- host side
struct Fp32Plan {
  int nfp, nfb, fbg, chunks, s_per_wave, nparts;
  bool table_per_tx;  // the TX table holds one transmit, rebuilt before each (else all transmits at once)
  bool geo;           // k_rx reads the geometry table (built once per call, shared by all transmits)
  size_t off_A, off_pp, off_pr, off_Q, off_part, off_scat, off_rt, off_la, off_qt, total;
};

bool plain_geo_table_possible(const Problem& p);  // simus_mfma.hip
void launch_plain_geo_table(const Problem& p, float4* scat, double* rt, float* la, float2* qt, hipStream_t st);

// Geometry table of k_rx: (S + 1) x E x n_sub terms of 24 B (plus 16 B per scatterer of inputs), used by
// the default plan up to this size.
constexpr size_t kPlainGeoTableMax = (size_t)16 << 30;
static size_t plain_geo_bytes(const Problem& p) { return (size_t)(p.S + 1) * p.E * p.NS * 24 + (size_t)p.S * 16; }

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// Enough waves per RX launch to fill 256 CUs about four times over at 4 waves per SIMD: a finer
// split shortens the tail of the last round (50 010 scatterers: 9.40 ms with 8192 waves, 9.05 ms
// with 16384, no further gain beyond); the partial sums grow with it (~265 MB there).
static int default_chunks(const Problem& p, int fbg) {
  const int etiles = (p.E + 63) / 64;
  const long want_waves = 16384;
  const long waves_per_chunk = (long)etiles * fbg * 4;
  const int chunks = (int)std::max<long>(1, want_waves / std::max<long>(waves_per_chunk, 1));
  const int max_chunks = std::max(1, (p.S + 4 * 16 - 1) / (4 * 16));  // at least 16 scatterers per wave
  return std::min(chunks, max_chunks);
}

static void fp32_layout(const Problem& p, int fbg, int chunks, bool table_per_tx, bool geo, Fp32Plan& P) {
  P.fbg = fbg;
  P.chunks = chunks;
  P.table_per_tx = table_per_tx;
  P.geo = geo;
  P.s_per_wave = (p.S + chunks * 4 - 1) / (chunks * 4);
  P.nparts = chunks * 4;
  const size_t per_block = (size_t)p.S * FB * sizeof(cf);
  size_t off = 0;
  P.off_A = off;    off = align_up(off + (size_t)(table_per_tx ? 1 : p.NT) * p.E * P.nfp * sizeof(cf));
  P.off_pp = off;   off = align_up(off + (size_t)P.nfp * sizeof(cf));
  P.off_pr = off;   off = align_up(off + (size_t)P.nfp * sizeof(float));
  P.off_Q = off;    off = align_up(off + per_block * P.fbg);
  P.off_part = off; off = align_up(off + (size_t)P.nparts * P.fbg * FB * p.E * sizeof(cf));
  P.off_scat = P.off_rt = P.off_la = P.off_qt = off;
  if (geo) {
    const size_t terms = (size_t)(p.S + 1) * p.E * p.NS;
    P.off_scat = off; off = align_up(off + (size_t)p.S * sizeof(float4));
    P.off_rt = off;   off = align_up(off + terms * sizeof(double));
    P.off_la = off;   off = align_up(off + terms * 2 * sizeof(float));
    P.off_qt = off;   off = align_up(off + terms * sizeof(float2));
  }
  P.total = off;
}

Fp32Plan fp32_plan(const Problem& p) {
  Fp32Plan P{};
  P.nfb = (p.NF + FB - 1) / FB;
  P.nfp = P.nfb * FB;
  // Q buffer holds fbg frequency blocks for all scatterers: cap it near 2 GiB (of 288 GB of HBM), so
  // that large problems still get several blocks per group and few scatterer partitions to reduce.
  const size_t per_block = (size_t)p.S * FB * sizeof(cf);
  const int fbg =
      (int)std::max<size_t>(1, std::min<size_t>((size_t)P.nfb, ((size_t)2 << 30) / std::max<size_t>(per_block, 1)));
  const bool geo = plain_geo_table_possible(p) && plain_geo_bytes(p) <= kPlainGeoTableMax;
  fp32_layout(p, fbg, default_chunks(p, fbg), false, geo, P);
  return P;
}

// The same kernels within a scratch of at most `budget` bytes: the TX table holds one transmit,
// then the largest frequency-block group (at most the default) whose default scatterer split
// fits, else one block and as many scatterer chunks as fit. Returns false when even one block
// with one chunk does not fit.
bool fp32_plan_within(const Problem& p, size_t budget, Fp32Plan* out) {
  Fp32Plan P{};
  P.nfb = (p.NF + FB - 1) / FB;
  P.nfp = P.nfb * FB;
  const Fp32Plan def = fp32_plan(p);
  for (int fbg = def.fbg; fbg >= 1; --fbg) {
    fp32_layout(p, fbg, default_chunks(p, fbg), true, false, P);
    if (P.total <= budget) {
      *out = P;
      return true;
    }
  }
  fp32_layout(p, 1, 1, true, false, P);
  if (P.total > budget) return false;
  const size_t per_chunk = (size_t)4 * FB * p.E * sizeof(cf);
  const int extra = (int)std::min<size_t>((budget - P.total) / per_chunk, (size_t)default_chunks(p, 1) - 1);
  fp32_layout(p, 1, 1 + extra, true, false, P);
  while (P.total > budget && P.chunks > 1) fp32_layout(p, 1, P.chunks - 1, true, false, P);  // alignment slack
  *out = P;
  return P.total <= budget;
}

// Two 32-frequency blocks per k_tx thread share its geometry but halve the grid and take about
// 64 more VGPRs (3 instead of 5 waves per SIMD): worth it where the halved grid still gives about
// 8 workgroups per CU of a 256-CU device (50 000 scatterers x 55 blocks: 3 % faster; 2 000
// scatterers x 52 blocks: 33 % slower, so one block there).
constexpr long kTwoBlockMinGroups = 2048;
static bool two_blocks_per_thread(int sx, int nb) { return (long)sx * ((nb + 1) / 2) >= kTwoBlockMinGroups; }

template <int NSUB, bool FD, bool GR>
static void launch_tx(const Problem& p, int tx, const Fp32Plan& P, int fb0, int nb, const cf* A, const cf* pp, cf* Q,
                      hipStream_t st) {
  const int sx = (p.S + 255) / 256;
  const int slot = timing_start(st);
  if constexpr (!FD && !GR) {
    if (two_blocks_per_thread(sx, nb)) {
      hipLaunchKernelGGL((k_tx<NSUB, false, 0, 2>), dim3(sx, (nb + 1) / 2), dim3(256), 0, st, p, tx, P.nfp, fb0, nb, A,
                         pp, Q, (float*)nullptr);
      timing_stop(slot, st);
      return;
    }
  }
  hipLaunchKernelGGL((k_tx<NSUB, FD, 0, 1, GR>), dim3(sx, nb), dim3(256), 0, st, p, tx, P.nfp, fb0, nb, A, pp, Q,
                     (float*)nullptr);
  timing_stop(slot, st);
}
template <int NSUB, bool FD, bool GR>
static void launch_rx(const Problem& p, const Fp32Plan& P, int fb0, int nb, const cf* Q, cf* part, const char* ws,
                      hipStream_t st) {
  dim3 grid((p.E + 63) / 64, nb, P.chunks);
  const int slot = timing_start(st);
  if constexpr (!FD && !GR) {
    if (P.geo) {
      hipLaunchKernelGGL((k_rx<NSUB, false, false, true>), grid, dim3(256), 0, st, p, P.fbg * FB, fb0, P.s_per_wave,
                         Q, part, (const double*)(ws + P.off_rt), (const float*)(ws + P.off_la),
                         (const float2*)(ws + P.off_qt));
      timing_stop(slot, st);
      return;
    }
  }
  hipLaunchKernelGGL((k_rx<NSUB, FD, GR>), grid, dim3(256), 0, st, p, P.fbg * FB, fb0, P.s_per_wave, Q, part,
                     (const double*)nullptr, (const float*)nullptr, (const float2*)nullptr);
  timing_stop(slot, st);
}

// NSUB sub-elements exactly, or (GR) any count in groups of NSUB
template <int NSUB, bool GR = false>
static int dispatch_fd(const Problem& p, int tx, const Fp32Plan& P, int fb0, int nb, const cf* A, const cf* pp, cf* Q,
                       cf* part, const char* ws, hipStream_t st) {
  if (p.fulldir) {
    launch_tx<NSUB, true, GR>(p, tx, P, fb0, nb, A, pp, Q, st);
    launch_rx<NSUB, true, GR>(p, P, fb0, nb, Q, part, ws, st);
  } else {
    launch_tx<NSUB, false, GR>(p, tx, P, fb0, nb, A, pp, Q, st);
    launch_rx<NSUB, false, GR>(p, P, fb0, nb, Q, part, ws, st);
  }
  return FS_OK;
}

static void launch_tx_table_n(const Problem& p, int nfp, cf* A, hipStream_t st) {
  const long n = (long)p.NT * p.E * nfp;
  const int blocks = (int)std::min<long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_tx_table, dim3(blocks), dim3(256), 0, st, p, nfp, A);
}

// All frequency-block groups of transmit `tx` of p (its TX table at A + tx E nfp), reduced into spect.
static int fp32_transmit(const Problem& p, const Fp32Plan& P, int tx, const cf* A, const cf* pp, const float* pr, cf* Q,
                         cf* part, cf* spect, const char* ws, hipStream_t st) {
  for (int fb0 = 0; fb0 < P.nfb; fb0 += P.fbg) {
    const int nb = std::min(P.fbg, P.nfb - fb0);
    int rc;
    switch (p.NS) {
      case 1: rc = dispatch_fd<1>(p, tx, P, fb0, nb, A, pp, Q, part, ws, st); break;
      case 2: rc = dispatch_fd<2>(p, tx, P, fb0, nb, A, pp, Q, part, ws, st); break;
      case 3: rc = dispatch_fd<3>(p, tx, P, fb0, nb, A, pp, Q, part, ws, st); break;
      case 4: rc = dispatch_fd<4>(p, tx, P, fb0, nb, A, pp, Q, part, ws, st); break;
      default: rc = dispatch_fd<4, true>(p, tx, P, fb0, nb, A, pp, Q, part, ws, st); break;
    }
    if (rc != FS_OK) return rc;
    const int kb = fb0 * FB, ke = std::min((fb0 + nb) * FB, p.NF);
    hipLaunchKernelGGL(k_reduce, dim3((p.E + 63) / 64, ke - kb), dim3(64, RG), 0, st, p, P.fbg * FB, P.nparts, kb,
                       part, pr, tx, spect);
  }
  return FS_OK;
}

int fp32_spectrum_plan(const Problem& p, const Fp32Plan& P, cf* spect, char* ws, hipStream_t st) {
  cf* A = (cf*)(ws + P.off_A);
  cf* pp = (cf*)(ws + P.off_pp);
  float* pr = (float*)(ws + P.off_pr);
  cf* Q = (cf*)(ws + P.off_Q);
  cf* part = (cf*)(ws + P.off_part);
  hipLaunchKernelGGL(k_spectra, dim3((P.nfp + 255) / 256), dim3(256), 0, st, p, P.nfp, pp, pr);
  if (P.geo)
    launch_plain_geo_table(p, (float4*)(ws + P.off_scat), (double*)(ws + P.off_rt), (float*)(ws + P.off_la),
                           (float2*)(ws + P.off_qt), st);
  if (!P.table_per_tx) {
    launch_tx_table_n(p, P.nfp, A, st);
    for (int tx = 0; tx < p.NT; ++tx) {
      const int rc = fp32_transmit(p, P, tx, A, pp, pr, Q, part, spect, ws, st);
      if (rc != FS_OK) return rc;
    }
  } else {
    // transmit t as a one-transmit problem: its delays and apodization, its output rows
    for (int tx = 0; tx < p.NT; ++tx) {
      Problem pt = p;
      pt.NT = 1;
      pt.delay = p.delay + (size_t)tx * p.E;
      pt.apod = p.apod + (size_t)tx * p.E;
      launch_tx_table_n(pt, P.nfp, A, st);
      const int rc = fp32_transmit(pt, P, 0, A, pp, pr, Q, part, spect + (size_t)tx * p.NF * p.E, ws, st);
      if (rc != FS_OK) return rc;
    }
  }
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

int fp32_spectrum(const Problem& p, cf* spect, char* ws, hipStream_t st) {
  return fp32_spectrum_plan(p, fp32_plan(p), spect, ws, st);
}

size_t fp32_workspace(const Problem& p) { return fp32_plan(p).total; }

// The plain method inside a fixed scratch of ws_bytes (no workspace query, no allocation):
// FS_ERR_SCRATCH when the problem does not fit even with one block and one scatterer chunk.
int fp32_spectrum_within(const Problem& p, cf* spect, char* ws, size_t ws_bytes, hipStream_t st) {
  Fp32Plan P;
  if (!fp32_plan_within(p, ws_bytes, &P)) return FS_ERR_SCRATCH;
  return fp32_spectrum_plan(p, P, spect, ws, st);
}

// ----------------------------------------------------------------------------- pfield
struct PfieldLayout {
  int nfb, nfp;
  size_t off_A, off_pp, off_pr, off_pw, total;
};

static PfieldLayout pfield_layout(const Problem& p) {
  PfieldLayout L{};
  L.nfb = (p.NF + FB - 1) / FB;
  L.nfp = L.nfb * FB;
  size_t off = 0;
  L.off_A = off;  off = align_up(off + (size_t)p.E * L.nfp * sizeof(cf));
  L.off_pp = off; off = align_up(off + (size_t)L.nfp * sizeof(cf));
  L.off_pr = off; off = align_up(off + (size_t)L.nfp * sizeof(float));
  L.off_pw = off; off = align_up(off + (size_t)L.nfb * std::max(p.S, 1) * sizeof(float));
  L.total = off;
  return L;
}

size_t pfield_workspace(const Problem& p) { return pfield_layout(p).total; }

template <int NSUB, bool GR = false>
static void launch_pfield(const Problem& p, const PfieldLayout& L, const cf* A, const cf* pp, float* pw,
                          hipStream_t st) {
  const int sx = (p.S + 255) / 256;
  if (p.fulldir)
    hipLaunchKernelGGL((k_tx<NSUB, true, 1, 1, GR>), dim3(sx, L.nfb), dim3(256), 0, st, p, 0, L.nfp, 0, L.nfb, A, pp,
                       (cf*)nullptr, pw);
  else if (GR)
    hipLaunchKernelGGL((k_tx<NSUB, false, 1, 1, GR>), dim3(sx, L.nfb), dim3(256), 0, st, p, 0, L.nfp, 0, L.nfb, A, pp,
                       (cf*)nullptr, pw);
  else if (two_blocks_per_thread(sx, L.nfb))
    hipLaunchKernelGGL((k_tx<NSUB, false, 1, 2>), dim3(sx, (L.nfb + 1) / 2), dim3(256), 0, st, p, 0, L.nfp, 0, L.nfb,
                       A, pp, (cf*)nullptr, pw);
  else
    hipLaunchKernelGGL((k_tx<NSUB, false, 1, 1>), dim3(sx, L.nfb), dim3(256), 0, st, p, 0, L.nfp, 0, L.nfb, A, pp,
                       (cf*)nullptr, pw);
}

int pfield_fp32(const Problem& p, float* out, char* ws, hipStream_t st) {
  const PfieldLayout L = pfield_layout(p);
  cf* A = (cf*)(ws + L.off_A);
  cf* pp = (cf*)(ws + L.off_pp);
  float* pr = (float*)(ws + L.off_pr);
  float* pw = (float*)(ws + L.off_pw);
  Problem q = p;
  q.NT = 1;
  {
    const long n = (long)q.E * L.nfp;
    hipLaunchKernelGGL(k_tx_table, dim3((int)std::min<long>((n + 255) / 256, 65536)), dim3(256), 0, st, q, L.nfp, A);
    hipLaunchKernelGGL(k_spectra, dim3((L.nfp + 255) / 256), dim3(256), 0, st, q, L.nfp, pp, pr);
  }
  switch (q.NS) {
    case 1: launch_pfield<1>(q, L, A, pp, pw, st); break;
    case 2: launch_pfield<2>(q, L, A, pp, pw, st); break;
    case 3: launch_pfield<3>(q, L, A, pp, pw, st); break;
    case 4: launch_pfield<4>(q, L, A, pp, pw, st); break;
    default: launch_pfield<4, true>(q, L, A, pp, pw, st); break;
  }
  hipLaunchKernelGGL(k_pfield_reduce, dim3((q.S + 255) / 256), dim3(256), 0, st, pw, L.nfb, q.S, q.corr, out);
  return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_LAUNCH;
}

}  // namespace fs
