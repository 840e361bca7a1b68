// C ABI entry points (declared in include/fastsimus.h): argument checks,
// problem set-up and dispatch to the kernels of this library.
#include <math.h>

#include <algorithm>
#include <mutex>

#include "fs_device.h"

namespace fs {
struct Fp32Plan;
int fp32_spectrum(const Problem& p, cf* spect, char* ws, hipStream_t st);
size_t fp32_workspace(const Problem& p);
int fp32_spectrum_within(const Problem& p, cf* spect, char* ws, size_t ws_bytes, hipStream_t st);
int mfma_spectrum(const Problem& p, cf* spect, char* ws, size_t ws_bytes, hipStream_t st);
size_t mfma_workspace(const Problem& p);
int mfma_spectrum_batch(const Problem& p, int P, const int64_t* off, cf* spect, char* ws, size_t ws_bytes,
                        hipStream_t st);
size_t mfma_workspace_batch(const Problem& p, int P, const int64_t* off);
fs_tuning msm_tuning(const Problem& p, int P);
int msm_tuning_set(const Problem& p, int P, const fs_tuning* t);
size_t rf_workspace(int NT, int NF, int N);
int spectrum_to_rf(const cf* spect, int NT, int NF, int E, int nff, int idx0, int N, cf* full, float* rf, float* rf_u,
                   char* ws, hipStream_t st);
size_t rf_adjoint_ws(int NT, int nkeep, int E);
int rf_adjoint(const float* u, const float* dy, int NT, int NF, int E, int idx0, int N, cf* dspect, char* ws,
               hipStream_t st);
int adjoint_fp32(const Problem& p, int method, const cf* dspect, float* d_rc, float* d_x, float* d_z, double* d_c,
                 char* ws, hipStream_t st);
size_t adjoint_workspace(const Problem& p, int method);
int pfield_fp32(const Problem& p, float* out, char* ws, hipStream_t st);
size_t pfield_workspace(const Problem& p);
int fill_zero(void* ptr, size_t n_words, hipStream_t st);
}  // namespace fs

using namespace fs;

// Sub-elements per element: up to 4 in registers, more in groups of 4 (plain and vector-adjoint
// kernels) or one at a time (matrix-core kernel); this bound only rejects absurd descriptors.
constexpr int kMaxSubElements = 1024;

// ---------------------------------------------------------------- static scratch of fs_simus_spectrum_static
// One array per GPU (each device holds its own copy at its own address). In host code `scratch` is
// only a stand-in for the symbol: its device address is looked up once per device with
// hipGetSymbolAddress and that pointer is what the kernels receive.
static __device__ __align__(256) char scratch[FS_STATIC_SCRATCH_BYTES];

constexpr int kMaxScratchDevices = 64;
static std::mutex scratch_mutex;
static char* scratch_ptr[kMaxScratchDevices] = {};

static int static_scratch(char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return FS_ERR_LAUNCH;
  if (dev < 0 || dev >= kMaxScratchDevices) return FS_ERR_RESOURCE;
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FS_ERR_LAUNCH;
    scratch_ptr[dev] = (char*)ptr;
  }
  *out = scratch_ptr[dev];
  return FS_OK;
}

static int make_problem(const fs_simus_desc* d, Problem* p, bool need_rc, bool need_tx) {
  if (!d) return FS_ERR_INVALID;
  if (d->n_scat < 0 || d->n_elem <= 0 || d->n_sub <= 0 || d->n_freq <= 0 || d->n_tx <= 0) return FS_ERR_INVALID;
  if (d->n_sub > kMaxSubElements) return FS_ERR_UNSUPPORTED;
  if (d->n_scat > 0 && (!d->scat_x || !d->scat_z || (need_rc && !d->rc))) return FS_ERR_INVALID;
  if (!d->sub_x || !d->sub_z || !d->elem_cos || !d->elem_sin) return FS_ERR_INVALID;
  if (need_tx && (!d->tx_delay || !d->tx_apod)) return FS_ERR_INVALID;
  if (!d->pulse_re || !d->pulse_im || !d->probe) return FS_ERR_INVALID;
  if (!(d->speed_of_sound > 0) || !(d->freq_center > 0) || !(d->seg_length > 0) || d->attenuation < 0)
    return FS_ERR_INVALID;
  if (d->baffle_kind < 0 || d->baffle_kind > 2) return FS_ERR_INVALID;
  Problem& q = *p;
  q.S = d->n_scat;
  q.E = d->n_elem;
  q.NS = d->n_sub;
  q.NF = d->n_freq;
  q.NT = d->n_tx;
  q.fulldir = d->full_frequency_directivity ? 1 : 0;
  q.baffle = d->baffle_kind;
  q.f0 = d->f0;
  q.df = d->df;
  q.inv_c = 1.0 / d->speed_of_sound;
  q.min_dist = d->speed_of_sound / d->freq_center / 2.0;
  q.att_coef = (float)(d->attenuation / kNeperToDb * 1e-4);
  q.kappa = (float)(M_PI * d->freq_center * d->seg_length / d->speed_of_sound);
  q.kappa_per_hz = (float)(M_PI * d->seg_length / d->speed_of_sound);
  q.baffle_b = (float)d->baffle_impedance;
  q.inv_nsub = 1.0f / (float)d->n_sub;
  q.radius = d->radius;
  q.apex = d->apex_offset;
  q.corr = (float)d->correction;
  q.x = d->scat_x;
  q.z = d->scat_z;
  q.rc = d->rc;
  q.sub_x = d->sub_x;
  q.sub_z = d->sub_z;
  q.ecos = d->elem_cos;
  q.esin = d->elem_sin;
  q.delay = d->tx_delay;
  q.apod = d->tx_apod;
  q.pulse_re = d->pulse_re;
  q.pulse_im = d->pulse_im;
  q.probe = d->probe;
  return FS_OK;
}

extern "C" {

int fs_version(void) { return 100; }

const char* fs_status_string(int s) {
  switch (s) {
    case FS_OK: return "ok";
    case FS_ERR_INVALID: return "invalid argument";
    case FS_ERR_UNSUPPORTED: return "unsupported configuration";
    case FS_ERR_WORKSPACE: return "workspace too small";
    case FS_ERR_LAUNCH: return "HIP launch or runtime error";
    case FS_ERR_RESOURCE: return "problem exceeds a kernel resource limit";
    case FS_ERR_SCRATCH: return "problem exceeds the static scratch of fs_simus_spectrum_static (64 MiB)";
    default: return "unknown status";
  }
}

size_t fs_simus_workspace_bytes(const fs_simus_desc* d, int method) {
  Problem p;
  if (make_problem(d, &p, true, true) != FS_OK) return 0;
  if (method == FS_METHOD_FP32) return fp32_workspace(p);
  if (method == FS_METHOD_MFMA_BF16X3) return mfma_workspace(p);
  return 0;
}

int fs_simus_spectrum(const fs_simus_desc* d, int method, float* spect, void* workspace, size_t workspace_bytes,
                      void* stream) {
  Problem p;
  int rc = make_problem(d, &p, true, true);
  if (rc != FS_OK) return rc;
  if (!spect) return FS_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  if (p.S == 0) {
    return fill_zero(spect, 2 * (size_t)p.NT * p.NF * p.E, st);
  }
  if (method == FS_METHOD_FP32) {
    if (!workspace || workspace_bytes < fp32_workspace(p)) return FS_ERR_WORKSPACE;
    return fp32_spectrum(p, (cf*)spect, (char*)workspace, st);
  }
  if (method == FS_METHOD_MFMA_BF16X3) {
    if (!workspace || workspace_bytes < mfma_workspace(p)) return FS_ERR_WORKSPACE;
    return mfma_spectrum(p, (cf*)spect, (char*)workspace, workspace_bytes, st);
  }
  return FS_ERR_INVALID;
}

int fs_simus_spectrum_static(const fs_simus_desc* d, float* spect, void* stream) {
  Problem p;
  int rc = make_problem(d, &p, true, true);
  if (rc != FS_OK) return rc;
  if (!spect) return FS_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  if (p.S == 0) return fill_zero(spect, 2 * (size_t)p.NT * p.NF * p.E, st);
  char* ws = nullptr;
  rc = static_scratch(&ws);
  if (rc != FS_OK) return rc;
  return fp32_spectrum_within(p, (cf*)spect, ws, FS_STATIC_SCRATCH_BYTES, st);
}

static bool valid_offsets(const Problem& p, int P, const int64_t* off) {
  if (P < 1 || P > 192 || !off || off[0] != 0 || off[P] != p.S) return false;
  for (int q = 0; q < P; ++q)
    if (off[q + 1] < off[q]) return false;
  return true;
}

// Phantom q of a batch as its own problem (scatterer pointers offset into the concatenation).
static Problem phantom_problem(const Problem& p, const int64_t* off, int q) {
  Problem s = p;
  s.S = (int)(off[q + 1] - off[q]);
  s.x = p.x + off[q];
  s.z = p.z + off[q];
  s.rc = p.rc + off[q];
  return s;
}

size_t fs_simus_batch_workspace_bytes(const fs_simus_desc* d, int method, int n_phantoms,
                                      const int64_t* phantom_offsets) {
  Problem p;
  if (make_problem(d, &p, true, true) != FS_OK || !valid_offsets(p, n_phantoms, phantom_offsets)) return 0;
  if (method == FS_METHOD_FP32) {
    size_t m = 256;
    for (int q = 0; q < n_phantoms; ++q) {
      const Problem s = phantom_problem(p, phantom_offsets, q);
      if (s.S > 0) m = std::max(m, fp32_workspace(s));
    }
    return m;
  }
  if (method == FS_METHOD_MFMA_BF16X3) return mfma_workspace_batch(p, n_phantoms, phantom_offsets);
  return 0;
}

int fs_simus_spectrum_batch(const fs_simus_desc* d, int method, int n_phantoms, const int64_t* phantom_offsets,
                            float* spect, void* workspace, size_t workspace_bytes, void* stream) {
  Problem p;
  int rc = make_problem(d, &p, true, true);
  if (rc != FS_OK) return rc;
  if (!spect) return FS_ERR_INVALID;
  if (n_phantoms > 192) return FS_ERR_UNSUPPORTED;
  if (!valid_offsets(p, n_phantoms, phantom_offsets)) return FS_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  if (method == FS_METHOD_FP32) {
    const size_t per = 2 * (size_t)p.NT * p.NF * p.E;
    for (int q = 0; q < n_phantoms; ++q) {
      const Problem s = phantom_problem(p, phantom_offsets, q);
      float* out = spect + q * per;
      if (s.S == 0) {
        rc = fill_zero(out, per, st);
      } else {
        if (!workspace || workspace_bytes < fp32_workspace(s)) return FS_ERR_WORKSPACE;
        rc = fp32_spectrum(s, (cf*)out, (char*)workspace, st);
      }
      if (rc != FS_OK) return rc;
    }
    return FS_OK;
  }
  if (method == FS_METHOD_MFMA_BF16X3) {
    if (!workspace || workspace_bytes < mfma_workspace_batch(p, n_phantoms, phantom_offsets))
      return FS_ERR_WORKSPACE;
    return mfma_spectrum_batch(p, n_phantoms, phantom_offsets, (cf*)spect, (char*)workspace, workspace_bytes, st);
  }
  return FS_ERR_INVALID;
}

int fs_tuning_set(const fs_simus_desc* d, int n_phantoms, const fs_tuning* t) {
  Problem p;
  int rc = make_problem(d, &p, false, false);
  if (rc != FS_OK) return rc;
  if (n_phantoms < 1 || n_phantoms > 192) return FS_ERR_INVALID;
  return msm_tuning_set(p, n_phantoms, t);
}

int fs_tuning_get(const fs_simus_desc* d, int n_phantoms, fs_tuning* t) {
  Problem p;
  int rc = make_problem(d, &p, false, false);
  if (rc != FS_OK) return rc;
  if (!t || n_phantoms < 1 || n_phantoms > 192) return FS_ERR_INVALID;
  *t = msm_tuning(p, n_phantoms);
  return FS_OK;
}

size_t fs_rf_workspace_bytes(int n_tx, int n_freq, int n_fft) {
  return (n_tx > 0 && n_freq > 0 && n_fft > 0) ? rf_workspace(n_tx, n_freq, n_fft) : 0;
}

int fs_spectrum_to_rf(const float* spect, int n_tx, int n_freq, int n_elem, int n_ff, int idx0, int n_fft,
                      float* spectrum_full, float* rf, float* rf_ungated, void* workspace, size_t workspace_bytes,
                      void* stream) {
  if (!spect || !rf || n_tx <= 0 || n_freq <= 0 || n_elem <= 0 || n_fft <= 0 || idx0 < 0) return FS_ERR_INVALID;
  if (spectrum_full && (n_ff <= 0 || idx0 + n_freq > n_ff)) return FS_ERR_INVALID;
  if (!workspace || workspace_bytes < fs_rf_workspace_bytes(n_tx, n_freq, n_fft)) return FS_ERR_WORKSPACE;
  return spectrum_to_rf((const cf*)spect, n_tx, n_freq, n_elem, n_ff, idx0, n_fft, (cf*)spectrum_full, rf, rf_ungated,
                        (char*)workspace, (hipStream_t)stream);
}

size_t fs_rf_adjoint_workspace_bytes(int n_tx, int n_keep, int n_elem) {
  if (n_tx <= 0 || n_keep <= 0 || n_elem <= 0) return 0;
  return rf_adjoint_ws(n_tx, n_keep, n_elem);
}

int fs_rf_adjoint(const float* rf_ungated, const float* d_rf, int n_tx, int n_freq, int n_elem, int idx0, int n_fft,
                  float* d_spect, void* workspace, size_t workspace_bytes, void* stream) {
  if (!rf_ungated || !d_rf || !d_spect || n_tx <= 0 || n_freq <= 0 || n_elem <= 0 || n_fft <= 0 || idx0 < 0)
    return FS_ERR_INVALID;
  const int nkeep = (n_fft + 1) / 2;
  if (!workspace || workspace_bytes < rf_adjoint_ws(n_tx, nkeep, n_elem)) return FS_ERR_WORKSPACE;
  return rf_adjoint(rf_ungated, d_rf, n_tx, n_freq, n_elem, idx0, n_fft, (cf*)d_spect, (char*)workspace,
                    (hipStream_t)stream);
}

static bool valid_adjoint_method(int m) { return m == FS_ADJOINT_AUTO || m == FS_ADJOINT_MATRIX || m == FS_ADJOINT_VECTOR; }

size_t fs_simus_adjoint_workspace_bytes(const fs_simus_desc* d, int adjoint_method) {
  Problem p;
  if (!valid_adjoint_method(adjoint_method) || make_problem(d, &p, true, true) != FS_OK) return 0;
  return adjoint_workspace(p, adjoint_method);
}

int fs_simus_adjoint(const fs_simus_desc* d, int adjoint_method, const float* d_spect, float* d_rc, float* d_x,
                     float* d_z, double* d_c, void* workspace, size_t workspace_bytes, void* stream) {
  Problem p;
  int rc = make_problem(d, &p, true, true);
  if (rc != FS_OK) return rc;
  if (!valid_adjoint_method(adjoint_method)) return FS_ERR_INVALID;
  // without scatterers the per-scatterer outputs are empty (their pointers may be null) and only d_c is written
  if (!d_c || (p.S > 0 && (!d_spect || !d_rc || !d_x || !d_z))) return FS_ERR_INVALID;
  if (p.S > 0 && (!workspace || workspace_bytes < adjoint_workspace(p, adjoint_method))) return FS_ERR_WORKSPACE;
  return adjoint_fp32(p, adjoint_method, (const cf*)d_spect, d_rc, d_x, d_z, d_c, (char*)workspace,
                      (hipStream_t)stream);
}

size_t fs_pfield_workspace_bytes(const fs_simus_desc* d) {
  Problem p;
  if (make_problem(d, &p, false, true) != FS_OK) return 0;
  return pfield_workspace(p);
}

int fs_pfield(const fs_simus_desc* d, float* out, void* workspace, size_t workspace_bytes, void* stream) {
  Problem p;
  int rc = make_problem(d, &p, false, true);
  if (rc != FS_OK) return rc;
  if (!out || p.NT != 1) return FS_ERR_INVALID;
  if (!workspace || workspace_bytes < pfield_workspace(p)) return FS_ERR_WORKSPACE;
  if (p.S == 0) return FS_OK;
  return pfield_fp32(p, out, (char*)workspace, (hipStream_t)stream);
}

}  // extern "C"
