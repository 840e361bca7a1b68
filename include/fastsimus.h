/*
 * fastsimus.h - C ABI of the MI355X (gfx950) SIMUS hot path.
 *
 * Every pointer is a device pointer owned by the caller; every call is
 * stream-ordered on the given hipStream_t (passed as void* so that the header
 * needs no HIP include). Inputs are never written. Outputs are fully
 * overwritten. Functions return FS_OK (0) or a status code; fs_status_string
 * names it. The library allocates no device memory: scratch is the caller's
 * `workspace` of at least fs_simus_workspace_bytes() bytes.
 *
 * Reference interfaces replaced (charlesbmi/FastSIMUS):
 *   fs_simus_spectrum     <- simus_cuda(scatterers, rc, params, plan, medium,
 *                            delays_clean, tx_apodization)
 *                            src/fast_simus/kernels/cuda_simus.py:232-321
 *                            (kernel simus_fused_kernel, simus_fused.cu:62-332),
 *                            times plan.correction_factor (simus.py:490-491)
 *   fs_spectrum_to_rf     <- _irfft_and_threshold, src/fast_simus/simus.py:336-367
 *   fs_simus_adjoint      <- no kernel in the reference; differentiates the
 *                            algebra of _simus_freq_outer_scan
 *                            (src/fast_simus/_simus_strategies.py:152-210);
 *                            its fs_adjoint_method argument selects its kernels
 *   fs_rf_adjoint         <- reverse of _irfft_and_threshold (simus.py:336-367)
 *   fs_pfield             <- pfield_metal / _freq_outer_python
 *                            (src/fast_simus/kernels/metal_pfield.py:89-219,
 *                             src/fast_simus/_pfield_strategies.py:53-88)
 *   fs_simus_spectrum_batch <- a loop of simus_cuda calls over phantoms (no batched API in the
 *                            reference; its harness loops, tests/_picmus_phantom_shared.py:319-367)
 *   fs_tuning_set / get   <- the fixed tuning constants of kernels/cuda_simus.py:39-45
 *   fs_max_distance       <- the O(points x sources) distance maximum of the plans,
 *                            src/fast_simus/simus.py:219-222 and pfield.py:271-282
 *   fs_rf2iq, fs_das      <- the consumers of the multi-firing RF stack in the reference's harness
 *                            (pymust.rf2iq and pymust.dasmtx per firing, summed over firings,
 *                            tests/_picmus_phantom_shared.py:407-460)
 */
#ifndef FASTSIMUS_H
#define FASTSIMUS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fs_status {
  FS_OK = 0,
  FS_ERR_INVALID = 1,      /* bad sizes / null pointers            -> ValueError   */
  FS_ERR_UNSUPPORTED = 2,  /* option not implemented by this method -> NotImplementedError */
  FS_ERR_WORKSPACE = 3,    /* workspace smaller than required       -> RuntimeError */
  FS_ERR_LAUNCH = 4,       /* HIP launch / runtime error            -> RuntimeError */
  FS_ERR_RESOURCE = 5,     /* size beyond a kernel resource limit   -> RuntimeError */
  FS_ERR_SCRATCH = 6       /* problem larger than the static scratch of fs_simus_spectrum_static -> RuntimeError */
};

enum fs_method {
  FS_METHOD_FP32 = 0,      /* per-transmit TX/RX contractions on FP32 VALU, FP64 reseed every 32 steps */
  FS_METHOD_MFMA_BF16X3 = 1 /* all transmits through one multistatic matrix per frequency, split-BF16 MFMA */
};

enum fs_baffle { FS_BAFFLE_SOFT = 0, FS_BAFFLE_RIGID = 1, FS_BAFFLE_IMPEDANCE = 2 };

/* One SIMUS problem: scatterers x probe x transmits x selected band. */
typedef struct fs_simus_desc {
  int32_t n_scat;   /* S  */
  int32_t n_elem;   /* E  */
  int32_t n_sub;    /* sub-elements per element */
  int32_t n_freq;   /* selected frequencies */
  int32_t n_tx;     /* transmits */
  int32_t full_frequency_directivity; /* 0: centre-frequency sinc, 1: sinc at every frequency */
  int32_t baffle_kind;                /* fs_baffle */
  int32_t pad0;
  /* scatterers, float32 [S] */
  const float* scat_x;
  const float* scat_z;
  const float* rc;
  /* sub-element centroids, float64 [E * n_sub] (element-major), element angle cos/sin [E] */
  const double* sub_x;
  const double* sub_z;
  const double* elem_cos;
  const double* elem_sin;
  /* transmits, float64 [n_tx * E], NaN already replaced (delay 0, apodization 0) */
  const double* tx_delay;
  const double* tx_apod;
  /* spectra at the selected frequencies, float64 [n_freq] */
  const double* pulse_re;
  const double* pulse_im;
  const double* probe;
  /* scalars */
  double f0;                /* first selected frequency (Hz) */
  double df;                /* frequency step (Hz) */
  double speed_of_sound;    /* m/s */
  double attenuation;       /* dB/cm/MHz */
  double freq_center;       /* Hz */
  double seg_length;        /* sub-element length (m) */
  double radius;            /* convex radius (m), +inf for linear */
  double apex_offset;       /* m */
  double baffle_impedance;  /* used when baffle_kind == FS_BAFFLE_IMPEDANCE */
  double correction;        /* plan correction factor (df * element width) */
} fs_simus_desc;

/* Library version (major*10000 + minor*100 + patch). */
int fs_version(void);
const char* fs_status_string(int status);

/* Bytes of workspace the method needs for this problem. */
size_t fs_simus_workspace_bytes(const fs_simus_desc* d, int method);

/* Selected-band spectrum of every transmit, already multiplied by d->correction:
 *   spect[t][k][e] (interleaved re,im float32), t < n_tx, k < n_freq, e < n_elem.
 * Replaces simus_cuda(...) * plan.correction_factor. */
int fs_simus_spectrum(const fs_simus_desc* d, int method, float* spect, void* workspace, size_t workspace_bytes,
                      void* stream);

/* Allocation-free variant of fs_simus_spectrum with FS_METHOD_FP32. No counterpart in the reference:
 * it is for callers that cannot allocate device memory or pass a workspace (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation). Its scratch is one static device
 * array of FS_STATIC_SCRATCH_BYTES per GPU, held from the library's load until the process ends;
 * the array's address on the current device is looked up on the first call there (make that first
 * call outside graph capture). The same kernels as fs_simus_spectrum run within it (the TX table is
 * rebuilt per transmit; fewer frequency blocks and scatterer chunks per launch where the default
 * plan does not fit), so results equal fs_simus_spectrum's up to FP32 summation order. Problems
 * that do not fit even at one 32-frequency block and one scatterer chunk (about 250 000 scatterers
 * at 128 elements and 1 900 frequencies) return FS_ERR_SCRATCH before any work. Every call on one
 * GPU shares that GPU's scratch: calls that overlap on one GPU (other streams or threads) race. */
#define FS_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
int fs_simus_spectrum_static(const fs_simus_desc* d, float* spect, void* stream);

/* Batch of phantoms sharing the probe, the transmits and the frequency plan of d: the scatterers of
 * all phantoms are concatenated in d->scat_x / scat_z / rc (d->n_scat in total) and phantom q is
 * [phantom_offsets[q], phantom_offsets[q + 1]) (host array of n_phantoms + 1 int64, starting at 0
 * and ending at d->n_scat; 1 <= n_phantoms <= 192). Writes spect[q][t][k][e] (interleaved complex
 * float32), each phantom exactly as fs_simus_spectrum would on its own scatterers. With
 * FS_METHOD_MFMA_BF16X3 all phantoms run in one persistent launch (largest work first) and the call
 * only enqueues kernels, so it can be captured in a graph. No counterpart in the reference, whose
 * harness calls simus_cuda once per phantom and firing (tests/_picmus_phantom_shared.py:319-367). */
size_t fs_simus_batch_workspace_bytes(const fs_simus_desc* d, int method, int n_phantoms,
                                      const int64_t* phantom_offsets);
int fs_simus_spectrum_batch(const fs_simus_desc* d, int method, int n_phantoms, const int64_t* phantom_offsets,
                            float* spect, void* workspace, size_t workspace_bytes, void* stream);

/* Tile-shape choices of the batched method's multistatic-matrix kernel for one problem signature
 * (device, element count class, sub-elements, directivity/baffle variant, frequencies, phantom count,
 * log2 of the scatterer tiles). Zero fields mean "default". Stored in a mutex-guarded process-wide
 * cache that fs_simus_spectrum / fs_simus_spectrum_batch consult; set by an autotuner
 * (fastsimus_amd.tune). No counterpart in the reference (its CUDA constants are fixed,
 * src/fast_simus/kernels/cuda_simus.py:39-45). */
typedef struct fs_tuning {
  int32_t msm_fbk;       /* frequencies per block: 0 (default 4), 2 or 4; 2 is forced where registers need it */
  int32_t msm_units;     /* single phantom: scatterer chunks (0 = chosen from the CU count) */
  int32_t msm_item_div;  /* batches: largest work item <= 1/msm_item_div of a queue slot's share (0 = 8) */
  int32_t msm_table;     /* geometry table (per scatterer, element, sub-element: FP64 distance, log2 amplitude,
                            frequency step; 24 B each, in the workspace): 0 = when possible and <= 32 GiB,
                            1 = when possible, 2 = never (geometry evaluated per frequency block) */
  int32_t reserved[4];
} fs_tuning;
/* t == NULL forgets the entry. n_phantoms >= 1 (1 for fs_simus_spectrum). */
int fs_tuning_set(const fs_simus_desc* d, int n_phantoms, const fs_tuning* t);
/* Entry the next call with this signature will use (zeros if none). */
int fs_tuning_get(const fs_simus_desc* d, int n_phantoms, fs_tuning* t);

/* Place the band into the full spectrum and compute gated RF per transmit:
 *   spectrum_full[t][i][e] (complex64, i < n_ff), zero outside [idx0, idx0 + n_freq);
 *   rf[t][n][e] = gate(irfft(conj(full), n_fft)[n]), n < (n_fft + 1) / 2.
 * workspace: >= fs_rf_workspace_bytes(n_tx, n_freq, n_fft) (the per-transmit peaks and the split-BF16
 * inverse-DFT matrix, (n_fft + 1) / 2 rows x 2 n_freq columns). May be called with spectrum_full == NULL.
 * rf_ungated, when not NULL, receives the pre-gate RF (needed by fs_rf_adjoint). */
size_t fs_rf_workspace_bytes(int n_tx, int n_freq, int n_fft);
int fs_spectrum_to_rf(const float* spect, int n_tx, int n_freq, int n_elem, int n_ff, int idx0, int n_fft,
                      float* spectrum_full, float* rf, float* rf_ungated, void* workspace, size_t workspace_bytes,
                      void* stream);

/* Reverse of fs_spectrum_to_rf: given dL/d(rf) of the gated RF and the pre-gate RF, writes
 * dL/d(spect) for the selected band (interleaved complex float32, [n_tx][n_freq][n_elem]);
 * the complex cotangent convention is g = dL/dRe + i dL/dIm. */
size_t fs_rf_adjoint_workspace_bytes(int n_tx, int n_keep, int n_elem);
int fs_rf_adjoint(const float* rf_ungated, const float* d_rf, int n_tx, int n_freq, int n_elem, int idx0, int n_fft,
                  float* d_spect, void* workspace, size_t workspace_bytes, void* stream);

/* Kernels of fs_simus_adjoint, chosen per call (the same value must be passed to the workspace query):
 * FS_ADJOINT_AUTO = matrix cores where they apply (n_elem <= 128, n_sub <= 2: the element contractions
 * as split-BF16 MFMA products, Green's functions generated once per group of 16 transmits), else FP32
 * vector kernels; FS_ADJOINT_MATRIX = matrix cores only (FS_ERR_UNSUPPORTED where they do not apply);
 * FS_ADJOINT_VECTOR = vector kernels only (Green's functions regenerated per transmit). No counterpart
 * in the reference. */
enum fs_adjoint_method { FS_ADJOINT_AUTO = 0, FS_ADJOINT_MATRIX = 1, FS_ADJOINT_VECTOR = 2 };

/* Gradients of L through fs_simus_spectrum (method FP32 algebra) given dL/d(spect) (the corrected
 * selected spectrum, interleaved complex): d_rc, d_x, d_z [S] float32 (overwritten),
 * d_c [1] float64 (overwritten). Green's functions are recomputed, not stored. With n_scat = 0 only
 * d_c is written (0; the other pointers and the workspace may be null). The workspace query
 * returns 0 for an invalid descriptor or method. */
size_t fs_simus_adjoint_workspace_bytes(const fs_simus_desc* d, int adjoint_method);
int fs_simus_adjoint(const fs_simus_desc* d, int adjoint_method, const float* d_spect, float* d_rc, float* d_x,
                     float* d_z, double* d_c, void* workspace, size_t workspace_bytes, void* stream);

/* RMS pressure field of one transmit on n_scat points (scat_x/scat_z reused as the grid; rc unused):
 *   out[p] = sqrt(correction * sum_k |pulse_k probe_k sum_{e,j} g A|^2), 0 where out of field. */
size_t fs_pfield_workspace_bytes(const fs_simus_desc* d);
int fs_pfield(const fs_simus_desc* d, float* out, void* workspace, size_t workspace_bytes, void* stream);

/* Largest distance from any of n_points points (x, z) to any of n_src sources (src_x, src_z),
 * all float64 device arrays, written to out[0] (device, float64). The frequency plans use it
 * (simus: element centres; pfield: sub-element centroids). One pass, no host synchronisation:
 * the caller reads out[0] when it needs the plan. n_points may be 0 (out = 0). */
int fs_max_distance(const double* x, const double* z, long long n_points, const double* src_x, const double* src_z,
                    int n_src, double* out, void* stream);

/* IQ demodulation of RF, rf[t][n][e] float32 -> interleaved complex float32 iq[t][n][e] (channel_major 0)
 * or iq[t][e][n] (channel_major != 0: each sequence contiguous, the fastest input of fs_das):
 *   iq = 2 LPF(rf(n) exp(-i 2 pi fc n / fs)),
 * the low-pass run forward and backward (zero phase) over the sequence extended by padlen samples of
 * its odd reflection at each end, each pass starting from the steady state of its first input value.
 * The filter is cascaded second-order sections (sos[k] = b0 b1 b2 a0 a1 a2) with zi[k] the section's
 * steady-state states (transposed direct form II) for a unit constant input to that section.
 * n_samples must exceed padlen. workspace >= fs_rf2iq_workspace_bytes(...). Replaces pymust.rf2iq
 * (Butterworth low-pass with filtfilt) as the reference's harness calls it per firing. */
typedef struct fs_iir {
  int32_t n_sections; /* 1 .. 8 */
  int32_t padlen;
  double sos[8][6];
  double zi[8][2];
} fs_iir;
size_t fs_rf2iq_workspace_bytes(int n_firings, int n_channels, const fs_iir* filter);
int fs_rf2iq(const float* rf, int n_firings, int n_samples, int n_channels, double fs_hz, double fc_hz,
             const fs_iir* filter, int channel_major, float* iq, void* workspace, size_t workspace_bytes,
             void* stream);

/* Delay-and-sum of IQ iq[t][n][e] (interleaved complex float32; iq[t][e][n] when channel_major != 0) onto
 * n_pixels pixels (pixel_x, pixel_z float32 device arrays): for firing t the transmit arrival is
 * tau_tx = min_e(tx_delay[t][e] + |p - x_e| / c) (NaN delays: elements that do not transmit), element e
 * samples tau = tau_tx + |p - x_e| / c; the IQ at
 * (tau - t0) fs is interpolated linearly, rotated by exp(i 2 pi fc tau) and summed over the elements with
 * |x_p - x_e| <= z_p / (2 f_number) (f_number 0: all elements). compound != 0 writes image[p], the sum over
 * firings; else image[t][p]. Elements elem_x / elem_z and tx_delay [t][e] are float64 device arrays;
 * n_samples >= 2.
 * Replaces pymust.dasmtx applied per firing (tests/_picmus_phantom_shared.py:407-460). */
int fs_das(const float* iq, int n_firings, int n_samples, int n_channels, int channel_major, const double* elem_x,
           const double* elem_z, const double* tx_delay, const float* pixel_x, const float* pixel_z, long long n_pixels,
           double fs_hz, double fc_hz, double speed_of_sound, double t0, double f_number, int compound, float* image,
           void* stream);

/* Optional timing of the dominant kernel launches. enable is a mask of classes:
 * FS_TIMING_SPECTRUM (1): k_msm for FS_METHOD_MFMA_BF16X3, k_tx + k_rx for FS_METHOD_FP32;
 * FS_TIMING_ADJOINT (2): k_adjm per transmit group, or k_adj_tx + k_adj_pos per transmit;
 * 0 turns timing off. While enabled, each such launch is bracketed by HIP events on its stream.
 * fs_kernel_time_ms waits for the recorded launches, returns their summed duration in milliseconds,
 * stores the launch count in *launches (if not NULL) and clears the record. Process-wide and not
 * thread-safe (meant for benchmarks). No counterpart in the reference. */
enum fs_timing_class { FS_TIMING_SPECTRUM = 1, FS_TIMING_ADJOINT = 2 };
int fs_set_kernel_timing(int enable);
double fs_kernel_time_ms(int* launches);

#ifdef __cplusplus
}
#endif

#endif /* FASTSIMUS_H */
