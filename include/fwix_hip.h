/*
 * fwix_hip.h -- C ABI of the MI355X (gfx950) one-way depth-stepping library
 * (libpyfwix_hip.so).
 *
 * Every entry point replaces one method of the reference's device operator
 * API: CudaOperator<M,D>::cu_forward / cu_adjoint (operator/src/CudaOperator.h:61-72)
 * and the concrete operators in operator/src and propagator/src.  Shapes are
 * passed by value, buffers are plain device pointers owned by the caller
 * (an operator borrows them, as the reference's operators borrow the
 * complex_vector* they are given, CudaOperator.h:29-53), and streams are
 * explicit hipStream_t values (0 = the default stream).
 *
 * Conventions
 *   - complex64 samples, interleaved (re, im), layout x fastest:
 *     wavefield (NS, NW, NY, NX), labels (nz, NW, NY, NX) uint8,
 *     reference slownesses (nz, nref, NW), slowness / model (nz, NW, ny, nx),
 *     traces (ntrace, NW) -- operator/src/complex_vector.h:10-20,
 *     propagator/src/injection.cu:26-27.
 *   - add = 0 overwrites the output, add != 0 accumulates into it
 *     (CudaOperator.h:74-120).
 *   - every function returns 0 on success and a negative value on error;
 *     fwix_last_error() then returns a message (thread-local).  Nothing
 *     calls exit() (the reference's gpuAssert does, complex_vector.h:22-30).
 *   - calls on one operator handle must not overlap in time (each handle owns
 *     one set of scratch buffers).
 */
#ifndef FWIX_HIP_H
#define FWIX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fwix_c64 { float re, im; } fwix_c64;
typedef void* fwix_stream;               /* a hipStream_t */
typedef struct fwix_op_s* fwix_op;       /* opaque operator handle */

/* Wavefield hypercube: the padded (x, y, w, s) domain of the step operators
 * (propagator/src/Propagator.cpp:61-73: NX = nx + padx, NY = ny + pady, the
 * padding appended at the high side, origin and spacing of the slowness). */
typedef struct fwix_dims {
  int NX, NY, NW, NS;       /* padded lateral grid, frequencies, shots */
  int nx, ny;               /* unpadded model grid, nx <= NX, ny <= NY */
  float ox, dx, oy, dy;     /* lateral axes (origin, spacing) */
  float ow, dw;             /* frequency axis in Hz: f = ow + iw*dw */
  float oz, dz;             /* depth axis */
  int nz;                   /* depth steps */
} fwix_dims;

/* Operator parameters: the reference's paramObj keys and defaults
 * (nref 1 OneStep.h:64, taperx/tapery 0 Taper.h:16-17, eps 0 OneStep.h:66,
 * eps_scat 0.04 Scatter.cpp:12, taylor 1 Scatter.cpp:14). */
typedef struct fwix_params {
  int nref;
  int taperx, tapery;
  float eps;
  float eps_scat;
  int taylor;
  int planes_per_chunk;     /* planes per fused launch; 0 = automatic */
} fwix_params;

/* step methods; FWIX_PSPI_PLAIN runs the PSPI step at the reference's
 * operator granularity (taper, 2-D FFT, per reference: phase shift, inverse
 * 2-D FFT, select; PSPI.cpp:66-103) with one kernel per operator and one
 * shot per launch -- the comparator of the fused path. */
enum { FWIX_PSPI = 0, FWIX_NSPS = 1, FWIX_PSPI_PLAIN = 2 };

/* ---------------------------------------------------------------- library */
int fwix_version(void);
const char* fwix_last_error(void);
void fwix_params_default(fwix_params* p);
/* 1 if an n-point lateral FFT is supported (2^a or 3*2^a, 8 <= n <= 1024) */
int fwix_fft_size_supported(int n);
/* release any operator handle */
int fwix_op_destroy(fwix_op op);

/* ---------------------------------------------------------------- cuFFT2d
 * Unitary batched 2-D C2C transform over (y, x), e^{-i} forward, adjoint =
 * inverse.  Replaces cuFFT2d (operator/src/FFT.h:10-37, FFT.cpp:5-69 with
 * the rsqrt(NX*NY) store callback, fft_callback.cu:7-22).  model == data is
 * allowed (the in-place cu_forward(data) / cu_adjoint(data), FFT.cpp:57-69). */
int fwix_fft2d_create(const fwix_dims* d, fwix_op* out);
int fwix_fft2d_forward(fwix_op op, int add, const fwix_c64* model, fwix_c64* data, fwix_stream s);
int fwix_fft2d_adjoint(fwix_op op, int add, fwix_c64* model, const fwix_c64* data, fwix_stream s);

/* ---------------------------------------------------------------- PhaseShift
 * PhaseShift(domain, dz, eps) (propagator/src/PhaseShift.h:11-72,
 * PhaseShift.cpp:6-48, phase_shift.cu).  set_slow copies the NW reference
 * slownesses of one reference from host memory (PhaseShift.h:21-23). */
int fwix_phase_shift_create(const fwix_dims* d, float dz, float eps, fwix_op* out);
int fwix_phase_shift_set_slow(fwix_op op, const fwix_c64* sref_host);
int fwix_phase_shift_forward(fwix_op op, int add, const fwix_c64* model, fwix_c64* data, fwix_stream s);
int fwix_phase_shift_adjoint(fwix_op op, int add, fwix_c64* model, const fwix_c64* data, fwix_stream s);
int fwix_phase_shift_inverse(fwix_op op, int add, fwix_c64* model, const fwix_c64* data, fwix_stream s);

/* ---------------------------------------------------------------- Selector
 * Selector(domain) + set_labels / set_value (propagator/src/Selector.h:8-58,
 * selector.cu:7-37).  Labels are int32 (NW, NY, NX) host memory, as the
 * reference's RefSampler hands them over (RefSampler.h:33-36). */
int fwix_selector_create(const fwix_dims* d, fwix_op* out);
int fwix_selector_set_labels(fwix_op op, const int32_t* labels_host);
int fwix_selector_set_value(fwix_op op, int value);
int fwix_selector_forward(fwix_op op, int add, const fwix_c64* model, fwix_c64* data, fwix_stream s);
int fwix_selector_adjoint(fwix_op op, int add, fwix_c64* model, const fwix_c64* data, fwix_stream s);

/* ---------------------------------------------------------------- Taper
 * Taper(domain, par) (propagator/src/Taper.h:8-42, taper.cu:7-77). */
int fwix_taper_create(const fwix_dims* d, int taperx, int tapery, fwix_op* out);
int fwix_taper_forward(fwix_op op, int add, const fwix_c64* model, fwix_c64* data, fwix_stream s);
int fwix_taper_adjoint(fwix_op op, int add, fwix_c64* model, const fwix_c64* data, fwix_stream s);

/* ---------------------------------------------------------------- OneStep: PSPI / NSPS
 * PSPI / NSPS(domain, slow_hyper, par) (propagator/src/OneStep.h:11-111,
 * PSPI.cpp, NSPS.cpp).  set_refs borrows device arrays for all depths:
 * sref (nz, nref, NW) complex64 and labels (nz, NW, NY, NX) uint8 -- the
 * reference's RefSampler output (RefSampler.h:29-36) resident in HBM instead
 * of copied host-to-device at every set_depth (OneStep.h:42-45).
 * forward: data (+)= P_z model; adjoint: model (+)= P_z^H data.  model == data
 * runs the step in place (PSPI::cu_forward(model) / cu_adjoint(data),
 * PSPI.cpp:66-103).  The PSPI adjoint always includes the taper (the
 * reference's non-in-place adjoint omits it, PSPI.cpp:26-44) and the NSPS
 * adjoint is the true adjoint (NSPS.cpp:16 passes swapped arguments). */
int fwix_onestep_create(int method, const fwix_dims* d, const fwix_params* p, fwix_op* out);  /* any FWIX_* method */
int fwix_onestep_set_refs(fwix_op op, const fwix_c64* sref_dev, const uint8_t* labels_dev);
int fwix_onestep_set_depth(fwix_op op, int iz);
/* Per-plane frequencies: NW values in Hz (host) replacing ow + iw * dw, so
 * that the planes of one launch may belong to shots with different frequency
 * sets (shots of one padded size batched along the plane axis).  No
 * counterpart in the reference, whose operators take a regular axis.  The
 * oneway and born variants below do the same for those handles. */
int fwix_onestep_set_frequencies(fwix_op op, const float* f_hz);
/* Shot batch on the fused path (NS == 1): shot k owns the next nplanes[k]
 * planes (sum = NW); references (nz, nref, nshot) and labels
 * (nz, nshot, NY, NX) are then per shot and must be set again.  No
 * counterpart in the reference (one shot geometry per operator). */
int fwix_onestep_set_shot_planes(fwix_op op, int nshot, const int* nplanes);
int fwix_onestep_forward(fwix_op op, int add, const fwix_c64* model, fwix_c64* data, fwix_stream s);
int fwix_onestep_adjoint(fwix_op op, int add, fwix_c64* model, const fwix_c64* data, fwix_stream s);
/* One PSPI or NSPS depth step that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same result as fwix_onestep_forward (adjoint = 0: data (+)= P model) or
 * fwix_onestep_adjoint (adjoint = 1: model (+)= P^H data) of a handle made
 * with the same method, dims and params, for one depth: sref (nref, NW) and
 * labels (NW, NY, NX) of that depth, device pointers.  The step's tables and
 * scratch live in one library-owned device array of
 * FWIX_NOALLOC_SCRATCH_BYTES per GPU (held from the first call on a GPU until
 * the process ends); the planes run in chunks that fit it, and a grid whose
 * tables plus (nref + 1) planes do not fit is rejected with an error.  Calls
 * that overlap in time on one GPU race on that array. */
#define FWIX_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int fwix_onestep_apply_noalloc(int method, int adjoint, const fwix_dims* d, const fwix_params* p,
                               const fwix_c64* sref_dev, const uint8_t* labels_dev, int add, fwix_c64* model,
                               fwix_c64* data, fwix_stream s);

/* ---------------------------------------------------------------- Injection
 * Injection(domain = traces (ntrace, NW), range = wavefield, oz, dz, cx, cy,
 * cz, ids) + set_coords / set_depth (propagator/src/Injection.h:10-55,
 * injection.cu).  Coordinates and shot ids are host arrays of ntrace values.
 * forward scatters traces into the wavefield at depth step iz (bilinear in
 * x, y, linear in z); adjoint records.  The forward is race-free (each node
 * sums its contributions in a fixed order; the reference's kernel races,
 * injection.cu:89-92). */
int fwix_injection_create(const fwix_dims* d, int ntrace, fwix_op* out);
int fwix_injection_set_coords(fwix_op op, const float* cx, const float* cy, const float* cz, const int* ids);
int fwix_injection_set_depth(fwix_op op, int iz);
int fwix_injection_forward(fwix_op op, int add, const fwix_c64* traces, fwix_c64* wfld, fwix_stream s);
int fwix_injection_adjoint(fwix_op op, int add, fwix_c64* traces, const fwix_c64* wfld, fwix_stream s);

/* ---------------------------------------------------------------- Scatter
 * Scatter(domain, slow, par) + set_depth (propagator/src/Scatter.h:12-104,
 * Scatter.cpp, scatter.cu): Born scattering, Taylor series of dkz/ds with
 * params.taylor terms and params.eps_scat.  set_slow borrows the unpadded
 * background slowness (nz, NW, ny, nx) in device memory (the reference copies
 * one host slice per set_depth, Scatter.cpp:68-74). */
int fwix_scatter_create(const fwix_dims* d, const fwix_params* p, fwix_op* out);
int fwix_scatter_set_slow(fwix_op op, const fwix_c64* slow_dev);
int fwix_scatter_set_depth(fwix_op op, int iz);
int fwix_scatter_forward(fwix_op op, int add, const fwix_c64* model, fwix_c64* data, fwix_stream s);
int fwix_scatter_adjoint(fwix_op op, int add, fwix_c64* model, const fwix_c64* data, fwix_stream s);

/* ---------------------------------------------------------------- ImagingCondition
 * ImagingCondition(domain = image (NW, ny, nx), range = wavefield)
 * (propagator/src/ImagingCondition.h:11-43, imaging_condition.cu).
 * set_background borrows the background wavefield slice U_z (device).
 * forward: wfld (+)= U_z * img on the model region; adjoint: img (+)=
 * sum_shots conj(U_z) * wfld, deterministic (no float atomics). */
int fwix_ic_create(const fwix_dims* d, fwix_op* out);
int fwix_ic_set_background(fwix_op op, const fwix_c64* bg_dev);
int fwix_ic_forward(fwix_op op, int add, const fwix_c64* img, fwix_c64* wfld, fwix_stream s);
int fwix_ic_adjoint(fwix_op op, int add, fwix_c64* img, const fwix_c64* wfld, fwix_stream s);

/* ---------------------------------------------------------------- Downward / Upward
 * Downward (direction 0) / Upward (direction 1) depth recursions
 * (propagator/src/OneWay.h:14-245, OneWay.cpp:6-79) over params.nref-reference
 * PSPI (method FWIX_PSPI, as the reference, OneWay.h:24-25) or NSPS steps.
 * forward: data (+)= P_{nz-1}...P_0 model (Downward) or P_0...P_{nz-1} model
 * (Upward); model is left untouched (the reference steps it in place).
 * ckpt (nullable, nz wavefields) receives the wavefield before every step
 * (OneWay::compress_slice, OneWay.h:52-58). */
int fwix_oneway_create(int direction, int method, const fwix_dims* d, const fwix_params* p, fwix_op* out);
int fwix_oneway_set_refs(fwix_op op, const fwix_c64* sref_dev, const uint8_t* labels_dev);
int fwix_oneway_set_frequencies(fwix_op op, const float* f_hz);
/* Shot batch (NS == 1): shot k owns the next nplanes[k] planes (the counts
 * add up to NW); references and labels are then per shot, (nz, nref, nshot)
 * and (nz, nshot, NY, NX), and must be set again.  No counterpart in the
 * reference, which runs one shot geometry per operator. */
int fwix_oneway_set_shot_planes(fwix_op op, int nshot, const int* nplanes);
/* Consecutive steps of the "shared spectrum" shape (PSPI forward, NSPS
 * adjoint) chain through the row spectrum: a step's select pass also leaves
 * F_x(taper * U_{z+1}) for the next step, which skips its first row pass.
 * on = 1 (default) chains, 0 runs every step whole; results agree to float32
 * rounding.  No counterpart in the reference. */
int fwix_oneway_set_step_chaining(fwix_op op, int on);
int fwix_oneway_forward(fwix_op op, int add, const fwix_c64* model, fwix_c64* data, fwix_c64* ckpt, fwix_stream s);
int fwix_oneway_adjoint(fwix_op op, int add, fwix_c64* model, const fwix_c64* data, fwix_stream s);

/* ---------------------------------------------------------------- Born (ExtendedBorn, down-scattering channel)
 * The background down-going modeling of Propagator::forward
 * (propagator/src/Propagator.cpp:155-180) and the linearized Born operator
 * of ExtendedBorn::downward_scattering_fwd/adj (ExtendedBorn.cpp:152-208)
 * with its migration image.  Background wavefields stay in HBM: every depth
 * (ckpt_every <= 1) or every ckpt_every-th depth with recomputation in the
 * adjoint -- lossless, replacing the ZFP WavefieldPool (WavefieldPool.cpp).
 *   sources / receivers: host coordinates (n values) and shot ids
 *   wavelet (nsrc, NW), data (nrec, NW), model ds (nz, NW, ny, nx),
 *   image (nz, ny, nx) = sum_w sum_s conj(U_z) P_z^H lam_{z+1}
 * fwix_born_model must run (after set_refs / set_sources / set_wavelet)
 * before forward / adjoint / migrate; it optionally records the direct data. */
int fwix_born_create(const fwix_dims* d, const fwix_params* p, int ckpt_every, fwix_op* out);
int fwix_born_set_refs(fwix_op op, const fwix_c64* sref_dev, const uint8_t* labels_dev);
int fwix_born_set_frequencies(fwix_op op, const float* f_hz);
int fwix_born_set_slow(fwix_op op, const fwix_c64* slow_dev);
int fwix_born_set_sources(fwix_op op, int n, const float* x, const float* y, const float* z, const int* ids);
int fwix_born_set_receivers(fwix_op op, int n, const float* x, const float* y, const float* z, const int* ids);
int fwix_born_set_wavelet(fwix_op op, const fwix_c64* wavelet_dev);
int fwix_born_model(fwix_op op, int add, fwix_c64* data, fwix_stream s);
int fwix_born_forward(fwix_op op, int add, const fwix_c64* dslow, fwix_c64* data, fwix_stream s);
int fwix_born_adjoint(fwix_op op, int add, fwix_c64* dslow, const fwix_c64* data, fwix_stream s);
int fwix_born_migrate(fwix_op op, int add, fwix_c64* image, const fwix_c64* data, fwix_stream s);
/* Shot batch: many shots of one padded grid share every launch (no
 * counterpart in the reference, which runs one shot geometry per operator;
 * StreamingPropagator.cpp:68-153 runs its batches one after another).  With
 * NS == 1, shot k owns the next nplanes[k] planes (sum = NW) and each plane
 * has its own frequency (fwix_born_set_frequencies).  Afterwards:
 *   slowness and model ds (nz, nshot, ny, nx), one slice per shot (shared by
 *   its planes); references (nz, nref, nshot); labels (nz, nshot, NY, NX);
 *   the adjoint returns ds summed over the shot's planes;
 *   trace ids are shots; trace t holds one sample per plane of its shot, the
 *   traces stored one after another (fwix_born_trace_samples gives the total).
 * Slowness, references, sources, receivers and wavelet must be set again; the
 * migration image and the reflection / extended channels are refused. */
int fwix_born_set_shot_planes(fwix_op op, int nshot, const int* nplanes);
/* FWIX_PSPI (default, fused depth steps) or FWIX_PSPI_PLAIN: the imaging
 * condition, the scattering operator (Scatter.cpp:90-116) and the PSPI step
 * (PSPI.cpp:66-103) as separate operators, one shot per launch, chained as
 * Scattering.cpp:43-61 and ExtendedBorn.cpp:152-208 do -- the comparator of the
 * fused path.  Same results; the background must be modeled again. */
int fwix_born_set_method(fwix_op op, int method);
/* samples of all receiver (receivers != 0) or source traces: nrec * NW
 * without a shot batch; -1 for an invalid handle */
long long fwix_born_trace_samples(fwix_op op, int receivers);
/* Checkpoint codec of the background wavefields (the reference stores them
 * ZFP-compressed at a fixed rate, WavefieldPool.cpp:13-284): 0 = lossless
 * complex64 (default), 1 = 16-bit block floating point (64 complex samples
 * share one float scale; 4.06 bytes per sample; |error| <= scale / 65534).
 * Forward and adjoint read the same decoded fields, so they stay adjoint.
 * Set before fwix_born_model. */
int fwix_born_set_codec(fwix_op op, int codec);
/* Arithmetic of the Born steps.  The first Taylor term of the scattering
 * operator weighs wavenumber k by k^2 / (w^2 s) (Scatter.h:73,
 * scatter.cu:23-59): up to ~10^3 at a few Hz on 10 m grids, so complex64
 * rounding of the amplified terms (in a transform, or a product in space)
 * lands on the propagating band the recordings see.  FWIX_PREC_F64_SCATTER
 * runs every pass of the Born adjoint step and the scattering + propagation
 * passes of the Born forward step in float64 arithmetic (values still stored
 * as complex64 between passes); FWIX_PREC_F32 (default) keeps float32.  No
 * counterpart in the reference (float32 only). */
enum { FWIX_PREC_F32 = 0, FWIX_PREC_F64_SCATTER = 1 };
int fwix_born_set_precision(fwix_op op, int precision);
/* Background recursions (fwix_born_model and the recomputed checkpoint
 * segments) chain consecutive PSPI steps through the row spectrum: a step's
 * select pass also leaves F_x(taper * U_{z+1}) for the next step, which then
 * skips its first row pass (one pass and one read of U fewer per step), except
 * where a source injects in between.  on = 1 (default) chains, 0 runs every
 * step whole.  Results agree to float32 rounding.  Costs one wavefield-sized
 * device buffer.  No counterpart in the reference. */
int fwix_born_set_step_chaining(fwix_op op, int on);
/* The codec on any device array of n complex samples: q holds n short2,
 * scales (n + 63) / 64 floats. */
int fwix_bfp16_pack(const fwix_c64* in, void* q, float* scales, size_t n, fwix_stream s);
int fwix_bfp16_unpack(const void* q, const float* scales, fwix_c64* out, size_t n, fwix_stream s);
/* Background density (nz, NW, ny, nx), borrowed; needed by the up pass and
 * the extended operator (Propagator.cpp:117-120 set_background_model). */
int fwix_born_set_density(fwix_op op, const fwix_c64* den_dev);
/* Up pass of Propagator::forward (Propagator.cpp:181-211): V_nz = 0,
 * V_z = P_z V_{z+1} + r_z U_z with r from (slowness, density); keeps every
 * V_z in HBM and records data (+)= Rec_z^T V_z when data != NULL.  Run after
 * fwix_born_model; together they give the nonlinear modeling's data. */
int fwix_born_model_up(fwix_op op, int add, fwix_c64* data, fwix_stream s);
/* ExtendedBorn::forward / adjoint with all four channels
 * (ExtendedBorn.cpp:80-150): model (dslow, dden), each (nz, NW, ny, nx);
 * down scattering, down-reflected, up scattering (scatter and PSPI at z-1
 * with V_z, ds_z) and back scattering through the linearized reflection
 * coefficient.  The linearized reflection is the exact derivative of
 * fwix_reflect_coef (lin_reflect.cu:69-79 uses (s1/s0)^(1/4) where the
 * derivative has (s1/s0)^(1/2)).  Needs fwix_born_model and
 * fwix_born_model_up; holds two more wavefields per depth in HBM. */
int fwix_born_extended_forward(fwix_op op, int add, const fwix_c64* dslow, const fwix_c64* dden, fwix_c64* data,
                               fwix_stream s);
int fwix_born_extended_adjoint(fwix_op op, int add, fwix_c64* dslow, fwix_c64* dden, const fwix_c64* data,
                               fwix_stream s);
/* Up-going reflection channel of the background modeling
 * (propagator/src/Propagator.cpp:185-208, Reflect.h:20-36, reflect.cu:8-47):
 * V = 0; for z = nz-1..0: V <- P_z V; V += refl_z * U_z on the model region;
 * data (+)= Rec_z^T V.  With refl = fwix_reflect_coef(slow, density) and
 * fwix_born_model's direct recording, this is Propagator::forward's data.
 * It is linear in refl (nz, NW, ny, nx); reflect_adjoint is its adjoint
 * (refl_z (+)= sum_s conj(U_z) mu_z with mu_z = Inj_z(data) + P_{z-1}^H mu_{z-1})
 * and reflect_migrate the migration image of surface data,
 * image_z (+)= sum_{w,s} conj(U_z) mu_z, image (nz, ny, nx). */
int fwix_born_reflect_forward(fwix_op op, int add, const fwix_c64* refl, fwix_c64* data, fwix_stream s);
int fwix_born_reflect_adjoint(fwix_op op, int add, fwix_c64* refl, const fwix_c64* data, fwix_stream s);
int fwix_born_reflect_migrate(fwix_op op, int add, fwix_c64* image, const fwix_c64* data, fwix_stream s);
/* Reflection coefficients r_z = (rho_{z+1} sqrt(s_z) - rho_z sqrt(s_{z+1})) /
 * (rho_{z+1} sqrt(s_z) + rho_z sqrt(s_{z+1})), principal square roots, r_{nz-1} = 0
 * (reflect.cu:42-45, Reflect.h:29-36); slow, den, refl (nz, NW, ny, nx). */
int fwix_reflect_coef(const fwix_dims* d, const fwix_c64* slow, const fwix_c64* den, fwix_c64* refl, fwix_stream s);
/* depth steps (PSPI applications of one wavefield) executed so far, -1 on a bad handle */
long long fwix_born_steps(fwix_op op);

/* ---------------------------------------------------------------- live pass timing
 * No counterpart in the reference (which times with std::chrono and pasted
 * Nsight numbers).  When enabled, every pass of the hot path records a pair
 * of HIP events on its launch stream; fwix_timing_collect synchronizes them
 * and accumulates per pass: milliseconds, launches and compulsory HBM bytes
 * (what the pass must read and write).  fwix_timing_report writes
 * {"pass": {"ms": .., "launches": .., "bytes": .., "total_launches": ..}, ...}
 * as JSON (ms, launches and bytes of the recorded launches; total_launches
 * counts every launch of the pass since the last reset).
 * fwix_timing_select restricts recording to one pass name (null: all
 * passes) and to every `every`-th launch of each pass, so a timed region
 * carries few event markers.  fwix_timing_overhead records `pairs` empty event
 * pairs on a stream and returns their mean elapsed time in microseconds
 * (the fixed cost inside every recorded pair). */
int fwix_timing_enable(int on);
int fwix_timing_select(const char* pass, int every);
int fwix_timing_overhead(fwix_stream s, int pairs, float* us);
int fwix_timing_collect(void);
int fwix_timing_reset(void);
int fwix_timing_report(char* buf, int len);

/* ---------------------------------------------------------------- column kernel tuning
 * No counterpart in the reference (cuFFT picks its own kernels).  Every
 * fused column kernel exists with three register budgets: 2, 4 or 6 waves
 * per SIMD.  kind 0 = fan-out transforms (PSPI forward, NSPS adjoint,
 * scattering terms), 1 = fan-in transforms; mode 0 / 1 = forward / adjoint
 * phase shift, 2 = scattering terms; n = column length (NY).
 * fwix_tune_columns_set returns 0, or -1 for an invalid argument;
 * fwix_tune_columns_get returns the budget in use or -1.  The table is
 * process-wide; pyfwix_amd/autotune.py sets it from timed launches. */
int fwix_tune_columns_set(int kind, int mode, int n, int waves_per_simd);
int fwix_tune_columns_get(int kind, int mode, int n);
/* Layout of the 1024-point column transforms (kind, mode as above; no
 * counterpart in the reference): 0 = one wave per column, 1 = two waves per
 * column (a radix-2 split into two 512-point transforms, twice the waves per
 * tile).  Any other n accepts only layout 0.  Results are identical up to
 * float rounding.  The float64 scattering mode always uses layout 0.
 * _set returns 0 or -1 for an invalid argument; _get the layout or -1. */
int fwix_tune_columns_layout_set(int kind, int mode, int n, int layout);
int fwix_tune_columns_layout_get(int kind, int mode, int n);
/* Layout of the fused passes' internal scratch (the spectrum K and the
 * per-reference / per-term planes W) for column length n (no counterpart in
 * the reference): 1 = 4 x 4 quad blocks (a 128-byte line holds 4 rows x 4
 * columns), 0 = row-major.  Default: quad blocks at n = 1024, row-major
 * elsewhere.  Results are bit-identical in both.  Takes effect for the next
 * operator call; _set returns 0 or -1 for an invalid argument, _get the
 * layout or -1. */
int fwix_tune_scratch_layout_set(int n, int quad);
int fwix_tune_scratch_layout_get(int n);

/* Device copy of `bytes` (a multiple of 16, 16-byte aligned buffers) with
 * 16-byte loads and stores: the stream-copy kernel that measures the
 * achievable HBM bandwidth beside the nominal peak.  No counterpart in the
 * reference. */
int fwix_copy(const void* in, void* out, size_t bytes, fwix_stream s);

#ifdef __cplusplus
}
#endif
#endif /* FWIX_HIP_H */
