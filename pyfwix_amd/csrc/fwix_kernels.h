// Shared kernel argument block, tile geometry and per-point physics for the
// PSPI / NSPS / Born depth-step kernels.
#pragma once
#include "fwix_fft.h"

namespace fwix {

// Supported lateral FFT sizes: 2^a and 3*2^a between 8 and 1024.
constexpr bool size_has3(int N) { return N % 3 == 0; }

// Rows per wavefront in the row passes (V = NB*N/64 values per lane).
constexpr int row_nb(int N) {
  const int base = size_has3(N) ? 768 : 512;
  return N >= base ? 1 : base / N;
}
// Column passes: 512-thread workgroups, 8 waves x CW/8 columns, with each
// lane holding V = CW*N/512 values: 8 (2^a, N <= 512), 16 (1024) or 12 (3*2^a).
// Two workgroups fit a CU, so one tile's loads overlap another's transforms.
// Default register budget (waves per SIMD) of the column kernels, from timed
// launches of every budget on MI355X: kind 0 the forward / fan-out
// transforms, kind 1 the fan-in transforms; mode 2 the scattering terms.
// (Every mode of a kind has the same winners so far; mode stays in the key.)
constexpr int col_wpe_default(int kind, int mode, int N) {
  const bool pow2 = !size_has3(N);
  (void)mode;
  if (kind == 0) return (pow2 && N <= 512) ? 6 : (N == 384 || N == 768) ? 4 : 2;
  return (pow2 && N <= 512) || N == 768 ? 4 : 2;
}
// 1024-point columns use 256-thread workgroups (4 waves) on 4-column tiles:
// a 1024 x 8 tile takes 74 KB of LDS and, with the 256-VGPR budget these
// transforms need, one 8-wave workgroup per CU, so nothing overlapped one
// tile's loads with another's transforms; two 4-wave workgroups share a CU.
constexpr int col_threads(int N) { return N == 1024 ? 256 : 512; }
constexpr int col_cw(int N) {
  if (N == 1024) return 4;
  const int cw = (size_has3(N) ? 6144 : 4096) / N;
  return cw < 8 ? 8 : cw;
}
constexpr int kColThreads = 512;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// LDS layouts -------------------------------------------------------------
// Stockham exchanges write with stride R inside a batch; one pad slot every
// 32 complex samples spreads those writes over all 64 LDS banks.
constexpr int padded(int N) { return N + N / 32; }
template <int N>
struct PadLds {  // one wave's NB batches, batch-major, padded
  __device__ __forceinline__ static int addr(int b, int i) { return b * padded(N) + i + (int)((unsigned)i >> 5); }
};

// One argument block for every step kernel (passed by value).
struct KArgs {
  const float2* in;        // input wavefield, already offset to the chunk's first plane
  float2* out;             // output wavefield, same offset
  float2* K;               // chunk spectrum scratch (np planes)
  float2* W;               // per-reference / per-Taylor-term scratch (nterm * np planes)
  const uint8_t* labels;   // (NW, NY, NX) reference labels of the current depth
  const uint32_t* rowmask; // (NW, NY) bit r set: reference r labels a node of that row (null: all)
  const uint32_t* planemask; // (NW) OR of rowmask over the plane's rows (null: all)
  const float2* sref;      // (nref, NW) reference slownesses of the current depth
  const int* mplane;       // (NW) model plane of every plane (null: the plane itself)
  const float* kx;         // (NX) wavenumbers
  const float* ky;         // (NY)
  const float* w2;         // (NW) (2 pi f)^2
  const float* w;          // (NW) 2 pi f
  const float* tapx;       // (NX) taper weights
  const float* tapy;       // (NY)
  const float2* twx;       // (NX) twiddles exp(-2 pi i k/NX)
  const float2* twy;       // (NY)
  const double2* twx64;    // (NX), (NY) twiddles and taper weights in float64, for
  const double2* twy64;    //   the passes that run in float64 arithmetic (f64)
  const double* tapx64;
  const double* tapy64;
  const float2* slow;      // (NW, ny, nx) unpadded slowness slice (Born scatter)
  const float2* dslow;     // (NW, ny, nx) model perturbation slice (Born IC)
  const float2* bg;        // background wavefield slice, chunk offset (Born IC)
  const float2* addin;     // extra wavefield added before the taper (Born: dU_z)
  float2* aux;             // extra output (checkpoint copy / image)
  int NX, NY, NW, NS, nx, ny;
  int nref, nterm;
  int p0, np;              // chunk: first global plane, number of planes
  long long wstride;       // elements between consecutive W terms (np*NX*NY)
  float dz, eps, eps_s;
  float scale;             // normalisation applied by the kernel (set per launch)
  int add;
  int f64;                 // the launch runs its transforms in float64 arithmetic
  int qin, qout;           // k_rows: in / out are K (in the scratch layout below)
  int quad;                // K and W in the quad-block layout (else row-major)
  int knext;               // k_rows_select: also K = scale * F_x(taper * out), the next step's spectrum
                           //   (2: K only, out itself is not needed)
};

// Quad-block layout of the column-pass scratch (K and the W terms, where the
// launch's a.quad says so; row-major otherwise): a 128-byte
// line holds a 4 x 4 block (rows 4q..4q+3 x columns 4p..4p+3, row-major inside),
// blocks row-major.  A column tile of 4 or 8 columns then loads and stores
// whole lines, and the 4 rows of a row-pass workgroup share every line they
// touch.  y counts rows over all planes of the chunk (NY is a multiple of 4).
__device__ __forceinline__ unsigned qidx(int y, int x, int NX) {
  return (unsigned)(((((y >> 2) * (NX >> 2)) + (x >> 2)) << 4) + ((y & 3) << 2) + (x & 3));
}
// element (y, x) of the K / W scratch in the launch's layout, as a row part
// (uniform over a row) plus a column part: quad blocks put (y, x) at
// ((y >> 2) NX + (y & 3)) 4 + ((x >> 2) << 4) + (x & 3)
__device__ __forceinline__ unsigned wrow(const KArgs& a, int y, int NX) {
  return a.quad ? (unsigned)(((y >> 2) * NX + (y & 3)) << 2) : (unsigned)(y * NX);
}
__device__ __forceinline__ unsigned wcol(const KArgs& a, int x) {
  const unsigned u = (unsigned)x;
  return a.quad ? ((u >> 2) << 4) + (u & 3u) : u;
}
__device__ __forceinline__ unsigned widx(const KArgs& a, int y, int x, int NX) {
  return wrow(a, y, NX) + wcol(a, x);
}

// twiddle tables of the arithmetic type T (float2 or double2)
__device__ __forceinline__ const float2* twx_of(const KArgs& a, float2*) { return a.twx; }
__device__ __forceinline__ const double2* twx_of(const KArgs& a, double2*) { return a.twx64; }
__device__ __forceinline__ const float2* twy_of(const KArgs& a, float2*) { return a.twy; }
__device__ __forceinline__ const double2* twy_of(const KArgs& a, double2*) { return a.twy64; }

__device__ __forceinline__ int plane_iw(const KArgs& a, int p_local) { return (a.p0 + p_local) % a.NW; }
// Index into the model-sized arrays (labels, presence masks, slowness, model
// perturbation): plane iw itself, or the shot it belongs to when planes of
// several shots share one launch.
// MAP is a template flag so that the per-element index of the unbatched
// kernels carries no branch (a per-element test of a.mplane split the
// unrolled loads into separate blocks and cost the row passes 7-17 %).
template <bool MAP>
__device__ __forceinline__ int model_iw(const KArgs& a, int iw) {
  if constexpr (MAP) return a.mplane[iw];
  else return iw;
}
// once-per-workgroup form (column kernels)
__device__ __forceinline__ int model_iw_any(const KArgs& a, int iw) { return a.mplane ? a.mplane[iw] : iw; }

// lowest set bit of m above position r (32 when none); r = -1 gives the lowest
__device__ __forceinline__ int next_bit(uint32_t m, int r) {
  const uint32_t rest = r >= 31 ? 0u : (m & ~((2u << r) - 1u));
  return rest ? __builtin_ctz(rest) : 32;
}

// sin and cos of x (radians) from the transcendental unit: v_sin_f32 /
// v_cos_f32 take revolutions, so x is scaled by 1/(2 pi) and reduced to
// [0, 1) with v_fract_f32.  Measured on gfx950 over [-2, 2) revolutions:
// |error| <= 1.3e-7 against double-precision sin / cos, float rounding level
// for the phases of a depth step (|kz dz| well below one revolution).
__device__ __forceinline__ void fast_sincos(float x, float* s, float* c) {
  const float r = __builtin_amdgcn_fractf(x * 0.159154943091895336f);
  *s = __builtin_amdgcn_sinf(r);
  *c = __builtin_amdgcn_cosf(r);
}

// Phase-shift multiplier exp(-i kz dz) for one reference slowness.
// Reference formula (propagator/src/phase_shift.cu:36-59): a = w2 Re s - k^2,
// b = w2 (Im s - eps Re s), kr = sqrt((c+a)/2), ki = sqrt((c-a)/2),
// c = sqrt(a^2+b^2).  Evaluated as the numerically stable complex square
// root (the smaller of kr, ki from kr*ki = |b|/2): identical in exact
// arithmetic, free of the float32 cancellation in c-a / c+a.
// MODE 0: forward  exp(-ki dz) exp(-i kr dz)
// MODE 1: adjoint  kr -> -kr if b > 0; exp(-ki dz) exp(+i kr dz)
// MODE 2: inverse  kr -> -kr if b > 0; exp(+ki dz) exp(+i kr dz)
// b depends only on (frequency, reference): callers test b == 0 once and
// use ps_real (kz real or purely imaginary) for the whole plane.
template <int MODE>
__device__ __forceinline__ float2 ps_real(float a, float dz) {
  const float t = __builtin_amdgcn_sqrtf(fabsf(a)) * dz;
  float sn, cs;
  fast_sincos(t, &sn, &cs);
  const float e = __expf(MODE == 2 ? t : -t);
  if (a >= 0.f) return MODE == 0 ? make_float2(cs, -sn) : make_float2(cs, sn);
  return make_float2(e, 0.f);
}
template <int MODE>
__device__ __forceinline__ float2 ps_general(float a, float b, float dz) {
  const float c = __builtin_amdgcn_sqrtf(fmaf(a, a, b * b));
  const float big = __builtin_amdgcn_sqrtf(0.5f * (c + fabsf(a)));
  const float small = 0.5f * fabsf(b) * __builtin_amdgcn_rcpf(big);
  float kr = a >= 0.f ? big : small;
  const float ki = a >= 0.f ? small : big;
  if (MODE != 0 && b > 0.f) kr = -kr;
  const float att = __expf((MODE == 2 ? ki : -ki) * dz);
  float sn, cs;
  fast_sincos(kr * dz, &sn, &cs);
  return MODE == 0 ? make_float2(att * cs, -att * sn) : make_float2(att * cs, att * sn);
}
// The same multipliers in float64 arithmetic (the float64 scattering passes:
// near the evanescent boundary a = w2 Re s - k^2 cancels, and a float32
// phase there is the largest error left once the transforms run in float64).
template <int MODE>
__device__ __forceinline__ double2 ps_real64(double a, double dz) {
  const double t = sqrt(fabs(a)) * dz;
  if (a >= 0.0) {
    double sn, cs;
    sincos(t, &sn, &cs);
    return MODE == 0 ? make_double2(cs, -sn) : make_double2(cs, sn);
  }
  return make_double2(exp(MODE == 2 ? t : -t), 0.0);
}
template <int MODE>
__device__ __forceinline__ double2 ps_general64(double a, double b, double dz) {
  const double c = sqrt(fma(a, a, b * b));
  const double big = sqrt(0.5 * (c + fabs(a)));
  const double small = 0.5 * fabs(b) / big;
  double kr = a >= 0.0 ? big : small;
  const double ki = a >= 0.0 ? small : big;
  if (MODE != 0 && b > 0.0) kr = -kr;
  const double att = exp((MODE == 2 ? ki : -ki) * dz);
  double sn, cs;
  sincos(kr * dz, &sn, &cs);
  return MODE == 0 ? make_double2(att * cs, -att * sn) : make_double2(att * cs, att * sn);
}

template <int MODE>
__device__ __forceinline__ float2 ps_mult(float w2, float sre, float sim, float k2, float dz, float eps) {
  const float a = fmaf(w2, sre, -k2);
  const float b = w2 * (sim - eps * sre);
  return b == 0.f ? ps_real<MODE>(a, dz) : ps_general<MODE>(a, b, dz);
}

}  // namespace fwix
