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
// Columns per 256-thread tile in the column passes (4 waves x CW/4 columns).
constexpr int col_cw(int N) {
  if (N < 128) return (size_has3(N) ? 3072 : 2048) / N;
  if (N <= 512) return 16;
  return 8;  // 768, 1024: 2 columns per wave keeps V <= 32 without spills
}
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// LDS layouts -------------------------------------------------------------
template <int N>
struct RowLds {  // one wave's NB rows, row-major
  __device__ __forceinline__ static int addr(int b, int i) { return b * N + i; }
};
template <int CWP>
struct ColLds {  // tile [N][CWP], base pointer already offset to the wave's first column
  __device__ __forceinline__ static int addr(int b, int i) { return i * CWP + b; }
};

// One argument block for every step kernel (passed by value).
struct KArgs {
  const float2* in;        // input wavefield, already offset to the chunk's first plane
  float2* out;             // output wavefield, same offset
  float2* K;               // chunk spectrum scratch (np planes)
  float2* W;               // per-reference / per-Taylor-term scratch (nterm * np planes)
  const uint8_t* labels;   // (NW, NY, NX) reference labels of the current depth
  const float2* sref;      // (nref, NW) reference slownesses of the current depth
  const float* kx;         // (NX) wavenumbers
  const float* ky;         // (NY)
  const float* w2;         // (NW) (2 pi f)^2
  const float* w;          // (NW) 2 pi f
  const float* tapx;       // (NX) taper weights
  const float* tapy;       // (NY)
  const float2* twx;       // (NX) twiddles exp(-2 pi i k/NX)
  const float2* twy;       // (NY)
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
};

__device__ __forceinline__ int plane_iw(const KArgs& a, int p_local) { return (a.p0 + p_local) % a.NW; }

// Phase-shift multiplier exp(-i kz dz) for one reference slowness.
// Reference formula (propagator/src/phase_shift.cu:36-59): a = w2 Re s - k^2,
// b = w2 (Im s - eps Re s), kr = sqrt((c+a)/2), ki = sqrt((c-a)/2),
// c = sqrt(a^2+b^2).  Evaluated here as the numerically stable complex square
// root (the smaller of kr, ki from kr*ki = |b|/2), identical in exact
// arithmetic but free of the float32 cancellation in c-a / c+a.
// MODE 0: forward  exp(-ki dz) exp(-i kr dz)
// MODE 1: adjoint  kr -> -kr if b > 0; exp(-ki dz) exp(+i kr dz)
// MODE 2: inverse  kr -> -kr if b > 0; exp(+ki dz) exp(+i kr dz)
template <int MODE>
__device__ __forceinline__ float2 ps_mult(float w2, float sre, float sim, float k2, float dz, float eps) {
  const float a = fmaf(w2, sre, -k2);
  const float b = w2 * (sim - eps * sre);
  const float c = sqrtf(fmaf(a, a, b * b));
  float kr, ki;
  if (a >= 0.f) {
    kr = sqrtf(0.5f * (c + a));
    ki = kr > 0.f ? 0.5f * fabsf(b) / kr : 0.f;
  } else {
    ki = sqrtf(0.5f * (c - a));
    kr = ki > 0.f ? 0.5f * fabsf(b) / ki : 0.f;
  }
  if (MODE != 0 && b > 0.f) kr = -kr;
  const float att = expf((MODE == 2 ? ki : -ki) * dz);
  float sn, cs;
  sincosf(kr * dz, &sn, &cs);
  return MODE == 0 ? make_float2(att * cs, -att * sn) : make_float2(att * cs, att * sn);
}

}  // namespace fwix
