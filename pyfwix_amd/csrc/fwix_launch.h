// Host-side launchers of the HIP kernels (internal to the shared library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fwix_kernels.h"

namespace fwix {

// Argument block of the point-wise operator kernels.
struct PwArgs {
  int NX, NY, NW, NS;
  const float* kx;
  const float* ky;
  const float* w2;
  const float* tapx;
  const float* tapy;
  const float2* sref;      // (NW) reference slowness of the phase shift
  const uint8_t* labels;   // (NW, NY, NX)
  int value;               // selector reference index
  float dz, eps;
  int add;
};

bool fft_size_supported(int n);

// fused step passes (step_kernels.hip)
hipError_t launch_rows(int kind, const KArgs& a, hipStream_t st);      // 0 fwd, 1 inv, 2 fwd+pre-taper, 3 inv+post-taper
hipError_t launch_rows_fanout(const KArgs& a, hipStream_t st);
hipError_t launch_rows_select(const KArgs& a, hipStream_t st);
hipError_t launch_cols(bool inv, const KArgs& a, hipStream_t st);
hipError_t launch_cols_fanout(int mode, const KArgs& a, hipStream_t st);
hipError_t launch_cols_fanin(int mode, const KArgs& a, hipStream_t st);

// point-wise kernels (ops_kernels.hip)
hipError_t launch_phase_shift(int mode, const float2* m, float2* d, const PwArgs& a, hipStream_t st);
hipError_t launch_select(const float2* m, float2* d, const PwArgs& a, hipStream_t st);
hipError_t launch_taper(const float2* m, float2* d, const PwArgs& a, hipStream_t st);
hipError_t launch_axpy(const float2* m, float2* d, size_t n, float alpha, int add, hipStream_t st);
hipError_t launch_zero(float2* d, size_t n, hipStream_t st);

}  // namespace fwix
