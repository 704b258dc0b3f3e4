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
int col_tile_width(int n);  // columns per column-pass workgroup for an n-point column FFT

// fused step passes (step_kernels.hip)
hipError_t launch_rows(int kind, const KArgs& a, hipStream_t st);      // 0 fwd, 1 inv, 2 fwd+pre-taper, 3 inv+post-taper
hipError_t launch_rows_fanout(const KArgs& a, hipStream_t st);
hipError_t launch_rows_select(const KArgs& a, hipStream_t st);
hipError_t launch_cols(bool inv, const KArgs& a, hipStream_t st);
hipError_t launch_cols_fanout(int mode, const KArgs& a, hipStream_t st);
hipError_t launch_cols_fanin(int mode, const KArgs& a, hipStream_t st);
hipError_t launch_presence(const uint8_t* labels, uint32_t* rowmask, uint32_t* planemask, int NX, int NY,
                           long long nplanes, hipStream_t st);
hipError_t launch_present_count(const uint32_t* rowmask, int* cnt, int NY, long long nplanes, int nref,
                                hipStream_t st);

// point-wise kernels (ops_kernels.hip)
hipError_t launch_phase_shift(int mode, const float2* m, float2* d, const PwArgs& a, hipStream_t st);
hipError_t launch_select(const float2* m, float2* d, const PwArgs& a, hipStream_t st);
hipError_t launch_taper(const float2* m, float2* d, const PwArgs& a, hipStream_t st);
hipError_t launch_axpy(const float2* m, float2* d, size_t n, float alpha, int add, hipStream_t st);
hipError_t launch_copy16(const void* in, void* out, size_t n16, hipStream_t st);  // n16 16-byte words
hipError_t launch_zero(float2* d, size_t n, hipStream_t st);
// 16-bit block floating point checkpoint codec (64 complex samples per block)
hipError_t launch_bfp16_pack(const float2* in, void* q, float* scale, size_t n, hipStream_t st);
hipError_t launch_bfp16_unpack(const void* q, const float* scale, float2* out, size_t n, hipStream_t st);

// Born / imaging / injection kernels (born_kernels.hip)
hipError_t launch_rows_ic_fwd(const KArgs& a, hipStream_t st);
hipError_t launch_rows_scatter(bool born, const KArgs& a, hipStream_t st);
hipError_t launch_rows_scatter_adj(bool fused, const KArgs& a, hipStream_t st);
hipError_t launch_rows_adj_out(bool ic, const KArgs& a, hipStream_t st);
// Reflection sources of the up-going channels (point-wise on the model region)
struct ReflArgs {
  float2* out;              // wavefield (NS, NW, NY, NX): out (+)= source
  const float2* U;          // background U_z
  const float2* dU;         // channel-1 wavefield dU_z (MODE 1)
  const float2* s0; const float2* s1;    // slowness slices z, z+1 (nw, ny, nx)
  const float2* r0; const float2* r1;    // density slices z, z+1
  const float2* ds0; const float2* ds1;  // slowness perturbation slices z, z+1
  const float2* dr0; const float2* dr1;  // density perturbation slices z, z+1
  float2* gs0; float2* gs1;              // adjoint: slowness gradient slices z, z+1
  float2* gr0; float2* gr1;              // adjoint: density gradient slices z, z+1
  int NX, NY, NW, NS, nx, ny;
  int last;                 // z = nz-1: the slice is paired with itself (r = 0, L = 0)
};
hipError_t launch_refl_src(int mode, const ReflArgs& a, hipStream_t st);
hipError_t launch_refl_adj(const ReflArgs& a, hipStream_t st);
hipError_t launch_reflect_coef(const float2* slow, const float2* den, float2* r, int nz, size_t slice, hipStream_t st);
hipError_t launch_image(const float2* U, const float2* R, float2* img, int NX, int NY, int NW, int NS, int nx,
                        int ny, int add, hipStream_t st);
hipError_t launch_ic_fwd(const float2* img, float2* D, const float2* U, int NX, int NY, int NW, int NS, int nx, int ny,
                         int add, hipStream_t st);
hipError_t launch_ic_adj(float2* img, const float2* D, const float2* U, int NX, int NY, int NW, int NS, int nx, int ny,
                         int add, hipStream_t st);
// Register budget (2, 4 or 6 waves per SIMD) of the column kernels, kind 0
// fan-out / 1 fan-in, per mode (0, 1, 2) and column length; -1 if invalid.
int col_wpe_get(int kind, int mode, int n);
int col_wpe_set(int kind, int mode, int n, int wpe);
// Column layout of the n-point transforms (kind, mode as above): 0 one wave
// per column, 1 two waves per column (n = 1024 only); -1 if invalid.
int col_layout_get(int kind, int mode, int n);
int col_layout_set(int kind, int mode, int n, int layout);
// Layout of the K / W scratch for column length n: 1 quad-block, 0 row-major; -1 if invalid.
int scratch_quad_get(int n);
int scratch_quad_set(int n, int quad);
// out[k] += sum of g over the planes of shot k (first[k], count[k]); m per plane
hipError_t launch_plane_sum(const float2* g, float2* out, const int* first, const int* count, int nshot, long long m,
                            hipStream_t st);
// traces: trace t's sample of its plane iw at toff[t] + iw (iw < its plane count)
hipError_t launch_inj_fwd(const float2* tr, float2* wf, const long long* seg_base, const int* seg_off,
                          const int* seg_np, const int* ent_trace, const float* ent_w, const long long* toff, int s0,
                          int nseg, int wmax, long long plane, hipStream_t st);
hipError_t launch_inj_adj(float2* tr, const float2* wf, const int* act_trace, const long long* act_base,
                          const float4* act_w, const int* act_np, const long long* toff, int a0, int nact, int wmax,
                          int NX, long long plane, hipStream_t st);

}  // namespace fwix
