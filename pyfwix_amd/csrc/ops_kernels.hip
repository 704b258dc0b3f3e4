// Plain per-operator kernels: the reference's operator granularity
// (PhaseShift, Selector, Taper, add/scale), used by the standalone operator
// entry points and by the plain per-shot comparator path.  One thread per
// complex sample, x fastest; the point-wise kernels run on a (plane samples,
// frequency, shot) grid so a thread finds its (x, y) with one 32-bit
// division (a plane holds at most 1024 x 1024 samples) and its frequency
// from the block index.
#include "fwix_kernels.h"
#include "fwix_launch.h"

namespace fwix {

static inline dim3 flat_grid(size_t n) {
  size_t g = (n + kThreads - 1) / kThreads;
  if (g > 1u << 30) g = 1u << 30;
  return dim3((unsigned)g);
}

// grid (plane samples / kThreads, NW, NS): j = sample within the plane
static inline dim3 plane_grid(const PwArgs& a) {
  return dim3((unsigned)(((size_t)a.NX * a.NY + kThreads - 1) / kThreads), (unsigned)a.NW, (unsigned)a.NS);
}
__device__ __forceinline__ size_t plane_base(const PwArgs& a) {
  return ((size_t)blockIdx.z * a.NW + blockIdx.y) * ((size_t)a.NX * a.NY);
}

// d = (add ? d : 0) + Phi * m  (propagator/src/phase_shift.cu:7-61,125-184,248-305)
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_phase_shift(const float2* __restrict__ m, float2* d, PwArgs a) {
  const unsigned j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= (unsigned)(a.NX * a.NY)) return;
  const unsigned iy = j / (unsigned)a.NX, ix = j - iy * (unsigned)a.NX;
  const int iw = blockIdx.y;
  const size_t idx = plane_base(a) + j;
  const float2 s = a.sref[iw];
  const float k2 = fmaf(a.kx[ix], a.kx[ix], a.ky[iy] * a.ky[iy]);
  float2 y = cmul(m[idx], ps_mult<MODE>(a.w2[iw], s.x, s.y, k2, a.dz, a.eps));
  if (a.add) y = cadd(y, d[idx]);
  d[idx] = y;
}

// d = (add ? d : 0) + [labels == value] * m  (propagator/src/selector.cu:7-37)
__global__ __launch_bounds__(kThreads) void k_select(const float2* __restrict__ m, float2* d, PwArgs a) {
  const unsigned j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= (unsigned)(a.NX * a.NY)) return;
  const size_t idx = plane_base(a) + j;
  float2 y = a.labels[(size_t)blockIdx.y * a.NX * a.NY + j] == a.value ? m[idx] : czero();
  if (a.add) y = cadd(y, d[idx]);
  d[idx] = y;
}

// d = (add ? d : 0) + wx(ix) wy(iy) m  (propagator/src/taper.cu:7-77)
__global__ __launch_bounds__(kThreads) void k_taper(const float2* __restrict__ m, float2* d, PwArgs a) {
  const unsigned j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= (unsigned)(a.NX * a.NY)) return;
  const unsigned iy = j / (unsigned)a.NX, ix = j - iy * (unsigned)a.NX;
  const size_t idx = plane_base(a) + j;
  float2 y = cscale(m[idx], a.tapx[ix] * a.tapy[iy]);
  if (a.add) y = cadd(y, d[idx]);
  d[idx] = y;
}

// d = (add ? d : 0) + alpha * m  (operator/src/complex_vector.cu:4-29 add / scale)
__global__ __launch_bounds__(kThreads) void k_axpy(const float2* __restrict__ m, float2* d, size_t n, float alpha, int add) {
  for (size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x; idx < n; idx += (size_t)gridDim.x * kThreads) {
    float2 y = cscale(m[idx], alpha);
    if (add) y = cadd(y, d[idx]);
    d[idx] = y;
  }
}

// plain stream copy, four 16-byte words per lane (coalesced across the
// workgroup, all loads issued before the stores): the achievable HBM
// bandwidth reference
__global__ __launch_bounds__(kThreads) void k_copy16(const float4* __restrict__ in, float4* __restrict__ out, size_t n) {
  const size_t base = (size_t)blockIdx.x * (4 * kThreads) + threadIdx.x;
  float4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t i = base + (size_t)k * kThreads;
    if (i < n) v[k] = in[i];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const size_t i = base + (size_t)k * kThreads;
    if (i < n) out[i] = v[k];
  }
}

// zero fill (used instead of hipMemsetAsync so that captured graphs replay correctly)
__global__ __launch_bounds__(kThreads) void k_zero(float2* d, size_t n) {
  for (size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x; idx < n; idx += (size_t)gridDim.x * kThreads)
    d[idx] = czero();
}

// Fixed-rate checkpoint codec: 16-bit block floating point.  Each block of 64
// complex samples (one wave) keeps one float scale m = max(|re|, |im|) and
// every component as round(v * 32767 / m) in int16, 4.06 bytes per complex
// sample instead of 8, with |error| <= m / 65534 (plus float rounding).
__global__ __launch_bounds__(kThreads) void k_bfp16_pack(const float2* __restrict__ in, short2* __restrict__ q,
                                                         float* __restrict__ scale, size_t n) {
  const size_t nblk = (n + 63) / 64;
  const int lane = threadIdx.x & 63;
  const size_t waves = (size_t)gridDim.x * (kThreads / 64);
  for (size_t blk = ((size_t)blockIdx.x * kThreads + threadIdx.x) / 64; blk < nblk; blk += waves) {
    const size_t i = blk * 64 + lane;
    const float2 v = i < n ? in[i] : czero();
    float m = fmaxf(fabsf(v.x), fabsf(v.y));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    const float inv = m > 0.f ? 32767.f / m : 0.f;
    if (i < n) q[i] = make_short2((short)__float2int_rn(v.x * inv), (short)__float2int_rn(v.y * inv));
    if (lane == 0) scale[blk] = m;
  }
}
__global__ __launch_bounds__(kThreads) void k_bfp16_unpack(const short2* __restrict__ q, const float* __restrict__ scale,
                                                           float2* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
    const float s = scale[i / 64] * (1.f / 32767.f);
    const short2 v = q[i];
    out[i] = make_float2((float)v.x * s, (float)v.y * s);
  }
}

hipError_t launch_bfp16_pack(const float2* in, void* q, float* scale, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bfp16_pack, flat_grid(n), dim3(kThreads), 0, st, in, (short2*)q, scale, n);
  return hipGetLastError();
}
hipError_t launch_bfp16_unpack(const void* q, const float* scale, float2* out, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bfp16_unpack, flat_grid(n), dim3(kThreads), 0, st, (const short2*)q, scale, out, n);
  return hipGetLastError();
}

hipError_t launch_phase_shift(int mode, const float2* m, float2* d, const PwArgs& a, hipStream_t st) {
  if (a.NW > 65535 || a.NS > 65535) return hipErrorInvalidValue;
  const dim3 g = plane_grid(a);
  if (mode == 0) hipLaunchKernelGGL(k_phase_shift<0>, g, dim3(kThreads), 0, st, m, d, a);
  else if (mode == 1) hipLaunchKernelGGL(k_phase_shift<1>, g, dim3(kThreads), 0, st, m, d, a);
  else hipLaunchKernelGGL(k_phase_shift<2>, g, dim3(kThreads), 0, st, m, d, a);
  return hipGetLastError();
}
hipError_t launch_select(const float2* m, float2* d, const PwArgs& a, hipStream_t st) {
  if (a.NW > 65535 || a.NS > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_select, plane_grid(a), dim3(kThreads), 0, st, m, d, a);
  return hipGetLastError();
}
hipError_t launch_taper(const float2* m, float2* d, const PwArgs& a, hipStream_t st) {
  if (a.NW > 65535 || a.NS > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_taper, plane_grid(a), dim3(kThreads), 0, st, m, d, a);
  return hipGetLastError();
}
hipError_t launch_copy16(const void* in, void* out, size_t n16, hipStream_t st) {
  if (n16 == 0) return hipSuccess;
  const size_t g = (n16 + 4 * kThreads - 1) / (4 * kThreads);
  if (g > 0x7fffffffu) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)g), dim3(kThreads), 0, st, (const float4*)in, (float4*)out, n16);
  return hipGetLastError();
}
hipError_t launch_axpy(const float2* m, float2* d, size_t n, float alpha, int add, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy, flat_grid(n), dim3(kThreads), 0, st, m, d, n, alpha, add);
  return hipGetLastError();
}
hipError_t launch_zero(float2* d, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero, flat_grid(n), dim3(kThreads), 0, st, d, n);
  return hipGetLastError();
}

}  // namespace fwix
