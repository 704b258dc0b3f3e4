// Wave-level Stockham FFT for gfx950 (CDNA4, 64-lane wavefronts).
//
// One wavefront transforms NB batches of N points held as V = NB*N/64 complex
// values per lane.  The radix plan (8, 4, 3, 2) is fixed at compile time from
// N and V so every loop unrolls and every index is a constant-folded shift or
// multiply.  Stages exchange data through the wave's own LDS region and only
// need wave-level ordering, never a workgroup barrier.
//
// Position conventions (Stockham autosort, decimation in time):
//   stage radix R, sub-transform size Ns (product of earlier radices):
//     butterfly (b, j), j < N/R, reads  x[b][j + r*N/R]
//     writes y[b][(j/Ns)*Ns*R + j%Ns + r*Ns]
//   The first stage therefore reads positions j + r*N/R0 and the last stage
//   writes positions j + r*N/RL, both contiguous across lanes: kernels load
//   the first stage straight from global memory and store the last stage
//   straight back.  The inverse plan is the forward plan reversed, so the
//   inverse's first stage reads exactly the positions the forward's last
//   stage wrote: a forward -> pointwise multiply -> inverse chain never leaves
//   registers between the two transforms.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fwix {

// ----------------------------------------------------------------- complex helpers
// float2 (complex64) is the storage and default arithmetic type; the same
// helpers exist for double2, so a transform can run in float64 arithmetic on
// values loaded from and stored to complex64 arrays.
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 czero() { return make_float2(0.f, 0.f); }

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// the arithmetic type T's scalar and constructor; conversions from / to the
// complex64 storage type
template <class T> struct CX;
template <> struct CX<float2> {
  using S = float;
  __device__ __forceinline__ static float2 mk(float x, float y) { return make_float2(x, y); }
};
template <> struct CX<double2> {
  using S = double;
  __device__ __forceinline__ static double2 mk(double x, double y) { return make_double2(x, y); }
};
template <class T> using real_of = typename CX<T>::S;
template <class T> __device__ __forceinline__ T cz() { return CX<T>::mk(0, 0); }
template <class T> __device__ __forceinline__ T from_f2(float2 a) { return CX<T>::mk(a.x, a.y); }
__device__ __forceinline__ float2 to_f2(float2 a) { return a; }
__device__ __forceinline__ float2 to_f2(double2 a) { return make_float2((float)a.x, (float)a.y); }

// multiply by sgn*i, sgn = -1 for the forward (e^{-i}) transform, +1 for the inverse
template <bool INV, class T>
__device__ __forceinline__ T mul_si(T a) {
  return INV ? CX<T>::mk(-a.y, a.x) : CX<T>::mk(a.y, -a.x);
}

// ----------------------------------------------------------------- compile-time plan
constexpr int kRadices[4] = {8, 4, 3, 2};

constexpr int plan_len(int N, int V) {
  int n = N, c = 0;
  for (int k = 0; k < 4; ++k) {
    const int R = kRadices[k];
    while (n % R == 0 && V % R == 0) { n /= R; ++c; }
  }
  return n == 1 ? c : -1;
}

constexpr int plan_radix(int N, int V, int s) {
  int n = N, c = 0;
  for (int k = 0; k < 4; ++k) {
    const int R = kRadices[k];
    while (n % R == 0 && V % R == 0) {
      if (c == s) return R;
      n /= R; ++c;
    }
  }
  return 0;
}

// radix of stage s of the forward (INV=false) or inverse (reversed) plan
template <int N, int V, bool INV>
struct Plan {
  static constexpr int S = plan_len(N, V);
  static_assert(S > 0, "FFT size has no radix plan for this register budget");
  static constexpr int radix(int s) { return plan_radix(N, V, INV ? (S - 1 - s) : s); }
  static constexpr int ns(int s) {  // product of the radices of stages < s
    int p = 1;
    for (int k = 0; k < s; ++k) p *= radix(k);
    return p;
  }
  static constexpr int first() { return radix(0); }
  static constexpr int last() { return radix(S - 1); }
};

// ----------------------------------------------------------------- in-register DFTs
template <int R, bool INV> struct Dft;

template <bool INV> struct Dft<2, INV> {
  template <class T>
  __device__ __forceinline__ static void run(T* x) {
    T a = x[0], b = x[1];
    x[0] = cadd(a, b); x[1] = csub(a, b);
  }
};

template <bool INV> struct Dft<4, INV> {
  template <class T>
  __device__ __forceinline__ static void run(T* x) {
    T a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    T c = cadd(x[1], x[3]), e = mul_si<INV>(csub(x[1], x[3]));
    x[0] = cadd(a, c); x[2] = csub(a, c);
    x[1] = cadd(b, e); x[3] = csub(b, e);
  }
};

template <bool INV> struct Dft<3, INV> {
  template <class T>
  __device__ __forceinline__ static void run(T* x) {
    using S = real_of<T>;
    const S h = (S)0.86602540378443864676;  // sin(2*pi/3)
    T t1 = cadd(x[1], x[2]);
    T t2 = mul_si<INV>(cscale(csub(x[1], x[2]), h));
    T m = CX<T>::mk(x[0].x - (S)0.5 * t1.x, x[0].y - (S)0.5 * t1.y);
    x[0] = cadd(x[0], t1);
    x[1] = cadd(m, t2);
    x[2] = csub(m, t2);
  }
};

template <bool INV> struct Dft<8, INV> {
  template <class T>
  __device__ __forceinline__ static void run(T* x) {
    using S = real_of<T>;
    T e[4] = {x[0], x[2], x[4], x[6]};
    T o[4] = {x[1], x[3], x[5], x[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    const S r = (S)0.70710678118654752440;
    // W8^k = exp(sgn*2*pi*i*k/8), sgn = -1 forward
    T o1 = cscale(cadd(o[1], mul_si<INV>(o[1])), r);           // (1 + sgn i)/sqrt2
    T o2 = mul_si<INV>(o[2]);                                   // sgn i
    T o3 = cscale(csub(mul_si<INV>(o[3]), o[3]), r);           // (-1 + sgn i)/sqrt2
    x[0] = cadd(e[0], o[0]); x[4] = csub(e[0], o[0]);
    x[1] = cadd(e[1], o1);   x[5] = csub(e[1], o1);
    x[2] = cadd(e[2], o2);   x[6] = csub(e[2], o2);
    x[3] = cadd(e[3], o3);   x[7] = csub(e[3], o3);
  }
};

// ----------------------------------------------------------------- wave ordering
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ----------------------------------------------------------------- positions
// Butterfly t = lane + 64*q of a radix-R stage owns batch b and index j.
// BFAST: batch index varies fastest across lanes (column tiles), else the
// position within the batch does (rows).
// (Unsigned arithmetic: t is never negative, and the lane index is opaque to
// the compiler inside a transform, so signed division would cost sign fix-ups.)
template <int N, int NB, int R, bool BFAST>
__device__ __forceinline__ void bj(int t, int& b, int& j) {
  const unsigned u = (unsigned)t;
  if (BFAST) { b = (int)(u % NB); j = (int)(u / NB); }
  else { b = (int)(u / (N / R)); j = (int)(u % (N / R)); }
}

// (batch, position) of value v = q*R + r at the input of a stage of radix R
// when it is the first stage, or at the output when it is the last stage.
template <int N, int NB, int R, bool BFAST>
__device__ __forceinline__ void edge_pos(int lane, int v, int& b, int& i) {
  const int q = v / R, r = v % R;
  int j;
  bj<N, NB, R, BFAST>(lane + 64 * q, b, j);
  i = j + r * (N / R);
}

// ----------------------------------------------------------------- the transform
// LDS layout policy: L::addr(b, i) gives the element offset of batch b, point i
// inside the wave's region.  T is the arithmetic type of the values, the LDS
// region and the twiddle table (float2, or double2 for float64 arithmetic).
template <int N, int NB, bool BFAST, bool INV, class L, class T = float2>
struct WaveFft {
  static constexpr int V = NB * N / 64;
  using P = Plan<N, V, INV>;
  static constexpr int S = P::S;

  template <int R, int Ns>
  __device__ __forceinline__ static void butterflies(T* v, const T* __restrict__ tw, int lane) {
#pragma unroll
    for (int q = 0; q < V / R; ++q) {
      T* x = v + q * R;
      if constexpr (Ns > 1) {
        int b, j;
        bj<N, NB, R, BFAST>(lane + 64 * q, b, j);
        const int k = (int)((unsigned)j % Ns);
        constexpr int step = N / (Ns * R);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          T w = tw[r * k * step];
          if (INV) w.y = -w.y;
          x[r] = cmul(x[r], w);
        }
      }
      Dft<R, INV>::run(x);
    }
  }

  template <int R, int Ns, int R2>
  __device__ __forceinline__ static void exchange(T* v, T* lds, int lane) {
#pragma unroll
    for (int q = 0; q < V / R; ++q) {
      int b, j;
      bj<N, NB, R, BFAST>(lane + 64 * q, b, j);
      const unsigned uj = (unsigned)j;
      const int base = (int)((uj / Ns) * Ns * R + (uj % Ns));
#pragma unroll
      for (int r = 0; r < R; ++r) lds[L::addr(b, base + r * Ns)] = v[q * R + r];
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < V / R2; ++q) {
      int b, j;
      bj<N, NB, R2, BFAST>(lane + 64 * q, b, j);
#pragma unroll
      for (int r = 0; r < R2; ++r) v[q * R2 + r] = lds[L::addr(b, j + r * (N / R2))];
    }
    wave_sync();
  }

  template <int s>
  __device__ __forceinline__ static void stages(T* v, T* lds, const T* __restrict__ tw, int lane) {
    if constexpr (s < S) {
      constexpr int R = P::radix(s);
      constexpr int Ns = P::ns(s);
      butterflies<R, Ns>(v, tw, lane);
      if constexpr (s + 1 < S) {
        exchange<R, Ns, P::radix(s + 1)>(v, lds, lane);
        stages<s + 1>(v, lds, tw, lane);
      }
    }
  }

  // v holds the values at first-stage input positions (edge_pos<..., P::first()>);
  // on return it holds the spectrum at last-stage output positions
  // (edge_pos<..., P::last()>).  Unnormalized.
  __device__ __forceinline__ static void run(T* v, T* lds, const T* __restrict__ tw, int lane) {
    // Keep the twiddles and the per-element twiddle / LDS addresses next to
    // their use: without these fences the compiler hoists every stage's
    // (loop-invariant) loads and addresses out of the per-reference loops
    // of the fused kernels and runs out of registers.  The lane index is
    // made opaque here, so the index arithmetic derived from it is redone
    // per transform (a few integer ops) instead of held in registers.
    asm volatile("" ::: "memory");
    asm volatile("" : "+v"(lane));
    stages<0>(v, lds, tw, lane);
  }
};

}  // namespace fwix
