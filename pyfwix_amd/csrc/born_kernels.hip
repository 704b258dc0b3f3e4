// Born scattering, imaging condition and injection kernels.
//
// Born forward step (propagator/src/ExtendedBorn.cpp:152-178 with
// ForwardScattering, Scattering.cpp:43-52):
//     dU_{z+1} = P_z [ dU_z + S_z(U_z * ds_z) ]
//   rows_ic_fwd       v = U_z * ds_z (zero in the padding), F_x        -> K
//   cols_fanout<2>    F_y, (k^2/w^2)^it, F_y^-1 for it = 1..T           -> W_it
//   rows_scatter      order 0 from U_z * ds_z, orders >= 1 from F_x^-1 W_it,
//                     * c_it s~^(-1/2-it), sum, * -i w dz/2, + dU_z, taper, F_x -> K
//   cols_fanout<0>, rows_select: the PSPI step                           -> dU_{z+1}
// Born adjoint step (ExtendedBorn.cpp:180-208, Scattering.cpp:54-61):
//   rows_fanout, cols_fanin<1>     the first two PSPI adjoint passes     -> K
//   rows_scatter_adj<fused>        mu = taper F_x^-1 K (written to lam),
//                                  * +i w dz/2, conj(c_it s~^(-1/2-it)), F_x  -> W_it (it >= 1)
//   cols_fanin<2>                  F_y, (k^2/w^2)^it, sum, F_y^-1        -> K
//   rows_adj_out<ic>               F_x^-1 K + order-0 term,
//                                  ds_z += conj(U_z) * (.) on the model region
// Taylor order 0 is the identity in wavenumber and never goes through a
// transform; only orders >= 1 use the column passes.
#include "fwix_kernels.h"
#include "fwix_launch.h"

#include <type_traits>

namespace fwix {

template <int N>
struct BRowCfg {
  static constexpr int NB = row_nb(N);
  static constexpr int V = NB * N / 64;
  static constexpr int LDSW = NB * padded(N);
};

// c_it * s~^(-1/2-it), s~ = Re s + i (Im s - eps_s Re s), from the edge-padded
// slowness slice (propagator/src/scatter.cu:61-106, pad_fwd :240-277,
// Taylor coefficients Scatter.h:73).  s^(-1/2) is the reciprocal of the
// principal square root, identical to |s|^(-1/2) e^{-i arg(s)/2}
// (scatter.cu:6-21); higher terms multiply by 1/s.
__device__ __forceinline__ void scatter_factors(const KArgs& a, int im, int iy, int ix, float2& rsq, float2& inv) {
  const int yy = iy < a.ny ? iy : a.ny - 1;
  const int xx = ix < a.nx ? ix : a.nx - 1;
  const float2 s = a.slow[(unsigned)((im * a.ny + yy) * a.nx + xx)];
  const float sr = s.x, si = s.y - a.eps_s * s.x;
  // principal sqrt q of (sr, si), from the hardware square root and
  // reciprocal (1 ulp each); |q|^2 = |s~| = r gives both factors from 1/r
  const float r = __builtin_amdgcn_sqrtf(fmaf(sr, sr, si * si));
  float qr, qi;
  if (sr >= 0.f) {
    qr = __builtin_amdgcn_sqrtf(0.5f * (r + sr));
    qi = qr > 0.f ? 0.5f * si * __builtin_amdgcn_rcpf(qr) : 0.f;
  } else {
    const float t = __builtin_amdgcn_sqrtf(0.5f * (r - sr));
    qr = 0.5f * fabsf(si) * __builtin_amdgcn_rcpf(t);
    qi = copysignf(t, si);
  }
  const float ir = __builtin_amdgcn_rcpf(r);
  rsq = make_float2(qr * ir, -qi * ir);                // 1/sqrt(s~) = conj(q)/|s~|
  const float is2 = ir * ir;
  inv = make_float2(sr * is2, -si * is2);              // 1/s~ = conj(s~)/|s~|^2
}
// the same factors in float64 arithmetic (from the complex64 slowness)
__device__ __forceinline__ void scatter_factors(const KArgs& a, int im, int iy, int ix, double2& rsq, double2& inv) {
  const int yy = iy < a.ny ? iy : a.ny - 1;
  const int xx = ix < a.nx ? ix : a.nx - 1;
  const float2 s = a.slow[(unsigned)((im * a.ny + yy) * a.nx + xx)];
  const double sr = s.x, si = (double)s.y - (double)a.eps_s * s.x;
  const double r = sqrt(fma(sr, sr, si * si));
  double qr, qi;
  if (sr >= 0.0) {
    qr = sqrt(0.5 * (r + sr));
    qi = qr > 0.0 ? 0.5 * si / qr : 0.0;
  } else {
    const double t = sqrt(0.5 * (r - sr));
    qr = 0.5 * fabs(si) / t;
    qi = copysign(t, si);
  }
  const double ir = 1.0 / r;
  rsq = make_double2(qr * ir, -qi * ir);
  const double is2 = ir * ir;
  inv = make_double2(sr * is2, -si * is2);
}
// taper weight of node (i, row j) in the arithmetic type T
__device__ __forceinline__ float tap_of(const KArgs& a, int i, int j, float2*) { return a.tapx[i] * a.tapy[j]; }
__device__ __forceinline__ double tap_of(const KArgs& a, int i, int j, double2*) { return a.tapx64[i] * a.tapy64[j]; }

__device__ __forceinline__ float taylor_coef(int it) {  // Scatter.h:73
  return it == 0 ? 1.f : it == 1 ? 0.5f : it == 2 ? 0.375f : it == 3 ? 0.3125f : 0.2734375f;
}

// K = scale * F_x( U * ds ) on the model region (x < nx, y < ny), zero elsewhere
template <int N, bool MAP>
__global__ __launch_bounds__(kThreads) void k_rows_ic_fwd(KArgs a) {
  using C = BRowCfg<N>;
  using F = WaveFft<N, C::NB, false, false, PadLds<N>>;
  __shared__ float2 lds[kWaves * C::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrows = a.np * a.NY;
  const int rb = (blockIdx.x * kWaves + wave) * C::NB;
  // twiddles in LDS: a global twiddle load would wait (one counter for all
  // vector memory operations) for every load still in flight
  __shared__ float2 twl[N];
  for (int i = threadIdx.x; i < N; i += kThreads) twl[i] = a.twx[i];
  __syncthreads();
  if (rb >= nrows) return;
  // rows of the padding (y >= ny) carry no product: their spectrum is zero
  bool any = false;
  for (int b = 0; b < C::NB; ++b) {
    const int lr = rb + b;
    if (lr < nrows && lr % a.NY < a.ny) any = true;
  }
  float2 v[C::V];
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
    const int lr = rb + b;
    v[e] = czero();
    if (lr < nrows) {
      const int p = lr / a.NY, iy = lr - p * a.NY;
      if (iy < a.ny && i < a.nx) {
        const float2 u = a.bg[(unsigned)(lr * N + i)];
        const float2 m = a.dslow[(unsigned)((model_iw<MAP>(a, plane_iw(a, p)) * a.ny + iy) * a.nx + i)];
        v[e] = cmul(u, m);
      }
    }
  }
  if (any) F::run(v, lds + wave * C::LDSW, twl, lane);
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
    const int lr = rb + b;
    if (lr < nrows) a.K[widx(a, lr, i, N)] = cscale(v[e], a.scale);
  }
}

// Sum of the Taylor terms in space:
//   y = -i (w dz / 2) [ c_0 s~^(-1/2) v + sum_{it>=1} c_it s~^(-1/2-it) F_x^-1(W_{it-1}) ]
// where order 0 needs no transform: v is the scattering source itself
// (BORN: U_z * ds_z on the model region, else the input rows).
// BORN = true : K = scale * F_x( taper * (y + addin) )   (feeds the PSPI column pass)
// BORN = false: out = (add ? out : 0) + y                (standalone Scatter forward)
// (T: arithmetic type, float2 or double2; loads and stores stay complex64)
// ONE: the launch has exactly one Taylor order (nterm == 2).  Its factors
// s~^(-1/2), s~^(-3/2) and the source v are then formed after the row
// transform, so that only the transformed row is live across it (the
// general form holds the running sum and two factor arrays across every
// transform).
template <int N, bool BORN, bool MAP, class T = float2, bool ONE = false>
__global__ __launch_bounds__(kThreads) void k_rows_scatter(KArgs a) {
  using C = BRowCfg<N>;
  using S = real_of<T>;
  using Fi = WaveFft<N, C::NB, false, true, PadLds<N>, T>;
  using F = WaveFft<N, C::NB, false, false, PadLds<N>, T>;
  __shared__ T lds[kWaves * C::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrows = a.np * a.NY;
  const int rb = (blockIdx.x * kWaves + wave) * C::NB;
  // twiddles in LDS: inside the term loops a global load would wait (one
  // counter for all vector memory operations) for the term stores / the next
  // term's loads still in flight
  __shared__ T twl[N];
  const T* twg = twx_of(a, (T*)nullptr);
  for (int i = threadIdx.x; i < N; i += kThreads) twl[i] = twg[i];
  __syncthreads();
  if (rb >= nrows) return;
  T* wl = lds + wave * C::LDSW;
  // scattering source v of node (lr, i): U_z * ds_z on the model region
  // (BORN) or the input row
  auto source = [&](int lr, int p, int iy, int i) -> T {
    if constexpr (BORN) {
      if (iy < a.ny && i < a.nx)
        return cmul(from_f2<T>(a.bg[(unsigned)(lr * N + i)]),
                    from_f2<T>(a.dslow[(unsigned)((model_iw<MAP>(a, plane_iw(a, p)) * a.ny + iy) * a.nx + i)]));
      return cz<T>();
    } else {
      return from_f2<T>(a.in[(unsigned)(lr * N + i)]);
    }
  };
  T acc[C::V];
  if constexpr (ONE) {
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, Fi::P::first(), false>(lane, e, b, i);
      const int lr = rb + b;
      acc[e] = lr < nrows ? from_f2<T>(a.W[widx(a, lr, i, N)]) : cz<T>();
    }
    Fi::run(acc, wl, twl, lane);
    const S c = (S)taylor_coef(1) * (S)a.scale;  // the inverse row transform's 1/sqrt(NX)
    // the lane index made opaque here: the per-element offsets are formed
    // after the transform instead of held in registers across it
    int ln = lane;
    asm volatile("" : "+v"(ln));
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, Fi::P::last(), false>(ln, e, b, i);
      const int lr = rb + b;
      if (lr < nrows) {
        const int p = lr / a.NY, iy = lr - p * a.NY;
        T pw, inv;
        scatter_factors(a, model_iw<MAP>(a, plane_iw(a, p)), iy, i, pw, inv);
        const T v = source(lr, p, iy, i);
        const T pw1 = cmul(pw, inv);
        acc[e] = cadd(cmul(v, pw), cscale(cmul(acc[e], pw1), c));  // c_0 = 1
      } else {
        acc[e] = cz<T>();
      }
    }
  } else {
    T pw[C::V], inv[C::V];
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, Fi::P::last(), false>(lane, e, b, i);
      const int lr = rb + b;
      acc[e] = cz<T>();
      pw[e] = cz<T>();
      inv[e] = cz<T>();
      if (lr < nrows) {
        const int p = lr / a.NY, iy = lr - p * a.NY;
        scatter_factors(a, model_iw<MAP>(a, plane_iw(a, p)), iy, i, pw[e], inv[e]);
        acc[e] = cmul(source(lr, p, iy, i), pw[e]);  // c_0 = 1
        pw[e] = cmul(pw[e], inv[e]);
      }
    }
    for (int it = 1; it < a.nterm; ++it) {
      const float2* W = a.W + (size_t)(it - 1) * a.wstride;
      T x[C::V];
#pragma unroll
      for (int e = 0; e < C::V; ++e) {
        int b, i;
        edge_pos<N, C::NB, Fi::P::first(), false>(lane, e, b, i);
        const int lr = rb + b;
        x[e] = lr < nrows ? from_f2<T>(W[widx(a, lr, i, N)]) : cz<T>();
      }
      Fi::run(x, wl, twl, lane);
      const S c = (S)taylor_coef(it) * (S)a.scale;  // the inverse row transform's 1/sqrt(NX)
#pragma unroll
      for (int e = 0; e < C::V; ++e) {
        acc[e] = cadd(acc[e], cscale(cmul(x[e], pw[e]), c));   // c_it s~^(-1/2-it) * term
        pw[e] = cmul(pw[e], inv[e]);
      }
    }
  }
  // acc *= -i * (w dz / 2)     (propagator/src/scatter.cu:155-195)
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, Fi::P::last(), false>(lane, e, b, i);
    const int lr = rb + b;
    if (lr < nrows) {
      const int p = lr / a.NY;
      const S h = (S)(0.5f * a.w[plane_iw(a, p)] * a.dz);
      acc[e] = CX<T>::mk(h * acc[e].y, -h * acc[e].x);
    }
  }
  if constexpr (!BORN) {
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, Fi::P::last(), false>(lane, e, b, i);
      const int lr = rb + b;
      if (lr < nrows) {
        const unsigned idx = lr * N + i;
        T y = acc[e];
        if (a.add) y = cadd(y, from_f2<T>(a.out[idx]));
        a.out[idx] = to_f2(y);
      }
    }
  } else {
    static_assert(Fi::P::last() == F::P::first(), "register hand-off needs mirrored plans");
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
      const int lr = rb + b;
      if (lr < nrows) {
        T y = cadd(acc[e], from_f2<T>(a.addin[(unsigned)(lr * N + i)]));
        acc[e] = cscale(y, tap_of(a, i, lr % a.NY, (T*)nullptr));
      }
    }
    F::run(acc, wl, twl, lane);
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
      const int lr = rb + b;
      if (lr < nrows) a.K[widx(a, lr, i, N)] = to_f2(cscale(acc[e], a.scale));
    }
  }
}

// Scatter adjoint, first half: u = +i (w dz / 2) * mu and, for Taylor orders
// it >= 1, W_{it-1} = scale * F_x( conj(c_it s~^(-1/2-it)) * u ).
// FUSED = true : mu = taper * scale * F_x^-1(K) is the end of the PSPI adjoint
//                step; it is written to out (the adjoint wavefield) as well.
// FUSED = false: mu = in (standalone Scatter adjoint).
// ONE: the launch has exactly one Taylor order (nterm == 2), a separate
// instantiation so that the general form's per-order factors cost it no registers.
template <int N, bool FUSED, bool MAP, class T = float2, bool ONE = false>
__global__ __launch_bounds__(kThreads) void k_rows_scatter_adj(KArgs a) {
  using C = BRowCfg<N>;
  using S = real_of<T>;
  using F = WaveFft<N, C::NB, false, false, PadLds<N>, T>;
  using Fi = WaveFft<N, C::NB, false, true, PadLds<N>, T>;
  static_assert(Fi::P::last() == F::P::first(), "register hand-off needs mirrored plans");
  __shared__ T lds[kWaves * C::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrows = a.np * a.NY;
  const int rb = (blockIdx.x * kWaves + wave) * C::NB;
  // twiddles in LDS: inside the term loops a global load would wait (one
  // counter for all vector memory operations) for the term stores / the next
  // term's loads still in flight
  __shared__ T twl[N];
  const T* twg = twx_of(a, (T*)nullptr);
  for (int i = threadIdx.x; i < N; i += kThreads) twl[i] = twg[i];
  __syncthreads();
  if (rb >= nrows) return;
  T* wl = lds + wave * C::LDSW;
  T u[C::V];
  if constexpr (FUSED) {
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, Fi::P::first(), false>(lane, e, b, i);
      const int lr = rb + b;
      u[e] = lr < nrows ? from_f2<T>(a.K[widx(a, lr, i, N)]) : cz<T>();
    }
    Fi::run(u, wl, twl, lane);
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, Fi::P::last(), false>(lane, e, b, i);
      const int lr = rb + b;
      if (lr < nrows) {
        u[e] = cscale(u[e], (S)a.scale * tap_of(a, i, lr % a.NY, (T*)nullptr));
        a.out[(unsigned)(lr * N + i)] = to_f2(u[e]);
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
      const int lr = rb + b;
      u[e] = lr < nrows ? from_f2<T>(a.in[(unsigned)(lr * N + i)]) : cz<T>();
    }
  }
  if (a.nterm < 2) return;
  if constexpr (ONE) {  // one Taylor order: x = conj(c_1 s~^(-3/2)) * u, nothing carried to a next order
    const S c = (S)taylor_coef(1);
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
      const int lr = rb + b;
      T x = cz<T>();
      if (lr < nrows) {
        const int p = lr / a.NY;
        const S h = (S)(0.5f * a.w[plane_iw(a, p)] * a.dz);   // scatter.cu:197-238
        const T v = CX<T>::mk(-h * u[e].y, h * u[e].x);
        T rsq, inv;
        scatter_factors(a, model_iw<MAP>(a, plane_iw(a, p)), lr - p * a.NY, i, rsq, inv);
        x = cscale(cmulc(v, cmul(rsq, inv)), c);
      }
      u[e] = x;
    }
    F::run(u, wl, twl, lane);
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
      const int lr = rb + b;
      if (lr < nrows) a.W[widx(a, lr, i, N)] = to_f2(cscale(u[e], a.scale));
    }
    return;
  } else {
    T pw[C::V], inv[C::V];
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
      const int lr = rb + b;
      pw[e] = cz<T>();
      inv[e] = cz<T>();
      if (lr < nrows) {
        const int p = lr / a.NY;
        const S h = (S)(0.5f * a.w[plane_iw(a, p)] * a.dz);   // scatter.cu:197-238
        u[e] = CX<T>::mk(-h * u[e].y, h * u[e].x);
        scatter_factors(a, model_iw<MAP>(a, plane_iw(a, p)), lr - p * a.NY, i, pw[e], inv[e]);
        pw[e] = cmul(pw[e], inv[e]);                          // start at order 1
      }
    }
    for (int it = 1; it < a.nterm; ++it) {
      T x[C::V];
      const S c = (S)taylor_coef(it);
#pragma unroll
      for (int e = 0; e < C::V; ++e) {
        x[e] = cscale(cmulc(u[e], pw[e]), c);                // conj(c_it s~^(-1/2-it)) * u
        pw[e] = cmul(pw[e], inv[e]);
      }
      F::run(x, wl, twl, lane);
      float2* W = a.W + (size_t)(it - 1) * a.wstride;
#pragma unroll
      for (int e = 0; e < C::V; ++e) {
        int b, i;
        edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
        const int lr = rb + b;
        if (lr < nrows) W[widx(a, lr, i, N)] = to_f2(cscale(x[e], a.scale));
      }
    }
  }
}

// Scatter adjoint, second half:
//   x = conj(c_0 s~^(-1/2)) * (+i w dz / 2) * in  +  scale * F_x^-1(K)   (K only for taylor >= 1)
// IC = true : ds[iw, y, x] (+)= conj(U) * x on the model region (into aux).
//             Launched over planes of a single shot, so every model sample
//             is owned by exactly one lane: no atomics, deterministic (the
//             reference's ic_adj uses float atomicAdd, imaging_condition.cu:103-104).
// IC = false: out = (add ? out : 0) + x  (standalone Scatter adjoint).
template <int N, bool IC, bool MAP, class T = float2>
__global__ __launch_bounds__(kThreads) void k_rows_adj_out(KArgs a) {
  using C = BRowCfg<N>;
  using S = real_of<T>;
  using Fi = WaveFft<N, C::NB, false, true, PadLds<N>, T>;
  __shared__ T lds[kWaves * C::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrows = a.np * a.NY;
  const int rb = (blockIdx.x * kWaves + wave) * C::NB;
  // twiddles in LDS: a global twiddle load would wait (one counter for all
  // vector memory operations) for every load still in flight
  __shared__ T twl[N];
  const T* twg = twx_of(a, (T*)nullptr);
  for (int i = threadIdx.x; i < N; i += kThreads) twl[i] = twg[i];
  __syncthreads();
  if (rb >= nrows) return;
  if constexpr (IC) {
    // the imaging condition writes the model region only: rows of the
    // padding (y >= ny) have nothing to produce, not even their transform
    bool any = false;
    for (int b = 0; b < C::NB; ++b) {
      const int lr = rb + b;
      if (lr < nrows && lr % a.NY < a.ny) any = true;
    }
    if (!any) return;
  }
  T x[C::V];
  if (a.nterm > 1) {
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, Fi::P::first(), false>(lane, e, b, i);
      const int lr = rb + b;
      x[e] = lr < nrows ? from_f2<T>(a.K[widx(a, lr, i, N)]) : cz<T>();
    }
    Fi::run(x, lds + wave * C::LDSW, twl, lane);
  } else {
#pragma unroll
    for (int e = 0; e < C::V; ++e) x[e] = cz<T>();
  }
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, Fi::P::last(), false>(lane, e, b, i);
    const int lr = rb + b;
    if (lr < nrows) {
      const int p = lr / a.NY, iy = lr - p * a.NY;
      const int iw = plane_iw(a, p);
      const unsigned idx = lr * N + i;
      T rsq, inv;
      scatter_factors(a, model_iw<MAP>(a, iw), iy, i, rsq, inv);
      const S h = (S)(0.5f * a.w[iw] * a.dz);
      const T d = from_f2<T>(a.in[idx]);
      const T u = CX<T>::mk(-h * d.y, h * d.x);
      const T y = cadd(cscale(x[e], a.scale), cmulc(u, rsq));
      if constexpr (IC) {
        if (iy < a.ny && i < a.nx) {
          const unsigned m = (iw * a.ny + iy) * a.nx + i;
          const T g = cmulc(y, from_f2<T>(a.bg[idx]));
          a.aux[m] = to_f2(a.add ? cadd(from_f2<T>(a.aux[m]), g) : g);
        }
      } else {
        a.out[idx] = to_f2(a.add ? cadd(from_f2<T>(a.out[idx]), y) : y);
      }
    }
  }
}

// ------------------------------------------------------------------ reflection coefficients
// principal square root of a complex number without cancellation: the
// larger part from sqrt((|z| + |x|)/2), the other as y / (2 * that)
__device__ __forceinline__ float2 csqrt_principal(float2 z) {
  const float m = hypotf(z.x, z.y);
  const float t = sqrtf(0.5f * (m + fabsf(z.x)));
  if (t == 0.f) return czero();
  const float o = 0.5f * fabsf(z.y) / t;
  return z.x >= 0.f ? make_float2(t, copysignf(o, z.y)) : make_float2(o, copysignf(t, z.y));
}
__device__ __forceinline__ float2 cdiv(float2 a, float2 b) {
  const float q = 1.f / fmaf(b.x, b.x, b.y * b.y);
  return make_float2(fmaf(a.x, b.x, a.y * b.y) * q, fmaf(a.y, b.x, -a.x * b.y) * q);
}
// r_z = (rho_{z+1} sqrt(s_z) - rho_z sqrt(s_{z+1})) / (rho_{z+1} sqrt(s_z) + rho_z sqrt(s_{z+1})),
// r_{nz-1} = 0 (the last slice is paired with itself); arrays (nz, NW, ny, nx)
__global__ __launch_bounds__(kThreads) void k_reflect_coef(const float2* __restrict__ slow, const float2* __restrict__ den,
                                                           float2* __restrict__ r, int nz, size_t slice) {
  const size_t n = (size_t)nz * slice;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
    if (i + slice >= n) { r[i] = czero(); continue; }
    const float2 a = cmul(den[i + slice], csqrt_principal(slow[i]));
    const float2 b = cmul(den[i], csqrt_principal(slow[i + slice]));
    r[i] = cdiv(csub(a, b), cadd(a, b));
  }
}

// Reflection coefficient and linearized reflection at one model point, from
// (s, rho) at depths z and z+1 (reflect.cu:42-45, lin_reflect.cu:45-90):
// r = (rho1 q0 - rho0 q1) / D, D = rho1 q0 + rho0 q1, q = sqrt(s);
// dr = a ds0 - b ds1 + c0 drho1 - c1 drho0 with A = rho0 rho1 / D^2,
// B = 2 q0 q1 / D^2, a = A q1/q0, b = A q0/q1, c0 = rho0 B, c1 = rho1 B
// (the exact derivative of r; the reference kernel's slowness terms use the
// square root of q1/q0).
struct ReflCoefs {
  float2 r, a, b, c0, c1;
  __device__ __forceinline__ ReflCoefs(float2 s0, float2 s1, float2 rho0, float2 rho1) {
    const float2 q0 = csqrt_principal(s0), q1 = csqrt_principal(s1);
    const float2 u = cmul(rho1, q0), v = cmul(rho0, q1);
    const float2 D = cadd(u, v);
    r = cdiv(csub(u, v), D);
    const float2 iD = cdiv(make_float2(1.f, 0.f), D), iD2 = cmul(iD, iD);
    const float2 A = cmul(cmul(rho0, rho1), iD2);
    const float2 B = cscale(cmul(cmul(q0, q1), iD2), 2.f);
    const float2 c = cdiv(q1, q0);
    a = cmul(A, c);
    b = cdiv(A, c);
    c0 = cmul(rho0, B);
    c1 = cmul(rho1, B);
  }
};
// MODE 0: out += r U (background up pass); MODE 1: out += r dU + U L(ds, drho)
// (down-reflected + back-scattering sources); MODE 2: out += conj(r) U
// (the channel-2 adjoint source, U = mu_z).  One thread per model point,
// looping over shots.
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_refl_src(ReflArgs a) {
  const size_t n = (size_t)a.nx * a.ny * a.NW;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < n; t += (size_t)gridDim.x * kThreads) {
    if (a.last) return;
    const int ix = (int)(t % a.nx);
    const int iy = (int)((t / a.nx) % a.ny);
    const int iw = (int)(t / ((size_t)a.nx * a.ny));
    const ReflCoefs c(a.s0[t], a.s1[t], a.r0[t], a.r1[t]);
    float2 L = czero();
    if constexpr (MODE == 1)
      L = csub(cadd(cmul(c.a, a.ds0[t]), cmul(c.c0, a.dr1[t])), cadd(cmul(c.b, a.ds1[t]), cmul(c.c1, a.dr0[t])));
    for (int is = 0; is < a.NS; ++is) {
      const size_t idx = (((size_t)is * a.NW + iw) * a.NY + iy) * a.NX + ix;
      float2 y = a.out[idx];
      if constexpr (MODE == 0) y = cadd(y, cmul(c.r, a.U[idx]));
      else if constexpr (MODE == 1) y = cadd(y, cadd(cmul(c.r, a.dU[idx]), cmul(a.U[idx], L)));
      else y = cadd(y, cmulc(a.U[idx], c.r));
      a.out[idx] = y;
    }
  }
}
// adjoint of the back-scattering source: g = sum_s conj(U) mu at a model point,
// gs0 += conj(a) g, gs1 -= conj(b) g, gr1 += conj(c0) g, gr0 -= conj(c1) g
__global__ __launch_bounds__(kThreads) void k_refl_adj(ReflArgs a) {
  const size_t n = (size_t)a.nx * a.ny * a.NW;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < n; t += (size_t)gridDim.x * kThreads) {
    if (a.last) return;
    const int ix = (int)(t % a.nx);
    const int iy = (int)((t / a.nx) % a.ny);
    const int iw = (int)(t / ((size_t)a.nx * a.ny));
    float2 g = czero();
    for (int is = 0; is < a.NS; ++is) {
      const size_t idx = (((size_t)is * a.NW + iw) * a.NY + iy) * a.NX + ix;
      g = cadd(g, cmulc(a.dU[idx], a.U[idx]));  // dU carries mu here
    }
    const ReflCoefs c(a.s0[t], a.s1[t], a.r0[t], a.r1[t]);
    a.gs0[t] = cadd(a.gs0[t], cmulc(g, c.a));
    a.gs1[t] = csub(a.gs1[t], cmulc(g, c.b));
    a.gr1[t] = cadd(a.gr1[t], cmulc(g, c.c0));
    a.gr0[t] = csub(a.gr0[t], cmulc(g, c.c1));
  }
}

// ------------------------------------------------------------------ point-wise Born helpers
// img (+)= sum_{s, w} conj(U) * R over the model region: one thread per (y, x),
// fixed summation order (the reduction of the migration image).
__global__ __launch_bounds__(kThreads) void k_image(const float2* __restrict__ U, const float2* __restrict__ R,
                                                    float2* img, int NX, int NY, int NW, int NS, int nx, int ny, int add) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= nx * ny) return;
  const int ix = t % nx, iy = t / nx;
  float2 acc = czero();
  for (int p = 0; p < NW * NS; ++p) {
    const size_t idx = ((size_t)p * NY + iy) * NX + ix;
    acc = cadd(acc, cmulc(R[idx], U[idx]));
  }
  img[t] = add ? cadd(img[t], acc) : acc;
}

// Standalone imaging condition (propagator/src/imaging_condition.cu):
// forward D = (add ? D : 0) + U * img on the model region (padding: 0 or kept)
__global__ __launch_bounds__(kThreads) void k_ic_fwd(const float2* __restrict__ img, float2* D, const float2* __restrict__ U,
                                                     int NX, int NY, int NW, int NS, int nx, int ny, int add) {
  const size_t n = (size_t)NX * NY * NW * NS;
  for (size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x; idx < n; idx += (size_t)gridDim.x * kThreads) {
    const int ix = (int)(idx % NX);
    const size_t t = idx / NX;
    const int iy = (int)(t % NY);
    const int iw = (int)((t / NY) % NW);
    float2 y = add ? D[idx] : czero();
    if (ix < nx && iy < ny) y = cadd(y, cmul(U[idx], img[((size_t)iw * ny + iy) * nx + ix]));
    D[idx] = y;
  }
}
// adjoint img[w, y, x] (+)= sum_s conj(U) * D, one thread per image sample
__global__ __launch_bounds__(kThreads) void k_ic_adj(float2* img, const float2* __restrict__ D, const float2* __restrict__ U,
                                                     int NX, int NY, int NW, int NS, int nx, int ny, int add) {
  const size_t n = (size_t)nx * ny * NW;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < n; t += (size_t)gridDim.x * kThreads) {
    const int ix = (int)(t % nx);
    const int iy = (int)((t / nx) % ny);
    const int iw = (int)(t / ((size_t)nx * ny));
    float2 acc = czero();
    for (int s = 0; s < NS; ++s) {
      const size_t idx = (((size_t)s * NW + iw) * NY + iy) * NX + ix;
      acc = cadd(acc, cmulc(D[idx], U[idx]));
    }
    img[t] = add ? cadd(img[t], acc) : acc;
  }
}

// ------------------------------------------------------------------ injection
// Forward (traces -> wavefield, propagator/src/injection.cu:7-96) as a gather
// over precomputed target nodes: every node of this depth level owns a
// segment of (trace, weight) contributions, summed in a fixed order, so the
// scatter is race-free and deterministic (the reference's kernel does a
// non-atomic read-modify-write, injection.cu:89-92).
__global__ __launch_bounds__(kThreads) void k_inj_fwd(const float2* __restrict__ tr, float2* wf, const long long* __restrict__ seg_base,
                                                      const int* __restrict__ seg_off, const int* __restrict__ seg_np,
                                                      const int* __restrict__ ent_trace, const float* __restrict__ ent_w,
                                                      const long long* __restrict__ toff, int s0, int nseg, int wmax,
                                                      long long plane) {
  const long long n = (long long)nseg * wmax;
  for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < n; t += (long long)gridDim.x * kThreads) {
    const int sg = s0 + (int)(t / wmax), iw = (int)(t % wmax);
    if (iw >= seg_np[sg]) continue;
    float2 acc = czero();
    for (int k = seg_off[sg]; k < seg_off[sg + 1]; ++k)
      acc = cadd(acc, cscale(tr[toff[ent_trace[k]] + iw], ent_w[k]));
    const long long idx = seg_base[sg] + (long long)iw * plane;
    wf[idx] = cadd(wf[idx], acc);
  }
}
// Adjoint (wavefield -> traces, injection.cu:98-187): each active trace gathers
// its four corners.
__global__ __launch_bounds__(kThreads) void k_inj_adj(float2* tr, const float2* __restrict__ wf, const int* __restrict__ act_trace,
                                                      const long long* __restrict__ act_base, const float4* __restrict__ act_w,
                                                      const int* __restrict__ act_np, const long long* __restrict__ toff,
                                                      int a0, int nact, int wmax, int NX, long long plane) {
  const long long n = (long long)nact * wmax;
  for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < n; t += (long long)gridDim.x * kThreads) {
    const int k = a0 + (int)(t / wmax), iw = (int)(t % wmax);
    if (iw >= act_np[k]) continue;
    const long long b = act_base[k] + (long long)iw * plane;
    const float4 w = act_w[k];
    float2 v = cscale(wf[b], w.x);
    v = cadd(v, cscale(wf[b + 1], w.y));
    v = cadd(v, cscale(wf[b + NX], w.z));
    v = cadd(v, cscale(wf[b + NX + 1], w.w));
    const long long o = toff[act_trace[k]] + iw;
    tr[o] = cadd(tr[o], v);
  }
}

// Shot sums of a per-plane gradient: out[k, j] += sum over the planes p of
// shot k, in plane order, of g[p, j] (shot k owns planes first[k] ..
// first[k] + count[k] - 1).
__global__ __launch_bounds__(kThreads) void k_plane_sum(const float2* __restrict__ g, float2* out,
                                                        const int* __restrict__ first, const int* __restrict__ count,
                                                        int nshot, long long m) {
  const long long n = (long long)nshot * m;
  for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < n; t += (long long)gridDim.x * kThreads) {
    const int k = (int)(t / m);
    const long long j = t - (long long)k * m;
    const float2* src = g + (size_t)first[k] * m + j;
    float2 acc = out[t];
    // eight planes' loads in flight per thread; the sum keeps plane order
    const int c = count[k];
    int p = 0;
    for (; p + 8 <= c; p += 8) {
      float2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = src[(size_t)(p + q) * m];
#pragma unroll
      for (int q = 0; q < 8; ++q) acc = cadd(acc, v[q]);
    }
    for (; p < c; ++p) acc = cadd(acc, src[(size_t)p * m]);
    out[t] = acc;
  }
}

// ------------------------------------------------------------------ launchers
template <class Fn>
static hipError_t bdispatch(int n, Fn&& fn) {
  switch (n) {
#define CASE(S) case S: return fn(std::integral_constant<int, S>{});
    CASE(8) CASE(12) CASE(16) CASE(24) CASE(32) CASE(48) CASE(64) CASE(96) CASE(128) CASE(192)
    CASE(256) CASE(384) CASE(512) CASE(768) CASE(1024)
#undef CASE
    default: return hipErrorInvalidValue;
  }
}
static dim3 brow_grid(int N, int nrows) {
  const int nb = row_nb(N);
  const int waves = (nrows + nb - 1) / nb;
  return dim3((waves + kWaves - 1) / kWaves);
}
static inline dim3 bflat(long long n) {
  long long g = (n + kThreads - 1) / kThreads;
  if (g > (1ll << 30)) g = 1ll << 30;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

// Born / scattering row passes; the shot-batch flavour (a.mplane set) is a
// separate instantiation (MAP).  ScatterOp and the unfused forms never map.
// F64: a float64-arithmetic instantiation exists (used when a.f64 is set;
// the fused Born forms only).
template <template <int, bool, bool, class> class K, bool FLAG, bool F64 = false>
static hipError_t launch_mapped(const KArgs& a, hipStream_t st) {
  return bdispatch(a.NX, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    const dim3 g = brow_grid(N, a.np * a.NY);
    if constexpr (F64) {
      if (a.f64) {
        if (a.mplane) hipLaunchKernelGGL((K<N, FLAG, true, double2>::fn), g, dim3(kThreads), 0, st, a);
        else hipLaunchKernelGGL((K<N, FLAG, false, double2>::fn), g, dim3(kThreads), 0, st, a);
        return hipGetLastError();
      }
    }
    if (a.f64) return hipErrorInvalidValue;
    if (a.mplane) hipLaunchKernelGGL((K<N, FLAG, true, float2>::fn), g, dim3(kThreads), 0, st, a);
    else hipLaunchKernelGGL((K<N, FLAG, false, float2>::fn), g, dim3(kThreads), 0, st, a);
    return hipGetLastError();
  });
}
template <int N, bool F, bool M, class T> struct KIcFwd { static constexpr auto fn = k_rows_ic_fwd<N, M>; };
template <int N, bool F, bool M, class T> struct KScatter { static constexpr auto fn = k_rows_scatter<N, F, M, T>; };
template <int N, bool F, bool M, class T> struct KScatter1 {
  static constexpr auto fn = k_rows_scatter<N, F, M, T, true>;
};
template <int N, bool F, bool M, class T> struct KScatterAdj { static constexpr auto fn = k_rows_scatter_adj<N, F, M, T>; };
template <int N, bool F, bool M, class T> struct KScatterAdj1 {
  static constexpr auto fn = k_rows_scatter_adj<N, F, M, T, true>;
};
template <int N, bool F, bool M, class T> struct KAdjOut { static constexpr auto fn = k_rows_adj_out<N, F, M, T>; };

hipError_t launch_rows_ic_fwd(const KArgs& a, hipStream_t st) {
  KArgs b = a;
  b.f64 = 0;  // the imaging-condition product and the first row transform keep float32 arithmetic
  return launch_mapped<KIcFwd, true>(b, st);
}
hipError_t launch_rows_scatter(bool born, const KArgs& a, hipStream_t st) {
  if (a.nterm == 2)
    return born ? launch_mapped<KScatter1, true, true>(a, st) : launch_mapped<KScatter1, false>(a, st);
  return born ? launch_mapped<KScatter, true, true>(a, st) : launch_mapped<KScatter, false>(a, st);
}
hipError_t launch_rows_scatter_adj(bool fused, const KArgs& a, hipStream_t st) {
  if (a.nterm == 2)
    return fused ? launch_mapped<KScatterAdj1, true, true>(a, st) : launch_mapped<KScatterAdj1, false>(a, st);
  return fused ? launch_mapped<KScatterAdj, true, true>(a, st) : launch_mapped<KScatterAdj, false>(a, st);
}
hipError_t launch_rows_adj_out(bool ic, const KArgs& a, hipStream_t st) {
  return ic ? launch_mapped<KAdjOut, true, true>(a, st) : launch_mapped<KAdjOut, false>(a, st);
}
hipError_t launch_refl_src(int mode, const ReflArgs& a, hipStream_t st) {
  const dim3 g = bflat((long long)a.nx * a.ny * a.NW);
  if (mode == 0) hipLaunchKernelGGL(k_refl_src<0>, g, dim3(kThreads), 0, st, a);
  else if (mode == 1) hipLaunchKernelGGL(k_refl_src<1>, g, dim3(kThreads), 0, st, a);
  else hipLaunchKernelGGL(k_refl_src<2>, g, dim3(kThreads), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_refl_adj(const ReflArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_refl_adj, bflat((long long)a.nx * a.ny * a.NW), dim3(kThreads), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_reflect_coef(const float2* slow, const float2* den, float2* r, int nz, size_t slice, hipStream_t st) {
  hipLaunchKernelGGL(k_reflect_coef, bflat((long long)nz * (long long)slice), dim3(kThreads), 0, st, slow, den, r, nz,
                     slice);
  return hipGetLastError();
}
hipError_t launch_image(const float2* U, const float2* R, float2* img, int NX, int NY, int NW, int NS, int nx,
                        int ny, int add, hipStream_t st) {
  hipLaunchKernelGGL(k_image, dim3((nx * ny + kThreads - 1) / kThreads), dim3(kThreads), 0, st, U, R, img, NX, NY, NW,
                     NS, nx, ny, add);
  return hipGetLastError();
}
hipError_t launch_ic_fwd(const float2* img, float2* D, const float2* U, int NX, int NY, int NW, int NS, int nx, int ny,
                         int add, hipStream_t st) {
  hipLaunchKernelGGL(k_ic_fwd, bflat((long long)NX * NY * NW * NS), dim3(kThreads), 0, st, img, D, U, NX, NY, NW, NS,
                     nx, ny, add);
  return hipGetLastError();
}
hipError_t launch_ic_adj(float2* img, const float2* D, const float2* U, int NX, int NY, int NW, int NS, int nx, int ny,
                         int add, hipStream_t st) {
  hipLaunchKernelGGL(k_ic_adj, bflat((long long)nx * ny * NW), dim3(kThreads), 0, st, img, D, U, NX, NY, NW, NS, nx,
                     ny, add);
  return hipGetLastError();
}
hipError_t launch_plane_sum(const float2* g, float2* out, const int* first, const int* count, int nshot, long long m,
                            hipStream_t st) {
  if (nshot <= 0 || m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_plane_sum, bflat((long long)nshot * m), dim3(kThreads), 0, st, g, out, first, count, nshot, m);
  return hipGetLastError();
}
hipError_t launch_inj_fwd(const float2* tr, float2* wf, const long long* seg_base, const int* seg_off,
                          const int* seg_np, const int* ent_trace, const float* ent_w, const long long* toff, int s0,
                          int nseg, int wmax, long long plane, hipStream_t st) {
  if (nseg <= 0 || wmax <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_inj_fwd, bflat((long long)nseg * wmax), dim3(kThreads), 0, st, tr, wf, seg_base, seg_off,
                     seg_np, ent_trace, ent_w, toff, s0, nseg, wmax, plane);
  return hipGetLastError();
}
hipError_t launch_inj_adj(float2* tr, const float2* wf, const int* act_trace, const long long* act_base,
                          const float4* act_w, const int* act_np, const long long* toff, int a0, int nact, int wmax,
                          int NX, long long plane, hipStream_t st) {
  if (nact <= 0 || wmax <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_inj_adj, bflat((long long)nact * wmax), dim3(kThreads), 0, st, tr, wf, act_trace, act_base, act_w,
                     act_np, toff, a0, nact, wmax, NX, plane);
  return hipGetLastError();
}

}  // namespace fwix
