// Fused one-way depth-step kernels: three passes per PSPI / NSPS step.
//
//   PSPI forward   (propagator/src/PSPI.cpp:6-24, :66-83)
//     rows_fwd(pre-taper)  ->  cols_fanout(Phi_r)  ->  rows_select
//   PSPI adjoint   (PSPI.cpp:85-103)
//     rows_fanout(mask)    ->  cols_fanin(Phi_r^adj) -> rows_inv(post-taper)
//   NSPS forward   (propagator/src/NSPS.cpp:27-45)
//     rows_fanout(mask)    ->  cols_fanin(Phi_r)    -> rows_inv
//   NSPS adjoint   (corrected; NSPS.cpp:6-25 passes swapped arguments)
//     rows_fwd             ->  cols_fanout(Phi_r^adj) -> rows_select
//
// Rows are transformed one wavefront per row (or per few short rows), loaded
// and stored straight from/to HBM in the positions the first/last Stockham
// stage needs.  Columns are staged through a [NY][CW+1] LDS tile so global
// accesses stay 128-byte coalesced.  The per-reference fan-out keeps the
// spectrum of a column tile in registers and runs all nref inverse
// transforms from it; the fan-in accumulates all nref spectra in registers
// before a single inverse transform.
#include "fwix_kernels.h"
#include "fwix_launch.h"

#include <type_traits>

namespace fwix {

template <int N>
struct RowCfg {
  static constexpr int NB = row_nb(N);
  static constexpr int V = NB * N / 64;
  static constexpr int LDSW = NB * padded(N);
};
// Column passes: 16 waves x NB columns.  The [N][CW+1] tile that stages the
// coalesced loads / stores is reused, between block barriers, as eight
// private padded exchange regions (one per wave) for the Stockham stages.
template <int N>
struct ColCfg {
  static constexpr int THREADS = kColThreads;
  static constexpr int WAVES = THREADS / 64;
  static constexpr int CW = col_cw(N);
  static constexpr int NB = CW / WAVES;
  static constexpr int V = NB * N / 64;
  static constexpr int CWP = CW + 1;
  static constexpr int NROW = N;
  static constexpr int TILE = N * CWP;
  static constexpr int PRIV = NB * padded(N);
  static constexpr int SMEM = TILE > WAVES * PRIV ? TILE : WAVES * PRIV;
  static_assert(NB >= 1 && V >= 4, "column tile too narrow");
};

// ------------------------------------------------------------------ row passes
// out = (add ? out : 0) + post(scale * F_x(pre * in))   along x, one row per wave
template <int N, bool INV, bool PRE_TAPER, bool POST_TAPER>
__global__ __launch_bounds__(kThreads) void k_rows(KArgs a) {
  using C = RowCfg<N>;
  using F = WaveFft<N, C::NB, false, INV, PadLds<N>>;
  __shared__ float2 lds[kWaves * C::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrows = a.np * a.NY;
  const int rb = (blockIdx.x * kWaves + wave) * C::NB;
  if (rb >= nrows) return;
  float2* wl = lds + wave * C::LDSW;
  float2 v[C::V];
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
    const int lr = rb + b;
    float2 x = czero();
    if (lr < nrows) {
      x = a.in[(size_t)lr * N + i];
      if (PRE_TAPER) x = cscale(x, a.tapx[i] * a.tapy[lr % a.NY]);
    }
    v[e] = x;
  }
  F::run(v, wl, a.twx, lane);
  const float s = a.scale;  // row normalisation 1/sqrt(NX)
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
    const int lr = rb + b;
    if (lr < nrows) {
      const size_t idx = (size_t)lr * N + i;
      float2 y = cscale(v[e], s);
      if (POST_TAPER) y = cscale(y, a.tapx[i] * a.tapy[lr % a.NY]);
      if (a.add) y = cadd(y, a.out[idx]);
      a.out[idx] = y;
    }
  }
}

// rows_fanout: W_r = scale * F_x([labels == r] * in) for every reference r
template <int N>
__global__ __launch_bounds__(kThreads) void k_rows_fanout(KArgs a) {
  using C = RowCfg<N>;
  using F = WaveFft<N, C::NB, false, false, PadLds<N>>;
  __shared__ float2 lds[kWaves * C::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrows = a.np * a.NY;
  const int rb = (blockIdx.x * kWaves + wave) * C::NB;
  if (rb >= nrows) return;
  float2* wl = lds + wave * C::LDSW;
  float2 u[C::V];
  uint8_t lab[C::V];
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
    const int lr = rb + b;
    u[e] = czero();
    lab[e] = 255;
    if (lr < nrows) {
      u[e] = a.in[(size_t)lr * N + i];
      const int p = lr / a.NY, iy = lr - p * a.NY;
      lab[e] = a.labels[((size_t)plane_iw(a, p) * a.NY + iy) * N + i];
    }
  }
  const float s = a.scale;
  uint32_t wm = 0xffffffffu;
  if (a.rowmask) {
    wm = 0;
    for (int b = 0; b < C::NB; ++b) {
      const int lr = rb + b;
      if (lr < nrows) {
        const int p = lr / a.NY;
        wm |= a.rowmask[plane_iw(a, p) * a.NY + (lr - p * a.NY)];
      }
    }
    wm = __builtin_amdgcn_readfirstlane(wm);
  }
  for (int r = 0; r < a.nref; ++r) {
    if (r < 32 && !((wm >> r) & 1u)) continue;  // rows without reference r: the column pass reads zeros
    float2 x[C::V];
#pragma unroll
    for (int e = 0; e < C::V; ++e) x[e] = (lab[e] == r) ? u[e] : czero();
    F::run(x, wl, a.twx, lane);
    float2* W = a.W + (size_t)r * a.wstride;
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
      const int lr = rb + b;
      if (lr < nrows) W[(size_t)lr * N + i] = cscale(x[e], s);
    }
  }
}

// rows_select: out = (add ? out : 0) + sum_r [labels == r] * scale * F_x^-1(W_r)
template <int N>
__global__ __launch_bounds__(kThreads) void k_rows_select(KArgs a) {
  using C = RowCfg<N>;
  using F = WaveFft<N, C::NB, false, true, PadLds<N>>;
  __shared__ float2 lds[kWaves * C::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrows = a.np * a.NY;
  const int rb = (blockIdx.x * kWaves + wave) * C::NB;
  if (rb >= nrows) return;
  float2* wl = lds + wave * C::LDSW;
  uint8_t lab[C::V];
  float2 acc[C::V];
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
    const int lr = rb + b;
    acc[e] = czero();
    lab[e] = 255;
    if (lr < nrows) {
      const int p = lr / a.NY, iy = lr - p * a.NY;
      lab[e] = a.labels[((size_t)plane_iw(a, p) * a.NY + iy) * N + i];
    }
  }
  // references present in this wave's rows (wave-uniform)
  uint32_t wm = 0xffffffffu;
  if (a.rowmask) {
    wm = 0;
    for (int b = 0; b < C::NB; ++b) {
      const int lr = rb + b;
      if (lr < nrows) {
        const int p = lr / a.NY;
        wm |= a.rowmask[plane_iw(a, p) * a.NY + (lr - p * a.NY)];
      }
    }
    wm = __builtin_amdgcn_readfirstlane(wm);
  }
  const uint32_t live = a.nref >= 32 ? wm : (wm & ((1u << a.nref) - 1u));
  float2 nxt[C::V];
  auto fetch = [&](int r) {
    const float2* W = a.W + (size_t)r * a.wstride;
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
      const int lr = rb + b;
      nxt[e] = lr < nrows ? W[(size_t)lr * N + i] : czero();
    }
  };
  // absent references contribute nothing to these rows: skip their loads and transforms
  int r = a.rowmask ? next_bit(live, -1) : 0;
  const int rend = a.rowmask ? 32 : a.nref;
  if (r < rend) fetch(r);
  while (r < rend) {
    float2 x[C::V];
#pragma unroll
    for (int e = 0; e < C::V; ++e) x[e] = nxt[e];
    const int rn = a.rowmask ? next_bit(live, r) : r + 1;
    if (rn < rend) fetch(rn);
    F::run(x, wl, a.twx, lane);
#pragma unroll
    for (int e = 0; e < C::V; ++e)
      if (lab[e] == r) acc[e] = cadd(acc[e], x[e]);
    r = rn;
  }
  const float s = a.scale;
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
    const int lr = rb + b;
    if (lr < nrows) {
      const size_t idx = (size_t)lr * N + i;
      float2 y = cscale(acc[e], s);
      if (a.add) y = cadd(y, a.out[idx]);
      a.out[idx] = y;
    }
  }
}

// ------------------------------------------------------------------ column passes
// Pointwise spectral multiplier of term r at (kx, ky) of frequency iw.
// MULT 0 / 1: forward / adjoint phase shift of reference slowness r
// (propagator/src/phase_shift.cu); MULT 2: Born scattering wavenumber factor
// (|k|/w)^(2r) = (k^2/w^2)^r (propagator/src/scatter.cu:23-59).
// The term's uniform part (reference slowness, real/complex branch) is set up
// once per (plane, term) so the unrolled per-sample loops carry no divergent
// or if-converted paths.
template <int MULT>
struct TermMult {
  float w2, wsre, b, iw2, dz;
  int r;
  bool real;
  __device__ __forceinline__ TermMult(const KArgs& a, int r_, int iw, float w2_) : w2(w2_), r(r_), dz(a.dz) {
    if constexpr (MULT == 2) {
      iw2 = 1.f / w2;
      real = true; wsre = 0.f; b = 0.f;
    } else {
      const float2 sr = a.sref[r * a.NW + iw];
      wsre = w2 * sr.x;
      b = w2 * (sr.y - a.eps * sr.x);
      real = b == 0.f;
      iw2 = 0.f;
    }
  }
  template <bool REAL>
  __device__ __forceinline__ float2 at(float kxv, float kyv) const {
    const float k2 = fmaf(kxv, kxv, kyv * kyv);
    if constexpr (MULT == 2) {
      const float q = k2 * iw2;
      float m = 1.f;
      for (int i = 0; i < r; ++i) m *= q;
      return make_float2(m, 0.f);
    } else if constexpr (REAL) {
      return ps_real<MULT>(wsre - k2, dz);
    } else {
      return ps_general<MULT>(wsre - k2, b, dz);
    }
  }
};

// Coalesced tile load / store: every thread first issues all of its global
// loads into registers (PER = N*CW/THREADS of them), then writes LDS, so the
// memory latency is paid once per tile rather than once per element.
template <class C>
struct TileRegs {
  static constexpr int PER = C::NROW * C::CW / C::THREADS;
  static_assert(PER * C::THREADS == C::NROW * C::CW, "tile must divide evenly over the threads");
  float2 t[PER];
  __device__ __forceinline__ void fetch(const float2* __restrict__ src, int NX, int c0) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + k * C::THREADS;
      const int i = e / C::CW, c = e % C::CW, col = c0 + c;
      t[k] = col < NX ? src[(size_t)i * NX + col] : czero();
    }
  }
  // rows whose mask bit is clear are known to be zero and are not loaded
  __device__ __forceinline__ void fetch_masked(const float2* __restrict__ src, int NX, int c0,
                                               const uint32_t* __restrict__ rm, uint32_t bit) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + k * C::THREADS;
      const int i = e / C::CW, c = e % C::CW, col = c0 + c;
      t[k] = (col < NX && (rm[i] & bit)) ? src[(size_t)i * NX + col] : czero();
    }
  }
  __device__ __forceinline__ void commit(float2* tile) const {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + k * C::THREADS;
      tile[(e / C::CW) * C::CWP + e % C::CW] = t[k];
    }
  }
};
template <class C>
__device__ __forceinline__ void tile_load(float2* tile, const float2* __restrict__ src, int NX, int c0) {
  TileRegs<C> r;
  r.fetch(src, NX, c0);
  r.commit(tile);
}
// store only the rows whose mask bit is set (rows that a later pass reads)
template <class C>
__device__ __forceinline__ void tile_store_masked(const float2* tile, float2* __restrict__ dst, int NX, int c0,
                                                  const uint32_t* __restrict__ rm, uint32_t bit) {
  constexpr int PER = C::NROW * C::CW / C::THREADS;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = threadIdx.x + k * C::THREADS;
    const int i = e / C::CW, col = c0 + e % C::CW;
    if (col < NX && (rm[i] & bit)) dst[(size_t)i * NX + col] = tile[i * C::CWP + e % C::CW];
  }
}
template <class C>
__device__ __forceinline__ void tile_store(const float2* tile, float2* __restrict__ dst, int NX, int c0, int add) {
  constexpr int PER = C::NROW * C::CW / C::THREADS;
  float2 t[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = threadIdx.x + k * C::THREADS;
    t[k] = tile[(e / C::CW) * C::CWP + e % C::CW];
  }
  if (add) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + k * C::THREADS;
      const int col = c0 + e % C::CW;
      if (col < NX) t[k] = cadd(t[k], dst[(size_t)(e / C::CW) * NX + col]);
    }
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = threadIdx.x + k * C::THREADS;
    const int col = c0 + e % C::CW;
    if (col < NX) dst[(size_t)(e / C::CW) * NX + col] = t[k];
  }
}
// this wave's columns at the first-stage positions of radix R, from the tile
template <int N, class C, int R>
__device__ __forceinline__ void tile_read(float2* v, const float2* tile, int wave, int lane) {
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, R, false>(lane, e, b, i);
    v[e] = tile[i * C::CWP + wave * C::NB + b];
  }
}
template <int N, class C, int R>
__device__ __forceinline__ void tile_write(float2* tile, const float2* v, int wave, int lane) {
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, R, false>(lane, e, b, i);
    tile[i * C::CWP + wave * C::NB + b] = v[e];
  }
}

// out = (add ? out : 0) + scale * F_y(in) on each plane (plain 2-D FFT column pass)
template <int N, bool INV>
__global__ __launch_bounds__(ColCfg<N>::THREADS) void k_cols(KArgs a) {
  using C = ColCfg<N>;
  using F = WaveFft<N, C::NB, false, INV, PadLds<N>>;
  __shared__ float2 smem[C::SMEM];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = blockIdx.y, c0 = blockIdx.x * C::CW;
  const size_t pb = (size_t)p * a.NX * N;
  float2* priv = smem + wave * C::PRIV;
  tile_load<C>(smem, a.in + pb, a.NX, c0);
  __syncthreads();
  float2 v[C::V];
  tile_read<N, C, F::P::first()>(v, smem, wave, lane);
  __syncthreads();
  F::run(v, priv, a.twy, lane);
  const float s = a.scale;  // column normalisation 1/sqrt(NY)
#pragma unroll
  for (int e = 0; e < C::V; ++e) v[e] = cscale(v[e], s);
  __syncthreads();
  tile_write<N, C, F::P::last()>(smem, v, wave, lane);
  __syncthreads();
  tile_store<C>(smem, a.out + pb, a.NX, c0, a.add);
}

// cols_fanout: W_r = F_y^-1[M_r * F_y(K)] / NY for every term r (M_r: term_mult<MODE>).
template <int N, int MODE>
__global__ __launch_bounds__(ColCfg<N>::THREADS) void k_cols_fanout(KArgs a) {
  using C = ColCfg<N>;
  using F = WaveFft<N, C::NB, false, false, PadLds<N>>;
  using Fi = WaveFft<N, C::NB, false, true, PadLds<N>>;
  static_assert(F::P::last() == Fi::P::first(), "register hand-off needs mirrored plans");
  __shared__ float2 smem[C::SMEM];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = blockIdx.y, c0 = blockIdx.x * C::CW;
  const size_t pb = (size_t)p * a.NX * N;
  float2* priv = smem + wave * C::PRIV;
  tile_load<C>(smem, a.K + pb, a.NX, c0);
  __syncthreads();
  float2 spec[C::V];
  tile_read<N, C, F::P::first()>(spec, smem, wave, lane);
  __syncthreads();
  F::run(spec, priv, a.twy, lane);
  const int iw = plane_iw(a, p);
  const float w2 = a.w2[iw];
  const float s2 = a.scale;  // 1/NY: forward and inverse column normalisation together
  const int colw = c0 + wave * C::NB;
  const uint32_t* rm = (MODE != 2 && a.rowmask) ? a.rowmask + iw * a.NY : nullptr;
  const uint32_t pm = (MODE != 2 && a.planemask) ? a.planemask[iw] : 0xffffffffu;
  for (int r = 0; r < a.nterm; ++r) {
    if (r < 32 && !((pm >> r) & 1u)) continue;  // reference absent from the plane: W_r is never read
    const TermMult<MODE> tm(a, r, iw, w2);
    float2 x[C::V];
    auto mult = [&](auto real) {
#pragma unroll
      for (int e = 0; e < C::V; ++e) {
        int b, i;
        edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
        const int col = colw + b;
        const float2 ph = tm.template at<decltype(real)::value>(a.kx[col < a.NX ? col : 0], a.ky[i]);
        x[e] = cscale(cmul(spec[e], ph), s2);
      }
    };
    if (tm.real) mult(std::true_type{});
    else mult(std::false_type{});
    Fi::run(x, priv, a.twy, lane);
    __syncthreads();
    tile_write<N, C, Fi::P::last()>(smem, x, wave, lane);
    __syncthreads();
    if (rm) tile_store_masked<C>(smem, a.W + (size_t)r * a.wstride + pb, a.NX, c0, rm, r < 32 ? 1u << r : 0xffffffffu);
    else tile_store<C>(smem, a.W + (size_t)r * a.wstride + pb, a.NX, c0, 0);
    __syncthreads();
  }
}

// cols_fanin: K = F_y^-1[sum_r M_r * F_y(W_r)] / NY (M_r: term_mult<MODE>).
template <int N, int MODE>
__global__ __launch_bounds__(ColCfg<N>::THREADS) void k_cols_fanin(KArgs a) {
  using C = ColCfg<N>;
  using F = WaveFft<N, C::NB, false, false, PadLds<N>>;
  using Fi = WaveFft<N, C::NB, false, true, PadLds<N>>;
  static_assert(F::P::last() == Fi::P::first(), "register hand-off needs mirrored plans");
  __shared__ float2 smem[C::SMEM];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = blockIdx.y, c0 = blockIdx.x * C::CW;
  const size_t pb = (size_t)p * a.NX * N;
  float2* priv = smem + wave * C::PRIV;
  const int iw = plane_iw(a, p);
  const float w2 = a.w2[iw];
  const int colw = c0 + wave * C::NB;
  float2 acc[C::V];
#pragma unroll
  for (int e = 0; e < C::V; ++e) acc[e] = czero();
  const uint32_t* rm = (MODE != 2 && a.rowmask) ? a.rowmask + iw * a.NY : nullptr;
  const uint32_t pm = (MODE != 2 && a.planemask) ? a.planemask[iw] : 0xffffffffu;
  const uint32_t live = a.nterm >= 32 ? pm : (pm & ((1u << a.nterm) - 1u));
  const bool masked = rm != nullptr;
  const int rend = masked ? 32 : a.nterm;
  auto fetch = [&](TileRegs<C>& tr, int r) {
    if (masked) tr.fetch_masked(a.W + (size_t)r * a.wstride + pb, a.NX, c0, rm, 1u << r);
    else tr.fetch(a.W + (size_t)r * a.wstride + pb, a.NX, c0);
  };
  TileRegs<C> next;
  int r = masked ? next_bit(live, -1) : 0;
  if (r < rend) fetch(next, r);
  while (r < rend) {
    next.commit(smem);
    __syncthreads();
    // the next present term's loads are in flight during this term's transform
    const int rn = masked ? next_bit(live, r) : r + 1;
    if (rn < rend) fetch(next, rn);
    float2 v[C::V];
    tile_read<N, C, F::P::first()>(v, smem, wave, lane);
    __syncthreads();
    F::run(v, priv, a.twy, lane);
    const TermMult<MODE> tm(a, r, iw, w2);
    auto mult = [&](auto real) {
#pragma unroll
      for (int e = 0; e < C::V; ++e) {
        int b, i;
        edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
        const int col = colw + b;
        const float2 ph = tm.template at<decltype(real)::value>(a.kx[col < a.NX ? col : 0], a.ky[i]);
        acc[e] = cadd(acc[e], cmul(v[e], ph));
      }
    };
    if (tm.real) mult(std::true_type{});
    else mult(std::false_type{});
    __syncthreads();
    r = rn;
  }
  const float s2 = a.scale;
#pragma unroll
  for (int e = 0; e < C::V; ++e) acc[e] = cscale(acc[e], s2);
  Fi::run(acc, priv, a.twy, lane);
  __syncthreads();
  tile_write<N, C, Fi::P::last()>(smem, acc, wave, lane);
  __syncthreads();
  tile_store<C>(smem, a.K + pb, a.NX, c0, 0);
}

// ------------------------------------------------------------------ launchers
#define FWIX_SIZES(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

bool fft_size_supported(int n) {
  switch (n) {
#define CASE(S) case S:
    FWIX_SIZES(CASE)
#undef CASE
    return true;
    default: return false;
  }
}

int col_tile_width(int n) { return col_cw(n); }

// Presence masks of reference labels: bit r of rowmask[row] is set when
// label r occurs in that (w, y) row; planemask is the OR over a plane's rows.
__global__ __launch_bounds__(kThreads) void k_presence(const uint8_t* __restrict__ labels, uint32_t* rowmask,
                                                       int NX, long long nrows) {
  const long long row = (long long)blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nrows) return;
  uint32_t m = 0;
  for (int x = lane; x < NX; x += 64) {
    const int l = labels[row * NX + x];
    if (l < 32) m |= 1u << l;
  }
  for (int off = 32; off > 0; off >>= 1) m |= (uint32_t)__shfl_xor((int)m, off);
  if (lane == 0) rowmask[row] = m;
}
__global__ __launch_bounds__(kThreads) void k_planemask(const uint32_t* __restrict__ rowmask, uint32_t* planemask,
                                                        int NY, long long nplanes) {
  const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= nplanes) return;
  uint32_t m = 0;
  for (int y = 0; y < NY; ++y) m |= rowmask[p * NY + y];
  planemask[p] = m;
}
hipError_t launch_presence(const uint8_t* labels, uint32_t* rowmask, uint32_t* planemask, int NX, int NY,
                           long long nplanes, hipStream_t st) {
  const long long nrows = nplanes * NY;
  hipLaunchKernelGGL(k_presence, dim3((unsigned)((nrows + kWaves - 1) / kWaves)), dim3(kThreads), 0, st, labels,
                     rowmask, NX, nrows);
  hipLaunchKernelGGL(k_planemask, dim3((unsigned)((nplanes + kThreads - 1) / kThreads)), dim3(kThreads), 0, st,
                     rowmask, planemask, NY, nplanes);
  return hipGetLastError();
}

template <class Fn>
static hipError_t dispatch_n(int n, Fn&& fn) {
  switch (n) {
#define CASE(S) case S: return fn(std::integral_constant<int, S>{});
    FWIX_SIZES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
}

static dim3 row_grid(int N, int nrows) {
  const int nb = row_nb(N);
  const int waves = (nrows + nb - 1) / nb;
  return dim3((waves + kWaves - 1) / kWaves);
}
static dim3 col_grid(int N, int NX, int np) {
  const int cw = col_cw(N);
  return dim3((NX + cw - 1) / cw, np);
}

// kind: 0 plain forward, 1 plain inverse, 2 forward with pre-taper,
//       3 inverse with post-taper
hipError_t launch_rows(int kind, const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NX, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    const dim3 g = row_grid(N, a.np * a.NY);
    switch (kind) {
      case 0: hipLaunchKernelGGL((k_rows<N, false, false, false>), g, dim3(kThreads), 0, st, a); break;
      case 1: hipLaunchKernelGGL((k_rows<N, true, false, false>), g, dim3(kThreads), 0, st, a); break;
      case 2: hipLaunchKernelGGL((k_rows<N, false, true, false>), g, dim3(kThreads), 0, st, a); break;
      case 3: hipLaunchKernelGGL((k_rows<N, true, false, true>), g, dim3(kThreads), 0, st, a); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  });
}

hipError_t launch_rows_fanout(const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NX, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    hipLaunchKernelGGL((k_rows_fanout<N>), row_grid(N, a.np * a.NY), dim3(kThreads), 0, st, a);
    return hipGetLastError();
  });
}

hipError_t launch_rows_select(const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NX, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    hipLaunchKernelGGL((k_rows_select<N>), row_grid(N, a.np * a.NY), dim3(kThreads), 0, st, a);
    return hipGetLastError();
  });
}

hipError_t launch_cols(bool inv, const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NY, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    const dim3 g = col_grid(N, a.NX, a.np);
    if (inv) hipLaunchKernelGGL((k_cols<N, true>), g, dim3(kColThreads), 0, st, a);
    else hipLaunchKernelGGL((k_cols<N, false>), g, dim3(kColThreads), 0, st, a);
    return hipGetLastError();
  });
}

hipError_t launch_cols_fanout(int mode, const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NY, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    const dim3 g = col_grid(N, a.NX, a.np);
    if (mode == 0) hipLaunchKernelGGL((k_cols_fanout<N, 0>), g, dim3(kColThreads), 0, st, a);
    else if (mode == 1) hipLaunchKernelGGL((k_cols_fanout<N, 1>), g, dim3(kColThreads), 0, st, a);
    else hipLaunchKernelGGL((k_cols_fanout<N, 2>), g, dim3(kColThreads), 0, st, a);
    return hipGetLastError();
  });
}

hipError_t launch_cols_fanin(int mode, const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NY, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    const dim3 g = col_grid(N, a.NX, a.np);
    if (mode == 0) hipLaunchKernelGGL((k_cols_fanin<N, 0>), g, dim3(kColThreads), 0, st, a);
    else if (mode == 1) hipLaunchKernelGGL((k_cols_fanin<N, 1>), g, dim3(kColThreads), 0, st, a);
    else hipLaunchKernelGGL((k_cols_fanin<N, 2>), g, dim3(kColThreads), 0, st, a);
    return hipGetLastError();
  });
}

}  // namespace fwix
