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
  static constexpr int THREADS = col_threads(N);
  static constexpr int WAVES = THREADS / 64;
  static constexpr int CW = col_cw(N);
  static constexpr int NB = CW / WAVES;
  static constexpr int V = NB * N / 64;
  static constexpr int CWP = CW + 1;
  static constexpr int NROW = N;
  static constexpr int TILE = N * CWP;
  static constexpr int PRIV = NB * padded(N);
  static constexpr int SMEM = TILE > WAVES * PRIV ? TILE : WAVES * PRIV;
  static constexpr bool SPLIT = false;  // tile rows in natural order
  static_assert(NB >= 1 && V >= 4, "column tile too narrow");
};
// 1024-point columns on two waves each (k_cols_fanout_split / _fanin_split):
// 512-thread workgroups, 4 columns, wave 2c + h holding rows 2i + h of
// column c as a 512-point transform (V = 8 values per lane).  A loaded tile
// keeps the even rows first and the odd rows after them (SPLIT), so each wave
// reads its 512 rows with the stride of a 512-point column; results leave
// through tile_store_r2.
struct SplitCfg {
  static constexpr int THREADS = 512;
  static constexpr int WAVES = 8;
  static constexpr int CW = 4;
  static constexpr int NB = 1;
  static constexpr int V = 8;
  static constexpr int CWP = CW + 1;
  static constexpr int NROW = 1024;
  static constexpr int TILE = NROW * CWP;
  static constexpr int PRIV = padded(512);
  static constexpr int SMEM = TILE > WAVES * PRIV ? TILE : WAVES * PRIV;
  static constexpr bool SPLIT = true;
};
// LDS row slot of tile row r in a loaded tile
template <class C>
__device__ __forceinline__ int tslot(int r) {
  if constexpr (C::SPLIT) return (r & 1) * (C::NROW / 2) + (r >> 1);
  else return r;
}

// Workgroups are handed to the 8 XCDs round-robin by linear id, and each
// XCD has its own L2.  A column tile reads CW * 8 bytes of every row; with
// CW = 8 (4) that is a half (quarter) of a 128-byte line, so G = 2 (4)
// neighbouring tiles share every line they read.  Linear ids L, L + 8, ...
// run on the same XCD back to back, so they take the G parts: within each
// group of 8G ids, 8G g + 8h + q gets the logical tile 8G g + G q + h (a
// trailing partial group keeps the identity).  Logical tile J is column
// tile J % gridDim.x of plane J / gridDim.x.
template <class C>
__device__ __forceinline__ void col_tile(int& plane, int& c0) {
  const int nb = gridDim.x;
  const int L = blockIdx.y * nb + blockIdx.x;
  int J = L;
  constexpr int G = 128 / (C::CW * 8);
  if constexpr (G > 1) {
    const int total = nb * gridDim.y;
    const int r = L % (8 * G);
    if (L - r + 8 * G <= total) J = L - r + G * (r % 8) + r / 8;
  }
  plane = J / nb;
  c0 = (J - plane * nb) * C::CW;
}

// ------------------------------------------------------------------ row passes
// A row workgroup covers rows R0 .. R0 + 4 NB - 1 (R0 = 4 NB blockIdx.x), wave
// w rows R0 + w NB .. + NB - 1, and its waves' LDS regions hold those rows
// in order: row R0 + rel at rel * padded(N) (+ the PadLds padding).  In the
// quad-block layout the same rows are one contiguous run of 4 NB N elements
// from R0 N, so the workgroup moves them between LDS and memory with
// consecutive lanes on consecutive elements (whole 128-byte lines).
static_assert(kWaves == 4, "quad-block row staging assumes 4 waves of rows");
template <int N>
__device__ __forceinline__ int rows_stage_addr(int j) {  // LDS slot of element j of the run
  const int qr = j / (4 * N), rem = j - qr * 4 * N;
  const int w = rem & 15, rel = 4 * qr + (w >> 2), col = ((rem >> 4) << 2) + (w & 3);
  return rel * padded(N) + col + (col >> 5);
}
template <int N>
__device__ __forceinline__ int rows_stage_row(int j) {  // row of element j relative to R0
  const int qr = j / (4 * N), rem = j - qr * 4 * N;
  return 4 * qr + ((rem & 15) >> 2);
}
template <int N>
__device__ __forceinline__ int rows_stage_col(int j) {
  const int rem = j % (4 * N);
  return ((rem >> 4) << 2) + (rem & 3);
}
// run (quad-block memory) -> LDS rows; rows at or beyond nrows read as zero
template <int N, int NB>
__device__ __forceinline__ void rows_quad_load(float2* lds, const float2* __restrict__ src, int R0, int nrows) {
  constexpr int TOT = 4 * NB * N, PER = TOT / kThreads;
  static_assert(PER * kThreads == TOT, "run must divide evenly");
  float2 t[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int j = threadIdx.x + k * kThreads;
    t[k] = R0 + rows_stage_row<N>(j) < nrows ? src[(size_t)R0 * N + j] : czero();
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) lds[rows_stage_addr<N>(threadIdx.x + k * kThreads)] = t[k];
}
// LDS rows -> run (quad-block memory): dst[j] = f(value, row, col) for rows below nrows
template <int N, int NB, class Fn>
__device__ __forceinline__ void rows_quad_store(const float2* lds, float2* __restrict__ dst, int R0, int nrows, Fn&& f) {
  constexpr int TOT = 4 * NB * N, PER = TOT / kThreads;
  static_assert(PER * kThreads == TOT, "run must divide evenly");
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int j = threadIdx.x + k * kThreads;
    const int row = R0 + rows_stage_row<N>(j);
    if (row < nrows) dst[(size_t)R0 * N + j] = f(lds[rows_stage_addr<N>(j)], row, rows_stage_col<N>(j), (size_t)R0 * N + j);
  }
}

// out = (add ? out : 0) + post(scale * F_x(pre * in))   along x, one row per wave
// (in / out are K when qin / qout: in the quad-block layout, staged through LDS)
template <int N, bool INV, bool PRE_TAPER, bool POST_TAPER>
__global__ __launch_bounds__(kThreads) void k_rows(KArgs a) {
  using C = RowCfg<N>;
  using F = WaveFft<N, C::NB, false, INV, PadLds<N>>;
  __shared__ float2 lds[kWaves * C::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrows = a.np * a.NY;
  const int R0 = blockIdx.x * kWaves * C::NB;
  const int rb = R0 + wave * C::NB;
  // K in the quad-block layout goes through LDS as whole lines; row-major K
  // is read and written like any wavefield
  const bool qin = a.qin && a.quad, qout = a.qout && a.quad;
  const bool staged = qin || qout;  // grid-uniform: every wave reaches the barriers
  if (rb >= nrows && !staged) return;
  float2* wl = lds + wave * C::LDSW;
  const float s = a.scale;  // row normalisation 1/sqrt(NX)
  float2 v[C::V];
  if (qin) {
    rows_quad_load<N, C::NB>(lds, a.in, R0, nrows);
    __syncthreads();
  }
  if (rb < nrows) {
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
      const int lr = rb + b;
      float2 x = czero();
      if (lr < nrows) {
        x = qin ? wl[PadLds<N>::addr(b, i)] : a.in[(unsigned)(lr * N + i)];
        if (PRE_TAPER) x = cscale(x, a.tapx[i] * a.tapy[lr % a.NY]);
      }
      v[e] = x;
    }
    if (qin) wave_sync();  // the transform reuses the region it was read from
    F::run(v, wl, a.twx, lane);
  }
  auto post = [&](float2 y, int lr, int i, unsigned idx) {
    y = cscale(y, s);
    if (POST_TAPER) y = cscale(y, a.tapx[i] * a.tapy[lr % a.NY]);
    if (a.add) y = cadd(y, a.out[idx]);
    return y;
  };
  if (qout) {
    if (rb < nrows) {
      wave_sync();
#pragma unroll
      for (int e = 0; e < C::V; ++e) {
        int b, i;
        edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
        wl[PadLds<N>::addr(b, i)] = v[e];
      }
    }
    __syncthreads();
    rows_quad_store<N, C::NB>(lds, a.out, R0, nrows, post);
    return;
  }
  if (rb >= nrows) return;
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
    const int lr = rb + b;
    if (lr < nrows) {
      const unsigned idx = lr * N + i;
      a.out[idx] = post(v[e], lr, i, idx);
    }
  }
}

// rows_fanout: W_r = scale * F_x([labels == r] * in) for every reference r
// (T: arithmetic type of the transforms, float2 or double2)
template <int N, bool MAP, class T = float2>
__global__ __launch_bounds__(kThreads) void k_rows_fanout(KArgs a) {
  using C = RowCfg<N>;
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
  T u[C::V];
  uint8_t lab[C::V];
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, F::P::first(), false>(lane, e, b, i);
    const int lr = rb + b;
    u[e] = cz<T>();
    lab[e] = 255;
    if (lr < nrows) {
      u[e] = from_f2<T>(a.in[(unsigned)(lr * N + i)]);
      const int p = lr / a.NY, iy = lr - p * a.NY;
      lab[e] = a.labels[(unsigned)((model_iw<MAP>(a, plane_iw(a, p)) * a.NY + iy) * N + i)];
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
        wm |= a.rowmask[model_iw<MAP>(a, plane_iw(a, p)) * a.NY + (lr - p * a.NY)];
      }
    }
    wm = __builtin_amdgcn_readfirstlane(wm);
  }
  for (int r = 0; r < a.nref; ++r) {
    if (r < 32 && !((wm >> r) & 1u)) continue;  // rows without reference r: the column pass reads zeros
    T x[C::V];
#pragma unroll
    for (int e = 0; e < C::V; ++e) x[e] = (lab[e] == r) ? u[e] : cz<T>();
    F::run(x, wl, twl, lane);
    float2* W = a.W + (size_t)r * a.wstride;
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::last(), false>(lane, e, b, i);
      const int lr = rb + b;
      if (lr < nrows) W[widx(a, lr, i, N)] = to_f2(cscale(x[e], s));
    }
  }
}

// rows_select: out = (add ? out : 0) + sum_r [labels == r] * scale * F_x^-1(W_r)
// KNEXT: also K = scale * F_x(taper * out) (quad-block), the first row pass of
// a next step that propagates out further: one transform more here instead of
// a pass that reads out back.
template <int N, bool MAP, bool KNEXT = false>
__global__ __launch_bounds__(kThreads) void k_rows_select(KArgs a) {
  using C = RowCfg<N>;
  using F = WaveFft<N, C::NB, false, true, PadLds<N>>;
  using Ff = WaveFft<N, C::NB, false, false, PadLds<N>>;
  static_assert(F::P::last() == Ff::P::first(), "register hand-off needs mirrored plans");
  __shared__ float2 lds[kWaves * C::LDSW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nrows = a.np * a.NY;
  const int rb = (blockIdx.x * kWaves + wave) * C::NB;
  // twiddles in LDS: inside the term loops a global load would wait (one
  // counter for all vector memory operations) for the term stores / the next
  // term's loads still in flight
  __shared__ float2 twl[N];
  for (int i = threadIdx.x; i < N; i += kThreads) twl[i] = a.twx[i];
  __syncthreads();
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
      lab[e] = a.labels[(unsigned)((model_iw<MAP>(a, plane_iw(a, p)) * a.NY + iy) * N + i)];
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
        wm |= a.rowmask[model_iw<MAP>(a, plane_iw(a, p)) * a.NY + (lr - p * a.NY)];
      }
    }
    wm = __builtin_amdgcn_readfirstlane(wm);
  }
  const uint32_t live = a.nref >= 32 ? wm : (wm & ((1u << a.nref) - 1u));
  float2 nxt[C::V];
  auto fetch = [&](int r) {
    const float2* W = a.W + (size_t)r * a.wstride;
    // the lane index made opaque here: the per-element offsets are recomputed
    // per reference (a few integer ops) instead of held in registers
    int ln = lane;
    asm volatile("" : "+v"(ln));
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, F::P::first(), false>(ln, e, b, i);
      const int lr = rb + b;
      nxt[e] = lr < nrows ? W[widx(a, lr, i, N)] : czero();
    }
  };
  // absent references contribute nothing to these rows: skip their loads and
  // transforms.  The next reference's row is fetched into registers during
  // this one's transform, except at 16 values per lane, where holding it
  // would cost a wave of occupancy
  constexpr bool PREFETCH = C::V < 16;
  int r = a.rowmask ? next_bit(live, -1) : 0;
  const int rend = a.rowmask ? 32 : a.nref;
  if (PREFETCH && r < rend) fetch(r);
  while (r < rend) {
    if constexpr (!PREFETCH) fetch(r);
    float2 x[C::V];
#pragma unroll
    for (int e = 0; e < C::V; ++e) x[e] = nxt[e];
    const int rn = a.rowmask ? next_bit(live, r) : r + 1;
    if (PREFETCH && rn < rend) fetch(rn);
    F::run(x, wl, twl, lane);
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
      const unsigned idx = lr * N + i;
      float2 y = cscale(acc[e], s);
      if (a.add) y = cadd(y, a.out[idx]);
      if (!KNEXT || a.knext != 2) a.out[idx] = y;
      if constexpr (KNEXT) acc[e] = cscale(y, a.tapx[i] * a.tapy[lr % a.NY]);
    }
  }
  if constexpr (KNEXT) {
    Ff::run(acc, wl, twl, lane);
#pragma unroll
    for (int e = 0; e < C::V; ++e) {
      int b, i;
      edge_pos<N, C::NB, Ff::P::last(), false>(lane, e, b, i);
      const int lr = rb + b;
      if (lr < nrows) a.K[widx(a, lr, i, N)] = cscale(acc[e], s);
    }
  }
}

// ------------------------------------------------------------------ column passes
// Pointwise spectral multiplier of term r at (kx, ky) of frequency iw.
// MULT 0 / 1: forward / adjoint phase shift of reference slowness r
// (propagator/src/phase_shift.cu); MULT 2: Born scattering wavenumber factor
// (|k|/w)^(2it) = (k^2/w^2)^it of Taylor order it = r + 1 (scatter.cu:23-59);
// order 0 is the identity and is applied in space by the row kernels.
// The term's uniform part (reference slowness, real/complex branch) is set up
// once per (plane, term) so the unrolled per-sample loops carry no divergent
// or if-converted paths.
template <int MULT>
struct TermMult {
  float w2, wsre, b, iw2, dz;
  double wsre64, b64;  // the same in float64 (the float64 passes)
  int r;
  bool real;
  __device__ __forceinline__ TermMult(const KArgs& a, int r_, int iw, float w2_)
      : TermMult(a, r_, w2_, MULT == 2 ? make_float2(0.f, 0.f) : a.sref[r_ * a.NW + iw]) {}
  // sr: the reference slowness of term r (already read, e.g. from LDS)
  __device__ __forceinline__ TermMult(const KArgs& a, int r_, float w2_, float2 sr) : w2(w2_), r(r_), dz(a.dz) {
    if constexpr (MULT == 2) {
      iw2 = 1.f / w2;
      real = true; wsre = 0.f; b = 0.f;
      wsre64 = 0.0; b64 = 0.0;
    } else {
      wsre = w2 * sr.x;
      b = w2 * (sr.y - a.eps * sr.x);
      real = b == 0.f;
      iw2 = 0.f;
      wsre64 = (double)w2 * sr.x;
      b64 = (double)w2 * ((double)sr.y - (double)a.eps * sr.x);
    }
  }
  template <bool REAL>
  __device__ __forceinline__ double2 at(double k2) const {
    if constexpr (MULT == 2) {
      const double q = k2 / (double)w2;
      double m = q;
      for (int i = 0; i < r; ++i) m *= q;
      return make_double2(m, 0.0);
    } else if constexpr (REAL) {
      return ps_real64<MULT>(wsre64 - k2, (double)dz);
    } else {
      return ps_general64<MULT>(wsre64 - k2, b64, (double)dz);
    }
  }
  template <bool REAL>
  __device__ __forceinline__ float2 at(float k2) const {
    if constexpr (MULT == 2) {  // term r carries Taylor order r + 1 (order 0 needs no transform)
      const float q = k2 * iw2;
      float m = q;
      for (int i = 0; i < r; ++i) m *= q;
      return make_float2(m, 0.f);
    } else if constexpr (REAL) {
      return ps_real<MULT>(wsre - k2, dz);
    } else {
      return ps_general<MULT>(wsre - k2, b, dz);
    }
  }
};

// k^2 = kx^2 + ky^2 at this wave's values after a column transform
// (positions of the last stage of radix R); loop-invariant over the terms.
template <int N, class C, int R, class K2>
__device__ __forceinline__ void column_k2(K2* k2v, const KArgs& a, int colw, int lane) {
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, R, false>(lane, e, b, i);
    const int col = colw + b;
    const float kxv = a.kx[col < a.NX ? col : 0], kyv = a.ky[i];
    if constexpr (sizeof(K2) == 8) k2v[e] = fma((double)kxv, (double)kxv, (double)kyv * kyv);
    else k2v[e] = fmaf(kxv, kxv, kyv * kyv);
  }
}
// multiplier of term tm at k^2 in the arithmetic type T
template <class T, bool REAL, class TM, class K2>
__device__ __forceinline__ T term_at(const TM& tm, K2 k2) {
  if constexpr (sizeof(T) == 16) return tm.template at<REAL>((double)k2);
  else return tm.template at<REAL>((float)k2);
}

// Coalesced tile load / store: every thread first issues all of its global
// loads into registers (PER = N*CW/THREADS of them), then writes LDS, so the
// memory latency is paid once per tile rather than once per element.
// One plane of a wavefield (or of a term buffer) as a buffer resource: the
// base address sits in scalar registers and each element costs one 32-bit
// byte offset.  Offsets at or beyond the plane's size read as zero and drop
// stores (hardware range check), which replaces the per-element bounds and
// presence branches: such elements get the offset kOff.
constexpr unsigned kOff = 0x80000000u;
struct PlaneBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ PlaneBuf(const float2* base, unsigned bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000)) {}
  __device__ __forceinline__ float2 load(unsigned off) const {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
  }
  __device__ __forceinline__ void store(float2 v, unsigned off) const {
    using U2 = __attribute__((ext_vector_type(2))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, (int)off, 0, 0);
  }
};
// Byte offsets of this thread's tile elements: element k sits in row
// threadIdx.x / CW + k * RS (RS = THREADS / CW), column c0 + threadIdx.x % CW,
// of a row-major plane or (quad) of one in the quad-block layout.  When RS is
// a multiple of 4, a step of RS rows moves the offset by RS rows of NX
// elements in both layouts; the narrow tiles of short columns (RS = 1, 2)
// add the quad-block row part per element.  Columns at or beyond NX get
// offsets at or beyond the plane (kOff + at most a plane).
template <class C>
struct TileOffs {
  static constexpr int PER = C::NROW * C::CW / C::THREADS;
  static constexpr int RS = C::THREADS / C::CW;
  static_assert(PER * C::THREADS == C::NROW * C::CW && C::THREADS % C::CW == 0, "tile must divide evenly");
  unsigned o0, step;
  int row0, nx4;
  bool quad;
  __device__ __forceinline__ TileOffs(int NX, int c0, bool quad_) : quad(quad_) {
    row0 = (int)threadIdx.x / C::CW;
    nx4 = NX >> 2;
    const int col = c0 + (int)threadIdx.x % C::CW;
    step = (unsigned)(RS * NX) * 8u;
    if (col >= NX) o0 = kOff;
    else if (!quad) o0 = (unsigned)(row0 * NX + col) * 8u;
    else if constexpr (RS % 4 == 0) o0 = qidx(row0, col, NX) * 8u;
    else o0 = (unsigned)(((col >> 2) << 4) + (col & 3)) * 8u;  // column part; rows added per element
  }
  __device__ __forceinline__ unsigned at(int k) const {
    if constexpr (RS % 4 != 0) {
      if (quad) {
        const int row = row0 + k * RS;
        return o0 + (unsigned)((((row >> 2) * nx4) << 4) + ((row & 3) << 2)) * 8u;
      }
    }
    return o0 + (unsigned)k * step;
  }
};
__device__ __forceinline__ unsigned plane_bytes(const KArgs& a) { return (unsigned)(a.NX * a.NY) * 8u; }

// presence masks of the tile rows, staged in LDS (all set when masks are off)
template <class C>
__device__ __forceinline__ void tile_row_masks(uint32_t* rml, const uint32_t* __restrict__ rm) {
  for (int i = threadIdx.x; i < C::NROW; i += C::THREADS) rml[i] = rm ? rm[i] : 0xffffffffu;
}
template <class C>
__device__ __forceinline__ uint32_t row_mask_of(const uint32_t* rml, int k) {
  return rml[(int)threadIdx.x / C::CW + k * (C::THREADS / C::CW)];
}
template <class C>
struct TileRegs {
  static constexpr int PER = TileOffs<C>::PER;
  float2 t[PER];
  __device__ __forceinline__ void fetch(const PlaneBuf& b, const TileOffs<C>& o) {
#pragma unroll
    for (int k = 0; k < PER; ++k) t[k] = b.load(o.at(k));
  }
  // rows whose mask bit is clear are known to be zero and are not loaded;
  // rmk holds the presence mask of every tile row
  __device__ __forceinline__ void fetch_masked(const PlaneBuf& b, const TileOffs<C>& o, const uint32_t* rmk,
                                               uint32_t bit) {
#pragma unroll
    for (int k = 0; k < PER; ++k) t[k] = b.load((row_mask_of<C>(rmk, k) & bit) ? o.at(k) : kOff);
  }
  __device__ __forceinline__ void commit(float2* tile) const {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + k * C::THREADS;
      tile[tslot<C>(e / C::CW) * C::CWP + e % C::CW] = t[k];
    }
  }
};
template <class C>
__device__ __forceinline__ void tile_load(float2* tile, const PlaneBuf& b, const TileOffs<C>& o) {
  TileRegs<C> r;
  r.fetch(b, o);
  r.commit(tile);
}
// store only the rows whose mask bit is set (rows that a later pass reads)
template <class C>
__device__ __forceinline__ void tile_store_masked(const float2* tile, const PlaneBuf& b, const TileOffs<C>& o,
                                                  const uint32_t* rmk, uint32_t bit) {
  static_assert(!C::SPLIT, "split tiles are stored by tile_store_r2");
#pragma unroll
  for (int k = 0; k < TileOffs<C>::PER; ++k) {
    const int e = threadIdx.x + k * C::THREADS;
    b.store(tile[(e / C::CW) * C::CWP + e % C::CW], (row_mask_of<C>(rmk, k) & bit) ? o.at(k) : kOff);
  }
}
template <class C>
__device__ __forceinline__ void tile_store(const float2* tile, const PlaneBuf& b, const TileOffs<C>& o, int add) {
  constexpr int PER = TileOffs<C>::PER;
  static_assert(!C::SPLIT, "split tiles are stored by tile_store_r2");
  float2 t[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = threadIdx.x + k * C::THREADS;
    t[k] = tile[(e / C::CW) * C::CWP + e % C::CW];
  }
  if (add) {
#pragma unroll
    for (int k = 0; k < PER; ++k) t[k] = cadd(t[k], b.load(o.at(k)));
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) b.store(t[k], o.at(k));
}
// this wave's columns at the first-stage positions of radix R, from the tile
template <int N, class C, int R, class T>
__device__ __forceinline__ void tile_read(T* v, const float2* tile, int wave, int lane) {
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, R, false>(lane, e, b, i);
    v[e] = from_f2<T>(tile[i * C::CWP + wave * C::NB + b]);
  }
}
template <int N, class C, int R, class T>
__device__ __forceinline__ void tile_write(float2* tile, const T* v, int wave, int lane) {
#pragma unroll
  for (int e = 0; e < C::V; ++e) {
    int b, i;
    edge_pos<N, C::NB, R, false>(lane, e, b, i);
    tile[i * C::CWP + wave * C::NB + b] = to_f2(v[e]);
  }
}

// LDS of a column workgroup in float2 units: the complex64 tile, reused
// between barriers as the waves' private exchange regions of type T
template <int N, class T>
constexpr int col_smem() {
  using C = ColCfg<N>;
  constexpr int priv = C::WAVES * C::PRIV * (int)(sizeof(T) / sizeof(float2));
  return C::TILE > priv ? C::TILE : priv;
}

// out = (add ? out : 0) + scale * F_y(in) on each plane (plain 2-D FFT column pass)
template <int N, bool INV>
__global__ __launch_bounds__(ColCfg<N>::THREADS) __attribute__((amdgpu_waves_per_eu(col_wpe_default(0, 0, N)))) void k_cols(KArgs a) {
  using C = ColCfg<N>;
  using F = WaveFft<N, C::NB, false, INV, PadLds<N>>;
  __shared__ float2 smem[C::SMEM];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int p, c0;
  col_tile<C>(p, c0);
  const size_t pb = (size_t)p * a.NX * N;
  float2* priv = smem + wave * C::PRIV;
  const TileOffs<C> to(a.NX, c0, false);
  tile_load<C>(smem, PlaneBuf(a.in + pb, plane_bytes(a)), to);
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
  tile_store<C>(smem, PlaneBuf(a.out + pb, plane_bytes(a)), to, a.add);
}

// Whether the twiddle table (N values) and the reference slownesses (up to
// 256) fit in LDS beside the tile and the row masks without lowering the
// workgroups a CU holds (at most WPE * 4 / waves per workgroup by
// registers; 160 KB of LDS).
template <int N, int WPE, class T = float2>
constexpr bool col_stage_lds() {
  using C = ColCfg<N>;
  constexpr int base = col_smem<N, T>() * 8 + C::NROW * 4;
  constexpr int extra = N * (int)sizeof(T) + 256 * 8;
  constexpr int by_regs = WPE * 4 / C::WAVES;
  constexpr int without = (160 * 1024) / base < by_regs ? (160 * 1024) / base : by_regs;
  constexpr int with = (160 * 1024) / (base + extra) < by_regs ? (160 * 1024) / (base + extra) : by_regs;
  return with >= without;
}

// cols_fanout: W_r = F_y^-1[M_r * F_y(K)] / NY for every term r (M_r: term_mult<MODE>).
// (T: arithmetic type of the transforms, float2 or double2)
template <int N, int MODE, int WPE, class T = float2>
__global__ __launch_bounds__(ColCfg<N>::THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_cols_fanout(KArgs a) {
  using C = ColCfg<N>;
  using F = WaveFft<N, C::NB, false, false, PadLds<N>, T>;
  using Fi = WaveFft<N, C::NB, false, true, PadLds<N>, T>;
  static_assert(F::P::last() == Fi::P::first(), "register hand-off needs mirrored plans");
  __shared__ __align__(16) float2 smem[col_smem<N, T>()];
  // The loop below stores each term's tile and then starts the next term; on
  // this ISA one wait counter covers loads and stores, so any global load
  // inside the loop (twiddles, reference slownesses) would also wait for the
  // previous term's stores to reach memory.  Both are staged in LDS first
  // where that costs no occupancy (STAGE).
  constexpr bool STAGE = col_stage_lds<N, WPE, T>();
  __shared__ T twl[STAGE ? N : 1];
  __shared__ float2 srl[STAGE ? 256 : 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int p, c0;
  col_tile<C>(p, c0);
  const size_t pb = (size_t)p * a.NX * N;
  T* priv = (T*)smem + wave * C::PRIV;
  const TileOffs<C> to(a.NX, c0, a.quad != 0);
  tile_load<C>(smem, PlaneBuf(a.K + pb, plane_bytes(a)), to);
  const int iw = plane_iw(a, p);
  const T* tw = twy_of(a, (T*)nullptr);
  if constexpr (STAGE) {
    for (int i = threadIdx.x; i < N; i += C::THREADS) twl[i] = tw[i];
    if constexpr (MODE != 2) {
      const int nt = a.nterm < 256 ? a.nterm : 256;
      for (int i = threadIdx.x; i < nt; i += C::THREADS) srl[i] = a.sref[i * a.NW + iw];
    }
    tw = twl;
  }
  __syncthreads();
  T spec[C::V];
  tile_read<N, C, F::P::first()>(spec, smem, wave, lane);
  __syncthreads();
  F::run(spec, priv, tw, lane);
  const float w2 = a.w2[iw];
  const float s2 = a.scale;  // 1/NY: forward and inverse column normalisation together
  const int colw = c0 + wave * C::NB;
  const int im = model_iw_any(a, iw);
  const uint32_t* rm = (MODE != 2 && a.rowmask) ? a.rowmask + im * a.NY : nullptr;
  const uint32_t pm = (MODE != 2 && a.planemask) ? a.planemask[im] : 0xffffffffu;
  __shared__ uint32_t rmk[C::NROW];
  tile_row_masks<C>(rmk, rm);
  real_of<T> k2v[C::V];
  column_k2<N, C, F::P::last()>(k2v, a, colw, lane);
  for (int r = 0; r < a.nterm; ++r) {
    if (r < 32 && !((pm >> r) & 1u)) continue;  // reference absent from the plane: W_r is never read
    float2 srv = make_float2(0.f, 0.f);  // nref <= 255
    if constexpr (MODE != 2) srv = STAGE ? srl[r] : a.sref[r * a.NW + iw];
    const TermMult<MODE> tm(a, r, w2, srv);
    T x[C::V];
    auto mult = [&](auto real) {
#pragma unroll
      for (int e = 0; e < C::V; ++e)
        x[e] = cscale(cmul(spec[e], term_at<T, decltype(real)::value>(tm, k2v[e])), s2);
    };
    if (tm.real) mult(std::true_type{});
    else mult(std::false_type{});
    Fi::run(x, priv, tw, lane);
    __syncthreads();
    tile_write<N, C, Fi::P::last()>(smem, x, wave, lane);
    __syncthreads();
    const PlaneBuf wb(a.W + (size_t)r * a.wstride + pb, plane_bytes(a));
    if (rm) tile_store_masked<C>(smem, wb, to, rmk, r < 32 ? 1u << r : 0xffffffffu);
    else tile_store<C>(smem, wb, to, 0);
    __syncthreads();
  }
}

// cols_fanin: K = F_y^-1[sum_r M_r * F_y(W_r)] / NY (M_r: term_mult<MODE>).
// (T: arithmetic type of the transforms, float2 or double2)
template <int N, int MODE, int WPE, class T = float2>
__global__ __launch_bounds__(ColCfg<N>::THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_cols_fanin(KArgs a) {
  using C = ColCfg<N>;
  using F = WaveFft<N, C::NB, false, false, PadLds<N>, T>;
  using Fi = WaveFft<N, C::NB, false, true, PadLds<N>, T>;
  static_assert(F::P::last() == Fi::P::first(), "register hand-off needs mirrored plans");
  __shared__ __align__(16) float2 smem[col_smem<N, T>()];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int p, c0;
  col_tile<C>(p, c0);
  const size_t pb = (size_t)p * a.NX * N;
  T* priv = (T*)smem + wave * C::PRIV;
  const TileOffs<C> to(a.NX, c0, a.quad != 0);
  const int iw = plane_iw(a, p);
  const float w2 = a.w2[iw];
  const int colw = c0 + wave * C::NB;
  T acc[C::V];
#pragma unroll
  for (int e = 0; e < C::V; ++e) acc[e] = cz<T>();
  const int im = model_iw_any(a, iw);
  const uint32_t* rm = (MODE != 2 && a.rowmask) ? a.rowmask + im * a.NY : nullptr;
  const uint32_t pm = (MODE != 2 && a.planemask) ? a.planemask[im] : 0xffffffffu;
  const uint32_t live = a.nterm >= 32 ? pm : (pm & ((1u << a.nterm) - 1u));
  const bool masked = rm != nullptr;
  const int rend = masked ? 32 : a.nterm;
  __shared__ uint32_t rmk[C::NROW];
  // twiddles and reference slownesses in LDS where that costs no occupancy
  // (STAGE): a global load in the loop would wait (one counter for all vector
  // memory operations) for the next term's tile, which is fetched into
  // registers while this term is transformed
  constexpr bool STAGE = col_stage_lds<N, WPE, T>();
  __shared__ T twl[STAGE ? N : 1];
  __shared__ float2 srl[STAGE ? 256 : 1];
  tile_row_masks<C>(rmk, rm);
  const T* tw = twy_of(a, (T*)nullptr);
  if constexpr (STAGE) {
    for (int i = threadIdx.x; i < N; i += C::THREADS) twl[i] = tw[i];
    if constexpr (MODE != 2) {
      const int nt = a.nterm < 256 ? a.nterm : 256;
      for (int i = threadIdx.x; i < nt; i += C::THREADS) srl[i] = a.sref[i * a.NW + iw];
    }
    tw = twl;
  }
  __syncthreads();
  real_of<T> k2v[C::V];
  column_k2<N, C, F::P::last()>(k2v, a, colw, lane);
  auto fetch = [&](TileRegs<C>& t, int rr) {
    const PlaneBuf wb(a.W + (size_t)rr * a.wstride + pb, plane_bytes(a));
    if (masked) t.fetch_masked(wb, to, rmk, 1u << rr);
    else t.fetch(wb, to);
  };
  // one term: transform its tile (already in LDS), multiply, accumulate;
  // after_read runs once the tile is in registers
  auto term = [&](int rc, auto&& after_read) {
    T v[C::V];
    tile_read<N, C, F::P::first()>(v, smem, wave, lane);
    after_read();
    __syncthreads();
    F::run(v, priv, tw, lane);
    float2 srv = make_float2(0.f, 0.f);  // nref <= 255
    if constexpr (MODE != 2) srv = STAGE ? srl[rc] : a.sref[rc * a.NW + iw];
    const TermMult<MODE> tm(a, rc, w2, srv);
    auto mult = [&](auto real) {
#pragma unroll
      for (int e = 0; e < C::V; ++e)
        acc[e] = cadd(acc[e], cmul(v[e], term_at<T, decltype(real)::value>(tm, k2v[e])));
    };
    if (tm.real) mult(std::true_type{});
    else mult(std::false_type{});
    __syncthreads();
  };
  int r = masked ? next_bit(live, -1) : 0;
  // the next term's tile is fetched into registers during this term's
  // transform where the register budget holds it without spilling
  if constexpr (N <= 512 || WPE <= 2) {
    TileRegs<C> tile;
    if (r < rend) fetch(tile, r);
    while (r < rend) {
      tile.commit(smem);
      __syncthreads();
      const int rn = masked ? next_bit(live, r) : r + 1;
      term(r, [&] { if (rn < rend) fetch(tile, rn); });
      r = rn;
    }
  } else {
    while (r < rend) {
      {
        TileRegs<C> cur;
        fetch(cur, r);
        cur.commit(smem);
      }
      __syncthreads();
      const int rn = masked ? next_bit(live, r) : r + 1;
      term(r, [] {});
      r = rn;
    }
  }
  const float s2 = a.scale;
#pragma unroll
  for (int e = 0; e < C::V; ++e) acc[e] = cscale(acc[e], s2);
  Fi::run(acc, priv, tw, lane);
  __syncthreads();
  tile_write<N, C, Fi::P::last()>(smem, acc, wave, lane);
  __syncthreads();
  tile_store<C>(smem, PlaneBuf(a.K + pb, plane_bytes(a)), to, 0);
}

// ------------------------------------------------------------------ 1024-point columns, two waves each
// Radix-2 across a wave pair.  Forward (decimation in time): wave h of column
// c transforms x[2i + h] (512 points: E for h = 0, O for h = 1); then X[n] =
// E[k] + W^k O[k] for n = k < 512 and E[k] - W^k O[k] for n = k + 512, W^k =
// exp(-2 pi i k / 1024).  Each wave keeps the spectrum samples of one parity,
// X[2j + h] at the positions j of a 512-point transform, so the multipliers
// are evaluated once per sample.  Inverse: wave h transforms its samples
// Y[2j + h] back (A for h = 0, B for h = 1, 512 points each) and the tile
// store combines x[m] = A[m] + W^-m B[m], x[m + 512] = A[m] - W^-m B[m].
struct SplitTw {
  float2 h[512];  // exp(-2 pi i j / 512): the 512-point transforms
  float2 c[512];  // exp(-2 pi i k / 1024), k < 512: the radix-2 combinations
};
__device__ __forceinline__ void split_tw_stage(SplitTw& tw, const float2* __restrict__ twy) {
  for (int i = threadIdx.x; i < 512; i += SplitCfg::THREADS) {
    tw.h[i] = twy[2 * i];
    tw.c[i] = twy[i];
  }
}
// Reference slowness r of the plane, by a scalar load (through a
// constant-address-space pointer): that load waits on the scalar counter
// only, never on the vector-memory counter that the tile stores and the
// prefetched tile hold, and every workgroup of the plane reads the same few
// values (scalar-cache hits).
__device__ __forceinline__ float2 split_sref(const KArgs& a, int r, int iw) {
  using CP = const __attribute__((address_space(4))) float*;
  const CP p = (CP)(const void*)a.sref + 2 * (r * a.NW + iw);
  return make_float2(p[0], p[1]);
}
// this wave's 512 samples (rows 2i + h of column cl) at the first-stage positions
template <int R>
__device__ __forceinline__ void split_read(float2* v, const float2* tile, int cl, int h, int lane) {
#pragma unroll
  for (int e = 0; e < SplitCfg::V; ++e) {
    int b, i;
    edge_pos<512, 1, R, false>(lane, e, b, i);
    v[e] = tile[(h * 512 + i) * SplitCfg::CWP + cl];
  }
}
// this wave's 512 values at positions R to tile slots 512 h + i of column cl
template <int R>
__device__ __forceinline__ void split_write(float2* tile, const float2* v, int cl, int h, int lane) {
#pragma unroll
  for (int e = 0; e < SplitCfg::V; ++e) {
    int b, i;
    edge_pos<512, 1, R, false>(lane, e, b, i);
    tile[(h * 512 + i) * SplitCfg::CWP + cl] = v[e];
  }
}
// spectrum index of value e of this wave at positions R
template <int R>
__device__ __forceinline__ int split_n(int lane, int e, int h) {
  int b, i;
  edge_pos<512, 1, R, false>(lane, e, b, i);
  return 2 * i + h;
}
// X[n] from E (tile slots k) and O (slots 512 + k), k = n mod 512
__device__ __forceinline__ float2 split_spec(const float2* tile, const SplitTw& tw, int cl, int n) {
  const int k = n & 511;
  const float2 E = tile[k * SplitCfg::CWP + cl];
  const float2 O = cmul(tile[(512 + k) * SplitCfg::CWP + cl], tw.c[k]);
  return n < 512 ? cadd(E, O) : csub(E, O);
}
// Store the tile's inverse transforms: A at slots m, B at slots 512 + m of
// each column.  Element k of this thread sits in row m = threadIdx.x / CW +
// k * (THREADS / CW): elements k and k + 4 are rows m and m + 512.  With rmk,
// only rows whose mask has `bit` set are stored.
__device__ __forceinline__ void tile_store_r2(const float2* tile, const SplitTw& tw, const PlaneBuf& b,
                                              const TileOffs<SplitCfg>& o, const uint32_t* rmk, uint32_t bit) {
  using C = SplitCfg;
  static_assert(TileOffs<C>::PER == 8 && 4 * (C::THREADS / C::CW) == 512, "rows k and k + 4 must be m and m + 512");
  const int c = (int)threadIdx.x % C::CW;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int m = (int)threadIdx.x / C::CW + k * (C::THREADS / C::CW);
    const float2 A = tile[m * C::CWP + c];
    const float2 B = cmulc(tile[(512 + m) * C::CWP + c], tw.c[m]);
    unsigned olo = o.at(k), ohi = o.at(k + 4);
    if (rmk) {
      if (!(row_mask_of<C>(rmk, k) & bit)) olo = kOff;
      if (!(row_mask_of<C>(rmk, k + 4) & bit)) ohi = kOff;
    }
    b.store(cadd(A, B), olo);
    b.store(csub(A, B), ohi);
  }
}
// k^2 at spectrum indices n of column colw
template <int R>
__device__ __forceinline__ void split_k2(float* k2v, const KArgs& a, int colw, int h, int lane) {
  const float kxv = a.kx[colw < a.NX ? colw : 0];
#pragma unroll
  for (int e = 0; e < SplitCfg::V; ++e) {
    const float kyv = a.ky[split_n<R>(lane, e, h)];
    k2v[e] = fmaf(kxv, kxv, kyv * kyv);
  }
}

// The phase-shift multipliers M_r(k^2) at this wave's 8 spectrum samples,
// evaluated at 4 of them.  Indices n and 1024 - n have the same k^2 (the ky
// table is odd-symmetric, bit for bit), and both lie in the same wave: wave h
// holds n = 2j + h at j = lane + 64 e (the positions of a last radix-8 stage),
// so value e >= 4 mirrors value 7 - e of lane 63 - lane (h = 1) or 64 - lane
// (h = 0).  Lane 0 of wave 0 holds n = 0 and n = 512, its own mirrors: its
// values 5..7 mirror its values 3..1, and value 4 (n = 512) is evaluated.
// msrc: byte address of the mirror lane for ds_bpermute.
template <bool REAL, class TM>
__device__ __forceinline__ void split_mults(float2* m, const TM& tm, const float* k2v, int msrc, bool lane0h0) {
  static_assert(SplitCfg::V == 8, "mirror map assumes 8 values per lane");
#pragma unroll
  for (int e = 0; e < 4; ++e) m[e] = tm.template at<REAL>(k2v[e]);
#pragma unroll
  for (int e = 4; e < 8; ++e) {
    m[e].x = __int_as_float(__builtin_amdgcn_ds_bpermute(msrc, __float_as_int(m[7 - e].x)));
    m[e].y = __int_as_float(__builtin_amdgcn_ds_bpermute(msrc, __float_as_int(m[7 - e].y)));
  }
  if (lane0h0) {
    m[4] = tm.template at<REAL>(k2v[4]);
    m[5] = m[3];
    m[6] = m[2];
    m[7] = m[1];
  }
}

// cols_fanout for 1024-point columns: W_r = F_y^-1[M_r * F_y(K)] / NY
template <int MODE, int WPE>
__global__ __launch_bounds__(SplitCfg::THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_cols_fanout_split(KArgs a) {
  using C = SplitCfg;
  using F = WaveFft<512, 1, false, false, PadLds<512>>;
  using Fi = WaveFft<512, 1, false, true, PadLds<512>>;
  static_assert(F::P::last() == Fi::P::first(), "register hand-off needs mirrored plans");
  constexpr int RS = F::P::last();
  __shared__ float2 smem[C::SMEM];
  __shared__ SplitTw tw;
  __shared__ uint32_t rmk[C::NROW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = wave >> 1, h = wave & 1;
  static_assert(RS == 8, "split_mults needs the positions of a last radix-8 stage");
  const int msrc = ((h ? 63 - lane : 64 - lane) & 63) * 4;
  const bool lane0h0 = h == 0 && lane == 0;
  int p, c0;
  col_tile<C>(p, c0);
  const size_t pb = (size_t)p * a.NX * C::NROW;
  float2* priv = smem + wave * C::PRIV;
  const TileOffs<C> to(a.NX, c0, a.quad != 0);
  {
    TileRegs<C> t;
    t.fetch(PlaneBuf(a.K + pb, plane_bytes(a)), to);
    t.commit(smem);
  }
  const int iw = plane_iw(a, p);
  split_tw_stage(tw, a.twy);
  const int im = model_iw_any(a, iw);
  const uint32_t* rm = (MODE != 2 && a.rowmask) ? a.rowmask + im * a.NY : nullptr;
  const uint32_t pm = (MODE != 2 && a.planemask) ? a.planemask[im] : 0xffffffffu;
  tile_row_masks<C>(rmk, rm);
  __syncthreads();
  float2 xs[C::V];
  split_read<F::P::first()>(xs, smem, cl, h, lane);
  __syncthreads();
  F::run(xs, priv, tw.h, lane);
  __syncthreads();
  split_write<RS>(smem, xs, cl, h, lane);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < C::V; ++e) xs[e] = split_spec(smem, tw, cl, split_n<RS>(lane, e, h));
  float k2v[C::V];
  split_k2<RS>(k2v, a, c0 + cl, h, lane);
  __syncthreads();  // the transforms' exchange regions overlap the tile
  const float w2 = a.w2[iw];
  const float s2 = a.scale;  // 1/NY: forward and inverse column normalisation together
  for (int r = 0; r < a.nterm; ++r) {
    if (r < 32 && !((pm >> r) & 1u)) continue;  // reference absent from the plane: W_r is never read
    float2 srv = make_float2(0.f, 0.f);
    if constexpr (MODE != 2) srv = split_sref(a, r, iw);
    const TermMult<MODE> tm(a, r, w2, srv);
    float2 y[C::V];
    auto mult = [&](auto real) {
      float2 m[C::V];
      if constexpr (MODE == 2) {
#pragma unroll
        for (int e = 0; e < C::V; ++e) m[e] = tm.template at<decltype(real)::value>(k2v[e]);
      } else {
        split_mults<decltype(real)::value>(m, tm, k2v, msrc, lane0h0);
      }
#pragma unroll
      for (int e = 0; e < C::V; ++e) y[e] = cscale(cmul(xs[e], m[e]), s2);
    };
    if (tm.real) mult(std::true_type{});
    else mult(std::false_type{});
    Fi::run(y, priv, tw.h, lane);
    __syncthreads();
    split_write<Fi::P::last()>(smem, y, cl, h, lane);
    __syncthreads();
    const PlaneBuf wb(a.W + (size_t)r * a.wstride + pb, plane_bytes(a));
    tile_store_r2(smem, tw, wb, to, rm ? rmk : nullptr, r < 32 ? 1u << r : 0xffffffffu);
    __syncthreads();
  }
}

// cols_fanin for 1024-point columns: K = F_y^-1[sum_r M_r * F_y(W_r)] / NY
template <int MODE, int WPE>
__global__ __launch_bounds__(SplitCfg::THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_cols_fanin_split(KArgs a) {
  using C = SplitCfg;
  using F = WaveFft<512, 1, false, false, PadLds<512>>;
  using Fi = WaveFft<512, 1, false, true, PadLds<512>>;
  static_assert(F::P::last() == Fi::P::first(), "register hand-off needs mirrored plans");
  constexpr int RS = F::P::last();
  __shared__ float2 smem[C::SMEM];
  __shared__ SplitTw tw;
  __shared__ uint32_t rmk[C::NROW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = wave >> 1, h = wave & 1;
  static_assert(RS == 8, "split_mults needs the positions of a last radix-8 stage");
  const int msrc = ((h ? 63 - lane : 64 - lane) & 63) * 4;
  const bool lane0h0 = h == 0 && lane == 0;
  int p, c0;
  col_tile<C>(p, c0);
  const size_t pb = (size_t)p * a.NX * C::NROW;
  float2* priv = smem + wave * C::PRIV;
  const TileOffs<C> to(a.NX, c0, a.quad != 0);
  const int iw = plane_iw(a, p);
  const float w2 = a.w2[iw];
  const int im = model_iw_any(a, iw);
  const uint32_t* rm = (MODE != 2 && a.rowmask) ? a.rowmask + im * a.NY : nullptr;
  const uint32_t pm = (MODE != 2 && a.planemask) ? a.planemask[im] : 0xffffffffu;
  const uint32_t live = a.nterm >= 32 ? pm : (pm & ((1u << a.nterm) - 1u));
  const bool masked = rm != nullptr;
  const int rend = masked ? 32 : a.nterm;
  tile_row_masks<C>(rmk, rm);
  split_tw_stage(tw, a.twy);
  __syncthreads();
  float k2v[C::V];
  split_k2<RS>(k2v, a, c0 + cl, h, lane);
  float2 acc[C::V];  // Y[2j + h]
#pragma unroll
  for (int e = 0; e < C::V; ++e) acc[e] = czero();
  auto fetch = [&](TileRegs<C>& t, int rr) {
    const PlaneBuf wb(a.W + (size_t)rr * a.wstride + pb, plane_bytes(a));
    if (masked) t.fetch_masked(wb, to, rmk, 1u << rr);
    else t.fetch(wb, to);
  };
  // the next term's tile is fetched into registers during this term's
  // transform where the budget holds it (2 waves per SIMD); at 4 the other
  // workgroup on the CU hides the load instead
  constexpr bool PREFETCH = WPE <= 2;
  int r = masked ? next_bit(live, -1) : 0;
  TileRegs<C> tile;
  if (PREFETCH && r < rend) fetch(tile, r);
  while (r < rend) {
    if constexpr (!PREFETCH) fetch(tile, r);
    tile.commit(smem);
    __syncthreads();
    float2 v[C::V];
    split_read<F::P::first()>(v, smem, cl, h, lane);
    const int rn = masked ? next_bit(live, r) : r + 1;
    if (PREFETCH && rn < rend) fetch(tile, rn);
    __syncthreads();
    F::run(v, priv, tw.h, lane);
    __syncthreads();
    split_write<RS>(smem, v, cl, h, lane);
    __syncthreads();
    float2 srv = make_float2(0.f, 0.f);
    if constexpr (MODE != 2) srv = split_sref(a, r, iw);
    const TermMult<MODE> tm(a, r, w2, srv);
    auto mult = [&](auto real) {
      float2 m[C::V];
      if constexpr (MODE == 2) {
#pragma unroll
        for (int e = 0; e < C::V; ++e) m[e] = tm.template at<decltype(real)::value>(k2v[e]);
      } else {
        split_mults<decltype(real)::value>(m, tm, k2v, msrc, lane0h0);
      }
#pragma unroll
      for (int e = 0; e < C::V; ++e)
        acc[e] = cadd(acc[e], cmul(split_spec(smem, tw, cl, split_n<RS>(lane, e, h)), m[e]));
    };
    if (tm.real) mult(std::true_type{});
    else mult(std::false_type{});
    __syncthreads();
    r = rn;
  }
  const float s2 = a.scale;
#pragma unroll
  for (int e = 0; e < C::V; ++e) acc[e] = cscale(acc[e], s2);
  Fi::run(acc, priv, tw.h, lane);
  __syncthreads();
  split_write<Fi::P::last()>(smem, acc, cl, h, lane);
  __syncthreads();
  tile_store_r2(smem, tw, PlaneBuf(a.K + pb, plane_bytes(a)), to, nullptr, 0);
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
// number of present (row, reference) pairs of every plane
__global__ __launch_bounds__(kThreads) void k_present_count(const uint32_t* __restrict__ rowmask, int* cnt, int NY,
                                                            long long nplanes, uint32_t refs) {
  const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= nplanes) return;
  int c = 0;
  for (int y = 0; y < NY; ++y) c += __builtin_popcount(rowmask[p * NY + y] & refs);
  cnt[p] = c;
}
hipError_t launch_present_count(const uint32_t* rowmask, int* cnt, int NY, long long nplanes, int nref,
                                hipStream_t st) {
  const uint32_t refs = nref >= 32 ? 0xffffffffu : ((1u << nref) - 1u);
  hipLaunchKernelGGL(k_present_count, dim3((unsigned)((nplanes + kThreads - 1) / kThreads)), dim3(kThreads), 0, st,
                     rowmask, cnt, NY, nplanes, refs);
  return hipGetLastError();
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
    const dim3 g = row_grid(N, a.np * a.NY);
    if (a.f64) {
      if (a.mplane) hipLaunchKernelGGL((k_rows_fanout<N, true, double2>), g, dim3(kThreads), 0, st, a);
      else hipLaunchKernelGGL((k_rows_fanout<N, false, double2>), g, dim3(kThreads), 0, st, a);
    } else {
      if (a.mplane) hipLaunchKernelGGL((k_rows_fanout<N, true>), g, dim3(kThreads), 0, st, a);
      else hipLaunchKernelGGL((k_rows_fanout<N, false>), g, dim3(kThreads), 0, st, a);
    }
    return hipGetLastError();
  });
}

hipError_t launch_rows_select(const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NX, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    const dim3 g = row_grid(N, a.np * a.NY);
    if (a.knext) {
      if (a.mplane) hipLaunchKernelGGL((k_rows_select<N, true, true>), g, dim3(kThreads), 0, st, a);
      else hipLaunchKernelGGL((k_rows_select<N, false, true>), g, dim3(kThreads), 0, st, a);
    } else {
      if (a.mplane) hipLaunchKernelGGL((k_rows_select<N, true>), g, dim3(kThreads), 0, st, a);
      else hipLaunchKernelGGL((k_rows_select<N, false>), g, dim3(kThreads), 0, st, a);
    }
    return hipGetLastError();
  });
}

hipError_t launch_cols(bool inv, const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NY, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    const dim3 g = col_grid(N, a.NX, a.np);
    if (inv) hipLaunchKernelGGL((k_cols<N, true>), g, dim3(col_threads(N)), 0, st, a);
    else hipLaunchKernelGGL((k_cols<N, false>), g, dim3(col_threads(N)), 0, st, a);
    return hipGetLastError();
  });
}

// ------------------------------------------------------------------ occupancy variants
// Every fan-out / fan-in column kernel is built for three register budgets:
// 2, 4 or 6 waves per SIMD (256, 128 or 80 VGPRs).  More waves hide more
// latency but may spill; which wins depends on the size and the transform,
// so the choice is a runtime table (kernel, mode, size) that an autotuner
// sets from timed launches; the defaults (col_wpe_default) are the winners
// of such timings.
static constexpr int kSizeList[] = {8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024};
static constexpr int kNumSizes = sizeof(kSizeList) / sizeof(int);
static int size_slot(int n) {
  for (int i = 0; i < kNumSizes; ++i)
    if (kSizeList[i] == n) return i;
  return -1;
}
static int g_wpe[2][3][kNumSizes];
static bool g_wpe_ready = false;
static void wpe_init() {
  if (g_wpe_ready) return;
  for (int k = 0; k < 2; ++k)
    for (int m = 0; m < 3; ++m)
      for (int i = 0; i < kNumSizes; ++i) g_wpe[k][m][i] = col_wpe_default(k, m, kSizeList[i]);
  g_wpe_ready = true;
}
int col_wpe_get(int kind, int mode, int n) {
  const int i = size_slot(n);
  if (kind < 0 || kind > 1 || mode < 0 || mode > 2 || i < 0) return -1;
  wpe_init();
  return g_wpe[kind][mode][i];
}
// Layout of the 1024-point column transforms, per kind (0 fan-out, 1 fan-in)
// and mode: 0 one wave per column (256-thread workgroups, 4 columns), 1 two
// waves per column (512-thread workgroups, 4 columns, radix-2 split).
static int g_layout[2][3] = {{1, 1, 1}, {1, 1, 1}};
// Layout of the K / W scratch per lateral size (the column length NY): 1
// quad-block, 0 row-major.  The quad-block layout pays where the column tiles
// are 4 columns wide (1024-point columns, whole lines instead of quarter
// lines); at 8 columns the row passes' partial-line accesses cost more than
// the column passes gain.
static int g_quad[16] = {};
static bool g_quad_init = false;
static void quad_init() {
  if (g_quad_init) return;
  g_quad[size_slot(1024)] = 1;
  g_quad_init = true;
}
int scratch_quad_get(int n) {
  const int i = size_slot(n);
  if (i < 0) return -1;
  quad_init();
  return g_quad[i];
}
int scratch_quad_set(int n, int quad) {
  const int i = size_slot(n);
  if (i < 0 || (quad != 0 && quad != 1)) return -1;
  quad_init();
  g_quad[i] = quad;
  return 0;
}
int col_layout_get(int kind, int mode, int n) {
  if (kind < 0 || kind > 1 || mode < 0 || mode > 2 || size_slot(n) < 0) return -1;
  return n == 1024 ? g_layout[kind][mode] : 0;
}
int col_layout_set(int kind, int mode, int n, int layout) {
  if (kind < 0 || kind > 1 || mode < 0 || mode > 2 || size_slot(n) < 0) return -1;
  if (layout != 0 && !(layout == 1 && n == 1024)) return -1;
  if (n == 1024) g_layout[kind][mode] = layout;
  return 0;
}
int col_wpe_set(int kind, int mode, int n, int wpe) {
  const int i = size_slot(n);
  if (kind < 0 || kind > 1 || mode < 0 || mode > 2 || i < 0) return -1;
  if (wpe != 2 && wpe != 4 && wpe != 6) return -1;
  wpe_init();
  g_wpe[kind][mode][i] = wpe;
  return 0;
}

template <class Fn>
static hipError_t dispatch_wpe(int wpe, Fn&& fn) {
  switch (wpe) {
    case 2: return fn(std::integral_constant<int, 2>{});
    case 4: return fn(std::integral_constant<int, 4>{});
    case 6: return fn(std::integral_constant<int, 6>{});
    default: return hipErrorInvalidValue;
  }
}

// float64 arithmetic (a.f64): one register budget, 2 waves per SIMD (256 VGPRs)
hipError_t launch_cols_fanout(int mode, const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NY, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    const dim3 g = col_grid(N, a.NX, a.np);
    if constexpr (N == 1024) {
      if (!a.f64 && col_layout_get(0, mode, N) == 1)
        return dispatch_wpe(col_wpe_get(0, mode, N), [&](auto wc) -> hipError_t {
          constexpr int W = decltype(wc)::value;
          if (mode == 0) hipLaunchKernelGGL((k_cols_fanout_split<0, W>), g, dim3(SplitCfg::THREADS), 0, st, a);
          else if (mode == 1) hipLaunchKernelGGL((k_cols_fanout_split<1, W>), g, dim3(SplitCfg::THREADS), 0, st, a);
          else hipLaunchKernelGGL((k_cols_fanout_split<2, W>), g, dim3(SplitCfg::THREADS), 0, st, a);
          return hipGetLastError();
        });
    }
    if (a.f64) {
      if (mode == 0) hipLaunchKernelGGL((k_cols_fanout<N, 0, 2, double2>), g, dim3(col_threads(N)), 0, st, a);
      else if (mode == 1) hipLaunchKernelGGL((k_cols_fanout<N, 1, 2, double2>), g, dim3(col_threads(N)), 0, st, a);
      else hipLaunchKernelGGL((k_cols_fanout<N, 2, 2, double2>), g, dim3(col_threads(N)), 0, st, a);
      return hipGetLastError();
    }
    return dispatch_wpe(col_wpe_get(0, mode, N), [&](auto wc) -> hipError_t {
      constexpr int W = decltype(wc)::value;
      if (mode == 0) hipLaunchKernelGGL((k_cols_fanout<N, 0, W>), g, dim3(col_threads(N)), 0, st, a);
      else if (mode == 1) hipLaunchKernelGGL((k_cols_fanout<N, 1, W>), g, dim3(col_threads(N)), 0, st, a);
      else hipLaunchKernelGGL((k_cols_fanout<N, 2, W>), g, dim3(col_threads(N)), 0, st, a);
      return hipGetLastError();
    });
  });
}

hipError_t launch_cols_fanin(int mode, const KArgs& a, hipStream_t st) {
  return dispatch_n(a.NY, [&](auto nc) -> hipError_t {
    constexpr int N = decltype(nc)::value;
    const dim3 g = col_grid(N, a.NX, a.np);
    if constexpr (N == 1024) {
      if (!a.f64 && col_layout_get(1, mode, N) == 1)
        return dispatch_wpe(col_wpe_get(1, mode, N), [&](auto wc) -> hipError_t {
          constexpr int W = decltype(wc)::value;
          if (mode == 0) hipLaunchKernelGGL((k_cols_fanin_split<0, W>), g, dim3(SplitCfg::THREADS), 0, st, a);
          else if (mode == 1) hipLaunchKernelGGL((k_cols_fanin_split<1, W>), g, dim3(SplitCfg::THREADS), 0, st, a);
          else hipLaunchKernelGGL((k_cols_fanin_split<2, W>), g, dim3(SplitCfg::THREADS), 0, st, a);
          return hipGetLastError();
        });
    }
    if (a.f64) {
      if (mode == 0) hipLaunchKernelGGL((k_cols_fanin<N, 0, 2, double2>), g, dim3(col_threads(N)), 0, st, a);
      else if (mode == 1) hipLaunchKernelGGL((k_cols_fanin<N, 1, 2, double2>), g, dim3(col_threads(N)), 0, st, a);
      else hipLaunchKernelGGL((k_cols_fanin<N, 2, 2, double2>), g, dim3(col_threads(N)), 0, st, a);
      return hipGetLastError();
    }
    return dispatch_wpe(col_wpe_get(1, mode, N), [&](auto wc) -> hipError_t {
      constexpr int W = decltype(wc)::value;
      if (mode == 0) hipLaunchKernelGGL((k_cols_fanin<N, 0, W>), g, dim3(col_threads(N)), 0, st, a);
      else if (mode == 1) hipLaunchKernelGGL((k_cols_fanin<N, 1, W>), g, dim3(col_threads(N)), 0, st, a);
      else hipLaunchKernelGGL((k_cols_fanin<N, 2, W>), g, dim3(col_threads(N)), 0, st, a);
      return hipGetLastError();
    });
  });
}

}  // namespace fwix
