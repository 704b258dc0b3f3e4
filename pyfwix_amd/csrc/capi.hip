// C ABI of libpyfwix_hip.so: operator handles over the HIP kernels.
// Declarations and reference citations: include/fwix_hip.h.
#include "fwix_hip.h"
#include "fwix_launch.h"
#include "capi_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace fwix;

namespace fwix {

thread_local std::string g_err;

void set_error(const std::string& s) { g_err = s; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ live pass timing
namespace {
struct EvRec {
  const char* name;
  double bytes;
  hipEvent_t a, b;
};
struct PassStat {
  double ms = 0, bytes = 0;
  long long n = 0;  // recorded launches
};
std::mutex g_tmu;
bool g_timing = false;
std::string g_only;        // time only this pass ("" = every pass)
long long g_every = 1;     // ... and only every g_every-th launch of each pass
std::map<std::string, long long> g_seen;  // launches per pass while enabled
std::vector<EvRec> g_pending;
std::vector<hipEvent_t> g_pool;
std::map<std::string, PassStat> g_stat;

hipEvent_t take_event() {
  if (g_pool.empty()) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  hipEvent_t e = g_pool.back();
  g_pool.pop_back();
  return e;
}
}  // namespace

PassTimer::PassTimer(const char* name, double bytes, hipStream_t s) {
  if (!g_timing || !name) return;
  // no event pairs inside a captured graph: its replays would never be collected
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return;
  std::lock_guard<std::mutex> g(g_tmu);
  const long long k = g_seen[name]++;
  if (!g_only.empty() && g_only != name) return;
  if (k % g_every) return;
  EvRec r{name, bytes, take_event(), take_event()};
  hip_check(hipEventRecord(r.a, s), "hipEventRecord");
  slot = (int)g_pending.size();
  st = s;
  g_pending.push_back(r);
}
PassTimer::~PassTimer() {
  if (slot < 0) return;
  std::lock_guard<std::mutex> g(g_tmu);
  (void)hipEventRecord(g_pending[slot].b, st);
}

void DevBuf::alloc(size_t bytes) {
  release();
  if (bytes == 0) return;
  hip_check(hipMalloc(&p, bytes), "hipMalloc");
  n = bytes;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  n = 0;
}

static void upload_bytes(DevBuf& b, const void* host, size_t bytes) {
  b.alloc(bytes);
  if (bytes) hip_check(hipMemcpy(b.p, host, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

static std::vector<float> k_table(int n, double d) {  // PhaseShift.h:43-57
  std::vector<float> k(n);
  const double dk = 2.0 * M_PI / (d * n);
  for (int i = 0; i < n; ++i) k[i] = (float)((i <= n / 2 ? i : i - n) * dk);
  return k;
}

static std::vector<double> taper_table64(int n, int t) {  // taper.cu:43-55
  std::vector<double> w(n, 1.0);
  if (t > 0)
    for (int i = 0; i < n; ++i) {
      if (i < t) w[i] = 0.5 * (1.0 - std::cos(M_PI * i / t));
      else if (i >= n - t) w[i] = 0.5 * (1.0 + std::cos(M_PI * (i - n + t) / t));
    }
  return w;
}
static std::vector<float> taper_table(int n, int t) {
  const auto w64 = taper_table64(n, t);
  return std::vector<float>(w64.begin(), w64.end());
}

static std::vector<double2> twiddle_table64(int n) {  // exp(-2 pi i k / n) in double
  std::vector<double2> t(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    t[k] = make_double2(std::cos(a), std::sin(a));
  }
  return t;
}
static std::vector<float2> twiddle_table(int n) {  // the float64 table rounded once
  std::vector<float2> t(n);
  const auto t64 = twiddle_table64(n);
  for (int k = 0; k < n; ++k) t[k] = make_float2((float)t64[k].x, (float)t64[k].y);
  return t;
}

void validate_dims(const fwix_dims* d, bool need_fft) {
  if (!d) throw Error("null dims");
  if (d->NX < 1 || d->NY < 1 || d->NW < 1 || d->NS < 1) throw Error("dims: NX, NY, NW, NS must be >= 1");
  if (d->nx < 1 || d->ny < 1 || d->nx > d->NX || d->ny > d->NY)
    throw Error("dims: need 1 <= nx <= NX and 1 <= ny <= NY");
  if (!(d->dx > 0.f) || !(d->dy > 0.f)) throw Error("dims: dx and dy must be > 0");
  if (need_fft && (!fft_size_supported(d->NX) || !fft_size_supported(d->NY)))
    throw Error("dims: NX=" + std::to_string(d->NX) + ", NY=" + std::to_string(d->NY) +
                ": lateral FFT sizes must be 2^a or 3*2^a between 8 and 1024");
  const double pts = (double)d->NX * d->NY * d->NW * d->NS;
  if (pts > 9.0e18 / 8) throw Error("dims: wavefield too large");
}

void validate_params(const fwix_params* p) {
  if (!p) throw Error("null params");
  if (p->nref < 1 || p->nref > 255) throw Error("params: nref must be in [1, 255]");
  if (p->taperx < 0 || p->tapery < 0) throw Error("params: taperx/tapery must be >= 0");
  if (p->taylor < 0 || p->taylor > 4) throw Error("params: taylor must be in [0, 4]");
  if (p->planes_per_chunk < 0) throw Error("params: planes_per_chunk must be >= 0");
}

void Tables::init(const fwix_dims& d, int tx, int ty) {
  if (tx > d.NX || ty > d.NY) throw Error("taper wider than the padded grid");
  auto kx = k_table(d.NX, d.dx), ky = k_table(d.NY, d.dy);
  std::vector<float> w2(d.NW), w(d.NW);
  for (int i = 0; i < d.NW; ++i) {  // PhaseShift.h:59-70, Scatter.h:91-103
    const double f = 2.0 * M_PI * ((double)d.ow + (double)i * d.dw);
    w2[i] = (float)(f * f);
    w[i] = (float)f;
  }
  auto tx_ = taper_table(d.NX, tx), ty_ = taper_table(d.NY, ty);
  auto tx64 = taper_table64(d.NX, tx), ty64 = taper_table64(d.NY, ty);
  upload_bytes(b_tapx64, tx64.data(), tx64.size() * 8);
  upload_bytes(b_tapy64, ty64.data(), ty64.size() * 8);
  upload_bytes(b_kx, kx.data(), kx.size() * 4);
  upload_bytes(b_ky, ky.data(), ky.size() * 4);
  upload_bytes(b_w2, w2.data(), w2.size() * 4);
  upload_bytes(b_w, w.data(), w.size() * 4);
  upload_bytes(b_tapx, tx_.data(), tx_.size() * 4);
  upload_bytes(b_tapy, ty_.data(), ty_.size() * 4);
  if (fft_size_supported(d.NX) && fft_size_supported(d.NY)) {
    auto twx = twiddle_table(d.NX), twy = twiddle_table(d.NY);
    upload_bytes(b_twx, twx.data(), twx.size() * 8);
    upload_bytes(b_twy, twy.data(), twy.size() * 8);
    auto twx64 = twiddle_table64(d.NX), twy64 = twiddle_table64(d.NY);
    upload_bytes(b_twx64, twx64.data(), twx64.size() * 16);
    upload_bytes(b_twy64, twy64.data(), twy64.size() * 16);
  }
  dims = d;
}

void Tables::set_frequencies(const float* f_hz, bool nonzero) {
  if (!f_hz) throw Error("null frequency list");
  std::vector<float> w2(dims.NW), w(dims.NW);
  for (int i = 0; i < dims.NW; ++i) {
    if (!std::isfinite(f_hz[i])) throw Error("frequencies must be finite");
    if (nonzero && f_hz[i] == 0.f) throw Error("Frequency is zero in the scattering operator!");
    const double f = 2.0 * M_PI * (double)f_hz[i];
    w2[i] = (float)(f * f);
    w[i] = (float)f;
  }
  upload_bytes(b_w2, w2.data(), w2.size() * 4);
  upload_bytes(b_w, w.data(), w.size() * 4);
}

KArgs Tables::kargs() const {
  KArgs a;
  std::memset(&a, 0, sizeof(a));
  a.kx = (const float*)b_kx.p;
  a.ky = (const float*)b_ky.p;
  a.w2 = (const float*)b_w2.p;
  a.w = (const float*)b_w.p;
  a.tapx = (const float*)b_tapx.p;
  a.tapy = (const float*)b_tapy.p;
  a.twx = (const float2*)b_twx.p;
  a.twy = (const float2*)b_twy.p;
  a.twx64 = (const double2*)b_twx64.p;
  a.twy64 = (const double2*)b_twy64.p;
  a.tapx64 = (const double*)b_tapx64.p;
  a.tapy64 = (const double*)b_tapy64.p;
  a.NX = dims.NX; a.NY = dims.NY; a.NW = dims.NW; a.NS = dims.NS;
  a.nx = dims.nx; a.ny = dims.ny;
  a.dz = dims.dz;
  a.quad = scratch_quad_get(dims.NY) == 1 ? 1 : 0;
  return a;
}

PwArgs Tables::pwargs() const {
  PwArgs a;
  std::memset(&a, 0, sizeof(a));
  a.NX = dims.NX; a.NY = dims.NY; a.NW = dims.NW; a.NS = dims.NS;
  a.kx = (const float*)b_kx.p;
  a.ky = (const float*)b_ky.p;
  a.w2 = (const float*)b_w2.p;
  a.tapx = (const float*)b_tapx.p;
  a.tapy = (const float*)b_tapy.p;
  return a;
}

int auto_chunk(const fwix_dims& d, int nterm, int requested) {
  const int planes = d.NW * d.NS;
  // The pass kernels index a chunk's planes with 32-bit element offsets
  // (row * NX + x), so one chunk stays below 2^31 complex samples.
  const int c_index = (int)std::min<long long>(planes, ((1LL << 31) - 1) / ((long long)d.NX * d.NY));
  if (requested > 0) return std::min(requested, c_index);
  // Enough planes that a column pass launches >= 32768 workgroups (128 per
  // CU), so every launch reaches a steady state of loads overlapping
  // transforms and the partly filled last round of workgroups is a small part
  // of it.  Born steps (PSPI, 4 references) per padded size, planes per
  // launch -> Gpts*freq/s: 1024^2 16 -> 32.1, 32 -> 33.8, 128 -> 34.9;
  // 768^2 43 -> 37.5, 128 -> 39.4, 256 -> 39.8; 512^2 64 -> 55.1, 256 ->
  // 57.6; 384^2 171 -> 31.0, 256 -> 32.0.  Steps with more than 4 fan-out
  // planes keep >= 4096 workgroups (16 per CU): NSPS at 768^2 with 8
  // references, 100 planes, ran 21.25-21.45 at 34 planes per launch and
  // 20.84-21.31 at 50 or 100.  At most 8 GiB of scratch (spectrum + nterm
  // fan-out planes).  Chunks are then balanced so the last one is not a
  // small tail.
  const double per_plane = (double)(1 + nterm) * d.NX * d.NY * 8.0;
  const int cw = fft_size_supported(d.NY) ? col_tile_width(d.NY) : 16;
  const int blocks_per_plane = (d.NX + cw - 1) / cw;
  const int target = nterm > 4 ? 4096 : 32768;
  int c = (target + blocks_per_plane - 1) / blocks_per_plane;
  const int c_max = (int)(8192.0 * 1024 * 1024 / per_plane);
  if (c > c_max) c = c_max;
  if (c < 1) c = 1;
  if (c > c_index) c = c_index;
  if (c >= planes) return planes;
  const int nchunks = (planes + c - 1) / c;
  return (planes + nchunks - 1) / nchunks;
}

// ------------------------------------------------------------------ one step engine
void StepEngine::init(int method_, const fwix_dims& d, const fwix_params& p, int wterms) {
  plain = method_ == FWIX_PSPI_PLAIN;
  method = plain ? FWIX_PSPI : method_;
  params = p;
  tables.init(d, method == FWIX_PSPI ? p.taperx : 0, method == FWIX_PSPI ? p.tapery : 0);
  nmodel = d.NW;
  mplane.clear();
  const size_t plane = (size_t)d.NX * d.NY;
  if (plain) {  // two whole-shot scratch wavefields
    if ((long long)d.NW * d.NX * d.NY >= (1LL << 31))
      throw Error("the per-operator path transforms a whole shot per launch: NW * NX * NY must stay below 2^31");
    chunk = d.NW;
    W.release();
    K.alloc(plane * d.NW * sizeof(float2));
    T.alloc(plane * d.NW * sizeof(float2));
    return;
  }
  T.release();
  if (wterms < p.nref) wterms = p.nref;
  chunk = auto_chunk(d, wterms, p.planes_per_chunk);
  K.alloc(plane * chunk * sizeof(float2));
  W.alloc(plane * chunk * wterms * sizeof(float2));
}

void StepEngine::set_model_map(int nshot, const int* nplanes) {
  const fwix_dims& d = tables.dims;
  if (plain) throw Error("shot batches run on the fused path only");
  if (d.NS != 1) throw Error("a shot batch needs NS == 1 (the shots are laid along the plane axis)");
  if (nshot < 1 || !nplanes) throw Error("a shot batch needs at least one shot");
  std::vector<int> map;
  map.reserve(d.NW);
  for (int k = 0; k < nshot; ++k) {
    if (nplanes[k] < 1) throw Error("every shot of a batch needs at least one plane");
    map.insert(map.end(), (size_t)nplanes[k], k);
    if ((long long)map.size() > d.NW) break;
  }
  if ((long long)map.size() != d.NW) throw Error("the shots' plane counts must add up to NW");
  mplane = map;
  nmodel = nshot;
  upload_bytes(mplane_dev, mplane.data(), mplane.size() * sizeof(int));
  sref = nullptr;  // references and labels are per shot now: set them again
  labels = nullptr;
  masks = false;
}

void StepEngine::set_refs(const float2* sref_, const uint8_t* labels_) {
  if (!sref_ || !labels_) throw Error("null reference slownesses or labels");
  const fwix_dims& d = tables.dims;
  sref = sref_;
  if (!mplane.empty()) {  // (nz, nref, nshot) per shot -> (nz, nref, NW) per plane
    const size_t nin = (size_t)d.nz * params.nref * nmodel, nout = (size_t)d.nz * params.nref * d.NW;
    std::vector<float2> hin(nin), hout(nout);
    hip_check(hipDeviceSynchronize(), "synchronize before reference expansion");
    hip_check(hipMemcpy(hin.data(), sref_, nin * sizeof(float2), hipMemcpyDeviceToHost), "references D2H");
    for (size_t zr = 0; zr < (size_t)d.nz * params.nref; ++zr)
      for (int w = 0; w < d.NW; ++w) hout[zr * d.NW + w] = hin[zr * nmodel + mplane[w]];
    upload_bytes(sref_exp, hout.data(), nout * sizeof(float2));
    sref = (const float2*)sref_exp.p;
  }
  labels = labels_;
  masks = false;
  if (params.nref > 32) return;  // presence skipping covers references 0..31
  const long long nplanes = (long long)d.nz * nmodel;
  rowmask.alloc((size_t)nplanes * d.NY * sizeof(uint32_t));
  planemask.alloc((size_t)nplanes * sizeof(uint32_t));
  hip_check(hipDeviceSynchronize(), "synchronize before presence masks");
  hip_check(launch_presence(labels, (uint32_t*)rowmask.p, (uint32_t*)planemask.p, d.NX, d.NY, nplanes, nullptr),
            "presence masks");
  DevBuf cnt;
  cnt.alloc((size_t)nplanes * sizeof(int));
  hip_check(launch_present_count((const uint32_t*)rowmask.p, (int*)cnt.p, d.NY, nplanes, params.nref, nullptr),
            "presence counts");
  present.assign((size_t)nplanes, 0);
  hip_check(hipMemcpy(present.data(), cnt.p, (size_t)nplanes * sizeof(int), hipMemcpyDeviceToHost), "counts D2H");
  masks = true;
}

double StepEngine::present_rows(int iz, int p0, int np) const {
  const fwix_dims& d = tables.dims;
  if (!masks) return (double)params.nref * np * d.NY;
  double c = 0;
  for (int p = 0; p < np; ++p) c += present[(size_t)iz * nmodel + model_of((p0 + p) % d.NW)];
  return c;
}

void StepEngine::check_ready(int iz) const {
  if (!sref || !labels) throw Error("reference slownesses / labels not set (call set_refs)");
  if (iz < 0 || iz >= tables.dims.nz) throw Error("depth index out of range");
}

KArgs StepEngine::base(int iz, int c0, int np) const {
  const fwix_dims& d = tables.dims;
  KArgs a = tables.kargs();
  a.K = (float2*)K.p;
  a.W = (float2*)W.p;
  a.labels = labels + (size_t)iz * nmodel * d.NY * d.NX;
  a.sref = sref + (size_t)iz * params.nref * d.NW;
  a.mplane = mplane.empty() ? nullptr : (const int*)mplane_dev.p;
  if (masks) {
    a.rowmask = (const uint32_t*)rowmask.p + (size_t)iz * nmodel * d.NY;
    a.planemask = (const uint32_t*)planemask.p + (size_t)iz * nmodel;
  }
  a.nref = params.nref;
  a.nterm = params.nref;
  a.p0 = c0;
  a.np = np;
  a.wstride = (long long)np * d.NX * d.NY;
  a.eps = params.eps;
  a.eps_s = params.eps_scat;
  return a;
}

// One depth step over every plane of `in`, run as chunks of `chunk` planes
// through the three fused passes.  `proto` carries the tables, scratch,
// depth slices and masks (its p0 / np / wstride are set per chunk);
// `rows(p0, np)` gives the per-reference rows a chunk moves, for the pass
// timers (null: no timers).
void step_passes(int method, bool adjoint, const fwix_dims& d, const fwix_params& p, const KArgs& proto, int chunk,
                 int add, const float2* in, float2* out, hipStream_t st,
                 const std::function<double(int, int)>* rows, float2* kall, bool k_ready, bool k_next,
                 bool out_needed) {
  const size_t plane = (size_t)d.NX * d.NY;
  const int planes = d.NW * d.NS;
  const float sx = (float)(1.0 / std::sqrt((double)d.NX));
  const float sy2 = (float)(1.0 / (double)d.NY);
  const bool pspi = method == FWIX_PSPI;
  const bool taper = pspi && (p.taperx > 0 || p.tapery > 0);
  for (int c0 = 0; c0 < planes; c0 += chunk) {
    const int np = (planes - c0) < chunk ? (planes - c0) : chunk;
    const double pb = 8.0 * np * plane;  // bytes of one chunk-sized complex64 array
    const double wb = rows ? 8.0 * d.NX * (*rows)(c0, np) : 0.0;  // per-reference rows moved
    auto timer = [&](const char* name, double bytes) { return PassTimer(rows ? name : nullptr, bytes, st); };
    KArgs a = proto;
    a.p0 = c0;
    a.np = np;
    a.wstride = (long long)np * plane;
    if (kall) a.K = kall + (size_t)c0 * plane;
    const float2* cin = in + (size_t)c0 * plane;
    float2* cout = out + (size_t)c0 * plane;
    // "shared spectrum" shape: PSPI forward and NSPS adjoint
    if ((pspi && !adjoint) || (!pspi && adjoint)) {
      if (!k_ready) {
        a.in = cin; a.out = a.K; a.qout = 1; a.add = 0; a.scale = sx;  // K in the quad-block layout
        { auto t = timer("rows_fwd", 2 * pb); hip_check(launch_rows(taper ? 2 : 0, a, st), "rows pass"); }
        a.qout = 0;
      }
      a.scale = sy2;
      { auto t = timer("cols_fanout", pb + wb); hip_check(launch_cols_fanout(adjoint ? 1 : 0, a, st), "column fan-out pass"); }
      a.out = cout; a.add = add; a.scale = sx; a.knext = k_next ? (out_needed ? 1 : 2) : 0;
      const double ob = (k_next && !out_needed) ? 0.0 : (1 + (add ? 1 : 0)) * pb;
      { auto t = timer("rows_select", wb + ob + (k_next ? pb : 0.0) + pb / 8); hip_check(launch_rows_select(a, st), "rows select pass"); }
      a.knext = 0;
    } else {  // "masked inputs" shape: PSPI adjoint and NSPS forward
      a.in = cin; a.scale = sx;
      { auto t = timer("rows_fanout", pb + pb / 8 + wb); hip_check(launch_rows_fanout(a, st), "rows fan-out pass"); }
      a.scale = sy2;
      { auto t = timer("cols_fanin", wb + pb); hip_check(launch_cols_fanin(adjoint ? 1 : 0, a, st), "column fan-in pass"); }
      a.in = a.K; a.qin = 1; a.out = cout; a.add = add; a.scale = sx;
      { auto t = timer("rows_inv", (2 + (add ? 1 : 0)) * pb); hip_check(launch_rows(taper ? 3 : 1, a, st), "rows inverse pass"); }
    }
  }
}

// out (+)= P_z in     (forward) or    out (+)= P_z^H in    (adjoint)
void StepEngine::apply(bool adjoint, int iz, int add, const float2* in, float2* out, hipStream_t st) {
  check_ready(iz);
  if (plain) return apply_plain(adjoint, iz, add, in, out, st);
  const std::function<double(int, int)> rows = [&](int c0, int np) { return present_rows(iz, c0, np); };
  step_passes(method, adjoint, tables.dims, params, base(iz, 0, chunk), chunk, add, in, out, st, &rows);
}

void StepEngine::apply_chained(int iz, const float2* in, float2* out, bool k_ready, bool k_next, hipStream_t st,
                               bool out_needed, bool adjoint) {
  if (!chains(adjoint)) return apply(adjoint, iz, 0, in, out, st);
  check_ready(iz);
  const fwix_dims& d = tables.dims;
  const size_t bytes = (size_t)d.NX * d.NY * d.NW * d.NS * sizeof(float2);
  if (Kall.n != bytes) {
    if (k_ready) throw Error("chained step: no row spectrum left by a previous step");
    Kall.alloc(bytes);
  }
  const std::function<double(int, int)> rows = [&](int c0, int np) { return present_rows(iz, c0, np); };
  step_passes(method, adjoint, d, params, base(iz, 0, chunk), chunk, 0, in, out, st, &rows, (float2*)Kall.p, k_ready,
              k_next, out_needed || !k_next);
}

// The PSPI step at the reference's operator granularity, one shot per launch
// (PSPI.cpp:66-103): taper, 2-D FFT (rows + columns), then per reference a
// phase shift, an inverse 2-D FFT and a select-accumulate; the adjoint mirrors
// it.  Same arithmetic as the fused path, one HBM round trip per operator.
void StepEngine::apply_plain(bool adjoint, int iz, int add, const float2* in, float2* out, hipStream_t st) {
  const fwix_dims& d = tables.dims;
  const size_t plane = (size_t)d.NX * d.NY, shot = plane * d.NW;
  const double sb = 8.0 * shot;
  const float sx = (float)(1.0 / std::sqrt((double)d.NX)), sy = (float)(1.0 / std::sqrt((double)d.NY));
  float2* Kb = (float2*)K.p;
  float2* Tb = (float2*)T.p;
  PwArgs pa = tables.pwargs();
  pa.NS = 1;
  pa.dz = d.dz;
  pa.eps = params.eps;
  pa.labels = labels + (size_t)iz * d.NW * d.NY * d.NX;
  auto fft = [&](bool inv, float2* buf) {  // in-place unitary 2-D FFT of one shot
    KArgs f = tables.kargs();
    f.NS = 1; f.p0 = 0; f.np = d.NW;
    f.in = buf; f.out = buf; f.add = 0; f.scale = sx;
    { PassTimer t("plain_fft_rows", 2 * sb, st); hip_check(launch_rows(inv ? 1 : 0, f, st), "plain fft rows"); }
    f.scale = sy;
    { PassTimer t("plain_fft_cols", 2 * sb, st); hip_check(launch_cols(inv, f, st), "plain fft columns"); }
  };
  for (int is = 0; is < d.NS; ++is) {
    const float2* sin_ = in + (size_t)is * shot;
    float2* sout = out + (size_t)is * shot;
    if (!adjoint) {
      pa.add = 0;
      { PassTimer t("plain_taper", 2 * sb, st); hip_check(launch_taper(sin_, Kb, pa, st), "plain taper"); }
      fft(false, Kb);
      if (!add) hip_check(launch_zero(sout, shot, st), "plain zero");
      for (int r = 0; r < params.nref; ++r) {
        pa.sref = sref + ((size_t)iz * params.nref + r) * d.NW;
        pa.add = 0;
        { PassTimer t("plain_ps", 2 * sb, st); hip_check(launch_phase_shift(0, Kb, Tb, pa, st), "plain phase shift"); }
        fft(true, Tb);
        pa.value = r; pa.add = 1;
        { PassTimer t("plain_select", 3 * sb + sb / 8, st); hip_check(launch_select(Tb, sout, pa, st), "plain select"); }
      }
    } else {
      hip_check(launch_zero(Kb, shot, st), "plain zero");
      for (int r = 0; r < params.nref; ++r) {
        pa.value = r; pa.add = 0;
        { PassTimer t("plain_select", 2 * sb + sb / 8, st); hip_check(launch_select(sin_, Tb, pa, st), "plain select"); }
        fft(false, Tb);
        pa.sref = sref + ((size_t)iz * params.nref + r) * d.NW;
        pa.add = 1;
        { PassTimer t("plain_ps", 3 * sb, st); hip_check(launch_phase_shift(1, Tb, Kb, pa, st), "plain phase shift"); }
      }
      fft(true, Kb);
      pa.add = add;
      { PassTimer t("plain_taper", (2 + (add ? 1 : 0)) * sb, st); hip_check(launch_taper(Kb, sout, pa, st), "plain taper"); }
    }
  }
}

}  // namespace fwix

// =================================================================== C entry points
namespace {

struct FftOp : fwix_op_s {
  Tables t;
  DevBuf K;
  int chunk = 1;
  FftOp() { kind = KIND_FFT; }
  void run(bool inv, int add, const float2* in, float2* out, hipStream_t st) {
    const fwix_dims& d = t.dims;
    const size_t plane = (size_t)d.NX * d.NY;
    const int planes = d.NW * d.NS;
    for (int c0 = 0; c0 < planes; c0 += chunk) {
      const int np = (planes - c0) < chunk ? (planes - c0) : chunk;
      KArgs a = t.kargs();
      a.p0 = c0; a.np = np;
      a.in = in + (size_t)c0 * plane; a.out = (float2*)K.p; a.add = 0;
      a.scale = (float)(1.0 / std::sqrt((double)d.NX));
      hip_check(launch_rows(inv ? 1 : 0, a, st), "fft rows");
      a.in = (const float2*)K.p; a.out = out + (size_t)c0 * plane; a.add = add;
      a.scale = (float)(1.0 / std::sqrt((double)d.NY));
      hip_check(launch_cols(inv, a, st), "fft columns");
    }
  }
};

struct PhaseShiftOp : fwix_op_s {
  Tables t;
  DevBuf sref;
  float dz = 0.f, eps = 0.f;
  PhaseShiftOp() { kind = KIND_PHASE_SHIFT; }
  void run(int mode, int add, const float2* in, float2* out, hipStream_t st) {
    PwArgs a = t.pwargs();
    a.sref = (const float2*)sref.p;
    a.dz = dz; a.eps = eps; a.add = add;
    hip_check(launch_phase_shift(mode, in, out, a, st), "phase shift");
  }
};

struct SelectorOp : fwix_op_s {
  Tables t;
  DevBuf labels;
  int value = 0;
  bool has_labels = false;
  SelectorOp() { kind = KIND_SELECTOR; }
  void run(int add, const float2* in, float2* out, hipStream_t st) {
    if (!has_labels) throw Error("selector labels not set");
    PwArgs a = t.pwargs();
    a.labels = (const uint8_t*)labels.p;
    a.value = value; a.add = add;
    hip_check(launch_select(in, out, a, st), "select");
  }
};

struct TaperOp : fwix_op_s {
  Tables t;
  TaperOp() { kind = KIND_TAPER; }
  void run(int add, const float2* in, float2* out, hipStream_t st) {
    PwArgs a = t.pwargs();
    a.add = add;
    hip_check(launch_taper(in, out, a, st), "taper");
  }
};

struct OneStepOp : fwix_op_s {
  StepEngine eng;
  int iz = 0;
  OneStepOp() { kind = KIND_ONESTEP; }
};

template <class T>
T* as(fwix_op op, int kind) {
  if (!op) throw Error("null operator handle");
  if (op->kind != kind) throw Error("operator handle of the wrong kind");
  return static_cast<T*>(op);
}

inline hipStream_t S(fwix_stream s) { return (hipStream_t)s; }

}  // namespace

extern "C" {

int fwix_version(void) { return 100; }
const char* fwix_last_error(void) { return fwix::g_err.c_str(); }

void fwix_params_default(fwix_params* p) {
  if (!p) return;
  p->nref = 1; p->taperx = 0; p->tapery = 0; p->eps = 0.f; p->eps_scat = 0.04f;
  p->taylor = 1; p->planes_per_chunk = 0;
}

int fwix_fft_size_supported(int n) { return fft_size_supported(n) ? 1 : 0; }

int fwix_op_destroy(fwix_op op) {
  return guard([&] { delete op; });
}

// ---------------------------------------------------------------- FFT
int fwix_fft2d_create(const fwix_dims* d, fwix_op* out) {
  return guard([&] {
    validate_dims(d, true);
    if (!out) throw Error("null output handle");
    auto* op = new FftOp();
    try {
      op->t.init(*d, 0, 0);
      op->chunk = auto_chunk(*d, 0, 0);
      op->K.alloc((size_t)d->NX * d->NY * op->chunk * sizeof(float2));
    } catch (...) { delete op; throw; }
    *out = op;
  });
}
int fwix_fft2d_forward(fwix_op op, int add, const fwix_c64* m, fwix_c64* d, fwix_stream s) {
  return guard([&] { as<FftOp>(op, KIND_FFT)->run(false, add, (const float2*)m, (float2*)d, S(s)); });
}
int fwix_fft2d_adjoint(fwix_op op, int add, fwix_c64* m, const fwix_c64* d, fwix_stream s) {
  return guard([&] { as<FftOp>(op, KIND_FFT)->run(true, add, (const float2*)d, (float2*)m, S(s)); });
}

// ---------------------------------------------------------------- PhaseShift
int fwix_phase_shift_create(const fwix_dims* d, float dz, float eps, fwix_op* out) {
  return guard([&] {
    validate_dims(d, false);
    if (!out) throw Error("null output handle");
    auto* op = new PhaseShiftOp();
    try {
      op->t.init(*d, 0, 0);
      op->dz = dz; op->eps = eps;
      std::vector<float2> ones(d->NW, make_float2(1.f, 0.f));
      upload_bytes(op->sref, ones.data(), ones.size() * 8);
    } catch (...) { delete op; throw; }
    *out = op;
  });
}
int fwix_phase_shift_set_slow(fwix_op op, const fwix_c64* sref) {
  return guard([&] {
    auto* p = as<PhaseShiftOp>(op, KIND_PHASE_SHIFT);
    if (!sref) throw Error("null slowness");
    hip_check(hipMemcpy(p->sref.p, sref, (size_t)p->t.dims.NW * 8, hipMemcpyHostToDevice), "set_slow");
  });
}
int fwix_phase_shift_forward(fwix_op op, int add, const fwix_c64* m, fwix_c64* d, fwix_stream s) {
  return guard([&] { as<PhaseShiftOp>(op, KIND_PHASE_SHIFT)->run(0, add, (const float2*)m, (float2*)d, S(s)); });
}
int fwix_phase_shift_adjoint(fwix_op op, int add, fwix_c64* m, const fwix_c64* d, fwix_stream s) {
  return guard([&] { as<PhaseShiftOp>(op, KIND_PHASE_SHIFT)->run(1, add, (const float2*)d, (float2*)m, S(s)); });
}
int fwix_phase_shift_inverse(fwix_op op, int add, fwix_c64* m, const fwix_c64* d, fwix_stream s) {
  return guard([&] { as<PhaseShiftOp>(op, KIND_PHASE_SHIFT)->run(2, add, (const float2*)d, (float2*)m, S(s)); });
}

// ---------------------------------------------------------------- Selector
int fwix_selector_create(const fwix_dims* d, fwix_op* out) {
  return guard([&] {
    validate_dims(d, false);
    if (!out) throw Error("null output handle");
    auto* op = new SelectorOp();
    try {
      op->t.init(*d, 0, 0);
      op->labels.alloc((size_t)d->NX * d->NY * d->NW);
    } catch (...) { delete op; throw; }
    *out = op;
  });
}
int fwix_selector_set_labels(fwix_op op, const int32_t* labels) {
  return guard([&] {
    auto* p = as<SelectorOp>(op, KIND_SELECTOR);
    if (!labels) throw Error("null labels");
    const fwix_dims& d = p->t.dims;
    const size_t n = (size_t)d.NX * d.NY * d.NW;
    std::vector<uint8_t> h(n);
    for (size_t i = 0; i < n; ++i) {
      if (labels[i] < 0 || labels[i] > 254) throw Error("labels must be in [0, 254]");
      h[i] = (uint8_t)labels[i];
    }
    hip_check(hipMemcpy(p->labels.p, h.data(), n, hipMemcpyHostToDevice), "set_labels");
    p->has_labels = true;
  });
}
int fwix_selector_set_value(fwix_op op, int value) {
  return guard([&] { as<SelectorOp>(op, KIND_SELECTOR)->value = value; });
}
int fwix_selector_forward(fwix_op op, int add, const fwix_c64* m, fwix_c64* d, fwix_stream s) {
  return guard([&] { as<SelectorOp>(op, KIND_SELECTOR)->run(add, (const float2*)m, (float2*)d, S(s)); });
}
int fwix_selector_adjoint(fwix_op op, int add, fwix_c64* m, const fwix_c64* d, fwix_stream s) {
  return guard([&] { as<SelectorOp>(op, KIND_SELECTOR)->run(add, (const float2*)d, (float2*)m, S(s)); });
}

// ---------------------------------------------------------------- Taper
int fwix_taper_create(const fwix_dims* d, int tx, int ty, fwix_op* out) {
  return guard([&] {
    validate_dims(d, false);
    if (!out) throw Error("null output handle");
    if (tx < 0 || ty < 0) throw Error("taper widths must be >= 0");
    auto* op = new TaperOp();
    try { op->t.init(*d, tx, ty); } catch (...) { delete op; throw; }
    *out = op;
  });
}
int fwix_taper_forward(fwix_op op, int add, const fwix_c64* m, fwix_c64* d, fwix_stream s) {
  return guard([&] { as<TaperOp>(op, KIND_TAPER)->run(add, (const float2*)m, (float2*)d, S(s)); });
}
int fwix_taper_adjoint(fwix_op op, int add, fwix_c64* m, const fwix_c64* d, fwix_stream s) {
  return guard([&] { as<TaperOp>(op, KIND_TAPER)->run(add, (const float2*)d, (float2*)m, S(s)); });
}

// ---------------------------------------------------------------- OneStep
int fwix_onestep_create(int method, const fwix_dims* d, const fwix_params* p, fwix_op* out) {
  return guard([&] {
    validate_dims(d, true);
    validate_params(p);
    if (method != FWIX_PSPI && method != FWIX_NSPS && method != FWIX_PSPI_PLAIN)
      throw Error("method must be FWIX_PSPI, FWIX_NSPS or FWIX_PSPI_PLAIN");
    if (d->nz < 1) throw Error("dims: nz must be >= 1");
    if (!out) throw Error("null output handle");
    auto* op = new OneStepOp();
    try { op->eng.init(method, *d, *p); } catch (...) { delete op; throw; }
    *out = op;
  });
}
int fwix_onestep_set_refs(fwix_op op, const fwix_c64* sref, const uint8_t* labels) {
  return guard([&] {
    auto* p = as<OneStepOp>(op, KIND_ONESTEP);
    p->eng.set_refs((const float2*)sref, labels);
  });
}
int fwix_onestep_set_shot_planes(fwix_op op, int nshot, const int* nplanes) {
  return guard([&] { as<OneStepOp>(op, KIND_ONESTEP)->eng.set_model_map(nshot, nplanes); });
}
int fwix_onestep_set_frequencies(fwix_op op, const float* f_hz) {
  return guard([&] { as<OneStepOp>(op, KIND_ONESTEP)->eng.tables.set_frequencies(f_hz, false); });
}
int fwix_onestep_set_depth(fwix_op op, int iz) {
  return guard([&] {
    auto* p = as<OneStepOp>(op, KIND_ONESTEP);
    if (iz < 0 || iz >= p->eng.tables.dims.nz) throw Error("depth index out of range");
    p->iz = iz;
  });
}
int fwix_onestep_forward(fwix_op op, int add, const fwix_c64* m, fwix_c64* d, fwix_stream s) {
  return guard([&] {
    auto* p = as<OneStepOp>(op, KIND_ONESTEP);
    p->eng.apply(false, p->iz, add, (const float2*)m, (float2*)d, S(s));
  });
}
int fwix_onestep_adjoint(fwix_op op, int add, fwix_c64* m, const fwix_c64* d, fwix_stream s) {
  return guard([&] {
    auto* p = as<OneStepOp>(op, KIND_ONESTEP);
    p->eng.apply(true, p->iz, add, (const float2*)d, (float2*)m, S(s));
  });
}

// ---------------------------------------------------------------- timing
int fwix_timing_enable(int on) {
  return guard([&] {
    std::lock_guard<std::mutex> g(g_tmu);
    if (on > 0)  // pre-create event pairs so the timed region does not allocate them
      while ((int)g_pool.size() < 2 * on) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        g_pool.push_back(e);
      }
    g_timing = on > 0;
  });
}
int fwix_tune_columns_set(int kind, int mode, int n, int waves_per_simd) {
  return guard([&] {
    if (col_wpe_set(kind, mode, n, waves_per_simd) != 0)
      throw Error("tune_columns: kind must be 0/1, mode 0..2, n a supported size, waves per SIMD 2, 4 or 6");
  });
}
int fwix_tune_columns_get(int kind, int mode, int n) { return col_wpe_get(kind, mode, n); }
int fwix_tune_columns_layout_set(int kind, int mode, int n, int layout) {
  return guard([&] {
    if (col_layout_set(kind, mode, n, layout) != 0)
      throw Error("tune_columns_layout: kind must be 0/1, mode 0..2, n a supported size, layout 0 (or 1 for n = 1024)");
  });
}
int fwix_tune_columns_layout_get(int kind, int mode, int n) { return col_layout_get(kind, mode, n); }
int fwix_tune_scratch_layout_set(int n, int quad) {
  return guard([&] {
    if (scratch_quad_set(n, quad) != 0)
      throw Error("tune_scratch_layout: n must be a supported size and quad 0 or 1");
  });
}
int fwix_tune_scratch_layout_get(int n) { return scratch_quad_get(n); }
int fwix_copy(const void* in, void* out, size_t bytes, fwix_stream s) {
  return guard([&] {
    if (bytes % 16) throw Error("fwix_copy: bytes must be a multiple of 16");
    if ((!in || !out) && bytes) throw Error("null buffer");
    if (((uintptr_t)in | (uintptr_t)out) % 16) throw Error("fwix_copy: buffers must be 16-byte aligned");
    hip_check(launch_copy16(in, out, bytes / 16, S(s)), "copy");
  });
}
int fwix_timing_select(const char* pass, int every) {
  return guard([&] {
    if (every < 1) throw Error("every must be >= 1");
    std::lock_guard<std::mutex> g(g_tmu);
    g_only = pass ? pass : "";
    g_every = every;
  });
}
int fwix_timing_overhead(fwix_stream s, int pairs, float* us) {
  return guard([&] {
    if (pairs < 1 || !us) throw Error("pairs must be >= 1 and us non-null");
    hipStream_t st = static_cast<hipStream_t>(s);
    std::vector<hipEvent_t> ev(2 * (size_t)pairs);
    for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    for (int i = 0; i < pairs; ++i) {
      hip_check(hipEventRecord(ev[2 * i], st), "hipEventRecord");
      hip_check(hipEventRecord(ev[2 * i + 1], st), "hipEventRecord");
    }
    double tot = 0;
    for (int i = 0; i < pairs; ++i) {
      hip_check(hipEventSynchronize(ev[2 * i + 1]), "hipEventSynchronize");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]), "hipEventElapsedTime");
      tot += ms;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *us = (float)(tot * 1e3 / pairs);
  });
}
int fwix_timing_collect(void) {
  return guard([&] {
    std::lock_guard<std::mutex> g(g_tmu);
    for (auto& r : g_pending) {
      hip_check(hipEventSynchronize(r.b), "hipEventSynchronize");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
      PassStat& ps = g_stat[r.name];
      ps.ms += ms;
      ps.bytes += r.bytes;
      ps.n += 1;
      g_pool.push_back(r.a);
      g_pool.push_back(r.b);
    }
    g_pending.clear();
  });
}
int fwix_timing_reset(void) {
  return guard([&] {
    std::lock_guard<std::mutex> g(g_tmu);
    g_stat.clear();
    g_seen.clear();
  });
}
int fwix_timing_report(char* buf, int len) {
  return guard([&] {
    std::lock_guard<std::mutex> g(g_tmu);
    std::string out = "{";
    bool first = true;
    for (auto& kv : g_stat) {
      char tmp[320];
      auto it = g_seen.find(kv.first);
      const long long total = it == g_seen.end() ? kv.second.n : it->second;
      std::snprintf(tmp, sizeof(tmp),
                    "%s\"%s\": {\"ms\": %.6f, \"launches\": %lld, \"bytes\": %.0f, \"total_launches\": %lld}",
                    first ? "" : ", ", kv.first.c_str(), kv.second.ms, kv.second.n, kv.second.bytes, total);
      out += tmp;
      first = false;
    }
    out += "}";
    if (!buf || len < (int)out.size() + 1) throw Error("timing report buffer too small");
    std::memcpy(buf, out.c_str(), out.size() + 1);
  });
}

}  // extern "C"
