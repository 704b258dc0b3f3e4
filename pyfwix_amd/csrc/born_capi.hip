// C ABI of the operators around the step: Injection, Scatter,
// ImagingCondition, Downward / Upward recursions and the Born operator
// (background modeling, linearized forward, adjoint, migration).
// Declarations and reference citations: include/fwix_hip.h.
#include "fwix_hip.h"
#include "fwix_launch.h"
#include "capi_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

using namespace fwix;

namespace fwix {

template <class T>
static void upload_vec(DevBuf& b, const std::vector<T>& v) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) hip_check(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

// ------------------------------------------------------------------ injection plan
void InjPlan::build(const fwix_dims& d, int nlev_, int ntrace_, const float* cx, const float* cy, const float* cz,
                    const int* ids, const std::vector<int>* first, const std::vector<int>* count) {
  if (ntrace_ < 0) throw Error("ntrace must be >= 0");
  if (ntrace_ > 0 && (!cx || !cy || !cz || !ids)) throw Error("null coordinates");
  ntrace = ntrace_;
  nlev = nlev_;
  NW = d.NW;
  NX = d.NX;
  plane = (long long)d.NX * d.NY;
  const int nshot = first ? (int)first->size() : d.NS;
  auto p_first = [&](int s) { return first ? (long long)(*first)[s] : (long long)s * d.NW; };
  auto p_count = [&](int s) { return count ? (*count)[s] : d.NW; };
  struct Ent { int lev; long long node; int trace; float w; };
  struct Act { int lev; int trace; long long base; float4 w; };
  std::vector<Ent> ents;
  std::vector<Act> acts;
  std::vector<long long> to(ntrace + 1, 0);
  wmax = 0;
  for (int t = 0; t < ntrace; ++t) {
    if (ids[t] < 0 || ids[t] >= nshot) throw Error("trace shot id out of range [0, NS)");
    to[t + 1] = to[t] + p_count(ids[t]);
    wmax = std::max(wmax, p_count(ids[t]));
    // propagator/src/injection.cu:38-48, float arithmetic as in the kernel
    const int iy = (int)floorf((cy[t] - d.oy) / d.dy);
    const float y = d.oy + iy * d.dy;
    const float ly = 1.f - (cy[t] - y) / d.dy;
    const int ix = (int)floorf((cx[t] - d.ox) / d.dx);
    const float x = d.ox + ix * d.dx;
    const float lx = 1.f - (cx[t] - x) / d.dx;
    const int iz = (int)floorf((cz[t] - d.oz) / d.dz);
    const float z = d.oz + iz * d.dz;
    const float lz = 1.f - (cz[t] - z) / d.dz;
    if (ix < 0 || ix >= d.NX - 1 || iy < 0 || iy >= d.NY - 1) continue;  // injection.cu:50-52
    for (int k = 0; k < 2; ++k) {
      const int lev = iz + k;
      if (lev < 0 || lev >= nlev) continue;
      const float f = k == 0 ? lz : 1 - lz;  // injection.cu:60-75
      float4 w;
      w.x = f * lx * ly;
      w.y = f * (1 - lx) * ly;
      w.z = f * lx * (1 - ly);
      w.w = f * (1 - lx) * (1 - ly);
      // first plane of the trace's shot: its nodes never meet another shot's
      const long long base = p_first(ids[t]) * plane + (long long)iy * d.NX + ix;
      acts.push_back({lev, t, base, w});
      ents.push_back({lev, base, t, w.x});
      ents.push_back({lev, base + 1, t, w.y});
      ents.push_back({lev, base + d.NX, t, w.z});
      ents.push_back({lev, base + d.NX + 1, t, w.w});
    }
  }
  nsamp = to[ntrace];
  std::stable_sort(acts.begin(), acts.end(), [](const Act& a, const Act& b) { return a.lev < b.lev; });
  std::stable_sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
    if (a.lev != b.lev) return a.lev < b.lev;
    if (a.node != b.node) return a.node < b.node;
    return a.trace < b.trace;
  });
  act_lvl.assign(nlev + 1, 0);
  seg_lvl.assign(nlev + 1, 0);
  std::vector<int> at(acts.size()), an(acts.size());
  std::vector<long long> ab(acts.size());
  std::vector<float4> aw(acts.size());
  for (size_t k = 0; k < acts.size(); ++k) {
    at[k] = acts[k].trace;
    ab[k] = acts[k].base;
    aw[k] = acts[k].w;
    an[k] = p_count(ids[acts[k].trace]);
    act_lvl[acts[k].lev + 1]++;
  }
  std::vector<long long> sb;
  std::vector<int> so, sn, et;
  std::vector<float> ew;
  for (size_t k = 0; k < ents.size(); ++k) {
    if (k == 0 || ents[k].lev != ents[k - 1].lev || ents[k].node != ents[k - 1].node) {
      sb.push_back(ents[k].node);
      so.push_back((int)et.size());
      sn.push_back(p_count(ids[ents[k].trace]));
      seg_lvl[ents[k].lev + 1]++;
    }
    et.push_back(ents[k].trace);
    ew.push_back(ents[k].w);
  }
  so.push_back((int)et.size());
  for (int l = 0; l < nlev; ++l) {
    act_lvl[l + 1] += act_lvl[l];
    seg_lvl[l + 1] += seg_lvl[l];
  }
  upload_vec(act_trace, at);
  upload_vec(act_base, ab);
  upload_vec(act_w, aw);
  upload_vec(act_np, an);
  upload_vec(seg_base, sb);
  upload_vec(seg_off, so);
  upload_vec(seg_np, sn);
  upload_vec(ent_trace, et);
  upload_vec(ent_w, ew);
  upload_vec(toff, to);
}

void InjPlan::forward(int iz, const float2* traces, float2* wfld, hipStream_t st) const {
  if (iz < 0 || iz >= nlev) return;
  hip_check(launch_inj_fwd(traces, wfld, (const long long*)seg_base.p, (const int*)seg_off.p, (const int*)seg_np.p,
                           (const int*)ent_trace.p, (const float*)ent_w.p, (const long long*)toff.p, seg_lvl[iz],
                           seg_lvl[iz + 1] - seg_lvl[iz], wmax, plane, st),
            "injection forward");
}

void InjPlan::adjoint(int iz, float2* traces, const float2* wfld, hipStream_t st) const {
  if (iz < 0 || iz >= nlev) return;
  hip_check(launch_inj_adj(traces, wfld, (const int*)act_trace.p, (const long long*)act_base.p,
                           (const float4*)act_w.p, (const int*)act_np.p, (const long long*)toff.p, act_lvl[iz],
                           act_lvl[iz + 1] - act_lvl[iz], wmax, NX, plane, st),
            "injection adjoint");
}

}  // namespace fwix

namespace {

inline hipStream_t S(fwix_stream s) { return (hipStream_t)s; }

template <class T>
T* as(fwix_op op, int kind) {
  if (!op) throw Error("null operator handle");
  if (op->kind != kind) throw Error("operator handle of the wrong kind");
  return static_cast<T*>(op);
}

size_t wsize(const fwix_dims& d) { return (size_t)d.NX * d.NY * d.NW * d.NS; }

// ------------------------------------------------------------------ Injection
struct InjectionOp : fwix_op_s {
  fwix_dims d{};
  int ntrace = 0, iz = 0;
  InjPlan plan;
  bool has_coords = false;
  InjectionOp() { kind = KIND_INJECTION; }
};

// ------------------------------------------------------------------ Scatter
struct ScatterOp : fwix_op_s {
  StepEngine eng;  // tables + K/W scratch (labels unused)
  const float2* slow = nullptr;
  int iz = 0;
  ScatterOp() { kind = KIND_SCATTER; }
  void run(bool adjoint, int add, const float2* in, float2* out, hipStream_t st) {
    if (!slow) throw Error("scatter slowness not set (call set_slow)");
    const fwix_dims& d = eng.tables.dims;
    const size_t plane = (size_t)d.NX * d.NY;
    const int planes = d.NW * d.NS;
    const float sx = (float)(1.0 / std::sqrt((double)d.NX));
    const float sy2 = (float)(1.0 / (double)d.NY);
    for (int c0 = 0; c0 < planes; c0 += eng.chunk) {
      const int np = std::min(eng.chunk, planes - c0);
      KArgs a = eng.tables.kargs();
      a.K = (float2*)eng.K.p;
      a.W = (float2*)eng.W.p;
      a.p0 = c0; a.np = np;
      a.wstride = (long long)np * plane;
      a.nterm = eng.params.taylor + 1;
      a.eps_s = eng.params.eps_scat;
      a.slow = slow + (size_t)iz * d.NW * d.ny * d.nx;
      const float2* cin = in + (size_t)c0 * plane;
      float2* cout = out + (size_t)c0 * plane;
      const int T = eng.params.taylor;
      const double pb = 8.0 * np * plane;
      if (!adjoint) {  // Scatter.cpp:90-102
        if (T >= 1) {
          a.in = cin; a.out = a.K; a.qout = 1; a.add = 0; a.scale = sx;  // K in the quad-block layout
          hip_check(launch_rows(0, a, st), "scatter rows");
          a.qout = 0; a.nterm = T; a.scale = sy2;
          PassTimer t("scatter_cols", (1 + T) * pb, st);
          hip_check(launch_cols_fanout(2, a, st), "scatter columns");
        }
        a.in = cin; a.out = cout; a.add = add; a.nterm = T + 1; a.scale = sx;
        hip_check(launch_rows_scatter(false, a, st), "scatter combine");
      } else {  // Scatter.cpp:104-116
        a.in = cin; a.nterm = T + 1; a.scale = sx;
        if (T >= 1) {
          hip_check(launch_rows_scatter_adj(false, a, st), "scatter adjoint rows");
          a.nterm = T; a.scale = sy2;
          PassTimer t("scatter_cols_adj", (1 + T) * pb, st);
          hip_check(launch_cols_fanin(2, a, st), "scatter adjoint columns");
        }
        a.in = cin; a.out = cout; a.add = add; a.nterm = T + 1; a.scale = sx;
        hip_check(launch_rows_adj_out(false, a, st), "scatter adjoint combine");
      }
    }
  }
};

// ------------------------------------------------------------------ ImagingCondition
struct ICOp : fwix_op_s {
  fwix_dims d{};
  const float2* bg = nullptr;
  ICOp() { kind = KIND_IC; }
};

// ------------------------------------------------------------------ Downward / Upward
struct OneWayOp : fwix_op_s {
  int direction = 0;  // 0 down, 1 up
  StepEngine eng;
  DevBuf u;
  OneWayOp() { kind = KIND_ONEWAY; }
  // d (+)= P_{last} ... P_{first} m  (OneWay.cpp:22-79)
  void run(bool adjoint, int add, const float2* in, float2* out, float2* ckpt, hipStream_t st) {
    const fwix_dims& d = eng.tables.dims;
    const size_t n = wsize(d);
    float2* w = (float2*)u.p;
    hip_check(launch_axpy(in, w, n, 1.f, 0, st), "copy");
    const int nz = d.nz;
    // Downward forward and Upward adjoint step z = 0..nz-1; the others z = nz-1..0
    const bool ascending = (direction == 0) != adjoint;
    // consecutive steps chain through the row spectrum (PSPI forward, NSPS
    // adjoint); w itself is written where a checkpoint or the result reads it
    bool kr = false;
    for (int k = 0; k < nz; ++k) {
      const int z = ascending ? k : nz - 1 - k;
      if (ckpt) hip_check(launch_axpy(w, ckpt + (size_t)z * n, n, 1.f, 0, st), "checkpoint");
      const bool kn = k + 1 < nz;
      eng.apply_chained(z, w, w, kr, kn, st, ckpt != nullptr || !kn, adjoint);
      kr = kn && eng.chains(adjoint);
    }
    hip_check(launch_axpy(w, out, n, 1.f, add, st), "accumulate");
  }
};

// ------------------------------------------------------------------ Born
struct BornOp : fwix_op_s {
  StepEngine eng;
  fwix_dims d{};
  fwix_params par{};
  const float2* slow = nullptr;     // (nz, NW, ny, nx) background slowness
  const float2* wavelet = nullptr;  // (nsrc, NW)
  InjPlan src, rec;
  int nsrc = 0, nrec = 0;
  int ckpt = 1;
  DevBuf store, seg, U, dU, lam;
  int codec = 0;   // checkpoints: 0 lossless complex64, 1 16-bit block floating point
  int f64 = 0;     // scattered-field transforms in float64 arithmetic (fwix_born_set_precision)
  DevBuf scales;  // codec 1: one float per 64 samples of every checkpoint
  DevBuf ubuf;    // codec 1, every depth kept: the decoded U_z
  DevBuf vstore;  // up-going background V_z of every depth (model_up)
  DevBuf xstore;  // dU_z (extended forward) or mu_z (extended adjoint) of every depth
  const float2* den = nullptr;  // (nz, NW, ny, nx) background density
  int seg_first = -1;  // first depth of the recomputed segment held in seg
  bool have_bg = false, have_up = false;
  long long steps = 0;  // depth steps executed (statistics)
  // shot batch: shot k owns planes [sfirst[k], sfirst[k] + scount[k]); the
  // slowness, references, labels and model are per shot, traces ragged
  std::vector<int> sfirst, scount;
  DevBuf gscr, sfirst_dev, scount_dev;  // per-plane gradient of one depth; plane ranges
  BornOp() { kind = KIND_BORN; }
  bool batched() const { return !sfirst.empty(); }
  // Reference-granularity form (FWIX_PSPI_PLAIN): the imaging condition, the
  // scattering operator and the PSPI step as separate operators, one shot per
  // launch, as Scattering.cpp:43-61 chains them -- the comparator of the fused
  // steps.  pa / pb: whole-wavefield scratch; pw: the scattering terms of one shot.
  DevBuf pa, pb, pw;
  void set_method(int method) {
    if (method != FWIX_PSPI && method != FWIX_PSPI_PLAIN) throw Error("method must be FWIX_PSPI or FWIX_PSPI_PLAIN");
    if (batched()) throw Error("shot batches run on the fused path only");
    const float2* sr = eng.sref;
    const uint8_t* lb = eng.labels;
    eng.init(method, d, par, par.taylor + 1);
    if (sr && lb) eng.set_refs(sr, lb);
    const size_t n = P() * sizeof(float2);
    if (eng.plain) {
      pa.alloc(n);
      pb.alloc(n);
      pw.alloc((size_t)std::max(par.taylor, 1) * d.NW * d.NX * d.NY * sizeof(float2));
    } else {
      pa.release();
      pb.release();
      pw.release();
    }
    have_bg = have_up = false;
  }
  // one shot's scattering operator (Scatter.cpp:90-116), forward or adjoint
  void scatter_shot(bool adjoint, int z, int is, const float2* in, float2* out, hipStream_t st) {
    const size_t plane = (size_t)d.NX * d.NY;
    const float sx = (float)(1.0 / std::sqrt((double)d.NX)), sy2 = (float)(1.0 / (double)d.NY);
    const double sb = 8.0 * d.NW * plane;
    KArgs a = eng.tables.kargs();
    a.K = (float2*)eng.K.p;
    a.W = (float2*)pw.p;
    a.p0 = is * d.NW; a.np = d.NW;
    a.wstride = (long long)d.NW * plane;
    a.eps_s = par.eps_scat;
    a.slow = slow + (size_t)z * MZ();
    const int T = par.taylor;
    if (!adjoint) {
      if (T >= 1) {
        a.in = in; a.out = a.K; a.qout = 1; a.add = 0; a.scale = sx;  // K in the quad-block layout
        { PassTimer t("plain_scatter_rows", 2 * sb, st); hip_check(launch_rows(0, a, st), "plain scatter rows"); }
        a.qout = 0; a.nterm = T; a.scale = sy2;
        { PassTimer t("plain_scatter_cols", (1 + T) * sb, st); hip_check(launch_cols_fanout(2, a, st), "plain scatter columns"); }
      }
      a.in = in; a.out = out; a.add = 0; a.nterm = T + 1; a.scale = sx;
      { PassTimer t("plain_scatter_combine", (T + 2) * sb, st); hip_check(launch_rows_scatter(false, a, st), "plain scatter combine"); }
    } else {
      a.in = in; a.nterm = T + 1; a.scale = sx;
      if (T >= 1) {
        { PassTimer t("plain_scatter_adj_rows", (1 + T) * sb, st); hip_check(launch_rows_scatter_adj(false, a, st), "plain scatter adjoint rows"); }
        a.nterm = T; a.scale = sy2;
        { PassTimer t("plain_scatter_adj_cols", (1 + T) * sb, st); hip_check(launch_cols_fanin(2, a, st), "plain scatter adjoint columns"); }
      }
      a.in = in; a.out = out; a.add = 0; a.nterm = T + 1; a.scale = sx;
      { PassTimer t("plain_scatter_adj_combine", 3 * sb, st); hip_check(launch_rows_adj_out(false, a, st), "plain scatter adjoint combine"); }
    }
  }
  // out = P_z[ addin + S_z(bg * ds_z) ] through separate operators
  void step_fwd_plain(int z, const float2* bg, const float2* dsz, const float2* addin, float2* out, hipStream_t st) {
    const size_t n = P(), shot = (size_t)d.NW * d.NX * d.NY;
    float2* a_ = (float2*)pa.p;
    float2* b_ = (float2*)pb.p;
    { PassTimer t("plain_ic", 24.0 * n, st);
      hip_check(launch_ic_fwd(dsz, a_, bg, d.NX, d.NY, d.NW, d.NS, d.nx, d.ny, 0, st), "plain imaging condition"); }
    for (int is = 0; is < d.NS; ++is) scatter_shot(false, z, is, a_ + is * shot, b_ + is * shot, st);
    { PassTimer t("plain_add", 24.0 * n, st); hip_check(launch_axpy(addin, b_, n, 1.f, 1, st), "plain add"); }
    eng.apply(false, z, 0, b_, out, st);
    ++steps;
  }
  // lam <- P_z^H lam; gz += conj(bg) S_z^H lam through separate operators
  void step_adj_plain(int z, float2* lam, const float2* bg, float2* gz, hipStream_t st) {
    const size_t n = P(), shot = (size_t)d.NW * d.NX * d.NY;
    float2* a_ = (float2*)pa.p;
    eng.apply(true, z, 0, lam, lam, st);
    for (int is = 0; is < d.NS; ++is) scatter_shot(true, z, is, lam + is * shot, a_ + is * shot, st);
    { PassTimer t("plain_ic_adj", 16.0 * n + 16.0 * MZ(), st);
      hip_check(launch_ic_adj(gz, a_, bg, d.NX, d.NY, d.NW, d.NS, d.nx, d.ny, 1, st), "plain imaging adjoint"); }
    ++steps;
  }
  void no_batch(const char* what) const {
    if (batched()) throw Error(std::string(what) + " is not available for a shot batch");
  }
  void set_shot_planes(int nshot, const int* nplanes) {
    eng.set_model_map(nshot, nplanes);
    sfirst.assign(nshot, 0);
    scount.assign(nplanes, nplanes + nshot);
    for (int k = 1; k < nshot; ++k) sfirst[k] = sfirst[k - 1] + scount[k - 1];
    upload_vec(sfirst_dev, sfirst);
    upload_vec(scount_dev, scount);
    gscr.alloc((size_t)d.NW * d.ny * d.nx * sizeof(float2));
    slow = nullptr;  // model-sized inputs change shape: set them again
    wavelet = nullptr;
    nsrc = nrec = 0;
    build_plan(src, 0, nullptr, nullptr, nullptr, nullptr);
    build_plan(rec, 0, nullptr, nullptr, nullptr, nullptr);
    have_bg = have_up = false;
  }
  void build_plan(InjPlan& pl, int n, const float* x, const float* y, const float* z, const int* ids) {
    if (batched()) pl.build(d, d.nz, n, x, y, z, ids, &sfirst, &scount);
    else pl.build(d, d.nz, n, x, y, z, ids);
  }

  size_t P() const { return wsize(d); }
  float2* bufU() { return (float2*)U.p; }

  size_t nblk() const { return (P() + 63) / 64; }
  size_t nck() const { return (size_t)((d.nz + ckpt - 1) / ckpt); }
  // checkpoint k into dst (decoded when the codec is on)
  void restore(int k, float2* dst, hipStream_t st) {
    const size_t n = P();
    if (codec)
      hip_check(launch_bfp16_unpack((const char*)store.p + (size_t)k * n * 4, (const float*)scales.p + (size_t)k * nblk(),
                                    dst, n, st), "restore");
    else
      hip_check(launch_axpy((const float2*)store.p + (size_t)k * n, dst, n, 1.f, 0, st), "restore");
  }
  // (re)allocate the checkpoint store for the codec
  void set_codec(int c) {
    if (c != 0 && c != 1) throw Error("codec must be 0 (lossless) or 1 (16-bit block floating point)");
    codec = c;
    const size_t n = P();
    store.alloc(n * (c ? 4 : 8) * nck());
    if (c) {
      scales.alloc(nblk() * nck() * sizeof(float));
      if (ckpt == 1) ubuf.alloc(n * sizeof(float2));
    } else {
      scales.release();
      ubuf.release();
    }
    have_bg = false;
    have_up = false;
    seg_first = -1;
  }

  // wavefield U_z of the background down-going recursion (Propagator.cpp:155-180)
  const float2* background(int z, hipStream_t st) {
    if (!have_bg) throw Error("background wavefield not modeled (call fwix_born_model)");
    const size_t n = P();
    if (ckpt == 1) {
      if (!codec) return (const float2*)store.p + (size_t)z * n;
      restore(z, (float2*)ubuf.p, st);
      return (const float2*)ubuf.p;
    }
    const int s0 = (z / ckpt) * ckpt;
    if (seg_first != s0) {
      // recompute U_{s0..s1-1} from the checkpoint at s0
      const int s1 = std::min(s0 + ckpt, d.nz);
      float2* sg = (float2*)seg.p;
      restore(s0 / ckpt, sg, st);
      bool kr = false;  // the row spectrum of prev is already in the engine (chained steps)
      for (int zz = s0 + 1; zz < s1; ++zz) {
        float2* prev = sg + (size_t)(zz - 1 - s0) * n;
        float2* cur = sg + (size_t)(zz - s0) * n;
        const bool kn = zz + 1 < s1 && !src.injects(zz);  // cur goes on unchanged into the next step
        eng.apply_chained(zz - 1, prev, cur, kr, kn, st);
        kr = kn && eng.chains();
        src.forward(zz, wavelet, cur, st);
        ++steps;
      }
      seg_first = s0;
    }
    return (const float2*)seg.p + (size_t)(z - s0) * n;
  }

  void model(float2* data, int add, hipStream_t st) {
    if (!wavelet) throw Error("wavelet not set (call fwix_born_set_wavelet)");
    if (!eng.sref) throw Error("reference slownesses / labels not set");
    const size_t n = P();
    if (data && !add) hip_check(launch_zero(data, (size_t)rec.nsamp, st), "zero data");
    // Lossless checkpoints of every depth: the recursion runs in the store
    // itself, each step writing U_{z+1} into its slot (no copies).
    const bool direct = ckpt == 1 && !codec;
    float2* u = direct ? (float2*)store.p : bufU();
    hip_check(launch_zero(u, n, st), "zero wavefield");
    // consecutive steps chain through the row spectrum while no source
    // injects in between (the receivers and checkpoints only read U)
    bool kr = false;
    auto step = [&](int z, const float2* in, float2* out) {
      const bool kn = z + 2 < d.nz && !src.injects(z + 1);
      // U_{z+1} itself is read only where it is stored or recorded (the next
      // chained step starts from its spectrum)
      const bool need = direct || (z + 1) % ckpt == 0 || (data && rec.records(z + 1));
      eng.apply_chained(z, in, out, kr, kn, st, need);
      kr = kn && eng.chains();
      ++steps;
    };
    for (int z = 0; z < d.nz; ++z) {
      src.forward(z, wavelet, u, st);
      if (data) rec.adjoint(z, data, u, st);
      if (direct) {
        if (z + 1 < d.nz) {
          step(z, u, u + n);
          u += n;
        }
        continue;
      }
      if (z % ckpt == 0) {
        PassTimer t("checkpoint_store", (codec ? 12.0 + 4.0 / 64 : 16.0) * n, st);
        const size_t k = (size_t)(z / ckpt);
        if (codec)
          hip_check(launch_bfp16_pack(u, (char*)store.p + k * n * 4, (float*)scales.p + k * nblk(), n, st), "store");
        else
          hip_check(launch_axpy(u, (float2*)store.p + k * n, n, 1.f, 0, st), "store");
      }
      if (z + 1 < d.nz) step(z, u, u);
    }
    seg_first = -1;
    have_bg = true;
  }

  KArgs born_args(int z, int p0, int np) const {
    KArgs a = eng.base(z, p0, np);
    a.nterm = par.taylor + 1;
    a.slow = slow + (size_t)z * MZ();
    a.eps_s = par.eps_scat;
    return a;
  }

  size_t MZ() const { return (size_t)eng.nmodel * d.ny * d.nx; }

  // out = P_z[ addin + S_z(bg * ds_z) ] over every plane (Scattering.cpp:43-52,
  // fused): addin and out may be the same wavefield.  The scattering operator
  // and the PSPI step use the slowness and references of depth z.
  void step_fwd(int z, const float2* bg, const float2* dsz, const float2* addin, float2* out, hipStream_t st) {
    if (eng.plain) return step_fwd_plain(z, bg, dsz, addin, out, st);
    const size_t plane = (size_t)d.NX * d.NY;
    const float sx = (float)(1.0 / std::sqrt((double)d.NX)), sy2 = (float)(1.0 / (double)d.NY);
    const int cw = std::min(eng.chunk, d.NW);
    for (int is = 0; is < d.NS; ++is)
      for (int w0 = 0; w0 < d.NW; w0 += cw) {
        const int p0 = is * d.NW + w0, np = std::min(cw, d.NW - w0);
        KArgs a = born_args(z, p0, np);
        const double pb = 8.0 * np * plane, mb = 8.0 * np * d.nx * d.ny;
        const double wb = 8.0 * d.NX * eng.present_rows(z, p0, np);
        const int T = par.taylor;
        a.dslow = dsz;
        a.bg = bg + (size_t)p0 * plane;
        a.wstride = (long long)np * plane;
        // with f64, the passes that carry the amplified scattering terms in
        // space or through a transform run in float64 arithmetic
        if (T >= 1) {
          a.scale = sx;
          { PassTimer t("born_rows_ic", 2 * pb + mb, st); hip_check(launch_rows_ic_fwd(a, st), "born ic"); }
          a.nterm = T; a.scale = sy2; a.f64 = f64;
          { PassTimer t("born_cols_kfac", (1 + T) * pb, st); hip_check(launch_cols_fanout(2, a, st), "born scatter columns"); }
        }
        a.addin = addin + (size_t)p0 * plane;
        a.nterm = T + 1; a.scale = sx; a.f64 = f64;
        { PassTimer t("born_rows_scatter", (T + 3) * pb + 2 * mb, st); hip_check(launch_rows_scatter(true, a, st), "born scatter rows"); }
        a.nterm = par.nref;
        a.scale = sy2;
        { PassTimer t("cols_fanout", pb + wb, st); hip_check(launch_cols_fanout(0, a, st), "born pspi columns"); }
        a.out = out + (size_t)p0 * plane;
        a.add = 0;
        a.scale = sx; a.f64 = 0;
        { PassTimer t("rows_select", wb + pb + pb / 8, st); hip_check(launch_rows_select(a, st), "born pspi rows"); }
      }
    ++steps;
  }

  // lam <- P_z^H lam in place, and gz += conj(bg) S_z^H (P_z^H lam) on the model
  // region (the adjoint of step_fwd's scattering term, Scattering.cpp:54-61)
  void step_adj(int z, float2* lam, const float2* bg, float2* gz, hipStream_t st) {
    if (eng.plain) return step_adj_plain(z, lam, bg, gz, st);
    const size_t plane = (size_t)d.NX * d.NY;
    const float sx = (float)(1.0 / std::sqrt((double)d.NX)), sy2 = (float)(1.0 / (double)d.NY);
    const int cw = std::min(eng.chunk, d.NW);
    for (int is = 0; is < d.NS; ++is)
      for (int w0 = 0; w0 < d.NW; w0 += cw) {
        const int p0 = is * d.NW + w0, np = std::min(cw, d.NW - w0);
        float2* lc = lam + (size_t)p0 * plane;
        KArgs a = born_args(z, p0, np);
        a.wstride = (long long)np * plane;
        a.f64 = f64;  // every pass of the adjoint step carries the amplified terms
        const double pb = 8.0 * np * plane, mb = 8.0 * np * d.nx * d.ny;
        const double wb = 8.0 * d.NX * eng.present_rows(z, p0, np);
        const int T = par.taylor;
        // mu = P_z^H lam (PSPI adjoint, written in place) fused with S_z^H's first row pass
        a.in = lc; a.scale = sx;
        { PassTimer t("rows_fanout", pb + pb / 8 + wb, st); hip_check(launch_rows_fanout(a, st), "born adjoint pspi rows"); }
        a.nterm = par.nref; a.scale = sy2;
        { PassTimer t("cols_fanin", wb + pb, st); hip_check(launch_cols_fanin(1, a, st), "born adjoint pspi columns"); }
        a.out = lc; a.nterm = T + 1; a.scale = sx;
        { PassTimer t("born_rows_inv_scatter_adj", (2 + T) * pb + mb, st); hip_check(launch_rows_scatter_adj(true, a, st), "born adjoint inverse + scatter rows"); }
        if (T >= 1) {
          a.nterm = T; a.scale = sy2;
          { PassTimer t("born_cols_kfac_adj", (1 + T) * pb, st); hip_check(launch_cols_fanin(2, a, st), "born adjoint scatter columns"); }
        }
        // gz += conj(bg) S_z^H mu (a shot batch: per plane, summed over each shot below)
        a.in = lc; a.bg = bg + (size_t)p0 * plane;
        a.aux = batched() ? (float2*)gscr.p : gz;
        a.add = batched() ? 0 : 1; a.nterm = T + 1; a.scale = sx;
        { PassTimer t("born_rows_adj_ic", 3 * pb + 3 * mb, st); hip_check(launch_rows_adj_out(true, a, st), "born adjoint imaging"); }
      }
    if (batched()) {
      PassTimer t("born_shot_sum", 8.0 * d.NW * d.ny * d.nx + 16.0 * MZ(), st);
      hip_check(launch_plane_sum((const float2*)gscr.p, gz, (const int*)sfirst_dev.p, (const int*)scount_dev.p,
                                 eng.nmodel, (long long)d.ny * d.nx, st), "shot gradient sum");
    }
    ++steps;
  }

  // dU_{z+1} = P_z [ dU_z + S_z(U_z * ds_z) ]   (ExtendedBorn.cpp:152-178); with
  // keep, dU_z of every depth is left in xstore (the down-reflected channel's input)
  void forward(int add, const float2* dslow, float2* data, hipStream_t st, bool keep = false) {
    if (!slow) throw Error("background slowness not set (call fwix_born_set_slow)");
    const size_t n = P();
    if (!add) hip_check(launch_zero(data, (size_t)rec.nsamp, st), "zero data");
    float2* du = keep ? (float2*)xstore.p : (float2*)dU.p;
    hip_check(launch_zero(du, n, st), "zero dU");
    for (int z = 0; z < d.nz; ++z) {
      if (z > 0) rec.adjoint(z, data, du, st);
      if (z + 1 == d.nz) break;  // dU_nz is never recorded
      float2* nxt = keep ? du + n : du;
      step_fwd(z, background(z, st), dslow + (size_t)z * MZ(), du, nxt, st);
      du = nxt;
    }
  }

  // adjoint (ExtendedBorn.cpp:180-208) or, with image != null, the migration
  // image I_z = sum_{w,s} conj(U_z) P_z^H lam_{z+1}.  With mu != null the
  // down-reflected channel's adjoint source conj(r_z) mu_z joins Inj_z(data).
  void adjoint(int add, float2* dslow, float2* image, const float2* data, hipStream_t st,
               const float2* mu = nullptr) {
    if (!slow && !image) throw Error("background slowness not set (call fwix_born_set_slow)");
    if (image || mu) no_batch("the migration image");
    const size_t n = P();
    if (!add) {
      if (dslow) hip_check(launch_zero(dslow, MZ() * d.nz, st), "zero model");
      if (image) hip_check(launch_zero(image, (size_t)d.ny * d.nx * d.nz, st), "zero image");
    }
    float2* l = (float2*)lam.p;
    hip_check(launch_zero(l, n, st), "zero lambda");
    for (int z = d.nz - 1; z >= 0; --z) {
      if (z + 1 < d.nz) {  // lam_{nz} = 0: nothing to propagate or image at the bottom
        const float2* bg = background(z, st);
        if (image) {
          eng.apply(true, z, 0, l, l, st);
          hip_check(launch_image(bg, l, image + (size_t)z * d.ny * d.nx, d.NX, d.NY, d.NW, d.NS, d.nx, d.ny, 1, st),
                    "image");
          ++steps;
        } else {
          step_adj(z, l, bg, dslow + (size_t)z * MZ(), st);
        }
      }
      rec.forward(z, data, l, st);  // adjoint source (ExtendedBorn.cpp:195-196)
      if (mu) {
        ReflArgs r = refl_args(z, l, mu + (size_t)z * n, nullptr);
        hip_check(launch_refl_src(2, r, st), "down-reflected adjoint source");
      }
    }
  }

  // point-wise reflection arguments at depth z (slices z and z+1 of the models)
  ReflArgs refl_args(int z, float2* out, const float2* U, const float2* dUz) const {
    ReflArgs r{};
    const size_t mz = MZ();
    const int z1 = z + 1 < d.nz ? z + 1 : z;
    r.out = out; r.U = U; r.dU = dUz;
    r.s0 = slow + (size_t)z * mz; r.s1 = slow + (size_t)z1 * mz;
    r.r0 = den + (size_t)z * mz; r.r1 = den + (size_t)z1 * mz;
    r.NX = d.NX; r.NY = d.NY; r.NW = d.NW; r.NS = d.NS; r.nx = d.nx; r.ny = d.ny;
    r.last = z + 1 >= d.nz;
    return r;
  }
  void need_up() const {
    no_batch("the up-going (reflection) channel");
    if (!den) throw Error("density not set (call fwix_born_set_density)");
    if (!slow) throw Error("background slowness not set (call fwix_born_set_slow)");
  }
  void alloc_stores(bool up) {
    const size_t bytes = P() * sizeof(float2) * (size_t)d.nz;
    if (up && vstore.n != bytes) vstore.alloc(bytes);
    if (xstore.n != bytes) xstore.alloc(bytes);
  }

  // Up pass of the nonlinear modeling (Propagator.cpp:181-211): V_nz = 0;
  // V_z = P_z V_{z+1} + r_z U_z with r from (slowness, density); every V_z is
  // kept in HBM (the up-scattering channel's background); data += Rec_z^T V_z.
  void model_up(float2* data, int add, hipStream_t st) {
    need_up();
    if (!have_bg) throw Error("background wavefield not modeled (call fwix_born_model)");
    const size_t n = P();
    const size_t bytes = n * sizeof(float2) * (size_t)d.nz;
    if (vstore.n != bytes) vstore.alloc(bytes);
    if (data && !add) hip_check(launch_zero(data, (size_t)rec.nsamp, st), "zero data");
    float2* V = (float2*)vstore.p;
    for (int z = d.nz - 1; z >= 0; --z) {
      float2* vz = V + (size_t)z * n;
      if (z + 1 < d.nz) {
        eng.apply(false, z, 0, vz + n, vz, st);
        ++steps;
      } else {
        hip_check(launch_zero(vz, n, st), "zero V");
      }
      ReflArgs r = refl_args(z, vz, background(z, st), nullptr);
      hip_check(launch_refl_src(0, r, st), "reflect");
      if (data) rec.adjoint(z, data, vz, st);
    }
    have_up = true;
  }

  // The four channels of ExtendedBorn::forward (ExtendedBorn.cpp:80-118) with
  // model (ds, drho): down scattering (channel 1, keeping every dU_z), then one
  // up-going recursion carrying the other three together, which is possible
  // because they are linear and share the same propagator and recording:
  //   T_nz = 0;  T_z = P_z[ T_{z+1} + S_z(V_{z+1} ds_{z+1}) ] + r_z dU_z + U_z L_z(ds, drho)
  // (up scattering :282-311 with its z-1 operators, down-reflected :210-238,
  // back scattering :352-389), data += Rec_z^T T_z.
  void ext_forward(int add, const float2* ds, const float2* dr, float2* data, hipStream_t st) {
    need_up();
    if (!have_up) throw Error("up-going background not modeled (call fwix_born_model_up)");
    alloc_stores(false);
    const size_t n = P(), mz = MZ();
    forward(add, ds, data, st, true);
    float2* T = (float2*)dU.p;
    const float2* V = (const float2*)vstore.p;
    const float2* dus = (const float2*)xstore.p;
    hip_check(launch_zero(T, n, st), "zero T");
    for (int z = d.nz - 1; z >= 0; --z) {
      if (z + 1 < d.nz) step_fwd(z, V + (size_t)(z + 1) * n, ds + (size_t)(z + 1) * mz, T, T, st);
      ReflArgs r = refl_args(z, T, background(z, st), dus + (size_t)z * n);
      const int z1 = z + 1 < d.nz ? z + 1 : z;
      r.ds0 = ds + (size_t)z * mz; r.ds1 = ds + (size_t)z1 * mz;
      r.dr0 = dr + (size_t)z * mz; r.dr1 = dr + (size_t)z1 * mz;
      hip_check(launch_refl_src(1, r, st), "reflection sources");
      rec.adjoint(z, data, T, st);
    }
  }

  // Adjoint (ExtendedBorn.cpp:122-150): the up-going residual
  // mu_z = Inj_z(data) + P_{z-1}^H mu_{z-1} (kept for every depth) feeds the
  // up-scattering adjoint (ds_z += conj(V_z) S_{z-1}^H P_{z-1}^H mu_{z-1}) and
  // the back-scattering adjoint (L_z^H of sum_s conj(U_z) mu_z); then the
  // down-scattering adjoint runs with the extra source conj(r_z) mu_z of the
  // down-reflected channel.
  void ext_adjoint(int add, float2* ds, float2* dr, const float2* data, hipStream_t st) {
    need_up();
    if (!have_up) throw Error("up-going background not modeled (call fwix_born_model_up)");
    alloc_stores(false);
    const size_t n = P(), mz = MZ();
    if (!add) {
      hip_check(launch_zero(ds, mz * d.nz, st), "zero model");
      hip_check(launch_zero(dr, mz * d.nz, st), "zero model");
    }
    float2* mu = (float2*)lam.p;
    float2* mus = (float2*)xstore.p;
    const float2* V = (const float2*)vstore.p;
    hip_check(launch_zero(mu, n, st), "zero mu");
    for (int z = 0; z < d.nz; ++z) {
      if (z > 0) step_adj(z - 1, mu, V + (size_t)z * n, ds + (size_t)z * mz, st);
      rec.forward(z, data, mu, st);
      hip_check(launch_axpy(mu, mus + (size_t)z * n, n, 1.f, 0, st), "keep mu");
      ReflArgs r = refl_args(z, nullptr, background(z, st), mu);
      const int z1 = z + 1 < d.nz ? z + 1 : z;
      r.gs0 = ds + (size_t)z * mz; r.gs1 = ds + (size_t)z1 * mz;
      r.gr0 = dr + (size_t)z * mz; r.gr1 = dr + (size_t)z1 * mz;
      hip_check(launch_refl_adj(r, st), "back-scattering adjoint");
    }
    adjoint(1, ds, nullptr, data, st, mus);
  }

  // Up-going reflection recording of the background (Propagator.cpp:185-208):
  // V = 0; for z = nz-1..0: V <- P_z V; V += refl_z * U_z (model region);
  // data += Rec_z^T V.  Linear in refl (nz, NW, ny, nx).
  void reflect_forward(int add, const float2* refl, float2* data, hipStream_t st) {
    no_batch("the reflection channel");
    const size_t n = P(), mz = (size_t)d.NW * d.ny * d.nx;
    if (!add) hip_check(launch_zero(data, (size_t)rec.nsamp, st), "zero data");
    float2* v = (float2*)dU.p;
    hip_check(launch_zero(v, n, st), "zero V");
    for (int z = d.nz - 1; z >= 0; --z) {
      if (z + 1 < d.nz) {  // V is still zero at the bottom
        eng.apply(false, z, 0, v, v, st);
        ++steps;
      }
      const float2* bg = background(z, st);
      {
        PassTimer t("reflect_ic", 24.0 * (double)n + 8.0 * mz, st);
        hip_check(launch_ic_fwd(refl + (size_t)z * mz, v, bg, d.NX, d.NY, d.NW, d.NS, d.nx, d.ny, 1, st), "reflect");
      }
      rec.adjoint(z, data, v, st);
    }
  }

  // Adjoint of reflect_forward in refl: mu_0 = Inj_0(data),
  // mu_z = Inj_z(data) + P_{z-1}^H mu_{z-1};  refl_z (+)= sum_s conj(U_z) mu_z.
  // With image != null: image_z (+)= sum_{w,s} conj(U_z) mu_z, the migration
  // image of surface data (source wavefield U, receiver wavefield mu).
  void reflect_adjoint(int add, float2* refl, float2* image, const float2* data, hipStream_t st) {
    no_batch("the reflection channel");
    const size_t n = P(), mz = (size_t)d.NW * d.ny * d.nx;
    if (!add) {
      if (refl) hip_check(launch_zero(refl, mz * d.nz, st), "zero model");
      if (image) hip_check(launch_zero(image, (size_t)d.ny * d.nx * d.nz, st), "zero image");
    }
    float2* mu = (float2*)lam.p;
    hip_check(launch_zero(mu, n, st), "zero mu");
    for (int z = 0; z < d.nz; ++z) {
      if (z > 0) {
        eng.apply(true, z - 1, 0, mu, mu, st);
        ++steps;
      }
      rec.forward(z, data, mu, st);
      const float2* bg = background(z, st);
      if (image) {
        PassTimer t("reflect_image", 16.0 * (double)n + 16.0 * d.nx * d.ny, st);
        hip_check(launch_image(bg, mu, image + (size_t)z * d.ny * d.nx, d.NX, d.NY, d.NW, d.NS, d.nx, d.ny, 1, st),
                  "reflect image");
      } else {
        PassTimer t("reflect_ic_adj", 16.0 * (double)d.nx * d.ny * d.NW * d.NS + 16.0 * mz, st);
        hip_check(launch_ic_adj(refl + (size_t)z * mz, mu, bg, d.NX, d.NY, d.NW, d.NS, d.nx, d.ny, 1, st),
                  "reflect adjoint");
      }
    }
  }
};

}  // namespace

extern "C" {

// ---------------------------------------------------------------- Injection
int fwix_injection_create(const fwix_dims* d, int ntrace, fwix_op* out) {
  return guard([&] {
    validate_dims(d, false);
    if (ntrace < 0) throw Error("ntrace must be >= 0");
    if (!out) throw Error("null output handle");
    auto* op = new InjectionOp();
    op->d = *d;
    op->ntrace = ntrace;
    *out = op;
  });
}
int fwix_injection_set_coords(fwix_op op, const float* cx, const float* cy, const float* cz, const int* ids) {
  return guard([&] {
    auto* p = as<InjectionOp>(op, KIND_INJECTION);
    int maxlev = std::max(p->d.nz, 1);
    for (int t = 0; t < p->ntrace; ++t) {
      const int iz = (int)floorf((cz[t] - p->d.oz) / p->d.dz);
      maxlev = std::max(maxlev, iz + 2);
    }
    p->plan.build(p->d, maxlev, p->ntrace, cx, cy, cz, ids);
    p->has_coords = true;
  });
}
int fwix_injection_set_depth(fwix_op op, int iz) {
  return guard([&] { as<InjectionOp>(op, KIND_INJECTION)->iz = iz; });
}
int fwix_injection_forward(fwix_op op, int add, const fwix_c64* traces, fwix_c64* wfld, fwix_stream s) {
  return guard([&] {
    auto* p = as<InjectionOp>(op, KIND_INJECTION);
    if (!p->has_coords) throw Error("injection coordinates not set");
    if (!add) hip_check(launch_zero((float2*)wfld, wsize(p->d), S(s)), "zero");
    p->plan.forward(p->iz, (const float2*)traces, (float2*)wfld, S(s));
  });
}
int fwix_injection_adjoint(fwix_op op, int add, fwix_c64* traces, const fwix_c64* wfld, fwix_stream s) {
  return guard([&] {
    auto* p = as<InjectionOp>(op, KIND_INJECTION);
    if (!p->has_coords) throw Error("injection coordinates not set");
    if (!add) hip_check(launch_zero((float2*)traces, (size_t)p->plan.nsamp, S(s)), "zero");
    p->plan.adjoint(p->iz, (float2*)traces, (const float2*)wfld, S(s));
  });
}

// ---------------------------------------------------------------- Scatter
int fwix_scatter_create(const fwix_dims* d, const fwix_params* par, fwix_op* out) {
  return guard([&] {
    validate_dims(d, true);
    validate_params(par);
    if (d->nz < 1) throw Error("dims: nz must be >= 1");
    for (int i = 0; i < d->NW; ++i)
      if ((double)d->ow + (double)i * d->dw == 0.0) throw Error("Frequency is zero in the scattering operator!");
    if (!out) throw Error("null output handle");
    auto* op = new ScatterOp();
    try {
      fwix_params p = *par;
      op->eng.init(FWIX_PSPI, *d, p, p.taylor + 1);
    } catch (...) { delete op; throw; }
    *out = op;
  });
}
int fwix_scatter_set_slow(fwix_op op, const fwix_c64* slow_dev) {
  return guard([&] {
    if (!slow_dev) throw Error("null slowness");
    as<ScatterOp>(op, KIND_SCATTER)->slow = (const float2*)slow_dev;
  });
}
int fwix_scatter_set_depth(fwix_op op, int iz) {
  return guard([&] {
    auto* p = as<ScatterOp>(op, KIND_SCATTER);
    if (iz < 0 || iz >= p->eng.tables.dims.nz) throw Error("depth index out of range");
    p->iz = iz;
  });
}
int fwix_scatter_forward(fwix_op op, int add, const fwix_c64* m, fwix_c64* d, fwix_stream s) {
  return guard([&] { as<ScatterOp>(op, KIND_SCATTER)->run(false, add, (const float2*)m, (float2*)d, S(s)); });
}
int fwix_scatter_adjoint(fwix_op op, int add, fwix_c64* m, const fwix_c64* d, fwix_stream s) {
  return guard([&] { as<ScatterOp>(op, KIND_SCATTER)->run(true, add, (const float2*)d, (float2*)m, S(s)); });
}

// ---------------------------------------------------------------- ImagingCondition
int fwix_ic_create(const fwix_dims* d, fwix_op* out) {
  return guard([&] {
    validate_dims(d, false);
    if (!out) throw Error("null output handle");
    auto* op = new ICOp();
    op->d = *d;
    *out = op;
  });
}
int fwix_ic_set_background(fwix_op op, const fwix_c64* bg_dev) {
  return guard([&] {
    if (!bg_dev) throw Error("null background wavefield");
    as<ICOp>(op, KIND_IC)->bg = (const float2*)bg_dev;
  });
}
int fwix_ic_forward(fwix_op op, int add, const fwix_c64* img, fwix_c64* wfld, fwix_stream s) {
  return guard([&] {
    auto* p = as<ICOp>(op, KIND_IC);
    if (!p->bg) throw Error("background wavefield not set");
    const fwix_dims& d = p->d;
    hip_check(launch_ic_fwd((const float2*)img, (float2*)wfld, p->bg, d.NX, d.NY, d.NW, d.NS, d.nx, d.ny, add, S(s)),
              "ic forward");
  });
}
int fwix_ic_adjoint(fwix_op op, int add, fwix_c64* img, const fwix_c64* wfld, fwix_stream s) {
  return guard([&] {
    auto* p = as<ICOp>(op, KIND_IC);
    if (!p->bg) throw Error("background wavefield not set");
    const fwix_dims& d = p->d;
    hip_check(launch_ic_adj((float2*)img, (const float2*)wfld, p->bg, d.NX, d.NY, d.NW, d.NS, d.nx, d.ny, add, S(s)),
              "ic adjoint");
  });
}

// ---------------------------------------------------------------- Downward / Upward
int fwix_oneway_create(int direction, int method, const fwix_dims* d, const fwix_params* par, fwix_op* out) {
  return guard([&] {
    validate_dims(d, true);
    validate_params(par);
    if (direction != 0 && direction != 1) throw Error("direction must be 0 (down) or 1 (up)");
    if (method != FWIX_PSPI && method != FWIX_NSPS) throw Error("method must be FWIX_PSPI or FWIX_NSPS");
    if (d->nz < 1) throw Error("dims: nz must be >= 1");
    if (!out) throw Error("null output handle");
    auto* op = new OneWayOp();
    try {
      op->direction = direction;
      op->eng.init(method, *d, *par);
      op->u.alloc(wsize(*d) * sizeof(float2));
    } catch (...) { delete op; throw; }
    *out = op;
  });
}
int fwix_oneway_set_shot_planes(fwix_op op, int nshot, const int* nplanes) {
  return guard([&] { as<OneWayOp>(op, KIND_ONEWAY)->eng.set_model_map(nshot, nplanes); });
}
int fwix_oneway_set_frequencies(fwix_op op, const float* f_hz) {
  return guard([&] { as<OneWayOp>(op, KIND_ONEWAY)->eng.tables.set_frequencies(f_hz, false); });
}
int fwix_oneway_set_step_chaining(fwix_op op, int on) {
  return guard([&] {
    auto* p = as<OneWayOp>(op, KIND_ONEWAY);
    p->eng.chain = on != 0;
    if (!on) p->eng.Kall.release();
  });
}
int fwix_oneway_set_refs(fwix_op op, const fwix_c64* sref, const uint8_t* labels) {
  return guard([&] {
    auto* p = as<OneWayOp>(op, KIND_ONEWAY);
    p->eng.set_refs((const float2*)sref, labels);
  });
}
int fwix_oneway_forward(fwix_op op, int add, const fwix_c64* m, fwix_c64* d, fwix_c64* ckpt, fwix_stream s) {
  return guard([&] {
    as<OneWayOp>(op, KIND_ONEWAY)->run(false, add, (const float2*)m, (float2*)d, (float2*)ckpt, S(s));
  });
}
int fwix_oneway_adjoint(fwix_op op, int add, fwix_c64* m, const fwix_c64* d, fwix_stream s) {
  return guard([&] {
    as<OneWayOp>(op, KIND_ONEWAY)->run(true, add, (const float2*)d, (float2*)m, nullptr, S(s));
  });
}

// ---------------------------------------------------------------- Born
int fwix_born_create(const fwix_dims* d, const fwix_params* par, int ckpt_every, fwix_op* out) {
  return guard([&] {
    validate_dims(d, true);
    validate_params(par);
    if (d->nz < 1) throw Error("dims: nz must be >= 1");
    if (ckpt_every < 0) throw Error("ckpt_every must be >= 0");
    for (int i = 0; i < d->NW; ++i)
      if ((double)d->ow + (double)i * d->dw == 0.0) throw Error("Frequency is zero in the scattering operator!");
    if (!out) throw Error("null output handle");
    auto* op = new BornOp();
    try {
      op->d = *d;
      op->par = *par;
      op->ckpt = ckpt_every <= 1 ? 1 : std::min(ckpt_every, d->nz);
      op->eng.init(FWIX_PSPI, *d, *par, par->taylor + 1);
      const size_t n = wsize(*d) * sizeof(float2);
      const size_t nck = (size_t)((d->nz + op->ckpt - 1) / op->ckpt);
      op->store.alloc(n * nck);
      if (op->ckpt > 1) op->seg.alloc(n * op->ckpt);
      op->U.alloc(n);
      op->dU.alloc(n);
      op->lam.alloc(n);
    } catch (...) { delete op; throw; }
    *out = op;
  });
}
int fwix_born_set_frequencies(fwix_op op, const float* f_hz) {
  return guard([&] {
    auto* p = as<BornOp>(op, KIND_BORN);
    p->eng.tables.set_frequencies(f_hz, true);
    p->have_bg = false;
    p->have_up = false;
  });
}
int fwix_born_set_refs(fwix_op op, const fwix_c64* sref, const uint8_t* labels) {
  return guard([&] {
    auto* p = as<BornOp>(op, KIND_BORN);
    p->eng.set_refs((const float2*)sref, labels);
    p->have_bg = false;
    p->have_up = false;
  });
}
int fwix_born_set_slow(fwix_op op, const fwix_c64* slow_dev) {
  return guard([&] {
    if (!slow_dev) throw Error("null slowness");
    auto* p = as<BornOp>(op, KIND_BORN);
    p->slow = (const float2*)slow_dev;
    p->have_up = false;
  });
}
int fwix_born_set_codec(fwix_op op, int codec) {
  return guard([&] { as<BornOp>(op, KIND_BORN)->set_codec(codec); });
}
int fwix_born_set_precision(fwix_op op, int precision) {
  return guard([&] {
    auto* p = as<BornOp>(op, KIND_BORN);
    if (precision != FWIX_PREC_F32 && precision != FWIX_PREC_F64_SCATTER)
      throw Error("precision must be FWIX_PREC_F32 or FWIX_PREC_F64_SCATTER");
    p->f64 = precision == FWIX_PREC_F64_SCATTER;
  });
}
int fwix_born_set_step_chaining(fwix_op op, int on) {
  return guard([&] {
    auto* p = as<BornOp>(op, KIND_BORN);
    p->eng.chain = on != 0;
    if (!on) p->eng.Kall.release();
  });
}
int fwix_bfp16_pack(const fwix_c64* in, void* q, float* scales, size_t n, fwix_stream s) {
  return guard([&] {
    if ((!in || !q || !scales) && n) throw Error("null buffer");
    hip_check(launch_bfp16_pack((const float2*)in, q, scales, n, S(s)), "bfp16 pack");
  });
}
int fwix_bfp16_unpack(const void* q, const float* scales, fwix_c64* out, size_t n, fwix_stream s) {
  return guard([&] {
    if ((!q || !scales || !out) && n) throw Error("null buffer");
    hip_check(launch_bfp16_unpack(q, scales, (float2*)out, n, S(s)), "bfp16 unpack");
  });
}
int fwix_born_set_density(fwix_op op, const fwix_c64* den_dev) {
  return guard([&] {
    if (!den_dev) throw Error("null density");
    auto* p = as<BornOp>(op, KIND_BORN);
    p->den = (const float2*)den_dev;
    p->have_up = false;
  });
}
int fwix_born_set_method(fwix_op op, int method) {
  return guard([&] { as<BornOp>(op, KIND_BORN)->set_method(method); });
}
int fwix_born_set_shot_planes(fwix_op op, int nshot, const int* nplanes) {
  return guard([&] { as<BornOp>(op, KIND_BORN)->set_shot_planes(nshot, nplanes); });
}
long long fwix_born_trace_samples(fwix_op op, int receivers) {
  if (!op || op->kind != KIND_BORN) return -1;
  auto* p = static_cast<BornOp*>(op);
  return receivers ? p->rec.nsamp : p->src.nsamp;
}
int fwix_born_set_sources(fwix_op op, int n, const float* x, const float* y, const float* z, const int* ids) {
  return guard([&] {
    auto* p = as<BornOp>(op, KIND_BORN);
    p->build_plan(p->src, n, x, y, z, ids);
    p->nsrc = n;
    p->have_bg = false;
    p->have_up = false;
  });
}
int fwix_born_set_receivers(fwix_op op, int n, const float* x, const float* y, const float* z, const int* ids) {
  return guard([&] {
    auto* p = as<BornOp>(op, KIND_BORN);
    p->build_plan(p->rec, n, x, y, z, ids);
    p->nrec = n;
  });
}
int fwix_born_set_wavelet(fwix_op op, const fwix_c64* wavelet_dev) {
  return guard([&] {
    if (!wavelet_dev) throw Error("null wavelet");
    auto* p = as<BornOp>(op, KIND_BORN);
    p->wavelet = (const float2*)wavelet_dev;
    p->have_bg = false;
    p->have_up = false;
  });
}
int fwix_born_model(fwix_op op, int add, fwix_c64* data, fwix_stream s) {
  return guard([&] {
    auto* p = as<BornOp>(op, KIND_BORN);
    p->model((float2*)data, add, S(s));
    p->have_up = false;
  });
}
int fwix_born_model_up(fwix_op op, int add, fwix_c64* data, fwix_stream s) {
  return guard([&] { as<BornOp>(op, KIND_BORN)->model_up((float2*)data, add, S(s)); });
}
int fwix_born_extended_forward(fwix_op op, int add, const fwix_c64* dslow, const fwix_c64* dden, fwix_c64* data,
                               fwix_stream s) {
  return guard([&] {
    if (!dslow || !dden || !data) throw Error("null model or data");
    as<BornOp>(op, KIND_BORN)->ext_forward(add, (const float2*)dslow, (const float2*)dden, (float2*)data, S(s));
  });
}
int fwix_born_extended_adjoint(fwix_op op, int add, fwix_c64* dslow, fwix_c64* dden, const fwix_c64* data,
                               fwix_stream s) {
  return guard([&] {
    if (!dslow || !dden || !data) throw Error("null model or data");
    as<BornOp>(op, KIND_BORN)->ext_adjoint(add, (float2*)dslow, (float2*)dden, (const float2*)data, S(s));
  });
}
int fwix_born_forward(fwix_op op, int add, const fwix_c64* dslow, fwix_c64* data, fwix_stream s) {
  return guard([&] { as<BornOp>(op, KIND_BORN)->forward(add, (const float2*)dslow, (float2*)data, S(s)); });
}
int fwix_born_adjoint(fwix_op op, int add, fwix_c64* dslow, const fwix_c64* data, fwix_stream s) {
  return guard([&] {
    if (!dslow) throw Error("null model");
    as<BornOp>(op, KIND_BORN)->adjoint(add, (float2*)dslow, nullptr, (const float2*)data, S(s));
  });
}
int fwix_born_migrate(fwix_op op, int add, fwix_c64* image, const fwix_c64* data, fwix_stream s) {
  return guard([&] {
    if (!image) throw Error("null image");
    as<BornOp>(op, KIND_BORN)->adjoint(add, nullptr, (float2*)image, (const float2*)data, S(s));
  });
}
int fwix_born_reflect_forward(fwix_op op, int add, const fwix_c64* refl, fwix_c64* data, fwix_stream s) {
  return guard([&] {
    if (!refl || !data) throw Error("null reflectivity or data");
    as<BornOp>(op, KIND_BORN)->reflect_forward(add, (const float2*)refl, (float2*)data, S(s));
  });
}
int fwix_born_reflect_adjoint(fwix_op op, int add, fwix_c64* refl, const fwix_c64* data, fwix_stream s) {
  return guard([&] {
    if (!refl || !data) throw Error("null reflectivity or data");
    as<BornOp>(op, KIND_BORN)->reflect_adjoint(add, (float2*)refl, nullptr, (const float2*)data, S(s));
  });
}
int fwix_born_reflect_migrate(fwix_op op, int add, fwix_c64* image, const fwix_c64* data, fwix_stream s) {
  return guard([&] {
    if (!image || !data) throw Error("null image or data");
    as<BornOp>(op, KIND_BORN)->reflect_adjoint(add, nullptr, (float2*)image, (const float2*)data, S(s));
  });
}
int fwix_reflect_coef(const fwix_dims* d, const fwix_c64* slow, const fwix_c64* den, fwix_c64* refl, fwix_stream s) {
  return guard([&] {
    validate_dims(d, false);  // point-wise: any lateral size
    if (d->nz < 1) throw Error("dims: nz must be >= 1");
    if (!slow || !den || !refl) throw Error("null slowness, density or output");
    const size_t slice = (size_t)d->NW * d->ny * d->nx;
    hip_check(launch_reflect_coef((const float2*)slow, (const float2*)den, (float2*)refl, d->nz, slice, S(s)),
              "reflection coefficients");
  });
}
long long fwix_born_steps(fwix_op op) {
  if (!op || op->kind != KIND_BORN) return -1;
  return static_cast<BornOp*>(op)->steps;
}

}  // extern "C"
