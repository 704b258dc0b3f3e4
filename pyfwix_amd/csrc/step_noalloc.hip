// One PSPI / NSPS depth step that allocates nothing (fwix_onestep_apply_noalloc).
//
// The handle-based step (fwix_onestep_create + fwix_onestep_forward/adjoint)
// keeps its tables and its spectrum / per-reference scratch in buffers the
// handle allocates.  This entry point runs the same three fused passes with
// all of that in one file-scope device array, `scratch`, so a caller that
// cannot call hipMalloc (code under HIP graph capture, serving loops,
// harnesses that forbid allocation) can still run a step.  Each GPU holds its
// own copy of the array from the first call on it until the process ends;
// calls that overlap on one GPU race on it.
#include "fwix_hip.h"
#include "fwix_launch.h"
#include "capi_internal.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

using namespace fwix;

namespace {

static __device__ __align__(256) unsigned char scratch[FWIX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<void*, kMaxDevices> g_scratch{};  // device address of `scratch`, per device

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current.  Host code uses only this address: the array's
// own name is a host-side stand-in.
void* scratch_ptr() {
  int dev = -1;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw Error("no-allocation step: device index out of range");
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw Error("hipGetSymbolAddress(scratch) returned a null address");
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

constexpr size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Byte offsets of the step's tables and scratch inside the array.
struct Layout {
  size_t kx, ky, w2, tapx, tapy, twx, twy, K, W, end;
  int chunk;
};

Layout layout(const fwix_dims& d, const fwix_params& p) {
  Layout l{};
  size_t o = 0;
  l.kx = o;   o = align256(o + 4 * (size_t)d.NX);
  l.ky = o;   o = align256(o + 4 * (size_t)d.NY);
  l.w2 = o;   o = align256(o + 4 * (size_t)d.NW);
  l.tapx = o; o = align256(o + 4 * (size_t)d.NX);
  l.tapy = o; o = align256(o + 4 * (size_t)d.NY);
  l.twx = o;  o = align256(o + 8 * (size_t)d.NX);
  l.twy = o;  o = align256(o + 8 * (size_t)d.NY);
  const size_t plane = (size_t)d.NX * d.NY * sizeof(float2);
  const size_t per_plane = plane * (1 + (size_t)p.nref);  // spectrum + one plane per reference
  const long long planes = (long long)d.NW * d.NS;
  long long c = o < FWIX_NOALLOC_SCRATCH_BYTES ? (long long)((FWIX_NOALLOC_SCRATCH_BYTES - o) / per_plane) : 0;
  if (c < 1)
    throw Error("no-allocation step: NX=" + std::to_string(d.NX) + ", NY=" + std::to_string(d.NY) +
                ", nref=" + std::to_string(p.nref) + " needs " + std::to_string(o + per_plane) +
                " bytes of scratch for one plane; the array holds " +
                std::to_string((unsigned long long)FWIX_NOALLOC_SCRATCH_BYTES));
  if (p.planes_per_chunk > 0 && c > p.planes_per_chunk) c = p.planes_per_chunk;
  if (c > planes) c = planes;
  l.chunk = (int)c;
  l.K = o;
  o = align256(o + plane * (size_t)c);
  l.W = o;
  o += plane * (size_t)c * p.nref;
  l.end = o;
  return l;
}

// The step's tables, from the same double-precision expressions as the
// handle's host-built tables (wavenumbers PhaseShift.h:43-57, frequencies
// PhaseShift.h:59-70, taper taper.cu:43-55, FFT twiddles), rounded once.
__global__ void k_step_tables(float* kx, float* ky, float* w2, float* tapx, float* tapy, float2* twx, float2* twy,
                              int NX, int NY, int NW, double dx, double dy, double ow, double dw, int tx, int ty) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  auto kval = [](int i, int n, double d) {
    const double dk = 2.0 * M_PI / (d * n);
    return (float)((i <= n / 2 ? i : i - n) * dk);
  };
  auto tap = [](int i, int n, int t) {
    if (t > 0 && i < t) return (float)(0.5 * (1.0 - cos(M_PI * i / t)));
    if (t > 0 && i >= n - t) return (float)(0.5 * (1.0 + cos(M_PI * (i - n + t) / t)));
    return 1.f;
  };
  auto tw = [](int k, int n) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    return make_float2((float)cos(a), (float)sin(a));
  };
  if (i < NX) {
    kx[i] = kval(i, NX, dx);
    tapx[i] = tap(i, NX, tx);
    twx[i] = tw(i, NX);
  }
  if (i < NY) {
    ky[i] = kval(i, NY, dy);
    tapy[i] = tap(i, NY, ty);
    twy[i] = tw(i, NY);
  }
  if (i < NW) {
    const double f = 2.0 * M_PI * (ow + (double)i * dw);
    w2[i] = (float)(f * f);
  }
}

}  // namespace

extern "C" {

int fwix_onestep_apply_noalloc(int method, int adjoint, const fwix_dims* d, const fwix_params* p,
                               const fwix_c64* sref, const uint8_t* labels, int add, fwix_c64* model,
                               fwix_c64* data, fwix_stream s) {
  return guard([&] {
    validate_dims(d, true);
    validate_params(p);
    if (method != FWIX_PSPI && method != FWIX_NSPS) throw Error("method must be FWIX_PSPI or FWIX_NSPS");
    if (!sref || !labels) throw Error("null reference slownesses or labels");
    if (!model || !data) throw Error("null model or data");
    const bool pspi = method == FWIX_PSPI;
    const int tx = pspi ? p->taperx : 0, ty = pspi ? p->tapery : 0;
    if (tx > d->NX || ty > d->NY) throw Error("taper wider than the padded grid");
    const Layout l = layout(*d, *p);  // rejects grids whose scratch does not fit
    hipStream_t st = (hipStream_t)s;
    unsigned char* base = (unsigned char*)scratch_ptr();
    float* kx = (float*)(base + l.kx);
    float* ky = (float*)(base + l.ky);
    float* w2 = (float*)(base + l.w2);
    float* tapx = (float*)(base + l.tapx);
    float* tapy = (float*)(base + l.tapy);
    float2* twx = (float2*)(base + l.twx);
    float2* twy = (float2*)(base + l.twy);
    const int n = std::max(d->NX, std::max(d->NY, d->NW));
    hipLaunchKernelGGL(k_step_tables, dim3((n + 255) / 256), dim3(256), 0, st, kx, ky, w2, tapx, tapy, twx, twy,
                       d->NX, d->NY, d->NW, (double)d->dx, (double)d->dy, (double)d->ow, (double)d->dw, tx, ty);
    hip_check(hipGetLastError(), "step tables");
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.kx = kx; a.ky = ky; a.w2 = w2; a.tapx = tapx; a.tapy = tapy; a.twx = twx; a.twy = twy;
    a.NX = d->NX; a.NY = d->NY; a.NW = d->NW; a.NS = d->NS; a.nx = d->nx; a.ny = d->ny;
    a.dz = d->dz;
    a.K = (float2*)(base + l.K);
    a.W = (float2*)(base + l.W);
    a.labels = labels;
    a.sref = (const float2*)sref;
    a.nref = a.nterm = p->nref;  // no presence masks: every reference on every row
    a.eps = p->eps;
    a.eps_s = p->eps_scat;
    const float2* in = (const float2*)(adjoint ? data : model);
    float2* out = (float2*)(adjoint ? model : data);
    step_passes(method, adjoint != 0, *d, *p, a, l.chunk, add, in, out, st, nullptr);
  });
}

}  // extern "C"
