// Internal host-side types of libpyfwix_hip.so (handles, device tables,
// scratch buffers, error plumbing).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "fwix_hip.h"
#include "fwix_launch.h"

enum {
  KIND_FFT = 1,
  KIND_PHASE_SHIFT,
  KIND_SELECTOR,
  KIND_TAPER,
  KIND_ONESTEP,
  KIND_ONEWAY,
  KIND_INJECTION,
  KIND_SCATTER,
  KIND_IC,
  KIND_BORN,
};

struct fwix_op_s {
  int kind = 0;
  virtual ~fwix_op_s() = default;
};

namespace fwix {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

extern thread_local std::string g_err;
void set_error(const std::string& s);
void hip_check(hipError_t e, const char* what);

// Runs f and converts any exception into the C ABI's error convention.
template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    set_error(e.what());
    return -1;
  } catch (...) {
    set_error("unknown error");
    return -1;
  }
}

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t bytes);
  void release();
};

// Device tables of one wavefield hypercube: wavenumbers, frequencies, taper
// weights and FFT twiddles.
struct Tables {
  fwix_dims dims{};
  DevBuf b_kx, b_ky, b_w2, b_w, b_tapx, b_tapy, b_twx, b_twy;
  DevBuf b_twx64, b_twy64, b_tapx64, b_tapy64;  // float64 twiddles and taper weights
  void init(const fwix_dims& d, int tx, int ty);
  // per-plane frequencies in Hz (NW of them) in place of ow + iw * dw
  void set_frequencies(const float* f_hz, bool nonzero);
  KArgs kargs() const;
  PwArgs pwargs() const;
};

void validate_dims(const fwix_dims* d, bool need_fft);

// Live per-pass timing (fwix_timing_*): HIP events recorded on the launch
// stream around every pass of the hot path, with the pass's compulsory HBM
// bytes.  A no-op unless enabled.
// A null name records nothing.
struct PassTimer {
  PassTimer(const char* name, double bytes, hipStream_t st);
  PassTimer(const PassTimer&) = delete;
  PassTimer& operator=(const PassTimer&) = delete;
  ~PassTimer();
  int slot = -1;
  hipStream_t st = nullptr;
};
void validate_params(const fwix_params* p);
int auto_chunk(const fwix_dims& d, int nterm, int requested);
// Step chaining (forward PSPI / adjoint NSPS only, the "shared spectrum"
// shape): with kall the chunks' row spectra live in kall (all planes) instead
// of the proto's K; k_ready skips the first row pass (kall already holds
// scale * F_x(taper * in), left by the previous step), and k_next makes the
// select pass leave scale * F_x(taper * out) there for the next step;
// out_needed = false (with k_next) skips writing out itself.
void step_passes(int method, bool adjoint, const fwix_dims& d, const fwix_params& p, const KArgs& proto, int chunk,
                 int add, const float2* in, float2* out, hipStream_t st,
                 const std::function<double(int, int)>* rows, float2* kall = nullptr, bool k_ready = false,
                 bool k_next = false, bool out_needed = true);

// One depth step (PSPI or NSPS) over a whole wavefield, forward or adjoint,
// run as chunks of planes through the three fused passes.
struct StepEngine {
  int method = FWIX_PSPI;
  bool plain = false;  // reference-granularity per-operator kernels (comparator path)
  DevBuf T;            // plain path: second whole-shot scratch
  fwix_params params{};
  Tables tables;
  const float2* sref = nullptr;     // borrowed (nz, nref, NW), or the expansion below
  const uint8_t* labels = nullptr;  // borrowed (nz, nmodel, NY, NX)
  DevBuf K, W;
  DevBuf rowmask, planemask;  // reference presence per (depth, model plane, y) row and per (depth, model plane)
  bool masks = false;
  std::vector<int> present;   // (nz, nmodel): present (row, reference) pairs per model plane
  // Shot batch: plane iw takes its labels / slowness / model from model plane
  // mplane[iw] (the shot it belongs to); nmodel == NW and no map otherwise.
  int nmodel = 0;
  std::vector<int> mplane;
  DevBuf mplane_dev, sref_exp;  // device map; references expanded to (nz, nref, NW)
  void set_model_map(int nshot, const int* nplanes);
  int model_of(int iw) const { return mplane.empty() ? iw : mplane[iw]; }
  // rows of the per-reference scratch a chunk actually reads or writes
  double present_rows(int iz, int p0, int np) const;
  int chunk = 1;
  // borrow refs / labels for all depths and derive the presence masks
  void set_refs(const float2* sref, const uint8_t* labels);
  // wterms: planes per chunk of the W scratch (>= nref; the Born path also
  // stores its taylor+1 scattering terms there)
  void init(int method, const fwix_dims& d, const fwix_params& p, int wterms = 0);
  void check_ready(int iz) const;
  KArgs base(int iz, int c0, int np) const;
  void apply(bool adjoint, int iz, int add, const float2* in, float2* out, hipStream_t st);
  void apply_plain(bool adjoint, int iz, int add, const float2* in, float2* out, hipStream_t st);
  // Steps of a recursion chained through the row spectrum (Kall, all
  // planes): k_ready = the previous chained step left F_x(taper * in) in Kall
  // (nothing changed in between), k_next = leave F_x(taper * out) for the next
  // step.  Falls back to apply() where chaining does not apply (plain path,
  // the masked-input shape: PSPI adjoint, NSPS forward).  chain = false turns
  // chaining off (every step whole).
  DevBuf Kall;
  bool chain = true;
  // the steps of this direction have the "shared spectrum" shape (PSPI
  // forward, NSPS adjoint): a row pass, the column fan-out, the select pass
  bool chains(bool adjoint = false) const { return chain && !plain && ((method == FWIX_PSPI) != adjoint); }
  // out_needed = false (only with k_next): nobody reads out before the next
  // chained step overwrites it, so it is not written at all
  void apply_chained(int iz, const float2* in, float2* out, bool k_ready, bool k_next, hipStream_t st,
                     bool out_needed = true, bool adjoint = false);
};


// Source / receiver injection plan (propagator/src/injection.cu): per depth
// level, the target nodes with their (trace, weight) contributions (forward,
// gathered per node) and the active traces with their four corners
// (adjoint).  Built on the host from float32 coordinates with the kernel's
// own float arithmetic.
//
// Traces are ragged when the shot ids map to plane ranges (a shot batch):
// trace t holds count[id_t] samples, one per plane of its shot, stored one
// trace after another; with no ranges shot id s owns planes [s NW, (s+1) NW)
// and every trace holds NW samples.
struct InjPlan {
  int ntrace = 0, nlev = 0, NW = 0, NX = 0;
  int wmax = 0;           // largest number of planes of one trace
  long long plane = 0;
  long long nsamp = 0;    // samples of all traces
  std::vector<int> seg_lvl, act_lvl;  // (nlev + 1) offsets
  DevBuf seg_base, seg_off, seg_np, ent_trace, ent_w, act_trace, act_base, act_w, act_np, toff;
  void build(const fwix_dims& d, int nlev, int ntrace, const float* cx, const float* cy, const float* cz,
             const int* ids, const std::vector<int>* first = nullptr, const std::vector<int>* count = nullptr);
  void forward(int iz, const float2* traces, float2* wfld, hipStream_t st) const;  // wfld += Inj_iz(traces)
  void adjoint(int iz, float2* traces, const float2* wfld, hipStream_t st) const;  // traces += Rec_iz^T wfld
  // whether forward(iz) changes the wavefield (some trace injects at level iz)
  bool injects(int iz) const { return iz >= 0 && iz < nlev && seg_lvl[iz + 1] > seg_lvl[iz]; }
  // whether adjoint(iz) reads the wavefield (some trace records at level iz)
  bool records(int iz) const { return iz >= 0 && iz < nlev && act_lvl[iz + 1] > act_lvl[iz]; }
};

}  // namespace fwix
