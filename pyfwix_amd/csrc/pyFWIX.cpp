// pyFWIX: the compiled Python module of the reference's front end
// (fwix/src/pyFWIX.cpp:9-15 = init_operator + init_propagator), with the
// same class names and argument lists (operator/src/python/pyCudaOperator.cpp:15-30,
// propagator/src/python/pyCudaWEM.cpp:22-242), over the C ABI of
// libpyfwix_hip.so through the C++ operator layer (fwix_operators.hpp).
//
// The SEP objects the reference's fwix/CudaWEM.py passes (hypercube,
// complexNDReg vectors, paramObj, each as `.cppMode`) are the stand-ins of
// pyfwix_amd/sep.py: a hypercube answers getAxis(i).n/.o/.d, a vector
// getNdArray() (a complex64 numpy array on the host, or a torch tensor on the
// GPU) and a paramObj getInt/getFloat(name, default).  Host vectors go
// through the operator's device buffers (CudaOperator.h:74-120); device
// vectors are used in place.
//
// Reference slownesses and labels of a slowness model come from
// pyfwix_amd.refsampler.select_refs (deterministic quantile references), the
// same selection the Python front end and the tests use.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "fwix_operators.hpp"

namespace py = pybind11;
using fwix::cfloat;
using fwix::DeviceVector;
using fwix::hostvec;

namespace {

// ------------------------------------------------------------------ stand-in accessors
struct Ax {
  int n;
  double o, d;
};
Ax axis_of(py::handle hyper, int i) {
  py::object a = hyper.attr("getAxis")(i);
  return {a.attr("n").cast<int>(), a.attr("o").cast<double>(), a.attr("d").cast<double>()};
}
int ndim_of(py::handle hyper) { return hyper.attr("getNdim")().cast<int>(); }
int par_int(py::handle par, const char* k, int def) { return par.attr("getInt")(k, def).cast<int>(); }
float par_float(py::handle par, const char* k, float def) { return par.attr("getFloat")(k, def).cast<float>(); }

// A vector's samples: a device pointer (torch tensor on the GPU) or a host
// numpy array (complex64, C order) read and written in place.
struct Vec {
  fwix_c64* dev = nullptr;
  py::array_t<std::complex<float>, py::array::c_style> host;
  size_t n = 0;
};
Vec vec_of(py::handle v) {
  py::object arr = v.attr("getNdArray")();
  Vec r;
  if (py::hasattr(arr, "is_cuda") && arr.attr("is_cuda").cast<bool>()) {
    if (!arr.attr("is_contiguous")().cast<bool>()) throw std::runtime_error("device vectors must be contiguous");
    if (py::str(arr.attr("dtype")).cast<std::string>() != "torch.complex64")
      throw std::runtime_error("device vectors must be complex64");
    r.dev = reinterpret_cast<fwix_c64*>(arr.attr("data_ptr")().cast<uintptr_t>());
    r.n = arr.attr("numel")().cast<size_t>();
    return r;
  }
  py::array a = py::array::ensure(arr);
  if (!a || !py::isinstance<py::array_t<std::complex<float>>>(a) || !(a.flags() & py::array::c_style))
    throw std::runtime_error("host vectors must be C-ordered complex64 numpy arrays");
  r.host = py::reinterpret_borrow<py::array_t<std::complex<float>, py::array::c_style>>(a);
  r.n = (size_t)r.host.size();
  return r;
}
hostvec to_host(const Vec& v) {
  hostvec h(v.n);
  std::memcpy(h.data(), v.host.data(), v.n * sizeof(cfloat));
  return h;
}
void from_host(Vec& v, const hostvec& h) { std::memcpy(v.host.mutable_data(), h.data(), v.n * sizeof(cfloat)); }
void sync() { fwix::hip_ok(hipDeviceSynchronize(), "synchronize"); }

// model (+)= A^H data or data (+)= A model on host or device vectors (both on one side)
template <class Op>
void apply(Op& op, bool forward, bool add, py::handle model, py::handle data) {
  Vec m = vec_of(model), d = vec_of(data);
  if (m.n != op.getDomainSize() || d.n != op.getRangeSize())
    throw std::runtime_error("vector sizes do not match the operator's domain / range");
  if (m.dev && d.dev) {
    if (forward) op.cu_forward(add, m.dev, d.dev);
    else op.cu_adjoint(add, m.dev, d.dev);
    sync();
    return;
  }
  if (m.dev || d.dev) throw std::runtime_error("model and data must both be on the host or both on the GPU");
  hostvec hm = to_host(m), hd = to_host(d);
  {
    py::gil_scoped_release nogil;
    if (forward) op.forward(add, hm, hd);
    else op.adjoint(add, hm, hd);
  }
  if (forward) from_host(d, hd);
  else from_host(m, hm);
}
// in place: v <- A v (forward) or v <- A^H v
template <class Op>
void apply_inplace(Op& op, bool forward, py::handle vec) {
  Vec v = vec_of(vec);
  if (v.n != op.getDomainSize()) throw std::runtime_error("vector size does not match the operator");
  if (v.dev) {
    if (forward) op.cu_forward(v.dev);
    else op.cu_adjoint(v.dev);
    sync();
    return;
  }
  hostvec h = to_host(v);
  {
    py::gil_scoped_release nogil;
    if (forward) op.forward(h);
    else op.adjoint(h);
  }
  from_host(v, h);
}

// wavefield hypercube (NX, NY, NW, NS) [+ slowness hypercube (nx, ny, nw, nz)] -> fwix_dims
fwix_dims dims_of(py::handle wfld, py::handle slow) {
  Ax a[4];
  for (int i = 0; i < 4; ++i) a[i] = i < ndim_of(wfld) ? axis_of(wfld, i + 1) : Ax{1, 0.0, 1.0};
  fwix_dims d{};
  d.NX = a[0].n; d.NY = a[1].n; d.NW = a[2].n; d.NS = a[3].n;
  d.nx = d.NX; d.ny = d.NY;
  d.ox = (float)a[0].o; d.dx = (float)a[0].d; d.oy = (float)a[1].o; d.dy = (float)a[1].d;
  d.ow = (float)a[2].o; d.dw = (float)a[2].d;
  d.oz = 0.f; d.dz = 1.f; d.nz = 1;
  if (!slow.is_none()) {
    const Ax sx = axis_of(slow, 1), sy = axis_of(slow, 2), sw = axis_of(slow, 3), sz = axis_of(slow, 4);
    if (sw.n != d.NW) throw std::runtime_error("slowness and wavefield frequency axes differ");
    if (sx.n > d.NX || sy.n > d.NY) throw std::runtime_error("wavefield grid must be the slowness grid plus padding");
    d.nx = sx.n; d.ny = sy.n;
    d.oz = (float)sz.o; d.dz = (float)sz.d; d.nz = sz.n;
  }
  return d;
}
fwix_params params_of(py::handle par, bool taper) {
  fwix_params p;
  fwix_params_default(&p);
  p.nref = par_int(par, "nref", 1);
  p.taperx = taper ? par_int(par, "taperx", 0) : 0;
  p.tapery = taper ? par_int(par, "tapery", 0) : 0;
  p.eps = par_float(par, "eps", 0.f);
  p.eps_scat = par_float(par, "eps_scat", 0.04f);
  p.taylor = par_int(par, "taylor", 1);
  return p;
}

// Reference slownesses and labels of a slowness array (nz, nw, ny, nx) on the device.
struct Refs {
  DeviceVector sref;
  uint8_t* labels = nullptr;
  size_t nlab = 0;
  Refs() = default;
  Refs(const Refs&) = delete;
  ~Refs() {
    if (labels) (void)hipFree(labels);
  }
  void build(py::handle slow_arr, int nref, int padx, int pady) {
    py::object sel = py::module_::import("pyfwix_amd.refsampler").attr("select_refs");
    py::tuple r = sel(slow_arr, nref, padx, pady);
    auto s = py::array_t<std::complex<float>, py::array::c_style | py::array::forcecast>(r[0]);
    auto l = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>(r[1]);
    sref.resize((size_t)s.size());
    fwix::hip_ok(hipMemcpy(sref.data(), s.data(), (size_t)s.size() * sizeof(cfloat), hipMemcpyHostToDevice), "refs");
    if (labels) (void)hipFree(labels);
    labels = nullptr;
    nlab = (size_t)l.size();
    fwix::hip_ok(hipMalloc(reinterpret_cast<void**>(&labels), nlab), "hipMalloc labels");
    fwix::hip_ok(hipMemcpy(labels, l.data(), nlab, hipMemcpyHostToDevice), "labels");
  }
};
py::array host_array(py::handle vec) {
  py::object arr = vec.attr("getNdArray")();
  if (py::hasattr(arr, "is_cuda")) arr = arr.attr("cpu")().attr("numpy")();
  return py::array_t<std::complex<float>, py::array::c_style | py::array::forcecast>(arr);
}

// ------------------------------------------------------------------ operator classes
// operator/src/FFT.cpp:5-69 (pyCudaOperator.cpp:15-30)
struct PyFFT {
  std::unique_ptr<fwix::cuFFT2d> op;
  explicit PyFFT(py::handle hyper) { op = std::make_unique<fwix::cuFFT2d>(dims_of(hyper, py::none())); }
};

// propagator/src/OneStep.h:11-111: PSPI / NSPS(domain, slow, par) + set_depth
template <int METHOD>
struct PyOneStep {
  Refs refs;
  std::unique_ptr<fwix::OneStep> op;
  PyOneStep(py::handle hyper, py::handle slow, py::handle par) {
    const fwix_dims d = dims_of(hyper, slow.attr("getHyper")());
    const fwix_params p = params_of(par, METHOD == FWIX_PSPI);
    refs.build(host_array(slow), p.nref, d.NX - d.nx, d.NY - d.ny);
    op = std::make_unique<fwix::OneStep>(METHOD, d, p);
    op->set_refs(refs.sref.data(), refs.labels);
  }
};

// propagator/src/OneWay.h:14-245: Downward / Upward(domain, slow, par)
template <int DIR>
struct PyOneWay {
  Refs refs;
  std::unique_ptr<fwix::OneWay> op;
  PyOneWay(py::handle hyper, py::handle slow, py::handle par) {
    const fwix_dims d = dims_of(hyper, slow.attr("getHyper")());
    const fwix_params p = params_of(par, true);
    refs.build(host_array(slow), p.nref, d.NX - d.nx, d.NY - d.ny);
    op = std::make_unique<fwix::OneWay>(DIR, d, p);
    op->set_refs(refs.sref.data(), refs.labels);
  }
};

// propagator/src/Injection.h:10-55: Injection(traces hyper, wavefield hyper, oz, dz, cx, cy, cz, ids)
struct PyInjection {
  std::unique_ptr<fwix::Injection> op;
  PyInjection(py::handle thyper, py::handle whyper, float oz, float dz, const std::vector<float>& cx,
              const std::vector<float>& cy, const std::vector<float>& cz, const std::vector<int>& ids) {
    fwix_dims d = dims_of(whyper, py::none());
    d.oz = oz;
    d.dz = dz;
    if (axis_of(thyper, 1).n != d.NW) throw std::runtime_error("trace axis 1 must be the wavefield's frequency axis");
    if (cy.size() != cx.size() || cz.size() != cx.size() || ids.size() != cx.size())
      throw std::runtime_error("coordinate lists differ in length");
    op = std::make_unique<fwix::Injection>(d, cx, cy, cz, ids);
  }
};

// ------------------------------------------------------------------ streaming modeling
// Source batches x frequency bands on ginsu windows of the model
// (propagator/src/StreamingPropagator.cpp:68-170 index maps, :327-401 windows,
// :438-464 accumulation); the batches run one after another on one GPU.
int next_fft_size(int n) {
  for (int m = std::max(8, n); m <= 1024; ++m)
    if (fwix_fft_size_supported(m)) return m;
  throw std::runtime_error("no supported FFT size >= " + std::to_string(n) + " (largest is 1024)");
}

struct Batch {
  std::vector<int> src, rec;  // trace indices
  int f0 = 0, nf = 0, x0 = 0, y0 = 0, wx = 0, wy = 0;
  fwix_dims dims{};
  std::unique_ptr<fwix::ExtendedBorn> op;
  DeviceVector wav, slow, den;
  Refs refs;
  bool have_bg = false;
};

struct PyStreaming {
  int nx = 0, ny = 0, nw = 0, nz = 0;
  double ox = 0, dx = 1, oy = 0, dy = 1, ow = 0, dw = 1, oz = 0, dz = 1;
  int nrec = 0, nref = 1;
  fwix_params params{};
  std::vector<std::unique_ptr<Batch>> batches;

  PyStreaming(py::handle wav_hyper, py::handle data_hyper, py::handle slow_hyper, py::handle wavelet,
              const std::vector<float>& sx, const std::vector<float>& sy, const std::vector<float>& sz,
              const std::vector<int>& s_ids, const std::vector<float>& rx, const std::vector<float>& ry,
              const std::vector<float>& rz, const std::vector<int>& r_ids, py::handle par,
              std::vector<int> nbatches, bool whole_model) {
    const Ax a1 = axis_of(slow_hyper, 1), a2 = axis_of(slow_hyper, 2), a3 = axis_of(slow_hyper, 3),
             a4 = axis_of(slow_hyper, 4);
    nx = a1.n; ny = a2.n; nw = a3.n; nz = a4.n;
    ox = a1.o; dx = a1.d; oy = a2.o; dy = a2.d; ow = a3.o; dw = a3.d; oz = a4.o; dz = a4.d;
    const int padx = par_int(par, "padx", -1), pady = par_int(par, "pady", -1);
    if (padx < 0 || pady < 0) throw std::runtime_error("Error: negative padx/pady values");
    params = params_of(par, true);
    nref = params.nref;
    const int ckpt = par_int(par, "ckpt_every", 1);
    const double gx = par_float(par, "ginsu_x", 0.f), gy = par_float(par, "ginsu_y", 0.f);
    nrec = (int)rx.size();
    if (axis_of(data_hyper, 1).n != nw || axis_of(wav_hyper, 1).n != nw)
      throw std::runtime_error("data and wavelet need the model's frequency axis");
    if (sy.size() != sx.size() || sz.size() != sx.size() || s_ids.size() != sx.size() || ry.size() != rx.size() ||
        rz.size() != rx.size() || r_ids.size() != rx.size())
      throw std::runtime_error("coordinate lists differ in length");
    if (nbatches.size() < 2) nbatches.resize(2, 1);
    // source batches: unique ids sorted, split evenly, the first batches taking the remainder
    std::vector<int> uniq(s_ids);
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    for (int id : r_ids)
      if (!std::binary_search(uniq.begin(), uniq.end(), id))
        throw std::runtime_error("Error: Found a receiver associated with a Shot ID not present in the source list.");
    const int nsb = std::max(1, std::min(nbatches[0], (int)uniq.size()));
    const int nfb = std::max(1, std::min(nbatches[1], nw));
    const Vec wv = vec_of(wavelet);  // (nsrc, nw)
    hostvec w_all(wv.n);
    if (wv.dev) fwix::hip_ok(hipMemcpy(w_all.data(), wv.dev, wv.n * sizeof(cfloat), hipMemcpyDeviceToHost), "wavelet");
    else w_all = to_host(wv);
    if (w_all.size() != sx.size() * (size_t)nw) throw std::runtime_error("wavelet must hold (nsrc, nw) samples");
    size_t k = 0;
    for (int b = 0; b < nsb; ++b) {
      const size_t cnt = uniq.size() / nsb + (b < (int)(uniq.size() % nsb) ? 1 : 0);
      std::vector<int> ids(uniq.begin() + k, uniq.begin() + k + cnt);
      k += cnt;
      std::vector<int> si, ri;
      for (int t = 0; t < (int)s_ids.size(); ++t)
        if (std::binary_search(ids.begin(), ids.end(), s_ids[t])) si.push_back(t);
      for (int t = 0; t < (int)r_ids.size(); ++t)
        if (std::binary_search(ids.begin(), ids.end(), r_ids[t])) ri.push_back(t);
      // ginsu window of the batch's sources and receivers (whole model for the plain Propagator)
      int x0 = 0, x1 = nx - 1, y0 = 0, y1 = ny - 1;
      if (!whole_model) {
        float mnx = 1e30f, mxx = -1e30f, mny = 1e30f, mxy = -1e30f;
        for (int t : si) { mnx = std::min(mnx, sx[t]); mxx = std::max(mxx, sx[t]); mny = std::min(mny, sy[t]); mxy = std::max(mxy, sy[t]); }
        for (int t : ri) { mnx = std::min(mnx, rx[t]); mxx = std::max(mxx, rx[t]); mny = std::min(mny, ry[t]); mxy = std::max(mxy, ry[t]); }
        x0 = std::max(0, (int)std::floor(((double)mnx - gx - ox) / dx));
        x1 = std::min(nx - 1, (int)std::ceil(((double)mxx + gx - ox) / dx));
        y0 = std::max(0, (int)std::floor(((double)mny - gy - oy) / dy));
        y1 = std::min(ny - 1, (int)std::ceil(((double)mxy + gy - oy) / dy));
        if (x1 <= x0) { x0 = 0; x1 = nx - 1; }
        if (y1 <= y0) { y0 = 0; y1 = ny - 1; }
      }
      // dense shot ids of the batch
      std::vector<int> sd(si.size()), rd(ri.size());
      for (size_t t = 0; t < si.size(); ++t)
        sd[t] = (int)(std::lower_bound(ids.begin(), ids.end(), s_ids[si[t]]) - ids.begin());
      for (size_t t = 0; t < ri.size(); ++t)
        rd[t] = (int)(std::lower_bound(ids.begin(), ids.end(), r_ids[ri[t]]) - ids.begin());
      int f0 = 0;
      for (int fb = 0; fb < nfb; ++fb) {
        const int nf = nw / nfb + (fb < nw % nfb ? 1 : 0);
        auto B = std::make_unique<Batch>();
        B->src = si; B->rec = ri; B->f0 = f0; B->nf = nf;
        B->x0 = x0; B->y0 = y0; B->wx = x1 - x0 + 1; B->wy = y1 - y0 + 1;
        fwix_dims& d = B->dims;
        d.NX = next_fft_size(B->wx + padx); d.NY = next_fft_size(B->wy + pady);
        d.NW = nf; d.NS = (int)ids.size(); d.nx = B->wx; d.ny = B->wy;
        d.ox = (float)(ox + x0 * dx); d.dx = (float)dx; d.oy = (float)(oy + y0 * dy); d.dy = (float)dy;
        d.ow = (float)(ow + f0 * dw); d.dw = (float)dw; d.oz = (float)oz; d.dz = (float)dz; d.nz = nz;
        B->op = std::make_unique<fwix::ExtendedBorn>(d, params, (int)ri.size(), ckpt);
        auto pick = [](const std::vector<float>& v, const std::vector<int>& idx) {
          std::vector<float> o(idx.size());
          for (size_t t = 0; t < idx.size(); ++t) o[t] = v[idx[t]];
          return o;
        };
        B->op->set_sources(pick(sx, si), pick(sy, si), pick(sz, si), sd);
        B->op->set_receivers(pick(rx, ri), pick(ry, ri), pick(rz, ri), rd);
        hostvec w(si.size() * (size_t)nf);
        for (size_t t = 0; t < si.size(); ++t)
          for (int f = 0; f < nf; ++f) w[t * nf + f] = w_all[(size_t)si[t] * nw + f0 + f];
        B->wav.resize(w.size());
        B->wav.upload(w, nullptr);
        B->op->set_wavelet(B->wav.data());
        batches.push_back(std::move(B));
        f0 += nf;
      }
    }
  }

  // window (nz, nw, ny, nx) -> (nz, nf, wy, wx) of a batch
  static hostvec window(const Batch& b, const std::complex<float>* a, int nw, int ny, int nx, int nz) {
    hostvec o((size_t)nz * b.nf * b.wy * b.wx);
    size_t k = 0;
    for (int z = 0; z < nz; ++z)
      for (int f = 0; f < b.nf; ++f)
        for (int y = 0; y < b.wy; ++y) {
          const std::complex<float>* row = a + (((size_t)z * nw + b.f0 + f) * ny + b.y0 + y) * nx + b.x0;
          std::copy(row, row + b.wx, o.begin() + k);
          k += b.wx;
        }
    return o;
  }

  // background model [slow, den] (host or device vectors) for every batch, then the down and up passes
  void set_background(const std::vector<py::object>& model) {
    if (model.empty()) throw std::runtime_error("model needs the slowness vector (and optionally the density)");
    py::array s = host_array(model[0]);
    py::array dn;
    if (model.size() > 1) dn = host_array(model[1]);
    const size_t n = (size_t)nz * nw * ny * nx;
    if ((size_t)s.size() != n || (dn && (size_t)dn.size() != n))
      throw std::runtime_error("model vectors must have the slowness hypercube (nx, ny, nw, nz)");
    const auto* sp = static_cast<const std::complex<float>*>(s.data());
    hostvec ones;
    if (!dn) ones.assign(n, cfloat(1.f, 0.f));
    const auto* dp = dn ? static_cast<const std::complex<float>*>(dn.data()) : ones.data();
    for (auto& bp : batches) {
      Batch& b = *bp;
      hostvec ws = window(b, sp, nw, ny, nx, nz), wd = window(b, dp, nw, ny, nx, nz);
      py::array_t<std::complex<float>> wa(std::vector<py::ssize_t>{nz, b.nf, b.wy, b.wx});
      std::memcpy(wa.mutable_data(), ws.data(), ws.size() * sizeof(cfloat));
      b.refs.build(wa, nref, b.dims.NX - b.dims.nx, b.dims.NY - b.dims.ny);
      b.slow.resize(ws.size());
      b.slow.upload(ws, nullptr);
      b.den.resize(wd.size());
      b.den.upload(wd, nullptr);
      b.op->set_refs(b.refs.sref.data(), b.refs.labels);
      b.op->set_slow(b.slow.data());
      b.op->set_density(b.den.data());
      b.have_bg = false;
    }
  }

  // data (nrec, nw) on the host: data (+)= direct + reflected recordings
  void forward(bool add, const std::vector<py::object>& model, py::handle data) {
    set_background(model);
    Vec d = vec_of(data);
    hostvec hd = d.dev ? hostvec(d.n) : to_host(d);
    if (d.dev) fwix::hip_ok(hipMemcpy(hd.data(), d.dev, d.n * sizeof(cfloat), hipMemcpyDeviceToHost), "data D2H");
    if (hd.size() != (size_t)nrec * nw) throw std::runtime_error("data must hold (nrec, nw) samples");
    if (!add) std::fill(hd.begin(), hd.end(), cfloat(0.f, 0.f));
    for (auto& bp : batches) {
      Batch& b = *bp;
      DeviceVector db(b.rec.size() * (size_t)b.nf);
      b.op->model(db.data(), false);
      b.op->model_up(db.data(), true);
      b.have_bg = true;
      hostvec out;
      db.download(out, nullptr);
      for (size_t t = 0; t < b.rec.size(); ++t)
        for (int f = 0; f < b.nf; ++f) hd[(size_t)b.rec[t] * nw + b.f0 + f] += out[t * b.nf + f];
    }
    if (d.dev) fwix::hip_ok(hipMemcpy(d.dev, hd.data(), d.n * sizeof(cfloat), hipMemcpyHostToDevice), "data H2D");
    else from_host(d, hd);
  }

  void model_background() {
    for (auto& bp : batches)
      if (!bp->have_bg) {
        bp->op->model(nullptr, false);
        bp->op->model_up(nullptr, false);
        bp->have_bg = true;
      }
  }

  // ExtendedBorn (all four channels) over the batches; model [ds, drho]
  // (nz, nw, ny, nx) each, data (nrec, nw)
  void born(bool forward, bool add, const std::vector<py::object>& model, py::handle data) {
    model_background();
    if (model.empty()) throw std::runtime_error("model needs [ds] or [ds, drho]");
    const size_t n = (size_t)nz * nw * ny * nx;
    std::vector<Vec> mv;
    std::vector<hostvec> mh;
    for (size_t i = 0; i < std::min<size_t>(model.size(), 2); ++i) {
      mv.push_back(vec_of(model[i]));
      if (mv.back().n != n) throw std::runtime_error("model vectors must have the slowness hypercube");
      hostvec h(n);
      if (mv.back().dev) fwix::hip_ok(hipMemcpy(h.data(), mv.back().dev, n * sizeof(cfloat), hipMemcpyDeviceToHost), "D2H");
      else h = to_host(mv.back());
      mh.push_back(std::move(h));
    }
    if (mh.size() < 2) mh.emplace_back(n, cfloat(0.f, 0.f));
    Vec d = vec_of(data);
    if (d.n != (size_t)nrec * nw) throw std::runtime_error("data must hold (nrec, nw) samples");
    hostvec hd(d.n);
    if (d.dev) fwix::hip_ok(hipMemcpy(hd.data(), d.dev, d.n * sizeof(cfloat), hipMemcpyDeviceToHost), "data D2H");
    else hd = to_host(d);
    if (forward && !add) std::fill(hd.begin(), hd.end(), cfloat(0.f, 0.f));
    if (!forward && !add)
      for (auto& h : mh) std::fill(h.begin(), h.end(), cfloat(0.f, 0.f));
    for (auto& bp : batches) {
      Batch& b = *bp;
      const size_t mw = (size_t)nz * b.nf * b.wy * b.wx, nd = b.rec.size() * (size_t)b.nf;
      DeviceVector ds(mw), dr(mw), db(nd);
      if (forward) {
        ds.upload(window(b, mh[0].data(), nw, ny, nx, nz), nullptr);
        dr.upload(window(b, mh[1].data(), nw, ny, nx, nz), nullptr);
        b.op->extended_forward(false, ds.data(), dr.data(), db.data());
        hostvec out;
        db.download(out, nullptr);
        for (size_t t = 0; t < b.rec.size(); ++t)
          for (int f = 0; f < b.nf; ++f) hd[(size_t)b.rec[t] * nw + b.f0 + f] += out[t * b.nf + f];
      } else {
        hostvec in(nd);
        for (size_t t = 0; t < b.rec.size(); ++t)
          for (int f = 0; f < b.nf; ++f) in[t * b.nf + f] = hd[(size_t)b.rec[t] * nw + b.f0 + f];
        db.upload(in, nullptr);
        b.op->extended_adjoint(false, ds.data(), dr.data(), db.data());
        hostvec gs, gr;
        ds.download(gs, nullptr);
        dr.download(gr, nullptr);
        for (int part = 0; part < 2; ++part) {
          const hostvec& g = part ? gr : gs;
          hostvec& out = mh[part];
          size_t k = 0;
          for (int z = 0; z < nz; ++z)
            for (int f = 0; f < b.nf; ++f)
              for (int y = 0; y < b.wy; ++y)
                for (int x = 0; x < b.wx; ++x)
                  out[(((size_t)z * nw + b.f0 + f) * ny + b.y0 + y) * nx + b.x0 + x] += g[k++];
        }
      }
    }
    if (forward) {
      if (d.dev) fwix::hip_ok(hipMemcpy(d.dev, hd.data(), d.n * sizeof(cfloat), hipMemcpyHostToDevice), "H2D");
      else from_host(d, hd);
    } else {
      for (size_t i = 0; i < mv.size(); ++i) {
        if (mv[i].dev) fwix::hip_ok(hipMemcpy(mv[i].dev, mh[i].data(), n * sizeof(cfloat), hipMemcpyHostToDevice), "H2D");
        else from_host(mv[i], mh[i]);
      }
    }
  }
};

// propagator/src/Propagator.cpp: one batch over the whole model
struct PyPropagator : PyStreaming {
  using PyStreaming::PyStreaming;
  double compression_ratio() const { return 1.0; }  // checkpoints are stored losslessly
};

// ExtendedBorn(model hyper, data hyper, [slow, den], propagator) and
// StreamingExtendedBorn(propagator, [slow, den]): the background is the
// propagator's batches on [slow, den]
struct PyExtendedBorn {
  std::shared_ptr<PyStreaming> prop;
  PyExtendedBorn(std::shared_ptr<PyStreaming> p, const std::vector<py::object>& slow_den) : prop(std::move(p)) {
    prop->set_background(slow_den);
    prop->model_background();
  }
};
struct PyStreamingExtendedBorn : PyExtendedBorn {
  using PyExtendedBorn::PyExtendedBorn;
};

}  // namespace

// operator/src/python/pyCudaOperator.cpp:15-30
static void init_operator(py::module_& m) {
  py::class_<PyFFT, std::shared_ptr<PyFFT>>(m, "cuFFT2d")
      .def(py::init<py::object>(), "Initialize cuFFT2d")
      .def("forward", [](PyFFT& s, bool add, py::object mo, py::object d) { apply(*s.op, true, add, mo, d); },
           "Forward operator of cuFFT2d")
      .def("adjoint", [](PyFFT& s, bool add, py::object mo, py::object d) { apply(*s.op, false, add, mo, d); },
           "Adjoint operator of cuFFT2d");
}

template <class T>
static void bind_onestep(py::module_& m, const char* name, bool one_arg_forward) {
  auto c = py::class_<T, std::shared_ptr<T>>(m, name)
               .def(py::init<py::object, py::object, py::object>())
               .def("forward", [](T& s, bool add, py::object mo, py::object d) { apply(*s.op, true, add, mo, d); })
               .def("adjoint", [](T& s, bool add, py::object mo, py::object d) { apply(*s.op, false, add, mo, d); })
               .def("set_depth", [](T& s, int iz) { s.op->set_depth(iz); });
  if (one_arg_forward) c.def("forward", [](T& s, py::object v) { apply_inplace(*s.op, true, v); });
}

template <class T>
static void bind_oneway(py::module_& m, const char* name, bool one_arg) {
  auto c = py::class_<T, std::shared_ptr<T>>(m, name)
               .def(py::init<py::object, py::object, py::object>())
               .def("forward", [](T& s, bool add, py::object mo, py::object d) { apply(*s.op, true, add, mo, d); })
               .def("adjoint", [](T& s, bool add, py::object mo, py::object d) { apply(*s.op, false, add, mo, d); });
  if (one_arg) {
    c.def("forward", [](T& s, py::object v) { apply_inplace(*s.op, true, v); });
    c.def("adjoint", [](T& s, py::object v) { apply_inplace(*s.op, false, v); });
  }
}

// propagator/src/python/pyCudaWEM.cpp:22-242
static void init_propagator(py::module_& m) {
  using Geo = const std::vector<float>&;
  using Ids = const std::vector<int>&;
  py::class_<PyStreaming, std::shared_ptr<PyStreaming>>(m, "StreamingPropagator")
      .def(py::init([](py::object wh, py::object dh, py::object sh, py::object wav, Geo sx, Geo sy, Geo sz, Ids sid,
                       Geo rx, Geo ry, Geo rz, Ids rid, py::object par, std::vector<int> nb) {
             return std::make_shared<PyStreaming>(wh, dh, sh, wav, sx, sy, sz, sid, rx, ry, rz, rid, par, nb, false);
           }),
           "Initialize StreamingPropagator")
      .def("forward", [](PyStreaming& s, bool add, std::vector<py::object> mo, py::object d) { s.forward(add, mo, d); },
           "Nonlinear forward operator of StreamingPropagator")
      .def("get_compression_ratio", [](PyStreaming&) { return 1.0; });
  py::class_<PyPropagator, PyStreaming, std::shared_ptr<PyPropagator>>(m, "Propagator")
      .def(py::init([](py::object wh, py::object dh, py::object sh, py::object wav, Geo sx, Geo sy, Geo sz, Ids sid,
                       Geo rx, Geo ry, Geo rz, Ids rid, py::object par) {
             return std::make_shared<PyPropagator>(wh, dh, sh, wav, sx, sy, sz, sid, rx, ry, rz, rid, par,
                                                   std::vector<int>{1, 1}, true);
           }),
           "Initialize Propagator")
      .def("forward", [](PyPropagator& s, bool add, std::vector<py::object> mo, py::object d) { s.forward(add, mo, d); },
           "Nonlinear forward operator of Propagator")
      .def("get_compression_ratio", &PyPropagator::compression_ratio, "Get compression ratio of Propagator");
  py::class_<PyExtendedBorn, std::shared_ptr<PyExtendedBorn>>(m, "ExtendedBorn")
      .def(py::init([](py::object, py::object, std::vector<py::object> sd, std::shared_ptr<PyPropagator> p) {
             return std::make_shared<PyExtendedBorn>(p, sd);
           }),
           "Initialize ExtendedBorn")
      .def("forward", [](PyExtendedBorn& s, bool add, std::vector<py::object> mo, py::object d) {
        s.prop->born(true, add, mo, d);
      })
      .def("adjoint", [](PyExtendedBorn& s, bool add, std::vector<py::object> mo, py::object d) {
        s.prop->born(false, add, mo, d);
      });
  py::class_<PyStreamingExtendedBorn, std::shared_ptr<PyStreamingExtendedBorn>>(m, "StreamingExtendedBorn")
      .def(py::init([](std::shared_ptr<PyStreaming> p, std::vector<py::object> sd) {
             return std::make_shared<PyStreamingExtendedBorn>(p, sd);
           }),
           "Initialize StreamingExtendedBorn")
      .def("forward", [](PyStreamingExtendedBorn& s, bool add, std::vector<py::object> mo, py::object d) {
        s.prop->born(true, add, mo, d);
      })
      .def("adjoint", [](PyStreamingExtendedBorn& s, bool add, std::vector<py::object> mo, py::object d) {
        s.prop->born(false, add, mo, d);
      });
  bind_onestep<PyOneStep<FWIX_PSPI>>(m, "PSPI", true);
  bind_onestep<PyOneStep<FWIX_NSPS>>(m, "NSPS", false);
  py::class_<PyInjection, std::shared_ptr<PyInjection>>(m, "Injection")
      .def(py::init<py::object, py::object, float, float, Geo, Geo, Geo, Ids>(), "Initialize Injection")
      .def("forward", [](PyInjection& s, bool add, py::object tr, py::object w) { apply(*s.op, true, add, tr, w); })
      .def("adjoint", [](PyInjection& s, bool add, py::object tr, py::object w) { apply(*s.op, false, add, tr, w); })
      .def("set_depth", [](PyInjection& s, int iz) { s.op->set_depth(iz); })
      .def("set_coords", [](PyInjection& s, Geo cx, Geo cy, Geo cz, Ids ids) { s.op->set_coords(cx, cy, cz, ids); });
  bind_oneway<PyOneWay<0>>(m, "Downward", true);
  bind_oneway<PyOneWay<1>>(m, "Upward", false);
}

PYBIND11_MODULE(_pyFWIX, m) {
  m.doc() = "pyfwix operators on MI355X (HIP, gfx950) behind the reference's pyFWIX classes";
  init_operator(m);
  init_propagator(m);
}
