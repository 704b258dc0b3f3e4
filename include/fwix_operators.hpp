// fwix_operators.hpp -- C++ operator layer over the C ABI (fwix_hip.h).
//
// The class names, constructor arguments, add semantics and host/device
// method pairs of the reference's operator API (operator/src/CudaOperator.h,
// operator/src/FFT.h, propagator/src/{PhaseShift,Selector,Taper,OneStep,
// Injection,Scatter,ImagingCondition,OneWay}.h, propagator/src/ExtendedBorn.h),
// so that code written against those classes keeps its shape:
//
//   * an operator borrows the device buffers it is given and allocates its
//     own model_vec / data_vec otherwise (CudaOperator.h:29-53);
//   * cu_forward / cu_adjoint(add, model, data) work on device buffers,
//     add = false overwrites, add = true accumulates (CudaOperator.h:61-72);
//   * forward / adjoint(add, model, data) on host vectors copy through
//     model_vec / data_vec (CudaOperator.h:74-120); the one-argument forms
//     propagate in place;
//   * dotTest() returns |Re(conj(<d, A m>) / <m, A^H d>) - 1| for add = false
//     and add = true (CudaOperator.h:190-225).
//
// The SEP hypercube is replaced by fwix_dims and the paramObj by
// fwix_params; host vectors are std::vector<std::complex<float>> laid out
// x fastest.  Errors throw std::runtime_error (never exit()).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fwix_hip.h"

namespace fwix {

using cfloat = std::complex<float>;
using hostvec = std::vector<cfloat>;

inline void check(int rc) {
  if (rc) throw std::runtime_error(fwix_last_error());
}
inline void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer of complex64 samples (the reference's complex_vector).
class DeviceVector {
 public:
  explicit DeviceVector(size_t n = 0) { resize(n); }
  DeviceVector(const DeviceVector&) = delete;
  DeviceVector& operator=(const DeviceVector&) = delete;
  ~DeviceVector() { release(); }
  void resize(size_t n) {
    release();
    n_ = n;
    if (n) hip_ok(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(fwix_c64)), "hipMalloc");
  }
  fwix_c64* data() { return p_; }
  const fwix_c64* data() const { return p_; }
  size_t size() const { return n_; }
  // returns once the host vector has been read (it may be released right after)
  void upload(const hostvec& h, hipStream_t s) {
    if (h.size() != n_) throw std::runtime_error("host vector size does not match the operator");
    hip_ok(hipMemcpyAsync(p_, h.data(), n_ * sizeof(fwix_c64), hipMemcpyHostToDevice, s), "upload");
    hip_ok(hipStreamSynchronize(s), "synchronize");
  }
  void download(hostvec& h, hipStream_t s) const {
    h.resize(n_);
    hip_ok(hipMemcpyAsync(h.data(), p_, n_ * sizeof(fwix_c64), hipMemcpyDeviceToHost, s), "download");
    hip_ok(hipStreamSynchronize(s), "synchronize");
  }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  fwix_c64* p_ = nullptr;
  size_t n_ = 0;
};

// sum(conj(a) * b) in double precision
inline std::complex<double> dot(const hostvec& a, const hostvec& b) {
  std::complex<double> s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += std::conj(std::complex<double>(a[i])) * std::complex<double>(b[i]);
  return s;
}

inline size_t wavefield_size(const fwix_dims& d) { return (size_t)d.NX * d.NY * d.NW * d.NS; }
inline size_t model_size(const fwix_dims& d) { return (size_t)d.nz * d.NW * d.ny * d.nx; }

class CudaOperator {
 public:
  CudaOperator(size_t domain, size_t range, DeviceVector* model, DeviceVector* data, hipStream_t stream)
      : stream_(stream), own_m_(model ? 0 : domain), own_d_(data ? 0 : range),
        model_vec(model ? model : &own_m_), data_vec(data ? data : &own_d_), domain_(domain), range_(range) {
    if (model_vec->size() != domain || data_vec->size() != range)
      throw std::runtime_error("borrowed vectors do not match the operator's domain / range");
  }
  virtual ~CudaOperator() = default;

  virtual void cu_forward(bool add, const fwix_c64* model, fwix_c64* data) = 0;
  virtual void cu_adjoint(bool add, fwix_c64* model, const fwix_c64* data) = 0;
  virtual void cu_forward(fwix_c64*) { throw std::runtime_error("cu_forward not implemented in the derived class."); }
  virtual void cu_adjoint(fwix_c64*) { throw std::runtime_error("cu_adjoint not implemented in the derived class."); }

  void forward(bool add, const hostvec& model, hostvec& data) {
    if (add) data_vec->upload(data, stream_);
    model_vec->upload(model, stream_);
    cu_forward(add, model_vec->data(), data_vec->data());
    data_vec->download(data, stream_);
  }
  void adjoint(bool add, hostvec& model, const hostvec& data) {
    if (add) model_vec->upload(model, stream_);
    data_vec->upload(data, stream_);
    cu_adjoint(add, model_vec->data(), data_vec->data());
    model_vec->download(model, stream_);
  }
  // in place: data <- A data (square operators)
  void forward(hostvec& data) {
    data_vec->upload(data, stream_);
    cu_forward(data_vec->data());
    data_vec->download(data, stream_);
  }
  void adjoint(hostvec& model) {
    model_vec->upload(model, stream_);
    cu_adjoint(model_vec->data());
    model_vec->download(model, stream_);
  }

  std::pair<double, double> dotTest(bool verbose = false, unsigned seed = 0) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    hostvec m(domain_), d(range_), m1(domain_), d1(range_);
    for (auto& v : m) v = cfloat(u(gen), u(gen));
    for (auto& v : d) v = cfloat(u(gen), u(gen));
    double err[2];
    for (int k = 0; k < 2; ++k) {
      const bool add = k == 1;
      forward(add, m, d1);
      adjoint(add, m1, d);
      const std::complex<double> lhs = std::conj(dot(d, d1)), rhs = dot(m, m1);
      err[k] = std::abs((lhs / rhs).real() - 1.0);
      if (verbose) std::printf("add=%d  <m,A'd> %.6e  <Am,d> %.6e  error %.3e\n", k, rhs.real(), lhs.real(), err[k]);
    }
    return {err[0], err[1]};
  }

  size_t getDomainSize() const { return domain_; }
  size_t getRangeSize() const { return range_; }

 protected:
  hipStream_t stream_;
  DeviceVector own_m_, own_d_;

 public:
  DeviceVector* model_vec;
  DeviceVector* data_vec;

 private:
  size_t domain_, range_;
};

// Owns one C-ABI handle.
class Handle {
 public:
  Handle() = default;
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  ~Handle() {
    if (h) fwix_op_destroy(h);
  }
  fwix_op h = nullptr;
};

// operator/src/FFT.cpp:5-69
class cuFFT2d : public CudaOperator {
 public:
  cuFFT2d(const fwix_dims& d, DeviceVector* model = nullptr, DeviceVector* data = nullptr, hipStream_t s = nullptr)
      : CudaOperator(wavefield_size(d), wavefield_size(d), model, data, s) {
    check(fwix_fft2d_create(&d, &op_.h));
  }
  using CudaOperator::cu_adjoint;
  using CudaOperator::cu_forward;
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override { check(fwix_fft2d_forward(op_.h, add, m, d, stream_)); }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_fft2d_adjoint(op_.h, add, m, d, stream_)); }

 private:
  Handle op_;
};

// propagator/src/PhaseShift.h:11-72
class PhaseShift : public CudaOperator {
 public:
  PhaseShift(const fwix_dims& d, float dz, float eps = 0.f, DeviceVector* model = nullptr, DeviceVector* data = nullptr,
             hipStream_t s = nullptr)
      : CudaOperator(wavefield_size(d), wavefield_size(d), model, data, s) {
    check(fwix_phase_shift_create(&d, dz, eps, &op_.h));
  }
  void set_slow(const std::vector<cfloat>& sref) {  // one reference slowness per frequency
    check(fwix_phase_shift_set_slow(op_.h, reinterpret_cast<const fwix_c64*>(sref.data())));
  }
  using CudaOperator::cu_adjoint;
  using CudaOperator::cu_forward;
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override { check(fwix_phase_shift_forward(op_.h, add, m, d, stream_)); }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_phase_shift_adjoint(op_.h, add, m, d, stream_)); }
  void cu_inverse(bool add, fwix_c64* m, const fwix_c64* d) { check(fwix_phase_shift_inverse(op_.h, add, m, d, stream_)); }

 private:
  Handle op_;
};

// propagator/src/Selector.h:8-58
class Selector : public CudaOperator {
 public:
  Selector(const fwix_dims& d, DeviceVector* model = nullptr, DeviceVector* data = nullptr, hipStream_t s = nullptr)
      : CudaOperator(wavefield_size(d), wavefield_size(d), model, data, s) {
    check(fwix_selector_create(&d, &op_.h));
  }
  void set_labels(const std::vector<int32_t>& labels) { check(fwix_selector_set_labels(op_.h, labels.data())); }
  void set_value(int v) { check(fwix_selector_set_value(op_.h, v)); }
  using CudaOperator::cu_adjoint;
  using CudaOperator::cu_forward;
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override { check(fwix_selector_forward(op_.h, add, m, d, stream_)); }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_selector_adjoint(op_.h, add, m, d, stream_)); }

 private:
  Handle op_;
};

// propagator/src/Taper.h:8-42
class Taper : public CudaOperator {
 public:
  Taper(const fwix_dims& d, const fwix_params& p, DeviceVector* model = nullptr, DeviceVector* data = nullptr,
        hipStream_t s = nullptr)
      : CudaOperator(wavefield_size(d), wavefield_size(d), model, data, s) {
    check(fwix_taper_create(&d, p.taperx, p.tapery, &op_.h));
  }
  using CudaOperator::cu_adjoint;
  using CudaOperator::cu_forward;
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override { check(fwix_taper_forward(op_.h, add, m, d, stream_)); }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_taper_adjoint(op_.h, add, m, d, stream_)); }

 private:
  Handle op_;
};

// propagator/src/OneStep.h:11-111 (PSPI, NSPS): one depth step, in place or not.
// set_refs borrows device arrays sref (nz, nref, NW) and labels (nz, NW, NY, NX) uint8.
class OneStep : public CudaOperator {
 public:
  OneStep(int method, const fwix_dims& d, const fwix_params& p, DeviceVector* model = nullptr,
          DeviceVector* data = nullptr, hipStream_t s = nullptr)
      : CudaOperator(wavefield_size(d), wavefield_size(d), model, data, s) {
    check(fwix_onestep_create(method, &d, &p, &op_.h));
  }
  void set_refs(const fwix_c64* sref_dev, const uint8_t* labels_dev) { check(fwix_onestep_set_refs(op_.h, sref_dev, labels_dev)); }
  void set_depth(int iz) { check(fwix_onestep_set_depth(op_.h, iz)); }
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override { check(fwix_onestep_forward(op_.h, add, m, d, stream_)); }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_onestep_adjoint(op_.h, add, m, d, stream_)); }
  void cu_forward(fwix_c64* m) override { check(fwix_onestep_forward(op_.h, 0, m, m, stream_)); }
  void cu_adjoint(fwix_c64* d) override { check(fwix_onestep_adjoint(op_.h, 0, d, d, stream_)); }

 private:
  Handle op_;
};
class PSPI : public OneStep {
 public:
  PSPI(const fwix_dims& d, const fwix_params& p, DeviceVector* model = nullptr, DeviceVector* data = nullptr,
       hipStream_t s = nullptr)
      : OneStep(FWIX_PSPI, d, p, model, data, s) {}
};
class NSPS : public OneStep {
 public:
  NSPS(const fwix_dims& d, const fwix_params& p, DeviceVector* model = nullptr, DeviceVector* data = nullptr,
       hipStream_t s = nullptr)
      : OneStep(FWIX_NSPS, d, p, model, data, s) {}
};

// propagator/src/Injection.h:10-55: model = traces (ntrace, NW), data = wavefield
class Injection : public CudaOperator {
 public:
  Injection(const fwix_dims& d, const std::vector<float>& cx, const std::vector<float>& cy, const std::vector<float>& cz,
            const std::vector<int>& ids, DeviceVector* model = nullptr, DeviceVector* data = nullptr,
            hipStream_t s = nullptr)
      : CudaOperator(cx.size() * (size_t)d.NW, wavefield_size(d), model, data, s) {
    check(fwix_injection_create(&d, (int)cx.size(), &op_.h));
    set_coords(cx, cy, cz, ids);
  }
  void set_coords(const std::vector<float>& cx, const std::vector<float>& cy, const std::vector<float>& cz,
                  const std::vector<int>& ids) {
    check(fwix_injection_set_coords(op_.h, cx.data(), cy.data(), cz.data(), ids.data()));
  }
  void set_depth(int iz) { check(fwix_injection_set_depth(op_.h, iz)); }
  using CudaOperator::cu_adjoint;
  using CudaOperator::cu_forward;
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override { check(fwix_injection_forward(op_.h, add, m, d, stream_)); }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_injection_adjoint(op_.h, add, m, d, stream_)); }

 private:
  Handle op_;
};

// propagator/src/Scatter.h:12-104: set_slow borrows the slowness (nz, NW, ny, nx) on the device
class Scatter : public CudaOperator {
 public:
  Scatter(const fwix_dims& d, const fwix_params& p, DeviceVector* model = nullptr, DeviceVector* data = nullptr,
          hipStream_t s = nullptr)
      : CudaOperator(wavefield_size(d), wavefield_size(d), model, data, s) {
    check(fwix_scatter_create(&d, &p, &op_.h));
  }
  void set_slow(const fwix_c64* slow_dev) { check(fwix_scatter_set_slow(op_.h, slow_dev)); }
  void set_depth(int iz) { check(fwix_scatter_set_depth(op_.h, iz)); }
  using CudaOperator::cu_adjoint;
  using CudaOperator::cu_forward;
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override { check(fwix_scatter_forward(op_.h, add, m, d, stream_)); }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_scatter_adjoint(op_.h, add, m, d, stream_)); }

 private:
  Handle op_;
};

// propagator/src/ImagingCondition.h:11-43: model = image (NW, ny, nx), data = wavefield
class ImagingCondition : public CudaOperator {
 public:
  ImagingCondition(const fwix_dims& d, DeviceVector* model = nullptr, DeviceVector* data = nullptr,
                   hipStream_t s = nullptr)
      : CudaOperator((size_t)d.NW * d.ny * d.nx, wavefield_size(d), model, data, s) {
    check(fwix_ic_create(&d, &op_.h));
  }
  void set_background(const fwix_c64* bg_dev) { check(fwix_ic_set_background(op_.h, bg_dev)); }
  using CudaOperator::cu_adjoint;
  using CudaOperator::cu_forward;
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override { check(fwix_ic_forward(op_.h, add, m, d, stream_)); }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_ic_adjoint(op_.h, add, m, d, stream_)); }

 private:
  Handle op_;
};

// propagator/src/OneWay.h:14-245: the depth recursion, down (0) or up (1)
class OneWay : public CudaOperator {
 public:
  OneWay(int direction, const fwix_dims& d, const fwix_params& p, DeviceVector* model = nullptr,
         DeviceVector* data = nullptr, hipStream_t s = nullptr)
      : CudaOperator(wavefield_size(d), wavefield_size(d), model, data, s) {
    check(fwix_oneway_create(direction, FWIX_PSPI, &d, &p, &op_.h));
  }
  void set_refs(const fwix_c64* sref_dev, const uint8_t* labels_dev) { check(fwix_oneway_set_refs(op_.h, sref_dev, labels_dev)); }
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override {
    check(fwix_oneway_forward(op_.h, add, m, d, nullptr, stream_));
  }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_oneway_adjoint(op_.h, add, m, d, stream_)); }
  void cu_forward(fwix_c64* m) override { check(fwix_oneway_forward(op_.h, 0, m, m, nullptr, stream_)); }
  void cu_adjoint(fwix_c64* d) override { check(fwix_oneway_adjoint(op_.h, 0, d, d, stream_)); }

 private:
  Handle op_;
};
class Downward : public OneWay {
 public:
  Downward(const fwix_dims& d, const fwix_params& p, DeviceVector* model = nullptr, DeviceVector* data = nullptr,
           hipStream_t s = nullptr)
      : OneWay(0, d, p, model, data, s) {}
};
class Upward : public OneWay {
 public:
  Upward(const fwix_dims& d, const fwix_params& p, DeviceVector* model = nullptr, DeviceVector* data = nullptr,
         hipStream_t s = nullptr)
      : OneWay(1, d, p, model, data, s) {}
};

// propagator/src/ExtendedBorn.h + Propagator.cpp: background modeling and the
// linearized operator.  cu_forward / cu_adjoint are the down-scattering
// channel with model = ds (nz, NW, ny, nx); extended_forward / _adjoint carry
// all four channels with model (ds, drho).  Data = receiver traces (nrec, NW).
// set_* borrow device arrays.
class ExtendedBorn : public CudaOperator {
 public:
  ExtendedBorn(const fwix_dims& d, const fwix_params& p, int nrec, int ckpt_every = 1, DeviceVector* model = nullptr,
               DeviceVector* data = nullptr, hipStream_t s = nullptr)
      : CudaOperator(model_size(d), (size_t)nrec * d.NW, model, data, s), dims_(d) {
    check(fwix_born_create(&d, &p, ckpt_every, &op_.h));
  }
  void set_refs(const fwix_c64* sref_dev, const uint8_t* labels_dev) { check(fwix_born_set_refs(op_.h, sref_dev, labels_dev)); }
  void set_slow(const fwix_c64* slow_dev) { check(fwix_born_set_slow(op_.h, slow_dev)); }
  void set_sources(const std::vector<float>& x, const std::vector<float>& y, const std::vector<float>& z,
                   const std::vector<int>& ids) {
    check(fwix_born_set_sources(op_.h, (int)x.size(), x.data(), y.data(), z.data(), ids.data()));
  }
  void set_receivers(const std::vector<float>& x, const std::vector<float>& y, const std::vector<float>& z,
                     const std::vector<int>& ids) {
    if (x.size() * (size_t)dims_.NW != getRangeSize()) throw std::runtime_error("receiver count does not match the range");
    check(fwix_born_set_receivers(op_.h, (int)x.size(), x.data(), y.data(), z.data(), ids.data()));
  }
  void set_wavelet(const fwix_c64* wavelet_dev) { check(fwix_born_set_wavelet(op_.h, wavelet_dev)); }
  // background modeling; records the direct data when data_dev != nullptr
  void model(fwix_c64* data_dev = nullptr, bool add = false) { check(fwix_born_model(op_.h, add, data_dev, stream_)); }
  using CudaOperator::cu_adjoint;
  using CudaOperator::cu_forward;
  void cu_forward(bool add, const fwix_c64* m, fwix_c64* d) override { check(fwix_born_forward(op_.h, add, m, d, stream_)); }
  void cu_adjoint(bool add, fwix_c64* m, const fwix_c64* d) override { check(fwix_born_adjoint(op_.h, add, m, d, stream_)); }
  void migrate(bool add, fwix_c64* image_dev, const fwix_c64* data_dev) {
    check(fwix_born_migrate(op_.h, add, image_dev, data_dev, stream_));
  }
  // density, the nonlinear up pass and the four-channel operator on (ds, drho)
  void set_density(const fwix_c64* den_dev) { check(fwix_born_set_density(op_.h, den_dev)); }
  void model_up(fwix_c64* data_dev = nullptr, bool add = false) { check(fwix_born_model_up(op_.h, add, data_dev, stream_)); }
  void extended_forward(bool add, const fwix_c64* ds, const fwix_c64* drho, fwix_c64* data) {
    check(fwix_born_extended_forward(op_.h, add, ds, drho, data, stream_));
  }
  void extended_adjoint(bool add, fwix_c64* ds, fwix_c64* drho, const fwix_c64* data) {
    check(fwix_born_extended_adjoint(op_.h, add, ds, drho, data, stream_));
  }

 private:
  fwix_dims dims_;
  Handle op_;
};

}  // namespace fwix
