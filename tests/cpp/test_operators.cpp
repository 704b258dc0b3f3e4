// Drives the C++ operator layer (include/fwix_operators.hpp) on the GPU:
// dot tests of every operator at the reference's tolerance (1e-5), host
// in-place vs out-of-place forms, and a PSPI forward written to <outdir> for
// the Python test to compare against the FP64 oracle.
//
//   test_operators <outdir>        exit status 0 when every check passes
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "fwix_operators.hpp"

using namespace fwix;

static int failures = 0;

static void report(const char* name, std::pair<double, double> e) {
  const bool ok = e.first < 1e-5 && e.second < 1e-5;
  std::printf("%-18s dot error add=false %.2e add=true %.2e %s\n", name, e.first, e.second, ok ? "ok" : "FAIL");
  if (!ok) ++failures;
}

template <class T>
static void write(const std::string& path, const std::vector<T>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

int main(int argc, char** argv) {
  const std::string out = argc > 1 ? argv[1] : ".";
  try {
    fwix_dims d{};
    d.NX = 32; d.NY = 24; d.NW = 3; d.NS = 2; d.nx = 24; d.ny = 16;
    d.ox = 0.f; d.dx = 10.f; d.oy = 0.f; d.dy = 10.f; d.ow = 5.f; d.dw = 2.f; d.oz = 0.f; d.dz = 10.f; d.nz = 4;
    fwix_params p;
    fwix_params_default(&p);
    p.nref = 2; p.taperx = 3; p.tapery = 2;
    hipStream_t s = nullptr;

    // two references per (depth, frequency) and a label pattern
    const size_t nlab = (size_t)d.nz * d.NW * d.NY * d.NX;
    std::vector<cfloat> sref((size_t)d.nz * p.nref * d.NW);
    for (int z = 0; z < d.nz; ++z)
      for (int r = 0; r < p.nref; ++r)
        for (int w = 0; w < d.NW; ++w) {
          const float v = 1500.f + 300.f * r + 50.f * z;
          sref[((size_t)z * p.nref + r) * d.NW + w] = cfloat(1.f / (v * v), 0.f);
        }
    std::vector<uint8_t> lab(nlab);
    for (size_t i = 0; i < nlab; ++i) {
      const size_t x = i % d.NX, y = (i / d.NX) % d.NY;
      lab[i] = (uint8_t)(((x / 5) + (y / 4)) % 2);
    }
    DeviceVector sref_d(sref.size());
    sref_d.upload(sref, s);
    uint8_t* lab_d = nullptr;
    hip_ok(hipMalloc(reinterpret_cast<void**>(&lab_d), nlab), "hipMalloc");
    hip_ok(hipMemcpy(lab_d, lab.data(), nlab, hipMemcpyHostToDevice), "labels");
    std::vector<cfloat> slow((size_t)d.nz * d.NW * d.ny * d.nx);
    for (size_t i = 0; i < slow.size(); ++i) {
      const float v = 1500.f + 400.f * (float)((i * 2654435761u) % 1000) / 1000.f;
      slow[i] = cfloat(1.f / (v * v), 0.f);
    }
    DeviceVector slow_d(slow.size());
    slow_d.upload(slow, s);
    hip_ok(hipDeviceSynchronize(), "sync");

    { cuFFT2d op(d); report("cuFFT2d", op.dotTest(false, 1)); }
    {
      PhaseShift op(d, d.dz, 0.f);
      op.set_slow(std::vector<cfloat>(d.NW, cfloat(1.f / (2000.f * 2000.f), 0.f)));
      report("PhaseShift", op.dotTest(false, 2));
    }
    {
      Selector op(d);
      std::vector<int32_t> l((size_t)d.NW * d.NY * d.NX);
      for (size_t i = 0; i < l.size(); ++i) l[i] = lab[i];
      op.set_labels(l);
      op.set_value(1);
      report("Selector", op.dotTest(false, 3));
    }
    { Taper op(d, p); report("Taper", op.dotTest(false, 4)); }
    {
      PSPI op(d, p);
      op.set_refs(sref_d.data(), lab_d);
      op.set_depth(2);
      report("PSPI", op.dotTest(false, 5));
      // out of place vs in place, and the vector for the oracle comparison
      hostvec m(wavefield_size(d)), y, u;
      std::mt19937 g(11);
      std::normal_distribution<float> n;
      for (auto& v : m) v = cfloat(n(g), n(g));
      op.forward(false, m, y);
      u = m;
      op.forward(u);
      double diff = 0, nrm = 0;
      for (size_t i = 0; i < y.size(); ++i) { diff += std::norm(y[i] - u[i]); nrm += std::norm(y[i]); }
      const bool ok = std::sqrt(diff / nrm) < 1e-6;
      std::printf("PSPI in place vs out of place rel %.2e %s\n", std::sqrt(diff / nrm), ok ? "ok" : "FAIL");
      if (!ok) ++failures;
      write(out + "/pspi_in.bin", m);
      write(out + "/pspi_out.bin", y);
      write(out + "/sref.bin", sref);
      write(out + "/labels.bin", lab);
    }
    {
      NSPS op(d, p);
      op.set_refs(sref_d.data(), lab_d);
      op.set_depth(1);
      report("NSPS", op.dotTest(false, 6));
    }
    {
      std::vector<float> cx{15.f, 95.f, 200.f}, cy{12.f, 40.f, 150.f}, cz{3.f, 12.f, 25.f};
      std::vector<int> ids{0, 1, 1};
      Injection op(d, cx, cy, cz, ids);
      op.set_depth(1);
      report("Injection", op.dotTest(false, 7));
    }
    {
      Scatter op(d, p);
      op.set_slow(slow_d.data());
      op.set_depth(3);
      report("Scatter", op.dotTest(false, 8));
    }
    {
      ImagingCondition op(d);
      DeviceVector bg(wavefield_size(d));
      hostvec b(wavefield_size(d));
      std::mt19937 g(12);
      std::normal_distribution<float> n;
      for (auto& v : b) v = cfloat(n(g), n(g));
      bg.upload(b, s);
      op.set_background(bg.data());
      report("ImagingCondition", op.dotTest(false, 9));
    }
    {
      Downward dn(d, p);
      dn.set_refs(sref_d.data(), lab_d);
      report("Downward", dn.dotTest(false, 10));
      Upward up(d, p);
      up.set_refs(sref_d.data(), lab_d);
      report("Upward", up.dotTest(false, 11));
    }
    {
      std::vector<float> rx, ry, rz;
      std::vector<int> rid;
      for (int j = 1; j < 15; j += 3)
        for (int i = 1; i < 23; i += 3) {
          rx.push_back(10.f * i); ry.push_back(10.f * j); rz.push_back(22.f); rid.push_back((i + j) % 2);
        }
      ExtendedBorn op(d, p, (int)rx.size(), 2);
      op.set_refs(sref_d.data(), lab_d);
      op.set_slow(slow_d.data());
      op.set_sources({110.f, 60.f}, {70.f, 90.f}, {2.f, 2.f}, {0, 1});
      op.set_receivers(rx, ry, rz, rid);
      DeviceVector wav((size_t)2 * d.NW);
      wav.upload(hostvec((size_t)2 * d.NW, cfloat(1.f, 0.5f)), s);
      op.set_wavelet(wav.data());
      op.model();
      report("ExtendedBorn", op.dotTest(false, 12));
      // all four channels on (ds, drho): dot test on the device
      std::vector<cfloat> den(slow.size(), cfloat(1.f, 0.f));
      for (size_t i = 0; i < den.size(); ++i) den[i] = cfloat(1.f + 0.3f * (float)((i * 40503u) % 97) / 97.f, 0.f);
      DeviceVector den_d(den.size());
      den_d.upload(den, s);
      op.set_density(den_d.data());
      op.model_up();
      const size_t nm = model_size(d), nd = rx.size() * (size_t)d.NW;
      std::mt19937 g(13);
      std::uniform_real_distribution<float> u(-1.f, 1.f);
      hostvec ms(nm), mr(nm), dd(nd);
      for (auto& v : ms) v = cfloat(u(g), u(g));
      for (auto& v : mr) v = cfloat(u(g), u(g));
      for (auto& v : dd) v = cfloat(u(g), u(g));
      DeviceVector ms_d(nm), mr_d(nm), dd_d(nd), am_d(nd), as_d(nm), ar_d(nm);
      ms_d.upload(ms, s); mr_d.upload(mr, s); dd_d.upload(dd, s);
      op.extended_forward(false, ms_d.data(), mr_d.data(), am_d.data());
      op.extended_adjoint(false, as_d.data(), ar_d.data(), dd_d.data());
      hostvec am, as, ar;
      am_d.download(am, s); as_d.download(as, s); ar_d.download(ar, s);
      const std::complex<double> lhs = dot(am, dd), rhs = dot(ms, as) + dot(mr, ar);
      const double e = std::abs((lhs / rhs).real() - 1.0);
      std::printf("%-18s dot error %.2e %s\n", "ExtendedBorn x4", e, e < 1e-5 ? "ok" : "FAIL");
      if (!(e < 1e-5)) ++failures;
    }
    // errors surface as exceptions, not exit()
    try {
      fwix_dims bad = d;
      bad.NX = 100;
      cuFFT2d op(bad);
      std::printf("bad FFT size accepted FAIL\n");
      ++failures;
    } catch (const std::runtime_error& e) {
      std::printf("bad FFT size rejected: %s ok\n", e.what());
    }
    (void)hipFree(lab_d);
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 2;
  }
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
  return failures ? 1 : 0;
}
