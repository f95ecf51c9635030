// Runs TAL-SH's own C++ wrapper (talsh::Tensor, talshxx.hpp / talshxx.cpp,
// compiled unchanged from the reference checkout by oracle/Makefile target
// cxx) on libtalsh_amd.so and the GPU:
//   1. contractAccumulate with constant tensors against the closed form
//      d0 + alpha * l * r * K (the reference's own check, test.cpp:193-226);
//   2. extractSlice of a tensor whose values are their own dlf positions;
//   3. insertSlice of a constant block, everything outside it unchanged;
//   4. contractAccumulateXL with complex constant tensors, cut into pieces by a
//      small device-memory grant, against the closed form of test.cpp:326-371,
//      and the full reduction into a scalar.
// Prints "cxx wrapper driver: OK" and exits 0 when every check holds.
#include "talshxx.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "talsh_amd.h"

static int fails = 0;
static void expect(bool ok, const char *what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++fails;
  }
}

int main() {
  std::size_t host_buffer = (std::size_t)1 << 28;
  if (talsh::initialize(&host_buffer) != TALSH_SUCCESS) {
    std::printf("talsh::initialize failed\n");
    return 2;
  }
  {
    const int n = 24;
    const std::vector<int> dims{n, n, n, n};
    talsh::Tensor d(dims, 0.25), l(dims, 0.001), r(dims, 0.01);
    const int e = d.contractAccumulate(nullptr, "D(a,b,c,d)+=L(d,i,c,j)*R(j,b,i,a)", l, r, DEV_NVIDIA_GPU, 0, 0.5);
    expect(e == TALSH_SUCCESS, "contractAccumulate returned an error");
    expect(d.sync(DEV_HOST, 0), "sync of the contraction result");
    const double *p = nullptr;
    expect(d.getDataAccessHostConst(&p), "host access to the contraction result");
    const double want = 0.25 + 0.5 * 0.001 * 0.01 * n * n;
    double worst = 0.0;
    for (std::size_t i = 0; p && i < (std::size_t)n * n * n * n; ++i) worst = std::fmax(worst, std::fabs(p[i] - want));
    expect(p && worst <= 1e-15, "contraction closed form");
  }
  {
    const std::vector<int> bd{7, 5, 6}, sd{3, 4, 2}, off{2, 1, 3};
    std::vector<double> vals(7 * 5 * 6);
    for (std::size_t i = 0; i < vals.size(); ++i) vals[i] = (double)i;
    talsh::Tensor big(bd, vals), small(sd, 0.0);
    expect(big.extractSlice(nullptr, small, off, DEV_NVIDIA_GPU, 0) == TALSH_SUCCESS, "extractSlice");
    expect(small.sync(DEV_HOST, 0), "sync of the slice");
    const double *s = nullptr;
    expect(small.getDataAccessHostConst(&s), "host access to the slice");
    bool ok = s != nullptr;
    for (int k = 0; ok && k < sd[2]; ++k)
      for (int j = 0; j < sd[1]; ++j)
        for (int i = 0; i < sd[0]; ++i) {
          const double pos = (i + off[0]) + (double)bd[0] * ((j + off[1]) + (double)bd[1] * (k + off[2]));
          if (s[i + sd[0] * (j + sd[1] * k)] != pos) ok = false;
        }
    expect(ok, "slice holds the positions of the block");
    talsh::Tensor block(sd, -1.0);
    expect(big.insertSlice(nullptr, block, off, DEV_NVIDIA_GPU, 0) == TALSH_SUCCESS, "insertSlice");
    expect(big.sync(DEV_HOST, 0), "sync after insert");
    const double *b = nullptr;
    expect(big.getDataAccessHostConst(&b), "host access after insert");
    ok = b != nullptr;
    for (int k = 0; ok && k < bd[2]; ++k)
      for (int j = 0; j < bd[1]; ++j)
        for (int i = 0; i < bd[0]; ++i) {
          const bool in = i >= off[0] && i < off[0] + sd[0] && j >= off[1] && j < off[1] + sd[1] && k >= off[2] &&
                          k < off[2] + sd[2];
          const std::size_t lin = i + (std::size_t)bd[0] * (j + (std::size_t)bd[1] * k);
          if (b[lin] != (in ? -1.0 : (double)lin)) ok = false;
        }
    expect(ok, "insert wrote exactly the block");
  }
  {
    const int o = 8, v = 16;
    const std::vector<int> dims{o, v, o, v};
    const float d_init = 1e-4f;
    talsh::Tensor r(dims, std::complex<float>{0.001f, 0.0f}), l(dims, std::complex<float>{0.01f, 0.0f});
    talsh::Tensor d(dims, std::complex<float>{d_init, 0.0f}), s(std::vector<int>{}, std::complex<float>{0.0f, 0.0f});
    expect(talshAmdSetContractXLMemory(64 << 10) == TALSH_SUCCESS, "XL memory grant");
    const int e = d.contractAccumulateXL(nullptr, "D(i,a,j,b)+=L(j,a,k,c)*R(k,b,i,c)", l, r, DEV_NVIDIA_GPU, 0,
                                         std::complex<float>{0.5f, 0.0f});
    expect(e == TALSH_SUCCESS, "contractAccumulateXL returned an error");
    expect(d.sync(DEV_HOST, 0), "sync of the XL result");
    const std::complex<float> *p = nullptr;
    expect(d.getDataAccessHostConst(&p), "host access to the XL result");
    const double want = (double)(o * v) * 1e-3 * 1e-2 * 0.5 + d_init;
    double worst = 0.0;
    for (std::size_t i = 0; p && i < (std::size_t)o * v * o * v; ++i) worst = std::fmax(worst, std::abs(p[i] - std::complex<float>((float)want, 0.0f)));
    expect(p && worst <= 1e-6 * want, "XL contraction closed form");
    const int e2 = s.contractAccumulateXL(nullptr, "D()+=L(i,a,j,b)*R(j,b,i,a)", l, r, DEV_NVIDIA_GPU, 0,
                                          std::complex<float>{1.0f, 0.0f});
    expect(e2 == TALSH_SUCCESS, "contractAccumulateXL into a scalar returned an error");
    expect(s.sync(DEV_HOST, 0), "sync of the XL scalar");
    const std::complex<float> *q = nullptr;
    expect(s.getDataAccessHostConst(&q), "host access to the XL scalar");
    const double sw = (double)(o * v) * (o * v) * 1e-3 * 1e-2;
    expect(q && std::abs(q[0] - std::complex<float>((float)sw, 0.0f)) <= 1e-5 * sw, "XL scalar closed form");
    talshAmdSetContractXLMemory(0);
  }
  talsh::shutdown();
  std::printf("cxx wrapper driver: %s\n", fails ? "FAIL" : "OK");
  return fails ? 1 : 0;
}
