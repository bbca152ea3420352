// Host harness for tests/test_buffer_limits.py: the size rule of the buffer-addressed matvec
// (orctm_amd/csrc/nh_kernels.h, bwd_buffers_fit), compiled for the host.
#include <cstdint>

#include "../orctm_amd/csrc/nh_kernels.h"

extern "C" int fits(long long plane, int nz, int fp32) {
  onh::Sys s{};
  s.plane = plane;
  s.nz = nz;
  return onh::bwd_buffers_fit(s, fp32 != 0) ? 1 : 0;
}

extern "C" unsigned long long max_bytes() { return onh::kBufMaxBytes; }
