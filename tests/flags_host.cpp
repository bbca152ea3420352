// Host harness for tests/test_flag_derivation.py: the flag plane derived from wet depths by the same
// header the device uses (orctm_amd/csrc/nh_flags.h), in the oracle's layout (flags[k*nx*ny + j*nx + i]).
#include <cstdint>

#include "../orctm_amd/csrc/nh_flags.h"

extern "C" void derive_flag_plane(const double* weto, int IE, int JE, int KE, int cyc_x, int cyc_y, uint16_t* out) {
  const int nx = IE - 2, ny = JE - 2;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      uint32_t w[3];
      onh::column_flag_words(weto, IE, JE, KE, i + 2, j + 2, cyc_x, cyc_y, w);
      for (int K = 1; K <= KE; ++K)
        out[(size_t)(K - 1) * nx * ny + (size_t)j * nx + i] = (uint16_t)onh::flags_from_words(K, KE, w);
    }
}
