// Sanitizer driver of the host-compiled parts of the library that the CPU tests load as harnesses
// (tests/test_sanitizers.py): the flag derivation shared with the device (nh_flags.h, through
// tests/flags_host.cpp) and the serving loop's step bookkeeping (nh_serve_step.h, through
// tests/serve_step_host.cpp), called on seeded inputs under AddressSanitizer / UndefinedBehaviorSanitizer.
// Prints "SANITIZED OK" when every call returned.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../flags_host.cpp"
#include "../serve_step_host.cpp"

int main() {
  std::mt19937_64 rng(20261202u);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  const int cases[][5] = {{19, 13, 7, 0, 0}, {16, 11, 5, 1, 0}, {14, 12, 6, 0, 1}, {12, 10, 1, 1, 1}, {3, 3, 2, 0, 0}};
  for (const auto& c : cases) {
    const int IE = c[0], JE = c[1], KE = c[2];
    std::vector<double> weto((size_t)IE * JE * KE);
    for (int j = 0; j < JE; ++j)
      for (int i = 0; i < IE; ++i) {
        const int wet = (int)(u(rng) * (KE + 1));  // leading wet levels of the column
        for (int k = 0; k < KE; ++k) weto[(size_t)i + (size_t)IE * ((size_t)j + (size_t)JE * k)] = k < wet ? 1.0 : 0.0;
      }
    std::vector<uint16_t> out((size_t)(IE - 2) * (JE - 2) * KE);
    derive_flag_plane(weto.data(), IE, JE, KE, c[3], c[4], out.data());
  }
  for (int rc : {0, 1, 3, 4, 5, 6, 7}) {
    const int n = 5;
    std::vector<orctm_nh_stats_t> st(n);
    std::vector<orctm_nh_request_stats_t> rs(n);
    std::vector<int32_t> nsteps = {1, 2, 3, 10, 10}, retire(n);
    for (int q = 0; q < n; ++q) st[q].status = q == 2 ? ORCTM_NH_ERR_NONFINITE : ORCTM_NH_OK;
    for (int step = 0; step < 3; ++step) serve_apply_step(rc, st.data(), n, nsteps.data(), rs.data(), retire.data());
  }
  printf("SANITIZED OK\n");
  return 0;
}
