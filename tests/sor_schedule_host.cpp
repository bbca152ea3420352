// Host harness for tests/test_sor_schedule.py: steps the device's temporally blocked SOR schedule
// (orctm_amd/csrc/nh_sor_tb.h, driven as orctm_amd/csrc/nh_sor.hip drives it) thread by thread on the
// CPU, barrier by barrier, and compares Z1O bit for bit with the oracle's sequential TRONEU
// (oracle/nh_oracle_sor.c, troneu.f90:48-131) on seeded random grids. Prints one line per case and
// "ALL BIT-EXACT" at the end when every case matches.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../orctm_amd/csrc/nh_sor_tb.h"

extern "C" {
#include "../oracle/nh_oracle_ocv.h"
#include "../oracle/nh_oracle_sor.h"
}

using namespace onh;

namespace {

struct Case {
  int IE, JE, IS, JS, cx, cy, iters;
};

struct Grid {
  Case c;
  std::vector<double> dlxp, dlyp, dtdxpo, dtdyo, weto1, dlxu, dlyu, deuto, dlxv, dlyv, deute;
  std::vector<double> u1o, uzo, v1e, vze;
  orc_sorgrid g;
};

void make(Grid& G, const Case& c, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> met(900.0, 2100.0), dep(0.0, 800.0), vel(-0.5, 0.5), u01(0.0, 1.0);
  const size_t np = (size_t)c.IE * c.JE, nu = (size_t)(c.IE - c.IS + 1) * c.JE, nv = (size_t)c.IE * (c.JE - c.JS + 1);
  const double dt = 60.0;
  G.c = c;
  auto fill = [&](std::vector<double>& v, size_t n, auto&& f) {
    v.resize(n);
    for (auto& x : v) x = f();
  };
  fill(G.dlxp, np, [&] { return met(rng); });
  fill(G.dlyp, np, [&] { return met(rng); });
  G.dtdxpo.resize(np);
  G.dtdyo.resize(np);
  for (size_t n = 0; n < np; ++n) {
    G.dtdxpo[n] = dt / G.dlxp[n];
    G.dtdyo[n] = dt / G.dlyp[n];
  }
  fill(G.weto1, np, [&] { return u01(rng) < 0.85 ? 1.0 : 0.0; });
  fill(G.dlxu, nu, [&] { return met(rng); });
  fill(G.dlyu, nu, [&] { return met(rng); });
  fill(G.deuto, nu, [&] { return u01(rng) < 0.1 ? 0.0 : dep(rng); });
  fill(G.dlxv, nv, [&] { return met(rng); });
  fill(G.dlyv, nv, [&] { return met(rng); });
  fill(G.deute, nv, [&] { return u01(rng) < 0.1 ? 0.0 : dep(rng); });
  fill(G.u1o, nu, [&] { return vel(rng); });
  fill(G.uzo, nu, [&] { return vel(rng); });
  fill(G.v1e, nv, [&] { return vel(rng); });
  fill(G.vze, nv, [&] { return vel(rng); });
  G.g = orc_sorgrid{c.IE,          c.JE,          c.IS,          c.JS,         c.cx,          c.cy,
                    c.iters,       G.dlxp.data(), G.dlyp.data(), G.dtdxpo.data(), G.dtdyo.data(), G.weto1.data(),
                    G.dlxu.data(), G.dlyu.data(), G.deuto.data(), G.dlxv.data(), G.dlyv.data(), G.deute.data(),
                    9.81,          0.5,           0.5,           1.0,          dt,            1.75};
}

// k_sor: one half-sweep of `colour` on z in place (the device's first half-sweep)
void half_sweep(const Case& c, const std::vector<double>& b1o, const std::vector<double>& ff,
                const std::vector<double>& uf, const std::vector<double>& vf, std::vector<double>& z, int colour,
                double sorpar) {
  const int IE = c.IE, NU = c.IE - c.IS + 1;
  auto P2 = [&](int i, int j) { return (size_t)(i - 1) + (size_t)IE * (j - 1); };
  auto U2 = [&](int i, int j) { return (size_t)(i - c.IS) + (size_t)NU * (j - 1); };
  auto V2 = [&](int i, int j) { return (size_t)(i - 1) + (size_t)IE * (j - c.JS); };
  const double sorpai = 1. - sorpar;
  for (int j = 2; j <= c.JE - 1; ++j)
    for (int i = 2; i <= c.IE - 1; ++i) {
      if ((i + j + colour) & 1) continue;
      const double s = b1o[P2(i, j)] + uf[U2(i, j)] * z[P2(i + 1, j)] + uf[U2(i - 1, j)] * z[P2(i - 1, j)] +
                       vf[V2(i, j)] * z[P2(i, j + 1)] + vf[V2(i, j - 1)] * z[P2(i, j - 1)];
      z[P2(i, j)] = sorpar * ff[P2(i, j)] * s + sorpai * z[P2(i, j)];
    }
}

template <int T, int RX, int RY, int NT, bool SEAM>
int run(const Grid& G, const std::vector<double>& ff, const std::vector<double>& uf, const std::vector<double>& vf,
        const std::vector<double>& b1o, const std::vector<double>& z0, const std::vector<double>& zref) {
  using K = SorTb<T, RX, RY, NT, SEAM>;
  const Case& c = G.c;
  TbGeom g{c.IE, c.JE, c.IS, c.JS, c.IE - c.IS + 1, c.cx, c.cy, 0};
  g.tiles_x = (c.IE - 2 + K::TXo - 1) / K::TXo;
  const int nb = g.tiles_x * ((c.JE - 2 + K::TYo - 1) / K::TYo);
  const double sorpar = G.g.sorpar, sorpai = 1. - sorpar;
  std::vector<double> buf[2] = {z0, z0};
  if (c.iters > 0) {
    const long long H = 2LL * c.iters - 1, nl = (H + T - 1) / T;
    const int s0 = (int)(nl & 1);
    half_sweep(c, b1o, ff, uf, vf, buf[s0], 0, sorpar);
    orc_bounds_exch(buf[s0].data(), c.IE, c.JE, 1, 1, 1, c.cx, c.cy);
    std::vector<typename K::Regs> regs(NT);
    std::vector<double> lz(K::LDS);
    for (long long l = 0; l < nl; ++l) {
      const int nh = (int)(H - l * T < T ? H - l * T : T);
      const int c0 = (int)((1 + l * T) & 1);
      const std::vector<double>& zin = buf[(s0 + l) & 1];
      std::vector<double>& zout = buf[(s0 + l + 1) & 1];
      std::vector<int> seen(nb, 0);
      for (int b = 0; b < nb; ++b) {
        const int t = tb_tile(b, nb);
        if (t < 0 || t >= nb || seen[t]++) {
          printf("block -> tile map is not a bijection\n");
          return 1;
        }
        int vx0, vy0;
        K::origin(g, t, vx0, vy0);
        std::fill(lz.begin(), lz.end(), -1e300);
        for (int tid = 0; tid < NT; ++tid)
          K::load(tid, vx0, vy0, g, zin.data(), b1o.data(), ff.data(), uf.data(), vf.data(), sorpar, lz.data(),
                  regs[tid]);
        for (int h = 0; h < nh; ++h) {
          if ((c0 + h) & 1) {
            for (int tid = 0; tid < NT; ++tid) K::template compute<1>(tid, sorpai, lz.data(), regs[tid]);
            for (int tid = 0; tid < NT; ++tid) K::template write<1>(tid, lz.data(), regs[tid]);
          } else {
            for (int tid = 0; tid < NT; ++tid) K::template compute<0>(tid, sorpai, lz.data(), regs[tid]);
            for (int tid = 0; tid < NT; ++tid) K::template write<0>(tid, lz.data(), regs[tid]);
          }
        }
        for (int tid = 0; tid < NT; ++tid) K::store(tid, vx0, vy0, g, lz.data(), zout.data());
      }
    }
    orc_bounds_exch(buf[0].data(), c.IE, c.JE, 1, 1, 1, c.cx, c.cy);
  }
  size_t bad = 0;
  for (size_t n = 0; n < zref.size(); ++n)
    if (memcmp(&buf[0][n], &zref[n], sizeof(double))) ++bad;
  printf("T=%2d seam=%d IE=%4d JE=%4d IS=%d JS=%d cyclic=%d%d sweeps=%d: %zu of %zu differ\n", T, (int)SEAM, c.IE,
         c.JE, c.IS, c.JS, c.cx, c.cy, c.iters, bad, zref.size());
  return bad != 0;
}

}  // namespace

int main() {
  const Case cases[] = {
      {72, 39, 1, 1, 0, 0, 25},  {66, 23, 0, 1, 1, 0, 25}, {42, 26, 1, 0, 1, 1, 25}, {98, 6, 1, 1, 0, 1, 25},
      {35, 31, 0, 0, 0, 0, 25},  {9, 7, 1, 1, 1, 1, 25},   {63, 20, 1, 1, 1, 0, 25}, {21, 30, 0, 0, 1, 1, 25},
      {3, 3, 1, 1, 1, 1, 5},     {4, 5, 1, 1, 0, 1, 7},    {150, 130, 1, 1, 1, 1, 30}, {130, 70, 0, 0, 0, 1, 3},
      {50, 40, 1, 1, 1, 1, 0},   {50, 40, 1, 1, 1, 1, 1},
  };
  int fails = 0;
  unsigned seed = 20261202u;
  for (const Case& c : cases) {
    Grid G;
    make(G, c, seed++);
    const size_t np = (size_t)c.IE * c.JE;
    std::vector<double> xx(np), ff(np), uf((size_t)(c.IE - c.IS + 1) * c.JE), vf((size_t)c.IE * (c.JE - c.JS + 1));
    orc_itprep(&G.g, xx.data(), uf.data(), vf.data(), ff.data());
    // the state before the sweeps (B1O, and Z1O with its boundary at B1O/XX), then the reference
    std::vector<double> b1o(np, 0.0), z0(np, 0.0), b1r(np, 0.0), zref(np, 0.0);
    orc_sorgrid g0 = G.g;
    g0.iterations = 0;
    orc_troneu(&g0, xx.data(), uf.data(), vf.data(), ff.data(), G.u1o.data(), G.uzo.data(), G.v1e.data(),
               G.vze.data(), b1o.data(), z0.data());
    orc_troneu(&G.g, xx.data(), uf.data(), vf.data(), ff.data(), G.u1o.data(), G.uzo.data(), G.v1e.data(),
               G.vze.data(), b1r.data(), zref.data());
    // the rule of nh_sor.hip: an odd cyclic period selects the seam kernel; the seam kernel is also
    // checked everywhere
    const bool seam = (c.cx && ((c.IE - 2) & 1)) || (c.cy && ((c.JE - 2) & 1));
#define X(T, RX, RY, NT)                                                                              \
  fails += seam ? run<T, RX, RY, NT, true>(G, ff, uf, vf, b1o, z0, zref)                              \
                : run<T, RX, RY, NT, false>(G, ff, uf, vf, b1o, z0, zref);                            \
  if (!seam) fails += run<T, RX, RY, NT, true>(G, ff, uf, vf, b1o, z0, zref);
    ORCTM_SOR_TB_VARIANTS(X)
#undef X
  }
  printf("%s\n", fails ? "MISMATCH" : "ALL BIT-EXACT");
  return fails != 0;
}
