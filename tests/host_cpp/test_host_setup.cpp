// The library's host C++ (descriptor validation, layout packing, the record,
// beam-moment and profile set-up, the coefficient tables) built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers.py,
// without a GPU:
//   * malformed descriptors are refused with the documented status codes;
//   * records from synthetic element tables are finite, and an energy outside
//     the stopping-power table is refused;
//   * beam moments and profiles for every order and a range of concentrations;
//   * pack_class / unpack_class round trips over every class in 1D, 2D, 3D.
// Prints "ok" and exits 0, or prints the failed check and exits 1.
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../pnepma_amd/csrc/pnx_host.hpp"

namespace pnx {
// the library defines this in its C ABI source; the test records the message
static std::string g_msg;
int set_error(int code, const std::string& msg) {
  g_msg = msg;
  return code;
}
}  // namespace pnx

using namespace pnx;

static int failures = 0;
#define CHECK(c)                                                 \
  do {                                                           \
    if (!(c)) {                                                  \
      std::printf("check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                \
    }                                                            \
  } while (0)

static pnx_desc desc(int N, int mx, int my, int mz) {
  pnx_desc d{};
  d.N = N;
  d.M[0] = mx;
  d.M[1] = my;
  d.M[2] = mz;
  for (int a = 0; a < 3; ++a) {
    d.lo[a] = -1e-4;
    d.hi[a] = d.M[a] > 1 ? 1e-4 : -1e-4;
  }
  d.n_elements = 2;
  d.n_lines = 2;
  d.n_steps = 12;
  d.n_beams = 1;
  d.E_initial = 15000.0;
  d.E_cutoff = 5000.0;
  return d;
}

static void test_descriptors() {
  Geometry g;
  std::string err;
  pnx_desc d = desc(5, 10, 8, 1);
  CHECK(make_geometry(d, g, err) == PNX_OK);
  d = desc(4, 10, 8, 1);
  CHECK(make_geometry(d, g, err) == PNX_ERR_UNSUPPORTED);  // even orders are not compiled
  d = desc(5, 10, 8, 1);
  d.n_elements = 9;
  CHECK(make_geometry(d, g, err) == PNX_ERR_ARG);
  d = desc(5, 10, 8, 1);
  d.n_lines = 0;
  CHECK(make_geometry(d, g, err) == PNX_ERR_ARG);
  d = desc(5, 10, 8, 1);
  d.n_steps = 0;
  CHECK(make_geometry(d, g, err) == PNX_ERR_ARG);
  d = desc(5, 10, 8, 1);
  d.E_cutoff = d.E_initial;
  CHECK(make_geometry(d, g, err) == PNX_ERR_ARG);
  d = desc(5, 1, 8, 1);  // a resolved axis after a singleton one
  CHECK(make_geometry(d, g, err) == PNX_ERR_ARG);
  d = desc(5, 10, 8, 1);
  d.hi[0] = d.lo[0];
  CHECK(make_geometry(d, g, err) == PNX_ERR_ARG);
  d = desc(5, 0, 1, 1);
  CHECK(make_geometry(d, g, err) == PNX_ERR_ARG);
}

static void test_records() {
  std::mt19937_64 rng(3);
  std::uniform_real_distribution<double> U(0.5, 2.0);
  double sp_E[6] = {500.0, 2e3, 5e3, 1e4, 2e4, 5e4}, sp[2][6], tc_E[5] = {100.0, 1e3, 1e4, 3e4, 1e5};
  std::vector<double> tc[2];
  for (int e = 0; e < 2; ++e) {
    for (int k = 0; k < 6; ++k) sp[e][k] = U(rng) * 1e8 / std::sqrt(sp_E[k]);
    tc[e].resize(5 * 10);
    for (int k = 0; k < 5; ++k)
      for (int l = 0; l < 10; ++l) tc[e][k * 10 + l] = U(rng) * l * (l + 1) * 1e4 / tc_E[k];
  }
  pnx_physics ph{};
  ph.n_knots = 6;
  ph.sp_E = sp_E;
  ph.n_tc = 5;
  ph.tc_stride = 10;
  ph.tc_E = tc_E;
  for (int e = 0; e < 2; ++e) {
    ph.sp[e] = sp[e];
    ph.tc[e] = tc[e].data();
  }
  ph.line_edge[0] = 7112.0;
  ph.line_edge[1] = 8332.8;
  ph.beam_E_mean = 14500.0;
  ph.beam_E_sd = 100.0;
  for (int N : {1, 3, 5, 7, 9}) {
    pnx_desc d = desc(N, 10, 8, 6);
    Geometry g;
    std::string err;
    CHECK(make_geometry(d, g, err) == PNX_OK);
    const size_t n = (size_t)(d.n_steps + 1) * g.rec_len;
    std::vector<double> fwd(n), adj(n), dens(n);
    CHECK(pnx_problem_records(&d, &ph, fwd.data(), adj.data(), dens.data()) == PNX_OK);
    for (size_t i = 0; i < n; ++i) CHECK(std::isfinite(fwd[i]) && std::isfinite(adj[i]) && std::isfinite(dens[i]));
    d.E_cutoff = 100.0;  // below the table
    CHECK(pnx_problem_records(&d, &ph, fwd.data(), adj.data(), dens.data()) == PNX_ERR_ARG);
    // beam moments and profiles
    const Tables t = make_tables(N);
    for (double kappa : {0.1, 5.0, 50.0, 500.0}) {
      double mm[8][16] = {};
      double* m8[8];
      for (int c = 0; c < 8; ++c) m8[c] = mm[c];
      const double dir[3] = {-0.8, 0.3, 0.2};
      CHECK(pnx_beam_moments(N, dir, kappa, m8) == PNX_OK);
      for (int c = 0; c < 8; ++c)
        for (int i = 0; i < t.n[c]; ++i) CHECK(std::isfinite(mm[c][i]) && std::fabs(mm[c][i]) < 1.0);
    }
    d = desc(N, 10, 8, 6);
    std::vector<double> pp[8];
    double* p8[8];
    for (int c = 0; c < 8; ++c) {
      pp[c].assign((size_t)g.class_n[c][1] * g.class_n[c][2], -1.0);
      p8[c] = pp[c].data();
    }
    const double pos[3] = {0.0, 0.1e-4, -0.2e-4};
    CHECK(pnx_beam_profiles(&d, pos, 50e-7, 0, p8) == PNX_OK);
    for (int c = 0; c < 8; ++c)
      for (double v : pp[c]) CHECK(odd_in(c, 0) ? (v >= 0.0 && std::isfinite(v)) : v == -1.0);
  }
}

static void test_packing() {
  std::mt19937_64 rng(5);
  for (int dims = 1; dims <= 3; ++dims) {
    pnx_desc d = desc(5, 70, dims >= 2 ? 9 : 1, dims >= 3 ? 6 : 1);
    Geometry g;
    std::string err;
    CHECK(make_geometry(d, g, err) == PNX_OK);
    for (int c = 0; c < 8; ++c) {
      if (!g.active(c)) continue;
      const long np = (long)g.class_n[c][0] * g.class_n[c][1] * g.class_n[c][2];
      const int vec = g.T.n[c];
      std::vector<double> src((size_t)np * vec), back((size_t)np * vec, 0.0), dev((size_t)vec * g.L.V, 0.0);
      for (double& x : src) x = (double)(rng() % 1000);
      pack_class(g, c, src.data(), vec, dev.data());
      unpack_class(g, c, dev.data(), vec, back.data());
      CHECK(src == back);
    }
  }
}

int main() {
  test_descriptors();
  test_records();
  test_packing();
  for (int N : {1, 3, 5, 7, 9}) CHECK(!build_coefficients(N, false).empty() && !build_coefficients(N, true).empty());
  if (failures) {
    std::printf("%d checks failed\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
