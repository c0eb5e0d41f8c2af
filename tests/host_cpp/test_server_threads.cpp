// The continuous-batching server's threads on the CPU, built with
// ThreadSanitizer or with AddressSanitizer + UndefinedBehaviorSanitizer by
// tests/test_sanitizers.py.  The server (pnx_server.cpp) and its C++ set-up
// are linked against fake_hip.cpp, which stands in for the HIP runtime and
// the kernel launchers (no GPU; kernels do nothing), so what runs is the host
// side: submission and its C++ set-up, admission under the HBM budget,
// planning and staging of ticks on the worker thread, completion, polling,
// waiting, cancellation and statistics from client threads.  The results are
// not checked (the kernels are no-ops); the sanitizers check the threads'
// accesses and the memory, and the program checks that
//   * every request ends completed or cancelled, the counts agree with the
//     server's statistics, and the server drains and is destroyed;
//   * with the n-th kernel launch failing, every request that had not
//     completed before it is reported with PNX_ERR_DEVICE, and the server is
//     still destroyed;
//   * with a device allocation failing at creation, pnx_server_create
//     returns PNX_ERR_MEMORY (the arena) or PNX_ERR_DEVICE (a coefficient
//     buffer) and frees what it had allocated (AddressSanitizer's leak check).
// Prints "ok" and exits 0, or prints the failed check and exits 1.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../pnepma_amd/csrc/pnx_host.hpp"

namespace pnx {
// the library defines this in its C ABI source; the test keeps the last message
static std::mutex g_mu;
static std::string g_msg;
int set_error(int code, const std::string& msg) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_msg = msg;
  return code;
}
}  // namespace pnx

void fake_fail_allocation(long k);
void fake_fail_launch(long k);

static std::atomic<int> failures{0};
#define CHECK(c)                                                 \
  do {                                                           \
    if (!(c)) {                                                  \
      std::printf("check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                \
    }                                                            \
  } while (0)

// Two elements and two lines; the stopping power and transport tables cover
// 100 eV - 100 keV.
struct Physics {
  double sp_E[6] = {500.0, 2e3, 5e3, 1e4, 2e4, 5e4}, sp[2][6], tc_E[5] = {100.0, 1e3, 1e4, 3e4, 1e5};
  std::vector<double> tc[2];
  pnx_physics ph{};
  Physics() {
    for (int e = 0; e < 2; ++e) {
      for (int k = 0; k < 6; ++k) sp[e][k] = (1.0 + 0.3 * e) * 1e8 / std::sqrt(sp_E[k]);
      tc[e].resize(5 * 10);
      for (int k = 0; k < 5; ++k)
        for (int l = 0; l < 10; ++l) tc[e][k * 10 + l] = (1.0 + 0.2 * e) * l * (l + 1) * 1e4 / tc_E[k];
    }
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
  }
};

// One request given as cell fractions, with its inputs and outputs.
struct Job {
  std::vector<double> cell_w, grad, I_out = std::vector<double>(2, 0.0);
  double rho_pure[2] = {7.87, 8.91}, embedding[2] = {0.5, 0.5}, mac[4] = {50.0, 80.0, 60.0, 40.0};
  double atom_mass[2] = {9.27e-23, 9.75e-23}, I_bar[2] = {1.0, 0.5};
  int32_t line_element[2] = {0, 1};
  pnx_request_spec spec{};

  Job(const pnx_physics* ph, int N, int mx, int my, int n_steps, bool want_grad) {
    pnx_desc& d = spec.desc;
    d.N = N;
    d.M[0] = mx;
    d.M[1] = my;
    d.M[2] = 1;
    d.lo[0] = -2e-4;
    d.hi[0] = 0.0;
    d.lo[1] = my > 1 ? -1e-4 : 0.0;
    d.hi[1] = my > 1 ? 1e-4 : 0.0;
    d.n_elements = 2;
    d.n_lines = 2;
    d.n_steps = n_steps;
    d.n_beams = 1;
    d.E_initial = 15000.0;
    d.E_cutoff = 5000.0;
    const int cells = mx * my;
    cell_w.resize(2 * cells);
    for (int c = 0; c < cells; ++c) {
      cell_w[2 * c] = 0.3 + 0.4 * (c % 3) / 2.0;
      cell_w[2 * c + 1] = 1.0 - cell_w[2 * c];
    }
    if (want_grad) grad.assign(2 * cells, 0.0);
    spec.physics = ph;
    spec.cell_w = cell_w.data();
    spec.rho_pure = rho_pure;
    spec.embedding = embedding;
    spec.mac = mac;
    spec.atom_mass = atom_mass;
    spec.line_element = line_element;
    spec.detector[0] = 1e-2;
    spec.detector[2] = 1e-2;
    spec.beam_sigma = 5e-6;
    spec.beam_direction[0] = -1.0;
    spec.beam_kappa = 10.0;
    spec.I_bar = I_bar;
    spec.I_out = I_out.data();
    spec.grad_out = want_grad ? grad.data() : nullptr;
  }
};

static std::vector<std::unique_ptr<Job>> make_jobs(const pnx_physics* ph, int n, unsigned seed) {
  std::mt19937 rng(seed);
  const int orders[3] = {3, 5, 7};
  std::vector<std::unique_ptr<Job>> jobs;
  for (int i = 0; i < n; ++i) {
    const int N = orders[rng() % 3];
    const bool is2d = rng() % 2;
    // 1D: one to three tiles of 256 cells; 2D: small planes
    const int mx = is2d ? 12 + (int)(rng() % 9) : 100 + (int)(rng() % 500);
    const int my = is2d ? 10 + (int)(rng() % 7) : 1;
    jobs.push_back(std::make_unique<Job>(ph, N, mx, my, 10 + (int)(rng() % 8), rng() % 3 != 0));
  }
  return jobs;
}

static pnx_context g_ctx;

// Four client threads submit, poll, cancel and wait; a fifth reads the
// statistics throughout.
static void test_threads() {
  Physics phys;
  pnx_server* s = nullptr;
  CHECK(pnx_server_create(&g_ctx, (int64_t)256 << 20, 64, &s) == PNX_OK);
  if (!s) return;
  constexpr int kClients = 4, kPer = 12;
  std::atomic<int> ok{0}, cancelled{0}, other{0};
  std::atomic<bool> stop{false};
  std::thread stats([&] {
    while (!stop) {
      int64_t t = 0, c = 0, pa = 0, pb = 0;
      CHECK(pnx_server_stats(s, &t, &c, &pa, &pb) == PNX_OK);
      CHECK(c >= 0 && c <= kClients * kPer && pa <= 64);
      std::this_thread::yield();
    }
  });
  std::vector<std::thread> clients;
  for (int ci = 0; ci < kClients; ++ci)
    clients.emplace_back([&, ci] {
      auto jobs = make_jobs(&phys.ph, kPer, 11 + ci);
      std::vector<int64_t> tickets;
      for (int i = 0; i < kPer; ++i) {
        jobs[i]->spec.arrival_s = 1e-4 * i;
        int64_t t = -1;
        CHECK(pnx_server_submit_spec(s, &jobs[i]->spec, &t) == PNX_OK);
        tickets.push_back(t);
      }
      for (int i = 0; i < kPer; ++i) {
        if (i % 5 == 2) (void)pnx_server_cancel(s, tickets[i]);  // pending, running or done: all allowed
        int32_t done = 0;
        pnx_timing tm{};
        CHECK(pnx_server_poll(s, tickets[i], &done, &tm) == PNX_OK);
        CHECK(pnx_server_wait(s, tickets[i], &tm) == PNX_OK);
        if (tm.status == PNX_OK)
          ++ok;
        else if (tm.status == PNX_ERR_CANCELLED)
          ++cancelled;
        else
          ++other;
        CHECK(pnx_server_wait(s, tickets[i], &tm) != PNX_OK);  // the ticket was released
      }
    });
  for (auto& c : clients) c.join();
  CHECK(pnx_server_drain(s) == PNX_OK);
  stop = true;
  stats.join();
  int64_t t = 0, c = 0, pa = 0, pb = 0;
  CHECK(pnx_server_stats(s, &t, &c, &pa, &pb) == PNX_OK);
  CHECK(other == 0);
  CHECK(ok + cancelled == kClients * kPer);
  CHECK(ok > 0);
  CHECK(c >= ok && c <= kClients * kPer);
  CHECK(pb <= ((int64_t)256 << 20));
  CHECK(pnx_server_destroy(s) == PNX_OK);
}

// A kernel launch fails: every waiter is told, and the server is destroyed.
static void test_launch_failure() {
  Physics phys;
  pnx_server* s = nullptr;
  CHECK(pnx_server_create(&g_ctx, (int64_t)256 << 20, 64, &s) == PNX_OK);
  if (!s) return;
  fake_fail_launch(5);
  auto jobs = make_jobs(&phys.ph, 8, 99);
  std::vector<int64_t> tickets;
  for (auto& j : jobs) {
    int64_t t = -1;
    CHECK(pnx_server_submit_spec(s, &j->spec, &t) == PNX_OK);
    tickets.push_back(t);
  }
  int errors = 0;
  for (int64_t t : tickets) {
    pnx_timing tm{};
    const int rc = pnx_server_wait(s, t, &tm);
    CHECK(rc == PNX_OK || rc == PNX_ERR_DEVICE);
    if (rc != PNX_OK) ++errors;
  }
  CHECK(errors > 0);
  CHECK(pnx_server_destroy(s) == PNX_OK);
  fake_fail_launch(0);
}

// A device allocation fails while the server is created: the arena (the
// first allocation) or a coefficient buffer (a later one).
static void test_create_failure() {
  const long which[2] = {1, 4};
  const int expect[2] = {PNX_ERR_MEMORY, PNX_ERR_DEVICE};
  for (int c = 0; c < 2; ++c) {
    pnx_server* s = nullptr;
    fake_fail_allocation(which[c]);
    CHECK(pnx_server_create(&g_ctx, (int64_t)64 << 20, 8, &s) == expect[c]);
    CHECK(s == nullptr);
    fake_fail_allocation(0);
  }
}

int main() {
  g_ctx.device = 0;
  if (hipStreamCreateWithFlags(&g_ctx.stream, 0) != hipSuccess) return 1;
  test_threads();
  test_launch_failure();
  test_create_failure();
  (void)hipStreamDestroy(g_ctx.stream);
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
