// Continuous-batching server of forward+adjoint requests (pnx_server_* in include/pnx.h).
//
// A scheduler thread admits arrived requests in arrival order while their
// device memory fits in the budget (one arena allocated at creation,
// first-fit sub-allocation; nothing is allocated per request on the device)
// and hands each to a lane.  A lane owns one HIP stream and advances its
// requests tick by tick, independently of the other lanes, so that the
// kernels of different lanes overlap on the device:
//   * 1D lane: one march launch per PN order per tick; a workgroup runs a
//     slice of up to kMarchOps operations of one request (march_next), with
//     workgroup barriers between the half steps;
//   * 2D/3D lanes (split by PN order): one operation per request per tick; for
//     every (PN order, dimension) group and every one of the seven step slots
//     one kernel runs over the work list of all requests in that
//     slot (the grouped kernels).
// Every request follows the operation sequence of pnx_forward_adjoint with
// one beam (march_next), with the same device functions, so its results are
// bitwise those of the request solved alone.  A lane keeps at most two ticks
// in flight: tick t's argument blocks go through pinned staging buffer t mod 2,
// reused once tick t-2 has completed; requests complete once the tick that
// copied their results back has completed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "pnx_host.hpp"
#include "pnx_march.hpp"

using namespace pnx;

namespace pnx {
int set_error(int code, const std::string& msg);
}

namespace {

constexpr size_t kAlign = 256;
constexpr int kSlots = 7;      // fwd odd, fwd even, fwd odd, dens, adj odd, adj even, adj odd
constexpr int kMarchOps = 48;  // operations per 1D request per march launch
constexpr int kLanes = 3;      // 1D; 2D/3D of orders <= 5; 2D/3D of orders 7, 9

int lane_of(int N, int dims) { return dims == 1 ? 0 : (N <= 5 ? 1 : 2); }

size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }
// First-fit sub-allocator over one device arena.
class Arena {
 public:
  int init(size_t bytes) {
    if (hipMalloc(&base_, bytes) != hipSuccess) {
      base_ = nullptr;
      (void)hipGetLastError();
      return -1;
    }
    size_ = bytes;
    free_[0] = bytes;
    return 0;
  }
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  // returns offset or SIZE_MAX
  size_t alloc(size_t bytes) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_ += bytes;
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  char* base() const { return base_; }
  size_t used() const { return used_; }
  size_t size() const { return size_; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

// doubles each device buffer of a request needs
struct Sizes {
  long state, ckpt, seg, rho, W, rec, prof, partial, rb, gen, small;
  long total() const { return 2 * state + ckpt + seg + rho + W + 3 * rec + prof + partial + rb + gen + small; }
};

Sizes request_sizes(const Geometry& g, const pnx_desc& d) {
  Sizes s{};
  s.state = g.L.beam_stride;
  s.ckpt = (long)g.nck * s.state;
  s.seg = (long)g.K * s.state;
  s.rho = g.rho_total;
  s.W = (long)d.n_lines * g.L.V;
  s.rec = (long)(d.n_steps + 1) * g.rec_len;
  s.prof = g.prof_stride;
  s.partial = g.partial_beam;
  s.rb = g.rho_total;
  s.gen = (long)d.n_lines * g.L.V;
  s.small = 8 * 16 + 4 * 8;  // mms, I, I_bar
  return s;
}

int default_ckpt(int n) { return std::max(1, (int)std::ceil(std::sqrt((double)n))); }

struct Request {
  int64_t ticket = 0;
  pnx_desc d{};
  Geometry g{};
  Sizes sz{};
  int lane = 0;
  // host copies of the inputs (packed to the device layout at submit)
  std::vector<double> h_rec, h_rho, h_W, h_prof, h_mms, h_ibar;
  // outputs
  double* I_out = nullptr;
  double* rho_bar[8] = {};
  double* gen_out = nullptr;
  double* pinned_out = nullptr;  // results land here, then go to the caller's arrays
  // device
  size_t off = 0, bytes = 0;
  double *state = nullptr, *lam = nullptr, *ckpt = nullptr, *seg = nullptr, *rho = nullptr, *W = nullptr,
         *rec_f = nullptr, *rec_a = nullptr, *rec_d = nullptr, *prof = nullptr, *partial = nullptr, *rb = nullptr,
         *gen = nullptr, *mms = nullptr, *I_dev = nullptr, *ibar = nullptr;
  // schedule
  MarchPos pos = march_start();
  int ops_left = 0;            // operations not yet issued
  int64_t done_tick = -1;      // lane tick that copied the results back
  pnx_timing timing{};
  bool done = false;
};

struct GroupKey {
  int N, dims;
  bool operator<(const GroupKey& o) const { return N != o.N ? N < o.N : dims < o.dims; }
};

struct Lane {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  char* pinned[2] = {nullptr, nullptr};
  size_t pinned_cap[2] = {0, 0};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int64_t ev_tick[2] = {-1, -1};  // tick whose completion ev[b] marks, -1 = none pending
  char* dev_args = nullptr;
  size_t dev_args_cap = 0;
  int64_t tick = 0;          // next tick
  int64_t done_through = -1; // every tick <= this one has completed on the device
  std::vector<std::shared_ptr<Request>> active;
};

}  // namespace

struct pnx_server {
  pnx_context* ctx = nullptr;
  Arena arena;
  int64_t budget = 0;
  int max_active = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<std::shared_ptr<Request>> pending;
  std::map<int64_t, std::shared_ptr<Request>> all;
  int n_active = 0;
  int64_t next_ticket = 1;
  std::atomic<bool> stop{false};
  std::thread worker;
  Lane lanes[kLanes];
  // coefficient buffers per order and direction
  std::map<int, double*> coef_f, coef_a;
  int64_t ticks = 0, completed = 0, peak_active = 0;
  size_t peak_bytes = 0;
  int error = 0;
  std::string error_msg;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

constexpr int kOrders[] = {1, 3, 5, 7, 9};

int upload_coefficients(pnx_server* s) {
  for (int N : kOrders)
    for (int adj = 0; adj < 2; ++adj) {
      std::vector<double> c = build_coefficients(N, adj != 0);
      double* p = nullptr;
      if (hipMalloc(&p, c.size() * sizeof(double)) != hipSuccess) return -1;
      (adj ? s->coef_a : s->coef_f)[N] = p;
      if (hipMemcpy(p, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return -1;
    }
  return 0;
}

const double* coef_for(pnx_server* s, int N, bool adjoint) { return (adjoint ? s->coef_a : s->coef_f).at(N); }

int fail_lane(pnx_server* s, const char* what) {
  (void)hipGetLastError();
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->error) {
    s->error = PNX_ERR_DEVICE;
    s->error_msg = what;
  }
  return -1;
}

// Carve the request's buffers out of its arena block and enqueue uploads + zero
// fills on its lane's stream.
int admit(pnx_server* s, Request& r) {
  hipStream_t st = s->lanes[r.lane].stream;
  const Sizes& z = r.sz;
  double* p = reinterpret_cast<double*>(s->arena.base() + r.off);
  auto take = [&](long n) {
    double* q = p;
    p += (align_up(n * sizeof(double)) / sizeof(double));
    return q;
  };
  r.state = take(z.state);
  r.lam = take(z.state);
  r.ckpt = take(z.ckpt);
  r.seg = take(z.seg);
  r.rho = take(z.rho);
  r.W = take(z.W);
  r.rec_f = take(z.rec);
  r.rec_a = take(z.rec);
  r.rec_d = take(z.rec);
  r.prof = take(z.prof);
  r.partial = take(z.partial);
  r.rb = take(z.rb);
  r.gen = take(z.gen);
  r.mms = take(8 * 16);
  r.I_dev = take(8);
  r.ibar = take(8);
  const long rec = z.rec;
  auto up = [&](double* dst, const double* src, size_t n) {
    return hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess;
  };
  if (!up(r.rec_f, r.h_rec.data(), rec) || !up(r.rec_a, r.h_rec.data() + rec, rec) ||
      !up(r.rec_d, r.h_rec.data() + 2 * rec, rec) || !up(r.rho, r.h_rho.data(), r.h_rho.size()) ||
      !up(r.W, r.h_W.data(), r.h_W.size()) || !up(r.prof, r.h_prof.data(), r.h_prof.size()) ||
      !up(r.mms, r.h_mms.data(), r.h_mms.size()) || !up(r.ibar, r.h_ibar.data(), r.h_ibar.size()))
    return -1;
  // the recompute writes whole segment slots; ghost slots are never read for
  // values, but keep them finite
  if (launch_fill(r.state, z.state, 0.0, st) != hipSuccess || launch_fill(r.lam, z.state, 0.0, st) != hipSuccess ||
      launch_fill(r.rb, z.rb, 0.0, st) != hipSuccess || launch_fill(r.gen, z.gen, 0.0, st) != hipSuccess ||
      launch_fill(r.seg, z.seg, 0.0, st) != hipSuccess)
    return -1;
  r.pos = march_start();
  r.ops_left = march_ops(r.d.n_steps);
  return 0;
}

HalfStepArgs make_half(pnx_server* s, const Request& r, int mode, const double* rec, double dt, const double* from,
                       const double* src, double* dst, double* partial) {
  HalfStepArgs a{};
  a.L = r.g.L;
  a.from = from;
  a.to_src = src;
  a.to_dst = dst;
  a.rho = r.rho;
  for (int c = 0; c < 8; ++c) a.rho_off[c] = r.g.rho_off[c];
  a.coef = coef_for(s, r.d.N, mode == kAdjoint);
  a.rec = rec;
  a.prof = r.prof;
  a.prof_stride = r.g.prof_stride;
  a.mms = r.mms;
  a.W = r.W;
  a.ibar = mode == kAdjoint ? r.ibar : nullptr;
  a.partial = partial;
  a.partial_stride = r.g.partial_beam;
  a.dt = dt;
  a.n_beams = 1;
  return a;
}

DensArgs make_dens(pnx_server* s, const Request& r, int i, const double* u, double int_fac) {
  DensArgs a{};
  a.L = r.g.L;
  a.u = u;
  a.lam = r.lam;
  a.rho = r.rho;
  for (int c = 0; c < 8; ++c) {
    a.rho_off[c] = r.g.rho_off[c];
    a.rb_off[c] = r.g.rho_off[c];
  }
  a.coef = coef_for(s, r.d.N, false);
  a.rec = r.rec_d + (long)i * r.g.rec_len;
  a.prof = r.prof;
  a.prof_stride = r.g.prof_stride;
  a.mms = r.mms;
  a.W = r.W;
  a.ibar = r.ibar;
  a.rho_bar = r.rb;
  a.rb_stride = r.g.rho_total;
  a.gen = r.gen;
  a.gen_stride = (long)r.d.n_lines * r.g.L.V;
  a.dE = r.g.dE;
  a.int_fac = int_fac;
  a.n_beams = 1;
  return a;
}

MarchArgs make_march(pnx_server* s, const Request& r) {
  const Geometry& g = r.g;
  MarchArgs m{};
  m.fwd = make_half(s, r, kForward, r.rec_f, g.dE, r.state, r.state, r.state, r.partial);
  m.adj = make_half(s, r, kAdjoint, r.rec_a, g.dE, r.lam, r.lam, r.lam, nullptr);
  m.dens = make_dens(s, r, 0, r.ckpt, 1.0);
  m.state = r.state;
  m.lam = r.lam;
  m.ckpt = r.ckpt;
  m.seg = r.seg;
  m.partial = r.partial;
  m.I_out = r.I_dev;
  m.rec_f = r.rec_f;
  m.rec_a = r.rec_a;
  m.rec_d = r.rec_d;
  m.partial_step = g.partial_step;
  m.rec_len = g.rec_len;
  m.sigma_off = g.sigma_off;
  m.n = r.d.n_steps;
  m.K = g.K;
  m.nck = g.nck;
  return m;
}

double* buffer_of(const Request& r, int buf, int idx) {
  const long st = r.g.L.beam_stride;
  if (buf == kBufCkpt) return r.ckpt + (long)idx * st;
  if (buf == kBufSeg) return r.seg + (long)idx * st;
  return r.state;
}

struct SlotWork {
  std::vector<HalfStepArgs> half;
  std::vector<DensArgs> dens;
  std::vector<int2> items;
};

struct TickPlan {
  std::map<GroupKey, SlotWork[kSlots]> groups;      // grid lanes
  std::map<int, std::vector<MarchArgs>> march;      // 1D lane, per order
  std::vector<std::pair<Request*, int>> copies;     // (request, checkpoint index) copies before the slots
  std::vector<Request*> reduce;                     // forward finished: reduce intensities after slot 2
  std::vector<Request*> copy_out;                   // results to copy back at the end of the tick
};

void add_half(SlotWork& w, const HalfStepArgs& a) {
  const int idx = (int)w.half.size();
  w.half.push_back(a);
  for (int t = 0; t < a.L.n_tiles; ++t) w.items.push_back(make_int2(idx, t));
}
void add_dens(SlotWork& w, const DensArgs& a) {
  const int idx = (int)w.dens.size();
  w.dens.push_back(a);
  for (int t = 0; t < a.L.n_tiles; ++t) w.items.push_back(make_int2(idx, t));
}

// Plan the next operation of request r on a grid lane.
void plan_grid(pnx_server* s, Request& r, TickPlan& plan) {
  const Geometry& g = r.g;
  SlotWork* w = plan.groups[GroupKey{r.d.N, g.L.dims}];
  const double h = g.dE / 2;
  const MarchOp o = march_next(r.pos, r.d.n_steps, g.K, g.nck);
  r.ops_left -= 1;
  const double* src = buffer_of(r, o.src_buf, o.src_idx);
  switch (o.kind) {
    case kOpFwd:
    case kOpRecompute: {
      if (o.ckpt_copy >= 0) plan.copies.push_back({&r, o.ckpt_copy});
      double* dst = buffer_of(r, o.dst_buf, o.dst_idx);
      const double* rec = r.rec_f + (long)o.step * g.rec_len;
      // the recompute's intensity partials go to the scratch slot of step 0
      double* part = o.kind == kOpFwd ? r.partial + (long)o.step * g.partial_step : r.partial;
      add_half(w[0], make_half(s, r, kForward, rec, h, src, src, dst, nullptr));
      add_half(w[1], make_half(s, r, kForward, rec, g.dE, dst, src, dst, part));
      add_half(w[2], make_half(s, r, kForward, rec, h, dst, dst, dst, nullptr));
      if (o.reduce) plan.reduce.push_back(&r);
      break;
    }
    case kOpReverse: {
      add_dens(w[3], make_dens(s, r, o.step, src, 1.0));
      const double* rec = r.rec_a + (long)o.step * g.rec_len;
      add_half(w[4], make_half(s, r, kAdjoint, rec, h, r.lam, r.lam, r.lam, nullptr));
      add_half(w[5], make_half(s, r, kAdjoint, rec, g.dE, r.lam, r.lam, r.lam, nullptr));
      add_half(w[6], make_half(s, r, kAdjoint, rec, h, r.lam, r.lam, r.lam, nullptr));
      break;
    }
    case kOpFinal:
      add_dens(w[3], make_dens(s, r, 0, src, 0.5));
      plan.copy_out.push_back(&r);
      break;
    default:
      break;
  }
  r.timing.ticks += 1;
}

// Plan the next slice of operations of request r on the 1D lane.
void plan_march(pnx_server* s, Request& r, TickPlan& plan) {
  MarchArgs m = make_march(s, r);
  m.pos = r.pos;
  m.n_ops = std::min(kMarchOps, r.ops_left);
  for (int i = 0; i < m.n_ops; ++i) (void)march_next(r.pos, r.d.n_steps, r.g.K, r.g.nck);
  r.ops_left -= m.n_ops;
  plan.march[r.d.N].push_back(m);
  if (r.ops_left == 0) plan.copy_out.push_back(&r);
  r.timing.ticks += 1;
}

int ensure_pinned(Lane& ln, int b, size_t bytes) {
  if (ln.pinned_cap[b] >= bytes) return 0;
  if (ln.pinned[b]) (void)hipHostFree(ln.pinned[b]);
  ln.pinned[b] = nullptr;
  ln.pinned_cap[b] = 0;
  const size_t cap = std::max(bytes, (size_t)1 << 20) * 2;
  if (hipHostMalloc(&ln.pinned[b], cap, hipHostMallocDefault) != hipSuccess) return -1;
  ln.pinned_cap[b] = cap;
  return 0;
}

int ensure_dev_args(Lane& ln, size_t bytes) {
  if (ln.dev_args_cap >= bytes) return 0;
  // the lane's earlier ticks read the old buffer
  if (hipStreamSynchronize(ln.stream) != hipSuccess) return -1;
  if (ln.dev_args) (void)hipFree(ln.dev_args);
  ln.dev_args = nullptr;
  ln.dev_args_cap = 0;
  const size_t cap = std::max(bytes, (size_t)1 << 20) * 2;
  if (hipMalloc(&ln.dev_args, cap) != hipSuccess) return -1;
  ln.dev_args_cap = cap;
  return 0;
}

// Serialise the plan's argument blocks and work lists into staging buffer
// `buf`, upload them, and enqueue the tick on the lane's stream.
int launch_plan(Lane& ln, TickPlan& plan, int buf) {
  hipStream_t st = ln.stream;
  struct Entry {
    GroupKey key;
    int slot;  // -1: march block of order key.N
    size_t n_off, args_off, items_off;
    int n_items;
  };
  std::vector<Entry> entries;
  size_t off = 0;
  for (auto& kv : plan.march) {
    Entry e{GroupKey{kv.first, 1}, -1, 0, 0, 0, (int)kv.second.size()};
    e.args_off = off;
    off = align_up(off + kv.second.size() * sizeof(MarchArgs));
    entries.push_back(e);
  }
  for (auto& kv : plan.groups)
    for (int sl = 0; sl < kSlots; ++sl) {
      SlotWork& w = kv.second[sl];
      if (w.items.empty()) continue;
      Entry e{kv.first, sl, 0, 0, 0, (int)w.items.size()};
      e.n_off = off;
      off = align_up(off + sizeof(int));
      e.args_off = off;
      off = align_up(off + (sl == 3 ? w.dens.size() * sizeof(DensArgs) : w.half.size() * sizeof(HalfStepArgs)));
      e.items_off = off;
      off = align_up(off + w.items.size() * sizeof(int2));
      entries.push_back(e);
    }
  if (off == 0) off = kAlign;
  if (ensure_pinned(ln, buf, off) || ensure_dev_args(ln, off)) return -1;
  char* h = ln.pinned[buf];
  for (const Entry& e : entries) {
    if (e.slot < 0) {
      const auto& v = plan.march[e.key.N];
      std::memcpy(h + e.args_off, v.data(), v.size() * sizeof(MarchArgs));
      continue;
    }
    SlotWork& w = plan.groups[e.key][e.slot];
    std::memcpy(h + e.n_off, &e.n_items, sizeof(int));
    if (e.slot == 3)
      std::memcpy(h + e.args_off, w.dens.data(), w.dens.size() * sizeof(DensArgs));
    else
      std::memcpy(h + e.args_off, w.half.data(), w.half.size() * sizeof(HalfStepArgs));
    std::memcpy(h + e.items_off, w.items.data(), w.items.size() * sizeof(int2));
  }
  if (hipMemcpyAsync(ln.dev_args, h, off, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
  for (auto& cp : plan.copies) {
    Request& r = *cp.first;
    const long stn = r.g.L.beam_stride;
    if (hipMemcpyAsync(r.ckpt + (long)cp.second * stn, r.state, stn * sizeof(double), hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      return -1;
  }
  for (const Entry& e : entries)
    if (e.slot < 0 &&
        launch_march(e.key.N, reinterpret_cast<const MarchArgs*>(ln.dev_args + e.args_off), e.n_items, st) !=
            hipSuccess)
      return -1;
  auto launch_slot = [&](int slot) -> int {
    for (const Entry& e : entries) {
      if (e.slot != slot) continue;
      const int grid = (e.n_items + 7) & ~7;  // one block per item, a multiple of 8 for the XCD mapping
      const int* nptr = reinterpret_cast<const int*>(ln.dev_args + e.n_off);
      const int2* items = reinterpret_cast<const int2*>(ln.dev_args + e.items_off);
      hipError_t err;
      if (slot == 3) {
        err = launch_dens_grouped(e.key.N, e.key.dims, reinterpret_cast<const DensArgs*>(ln.dev_args + e.args_off),
                                  items, nptr, grid, st);
      } else {
        const int mode = slot < 3 ? kForward : kAdjoint;
        const int parity = (slot == 1 || slot == 5) ? 0 : 1;
        err = launch_half_step_grouped(e.key.N, e.key.dims, parity, mode,
                                       reinterpret_cast<const HalfStepArgs*>(ln.dev_args + e.args_off), items, nptr,
                                       grid, st);
      }
      if (err != hipSuccess) return -1;
    }
    return 0;
  };
  for (int sl = 0; sl < 3; ++sl)
    if (launch_slot(sl)) return -1;
  for (Request* r : plan.reduce) {
    const Geometry& g = r->g;
    if (launch_reduce_intensity(r->partial, g.partial_step, g.partial_beam, g.L.n_tiles, r->rec_f, g.rec_len,
                                g.sigma_off, r->d.n_steps, r->d.n_lines, g.dE, g.L.dvol, 1, r->I_dev, st) !=
        hipSuccess)
      return -1;
  }
  for (int sl = 3; sl < kSlots; ++sl)
    if (launch_slot(sl)) return -1;
  for (Request* r : plan.copy_out) {
    const long nI = r->d.n_lines, nrb = r->sz.rb, ngen = r->sz.gen;
    if (hipMemcpyAsync(r->pinned_out, r->I_dev, nI * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(r->pinned_out + nI, r->rb, nrb * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(r->pinned_out + nI + nrb, r->gen, ngen * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess)
      return -1;
    r->done_tick = ln.tick;
  }
  return 0;
}

void finish_request(pnx_server* s, Request& r) {
  const long nI = r.d.n_lines, nrb = r.sz.rb;
  const double* out = r.pinned_out;
  if (r.I_out) std::memcpy(r.I_out, out, nI * sizeof(double));
  for (int c = 0; c < 8; ++c) {
    if (!r.rho_bar[c]) continue;
    if (r.g.active(c)) {
      unpack_class(r.g, c, out + nI + r.g.rho_off[c], r.d.n_elements, r.rho_bar[c]);
    } else {
      const long np = (long)r.g.class_n[c][0] * r.g.class_n[c][1] * r.g.class_n[c][2];
      std::fill(r.rho_bar[c], r.rho_bar[c] + np * r.d.n_elements, 0.0);
    }
  }
  if (r.gen_out) unpack_lines(r.g, out + nI + nrb, r.gen_out);
  (void)hipHostFree(r.pinned_out);
  r.pinned_out = nullptr;
  s->arena.release(r.off, r.bytes);
  r.timing.done_s = s->now();
  r.timing.status = PNX_OK;
  r.done = true;
  std::vector<double>().swap(r.h_rec);
  std::vector<double>().swap(r.h_rho);
  std::vector<double>().swap(r.h_W);
  std::vector<double>().swap(r.h_prof);
}

// Records which of the lane's in-flight ticks have completed.
int poll_lane(Lane& ln) {
  for (int b = 0; b < 2; ++b) {
    if (ln.ev_tick[b] < 0) continue;
    const hipError_t e = hipEventQuery(ln.ev[b]);
    if (e == hipSuccess) {
      ln.done_through = std::max(ln.done_through, ln.ev_tick[b]);
      ln.ev_tick[b] = -1;
    } else if (e != hipErrorNotReady) {
      return -1;
    }
  }
  return 0;
}

// Plans and enqueues the lane's next tick if its staging buffer is free.
// Returns 1 if a tick was launched, 0 if not, -1 on error.
int step_lane(pnx_server* s, int li) {
  Lane& ln = s->lanes[li];
  const int buf = (int)(ln.tick & 1);
  if (ln.ev_tick[buf] >= 0) return 0;  // tick - 2 still in flight
  TickPlan plan;
  bool any = false;
  for (auto& r : ln.active) {
    if (r->ops_left == 0) continue;
    any = true;
    if (li == 0)
      plan_march(s, *r, plan);
    else
      plan_grid(s, *r, plan);
  }
  if (!any) return 0;
  if (launch_plan(ln, plan, buf)) return -1;
  if (hipEventRecord(ln.ev[buf], ln.stream) != hipSuccess) return -1;
  ln.ev_tick[buf] = ln.tick;
  ln.tick += 1;
  return 1;
}

void worker_loop(pnx_server* s) {
  (void)hipSetDevice(s->ctx->device);
  while (true) {
    std::vector<std::shared_ptr<Request>> admitted;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->error) break;
      const double now = s->now();
      // admission in arrival order while memory and slots allow
      while (!s->pending.empty() && s->n_active < s->max_active) {
        auto& r = s->pending.front();
        if (r->timing.arrival_s > now) break;
        const size_t off = s->arena.alloc(r->bytes);
        if (off == SIZE_MAX) break;
        r->off = off;
        r->timing.admit_s = now;
        admitted.push_back(r);
        s->n_active += 1;
        s->pending.pop_front();
      }
      if (s->n_active == 0) {
        if (s->stop && s->pending.empty()) break;
        if (s->pending.empty()) {
          s->cv.wait_for(lk, std::chrono::milliseconds(5));
        } else {
          const double wait = s->pending.front()->timing.arrival_s - s->now();
          if (wait > 0) s->cv.wait_for(lk, std::chrono::microseconds((int64_t)(std::min(wait, 0.005) * 1e6) + 1));
        }
        continue;
      }
      s->peak_active = std::max<int64_t>(s->peak_active, (int64_t)s->n_active);
      s->peak_bytes = std::max(s->peak_bytes, s->arena.used());
    }
    for (auto& r : admitted) {
      if (admit(s, *r)) {
        fail_lane(s, "request upload failed");
        break;
      }
      s->lanes[r->lane].active.push_back(r);
    }
    bool launched = false;
    std::vector<std::shared_ptr<Request>> finished;
    for (int li = 0; li < kLanes; ++li) {
      Lane& ln = s->lanes[li];
      if (ln.active.empty()) continue;
      if (poll_lane(ln)) {
        fail_lane(s, "device error in a served tick");
        break;
      }
      // requests whose copy-out tick has completed
      auto it = std::partition(ln.active.begin(), ln.active.end(), [&](const std::shared_ptr<Request>& r) {
        return !(r->done_tick >= 0 && r->done_tick <= ln.done_through);
      });
      finished.insert(finished.end(), it, ln.active.end());
      ln.active.erase(it, ln.active.end());
      const int rc = step_lane(s, li);
      if (rc < 0) {
        fail_lane(s, "tick launch failed");
        break;
      }
      launched |= rc > 0;
    }
    if (!finished.empty()) {
      {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto& r : finished) {
          finish_request(s, *r);
          s->completed += 1;
          s->n_active -= 1;
        }
      }
      s->done_cv.notify_all();
    }
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->ticks = s->lanes[0].tick + s->lanes[1].tick + s->lanes[2].tick;
    }
    if (!launched && finished.empty()) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  for (Lane& ln : s->lanes)
    if (ln.stream) (void)hipStreamSynchronize(ln.stream);
  s->done_cv.notify_all();
}

}  // namespace

extern "C" {

int64_t pnx_request_device_bytes(const pnx_desc* desc) {
  if (!desc) return -1;
  pnx_desc d = *desc;
  if (d.ckpt_every <= 0) d.ckpt_every = default_ckpt(d.n_steps);
  Geometry g;
  std::string err;
  if (make_geometry(d, g, err)) return set_error(PNX_ERR_ARG, err);
  Sizes z = request_sizes(g, d);
  // each buffer is aligned separately
  return (int64_t)(z.total() * sizeof(double) + 20 * kAlign);
}

static void release_server(pnx_server* s) {
  for (auto& kv : s->coef_f) (void)hipFree(kv.second);
  for (auto& kv : s->coef_a) (void)hipFree(kv.second);
  for (Lane& ln : s->lanes) {
    for (int b = 0; b < 2; ++b) {
      if (ln.pinned[b]) (void)hipHostFree(ln.pinned[b]);
      if (ln.ev[b]) (void)hipEventDestroy(ln.ev[b]);
    }
    if (ln.dev_args) (void)hipFree(ln.dev_args);
    if (ln.stream && ln.own_stream) (void)hipStreamDestroy(ln.stream);
  }
  for (auto& kv : s->all)
    if (kv.second->pinned_out) (void)hipHostFree(kv.second->pinned_out);
  delete s;
}

int pnx_server_create(pnx_context* ctx, int64_t hbm_budget_bytes, int32_t max_active, pnx_server** out) {
  if (!ctx || !out || hbm_budget_bytes <= 0 || max_active < 1) return set_error(PNX_ERR_ARG, "bad argument");
  if (hipSetDevice(ctx->device) != hipSuccess) return set_error(PNX_ERR_DEVICE, "hipSetDevice failed");
  auto* s = new pnx_server;
  s->ctx = ctx;
  s->budget = hbm_budget_bytes;
  s->max_active = max_active;
  if (s->arena.init((size_t)hbm_budget_bytes)) {
    release_server(s);
    return set_error(PNX_ERR_MEMORY, "could not allocate the server arena");
  }
  if (upload_coefficients(s)) {
    release_server(s);
    return set_error(PNX_ERR_DEVICE, "coefficient upload failed");
  }
  // The 1D lane runs on the context's stream: with the runtime's default of
  // four hardware queues per process, two more streams leave every lane a
  // queue of its own next to the default stream.
  for (int li = 0; li < kLanes; ++li) {
    Lane& ln = s->lanes[li];
    if (li == 0) {
      ln.stream = ctx->stream;
    } else if (hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking) != hipSuccess) {
      ln.stream = nullptr;
      release_server(s);
      return set_error(PNX_ERR_DEVICE, "hipStreamCreate failed");
    } else {
      ln.own_stream = true;
    }
    for (int b = 0; b < 2; ++b)
      if (hipEventCreateWithFlags(&ln.ev[b], hipEventDisableTiming) != hipSuccess) {
        ln.ev[b] = nullptr;
        release_server(s);
        return set_error(PNX_ERR_DEVICE, "hipEventCreate failed");
      }
  }
  s->t0 = std::chrono::steady_clock::now();
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return PNX_OK;
}

int pnx_server_destroy(pnx_server* s) {
  if (!s) return PNX_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->ctx->device);
  release_server(s);
  return PNX_OK;
}

double pnx_server_clock(const pnx_server* s) { return s ? s->now() : 0.0; }

int pnx_server_submit(pnx_server* s, const pnx_request* q, int64_t* ticket) {
  if (!s || !q || !ticket) return set_error(PNX_ERR_ARG, "null argument");
  if (q->desc.n_beams != 1) return set_error(PNX_ERR_ARG, "served requests carry exactly one beam");
  if (!q->fwd || !q->adj || !q->dens || !q->W || !q->I_bar || !q->I_out)
    return set_error(PNX_ERR_ARG, "records, line weights, seed and output are required");
  auto r = std::make_shared<Request>();
  r->d = q->desc;
  if (r->d.ckpt_every <= 0) r->d.ckpt_every = default_ckpt(r->d.n_steps);
  std::string err;
  int rc = make_geometry(r->d, r->g, err);
  if (rc) return set_error(rc, err);
  r->sz = request_sizes(r->g, r->d);
  r->lane = lane_of(r->d.N, r->g.L.dims);
  r->bytes = (size_t)r->sz.total() * sizeof(double) + 20 * kAlign;
  if ((int64_t)r->bytes > s->budget) return set_error(PNX_ERR_MEMORY, "request exceeds the server's HBM budget");
  const Geometry& g = r->g;
  const long rec = r->sz.rec;
  r->h_rec.resize(3 * rec);
  std::memcpy(r->h_rec.data(), q->fwd, rec * sizeof(double));
  std::memcpy(r->h_rec.data() + rec, q->adj, rec * sizeof(double));
  std::memcpy(r->h_rec.data() + 2 * rec, q->dens, rec * sizeof(double));
  r->h_rho.assign(g.rho_total, 0.0);
  for (int c = 0; c < 8; ++c) {
    if (!g.active(c)) continue;
    if (!q->rho[c]) return set_error(PNX_ERR_ARG, "densities missing for an updated class");
    pack_class(g, c, q->rho[c], r->d.n_elements, r->h_rho.data() + g.rho_off[c]);
  }
  r->h_W.assign((size_t)r->d.n_lines * g.L.V, 0.0);
  pack_lines(g, q->W, r->h_W.data());
  r->h_prof.assign(g.prof_stride, 0.0);
  for (int c = 0; c < 8; ++c)
    if (odd_in(c, 0) && g.active(c) && (!q->ppos[c] || !q->mms[c]))
      return set_error(PNX_ERR_ARG, "beam data missing for an x-odd class");
  pack_profiles(g, q->ppos, r->h_prof.data());
  r->h_mms.assign(8 * 16, 0.0);
  for (int c = 0; c < 8; ++c)
    if (odd_in(c, 0) && g.active(c))
      for (int m = 0; m < g.T.n[c]; ++m) r->h_mms[c * 16 + m] = q->mms[c][m];
  r->h_ibar.assign(8, 0.0);
  for (int l = 0; l < r->d.n_lines; ++l) r->h_ibar[l] = q->I_bar[l];
  r->I_out = q->I_out;
  for (int c = 0; c < 8; ++c) r->rho_bar[c] = q->rho_bar[c];
  r->gen_out = q->gen_out;
  const size_t nout = (size_t)(r->d.n_lines + r->sz.rb + r->sz.gen);
  if (hipHostMalloc(&r->pinned_out, nout * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return set_error(PNX_ERR_MEMORY, "pinned result buffer allocation failed");
  r->timing.arrival_s = q->arrival_s;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    // keep the pending queue ordered by arrival time
    auto it = s->pending.end();
    while (it != s->pending.begin() && (*std::prev(it))->timing.arrival_s > q->arrival_s) --it;
    s->pending.insert(it, r);
    s->all[r->ticket] = r;
    *ticket = r->ticket;
  }
  s->cv.notify_all();
  return PNX_OK;
}

int pnx_server_poll(pnx_server* s, int64_t ticket, int32_t* done, pnx_timing* t) {
  if (!s || !done) return set_error(PNX_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return set_error(PNX_ERR_ARG, "unknown ticket");
  *done = it->second->done ? 1 : 0;
  if (t) *t = it->second->timing;
  if (s->error) return set_error(s->error, s->error_msg);
  return PNX_OK;
}

int pnx_server_wait(pnx_server* s, int64_t ticket, pnx_timing* t) {
  if (!s) return set_error(PNX_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return set_error(PNX_ERR_ARG, "unknown ticket");
  auto r = it->second;
  s->done_cv.wait(lk, [&] { return r->done || s->error != 0; });
  if (t) *t = r->timing;
  if (s->error) return set_error(s->error, s->error_msg);
  return PNX_OK;
}

int pnx_server_drain(pnx_server* s) {
  if (!s) return set_error(PNX_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->done_cv.wait(lk, [&] { return (s->pending.empty() && s->n_active == 0) || s->error != 0; });
  if (s->error) return set_error(s->error, s->error_msg);
  return PNX_OK;
}

int pnx_server_stats(const pnx_server* s, int64_t* ticks, int64_t* completed, int64_t* peak_active,
                     int64_t* peak_bytes) {
  if (!s) return set_error(PNX_ERR_ARG, "null server");
  if (ticks) *ticks = s->ticks;
  if (completed) *completed = s->completed;
  if (peak_active) *peak_active = s->peak_active;
  if (peak_bytes) *peak_bytes = (int64_t)s->peak_bytes;
  return PNX_OK;
}

}  // extern "C"
