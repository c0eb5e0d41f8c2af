// Continuous-batching server of forward+adjoint requests (pnx_server_* in include/pnx.h).
//
// A scheduler thread admits arrived requests in arrival order while their
// device memory fits in the budget (one arena allocated at creation,
// first-fit sub-allocation; nothing is allocated per request on the device)
// and hands each to a lane.  A lane owns one HIP stream and advances its
// requests tick by tick, independently of the other lanes, so that the
// kernels of different lanes overlap on the device:
//   * 1D lane: one march launch per tick for all PN orders; a workgroup runs a
//     slice of operations of one request (march_next; more for lower orders), with
//     workgroup barriers between the half steps;
//   * 2D/3D lanes (split by PN order): one operation per request per tick; for
//     every (PN order, dimension) group and every launch phase one kernel runs
//     over the work list of all requests' passes in that phase (the grouped
//     and phase kernels).  kGridTicks ticks are planned and uploaded at once, so
//     the argument upload's latency is paid once per kGridTicks ticks.
// Every request follows the operation sequence of pnx_forward_adjoint with
// one beam (march_next), with the same device functions, so its results are
// bitwise those of the request solved alone.  A lane keeps at most two
// launch batches ("lane ticks") in flight: batch t's argument blocks go through
// pinned staging buffer t mod 2, reused once batch t-2 has completed; requests
// complete once the batch that copied their results back has completed.
// Request inputs are staged at submit in a pinned block laid out like the
// device buffers, so admission is one upload and one zero fill.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "pnx_admission.hpp"
#include "pnx_host.hpp"
#include "pnx_march.hpp"

using namespace pnx;

namespace pnx {
int set_error(int code, const std::string& msg);
}

namespace {

// Launch phases of a sub-tick, in launch order.  Each request's operation
// puts one pass in each phase it uses, so the phases are its passes in order,
// and requests in one phase launch are independent:
//   first odd:  the forward step's first odd half step (first step of a run);
//   phase A:    forward even half step | fused reverse pass;
//   pair:       fused odd pair (a step that is not the last of its run);
//   phase B:    last odd half step of a run | adjoint even half step.
// (The pair and phase B are independent of each other.)
enum Phase : int { kPhaseFirstOdd, kPhaseA, kPhasePair, kPhaseB, kPhases };
constexpr int kMarchOps = 48;  // operations per P9 1D request per march launch
// Operations per march launch by PN order: one launch serves every order and
// lasts as long as its slowest request, and an operation's time grows with the
// moment count (1D: 6, 12, 20, 30 for P3-P9), so lower orders take more steps.
int march_ops_per_launch(int N) {
  return N <= 3 ? 5 * kMarchOps : (N == 5 ? 5 * kMarchOps / 2 : (N == 7 ? 3 * kMarchOps / 2 : kMarchOps));
}
// A march launch with a request of two or more tiles runs kMarchWide-thread
// workgroups (two tiles of a pass side by side) while it holds at most this
// many requests.  A wide workgroup takes a whole CU's registers, against half
// a CU for a kBlock one, and a two-tile request runs 1.4x faster in it, so it
// costs about 1.4x the CU time (and a one-tile request in a wide launch 2x);
// under a large 1D load the narrow form is kept.  On the 4096-request serving
// stream the wide form halves the 1D requests' p50 and p99 latency for about
// 1 % of the throughput.
constexpr int kMarchWideMax = 16;
constexpr int kGridTicks = 4;  // operations per 2D/3D request per grid-lane launch batch
// Active requests that keep a lane's launches full: beyond these a request
// that cannot be taped waits for memory rather than run checkpointed.
constexpr int kBusyGridLane = 32;
constexpr int kBusyMarchLane = 128;
// 1D, and up to four 2D/3D lanes.  (Six lanes, with orders 7 and 9 split by
// grid size, measured slower than four: the grouped launches get smaller.)
constexpr int kMaxLanes = 5;

// Lane of a request.  With two 2D/3D lanes: orders <= 5 and orders 7, 9; with
// three: 1-3, 5-7, 9; with four: 1-3, 5, 7, 9.
int lane_of(int N, int dims, int grid_lanes) {
  if (dims == 1) return 0;
  switch (grid_lanes) {
    case 1: return 1;
    case 2: return N <= 5 ? 1 : 2;
    case 3: return N <= 3 ? 1 : (N <= 7 ? 2 : 3);
    default: return N <= 3 ? 1 : (N == 5 ? 2 : (N == 7 ? 3 : 4));
  }
}

// 2D/3D lanes: one stream each, besides the context stream (1D lane) and the
// default stream, so that every stream has a hardware queue of its own (the
// runtime gives a process GPU_MAX_HW_QUEUES queues, 4 by default; streams
// beyond that share queues and their kernels serialise).
int grid_lanes_for_queues() {
  int q = 4;
  if (const char* e = std::getenv("GPU_MAX_HW_QUEUES")) q = std::atoi(e);
  return std::max(1, std::min(kMaxLanes - 1, q - 2));
}

// The device arena: one allocation, sub-allocated first-fit.
class Arena {
 public:
  int init(size_t bytes) {
    if (hipMalloc(&base_, bytes) != hipSuccess) {
      base_ = nullptr;
      (void)hipGetLastError();
      return -1;
    }
    fl_ = FreeList(bytes);
    return 0;
  }
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  size_t alloc(size_t bytes) { return fl_.alloc(bytes); }
  void release(size_t off, size_t bytes) { fl_.release(off, bytes); }
  char* base() const { return base_; }
  size_t used() const { return fl_.used(); }

 private:
  char* base_ = nullptr;
  FreeList fl_;
};

// Pinned host memory for request staging, in chunks that are sub-allocated
// and kept: pinning and unpinning pages per request would cost milliseconds
// in the submitting thread and in the scheduler.  Thread-safe.
class PinnedPool {
 public:
  ~PinnedPool() {
    for (auto& c : chunks_) (void)hipHostFree(c.base);
  }
  double* alloc(size_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    bytes = align_up(bytes);
    for (auto& c : chunks_) {
      const size_t off = c.fl.alloc(bytes);
      if (off != SIZE_MAX) return reinterpret_cast<double*>(c.base + off);
    }
    const size_t cap = std::max(bytes, kChunk);
    char* p = nullptr;
    if (hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess) return nullptr;
    chunks_.push_back(Chunk{p, cap, FreeList(cap)});
    return reinterpret_cast<double*>(p + chunks_.back().fl.alloc(bytes));
  }
  void release(double* ptr, size_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    char* p = reinterpret_cast<char*>(ptr);
    for (auto& c : chunks_)
      if (p >= c.base && p < c.base + c.size) {
        c.fl.release((size_t)(p - c.base), bytes);
        return;
      }
  }

 private:
  static constexpr size_t kChunk = (size_t)256 << 20;
  struct Chunk {
    char* base;
    size_t size;
    FreeList fl;
  };
  std::mutex mu_;
  std::vector<Chunk> chunks_;
};

// doubles each device buffer of a request needs
struct Sizes {
  long state, ckpt, seg, tape, rho, W, rec, prof, partial, rb, gen, small;
  // requests given as cell fractions (pnx_server_submit_spec): the fractions,
  // the absorption inputs, the absorption table and factors, the gradient
  long cell, abs, mu, ac, grad, extra;
};

// Doubles of the absorption inputs of a request given as cell fractions:
// even points and box edges per axis, mac, atomic masses, embedding,
// pure-element densities, and the line elements (eight int32 in four doubles).
long abs_inputs(const Geometry& g, const pnx_desc& d) {
  long n = 0;
  for (int a = 0; a < 3; ++a) n += 2L * d.M[a] + 1;
  return n + (long)d.n_lines * d.n_elements + 3L * d.n_elements + 4;
}

Sizes request_sizes(const Geometry& g, const pnx_desc& d, bool spec = false) {
  Sizes s{};
  if (spec) {
    const long n_cells = (long)d.M[0] * d.M[1] * d.M[2];
    s.cell = n_cells * d.n_elements;
    s.abs = abs_inputs(g, d);
    s.mu = (n_cells + 1) * d.n_lines;
    s.ac = (long)d.n_lines * g.L.V;
    s.grad = n_cells * d.n_elements;
    s.extra = (long)d.n_elements * g.L.V;
  }
  s.state = g.L.beam_stride;
  s.ckpt = (long)g.nck * s.state;
  s.seg = (long)g.K * s.state;
  s.tape = 0;
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

// The same request taped: no checkpoints or recompute slots, the n forward states instead.
Sizes taped_sizes(const Sizes& z, int n) {
  Sizes s = z;
  s.ckpt = 0;
  s.seg = 0;
  s.tape = (long)n * z.state;
  return s;
}

// Offsets (doubles, each buffer 256-byte aligned) of a request's buffers in
// its arena block.  The inputs come first, in the same layout as the pinned
// host block they are staged in, so admission is one copy; the buffers that
// start at zero follow, so admission clears them with one fill.  The tape is
// not cleared: its slots are written before they are read, and the kernels
// read its ghost slots only to replace them by rebuilt boundary values.
struct Offsets {
  long rec_f, rec_a, rec_d, rho, W, prof, mms, ibar, cell, abs, in_end;
  long state, lam, rb, gen, seg, ac, extra, zero_end;
  long ckpt, tape, partial, I, mu, grad, total;
};

Offsets request_offsets(const Sizes& z) {
  Offsets o{};
  long p = 0;
  auto take = [&](long n) {
    const long q = p;
    p += (long)(align_up((size_t)n * sizeof(double)) / sizeof(double));
    return q;
  };
  o.rec_f = take(z.rec);
  o.rec_a = take(z.rec);
  o.rec_d = take(z.rec);
  o.rho = take(z.rho);
  o.W = take(z.W);
  o.prof = take(z.prof);
  o.mms = take(8 * 16);
  o.ibar = take(8);
  o.cell = take(z.cell);
  o.abs = take(z.abs);
  o.in_end = p;
  o.state = take(z.state);
  o.lam = take(z.state);
  o.rb = take(z.rb);
  o.gen = take(z.gen);
  o.seg = take(z.seg);
  o.ac = take(z.ac);
  o.extra = take(z.extra);
  o.zero_end = p;
  o.ckpt = take(z.ckpt);
  o.tape = take(z.tape);
  o.partial = take(z.partial);
  o.I = take(8);
  o.mu = take(z.mu);
  o.grad = take(z.grad);
  o.total = p;
  return o;
}

int default_ckpt(int n) { return std::max(1, (int)std::ceil(std::sqrt((double)n))); }

struct Request {
  int64_t ticket = 0;
  pnx_desc d{};
  Geometry g{};
  Sizes sz{};
  int lane = 0;
  Offsets off_{};
  Offsets off_tape{};            // layout if admitted taped
  size_t bytes_tape = 0;
  // pinned host block: the inputs packed in the device layout (off_.in_end
  // doubles), then the results (n_lines + rho_bar + gen doubles)
  double* pinned = nullptr;
  size_t pinned_bytes = 0;
  // outputs
  double* I_out = nullptr;
  double* rho_bar[8] = {};
  double* gen_out = nullptr;
  double* grad_out = nullptr;    // spec requests: the whole material gradient [cell][element]
  double* pinned_out = nullptr;  // results land here (inside `pinned`), then go to the caller's arrays
  // given as cell fractions: the material and absorption are set on the device at admission
  bool spec = false;
  // 2D cell-fraction request whose caller takes no per-class density
  // gradient: its reverse passes add the classes of a cell into one per-cell
  // rho_bar (DensArgs::rb_cell), which only the material pullback reads
  bool rb_cell = false;
  long abs_pts[3] = {}, abs_edges[3] = {}, abs_mac = 0, abs_am = 0, abs_emb = 0, abs_rp = 0, abs_el = 0;
  double inv_span[3] = {}, detector[3] = {}, beam_x0 = 0.0;
  AbsorptionGeometry ag{};
  // device
  size_t off = 0, bytes = 0;
  double *state = nullptr, *lam = nullptr, *ckpt = nullptr, *seg = nullptr, *rho = nullptr, *W = nullptr,
         *rec_f = nullptr, *rec_a = nullptr, *rec_d = nullptr, *prof = nullptr, *partial = nullptr, *rb = nullptr,
         *gen = nullptr, *mms = nullptr, *I_dev = nullptr, *ibar = nullptr, *cell_w = nullptr, *abs = nullptr,
         *mu = nullptr, *ac = nullptr, *grad = nullptr, *extra = nullptr;
  // schedule
  MarchPos pos = march_start();
  int ops_left = 0;            // operations not yet issued
  int64_t done_tick = -1;      // lane tick that copied the results back
  std::atomic<bool> cancel{false};  // set by pnx_server_cancel: issue nothing more
  // Ticks taken so far, counted by the worker outside the server mutex; poll
  // and wait report it as timing.ticks.  The other timing fields are written
  // and read under the mutex.
  std::atomic<int32_t> ticks{0};
  pnx_timing timing{};
  bool done = false;
  pnx_timing report() const {
    pnx_timing t = timing;
    t.ticks = ticks.load(std::memory_order_relaxed);
    return t;
  }
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
  int64_t tape_limit = 0;  // admit a request taped while the arena stays within this
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
  Lane lanes[kMaxLanes];
  int n_lanes = 3;
  PinnedPool pinned_pool;
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

// Point the request's buffers into its arena block and enqueue the input
// upload (one copy from the pinned block) and the zero fill on its lane's stream.
int admit(pnx_server* s, Request& r) {
  hipStream_t st = s->lanes[r.lane].stream;
  const Offsets& o = r.off_;
  double* base = reinterpret_cast<double*>(s->arena.base() + r.off);
  r.rec_f = base + o.rec_f;
  r.rec_a = base + o.rec_a;
  r.rec_d = base + o.rec_d;
  r.rho = base + o.rho;
  r.W = base + o.W;
  r.prof = base + o.prof;
  r.mms = base + o.mms;
  r.ibar = base + o.ibar;
  r.state = base + o.state;
  r.lam = base + o.lam;
  r.rb = base + o.rb;
  r.gen = base + o.gen;
  r.seg = base + (r.g.nck == 0 ? o.tape : o.seg);
  r.ckpt = base + o.ckpt;
  r.partial = base + o.partial;
  r.I_dev = base + o.I;
  if (hipMemcpyAsync(base, r.pinned, (size_t)o.in_end * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
    return -1;
  // state, lam, rho_bar, gen start at zero; the segment states are written
  // before they are read, but their ghost slots are kept finite
  if (launch_fill(r.state, o.zero_end - o.state, 0.0, st) != hipSuccess) return -1;
  if (r.spec) {
    // the densities of every class and the absorption factors and line
    // weights from the cell fractions (pnx_problem_set_cell_fractions and
    // pnx_problem_set_absorption on the request's own buffers)
    r.cell_w = base + o.cell;
    r.abs = base + o.abs;
    r.mu = base + o.mu;
    r.ac = base + o.ac;
    r.grad = base + o.grad;
    r.extra = base + o.extra;
    AbsorptionGeometry& ag = r.ag;
    for (int a = 0; a < 3; ++a) {
      ag.points[a] = r.abs + r.abs_pts[a];
      ag.edges[a] = r.abs + r.abs_edges[a];
      ag.n_boxes[a] = r.d.M[a];
      ag.inv_span[a] = r.inv_span[a];
      ag.detector[a] = r.detector[a];
    }
    ag.beam_x0 = r.beam_x0;
    const Geometry& g = r.g;
    if (launch_material_densities(g.L, g.class_n, g.rho_off, g.T.n, r.cell_w, r.abs + r.abs_rp, r.rho, st) !=
            hipSuccess ||
        launch_absorption(ag, g.L, r.cell_w, r.abs + r.abs_emb, r.abs + r.abs_rp, r.abs + r.abs_mac,
                          r.rho + g.rho_off[0], r.abs + r.abs_am, reinterpret_cast<const int*>(r.abs + r.abs_el),
                          r.mu, r.ac, r.W, st) != hipSuccess)
      return -1;
  }
  r.pos = march_start();
  r.ops_left = march_ops(r.d.n_steps, r.g.nck);
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
  density_read_offsets(r.g, r.spec, a.rho_off);
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
  density_read_offsets(r.g, r.spec, a.rho_off);
  for (int c = 0; c < 8; ++c) a.rb_off[c] = r.g.rho_off[c];
  a.rb_cell = r.rb_cell ? 1 : 0;
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
  m.N = r.d.N;
  return m;
}

double* buffer_of(const Request& r, int buf, int idx) {
  const long st = r.g.L.beam_stride;
  if (buf == kBufCkpt) return r.ckpt + (long)idx * st;
  if (buf == kBufSeg) return r.seg + (long)idx * st;
  return r.state;
}

// One phase launch of one (order, dimension) group: argument blocks and the
// work list (item.x = kind << kItemKindShift | argument index, item.y = tile).
struct PhaseWork {
  std::vector<HalfStepArgs> half;
  std::vector<RevArgs> rev;
  std::vector<int2> items;
};

// One operation of every planned request of a grid lane.
struct SubTick {
  std::map<GroupKey, PhaseWork[kPhases]> groups;
  std::vector<std::pair<Request*, int>> copies;     // (request, checkpoint index) copies before the phases
  std::vector<Request*> reduce;                     // forward finished: reduce intensities after the phases
};

// What a lane enqueues at once: kGridTicks sub-ticks (grid lanes) or one
// march slice per request (1D lane), with one argument upload.
struct TickPlan {
  std::vector<SubTick> sub;                         // grid lanes
  std::vector<MarchArgs> march;                     // 1D lane, all orders
  std::vector<Request*> copy_out;                   // results to copy back at the end
};

void add_half(PhaseWork& w, const HalfStepArgs& a, int kind) {
  const int idx = (int)w.half.size();
  w.half.push_back(a);
  for (int t = 0; t < a.L.n_tiles; ++t) w.items.push_back(make_int2(kind << kItemKindShift | idx, t));
}
void add_rev(PhaseWork& w, const RevArgs& a) {
  const int idx = (int)w.rev.size();
  w.rev.push_back(a);
  // one class per wavefront, 2D: one block per two 64-cell sub-tiles of a
  // cw tile; 3D: two blocks per two sub-tiles (classes 0-3 and 4-7)
  const int dims = a.adj.L.dims;
  const int n = dims == 2 ? a.adj.L.n_cw_items / 2 : (dims == 3 ? a.adj.L.n_cw_items : a.adj.L.n_tiles);
  for (int t = 0; t < n; ++t) w.items.push_back(make_int2(kItemRev << kItemKindShift | idx, t));
}

// Plan the next operation of request r on a grid lane.
void plan_grid(pnx_server* s, Request& r, SubTick& plan, TickPlan& whole) {
  const Geometry& g = r.g;
  PhaseWork* w = plan.groups[GroupKey{r.d.N, g.L.dims}];
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
      if (o.first) add_half(w[kPhaseFirstOdd], make_half(s, r, kForward, rec, h, src, src, dst, nullptr), kItemFwdOdd);
      add_half(w[kPhaseA], make_half(s, r, kForward, rec, g.dE, dst, src, dst, part), kItemFwdEven);
      if (o.fuse) {
        double* dst2 = buffer_of(r, o.dst2_buf, o.dst2_idx);
        HalfStepArgs a = make_half(s, r, kForward, rec, h, dst, dst, dst2 == dst ? nullptr : dst, nullptr);
        a.rec2 = rec + g.rec_len;
        a.dt2 = h;
        a.to_dst2 = dst2;
        add_half(w[kPhasePair], a, kItemPair);
      } else {
        add_half(w[kPhaseB], make_half(s, r, kForward, rec, h, dst, dst, dst, nullptr), kItemFwdOdd);
      }
      if (o.reduce) plan.reduce.push_back(&r);
      break;
    }
    case kOpReverse: {
      // A3(i+1) + pullback(i) + A1(i) fused, then AE(i)
      const double* rec = r.rec_a + (long)o.step * g.rec_len;
      RevArgs ra{};
      ra.adj = make_half(s, r, kAdjoint, rec, h, r.lam, r.lam, r.lam, nullptr);
      ra.dens = make_dens(s, r, o.step, src, 1.0);
      ra.rec3 = o.step < r.d.n_steps ? rec + g.rec_len : nullptr;
      ra.a1 = 1;
      add_rev(w[kPhaseA], ra);
      add_half(w[kPhaseB], make_half(s, r, kAdjoint, rec, g.dE, r.lam, r.lam, r.lam, nullptr), kItemAdjEven);
      break;
    }
    case kOpFinal: {
      // A3(1) + the pullback at u_0 with weight 0.5
      RevArgs ra{};
      ra.adj = make_half(s, r, kAdjoint, r.rec_a, h, r.lam, r.lam, r.lam, nullptr);
      ra.dens = make_dens(s, r, 0, src, 0.5);
      ra.rec3 = r.rec_a + g.rec_len;
      ra.a1 = 0;
      add_rev(w[kPhaseA], ra);
      whole.copy_out.push_back(&r);
      break;
    }
    default:
      break;
  }
  r.ticks.fetch_add(1, std::memory_order_relaxed);
}

// Plan the next slice of operations of request r on the 1D lane.
void plan_march(pnx_server* s, Request& r, TickPlan& plan) {
  MarchArgs m = make_march(s, r);
  m.pos = r.pos;
  m.n_ops = std::min(march_ops_per_launch(r.d.N), r.ops_left);
  for (int i = 0; i < m.n_ops; ++i) (void)march_next(r.pos, r.d.n_steps, r.g.K, r.g.nck);
  r.ops_left -= m.n_ops;
  plan.march.push_back(m);
  if (r.ops_left == 0) plan.copy_out.push_back(&r);
  r.ticks.fetch_add(1, std::memory_order_relaxed);
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
// `buf`, upload them, and enqueue the plan on the lane's stream.
int launch_plan(Lane& ln, TickPlan& plan, int buf) {
  hipStream_t st = ln.stream;
  struct Entry {
    int sub;    // sub-tick (grid lanes)
    GroupKey key;
    int phase;  // -1: the march block
    size_t n_off, half_off, rev_off, items_off;
    int n_items;
  };
  std::vector<Entry> entries;
  size_t off = 0;
  if (!plan.march.empty()) {
    Entry e{0, GroupKey{0, 1}, -1, 0, 0, 0, 0, (int)plan.march.size()};
    e.half_off = off;
    off = align_up(off + plan.march.size() * sizeof(MarchArgs));
    entries.push_back(e);
  }
  for (int t = 0; t < (int)plan.sub.size(); ++t)
    for (auto& kv : plan.sub[t].groups)
      for (int ph = 0; ph < kPhases; ++ph) {
        PhaseWork& w = kv.second[ph];
        if (w.items.empty()) continue;
        Entry e{t, kv.first, ph, 0, 0, 0, 0, (int)w.items.size()};
        e.n_off = off;
        off = align_up(off + sizeof(int));
        e.half_off = off;
        off = align_up(off + w.half.size() * sizeof(HalfStepArgs));
        e.rev_off = off;
        off = align_up(off + w.rev.size() * sizeof(RevArgs));
        e.items_off = off;
        off = align_up(off + w.items.size() * sizeof(int2));
        entries.push_back(e);
      }
  if (off == 0) off = kAlign;
  if (ensure_pinned(ln, buf, off) || ensure_dev_args(ln, off)) return -1;
  char* h = ln.pinned[buf];
  for (const Entry& e : entries) {
    if (e.phase < 0) {
      std::memcpy(h + e.half_off, plan.march.data(), plan.march.size() * sizeof(MarchArgs));
      continue;
    }
    PhaseWork& w = plan.sub[e.sub].groups[e.key][e.phase];
    std::memcpy(h + e.n_off, &e.n_items, sizeof(int));
    if (!w.half.empty()) std::memcpy(h + e.half_off, w.half.data(), w.half.size() * sizeof(HalfStepArgs));
    if (!w.rev.empty()) std::memcpy(h + e.rev_off, w.rev.data(), w.rev.size() * sizeof(RevArgs));
    std::memcpy(h + e.items_off, w.items.data(), w.items.size() * sizeof(int2));
  }
  if (hipMemcpyAsync(ln.dev_args, h, off, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
  int march_tiles = 0;
  for (const MarchArgs& m : plan.march) march_tiles = std::max(march_tiles, m.fwd.L.n_tiles);
  const int march_threads = march_tiles >= 2 && (int)plan.march.size() <= kMarchWideMax ? kMarchWide : kBlock;
  for (const Entry& e : entries)
    if (e.phase < 0 && launch_march(reinterpret_cast<const MarchArgs*>(ln.dev_args + e.half_off), e.n_items,
                                    march_threads, st) != hipSuccess)
      return -1;
  auto launch_phase = [&](int t, int phase) -> int {
    for (const Entry& e : entries) {
      if (e.phase != phase || e.sub != t) continue;
      const int grid = (e.n_items + 7) & ~7;  // one block per item, a multiple of 8 for the XCD mapping
      const int* nptr = reinterpret_cast<const int*>(ln.dev_args + e.n_off);
      const auto* half = reinterpret_cast<const HalfStepArgs*>(ln.dev_args + e.half_off);
      const auto* rev = reinterpret_cast<const RevArgs*>(ln.dev_args + e.rev_off);
      const int2* items = reinterpret_cast<const int2*>(ln.dev_args + e.items_off);
      hipError_t err;
      if (phase == kPhaseFirstOdd)  // kind kItemFwdOdd = 0: the items are (argument index, tile)
        err = launch_half_step_grouped(e.key.N, e.key.dims, 1, kForward, half, items, nptr, grid, st);
      else if (phase == kPhaseA)
        err = launch_phase_a_grouped(e.key.N, e.key.dims, half, rev, items, nptr, grid, st);
      else if (phase == kPhasePair)
        err = launch_pair_grouped(e.key.N, e.key.dims, half, items, nptr, grid, st);
      else
        err = launch_phase_b_grouped(e.key.N, e.key.dims, half, items, nptr, grid, st);
      if (err != hipSuccess) return -1;
    }
    return 0;
  };
  for (int t = 0; t < (int)plan.sub.size(); ++t) {
    SubTick& sp = plan.sub[t];
    for (auto& cp : sp.copies) {
      Request& r = *cp.first;
      const long stn = r.g.L.beam_stride;
      if (launch_copy(r.ckpt + (long)cp.second * stn, r.state, stn, st) != hipSuccess) return -1;
    }
    for (int ph = 0; ph < kPhases; ++ph)
      if (launch_phase(t, ph)) return -1;
    for (Request* r : sp.reduce) {
      const Geometry& g = r->g;
      if (launch_reduce_intensity(r->partial, g.partial_step, g.partial_beam, g.L.n_tiles * kWavesPerBlock, r->rec_f, g.rec_len,
                                  g.sigma_off, r->d.n_steps, r->d.n_lines, g.dE, g.L.dvol, 1, r->I_dev, st) !=
          hipSuccess)
        return -1;
    }
  }
  for (Request* r : plan.copy_out) {
    const long nI = r->d.n_lines, nrb = r->sz.rb, ngen = r->sz.gen;
    if (hipMemcpyAsync(r->pinned_out, r->I_dev, nI * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        (!r->rb_cell &&
         hipMemcpyAsync(r->pinned_out + nI, r->rb, nrb * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipMemcpyAsync(r->pinned_out + nI + nrb, r->gen, ngen * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess)
      return -1;
    if (r->spec && r->grad_out) {
      // the material pullbacks of the reference adjoint (pn_intensities.jl:406-410)
      // as pnx_problem_material_gradient: numbers of atoms, densities, absorption
      const Geometry& g = r->g;
      const long gs = (long)r->d.n_lines * g.L.V;
      const double* am = r->abs + r->abs_am;
      const int* el = reinterpret_cast<const int*>(r->abs + r->abs_el);
      const long n_cells = (long)g.L.Mx * g.L.My * g.L.Mz;
      if (launch_atoms_bar(r->ag, g.L, r->ac, r->gen, gs, 1, am, el, r->extra, st) != hipSuccess ||
          launch_fill(r->grad, n_cells * r->d.n_elements, 0.0, st) != hipSuccess ||
          launch_absorption_pullback(r->ag, g.L, r->abs + r->abs_mac, r->ac, r->rho + g.rho_off[0], am, el, r->gen,
                                     gs, 1, r->grad, st) != hipSuccess ||
          launch_material_pullback(g.L, g.class_n, g.rho_off, g.T.n, r->cell_w, r->abs + r->abs_rp, r->rb,
                                   g.rho_total, 1, r->extra, true, r->grad, st, r->rb_cell) != hipSuccess ||
          hipMemcpyAsync(r->pinned_out + nI + nrb + ngen, r->grad, r->sz.grad * sizeof(double),
                         hipMemcpyDeviceToHost, st) != hipSuccess)
        return -1;
    }
    r->done_tick = ln.tick;
  }
  return 0;
}

// A request cancelled after admission: once its lane's stream has drained
// (its upload and every launched pass read or write its memory) the memory is
// returned; its outputs are left untouched.
void finish_cancelled(pnx_server* s, Request& r) {
  (void)hipStreamSynchronize(s->lanes[r.lane].stream);
  s->pinned_pool.release(r.pinned, r.pinned_bytes);
  r.pinned = nullptr;
  r.pinned_out = nullptr;
  s->arena.release(r.off, r.bytes);
  r.timing.done_s = s->now();
  r.timing.status = PNX_ERR_CANCELLED;
  r.done = true;
}

void finish_request(pnx_server* s, Request& r) {
  if (r.done_tick < 0) {  // cancelled before its results were copied back
    finish_cancelled(s, r);
    return;
  }
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
  if (r.spec && r.grad_out) std::memcpy(r.grad_out, out + nI + nrb + r.sz.gen, r.sz.grad * sizeof(double));
  s->pinned_pool.release(r.pinned, r.pinned_bytes);
  r.pinned = nullptr;
  r.pinned_out = nullptr;
  s->arena.release(r.off, r.bytes);
  r.timing.done_s = s->now();
  r.timing.status = PNX_OK;
  r.done = true;
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
  if (li == 0) {
    for (auto& r : ln.active)
      if (r->ops_left > 0 && !r->cancel) {
        plan_march(s, *r, plan);
        any = true;
      }
  } else {
    for (int t = 0; t < kGridTicks; ++t) {
      SubTick sub;
      bool here = false;
      for (auto& r : ln.active)
        if (r->ops_left > 0 && !r->cancel) {
          plan_grid(s, *r, sub, plan);
          here = true;
        }
      if (!here) break;
      plan.sub.push_back(std::move(sub));
      any = true;
    }
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
      int lane_n[kMaxLanes];
      for (int li = 0; li < kMaxLanes; ++li) lane_n[li] = (int)s->lanes[li].active.size();
      while (!s->pending.empty() && s->n_active < s->max_active) {
        auto& r = s->pending.front();
        if (r->timing.arrival_s > now) break;
        // Taped (no recompute) while the tape fits under the tape limit.  Else
        // checkpointed, but only into a lane short of work: a lane that is
        // busy lets the request wait until memory frees for its tape.
        size_t off = SIZE_MAX;
        const Admission how =
            admission(s->arena.used(), (size_t)s->tape_limit, r->bytes_tape, lane_n[r->lane],
                      r->lane == 0 ? kBusyMarchLane : kBusyGridLane);
        if (how == kAdmitTaped) off = s->arena.alloc(r->bytes_tape);
        if (off != SIZE_MAX) {
          r->off_ = r->off_tape;
          r->bytes = r->bytes_tape;
          r->g.K = r->d.n_steps;
          r->g.nck = 0;
          r->timing.taped = 1;
        } else {
          // not taped (or its taped layout found no free range): checkpointed
          // unless its lane is busy
          // (kAdmitWait implies a busy lane)
          if (s->tape_limit > 0 && lane_n[r->lane] >= (r->lane == 0 ? kBusyMarchLane : kBusyGridLane)) break;
          off = s->arena.alloc(r->bytes);
        }
        if (off == SIZE_MAX) break;
        lane_n[r->lane] += 1;
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
    for (int li = 0; li < s->n_lanes; ++li) {
      Lane& ln = s->lanes[li];
      if (ln.active.empty()) continue;
      if (poll_lane(ln)) {
        fail_lane(s, "device error in a served tick");
        break;
      }
      // requests whose copy-out tick has completed
      auto it = std::partition(ln.active.begin(), ln.active.end(), [&](const std::shared_ptr<Request>& r) {
        if (r->done_tick >= 0) return !(r->done_tick <= ln.done_through);
        return !r->cancel;  // cancelled before its results were copied back
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
      s->ticks = 0;
      for (int li = 0; li < s->n_lanes; ++li) s->ticks += s->lanes[li].tick;
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
  return (int64_t)(request_offsets(request_sizes(g, d)).total * sizeof(double));
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
  delete s;
}

int pnx_server_create(pnx_context* ctx, int64_t hbm_budget_bytes, int32_t max_active, pnx_server** out) {
  if (!ctx || !out || hbm_budget_bytes <= 0 || max_active < 1) return set_error(PNX_ERR_ARG, "bad argument");
  if (hipSetDevice(ctx->device) != hipSuccess) return set_error(PNX_ERR_DEVICE, "hipSetDevice failed");
  auto* s = new pnx_server;
  s->ctx = ctx;
  s->budget = hbm_budget_bytes;
  s->tape_limit = hbm_budget_bytes / 20 * 19;
  s->max_active = max_active;
  if (s->arena.init((size_t)hbm_budget_bytes)) {
    release_server(s);
    return set_error(PNX_ERR_MEMORY, "could not allocate the server arena");
  }
  if (upload_coefficients(s)) {
    release_server(s);
    return set_error(PNX_ERR_DEVICE, "coefficient upload failed");
  }
  // The 1D lane runs on the context's stream, the 2D/3D lanes on streams of
  // their own (grid_lanes_for_queues).
  s->n_lanes = 1 + grid_lanes_for_queues();
  for (int li = 0; li < s->n_lanes; ++li) {
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

int pnx_server_set_tape_limit(pnx_server* s, int64_t bytes) {
  if (!s || bytes < 0) return set_error(PNX_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  s->tape_limit = std::min(bytes, s->budget);
  return PNX_OK;
}

int pnx_march_sequence(int32_t n, int32_t K, int32_t* out, int32_t capacity) {
  if (n < 1 || K < 0) return set_error(PNX_ERR_ARG, "n must be positive and K non-negative");
  const int Kc = K == 0 ? n : std::min(K, n), nck = K == 0 ? 0 : (n + Kc - 1) / Kc;
  const int ops = march_ops(n, nck);
  if (!out) return ops;
  if (capacity < 9 * ops) return set_error(PNX_ERR_ARG, "output too small: 9 integers per operation needed");
  MarchPos pos = march_start();
  for (int t = 0; t < ops; ++t) {
    const MarchOp o = march_next(pos, n, Kc, nck);
    const int32_t v[9] = {o.kind, o.step, o.ckpt_copy, o.src_buf, o.src_idx, o.dst_buf, o.dst_idx, o.first, o.fuse};
    for (int i = 0; i < 9; ++i) out[9 * t + i] = v[i];
  }
  return ops;
}

}  // extern "C"

namespace {

// Lane, memory layouts (checkpointed and taped) and sizes of a request whose
// descriptor, geometry and buffer sizes are set.
int prepare_request(pnx_server* s, Request& r) {
  r.lane = lane_of(r.d.N, r.g.L.dims, s->n_lanes - 1);
  r.off_ = request_offsets(r.sz);
  r.off_tape = request_offsets(taped_sizes(r.sz, r.d.n_steps));
  r.bytes = (size_t)r.off_.total * sizeof(double);
  r.bytes_tape = (size_t)r.off_tape.total * sizeof(double);
  if ((int64_t)r.bytes > s->budget) return set_error(PNX_ERR_MEMORY, "request exceeds the server's HBM budget");
  return PNX_OK;
}

// Queues a prepared request in arrival order and returns its ticket.
int enqueue(pnx_server* s, const std::shared_ptr<Request>& r, int64_t* ticket) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    auto it = s->pending.end();
    while (it != s->pending.begin() && (*std::prev(it))->timing.arrival_s > r->timing.arrival_s) --it;
    s->pending.insert(it, r);
    s->all[r->ticket] = r;
    *ticket = r->ticket;
  }
  s->cv.notify_all();
  return PNX_OK;
}

}  // namespace

extern "C" {

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
  if ((rc = prepare_request(s, *r))) return rc;
  const Offsets& o = r->off_;
  const Geometry& g = r->g;
  for (int c = 0; c < 8; ++c) {
    if (!g.active(c)) continue;
    if (!q->rho[c]) return set_error(PNX_ERR_ARG, "densities missing for an updated class");
    if (odd_in(c, 0) && (!q->ppos[c] || !q->mms[c]))
      return set_error(PNX_ERR_ARG, "beam data missing for an x-odd class");
  }
  const size_t nout = (size_t)(r->d.n_lines + r->sz.rb + r->sz.gen);
  const size_t nh = (size_t)o.in_end + nout;
  r->pinned_bytes = nh * sizeof(double);
  r->pinned = s->pinned_pool.alloc(r->pinned_bytes);
  if (!r->pinned) return set_error(PNX_ERR_MEMORY, "pinned staging allocation failed");
  double* h = r->pinned;
  std::memset(h, 0, (size_t)o.in_end * sizeof(double));
  const long rec = r->sz.rec;
  std::memcpy(h + o.rec_f, q->fwd, rec * sizeof(double));
  std::memcpy(h + o.rec_a, q->adj, rec * sizeof(double));
  std::memcpy(h + o.rec_d, q->dens, rec * sizeof(double));
  for (int c = 0; c < 8; ++c)
    if (g.active(c)) pack_class(g, c, q->rho[c], r->d.n_elements, h + o.rho + g.rho_off[c]);
  pack_lines(g, q->W, h + o.W);
  pack_profiles(g, q->ppos, h + o.prof);
  for (int c = 0; c < 8; ++c)
    if (odd_in(c, 0) && g.active(c))
      for (int m = 0; m < g.T.n[c]; ++m) h[o.mms + c * 16 + m] = q->mms[c][m];
  for (int l = 0; l < r->d.n_lines; ++l) h[o.ibar + l] = q->I_bar[l];
  r->pinned_out = h + o.in_end;
  r->I_out = q->I_out;
  for (int c = 0; c < 8; ++c) r->rho_bar[c] = q->rho_bar[c];
  r->gen_out = q->gen_out;
  r->timing.arrival_s = q->arrival_s;
  return enqueue(s, r, ticket);
}

int pnx_server_submit_spec(pnx_server* s, const pnx_request_spec* q, int64_t* ticket) {
  if (!s || !q || !ticket) return set_error(PNX_ERR_ARG, "null argument");
  if (q->desc.n_beams != 1) return set_error(PNX_ERR_ARG, "served requests carry exactly one beam");
  if (!q->physics || !q->cell_w || !q->rho_pure || !q->embedding || !q->mac || !q->atom_mass || !q->line_element ||
      !q->I_bar || !q->I_out)
    return set_error(PNX_ERR_ARG, "tables, fractions, absorption inputs, seed and output are required");
  auto r = std::make_shared<Request>();
  r->d = q->desc;
  if (r->d.ckpt_every <= 0) r->d.ckpt_every = default_ckpt(r->d.n_steps);
  std::string err;
  int rc = make_geometry(r->d, r->g, err);
  if (rc) return set_error(rc, err);
  const pnx_desc& d = r->d;
  const int ne = d.n_elements, nl = d.n_lines;
  for (int e = 0; e < ne; ++e)
    if (!(q->rho_pure[e] > 0.0)) return set_error(PNX_ERR_ARG, "pure-element densities must be positive");
  for (int l = 0; l < nl; ++l)
    if (q->line_element[l] < 0 || q->line_element[l] >= ne)
      return set_error(PNX_ERR_ARG, "line element index out of range");
  r->spec = true;
  r->sz = request_sizes(r->g, d, true);
  if ((rc = prepare_request(s, *r))) return rc;
  const Geometry& g = r->g;
  const Offsets& o = r->off_;
  const size_t nout = (size_t)(nl + r->sz.rb + r->sz.gen + (q->grad_out ? r->sz.grad : 0));
  const size_t nh = (size_t)o.in_end + nout;
  r->pinned_bytes = nh * sizeof(double);
  r->pinned = s->pinned_pool.alloc(r->pinned_bytes);
  if (!r->pinned) return set_error(PNX_ERR_MEMORY, "pinned staging allocation failed");
  double* h = r->pinned;
  std::memset(h, 0, (size_t)o.in_end * sizeof(double));  // densities and line weights are set on the device
  auto give_back = [&](int code) {
    s->pinned_pool.release(r->pinned, r->pinned_bytes);
    return code;
  };
  // records, beam moments and profiles: the library's C++ set-up
  if ((rc = pnx_problem_records(&d, q->physics, h + o.rec_f, h + o.rec_a, h + o.rec_d))) return give_back(rc);
  {
    double mm[8][16] = {};
    double* m8[8];
    std::vector<double> pp[8];
    double* p8[8];
    for (int c = 0; c < 8; ++c) {
      m8[c] = odd_in(c, 0) ? mm[c] : nullptr;
      pp[c].assign(odd_in(c, 0) ? (size_t)g.class_n[c][1] * g.class_n[c][2] : 0, 0.0);
      p8[c] = odd_in(c, 0) ? pp[c].data() : nullptr;
    }
    if ((rc = pnx_beam_moments(d.N, q->beam_direction, q->beam_kappa, m8)) ||
        (rc = pnx_beam_profiles(&d, q->beam_position, q->beam_sigma, q->literal_source, p8)))
      return give_back(rc);
    pack_profiles(g, p8, h + o.prof);
    for (int c = 0; c < 8; ++c)
      if (odd_in(c, 0) && g.active(c))
        for (int m = 0; m < g.T.n[c]; ++m) h[o.mms + c * 16 + m] = mm[c][m];
  }
  for (int l = 0; l < nl; ++l) h[o.ibar + l] = q->I_bar[l];
  std::memcpy(h + o.cell, q->cell_w, (size_t)r->sz.cell * sizeof(double));
  // absorption inputs: even points and box edges per axis, then the tables
  {
    double* a0 = h + o.abs;
    long k = 0;
    for (int a = 0; a < 3; ++a) {
      const int M = d.M[a];
      r->abs_pts[a] = k;
      for (int i = 0; i < M; ++i) a0[k++] = axis_point(d, a, 0, i);
      r->abs_edges[a] = k;
      if (M == 1) {
        a0[k++] = d.lo[a] - 0.5;
        a0[k++] = d.lo[a] + 0.5;
      } else {
        for (int i = 0; i <= M; ++i) a0[k++] = axis_point(d, a, 1, i);
      }
      r->inv_span[a] = (double)M / (a0[r->abs_edges[a] + M] - a0[r->abs_edges[a]]);
    }
    r->abs_mac = k;
    for (int i = 0; i < nl * ne; ++i) a0[k++] = q->mac[i];
    r->abs_am = k;
    for (int e = 0; e < ne; ++e) a0[k++] = q->atom_mass[e];
    r->abs_emb = k;
    for (int e = 0; e < ne; ++e) a0[k++] = q->embedding[e];
    r->abs_rp = k;
    for (int e = 0; e < ne; ++e) a0[k++] = q->rho_pure[e];
    r->abs_el = k;
    std::memcpy(a0 + k, q->line_element, nl * sizeof(int32_t));
  }
  for (int a = 0; a < 3; ++a) r->detector[a] = q->detector[a];
  r->beam_x0 = q->beam_position[0];
  r->pinned_out = h + o.in_end;
  r->I_out = q->I_out;
  bool any_rb = false;
  for (int c = 0; c < 8; ++c) {
    r->rho_bar[c] = q->rho_bar[c];
    any_rb |= q->rho_bar[c] != nullptr;
  }
  r->rb_cell = !any_rb && g.L.dims == 2 && ne <= kCellE;
  r->gen_out = q->gen_out;
  r->grad_out = q->grad_out;
  r->timing.arrival_s = q->arrival_s;
  return enqueue(s, r, ticket);
}

int pnx_server_poll(pnx_server* s, int64_t ticket, int32_t* done, pnx_timing* t) {
  if (!s || !done) return set_error(PNX_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return set_error(PNX_ERR_ARG, "unknown ticket");
  *done = it->second->done ? 1 : 0;
  if (t) *t = it->second->report();
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
  if (t) *t = r->report();
  if (s->error) return set_error(s->error, s->error_msg);
  s->all.erase(ticket);  // the ticket is released: its request is complete
  return PNX_OK;
}

int pnx_server_cancel(pnx_server* s, int64_t ticket) {
  if (!s) return set_error(PNX_ERR_ARG, "null server");
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return set_error(PNX_ERR_ARG, "unknown ticket");
    r = it->second;
    if (r->done) return PNX_OK;
    auto pit = std::find(s->pending.begin(), s->pending.end(), r);
    if (pit != s->pending.end()) {
      // not admitted: it holds only its pinned staging block
      s->pending.erase(pit);
      s->pinned_pool.release(r->pinned, r->pinned_bytes);
      r->pinned = nullptr;
      r->pinned_out = nullptr;
      r->timing.done_s = s->now();
      r->timing.status = PNX_ERR_CANCELLED;
      r->done = true;
    } else {
      // admitted: the scheduler issues no more of its work and completes it
      // once its launched ticks have passed
      r->cancel = true;
    }
  }
  s->done_cv.notify_all();
  s->cv.notify_all();
  return PNX_OK;
}

int pnx_server_drain(pnx_server* s) {
  if (!s) return set_error(PNX_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->done_cv.wait(lk, [&] { return (s->pending.empty() && s->n_active == 0) || s->error != 0; });
  if (s->error) return set_error(s->error, s->error_msg);
  return PNX_OK;
}

int pnx_server_stats(pnx_server* s, int64_t* ticks, int64_t* completed, int64_t* peak_active,
                     int64_t* peak_bytes) {
  if (!s) return set_error(PNX_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (ticks) *ticks = s->ticks;
  if (completed) *completed = s->completed;
  if (peak_active) *peak_active = s->peak_active;
  if (peak_bytes) *peak_bytes = (int64_t)s->peak_bytes;
  return PNX_OK;
}

}  // extern "C"
