// C ABI of the PN-CSD solver: contexts, problems, and the energy-march drivers.
//
// The drivers follow the reference loops: calc_intensities
// (pn_solver/pn_intensities.jl:325-351) and its Zygote adjoint
// (pn_solver/pn_intensities.jl:353-413), with the full state tape of the
// reference replaced by checkpoints every `ckpt_every` energy steps and a
// forward recomputation of each segment during the reverse sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pnx_host.hpp"

// Device scratch of the allocation-free forward (pnx_forward_noalloc): the
// intensity partial sums of every beam and energy step, then the reduced
// intensities.  Each device the library runs on holds its own copy until the
// process ends.  Host code never uses this name as an address: it uses the
// per-device pointer from hipGetSymbolAddress (pnx_context.scratch).
constexpr size_t kScratchDoubles = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace pnx {
double advection_coefficient(int l, int k, int lf, int kf, int d);

namespace {
thread_local std::string g_err;
}

int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace pnx

using namespace pnx;

namespace {

int fail(int code, const std::string& msg) { return set_error(code, msg); }

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of the current device's `scratch`, resolved once per
// device (the device must be current).
int resolve_scratch(int device, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return fail(PNX_ERR_ARG, "device index out of range");
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(PNX_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch[device];
  return PNX_OK;
}

#define HIP_TRY(expr)                                                                               \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return fail(PNX_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  double* p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    if (p && n >= count) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) return 0;
    if (hipMalloc(&p, count * sizeof(double)) != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return fail(PNX_ERR_MEMORY, "hipMalloc of " + std::to_string(count * sizeof(double)) + " bytes failed");
    }
    n = count;
    return 0;
  }
};

}  // namespace

struct pnx_problem {
  pnx_context* ctx = nullptr;
  pnx_desc d{};
  Geometry g{};
  Layout& L = g.L;
  Tables& T = g.T;
  long rb_stride = 0;   // per beam rho_bar doubles
  bool have_rec = false, have_adj = false, have_mat = false, have_beams = false;
  DevBuf state, rho, W, coef_f, coef_a, rec_f, rec_a, rec_d, prof, mms, partial, I_dev;
  DevBuf lam, ckpt, seg, rho_bar_b, gen_b, ibar, sum_buf, w1, w2;
  DevBuf cell_w, rho_pure, extra, grad;  // device material layer
  DevBuf abs_small, mu, ac;             // device absorption: points/edges/mac/masses/embedding, table, factors
  int* line_el = nullptr;
  AbsorptionGeometry abs{};
  long abs_off_mac = 0, abs_off_am = 0, abs_off_emb = 0;
  bool have_cells = false, have_abs = false, have_rho_bar = false;
  hipGraphExec_t fwd_graph = nullptr;
  // The problem's own stream: its launches, copies and graph capture never
  // share a stream with a server attached to the same context.
  hipStream_t stream = nullptr;
  ~pnx_problem() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static void destroy_graph(pnx_problem* p) {
  if (p->fwd_graph) {
    (void)hipGraphExecDestroy(p->fwd_graph);
    p->fwd_graph = nullptr;
  }
}

extern "C" {

const char* pnx_last_error(void) { return g_err.c_str(); }
int pnx_version(void) { return 1; }

int pnx_context_create(int device, pnx_context** out) {
  if (!out) return fail(PNX_ERR_ARG, "null output pointer");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(PNX_ERR_ARG, "device index out of range");
  HIP_TRY(hipSetDevice(device));
  double* scr = nullptr;
  if (int rc = resolve_scratch(device, &scr)) return rc;
  auto* c = new pnx_context;
  c->device = device;
  c->scratch = scr;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(PNX_ERR_DEVICE, "hipStreamCreate failed");
  }
  *out = c;
  return PNX_OK;
}

int pnx_context_destroy(pnx_context* ctx) {
  if (!ctx) return PNX_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PNX_OK;
}

int pnx_context_synchronize(pnx_context* ctx) {
  if (!ctx) return fail(PNX_ERR_ARG, "null context");
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PNX_OK;
}

int pnx_coefficients(int N, int adjoint, double* out, int capacity) {
  if (N < 0 || N > kMaxOrder) return fail(PNX_ERR_ARG, "order out of range");
  std::vector<double> c = build_coefficients(N, adjoint != 0);
  if (!out) return (int)c.size();
  if (capacity < (int)c.size()) return fail(PNX_ERR_ARG, "capacity too small");
  std::copy(c.begin(), c.end(), out);
  return (int)c.size();
}

int pnx_advection_block(int N, int c, int d, double* out) {
  if (N < 0 || N > kMaxOrder || c < 0 || c > 7 || d < 0 || d > 2 || !out) return fail(PNX_ERR_ARG, "bad argument");
  const Tables t = make_tables(N);
  const int f = switch_axis(c, d);
  for (int i = 0; i < t.n[c]; ++i)
    for (int j = 0; j < t.n[f]; ++j)
      out[i * t.n[f] + j] = advection_coefficient(t.l[c][i], t.k[c][i], t.l[f][j], t.k[f][j], d);
  return PNX_OK;
}

int pnx_boundary_block(int N, int c, int d, int adjoint, double* out) {
  if (N < 0 || N > kMaxOrder || c < 0 || c > 7 || d < 0 || d > 2 || !out || !odd_in(c, d))
    return fail(PNX_ERR_ARG, "bad argument");
  const Tables t = make_tables(N);
  std::vector<double> co = build_coefficients(N, adjoint != 0);
  const int n = t.n[c] * t.n[switch_axis(c, d)];
  std::copy(co.begin() + t.b_off[c][d], co.begin() + t.b_off[c][d] + n, out);
  return PNX_OK;
}

int pnx_problem_create(pnx_context* ctx, const pnx_desc* desc, pnx_problem** out) {
  if (!ctx || !desc || !out) return fail(PNX_ERR_ARG, "null argument");
  const pnx_desc& d = *desc;
  HIP_TRY(hipSetDevice(ctx->device));
  auto* p = new pnx_problem;
  p->ctx = ctx;
  p->d = d;
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    p->stream = nullptr;
    delete p;
    return fail(PNX_ERR_DEVICE, "hipStreamCreate failed");
  }
  std::string err;
  if (int e = make_geometry(d, p->g, err)) {
    delete p;
    return fail(e, err);
  }
  const Layout& L = p->L;
  const long ro = p->g.rho_total;
  p->rb_stride = ro;
  int rc = 0;
  if ((rc = p->state.alloc((size_t)L.beam_stride * d.n_beams)) || (rc = p->rho.alloc(ro)) ||
      (rc = p->W.alloc((size_t)d.n_lines * L.V)) || (rc = p->prof.alloc((size_t)p->g.prof_stride * d.n_beams)) ||
      (rc = p->mms.alloc(8 * 16)) || (rc = p->partial.alloc((size_t)p->g.partial_beam * d.n_beams)) ||
      (rc = p->I_dev.alloc((size_t)d.n_beams * d.n_lines)) ||
      (rc = p->rec_f.alloc((size_t)(d.n_steps + 1) * p->g.rec_len))) {
    delete p;
    return rc;
  }
  std::vector<double> cf = build_coefficients(d.N, false), ca = build_coefficients(d.N, true);
  if ((rc = p->coef_f.alloc(cf.size())) || (rc = p->coef_a.alloc(ca.size()))) {
    delete p;
    return rc;
  }
  hipStream_t s = ctx->stream;
  if (hipMemcpyAsync(p->coef_f.p, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(p->coef_a.p, ca.data(), ca.size() * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
      launch_fill(p->state.p, (long)p->state.n, 0.0, s) != hipSuccess ||
      launch_fill(p->rho.p, (long)p->rho.n, 0.0, s) != hipSuccess ||
      launch_fill(p->W.p, (long)p->W.n, 0.0, s) != hipSuccess ||
      launch_fill(p->prof.p, (long)p->prof.n, 0.0, s) != hipSuccess ||
      launch_fill(p->mms.p, (long)p->mms.n, 0.0, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    delete p;
    return fail(PNX_ERR_DEVICE, "problem initialisation failed");
  }
  *out = p;
  return PNX_OK;
}

int pnx_problem_destroy(pnx_problem* p) {
  if (!p) return PNX_OK;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->stream);
  destroy_graph(p);
  if (p->line_el) (void)hipFree(p->line_el);
  delete p;
  return PNX_OK;
}

int64_t pnx_problem_device_bytes(const pnx_problem* p) {
  if (!p) return 0;
  size_t n = 0;
  for (const DevBuf* b : {&p->state, &p->rho, &p->W, &p->coef_f, &p->coef_a, &p->rec_f, &p->rec_a, &p->rec_d,
                          &p->prof, &p->mms, &p->partial, &p->I_dev, &p->lam, &p->ckpt, &p->seg, &p->rho_bar_b,
                          &p->gen_b, &p->ibar, &p->sum_buf, &p->w1, &p->w2, &p->cell_w, &p->rho_pure, &p->extra,
                          &p->grad, &p->abs_small, &p->mu, &p->ac})
    n += b->n;
  return (int64_t)(n * sizeof(double) + (p->line_el ? 8 * sizeof(int) : 0));
}

int pnx_problem_set_records(pnx_problem* p, const double* fwd, const double* adj, const double* dens) {
  if (!p || !fwd) return fail(PNX_ERR_ARG, "null argument");
  if ((adj == nullptr) != (dens == nullptr)) return fail(PNX_ERR_ARG, "adj and dens must be given together");
  HIP_TRY(hipSetDevice(p->ctx->device));
  const size_t n = (size_t)(p->d.n_steps + 1) * p->g.rec_len;
  hipStream_t s = p->stream;
  HIP_TRY(hipMemcpyAsync(p->rec_f.p, fwd, n * sizeof(double), hipMemcpyHostToDevice, s));
  if (adj) {
    int rc;
    if ((rc = p->rec_a.alloc(n)) || (rc = p->rec_d.alloc(n))) return rc;
    HIP_TRY(hipMemcpyAsync(p->rec_a.p, adj, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(p->rec_d.p, dens, n * sizeof(double), hipMemcpyHostToDevice, s));
    p->have_adj = true;
  }
  HIP_TRY(hipStreamSynchronize(s));
  p->have_rec = true;
  return PNX_OK;
}

int pnx_problem_set_material(pnx_problem* p, const double* const rho[8], const double* W) {
  if (!p || !rho || !W) return fail(PNX_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(p->ctx->device));
  const Geometry& g = p->g;
  std::vector<double> h((size_t)g.rho_total, 0.0);
  for (int c = 0; c < 8; ++c) {
    if (!g.active(c)) continue;
    if (!rho[c]) return fail(PNX_ERR_ARG, "densities missing for an updated class");
    pack_class(g, c, rho[c], p->d.n_elements, h.data() + g.rho_off[c]);
  }
  std::vector<double> w((size_t)p->d.n_lines * g.L.V, 0.0);
  pack_lines(g, W, w.data());
  hipStream_t s = p->stream;
  HIP_TRY(hipMemcpyAsync(p->rho.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->W.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  // densities per class again: the kernels read each class's own array
  if (p->have_cells) destroy_graph(p);
  p->have_mat = true;
  p->have_cells = false;
  return PNX_OK;
}

int pnx_problem_set_cell_fractions(pnx_problem* p, const double* w, const double* rho_pure, const double* W) {
  if (!p || !w || !rho_pure || !W) return fail(PNX_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(p->ctx->device));
  const Geometry& g = p->g;
  const int ne = p->d.n_elements;
  for (int e = 0; e < ne; ++e)
    if (!(rho_pure[e] > 0.0)) return fail(PNX_ERR_ARG, "pure-element densities must be positive");
  const size_t n_cells = (size_t)g.L.Mx * g.L.My * g.L.Mz;
  int rc;
  if ((rc = p->cell_w.alloc(n_cells * ne)) || (rc = p->rho_pure.alloc(ne))) return rc;
  hipStream_t s = p->stream;
  HIP_TRY(hipMemcpyAsync(p->cell_w.p, w, n_cells * ne * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->rho_pure.p, rho_pure, ne * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(launch_fill(p->rho.p, g.rho_total, 0.0, s));
  HIP_TRY(launch_material_densities(g.L, g.class_n, g.rho_off, p->T.n, p->cell_w.p, p->rho_pure.p, p->rho.p, s));
  std::vector<double> wl((size_t)p->d.n_lines * g.L.V, 0.0);
  pack_lines(g, W, wl.data());
  HIP_TRY(hipMemcpyAsync(p->W.p, wl.data(), wl.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  destroy_graph(p);
  p->have_mat = true;
  p->have_cells = true;
  return PNX_OK;
}

int pnx_problem_cell_fraction_gradient(pnx_problem* p, const double* extra_eee, double* grad) {
  if (!p || !grad) return fail(PNX_ERR_ARG, "null argument");
  if (!p->have_cells) return fail(PNX_ERR_STATE, "the material was not set from cell fractions");
  if (!p->have_rho_bar) return fail(PNX_ERR_STATE, "no density gradient yet: run pnx_forward_adjoint first");
  HIP_TRY(hipSetDevice(p->ctx->device));
  const Geometry& g = p->g;
  const int ne = p->d.n_elements;
  const size_t n_cells = (size_t)g.L.Mx * g.L.My * g.L.Mz;
  hipStream_t s = p->stream;
  int rc;
  if ((rc = p->grad.alloc(n_cells * ne))) return rc;
  std::vector<double> h;
  if (extra_eee) {
    if ((rc = p->extra.alloc((size_t)ne * g.L.V))) return rc;
    h.assign((size_t)ne * g.L.V, 0.0);
    pack_class(g, 0, extra_eee, ne, h.data());
    HIP_TRY(hipMemcpyAsync(p->extra.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  }
  HIP_TRY(launch_material_pullback(g.L, g.class_n, g.rho_off, p->T.n, p->cell_w.p, p->rho_pure.p, p->rho_bar_b.p,
                                   p->rb_stride, p->d.n_beams, extra_eee ? p->extra.p : nullptr, false, p->grad.p, s));
  HIP_TRY(hipMemcpyAsync(grad, p->grad.p, n_cells * ne * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return PNX_OK;
}

int pnx_problem_set_absorption(pnx_problem* p, const double* mac, const double* atom_mass,
                               const int32_t* line_element, const double* detector, double beam_x0,
                               const double* const axis_points[3], const double* const cell_edges[3],
                               const double* embedding, double* ac_out) {
  if (!p || !mac || !atom_mass || !line_element || !detector || !axis_points || !cell_edges || !embedding)
    return fail(PNX_ERR_ARG, "null argument");
  if (!p->have_cells) return fail(PNX_ERR_STATE, "set the cell fractions first (pnx_problem_set_cell_fractions)");
  const Geometry& g = p->g;
  const int ne = p->d.n_elements, nl = p->d.n_lines;
  const int M[3] = {g.L.Mx, g.L.My, g.L.Mz};
  for (int r = 0; r < nl; ++r)
    if (line_element[r] < 0 || line_element[r] >= ne) return fail(PNX_ERR_ARG, "line element index out of range");
  for (int a = 0; a < 3; ++a)
    if (!axis_points[a] || !cell_edges[a]) return fail(PNX_ERR_ARG, "null axis points or edges");
  HIP_TRY(hipSetDevice(p->ctx->device));
  // small inputs in one buffer: points and edges per axis, mac, atom masses, embedding
  std::vector<double> h;
  long off_pts[3], off_edges[3];
  for (int a = 0; a < 3; ++a) {
    off_pts[a] = (long)h.size();
    h.insert(h.end(), axis_points[a], axis_points[a] + M[a]);
    off_edges[a] = (long)h.size();
    h.insert(h.end(), cell_edges[a], cell_edges[a] + M[a] + 1);  // one box per point (a singleton axis: one)
  }
  p->abs_off_mac = (long)h.size();
  h.insert(h.end(), mac, mac + (size_t)nl * ne);
  p->abs_off_am = (long)h.size();
  h.insert(h.end(), atom_mass, atom_mass + ne);
  p->abs_off_emb = (long)h.size();
  h.insert(h.end(), embedding, embedding + ne);
  int rc;
  const long n_boxes = (long)M[0] * M[1] * M[2];
  if ((rc = p->abs_small.alloc(h.size())) || (rc = p->mu.alloc((size_t)(n_boxes + 1) * nl)) ||
      (rc = p->ac.alloc((size_t)nl * g.L.V)))
    return rc;
  if (!p->line_el && hipMalloc(&p->line_el, 8 * sizeof(int)) != hipSuccess) {
    p->line_el = nullptr;
    return fail(PNX_ERR_MEMORY, "hipMalloc failed");
  }
  hipStream_t s = p->stream;
  HIP_TRY(hipMemcpyAsync(p->abs_small.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->line_el, line_element, nl * sizeof(int), hipMemcpyHostToDevice, s));
  AbsorptionGeometry& ag = p->abs;
  for (int a = 0; a < 3; ++a) {
    ag.points[a] = p->abs_small.p + off_pts[a];
    ag.edges[a] = p->abs_small.p + off_edges[a];
    ag.n_boxes[a] = M[a];
    const double* e = cell_edges[a];
    ag.inv_span[a] = (double)M[a] / (e[M[a]] - e[0]);
    ag.detector[a] = detector[a];
  }
  ag.beam_x0 = beam_x0;
  HIP_TRY(launch_fill(p->ac.p, (long)nl * g.L.V, 0.0, s));
  HIP_TRY(launch_fill(p->W.p, (long)nl * g.L.V, 0.0, s));
  HIP_TRY(launch_absorption(ag, g.L, p->cell_w.p, p->abs_small.p + p->abs_off_emb, p->rho_pure.p,
                            p->abs_small.p + p->abs_off_mac, p->rho.p + g.rho_off[0], p->abs_small.p + p->abs_off_am,
                            p->line_el, p->mu.p, p->ac.p, p->W.p, s));
  if (ac_out) {
    std::vector<double> a((size_t)nl * g.L.V);
    HIP_TRY(hipMemcpyAsync(a.data(), p->ac.p, a.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    unpack_lines(g, a.data(), ac_out);
  }
  HIP_TRY(hipStreamSynchronize(s));
  destroy_graph(p);
  p->have_abs = true;
  return PNX_OK;
}

int pnx_problem_material_gradient(pnx_problem* p, double* grad) {
  if (!p || !grad) return fail(PNX_ERR_ARG, "null argument");
  if (!p->have_cells || !p->have_abs)
    return fail(PNX_ERR_STATE, "set the cell fractions and the absorption on the device first");
  if (!p->have_rho_bar) return fail(PNX_ERR_STATE, "no density gradient yet: run pnx_forward_adjoint first");
  HIP_TRY(hipSetDevice(p->ctx->device));
  const Geometry& g = p->g;
  const int ne = p->d.n_elements;
  const size_t n_cells = (size_t)g.L.Mx * g.L.My * g.L.Mz;
  hipStream_t s = p->stream;
  int rc;
  if ((rc = p->grad.alloc(n_cells * ne)) || (rc = p->extra.alloc((size_t)ne * g.L.V))) return rc;
  const long gs = (long)p->d.n_lines * g.L.V;
  const double* am = p->abs_small.p + p->abs_off_am;
  HIP_TRY(launch_fill(p->extra.p, (long)ne * g.L.V, 0.0, s));
  HIP_TRY(launch_atoms_bar(p->abs, g.L, p->ac.p, p->gen_b.p, gs, p->d.n_beams, am, p->line_el, p->extra.p, s));
  // the absorption pullback gathers its density gradient per box in grad,
  // and the material pullback adds the MassFraction pullback of it
  HIP_TRY(launch_fill(p->grad.p, n_cells * ne, 0.0, s));
  HIP_TRY(launch_absorption_pullback(p->abs, g.L, p->abs_small.p + p->abs_off_mac, p->ac.p, p->rho.p + g.rho_off[0],
                                     am, p->line_el, p->gen_b.p, gs, p->d.n_beams, p->grad.p, s));
  HIP_TRY(launch_material_pullback(g.L, g.class_n, g.rho_off, p->T.n, p->cell_w.p, p->rho_pure.p, p->rho_bar_b.p,
                                   p->rb_stride, p->d.n_beams, p->extra.p, true, p->grad.p, s));
  HIP_TRY(hipMemcpyAsync(grad, p->grad.p, n_cells * ne * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return PNX_OK;
}

int pnx_problem_set_beams(pnx_problem* p, const double* const mms[8], const double* const* ppos) {
  if (!p || !mms || !ppos) return fail(PNX_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(p->ctx->device));
  const Geometry& g = p->g;
  std::vector<double> m(8 * 16, 0.0), pr((size_t)g.prof_stride * p->d.n_beams, 0.0);
  for (int c = 0; c < 8; ++c) {
    if (!odd_in(c, 0) || !g.active(c)) continue;
    if (!mms[c]) return fail(PNX_ERR_ARG, "beam moments missing for an x-odd class");
    for (int i = 0; i < g.T.n[c]; ++i) m[c * 16 + i] = mms[c][i];
    for (int b = 0; b < p->d.n_beams; ++b)
      if (!ppos[b * 8 + c]) return fail(PNX_ERR_ARG, "beam profile missing for an x-odd class");
  }
  for (int b = 0; b < p->d.n_beams; ++b) pack_profiles(g, ppos + b * 8, pr.data() + b * g.prof_stride);
  hipStream_t s = p->stream;
  HIP_TRY(hipMemcpyAsync(p->mms.p, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(p->prof.p, pr.data(), pr.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  p->have_beams = true;
  return PNX_OK;
}

}  // extern "C"

namespace {

HalfStepArgs half_args(pnx_problem* p, int mode, const double* rec, double dt, const double* from, const double* src,
                       double* dst, double* partial) {
  HalfStepArgs a{};
  a.L = p->L;
  a.from = from;
  a.to_src = src;
  a.to_dst = dst;
  a.rho = p->rho.p;
  density_read_offsets(p->g, p->have_cells, a.rho_off);
  a.coef = mode == kForward ? p->coef_f.p : p->coef_a.p;
  a.rec = rec;
  a.prof = p->prof.p;
  a.prof_stride = p->g.prof_stride;
  a.mms = p->mms.p;
  a.W = p->W.p;
  a.ibar = mode == kAdjoint ? p->ibar.p : nullptr;
  a.partial = partial;
  a.partial_stride = p->g.partial_beam;
  a.dt = dt;
  a.n_beams = p->d.n_beams;
  return a;
}

// One energy step of the forward problem (step_pde!): odd half, even full, odd half.
// src -> dst (dst may equal src); partial receives the intensity partials of this step.
int forward_step(pnx_problem* p, int k, const double* src, double* dst, double* partial) {
  hipStream_t s = p->stream;
  const double* rec = p->rec_f.p + (long)k * p->g.rec_len;
  const int N = p->d.N, dims = p->L.dims;
  const double h = p->g.dE / 2;
  HIP_TRY(launch_half_step(N, dims, 1, kForward, half_args(p, kForward, rec, h, src, src, dst, nullptr), s));
  HIP_TRY(launch_half_step(N, dims, 0, kForward, half_args(p, kForward, rec, p->g.dE, dst, src, dst, partial), s));
  HIP_TRY(launch_half_step(N, dims, 1, kForward, half_args(p, kForward, rec, h, dst, dst, dst, nullptr), s));
  return PNX_OK;
}

// Forward energy steps ka..kb with the odd half steps between consecutive
// steps fused into one pass (the even classes they read do not change in
// between): O1(ka), E(ka), [O3(ka) O1(ka+1)], E(ka+1), ..., E(kb), O3(kb).
// Step k reads the state after step k-1 (src0 for k = ka) and writes dst(k);
// dst(k) may equal dst(k-1) (in place).  partial(k) receives the intensity
// partials of step k.
template <class Dst, class Part>
int forward_run(pnx_problem* p, int ka, int kb, const double* src0, Dst dst, Part partial) {
  hipStream_t s = p->stream;
  const int N = p->d.N, dims = p->L.dims;
  const double h = p->g.dE / 2;
  auto rec = [&](int k) { return p->rec_f.p + (long)k * p->g.rec_len; };
  HIP_TRY(launch_half_step(N, dims, 1, kForward, half_args(p, kForward, rec(ka), h, src0, src0, dst(ka), nullptr), s));
  for (int k = ka; k <= kb; ++k) {
    const double* prev = k == ka ? src0 : dst(k - 1);
    double* cur = dst(k);
    HIP_TRY(launch_half_step(N, dims, 0, kForward, half_args(p, kForward, rec(k), p->g.dE, cur, prev, cur, partial(k)),
                             s));
    if (k < kb) {
      HalfStepArgs a = half_args(p, kForward, rec(k), h, cur, cur, dst(k + 1) == cur ? nullptr : cur, nullptr);
      a.rec2 = rec(k + 1);
      a.dt2 = h;
      a.to_dst2 = dst(k + 1);
      HIP_TRY(launch_half_step2(N, dims, kForward, a, s));
    } else {
      HIP_TRY(launch_half_step(N, dims, 1, kForward, half_args(p, kForward, rec(k), h, cur, cur, cur, nullptr), s));
    }
  }
  return PNX_OK;
}

int adjoint_step(pnx_problem* p, int i, double* lam) {
  hipStream_t s = p->stream;
  const double* rec = p->rec_a.p + (long)i * p->g.rec_len;
  const int N = p->d.N, dims = p->L.dims;
  const double h = p->g.dE / 2;
  HIP_TRY(launch_half_step(N, dims, 1, kAdjoint, half_args(p, kAdjoint, rec, h, lam, lam, lam, nullptr), s));
  HIP_TRY(launch_half_step(N, dims, 0, kAdjoint, half_args(p, kAdjoint, rec, p->g.dE, lam, lam, lam, nullptr), s));
  HIP_TRY(launch_half_step(N, dims, 1, kAdjoint, half_args(p, kAdjoint, rec, h, lam, lam, lam, nullptr), s));
  return PNX_OK;
}

// The forward march of every beam; the intensity partial sums go to
// `partial` (partial_beam doubles per beam) and the intensities to I_dev.
int enqueue_forward(pnx_problem* p, double* partial, double* I_dev) {
  hipStream_t s = p->stream;
  HIP_TRY(launch_fill(p->state.p, (long)p->L.beam_stride * p->d.n_beams, 0.0, s));
  double* st = p->state.p;
  int rc = forward_run(
      p, 1, p->d.n_steps, st, [&](int) { return st; },
      [&](int k) { return partial + (long)k * p->g.partial_step; });
  if (rc) return rc;
  HIP_TRY(launch_reduce_intensity(partial, p->g.partial_step, p->g.partial_beam, p->L.n_tiles * kWavesPerBlock,
                                  p->rec_f.p, p->g.rec_len, p->g.sigma_off, p->d.n_steps, p->d.n_lines, p->g.dE,
                                  p->L.dvol, p->d.n_beams, I_dev, s));
  return PNX_OK;
}

int enqueue_forward(pnx_problem* p) { return enqueue_forward(p, p->partial.p, p->I_dev.p); }

int ensure_graph(pnx_problem* p) {
  if (p->fwd_graph) return PNX_OK;
  hipStream_t s = p->stream;
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_forward(p);
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(PNX_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&p->fwd_graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    p->fwd_graph = nullptr;
    return fail(PNX_ERR_DEVICE, std::string("graph instantiate failed: ") + hipGetErrorString(e));
  }
  return PNX_OK;
}

int check_ready(pnx_problem* p, bool adjoint) {
  if (!p) return fail(PNX_ERR_ARG, "null problem");
  if (!p->have_rec || !p->have_mat || !p->have_beams)
    return fail(PNX_ERR_STATE, "records, material and beams must be set before solving");
  if (adjoint && !p->have_adj) return fail(PNX_ERR_STATE, "adjoint records were not set");
  if (hipSetDevice(p->ctx->device) != hipSuccess) return fail(PNX_ERR_DEVICE, "hipSetDevice failed");
  return PNX_OK;
}

int alloc_adjoint(pnx_problem* p, bool exact = false) {
  const size_t st = (size_t)p->L.beam_stride * p->d.n_beams;
  int rc;
  if ((rc = p->lam.alloc(st)) || (rc = p->ckpt.alloc(st * p->g.nck)) ||
      (rc = p->seg.alloc(st * p->g.K * (exact ? 2 : 1))) ||
      (rc = p->rho_bar_b.alloc((size_t)p->rb_stride * p->d.n_beams)) ||
      (rc = p->gen_b.alloc((size_t)p->d.n_lines * p->L.V * p->d.n_beams)) ||
      (rc = p->ibar.alloc((size_t)p->d.n_beams * p->d.n_lines)) ||
      (rc = p->sum_buf.alloc(std::max((size_t)p->rb_stride, (size_t)p->d.n_lines * p->L.V))))
    return rc;
  if (exact && ((rc = p->w1.alloc(st)) || (rc = p->w2.alloc(st)))) return rc;
  return PNX_OK;
}

DensArgs dens_args(pnx_problem* p, int i, const double* u, double int_fac) {
  DensArgs a{};
  a.L = p->L;
  a.u = u;
  a.lam = p->lam.p;
  a.rho = p->rho.p;
  density_read_offsets(p->g, p->have_cells, a.rho_off);
  for (int c = 0; c < 8; ++c) a.rb_off[c] = p->g.rho_off[c];
  a.coef = p->coef_f.p;
  a.rec = p->rec_d.p + (long)i * p->g.rec_len;
  a.prof = p->prof.p;
  a.prof_stride = p->g.prof_stride;
  a.mms = p->mms.p;
  a.W = p->W.p;
  a.ibar = p->ibar.p;
  a.rho_bar = p->rho_bar_b.p;
  a.rb_stride = p->rb_stride;
  a.gen = p->gen_b.p;
  a.gen_stride = (long)p->d.n_lines * p->L.V;
  a.dE = p->g.dE;
  a.int_fac = int_fac;
  a.n_beams = p->d.n_beams;
  return a;
}

const double* rec_adj(pnx_problem* p, int i) { return p->rec_a.p + (long)i * p->g.rec_len; }

// Fused reverse pass at step i on u: the last odd adjoint half step of step
// i + 1 first (if with_a3), the pullback with weight int_fac, the first odd
// adjoint half step of step i last (if with_a1).
RevArgs rev_args(pnx_problem* p, int i, const double* u, double int_fac, bool with_a3, bool with_a1) {
  RevArgs ra{};
  ra.adj = half_args(p, kAdjoint, rec_adj(p, i), p->g.dE / 2, p->lam.p, p->lam.p, p->lam.p, nullptr);
  ra.dens = dens_args(p, i, u, int_fac);
  ra.rec3 = with_a3 ? rec_adj(p, i + 1) : nullptr;
  ra.a1 = with_a1 ? 1 : 0;
  return ra;
}

// Forward pass with checkpoints, then the reverse sweep of the reference
// pullback (pn_intensities.jl:381-404).
int enqueue_forward_adjoint(pnx_problem* p) {
  hipStream_t s = p->stream;
  const long st = p->L.beam_stride * p->d.n_beams;
  const int n = p->d.n_steps, K = p->g.K, N = p->d.N, dims = p->L.dims;
  int rc;
  HIP_TRY(launch_fill(p->state.p, st, 0.0, s));
  double* state = p->state.p;
  auto part_k = [&](int k) { return p->partial.p + (long)k * p->g.partial_step; };
  // One segment (checkpoints every n steps or more): the forward run writes
  // every state into the segment buffer, which the reverse sweep then reads
  // as a tape, so the segment is not recomputed (2n + 1 sweeps instead of
  // 3n + 1, with the same memory).
  const bool taped = p->g.nck == 1;
  if (taped) {
    HIP_TRY(launch_fill(p->ckpt.p, st, 0.0, s));  // u_0 = 0 for the final pullback
    if ((rc = forward_run(
             p, 1, n, state, [&](int k) { return p->seg.p + (long)(k - 1) * st; }, part_k)))
      return rc;
  }
  // otherwise one fused run per checkpoint segment: the state before each segment is stored
  for (int sgi = 0; !taped && sgi < p->g.nck; ++sgi) {
    const int k0 = sgi * K, k1 = std::min(n, k0 + K);
    HIP_TRY(launch_copy(p->ckpt.p + (long)sgi * st, state, st, s));
    if ((rc = forward_run(p, k0 + 1, k1, state, [&](int) { return state; }, part_k))) return rc;
  }
  HIP_TRY(launch_reduce_intensity(p->partial.p, p->g.partial_step, p->g.partial_beam, p->L.n_tiles * kWavesPerBlock, p->rec_f.p, p->g.rec_len,
                                  p->g.sigma_off, n, p->d.n_lines, p->g.dE, p->L.dvol, p->d.n_beams, p->I_dev.p, s));
  HIP_TRY(launch_fill(p->lam.p, st, 0.0, s));
  HIP_TRY(launch_fill(p->rho_bar_b.p, (long)p->rho_bar_b.n, 0.0, s));
  HIP_TRY(launch_fill(p->gen_b.p, (long)p->gen_b.n, 0.0, s));
  for (int sgi = p->g.nck - 1; sgi >= 0; --sgi) {
    const int k0 = sgi * K, k1 = std::min(n, k0 + K);
    // recompute u_{k0+1} .. u_{k1} into the segment buffer (the partials go to
    // the scratch slot of step 0)
    if (!taped && (rc = forward_run(
                       p, k0 + 1, k1, p->ckpt.p + (long)sgi * st,
                       [&](int k) { return p->seg.p + (long)(k - k0 - 1) * st; }, [&](int) { return p->partial.p; })))
      return rc;
    // per reverse step i: A3(i+1) + dens(i) + A1(i) fused, then AE(i)
    for (int i = k1; i > k0; --i) {
      const double* u = p->seg.p + (long)(i - k0 - 1) * st;
      HIP_TRY(launch_rev(N, dims, rev_args(p, i, u, 1.0, i < n, true), s));
      HIP_TRY(launch_half_step(N, dims, 0, kAdjoint,
                               half_args(p, kAdjoint, rec_adj(p, i), p->g.dE, p->lam.p, p->lam.p, p->lam.p, nullptr),
                               s));
    }
  }
  // A3(1) + the final accumulation at u_0 = 0 with weight 0.5 (pn_intensities.jl:402-404)
  HIP_TRY(launch_rev(N, dims, rev_args(p, 0, p->ckpt.p, 0.5, true, false), s));
  return PNX_OK;
}

DiscArgs disc_args(pnx_problem* p, int k, const double* s_from, const double* s_own, const double* w_in,
                   double* w_out, double tau, const double* u_direct) {
  DiscArgs a{};
  a.L = p->L;
  a.s_from = s_from;
  a.s_own = s_own;
  a.lam = p->lam.p;
  a.w_in = w_in;
  a.w_out = w_out;
  a.rho = p->rho.p;
  for (int c = 0; c < 8; ++c) a.rho_off[c] = p->g.rho_off[c];
  a.coef = p->coef_f.p;
  a.rec = p->rec_f.p + (long)k * p->g.rec_len;
  a.prof = p->prof.p;
  a.prof_stride = p->g.prof_stride;
  a.mms = p->mms.p;
  a.W = p->W.p;
  a.ibar = p->ibar.p;
  a.u_direct = u_direct;
  a.rho_bar = p->rho_bar_b.p;
  a.rb_stride = p->rb_stride;
  a.gen = p->gen_b.p;
  a.gen_stride = (long)p->d.n_lines * p->L.V;
  a.tau = tau;
  a.dE = p->g.dE;
  a.n_beams = p->d.n_beams;
  return a;
}

// Forward pass with checkpoints, then the exact adjoint of the discrete march:
// per segment, recompute the states after every step and after its first odd
// half step, and back-propagate each step as three fused passes.
int enqueue_forward_adjoint_exact(pnx_problem* p) {
  hipStream_t s = p->stream;
  const long st = p->L.beam_stride * p->d.n_beams;
  const int n = p->d.n_steps, K = p->g.K, N = p->d.N, dims = p->L.dims;
  const double dE = p->g.dE, h = dE / 2;
  int rc;
  HIP_TRY(launch_fill(p->state.p, st, 0.0, s));
  for (int k = 1; k <= n; ++k) {
    if ((k - 1) % K == 0)
      HIP_TRY(hipMemcpyAsync(p->ckpt.p + (long)((k - 1) / K) * st, p->state.p, st * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
    if ((rc = forward_step(p, k, p->state.p, p->state.p, p->partial.p + (long)k * p->g.partial_step))) return rc;
  }
  HIP_TRY(launch_reduce_intensity(p->partial.p, p->g.partial_step, p->g.partial_beam, p->L.n_tiles * kWavesPerBlock, p->rec_f.p,
                                  p->g.rec_len, p->g.sigma_off, n, p->d.n_lines, dE, p->L.dvol, p->d.n_beams,
                                  p->I_dev.p, s));
  HIP_TRY(launch_fill(p->lam.p, st, 0.0, s));
  HIP_TRY(launch_fill(p->w1.p, st, 0.0, s));
  HIP_TRY(launch_fill(p->w2.p, st, 0.0, s));
  HIP_TRY(launch_fill(p->rho_bar_b.p, (long)p->rho_bar_b.n, 0.0, s));
  HIP_TRY(launch_fill(p->gen_b.p, (long)p->gen_b.n, 0.0, s));
  double* segU = p->seg.p;
  double* segM = p->seg.p + (long)K * st;
  for (int sgi = p->g.nck - 1; sgi >= 0; --sgi) {
    const int k0 = sgi * K, k1 = std::min(n, k0 + K);
    const double* ck = p->ckpt.p + (long)sgi * st;
    for (int k = k0 + 1; k <= k1; ++k) {
      const double* src = k == k0 + 1 ? ck : segU + (long)(k - k0 - 2) * st;
      double* dst = segU + (long)(k - k0 - 1) * st;
      double* mid = segM + (long)(k - k0 - 1) * st;
      const double* rec = p->rec_f.p + (long)k * p->g.rec_len;
      HIP_TRY(launch_half_step(N, dims, 1, kForward, half_args(p, kForward, rec, h, src, src, mid, nullptr), s));
      HIP_TRY(launch_half_step(N, dims, 0, kForward, half_args(p, kForward, rec, dE, mid, src, dst, p->partial.p), s));
      HIP_TRY(launch_half_step(N, dims, 1, kForward, half_args(p, kForward, rec, h, dst, mid, dst, nullptr), s));
    }
    for (int i = k1; i > k0; --i) {
      const double* prev = i == k0 + 1 ? ck : segU + (long)(i - k0 - 2) * st;
      const double* cur = segU + (long)(i - k0 - 1) * st;
      const double* mid = segM + (long)(i - k0 - 1) * st;
      // transposed first odd half of step i+1 and intensity terms of step i; odd half 3 of step i
      HIP_TRY(launch_disc(N, dims, 0, 1, disc_args(p, i, cur, mid, p->w1.p, p->w2.p, h, cur), s));
      // transposed odd half 3; even half 2 (input: even of step i-1, odd after half 1)
      HIP_TRY(launch_disc(N, dims, 0, 0, disc_args(p, i, mid, prev, p->w2.p, p->w2.p, dE, nullptr), s));
      // transposed even half 2; odd half 1 (input: state of step i-1)
      HIP_TRY(launch_disc(N, dims, 1, 1, disc_args(p, i, prev, prev, p->w2.p, p->w1.p, h, nullptr), s));
    }
  }
  return PNX_OK;
}

}  // namespace

extern "C" {

int pnx_forward(pnx_problem* p, double* I_out) {
  int rc = check_ready(p, false);
  if (rc) return rc;
  if (!I_out) return fail(PNX_ERR_ARG, "null output");
  if ((rc = ensure_graph(p))) return rc;
  hipStream_t s = p->stream;
  HIP_TRY(hipGraphLaunch(p->fwd_graph, s));
  HIP_TRY(hipMemcpyAsync(I_out, p->I_dev.p, sizeof(double) * p->d.n_beams * p->d.n_lines, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return PNX_OK;
}

static int forward_adjoint_impl(pnx_problem* p, bool exact, const double* I_bar, double* I_out,
                                double* const rho_bar[8], double* gen_out) {
  int rc = check_ready(p, !exact);
  if (rc) return rc;
  if (!I_bar || !I_out) return fail(PNX_ERR_ARG, "null argument");
  if ((rc = alloc_adjoint(p, exact))) return rc;
  hipStream_t s = p->stream;
  HIP_TRY(hipMemcpyAsync(p->ibar.p, I_bar, sizeof(double) * p->d.n_beams * p->d.n_lines, hipMemcpyHostToDevice, s));
  if ((rc = exact ? enqueue_forward_adjoint_exact(p) : enqueue_forward_adjoint(p))) return rc;
  p->have_rho_bar = true;
  HIP_TRY(hipMemcpyAsync(I_out, p->I_dev.p, sizeof(double) * p->d.n_beams * p->d.n_lines, hipMemcpyDeviceToHost, s));
  const Layout& L = p->L;
  const int ne = p->d.n_elements;
  if (rho_bar) {
    HIP_TRY(launch_beam_sum(p->rho_bar_b.p, p->rb_stride, p->d.n_beams, p->rb_stride, p->sum_buf.p, s));
    std::vector<double> h((size_t)p->rb_stride);
    HIP_TRY(hipMemcpyAsync(h.data(), p->sum_buf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int c = 0; c < 8; ++c) {
      if (!rho_bar[c]) continue;
      if (p->g.active(c)) {
        unpack_class(p->g, c, h.data() + p->g.rho_off[c], ne, rho_bar[c]);
      } else {
        const long np = (long)p->g.class_n[c][0] * p->g.class_n[c][1] * p->g.class_n[c][2];
        std::fill(rho_bar[c], rho_bar[c] + np * ne, 0.0);
      }
    }
  }
  if (gen_out) {
    const long gs = (long)p->d.n_lines * L.V;
    HIP_TRY(launch_beam_sum(p->gen_b.p, gs, p->d.n_beams, gs, p->sum_buf.p, s));
    std::vector<double> h((size_t)gs);
    HIP_TRY(hipMemcpyAsync(h.data(), p->sum_buf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    unpack_lines(p->g, h.data(), gen_out);
  }
  HIP_TRY(hipStreamSynchronize(s));
  return PNX_OK;
}

int pnx_forward_noalloc(pnx_problem* p, double* I_out) {
  int rc = check_ready(p, false);
  if (rc) return rc;
  if (!I_out) return fail(PNX_ERR_ARG, "null output");
  double* sp = p->ctx->scratch;
  if (!sp) return fail(PNX_ERR_STATE, "the context has no scratch address");
  // partial sums [beam][step][wavefront][line], then the intensities 256-byte aligned
  const size_t n_part = (size_t)p->g.partial_beam * p->d.n_beams;
  const size_t off_I = (n_part + 31) / 32 * 32;
  const size_t need = off_I + (size_t)p->d.n_beams * p->d.n_lines;
  if (need > kScratchDoubles)
    return fail(PNX_ERR_ARG, "the allocation-free forward needs " + std::to_string(need * sizeof(double)) +
                                 " bytes of scratch for this problem; its maximum is " +
                                 std::to_string(kScratchDoubles * sizeof(double)) +
                                 " bytes (fewer beams or energy steps, or use pnx_forward)");
  if ((rc = enqueue_forward(p, sp, sp + off_I))) return rc;
  hipStream_t s = p->stream;
  HIP_TRY(hipMemcpyAsync(I_out, sp + off_I, sizeof(double) * p->d.n_beams * p->d.n_lines, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return PNX_OK;
}

int pnx_problem_reserve_adjoint(pnx_problem* p, int exact) {
  if (!p) return fail(PNX_ERR_ARG, "null problem");
  if (hipSetDevice(p->ctx->device) != hipSuccess) return fail(PNX_ERR_DEVICE, "hipSetDevice failed");
  return alloc_adjoint(p, exact != 0);
}

int pnx_forward_adjoint(pnx_problem* p, const double* I_bar, double* I_out, double* const rho_bar[8], double* gen_out) {
  return forward_adjoint_impl(p, false, I_bar, I_out, rho_bar, gen_out);
}

int pnx_forward_adjoint_exact(pnx_problem* p, const double* I_bar, double* I_out, double* const rho_bar[8],
                              double* gen_out) {
  return forward_adjoint_impl(p, true, I_bar, I_out, rho_bar, gen_out);
}

int pnx_step(pnx_problem* p, int adjoint, int record, int beam, const double* I_bar, double* const u[8]) {
  int rc = check_ready(p, adjoint != 0);
  if (rc) return rc;
  if (!u || beam < 0 || beam >= p->d.n_beams || record < 0 || record > p->d.n_steps)
    return fail(PNX_ERR_ARG, "bad argument");
  const Layout& L = p->L;
  hipStream_t s = p->stream;
  std::vector<double> h((size_t)L.beam_stride, 0.0);
  for (int c = 0; c < 8; ++c) {
    if (!p->g.active(c)) continue;
    if (!u[c]) return fail(PNX_ERR_ARG, "state missing for an updated class");
    pack_class(p->g, c, u[c], p->T.n[c], h.data() + L.cls_off[c]);
  }
  double* sb = p->state.p + (long)beam * L.beam_stride;
  HIP_TRY(hipMemcpyAsync(sb, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  if (adjoint) {
    if ((rc = alloc_adjoint(p))) return rc;
    std::vector<double> ib((size_t)p->d.n_beams * p->d.n_lines, 0.0);
    if (I_bar)
      for (int r = 0; r < p->d.n_lines; ++r) ib[beam * p->d.n_lines + r] = I_bar[r];
    HIP_TRY(hipMemcpyAsync(p->ibar.p, ib.data(), ib.size() * sizeof(double), hipMemcpyHostToDevice, s));
    if ((rc = adjoint_step(p, record, p->state.p))) return rc;
  } else {
    if ((rc = forward_step(p, record, p->state.p, p->state.p, p->partial.p))) return rc;
  }
  HIP_TRY(hipMemcpyAsync(h.data(), sb, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int c = 0; c < 8; ++c)
    if (p->g.active(c) && u[c]) unpack_class(p->g, c, h.data() + L.cls_off[c], p->T.n[c], u[c]);
  return PNX_OK;
}

int pnx_forward_timed(pnx_problem* p, int reps, int use_graph, double* ms_out) {
  int rc = check_ready(p, false);
  if (rc) return rc;
  if (reps < 1 || !ms_out) return fail(PNX_ERR_ARG, "bad argument");
  if (use_graph && (rc = ensure_graph(p))) return rc;
  hipStream_t s = p->stream;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) {
    if (use_graph) {
      HIP_TRY(hipGraphLaunch(p->fwd_graph, s));
    } else if ((rc = enqueue_forward(p))) {
      return rc;
    }
  }
  HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_out = ms / reps;
  return PNX_OK;
}

int pnx_forward_adjoint_timed(pnx_problem* p, int reps, double* ms_out) {
  int rc = check_ready(p, true);
  if (rc) return rc;
  if (reps < 1 || !ms_out) return fail(PNX_ERR_ARG, "bad argument");
  if ((rc = alloc_adjoint(p))) return rc;
  hipStream_t s = p->stream;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r)
    if ((rc = enqueue_forward_adjoint(p))) return rc;
  HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_out = ms / reps;
  return PNX_OK;
}

int pnx_rev_bench(pnx_problem* p, int reps, double* ms_out, double* bytes_out) {
  int rc = check_ready(p, true);
  if (rc) return rc;
  if (reps < 1 || p->d.n_steps < 2) return fail(PNX_ERR_ARG, "bad argument");
  const Layout& L = p->L;
  // algorithmic bytes: u and lambda of every class read once, odd lambda
  // written, densities read and rho_bar read and written at every class
  // point, generation sums read and written at eee points
  double bytes = 0;
  for (int c = 0; c < 8; ++c) {
    if (!(class_active(c, L.dims) && p->T.n[c] > 0)) continue;
    double np = 1;
    for (int a = 0; a < L.dims; ++a) np *= odd_in(c, a) ? p->d.M[a] - 1 : p->d.M[a];
    bytes += np * p->T.n[c] * 8.0 * (parity_of(c) == 1 ? 3 : 2);
    bytes += np * p->d.n_elements * 24.0;
    if (c == 0) bytes += np * p->d.n_lines * 16.0;
  }
  bytes *= p->d.n_beams;
  if (bytes_out) *bytes_out = bytes;
  if (!ms_out) return PNX_OK;
  hipStream_t s = p->stream;
  // lambda, rho_bar, the generation sums and the seed are allocated by the
  // first adjoint solve; a problem that has not had one gets them here
  if ((rc = alloc_adjoint(p))) return rc;
  HIP_TRY(launch_fill(p->lam.p, (long)p->lam.n, 0.0, s));
  HIP_TRY(launch_fill(p->rho_bar_b.p, (long)p->rho_bar_b.n, 0.0, s));
  HIP_TRY(launch_fill(p->gen_b.p, (long)p->gen_b.n, 0.0, s));
  HIP_TRY(launch_fill(p->ibar.p, (long)p->ibar.n, 0.0, s));
  // step 1 with the last odd half step of step 2 folded in, on the forward state
  const RevArgs ra = rev_args(p, 1, p->state.p, 1.0, true, true);
  HIP_TRY(launch_rev(p->d.N, L.dims, ra, s));  // warm-up
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) HIP_TRY(launch_rev(p->d.N, L.dims, ra, s));
  HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_out = ms / reps;
  return PNX_OK;
}

int pnx_half_step_bench(pnx_problem* p, int parity, int reps, double* ms_out, double* bytes_out) {
  int rc = check_ready(p, false);
  if (rc) return rc;
  if (reps < 1 || (parity != 0 && parity != 1)) return fail(PNX_ERR_ARG, "bad argument");
  const Layout& L = p->L;
  // algorithmic bytes: read+write of the updated classes' interior, one read of
  // every non-ghost point of the other parity, the densities of the updated
  // points, and (even step) the line weights at eee points
  double upd = 0, src = 0, pts = 0;
  for (int c = 0; c < 8; ++c) {
    if (!(class_active(c, L.dims) && p->T.n[c] > 0)) continue;
    double np = 1;
    for (int a = 0; a < L.dims; ++a) np *= odd_in(c, a) ? p->d.M[a] - 1 : p->d.M[a];
    if (parity_of(c) == parity) {
      upd += np * p->T.n[c] * 8.0;
      pts += np;
    } else {
      src += np * p->T.n[c] * 8.0;
    }
  }
  double bytes = 2 * upd + src + 8.0 * p->d.n_elements * pts;
  if (parity == 0) bytes += 8.0 * L.n_cells * p->d.n_lines;
  bytes *= p->d.n_beams;
  if (bytes_out) *bytes_out = bytes;
  if (!ms_out) return PNX_OK;
  hipStream_t s = p->stream;
  const double* rec = p->rec_f.p + p->g.rec_len;
  const double dt = parity == 0 ? p->g.dE : p->g.dE / 2;
  HalfStepArgs a = half_args(p, kForward, rec, dt, p->state.p, p->state.p, p->state.p,
                             parity == 0 ? p->partial.p + p->g.partial_step : nullptr);
  HIP_TRY(launch_half_step(p->d.N, L.dims, parity, kForward, a, s));  // warm-up
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) HIP_TRY(launch_half_step(p->d.N, L.dims, parity, kForward, a, s));
  HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_out = ms / reps;
  return PNX_OK;
}

}  // extern "C"
