/* pnx -- MI355X-native PN-CSD forward solve and adjoint (C ABI).
 *
 * Drop-in for the reference's k-ratio hot path.  A Julia host binds these with
 * ccall; the Python package pnepma_amd binds them with
 * ctypes.  Plain C types only; every call returns 0 on success and a negative
 * status on failure, with a thread-local message in pnx_last_error().
 *
 * Layouts at this boundary are the reference's: per-class arrays of points
 * (x fastest, then y, then z; the class point counts are the reference's
 * staggered sizes) with the per-point vector (elements, moments or lines)
 * fastest.  The library converts to its device layout internally.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   pnx_problem_create      ForwardPNProblem{N,T}(...)          pn_solver/pn_problem.jl:72-106
 *   pnx_forward             calc_intensities(problem, m, ...)   pn_solver/pn_intensities.jl:325-351
 *   pnx_forward_adjoint     Zygote.@adjoint calc_intensities    pn_solver/pn_intensities.jl:353-413
 *   pnx_step                step_pde!(problem, u, du, t, dt)    starmap/src/solver.jl:171-196
 */
#ifndef PNX_H
#define PNX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PNX_OK = 0,
  PNX_ERR_ARG = -1,      /* invalid argument */
  PNX_ERR_DEVICE = -2,   /* HIP runtime error */
  PNX_ERR_MEMORY = -3,   /* device allocation failed or budget exceeded */
  PNX_ERR_STATE = -4,    /* call out of order (e.g. solve before inputs are set) */
  PNX_ERR_UNSUPPORTED = -5,
  PNX_ERR_CANCELLED = -6 /* pnx_timing.status of a served request cancelled before completion */
};

typedef struct pnx_context pnx_context;
typedef struct pnx_problem pnx_problem;

/* Geometry and sizes of one solve batch: n_beams beams that share grid,
 * material and energy grid (ForwardPNProblem{N,Float64}). */
typedef struct {
  int32_t N;          /* PN order, 1..9 (odd orders are compiled) */
  int32_t M[3];       /* grid points per axis; 1 = singleton; non-singleton axes first */
  double lo[3];       /* axis extents in cm (singleton: position) */
  double hi[3];
  int32_t n_elements; /* 1..8 */
  int32_t n_lines;    /* 1..8 */
  int32_t n_steps;    /* energy steps */
  int32_t n_beams;    /* beams in the batch */
  double E_initial;   /* eV */
  double E_cutoff;    /* eV */
  int32_t ckpt_every; /* adjoint checkpoint interval in energy steps; 0 or >= n_steps: the whole
                         forward march is kept as a tape and not recomputed */
} pnx_desc;

const char* pnx_last_error(void);
int pnx_version(void);

/* Context: one device, one HIP stream. */
int pnx_context_create(int device, pnx_context** out);
int pnx_context_destroy(pnx_context* ctx);
int pnx_context_synchronize(pnx_context* ctx);

/* Problem: allocates the device state of the batch. */
int pnx_problem_create(pnx_context* ctx, const pnx_desc* desc, pnx_problem** out);
int pnx_problem_destroy(pnx_problem* p);
/* Device bytes held by the problem (state, checkpoints, inputs). */
int64_t pnx_problem_device_bytes(const pnx_problem* p);

/* Energy-point records (n_steps + 1 each, record length
 * n_e + n_e*(N+1) + 1 + n_lines doubles: s[e], c[e][l], gfac, sigma[r]).
 *   fwd[k]  : coefficients at E_{k-1}-dE/2, sigma at E_k     (pn_intensities.jl:341-349)
 *   adj[i]  : adjoint coefficients and sigma at E_i + dE/2   (pn_intensities.jl:397-399)
 *   dens[i] : forward coefficients, gfac, sigma at E_i       (pn_intensities.jl:394-395)
 * adj and dens may be NULL for forward-only use. */
int pnx_problem_set_records(pnx_problem* p, const double* fwd, const double* adj, const double* dens);

/* Physical tables of a problem's elements and lines, from which the library
 * builds the energy records (pnx_problem_records).  Per element e: the
 * stopping power sp[e] (eV cm^2/g) at the knots sp_E (eV, ascending; the
 * reference's ICRU-77 table divided by the density, stopping_power.jl:37-44)
 * and the transport coefficients tc[e] ([n_tc][tc_stride] cm^2/g, degree l
 * in column l, l = 0..N) at the energies tc_E (eV, ascending), interpolated
 * linearly in log E and clamped to the table.  Per line r: line_edge[r], the
 * K edge (eV) of its element.  The beam energy distribution is
 * Normal(beam_E_mean, beam_E_sd). */
typedef struct {
  int32_t n_knots;
  const double* sp_E;
  const double* sp[8];
  int32_t n_tc;
  int32_t tc_stride;
  const double* tc_E;
  const double* tc[8];
  double line_edge[8];
  double beam_E_mean;
  double beam_E_sd;
} pnx_physics;

/* The energy records of pnx_problem_set_records built on the host in C++
 * (update_coefficients!, pn_solver/pn_problem.jl:157-190;
 * update_boundary_condition!, :205-211; the energies of
 * pn_solver/pn_intensities.jl:10-30, 341-401): fwd, adj and dens are
 * (n_steps + 1) * record length doubles each; adj and dens may be NULL
 * (forward only).  The emission cross section is the library's K-shell
 * stand-in 6.51e-14 * 2 * 0.9 ln(U) / (U E_K^2), U = E / E_K (the
 * reference's NeXLCore Bote2009 table is not available).  Fails with
 * PNX_ERR_ARG if an energy lies outside the stopping-power table. */
int pnx_problem_records(const pnx_desc* desc, const pnx_physics* physics, double* fwd, double* adj, double* dens);

/* Beam moments of a von Mises-Fisher beam direction (beam_moments,
 * pn_solver/pn_matrices.jl:331-338) in the Funk-Hecke closed form
 * -Y_lk(direction) i_l(kappa) / i_0(kappa): mms[c] receives the n_c values of
 * every x-odd class c (NULL for the others). */
int pnx_beam_moments(int N, const double* direction, double kappa, double* const mms[8]);

/* Beam position profiles of one beam (setup_boundary_condition!,
 * pn_solver/pn_problem.jl:192-203): ppos[c] receives the isotropic normal
 * pdf (standard deviation sigma) on the (y, z) points of x-odd class c, y
 * fastest (classes whose ppos[c] is NULL are skipped).  literal != 0 evaluates it at x = the first x point,
 * as the reference does (the factor exp(-(x_min - x_b)^2 / (2 sigma^2))
 * underflows for deep slabs); literal = 0 drops that global factor. */
int pnx_beam_profiles(const pnx_desc* desc, const double* beam_position, double sigma, int literal,
                      double* const ppos[8]);

/* Material inputs.  rho[c] (c = 0..7, eee..ooo) is [point][element] over the
 * reference's class-c points; entries for classes the grid never updates may be
 * NULL.  W is [eee point][line] = absorption factor * number of atoms of the
 * line's element (x_ray_intensities.jl:183). */
int pnx_problem_set_material(pnx_problem* p, const double* const rho[8], const double* W);

/* Material inputs from per-cell mass fractions, with the densities computed on
 * the device.  Replaces the host side of component_densities!
 * (pn_solver/pn_parametrizations/pn_material.jl:233-264) for a MassFraction
 * material whose box parametrization (pn_parametrization_boxed.jl:137-175)
 * has one box per grid cell: boxes centred on the even points, edges at the
 * odd points.  w is [cell][element] with cells x fastest (Mx*My*Mz cells),
 * rho_pure[element] the pure-element densities, W as pnx_problem_set_material. */
int pnx_problem_set_cell_fractions(pnx_problem* p, const double* w, const double* rho_pure, const double* W);

/* Gradient of sum(rho_bar * rho) with respect to the cell fractions, through
 * the same map (component_densities_pullback!, pn_material.jl, and the box
 * parametrization's pullback).  rho_bar is the density gradient of the last
 * pnx_forward_adjoint, summed over beams; extra_eee ([eee point][element],
 * reference layout, may be NULL) is added at eee points (the number-of-atoms
 * term of pn_intensities.jl:406-410).  grad is [cell][element].  Requires
 * pnx_problem_set_cell_fractions. */
int pnx_problem_cell_fraction_gradient(pnx_problem* p, const double* extra_eee, double* grad);

/* Absorption factors on the device (compute_absorption_coefficients,
 * pn_solver/x_ray_intensities.jl:100-128): from every eee point a 100-sample
 * march towards `detector`, cut at the plane x = beam_x0, through the boxes of
 * pnx_problem_set_cell_fractions (the embedding fractions outside them).  Also
 * sets the line weights W = ac * number of atoms of the line's element.
 * mac[line][element] mass absorption coefficients, atom_mass[element] (g),
 * line_element[line] element index of each line, axis_points[a] the M_a even
 * points of axis a, cell_edges[a] its M_a + 1 box edges (a singleton axis:
 * one box, two edges).  ac_out ([eee point][line], may be NULL) receives the
 * factors.  Requires pnx_problem_set_cell_fractions. */
int pnx_problem_set_absorption(pnx_problem* p, const double* mac, const double* atom_mass,
                               const int32_t* line_element, const double* detector, double beam_x0,
                               const double* const axis_points[3], const double* const cell_edges[3],
                               const double* embedding, double* ac_out);

/* The whole gradient of sum_{b,r} I_bar I with respect to the cell fractions
 * after pnx_forward_adjoint (the material pullbacks of pn_intensities.jl:406-410:
 * densities, number of atoms and absorption factors,
 * x_ray_intensities.jl:131-153, with its fixed-factor approximation), on the
 * device.  grad is [cell][element], cells x fastest.  The absorption term is
 * accumulated with atomic additions, so its rounding depends on their order.
 * Requires pnx_problem_set_cell_fractions and pnx_problem_set_absorption. */
int pnx_problem_material_gradient(pnx_problem* p, double* grad);

/* Beam source.  mms[c] = beam moments of each x-odd class c (n_c values),
 * ppos[b*8 + c] = position pdf over the (y, z) points of class c for beam b
 * (y fastest); NULL for x-even classes (pn_problem.jl:192-203). */
int pnx_problem_set_beams(pnx_problem* p, const double* const mms[8], const double* const* ppos);

/* calc_intensities for every beam: I_out[b*n_lines + r]. */
int pnx_forward(pnx_problem* p, double* I_out);

/* pnx_forward that allocates nothing: no hipMalloc, hipFree or graph, only
 * kernel launches, one device-to-host copy of I_out and a stream
 * synchronisation.  No counterpart in the reference; it is for callers that
 * cannot allocate device memory while they run (code under HIP graph capture
 * of their own, tight serving loops, harnesses that forbid allocation).  The
 * intensity partial sums go to a fixed device scratch array of 64 MiB per
 * device, held until the process ends.  A problem needs (n_steps + 1) *
 * n_beams * n_tiles * 4 * n_lines + n_beams * n_lines doubles of it, n_tiles
 * being the number of 256-cell tiles covering the grid (256 x 1 in 1D, 32 x 8
 * per z-plane in 2D and 3D): the 2D P7 128x96 raster with 32 beams, 3 lines
 * and 400 steps needs 59 MB.  A larger problem returns PNX_ERR_ARG without
 * launching anything.  The scratch needs no clearing: every
 * value is written before it is read.  Calls on one device share that array,
 * so calls that overlap on one device (different threads or streams) race;
 * serialise them.  Results equal pnx_forward's. */
int pnx_forward_noalloc(pnx_problem* p, double* I_out);

/* Forward pass, then the reference adjoint sweep with checkpointing every
 * ckpt_every steps and forward recomputation between checkpoints.
 *   I_bar[b*n_lines + r]   seed of the pullback per beam
 *   I_out[b*n_lines + r]   intensities
 *   rho_bar[c]             [point][element] per class, summed over beams (NULL to skip a class)
 *   gen_out                [eee point][line], summed over beams:
 *                          sum_i u_i,eee(p,0) * dE*dvol*sigma_r(E_i)*int_fac_i*I_bar_r,
 *                          from which ac_bar = noa[el_r] * gen and noa_bar[el_r] += ac * gen
 *                          (x_ray_intensities.jl:193-223). May be NULL. */
int pnx_forward_adjoint(pnx_problem* p, const double* I_bar, double* I_out, double* const rho_bar[8], double* gen_out);

/* Allocates the adjoint sweep's device buffers (adjoint state, checkpoints,
 * segment or tape, gradients) now rather than in the first
 * pnx_forward_adjoint (exact = 0) or pnx_forward_adjoint_exact (exact = 1).
 * With ckpt_every >= n_steps the whole forward march is kept (a tape of
 * n_steps states) and the reverse sweep needs no recompute. */
int pnx_problem_reserve_adjoint(pnx_problem* p, int exact);

/* Exact gradient of sum_{b,r} I_bar[b][r] I[b][r] of the discrete forward
 * march (the transpose of every half step, including the exponential factor
 * and the boundary ghosts), with the same checkpointing.  Outputs as
 * pnx_forward_adjoint; gen_out here is d/dW (sigma at the forward energies).
 * The reference's adjoint (pnx_forward_adjoint) is a discretised continuous
 * adjoint and agrees with finite differences only to discretisation error;
 * this variant agrees with central differences of pnx_forward to their
 * truncation error.  It has no counterpart in the reference. */
int pnx_forward_adjoint_exact(pnx_problem* p, const double* I_bar, double* I_out, double* const rho_bar[8],
                              double* gen_out);

/* One energy step (step_pde!) on a host state for beam `beam`: u[c] is the
 * reference's [point][moment] array of class c (ghost cells ignored on input,
 * undefined on output); classes the grid never updates may be NULL.
 * adjoint != 0 runs the adjoint problem with record index `record` of adj and
 * the eee m=0 source -sum_r sigma_r W_r I_bar_r (I_bar may be NULL for none);
 * otherwise the forward problem with record `record` of fwd. */
int pnx_step(pnx_problem* p, int adjoint, int record, int beam, const double* I_bar, double* const u[8]);

/* Timing support: run the forward march `reps` times on device-resident
 * inputs without host copies; returns the elapsed milliseconds of the last
 * run measured with HIP events on the problem's stream. */
int pnx_forward_timed(pnx_problem* p, int reps, int use_graph, double* ms_out);
int pnx_forward_adjoint_timed(pnx_problem* p, int reps, double* ms_out);

/* Average duration (ms) of one launch of the fused half-step kernel of the
 * given parity (0 even, 1 odd), measured with HIP events around `reps`
 * back-to-back launches on the problem's stream, and the algorithmic bytes
 * of one such launch. */
int pnx_half_step_bench(pnx_problem* p, int parity, int reps, double* ms_out, double* bytes_out);
/* The same for the fused reverse pass of the reference adjoint (the last odd
 * adjoint half step of step 2, the density pullback of step 1 and the first
 * odd adjoint half step of step 1) on the problem's forward state; needs the
 * adjoint records and at least two energy steps. */
int pnx_rev_bench(pnx_problem* p, int reps, double* ms_out, double* bytes_out);

/* ------------------------------------------------------------------ serving
 * A continuous-batching server of forward+adjoint requests with mixed PN
 * order, grid and energy range.  Requests are admitted, in arrival order, as
 * soon as their arrival time has passed and their device memory fits in the
 * HBM budget (taped or checkpointed, see pnx_server_set_tape_limit).  Every
 * scheduler tick advances each admitted 2D/3D request by one operation of its
 * sequence (a forward, recomputed or reverse energy step) in grouped kernels
 * that batch all requests of one PN order and dimension, and each 1D request
 * by a slice of operations in one march kernel.  The reference has no
 * counterpart (it solves one problem at a time); each request runs the march
 * of pnx_forward_adjoint with one beam. */
typedef struct pnx_server pnx_server;

typedef struct {
  pnx_desc desc;                 /* n_beams must be 1; ckpt_every 0 = about sqrt(n_steps) */
  const double* fwd;             /* records as pnx_problem_set_records */
  const double* adj;
  const double* dens;
  const double* rho[8];          /* as pnx_problem_set_material */
  const double* W;
  const double* mms[8];          /* as pnx_problem_set_beams, one beam */
  const double* ppos[8];
  const double* I_bar;           /* n_lines pullback seed */
  double arrival_s;              /* admission not before this server-clock time */
  double* I_out;                 /* n_lines, written at completion */
  double* rho_bar[8];            /* optional [point][element] per class, written at completion */
  double* gen_out;               /* optional [eee point][line], written at completion */
} pnx_request;

/* A served request given as what the reference's calc_intensities takes
 * (pn_solver/pn_intensities.jl:325-335): the problem's geometry and energy
 * grid, its element tables, a MassFraction material with one box per grid
 * cell (boxes centred on the even points, edges at the odd points) and the
 * detector.  At submit the library builds the records, beam moments and beam
 * profiles (as pnx_problem_records, pnx_beam_moments, pnx_beam_profiles); at
 * admission the densities, absorption factors and line weights are computed
 * on the device (as pnx_problem_set_cell_fractions and
 * pnx_problem_set_absorption); at completion, if grad_out is given, the whole
 * material gradient (as pnx_problem_material_gradient).  Inputs are copied at
 * submit. */
typedef struct {
  pnx_desc desc;                 /* n_beams must be 1; ckpt_every 0 = about sqrt(n_steps) */
  const pnx_physics* physics;
  const double* cell_w;          /* [cell][element] mass fractions, cells x fastest (Mx*My*Mz) */
  const double* rho_pure;        /* [element] pure-element densities (g/cm^3) */
  const double* embedding;       /* [element] fractions outside the boxes (for the absorption rays) */
  const double* mac;             /* [line][element] mass absorption coefficients (cm^2/g) */
  const double* atom_mass;       /* [element] atomic masses (g) */
  const int32_t* line_element;   /* [line] element index of each line */
  double detector[3];            /* detector position (cm) */
  double beam_position[3];       /* beam centre; x = the surface, where the absorption rays are cut */
  double beam_sigma;             /* beam profile standard deviation (cm) */
  double beam_direction[3];      /* von Mises-Fisher mean direction */
  double beam_kappa;             /* and concentration */
  int32_t literal_source;        /* 1: the reference's beam pdf at x_min (see pnx_beam_profiles) */
  const double* I_bar;           /* n_lines pullback seed */
  double arrival_s;              /* admission not before this server-clock time */
  double* I_out;                 /* n_lines, written at completion */
  double* rho_bar[8];            /* optional [point][element] per class; when all are null on a 2D
                                    request of at most 4 elements, the reverse passes sum a cell's
                                    classes into one per-cell gradient that only grad_out uses */
  double* gen_out;               /* optional [eee point][line] */
  double* grad_out;              /* optional [cell][element]: d(I_bar . I)/d(cell fractions) */
} pnx_request_spec;

typedef struct {
  double arrival_s;   /* server clock, seconds */
  double admit_s;
  double done_s;
  int32_t status;     /* PNX_OK or an error code */
  int32_t ticks;      /* scheduler ticks the request took part in */
  int32_t taped;      /* 1: admitted taped (every forward state kept, no recompute) */
} pnx_timing;

/* Inputs are copied at submit; output pointers must stay valid until done. */
int pnx_server_create(pnx_context* ctx, int64_t hbm_budget_bytes, int32_t max_active, pnx_server** out);
int pnx_server_destroy(pnx_server* s);
double pnx_server_clock(const pnx_server* s);
/* Memory-for-work trade of the admission: a request is admitted taped (its n
 * forward states kept in HBM, so the reverse sweep needs no recompute: 2n + 1
 * state sweeps instead of 3n + 1) while the arena's bytes in use plus its
 * taped size stay within `bytes`.  Otherwise it is admitted checkpointed if
 * its lane is short of work (fewer than 32 active 2D/3D or 128 active 1D
 * requests), and waits for memory if not.  0 disables taping (every request
 * checkpointed, admitted as soon as it fits).  Default: 95 % of the HBM
 * budget (on the serving stream 183 solves/s, against 172 at 75 %).  Results agree with the
 * checkpointed march to rounding (the fused odd pairs run over one run
 * instead of one run per segment).  No counterpart in the reference, whose
 * pullback always keeps the whole tape (pn_solver/pn_intensities.jl:353-404). */
int pnx_server_set_tape_limit(pnx_server* s, int64_t bytes);
int pnx_server_submit(pnx_server* s, const pnx_request* req, int64_t* ticket);
/* Submits a request given as cell fractions (pnx_request_spec): its
 * host preparation is the library's C++ set-up in the caller's thread. */
int pnx_server_submit_spec(pnx_server* s, const pnx_request_spec* req, int64_t* ticket);
int pnx_server_poll(pnx_server* s, int64_t ticket, int32_t* done, pnx_timing* t);
/* Blocks until the request is complete; this releases the ticket (a later
 * poll, wait or cancel of it returns PNX_ERR_ARG). */
int pnx_server_wait(pnx_server* s, int64_t ticket, pnx_timing* t);
/* Cancels a request: a pending one is dropped at once, an admitted one issues
 * no further passes and completes after its lane's launched work has drained.
 * Its outputs are not written and its pnx_timing.status is PNX_ERR_CANCELLED.
 * A request that has already completed is left as it is.  The reference has no
 * counterpart (it runs one blocking solve per call). */
int pnx_server_cancel(pnx_server* s, int64_t ticket);
/* Block until every submitted request has completed. */
int pnx_server_drain(pnx_server* s);
/* ticks, completed requests, peak concurrently admitted requests, peak device bytes in use */
int pnx_server_stats(pnx_server* s, int64_t* ticks, int64_t* completed, int64_t* peak_active, int64_t* peak_bytes);
/* Device bytes a request would hold while admitted. */
int64_t pnx_request_device_bytes(const pnx_desc* desc);
/* The operation sequence the server runs for a request with n energy steps and
 * checkpoints every K (K = 0: taped, 2n + 1 operations; else 3n + 1), for
 * inspection: for each operation t,
 * out[9t .. 9t+8] = kind (1 forward, 2 recompute, 3 reverse, 4 final),
 * energy step, checkpoint slot written before the step (-1: none), source
 * buffer and slot, destination buffer and slot (0 state, 1 checkpoints,
 * 2 segment), first (the step starts a forward run), fuse (the step ends with
 * the fused odd pair).  Returns the operation count (also when out is NULL). */
int pnx_march_sequence(int32_t n, int32_t K, int32_t* out, int32_t capacity);

/* Matrices the kernels use, for inspection: the coefficient buffer of order N
 * (adjoint = 0 forward, 1 adjoint): all advection nonzeros, then dense
 * boundary blocks.  Returns the number of doubles when out is NULL. */
int pnx_coefficients(int N, int adjoint, double* out, int capacity);
/* Dense advection block A_d[c] (n_c x n_switch, forward sign) and boundary
 * block B_d[c] (n_c x n_switch, c odd in d) as the library builds them. */
int pnx_advection_block(int N, int c, int d, double* out);
int pnx_boundary_block(int N, int c, int d, int adjoint, double* out);

#ifdef __cplusplus
}
#endif

#endif
