/* CPU oracle of the PN-CSD forward solve and its adjoint -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference algorithm, kept structurally close to
 * it: array-of-structs state (point-major, moment fastest, x fastest), CSC
 * advection matrices applied column by column, S and C rebuilt at every point
 * every step, ghost cells stored and filled by the boundary pass, a separate
 * update pass, and a full state tape for the adjoint.  Only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it; the product
 * never does.
 *
 * Reference files followed (paths relative to the reference root):
 *   starmap/src/solver_core.jl, solver.jl, boundary_conditions.jl,
 *   staggered_grid.jl, even_odd_classification.jl;
 *   pn_solver/moment_specification.jl, pn_matrices.jl, pn_problem.jl,
 *   pn_intensities.jl, x_ray_intensities.jl.
 *
 * Energy-point record (one per energy at which the reference evaluates its
 * coefficients), n_e + n_e*(N+1) + 1 + n_lines doubles:
 *   s[e]            stopping power sp_e(E)
 *   c[e][l]         per-degree coefficient (forward: -sp_e'(E) - (tc_e[l]-tc_e[0]),
 *                   adjoint: -(tc_e[l]-tc_e[0]))
 *   gfac            beam energy bin factor (Phi(E+dE)-Phi(E))/dE (forward only)
 *   sigma[r]        emission cross section of line r at the record's sigma energy
 */
#ifndef PNX_REF_H
#define PNX_REF_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int N;           /* PN order */
  int M[3];        /* grid points per axis, 1 = singleton */
  double lo[3];    /* axis extents (singleton: position) */
  double hi[3];
  int n_e;         /* elements */
  int n_lines;     /* x-ray lines */
  int n_steps;     /* energy steps */
  double E_initial;
  double E_cutoff;
} ref_desc;

/* Number of moments in class c (0..7 = eee..ooo) and the class point count. */
int ref_class_moments(int N, int c);
int ref_class_points(const ref_desc* d, int c);
/* Doubles in one energy-point record. */
int ref_record_len(const ref_desc* d);

/* Advection block A_d[c] as dense row-major n_c x n_switch (forward sign),
 * boundary block B_d[c] (forward: L*A, adjoint: -L^T*A) for c odd in d. */
int ref_advection_dense(int N, int c, int d, double* out);
int ref_boundary_dense(int N, int c, int d, int adjoint, double* out);
int ref_boundary_L(int N, int c, int d, double* out);

/* Inputs shared by the drivers.
 *   rho[c]    class-c densities, [point][e] (all 8 classes)
 *   ac        absorption factor at eee points, [p][r]
 *   noa       number of atoms at eee points, [p][e]
 *   el_idx    element index of each line
 *   mms[c]    beam moments of each x-odd class c (NULL for x-even classes)
 *   ppos[c]   position pdf on the (y, z) points of class c, y fastest
 *             (NULL for x-even classes) */
typedef struct {
  const double* rho[8];
  const double* ac;
  const double* noa;
  const int* el_idx;
  const double* mms[8];
  const double* ppos[8];
} ref_inputs;

/* calc_intensities (pn_solver/pn_intensities.jl:325-351).
 * fwd_rec: n_steps+1 records; record k >= 1 holds the coefficients at
 * E_{k-1} - dE/2 and sigma at E_k; record 0 holds sigma at E_initial. */
int ref_forward(const ref_desc* d, const ref_inputs* in, const double* fwd_rec, double* I_out);

/* The Zygote adjoint of calc_intensities (pn_solver/pn_intensities.jl:353-413),
 * up to the material pullbacks.
 * adj_rec[i], i = 1..n_steps: adjoint coefficients at E_i + dE/2, sigma at E_i + dE/2.
 * dens_rec[i], i = 0..n_steps: forward coefficients, gfac and sigma at E_i,
 * where E_i is the reverse-sweep energy of step i (E_n = E_cutoff).
 * Outputs (accumulated from zero): I_out[r], rho_bar[c] ([point][e] per class),
 * ac_bar[p][r], noa_bar[p][e]. */
int ref_forward_adjoint(const ref_desc* d, const ref_inputs* in, const double* fwd_rec, const double* adj_rec,
                        const double* dens_rec, const double* I_bar, double* I_out, double* const rho_bar[8],
                        double* ac_bar, double* noa_bar);

/* Test infrastructure: plants a known mistake in the adjoint so that a test
 * can show that its check detects it.  1: the adjoint boundary blocks with the
 * wrong sign (+L^T A); 2: the density pullback without its C u term; 3: the
 * adjoint source with the wrong sign.  0 restores the faithful restatement. */
void ref_set_plant(int k);

/* One energy step (step_pde!, starmap/src/solver.jl:171-196) on a caller-owned
 * state u[c] ([point][moment] per class, ghost cells included) with the
 * coefficients of record `rec`.  adjoint != 0 uses the adjoint problem
 * (negated A, adjoint boundary blocks, zero boundary data) and the eee m=0
 * source q_eee (NULL for none). */
int ref_step(const ref_desc* d, int adjoint, const double* rec, const double* const rho[8], const double* const mms[8],
             const double* const ppos[8], const double* q_eee, double* const u[8]);

/* ------------------------------------------------------------- material
 * MassFraction material over a box parametrization whose slices are
 * range(lo[a], hi[a], length = n[a] + 1), with edges[a] its n[a] + 1 edge
 * coordinates (pnx_ref_material.c).  Points are [point][3]; densities,
 * gradients [point][element]; vals [box][element] with boxes counted x
 * fastest.  n[a] = 0 on every axis: the embedding everywhere. */
typedef struct {
  int n[3];
  double lo[3], hi[3];
  const double* edges[3];
  int n_e;
  const double* vals;
  const double* embedding;
  const double* rho_pure;
} ref_boxes;

/* The box of a point (-1: the embedding) in exact arithmetic, as the oracle's
 * densities use it, and as the reference's floor((x - lo) / step) computes it
 * in floating point; they differ only for points on interior edges. */
long ref_box_index(const ref_boxes* B, const double* x);
long ref_box_index_floor(const ref_boxes* B, const double* x);
int ref_component_densities(const ref_boxes* B, const double* pts, long n_pts, double* rho);
int ref_component_densities_pullback(const ref_boxes* B, const double* pts, long n_pts, const double* rho_bar,
                                     double* vals_bar, double* emb_bar);
int ref_absorption(const ref_boxes* B, const double* pts, long n_pts, const double* mac, int n_lines,
                   const double* x_d, double x_b0, double* ac);
int ref_absorption_pullback(const ref_boxes* B, const double* pts, long n_pts, const double* mac, int n_lines,
                            const double* x_d, double x_b0, const double* ac, const double* ac_bar, double* vals_bar,
                            double* emb_bar);
int ref_number_of_atoms(const ref_boxes* B, const double* pts, long n_pts, const double* atom_mass, double* noa);
int ref_number_of_atoms_pullback(const ref_boxes* B, const double* pts, long n_pts, const double* atom_mass,
                                 const double* noa_bar, double* vals_bar, double* emb_bar);

/* -------------------------------------------------------------- records
 * Per-energy coefficients (pnx_ref_records.c) from the element tables: the
 * stopping power sp[e] at the knots sp_E, the transport coefficients tc[e]
 * ([n_tc][tc_stride], degree l in column l) at the energies tc_E, the K edge
 * of each line, the beam energy distribution. */
typedef struct {
  int n_knots;
  const double* sp_E;
  const double* sp[8];
  int n_tc, tc_stride;
  const double* tc_E;
  const double* tc[8];
  double line_edge[8];
  double beam_E_mean, beam_E_sd;
} ref_tables;

/* fwd, adj, dens: (n_steps + 1) records each, as the solver takes them. */
int ref_records(const ref_desc* d, const ref_tables* T, double* fwd, double* adj, double* dens);
/* One record with coefficients (and beam bin) at Ec and cross sections at Es. */
int ref_record_at(const ref_desc* d, const ref_tables* T, double Ec, double Es, int adjoint, double* rec);

#ifdef __cplusplus
}
#endif

#endif
