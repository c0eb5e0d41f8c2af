/* CPU oracle of the per-energy coefficients of the march -- TEST
 * INFRASTRUCTURE.  A plain-C restatement of what the reference evaluates at
 * every energy of calc_intensities and its pullback, laid out as the energy
 * records the solver consumes (s[e], c[e][l], gfac, sigma[r]).  Only the tests
 * load it.
 *
 * Reference files followed (paths relative to the reference root):
 *   pn_solver/pn_intensities.jl:10-30, 325-413 (energies of the march and of
 *     the reverse sweep: E - dE/2 forward, E + dE/2 adjoint, E for the
 *     density pullback; cross sections at the post-step E)
 *   pn_solver/pn_problem.jl:157-190 (update_coefficients!, forward and adjoint)
 *   pn_solver/pn_problem.jl:205-211 (update_boundary_condition!: the beam
 *     energy bin (cdf(E + dE) - cdf(E)) / dE)
 *   pn_solver/stopping_power.jl:37-44 (linear interpolant and its derivative)
 *   pn_solver/x_ray_intensities.jl:32-37 (emissioncrosssection, without the
 *     fluorescence yield; the cross section itself is the package's stand-in)
 */
#include <math.h>
#include <stdlib.h>

#include "pnx_ref.h"

/* stopping_power_func: the linear interpolant between knots and its slope;
 * a knot belongs to the segment on its right, the last knot to the last
 * segment.  Returns -1 outside the table. */
static int sp_segment(const ref_tables* T, double E) {
  const int n = T->n_knots;
  if (!(E >= T->sp_E[0] && E <= T->sp_E[n - 1])) return -1;
  int i = 0;
  while (i < n - 2 && T->sp_E[i + 1] <= E) ++i;
  return i;
}

static double stopping_power(const ref_tables* T, int e, int i, double E) {
  const double x0 = T->sp_E[i], x1 = T->sp_E[i + 1], y0 = T->sp[e][i], y1 = T->sp[e][i + 1];
  return y0 + (E - x0) / (x1 - x0) * (y1 - y0);
}

static double stopping_power_derivative(const ref_tables* T, int e, int i) {
  return (T->sp[e][i + 1] - T->sp[e][i]) / (T->sp_E[i + 1] - T->sp_E[i]);
}

/* Transport coefficients: linear in log E between the table's energies,
 * clamped to the table.  tc_at finds the interval and weight of E once; the
 * coefficient of element e and degree l is then tc_value. */
static void tc_at(const ref_tables* T, double E, int* i_out, double* t_out) {
  const int n = T->n_tc;
  const double Ec = E < T->tc_E[0] ? T->tc_E[0] : (E > T->tc_E[n - 1] ? T->tc_E[n - 1] : E);
  const double x = log(Ec);
  int lo = 0, hi = n - 1; /* largest i <= n - 2 with log(tc_E[i]) <= x */
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (log(T->tc_E[mid]) <= x) lo = mid; else hi = mid;
  }
  const double a = log(T->tc_E[lo]), b = log(T->tc_E[lo + 1]);
  *i_out = lo;
  *t_out = (x - a) / (b - a);
}

static double tc_value(const ref_tables* T, int e, int l, int i, double t) {
  const double* r0 = T->tc[e] + (size_t)i * T->tc_stride;
  const double* r1 = r0 + T->tc_stride;
  return (1.0 - t) * r0[l] + t * r1[l];
}

static double normal_cdf(double x, double mu, double sd) { return 0.5 * erfc((mu - x) / (sd * sqrt(2.0))); }

static double cross_section(double E, double edge) {
  const double U = E / edge;
  return U > 1.0 ? 6.51e-14 * 2.0 * 0.9 * log(U) / (U * (edge * edge)) : 0.0;
}

/* One record: update_coefficients! at Ec (adjoint: without the stopping-power
 * derivative), the beam bin at Ec (zero in the adjoint problem, which has no
 * source), the cross sections at Es. */
static int record(const ref_desc* d, const ref_tables* T, double Ec, double Es, double dE, int adjoint, double* rec) {
  const int ne = d->n_e, N = d->N;
  const int i = sp_segment(T, Ec);
  if (i < 0) return -1;
  int ti;
  double tt;
  tc_at(T, Ec, &ti, &tt);
  for (int e = 0; e < ne; ++e) {
    rec[e] = stopping_power(T, e, i, Ec);
    const double tc0 = tc_value(T, e, 0, ti, tt);
    for (int l = 0; l <= N; ++l) {
      const double dtc = tc_value(T, e, l, ti, tt) - tc0;
      rec[ne + e * (N + 1) + l] = adjoint ? -dtc : -stopping_power_derivative(T, e, i) - dtc;
    }
  }
  double* g = rec + ne + ne * (N + 1);
  *g = adjoint ? 0.0 : (normal_cdf(Ec + dE, T->beam_E_mean, T->beam_E_sd) - normal_cdf(Ec, T->beam_E_mean, T->beam_E_sd)) / dE;
  for (int r = 0; r < d->n_lines; ++r) g[1 + r] = cross_section(Es, T->line_edge[r]);
  return 0;
}

int ref_record_at(const ref_desc* d, const ref_tables* T, double Ec, double Es, int adjoint, double* rec) {
  return record(d, T, Ec, Es, (d->E_initial - d->E_cutoff) / d->n_steps, adjoint, rec);
}

int ref_records(const ref_desc* d, const ref_tables* T, double* fwd, double* adj, double* dens) {
  const int L = ref_record_len(d), n = d->n_steps;
  const double dE = (d->E_initial - d->E_cutoff) / n;
  /* calc_intensities: E starts at E_initial; each step converts E to the
   * forward time t = E_initial - E, evaluates the problem at E - dE/2,
   * advances t by dE and converts back (pn_intensities.jl:10-14,341-349) */
  double E = d->E_initial;
  if (record(d, T, E, E, dE, 0, fwd)) return -1;
  for (int k = 1; k <= n; ++k) {
    const double Ec = E - dE / 2;
    const double t = (d->E_initial - E) + dE;
    E = d->E_initial - t;
    if (record(d, T, Ec, E, dE, 0, fwd + (size_t)k * L)) return -1;
  }
  if (!adj) return 0;
  /* the pullback: E starts at E_cutoff; step i uses the adjoint problem and
   * sigma at E + dE/2 and the forward coefficients at E, then advances the
   * adjoint time t = E - E_cutoff by dE (pn_intensities.jl:391-404) */
  E = d->E_cutoff;
  for (int i = n; i >= 1; --i) {
    if (record(d, T, E + dE / 2, E + dE / 2, dE, 1, adj + (size_t)i * L) ||
        record(d, T, E, E, dE, 0, dens + (size_t)i * L))
      return -1;
    const double t = (E - d->E_cutoff) + dE;
    E = d->E_cutoff + t;
  }
  if (record(d, T, E + dE / 2, E + dE / 2, dE, 1, adj) || record(d, T, E, E, dE, 0, dens)) return -1;
  return 0;
}
