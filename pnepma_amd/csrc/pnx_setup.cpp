// Host set-up of one problem in C++: the per-energy records, the beam moments
// and the beam position profiles (pnx_problem_records, pnx_beam_moments,
// pnx_beam_profiles in include/pnx.h).
//
// Records follow the energy bookkeeping of the reference march and its
// pullback (pn_solver/pn_intensities.jl:10-30, 325-413), with the
// coefficients of update_coefficients! (pn_solver/pn_problem.jl:157-190), the
// beam energy bin of update_boundary_condition! (pn_problem.jl:205-211) and
// the emission cross section of emissioncrosssection
// (pn_solver/x_ray_intensities.jl:32-37); the stopping power is the linear
// interpolant of the caller's table and its slope
// (pn_solver/stopping_power.jl:37-44), the transport coefficients are
// interpolated linearly in log E from the caller's table.
#include <cmath>
#include <string>
#include <vector>

#include "pnx_host.hpp"

namespace pnx {
namespace {

constexpr double kPi = 3.14159265358979323846;

int arg_error(const std::string& msg) { return set_error(PNX_ERR_ARG, msg); }

// Segment i of an ascending knot table with knots[i] <= E < knots[i + 1]
// (the last segment for E at the last knot); -1 outside the table.
int segment(const double* knots, int n, double E) {
  if (!(E >= knots[0] && E <= knots[n - 1])) return -1;
  int lo = 0, hi = n;  // first knot > E, minus one
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (knots[mid] <= E) lo = mid; else hi = mid;
  }
  return std::min(lo, n - 2);
}

struct Coef {
  const pnx_physics* ph;
  int ne, N;
  std::vector<double> log_tc_E;

  // s[e] and c[e][l] at energy E (forward or adjoint coefficients); false if
  // E is outside the stopping-power table.
  bool at(double E, bool adjoint, double* s, double* c) const {
    const int nk = ph->n_knots;
    const int i = segment(ph->sp_E, nk, E);
    if (i < 0) return false;
    const int nt = ph->n_tc;
    const double x = std::log(std::min(std::max(E, ph->tc_E[0]), ph->tc_E[nt - 1]));
    int j = 0;
    {
      int lo = 0, hi = nt;
      while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (log_tc_E[mid] <= x) lo = mid; else hi = mid;
      }
      j = std::min(lo, nt - 2);
    }
    const double tj = (x - log_tc_E[j]) / (log_tc_E[j + 1] - log_tc_E[j]);
    for (int e = 0; e < ne; ++e) {
      const double* sp = ph->sp[e];
      const double x0 = ph->sp_E[i], x1 = ph->sp_E[i + 1];
      const double t = (E - x0) / (x1 - x0);
      s[e] = sp[i] + t * (sp[i + 1] - sp[i]);
      const double slope = (sp[i + 1] - sp[i]) / (x1 - x0);
      const double* tc0 = ph->tc[e] + (size_t)j * ph->tc_stride;
      const double* tc1 = tc0 + ph->tc_stride;
      const double t0 = (1.0 - tj) * tc0[0] + tj * tc1[0];
      for (int l = 0; l <= N; ++l) {
        const double tl = (1.0 - tj) * tc0[l] + tj * tc1[l];
        c[e * (N + 1) + l] = adjoint ? 0.0 - (tl - t0) : -slope - (tl - t0);
      }
    }
    return true;
  }
};

double normal_cdf(double x, double mu, double sd) { return 0.5 * std::erfc(-(x - mu) / (sd * std::sqrt(2.0))); }

// K-shell ionisation cross section stand-in: 6.51e-14 n_K b_K ln(U) / (U E_K^2),
// n_K = 2, b_K = 0.9, U = E / E_K, zero at or below the edge.
double cross_section(double E, double edge) {
  const double U = E / edge;
  if (!(U > 1.0)) return 0.0;
  return 6.51e-14 * 2.0 * 0.9 * std::log(U) / (U * (edge * edge));
}

}  // namespace
}  // namespace pnx

using namespace pnx;

extern "C" {

int pnx_problem_records(const pnx_desc* d, const pnx_physics* ph, double* fwd, double* adj, double* dens) {
  if (!d || !ph || !fwd) return arg_error("null argument");
  Geometry g;
  std::string err;
  if (int rc = make_geometry(*d, g, err)) return set_error(rc, err);
  const int ne = d->n_elements, nl = d->n_lines, N = d->N, n = d->n_steps;
  if (ph->n_knots < 2 || !ph->sp_E || ph->n_tc < 2 || !ph->tc_E || ph->tc_stride < N + 1)
    return arg_error("stopping-power or transport tables missing or too short");
  for (int e = 0; e < ne; ++e)
    if (!ph->sp[e] || !ph->tc[e]) return arg_error("tables missing for an element");
  if (!(ph->beam_E_sd > 0.0)) return arg_error("beam energy spread must be positive");
  Coef co{ph, ne, N, {}};
  co.log_tc_E.resize(ph->n_tc);
  for (int t = 0; t < ph->n_tc; ++t) co.log_tc_E[t] = std::log(ph->tc_E[t]);
  const int L = g.rec_len, sig = g.sigma_off;
  const double dE = (d->E_initial - d->E_cutoff) / n;
  auto fill = [&](double* rec, double Ec, double Es, bool adjoint) {
    if (!co.at(Ec, adjoint, rec, rec + ne)) return false;
    rec[sig - 1] = adjoint ? 0.0
                           : (normal_cdf(Ec + dE, ph->beam_E_mean, ph->beam_E_sd) -
                              normal_cdf(Ec, ph->beam_E_mean, ph->beam_E_sd)) / dE;
    for (int r = 0; r < nl; ++r) rec[sig + r] = cross_section(Es, ph->line_edge[r]);
    return true;
  };
  // forward: record 0 at E_initial; record k with the coefficients at
  // E_{k-1} - dE/2 and the cross sections at E_k (pn_intensities.jl:341-349)
  double E = d->E_initial;
  bool ok = fill(fwd, E, E, false);
  for (int k = 1; ok && k <= n; ++k) {
    double t = d->E_initial - E;  // E_to_t (forward)
    const double Ec = E - dE / 2;
    t = t + dE;
    E = d->E_initial - t;  // t_to_E (forward)
    ok = fill(fwd + (size_t)k * L, Ec, E, false);
  }
  if (ok && adj) {
    if (!dens) return arg_error("adj and dens must be given together");
    // reverse sweep energies E^_i, i = n..0 (pn_intensities.jl:391-401)
    std::vector<double> Er(n + 1);
    double Eb = d->E_cutoff;
    for (int i = n; i >= 1; --i) {
      Er[i] = Eb;
      double t = Eb - d->E_cutoff;  // E_to_t (adjoint)
      t = t + dE;
      Eb = d->E_cutoff + t;  // t_to_E (adjoint)
    }
    Er[0] = Eb;
    for (int i = 0; ok && i <= n; ++i) {
      ok = fill(adj + (size_t)i * L, Er[i] + dE / 2, Er[i] + dE / 2, true) &&
           fill(dens + (size_t)i * L, Er[i], Er[i], false);
    }
  }
  if (!ok) return arg_error("an energy of the march lies outside the stopping-power table");
  return PNX_OK;
}

int pnx_beam_moments(int N, const double* direction, double kappa, double* const mms[8]) {
  if (N < 0 || N > kMaxOrder || !direction || !mms || !(kappa > 0.0)) return arg_error("bad argument");
  double dv[3] = {direction[0], direction[1], direction[2]};
  const double nrm = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
  if (!(nrm > 0.0)) return arg_error("beam direction must be nonzero");
  for (double& v : dv) v /= nrm;
  // The harmonics are evaluated from the direction's components rather than
  // from its angles: with x = d_z, P_l^m(x) = sin^m(theta) Q_l^m(x) and
  // sin^m(theta) (cos, sin)(m phi) = (Re, Im) (d_x + i d_y)^m, so a component
  // that is exactly zero makes the moments that vanish by symmetry exact zeros
  // (angles would leave rounding residues such as cos(pi/2)).
  const double x = dv[2];
  // i_l(kappa) / i_0(kappa) by backward recurrence (Miller): f_{l-1} = f_{l+1} + (2l+1)/kappa f_l
  const int top = N + 40 + (int)kappa;
  std::vector<double> f(top + 2, 0.0);
  f[top + 1] = 0.0;
  f[top] = 1.0;
  for (int l = top; l >= 1; --l) {
    f[l - 1] = f[l + 1] + (2.0 * l + 1.0) / kappa * f[l];
    if (f[l - 1] > 1e250) {  // rescale to stay in range
      for (int q = l - 1; q <= top + 1; ++q) f[q] *= 1e-250;
    }
  }
  // Q_l^m (associated Legendre functions without the Condon-Shortley phase,
  // divided by sin^m(theta)): Q_m^m = (2m - 1)!!, then the recurrence in l
  double Q[kMaxOrder + 1][kMaxOrder + 1] = {};
  for (int m = 0; m <= N; ++m) {
    double qmm = 1.0;
    for (int q = 1; q <= m; ++q) qmm *= 2.0 * q - 1.0;
    Q[m][m] = qmm;
    if (m + 1 <= N) Q[m + 1][m] = x * (2.0 * m + 1.0) * qmm;
    for (int l = m + 2; l <= N; ++l) Q[l][m] = ((2.0 * l - 1.0) * x * Q[l - 1][m] - (l + m - 1.0) * Q[l - 2][m]) / (l - m);
  }
  double cm[kMaxOrder + 1], sm[kMaxOrder + 1];  // (d_x + i d_y)^m
  cm[0] = 1.0;
  sm[0] = 0.0;
  for (int m = 1; m <= N; ++m) {
    cm[m] = cm[m - 1] * dv[0] - sm[m - 1] * dv[1];
    sm[m] = sm[m - 1] * dv[0] + cm[m - 1] * dv[1];
  }
  const Tables t = make_tables(N);
  for (int c = 0; c < 8; ++c) {
    if (!odd_in(c, 0) || t.n[c] == 0) continue;
    if (!mms[c]) return arg_error("output missing for an x-odd class");
    for (int i = 0; i < t.n[c]; ++i) {
      const int l = t.l[c][i], k = t.k[c][i], ak = std::abs(k);
      double fac = 1.0;  // (l - |k|)! / (l + |k|)!
      for (int q = l - ak + 1; q <= l + ak; ++q) fac /= q;
      const double Nlk = std::sqrt((2.0 * l + 1.0) / (4.0 * kPi) * fac);
      double Y = Nlk * Q[l][ak];
      if (k > 0) Y *= std::sqrt(2.0) * cm[ak];
      if (k < 0) Y *= std::sqrt(2.0) * sm[ak];
      mms[c][i] = -Y * (f[l] / f[0]);
    }
  }
  return PNX_OK;
}

int pnx_beam_profiles(const pnx_desc* d, const double* beam_position, double sigma, int literal,
                      double* const ppos[8]) {
  if (!d || !beam_position || !ppos || !(sigma > 0.0)) return arg_error("bad argument");
  Geometry g;
  std::string err;
  if (int rc = make_geometry(*d, g, err)) return set_error(rc, err);
  // axis points of the reference's points(eo, gspec) (starmap/src/staggered_grid.jl:67-78)
  auto point = [&](int a, int odd, int i) { return axis_point(*d, a, odd, i); };
  const double norm = std::pow(2.0 * kPi * sigma * sigma, -1.5);
  const double x0 = point(0, 0, 0), xb = beam_position[0] - x0;
  for (int c = 0; c < 8; ++c) {
    if (!odd_in(c, 0) || !ppos[c]) continue;
    const int ny = g.class_n[c][1], nz = g.class_n[c][2];
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j) {
        const double dy = point(1, odd_in(c, 1), j) - beam_position[1];
        const double dz = point(2, odd_in(c, 2), k) - beam_position[2];
        double r2 = dy * dy + dz * dz;
        if (literal) r2 = r2 + xb * xb;
        ppos[c][j + (size_t)ny * k] = norm * std::exp(-r2 / (2.0 * sigma * sigma));
      }
  }
  return PNX_OK;
}

}  // extern "C"
