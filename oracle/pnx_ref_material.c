/* CPU oracle of the material layer at the boundary of the hot path -- TEST
 * INFRASTRUCTURE.  A plain-C restatement of the reference's MassFraction
 * material over a box parametrization, its number-of-atoms and absorption
 * factors, and their pullbacks.  Only the tests load it.
 *
 * Reference files followed (paths relative to the reference root):
 *   pn_solver/pn_parametrizations/pn_material.jl:199-264      (MassFraction)
 *   pn_solver/pn_parametrizations/pn_parametrization_boxed.jl:1-175 (Boxed)
 *   pn_solver/pn_parametrizations/biliner_interp.jl:5-15,45-47  (box lookup)
 *   pn_solver/x_ray_intensities.jl:43-153    (number of atoms, absorption)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pnx_ref.h"

/* find_index_smaller_and_partition on a range slice vector
 * (biliner_interp.jl:5-15): the slices are range(lo, hi, length = n + 1), so
 * i = floor((x - lo) / step) counted from the first box, and the partition
 * (d - i step) / step is returned through part. */
static long slice_index(double lo, double hi, int n, double x, double* part) {
  if (n < 1) {
    *part = 0.0;
    return 0;
  }
  const double step = (hi - lo) / n;
  const double d = x - lo;
  const long i = (long)floor(d / step);
  *part = (d - (double)i * step) / step;
  return i;
}

/* The box of a point as the reference computes it in floating point
 * (parametrization!, pn_parametrization_boxed.jl:137-156): per axis the slice
 * index; a point on the upper outer edge (index n with a partition of
 * magnitude below 1e-20, isapproxzero) belongs to the last box; any index
 * outside the boxes selects the embedding (-1).  Boxes are counted x fastest.
 * A point lying exactly on an interior edge gets either neighbour, depending
 * on the rounding of (x - lo) / step. */
static long box_of_floor(const ref_boxes* B, const double* x) {
  long idx[3];
  for (int a = 0; a < 3; ++a) {
    double part;
    long i = slice_index(B->lo[a], B->hi[a], B->n[a], x[a], &part);
    if (i == B->n[a] && fabs(part) < 1e-20) i -= 1;
    if (i < 0 || i > B->n[a] - 1) return -1;
    idx[a] = i;
  }
  return idx[0] + (long)B->n[0] * (idx[1] + (long)B->n[1] * idx[2]);
}

long ref_box_index_floor(const ref_boxes* B, const double* x) { return box_of_floor(B, x); }

/* The same lookup in exact arithmetic, which the oracle uses: per axis the box
 * k whose lower edge edges[k] <= x < edges[k + 1] (the floor of
 * (x - lo) / step when the edges are lo + k step exactly), the upper outer edge
 * in the last box, the embedding outside [edges[0], edges[n]].  It differs
 * from box_of_floor only for points on interior edges. */
static long box_of(const ref_boxes* B, const double* x) {
  long idx[3];
  for (int a = 0; a < 3; ++a) {
    const int n = B->n[a];
    if (n < 1) return -1;
    const double* e = B->edges[a];
    if (!(x[a] >= e[0] && x[a] <= e[n])) return -1;
    long k = 0, hi = n;  /* largest k in [0, n) with e[k] <= x */
    while (hi - k > 1) {
      const long mid = (k + hi) / 2;
      if (e[mid] <= x[a]) k = mid; else hi = mid;
    }
    idx[a] = k;
  }
  return idx[0] + (long)B->n[0] * (idx[1] + (long)B->n[1] * idx[2]);
}

long ref_box_index(const ref_boxes* B, const double* x) { return box_of(B, x); }

static const double* fractions_at(const ref_boxes* B, const double* x) {
  const long b = box_of(B, x);
  return b < 0 ? B->embedding : B->vals + (size_t)b * B->n_e;
}

/* component_densities! of MassFractionPNMaterial (pn_material.jl:236-240):
 * rho = w / R with R = sum(w ./ rho_pure). */
int ref_component_densities(const ref_boxes* B, const double* pts, long n_pts, double* rho) {
  const int ne = B->n_e;
  for (long p = 0; p < n_pts; ++p) {
    const double* w = fractions_at(B, pts + 3 * p);
    double R = 0.0;
    for (int e = 0; e < ne; ++e) R += w[e] / B->rho_pure[e];
    for (int e = 0; e < ne; ++e) rho[p * ne + e] = w[e] / R;
  }
  return 0;
}

/* component_densities_pullback! (pn_material.jl:249-256) followed by the box
 * parametrization's pullback (pn_parametrization_boxed.jl:158-175):
 * R_ = -dot(w, rho_) / R^2, w_ = rho_ ./ R + R_ ./ rho_pure, added to the
 * box of the point (or to the embedding). */
int ref_component_densities_pullback(const ref_boxes* B, const double* pts, long n_pts, const double* rho_bar,
                                     double* vals_bar, double* emb_bar) {
  const int ne = B->n_e;
  for (long p = 0; p < n_pts; ++p) {
    const long b = box_of(B, pts + 3 * p);
    const double* w = b < 0 ? B->embedding : B->vals + (size_t)b * ne;
    double R = 0.0, dot = 0.0;
    for (int e = 0; e < ne; ++e) {
      R += w[e] / B->rho_pure[e];
      dot += w[e] * rho_bar[p * ne + e];
    }
    const double R_ = -dot / (R * R);
    double* out = b < 0 ? emb_bar : vals_bar + (size_t)b * ne;
    for (int e = 0; e < ne; ++e) out[e] += rho_bar[p * ne + e] / R + R_ / B->rho_pure[e];
  }
  return 0;
}

/* The ray of compute_absorption_coefficient! (x_ray_intensities.jl:100-113)
 * from `origin` towards the detector x_d, cut where it crosses x = x_b[0]:
 * lambda = max((x_b[0] - origin[0]) / dir[0], 0), dx = |dir| lambda / n, and
 * the n sample parameters range(0, lambda, length = n). */
static void ray(const double* origin, const double* x_d, double x_b0, int n, double* dir, double* lam, double* dx) {
  for (int a = 0; a < 3; ++a) dir[a] = x_d[a] - origin[a];
  *lam = fmax((x_b0 - origin[0]) / dir[0], 0.0);
  *dx = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]) * *lam / n;
}

/* compute_absorption_coefficients (x_ray_intensities.jl:100-128) with the
 * reference's 100 samples: ac[p][r] = exp(-dx sum_samples dot(mac[r], rho)).
 * mac is [line][element]. */
int ref_absorption(const ref_boxes* B, const double* pts, long n_pts, const double* mac, int n_lines,
                   const double* x_d, double x_b0, double* ac) {
  const int ne = B->n_e, ns = 100;
  double* rho = malloc(sizeof(double) * ne);
  if (!rho) return -1;
  for (long p = 0; p < n_pts; ++p) {
    const double* o = pts + 3 * p;
    double dir[3], lam, dx;
    ray(o, x_d, x_b0, ns, dir, &lam, &dx);
    double* a = ac + p * n_lines;
    for (int r = 0; r < n_lines; ++r) a[r] = 0.0;
    for (int k = 0; k < ns; ++k) {
      const double l = lam * k / (ns - 1);
      double q[3];
      for (int d = 0; d < 3; ++d) q[d] = o[d] + l * dir[d];
      ref_component_densities(B, q, 1, rho);
      for (int r = 0; r < n_lines; ++r) {
        double s = 0.0;
        for (int e = 0; e < ne; ++e) s += mac[r * ne + e] * rho[e];
        a[r] -= s;
      }
    }
    for (int r = 0; r < n_lines; ++r) a[r] = exp(a[r] * dx);
  }
  free(rho);
  return 0;
}

/* compute_absorption_coefficient_pullback (x_ray_intensities.jl:131-153):
 * ac_adj <- -ac dx ac_adj, then at every sample the density pullback of
 * rho_adj = mac^T ac_adj. */
int ref_absorption_pullback(const ref_boxes* B, const double* pts, long n_pts, const double* mac, int n_lines,
                            const double* x_d, double x_b0, const double* ac, const double* ac_bar, double* vals_bar,
                            double* emb_bar) {
  const int ne = B->n_e, ns = 100;
  double* rho_adj = malloc(sizeof(double) * ne);
  double* a_ = malloc(sizeof(double) * n_lines);
  if (!rho_adj || !a_) {
    free(rho_adj);
    free(a_);
    return -1;
  }
  for (long p = 0; p < n_pts; ++p) {
    const double* o = pts + 3 * p;
    double dir[3], lam, dx;
    ray(o, x_d, x_b0, ns, dir, &lam, &dx);
    for (int r = 0; r < n_lines; ++r) a_[r] = -ac[p * n_lines + r] * dx * ac_bar[p * n_lines + r];
    for (int e = 0; e < ne; ++e) {
      rho_adj[e] = 0.0;
      for (int r = 0; r < n_lines; ++r) rho_adj[e] += mac[r * ne + e] * a_[r];
    }
    for (int k = 0; k < ns; ++k) {
      const double l = lam * k / (ns - 1);
      double q[3];
      for (int d = 0; d < 3; ++d) q[d] = o[d] + l * dir[d];
      ref_component_densities_pullback(B, q, 1, rho_adj, vals_bar, emb_bar);
    }
  }
  free(rho_adj);
  free(a_);
  return 0;
}

/* compute_number_of_atoms (x_ray_intensities.jl:43-55): rho ./ AM at eee
 * points; its pullback (:85-96) is the density pullback of noa_ ./ AM. */
int ref_number_of_atoms(const ref_boxes* B, const double* pts, long n_pts, const double* atom_mass, double* noa) {
  ref_component_densities(B, pts, n_pts, noa);
  for (long p = 0; p < n_pts; ++p)
    for (int e = 0; e < B->n_e; ++e) noa[p * B->n_e + e] /= atom_mass[e];
  return 0;
}

int ref_number_of_atoms_pullback(const ref_boxes* B, const double* pts, long n_pts, const double* atom_mass,
                                 const double* noa_bar, double* vals_bar, double* emb_bar) {
  const int ne = B->n_e;
  double* t = malloc(sizeof(double) * (size_t)(n_pts > 0 ? n_pts : 1) * ne);
  if (!t) return -1;
  for (long p = 0; p < n_pts; ++p)
    for (int e = 0; e < ne; ++e) t[p * ne + e] = noa_bar[p * ne + e] / atom_mass[e];
  ref_component_densities_pullback(B, pts, n_pts, t, vals_bar, emb_bar);
  free(t);
  return 0;
}
