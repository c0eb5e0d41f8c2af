/* Runs every entry point of the C oracle on a small synthetic problem, built
 * with AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_sanitizers.py: records from synthetic element tables, the
 * forward march, the reference adjoint, single forward and adjoint steps, the
 * material densities, absorption factors and their pullbacks.  Prints
 * intensities, a gradient checksum and "ok"; exits 1 if an entry point fails
 * or a result is not finite. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/pnx_ref.h"

static int finite_all(const double* a, long n) {
  for (long i = 0; i < n; ++i)
    if (!isfinite(a[i])) return 0;
  return 1;
}

int main(void) {
  ref_desc d = {3, {6, 5, 1}, {-0.3e-4, -0.2e-4, 0.0}, {0.0, 0.2e-4, 0.0}, 2, 2, 6, 15000.0, 12000.0};
  const int ne = 2, nl = 2, L = ref_record_len(&d), n = d.n_steps;
  /* element tables: stopping power falling with energy, transport growing with degree */
  double sp_E[5] = {1e3, 5e3, 1e4, 2e4, 5e4}, sp0[5], sp1[5], tc_E[4] = {500.0, 5e3, 2e4, 1e5};
  double tc0[4 * 4], tc1[4 * 4];
  for (int k = 0; k < 5; ++k) {
    sp0[k] = 4e8 / sqrt(sp_E[k]);
    sp1[k] = 3e8 / sqrt(sp_E[k]);
  }
  for (int k = 0; k < 4; ++k)
    for (int l = 0; l < 4; ++l) {
      tc0[k * 4 + l] = 1e4 * l * (l + 1) / tc_E[k];
      tc1[k * 4 + l] = 2e4 * l * (l + 1) / tc_E[k];
    }
  ref_tables T = {5, sp_E, {sp0, sp1}, 4, 4, tc_E, {tc0, tc1}, {7112.0, 8332.8}, 14500.0, 100.0};
  double* fwd = calloc((size_t)(n + 1) * L, sizeof(double));
  double* adj = calloc((size_t)(n + 1) * L, sizeof(double));
  double* dens = calloc((size_t)(n + 1) * L, sizeof(double));
  if (ref_records(&d, &T, fwd, adj, dens)) return 1;
  /* material: 6 x 5 boxes with edges at the odd points */
  double ex[7], ey[6], ez[2] = {-0.5, 0.5}, vals[6 * 5 * 2], emb[2] = {0.5, 0.5}, rho_pure[2] = {7.874, 8.908};
  const double hx = 0.3e-4 / 5, hy = 0.4e-4 / 4;
  for (int i = 0; i < 7; ++i) ex[i] = -0.3e-4 - hx / 2 + i * hx;
  for (int j = 0; j < 6; ++j) ey[j] = -0.2e-4 - hy / 2 + j * hy;
  for (int b = 0; b < 30; ++b) {
    vals[2 * b] = 0.2 + 0.02 * b;
    vals[2 * b + 1] = 1.0 - vals[2 * b];
  }
  ref_boxes B = {{6, 5, 1}, {ex[0], ey[0], ez[0]}, {ex[6], ey[5], ez[1]}, {ex, ey, ez}, 2, vals, emb, rho_pure};
  double* rho[8];
  double* mms[8];
  double* ppos[8];
  for (int c = 0; c < 8; ++c) {
    const int np = ref_class_points(&d, c);
    rho[c] = calloc((size_t)np * ne + 1, sizeof(double));
    int odd_x = (c >> 2) & 1, odd_y = (c >> 1) & 1;
    int nx = odd_x ? 7 : 6, ny = odd_y ? 6 : 5;
    double* pts = malloc(sizeof(double) * 3 * np);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        double* q = pts + 3 * (i + nx * j);
        q[0] = odd_x ? ex[i] : -0.3e-4 + i * hx;
        q[1] = odd_y ? ey[j] : -0.2e-4 + j * hy;
        q[2] = 0.0;
      }
    ref_component_densities(&B, pts, np, rho[c]);
    free(pts);
    const int nm = ref_class_moments(d.N, c);
    mms[c] = odd_x ? calloc((size_t)nm + 1, sizeof(double)) : NULL;
    ppos[c] = odd_x ? calloc((size_t)ny + 1, sizeof(double)) : NULL;
    for (int m = 0; odd_x && m < nm; ++m) mms[c][m] = -0.3 / (m + 1);
    for (int j = 0; odd_x && j < ny; ++j) ppos[c][j] = exp(-(double)(j - 2) * (j - 2));
  }
  const int np0 = ref_class_points(&d, 0);
  double* eee = malloc(sizeof(double) * 3 * np0);
  for (int j = 0; j < 5; ++j)
    for (int i = 0; i < 6; ++i) {
      double* q = eee + 3 * (i + 6 * j);
      q[0] = -0.3e-4 + i * hx;
      q[1] = -0.2e-4 + j * hy;
      q[2] = 0.0;
    }
  double mac[4] = {70.0, 85.0, 380.0, 60.0}, am[2] = {9.27e-23, 9.75e-23}, det[3] = {0.64, 0.77, 0.0};
  double* ac = calloc((size_t)np0 * nl, sizeof(double));
  double* noa = calloc((size_t)np0 * ne, sizeof(double));
  if (ref_absorption(&B, eee, np0, mac, nl, det, 0.0, ac) || ref_number_of_atoms(&B, eee, np0, am, noa)) return 1;
  int el[2] = {0, 1};
  ref_inputs in;
  for (int c = 0; c < 8; ++c) {
    in.rho[c] = rho[c];
    in.mms[c] = mms[c];
    in.ppos[c] = ppos[c];
  }
  in.ac = ac;
  in.noa = noa;
  in.el_idx = el;
  double I[2], I2[2], ib[2] = {1.0, -0.5};
  double* rb[8];
  for (int c = 0; c < 8; ++c) rb[c] = calloc((size_t)ref_class_points(&d, c) * ne + 1, sizeof(double));
  double* acb = calloc((size_t)np0 * nl, sizeof(double));
  double* noab = calloc((size_t)np0 * ne, sizeof(double));
  if (ref_forward(&d, &in, fwd, I) || ref_forward_adjoint(&d, &in, fwd, adj, dens, ib, I2, rb, acb, noab)) return 1;
  /* single steps on the final-like state */
  double* u[8];
  for (int c = 0; c < 8; ++c) {
    const long m = (long)ref_class_points(&d, c) * (ref_class_moments(d.N, c) + 1);
    u[c] = calloc((size_t)m + 1, sizeof(double));
    for (long k = 0; k < m; ++k) u[c][k] = sin(0.1 * k + c);
  }
  double* q = calloc((size_t)np0, sizeof(double));
  for (int p = 0; p < np0; ++p) q[p] = cos(p);
  if (ref_step(&d, 0, fwd + L, (const double* const*)rho, (const double* const*)mms, (const double* const*)ppos,
               NULL, u) ||
      ref_step(&d, 1, adj + 2 * L, (const double* const*)rho, (const double* const*)mms,
               (const double* const*)ppos, q, u))
    return 1;
  /* material pullbacks */
  double vb[60] = {0}, eb[2] = {0};
  if (ref_component_densities_pullback(&B, eee, np0, noab, vb, eb) ||
      ref_absorption_pullback(&B, eee, np0, mac, nl, det, 0.0, ac, acb, vb, eb) ||
      ref_number_of_atoms_pullback(&B, eee, np0, am, noab, vb, eb))
    return 1;
  double sum = 0.0;
  for (int k = 0; k < 60; ++k) sum += vb[k];
  int ok = finite_all(I, 2) && finite_all(I2, 2) && finite_all(vb, 60) && fabs(I[0] - I2[0]) <= 1e-12 * fabs(I[0]);
  for (int c = 0; c < 8; ++c) ok = ok && finite_all(rb[c], (long)ref_class_points(&d, c) * ne);
  printf("I %.6e %.6e gradient checksum %.6e\n", I[0], I[1], sum);
  for (int c = 0; c < 8; ++c) {
    free(rho[c]);
    free(mms[c]);
    free(ppos[c]);
    free(rb[c]);
    free(u[c]);
  }
  free(fwd);
  free(adj);
  free(dens);
  free(eee);
  free(ac);
  free(noa);
  free(acb);
  free(noab);
  free(q);
  if (!ok) return 1;
  printf("ok\n");
  return 0;
}
