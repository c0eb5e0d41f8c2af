/* CPU oracle of the PN-CSD forward solve and its adjoint -- TEST INFRASTRUCTURE.
 * See pnx_ref.h for the contract.  Every function names the reference lines it
 * restates (paths relative to the reference root). */
#include "pnx_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef long double ld;

#define MAXC 16

/* ------------------------------------------------------------------ moments */
/* pn_solver/moment_specification.jl:26-36 */
static int x_odd(int k) { return ((k < 0 && k % 2 != 0) || (k >= 0 && k % 2 == 0)) ? 0 : 1; }
static int y_odd(int k) { return k >= 0 ? 0 : 1; }
static int z_odd(int l, int k) { return ((l + k) % 2 == 0) ? 0 : 1; }
static int class_of(int l, int k) { return 4 * x_odd(k) + 2 * y_odd(k) + z_odd(l, k); }
/* starmap/src/even_odd_classification.jl:28-88 */
static int odd_in(int c, int d) { return (c >> (2 - d)) & 1; }
static int sw(int c, int d) { return c ^ (4 >> d); }
static int parity_of(int c) { return (odd_in(c, 0) + odd_in(c, 1) + odd_in(c, 2)) & 1; }

typedef struct {
  int n;
  int l[MAXC], k[MAXC];
} cls_t;

/* make_moments (pn_solver/pn_variable.jl:12): all_moments order filtered by class */
static void make_classes(int N, cls_t cls[8]) {
  for (int c = 0; c < 8; ++c) cls[c].n = 0;
  for (int l = 0; l <= N; ++l)
    for (int k = -l; k <= l; ++k) {
      cls_t* t = &cls[class_of(l, k)];
      t->l[t->n] = l;
      t->k[t->n] = k;
      t->n++;
    }
}

int ref_class_moments(int N, int c) {
  cls_t cls[8];
  if (N < 0 || N > 9 || c < 0 || c > 7) return -1;
  make_classes(N, cls);
  return cls[c].n;
}

/* ------------------------------------------------------ advection closed form */
/* pn_solver/pn_matrices.jl:24-144 */
static int Theta(int k) { return k < 0 ? -1 : 1; }
static int kplus(int k) { return k == 0 ? 1 : k + (k > 0 ? 1 : -1); }
static int kminus(int k) { return k == 0 ? -1 : k - (k > 0 ? 1 : -1); }
static double rt(long a, long b) { return sqrt((double)a / (double)b); }
static double ca(int k, int l) { return rt((long)(l - k + 1) * (l + k + 1), (long)(2 * l + 3) * (2 * l + 1)); }
static double cb(int k, int l) { return rt((long)(l - k) * (l + k), (long)(2 * l + 1) * (2 * l - 1)); }
static double cc(int k, int l) {
  double v = rt((long)(l + k + 1) * (l + k + 2), (long)(2 * l + 3) * (2 * l + 1));
  if (k < 0) return 0.0;
  return k > 0 ? v : v * sqrt(2.0);
}
static double cd(int k, int l) {
  double v = rt((long)(l - k) * (l - k - 1), (long)(2 * l + 1) * (2 * l - 1));
  if (k < 0) return 0.0;
  return k > 0 ? v : v * sqrt(2.0);
}
static double ce(int k, int l) {
  double v = rt((long)(l - k + 1) * (l - k + 2), (long)(2 * l + 3) * (2 * l + 1));
  return k == 1 ? v * sqrt(2.0) : v;
}
static double cf(int k, int l) {
  double v = rt((long)(l + k) * (l + k - 1), (long)(2 * l + 1) * (2 * l - 1));
  return k == 1 ? v * sqrt(2.0) : v;
}
static double A_minus(int l, int k, int kp, int d) {
  int ak = abs(k);
  if (d == 0) {
    if (kp == kminus(k) && k != -1) return 0.5 * cc(ak - 1, l - 1);
    if (kp == kplus(k)) return -0.5 * ce(ak + 1, l - 1);
    return 0.0;
  }
  if (d == 1) {
    if (kp == -kminus(k) && k != 1) return (-Theta(k) / 2.0) * cc(ak - 1, l - 1);
    if (kp == -kplus(k)) return (-Theta(k) / 2.0) * ce(ak + 1, l - 1);
    return 0.0;
  }
  return kp == k ? ca(k, l - 1) : 0.0;
}
static double A_plus(int l, int k, int kp, int d) {
  int ak = abs(k);
  if (d == 0) {
    if (kp == kminus(k) && k != -1) return -0.5 * cd(ak - 1, l + 1);
    if (kp == kplus(k)) return 0.5 * cf(ak + 1, l + 1);
    return 0.0;
  }
  if (d == 1) {
    if (kp == -kminus(k) && k != 1) return (Theta(k) / 2.0) * cd(ak - 1, l + 1);
    if (kp == -kplus(k)) return (Theta(k) / 2.0) * cf(ak + 1, l + 1);
    return 0.0;
  }
  return kp == k ? cb(k, l + 1) : 0.0;
}
static double get_coefficient(int l, int k, int lp, int kp, int d) {
  if (l == lp - 1) return A_plus(l, k, kp, d);
  if (l == lp + 1) return A_minus(l, k, kp, d);
  return 0.0;
}

/* ------------------------------------------------------- boundary closed form */
/* pn_solver/pn_matrices.jl:191-275 */
static ld lfact(int n) {
  ld r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
static ld ldfact(int n) {
  if (n < 0) return 1;
  ld r = 1;
  for (int i = n; i > 1; i -= 2) r *= i;
  return r;
}
static ld lbinom(int n, int k) {
  if (k < 0 || k > n) return 0;
  ld r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}
static int fl2(int a) { return (int)floorl((ld)a / 2.0L); }
static ld sgn1(int e) { return (e % 2 == 0) ? 1.0L : -1.0L; }
static const ld PI_L = 3.14159265358979323846264338327950288L;
static ld Cn(int l, int k) {
  ld v = sqrtl((2.0L * l + 1.0L) / (2.0L * PI_L) * lfact(l - abs(k)) / lfact(l + abs(k)));
  return (abs(k) % 2 == 0) ? v : -v;
}
static double get_boundary_coefficient(int l, int k, int lp, int kp, int d) {
  int ak = abs(k), akp = abs(kp);
  if (d == 2) {
    if (k != kp) return 0.0;
    ld s = 0;
    for (int j = 0; j <= fl2(l - ak); ++j)
      for (int q = 0; q <= fl2(lp - akp); ++q)
        s += sgn1(q + j) / powl(2.0L, q + j) / (lfact(q) * lfact(j)) * ldfact(2 * (l - j) - 1) / lfact(l - ak - 2 * j) *
             ldfact(2 * (lp - q) - 1) / lfact(lp - akp - 2 * q) * tgammal((l + lp) / 2.0L - q - j - ak) /
             tgammal((l + lp) / 2.0L + 1 - q - j);
    return (double)(s * PI_L * Cn(l, ak) * Cn(lp, akp) * lfact(ak));
  }
  ld sjq = 0;
  for (int j = 0; j <= fl2(l - ak); ++j)
    for (int q = 0; q <= fl2(lp - akp); ++q) {
      ld sis = 0;
      for (int i = 0; i <= fl2(ak - 1); ++i)
        for (int s = 0; s <= fl2(akp - 1); ++s) {
          if (d == 0) {
            if (k >= 0 && kp >= 0 && (k % 2) != 0 && (kp % 2) != 0)
              sis += sgn1(s + i) * lbinom(k, 2 * i) * lbinom(kp, 2 * s) * tgammal(0.5L + s + i) *
                     tgammal((k + kp) / 2.0L - (s + i));
            else if (k < 0 && kp < 0 && (k % 2) == 0 && (kp % 2) == 0)
              sis += sgn1(s + i) * lbinom(ak, 2 * i + 1) * lbinom(akp, 2 * s + 1) * tgammal(1.5L + s + i) *
                     tgammal((ak + akp) / 2.0L - (s + i) - 1);
          } else {
            sis += sgn1(s + i) * lfact(s + i) * lbinom(ak, 2 * i + 1) * lbinom(akp, 2 * s + 1) *
                   tgammal((ak + akp) / 2.0L - s - i - 0.5L);
          }
        }
      sjq += sis * sgn1(q + j) / (powl(2.0L, j + q) * lfact(j) * lfact(q)) * ldfact(2 * (l - j) - 1) /
             lfact(l - ak - 2 * j) * ldfact(2 * (lp - q) - 1) / lfact(lp - akp - 2 * q) *
             tgammal((l + lp + 1 - 2 * (q + j) - (ak + akp)) / 2.0L) / tgammal((l + lp) / 2.0L + 1 - (q + j));
    }
  return (double)(Cn(l, ak) * Cn(lp, akp) * (1 + sgn1(l + lp)) * sgn1(ak + akp) * sjq);
}

int ref_advection_dense(int N, int c, int d, double* out) {
  cls_t cls[8];
  if (N < 0 || N > 9 || c < 0 || c > 7 || d < 0 || d > 2) return -1;
  make_classes(N, cls);
  int f = sw(c, d);
  for (int i = 0; i < cls[c].n; ++i)
    for (int j = 0; j < cls[f].n; ++j)
      out[i * cls[f].n + j] = get_coefficient(cls[c].l[i], cls[c].k[i], cls[f].l[j], cls[f].k[j], d);
  return 0;
}

/* Half-range coupling L over the moments of class c (pn_matrices.jl:367-370). */
int ref_boundary_L(int N, int c, int d, double* out) {
  cls_t cls[8];
  if (N < 0 || N > 9 || c < 0 || c > 7 || d < 0 || d > 2) return -1;
  make_classes(N, cls);
  for (int i = 0; i < cls[c].n; ++i)
    for (int j = 0; j < cls[c].n; ++j)
      out[i * cls[c].n + j] = get_boundary_coefficient(cls[c].l[i], cls[c].k[i], cls[c].l[j], cls[c].k[j], d);
  return 0;
}

/* make_boundary_matrix / _adjoint (pn_solver/pn_matrices.jl:277-293) */
int ref_boundary_dense(int N, int c, int d, int adjoint, double* out) {
  cls_t cls[8];
  if (N < 0 || N > 9 || c < 0 || c > 7 || d < 0 || d > 2 || !odd_in(c, d)) return -1;
  make_classes(N, cls);
  int f = sw(c, d), nc = cls[c].n, nf = cls[f].n;
  double* L = malloc(sizeof(double) * nc * nc);
  double* A = malloc(sizeof(double) * nc * nf);
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) L[i * nc + j] = get_boundary_coefficient(cls[c].l[i], cls[c].k[i], cls[c].l[j], cls[c].k[j], d);
  ref_advection_dense(N, c, d, A);
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nf; ++j) {
      double s = 0.0;
      for (int m = 0; m < nc; ++m) s += (adjoint ? L[m * nc + i] : L[i * nc + m]) * A[m * nf + j];
      out[i * nf + j] = adjoint ? -s : s;
    }
  free(L);
  free(A);
  return 0;
}

/* ------------------------------------------------------------------ problem */
typedef struct {
  int ncol;
  int* colptr; /* ncol + 1 */
  int* rowval;
  double* nzval;
} csc_t;

typedef struct {
  ref_desc d;
  cls_t cls[8];
  int n[8][3];   /* class sizes per axis */
  int np[8];     /* points per class */
  int single[3]; /* singleton axes */
  double h[3];
  double dvol;
  int adjoint;
  csc_t A[8][3];
  double* B[8][3];
  double* gx[8];    /* x boundary data per x-odd class: [face 0/1][j][k][m] (current step) */
  double* gset[8];  /* setup g = mms * ppos on the right face: [j][k][m] */
  double* S[8];     /* capacity, [point][m] */
  double* C[8];     /* homogeneous term, [point][m] */
  double* Q[8];     /* source (adjoint only), [point][m] */
  double s_cur[8];  /* per element, at most 8 elements */
} prob_t;

static int rec_len(const ref_desc* d) { return d->n_e + d->n_e * (d->N + 1) + 1 + d->n_lines; }
int ref_record_len(const ref_desc* d) { return rec_len(d); }

/* staggered_grid.jl:30-60: even axis M points, odd axis M+1 points, singleton 1 */
static void class_sizes(const ref_desc* d, int c, int n[3]) {
  for (int a = 0; a < 3; ++a) n[a] = d->M[a] == 1 ? 1 : (odd_in(c, a) ? d->M[a] + 1 : d->M[a]);
}

int ref_class_points(const ref_desc* d, int c) {
  int n[3];
  class_sizes(d, c, n);
  return n[0] * n[1] * n[2];
}

static void csc_build(csc_t* A, const cls_t* to, const cls_t* from, int d, double sign) {
  /* make_advection_matrix (pn_matrices.jl:146-152): stored entries are the nonzeros */
  A->ncol = from->n;
  A->colptr = malloc(sizeof(int) * (from->n + 1));
  A->rowval = malloc(sizeof(int) * (to->n * from->n + 1));
  A->nzval = malloc(sizeof(double) * (to->n * from->n + 1));
  int nz = 0;
  for (int j = 0; j < from->n; ++j) {
    A->colptr[j] = nz;
    for (int i = 0; i < to->n; ++i) {
      double v = get_coefficient(to->l[i], to->k[i], from->l[j], from->k[j], d);
      if (v != 0.0) {
        A->rowval[nz] = i;
        A->nzval[nz] = sign * v;
        nz++;
      }
    }
  }
  A->colptr[from->n] = nz;
}

/* Planted mistakes, for the tests that show the oracle's pins detect them
 * (0 = the faithful restatement; see ref_set_plant in pnx_ref.h). */
static int g_plant = 0;
void ref_set_plant(int k) { g_plant = k; }

static int prob_init(prob_t* P, const ref_desc* d, int adjoint, const ref_inputs* in) {
  memset(P, 0, sizeof(*P));
  if (d->N < 0 || d->N > 9 || d->n_e < 1 || d->n_e > 8) return -1;
  P->d = *d;
  P->adjoint = adjoint;
  make_classes(d->N, P->cls);
  for (int a = 0; a < 3; ++a) {
    P->single[a] = d->M[a] == 1;
    P->h[a] = P->single[a] ? 1.0 : (d->hi[a] - d->lo[a]) / (d->M[a] - 1);
  }
  P->dvol = P->h[0] * P->h[1] * P->h[2];
  for (int c = 0; c < 8; ++c) {
    class_sizes(d, c, P->n[c]);
    P->np[c] = P->n[c][0] * P->n[c][1] * P->n[c][2];
    int nm = P->cls[c].n;
    P->S[c] = calloc((size_t)P->np[c] * nm + 1, sizeof(double));
    P->C[c] = calloc((size_t)P->np[c] * nm + 1, sizeof(double));
    if (adjoint) P->Q[c] = calloc((size_t)P->np[c] * nm + 1, sizeof(double));
    for (int a = 0; a < 3; ++a) {
      csc_build(&P->A[c][a], &P->cls[c], &P->cls[sw(c, a)], a, adjoint ? -1.0 : 1.0);
      if (odd_in(c, a)) {
        P->B[c][a] = malloc(sizeof(double) * nm * P->cls[sw(c, a)].n + 1);
        ref_boundary_dense(d->N, c, a, adjoint, P->B[c][a]);
        if (adjoint && g_plant == 1)
          for (int q = 0; q < nm * P->cls[sw(c, a)].n; ++q) P->B[c][a][q] = -P->B[c][a][q];
      }
    }
    if (!adjoint && odd_in(c, 0)) {
      /* setup_boundary_condition! (pn_problem.jl:192-203): right face = mms .* ppos */
      int ny = P->n[c][1], nz = P->n[c][2];
      P->gx[c] = calloc((size_t)2 * ny * nz * nm + 1, sizeof(double));
      P->gset[c] = calloc((size_t)ny * nz * nm + 1, sizeof(double));
      if (in && in->mms[c] && in->ppos[c])
        for (int k = 0; k < nz; ++k)
          for (int j = 0; j < ny; ++j)
            for (int m = 0; m < nm; ++m) P->gset[c][(k * ny + j) * nm + m] = in->mms[c][m] * in->ppos[c][k * ny + j];
    }
  }
  return 0;
}

static void prob_free(prob_t* P) {
  for (int c = 0; c < 8; ++c) {
    free(P->S[c]);
    free(P->C[c]);
    free(P->Q[c]);
    free(P->gx[c]);
    free(P->gset[c]);
    for (int a = 0; a < 3; ++a) {
      free(P->A[c][a].colptr);
      free(P->A[c][a].rowval);
      free(P->A[c][a].nzval);
      free(P->B[c][a]);
    }
  }
}

/* update_coefficients! + update_boundary_condition! + weighted_sum!
 * (pn_problem.jl:157-271) from one energy-point record */
static void update_problem(prob_t* P, const double* rec, const double* const rho[8]) {
  const int ne = P->d.n_e, N = P->d.N;
  const double* s = rec;
  const double* cl = rec + ne;
  const double gfac = rec[ne + ne * (N + 1)];
  for (int c = 0; c < 8; ++c) {
    const int nm = P->cls[c].n;
    for (int p = 0; p < P->np[c]; ++p) {
      const double* r = rho[c] + (size_t)p * ne;
      double* Sp = P->S[c] + (size_t)p * nm;
      double* Cp = P->C[c] + (size_t)p * nm;
      for (int m = 0; m < nm; ++m) {
        double sacc = 0.0, cacc = 0.0;
        for (int e = 0; e < ne; ++e) {
          sacc += r[e] * s[e];
          cacc += r[e] * cl[e * (N + 1) + P->cls[c].l[m]];
        }
        Sp[m] = sacc;
        Cp[m] = cacc;
      }
    }
    if (!P->adjoint && P->gx[c]) {
      /* problem.Bx.g[eo] .= problem.gx_temp[eo] * c: face 1 stays zero */
      int ny = P->n[c][1], nz = P->n[c][2];
      size_t face = (size_t)ny * nz * nm;
      for (size_t q = 0; q < face; ++q) {
        P->gx[c][q] = 0.0;
        P->gx[c][face + q] = P->gset[c][q] * gfac;
      }
    }
  }
}

static size_t pidx(const prob_t* P, int c, int i, int j, int k) {
  return (size_t)i + (size_t)P->n[c][0] * ((size_t)j + (size_t)P->n[c][1] * k);
}

/* apply_boundary_source! / apply_boundary_zero! (boundary_conditions.jl:67-102) */
static void boundary_one(prob_t* P, double* const u[8], int c, int a) {
  const int v = sw(c, a), nm = P->cls[c].n, nv = P->cls[v].n;
  const double* M = P->B[c][a];
  const int src = (!P->adjoint && a == 0);
  int n[3] = {P->n[c][0], P->n[c][1], P->n[c][2]};
  double tmp[MAXC];
  for (int side = 0; side < 2; ++side) {
    int lo[3] = {0, 0, 0}, hi[3] = {n[0], n[1], n[2]};
    lo[a] = side == 0 ? 0 : n[a] - 1;
    hi[a] = lo[a] + 1;
    for (int k = lo[2]; k < hi[2]; ++k)
      for (int j = lo[1]; j < hi[1]; ++j)
        for (int i = lo[0]; i < hi[0]; ++i) {
          int iv[3] = {i, j, k}, in[3] = {i, j, k};
          if (side == 0) {
            in[a] = 1;                 /* right_index */
          } else {
            iv[a] = n[a] - 2;          /* v at left_index */
            in[a] = n[a] - 2;          /* u at left_index */
          }
          const double* vp = u[v] + pidx(P, v, iv[0], iv[1], iv[2]) * nv;
          double* up = u[c] + pidx(P, c, i, j, k) * nm;
          const double* un = u[c] + pidx(P, c, in[0], in[1], in[2]) * nm;
          for (int r = 0; r < nm; ++r) {
            double s = 0.0;
            for (int q = 0; q < nv; ++q) s += M[r * nv + q] * vp[q];
            tmp[r] = s;
          }
          for (int r = 0; r < nm; ++r) {
            if (src) {
              const double* g = P->gx[c] + ((size_t)side * n[1] * n[2] + (size_t)k * n[1] + j) * nm;
              up[r] = side == 0 ? 2. * (-tmp[r] + g[r]) - un[r] : 2. * (tmp[r] + g[r]) - un[r];
            } else {
              up[r] = side == 0 ? -2. * tmp[r] - un[r] : 2. * tmp[r] - un[r];
            }
          }
        }
  }
}

/* apply_boundary! (solver.jl:107-120) */
static void apply_boundary(prob_t* P, int parity, double* const u[8]) {
  for (int a = 0; a < 3; ++a) {
    if (P->single[a]) continue;
    for (int c = 0; c < 8; ++c)
      if (odd_in(c, a) && parity_of(c) == parity) boundary_one(P, u, c, a);
  }
}

static double expm1div(double c) { /* solver_core.jl:38-44 */
  if (fabs(c) <= 0.0002) return 1.0 + 0.5 * c + 1.0 / 6.0 * c * c;
  return (exp(c) - 1) / c;
}

/* kernel! (solver_core.jl:108-137) for one to-class */
static void kernel_class(prob_t* P, int c, double* const du[8], double* const u[8], double dt, int use_phi) {
  const int nm = P->cls[c].n;
  int n[3] = {P->n[c][0], P->n[c][1], P->n[c][2]};
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) { /* eo_range: even 1:n, odd 2:n-1 (1-based) */
    lo[a] = odd_in(c, a) ? 1 : 0;
    hi[a] = odd_in(c, a) ? n[a] - 1 : n[a];
  }
  for (int k = lo[2]; k < hi[2]; ++k)
    for (int j = lo[1]; j < hi[1]; ++j)
      for (int i = lo[0]; i < hi[0]; ++i) {
        size_t p = pidx(P, c, i, j, k);
        double* d = du[c] + p * nm;
        for (int m = 0; m < nm; ++m) d[m] = 0.0;
        for (int a = 0; a < 3; ++a) {
          if (n[a] == 1) continue; /* singleton dimension */
          int f = sw(c, a), nf = P->cls[f].n;
          int iR[3] = {i, j, k}, iL[3] = {i, j, k};
          int idx = a == 0 ? i : (a == 1 ? j : k);
          if (odd_in(c, a)) { /* eo_index_right/left (solver_core.jl:21-35) */
            iR[a] = idx;
            iL[a] = idx - 1;
          } else {
            iR[a] = idx + 1;
            iL[a] = idx;
          }
          const double* uR = u[f] + pidx(P, f, iR[0], iR[1], iR[2]) * nf;
          const double* uL = u[f] + pidx(P, f, iL[0], iL[1], iL[2]) * nf;
          const csc_t* A = &P->A[c][a];
          for (int m = 0; m < A->ncol; ++m) { /* matmuldiff! (solver_core.jl:59-69) */
            double dd = (uR[m] - uL[m]) / P->h[a];
            for (int z = A->colptr[m]; z < A->colptr[m + 1]; ++z) d[A->rowval[z]] += A->nzval[z] * dd;
          }
        }
        const double* up = u[c] + p * nm;
        const double* S = P->S[c] + p * nm;
        const double* C = P->C[c] + p * nm;
        const double* Q = P->Q[c] ? P->Q[c] + p * nm : NULL;
        for (int m = 0; m < nm; ++m) { /* right_hand_side! (solver_core.jl:92-106) */
          double q = Q ? Q[m] : 0.0;
          double t = Q ? (q - d[m] - C[m] * up[m]) : (-d[m] - C[m] * up[m]);
          t = t / S[m];
          if (use_phi) t = t * expm1div(-C[m] / S[m] * dt);
          d[m] = t;
        }
      }
}

/* calc_du! and update! (solver.jl:126-168) */
static void calc_du(prob_t* P, int parity, double* const du[8], double* const u[8], double dt, int use_phi) {
  for (int c = 0; c < 8; ++c)
    if (parity_of(c) == parity) kernel_class(P, c, du, u, dt, use_phi);
}
static void update(prob_t* P, int parity, double* const u[8], double* const du[8], double dt) {
  for (int c = 0; c < 8; ++c) {
    if (parity_of(c) != parity) continue;
    size_t tot = (size_t)P->np[c] * P->cls[c].n;
    for (size_t q = 0; q < tot; ++q) u[c][q] = u[c][q] + dt * du[c][q];
  }
}

/* step_pde! (solver.jl:171-196): even boundary, odd half step, odd boundary,
 * even full step, even boundary, odd half step */
static void step_pde(prob_t* P, double* const u[8], double* const du[8], double dt) {
  double h = dt / 2;
  apply_boundary(P, 0, u);
  calc_du(P, 1, du, u, h, 1);
  update(P, 1, u, du, h);
  apply_boundary(P, 1, u);
  calc_du(P, 0, du, u, dt, 1);
  update(P, 0, u, du, dt);
  apply_boundary(P, 0, u);
  calc_du(P, 1, du, u, h, 1);
  update(P, 1, u, du, h);
}

static void alloc_state(const prob_t* P, double* u[8]) {
  for (int c = 0; c < 8; ++c) u[c] = calloc((size_t)P->np[c] * P->cls[c].n + 1, sizeof(double));
}
static void free_state(double* u[8]) {
  for (int c = 0; c < 8; ++c) free(u[c]);
}
static void copy_state(const prob_t* P, double* dst[8], double* const src[8]) {
  for (int c = 0; c < 8; ++c) memcpy(dst[c], src[c], sizeof(double) * P->np[c] * P->cls[c].n);
}

/* integrate_step_and_add! (x_ray_intensities.jl:172-191) */
static void integrate_step_and_add(const prob_t* P, const ref_inputs* in, double* I, double* const u[8],
                                   const double* sigma, double dE, double int_fac) {
  const int nl = P->d.n_lines, ne = P->d.n_e, n0 = P->cls[0].n;
  for (int r = 0; r < nl; ++r) {
    double integral = 0.0;
    for (int p = 0; p < P->np[0]; ++p)
      integral += in->ac[(size_t)p * nl + r] * in->noa[(size_t)p * ne + in->el_idx[r]] * u[0][(size_t)p * n0];
    I[r] += dE * P->dvol * sigma[r] * integral * int_fac;
  }
}

int ref_step(const ref_desc* d, int adjoint, const double* rec, const double* const rho[8], const double* const mms[8],
             const double* const ppos[8], const double* q_eee, double* const u[8]) {
  prob_t P;
  ref_inputs in;
  memset(&in, 0, sizeof(in));
  for (int c = 0; c < 8; ++c) {
    in.mms[c] = mms ? mms[c] : NULL;
    in.ppos[c] = ppos ? ppos[c] : NULL;
  }
  if (prob_init(&P, d, adjoint, &in)) return -1;
  update_problem(&P, rec, rho);
  if (adjoint && q_eee) {
    int n0 = P.cls[0].n;
    for (int p = 0; p < P.np[0]; ++p) P.Q[0][(size_t)p * n0] = q_eee[p];
  }
  double* du[8];
  alloc_state(&P, du);
  step_pde(&P, u, du, (d->E_initial - d->E_cutoff) / d->n_steps);
  free_state(du);
  prob_free(&P);
  return 0;
}

int ref_forward(const ref_desc* d, const ref_inputs* in, const double* fwd_rec, double* I_out) {
  prob_t P;
  if (prob_init(&P, d, 0, in)) return -1;
  const int L = rec_len(d), ne = d->n_e, N = d->N;
  const double dE = (d->E_initial - d->E_cutoff) / d->n_steps;
  double *u[8], *du[8];
  alloc_state(&P, u);
  alloc_state(&P, du);
  for (int r = 0; r < d->n_lines; ++r) I_out[r] = 0.0;
  const int sig = ne + ne * (N + 1) + 1;
  integrate_step_and_add(&P, in, I_out, u, fwd_rec + sig, dE, 0.5);
  for (int i = 1; i <= d->n_steps; ++i) { /* pn_intensities.jl:341-349 */
    const double* rec = fwd_rec + (size_t)i * L;
    update_problem(&P, rec, in->rho);
    step_pde(&P, u, du, dE);
    /* int_fac = i == N_steps ? 0.5 : 1.0 compares an Int with a function: always 1.0 */
    integrate_step_and_add(&P, in, I_out, u, rec + sig, dE, 1.0);
  }
  free_state(u);
  free_state(du);
  prob_free(&P);
  return 0;
}

/* integrate_adjoint_densities! (pn_intensities.jl:129-187) */
static void integrate_adjoint_densities(prob_t* F, const ref_inputs* in, const double* rec, double* const u[8],
                                        double* const du[8], double* const lam[8], double dE, double* const rho_bar[8],
                                        double int_fac) {
  update_problem(F, rec, in->rho);
  apply_boundary(F, 0, u);
  apply_boundary(F, 1, u);
  calc_du(F, 0, du, u, dE, 0);
  calc_du(F, 1, du, u, dE, 0);
  const int ne = F->d.n_e, N = F->d.N;
  const double* s = rec;
  const double* cl = rec + ne;
  for (int c = 0; c < 8; ++c) {
    const int nm = F->cls[c].n;
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = odd_in(c, a) ? 1 : 0;
      hi[a] = odd_in(c, a) ? F->n[c][a] - 1 : F->n[c][a];
    }
    for (int i = lo[0]; i < hi[0]; ++i)
      for (int j = lo[1]; j < hi[1]; ++j)
        for (int k = lo[2]; k < hi[2]; ++k) {
          size_t p = pidx(F, c, i, j, k);
          for (int e = 0; e < ne; ++e)
            for (int m = 0; m < nm; ++m) {
              double ce = g_plant == 2 ? 0.0 : cl[e * (N + 1) + F->cls[c].l[m]];
              rho_bar[c][p * ne + e] +=
                  lam[c][p * nm + m] * (s[e] * du[c][p * nm + m] + ce * u[c][p * nm + m]) * F->dvol * dE * int_fac;
            }
        }
  }
}

/* integrate_step_and_add_pullback! (x_ray_intensities.jl:193-223) */
static void intensity_pullback(const prob_t* P, const ref_inputs* in, double* const u[8], const double* sigma, double dE,
                               double int_fac, const double* I_bar, double* noa_bar, double* ac_bar) {
  const int nl = P->d.n_lines, ne = P->d.n_e, n0 = P->cls[0].n;
  double integral_[64];
  for (int r = 0; r < nl; ++r) integral_[r] = dE * P->dvol * sigma[r] * int_fac * I_bar[r];
  for (int p = 0; p < P->np[0]; ++p)
    for (int r = 0; r < nl; ++r) {
      int e = in->el_idx[r];
      ac_bar[(size_t)p * nl + r] += in->noa[(size_t)p * ne + e] * u[0][(size_t)p * n0] * integral_[r];
      noa_bar[(size_t)p * ne + e] += in->ac[(size_t)p * nl + r] * u[0][(size_t)p * n0] * integral_[r];
    }
}

/* compute_adjoint_source! (x_ray_intensities.jl:155-170) */
static void adjoint_source(prob_t* A, const ref_inputs* in, const double* sigma, const double* I_bar) {
  const int nl = A->d.n_lines, ne = A->d.n_e, n0 = A->cls[0].n;
  for (int p = 0; p < A->np[0]; ++p) {
    double q = 0.0;
    for (int r = 0; r < nl; ++r)
      q -= sigma[r] * in->ac[(size_t)p * nl + r] * in->noa[(size_t)p * ne + in->el_idx[r]] * I_bar[r];
    A->Q[0][(size_t)p * n0] = g_plant == 3 ? -q : q;
  }
}

int ref_forward_adjoint(const ref_desc* d, const ref_inputs* in, const double* fwd_rec, const double* adj_rec,
                        const double* dens_rec, const double* I_bar, double* I_out, double* const rho_bar[8],
                        double* ac_bar, double* noa_bar) {
  if (d->n_lines > 64) return -1;
  prob_t F, A;
  if (prob_init(&F, d, 0, in)) return -1;
  if (prob_init(&A, d, 1, in)) {
    prob_free(&F);
    return -1;
  }
  const int L = rec_len(d), ne = d->n_e, N = d->N, nst = d->n_steps;
  const int sig = ne + ne * (N + 1) + 1;
  const double dE = (d->E_initial - d->E_cutoff) / nst;
  double *u[8], *du[8], *lam[8], *dlam[8];
  alloc_state(&F, u);
  alloc_state(&F, du);
  alloc_state(&F, lam);
  alloc_state(&F, dlam);
  double*** tape = malloc(sizeof(double**) * (nst + 1));
  for (int i = 0; i <= nst; ++i) {
    tape[i] = malloc(sizeof(double*) * 8);
    alloc_state(&F, tape[i]);
  }
  for (int r = 0; r < d->n_lines; ++r) I_out[r] = 0.0;
  copy_state(&F, tape[0], u);
  integrate_step_and_add(&F, in, I_out, u, fwd_rec + sig, dE, 0.5);
  for (int i = 1; i <= nst; ++i) {
    const double* rec = fwd_rec + (size_t)i * L;
    update_problem(&F, rec, in->rho);
    step_pde(&F, u, du, dE);
    copy_state(&F, tape[i], u);
    integrate_step_and_add(&F, in, I_out, u, rec + sig, dE, 1.0);
  }
  /* pullback (pn_intensities.jl:381-411) */
  for (int c = 0; c < 8; ++c) memset(rho_bar[c], 0, sizeof(double) * F.np[c] * ne);
  memset(ac_bar, 0, sizeof(double) * F.np[0] * d->n_lines);
  memset(noa_bar, 0, sizeof(double) * F.np[0] * ne);
  for (int i = nst; i >= 1; --i) {
    const double* drec = dens_rec + (size_t)i * L;
    const double* arec = adj_rec + (size_t)i * L;
    integrate_adjoint_densities(&F, in, drec, tape[i], du, lam, dE, rho_bar, 1.0);
    intensity_pullback(&F, in, tape[i], drec + sig, dE, 1.0, I_bar, noa_bar, ac_bar);
    adjoint_source(&A, in, arec + sig, I_bar);
    update_problem(&A, arec, in->rho);
    step_pde(&A, lam, dlam, dE);
  }
  integrate_adjoint_densities(&F, in, dens_rec, tape[0], du, lam, dE, rho_bar, 0.5);
  intensity_pullback(&F, in, tape[0], dens_rec + sig, dE, 0.5, I_bar, noa_bar, ac_bar);
  for (int i = 0; i <= nst; ++i) {
    free_state(tape[i]);
    free(tape[i]);
  }
  free(tape);
  free_state(u);
  free_state(du);
  free_state(lam);
  free_state(dlam);
  prob_free(&F);
  prob_free(&A);
  return 0;
}
