// Host-side closed forms of the PN advection and boundary matrices, packed into
// the coefficient buffer layout of pnx_tables.hpp.
//
// Advection entries follow pn_solver/pn_matrices.jl:24-152 (helper functions
// a..f, A_minus / A_plus per axis, get_coefficient).  Boundary blocks follow
// pn_solver/pn_matrices.jl:191-293: B = L * A for the forward problem and
// B = -L^T * A for the adjoint problem, with L the half-range coupling of the
// to-class moments.  The adjoint problem also negates A
// (pn_solver/pn_problem.jl:113-118).
//
// The reference evaluates the boundary sums with BigInt factorials and Float64
// gamma values; here they are evaluated in long double (64-bit mantissa),
// which reproduces the same closed form to well below double rounding.
#include <cmath>
#include <vector>

#include "pnx_tables.hpp"

namespace pnx {
namespace {

using ld = long double;

int floor_div2(int a) { return a >= 0 ? a / 2 : -((-a + 1) / 2); }

ld fact(int n) {
  ld r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
// (n)!! for n >= 0, and 1 for negative n (pn_solver/pn_matrices.jl:200-206)
ld dfact(int n) {
  if (n < 0) return 1;
  ld r = 1;
  for (int i = n; i > 1; i -= 2) r *= i;
  return r;
}
ld binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  ld r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}
ld ipow_m1(int e) { return (e % 2 == 0) ? 1.0L : -1.0L; }

double sq(long num, long den) { return std::sqrt(static_cast<double>(num) / static_cast<double>(den)); }

double fa(int k, int l) { return sq((long)(l - k + 1) * (l + k + 1), (long)(2 * l + 3) * (2 * l + 1)); }
double fb(int k, int l) { return sq((long)(l - k) * (l + k), (long)(2 * l + 1) * (2 * l - 1)); }
double fc(int k, int l) {
  const double v = sq((long)(l + k + 1) * (l + k + 2), (long)(2 * l + 3) * (2 * l + 1));
  return k < 0 ? 0.0 : (k > 0 ? v : v * std::sqrt(2.0));
}
double fd(int k, int l) {
  const double v = sq((long)(l - k) * (l - k - 1), (long)(2 * l + 1) * (2 * l - 1));
  return k < 0 ? 0.0 : (k > 0 ? v : v * std::sqrt(2.0));
}
double fe(int k, int l) {
  const double v = sq((long)(l - k + 1) * (l - k + 2), (long)(2 * l + 3) * (2 * l + 1));
  return k == 1 ? v * std::sqrt(2.0) : v;
}
double ff(int k, int l) {
  const double v = sq((long)(l + k) * (l + k - 1), (long)(2 * l + 1) * (2 * l - 1));
  return k == 1 ? v * std::sqrt(2.0) : v;
}
int theta(int k) { return k < 0 ? -1 : 1; }

double a_minus(int l, int k, int kf, int d) {
  const int ak = k < 0 ? -k : k;
  if (d == 0) {
    if (kf == minus_k(k) && k != -1) return 0.5 * fc(ak - 1, l - 1);
    if (kf == plus_k(k)) return -0.5 * fe(ak + 1, l - 1);
    return 0.0;
  }
  if (d == 1) {
    if (kf == -minus_k(k) && k != 1) return (-theta(k) / 2.0) * fc(ak - 1, l - 1);
    if (kf == -plus_k(k)) return (-theta(k) / 2.0) * fe(ak + 1, l - 1);
    return 0.0;
  }
  return kf == k ? fa(k, l - 1) : 0.0;
}

double a_plus(int l, int k, int kf, int d) {
  const int ak = k < 0 ? -k : k;
  if (d == 0) {
    if (kf == minus_k(k) && k != -1) return -0.5 * fd(ak - 1, l + 1);
    if (kf == plus_k(k)) return 0.5 * ff(ak + 1, l + 1);
    return 0.0;
  }
  if (d == 1) {
    if (kf == -minus_k(k) && k != 1) return (theta(k) / 2.0) * fd(ak - 1, l + 1);
    if (kf == -plus_k(k)) return (theta(k) / 2.0) * ff(ak + 1, l + 1);
    return 0.0;
  }
  return kf == k ? fb(k, l + 1) : 0.0;
}

}  // namespace

// Advection coefficient between to-moment (l, k) and from-moment (lf, kf) along axis d
// (pn_solver/pn_matrices.jl:132-144).
double advection_coefficient(int l, int k, int lf, int kf, int d) {
  if (l == lf - 1) return a_plus(l, k, kf, d);
  if (l == lf + 1) return a_minus(l, k, kf, d);
  return 0.0;
}

namespace {

ld cnorm(int l, int k) {  // pn_solver/pn_matrices.jl:191-198, k >= 0
  const ld v = std::sqrt((2.0L * l + 1.0L) / (2.0L * 3.14159265358979323846264338327950288L) * fact(l - k) / fact(l + k));
  return (k % 2 == 0) ? v : -v;
}

}  // namespace

// Half-range boundary coupling between moments (l, k) and (lf, kf) along axis d
// (pn_solver/pn_matrices.jl:212-275).
double boundary_coefficient(int l, int k, int lf, int kf, int d) {
  const int ak = k < 0 ? -k : k, akf = kf < 0 ? -kf : kf;
  if (d == 2) {
    if (k != kf) return 0.0;
    ld sum = 0;
    for (int j = 0; j <= floor_div2(l - ak); ++j)
      for (int q = 0; q <= floor_div2(lf - akf); ++q) {
        const ld sgn = ipow_m1(q + j) / std::pow(2.0L, q + j);
        sum += sgn / (fact(q) * fact(j)) * dfact(2 * (l - j) - 1) / fact(l - ak - 2 * j) *
               dfact(2 * (lf - q) - 1) / fact(lf - akf - 2 * q) *
               std::tgamma((l + lf) / 2.0L - q - j - ak) / std::tgamma((l + lf) / 2.0L + 1 - q - j);
      }
    return static_cast<double>(sum * 3.14159265358979323846264338327950288L * cnorm(l, ak) * cnorm(lf, akf) * fact(ak));
  }
  ld sum_jq = 0;
  for (int j = 0; j <= floor_div2(l - ak); ++j)
    for (int q = 0; q <= floor_div2(lf - akf); ++q) {
      ld sum_is = 0;
      for (int i = 0; i <= floor_div2(ak - 1); ++i)
        for (int s = 0; s <= floor_div2(akf - 1); ++s) {
          if (d == 0) {
            if (k >= 0 && kf >= 0 && k % 2 != 0 && kf % 2 != 0)
              sum_is += ipow_m1(s + i) * binom(k, 2 * i) * binom(kf, 2 * s) * std::tgamma(0.5L + s + i) *
                        std::tgamma((k + kf) / 2.0L - (s + i));
            else if (k < 0 && kf < 0 && k % 2 == 0 && kf % 2 == 0)
              sum_is += ipow_m1(s + i) * binom(ak, 2 * i + 1) * binom(akf, 2 * s + 1) * std::tgamma(1.5L + s + i) *
                        std::tgamma((ak + akf) / 2.0L - (s + i) - 1);
          } else {
            sum_is += ipow_m1(s + i) * fact(s + i) * binom(ak, 2 * i + 1) * binom(akf, 2 * s + 1) *
                      std::tgamma((ak + akf) / 2.0L - s - i - 0.5L);
          }
        }
      sum_jq += sum_is * ipow_m1(q + j) / (std::pow(2.0L, j + q) * fact(j) * fact(q)) * dfact(2 * (l - j) - 1) /
                fact(l - ak - 2 * j) * dfact(2 * (lf - q) - 1) / fact(lf - akf - 2 * q) *
                std::tgamma((l + lf + 1 - 2 * (q + j) - (ak + akf)) / 2.0L) / std::tgamma((l + lf) / 2.0L + 1 - (q + j));
    }
  return static_cast<double>(cnorm(l, ak) * cnorm(lf, akf) * (1 + ipow_m1(l + lf)) * ipow_m1(ak + akf) * sum_jq);
}

// Fill the coefficient buffer (layout of make_tables(N)) for the forward
// (adjoint = false) or adjoint (adjoint = true) problem.
std::vector<double> build_coefficients(int N, bool adjoint) {
  const Tables t = make_tables(N);
  std::vector<double> out(t.n_coef, 0.0);
  const double sign = adjoint ? -1.0 : 1.0;
  for (int c = 0; c < 8; ++c)
    for (int d = 0; d < 3; ++d) {
      const int f = switch_axis(c, d);
      for (int i = 0; i < t.n[c]; ++i)
        for (int j = 0; j < t.n[f]; ++j) {
          const int z = t.nz[c][d][i][j];
          if (z >= 0) out[z] = sign * advection_coefficient(t.l[c][i], t.k[c][i], t.l[f][j], t.k[f][j], d);
        }
    }
  for (int d = 0; d < 3; ++d)
    for (int c = 0; c < 8; ++c) {
      if (!odd_in(c, d)) continue;
      const int f = switch_axis(c, d), nc = t.n[c], nf = t.n[f];
      // L over the to-class moments; A = forward advection block c <- f (unsigned)
      std::vector<double> L(nc * nc), A(nc * nf);
      for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nc; ++j) L[i * nc + j] = boundary_coefficient(t.l[c][i], t.k[c][i], t.l[c][j], t.k[c][j], d);
      for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nf; ++j) A[i * nf + j] = advection_coefficient(t.l[c][i], t.k[c][i], t.l[f][j], t.k[f][j], d);
      double* B = out.data() + t.b_off[c][d];
      for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nf; ++j) {
          double acc = 0.0;
          for (int m = 0; m < nc; ++m) acc += (adjoint ? L[m * nc + i] : L[i * nc + m]) * A[m * nf + j];
          B[i * nf + j] = adjoint ? -acc : acc;
        }
    }
  return out;
}

}  // namespace pnx
