// Compile-time moment, class and sparsity tables for the PN-CSD staggered solver.
//
// Everything here is integer bookkeeping that both the host setup code and the
// gfx950 kernels need to agree on, so it lives in one constexpr header:
//
//  * real spherical-harmonic moments (l, k), 0 <= l <= N, -l <= k <= l, in the
//    order l ascending then k ascending (pn_solver/moment_specification.jl:42-44);
//  * their even/odd classification per spatial axis
//    (pn_solver/moment_specification.jl:26-36) and the class index
//    4*ox + 2*oy + oz, which is the eee, eeo, eoe, eoo, oee, oeo, ooe, ooo
//    order of starmap/src/even_odd_classification.jl:28-37 / all_eos();
//  * the parity of a class (starmap/src/even_odd_classification.jl:72-88):
//    even parity = {eee, ooe, oeo, eoo}, i.e. an even number of odd axes;
//  * the sparsity pattern of the advection blocks A_d[c] (rows: moments of c,
//    columns: moments of the class switched in axis d), following the k'
//    selection rules of pn_solver/pn_matrices.jl:76-144, enumerated column by
//    column (CSC order, as starmap/src/solver_core.jl:59-69 traverses them);
//  * the offsets of every value in the per-(N, direction) coefficient buffer
//    that the kernels read: all A nonzeros, then the dense boundary blocks.
#pragma once

#include <type_traits>
#include <utility>

#if defined(__HIPCC__)
#define PNX_HD __host__ __device__
#else
#define PNX_HD
#endif

namespace pnx {

constexpr int kMaxOrder = 9;    // largest PN order compiled into the kernels
constexpr int kMaxClass = 15;   // moments in the largest class at N = 9

PNX_HD constexpr int x_odd(int k) { return ((k < 0 && k % 2 != 0) || (k >= 0 && k % 2 == 0)) ? 0 : 1; }
PNX_HD constexpr int y_odd(int k) { return k >= 0 ? 0 : 1; }
PNX_HD constexpr int z_odd(int l, int k) { return ((l + k) % 2 == 0) ? 0 : 1; }
PNX_HD constexpr int class_of(int l, int k) { return 4 * x_odd(k) + 2 * y_odd(k) + z_odd(l, k); }
// odd-ness of class c along axis d (d = 0: x, 1: y, 2: z)
PNX_HD constexpr int odd_in(int c, int d) { return (c >> (2 - d)) & 1; }
// class obtained by switching even/odd along axis d
PNX_HD constexpr int switch_axis(int c, int d) { return c ^ (4 >> d); }
// 0 = even parity {eee, ooe, oeo, eoo}, 1 = odd parity {oee, eoe, eeo, ooo}
PNX_HD constexpr int parity_of(int c) { return (odd_in(c, 0) + odd_in(c, 1) + odd_in(c, 2)) & 1; }
// A class is updated by the solver only if it is even in every singleton axis:
// for an axis of length 1 the odd interior range 2:n-1 is empty
// (starmap/src/solver_core.jl:17-19).  dims = number of leading non-singleton axes.
PNX_HD constexpr bool class_active(int c, int dims) {
  return (dims >= 2 || odd_in(c, 1) == 0) && (dims >= 3 || odd_in(c, 2) == 0);
}

PNX_HD constexpr int plus_k(int k) { return k == 0 ? 1 : k + (k > 0 ? 1 : -1); }
PNX_HD constexpr int minus_k(int k) { return k == 0 ? -1 : k - (k > 0 ? 1 : -1); }

// Structural nonzero test of the advection coefficient between to-moment (l, k)
// and from-moment (lf, kf) along axis d (pn_solver/pn_matrices.jl:76-144).
PNX_HD constexpr bool a_pattern(int l, int k, int lf, int kf, int d) {
  if (!(l == lf - 1 || l == lf + 1)) return false;
  if (d == 0) return (kf == minus_k(k) && k != -1) || kf == plus_k(k);
  if (d == 1) return (kf == -minus_k(k) && k != 1) || kf == -plus_k(k);
  return kf == k;
}

// ---------------------------------------------------------------------------
// Compile-time advection values.  The closed forms of pn_solver/pn_matrices.jl:24-144
// are sqrt((integer)/(integer)) times 1/2, sqrt(2) or a sign; cx_sqrt returns the
// correctly rounded double square root (Newton in double, then the neighbour
// whose exact square residual is smallest), so these constants equal the
// runtime std::sqrt values bit for bit (checked by the host tests).
PNX_HD constexpr double cx_abs(double x) { return x < 0 ? -x : x; }

// exact a*b = p + e (Dekker's product; no FMA needed)
PNX_HD constexpr void cx_two_prod(double a, double b, double& p, double& e) {
  p = a * b;
  const double sa = 134217729.0 * a, ah = sa - (sa - a), al = a - ah;
  const double sb = 134217729.0 * b, bh = sb - (sb - b), bl = b - bh;
  e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
}

PNX_HD constexpr double cx_ulp(double r) {  // r > 0, normal
  double u = 1.0;
  while (u <= r) u *= 2.0;
  while (u > r) u *= 0.5;  // u = 2^floor(log2 r)
  return u * 2.220446049250313e-16;
}

PNX_HD constexpr double cx_sqrt(double x) {
  if (!(x > 0.0)) return 0.0;
  double r = x > 1.0 ? x : 1.0;
  for (int it = 0; it < 200; ++it) {
    const double nr = 0.5 * (r + x / r);
    if (nr == r) break;
    r = nr;
  }
  double best = r, bestd = -1.0;
  for (int k = -2; k <= 2; ++k) {
    const double c = r + k * cx_ulp(r);
    double pp = 0.0, ee = 0.0;
    cx_two_prod(c, c, pp, ee);
    const double d = cx_abs((x - pp) - ee);
    if (bestd < 0.0 || d < bestd) {
      bestd = d;
      best = c;
    }
  }
  return best;
}

PNX_HD constexpr double cx_rt(long a, long b) { return cx_sqrt((double)a / (double)b); }
PNX_HD constexpr double cx_sqrt2() { return cx_sqrt(2.0); }
PNX_HD constexpr double cx_a(int k, int l) { return cx_rt((long)(l - k + 1) * (l + k + 1), (long)(2 * l + 3) * (2 * l + 1)); }
PNX_HD constexpr double cx_b(int k, int l) { return cx_rt((long)(l - k) * (l + k), (long)(2 * l + 1) * (2 * l - 1)); }
PNX_HD constexpr double cx_c(int k, int l) {
  const double v = cx_rt((long)(l + k + 1) * (l + k + 2), (long)(2 * l + 3) * (2 * l + 1));
  return k < 0 ? 0.0 : (k > 0 ? v : v * cx_sqrt2());
}
PNX_HD constexpr double cx_d(int k, int l) {
  const double v = cx_rt((long)(l - k) * (l - k - 1), (long)(2 * l + 1) * (2 * l - 1));
  return k < 0 ? 0.0 : (k > 0 ? v : v * cx_sqrt2());
}
PNX_HD constexpr double cx_e(int k, int l) {
  const double v = cx_rt((long)(l - k + 1) * (l - k + 2), (long)(2 * l + 3) * (2 * l + 1));
  return k == 1 ? v * cx_sqrt2() : v;
}
PNX_HD constexpr double cx_f(int k, int l) {
  const double v = cx_rt((long)(l + k) * (l + k - 1), (long)(2 * l + 1) * (2 * l - 1));
  return k == 1 ? v * cx_sqrt2() : v;
}
PNX_HD constexpr double cx_advection(int l, int k, int lf, int kf, int d) {
  const int ak = k < 0 ? -k : k;
  const double th = k < 0 ? -1.0 : 1.0;
  if (l == lf + 1) {  // A_minus
    if (d == 0) {
      if (kf == minus_k(k) && k != -1) return 0.5 * cx_c(ak - 1, l - 1);
      if (kf == plus_k(k)) return -0.5 * cx_e(ak + 1, l - 1);
      return 0.0;
    }
    if (d == 1) {
      if (kf == -minus_k(k) && k != 1) return (-th / 2.0) * cx_c(ak - 1, l - 1);
      if (kf == -plus_k(k)) return (-th / 2.0) * cx_e(ak + 1, l - 1);
      return 0.0;
    }
    return kf == k ? cx_a(k, l - 1) : 0.0;
  }
  if (l == lf - 1) {  // A_plus
    if (d == 0) {
      if (kf == minus_k(k) && k != -1) return -0.5 * cx_d(ak - 1, l + 1);
      if (kf == plus_k(k)) return 0.5 * cx_f(ak + 1, l + 1);
      return 0.0;
    }
    if (d == 1) {
      if (kf == -minus_k(k) && k != 1) return (th / 2.0) * cx_d(ak - 1, l + 1);
      if (kf == -plus_k(k)) return (th / 2.0) * cx_f(ak + 1, l + 1);
      return 0.0;
    }
    return kf == k ? cx_b(k, l + 1) : 0.0;
  }
  return 0.0;
}

struct Tables {
  int N;
  int n[8];                          // moments per class
  int l[8][kMaxClass];               // degree of each class moment
  int k[8][kMaxClass];               // order of each class moment
  int nz[8][3][kMaxClass][kMaxClass];  // index into coefficient buffer, -1 if structurally zero
  int n_a;                           // number of A values
  int b_off[8][3];                   // offset of dense boundary block B_d[c] (row-major n[c] x n[switch]), -1 if c even in d
  double a[8][3][kMaxClass][kMaxClass];  // advection values (forward sign) at the structural nonzeros
  int n_coef;                        // total doubles in the coefficient buffer
};

PNX_HD constexpr Tables make_tables(int N) {
  Tables t{};
  t.N = N;
  for (int c = 0; c < 8; ++c) t.n[c] = 0;
  for (int l = 0; l <= N; ++l)
    for (int k = -l; k <= l; ++k) {
      const int c = class_of(l, k);
      t.l[c][t.n[c]] = l;
      t.k[c][t.n[c]] = k;
      t.n[c] += 1;
    }
  int idx = 0;
  for (int c = 0; c < 8; ++c)
    for (int d = 0; d < 3; ++d) {
      const int f = switch_axis(c, d);
      for (int i = 0; i < kMaxClass; ++i)
        for (int j = 0; j < kMaxClass; ++j) {
          t.nz[c][d][i][j] = -1;
          t.a[c][d][i][j] = 0.0;
        }
      for (int j = 0; j < t.n[f]; ++j)       // column (from-moment) major: CSC order
        for (int i = 0; i < t.n[c]; ++i)
          if (a_pattern(t.l[c][i], t.k[c][i], t.l[f][j], t.k[f][j], d)) {
            t.nz[c][d][i][j] = idx++;
            t.a[c][d][i][j] = cx_advection(t.l[c][i], t.k[c][i], t.l[f][j], t.k[f][j], d);
          }
    }
  t.n_a = idx;
  for (int d = 0; d < 3; ++d)
    for (int c = 0; c < 8; ++c) {
      if (odd_in(c, d)) {
        t.b_off[c][d] = idx;
        idx += t.n[c] * t.n[switch_axis(c, d)];
      } else {
        t.b_off[c][d] = -1;
      }
    }
  t.n_coef = idx;
  return t;
}

template <int N>
struct Tab {
  static constexpr Tables t = make_tables(N);
};

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, class F>
PNX_HD constexpr void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

}  // namespace pnx
