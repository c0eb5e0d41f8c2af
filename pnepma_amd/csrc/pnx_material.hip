// Material layer on the device: partial densities of every staggered class
// from per-cell mass fractions, and the pullback of a density gradient to
// those fractions.
//
// Material: MassFraction (pn_solver/pn_parametrizations/pn_material.jl:199-264),
// rho_e = w_e / sum_e'(w_e' / rho_pure_e'), with its pullback
// (component_densities_pullback!, same file).  Parametrization: boxes
// (pn_parametrization_boxed.jl:137-175) whose boxes are the grid's cells, i.e.
// one box per even point with edges at the odd points.  A class point then
// lies in the box with the same index along an even axis; along an odd axis
// point i sits on the lower edge of box i, and the last point (i = M) on the
// upper outer edge of box M-1, which the reference assigns to the last box.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pnx_kernels.hpp"

namespace pnx {
namespace {

constexpr int kMaxE = 8;

struct ClassGeom {
  int n[3];     // class points per axis
  int odd[3];   // class odd along the axis (and the axis resolved)
  long off;     // offset of the class in a [c][e][V] buffer
};

struct MatGeom {
  int M[3];      // grid points per axis (= cells)
  int X, Y;      // padded extents (point (i, j, k) -> i + X (j + Y k))
  long V;
  int n_e;
  ClassGeom cls[8];
  int n_cls;
  int cls_id[8];
};

__device__ __forceinline__ int cell_of(int i, int odd, int M) { return odd ? (i < M - 1 ? i : M - 1) : i; }

// One thread per point of class cls_id[blockIdx.y].
__global__ void densities_kernel(MatGeom g, const double* __restrict__ w, const double* __restrict__ rho_pure,
                                 double* __restrict__ rho) {
  const ClassGeom& c = g.cls[blockIdx.y];
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nq = (long)c.n[0] * c.n[1] * c.n[2];
  if (q >= nq) return;
  const int i = (int)(q % c.n[0]);
  const int j = (int)((q / c.n[0]) % c.n[1]);
  const int k = (int)(q / ((long)c.n[0] * c.n[1]));
  const long cell = cell_of(i, c.odd[0], g.M[0]) +
                    (long)g.M[0] * (cell_of(j, c.odd[1], g.M[1]) + (long)g.M[1] * cell_of(k, c.odd[2], g.M[2]));
  const double* __restrict__ wc = w + cell * g.n_e;
  double R = 0.0;
  for (int e = 0; e < g.n_e; ++e) R += wc[e] / rho_pure[e];
  const long p = i + (long)g.X * (j + (long)g.Y * k);
  for (int e = 0; e < g.n_e; ++e) rho[c.off + e * g.V + p] = wc[e] / R;
}

// One thread per cell: gathers the class points that lie in the cell (at most
// two per odd axis), sums their density gradient over beams in beam order
// and adds the MassFraction pullback of each point to the cell's fractions.
// With box_rb_in_grad, grad holds on entry the density gradient that the
// absorption pullback gathered per box (absorb_pullback_kernel), whose
// MassFraction pullback is added last.
// With rb_cell, rho_bar holds one gradient per cell ([b][e][V] at the cell's
// eee index), already summed over the cell's class points.
__global__ void pullback_kernel(MatGeom g, const double* __restrict__ w, const double* __restrict__ rho_pure,
                                const double* __restrict__ rho_bar, long rb_stride, int n_beams,
                                const double* __restrict__ extra_eee, int box_rb_in_grad, int rb_cell, double* grad) {
  const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n_cells = (long)g.M[0] * g.M[1] * g.M[2];
  if (cell >= n_cells) return;
  const int cx = (int)(cell % g.M[0]);
  const int cy = (int)((cell / g.M[0]) % g.M[1]);
  const int cz = (int)(cell / ((long)g.M[0] * g.M[1]));
  const int ne = g.n_e;
  double wc[kMaxE], acc[kMaxE];
  double R = 0.0;
  for (int e = 0; e < ne; ++e) {
    wc[e] = w[cell * ne + e];
    R += wc[e] / rho_pure[e];
    acc[e] = 0.0;
  }
  for (int ci = 0; ci < (rb_cell ? 1 : g.n_cls); ++ci) {
    const ClassGeom& c = g.cls[ci];
    const int cc[3] = {cx, cy, cz};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = cc[a];
      // odd axis: the last box also has point M (a per-cell gradient sits at the eee index)
      hi[a] = (!rb_cell && c.odd[a] && cc[a] == g.M[a] - 1) ? g.M[a] : cc[a];
      if (c.n[a] == 1) lo[a] = hi[a] = 0;
    }
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
          const long p = i + (long)g.X * (j + (long)g.Y * k);
          double rb[kMaxE];
          double dot = 0.0;
          for (int e = 0; e < ne; ++e) {
            double s = 0.0;
            for (int b = 0; b < n_beams; ++b) s += rho_bar[b * rb_stride + (rb_cell ? 0 : c.off) + e * g.V + p];
            if (g.cls_id[ci] == 0 && extra_eee) s += extra_eee[e * g.V + p];
            rb[e] = s;
            dot += wc[e] * s;
          }
          const double R_bar = -dot / (R * R);
          for (int e = 0; e < ne; ++e) acc[e] += rb[e] / R + R_bar / rho_pure[e];
        }
  }
  if (box_rb_in_grad) {
    double rb[kMaxE];
    double dot = 0.0;
    for (int e = 0; e < ne; ++e) {
      rb[e] = grad[cell * ne + e];
      dot += wc[e] * rb[e];
    }
    const double R_bar = -dot / (R * R);
    for (int e = 0; e < ne; ++e) acc[e] += rb[e] / R + R_bar / rho_pure[e];
  }
  for (int e = 0; e < ne; ++e) grad[cell * ne + e] = acc[e];
}

MatGeom make_mat_geom(const Layout& L, const int class_n[8][3], const long rho_off[8], bool (*active)(int, int),
                      const int* ncls_of) {
  MatGeom g{};
  g.M[0] = L.Mx;
  g.M[1] = L.My;
  g.M[2] = L.Mz;
  g.X = L.X;
  g.Y = L.Y;
  g.V = L.V;
  g.n_e = L.n_e;
  for (int c = 0; c < 8; ++c) {
    if (!(active(c, L.dims) && ncls_of[c] > 0)) continue;
    ClassGeom& cg = g.cls[g.n_cls];
    for (int a = 0; a < 3; ++a) {
      cg.n[a] = class_n[c][a];
      cg.odd[a] = ((c >> (2 - a)) & 1) && (a == 0 ? L.Mx : (a == 1 ? L.My : L.Mz)) > 1;
    }
    cg.off = rho_off[c];
    g.cls_id[g.n_cls] = c;
    ++g.n_cls;
  }
  return g;
}

bool class_active_dims(int c, int dims) { return (dims >= 2 || ((c >> 1) & 1) == 0) && (dims >= 3 || (c & 1) == 0); }

// ---------------------------------------------------------------- absorption
//
// compute_absorption_coefficients (pn_solver/x_ray_intensities.jl:100-128):
// from every eee point a 100-sample march towards the detector, cut where the
// ray crosses the plane x = beam_x0; each sample adds the attenuation
// sum_e mac[r][e] rho_e of the box it falls in (boxes found by arithmetic on
// the uniform edges; samples outside the boxes use the embedding fractions).
// The arithmetic follows the host restatement step by step, without
// contracting products into fused multiply-adds.

struct AbsGeom {
  const double* pts[3];   // even points per axis (eee point coordinates)
  const double* edges[3]; // box edges per axis (n[a] + 1 values)
  int n[3];               // boxes per axis
  double inv[3];          // n / (last edge - first edge)
  int M[3];
  int X, Y;
  long V;
  int n_e, n_lines;
  double det[3];
  double beam_x0;
};

AbsGeom abs_geom(const AbsorptionGeometry& a, const Layout& L) {
  AbsGeom g{};
  for (int d = 0; d < 3; ++d) {
    g.pts[d] = a.points[d];
    g.edges[d] = a.edges[d];
    g.n[d] = a.n_boxes[d];
    g.inv[d] = a.inv_span[d];
    g.det[d] = a.detector[d];
  }
  g.M[0] = L.Mx;
  g.M[1] = L.My;
  g.M[2] = L.Mz;
  g.X = L.X;
  g.Y = L.Y;
  g.V = L.V;
  g.n_e = L.n_e;
  g.n_lines = L.n_lines;
  g.beam_x0 = a.beam_x0;
  return g;
}

// attenuation table: rows = boxes (x fastest), then the embedding row
__global__ void mu_table_kernel(AbsGeom g, const double* __restrict__ w, const double* __restrict__ emb,
                                const double* __restrict__ rho_pure, const double* __restrict__ mac,
                                double* __restrict__ mu) {
#pragma clang fp contract(off)
  const long n_rows = (long)g.n[0] * g.n[1] * g.n[2] + 1;
  const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n_rows) return;
  const double* wr = row == n_rows - 1 ? emb : w + row * g.n_e;
  double R = 0.0;
  for (int e = 0; e < g.n_e; ++e) R += wr[e] / rho_pure[e];
  for (int r = 0; r < g.n_lines; ++r) {
    double m = 0.0;
    for (int e = 0; e < g.n_e; ++e) m += (wr[e] / R) * mac[r * g.n_e + e];
    mu[row * g.n_lines + r] = m;
  }
}

__device__ __forceinline__ void ray(const AbsGeom& g, int i, int j, int k, double pt[3], double d[3], double& lam,
                                    double& dx) {
#pragma clang fp contract(off)
  pt[0] = g.pts[0][i];
  pt[1] = g.pts[1][j];
  pt[2] = g.pts[2][k];
  for (int a = 0; a < 3; ++a) d[a] = g.det[a] - pt[a];
  lam = (g.beam_x0 - pt[0]) / d[0];
  if (!(lam > 0.0)) lam = 0.0;
  dx = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * lam / 100.0;
}

__device__ __forceinline__ double sample_t(int s) { return s == 99 ? 1.0 : s * (1.0 / 99.0); }

// One thread per eee point: ac[r][p] and the line weights W[r][p] = ac * noa.
__global__ void absorb_kernel(AbsGeom g, const double* __restrict__ mu, const double* __restrict__ rho0,
                              const double* __restrict__ atom_mass, const int* __restrict__ line_el,
                              double* __restrict__ ac, double* __restrict__ W) {
#pragma clang fp contract(off)
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nq = (long)g.M[0] * g.M[1] * g.M[2];
  if (q >= nq) return;
  const int i = (int)(q % g.M[0]), j = (int)((q / g.M[0]) % g.M[1]), k = (int)(q / ((long)g.M[0] * g.M[1]));
  double pt[3], d[3], lam, dx;
  ray(g, i, j, k, pt, d, lam, dx);
  const long emb = (long)g.n[0] * g.n[1] * g.n[2];
  double acc[kMaxE];
  for (int r = 0; r < g.n_lines; ++r) acc[r] = 0.0;
  for (int s = 0; s < 100; ++s) {
    const double tl = sample_t(s) * lam;
    long row = 0;
    bool inside = true;
    long mult = 1;
    for (int a = 0; a < 3; ++a) {
      const double x = pt[a] + tl * d[a];
      const double f = floor((x - g.edges[a][0]) * g.inv[a]);
      int ia = 0;
      if (f >= 0.0 && f < (double)g.n[a]) {
        ia = (int)f;
      } else {
        inside = false;
      }
      row += ia * mult;
      mult *= g.n[a];
    }
    if (!inside) row = emb;
    for (int r = 0; r < g.n_lines; ++r) acc[r] -= mu[row * g.n_lines + r];
  }
  const long p = i + (long)g.X * (j + (long)g.Y * k);
  for (int r = 0; r < g.n_lines; ++r) {
    const double a = exp(acc[r] * dx);
    const int e = line_el[r];
    ac[r * g.V + p] = a;
    W[r * g.V + p] = a * (rho0[e * g.V + p] / atom_mass[e]);
  }
}

// noa_bar / atom mass at eee points ([e][V]) from the generation sums:
// noa_bar[e] = sum over the lines of element e of ac * gen (summed over beams).
__global__ void atoms_bar_kernel(AbsGeom g, const double* __restrict__ ac, const double* __restrict__ gen_b,
                                 long gen_stride, int n_beams, const double* __restrict__ atom_mass,
                                 const int* __restrict__ line_el, double* __restrict__ extra) {
#pragma clang fp contract(off)
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nq = (long)g.M[0] * g.M[1] * g.M[2];
  if (q >= nq) return;
  const int i = (int)(q % g.M[0]), j = (int)((q / g.M[0]) % g.M[1]), k = (int)(q / ((long)g.M[0] * g.M[1]));
  const long p = i + (long)g.X * (j + (long)g.Y * k);
  double nb[kMaxE];
  for (int e = 0; e < g.n_e; ++e) nb[e] = 0.0;
  for (int r = 0; r < g.n_lines; ++r) {
    double gs = 0.0;
    for (int b = 0; b < n_beams; ++b) gs += gen_b[b * gen_stride + r * g.V + p];
    nb[line_el[r]] += ac[r * g.V + p] * gs;
  }
  for (int e = 0; e < g.n_e; ++e) extra[e * g.V + p] = nb[e] / atom_mass[e];
}

// compute_absorption_coefficient_pullback (x_ray_intensities.jl:131-153), with
// its approximation: the march is traversed again with the factor fixed, and
// at every sample inside the boxes the MassFraction pullback of
// rho_bar = sum_r (-ac dx ac_bar)[r] mac[r] is added to that box (boxes found
// by search on the edges; a sample on the last edge belongs to the last box).
// ac_bar = noa[el(r)] * gen.  That pullback is linear in rho_bar for the
// box's fixed fractions, so this kernel adds only rho_bar, times the number
// of samples, into the box's row of box_rb (zeroed; n_e per box), and
// pullback_kernel applies the MassFraction pullback once per box.  The march
// then waits for no load of the box's fractions.  Samples from many points
// land in one box, so the additions are atomic (their rounding depends on
// their order).
__global__ void absorb_pullback_kernel(AbsGeom g, const double* __restrict__ mac, const double* __restrict__ ac,
                                       const double* __restrict__ rho0, const double* __restrict__ atom_mass,
                                       const int* __restrict__ line_el, const double* __restrict__ gen_b,
                                       long gen_stride, int n_beams, double* __restrict__ box_rb) {
#pragma clang fp contract(off)
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nq = (long)g.M[0] * g.M[1] * g.M[2];
  if (q >= nq) return;
  const int i = (int)(q % g.M[0]), j = (int)((q / g.M[0]) % g.M[1]), k = (int)(q / ((long)g.M[0] * g.M[1]));
  const long p = i + (long)g.X * (j + (long)g.Y * k);
  double pt[3], d[3], lam, dx;
  ray(g, i, j, k, pt, d, lam, dx);
  double rb[kMaxE];
  for (int e = 0; e < g.n_e; ++e) rb[e] = 0.0;
  for (int r = 0; r < g.n_lines; ++r) {
    double gs = 0.0;
    for (int b = 0; b < n_beams; ++b) gs += gen_b[b * gen_stride + r * g.V + p];
    const int el = line_el[r];
    const double ac_bar = (rho0[el * g.V + p] / atom_mass[el]) * gs;
    const double a_ = -ac[r * g.V + p] * dx * ac_bar;
    for (int e = 0; e < g.n_e; ++e) rb[e] += a_ * mac[r * g.n_e + e];
  }
  // Consecutive samples of a ray often fall in one box: each run of samples
  // in one box is added with one atomic (count times the value).
  long run_cell = -1;
  int run_len = 0;
  auto flush = [&]() {
    if (run_cell < 0) return;
    for (int e = 0; e < g.n_e; ++e) atomicAdd(&box_rb[run_cell * g.n_e + e], run_len * rb[e]);
  };
  for (int s = 0; s < 100; ++s) {
    const double tl = sample_t(s) * lam;
    long cell = 0, mult = 1;
    bool inside = true;
    for (int a = 0; a < 3; ++a) {
      const double x = pt[a] + tl * d[a];
      const double* __restrict__ ed = g.edges[a];
      const int n = g.n[a];
      int ia;
      if (x == ed[n]) {
        ia = n - 1;
      } else {
        // searchsorted(edges, x, side="right") - 1: the largest ia in [-1, n]
        // with edges[ia] <= x.  The floor on the uniform spacing starts within
        // one box of it, and the two corrections move to it on the edges
        // themselves (the largest such index for any non-decreasing edges,
        // NaN included: it gives -1), so the result is the search's while
        // the loads are independent instead of a chain of eight per axis.
        const double f = floor((x - ed[0]) * g.inv[a]);
        ia = f >= (double)n ? n : (f >= 0.0 ? (int)f : -1);
        while (ia < n && ed[ia + 1] <= x) ++ia;
        while (ia >= 0 && ed[ia] > x) --ia;
      }
      if (ia < 0 || ia > n - 1) inside = false;
      cell += (long)(ia < 0 ? 0 : ia) * mult;
      mult *= n;
    }
    if (!inside) cell = -1;
    if (cell != run_cell) {
      flush();
      run_cell = cell;
      run_len = 0;
    }
    run_len += 1;
  }
  flush();
}

}  // namespace

hipError_t launch_material_densities(const Layout& L, const int class_n[8][3], const long rho_off[8],
                                     const int ncls_of[8], const double* w, const double* rho_pure, double* rho,
                                     hipStream_t s) {
  if (L.n_e > kMaxE) return hipErrorInvalidValue;
  MatGeom g = make_mat_geom(L, class_n, rho_off, class_active_dims, ncls_of);
  long most = 1;
  for (int ci = 0; ci < g.n_cls; ++ci)
    most = std::max(most, (long)g.cls[ci].n[0] * g.cls[ci].n[1] * g.cls[ci].n[2]);
  const dim3 grid((unsigned)((most + 255) / 256), (unsigned)g.n_cls), block(256);
  hipLaunchKernelGGL(densities_kernel, grid, block, 0, s, g, w, rho_pure, rho);
  return hipGetLastError();
}

hipError_t launch_material_pullback(const Layout& L, const int class_n[8][3], const long rho_off[8],
                                    const int ncls_of[8], const double* w, const double* rho_pure,
                                    const double* rho_bar, long rb_stride, int n_beams, const double* extra_eee,
                                    bool box_rb_in_grad, double* grad, hipStream_t s, bool rb_cell) {
  if (L.n_e > kMaxE) return hipErrorInvalidValue;
  MatGeom g = make_mat_geom(L, class_n, rho_off, class_active_dims, ncls_of);
  const long n_cells = (long)L.Mx * L.My * L.Mz;
  hipLaunchKernelGGL(pullback_kernel, dim3((unsigned)((n_cells + 255) / 256)), dim3(256), 0, s, g, w, rho_pure,
                     rho_bar, rb_stride, n_beams, extra_eee, box_rb_in_grad ? 1 : 0, rb_cell ? 1 : 0, grad);
  return hipGetLastError();
}


hipError_t launch_absorption(const AbsorptionGeometry& a, const Layout& L, const double* w, const double* emb,
                             const double* rho_pure, const double* mac, const double* rho, const double* atom_mass,
                             const int* line_el, double* mu_scratch, double* ac, double* W, hipStream_t s) {
  if (L.n_e > kMaxE || L.n_lines > kMaxE) return hipErrorInvalidValue;
  AbsGeom g = abs_geom(a, L);
  const long rows = (long)g.n[0] * g.n[1] * g.n[2] + 1;
  hipLaunchKernelGGL(mu_table_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, g, w, emb, rho_pure, mac,
                     mu_scratch);
  const long nq = (long)L.Mx * L.My * L.Mz;
  hipLaunchKernelGGL(absorb_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, g, mu_scratch, rho, atom_mass,
                     line_el, ac, W);
  return hipGetLastError();
}

hipError_t launch_atoms_bar(const AbsorptionGeometry& a, const Layout& L, const double* ac, const double* gen_b,
                            long gen_stride, int n_beams, const double* atom_mass, const int* line_el, double* extra,
                            hipStream_t s) {
  AbsGeom g = abs_geom(a, L);
  const long nq = (long)L.Mx * L.My * L.Mz;
  hipLaunchKernelGGL(atoms_bar_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, g, ac, gen_b, gen_stride,
                     n_beams, atom_mass, line_el, extra);
  return hipGetLastError();
}

hipError_t launch_absorption_pullback(const AbsorptionGeometry& a, const Layout& L, const double* mac,
                                      const double* ac, const double* rho, const double* atom_mass,
                                      const int* line_el, const double* gen_b, long gen_stride, int n_beams,
                                      double* box_rb, hipStream_t s) {
  if (L.n_e > kMaxE || L.n_lines > kMaxE) return hipErrorInvalidValue;
  AbsGeom g = abs_geom(a, L);
  const long nq = (long)L.Mx * L.My * L.Mz;
  hipLaunchKernelGGL(absorb_pullback_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, g, mac, ac, rho,
                     atom_mass, line_el, gen_b, gen_stride, n_beams, box_rb);
  return hipGetLastError();
}

}  // namespace pnx
