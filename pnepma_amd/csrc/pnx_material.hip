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
__global__ void pullback_kernel(MatGeom g, const double* __restrict__ w, const double* __restrict__ rho_pure,
                                const double* __restrict__ rho_bar, long rb_stride, int n_beams,
                                const double* __restrict__ extra_eee, double* __restrict__ grad) {
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
  for (int ci = 0; ci < g.n_cls; ++ci) {
    const ClassGeom& c = g.cls[ci];
    const int cc[3] = {cx, cy, cz};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = cc[a];
      hi[a] = (c.odd[a] && cc[a] == g.M[a] - 1) ? g.M[a] : cc[a];  // odd axis: the last box also has point M
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
            for (int b = 0; b < n_beams; ++b) s += rho_bar[b * rb_stride + c.off + e * g.V + p];
            if (g.cls_id[ci] == 0 && extra_eee) s += extra_eee[e * g.V + p];
            rb[e] = s;
            dot += wc[e] * s;
          }
          const double R_bar = -dot / (R * R);
          for (int e = 0; e < ne; ++e) acc[e] += rb[e] / R + R_bar / rho_pure[e];
        }
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
                                    double* grad, hipStream_t s) {
  if (L.n_e > kMaxE) return hipErrorInvalidValue;
  MatGeom g = make_mat_geom(L, class_n, rho_off, class_active_dims, ncls_of);
  const long n_cells = (long)L.Mx * L.My * L.Mz;
  hipLaunchKernelGGL(pullback_kernel, dim3((unsigned)((n_cells + 255) / 256)), dim3(256), 0, s, g, w, rho_pure,
                     rho_bar, rb_stride, n_beams, extra_eee, grad);
  return hipGetLastError();
}

}  // namespace pnx
