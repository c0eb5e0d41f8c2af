#include "pnx_host.hpp"

#include <algorithm>

namespace pnx {

static bool order_supported(int N) { return N == 1 || N == 3 || N == 5 || N == 7 || N == 9; }

double axis_point(const pnx_desc& d, int a, int odd, int i) {
  const int M = d.M[a];
  if (M == 1) return d.lo[a];
  const double h = (d.hi[a] - d.lo[a]) / (M - 1);
  const double lo = odd ? d.lo[a] - h / 2 : d.lo[a], hi = odd ? d.hi[a] + h / 2 : d.hi[a];
  const int n = odd ? M + 1 : M;
  if (i == n - 1) return hi;
  return i * ((hi - lo) / (n - 1)) + lo;
}

int make_geometry(const pnx_desc& d, Geometry& g, std::string& err) {
  if (!order_supported(d.N)) {
    err = "PN order must be one of 1, 3, 5, 7, 9";
    return PNX_ERR_UNSUPPORTED;
  }
  if (d.n_elements < 1 || d.n_elements > 8) {
    err = "n_elements must be 1..8";
    return PNX_ERR_ARG;
  }
  if (d.n_lines < 1 || d.n_lines > 8) {
    err = "n_lines must be 1..8";
    return PNX_ERR_ARG;
  }
  if (d.n_steps < 1 || d.n_beams < 1) {
    err = "n_steps and n_beams must be positive";
    return PNX_ERR_ARG;
  }
  if (!(d.E_initial > d.E_cutoff)) {
    err = "E_initial must exceed E_cutoff";
    return PNX_ERR_ARG;
  }
  int dims = 0;
  for (int a = 0; a < 3; ++a) {
    if (d.M[a] < 1) {
      err = "grid sizes must be >= 1";
      return PNX_ERR_ARG;
    }
    if (d.M[a] == 1) continue;
    if (a != dims) {
      err = "non-singleton axes must come first (x, then y, then z)";
      return PNX_ERR_ARG;
    }
    if (!(d.hi[a] > d.lo[a])) {
      err = "axis extent must be positive";
      return PNX_ERR_ARG;
    }
    ++dims;
  }
  if (dims == 0) {
    err = "at least the x axis must be resolved";
    return PNX_ERR_ARG;
  }
  g.T = make_tables(d.N);
  Layout& L = g.L;
  L.Mx = d.M[0];
  L.My = d.M[1];
  L.Mz = d.M[2];
  L.dims = dims;
  L.N = d.N;
  L.n_e = d.n_elements;
  L.n_lines = d.n_lines;
  // x pitch padded to 16 doubles (128 B cache lines) once rows are long enough
  const int xneed = L.Mx + 1;
  L.X = xneed >= 64 ? (xneed + 15) / 16 * 16 : xneed;
  L.Y = L.My == 1 ? 1 : L.My + 1;
  L.Z = L.Mz == 1 ? 1 : L.Mz + 1;
  L.V = (long)L.X * L.Y * L.Z;
  long off = 0, ro = 0;
  for (int c = 0; c < 8; ++c) {
    const bool act = class_active(c, dims) && g.T.n[c] > 0;
    L.cls_off[c] = act ? off : 0;
    g.rho_off[c] = act ? ro : 0;
    if (act) {
      off += g.T.n[c] * L.V;
      ro += (long)d.n_elements * L.V;
    }
  }
  L.beam_stride = off;
  g.rho_total = ro;
  double h[3];
  for (int a = 0; a < 3; ++a) {
    h[a] = d.M[a] == 1 ? 1.0 : (d.hi[a] - d.lo[a]) / (d.M[a] - 1);
    L.inv_h[a] = 1.0 / h[a];
    for (int c = 0; c < 8; ++c) g.class_n[c][a] = d.M[a] == 1 ? 1 : (odd_in(c, a) ? d.M[a] + 1 : d.M[a]);
  }
  L.dvol = h[0] * h[1] * h[2];
  L.n_cells = L.Mx * L.My * L.Mz;
  {
    const int tx = tile_x(dims), ty = kBlock / tx;
    L.tiles_x = (L.Mx + tx - 1) / tx;
    L.tiles_y = (L.My + ty - 1) / ty;
    L.n_tiles = L.tiles_x * L.tiles_y * L.Mz;
    L.cw_tiles_x = (L.Mx + kCwTileX - 1) / kCwTileX;
    L.n_cw_items = dims >= 2 ? 4 * L.cw_tiles_x * ((L.My + kCwTileX - 1) / kCwTileX) * L.Mz : 0;
  }
  g.rec_len = d.n_elements + d.n_elements * (d.N + 1) + 1 + d.n_lines;
  L.rec_len = g.rec_len;
  g.sigma_off = d.n_elements + d.n_elements * (d.N + 1) + 1;
  g.dE = (d.E_initial - d.E_cutoff) / d.n_steps;
  g.K = d.ckpt_every > 0 ? std::min(d.ckpt_every, d.n_steps) : d.n_steps;
  g.nck = (d.n_steps + g.K - 1) / g.K;
  g.prof_stride = 8L * L.Y * L.Z;
  g.partial_step = (long)L.n_tiles * kWavesPerBlock * d.n_lines;
  g.partial_beam = (long)(d.n_steps + 1) * g.partial_step;
  return PNX_OK;
}

void pack_class(const Geometry& g, int c, const double* src, int vec, double* dst) {
  const Layout& L = g.L;
  const int nx = g.class_n[c][0], ny = g.class_n[c][1], nz = g.class_n[c][2];
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long q = i + (long)nx * (j + (long)ny * k);
        const long pd = i + (long)L.X * (j + (long)L.Y * k);
        for (int e = 0; e < vec; ++e) dst[e * L.V + pd] = src[q * vec + e];
      }
}

void unpack_class(const Geometry& g, int c, const double* src, int vec, double* dst) {
  const Layout& L = g.L;
  const int nx = g.class_n[c][0], ny = g.class_n[c][1], nz = g.class_n[c][2];
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long q = i + (long)nx * (j + (long)ny * k);
        const long pd = i + (long)L.X * (j + (long)L.Y * k);
        for (int e = 0; e < vec; ++e) dst[q * vec + e] = src[e * L.V + pd];
      }
}

void pack_lines(const Geometry& g, const double* src, double* dst) { pack_class(g, 0, src, g.L.n_lines, dst); }
void unpack_lines(const Geometry& g, const double* src, double* dst) { unpack_class(g, 0, src, g.L.n_lines, dst); }

void pack_profiles(const Geometry& g, const double* const* ppos, double* dst) {
  const Layout& L = g.L;
  for (int c = 0; c < 8; ++c) {
    if (!odd_in(c, 0) || !g.active(c) || !ppos[c]) continue;
    const int ny = g.class_n[c][1], nz = g.class_n[c][2];
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j) dst[c * (long)L.Y * L.Z + j + (long)L.Y * k] = ppos[c][j + ny * k];
  }
}

}  // namespace pnx
