// Host-side geometry of the propagation path.
#include "hostgeom.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

#include "common.hpp"

namespace csx {

std::vector<double> grid_symmetric(int n, double d, double origin) {
  std::vector<double> x(n);
  double phys = d * n;
  for (int i = 0; i < n; i++) x[i] = -(phys - d) / 2.0 + i * d - origin;
  return x;
}

std::vector<double> grid_offset(int n, double d) {
  std::vector<double> x(n);
  for (int i = 0; i < n; i++) x[i] = ((double)i - ((double)n) / 2.) * d;
  return x;
}

std::vector<InterpTable> interp_table(const std::vector<double>& xg, const std::vector<double>& xi) {
  std::vector<InterpTable> t(xi.size());
  const int ng = (int)xg.size();
  double lo_v = *std::min_element(xg.begin(), xg.end());
  double hi_v = *std::max_element(xg.begin(), xg.end());
  int a = 0;
  for (size_t i = 0; i < xi.size(); i++) {
    double v = xi[i];
    if (v < lo_v || v > hi_v) { t[i] = InterpTable{-1, -1, 0.0}; continue; }
    // nearest node, scanning forward from the previous one (xi ascending)
    double aerr = std::numeric_limits<double>::infinity();
    for (int j = a; j < ng; j++) {
      double e = std::fabs(xg[j] - v);
      if (e >= aerr) break;
      aerr = e;
      a = j;
    }
    int b = (xg[a] - v <= 0) ? std::min(a + 1, ng - 1) : std::max(a - 1, 0);
    double berr = std::fabs(xg[b] - v);
    int lo = a, hi = b;
    double le = aerr, he = berr;
    if (lo > hi) { std::swap(lo, hi); std::swap(le, he); }
    t[i] = InterpTable{lo, hi, le > 0 ? le / (le + he) : 0.0};
  }
  return t;
}

int hex_centers(int layout, int n_rings, double step, std::vector<double>& hx, std::vector<double>& hy) {
  hx.clear(); hy.clear();
  if (layout == 0) {
    // lattice rows (3i, sqrt3 j) and the half-offset row; keep (int)|c| < n_rings
    struct Cand { double x, y; int ring; };
    std::vector<Cand> cands;
    int imax = n_rings / 3 + 2;
    int jmax = (int)(n_rings / std::sqrt(3.0) + 2);
    for (int i = -imax; i < imax; i++)
      for (int j = -jmax; j < jmax; j++)
        for (int odd = 0; odd < 2; odd++) {
          double x = step * i * 3.0, y = step * j * std::sqrt(3.0);
          if (odd) { x += step * 1.5; y += step * std::sqrt(3.0) / 2.0; }
          cands.push_back({x, y, (int)std::sqrt(x * x + y * y)});
        }
    // numbered ring by ring, in generation order within a ring
    for (int r = 0; r < n_rings; r++)
      for (const Cand& c : cands)
        if (c.ring == r) { hx.push_back(c.x); hy.push_back(c.y); }
  } else {
    // standard rings: p a1 + q a2, a1 = (1.5, sqrt3/2), a2 = (0, sqrt3); ring = max(|p|,|q|,|p+q|)
    for (int r = 0; r <= n_rings; r++) {
      std::vector<std::pair<double, std::pair<double, double>>> ringv;
      for (int p = -r; p <= r; p++)
        for (int q = -r; q <= r; q++) {
          int d = std::max(std::abs(p), std::max(std::abs(q), std::abs(p + q)));
          if (d != r) continue;
          double x = step * 1.5 * p, y = step * (std::sqrt(3.0) / 2.0 * p + std::sqrt(3.0) * q);
          double th = std::atan2(y, x);
          if (th < 0) th += 2 * M_PI;
          ringv.push_back({r == 0 ? 0.0 : th, {x, y}});
        }
      std::stable_sort(ringv.begin(), ringv.end(),
                       [](const auto& a, const auto& b) { return a.first < b.first; });
      for (auto& e : ringv) { hx.push_back(e.second.first); hy.push_back(e.second.second); }
    }
  }
  return (int)hx.size();
}

double hex_layout_radius(int layout, int n_rings, double step) {
  if (layout == 0) return n_rings;
  // outermost centre at n_rings * sqrt3 * step plus the circumradius (= step)
  return n_rings * std::sqrt(3.0) * step + step;
}

int hex_rasterize(int layout, int n_rings, double step, int A, double scale, double gap, int S, int nsp,
                  double pixel_scale, double fpm_radius, HexRaster& out) {
  if (A <= 0 || S <= 0 || nsp <= 0 || n_rings < 0 || pixel_scale <= 0 || fpm_radius <= 0)
    return fail(CSX_ERR_ARG, "hex_rasterize: bad geometry");
  std::vector<double> hx, hy;
  int nhex = hex_centers(layout, n_rings, step, hx, hy);
  // design array: hex index per design pixel, -1 outside (make_hex_array)
  double R = hex_layout_radius(layout, n_rings, step);
  double stepPix = (0.5 * A) / R * scale;
  double inner = stepPix * (1.0 - gap) * (std::sqrt(3.0) / 2.0);
  double outer = 2 * inner / std::sqrt(3.0);
  std::vector<int32_t> design((size_t)A * A, -1);
  const double vx = std::sqrt(3) / 2, vy = 0.5;
  for (int h = 0; h < nhex; h++) {
    double xc = 0.5 * A + hx[h] * stepPix, yc = 0.5 * A + hy[h] * stepPix;
    int c0 = std::max(0, (int)std::floor(xc - outer) - 1), c1 = std::min(A - 1, (int)std::ceil(xc + outer) + 1);
    int r0 = std::max(0, (int)std::floor(yc - outer) - 1), r1 = std::min(A - 1, (int)std::ceil(yc + outer) + 1);
    for (int jj = c0; jj <= c1; jj++)
      for (int ii = r0; ii <= r1; ii++) {
        double dx = (double)jj - xc, dy = (double)ii - yc;
        if (dx * dx + dy * dy > outer * outer) continue;
        double px = dx * vx, py = dy * vy;
        if (std::fabs(dy) < inner && std::fabs(px + py) < inner && std::fabs(px - py) < inner)
          design[(size_t)jj * A + ii] = h;
      }
  }
  // sub-pixel sampling (make_interpolation_data): slot 0 also counts the
  // off-design / off-hex samples, as in the reference.
  out.S = S; out.nsubpix = nsp; out.nhex = nhex;
  size_t plane = (size_t)S * S;
  out.nsub.assign(3 * plane, 0);
  out.hexnum.assign(3 * plane, -1);
  const double ac1 = -S / 2.0 - 0.5, xc1 = (0.5 / fpm_radius) * scale;
  auto design_index = [&](int pix, int sub) {
    double x = (pix + ac1 + (0.5 + sub) / ((double)nsp)) * pixel_scale;
    return (int)((0.5 + x * xc1) * A + 0.5);
  };
  for (int j = 0; j < S; j++) {
    for (int i = 0; i < S; i++) {
      int mid = (int)(nsp / 2.0);
      int ci = design_index(i, mid), cj = design_index(j, mid);
      if (ci < 0 || ci >= A || cj < 0 || cj >= A) continue;
      size_t p = (size_t)j * S + i;
      int list[3], nl = 0;
      for (int a = 0; a < nsp; a++) {
        int i1 = design_index(i, a);
        for (int b = 0; b < nsp; b++) {
          int j1 = design_index(j, b);
          if (i1 < 0 || i1 >= A || j1 < 0 || j1 >= A) { out.nsub[p]++; continue; }
          int h = design[(size_t)j1 * A + i1];
          if (h < 0) { out.nsub[p]++; continue; }
          int slot = -1;
          for (int q = 0; q < nl; q++) if (list[q] == h) slot = q;
          if (slot < 0) {
            if (nl == 3) continue;
            list[nl] = h; slot = nl++;
          }
          out.nsub[p + slot * plane]++;
          out.hexnum[p + slot * plane] = h;
        }
      }
    }
  }
  for (size_t p = 0; p < plane; p++)
    if (out.nsub[p] + out.nsub[p + plane] + out.nsub[p + 2 * plane] == 0) out.nsub[p] = nsp * nsp;
  return CSX_OK;
}

}  // namespace csx
