// Host-side geometry: csim grids, Armadillo-style linear interpolation tables,
// hex-cell FPM layout and rasterisation.
#pragma once
#include <cstdint>
#include <vector>

#include "ops.hpp"

namespace csx {

// arrayGeom::set_xy (src/lib/csim_geom.cpp:113-140): x_i = -(n d - d)/2 + i d - origin
std::vector<double> grid_symmetric(int n, double d, double origin = 0.0);
// (i - n/2) d  (src/elements/doubleFraunhoferFPM.cpp:59-64)
std::vector<double> grid_offset(int n, double d);

// Linear-interpolation table for interp1(xg -> xi, "*linear", extrap 0) as
// Armadillo 8.500.1 evaluates it (nearest node, neighbour across, weight by
// distances); lo = -1 marks extrapolated points.
std::vector<InterpTable> interp_table(const std::vector<double>& xg, const std::vector<double>& xi);

// complexHexMaskFPM hex centres: layout 0 = csim radial cut-off
// (complexHexMaskFPM.cpp:236-315), 1 = standard rings 1 + 3n(n+1).
int hex_centers(int layout, int n_rings, double hex_step, std::vector<double>& hx, std::vector<double>& hy);
// Lattice radius that maps to half the design array (maskNRings for csim's layout).
double hex_layout_radius(int layout, int n_rings, double hex_step);

struct HexRaster {
  int S = 0, nsubpix = 0, nhex = 0;
  std::vector<int32_t> nsub, hexnum;   // S x S x 3, column-major
};
// make_hex_array + make_interpolation_data (complexHexMaskFPM.cpp:317-469).
int hex_rasterize(int layout, int n_rings, double hex_step, int fpm_array_size, double fpm_scale_factor,
                  double hex_gap, int fpm_mask_size, int n_sub_pix, double pixel_scale, double fpm_radius,
                  HexRaster& out);

}  // namespace csx
