// io.hpp -- rasters, CSV series and grid geometry of the RIFT driver.
//
//  * ESRI-ASCII rasters (io.cc:28-361): six header lines, then rows north
//    first; stored with row 0 = south.  NODATA is an error unless allowed,
//    in which case it becomes GridConfig's nodata (-9999).
//  * Output rasters (io.cc:331-361): the interior of the internal grid,
//    header with 15 significant digits, rows north first, values followed by
//    a space.
//  * CSV series (source.cc:11-63): `hours,value` lines, `#` comments, float
//    arithmetic, a zero-rate point appended when the last rate is nonzero.
//  * GridConfig (grid_config.cpp:67-171): internal extent b+3, cell size in
//    metres or degrees, Coord2Index, and the 82-byte binary header of the
//    hotstart file.
#pragma once

#include <cstdint>
#include <iosfwd>
#include <stdexcept>
#include <string>
#include <vector>

namespace rift {

struct IoError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Raster {
    int ncols = 0, nrows = 0;
    double xll = 0, yll = 0, cellsize = 0, nodata = -9999.0;
    std::vector<double> data;   // nrows x ncols, row 0 = south
};

// Reads an ESRI-ASCII raster.  With allow_nodata, file NODATA values become
// `nodata_out`; without, a NODATA value is an error.
Raster read_raster(const std::string& path, bool allow_nodata = false,
                   double nodata_out = -9999.0);

struct GridConfig {
    static constexpr double kNodata = -9999.0;   // grid_config.cpp:14
    bool square_cells = true, projected = false;
    int b_nx = 0, b_ny = 0;
    double b_xll = 0, b_yll = 0, cellsize_original = 0;
    int h_nx = 0, h_ny = 0;
    double h_xll = 0, h_yll = 0, cellsize = 0, h_dx = 0, h_dy = 0;

    // from a DEM header (InitBathymetry, io.cc:42-72, grid_config.cpp:67-86)
    void from_dem(const Raster& dem);
    void compute_cell_size();
    // Coord2Index (grid_config.cpp:104-115); the reference's range check is
    // not used by its callers
    void coord_to_index(double X, double Y, int& i, int& j) const;
    void write(std::ostream& out) const;                      // 82 bytes
    void read(std::istream& in, bool& square, bool& proj);
};

// Writes the interior of an internal-grid field (h_nx x h_ny, row 0 = south).
void write_grid(const std::string& path, const double* x, const GridConfig& gc);

struct Series {
    std::vector<float> time;   // seconds
    std::vector<float> rate;
};

Series read_series(const std::string& path);

}  // namespace rift
