// io.cpp -- ESRI-ASCII rasters, CSV series, GridConfig.
#include "io.hpp"

#include <cmath>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>

namespace rift {

Raster read_raster(const std::string& path, bool allow_nodata, double nodata_out)
{
    std::ifstream ifs(path.c_str());
    if (!ifs) throw IoError(path + ": error: cannot open");
    Raster r;
    std::string word;
    int line = 1;
    auto header = [&](auto& v) {
        if (!(ifs >> word >> v)) throw IoError(path + ": read error at line " + std::to_string(line));
        ++line;
    };
    header(r.ncols);
    header(r.nrows);
    header(r.xll);
    header(r.yll);
    header(r.cellsize);
    header(r.nodata);
    if (r.ncols <= 0 || r.nrows <= 0) throw IoError(path + ": bad raster size");
    r.data.assign((size_t)r.ncols * r.nrows, 0.0);
    // rows are stored north first; row 0 of the array is the south row
    for (int j = r.nrows - 1; j >= 0; --j, ++line) {
        for (int i = 0; i < r.ncols; ++i) {
            double v;
            if (!(ifs >> v)) throw IoError(path + ": read error at line " + std::to_string(line));
            if (v == r.nodata) {
                if (!allow_nodata) throw IoError(path + ": NODATA values present");
                v = nodata_out;
            }
            r.data[(size_t)j * r.ncols + i] = v;
        }
    }
    return r;
}

void GridConfig::from_dem(const Raster& dem)
{
    b_nx = dem.ncols;
    b_ny = dem.nrows;
    b_xll = dem.xll;
    b_yll = dem.yll;
    cellsize_original = dem.cellsize;
    h_xll = b_xll + cellsize_original / 2.0;
    h_yll = b_yll + cellsize_original / 2.0;
    compute_cell_size();
}

// Haversine distance in km (grid_config.cpp:27-46)
static double haversine_km(double long1, double lat1, double long2, double lat2)
{
    const double R = 6372.8;
    const double rad = M_PI / 180.0;
    const double la1 = lat1 * rad, la2 = lat2 * rad, dla = la2 - la1, dlo = (long2 - long1) * rad;
    const double c = std::asin(std::sqrt(std::sin(dla / 2) * std::sin(dla / 2) +
                                         std::cos(la1) * std::cos(la2) * std::sin(dlo / 2) *
                                             std::sin(dlo / 2)));
    return 2 * R * c;
}

void GridConfig::compute_cell_size()
{
    if (!projected) {
        cellsize = cellsize_original * 6378137.0 * M_PI / 180.0;
        if (square_cells) {
            h_dx = (double)cellsize_original * 6378137.f * (double)M_PI / 180.f;
            h_dy = h_dx;
        } else {
            // ComputeNonSquare (grid_config.cpp:48-64)
            const double w = b_nx * cellsize_original, h = b_ny * cellsize_original;
            const double xc = h_xll + w / 2.0, xmax = h_xll + w;
            const double yc = h_yll + h / 2.0, ymax = h_yll + h;
            h_dx = haversine_km(h_xll, yc, xmax, yc) / b_nx * 1000.0;
            h_dy = haversine_km(xc, h_yll, xc, ymax) / b_ny * 1000.0;
        }
    } else {
        cellsize = cellsize_original;
        h_dx = cellsize;
        h_dy = cellsize;
    }
    h_nx = b_nx + 4 - 1;
    h_ny = b_ny + 4 - 1;
}

void GridConfig::coord_to_index(double X, double Y, int& i, int& j) const
{
    i = (int)((X - h_xll) / cellsize_original + 0.5);
    j = (int)((Y - h_yll) / cellsize_original + 0.5);
}

void GridConfig::write(std::ostream& o) const
{
    auto w = [&](const auto& v) { o.write(reinterpret_cast<const char*>(&v), sizeof(v)); };
    w(b_nx); w(b_ny); w(b_xll); w(b_yll); w(cellsize_original);
    w(h_nx); w(h_ny); w(h_xll); w(h_yll); w(cellsize); w(h_dx); w(h_dy);
    w(square_cells); w(projected);
}

void GridConfig::read(std::istream& in, bool& square, bool& proj)
{
    auto r = [&](auto& v) { in.read(reinterpret_cast<char*>(&v), sizeof(v)); };
    r(b_nx); r(b_ny); r(b_xll); r(b_yll); r(cellsize_original);
    r(h_nx); r(h_ny); r(h_xll); r(h_yll); r(cellsize); r(h_dx); r(h_dy);
    r(square); r(proj);
}

void write_grid(const std::string& path, const double* x, const GridConfig& gc)
{
    std::ofstream out(path.c_str());
    if (!out.is_open()) throw IoError(path + ": error: cannot open for writing");
    out.precision(std::numeric_limits<double>::digits10);
    out << "ncols         " << gc.h_nx - 4 << std::endl;
    out << "nrows         " << gc.h_ny - 4 << std::endl;
    out << "xllcorner     " << gc.h_xll << std::endl;
    out << "yllcorner     " << gc.h_yll << std::endl;
    out << "cellsize      " << gc.cellsize_original << std::endl;
    out << "NODATA_value  " << GridConfig::kNodata << std::endl;
    for (int j = gc.h_ny - 3; j >= 2; --j) {
        for (int i = 2; i < gc.h_nx - 2; ++i) out << x[(size_t)j * gc.h_nx + i] << " ";
        out << std::endl;
    }
}

Series read_series(const std::string& path)
{
    std::ifstream file(path.c_str());
    if (!file.is_open()) throw IoError(path + ": error: cannot open");
    Series s;
    std::string line;
    while (std::getline(file, line)) {
        if (!line.empty() && line[0] == '#') continue;
        std::stringstream ls(line);
        std::string field;
        int count = 0;
        float value = 0.f;
        while (std::getline(ls, field, ',')) {
            std::stringstream(field) >> value;
            if (count == 0) s.time.push_back(value * 60 * 60);
            else s.rate.push_back(value);
            ++count;
        }
    }
    if (s.rate.empty()) throw IoError(path + ": no data");
    if (s.rate.back() != 0.f) {
        float est = 0.f;
        for (size_t i = 1; i < s.time.size(); ++i) est += s.time[i] - s.time[i - 1];
        est /= (float)s.time.size() - 1;
        s.time.push_back(s.time.back() + est);
        s.rate.push_back(0.f);
    }
    return s;
}

}  // namespace rift
