// geometry.cpp -- internal-grid construction of Simulator::InitSimulation
// (simulator.cu:267-436), the kappa formula (simulator.cu:281) and the
// corner-to-cell rule of the gridded forcing series (GridSeries.cu:91-122).
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/rift_host.h"

extern "C" double rift_kappa(double dx, double dy)
{
    return (double)std::sqrt(0.01f * std::fmax(1.f, std::fmin((float)dx, (float)dy)));
}

extern "C" int rift_geometry(int b_nx, int b_ny, const double* dem, const double* h0c,
                             const double* nc, double n_const, const double* Kc, double* BX,
                             double* BY, double* h0, double* n, double* K)
{
    if (b_nx < 3 || b_ny < 3 || !dem || !BX || !BY || !n) return RIFT_E_ARG;
    const int nx = b_nx + 3, ny = b_ny + 3;   // grid_config.cpp:84-85
    std::memset(BX, 0, sizeof(double) * (size_t)(nx + 1) * ny);
    std::memset(BY, 0, sizeof(double) * (size_t)nx * (ny + 1));
    std::memset(n, 0, sizeof(double) * (size_t)nx * ny);
    if (h0) std::memset(h0, 0, sizeof(double) * (size_t)nx * ny);
    if (K) std::memset(K, 0, sizeof(double) * (size_t)nx * ny);
    auto bx = [&](int i, int j) -> double& { return BX[(size_t)j * (nx + 1) + i]; };
    auto by = [&](int i, int j) -> double& { return BY[(size_t)j * nx + i]; };
    for (int j = 2; j < ny - 2; ++j) {
        for (int i = 2; i < nx - 2; ++i) {
            const int jt = j - 2, it = i - 2;
            const size_t ll = (size_t)jt * b_nx + it, ul = (size_t)(jt + 1) * b_nx + it;
            const size_t ur = (size_t)(jt + 1) * b_nx + it + 1, lr = (size_t)jt * b_nx + it + 1;
            bx(i + 1, j) = 0.5f * (dem[ur] + dem[lr]);
            bx(i, j) = 0.5f * (dem[ll] + dem[ul]);
            by(i, j + 1) = 0.5f * (dem[ul] + dem[ur]);
            by(i, j) = 0.5f * (dem[ll] + dem[lr]);
            // ghost faces repeat the nearest interior faces (simulator.cu:404-414)
            if (i == 2) {
                bx(i - 1, j) = bx(i + 1, j);
                bx(i - 2, j) = bx(i, j);
            }
            if (i == nx - 3) bx(i + 2, j) = bx(i, j);
            if (j == 2) {
                by(i, j - 1) = by(i, j + 1);
                by(i, j - 2) = by(i, j);
            }
            if (j == ny - 3) by(i, j + 2) = by(i, j);
            const size_t c = (size_t)j * nx + i;
            if (h0 && h0c) h0[c] = 0.25f * (h0c[ur] + h0c[lr] + h0c[ll] + h0c[ul]);
            n[c] = nc ? 0.25f * (nc[ur] + nc[lr] + nc[ll] + nc[ul]) : n_const;
            if (K && Kc) {
                double k = 0.25f * (Kc[ur] + Kc[lr] + Kc[ll] + Kc[ul]);
                K[c] = k / (3600.f * 1000.f);
            }
        }
    }
    return RIFT_OK;
}

// Corners from the face means: walked from D(0, 0) = 0 along the first row
// (BY: D(i+1) = 2 BY - D(i)) and up every column (BX), then the
// checkerboard's mean is taken out (a DEM has almost none, so the corners
// come out close to its points and round to FP32 well).  Every face is then
// checked against the corners' mean, and the ghost faces against the faces
// two cells in (simulator.cu:404-414).
extern "C" int rift_corners_from_faces(int nx, int ny, const double* BX, const double* BY, double* D)
{
    if (nx < 6 || ny < 6 || !BX || !BY || !D) return RIFT_E_ARG;
    const int W = nx - 3, H = ny - 3;   // corner columns / rows
    std::vector<double> c((size_t)W * H, 0.0);
    auto bx = [&](int i, int j) { return BX[(size_t)j * (nx + 1) + i]; };
    auto by = [&](int i, int j) { return BY[(size_t)j * nx + i]; };
    auto at = [&](int ci, int cj) -> double& { return c[(size_t)cj * W + ci]; };
    for (int ci = 0; ci + 1 < W; ++ci) at(ci + 1, 0) = 2.0 * by(ci + 2, 2) - at(ci, 0);
    for (int cj = 0; cj + 1 < H; ++cj)
        for (int ci = 0; ci < W; ++ci) at(ci, cj + 1) = 2.0 * bx(ci + 2, cj + 2) - at(ci, cj);
    double cb = 0.0;
    for (int cj = 0; cj < H; ++cj)
        for (int ci = 0; ci < W; ++ci) cb += ((ci + cj) & 1) ? -at(ci, cj) : at(ci, cj);
    cb /= (double)W * (double)H;
    for (int cj = 0; cj < H; ++cj)
        for (int ci = 0; ci < W; ++ci) at(ci, cj) -= ((ci + cj) & 1) ? -cb : cb;
    auto near = [](double a, double b) { return std::fabs(a - b) <= 1e-9 * (1.0 + std::fabs(b)); };
    bool ok = true;
    for (int j = 2; j <= ny - 3 && ok; ++j) {
        for (int i = 2; i <= nx - 2 && ok; ++i) ok = near(0.5 * (at(i - 2, j - 2) + at(i - 2, j - 1)), bx(i, j));
        ok = ok && bx(1, j) == bx(3, j) && bx(0, j) == bx(2, j) && bx(nx - 1, j) == bx(nx - 3, j);
    }
    for (int j = 2; j <= ny - 2 && ok; ++j)
        for (int i = 2; i <= nx - 3 && ok; ++i) ok = near(0.5 * (at(i - 2, j - 2) + at(i - 1, j - 2)), by(i, j));
    for (int i = 2; i <= nx - 3 && ok; ++i)
        ok = by(i, 1) == by(i, 3) && by(i, 0) == by(i, 2) && by(i, ny - 1) == by(i, ny - 3);
    if (!ok) return RIFT_E_ARG;
    std::memcpy(D, c.data(), c.size() * sizeof(double));
    return RIFT_OK;
}

// GridSeries::p_interp (GridSeries.cu:91-122): a cell takes the mean of its
// four corner values when more than two of them hold data, times the series
// scale; every other cell (ghosts included) holds the fill value, nodata
// when the series allows it, else 0.
extern "C" int rift_map_cells(int b_nx, int b_ny, const double* corners, double scale,
                              int allow_nodata, double nodata, double* cells)
{
    if (b_nx < 2 || b_ny < 2 || !corners || !cells) return RIFT_E_ARG;
    const int nx = b_nx + 3, ny = b_ny + 3;
    const double fill = allow_nodata ? nodata : 0.0;
    for (size_t k = 0; k < (size_t)nx * ny; ++k) cells[k] = fill;
    for (int j = 2; j < ny - 2; ++j) {
        for (int i = 2; i < nx - 2; ++i) {
            const int jt = j - 2, it = i - 2;
            int nnd = 0;
            double vsum = 0.0;
            for (int ni = 0; ni < 2; ++ni)
                for (int nj = 0; nj < 2; ++nj) {
                    const double v = corners[(size_t)(jt + nj) * b_nx + it + ni];
                    if (v != nodata) {
                        ++nnd;
                        vsum += v;
                    }
                }
            if (nnd > 2) {
                double v = vsum / (double)nnd;
                v *= scale;
                cells[(size_t)j * nx + i] = v;
            }
        }
    }
    return RIFT_OK;
}
