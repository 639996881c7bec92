// simulator.cpp -- RIFT's driver (simulator.cu) over the device step.
#include "simulator.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "../../include/rift_host.h"

namespace rift {

// ---------------------------------------------------------------------------
// MapSeries (GridSeries.cu)
// ---------------------------------------------------------------------------
MapSeries::MapSeries(rift_grid* g, int which, std::string prefix, double scale, int deltat,
                     double tmax, const GridConfig& gc, bool interpolated, bool allow_nodata)
    : g_(g), which_(which), prefix_(std::move(prefix)), scale_(scale), deltat_(deltat), tmax_(tmax),
      gc_(gc), interpolated_(interpolated), allow_nodata_(allow_nodata),
      ibuf_((size_t)gc.h_nx * gc.h_ny, 0.0)
{
    if (deltat_ <= 0) throw IoError(prefix_ + ": map series time step must be a positive integer");
    // the surge series lives on the device: two slots, maps loaded as
    // p_update reads them
    device_ = interpolated_ && which_ == RIFT_MAP_SURGE;
    if (device_) {
        if (rift_map_series_create(g_, which_, 2, deltat_, tmax_) != RIFT_OK)
            throw IoError(std::string("map series: ") + rift_last_error());
        return;
    }
    // p_init_dev: planes start at the fill value (nodata when allowed)
    std::vector<double> fill(ibuf_.size(), allow_nodata_ ? GridConfig::kNodata : 0.0);
    for (int plane = 0; plane < 3; ++plane)
        if (interpolated_ || plane == 0) rift_set_map(g_, which_, plane, fill.data());
}

// p_read_grid + p_interp (GridSeries.cu:91-139): four-corner mean where more
// than two corners hold data, times the series scale.
void MapSeries::read_map(int index)
{
    std::ostringstream name;
    name << prefix_ << "-" << std::setw(3) << std::setfill('0') << index << ".txt";
    Raster r = read_raster(name.str(), allow_nodata_, GridConfig::kNodata);
    if (r.ncols != gc_.b_nx || r.nrows != gc_.b_ny)
        throw IoError(name.str() + ": incorrect size");
    rift_map_cells(gc_.b_nx, gc_.b_ny, r.data.data(), scale_, allow_nodata_ ? 1 : 0, GridConfig::kNodata,
                   ibuf_.data());
    std::cout << "Reading from " << name.str() << " ..." << std::endl;
}

// Map `index` into its slot of the device series.
void MapSeries::load_device(int index)
{
    read_map(index);
    if (rift_map_series_load(g_, which_, index, ibuf_.data()) != RIFT_OK)
        throw IoError(std::string("map series load: ") + rift_last_error());
}

void MapSeries::upload(int plane)
{
    if (rift_set_map(g_, which_, plane, ibuf_.data()) != RIFT_OK)
        throw IoError(std::string("map upload: ") + rift_last_error());
}

// GridSeries::p_update (GridSeries.cu:144-178)
void MapSeries::base_update(double t)
{
    bool sendit = false;
    if (in_time_ < 0.0) {
        in_time_ = t;
        if (in_time_ < tmax_) {
            read_map((int)std::trunc(in_time_ / deltat_));
            sendit = true;
        }
    }
    if (t >= tmax_) {
        if (!done_) {
            std::fill(ibuf_.begin(), ibuf_.end(), 0.0);
            sendit = true;
        }
        done_ = true;
    } else if (t >= in_time_ + deltat_) {
        in_time_ += deltat_;
        read_map((int)std::trunc(in_time_ / deltat_));
        sendit = true;
    }
    if (sendit) upload(0);
}

// HyetographGridSeries::p_update (GridSeries.cu:261-283)
void MapSeries::held_update(double t)
{
    if (in_time_ < 0.0) base_update(t);
    if (t >= tmax_) {
        base_update(t);
    } else if (t >= in_time_) {
        in_time_ += deltat_;
        read_map((int)std::trunc(in_time_ / deltat_));
        upload(0);
        double s = 0.0;
        rift_sum_interior(g_, which_ == RIFT_MAP_RAIN ? RIFT_RAINMAP : RIFT_DRAINMAP, &s);
        sum_cache_ = s;
    }
}

// InterpolatedGridSeries::p_update (GridSeries.cu:325-384) with the device
// series: the device advances the window and interpolates at every step; the
// host keeps the same window only to read each map when p_update would,
// into the slot of the two-slot series that the window has just left.
void MapSeries::device_update(double t)
{
    if (in_time_ < 0.0) {
        in_time_ = t;
        const int index = (int)std::trunc(in_time_ / deltat_);
        load_device(index);
        in_time_ = index * (double)deltat_;
        next_time_ = in_time_ + deltat_;
        load_device(index + 1);
        if (rift_map_series_start(g_, which_, t) != RIFT_OK)
            throw IoError(std::string("map series start: ") + rift_last_error());
    } else if (t >= tmax_) {
        if (!done_) {
            in_time_ = next_time_;
            next_time_ = in_time_ + deltat_;
        }
        done_ = true;
    } else if (t >= next_time_) {
        in_time_ = next_time_;
        next_time_ = in_time_ + deltat_;
        load_device((int)std::trunc(next_time_ / deltat_));
    }
}

// InterpolatedGridSeries::p_update (GridSeries.cu:325-384)
void MapSeries::interp_update(double t)
{
    if (in_time_ < 0.0) {
        in_time_ = t;
        const int index = (int)std::trunc(in_time_ / deltat_);
        read_map(index);
        upload(1);
        in_time_ = index * (double)deltat_;
        next_time_ = in_time_ + deltat_;
        read_map(index + 1);
        upload(2);
    } else if (t >= tmax_) {
        if (!done_) {
            in_time_ = next_time_;
            next_time_ = in_time_ + deltat_;
            rift_copy_map(g_, which_, 1, 2);
        }
        done_ = true;
    } else if (t >= next_time_) {
        rift_copy_map(g_, which_, 1, 2);
        in_time_ = next_time_;
        next_time_ = in_time_ + deltat_;
        read_map((int)std::trunc(next_time_ / deltat_));
        upload(2);
    }
    const double factor = (t - in_time_) / (next_time_ - in_time_);
    rift_interp_map(g_, which_, factor);
}

void MapSeries::update(double t)
{
    if (device_) device_update(t);
    else if (interpolated_) interp_update(t);
    else held_update(t);
}

// ---------------------------------------------------------------------------
// Simulator
// ---------------------------------------------------------------------------
Simulator::~Simulator()
{
    if (grid_) rift_destroy(grid_);
}

void Simulator::check(int rc, const char* what) const
{
    if (rc != RIFT_OK) throw std::runtime_error(std::string(what) + ": " + rift_last_error());
}

// Simulator::ReadUserParams (simulator.cu:25-265): keywords and defaults.
void Simulator::ReadUserParams(const std::string& path)
{
    Config cfg(path);
    dem_file_ = cfg.get<std::string>("DEM");
    t0 = cfg.get<double>("t0");
    tf = cfg.get<double>("tf");
    t = t0;
    dt = (double)0.0000001f;
    euler_ = cfg.has("euler_integration") ? cfg.get<bool>("euler_integration") : false;
    output_dir = cfg.has("output") ? cfg.get<std::string>("output") : std::string("./");
    n_gridded_ = false;
    n_const_ = 0.0;
    if (cfg.has("n")) {
        if (cfg.is_number("n")) n_const_ = cfg.get<double>("n");
        else { n_gridded_ = true; n_file_ = cfg.get<std::string>("n"); }
    }
    courant_ = 0.25;
    if (cfg.has("courant_limit")) {
        courant_ = cfg.get<double>("courant_limit");
        if (courant_ <= 0 || courant_ >= 1.0) {
            std::cerr << "Specifed Courant_limit (" << courant_ << ") is out of range, reverting to default"
                      << std::endl;
            courant_ = 0.25;
        }
    }
    infiltration_ = cfg.has("K");
    if (infiltration_) {
        K_file_ = cfg.get<std::string>("K");
        psi_ = cfg.get<double>("psi");
        dtheta_ = cfg.get<double>("dtheta");
    }
    h_init_ = cfg.has("h0");
    if (h_init_) h_file_ = cfg.get<std::string>("h0");
    dambreak_ = cfg.has("hydrograph");
    if (dambreak_) {
        hydro_file_ = cfg.get<std::string>("hydrograph");
        if (cfg.has("source_rect_X")) {
            auto X = cfg.get<std::complex<double>>("source_rect_X");
            auto Y = cfg.get<std::complex<double>>("source_rect_Y");
            source_X0_ = X.real(); source_X1_ = X.imag();
            source_Y0_ = Y.real(); source_Y1_ = Y.imag();
        } else if (cfg.has("source_pt")) {
            auto p = cfg.get<std::complex<double>>("source_pt");
            source_X0_ = source_X1_ = p.real();
            source_Y0_ = source_Y1_ = p.imag();
        } else {
            source_X0_ = source_X1_ = cfg.get<double>("source_X");
            source_Y0_ = source_Y1_ = cfg.get<double>("source_Y");
        }
    }
    rain_avg_ = cfg.has("hyetograph");
    if (rain_avg_) hyeto_file_ = cfg.get<std::string>("hyetograph");
    rain_grid_ = cfg.has("hyetograph_prefix");
    if (rain_grid_) {
        hyeto_prefix_ = cfg.get<std::string>("hyetograph_prefix");
        rain_interp_ = cfg.has("hyetograph_interp") ? cfg.get<bool>("hyetograph_interp") : false;
        hyeto_dt_ = cfg.get<double>("hyetograph_dt");
        hyeto_tf_ = cfg.get<double>("hyetograph_tf");
    }
    drain_avg_ = cfg.has("drain");
    if (drain_avg_) drain_file_ = cfg.get<std::string>("drain");
    drain_grid_ = cfg.has("drain_prefix");
    if (drain_grid_) {
        drain_prefix_ = cfg.get<std::string>("drain_prefix");
        drain_dt_ = cfg.get<double>("drain_dt");
        drain_tf_ = cfg.get<double>("drain_tf");
    }
    surge_grid_ = cfg.has("surge_prefix");
    if (surge_grid_) {
        surge_prefix_ = cfg.get<std::string>("surge_prefix");
        surge_dt_ = cfg.get<double>("surge_dt");
        surge_tf_ = cfg.get<double>("surge_tf");
    }
    auto flag = [&](const char* k) { return cfg.has(k) ? cfg.get<bool>(k) : false; };
    h_print_ = flag("h_print");
    ws_print_ = flag("ws_print");
    save_max_ = flag("save_max");
    save_arrival_ = flag("save_arrival");
    q_print_ = flag("q_print");
    check_volume_ = flag("check_volume");
    save_hotstart_ = flag("save_hotstart");
    read_hotstart_ = cfg.has("read_hotstart") ? cfg.get<std::string>("read_hotstart") : std::string();
    if (h_print_ || q_print_ || ws_print_ || check_volume_ || save_hotstart_) {
        t_print_ = t0;
        dt_print_ = cfg.get<double>("dt_print");
        count_print = 0;
    }
    b_print_ = cfg.has("b_print");
    n_print_ = cfg.has("n_print");
    device_ = cfg.has("device") ? cfg.get<int>("device") : 0;
    gc_.square_cells = cfg.has("square_cells") ? cfg.get<bool>("square_cells") : true;
    gc_.projected = cfg.has("projected") ? cfg.get<bool>("projected") : false;
    tile_acceleration_ = cfg.has("tile_acceleration") ? cfg.get<bool>("tile_acceleration") : true;
    V_added_ = 0.0;
}

// InitializeSources (simulator.cu:858-889): the rectangle of cells between
// the two coordinate corners, i outer and j inner.
void Simulator::InitializeSources()
{
    int imin, imax, jmin, jmax;
    gc_.coord_to_index(source_X0_, source_Y0_, imin, jmin);
    gc_.coord_to_index(source_X1_, source_Y1_, imax, jmax);
    source_idx_.clear();
    for (int i = imin; i <= imax; ++i)
        for (int j = jmin; j <= jmax; ++j) source_idx_.push_back(j * gc_.h_nx + i);
}

// InitSimulation (simulator.cu:267-493)
void Simulator::InitSimulation()
{
    Raster dem = read_raster(dem_file_);
    gc_.from_dem(dem);
    std::cout << "Computed Cell Sizes: " << gc_.h_dx << ", " << gc_.h_dy << std::endl;
    kappa_ = rift_kappa(gc_.h_dx, gc_.h_dy);
    auto corner = [&](const std::string& f) {
        Raster r = read_raster(f);
        if (r.ncols != gc_.b_nx || r.nrows != gc_.b_ny)
            throw IoError(f + ": incorrect size");
        return r.data;
    };
    std::vector<double> h0c, nc, Kc;
    if (h_init_) h0c = corner(h_file_);
    if (n_gridded_) nc = corner(n_file_);
    if (infiltration_) Kc = corner(K_file_);
    if (dambreak_) InitializeSources();

    unsigned flags = 0;
    if (dambreak_) flags |= RIFT_HYDROGRAPH;
    if (rain_avg_) flags |= RIFT_RAIN_AVG;
    if (rain_grid_) flags |= RIFT_RAIN_GRID;
    if (drain_avg_) flags |= RIFT_DRAIN_AVG;
    if (drain_grid_) flags |= RIFT_DRAIN_GRID;
    if (infiltration_) flags |= RIFT_INFILTRATION;
    if (surge_grid_) flags |= RIFT_SURGE_GRID;
    if (euler_) flags |= RIFT_EULER;
    if (save_max_) flags |= RIFT_SAVE_MAX;
    if (save_arrival_) flags |= RIFT_SAVE_ARRIVAL;
    if (tile_acceleration_) flags |= RIFT_TILE_ACCEL;
    if (h_init_) flags |= RIFT_H_INIT;
    if (rain_interp_) flags |= RIFT_RAIN_INTERP;
    if (check_volume_) flags |= RIFT_MASS_BUDGET;
    rift_desc d{};
    d.nx = gc_.h_nx;
    d.ny = gc_.h_ny;
    d.dx = gc_.h_dx;
    d.dy = gc_.h_dy;
    d.kappa = kappa_;
    d.courant = courant_;
    d.psi = psi_;
    d.dtheta = dtheta_;
    d.nodata = GridConfig::kNodata;
    d.flags = flags;
    d.device = device_;
    check(rift_create(&d, &grid_), "rift_create");

    if (rain_avg_) { Series s = read_series(hyeto_file_); check(rift_set_series(grid_, RIFT_SER_HYETOGRAPH, s.time.data(), s.rate.data(), (int)s.time.size()), "hyetograph"); }
    if (drain_avg_) { Series s = read_series(drain_file_); check(rift_set_series(grid_, RIFT_SER_DRAIN, s.time.data(), s.rate.data(), (int)s.time.size()), "drain"); }
    if (dambreak_) {
        Series s = read_series(hydro_file_);
        check(rift_set_series(grid_, RIFT_SER_HYDROGRAPH, s.time.data(), s.rate.data(), (int)s.time.size()), "hydrograph");
        check(rift_set_source_cells(grid_, source_idx_.data(), (long)source_idx_.size()), "source cells");
    }
    // map series; the series time steps are integers (GridSeries.cuh:29)
    const double rain_scale = 1.0 / 3600.0 / 1000.0;
    if (rain_grid_) {
        hyeto_series_.reset(new MapSeries(grid_, RIFT_MAP_RAIN, hyeto_prefix_, rain_scale, (int)hyeto_dt_,
                                          hyeto_tf_, gc_, rain_interp_, false));
        hyeto_series_->update(t0);
    }
    if (drain_grid_) {
        drain_series_.reset(new MapSeries(grid_, RIFT_MAP_DRAIN, drain_prefix_, rain_scale, (int)drain_dt_,
                                          drain_tf_, gc_, false, false));
        drain_series_->update(t0);
    }

    const size_t N = (size_t)gc_.h_nx * gc_.h_ny;
    std::vector<double> BX((size_t)(gc_.h_nx + 1) * gc_.h_ny), BY((size_t)gc_.h_nx * (gc_.h_ny + 1));
    std::vector<double> h0(N), n(N), K(N);
    check(rift_geometry(gc_.b_nx, gc_.b_ny, dem.data.data(), h_init_ ? h0c.data() : nullptr,
                        n_gridded_ ? nc.data() : nullptr, n_const_, infiltration_ ? Kc.data() : nullptr,
                        BX.data(), BY.data(), h0.data(), n.data(), K.data()),
          "rift_geometry");
    check(rift_upload_static(grid_, BX.data(), BY.data(), n.data(), infiltration_ ? K.data() : nullptr,
                             h_init_ ? h0.data() : nullptr),
          "rift_upload_static");
    check(rift_init(grid_, t0), "rift_init");
    // the surge series starts its window on the initialised device state
    if (surge_grid_) {
        surge_series_.reset(new MapSeries(grid_, RIFT_MAP_SURGE, surge_prefix_, 1.0, (int)surge_dt_,
                                          surge_tf_, gc_, true, true));
        surge_series_->update(t0);
    }
    bc_.assign(N, 0.0);
    check(rift_download(grid_, RIFT_BC, bc_.data()), "download BC");
    if (b_print_) write_grid(output_dir + "/bc.txt", bc_.data(), gc_);
    if (n_print_) write_grid(output_dir + "/n.txt", n.data(), gc_);
}

void Simulator::OpenSimulation(const std::string& config_file, double)
{
    ReadUserParams(config_file);
    InitSimulation();
    if (!read_hotstart_.empty()) ReadHotstart(read_hotstart_);
    check(rift_get_time(grid_, &t, &dt), "rift_get_time");
}

double Simulator::RunSimulation()
{
    check(rift_step(grid_, 1), "rift_step");
    check(rift_get_time(grid_, &t, &dt), "rift_get_time");
    if (hyeto_series_) {
        hyeto_series_->update(t);
        V_added_ += hyeto_series_->sum() * dt;
    }
    if (drain_series_) drain_series_->update(t);
    if (surge_series_) surge_series_->update(t);
    if (h_print_ || q_print_ || ws_print_ || check_volume_) {
        if (t > t_print_) {
            PrintData();
            t_print_ += dt_print_;
            std::cout << t << "\t" << dt << std::endl;
        }
    }
    return t;
}

void Simulator::setSourceValue(double value) { check(rift_set_source_value(grid_, value), "source value"); }

// PrintData (simulator.cu:568-648)
void Simulator::PrintData()
{
    const size_t N = (size_t)gc_.h_nx * gc_.h_ny;
    std::vector<double> h(N);
    check(rift_download(grid_, RIFT_DEPTH, h.data()), "download depth");
    if (save_max_) PrintSummaryData();
    std::ostringstream name;
    if (h_print_) write_grid(output_dir + "/h" + std::to_string(count_print) + ".txt", h.data(), gc_);
    if (ws_print_) {
        std::vector<double> ws(N);
        for (size_t k = 0; k < N; ++k) ws[k] = h[k] > 0.0 ? h[k] + bc_[k] : GridConfig::kNodata;
        write_grid(output_dir + "/ws" + std::to_string(count_print) + ".txt", ws.data(), gc_);
    }
    if (save_hotstart_) WriteHotstart(output_dir + "/hot" + std::to_string(count_print) + ".bin");
    if (q_print_) {
        std::vector<double> q(N);
        check(rift_download(grid_, RIFT_Q, q.data()), "download q");
        write_grid(output_dir + "/q" + std::to_string(count_print) + ".txt", q.data(), gc_);
    }
    if (check_volume_) {
        // the FP64 device budget (m^3) and its relative closure
        double b[9];
        check(rift_budget(grid_, b), "volume budget");
        const double scale = std::max({std::fabs(b[0]) + std::fabs(b[1]), std::fabs(b[2]), 1e-300});
        std::cout << "volume stored " << b[0] << " infiltrated " << b[1] << " rain " << b[3]
                  << " source " << b[4] << " drain " << b[5] << " outflow " << b[6] << " exchange "
                  << b[7] << " residual " << b[8] << " (" << b[8] / scale << ")" << std::endl;
    }
    ++count_print;
}

// PrintSummaryData (simulator.cu:650-697)
void Simulator::PrintSummaryData()
{
    const size_t N = (size_t)gc_.h_nx * gc_.h_ny;
    std::vector<double> a(N);
    if (save_max_) {
        const std::pair<int, const char*> maps[] = {{RIFT_HMAX, "peak_flood_depth.txt"},
                                                    {RIFT_QMAX, "peak_unit_flow.txt"},
                                                    {RIFT_TPEAK, "time_to_peak.txt"},
                                                    {RIFT_TDRY, "time_to_dry.txt"}};
        for (auto& m : maps) {
            check(rift_download(grid_, m.first, a.data()), m.second);
            write_grid(output_dir + "/" + m.second, a.data(), gc_);
        }
    }
    if (save_arrival_) {
        check(rift_download(grid_, RIFT_TWET, a.data()), "arrival");
        write_grid(output_dir + "/flood_wave_arrival.txt", a.data(), gc_);
    }
}

void Simulator::CloseSimulation()
{
    if (save_max_) PrintSummaryData();
    if (grid_) rift_destroy(grid_);
    grid_ = nullptr;
}

// Tile arrays of the hotstart: GrowGridSize entries (grid.cu:472-489).
static size_t grow_grid_size(int nbx, int nby)
{
    const size_t gx = (size_t)std::ceil((double)nbx / 16.0), gy = (size_t)std::ceil((double)nby / 14.0);
    return (gx * 16) * (gy * 14);
}

// WriteHotstart (simulator.cu:912-968): t, dt, GridConfig, w, hu, hv, [F],
// wet_blocks, active_blocks.
void Simulator::WriteHotstart(const std::string& path)
{
    std::ofstream o(path, std::ios::binary);
    if (!o.good()) throw IoError("Cannot open open hotstart file \"" + path + "\" for writing");
    check(rift_get_time(grid_, &t, &dt), "rift_get_time");
    o.write(reinterpret_cast<const char*>(&t), sizeof t);
    o.write(reinterpret_cast<const char*>(&dt), sizeof dt);
    gc_.write(o);
    const size_t N = (size_t)gc_.h_nx * gc_.h_ny;
    std::vector<double> buf(N);
    std::vector<int> fields = {RIFT_W, RIFT_HU, RIFT_HV};
    if (infiltration_) fields.push_back(RIFT_F);
    for (int f : fields) {
        check(rift_download(grid_, f, buf.data()), "hotstart download");
        o.write(reinterpret_cast<const char*>(buf.data()), N * sizeof(double));
    }
    int nbx, nby;
    check(rift_get_tile_dims(grid_, &nbx, &nby), "tile dims");
    const size_t nt = (size_t)nbx * nby, gs = grow_grid_size(nbx, nby);
    std::vector<unsigned char> wet(nt), act(nt);
    int nact = 0;
    check(rift_get_tiles(grid_, wet.data(), act.data(), &nact), "tiles");
    // tile-indexed: active tiles hold their index, inactive ones its negative
    // (the Grow_k encoding, grid.cu:1216-1220); entries past the tile count
    // are padding and read as inactive
    std::vector<char> wb(gs, 0);
    std::vector<int> ab(gs);
    for (size_t k = 0; k < gs; ++k) {
        wb[k] = (k < nt && wet[k]) ? 1 : 0;
        ab[k] = (k < nt && act[k]) ? (int)k : -(int)k;
    }
    o.write(wb.data(), (std::streamsize)gs);
    o.write(reinterpret_cast<const char*>(ab.data()), (std::streamsize)(gs * sizeof(int)));
}

// ReadHotstart (simulator.cu:974-1090)
void Simulator::ReadHotstart(const std::string& path)
{
    std::ifstream in(path, std::ios::binary | std::ios::in);
    if (!in.good()) throw IoError("Cannot open open hotstart file \"" + path + "\" for reading");
    double tt, dtt;
    in.read(reinterpret_cast<char*>(&tt), sizeof tt);
    in.read(reinterpret_cast<char*>(&dtt), sizeof dtt);
    GridConfig h;
    bool square, projected;
    h.read(in, square, projected);
    if (h.b_nx != gc_.b_nx || h.b_ny != gc_.b_ny || h.h_dx != gc_.h_dx || h.h_dy != gc_.h_dy ||
        h.cellsize_original != gc_.cellsize_original || h.cellsize != gc_.cellsize ||
        square != gc_.square_cells || projected != gc_.projected)
        throw IoError("Invalid hotstart");
    const size_t N = (size_t)gc_.h_nx * gc_.h_ny;
    std::vector<double> buf(N);
    std::vector<int> fields = {RIFT_W, RIFT_HU, RIFT_HV};
    if (infiltration_) fields.push_back(RIFT_F);
    for (int f : fields) {
        in.read(reinterpret_cast<char*>(buf.data()), N * sizeof(double));
        if (!in) throw IoError("Invalid hotstart: truncated");
        check(rift_upload(grid_, f, buf.data()), "hotstart upload");
    }
    int nbx, nby;
    check(rift_get_tile_dims(grid_, &nbx, &nby), "tile dims");
    const size_t nt = (size_t)nbx * nby, gs = grow_grid_size(nbx, nby);
    std::vector<char> wb(gs);
    std::vector<int> ab(gs);
    in.read(wb.data(), (std::streamsize)gs);
    in.read(reinterpret_cast<char*>(ab.data()), (std::streamsize)(gs * sizeof(int)));
    if (!in) throw IoError("Invalid hotstart: truncated tile arrays");
    std::vector<unsigned char> wet(nt), act(nt, 0);
    for (size_t k = 0; k < nt; ++k) {
        wet[k] = wb[k] ? 1 : 0;
        const int v = ab[k];
        if (v >= 0 && (size_t)v < nt) act[(size_t)v] = 1;
    }
    check(rift_set_tiles(grid_, wet.data(), act.data()), "tiles");
    check(rift_set_time(grid_, tt, dtt), "time");
    check(rift_apply_boundaries(grid_), "boundaries");
    t = tt;
    dt = dtt;
    std::cout << "Using hotstart \"" << path << "\" created at simulation time " << t << std::endl;
}

}  // namespace rift
