// simulator.hpp -- the RIFT driver over the MI355X step.
//
// Same public behaviour as the reference's Simulator (simulator.h:14-173):
// OpenSimulation reads the keyword file, rasters and series, builds the
// internal grid and initial state; RunSimulation advances one step and
// returns t; CloseSimulation writes the summary maps.  The step itself runs
// on the device through include/rift_kp.h; this class keeps the host work:
// configuration, file I/O, gridded forcing series, printing and hotstart.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/rift_kp.h"
#include "config.hpp"
#include "io.hpp"

namespace rift {

// A series of forcing maps `prefix-###.txt` (GridSeries.cu:33-404).  Held
// series (rain, drain) hold the map of the next map time on plane 0;
// interpolated series keep two planes and interpolate in time on the
// device: rain with hyetograph_interp through map planes 1 and 2 and one
// interpolation launch per step, surge through the device series (the
// device advances the window; the host loads each map into a slot).
class MapSeries {
public:
    MapSeries(rift_grid* g, int which, std::string prefix, double scale, int deltat, double tmax,
              const GridConfig& gc, bool interpolated, bool allow_nodata);
    void update(double t);
    double sum() const { return sum_cache_; }   // interior sum of the held map

private:
    void read_map(int index);                    // p_read_grid + p_interp
    void load_device(int index);                 // read_map into the device series
    void upload(int plane);
    void device_update(double t);
    void base_update(double t);
    void held_update(double t);
    void interp_update(double t);

    rift_grid* g_;
    int which_;
    std::string prefix_;
    double scale_;
    int deltat_;
    double tmax_;
    GridConfig gc_;
    bool interpolated_, allow_nodata_;
    bool device_ = false;                        // interpolated surge: device series
    double in_time_ = -9999.0, next_time_ = 0.0, sum_cache_ = 0.0;
    bool done_ = false;
    std::vector<double> ibuf_;                   // internal grid
};

class Simulator {
public:
    Simulator() = default;
    ~Simulator();
    Simulator(const Simulator&) = delete;
    Simulator& operator=(const Simulator&) = delete;

    void OpenSimulation(const std::string& config_file, double flow_rate = 0.0);
    double RunSimulation();
    void CloseSimulation();
    void setSourceValue(double discharge_m3s);
    void WriteHotstart(const std::string& path);
    void ReadHotstart(const std::string& path);

    double t = 0, tf = 0, dt = 0, t0 = 0;
    bool steady_state = true;
    rift_grid* grid() const { return grid_; }
    const GridConfig& grid_config() const { return gc_; }
    std::string output_dir = "./";
    int count_print = 0;

private:
    void ReadUserParams(const std::string& config_file);
    void InitSimulation();
    void InitializeSources();
    void PrintData();
    void PrintSummaryData();
    void check(int rc, const char* what) const;

    rift_grid* grid_ = nullptr;
    GridConfig gc_;
    std::string dem_file_, n_file_, K_file_, h_file_, hydro_file_, hyeto_file_, drain_file_;
    std::string hyeto_prefix_, drain_prefix_, surge_prefix_, read_hotstart_;
    bool euler_ = false, n_gridded_ = false, infiltration_ = false, h_init_ = false;
    bool dambreak_ = false, rain_avg_ = false, rain_grid_ = false, rain_interp_ = false;
    bool drain_avg_ = false, drain_grid_ = false, surge_grid_ = false;
    bool h_print_ = false, ws_print_ = false, q_print_ = false, save_max_ = false;
    bool save_arrival_ = false, check_volume_ = false, save_hotstart_ = false;
    bool b_print_ = false, n_print_ = false, tile_acceleration_ = true;
    double n_const_ = 0, psi_ = 0, dtheta_ = 0, courant_ = 0.25, kappa_ = 0;
    double hyeto_dt_ = 0, hyeto_tf_ = 0, drain_dt_ = 0, drain_tf_ = 0, surge_dt_ = 0, surge_tf_ = 0;
    double source_X0_ = 0, source_X1_ = 0, source_Y0_ = 0, source_Y1_ = 0;
    double t_print_ = 0, dt_print_ = 0, V_added_ = 0;
    int device_ = 0;
    std::vector<int> source_idx_;
    std::unique_ptr<MapSeries> hyeto_series_, drain_series_, surge_series_;
    std::vector<double> bc_;
};

}  // namespace rift
