// host_capi.cpp -- C ABI over the host C++: keyword configuration, file
// readers and writers, and rift::Simulator (include/rift_host.h).
#include <algorithm>
#include <complex>
#include <cstring>
#include <exception>
#include <string>

#include "../../include/rift_host.h"
#include "simulator.hpp"

namespace rift {
int set_last_error(int code, const std::string& msg);
}

struct rift_sim {
    rift::Simulator sim;
};

namespace {

template <class F>
int guarded(const char* what, F&& f)
{
    try {
        f();
        return RIFT_OK;
    } catch (const rift::ConfigError& e) {
        return rift::set_last_error(RIFT_E_ARG, std::string(what) + ": " + e.what());
    } catch (const rift::IoError& e) {
        return rift::set_last_error(RIFT_E_ARG, std::string(what) + ": " + e.what());
    } catch (const std::exception& e) {
        return rift::set_last_error(RIFT_E_STATE, std::string(what) + ": " + e.what());
    }
}

}  // namespace

extern "C" int rift_sim_open(const char* config_path, double flow_rate, rift_sim** out)
{
    if (!config_path || !out) return rift::set_last_error(RIFT_E_ARG, "rift_sim_open: null argument");
    *out = nullptr;
    rift_sim* s = new rift_sim;
    int rc = guarded("rift_sim_open", [&] { s->sim.OpenSimulation(config_path, flow_rate); });
    if (rc != RIFT_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return RIFT_OK;
}

extern "C" int rift_sim_step(rift_sim* s, double* t)
{
    if (!s) return rift::set_last_error(RIFT_E_ARG, "rift_sim_step: null handle");
    return guarded("rift_sim_step", [&] {
        const double tn = s->sim.RunSimulation();
        if (t) *t = tn;
    });
}

extern "C" int rift_sim_run(rift_sim* s, double t_end)
{
    if (!s) return rift::set_last_error(RIFT_E_ARG, "rift_sim_run: null handle");
    const double te = t_end < 0 ? s->sim.tf : t_end;
    return guarded("rift_sim_run", [&] {
        while (s->sim.t < te && s->sim.steady_state) s->sim.RunSimulation();
    });
}

extern "C" int rift_sim_time(rift_sim* s, double* t, double* dt, double* tf)
{
    if (!s) return rift::set_last_error(RIFT_E_ARG, "rift_sim_time: null handle");
    if (t) *t = s->sim.t;
    if (dt) *dt = s->sim.dt;
    if (tf) *tf = s->sim.tf;
    return RIFT_OK;
}

extern "C" rift_grid* rift_sim_grid(rift_sim* s) { return s ? s->sim.grid() : nullptr; }

extern "C" int rift_sim_set_source_value(rift_sim* s, double discharge)
{
    if (!s) return rift::set_last_error(RIFT_E_ARG, "rift_sim_set_source_value: null handle");
    return guarded("rift_sim_set_source_value", [&] { s->sim.setSourceValue(discharge); });
}

extern "C" int rift_sim_write_hotstart(rift_sim* s, const char* path)
{
    if (!s || !path) return rift::set_last_error(RIFT_E_ARG, "rift_sim_write_hotstart: null argument");
    return guarded("rift_sim_write_hotstart", [&] { s->sim.WriteHotstart(path); });
}

extern "C" int rift_sim_read_hotstart(rift_sim* s, const char* path)
{
    if (!s || !path) return rift::set_last_error(RIFT_E_ARG, "rift_sim_read_hotstart: null argument");
    return guarded("rift_sim_read_hotstart", [&] { s->sim.ReadHotstart(path); });
}

extern "C" int rift_sim_close(rift_sim* s)
{
    if (!s) return rift::set_last_error(RIFT_E_ARG, "rift_sim_close: null handle");
    int rc = guarded("rift_sim_close", [&] { s->sim.CloseSimulation(); });
    delete s;
    return rc;
}

// ---------------------------------------------------------------------------
// Keyword configuration
// ---------------------------------------------------------------------------
struct rift_config {
    explicit rift_config(const std::string& p) : cfg(p) {}
    rift::Config cfg;
};

namespace {

// Typed read without the missing-key default: a missing key is an error.
template <class T>
int config_get(const rift_config* c, const char* key, const char* what, T& out)
{
    if (!c || !key) return rift::set_last_error(RIFT_E_ARG, std::string(what) + ": null argument");
    if (!c->cfg.has(key)) return rift::set_last_error(RIFT_E_ARG, std::string(what) + ": missing key '" + key + "'");
    return guarded(what, [&] { out = c->cfg.get<T>(key); });
}

}  // namespace

extern "C" int rift_config_open(const char* path, rift_config** out)
{
    if (!path || !out) return rift::set_last_error(RIFT_E_ARG, "rift_config_open: null argument");
    *out = nullptr;
    return guarded("rift_config_open", [&] { *out = new rift_config(path); });
}

extern "C" void rift_config_close(rift_config* c) { delete c; }

extern "C" int rift_config_has(const rift_config* c, const char* key)
{
    return (c && key && c->cfg.has(key)) ? 1 : 0;
}

extern "C" int rift_config_is_number(const rift_config* c, const char* key, int* out)
{
    if (!c || !key || !out) return rift::set_last_error(RIFT_E_ARG, "rift_config_is_number: null argument");
    return guarded("rift_config_is_number", [&] { *out = c->cfg.is_number(key) ? 1 : 0; });
}

extern "C" int rift_config_string(const rift_config* c, const char* key, char* buf, int cap, int* len)
{
    std::string v;
    int rc = config_get(c, key, "rift_config_string", v);
    if (rc != RIFT_OK) return rc;
    if (len) *len = (int)v.size();
    if (buf && cap > 0) {
        const size_t n = std::min(v.size(), (size_t)cap - 1);
        std::memcpy(buf, v.data(), n);
        buf[n] = '\0';
    }
    return RIFT_OK;
}

extern "C" int rift_config_double(const rift_config* c, const char* key, double* out)
{
    if (!out) return rift::set_last_error(RIFT_E_ARG, "rift_config_double: null argument");
    return config_get(c, key, "rift_config_double", *out);
}

extern "C" int rift_config_bool(const rift_config* c, const char* key, int* out)
{
    if (!out) return rift::set_last_error(RIFT_E_ARG, "rift_config_bool: null argument");
    bool b = false;
    int rc = config_get(c, key, "rift_config_bool", b);
    if (rc == RIFT_OK) *out = b ? 1 : 0;
    return rc;
}

extern "C" int rift_config_complex(const rift_config* c, const char* key, double* re, double* im)
{
    if (!re || !im) return rift::set_last_error(RIFT_E_ARG, "rift_config_complex: null argument");
    std::complex<double> z;
    int rc = config_get(c, key, "rift_config_complex", z);
    if (rc == RIFT_OK) {
        *re = z.real();
        *im = z.imag();
    }
    return rc;
}

// ---------------------------------------------------------------------------
// File formats
// ---------------------------------------------------------------------------
extern "C" int rift_read_series(const char* path, float* time_s, float* rate, int cap, int* n)
{
    if (!path || !n || (cap > 0 && (!time_s || !rate)))
        return rift::set_last_error(RIFT_E_ARG, "rift_read_series: null argument");
    return guarded("rift_read_series", [&] {
        rift::Series s = rift::read_series(path);
        *n = (int)s.time.size();
        for (int k = 0; k < *n && k < cap; ++k) {
            time_s[k] = s.time[k];
            rate[k] = s.rate[k];
        }
    });
}

extern "C" int rift_read_raster(const char* path, int allow_nodata, int* ncols, int* nrows,
                                double* hdr4, double* data, long cap)
{
    if (!path || !ncols || !nrows) return rift::set_last_error(RIFT_E_ARG, "rift_read_raster: null argument");
    return guarded("rift_read_raster", [&] {
        rift::Raster r = rift::read_raster(path, allow_nodata != 0, rift::GridConfig::kNodata);
        *ncols = r.ncols;
        *nrows = r.nrows;
        if (hdr4) {
            hdr4[0] = r.xll;
            hdr4[1] = r.yll;
            hdr4[2] = r.cellsize;
            hdr4[3] = r.nodata;
        }
        if (data) {
            if ((long)r.data.size() > cap) throw rift::IoError(std::string(path) + ": buffer too small");
            std::memcpy(data, r.data.data(), r.data.size() * sizeof(double));
        }
    });
}

extern "C" int rift_write_grid(const char* path, const double* field, int nx, int ny, double h_xll,
                               double h_yll, double cellsize)
{
    if (!path || !field || nx < 5 || ny < 5)
        return rift::set_last_error(RIFT_E_ARG, "rift_write_grid: bad argument");
    return guarded("rift_write_grid", [&] {
        rift::GridConfig gc;
        gc.h_nx = nx;
        gc.h_ny = ny;
        gc.h_xll = h_xll;
        gc.h_yll = h_yll;
        gc.cellsize_original = cellsize;
        rift::write_grid(path, field, gc);
    });
}
