// ref_host_driver.cpp -- a C-ABI shim over the reference's own host sources
// that compile without CUDA: Source (source.cc:11-73) and ConfigFile
// (config.h:48-209).  Built into oracle/_ref/ by oracle/Makefile from the
// reference tree in place; used only by tests to pin this repository's C++
// readers against the reference's behaviour.  Test infrastructure only.
#include <complex>
#include <cstring>
#include <string>

#include "config.h"
#include "source.h"

extern "C" {

// Reads a CSV series with Source::ReadSource; copies up to cap points.
// Returns the number of points (after the reference's zero-append rule).
int ref_source_read(const char* path, float* time_s, float* rate, int cap)
{
    Source s;
    s.ReadSource(path);
    int n = (int)s.time.size();
    for (int k = 0; k < n && k < cap; ++k) {
        time_s[k] = s.time[k];
        rate[k] = s.rate[k];
    }
    return n;
}

// Replays Source::InterpolateRate over `steps` times, starting from an
// interpolated_rate of `initial` and dividing by `unit` after each call, as
// Simulator::UpdateSource does (simulator.cu:504-536).
int ref_source_replay(const char* path, const double* t, int steps, float unit, float* out)
{
    Source s;
    s.ReadSource(path);
    for (int k = 0; k < steps; ++k) {
        s.InterpolateRate((float)t[k]);
        s.interpolated_rate /= unit;
        out[k] = s.interpolated_rate;
    }
    return 0;
}

// ConfigFile accessors.  kind: 0 = key exists, 1 = keyIsNumber,
// 2 = double value, 3 = bool value, 4 = complex value (re, im).
int ref_config_query(const char* path, const char* key, int kind, double* out)
{
    ConfigFile cfg(path);
    std::string k(key);
    switch (kind) {
    case 0: out[0] = cfg.keyExists(k) ? 1.0 : 0.0; return 0;
    case 1: out[0] = cfg.keyIsNumber(k) ? 1.0 : 0.0; return 0;
    case 2: out[0] = cfg.getValueOfKey<double>(k); return 0;
    case 3: out[0] = cfg.getValueOfKey<bool>(k) ? 1.0 : 0.0; return 0;
    case 4: {
        std::complex<double> c = cfg.getValueOfKey<std::complex<double> >(k);
        out[0] = c.real();
        out[1] = c.imag();
        return 0;
    }
    default: return -1;
    }
}

// String value of a key (getValueOfKey<std::string>), NUL-terminated.
int ref_config_string(const char* path, const char* key, char* out, int cap)
{
    ConfigFile cfg(path);
    std::string v = cfg.getValueOfKey<std::string>(key);
    std::strncpy(out, v.c_str(), (size_t)cap - 1);
    out[cap - 1] = '\0';
    return (int)v.size();
}

}  // extern "C"
