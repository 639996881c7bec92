// rift_main.cpp -- command-line driver: `rift <config.cfg>` runs the
// simulation of the keyword file to its final time (main.cc:40-67).
#include <cstdio>

#include "../../include/rift_host.h"

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <config file>\n", argv[0]);
        return 2;
    }
    rift_sim* s = nullptr;
    if (rift_sim_open(argv[1], 0.0, &s) != RIFT_OK) {
        std::fprintf(stderr, "%s\n", rift_last_error());
        return 1;
    }
    double t, dt, tf;
    rift_sim_time(s, &t, &dt, &tf);
    std::printf("This will run for %g seconds\n", tf);
    if (rift_sim_run(s, -1.0) != RIFT_OK) {
        std::fprintf(stderr, "%s\n", rift_last_error());
        rift_sim_close(s);
        return 1;
    }
    if (rift_sim_close(s) != RIFT_OK) {
        std::fprintf(stderr, "%s\n", rift_last_error());
        return 1;
    }
    return 0;
}
