#pragma once
// Solve options of the pdhg engine slot.  Same fields, order, defaults and
// meaning as the reference's igaos::Options (include/igaos/options.h:5-13):
// the pdhg engine reads time_limit_s, tolerance, max_iterations (only when
// time_limit_s <= 0 or >= 1e6, src/pdhg/pdhg.cu:669-671) and verbosity.

namespace igaos {

struct Options {
    double time_limit_s = 120.0;
    double tolerance = 1e-4;
    long max_iterations = 500000;
    int seed = 0;
    bool presolve = true;
    bool restarts = false;
    int verbosity = 0;
};

}  // namespace igaos
