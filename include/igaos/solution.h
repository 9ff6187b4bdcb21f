#pragma once
// Solve result, same fields as the reference's include/igaos/solution.h:10-22.
// The pdhg engine fills status, message, objective, pinf, dinf, rel_gap,
// iterations, solve_time_ms (setup excluded), x (n) and y (m, sign
// convention d = c + A'y), and row_activity = A x (m) computed as the
// reference simplex computes it (the reference's pdhg leaves it empty).

#include <string>
#include <vector>

#include "igaos/status.h"

namespace igaos {

struct Solution {
    Status status = Status::Error;
    double objective = 0.0;
    std::vector<double> x;
    std::vector<double> y;
    std::vector<double> row_activity;
    double pinf = 0.0;
    double dinf = 0.0;
    double rel_gap = 0.0;
    long iterations = 0;
    double solve_time_ms = 0.0;
    std::string message;
};

}  // namespace igaos
