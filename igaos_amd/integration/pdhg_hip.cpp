// Drop-in replacement for the reference's src/pdhg/pdhg.cu: implements
// igaos::pdhg::solve (src/pdhg/pdhg.hpp:9) on the MI355X engine through its
// C-ABI (include/igaos_pdhg.h).  It is compiled inside the reference's
// build, against the reference's own headers (io::Model, Options,
// Solution, Status), and linked with libigaos_pdhg.so; nothing else in the
// reference changes (solve_with_engine, the CLI and the Python bindings call
// pdhg::solve as before).
#include <string>
#include <vector>

#include "igaos_pdhg.h"
#include "pdhg.hpp"

namespace igaos::pdhg {

Solution solve(const io::Model& md, const Options& o) {
    Solution sol;
    igaos_model* h = igaos_model_from_csr_csc(
        md.m, md.n, md.c.data(), md.cl.data(), md.cu.data(), md.rmin.data(),
        md.rmax.data(), md.ap.data(), md.ai.data(), md.ax.data(), md.cp.data(),
        md.ci.data(), md.acx.data(), md.obj_const);
    if (!h) {
        sol.status = Status::Error;
        sol.message = igaos_last_error();
        return sol;
    }
    igaos_options opt;
    igaos_options_default(&opt);
    opt.time_limit_s = o.time_limit_s;
    opt.tolerance = o.tolerance;
    opt.max_iterations = o.max_iterations;
    opt.seed = o.seed;
    opt.presolve = o.presolve ? 1 : 0;
    opt.restarts = o.restarts ? 1 : 0;
    opt.verbosity = o.verbosity;
    igaos_solve_info info;
    std::vector<char> msg(1024, '\0');
    sol.x.assign(md.n, 0.0);
    sol.y.assign(md.m, 0.0);
    // cfg NULL: the plain FP64 engine
    const int rc = igaos_pdhg_solve(h, &opt, nullptr, sol.x.data(), sol.y.data(), &info,
                                    msg.data(), static_cast<int>(msg.size()));
    if (rc != 0) {
        igaos_model_free(h);
        sol.status = Status::Error;
        sol.message = igaos_last_error();
        return sol;
    }
    sol.status = static_cast<Status>(info.status);  // same enumerator order
    if (sol.status != Status::Error) {  // A x, as the reference simplex fills it
        sol.row_activity.assign(md.m, 0.0);
        igaos_model_row_activity(h, sol.x.data(), sol.row_activity.data());
    } else {  // no point to report, as in a build without the engine
        sol.x.clear();
        sol.y.clear();
    }
    igaos_model_free(h);
    sol.objective = info.objective;
    sol.pinf = info.pinf;
    sol.dinf = info.dinf;
    sol.rel_gap = info.rel_gap;
    sol.iterations = info.iterations;
    sol.solve_time_ms = info.solve_time_ms;
    sol.message = msg.data();
    return sol;
}

}  // namespace igaos::pdhg
