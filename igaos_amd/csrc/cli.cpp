// `igaos info|solve <model.mps> [flags]` for the pdhg engine slot.
// Flags, JSON fields and number formats follow the reference CLI
// (src/api/igaos_cli.cpp:19-35, 45-90, 115-150); `--engine pdhg` runs the
// MI355X engine.  Engines outside the pdhg slot (simplex, milp, qp, auto)
// are not part of this build and report status "error".  Additions:
// `--precision 32|64` (FP32 + FP64 refinement, or plain FP64) and a
// "y" array in the JSON output.  `--seed` parses into an int (the
// reference reinterprets the int as a long).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>

#include "igaos/options.h"
#include "igaos/solution.h"
#include "igaos/status.h"
#include "model.hpp"
#include "pdhg.hpp"

namespace {

void usage() {
    std::printf(
        "igaos - MI355X PDHG engine\n\n"
        "USAGE:\n"
        "  igaos info   <model.mps>\n"
        "  igaos solve  <model.mps> [flags]\n\n"
        "FLAGS (solve):\n"
        "  --time-limit <sec>      wall-clock budget (default 120)\n"
        "  --tol <float>           termination tolerance (default 1e-4)\n"
        "  --max-iter <int>        iteration cap (default 500000)\n"
        "  --presolve on|off       accepted for compatibility (pdhg ignores it)\n"
        "  --seed <int>            accepted for compatibility (pdhg ignores it)\n"
        "  --out <file.json>       write solution JSON here (else stdout)\n"
        "  --engine <name>         pdhg (auto|simplex|milp|qp are not in this build)\n"
        "  --precision 32|64       FP32 + FP64 refinement, or plain FP64 (default 64)\n"
        "  --verbose               engine chatter to stderr\n");
}

struct Args {
    std::string cmd, model_path, out_path, engine = "auto";
    igaos::Options opts;
    int precision = 64;
    bool ok = true;
};

Args parse_args(int argc, char** argv) {
    Args a;
    if (argc < 3) {
        a.ok = false;
        return a;
    }
    a.cmd = argv[1];
    a.model_path = argv[2];
    if (a.cmd != "info" && a.cmd != "solve") a.ok = false;
    for (int i = 3; i < argc && a.ok; ++i) {
        const std::string f = argv[i];
        const bool has = i + 1 < argc;
        if (f == "--time-limit" && has) a.opts.time_limit_s = std::atof(argv[++i]);
        else if (f == "--tol" && has) a.opts.tolerance = std::atof(argv[++i]);
        else if (f == "--max-iter" && has) a.opts.max_iterations = std::atol(argv[++i]);
        else if (f == "--seed" && has) a.opts.seed = std::atoi(argv[++i]);
        else if (f == "--presolve" && has) a.opts.presolve = std::string(argv[++i]) == "on";
        else if (f == "--verbose") ++a.opts.verbosity;
        else if (f == "--out" && has) a.out_path = argv[++i];
        else if (f == "--precision" && has) {
            a.precision = std::atoi(argv[++i]);
            if (a.precision != 32 && a.precision != 64) a.ok = false;
        } else if (f == "--engine" && has) {
            a.engine = argv[++i];
            if (a.engine != "auto" && a.engine != "simplex" && a.engine != "pdhg" &&
                a.engine != "milp" && a.engine != "qp")
                a.ok = false;
        } else {
            a.ok = false;
        }
    }
    return a;
}

std::string fmt(const char* f, double v) {
    if (!std::isfinite(v)) return "null";
    char b[64];
    std::snprintf(b, sizeof(b), f, v);
    return b;
}

std::string esc(const std::string& s) {
    std::string r;
    for (char ch : s) {
        if (ch == '"' || ch == '\\') r += '\\';
        r += ch;
    }
    return r;
}

void emit_json(const igaos::io::Model& md, const igaos::Solution& sol,
               const std::string& instance, const std::string& engine,
               const std::string& out_path) {
    std::string j = "{\n";
    j += "  \"instance\": \"" + esc(instance) + "\",\n";
    j += "  \"m\": " + std::to_string(md.m) + ",\n";
    j += "  \"n\": " + std::to_string(md.n) + ",\n";
    j += "  \"nnz\": " + std::to_string(md.nnz()) + ",\n";
    j += std::string("  \"status\": \"") + igaos::status_name(sol.status) + "\",\n";
    j += "  \"objective\": " + fmt("%.10g", sol.objective) + ",\n";
    j += "  \"pinf\": " + fmt("%.3e", sol.pinf) + ",\n";
    j += "  \"dinf\": " + fmt("%.3e", sol.dinf) + ",\n";
    j += "  \"rel_gap\": " + fmt("%.3e", sol.rel_gap) + ",\n";
    j += "  \"iterations\": " + std::to_string(sol.iterations) + ",\n";
    char tb[64];
    std::snprintf(tb, sizeof(tb), "%.1f", sol.solve_time_ms);
    j += "  \"solve_time_ms\": " + std::string(tb) + ",\n";
    j += "  \"engine\": \"" + esc(engine) + "\",\n";
    for (int pass = 0; pass < 2; ++pass) {
        const auto& v = pass == 0 ? sol.x : sol.y;
        j += pass == 0 ? "  \"x\": [" : "  \"y\": [";
        for (size_t k = 0; k < v.size(); ++k) {
            char xb[64];
            std::snprintf(xb, sizeof(xb), "%.10g", v[k]);
            j += xb;
            if (k + 1 < v.size()) j += ",";
        }
        j += "],\n";
    }
    j += "  \"message\": \"" + esc(sol.message) + "\"\n";
    j += "}\n";
    if (out_path.empty()) {
        std::fputs(j.c_str(), stdout);
    } else {
        std::ofstream of(out_path);
        of << j;
    }
}

int cmd_info(const igaos::io::Model& md) {
    int nE = 0, nL = 0, nG = 0, nR = 0;
    for (int i = 0; i < md.m; ++i) {
        if (md.rmin[i] == md.rmax[i]) {
            ++nE;
            continue;
        }
        const bool fl = std::isfinite(md.rmin[i]), fu = std::isfinite(md.rmax[i]);
        if (fl) ++nG;
        if (fu) ++nL;
        if (fl && fu) ++nR;
    }
    int nfree = 0, nfixed = 0, nboxed = 0, none = 0;
    for (int j = 0; j < md.n; ++j) {
        const bool fl = std::isfinite(md.cl[j]), fu = std::isfinite(md.cu[j]);
        if (!fl && !fu) ++nfree;
        else if (fl && fu && md.cl[j] == md.cu[j]) ++nfixed;
        else if (fl && fu) ++nboxed;
        else ++none;
    }
    double cmax = 0;
    for (double v : md.c) cmax = std::max(cmax, std::fabs(v));
    std::printf("rows           %d\n", md.m);
    std::printf("cols           %d\n", md.n);
    std::printf("nnz            %d\n", md.nnz());
    std::printf("row types      E=%d L=%d G=%d ranged=%d\n", nE, nL - nR > 0 ? nL - nR : 0, nG, nR);
    std::printf("col bounds     free=%d fixed=%d boxed=%d one-sided=%d\n", nfree, nfixed, nboxed, none);
    std::printf("objective      |c|max=%.6g\n", cmax);
    std::printf("parsed         FR=%d RANGES=%d obj_const=%.6g\n", md.n_fr_parsed,
                md.n_ranges_parsed, md.obj_const);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        usage();
        return 2;
    }
    if (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0) {
        usage();
        return 0;
    }
    Args a = parse_args(argc, argv);
    if (!a.ok) {
        usage();
        return 2;
    }
    igaos::io::Model md;
    try {
        igaos::io::parse_mps(a.model_path, md);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    if (a.cmd == "info") return cmd_info(md);
    igaos::Solution sol;
    if (a.engine == "pdhg") {
        igaos::pdhg::EngineConfig cfg;
        cfg.precision = a.precision;
        sol = igaos::pdhg::solve_ex(md, a.opts, cfg, nullptr);
    } else {
        sol.status = igaos::Status::Error;
        sol.message = "engine '" + a.engine + "' is not part of this build (pdhg engine slot only)";
    }
    emit_json(md, sol, a.model_path, a.engine, a.out_path);
    return 0;
}
