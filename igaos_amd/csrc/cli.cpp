// `igaos info|solve <model.mps> [flags]`: the command-line front end of
// the pdhg engine slot.
//
// The command line and the output are the reference CLI's contract and are
// kept as it defines them: the flags (src/api/igaos_cli.cpp:19-35), the
// JSON document's fields, their order and number formats (115-150; values
// that are not finite print as null), the `info` report (152-185) and the
// exit codes (2: bad command line, 3: unreadable model).  `--engine pdhg`
// runs the MI355X engine; the engines outside the pdhg slot (auto, simplex,
// milp, qp) are not part of this build and report status "error".
// Additions: `--precision 32|64` (FP32 iterations with FP64 refinement, or
// the plain FP64 engine) and a "y" array after "x".  `--seed` is read as an
// int (the reference reinterprets the int as a long).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

#include "igaos/options.h"
#include "igaos/solution.h"
#include "igaos/status.h"
#include "model.hpp"
#include "pdhg.hpp"

namespace {

// ------------------------------------------------------------ command line

struct Invocation {
    std::string command;  // "info" or "solve"
    std::string model;    // MPS path
    std::string out;      // JSON destination ("" = stdout)
    std::string engine = "auto";
    igaos::Options opt;
    int precision = 64;
};

// A flag that takes a value: its name, its help line and what it sets.
struct ValueFlag {
    const char* name;
    const char* help;
    std::function<bool(Invocation&, const char*)> set;
};

bool known_engine(const std::string& e) {
    for (const char* k : {"auto", "simplex", "pdhg", "milp", "qp"})
        if (e == k) return true;
    return false;
}

const std::vector<ValueFlag>& value_flags() {
    static const std::vector<ValueFlag> flags = {
        {"--time-limit", "<sec>      wall-clock budget (default 120)",
         [](Invocation& v, const char* s) { v.opt.time_limit_s = std::atof(s); return true; }},
        {"--tol", "<float>           termination tolerance (default 1e-4)",
         [](Invocation& v, const char* s) { v.opt.tolerance = std::atof(s); return true; }},
        {"--max-iter", "<int>        iteration cap (default 500000)",
         [](Invocation& v, const char* s) { v.opt.max_iterations = std::atol(s); return true; }},
        {"--presolve", "on|off       accepted for compatibility (pdhg ignores it)",
         [](Invocation& v, const char* s) { v.opt.presolve = std::strcmp(s, "on") == 0; return true; }},
        {"--seed", "<int>            accepted for compatibility (pdhg ignores it)",
         [](Invocation& v, const char* s) { v.opt.seed = std::atoi(s); return true; }},
        {"--out", "<file.json>       write solution JSON here (else stdout)",
         [](Invocation& v, const char* s) { v.out = s; return true; }},
        {"--engine", "<name>         pdhg (auto|simplex|milp|qp are not in this build)",
         [](Invocation& v, const char* s) { v.engine = s; return known_engine(v.engine); }},
        {"--precision", "32|64       FP32 + FP64 refinement, or plain FP64 (default 64)",
         [](Invocation& v, const char* s) {
             v.precision = std::atoi(s);
             return v.precision == 32 || v.precision == 64;
         }},
    };
    return flags;
}

void print_help() {
    std::string t =
        "igaos - MI355X PDHG engine\n\n"
        "USAGE:\n"
        "  igaos info   <model.mps>\n"
        "  igaos solve  <model.mps> [flags]\n\n"
        "FLAGS (solve):\n";
    for (const ValueFlag& f : value_flags()) t += std::string("  ") + f.name + " " + f.help + "\n";
    t += "  --verbose               engine chatter to stderr\n";
    std::fputs(t.c_str(), stdout);
}

// false for a command line the reference CLI would reject
bool read_command_line(int argc, char** argv, Invocation& inv) {
    if (argc < 3) return false;
    inv.command = argv[1];
    inv.model = argv[2];
    if (inv.command != "info" && inv.command != "solve") return false;
    for (int i = 3; i < argc; ++i) {
        if (std::strcmp(argv[i], "--verbose") == 0) {
            ++inv.opt.verbosity;
            continue;
        }
        const ValueFlag* hit = nullptr;
        for (const ValueFlag& f : value_flags())
            if (std::strcmp(argv[i], f.name) == 0) hit = &f;
        if (!hit || i + 1 >= argc) return false;
        if (!hit->set(inv, argv[++i])) return false;
    }
    return true;
}

// ---------------------------------------------------------------- output

std::string printf_str(const char* format, double v) {
    char b[64];
    std::snprintf(b, sizeof(b), format, v);
    return b;
}

std::string quoted(const std::string& s) {
    std::string r = "\"";
    for (char ch : s) {
        if (ch == '"' || ch == '\\') r += '\\';
        r += ch;
    }
    return r + "\"";
}

// A flat JSON object written one "  key: value" line per field, in the
// order the fields are added.
class JsonObject {
  public:
    void raw(const char* key, const std::string& value) {
        body_ += body_.empty() ? "" : ",\n";
        body_ += std::string("  \"") + key + "\": " + value;
    }
    void integer(const char* key, long v) { raw(key, std::to_string(v)); }
    void text(const char* key, const std::string& v) { raw(key, quoted(v)); }
    // a finite number in the given format, else null
    void number(const char* key, const char* format, double v) {
        raw(key, std::isfinite(v) ? printf_str(format, v) : std::string("null"));
    }
    void numbers(const char* key, const std::vector<double>& v) {
        std::string a = "[";
        for (size_t k = 0; k < v.size(); ++k) a += (k ? "," : "") + printf_str("%.10g", v[k]);
        raw(key, a + "]");
    }
    std::string str() const { return "{\n" + body_ + "\n}\n"; }

  private:
    std::string body_;
};

void write_solution(const igaos::io::Model& md, const igaos::Solution& sol,
                    const Invocation& inv) {
    JsonObject j;
    j.text("instance", inv.model);
    j.integer("m", md.m);
    j.integer("n", md.n);
    j.integer("nnz", md.nnz());
    j.text("status", igaos::status_name(sol.status));
    j.number("objective", "%.10g", sol.objective);
    j.number("pinf", "%.3e", sol.pinf);
    j.number("dinf", "%.3e", sol.dinf);
    j.number("rel_gap", "%.3e", sol.rel_gap);
    j.integer("iterations", sol.iterations);
    j.raw("solve_time_ms", printf_str("%.1f", sol.solve_time_ms));
    j.text("engine", inv.engine);
    j.numbers("x", sol.x);
    j.numbers("y", sol.y);
    j.text("message", sol.message);
    const std::string doc = j.str();
    if (inv.out.empty()) {
        std::fputs(doc.c_str(), stdout);
    } else {
        std::ofstream(inv.out) << doc;
    }
}

// The `info` report: sizes, row and column kinds, |c|max and what the
// reader parsed.
void report(const igaos::io::Model& md) {
    struct {
        int eq = 0, le = 0, ge = 0, ranged = 0;
    } rows;
    for (int i = 0; i < md.m; ++i) {
        const bool lo = std::isfinite(md.rmin[i]), hi = std::isfinite(md.rmax[i]);
        if (md.rmin[i] == md.rmax[i]) {
            ++rows.eq;
        } else {
            rows.ge += lo;
            rows.le += hi;
            rows.ranged += lo && hi;
        }
    }
    struct {
        int free = 0, fixed = 0, boxed = 0, one_sided = 0;
    } cols;
    for (int j = 0; j < md.n; ++j) {
        const bool lo = std::isfinite(md.cl[j]), hi = std::isfinite(md.cu[j]);
        if (lo && hi) (md.cl[j] == md.cu[j] ? cols.fixed : cols.boxed) += 1;
        else if (lo || hi) ++cols.one_sided;
        else ++cols.free;
    }
    double cmax = 0.0;
    for (double v : md.c) cmax = std::fmax(cmax, std::fabs(v));
    const int le_only = rows.le > rows.ranged ? rows.le - rows.ranged : 0;
    std::printf("rows           %d\n", md.m);
    std::printf("cols           %d\n", md.n);
    std::printf("nnz            %d\n", md.nnz());
    std::printf("row types      E=%d L=%d G=%d ranged=%d\n", rows.eq, le_only, rows.ge,
                rows.ranged);
    std::printf("col bounds     free=%d fixed=%d boxed=%d one-sided=%d\n", cols.free,
                cols.fixed, cols.boxed, cols.one_sided);
    std::printf("objective      |c|max=%.6g\n", cmax);
    std::printf("parsed         FR=%d RANGES=%d obj_const=%.6g\n", md.n_fr_parsed,
                md.n_ranges_parsed, md.obj_const);
}

igaos::Solution run_engine(const igaos::io::Model& md, const Invocation& inv) {
    igaos::Solution sol;
    if (inv.engine != "pdhg") {
        sol.status = igaos::Status::Error;
        sol.message = "engine '" + inv.engine + "' is not part of this build (pdhg engine slot only)";
        return sol;
    }
    igaos::pdhg::EngineConfig cfg;
    cfg.precision = inv.precision;
    return igaos::pdhg::solve_ex(md, inv.opt, cfg, nullptr);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0)) {
        print_help();
        return 0;
    }
    Invocation inv;
    if (!read_command_line(argc, argv, inv)) {
        print_help();
        return 2;
    }
    igaos::io::Model md;
    try {
        igaos::io::parse_mps(inv.model, md);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    if (inv.command == "info") {
        report(md);
        return 0;
    }
    write_solution(md, run_engine(md, inv), inv);
    return 0;
}
