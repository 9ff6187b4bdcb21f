// Minimal reference-side driver for the drop-in check: parses an MPS file
// with the reference's io::parse_mps (src/io/model.hpp:69) and solves it
// through igaos::pdhg::solve, i.e. through pdhg_hip.cpp and the MI355X
// engine.  Prints "status objective iterations | message".
#include <cstdio>
#include <cstdlib>

#include "igaos/status.h"
#include "model.hpp"
#include "pdhg.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s file.mps [tolerance]\n", argv[0]);
        return 2;
    }
    igaos::io::Model md;
    igaos::io::parse_mps(argv[1], md);
    igaos::Options o;
    if (argc > 2) o.tolerance = std::atof(argv[2]);
    const igaos::Solution s = igaos::pdhg::solve(md, o);
    std::printf("%s %.12g %ld | %s\n", igaos::status_name(s.status), s.objective,
                s.iterations, s.message.c_str());
    return 0;
}
