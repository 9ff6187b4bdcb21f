// MPS reader and triplet sort for io::Model.
//
// Section semantics follow the reference's parse_mps (src/io/model.hpp:
// 69-218): the first N row is the objective and other N rows are dropped;
// RHS on the objective row is a negated constant; RANGES turn E/G/L rows
// into two-sided rows by the MPS sign rules; BOUNDS types UP (a negative
// UP with a zero lower bound frees the lower bound), LO, FX, FR, MI, PL,
// BV, LI, UI; 'MARKER' INTORG/INTEND lines flag integer columns.  The
// tokenizer splits each line on blanks itself instead of going through an
// istringstream per line.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "model.hpp"

namespace igaos::io {

void counting_sort(const std::vector<Triplet>& trips, int rows, int cols,
                   std::vector<int>& ap, std::vector<int>& ai,
                   std::vector<double>& ax, std::vector<int>& cp,
                   std::vector<int>& ci, std::vector<double>& acx) {
    ap.assign(rows + 1, 0);
    cp.assign(cols + 1, 0);
    for (const Triplet& t : trips) {
        ap[t.row + 1]++;
        cp[t.col + 1]++;
    }
    for (int i = 0; i < rows; ++i) ap[i + 1] += ap[i];
    for (int j = 0; j < cols; ++j) cp[j + 1] += cp[j];
    std::vector<int> rpos(ap.begin(), ap.end() - 1);
    std::vector<int> cpos(cp.begin(), cp.end() - 1);
    const size_t nz = trips.size();
    ai.resize(nz);
    ax.resize(nz);
    ci.resize(nz);
    acx.resize(nz);
    for (const Triplet& t : trips) {
        int pr = rpos[t.row]++;
        ai[pr] = t.col;
        ax[pr] = t.val;
        int pc = cpos[t.col]++;
        ci[pc] = t.row;
        acx[pc] = t.val;
    }
}

namespace {

void split_blanks(const std::string& line, std::vector<std::string>& tok) {
    tok.clear();
    size_t i = 0, L = line.size();
    while (i < L) {
        while (i < L && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
        size_t j = i;
        while (j < L && !std::isspace(static_cast<unsigned char>(line[j]))) ++j;
        if (j > i) tok.emplace_back(line, i, j - i);
        i = j;
    }
}

double to_num(const std::string& s) {
    size_t used = 0;
    double v = std::stod(s, &used);  // throws std::invalid_argument like stod
    return v;
}

}  // namespace

void parse_mps(const std::string& path, Model& model) {
    const double INF = std::numeric_limits<double>::infinity();
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line, section, objname;
    std::unordered_map<std::string, int> rowidx, colidx;
    std::vector<char> rsense;
    std::vector<double> rhs;
    std::vector<Triplet> trips;
    std::vector<std::pair<int, double>> ranges;
    std::vector<std::string> t;
    bool pending_int = false;

    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '*') continue;
        const bool header = !(line[0] == ' ' || line[0] == '\t');
        split_blanks(line, t);
        if (t.empty()) continue;
        if (header) {
            section = t[0];
            continue;
        }
        if (section == "ROWS") {
            const char s = t[0][0];
            if (t.size() < 2) continue;
            if (s == 'N') {
                if (objname.empty()) objname = t[1];
                rowidx[t[1]] = -1;
            } else {
                rowidx[t[1]] = static_cast<int>(rsense.size());
                rsense.push_back(s);
                rhs.push_back(0.0);
            }
        } else if (section == "COLUMNS") {
            if (line.find("'MARKER'") != std::string::npos) {
                pending_int = line.find("INTORG") != std::string::npos;
                continue;
            }
            auto it = colidx.find(t[0]);
            int j;
            if (it == colidx.end()) {
                j = static_cast<int>(colidx.size());
                colidx.emplace(t[0], j);
                model.c.push_back(0.0);
                model.cl.push_back(0.0);
                model.cu.push_back(INF);
                model.integ.push_back(pending_int ? 1 : 0);
            } else {
                j = it->second;
            }
            for (size_t k = 1; k + 1 < t.size(); k += 2) {
                const double v = to_num(t[k + 1]);
                if (t[k] == objname) {
                    model.c[j] += v;
                    continue;
                }
                auto r = rowidx.find(t[k]);
                if (r == rowidx.end() || r->second < 0) continue;
                trips.push_back({r->second, j, v});
            }
        } else if (section == "RHS") {
            for (size_t k = 1; k + 1 < t.size(); k += 2) {
                if (t[k] == objname) {
                    model.obj_const -= to_num(t[k + 1]);
                    continue;
                }
                auto r = rowidx.find(t[k]);
                if (r == rowidx.end() || r->second < 0) continue;
                rhs[r->second] = to_num(t[k + 1]);
            }
        } else if (section == "RANGES") {
            for (size_t k = 1; k + 1 < t.size(); k += 2) {
                auto r = rowidx.find(t[k]);
                if (r == rowidx.end() || r->second < 0) continue;
                ranges.emplace_back(r->second, to_num(t[k + 1]));
                ++model.n_ranges_parsed;
            }
        } else if (section == "QUADOBJ" || section == "QSECTION" ||
                   section == "QMATRIX") {
            for (size_t k = 0; k + 2 < t.size(); k += 3) {
                auto c1 = colidx.find(t[k]);
                auto c2 = colidx.find(t[k + 1]);
                if (c1 == colidx.end() || c2 == colidx.end()) continue;
                model.q_i.push_back(c1->second);
                model.q_j.push_back(c2->second);
                model.q_v.push_back(to_num(t[k + 2]));
            }
        } else if (section == "BOUNDS") {
            if (t.size() < 3) throw std::out_of_range("short BOUNDS line");
            const std::string& type = t[0];
            const int j = colidx.at(t[2]);
            const double v = t.size() > 3 ? to_num(t[3]) : 0.0;
            if (type == "UP") {
                model.cu[j] = v;
                if (v < 0 && model.cl[j] == 0.0) model.cl[j] = -INF;
            } else if (type == "LO") {
                model.cl[j] = v;
            } else if (type == "FX") {
                model.cl[j] = model.cu[j] = v;
            } else if (type == "FR") {
                model.cl[j] = -INF;
                model.cu[j] = INF;
                ++model.n_fr_parsed;
            } else if (type == "MI") {
                model.cl[j] = -INF;
            } else if (type == "PL") {
                model.cu[j] = INF;
            } else if (type == "BV") {
                model.cl[j] = 0.0;
                model.cu[j] = 1.0;
                model.integ[j] = 1;
            } else if (type == "LI") {
                model.cl[j] = v;
                model.integ[j] = 1;
            } else if (type == "UI") {
                model.cu[j] = v;
                model.integ[j] = 1;
            }
        }
    }

    model.m = static_cast<int>(rsense.size());
    model.n = static_cast<int>(colidx.size());
    model.rmin.assign(model.m, -INF);
    model.rmax.assign(model.m, INF);
    for (int i = 0; i < model.m; ++i) {
        if (rsense[i] == 'E') model.rmin[i] = model.rmax[i] = rhs[i];
        else if (rsense[i] == 'L') model.rmax[i] = rhs[i];
        else model.rmin[i] = rhs[i];
    }
    for (const auto& pr : ranges) {
        const int i = pr.first;
        const double r = pr.second, b = rhs[i];
        if (rsense[i] == 'E') {
            if (r >= 0) { model.rmin[i] = b; model.rmax[i] = b + r; }
            else { model.rmin[i] = b + r; model.rmax[i] = b; }
        } else if (rsense[i] == 'G') {
            model.rmin[i] = b;
            model.rmax[i] = b + std::fabs(r);
        } else {
            model.rmin[i] = b - std::fabs(r);
            model.rmax[i] = b;
        }
    }
    counting_sort(trips, model.m, model.n, model.ap, model.ai, model.ax,
                  model.cp, model.ci, model.acx);
}

std::string check_compressed(int rows, int cols, const int* ptr, const int* idx) {
    if (rows < 0 || cols < 0) return "negative size";
    if (!ptr) return "null pointer array";
    if (ptr[0] != 0) return "pointer array does not start at 0";
    for (int i = 0; i < rows; ++i)
        if (ptr[i + 1] < ptr[i])
            return "pointer array decreases at " + std::to_string(i);
    const int nz = ptr[rows];
    if (nz > 0 && !idx) return "null index array";
    for (int p = 0; p < nz; ++p)
        if (idx[p] < 0 || idx[p] >= cols)
            return "index " + std::to_string(idx[p]) + " at entry " + std::to_string(p) +
                   " is outside [0, " + std::to_string(cols) + ")";
    return {};
}

std::string check_csr_csc(int m, int n, const int* ap, const int* ai, const int* cp,
                          const int* ci) {
    std::string e = check_compressed(m, n, ap, ai);
    if (!e.empty()) return "CSR: " + e;
    e = check_compressed(n, m, cp, ci);
    if (!e.empty()) return "CSC: " + e;
    if (ap[m] != cp[n]) return "CSR and CSC nonzero counts differ";
    std::vector<int> per_col(n, 0), per_row(m, 0);
    for (int p = 0; p < ap[m]; ++p) per_col[ai[p]]++;
    for (int p = 0; p < cp[n]; ++p) per_row[ci[p]]++;
    for (int j = 0; j < n; ++j)
        if (per_col[j] != cp[j + 1] - cp[j])
            return "CSR and CSC differ in the entries of column " + std::to_string(j);
    for (int i = 0; i < m; ++i)
        if (per_row[i] != ap[i + 1] - ap[i])
            return "CSR and CSC differ in the entries of row " + std::to_string(i);
    return {};
}

}  // namespace igaos::io
