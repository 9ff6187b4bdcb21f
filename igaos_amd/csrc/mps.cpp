// MPS reader and triplet sort for io::Model.
//
// Section semantics follow the reference's parse_mps (src/io/model.hpp:
// 69-218): the first N row is the objective and other N rows are dropped;
// RHS on the objective row is a negated constant; RANGES turn E/G/L rows
// into two-sided rows by the MPS sign rules; BOUNDS types UP (a negative
// UP with a zero lower bound frees the lower bound), LO, FX, FR, MI, PL,
// BV, LI, UI; 'MARKER' INTORG/INTEND lines flag integer columns.  The
// file is read whole and tokenised in place (string views into the buffer,
// name tables keyed by those views, numbers converted where they lie)
// instead of going through an istringstream and string copies per line.
#include <algorithm>
#include <charconv>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <stdexcept>
#include <string>
#include <string_view>
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

// The whole file in one string (NUL-terminated, so strtod stops at its end).
std::string slurp(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string buf;
    in.seekg(0, std::ios::end);
    const std::streamoff size = in.tellg();
    if (size > 0) {
        buf.resize(static_cast<size_t>(size));
        in.seekg(0, std::ios::beg);
        in.read(&buf[0], size);
        buf.resize(static_cast<size_t>(in.gcount()));
    } else {  // not seekable (a pipe): read it as a stream
        in.clear();
        in.seekg(0, std::ios::beg);
        buf.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    }
    return buf;
}

// Token separators: the blanks of the classic locale (space, \t, \n, \v,
// \f, \r), the ones a default-constructed istringstream splits words on.
struct Blanks {
    bool is[256] = {};
    Blanks() {
        for (unsigned char c : {' ', '\t', '\n', '\v', '\f', '\r'}) is[c] = true;
    }
    bool operator()(char c) const { return is[static_cast<unsigned char>(c)]; }
};

void split_blanks(std::string_view line, const Blanks& blank,
                  std::vector<std::string_view>& tok) {
    tok.clear();
    size_t i = 0;
    const size_t L = line.size();
    while (i < L) {
        while (i < L && blank(line[i])) ++i;
        size_t j = i;
        while (j < L && !blank(line[j])) ++j;
        if (j > i) tok.push_back(line.substr(i, j - i));
        i = j;
    }
}

// Name -> index table with open addressing (linear probing over a
// power-of-two array, load <= 1/2): keys are views into the file buffer,
// each slot keeps the key's 64-bit FNV-1a hash so most probes compare one
// word.  A chained std::unordered_map spent most of a large file's reading
// time walking its bucket nodes.
class NameTable {
  public:
    // pointer to the value of `key`, or nullptr
    int* find(std::string_view key) {
        const uint64_t h = hash(key);
        for (size_t s = h & mask_;; s = (s + 1) & mask_) {
            Slot& e = slots_[s];
            if (!e.used) return nullptr;
            if (e.h == h && e.key == key) return &e.val;
        }
    }
    // value of `key`, inserted as `val` when absent; `inserted` tells which
    int& get(std::string_view key, int val, bool& inserted) {
        if (2 * (count_ + 1) > slots_.size()) grow();
        const uint64_t h = hash(key);
        size_t s = h & mask_;
        for (;; s = (s + 1) & mask_) {
            Slot& e = slots_[s];
            if (!e.used) break;
            if (e.h == h && e.key == key) {
                inserted = false;
                return e.val;
            }
        }
        slots_[s] = Slot{key, h, val, true};
        ++count_;
        inserted = true;
        return slots_[s].val;
    }
    size_t size() const { return count_; }

  private:
    struct Slot {
        std::string_view key;
        uint64_t h = 0;
        int val = 0;
        bool used = false;
    };
    static uint64_t hash(std::string_view k) {
        uint64_t h = 1469598103934665603ull;
        for (char c : k) h = (h ^ static_cast<unsigned char>(c)) * 1099511628211ull;
        return h;
    }
    void grow() {
        std::vector<Slot> old;
        old.swap(slots_);
        slots_.assign(old.empty() ? 1024 : 2 * old.size(), Slot{});
        mask_ = slots_.size() - 1;
        for (const Slot& e : old) {
            if (!e.used) continue;
            size_t s = e.h & mask_;
            while (slots_[s].used) s = (s + 1) & mask_;
            slots_[s] = e;
        }
    }
    std::vector<Slot> slots_;
    size_t mask_ = 0, count_ = 0;
};

// std::stod on the token.  Plain decimals go through std::from_chars
// (correctly rounded, as strtod is, so the value is the same double); a
// result that is not a normal nonzero number, or a form from_chars does not
// take (a leading '+', hexadecimal, inf, nan), goes through strtod from the
// token's first character (a token ends at a blank or at the end of the
// buffer, where strtod stops as well) with stod's exceptions for no
// conversion and for ERANGE.
double to_num(std::string_view s) {
    double v = 0.0;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), v);
    if (r.ec == std::errc() && std::isnormal(v)) return v;
    char* end = nullptr;
    errno = 0;
    v = std::strtod(s.data(), &end);
    if (end == s.data()) throw std::invalid_argument("stod");
    if (errno == ERANGE) throw std::out_of_range("stod");
    return v;
}

}  // namespace

void parse_mps(const std::string& path, Model& model) {
    const double INF = std::numeric_limits<double>::infinity();
    const std::string buf = slurp(path);
    const Blanks blank;
    std::string_view section, objname;
    NameTable rowidx, colidx;
    bool fresh = false;
    std::string_view last_col;  // COLUMNS lines come grouped by column
    int last_j = -1;
    std::vector<char> rsense;
    std::vector<double> rhs;
    std::vector<Triplet> trips;
    std::vector<std::pair<int, double>> ranges;
    std::vector<std::string_view> t;
    bool pending_int = false;

    const char* const base = buf.data();
    size_t pos = 0;
    const size_t size = buf.size();
    while (pos < size) {
        const void* nl = std::memchr(base + pos, '\n', size - pos);
        const size_t eol = nl ? static_cast<size_t>(static_cast<const char*>(nl) - base) : size;
        const std::string_view line(base + pos, eol - pos);
        pos = eol + 1;
        if (line.empty() || line[0] == '*') continue;
        const bool header = !(line[0] == ' ' || line[0] == '\t');
        split_blanks(line, blank, t);
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
                rowidx.get(t[1], -1, fresh) = -1;
            } else {
                const int i = static_cast<int>(rsense.size());
                rowidx.get(t[1], i, fresh) = i;  // a repeated name now means the new row
                rsense.push_back(s);
                rhs.push_back(0.0);
            }
        } else if (section == "COLUMNS") {
            if (line.find("'MARKER'") != std::string_view::npos) {
                pending_int = line.find("INTORG") != std::string_view::npos;
                continue;
            }
            int j = last_j;
            if (last_j < 0 || t[0] != last_col) {
                j = colidx.get(t[0], static_cast<int>(colidx.size()), fresh);
                if (fresh) {
                    model.c.push_back(0.0);
                    model.cl.push_back(0.0);
                    model.cu.push_back(INF);
                    model.integ.push_back(pending_int ? 1 : 0);
                }
                last_col = t[0];
                last_j = j;
            }
            for (size_t k = 1; k + 1 < t.size(); k += 2) {
                const double v = to_num(t[k + 1]);
                if (t[k] == objname) {
                    model.c[j] += v;
                    continue;
                }
                const int* r = rowidx.find(t[k]);
                if (!r || *r < 0) continue;
                trips.push_back({*r, j, v});
            }
        } else if (section == "RHS") {
            for (size_t k = 1; k + 1 < t.size(); k += 2) {
                if (t[k] == objname) {
                    model.obj_const -= to_num(t[k + 1]);
                    continue;
                }
                const int* r = rowidx.find(t[k]);
                if (!r || *r < 0) continue;
                rhs[*r] = to_num(t[k + 1]);
            }
        } else if (section == "RANGES") {
            for (size_t k = 1; k + 1 < t.size(); k += 2) {
                const int* r = rowidx.find(t[k]);
                if (!r || *r < 0) continue;
                ranges.emplace_back(*r, to_num(t[k + 1]));
                ++model.n_ranges_parsed;
            }
        } else if (section == "QUADOBJ" || section == "QSECTION" ||
                   section == "QMATRIX") {
            for (size_t k = 0; k + 2 < t.size(); k += 3) {
                const int* c1 = colidx.find(t[k]);
                const int* c2 = colidx.find(t[k + 1]);
                if (!c1 || !c2) continue;
                model.q_i.push_back(*c1);
                model.q_j.push_back(*c2);
                model.q_v.push_back(to_num(t[k + 2]));
            }
        } else if (section == "BOUNDS") {
            if (t.size() < 3) throw std::out_of_range("short BOUNDS line");
            const std::string_view type = t[0];
            const int* jp = colidx.find(t[2]);
            if (!jp) throw std::out_of_range("BOUNDS names an unknown column");
            const int j = *jp;
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

void row_activity(const Model& model, const double* x, double* out) {
    for (int i = 0; i < model.m; ++i) {
        double s = 0.0;
        for (int p = model.ap[i]; p < model.ap[i + 1]; ++p) s += model.ax[p] * x[model.ai[p]];
        out[i] = s;
    }
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

std::string check_transpose(int m, int n, const int* ap, const int* ai, const double* ax,
                            const int* cp, const int* ci, const double* acx) {
    std::string e = check_csr_csc(m, n, ap, ai, cp, ci);
    if (!e.empty()) return e;
    const int nz = ap[m];
    // the CSR's transpose with each column's entries in row order
    std::vector<int> fill(cp, cp + n), tr(nz);
    std::vector<double> tv(nz);
    for (int i = 0; i < m; ++i)
        for (int p = ap[i]; p < ap[i + 1]; ++p) {
            const int q = fill[ai[p]]++;
            tr[q] = i;
            tv[q] = ax[p];
        }
    auto same = [](double a, double b) { return a == b || (a != a && b != b); };
    std::vector<std::pair<int, double>> u, w;
    for (int j = 0; j < n; ++j) {
        bool sorted = true;
        for (int q = cp[j] + 1; q < cp[j + 1] && sorted; ++q) sorted = ci[q - 1] <= ci[q];
        bool ok = true;
        if (sorted) {
            // equal rows (duplicates) may come in either order: compare
            // runs of one row as sorted value lists below if they differ
            for (int q = cp[j]; q < cp[j + 1] && ok; ++q)
                ok = ci[q] == tr[q] && same(acx[q], tv[q]);
        }
        if (!sorted || !ok) {
            u.clear();
            w.clear();
            for (int q = cp[j]; q < cp[j + 1]; ++q) {
                u.emplace_back(ci[q], acx[q]);
                w.emplace_back(tr[q], tv[q]);
            }
            std::sort(u.begin(), u.end());
            std::sort(w.begin(), w.end());
            ok = true;
            for (size_t k = 0; k < u.size() && ok; ++k)
                ok = u[k].first == w[k].first && same(u[k].second, w[k].second);
        }
        if (!ok) return "CSC is not the transpose of CSR in column " + std::to_string(j);
    }
    return {};
}

}  // namespace igaos::io
