// PDHG -> simplex crossover: a starting basis for the reference's revised
// simplex (simplex::WarmStart, src/simplex/simplex.hpp:10-15) read off a
// near-optimal primal-dual pair of this engine, so that the simplex can
// finish at a vertex in a few pivots instead of solving from scratch (the
// reference's algorithm sheet lists this crossover as the step after PDHG,
// docs/research/pdlp-algorithm-sheet.md:171).
//
// Variables are numbered as the simplex numbers them: structural columns
// [0, n), row slacks n + i whose value is the row activity (A x)_i with
// bounds [rmin_i, rmax_i].  A variable whose value lies within `tol`
// (relative to 1 + |bound|) of a finite bound is nonbasic at that bound,
// unless its reduced cost (d = c + A'y for a column, y_i for a slack, in
// this engine's sign convention) is zero to the same tolerance while the
// other side is open, in which case it stays a basic candidate.  Basic
// candidates are ranked by how far inside their bounds they lie (relative),
// free variables first, and the best m fill the basis; missing slots are
// left at -1, which the simplex fills with artificial columns.  The states
// use the simplex's encoding: 0 at lower, 1 at upper, 2 basic, 3 free and
// nonbasic.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "model.hpp"

namespace igaos::pdhg {

int crossover_basis(const io::Model& md, const double* x, const double* y, double tol,
                    int* basis, unsigned char* nb_state, bool fill) {
    const int m = md.m, n = md.n;
    const double INF = std::numeric_limits<double>::infinity();
    std::vector<double> act(m, 0.0), d(md.c.begin(), md.c.end());
    for (int i = 0; i < m; ++i)
        for (int p = md.ap[i]; p < md.ap[i + 1]; ++p) act[i] += md.ax[p] * x[md.ai[p]];
    for (int j = 0; j < n; ++j)
        for (int p = md.cp[j]; p < md.cp[j + 1]; ++p) d[j] += md.acx[p] * y[md.ci[p]];
    struct Cand {
        double score;
        int var;
    };
    std::vector<Cand> cand;
    cand.reserve(static_cast<size_t>(n) + m);
    // value v in [lo, hi] with reduced cost r: state or a basic candidate
    auto classify = [&](int var, double v, double lo, double hi, double r) {
        const bool fl = std::isfinite(lo), fu = std::isfinite(hi);
        const double dl = fl ? (v - lo) / (1.0 + std::fabs(lo)) : INF;
        const double du = fu ? (hi - v) / (1.0 + std::fabs(hi)) : INF;
        const double rt = tol * (1.0 + std::fabs(r));
        if (!fl && !fu) {
            cand.push_back({INF, var});
            nb_state[var] = 3;
            return;
        }
        if (fl && fu && lo == hi) {  // fixed (an equality row's slack)
            nb_state[var] = 0;
            return;
        }
        if (dl <= tol && !(std::fabs(r) <= rt && du > tol)) {
            nb_state[var] = 0;
        } else if (du <= tol && !(std::fabs(r) <= rt && dl > tol)) {
            nb_state[var] = 1;
        } else {
            cand.push_back({std::min(dl, du), var});
            nb_state[var] = dl <= du ? 0 : 1;  // if it does not make the basis
        }
    };
    for (int j = 0; j < n; ++j) classify(j, x[j], md.cl[j], md.cu[j], d[j]);
    for (int i = 0; i < m; ++i) classify(n + i, act[i], md.rmin[i], md.rmax[i], y[i]);
    std::stable_sort(cand.begin(), cand.end(),
                     [](const Cand& a, const Cand& b) { return a.score > b.score; });
    int nb = static_cast<int>(std::min<size_t>(cand.size(), static_cast<size_t>(m)));
    for (int s = 0; s < m; ++s) basis[s] = -1;
    for (int s = 0; s < nb; ++s) {
        basis[s] = cand[s].var;
        nb_state[cand[s].var] = 2;
    }
    if (fill && nb < m) {
        // complete the basis with the slacks of the rows whose duals are
        // smallest (the least binding constraints), so that the simplex
        // starts from a full basis rather than artificial columns
        std::vector<int> rows;
        for (int i = 0; i < m; ++i)
            if (nb_state[n + i] != 2) rows.push_back(i);
        std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) {
            return std::fabs(y[a]) < std::fabs(y[b]);
        });
        for (size_t k = 0; k < rows.size() && nb < m; ++k) {
            basis[nb++] = n + rows[k];
            nb_state[n + rows[k]] = 2;
        }
    }
    return nb;
}

}  // namespace igaos::pdhg
