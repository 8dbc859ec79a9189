// Merge walk of the OT misfit over one cell of the synthetic CDF: the exact
// integrals of one piecewise-constant cell against the other distribution's
// quantile function.  Plain C++ on doubles, shared by the HIP kernel (ot.hpp)
// and a CPU build of the same code that the tests compile with g++.
#pragma once

#ifdef __HIPCC__
#define SFWI_HD __host__ __device__
#else
#include <cmath>
#define SFWI_HD
#endif

namespace sfwi {

// Largest k in [0, n-1] with Q[k] <= y (Q[0] = 0 <= y).
SFWI_HD inline int ot_floor_index(const double *Q, int n, double y) {
    int lo = 0, hi = n;   // invariant: Q[lo] <= y, answer in [lo, hi)
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (Q[mid] <= y) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Cell i of F, covering the levels [Pi, Pe) with Pe - Pi = pp > 0, against
// the cells of G with cumulative levels Q[0..n] (Q[0] = 0, Q[n] = 1).  Adds
// the cell's share of the W2 integral (times 12) to w and returns its A and B
// moments of h(x) = 2 (x - G^-1(F(x))) in a, b.  Every pass either ends the
// cell or moves k up by one, so a walk over valid levels (Pe <= Q[n]) ends
// within n passes; the bound only ends a walk over levels that break that
// precondition, and the function then returns false (the caller reports it).
SFWI_HD inline bool ot_cell_walk(const double *Q, int n, int i, double Pi, double Pe, double &w, double &a,
                                 double &b) {
    const double pp = Pe - Pi;
    int k = ot_floor_index(Q, n, Pi);
    double ya = Pi;
    a = 0.0;
    b = 0.0;
    for (int pass = 0; pass <= n; ++pass) {
        const double Qk = Q[k], Qn = Q[k + 1];
        const double yb = fmin(Qn, Pe);
        if (yb > ya) {
            const double qq = Qn - Qk;
            const double xa = i + (ya - Pi) / pp, xb = i + (yb - Pi) / pp;
            const double ha = 2.0 * (xa - (k + (ya - Qk) / qq));
            const double hb = 2.0 * (xb - (k + (yb - Qk) / qq));
            const double dx = xb - xa, ua = xa - i, ub = xb - i;
            w += (yb - ya) * (ha * ha + ha * hb + hb * hb);
            a += dx * (ha + hb) / 2.0;
            b += dx * (ha * (2.0 * ua + ub) + hb * (ua + 2.0 * ub)) / 6.0;
        }
        ya = yb;
        if (ya >= Pe) return true;
        while (k < n - 1 && Q[k + 1] <= ya) ++k;
    }
    return false;
}

}  // namespace sfwi
