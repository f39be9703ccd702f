// A C++ caller of the reference's header API, built against include/sfac.hpp
// and linked to libsfac.so the way a user of the reference links libsfac.
// The first part is the reference's own test.cpp (test.cpp:1-18): a table of
// the order-4 cardinal B-spline through BSpline::bspl_coef.  The rest needs
// no GPU: BSpline::dbspl_coef, the constructors' argument rejection (the C
// entry point returns NULL, the reference's C++ constructor symbol throws).
#include <stdio.h>

#include <stdexcept>

#include "sfac.hpp"

int main() {
    int order = 4;
    double dx = 0.1;
    BSpline B(order);

    int nmax = order / dx;

    // (the reference's loop runs to i = nmax, where jj = order indexes one row
    // past the coefficient matrix; this one stops before it)
    for (int i = 0; i < nmax; i++) {
        double x = i * dx;
        int jj = (int)x;
        double c = B.bspl_coef(x - jj, jj);
        printf("%.2f %f\n", x, c);
    }
    // value and derivative at a few points
    for (int j = 0; j < order; ++j) {
        double b, db;
        B.dbspl_coef(&b, &db, 0.3, j);
        printf("d %d %.17g %.17g\n", j, b, db);
    }
    double L[6] = {10, 10, 10, 0, 0, 0};
    int K[3] = {8, 8, 8};
    void *h = sfac_ctor(L, K, 3);
    printf("ctor_odd_order %s %s\n", h ? "handle" : "NULL", sfac_last_error());
    int K0[3] = {8, 0, 8};
    h = sfac_ctor(L, K0, 4);
    printf("ctor_bad_K %s %s\n", h ? "handle" : "NULL", sfac_last_error());
    try {
        int32_t Ki[3] = {8, 8, 8};
        SFac s(L, Ki, 5);
        printf("cxx_ctor returned\n");
    } catch (const std::invalid_argument &e) {
        printf("cxx_ctor throws %s\n", e.what());
    }
    return 0;
}
