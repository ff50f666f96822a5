/*
 * oracle/kepler_oracle.c -- CPU restatement of `kepler.solve(M, e) -> E`.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is part of the oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or as the timed CPU baseline.  The product path never
 * links it.
 *
 * What it restates.  The reference's Keplerian templates call the third-party
 * C++ extension kepler.py 0.0.7 (requirements.txt:143) as
 *     E = kepler.solve(M, np.repeat(ecc, len(M)))
 * (support/models/kep00.model:6, kep01.model:13, kep02.model:20,
 *  kep03.model:5, kep04.model:14, kep06.model:8, kep07.model:16).
 * kepler.py is not vendored in the reference and is absent here, so its
 * output is pinned by the mathematics instead: E is the unique root of
 * E - e sin E = M (0 <= e < 1).  This solver deliberately uses a different
 * published method from the GPU kernels (which use Markley 1995): a
 * bracketed Newton-Raphson iteration with bisection fallback, run to
 * convergence.  For M reduced to m in [0, pi], the root lies in
 * [m, min(m + e, pi)] because E - m = e sin E lies in [0, e].
 *
 * Output range: E in [0, 2 pi).  Only E mod 2 pi enters the RV model
 * (tan(E/2) and cos/sin E are 2 pi periodic), so the range convention of
 * kepler.py does not affect any likelihood.
 */
#include <math.h>
#include <stdint.h>

static const double TWO_PI = 6.283185307179586476925286766559;
static const double PI_D   = 3.141592653589793238462643383279;

/* Root of E - e sin E = m for m in [0, pi], 0 <= e < 1. */
static double solve_half(double m, double e)
{
    if (e == 0.0 || m == 0.0 || m == PI_D) return m;
    double lo = m, hi = m + e;
    if (hi > PI_D) hi = PI_D;
    /* Danby (1987) style starter clipped into the bracket. */
    double E = m + 0.85 * e;
    if (E > hi) E = 0.5 * (lo + hi);
    for (int it = 0; it < 100; ++it) {
        double s = sin(E), c = cos(E);
        double f = E - e * s - m;
        if (f > 0.0) hi = E; else lo = E;
        double fp = 1.0 - e * c;
        double En = E - f / fp;
        if (!(En > lo && En < hi)) En = 0.5 * (lo + hi);   /* bisection fallback */
        double d = fabs(En - E);
        E = En;
        if (d <= 4.0e-16 * (E > 1.0 ? E : 1.0) || hi - lo <= 4.0e-16 * (hi > 1.0 ? hi : 1.0)) break;
    }
    /* One last Newton polish from the converged point. */
    {
        double s = sin(E), c = cos(E);
        E = E - (E - e * s - m) / (1.0 - e * c);
    }
    return E;
}

double kepler_oracle_one(double M, double e)
{
    double m = fmod(M, TWO_PI);
    if (m < 0.0) m += TWO_PI;
    if (m <= PI_D) return solve_half(m, e);
    return TWO_PI - solve_half(TWO_PI - m, e);
}

/* Vectorised form with kepler.solve's signature: E[i] = solve(M[i], e[i]). */
void kepler_oracle_solve(const double* M, const double* e, double* E, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) E[i] = kepler_oracle_one(M[i], e[i]);
}
