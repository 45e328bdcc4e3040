/*
 * F-distribution tail and its inverse.
 *
 * The reference takes F thresholds only (-F / -I, src/main.c:158-171). The
 * benchmark workloads are specified as p-value thresholds, so the host maps
 * p -> F with the survival function of F(d1, d2):
 *     P(F > f) = I_x(d2/2, d1/2),  x = d2 / (d2 + d1 f),
 * where I_x(a, b) is the regularised incomplete beta function. I_x is
 * evaluated from its continued fraction
 *     I_x(a,b) = x^a (1-x)^b / (a B(a,b)) * 1/(1+ d1/(1+ d2/(1+ ...)))
 *     d_{2m+1} = -(a+m)(a+b+m) x / ((a+2m)(a+2m+1))
 *     d_{2m}   =  m (b-m) x / ((a+2m-1)(a+2m))
 * with the modified Lentz algorithm (Lentz 1976; Thompson & Barnett 1986),
 * using the symmetry I_x(a,b) = 1 - I_{1-x}(b,a) where the fraction converges
 * slowly. The inverse is found by bisection in log(F).
 */
#include <math.h>
#include <float.h>

#include "epihost.h"

static double beta_cf(double a, double b, double x)
{
    const double tiny = 1e-300, eps = 1e-16;
    double c = 1.0, d = 1.0 - (a + b) * x / (a + 1.0);
    if (fabs(d) < tiny) d = tiny;
    d = 1.0 / d;
    double h = d;
    for (int m = 1; m <= 10000; m++) {
        double m2 = 2.0 * m;
        double num = m * (b - m) * x / ((a + m2 - 1.0) * (a + m2));
        d = 1.0 + num * d;
        if (fabs(d) < tiny) d = tiny;
        c = 1.0 + num / c;
        if (fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        h *= d * c;
        num = -(a + m) * (a + b + m) * x / ((a + m2) * (a + m2 + 1.0));
        d = 1.0 + num * d;
        if (fabs(d) < tiny) d = tiny;
        c = 1.0 + num / c;
        if (fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < eps) break;
    }
    return h;
}

double epi_betai(double a, double b, double x)
{
    if (!(a > 0.0) || !(b > 0.0) || x != x) return NAN;
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    double lbt = lgamma(a + b) - lgamma(a) - lgamma(b) + a * log(x) + b * log1p(-x);
    if (x < (a + 1.0) / (a + b + 2.0))
        return exp(lbt) * beta_cf(a, b, x) / a;
    return 1.0 - exp(lbt) * beta_cf(b, a, 1.0 - x) / b;
}

double epi_f_sf(double F, double d1, double d2)
{
    if (F != F) return NAN;
    if (F <= 0.0) return 1.0;
    if (isinf(F)) return 0.0;
    double x = d2 / (d2 + d1 * F);
    return epi_betai(0.5 * d2, 0.5 * d1, x);
}

double epi_f_isf(double p, double d1, double d2)
{
    if (!(p > 0.0) || !(p < 1.0) || !(d1 > 0.0) || !(d2 > 0.0)) return NAN;
    double lo = 0.0, hi = 1.0;
    while (epi_f_sf(hi, d1, d2) > p) {
        lo = hi;
        hi *= 2.0;
        if (hi > 1e300) return INFINITY;
    }
    /* sf(lo) > p >= sf(hi); bisect, geometrically once lo > 0. */
    for (int it = 0; it < 400; it++) {
        double mid = lo > 0.0 ? sqrt(lo * hi) : 0.5 * hi;
        if (mid <= lo || mid >= hi) break;
        if (epi_f_sf(mid, d1, d2) > p)
            lo = mid;
        else
            hi = mid;
        if (hi - lo <= 4.0 * DBL_EPSILON * hi) break;
    }
    return 0.5 * (lo + hi);
}
