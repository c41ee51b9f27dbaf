// discretizeGamma (Miscellaneous.hpp:22-37) without Boost: the quantiles of
// Gamma(shape, scale 1/shape) at (2i+1)/(2K) are x_i / shape with
// P(shape, x_i) = q_i, P the regularised lower incomplete gamma function.
// P(a,x) is evaluated by its power series (DLMF 8.11.4) below x = a+1 and by
// the continued fraction of Q(a,x) (DLMF 8.9.2, modified Lentz) above; the
// inverse is Newton's method safeguarded by a shrinking bracket.
#include <cmath>

#include "internal.hpp"

namespace smccat {

namespace {

double reg_lower_gamma(double a, double x)
{
    if (x <= 0.0) return 0.0;
    const double lpref = a * std::log(x) - x - std::lgamma(a);
    if (x < a + 1.0) {
        double term = 1.0 / a, sum = term;
        for (int n = 1; n < 100000; ++n) {
            term *= x / (a + n);
            sum += term;
            if (std::fabs(term) < std::fabs(sum) * 1e-17) break;
        }
        return sum * std::exp(lpref);
    }
    const double tiny = 1e-300;
    double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int n = 1; n < 100000; ++n) {
        const double an = -n * (n - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (std::fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) < 1e-17) break;
    }
    return 1.0 - std::exp(lpref) * h;
}

double inv_reg_lower_gamma(double a, double p)
{
    if (p <= 0.0) return 0.0;
    // starting point: Wilson-Hilferty for a >= 1, small-x series inversion otherwise
    double x;
    if (a >= 1.0) {
        // normal quantile from Abramowitz & Stegun 26.2.23 (|error| < 4.5e-4), then
        // Wilson-Hilferty; Newton below refines to full precision
        const double t = std::sqrt(-2.0 * std::log(p < 0.5 ? p : 1.0 - p));
        double z = t - (2.515517 + 0.802853 * t + 0.010328 * t * t) /
                           (1.0 + 1.432788 * t + 0.189269 * t * t + 0.001308 * t * t * t);
        if (p < 0.5) z = -z;
        const double c = 1.0 / (9.0 * a);
        x = a * std::pow(1.0 - c + z * std::sqrt(c), 3.0);
        if (!(x > 0.0)) x = 0.5 * a;
    } else {
        x = std::exp((std::log(p) + std::lgamma(a + 1.0)) / a);
        if (!(x > 0.0)) x = 1e-300;
    }
    double lo = 0.0, hi = INFINITY;
    for (int it = 0; it < 200; ++it) {
        const double f = reg_lower_gamma(a, x) - p;
        if (f == 0.0) break;
        if (f < 0.0)
            lo = x;
        else
            hi = x;
        const double dens = std::exp((a - 1.0) * std::log(x) - x - std::lgamma(a));
        double xn = (dens > 0.0) ? x - f / dens : x;
        if (!(xn > lo && xn < hi)) xn = std::isfinite(hi) ? 0.5 * (lo + hi) : 2.0 * x + 1.0;
        if (std::fabs(xn - x) <= 1e-16 * std::fabs(x)) {
            x = xn;
            break;
        }
        x = xn;
    }
    return x;
}

}  // namespace

void discretize_gamma(double shape, int num, double *out)
{
    const double interval = 1.0 / (2.0 * num);
    double sum = 0.0;
    for (int i = 0; i < num; ++i) {
        out[i] = inv_reg_lower_gamma(shape, (i * 2.0 + 1.0) * interval) / shape;
        sum += out[i];
    }
    const double factor = num / sum;
    for (int i = 0; i < num; ++i) out[i] *= factor;
}

}  // namespace smccat
