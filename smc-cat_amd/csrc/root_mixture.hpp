// Root mixture of one site: rate log-sum-exp and the invariant-site term.
//
// Reference formula (Particle.cpp:653-670): l_r = log(pi . x_r) + scale_r,
// l = LSE_r l_r with rate weights 1, L = exp(l) (1 - p) + p pi[c] isInv, log L.
// SMC_MIX_REFERENCE keeps it exactly, including its failure modes: exp(l)
// underflows to 0 below l = -745 (log L = -inf for a variable site) and a rate
// term of -inf or NaN turns the LSE into NaN.  SMC_MIX_LOGSPACE evaluates the same
// quantity without leaving log space: -inf and NaN rate terms contribute 0 and
// log L = logaddexp(l + log(1 - p), log(p pi[c] isInv)).  A NaN rate term is one
// whose partials underflowed: at a rate near 0 (a small gamma shape) P is the
// identity up to round-off, a variable column's partial at some node is then 0
// or a negative round-off in every state, and the reference's scaling by the max
// makes it 0/0.  Its true likelihood is below the range of a double, so it adds 0.
#pragma once

#include "internal.hpp"
#include "smccat_hip.h"

// Every kernel that forms a site's mixture (the pruning roots, the walks, the
// site-class scoring) evaluates these expressions with no multiply-add contraction,
// so the rounding of a site's log-likelihood does not depend on the code around it
// and the kernels agree bit for bit.
namespace smccat {

// log of one rate's site likelihood: pi . x scaled by 2^exponent
__device__ __forceinline__ double rate_log_term(double dot, int exponent)
{
#pragma clang fp contract(off)
    return log(dot) + (double)exponent * 0.69314718055994530942;
}

__device__ __forceinline__ void lse_push(double &m, double &s, double l, int mode)
{
#pragma clang fp contract(off)
    if (mode == SMC_MIX_LOGSPACE && !(l > -INFINITY)) return;  // -inf and NaN
    if (l > m) {
        s = s * exp(m - l) + 1.0;
        m = l;
    } else {
        s += exp(l - m);
    }
}

__device__ __forceinline__ double site_log_likelihood(double m, double s, double pv, double pi_c,
                                                      double is_inv, int mode)
{
#pragma clang fp contract(off)
    const double lsite = m + log(s);
    if (mode != SMC_MIX_LOGSPACE) return log(exp(lsite) * (1.0 - pv) + pv * pi_c * is_inv);
    const double a = lsite + log1p(-pv);
    const double b = (is_inv != 0.0) ? log(pv * pi_c * is_inv) : -INFINITY;
    const double hi = fmax(a, b), lo = fmin(a, b);
    if (hi == -INFINITY) return -INFINITY;
    return hi + log1p(exp(lo - hi));
}

}  // namespace smccat
