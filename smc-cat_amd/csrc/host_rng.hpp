// Host-side random numbers shared by the move drivers (not installed).
#pragma once

#include <cmath>
#include <cstdint>

namespace smccat {

// Philox4x32-10 (Salmon et al., SC 2011), counter-based: one stream per slot.
struct Rng {
    uint32_t key[2];
    uint64_t counter = 0;
    Rng(uint64_t seed, uint64_t stream)
    {
        key[0] = (uint32_t)(seed ^ (stream * 0xD1B54A32D192ED03ull));
        key[1] = (uint32_t)((seed >> 32) ^ (stream * 0x9E3779B9u) ^ 0x6d63u);
    }
    double uniform()
    {
        uint32_t c[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), 0x3c6ef372u, 0x13579bdu};
        uint32_t k0 = key[0], k1 = key[1];
        for (int r = 0; r < 10; r++) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
            const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
            const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
            c[0] = n0;
            c[1] = (uint32_t)p1;
            c[2] = n2;
            c[3] = (uint32_t)p0;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        counter++;
        return (double)(((uint64_t)c[0] << 32 | c[1]) >> 11) * (1.0 / 9007199254740992.0);
    }
    double normal()
    {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(6.283185307179586 * u2);
    }
    // Gamma(a, 1): Marsaglia & Tsang (2000), with the a < 1 boost U^(1/a)
    double gamma(double a)
    {
        if (a < 1.0) return gamma(a + 1.0) * std::pow(uniform(), 1.0 / a);
        const double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x, v;
            do {
                x = normal();
                v = 1.0 + c * x;
            } while (v <= 0.0);
            v = v * v * v;
            const double u = uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
            if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
        }
    }
    double beta(double a, double b)
    {
        const double x = gamma(a), y = gamma(b);
        return x / (x + y);
    }
};

}  // namespace smccat
