// MCMC proposals on a particle's parameters (host code, shared by the batched
// rejuvenation in mcmc.cpp and the C++ facade's Particle moves).
//
// Reference: Particle::treeMove / shapeMove / invarMove / rateMatrixMove
// (Particle.cpp:677-783) and the tree moves of Tree.cpp:553-811.  Each proposal
// edits the parameters in place, marks node flags (bit 0 updateCL, bit 1
// updateTP, Tree.hpp:26-27) and returns the log Hastings ratio.  Differences from
// the reference: adaptive NNI proposes a plain NNI (the split posteriors that bias
// it live outside this engine); SPR marks the old and the new path to the root
// dirty (the reference marks only the new one, Tree.cpp:795-800).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "host_rng.hpp"
#include "smccat_hip.h"

namespace smccat {

// The parameters a move changes: children [T-1][2], branch lengths [2T-1] by
// node id, the 190 log-exchangeabilities, pInvar and the gamma shape.
struct MoveParams {
    std::vector<int32_t> ch;
    std::vector<double> bl, base;
    double pinv = 0.0, shape = 1.0;
};

inline double log_beta_pdf(double a, double b, double x)
{
    return (a - 1.0) * std::log(x) + (b - 1.0) * std::log1p(-x) + std::lgamma(a + b) - std::lgamma(a) -
           std::lgamma(b);
}

inline void move_parents(const std::vector<int32_t> &ch, int T, std::vector<int> &par)
{
    par.assign(2 * T - 1, -1);
    for (int k = 0; k < T - 1; ++k) {
        par[ch[2 * k]] = T + k;
        par[ch[2 * k + 1]] = T + k;
    }
}

inline void mark_path(std::vector<uint8_t> &f, const std::vector<int> &par, int n)
{
    for (; n >= 0; n = par[n]) f[n] |= 1;
}

// One tree move (Particle::treeMove mixture, Particle.cpp:680-701).
inline double tree_move(Rng &g, MoveParams &q, int T, std::vector<uint8_t> &f,
                        const smc_rejuvenation_options &o)
{
    const int nn = 2 * T - 1;
    std::vector<int> par;
    move_parents(q.ch, T, par);
    int root = -1;
    for (int n = T; n < nn; ++n)
        if (par[n] < 0) root = n;
    auto pick = [&](bool internal) {
        for (;;) {
            const int n = std::min(nn - 1, (int)(g.uniform() * nn));
            if (n != root && (!internal || n >= T)) return n;
        }
    };
    const double u = g.uniform();
    if (u < 0.4) {  // NNI 0.1 and adaptive NNI 0.3 (Tree.cpp:625-666)
        const int p = pick(true), a = par[p];
        const int n1 = q.ch[2 * (p - T) + (g.uniform() < 0.5 ? 0 : 1)];
        const int n2 = q.ch[2 * (a - T)] == p ? q.ch[2 * (a - T) + 1] : q.ch[2 * (a - T)];
        for (int c = 0; c < 2; ++c) {
            if (q.ch[2 * (p - T) + c] == n1) q.ch[2 * (p - T) + c] = n2;
            if (q.ch[2 * (a - T) + c] == n2) q.ch[2 * (a - T) + c] = n1;
        }
        move_parents(q.ch, T, par);
        mark_path(f, par, p);
        return 0.0;
    }
    if (u < 0.5) {  // SPR (Tree.cpp:753-811)
        const int p = pick(true);
        std::vector<char> in_sub(nn, 0);
        std::vector<int> st{p};
        while (!st.empty()) {
            const int x = st.back();
            st.pop_back();
            in_sub[x] = 1;
            if (x >= T) {
                st.push_back(q.ch[2 * (x - T)]);
                st.push_back(q.ch[2 * (x - T) + 1]);
            }
        }
        int tgt = -1;
        for (int tries = 0; tries < 64 && tgt < 0; ++tries) {
            const int a = std::min(nn - 1, (int)(g.uniform() * nn));
            if (!in_sub[a] && a != root) tgt = a;
        }
        if (tgt < 0) return 0.0;
        const int s = q.ch[2 * (p - T) + (g.uniform() < 0.5 ? 0 : 1)];
        const int sib = q.ch[2 * (p - T)] == s ? q.ch[2 * (p - T) + 1] : q.ch[2 * (p - T)];
        const int old = par[p];
        if (tgt == sib) return 0.0;  // regrafting onto its own sibling changes nothing
        for (int c = 0; c < 2; ++c)
            if (q.ch[2 * (old - T) + c] == p) q.ch[2 * (old - T) + c] = sib;
        const int tpar = par[tgt] == p ? old : par[tgt];
        q.ch[2 * (p - T)] = s;
        q.ch[2 * (p - T) + 1] = tgt;
        for (int c = 0; c < 2; ++c)
            if (q.ch[2 * (tpar - T) + c] == tgt) q.ch[2 * (tpar - T) + c] = p;
        move_parents(q.ch, T, par);
        mark_path(f, par, p);
        mark_path(f, par, old);
        return 0.0;
    }
    if (u < 0.9) {  // branch-length scale (Tree.cpp:553-577)
        const int n = pick(false);
        const double sc = std::exp(o.scale_delta * (g.uniform() - 0.5));
        q.bl[n] *= sc;
        f[n] |= 2;
        mark_path(f, par, par[n]);
        return std::log(sc);
    }
    // subtree scale (Tree.cpp:579-612)
    const int n = pick(false);
    const double sc = std::exp(o.subtree_scale_delta * (g.uniform() - 0.5));
    int nb = 0;
    std::vector<int> st{n};
    while (!st.empty()) {
        const int x = st.back();
        st.pop_back();
        q.bl[x] *= sc;
        f[x] |= 3;
        nb++;
        if (x >= T) {
            st.push_back(q.ch[2 * (x - T)]);
            st.push_back(q.ch[2 * (x - T) + 1]);
        }
    }
    mark_path(f, par, par[n]);
    return nb * std::log(sc);
}

// Particle::shapeMove (Particle.cpp:768-783): every P changes (Tree::updateAll).
inline double shape_move(Rng &g, MoveParams &q, std::vector<uint8_t> &f, double shape_delta)
{
    const double sc = std::exp(shape_delta * (g.uniform() - 0.5));
    q.shape *= sc;
    std::fill(f.begin(), f.end(), 3);
    return std::log(sc);
}

// Particle::invarMove (Particle.cpp:744-766): Beta proposal around the current pInvar.
inline double invar_move(Rng &g, MoveParams &q, double invar_alpha)
{
    const double a = invar_alpha + 1.0, b = invar_alpha / q.pinv - a + 2.0;
    const double np = g.beta(a, b);
    const double a2 = invar_alpha + 1.0, b2 = invar_alpha / np - a2 + 2.0;
    const double h = log_beta_pdf(a2, b2, q.pinv) - log_beta_pdf(a, b, np);
    q.pinv = np;
    return h;
}

// Particle::rateMatrixMove (Particle.cpp:706-732): a full 190-dimensional step
// (base += L z) with probability 0.25, else one coordinate.  The caller
// recomputes every class's eigensystem (the reference keeps stale eigenvectors,
// Q13).  Hastings ratio 0.
inline double rate_matrix_move(Rng &g, MoveParams &q, const smc_rejuvenation_options &o)
{
    if (g.uniform() < 0.25) {
        double z[190];
        for (double &v : z) v = g.normal();
        if (o.rate_cholesky) {
            for (int a = 0; a < 190; ++a) {
                double acc = 0.0;
                for (int b = 0; b <= a; ++b) acc += o.rate_cholesky[a * 190 + b] * z[b];
                q.base[a] += acc;
            }
        } else {
            for (int a = 0; a < 190; ++a) q.base[a] += o.rate_full_scale * z[a];
        }
    } else {
        const int idx = std::min(189, (int)(190 * g.uniform()));
        q.base[idx] += g.normal() * o.rate_delta;
    }
    return 0.0;
}

}  // namespace smccat
