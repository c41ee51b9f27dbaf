/*
 * smccat_oracle.c -- TEST INFRASTRUCTURE ONLY (see smccat_oracle.h).
 *
 * CPU FP64 restatement of the reference's particle hot path.  The loop order
 * of the likelihood follows Particle.cpp:577-675 (node -> child -> class ->
 * rate -> member site, then per-(site,rate) max-rescale with log, then the
 * root mixture that sums every node's log-scale), so the same code is also
 * the timed CPU baseline.  The product library never links this file.
 */
#include "smccat_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NS 20 /* amino-acid states */

/* ------------------------------------------------------------------ a1 */
/* RateMatrices.hpp:97-111: c runs over i = 1..19, j = 0..i-1 row by row. */
void orc_lower_triangle_coord(int c, int *i, int *j)
{
    int row = 1;
    while (c >= row) {
        c -= row;
        row++;
    }
    *i = row;
    *j = c;
}

/* -------------------------------------------------------------- a3, a4 */
/* TransitionProbabilityClass.cpp:101-109 (softmax without max shift) and
 * :50-75 (Q_ij = exp(b_c) pi_j, Q_ii = -sum_{j!=i} Q_ij, Q /= sum_i pi_i(-Q_ii)). */
void orc_build_q(const double base[190], const double logits[20], double pi[20], double Q[400])
{
    double tot = 0.0;
    for (int i = 0; i < NS; i++) {
        pi[i] = exp(logits[i]);
        tot += pi[i];
    }
    for (int i = 0; i < NS; i++) pi[i] /= tot;

    memset(Q, 0, sizeof(double) * 400);
    for (int c = 0; c < 190; c++) {
        int i, j;
        orc_lower_triangle_coord(c, &i, &j);
        double v = exp(base[c]);
        Q[i * NS + j] = v;
        Q[j * NS + i] = v;
    }
    for (int i = 0; i < NS; i++)
        for (int j = 0; j < NS; j++) Q[i * NS + j] *= pi[j];
    for (int i = 0; i < NS; i++) {
        double off = 0.0;
        for (int j = 0; j < NS; j++)
            if (j != i) off += Q[i * NS + j];
        Q[i * NS + i] = -off;
    }
    double mean_rate = 0.0;
    for (int i = 0; i < NS; i++) mean_rate += -Q[i * NS + i] * pi[i];
    for (int k = 0; k < 400; k++) Q[k] /= mean_rate;
}

/* Serial cyclic Jacobi on a symmetric 20x20 matrix (Golub & Van Loan,
 * "Matrix Computations", Alg. 8.4.2/8.4.3).  S is overwritten; V accumulates
 * the rotations so that S_in = V diag(S_out) V^T. */
static void jacobi_sym(double S[400], double V[400])
{
    for (int i = 0; i < 400; i++) V[i] = 0.0;
    for (int i = 0; i < NS; i++) V[i * NS + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < NS; i++)
            for (int j = 0; j < NS; j++) {
                double v = S[i * NS + j] * S[i * NS + j];
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * tot) break;  /* the rounding floor is about 1e-32 of tot */
        for (int p = 0; p < NS - 1; p++) {
            for (int q = p + 1; q < NS; q++) {
                double apq = S[p * NS + q];
                if (apq == 0.0) continue;
                double tau = (S[q * NS + q] - S[p * NS + p]) / (2.0 * apq);
                double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double c = 1.0 / sqrt(1.0 + t * t);
                double s = t * c;
                for (int k = 0; k < NS; k++) { /* columns p,q: S <- S J */
                    double skp = S[k * NS + p], skq = S[k * NS + q];
                    S[k * NS + p] = c * skp - s * skq;
                    S[k * NS + q] = s * skp + c * skq;
                }
                for (int k = 0; k < NS; k++) { /* rows p,q: S <- J^T S */
                    double spk = S[p * NS + k], sqk = S[q * NS + k];
                    S[p * NS + k] = c * spk - s * sqk;
                    S[q * NS + k] = s * spk + c * sqk;
                }
                for (int k = 0; k < NS; k++) {
                    double vkp = V[k * NS + p], vkq = V[k * NS + q];
                    V[k * NS + p] = c * vkp - s * vkq;
                    V[k * NS + q] = s * vkp + c * vkq;
                }
            }
        }
    }
}

/* a4: TransitionProbabilityClass::recomputeEigens (TransitionProbabilityClass.cpp:50-81).
 * The reference uses a general complex eigensolver; Q is reversible, so the
 * decomposition is computed here on Pi^{1/2} Q Pi^{-1/2} (symmetric) instead. */
void orc_eigen(const double base[190], const double logits[20],
               double A[400], double lam[20], double B[400], double pi[20])
{
    double Q[400], S[400], U[400], sp[NS];
    orc_build_q(base, logits, pi, Q);
    for (int i = 0; i < NS; i++) sp[i] = sqrt(pi[i]);
    for (int i = 0; i < NS; i++)
        for (int j = 0; j < NS; j++) S[i * NS + j] = (i == j) ? Q[i * NS + i] : 0.5 * (sp[i] * Q[i * NS + j] / sp[j] + sp[j] * Q[j * NS + i] / sp[i]);
    jacobi_sym(S, U);
    for (int k = 0; k < NS; k++) lam[k] = S[k * NS + k];
    for (int i = 0; i < NS; i++)
        for (int k = 0; k < NS; k++) {
            A[i * NS + k] = U[i * NS + k] / sp[i];
            B[k * NS + i] = U[i * NS + k] * sp[i];
        }
}

/* a5: TransitionProbabilityClass::recomputeTransitionProbs (:83-95). */
void orc_pmatrix(const double A[400], const double lam[20], const double B[400],
                 double t, double r, double P[400])
{
    double e[NS];
    for (int k = 0; k < NS; k++) e[k] = exp(lam[k] * t * r);
    for (int i = 0; i < NS; i++)
        for (int j = 0; j < NS; j++) {
            double acc = 0.0;
            for (int k = 0; k < NS; k++) acc += A[i * NS + k] * e[k] * B[k * NS + j];
            P[i * NS + j] = acc;
        }
}

/* ------------------------------------------------------------------ a6 */
/* P(a,x): power series (DLMF 8.11.4) for x < a+1, otherwise the continued
 * fraction of Q(a,x) (DLMF 8.9.2) evaluated with the modified Lentz method. */
double orc_gamma_p(double a, double x)
{
    if (x <= 0.0) return 0.0;
    double lpref = a * log(x) - x - lgamma(a);
    if (x < a + 1.0) {
        double term = 1.0 / a, sum = term;
        for (int n = 1; n < 100000; n++) {
            term *= x / (a + n);
            sum += term;
            if (fabs(term) < fabs(sum) * 1e-17) break;
        }
        return sum * exp(lpref);
    }
    const double tiny = 1e-300;
    double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int n = 1; n < 100000; n++) {
        double an = -n * (n - a);
        b += 2.0;
        d = an * d + b;
        if (fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < 1e-17) break;
    }
    return 1.0 - exp(lpref) * h;
}

/* Inverse of P(a, .) by bracketing and bisection to full double precision. */
double orc_gamma_p_inv(double a, double p)
{
    if (p <= 0.0) return 0.0;
    double lo = 0.0, hi = a > 1.0 ? a : 1.0;
    while (orc_gamma_p(a, hi) < p) {
        lo = hi;
        hi *= 2.0;
    }
    for (int it = 0; it < 2000; it++) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (orc_gamma_p(a, mid) < p)
            lo = mid;
        else
            hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* discretizeGamma (Miscellaneous.hpp:22-37): quantiles of Gamma(shape, scale
 * 1/shape) at (2i+1)/(2K), rescaled to mean 1. */
void orc_discretize_gamma(double shape, int num, double *out)
{
    double interval = 1.0 / (2.0 * num);
    double sum = 0.0;
    for (int i = 0; i < num; i++) {
        out[i] = orc_gamma_p_inv(shape, (i * 2.0 + 1.0) * interval) / shape;
        sum += out[i];
    }
    double factor = num / sum;
    for (int i = 0; i < num; i++) out[i] *= factor;
}

/* ------------------------------------------------------------------ a7 */
/* Particle.cpp:44-72: a column is invariant iff every taxon has exactly the
 * code of taxon 0 (gaps included, Q5); its character is that code if < 20, else 0. */
void orc_invariant_sites(const int8_t *codes, int taxa, int sites,
                         double *is_invariant, int32_t *invariant_char)
{
    for (int s = 0; s < sites; s++) {
        int c1 = codes[s];
        double inv = 1.0;
        for (int t = 0; t < taxa; t++)
            if (codes[(size_t)t * sites + s] != c1) inv = 0.0;
        is_invariant[s] = inv;
        invariant_char[s] = (inv == 1.0 && c1 < 20) ? c1 : 0;
    }
}

/* ------------------------------------------------------------ tree helpers */
static int tree_root(const orc_particle *p)
{
    int T = p->taxa, nn = 2 * T - 1;
    char *has_parent = (char *)calloc(nn, 1);
    for (int i = 0; i < T - 1; i++) {
        has_parent[p->children[2 * i]] = 1;
        has_parent[p->children[2 * i + 1]] = 1;
    }
    int root = -1;
    for (int n = T; n < nn; n++)
        if (!has_parent[n]) root = n;
    free(has_parent);
    return root;
}

/* Post-order of internal nodes (children before parents), Tree.cpp:462-476. */
static int internal_postorder(const orc_particle *p, int *order)
{
    int T = p->taxa;
    int *stack = (int *)malloc(sizeof(int) * 2 * (2 * T));
    int sp = 0, cnt = 0;
    stack[sp++] = tree_root(p);
    stack[sp++] = 0;
    while (sp > 0) {
        int state = stack[--sp];
        int n = stack[--sp];
        if (n < T) continue;
        if (state == 0) {
            stack[sp++] = n;
            stack[sp++] = 1;
            for (int k = 1; k >= 0; k--) {
                stack[sp++] = p->children[2 * (n - T) + k];
                stack[sp++] = 0;
            }
        } else {
            order[cnt++] = n;
        }
    }
    free(stack);
    return cnt;
}

/* --------------------------------------------------------------- a8-a10 */
typedef struct {
    double A[400], B[400], lam[20], pi[20];
} orc_class_eig;

double orc_particle_lnl_sites_mode(const orc_particle *p, double *site_lnl, int logspace)
{
    const int T = p->taxa, S = p->sites, R = p->num_rates, K = p->num_classes;
    const int nn = 2 * T - 1;
    const size_t node_sp = (size_t)S * R; /* nodeSpacer, Particle.cpp:577 */

    double *cl = (double *)malloc(sizeof(double) * NS * node_sp * nn);
    double *resc = (double *)calloc(node_sp * nn, sizeof(double));
    double *P = (double *)malloc(sizeof(double) * 400 * (size_t)K * nn * R);
    orc_class_eig *eig = (orc_class_eig *)malloc(sizeof(orc_class_eig) * K);
    double *is_inv = (double *)malloc(sizeof(double) * S);
    int32_t *inv_char = (int32_t *)malloc(sizeof(int32_t) * S);
    int *order = (int *)malloc(sizeof(int) * T);

    /* class membership lists in ascending site order (std::set<int> members) */
    int *cnt = (int *)calloc(K, sizeof(int));
    int *start = (int *)calloc(K + 1, sizeof(int));
    int *members = (int *)malloc(sizeof(int) * S);
    for (int s = 0; s < S; s++) cnt[p->assignment[s]]++;
    for (int k = 0; k < K; k++) start[k + 1] = start[k] + cnt[k];
    memset(cnt, 0, sizeof(int) * K);
    for (int s = 0; s < S; s++) {
        int k = p->assignment[s];
        members[start[k] + cnt[k]++] = s;
    }

    /* tip encoding, Particle.cpp:44-64: one-hot for codes < 20, all ones otherwise */
    for (int t = 0; t < T; t++)
        for (int s = 0; s < S; s++) {
            int code = p->codes[(size_t)t * S + s];
            for (int r = 0; r < R; r++) {
                double *x = cl + NS * ((size_t)t * node_sp + (size_t)s * R + r);
                for (int i = 0; i < NS; i++) x[i] = (code < 20) ? (i == code ? 1.0 : 0.0) : 1.0;
            }
        }
    orc_invariant_sites(p->codes, T, S, is_inv, inv_char);

    /* P refresh, Particle.cpp:537-548 (forced: every class, node, rate) */
    for (int k = 0; k < K; k++) {
        orc_eigen(p->base, p->logits + 20 * k, eig[k].A, eig[k].lam, eig[k].B, eig[k].pi);
        for (int n = 0; n < nn; n++)
            for (int r = 0; r < R; r++)
                orc_pmatrix(eig[k].A, eig[k].lam, eig[k].B, p->branch_len[n], p->rates[r],
                            P + 400 * (((size_t)k * nn + n) * R + r));
    }

    /* post-order pruning, Particle.cpp:581-626 */
    int ni = internal_postorder(p, order);
    for (int oi = 0; oi < ni; oi++) {
        int n = order[oi];
        double *pN = cl + NS * (size_t)n * node_sp;
        for (size_t i = 0; i < NS * node_sp; i++) pN[i] = 1.0;
        for (int ci = 0; ci < 2; ci++) {
            int d = p->children[2 * (n - T) + ci];
            const double *pD = cl + NS * (size_t)d * node_sp;
            for (int k = 0; k < K; k++)
                for (int r = 0; r < R; r++) {
                    const double *Pm = P + 400 * (((size_t)k * nn + d) * R + r);
                    for (int m = start[k]; m < start[k + 1]; m++) {
                        int c = members[m];
                        const double *x = pD + NS * ((size_t)c * R + r);
                        double *y = pN + NS * ((size_t)c * R + r);
                        for (int i = 0; i < NS; i++) {
                            double w = 0.0;
                            for (int j = 0; j < NS; j++) w += Pm[i * NS + j] * x[j];
                            y[i] *= w;
                        }
                    }
                }
        }
        double *rp = resc + (size_t)n * node_sp;
        for (size_t c = 0; c < node_sp; c++) {
            double *y = pN + NS * c;
            double mx = y[0];
            for (int i = 1; i < NS; i++)
                if (y[i] > mx) mx = y[i];
            for (int i = 0; i < NS; i++) y[i] /= mx;
            rp[c] = log(mx);
        }
    }

    /* root mixture, Particle.cpp:634-674 (rate weights 1: Q4) */
    int root = order[ni - 1];
    const double *pR = cl + NS * (size_t)root * node_sp;
    double lnl = 0.0;
    double inv_scaler = 1.0 - p->pinv;
    double logs[64];
    for (int k = 0; k < K; k++) {
        const double *pi = eig[k].pi;
        for (int m = start[k]; m < start[k + 1]; m++) {
            int c = members[m];
            double mx = -INFINITY;
            for (int r = 0; r < R; r++) {
                const double *x = pR + NS * ((size_t)c * R + r);
                double dot = 0.0;
                for (int i = 0; i < NS; i++) dot += pi[i] * x[i];
                double l = log(dot);
                for (int n = 0; n < nn; n++) l += resc[(size_t)n * node_sp + (size_t)c * R + r];
                logs[r] = l;
                if (l > mx) mx = l;
            }
            double se = 0.0, lf;
            if (!logspace) {
                for (int r = 0; r < R; r++) se += exp(logs[r] - mx);
                double lsite = mx + log(se);
                double fin = exp(lsite) * inv_scaler + p->pinv * pi[inv_char[c]] * is_inv[c];
                lf = log(fin);
            } else {
                /* the same quantity evaluated in log space: -inf and NaN rate terms (a
                 * rate whose partials underflowed, 0/0 in the scaling above) drop out and
                 * log(e^l (1-p) + p pi isInv) = logaddexp(l + log(1-p), log(p pi isInv)) */
                for (int r = 0; r < R; r++)
                    if (logs[r] > -INFINITY) se += exp(logs[r] - mx);
                double a = mx + log(se) + log1p(-p->pinv);
                double b = is_inv[c] != 0.0 ? log(p->pinv * pi[inv_char[c]] * is_inv[c]) : -INFINITY;
                double hi = a > b ? a : b, lo = a > b ? b : a;
                lf = (hi == -INFINITY) ? -INFINITY : hi + log1p(exp(lo - hi));
            }
            if (site_lnl) site_lnl[c] = lf;
            lnl += lf;
        }
    }

    free(cl); free(resc); free(P); free(eig); free(is_inv); free(inv_char);
    free(order); free(cnt); free(start); free(members);
    return lnl;
}

double orc_particle_lnl_sites(const orc_particle *p, double *site_lnl)
{
    return orc_particle_lnl_sites_mode(p, site_lnl, 0);
}

double orc_particle_lnl(const orc_particle *p) { return orc_particle_lnl_sites_mode(p, NULL, 0); }

/* main.cpp:295: one heavy particle per OpenMP thread, dynamic schedule. */
int orc_lnl_batch(const orc_particle *ps, int n, double *lnl_out, int threads)
{
    int used = 1;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
    used = threads;
#pragma omp parallel for schedule(dynamic) num_threads(threads)
#endif
    for (int i = 0; i < n; i++) lnl_out[i] = orc_particle_lnl(&ps[i]);
    return used;
}

/* ------------------------------------------------------------------ a13 */
static double log_gamma_pdf(double k, double theta, double x)
{
    return (k - 1.0) * log(x) - x / theta - lgamma(k) - k * log(theta);
}
static double log_std_normal_pdf(double z) { return -0.5 * z * z - 0.5 * log(2.0 * M_PI); }

/* Particle::lnPrior (Particle.cpp:484-517). */
double orc_ln_prior(const orc_particle *p, double shape, int using_lg, double dpp_alpha)
{
    const int T = p->taxa, nn = 2 * T - 1, K = p->num_classes;
    int root = tree_root(p);
    double lnp = 0.0, len = 0.0;
    for (int n = 0; n < nn; n++)
        if (n != root) len += p->branch_len[n]; /* root length is 0 in the reference */
    lnp += log_gamma_pdf((double)(nn - 1), 1.0 / 10.0, len);
    if (p->num_rates > 1) lnp += log_gamma_pdf(3.0, 2.0, shape);
    if (!using_lg)
        for (int i = 0; i < 190; i++) lnp += log_std_normal_pdf(p->base[i]);
    lnp += log(dpp_alpha) * K;
    int *sz = (int *)calloc(K, sizeof(int));
    for (int s = 0; s < p->sites; s++) sz[p->assignment[s]]++;
    for (int k = 0; k < K; k++) {
        for (int i = 0; i < NS; i++) lnp += log_std_normal_pdf(p->logits[20 * k + i]);
        lnp += lgamma((double)sz[k]);
    }
    free(sz);
    lnp += lgamma(dpp_alpha) - lgamma(dpp_alpha + p->sites);
    return lnp;
}

/* ------------------------------------------------------------------ a16 */
static double ess_at(int n, const double *w, const double *l, double temp, double last, double *tmp)
{
    for (int i = 0; i < n; i++) tmp[i] = w[i] + (temp - last) * l[i];
    double mx = tmp[0];
    for (int i = 1; i < n; i++)
        if (tmp[i] > mx) mx = tmp[i];
    double tot = 0.0;
    for (int i = 0; i < n; i++) {
        tmp[i] = exp(tmp[i] - mx);
        tot += tmp[i];
    }
    double sq = 0.0;
    for (int i = 0; i < n; i++) {
        tmp[i] /= tot;
        sq += tmp[i] * tmp[i];
    }
    return 1.0 / sq;
}

/* computeNextStep (Miscellaneous.hpp:42-109).  target_ess is an int (Q17). */
void orc_compute_next_step(int n, double *raw_weights, const double *raw_lnl,
                           double *normalized, double *ess, int target_ess,
                           double *current_temp, double last_temp)
{
    double *tmp = (double *)malloc(sizeof(double) * n);
    double low = last_temp, high = 1.0;
    while (high - low > 1e-5) {
        double mid = 0.5 * (low + high);
        double e = ess_at(n, raw_weights, raw_lnl, mid, last_temp, tmp);
        if (e > target_ess)
            low = mid;
        else
            high = mid;
    }
    *current_temp = high;
    for (int i = 0; i < n; i++) raw_weights[i] = raw_weights[i] + (high - last_temp) * raw_lnl[i];
    double mx = raw_weights[0];
    for (int i = 1; i < n; i++)
        if (raw_weights[i] > mx) mx = raw_weights[i];
    double tot = 0.0;
    for (int i = 0; i < n; i++) {
        normalized[i] = exp(raw_weights[i] - mx);
        tot += normalized[i];
    }
    double sq = 0.0;
    for (int i = 0; i < n; i++) {
        normalized[i] /= tot;
        sq += normalized[i] * normalized[i];
    }
    *ess = 1.0 / sq;
    free(tmp);
}

/* ------------------------------------------------------------------ a17 */
/* main.cpp:179-195. */
void orc_systematic_resample(int n, const double *normalized, double u, int32_t *assignments)
{
    double cumulative = normalized[0];
    int k = 0, cnt = 0;
    double ruler = u / (double)n;
    double step = 1.0 / (double)n;
    while (cnt < n) {
        if (ruler <= cumulative) {
            assignments[cnt++] = k;
            ruler += step;
        } else {
            if (++k >= n) {
                k = n - 1;
                cumulative = 1.0;
            } else
                cumulative += normalized[k];
        }
    }
}

/* ------------------------------------------------------------------ a15 */
/* Single-site pruning under each candidate profile (Particle.cpp:857-941),
 * with the tip read at the site's own column for every rate (the reference
 * reads column `randomSite` of the flattened [site][rate] array: Q15(i)) and
 * the rescale read with the stride it was written with (Q15(ii)). */
void orc_site_class_lnl(const orc_particle *p, int site, int num_cand,
                        const double *cand_logits, double tempering, double *out)
{
    const int T = p->taxa, S = p->sites, R = p->num_rates, nn = 2 * T - 1;
    double *cl = (double *)malloc(sizeof(double) * NS * (size_t)R * nn);
    double *resc = (double *)calloc((size_t)R * nn, sizeof(double));
    int *order = (int *)malloc(sizeof(int) * T);
    int ni = internal_postorder(p, order);
    int root = order[ni - 1];
    double is_inv;
    int32_t inv_char;
    {
        int c1 = p->codes[site];
        is_inv = 1.0;
        for (int t = 0; t < T; t++)
            if (p->codes[(size_t)t * S + site] != c1) is_inv = 0.0;
        inv_char = (is_inv == 1.0 && c1 < 20) ? c1 : 0;
    }
    for (int c = 0; c < num_cand; c++) {
        orc_class_eig e;
        orc_eigen(p->base, cand_logits + 20 * c, e.A, e.lam, e.B, e.pi);
        for (int t = 0; t < T; t++) {
            int code = p->codes[(size_t)t * S + site];
            for (int r = 0; r < R; r++)
                for (int i = 0; i < NS; i++)
                    cl[NS * ((size_t)t * R + r) + i] = (code < 20) ? (i == code ? 1.0 : 0.0) : 1.0;
        }
        memset(resc, 0, sizeof(double) * R * nn);
        for (int oi = 0; oi < ni; oi++) {
            int n = order[oi];
            for (int r = 0; r < R; r++) {
                double *y = cl + NS * ((size_t)n * R + r);
                for (int i = 0; i < NS; i++) y[i] = 1.0;
                for (int ci = 0; ci < 2; ci++) {
                    int d = p->children[2 * (n - T) + ci];
                    double Pm[400];
                    orc_pmatrix(e.A, e.lam, e.B, p->branch_len[d], p->rates[r], Pm);
                    const double *x = cl + NS * ((size_t)d * R + r);
                    for (int i = 0; i < NS; i++) {
                        double w = 0.0;
                        for (int j = 0; j < NS; j++) w += Pm[i * NS + j] * x[j];
                        y[i] *= w;
                    }
                }
                double mx = y[0];
                for (int i = 1; i < NS; i++)
                    if (y[i] > mx) mx = y[i];
                for (int i = 0; i < NS; i++) y[i] /= mx;
                resc[(size_t)n * R + r] = log(mx);
            }
        }
        double logs[64], mx = -INFINITY;
        for (int r = 0; r < R; r++) {
            const double *x = cl + NS * ((size_t)root * R + r);
            double dot = 0.0;
            for (int i = 0; i < NS; i++) dot += e.pi[i] * x[i];
            double l = log(dot);
            for (int n = 0; n < nn; n++) l += resc[(size_t)n * R + r];
            logs[r] = l;
            if (l > mx) mx = l;
        }
        double se = 0.0;
        for (int r = 0; r < R; r++) se += exp(logs[r] - mx);
        double lsite = mx + log(se);
        double fin = exp(lsite) * (1.0 - p->pinv) + p->pinv * e.pi[inv_char] * is_inv;
        out[c] = log(fin) * tempering;
    }
    free(cl);
    free(resc);
    free(order);
}
