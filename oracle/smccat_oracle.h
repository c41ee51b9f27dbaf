/*
 * smccat_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the Smc-CAT particle-likelihood path, used as the
 * checker for the HIP engine and as the timed CPU baseline.  Nothing in the
 * product library (libsmccat_hip.so) links, loads or calls this code; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may do so.
 *
 * Every function states the reference file:line it restates (paths relative
 * to the Smc-CAT source tree).  Parity pins are listed in the test suite:
 * scipy golden vectors (expm, gammaincinv, gamma/normal log-pdfs), brute-force
 * ancestral-state enumeration on tiny trees, closed forms and invariants.
 */
#ifndef SMCCAT_ORACLE_H
#define SMCCAT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One particle's state in the reference's own terms.
 * Node ids follow Tree.cpp:62-67,138-157: tips 0..T-1 (= alignment row),
 * internal nodes T..2T-2; the root is the internal node that is nobody's child. */
typedef struct {
    int taxa;                    /* T */
    int sites;                   /* S (numChar, no pattern compression: Q7) */
    int num_rates;               /* K_r */
    const int8_t *codes;         /* [T][S] alignment codes 0..22 (Alignment.cpp:7-31) */
    const int32_t *children;     /* [T-1][2] children of internal node T+i */
    const double *branch_len;    /* [2T-1] indexed by node id (root entry unused) */
    const double *base;          /* [190] log-exchangeabilities, RateMatrices.hpp:97-111 order */
    int num_classes;             /* K CAT classes */
    const double *logits;        /* [K][20] stationary logits */
    const int32_t *assignment;   /* [S] class of every site */
    double pinv;                 /* proportion of invariant sites */
    const double *rates;         /* [K_r] discrete-gamma rates */
} orc_particle;

/* a1: index c of the 190-vector <-> (i,j), i=1..19, j=0..i-1 (RateMatrices.hpp:97-111). */
void orc_lower_triangle_coord(int c, int *i, int *j);

/* a3+a4: pi = softmax(logits) (TransitionProbabilityClass.cpp:101-109); Q built from
 * exp(base) (TransitionProbabilityClass.cpp:50-75).  Q row-major [i*20+j]. */
void orc_build_q(const double base[190], const double logits[20], double pi[20], double Q[400]);

/* a4: eigen-decomposition of the reversible Q through its symmetric form.
 * Q = A diag(lam) B with A = Pi^{-1/2} U, B = U^T Pi^{1/2}.  Row-major. */
void orc_eigen(const double base[190], const double logits[20],
               double A[400], double lam[20], double B[400], double pi[20]);

/* a5: P(t*r) = Re(V diag(exp(lam t r)) V^{-1}) (TransitionProbabilityClass.cpp:83-95).
 * P[i*20+j] = Pr(child state j | parent state i). */
void orc_pmatrix(const double A[400], const double lam[20], const double B[400],
                 double t, double r, double P[400]);

/* Regularised lower incomplete gamma P(a,x) and its inverse (bisection). */
double orc_gamma_p(double a, double x);
double orc_gamma_p_inv(double a, double p);

/* a6: discretizeGamma (Miscellaneous.hpp:22-37). */
void orc_discretize_gamma(double shape, int num, double *out);

/* a7: invariant-site detection (Particle.cpp:41-73). */
void orc_invariant_sites(const int8_t *codes, int taxa, int sites,
                         double *is_invariant, int32_t *invariant_char);

/* a8-a10: Particle::refreshLikelihood(true) (Particle.cpp:530-675) including the
 * tip encoding of Particle.cpp:41-73.  Returns the particle log-likelihood. */
double orc_particle_lnl(const orc_particle *p);

/* Same, also returning per-site log-likelihood contributions site_lnl[S]
 * (the log(finalSiteLike) of Particle.cpp:667-670). */
double orc_particle_lnl_sites(const orc_particle *p, double *site_lnl);

/* logspace = 1: the same mixture evaluated in log space (finite where
 * exp(logSiteLike) of Particle.cpp:667 underflows; a rate term that is -inf or NaN,
 * i.e. whose partials underflowed, adds 0 to the sum over rates); logspace = 0 is
 * the above. */
double orc_particle_lnl_sites_mode(const orc_particle *p, double *site_lnl, int logspace);

/* Batch of particles, one particle per OpenMP thread (main.cpp:295 pattern).
 * threads<=0 uses the OpenMP default.  Returns the thread count used. */
int orc_lnl_batch(const orc_particle *ps, int n, double *lnl_out, int threads);

/* a13: Particle::lnPrior (Particle.cpp:484-517).  using_lg: no GTR prior term.
 * class_sizes[K] member counts. */
double orc_ln_prior(const orc_particle *p, double shape, int using_lg, double dpp_alpha);

/* a16: computeNextStep (Miscellaneous.hpp:42-109). */
void orc_compute_next_step(int n, double *raw_weights, const double *raw_lnl,
                           double *normalized, double *ess, int target_ess,
                           double *current_temp, double last_temp);

/* a17: systematic resampling (main.cpp:179-195). */
void orc_systematic_resample(int n, const double *normalized, double u, int32_t *assignments);

/* a15 (corrected Algorithm-8 form of Particle.cpp:857-941): tempered site
 * log-likelihood of `site` under each candidate profile.  cand_logits[C][20]
 * are candidate stationary logits (existing classes and auxiliaries);
 * out[C] = tempering * log L_site(c).  Uses the correct per-rate tip index
 * (Q15(i) fixed) and the correct rescale stride (Q15(ii) fixed). */
void orc_site_class_lnl(const orc_particle *p, int site, int num_cand,
                        const double *cand_logits, double tempering, double *out);

#ifdef __cplusplus
}
#endif
#endif
