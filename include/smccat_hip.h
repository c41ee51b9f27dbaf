/*
 * smccat_hip.h -- C-ABI of the MI355X particle-likelihood engine (libsmccat_hip.so).
 *
 * The reference (Smc-CAT) is one C++ executable with no FFI; its seam is the
 * C++ Particle API (src/analysis/Particle.hpp:51-83) plus the free SMC
 * functions of src/core/Miscellaneous.hpp.  Each entry point below names the
 * reference interface it replaces.  The C++ facade in
 * smc-cat_amd/csrc/host/particle.hpp keeps the reference's call shapes on
 * top of these functions.
 *
 * Conventions: every call returns an int status (SMC_OK == 0) and leaves a
 * message in smc_engine_last_error(); nothing throws across the ABI.  Host
 * buffers are copied in or out before the call returns; the engine never
 * keeps a host pointer.  One host thread per engine (the reference's
 * Particle is thread-confined in the same way, main.cpp:295-298).
 */
#ifndef SMCCAT_HIP_H
#define SMCCAT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SMC_OK = 0,
    SMC_ERR_ARG = 1,    /* invalid argument or size */
    SMC_ERR_HIP = 2,    /* HIP runtime error (message in last_error) */
    SMC_ERR_NOMEM = 3,  /* device allocation failed */
    SMC_ERR_STATE = 4   /* call order violated (e.g. no alignment yet) */
};

enum { SMC_PREC_FP64 = 0, SMC_PREC_FP32 = 1 };

/* Kernel families reported by smc_engine_kernel_stats. */
enum {
    SMC_K_EIGEN = 0,   /* 20x20 eigen-decomposition per (particle, class)    */
    SMC_K_PMAT = 1,    /* P(t) back-transform per (particle, class, node, rate) */
    SMC_K_PRUNE = 2,   /* level-ordered pruning, root mixture fused          */
    SMC_K_REDUCE = 3,  /* per-particle log-likelihood reduction              */
    SMC_K_SMC = 4,     /* reweight / ESS bisection / resampling              */
    SMC_K_SITE = 5,    /* single-site x candidate-profile scoring            */
    SMC_K_COUNT = 6
};

typedef struct smc_engine smc_engine;

typedef struct {
    int32_t device;        /* HIP device ordinal                                     */
    int32_t taxa;          /* T: Alignment::getNumTaxa (Alignment.hpp)               */
    int32_t sites;         /* S: numChar, one pattern per column (Particle.cpp:22)   */
    int32_t num_rates;     /* K_r: Particle::numRates (Particle.hpp:101)             */
    int32_t max_particles; /* resident particle slots                                */
    int32_t max_classes;   /* CAT classes per particle (1 without CAT)               */
    int32_t precision;     /* SMC_PREC_FP64 (plain) or SMC_PREC_FP32 (fast path)     */
    int32_t partial_slots; /* particles whose partials are resident at once; a sweep
                              over more runs in waves (0: max_particles)             */
    int64_t pmat_classes;  /* classes whose P(t) are resident at once
                              (0: partial_slots * max_classes)                       */
    int32_t workspaces;    /* partial workspaces per particle: 2 enables the
                              incremental refresh with accept/reject (0 or 1: one)    */
    int32_t reserved;      /* must be 0                                              */
} smc_engine_config;

/* Engine lifetime.  Replaces the per-thread Particle construction of
 * main.cpp:112-115 / Particle::Particle (Particle.cpp:21-40): all device
 * buffers are allocated here -- state for max_particles particles, partial
 * likelihoods for partial_slots of them, P(t) for pmat_classes classes.  A full
 * sweep in several waves later adds a second P(t) buffer of the same size, once,
 * when the device keeps 16 GB free after it, to overlap the next wave's P with the
 * current wave's pruning. */
int smc_engine_create(const smc_engine_config *cfg, smc_engine **out);
int smc_engine_destroy(smc_engine *e);
const char *smc_engine_last_error(const smc_engine *e);

/* Alignment upload: codes[taxon][site] in the Alignment code table
 * (Alignment.cpp:7-31: 0..19 amino acids, 20 '-', 21 'X', 22 '?').
 * Replaces the tip encoding and invariant detection of Particle.cpp:41-73. */
int smc_set_alignment(smc_engine *e, const int8_t *codes);

/* Particle state upload for slots [first, first+count).  Replaces
 * Particle::copyFromSerialized (Particle.cpp:218-300) on already-parsed state:
 *   children    [count][T-1][2]  children of internal node T+i (Tree.cpp:62-67 ids)
 *   branch_len  [count][2T-1]    by node id; the root entry is ignored
 *   base        [count][190]     log-exchangeabilities (RateMatrices.hpp:97-111 order)
 *   num_classes [count]          K per particle, 1..max_classes
 *   logits      [count][max_classes][20]  stationary logits per class
 *   assignment  [count][S]       class index of every site
 *   pinv        [count]          proportion of invariant sites
 *   shape       [count]          gamma shape; rates = discretizeGamma(shape, K_r)
 *                                (Miscellaneous.hpp:22-37), 1.0 when K_r == 1 */
int smc_set_particles(smc_engine *e, int first, int count, const int32_t *children,
                      const double *branch_len, const double *base, const int32_t *num_classes,
                      const double *logits, const int32_t *assignment, const double *pinv,
                      const double *shape);

/* Full likelihood of slots [first, first+count): Particle::refreshLikelihood(true)
 * followed by Particle::lnLikelihood() (Particle.cpp:530-675, 480-482).
 * lnl_out[count] and nonfinite_out[count] may be NULL.  nonfinite_out[i] = 1
 * when the particle's lnL is NaN/Inf (the reference's NaN guard, main.cpp:158-174). */
int smc_refresh_likelihood(smc_engine *e, int first, int count, double *lnl_out,
                           uint8_t *nonfinite_out);

/* Incremental likelihood with accept/reject (engines created with workspaces = 2).
 * The slots' partials must be resident together: count <= partial_slots and
 * their classes fit pmat_classes.
 *
 * smc_update_params: new topology / branch lengths / log-exchangeabilities /
 * pInvar / gamma shape for slots [first, first+count) (classes untouched) --
 * what a tree, rate-matrix, shape or invariant move changes (Particle.cpp:677-783,
 * Tree.cpp:553-811).
 *
 * smc_refresh_incremental: Particle::refreshLikelihood(false) (Particle.cpp:530-675).
 * node_flags[count][2T-1]: bit 0 = updateCL, bit 1 = updateTP (Tree.hpp:26-27);
 * eig_flags[count] (nullable): the exchangeabilities changed, so every class is
 * re-decomposed and every P recomputed.  Dirty internal nodes (flagged, parents
 * of changed branches, and their ancestors) write their other workspace
 * (Particle.cpp:566-571); the rest keep theirs; P is recomputed for particles
 * with any changed branch.
 *
 * smc_accept_reject: Particle::accept / reject (Particle.cpp:343-478) per slot:
 * accept[i] != 0 commits flags, lnL and parameters; 0 returns to the last
 * committed state (flags, lnL, parameters, eigensystems and P).  A forced
 * refresh (smc_refresh_likelihood) also commits. */
int smc_update_params(smc_engine *e, int first, int count, const int32_t *children,
                      const double *branch_len, const double *base, const double *pinv,
                      const double *shape);
int smc_refresh_incremental(smc_engine *e, int first, int count, const uint8_t *node_flags,
                            const uint8_t *eig_flags, double *lnl_out, uint8_t *nonfinite_out);
int smc_accept_reject(smc_engine *e, int first, int count, const uint8_t *accept);

/* The resampled copies of main.cpp:300,309,313 (each slot deserialises its parent's
 * record, Particle::copyFromSerialized, Particle.cpp:218-276) on the device: slot i of
 * [0, n) takes the state of slot parents[i] -- tree, branch lengths, exchangeabilities,
 * classes, allocation, pInvar, shape, eigensystems and committed lnL -- by gathers of
 * device rows, without an upload (0 <= parents[i] < n, any order, repeats allowed).
 * Copies are exact (the reference rounds branch lengths to 6 decimals through Newick
 * text, Q11).  A slot that keeps its own state (parents[i] == i) keeps its resident
 * partial likelihoods, so rejuvenation recomputes only the copied slots. */
int smc_copy_particles(smc_engine *e, int n, const int32_t *parents);

typedef struct smc_mlp smc_mlp;  /* learned CAT-allocation proposal (below) */

/* Rejuvenation settings (main.cpp:54-105, Particle.hpp:87-94, Settings.hpp:19-29). */
typedef struct {
    int32_t iterations;          /* MCMC iterations r per particle (Settings::rejuvenationIterations) */
    int32_t using_lg;            /* LG particles: no rate-matrix moves and no GTR prior           */
    int32_t invariant;           /* propose pInvar moves                                          */
    int32_t reserved;            /* must be 0                                                     */
    uint64_t seed;               /* Philox stream key (streams per slot)                          */
    double scale_delta;          /* Particle::scaleDelta (1.0)                                    */
    double subtree_scale_delta;  /* Particle::subtreeScaleDelta (1.0)                             */
    double shape_delta;          /* Particle::shapeDelta (1.0)                                    */
    double rate_delta;           /* Particle::rateMatrixDelta (0.25)                              */
    double rate_full_scale;      /* diagonal of Particle::rateMatrixCholesky (0.5)                */
    double invar_alpha;          /* Particle::invarAlpha (100)                                    */
    double gibbs_probability;    /* Settings::alg8Probability (0.1); needs CAT to matter          */
    double dpp_alpha;            /* CRP concentration (Particle.hpp:106: 0.01)                    */
    const double *rate_cholesky; /* [190][190] row-major lower factor of the full rate-matrix move
                                    (Particle::rateMatrixCholesky, Particle.cpp:710-716); NULL:
                                    rate_full_scale * identity                                    */
    smc_mlp *allocation_proposal;/* non-NULL: the CAT move is the learned Metropolis-Hastings move
                                    (smc_learned_partition_move) with this network, or, with
                                    record_allocations, the Gibbs move recording its draws into
                                    the network's examples; NULL: the Gibbs move                 */
    int32_t record_allocations;
    int32_t reserved2;           /* must be 0                                                     */
    const int32_t *stream_keys;  /* [count] (nullable): the random streams of slot first+i are keyed
                                    by stream_keys[i] instead of by the slot; the move types of an
                                    iteration come from a stream of their own, so a key fixes its
                                    chain's number of MH steps (smc_rejuvenation_steps)           */
} smc_rejuvenation_options;

/* Rejuvenation of slots [first, first+count) at tempering T, replacing the
 * per-particle loop of main.cpp:296-311: a forced refresh of the (copied)
 * particles, `iterations` MCMC iterations (Mcmc::operator(), Mcmc.cpp:25-73:
 * move by weight -- tree 3, shape 1, invariant 1, rate matrix 1 -- repeated
 * T*0.25 / 2 / 2 / 5 times, each repetition a Metropolis-Hastings step on
 * T lnL + lnPrior with the move's Hastings term), then with probability
 * gibbs_probability the CAT Gibbs move and a forced refresh.  All particles
 * step in lockstep: one incremental refresh evaluates every proposal of a
 * round.  Needs workspaces = 2 and the range resident (count <= partial_slots).
 * lnl_out[count] (nullable): final lnL; stats_out[8] (nullable): proposals and
 * acceptances per move kind (tree, shape, invariant, rate matrix). */
int smc_rejuvenate(smc_engine *e, int first, int count, double tempering,
                   const smc_rejuvenation_options *opt, double *lnl_out, int64_t *stats_out);

/* The number of MH steps (proposals) smc_rejuvenate will make for each of `count`
 * stream keys under these options: the sum over its iterations of the drawn move's
 * repetitions.  No counterpart in the reference: lockstep batches of chains with
 * similar step counts finish together (smc_run_sampler assigns keys to slots so). */
int smc_rejuvenation_steps(smc_engine *e, const smc_rejuvenation_options *opt, int count,
                           const int32_t *keys, int32_t *steps_out);

/* The SMC sampler loop of main.cpp:143-336 over slots [0, num_particles),
 * which must hold the initial population (main.cpp:129-139): per iteration
 * computeNextStep, and when ESS <= threshold * N (and T < 1) systematic
 * resampling, proposal tuning from the resampled population (main.cpp:197-289),
 * the copy of each parent into its new slot (on the engine, exact values), and
 * the rejuvenation of every slot (smc_rejuvenate) in lockstep batches of
 * `batch` particles.  With batch > 1 the CAT moves of all batches run in one
 * lockstep after the last batch, followed by one refresh of the population (the
 * particles end the same: each slot has its own random streams).  Differences: copies are exact (the reference rounds
 * branch lengths to 6 decimals through Newick text, Tree.cpp:429,433); split
 * posteriors and consensus (main.cpp:203-207,338-425) are not computed. */
typedef struct {
    int32_t num_particles;          /* N (Settings::numParticles)                                */
    int32_t rejuvenation_iterations;/* r (Settings::rejuvenationIterations)                      */
    double rejuvenation_threshold;  /* resample when ESS <= threshold * N (0.6)                  */
    double alg8_probability;        /* CAT Gibbs move probability per particle (0.1)             */
    int32_t using_lg;               /* -LG                                                       */
    int32_t invariant;              /* -I                                                        */
    int32_t cat;                    /* CAT classes (Gibbs move enabled)                          */
    int32_t batch;                  /* particles per lockstep rejuvenation batch; 0 = as many as
                                       the engine's partial slots hold; 1 = one particle at a
                                       time, the reference's per-thread shape                    */
    int32_t max_iterations;         /* stop after this many SMC iterations; 0 = until T = 1      */
    int32_t reserved;               /* must be 0                                                 */
    uint64_t seed;                  /* Philox key of the sampler                                 */
    double dpp_alpha;               /* CRP concentration                                         */
    smc_mlp *allocation_proposal;   /* non-NULL: the learned CAT move; the first
                                       proposal_warmup resampling steps run the Gibbs move and
                                       record its draws, the network is trained on them
                                       (train_epochs Adam epochs at train_lr) and later steps
                                       use the learned move                                      */
    int32_t proposal_warmup;
    int32_t train_epochs;
    double train_lr;
    /* called after every computeNextStep (nullable): iteration, tempering, ESS */
    void (*on_iteration)(void *user, int32_t iteration, double temp, double ess);
    void *user;
} smc_sampler_options;

typedef struct {
    int32_t iterations;             /* SMC iterations run                                        */
    int32_t resamples;              /* resample + rejuvenation steps                             */
    double final_temp;              /* tempering reached                                         */
    double final_ess;               /* ESS of the last computeNextStep (main.cpp:154-156)        */
    double ess_sum;                 /* sum of the ESS after each reweight                        */
    double seconds;                 /* wall time of the loop, initial likelihoods included       */
    double rejuvenation_seconds;    /* of which rejuvenation                                     */
    double training_seconds;        /* of which training the proposal network                    */
    int64_t move_stats[8];          /* proposed / accepted: tree, shape, invariant, rate matrix  */
    int64_t cat_stats[3];           /* CAT move: site updates, site moves, Gibbs examples kept   */
} smc_sampler_result;

/* lnl_out[N], weights_out[N] (nullable): final raw lnL and raw log weights. */
int smc_run_sampler(smc_engine *e, const smc_sampler_options *opt, smc_sampler_result *result,
                    double *lnl_out, double *weights_out);

/* ---- Learned CAT-allocation proposal (no counterpart in the reference) ----
 * A 40 -> 128 -> 128 (GELU) network scores the candidates of Neal's Algorithm 8
 * for one site (the context of Particle::gibbsPartitionMove, Particle.cpp:785-995):
 *   x = [site amino-acid frequencies (20), size-weighted mean class profile (20)],
 *   l_c = h2 . (U log pi_c + u), l_new = h2 . v + b_new,
 *   logits tau * l_c + log n_c and tau * l_new + log alpha over K + 1 slots.
 * Trained on the GPU (hand-written forward, backward and Adam kernels) on the
 * Gibbs move's own draws; used by smc_learned_partition_move as a
 * Metropolis-Hastings proposal under the exact site likelihood.
 * Example layout: header[44] = x[40], tau, log alpha, 0, 0; class entries[24] =
 * log pi[20], log n, 0, 0, 0; label in [0, K] (K = a new class). */
int smc_mlp_num_params(void);
int smc_mlp_create(smc_engine *e, uint64_t seed, smc_mlp **out);
int smc_mlp_destroy(smc_mlp *m);
int smc_mlp_get_params(smc_mlp *m, float *theta);
int smc_mlp_set_params(smc_mlp *m, const float *theta);   /* resets the optimiser state */
/* Appends n examples; cls_off[i] indexes the class entries cls[n_cls][24]. */
int smc_mlp_add_examples(smc_mlp *m, int n, const float *header, const int32_t *num_classes,
                         const int32_t *label, const int64_t *cls_off, const float *cls, int64_t n_cls);
/* Runs the Gibbs allocation move (smc_gibbs_partition_move) on the slots and
 * records every draw as an example. */
int smc_mlp_collect_gibbs(smc_mlp *m, int count, const int32_t *slots, double tempering, uint64_t seed,
                          int num_iter, double dpp_alpha, int64_t *stats_out);
int64_t smc_mlp_num_examples(smc_mlp *m);
int smc_mlp_clear_examples(smc_mlp *m);
/* header[44], *num_classes, *label, cls[K][24] of example i (any may be NULL). */
int smc_mlp_get_example(smc_mlp *m, int64_t i, float *header, int32_t *num_classes, int32_t *label,
                        float *cls);
/* Mean cross-entropy of examples [first, first+n) and its gradient (no update). */
int smc_mlp_loss_grad(smc_mlp *m, int first, int n, double *loss, float *grad);
/* log q over the K + 1 slots of examples [first, first+n): logq[n][out_stride]. */
int smc_mlp_infer(smc_mlp *m, int first, int n, int out_stride, float *logq);
/* Adam epochs over all examples (mini-batches of `batch`, reshuffled each
 * epoch); epoch_loss[epochs] (nullable) receives each epoch's mean loss. */
int smc_mlp_train(smc_mlp *m, int epochs, int batch, double lr, uint64_t seed, double *epoch_loss);
/* The learned allocation move on the engine's slots: num_iter (0: S/2) random
 * sites per particle, each one Metropolis-Hastings step of Algorithm 8 (one
 * auxiliary class) with the network's proposal, tempered by
 * proposal_tempering (< 0: the target's), accepted under the exact tempered
 * site likelihood.  stats_out[5]: site updates, proposals of another slot,
 * accepted, classes created, classes erased. */
int smc_learned_partition_move(smc_mlp *m, int count, const int32_t *slots, double tempering,
                               uint64_t seed, int num_iter, double dpp_alpha, double proposal_tempering,
                               int64_t *stats_out);

/* Launch shape of the full sweep (no counterpart in the reference: the
 * reference runs one particle per OpenMP thread).  The pruning launches of a
 * sweep are persistent grid-stride kernels (blocks_per_cu workgroups per
 * compute unit) that read their item ranges from the device, so a whole sweep
 * (eigen, P, every level bucket, roots, lnL reduction) is captured once as a
 * hipGraph and replayed while the waves and level counts stay the same.
 * unroll[k] is the number of 16-site column blocks a wavefront carries at
 * once for items with k children read as stored partials (tips and cherries,
 * internal nodes with two tip children, are built in registers). */
typedef struct {
    int32_t blocks_per_cu;   /* 1..64 (default 32)         */
    int32_t unroll[3];       /* 1, 2 or 4 (default 1, 1, 2 for FP64, 2, 2, 2 for FP32) */
    double ms_per_sweep;     /* autotune: best measured     */
    int32_t configs_tried;   /* autotune: candidates timed  */
    int32_t xcd_groups;      /* 1, 2, 4 or 8 (default 8): the workgroups of one item are
                                kept on one XCD (workgroup i runs on XCD i mod 8), so the
                                P matrices and codes it reads are fetched into one L2 */
    int32_t unroll_runtime;  /* 1, 2 or 4 (default 2): unroll of the launches that decide the
                                child modes per item (roots, and the merged small buckets of
                                incremental refreshes) */
    int32_t chain_max_nodes; /* >= 0 (default 48): an incremental refresh runs a particle's
                                dirty nodes as chains (SMC_PRUNE_MFMA) when at most this many
                                of them are stored partials and the call's moving particles
                                are too few to fill per-level launches (8 workgroups per CU:
                                one-particle rounds, the tail of a lockstep batch); else
                                level by level */
} smc_tuning;

int smc_engine_get_tuning(smc_engine *e, smc_tuning *t);
int smc_engine_set_tuning(smc_engine *e, const smc_tuning *t);
/* Times candidate launch shapes on full sweeps of slots [first, first+count)
 * (coordinate descent over blocks_per_cu, then each unroll; `sweeps` replays per
 * candidate) and keeps the fastest.  Recomputes those particles' lnL. */
int smc_engine_autotune(smc_engine *e, int first, int count, int sweeps, smc_tuning *out);
/* enable = 0 runs sweeps as eager launches (the default captures graphs);
 * per-family kernel timing also runs eagerly. */
int smc_engine_set_graphs(smc_engine *e, int enable);
int smc_engine_graph_stats(smc_engine *e, int64_t *captures, int64_t *launches);

/* Host wall time spent in each entry point (diagnostics; not reset by
 * smc_engine_reset_stats).  ms[SMC_PHASE_COUNT], calls[SMC_PHASE_COUNT]; either
 * may be NULL.  Nested calls count in both phases. */
enum {
    SMC_PHASE_SET_PARTICLES = 0,
    SMC_PHASE_REFRESH = 1,
    SMC_PHASE_UPDATE_PARAMS = 2,
    SMC_PHASE_REFRESH_INCREMENTAL = 3,
    SMC_PHASE_ACCEPT_REJECT = 4,
    SMC_PHASE_PROPOSE = 5,      /* host proposals and priors of smc_rejuvenate   */
    SMC_PHASE_CAT_MOVE = 6,     /* Gibbs or learned allocation moves              */
    SMC_PHASE_SITE_CLASS = 7,
    SMC_PHASE_COUNT = 8
};
int smc_engine_phase_stats(smc_engine *e, double *ms, int64_t *calls);
int smc_engine_reset_phase_stats(smc_engine *e);

/* Particle::lnPrior (Particle.cpp:484-517) of an uploaded slot: Gamma(nodes-1,
 * scale 0.1) on the tree length, Gamma(3, 2) on the shape when K_r > 1, N(0,1)
 * on the 190 exchangeabilities unless using_lg, and the CRP/class terms. */
int smc_ln_prior(smc_engine *e, int slot, int using_lg, double dpp_alpha, double *out);

/* computeNextStep (Miscellaneous.hpp:42-109) on the GPU.  raw_weights is
 * updated in place, normalized[n] and *ess / *current_temp are written. */
int smc_compute_next_step(smc_engine *e, int n, double *raw_weights, const double *raw_lnl,
                          double *normalized, double *ess, int target_ess,
                          double *current_temp, double last_temp);

/* No counterpart in the reference: smc_compute_next_step for callers that cannot
 * allocate device memory (tight sampler loops, harnesses that forbid hipMalloc).
 * Same arguments and results, bit for bit; its device staging is a fixed array
 * of the library (one per GPU, held until the process ends) instead of engine
 * buffers grown on demand, so it never calls hipMalloc or hipFree.  n at most
 * 65536 (SMC_ERR_ARG above).  Calls on one GPU share that array: calls that
 * overlap on one GPU (two engines on two streams or threads) race. */
int smc_compute_next_step_noalloc(smc_engine *e, int n, double *raw_weights, const double *raw_lnl,
                                  double *normalized, double *ess, int target_ess,
                                  double *current_temp, double last_temp);

/* Systematic resampling (main.cpp:179-195) on the GPU: u in [0,1). */
int smc_systematic_resample(smc_engine *e, int n, const double *normalized, double u,
                            int32_t *assignments);

/* Tempered single-site log-likelihoods under candidate CAT profiles: the
 * per-site, per-class conditional of Particle::gibbsPartitionMove
 * (Particle.cpp:857-941, with the per-rate tip index and rescale stride of
 * Q15 corrected), batched over entries.  Entry i scores alignment column
 * sites[i] of particle slot particles[i] (its tree, branch lengths, base
 * matrix, rates and pInvar) under the ncand[i] <= n_cand profiles
 * cand_logits[i][c][20]:
 *   out[i][c] = tempering * log L_site(profile c)   (NaN for c >= ncand[i]).
 * The root mixture follows smc_engine_set_mixture. */
int smc_site_class_lnl(smc_engine *e, int count, const int32_t *particles, const int32_t *sites,
                       int n_cand, const int32_t *ncand, const double *cand_logits,
                       double tempering, double *out);

/* Particle::gibbsPartitionMove(tempering) (Particle.cpp:785-995) for count
 * particle slots at once: Neal's Algorithm 8 with m = 5 auxiliary profiles
 * drawn from the logit prior, num_iter random-site updates (<= 0: numChar/2,
 * :791), CRP concentration dpp_alpha (the reference fixes 0.01,
 * Particle.hpp:106).  Each iteration scores one site per particle under all
 * its classes and auxiliaries in one batched smc_site_class_lnl launch.  An
 * emptied class is erased (:817-821).  Auxiliaries are capped so a particle
 * never exceeds max_classes.  Random numbers: Philox4x32-10 keyed by
 * (seed, slot).  The slots are re-uploaded with their new allocations;
 * stats_out[3] (nullable) = sites moved, classes created, classes erased. */
int smc_gibbs_partition_move(smc_engine *e, int count, const int32_t *slots, double tempering,
                             uint64_t seed, int num_iter, double dpp_alpha, int64_t *stats_out);

/* Current allocation state of a slot (Particle::getAssignments / getCategories,
 * Particle.hpp:73-74): num_classes, logits[max_classes][20], assignment[S]. */
int smc_get_particle_state(smc_engine *e, int slot, int32_t *num_classes, double *logits,
                           int32_t *assignment);

/* Current (last uploaded or committed) parameters of a slot: children[T-1][2],
 * branch_len[2T-1], base[190], pinv, shape (what writeToSerialized reads,
 * Particle.cpp:302-320).  Any pointer may be NULL. */
int smc_get_particle_params(smc_engine *e, int slot, int32_t *children, double *branch_len,
                            double *base, double *pinv, double *shape);

/* discretizeGamma (Miscellaneous.hpp:22-37): out[num] mean-one gamma rates. */
int smc_discretize_gamma(double shape, int num, double *out);

/* The LG exchangeabilities as the reference's LG particle stores them
 * (RateMatrices.hpp:7-95 read in the order of :97-111, Particle.cpp:84-93). */
void smc_lg_base_matrix(double out[190]);

/* Measurement (no reference counterpart; the reference only has TIME_PROFILE
 * prints, Particle.cpp:533-631): when enabled, every kernel launch is bracketed
 * by HIP events on the engine's stream and its duration accumulated per family. */
int smc_engine_set_timing(smc_engine *e, int enabled);
int smc_engine_kernel_stats(smc_engine *e, int family, int64_t *launches, double *total_ms);
int smc_engine_reset_stats(smc_engine *e);

/* Kernel variant of a family (tuning; no reference counterpart).  For
 * SMC_K_PRUNE: SMC_PRUNE_MFMA (matrix cores, the default) or SMC_PRUNE_VALU
 * (vector FMAs with the P matrix as scalar operands).  Results agree to rounding.
 * SMC_PRUNE_MFMA runs the short dirty paths of an incremental refresh as chains
 * (one launch walks each path bottom-up with the partial in registers) when few
 * particles move in the call (see smc_tuning.chain_max_nodes), and walks them all
 * in one launch when many do (whole trees go level by level);
 * SMC_PRUNE_MFMA_LEVELS launches every dirty node level by level instead, with the
 * same partials bit for bit.  SMC_PRUNE_MFMA_WALK runs every changed particle of an
 * incremental refresh in one launch: each workgroup walks one tile of a particle's
 * dirty nodes in post-order with the partial carried in registers, the root
 * mixture included (same partials and lnL bit for bit). */
enum { SMC_PRUNE_MFMA = 0, SMC_PRUNE_VALU = 1, SMC_PRUNE_MFMA_LEVELS = 2, SMC_PRUNE_MFMA_WALK = 3 };
int smc_engine_set_variant(smc_engine *e, int family, int variant);

/* Scheduling of lockstep rejuvenation (no reference counterpart: the reference
 * runs one particle at a time, main.cpp:296-311).  Enabled (the default), a batch's
 * whole-tree steps (shape and rate-matrix moves) are refreshed on a stream of their
 * own and decided when done, while the batch's other particles go on with their
 * rounds; disabled, every round waits for all its particles.  The chains, and so
 * the results, are the same either way. */
int smc_engine_set_lane(smc_engine *e, int enabled);

/* Site mixture at the root.  SMC_MIX_REFERENCE (default) is the reference's
 * formula exactly (Particle.cpp:653-670): log(exp(l)(1-p) + p pi[c] isInv), which
 * underflows to -inf below l = -745 and is NaN when a rate term is -inf, as in
 * the reference.  SMC_MIX_LOGSPACE computes the same value in log space: equal
 * to rounding wherever the reference formula is finite, finite where it
 * underflows; a rate term whose partials underflowed to zero (NaN through the
 * reference's scaling, e.g. the lowest gamma category of a tiny shape on a
 * variable column) contributes 0 to the site's sum over rates.  No reference
 * counterpart. */
enum { SMC_MIX_REFERENCE = 0, SMC_MIX_LOGSPACE = 1 };
int smc_engine_set_mixture(smc_engine *e, int mode);

/* Proposal tuning from the resampled population (main.cpp:213-289) on the GPU, over
 * slots [0, n): out[0] variance of log shape, out[1] variance of log branch length
 * (non-root nodes), out[2] mean pInvar, out[3] the shrinkage diagonal (the base for
 * rateMatrixDelta); with using_lg = 0 and rate_cholesky non-NULL, rate_cholesky[190][190]
 * receives the lower Cholesky factor (row-major, upper part zero) of the uncentred second
 * moment of the 190 log-exchangeabilities (Q21) scaled by 2.38^2 / 190 * 0.21 and shrunk
 * by 0.2 towards its mean diagonal, plus 1e-6 on the diagonal.  Variances are the mean
 * of squares minus the squared mean, as the reference's accumulators.  SMC_ERR_STATE
 * when the matrix is not positive definite. */
int smc_proposal_tuning(smc_engine *e, int n, int using_lg, double *out, double *rate_cholesky);

/* Split posteriors, computeSplitPosteriors (Miscellaneous.hpp:115-150), on the GPU:
 * the split sets of slots [0, n) (Tree::getSplitSet, Tree.cpp:485-535: the tips below
 * every non-root node, flipped when they do not hold taxon 0) weighted by weights[n]
 * (normalized weights, main.cpp:338-342, or 1/n per resampled copy, main.cpp:203-207).
 * A split is W = ceil(taxa / 64) 64-bit words, bit t = taxon t.  Writes the number of
 * distinct splits to *n_splits and, when it is at most max_splits, their words
 * split_words[n_splits][W], posterior[n_splits] (the sum of the weights of the
 * particles holding the split) and mean_length[n_splits] (the weighted mean of its
 * branch length over those particles, Tree::getSplitBranchMap, main.cpp:367-383),
 * ordered by a 64-bit hash of the words; any output may be NULL.  Two different
 * splits sharing a hash fail with SMC_ERR_STATE. */
int smc_split_posteriors(smc_engine *e, int n, const double *weights, int64_t max_splits,
                         int64_t *n_splits, uint64_t *split_words, double *posterior,
                         double *mean_length);
/* The greedy consensus (main.cpp:344-365) of the same splits: in decreasing posterior,
 * each split kept when it is compatible (four-point condition) with every split kept
 * before, until the tree is resolved.  Writes *n_selected (at most 2 taxa - 3, the
 * trivial splits included) and the kept splits' words, posteriors and mean lengths in
 * the order they were kept (arrays of 2 taxa - 3 entries; any may be NULL). */
int smc_consensus(smc_engine *e, int n, const double *weights, int32_t *n_selected,
                  uint64_t *split_words, double *posterior, double *mean_length);

/* Site patterns (no reference counterpart: the reference evaluates one pattern per
 * column, Alignment.cpp:83, Particle.cpp:22).  With compress = 1 (the default) the
 * identical columns of one CAT class share one slot whose log-likelihood is counted
 * once per column; 0 keeps one slot per column.  Applies to particles uploaded
 * afterwards.  Columns are identical when every taxon has the same code, gaps and
 * unknowns included, so the invariant-site terms (Q5) agree as well. */
int smc_engine_set_patterns(smc_engine *e, int compress);
/* Slots (site patterns within classes) of slots [first, first+count): the columns a
 * sweep prunes; -1 on invalid arguments. */
int64_t smc_pattern_count(const smc_engine *e, int first, int count);

/* Bytes of device memory held by the engine. */
int64_t smc_engine_device_bytes(const smc_engine *e);

#ifdef __cplusplus
}
#endif
#endif
