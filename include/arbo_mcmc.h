/* arbo_mcmc.h — C ABI of the batched MCMC driver.
 *
 * The reference fits its model with one Metropolis chain whose every step
 * runs C_MAXSIMRUN stochastic simulations of the proposal
 * (main.cpp:72-349, MCMC.cpp:108-351):
 *   covariance    MCMC::calcCov         MCMC.cpp:89-105
 *   propose       MCMC::propose         MCMC.cpp:108-161
 *   checkParams   MCMC::checkParams     MCMC.cpp:164-186
 *   storeIncidence (weekly binning)     MCMC.cpp:190-213
 *   likelihood    negbinomial_pdf, MCMC::calcEachLogLikelihood,
 *                 MCMC::calcLogLikelihood                MCMC.cpp:216-282
 *   accept/reject MCMC::nextStep, MCMC::reject           MCMC.cpp:284-380
 * This driver advances n_chains chains together: the proposals of one
 * iteration are simulated as one engine batch (include/arbo_c.h), so a step
 * of all chains costs one pass of the day loop over the batch.
 *
 * Parameter vector (ARBO_MCMC_PARAMS = 7), by ArboMcmcDesc.param_set:
 *   ARBO_FIT_SEASONAL: 0 beta_hm, 1 beta_mh, 2 log10(long-distance
 *     fraction), 3 seasonal amplitude, 4 seasonal phase (circular); 0-4 are
 *     set in ArboParams of the base parameter set.
 *   ARBO_FIT_CLIMATE (the reference's fitted set, MCMC.h:37-50 and
 *     MCMC.cpp:79-86): 0 scale_eip, 1 scale_mle, 2 scale_cc, 3
 *     scale_humidity, 4 min_mos_to_human; 0-4 are set in the ArboClimate of
 *     the shape (which must have one).
 *   In both: 5 the reporting rate (observationRate), which scales the
 *   simulated weekly incidence, and 6 the negative-binomial dispersion d
 *   (dispersion): size 1/d, success probability d mu / (1 + d mu)
 *   (MCMC.cpp:240-243).  A parameter with an ARBO_PRIOR_FIXED prior keeps
 *   its value (the configured fit holds d at 0.1, size 10).
 *
 * Pseudo-marginal seeds (ArboMcmcDesc.n_blocks).  Every simulation of a
 * proposal has its own Philox seed.  With n_blocks = 0 every proposal draws
 * fresh seeds for all n_particles simulations (the reference's scheme,
 * main.cpp:262-280).  With n_blocks = G >= 1 each chain keeps its particles'
 * seeds in its state, and each proposal refreshes only the particles p with
 * p mod G = b, b uniform in [0, G), the rest keep theirs: the block
 * pseudo-marginal method of Tran, Kohn, Quiroz & Villani (2016).  The
 * seed move is reversible with respect to the seeds' own law, so the
 * chains still target the exact posterior; neighbouring proposals share
 * their random numbers, which correlates their likelihood estimates.
 *
 * Conventions as include/arbo_c.h: 0 or a negative ARBO_E_* code, message in
 * arbo_mcmc_last_error(); one driver per thread.
 */
#ifndef ARBO_MCMC_H
#define ARBO_MCMC_H
#include <stdint.h>

#include "arbo_c.h"
#include "arbo_surrogate.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ARBO_MCMC_PARAMS 7u
#define ARBO_MCMC_SUR_INPUTS 6u   /* surrogate inputs of parameters 0-4 (see arbo_mcmc_features) */
#define ARBO_MCMC_MAX_CHAINS 4096u

#define ARBO_FIT_SEASONAL 0u
#define ARBO_FIT_CLIMATE 1u

enum {
    ARBO_PRIOR_UNIFORM = 0,   /* U[a, b] */
    ARBO_PRIOR_CIRCULAR = 1,  /* U[a, b) on a circle: proposals wrap into [a, b) */
    ARBO_PRIOR_BETA = 2,      /* Beta(a, b) on (0, 1) */
    ARBO_PRIOR_TRUNC_NORMAL = 3, /* N(a, b^2) truncated to (0, 10000] (TruncatedNormal, MCMC.h:37-46) */
    ARBO_PRIOR_FIXED = 4         /* the point a: the parameter is not fitted and keeps that value */
};

#define ARBO_MCMC_F_ZERO_WEEK_FLOOR 1u   /* a week whose simulated mean is 0 scores week_floor even when its
                                            count is 0, as the reference's fmax(-250, NaN) does
                                            (MCMC.cpp:216-243); without it such a week scores log 1 = 0 */

typedef struct ArboPrior {
    uint32_t kind;
    uint32_t reserved;
    double a, b;
} ArboPrior;

typedef struct ArboMcmcDesc {
    uint32_t abi_version;     /* ARBO_ABI_VERSION */
    uint32_t n_chains;        /* 1..ARBO_MCMC_MAX_CHAINS */
    uint32_t n_particles;     /* simulations per proposal, averaged as in MCMC.cpp:252-282 (C_MAXSIMRUN) */
    uint32_t n_weeks;         /* weekly counts; each simulation runs 7 n_weeks - 1 days (main.cpp:34) */
    uint32_t skip_weeks;      /* leading weeks left out of the likelihood (MCMC.cpp:238 skips one) */
    uint32_t adapt_after;     /* pooled chain samples before the adapted covariance is used */
    uint32_t param_set;       /* ARBO_FIT_SEASONAL or ARBO_FIT_CLIMATE */
    uint32_t flags;           /* ARBO_MCMC_F_* */
    uint32_t n_blocks;        /* pseudo-marginal seed blocks (see above): 0 fresh seeds for every proposal */
    uint32_t reserved0;       /* 0 */
    double week_floor;        /* lower clamp of each week's log-likelihood (MCMC.cpp:243: -250) */
    double mix_beta;          /* probability of the fixed proposal (MCMC.cpp:127: 0.05) */
    double init_scale;        /* fixed proposal sd per parameter, as a fraction of its prior sd, over sqrt(d) */
    uint64_t seed;            /* proposals, acceptances and simulation seeds */
    ArboPrior priors[ARBO_MCMC_PARAMS];
} ArboMcmcDesc;

/* The engine entry points the driver calls (a caller may point them at any
 * library exporting include/arbo_c.h). */
typedef struct ArboEngineApi {
    int (*batch_init)(ArboEngine* e, uint32_t n_sims, const ArboShape* shapes, const ArboParams* params,
                      const uint64_t* seeds, uint32_t max_days);
    int (*batch_step)(ArboEngine* e, uint32_t n_days);
    int (*read_series)(ArboEngine* e, uint32_t sim, ArboSeries* out);
    const char* (*last_error)(const ArboEngine* e);
} ArboEngineApi;

/* A caller-supplied log-likelihood of n parameter vectors (theta[n * 7]);
 * returns 0 or a negative code. */
typedef int (*ArboLogLikFn)(void* ctx, uint32_t n, const double* theta, double* loglik);

typedef struct ArboMcmc ArboMcmc;

/* Delayed acceptance with the surrogate (no counterpart in the reference).
 * The surrogate posterior is the prior x the negative-binomial likelihood of
 * the surrogate's mean incidence.  Two first stages:
 *
 * ARBO_DA_LANGEVIN: each iteration every chain runs sub_steps
 *   Metropolis-adjusted Langevin steps on the surrogate posterior,
 *   preconditioned by the chains' pooled covariance, then the simulator
 *   judges the end point:
 *     alpha2 = min(1, L(theta') L_s(theta) / (L(theta) L_s(theta')))
 *   (the surrogate-transition method of Liu 2001).  The step size adapts
 *   (diminishing Robbins-Monro) to target_accept of the surrogate steps.
 * ARBO_DA_SCREEN: each iteration every chain draws the adaptive-Metropolis
 *   proposal of arbo_mcmc_run; the surrogate screens it,
 *     alpha1 = min(1, pi_s(theta') / pi_s(theta)),
 *   and only proposals that pass are simulated, for the same alpha2
 *   (Christen & Fox 2005).  An iteration whose proposals all fail the screen
 *   runs no simulation.
 *
 * Either way the posterior under the simulator's likelihood stays exact. */
#define ARBO_DA_LANGEVIN 0u
#define ARBO_DA_SCREEN 1u
typedef struct ArboDaDesc {
    uint32_t sub_steps;       /* ARBO_DA_LANGEVIN: surrogate steps per simulator evaluation, 1..1000 */
    uint32_t mode;            /* ARBO_DA_LANGEVIN or ARBO_DA_SCREEN */
    double step0;             /* ARBO_DA_LANGEVIN: initial step in preconditioner units */
    double target_accept;     /* ARBO_DA_LANGEVIN: surrogate-step acceptance the step size adapts to (0.574) */
} ArboDaDesc;

/* Create the driver.  shape: the model's lattice and humans per community
 * (copied); base: every simulation parameter not in the vector; data:
 * n_weeks observed weekly counts; theta0: n_chains * 7 starting values, or
 * NULL for the prior means. */
ARBO_API int arbo_mcmc_create(const ArboMcmcDesc* desc, const ArboShape* shape, const ArboParams* base,
                              const uint32_t* data, const double* theta0, ArboMcmc** out);

/* n_iters iterations of every chain, the likelihood from simulations on the
 * engine (api NULL: this library's engine).  The first call also evaluates
 * the starting values. */
ARBO_API int arbo_mcmc_run(ArboMcmc* m, ArboEngine* engine, const ArboEngineApi* api, uint32_t n_iters);

/* The same iterations with a caller-supplied log-likelihood. */
ARBO_API int arbo_mcmc_run_target(ArboMcmc* m, ArboLogLikFn fn, void* ctx, uint32_t n_iters);

/* n_iters delayed-acceptance iterations: surrogate `sur` (inputs:
 * arbo_mcmc_features, outputs: n_weeks log-mean weekly incidences) first,
 * simulations on the engine second. */
ARBO_API int arbo_mcmc_run_da(ArboMcmc* m, ArboEngine* engine, const ArboEngineApi* api, ArboSurrogate* sur,
                              const ArboDaDesc* da, uint32_t n_iters);

/* The same with a caller-supplied second-stage log-likelihood. */
ARBO_API int arbo_mcmc_run_da_target(ArboMcmc* m, ArboLogLikFn fn, void* ctx, ArboSurrogate* sur,
                                     const ArboDaDesc* da, uint32_t n_iters);

/* Surrogate inputs of n vectors (x[n * ARBO_MCMC_SUR_INPUTS]): parameters 0-4 in order, each
 * by its prior: uniform scaled from its range to [-1, 1], circular as (cos,
 * sin), truncated normal as (x - a) / b, Beta standardised by its mean and
 * sd; unused inputs 0.  The reporting rate (5) and the dispersion (6)
 * are no inputs: they enter the likelihood, not the simulation. */
ARBO_API int arbo_mcmc_features(const ArboMcmc* m, uint32_t n, const double* theta, float* x);

/* Surrogate log-posterior and its gradient for n vectors (out[n],
 * grad[n * 7]; a fixed parameter's component is 0). */
ARBO_API int arbo_mcmc_surrogate_logpost(ArboMcmc* m, ArboSurrogate* sur, uint32_t n, const double* theta,
                                         double* out, double* grad);

/* Delayed-acceptance counters: surrogate steps proposed and accepted,
 * seconds in the surrogate, current step size. */
ARBO_API int arbo_mcmc_da_counters(const ArboMcmc* m, uint64_t* steps, uint64_t* accepted, double* seconds,
                                   double* step);

/* Iterations recorded so far (not counting the starting values). */
ARBO_API int arbo_mcmc_iterations(const ArboMcmc* m, uint32_t* out);

/* History of iterations [first, first + count): theta[count * n_chains * 7]
 * (state after each iteration), loglik[count * n_chains] (its likelihood
 * estimate), accepted[count * n_chains]; any output may be NULL. */
ARBO_API int arbo_mcmc_read(const ArboMcmc* m, uint32_t first, uint32_t count, double* theta, double* loglik,
                            uint8_t* accepted);

/* Candidates of iterations [first, first + count), the records of
 * MCMC::writeCandidates (MCMC.cpp:443-484): theta[count * n_chains * 7]
 * (each chain's candidate of the iteration) and logpost[count * n_chains]
 * (its log-likelihood estimate plus log prior, the value MCMC::nextStep
 * compares, MCMC.cpp:284-300; -inf outside the prior support or for a
 * vector the engine cannot simulate, NaN when a delayed-acceptance first
 * stage stopped it before the simulator).  Any output may be NULL. */
ARBO_API int arbo_mcmc_read_candidates(const ArboMcmc* m, uint32_t first, uint32_t count, double* theta,
                                       double* logpost);

/* Counters: proposals made, proposals inside the prior support, simulations
 * run, proposals accepted, seconds spent in likelihood evaluation. */
ARBO_API int arbo_mcmc_counters(const ArboMcmc* m, uint64_t* proposals, uint64_t* in_support,
                                uint64_t* simulations, uint64_t* accepted, double* eval_seconds);

/* Proposals inside the prior support that the engine cannot simulate (its
 * batch_init checks reject them, e.g. more mosquito slots per community
 * than ARBO_MAX_COMMUNITY_SLOTS): each gets log-likelihood -inf and is
 * rejected, where the reference would simulate it. */
ARBO_API int arbo_mcmc_unsimulable(const ArboMcmc* m, uint64_t* out);

/* Climate of one vector of the ARBO_FIT_CLIMATE set: base with its five
 * coefficients set (MCMC::copyParams, MCMC.cpp:79-86). */
ARBO_API int arbo_mcmc_sim_climate(const double* theta, const ArboClimate* base, ArboClimate* out);

/* Simulation parameters of one vector: base with the vector's fields set. */
ARBO_API int arbo_mcmc_sim_params(const double* theta, const ArboParams* base, ArboParams* out);

/* log prior density of one vector (-inf outside the support is reported as
 * ARBO_E_INVALID). */
ARBO_API int arbo_mcmc_log_prior(const ArboMcmc* m, const double* theta, double* out);

ARBO_API void arbo_mcmc_destroy(ArboMcmc* m);
ARBO_API const char* arbo_mcmc_last_error(const ArboMcmc* m);

/* Weekly incidence from a series (MCMC.cpp:199-209): week w sums rows
 * 7w .. 7w+6 (rows <= rows-1) over all serotypes; incidence is [rows * S]. */
ARBO_API int arbo_weekly_incidence(const uint32_t* incidence, uint32_t rows, uint32_t serotypes, uint32_t n_weeks,
                                   uint32_t* out);

/* Negative-binomial log-likelihood of observed weekly counts given simulated
 * weekly incidence (MCMC.cpp:216-248): weeks skip..n_weeks-1, mean
 * rho * simulated, size r, each week clamped below at floor (MCMC.cpp:243). */
ARBO_API int arbo_nb_loglik(const uint32_t* sim_weekly, const uint32_t* data, uint32_t n_weeks, uint32_t skip,
                            double rho, double r, double floor, double* out);

#ifdef __cplusplus
}
#endif
#endif /* ARBO_MCMC_H */
