/* arbo_c.h — C ABI of the MI355X dengue individual-based-model engine.
 *
 * One engine owns the device memory for a batch of independent simulations
 * (each its own lattice of communities, parameters and Philox seed) and
 * advances all of them together, one day per step.  The same ABI is exported
 * by the GPU engine (libarbo_hip.so) and, under the same names, by the
 * serial/OpenMP CPU oracle (test infrastructure only).
 *
 * The reference crosses the host/device boundary through the host launchers
 * called by simulation() once per day (simulation.cpp:115-155):
 *   demographicInitialization  demographicinitial.h:13-44
 *   epidemicInitial            epidemicinitial.h:14-34
 *   nDemographic               demographic.h:17-31
 *   mDemographic               demographic.h:36-51
 *   epidemic                   epidemic.h:17-53
 *   dataCollect                datacollect.h:14-29
 *   setupCudaRNG               setuprng.h:11-12
 * and the caller reads d_nIncidence back (MCMC.cpp:190-213,
 * writeoutput.cpp:13-98).  Each entry point below names the launchers it
 * replaces.
 *
 * Conventions: every function returns 0 on success and a negative ARBO_E_*
 * code on failure; arbo_last_error() then holds a message.  An engine is
 * single-threaded (one HIP stream); distinct engines may run concurrently.
 * Plain pointers and sizes only; all host buffers are owned by the caller.
 */
#ifndef ARBO_C_H
#define ARBO_C_H
#include <stdint.h>

#if defined(__GNUC__)
#define ARBO_API __attribute__((visibility("default")))
#else
#define ARBO_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define ARBO_ABI_VERSION 7u
#define ARBO_MAX_SEROTYPES 4u
#define ARBO_MAX_COMMUNITY_HUMANS 65535u
#define ARBO_MAX_COMMUNITY_SLOTS 65535u

enum {
    ARBO_OK = 0,
    ARBO_E_INVALID = -1,   /* bad argument or parameter */
    ARBO_E_DEVICE = -2,    /* HIP runtime failure */
    ARBO_E_ALLOC = -3,     /* device or host allocation failure */
    ARBO_E_STATE = -4,     /* call out of order (e.g. step before init) */
    ARBO_E_CAPACITY = -5   /* more days requested than the series can hold */
};

/* Engine creation options. */
typedef struct ArboEngineDesc {
    uint32_t abi_version;  /* must be ARBO_ABI_VERSION */
    int32_t device;        /* HIP device ordinal (ignored by the oracle) */
    uint32_t flags;        /* ARBO_F_* */
    uint32_t threads;      /* oracle only: OpenMP threads (0 = runtime default) */
} ArboEngineDesc;

#define ARBO_F_GRAPH 1u        /* capture day loops in hipGraphs */
#define ARBO_F_TIMING 2u       /* time each day kernel with HIP events */
#define ARBO_F_SPATIAL 4u      /* keep the per-community series of arbo_read_spatial */
#define ARBO_F_HISTORY 8u      /* keep each human's infection history (arbo_read_history) */
#define ARBO_F_AOI 16u         /* keep the age-of-infection rings of arbo_read_aoi (implies ARBO_F_HISTORY) */
#define ARBO_AOI_RECORD 5000u  /* ages per ring (C_NAOIRECORD, constant.h:22) */

/* Model parameters of one simulation: the fields of SimulationParam that reach
 * the kernels (SimulationParam.h:16-54) plus the seasonal, incubation and
 * seeding extensions of the configured workloads. */
typedef struct ArboParams {
    uint32_t n_serotypes;       /* 1..4 (C_STRAINS, constant.h:17) */
    /* transmission (nepidemic.cu:66-70, mepidemic.cu:65-66) */
    float beta_hm;              /* events per infectious human per day per mosquito/human ratio:
                                   bitingRate * nmBitingSuccess */
    float beta_mh;              /* events per infectious mosquito per day: bitingRate * mnBitingSuccess */
    float long_distance;        /* fraction of events sent to a uniform random community */
    float external_infection;   /* introductions per day per serotype per simulation */
    /* dispersal network (network.cpp:31-41) */
    uint32_t neighbourhood;     /* 4 (von Neumann) or 8 (Moore) */
    uint32_t include_self;      /* 1: the community itself is a CSR entry */
    float dispersal_scale;      /* weights exp(-d/scale); 0 = unit weights */
    /* natural history */
    uint32_t iip_min, iip_max;  /* human intrinsic incubation, uniform integer days */
    float recovery;             /* human infectious period in days (stochastic rounding) */
    float eip;                  /* mosquito extrinsic incubation in days (stochastic rounding) */
    /* human bi-Weibull lifespan (ndemographic.cu:13-29): F(a) = 1-exp(-[(a*infant_scale)^infant_shape
       + 1{a>=loc}((a-loc)/scale)^shape]), a in days */
    float h_infant_shape, h_infant_scale, h_loc, h_scale, h_shape;
    /* mosquito Weibull lifespan (mdemographic.cu:13-17): F(a) = 1-exp(-(a/scale)^shape) */
    float m_scale, m_shape;
    /* mosquito emergence target: m_ratio * N_c * (1 + season_amp*sin(2 pi t/365 + season_phase)) */
    float m_ratio, season_amp, season_phase;
    /* t=0 seeding: independent Bernoulli per individual (alive mosquitoes only) */
    float seed_frac_h, seed_frac_m;
    /* t=0 immunity per serotype: 0 = probability sero_prev everywhere,
       1 = fraction drawn Beta(2,5) per community and serotype */
    uint32_t immunity_mode;
    float sero_prev;
    /* fraction of a source's events that stay in the source community before
       the long-distance split (noDistance, nepidemic.cu:98-100); 0 when the
       community is on its own CSR row (include_self) */
    float no_distance;
    /* law of a source's daily event count (nepidemic.cu:64-104), rate lam:
       ARBO_COUNT_POISSON: stay ~ Poisson(lam no_distance), long-distance
         ~ Poisson(lam (1 - no_distance) long_distance), the rest local, all
         independent Poisson;
       ARBO_COUNT_STOCH_ROUND: the reference's rule, n = stochRound(lam),
         stay = stochRound(n no_distance), long-distance =
         stochRound((n - stay) long_distance), the rest local */
    uint32_t count_law;
} ArboParams;

#define ARBO_COUNT_POISSON 0u
#define ARBO_COUNT_STOCH_ROUND 1u

/* Daily climate of one simulation and the coefficients that turn it into
 * the day's parameters (SimulationParam::adjustClimate,
 * SimulationParam.cpp:92-137; the logistic emergence target,
 * simulation.cpp:87,129-131).  Day t of the simulation reads row t + 1 of
 * the climate, row 0 sets the initial target; days >= max_days + 1.
 * With a climate, per day t:
 *   beta_hm(t) = biting_rate * nmBitingSuccess(T), beta_mh(t) = biting_rate *
 *     mnBitingSuccess(T), both Briere curves of the temperature T (deg C);
 *   eip(t) = 1 / (scale_eip * 24 * 0.003359 (TK/298) exp(15000/1.986
 *     (1/298 - 1/TK))), TK = T + 273.15;
 *   mosquito Weibull scale(t) = 1 / (scale_mle * mu(T) * (1 - H)^scale_humidity),
 *     mu a quartic death rate in T; mosquitoes born on day b die on the first
 *     day b + a with PLE <= F(a; scale(b + a)), m_shape unchanged (the t = 0
 *     census keeps m_scale);
 *   the whole simulation's target M(t+1) = M(t) + mos_net_growth M(t) (1 -
 *     M(t) / K(t+1)), K(t) = min_mos_to_human N (1 + P(t))^scale_cc, M(0) =
 *     K(0); community c's share is M(t) humans[c] / N.
 * beta_hm, beta_mh, eip, m_scale, season_amp and season_phase of
 * ArboParams are then not used by the day loop. */
typedef struct ArboClimate {
    const float* temperature;    /* [days] deg C (the climate file's T column) */
    const float* humidity;       /* [days] relative humidity in [0, 1) (H) */
    const float* precipitation;  /* [days] precipitation, > -1 (P) */
    uint32_t days;
    float biting_rate;           /* bites per mosquito per day (bitingRate) */
    float scale_eip, scale_mle, scale_cc, scale_humidity;
    float min_mos_to_human;      /* minMosToHuman */
    float mos_net_growth;        /* mosNetGrowth */
} ArboClimate;

/* Shape of one simulation: a grid_w x grid_h non-wrapping lattice of
 * communities (community id = y*grid_w + x) and the humans of each. */
typedef struct ArboShape {
    uint32_t grid_w, grid_h;
    const uint32_t* humans;      /* [grid_w*grid_h], each 1..ARBO_MAX_COMMUNITY_HUMANS */
    const uint32_t* mosq_slots;  /* [grid_w*grid_h] slot capacity, or NULL for
                                    ceil(m_ratio*(1+season_amp)*N*share[c]), with a
                                    climate ceil(max_t M(t) share[c]) + 1 */
    const ArboClimate* climate;  /* NULL: the seasonal model of ArboParams */
    const float* mosq_share;     /* [grid_w*grid_h] each community's share of the
                                    simulation's mosquitoes (positive, summing to 1
                                    within 1e-3; mHeteroPopSize sizes,
                                    network.cpp:138-159), or NULL: humans[c] / N */
} ArboShape;

/* Time series of one simulation, layout of the reference's d_n* arrays
 * (datacollect.cu:10-74): row t is the state at the start of day t,
 * rows 0..days_done; per-serotype arrays are [t*S + s]. */
typedef struct ArboSeries {
    uint32_t* n_count;        /* [T+1]  humans alive              (d_nCount) */
    uint32_t* m_count;        /* [T+1]  mosquitoes alive          (d_mCount) */
    uint32_t* n_infected;     /* [(T+1)*S] infectious humans      (d_nInfectedCount) */
    uint32_t* n_incidence;    /* [(T+1)*S] E->I onsets of day t-1 (d_nIncidence) */
    uint32_t* n_one_subpop;   /* [(T+1)*S] infectious humans in community 0 (d_nOneSubPopInfectedCount) */
} ArboSeries;

/* Packed census of one simulation, community-contiguous (no padding).
 * Human word: bits 0-1 status (0 S, 1 E, 2 I), 2-3 strain, 4-7 immune mask,
 *             8-31 death day (dies during day t iff t >= death day).
 * Mosquito word: bits 0-1 status (0 S, 1 E, 2 I, 3 dead slot), 2-3 strain,
 *             4-15 death day mod 4096.
 * next_day: E->I day (or I->S day for humans), mod 2^16; the transition fires
 *           on the first day t with t > next_day. */
typedef struct ArboState {
    uint32_t* h_word;     /* [N] */
    uint16_t* h_next;     /* [N] */
    int32_t* h_birth;     /* [N] day of birth (negative: before day 0) */
    uint16_t* m_word;     /* [M slots] */
    uint16_t* m_next;     /* [M slots] */
} ArboState;

typedef struct ArboEngine ArboEngine;

/* Create an engine on one device.  Replaces the per-MCMC-step cudaMalloc of
 * main.cpp:198-249 (buffers are sized by arbo_batch_init and reused). */
ARBO_API int arbo_create(const ArboEngineDesc* desc, ArboEngine** out);

/* Build the dispersal CSR and alias tables, the lifespan tables, the seasonal
 * targets and the t=0 census of n_sims simulations, room for max_days days.
 * Replaces setupCudaRNG, demographicInitialization, epidemicInitial and the
 * first dataCollect (simulation.cpp:88-113). */
ARBO_API int arbo_batch_init(ArboEngine* e, uint32_t n_sims, const ArboShape* shapes,
                    const ArboParams* params, const uint64_t* seeds, uint32_t max_days);

/* Advance every simulation by n_days days.  One day = nDemographic,
 * mDemographic, epidemic and dataCollect of the reference's day loop
 * (simulation.cpp:115-155). */
ARBO_API int arbo_batch_step(ArboEngine* e, uint32_t n_days);

/* Wait for queued work. */
ARBO_API int arbo_sync(ArboEngine* e);

/* Days simulated so far. */
ARBO_API int arbo_days_done(const ArboEngine* e, uint32_t* out);

/* Sizes of simulation `sim`: communities, humans, mosquito slots, serotypes. */
ARBO_API int arbo_sim_sizes(const ArboEngine* e, uint32_t sim, uint32_t* communities,
                   uint64_t* humans, uint64_t* slots, uint32_t* serotypes);

/* Copy the time series of one simulation (rows 0..days_done) to the host.
 * Replaces the D2H copies of MCMC::storeIncidence and writeIncidence. */
ARBO_API int arbo_read_series(ArboEngine* e, uint32_t sim, ArboSeries* out);

/* Copy the census of one simulation to the host / restore it (checkpoint). */
ARBO_API int arbo_read_state(ArboEngine* e, uint32_t sim, ArboState* out);
ARBO_API int arbo_write_state(ArboEngine* e, uint32_t sim, const ArboState* in);

/* Set the day counter of the whole batch (resume from a checkpoint written by
 * arbo_read_state at that day); series rows after `day` are cleared. */
ARBO_API int arbo_set_day(ArboEngine* e, uint32_t day);

/* Infectious humans per community and serotype at the start of every day
 * simulated so far (engines created with ARBO_F_SPATIAL): out[(t * S + s) * C
 * + c] for t = 0..days, the layout of d_nInfectedSubPopCountSeries
 * (datacollect.cu:10-23) that writeSpatial prints (writeoutput.cpp:194-241);
 * out holds (days + 1) * S * C words. */
ARBO_API int arbo_read_spatial(ArboEngine* e, uint32_t sim, uint32_t* out);

/* Infection history of every human of a simulation (engines created with
 * ARBO_F_HISTORY), the reference's d_nHistory: out[s * N + i] for serotype s
 * and human i (community order), -1 if never infected by s, -2 if immune to s
 * from the t = 0 census (prior immunity or a seeded infection), otherwise
 * the age in days on the day of the infection (demographicintial.cu:164-169,
 * epidemicinitial.cu:59, nepidemic.cu:248,298; a rebirth resets it,
 * ndemographic.cu:100).  out holds S * N words. */
ARBO_API int arbo_read_history(ArboEngine* e, uint32_t sim, int32_t* out);

/* Age-of-infection rings of a simulation (engines created with ARBO_F_AOI),
 * the reference's d_nAgeOfInfection (nepidemic.cu:229-243): every infection
 * by a mosquito event (not an external introduction) on a day t with
 * max_days - t <= 3650 writes the age at infection into ring k, k = the
 * serotypes the human had been infected with before, at the ring's next
 * position (cycling through ARBO_AOI_RECORD slots).  The infections of one
 * day enter in human order (community order, then index).  Rings start at
 * zero.  out[k * ARBO_AOI_RECORD + r], 4 * ARBO_AOI_RECORD words. */
ARBO_API int arbo_read_aoi(ArboEngine* e, uint32_t sim, uint32_t* out);

/* The parameters climate row `row` gives a day (SimulationParam::adjustClimate,
 * SimulationParam.cpp:92-137), before their rounding to float: out[4] =
 * {beta_hm, beta_mh, eip, mosquito Weibull scale}.  Day t of a simulation
 * uses row t + 1. */
ARBO_API int arbo_climate_day(const ArboClimate* climate, uint32_t row, double* out);

/* Kernel timing (ARBO_F_TIMING): total milliseconds and launches of the
 * demography kernel (kind 0) and event-application kernel (kind 1) since the
 * last reset. */
ARBO_API int arbo_kernel_time(ArboEngine* e, uint32_t kind, double* ms, uint64_t* launches);
ARBO_API int arbo_reset_timing(ArboEngine* e);

/* Census traffic of the day loop since arbo_batch_init or the last
 * arbo_reset_timing: the 16-byte vectors of packed human words (4 humans)
 * and of packed mosquito words (8 slots) that the demography kernel read.
 * The kernel skips exactly what cannot change on a day (a community's
 * humans before its earliest death or state change, mosquito slots above
 * its highest alive one), so these count the census bytes the algorithm
 * needs; every mosquito vector read is also written back.  The oracle,
 * which scans every member every day, reports the full census.  No
 * counterpart in the reference (test and measurement only). */
ARBO_API int arbo_scan_counts(ArboEngine* e, uint64_t* human_vectors, uint64_t* mosquito_vectors);

/* Launch knobs of the day kernels.  No knob changes a result: every random
 * draw is keyed by model entities (simulation, community, individual, day,
 * event), never by workgroup or lane, so runs under any tuning are
 * bit-identical.  No counterpart in the reference, whose launch shapes come
 * from the device properties (cudaidentities.h:14-37). */
typedef struct ArboTuning {
    uint32_t census_unroll;   /* demography kernel: 16-byte census loads in flight per lane (2, 4, 8) */
    uint32_t early_mosq;      /* of the first batch, mosquito loads issued before the human scan */
    uint32_t xcd_order;       /* bit 0 demography, bit 1 event kernel: the workgroups one XCD receives
                                 take a contiguous range of communities */
    uint32_t ld_split;        /* long-distance kernel: workgroups per simulation, 1..64 */
    uint32_t waves_per_simd;  /* demography kernel: 0 the compiler's register budget, 8 one that
                                 keeps 8 waves per SIMD (with census_unroll/early_mosq 4/2 or
                                 2/2): more resident waves for batches of few communities */
    uint32_t waves_per_comm;  /* demography kernel: waves per community, 1, or 2 (with 4/2) or 4
                                 (with 2/2): shorter waves for batches of few communities */
} ArboTuning;

/* The tunings arbo_autotune tries (out may be NULL to count them). */
ARBO_API int arbo_tuning_candidates(ArboTuning* out, uint32_t cap, uint32_t* n);
/* Set or read the engine's tuning (default {4, 2, 0, 8, 0, 1}); ARBO_E_INVALID for a
 * combination the library was not built with. */
ARBO_API int arbo_set_tuning(ArboEngine* e, const ArboTuning* t);
ARBO_API int arbo_get_tuning(const ArboEngine* e, ArboTuning* out);
/* Time every candidate on the engine's own batch for days_each days each
 * (those days are simulated: the run advances by candidates x days_each days,
 * with the same results as under any fixed tuning; the candidates take turns
 * a day at a time, forward and backward in alternate rounds) and keep the
 * fastest demography knobs and the fastest event-kernel order; ms_per_day
 * gets the sum of their day times.  chosen and ms_per_day may be NULL. */
ARBO_API int arbo_autotune(ArboEngine* e, uint32_t days_each, ArboTuning* chosen, double* ms_per_day);

ARBO_API void arbo_destroy(ArboEngine* e);
ARBO_API const char* arbo_last_error(const ArboEngine* e);

#ifdef __cplusplus
}
#endif
#endif /* ARBO_C_H */
