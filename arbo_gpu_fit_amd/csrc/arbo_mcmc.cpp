// arbo_mcmc.cpp — batched adaptive-Metropolis driver over the engine.
//
// The reference runs one chain; each step proposes, checks the support,
// simulates C_MAXSIMRUN runs of the proposal one after another, bins their
// incidence by week and accepts by the Metropolis ratio of the averaged
// negative-binomial likelihood times the prior (main.cpp:72-349,
// MCMC.cpp:89-380).  Here n_chains chains step together and the
// n_chains x n_particles simulations of an iteration are one engine batch,
// so an iteration costs one pass of the day loop over that batch.
//
// Proposal (Roberts & Rosenthal 2009, "Examples of adaptive MCMC", the
// mixture the reference's MCMC.cpp:108-161 approximates by summing two
// draws): with probability mix_beta, or while fewer than adapt_after samples
// have been pooled, theta + N(0, diag((init_scale sd_i)^2 / d)); otherwise
// theta + N(0, 2.38^2 / d Sigma_n), Sigma_n the covariance of all chains'
// states so far (inter-chain adaptation, Craiu, Rosenthal & Yang 2009),
// updated incrementally instead of recomputed (MCMC.cpp:89-105).
#include <hip/hip_runtime.h>

#include <chrono>
#include <new>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "arbo_c.h"
#include "arbo_mcmc.h"
#include "arbo_rng.h"
#include "arbo_tables.h"

using namespace arbo;

namespace {

constexpr uint32_t D = ARBO_MCMC_PARAMS;
constexpr uint32_t kModel = 5;                      // parameters 0..4 set the simulation
constexpr uint32_t kRho = 5, kDisp = 6;             // reporting rate, NB dispersion
constexpr uint32_t kSurIn = ARBO_MCMC_SUR_INPUTS;   // surrogate inputs of parameters 0..4
constexpr double kTwoPi = 6.283185307179586;
constexpr double kTruncHi = 10000.0;   // upper end of the truncated-normal priors (MCMC.h:37-46)

// Philox purposes of the driver (key: the driver's seed).
enum : uint32_t { Q_MIX = 1, Q_NORMAL = 2, Q_ACCEPT = 3, Q_SIM_SEED = 4, Q_BLOCK = 8 };
constexpr uint32_t kInitIter = 0xFFFFFFFFu;

double prior_mean(const ArboPrior& p) {
    if (p.kind == ARBO_PRIOR_BETA) return p.a / (p.a + p.b);
    if (p.kind == ARBO_PRIOR_FIXED) return p.a;
    return 0.5 * (p.a + p.b);
}

double prior_sd(const ArboPrior& p) {
    if (p.kind == ARBO_PRIOR_FIXED) return 0.0;
    if (p.kind == ARBO_PRIOR_BETA) {
        const double s = p.a + p.b;
        return std::sqrt(p.a * p.b / (s * s * (s + 1.0)));
    }
    return (p.b - p.a) / std::sqrt(12.0);
}

// log density; -inf outside the support
double prior_logpdf(const ArboPrior& p, double x) {
    switch (p.kind) {
        case ARBO_PRIOR_UNIFORM:
            return (x >= p.a && x <= p.b) ? -std::log(p.b - p.a) : -INFINITY;
        case ARBO_PRIOR_CIRCULAR:
            return (x >= p.a && x < p.b) ? -std::log(p.b - p.a) : -INFINITY;
        case ARBO_PRIOR_BETA:
            if (!(x > 0.0 && x < 1.0)) return -INFINITY;
            return (p.a - 1.0) * std::log(x) + (p.b - 1.0) * std::log1p(-x) -
                   (std::lgamma(p.a) + std::lgamma(p.b) - std::lgamma(p.a + p.b));
        case ARBO_PRIOR_TRUNC_NORMAL: {   // N(a, b^2) on (0, kTruncHi], normalised
            if (!(x > 0.0 && x <= kTruncHi)) return -INFINITY;
            const double z = (x - p.a) / p.b;
            const double mass = 0.5 * (std::erfc(-(kTruncHi - p.a) / (p.b * std::sqrt(2.0))) -
                                       std::erfc(p.a / (p.b * std::sqrt(2.0))));
            return -0.5 * z * z - std::log(p.b) - 0.5 * std::log(2.0 * M_PI) - std::log(mass);
        }
        case ARBO_PRIOR_FIXED:   // a point mass: the parameter is not fitted
            return x == p.a ? 0.0 : -INFINITY;
    }
    return -INFINITY;
}

bool valid_prior(const ArboPrior& p) {
    if (!(std::isfinite(p.a) && std::isfinite(p.b))) return false;
    if (p.kind == ARBO_PRIOR_BETA) return p.a > 0.0 && p.b > 0.0;
    if (p.kind == ARBO_PRIOR_TRUNC_NORMAL) return p.b > 0.0 && p.a > -10.0 * p.b && p.a < kTruncHi;
    if (p.kind == ARBO_PRIOR_UNIFORM || p.kind == ARBO_PRIOR_CIRCULAR) return p.b > p.a;
    if (p.kind == ARBO_PRIOR_FIXED) return true;
    return false;
}

// Inputs the parameters 0..4 take: 1 each, 2 for a circular one.
uint32_t n_inputs(const ArboPrior* priors) {
    uint32_t n = 0;
    for (uint32_t i = 0; i < kModel; ++i) n += priors[i].kind == ARBO_PRIOR_CIRCULAR ? 2u : 1u;
    return n;
}

// Digamma by the recurrence psi(x) = psi(x + 1) - 1/x up to x >= 6, then
// the asymptotic series (Abramowitz & Stegun 6.3.18); x > 0.
double digamma(double x) {
    double acc = 0.0;
    while (x < 6.0) {
        acc -= 1.0 / x;
        x += 1.0;
    }
    const double f = 1.0 / (x * x);
    return acc + std::log(x) - 0.5 / x -
           f * (1.0 / 12.0 - f * (1.0 / 120.0 - f * (1.0 / 252.0 - f * (1.0 / 240.0 - f / 132.0))));
}

// Cholesky factor (lower) of a d x d symmetric positive-definite matrix;
// false when a pivot is not positive.
bool cholesky(const double* A, double* L) {
    std::memset(L, 0, sizeof(double) * D * D);
    for (uint32_t i = 0; i < D; ++i)
        for (uint32_t j = 0; j <= i; ++j) {
            double s = A[i * D + j];
            for (uint32_t k = 0; k < j; ++k) s -= L[i * D + k] * L[j * D + k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                L[i * D + i] = std::sqrt(s);
            } else {
                L[i * D + j] = s / L[j * D + j];
            }
        }
    return true;
}

}  // namespace

struct ArboMcmc {
    ArboMcmcDesc desc{};
    uint32_t grid_w = 0, grid_h = 0;
    std::vector<uint32_t> humans, slots;
    bool has_slots = false;
    // the climate every proposal's simulations run under (copied), if any
    bool has_climate = false;
    std::vector<float> share;   // mosquito shares per community, if given
    std::vector<float> clim_t, clim_h, clim_p;
    ArboClimate climate{};
    ArboParams base{};
    std::vector<uint32_t> data;
    Key key{0, 0};
    uint32_t C = 0;
    // chain state: parameters, likelihood estimate, log prior, and the Philox
    // seeds of the chain's particles (the pseudo-marginal auxiliary state)
    std::vector<double> theta, loglik, logprior;
    std::vector<uint64_t> useed;
    bool started = false;
    uint32_t iter = 0;
    // pooled adaptation (Welford over every chain state after every iteration)
    double n_pool = 0.0;
    double mean[D] = {};
    double m2[D * D] = {};
    double sd0[D] = {};
    // history: states after each iteration, and each iteration's candidates
    std::vector<double> h_theta, h_loglik;
    std::vector<uint8_t> h_acc;
    std::vector<double> h_cand, h_cand_lp;
    // counters
    uint64_t n_prop = 0, n_support = 0, n_sims = 0, n_acc = 0, n_unsim = 0;
    double eval_s = 0.0;
    // delayed acceptance
    uint64_t da_steps = 0, da_acc = 0;
    double sur_s = 0.0, da_log_eps = 0.0, da_t = 0.0;
    bool da_init = false;
    std::string err;
};

namespace {

int fail(ArboMcmc* m, int code, const std::string& msg) {
    if (m) m->err = msg;
    return code;
}

double u01(uint32_t a, uint32_t b) { return u53(a, b); }

// D standard normals of (chain, iteration) by Box-Muller on 53-bit
// uniforms; z holds D rounded up to even.
void normals(const ArboMcmc* m, uint32_t c, uint32_t it, double* z) {
    for (uint32_t k = 0; k < (D + 1) / 2; ++k) {
        const U4 r = draw(m->key, c, it, Q_NORMAL, k);
        const double u1 = 1.0 - u01(r.x, r.y);   // (0, 1]
        const double u2 = u01(r.z, r.w);
        const double rad = std::sqrt(-2.0 * std::log(u1));
        z[2 * k] = rad * std::cos(kTwoPi * u2);
        z[2 * k + 1] = rad * std::sin(kTwoPi * u2);
    }
}

double log_prior(const ArboMcmc* m, const double* th) {
    double s = 0.0;
    for (uint32_t i = 0; i < D; ++i) s += prior_logpdf(m->desc.priors[i], th[i]);
    return s;
}

void pool_add(ArboMcmc* m, const double* th) {
    m->n_pool += 1.0;
    double delta[D];
    for (uint32_t i = 0; i < D; ++i) {
        delta[i] = th[i] - m->mean[i];
        m->mean[i] += delta[i] / m->n_pool;
    }
    for (uint32_t i = 0; i < D; ++i)
        for (uint32_t j = 0; j < D; ++j) m->m2[i * D + j] += delta[i] * (th[j] - m->mean[j]);
}

// Proposal of chain c at iteration it (see file header).
// Parameters with a fixed prior keep their value.
void pin_fixed(const ArboMcmc* m, double* th) {
    for (uint32_t i = 0; i < D; ++i)
        if (m->desc.priors[i].kind == ARBO_PRIOR_FIXED) th[i] = m->desc.priors[i].a;
}

void propose(const ArboMcmc* m, uint32_t c, uint32_t it, const double* L, bool adapted, double* out) {
    const double* th = &m->theta[(size_t)c * D];
    double z[D + 1];
    normals(m, c, it, z);
    const U4 r = draw(m->key, c, it, Q_MIX, 0);
    const bool fixed = !adapted || u01(r.x, r.y) < m->desc.mix_beta;
    for (uint32_t i = 0; i < D; ++i) {
        double step = 0.0;
        if (fixed) {
            step = m->sd0[i] * z[i];
        } else {
            for (uint32_t k = 0; k <= i; ++k) step += L[i * D + k] * z[k];
        }
        out[i] = th[i] + step;
    }
    for (uint32_t i = 0; i < D; ++i) {
        const ArboPrior& p = m->desc.priors[i];
        if (p.kind == ARBO_PRIOR_CIRCULAR) {
            const double w = p.b - p.a;
            double x = std::fmod(out[i] - p.a, w);
            if (x < 0.0) x += w;
            out[i] = p.a + x;
            if (out[i] >= p.b) out[i] = p.a;
        }
    }
    pin_fixed(m, out);
}

uint64_t sim_seed(const ArboMcmc* m, uint32_t c, uint32_t it, uint32_t particle) {
    const U4 r = draw(m->key, c, it, Q_SIM_SEED, particle);
    return ((uint64_t)r.y << 32) | r.x;
}

// Seeds of chain c's particles for its proposal of iteration it: the
// starting values use the chain's initial seeds; with n_blocks = 0 every
// particle is fresh; with G blocks the particles p = b (mod G) of one
// uniform block b are fresh and the rest keep the chain's seeds.
void particle_seeds(const ArboMcmc* m, uint32_t c, uint32_t it, uint64_t* out) {
    const uint32_t P = m->desc.n_particles, G = m->desc.n_blocks;
    const uint64_t* cur = &m->useed[(size_t)c * P];
    if (it == kInitIter) {
        std::memcpy(out, cur, sizeof(uint64_t) * P);
        return;
    }
    const uint32_t b = G ? mulhi(draw(m->key, c, it, Q_BLOCK, 0).x, G) : 0u;
    for (uint32_t p = 0; p < P; ++p) out[p] = (G == 0 || p % G == b) ? sim_seed(m, c, it, p) : cur[p];
}

// NB log pmf of count k, mean mu, size r (MCMC.cpp:216-243 with r = 1 /
// dispersion).  A zero mean gives -inf for k > 0; for k = 0 it gives 0, or
// -inf under ARBO_MCMC_F_ZERO_WEEK_FLOOR (the reference computes 0 log 0 =
// NaN there and its fmax(-250, NaN) keeps the floor).
double nb_week(uint32_t k, double mu, double r, bool zero_floor = false) {
    if (!(mu > 0.0)) return (k == 0 && !zero_floor) ? 0.0 : -INFINITY;
    // lgamma(k + r) - lgamma(k + 1) - lgamma(r) + k log(mu/(r+mu)) + r log(r/(r+mu))
    const double lrm = std::log(r + mu);
    return std::lgamma((double)k + r) - std::lgamma((double)k + 1.0) - std::lgamma(r) +
           (double)k * (std::log(mu) - lrm) + r * (std::log(r) - lrm);
}

double nb_loglik(const uint32_t* sim_weekly, const uint32_t* data, uint32_t W, uint32_t skip, double rho,
                 double r, double floor_v, bool zero_floor = false) {
    double s = 0.0;
    for (uint32_t w = skip; w < W; ++w) {
        const double v = nb_week(data[w], rho * (double)sim_weekly[w], r, zero_floor);
        s += v > floor_v ? v : floor_v;
    }
    return s;
}

void weekly(const uint32_t* inc, uint32_t rows, uint32_t S, uint32_t W, uint32_t* out) {
    for (uint32_t w = 0; w < W; ++w) {
        uint64_t acc = 0;
        for (uint32_t t = 7 * w; t < 7 * w + 7 && t < rows; ++t)
            for (uint32_t s = 0; s < S; ++s) acc += inc[(size_t)t * S + s];
        out[w] = (uint32_t)acc;
    }
}

// log-likelihoods of n vectors from simulations on the engine; particles of a
// vector are combined by log-mean-exp (MCMC.cpp:252-282).  Vectors the
// engine's own checks would reject (arbo::validate, the test batch_init
// applies) are not simulated: they get -inf, so a proposal the engine cannot
// represent is rejected instead of failing the whole batch.
int eval_engine(ArboMcmc* m, ArboEngine* eng, const ArboEngineApi& api, uint32_t n, const double* th,
                const uint32_t* chains, uint32_t it, double* ll) {
    if (n == 0) return ARBO_OK;
    const uint32_t P = m->desc.n_particles, W = m->desc.n_weeks;
    const uint32_t T = 7 * W - 1, rows = T + 1;
    const uint32_t S = m->base.n_serotypes;
    const bool fit_climate = m->desc.param_set == ARBO_FIT_CLIMATE;
    std::vector<ArboClimate> climates(fit_climate ? n : 0);   // each vector's climate coefficients
    std::vector<ArboParams> vparams(n);
    std::vector<uint32_t> sim_of(n, UINT32_MAX);               // batch index of each vector's first particle
    uint32_t nv = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (fit_climate) {
            arbo_mcmc_sim_climate(th + (size_t)i * D, &m->climate, &climates[i]);
            vparams[i] = m->base;
        } else {
            arbo_mcmc_sim_params(th + (size_t)i * D, &m->base, &vparams[i]);
        }
        const ArboClimate* cl = fit_climate ? &climates[i] : (m->has_climate ? &m->climate : nullptr);
        const ArboShape sh{m->grid_w, m->grid_h, m->humans.data(), m->has_slots ? m->slots.data() : nullptr, cl,
                           m->share.empty() ? nullptr : m->share.data()};
        // the observation model needs a reporting rate >= 0 and a dispersion > 0
        const bool observable = th[(size_t)i * D + kRho] >= 0.0 && th[(size_t)i * D + kDisp] > 0.0;
        if (!observable || !validate(sh, vparams[i], T).empty()) {
            ll[i] = -INFINITY;
            m->n_unsim++;
            continue;
        }
        sim_of[i] = nv * P;
        ++nv;
    }
    if (nv == 0) return ARBO_OK;
    const uint32_t sims = nv * P;
    std::vector<ArboShape> shapes(sims);
    std::vector<ArboParams> params(sims);
    std::vector<uint64_t> seeds(sims);
    for (uint32_t i = 0; i < n; ++i) {
        if (sim_of[i] == UINT32_MAX) continue;
        const ArboClimate* cl = fit_climate ? &climates[i] : (m->has_climate ? &m->climate : nullptr);
        particle_seeds(m, chains[i], it, &seeds[sim_of[i]]);
        for (uint32_t p = 0; p < P; ++p) {
            const uint32_t k = sim_of[i] + p;
            shapes[k] = ArboShape{m->grid_w, m->grid_h, m->humans.data(), m->has_slots ? m->slots.data() : nullptr,
                                  cl, m->share.empty() ? nullptr : m->share.data()};
            params[k] = vparams[i];
        }
    }
    int rc = api.batch_init(eng, sims, shapes.data(), params.data(), seeds.data(), T);
    if (rc == ARBO_OK) rc = api.batch_step(eng, T);
    if (rc != ARBO_OK) return fail(m, rc, std::string("engine: ") + (api.last_error ? api.last_error(eng) : ""));
    std::vector<uint32_t> nc(rows), mc(rows), inf((size_t)rows * S), inc((size_t)rows * S), one((size_t)rows * S);
    std::vector<uint32_t> wk(W);
    std::vector<double> part(P);
    const bool zero_floor = (m->desc.flags & ARBO_MCMC_F_ZERO_WEEK_FLOOR) != 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (sim_of[i] == UINT32_MAX) continue;
        const double rho = th[(size_t)i * D + kRho];
        const double r = 1.0 / th[(size_t)i * D + kDisp];
        for (uint32_t p = 0; p < P; ++p) {
            ArboSeries ser{nc.data(), mc.data(), inf.data(), inc.data(), one.data()};
            rc = api.read_series(eng, sim_of[i] + p, &ser);
            if (rc != ARBO_OK) return fail(m, rc, std::string("engine: ") + (api.last_error ? api.last_error(eng) : ""));
            weekly(inc.data(), rows, S, W, wk.data());
            part[p] = nb_loglik(wk.data(), m->data.data(), W, m->desc.skip_weeks, rho, r, m->desc.week_floor,
                                zero_floor);
        }
        double mx = part[0];
        for (uint32_t p = 1; p < P; ++p) mx = part[p] > mx ? part[p] : mx;
        double se = 0.0;
        for (uint32_t p = 0; p < P; ++p) se += std::exp(part[p] - mx);
        ll[i] = mx + std::log(se) - std::log((double)P);
    }
    m->n_sims += sims;
    return ARBO_OK;
}

// A chain's move to its candidate: parameters, likelihood, prior and, for
// the pseudo-marginal state, the seeds its likelihood came from.
void accept_move(ArboMcmc* m, uint32_t c, uint32_t it, const double* th, double ll, double lp) {
    std::memcpy(&m->theta[(size_t)c * D], th, sizeof(double) * D);
    m->loglik[c] = ll;
    m->logprior[c] = lp;
    const uint32_t P = m->desc.n_particles;
    std::vector<uint64_t> s(P);
    particle_seeds(m, c, it, s.data());
    std::memcpy(&m->useed[(size_t)c * P], s.data(), sizeof(uint64_t) * P);
}

// Appends the iteration's states and candidates to the history.
void record_iteration(ArboMcmc* m, const std::vector<uint8_t>& acc, const std::vector<double>& cand,
                      const std::vector<double>& cand_lp) {
    m->h_theta.insert(m->h_theta.end(), m->theta.begin(), m->theta.end());
    m->h_loglik.insert(m->h_loglik.end(), m->loglik.begin(), m->loglik.end());
    m->h_acc.insert(m->h_acc.end(), acc.begin(), acc.end());
    m->h_cand.insert(m->h_cand.end(), cand.begin(), cand.end());
    m->h_cand_lp.insert(m->h_cand_lp.end(), cand_lp.begin(), cand_lp.end());
}

// Cholesky factor of the adapted proposal covariance 2.38^2/d Sigma_n (+ a
// small ridge); false while the pool is too small (the fixed proposal).
bool am_factor(const ArboMcmc* m, double* L) {
    if (!(m->n_pool >= (double)m->desc.adapt_after && m->n_pool > 1.0)) return false;
    double S[D * D];
    for (uint32_t i = 0; i < D; ++i)
        for (uint32_t j = 0; j < D; ++j) {
            S[i * D + j] = 2.38 * 2.38 / D * m->m2[i * D + j] / (m->n_pool - 1.0);
            if (i == j) S[i * D + j] += 1e-12 + 1e-10 * m->sd0[i] * m->sd0[i];
        }
    return cholesky(S, L);
}

template <class Eval>
int run_iterations(ArboMcmc* m, uint32_t n_iters, Eval eval) {
    const uint32_t C = m->C;
    std::vector<uint32_t> all(C);
    for (uint32_t c = 0; c < C; ++c) all[c] = c;
    using clk = std::chrono::steady_clock;
    if (!m->started) {
        const auto t0 = clk::now();
        int rc = eval(C, m->theta.data(), all.data(), kInitIter, m->loglik.data());
        m->eval_s += std::chrono::duration<double>(clk::now() - t0).count();
        if (rc != ARBO_OK) return rc;
        for (uint32_t c = 0; c < C; ++c) {
            m->logprior[c] = log_prior(m, &m->theta[(size_t)c * D]);
            pool_add(m, &m->theta[(size_t)c * D]);
        }
        m->started = true;
    }
    std::vector<double> prop((size_t)C * D), sub((size_t)C * D), sub_ll(C), cand_lp(C);
    std::vector<uint32_t> sub_chain(C);
    std::vector<uint8_t> acc(C);
    for (uint32_t step = 0; step < n_iters; ++step) {
        const uint32_t it = m->iter;
        double L[D * D];
        const bool adapted = am_factor(m, L);
        uint32_t ns = 0;
        for (uint32_t c = 0; c < C; ++c) {
            propose(m, c, it, L, adapted, &prop[(size_t)c * D]);
            acc[c] = 0;
            cand_lp[c] = -INFINITY;
            if (std::isfinite(log_prior(m, &prop[(size_t)c * D]))) {  // checkParams (MCMC.cpp:164-186)
                std::memcpy(&sub[(size_t)ns * D], &prop[(size_t)c * D], sizeof(double) * D);
                sub_chain[ns++] = c;
            }
        }
        m->n_prop += C;
        m->n_support += ns;
        const auto t0 = clk::now();
        int rc = eval(ns, sub.data(), sub_chain.data(), it, sub_ll.data());
        m->eval_s += std::chrono::duration<double>(clk::now() - t0).count();
        if (rc != ARBO_OK) return rc;
        for (uint32_t k = 0; k < ns; ++k) {
            const uint32_t c = sub_chain[k];
            const double* th = &sub[(size_t)k * D];
            const double lp = log_prior(m, th);
            cand_lp[c] = sub_ll[k] + lp;
            const double log_a = (sub_ll[k] + lp) - (m->loglik[c] + m->logprior[c]);
            const U4 r = draw(m->key, c, it, Q_ACCEPT, 0);
            const double u = 1.0 - u01(r.x, r.y);   // (0, 1]
            if (std::log(u) < log_a) {               // MCMC.cpp:306-344
                accept_move(m, c, it, th, sub_ll[k], lp);
                acc[c] = 1;
                m->n_acc++;
            }
        }
        for (uint32_t c = 0; c < C; ++c) pool_add(m, &m->theta[(size_t)c * D]);
        record_iteration(m, acc, prop, cand_lp);
        m->iter++;
    }
    return ARBO_OK;
}

}  // namespace

extern "C" {

ARBO_API int arbo_mcmc_sim_params(const double* th, const ArboParams* base, ArboParams* out) {
    if (!th || !base || !out) return ARBO_E_INVALID;
    *out = *base;
    out->beta_hm = (float)th[0];
    out->beta_mh = (float)th[1];
    out->long_distance = (float)std::pow(10.0, th[2]);
    out->season_amp = (float)th[3];
    out->season_phase = (float)th[4];
    return ARBO_OK;
}

ARBO_API int arbo_mcmc_sim_climate(const double* th, const ArboClimate* base, ArboClimate* out) {
    if (!th || !base || !out) return ARBO_E_INVALID;
    *out = *base;
    out->scale_eip = (float)th[0];
    out->scale_mle = (float)th[1];
    out->scale_cc = (float)th[2];
    out->scale_humidity = (float)th[3];
    out->min_mos_to_human = (float)th[4];
    return ARBO_OK;
}

ARBO_API int arbo_mcmc_create(const ArboMcmcDesc* desc, const ArboShape* shape, const ArboParams* base,
                              const uint32_t* data, const double* theta0, ArboMcmc** out) {
    if (!out) return ARBO_E_INVALID;
    *out = nullptr;
    if (!desc || !shape || !base || !data || !shape->humans) return ARBO_E_INVALID;
    if (desc->abi_version != ARBO_ABI_VERSION) return ARBO_E_INVALID;
    if (desc->n_chains == 0 || desc->n_chains > ARBO_MCMC_MAX_CHAINS || desc->n_particles == 0 ||
        desc->n_weeks < 2 || desc->n_weeks > (1u << 20) || desc->skip_weeks >= desc->n_weeks ||
        !(desc->mix_beta >= 0.0 && desc->mix_beta <= 1.0) || !(desc->init_scale > 0.0))
        return ARBO_E_INVALID;
    if ((uint64_t)desc->n_chains * desc->n_particles > (1u << 20)) return ARBO_E_INVALID;
    if (desc->n_blocks > desc->n_particles * 1024u || (desc->flags & ~ARBO_MCMC_F_ZERO_WEEK_FLOOR) ||
        desc->reserved0 != 0)
        return ARBO_E_INVALID;
    for (uint32_t i = 0; i < D; ++i)
        if (!valid_prior(desc->priors[i])) return ARBO_E_INVALID;
    // at most six surrogate inputs; the reporting rate and the dispersion are no circular parameters
    if (n_inputs(desc->priors) > kSurIn || desc->priors[kRho].kind == ARBO_PRIOR_CIRCULAR ||
        desc->priors[kDisp].kind == ARBO_PRIOR_CIRCULAR)
        return ARBO_E_INVALID;
    if (desc->param_set > ARBO_FIT_CLIMATE || (desc->param_set == ARBO_FIT_CLIMATE && !shape->climate))
        return ARBO_E_INVALID;
    ArboMcmc* m = new (std::nothrow) ArboMcmc();
    if (!m) return ARBO_E_ALLOC;
    m->desc = *desc;
    m->grid_w = shape->grid_w;
    m->grid_h = shape->grid_h;
    const size_t C_comm = (size_t)shape->grid_w * shape->grid_h;
    m->humans.assign(shape->humans, shape->humans + C_comm);
    if (shape->mosq_share) m->share.assign(shape->mosq_share, shape->mosq_share + C_comm);
    if (shape->climate) {
        const ArboClimate& c = *shape->climate;
        if (!c.temperature || !c.humidity || !c.precipitation) {
            delete m;
            return ARBO_E_INVALID;
        }
        m->clim_t.assign(c.temperature, c.temperature + c.days);
        m->clim_h.assign(c.humidity, c.humidity + c.days);
        m->clim_p.assign(c.precipitation, c.precipitation + c.days);
        m->climate = c;
        m->climate.temperature = m->clim_t.data();
        m->climate.humidity = m->clim_h.data();
        m->climate.precipitation = m->clim_p.data();
        m->has_climate = true;
    }
    if (shape->mosq_slots) {
        m->slots.assign(shape->mosq_slots, shape->mosq_slots + C_comm);
        m->has_slots = true;
    }
    m->base = *base;
    m->data.assign(data, data + desc->n_weeks);
    m->key = Key{(uint32_t)desc->seed, (uint32_t)(desc->seed >> 32)};
    m->C = desc->n_chains;
    m->theta.resize((size_t)m->C * D);
    m->loglik.assign(m->C, -INFINITY);
    m->logprior.assign(m->C, 0.0);
    m->useed.resize((size_t)m->C * desc->n_particles);
    for (uint32_t c = 0; c < m->C; ++c)
        for (uint32_t p = 0; p < desc->n_particles; ++p)
            m->useed[(size_t)c * desc->n_particles + p] = sim_seed(m, c, kInitIter, p);
    for (uint32_t i = 0; i < D; ++i) m->sd0[i] = desc->init_scale * prior_sd(desc->priors[i]) / std::sqrt((double)D);
    for (uint32_t c = 0; c < m->C; ++c)
        for (uint32_t i = 0; i < D; ++i)
            m->theta[(size_t)c * D + i] = theta0 ? theta0[(size_t)c * D + i] : prior_mean(desc->priors[i]);
    for (uint32_t c = 0; c < m->C; ++c)
        if (!std::isfinite(log_prior(m, &m->theta[(size_t)c * D]))) {
            delete m;
            return ARBO_E_INVALID;
        }
    *out = m;
    return ARBO_OK;
}

ARBO_API int arbo_mcmc_run(ArboMcmc* m, ArboEngine* engine, const ArboEngineApi* api, uint32_t n_iters) {
    if (!m) return ARBO_E_INVALID;
    if (!engine) return fail(m, ARBO_E_INVALID, "no engine");
    ArboEngineApi own{arbo_batch_init, arbo_batch_step, arbo_read_series, arbo_last_error};
    const ArboEngineApi a = api ? *api : own;
    if (!a.batch_init || !a.batch_step || !a.read_series) return fail(m, ARBO_E_INVALID, "incomplete engine API");
    return run_iterations(m, n_iters, [&](uint32_t n, const double* th, const uint32_t* chains, uint32_t it, double* ll) {
        return eval_engine(m, engine, a, n, th, chains, it, ll);
    });
}

ARBO_API int arbo_mcmc_run_target(ArboMcmc* m, ArboLogLikFn fn, void* ctx, uint32_t n_iters) {
    if (!m) return ARBO_E_INVALID;
    if (!fn) return fail(m, ARBO_E_INVALID, "no log-likelihood");
    return run_iterations(m, n_iters, [&](uint32_t n, const double* th, const uint32_t*, uint32_t, double* ll) {
        if (n == 0) return (int)ARBO_OK;
        const int rc = fn(ctx, n, th, ll);
        return rc == ARBO_OK ? (int)ARBO_OK : fail(m, rc, "log-likelihood callback failed");
    });
}

ARBO_API int arbo_mcmc_iterations(const ArboMcmc* m, uint32_t* out) {
    if (!m || !out) return ARBO_E_INVALID;
    *out = m->iter;
    return ARBO_OK;
}

ARBO_API int arbo_mcmc_read(const ArboMcmc* m, uint32_t first, uint32_t count, double* theta, double* loglik,
                            uint8_t* accepted) {
    if (!m) return ARBO_E_INVALID;
    if ((uint64_t)first + count > m->iter) return ARBO_E_INVALID;
    const size_t C = m->C;
    if (theta) std::memcpy(theta, m->h_theta.data() + (size_t)first * C * D, sizeof(double) * count * C * D);
    if (loglik) std::memcpy(loglik, m->h_loglik.data() + (size_t)first * C, sizeof(double) * count * C);
    if (accepted) std::memcpy(accepted, m->h_acc.data() + (size_t)first * C, (size_t)count * C);
    return ARBO_OK;
}

ARBO_API int arbo_mcmc_read_candidates(const ArboMcmc* m, uint32_t first, uint32_t count, double* theta,
                                       double* logpost) {
    if (!m) return ARBO_E_INVALID;
    if ((uint64_t)first + count > m->iter) return ARBO_E_INVALID;
    const size_t C = m->C;
    if (theta) std::memcpy(theta, m->h_cand.data() + (size_t)first * C * D, sizeof(double) * count * C * D);
    if (logpost) std::memcpy(logpost, m->h_cand_lp.data() + (size_t)first * C, sizeof(double) * count * C);
    return ARBO_OK;
}

ARBO_API int arbo_mcmc_unsimulable(const ArboMcmc* m, uint64_t* out) {
    if (!m || !out) return ARBO_E_INVALID;
    *out = m->n_unsim;
    return ARBO_OK;
}

ARBO_API int arbo_mcmc_counters(const ArboMcmc* m, uint64_t* proposals, uint64_t* in_support, uint64_t* simulations,
                                uint64_t* accepted, double* eval_seconds) {
    if (!m) return ARBO_E_INVALID;
    if (proposals) *proposals = m->n_prop;
    if (in_support) *in_support = m->n_support;
    if (simulations) *simulations = m->n_sims;
    if (accepted) *accepted = m->n_acc;
    if (eval_seconds) *eval_seconds = m->eval_s;
    return ARBO_OK;
}

ARBO_API int arbo_mcmc_log_prior(const ArboMcmc* m, const double* theta, double* out) {
    if (!m || !theta || !out) return ARBO_E_INVALID;
    *out = log_prior(m, theta);
    return std::isfinite(*out) ? ARBO_OK : ARBO_E_INVALID;
}

ARBO_API void arbo_mcmc_destroy(ArboMcmc* m) { delete m; }

ARBO_API const char* arbo_mcmc_last_error(const ArboMcmc* m) { return m ? m->err.c_str() : "null driver"; }

ARBO_API int arbo_weekly_incidence(const uint32_t* incidence, uint32_t rows, uint32_t serotypes, uint32_t n_weeks,
                                   uint32_t* out) {
    if (!incidence || !out || serotypes == 0 || serotypes > ARBO_MAX_SEROTYPES) return ARBO_E_INVALID;
    weekly(incidence, rows, serotypes, n_weeks, out);
    return ARBO_OK;
}

ARBO_API int arbo_nb_loglik(const uint32_t* sim_weekly, const uint32_t* data, uint32_t n_weeks, uint32_t skip,
                            double rho, double r, double floor_v, double* out) {
    if (!sim_weekly || !data || !out || !(r > 0.0) || !(rho >= 0.0)) return ARBO_E_INVALID;
    *out = nb_loglik(sim_weekly, data, n_weeks, skip, rho, r, floor_v);
    return ARBO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Delayed acceptance with the surrogate (see include/arbo_mcmc.h).

namespace {

enum : uint32_t { Q_MALA_NORMAL = 5, Q_MALA_ACCEPT = 6, Q_SCREEN = 7 };

// Centre and scale of a parameter's surrogate input: (theta - c) / s.
void input_scale(const ArboPrior& p, double& c, double& s) {
    switch (p.kind) {
        case ARBO_PRIOR_TRUNC_NORMAL: c = p.a; s = p.b; return;
        case ARBO_PRIOR_BETA:
            c = p.a / (p.a + p.b);
            s = std::sqrt(p.a * p.b / ((p.a + p.b) * (p.a + p.b) * (p.a + p.b + 1.0)));
            return;
        default: c = 0.5 * (p.a + p.b); s = 0.5 * (p.b - p.a); return;
    }
}

// Surrogate inputs: parameters 0..4 by their priors (a circular one as cos,
// sin), zeros after; at most kSurIn inputs (checked at creation).
void features_of(const ArboMcmc* m, const double* th, float* x) {
    uint32_t j = 0;
    for (uint32_t i = 0; i < kModel; ++i) {
        const ArboPrior& p = m->desc.priors[i];
        if (p.kind == ARBO_PRIOR_CIRCULAR) {
            x[j++] = (float)std::cos(th[i]);
            x[j++] = (float)std::sin(th[i]);
        } else {
            double c, s;
            input_scale(p, c, s);
            x[j++] = (float)((th[i] - c) / s);
        }
    }
    while (j < kSurIn) x[j++] = 0.0f;
}


// log prior gradient (zero for the uniform, circular and fixed priors)
double prior_grad(const ArboPrior& p, double x) {
    if (p.kind == ARBO_PRIOR_BETA) return (p.a - 1.0) / x - (p.b - 1.0) / (1.0 - x);
    if (p.kind == ARBO_PRIOR_TRUNC_NORMAL) return -(x - p.a) / (p.b * p.b);
    return 0.0;
}

// Surrogate log posterior and gradient of n vectors: the prior times the
// NB likelihood (size 1 / dispersion) of the data given the reporting rate
// times the surrogate's mean incidence.  A fixed parameter's gradient is 0.
int sur_logpost(ArboMcmc* m, ArboSurrogate* sur, uint32_t n, const double* th, double* lp, double* grad) {
    if (n == 0) return ARBO_OK;
    const uint32_t W = m->desc.n_weeks;
    std::vector<float> x((size_t)n * kSurIn), f((size_t)n * W), dout((size_t)n * W, 0.0f), dx((size_t)n * kSurIn);
    std::vector<double> drho(n, 0.0), dr(n, 0.0);
    for (uint32_t i = 0; i < n; ++i) features_of(m, th + (size_t)i * D, &x[(size_t)i * kSurIn]);
    int rc = arbo_sur_forward(sur, n, x.data(), f.data());
    if (rc != ARBO_OK) return fail(m, rc, std::string("surrogate: ") + arbo_sur_last_error(sur));
    for (uint32_t i = 0; i < n; ++i) {
        const double rho = th[(size_t)i * D + kRho];
        const double r = 1.0 / th[(size_t)i * D + kDisp];
        if (!(rho >= 0.0) || !(th[(size_t)i * D + kDisp] > 0.0)) {   // no observation model
            lp[i] = -INFINITY;
            continue;
        }
        double ll = 0.0;
        for (uint32_t w = m->desc.skip_weeks; w < W; ++w) {
            const double mu = rho * std::exp((double)f[(size_t)i * W + w]);
            const uint32_t k = m->data[w];
            const double v = nb_week(k, mu, r);
            if (v > m->desc.week_floor) {
                ll += v;
                const double d = (double)k - ((double)k + r) * mu / (r + mu);   // d ll / d log mu
                dout[(size_t)i * W + w] = (float)d;
                drho[i] += d / rho;
                if (grad)   // d ll / d r
                    dr[i] += digamma((double)k + r) - digamma(r) + std::log(r) + 1.0 - std::log(r + mu) -
                             ((double)k + r) / (r + mu);
            } else {
                ll += m->desc.week_floor;
            }
        }
        lp[i] = ll + log_prior(m, th + (size_t)i * D);
    }
    if (grad) {
        rc = arbo_sur_input_grad(sur, n, x.data(), dout.data(), nullptr, dx.data());
        if (rc != ARBO_OK) return fail(m, rc, std::string("surrogate: ") + arbo_sur_last_error(sur));
        for (uint32_t i = 0; i < n; ++i) {
            const double* t = th + (size_t)i * D;
            double* g = grad + (size_t)i * D;
            const float* d = &dx[(size_t)i * kSurIn];
            uint32_t j = 0;   // the inputs of features_of, in order
            for (uint32_t k = 0; k < kModel; ++k) {
                const ArboPrior& p = m->desc.priors[k];
                if (p.kind == ARBO_PRIOR_CIRCULAR) {
                    g[k] = -std::sin(t[k]) * (double)d[j] + std::cos(t[k]) * (double)d[j + 1];
                    j += 2;
                } else {
                    double c, s;
                    input_scale(p, c, s);
                    g[k] = (double)d[j++] / s;
                }
                g[k] += prior_grad(p, t[k]);
            }
            g[kRho] = drho[i] + prior_grad(m->desc.priors[kRho], t[kRho]);
            // r = 1 / d: d ll / d d = -(d ll / d r) / d^2
            g[kDisp] = -dr[i] / (t[kDisp] * t[kDisp]) + prior_grad(m->desc.priors[kDisp], t[kDisp]);
            for (uint32_t k = 0; k < D; ++k)
                if (m->desc.priors[k].kind == ARBO_PRIOR_FIXED) g[k] = 0.0;
        }
    }
    return ARBO_OK;
}

// solves L y = v (lower triangular)
void lower_solve(const double* L, const double* v, double* y) {
    for (uint32_t i = 0; i < D; ++i) {
        double s = v[i];
        for (uint32_t k = 0; k < i; ++k) s -= L[i * D + k] * y[k];
        y[i] = s / L[i * D + i];
    }
}

// -|L^-1 (a - mean)|^2 / (2 eps^2), the circular coordinate by its shortest arc
double log_q(const ArboMcmc* m, const double* L, double eps, const double* a, const double* mean) {
    double d[D], y[D];
    for (uint32_t i = 0; i < D; ++i) {
        d[i] = a[i] - mean[i];
        const ArboPrior& p = m->desc.priors[i];
        if (p.kind == ARBO_PRIOR_CIRCULAR) {
            const double w = p.b - p.a;
            d[i] -= w * std::nearbyint(d[i] / w);
        }
    }
    lower_solve(L, d, y);
    double s = 0.0;
    for (uint32_t i = 0; i < D; ++i) s += y[i] * y[i];
    return -s / (2.0 * eps * eps);
}

void wrap(const ArboMcmc* m, double* th) {
    for (uint32_t i = 0; i < D; ++i) {
        const ArboPrior& p = m->desc.priors[i];
        if (p.kind != ARBO_PRIOR_CIRCULAR) continue;
        const double w = p.b - p.a;
        double x = std::fmod(th[i] - p.a, w);
        if (x < 0.0) x += w;
        th[i] = p.a + x;
        if (th[i] >= p.b) th[i] = p.a;
    }
}

}  // namespace

extern "C" {

ARBO_API int arbo_mcmc_features(const ArboMcmc* m, uint32_t n, const double* theta, float* x) {
    if (!m || !theta || !x) return ARBO_E_INVALID;
    for (uint32_t i = 0; i < n; ++i) features_of(m, theta + (size_t)i * D, x + (size_t)i * kSurIn);
    return ARBO_OK;
}

ARBO_API int arbo_mcmc_surrogate_logpost(ArboMcmc* m, ArboSurrogate* sur, uint32_t n, const double* theta,
                                         double* out, double* grad) {
    if (!m || !sur || !theta || !out) return ARBO_E_INVALID;
    uint32_t in = 0, outw = 0;
    if (arbo_sur_shape(sur, &in, &outw) != ARBO_OK || in != kSurIn || outw != m->desc.n_weeks)
        return fail(m, ARBO_E_INVALID, "surrogate shape does not match (6 inputs, n_weeks outputs)");
    return sur_logpost(m, sur, n, theta, out, grad);
}

}  // extern "C"

namespace {

int check_da(ArboMcmc* m, ArboSurrogate* sur, const ArboDaDesc* da) {
    if (!sur || !da) return fail(m, ARBO_E_INVALID, "surrogate and descriptor are required");
    if (da->mode > ARBO_DA_SCREEN) return fail(m, ARBO_E_INVALID, "bad delayed-acceptance mode");
    if (da->mode == ARBO_DA_LANGEVIN && (da->sub_steps == 0 || da->sub_steps > 1000 || !(da->step0 > 0.0) ||
                                         !(da->target_accept > 0.0 && da->target_accept < 1.0)))
        return fail(m, ARBO_E_INVALID, "bad delayed-acceptance descriptor");
    uint32_t in = 0, outw = 0;
    if (arbo_sur_shape(sur, &in, &outw) != ARBO_OK || in != kSurIn || outw != m->desc.n_weeks)
        return fail(m, ARBO_E_INVALID, "surrogate shape does not match (6 inputs, n_weeks outputs)");
    return ARBO_OK;
}

// Screened delayed acceptance (ARBO_DA_SCREEN): the adaptive-Metropolis
// proposal, a surrogate first stage, the simulator only for the proposals
// that pass it.  `ibm` gives the second stage's log-likelihoods.
template <class Ibm>
int run_screen_impl(ArboMcmc* m, ArboSurrogate* sur, uint32_t n_iters, Ibm ibm) {
    using clk = std::chrono::steady_clock;
    const uint32_t C = m->C;
    std::vector<double> s_lp(C);   // surrogate log posterior of the current states
    {
        const auto t0 = clk::now();
        int rc = sur_logpost(m, sur, C, m->theta.data(), s_lp.data(), nullptr);
        m->sur_s += std::chrono::duration<double>(clk::now() - t0).count();
        if (rc != ARBO_OK) return rc;
    }
    std::vector<double> prop((size_t)C * D), sub((size_t)C * D), sub_slp(C), pass((size_t)C * D), pass_slp(C),
        pass_ll(C), cand_lp(C);
    std::vector<uint32_t> sub_chain(C), pass_chain(C);
    std::vector<uint8_t> acc(C);
    for (uint32_t step = 0; step < n_iters; ++step) {
        const uint32_t it = m->iter;
        double L[D * D];
        const bool adapted = am_factor(m, L);
        uint32_t ns = 0;
        for (uint32_t c = 0; c < C; ++c) {
            propose(m, c, it, L, adapted, &prop[(size_t)c * D]);
            acc[c] = 0;
            cand_lp[c] = -INFINITY;
            if (std::isfinite(log_prior(m, &prop[(size_t)c * D]))) {   // checkParams (MCMC.cpp:164-186)
                cand_lp[c] = NAN;   // until the simulator judges it
                std::memcpy(&sub[(size_t)ns * D], &prop[(size_t)c * D], sizeof(double) * D);
                sub_chain[ns++] = c;
            }
        }
        m->n_prop += C;
        m->n_support += ns;
        // first stage: the surrogate posterior ratio
        const auto t0 = clk::now();
        int rc = sur_logpost(m, sur, ns, sub.data(), sub_slp.data(), nullptr);
        m->sur_s += std::chrono::duration<double>(clk::now() - t0).count();
        if (rc != ARBO_OK) return rc;
        uint32_t np = 0;
        for (uint32_t k = 0; k < ns; ++k) {
            const uint32_t c = sub_chain[k];
            const U4 r = draw(m->key, c, it, Q_SCREEN, 0);
            if (std::log(1.0 - u01(r.x, r.y)) < sub_slp[k] - s_lp[c]) {
                std::memcpy(&pass[(size_t)np * D], &sub[(size_t)k * D], sizeof(double) * D);
                pass_slp[np] = sub_slp[k];
                pass_chain[np++] = c;
            }
        }
        m->da_steps += ns;
        m->da_acc += np;
        // second stage: the simulator at the proposals that passed
        rc = ibm(np, pass.data(), pass_chain.data(), it, pass_ll.data());
        if (rc != ARBO_OK) return rc;
        for (uint32_t k = 0; k < np; ++k) {
            const uint32_t c = pass_chain[k];
            const double* tp = &pass[(size_t)k * D];
            const double lp_new = log_prior(m, tp);
            // L(theta') L_s(theta) / (L(theta) L_s(theta')); the surrogate terms carry the prior
            const double log_a = (pass_ll[k] + lp_new - pass_slp[k]) - (m->loglik[c] + m->logprior[c] - s_lp[c]);
            cand_lp[c] = pass_ll[k] + lp_new;
            const U4 r = draw(m->key, c, it, Q_ACCEPT, 0);
            if (std::log(1.0 - u01(r.x, r.y)) < log_a) {
                accept_move(m, c, it, tp, pass_ll[k], lp_new);
                s_lp[c] = pass_slp[k];
                acc[c] = 1;
                m->n_acc++;
            }
        }
        for (uint32_t c = 0; c < C; ++c) pool_add(m, &m->theta[(size_t)c * D]);
        record_iteration(m, acc, prop, cand_lp);
        m->iter++;
    }
    return ARBO_OK;
}

// Delayed-acceptance iterations; `eval` gives the second stage's
// log-likelihoods (n vectors, their chains, the iteration).
template <class Eval>
int run_da_impl(ArboMcmc* m, ArboSurrogate* sur, const ArboDaDesc* da, uint32_t n_iters, Eval eval) {
    using clk = std::chrono::steady_clock;
    const uint32_t C = m->C;
    auto ibm = [&](uint32_t n, const double* th, const uint32_t* chains, uint32_t it, double* ll) {
        if (n == 0) return (int)ARBO_OK;
        const auto t0 = clk::now();
        const int rc = eval(n, th, chains, it, ll);
        m->eval_s += std::chrono::duration<double>(clk::now() - t0).count();
        return rc;
    };
    std::vector<uint32_t> all(C);
    for (uint32_t c = 0; c < C; ++c) all[c] = c;
    if (!m->started) {
        int rc = ibm(C, m->theta.data(), all.data(), kInitIter, m->loglik.data());
        if (rc != ARBO_OK) return rc;
        for (uint32_t c = 0; c < C; ++c) {
            m->logprior[c] = log_prior(m, &m->theta[(size_t)c * D]);
            pool_add(m, &m->theta[(size_t)c * D]);
        }
        m->started = true;
    }
    if (da->mode == ARBO_DA_SCREEN) return run_screen_impl(m, sur, n_iters, ibm);
    if (!m->da_init) {
        m->da_log_eps = std::log(da->step0);
        m->da_init = true;
    }
    // surrogate log posterior and gradient at the chains' states
    std::vector<double> s_lp(C), s_g((size_t)C * D);
    {
        const auto t0 = clk::now();
        int rc = sur_logpost(m, sur, C, m->theta.data(), s_lp.data(), s_g.data());
        m->sur_s += std::chrono::duration<double>(clk::now() - t0).count();
        if (rc != ARBO_OK) return rc;
    }
    std::vector<double> cur((size_t)C * D), cur_lp(C), cur_g((size_t)C * D), prop((size_t)C * D), p_lp(C),
        p_g((size_t)C * D), sub((size_t)C * D), sub_ll(C), cand_lp(C);
    std::vector<uint32_t> idx(C), sub_chain(C);
    std::vector<uint8_t> moved(C), acc(C);
    for (uint32_t step = 0; step < n_iters; ++step) {
        const uint32_t it = m->iter;
        // preconditioner: the pooled covariance once adapted, else the fixed proposal's
        double L[D * D];
        bool adapted = m->n_pool >= (double)m->desc.adapt_after && m->n_pool > 1.0;
        if (adapted) {
            double S[D * D];
            for (uint32_t i = 0; i < D; ++i)
                for (uint32_t j = 0; j < D; ++j) {
                    S[i * D + j] = m->m2[i * D + j] / (m->n_pool - 1.0);
                    if (i == j) S[i * D + j] += 1e-12 + 1e-10 * m->sd0[i] * m->sd0[i];
                }
            adapted = cholesky(S, L);
        }
        if (!adapted) {
            std::memset(L, 0, sizeof L);
            for (uint32_t i = 0; i < D; ++i) L[i * D + i] = m->sd0[i] * std::sqrt((double)D);
        }
        // fixed parameters: an identity row and column (their gradient is 0 and
        // every proposal is pinned back, so they never move)
        for (uint32_t i = 0; i < D; ++i) {
            if (m->desc.priors[i].kind != ARBO_PRIOR_FIXED) continue;
            for (uint32_t k = 0; k < D; ++k) L[i * D + k] = L[k * D + i] = 0.0;
            L[i * D + i] = 1.0;
        }
        cur = m->theta;
        cur_lp = s_lp;
        cur_g = s_g;
        std::fill(moved.begin(), moved.end(), 0);
        for (uint32_t j = 0; j < da->sub_steps; ++j) {
            const double eps = std::exp(m->da_log_eps);
            const uint32_t sj = it * da->sub_steps + j;
            uint32_t nv = 0;
            for (uint32_t c = 0; c < C; ++c) {
                const double* th = &cur[(size_t)c * D];
                const double* g = &cur_g[(size_t)c * D];
                double mg[D], z[D + 1], mean[D];
                // M g = L (L^T g)
                double ltg[D];
                for (uint32_t i = 0; i < D; ++i) {
                    double s = 0.0;
                    for (uint32_t k = i; k < D; ++k) s += L[k * D + i] * g[k];
                    ltg[i] = s;
                }
                for (uint32_t i = 0; i < D; ++i) {
                    double s = 0.0;
                    for (uint32_t k = 0; k <= i; ++k) s += L[i * D + k] * ltg[k];
                    mg[i] = s;
                }
                for (uint32_t k = 0; k < (D + 1) / 2; ++k) {
                    const U4 r = draw(m->key, c, sj, Q_MALA_NORMAL, k);
                    const double u1 = 1.0 - u01(r.x, r.y), u2 = u01(r.z, r.w);
                    const double rad = std::sqrt(-2.0 * std::log(u1));
                    z[2 * k] = rad * std::cos(kTwoPi * u2);
                    z[2 * k + 1] = rad * std::sin(kTwoPi * u2);
                }
                double* out = &prop[(size_t)c * D];
                for (uint32_t i = 0; i < D; ++i) {
                    mean[i] = th[i] + 0.5 * eps * eps * mg[i];
                    double s = 0.0;
                    for (uint32_t k = 0; k <= i; ++k) s += L[i * D + k] * z[k];
                    out[i] = mean[i] + eps * s;
                }
                wrap(m, out);
                pin_fixed(m, out);
                if (std::isfinite(log_prior(m, out))) {
                    std::memcpy(&sub[(size_t)nv * D], out, sizeof(double) * D);
                    idx[nv++] = c;
                }
            }
            const auto t0 = clk::now();
            int rc = sur_logpost(m, sur, nv, sub.data(), sub_ll.data(), p_g.data());
            m->sur_s += std::chrono::duration<double>(clk::now() - t0).count();
            if (rc != ARBO_OK) return rc;
            uint32_t n_acc = 0;
            for (uint32_t k = 0; k < nv; ++k) {
                const uint32_t c = idx[k];
                const double* th = &cur[(size_t)c * D];
                const double* tp = &sub[(size_t)k * D];
                const double* gp = &p_g[(size_t)k * D];
                const double* g = &cur_g[(size_t)c * D];
                // forward and backward Langevin means
                double mf[D], mb[D], ltg[D], mg[D];
                for (int pass = 0; pass < 2; ++pass) {
                    const double* gg = pass == 0 ? g : gp;
                    const double* base = pass == 0 ? th : tp;
                    double* mm = pass == 0 ? mf : mb;
                    for (uint32_t i = 0; i < D; ++i) {
                        double s = 0.0;
                        for (uint32_t q = i; q < D; ++q) s += L[q * D + i] * gg[q];
                        ltg[i] = s;
                    }
                    for (uint32_t i = 0; i < D; ++i) {
                        double s = 0.0;
                        for (uint32_t q = 0; q <= i; ++q) s += L[i * D + q] * ltg[q];
                        mg[i] = s;
                        mm[i] = base[i] + 0.5 * eps * eps * mg[i];
                    }
                }
                const double log_a = (sub_ll[k] - cur_lp[c]) + log_q(m, L, eps, th, mb) - log_q(m, L, eps, tp, mf);
                const U4 r = draw(m->key, c, sj, Q_MALA_ACCEPT, 0);
                if (std::log(1.0 - u01(r.x, r.y)) < log_a) {
                    std::memcpy(&cur[(size_t)c * D], tp, sizeof(double) * D);
                    cur_lp[c] = sub_ll[k];
                    std::memcpy(&cur_g[(size_t)c * D], gp, sizeof(double) * D);
                    moved[c] = 1;
                    ++n_acc;
                }
            }
            m->da_steps += C;
            m->da_acc += n_acc;
            // step size: diminishing Robbins-Monro on the surrogate acceptance
            m->da_t += 1.0;
            m->da_log_eps += std::pow(m->da_t + 10.0, -0.6) * ((double)n_acc / (double)C - da->target_accept);
        }
        // second stage: the simulator at the chains that moved
        uint32_t ns = 0;
        for (uint32_t c = 0; c < C; ++c) {
            acc[c] = 0;
            cand_lp[c] = NAN;   // a chain that did not move has no candidate for the simulator
            if (!moved[c]) continue;
            std::memcpy(&sub[(size_t)ns * D], &cur[(size_t)c * D], sizeof(double) * D);
            sub_chain[ns++] = c;
        }
        m->n_prop += C;
        m->n_support += ns;
        int rc = ibm(ns, sub.data(), sub_chain.data(), it, sub_ll.data());
        if (rc != ARBO_OK) return rc;
        for (uint32_t k = 0; k < ns; ++k) {
            const uint32_t c = sub_chain[k];
            const double* tp = &cur[(size_t)c * D];
            const double lp_new = log_prior(m, tp);
            // L(theta') L_s(theta) / (L(theta) L_s(theta')); the surrogate terms carry the prior
            const double log_a = (sub_ll[k] + lp_new - cur_lp[c]) - (m->loglik[c] + m->logprior[c] - s_lp[c]);
            cand_lp[c] = sub_ll[k] + lp_new;
            const U4 r = draw(m->key, c, it, Q_ACCEPT, 0);
            if (std::log(1.0 - u01(r.x, r.y)) < log_a) {
                accept_move(m, c, it, tp, sub_ll[k], lp_new);
                s_lp[c] = cur_lp[c];
                std::memcpy(&s_g[(size_t)c * D], &cur_g[(size_t)c * D], sizeof(double) * D);
                acc[c] = 1;
                m->n_acc++;
            }
        }
        for (uint32_t c = 0; c < C; ++c) pool_add(m, &m->theta[(size_t)c * D]);
        record_iteration(m, acc, cur, cand_lp);
        m->iter++;
    }
    return ARBO_OK;
}

}  // namespace

extern "C" {

ARBO_API int arbo_mcmc_run_da(ArboMcmc* m, ArboEngine* engine, const ArboEngineApi* api, ArboSurrogate* sur,
                              const ArboDaDesc* da, uint32_t n_iters) {
    if (!m) return ARBO_E_INVALID;
    if (!engine) return fail(m, ARBO_E_INVALID, "no engine");
    if (int rc = check_da(m, sur, da)) return rc;
    ArboEngineApi own{arbo_batch_init, arbo_batch_step, arbo_read_series, arbo_last_error};
    const ArboEngineApi a = api ? *api : own;
    if (!a.batch_init || !a.batch_step || !a.read_series) return fail(m, ARBO_E_INVALID, "incomplete engine API");
    return run_da_impl(m, sur, da, n_iters,
                       [&](uint32_t n, const double* th, const uint32_t* chains, uint32_t it, double* ll) {
                           return eval_engine(m, engine, a, n, th, chains, it, ll);
                       });
}

ARBO_API int arbo_mcmc_run_da_target(ArboMcmc* m, ArboLogLikFn fn, void* ctx, ArboSurrogate* sur,
                                     const ArboDaDesc* da, uint32_t n_iters) {
    if (!m) return ARBO_E_INVALID;
    if (!fn) return fail(m, ARBO_E_INVALID, "no log-likelihood");
    if (int rc = check_da(m, sur, da)) return rc;
    return run_da_impl(m, sur, da, n_iters, [&](uint32_t n, const double* th, const uint32_t*, uint32_t, double* ll) {
        const int rc = fn(ctx, n, th, ll);
        return rc == ARBO_OK ? (int)ARBO_OK : fail(m, rc, "log-likelihood callback failed");
    });
}

ARBO_API int arbo_mcmc_da_counters(const ArboMcmc* m, uint64_t* steps, uint64_t* accepted, double* seconds,
                                   double* step) {
    if (!m) return ARBO_E_INVALID;
    if (steps) *steps = m->da_steps;
    if (accepted) *accepted = m->da_acc;
    if (seconds) *seconds = m->sur_s;
    if (step) *step = std::exp(m->da_log_eps);
    return ARBO_OK;
}

}  // extern "C"
