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

using namespace arbo;

namespace {

constexpr uint32_t D = ARBO_MCMC_PARAMS;
constexpr double kTwoPi = 6.283185307179586;

// Philox purposes of the driver (key: the driver's seed).
enum : uint32_t { Q_MIX = 1, Q_NORMAL = 2, Q_ACCEPT = 3, Q_SIM_SEED = 4 };
constexpr uint32_t kInitIter = 0xFFFFFFFFu;

double prior_mean(const ArboPrior& p) {
    if (p.kind == ARBO_PRIOR_BETA) return p.a / (p.a + p.b);
    return 0.5 * (p.a + p.b);
}

double prior_sd(const ArboPrior& p) {
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
    }
    return -INFINITY;
}

bool valid_prior(const ArboPrior& p) {
    if (!(std::isfinite(p.a) && std::isfinite(p.b))) return false;
    if (p.kind == ARBO_PRIOR_BETA) return p.a > 0.0 && p.b > 0.0;
    if (p.kind == ARBO_PRIOR_UNIFORM || p.kind == ARBO_PRIOR_CIRCULAR) return p.b > p.a;
    return false;
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
    ArboParams base{};
    std::vector<uint32_t> data;
    Key key{0, 0};
    uint32_t C = 0;
    // chain state
    std::vector<double> theta, loglik, logprior;
    bool started = false;
    uint32_t iter = 0;
    // pooled adaptation (Welford over every chain state after every iteration)
    double n_pool = 0.0;
    double mean[D] = {};
    double m2[D * D] = {};
    double sd0[D] = {};
    // history
    std::vector<double> h_theta, h_loglik;
    std::vector<uint8_t> h_acc;
    // counters
    uint64_t n_prop = 0, n_support = 0, n_sims = 0, n_acc = 0;
    double eval_s = 0.0;
    std::string err;
};

namespace {

int fail(ArboMcmc* m, int code, const std::string& msg) {
    if (m) m->err = msg;
    return code;
}

double u01(uint32_t a, uint32_t b) { return u53(a, b); }

// Six standard normals of (chain, iteration) by Box-Muller on 53-bit uniforms.
void normals(const ArboMcmc* m, uint32_t c, uint32_t it, double* z) {
    for (uint32_t k = 0; k < D / 2; ++k) {
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
void propose(const ArboMcmc* m, uint32_t c, uint32_t it, const double* L, bool adapted, double* out) {
    const double* th = &m->theta[(size_t)c * D];
    double z[D];
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
}

uint64_t sim_seed(const ArboMcmc* m, uint32_t c, uint32_t it, uint32_t particle) {
    const U4 r = draw(m->key, c, it, Q_SIM_SEED, particle);
    return ((uint64_t)r.y << 32) | r.x;
}

double nb_week(uint32_t k, double mu, double r) {
    if (!(mu > 0.0)) return k == 0 ? 0.0 : -INFINITY;
    // lgamma(k + r) - lgamma(k + 1) - lgamma(r) + k log(mu/(r+mu)) + r log(r/(r+mu))
    const double lrm = std::log(r + mu);
    return std::lgamma((double)k + r) - std::lgamma((double)k + 1.0) - std::lgamma(r) +
           (double)k * (std::log(mu) - lrm) + r * (std::log(r) - lrm);
}

double nb_loglik(const uint32_t* sim_weekly, const uint32_t* data, uint32_t W, uint32_t skip, double rho,
                 double r, double floor_v) {
    double s = 0.0;
    for (uint32_t w = skip; w < W; ++w) {
        const double v = nb_week(data[w], rho * (double)sim_weekly[w], r);
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
// vector are combined by log-mean-exp (MCMC.cpp:252-282)
int eval_engine(ArboMcmc* m, ArboEngine* eng, const ArboEngineApi& api, uint32_t n, const double* th,
                const uint32_t* chains, uint32_t it, double* ll) {
    if (n == 0) return ARBO_OK;
    const uint32_t P = m->desc.n_particles, W = m->desc.n_weeks;
    const uint32_t T = 7 * W - 1, rows = T + 1;
    const uint32_t S = m->base.n_serotypes;
    const uint32_t sims = n * P;
    std::vector<ArboShape> shapes(sims);
    std::vector<ArboParams> params(sims);
    std::vector<uint64_t> seeds(sims);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t p = 0; p < P; ++p) {
            const uint32_t k = i * P + p;
            shapes[k] = ArboShape{m->grid_w, m->grid_h, m->humans.data(), m->has_slots ? m->slots.data() : nullptr};
            arbo_mcmc_sim_params(th + (size_t)i * D, &m->base, &params[k]);
            seeds[k] = sim_seed(m, chains[i], it, p);
        }
    int rc = api.batch_init(eng, sims, shapes.data(), params.data(), seeds.data(), T);
    if (rc == ARBO_OK) rc = api.batch_step(eng, T);
    if (rc != ARBO_OK) return fail(m, rc, std::string("engine: ") + (api.last_error ? api.last_error(eng) : ""));
    std::vector<uint32_t> nc(rows), mc(rows), inf((size_t)rows * S), inc((size_t)rows * S), one((size_t)rows * S);
    std::vector<uint32_t> wk(W);
    std::vector<double> part(P);
    for (uint32_t i = 0; i < n; ++i) {
        const double rho = th[(size_t)i * D + 5];
        for (uint32_t p = 0; p < P; ++p) {
            ArboSeries ser{nc.data(), mc.data(), inf.data(), inc.data(), one.data()};
            rc = api.read_series(eng, i * P + p, &ser);
            if (rc != ARBO_OK) return fail(m, rc, std::string("engine: ") + (api.last_error ? api.last_error(eng) : ""));
            weekly(inc.data(), rows, S, W, wk.data());
            part[p] = nb_loglik(wk.data(), m->data.data(), W, m->desc.skip_weeks, rho, m->desc.nb_size,
                                m->desc.week_floor);
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
    std::vector<double> prop((size_t)C * D), sub((size_t)C * D), sub_ll(C);
    std::vector<uint32_t> sub_chain(C);
    std::vector<uint8_t> acc(C);
    for (uint32_t step = 0; step < n_iters; ++step) {
        const uint32_t it = m->iter;
        // adapted proposal covariance 2.38^2/d Sigma_n (+ a small ridge)
        double L[D * D];
        bool adapted = m->n_pool >= (double)m->desc.adapt_after && m->n_pool > 1.0;
        if (adapted) {
            double S[D * D];
            for (uint32_t i = 0; i < D; ++i)
                for (uint32_t j = 0; j < D; ++j) {
                    S[i * D + j] = 2.38 * 2.38 / D * m->m2[i * D + j] / (m->n_pool - 1.0);
                    if (i == j) S[i * D + j] += 1e-12 + 1e-10 * m->sd0[i] * m->sd0[i];
                }
            adapted = cholesky(S, L);
        }
        uint32_t ns = 0;
        for (uint32_t c = 0; c < C; ++c) {
            propose(m, c, it, L, adapted, &prop[(size_t)c * D]);
            acc[c] = 0;
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
            const double log_a = (sub_ll[k] + lp) - (m->loglik[c] + m->logprior[c]);
            const U4 r = draw(m->key, c, it, Q_ACCEPT, 0);
            const double u = 1.0 - u01(r.x, r.y);   // (0, 1]
            if (std::log(u) < log_a) {               // MCMC.cpp:306-344
                std::memcpy(&m->theta[(size_t)c * D], th, sizeof(double) * D);
                m->loglik[c] = sub_ll[k];
                m->logprior[c] = lp;
                acc[c] = 1;
                m->n_acc++;
            }
        }
        for (uint32_t c = 0; c < C; ++c) pool_add(m, &m->theta[(size_t)c * D]);
        m->h_theta.insert(m->h_theta.end(), m->theta.begin(), m->theta.end());
        m->h_loglik.insert(m->h_loglik.end(), m->loglik.begin(), m->loglik.end());
        m->h_acc.insert(m->h_acc.end(), acc.begin(), acc.end());
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

ARBO_API int arbo_mcmc_create(const ArboMcmcDesc* desc, const ArboShape* shape, const ArboParams* base,
                              const uint32_t* data, const double* theta0, ArboMcmc** out) {
    if (!out) return ARBO_E_INVALID;
    *out = nullptr;
    if (!desc || !shape || !base || !data || !shape->humans) return ARBO_E_INVALID;
    if (desc->abi_version != ARBO_ABI_VERSION) return ARBO_E_INVALID;
    if (desc->n_chains == 0 || desc->n_chains > ARBO_MCMC_MAX_CHAINS || desc->n_particles == 0 ||
        desc->n_weeks < 2 || desc->n_weeks > (1u << 20) || desc->skip_weeks >= desc->n_weeks ||
        !(desc->nb_size > 0.0) || !(desc->mix_beta >= 0.0 && desc->mix_beta <= 1.0) || !(desc->init_scale > 0.0))
        return ARBO_E_INVALID;
    if ((uint64_t)desc->n_chains * desc->n_particles > (1u << 20)) return ARBO_E_INVALID;
    for (uint32_t i = 0; i < D; ++i)
        if (!valid_prior(desc->priors[i])) return ARBO_E_INVALID;
    ArboMcmc* m = new (std::nothrow) ArboMcmc();
    if (!m) return ARBO_E_ALLOC;
    m->desc = *desc;
    m->grid_w = shape->grid_w;
    m->grid_h = shape->grid_h;
    const size_t C_comm = (size_t)shape->grid_w * shape->grid_h;
    m->humans.assign(shape->humans, shape->humans + C_comm);
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
