// arbo_oracle.cpp — TEST INFRASTRUCTURE ONLY: the serial / OpenMP CPU oracle
// of the dengue IBM day loop.  It exports the engine's C ABI (include/arbo_c.h)
// so tests can drive it exactly like the GPU engine, and it is what the GPU
// engine must match bit-for-bit on every census word and series value.  No
// product code links or loads it.
//
// It restates the reference's per-day algorithm (simulation.cpp:115-155):
//   human ageing, death/rebirth, E->I and I->S   ndemographic.cu:32-105
//   mosquito ageing, death, E->I, emergence       mdemographic.cu:22-150
//   event counts and dispersal                    nepidemic.cu:28-142, mepidemic.cu:28-138
//   event application onto mosquitoes             mepidemic.cu:143-246
//   event application onto humans + introductions nepidemic.cu:146-319
//   time-series counters                          datacollect.cu:10-189
//   t=0 census                                    demographicintial.cu:86-297, epidemicinitial.cu:27-99
//   per-community series, infection history,      datacollect.cu:10-23, demographicintial.cu:158-171,
//   age-of-infection rings                        nepidemic.cu:229-298, ndemographic.cu:96-102
//   climate-driven days and the logistic target   SimulationParam.cpp:92-137, simulation.cpp:84-132
// with these deliberate changes, the same in the engine:
//   - stateless Philox keyed on (seed; entity, day, purpose, index) instead of
//     curand states picked by SM/warp/lane id;
//   - lifetimes as integer death days: the reference's daily test
//     PLE > F(age) becomes age >= death_age with death_age = min{a>=1: F(a) >= PLE},
//     F tabulated once in double, PLE a 24-bit uniform;
//   - Poisson event counts (local ~ Poisson(lambda(1-w)), long-distance
//     ~ Poisson(lambda w)) instead of decimalResolve splits; local
//     destinations from per-row alias tables; the long-distance events of a
//     simulation are drawn as one Poisson total (superposition) with uniform
//     destinations, so only their count per destination is kept;
//   - emergence per community into its lowest-index dead slots;
//   - mosquito targets uniform over alive slots (the reference retries dead
//     picks until an alive one is found: the same law);
//   - fixed event order per destination: serotype rotation from a keyed
//     start, local events by source community then event index, then the
//     long-distance arrivals; the k-th event of a (source, serotype, slot)
//     segment draws its target from word k%4 of Philox block k/4.
// Serial semantics throughout: each event checks the target's current state.
#include "arbo_c.h"
#include "oracle_rng.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#else
static int omp_get_max_threads() { return 1; }
#endif

using namespace orc;

namespace {

constexpr uint32_t HUMAN_AGE_TABLE = 65536;  // death ages 1..65535
constexpr uint32_t MOSQ_AGE_TABLE = 2048;    // death ages 1..2047
constexpr uint32_t ONE24 = 1u << 24;

struct Sim {
    uint32_t W = 0, H = 0, C = 0, S = 0;
    Seed key{0, 0};
    ArboParams p{};
    std::vector<uint32_t> N, cap;
    std::vector<uint64_t> h_off, m_off;
    uint64_t NH = 0, NM = 0;
    std::vector<float> base;
    std::vector<uint32_t> row_ptr, to, thr, alias, self_col;
    std::vector<uint32_t> in_ptr, in_src, in_slot;
    std::vector<uint32_t> Kh, Km, Thy, Tmd;
    std::vector<float> dens;
    // climate: per-day beta_hm, beta_mh, eip and the lifetime table of each
    // birth day (the t = 0 census keeps Km)
    bool climate = false;
    std::vector<float> day_bhm, day_bmh, day_eip;
    std::vector<std::vector<uint32_t>> Kmb;
    std::vector<uint32_t> imm_thr;
    uint32_t seed_thr_h = 0, seed_thr_m = 0;
    // census
    std::vector<uint32_t> hw;
    std::vector<uint16_t> hn;
    std::vector<int32_t> hb;
    std::vector<uint16_t> mw, mn;
    // per-day
    std::vector<uint32_t> cnt;      // [C][2][4][9]
    std::vector<uint32_t> Ih, Im, inc, Mc;
    std::vector<uint64_t> ldfx;     // [C][2][4] long-distance rate, 2^-24 fixed point
    std::vector<uint64_t> lam;      // [2][4] summed long-distance rate
    std::vector<uint32_t> nld;      // [C][2][4] long-distance arrivals of the day
    // series
    std::vector<uint32_t> n_count, m_count, n_inf, n_inc, n_one;
    std::vector<uint32_t> spatial;   // [(T+1)][S][C] infectious humans per community (ARBO_F_SPATIAL)
    std::vector<int32_t> hist;       // [human][4] infection history (ARBO_F_HISTORY), else empty
    // age-of-infection rings (ARBO_F_AOI): marks of today's infections by
    // mosquito events, rings [4][AOI_RECORD] and their next positions
    std::vector<uint8_t> aoi_mark;
    std::vector<uint32_t> aoi;
    uint32_t aoi_pos[4] = {0, 0, 0, 0};
    uint32_t aoi_from = 0;
};

constexpr uint32_t AOI_RECORD = 5000;   // C_NAOIRECORD (constant.h:22)
constexpr uint32_t AOI_WINDOW = 3650;   // 10 * C_YEAR (nepidemic.cu:229)

}  // namespace

struct ArboEngine {
    std::vector<Sim> sims;
    uint32_t day = 0, max_days = 0;
    int threads = 0;
    bool ready = false;
    std::string err;
    ArboTuning tuning{4, 2, 0, 8, 0, 1};   // launch knobs: accepted and reported, no effect on the CPU
    uint32_t flags = 0;
    uint64_t scan_h = 0, scan_m = 0;    // census vectors of the days stepped (every member, every day)
    uint64_t full_h = 0, full_m = 0;    // one full scan: 4-human and 8-slot vectors of every community
};

namespace {

int fail(ArboEngine* e, int code, const char* msg) {
    if (e) e->err = msg;
    return code;
}

double pw(double x, double y) { return x > 0.0 ? exp_d(y * log_d(x)) : 0.0; }

// Human bi-Weibull CDF of ndemographic.cu:13-29 (b multiplies the age).
double cdf_h(const ArboParams& p, double a) {
    double h = pw(a * (double)p.h_infant_scale, (double)p.h_infant_shape);
    if (a >= (double)p.h_loc) h = h + pw((a - (double)p.h_loc) / (double)p.h_scale, (double)p.h_shape);
    return 1.0 - exp_d(-h);
}

// Mosquito Weibull CDF of mdemographic.cu:13-17.
double cdf_m(const ArboParams& p, double a) {
    return 1.0 - exp_d(-pw(a / (double)p.m_scale, (double)p.m_shape));
}

uint32_t to24(double f) {
    double v = std::floor(f * 16777216.0);
    if (!(v > 0.0)) return 0;
    if (v >= 16777216.0) return ONE24;
    return (uint32_t)v;
}

// ---- climate-driven days (SimulationParam::adjustClimate,
// SimulationParam.cpp:92-137; the logistic target, simulation.cpp:84-132) ----

// Briere curve c T (T - lo) sqrt(hi - T), zero outside [lo, hi]
// (SimulationParam.cpp:114-136).
double briere(double T, double c, double lo, double hi) {
    return (T < lo || T > hi) ? 0.0 : c * T * (T - lo) * std::sqrt(hi - T);
}

// {beta_hm, beta_mh, eip, mosquito Weibull scale} of climate row r.
void clim_day(const ArboClimate& c, uint32_t r, double out[4]) {
    const double T = (double)c.temperature[r], H = (double)c.humidity[r];
    const double TK = T + 273.15;
    // EIP from the temperature (SimulationParam.cpp:97-100)
    const double rate = 24.0 * 0.003359 * (TK / 298.0) * exp_d((15000.0 / 1.986) * ((1.0 / 298.0) - (1.0 / TK)));
    // mosquito death rate quartic and humidity (SimulationParam.cpp:104-107)
    const double mu = 0.8692 - 0.159 * T + 0.01116 * (T * T) - 0.0003408 * (T * T * T) + 0.000003809 * (T * T * T * T);
    out[0] = (double)c.biting_rate * briere(T, 0.000491, 12.22, 37.46);
    out[1] = (double)c.biting_rate * briere(T, 0.000849, 17.05, 35.83);
    out[2] = 1.0 / (rate * (double)c.scale_eip);
    out[3] = 1.0 / (mu * (double)c.scale_mle * pw(1.0 - H, (double)c.scale_humidity));
}

// minMosToHuman N (1 + P)^scaleCC (simulation.cpp:87,131).
double clim_capacity(const ArboClimate& c, uint32_t r, double N) {
    return (double)c.min_mos_to_human * N * pw(1.0 + (double)c.precipitation[r], (double)c.scale_cc);
}

// h_mExpectedPopSize: M(0) = K(0), M(t+1) = M(t) + g M(t) (1 - M(t)/K(t+1)), in float.
std::vector<float> clim_target(const ArboClimate& c, double N, uint32_t max_days) {
    std::vector<float> M(max_days + 1);
    M[0] = (float)clim_capacity(c, 0, N);
    for (uint32_t t = 0; t < max_days; ++t) {
        double m = (double)M[t];
        M[t + 1] = (float)(m + (double)c.mos_net_growth * m * (1.0 - m / clim_capacity(c, t + 1, N)));
    }
    return M;
}

const char* clim_invalid(const ArboClimate& c, uint32_t max_days) {
    if (!c.temperature || !c.humidity || !c.precipitation) return "climate arrays missing";
    if (c.days < max_days + 1) return "climate shorter than max_days + 1 days";
    if (!(c.biting_rate >= 0.0f) || !(c.scale_eip > 0.0f) || !(c.scale_mle > 0.0f) || !(c.min_mos_to_human > 0.0f) ||
        !(c.mos_net_growth >= 0.0f) || !std::isfinite(c.scale_cc) || !std::isfinite(c.scale_humidity))
        return "bad climate coefficients";
    for (uint32_t r = 0; r <= max_days; ++r) {
        if (!std::isfinite(c.temperature[r]) || !(c.humidity[r] >= 0.0f && c.humidity[r] < 1.0f) ||
            !(c.precipitation[r] > -1.0f) || !std::isfinite(c.precipitation[r]))
            return "climate row out of range";
        double d[4];
        clim_day(c, r, d);
        if (!(d[2] > 0.0 && d[2] < 1000.0)) return "climate gives eip out of range";
        if (!(d[3] > 0.0 && d[3] < 1.0e4)) return "climate gives a mosquito lifespan scale out of range";
    }
    return nullptr;
}

// Climate tables of a simulation: the days' parameters (day t reads row
// t + 1), the target M(t)/N, and for each birth day b the first-passage
// table of mdemographic.cu:72 under the day's scale: entry a is the largest
// floor(F(a'; scale(b + a')) 2^24) over a' <= a (days past the run keep the
// last day's scale), up to its first 2^24 entry.
void init_climate(Sim& s, const ArboClimate& c, uint32_t max_days) {
    s.climate = true;
    s.day_bhm.resize(max_days); s.day_bmh.resize(max_days); s.day_eip.resize(max_days);
    std::vector<double> scale(max_days);
    for (uint32_t t = 0; t < max_days; ++t) {
        double d[4];
        clim_day(c, t + 1, d);
        s.day_bhm[t] = (float)d[0]; s.day_bmh[t] = (float)d[1]; s.day_eip[t] = (float)d[2];
        scale[t] = d[3];
    }
    std::vector<float> M = clim_target(c, (double)s.NH, max_days);
    s.dens.resize(max_days + 1);
    for (uint32_t t = 0; t <= max_days; ++t) s.dens[t] = (float)((double)M[t] / (double)s.NH);
    s.Kmb.assign(max_days, {});
    for (uint32_t b = 0; b < max_days; ++b) {
        std::vector<uint32_t>& K = s.Kmb[b];
        K.assign(1, 0u);
        uint32_t run = 0;
        for (uint32_t a = 1; a < MOSQ_AGE_TABLE && run < ONE24; ++a) {
            const uint32_t day = (uint32_t)std::min<uint64_t>((uint64_t)b + a, max_days - 1);
            run = std::max(run, to24(1.0 - exp_d(-pw((double)a / scale[day], (double)s.p.m_shape))));
            K.push_back(run);
        }
    }
}

// Lifetime threshold table: K[a] = floor(F(a) 2^24), non-decreasing.
template <class F>
std::vector<uint32_t> death_table(uint32_t n, F cdf) {
    std::vector<uint32_t> K(n);
    for (uint32_t a = 0; a < n; ++a) {
        K[a] = to24(cdf((double)a));
        if (a > 0 && K[a] < K[a - 1]) K[a] = K[a - 1];
    }
    return K;
}

// Initial-age table from the cumulative survival sums of
// demographicintial.cu:48-83: T[i] = floor(S[i]/S[last] 2^24).
template <class F>
std::vector<uint32_t> age_table(uint32_t n, uint32_t step, F cdf) {
    std::vector<double> S(n);
    double acc = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        acc = acc + (1.0 - cdf((double)(i * step)));
        S[i] = acc;
    }
    std::vector<uint32_t> T(n);
    for (uint32_t i = 0; i < n; ++i) T[i] = to24(S[i] / S[n - 1]);
    T[n - 1] = ONE24;
    for (uint32_t i = 1; i < n; ++i)
        if (T[i] < T[i - 1]) T[i] = T[i - 1];
    return T;
}

// min{a in [1, n-1] : K[a] >= k}, else n-1 (K is non-decreasing).
uint32_t death_age(const std::vector<uint32_t>& K, uint32_t k) {
    auto it = std::lower_bound(K.begin() + 1, K.end(), k);
    if (it == K.end()) return (uint32_t)K.size() - 1;
    return (uint32_t)(it - K.begin());
}

// min{i : k <= T[i]} (T is non-decreasing and ends at 2^24).
uint32_t pick_index(const std::vector<uint32_t>& T, uint32_t k) {
    auto it = std::lower_bound(T.begin(), T.end(), k);
    if (it == T.end()) return (uint32_t)T.size() - 1;
    return (uint32_t)(it - T.begin());
}

// Death age conditional on having reached age a0: PLE uniform on (F(a0-1), 1]
// (demographicintial.cu:143-152, 266-275).
uint32_t death_age_given(const std::vector<uint32_t>& K, uint32_t a0, uint32_t r) {
    uint32_t prev = K[a0 >= 1 ? a0 - 1 : 0];
    if (prev >= ONE24) return a0;
    uint32_t k = prev + 1 + range(r, ONE24 - prev);
    return death_age(K, k);
}

// Non-wrapping lattice CSR (network.cpp:31-41, rows in ascending column order
// as network.cpp:95-113) and Vose alias tables of the dispersal weights.
void build_network(Sim& s) {
    const ArboParams& p = s.p;
    s.row_ptr.assign(1, 0);
    s.self_col.clear();
    s.to.clear(); s.thr.clear(); s.alias.clear();
    for (uint32_t y = 0; y < s.H; ++y) {
        for (uint32_t x = 0; x < s.W; ++x) {
            std::vector<uint32_t> cols;
            std::vector<double> w;
            for (int dy = -1; dy <= 1; ++dy) {
                for (int dx = -1; dx <= 1; ++dx) {
                    int nx = (int)x + dx, ny = (int)y + dy;
                    if (nx < 0 || ny < 0 || nx >= (int)s.W || ny >= (int)s.H) continue;
                    int manh = std::abs(dx) + std::abs(dy);
                    // self: a dispersal target with include_self; with noDistance > 0
                    // also the column staying events are counted in (weight 0
                    // when it is no dispersal target)
                    const bool stay_col = p.no_distance > 0.0f;
                    if (manh == 0 && !p.include_self && !stay_col) continue;
                    if (p.neighbourhood == 4 && manh == 2) continue;
                    double dist = manh == 0 ? 0.0 : (manh == 1 ? 1.0 : 1.4142135623730951);
                    double wt = p.dispersal_scale > 0.0f ? exp_d(-dist / (double)p.dispersal_scale) : 1.0;
                    if (manh == 0 && !p.include_self) wt = 0.0;
                    if (manh == 0) s.self_col.push_back((uint32_t)cols.size());
                    cols.push_back((uint32_t)ny * s.W + (uint32_t)nx);
                    w.push_back(wt);
                }
            }
            if (cols.empty() || (cols.size() == 1 && w[0] == 0.0)) {  // isolated: every local event stays home
                if (cols.empty()) s.self_col.push_back(0);
                cols.assign(1, y * s.W + x);
                w.assign(1, 1.0);
            }
            if (s.self_col.size() < s.row_ptr.size()) s.self_col.push_back(0);   // no own column: never used
            uint32_t n = (uint32_t)cols.size();
            double tot = 0.0;
            for (double v : w) tot = tot + v;
            std::vector<double> q(n), prob(n, 1.0);
            std::vector<uint32_t> al(n);
            for (uint32_t i = 0; i < n; ++i) { q[i] = w[i] / tot * (double)n; al[i] = i; }
            std::vector<uint32_t> small, large;
            for (uint32_t i = 0; i < n; ++i) (q[i] < 1.0 ? small : large).push_back(i);
            while (!small.empty() && !large.empty()) {
                uint32_t l = small.back(); small.pop_back();
                uint32_t g = large.back(); large.pop_back();
                prob[l] = q[l];
                al[l] = g;
                q[g] = (q[g] + q[l]) - 1.0;
                (q[g] < 1.0 ? small : large).push_back(g);
            }
            for (uint32_t i = 0; i < n; ++i) {
                s.to.push_back(cols[i]);
                s.thr.push_back(to24(prob[i]));
                s.alias.push_back(al[i]);
            }
            s.row_ptr.push_back((uint32_t)s.to.size());
        }
    }
    // in-edges of each destination, ascending source
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> in(s.C);
    for (uint32_t c = 0; c < s.C; ++c)
        for (uint32_t e = s.row_ptr[c]; e < s.row_ptr[c + 1]; ++e)
            in[s.to[e]].push_back({c, e - s.row_ptr[c]});
    s.in_ptr.assign(1, 0);
    s.in_src.clear(); s.in_slot.clear();
    for (uint32_t d = 0; d < s.C; ++d) {
        std::sort(in[d].begin(), in[d].end());
        for (auto& pr : in[d]) { s.in_src.push_back(pr.first); s.in_slot.push_back(pr.second); }
        s.in_ptr.push_back((uint32_t)s.in_src.size());
    }
}

int validate(ArboEngine* e, const ArboShape& sh, const ArboParams& p, uint32_t max_days) {
    if (p.n_serotypes < 1 || p.n_serotypes > ARBO_MAX_SEROTYPES) return fail(e, ARBO_E_INVALID, "n_serotypes must be 1..4");
    if (sh.grid_w < 1 || sh.grid_h < 1 || !sh.humans) return fail(e, ARBO_E_INVALID, "empty lattice");
    if ((uint64_t)sh.grid_w * sh.grid_h > (1u << 24)) return fail(e, ARBO_E_INVALID, "too many communities");
    if (p.neighbourhood != 4 && p.neighbourhood != 8) return fail(e, ARBO_E_INVALID, "neighbourhood must be 4 or 8");
    if (!(p.beta_hm >= 0.0f) || !(p.beta_mh >= 0.0f)) return fail(e, ARBO_E_INVALID, "beta must be >= 0");
    if (!(p.long_distance >= 0.0f && p.long_distance <= 1.0f)) return fail(e, ARBO_E_INVALID, "long_distance must be in [0,1]");
    if (!(p.season_amp >= 0.0f && p.season_amp <= 1.0f)) return fail(e, ARBO_E_INVALID, "season_amp must be in [0,1]");
    if (!(p.recovery >= 0.0f && p.recovery < 1000.0f) || !(p.eip >= 0.0f && p.eip < 1000.0f)) return fail(e, ARBO_E_INVALID, "recovery/eip out of range");
    if (p.iip_max < p.iip_min || p.iip_max > 1000) return fail(e, ARBO_E_INVALID, "bad incubation range");
    if (!(p.m_scale > 0.0f) || !(p.m_shape > 0.0f) || !(p.h_scale > 0.0f) || !(p.h_shape > 0.0f)) return fail(e, ARBO_E_INVALID, "bad lifespan parameters");
    // integer lifetimes end at the tables' last ages: the tail past them must be
    // below the 24-bit resolution of the lifetime draws, F >= 1 - 2^-24 there
    if (to24(cdf_h(p, (double)(HUMAN_AGE_TABLE - 1))) < ONE24 - 1)
        return fail(e, ARBO_E_INVALID, "human lifespan beyond the 65,535-day lifetime table");
    if (to24(cdf_m(p, (double)(MOSQ_AGE_TABLE - 1))) < ONE24 - 1)
        return fail(e, ARBO_E_INVALID, "mosquito lifespan beyond the 2,047-day lifetime table");
    if (sh.climate && !clim_invalid(*sh.climate, max_days)) {
        // every birth day's first-passage row lies above F(a; the largest daily scale)
        double smax = 0.0;
        for (uint32_t r = 1; r <= max_days; ++r) {
            double d[4];
            clim_day(*sh.climate, r, d);
            smax = std::max(smax, d[3]);
        }
        if (to24(1.0 - exp_d(-pw((double)(MOSQ_AGE_TABLE - 1) / smax, (double)p.m_shape))) < ONE24 - 1)
            return fail(e, ARBO_E_INVALID, "climate gives a mosquito lifespan beyond the 2,047-day lifetime table");
    }
    if (!(p.m_ratio >= 0.0f)) return fail(e, ARBO_E_INVALID, "m_ratio must be >= 0");
    if (!(p.external_infection >= 0.0f)) return fail(e, ARBO_E_INVALID, "external_infection must be >= 0");
    if (!(p.seed_frac_h >= 0.0f && p.seed_frac_h <= 1.0f) || !(p.seed_frac_m >= 0.0f && p.seed_frac_m <= 1.0f)) return fail(e, ARBO_E_INVALID, "seed fractions must be in [0,1]");
    if (p.immunity_mode > 1 || !(p.sero_prev >= 0.0f && p.sero_prev <= 1.0f)) return fail(e, ARBO_E_INVALID, "bad immunity settings");
    if (p.count_law > ARBO_COUNT_STOCH_ROUND) return fail(e, ARBO_E_INVALID, "bad count_law");
    if (!(p.no_distance >= 0.0f && p.no_distance <= 1.0f)) return fail(e, ARBO_E_INVALID, "no_distance must be in [0,1]");
    if (!(p.dispersal_scale >= 0.0f)) return fail(e, ARBO_E_INVALID, "dispersal_scale must be >= 0");
    return ARBO_OK;
}

void init_sim(Sim& s, uint32_t max_days) {
    const ArboParams& p = s.p;
    const uint32_t S = s.S;
    // lifespan and initial-age tables
    s.Kh = death_table(HUMAN_AGE_TABLE, [&](double a) { return cdf_h(p, a); });
    s.Km = death_table(MOSQ_AGE_TABLE, [&](double a) { return cdf_m(p, a); });
    s.Thy = age_table(151, 365, [&](double a) { return cdf_h(p, a); });
    s.Tmd = age_table(301, 1, [&](double a) { return cdf_m(p, a); });
    // seasonal density 1 + A sin(2 pi t/365 + phi) (a climate set it already)
    if (!s.climate) {
        s.dens.resize(max_days + 1);
        for (uint32_t t = 0; t <= max_days; ++t) {
            double arg = 6.283185307179586 * (double)t / 365.0 + (double)p.season_phase;
            s.dens[t] = (float)(1.0 + (double)p.season_amp * sin_d(arg));
        }
    }
    // immunity thresholds per (community, serotype)
    s.imm_thr.assign((size_t)s.C * S, 0);
    for (uint32_t c = 0; c < s.C; ++c)
        for (uint32_t v = 0; v < S; ++v) {
            if (p.immunity_mode == 0) {
                s.imm_thr[(size_t)c * S + v] = to24((double)p.sero_prev);
            } else {  // Beta(2,5) = 2nd smallest of 6 uniforms
                W4 a = rnd(s.key, c, DAY_INIT, tag(IMMUNE, 0, v, 0), 0);
                W4 b = rnd(s.key, c, DAY_INIT, tag(IMMUNE, 0, v, 0), 1);
                uint32_t u[6] = {a.v[0] >> 8, a.v[1] >> 8, a.v[2] >> 8, a.v[3] >> 8, b.v[0] >> 8, b.v[1] >> 8};
                std::sort(u, u + 6);
                s.imm_thr[(size_t)c * S + v] = u[1];
            }
        }
    s.seed_thr_h = to24((double)p.seed_frac_h);
    s.seed_thr_m = to24((double)p.seed_frac_m);
    build_network(s);

    s.hw.assign(s.NH, 0); s.hn.assign(s.NH, 0); s.hb.assign(s.NH, 0);
    s.mw.assign(s.NM, 3); s.mn.assign(s.NM, 0);
    s.cnt.assign((size_t)s.C * 72, 0);
    s.Ih.assign((size_t)s.C * S, 0); s.Im.assign((size_t)s.C * S, 0);
    s.inc.assign((size_t)s.C * S, 0); s.Mc.assign(s.C, 0);
    s.ldfx.assign((size_t)s.C * 8, 0);
    s.lam.assign(8, 0);
    s.nld.assign((size_t)s.C * 8, 0);
    s.n_count.assign(max_days + 1, 0); s.m_count.assign(max_days + 1, 0);
    s.n_inf.assign((size_t)(max_days + 1) * S, 0);
    s.n_inc.assign((size_t)(max_days + 1) * S, 0);
    s.n_one.assign((size_t)(max_days + 1) * S, 0);
    s.spatial.clear();
}

// t=0 census of one community (demographicintial.cu:86-297 and
// epidemicinitial.cu:27-99, with Philox-chosen seed infections).
void init_community(Sim& s, uint32_t c) {
    const ArboParams& p = s.p;
    const uint32_t S = s.S;
    uint32_t Ih[4] = {0, 0, 0, 0}, Im[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < s.N[c]; ++i) {
        uint32_t h = (uint32_t)(s.h_off[c] + i);
        W4 r0 = rnd(s.key, h, DAY_INIT, INIT_H, 0);
        uint32_t y = pick_index(s.Thy, (r0.v[0] >> 8) + 1);
        uint32_t a0 = 365 * y + range(r0.v[1], 365);
        uint32_t da = death_age_given(s.Kh, a0, r0.v[2]);
        int32_t ddi = (int32_t)da - (int32_t)(a0 + 1);
        uint32_t dd = ddi < 0 ? 0u : (uint32_t)ddi;
        W4 r1 = rnd(s.key, h, DAY_INIT, INIT_H, 1);
        uint32_t imm = 0;
        for (uint32_t v = 0; v < S; ++v)
            if ((r1.v[v] >> 8) < s.imm_thr[(size_t)c * S + v]) imm |= 1u << v;
        uint32_t st = 0, strain = 0;
        uint16_t nxt = 0;
        if ((r0.v[3] >> 8) < s.seed_thr_h) {
            W4 r2 = rnd(s.key, h, DAY_INIT, INIT_H, 2);
            strain = range(r2.v[0], S);
            st = 2;
            imm |= 1u << strain;
            uint32_t R = sround(p.recovery, r2.v[2]);
            if (R < 1) R = 1;
            nxt = (uint16_t)range(r2.v[1], R);
            Ih[strain]++;
        }
        s.hw[h] = (dd << 8) | (imm << 4) | (strain << 2) | st;
        s.hn[h] = nxt;
        s.hb[h] = -(int32_t)(a0 + 1);
        if (!s.hist.empty())   // demographicintial.cu:158-171 (-2 immune, -1 naive), epidemicinitial.cu:59
            for (uint32_t v = 0; v < 4; ++v) s.hist[(size_t)h * 4 + v] = ((imm >> v) & 1u) ? -2 : -1;
    }
    W4 rt = rnd(s.key, c, DAY_INIT, INIT_MT, 0);
    uint32_t alive = sround(s.dens[0] * s.base[c], rt.v[0]);
    if (alive > s.cap[c]) alive = s.cap[c];
    for (uint32_t j = 0; j < s.cap[c]; ++j) {
        uint32_t m = (uint32_t)(s.m_off[c] + j);
        s.mn[m] = 0;
        if (j >= alive) { s.mw[m] = 3; continue; }
        W4 r0 = rnd(s.key, m, DAY_INIT, INIT_M, 0);
        uint32_t a0 = pick_index(s.Tmd, (r0.v[0] >> 8) + 1);
        uint32_t da = death_age_given(s.Km, a0, r0.v[2]);
        int32_t ddi = (int32_t)da - (int32_t)(a0 + 1);
        uint32_t dd12 = (uint32_t)ddi & 0xFFFu;
        uint32_t st = 0, strain = 0;
        if ((r0.v[3] >> 8) < s.seed_thr_m) {
            st = 2;
            strain = range(r0.v[1], S);
            Im[strain]++;
        }
        s.mw[m] = (uint16_t)((dd12 << 4) | (strain << 2) | st);
    }
    s.Mc[c] = alive;
    for (uint32_t v = 0; v < S; ++v) {
        s.Ih[(size_t)c * S + v] = Ih[v];
        s.Im[(size_t)c * S + v] = Im[v];
        s.inc[(size_t)c * S + v] = 0;
    }
}

// Series row t from the per-community counters (datacollect.cu:79-189 as
// plain sums; community 0 is the reference's h_specialSubPop).
void record_row(Sim& s, uint32_t t) {
    const uint32_t S = s.S;
    uint64_t nc = 0, mc = 0;
    uint64_t inf[4] = {0, 0, 0, 0}, inc[4] = {0, 0, 0, 0};
    for (uint32_t c = 0; c < s.C; ++c) {
        nc += s.N[c];
        mc += s.Mc[c];
        for (uint32_t v = 0; v < S; ++v) {
            inf[v] += s.Ih[(size_t)c * S + v];
            inc[v] += s.inc[(size_t)c * S + v];
        }
    }
    s.n_count[t] = (uint32_t)nc;
    s.m_count[t] = (uint32_t)mc;
    for (uint32_t v = 0; v < S; ++v) {
        s.n_inf[(size_t)t * S + v] = (uint32_t)inf[v];
        s.n_inc[(size_t)t * S + v] = (uint32_t)inc[v];
        s.n_one[(size_t)t * S + v] = s.Ih[v];
    }
    if (!s.spatial.empty())   // d_nInfectedSubPopCountSeries (datacollect.cu:10-23)
        for (uint32_t v = 0; v < S; ++v)
            for (uint32_t c = 0; c < s.C; ++c) s.spatial[((size_t)t * S + v) * s.C + c] = s.Ih[(size_t)c * S + v];
}

// Daily event count of source community c, direction dir and serotype v with
// rate lam > 0 (nVisitingInfected, nepidemic.cu:49-136; mVisitingInfected,
// mepidemic.cu:48-132).  Returns the events dispersed over the CSR row; stay
// gets those staying at the source (only when nd > 0) and fx the
// long-distance part: under ARBO_COUNT_STOCH_ROUND an event count, under
// ARBO_COUNT_POISSON the rate in 2^-24 units (the simulation's sum is drawn
// once).
//   stochastic rounding (nepidemic.cu:67-70, 98-104): n = stochRound(lam),
//   stay = stochRound(n * noDistance), long-distance =
//   stochRound((n - stay) * longDistance), the rest dispersed;
//   Poisson: stay ~ Poisson(lam noDistance), and of the rest
//   g = lam (1 - noDistance): dispersed ~ Poisson(g (1 - longDistance)),
//   long-distance ~ Poisson(g longDistance), all independent.
uint32_t source_counts(uint32_t law, float lam, float omega, float nd, Seed key, uint32_t c, uint32_t t,
                       uint32_t dir, uint32_t v, uint64_t& fx, uint32_t& stay) {
    const bool staying = nd > 0.0f;
    stay = 0;
    if (law == ARBO_COUNT_STOCH_ROUND) {
        W4 r = rnd(key, c, t, tag(NSR, dir, v, 0), 0);
        uint32_t all = sround(lam, r.v[0]);
        if (staying) stay = sround((float)all * nd, r.v[1]);
        all -= stay;
        const uint32_t ld = sround((float)all * omega, r.v[2]);
        fx = ld;
        return all - ld;
    }
    const float go = staying ? lam * (1.0f - nd) : lam;
    const float mu_loc = go * (1.0f - omega);
    const float mu_ld = go * omega;
    fx = (uint64_t)std::floor((double)mu_ld * 16777216.0);
    if (staying) stay = poisson_d((double)(lam * nd), key, c, t, tag(NSTAY, dir, v, 0));
    return poisson_d((double)mu_loc, key, c, t, tag(NLOC, dir, v, 0));
}

// Phase A of day t for one community: demography, counters, event counts.
void phase_a(Sim& s, uint32_t c, uint32_t t) {
    const ArboParams& p = s.p;
    const uint32_t S = s.S;
    uint32_t Ih[4] = {0, 0, 0, 0}, Im[4] = {0, 0, 0, 0}, inc[4] = {0, 0, 0, 0};
    // humans: nBlockBirthDeath (ndemographic.cu:32-105)
    for (uint32_t i = 0; i < s.N[c]; ++i) {
        uint32_t h = (uint32_t)(s.h_off[c] + i);
        uint32_t w = s.hw[h];
        uint32_t st = w & 3u, strain = (w >> 2) & 3u;
        if (t >= (w >> 8)) {  // death and instant rebirth
            W4 r = rnd(s.key, h, t, H_BIRTH, 0);
            uint32_t da = death_age(s.Kh, (r.v[0] >> 8) + 1);
            s.hw[h] = (t + da) << 8;
            s.hb[h] = (int32_t)t;
            if (!s.hist.empty())   // ndemographic.cu:96-102: the newborn has no history
                for (uint32_t v = 0; v < 4; ++v) s.hist[(size_t)h * 4 + v] = -1;
            continue;
        }
        if (st == 2) {
            if ((int16_t)(uint16_t)(t - s.hn[h]) > 0) {
                s.hw[h] = w & ~0xFu;   // recovered: susceptible to the other serotypes
            } else {
                Ih[strain]++;
            }
        } else if (st == 1) {
            if ((int16_t)(uint16_t)(t - s.hn[h]) > 0) {
                W4 r = rnd(s.key, h, t, H_REC, 0);
                s.hn[h] = (uint16_t)(s.hn[h] + sround(p.recovery, r.v[0]));
                s.hw[h] = (w & ~3u) | 2u;
                inc[strain]++;
                Ih[strain]++;
            }
        }
    }
    // mosquitoes: mBlockBirthDeath (mdemographic.cu:22-150) per community
    const uint32_t cap = s.cap[c];
    thread_local std::vector<uint8_t> dead, died;
    dead.assign(cap, 0);
    died.assign(cap, 0);
    uint32_t alive = 0, ndead = 0;
    for (uint32_t j = 0; j < cap; ++j) {
        uint32_t m = (uint32_t)(s.m_off[c] + j);
        uint32_t w = s.mw[m];
        uint32_t st = w & 3u;
        if (st == 3) { dead[j] = 1; ndead++; continue; }
        if (((t - (w >> 4)) & 0xFFFu) < 2048u) { dead[j] = 1; died[j] = 1; ndead++; continue; }
        if (st == 1 && (int16_t)(uint16_t)(t - s.mn[m]) > 0) {
            st = 2;
            s.mw[m] = (uint16_t)((w & ~3u) | 2u);
        }
        alive++;
        if (st == 2) Im[(w >> 2) & 3u]++;
    }
    // the day-t target is the expected population of the next day, E[t+1]
    // (simulation.cpp:129-132 computes it before the day's demography)
    W4 rt = rnd(s.key, c, t, M_TARGET, 0);
    uint32_t target = sround(s.dens[t + 1] * s.base[c], rt.v[0]);
    uint32_t births = target > alive ? target - alive : 0;
    if (births > ndead) births = ndead;
    uint32_t rank = 0;
    for (uint32_t j = 0; j < cap; ++j) {
        if (!dead[j]) continue;
        uint32_t m = (uint32_t)(s.m_off[c] + j);
        if (rank < births) {
            // birth `rank` of the community draws word rank%4 of block rank/4
            W4 r = rnd(s.key, c, t, M_BIRTH, rank >> 2);
            uint32_t da = death_age(s.climate ? s.Kmb[t] : s.Km, (r.v[rank & 3u] >> 8) + 1);
            s.mw[m] = (uint16_t)(((t + da) & 0xFFFu) << 4);
        } else if (died[j]) {
            s.mw[m] = 3;
        }
        rank++;
    }
    const uint32_t Mc = alive + births;
    s.Mc[c] = Mc;
    for (uint32_t v = 0; v < S; ++v) {
        s.Ih[(size_t)c * S + v] = Ih[v];
        s.Im[(size_t)c * S + v] = Im[v];
        s.inc[(size_t)c * S + v] = inc[v];
    }
    // event counts and alias destinations (nVisitingInfected / mVisitingInfected)
    uint32_t* cnt = &s.cnt[(size_t)c * 72];
    std::fill(cnt, cnt + 72, 0u);
    const uint32_t r0 = s.row_ptr[c], rl = s.row_ptr[c + 1] - r0;
    for (uint32_t dir = 0; dir < 2; ++dir) {
        for (uint32_t v = 0; v < 4; ++v) {
            uint64_t fx = 0;
            if (v < S) {
                float lam = 0.0f;
                const float bhm = s.climate ? s.day_bhm[t] : p.beta_hm;   // the day's (adjustClimate)
                const float bmh = s.climate ? s.day_bmh[t] : p.beta_mh;
                if (dir == HM && Ih[v] > 0) lam = (float)Mc / (float)s.N[c] * bhm * (float)Ih[v];
                if (dir == MH && Im[v] > 0) lam = bmh * (float)Im[v];
                if (lam > 0.0f) {
                    uint32_t stay;
                    const uint32_t n = source_counts(p.count_law, lam, p.long_distance, p.no_distance, s.key, c, t,
                                                     dir, v, fx, stay);
                    cnt[(dir * 4 + v) * 9 + s.self_col[c]] += stay;
                    // event k takes word k%4 of block k/4; the product word * row
                    // length gives the column (high half) and the alias test
                    // uniform (low half)
                    for (uint32_t k = 0; k < n; ++k) {
                        W4 r = rnd(s.key, c, t, tag(DEST, dir, v, 0), k >> 2);
                        const uint64_t prod = (uint64_t)r.v[k & 3] * rl;
                        uint32_t col = (uint32_t)(prod >> 32);
                        if (((uint32_t)prod >> 8) >= s.thr[r0 + col]) col = s.alias[r0 + col];
                        cnt[(dir * 4 + v) * 9 + col]++;
                    }
                }
            }
            s.ldfx[(size_t)c * 8 + dir * 4 + v] = fx;
        }
    }
}

struct Event { uint32_t s, x0, x2, k; };

// Ordered event list of destination d in direction dir (see file header).
void events_of(const Sim& s, uint32_t d, uint32_t dir, uint32_t s0,
               const uint32_t* nld, std::vector<std::vector<Event>>& per_pos) {
    const uint32_t S = s.S;
    per_pos.assign(S, {});
    for (uint32_t pos = 0; pos < S; ++pos) {
        uint32_t v = (s0 + pos) % S;
        for (uint32_t e = s.in_ptr[d]; e < s.in_ptr[d + 1]; ++e) {
            uint32_t src = s.in_src[e], slot = s.in_slot[e];
            uint32_t n = s.cnt[(size_t)src * 72 + (dir * 4 + v) * 9 + slot];
            for (uint32_t k = 0; k < n; ++k) per_pos[pos].push_back({v, src, tag(TGT, dir, v, slot), k});
        }
        for (uint32_t k = 0; k < nld[v]; ++k) per_pos[pos].push_back({v, d, tag(TGT_LD, dir, v, 0), k});
    }
}

void infect_human(Sim& s, uint32_t h, uint32_t v, uint32_t t) {
    W4 r = rnd(s.key, h, t, H_IIP, 0);
    uint32_t iip = s.p.iip_min + range(r.v[0], s.p.iip_max - s.p.iip_min + 1);
    s.hw[h] = (s.hw[h] & ~0xFu) | (1u << (4 + v)) | (v << 2) | 1u;
    s.hn[h] = (uint16_t)(t + iip);
    if (!s.hist.empty()) s.hist[(size_t)h * 4 + v] = (int32_t)t - s.hb[h];   // age at exposure (nepidemic.cu:248,298)
}

bool human_eligible(uint32_t w, uint32_t v) { return (w & 3u) == 0 && !((w >> (4 + v)) & 1u); }

// Phase B of day t for destination d: nmTransmission then mnTransmission.
void phase_b(Sim& s, uint32_t d, uint32_t t) {
    const ArboParams& p = s.p;
    const uint32_t S = s.S;
    uint32_t nld[2][4];
    for (uint32_t k = 0; k < 8; ++k) nld[k >> 2][k & 3] = s.nld[(size_t)d * 8 + k];
    std::vector<std::vector<Event>> ev;
    // onto mosquitoes (mepidemic.cu:143-246)
    {
        uint32_t s0 = range(rnd(s.key, d, t, tag(ROT, HM, 0, 0), 0).v[0], S);
        events_of(s, d, HM, s0, nld[HM], ev);
        std::vector<uint32_t> alive_slots;
        for (uint32_t j = 0; j < s.cap[d]; ++j)
            if ((s.mw[s.m_off[d] + j] & 3u) != 3u) alive_slots.push_back(j);
        const uint32_t Md = (uint32_t)alive_slots.size();
        if (Md > 0) {
            for (uint32_t pos = 0; pos < S; ++pos)
                for (const Event& E : ev[pos]) {
                    W4 r = rnd(s.key, E.x0, t, E.x2, E.k >> 2);
                    uint32_t m = (uint32_t)(s.m_off[d] + alive_slots[range(r.v[E.k & 3], Md)]);
                    uint32_t w = s.mw[m];
                    if ((w & 3u) != 0) continue;
                    W4 r2 = rnd(s.key, m, t, M_EIP, 0);
                    s.mw[m] = (uint16_t)((w & ~0xFu) | (E.s << 2) | 1u);
                    s.mn[m] = (uint16_t)(t + sround(s.climate ? s.day_eip[t] : p.eip, r2.v[0]));
                }
        }
    }
    // onto humans (nepidemic.cu:146-319)
    {
        uint32_t s0 = range(rnd(s.key, d, t, tag(ROT, MH, 0, 0), 0).v[0], S);
        events_of(s, d, MH, s0, nld[MH], ev);
        const uint32_t Nd = s.N[d];
        for (uint32_t pos = 0; pos < S; ++pos) {
            uint32_t v = (s0 + pos) % S;
            for (const Event& E : ev[pos]) {
                W4 r = rnd(s.key, E.x0, t, E.x2, E.k >> 2);
                uint32_t h = (uint32_t)(s.h_off[d] + range(r.v[E.k & 3], Nd));
                if (human_eligible(s.hw[h], v)) {
                    infect_human(s, h, v, t);
                    if (!s.aoi_mark.empty() && t >= s.aoi_from) s.aoi_mark[h] = 1;   // nepidemic.cu:229-243
                }
            }
            if (p.external_infection > 0.0f) {  // nepidemic.cu:261-309
                uint32_t x = sround(p.external_infection / (float)s.C,
                                    rnd(s.key, d, t, tag(INTRO, 0, v, 0), 0).v[0]);
                for (uint32_t att = 0; x > 0 && att < 1000; ++att) {
                    W4 r = rnd(s.key, d, t, tag(INTRO, 0, v, 0), 1 + att);
                    uint32_t h = (uint32_t)(s.h_off[d] + range(r.v[0], Nd));
                    if (human_eligible(s.hw[h], v)) { infect_human(s, h, v, t); x--; }
                }
            }
        }
    }
}

// Long-distance arrivals of day t (nepidemic.cu:118-126: each long-distance
// event goes to a uniform random community).  The simulation's long-distance
// events of one direction and serotype are L ~ Poisson(sum of the sources'
// long-distance rates); event l lands on a uniform community.
void long_distance(Sim& s, uint32_t t) {
    std::fill(s.nld.begin(), s.nld.end(), 0u);
    for (uint32_t dir = 0; dir < 2; ++dir)
        for (uint32_t v = 0; v < s.S; ++v) {
            const uint32_t k = dir * 4 + v;
            const uint32_t L = s.p.count_law == ARBO_COUNT_STOCH_ROUND
                                   ? (uint32_t)s.lam[k]   // the sources' own long-distance counts
                                   : poisson_d((double)s.lam[k] * 0x1p-24, s.key, DAY_INIT, t, tag(NLD, dir, v, 0));
            for (uint32_t l = 0; l < L; ++l) {
                const uint32_t d = range(rnd(s.key, l, t, tag(LD_DEST, dir, v, 0), 0).v[0], s.C);
                s.nld[(size_t)d * 8 + k]++;
            }
        }
}

// The day's infections by mosquito events enter the age-of-infection rings
// (nepidemic.cu:229-243) in human order: ring k = the serotypes the human had
// been infected with before (its history words other than today's, which is
// the only one set today), slot = the ring's next position, cycling through
// AOI_RECORD slots like atomicInc(&count, C_NAOIRECORD - 1).
void aoi_day(Sim& s, uint32_t t) {
    for (uint64_t h = 0; h < s.NH; ++h) {
        if (!s.aoi_mark[h]) continue;
        s.aoi_mark[h] = 0;
        uint32_t n = 0;
        for (uint32_t v = 0; v < 4; ++v) n += s.hist[h * 4 + v] != -1;
        const uint32_t k = n - 1;
        s.aoi[(size_t)k * AOI_RECORD + s.aoi_pos[k]] = (uint32_t)((int32_t)t - s.hb[h]);
        s.aoi_pos[k] = (s.aoi_pos[k] + 1) % AOI_RECORD;
    }
}

void sim_day(Sim& s, uint32_t t, int threads) {
    const int C = (int)s.C;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads) if (threads > 1)
    for (int c = 0; c < C; ++c) phase_a(s, (uint32_t)c, t);
    std::fill(s.lam.begin(), s.lam.end(), 0ull);
    for (uint32_t c = 0; c < s.C; ++c)
        for (uint32_t k = 0; k < 8; ++k) s.lam[k] += s.ldfx[(size_t)c * 8 + k];
    record_row(s, t + 1);
    long_distance(s, t);
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads) if (threads > 1)
    for (int d = 0; d < C; ++d) phase_b(s, (uint32_t)d, t);
    if (!s.aoi_mark.empty()) aoi_day(s, t);
}

int thread_count(const ArboEngine* e) { return e->threads > 0 ? e->threads : omp_get_max_threads(); }

// Simulations are independent: many small ones run one per thread, a few
// large ones run with their communities spread over the threads.  Both orders
// give identical results.
void day_step(ArboEngine* e, uint32_t t) {
    const int nt = thread_count(e);
    const int n = (int)e->sims.size();
    if (n >= 2 * nt) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
        for (int i = 0; i < n; ++i) sim_day(e->sims[i], t, 1);
    } else {
        for (Sim& s : e->sims) sim_day(s, t, nt);
    }
}

}  // namespace

extern "C" {

int arbo_create(const ArboEngineDesc* desc, ArboEngine** out) {
    if (!out) return ARBO_E_INVALID;
    *out = nullptr;
    if (!desc || desc->abi_version != ARBO_ABI_VERSION) return ARBO_E_INVALID;
    ArboEngine* e = new ArboEngine();
    e->threads = (int)desc->threads;
    e->flags = desc->flags;
    *out = e;
    return ARBO_OK;
}

int arbo_batch_init(ArboEngine* e, uint32_t n_sims, const ArboShape* shapes,
                    const ArboParams* params, const uint64_t* seeds, uint32_t max_days) {
    if (!e) return ARBO_E_INVALID;
    if (n_sims == 0 || !shapes || !params || !seeds) return fail(e, ARBO_E_INVALID, "empty batch");
    if (max_days == 0 || max_days > (1u << 23)) return fail(e, ARBO_E_INVALID, "max_days must be 1..2^23");
    e->ready = false;
    for (uint32_t i = 0; i < n_sims; ++i) {
        int rc = validate(e, shapes[i], params[i], max_days);
        if (rc) return rc;
    }
    e->sims.clear();
    e->sims.resize(n_sims);
    for (uint32_t i = 0; i < n_sims; ++i) {
        Sim& s = e->sims[i];
        const ArboShape& sh = shapes[i];
        s.p = params[i];
        s.W = sh.grid_w; s.H = sh.grid_h; s.C = s.W * s.H; s.S = s.p.n_serotypes;
        s.key = {(uint32_t)seeds[i], (uint32_t)(seeds[i] >> 32)};
        s.N.assign(sh.humans, sh.humans + s.C);
        s.cap.resize(s.C);
        s.base.resize(s.C);
        s.h_off.resize(s.C); s.m_off.resize(s.C);
        uint64_t nh = 0, nm = 0;
        for (uint32_t c = 0; c < s.C; ++c) {
            if (s.N[c] < 1 || s.N[c] > ARBO_MAX_COMMUNITY_HUMANS) return fail(e, ARBO_E_INVALID, "humans per community out of range");
            nh += s.N[c];
        }
        s.climate = false;
        double peak = 0.0;   // largest M(t)/N under a climate
        if (sh.climate) {
            if (const char* why = clim_invalid(*sh.climate, max_days)) return fail(e, ARBO_E_INVALID, why);
            s.NH = nh;
            init_climate(s, *sh.climate, max_days);
            for (float m : clim_target(*sh.climate, (double)nh, max_days)) {
                if (!std::isfinite(m) || !(m >= 0.0f)) return fail(e, ARBO_E_INVALID, "climate gives an invalid mosquito target");
                peak = std::max(peak, (double)m / (double)nh);
            }
        }
        // mosquito shares (mHeteroPopSize sizes, network.cpp:138-159), else by humans
        if (sh.mosq_share) {
            double sum = 0.0;
            for (uint32_t c = 0; c < s.C; ++c) {
                if (!(sh.mosq_share[c] > 0.0f) || !std::isfinite(sh.mosq_share[c]))
                    return fail(e, ARBO_E_INVALID, "mosquito shares must be positive");
                sum += (double)sh.mosq_share[c];
            }
            if (!(std::fabs(sum - 1.0) <= 1e-3)) return fail(e, ARBO_E_INVALID, "mosquito shares must sum to 1");
        }
        const double Ntot = (double)nh;
        nh = 0;
        for (uint32_t c = 0; c < s.C; ++c) {
            const double hs = sh.mosq_share ? Ntot * (double)sh.mosq_share[c] : (double)s.N[c];
            s.cap[c] = sh.mosq_slots ? sh.mosq_slots[c]
                       : s.climate ? (uint32_t)(std::ceil(peak * hs) + 1.0)
                                   : (uint32_t)std::ceil((double)s.p.m_ratio * (1.0 + (double)s.p.season_amp) * hs);
            if (s.cap[c] > ARBO_MAX_COMMUNITY_SLOTS) return fail(e, ARBO_E_INVALID, "mosquito slots per community out of range");
            if (!sh.mosq_share) s.base[c] = s.climate ? (float)s.N[c] : s.p.m_ratio * (float)s.N[c];
            else s.base[c] = s.climate ? (float)hs : (float)((double)s.p.m_ratio * hs);
            s.h_off[c] = nh; s.m_off[c] = nm;
            nh += s.N[c]; nm += s.cap[c];
        }
        if (nh >= (1ull << 32) || nm >= (1ull << 32)) return fail(e, ARBO_E_INVALID, "simulation too large");
        s.NH = nh; s.NM = nm;
    }
    const int nt = thread_count(e);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int i = 0; i < (int)n_sims; ++i) init_sim(e->sims[i], max_days);
    if (e->flags & ARBO_F_SPATIAL)
        for (Sim& s : e->sims) s.spatial.assign(((size_t)max_days + 1) * s.S * s.C, 0u);
    if (e->flags & (ARBO_F_HISTORY | ARBO_F_AOI))
        for (Sim& s : e->sims) s.hist.assign((size_t)s.NH * 4, -1);
    if (e->flags & ARBO_F_AOI)
        for (Sim& s : e->sims) {
            s.aoi_mark.assign(s.NH, 0);
            s.aoi.assign((size_t)4 * AOI_RECORD, 0u);
            std::fill(s.aoi_pos, s.aoi_pos + 4, 0u);
            s.aoi_from = max_days > AOI_WINDOW ? max_days - AOI_WINDOW : 0u;
        }
    if ((int)n_sims >= 2 * nt) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
        for (int i = 0; i < (int)n_sims; ++i) {
            Sim& s = e->sims[i];
            for (uint32_t c = 0; c < s.C; ++c) init_community(s, c);
            record_row(s, 0);
        }
    } else {
        for (Sim& s : e->sims) {
            const int C = (int)s.C;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
            for (int c = 0; c < C; ++c) init_community(s, (uint32_t)c);
            record_row(s, 0);
        }
    }
    e->day = 0;
    e->max_days = max_days;
    e->scan_h = e->scan_m = e->full_h = e->full_m = 0;
    for (const Sim& s : e->sims)
        for (uint32_t c = 0; c < s.C; ++c) {
            e->full_h += (s.N[c] + 3) / 4;
            e->full_m += (s.cap[c] + 7) / 8;
        }
    e->ready = true;
    return ARBO_OK;
}

int arbo_batch_step(ArboEngine* e, uint32_t n_days) {
    if (!e) return ARBO_E_INVALID;
    if (!e->ready) return fail(e, ARBO_E_STATE, "arbo_batch_init has not been called");
    if ((uint64_t)e->day + n_days > e->max_days) return fail(e, ARBO_E_CAPACITY, "series capacity exceeded");
    for (uint32_t k = 0; k < n_days; ++k) {
        day_step(e, e->day);
        e->day++;
        e->scan_h += e->full_h;
        e->scan_m += e->full_m;
    }
    return ARBO_OK;
}

int arbo_sync(ArboEngine* e) { return e ? ARBO_OK : ARBO_E_INVALID; }

int arbo_days_done(const ArboEngine* e, uint32_t* out) {
    if (!e || !out) return ARBO_E_INVALID;
    *out = e->day;
    return ARBO_OK;
}

int arbo_sim_sizes(const ArboEngine* e, uint32_t sim, uint32_t* communities, uint64_t* humans,
                   uint64_t* slots, uint32_t* serotypes) {
    if (!e || !e->ready || sim >= e->sims.size()) return ARBO_E_INVALID;
    const Sim& s = e->sims[sim];
    if (communities) *communities = s.C;
    if (humans) *humans = s.NH;
    if (slots) *slots = s.NM;
    if (serotypes) *serotypes = s.S;
    return ARBO_OK;
}

int arbo_read_series(ArboEngine* e, uint32_t sim, ArboSeries* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->sims.size()) return fail(e, ARBO_E_INVALID, "bad simulation index");
    const Sim& s = e->sims[sim];
    const size_t rows = e->day + 1;
    if (out->n_count) std::memcpy(out->n_count, s.n_count.data(), rows * 4);
    if (out->m_count) std::memcpy(out->m_count, s.m_count.data(), rows * 4);
    if (out->n_infected) std::memcpy(out->n_infected, s.n_inf.data(), rows * s.S * 4);
    if (out->n_incidence) std::memcpy(out->n_incidence, s.n_inc.data(), rows * s.S * 4);
    if (out->n_one_subpop) std::memcpy(out->n_one_subpop, s.n_one.data(), rows * s.S * 4);
    return ARBO_OK;
}

int arbo_read_spatial(ArboEngine* e, uint32_t sim, uint32_t* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->sims.size()) return fail(e, ARBO_E_INVALID, "bad simulation index");
    if (!(e->flags & ARBO_F_SPATIAL)) return fail(e, ARBO_E_STATE, "engine created without ARBO_F_SPATIAL");
    const Sim& s = e->sims[sim];
    std::memcpy(out, s.spatial.data(), ((size_t)e->day + 1) * s.S * s.C * 4);
    return ARBO_OK;
}

int arbo_read_history(ArboEngine* e, uint32_t sim, int32_t* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->sims.size()) return fail(e, ARBO_E_INVALID, "bad simulation index");
    if (!(e->flags & (ARBO_F_HISTORY | ARBO_F_AOI)))
        return fail(e, ARBO_E_STATE, "engine created without ARBO_F_HISTORY");
    const Sim& s = e->sims[sim];
    for (uint64_t h = 0; h < s.NH; ++h)
        for (uint32_t v = 0; v < s.S; ++v) out[v * s.NH + h] = s.hist[h * 4 + v];
    return ARBO_OK;
}

int arbo_climate_day(const ArboClimate* c, uint32_t row, double* out) {
    if (!c || !out || !c->temperature || !c->humidity || !c->precipitation || row >= c->days) return ARBO_E_INVALID;
    clim_day(*c, row, out);
    return ARBO_OK;
}

int arbo_read_aoi(ArboEngine* e, uint32_t sim, uint32_t* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->sims.size()) return fail(e, ARBO_E_INVALID, "bad simulation index");
    if (!(e->flags & ARBO_F_AOI)) return fail(e, ARBO_E_STATE, "engine created without ARBO_F_AOI");
    std::memcpy(out, e->sims[sim].aoi.data(), (size_t)4 * AOI_RECORD * 4);
    return ARBO_OK;
}

int arbo_read_state(ArboEngine* e, uint32_t sim, ArboState* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->sims.size()) return fail(e, ARBO_E_INVALID, "bad simulation index");
    const Sim& s = e->sims[sim];
    if (out->h_word) std::memcpy(out->h_word, s.hw.data(), s.NH * 4);
    if (out->h_next) std::memcpy(out->h_next, s.hn.data(), s.NH * 2);
    if (out->h_birth) std::memcpy(out->h_birth, s.hb.data(), s.NH * 4);
    if (out->m_word) std::memcpy(out->m_word, s.mw.data(), s.NM * 2);
    if (out->m_next) std::memcpy(out->m_next, s.mn.data(), s.NM * 2);
    return ARBO_OK;
}

int arbo_write_state(ArboEngine* e, uint32_t sim, const ArboState* in) {
    if (!e || !in) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->sims.size()) return fail(e, ARBO_E_INVALID, "bad simulation index");
    if (!in->h_word || !in->h_next || !in->h_birth || !in->m_word || !in->m_next)
        return fail(e, ARBO_E_INVALID, "every census array is required");
    Sim& s = e->sims[sim];
    std::memcpy(s.hw.data(), in->h_word, s.NH * 4);
    std::memcpy(s.hn.data(), in->h_next, s.NH * 2);
    std::memcpy(s.hb.data(), in->h_birth, s.NH * 4);
    std::memcpy(s.mw.data(), in->m_word, s.NM * 2);
    std::memcpy(s.mn.data(), in->m_next, s.NM * 2);
    return ARBO_OK;
}

int arbo_set_day(ArboEngine* e, uint32_t day) {
    if (!e) return ARBO_E_INVALID;
    if (!e->ready) return fail(e, ARBO_E_STATE, "arbo_batch_init has not been called");
    if (day > e->max_days) return fail(e, ARBO_E_CAPACITY, "day beyond series capacity");
    e->day = day;
    return ARBO_OK;
}

int arbo_kernel_time(ArboEngine* e, uint32_t, double* ms, uint64_t* launches) {
    if (!e) return ARBO_E_INVALID;
    if (ms) *ms = 0.0;
    if (launches) *launches = 0;
    return ARBO_OK;
}

int arbo_reset_timing(ArboEngine* e) {
    if (!e) return ARBO_E_INVALID;
    e->scan_h = e->scan_m = 0;
    return ARBO_OK;
}

// The oracle scans every member every day: the full census per day.
int arbo_scan_counts(ArboEngine* e, uint64_t* human_vectors, uint64_t* mosquito_vectors) {
    if (!e) return ARBO_E_INVALID;
    if (!e->ready) return fail(e, ARBO_E_STATE, "arbo_batch_init has not been called");
    if (human_vectors) *human_vectors = e->scan_h;
    if (mosquito_vectors) *mosquito_vectors = e->scan_m;
    return ARBO_OK;
}

// Launch tunings: the same candidate set as the GPU library, so a caller's
// tuning calls run unchanged; the CPU engine's results never depend on them.
static const uint32_t kShapes[4][2] = {{4, 4}, {4, 2}, {8, 4}, {2, 2}};
static const uint32_t kOcc8Shapes[2][2] = {{4, 2}, {2, 2}};   // also offered with 8 waves per SIMD
static const uint32_t kWideShapes[2][3] = {{4, 2, 2}, {2, 2, 4}};   // several waves per community

int arbo_tuning_candidates(ArboTuning* out, uint32_t cap, uint32_t* n) {
    if (!n) return ARBO_E_INVALID;
    uint32_t k = 0;
    for (const auto& d : kShapes)
        for (uint32_t x = 0; x <= 3; x += 3) {
            if (out && k < cap) out[k] = ArboTuning{d[0], d[1], x, 8, 0, 1};
            ++k;
        }
    for (const auto& d : kOcc8Shapes) {
        if (out && k < cap) out[k] = ArboTuning{d[0], d[1], 0, 8, 8, 1};
        ++k;
    }
    for (const auto& d : kWideShapes) {
        if (out && k < cap) out[k] = ArboTuning{d[0], d[1], 0, 8, 0, d[2]};
        ++k;
    }
    *n = k;
    return ARBO_OK;
}

int arbo_set_tuning(ArboEngine* e, const ArboTuning* t) {
    if (!e || !t) return ARBO_E_INVALID;
    bool shape = false;
    if (t->waves_per_comm > 1) {
        for (const auto& d : kWideShapes)
            shape = shape || (d[0] == t->census_unroll && d[1] == t->early_mosq && d[2] == t->waves_per_comm &&
                              t->waves_per_simd == 0);
    } else if (t->waves_per_comm != 1) {
        shape = false;
    } else if (t->waves_per_simd == 8) {
        for (const auto& d : kOcc8Shapes) shape = shape || (d[0] == t->census_unroll && d[1] == t->early_mosq);
    } else if (t->waves_per_simd == 0) {
        for (const auto& d : kShapes) shape = shape || (d[0] == t->census_unroll && d[1] == t->early_mosq);
    }
    if (!shape || t->xcd_order > 3 || t->ld_split < 1 || t->ld_split > 64) return fail(e, ARBO_E_INVALID, "unsupported tuning");
    e->tuning = *t;
    return ARBO_OK;
}

int arbo_get_tuning(const ArboEngine* e, ArboTuning* out) {
    if (!e || !out) return ARBO_E_INVALID;
    *out = e->tuning;
    return ARBO_OK;
}

int arbo_autotune(ArboEngine* e, uint32_t days_each, ArboTuning* chosen, double* ms_per_day) {
    if (!e || days_each == 0) return ARBO_E_INVALID;
    if (!e->ready) return fail(e, ARBO_E_STATE, "arbo_batch_init has not been called");
    uint32_t n = 0;
    arbo_tuning_candidates(nullptr, 0, &n);
    if ((uint64_t)e->day + (uint64_t)n * days_each > e->max_days)
        return fail(e, ARBO_E_CAPACITY, "autotuning needs candidates x days_each days of series capacity");
    const double t0 = omp_get_wtime();
    const int rc = arbo_batch_step(e, n * days_each);
    if (rc) return rc;
    if (chosen) *chosen = e->tuning;
    if (ms_per_day) *ms_per_day = (omp_get_wtime() - t0) * 1e3 / (n * days_each);
    return ARBO_OK;
}

void arbo_destroy(ArboEngine* e) { delete e; }

const char* arbo_last_error(const ArboEngine* e) { return e ? e->err.c_str() : "null engine"; }

// Oracle-only inspection hooks for the tests.
int oracle_tables(ArboEngine* e, uint32_t sim, uint32_t which, uint32_t* out, uint32_t cap, uint32_t* n) {
    if (!e || !e->ready || sim >= e->sims.size() || !n) return ARBO_E_INVALID;
    const Sim& s = e->sims[sim];
    const std::vector<uint32_t>* v = nullptr;
    switch (which) {
        case 0: v = &s.Kh; break;
        case 1: v = &s.Km; break;
        case 2: v = &s.Thy; break;
        case 3: v = &s.Tmd; break;
        case 4: v = &s.row_ptr; break;
        case 5: v = &s.to; break;
        case 6: v = &s.thr; break;
        case 7: v = &s.alias; break;
        case 8: v = &s.in_ptr; break;
        case 9: v = &s.in_src; break;
        case 10: v = &s.in_slot; break;
        case 11: v = &s.imm_thr; break;
        case 12: v = &s.cap; break;
        case 13: v = &s.cnt; break;
        default: return ARBO_E_INVALID;
    }
    *n = (uint32_t)v->size();
    if (out) std::memcpy(out, v->data(), std::min<size_t>(cap, v->size()) * 4);
    return ARBO_OK;
}

int oracle_dens(ArboEngine* e, uint32_t sim, float* out, uint32_t cap) {
    if (!e || !e->ready || sim >= e->sims.size()) return ARBO_E_INVALID;
    const Sim& s = e->sims[sim];
    std::memcpy(out, s.dens.data(), std::min<size_t>(cap, s.dens.size()) * 4);
    return ARBO_OK;
}

// Primitive hooks, so the tests can pin the oracle's own building blocks.
void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    W4 r = philox10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    std::memcpy(out, r.v, 16);
}
double oracle_exp(double x) { return exp_d(x); }
double oracle_log(double x) { return log_d(x); }
double oracle_sin(double x) { return sin_d(x); }
double oracle_logfact(double k) { return logfact_d(k); }
uint32_t oracle_poisson(double mu, uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2) {
    Seed k{(uint32_t)seed, (uint32_t)(seed >> 32)};
    return poisson_d(mu, k, c0, c1, c2);
}
uint32_t oracle_sround(float x, uint32_t r) { return sround(x, r); }
uint32_t oracle_source_counts(uint32_t law, float lam, float omega, float nd, uint64_t seed, uint32_t c,
                              uint32_t t, uint32_t dir, uint32_t v, uint64_t* fx, uint32_t* stay) {
    Seed k{(uint32_t)seed, (uint32_t)(seed >> 32)};
    return source_counts(law, lam, omega, nd, k, c, t, dir, v, *fx, *stay);
}

}  // extern "C"
