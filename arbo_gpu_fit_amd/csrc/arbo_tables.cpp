// arbo_tables.cpp — host tables of one simulation (see arbo_tables.h).
//
// Lifetimes: the reference draws PLE ~ U(0,1] at birth and kills on the first
// day with F(age) >= PLE (ndemographic.cu:62, mdemographic.cu:72).  Here F is
// tabulated once per parameter set as 24-bit thresholds K[a] = floor(F(a) 2^24)
// and the kernels search K for a 24-bit PLE: an exact integer version of the
// same law.  Initial ages follow the cumulative survival sums of
// demographicintial.cu:48-83.  The network is the reference's non-wrapping
// lattice (network.cpp:31-41) with rows in ascending column order
// (network.cpp:95-113), generalised to Moore neighbourhoods, a self entry and
// exp(-d/scale) weights; destinations are drawn from Vose alias tables.
// All floating point is the deterministic double arithmetic of arbo_rng.h.
#include "arbo_tables.h"

#include <algorithm>
#include <cmath>

#include "arbo_rng.h"

namespace arbo {

uint32_t fixed24(double f) {
    const double v = floor(f * 16777216.0);
    if (!(v > 0.0)) return 0u;
    if (v >= 16777216.0) return kOne24;
    return (uint32_t)v;
}

static double powd(double x, double y) { return x > 0.0 ? dexp(y * dlog(x)) : 0.0; }

// ndemographic.cu:13-29: F(a) = 1 - exp(-[(a b)^alpha + 1{a >= L} ((a-L)/d)^c]).
static double human_cdf(const ArboParams& p, double a) {
    double haz = powd(a * (double)p.h_infant_scale, (double)p.h_infant_shape);
    if (a >= (double)p.h_loc) haz = haz + powd((a - (double)p.h_loc) / (double)p.h_scale, (double)p.h_shape);
    return 1.0 - dexp(-haz);
}

// mdemographic.cu:13-17: F(a) = 1 - exp(-(a/scale)^shape).
static double weibull_cdf(double a, double scale, double shape) { return 1.0 - dexp(-powd(a / scale, shape)); }

static double mosq_cdf(const ArboParams& p, double a) { return weibull_cdf(a, (double)p.m_scale, (double)p.m_shape); }

// ---- climate (SimulationParam.cpp:92-137, simulation.cpp:87,129-131) ----

// c T (T - lo) sqrt(hi - T) inside [lo, hi], else 0 (SimulationParam.cpp:114-136).
static double briere(double T, double c, double lo, double hi) {
    if (T < lo || T > hi) return 0.0;
    return c * T * (T - lo) * std::sqrt(hi - T);
}

void climate_day(const ArboClimate& c, uint32_t row, double out[4]) {
    const double T = (double)c.temperature[row], H = (double)c.humidity[row];
    const double TK = T + 273.15;
    const double rate = 24.0 * 0.003359 * (TK / 298.0) * dexp((15000.0 / 1.986) * ((1.0 / 298.0) - (1.0 / TK)));
    const double mu = 0.8692 - 0.159 * T + 0.01116 * (T * T) - 0.0003408 * (T * T * T) + 0.000003809 * (T * T * T * T);
    out[0] = (double)c.biting_rate * briere(T, 0.000491, 12.22, 37.46);   // nmBitingSuccess
    out[1] = (double)c.biting_rate * briere(T, 0.000849, 17.05, 35.83);   // mnBitingSuccess
    out[2] = 1.0 / (rate * (double)c.scale_eip);
    out[3] = 1.0 / (mu * (double)c.scale_mle * powd(1.0 - H, (double)c.scale_humidity));
}

// Carrying capacity of climate row `row` (simulation.cpp:87,131):
// minMosToHuman N (1 + P)^scaleCC.
static double climate_capacity(const ArboClimate& c, uint32_t row, double N) {
    return (double)c.min_mos_to_human * N * powd(1.0 + (double)c.precipitation[row], (double)c.scale_cc);
}

// The expected mosquitoes of the whole simulation, M(0..max_days), kept in
// float like h_mExpectedPopSize (simulation.cpp:84-132).
static std::vector<float> climate_target(const ArboClimate& c, double N, uint32_t max_days) {
    std::vector<float> M((size_t)max_days + 1);
    M[0] = (float)climate_capacity(c, 0, N);
    for (uint32_t t = 0; t < max_days; ++t) {
        const double m = (double)M[t];
        M[t + 1] = (float)(m + (double)c.mos_net_growth * m * (1.0 - m / climate_capacity(c, t + 1, N)));
    }
    return M;
}

static std::string validate_climate(const ArboClimate& c, uint32_t max_days) {
    if (!c.temperature || !c.humidity || !c.precipitation) return "climate arrays missing";
    if (c.days < max_days + 1) return "climate shorter than max_days + 1 days";
    if (!(c.biting_rate >= 0.0f) || !(c.scale_eip > 0.0f) || !(c.scale_mle > 0.0f) || !(c.min_mos_to_human > 0.0f) ||
        !(c.mos_net_growth >= 0.0f) || !std::isfinite(c.scale_cc) || !std::isfinite(c.scale_humidity))
        return "bad climate coefficients";
    for (uint32_t r = 0; r <= max_days; ++r) {
        if (!std::isfinite(c.temperature[r]) || !(c.humidity[r] >= 0.0f && c.humidity[r] < 1.0f) ||
            !(c.precipitation[r] > -1.0f) || !std::isfinite(c.precipitation[r]))
            return "climate row " + std::to_string(r) + " out of range";
        double d[4];
        climate_day(c, r, d);
        if (!(d[2] > 0.0 && d[2] < 1000.0)) return "climate row " + std::to_string(r) + " gives eip outside (0, 1000) days";
        if (!(d[3] > 0.0 && d[3] < 1.0e4))
            return "climate row " + std::to_string(r) + " gives a mosquito lifespan scale outside (0, 1e4) days";
    }
    return std::string();
}

template <class Cdf>
static void threshold_table(uint32_t n, Cdf cdf, std::vector<uint32_t>& K) {
    K.resize(n);
    uint32_t prev = 0;
    for (uint32_t a = 0; a < n; ++a) {
        uint32_t k = fixed24(cdf((double)a));
        if (k < prev) k = prev;
        K[a] = k;
        prev = k;
    }
}

template <class Cdf>
static void initial_age_table(uint32_t n, uint32_t step, Cdf cdf, std::vector<uint32_t>& T) {
    std::vector<double> cum(n);
    double acc = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        acc = acc + (1.0 - cdf((double)(i * step)));
        cum[i] = acc;
    }
    T.resize(n);
    for (uint32_t i = 0; i < n; ++i) T[i] = fixed24(cum[i] / cum[n - 1]);
    T[n - 1] = kOne24;
    for (uint32_t i = 1; i < n; ++i) T[i] = std::max(T[i], T[i - 1]);
}

std::string validate(const ArboShape& sh, const ArboParams& p, uint32_t max_days) {
    if (p.n_serotypes < 1 || p.n_serotypes > ARBO_MAX_SEROTYPES) return "n_serotypes must be 1..4";
    if (sh.grid_w < 1 || sh.grid_h < 1 || !sh.humans) return "empty lattice";
    if ((uint64_t)sh.grid_w * sh.grid_h > (1u << 24)) return "too many communities";
    if (p.neighbourhood != 4 && p.neighbourhood != 8) return "neighbourhood must be 4 or 8";
    if (!(p.beta_hm >= 0.0f) || !(p.beta_mh >= 0.0f)) return "beta must be >= 0";
    if (!(p.long_distance >= 0.0f && p.long_distance <= 1.0f)) return "long_distance must be in [0,1]";
    if (!(p.season_amp >= 0.0f && p.season_amp <= 1.0f)) return "season_amp must be in [0,1]";
    if (!(p.recovery >= 0.0f && p.recovery < 1000.0f) || !(p.eip >= 0.0f && p.eip < 1000.0f)) return "recovery/eip out of range";
    if (p.iip_max < p.iip_min || p.iip_max > 1000) return "bad incubation range";
    if (!(p.m_scale > 0.0f) || !(p.m_shape > 0.0f) || !(p.h_scale > 0.0f) || !(p.h_shape > 0.0f)) return "bad lifespan parameters";
    // Lifetimes are integer ages up to the tables' last entries (a human death
    // day fits 16 bits of age, a mosquito one 11), drawn by 24-bit uniforms:
    // the lifetimes past the table (which would take its last age) must be
    // below that resolution, F(last age) >= 1 - 2^-24 in 24-bit fixed point.
    if (fixed24(human_cdf(p, (double)(kHumanAgeTable - 1))) < kOne24 - 1)
        return "human lifespan beyond the 65,535-day lifetime table";
    if (fixed24(mosq_cdf(p, (double)(kMosqAgeTable - 1))) < kOne24 - 1)
        return "mosquito lifespan beyond the 2,047-day lifetime table";
    if (!(p.m_ratio >= 0.0f)) return "m_ratio must be >= 0";
    if (!(p.external_infection >= 0.0f)) return "external_infection must be >= 0";
    if (!(p.seed_frac_h >= 0.0f && p.seed_frac_h <= 1.0f) || !(p.seed_frac_m >= 0.0f && p.seed_frac_m <= 1.0f)) return "seed fractions must be in [0,1]";
    if (p.immunity_mode > 1 || !(p.sero_prev >= 0.0f && p.sero_prev <= 1.0f)) return "bad immunity settings";
    if (p.count_law > ARBO_COUNT_STOCH_ROUND) return "count_law must be ARBO_COUNT_POISSON or ARBO_COUNT_STOCH_ROUND";
    if (!(p.no_distance >= 0.0f && p.no_distance <= 1.0f)) return "no_distance must be in [0,1]";
    if (!(p.dispersal_scale >= 0.0f)) return "dispersal_scale must be >= 0";
    const uint64_t C = (uint64_t)sh.grid_w * sh.grid_h;
    uint64_t nh = 0, nm = 0;
    for (uint64_t c = 0; c < C; ++c) {
        const uint32_t n = sh.humans[c];
        if (n < 1 || n > ARBO_MAX_COMMUNITY_HUMANS) return "humans per community out of range";
        nh += n;
    }
    double clim_peak = 0.0;   // largest M(t) / N of a climate
    if (sh.climate) {
        const std::string why = validate_climate(*sh.climate, max_days);
        if (!why.empty()) return why;
        // Births on day t search the running maximum of F(a; scale(t + a)),
        // which is at least F(a; largest scale of the run): that must reach 1
        // within the table for every birth day's row to.
        double smax = 0.0;
        for (uint32_t r = 1; r <= max_days; ++r) {
            double d[4];
            climate_day(*sh.climate, r, d);
            smax = std::max(smax, d[3]);
        }
        if (fixed24(weibull_cdf((double)(kMosqAgeTable - 1), smax, (double)p.m_shape)) < kOne24 - 1)
            return "climate gives a mosquito lifespan beyond the 2,047-day lifetime table";
        for (float m : climate_target(*sh.climate, (double)nh, max_days)) {
            if (!std::isfinite(m) || !(m >= 0.0f)) return "climate gives an invalid mosquito target";
            clim_peak = std::max(clim_peak, (double)m / (double)nh);
        }
    }
    if (sh.mosq_share) {
        double sum = 0.0;
        for (uint64_t c = 0; c < C; ++c) {
            if (!(sh.mosq_share[c] > 0.0f) || !std::isfinite(sh.mosq_share[c])) return "mosquito shares must be positive";
            sum += (double)sh.mosq_share[c];
        }
        if (!(std::fabs(sum - 1.0) <= 1e-3)) return "mosquito shares must sum to 1";
    }
    const double N = (double)nh;
    nh = 0;
    for (uint64_t c = 0; c < C; ++c) {
        const uint32_t n = sh.humans[c];
        // humans-equivalent of the community's mosquito share
        const double hs = sh.mosq_share ? N * (double)sh.mosq_share[c] : (double)n;
        const double cap = sh.mosq_slots ? (double)sh.mosq_slots[c]
                         : sh.climate ? ceil(clim_peak * hs) + 1.0
                                      : ceil((double)p.m_ratio * (1.0 + (double)p.season_amp) * hs);
        if (cap > (double)ARBO_MAX_COMMUNITY_SLOTS) return "mosquito slots per community out of range";
        nh += n;
        nm += (uint64_t)cap;
    }
    if (nh >= (1ull << 32) || nm >= (1ull << 32)) return "simulation too large";
    return std::string();
}

static void build_network(const ArboParams& p, SimTables& t) {
    t.row_ptr.assign(1, 0u);
    t.to.clear();
    t.thr.clear();
    t.alias.clear();
    for (uint32_t y = 0; y < t.H; ++y) {
        for (uint32_t x = 0; x < t.W; ++x) {
            uint32_t col[kMaxRow];
            double w[kMaxRow];
            uint32_t n = 0;
            for (int dy = -1; dy <= 1; ++dy) {
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = (int)x + dx, yy = (int)y + dy;
                    if (xx < 0 || yy < 0 || xx >= (int)t.W || yy >= (int)t.H) continue;
                    const int m = (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy);
                    // the community itself: a dispersal target with include_self, and
                    // the column its staying events (no_distance) are counted in,
                    // with weight 0 when it is no dispersal target
                    if (m == 0 && !p.include_self && !(p.no_distance > 0.0f)) continue;
                    if (m == 2 && p.neighbourhood == 4) continue;
                    const double dist = (m == 0) ? 0.0 : ((m == 1) ? 1.0 : 1.4142135623730951);
                    col[n] = (uint32_t)yy * t.W + (uint32_t)xx;
                    w[n] = (p.dispersal_scale > 0.0f) ? dexp(-dist / (double)p.dispersal_scale) : 1.0;
                    if (m == 0 && !p.include_self) w[n] = 0.0;
                    ++n;
                }
            }
            if (n == 0 || (n == 1 && w[0] == 0.0)) {  // isolated community: local events stay at home
                col[0] = y * t.W + x;
                w[0] = 1.0;
                n = 1;
            }
            // Vose's alias method: scaled weights q = n w / sum(w); a stack of
            // under-full and of over-full columns, paired until one is empty.
            double sum = 0.0;
            for (uint32_t i = 0; i < n; ++i) sum = sum + w[i];
            double q[kMaxRow], prob[kMaxRow];
            uint32_t al[kMaxRow], small[kMaxRow], large[kMaxRow];
            uint32_t ns = 0, nl = 0;
            for (uint32_t i = 0; i < n; ++i) {
                q[i] = w[i] / sum * (double)n;
                prob[i] = 1.0;
                al[i] = i;
                if (q[i] < 1.0) small[ns++] = i; else large[nl++] = i;
            }
            while (ns > 0 && nl > 0) {
                const uint32_t l = small[--ns];
                const uint32_t g = large[--nl];
                prob[l] = q[l];
                al[l] = g;
                q[g] = (q[g] + q[l]) - 1.0;
                if (q[g] < 1.0) small[ns++] = g; else large[nl++] = g;
            }
            for (uint32_t i = 0; i < n; ++i) {
                t.to.push_back(col[i]);
                t.thr.push_back(fixed24(prob[i]));
                t.alias.push_back(al[i]);
            }
            t.row_ptr.push_back((uint32_t)t.to.size());
        }
    }
    // transpose: in-edges per destination, sources ascending (a source visits
    // each destination at most once, and sources are enumerated in order)
    std::vector<uint32_t> deg(t.C + 1, 0u);
    for (uint32_t e = 0; e < t.to.size(); ++e) deg[t.to[e] + 1]++;
    for (uint32_t c = 0; c < t.C; ++c) deg[c + 1] += deg[c];
    t.in_ptr = deg;
    t.in_src.assign(t.to.size(), 0u);
    t.in_slot.assign(t.to.size(), 0u);
    std::vector<uint32_t> fill(t.in_ptr.begin(), t.in_ptr.end() - 1);
    for (uint32_t c = 0; c < t.C; ++c) {
        for (uint32_t e = t.row_ptr[c]; e < t.row_ptr[c + 1]; ++e) {
            const uint32_t d = t.to[e];
            t.in_src[fill[d]] = c;
            t.in_slot[fill[d]] = e - t.row_ptr[c];
            fill[d]++;
        }
    }
}

// Mosquito lifetime rows and their guides.  Row 0 is the tabulated F of
// m_scale.  With a climate, a mosquito born on day b dies on the first day
// b + a whose F(a; scale(b + a)) reaches its PLE (mdemographic.cu:72 with the
// day's mScaleLifeExpectancy); the running maximum over a of those
// thresholds turns that first passage into a search of a monotone row.
// Days past the run use the last day's scale.  Every row ends at its first
// 2^24 entry, and all rows share the largest such length.
static void build_mosquito_rows(const ArboShape& sh, const ArboParams& p, uint32_t max_days, SimTables& t) {
    std::vector<std::vector<uint32_t>> rows;
    rows.emplace_back();
    threshold_table(kMosqAgeTable, [&](double a) { return mosq_cdf(p, a); }, rows[0]);
    if (t.climate) {
        std::vector<double> scale(max_days);
        for (uint32_t d = 0; d < max_days; ++d) {
            double x[4];
            climate_day(*sh.climate, d + 1, x);
            scale[d] = x[3];
        }
        for (uint32_t b = 0; b < max_days; ++b) {
            std::vector<uint32_t> K(1, 0u);
            uint32_t run = 0;
            for (uint32_t a = 1; a < kMosqAgeTable && run < kOne24; ++a) {
                const uint64_t day = std::min<uint64_t>((uint64_t)b + a, max_days - 1);
                run = std::max(run, fixed24(weibull_cdf((double)a, scale[day], (double)p.m_shape)));
                K.push_back(run);
            }
            rows.push_back(std::move(K));
        }
    }
    // the largest death age any row returns: its first 2^24 entry, else the table end
    uint32_t dmax = 1;
    for (const auto& K : rows) {
        uint32_t first = kMosqAgeTable - 1;
        for (uint32_t a = 1; a < K.size(); ++a)
            if (K[a] >= kOne24) { first = a; break; }
        dmax = std::max(dmax, first);
    }
    t.death_m_max = dmax;
    t.m_rows = (uint32_t)rows.size();
    t.death_m.assign((size_t)t.m_rows * (dmax + 1), kOne24);
    t.death_m_guide.assign((size_t)t.m_rows * kGuideBuckets, dmax);
    for (uint32_t r = 0; r < t.m_rows; ++r) {
        uint32_t* K = t.death_m.data() + (size_t)r * (dmax + 1);
        for (uint32_t a = 0; a <= dmax && a < rows[r].size(); ++a) K[a] = rows[r][a];
        K[dmax] = kOne24;
        // the smallest death age of every life_bucket() of 24-bit draws (padding: the largest)
        for (uint32_t b = 0; b < kLifeBuckets; ++b) {
            const uint32_t k = life_bucket_kmin(b);
            uint32_t a = 1;
            while (a < dmax && K[a] < k) ++a;
            t.death_m_guide[(size_t)r * kGuideBuckets + b] = a;
        }
    }
}

void build_tables(const ArboShape& sh, const ArboParams& p, uint64_t seed, uint32_t max_days,
                  SimTables& t) {
    t.W = sh.grid_w;
    t.H = sh.grid_h;
    t.C = t.W * t.H;
    t.S = p.n_serotypes;
    t.humans.assign(sh.humans, sh.humans + t.C);
    t.slots.resize(t.C);
    t.base.resize(t.C);
    t.n_humans = 0;
    t.n_slots = 0;
    for (uint32_t c = 0; c < t.C; ++c) t.n_humans += t.humans[c];
    t.climate = sh.climate != nullptr;
    double clim_peak = 0.0;
    if (t.climate) {   // the target M(t) / N, each community's share by its humans
        const ArboClimate& cl = *sh.climate;
        const std::vector<float> M = climate_target(cl, (double)t.n_humans, max_days);
        t.dens.resize(M.size());
        for (size_t d = 0; d < M.size(); ++d) {
            t.dens[d] = (float)((double)M[d] / (double)t.n_humans);
            clim_peak = std::max(clim_peak, (double)M[d] / (double)t.n_humans);
        }
        t.clim.resize((size_t)max_days * 3);
        for (uint32_t d = 0; d < max_days; ++d) {   // day d reads row d + 1 (simulation.cpp:122)
            double x[4];
            climate_day(cl, d + 1, x);
            t.clim[3 * (size_t)d] = (float)x[0];
            t.clim[3 * (size_t)d + 1] = (float)x[1];
            t.clim[3 * (size_t)d + 2] = (float)x[2];
        }
    }
    for (uint32_t c = 0; c < t.C; ++c) {
        // humans-equivalent of the community's mosquito share (its humans without shares)
        const double hs = sh.mosq_share ? (double)t.n_humans * (double)sh.mosq_share[c] : (double)t.humans[c];
        t.slots[c] = sh.mosq_slots ? sh.mosq_slots[c]
                     : t.climate ? (uint32_t)(ceil(clim_peak * hs) + 1.0)
                                 : (uint32_t)ceil((double)p.m_ratio * (1.0 + (double)p.season_amp) * hs);
        if (!sh.mosq_share) t.base[c] = t.climate ? (float)t.humans[c] : p.m_ratio * (float)t.humans[c];
        else t.base[c] = t.climate ? (float)hs : (float)((double)p.m_ratio * hs);
        t.n_slots += t.slots[c];
    }
    threshold_table(kHumanAgeTable, [&](double a) { return human_cdf(p, a); }, t.death_h);
    build_mosquito_rows(sh, p, max_days, t);
    initial_age_table(kHumanYears, 365, [&](double a) { return human_cdf(p, a); }, t.init_h);
    initial_age_table(kMosqInitDays, 1, [&](double a) { return mosq_cdf(p, a); }, t.init_m);
    if (!t.climate) {
        t.dens.resize((size_t)max_days + 1);
        for (uint32_t d = 0; d <= max_days; ++d) {
            const double arg = 6.283185307179586 * (double)d / 365.0 + (double)p.season_phase;
            t.dens[d] = (float)(1.0 + (double)p.season_amp * dsin(arg));
        }
    }
    const Key key{(uint32_t)seed, (uint32_t)(seed >> 32)};
    t.immune.assign((size_t)t.C * t.S, 0u);
    for (uint32_t c = 0; c < t.C; ++c) {
        for (uint32_t s = 0; s < t.S; ++s) {
            uint32_t thr;
            if (p.immunity_mode == 0) {
                thr = fixed24((double)p.sero_prev);
            } else {
                // Beta(2,5) as the 2nd order statistic of 6 uniforms
                const U4 a = draw(key, c, INIT_DAY, ctr_tag(P_IMMUNE, 0, s, 0), 0);
                const U4 b = draw(key, c, INIT_DAY, ctr_tag(P_IMMUNE, 0, s, 0), 1);
                uint32_t u[6] = {a.x >> 8, a.y >> 8, a.z >> 8, a.w >> 8, b.x >> 8, b.y >> 8};
                std::sort(u, u + 6);
                thr = u[1];
            }
            t.immune[(size_t)c * t.S + s] = thr;
        }
    }
    t.seed_h = fixed24((double)p.seed_frac_h);
    t.seed_m = fixed24((double)p.seed_frac_m);
    build_network(p, t);
}

}  // namespace arbo
