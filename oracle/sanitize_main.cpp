// sanitize_main.cpp — a driver for sanitizer builds of the oracle (test
// infrastructure; built only by `make asan` / `make tsan`).
//
// It runs representative configurations through the oracle's C ABI: every
// model switch the parity cases use (both count laws, a staying fraction,
// von Neumann and Moore lattices with and without self entries, 1-4
// serotypes, Beta(2,5) immunity, introductions, a climate, mosquito shares,
// ragged communities up to the 65,535 maximum and one-human ones), the kept
// outputs (history, age-of-infection rings, per-community series), reads of
// every output, a checkpoint written back at a later day, and a batch of
// several simulations for the OpenMP paths.  Any sanitizer report ends the
// process with a non-zero status.
//
//   oracle_asan [threads]      oracle_tsan [threads]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "arbo_c.h"

namespace {

int g_fail = 0;

void check(int rc, ArboEngine* e, const char* what) {
    if (rc != ARBO_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, e ? arbo_last_error(e) : "");
        g_fail = 1;
        std::exit(2);
    }
}

ArboParams base_params() {
    ArboParams p;
    std::memset(&p, 0, sizeof p);
    p.n_serotypes = 4;
    p.beta_hm = 0.3f;
    p.beta_mh = 0.3f;
    p.long_distance = 0.01f;
    p.neighbourhood = 8;
    p.include_self = 1;
    p.dispersal_scale = 1.0f;
    p.iip_min = 4;
    p.iip_max = 7;
    p.recovery = 5.0f;
    p.eip = 10.0f;
    p.h_infant_shape = 0.4f;
    p.h_scale = (float)(70.0 * 365.0 / std::tgamma(1.0 + 1.0 / 6.0));
    p.h_shape = 6.0f;
    p.m_scale = 10.0f;
    p.m_shape = 1.0f;
    p.m_ratio = 2.0f;
    p.season_amp = 0.5f;
    p.seed_frac_h = 0.01f;
    p.seed_frac_m = 0.01f;
    p.count_law = ARBO_COUNT_POISSON;
    return p;
}

// The reference's own defaults (SimulationParam.cpp:11-47) with seeding.
ArboParams reference_params() {
    ArboParams p = base_params();
    p.n_serotypes = 1;
    p.beta_hm = p.beta_mh = 0.4f;
    p.external_infection = 5.0f;
    p.neighbourhood = 4;
    p.include_self = 0;
    p.dispersal_scale = 0.0f;
    p.iip_min = p.iip_max = 7;
    p.recovery = 4.0f;
    p.eip = 6.0f;
    p.h_infant_scale = 1.0f / 100000.0f / 365.0f;
    p.h_loc = 8.0f * 365.0f;
    p.h_scale = 75.0f * 365.0f;
    p.m_scale = 23.0f;
    p.m_shape = 4.0f;
    p.m_ratio = 1.5f;
    p.season_amp = 0.0f;
    p.seed_frac_h = p.seed_frac_m = 0.02f;
    p.no_distance = 0.75f;
    p.count_law = ARBO_COUNT_STOCH_ROUND;
    return p;
}

struct Sim {
    ArboShape shape;
    ArboParams params;
    uint64_t seed;
    std::vector<uint32_t> humans, slots;
    std::vector<float> share;
};

Sim make_sim(uint32_t w, uint32_t h, uint32_t n, const ArboParams& p, uint64_t seed) {
    Sim s;
    s.humans.assign((size_t)w * h, n);
    s.params = p;
    s.seed = seed;
    std::memset(&s.shape, 0, sizeof s.shape);
    s.shape.grid_w = w;
    s.shape.grid_h = h;
    return s;
}

// Runs a batch for `days`, reading every output on the way, then writes
// simulation 0's census back at a later day and runs on.
void run(const char* name, std::vector<Sim>& sims, uint32_t days, uint32_t flags, int threads) {
    std::fprintf(stderr, "[sanitize] %s: %zu simulations, %u days\n", name, sims.size(), days);
    ArboEngineDesc d{ARBO_ABI_VERSION, 0, flags, (uint32_t)threads};
    ArboEngine* e = nullptr;
    check(arbo_create(&d, &e), nullptr, "arbo_create");
    std::vector<ArboShape> shapes;
    std::vector<ArboParams> params;
    std::vector<uint64_t> seeds;
    for (Sim& s : sims) {
        s.shape.humans = s.humans.data();
        s.shape.mosq_slots = s.slots.empty() ? nullptr : s.slots.data();
        s.shape.mosq_share = s.share.empty() ? nullptr : s.share.data();
        shapes.push_back(s.shape);
        params.push_back(s.params);
        seeds.push_back(s.seed);
    }
    const uint32_t T = days + 10;
    check(arbo_batch_init(e, (uint32_t)sims.size(), shapes.data(), params.data(), seeds.data(), T), e, "batch_init");
    check(arbo_batch_step(e, days / 2), e, "step");
    check(arbo_batch_step(e, days - days / 2), e, "step");
    for (uint32_t i = 0; i < sims.size(); ++i) {
        uint32_t C = 0, S = 0;
        uint64_t N = 0, M = 0;
        check(arbo_sim_sizes(e, i, &C, &N, &M, &S), e, "sim_sizes");
        const size_t rows = days + 1;
        std::vector<uint32_t> nc(rows), mc(rows), inf(rows * S), inc(rows * S), one(rows * S);
        ArboSeries ser{nc.data(), mc.data(), inf.data(), inc.data(), one.data()};
        check(arbo_read_series(e, i, &ser), e, "read_series");
        std::vector<uint32_t> hw(N);
        std::vector<uint16_t> hn(N), mw(M), mn(M);
        std::vector<int32_t> hb(N);
        ArboState st{hw.data(), hn.data(), hb.data(), mw.data(), mn.data()};
        check(arbo_read_state(e, i, &st), e, "read_state");
        if (flags & ARBO_F_SPATIAL) {
            std::vector<uint32_t> sp(rows * S * C);
            check(arbo_read_spatial(e, i, sp.data()), e, "read_spatial");
        }
        if (flags & (ARBO_F_HISTORY | ARBO_F_AOI)) {
            std::vector<int32_t> hist((size_t)S * N);
            check(arbo_read_history(e, i, hist.data()), e, "read_history");
        }
        if (flags & ARBO_F_AOI) {
            std::vector<uint32_t> aoi(4 * ARBO_AOI_RECORD);
            check(arbo_read_aoi(e, i, aoi.data()), e, "read_aoi");
        }
        if (i == 0) {   // checkpoint written back at a later day, then resumed
            check(arbo_set_day(e, days + 3), e, "set_day");
            check(arbo_write_state(e, 0, &st), e, "write_state");
            check(arbo_batch_step(e, 5), e, "step after resume");
            check(arbo_set_day(e, days), e, "set_day back");
        }
    }
    arbo_destroy(e);
}

}  // namespace

int main(int argc, char** argv) {
    const int threads = argc > 1 ? std::atoi(argv[1]) : 1;
    {   // the configured model on a Moore lattice, every output kept
        std::vector<Sim> v{make_sim(6, 5, 300, base_params(), 11)};
        run("lattice, outputs", v, 60, ARBO_F_SPATIAL | ARBO_F_AOI, threads);
    }
    {   // the reference's parameter set: stochastic rounding, staying fraction, introductions
        std::vector<Sim> v{make_sim(4, 4, 200, reference_params(), 12)};
        run("reference parameters", v, 80, ARBO_F_HISTORY, threads);
    }
    {   // ragged communities: 1 human, the maxima of 65,535 humans and slots, no mosquito slots
        ArboParams p = base_params();
        p.n_serotypes = 3;
        p.long_distance = 0.2f;
        Sim s = make_sim(5, 3, 1, p, 13);
        const uint32_t hum[15] = {1, 7, 8192, 13, 64, 65, 255, 256, 257, 1000, 2, 3, 65535, 5, 999};
        const uint32_t sl[15] = {0, 16, 8192, 30, 100, 200, 511, 512, 513, 3000, 1, 8, 9, 65535, 2500};
        s.humans.assign(hum, hum + 15);
        s.slots.assign(sl, sl + 15);
        std::vector<Sim> v{s};
        run("ragged", v, 40, 0, threads);
    }
    {   // Beta(2,5) immunity, Poisson counts with a staying fraction, von Neumann, shares
        ArboParams a = base_params();
        a.immunity_mode = 1;
        a.no_distance = 0.3f;
        a.include_self = 0;
        ArboParams b = base_params();
        b.no_distance = 0.6f;
        b.neighbourhood = 4;
        b.dispersal_scale = 0.0f;
        b.count_law = ARBO_COUNT_STOCH_ROUND;
        b.n_serotypes = 2;
        Sim sb = make_sim(4, 6, 400, b, 18);
        sb.share.assign(24, 1.0f / 24.0f);
        sb.share[0] = 2.0f / 24.0f;
        sb.share[1] = 0.0f / 24.0f + 1e-6f;
        float sum = 0.0f;
        for (float x : sb.share) sum += x;
        for (float& x : sb.share) x /= sum;
        std::vector<Sim> v{make_sim(5, 5, 600, a, 17), sb, make_sim(3, 3, 2000, base_params(), 15)};
        run("mixed batch", v, 50, 0, threads);
    }
    {   // a climate: per-day transmission, EIP, mosquito lifetimes, logistic target
        ArboParams p = reference_params();
        Sim s = make_sim(5, 5, 300, p, 31);
        const uint32_t days = 60;
        static std::vector<float> T(days + 11), H(days + 11), P(days + 11);
        for (size_t t = 0; t < T.size(); ++t) {
            const double x = std::sin(6.283185307179586 * (double)t / 365.0);
            T[t] = (float)(26.0 + 4.0 * x);
            H[t] = (float)(0.7 + 0.1 * x);
            P[t] = (float)(0.4 + 0.4 * x);
        }
        static ArboClimate cl{T.data(), H.data(), P.data(), (uint32_t)T.size(), 0.5f, 2.5f, 3.0f, 1.0f, 1.0f, 3.0f,
                              0.5f};
        s.shape.climate = &cl;
        std::vector<Sim> v{s};
        run("climate", v, days, 0, threads);
    }
    std::fprintf(stderr, "[sanitize] all cases ran clean\n");
    return g_fail;
}
