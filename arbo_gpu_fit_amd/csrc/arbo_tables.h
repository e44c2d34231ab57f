// arbo_tables.h — host-side construction of everything a simulation needs
// before day 0: lifespan and initial-age tables, the seasonal emergence
// targets, per-community immunity thresholds and the dispersal network (CSR
// rows with Vose alias tables, plus each destination's in-edges).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "arbo_c.h"

namespace arbo {

constexpr uint32_t kHumanAgeTable = 65536;  // death ages 1..65535
constexpr uint32_t kMosqAgeTable = 2048;    // death ages 1..2047 (12-bit death day)
constexpr uint32_t kHumanYears = 151;       // C_NMAXINITIALAGE + 1 (constant.h:19)
constexpr uint32_t kMosqInitDays = 301;     // C_MMAXINITIALAGE + 1 (constant.h:18)
constexpr uint32_t kOne24 = 1u << 24;
constexpr uint32_t kMaxRow = 9;             // Moore-8 + self
constexpr uint32_t kGuideBuckets = 256;     // life_bucket() range (176), padded

struct SimTables {
    uint32_t W = 0, H = 0, C = 0, S = 0;
    std::vector<uint32_t> humans, slots;   // per community
    std::vector<float> base;               // m_ratio * humans, f32
    std::vector<uint32_t> death_h;         // [kHumanAgeTable]  floor(F_h(a) 2^24)
    // mosquito lifetimes, m_rows rows of death_m_max + 1 thresholds: row 0
    // floor(F_m(a) 2^24) of m_scale (the t = 0 census; every birth without a
    // climate); with a climate, row 1 + b the running maximum over a' <= a of
    // floor(F_m(a'; scale(b + a')) 2^24) for births on day b
    std::vector<uint32_t> death_m;
    uint32_t death_m_max = 0;              // largest death age any row returns (its first 2^24 entry)
    uint32_t m_rows = 1;
    std::vector<uint32_t> death_m_guide;   // [m_rows][kGuideBuckets] death age of life_bucket_kmin(b)
    bool climate = false;
    std::vector<float> clim;               // with a climate: [max_days][3] beta_hm, beta_mh, eip of day t
    std::vector<uint32_t> init_h;          // [kHumanYears]     initial-age year thresholds
    std::vector<uint32_t> init_m;          // [kMosqInitDays]   initial-age day thresholds
    std::vector<float> dens;               // [max_days+1] 1 + A sin(2 pi t/365 + phi), or M(t) / N
    std::vector<uint32_t> immune;          // [C*S] 24-bit immunity thresholds
    uint32_t seed_h = 0, seed_m = 0;       // 24-bit seeding thresholds
    // dispersal rows: row_ptr[C+1]; entry e: to[e] (community), thr[e] (24-bit
    // alias threshold, 2^24 = always keep), alias[e] (slot within the row)
    std::vector<uint32_t> row_ptr, to, thr, alias;
    // in-edges of each destination, ascending source: in_ptr[C+1],
    // in_src[e] (source community), in_slot[e] (slot of the destination in
    // the source's row)
    std::vector<uint32_t> in_ptr, in_src, in_slot;
    uint64_t n_humans = 0, n_slots = 0;
};

// Returns an empty string on success, else the reason the input is invalid.
std::string validate(const ArboShape& shape, const ArboParams& p, uint32_t max_days);

// The parameters climate row `row` gives a day (SimulationParam.cpp:92-137):
// out = {beta_hm, beta_mh, eip, mosquito Weibull scale}, in double.
void climate_day(const ArboClimate& c, uint32_t row, double out[4]);

// Builds every table of one simulation.  validate() must have passed.
void build_tables(const ArboShape& shape, const ArboParams& p, uint64_t seed,
                  uint32_t max_days, SimTables& out);

// 24-bit fixed point of a probability: floor(f 2^24) clamped to [0, 2^24].
uint32_t fixed24(double f);

}  // namespace arbo
