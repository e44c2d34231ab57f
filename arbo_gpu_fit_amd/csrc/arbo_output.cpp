// arbo_output.cpp — the reference's output files from a simulation's series:
// nData of writeIncidence (writeoutput.cpp:13-98), nAlive of writeCensus
// (writeoutput.cpp:100-148), nAOI of writeAOI (writeoutput.cpp:152-191) and
// nSpatial of writeSpatial (writeoutput.cpp:194-241).
#include <cstdio>
#include <string>

#include "arbo_c.h"
#include "arbo_output.h"

namespace {
constexpr uint32_t kCensusRecord = 5000;   // humans per census (C_NSIZERECORD, constant.h:21)
}

extern "C" {

ARBO_API int arbo_write_ndata(const char* path, const ArboSeries* s, uint32_t rows, uint32_t S, uint32_t sim_run) {
    if (!path || !s || !s->n_count || !s->n_incidence || !s->n_infected || !s->n_one_subpop || rows == 0 ||
        S == 0 || S > ARBO_MAX_SEROTYPES)
        return ARBO_E_INVALID;
    FILE* f = std::fopen(path, sim_run == 0 ? "wb" : "ab");
    if (!f) return ARBO_E_INVALID;
    std::string out;
    out.reserve((size_t)rows * (16 + 36 * S));
    if (sim_run == 0) {
        out += "t,popSize";
        for (uint32_t v = 0; v < S; ++v) out += ",incidence_DENV" + std::to_string(v + 1);
        for (uint32_t v = 0; v < S; ++v) out += ",prevalence_DENV" + std::to_string(v + 1);
        for (uint32_t v = 0; v < S; ++v) out += ",prevalence_DENV" + std::to_string(v + 1) + "_random_community";
        out += ",simRun";
    }
    for (uint32_t t = 0; t < rows; ++t) {
        out += "\n" + std::to_string(t) + "," + std::to_string(s->n_count[t]);
        for (uint32_t v = 0; v < S; ++v) out += "," + std::to_string(s->n_incidence[(size_t)t * S + v]);
        for (uint32_t v = 0; v < S; ++v) out += "," + std::to_string(s->n_infected[(size_t)t * S + v]);
        for (uint32_t v = 0; v < S; ++v) out += "," + std::to_string(s->n_one_subpop[(size_t)t * S + v]);
        out += "," + std::to_string(sim_run);
    }
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? ARBO_OK : ARBO_E_INVALID;
}

ARBO_API int arbo_write_nspatial(const char* path, const uint32_t* sp, uint32_t rows, uint32_t S, uint32_t C,
                                 uint32_t sim_run) {
    if (!path || !sp || rows == 0 || C == 0 || S == 0 || S > ARBO_MAX_SEROTYPES) return ARBO_E_INVALID;
    FILE* f = std::fopen(path, sim_run == 0 ? "wb" : "ab");
    if (!f) return ARBO_E_INVALID;
    bool ok = true;
    std::string out;
    if (sim_run == 0) {
        out += "t,subPop";
        for (uint32_t v = 0; v < S; ++v) out += ",prevalence.DENV" + std::to_string(v + 1);
        out += ",simRun";
    }
    const std::string run = "," + std::to_string(sim_run);
    for (uint32_t t = 0; t < rows && ok; ++t) {
        const std::string ts = "\n" + std::to_string(t) + ",";
        for (uint32_t c = 0; c < C; ++c) {
            out += ts + std::to_string(c);
            for (uint32_t v = 0; v < S; ++v) out += "," + std::to_string(sp[((size_t)t * S + v) * C + c]);
        }
        out += run;   // once per day, after its last community (writeoutput.cpp:237)
        if (out.size() > (1u << 24)) {   // write in pieces: a long run has many lines
            ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
            out.clear();
        }
    }
    if (ok && !out.empty()) ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? ARBO_OK : ARBO_E_INVALID;
}

ARBO_API int arbo_write_nalive(const char* path, const uint32_t* age, const int32_t* hist, uint32_t n, uint32_t S,
                               uint32_t sim_run) {
    if (!path || (n && (!age || !hist)) || S == 0 || S > ARBO_MAX_SEROTYPES) return ARBO_E_INVALID;
    FILE* f = std::fopen(path, sim_run == 0 ? "wb" : "ab");
    if (!f) return ARBO_E_INVALID;
    std::string out;
    if (sim_run == 0) {
        out += "age";
        for (uint32_t v = 0; v < S; ++v) out += ",history" + std::to_string(v + 1);
        out += ",simRun";
    }
    const uint32_t rec = n < kCensusRecord ? n : kCensusRecord;
    const std::string run = "," + std::to_string(sim_run);
    for (uint32_t i = 0; i < rec; ++i) {
        out += "\n" + std::to_string(age[i]);
        for (uint32_t v = 0; v < S; ++v) out += "," + std::to_string(hist[(size_t)v * n + i]);
        out += run;
    }
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? ARBO_OK : ARBO_E_INVALID;
}

ARBO_API int arbo_write_naoi(const char* path, const uint32_t* aoi, uint32_t sim_run) {
    if (!path || !aoi) return ARBO_E_INVALID;
    FILE* f = std::fopen(path, sim_run == 0 ? "wb" : "ab");
    if (!f) return ARBO_E_INVALID;
    std::string out;
    if (sim_run == 0) out += "first,second,third,fourth,simRun";
    const std::string run = "," + std::to_string(sim_run);
    for (uint32_t r = 0; r < ARBO_AOI_RECORD; ++r) {
        out += "\n" + std::to_string(aoi[r]);
        for (uint32_t k = 1; k < 4; ++k) out += "," + std::to_string(aoi[k * ARBO_AOI_RECORD + r]);
        out += run;
    }
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? ARBO_OK : ARBO_E_INVALID;
}

}  // extern "C"
