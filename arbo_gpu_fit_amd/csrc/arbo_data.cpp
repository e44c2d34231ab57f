// arbo_data.cpp — the reference's input files (ClimateData.cpp:53-126,
// DengueData.cpp:49-112) and degree-heterogeneous community sizes
// (network.cpp:116-164).
#include "arbo_data.h"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "arbo_rng.h"

namespace {

std::string unquote(const std::string& s) {
    if (s.size() >= 2 && s.front() == '"') return s.substr(1, s.size() - 2);
    return s;
}

// Drops the carriage return of a line saved with CRLF endings.
void chomp(std::string& line) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
}

// Reads the named columns of a csv file: out[k][row] for names[k].  Entry j
// of a data line belongs to header column j mod (header columns).
int read_columns(const char* path, const std::vector<std::string>& names, std::vector<std::vector<float>>& out) {
    if (!path) return ARBO_E_INVALID;
    std::ifstream f(path);
    if (!f.is_open()) return ARBO_E_INVALID;
    std::string line, cell;
    if (!std::getline(f, line)) return ARBO_E_INVALID;
    chomp(line);
    std::vector<int> col(names.size(), -1);
    int ncol = 0;
    {
        std::istringstream in(line);
        while (std::getline(in, cell, ',')) {
            const std::string h = unquote(cell);
            for (size_t k = 0; k < names.size(); ++k)
                if (h == names[k]) col[k] = ncol;
            ++ncol;
        }
    }
    for (int c : col)
        if (c < 0) return ARBO_E_INVALID;
    out.assign(names.size(), {});
    while (std::getline(f, line)) {
        chomp(line);
        std::istringstream in(line);
        int j = 0;
        while (std::getline(in, cell, ',')) {
            const int c = j % ncol;
            for (size_t k = 0; k < names.size(); ++k) {
                if (c != col[k]) continue;
                const std::string v = unquote(cell);
                char* end = nullptr;
                const float x = std::strtof(v.c_str(), &end);
                if (end == v.c_str()) return ARBO_E_INVALID;   // not a number
                out[k].push_back(x);
            }
            ++j;
        }
    }
    return ARBO_OK;
}

}  // namespace

extern "C" {

ARBO_API int arbo_read_climate_csv(const char* path, float* temperature, float* humidity, float* precipitation,
                                   uint32_t cap, uint32_t* rows) {
    if (!rows) return ARBO_E_INVALID;
    std::vector<std::vector<float>> v;
    const int rc = read_columns(path, {"T", "H", "P"}, v);
    if (rc) return rc;
    *rows = (uint32_t)v[0].size();
    float* dst[3] = {temperature, humidity, precipitation};
    for (int k = 0; k < 3; ++k)
        if (dst[k])
            for (uint32_t i = 0; i < cap && i < v[k].size(); ++i) dst[k][i] = v[k][i];
    return ARBO_OK;
}

ARBO_API int arbo_read_dengue_csv(const char* path, float* incidence, uint32_t cap, uint32_t* rows) {
    if (!rows) return ARBO_E_INVALID;
    std::vector<std::vector<float>> v;
    const int rc = read_columns(path, {"incidence"}, v);
    if (rc) return rc;
    *rows = (uint32_t)v[0].size();
    if (incidence)
        for (uint32_t i = 0; i < cap && i < v[0].size(); ++i) incidence[i] = v[0][i];
    return ARBO_OK;
}

ARBO_API int arbo_community_sizes(uint32_t W, uint32_t H, uint32_t nb, uint32_t include_self, uint64_t total,
                                  float hetero, uint64_t seed, uint32_t* out) {
    if (!out || W < 1 || H < 1 || (nb != 4 && nb != 8) || !std::isfinite(hetero)) return ARBO_E_INVALID;
    const uint64_t C = (uint64_t)W * H;
    if (C > (1u << 24) || total < C || total > 0xffffffffull) return ARBO_E_INVALID;
    // cumulative relative sizes degree^hetero, normalised (network.cpp:138-159), as 24-bit thresholds
    std::vector<double> cum(C);
    double acc = 0.0;
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            uint32_t deg = include_self ? 1u : 0u;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if ((dx == 0 && dy == 0) || (nb == 4 && dx != 0 && dy != 0)) continue;
                    const int nx = (int)x + dx, ny = (int)y + dy;
                    if (nx >= 0 && ny >= 0 && nx < (int)W && ny < (int)H) ++deg;
                }
            // an isolated community (1x1 lattice without a self entry) counts as degree 1
            const double d = deg ? (double)deg : 1.0;
            acc = acc + (hetero == 0.0f ? 1.0 : arbo::dexp((double)hetero * arbo::dlog(d)));
            cum[(size_t)y * W + x] = acc;
        }
    std::vector<uint32_t> thr(C);
    for (uint64_t c = 0; c < C; ++c) {
        const double v = std::floor(cum[c] / acc * 16777216.0);
        thr[c] = v >= 16777216.0 ? (1u << 24) : (uint32_t)v;
    }
    thr[C - 1] = 1u << 24;
    for (uint64_t c = 0; c < C; ++c) out[c] = 1;
    const arbo::Key key{(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint64_t rest = total - C;
    for (uint64_t i = 0; i < rest; i += 4) {
        const arbo::U4 r = arbo::draw(key, (uint32_t)(i >> 2), arbo::INIT_DAY, arbo::P_SIZES, (uint32_t)(i >> 34));
        const uint32_t w[4] = {r.x, r.y, r.z, r.w};
        for (uint64_t j = 0; j < 4 && i + j < rest; ++j) {
            const uint32_t k = (w[j] >> 8) + 1;   // (0, 2^24]
            uint64_t lo = 0, hi = C - 1;           // min{c : k <= thr[c]}
            while (lo < hi) {
                const uint64_t mid = (lo + hi) / 2;
                if (thr[mid] >= k) hi = mid; else lo = mid + 1;
            }
            out[lo]++;
        }
    }
    return ARBO_OK;
}

}  // extern "C"
