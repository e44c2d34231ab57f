// arbo_device.h — HBM layout of a batch of simulations, as the kernels see it.
//
// Every simulation of the batch is flattened into global arrays:
//   human words   u32  h_word[]  (4-element aligned per community)
//   human next    u16  h_next[]
//   human birth   i32  h_birth[]   (cold: written at rebirth only)
//   mosquito word u16  m_word[]  (32-element aligned per community)
//   mosquito next u16  m_next[]
//   alive bitmap  u32  alive[]   (bit j of word m_off/32 + j/32: slot alive after emergence)
// A community is the unit of work of both day kernels; its descriptor gives
// the offsets of its members and of its CSR row and in-edges.
#pragma once
#include <stdint.h>

namespace arbo {

constexpr uint32_t kWave = 64;            // threads per workgroup: one wavefront per community
constexpr uint32_t kCntPerComm = 72;      // event counts of a source: [dir 2][serotype 4][row slot 9]
constexpr uint32_t kInboxEdges = 9;       // in-edges per destination (Moore-8 + self)
constexpr uint32_t kHumanAlign = 4;       // humans per 16-byte vector
constexpr uint32_t kMosqAlign = 128;      // mosquito slots per 16-byte group of alive-bitmap words
constexpr uint32_t kAoiRecord = 5000;     // ages per age-of-infection ring
// Per-community partials of a day, written by the demography kernel and
// summed per simulation by k_day_reduce: word 0 mosquitoes alive after the
// births, 1 humans, 2-3 infectious humans (two 16-bit serotype pairs), 4-5
// E->I onsets (pairs), 6-7 the 16-byte census vectors of humans and of
// mosquitoes the day's scan read, 8-23 the eight long-distance rates or
// counts (u64).
constexpr uint32_t kPartWords = 24;
constexpr uint32_t kMaxReduceSplit = 64;  // workgroups per simulation of k_day_reduce: about one
                                          // community per thread, 1..64, about 2,048 in all
constexpr uint32_t kAoiWindow = 3650;     // rings record the days t with max_days - t <= this

struct CommDesc {
    uint64_t h_off;       // global offset of the first human
    uint64_t m_off;       // global offset of the first mosquito slot
    uint32_t sim;         // simulation index
    uint32_t c;           // community index within the simulation
    uint32_t n;           // humans
    uint32_t cap;         // mosquito slots
    uint32_t h_base;      // simulation-local index of the first human (Philox entity)
    uint32_t m_base;      // simulation-local index of the first slot (Philox entity)
    uint32_t row_off;     // global offset of the community's CSR row
    uint32_t row_len;
    uint32_t in_off;      // global offset of its in-edges
    uint32_t in_len;
    float base;           // m_ratio * n (f32)
    uint32_t flags;       // kCommIntro | kCommStay | self column << kCommSelfShift
};

constexpr uint32_t kCommIntro = 1u;      // the simulation has external introductions
constexpr uint32_t kCommStay = 2u;       // a fraction no_distance of the events stays at the source
constexpr uint32_t kCommSelfShift = 8u;  // bits 8..11: the source's own column on its CSR row

struct SimDesc {
    uint32_t k0, k1;              // Philox key = 64-bit seed
    uint32_t C, S;                // communities, serotypes
    uint32_t comm_off;            // global index of community 0
    float beta_hm, beta_mh, omega, recovery, eip, intro;
    uint32_t iip_min, iip_span;   // incubation uniform on [iip_min, iip_min + iip_span)
    uint32_t death_h_off;         // offsets into the shared table pools
    uint32_t death_m_off;
    uint32_t init_h_off;
    uint32_t init_m_off;
    uint32_t dens_off;
    uint32_t immune_off;          // [C*S] thresholds
    uint32_t seed_h, seed_m;
    uint32_t death_m_max;         // largest death age the mosquito table can return
    uint32_t death_m_guide_off;   // [256] smallest death age of each life_bucket() of draws
    uint32_t count_law;           // ARBO_COUNT_POISSON or ARBO_COUNT_STOCH_ROUND
    float no_distance;            // fraction of events staying at the source (communities with kCommStay)
    uint64_t series_off;          // this simulation's block in the series array
    uint64_t spatial_off;         // its block [(T+1)][S][C] of the per-community series (if kept)
    uint64_t h_begin, h_end;      // its humans: [h_begin, h_end) of the padded census, h_end a multiple of 4
    uint32_t climate;             // 1: per-day beta/eip at clim_off and per-birth-day lifetime rows
    uint32_t clim_off;            // [max_days][3] beta_hm, beta_mh, eip in the dens pool
};

// Parameters of day t: the climate's or the constant ones.
struct DayParams {
    float beta_hm, beta_mh, eip;
};

__host__ __device__ inline DayParams day_params(const SimDesc& sd, const float* dens, uint32_t t) {
    if (!sd.climate) return DayParams{sd.beta_hm, sd.beta_mh, sd.eip};
    const float* d = dens + sd.clim_off + 3 * (size_t)t;
    return DayParams{d[0], d[1], d[2]};
}

// Lifetime row of mosquitoes born on day t (row 0: the t = 0 census).
__host__ __device__ inline uint32_t birth_row(const SimDesc& sd, uint32_t t) { return sd.climate ? t + 1 : 0u; }

// Series block of one simulation with room for T days (T+1 rows):
//   n_count[T+1] m_count[T+1] n_inf[(T+1)S] n_inc[(T+1)S] n_one[(T+1)S]
struct SeriesIdx {
    uint64_t n_count, m_count, n_inf, n_inc, n_one;
};

__host__ __device__ inline SeriesIdx series_index(uint64_t off, uint32_t T, uint32_t S) {
    SeriesIdx x;
    const uint64_t R = (uint64_t)T + 1;
    x.n_count = off;
    x.m_count = off + R;
    x.n_inf = off + 2 * R;
    x.n_inc = off + 2 * R + R * S;
    x.n_one = off + 2 * R + 2 * R * S;
    return x;
}

__host__ __device__ inline uint64_t series_words(uint32_t T, uint32_t S) {
    return ((uint64_t)T + 1) * (2 + 3 * (uint64_t)S);
}

struct EngineView {
    const CommDesc* comm;
    const SimDesc* sims;
    uint32_t* h_word;
    uint16_t* h_next;
    int32_t* h_birth;
    uint16_t* m_word;
    uint16_t* m_next;
    uint32_t* alive;               // [M/32] alive-slot bitmap of the current day
    uint32_t* cnt;                 // [G][9 in-edges][8 = dir*4 + serotype] local events arriving today
    uint32_t* nld;                 // [G][8] long-distance arrivals of the current day
    uint32_t* active;              // [G] 1 when events arrive today (set by the sources, cleared by the
                                   // destination's event kernel)
    uint32_t* hsum;                // [G][4] summary: earliest day any human has an event (0 = unknown),
                                   // infectious humans by serotype (two 16-bit pairs), mosquito scan
                                   // limit + 1 in 8-slot vectors (0 = unknown)
    unsigned long long* lam;       // [2 parity][n_sims][8] long-distance events of the day: the rate in
                                   // 2^-24 units (Poisson counts) or the count (stochastic rounding)
    uint32_t* series;
    uint32_t* part;                // [G][kPartWords] per-community partials of the day (see kPartWords)
    unsigned long long* scanned;   // [n_sims][2] census vectors read by the demography kernel (humans,
                                   // mosquitoes), summed over days by k_day_reduce
    uint32_t* spatial;             // per-community infectious humans [(T+1)][S][C] per simulation, or null
    int32_t* hist;                 // per human 4 infection-history words (one per serotype), or null
    uint8_t* aoi_mark;             // per human: 1 when a mosquito event infected it today (age-of-infection
                                   // rings kept), else null
    uint32_t* aoi;                 // [n_sims][4][kAoiRecord] age-of-infection rings
    uint32_t* aoi_pos;             // [n_sims][4] next position of each ring
    uint32_t aoi_from;             // first day whose infections enter the rings
    const uint32_t* csr_inbox;     // per CSR entry: inbox row (destination * 9 + in-edge index)
    const uint32_t* csr_thr;       // bits 0..24 threshold, 27..31 alias slot
    const uint32_t* in_edge;       // global source community
    const uint32_t* in_slot;       // slot of the destination in the source's row
    const uint32_t* tables;        // death / initial-age / immunity pools
    const float* dens;
    const uint32_t* day;           // device day counter (graph replay)
    uint32_t day_offset;           // the launch's day is *day + day_offset (a captured block of days
                                   // reads one counter, advanced once per block)
    uint32_t n_sims;
    uint32_t n_comm;
    uint32_t max_days;
    uint32_t death_m_lds;          // entries of the largest mosquito death table (LDS staging)
    uint32_t reduce_split;         // workgroups per simulation of k_day_reduce
    uint32_t max_words;            // alive-bitmap words of the batch's largest community
};

// Launch knobs of the day kernels.  None of them changes a result: every
// random draw is keyed by model entities, never by workgroup or lane.
struct Tuning {
    uint32_t unroll;     // demography: 16-byte census loads in flight per lane
    uint32_t early_m;    // demography: of the first batch, mosquito loads issued before the humans
    uint32_t xcd;        // bit 0 demography, bit 1 events: the workgroups one XCD receives take a
                         // contiguous range of communities
    uint32_t ld_split;   // long-distance kernel: workgroups per simulation
    uint32_t wps;        // demography: 0 the compiler's register budget, 8 eight waves per SIMD
    uint32_t nw;         // demography: waves per community (1, 2, 4)
};
constexpr uint32_t kDemographyShapes[][2] = {{4, 4}, {4, 2}, {8, 4}, {2, 2}};   // (unroll, early_m)
constexpr uint32_t kOccShapes[][2] = {{4, 2}, {2, 2}};   // shapes also built for 8 waves per SIMD
constexpr uint32_t kWideShapes[][3] = {{4, 2, 2}, {2, 2, 4}};   // (unroll, early_m, waves per community)
constexpr uint32_t kMaxLdSplit = 64;
constexpr uint32_t kXcds = 8;   // gfx950: workgroups go to the 8 XCDs in turn

// Community of workgroup b of n when the workgroups of XCD b % 8 take the
// contiguous range [x q + min(x, r), ...) of q + (x < r) communities, where
// n = 8 q + r.  A bijection of [0, n).
__host__ __device__ inline uint32_t xcd_block(uint32_t b, uint32_t n) {
    const uint32_t x = b % kXcds, i = b / kXcds, q = n / kXcds, r = n % kXcds;
    return x * q + (x < r ? x : r) + i;
}

hipError_t launch_init(const EngineView& v, hipStream_t st);
hipError_t launch_demography(const EngineView& v, const Tuning& tu, hipStream_t st);
hipError_t launch_events(const EngineView& v, const Tuning& tu, hipStream_t st);
hipError_t launch_longdist(const EngineView& v, const Tuning& tu, hipStream_t st);
hipError_t launch_advance(uint32_t* day, uint32_t n, hipStream_t st);
hipError_t launch_fill(uint32_t* p, uint64_t n, uint32_t value, hipStream_t st);
bool tuning_supported(const Tuning& tu);

}  // namespace arbo
