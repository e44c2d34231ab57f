// arbo_kernels.hip — the day-loop kernels of the dengue IBM for gfx950.
//
// One simulated day is two launches over every community of every simulation
// of the batch, one wave64 workgroup per community (several communities in
// flight per SIMD, so the dependent-latency chain of one community is hidden
// by the others), plus a small per-simulation long-distance kernel between
// them:
//
//  k_day_demography  ndemographic.cu:32-105 + mdemographic.cu:22-150 +
//                    nVisitingInfected/mVisitingInfected (nepidemic.cu:28-142,
//                    mepidemic.cu:28-138) + dataCollect (datacollect.cu:79-189)
//     streams the community's packed census once (16-byte vector loads, four
//     in flight per lane, of u32 human words and u16 mosquito words), applies
//     deaths, rebirths and E->I / I->S transitions, recounts the infectious
//     per serotype with wave reductions, fills the lowest-index dead mosquito
//     slots up to the seasonal target with a wave prefix scan over the
//     dead-slot bitmap (newborn lifetimes from an LDS copy of the mosquito
//     lifetime table), publishes the alive-slot bitmap, draws the Poisson
//     event counts of the community as a source, spreads the local events
//     over its CSR row with the alias table into an LDS histogram, and adds
//     its counters into the series row of the next day.
//
//  k_day_events      mepidemic.cu:143-246 (nmTransmission) +
//                    nepidemic.cu:146-319 (mnTransmission)
//     gathers the event counts of the community's in-edges and its
//     long-distance arrivals, orders each direction's events (serotype
//     rotation, source, index) and applies them in chunks with an LDS
//     claim-min hash table: the earliest event whose target was eligible at
//     the start of the chunk wins, which is exactly the serial outcome.
//     Mosquito targets are the q-th alive slot, found from the alive bitmap
//     through a coarse rank index.  Events onto mosquitoes and onto humans
//     have disjoint targets, so they share one chunk sequence (and, without
//     external introductions, so do the serotypes).  No global atomics on
//     the census.
//
// All randomness is Philox keyed by (seed; entity, day, purpose, index), so a
// simulation's results do not depend on launch shape or batch composition.
#include <hip/hip_runtime.h>

#include "arbo_c.h"
#include "arbo_device.h"
#include "arbo_rng.h"

namespace arbo {

constexpr uint32_t kWordsPerLane = 4;                          // alive-bitmap words per thread
constexpr uint32_t kWaveWords = kWave * kWordsPerLane;         // bitmap words of one wave: 8,192 slots
constexpr uint32_t kMaxWords = (ARBO_MAX_COMMUNITY_SLOTS + 31) / 32;   // bitmap words of the largest community
constexpr uint32_t kMaxWaves = (kMaxWords + kWaveWords - 1) / kWaveWords;   // waves that cover it: 8
// The demography kernel's census loads in flight per lane (UNROLL) and, of
// the first batch, the mosquito vectors issued before the human scan
// (EARLY_M) are template parameters chosen per engine (Tuning).
constexpr uint32_t kChunk = 256;                               // events per claim pass
constexpr uint32_t kPerLane = kChunk / kWave;                  // 4
constexpr uint32_t kHash = 512;                                // claim-table entries (2 x chunk)
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kKeyBits = 17;                              // member index < 2^17
constexpr uint32_t kKeyMask = (1u << kKeyBits) - 1u;
constexpr uint32_t kHumanTable = 65536;
constexpr uint32_t kMaxRowLen = 9;                             // Moore-8 + self
constexpr uint32_t kGuide = 256;                               // death-age guide entries (life_bucket, padded)
static_assert(kMaxWaves == 8, "eight waves cover the largest community's bitmap");
static_assert(kLifeBuckets <= kGuide, "guide table holds every life bucket");
static_assert(ARBO_MAX_COMMUNITY_SLOTS <= (1u << kKeyBits) && ARBO_MAX_COMMUNITY_HUMANS <= (1u << kKeyBits),
              "member index must fit the claim key");

// Wavefront reductions, every lane active: DPP within each row of 16 lanes
// (quad_perm 1,0,3,2; quad_perm 2,3,0,1; row_half_mirror; row_mirror), then
// the four row results read from lanes 0, 16, 32 and 48.  The result is
// wave-uniform.
template <class Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t x, Op op) {
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false));
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false));
    return op(op((uint32_t)__builtin_amdgcn_readlane((int)x, 0), (uint32_t)__builtin_amdgcn_readlane((int)x, 16)),
              op((uint32_t)__builtin_amdgcn_readlane((int)x, 32), (uint32_t)__builtin_amdgcn_readlane((int)x, 48)));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    return wave_reduce(x, [](uint32_t a, uint32_t b) { return a + b; });
}

__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
    return wave_reduce(x, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
}

__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
    return wave_reduce(x, [](uint32_t a, uint32_t b) { return a < b ? a : b; });
}

__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// Two 16-bit differences a - b (mod 2^16) in one word (v_pk_sub_u16).
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t sub_u16x2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) - __builtin_bit_cast(u16x2, b));
}

// Exclusive prefix sum over the wavefront, every lane active; `total` gets
// the wave sum.  An inclusive scan by DPP: row_shr 1, 2, 4, 8 within each
// row of 16 lanes (lanes shifted in from outside the row add 0), then lane
// 15 of rows 0 and 2 added to rows 1 and 3 (row_bcast:15) and lane 31 to
// rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_exclusive_scan(uint32_t x, uint32_t& total) {
    uint32_t v = x;
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);   // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);   // row_bcast:31
    total = (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
    return v - x;
}

// min{a in [1, hi] : K[a] >= k}; hi when K[hi] < k (hi = table end).
template <class P>
__device__ __forceinline__ uint32_t death_age(P K, uint32_t hi, uint32_t k) {
    if (K[hi] < k) return hi;
    uint32_t lo = 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (K[mid] >= k) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// The same minimum, started from the smallest death age of k's bucket:
// guide[b] = min{a >= 1 : K[a] >= life_bucket_kmin(b)} <= result.
__device__ __forceinline__ uint32_t death_age_guided(const uint32_t* K, const uint32_t* guide, uint32_t hi,
                                                     uint32_t k) {
    uint32_t a = guide[life_bucket(k)];
    while (a < hi && K[a] < k) ++a;
    return a;
}

// Death age given survival to age a0: PLE uniform on (F(a0-1), 1].
__device__ __forceinline__ uint32_t death_age_given(const uint32_t* K, uint32_t hi, uint32_t a0, uint32_t r) {
    const uint32_t prev = K[a0 >= 1 ? a0 - 1 : 0];
    if (prev >= (1u << 24)) return a0;
    return death_age(K, hi, prev + 1 + mulhi(r, (1u << 24) - prev));
}

// min{i : k <= T[i]} over a table ending at 2^24.
__device__ __forceinline__ uint32_t pick_index(const uint32_t* T, uint32_t n, uint32_t k) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (T[mid] >= k) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ void add_pair(uint32_t& lo, uint32_t& hi, uint32_t s) {
    const uint32_t inc = 1u << ((s & 1u) * 16u);
    if (s < 2) lo += inc; else hi += inc;
}

__device__ __forceinline__ uint32_t unpack_pair(uint32_t lo, uint32_t hi, uint32_t s) {
    const uint32_t v = (s < 2) ? lo : hi;
    return (v >> ((s & 1u) * 16u)) & 0xFFFFu;
}

// Bits of bitmap word w that are slots below cap.
__device__ __forceinline__ uint32_t valid_bits(uint32_t w, uint32_t cap) {
    const uint32_t first = 32u * w;
    if (first + 32u <= cap) return 0xFFFFFFFFu;
    if (first >= cap) return 0u;
    return (1u << (cap - first)) - 1u;
}

// Position of the k-th (0-based) set bit of x; k < popcount(x).
__device__ __forceinline__ uint32_t select_in_word(uint32_t x, uint32_t k) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t width = 16; width >= 1; width >>= 1) {
        const uint32_t c = __popc(x & ((1u << width) - 1u));
        if (k >= c) { k -= c; x >>= width; pos += width; }
    }
    return pos;
}

// The q-th set bit of a bitmap given its words and their exclusive popcount
// prefix: word index w and bit position pos.
template <class Pre>
__device__ __forceinline__ void select_bit(uint32_t q, const uint32_t* word, const Pre* wpre, uint32_t nwords,
                                           uint32_t& w, uint32_t& pos) {
    uint32_t lo = 0, hi = nwords - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (wpre[mid] <= q) lo = mid; else hi = mid - 1;
    }
    w = lo;
    pos = select_in_word(word[lo], q - wpre[lo]);
}

// ---------------------------------------------------------------------------
// t = 0 census: demographicintial.cu:86-297, epidemicinitial.cu:27-99.
__global__ __launch_bounds__(kWave) void k_init(EngineView v) {
    const uint32_t g = blockIdx.x, lane = threadIdx.x;
    const CommDesc cd = v.comm[g];
    const SimDesc sd = v.sims[cd.sim];
    const Key key{sd.k0, sd.k1};
    const uint32_t S = sd.S;
    const uint32_t* death_h = v.tables + sd.death_h_off;
    const uint32_t* death_m = v.tables + sd.death_m_off;
    const uint32_t* init_h = v.tables + sd.init_h_off;
    const uint32_t* init_m = v.tables + sd.init_m_off;
    const uint32_t* immune = v.tables + sd.immune_off + (size_t)cd.c * S;
    uint32_t alive0;
    {
        const U4 r = draw(key, cd.c, INIT_DAY, P_INIT_MT, 0);
        const uint32_t a = stoch_round(v.dens[sd.dens_off] * cd.base, r.x);
        alive0 = a < cd.cap ? a : cd.cap;
    }
    uint32_t ih01 = 0, ih23 = 0;
    for (uint32_t i = lane; i < cd.n; i += kWave) {
        const uint32_t h = cd.h_base + i;
        const U4 r0 = draw(key, h, INIT_DAY, P_INIT_H, 0);
        const uint32_t y = pick_index(init_h, 151, (r0.x >> 8) + 1);
        const uint32_t a0 = 365u * y + mulhi(r0.y, 365);
        const uint32_t da = death_age_given(death_h, kHumanTable - 1, a0, r0.z);
        const int32_t ddi = (int32_t)da - (int32_t)(a0 + 1);
        const uint32_t dd = ddi < 0 ? 0u : (uint32_t)ddi;
        const U4 r1 = draw(key, h, INIT_DAY, P_INIT_H, 1);
        const uint32_t rr[4] = {r1.x, r1.y, r1.z, r1.w};
        uint32_t imm = 0;
        for (uint32_t s = 0; s < S; ++s)
            if ((rr[s] >> 8) < immune[s]) imm |= 1u << s;
        uint32_t st = 0, strain = 0, nxt = 0;
        if ((r0.w >> 8) < sd.seed_h) {
            const U4 r2 = draw(key, h, INIT_DAY, P_INIT_H, 2);
            strain = mulhi(r2.x, S);
            st = 2;
            imm |= 1u << strain;
            uint32_t R = stoch_round(sd.recovery, r2.z);
            if (R < 1) R = 1;
            nxt = mulhi(r2.y, R);
            add_pair(ih01, ih23, strain);
        }
        const uint64_t gi = cd.h_off + i;
        v.h_word[gi] = (dd << 8) | (imm << 4) | (strain << 2) | st;
        v.h_next[gi] = (uint16_t)nxt;
        v.h_birth[gi] = -(int32_t)(a0 + 1);
        if (v.hist)   // immune at t = 0 (prior immunity or the seeded infection): -2, naive: -1
            reinterpret_cast<int4*>(v.hist)[gi] = make_int4((imm & 1u) ? -2 : -1, (imm & 2u) ? -2 : -1,
                                                            (imm & 4u) ? -2 : -1, (imm & 8u) ? -2 : -1);
    }
    for (uint32_t j = lane; j < cd.cap; j += kWave) {
        const uint64_t gm = cd.m_off + j;
        v.m_next[gm] = 0;
        if (j >= alive0) { v.m_word[gm] = 3; continue; }
        const uint32_t m = cd.m_base + j;
        const U4 r0 = draw(key, m, INIT_DAY, P_INIT_M, 0);
        const uint32_t a0 = pick_index(init_m, 301, (r0.x >> 8) + 1);
        const uint32_t da = death_age_given(death_m, sd.death_m_max, a0, r0.z);
        const uint32_t dd12 = (uint32_t)((int32_t)da - (int32_t)(a0 + 1)) & 0xFFFu;
        uint32_t st = 0, strain = 0;
        if ((r0.w >> 8) < sd.seed_m) {
            st = 2;
            strain = mulhi(r0.y, S);
        }
        v.m_word[gm] = (uint16_t)((dd12 << 4) | (strain << 2) | st);
    }
    ih01 = wave_sum(ih01);
    ih23 = wave_sum(ih23);
    const SeriesIdx si = series_index(sd.series_off, v.max_days, S);
    if (lane == 0) {
        atomicAdd(&v.series[si.n_count], cd.n);
        atomicAdd(&v.series[si.m_count], alive0);
    }
    if (lane < S) {
        const uint32_t ih = unpack_pair(ih01, ih23, lane);
        atomicAdd(&v.series[si.n_inf + lane], ih);
        if (cd.c == 0) v.series[si.n_one + lane] = ih;
        if (v.spatial) v.spatial[sd.spatial_off + (uint64_t)lane * sd.C + cd.c] = ih;   // row 0
    }
}

// ---------------------------------------------------------------------------
// Day kernel A: demography, counters, event counts, series row t+1.

// Per-lane tallies of the exposed and infectious humans: infectious (ih)
// and E -> I onsets today (inc) as pairs of 16-bit serotype counts, and the
// earliest next state-change day of those members (minev).
struct HumanCount { uint32_t ih01, ih23, inc01, inc23, minev; };

__device__ __forceinline__ uint32_t day_of(uint32_t t, uint32_t nx) {   // absolute day of a 16-bit day near t
    return (uint32_t)((int32_t)t + (int32_t)(int16_t)(uint16_t)(nx - t));
}

// The members of one vector of humans (ndemographic.cu:57-103; four packed
// words, index 4q..4q+3) that are exposed or infectious and survive today
// (word >= tt).  An I -> R recovery is made here.  An E -> I onset is counted
// here and returned (bit e: member e): its infectious period is drawn later,
// with the wave's other deferred changes (human_deferred); its E -> I day
// goes to onx[e].  The counts and every other member's next state-change day
// go to the lane's tallies.  The next-event days of the vector are read with
// one 8-byte load.
__device__ __forceinline__ uint32_t human_states(const EngineView& v, const CommDesc& cd, uint32_t t, uint32_t tt,
                                                 uint32_t q, const uint32_t (&ws)[4], HumanCount& hc,
                                                 uint16_t* onx) {
    // The tallies are updated in place only on this (rare) path.  The empty
    // asm hides their values from the compiler, which would otherwise carry
    // each increment out of the branch and add it after the join: a register
    // cleared for every count on every vector scanned, infected or not.
    asm volatile("" : "+v"(hc.ih01), "+v"(hc.ih23), "+v"(hc.inc01), "+v"(hc.inc23), "+v"(hc.minev));
    const uint2 nx2 = *reinterpret_cast<const uint2*>(v.h_next + cd.h_off + 4 * q);
    const uint32_t nxs[4] = {nx2.x & 0xFFFFu, nx2.x >> 16, nx2.y & 0xFFFFu, nx2.y >> 16};
    uint32_t ons = 0;
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
        const uint32_t w = ws[e];
        const uint32_t st = w & 3u, strain = (w >> 2) & 3u;
        if (st == 0u || w < tt) continue;
        const uint32_t nx = nxs[e];
        const bool due = (int16_t)(uint16_t)(t - nx) > 0;
        uint32_t ev = 0xFFFFFFFFu;                // next state change: the day after `next`
        if (st == 2) {
            if (due) {
                v.h_word[cd.h_off + 4 * q + e] = w & ~0xFu;   // recovered
            } else {
                add_pair(hc.ih01, hc.ih23, strain);
                ev = day_of(t, nx) + 1;
            }
        } else if (due) {                        // E -> I: counted now, its infectious period drawn later
            add_pair(hc.inc01, hc.inc23, strain);
            add_pair(hc.ih01, hc.ih23, strain);
            ons |= 1u << e;
            onx[e] = (uint16_t)nx;
        } else {
            ev = day_of(t, nx) + 1;
        }
        hc.minev = ev < hc.minev ? ev : hc.minev;
    }
    return ons;
}

// Queue of the wave's deferred human changes (in LDS): entry k holds the
// member index (bits 0-15) and, for an onset, its E -> I day (bits 16-31) in
// qi[k], the member's word in qw[k].  A death is told from an onset by the
// word: the member dies today iff its death day (bits 8-31) is at most t.
constexpr uint32_t kHumanQueue = 256;
static_assert(ARBO_MAX_COMMUNITY_HUMANS <= 65536u, "queue entries keep the member index in 16 bits");

// The deferred changes, one per lane: a death draws the newborn's lifetime
// (P_H_BIRTH) and rebirths the member; an onset draws the infectious period
// (P_H_REC) and makes the member infectious.  The same draws and writes as
// made in place, so the census is unchanged by the deferral.
template <bool HIST>
__device__ __forceinline__ void human_deferred(const EngineView& v, const CommDesc& cd, const SimDesc& sd, Key key,
                                               uint32_t t, const uint32_t* qi, const uint32_t* qw, uint32_t n,
                                               uint32_t lane, uint32_t& minev) {
    for (uint32_t k = lane; k < n; k += kWave) {
        const uint32_t e = qi[k], i = e & 0xFFFFu, w = qw[k];
        const uint64_t gi = cd.h_off + i;
        if (w < ((t + 1u) << 8)) {   // death and instant rebirth
            const U4 r = draw(key, cd.h_base + i, t, P_H_BIRTH, 0);
            const uint32_t da = death_age(v.tables + sd.death_h_off, kHumanTable - 1, (r.x >> 8) + 1);
            v.h_word[gi] = (t + da) << 8;
            v.h_birth[gi] = (int32_t)t;
            if (HIST && v.hist) reinterpret_cast<int4*>(v.hist)[gi] = make_int4(-1, -1, -1, -1);   // naive newborn
            minev = t + da < minev ? t + da : minev;
        } else {                   // E -> I
            const U4 r = draw(key, cd.h_base + i, t, P_H_REC, 0);
            const uint32_t nn = (e >> 16) + stoch_round(sd.recovery, r.x);
            v.h_next[gi] = (uint16_t)nn;
            v.h_word[gi] = (w & ~3u) | 2u;
            const uint32_t ev = day_of(t, nn) + 1;
            minev = ev < minev ? ev : minev;
        }
    }
}

// HIST: the infection history or the per-community series may be kept
// (v.hist or v.spatial not null); a separate instantiation keeps their
// stores out of the plain day loop.
// WPS: 0, or the resident waves per SIMD the register budget must allow.
// NW: waves per community (workgroup of 64 NW threads).  The census loops
// stride over the workgroup, each wave queues its own deferred human changes,
// and the counts, the dead-slot prefix and the high-water mark are combined
// across the waves through LDS; a community may hold up to NW times
// kWaveWords bitmap words of slots (four per thread).
template <uint32_t UNROLL, uint32_t EARLY_M, bool XCD, bool HIST, uint32_t WPS, uint32_t NW>
__global__ __launch_bounds__(kWave * NW) __attribute__((amdgpu_waves_per_eu(WPS ? WPS : 1)))
void k_day_demography(EngineView v) {
    static_assert(EARLY_M >= 1 && EARLY_M <= UNROLL, "early mosquito batch within the first unrolled batch");
    static_assert(NW == 1 || NW == 2 || NW == 4 || NW == 8, "one to eight waves per community");
    constexpr uint32_t NT = kWave * NW;   // threads per community
    const uint32_t g = XCD ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x, tid = threadIdx.x;
    const uint32_t lane = NW == 1 ? tid : tid & (kWave - 1), wv = NW == 1 ? 0u : tid / kWave;
    const uint32_t t = *v.day + v.day_offset;
    const CommDesc cd = v.comm[g];
    // Humans are scanned only on days when one of them may die or change
    // state: the summary holds the earliest such day (0 forces a scan) and the
    // infectious counts, which nothing else changes in between.
    const uint4 hs = reinterpret_cast<const uint4*>(v.hsum)[g];
    const bool scan_h = t >= hs.x;
    // the first census batches go out before anything else
    const uint32_t nvec = (cd.n + 3) >> 2;
    const uint32_t mvec = (cd.cap + 7) >> 3;
    // Mosquito slots from vector mlim on are all empty (births fill the
    // lowest empty slots, so the alive slots stay packed low): they are not
    // read.  The summary keeps mlim + 1 (0 = unknown: scan every slot).
    const uint32_t mlim = (hs.w == 0u || hs.w - 1u > mvec) ? mvec : hs.w - 1u;
    const uint4* w4p = reinterpret_cast<const uint4*>(v.h_word + cd.h_off);
    const uint4* m4p = reinterpret_cast<const uint4*>(v.m_word + cd.m_off);
    // lanes past the last vector of humans read that vector again (its words
    // are not counted), so every lane tests defined words
    uint4 hb[UNROLL], mb[EARLY_M];
#pragma unroll
    for (uint32_t u = 0; u < UNROLL; ++u) {
        const uint32_t q = u * NT + tid;
        if (scan_h) hb[u] = w4p[umin(q, nvec - 1)];
    }
#pragma unroll
    for (uint32_t u = 0; u < EARLY_M; ++u) {
        const uint32_t q = u * NT + tid;
        if (q < mlim) mb[u] = m4p[q];
    }
    const SimDesc sd = v.sims[cd.sim];
    const Key key{sd.k0, sd.k1};
    const uint32_t S = sd.S;

    extern __shared__ uint32_t s_dtab[];                            // mosquito lifetime thresholds
    __shared__ uint32_t s_dead[kWaveWords * NW];                    // dead-slot bitmap
    __shared__ uint32_t s_died[kWaveWords * NW];                    // died today
    __shared__ __align__(16) uint32_t s_dpre[kWaveWords * NW];      // dead slots before each word
    __shared__ uint32_t s_red[NW][8];                               // per-wave counts (NW > 1)
    __shared__ uint32_t s_wtot[NW];                                 // per-wave dead slots (NW > 1)
    __shared__ uint32_t s_wtop[NW];                                 // per-wave high-water marks (NW > 1)
    __shared__ uint32_t s_mguide[kLifeBuckets];                     // death-age guide of the lifetime table
    __shared__ uint32_t s_row[kMaxRowLen];                          // alias row of the community
    __shared__ uint32_t s_n[8];
    __shared__ uint32_t s_hist[kCntPerComm];
    uint8_t* s_dead8 = reinterpret_cast<uint8_t*>(s_dead);
    // E -> I days of each lane's onsets in the vector being scanned: the
    // dead-slot prefix array, not used before the mosquito scan
    uint16_t* s_onx = reinterpret_cast<uint16_t*>(s_dpre);
    static_assert(4 * kWave * sizeof(uint16_t) <= kWaveWords * sizeof(uint32_t), "onset days fit the prefix array");
    uint8_t* s_died8 = reinterpret_cast<uint8_t*>(s_died);

    // Every small operand of the day goes out now, behind the first census
    // batches, so its latency is spent while the census streams in: the
    // mosquito lifetime table and its guide (straight to LDS), the alias row
    // and the inbox rows of the community's CSR entries, the next day's
    // seasonal density.
    {
        const uint32_t nd = sd.death_m_max + 1;
        const uint32_t row = birth_row(sd, t);   // today's births
        const uint32_t* dt = v.tables + sd.death_m_off + (size_t)row * nd;
        const uint32_t* gt = v.tables + sd.death_m_guide_off + (size_t)row * kGuide;
        constexpr uint32_t kT = (4 + NW - 1) / NW, kG = (kLifeBuckets + NT - 1) / NT;
        uint32_t tv[kT], gv[kG];
#pragma unroll
        for (uint32_t k = 0; k < kT; ++k) {
            const uint32_t a = k * NT + tid;
            tv[k] = a < nd ? dt[a] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kG; ++k) {
            const uint32_t b = k * NT + tid;
            gv[k] = b < kLifeBuckets ? gt[b] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kT; ++k) {
            const uint32_t a = k * NT + tid;
            if (a < nd) s_dtab[a] = tv[k];
        }
        for (uint32_t a = kT * NT + tid; a < nd; a += NT) s_dtab[a] = dt[a];   // long lifetimes
#pragma unroll
        for (uint32_t k = 0; k < kG; ++k) {
            const uint32_t b = k * NT + tid;
            if (b < kLifeBuckets) s_mguide[b] = gv[k];
        }
    }
    if (tid < cd.row_len) s_row[tid] = v.csr_thr[cd.row_off + tid];
    const uint32_t inboxent = tid < cd.row_len ? v.csr_inbox[cd.row_off + tid] : 0u;
    for (uint32_t i = tid; i < kCntPerComm; i += NT) s_hist[i] = 0;
    // emergence target of the day: the next day's expected population E[t+1]
    // (simulation.cpp:129-132), drawn once the census is counted
    const float dens_next = v.dens[sd.dens_off + t + 1];

    // ---- humans ----
    HumanCount hc{0, 0, 0, 0, 0xFFFFFFFFu};
    // Changes that need a random draw queue up in LDS (the dead-slot bitmap
    // arrays, not used before the mosquito scan; one queue per wave) and are
    // processed a lane each when the queue could overflow and after the scan.
    // With one wave the queue is guarded by workgroup barriers; with several
    // the waves flush at different times, so each wave waits for its own LDS
    // operations instead (s_waitcnt lgkmcnt(0)), behind a fence that keeps
    // the compiler from moving memory accesses across.
    uint32_t* hq_i = s_dead + wv * kHumanQueue;
    uint32_t* hq_w = s_died + wv * kHumanQueue;
    uint32_t hq_n = 0;   // wave-uniform
    auto queue_sync = []() {
        if (NW == 1) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_s_waitcnt(0xC07F);   // gfx9 encoding: lgkmcnt(0), vmcnt and expcnt unconstrained
            __builtin_amdgcn_wave_barrier();
        }
    };
    static_assert(kWaveWords >= kHumanQueue, "each wave's human queue fits its share of the bitmap arrays");
    // Humans of ndemographic.cu:57-103, four to a lane (one 16-byte vector of
    // packed words, index 4q..4q+3).  A member dies today iff its word is
    // below tt (death day <= t; t < 2^23, so no overflow).  The words past
    // the community's last member in its last vector read as kPadWord: a
    // member that neither dies nor changes state.  Only a vector with an
    // exposed or infectious member reads its next-event days
    // (human_states).  The changes that need a random draw (a death with its
    // rebirth's lifetime, an E -> I onset with its infectious period) are
    // queued: member e of each lane goes to hq_n + (changes of members < e in
    // the wave) + (lanes below with a change of member e).
    constexpr uint32_t kPadWord = 0xFFFFFF00u;
    const uint32_t tt = (t + 1u) << 8;
    const uint32_t nlast = (cd.n & 3u) ? cd.n >> 2 : 0xFFFFFFFFu;   // the vector with fewer than four members
    uint32_t minw = 0xFFFFFFFFu;   // smallest word of the lane's members that survive today
    auto scan_vector = [&](uint32_t q, uint4 w4) {   // every lane calls it; lanes past nvec count nothing
        uint32_t ws[4] = {w4.x, w4.y, w4.z, w4.w};
        const bool in = q < nvec;
        if (q == nlast) {
#pragma unroll
            for (uint32_t e = 1; e < 4; ++e)
                if (4 * q + e >= cd.n) ws[e] = kPadWord;
        }
        // deaths, straight from the compares; the smallest surviving word
        uint64_t bal[4];
        const uint64_t bin = __builtin_amdgcn_ballot_w64(in);
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) bal[e] = __builtin_amdgcn_ballot_w64(ws[e] < tt) & bin;
        const uint32_t m01 = umin(ws[0] < tt ? 0xFFFFFFFFu : ws[0], ws[1] < tt ? 0xFFFFFFFFu : ws[1]);
        const uint32_t m23 = umin(ws[2] < tt ? 0xFFFFFFFFu : ws[2], ws[3] < tt ? 0xFFFFFFFFu : ws[3]);
        minw = in ? umin(minw, umin(m01, m23)) : minw;
        // exposed and infectious members (rare): their onsets join the deaths
        uint32_t ons = 0;
        if (in && ((ws[0] | ws[1] | ws[2] | ws[3]) & 3u) != 0u)
            ons = human_states(v, cd, t, tt, q, ws, hc, s_onx + 4 * tid);
        if (__builtin_amdgcn_ballot_w64(ons != 0u)) {
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) bal[e] |= __builtin_amdgcn_ballot_w64((ons >> e) & 1u);
        }
        uint32_t before[4], many = 0;
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            before[e] = many;
            many += (uint32_t)__popcll(bal[e]);
        }
        if (many == 0) return;
        if (hq_n + many > kHumanQueue) {
            queue_sync();
            human_deferred<HIST>(v, cd, sd, key, t, hq_i, hq_w, hq_n, lane, hc.minev);
            queue_sync();
            hq_n = 0;
        }
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            if (!((in && ws[e] < tt) || ((ons >> e) & 1u))) continue;
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal[e] >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)bal[e], 0u));
            const uint32_t pos = hq_n + before[e] + below;
            hq_i[pos] = (4u * q + e) | (((ons >> e) & 1u) ? (uint32_t)s_onx[4 * tid + e] << 16 : 0u);
            hq_w[pos] = ws[e];
        }
        hq_n += many;
    };
    if (scan_h) {
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; ++u) scan_vector(u * NT + tid, hb[u]);
        for (uint32_t q0 = UNROLL * NT; q0 < nvec; q0 += UNROLL * NT) {
            uint4 w4[UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < UNROLL; ++u) w4[u] = w4p[umin(q0 + u * NT + tid, nvec - 1)];
#pragma unroll
            for (uint32_t u = 0; u < UNROLL; ++u) scan_vector(q0 + u * NT + tid, w4[u]);
        }
        if (hq_n) {
            queue_sync();
            human_deferred<HIST>(v, cd, sd, key, t, hq_i, hq_w, hq_n, lane, hc.minev);
            queue_sync();
        }
    }
    if (NW > 1) __syncthreads();   // every wave's queue is done before the bitmap bytes are written

    // ---- mosquitoes: death and E->I (mdemographic.cu:68-98), eight slots per u32 op ----
    // A 16-byte vector holds slots 0..7 as four u32 pairs.  The status and
    // strain fields and the death test of all eight are gathered into one
    // word each, slot e at bit position 16*(e&1) + 2*(e>>1), and tested
    // together.
    uint32_t im01 = 0, im23 = 0, alive = 0;
    const uint32_t nwords = (cd.cap + 31) >> 5;
    {
        // per 16-bit half x = death day << 4 | strain << 2 | status:
        // ((t mod 4096) << 4 | 15) - x = ((t - death day) mod 4096) << 4 | (15 -
        // the low nibble) mod 2^16, whose bit 15 is clear iff (t - death day)
        // mod 4096 < 2048, i.e. the slot's death day is today or past
        const uint32_t Th = ((((t & 0xFFFu) << 4) | 0xFu) * 0x00010001u);
        const uint32_t T16 = (t & 0xFFFFu) * 0x00010001u;   // t mod 2^16 in both halves
        // the valid slots of a vector: all eight below mfull, else those below cap
        const uint32_t mfull = cd.cap >> 3, pb = (1u << (cd.cap & 7u)) - 1u;
        const uint32_t mpart = (pb & 0x55u) | ((pb & 0xAAu) << 15);
        uint32_t ic[4] = {0u, 0u, 0u, 0u};   // infectious slots per serotype
        for (uint32_t q0 = 0; q0 < mlim; q0 += UNROLL * NT) {
            uint4 raw[UNROLL];
            if (q0 == 0) {
#pragma unroll
                for (uint32_t u = 0; u < UNROLL; ++u) {
                    const uint32_t q = u * NT + tid;
                    if (u < EARLY_M) raw[u] = mb[u < EARLY_M ? u : 0];
                    else if (q < mlim) raw[u] = m4p[q];
                }
            } else {
#pragma unroll
                for (uint32_t u = 0; u < UNROLL; ++u) {
                    const uint32_t q = q0 + u * NT + tid;
                    if (q < mlim) raw[u] = m4p[q];
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < UNROLL; ++u) {
                const uint32_t q = q0 + u * NT + tid;
                if (q >= mlim) continue;
                const uint32_t rw[4] = {raw[u].x, raw[u].y, raw[u].z, raw[u].w};
                uint32_t st = 0, ndy = 0;
#pragma unroll
                for (uint32_t p = 0; p < 4; ++p) {
                    const uint32_t x = rw[p];
                    st |= (x & 0x00030003u) << (2 * p);
                    ndy |= (sub_u16x2(Th, x) >> (15 - 2 * p)) & (0x00010001u << (2 * p));
                }
                const uint32_t valid = q < mfull ? 0x00550055u : mpart;
                const uint32_t b0 = st & 0x00550055u, b1 = (st >> 1) & 0x00550055u;
                const uint32_t empty = b0 & b1;                       // status 3
                const uint32_t died = ~ndy & ~empty & valid;
                const uint32_t dead = (empty | died) & valid;
                const uint32_t live = valid & ~dead;
                uint32_t em = b0 & ~b1 & live;                        // status 1
                uint32_t im = b1 & ~b0 & live;                        // status 2
                if (em) {  // E -> I when the extrinsic incubation is over; one 16-byte read of the next days
                    const uint4 nx4 = *reinterpret_cast<const uint4*>(v.m_next + cd.m_off + 8 * q);
                    const uint32_t nw[4] = {nx4.x, nx4.y, nx4.z, nx4.w};
                    // due: (int16)(t - next) > 0, i.e. bit 15 clear in both d = t - next and
                    // d - 1 (per 16-bit half, mod 2^16), for all eight slots at once
                    uint32_t due = 0;
#pragma unroll
                    for (uint32_t p = 0; p < 4; ++p) {
                        const uint32_t d = sub_u16x2(T16, nw[p]), d1 = sub_u16x2(d, 0x00010001u);
                        due |= (~(d | d1) >> (15 - 2 * p)) & (0x00010001u << (2 * p));
                    }
                    uint32_t go = em & due;
                    im |= go;
                    while (go) {
                        const uint32_t b = __ffs(go) - 1;
                        go &= go - 1;
                        const uint32_t e = b < 16 ? b : b - 15;           // slot within the vector
                        const uint32_t w = (rw[e >> 1] >> ((e & 1u) * 16u)) & 0xFFFFu;
                        v.m_word[cd.m_off + 8 * q + e] = (uint16_t)((w & ~3u) | 2u);
                    }
                }
                alive += __popc(live);
                if (im) {
                    uint32_t sn = 0;
#pragma unroll
                    for (uint32_t p = 0; p < 4; ++p) sn |= ((rw[p] >> 2) & 0x00030003u) << (2 * p);
                    const uint32_t sa = sn & 0x00550055u, sb = (sn >> 1) & 0x00550055u;
                    ic[0] += __popc(im & ~sa & ~sb);
                    ic[1] += __popc(im & sa & ~sb);
                    ic[2] += __popc(im & ~sa & sb);
                    ic[3] += __popc(im & sa & sb);
                }
                s_dead8[q] = (uint8_t)((dead & 0x55u) | ((dead >> 15) & 0xAAu));
                s_died8[q] = (uint8_t)((died & 0x55u) | ((died >> 15) & 0xAAu));
            }
        }
        im01 = ic[0] | (ic[1] << 16);   // a lane's counts are below 2^16
        im23 = ic[2] | (ic[3] << 16);
        // the empty slots not read (all dead, none died today) and the bytes
        // past the capacity in the last bitmap word (a few iterations at most)
#pragma unroll 1
        for (uint32_t q = mlim + tid; q < 4 * nwords; q += NT) {
            const uint32_t left = q < mvec ? cd.cap - 8 * q : 0u;
            s_dead8[q] = (uint8_t)(left >= 8 ? 0xFFu : ((1u << left) - 1u));
            s_died8[q] = 0;
        }
    }
    // the wave's human counts (pairs of 16-bit serotype counts; a wave's
    // counts are below 2^16) and earliest event day
    uint32_t ih01, ih23, inc01, inc23, minev;
    if (scan_h) {
        ih01 = wave_sum(hc.ih01); ih23 = wave_sum(hc.ih23);
        inc01 = wave_sum(hc.inc01); inc23 = wave_sum(hc.inc23);
        minev = wave_min(umin(hc.minev, minw >> 8));
    } else {
        ih01 = wv == 0 ? hs.y : 0u;
        ih23 = wv == 0 ? hs.z : 0u;
        inc01 = inc23 = 0;
        minev = hs.x;
    }
    im01 = wave_sum(im01);
    im23 = wave_sum(im23);
    uint32_t alive_c = wave_sum(alive);
    if (NW > 1) {   // the waves' counts through LDS (16-bit pair halves cannot carry: each is below 2^16)
        if (lane == 0) {
            s_red[wv][0] = ih01; s_red[wv][1] = ih23; s_red[wv][2] = inc01; s_red[wv][3] = inc23;
            s_red[wv][4] = im01; s_red[wv][5] = im23; s_red[wv][6] = minev; s_red[wv][7] = alive_c;
        }
        __syncthreads();
        ih01 = ih23 = inc01 = inc23 = im01 = im23 = alive_c = 0;
        minev = 0xFFFFFFFFu;
#pragma unroll
        for (uint32_t w = 0; w < NW; ++w) {
            ih01 += s_red[w][0]; ih23 += s_red[w][1]; inc01 += s_red[w][2]; inc23 += s_red[w][3];
            im01 += s_red[w][4]; im23 += s_red[w][5]; alive_c += s_red[w][7];
            minev = s_red[w][6] < minev ? s_red[w][6] : minev;
        }
    }
    const uint32_t dead_c = cd.cap - alive_c;

    // ---- emergence into the lowest-index dead slots (mdemographic.cu:104-142) ----
    uint32_t births;
    {
        const U4 r = draw(key, cd.c, t, P_M_TARGET, 0);
        const uint32_t target = stoch_round(dens_next * cd.base, r.x);
        const uint32_t b = target > alive_c ? target - alive_c : 0u;
        births = b < dead_c ? b : dead_c;
    }
    __syncthreads();
    uint32_t dw[kWordsPerLane], kw[kWordsPerLane], wpre[kWordsPerLane];
    {
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t k = 0; k < kWordsPerLane; ++k) {
            const uint32_t w = kWordsPerLane * tid + k;
            dw[k] = w < nwords ? s_dead[w] : 0u;
            kw[k] = w < nwords ? s_died[w] : 0u;
            cnt += __popc(dw[k]);
        }
        uint32_t total;
        uint32_t pre = wave_exclusive_scan(cnt, total);
        if (NW > 1) {   // plus the dead slots of the waves before
            if (lane == 0) s_wtot[wv] = total;
            __syncthreads();
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w) pre += w < wv ? s_wtot[w] : 0u;
        }
        // every thread's four prefix entries are written (words past nwords
        // hold the total, above every rank), so a thread's four can be read
        // as one vector
#pragma unroll
        for (uint32_t k = 0; k < kWordsPerLane; ++k) {
            wpre[k] = pre;
            s_dpre[kWordsPerLane * tid + k] = pre;
            pre += __popc(dw[k]);
        }
        if (NW == 1) {
            // Owners of the first 64 birth quads: quad g starts at rank 4g, in
            // the words of the lane whose ranks [wpre[0], pre) hold it.  Lane L
            // marks s_own[g] = L + 1 at the first quad that starts in its words
            // (if any); quad g's owner is the last mark at or before g, since
            // the owners ascend with g.  s_died is free: kw[] holds its words.
            s_died[lane] = 0u;
            const uint32_t qa = (wpre[0] + 3) >> 2;
            if (4 * qa < pre && qa < kWave) s_died[qa] = lane + 1;
        }
    }
    __syncthreads();
    // births i = 0..births-1 fill the i-th dead slot; births 4g..4g+3 share one Philox block
    for (uint32_t g4 = tid; 4 * g4 < births; g4 += NT) {
        const U4 r = draw(key, cd.c, t, P_M_BIRTH, g4);
        const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
        uint32_t w, pos;
        if (NW == 1 && g4 < kWave) {   // the first 64 quads (one per lane): from the owner marks
            const uint64_t marked = __builtin_amdgcn_ballot_w64(s_died[g4] != 0u);
            const uint64_t upto = marked & (~0ull >> (63 - g4));   // marks at quads <= g4 (quad 0 is marked)
            const uint32_t j = 63u - (uint32_t)__clzll(upto);
            const uint32_t o = s_died[j] - 1;                        // the owner lane
            const uint4 p = reinterpret_cast<const uint4*>(s_dpre)[o];
            const uint32_t rk = 4 * g4;
            const uint32_t k = (p.y <= rk ? 1u : 0u) + (p.z <= rk ? 1u : 0u) + (p.w <= rk ? 1u : 0u);
            const uint32_t pk = k == 0 ? p.x : k == 1 ? p.y : k == 2 ? p.z : p.w;
            w = kWordsPerLane * o + k;
            pos = select_in_word(s_dead[w], rk - pk);
        } else {
            select_bit(4 * g4, s_dead, s_dpre, nwords, w, pos);
        }
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            if (4 * g4 + j >= births) break;
            if (j > 0) {  // next dead slot
                uint32_t bits = s_dead[w] & ~((2u << pos) - 1u);
                while (!bits) bits = s_dead[++w];
                pos = __ffs(bits) - 1;
            }
            const uint32_t da = death_age_guided(s_dtab, s_mguide, sd.death_m_max, (rr[j] >> 8) + 1);
            v.m_word[cd.m_off + 32u * w + pos] = (uint16_t)(((t + da) & 0xFFFu) << 4);
        }
    }
    // slots that died today and stay empty; the alive bitmap of the day
    uint32_t top = 0;
    uint32_t awv[kWordsPerLane] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (uint32_t k = 0; k < kWordsPerLane; ++k) {
        const uint32_t w = kWordsPerLane * tid + k;
        if (w >= nwords) continue;
        const uint32_t d = dw[k], c = __popc(d);
        uint32_t reborn;
        if (wpre[k] + c <= births) reborn = d;
        else if (wpre[k] >= births) reborn = 0u;
        else reborn = d & ((1u << select_in_word(d, births - wpre[k])) - 1u);
        uint32_t stay = kw[k] & ~reborn;
        while (stay) {
            const uint32_t e = __ffs(stay) - 1;
            stay &= stay - 1;
            v.m_word[cd.m_off + 32u * w + e] = 3;
        }
        const uint32_t aw = ~(d & ~reborn) & valid_bits(w, cd.cap);
        awv[k] = aw;
        if (aw) top = 32u * w + 32u - (uint32_t)__clz((int)aw);   // one past the highest alive slot
    }
    // the lane's four bitmap words in one store (communities start on 128-slot
    // boundaries and own whole multiples of 128 slots)
    if (kWordsPerLane * tid < nwords)
        reinterpret_cast<uint4*>(v.alive + (cd.m_off >> 5))[tid] = make_uint4(awv[0], awv[1], awv[2], awv[3]);
    top = wave_max(top);
    if (NW > 1) {
        if (lane == 0) s_wtop[wv] = top;
        __syncthreads();
#pragma unroll
        for (uint32_t w = 0; w < NW; ++w) top = s_wtop[w] > top ? s_wtop[w] : top;
    }
    if (tid == 0)   // the day's summary: human part (unchanged when not scanned) and mosquito limit
        reinterpret_cast<uint4*>(v.hsum)[g] = make_uint4(minev, ih01, ih23, ((top + 7) >> 3) + 1u);
    const uint32_t Mc = alive_c + births;

    // ---- event counts of this community as a source (nepidemic.cu:49-136) ----
    if (tid < 8) {
        const uint32_t dir = tid >> 2, s = tid & 3u;
        uint32_t n = 0;
        unsigned long long fx = 0;
        if (s < S) {
            const uint32_t I = dir == DIR_HM ? unpack_pair(ih01, ih23, s) : unpack_pair(im01, im23, s);
            if (I > 0) {
                const DayParams dp = day_params(sd, v.dens, t);
                const float lam = dir == DIR_HM ? (float)Mc / (float)cd.n * dp.beta_hm * (float)I
                                                : dp.beta_mh * (float)I;
                const bool staying = (cd.flags & kCommStay) != 0;
                uint32_t stay = 0;
                if (sd.count_law == ARBO_COUNT_STOCH_ROUND) {   // nepidemic.cu:67-104
                    const U4 r = draw(key, cd.c, t, ctr_tag(P_NSR, dir, s, 0), 0);
                    uint32_t all = stoch_round(lam, r.x);
                    if (staying) stay = stoch_round((float)all * sd.no_distance, r.y);
                    all -= stay;
                    const uint32_t ld = stoch_round((float)all * sd.omega, r.z);
                    n = all - ld;
                    fx = ld;   // an event count: the long-distance kernel sends exactly these
                } else {
                    const float go = staying ? lam * (1.0f - sd.no_distance) : lam;
                    const float mu_loc = go * (1.0f - sd.omega);
                    const float mu_ld = go * sd.omega;
                    fx = (unsigned long long)floor((double)mu_ld * 16777216.0);
                    const Stream st{key, cd.c, t, ctr_tag(P_NLOC, dir, s, 0)};
                    n = poisson((double)mu_loc, st);
                    if (staying) {
                        const Stream ss{key, cd.c, t, ctr_tag(P_NSTAY, dir, s, 0)};
                        stay = poisson((double)(lam * sd.no_distance), ss);
                    }
                }
                // staying events join the count of the source's own column
                if (stay) atomicAdd(&s_hist[tid * 9 + ((cd.flags >> kCommSelfShift) & 15u)], stay);
            }
        }
        s_n[tid] = n;
        // summed over the simulation by k_day_reduce (no same-address atomics
        // across the simulation's communities)
        reinterpret_cast<unsigned long long*>(v.part + (size_t)g * kPartWords + 8)[tid] = fx;
    }
    __syncthreads();
    // Destinations: event i of source segment k takes word i%4 of Philox
    // block i/4 (one block per lane and quad of events); the product
    // word * row length gives the column (high half) and the alias test
    // uniform (low half).
    const uint32_t rl = cd.row_len;
    uint32_t qtot = 0;   // quads of the 8 segments (the segment of a quad is found by walking s_n)
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) qtot += (s_n[k] + 3) >> 2;
    for (uint32_t gq = tid; gq < qtot; gq += NT) {
        uint32_t k = 0, base = 0, nq = (s_n[0] + 3) >> 2;
        while (gq >= base + nq) {   // ends: gq < qtot
            base += nq;
            nq = (s_n[++k] + 3) >> 2;
        }
        const uint32_t j = gq - base;
        const uint32_t n = s_n[k];
        const U4 r = draw(key, cd.c, t, ctr_tag(P_DEST, k >> 2, k & 3u, 0), j);
        const uint32_t w4[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            if (4 * j + e >= n) break;
            const uint64_t prod = (uint64_t)w4[e] * rl;
            uint32_t col = (uint32_t)(prod >> 32);
            const uint32_t en = s_row[col];
            if ((((uint32_t)prod) >> 8) >= (en & 0x1FFFFFFu)) col = en >> 27;
            atomicAdd(&s_hist[k * 9 + col], 1u);
        }
    }
    __syncthreads();
    // push the counts into each destination's inbox: 8 contiguous words per (destination, in-edge)
    for (uint32_t i0 = 0; wv == 0 && i0 < 8 * rl; i0 += kWave) {   // wave 0, every lane shuffles
        const uint32_t i = i0 + lane, col = i >> 3, k = i & 7u;
        const uint32_t dst = __shfl(inboxent, col < rl ? col : 0u, kWave);
        if (i < 8 * rl) {
            const uint32_t n = s_hist[k * 9 + col];
            v.cnt[(size_t)dst * 8 + k] = n;
            if (n) v.active[dst / kInboxEdges] = 1u;   // the destination has events today
        }
    }

    // ---- series row t+1 (datacollect.cu:79-189): this community's part ----
    const SeriesIdx si = series_index(sd.series_off, v.max_days, S);
    const uint32_t row = t + 1;
    if (tid == 0)
        reinterpret_cast<uint4*>(v.part + (size_t)g * kPartWords)[0] = make_uint4(Mc, cd.n, ih01, ih23);
    else if (tid == 1)   // with the census vectors this scan read (the exact skips leave the rest)
        reinterpret_cast<uint4*>(v.part + (size_t)g * kPartWords + 4)[0] =
            make_uint4(inc01, inc23, scan_h ? nvec : 0u, mlim);
    if (tid < S) {
        const uint32_t ih = unpack_pair(ih01, ih23, tid);
        if (cd.c == 0) v.series[si.n_one + (size_t)row * S + tid] = ih;
        if (HIST && v.spatial) v.spatial[sd.spatial_off + ((uint64_t)row * S + tid) * sd.C + cd.c] = ih;
    }
}

// ---------------------------------------------------------------------------
// Day kernel B: application of the day's events at one destination.
//
// The events of a destination and direction form segments in a fixed order:
// for each serotype of the rotation, one segment per in-edge (ascending
// source) and then the long-distance arrivals.  Event k of a segment draws its
// target from word k%4 of the segment's Philox block k/4, so a lane handles a
// quad of consecutive events with one block.  Quads are applied in chunks of
// kChunkQuads with an LDS claim-min table.

constexpr uint32_t kChunkQuads = kChunk / 4;   // 64 quads = up to 256 events
constexpr uint32_t kQuadsPerLane = kChunkQuads / kWave;  // 1

struct Quad {
    uint32_t s;       // serotype
    uint32_t x0, x2;  // Philox counter words of the segment
    uint32_t j;       // quad index within the segment (Philox block)
    uint32_t rank;    // rank of the quad's first event within the direction
    uint32_t count;   // events in the quad (1..4)
};

__device__ __forceinline__ Quad decode_quad(uint32_t gq, const uint32_t* seg, const uint32_t* qseg, uint32_t nseg,
                                            uint32_t in_len, uint32_t rot, uint32_t S, uint32_t dir,
                                            const uint32_t* src_local, const uint32_t* slot, uint32_t self_c) {
    uint32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (qseg[mid] <= gq) lo = mid; else hi = mid - 1;
    }
    const uint32_t pos = lo / (in_len + 1), e = lo % (in_len + 1);
    Quad q;
    q.s = (rot + pos) % S;
    q.j = gq - qseg[lo];
    q.rank = seg[lo] + 4 * q.j;
    const uint32_t left = seg[lo + 1] - q.rank;
    q.count = left < 4 ? left : 4;
    if (e < in_len) {
        q.x0 = src_local[e];
        q.x2 = ctr_tag(P_TGT, dir, q.s, slot[e]);
    } else {
        q.x0 = self_c;
        q.x2 = ctr_tag(P_TGT_LD, dir, q.s, 0);
    }
    return q;
}

__device__ __forceinline__ uint32_t claim_slot(uint32_t key) { return (key * 2654435761u) >> 23; }

// Record event `rank` of the chunk on member `key`; the table keeps the
// smallest rank per member (open addressing, at most kChunk keys in kHash).
__device__ __forceinline__ void claim(uint32_t* tab, uint32_t key, uint32_t rank) {
    const uint32_t packed = (rank << kKeyBits) | key;
    uint32_t h = claim_slot(key);
    while (true) {
        uint32_t cur = tab[h];
        if (cur == kEmpty) {
            cur = atomicCAS(&tab[h], kEmpty, packed);
            if (cur == kEmpty) return;
        }
        if ((cur & kKeyMask) == key) {
            atomicMin(&tab[h], packed);
            return;
        }
        h = (h + 1) & (kHash - 1);
    }
}

__device__ __forceinline__ uint32_t claimed_rank(const uint32_t* tab, uint32_t key) {
    uint32_t h = claim_slot(key);
    while (tab[h] == kEmpty || (tab[h] & kKeyMask) != key) h = (h + 1) & (kHash - 1);
    return tab[h] >> kKeyBits;
}

constexpr uint32_t kEligible = 1u << 31;
constexpr uint32_t kHumanKey = ARBO_MAX_COMMUNITY_SLOTS;   // claim keys of humans start here
static_assert(ARBO_MAX_COMMUNITY_SLOTS + ARBO_MAX_COMMUNITY_HUMANS <= (1u << kKeyBits),
              "mosquito and human claim keys fit the key field");

// MW: bitmap words of a community in units of kWaveWords (1, or kMaxWaves
// for batches with communities above 8,192 slots): the alive-rank index is
// built in MW passes of four words per lane.
template <bool XCD, bool HIST, uint32_t MW>
__global__ __launch_bounds__(kWave) void k_day_events(EngineView v) {
    const uint32_t g = XCD ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x, lane = threadIdx.x;
    const uint32_t t = *v.day + v.day_offset;
    const CommDesc cd = v.comm[g];
    const uint32_t act = v.active[g];
    if (!act && !(cd.flags & kCommIntro)) return;   // no events arrive and no introductions
    if (act && lane == 0) v.active[g] = 0u;         // consumed
    const SimDesc sd = v.sims[cd.sim];
    const Key key{sd.k0, sd.k1};
    const uint32_t S = sd.S;
    const uint32_t in_len = cd.in_len;
    const uint32_t nseg = S * (in_len + 1);
    const float eip = day_params(sd, v.dens, t).eip;   // extrinsic incubation of today's infections

    __shared__ uint32_t s_src[9], s_slot[9];
    __shared__ uint32_t s_ecnt[8][9];
    __shared__ uint32_t s_nld[8];
    __shared__ uint32_t s_rot[2];
    __shared__ uint32_t s_seg[2][4 * 10 + 1];
    __shared__ uint32_t s_qseg[2][4 * 10 + 1];
    __shared__ uint32_t s_word[kWaveWords * MW];
    __shared__ uint16_t s_wpre[kWaveWords * MW];   // alive slots before each word (below 2^16)
    __shared__ uint16_t s_sel[kWaveWords * MW];    // word holding alive rank 32k, k = 0..(Md-1)/32
    __shared__ uint32_t s_tab[kHash];

    if (lane < in_len) {
        s_src[lane] = v.in_edge[cd.in_off + lane];
        s_slot[lane] = v.in_slot[cd.in_off + lane];
    }
    uint32_t tot_dir[2] = {0, 0};
    if (lane < 8) {  // long-distance arrivals of the day, consumed
        const uint32_t n = v.nld[(size_t)g * 8 + lane];
        s_nld[lane] = n;
        if (n) v.nld[(size_t)g * 8 + lane] = 0;
        tot_dir[lane >> 2] += n;
    }
    __syncthreads();
    for (uint32_t i = lane; i < 8 * in_len; i += kWave) {
        const uint32_t k = i / in_len, e = i % in_len;      // k = dir*4 + s
        const uint32_t n = ((k & 3u) < S) ? v.cnt[((size_t)g * kInboxEdges + e) * 8 + k] : 0u;
        s_ecnt[k][e] = n;
        tot_dir[k >> 2] += n;
    }
    tot_dir[0] = wave_sum(tot_dir[0]);
    tot_dir[1] = wave_sum(tot_dir[1]);
    if (tot_dir[0] == 0 && tot_dir[1] == 0 && !(sd.intro > 0.0f)) return;   // nothing arrives here today
    if (lane < 2) s_rot[lane] = mulhi(draw(key, cd.c, t, ctr_tag(P_ROT, lane, 0, 0), 0).x, S);
    __syncthreads();
    // segment starts in events and in quads (one lane per segment, wave scans)
#pragma unroll
    for (uint32_t dir = 0; dir < 2; ++dir) {
        uint32_t cnt = 0;
        if (lane < nseg) {
            const uint32_t pos = lane / (in_len + 1), e = lane % (in_len + 1);
            const uint32_t s = (s_rot[dir] + pos) % S;
            cnt = (e < in_len) ? s_ecnt[dir * 4 + s][e] : s_nld[dir * 4 + s];
        }
        uint32_t total, qtotal;
        const uint32_t start = wave_exclusive_scan(cnt, total);
        const uint32_t qstart = wave_exclusive_scan((cnt + 3) >> 2, qtotal);
        if (lane < nseg) { s_seg[dir][lane] = start; s_qseg[dir][lane] = qstart; }
        if (lane == 0) { s_seg[dir][nseg] = total; s_qseg[dir][nseg] = qtotal; }
    }
    if (lane < in_len) s_src[lane] -= sd.comm_off;   // simulation-local source ids
    __syncthreads();

    // ---- onto mosquitoes: uniform over alive slots, success iff susceptible ----
    const uint32_t tq_m = s_qseg[DIR_HM][nseg], tq_h = s_qseg[DIR_MH][nseg];
    const uint32_t nwords = (cd.cap + 31) >> 5;
    uint32_t Md = 0;
    for (uint32_t c0 = 0; tq_m > 0 && c0 < MW && kWaveWords * c0 < nwords; ++c0) {   // wave-uniform
        const uint32_t i4 = kWave * c0 + lane;   // the lane's 16-byte group of words
        uint32_t wv[kWordsPerLane] = {0u, 0u, 0u, 0u}, c = 0;
        if (kWordsPerLane * i4 < nwords) {   // one 16-byte load (words past the capacity are zero)
            const uint4 a4 = reinterpret_cast<const uint4*>(v.alive + (cd.m_off >> 5))[i4];
            wv[0] = a4.x; wv[1] = a4.y; wv[2] = a4.z; wv[3] = a4.w;
        }
#pragma unroll
        for (uint32_t k = 0; k < kWordsPerLane; ++k) c += __popc(wv[k]);
        uint32_t chunk;
        uint32_t pre = Md + wave_exclusive_scan(c, chunk);
#pragma unroll
        for (uint32_t k = 0; k < kWordsPerLane; ++k) {
            const uint32_t w = kWordsPerLane * i4 + k;
            const uint32_t cw = __popc(wv[k]);
            if (w < nwords) {
                s_word[w] = wv[k];
                s_wpre[w] = (uint16_t)pre;
                const uint32_t m = (pre + 31u) & ~31u;   // the multiple of 32 among this word's ranks, if any
                if (m < pre + cw) s_sel[m >> 5] = (uint16_t)w;
            }
            pre += cw;
        }
        Md += chunk;
    }
    const uint32_t mq = Md > 0 ? tq_m : 0u;   // with no alive mosquito every event onto one fails
    // ---- onto humans: uniform target, success iff S and never infected by s ----
    const uint32_t N = cd.n;
    bool infected = false;

    // One claim sequence: mosquito quads [m_lo, m_hi) of their direction's
    // list, then human quads [h_lo, h_hi) of theirs.  Mosquito slots and
    // humans (keyed from ARBO_MAX_COMMUNITY_SLOTS up) never share a claim,
    // so the two directions' events can share chunks; within a direction the
    // chunk rank 4 p + e keeps the list order.
    auto apply_run = [&](uint32_t m_lo, uint32_t m_hi, uint32_t h_lo, uint32_t h_hi) {
        const uint32_t lm = m_hi - m_lo, len = lm + (h_hi - h_lo);
        for (uint32_t c0 = 0; c0 < len; c0 += kChunkQuads) {
            for (uint32_t i = lane; i < kHash; i += kWave) s_tab[i] = kEmpty;
            __syncthreads();
            uint32_t cache[kPerLane], word[kPerLane];
#pragma unroll
            for (uint32_t u = 0; u < kQuadsPerLane; ++u) {
#pragma unroll
                for (uint32_t e = 0; e < 4; ++e) cache[4 * u + e] = 0;
                const uint32_t p = c0 + u * kWave + lane;
                if (p >= len) continue;
                const bool hum = p >= lm;
                const uint32_t dir = hum ? DIR_MH : DIR_HM;
                const Quad q = decode_quad(hum ? h_lo + (p - lm) : m_lo + p, s_seg[dir], s_qseg[dir], nseg, in_len,
                                           s_rot[dir], S, dir, s_src, s_slot, cd.c);
                const U4 r = draw(key, q.x0, t, q.x2, q.j);
                const uint32_t w4[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
                for (uint32_t e = 0; e < 4; ++e) {
                    if (e >= q.count) break;
                    const uint32_t rank = 4 * (u * kWave + lane) + e;
                    if (hum) {
                        const uint32_t h = mulhi(w4[e], N);
                        const uint32_t w = v.h_word[cd.h_off + h];
                        if ((w & 3u) == 0 && !((w >> (4 + q.s)) & 1u)) {
                            claim(s_tab, kHumanKey + h, rank);
                            cache[4 * u + e] = kEligible | (q.s << kKeyBits) | (kHumanKey + h);
                            word[4 * u + e] = w;
                        }
                    } else {
                        const uint32_t rk = mulhi(w4[e], Md);   // alive rank of the target
                        uint32_t w = s_sel[rk >> 5];
                        while (w + 1 < nwords && s_wpre[w + 1] <= rk) ++w;
                        const uint32_t slot = 32u * w + select_in_word(s_word[w], rk - s_wpre[w]);
                        const uint32_t mw = v.m_word[cd.m_off + slot];
                        if ((mw & 3u) == 0) {
                            claim(s_tab, slot, rank);
                            cache[4 * u + e] = kEligible | (q.s << kKeyBits) | slot;
                            word[4 * u + e] = mw;
                        }
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t i = 0; i < kPerLane; ++i) {
                if (!(cache[i] & kEligible)) continue;
                const uint32_t k = cache[i] & kKeyMask;
                if (claimed_rank(s_tab, k) != 4 * ((i >> 2) * kWave + lane) + (i & 3u)) continue;
                const uint32_t s = (cache[i] >> kKeyBits) & 3u;
                const uint32_t w = word[i];   // as read by the claim pass: only the winner writes it
                if (k >= kHumanKey) {
                    const uint32_t h = k - kHumanKey;
                    const uint64_t gi = cd.h_off + h;
                    v.h_word[gi] = (w & ~0xFu) | (1u << (4 + s)) | (s << 2) | 1u;
                    if (HIST) {
                        v.hist[4 * gi + s] = (int32_t)t - v.h_birth[gi];   // age at infection
                        if (v.aoi_mark && t >= v.aoi_from) v.aoi_mark[gi] = 1;
                    }
                    const U4 r2 = draw(key, cd.h_base + h, t, P_H_IIP, 0);
                    v.h_next[gi] = (uint16_t)(t + sd.iip_min + mulhi(r2.x, sd.iip_span));
                    infected = true;
                } else {
                    const uint64_t gm = cd.m_off + k;
                    v.m_word[gm] = (uint16_t)((w & ~0xFu) | (s << 2) | 1u);
                    const U4 r2 = draw(key, cd.m_base + k, t, P_M_EIP, 0);
                    v.m_next[gm] = (uint16_t)(t + stoch_round(eip, r2.x));
                }
            }
            __syncthreads();
        }
    };

    if (!(sd.intro > 0.0f)) {
        apply_run(0, mq, 0, tq_h);   // mosquitoes, then every serotype's human events
    } else {
        // Introductions follow each serotype's human events, so the serotype
        // positions run one by one after the mosquitoes.
        apply_run(0, mq, 0, 0);
        for (uint32_t pos = 0; pos < S; ++pos) {
            apply_run(0, 0, s_qseg[DIR_MH][pos * (in_len + 1)], s_qseg[DIR_MH][(pos + 1) * (in_len + 1)]);
            const uint32_t s = (s_rot[DIR_MH] + pos) % S;
            if (lane == 0) {  // external introductions (nepidemic.cu:261-309)
                const uint32_t tagi = ctr_tag(P_INTRO, 0, s, 0);
                uint32_t x = stoch_round(sd.intro / (float)sd.C, draw(key, cd.c, t, tagi, 0).x);
                for (uint32_t att = 0; x > 0 && att < 1000; ++att) {
                    const uint32_t h = mulhi(draw(key, cd.c, t, tagi, 1 + att).x, N);
                    const uint64_t gi = cd.h_off + h;
                    const uint32_t w = v.h_word[gi];
                    if ((w & 3u) == 0 && !((w >> (4 + s)) & 1u)) {
                        v.h_word[gi] = (w & ~0xFu) | (1u << (4 + s)) | (s << 2) | 1u;
                        if (HIST) v.hist[4 * gi + s] = (int32_t)t - v.h_birth[gi];
                        const U4 r2 = draw(key, cd.h_base + h, t, P_H_IIP, 0);
                        v.h_next[gi] = (uint16_t)(t + sd.iip_min + mulhi(r2.x, sd.iip_span));
                        infected = true;
                        x--;
                    }
                }
            }
            __syncthreads();
        }
    }
    // a human exposed today turns infectious on day t + iip + 1 at the earliest
    if (__any(infected) && lane == 0) {
        const uint32_t ev = t + sd.iip_min + 1;
        if (ev < v.hsum[4 * (size_t)g]) v.hsum[4 * (size_t)g] = ev;
    }
}

// ---------------------------------------------------------------------------
// Long-distance events of the day (nepidemic.cu:118-126, mepidemic.cu:117-122:
// a long-distance event lands on a uniform random community).  Per simulation,
// direction and serotype the events are L ~ Poisson(summed long-distance rate
// of all sources); event l lands on community mulhi(Philox(l), C).  Only the
// number landing on each destination matters (their targets are drawn at the
// destination), so the destinations are counted with integer atomics.
// Part `part` of `split` of simulation `sim`'s long-distance events of day t
// (all 256 threads of the workgroup call it).
__device__ __forceinline__ void longdist_part(const EngineView& v, uint32_t sim, uint32_t t, uint32_t part,
                                              uint32_t split, uint32_t tid) {
    const SimDesc sd = v.sims[sim];
    const Key key{sd.k0, sd.k1};
    __shared__ uint32_t s_L[8];
    if (tid < 8) {
        const uint32_t dir = tid >> 2, s = tid & 3u;
        uint32_t L = 0;
        if (s < sd.S) {
            const unsigned long long fx = v.lam[((size_t)(t & 1u) * v.n_sims + sim) * 8 + tid];   // summed by k_day_reduce
            if (sd.count_law == ARBO_COUNT_STOCH_ROUND) {
                L = (uint32_t)fx;   // the sources' own long-distance counts
            } else {
                const Stream st{key, INIT_DAY, t, ctr_tag(P_NLD, dir, s, 0)};
                L = poisson((double)fx * 0x1p-24, st);
            }
        }
        s_L[tid] = L;
    }
    __syncthreads();
    uint32_t pre[9];
    pre[0] = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) pre[k + 1] = pre[k] + s_L[k];
    for (uint32_t l = part * 256 + tid; l < pre[8]; l += split * 256) {
        uint32_t k = 0;
#pragma unroll
        for (uint32_t q = 1; q < 8; ++q) k += (l >= pre[q]) ? 1u : 0u;
        const uint32_t idx = l - pre[k];
        const uint32_t d = mulhi(draw(key, idx, t, ctr_tag(P_LD_DEST, k >> 2, k & 3u, 0), 0).x, sd.C);
        atomicAdd(&v.nld[(size_t)(sd.comm_off + d) * 8 + k], 1u);
        v.active[sd.comm_off + d] = 1u;
    }
    // the long-distance sums of the next day start at zero
    if (part == 0 && tid < 8) v.lam[((size_t)((t + 1) & 1u) * v.n_sims + sim) * 8 + tid] = 0ull;
}

__global__ __launch_bounds__(256) void k_day_longdist(EngineView v, uint32_t split) {
    // split workgroups per simulation
    const uint32_t sim = blockIdx.x / split, part = blockIdx.x % split;
    longdist_part(v, sim, *v.day + v.day_offset, part, split, threadIdx.x);
}

// Sums the day's per-community partials of a simulation: `split`
// workgroups of 256 threads (about one community per thread, at most 64
// per simulation and about 2,048 in all) each take a share of its
// communities and add
// their sums to the series row t + 1 (datacollect.cu:79-189, zero until
// then) and to the long-distance totals the long-distance kernel reads
// (zeroed by the previous day's long-distance kernel), `split` integer
// atomics per word, so the order does not matter.
__global__ __launch_bounds__(256) void k_day_reduce(EngineView v, uint32_t split) {
    const uint32_t sim = blockIdx.x / split, part = blockIdx.x % split;
    const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    const uint32_t t = *v.day + v.day_offset;
    const SimDesc& sd = v.sims[sim];
    __shared__ unsigned long long s_fx[8];
    __shared__ uint32_t s_cnt[4][10];   // per wave: humans, mosquitoes, infectious[4], onsets[4]
    __shared__ unsigned long long s_scan[2];
    if (tid < 8) s_fx[tid] = 0ull;
    if (tid < 2) s_scan[tid] = 0ull;
    __syncthreads();
    uint32_t nh = 0, nm = 0, inf[4] = {0, 0, 0, 0}, inc[4] = {0, 0, 0, 0};
    uint32_t sv_h = 0, sv_m = 0;   // census vectors scanned (fits 32 bits: at most 64 communities per thread)
    unsigned long long fx[8] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
    for (uint32_t c = part * 256 + tid; c < sd.C; c += split * 256) {
        const uint32_t* pw = v.part + (size_t)(sd.comm_off + c) * kPartWords;
        const uint4 a = reinterpret_cast<const uint4*>(pw)[0];
        const uint4 b = reinterpret_cast<const uint4*>(pw + 4)[0];
        sv_h += b.z;
        sv_m += b.w;
        nm += a.x;
        nh += a.y;
        inf[0] += a.z & 0xFFFFu; inf[1] += a.z >> 16; inf[2] += a.w & 0xFFFFu; inf[3] += a.w >> 16;
        inc[0] += b.x & 0xFFFFu; inc[1] += b.x >> 16; inc[2] += b.y & 0xFFFFu; inc[3] += b.y >> 16;
        const unsigned long long* f = reinterpret_cast<const unsigned long long*>(pw + 8);
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) fx[k] += f[k];
    }
    const uint32_t cnt[10] = {nh, nm, inf[0], inf[1], inf[2], inf[3], inc[0], inc[1], inc[2], inc[3]};
#pragma unroll
    for (uint32_t k = 0; k < 10; ++k) {
        const uint32_t w = wave_sum(cnt[k]);
        if (lane == 0) s_cnt[wv][k] = w;
    }
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {   // butterfly over the wave, then one LDS add per wave
        unsigned long long x = fx[k];
#pragma unroll
        for (uint32_t d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
        if (lane == 0 && x) atomicAdd(&s_fx[k], x);
    }
    {
        const unsigned long long h = wave_sum(sv_h), m = wave_sum(sv_m);
        if (lane == 0) {
            atomicAdd(&s_scan[0], h);
            atomicAdd(&s_scan[1], m);
        }
    }
    __syncthreads();
    const uint32_t S = sd.S, row = t + 1;
    const SeriesIdx si = series_index(sd.series_off, v.max_days, S);
    if (tid < 10) {
        const uint32_t k = tid;
        const uint32_t x = s_cnt[0][k] + s_cnt[1][k] + s_cnt[2][k] + s_cnt[3][k];
        uint32_t* dst = nullptr;
        if (k == 0) dst = &v.series[si.n_count + row];
        else if (k == 1) dst = &v.series[si.m_count + row];
        else if (k < 6) { if (k - 2 < S) dst = &v.series[si.n_inf + (size_t)row * S + (k - 2)]; }
        else if (k - 6 < S) dst = &v.series[si.n_inc + (size_t)row * S + (k - 6)];
        if (dst && x) atomicAdd(dst, x);
    }
    if (tid >= 32 && tid < 40 && s_fx[tid - 32])
        atomicAdd(&v.lam[((size_t)(t & 1u) * v.n_sims + sim) * 8 + (tid - 32)], s_fx[tid - 32]);
    if (tid >= 64 && tid < 66 && s_scan[tid - 64]) atomicAdd(&v.scanned[2 * (size_t)sim + (tid - 64)], s_scan[tid - 64]);
}

// Age-of-infection rings (nepidemic.cu:229-243), one workgroup per
// simulation after the event kernel.  The reference appends each novel
// infection by a mosquito event to ring k (k = serotypes the human had been
// infected with before) at atomicInc positions, in whatever order threads
// arrive.  Here a day's infections enter in human order, which makes the
// rings deterministic: a counting pass gives the day's total per ring, and
// the writing pass stores only the day's last kAoiRecord records of each
// ring, so no slot is written twice in one launch.
__global__ __launch_bounds__(256) void k_day_aoi(EngineView v) {
    const uint32_t t = *v.day + v.day_offset;
    if (t < v.aoi_from) return;
    const uint32_t sim = blockIdx.x, tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    constexpr uint32_t kWaves = 256 / kWave, kChunk = 4 * 256;
    const SimDesc& sd = v.sims[sim];
    __shared__ uint32_t s_part[kWaves][4];
    __shared__ uint32_t s_tot[4], s_pos[4];
    // previous infections of the human infected today: its history words
    // other than today's (one infection per human and day)
    auto ring_of = [&](uint64_t gi) -> uint32_t {
        const int4 h = reinterpret_cast<const int4*>(v.hist)[gi];
        const uint32_t n = (uint32_t)(h.x != -1) + (uint32_t)(h.y != -1) + (uint32_t)(h.z != -1) + (uint32_t)(h.w != -1);
        return n ? n - 1u : 0u;   // n >= 1: today's word is set
    };
    // counting pass
    uint32_t c[4] = {0, 0, 0, 0};
    for (uint64_t base = sd.h_begin; base < sd.h_end; base += kChunk) {
        const uint64_t gi = base + 4 * tid;
        if (gi >= sd.h_end) continue;
        const uint32_t m = *reinterpret_cast<const uint32_t*>(v.aoi_mark + gi);
        if (!m) continue;
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            if (!((m >> (8 * e)) & 0xFFu)) continue;
            const uint32_t r = ring_of(gi + e);
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) c[k] += r == k;   // constant indices: the counters stay in registers
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t w = wave_sum(c[k]);
        if (lane == 0) s_part[wv][k] = w;
    }
    __syncthreads();
    if (tid < 4) {
        uint32_t n = 0;
        for (uint32_t w = 0; w < kWaves; ++w) n += s_part[w][tid];
        s_tot[tid] = n;
        s_pos[tid] = v.aoi_pos[4 * sim + tid];
    }
    __syncthreads();
    if ((s_tot[0] | s_tot[1] | s_tot[2] | s_tot[3]) == 0) return;
    // writing pass: record j of the day (per ring) goes to slot (pos + j) mod kAoiRecord
    uint32_t* ring = v.aoi + (uint64_t)sim * 4 * kAoiRecord;
    uint32_t done[4] = {0, 0, 0, 0};   // records of the ring in earlier chunks (same in every thread)
    for (uint64_t base = sd.h_begin; base < sd.h_end; base += kChunk) {
        const uint64_t gi = base + 4 * tid;
        uint32_t m = 0;
        if (gi < sd.h_end) m = *reinterpret_cast<const uint32_t*>(v.aoi_mark + gi);
        uint32_t k4[4] = {4, 4, 4, 4};
        uint32_t mine[4] = {0, 0, 0, 0};
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            if (!((m >> (8 * e)) & 0xFFu)) continue;
            k4[e] = ring_of(gi + e);
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) mine[k] += k4[e] == k;
        }
        uint32_t before[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {   // exclusive prefix over the workgroup
            uint32_t x = mine[k];
#pragma unroll
            for (uint32_t d = 1; d < kWave; d <<= 1) {
                const uint32_t y = __shfl_up(x, d, kWave);
                if (lane >= d) x += y;
            }
            if (lane == kWave - 1) s_part[wv][k] = x;
            before[k] = x - mine[k];
        }
        __syncthreads();
        uint32_t chunk_tot[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            uint32_t off = 0, all = 0;
            for (uint32_t w = 0; w < kWaves; ++w) {
                off += w < wv ? s_part[w][k] : 0u;
                all += s_part[w][k];
            }
            before[k] += off + done[k];
            chunk_tot[k] = all;
        }
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                if (k4[e] != k) continue;
                const uint32_t j = before[k]++;
                if (j + kAoiRecord >= s_tot[k])   // among the day's last kAoiRecord records of ring k
                    ring[k * kAoiRecord + (s_pos[k] + j) % kAoiRecord] = (uint32_t)((int32_t)t - v.h_birth[gi + e]);
            }
        }
        if (m) *reinterpret_cast<uint32_t*>(v.aoi_mark + gi) = 0u;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) done[k] += chunk_tot[k];
        __syncthreads();   // s_part is rewritten by the next chunk
    }
    if (tid < 4) v.aoi_pos[4 * sim + tid] = (s_pos[tid] + s_tot[tid]) % kAoiRecord;
}

__global__ void k_advance_day(uint32_t* day, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *day = *day + n;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---------------------------------------------------------------------------
// Launchers (host).
hipError_t launch_init(const EngineView& v, hipStream_t st) {
    hipLaunchKernelGGL(k_init, dim3(v.n_comm), dim3(kWave), 0, st, v);
    return hipGetLastError();
}

// Demography variants: (UNROLL, EARLY_M) pairs of kDemographyShapes with one
// wave per community, the kOccShapes pairs under an 8-waves-per-SIMD register
// budget, and the kWideShapes with several waves per community.
template <bool XCD, bool HIST>
static hipError_t launch_demography_xh(const EngineView& v, const Tuning& tu, hipStream_t st) {
    const size_t lds = (size_t)v.death_m_lds * 4;
    if (v.max_words > kWaveWords * tu.nw) {   // a community's bitmap needs more waves than the tuning's
        hipLaunchKernelGGL((k_day_demography<2, 2, false, HIST, 0, kMaxWaves>), dim3(v.n_comm),
                           dim3(kWave * kMaxWaves), lds, st, v);
        return hipGetLastError();
    }
    const dim3 grid(v.n_comm), block(kWave * tu.nw);
    if (tu.nw == 2) {
        if (tu.unroll == 4 && tu.early_m == 2 && tu.wps == 0)
            hipLaunchKernelGGL((k_day_demography<4, 2, XCD, HIST, 0, 2>), grid, block, lds, st, v);
        else
            return hipErrorInvalidValue;
    } else if (tu.nw == 4) {
        if (tu.unroll == 2 && tu.early_m == 2 && tu.wps == 0)
            hipLaunchKernelGGL((k_day_demography<2, 2, XCD, HIST, 0, 4>), grid, block, lds, st, v);
        else
            return hipErrorInvalidValue;
    } else if (tu.nw != 1) {
        return hipErrorInvalidValue;
    } else if (tu.wps == 8) {
        if (tu.unroll == 4 && tu.early_m == 2)
            hipLaunchKernelGGL((k_day_demography<4, 2, XCD, HIST, 8, 1>), grid, block, lds, st, v);
        else if (tu.unroll == 2 && tu.early_m == 2)
            hipLaunchKernelGGL((k_day_demography<2, 2, XCD, HIST, 8, 1>), grid, block, lds, st, v);
        else
            return hipErrorInvalidValue;
    } else if (tu.unroll == 4 && tu.early_m == 4)
        hipLaunchKernelGGL((k_day_demography<4, 4, XCD, HIST, 0, 1>), grid, block, lds, st, v);
    else if (tu.unroll == 4 && tu.early_m == 2)
        hipLaunchKernelGGL((k_day_demography<4, 2, XCD, HIST, 0, 1>), grid, block, lds, st, v);
    else if (tu.unroll == 8 && tu.early_m == 4)
        hipLaunchKernelGGL((k_day_demography<8, 4, XCD, HIST, 0, 1>), grid, block, lds, st, v);
    else if (tu.unroll == 2 && tu.early_m == 2)
        hipLaunchKernelGGL((k_day_demography<2, 2, XCD, HIST, 0, 1>), grid, block, lds, st, v);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <bool XCD>
static hipError_t launch_demography_x(const EngineView& v, const Tuning& tu, hipStream_t st) {
    return (v.hist || v.spatial) ? launch_demography_xh<XCD, true>(v, tu, st)
                                 : launch_demography_xh<XCD, false>(v, tu, st);
}

bool tuning_supported(const Tuning& tu) {
    bool shape = false;
    if (tu.nw > 1) {
        for (const auto& d : kWideShapes)
            shape = shape || (d[0] == tu.unroll && d[1] == tu.early_m && d[2] == tu.nw && tu.wps == 0);
    } else if (tu.nw != 1) {
        shape = false;
    } else if (tu.wps == 8) {
        for (const auto& d : kOccShapes) shape = shape || (d[0] == tu.unroll && d[1] == tu.early_m);
    } else if (tu.wps == 0) {
        for (const auto& d : kDemographyShapes) shape = shape || (d[0] == tu.unroll && d[1] == tu.early_m);
    }
    return shape && tu.xcd <= 3 && tu.ld_split >= 1 && tu.ld_split <= kMaxLdSplit;
}

hipError_t launch_demography(const EngineView& v, const Tuning& tu, hipStream_t st) {
    return (tu.xcd & 1u) ? launch_demography_x<true>(v, tu, st) : launch_demography_x<false>(v, tu, st);
}

hipError_t launch_events(const EngineView& v, const Tuning& tu, hipStream_t st) {
    const dim3 grid(v.n_comm), block(kWave);
    if (v.max_words > kWaveWords) {   // communities above 8,192 slots
        if (v.hist) hipLaunchKernelGGL((k_day_events<false, true, kMaxWaves>), grid, block, 0, st, v);
        else hipLaunchKernelGGL((k_day_events<false, false, kMaxWaves>), grid, block, 0, st, v);
    } else if (tu.xcd & 2u) {
        if (v.hist) hipLaunchKernelGGL((k_day_events<true, true, 1>), grid, block, 0, st, v);
        else hipLaunchKernelGGL((k_day_events<true, false, 1>), grid, block, 0, st, v);
    } else {
        if (v.hist) hipLaunchKernelGGL((k_day_events<false, true, 1>), grid, block, 0, st, v);
        else hipLaunchKernelGGL((k_day_events<false, false, 1>), grid, block, 0, st, v);
    }
    if (v.aoi_mark) hipLaunchKernelGGL(k_day_aoi, dim3(v.n_sims), dim3(256), 0, st, v);
    return hipGetLastError();
}

hipError_t launch_longdist(const EngineView& v, const Tuning& tu, hipStream_t st) {
    hipLaunchKernelGGL(k_day_reduce, dim3(v.n_sims * v.reduce_split), dim3(256), 0, st, v, v.reduce_split);
    hipLaunchKernelGGL(k_day_longdist, dim3(v.n_sims * tu.ld_split), dim3(256), 0, st, v, tu.ld_split);
    return hipGetLastError();
}

hipError_t launch_advance(uint32_t* day, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_day, dim3(1), dim3(64), 0, st, day, n);
    return hipGetLastError();
}

hipError_t launch_fill(uint32_t* p, uint64_t n, uint32_t value, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, st, p, n, value);
    return hipGetLastError();
}

}  // namespace arbo
