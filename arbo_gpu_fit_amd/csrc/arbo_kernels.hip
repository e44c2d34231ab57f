// arbo_kernels.hip — the day-loop kernels of the dengue IBM for gfx950.
//
// One simulated day is two launches over every community of every simulation
// of the batch (one 256-thread workgroup = four wave64s per community):
//
//  k_day_demography  ndemographic.cu:32-105 + mdemographic.cu:22-150 +
//                    nVisitingInfected/mVisitingInfected (nepidemic.cu:28-142,
//                    mepidemic.cu:28-138) + dataCollect (datacollect.cu:79-189)
//     streams the community's packed census once (16-byte vector loads of
//     u32 human words and u16 mosquito words), applies deaths, rebirths and
//     E->I / I->S transitions, recounts the infectious per serotype with wave
//     reductions, fills the lowest-index dead mosquito slots up to the seasonal
//     target with a workgroup prefix scan over the dead-slot bitmap, draws the
//     Poisson event counts of the community as a source, spreads the local
//     events over its CSR row with the alias table into an LDS histogram, and
//     adds its counters into the series row of the next day.
//
//  k_day_events      mepidemic.cu:143-246 (nmTransmission) +
//                    nepidemic.cu:146-319 (mnTransmission)
//     gathers the event counts of the community's in-edges, draws its
//     long-distance arrivals, orders all its events (serotype rotation,
//     source, index) and applies them with an LDS claim-min: the earliest
//     event whose target was eligible at the start of the pass wins, which is
//     exactly the serial outcome, with no global atomics on the census.
//
// All randomness is Philox keyed by (seed; entity, day, purpose, index), so a
// simulation's results do not depend on launch shape or batch composition.
#include <hip/hip_runtime.h>

#include "arbo_c.h"
#include "arbo_device.h"
#include "arbo_rng.h"

namespace arbo {

constexpr uint32_t kSlotsPerThread = ARBO_MAX_COMMUNITY_SLOTS / kBlock;  // 32
static_assert(kSlotsPerThread == 32, "one u32 bitmap word per thread");

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Exclusive scan of one value per thread over the 256-thread workgroup.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* s_wave, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t incl = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += y;
    }
    if (lane == 63) s_wave[wid] = incl;
    __syncthreads();
    uint32_t base = 0;
    total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        const uint32_t v = s_wave[w];
        base += (w < wid) ? v : 0u;
        total += v;
    }
    __syncthreads();
    return base + incl - x;
}

// min{a in [1, hi] : K[a] >= k}; hi when K[hi] < k (hi = table end).
__device__ __forceinline__ uint32_t death_age(const uint32_t* K, uint32_t hi, uint32_t k) {
    if (K[hi] < k) return hi;
    uint32_t lo = 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (K[mid] >= k) hi = mid; else lo = mid + 1;
    }
    return lo;
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

constexpr uint32_t kHumanTable = 65536;

struct TableRefs {
    const uint32_t* death_h;
    const uint32_t* death_m;
    uint32_t death_h_max;
    uint32_t death_m_max;
};

__device__ __forceinline__ TableRefs table_refs(const EngineView& v, const SimDesc& sd) {
    TableRefs t;
    t.death_h = v.tables + sd.death_h_off;
    t.death_m = v.tables + sd.death_m_off;
    t.death_h_max = kHumanTable - 1;
    t.death_m_max = sd.death_m_max;
    return t;
}

// ---------------------------------------------------------------------------
// t = 0 census: demographicintial.cu:86-297, epidemicinitial.cu:27-99.
__global__ __launch_bounds__(kBlock) void k_init(EngineView v) {
    const uint32_t g = blockIdx.x;
    const CommDesc cd = v.comm[g];
    const SimDesc sd = v.sims[cd.sim];
    const Key key{sd.k0, sd.k1};
    const uint32_t S = sd.S, tid = threadIdx.x;
    const TableRefs tb = table_refs(v, sd);
    const uint32_t* init_h = v.tables + sd.init_h_off;
    const uint32_t* init_m = v.tables + sd.init_m_off;
    const uint32_t* immune = v.tables + sd.immune_off + (size_t)cd.c * S;
    __shared__ uint32_t s_red[8];
    __shared__ uint32_t s_alive0;
    if (tid < 8) s_red[tid] = 0;
    if (tid == 0) {
        const U4 r = draw(key, cd.c, INIT_DAY, P_INIT_MT, 0);
        uint32_t a = stoch_round(v.dens[sd.dens_off] * cd.base, r.x);
        s_alive0 = a < cd.cap ? a : cd.cap;
    }
    __syncthreads();
    uint32_t ih01 = 0, ih23 = 0, im01 = 0, im23 = 0;
    for (uint32_t i = tid; i < cd.n; i += kBlock) {
        const uint32_t h = cd.h_base + i;
        const U4 r0 = draw(key, h, INIT_DAY, P_INIT_H, 0);
        const uint32_t y = pick_index(init_h, 151, (r0.x >> 8) + 1);
        const uint32_t a0 = 365u * y + mulhi(r0.y, 365);
        const uint32_t da = death_age_given(tb.death_h, tb.death_h_max, a0, r0.z);
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
    }
    const uint32_t alive0 = s_alive0;
    for (uint32_t j = tid; j < cd.cap; j += kBlock) {
        const uint64_t gm = cd.m_off + j;
        v.m_next[gm] = 0;
        if (j >= alive0) { v.m_word[gm] = 3; continue; }
        const uint32_t m = cd.m_base + j;
        const U4 r0 = draw(key, m, INIT_DAY, P_INIT_M, 0);
        const uint32_t a0 = pick_index(init_m, 301, (r0.x >> 8) + 1);
        const uint32_t da = death_age_given(tb.death_m, tb.death_m_max, a0, r0.z);
        const uint32_t dd12 = (uint32_t)((int32_t)da - (int32_t)(a0 + 1)) & 0xFFFu;
        uint32_t st = 0, strain = 0;
        if ((r0.w >> 8) < sd.seed_m) {
            st = 2;
            strain = mulhi(r0.y, S);
            add_pair(im01, im23, strain);
        }
        v.m_word[gm] = (uint16_t)((dd12 << 4) | (strain << 2) | st);
    }
    ih01 = wave_sum(ih01); ih23 = wave_sum(ih23);
    if ((tid & 63u) == 0) { atomicAdd(&s_red[0], ih01); atomicAdd(&s_red[1], ih23); }
    __syncthreads();
    const SeriesIdx si = series_index(sd.series_off, v.max_days, S);
    if (tid == 0) {
        atomicAdd(&v.series[si.n_count], cd.n);
        atomicAdd(&v.series[si.m_count], alive0);
    }
    if (tid < S) {
        const uint32_t ih = unpack_pair(s_red[0], s_red[1], tid);
        atomicAdd(&v.series[si.n_inf + tid], ih);
        if (cd.c == 0) v.series[si.n_one + tid] = ih;
    }
    (void)im01; (void)im23;
}

// ---------------------------------------------------------------------------
// Day kernel A: demography, counters, event counts, series row t+1.
__global__ __launch_bounds__(kBlock) void k_day_demography(EngineView v) {
    const uint32_t g = blockIdx.x;
    const uint32_t t = *v.day;
    const CommDesc cd = v.comm[g];
    const SimDesc sd = v.sims[cd.sim];
    const Key key{sd.k0, sd.k1};
    const uint32_t S = sd.S, tid = threadIdx.x;
    const TableRefs tb = table_refs(v, sd);

    __shared__ uint32_t s_red[8];     // ih01 ih23 inc01 inc23 im01 im23 alive dead
    __shared__ uint8_t s_dead[ARBO_MAX_COMMUNITY_SLOTS / 8];
    __shared__ uint8_t s_died[ARBO_MAX_COMMUNITY_SLOTS / 8];
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t s_births;
    __shared__ uint32_t s_n[8];
    __shared__ uint32_t s_hist[kCntPerComm];
    if (tid < 8) s_red[tid] = 0;
    if (tid < kCntPerComm) s_hist[tid] = 0;
    __syncthreads();

    // ---- humans: death/rebirth, I->S, E->I (ndemographic.cu:57-103) ----
    uint32_t ih01 = 0, ih23 = 0, inc01 = 0, inc23 = 0;
    {
        const uint4* w4p = reinterpret_cast<const uint4*>(v.h_word + cd.h_off);
        const uint32_t nvec = (cd.n + 3) >> 2;
        for (uint32_t q = tid; q < nvec; q += kBlock) {
            const uint4 w4 = w4p[q];
            const uint32_t ws[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                const uint32_t i = 4 * q + e;
                if (i >= cd.n) break;
                const uint32_t w = ws[e];
                const uint32_t st = w & 3u;
                const uint64_t gi = cd.h_off + i;
                if (t >= (w >> 8)) {
                    const U4 r = draw(key, cd.h_base + i, t, P_H_BIRTH, 0);
                    const uint32_t da = death_age(tb.death_h, tb.death_h_max, (r.x >> 8) + 1);
                    v.h_word[gi] = (t + da) << 8;
                    v.h_birth[gi] = (int32_t)t;
                } else if (st != 0) {
                    const uint32_t strain = (w >> 2) & 3u;
                    const uint16_t nx = v.h_next[gi];
                    const bool due = (int16_t)(uint16_t)(t - nx) > 0;
                    if (st == 2) {
                        if (due) v.h_word[gi] = w & ~0xFu;
                        else add_pair(ih01, ih23, strain);
                    } else if (due) {
                        const U4 r = draw(key, cd.h_base + i, t, P_H_REC, 0);
                        v.h_next[gi] = (uint16_t)(nx + stoch_round(sd.recovery, r.x));
                        v.h_word[gi] = (w & ~3u) | 2u;
                        add_pair(inc01, inc23, strain);
                        add_pair(ih01, ih23, strain);
                    }
                }
            }
        }
    }

    // ---- mosquitoes: death and E->I (mdemographic.cu:68-98) ----
    uint32_t im01 = 0, im23 = 0, alive = 0, ndead = 0;
    const uint32_t mvec = (cd.cap + 7) >> 3;
    {
        const uint4* m4p = reinterpret_cast<const uint4*>(v.m_word + cd.m_off);
        for (uint32_t q = tid; q < mvec; q += kBlock) {
            const uint4 raw = m4p[q];
            const uint32_t rw[4] = {raw.x, raw.y, raw.z, raw.w};
            uint32_t dmask = 0, kmask = 0;
#pragma unroll
            for (uint32_t e = 0; e < 8; ++e) {
                const uint32_t j = 8 * q + e;
                if (j >= cd.cap) break;
                const uint32_t w = (rw[e >> 1] >> ((e & 1u) * 16u)) & 0xFFFFu;
                uint32_t st = w & 3u;
                if (st == 3) { dmask |= 1u << e; continue; }
                if (((t - (w >> 4)) & 0xFFFu) < 2048u) { dmask |= 1u << e; kmask |= 1u << e; continue; }
                if (st == 1) {
                    const uint64_t gm = cd.m_off + j;
                    if ((int16_t)(uint16_t)(t - v.m_next[gm]) > 0) {
                        st = 2;
                        v.m_word[gm] = (uint16_t)((w & ~3u) | 2u);
                    }
                }
                alive++;
                if (st == 2) add_pair(im01, im23, (w >> 2) & 3u);
            }
            s_dead[q] = (uint8_t)dmask;
            s_died[q] = (uint8_t)kmask;
            ndead += __popc(dmask);
        }
    }
    ih01 = wave_sum(ih01); ih23 = wave_sum(ih23);
    inc01 = wave_sum(inc01); inc23 = wave_sum(inc23);
    im01 = wave_sum(im01); im23 = wave_sum(im23);
    alive = wave_sum(alive); ndead = wave_sum(ndead);
    if ((tid & 63u) == 0) {
        atomicAdd(&s_red[0], ih01); atomicAdd(&s_red[1], ih23);
        atomicAdd(&s_red[2], inc01); atomicAdd(&s_red[3], inc23);
        atomicAdd(&s_red[4], im01); atomicAdd(&s_red[5], im23);
        atomicAdd(&s_red[6], alive); atomicAdd(&s_red[7], ndead);
    }
    __syncthreads();
    const uint32_t alive_c = s_red[6], dead_c = s_red[7];

    // ---- emergence into the lowest-index dead slots (mdemographic.cu:104-142) ----
    if (tid == 0) {
        const U4 r = draw(key, cd.c, t, P_M_TARGET, 0);
        const uint32_t target = stoch_round(v.dens[sd.dens_off + t] * cd.base, r.x);
        uint32_t b = target > alive_c ? target - alive_c : 0u;
        s_births = b < dead_c ? b : dead_c;
    }
    __syncthreads();
    const uint32_t births = s_births;
    if (dead_c > 0) {
        // thread tid owns bitmap bytes [4 tid, 4 tid + 4): slots [32 tid, 32 tid + 32)
        uint32_t dword = 0, kword = 0;
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
            const uint32_t q = 4 * tid + b;
            if (q < mvec) {
                dword |= (uint32_t)s_dead[q] << (8 * b);
                kword |= (uint32_t)s_died[q] << (8 * b);
            }
        }
        uint32_t total;
        uint32_t rank = block_exclusive_scan(__popc(dword), s_wave, total);
        if ((rank < births) || kword) {
            uint32_t bits = dword;
            while (bits) {
                const uint32_t e = __ffs(bits) - 1;
                bits &= bits - 1;
                const uint32_t jj = 32 * tid + e;
                const uint64_t gm = cd.m_off + jj;
                if (rank < births) {
                    const U4 r = draw(key, cd.m_base + jj, t, P_M_BIRTH, 0);
                    const uint32_t da = death_age(tb.death_m, tb.death_m_max, (r.x >> 8) + 1);
                    v.m_word[gm] = (uint16_t)(((t + da) & 0xFFFu) << 4);
                } else if ((kword >> e) & 1u) {
                    v.m_word[gm] = 3;
                }
                rank++;
            }
        }
    }
    const uint32_t Mc = alive_c + births;

    // ---- event counts of this community as a source (nepidemic.cu:49-136) ----
    if (tid < 8) {
        const uint32_t dir = tid >> 2, s = tid & 3u;
        uint32_t n = 0;
        unsigned long long fx = 0;
        if (s < S) {
            const uint32_t I = dir == DIR_HM ? unpack_pair(s_red[0], s_red[1], s) : unpack_pair(s_red[4], s_red[5], s);
            if (I > 0) {
                const float lam = dir == DIR_HM ? (float)Mc / (float)cd.n * sd.beta_hm * (float)I
                                                : sd.beta_mh * (float)I;
                const float mu_loc = lam * (1.0f - sd.omega);
                const float mu_ld = lam * sd.omega;
                fx = (unsigned long long)floor((double)mu_ld * 16777216.0);
                const Stream st{key, cd.c, t, ctr_tag(P_NLOC, dir, s, 0)};
                n = poisson((double)mu_loc, st);
            }
        }
        s_n[tid] = n;
        if (fx) atomicAdd(&v.lam[((size_t)(t & 1u) * v.n_sims + cd.sim) * 8 + tid], fx);
    }
    __syncthreads();
    uint32_t pre[9];
    pre[0] = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) pre[k + 1] = pre[k] + s_n[k];
    const uint32_t E = pre[8];
    for (uint32_t ev = tid; ev < E; ev += kBlock) {
        uint32_t k = 0;
#pragma unroll
        for (uint32_t q = 1; q < 8; ++q) k += (ev >= pre[q]) ? 1u : 0u;
        const uint32_t i = ev - pre[k];
        const U4 r = draw(key, cd.c, t, ctr_tag(P_DEST, k >> 2, k & 3u, 0), i);
        uint32_t col = mulhi(r.x, cd.row_len);
        const uint32_t ent = v.csr_thr[cd.row_off + col];
        if ((r.y >> 8) >= (ent & 0x1FFFFFFu)) col = ent >> 27;
        atomicAdd(&s_hist[k * 9 + col], 1u);
    }
    __syncthreads();
    if (tid < kCntPerComm) v.cnt[(size_t)g * kCntPerComm + tid] = s_hist[tid];

    // ---- series row t+1 (datacollect.cu:79-189) ----
    const SeriesIdx si = series_index(sd.series_off, v.max_days, S);
    const uint32_t row = t + 1;
    if (tid == 0) {
        atomicAdd(&v.series[si.n_count + row], cd.n);
        atomicAdd(&v.series[si.m_count + row], Mc);
    }
    if (tid < S) {
        const uint32_t ih = unpack_pair(s_red[0], s_red[1], tid);
        const uint32_t ic = unpack_pair(s_red[2], s_red[3], tid);
        atomicAdd(&v.series[si.n_inf + (size_t)row * S + tid], ih);
        if (ic) atomicAdd(&v.series[si.n_inc + (size_t)row * S + tid], ic);
        if (cd.c == 0) v.series[si.n_one + (size_t)row * S + tid] = ih;
    }
}

// ---------------------------------------------------------------------------
// Day kernel B: application of the day's events at one destination.
struct EventRef { uint32_t s, x0, x2, k; };

__device__ __forceinline__ EventRef decode_event(uint32_t r, const uint32_t* seg, uint32_t nseg, uint32_t in_len,
                                                 uint32_t rot, uint32_t S, uint32_t dir, const uint32_t* src_local,
                                                 const uint32_t* slot, uint32_t self_c) {
    uint32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= r) lo = mid; else hi = mid - 1;
    }
    const uint32_t pos = lo / (in_len + 1), e = lo % (in_len + 1);
    EventRef ev;
    ev.s = (rot + pos) % S;
    ev.k = r - seg[lo];
    if (e < in_len) {
        ev.x0 = src_local[e];
        ev.x2 = ctr_tag(P_TGT, dir, ev.s, slot[e]);
    } else {
        ev.x0 = self_c;
        ev.x2 = ctr_tag(P_TGT_LD, dir, ev.s, 0);
    }
    return ev;
}

__global__ __launch_bounds__(kBlock) void k_day_events(EngineView v) {
    const uint32_t g = blockIdx.x;
    const uint32_t t = *v.day;
    const CommDesc cd = v.comm[g];
    const SimDesc sd = v.sims[cd.sim];
    const Key key{sd.k0, sd.k1};
    const uint32_t S = sd.S, tid = threadIdx.x;
    const uint32_t in_len = cd.in_len;
    const uint32_t nseg = S * (in_len + 1);

    extern __shared__ uint32_t s_dyn[];
    uint32_t* s_claim = s_dyn;                                         // [max(n, cap)]
    uint16_t* s_alive = reinterpret_cast<uint16_t*>(s_dyn + ARBO_MAX_COMMUNITY_SLOTS);  // [cap]
    __shared__ uint32_t s_src[9], s_slot[9];
    __shared__ uint32_t s_ecnt[2][4][9];
    __shared__ uint32_t s_nld[8];
    __shared__ uint32_t s_rot[2];
    __shared__ uint32_t s_seg[2][4 * 10 + 1];
    __shared__ uint32_t s_wave[kBlock / 64];

    if (tid < in_len) {
        const uint32_t src = v.in_edge[cd.in_off + tid];
        s_src[tid] = src;
        s_slot[tid] = v.in_slot[cd.in_off + tid];
    }
    __syncthreads();
    for (uint32_t i = tid; i < 8 * in_len; i += kBlock) {
        const uint32_t k = i / in_len, e = i % in_len;      // k = dir*4 + s
        s_ecnt[k >> 2][k & 3u][e] = ((k & 3u) < S) ? v.cnt[(size_t)s_src[e] * kCntPerComm + k * 9 + s_slot[e]] : 0u;
    }
    if (tid < 8) {
        const uint32_t dir = tid >> 2, s = tid & 3u;
        uint32_t n = 0;
        if (s < S) {
            const unsigned long long fx = v.lam[((size_t)(t & 1u) * v.n_sims + cd.sim) * 8 + tid];
            const double mu = (double)fx * 0x1p-24 / (double)sd.C;
            const Stream st{key, cd.c, t, ctr_tag(P_NLD, dir, s, 0)};
            n = poisson(mu, st);
        }
        s_nld[tid] = n;
    }
    if (tid < 2) s_rot[tid] = mulhi(draw(key, cd.c, t, ctr_tag(P_ROT, tid, 0, 0), 0).x, S);
    if (cd.c == 0 && tid < 8) v.lam[((size_t)((t + 1) & 1u) * v.n_sims + cd.sim) * 8 + tid] = 0ull;
    __syncthreads();
    if (tid < 2) {
        const uint32_t dir = tid;
        uint32_t acc = 0, idx = 0;
        for (uint32_t pos = 0; pos < S; ++pos) {
            const uint32_t s = (s_rot[dir] + pos) % S;
            for (uint32_t e = 0; e <= in_len; ++e) {
                s_seg[dir][idx++] = acc;
                acc += (e < in_len) ? s_ecnt[dir][s][e] : s_nld[dir * 4 + s];
            }
        }
        s_seg[dir][idx] = acc;
    }
    if (tid < in_len) s_src[tid] -= sd.comm_off;   // simulation-local source ids
    __syncthreads();

    // ---- onto mosquitoes: uniform over alive slots, success iff susceptible ----
    const uint32_t tot_m = s_seg[DIR_HM][nseg];
    if (tot_m > 0) {
        uint32_t aw[kSlotsPerThread / 2];
        uint32_t nalive = 0;
        const uint32_t j0 = kSlotsPerThread * tid;
        {
            const uint4* m4p = reinterpret_cast<const uint4*>(v.m_word + cd.m_off + j0);
#pragma unroll
            for (uint32_t b = 0; b < kSlotsPerThread / 8; ++b) {
                uint4 raw = make_uint4(0x00030003u, 0x00030003u, 0x00030003u, 0x00030003u);
                if (j0 + 8 * b < cd.cap) raw = m4p[b];
                aw[4 * b + 0] = raw.x; aw[4 * b + 1] = raw.y; aw[4 * b + 2] = raw.z; aw[4 * b + 3] = raw.w;
            }
#pragma unroll
            for (uint32_t e = 0; e < kSlotsPerThread; ++e) {
                const uint32_t w = (aw[e >> 1] >> ((e & 1u) * 16u)) & 0xFFFFu;
                if (j0 + e < cd.cap && (w & 3u) != 3u) nalive++;
            }
        }
        uint32_t Md;
        uint32_t q = block_exclusive_scan(nalive, s_wave, Md);
#pragma unroll
        for (uint32_t e = 0; e < kSlotsPerThread; ++e) {
            const uint32_t w = (aw[e >> 1] >> ((e & 1u) * 16u)) & 0xFFFFu;
            if (j0 + e < cd.cap && (w & 3u) != 3u) {
                s_alive[q] = (uint16_t)((j0 + e) | ((w & 3u) == 0 ? 0x8000u : 0u));
                s_claim[q] = 0xFFFFFFFFu;
                q++;
            }
        }
        __syncthreads();
        if (Md > 0) {
            for (uint32_t r = tid; r < tot_m; r += kBlock) {
                const EventRef ev = decode_event(r, s_seg[DIR_HM], nseg, in_len, s_rot[DIR_HM], S, DIR_HM, s_src, s_slot, cd.c);
                const uint32_t qq = mulhi(draw(key, ev.x0, t, ev.x2, ev.k).x, Md);
                if (s_alive[qq] & 0x8000u) atomicMin(&s_claim[qq], r);
            }
            __syncthreads();
            for (uint32_t r = tid; r < tot_m; r += kBlock) {
                const EventRef ev = decode_event(r, s_seg[DIR_HM], nseg, in_len, s_rot[DIR_HM], S, DIR_HM, s_src, s_slot, cd.c);
                const uint32_t qq = mulhi(draw(key, ev.x0, t, ev.x2, ev.k).x, Md);
                if (s_claim[qq] == r) {
                    const uint32_t jj = s_alive[qq] & 0x7FFFu;
                    const uint64_t gm = cd.m_off + jj;
                    const uint32_t w = v.m_word[gm];
                    v.m_word[gm] = (uint16_t)((w & ~0xFu) | (ev.s << 2) | 1u);
                    const U4 r2 = draw(key, cd.m_base + jj, t, P_M_EIP, 0);
                    v.m_next[gm] = (uint16_t)(t + stoch_round(sd.eip, r2.x));
                }
            }
        }
        __syncthreads();
    }

    // ---- onto humans: uniform target, success iff S and never infected by s ----
    const uint32_t tot_h = s_seg[DIR_MH][nseg];
    if (tot_h > 0 || sd.intro > 0.0f) {
        const uint32_t N = cd.n;
        for (uint32_t pos = 0; pos < S; ++pos) {
            const uint32_t s = (s_rot[DIR_MH] + pos) % S;
            const uint32_t rb = s_seg[DIR_MH][pos * (in_len + 1)];
            const uint32_t re = s_seg[DIR_MH][(pos + 1) * (in_len + 1)];
            if (re > rb) {
                for (uint32_t h = tid; h < N; h += kBlock) s_claim[h] = 0xFFFFFFFFu;
                __syncthreads();
                for (uint32_t r = rb + tid; r < re; r += kBlock) {
                    const EventRef ev = decode_event(r, s_seg[DIR_MH], nseg, in_len, s_rot[DIR_MH], S, DIR_MH, s_src, s_slot, cd.c);
                    const uint32_t h = mulhi(draw(key, ev.x0, t, ev.x2, ev.k).x, N);
                    const uint32_t w = v.h_word[cd.h_off + h];
                    if ((w & 3u) == 0 && !((w >> (4 + s)) & 1u)) atomicMin(&s_claim[h], r);
                }
                __syncthreads();
                for (uint32_t r = rb + tid; r < re; r += kBlock) {
                    const EventRef ev = decode_event(r, s_seg[DIR_MH], nseg, in_len, s_rot[DIR_MH], S, DIR_MH, s_src, s_slot, cd.c);
                    const uint32_t h = mulhi(draw(key, ev.x0, t, ev.x2, ev.k).x, N);
                    if (s_claim[h] == r) {
                        const uint64_t gi = cd.h_off + h;
                        const uint32_t w = v.h_word[gi];
                        v.h_word[gi] = (w & ~0xFu) | (1u << (4 + s)) | (s << 2) | 1u;
                        const U4 r2 = draw(key, cd.h_base + h, t, P_H_IIP, 0);
                        v.h_next[gi] = (uint16_t)(t + sd.iip_min + mulhi(r2.x, sd.iip_span));
                    }
                }
                __syncthreads();
            }
            if (sd.intro > 0.0f) {  // external introductions (nepidemic.cu:261-309)
                if (tid == 0) {
                    const uint32_t tagi = ctr_tag(P_INTRO, 0, s, 0);
                    uint32_t x = stoch_round(sd.intro / (float)sd.C, draw(key, cd.c, t, tagi, 0).x);
                    for (uint32_t att = 0; x > 0 && att < 1000; ++att) {
                        const uint32_t h = mulhi(draw(key, cd.c, t, tagi, 1 + att).x, N);
                        const uint64_t gi = cd.h_off + h;
                        const uint32_t w = v.h_word[gi];
                        if ((w & 3u) == 0 && !((w >> (4 + s)) & 1u)) {
                            v.h_word[gi] = (w & ~0xFu) | (1u << (4 + s)) | (s << 2) | 1u;
                            const U4 r2 = draw(key, cd.h_base + h, t, P_H_IIP, 0);
                            v.h_next[gi] = (uint16_t)(t + sd.iip_min + mulhi(r2.x, sd.iip_span));
                            x--;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

__global__ void k_advance_day(uint32_t* day) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *day = *day + 1u;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---------------------------------------------------------------------------
// Launchers (host).
size_t events_lds_bytes() {
    return (size_t)ARBO_MAX_COMMUNITY_SLOTS * 4 + (size_t)ARBO_MAX_COMMUNITY_SLOTS * 2;
}

hipError_t launch_init(const EngineView& v, hipStream_t st) {
    hipLaunchKernelGGL(k_init, dim3(v.n_comm), dim3(kBlock), 0, st, v);
    return hipGetLastError();
}

hipError_t launch_demography(const EngineView& v, hipStream_t st) {
    hipLaunchKernelGGL(k_day_demography, dim3(v.n_comm), dim3(kBlock), 0, st, v);
    return hipGetLastError();
}

hipError_t launch_events(const EngineView& v, hipStream_t st) {
    hipLaunchKernelGGL(k_day_events, dim3(v.n_comm), dim3(kBlock), events_lds_bytes(), st, v);
    return hipGetLastError();
}

hipError_t launch_advance(uint32_t* day, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_day, dim3(1), dim3(64), 0, st, day);
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
