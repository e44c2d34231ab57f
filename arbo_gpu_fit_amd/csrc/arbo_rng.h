// arbo_rng.h — counter-based randomness and deterministic floating-point
// primitives shared by the HIP kernels and the engine's host-side table
// builders.
//
// Everything here is written so that the same inputs give the same bits on
// x86-64 and on gfx950: only IEEE-754 +, -, *, / (correctly rounded on both),
// integer arithmetic and bit manipulation are used, and every translation unit
// is compiled with -ffp-contract=off so no multiply-add is fused behind our
// back.  No libm transcendental feeds simulation state.
//
// Replaces the curand XORWOW state-per-hardware-thread scheme of the reference
// (setuprng.cu:11-40, cudaidentities.h:14-37) and its curand_uniform /
// `curand() % n` call sites (e.g. ndemographic.cu:94, nepidemic.cu:104,197,210).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ARBO_HD __host__ __device__ __forceinline__

namespace arbo {

struct U4 { uint32_t x, y, z, w; };

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Multipliers and Weyl key increments are the published
// constants of the generator.
ARBO_HD U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        U4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Purpose codes (low byte of counter word 2).  Bits 8.. carry direction,
// serotype and CSR slot, see ctr_tag().
enum Purpose : uint32_t {
    P_INIT_H = 1,      // human t=0 census (x3 = 0,1,2)
    P_INIT_M = 2,      // mosquito slot t=0 census
    P_INIT_MT = 3,     // community t=0 mosquito count
    P_IMMUNE = 4,      // per-(community, serotype) Beta(2,5) immunity fraction
    P_H_BIRTH = 10,    // human rebirth: death age of the newborn
    P_H_REC = 11,      // human E->I: length of the infectious period
    P_M_BIRTH = 12,    // mosquito emergence: death ages of births 4i..4i+3 of a community (x3 = i)
    P_M_TARGET = 13,   // community emergence target rounding
    P_NLOC = 20,       // local event count of a source (community, dir, serotype)
    P_DEST = 21,       // alias-table destination of the k-th local event
    P_NLD = 22,        // long-distance events of a simulation (x0 = INIT_DAY)
    P_ROT = 23,        // serotype rotation start at a destination
    P_TGT = 24,        // target of the k-th local event on CSR slot j
    P_TGT_LD = 25,     // target of the k-th long-distance arrival
    P_H_IIP = 26,      // newly infected human: intrinsic incubation
    P_M_EIP = 27,      // newly infected mosquito: extrinsic incubation
    P_INTRO = 28,      // external introductions (x3 = 0 count, 1+a attempt a)
    P_LD_DEST = 29,    // destination of the l-th long-distance event (x0 = l)
    P_NSR = 30,        // stochastic-rounding event count of a source and its splits
    P_NSTAY = 31,      // Poisson count of a source's events that stay at the source
    P_SIZES = 32,      // community of individual 4 x0 + word (x1 = INIT_DAY) of arbo_community_sizes
};

// Direction of a transmission event.
enum Dir : uint32_t { DIR_HM = 0,   // infectious human -> mosquito ("d_nTransmission")
                      DIR_MH = 1 }; // infectious mosquito -> human ("d_mTransmission")

ARBO_HD uint32_t ctr_tag(uint32_t purpose, uint32_t dir, uint32_t s, uint32_t slot) {
    return purpose | (dir << 8) | (s << 12) | (slot << 16);
}

constexpr uint32_t INIT_DAY = 0xFFFFFFFFu;

struct Key { uint32_t k0, k1; };

ARBO_HD U4 draw(Key k, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3) {
    U4 c; c.x = x0; c.y = x1; c.z = x2; c.w = x3;
    return philox(c, k.k0, k.k1);
}

// floor(r * n / 2^32): an exact integer draw in [0, n).
ARBO_HD uint32_t mulhi(uint32_t r, uint32_t n) {
    return (uint32_t)(((uint64_t)r * (uint64_t)n) >> 32);
}

// 24-bit uniform in (0, 1], exactly representable in f32.
ARBO_HD float u24f(uint32_t r) { return (float)((r >> 8) + 1u) * 0x1p-24f; }

// Stochastic rounding of the reference's decimalResolve (nepidemic.cu:13-24):
// ceil(x) with probability frac(x), else floor(x).  x must be >= 0.
ARBO_HD uint32_t stoch_round(float x, uint32_t r) {
    const float fl = floorf(x);
    const float frac = x - fl;
    return (u24f(r) <= frac) ? (uint32_t)ceilf(x) : (uint32_t)fl;
}

// Guide buckets of a 24-bit lifetime draw k in [1, 2^24]: the survival
// c = 2^24 - k on a log scale (its leading bit and the three bits after it),
// so the draws of one bucket have survivals within a factor 9/8 of each other
// (c < 8 has a bucket per value).  A table search started at the bucket's
// smallest death age then walks at most a few entries, also in the tail.
constexpr uint32_t kLifeBuckets = 176;
ARBO_HD uint32_t life_bucket(uint32_t k) {
    const uint32_t c = (1u << 24) - k;
    if (c < 8u) return c;
    const uint32_t e = 31u - (uint32_t)__builtin_clz(c);  // 3..23
    return 8u + (e - 3u) * 8u + ((c >> (e - 3u)) & 7u);
}
// The smallest draw of bucket b (its largest survival).
ARBO_HD uint32_t life_bucket_kmin(uint32_t b) {
    if (b < 8u) return (1u << 24) - b;
    const uint32_t e = 3u + (b - 8u) / 8u, m = (b - 8u) % 8u;
    return (1u << 24) - (((9u + m) << (e - 3u)) - 1u);
}

// 53-bit uniform in [0, 1) from two words.
ARBO_HD double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 0x1p-53;
}

ARBO_HD double bits_to_f64(uint64_t b) { return __builtin_bit_cast(double, b); }
ARBO_HD uint64_t f64_to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

// 2^k for integer k in [-1074, 1023].
ARBO_HD double pow2i(int k) {
    if (k >= -1022) return bits_to_f64((uint64_t)(k + 1023) << 52);
    return bits_to_f64((uint64_t)(k + 1022 + 1023) << 52) * 0x1p-1022;
}

// exp(x): Cody & Waite (1980) reduction x = k ln2 + r with the two-part ln2 of
// fdlibm, then the degree-13 Taylor polynomial of e^r on |r| <= ln2/2
// (truncation error < 5e-18 relative).
ARBO_HD double dexp(double x) {
    if (!(x > -745.0)) return 0.0;
    if (x > 709.0) return bits_to_f64(0x7FF0000000000000ull);
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double kd = floor(x * inv_ln2 + 0.5);
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 6227020800.0;           // 1/13!
    p = p * r + 1.0 / 479001600.0;           // 1/12!
    p = p * r + 1.0 / 39916800.0;            // 1/11!
    p = p * r + 1.0 / 3628800.0;             // 1/10!
    p = p * r + 1.0 / 362880.0;              // 1/9!
    p = p * r + 1.0 / 40320.0;               // 1/8!
    p = p * r + 1.0 / 5040.0;                // 1/7!
    p = p * r + 1.0 / 720.0;                 // 1/6!
    p = p * r + 1.0 / 120.0;                 // 1/5!
    p = p * r + 1.0 / 24.0;                  // 1/4!
    p = p * r + 1.0 / 6.0;                   // 1/3!
    p = p * r + 0.5;                         // 1/2!
    p = p * r + 1.0;
    p = p * r + 1.0;
    const int k = (int)kd;
    if (k < -1021) return (p * pow2i(k + 600)) * 0x1p-600;
    return p * pow2i(k);
}

// log(x): x = 2^e m with m in [sqrt(1/2), sqrt(2)], log m = 2 atanh(s),
// s = (m-1)/(m+1), summed as the odd power series of atanh to s^21
// (|s| <= 0.1716, truncation < 1e-18 relative).
ARBO_HD double dlog(double x) {
    if (!(x > 0.0)) return -bits_to_f64(0x7FF0000000000000ull);
    uint64_t b = f64_to_bits(x);
    int e = (int)((b >> 52) & 0x7FF);
    if (e == 0x7FF) return x;                 // +inf
    if (e == 0) {                              // subnormal
        x = x * 0x1p54;
        b = f64_to_bits(x);
        e = (int)((b >> 52) & 0x7FF) - 54;
    }
    e -= 1023;
    double m = bits_to_f64((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    p = p * z + 1.0;
    const double logm = 2.0 * s * p;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double ed = (double)e;
    return ed * ln2_hi + (ed * ln2_lo + logm);
}

// sin(x): reduction by 2*pi in two parts to r in [-pi, pi], then the Taylor
// series of sin to r^33 (alternating; absolute error ~1e-15).
ARBO_HD double dsin(double x) {
    const double twopi_hi = 6.28318530717958623200e+00;
    const double twopi_lo = 2.44929359829470635445e-16;
    const double inv_twopi = 1.59154943091895345555e-01;
    const double kd = floor(x * inv_twopi + 0.5);
    const double r = (x - kd * twopi_hi) - kd * twopi_lo;
    const double z = r * r;
    // sum_{n=0}^{16} (-1)^n r^(2n+1)/(2n+1)!
    double p = 1.0 / 8.68331761881188649551e+36;          // 1/33!
    p = -p * z + 1.0 / 8.22283865417792281773e+33;        // 1/31!
    p = -p * z + 1.0 / 8.84176199373970195455e+30;        // 1/29!
    p = -p * z + 1.0 / 1.08888694504183521607e+28;        // 1/27!
    p = -p * z + 1.0 / 1.55112100433309859840e+25;        // 1/25!
    p = -p * z + 1.0 / 2.58520167388849766400e+22;        // 1/23!
    p = -p * z + 1.0 / 5.10909421717094400000e+19;        // 1/21!
    p = -p * z + 1.0 / 1.21645100408832000000e+17;        // 1/19!
    p = -p * z + 1.0 / 355687428096000.0;                 // 1/17!
    p = -p * z + 1.0 / 1307674368000.0;                   // 1/15!
    p = -p * z + 1.0 / 6227020800.0;                      // 1/13!
    p = -p * z + 1.0 / 39916800.0;                        // 1/11!
    p = -p * z + 1.0 / 362880.0;                          // 1/9!
    p = -p * z + 1.0 / 5040.0;                            // 1/7!
    p = -p * z + 1.0 / 120.0;                             // 1/5!
    p = -p * z + 1.0 / 6.0;                               // 1/3!
    p = -p * z + 1.0;
    return r * p;
}

// log(k!) for integer k >= 0: exact sums of dlog below 16, Stirling's series
// for log Gamma(k+1) above (error < 1e-15 at k = 16).
ARBO_HD double dlogfact(double k) {
    if (k < 16.0) {
        double s = 0.0;
        for (double i = 2.0; i <= k; i += 1.0) s += dlog(i);
        return s;
    }
    const double x = k + 1.0;
    const double ix = 1.0 / x;
    const double ix2 = ix * ix;
    const double half_log_2pi = 0.91893853320467274178;
    double series = -1.0 / 1680.0;
    series = series * ix2 + 1.0 / 1260.0;
    series = series * ix2 - 1.0 / 360.0;
    series = series * ix2 + 1.0 / 12.0;
    return (x - 0.5) * dlog(x) - x + half_log_2pi + series * ix;
}

// Counter base of one Poisson variate: x3 is the draw index.
struct Stream { Key key; uint32_t x0, x1, x2; };

ARBO_HD U4 stream_draw(const Stream& s, uint32_t i) {
    return draw(s.key, s.x0, s.x1, s.x2, i);
}

constexpr uint32_t POIS_INV_MAX = 200;     // inversion search cap (mu < 10)
constexpr uint32_t POIS_PTRS_MAXIT = 256;  // rejection loop cap (mu >= 10)

// Poisson(mu).  mu < 10: sequential-search inversion of one 53-bit uniform.
// mu >= 10: PTRS, the transformed rejection with squeeze of W. Hoermann,
// "The transformed rejection method for generating Poisson random variables",
// Insurance: Mathematics and Economics 12 (1993), one Philox block per trial.
// Both branches are a fixed sequence of IEEE operations, so the result is
// bit-identical on every platform that runs this code.
ARBO_HD uint32_t poisson(double mu, const Stream& st) {
    if (!(mu > 0.0)) return 0u;
    if (mu < 10.0) {
        const U4 r = stream_draw(st, 0);
        const double u = u53(r.x, r.y);
        double p = dexp(-mu);
        double F = p;
        uint32_t k = 0;
        while (u >= F && k < POIS_INV_MAX) {
            k += 1u;
            p = p * mu / (double)k;
            F = F + p;
        }
        return k;
    }
    const double loglam = dlog(mu);
    const double slam = dexp(0.5 * loglam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (uint32_t it = 0; it < POIS_PTRS_MAXIT; ++it) {
        const U4 r = stream_draw(st, it);
        const double U = u53(r.x, r.y) - 0.5;
        const double V = u53(r.z, r.w);
        const double us = 0.5 - (U < 0.0 ? -U : U);
        if (!(us > 0.0)) continue;
        const double kf = floor((2.0 * a / us + b) * U + mu + 0.43);
        if (us >= 0.07 && V <= vr) {
            if (kf < 0.0) continue;
            return (uint32_t)kf;
        }
        if (kf < 0.0 || (us < 0.013 && V > us)) continue;
        if (kf > 4.0e9) continue;
        const double lhs = dlog(V) + dlog(invalpha) - dlog(a / (us * us) + b);
        const double rhs = -mu + kf * loglam - dlogfact(kf);
        if (lhs <= rhs) return (uint32_t)kf;
    }
    return (uint32_t)floor(mu);
}

}  // namespace arbo
