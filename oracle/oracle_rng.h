// oracle_rng.h — TEST INFRASTRUCTURE ONLY (the CPU oracle).  Nothing in the
// product path includes this file; tests, smoke() and the bench's CPU baseline
// use it through liboracle.
//
// Counter-based randomness and deterministic double-precision functions of the
// oracle.  They restate the engine's randomness specification independently:
//  - Philox4x32-10, Salmon et al., SC'11 (pinned by the Random123 known-answer
//    vectors and by rocRAND's open-source philox4x32_10 engine in the tests);
//  - exp: Cody-Waite reduction by ln2 + degree-13 Taylor polynomial;
//  - log: 2 atanh((m-1)/(m+1)) series to s^21 after exponent extraction;
//  - sin: reduction by 2 pi + Taylor series to r^33;
//  - log k!: sums of log below 16, Stirling series above;
//  - Poisson: inversion below mean 10, Hoermann's PTRS (1993) at and above 10.
// The exp/log/Poisson accuracy is pinned against numpy/scipy in the tests.
// They replace curand's XORWOW streams (setuprng.cu:11-40) and the
// curand_uniform / `curand() % n` draws of the reference kernels.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

struct W4 { uint32_t v[4]; };

inline W4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    uint32_t c[4] = {c0, c1, c2, c3};
    for (int round = 0; round < 10; ++round) {
        uint64_t m0 = (uint64_t)c[0] * 0xD2511F53ull;
        uint64_t m1 = (uint64_t)c[2] * 0xCD9E8D57ull;
        uint32_t n0 = (uint32_t)(m1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)m1;
        uint32_t n2 = (uint32_t)(m0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)m0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    W4 out;
    std::memcpy(out.v, c, sizeof c);
    return out;
}

// Counter word 2: purpose | dir << 8 | serotype << 12 | slot << 16.
enum : uint32_t {
    INIT_H = 1, INIT_M = 2, INIT_MT = 3, IMMUNE = 4,
    H_BIRTH = 10, H_REC = 11, M_BIRTH = 12, M_TARGET = 13,
    NLOC = 20, DEST = 21, NLD = 22, ROT = 23, TGT = 24, TGT_LD = 25,
    H_IIP = 26, M_EIP = 27, INTRO = 28, LD_DEST = 29, NSR = 30, NSTAY = 31
};
enum : uint32_t { HM = 0, MH = 1 };
constexpr uint32_t DAY_INIT = 0xFFFFFFFFu;

inline uint32_t tag(uint32_t purpose, uint32_t dir, uint32_t s, uint32_t slot) {
    return purpose + (dir << 8) + (s << 12) + (slot << 16);
}

struct Seed { uint32_t lo, hi; };

inline W4 rnd(Seed k, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return philox10(a, b, c, d, k.lo, k.hi);
}

inline uint32_t range(uint32_t r, uint32_t n) { return (uint32_t)(((uint64_t)r * n) >> 32); }

inline float unit24(uint32_t r) { return (float)((r >> 8) + 1u) / 16777216.0f; }

// decimalResolve of nepidemic.cu:13-24.
inline uint32_t sround(float x, uint32_t r) {
    float lo = std::floor(x);
    float fr = x - lo;
    if (unit24(r) <= fr) return (uint32_t)std::ceil(x);
    return (uint32_t)lo;
}

inline double unit53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

inline double from_bits(uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; }
inline uint64_t to_bits(double d) { uint64_t b; std::memcpy(&b, &d, 8); return b; }

inline double two_to(int k) {
    if (k >= -1022) return from_bits((uint64_t)(k + 1023) << 52);
    return from_bits((uint64_t)(k + 2045) << 52) * from_bits((uint64_t)(1) << 52);  // * 2^-1022
}

inline double exp_d(double x) {
    if (!(x > -745.0)) return 0.0;
    if (x > 709.0) return HUGE_VAL;
    const double hi = 6.93147180369123816490e-01, lo = 1.90821492927058770002e-10;
    double kd = std::floor(x * 1.44269504088896338700e+00 + 0.5);
    double r = (x - kd * hi) - kd * lo;
    // Horner from 1/13! down to 1/0!
    static const double inv_fact[14] = {
        1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0, 1.0 / 5040.0,
        1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0, 1.0 / 39916800.0,
        1.0 / 479001600.0, 1.0 / 6227020800.0};
    double p = inv_fact[13];
    for (int i = 12; i >= 0; --i) p = p * r + inv_fact[i];
    int k = (int)kd;
    if (k < -1021) return (p * two_to(k + 600)) * 0x1p-600;
    return p * two_to(k);
}

inline double log_d(double x) {
    if (!(x > 0.0)) return -HUGE_VAL;
    uint64_t b = to_bits(x);
    int e = (int)((b >> 52) & 0x7FF);
    if (e == 0x7FF) return x;
    if (e == 0) {
        x = x * 0x1p54;
        b = to_bits(x);
        e = (int)((b >> 52) & 0x7FF) - 54;
    }
    e -= 1023;
    double m = from_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 1.0 / 21.0;
    for (int n = 19; n >= 1; n -= 2) p = p * z + 1.0 / (double)n;
    double lm = 2.0 * s * p;
    double ed = (double)e;
    return ed * 6.93147180369123816490e-01 + (ed * 1.90821492927058770002e-10 + lm);
}

inline double sin_d(double x) {
    double kd = std::floor(x * 1.59154943091895345555e-01 + 0.5);
    double r = (x - kd * 6.28318530717958623200e+00) - kd * 2.44929359829470635445e-16;
    double z = r * r;
    static const double fact_odd[17] = {  // (2n+1)!, n = 0..16
        1.0, 6.0, 120.0, 5040.0, 362880.0, 39916800.0, 6227020800.0, 1307674368000.0,
        355687428096000.0, 1.21645100408832000000e+17, 5.10909421717094400000e+19,
        2.58520167388849766400e+22, 1.55112100433309859840e+25, 1.08888694504183521607e+28,
        8.84176199373970195455e+30, 8.22283865417792281773e+33, 8.68331761881188649551e+36};
    double p = 1.0 / fact_odd[16];
    for (int n = 15; n >= 0; --n) p = -p * z + 1.0 / fact_odd[n];
    return r * p;
}

inline double logfact_d(double k) {
    if (k < 16.0) {
        double s = 0.0;
        for (double i = 2.0; i <= k; i += 1.0) s += log_d(i);
        return s;
    }
    double x = k + 1.0;
    double ix = 1.0 / x;
    double ix2 = ix * ix;
    double ser = -1.0 / 1680.0;
    ser = ser * ix2 + 1.0 / 1260.0;
    ser = ser * ix2 - 1.0 / 360.0;
    ser = ser * ix2 + 1.0 / 12.0;
    return (x - 0.5) * log_d(x) - x + 0.91893853320467274178 + ser * ix;
}

// Poisson variate; draw i of the stream uses counter word 3 = i.
inline uint32_t poisson_d(double mu, Seed k, uint32_t c0, uint32_t c1, uint32_t c2) {
    if (!(mu > 0.0)) return 0;
    if (mu < 10.0) {
        W4 r = rnd(k, c0, c1, c2, 0);
        double u = unit53(r.v[0], r.v[1]);
        double p = exp_d(-mu);
        double cdf = p;
        uint32_t n = 0;
        while (u >= cdf && n < 200u) {
            n += 1;
            p = p * mu / (double)n;
            cdf = cdf + p;
        }
        return n;
    }
    double loglam = log_d(mu);
    double slam = exp_d(0.5 * loglam);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (uint32_t it = 0; it < 256u; ++it) {
        W4 r = rnd(k, c0, c1, c2, it);
        double U = unit53(r.v[0], r.v[1]) - 0.5;
        double V = unit53(r.v[2], r.v[3]);
        double us = 0.5 - std::fabs(U);
        if (!(us > 0.0)) continue;
        double kk = std::floor((2.0 * a / us + b) * U + mu + 0.43);
        if (us >= 0.07 && V <= vr) {
            if (kk < 0.0) continue;
            return (uint32_t)kk;
        }
        if (kk < 0.0 || (us < 0.013 && V > us)) continue;
        if (kk > 4.0e9) continue;
        double lhs = log_d(V) + log_d(invalpha) - log_d(a / (us * us) + b);
        double rhs = -mu + kk * loglam - logfact_d(kk);
        if (lhs <= rhs) return (uint32_t)kk;
    }
    return (uint32_t)std::floor(mu);
}

}  // namespace orc
