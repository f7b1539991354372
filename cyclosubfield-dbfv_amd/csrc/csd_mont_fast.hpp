// Montgomery steps with cheaper high products, for the tails of the conversion kernels.
//
// Each function returns exactly the bits of its csd_arith.hpp counterpart (red_words,
// csd_mont_mul, csd_double_red128), within the stated operand ranges; only the way the high
// words of the 64 x 64-bit products are formed differs:
//  * a value x with |x| < 2^63 - 2^31 is split as x = hi 2^32 + lo with both halves signed 32-bit
//    ("balanced"), so the four partial products of floor(x c / 2^64) are v_mad_i64_i32 steps and no
//    sign correction is needed;
//  * floor(m N / 2^64) for an arbitrary 64-bit m (the Montgomery quotient) splits m as signed high
//    word and unsigned low word and N both ways (N = n1 2^32 + n0 unsigned, and with the high word
//    n1b = n1 + [n0 >= 2^31] of its balanced split): m N = m1 n1b 2^64 + (m1 (int32)n0 + m0 n1) 2^32
//    + m0 n0, each term a single multiply-add that cannot overflow for N < 2^62;
//  * for a fixed multiplier c the Montgomery quotient lo(a c) ninv mod 2^64 is lo(a (c ninv)), and
//    [lo(a c) != 0] = [quotient != 0] (ninv is odd), so the low word of a c is never formed.
#pragma once
#include "csd_arith.hpp"

struct CsdBal {
    int32_t lo, hi;
};

// the balanced halves of x: needs |x| < 2^63 - 2^31
CSD_HD CsdBal csd_bal(int64_t x) {
    CsdBal b;
    b.lo = (int32_t)(uint32_t)(uint64_t)x;
    b.hi = (int32_t)((uint32_t)((uint64_t)x >> 32) + ((uint32_t)(uint64_t)x >> 31));
    return b;
}

// floor(a b / 2^64) of two balanced values
CSD_HD int64_t csd_mulhi_bal(CsdBal a, CsdBal b) {
    const int64_t p = (int64_t)a.lo * b.lo;
    int64_t q = (int64_t)a.hi * b.lo + (p >> 32);
    q = (int64_t)a.lo * b.hi + q;
    return (int64_t)a.hi * b.hi + (q >> 32);
}

// a b mod 2^64
CSD_HD uint64_t csd_mullo64(uint64_t a, uint64_t b) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    return (uint64_t)a0 * b0 + ((uint64_t)(a0 * b1 + a1 * b0) << 32);
}

// a modulus 0 < N < 2^62 for these steps
struct CsdFastMod {
    uint64_t ninv;         // -N^-1 mod 2^64
    uint32_t n0, n1, n1b;  // N = n1 2^32 + n0; n1b = n1 + [n0 >= 2^31]
};

CSD_HD CsdFastMod csd_fast_mod(int64_t N, int64_t ninv) {
    const uint32_t n0 = (uint32_t)(uint64_t)N, n1 = (uint32_t)((uint64_t)N >> 32);
    return CsdFastMod{(uint64_t)ninv, n0, n1, n1 + (n0 >> 31)};
}

// a fixed multiplier |c| < 2^63 - 2^31: balanced halves and c ninv mod 2^64
struct CsdFastConst {
    CsdBal c;
    uint64_t cn;
};

CSD_HD CsdFastConst csd_fast_const(int64_t c, int64_t ninv) { return CsdFastConst{csd_bal(c), (uint64_t)c * (uint64_t)ninv}; }

// floor(m N / 2^64), m taken as a signed 64-bit value (csd_smulhi_pos(m, N))
CSD_HD int64_t csd_mulhi_mn(const CsdFastMod& f, uint64_t m) {
    const uint32_t m0 = (uint32_t)m;
    const int32_t m1 = (int32_t)(uint32_t)(m >> 32);
    const uint64_t t = ((uint64_t)m0 * f.n0) >> 32;
    const uint64_t v = (uint64_t)m0 * f.n1 + t;
    const int64_t w = (int64_t)m1 * (int32_t)f.n0 + (int64_t)v;
    return (int64_t)m1 * (int64_t)(int32_t)f.n1b + (w >> 32);
}

// red_words(lo, hi) given its quotient mm = lo ninv mod 2^64
CSD_HD int64_t csd_redc_q(const CsdFastMod& f, int64_t hi, uint64_t mm) {
    return (int64_t)((uint64_t)hi + (uint64_t)csd_mulhi_mn(f, mm) + (mm != 0 ? 1u : 0u));
}

// csd_mont_mul(a, c) for |a| < 2^63 - 2^31
CSD_HD int64_t csd_mont_mul_fc(const CsdFastMod& f, int64_t a, const CsdFastConst& c) {
    return csd_redc_q(f, csd_mulhi_bal(csd_bal(a), c.c), csd_mullo64((uint64_t)a, c.cn));
}

// csd_double_red128 of x = hi 2^64 + lo for |x| < 2^126 (the first reduction then fits an i64)
CSD_HD int64_t csd_double_red_fast(const CsdFastMod& f, uint64_t lo, int64_t hi) {
    const int64_t r1 = csd_redc_q(f, hi, csd_mullo64(lo, f.ninv));
    return csd_redc_q(f, r1 >> 63, csd_mullo64((uint64_t)r1, f.ninv));
}
