// 64-bit modular arithmetic shared by the host plan builders and the gfx950 kernels.
//
// Montgomery reduction with R = 2^64 (hal/mont.rs:38-76) and Barrett multiplication by a
// constant (hal/ntt.rs:126-158). The device forms avoid forming full 128-bit sums where the
// low word is known: for x = lo + 2^64 hi and m = lo * (-N^-1) mod 2^64, the low word of
// m * N is -lo mod 2^64, so (x + m N) >> 64 = hi + mulhi(m, N) + [lo != 0].
#pragma once
#include <stdint.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

#if defined(__HIPCC__)
#define CSD_HD __host__ __device__ __forceinline__
#else
#define CSD_HD inline
#endif

struct CsdMod {
    int64_t N;
    int64_t ninv;  // -N^-1 mod 2^64
    int64_t r2;    // 2^128 mod N (0 mapped to N)
    int64_t r3;    // 2^192 mod N (0 mapped to N)
};

// a multiplication by a fixed constant z: u = round(-z 2^64 / N)
struct CsdBarrett {
    int64_t z;
    int64_t u;
};

// The signed 64 x 64 -> 128 products below are formed as one unsigned product (four 32-bit
// multiply-adds on gfx950) plus sign corrections of the high word, instead of the compiler's
// sign-extended 128 x 128 multiply.
CSD_HD uint64_t csd_umulhi(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * (u128)b) >> 64); }

// (low, high) words of the signed product a * b
CSD_HD void csd_smul_wide(int64_t a, int64_t b, uint64_t& lo, int64_t& hi) {
    u128 p = (u128)(uint64_t)a * (uint64_t)b;
    lo = (uint64_t)p;
    hi = (int64_t)((uint64_t)(p >> 64) - (a < 0 ? (uint64_t)b : 0) - (b < 0 ? (uint64_t)a : 0));
}

// high word of the signed 128-bit product
CSD_HD int64_t csd_smulhi(int64_t a, int64_t b) {
    uint64_t lo;
    int64_t hi;
    csd_smul_wide(a, b, lo, hi);
    return hi;
}

// high word of a signed * non-negative product
CSD_HD int64_t csd_smulhi_pos(int64_t a, int64_t b_nonneg) {
    uint64_t h = csd_umulhi((uint64_t)a, (uint64_t)b_nonneg);
    return (int64_t)(h - (a < 0 ? (uint64_t)b_nonneg : 0));
}

// red128: |x| <= 2^63 N + N  ->  [-N, N], congruent to 2^-64 x (mont.rs:38-45)
CSD_HD int64_t csd_red_words(const CsdMod& m, uint64_t lo, int64_t hi) {
    int64_t mm = (int64_t)(lo * (uint64_t)m.ninv);
    return (int64_t)((uint64_t)hi + (uint64_t)csd_smulhi_pos(mm, m.N) + (lo != 0 ? 1u : 0u));
}
CSD_HD int64_t csd_red128(const CsdMod& m, i128 x) { return csd_red_words(m, (uint64_t)x, (int64_t)(x >> 64)); }

// red128(a * b): the Montgomery product of two i64 values
CSD_HD int64_t csd_mont_mul(const CsdMod& m, int64_t a, int64_t b) {
    uint64_t lo;
    int64_t hi;
    csd_smul_wide(a, b, lo, hi);
    return csd_red_words(m, lo, hi);
}

// double_red128: any i128 -> [-N, N], congruent to 2^-128 x (mont.rs:57-66)
CSD_HD int64_t csd_double_red128(const CsdMod& m, i128 x) {
    uint64_t lo = (uint64_t)x;
    int64_t hi = (int64_t)(x >> 64);
    int64_t mm = (int64_t)(lo * (uint64_t)m.ninv);
    int64_t t_hi = csd_smulhi_pos(mm, m.N);
    // r1 = hi + t_hi + carry(lo + lo(t)), lo(t) = -lo mod 2^64, as an i128
    i128 r1 = (i128)hi + (i128)t_hi + (lo != 0 ? 1 : 0);
    return csd_red128(m, r1);
}

CSD_HD int64_t csd_to_mont64(const CsdMod& m, int64_t x) { return csd_mont_mul(m, x, m.r2); }
CSD_HD int64_t csd_to_mont128(const CsdMod& m, i128 x) { return csd_mont_mul(m, csd_red128(m, x), m.r3); }

// cmod_pow2 (ring.rs:86-91) as bconv_correct_pow2 uses it (cpu.rs:982-1057): x mod 2^k in
// (-2^(k-1), 2^(k-1)]
CSD_HD i128 csd_pow2_centre(i128 x, int k) {
    u128 mask = ((u128)1 << k) - 1;
    i128 reduced = (i128)((u128)x & mask);
    i128 modulus = (i128)1 << k;
    return reduced > modulus / 2 ? reduced - modulus : reduced;
}

// BarrettMul::mul (ntt.rs:149-152): |x| <= 2^63 -> |t| <= 3N/4, t = z x mod N
CSD_HD int64_t csd_barrett(CsdBarrett b, int64_t x, int64_t n) {
    uint64_t lo;
    int64_t hi;
    csd_smul_wide(b.u, x, lo, hi);
    // q = (u x + 2^63) >> 64 = hi + [lo >= 2^63]
    uint64_t q = (uint64_t)hi + (lo >> 63);
    return (int64_t)((uint64_t)b.z * (uint64_t)x + q * (uint64_t)n);
}

// The same product with the sign correction of u folded into the constant. The signed high word
// is hi(u x) = hi_uu - [u < 0] x - [x < 0] u (mod 2^64), hi_uu that of the unsigned product, so
//   z x + q N = (z - [u < 0] N) x + (hi_uu + [lo >= 2^63] - [x < 0] u) N   (mod 2^64).
// The device tables hold z' = z - [u < 0] N (csd_barrett_fold); csd_barrett_folded returns the
// bits csd_barrett returns for (z, u), with one select and one 64-bit add fewer.
CSD_HD CsdBarrett csd_barrett_fold(CsdBarrett b, int64_t n) {
    return CsdBarrett{b.u < 0 ? (int64_t)((uint64_t)b.z - (uint64_t)n) : b.z, b.u};
}
CSD_HD int64_t csd_barrett_folded(CsdBarrett f, int64_t x, int64_t n) {
    const u128 p = (u128)(uint64_t)f.u * (uint64_t)x;
    const uint64_t lo = (uint64_t)p;
    const uint64_t q = (uint64_t)(p >> 64) + (lo >> 63) - (x < 0 ? (uint64_t)f.u : 0);
    return (int64_t)((uint64_t)f.z * (uint64_t)x + q * (uint64_t)n);
}
// the folded form of the constant 1 (z = 1 has u < 0)
CSD_HD bool csd_folded_is_one(const CsdBarrett& f, int64_t n) { return f.z == (int64_t)(1 - (uint64_t)n); }
