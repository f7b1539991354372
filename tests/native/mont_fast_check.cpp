// Checks the fast Montgomery forms (csd_mont_fast.hpp) against the general ones (csd_arith.hpp)
// on random and boundary operands for several moduli below 2^57; prints the mismatch count.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../cyclosubfield-dbfv_amd/csrc/csd_mont_fast.hpp"

static CsdMod make_mod(int64_t N) {
    CsdMod m;
    m.N = N;
    uint64_t inv = 1;
    for (int i = 0; i < 7; i++) inv *= 2 - (uint64_t)N * inv;  // N^-1 mod 2^64 by Newton steps
    m.ninv = (int64_t)(0 - inv);
    m.r2 = (int64_t)((((u128)1 << 127) % (u128)N) * 2 % (u128)N);
    m.r3 = 0;
    return m;
}

int main(int argc, char** argv) {
    const long iters = argc > 1 ? atol(argv[1]) : 1000000;
    std::mt19937_64 g(1234);
    long bad = 0, n = 0;
    const int64_t mods[] = {(int64_t)0x1FFFFFFFFFFFFFFLL - 58, (int64_t)0x100000000000001LL + 6, (int64_t)0xFFFFFFFFFFFFC5LL,
                            (int64_t)0x1FFFFFFFFFFFF07LL, 1000003, (int64_t)0x3FFFFFFFFFFFFFFLL - 56};
    for (int64_t N : mods) {
        if (N % 2 == 0) N++;
        const CsdMod m = make_mod(N);
        const CsdFastMod f = csd_fast_mod(m.N, m.ninv);
        for (long it = 0; it < iters; it++) {
            // double_red128 over |x| < 2^126, boundaries included
            i128 S = ((i128)(int64_t)g() << 61) ^ (i128)(uint64_t)g();
            if (it % 7 == 0) S = (i128)(int64_t)g() >> (g() % 64);
            if (it % 11 == 0) S = -S;
            if (it % 13 == 0) S = ((i128)1 << 125) - 1 - (it & 5);
            if (it % 17 == 0) S = -((i128)1 << 125) + (it & 3);
            if (it % 29 == 0) S = (i128)(it & 3) - 1;
            bad += csd_double_red128(m, S) != csd_double_red_fast(f, (uint64_t)S, (int64_t)(S >> 64));
            // Montgomery products by a fixed multiplier, |a| and |c| below 2^63 - 2^31
            int64_t a = (int64_t)(g() >> 1) * (it % 2 ? 1 : -1);
            if (it % 3 == 0) a = (int64_t)(g() % (uint64_t)(4 * N)) - 2 * N;
            if (it % 19 == 0) a = INT64_MAX - ((int64_t)1 << 31) - (it & 7);
            if (it % 23 == 0) a = INT64_MIN + ((int64_t)1 << 31) + 1 + (it & 7);
            if (it % 31 == 0) a = 0;
            int64_t c = (int64_t)(g() % (uint64_t)N) - (it % 2 ? N / 2 : 0);
            if (it % 37 == 0) c = N;
            if (it % 41 == 0) c = m.r2;
            bad += csd_mont_mul(m, a, c) != csd_mont_mul_fc(f, a, csd_fast_const(c, m.ninv));
            // the reduction of any low word (quotients near +-2^63 included) and the high product
            uint64_t lo = g();
            if (it % 43 == 0) lo = ((uint64_t)1 << 63) - 1 - (it & 0xffff);
            if (it % 47 == 0) lo = ((uint64_t)1 << 63) + (it & 0xffff);
            const int64_t hi = (int64_t)(g() % (uint64_t)N) - N / 2;
            bad += csd_red_words(m, lo, hi) != csd_redc_q(f, hi, lo * (uint64_t)m.ninv);
            uint64_t mm = g();
            if (it % 53 == 0) mm = ((uint64_t)1 << 63) - 1 - (it & 0xffff);
            if (it % 59 == 0) mm = ((uint64_t)1 << 63) + (it & 0xffff);
            bad += csd_smulhi_pos((int64_t)mm, m.N) != csd_mulhi_mn(f, mm);
            n += 4;
        }
    }
    printf("checked %ld mismatches %ld\n", n, bad);
    return bad != 0;
}
