// Host-side check of the product's 64-bit modular arithmetic (cyclosubfield-dbfv_amd/csrc/
// csd_arith.hpp, the same source the gfx950 kernels compile) against the oracle's exported
// restatement of hal/mont.rs and BarrettMul (hal/ntt.rs:126-158), on random and extreme
// inputs. Built and run by tests/test_arith_host.py; prints "ok <count>" or the first mismatch.
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "../cyclosubfield-dbfv_amd/csrc/csd_arith.hpp"

extern "C" {
typedef struct {
    int64_t N, neg_inv_N, r2, r3;
} orc_modulus;
int orc_create_modulus(int64_t prime, orc_modulus* out);
int orc_red128(const orc_modulus* m, int64_t lo, int64_t hi, int64_t* out);
int orc_double_red128(const orc_modulus* m, int64_t lo, int64_t hi, int64_t* out);
int orc_barrett_new(int64_t value, int64_t n, int64_t* z, int64_t* u);
int orc_barrett_mul(int64_t z, int64_t u, int64_t x, int64_t n, int64_t* out);
}

static uint64_t s = 0x243F6A8885A308D3ull;
static uint64_t rnd() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int64_t edge(int k) {
    static const int64_t e[] = {0, 1, -1, INT64_MAX, INT64_MIN, INT64_MIN + 1, 2, -2, (int64_t)1 << 62, -((int64_t)1 << 62)};
    return e[k % 10];
}

int main(int argc, char** argv) {
    long iters = argc > 1 ? atol(argv[1]) : 1000000;
    const int64_t moduli[] = {97, 65537, 144115188075835777, 144115176243904321, 4611686018427379009, 3074457345618256921};
    long count = 0;
    for (int64_t N : moduli) {
        orc_modulus om;
        orc_create_modulus(N, &om);
        CsdMod m{om.N, om.neg_inv_N, om.r2, om.r3};
        for (long i = 0; i < iters; i++) {
            int64_t lo = i < 100 ? edge((int)i) : (int64_t)rnd();
            int64_t hi = i < 100 ? edge((int)(i / 10)) : (int64_t)rnd() >> (rnd() % 64);
            i128 x = (i128)(((u128)(uint64_t)hi << 64) | (uint64_t)lo);
            int64_t want, got;
            orc_red128(&om, lo, hi, &want);
            got = csd_red128(m, x);
            if (want != got) {
                printf("red128 mismatch N=%lld lo=%lld hi=%lld: %lld vs %lld\n", (long long)N, (long long)lo, (long long)hi,
                       (long long)want, (long long)got);
                return 1;
            }
            orc_double_red128(&om, lo, hi, &want);
            got = csd_double_red128(m, x);
            if (want != got) {
                printf("double_red128 mismatch N=%lld\n", (long long)N);
                return 1;
            }
            int64_t a = i < 100 ? edge((int)i) : (int64_t)rnd(), b = i < 100 ? edge((int)(i / 10)) : (int64_t)rnd();
            orc_red128(&om, (int64_t)((uint64_t)a * (uint64_t)b), (int64_t)(((i128)a * (i128)b) >> 64), &want);
            got = csd_mont_mul(m, a, b);
            if (want != got) {
                printf("mont_mul mismatch N=%lld a=%lld b=%lld\n", (long long)N, (long long)a, (long long)b);
                return 1;
            }
            if (N < ((int64_t)1 << 62)) {
                int64_t z, u;
                const int64_t value = i % 64 == 0 ? 1 : (int64_t)(rnd() % (uint64_t)N);
                orc_barrett_new(value, N, &z, &u);
                int64_t xx = i < 100 ? edge((int)i) : (int64_t)rnd();
                orc_barrett_mul(z, u, xx, N, &want);
                got = csd_barrett(CsdBarrett{z, u}, xx, N);
                if (want != got) {
                    printf("barrett mismatch N=%lld z=%lld x=%lld\n", (long long)N, (long long)z, (long long)xx);
                    return 1;
                }
                // the device tables' folded constant gives the same bits; only the constant 1 reads as one
                const CsdBarrett f = csd_barrett_fold(CsdBarrett{z, u}, N);
                got = csd_barrett_folded(f, xx, N);
                if (want != got || csd_folded_is_one(f, N) != (value == 1)) {
                    printf("folded barrett mismatch N=%lld z=%lld x=%lld\n", (long long)N, (long long)z, (long long)xx);
                    return 1;
                }
            }
            count++;
        }
    }
    printf("ok %ld\n", count);
    return 0;
}
